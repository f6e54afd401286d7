"""The oracle against the reference's own golden fixture (box/gas-jump, HLLC order 2 +
WENO-XYZ + Wray RK3 with the reference's time-step rules), on the CPU."""
import numpy as np

from oracle.oracle import Oracle
from tests import gasjump


def test_oracle_matches_reference_golden_gas_jump():
    g = gasjump.load()
    m, num, cv0 = gasjump.case()
    cv, dv, st = Oracle(m, num).run_until(cv0, gasjump.END_TIME)
    assert st["rc"] == 0
    assert abs(st["time"] - g["time"]) <= 1e-9 * g["time"]
    prof, spread = gasjump.columns(m, cv)
    assert spread <= 1e-12 * np.abs(cv).max()      # y/z invariance, as in the golden
    worst, fails = gasjump.compare(m, cv, g)
    assert not fails, fails[:5]
    assert worst < 1e-13, worst


def test_golden_totals_conserved():
    """Mass and energy of the closed box equal their initial values (checks.py:351-369)."""
    g = gasjump.load()
    m, num, cv0 = gasjump.case()
    cv, _, _ = Oracle(m, num).run_until(cv0, gasjump.END_TIME)
    vol = m.vol
    for k, key in ((0, "gas_rho"), (4, "gas_rhoE")):
        tot0 = float((cv0[:, k] * vol).sum())
        tot = float((cv[:, k] * vol).sum())
        assert abs(tot - tot0) <= 1e-13 * abs(tot0)
        assert abs(tot - g["totals"][key]) <= 1e-10 * abs(g["totals"][key])
