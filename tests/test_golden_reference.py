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


def test_oracle_near_reference_golden_gas_jump_constrained():
    """The same case with constrained gradients at the walls (CRECONSTC 1), golden
    tests/goldens/box/gas-jump/closed/np2.  Which wall faces become constraint rows
    depends on the order of the stencil members and on how equidistant faces are
    ranked (RFLU_AddBFaces collects at most 18 in member order, RFLU_SortBFaces keeps
    9), and that order follows the reference's cell numbering, which comes from its
    shipped mesh generator and cannot be rebuilt here; permuting the member order moves
    the result by ~1.5e-4, and so does the rounding of the geometry: equidistant wall
    faces are ranked by squared distances that differ in the last bits, and a box whose
    hexahedra list their vertices in another order (another start of each face's
    vertex cycle) ranks them differently (measured worst: rho u 8.7e-4 and 1.0e-3, rho E
    6.4e-5 and 1.2e-4 for two vertex orders).  So the golden pins the constrained run
    only to that band: per x column, the sorted values of every field agree within
    1.5e-3 of the field's scale (density and energy within 2e-4), and the conserved
    totals to 1e-10.  The
    constrained fit itself is pinned by tests/test_oracle_pins.py
    (test_constrained_gradient_exact_on_admissible_linear_fields)."""
    g = gasjump.load("gas_jump_closed_np2")
    m, num, cv0 = gasjump.case()
    num.creconst = 1
    cv, dv, st = Oracle(m, num).run_until(cv0, gasjump.END_TIME)
    assert st["rc"] == 0
    assert abs(st["time"] - g["time"]) <= 1e-9 * g["time"]
    gas = g["gas"]
    allq = np.concatenate([np.array([gas[r]["q%d" % k] for k in range(1, 6)]).T for r in gas])
    scale = np.abs(allq).max(0)
    scale[1:4] = scale[1:4].max()
    ix = np.round(m.cofg[:, 0] / 0.001 - 0.5).astype(int)
    worst = np.zeros(5)
    for r, base in (("2", 0), ("1", 4)):
        q = np.array([gas[r]["q%d" % k] for k in range(1, 6)]).T
        for col in range(4):
            for f in range(5):
                gk = np.sort(q[np.arange(q.shape[0]) % 4 == col, f])
                ours = np.sort(cv[ix == base + col, f])
                worst[f] = max(worst[f], np.abs(gk - ours).max() / scale[f])
    print("per-field worst value-set difference / scale", worst)
    assert worst.max() <= 1.5e-3 and worst[0] <= 2e-4 and worst[4] <= 2e-4, worst
    vol = m.vol
    for k, key in ((0, "gas_rho"), (4, "gas_rhoE")):
        tot = float((cv[:, k] * vol).sum())
        assert abs(tot - g["totals"][key]) <= 1e-10 * abs(g["totals"][key])
