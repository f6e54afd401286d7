"""The autotuner on GPU regions: its exact choice gives the bits of the default
variants, the choice is cached per mesh/scheme signature (in memory and in the file
named by ROCFLU_AMD_TUNE_CACHE) and applied again from there, and a mixed region's
choice among all variants, the FP32 tile records included, stays within the mixed
mode's 1e-6 of the FP64 run after one limited step."""
import numpy as np
import pytest

from rocflu_amd import autotune, cases
from rocflu_amd.solver import GpuRegion, Numerics

pytestmark = pytest.mark.gpu

KEYS = ("xcd", "face_block", "cell_block", "grad_tile", "grad_lds", "grad_unroll",
        "grad_weights", "grad_fast", "face_tile", "update_tile")


def _run(m, num, cv0, steps, tune=None):
    g = GpuRegion(m, num)
    try:
        choice = None
        if tune is not None:
            g.upload(cv0)
            choice = autotune.tune(g, reps=3, **tune)
        g.upload(cv0)
        g.time_steps(steps, graph=True)
        return g.download()[0], choice
    finally:
        g.close()


def test_exact_choice_keeps_the_bits_and_is_cached(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv("ROCFLU_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune._cache.clear()
    m = cases.tet_box(n=12, jitter=0.08, seed=2)
    num = Numerics(discr=3, order=2, reconst=11, rk=4, cfl=0.5, dt_imposed=1e-3)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    print("case: autotuner, 12^3 x 6 FP64, HLLC-2 + Venkatakrishnan, 10 steps", flush=True)
    ref, _ = _run(m, num, cv0, 10)
    out, ch = _run(m, num, cv0, 10, tune=dict(exact=True))
    print("choice:", {k: ch[k] for k in KEYS}, flush=True)
    assert not ch["cached"] and ch["timings_ms"]
    assert ch["grad_fast"] == 0 and ch["grad_lds"] in (0, 1)
    assert np.array_equal(out, ref)
    # the same signature again: from memory, then from the file
    for clear in (False, True):
        if clear:
            autotune._cache.clear()
        out2, ch2 = _run(m, num, cv0, 10, tune=dict(exact=True))
        assert ch2["cached"]
        assert {k: ch2[k] for k in KEYS} == {k: ch[k] for k in KEYS}
        assert np.array_equal(out2, ref)
    autotune._cache.clear()


def test_mixed_choice_with_fp32_records(gpu, tmp_path, monkeypatch):
    monkeypatch.setenv("ROCFLU_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune._cache.clear()
    m = cases.tet_box(n=12, jitter=0.08, seed=2)
    out = {}
    for prec in (0, 1):
        num = Numerics(discr=3, order=2, reconst=11, rk=4, cfl=0.5, dt_imposed=1e-3,
                       precision=prec)
        cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
        print(f"case: autotuner, 12^3 x 6 precision {prec}, 1 step", flush=True)
        out[prec], ch = _run(m, num, cv0, 1,
                             tune=None if prec == 0 else dict(exact=False, fp32_records=True))
        if prec:
            print("choice:", {k: ch[k] for k in KEYS}, flush=True)
            assert not ch["cached"]
            # the FP32 records were timed beside the FP64-record variants
            assert any(k.startswith("grad_t") and "_lds2_" in k for k in ch["timings_ms"])
            assert any(k.startswith("grad_t") and "_lds1_" in k for k in ch["timings_ms"])
    err = np.array([np.linalg.norm(out[1][:, k] - out[0][:, k]) / np.linalg.norm(out[0][:, k])
                    for k in range(5)])
    print("tuned mixed vs FP64:", err, flush=True)
    assert np.all(err <= 1e-6), err
    autotune._cache.clear()
