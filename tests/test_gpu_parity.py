"""GPU parity of the HIP path against the CPU oracle on the same seeded inputs.

Bar: FP64 runs agree with the oracle to 1e-12 relative L2 per conserved variable
(the FP64 tolerance of the build's target); mixed precision is compared with the
FP64 GPU run.  All GPU calls go through the C-ABI (librocflu_amd.so).
"""
import numpy as np
import pytest

from rocflu_amd import cases, mesh as rmesh
from rocflu_amd.solver import GpuBatch, GpuRegion, Numerics
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for k in range(a.shape[1]):
        den = np.linalg.norm(b[:, k])
        num = np.linalg.norm(a[:, k] - b[:, k])
        out.append(num / den if den > 0 else num)
    return np.array(out)


def run_both(m, num, cv0, nsteps, mode="fused", graph=False):
    print(f"case: cells={m.n_cells} discr={num.discr} order={num.order} reconst={num.reconst} "
          f"rk={num.rk} prec={num.precision} steps={nsteps} mode={mode}", flush=True)
    o = Oracle(m, num)
    cvo, dvo, st = o.run(cv0, nsteps)
    assert st["rc"] == 0
    g = GpuRegion(m, num)
    try:
        g.upload(cv0)
        if mode == "fused":
            g.time_steps(nsteps, graph=graph)
        else:
            for _ in range(nsteps):
                g.runge_kutta()
        cvg, dvg = g.download()
        gst = g.stats()
    finally:
        g.close()
    return cvo, dvo, st, cvg, dvg, gst


SOD_NUM = dict(order=1, rk=2, cfl=0.5, dt_imposed=1e-3, dt_min_limit=0.0)


@pytest.mark.parametrize("discr", [1, 3, 4])
def test_sod_first_order_100_steps(gpu, discr):
    """Config 1: 64x16x16 jittered hex box, Sod along x, slip walls + outflow ends."""
    m = cases.sod_box()
    num = Numerics(discr=discr, **SOD_NUM)
    cv0 = cases.sod(m, num, 32.0)
    cvo, dvo, st, cvg, dvg, gst = run_both(m, num, cv0, 100)
    err = rel_l2(cvg, cvo)
    print("rel L2 per variable:", err, "max abs:", np.abs(cvg - cvo).max(), flush=True)
    assert np.all(err <= TOL_FP64), err
    assert gst["steps"] == 100
    assert abs(gst["time"] - st["time"]) <= 1e-12 * st["time"]


def test_sod_entry_points_match_fused(gpu):
    """The per-callee entry points (RungeKuttaMP order) give the fused loop's state."""
    m = cases.sod_box(nx=32, ny=8, nz=8, subsonic=True)
    num = Numerics(discr=3, **SOD_NUM)
    cv0 = cases.sod(m, num, 16.0)
    cvo, _, _, cvg, _, _ = run_both(m, num, cv0, 20, mode="steps")
    g = GpuRegion(m, num)
    g.upload(cv0)
    g.time_steps(20)
    cvf, _ = g.download()
    g.close()
    assert np.array_equal(cvg, cvf)
    assert np.all(rel_l2(cvg, cvo) <= TOL_FP64)


def test_graph_replay_matches_eager(gpu):
    m = cases.sod_box(nx=32, ny=8, nz=8)
    num = Numerics(discr=1, **SOD_NUM)
    cv0 = cases.sod(m, num, 16.0)
    res = []
    for graph in (False, True):
        g = GpuRegion(m, num)
        g.upload(cv0)
        for _ in range(3):  # several replays with host work between them
            g.time_steps(5, graph=graph)
            _ = g.stats()
        res.append(g.download()[0])
        g.close()
    assert np.array_equal(res[0], res[1])


@pytest.mark.parametrize("rk", [1, 3, 4])
def test_other_integrators(gpu, rk):
    m = cases.sod_box(nx=32, ny=8, nz=8, subsonic=True)
    num = Numerics(discr=3, order=1, rk=rk, cfl=0.5 if rk != 3 else 0.2, dt_imposed=1e-3)
    cv0 = cases.sod(m, num, 16.0)
    cvo, _, _, cvg, _, _ = run_both(m, num, cv0, 30)
    assert np.all(rel_l2(cvg, cvo) <= TOL_FP64)


def small_tet(n=8, bc=rmesh.BC_SLIPWALL, seed=5):
    return cases.tet_box(n=n, jitter=0.08, seed=seed, bc=bc)


@pytest.mark.parametrize("discr,reconst", [(3, 11), (4, 11), (1, 11), (3, 2), (3, 0), (3, 10),
                                           (1, 10)])
def test_second_order_tets(gpu, discr, reconst):
    """Config 2's scheme on a reduced 8^3 x 6 Kuhn mesh (3,072 tets), Mach-2 shock."""
    m = small_tet()
    num = Numerics(discr=discr, order=2, reconst=reconst, rk=2, cfl=0.5, dt_imposed=1e-3)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    cvo, _, _, cvg, _, _ = run_both(m, num, cv0, 20)
    err = rel_l2(cvg, cvo)
    print("rel L2 per variable:", err, flush=True)
    assert np.all(err <= TOL_FP64), err


def test_second_order_hex_weno_outflow(gpu):
    """WENO-XYZ + clip on hexes with subsonic outflow ends, first- and second-order patches."""
    m = cases.sod_box(nx=24, ny=6, nz=6, subsonic=True, stencils=True)
    for po in ([1] * 6, [2, 2, 1, 1, 1, 1]):
        num = Numerics(discr=3, order=2, reconst=2, rk=2, cfl=0.4, dt_imposed=1e-3,
                       patch_order=po)
        cv0 = cases.sod(m, num, 12.0)
        cvo, _, _, cvg, _, _ = run_both(m, num, cv0, 25)
        err = rel_l2(cvg, cvo)
        print("patch order", po, "rel L2:", err, flush=True)
        assert np.all(err <= TOL_FP64), err


def test_residual_matches_oracle(gpu):
    m = small_tet(n=6)
    for order, reconst in ((1, 0), (2, 11), (2, 2)):
        num = Numerics(discr=3, order=order, reconst=reconst, rk=2)
        cv0 = cases.planar_shock(m, num, 0.5 * m.dims[0], mach=2.0, noise=0.01, seed=3)
        o = Oracle(m, num)
        ro = o.residual(cv0)
        g = GpuRegion(m, num)
        g.upload(cv0)
        rg = g.residual()
        g.close()
        scale = np.abs(ro).max(axis=0)
        assert np.all(np.abs(rg - ro).max(axis=0) <= 1e-13 * scale), (order, reconst)


def test_time_step_matches_oracle(gpu):
    m = cases.sod_box(nx=16, ny=8, nz=8)
    num = Numerics(discr=3, order=1, rk=2, cfl=0.5, dt_imposed=1e-3)
    cv0 = cases.sod(m, num, 8.0)
    o = Oracle(m, num)
    g = GpuRegion(m, num)
    # first step: imposed dt (time <= 1e-8); later: CFL dt
    for t in (0.0, 1.0):
        g.upload(cv0)
        dtg = g.time_step(t)
        dto = o.min_timestep(o.local_dt(cv0), t)
        assert dtg == dto, (t, dtg, dto)
    g.close()


def _mixed_vs_fp64(m, reconst, steps, floor=False, mixed_opts=(), keys=("fp64", "mixed")):
    """Final states of the FP64 and mixed runs (and, with floor, of an FP64 run whose
    gradient sums in another order: the moment form) from the same Mach-2 shock with
    1% density noise, HLLC order 2, SSP-RK3.  mixed_opts: (option, value) pairs set on
    the mixed run; keys: the runs to make ("fp64", "mixed", "mixed_opts")."""
    out = {}
    runs = [(k, 0 if k == "fp64" else 1, 0) for k in keys]
    if floor:
        runs.append(("reorder", 0, 1))
    for key, prec, fast in runs:
        num = Numerics(discr=3, order=2, reconst=reconst, rk=4, cfl=0.5, dt_imposed=1e-3,
                       precision=prec)
        cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
        g = GpuRegion(m, num)
        try:
            if fast:
                g.set_option(GpuRegion.OPT_GRAD_FAST, 1)
            if key == "mixed_opts":
                for opt, val in mixed_opts:
                    g.set_option(opt, val)
                    assert g.get_option(opt) == val, (opt, val)
            g.upload(cv0)
            g.time_steps(steps)
            out[key] = g.download()[0]
        finally:
            g.close()
        if key != "reorder":
            out[key + "_totals"] = [(float(np.dot(cv0[:, k], m.vol)),
                                     float(np.dot(out[key][:, k], m.vol))) for k in (0, 4)]
    return out


@pytest.mark.parametrize("reconst,steps", [(0, 100), (2, 100), (11, 1)])
def test_mixed_precision_error_vs_fp64(gpu, reconst, steps):
    """The mixed mode (FP32 gradients and reconstruction increments, FP64 fluxes and
    state) within 1e-6 relative L2 of the FP64 run in every conserved variable: after
    100 steps without a modifier and with WENO-XYZ, after one step with the
    Venkatakrishnan limiter."""
    m = cases.tet_box(n=12, jitter=0.08, seed=2)
    print(f"case: mixed vs FP64, 12^3 x 6, reconst={reconst}, {steps} steps", flush=True)
    r = _mixed_vs_fp64(m, reconst, steps)
    err = rel_l2(r["mixed"], r["fp64"])
    print("mixed vs FP64 rel L2 per variable:", err, flush=True)
    assert np.all(err <= 1e-6), err
    assert np.any(err > 0.0)


def test_mixed_precision_with_the_limiter_over_steps(gpu):
    """With the Venkatakrishnan limiter, rounding-level differences in the gradients
    grow with the steps in the momenta normal to the shock (they start at zero; the
    limiter's face minimum and its additive TINY switch on differences of that size):
    after 30 steps the mixed run stays within 1e-6 of the FP64 run in rho, rho u and
    rho E and within 1e-4 in rho v and rho w.  An FP64 run whose gradient only sums in
    another order (the moment form) is reported beside it."""
    m = cases.tet_box(n=12, jitter=0.08, seed=2)
    print("case: mixed vs FP64 and FP64 reorder, 12^3 x 6, Venkatakrishnan, 30 steps",
          flush=True)
    r = _mixed_vs_fp64(m, 11, 30, floor=True)
    err = rel_l2(r["mixed"], r["fp64"])
    flo = rel_l2(r["reorder"], r["fp64"])
    print("mixed vs FP64:", err, "FP64 reorder:", flo, flush=True)
    assert np.all(err[[0, 1, 4]] <= 1e-6), err
    assert np.all(err[[2, 3]] <= 1e-4), err
    assert np.all(flo > 0.0), flo


@pytest.mark.parametrize("tile,unroll,reconst,steps",
                         [(256, 4, 11, 1), (512, 8, 11, 1), (128, 4, 0, 100), (256, 8, 2, 100)])
def test_mixed_precision_fp32_records(gpu, tile, unroll, reconst, steps):
    """GRAD_LDS 2, the mixed mode's gradient over FP32 tile records (values less the
    tile's first cell's, FP32 member loop): within 1e-6 relative L2 of the FP64 run
    under the same bars as the FP64-record mixed mode (one step with the
    Venkatakrishnan limiter; 100 steps without a modifier and with WENO-XYZ); the
    closed box keeps its totals of rho and rho E to 1e-12; the records change the bits
    of the FP64-record mixed run (the kernel ran), and the option is refused in FP64."""
    m = cases.tet_box(n=12, jitter=0.08, seed=2)
    print(f"case: mixed FP32 records, 12^3 x 6, tile {tile}, unroll {unroll}, "
          f"reconst {reconst}, {steps} steps", flush=True)
    opts = ((GpuRegion.OPT_GRAD_TILE, tile), (GpuRegion.OPT_GRAD_LDS, 2),
            (GpuRegion.OPT_GRAD_UNROLL, unroll))
    r = _mixed_vs_fp64(m, reconst, steps, mixed_opts=opts,
                       keys=("fp64", "mixed", "mixed_opts"))
    err = rel_l2(r["mixed_opts"], r["fp64"])
    print("FP32 records vs FP64:", err, "FP64 records vs FP64:",
          rel_l2(r["mixed"], r["fp64"]), flush=True)
    assert np.all(err <= 1e-6), err
    assert not np.array_equal(r["mixed_opts"], r["mixed"])
    for t0, t in r["mixed_opts_totals"]:
        assert abs(t - t0) <= 1e-12 * abs(t0), (t0, t)
    g = GpuRegion(m, Numerics(discr=3, order=2, reconst=reconst, rk=4, precision=0))
    try:
        with pytest.raises(Exception):
            g.set_option(GpuRegion.OPT_GRAD_LDS, 2)
        assert g.get_option(GpuRegion.OPT_GRAD_LDS) == 1
    finally:
        g.close()
    g = GpuRegion(m, Numerics(discr=3, order=2, reconst=reconst, rk=4, precision=1))
    try:
        g.set_option(GpuRegion.OPT_GRAD_LDS, 2)
        g.set_option(GpuRegion.OPT_GRAD_FAST, 1)   # the moment form stages FP64 records
        assert g.get_option(GpuRegion.OPT_GRAD_LDS) == 1
        g.set_option(GpuRegion.OPT_GRAD_FAST, 0)
        assert g.get_option(GpuRegion.OPT_GRAD_LDS) == 2
    finally:
        g.close()


def test_fp32_flux_split(gpu):
    """The FP32-flux diagnostic split (FP64 gradients, FP32 interior-flux arithmetic with
    the pressure terms in FP64, FP32 mass and energy fluxes stored beside FP64 momenta):
    a closed box keeps its totals of rho and rho E to 1e-12 (each stored flux enters
    both cells of its face), and rho, rho u and rho E stay within 1e-6 of the FP64 run
    after 50 steps without a modifier."""
    m = cases.tet_box(n=12, jitter=0.08, seed=2)
    out = {}
    for prec in (0, 3):
        num = Numerics(discr=3, order=2, reconst=0, rk=4, cfl=0.5, dt_imposed=1e-3,
                       precision=prec)
        cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
        print(f"case: 12^3 x 6, precision {prec}, 50 steps", flush=True)
        g = GpuRegion(m, num)
        try:
            g.upload(cv0)
            g.time_steps(50)
            out[prec] = g.download()[0]
        finally:
            g.close()
        for k in (0, 4):
            t0, t = float(np.dot(cv0[:, k], m.vol)), float(np.dot(out[prec][:, k], m.vol))
            assert abs(t - t0) <= 1e-12 * abs(t0), (prec, k, t0, t)
    err = rel_l2(out[3], out[0])
    print("FP32-flux split vs FP64:", err, flush=True)
    assert np.all(err[[0, 1, 4]] <= 1e-6), err
    assert np.all(err > 0.0), err


def test_gpu_matches_reference_golden_gas_jump(gpu):
    """The GPU path against the reference's golden (box/gas-jump/closed-creconst0):
    fused loop with the end-time clamp, and the per-callee entry points."""
    from tests import gasjump
    g = gasjump.load()
    m, num, cv0 = gasjump.case()
    r = GpuRegion(m, num)
    r.upload(cv0)
    r.set_end_time(gasjump.END_TIME)
    r.time_steps(700, graph=True)
    st = r.stats()
    cvf, _ = r.download()
    r.close()
    assert st["finished"] == 1 and st["steps"] == 569, st
    assert abs(st["time"] - g["time"]) <= 1e-9 * g["time"]
    worst, fails = gasjump.compare(m, cvf, g)
    print("fused: worst error / field scale:", worst, flush=True)
    assert not fails, fails[:5]
    r = GpuRegion(m, num)
    r.upload(cv0)
    n = 0
    while True:
        _, done = r.runge_kutta(end_time=gasjump.END_TIME)
        n += 1
        if done:
            break
    cve, _ = r.download()
    r.close()
    assert n == 569
    assert np.array_equal(cve, cvf)


@pytest.mark.parametrize("reconst", [11, 2])
def test_tuning_options_do_not_change_results(gpu, reconst):
    """Every tuning option (LDS tiles or global gathers, member unroll, XCD-aware block
    order, launch sizes) gives the same bits, and the same bits as the oracle's run."""
    m = small_tet(n=10, bc=rmesh.BC_OUTFLOW)
    num = Numerics(discr=3, order=2, reconst=reconst, rk=2, cfl=0.5, dt_imposed=1e-3)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=3)
    o = Oracle(m, num)
    cvo, _, st = o.run(cv0, 8)
    assert st["rc"] == 0
    ref = None
    variants = [dict(lds=1, unroll=4, xcd=8), dict(lds=0, unroll=4, xcd=8),
                dict(lds=1, unroll=1, xcd=1), dict(lds=1, unroll=2, xcd=3),
                dict(lds=0, unroll=2, xcd=1, blocks=(64, 512, 128)),
                dict(lds=1, unroll=4, xcd=8, weights=1), dict(lds=1, unroll=1, xcd=1, weights=1),
                dict(lds=1, unroll=2, xcd=8, weights=1),
                dict(lds=1, unroll=4, xcd=8, tile=128), dict(lds=1, unroll=2, xcd=8, tile=512),
                dict(lds=1, unroll=8, xcd=8), dict(lds=1, unroll=8, xcd=1, tile=512),
                dict(lds=1, unroll=1, xcd=8, tile=128, weights=1),
                dict(lds=1, unroll=2, xcd=1, tile=512, weights=1),
                # weights formed for 256-cell tiles, then the tiles rebuilt at 128
                dict(lds=1, unroll=2, xcd=8, weights=1, tile_after=128),
                # stored inverse distances, at two tile sizes and rebuilt with the tiles
                dict(lds=1, unroll=4, xcd=8, weights=2), dict(lds=1, unroll=4, xcd=1, tile=512, weights=2),
                dict(lds=1, unroll=4, xcd=8, weights=2, tile_after=128),
                # interior fluxes from LDS face tiles
                dict(lds=1, unroll=4, xcd=8, face_tile=128), dict(lds=1, unroll=4, xcd=1, face_tile=64),
                dict(lds=0, unroll=2, xcd=8, face_tile=128, tile=128),
                # face tiles by blocks of 512 threads
                dict(lds=1, unroll=4, xcd=8, face_tile=256, face_threads=512),
                dict(lds=1, unroll=4, xcd=1, face_tile=64, face_threads=512)]
    for v in variants:
        print("variant", v, flush=True)
        g = GpuRegion(m, num)
        try:
            g.set_option(GpuRegion.OPT_GRAD_LDS, v["lds"])
            g.set_option(GpuRegion.OPT_GRAD_UNROLL, v["unroll"])
            g.set_option(GpuRegion.OPT_XCD, v["xcd"])
            if "blocks" in v:
                g.set_launch(*v["blocks"])
            if "tile" in v:
                g.set_option(GpuRegion.OPT_GRAD_TILE, v["tile"])
            if v.get("weights"):
                g.set_option(GpuRegion.OPT_GRAD_WEIGHTS, v["weights"])
                assert g.get_option(GpuRegion.OPT_GRAD_WEIGHTS) == v["weights"]
            if "face_tile" in v:
                with pytest.raises(Exception):   # tiles of 64, 128 or 256 cells only
                    g.set_option(GpuRegion.OPT_FACE_TILE, 100)
                assert g.get_option(GpuRegion.OPT_FACE_TILE) == 0
                g.set_option(GpuRegion.OPT_FACE_TILE, v["face_tile"])
                assert g.get_option(GpuRegion.OPT_FACE_TILE) == v["face_tile"]
            if "face_threads" in v:
                with pytest.raises(Exception):   # blocks of 256 or 512 threads only
                    g.set_option(GpuRegion.OPT_FACE_THREADS, 384)
                g.set_option(GpuRegion.OPT_FACE_THREADS, v["face_threads"])
                assert g.get_option(GpuRegion.OPT_FACE_THREADS) == v["face_threads"]
            if "tile_after" in v:
                g.set_option(GpuRegion.OPT_GRAD_TILE, v["tile_after"])
                assert g.get_option(GpuRegion.OPT_GRAD_TILE) == v["tile_after"]
                assert g.get_option(GpuRegion.OPT_GRAD_WEIGHTS) == v["weights"]
            g.upload(cv0)
            g.time_steps(8, graph=True)
            cvg, _ = g.download()
        finally:
            g.close()
        if ref is None:
            ref = cvg
            g2 = GpuRegion(m, num)
            try:
                with pytest.raises(Exception):
                    g2.set_option(GpuRegion.OPT_GRAD_TILE, 100)
                assert g2.get_option(GpuRegion.OPT_GRAD_TILE) == 256
            finally:
                g2.close()
            err = rel_l2(cvg, cvo)
            print("rel L2 vs oracle:", err, flush=True)
            assert np.all(err <= TOL_FP64), err
        else:
            assert np.array_equal(cvg, ref), v


@pytest.mark.parametrize("discr,prec", [(1, 0), (4, 0), (3, 1), (4, 1)])
def test_face_tiles_give_the_same_bits(gpu, discr, prec):
    """The LDS face-tile flux (FACE_TILE) and the tiled update (UPDATE_TILE) against the
    one-thread-per-face and one-thread-per-cell kernels, fused and per-callee steps (the
    fused first stage also stores the face spectral radii), FP64 and mixed precision,
    Roe and AUSM+ as well as HLLC."""
    m = small_tet(n=9, bc=rmesh.BC_OUTFLOW)
    num = Numerics(discr=discr, order=2, reconst=11, rk=4, cfl=0.5, dt_imposed=1e-3,
                   precision=prec)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=5)
    outs = {}
    for ft, ut, nt in ((0, 0, 256), (128, 0, 256), (64, 1, 256), (128, 1, 256), (256, 1, 512)):
        for mode in ("fused", "callee"):
            print("face_tile", ft, "update_tile", ut, "threads", nt, mode, flush=True)
            g = GpuRegion(m, num)
            try:
                g.set_option(GpuRegion.OPT_FACE_THREADS, nt)
                g.set_option(GpuRegion.OPT_FACE_TILE, ft)
                g.set_option(GpuRegion.OPT_UPDATE_TILE, ut)
                assert g.get_option(GpuRegion.OPT_UPDATE_TILE) == (1 if ft and ut else 0)
                g.upload(cv0)
                if mode == "fused":
                    g.time_steps(6)
                else:
                    for _ in range(6):
                        g.runge_kutta()
                outs[(ft, ut, mode)] = g.download()[0]
            finally:
                g.close()
    for key, cv in outs.items():
        assert np.array_equal(cv, outs[(0, 0, key[2])]), key


def test_face_tiles_in_a_batch(gpu):
    """Face tiles (and the tiled update) that straddle the cases of a batch."""
    ens = cases.ensemble(n_cases=4, min_cells=600, max_cells=2500, seed=21, dt_imposed=1e-3)
    outs = []
    for ft in (0, 128):
        b = GpuBatch([(m, num) for m, num, _ in ens])
        try:
            b.set_option(GpuRegion.OPT_FACE_TILE, ft)
            b.set_option(GpuRegion.OPT_UPDATE_TILE, 1)
            b.upload([cv0 for _, _, cv0 in ens])
            b.time_steps(5, graph=True)
            outs.append(b.download()[0])
        finally:
            b.close()
    for a, c in zip(outs[0], outs[1]):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("order,opt,discr", [(1, 1, 3), (1, 0, 1), (2, 1, 3), (2, 0, 4)])
def test_inflow_patch(gpu, order, opt, discr):
    """Velocity + temperature inflow (BC_INFLOW_VELTEMP) at x-, subsonic outflow at x+,
    slip walls elsewhere; a blast inside the box."""
    m = rmesh.box(rmesh.TET6, 6, 5, 5, lengths=(6.0, 5.0, 5.0), jitter=0.08, seed=4,
                  bc_type=(rmesh.BC_INFLOW_VELTEMP, rmesh.BC_OUTFLOW) + (rmesh.BC_SLIPWALL,) * 4,
                  bc_opt=(opt, 1, 0, 0, 0, 0), bc_pout=(0.0, 1.0, 0, 0, 0, 0),
                  stencils=order == 2)
    base = Numerics()
    rg = base.gas_prime()[1]
    vals = [[0.4, 0.05, 0.0, 1.1 / rg, 1.2]] + [[0.0] * 5] * 5
    num = Numerics(discr=discr, order=order, reconst=11 if order == 2 else 0, rk=4, cfl=0.5,
                   dt_imposed=1e-3, patch_vals=vals,
                   patch_order=[2, 2, 1, 1, 1, 1] if order == 2 else None)
    cv0 = cases.blast(m, num, (3.0, 2.5, 2.5), 1.0, p_in=5.0)
    cvo, _, st, cvg, _, gst = run_both(m, num, cv0, 30)
    err = rel_l2(cvg, cvo)
    print("rel L2 per variable:", err, "mass in/out", gst["mass_in"], gst["mass_out"], flush=True)
    assert np.all(err <= TOL_FP64), err
    assert abs(gst["mass_in"] - st["mass_in"]) <= 1e-12 * max(st["mass_in"], 1e-30)
    assert st["mass_in"] > 0.0


@pytest.mark.parametrize("batch", [False, True])
def test_mass_bookkeeping_stops_at_the_end_time(gpu, batch):
    """Mass in/out count only the steps a run takes: asking the fused loop for more steps
    than reach MAXTIME leaves them at the oracle's values for the run to MAXTIME
    (regression: steps after the end once added the last stored boundary fluxes)."""
    m = rmesh.box(rmesh.TET6, 6, 5, 5, lengths=(6.0, 5.0, 5.0), jitter=0.08, seed=4,
                  bc_type=(rmesh.BC_INFLOW_VELTEMP, rmesh.BC_OUTFLOW) + (rmesh.BC_SLIPWALL,) * 4,
                  bc_opt=(1, 1, 0, 0, 0, 0), bc_pout=(0.0, 1.0, 0, 0, 0, 0))
    rg = Numerics().gas_prime()[1]
    vals = [[0.4, 0.05, 0.0, 1.1 / rg, 1.2]] + [[0.0] * 5] * 5
    num = Numerics(discr=3, order=2, reconst=11, rk=4, cfl=0.5, dt_imposed=1e-3, patch_vals=vals)
    cv0 = cases.blast(m, num, (3.0, 2.5, 2.5), 1.0, p_in=5.0)
    end = 0.06
    cvo, _, st = Oracle(m, num).run_until(cv0, end)
    assert st["rc"] == 0 and st["mass_in"] > 0.0 and st["mass_out"] > 0.0
    nsteps = 3 * st["steps"] + 2
    if batch:
        # a second case with smaller steps keeps the batch stepping after the first one
        # finishes
        num2 = Numerics(discr=3, order=2, reconst=11, rk=4, cfl=0.2, dt_imposed=1e-3,
                        patch_vals=vals)
        g = GpuBatch([(m, num), (m, num2)])
        try:
            g.upload([cv0, cv0])
            g.set_end_time(end)
            g.time_steps(nsteps, graph=True)
            gst = g.case_stats(0)
            cvg = g.download()[0][0]
        finally:
            g.close()
    else:
        g = GpuRegion(m, num)
        try:
            g.upload(cv0)
            g.set_end_time(end)
            g.time_steps(nsteps, graph=True)
            gst = g.stats()
            cvg = g.download()[0]
        finally:
            g.close()
    print("oracle", st["steps"], st["mass_in"], st["mass_out"], "gpu", gst["steps"],
          gst["mass_in"], gst["mass_out"], flush=True)
    assert gst["finished"] and gst["steps"] == st["steps"]
    assert np.all(rel_l2(cvg, cvo) <= TOL_FP64)
    for k in ("mass_in", "mass_out"):
        assert abs(gst[k] - st[k]) <= 1e-12 * st[k], k


@pytest.mark.parametrize("kind,reconst,discr", [(rmesh.HEX, 2, 3), (rmesh.TET6, 11, 3),
                                                (rmesh.TET6, 0, 4), (rmesh.TET6, 10, 3)])
def test_constrained_gradients(gpu, kind, reconst, discr):
    """CRECONSTC 1 (wall-constrained velocity gradients): the GPU precomputes the rows
    with a Jacobi SVD and the reference's truncation, the oracle re-solves each cell's
    problem with Householder QR; both must agree."""
    n = 8 if kind == rmesh.HEX else 6
    m = rmesh.box(kind, n, n, n, lengths=(float(n),) * 3, jitter=0.08, seed=6,
                  bc_type=(rmesh.BC_SLIPWALL,) * 6)
    num = Numerics(discr=discr, order=2, reconst=reconst, rk=2, cfl=0.4, dt_imposed=1e-3,
                   creconst=1)
    cv0 = cases.blast(m, num, (0.4 * n, 0.5 * n, 0.45 * n), 0.25 * n, p_in=4.0)
    cvo, _, _, cvg, _, _ = run_both(m, num, cv0, 25)
    err = rel_l2(cvg, cvo)
    print("rel L2 per variable:", err, flush=True)
    assert np.all(err <= TOL_FP64), err
    # the constraints act: the unconstrained run differs
    num0 = Numerics(discr=discr, order=2, reconst=reconst, rk=2, cfl=0.4, dt_imposed=1e-3)
    cv_free, _, _ = Oracle(m, num0).run(cv0, 25)
    assert rel_l2(cvg, cv_free)[1] > 1e-8


@pytest.mark.parametrize("prec", [0, 1])
def test_full_size_closed_box_conserves(gpu, prec):
    """Config 2's mesh at full size (80^3 x 6 = 3,072,000 tets, jitter 8%) closed by slip
    walls, its Mach-2 shock with 1% density noise, HLLC order 2 + Venkatakrishnan,
    SSP-RK3: slip walls carry no mass or energy, so the totals of rho and rho E are
    those of the initial state to round-off after 20 steps (FP64 state and fluxes in
    both precision modes; each face flux enters both cells of the face with the same
    bits, so the fluxes cancel in the totals)."""
    m = cases.tet_box(n=80)
    num = Numerics(discr=3, order=2, reconst=11, rk=4, cfl=0.5, dt_imposed=1e-3,
                   precision=prec)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    vol = m.vol
    tot0 = [float(np.dot(cv0[:, k], vol)) for k in (0, 4)]
    print(f"full size closed box: cells={m.n_cells} prec={prec}", flush=True)
    g = GpuRegion(m, num)
    try:
        g.upload(cv0)
        g.time_steps(20)
        cv, _ = g.download()
        st = g.stats()
    finally:
        g.close()
    assert st["n_invalid"] == 0 and st["n_negative"] == 0
    for k, t0 in zip((0, 4), tot0):
        t = float(np.dot(cv[:, k], vol))
        print(f"variable {k}: total {t0:.17g} -> {t:.17g} (rel {abs(t - t0) / abs(t0):.2e})",
              flush=True)
        assert abs(t - t0) <= 1e-12 * abs(t0)
    assert not np.array_equal(cv, cv0)


@pytest.mark.parametrize("discr,reconst", [(3, 0), (4, 2), (1, 0)])
def test_fast_gradient_within_fp64_bar(gpu, discr, reconst):
    """The moment form of the gradient (GRAD_FAST: one reciprocal of |d|^2 and fused
    multiply-adds per stencil member, the inverse moment matrix once per cell) is the
    same least-squares gradient with other rounding: without a modifier or with the
    WENO-XYZ blend, a Mach-2 shock with 1% noise on a 16^3 x 6 mesh (SSP-RK3) stays
    within 1e-12 relative L2 per variable of the oracle after 100 steps, and differs
    from it (it is not the bit-identical path)."""
    m = cases.tet_box(n=16, jitter=0.08, seed=2)
    num = Numerics(discr=discr, order=2, reconst=reconst, rk=4, cfl=0.5, dt_imposed=1e-4)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    print(f"case: fast gradient, cells={m.n_cells} discr={discr} reconst={reconst} steps=100",
          flush=True)
    from oracle import oracle as orc
    orc.set_threads(8)
    try:
        cvo, _, st = Oracle(m, num).run(cv0, 100)
    finally:
        orc.set_threads(1)
    assert st["rc"] == 0
    outs = {}
    for fast in (0, 1):
        g = GpuRegion(m, num)
        try:
            g.set_option(GpuRegion.OPT_GRAD_FAST, fast)
            assert g.get_option(GpuRegion.OPT_GRAD_FAST) == fast
            g.upload(cv0)
            g.time_steps(100)
            outs[fast] = g.download()[0]
        finally:
            g.close()
    assert np.array_equal(outs[0], cvo)
    err = rel_l2(outs[1], cvo)
    print("fast gradient rel L2 per variable after 100 steps:", err, flush=True)
    assert np.all(err <= TOL_FP64), err
    assert np.any(err > 0.0)


def test_fast_gradient_with_venkatakrishnan_departs_from_the_bar(gpu):
    """With the Venkatakrishnan limiter, differences at the rounding level grow: the
    reference's additive TINY = 1e-12 in the limiter's denominator
    (RFLU_ModLimiters.F90:505-518) makes phi change by O(1e-4) for rounding-level changes
    of Delta_2 on faces where Delta_2 is below ~1e-12.  The moment form therefore leaves
    the 1e-12 bar within 10 steps (so the tuner offers it only where results may differ
    from the reference's), while the bit-identical path stays on the oracle."""
    m = cases.tet_box(n=16, jitter=0.08, seed=2)
    num = Numerics(discr=3, order=2, reconst=11, rk=4, cfl=0.5, dt_imposed=1e-4)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    print(f"case: fast gradient + Venkatakrishnan, cells={m.n_cells}", flush=True)
    outs = {}
    for fast, steps in ((0, 1), (1, 1), (0, 10), (1, 10)):
        g = GpuRegion(m, num)
        try:
            g.set_option(GpuRegion.OPT_GRAD_FAST, fast)
            g.upload(cv0)
            g.time_steps(steps)
            outs[fast, steps] = g.download()[0]
        finally:
            g.close()
    e1 = rel_l2(outs[1, 1], outs[0, 1])
    e10 = rel_l2(outs[1, 10], outs[0, 10])
    print("one step:", e1, "ten steps:", e10, flush=True)
    assert np.all(e1 <= 1e-14)
    assert e10.max() > 1e3 * e1.max()
