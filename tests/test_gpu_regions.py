"""Regions of one problem on one GPU, coupled through virtual cells and the device halo
exchange (rflu_gpu_set_borders / rflu_gpu_exchange; RungeKuttaMP.F90:229-236,440-459):
K regions stepped through the per-callee entry points give the single region's states
bit for bit (same per-cell face, stencil and limiter-neighbour orders; dt the minimum
over the regions), in FP64 and in the mixed mode; mass in/out summed over the regions
to 1e-12; the single region itself against the oracle to 1e-12."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from rocflu_amd import cases, mesh as rmesh
from rocflu_amd import regions as rg
from rocflu_amd.solver import GpuRegion, Numerics

pytestmark = pytest.mark.gpu


def setup(order, discr, reconst, prec=0, n=(12, 5, 5)):
    m = rmesh.box(rmesh.TET6, *n, lengths=tuple(float(k) for k in n), jitter=0.08, seed=5,
                  bc_type=(rmesh.BC_INFLOW_VELTEMP, rmesh.BC_OUTFLOW) + (rmesh.BC_SLIPWALL,) * 4,
                  bc_opt=(1, 1, 0, 0, 0, 0), bc_pout=(0.0, 1.0, 0, 0, 0, 0),
                  stencils=order == 2)
    rg_ = Numerics().gas_prime()[1]
    vals = [[0.4, 0.05, -0.02, 1.1 / rg_, 1.2]] + [[0.0] * 5] * 5
    num = Numerics(discr=discr, order=order, reconst=reconst if order == 2 else 0, rk=4,
                   cfl=0.5, dt_imposed=1e-3, patch_vals=vals, precision=prec,
                   patch_order=[2, 2, 1, 1, 1, 1] if order == 2 else None)
    cv0 = cases.blast(m, num, (0.35 * n[0], 0.5 * n[1], 0.45 * n[2]), 1.5, p_in=4.0)
    return m, num, cv0


def partition_map(m, kind, k=3):
    c = m.cofg - m.cofg.mean(axis=0)
    if kind == "slabs":
        x = m.cofg[:, 0]
        return np.minimum((k * (x - x.min()) / (np.ptp(x) + 1e-12)).astype(int), k - 1)
    return (c[:, 0] + 0.4 * c[:, 1] > 0).astype(int) + 2 * (c[:, 2] - 0.3 * c[:, 0] > 0)


@pytest.mark.parametrize("order,discr,reconst,prec,kind,blend", [
    (1, 1, 0, 0, "slabs", True), (1, 3, 0, 0, "blocks", True),
    (2, 3, 11, 0, "slabs", True), (2, 3, 2, 0, "blocks", True), (2, 4, 0, 0, "slabs", True),
    (2, 1, 11, 0, "blocks", True), (2, 3, 11, 1, "slabs", True), (2, 3, 2, 1, "blocks", True),
    (2, 3, 11, 0, "blocks", False), (2, 4, 0, 1, "slabs", False), (2, 3, 10, 0, "blocks", False),
    (2, 1, 10, 1, "slabs", True)])
def test_regions_match_the_single_region(gpu, order, discr, reconst, prec, kind, blend):
    """blend=False: one virtual layer without stencils (enough without WENO-XYZ)."""
    m, num, cv0 = setup(order, discr, reconst, prec)
    part = partition_map(m, kind)
    parts = rg.partition(m, part, order=order, blend=blend)
    nsteps = 12
    print(f"case: order={order} discr={discr} reconst={reconst} prec={prec} {kind} "
          f"regions={len(parts)} cells={m.n_cells} virtual={[p.mesh.n_virtual for p in parts]}",
          flush=True)
    g = GpuRegion(m, num)
    try:
        g.upload(cv0)
        dts1 = [g.runge_kutta()[0] for _ in range(nsteps)]
        cv1, dv1 = g.download()
        st1 = g.stats()
    finally:
        g.close()
    R = rg.GpuRegions(parts, num)
    try:
        R.upload(cv0)
        dtsK = [R.runge_kutta()[0] for _ in range(nsteps)]
        cvK, dvK = R.download()
        stK = R.stats()
    finally:
        R.close()
    diff = np.abs(cvK - cv1).max()
    print("dt equal:", dts1 == dtsK, "max |cv diff|:", diff,
          "mass in/out", st1["mass_in"], stK["mass_in"], st1["mass_out"], stK["mass_out"],
          flush=True)
    assert dts1 == dtsK
    assert np.array_equal(cvK, cv1)
    assert np.array_equal(dvK, dv1)
    for key in ("mass_in", "mass_out"):
        assert abs(stK[key] - st1[key]) <= 1e-12 * max(abs(st1[key]), 1e-30)
    assert st1["mass_in"] > 0.0
    if prec == 0 and kind == "slabs":
        cvo, _, _ = Oracle(m, num).run(cv0, nsteps)
        for k in range(5):
            err = np.linalg.norm(cvK[:, k] - cvo[:, k]) / np.linalg.norm(cvo[:, k])
            assert err <= 1e-12, (k, err)


def test_exchange_checks_its_arguments(gpu):
    m, num, cv0 = setup(2, 3, 11)
    parts = rg.partition(m, partition_map(m, "slabs"), order=2)
    R = rg.GpuRegions(parts, num)
    try:
        R.upload(cv0)
        L = R._lib
        import ctypes
        from rocflu_amd import _native as nat
        one = np.asarray(R._handles[:1], dtype=np.int32)
        # a border's neighbour missing from the exchange
        assert L.rflu_gpu_exchange(1, nat.ptr(one, ctypes.c_int32), rg.XCHG_STATE) != 0
        assert b"neighbour" in L.rflu_gpu_last_error()
        assert L.rflu_gpu_exchange(len(parts), nat.ptr(R._handles, ctypes.c_int32), 7) != 0
        # receive cells must be virtual cells
        g = R.regions[0]
        z = np.zeros(2, dtype=np.int32)
        nbr = np.asarray([R._handles[1]], dtype=np.int32)
        ptr = np.asarray([0, 1], dtype=np.int32)
        assert L.rflu_gpu_set_borders(g.handle, 1, nat.ptr(nbr, ctypes.c_int32),
                                      nat.ptr(z, ctypes.c_int32), nat.ptr(ptr, ctypes.c_int32),
                                      nat.ptr(z, ctypes.c_int32), nat.ptr(ptr, ctypes.c_int32),
                                      nat.ptr(z, ctypes.c_int32)) != 0
        assert b"virtual" in L.rflu_gpu_last_error()
        # the fused loop needs the exchanges between stages: refused
        assert L.rflu_gpu_time_steps(g.handle, 1, 0) != 0
        # the earlier borders still hold: a step still matches
        R.runge_kutta()
    finally:
        R.close()
