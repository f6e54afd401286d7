"""Fluid dumps of device regions (rflu_gpu_write_dump / rflu_gpu_read_dump): a run
restarted from a dump continues with the bits of the run that was not interrupted
(the reference's restart contract, modflu/RFLU_ModRestartData.F90:693-872), and a
case of a batch dumps and restarts alone."""
import numpy as np
import pytest

from oracle import dump as odump
from rocflu_amd import cases, restart
from rocflu_amd.solver import GpuBatch, GpuRegion, Numerics

pytestmark = pytest.mark.gpu


def tet_case():
    m = cases.tet_box(n=8, jitter=0.08, seed=3)
    num = Numerics(discr=3, order=2, reconst=11, rk=4, cfl=0.5, dt_imposed=1e-3)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    return m, num, cv0


@pytest.mark.parametrize("fused", [True, False])
def test_restart_continues_bit_identically(gpu, tmp_path, fused):
    m, num, cv0 = tet_case()
    path = restart.dump_name(str(tmp_path) + "/", 6, 0)
    (tmp_path / "restart_data").mkdir()
    a = GpuRegion(m, num)
    b = GpuRegion(m, num)
    try:
        a.upload(cv0)
        for _ in range(6):
            if fused:
                a.time_steps(1)
            else:
                a.runge_kutta()
        t6 = a.stats()["time"] if fused else a.time
        a.write_dump(path, 6, t6, 0)
        cv6, _ = a.download()
        # the file holds the state in the caller's cell order, bit for bit
        d = odump.decode(open(path, "rb").read())
        assert d["step"] == 6 and d["time"] == t6 and d["form"] == 1
        assert d["cv"].tobytes() == cv6.tobytes()
        step, t, region = b.read_dump(path)
        assert (step, t, region) == (6, t6, 0)
        assert b.stats()["time"] == t6
        cvb, dvb = b.download()
        assert cvb.tobytes() == cv6.tobytes()
        assert np.array_equal(dvb, a.download()[1])
        for _ in range(6):
            if fused:
                a.time_steps(1)
                b.time_steps(1)
            else:
                a.runge_kutta()
                b.runge_kutta()
        ca, cb = a.download()[0], b.download()[0]
        print("restart max |diff|", np.abs(ca - cb).max(), flush=True)
        assert ca.tobytes() == cb.tobytes()
        if fused:
            assert a.stats()["time"] == b.stats()["time"]
        else:
            assert a.time == b.time
    finally:
        a.close()
        b.close()


def test_batch_case_dump_and_restart(gpu, tmp_path):
    ens = cases.ensemble(n_cases=3, min_cells=700, max_cells=3000, seed=11, dt_imposed=1e-3)
    b = GpuBatch([(m, num) for m, num, _ in ens])
    try:
        b.upload([cv0 for _, _, cv0 in ens])
        b.time_steps(5, graph=True)
        outs, _ = b.download()
        times = [b.case_stats(i)["time"] for i in range(3)]
        p = str(tmp_path / "case1.dat")
        b.write_dump(p, 5, times[1], 1, icase=1)
        cv, step, t, region = restart.read_dump(p, ens[1][0].n_cells)
        assert cv.tobytes() == outs[1].tobytes() and (step, t, region) == (5, times[1], 1)
        # case 1 alone, from the dump, against the batch's own continuation
        g = GpuRegion(ens[1][0], ens[1][1])
        try:
            g.read_dump(p)
            g.time_steps(4, graph=True)
            single = g.download()[0]
        finally:
            g.close()
        b.time_steps(4, graph=True)
        outs2, _ = b.download()
        assert single.tobytes() == outs2[1].tobytes()
        # reading case 0's old dump back into the batch rewinds case 0 only
        p0 = str(tmp_path / "case0.dat")
        restart.write_dump(p0, outs[0], 5, times[0], 0)
        b.read_dump(p0, icase=0)
        outs3, _ = b.download()
        assert outs3[0].tobytes() == outs[0].tobytes()
        assert outs3[1].tobytes() == outs2[1].tobytes() and outs3[2].tobytes() == outs2[2].tobytes()
        assert b.case_stats(0)["time"] == times[0]
    finally:
        b.close()
