"""Partition of a region grid into regions coupled through virtual cells (host side):
every region holds the grid's faces of its actual cells with their orientation and
geometry, the stencils of its actual and first-layer virtual cells in grid member
order, the per-cell face order of the grid, and borders whose send and receive lists
pair up cell for cell (modflu/RFLU_ModMPI.F90:332-390 copies send(i) to recv(i))."""
import numpy as np
import pytest

from rocflu_amd import mesh as rmesh
from rocflu_amd import regions as rg


def grid(n=(9, 4, 4), order=2):
    return rmesh.box(rmesh.TET6, *n, lengths=tuple(float(k) for k in n), jitter=0.08, seed=5,
                     bc_type=(rmesh.BC_INFLOW_VELTEMP, rmesh.BC_OUTFLOW) + (rmesh.BC_SLIPWALL,) * 4,
                     bc_opt=(1, 1, 0, 0, 0, 0), bc_pout=(0.0, 1.0, 0, 0, 0, 0),
                     stencils=order == 2)


def slabs(m, k):
    x = m.cofg[:, 0]
    return np.minimum((k * (x - x.min()) / (np.ptp(x) + 1e-12)).astype(int), k - 1)


def blocks(m, seed=1):
    """An irregular map: cells by the sign pattern of two tilted planes (4 regions)."""
    c = m.cofg - m.cofg.mean(axis=0)
    return ((c[:, 0] + 0.4 * c[:, 1] > 0).astype(int) + 2 * (c[:, 2] - 0.3 * c[:, 0] > 0))


@pytest.mark.parametrize("order,kind,blend", [(1, "slabs", True), (2, "slabs", True),
                                              (2, "blocks", True), (2, "blocks", False)])
def test_partition_invariants(order, kind, blend):
    m = grid(order=order)
    part = slabs(m, 3) if kind == "slabs" else blocks(m)
    parts = rg.partition(m, part, order=order, blend=blend)
    K = len(parts)
    f2c = m.f2c.reshape(-1, 2)
    # every actual cell in exactly one region, in grid order
    owned = np.concatenate([p.cells[:p.n_act] for p in parts])
    assert np.array_equal(np.sort(owned), np.arange(m.n_cells))
    for r, p in enumerate(parts):
        assert np.all(part[p.cells[:p.n_act]] == r)
        assert np.all(np.diff(p.cells[:p.n_act]) > 0)
        assert np.all(part[p.cells[p.n_act:]] != r)
        a, cells = p.mesh.a, p.cells
        # faces: the grid's faces with an actual cell, in grid order and orientation
        lf = cells[a["f2c"]]
        want = np.flatnonzero((part[f2c[:, 0]] == r) | (part[f2c[:, 1]] == r))
        gid = np.array([np.flatnonzero((f2c[:, 0] == u) & (f2c[:, 1] == v))[0] for u, v in lf])
        assert np.array_equal(gid, want)
        assert np.array_equal(a["fn"], m.fn.reshape(-1, 4)[want])
        assert np.array_equal(a["fc"], m.fc.reshape(-1, 3)[want])
        # boundary faces: those of actual cells, patch by patch
        bp = m.bpatch_ptr
        for q in range(m.n_patches):
            gb = np.arange(bp[q], bp[q + 1])
            gb = gb[part[m.bf2c[gb]] == r]
            lb = np.arange(a["bpatch_ptr"][q], a["bpatch_ptr"][q + 1])
            assert np.array_equal(cells[a["bf2c"][lb]], m.bf2c[gb])
            assert np.array_equal(a["bfn"][lb], m.bfn.reshape(-1, 4)[gb])
        # per-cell face order of actual cells = the grid's, face for face
        for i in range(0, p.n_act, 7):
            g = cells[i]
            gl = m.c2f_idx[m.c2f_ptr[g]:m.c2f_ptr[g + 1]]
            ll = a["c2f_idx"][a["c2f_ptr"][i]:a["c2f_ptr"][i + 1]]
            assert len(gl) == len(ll)
            Fg, Fl = m.n_faces, p.mesh.n_faces
            for s, t in zip(gl, ll):
                if abs(s) <= Fg:
                    assert abs(t) <= Fl and np.sign(s) == np.sign(t)
                    assert want[abs(t) - 1] == abs(s) - 1
                else:
                    assert t > Fl
        if order == 2:
            # stencils of actual and first-layer cells in grid member order
            sp, si = m.c2cs_ptr, m.c2cs_idx
            lp, li = a["c2cs_ptr"], a["c2cs_idx"]
            n_full = 0
            for i, g in enumerate(cells):
                loc = cells[li[lp[i]:lp[i + 1]]]
                if lp[i + 1] > lp[i]:
                    assert np.array_equal(loc, si[sp[g]:sp[g + 1]])
                    n_full += 1
            # every actual cell and every face neighbour of one has its stencil
            assert n_full >= p.n_act
        else:
            assert p.mesh.n_virtual == len(np.setdiff1d(np.unique(lf), cells[:p.n_act]))
    # borders pair up: send list of r to s = receive list of s from r, cell for cell
    recv_count = [np.zeros(len(p.cells), dtype=int) for p in parts]
    for r, p in enumerate(parts):
        for b in p.borders:
            q = parts[b.nbr]
            e = q.borders[b.nbr_border]
            assert e.nbr == r
            assert np.array_equal(p.cells[b.send], q.cells[e.recv])
            assert np.all(b.send < p.n_act) and np.all(e.recv >= q.n_act)
            recv_count[b.nbr][e.recv] += 1
    for p, cnt in zip(parts, recv_count):
        assert np.all(cnt[:p.n_act] == 0) and np.all(cnt[p.n_act:] == 1)
    assert K == int(part.max()) + 1


def test_partition_rejects_bad_maps():
    m = grid()
    with pytest.raises(ValueError):
        rg.partition(m, np.zeros(m.n_cells - 1, dtype=int))
    part = np.zeros(m.n_cells, dtype=int)
    part[:10] = 2                              # region 1 has no cells
    with pytest.raises(ValueError):
        rg.partition(m, part)


def test_virtual_cell_arguments_are_checked_without_a_device():
    """Descriptor and border checks that report before any device call."""
    import ctypes

    from rocflu_amd import _native as nat
    from rocflu_amd.solver import GpuRegion, Numerics
    m = grid()
    parts = rg.partition(m, slabs(m, 2), order=2)
    sub = parts[0].mesh
    vals = [[0.4, 0.0, 0.0, 1.1 / Numerics().gas_prime()[1], 1.2]] + [[0.0] * 5] * 5

    def code(mesh, **kw):
        with pytest.raises(nat.RfluError) as e:
            GpuRegion(mesh, Numerics(order=2, reconst=11, patch_vals=vals, **kw))
        return e.value.code

    # partitioned regions take no constrained gradients (their wall lists end at the region)
    with pytest.raises(ValueError, match="CRECONSTC"):
        rg.check_parts(parts, Numerics(order=2, reconst=11, creconst=1, patch_vals=vals))
    sub_bad = rg.Mesh(**{**sub.__dict__, "n_virtual": -1})
    assert code(sub_bad) == nat.RFLU_ERR_ARG
    L = nat.lib()
    # a batch takes no virtual cells
    from rocflu_amd.solver import region_desc
    num = Numerics(order=2, reconst=11, patch_vals=vals)
    descs = (nat.RegionDesc * 2)()
    keep = [region_desc(sub, num, descs[0]), region_desc(sub, num, descs[1])]
    h = ctypes.c_int32(0)
    assert L.rflu_gpu_create_batch(2, descs, ctypes.byref(h)) == nat.RFLU_ERR_ARG
    del keep
    # unknown handles
    hs = np.asarray([777], dtype=np.int32)
    assert L.rflu_gpu_exchange(1, nat.ptr(hs, ctypes.c_int32), 1) == nat.RFLU_ERR_HANDLE
    z = np.zeros(2, dtype=np.int32)
    p = nat.ptr(z, ctypes.c_int32)
    assert L.rflu_gpu_set_borders(777, 1, p, p, p, p, p, p) == nat.RFLU_ERR_HANDLE
    if L.rflu_gpu_device_count() == 0:
        # a valid region with virtual cells stops only at the missing device
        assert code(sub) == nat.RFLU_ERR_HIP


def test_weno_needs_blend_partitions():
    """WENO-XYZ blends raw gradients of the virtual cells in actual cells' stencils
    before the gradient exchange: a partition built with blend=False gives those cells
    no stencils and is refused before any device call; blend=True is accepted, and a
    first-order partition is refused for a second-order scheme."""
    from rocflu_amd.solver import Numerics
    m = grid()
    weno = Numerics(order=2, reconst=2)
    with pytest.raises(ValueError, match="blend=True"):
        rg.check_parts(rg.partition(m, slabs(m, 2), order=2, blend=False), weno)
    rg.check_parts(rg.partition(m, slabs(m, 2), order=2, blend=True), weno)
    rg.check_parts(rg.partition(m, slabs(m, 2), order=2, blend=False),
                   Numerics(order=2, reconst=11))
    with pytest.raises(ValueError, match="order 1"):
        rg.check_parts(rg.partition(m, slabs(m, 2), order=1), Numerics(order=2, reconst=11))
    with pytest.raises(ValueError, match="blend=True"):
        rg.GpuRegions(rg.partition(m, slabs(m, 2), order=2, blend=False), weno)
