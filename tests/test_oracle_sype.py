"""Symmetry and periodic patches as virtual cells (rocflu_amd/csrc/sype.cpp, restating
the reference's RFLU_SYPE_AddVirtualCells, modflu/RFLU_ModSymmetryPeriodic.F90:148-1421,
and the exchange + transforms of libfloflu/RungeKuttaMP.F90:229-236, 440-459 and
modflu/RFLU_ModRelatedPatches.F90:380-435, 673-735), on the oracle.

* Structure: every actual cell of a fully periodic box keeps all 26 vertex neighbours
  and no boundary face; virtual cells are their sources moved by whole periods, or
  mirrored (symmetry), with the same volumes; the symmetry box's transforms are the
  eight products of the three reflections.
* Mathematics: a box periodic in x is the middle copy of a domain made of nine copies
  of it, and a box with symmetry patches at both x ends the middle copy of the domain
  mirrored about them, until the far ends' influence arrives: the oracle's states agree
  to rounding (HLLC order 2, with and without WENO-XYZ).  Turning the symmetry copies
  by the identity instead of the mirror breaks the agreement at O(1e-4).
* The reference's goldens tests/goldens/box/gas-jump/{periodic,symmetry}/np1 (8x8x8
  box, all six patches periodic or symmetric, run to 105 us): mass and energy to the
  golden's totals, and the x-column means within a band.  The goldens themselves are
  not invariant in y and z (the case is: its initial state depends on x only) and the
  periodic one is not antisymmetric in x as it should be; their defects are of the
  size of the band (measured below), while this construction keeps both symmetries to
  1e-12.  Parity with the golden is therefore pinned to that band only.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as orc
from oracle.oracle import Oracle
from rocflu_amd import mesh as rm
from rocflu_amd.solver import Numerics

HERE = os.path.dirname(os.path.abspath(__file__))
S = rm.BC_SLIPWALL
L = 0.008


@pytest.fixture(autouse=True, scope="module")
def threads():
    orc.set_threads(8)
    yield
    orc.set_threads(1)


def test_periodic_box_structure():
    g = rm.box_grid(rm.HEX, 8, 8, 8, lengths=(L,) * 3)
    m = g.build_sype([110] * 6, [0] * 6, [0.0] * 6, [2, 1, 4, 3, 6, 5], [1, 1, 2, 2, 3, 3])
    N = m.n_cells
    assert N == 512 and m.n_virtual > 0 and list(m.a["patch_map"]) == [-1]
    assert np.array_equal(np.diff(m.c2cs_ptr)[:N], np.full(N, 26))
    assert not np.isin(m.bf2c, np.arange(N)).any()
    cnt = np.bincount(m.f2c.reshape(-1), minlength=N + m.n_virtual)
    assert (cnt[:N] == 6).all()
    src = m.a["virt_src"]
    shift = (m.cofg[N:] - m.cofg[src]) / L
    assert np.abs(shift - np.round(shift)).max() <= 1e-12
    assert (np.abs(np.round(shift)).sum(1) >= 1).all()
    assert np.abs(m.vol[N:] - m.vol[src]).max() <= 1e-14 * m.vol.max()
    assert np.array_equal(m.a["xforms"].reshape(-1, 9), [np.eye(3).reshape(9)])


def test_symmetry_box_structure():
    g = rm.box_grid(rm.TET6, 4, 4, 4, lengths=(L,) * 3, jitter=0.1, seed=3)
    m = g.build_sype([100] * 6, [0] * 6, [0.0] * 6, [0] * 6, [0] * 6)
    N = m.n_cells
    assert not np.isin(m.bf2c, np.arange(N)).any()
    mats = m.a["xforms"].reshape(-1, 3, 3)
    diag = {tuple(np.diag(x)) for x in mats}
    assert len(mats) == 8 and diag == {(a, b, c) for a in (1, -1) for b in (1, -1) for c in (1, -1)}
    src, xf = m.a["virt_src"], m.a["virt_xf"]
    # a mirrored copy: each coordinate reflected about 0 or L an odd number of times or not
    for k in range(len(mats)):
        sel = xf == k
        if not sel.any():        # the identity: no copy of a symmetry box keeps it
            continue
        for d in range(3):
            a, b = m.cofg[N:][sel, d], m.cofg[src[sel], d]
            if mats[k][d, d] > 0:
                assert np.abs(a - b).max() <= 1e-12
            else:
                assert np.minimum(np.abs(a + b), np.abs(a + b - 2 * L)).max() <= 1e-12
    assert np.abs(m.vol[N:] - m.vol[src]).max() <= 1e-14 * m.vol.max()


def numerics(reconst):
    # a fixed step (DTMINLIMIT above the CFL step) keeps both runs on the same dt
    return Numerics(discr=3, order=2, reconst=reconst, rk=2, cfl=0.01, dt_imposed=1e-9,
                    dt_min_limit=1.5e-7, gamma=1.4, cp=1004.64)


def jump_state(x, num):
    gp = num.gas_prime()[2]
    left = x < 0.0043
    rho = np.where(left, 1.2, 0.97)
    p = np.where(left, 101325.0, 82700.0)
    cv = np.zeros((len(x), 5))
    cv[:, 0] = rho
    cv[:, 4] = rho * (p / (rho * (gp - 1.0)) + 0.0)
    return cv


def equivalent(kind, cell_kind, reconst, mirror_identity=False, ntile=9, t_end=9e-6):
    num = numerics(reconst)
    g = rm.box_grid(cell_kind, 8, 4, 4, lengths=(L, 0.004, 0.004))
    if kind == "periodic":
        m = g.build_sype([110, 110, S, S, S, S], [0] * 6, [0.0] * 6, [2, 1, 0, 0, 0, 0],
                         [1, 1, 0, 0, 0, 0])
    else:
        m = g.build_sype([100, 100, S, S, S, S], [0] * 6, [0.0] * 6, [0] * 6, [0] * 6)
        if mirror_identity:
            a = dict(m.a)
            a["xforms"] = np.tile(np.eye(3).reshape(1, 9), (len(a["xforms"]), 1))
            m = rm.Mesh(**{**m.__dict__, "a": a})
    big = rm.box(cell_kind, 8 * ntile, 4, 4, lengths=(L * ntile, 0.004, 0.004),
                 origin=(-L * (ntile // 2), 0.0, 0.0))
    xb = big.cofg[:, 0]
    if kind == "periodic":
        xr = np.mod(xb, L)
    else:
        xr = np.mod(xb, 2 * L)
        xr = np.where(xr > L, 2 * L - xr, xr)
    cvp, _, st1 = Oracle(m, num).run_until(jump_state(m.cofg[:m.n_cells, 0], num), t_end)
    cvb, _, st2 = Oracle(big, num).run_until(jump_state(xr, num), t_end)
    assert st1["rc"] == 0 and st2["rc"] == 0 and st1["steps"] == st2["steps"] == 60
    key = {tuple(np.round(c / 1e-5).astype(np.int64)): i for i, c in enumerate(big.cofg)}
    sel = np.array([key[tuple(np.round(c / 1e-5).astype(np.int64))] for c in m.cofg[:m.n_cells]])
    ref = cvb[sel]
    scale = np.abs(cvb).max(0)
    scale[1:4] = scale[1]          # momentum on the x-momentum scale
    return np.abs(cvp - ref).max(0) / scale


# hexahedra only: a hexahedron's 26 vertex neighbours all join its stencil, while a
# tetrahedron's stencil keeps the 40 best-ranked of more, ties broken in cell-number
# order, and the numbering of virtual cells is not that of the copies (the reference's
# rule, RFLU_BuildC2CStencilBasic, does the same)
@pytest.mark.parametrize("cell_kind,reconst", [(rm.HEX, 2), (rm.HEX, 0)])
def test_periodic_box_is_the_middle_of_its_copies(cell_kind, reconst):
    err = equivalent("periodic", cell_kind, reconst)
    assert err.max() <= 1e-12, err


@pytest.mark.parametrize("reconst", [2, 0])
def test_symmetry_box_is_the_middle_of_its_mirror_images(reconst):
    err = equivalent("symmetry", rm.HEX, reconst)
    assert err.max() <= 1e-12, err


def test_symmetry_needs_the_mirror():
    err = equivalent("symmetry", rm.HEX, 2, mirror_identity=True)
    assert err.max() > 1e-5, err


def gas_jump(kind):
    g = rm.box_grid(rm.HEX, 8, 8, 8, lengths=(L,) * 3)
    if kind == "periodic":
        m = g.build_sype([110] * 6, [0] * 6, [0.0] * 6, [2, 1, 4, 3, 6, 5], [1, 1, 2, 2, 3, 3])
    else:
        m = g.build_sype([100] * 6, [0] * 6, [0.0] * 6, [0] * 6, [0] * 6)
    # the template's numerics (tests/templates/rectshktb.inp) with the case's jump
    num = Numerics(discr=3, order=2, reconst=2, rk=2, cfl=0.4, dt_imposed=1e-9,
                   dt_min_limit=2e-10, gamma=1.4, cp=1004.64, creconst=1)
    return m, num, jump_state(m.cofg[:m.n_cells, 0], num)


# measured: column-mean difference / field scale (rho, rho u on the momentum scale, rho E)
BANDS = {"periodic": (6.6e-5, 1.8e-2, 9.7e-6), "symmetry": (2.7e-4, 2.0e-3, 3.8e-4)}


@pytest.mark.parametrize("kind", ["periodic", "symmetry"])
def test_reference_golden_gas_jump(kind):
    with open(os.path.join(HERE, "golden", f"gas_jump_{kind}_np1.json")) as f:
        gold = json.load(f)
    m, num, cv0 = gas_jump(kind)
    cv, _, st = Oracle(m, num).run_until(cv0, 1.05e-4)
    assert st["rc"] == 0 and abs(st["time"] - gold["time"]) <= 1e-9 * gold["time"]
    vol = m.vol[:m.n_cells]
    for k, name in ((0, "gas_rho"), (4, "gas_rhoE")):
        tot = float((cv[:, k] * vol).sum())
        assert abs(tot - gold["totals"][name]) <= 1e-12 * abs(gold["totals"][name])
        assert abs(tot - float((cv0[:, k] * vol).sum())) <= 1e-13 * abs(tot)
    q = np.array([gold["gas"]["0"][f"q{k}"] for k in range(1, 6)]).T      # x fastest
    ix = np.round(m.cofg[:m.n_cells, 0] / 0.001 - 0.5).astype(int)
    gx = np.arange(512) % 8
    scale = np.abs(q).max(0)
    scale[1:4] = scale[1:4].max()
    ours = np.array([[cv[ix == i, f].mean() for f in range(5)] for i in range(8)])
    theirs = np.array([[q[gx == i, f].mean() for f in range(5)] for i in range(8)])
    diff = np.abs(ours - theirs).max(0) / scale
    print(kind, "column-mean difference / scale", diff)
    band = BANDS[kind]
    assert diff[0] <= 1.5 * band[0] and diff[1] <= 1.5 * band[1] and diff[4] <= 1.5 * band[2]
    assert diff[2] <= 1e-12 and diff[3] <= 1e-12      # no mean transverse momentum
    # this construction: invariant in y and z
    spread = max(np.ptp(cv[ix == i], axis=0).max() / np.abs(cv).max() for i in range(8))
    assert spread <= 1e-12
    # the golden: not invariant in y and z, by about the band in density
    gspread = max(np.ptp(q[gx == i, 0]) for i in range(8)) / scale[0]
    print(kind, "golden y/z spread of rho / scale", gspread)
    assert gspread >= 1e-5
