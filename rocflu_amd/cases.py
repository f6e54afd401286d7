"""Initial states of the benchmark cases (the role of the reference's rfluinit
hard-coded cases, utilities/init/RFLU_InitFlowHardCode.F90:862-920).

States are built in primitive form and converted with the dependent-variable
gamma' (libflu/RFLU_SetGasVars.F90:172-178), so that the pressure the solver
recovers is the one given here.  All return cv with shape (nCells, 5).
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh, box, HEX, TET6, BC_OUTFLOW, BC_SLIPWALL, BCOPT_SUBSONIC, BCOPT_SUPERSONIC
from .solver import Numerics


def prim_to_cons(rho, u, v, w, p, gamma_p):
    rho = np.asarray(rho, dtype=np.float64)
    cv = np.empty(rho.shape + (5,))
    cv[..., 0] = rho
    cv[..., 1] = rho * u
    cv[..., 2] = rho * v
    cv[..., 3] = rho * w
    cv[..., 4] = p / (gamma_p - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return cv


def sod(mesh: Mesh, num: Numerics, x_split: float, left=(1.0, 0.0, 1.0),
        right=(0.125, 0.0, 0.1)) -> np.ndarray:
    """Shock tube along x: (rho, u, p) left of x_split and right of it
    (the rectshktb case, RFLU_InitFlowHardCode.F90:862-920)."""
    x = mesh.cofg[:, 0]
    gp = num.gas_prime()[2]
    L = x < x_split
    rho = np.where(L, left[0], right[0])
    u = np.where(L, left[1], right[1])
    p = np.where(L, left[2], right[2])
    z = np.zeros_like(rho)
    return prim_to_cons(rho, u, z, z, p, gp)


def normal_shock(mach: float, gamma: float, rho1: float = 1.0, p1: float = 1.0):
    """Post-shock (rho2, u2, p2) of a shock moving at Mach `mach` into gas at rest
    (Rankine-Hugoniot)."""
    g = gamma
    m2 = mach * mach
    p2 = p1 * (2.0 * g * m2 - (g - 1.0)) / (g + 1.0)
    rho2 = rho1 * (g + 1.0) * m2 / ((g - 1.0) * m2 + 2.0)
    a1 = np.sqrt(g * p1 / rho1)
    u2 = mach * a1 * (1.0 - rho1 / rho2)
    return rho2, u2, p2


def planar_shock(mesh: Mesh, num: Numerics, x_shock: float, mach: float = 2.0,
                 rho1: float = 1.0, p1: float = 1.0, noise: float = 0.01,
                 seed: int = 0) -> np.ndarray:
    """Mach-`mach` planar shock at x = x_shock moving into quiescent gas, with a seeded
    relative density perturbation of amplitude `noise` everywhere."""
    gp = num.gas_prime()[2]
    rho2, u2, p2 = normal_shock(mach, gp, rho1, p1)
    x = mesh.cofg[:mesh.n_cells, 0]
    behind = x < x_shock
    rng = np.random.default_rng(seed)
    pert = 1.0 + noise * (2.0 * rng.random(mesh.n_cells) - 1.0)
    rho = np.where(behind, rho2, rho1) * pert
    u = np.where(behind, u2, 0.0)
    p = np.where(behind, p2, p1)
    z = np.zeros_like(rho)
    return prim_to_cons(rho, u, z, z, p, gp)


def blast(mesh: Mesh, num: Numerics, center, radius: float, p_in: float = 100.0,
          p_out: float = 1.0, rho: float = 1.0) -> np.ndarray:
    """Spherical blast: p_in inside `radius` of `center`, p_out outside, uniform rho."""
    gp = num.gas_prime()[2]
    d = mesh.cofg - np.asarray(center)[None, :]
    r = np.sqrt((d * d).sum(axis=1))
    p = np.where(r < radius, p_in, p_out)
    rr = np.full(mesh.n_cells, rho)
    z = np.zeros_like(rr)
    return prim_to_cons(rr, z, z, z, p, gp)


# ---- the benchmark configurations -----------------------------------------------

def sod_box(nx=64, ny=16, nz=16, jitter=0.1, seed=1, subsonic=False, stencils=False):
    """64x16x16 unit-spacing hex box with slip walls and outflow ends."""
    opt = BCOPT_SUBSONIC if subsonic else BCOPT_SUPERSONIC
    return box(HEX, nx, ny, nz, lengths=(float(nx), float(ny), float(nz)), jitter=jitter,
               seed=seed,
               bc_type=(BC_OUTFLOW, BC_OUTFLOW, BC_SLIPWALL, BC_SLIPWALL, BC_SLIPWALL, BC_SLIPWALL),
               bc_opt=(opt, opt, 0, 0, 0, 0), bc_pout=(1.0, 0.1, 0, 0, 0, 0), stencils=stencils)


def tet_box(n=80, jitter=0.08, seed=2, stencils=True, bc=BC_SLIPWALL):
    """n^3 unit-spaced box [0,n]^3 split into 6 Kuhn tetrahedra per box cell (6 n^3 cells).

    Unit spacing keeps the Venkatakrishnan eps^2 = vol far above the limiter's additive
    TINY = 1e-12, whose denominator d2*(...) + TINY has a pole for d2 < 0
    (modflu/RFLU_ModLimiters.F90:505-518) that a unit cube at n = 80 reaches."""
    return box(TET6, n, n, n, lengths=(float(n), float(n), float(n)), jitter=jitter, seed=seed,
               bc_type=(bc,) * 6, bc_opt=(0,) * 6, bc_pout=(0.0,) * 6, stencils=stencils)


def ensemble(n_cases=64, min_cells=20_000, max_cells=500_000, mach=(1.2, 3.0),
             gamma=(1.3, 1.67), seed=5, discr=3, order=2, reconst=11, rk=4, precision=1,
             cfl=0.5, dt_imposed=1e-4):
    """Independent shock-tube and blast cases for a batch: case k is a Kuhn-tet box
    whose size is drawn so that its cell count is uniform in [min_cells, max_cells],
    with its own shock Mach number and gamma drawn uniformly from the given ranges.
    Even cases are planar shocks (x = 0.3 L), odd cases spherical blasts at the box
    centre of radius 0.1 L with the pressure ratio of a normal shock at that Mach
    number.  Returns [(mesh, numerics, cv0)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_cases):
        ncell = rng.uniform(min_cells, max_cells)
        n = max(2, int(round((ncell / 6.0) ** (1.0 / 3.0))))
        ms = float(rng.uniform(*mach))
        g = float(rng.uniform(*gamma))
        m = tet_box(n=n, jitter=0.08, seed=100 + k)
        num = Numerics(discr=discr, order=order, reconst=reconst, rk=rk, precision=precision,
                       cfl=cfl, dt_imposed=dt_imposed, gamma=g)
        if k % 2 == 0:
            cv0 = planar_shock(m, num, 0.3 * n, mach=ms, noise=0.01, seed=200 + k)
        else:
            p2 = normal_shock(ms, num.gas_prime()[2])[2]
            cv0 = blast(m, num, (0.5 * n,) * 3, 0.1 * n, p_in=p2, p_out=1.0)
        out.append((m, num, cv0))
    return out


def adjoint_case(n=55, discr=3, p_in=100.0, stream=0.3, noise=0.01, seed=13, jitter=0.08,
                 order=1, reconst=11):
    """Spherical blast (p = p_in, rho = 1 inside radius 0.1 L; p = 1 outside) in an
    n^3 x 6-tet box with a velocity/temperature inflow at x- carrying a slow stream
    (u = stream, T of rho = 1, p = 1), subsonic outflow (p = 1) at x+ and slip walls
    elsewhere; SSP-RK3, first order or second order (`reconst`: 11 Venkatakrishnan or 0;
    inflow and outflow patches second order, walls first order).  Seeded relative noise
    of amplitude `noise` on density, pressure and velocity keeps states off the kinks of
    the flux functions (ties in min/max, zero speeds), where derivatives have no value.
    Returns (mesh, numerics, cv0); patch 2 (y-) is the wall of the objective."""
    from .mesh import BC_INFLOW_VELTEMP
    m = box(TET6, n, n, n, lengths=(float(n),) * 3, jitter=jitter, seed=seed,
            bc_type=(BC_INFLOW_VELTEMP, BC_OUTFLOW) + (BC_SLIPWALL,) * 4,
            bc_opt=(BCOPT_SUBSONIC, BCOPT_SUBSONIC, 0, 0, 0, 0), bc_pout=(0.0, 1.0, 0, 0, 0, 0),
            stencils=order == 2)
    base = Numerics()
    rg = base.gas_prime()[1]
    vals = [[stream, 0.0, 0.0, 1.0 / rg, 1.0]] + [[0.0] * 5] * 5
    num = Numerics(discr=discr, order=order, reconst=reconst if order == 2 else 0, rk=4, cfl=0.5,
                   dt_imposed=1e-3, patch_vals=vals,
                   patch_order=[2, 2, 1, 1, 1, 1] if order == 2 else None)
    gp = num.gas_prime()[2]
    rng = np.random.default_rng(seed)
    d = m.cofg - 0.5 * n
    r = np.sqrt((d * d).sum(axis=1))
    p = np.where(r < 0.1 * n, p_in, 1.0) * (1.0 + noise * rng.standard_normal(m.n_cells))
    rho = 1.0 + noise * rng.standard_normal(m.n_cells)
    vel = np.array([stream, 0.0, 0.0]) + noise * rng.standard_normal((m.n_cells, 3))
    return m, num, prim_to_cons(rho, vel[:, 0], vel[:, 1], vel[:, 2], p, gp)
