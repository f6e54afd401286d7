#!/usr/bin/env python3
"""Throughput of the explicit residual/update path on one MI355X.

Metric: cell-updates/s per RK stage (real cells x stages / wall time of the time-step
loop: dt + all stages), FP64 and mixed precision, with the mixed run's error against
the FP64 GPU run per conserved variable, and the dominant kernel's HBM roofline
fraction.  Headline workload (the largest single-GPU configuration): a 160^3 box split
into 6 Kuhn tetrahedra per box cell (24,576,000 tets), interior nodes jittered +-8%,
Mach-2 planar shock into quiescent gas with 1% seeded density noise, second-order HLLC
with WLSQ gradients and the Venkatakrishnan limiter, SSP-RK3 (three stages), slip
walls; FP64 through the bit-identical kernels.  Beside it: the same mesh in mixed
precision, the same run through the per-callee (drop-in) entry points, the 80^3 x 6
configuration (HLLC and AUSM+, FP64 and mixed), the 64-case batch, the adjoint and the
CPU oracle.  A step is one time step (dt + 3 stages) over the whole mesh; inputs are
resident in HBM before the timed region.  Kernel variants are chosen by the project's
autotuner before timing (bit-identical ones only for FP64).

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32

KERNELS = ["face_flux", "update", "gradient", "blend", "time_step", "bface_flux"]


def algorithmic_bytes(m, num, kernel: str, tuning: dict, compulsory: bool = False) -> float:
    """Bytes of one launch: every array element the kernel needs, once, at its stored
    width (double4 geometry counted at 32 B; re-reads of neighbours' data through L2/LDS
    are not counted).  compulsory=True leaves out stored stencil weights (a design
    choice that trades HBM bytes for arithmetic) and counts what forming them needs
    (centroids and R-factor moments) instead."""
    if compulsory and tuning.get("grad_weights"):
        tuning = dict(tuning, grad_weights=0)
    N, F, B = int(m.n_cells), int(m.n_faces), int(m.n_bfaces)
    S = int(m.c2cs_idx.size) if num.order == 2 else 0
    K = int(m.c2f_idx.size)               # cell-to-face entries (2F + B)
    fpf = 32 if num.precision == 3 else 40     # flux buffer bytes per face (FP32 mass/energy
                                               # and FP64 momenta in the FP32-flux split)
    gb = 4 if num.precision in (1, 2) else 8   # gradient width
    if kernel == "gradient":
        b = N * (40 + 15 * gb) + 4 * (N + 1)   # own state in, gradient out, stencil counts
        if tuning.get("grad_lds"):
            b += 2 * S                          # 16-bit LDS slot of every member
        else:
            b += 4 * S                          # member indices
        if tuning.get("grad_weights") == 1:
            b += 3 * gb * S                     # stored stencil weights
        else:
            b += N * (32 + 48)                  # centroids, R-factor moments
            if tuning.get("grad_weights") == 2:
                b += gb * S                     # stored inverse distances
        if num.reconst == 11:
            b += 8 * N + 2 * F * (24 + 2)       # eps^2; cofg - fc and slot per face side
        return float(b)
    if kernel == "face_flux":
        per_face = 8 + 32 + (32 if num.order == 2 else 0) + fpf
        per_cell = (40 + 32 + 15 * gb) if num.order == 2 else (40 + 24)
        return float(F * per_face + N * per_cell)
    if kernel == "update":
        # c2f, flux (each face value once), cv in/out, RK storage (SSP-RK3: cvOld once;
        # Wray: rhsSum written, read+written, read over 3 stages), vol, dv out, prim out
        rk_store = 40 if num.rk == 4 else 40 * 4 / 3
        b = 4 * (N + 1) + 4 * K + fpf * (F + B) + N * (40 + 40 + rk_store + 8 + 24)
        if num.order == 2:
            b += 40 * N
        return float(b)
    if kernel == "time_step":
        return float(4 * (N + 1) + 4 * K + 8 * (F + B) + 8 * N)
    return 0.0


def grad_variant(t: dict) -> str:
    v = "grad_t{}_lds{}_u{}_w{}".format(t.get("grad_tile", 128), t["grad_lds"],
                                        t["grad_unroll"], t["grad_weights"])
    return v + ("_fast" if t.get("grad_fast") else "")


def gradient_flops(m, num, tuning: dict) -> float:
    """Floating-point operations of one gradient launch (sqrt and divide counted as
    one): per stencil member 31 to form its weights (when not stored) and 35 to
    accumulate 5 variables; per cell 30 for the R-factor coefficients and applying the
    limiter; per interior face side 115 for the limiter's min/max and 5 Venkatakrishnan
    terms."""
    N, F = int(m.n_cells), int(m.n_faces)
    S = int(m.c2cs_idx.size)
    # weights stored: none formed; inverse distances stored: 24 of the 31 (no |d|^2,
    # square root or quotient)
    per_member = 35 + {1: 0, 2: 24}.get(tuning.get("grad_weights", 0), 31)
    if tuning.get("grad_fast"):
        per_member = 28              # moment form: |d|^2, reciprocal, 3 scalings, 15 FMAs
    fl = S * per_member + 30 * N
    if num.reconst == 11:
        fl += 2 * F * 115
    return float(fl)


def build_case(args, precision, discr=None):
    from rocflu_amd.solver import Numerics
    num = Numerics(discr=args.discr if discr is None else discr, order=2,
                   reconst=args.reconst, rk=4, cfl=0.5, dt_imposed=1e-4, dt_min_limit=0.0,
                   precision=precision)
    return num


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_topology() -> dict:
    """lscpu's model, sockets, cores, threads and caches (empty where it is missing)."""
    keep = ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
            "CPU max MHz", "L2 cache", "L3 cache", "NUMA node(s)", "Flags")
    out = {}
    try:
        p = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                           text=True, timeout=20)
        for line in p.stdout.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in keep:
                v = v.strip()
                if k == "Flags":   # the vector extensions only
                    v = " ".join(f for f in v.split() if f.startswith(("avx", "fma", "sse4")))
                out[k] = v
    except (OSError, subprocess.SubprocessError):
        pass
    return out


def host_threads():
    """Threads the host grants this process (OMP_NUM_THREADS, else the affinity mask)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(args):
    """The oracle (the C restatement of the reference loops) on a bounded sample of the
    same workload (a cpu_n^3 x 6-tet mesh, same scheme, cpu_steps steps), once with one
    thread and once with every thread the host grants (OpenMP over cells, gathers in
    place of the face scatters: the same bits).  value is the all-threads rate."""
    from rocflu_amd import cases
    from oracle import oracle as orc
    from oracle.oracle import Oracle
    n = args.cpu_n
    m = cases.tet_box(n=n, jitter=0.08, seed=2)
    num = build_case(args, 0)
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    o = Oracle(m, num)
    steps = args.cpu_steps
    rates, times = {}, {}
    finals = {}
    for th in (1, host_threads()):
        used = orc.set_threads(th)
        o.run(cv0, 1)  # warm-up
        t0 = time.perf_counter()
        finals[used], _, _ = o.run(cv0, steps)
        times[used] = time.perf_counter() - t0
        rates[used] = m.n_cells * 3 * steps / times[used]
    orc.set_threads(1)
    tmax = max(rates)
    same = all(np.array_equal(finals[1], f) for f in finals.values())
    return {"value": rates[tmax], "unit": "cell-updates/s", "cores": tmax, "kind": "port",
            "value_1thread": rates[1], "thread_speedup": rates[tmax] / rates[1],
            "threads_same_bits": bool(same), "cpu_model": cpu_model(),
            "lscpu": cpu_topology(),
            "sample": f"{n}^3x6 = {m.n_cells} tets, {steps} SSP-RK3 steps of the same scheme: "
                      f"{times[1]:.1f} s on 1 thread, {times[tmax]:.1f} s on {tmax} threads"}


def time_region(args, m, num, exact: bool, steps: int, err_steps: int, dropin: bool = False,
                fp32_records: bool = False):
    """One region of mesh m with numerics num: autotuned (exact: bit-identical variants
    only; fp32_records: the FP32 gradient tile records among the candidates), then `steps` fused steps timed three ways -- eager (the value), eager with HIP
    events around every launch (per-kernel milliseconds), and as a replayed graph -- and
    with dropin the same steps through the per-callee entry points (the reference
    host's call sequence); finally err_steps graph steps from the initial state for the
    accuracy comparison.  Every timed run starts from the same initial state."""
    from rocflu_amd import cases, autotune
    from rocflu_amd.solver import GpuRegion
    import torch  # device memory and streams are owned by the C-ABI library
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    g = GpuRegion(m, num)
    try:
        g.upload(cv0)
        tuning = autotune.tune(g, exact=exact, fp32_records=fp32_records)
        tuning = {k: v for k, v in tuning.items() if k != "timings_ms"}

        def timed(fn, n):
            fn(args.warmup)
            g.synchronize()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            fn(n)
            g.synchronize()
            torch.cuda.synchronize()
            return time.perf_counter() - t1

        g.upload(cv0)
        el = timed(lambda n: g.time_steps(n), steps)
        g.upload(cv0)
        g.set_profiling(True)
        el_prof = timed(lambda n: g.time_steps(n), steps)
        kt = g.kernel_times()
        g.set_profiling(False)
        g.upload(cv0)
        elg = timed(lambda n: g.time_steps(n, graph=True), steps)
        n_upd = m.n_cells * 3 * steps
        res = {"value": n_upd / el, "ms_per_step": 1e3 * el / steps, "graph_value": n_upd / elg,
               "graph_ms_per_step": 1e3 * elg / steps,
               "profiled_ms_per_step": 1e3 * el_prof / steps,
               "kernels": kt, "num": num, "tuning": tuning}
        if dropin:
            def callee(n):
                for _ in range(n):
                    g.runge_kutta()
                g.check_status()
            g.upload(cv0)
            eld = timed(callee, steps)
            res["dropin"] = {"value": n_upd / eld, "ms_per_step": 1e3 * eld / steps,
                             "vs_fused": el / eld}
        res["final"] = None
        if err_steps:
            g.upload(cv0)
            g.time_steps(err_steps, graph=True)
            res["final"] = g.download()[0]
        res["stats"] = g.stats()
    finally:
        g.close()
    print(f"# {num.discr}/{num.precision} on {m.n_cells}: {res['value']:.4g} cell-upd/s "
          f"({res['ms_per_step']:.3f} ms/step), graph {res['graph_value']:.4g}"
          + (f", per-callee {res['dropin']['ms_per_step']:.3f} ms/step" if dropin else "")
          + f"; tuning {tuning}; kernels {kt}", file=sys.stderr, flush=True)
    return res


def final_state(m, num, steps, fast=False, tuning=None):
    """State after `steps` fused steps from the headline initial state, with the kernel
    variants of `tuning` (default ones without; fast: the FP64 moment-form gradient,
    which only sums in another order)."""
    from rocflu_amd import cases, autotune
    from rocflu_amd.solver import GpuRegion
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    g = GpuRegion(m, num)
    try:
        if tuning:
            autotune.apply(g, tuning)
        if fast:
            g.set_option(GpuRegion.OPT_GRAD_FAST, 1)
        g.upload(cv0)
        g.time_steps(steps, graph=True)
        return g.download()[0]
    finally:
        g.close()


def one_stage_state(m, num, tuning=None):
    """State after the first RK stage from the headline initial state, through the
    per-callee entry points (RungeKuttaMP's order within a stage, the FP64 time step)."""
    from rocflu_amd import cases, autotune
    from rocflu_amd.solver import GpuRegion
    cv0 = cases.planar_shock(m, num, 0.3 * m.dims[0], mach=2.0, noise=0.01, seed=7)
    g = GpuRegion(m, num)
    try:
        if tuning:
            autotune.apply(g, tuning)
        g.upload(cv0)
        dt = g.time_step()
        g.rk_init(1)
        g.cell_gradients()
        g.convective_fluxes(1)
        g.rk_update(1, dt)
        g.set_dependent_vars()
        g.after_update(1)
        return g.download()[0]
    finally:
        g.close()


def mixed_accuracy(args, m, final64, tuning, keep=None):
    """The mixed run's error per conserved variable beside what sets it: after one stage
    and one step (the stage's own rounding), after err_steps without a modifier (no
    amplification),
    and after err_steps with the limiter next to an FP64 run whose gradient only sums in
    another order (the moment form) -- the limiter amplifies rounding-level differences
    of the gradients in the transverse momenta whatever causes them."""
    from rocflu_amd.solver import Numerics
    out = {}
    mk = lambda prec, reconst: Numerics(discr=args.discr, order=2, reconst=reconst, rk=4,
                                        cfl=0.5, dt_imposed=1e-4, dt_min_limit=0.0,
                                        precision=prec)
    # the mixed runs use the timed mixed run's kernel variants (its error is the one that
    # goes with its speed)
    a = one_stage_state(m, mk(1, args.reconst), tuning)
    b = one_stage_state(m, mk(0, args.reconst))
    if keep is not None:
        keep["one_stage64"] = b   # the FP64 side of the FP32-record line's one-stage error
    out["one_stage"] = rel_l2_per_variable(a, b)
    a = final_state(m, mk(1, args.reconst), 1, tuning=tuning)
    b = final_state(m, mk(0, args.reconst), 1)
    out["one_step"] = rel_l2_per_variable(a, b)
    r = final_state(m, mk(0, args.reconst), args.err_steps, fast=True)
    out["fp64_reorder_drift"] = rel_l2_per_variable(r, final64)
    a = final_state(m, mk(1, 0), args.err_steps, tuning=tuning)
    b = final_state(m, mk(0, 0), args.err_steps)
    out["no_modifier"] = rel_l2_per_variable(a, b)
    print(f"# mixed accuracy: {out}", file=sys.stderr, flush=True)
    return out


def rel_l2_per_variable(a, b):
    return [float(np.linalg.norm(a[:, k] - b[:, k]) / np.linalg.norm(b[:, k])) for k in range(5)]


def persistent_vs_launches(b, cvs, args, n_tot):
    """One time step of every case as one persistent launch (PERSISTENT: grid barriers
    between the phases) against the same step as separate launches, both replayed as a
    graph with the persistent step's kernel choices (256-cell gradient tiles with the
    moment form, 128-cell face tiles); the final states compared bit for bit."""
    from rocflu_amd import autotune
    from rocflu_amd.solver import GpuRegion
    autotune._set_tile(b, 256)
    b.set_option(GpuRegion.OPT_GRAD_LDS, 1)
    b.set_option(GpuRegion.OPT_GRAD_WEIGHTS, 0)
    b.set_option(GpuRegion.OPT_FACE_TILE, 128)
    b.set_option(GpuRegion.OPT_UPDATE_TILE, 1)
    b.set_option(GpuRegion.OPT_GRAD_FAST, 1)
    out, finals = {}, {}
    for pers in (0, 1):
        b.set_option(GpuRegion.OPT_PERSISTENT, pers)
        b.upload(cvs)
        b.time_steps(args.warmup, graph=True)
        b.synchronize()
        t1 = time.perf_counter()
        b.time_steps(args.batch_steps, graph=True)
        b.synchronize()
        el = time.perf_counter() - t1
        key = "persistent" if pers else "separate_launches"
        out[key] = {"ms_per_step": 1e3 * el / args.batch_steps,
                    "value": n_tot * 3 * args.batch_steps / el}
        finals[pers] = b.download()[0]
    b.set_option(GpuRegion.OPT_PERSISTENT, 0)
    out["same_bits"] = bool(all(np.array_equal(x, y) for x, y in zip(finals[0], finals[1])))
    print(f"# batch persistent: {out}", file=sys.stderr, flush=True)
    return out


def batch_leg(args):
    """The batched ensemble: independent shock-tube and blast cases of 20k-500k tets
    each, Mach 1.2-3.0, gamma 1.3-1.67, mixed precision HLLC at second order with
    the Venkatakrishnan limiter, SSP-RK3; every kernel launch covers all cases and one
    step of all of them is one captured graph.  Also timed: the same cases advanced
    one after the other as regions of their own (each with its own graph)."""
    from rocflu_amd import cases, autotune
    from rocflu_amd.solver import GpuBatch, GpuRegion
    t0 = time.time()
    ens = cases.ensemble(n_cases=args.batch_cases, seed=5)
    n_tot = sum(int(m.n_cells) for m, _, _ in ens)
    print(f"# batch: {len(ens)} cases, {n_tot} tets, built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    b = GpuBatch([(m, num) for m, num, _ in ens])
    cvs = [cv0 for _, _, cv0 in ens]
    b.upload(cvs)
    tuning = autotune.tune(b, exact=False)
    b.upload(cvs)
    b.time_steps(args.warmup, graph=True)
    b.synchronize()
    t1 = time.perf_counter()
    b.time_steps(args.batch_steps, graph=True)
    b.synchronize()
    el = time.perf_counter() - t1
    st = [b.case_stats(i) for i in range(len(ens))]
    pers = persistent_vs_launches(b, cvs, args, n_tot)
    b.close()
    # the same cases one region at a time
    t_seq = 0.0
    for m, num, cv0 in ens:
        g = GpuRegion(m, num)
        g.upload(cv0)
        autotune.apply(g, tuning)
        g.time_steps(args.warmup, graph=True)
        g.synchronize()
        t2 = time.perf_counter()
        g.time_steps(args.batch_steps, graph=True)
        g.synchronize()
        t_seq += time.perf_counter() - t2
        g.close()
    n_upd = n_tot * 3 * args.batch_steps
    out = {"value": n_upd / el, "unit": "cell-updates/s", "cases": len(ens), "cells": n_tot,
           "steps": args.batch_steps, "ms_per_step": 1e3 * el / args.batch_steps,
           "sequential_value": n_upd / t_seq, "speedup_vs_sequential": t_seq / el,
           "dtype": "f32/f64", "failed_cases": sum(1 for x in st if x["n_invalid"] or x["n_negative"]),
           "tuning": {k: v for k, v in tuning.items() if k != "timings_ms"},
           "persistent": pers,
           "workload": "64 Kuhn-tet boxes, cell counts U(20k, 500k), shock Mach U(1.2, 3.0), "
                       "gamma U(1.3, 1.67), planar shocks and blasts, HLLC order 2 + "
                       "Venkatakrishnan, SSP-RK3, mixed precision, slip walls"}
    print(f"# batch: {out['value']:.4g} cell-upd/s batched, {out['sequential_value']:.4g} one "
          f"region at a time", file=sys.stderr, flush=True)
    return out


def regions_leg(args):
    """Regions coupled through virtual cells on one device (--regions K): the n^3 x 6
    shock case split into K x-slabs, stepped through the per-callee entry points with the
    halo exchanges, against the single region through the same entry points (states
    compared bit for bit); the exchanges alone timed, with the bytes they move."""
    from rocflu_amd import cases, regions as rgs
    from rocflu_amd.solver import GpuRegion, n_stages
    n, K, steps = args.regions_n, args.regions, args.regions_steps
    m = cases.tet_box(n=n, jitter=0.08, seed=2)
    num = build_case(args, 0)
    cv0 = cases.planar_shock(m, num, 0.3 * n, mach=2.0, noise=0.01, seed=7)
    x = m.cofg[:, 0]
    part = np.minimum((K * (x - x.min()) / np.ptp(x)).astype(int), K - 1)
    t0 = time.time()
    # the WENO-XYZ blend reads virtual cells' raw gradients: two layers, else one
    parts = rgs.partition(m, part, order=2, blend=args.reconst == 2)
    print(f"# regions: {K} slabs of {m.n_cells} tets, virtual cells "
          f"{[p.mesh.n_virtual for p in parts]}, built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)

    def timed(stepper, sync):
        for _ in range(args.warmup):
            stepper()
        sync()
        t1 = time.perf_counter()
        for _ in range(steps):
            stepper()
        sync()
        return time.perf_counter() - t1

    g = GpuRegion(m, num)
    g.upload(cv0)
    el1 = timed(g.runge_kutta, g.synchronize)
    cv1 = g.download()[0]
    g.close()
    R = rgs.GpuRegions(parts, num)
    try:
        R.upload(cv0)
        sync = lambda: [r.synchronize() for r in R.regions]  # noqa: E731
        elK = timed(R.runge_kutta, sync)
        cvK = R.download()[0]
        reps = 20
        t1 = time.perf_counter()
        for _ in range(reps):
            R.exchange(rgs.XCHG_STATE)
        sync()
        el_state = (time.perf_counter() - t1) / reps
        t1 = time.perf_counter()
        for _ in range(reps):
            R.exchange(rgs.XCHG_GRAD)
        sync()
        el_grad = (time.perf_counter() - t1) / reps
    finally:
        R.close()
    nv = int(sum(p.mesh.n_virtual for p in parts))
    st = n_stages(num.rk)
    return {"workload": f"{n}^3 x 6 Kuhn tets ({m.n_cells} cells) in {K} x-slab regions, "
                        f"HLLC order 2 + Venkatakrishnan, SSP-RK3, FP64, per-callee entry points",
            "regions": K, "virtual_cells": nv, "steps": steps,
            "identical_to_single_region": bool(np.array_equal(cvK, cv1)),
            "single_region_ms_per_step": 1e3 * el1 / steps,
            "regions_ms_per_step": 1e3 * elK / steps,
            "exchange_state_ms": 1e3 * el_state, "exchange_grad_ms": 1e3 * el_grad,
            "halo_bytes_per_step": nv * (40 + 120) * st,
            "value": m.n_cells * st * steps / elK}


def adjoint_leg(args, order=1, p_in=100.0, dirs=32, steps=None, checkpoint=None,
                extra_eps=()):
    """The discrete adjoint: a spherical blast (p = p_in inside 0.1 L) in a 55^3 x 6-tet
    box (~1M cells) with a velocity/temperature inflow, HLLC of the given order (2: WLSQ
    gradients + Venkatakrishnan), SSP-RK3 at the solver's own time steps; J =
    time-integrated pressure on a wall patch over 200 steps; gradient with respect to the
    initial state and the inflow values by reverse mode through the adjoint kernels,
    checkpointed every 20 steps, compared with central differences (eps = 1e-6) along
    `dirs` seeded directions (and, for each of extra_eps, a second difference along the
    same directions) and with the forward-mode tangent along 2 of them."""
    steps = args.adj_steps if steps is None else steps
    checkpoint = args.adj_checkpoint if checkpoint is None else checkpoint
    import torch
    import torch.autograd.forward_ad as fwAD
    from rocflu_amd import cases, adjoint as adj
    from rocflu_amd.solver import GpuRegion
    m, num, cv0 = cases.adjoint_case(n=args.adj_n, order=order, p_in=p_in)
    g = GpuRegion(m, num)
    try:
        t0 = time.perf_counter()
        dts, _ = adj.record_time_steps(g, cv0, steps)
        t_solver = time.perf_counter() - t0
        dr = adj.DiffRegion(g)
        gp = num.gas_prime()[2]
        u0 = dr.to_internal(cv0)
        vals0 = torch.tensor(num.patch_vals, dtype=torch.float64, device="cuda")
        mask = torch.zeros_like(vals0)
        mask[0] = 1.0

        def J(u, v, ce=checkpoint):
            return adj.time_integrated_wall_pressure(dr, u, v, dts, 2, gp, ce)[0]

        u = u0.clone().requires_grad_(True)
        v = vals0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        Jv = J(u, v)
        Jv.backward()
        torch.cuda.synchronize()
        t_grad = time.perf_counter() - t1
        gu, gv = u.grad, v.grad
        gen = torch.Generator(device="cpu").manual_seed(4)
        eps = 1e-6
        rel_fd, rel_tan = [], []
        rel_extra = {e: [] for e in extra_eps}
        t2 = time.perf_counter()
        for k in range(dirs):
            du = torch.randn(u0.shape, generator=gen, dtype=torch.float64).to(u0.device) * u0
            dv = torch.randn(vals0.shape, generator=gen, dtype=torch.float64).to(u0.device) * vals0 * mask
            ad = float((gu * du).sum() + (gv * dv).sum())
            with torch.no_grad():
                fd = (float(J(u0 + eps * du, vals0 + eps * dv, None)) -
                      float(J(u0 - eps * du, vals0 - eps * dv, None))) / (2 * eps)
            rel_fd.append(abs(ad - fd) / abs(fd))
            for e in extra_eps:
                with torch.no_grad():
                    fe = (float(J(u0 + e * du, vals0 + e * dv, None)) -
                          float(J(u0 - e * du, vals0 - e * dv, None))) / (2 * e)
                rel_extra[e].append(abs(ad - fe) / abs(fe))
            if k < 2:
                with fwAD.dual_level(), torch.no_grad():
                    jt = fwAD.unpack_dual(J(fwAD.make_dual(u0, du), fwAD.make_dual(vals0, dv),
                                            None)).tangent
                rel_tan.append(abs(ad - float(jt)) / abs(float(jt)))
        t_fd = time.perf_counter() - t2
    finally:
        g.close()
    out = {"workload": f"{args.adj_n}^3 x 6 Kuhn tets ({m.n_cells} cells), blast p={p_in:g} "
                       f"in 0.1 L, inflow u=0.3 at x-, HLLC order {order}"
                       f"{' + Venkatakrishnan' if order == 2 else ''}, SSP-RK3, "
                       f"1% seeded noise",
           "cells": int(m.n_cells), "steps": steps,
           "checkpoint_every": checkpoint, "J": float(Jv),
           "grad_inflow_values": [float(x) for x in gv[0]],
           "directions": dirs, "fd_eps": eps,
           "rel_err_vs_fd_max": max(rel_fd), "rel_err_vs_fd_median": float(np.median(rel_fd)),
           "other_fd_eps": {f"{e:g}": {"max": max(v), "median": float(np.median(v))}
                            for e, v in rel_extra.items()},
           "rel_err_vs_tangent_max": max(rel_tan) if rel_tan else None,
           "solver_s": t_solver, "gradient_s": t_grad, "fd_s": t_fd}
    print(f"# adjoint: J={float(Jv):.6e} grad in {t_grad:.1f}s, vs FD max {max(rel_fd):.2e} "
          f"median {np.median(rel_fd):.2e}, vs tangent {out['rel_err_vs_tangent_max']}",
          file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=160,
                    help="headline box cells per side (6 tets each): the largest single-GPU "
                         "configuration, 160^3 x 6 = 24.6M tets")
    ap.add_argument("--discr", type=int, default=3, help="1 Roe, 3 HLLC, 4 AUSM+")
    ap.add_argument("--reconst", type=int, default=11)
    ap.add_argument("--err-steps", type=int, default=100,
                    help="steps of the mixed-vs-FP64 comparison on the headline mesh")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="time the mixed legs without their error runs (profiling)")
    ap.add_argument("--c2-n", type=int, default=80, help="the 80^3 x 6 configuration")
    ap.add_argument("--c2-err-steps", type=int, default=200)
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=48)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-cases", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--adj-n", type=int, default=55)
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--adj-checkpoint", type=int, default=20)
    ap.add_argument("--adj-dirs", type=int, default=32)
    ap.add_argument("--adj2-dirs", type=int, default=8,
                    help="directions of the second-order adjoint leg (blast p = 5)")
    ap.add_argument("--adj2-steps", type=int, default=3)
    ap.add_argument("--adj2-checkpoint", type=int, default=1)
    ap.add_argument("--no-adjoint2", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--regions", type=int, default=0,
                    help="also run K regions coupled through virtual cells (0: off)")
    ap.add_argument("--regions-n", type=int, default=40)
    ap.add_argument("--regions-steps", type=int, default=5)
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain device copy of 512 MiB three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against known bytes)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note only)"}))
        sys.exit(2)
    ws = int(os.environ.get("WORLD_SIZE", "1"))
    if ws != 1:
        print(json.dumps({"error": "run with one rank"}))
        sys.exit(2)

    from rocflu_amd import cases
    flux_name = {1: "Roe", 3: "HLLC", 4: "AUSM+"}
    lim_name = "Venkatakrishnan" if args.reconst == 11 else "WENO-XYZ"
    t0 = time.time()
    m = cases.tet_box(n=args.n, jitter=0.08, seed=2)
    print(f"# mesh {args.n}^3x6: {m.n_cells} tets, {m.n_faces} faces, "
          f"{m.c2cs_idx.size} stencil entries, built in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    fp = time_region(args, m, build_case(args, 0), exact=True, steps=args.steps,
                     err_steps=0 if args.no_mixed or args.no_accuracy else args.err_steps,
                     dropin=True)
    kt = fp["kernels"]
    dom = max(KERNELS, key=lambda k: kt[k][0])
    ms_launch = kt[dom][0] / max(kt[dom][1], 1)
    # algorithmic bytes without stored stencil data (weights or inverse distances, a
    # design choice that trades HBM bytes for arithmetic); what the variant stores is
    # reported beside them
    ab = algorithmic_bytes(m, fp["num"], dom, fp["tuning"], compulsory=True)
    stored = algorithmic_bytes(m, fp["num"], dom, fp["tuning"]) - ab
    achieved = ab / (ms_launch * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_tet{args.n}_prec0.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        # measured for the same kernel variant only
        rec = tj.get("variants", {}).get(grad_variant(fp["tuning"]) if dom == "gradient" else dom)
        if rec:
            traffic = rec.get("bytes_per_launch")

    def step_hbm(mesh, rr):
        """Compulsory bytes of a whole step (three stages of gradient, fluxes and
        update, one time step) over the timed step, against the HBM peak."""
        per_stage = sum(algorithmic_bytes(mesh, rr["num"], k, rr["tuning"], compulsory=True)
                        for k in ("gradient", "face_flux", "update"))
        total = 3 * per_stage + algorithmic_bytes(mesh, rr["num"], "time_step", rr["tuning"])
        gbs = total / (rr["ms_per_step"] * 1e-3) / 1e9
        return {"compulsory_bytes_per_step": total, "achieved": gbs, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}

    out = {
        "metric": "cell-updates/s per RK stage, FP64 vs mixed at 1e-6 rel L2; % of HBM peak",
        "value": fp["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.n}^3 x 6 Kuhn tets ({m.n_cells} cells, {m.n_faces} "
                               f"faces), jitter 8%, Mach-2 shock + 1% density noise, "
                               f"{flux_name[args.discr]} order 2, WLSQ + {lim_name}, SSP-RK3, "
                               f"slip walls, FP64 (bit-identical kernels)",
                   "cells": int(m.n_cells), "stages_per_step": 3},
        "graph_value": fp["graph_value"],
        "kernel_ms": {k: [v[0], v[1]] for k, v in kt.items()},
        "tuning": fp["tuning"],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": ab,
                     "stored_stencil_bytes_per_launch": stored, "ms_per_launch": ms_launch,
                     "variant": grad_variant(fp["tuning"]) if dom == "gradient" else dom},
        "step_hbm": step_hbm(m, fp),
        "dropin": fp["dropin"],
    }
    if dom == "gradient":
        # the other side of the kernel: its floating-point work against the vector peak
        fl = gradient_flops(m, fp["num"], fp["tuning"])
        tf = fl / (ms_launch * 1e-3) / 1e12
        out["roofline"]["compute"] = {"flops_per_launch": fl, "achieved": tf,
                                      "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                      "frac": tf / FP64_PEAK_TFLOPS}
    if not args.no_mixed:
        acc = not args.no_accuracy
        keep = {}
        mx = time_region(args, m, build_case(args, 1), exact=False, steps=args.steps,
                         err_steps=args.err_steps if acc else 0)
        out["mixed"] = {"value": mx["value"], "ms_per_step": mx["ms_per_step"],
                        "graph_value": mx["graph_value"], "step_hbm": step_hbm(m, mx),
                        "speedup_vs_fp64": mx["value"] / fp["value"],
                        "dtype": "f32 gradients and reconstruction increments, f64 fluxes/"
                                 "state/residual/limiter/positivity",
                        "err_steps": args.err_steps,
                        "rel_l2_per_variable": rel_l2_per_variable(mx["final"], fp["final"])
                        if acc else None,
                        "accuracy": mixed_accuracy(args, m, fp["final"], mx["tuning"], keep)
                        if acc else None,
                        "kernel_ms": {k: [v[0], v[1]] for k, v in mx["kernels"].items()},
                        "tuning": mx["tuning"]}
        # the FP32 tile records (GRAD_LDS 2): faster gradients, further from FP64
        f32 = time_region(args, m, build_case(args, 1), exact=False, steps=args.steps,
                          err_steps=args.err_steps if acc else 0, fp32_records=True)
        out["mixed"]["fp32_records"] = {
            "value": f32["value"], "ms_per_step": f32["ms_per_step"],
            "speedup_vs_fp64": f32["value"] / fp["value"],
            "rel_l2_per_variable": rel_l2_per_variable(f32["final"], fp["final"])
            if acc else None,
            "one_stage": rel_l2_per_variable(
                one_stage_state(m, build_case(args, 1), f32["tuning"]),
                keep["one_stage64"]) if acc else None,
            "kernel_ms": {k: [v[0], v[1]] for k, v in f32["kernels"].items()},
            "tuning": f32["tuning"]}
    # the 80^3 x 6 configuration: HLLC and AUSM+ in FP64, HLLC mixed
    if not args.no_c2:
        try:
            t0 = time.time()
            m2 = cases.tet_box(n=args.c2_n, jitter=0.08, seed=2)
            print(f"# mesh {args.c2_n}^3x6 built in {time.time() - t0:.1f}s", file=sys.stderr,
                  flush=True)
            c2 = {"workload": f"{args.c2_n}^3 x 6 Kuhn tets ({m2.n_cells} cells), jitter 8%, "
                              f"Mach-2 shock + 1% density noise, order 2, WLSQ + {lim_name}, "
                              f"SSP-RK3, slip walls", "cells": int(m2.n_cells)}
            finals = {}
            for label, discr, prec, exact, es in (("hllc_fp64", 3, 0, True, args.c2_err_steps),
                                                  ("ausm_fp64", 4, 0, True, 0),
                                                  ("hllc_mixed", 3, 1, False, args.c2_err_steps)):
                r = time_region(args, m2, build_case(args, prec, discr), exact, args.steps, es)
                finals[label] = r["final"]
                c2[label] = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                             "graph_value": r["graph_value"], "step_hbm": step_hbm(m2, r),
                             "kernel_ms": {k: [v[0], v[1]] for k, v in r["kernels"].items()},
                             "tuning": r["tuning"]}
            c2["hllc_mixed"]["err_steps"] = args.c2_err_steps
            c2["hllc_mixed"]["rel_l2_per_variable"] = rel_l2_per_variable(
                finals["hllc_mixed"], finals["hllc_fp64"])
            out["config_80"] = c2
        except Exception as e:  # noqa: BLE001
            out["config_80"] = {"error": repr(e)}
    # secondary workloads: a failure is reported in the line instead of losing it
    if not args.no_batch:
        try:
            out["batch"] = batch_leg(args)
        except Exception as e:  # noqa: BLE001
            out["batch"] = {"error": repr(e)}
    if not args.no_adjoint:
        try:
            out["adjoint"] = adjoint_leg(args, order=1, p_in=100.0, dirs=args.adj_dirs)
        except Exception as e:  # noqa: BLE001
            out["adjoint"] = {"error": repr(e)}
    if not args.no_adjoint2:
        # second order: the reference's limiter does not keep a p = 100 blast's
        # reconstructed states positive (NaN within the first step), p = 5 runs.  The
        # limited residual is piecewise smooth with many kinks per cell at 1M cells, so a
        # central difference at eps = 1e-6 crosses some of them; it is reported with one at
        # eps = 1e-9 beside it, on a short horizon
        try:
            out["adjoint_order2"] = adjoint_leg(args, order=2, p_in=5.0, dirs=args.adj2_dirs,
                                                steps=args.adj2_steps,
                                                checkpoint=args.adj2_checkpoint,
                                                extra_eps=(1e-9,))
        except Exception as e:  # noqa: BLE001
            out["adjoint_order2"] = {"error": repr(e)}
    if args.regions > 1:
        try:
            out["regions"] = regions_leg(args)
        except Exception as e:  # noqa: BLE001
            out["regions"] = {"error": repr(e)}
    if not args.no_cpu:
        cb = cpu_baseline(args)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = fp["value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
