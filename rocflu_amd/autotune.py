"""Autotuner of a region's launch variants.

By default every option it sets leaves results bit-for-bit unchanged (tested in
tests/test_gpu_parity.py::test_tuning_options_do_not_change_results), so the choice is
made on time alone: each candidate is launched on the region's own stream and timed
with HIP events (``rflu_gpu_time_kernel``), and the fastest is kept.  ``exact=False``
adds the moment form of the gradient (``GRAD_FAST``), whose results differ from the
reference's in the last bits (held to 1e-12 relative L2 per variable after 100 steps
by tests/test_gpu_parity.py::test_fast_gradient_within_fp64_bar); with FP32 gradients,
``fp32_records=True`` also adds the FP32 tile records (``GRAD_LDS 2``), which drift
further from the FP64 run over many limited steps.  Choices are cached per mesh/scheme
signature and those flags, in memory and optionally in a JSON file named by
``ROCFLU_AMD_TUNE_CACHE``.

Timing launches run the kernels on the region's current state (the update kernel with a
negligible dt), so upload the initial state after tuning.
"""
from __future__ import annotations

import json
import os
from typing import Dict, List, Optional

from . import _native as nat
from .solver import GpuRegion

K_FACE, K_UPDATE, K_GRAD = 0, 1, 2

_cache: Dict[str, dict] = {}


def signature(g: GpuRegion) -> str:
    return g.signature()


def _load_file() -> dict:
    path = os.environ.get("ROCFLU_AMD_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _store_file(sig: str, choice: dict):
    path = os.environ.get("ROCFLU_AMD_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data[sig] = choice
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def apply(g: GpuRegion, choice: dict):
    """Sets the options of a tuning choice on a region."""
    g.set_option(GpuRegion.OPT_XCD, choice["xcd"])
    g.set_launch(face_block=choice["face_block"], cell_block=choice["cell_block"])
    if g.num.order == 2:
        # choices cached before tile sizes were tuned were made with 128-cell tiles
        _set_tile(g, choice.get("grad_tile", 128))
        g.set_option(GpuRegion.OPT_GRAD_LDS, choice["grad_lds"])
        g.set_option(GpuRegion.OPT_GRAD_UNROLL, choice["grad_unroll"])
        g.set_option(GpuRegion.OPT_GRAD_WEIGHTS, choice["grad_weights"])
        g.set_option(GpuRegion.OPT_GRAD_FAST, choice.get("grad_fast", 0))
        _set_face_tile(g, choice.get("face_tile", 0))
        g.set_option(GpuRegion.OPT_FACE_THREADS, choice.get("face_threads", 256))
        g.set_option(GpuRegion.OPT_UPDATE_TILE, choice.get("update_tile", 0))


def _set_face_tile(g: GpuRegion, ts: int) -> bool:
    """Sets the face tiles of the second-order flux (False if they do not fit LDS)."""
    if g.get_option(GpuRegion.OPT_FACE_TILE) == ts:
        return True
    try:
        g.set_option(GpuRegion.OPT_FACE_TILE, ts)
    except nat.RfluError as e:
        if e.code in (nat.RFLU_ERR_ARG, nat.RFLU_ERR_NOMEM):
            return False
        raise
    return True


FACE_TILES = (0, 128, 64, 256)


def _set_tile(g: GpuRegion, ts: int) -> bool:
    """Rebuilds the gradient tiles with ts cells (False if they do not fit LDS)."""
    if g.get_option(GpuRegion.OPT_GRAD_TILE) == ts:
        return True
    try:
        g.set_option(GpuRegion.OPT_GRAD_TILE, ts)
    except nat.RfluError as e:
        if e.code in (nat.RFLU_ERR_ARG, nat.RFLU_ERR_NOMEM):
            return False
        raise
    return True


def _time(g: GpuRegion, which: int, reps: int) -> float:
    g.time_kernel(which, 2)  # warm-up (first launch of a variant loads its code object)
    return min(g.time_kernel(which, reps) for _ in range(2))


TILE_SIZES = (256, 128, 512)


def gradient_candidates(g: GpuRegion, allow_weights: bool = True,
                        exact: bool = True, fp32_records: bool = False) -> List[dict]:
    """Gradient variants, grouped by tile size (a size change rebuilds the tiles) and,
    within a size, formed weights before stored ones (and, unless exact, the moment
    form last; with fp32_records and FP32 gradients, the FP32 tile records, whose
    differences carry the rounding of each value's deviation from its tile's first cell
    and so drift further from the FP64 run over limited steps)."""
    c = []
    if g.get_option(GpuRegion.OPT_TILE_CAP) > 0:
        for ts in TILE_SIZES:
            for u in ((2, 4, 8) if ts == 256 else (2, 4)):
                c.append(dict(grad_tile=ts, grad_lds=1, grad_unroll=u, grad_weights=0,
                              grad_fast=0))
            if allow_weights:
                c.append(dict(grad_tile=ts, grad_lds=1, grad_unroll=4, grad_weights=2,
                              grad_fast=0))
                for u in (1, 2):
                    c.append(dict(grad_tile=ts, grad_lds=1, grad_unroll=u, grad_weights=1,
                                  grad_fast=0))
            if not exact:
                for u in (1, 2, 4):
                    c.append(dict(grad_tile=ts, grad_lds=1, grad_unroll=u, grad_weights=0,
                                  grad_fast=1))
                if fp32_records and g.num.precision in (1, 2):
                    for u in (4, 8):
                        c.append(dict(grad_tile=ts, grad_lds=2, grad_unroll=u, grad_weights=0,
                                      grad_fast=0))
    c.append(dict(grad_tile=g.get_option(GpuRegion.OPT_GRAD_TILE), grad_lds=0, grad_unroll=4,
                  grad_weights=0, grad_fast=0))
    return c


def tune(g: GpuRegion, reps: int = 10, allow_weights: bool = True, use_cache: bool = True,
         verbose: bool = False, exact: bool = True, fp32_records: bool = False) -> dict:
    """Chooses and applies the fastest variant of each kernel of the stage for region g
    (exact=False: also the gradient's moment form, not bit-identical; fp32_records with
    exact=False: also the FP32 tile records of FP32 gradients).  Returns the choice with
    the measured milliseconds of every candidate."""
    sig = signature(g) + ("" if exact else "|fast") + ("|f32rec" if fp32_records else "")
    if use_cache:
        if not _cache:
            _cache.update(_load_file())
        if sig in _cache:
            apply(g, _cache[sig])
            return dict(_cache[sig], cached=True)
    choice = dict(xcd=8, face_block=256, cell_block=256, grad_tile=256, grad_lds=0,
                  grad_unroll=4, grad_weights=0, grad_fast=0, face_tile=0, face_threads=256,
                  update_tile=0)
    timings = {}
    # XCD-aware block order and block sizes of the face and cell kernels
    best = None
    for xcd in (8, 1):
        for blk in (256, 512):
            g.set_option(GpuRegion.OPT_XCD, xcd)
            g.set_launch(face_block=blk, cell_block=blk)
            t = _time(g, K_FACE, reps) + _time(g, K_UPDATE, reps)
            timings[f"xcd{xcd}_blk{blk}"] = t
            if best is None or t < best[0]:
                best = (t, xcd, blk)
    choice.update(xcd=best[1], face_block=best[2], cell_block=best[2])
    g.set_option(GpuRegion.OPT_XCD, choice["xcd"])
    g.set_launch(face_block=choice["face_block"], cell_block=choice["cell_block"])
    if g.num.order == 2:
        # interior flux: one thread per face, or face tiles staged in LDS by blocks of
        # 256 or 512 threads
        best = None
        for ft in FACE_TILES:
            if not _set_face_tile(g, ft):
                continue
            for nt in ((256, 512) if ft else (256,)):
                g.set_option(GpuRegion.OPT_FACE_THREADS, nt)
                t = _time(g, K_FACE, reps)
                timings[f"face_tile{ft}" + (f"_nt{nt}" if ft else "")] = t
                if best is None or t < best[0]:
                    best = (t, ft, nt)
        choice["face_tile"], choice["face_threads"] = best[1], best[2]
        _set_face_tile(g, best[1])
        g.set_option(GpuRegion.OPT_FACE_THREADS, best[2])
        if best[1] > 0:
            # the update over the same tiles, its own faces' fluxes from LDS
            best = None
            for ut in (0, 1):
                g.set_option(GpuRegion.OPT_UPDATE_TILE, ut)
                t = _time(g, K_UPDATE, reps)
                timings[f"update_tile{ut}"] = t
                if best is None or t < best[0]:
                    best = (t, ut)
            choice["update_tile"] = best[1]
            g.set_option(GpuRegion.OPT_UPDATE_TILE, best[1])
        best = None
        for cand in gradient_candidates(g, allow_weights, exact, fp32_records):
            if g.get_option(GpuRegion.OPT_GRAD_TILE) != cand["grad_tile"]:
                g.set_option(GpuRegion.OPT_GRAD_WEIGHTS, 0)   # no weights rebuilt in vain
                if not _set_tile(g, cand["grad_tile"]):
                    continue
            try:
                g.set_option(GpuRegion.OPT_GRAD_WEIGHTS, cand["grad_weights"])
            except nat.RfluError as e:
                if e.code == nat.RFLU_ERR_NOMEM:
                    continue  # no room for the stored weights
                raise
            g.set_option(GpuRegion.OPT_GRAD_LDS, cand["grad_lds"])
            g.set_option(GpuRegion.OPT_GRAD_UNROLL, cand["grad_unroll"])
            g.set_option(GpuRegion.OPT_GRAD_FAST, cand["grad_fast"])
            t = _time(g, K_GRAD, reps)
            key = "grad_t{grad_tile}_lds{grad_lds}_u{grad_unroll}_w{grad_weights}".format(**cand)
            if cand["grad_fast"]:
                key += "_fast"
            timings[key] = t
            if verbose:
                print(f"autotune {sig}: {key} {t:.3f} ms", flush=True)
            if best is None or t < best[0]:
                best = (t, cand)
        choice.update(best[1])
        g.set_option(GpuRegion.OPT_GRAD_WEIGHTS, 0)
        g.set_option(GpuRegion.OPT_GRAD_FAST, 0)
        apply(g, choice)
        # the block order again, now over the chosen tiles and with the gradient: the
        # first pass timed the one-thread-per-face flux without tiles
        best = None
        for xcd in (8, 1):
            g.set_option(GpuRegion.OPT_XCD, xcd)
            t = sum(_time(g, k, reps) for k in (K_GRAD, K_FACE, K_UPDATE))
            timings[f"stage_xcd{xcd}"] = t
            if best is None or t < best[0]:
                best = (t, xcd)
        choice["xcd"] = best[1]
        g.set_option(GpuRegion.OPT_XCD, best[1])
    _cache[sig] = dict(choice)
    _store_file(sig, choice)
    return dict(choice, timings_ms=timings, cached=False)
