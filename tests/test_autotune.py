"""The autotuner's candidate lists (CPU): bit-identical variants only unless asked, the
moment form only with exact=False, the FP32 tile records only with fp32_records and FP32
gradients, a global-gather fallback always."""
from types import SimpleNamespace

import pytest

from rocflu_amd import autotune
from rocflu_amd.solver import GpuRegion


class FakeRegion:
    """The two options gradient_candidates reads, and the numerics' precision."""

    def __init__(self, precision, tile_cap=1200, tile=256):
        self.num = SimpleNamespace(precision=precision, order=2)
        self.opts = {GpuRegion.OPT_TILE_CAP: tile_cap, GpuRegion.OPT_GRAD_TILE: tile}

    def get_option(self, opt):
        return self.opts[opt]


@pytest.mark.parametrize("precision", [0, 1])
def test_exact_candidates_are_the_bit_identical_ones(precision):
    c = autotune.gradient_candidates(FakeRegion(precision), exact=True, fp32_records=True)
    assert c and all(x["grad_fast"] == 0 and x["grad_lds"] in (0, 1) for x in c)
    assert {x["grad_tile"] for x in c if x["grad_lds"]} == set(autotune.TILE_SIZES)
    assert c[-1]["grad_lds"] == 0      # the global-gather kernel stays a candidate


def test_inexact_candidates():
    fast = autotune.gradient_candidates(FakeRegion(1), exact=False)
    assert any(x["grad_fast"] for x in fast)
    assert not any(x["grad_lds"] == 2 for x in fast)
    rec = autotune.gradient_candidates(FakeRegion(1), exact=False, fp32_records=True)
    assert {x["grad_tile"] for x in rec if x["grad_lds"] == 2} == set(autotune.TILE_SIZES)
    assert all(x["grad_weights"] == 0 and x["grad_fast"] == 0 and x["grad_unroll"] in (4, 8)
               for x in rec if x["grad_lds"] == 2)
    # FP64 gradients have no FP32 records
    f64 = autotune.gradient_candidates(FakeRegion(0), exact=False, fp32_records=True)
    assert not any(x["grad_lds"] == 2 for x in f64)


def test_no_tiles_leaves_the_global_gather_only():
    c = autotune.gradient_candidates(FakeRegion(1, tile_cap=0), exact=False, fp32_records=True)
    assert len(c) == 1 and c[0]["grad_lds"] == 0
