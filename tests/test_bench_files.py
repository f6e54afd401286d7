"""The measured-traffic record the bench line reads: it names a gradient variant in the
form the bench's tuning report uses, so that the bench can match it to the variant it
runs, and it points at a committed profile summary."""
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_traffic_record_names_a_variant_the_bench_can_match():
    import glob
    import sys
    sys.path.insert(0, ROOT)
    import bench
    files = sorted(glob.glob(os.path.join(ROOT, "profiles", "traffic_gradient_tet*_prec*.json")))
    assert any("tet160_prec0" in f for f in files)      # the headline configuration
    for p in files:
        d = json.load(open(p))
        assert d["variants"]
        for v, rec in d["variants"].items():
            m = re.fullmatch(r"grad_t(\d+)_lds([012])_u(\d)_w([012])(_fast)?", v)
            assert m, v
            t = {"grad_tile": int(m.group(1)), "grad_lds": int(m.group(2)),
                 "grad_unroll": int(m.group(3)), "grad_weights": int(m.group(4)),
                 "grad_fast": 1 if m.group(5) else 0}
            assert bench.grad_variant(t) == v
            assert rec["bytes_per_launch"] > 0
        assert os.path.exists(os.path.join(ROOT, d["source"]))
