"""The oracle's results on the 16^3 scene match the committed fingerprint
(tests/golden/c0_summary.json, written by tests/golden/make_golden.py)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


def test_c0_fingerprint(orc, pkg):
    from make_golden import fingerprint
    ref = json.load(open(os.path.join(HERE, "golden", "c0_summary.json")))
    got = fingerprint(orc, pkg)
    for k in ("n", "n_tets", "nnz", "index_sha256"):
        assert got[k] == ref[k], k
    for k in ("vals_abs_sum", "diag_trace"):
        assert abs(got[k] - ref[k]) <= 1e-12 * abs(ref[k]), k
    for B in (16, 32):
        g, r = got[f"mas{B}"], ref[f"mas{B}"]
        assert (g["levels"], g["offsets"], g["counts"]) == (r["levels"], r["offsets"], r["counts"])
        for mode in (0, 1):
            gi, ri = g[f"pcg_stop{mode}"], r[f"pcg_stop{mode}"]
            assert abs(gi["iterations"] - ri["iterations"]) <= 1
            if mode == 1:
                assert gi["rel_residual"] <= ri["tol"]
