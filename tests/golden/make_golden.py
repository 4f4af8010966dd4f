"""Writes tests/golden/c0_summary.json: a fingerprint of the oracle's results
on the 16^3 scene (assembled BCOO, MAS hierarchy, PCG iteration counts under
both stop rules).  The reference ships no golden vectors, so these are
regression fixtures of the oracle itself, re-checked by
tests/test_golden.py; regenerate only after a deliberate oracle change:
    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def fingerprint(orc, pkg):
    s = pkg.scenes.scene_c0()
    vals, rows, cols = orc.assemble(s)
    out = {
        "n": int(s.n), "n_tets": int(s.n_tets), "nnz": int(rows.size),
        "index_sha256": hashlib.sha256(rows.tobytes() + cols.tobytes()).hexdigest(),
        "vals_abs_sum": float(np.abs(vals).sum()),
        "diag_trace": float(sum(vals[k, 0] + vals[k, 4] + vals[k, 8] for k in np.flatnonzero(rows == cols))),
    }
    for B in (16, 32):
        m = orc.Mas(s, B)
        m.build(vals, rows, cols, 0, s.pairs)
        L, off, cnt, total = m.levels()
        out[f"mas{B}"] = {"levels": int(L), "offsets": [int(v) for v in off], "counts": [int(v) for v in cnt]}
        for mode in (0, 1):
            tol = 1e-4 if mode == 0 else 1e-5
            x, it, res = orc.pcg(vals, rows, cols, s.n, s.rhs, m, tol=tol, stop_mode=mode)
            out[f"mas{B}"][f"pcg_stop{mode}"] = {"tol": tol, "iterations": int(it), "rel_residual": float(res)}
        m.close()
    return out


if __name__ == "__main__":
    import __graft_entry__ as ge
    fp = fingerprint(ge.load_oracle(), ge.load_package())
    with open(os.path.join(HERE, "c0_summary.json"), "w") as f:
        json.dump(fp, f, indent=1, sort_keys=True)
    print(json.dumps(fp, indent=1))
