"""The C-ABI library loads without a GPU, exports every entry point that
include/sgipc.h declares, and the ctypes binding declares exactly those; host
argument checks return status codes without touching a device; the scene
generators produce the shapes and counts the reference configs fix."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sgipc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sgipc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(pkg):
    lib = ctypes.CDLL(pkg._lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding covers the whole ABI, and nothing else
    assert sorted(pkg._lib.exported_symbols()) == names


def test_argument_checks_without_gpu(pkg):
    L = pkg._lib.lib()
    # null handles and negative sizes are rejected before any HIP call
    BAD_ARG = 6
    assert L.sgipc_spmv_srbk(None, 1.0, None, None, None, 0, None, 0.0, None, 0) == BAD_ARG
    assert L.sgipc_convert_srbk(None, None, None, None, 1, None, None, None, 1, None) == BAD_ARG
    assert L.sgipc_mas_destroy(None) == BAD_ARG
    assert L.sgipc_system_destroy(None) == BAD_ARG
    assert L.sgipc_spmat_destroy(None) == BAD_ARG
    assert L.sgipc_abd_jacobi_apply(None, None, 1, None, None) == BAD_ARG
    opts = pkg._lib.PcgOpts()
    L.sgipc_pcg_default_opts(ctypes.byref(opts))
    assert opts.tol > 0 and opts.max_iter == 0 and opts.fuse == 1
    assert L.sgipc_abi_version() >= 1


def test_scene_configs(pkg):
    s = pkg.scenes.scene_c0()
    assert (s.n, s.n_tets, s.nbr_list.size // 2 + s.n) == (4096, 20250, 29791)
    assert np.all(s.vol > 0) and s.rhs.size == 3 * s.n
    c = pkg.scenes.scene_cloth_tiny(N=10, n_pairs=8)
    assert (c.n, c.n_tris, c.n_hinges, int((c.adj[:, 1] < 0).sum())) == (100, 162, 261, 36)
    h = pkg.scenes.scene_hybrid_tiny(N=6, n_bodies=4, n_pairs=10)
    assert h.n_sys == 4 * 4 + 216 and h.rhs.size == 3 * h.n_sys
    assert np.all(h.pairs[:, 0] < 0) and np.all(-h.pairs[:, 0] - 1 < h.n_abd_points)
    assert np.all(h.pairs[:, 1:] >= h.n_abd_points)
