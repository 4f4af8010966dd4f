"""Feasible step sizes, the Newton objective and the line search on the GPU
against the oracle through the C ABI.

Tolerances: CCD, ground and injectivity steps within 1e-9 relative (the
same conservative-advancement arithmetic; the step count may differ by the
last rounding of a distance); energies within 1e-10 relative (fixed-order
sums, different order); the line search's step equal to the oracle's and its
energies within 1e-10 relative.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_ccd_self_step_matches_oracle(pkg, orc, gpu_ctx):
    L, h = pkg._lib.lib(), gpu_ctx.h
    geo = pkg.scenes.contact_sheets(24)
    n1 = geo.n // 2
    d = np.zeros_like(geo.x)
    d[n1:, 2] = 2 * geo.dhat
    d += np.random.default_rng(3).uniform(-1e-4, 1e-4, d.shape)
    want, nc = orc.ccd_self_step(geo.x, d, 1.0, geo.faces, geo.edges, geo.surf, geo.dhat ** 2, 0.8)
    dx, dd = gpu_ctx.to_device(geo.x.reshape(-1)), gpu_ctx.to_device(d.reshape(-1))
    f, e, sv = (gpu_ctx.to_device(geo.faces.reshape(-1), np.uint32), gpu_ctx.to_device(geo.edges.reshape(-1), np.uint32),
                gpu_ctx.to_device(geo.surf, np.uint32))
    step, cnt = ctypes.c_double(), ctypes.c_int()
    pkg._lib.check(L.sgipc_ccd_self_step(h, dx.vp, dd.vp, 1.0, f.vp, geo.faces.shape[0], e.vp, geo.edges.shape[0],
                                         sv.vp, geo.surf.size, None, None, geo.dhat ** 2, 0.8, ctypes.byref(step),
                                         ctypes.byref(cnt)), h)
    print(f"ccd: step {step.value:.6f} (oracle {want:.6f}), {cnt.value} candidates (oracle {nc})")
    assert cnt.value == nc
    assert step.value == pytest.approx(want, rel=1e-9)


def test_ground_and_injective_steps_match_oracle(pkg, orc, gpu_ctx):
    L, h = pkg._lib.lib(), gpu_ctx.h
    s = pkg.scenes.tet_grid_scene("step_grid8", 8, n_pairs=0)
    rng = np.random.default_rng(5)
    d = rng.uniform(-4e-3, 4e-3, s.x.shape)
    want = orc.injective_step(s.x, s.tets, d, 0.8)
    dx, dd = gpu_ctx.to_device(s.x.reshape(-1)), gpu_ctx.to_device(d.reshape(-1))
    dt = gpu_ctx.to_device(s.tets.reshape(-1), np.uint32)
    step = ctypes.c_double()
    pkg._lib.check(L.sgipc_injective_step(h, dx.vp, dt.vp, s.n_tets, dd.vp, 0.8, 1e-20, ctypes.byref(step)), h)
    print(f"injective step {step.value:.6f} (oracle {want:.6f})")
    assert 0 < want < 1 and step.value == pytest.approx(want, rel=1e-9)
    nrm = np.array([0.0, 1.0, 0.0])
    off = s.x[:, 1].min() - 1e-3
    surf = np.arange(s.n, dtype=np.int32)
    dg = np.zeros_like(s.x)
    dg[:, 1] = rng.uniform(0, 2e-3, s.n)
    want = orc.ground_step(s.x, surf, nrm, off, dg, 0.8)
    dn, dsv, ddg = gpu_ctx.to_device(nrm), gpu_ctx.to_device(surf, np.uint32), gpu_ctx.to_device(dg.reshape(-1))
    pkg._lib.check(L.sgipc_ground_step(h, dx.vp, dsv.vp, s.n, dn.vp, off, ddg.vp, 0.8, ctypes.byref(step)), h)
    assert step.value == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("kind", ["tets", "cloth_contact"])
def test_energy_and_line_search_match_oracle(pkg, orc, kind):
    if kind == "tets":
        s = pkg.scenes.tet_grid_scene("ls_grid8", 8, n_pairs=0)
    else:
        s = pkg.scenes.cloth_contact_scene(12, kappa=100.0)
    L = pkg._lib.lib()
    _, _, _, g = orc.triplets(s, with_grad=True)
    xt = s.x + np.random.default_rng(2).uniform(-1e-5, 1e-5, s.x.shape)
    s2 = s.__class__(**{**s.__dict__, "rhs": g + np.repeat(s.mass, 3) * (s.x - xt).reshape(-1)})
    d = orc.solve_scene(s2, 32, tol=1e-8).x.reshape(-1, 3)
    ls = pkg.LinearSystem(s, bank_size=32)
    try:
        ctx = ls.ctx
        dxt, dd = ctx.to_device(xt.reshape(-1)), ctx.to_device(d.reshape(-1))
        xo = ctx.empty(3 * s.n, np.float64)
        e = ctypes.c_double()
        pkg._lib.check(L.sgipc_system_energy(ls.sys.h, ctypes.c_void_p(ls.sys._keep["x"].ptr), dxt.vp,
                                             ctypes.byref(e)), ctx.h)
        want = orc.scene_energy(s, s.x, xt)
        assert e.value == pytest.approx(want, rel=1e-10)
        a, en, st = ctypes.c_double(), (ctypes.c_double * 2)(), ctypes.c_int()
        pkg._lib.check(L.sgipc_system_line_search(ls.sys.h, dd.vp, dxt.vp, 1.0, 0.8, xo.vp, ctypes.byref(a), en,
                                                  ctypes.byref(st)), ctx.h)
        wa, wx, we0, we, wst = orc.line_search(s, d, xt, 1.0, 0.8)
        print(f"{kind}: alpha {a.value:.6g} (oracle {wa:.6g}), E {en[0]:.6e} -> {en[1]:.6e}")
        assert a.value == pytest.approx(wa, rel=1e-9) and st.value == int(wst)
        assert en[0] == pytest.approx(we0, rel=1e-10) and en[1] == pytest.approx(we, rel=1e-9)
        np.testing.assert_allclose(xo.download().reshape(-1, 3), wx, rtol=0, atol=1e-12 * np.abs(wx).max())
        assert en[1] < en[0]
    finally:
        ls.close()
