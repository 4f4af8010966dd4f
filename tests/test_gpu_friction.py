"""Lagged friction on the GPU against the oracle through the C ABI.

Tolerances: normal forces, closest-point coordinates and tangent bases
within 1e-10 relative (same formulas, different contraction of products);
friction blocks within 1e-9 of the largest block of their pair, gradients
within 1e-10 relative; ground blocks and gradients within 1e-12 relative.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
vp = ctypes.c_void_p


def _blockscale(ov, counts):
    nt = ov.shape[0]
    scale = np.zeros(nt)
    start = 0
    for sz, cnt in ((16, counts[4]), (9, counts[3]), (4, counts[2])):
        if cnt:
            blk = np.abs(ov[start:start + sz * cnt]).reshape(cnt, sz * 9).max(1)
            scale[start:start + sz * cnt] = np.repeat(blk, sz)
        start += sz * cnt
    return scale


@pytest.mark.parametrize("eps", [5e-7, 2e-5])
def test_pair_friction_matches_oracle(pkg, orc, gpu_ctx, eps):
    L, h = pkg._lib.lib(), gpu_ctx.h
    geo = pkg.scenes.contact_sheets(24)
    pairs, mi, counts = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2)
    lam, coord, basis = orc.friction_prepare(geo.x, pairs, geo.dhat ** 2, geo.kappa)
    x = geo.x + np.random.default_rng(11).uniform(-2e-6, 2e-6, geo.x.shape)
    mu = 0.4
    ov, orr, oc, og = orc.friction(x, geo.x, pairs, mi, counts, lam, coord, basis, eps * eps, mu, hess_offset=2)
    n = pairs.shape[0]
    dxt, dx = gpu_ctx.to_device(geo.x.reshape(-1)), gpu_ctx.to_device(x.reshape(-1))
    dp, dm = gpu_ctx.to_device(pairs.reshape(-1), np.int32), gpu_ctx.to_device(mi, np.int32)
    dl, dc, db = gpu_ctx.empty(n, np.float64), gpu_ctx.empty(2 * n, np.float64), gpu_ctx.empty(6 * n, np.float64)
    pkg._lib.check(L.sgipc_friction_prepare(h, dxt.vp, dp.vp, n, geo.dhat ** 2, geo.kappa, dl.vp, dc.vp, db.vp), h)
    for got, want in ((dl.download(), lam), (dc.download(), coord.reshape(-1)), (db.download(), basis.reshape(-1))):
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    nt = orr.size
    tv, tr, tc = gpu_ctx.empty((nt, 9), np.float64), gpu_ctx.empty(nt, np.int32), gpu_ctx.empty(nt, np.int32)
    g = gpu_ctx.to_device(np.zeros(x.size))
    c5 = (ctypes.c_int * 5)(*[int(v) for v in counts])
    pkg._lib.check(L.sgipc_friction_gradient_hessian(h, dx.vp, dxt.vp, dp.vp, dm.vp, n, c5, dl.vp, dc.vp, db.vp,
                                                     eps * eps, mu, 2, g.vp, tv.vp, tr.vp, tc.vp), h)
    np.testing.assert_array_equal(tr.download(), orr)
    np.testing.assert_array_equal(tc.download(), oc)
    err = np.abs(tv.download() - ov).max(1) / np.maximum(_blockscale(ov, counts), 1e-300)
    assert float(err.max()) <= 1e-9, float(err.max())
    gg = g.download()
    assert np.linalg.norm(gg - og) <= 1e-10 * np.linalg.norm(og)
    print(f"friction eps {eps:g}: {n} pairs, max block err {err.max():.2e}")


def test_ground_friction_matches_oracle(pkg, orc, gpu_ctx):
    L, h = pkg._lib.lib(), gpu_ctx.h
    rng = np.random.default_rng(4)
    n = 5000
    nrm = np.array([0.0, 1.0, 0.0])
    dsq, kappa, mu, offset = 1e-6, 1e4, 0.5, -0.2
    xt = rng.uniform(-1, 1, (n, 3))
    xt[:, 1] = offset + rng.uniform(0.1, 0.9, n) * 1e-3
    verts = np.arange(n, dtype=np.int32)
    x = xt + rng.uniform(-2e-6, 2e-6, xt.shape)
    lam = orc.ground_friction_prepare(xt, verts, nrm, offset, dsq, kappa)
    eps2 = 1e-12
    ov, orr, oc, og = orc.ground_friction(x, xt, verts, nrm, lam, eps2, mu, hess_offset=1)
    dxt, dx, dv, dn = (gpu_ctx.to_device(xt.reshape(-1)), gpu_ctx.to_device(x.reshape(-1)),
                       gpu_ctx.to_device(verts, np.int32), gpu_ctx.to_device(nrm))
    dl = gpu_ctx.empty(n, np.float64)
    pkg._lib.check(L.sgipc_ground_friction_prepare(h, dxt.vp, dv.vp, n, dn.vp, offset, dsq, kappa, dl.vp), h)
    assert np.abs(dl.download() - lam).max() <= 1e-12 * np.abs(lam).max()
    tv, tr, tc = gpu_ctx.empty((n, 9), np.float64), gpu_ctx.empty(n, np.int32), gpu_ctx.empty(n, np.int32)
    g = gpu_ctx.to_device(np.zeros(3 * n))
    pkg._lib.check(L.sgipc_ground_friction_gradient_hessian(h, dx.vp, dxt.vp, dv.vp, n, dn.vp, dl.vp, eps2, mu, 1,
                                                            g.vp, tv.vp, tr.vp, tc.vp), h)
    np.testing.assert_array_equal(tr.download(), orr)
    assert np.abs(tv.download() - ov).max() <= 1e-12 * np.abs(ov).max()
    assert np.linalg.norm(g.download() - og) <= 1e-12 * np.linalg.norm(og)
