"""Contact detection (LBVH) and the barrier of every pair type on the GPU,
against the oracle through the C ABI.

Tolerances:
  pairs, MatIndex, class counts      bit-exact, in the same order (the oracle
                                      tests every candidate by brute force;
                                      the device's order is fixed by sorting)
  two detections of the same input   identical
  barrier blocks                     within 1e-9 of the largest block of the
                                      pair (nearly parallel edge pairs amplify
                                      rounding in their closest-point
                                      parameters); gradient within 1e-10 relative
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def upload_geo(ctx, geo):
    d = dict(x=ctx.to_device(geo.x.reshape(-1)), rest=ctx.to_device(geo.rest.reshape(-1)),
             faces=ctx.to_device(geo.faces.reshape(-1), np.uint32), edges=ctx.to_device(geo.edges.reshape(-1), np.uint32),
             surf=ctx.to_device(geo.surf, np.uint32))
    d["body"] = ctx.to_device(geo.body_id, np.int32) if geo.body_id is not None else None
    return d


def gpu_detect(pkg, ctx, geo, d, capacity=None):
    det = pkg.ContactDetector(ctx)
    return det, det.detect(d["x"], d["rest"], d["faces"], geo.faces.shape[0], d["edges"], geo.edges.shape[0],
                           d["surf"], geo.surf.size, geo.n, geo.dhat ** 2, body_id=d["body"], capacity=capacity)


@pytest.mark.parametrize("N,bodies", [(16, False), (24, False), (48, False), (20, True)])
def test_detection_matches_oracle(pkg, orc, gpu_ctx, N, bodies):
    geo = pkg.scenes.contact_sheets(N, body_ids=bodies)
    want, wmi, wc = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2,
                                       body_id=geo.body_id)
    d = upload_geo(gpu_ctx, geo)
    _, (pairs, mi, counts) = gpu_detect(pkg, gpu_ctx, geo, d)
    n = int(counts[0])
    print(f"sheets{N}: {n} pairs (PP {counts[2]}, PE {counts[3]}, PT+EE {counts[4]})")
    np.testing.assert_array_equal(counts, wc)
    np.testing.assert_array_equal(pairs.download()[:n], want)
    np.testing.assert_array_equal(mi.download()[:n], wmi)
    # the same input again: the same pairs in the same order
    _, (p2, m2, c2) = gpu_detect(pkg, gpu_ctx, geo, d)
    np.testing.assert_array_equal(p2.download()[:n], pairs.download()[:n])
    np.testing.assert_array_equal(m2.download()[:n], mi.download()[:n])


def test_detection_capacity(pkg, orc, gpu_ctx):
    geo = pkg.scenes.contact_sheets(12)
    want, _, _ = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2)
    d = upload_geo(gpu_ctx, geo)
    with pytest.raises(pkg.SgipcError) as ei:
        gpu_detect(pkg, gpu_ctx, geo, d, capacity=want.shape[0] - 1)
    assert ei.value.status == pkg._lib.SGIPC_CAPACITY


def test_detection_large_sheet_properties(pkg, orc, gpu_ctx):
    """A 2 x 256^2-vertex scene (too large for the brute-force oracle): the
    class counts agree with the encodings, and the oracle's barrier energy of
    the reported pairs is finite and positive (every pair lies inside dhat,
    where ln(I5) is defined)."""
    geo = pkg.scenes.contact_sheets(256)
    d = upload_geo(gpu_ctx, geo)
    _, (pairs, mi, counts) = gpu_detect(pkg, gpu_ctx, geo, d)
    n = int(counts[0])
    pr = pairs.download()[:n]
    kind = np.where(pr[:, 0] >= 0, 4, np.where(pr[:, 2] < 0, 2, np.where(pr[:, 3] < 0, 3, 4)))
    for c in (2, 3, 4):
        assert (kind == c).sum() == counts[c]
    e = orc.barrier_energy(geo.x, pr, geo.dhat ** 2, geo.kappa)
    assert np.isfinite(e) and e > 0
    print(f"sheets256: {n} pairs, barrier energy {e:.3e}")
    assert n > 4 * geo.n


@pytest.mark.parametrize("N", [12, 32])
def test_barrier_matches_oracle(pkg, orc, gpu_ctx, N):
    geo = pkg.scenes.contact_sheets(N)
    pairs, mi, counts = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2)
    ov, orr, oc, og = orc.barrier(geo.x, pairs, mi, counts, geo.dhat ** 2, geo.kappa, hess_offset=3)
    det = pkg.ContactDetector(gpu_ctx)
    x = gpu_ctx.to_device(geo.x.reshape(-1))
    tv, tr, tc, g = det.barrier(x, gpu_ctx.to_device(pairs.reshape(-1), np.int32), gpu_ctx.to_device(mi, np.int32),
                                counts, geo.dhat ** 2, geo.kappa, geo.n, hess_offset=3)
    nt = orr.size
    np.testing.assert_array_equal(tr.download()[:nt], orr)
    np.testing.assert_array_equal(tc.download()[:nt], oc)
    gv = tv.download()[:nt]
    # per pair, against its largest block
    scale = np.zeros(nt)
    start = 0
    for sz, cnt in ((16, counts[4]), (9, counts[3]), (4, counts[2])):
        blk = np.abs(ov[start:start + sz * cnt]).reshape(cnt, sz * 9).max(1) if cnt else np.zeros(0)
        scale[start:start + sz * cnt] = np.repeat(blk, sz)
        start += sz * cnt
    err = np.abs(gv - ov).max(1) / np.maximum(scale, 1e-300)
    assert float(err.max()) <= 1e-9, float(err.max())
    gg = g.download()
    assert np.linalg.norm(gg - og) <= 1e-10 * np.linalg.norm(og)


@pytest.mark.parametrize("N", [12, 24])
def test_system_with_detected_contacts_matches_oracle(pkg, orc, N):
    """Two draped cloth sheets whose contacts are detected on every solve:
    the system's pairs equal the oracle's detection, the assembled BCOO equals
    the oracle's (indices exact, blocks within 1e-10 of their row's largest),
    12 fixed FP64 PCG iterations agree to 1e-8 and the solve reaches 1e-5.
    kappa = 100 keeps the sheets' condition number near 1e8 (at kappa = 1e6
    the barrier blocks push it past 5e9, where rounding differences grow
    along the PCG trajectory; that case is checked without the trajectory
    below)."""
    s = pkg.scenes.cloth_contact_scene(N, kappa=100.0)
    ref = orc.assemble(s)
    pr, mi, counts = orc.scene_pairs(s, with_index=True)
    ls = pkg.LinearSystem(s, bank_size=32)
    try:
        fixed = ls.solve(tol=1e-30, max_iter=12, precision="fp64")
        gp, gm = ls.sys.contacts()
        np.testing.assert_array_equal(gp, pr)
        np.testing.assert_array_equal(gm, mi)
        gv, gr, gc = ls.sys.bcoo()
        ov, orr, oc = ref
        np.testing.assert_array_equal(gr, orr)
        np.testing.assert_array_equal(gc, oc)
        rowmax = np.zeros(int(orr.max()) + 1)
        np.maximum.at(rowmax, orr, np.abs(ov).max(1))
        assert float((np.abs(gv - ov).max(1) / rowmax[orr]).max()) <= 1e-10
        xr = orc.solve_scene(s, 32, tol=1e-30, max_iter=12).x
        e = np.linalg.norm(fixed.x - xr) / np.linalg.norm(xr)
        res = ls.solve(tol=1e-5, precision="fp64")
        print(f"cloth_contact{N}: {pr.shape[0]} pairs (PP {counts[2]}, PE {counts[3]}, PT+EE {counts[4]}), "
              f"fixed-12 rel diff {e:.2e}, solve {res.iterations} it res {res.true_rel_residual:.2e}")
        assert e <= 1e-8
        assert res.status == 0 and res.true_rel_residual <= 1e-5
    finally:
        ls.close()


def test_system_with_stiff_detected_contacts(pkg, orc):
    """kappa = 1e6 (condition number above 5e9): the detected pairs and the
    assembled BCOO equal the oracle's; the two-grid FP64 solve reaches 1e-5,
    checked with the oracle's SpMV on the oracle's matrix."""
    s = pkg.scenes.cloth_contact_scene(24)
    ov, orr, oc = orc.assemble(s)
    pr, mi, _ = orc.scene_pairs(s, with_index=True)
    ls = pkg.LinearSystem(s, bank_size=32, coarse_mode=2)
    try:
        res = ls.solve(tol=1e-5, precision="fp64", max_iter=20000)
        gp, gm = ls.sys.contacts()
        np.testing.assert_array_equal(gp, pr)
        np.testing.assert_array_equal(gm, mi)
        gv, gr, gc = ls.sys.bcoo()
        np.testing.assert_array_equal(gr, orr)
        np.testing.assert_array_equal(gc, oc)
        ax = orc.spmv(1.0, ov, orr, oc, res.x, 0.0, np.zeros(3 * s.n), s.n)
        r = np.linalg.norm(s.rhs - ax) / np.linalg.norm(s.rhs)
        print(f"cloth_contact24 kappa 1e6: {pr.shape[0]} pairs, mode {res.coarse_mode} {res.iterations} it, res {r:.2e}")
        assert res.status == 0 and r <= 1.01e-5
    finally:
        ls.close()
