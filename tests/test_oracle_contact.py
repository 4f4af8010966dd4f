"""Pins of the oracle's contact detection and barrier of every pair type
(oracle/contact_oracle.c), independent of its formulas:

* detection: an independent numpy restatement -- exact closest points by
  Voronoi regions (Ericson, Real-Time Collision Detection, 5.1.5
  point-triangle and 5.1.9 segment-segment) over every candidate -- must give
  the same pairs, in the same order, with the same encodings and classes;
* barrier gradient: central differences of the pairs' barrier energy;
* barrier Hessian: lambda0 g g^T with g the central-difference gradient of
  s = dist / sqrt(dHat) and lambda0 = d2E/ds2 by central differences of E(s);
* a planted wrong point-edge weight is caught by the gradient check.
"""
import numpy as np
import pytest


def _closest_tri_region(p, a, b, c):
    """Ericson 5.1.5: closest point of triangle abc to p and its Voronoi
    region: 0/1/2 vertex a/b/c, 3 edge ab, 4 edge bc, 5 edge ca, 6 face."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a, 0
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b, 1
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab, 3
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c, 2
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac, 5
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b), 4
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den), 6


def _closest_seg_seg(p1, q1, p2, q2):
    """Ericson 5.1.9: parameters s, t of the closest points of segments
    p1q1 and p2q2 (clamped to [0, 1])."""
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0.0, 1.0) if den != 0 else 0.0
    t = (b * s + f) / e
    if t < 0:
        t, s = 0.0, np.clip(-c / a, 0.0, 1.0)
    elif t > 1:
        t, s = 1.0, np.clip((b - c) / a, 0.0, 1.0)
    return s, t


def restated_detection(geo):
    """Every (surface vertex, face) and (edge, larger edge) candidate, in
    query order, classified by its exact closest feature; the reference's
    filters and AABB gap test apply."""
    x, rest, dsq = geo.x, geo.rest, geo.dhat ** 2
    gap = geo.dhat
    out, cls = [], []
    fx = x[geo.faces]
    flo, fhi = fx.min(1), fx.max(1)
    for p in geo.surf:
        P = x[p]
        cand = np.flatnonzero(~((flo - P >= gap).any(1) | (P - fhi >= gap).any(1)))
        for f in cand:
            t0, t1, t2 = (int(v) for v in geo.faces[f])
            if geo.body_id is not None and geo.body_id[p] == geo.body_id[t0] and geo.body_id[p] != -1:
                continue
            if p in (t0, t1, t2):
                continue
            q, reg = _closest_tri_region(P, x[t0], x[t1], x[t2])
            if (P - q) @ (P - q) >= dsq:
                continue
            m = -int(p) - 1
            enc = {0: ((m, t0, -1, -1), 2), 1: ((m, t1, -1, -1), 2), 2: ((m, t2, -1, -1), 2),
                   3: ((m, t0, t1, -1), 3), 4: ((m, t1, t2, -1), 3), 5: ((m, t2, t0, -1), 3),
                   6: ((m, t0, t1, t2), 4)}[reg]
            out.append(enc[0])
            cls.append(enc[1])
    ex = x[geo.edges]
    elo, ehi = ex.min(1), ex.max(1)
    for e in range(geo.edges.shape[0]):
        a0, a1 = (int(v) for v in geo.edges[e])
        sep = (elo - ehi[e] >= gap).any(1) | (elo[e] - ehi >= gap).any(1)
        for o in np.flatnonzero(~sep):
            if o <= e:
                continue
            b0, b1 = (int(v) for v in geo.edges[o])
            if geo.body_id is not None and geo.body_id[a0] == geo.body_id[b0] and geo.body_id[a0] != -1:
                continue
            if len({a0, a1, b0, b1}) < 4:
                continue
            s, t = _closest_seg_seg(x[a0], x[a1], x[b0], x[b1])
            d = x[a0] + s * (x[a1] - x[a0]) - x[b0] - t * (x[b1] - x[b0])
            if d @ d >= dsq:
                continue
            cr = np.cross(x[a0] - x[a1], x[b0] - x[b1])
            ex_ = 1e-3 * ((rest[a0] - rest[a1]) @ (rest[a0] - rest[a1])) * ((rest[b0] - rest[b1]) @ (rest[b0] - rest[b1]))
            w = -int(o) - 2 if cr @ cr < ex_ else -1
            se = 0 if s == 0 else (1 if s == 1 else 2)
            te = 0 if t == 0 else (1 if t == 1 else 2)
            enc = {(0, 0): ((-a0 - 1, b0, -1, w), 2), (0, 1): ((-a0 - 1, b1, -1, w), 2),
                   (0, 2): ((-a0 - 1, b0, b1, w), 3), (1, 0): ((-a1 - 1, b0, -1, w), 2),
                   (1, 1): ((-a1 - 1, b1, -1, w), 2), (1, 2): ((-a1 - 1, b0, b1, w), 3),
                   (2, 0): ((-b0 - 1, a0, a1, w), 3), (2, 1): ((-b1 - 1, a0, a1, w), 3),
                   (2, 2): ((a0, a1, b0, b1), 4)}[(se, te)]
            out.append(enc[0])
            cls.append(enc[1])
    return np.array(out, dtype=np.int32).reshape(-1, 4), np.array(cls)


@pytest.mark.parametrize("N,bodies", [(10, False), (16, False), (12, True)])
def test_detection_matches_exact_closest_features(pkg, orc, N, bodies):
    geo = pkg.scenes.contact_sheets(N, body_ids=bodies)
    pairs, mi, counts = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2,
                                           body_id=geo.body_id)
    want, cls = restated_detection(geo)
    assert pairs.shape[0] > 0
    np.testing.assert_array_equal(pairs, want)
    # every class occurs, MatIndex numbers each class in order
    for c in (2, 3, 4):
        assert counts[c] == (cls == c).sum() > 0
        np.testing.assert_array_equal(mi[cls == c], np.arange(counts[c]))
    assert counts[0] == pairs.shape[0] == counts[2] + counts[3] + counts[4]


def test_detection_capacity_reports_total(pkg, orc):
    geo = pkg.scenes.contact_sheets(8)
    full, _, c = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2)
    _, _, c2 = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2,
                                  capacity=full.shape[0] - 1)
    assert c2[0] == full.shape[0]


def _pair_kind(pr):
    if pr[0] >= 0:
        return "EE"
    return "PP" if pr[2] < 0 else ("PE" if pr[3] < 0 else "PT")


def _vids(pr):
    if pr[0] >= 0:
        return list(pr)
    v = [-pr[0] - 1, pr[1]]
    if pr[2] >= 0:
        v.append(pr[2])
        if pr[3] >= 0:
            v.append(pr[3])
    return v


def _pairs_by_kind(pkg, orc):
    geo = pkg.scenes.contact_sheets(10)
    pairs, mi, counts = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2)
    pick = {}
    for k, pr in enumerate(pairs):
        pick.setdefault(_pair_kind(pr), []).append(k)
    return geo, pairs, mi, counts, pick


def _fd_grad(f, x, vids, h):
    g = np.zeros((len(vids), 3))
    for a, v in enumerate(vids):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[v, d] += h
            xm[v, d] -= h
            g[a, d] = (f(xp) - f(xm)) / (2 * h)
    return g.ravel()


def test_barrier_gradient_matches_fd_for_every_pair_type(pkg, orc):
    geo, pairs, mi, counts, pick = _pairs_by_kind(pkg, orc)
    dsq, kappa = geo.dhat ** 2, geo.kappa
    assert set(pick) == {"PP", "PE", "PT", "EE"}
    for kind, ks in pick.items():
        for k in ks[:4]:
            one = pairs[k:k + 1]
            c1 = np.zeros(5, np.int32)
            c1[{"PP": 2, "PE": 3, "PT": 4, "EE": 4}[kind]] = 1
            _, _, _, g = orc.barrier(geo.x, one, np.zeros(1, np.int32), c1, dsq, kappa)
            vids = _vids(one[0])
            fd = _fd_grad(lambda xx: orc.barrier_energy(xx, one, dsq, kappa), geo.x, vids, 1e-9)
            got = g.reshape(-1, 3)[vids].ravel()
            np.testing.assert_allclose(got, fd, rtol=2e-5, atol=1e-6 * np.abs(fd).max(), err_msg=kind)


def _s_of(xx, one, dsq, vids):
    # s = dist / sqrt(dHat) from the energy's own distance: E = kappa dHat^2 (I5-1)^2 ln(I5)^2 is monotone in
    # I5 on (0, 1), so I5 is recovered from the distance functions the tests restate below
    pr = one[0]
    P = xx[vids]
    kind = _pair_kind(pr)
    if kind == "PP":
        d = np.linalg.norm(P[0] - P[1])
    elif kind == "PE":
        e = P[2] - P[1]
        d = np.linalg.norm(np.cross(P[0] - P[1], e)) / np.linalg.norm(e)
    elif kind == "PT":
        nv = np.cross(P[2] - P[1], P[3] - P[1])
        d = abs((P[0] - P[1]) @ nv) / np.linalg.norm(nv)
    else:
        nv = np.cross(P[1] - P[0], P[3] - P[2])
        d = abs((P[2] - P[0]) @ nv) / np.linalg.norm(nv)
    return d / np.sqrt(dsq)


def test_barrier_hessian_is_lambda0_grad_s_outer(pkg, orc):
    geo, pairs, mi, counts, pick = _pairs_by_kind(pkg, orc)
    dsq, kappa = geo.dhat ** 2, geo.kappa
    for kind, ks in pick.items():
        for k in ks[:3]:
            one = pairs[k:k + 1]
            c1 = np.zeros(5, np.int32)
            c1[{"PP": 2, "PE": 3, "PT": 4, "EE": 4}[kind]] = 1
            v, r, c, _ = orc.barrier(geo.x, one, np.zeros(1, np.int32), c1, dsq, kappa)
            vids = _vids(one[0])
            nv = len(vids)
            gs = _fd_grad(lambda xx: _s_of(xx, one, dsq, vids), geo.x, vids, 1e-10)
            s0 = _s_of(geo.x, one, dsq, vids)
            h = 1e-5 * s0
            E = lambda s: orc.lib().orc_barrier_energy(s * s, dsq, kappa)
            lam0 = (E(s0 + h) - 2 * E(s0) + E(s0 - h)) / h ** 2
            want = lam0 * np.outer(gs, gs)
            # dense 3nv x 3nv from the upper blocks (lower ones are zeroed)
            H = np.zeros((3 * nv, 3 * nv))
            pos = {vv: a for a, vv in enumerate(vids)}
            for blk, rr, cc in zip(v, r, c):
                M = blk.reshape(3, 3).T
                i, j = pos[rr], pos[cc]
                H[3 * i:3 * i + 3, 3 * j:3 * j + 3] += M
                if i != j:
                    H[3 * j:3 * j + 3, 3 * i:3 * i + 3] += M.T
            assert r.size == nv * nv
            np.testing.assert_allclose(H, want, rtol=1e-3, atol=1e-4 * np.abs(want).max(), err_msg=kind)


def test_barrier_arena_layout(pkg, orc):
    geo, pairs, mi, counts, _ = _pairs_by_kind(pkg, orc)
    v, r, c, _ = orc.barrier(geo.x, pairs, mi, counts, geo.dhat ** 2, geo.kappa)
    n4, n3, n2 = counts[4], counts[3], counts[2]
    assert r.size == 16 * n4 + 9 * n3 + 4 * n2
    assert (r <= c).all()
    for k, pr in enumerate(pairs):
        vids = _vids(pr)
        nv = len(vids)
        base = 16 * mi[k] if nv == 4 else (16 * n4 + 9 * mi[k] if nv == 3 else 16 * n4 + 9 * n3 + 4 * mi[k])
        got = set(zip(r[base:base + nv * nv], c[base:base + nv * nv]))
        assert got == {(min(a, b), max(a, b)) for a in vids for b in vids}


def test_planted_point_edge_weight_is_caught(pkg, orc):
    geo, pairs, mi, counts, pick = _pairs_by_kind(pkg, orc)
    k = pick["PE"][0]
    one = pairs[k:k + 1]
    c1 = np.array([1, 0, 0, 1, 0], np.int32)
    vids = _vids(one[0])
    dsq, kappa = geo.dhat ** 2, geo.kappa
    fd = _fd_grad(lambda xx: orc.barrier_energy(xx, one, dsq, kappa), geo.x, vids, 1e-9)
    orc.set_fault(5)
    try:
        _, _, _, g = orc.barrier(geo.x, one, np.zeros(1, np.int32), c1, dsq, kappa)
    finally:
        orc.set_fault(0)
    got = g.reshape(-1, 3)[vids].ravel()
    assert not np.allclose(got, fd, rtol=2e-5, atol=1e-6 * np.abs(fd).max())
