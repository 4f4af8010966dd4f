"""Pins of the oracle's lagged friction (oracle/friction_oracle.c), independent
of its formulas:

* normal force: lambda = -2 sqrt(d) dB/dd with B the barrier energy, by
  central differences in d (pairs: the RANK 2 barrier the barrier oracle
  uses; ground: -kappa (d - dHat)^2 ln(d / dHat));
* closest-point coordinates: the gap vector between the closest points is
  orthogonal to the pair's edges / triangle (the defining property);
* tangent basis: orthonormal and orthogonal to the contact normal;
* gradient: central differences of the friction energy mu lambda f0(|u|);
* Hessian: central differences of that gradient (the lagged quantities
  fixed), in both regimes of the static-friction clamp (|u| < eps, > eps).
"""
import numpy as np
import pytest


def _kind(pr):
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


@pytest.fixture(scope="module")
def setup(pkg, orc):
    geo = pkg.scenes.contact_sheets(10)
    pairs, mi, counts = orc.contact_detect(geo.x, geo.rest, geo.faces, geo.edges, geo.surf, geo.dhat ** 2)
    lam, coord, basis = orc.friction_prepare(geo.x, pairs, geo.dhat ** 2, geo.kappa)
    rng = np.random.default_rng(11)
    x = geo.x + rng.uniform(-2e-6, 2e-6, geo.x.shape)
    return geo, pairs, mi, counts, lam, coord, basis, x


def _d2(kind, P):
    if kind == "PP":
        return (P[0] - P[1]) @ (P[0] - P[1])
    if kind == "PE":
        c = np.cross(P[1] - P[0], P[2] - P[0])
        return (c @ c) / ((P[2] - P[1]) @ (P[2] - P[1]))
    if kind == "PT":
        nv = np.cross(P[2] - P[1], P[3] - P[1])
        return ((P[0] - P[1]) @ nv) ** 2 / (nv @ nv)
    nv = np.cross(P[1] - P[0], P[3] - P[2])
    return ((P[2] - P[0]) @ nv) ** 2 / (nv @ nv)


def test_normal_force_is_barrier_derivative(orc, setup):
    geo, pairs, _, _, lam, _, _, _ = setup
    dsq, kappa = geo.dhat ** 2, geo.kappa
    E = lambda d: orc.lib().orc_barrier_energy(d / dsq, dsq, kappa)
    for k, pr in enumerate(pairs[:200]):
        d = _d2(_kind(pr), geo.x[_vids(pr)])
        h = 1e-6 * d
        want = -2.0 * np.sqrt(d) * (E(d + h) - E(d - h)) / (2 * h)
        assert lam[k] == pytest.approx(want, rel=1e-6)
        assert lam[k] > 0


def test_closest_points_and_tangent_basis(setup):
    geo, pairs, _, _, _, coord, basis, _ = setup
    seen = set()
    for k, pr in enumerate(pairs):
        kind = _kind(pr)
        seen.add(kind)
        P = geo.x[_vids(pr)]
        c = coord[k]
        B = basis[k].reshape(2, 3)
        np.testing.assert_allclose(B @ B.T, np.eye(2), atol=1e-12)
        if kind == "PT":
            e1, e2 = P[2] - P[1], P[3] - P[1]
            gapv = P[0] - (P[1] + c[0] * e1 + c[1] * e2)
            assert abs(gapv @ e1) <= 1e-12 * np.linalg.norm(e1) * np.linalg.norm(gapv) + 1e-30
            assert abs(gapv @ e2) <= 1e-12 * np.linalg.norm(e2) * np.linalg.norm(gapv) + 1e-30
            normal = np.cross(e1, e2)
        elif kind == "EE":
            ea, eb = P[1] - P[0], P[3] - P[2]
            gapv = (P[0] + c[0] * ea) - (P[2] + c[1] * eb)
            scale = np.linalg.norm(gapv) + 1e-30
            assert abs(gapv @ ea) <= 1e-9 * np.linalg.norm(ea) * scale
            assert abs(gapv @ eb) <= 1e-9 * np.linalg.norm(eb) * scale
            normal = np.cross(ea, eb)
        elif kind == "PE":
            e = P[2] - P[1]
            gapv = P[0] - (P[1] + c[0] * e)
            assert abs(gapv @ e) <= 1e-12 * np.linalg.norm(e) * np.linalg.norm(gapv) + 1e-30
            normal = gapv
        else:
            normal = P[1] - P[0]
        nrm = normal / np.linalg.norm(normal)
        assert np.abs(B @ nrm).max() <= 1e-9
    assert seen == {"PP", "PE", "PT", "EE"}


def _fd(f, x, vids, h):
    out = []
    for v in vids:
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[v, d] += h
            xm[v, d] -= h
            out.append((f(xp) - f(xm)) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("eps", [5e-7, 2e-5])
def test_gradient_and_hessian_match_fd(orc, setup, eps):
    """eps = 5e-7: most pairs slide (|u| > eps); 2e-5: most stick."""
    geo, pairs, mi, counts, lam, coord, basis, x = setup
    mu = 0.4
    pick = {}
    for k, pr in enumerate(pairs):
        pick.setdefault(_kind(pr), []).append(k)
    for kind, ks in pick.items():
        for k in ks[:3]:
            one = pairs[k:k + 1]
            c1 = np.zeros(5, np.int32)
            c1[{"PP": 2, "PE": 3, "PT": 4, "EE": 4}[kind]] = 1
            args = (lam[k:k + 1], coord[k:k + 1], basis[k:k + 1], eps * eps, mu)
            v, r, c, g = orc.friction(x, geo.x, one, np.zeros(1, np.int32), c1, *args)
            vids = _vids(one[0])
            fd = _fd(lambda xx: orc.friction_energy(xx, geo.x, one, *args), x, vids, 1e-10)
            got = g.reshape(-1, 3)[vids].ravel()
            np.testing.assert_allclose(got, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max(), err_msg=kind)
            # Hessian: columns are differences of the gradient
            nv = len(vids)
            H = np.zeros((3 * nv, 3 * nv))
            pos = {vv: a for a, vv in enumerate(vids)}
            for blk, rr, cc in zip(v, r, c):
                M = blk.reshape(3, 3).T
                i, j = pos[rr], pos[cc]
                H[3 * i:3 * i + 3, 3 * j:3 * j + 3] += M
                if i != j:
                    H[3 * j:3 * j + 3, 3 * i:3 * i + 3] += M.T
            Hfd = np.zeros_like(H)
            h = 1e-11
            for a, vv in enumerate(vids):
                for d in range(3):
                    xp, xm = x.copy(), x.copy()
                    xp[vv, d] += h
                    xm[vv, d] -= h
                    gp = orc.friction(xp, geo.x, one, np.zeros(1, np.int32), c1, *args)[3].reshape(-1, 3)[vids].ravel()
                    gm = orc.friction(xm, geo.x, one, np.zeros(1, np.int32), c1, *args)[3].reshape(-1, 3)[vids].ravel()
                    Hfd[:, 3 * a + d] = (gp - gm) / (2 * h)
            np.testing.assert_allclose(H, Hfd, rtol=1e-4, atol=1e-4 * np.abs(Hfd).max(), err_msg=kind)


@pytest.mark.parametrize("normal", [(0.0, 1.0, 0.0), (0.3, 0.9, 0.1)])
def test_ground_friction_matches_fd(orc, normal):
    rng = np.random.default_rng(4)
    n = 30
    nrm = np.array(normal) / np.linalg.norm(normal)
    dsq, kappa, mu, offset = 1e-6, 1e4, 0.5, -0.2
    xt = rng.uniform(-1, 1, (n, 3))
    # put every vertex within sqrt(dHat) above the plane n.x = offset
    xt += ((offset + rng.uniform(0.1, 0.9, n) * 1e-3) - xt @ nrm)[:, None] * nrm
    verts = np.arange(n, dtype=np.int32)
    lam = orc.ground_friction_prepare(xt, verts, nrm, offset, dsq, kappa)
    B0 = lambda d2: -kappa * (d2 - dsq) ** 2 * np.log(d2 / dsq)
    for i in range(5):
        d2 = (xt[i] @ nrm - offset) ** 2
        h = 1e-6 * d2
        assert lam[i] == pytest.approx(-2 * np.sqrt(d2) * (B0(d2 + h) - B0(d2 - h)) / (2 * h), rel=1e-6)
    for eps in (1e-7, 1e-4):
        x = xt + rng.uniform(-2e-6, 2e-6, xt.shape)
        v, r, c, g = orc.ground_friction(x, xt, verts, nrm, lam, eps * eps, mu)
        for i in range(n):
            f = lambda xx: orc.ground_friction_energy(xx, xt, verts, nrm, lam, eps * eps, mu)
            fd = _fd(f, x, [i], 1e-9 if eps > 1e-6 else 1e-11)
            np.testing.assert_allclose(g.reshape(-1, 3)[i], fd, rtol=1e-4, atol=1e-5 * np.abs(fd).max())
            assert r[i] == c[i] == i
            Hfd = np.zeros((3, 3))
            h = 1e-11
            for d in range(3):
                xp, xm = x.copy(), x.copy()
                xp[i, d] += h
                xm[i, d] -= h
                Hfd[:, d] = (orc.ground_friction(xp, xt, verts, nrm, lam, eps * eps, mu)[3].reshape(-1, 3)[i] -
                             orc.ground_friction(xm, xt, verts, nrm, lam, eps * eps, mu)[3].reshape(-1, 3)[i]) / (2 * h)
            np.testing.assert_allclose(v[i].reshape(3, 3).T, Hfd, rtol=1e-4, atol=1e-4 * np.abs(Hfd).max())
