"""Pins the oracle's conversion, SpMV, MAS and PCG against independent
restatements (numpy lexsort/reduceat, scipy.sparse, scipy's CG, and a
set-based numpy/scipy MAS built from the reference's rules), plus planted
faults that those checks must catch."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as csg
import scipy.sparse.linalg as spla


def bcoo_to_csr(vals, rows, cols, n):
    """Full symmetric scipy matrix from upper 3x3 blocks (column-major)."""
    blocks = vals.reshape(-1, 3, 3).transpose(0, 2, 1)  # -> row-major M
    r = (3 * rows[:, None, None] + np.arange(3)[None, :, None]).repeat(3, axis=2)
    c = (3 * cols[:, None, None] + np.arange(3)[None, None, :]).repeat(3, axis=1)
    U = sp.coo_matrix((blocks.reshape(-1), (r.reshape(-1), c.reshape(-1))), shape=(3 * n, 3 * n)).tocsr()
    D = sp.coo_matrix(
        (blocks[rows == cols].reshape(-1),
         (r[rows == cols].reshape(-1), c[rows == cols].reshape(-1))), shape=(3 * n, 3 * n)).tocsr()
    return (U + U.T - D).tocsr()


# ----------------------------------------------------------------- convert
def test_convert_matches_lexsort_reduceat(orc):
    rng = np.random.default_rng(0)
    m, n = 5000, 300
    r = rng.integers(0, n, m).astype(np.int32)
    c = rng.integers(0, n, m).astype(np.int32)
    r, c = np.minimum(r, c), np.maximum(r, c)
    v = rng.standard_normal((m, 9))
    ov, orr, oc = orc.convert(v, r, c)
    order = np.lexsort((c, r))
    key = r[order].astype(np.int64) * n + c[order]
    heads = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    np.testing.assert_array_equal(orr, r[order][heads])
    np.testing.assert_array_equal(oc, c[order][heads])
    np.testing.assert_allclose(ov, np.add.reduceat(v[order], heads, axis=0), rtol=1e-15, atol=1e-15)


def test_assembly_counts(orc, scene_c0):
    s = scene_c0
    vals, rows, cols = orc.assemble(s)
    # one block per vertex + one per unique tet edge (Kuhn grid: 25,695 at 16^3)
    assert rows.size == s.n + s.nbr_list.size // 2 == 29791
    assert np.all(rows <= cols)
    key = rows.astype(np.int64) * s.n + cols
    assert np.all(np.diff(key) > 0)


# -------------------------------------------------------------------- SpMV
@pytest.mark.parametrize("which", ["tiny", "small"])
def test_spmv_matches_scipy(orc, which, scene_tiny, scene_small):
    s = {"tiny": scene_tiny, "small": scene_small}[which]
    vals, rows, cols = orc.assemble(s)
    A = bcoo_to_csr(vals, rows, cols, s.n)
    rng = np.random.default_rng(2)
    x, y0 = rng.standard_normal(3 * s.n), rng.standard_normal(3 * s.n)
    got = orc.spmv(0.7, vals, rows, cols, x, -1.3, y0, s.n)
    ref = 0.7 * (A @ x) - 1.3 * y0
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    # the assembled system is symmetric positive definite
    if s.n <= 300:
        w = np.linalg.eigvalsh(A.toarray())
        assert w.min() > 0


def test_planted_spmv_fault_is_caught(orc, scene_tiny):
    s = scene_tiny
    vals, rows, cols = orc.assemble(s)
    A = bcoo_to_csr(vals, rows, cols, s.n)
    x = np.random.default_rng(4).standard_normal(3 * s.n)
    orc.set_fault(1)
    try:
        got = orc.spmv(1.0, vals, rows, cols, x, 0.0, np.zeros(3 * s.n), s.n)
    finally:
        orc.set_fault(0)
    assert np.max(np.abs(got - A @ x)) > 1e-3 * np.max(np.abs(A @ x))


# --------------------------------------------------------------------- MAS
def num_levels(n, B):
    """computeNumLevels, MASPreconditioner.cu:1731-1750."""
    nl, sz = 1, (n + B - 1) // B * B
    while sz > B:
        sz //= B
        nl += 1
        sz = (sz + B - 1) // B * B
    return min(nl + 1, 6)


def decode_pair(pr):
    """Vertex ids of a contact pair in the reference int4 encoding."""
    x, y, z, w = (int(t) for t in pr)
    if x >= 0:
        return [x, y, z, -w - 1 if w < 0 else w]
    x = -x - 1
    if z < 0:
        if y < 0:
            return [x, -y - 1, -z - 1, -w - 1]
        return [x, y]
    if w < 0:
        if y < 0:
            return [x, -y - 1, z, -w - 1]
        return [x, y, z]
    return [x, y, z, w]


def mas_reference(s, vals, rows, cols, B):
    """Set-based restatement of the MAS hierarchy and Galerkin operators.

    Level l+1 nodes = connected components, inside each level-l domain of B
    consecutive ids, of the graph of mesh edges and contact cliques mapped to
    level l; numbered domain by domain, then by lowest member.  Returns the
    per-level maps (vertex -> local id), counts, offsets and the dense domain
    blocks of P_l^T A P_l."""
    n = s.n
    L = num_levels(n, B)
    src = np.repeat(np.arange(n), s.nbr_num.astype(np.int64))
    dst = s.nbr_list.astype(np.int64)
    cp = [decode_pair(p) for p in s.pairs]
    cedges = [(a, b) for vs in cp for i, a in enumerate(vs) for b in vs[i + 1:]]
    ce = np.array(cedges, np.int64).reshape(-1, 2)
    maps = [np.arange(n)]
    counts = [n]
    for l in range(L):
        m = maps[-1]
        c = counts[-1]
        a = np.concatenate([m[src], m[ce[:, 0]], m[ce[:, 1]]]) if ce.size else m[src]
        b = np.concatenate([m[dst], m[ce[:, 1]], m[ce[:, 0]]]) if ce.size else m[dst]
        keep = (a != b) & (a // B == b // B)
        G = sp.coo_matrix((np.ones(keep.sum()), (a[keep], b[keep])), shape=(c, c))
        _, lab = csg.connected_components(G, directed=False)
        # renumber: components in order of their lowest member (domains are contiguous)
        first = {}
        new = np.empty(c, np.int64)
        nxt = 0
        for node in range(c):
            comp = lab[node]
            if comp not in first:
                first[comp] = nxt
                nxt += 1
            new[node] = first[comp]
        if l + 1 < L:
            maps.append(new[m])
        counts.append(nxt)
    offsets = [0]
    for l in range(L):
        offsets.append(offsets[-1] + (counts[l] + B - 1) // B * B)
    A = bcoo_to_csr(vals, rows, cols, n)
    Aabs = abs(A)
    dom_blocks = []  # dense (3B x 3B) Galerkin matrix per domain, all levels
    dom_abs = []     # P^T |A| P: the magnitude the sums are formed from
    for l in range(L):
        P = sp.coo_matrix((np.ones(n), (np.arange(n), maps[l])), shape=(n, counts[l]))
        P3 = sp.kron(P, sp.eye(3)).tocsr()
        Al = (P3.T @ A @ P3).toarray()
        Ab = (P3.T @ Aabs @ P3).toarray()
        ndom = (offsets[l + 1] - offsets[l]) // B
        for d in range(ndom):
            M = np.zeros((3 * B, 3 * B))
            Mb = np.zeros((3 * B, 3 * B))
            lo, hi = d * B, min((d + 1) * B, counts[l])
            M[:3 * (hi - lo), :3 * (hi - lo)] = Al[3 * lo:3 * hi, 3 * lo:3 * hi]
            Mb[:3 * (hi - lo), :3 * (hi - lo)] = Ab[3 * lo:3 * hi, 3 * lo:3 * hi]
            dom_blocks.append(M)
            dom_abs.append(Mb)
    return maps, counts[:L], offsets, dom_blocks, dom_abs


def blocks_to_dense(ub, B):
    M = np.zeros((3 * B, 3 * B))
    k = 0
    for r in range(B):
        for c in range(r, B):
            blk = ub[k].reshape(3, 3).T
            M[3 * r:3 * r + 3, 3 * c:3 * c + 3] = blk
            if c != r:
                M[3 * c:3 * c + 3, 3 * r:3 * r + 3] = blk.T
            k += 1
    return M


@pytest.mark.parametrize("B", [16, 32])
@pytest.mark.parametrize("which", ["tiny", "small"])
def test_mas_matches_set_restatement(orc, B, which, scene_tiny, scene_small):
    s = {"tiny": scene_tiny, "small": scene_small}[which]
    vals, rows, cols = orc.assemble(s)
    m = orc.Mas(s, B)
    m.build(vals, rows, cols, 0, s.pairs)
    L, off, cnt, total = m.levels()
    maps, counts, offsets, dom, dom_abs = mas_reference(s, vals, rows, cols, B)
    assert L == len(maps)
    np.testing.assert_array_equal(cnt, counts)
    np.testing.assert_array_equal(off, offsets)
    ct, _ = m.tables()
    for l in range(1, L):
        np.testing.assert_array_equal(ct[:, l - 1], offsets[l] + maps[l])
    gal = m.matrices(inverse=False)
    inv = m.matrices(inverse=True)
    assert gal.shape[0] == len(dom)
    for d, Md in enumerate(dom):
        G = blocks_to_dense(gal[d], B)
        # entrywise rounding bound of a k-term sum: (k-1) u sum|terms|, with k
        # up to a few hundred terms per coarse entry
        assert np.all(np.abs(G - Md) <= 1e-12 * dom_abs[d] + 1e-300), d
        # invert the oracle's own Galerkin block: on the coarsest levels the
        # stiffness sums cancel to the mass, so G carries rounding relative
        # to M_d that is not the inversion's doing
        Mi = G.copy()
        dz = np.diag(Mi) == 0
        Mi[dz, dz] = 1.0
        ref = np.linalg.inv(Mi)  # LU with pivoting, an independent method
        got = blocks_to_dense(inv[d], B)
        # both inverses are accurate to ~eps * cond(M); stiff domains are ill-conditioned
        tol = max(1e-10, 50 * np.finfo(float).eps * np.linalg.cond(Mi))
        assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref)), (d, tol)
    m.close()


def test_mas_apply_is_sum_of_level_solves_and_spd(orc, scene_tiny):
    s = scene_tiny
    B = 16
    vals, rows, cols = orc.assemble(s)
    m = orc.Mas(s, B)
    m.build(vals, rows, cols, 0, s.pairs)
    L, off, cnt, total = m.levels()
    ct, _ = m.tables()
    inv = m.matrices(inverse=True)
    n = s.n
    r = np.random.default_rng(8).standard_normal(3 * n)
    z = np.zeros(3 * n)
    slot = np.concatenate([np.arange(n)[:, None], ct], axis=1)  # global slot per level
    for l in range(L):
        R = np.zeros(3 * total)
        for v in range(n):
            R[3 * slot[v, l]:3 * slot[v, l] + 3] += r[3 * v:3 * v + 3]
        Z = np.zeros(3 * total)
        for d in range(off[l] // B, off[l + 1] // B):
            Z[3 * B * d:3 * B * (d + 1)] = blocks_to_dense(inv[d], B) @ R[3 * B * d:3 * B * (d + 1)]
        for v in range(n):
            z[3 * v:3 * v + 3] += Z[3 * slot[v, l]:3 * slot[v, l] + 3]
    got = m.apply(r)
    assert np.max(np.abs(got - z)) <= 1e-12 * np.max(np.abs(z))
    # M^-1 as a dense operator: symmetric positive definite
    Minv = np.stack([m.apply(e) for e in np.eye(3 * n)], axis=1)
    assert np.max(np.abs(Minv - Minv.T)) <= 1e-9 * np.max(np.abs(Minv))
    assert np.linalg.eigvalsh(0.5 * (Minv + Minv.T)).min() > 0
    m.close()


def test_planted_galerkin_fault_is_caught(orc, scene_small):
    s = scene_small
    B = 16
    vals, rows, cols = orc.assemble(s)
    _, _, _, dom, _ = mas_reference(s, vals, rows, cols, B)
    orc.set_fault(2)
    try:
        m = orc.Mas(s, B)
        m.build(vals, rows, cols, 0, s.pairs)
        gal = m.matrices(inverse=False)
    finally:
        orc.set_fault(0)
    worst = max(np.max(np.abs(blocks_to_dense(gal[d], B) - Md)) / max(np.max(np.abs(Md)), 1e-300)
                for d, Md in enumerate(dom))
    assert worst > 1e-3
    m.close()


# --------------------------------------------------------------------- PCG
def test_pcg_iterates_match_scipy_cg(orc, scene_small):
    s = scene_small
    vals, rows, cols = orc.assemble(s)
    A = bcoo_to_csr(vals, rows, cols, s.n)
    m = orc.Mas(s, 32)
    m.build(vals, rows, cols, 0, s.pairs)
    M = spla.LinearOperator(A.shape, matvec=lambda v: m.apply(np.asarray(v).reshape(-1)))
    for k in (3, 10, 25):
        xr, it, _ = orc.pcg(vals, rows, cols, s.n, s.rhs, m, tol=1e-30, stop_mode=1, max_iter=k + 1)
        assert it == k
        xs, _ = spla.cg(A, s.rhs, x0=np.zeros_like(s.rhs), rtol=1e-30, atol=0.0, maxiter=k, M=M)
        assert np.linalg.norm(xr - xs) <= 1e-8 * np.linalg.norm(xs), k
    x, it, res = orc.pcg(vals, rows, cols, s.n, s.rhs, m, tol=1e-5, stop_mode=1)
    assert np.linalg.norm(s.rhs - A @ x) / np.linalg.norm(s.rhs) <= 1e-5 + 1e-12
    assert abs(res - np.linalg.norm(s.rhs - A @ x) / np.linalg.norm(s.rhs)) <= 1e-12
    m.close()


def test_pcg_reference_stop_rule(orc, scene_c0):
    """|r.z| <= tol r0.z0, checked with the rz that formed alpha: the reference
    rule stops long before the true residual reaches the same tolerance."""
    s = scene_c0
    vals, rows, cols = orc.assemble(s)
    m = orc.Mas(s, 32)
    m.build(vals, rows, cols, 0, s.pairs)
    _, it_ref, res_ref = orc.pcg(vals, rows, cols, s.n, s.rhs, m, tol=1e-4, stop_mode=0)
    _, it_true, res_true = orc.pcg(vals, rows, cols, s.n, s.rhs, m, tol=1e-4, stop_mode=1)
    assert it_ref < it_true and res_ref > res_true
    m.close()


def test_convert_long_blocks_in_chunks(orc):
    """A block merging more than 64 triplets is the ordered sum of its
    64-triplet chunk sums (the library's fixed order), each chunk summed in
    sorted (= arena) order: checked bit for bit against that order written
    out, and to rounding against a plain sum."""
    rng = np.random.default_rng(1)
    m = 1000
    r = np.concatenate([np.zeros(m, np.int32), np.ones(10, np.int32)])
    c = np.concatenate([np.full(m, 2, np.int32), np.full(10, 3, np.int32)])
    v = rng.standard_normal((r.size, 9)) * 10.0 ** rng.uniform(-8, 8, (r.size, 1))
    ov, orr, oc = orc.convert(v, r, c)
    assert list(orr) == [0, 1] and list(oc) == [2, 3]
    want = np.zeros(9)
    for k in range(0, m, 64):
        part = np.zeros(9)
        for row in v[k:min(k + 64, m)]:
            part = part + row
        want = want + part
    np.testing.assert_array_equal(ov[0], want)
    short = np.zeros(9)
    for row in v[m:]:
        short = short + row
    np.testing.assert_array_equal(ov[1], short)
    np.testing.assert_allclose(ov[0], v[:m].sum(axis=0), rtol=1e-6, atol=1e-6 * np.abs(v[:m]).sum(axis=0).max())
