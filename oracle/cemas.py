"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY.

CEMAS (partition-driven MAS) restated from the reference in operator form,
independent of the product library and of the identity-map C oracle:

* slot maps: ``set_mas_partition`` (app/common/sim_bootstrap.cpp:90-113),
  its two loops as written (``partition_maps_ref``), valid for part ids in
  non-decreasing order, the order MeshProcess writes;
* hierarchy: level-0 nodes are the slots of the real vertices
  (``real_map_partId``), domains the banks of B slots.  Mesh edges
  (``_buildCML0_new``, MASPreconditioner.cu:63-100) and contact cliques
  (``_buildCollisionConnection_new``, :1132-1548, pair encodings decoded as
  there) join nodes of one domain; each domain's connected components
  (``_preparePrefixSumL0_new`` :155-210) become the level-1 nodes, numbered
  domain by domain and by lowest lane inside a domain (``_buildLevel1_new``
  :270-325).  Levels >= 1 repeat this over the level-l node ids with domains
  of B consecutive ids; the level count is ``computeNumLevels``
  (:1731-1750) of the slot count;
* operator: the level-l domain matrices are the Galerkin blocks
  P_l^T A P_l restricted to node pairs of one domain (``PrepareHessian_bcoo``
  :1829-2106: level 0 through real_map_partId, coarse levels through
  goingNext, H^T for lower-triangle positions and H + H^T on merged
  diagonals), inverted with zero diagonals set to 1 (``__inverse6_P96x96``,
  :630-726), and the preconditioner is z = sum_l P_l D_l^-1 P_l^T r
  (``preconditioning`` :2248-2289 with ``__buildMultiLevelR_optimized_new``
  :734-850 and ``__collectFinalZ_new`` :853-880, in FP64).

The pins live in tests/test_oracle_cemas.py: with the partition
part = vertex // B the operator equals the identity-map C oracle's MAS
apply; it is symmetric positive definite; relabelling the vertices and the
parts together permutes it; a one-domain system gets the exact inverse.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def partition_maps_ref(part_id, bank):
    """The reference's two loops of set_mas_partition, as written."""
    part = [int(p) for p in part_id]
    n = len(part)
    part_offset = part[-1] + 1
    p2r = [-1] * (part_offset * bank)
    r2p = [0] * n
    index = 0
    for i in range(n):
        p2r[bank * part[i] + index] = i
        index += 1
        if i <= n - 2 and part[i + 1] != part[i]:
            index = 0
    index = 0
    for i in range(len(p2r)):
        if p2r[i] == index:
            r2p[index] = i
            index += 1
    return np.array(p2r, dtype=np.int64), np.array(r2p, dtype=np.int64)


def num_levels(n, B, max_levels=6):
    """computeNumLevels (MASPreconditioner.cu:1731-1750)."""
    nl, sz = 1, (n + B - 1) // B * B
    while sz > B:
        sz //= B
        nl += 1
        sz = (sz + B - 1) // B * B
    return min(nl + 1, max_levels)


def pair_vertices(p):
    """Vertex ids of one contact pair in the reference's int4 encodings
    (MASPreconditioner.cu:1145-1548)."""
    x, y, z, w = (int(v) for v in p)
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


def _components(ids, edges, B):
    """Next-level ids of nodes 0..len(ids)-1 grouped by domain id // B over
    the edges joining nodes of one domain: components numbered domain by
    domain, by lowest member inside a domain."""
    m = int(ids.max()) + 1 if ids.size else 0
    parent = np.arange(m)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for a, b in edges:
        a, b = ids[a], ids[b]
        if a != b and a // B == b // B:
            ra, rb = find(a), find(b)
            if ra != rb:
                parent[max(ra, rb)] = min(ra, rb)
    present = np.zeros(m, bool)
    present[ids] = True
    root = np.array([find(a) for a in range(m)])
    nxt = -np.ones(m, np.int64)
    k = 0
    for a in range(m):  # ascending: domains in order, lowest member first
        if present[a] and root[a] == a:
            nxt[a] = k
            k += 1
    return nxt[root[ids]], k


def hierarchy(n, nbr, r2p, map_size, B, pairs=None):
    """Level-l node of every vertex, l = 0..L-1 (level 0 = slots), and the node counts."""
    start, num, lst = (np.asarray(a, dtype=np.int64) for a in nbr)
    src = np.repeat(np.arange(n), num)
    dst = lst[np.concatenate([np.arange(s, s + c) for s, c in zip(start, num)])] if num.sum() else np.zeros(0, np.int64)
    edges = list(zip(src.tolist(), dst.tolist()))
    if pairs is not None:
        for p in np.asarray(pairs).reshape(-1, 4):
            vs = [v for v in pair_vertices(p) if 0 <= v < n]
            edges += [(vs[i], vs[j]) for i in range(len(vs)) for j in range(i + 1, len(vs))]
    L = num_levels(map_size, B)
    node = [np.asarray(r2p, dtype=np.int64)]
    counts = [map_size]
    vid = np.arange(n)
    for _ in range(1, L):
        ids = node[-1]
        # components of the current level, expressed per vertex
        nxt, k = _components(ids, [(a, b) for a, b in edges], B)
        node.append(nxt[vid])
        counts.append(k)
    return node, counts


def _bcoo_to_csr(vals, rows, cols, n):
    """Full symmetric 3n x 3n CSR of an upper BCOO (column-major 3x3 blocks)."""
    vals = np.asarray(vals).reshape(-1, 3, 3).transpose(0, 2, 1)  # row-major blocks
    rows, cols = np.asarray(rows, np.int64), np.asarray(cols, np.int64)
    r = (3 * rows[:, None, None] + np.arange(3)[None, :, None]).repeat(3, 2)
    c = (3 * cols[:, None, None] + np.arange(3)[None, None, :]).repeat(3, 1)
    A = sp.coo_matrix((vals.ravel(), (r.ravel(), c.ravel())), shape=(3 * n, 3 * n)).tocsr()
    off = rows != cols
    At = sp.coo_matrix((vals[off].ravel(), (r[off].ravel(), c[off].ravel())),
                       shape=(3 * n, 3 * n)).tocsr()
    return A + At.T


class CemasOperator:
    """The CEMAS preconditioner of an upper BCOO matrix (vertex space)."""

    def __init__(self, n, nbr, part_id, B, vals, rows, cols, pairs=None, maps=None):
        if maps is None:
            p2r, r2p = partition_maps_ref(part_id, B)
        else:
            p2r, r2p = maps
        self.p2r, self.r2p = np.asarray(p2r), np.asarray(r2p)
        self.n, self.B = n, B
        self.node, self.counts = hierarchy(n, nbr, self.r2p, self.p2r.size, B, pairs)
        A = _bcoo_to_csr(vals, rows, cols, n)
        self.levels = []
        for ids, c in zip(self.node, self.counts):
            m = (c + B - 1) // B * B
            P = sp.csr_matrix((np.ones(3 * n), (np.arange(3 * n), (3 * ids[:, None] + np.arange(3)).ravel())),
                              shape=(3 * n, 3 * m))
            Al = (P.T @ A @ P).tocsr()
            inv = np.zeros((m // B, 3 * B, 3 * B))
            for d in range(m // B):
                sl = slice(3 * B * d, 3 * B * (d + 1))
                D = Al[sl, sl].toarray()
                dg = np.diag(D).copy()
                dg[dg == 0.0] = 1.0
                np.fill_diagonal(D, dg)
                inv[d] = np.linalg.inv(D)
            self.levels.append((P, inv))

    def apply(self, r):
        r = np.asarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        for P, inv in self.levels:
            R = (P.T @ r).reshape(inv.shape[0], -1)
            Z = np.einsum("dij,dj->di", inv, R).reshape(-1)
            z += P @ Z
        return z

    def dense(self):
        e = np.eye(3 * self.n)
        return np.stack([self.apply(e[:, j]) for j in range(3 * self.n)], axis=1)


def pcg(A, b, M, tol=1e-4, stop_mode=0, max_iter=None):
    """PCG in the reference's form (pcg_solver.cu:191-332, the loop of the C
    oracle's orc_pcg2): x0 = 0; after the x/r update of iteration k stop when
    |r.z| <= tol r0.z0 with the previous iterate's r.z (stop_mode 0) or
    ||r|| <= tol ||b|| (stop_mode 1); k < max_iter (default 0.3 x unknowns).
    A: scipy matrix; M: callable.  Returns (x, iterations)."""
    b = np.asarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    r = b.copy()
    nb = np.linalg.norm(b)
    max_iter = max_iter or max(2, int(0.3 * b.size))
    iters = 0
    if nb == 0:
        return x, 0
    z = M(r)
    rz = r @ z
    rz0 = rz
    p = z.copy()
    for k in range(1, max_iter):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        iters = k
        if stop_mode == 0:
            if abs(rz) <= tol * rz0:
                break
        elif np.linalg.norm(r) <= tol * nb:
            break
        z = M(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    return x, iters
