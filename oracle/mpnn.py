"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY.

Learned coarse level: a message-passing network over the MAS level-1
aggregate graph that adds a coarse correction to the MAS preconditioner
inside flexible PCG.  The reference has no learned component; this module
restates the model that the library's kernels implement so that tests can
check the GPU forward pass, gradients and training against it.  Only
``tests/`` may import it.

Model (C channels, L layers, rows = 3 * nodes, one row per node component):

* Graph.  A_1 = P^T A P with P the level-1 aggregation (row v of the fine
  system maps to its level-1 node, ``MASPreconditioner.cu:1551-1728`` builds
  the same aggregates).  s_k = (tr(A_1[k,k]) / 3)^(-1/2) and
  Ahat = S A_1 S (S = diag(s) repeated per component) has a unit diagonal
  on average.
* Inputs per row (3 features): ghat = S P^T r, zhat = S^-1 zc and
  rho = ghat - Ahat zhat, where zc is the coarse part (levels >= 1) of the
  MAS solution restricted to level 1.
* V0 = X W_in; for each layer l: U = [V, Ahat V], Y = U W_l,
  V <- V + act(Y); chat = V w_out, act = relu or the identity.  No biases,
  so the network is positively homogeneous: f(a X) = a f(X) for a > 0 (and
  linear in X with the identity activation).
* Correction c = S chat, added to the MAS output as P c.
* Loss per sample: (0.5 chat^T Ahat chat - chat^T qhat) / rows with
  qhat = S P^T (r - A M r).  In original units this is the A-norm error of
  the preconditioned vector after the correction, minus a constant:
  ||e - M r - P c||_A^2 = const - 2 c^T P^T (r - A M r) + c^T A_1 c with
  A e = r.

GEMM operands (U, W, dY) are rounded to bfloat16 when ``bf16=True``, the
way the MFMA kernels consume them; accumulation is in float64 here.

What pins this module: finite differences of the loss (gradients), the
closed-form minimiser chat = Ahat^-1 qhat of the loss, positive homogeneity,
and scipy for the Galerkin product.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import scipy.sparse as sp


# ------------------------------------------------------------------ helpers
def bf16(x):
    """Round to bfloat16 (nearest, ties to even); returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bcoo_to_csr(vals, rows, cols, n):
    """Full symmetric scipy matrix from upper 3x3 blocks (column-major)."""
    blocks = np.asarray(vals, np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)
    r = (3 * rows[:, None, None] + np.arange(3)[None, :, None]).repeat(3, axis=2)
    c = (3 * cols[:, None, None] + np.arange(3)[None, None, :]).repeat(3, axis=1)
    U = sp.coo_matrix((blocks.reshape(-1), (r.reshape(-1), c.reshape(-1))), shape=(3 * n, 3 * n)).tocsr()
    d = rows == cols
    D = sp.coo_matrix((blocks[d].reshape(-1), (r[d].reshape(-1), c[d].reshape(-1))), shape=(3 * n, 3 * n)).tocsr()
    return (U + U.T - D).tocsr()


def prolongation(map1, n_coarse):
    """P (3n x 3c): P[3v+d, 3 map1[v] + d] = 1."""
    map1 = np.asarray(map1, np.int64)
    n = map1.size
    r = np.arange(3 * n)
    c = 3 * np.repeat(map1, 3) + np.tile(np.arange(3), n)
    return sp.csr_matrix((np.ones(3 * n), (r, c)), shape=(3 * n, 3 * n_coarse))


@dataclasses.dataclass
class CoarseGraph:
    A1: sp.csr_matrix      # level-1 Galerkin operator (3c x 3c)
    s: np.ndarray          # per-node scale (c,)
    Ahat: sp.csr_matrix    # S A1 S
    P: sp.csr_matrix       # prolongation (3n x 3c)

    @property
    def n(self):
        return self.s.size


def coarse_graph(A, map1, n_coarse):
    """Level-1 operator, scale and normalised operator from the fine matrix."""
    P = prolongation(map1, n_coarse)
    A1 = (P.T @ A @ P).tocsr()
    d = A1.diagonal().reshape(-1, 3).sum(axis=1) / 3.0
    s = np.where(d > 0, 1.0 / np.sqrt(np.where(d > 0, d, 1.0)), 0.0)
    S = sp.diags(np.repeat(s, 3))
    return CoarseGraph(A1, s, (S @ A1 @ S).tocsr(), P)


def coarse_upper_bcoo(A1, n_coarse):
    """Upper 3x3 blocks (column-major values, sorted by (row, col)) of A1."""
    B = sp.bsr_matrix(A1, blocksize=(3, 3))
    B.sort_indices()
    rows = np.repeat(np.arange(n_coarse), np.diff(B.indptr))
    cols = B.indices
    keep = cols >= rows
    blocks = B.data[keep]  # row-major 3x3
    return blocks.transpose(0, 2, 1).reshape(-1).copy(), rows[keep].astype(np.int32), cols[keep].astype(np.int32)


def features(g, zc, qfull, graph):
    """Row features X (3c x 3) and the loss target qhat (3c,) of one sample.

    g = P^T r, zc = coarse MAS solution at level 1 (primal), qfull =
    P^T (r - A M r).  Returned unnormalised: callers divide a sample by its
    own RMS of ghat so that every sample weighs the same in the loss."""
    s3 = np.repeat(graph.s, 3)
    ghat = s3 * g
    zhat = np.where(s3 > 0, zc / np.where(s3 > 0, s3, 1.0), 0.0)
    rho = ghat - graph.Ahat @ zhat
    return np.stack([ghat, zhat, rho], axis=1), s3 * qfull


# -------------------------------------------------------------------- model
@dataclasses.dataclass
class Params:
    W_in: np.ndarray   # (3, C)
    W: np.ndarray      # (L, 2C, C)
    w_out: np.ndarray  # (C,)

    @property
    def C(self):
        return self.w_out.size

    @property
    def L(self):
        return self.W.shape[0]

    def flat(self):
        return np.concatenate([self.W_in.reshape(-1), self.W.reshape(-1), self.w_out])

    @staticmethod
    def from_flat(v, C, L):
        v = np.asarray(v, np.float64)
        a = 3 * C
        b = a + L * 2 * C * C
        return Params(v[:a].reshape(3, C).copy(), v[a:b].reshape(L, 2 * C, C).copy(), v[b:b + C].copy())

    @staticmethod
    def size(C, L):
        return 3 * C + L * 2 * C * C + C


def init_params(C=64, L=3, seed=0):
    """He-style initial weights (zero-mean normal); the output layer starts
    small so the initial correction is close to zero."""
    rng = np.random.default_rng(seed)
    W_in = rng.standard_normal((3, C)) * np.sqrt(1.0 / 3.0)
    W = rng.standard_normal((L, 2 * C, C)) * np.sqrt(1.0 / (2 * C)) * 0.5
    w_out = rng.standard_normal(C) * np.sqrt(1.0 / C) * 0.1
    return Params(W_in, W, w_out)


def _r(x, on):
    return bf16(x) if on else x


def forward(p: Params, Ahat, X, bf16_ops=False, linear=False):
    """chat for one sample; returns (chat, cache).  linear=True uses the
    identity activation."""
    V = X @ p.W_in
    cache = []
    for l in range(p.L):
        U = np.concatenate([V, Ahat @ V], axis=1)
        Ub = _r(U, bf16_ops)
        Y = Ub @ _r(p.W[l], bf16_ops)
        cache.append((V, Ub, Y))
        V = V + (Y if linear else np.maximum(Y, 0.0))
    chat = V @ p.w_out
    cache.append((V,))
    return chat, cache


def loss(chat, Ahat, qhat):
    rows = chat.size
    return (0.5 * chat @ (Ahat @ chat) - chat @ qhat) / rows


def loss_grad(p: Params, Ahat, samples, bf16_ops=False, linear=False):
    """Mean loss over samples [(X, qhat), ...] and its gradient (Params)."""
    gW_in = np.zeros_like(p.W_in)
    gW = np.zeros_like(p.W)
    gw = np.zeros_like(p.w_out)
    total = 0.0
    ns = len(samples)
    for X, qhat in samples:
        chat, cache = forward(p, Ahat, X, bf16_ops, linear)
        rows = chat.size
        total += loss(chat, Ahat, qhat)
        dc = (Ahat @ chat - qhat) / (rows * ns)
        VL = cache[-1][0]
        gw += VL.T @ dc
        G = np.outer(dc, p.w_out)
        for l in range(p.L - 1, -1, -1):
            V, Ub, Y = cache[l]
            dY = G if linear else G * (Y > 0)
            dYb = _r(dY, bf16_ops)
            gW[l] += Ub.T @ dYb
            dU = dYb @ _r(p.W[l], bf16_ops).T
            C = p.C
            G = G + dU[:, :C] + Ahat @ dU[:, C:]
        gW_in += X.T @ G
    return total / ns, Params(gW_in, gW, gw)


def optimal_loss(Ahat, qhat):
    """Closed-form minimum of the loss over chat: -0.5 q^T Ahat^-1 q / rows."""
    import scipy.sparse.linalg as spla
    c = spla.spsolve(Ahat.tocsc(), qhat)
    return loss(c, Ahat, qhat), c


# ----------------------------------------------------------------- training
def adam(p: Params, Ahat, samples, epochs, batch, lr=1e-3, seed=0, b1=0.9, b2=0.999, eps=1e-8, bf16_ops=False,
         linear=False):
    """Plain Adam over shuffled minibatches; returns (params, loss history)."""
    v = p.flat()
    m = np.zeros_like(v)
    s2 = np.zeros_like(v)
    rng = np.random.default_rng(seed)
    hist = []
    t = 0
    C, L = p.C, p.L
    for _ in range(epochs):
        order = rng.permutation(len(samples))
        for b0 in range(0, len(order), batch):
            idx = order[b0:b0 + batch]
            cur = Params.from_flat(v, C, L)
            lv, g = loss_grad(cur, Ahat, [samples[i] for i in idx], bf16_ops, linear)
            g = g.flat()
            t += 1
            m = b1 * m + (1 - b1) * g
            s2 = b2 * s2 + (1 - b2) * g * g
            mh = m / (1 - b1 ** t)
            sh = s2 / (1 - b2 ** t)
            v = v - lr * mh / (np.sqrt(sh) + eps)
            hist.append(lv)
    return Params.from_flat(v, C, L), np.array(hist)


# ------------------------------------------------------- MAS parts in numpy
def _domain_dense(ub, B):
    """Dense (3B x 3B) domain matrices from the upper blocks (ndom, nblk, 9)
    in the MAS storage order (row r, then columns c >= r; column-major 3x3)."""
    nd = ub.shape[0]
    M = np.zeros((nd, 3 * B, 3 * B))
    k = 0
    for r in range(B):
        for c in range(r, B):
            blk = ub[:, k].reshape(nd, 3, 3).transpose(0, 2, 1)
            M[:, 3 * r:3 * r + 3, 3 * c:3 * c + 3] = blk
            if c != r:
                M[:, 3 * c:3 * c + 3, 3 * r:3 * r + 3] = blk.transpose(0, 2, 1)
            k += 1
    return M


class MasParts:
    """Level-by-level MAS apply of an oracle ``Mas`` in numpy: the total
    z = M r and the coarse part (levels >= 1) restricted to level 1."""

    def __init__(self, mas):
        L, off, cnt, total = mas.levels()
        ct, _ = mas.tables()
        self.L, self.off, self.cnt, self.total, self.B, self.n = L, off, cnt, total, mas.B, mas.n
        self.slot = np.concatenate([np.arange(mas.n)[:, None], ct], axis=1)  # (n, L)
        self.inv = _domain_dense(mas.matrices(inverse=True), mas.B)
        self.map1 = (self.slot[:, 1] - off[1]) if L > 1 else np.arange(mas.n)
        # first fine member of every level-1 node: its ancestors are the node's
        first = np.full(cnt[1] if L > 1 else mas.n, -1, np.int64)
        order = np.arange(mas.n)[::-1]
        first[self.map1[order]] = order
        self.first = first

    def apply(self, r):
        n, B, total = self.n, self.B, self.total
        r3 = r.reshape(n, 3)
        R = np.zeros((total, 3))
        for l in range(self.L):
            np.add.at(R, self.slot[:, l], r3)
        Z = np.einsum("dij,dj->di", self.inv, R.reshape(-1, 3 * B)).reshape(total, 3)
        z = np.zeros((n, 3))
        for l in range(self.L):
            z += Z[self.slot[:, l]]
        zc = np.zeros((self.first.size, 3))
        for l in range(1, self.L):
            zc += Z[self.slot[self.first, l]]
        return z.reshape(-1), zc.reshape(-1)


def level2_operator(parts: MasParts, graph):
    """Dense Galerkin operator of level 2, P_12^T A_1 P_12."""
    off = parts.off
    map2 = parts.slot[parts.first, 2] - off[2]
    P12 = prolongation(map2, parts.cnt[2])
    return (P12.T @ graph.A1 @ P12).toarray()


def sample(A, graph, parts, r):
    """Features and target of one residual r (rows of the fine system)."""
    z, zc = parts.apply(r)
    g = graph.P.T @ r
    q = graph.P.T @ (r - A @ z)
    X, qhat = features(g, zc, q, graph)
    sc = np.sqrt(np.mean(X[:, 0] ** 2))
    sc = sc if sc > 0 else 1.0
    return X / sc, qhat / sc, z


# ------------------------------------------------------- level-1 coarse cycle
def float_operator(graph):
    """A_1 as the device applies it: S^-1 Ahat S^-1 with s rounded to float32
    first and the scaled blocks (float32(s_i s_j A_ij)) rounded to float32."""
    s32 = graph.s.astype(np.float32).astype(np.float64)
    S = sp.diags(np.repeat(s32, 3))
    blk = (S @ graph.A1 @ S).tocsr()
    blk.data = blk.data.astype(np.float32).astype(np.float64)
    w = np.where(s32 > 0, 1.0 / np.where(s32 > 0, s32, 1.0), 0.0)
    W = sp.diags(np.repeat(w, 3))
    return (W @ blk @ W).tocsr()


def _domain_solve(parts: MasParts, R, lo_slot, hi_slot):
    """Domain solves of the slot range [lo_slot, hi_slot) (whole domains)."""
    B, total = parts.B, parts.total
    Z = np.zeros((total, 3))
    d0, d1 = lo_slot // B, hi_slot // B
    Z.reshape(-1, 3 * B)[d0:d1] = np.einsum("dij,dj->di", parts.inv[d0:d1], R.reshape(-1, 3 * B)[d0:d1])
    return Z


def level0_apply(parts: MasParts, r):
    """The level-0 domain solves alone, M_0 r (the identity slot map of level 0)."""
    n = parts.n
    R = np.zeros((parts.total, 3))
    R[:n] = r.reshape(n, 3)
    return _domain_solve(parts, R, 0, parts.off[1])[:n].reshape(-1)


def level1_cycle(parts: MasParts, A1f, g, A2=None):
    """The level-1 V(1,1) cycle on the level-1 right-hand side g (c1 x 3,
    the restriction P^T r): pre-smoothing with the level-1 domain solves, a
    coarse correction of the residual, post-smoothing; residuals with A1f.
    The coarse correction is the exact solve with A2 (the level-2 Galerkin
    operator, dense) when given, else the additive levels >= 2.  Returns the
    level-1 correction c (c1 x 3)."""
    off, L, total = parts.off, parts.L, parts.total
    c1 = parts.cnt[1]
    R = np.zeros((total, 3))
    R[off[1]:off[1] + c1] = g.reshape(-1, 3)
    c1v = _domain_solve(parts, R, off[1], off[2])[off[1]:off[1] + c1].reshape(-1)
    gv = g.reshape(-1)
    rho = gv - A1f @ c1v
    slots1 = parts.slot[parts.first]           # ancestors of every level-1 node
    C2 = c1v.reshape(-1, 3).copy()
    if A2 is not None:
        c2 = parts.cnt[2]
        R2 = np.zeros((c2, 3))
        np.add.at(R2, slots1[:, 2] - off[2], rho.reshape(-1, 3))
        e2 = np.linalg.solve(A2, R2.reshape(-1)).reshape(-1, 3)
        C2 += e2[slots1[:, 2] - off[2]]
    else:
        Rc = np.zeros((total, 3))
        for l in range(2, L):
            np.add.at(Rc, slots1[:, l], rho.reshape(-1, 3))
        Zc = _domain_solve(parts, Rc, off[2], off[L])
        for l in range(2, L):
            C2 += Zc[slots1[:, l]]
    rho2 = gv - A1f @ C2.reshape(-1)
    R2 = np.zeros((total, 3))
    R2[off[1]:off[1] + c1] = rho2.reshape(-1, 3)
    Z1 = _domain_solve(parts, R2, off[1], off[2])[off[1]:off[1] + c1]
    return C2 + Z1


def mas_cycle_apply(parts: MasParts, A1f, r, A2=None):
    """MAS apply with the level-1 V(1,1) cycle (the library's coarse mode 1):
    level 0 additive, the level-1 cycle on P^T r prolonged."""
    n = parts.n
    g = np.zeros((parts.cnt[1], 3))
    np.add.at(g, parts.map1, r.reshape(n, 3))
    c = level1_cycle(parts, A1f, g, A2)
    return level0_apply(parts, r) + c[parts.map1].reshape(-1)


def ap_operator(A, map1, n_coarse):
    """A P (P the level-1 aggregation), blocks summed in float64 and rounded to
    float32, as the library stores it for the post-smoothing residual."""
    AP = (A @ prolongation(map1, n_coarse)).tocsr()
    AP.data = AP.data.astype(np.float32).astype(np.float64)
    return AP


def mas_vcycle_apply(parts: MasParts, A1f, A, APf, r, A2=None, omega=0.5):
    """The library's coarse mode 2, a symmetric two-grid cycle on level 0
    around the level-1 cycle (no counterpart in the reference):
        z0 = omega M0 r,  r1 = r - A z0,  c = level-1 cycle of P^T r1,
        r2 = r1 - A P c,  z = z0 + P c + omega M0 r2,
    with M0 the level-0 domain solves and A P in float32 (APf).  The library
    forms M0 (r + r2) in one pass of solves; that is the same z."""
    n = parts.n
    z0 = omega * level0_apply(parts, r)
    r1 = r - A @ z0
    g = np.zeros((parts.cnt[1], 3))
    np.add.at(g, parts.map1, r1.reshape(n, 3))
    c = level1_cycle(parts, A1f, g, A2)
    r2 = r1 - APf @ c.reshape(-1)
    return z0 + c[parts.map1].reshape(-1) + omega * level0_apply(parts, r2)


def half_level0_inverse(inv, n_dom0):
    """The FP16 storage of the level-0 domain inverses that the float applies
    of coarse mode 2 read (k_half_level0 in mas.hip; no counterpart in the
    reference): per slot d = sqrt(Minv_ii) and 1/d in float32, the scaled
    entries C = (Minv_ij / d_i) / d_j rounded to float32 and then to float16
    (|C_ij| <= 1 for an SPD inverse), and the matrix the apply uses,
    D C D.  Slots with a non-positive diagonal (padding) get d = 0.
    ``inv``: (domains, 3B, 3B) dense inverses; domains below ``n_dom0`` are
    level 0 and are replaced, the rest are returned unchanged."""
    out = inv.copy()
    for k in range(n_dom0):
        M = inv[k].astype(np.float32)
        v = np.diag(M).copy()
        pos = v > 0
        d = np.where(pos, np.sqrt(np.where(pos, v, 1)).astype(np.float32), np.float32(0))
        dinv = np.where(pos, np.float32(1) / np.where(pos, d, 1), np.float32(0)).astype(np.float32)
        C = ((M * dinv[:, None]).astype(np.float32) * dinv[None, :]).astype(np.float32)
        C = np.clip(C, -65504, 65504).astype(np.float16).astype(np.float64)
        out[k] = d[:, None].astype(np.float64) * C * d[None, :].astype(np.float64)
    return out
