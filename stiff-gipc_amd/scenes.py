"""Seeded synthetic scenes for the Newton linear-solve path.

Each generator returns a :class:`TetScene` holding exactly the per-element and
per-vertex inputs the reference's assembly kernels read, computed the way the
reference's headless bootstrap computes them:

* masses, rest ``DmInverse`` (row-major 3x3) and rest volume per tet, and the
  per-tet Lame pair ``mu' = 4 mu / 3``, ``lambda' = lambda + 5 mu / 6``
  (``app/common/sim_bootstrap.cpp:22-88``, ``init_fem``);
* ``Ds`` columns ``x1-x0, x2-x0, x3-x0`` (``StiffGIPC/femEnergy.cu:324-350``);
* the symmetric vertex-neighbour CSR built from the unique tet edges
  (``StiffGIPC/load_mesh.cpp:934-1010``, ``getVertNeighbors``), which is what the
  MAS hierarchy reads;
* point-triangle contact pairs in the reference's ``int4`` encoding
  ``(-v-1, t0, t1, t2)`` (``StiffGIPC/mlbvh.cu:539-760``).

Grids use the Kuhn (Freudenthal) 6-tet split of each cube, which is conforming,
and vertices are relabelled in Morton (Z-curve) order so that MAS domains of
consecutive vertices are compact sub-boxes.  Rest positions are the regular
grid; the current positions are the rest positions jittered by a seeded uniform
offset, so the deformation gradients differ from the identity and the PSD
projection of the elastic Hessian is exercised.

The system matrix follows the reference scaling ``H = M + dt^2 (K + B)``
(``StiffGIPC/GIPC.cu:10466-10485``, ``StiffGIPC/femEnergy.cu:1820``): it differs
from ``M/dt^2 + K`` by the constant factor ``dt^2`` only, so relative residuals
and relative solution differences are unchanged.
"""
from __future__ import annotations

import dataclasses
import numpy as np

__all__ = [
    "TetScene",
    "HybridScene",
    "hybrid_scene",
    "scene_c2",
    "scene_hybrid_tiny",
    "ClothScene",
    "cloth_grid_scene",
    "scene_cloth_tiny",
    "scene_c4",
    "morton2",
    "lame_from_young",
    "morton3",
    "kuhn_grid",
    "tet_grid_scene",
    "tet_mesh_scene",
    "cloth_mesh_scene",
    "cloth_rest_dm",
    "scene_c0",
    "scene_c1",
    "scene_tiny",
    "seed_pt_pairs",
]


@dataclasses.dataclass
class TetScene:
    """Inputs of one Newton linear solve on a tet mesh (+ optional contacts)."""

    name: str
    x: np.ndarray            # (n, 3) float64 current positions
    rest: np.ndarray         # (n, 3) float64 rest positions
    tets: np.ndarray         # (m, 4) int32, positively oriented at rest
    dm_inv: np.ndarray       # (m, 3, 3) float64 row-major rest Dm^{-1}
    vol: np.ndarray          # (m,) float64 rest volume
    mu: np.ndarray           # (m,) float64 lengthRate  = 4 mu / 3
    lam: np.ndarray          # (m,) float64 volumeRate  = lambda + 5 mu / 6
    mass: np.ndarray         # (n,) float64 lumped mass
    dt: float
    nbr_start: np.ndarray    # (n,) uint32 CSR start of the vertex-neighbour list
    nbr_num: np.ndarray      # (n,) uint32 neighbour count
    nbr_list: np.ndarray     # (nnz,) uint32 neighbour ids
    pairs: np.ndarray        # (p, 4) int32 contact pairs, reference int4 encoding
    pair_x: np.ndarray | None  # (p, 4, 3) float64 local contact geometry or None
    dhat: float              # barrier activation distance (m); reference dHat = dhat^2
    kappa: float
    rhs: np.ndarray          # (3n,) float64 right-hand side
    seeds: dict
    boundary_type: np.ndarray | None = None  # (n,) int32, nonzero = Dirichlet vertex (TetMesh.BoundaryType)
    part_id: np.ndarray | None = None  # (n,) int32 CEMAS partition (MAS slot maps), or None

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_tets(self) -> int:
        return int(self.tets.shape[0])

    @property
    def n_pairs(self) -> int:
        return int(self.pairs.shape[0])

    @property
    def n_tris(self) -> int:
        return 0

    @property
    def n_hinges(self) -> int:
        return 0

    @property
    def n_triplets(self) -> int:
        """Triplet count written by the assembly kernels (16/pair, 10/tet, 1/vertex)."""
        return 16 * self.n_pairs + 10 * self.n_tets + self.n


@dataclasses.dataclass
class ClothScene:
    """Inputs of one Newton linear solve on a cloth sheet (+ optional contacts).

    Per-triangle inputs follow the reference's cloth set-up
    (``StiffGIPC/gl_main.cu:578-584,622-641``): ``dm_inv2`` is the inverse of
    the 2x2 rest edge matrix after rotating the rest triangle into the xy
    plane (``femEnergy.cu:256-300``), ``area`` is the rest area times the
    cloth thickness, the lumped mass is density * area / 3 per corner,
    ``k_stretch = E / (2 (1 + nu))`` and ``k_shear = 0.03 k_stretch rate``.
    Hinges list each interior edge with the two opposite vertices
    (``adj[:, 1] == -1`` on the boundary), the layout
    ``_calculate_quad_bending_gradient_hessian`` reads (femEnergy.cu:2623-2754).
    """

    name: str
    x: np.ndarray            # (n, 3) current positions
    rest: np.ndarray         # (n, 3) rest positions
    tris: np.ndarray         # (m, 3) int32
    dm_inv2: np.ndarray      # (m, 2, 2) row-major
    area: np.ndarray         # (m,) rest area * thickness
    k_stretch: float
    k_shear: float
    strain_rate: float
    edges: np.ndarray        # (e, 2) int32 hinge edge
    adj: np.ndarray          # (e, 2) int32 opposite vertices, -1 on the boundary
    k_bend: float
    mass: np.ndarray
    dt: float
    nbr_start: np.ndarray
    nbr_num: np.ndarray
    nbr_list: np.ndarray
    pairs: np.ndarray
    pair_x: np.ndarray | None
    dhat: float
    kappa: float
    rhs: np.ndarray
    seeds: dict
    boundary_type: np.ndarray | None = None  # (n,) int32, nonzero = Dirichlet vertex
    # surface whose contacts are detected on every solve (pairs then unused)
    contact_surface: "ContactGeometry | None" = None
    part_id: np.ndarray | None = None  # (n,) int32 CEMAS partition (MAS slot maps), or None

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def n_tets(self) -> int:
        return 0

    @property
    def n_tris(self) -> int:
        return int(self.tris.shape[0])

    @property
    def n_hinges(self) -> int:
        return int(self.edges.shape[0])

    @property
    def n_pairs(self) -> int:
        return int(self.pairs.shape[0])

    @property
    def n_triplets(self) -> int:
        """16/pair, 10/hinge, 6/triangle, 1/vertex."""
        return 16 * self.n_pairs + 10 * self.n_hinges + 6 * self.n_tris + self.n


def lame_from_young(E, nu):
    """Return the reference's (lengthRate, volumeRate) pair for Young's modulus E.

    ``app/common/sim_bootstrap.cpp:27-31,57-62``: mu = E / (2(1+nu)),
    lambda = E nu / ((1+nu)(1-2nu)); lengthRate = 4 mu / 3 and
    volumeRate = lambda + 5 mu / 6 (the stable Neo-Hookean re-parameterisation).
    """
    E = np.asarray(E, dtype=np.float64)
    mu = E / (2.0 * (1.0 + nu))
    lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return 4.0 * mu / 3.0, lam + 5.0 * mu / 6.0


def _part1by2(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & np.uint64(0x1FFFFF)
    v = (v | (v << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x1249249249249249)
    return v


def _part1by1(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.uint64) & np.uint64(0xFFFFFFFF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x0000FFFF0000FFFF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF00FF00FF)
    v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F0F0F0F0F)
    v = (v | (v << np.uint64(2))) & np.uint64(0x3333333333333333)
    v = (v | (v << np.uint64(1))) & np.uint64(0x5555555555555555)
    return v


def morton2(i, j) -> np.ndarray:
    """64-bit 2D Morton code with i in the lowest bit of each pair."""
    return _part1by1(i) | (_part1by1(j) << np.uint64(1))


def morton3(i, j, k) -> np.ndarray:
    """64-bit Morton code with x in the lowest bit of each triple."""
    return _part1by2(i) | (_part1by2(j) << np.uint64(1)) | (_part1by2(k) << np.uint64(2))


# Kuhn split: for each axis permutation (a, b, c) the tet 000 -> e_a -> e_a+e_b -> 111.
_KUHN_PERMS = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
_PERM_SIGN = {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1, (0, 2, 1): -1, (1, 0, 2): -1, (2, 1, 0): -1}


def kuhn_grid(nx: int, ny: int, nz: int):
    """Vertex lattice coordinates (n,3) int and positively oriented tets (m,4) int64.

    Vertices are numbered lexicographically (x fastest); relabel afterwards.
    """
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    lat = np.stack([ii.ravel(order="F"), jj.ravel(order="F"), kk.ravel(order="F")], axis=1)

    def vid(i, j, k):
        return i + nx * (j + ny * k)

    ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    ci, cj, ck = ci.ravel(order="F"), cj.ravel(order="F"), ck.ravel(order="F")
    tets = []
    for perm in _KUHN_PERMS:
        a, b, _ = perm
        o = np.zeros(3, dtype=np.int64)
        p1 = o.copy(); p1[a] = 1
        p2 = p1.copy(); p2[b] = 1
        p3 = np.ones(3, dtype=np.int64)
        corners = [o, p1, p2, p3]
        ids = [vid(ci + c[0], cj + c[1], ck + c[2]) for c in corners]
        if _PERM_SIGN[perm] < 0:
            ids[1], ids[2] = ids[2], ids[1]
        tets.append(np.stack(ids, axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4)  # 6 consecutive tets per cube
    return lat, tets


def _edges_csr(n: int, tets: np.ndarray):
    """Symmetric vertex-neighbour CSR from the unique tet edges."""
    e = np.concatenate([tets[:, [a, b]] for a, b in ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))])
    e = np.sort(e, axis=1)
    key = np.unique(e[:, 0].astype(np.int64) * n + e[:, 1])
    u, v = key // n, key % n
    src = np.concatenate([u, v])
    dst = np.concatenate([v, u])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    num = np.bincount(src, minlength=n).astype(np.uint32)
    start = np.zeros(n, dtype=np.uint32)
    start[1:] = np.cumsum(num)[:-1]
    return start, num, dst.astype(np.uint32), int(key.size)


def _surface_triangles(tets: np.ndarray) -> np.ndarray:
    """Boundary faces of a tet mesh (faces referenced by exactly one tet)."""
    faces = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 3, 2]], tets[:, [0, 1, 3]], tets[:, [0, 2, 1]]])
    s = np.sort(faces, axis=1)
    _, idx, cnt = np.unique(s, axis=0, return_index=True, return_counts=True)
    return faces[idx[cnt == 1]]


def seed_pt_pairs(x: np.ndarray, tets: np.ndarray, n_pairs: int, dhat: float, seed: int):
    """Seeded point-triangle pairs on the boundary of a tet mesh; see
    :func:`seed_pt_pairs_on`."""
    return seed_pt_pairs_on(x, _surface_triangles(tets), n_pairs, dhat, seed)


def seed_pt_pairs_on(x: np.ndarray, tris: np.ndarray, n_pairs: int, dhat: float, seed: int):
    """Seeded point-triangle pairs with prescribed squared distance d^2 ~ U(0, dhat^2).

    Vertex p is a random vertex of ``tris`` and (t0,t1,t2) a random triangle
    not incident to p.  The barrier Hessian depends on the relative geometry
    only, so each pair carries its own local geometry: the triangle at its
    current positions and the point placed at distance d along the triangle
    normal above a random interior barycentric point (so the point-triangle
    distance is the point-plane distance, as for a reference PT pair).
    Returns (pairs (p,4) int32 in the reference encoding, pair_x (p,4,3)).
    """
    rng = np.random.default_rng(seed)
    surf_v = np.unique(tris)
    pv = surf_v[rng.integers(0, surf_v.size, size=n_pairs)]
    ti = rng.integers(0, tris.shape[0], size=n_pairs)
    t = tris[ti]
    bad = (t == pv[:, None]).any(axis=1)
    while bad.any():
        ti[bad] = rng.integers(0, tris.shape[0], size=int(bad.sum()))
        t = tris[ti]
        bad = (t == pv[:, None]).any(axis=1)
    d2 = rng.uniform(0.0, 1.0, size=n_pairs) * dhat * dhat
    d2 = np.maximum(d2, 1e-12 * dhat * dhat)
    bary = rng.uniform(0.1, 1.0, size=(n_pairs, 3))
    bary /= bary.sum(axis=1, keepdims=True)
    side = np.where(rng.uniform(size=n_pairs) < 0.5, -1.0, 1.0)
    t0, t1, t2 = x[t[:, 0]], x[t[:, 1]], x[t[:, 2]]
    nrm = np.cross(t1 - t0, t2 - t0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    foot = bary[:, 0:1] * t0 + bary[:, 1:2] * t1 + bary[:, 2:3] * t2
    p = foot + (side * np.sqrt(d2))[:, None] * nrm
    pairs = np.stack([-pv - 1, t[:, 0], t[:, 1], t[:, 2]], axis=1).astype(np.int32)
    pair_x = np.stack([p, t0, t1, t2], axis=1)
    return pairs, pair_x


def tet_grid_scene(name: str, N: int, *, spacing=0.01, jitter=0.1, E=1e5, E_loguniform=None,
                   nu=0.45, density=1e3, dt=0.01, n_pairs=0, dhat=1e-3, kappa=1e6,
                   seed_jitter=0, seed_rhs=1, seed_E=2, seed_pairs=3) -> TetScene:
    """Regular N^3 vertex grid, Kuhn split, Morton order, seeded jitter/E/pairs/rhs."""
    lat, tets = kuhn_grid(N, N, N)
    code = morton3(lat[:, 0], lat[:, 1], lat[:, 2])
    order = np.argsort(code, kind="stable")          # new id -> old id
    relabel = np.empty_like(order)
    relabel[order] = np.arange(order.size)            # old id -> new id
    lat = lat[order]
    tets = relabel[tets]
    rest = lat.astype(np.float64) * spacing
    rng = np.random.default_rng(seed_jitter)
    x = rest + rng.uniform(-jitter * spacing, jitter * spacing, size=rest.shape)
    return _tet_scene(name, rest, x, tets, E=E, E_loguniform=E_loguniform, nu=nu, density=density, dt=dt,
                      n_pairs=n_pairs, dhat=dhat, kappa=kappa, seed_rhs=seed_rhs, seed_E=seed_E,
                      seed_pairs=seed_pairs, seeds=dict(jitter=seed_jitter, rhs=seed_rhs, E=seed_E, pairs=seed_pairs))


def _tet_scene(name, rest, x, tets, *, E, E_loguniform, nu, density, dt, n_pairs, dhat, kappa, seed_rhs, seed_E,
               seed_pairs, seeds, part_id=None) -> TetScene:
    n = rest.shape[0]
    Dm = np.stack([rest[tets[:, 1]] - rest[tets[:, 0]],
                   rest[tets[:, 2]] - rest[tets[:, 0]],
                   rest[tets[:, 3]] - rest[tets[:, 0]]], axis=2)   # columns
    detm = np.linalg.det(Dm)
    assert (detm > 0).all()
    dm_inv = np.linalg.inv(Dm)
    vol = np.abs(detm) / 6.0
    mass = np.zeros(n)
    np.add.at(mass, tets.ravel(), np.repeat(vol * density / 4.0, 4))

    m = tets.shape[0]
    if E_loguniform is not None:
        lo, hi = E_loguniform
        Es = 10.0 ** np.random.default_rng(seed_E).uniform(np.log10(lo), np.log10(hi), size=m)
    else:
        Es = np.full(m, float(E))
    mu, lam = lame_from_young(Es, nu)

    start, num, lst, _ = _edges_csr(n, tets)
    if n_pairs:
        pairs, pair_x = seed_pt_pairs(x, tets, n_pairs, dhat, seed_pairs)
    else:
        pairs, pair_x = np.zeros((0, 4), dtype=np.int32), None
    rhs = np.random.default_rng(seed_rhs).standard_normal(3 * n)
    return TetScene(name=name, x=np.ascontiguousarray(x), rest=rest, tets=tets.astype(np.int32),
                    dm_inv=np.ascontiguousarray(dm_inv), vol=vol, mu=np.ascontiguousarray(mu),
                    lam=np.ascontiguousarray(lam), mass=mass, dt=dt, nbr_start=start, nbr_num=num,
                    nbr_list=lst, pairs=pairs, pair_x=pair_x, dhat=dhat, kappa=kappa, rhs=rhs,
                    seeds=seeds, part_id=part_id)


def _mean_edge(x: np.ndarray, el: np.ndarray) -> float:
    k = el.shape[1]
    return float(np.mean([np.linalg.norm(x[el[:, i]] - x[el[:, (i + 1) % k]], axis=1).mean() for i in range(k)]))


def tet_mesh_scene(name: str, rest: np.ndarray, tets: np.ndarray, *, part_id=None, jitter=0.1, E=1e5,
                   E_loguniform=None, nu=0.45, density=1e3, dt=0.01, n_pairs=0, dhat=None, kappa=1e6,
                   seed_jitter=0, seed_rhs=1, seed_E=2, seed_pairs=3) -> TetScene:
    """A loaded tet mesh at rest, jittered by a seeded uniform offset of
    ``jitter`` times its mean edge length; negatively oriented tets get their
    last two vertices swapped.  ``part_id``: the CEMAS partition (per vertex)
    the system uses for its MAS slot maps, or None."""
    rest = np.ascontiguousarray(rest, dtype=np.float64)
    tets = np.asarray(tets, dtype=np.int64).copy()
    d = np.stack([rest[tets[:, i]] - rest[tets[:, 0]] for i in (1, 2, 3)], axis=2)
    neg = np.linalg.det(d) < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    h = _mean_edge(rest, tets)
    rng = np.random.default_rng(seed_jitter)
    x = rest + rng.uniform(-jitter * h, jitter * h, size=rest.shape)
    if dhat is None:
        dhat = 0.1 * h
    return _tet_scene(name, rest, x, tets, E=E, E_loguniform=E_loguniform, nu=nu, density=density, dt=dt,
                      n_pairs=n_pairs, dhat=dhat, kappa=kappa, seed_rhs=seed_rhs, seed_E=seed_E,
                      seed_pairs=seed_pairs, seeds=dict(jitter=seed_jitter, rhs=seed_rhs, E=seed_E, pairs=seed_pairs),
                      part_id=None if part_id is None else np.ascontiguousarray(part_id, dtype=np.int32))


def scene_c0() -> TetScene:
    """16^3 grid, E = 1e5 Pa, no contact (the CPU-runnable case)."""
    return tet_grid_scene("c0_grid16", 16)


def scene_c1(n_pairs: int = 50_000) -> TetScene:
    """64^3 grid, per-tet E log-uniform in [1e4, 1e9] Pa, 50k point-triangle pairs."""
    return tet_grid_scene("c1_grid64_contact", 64, E_loguniform=(1e4, 1e9), n_pairs=n_pairs)


def scene_tiny(N: int = 6, n_pairs: int = 12, E_loguniform=(1e4, 1e9)) -> TetScene:
    """Small grid with contacts and a wide stiffness range, for dense checks."""
    return tet_grid_scene(f"tiny_grid{N}", N, E_loguniform=E_loguniform, n_pairs=n_pairs)


def _tri_edges_csr(n: int, tris: np.ndarray):
    """Unique triangle edges with their opposite vertices, and the symmetric
    vertex-neighbour CSR of the edges."""
    m = tris.shape[0]
    e = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]])
    opp = np.concatenate([tris[:, 2], tris[:, 0], tris[:, 1]])
    tid = np.tile(np.arange(m), 3)
    lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
    key = lo.astype(np.int64) * n + hi
    order = np.lexsort((tid, key))
    key, opp, lo, hi = key[order], opp[order], lo[order], hi[order]
    head = np.r_[True, key[1:] != key[:-1]]
    first = np.flatnonzero(head)
    cnt = np.diff(np.r_[first, key.size])
    assert cnt.max() <= 2, "non-manifold edge"
    edges = np.stack([lo[first], hi[first]], axis=1)
    adj = np.full((first.size, 2), -1, dtype=np.int64)
    adj[:, 0] = opp[first]
    two = cnt == 2
    adj[two, 1] = opp[first[two] + 1]
    src = np.concatenate([edges[:, 0], edges[:, 1]])
    dst = np.concatenate([edges[:, 1], edges[:, 0]])
    o = np.lexsort((dst, src))
    src, dst = src[o], dst[o]
    num = np.bincount(src, minlength=n).astype(np.uint32)
    start = np.zeros(n, dtype=np.uint32)
    start[1:] = np.cumsum(num)[:-1]
    return edges.astype(np.int32), adj.astype(np.int32), start, num, dst.astype(np.uint32)


def cloth_grid_scene(name: str, N: int, *, spacing=1e-3, jitter=0.1, E=1e6, nu=0.49, thickness=1e-3,
                     density=2e2, strain_rate=100.0, k_bend=1e-4, drape_radius=1.0, dt=0.01, n_pairs=0,
                     dhat=1e-3, kappa=1e6, seed_jitter=0, seed_rhs=1, seed_pairs=3) -> ClothScene:
    """N x N vertex sheet (2 triangles per quad, 2D Morton order), flat at rest
    in the xy plane, currently draped over a sphere of radius
    drape_radius * side length (z = -(u^2+v^2) / (2 R)) with a seeded in-plane
    jitter, so stretch, compression, shear and bending all appear."""
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    code = morton2(ii, jj)
    order = np.argsort(code, kind="stable")
    relabel = np.empty_like(order)
    relabel[order] = np.arange(order.size)
    n = N * N

    def vid(i, j):
        return relabel[i * N + j]

    qi, qj = np.meshgrid(np.arange(N - 1), np.arange(N - 1), indexing="ij")
    qi, qj = qi.ravel(), qj.ravel()
    v00, v10, v01, v11 = vid(qi, qj), vid(qi + 1, qj), vid(qi, qj + 1), vid(qi + 1, qj + 1)
    tris = np.concatenate([np.stack([v00, v10, v11], axis=1), np.stack([v00, v11, v01], axis=1)]).astype(np.int32)
    lat = np.stack([ii[order], jj[order]], axis=1).astype(np.float64)
    side = (N - 1) * spacing
    rest = np.zeros((n, 3))
    rest[:, :2] = lat * spacing - 0.5 * side
    rng = np.random.default_rng(seed_jitter)
    x = rest.copy()
    x[:, :2] += rng.uniform(-jitter * spacing, jitter * spacing, size=(n, 2))
    R = drape_radius * side
    x[:, 2] = -(x[:, 0] ** 2 + x[:, 1] ** 2) / (2.0 * R)

    a, b, c = rest[tris[:, 0]], rest[tris[:, 1]], rest[tris[:, 2]]
    Dm = np.stack([(b - a)[:, :2], (c - a)[:, :2]], axis=2)   # columns u0, u1
    detm = np.linalg.det(Dm)
    assert (detm > 0).all()
    dm_inv2 = np.linalg.inv(Dm)
    area = 0.5 * np.abs(detm) * thickness
    mass = np.zeros(n)
    np.add.at(mass, tris.ravel(), np.repeat(density * area / 3.0, 3))
    k_stretch = E / (2.0 * (1.0 + nu))
    k_shear = 0.03 * k_stretch * strain_rate
    edges, adj, start, num, lst = _tri_edges_csr(n, tris)
    if n_pairs:
        pairs, pair_x = seed_pt_pairs_on(x, tris, n_pairs, dhat, seed_pairs)
    else:
        pairs, pair_x = np.zeros((0, 4), dtype=np.int32), None
    rhs = np.random.default_rng(seed_rhs).standard_normal(3 * n)
    return ClothScene(name=name, x=np.ascontiguousarray(x), rest=rest, tris=tris,
                      dm_inv2=np.ascontiguousarray(dm_inv2), area=area, k_stretch=k_stretch, k_shear=k_shear,
                      strain_rate=strain_rate, edges=edges, adj=adj, k_bend=k_bend, mass=mass, dt=dt,
                      nbr_start=start, nbr_num=num, nbr_list=lst, pairs=pairs, pair_x=pair_x, dhat=dhat,
                      kappa=kappa, rhs=rhs, seeds=dict(jitter=seed_jitter, rhs=seed_rhs, pairs=seed_pairs))


def cloth_rest_dm(rest: np.ndarray, tris: np.ndarray) -> np.ndarray:
    """2x2 rest edge matrices [u0 u1] of each triangle after the rotation
    that takes its unit normal n onto +z (``__calculateDm2D_double``,
    StiffGIPC/femEnergy.cu:256-300): R = I + [v]x + [v]x^2 / (1 + c) with
    v = n x z, c = n.z, and diag(-1, -1, 1) when n = -z; the rotated corners'
    xy coordinates give u0 = p1 - p0, u1 = p2 - p0."""
    a, b, c = rest[tris[:, 0]], rest[tris[:, 1]], rest[tris[:, 2]]
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    z = np.array([0.0, 0.0, 1.0])
    v = np.cross(nrm, z)
    cs = nrm @ z
    m = tris.shape[0]
    K = np.zeros((m, 3, 3))
    K[:, 0, 1], K[:, 0, 2] = -v[:, 2], v[:, 1]
    K[:, 1, 0], K[:, 1, 2] = v[:, 2], -v[:, 0]
    K[:, 2, 0], K[:, 2, 1] = -v[:, 1], v[:, 0]
    R = np.broadcast_to(np.eye(3), (m, 3, 3)).copy()
    flip = cs + 1 == 0
    ok = ~flip
    R[ok] = R[ok] + K[ok] + (K[ok] @ K[ok]) / (1.0 + cs[ok])[:, None, None]
    R[flip, 0, 0] = -1.0
    R[flip, 1, 1] = -1.0
    p0 = np.einsum("mij,mj->mi", R, a)[:, :2]
    p1 = np.einsum("mij,mj->mi", R, b)[:, :2]
    p2 = np.einsum("mij,mj->mi", R, c)[:, :2]
    return np.stack([p1 - p0, p2 - p0], axis=2)


def cloth_mesh_scene(name: str, rest: np.ndarray, tris: np.ndarray, *, part_id=None, jitter=0.1, E=1e6, nu=0.49,
                     thickness=1e-3, density=2e2, strain_rate=100.0, k_bend=1e-4, drape_radius=1.0, dt=0.01,
                     seed_jitter=0, seed_rhs=1) -> ClothScene:
    """A loaded triangle mesh at rest (rest frames per ``cloth_rest_dm``),
    currently jittered in place by a seeded uniform offset of ``jitter``
    times its mean edge length and bent along its normal direction of least
    extent into a bowl of radius drape_radius * its largest extent.
    ``part_id``: the CEMAS partition (per vertex), or None."""
    rest = np.ascontiguousarray(rest, dtype=np.float64)
    tris = np.asarray(tris, dtype=np.int32)
    n = rest.shape[0]
    h = _mean_edge(rest, tris)
    rng = np.random.default_rng(seed_jitter)
    x = rest + rng.uniform(-jitter * h, jitter * h, size=rest.shape)
    ext = rest.max(axis=0) - rest.min(axis=0)
    ax = int(np.argmin(ext))
    ctr = 0.5 * (rest.max(axis=0) + rest.min(axis=0))
    uv = [i for i in range(3) if i != ax]
    R = drape_radius * float(ext.max())
    x[:, ax] = rest[:, ax] - ((x[:, uv[0]] - ctr[uv[0]]) ** 2 + (x[:, uv[1]] - ctr[uv[1]]) ** 2) / (2.0 * R)
    Dm = cloth_rest_dm(rest, tris)
    detm = np.linalg.det(Dm)
    assert (detm > 0).all()
    dm_inv2 = np.linalg.inv(Dm)
    area = 0.5 * np.abs(detm) * thickness
    mass = np.zeros(n)
    np.add.at(mass, tris.ravel(), np.repeat(density * area / 3.0, 3))
    k_stretch = E / (2.0 * (1.0 + nu))
    k_shear = 0.03 * k_stretch * strain_rate
    edges, adj, start, num, lst = _tri_edges_csr(n, tris)
    rhs = np.random.default_rng(seed_rhs).standard_normal(3 * n)
    return ClothScene(name=name, x=np.ascontiguousarray(x), rest=rest, tris=tris,
                      dm_inv2=np.ascontiguousarray(dm_inv2), area=area, k_stretch=k_stretch, k_shear=k_shear,
                      strain_rate=strain_rate, edges=edges, adj=adj, k_bend=k_bend, mass=mass, dt=dt,
                      nbr_start=start, nbr_num=num, nbr_list=lst, pairs=np.zeros((0, 4), dtype=np.int32),
                      pair_x=None, dhat=h, kappa=0.0, rhs=rhs, seeds=dict(jitter=seed_jitter, rhs=seed_rhs),
                      part_id=None if part_id is None else np.ascontiguousarray(part_id, dtype=np.int32))


@dataclasses.dataclass
class ContactGeometry:
    """A surface for contact detection: positions, rest positions, triangles,
    unique edges and surface vertices (all vertices of a sheet), with the
    barrier parameters (dhat is a distance; the reference's dHat = dhat^2)."""

    name: str
    x: np.ndarray          # (n, 3)
    rest: np.ndarray       # (n, 3)
    faces: np.ndarray      # (m, 3) int32
    edges: np.ndarray      # (k, 2) int32
    surf: np.ndarray       # (s,) int32
    dhat: float
    kappa: float
    body_id: np.ndarray | None = None

    @property
    def n(self) -> int:
        return self.x.shape[0]


def _sheet(N, spacing, rng, angle, shift, z, zjit, xyjit):
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    order = np.argsort(morton2(ii, jj), kind="stable")
    relabel = np.empty_like(order)
    relabel[order] = np.arange(order.size)
    lat = np.stack([ii[order], jj[order]], axis=1).astype(np.float64) * spacing
    lat -= lat.mean(axis=0)
    c, s_ = np.cos(angle), np.sin(angle)
    rest = np.zeros((N * N, 3))
    rest[:, 0] = c * lat[:, 0] - s_ * lat[:, 1] + shift[0]
    rest[:, 1] = s_ * lat[:, 0] + c * lat[:, 1] + shift[1]
    rest[:, 2] = z
    x = rest.copy()
    x[:, :2] += rng.uniform(-xyjit, xyjit, size=(N * N, 2))
    x[:, 2] += rng.uniform(-zjit, zjit, size=N * N)
    qi, qj = np.meshgrid(np.arange(N - 1), np.arange(N - 1), indexing="ij")
    qi, qj = qi.ravel(), qj.ravel()
    vid = lambda i, j: relabel[i * N + j]
    v00, v10, v01, v11 = vid(qi, qj), vid(qi + 1, qj), vid(qi, qj + 1), vid(qi + 1, qj + 1)
    tris = np.concatenate([np.stack([v00, v10, v11], axis=1), np.stack([v00, v11, v01], axis=1)])
    return x, rest, tris.astype(np.int32)


def contact_sheets(N: int = 24, *, spacing=1e-3, dhat=4e-4, gap=2.5e-4, kappa=1e6, seed=5, angle=0.12,
                   body_ids=False) -> ContactGeometry:
    """Two N x N sheets, the second turned by `angle` and shifted, `gap` above
    the first with seeded jitter: vertices over triangles, vertices near
    edges and vertices, and crossing edges all come within dhat, so every
    pair type occurs.  With body_ids each sheet is its own body (no
    self-contact within a sheet)."""
    rng = np.random.default_rng(seed)
    xa, ra, ta = _sheet(N, spacing, rng, 0.0, (0.0, 0.0), 0.0, 0.35 * gap, 0.08 * spacing)
    xb, rb, tb = _sheet(N, spacing, rng, angle, (0.31 * spacing, 0.17 * spacing), gap, 0.35 * gap, 0.08 * spacing)
    n1 = N * N
    x = np.concatenate([xa, xb])
    rest = np.concatenate([ra, rb])
    faces = np.concatenate([ta, tb + n1]).astype(np.int32)
    edges = _tri_edges_csr(2 * n1, faces)[0]
    body = np.repeat(np.arange(2, dtype=np.int32), n1) if body_ids else None
    return ContactGeometry(name=f"sheets{N}", x=x, rest=rest, faces=faces, edges=edges,
                           surf=np.arange(2 * n1, dtype=np.int32), dhat=dhat, kappa=kappa, body_id=body)


def cloth_contact_scene(N: int = 24, *, gap=2.5e-4, dhat=4e-4, kappa=1e6, angle=0.12, seed=9) -> ClothScene:
    """Two draped N x N cloth sheets, the second turned by `angle`, shifted
    and `gap` above the first (same drape), with seeded jitter: their contacts
    (every pair type) are detected on each solve from the contact surface
    (all triangles, all edges, all vertices)."""
    a = cloth_grid_scene(f"sheetA{N}", N, n_pairs=0, seed_jitter=seed, seed_rhs=seed + 1)
    b = cloth_grid_scene(f"sheetB{N}", N, n_pairs=0, seed_jitter=seed + 2, seed_rhs=seed + 3)
    n1 = a.n
    rng = np.random.default_rng(seed + 4)
    side = (N - 1) * 1e-3
    c, s_ = np.cos(angle), np.sin(angle)
    xb = b.x.copy()
    xb[:, 0], xb[:, 1] = c * b.x[:, 0] - s_ * b.x[:, 1] + 0.31e-3, s_ * b.x[:, 0] + c * b.x[:, 1] + 0.17e-3
    xb[:, 2] = -(xb[:, 0] ** 2 + xb[:, 1] ** 2) / (2.0 * side) + gap + rng.uniform(-0.35 * gap, 0.35 * gap, n1)
    adj_b = np.where(b.adj >= 0, b.adj + n1, -1).astype(np.int32)
    tris = np.concatenate([a.tris, b.tris + n1]).astype(np.int32)
    edges = np.concatenate([a.edges, b.edges + n1]).astype(np.int32)
    x = np.concatenate([a.x, xb])
    rest = np.concatenate([a.rest, b.rest])
    surf = ContactGeometry(name=f"cloth_contact{N}", x=x, rest=rest, faces=tris, edges=edges,
                           surf=np.arange(2 * n1, dtype=np.int32), dhat=dhat, kappa=kappa)
    return ClothScene(name=f"cloth_contact{N}", x=x, rest=rest, tris=tris,
                      dm_inv2=np.concatenate([a.dm_inv2, b.dm_inv2]), area=np.concatenate([a.area, b.area]),
                      k_stretch=a.k_stretch, k_shear=a.k_shear, strain_rate=a.strain_rate, edges=edges,
                      adj=np.concatenate([a.adj, adj_b]), k_bend=a.k_bend, mass=np.concatenate([a.mass, b.mass]),
                      dt=a.dt, nbr_start=np.concatenate([a.nbr_start, b.nbr_start + a.nbr_list.size]).astype(np.uint32),
                      nbr_num=np.concatenate([a.nbr_num, b.nbr_num]).astype(np.uint32),
                      nbr_list=np.concatenate([a.nbr_list, b.nbr_list + n1]).astype(np.uint32),
                      pairs=np.zeros((0, 4), np.int32), pair_x=None, dhat=dhat, kappa=kappa,
                      rhs=np.random.default_rng(seed + 5).standard_normal(6 * n1),
                      seeds=dict(jitter=seed, pairs=None), contact_surface=surf)


def scene_cloth_tiny(N: int = 10, n_pairs: int = 16) -> ClothScene:
    """Small draped sheet with contacts, for dense checks."""
    return cloth_grid_scene(f"cloth_tiny{N}", N, n_pairs=n_pairs)


def scene_c4(n_pairs: int = 200_000) -> ClothScene:
    """1024 x 1024 sheet, 1 mm spacing, E = 1e6 Pa, k_bend = 1e-4, 200k pairs."""
    return cloth_grid_scene("c4_cloth1024", 1024, n_pairs=n_pairs)


@dataclasses.dataclass
class HybridScene(TetScene):
    """A tet mesh plus affine bodies (ABD), numbered as the reference numbers
    them (``GIPC.cu:10329-10330,10482-10483``): system rows are
    [4 per body | FEM vertices], and global point ids are [body points | FEM
    vertices].  Contact pairs use global point ids; ``tets``, ``nbr_*`` and
    ``mass`` use FEM-local ids.  Each body carries q = [p, a1, a2, a3], the
    previous-step prediction q_tilde, its dyadic mass (m, m xb, m xb xb^T) over
    its mass points, its volume and a fixed flag
    (``abd_system/abd_jacobi_matrix.h:162-199``)."""

    n_bodies: int = 0
    q: np.ndarray | None = None            # (nb, 12)
    q_tilde: np.ndarray | None = None      # (nb, 12)
    mass10: np.ndarray | None = None       # (nb, 10): m, m xb (3), m xb xb^T (xx, xy, xz, yy, yz, zz)
    body_volume: np.ndarray | None = None  # (nb,)
    kappa_abd: float = 1e8
    body_fixed: np.ndarray | None = None   # (nb,) int32
    n_abd_points: int = 0
    point_body: np.ndarray | None = None   # (n_abd_points,) int32
    x_bar: np.ndarray | None = None        # (n_abd_points, 3) rest coordinates in the body frame
    motor_speed: float = 31.4              # rad/s (abd_system_parms.h:13)
    motor_strength: float = 10.0           # multiple of the body mass block (abd_system_parms.h:14)

    @property
    def n_sys(self) -> int:
        return 4 * self.n_bodies + self.n

    @property
    def n_triplets(self) -> int:
        """Upper bound: 16 vertex blocks per pair, each up to 16 body blocks."""
        return 256 * self.n_pairs + 10 * self.n_bodies + 10 * self.n_tets + self.n


def hybrid_scene(name: str, N: int, n_bodies: int, n_pairs: int, *, E=1e5, body_side_cells=2.0,
                 kappa_abd=1e8, density=1e3, n_fixed=2, n_motor=0, spacing=0.01, dhat=1e-3, kappa=1e6, dt=0.01,
                 seed_bodies=4, seed_pairs=3, seed_rhs=1) -> HybridScene:
    """N^3 Neo-Hookean grid (Kuhn split, Morton order) with n_bodies affine
    cubes resting above its top face; n_pairs seeded point-triangle pairs
    between body corner points and the grid's top-face triangles, each with
    its own local geometry at d^2 ~ U(0, dhat^2)."""
    base = tet_grid_scene(name, N, E=E, spacing=spacing, n_pairs=0, dt=dt)
    rng = np.random.default_rng(seed_bodies)
    side = body_side_cells * spacing
    corners = np.array([[sx, sy, sz] for sx in (-0.5, 0.5) for sy in (-0.5, 0.5) for sz in (-0.5, 0.5)]) * side
    nb = n_bodies
    top = base.x[:, 2].max()
    grid_len = (N - 1) * spacing
    cols_ = int(np.ceil(np.sqrt(nb)))
    q = np.zeros((nb, 12))
    for b in range(nb):
        cx = (0.5 + b % cols_) / cols_ * grid_len
        cy = (0.5 + b // cols_) / cols_ * grid_len
        q[b, 0:3] = [cx, cy, top + 0.5 * side + 1e-4]
        A = np.eye(3) + 0.01 * rng.standard_normal((3, 3))
        q[b, 3:12] = A.reshape(-1)           # rows a1, a2, a3
    q_tilde = q + 1e-4 * rng.standard_normal(q.shape)
    m = density * side ** 3
    mass10 = np.zeros((nb, 10))
    mass10[:, 0] = m
    d = m * (0.5 * side) ** 2             # sum over 8 corners of (m/8) xb xb^T = m (side/2)^2 I
    mass10[:, 4] = d
    mass10[:, 7] = d
    mass10[:, 9] = d
    body_volume = np.full(nb, side ** 3)
    body_fixed = np.zeros(nb, np.int32)   # BodyBoundaryType: 0 free, 1 fixed, 2 motor
    body_fixed[:n_fixed] = 1
    body_fixed[n_fixed:n_fixed + n_motor] = 2
    n_abd_points = 8 * nb
    point_body = np.repeat(np.arange(nb), 8).astype(np.int32)
    x_bar = np.tile(corners, (nb, 1))
    # body point positions x = p + A xb
    A = q[:, 3:12].reshape(nb, 3, 3)
    xpts = q[point_body, 0:3] + np.einsum("bij,bj->bi", A[point_body], x_bar)
    # pairs: body point vs FEM top-face triangle (global ids)
    tris = _surface_triangles(base.tets)
    ztop = base.rest[:, 2].max()
    tris = tris[np.all(base.rest[tris][:, :, 2] >= ztop - 1e-12, axis=1)]
    rp = np.random.default_rng(seed_pairs)
    pv = rp.integers(0, n_abd_points, size=n_pairs)
    ti = rp.integers(0, tris.shape[0], size=n_pairs)
    t = tris[ti]
    d2 = np.maximum(rp.uniform(0.0, 1.0, size=n_pairs) * dhat * dhat, 1e-12 * dhat * dhat)
    bary = rp.uniform(0.1, 1.0, size=(n_pairs, 3))
    bary /= bary.sum(axis=1, keepdims=True)
    t0, t1, t2 = base.x[t[:, 0]], base.x[t[:, 1]], base.x[t[:, 2]]
    nrm = np.cross(t1 - t0, t2 - t0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    foot = bary[:, 0:1] * t0 + bary[:, 1:2] * t1 + bary[:, 2:3] * t2
    ppos = foot + np.sqrt(d2)[:, None] * nrm
    pairs = np.stack([-pv - 1, t[:, 0] + n_abd_points, t[:, 1] + n_abd_points, t[:, 2] + n_abd_points],
                     axis=1).astype(np.int32)
    pair_x = np.stack([ppos, t0, t1, t2], axis=1)
    n_sys = 4 * nb + base.n
    rhs = np.random.default_rng(seed_rhs).standard_normal(3 * n_sys)
    fields = {f.name: getattr(base, f.name) for f in dataclasses.fields(TetScene)}
    fields.update(name=name, pairs=pairs, pair_x=pair_x, dhat=dhat, kappa=kappa, rhs=rhs,
                  seeds=dict(base.seeds, bodies=seed_bodies, pairs=seed_pairs))
    return HybridScene(**fields, n_bodies=nb, q=q, q_tilde=q_tilde, mass10=mass10, body_volume=body_volume,
                       kappa_abd=kappa_abd, body_fixed=body_fixed, n_abd_points=n_abd_points,
                       point_body=point_body, x_bar=x_bar)


def scene_hybrid_tiny(N: int = 8, n_bodies: int = 6, n_pairs: int = 40) -> HybridScene:
    """Small hybrid scene for dense checks."""
    return hybrid_scene(f"hybrid_tiny{N}", N, n_bodies, n_pairs)


def scene_c2(n_pairs: int = 30_000) -> HybridScene:
    """48^3 Neo-Hookean grid (E = 1e5) + 200 affine bodies (kappa 1e8), 30k pairs."""
    return hybrid_scene("c2_hybrid48", 48, 200, n_pairs)


def with_dirichlet(scene, which="bottom"):
    """A copy of the scene whose lowest layer of vertices (z within 1e-9 of
    the minimum rest z; ``which="bottom"``) is Dirichlet, as a scene's
    BoundaryType marks pinned vertices."""
    z = scene.rest[:, 2]
    bt = (z <= z.min() + 1e-9).astype(np.int32)
    return dataclasses.replace(scene, boundary_type=bt)
