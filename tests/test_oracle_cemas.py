"""CEMAS partitions on the CPU: the slot maps and the partitioner of the
library's host functions, and the pins of the oracle's CEMAS operator
(oracle/cemas.py).

* slot maps: the library's sgipc_mas_partition_maps equals the reference's
  two loops of set_mas_partition (restated as written) on the reference's
  own METIS-sorted part files (cube, cloth_high, bunny2), and gives a valid
  map (bank p holds part p's vertices in vertex order) for unsorted ids;
* operator: with the partition part = vertex // B it equals the
  identity-map C oracle's MAS apply (pinned on its own); on a one-part
  system it is A^-1 + P1 (P1^T A P1)^-1 P1^T (two levels, one domain each);
  for a real partition it is symmetric positive definite;
* partitioner: parts of 1..B vertices, every vertex placed,
  deterministic, and an edge cut within 20 % of the reference's METIS
  partitions of the same meshes.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sorted16_meshes.npz")


@pytest.fixture(scope="module")
def meshes():
    return np.load(GOLD)


def _nbr(pkg, x, el):
    if el.shape[1] == 4:
        st, nn, nl, _ = pkg.scenes._edges_csr(x.shape[0], el.astype(np.int64))
    else:
        _, _, st, nn, nl = pkg.scenes._tri_edges_csr(x.shape[0], el.astype(np.int32))
    return st, nn, nl


@pytest.mark.parametrize("name", ["cube", "cloth", "bunny"])
def test_slot_maps_equal_reference_loops(pkg, orc, meshes, name):
    part = meshes[f"{name}_part"].astype(np.int32)
    p2r, r2p = pkg.partition.mas_partition_maps(part, 16)
    q2r, s2p = orc.cemas.partition_maps_ref(part, 16)
    np.testing.assert_array_equal(p2r, q2r)
    np.testing.assert_array_equal(r2p, s2p)


def test_slot_maps_unsorted_and_errors(pkg):
    rng = np.random.default_rng(4)
    part = rng.integers(0, 40, 300).astype(np.int32)
    part[:40] = np.arange(40)  # every part used
    sizes = np.bincount(part)
    part = part[np.argsort(rng.random(300))]
    p2r, r2p = pkg.partition.mas_partition_maps(part, 16) if sizes.max() <= 16 else (None, None)
    assert p2r.size == 40 * 16
    np.testing.assert_array_equal(p2r[r2p], np.arange(300))
    assert (r2p // 16 == part).all()
    for p in range(40):  # vertex order inside a bank
        v = np.flatnonzero(part == p)
        np.testing.assert_array_equal(r2p[v], 16 * p + np.arange(v.size))
    assert (p2r >= 0).sum() == 300
    with pytest.raises(pkg.SgipcError, match="CAPACITY"):
        pkg.partition.mas_partition_maps(np.zeros(17, np.int32), 16)
    with pytest.raises(pkg.SgipcError, match="BAD_INDEX"):
        pkg.partition.mas_partition_maps(np.array([0, -1], np.int32), 16)


@pytest.mark.parametrize("kind", ["tets", "cloth"])
@pytest.mark.parametrize("B", [16, 32])
def test_identity_partition_is_the_identity_map_mas(pkg, orc, kind, B):
    # (a heavy cloth keeps the coarsest Galerkin operators well conditioned:
    # the two inverses, Gauss-Jordan without pivoting and LU, then agree)
    s = (pkg.scenes.scene_tiny(N=7, n_pairs=20) if kind == "tets" else
         pkg.scenes.cloth_grid_scene("cloth11_heavy", 11, n_pairs=16, density=2e5))
    vals, rows, cols = orc.assemble(s)
    pairs = orc.scene_pairs(s)
    m = orc.Mas(s, B)
    m.build(vals, rows, cols, 0, pairs)
    op = orc.cemas.CemasOperator(s.n, (s.nbr_start, s.nbr_num, s.nbr_list), np.arange(s.n) // B, B, vals, rows,
                                 cols, pairs)
    rng = np.random.default_rng(1)
    for _ in range(3):
        r = rng.standard_normal(3 * s.n)
        z0, z1 = m.apply(r), op.apply(r)
        # Gauss-Jordan vs LU on coarse operators of condition up to ~2e7
        np.testing.assert_allclose(z1, z0, rtol=0, atol=1e-8 * np.abs(z0).max())
    m.close()


def test_one_part_closed_form(pkg, orc):
    s = pkg.scenes.scene_tiny(N=2, n_pairs=0)  # 8 vertices, one bank
    vals, rows, cols = orc.assemble(s)
    op = orc.cemas.CemasOperator(s.n, (s.nbr_start, s.nbr_num, s.nbr_list), np.zeros(s.n, np.int32), 16, vals,
                                 rows, cols)
    A = orc.cemas._bcoo_to_csr(vals, rows, cols, s.n).toarray()
    assert len(op.levels) == 2 and op.counts[1] == 1
    P1 = np.tile(np.eye(3), (s.n, 1))
    want = np.linalg.inv(A) + P1 @ np.linalg.inv(P1.T @ A @ P1) @ P1.T
    np.testing.assert_allclose(op.dense(), want, rtol=0, atol=1e-10 * np.abs(want).max())


def test_partitioned_operator_is_spd(pkg, orc):
    s = pkg.scenes.cloth_grid_scene("cloth12_heavy", 12, n_pairs=20, density=2e5)
    part = pkg.partition.partition_vertices(s.nbr_start, s.nbr_num, s.nbr_list, 16)
    order, sp, tris = pkg.mesh_io.sort_by_part(part, s.tris)
    assert np.all(np.diff(sp) >= 0)
    vals, rows, cols = orc.assemble(s)
    maps = pkg.partition.mas_partition_maps(part, 16)
    op = orc.cemas.CemasOperator(s.n, (s.nbr_start, s.nbr_num, s.nbr_list), part, 16, vals, rows, cols,
                                 orc.scene_pairs(s), maps=maps)
    assert op.counts[0] == 16 * (part.max() + 1) and len(op.levels) >= 2
    M = op.dense()
    np.testing.assert_allclose(M, M.T, rtol=0, atol=1e-9 * np.abs(M).max())  # inverses of condition ~1e7
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


@pytest.mark.parametrize("name,bound", [("cloth", 1.2), ("bunny", 1.2)])
def test_partitioner_against_metis(pkg, meshes, name, bound):
    x, el, metis = meshes[f"{name}_x"], meshes[f"{name}_el"], meshes[f"{name}_part"].astype(np.int32)
    st, nn, nl = _nbr(pkg, x, el)
    part = pkg.partition.partition_vertices(st, nn, nl, 16)
    assert part.min() == 0 and np.bincount(part).max() <= 16 and np.bincount(part).min() >= 1
    np.testing.assert_array_equal(part, pkg.partition.partition_vertices(st, nn, nl, 16))
    cut, ref = pkg.partition.edge_cut(part, st, nn, nl), pkg.partition.edge_cut(metis, st, nn, nl)
    assert cut <= bound * ref, (cut, ref)
    assert part.max() + 1 <= 1.1 * (metis.max() + 1)


def test_part_file_round_trip(pkg, meshes, tmp_path):
    part = meshes["cloth_part"].astype(np.int32)
    f = tmp_path / "m.part"
    pkg.mesh_io.write_part(str(f), part)
    np.testing.assert_array_equal(pkg.mesh_io.read_part(str(f), part.size), part)
    np.testing.assert_array_equal(pkg.mesh_io.read_part(str(f), part.size, part_offset=7), part + 7)
    assert pkg.mesh_io.next_part_offset(part) == 282
    with pytest.raises(ValueError):
        pkg.mesh_io.read_part(str(f), part.size + 1)


@pytest.mark.parametrize("name", ["cloth", "bunny"])
def test_partitioner_iterations_against_metis(pkg, orc, meshes, name):
    """CEMAS PCG iterations (oracle operator, true residual 1e-5) with the
    library's partition within 15 % of those with the reference's METIS
    partition of the same mesh."""
    x, el, metis = meshes[f"{name}_x"], meshes[f"{name}_el"], meshes[f"{name}_part"].astype(np.int32)
    if name == "cloth":
        s = pkg.scenes.cloth_mesh_scene("cloth_high", x, el.astype(np.int32), density=2e4)
    else:
        s = pkg.scenes.tet_mesh_scene("bunny2", x, el.astype(np.int64), E_loguniform=(1e4, 1e8))
    vals, rows, cols = orc.assemble(s)
    A = orc.cemas._bcoo_to_csr(vals, rows, cols, s.n)
    its = {}
    for lab, part in (("metis", metis),
                      ("library", pkg.partition.partition_vertices(s.nbr_start, s.nbr_num, s.nbr_list, 16))):
        op = orc.cemas.CemasOperator(s.n, (s.nbr_start, s.nbr_num, s.nbr_list), part, 16, vals, rows, cols,
                                     maps=pkg.partition.mas_partition_maps(part, 16))
        its[lab] = orc.cemas.pcg(A, s.rhs, op.apply, tol=1e-5, stop_mode=1)[1]
    assert its["library"] <= 1.15 * its["metis"], its
