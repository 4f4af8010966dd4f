"""CEMAS partitions on the GPU: the system solves in the slot space of a
vertex partition (scene part_id) and is compared with the oracle's CEMAS
operator (oracle/cemas.py, pinned in tests/test_oracle_cemas.py):

* the MAS apply of a partitioned system equals the operator on the
  reference's own METIS partition of cloth_high (Assets/sorted_mesh);
* PCG under the reference stop rule takes the oracle's iteration count
  (within one) and reaches its solution;
* the partition part = vertex // B reproduces the unpartitioned solve;
* a detected-contact scene with a partition from sgipc_partition_vertices
  solves to the unpartitioned solution;
* on bunny2 (19193 vertices) the METIS parts, this library's partitioner
  and no partition (consecutive vertex domains) all converge; their
  iteration counts are printed.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sorted16_meshes.npz")


@pytest.fixture(scope="module")
def meshes():
    return np.load(GOLD)


def _cloth_high(pkg, z, part=True, **kw):
    return pkg.scenes.cloth_mesh_scene("cloth_high", z["cloth_x"], z["cloth_el"].astype(np.int32),
                                       part_id=z["cloth_part"].astype(np.int32) if part else None, **kw)


def _slot_apply(pkg, ls, r, r2p, n_slots):
    lib, check = pkg._lib.lib, pkg._lib.check
    rs = np.zeros((n_slots, 3))
    rs[r2p] = r.reshape(-1, 3)
    d_r = ls.ctx.to_device(rs.reshape(-1), np.float64)
    d_z = ls.ctx.empty(3 * n_slots, np.float64)
    check(lib().sgipc_mas_preconditioning(ls.sys.mas_handle(), d_r.vp, d_z.vp), ls.ctx.h)
    return d_z.download().reshape(-1, 3)[r2p].reshape(-1)


@pytest.mark.gpu
def test_cemas_apply_and_pcg_match_oracle(pkg, orc, meshes):
    s = _cloth_high(pkg, meshes, density=2e4)
    print(f"[cemas] cloth_high n={s.n} parts={s.part_id.max() + 1}", flush=True)
    ls = pkg.LinearSystem(s, bank_size=16)
    try:
        res = ls.solve(tol=1e-4, stop_mode=pkg.STOP_REF_RZ)
        vals, rows, cols = orc.assemble(s)
        op = orc.cemas.CemasOperator(s.n, (s.nbr_start, s.nbr_num, s.nbr_list), s.part_id, 16, vals, rows, cols)
        n_slots = op.p2r.size
        rng = np.random.default_rng(2)
        for _ in range(2):
            r = rng.standard_normal(3 * s.n)
            z_g = _slot_apply(pkg, ls, r, op.r2p, n_slots)
            z_o = op.apply(r)
            err = np.abs(z_g - z_o).max() / np.abs(z_o).max()
            print(f"[cemas] apply rel err {err:.2e}", flush=True)
            assert err < 1e-9
        A = orc.cemas._bcoo_to_csr(vals, rows, cols, s.n)
        x_o, it_o = orc.cemas.pcg(A, s.rhs, op.apply, tol=1e-4, stop_mode=0)
        dx = np.linalg.norm(res.x - x_o) / np.linalg.norm(x_o)
        print(f"[cemas] iterations gpu {res.iterations} oracle {it_o}, |dx|/|x| {dx:.2e}", flush=True)
        assert abs(res.iterations - it_o) <= 1
        assert dx < 1e-6
        # the same system without the partition has the same solution
        ls2 = pkg.LinearSystem(_cloth_high(pkg, meshes, part=False, density=2e4), bank_size=16)
        try:
            r1 = ls.solve(tol=1e-10)
            r2 = ls2.solve(tol=1e-10)
            print(f"[cemas] tight solve iterations partitioned {r1.iterations} consecutive {r2.iterations}",
                  flush=True)
            assert np.linalg.norm(r1.x - r2.x) / np.linalg.norm(r2.x) < 1e-8
        finally:
            ls2.close()
    finally:
        ls.close()


@pytest.mark.gpu
@pytest.mark.parametrize("N,B", [(12, 16), (11, 16), (11, 32)])
def test_identity_partition_reproduces_consecutive_domains(pkg, N, B):
    s = pkg.scenes.tet_grid_scene(f"grid{N}", N, E_loguniform=(1e4, 1e9), n_pairs=8 * N)
    s2 = pkg.scenes.TetScene(**{**s.__dict__, "part_id": (np.arange(s.n) // B).astype(np.int32)})
    print(f"[cemas] identity partition N={N} B={B}", flush=True)
    a, b = pkg.LinearSystem(s, bank_size=B), pkg.LinearSystem(s2, bank_size=B)
    try:
        for prec in ("fp64", "fp32_mas"):
            ra = a.solve(tol=1e-8, precision=prec)
            rb = b.solve(tol=1e-8, precision=prec)
            print(f"[cemas]   {prec}: {ra.iterations} vs {rb.iterations}", flush=True)
            assert ra.iterations == rb.iterations
            assert np.linalg.norm(ra.x - rb.x) / np.linalg.norm(ra.x) < 1e-9
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
def test_partitioned_contact_scene(pkg, orc):
    s = pkg.scenes.cloth_contact_scene(24, kappa=100.0)
    part = pkg.partition.partition_vertices(s.nbr_start, s.nbr_num, s.nbr_list, 16)
    sp = pkg.scenes.ClothScene(**{**s.__dict__, "part_id": part})
    print(f"[cemas] contact scene n={s.n} parts={part.max() + 1}", flush=True)
    a, b = pkg.LinearSystem(s, bank_size=16), pkg.LinearSystem(sp, bank_size=16)
    try:
        ra = a.solve(tol=1e-9)
        rb = b.solve(tol=1e-9)
        pairs, _ = b.sys.contacts()
        print(f"[cemas]   pairs {pairs.shape[0]}, iterations {ra.iterations} vs {rb.iterations}", flush=True)
        assert pairs.shape[0] > 0
        assert np.linalg.norm(ra.x - rb.x) / np.linalg.norm(ra.x) < 1e-7
        # reference stop rule: the oracle's CEMAS PCG with the detected pairs
        rr = b.solve(tol=1e-4, stop_mode=pkg.STOP_REF_RZ)
        vals, rows, cols = orc.assemble(s)
        maps = pkg.partition.mas_partition_maps(part, 16)
        op = orc.cemas.CemasOperator(s.n, (s.nbr_start, s.nbr_num, s.nbr_list), part, 16, vals, rows, cols,
                                     pairs, maps=maps)
        A = orc.cemas._bcoo_to_csr(vals, rows, cols, s.n)
        _, it_o = orc.cemas.pcg(A, s.rhs, op.apply, tol=1e-4, stop_mode=0)
        print(f"[cemas]   ref-rule iterations gpu {rr.iterations} oracle {it_o}", flush=True)
        assert abs(rr.iterations - it_o) <= 1
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
def test_bunny_partitions(pkg, meshes):
    x, tets, metis = meshes["bunny_x"], meshes["bunny_el"].astype(np.int64), meshes["bunny_part"].astype(np.int32)
    base = pkg.scenes.tet_mesh_scene("bunny2", x, tets, E_loguniform=(1e4, 1e8))
    ours = pkg.partition.partition_vertices(base.nbr_start, base.nbr_num, base.nbr_list, 16)
    out = {}
    for name, part in (("consecutive", None), ("metis", metis), ("library", ours)):
        s = pkg.scenes.TetScene(**{**base.__dict__, "part_id": part})
        ls = pkg.LinearSystem(s, bank_size=16)
        try:
            r0 = ls.solve(tol=1e-4, stop_mode=pkg.STOP_REF_RZ)
            r1 = ls.solve(tol=1e-5)
            out[name] = (r0.iterations, r1.iterations, r1.true_rel_residual)
            print(f"[cemas] bunny2 {name}: ref-rule {r0.iterations} its, true-residual 1e-5 {r1.iterations} its "
                  f"(res {r1.true_rel_residual:.2e})", flush=True)
            assert r1.status == 0 and r1.true_rel_residual <= 1e-5
        finally:
            ls.close()
    assert out["metis"][1] <= out["consecutive"][1] and out["library"][1] <= 1.15 * out["metis"][1]
