#!/usr/bin/env python3
"""Benchmark: ms per Newton linear solve to true relative residual 1e-5 on the
64^3 stiff-contact scene, plus SpMV HBM bandwidth against MI355X peak.

One step = one whole Newton linear solve on the GPU with every input already
resident in HBM: barrier + tet Hessian assembly into the triplet arena,
conversion to unique upper BCOO, MAS rebuild (hierarchy, Galerkin, FP64
inverses, dense FP64 level-2 inverse) and PCG until ||b - A x|| <= 1e-5 ||b||
(checked in FP64).

Default configuration: MAS B=32 with the level-1 V(1,1) cycle as the coarse
solve of a symmetric two-grid cycle on level 0 (level-0 domain solves
weighted 0.4 before and after it); mixed
precision -- the SpMV reads FP32 matrix values and the MAS applies FP32
inverse blocks, while vectors, dot products, alpha/beta and the MAS
factorisations stay FP64, inside FP64 refinement (the recursive residual is
replaced by the FP64 one every 8 iterations, and the final residual is
checked in FP64); the learned coarse level in automatic mode (trained before
the timed solves on a separately seeded scene; used only when it lowers the
iteration count).  The plain FP64 additive-MAS solve is timed in the same
run.

Prints one JSON line (rank 0).  Only one GPU is used: ``--gpus N`` with N > 1
exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def spmv_bytes(nnz, n_rows, f32_values=False):
    """Algorithmic bytes of one symmetric SpMV over `nnz` unique upper blocks:
    every upper block's values + row/col index once, x read once and y written
    once, vectors in FP64: nb*80 + 48*n_rows with FP64 values, nb*44 + 48*n_rows
    with FP32 values (the mixed-precision PCG's SpMV)."""
    return nnz * ((36 if f32_values else 72) + 8) + 2 * 24 * n_rows


def layout_bytes(stored, n_rows, f32_values=False):
    """Bytes the tiled layout actually streams: every stored block (upper
    blocks + cross-tile transposed copies) with its 4-byte column word, the
    per-row slice metadata, x once and y once."""
    v = 36 if f32_values else 72
    return stored * (v + 4) + 12 * n_rows + 2 * 24 * n_rows


class _Ptr:
    """Device pointer (ctypes) in the form the Python handles take."""

    def __init__(self, p):
        self.ptr = p.value


def measured_traffic(kernel_key):
    """HBM bytes per SpMV launch from the FETCH_SIZE / WRITE_SIZE profile
    summary committed under profiles/ (None when absent)."""
    path = os.path.join(ROOT, "profiles", "spmv_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", default="fp32_ir", choices=["fp64", "fp32_mas", "fp32_ir"])
    ap.add_argument("--bank", type=int, default=32, choices=[16, 32])
    ap.add_argument("--pairs", type=int, default=50_000)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--coarse", type=int, default=2, choices=[0, 1, 2],
                    help="MAS coarse levels: 0 additive (the reference), 1 V(1,1) cycle on level 1, "
                         "2 that cycle inside a two-grid cycle on level 0")
    ap.add_argument("--level0-half", type=int, default=1, choices=[0, 1],
                    help="--coarse 2 float applies: level-0 domain inverses as unit-diagonal scaled FP16 (1) or float (0)")
    ap.add_argument("--omega", type=float, default=None,
                    help="level-0 smoothing weight of --coarse 2 (library default 0.4)")
    ap.add_argument("--learned", type=int, default=2, choices=[0, 1, 2],
                    help="learned coarse level: 0 off, 1 always, 2 automatic (falls back to plain MAS when worse); "
                         "the network is trained on a separately seeded scene before the timed solves")
    ap.add_argument("--train-samples", type=int, default=600)
    ap.add_argument("--train-epochs", type=int, default=10)
    ap.add_argument("--compare-plain", type=int, default=3,
                    help="steps of plain FP64 additive-MAS (B=32) solves timed in the same run for the speedup")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1],
                    help="1: PCG replays its full chunks of iterations from a captured HIP graph")
    ap.add_argument("--ir-replace", type=int, default=8,
                    help="fp32_ir: iterations between replacements of the recursive residual by the FP64 one")
    ap.add_argument("--fuse", type=int, default=1, choices=[0, 1, 2],
                    help="0: separate x/r and p update kernels; 1: x/r update inside the first MAS kernel; "
                         "2: also the p update inside the SpMV")
    ap.add_argument("--cpu-reps", type=int, default=3,
                    help="full CPU-baseline solves timed (the median is reported)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (0: every core this process may use)")
    ap.add_argument("--only-cpu-baseline", action="store_true",
                    help="build the scene and time the CPU baseline only (no GPU work)")
    ap.add_argument("--only-spmv", action="store_true", help="time the SpMV kernel only (for profiling runs)")
    ap.add_argument("--scene", default="c1", choices=["c1", "c2", "c4"],
                    help="c1: the 64^3 stiff-contact scene (the headline line, with roofline and CPU baseline); "
                         "c2: 48^3 grid + 200 affine bodies + 30k pairs; c4: cloth 1024^2 + 200k pairs "
                         "(solve times only)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import __graft_entry__ as ge
    pkg = ge.load_package()
    if args.scene != "c1":
        bench_scene(pkg, args)
        return

    t0 = time.perf_counter()
    if args.grid == 64:
        scene = pkg.scenes.scene_c1(n_pairs=args.pairs)
    else:
        scene = pkg.scenes.tet_grid_scene(f"grid{args.grid}", args.grid, E_loguniform=(1e4, 1e9),
                                          n_pairs=args.pairs)
    log(f"scene {scene.name}: n={scene.n} tets={scene.n_tets} pairs={scene.n_pairs} "
        f"({time.perf_counter() - t0:.1f}s to generate)")

    if args.only_cpu_baseline:
        print(json.dumps({"cpu_baseline": cpu_baseline(ge, scene, 32, args.cpu_reps, args.cpu_threads)}), flush=True)
        return

    ls = pkg.LinearSystem(scene, bank_size=args.bank, coarse_mode=args.coarse)
    if args.omega is not None:
        ls.set_smoothing_weight(args.omega)
    ls.set_level0_half(args.level0_half)
    ctx = ls.ctx
    L = pkg._lib.lib()
    import ctypes

    out = {}
    learned_info = None
    net = None
    if args.learned and not args.only_spmv:
        learned_info, net = train_learned(pkg, ls, args)
    if not args.only_spmv:
        # ---- warm-up solves (also sizes every buffer)
        for w in range(args.warmup):
            r = ls.solve(tol=1e-5, precision=args.precision, fuse=args.fuse,
                         ir_replace_every=args.ir_replace, graph=bool(args.graph))
            log(f"warmup {w}: {r.timing['ms_total']:.2f} ms iters={r.iterations} "
                f"outer={r.outer_iterations} res={r.true_rel_residual:.2e}")
        # ---- timed steps
        ctx.synchronize()
        times, iters, res_list, phases = [], [], [], []
        tt0 = time.perf_counter()
        for k in range(args.steps):
            s0 = time.perf_counter()
            r = ls.solve(tol=1e-5, precision=args.precision, fuse=args.fuse,
                         ir_replace_every=args.ir_replace, graph=bool(args.graph))
            times.append((time.perf_counter() - s0) * 1e3)
            iters.append(r.iterations)
            res_list.append(r.true_rel_residual)
            phases.append(r.timing)
            log(f"step {k}: {times[-1]:.2f} ms iters={r.iterations} outer={r.outer_iterations} "
                f"res={r.true_rel_residual:.2e} assemble={r.timing['ms_assemble']:.2f} "
                f"convert={r.timing['ms_convert']:.2f} mas={r.timing['ms_mas_build']:.2f} "
                f"pcg={r.timing['ms_pcg']:.2f} mode={r.coarse_mode} l2exact={int(r.level2_exact)} "
                f"omega={r.omega:.3g}")
        ctx.synchronize()
        total_ms = (time.perf_counter() - tt0) * 1e3
        ms_step = total_ms / args.steps
        assert max(res_list) <= 1e-5, f"solve missed the tolerance: {max(res_list)}"
        out["ms"] = ms_step
        out["iters"] = float(np.mean(iters))
        out["res"] = float(max(res_list))
        out["phases"] = {k: float(np.mean([p[k] for p in phases]))
                         for k in ("ms_assemble", "ms_convert", "ms_mas_build", "ms_pcg")}
        out["nnz"] = int(phases[-1]["nnz"])
        out["n_triplets"] = int(phases[-1]["n_triplets"])
        out["learned_used"] = bool(r.learned)
        # error of the timed configuration's solution against a tight FP64 solve
        # of the same system (plain FP64 SpMV and MAS, ||b-Ax|| <= 1e-10 ||b||)
        x_cfg = r.x
        ref = ls.solve(tol=1e-10, precision="fp64", fuse=args.fuse)
        nrm = float(np.linalg.norm(ref.x))
        out["err_vs_fp64"] = float(np.linalg.norm(x_cfg - ref.x) / nrm) if nrm > 0 else 0.0
        log(f"solution error vs FP64 solve to 1e-10: {out['err_vs_fp64']:.3e} "
            f"(FP64 reference: {ref.iterations} iterations, residual {ref.true_rel_residual:.1e})")

    # ---- SpMV kernel on the assembled matrix, HIP events recorded by the launches themselves
    if args.only_spmv:
        ls.solve(tol=1e-5, precision="fp64", max_iter=2)
    v, rr, cc, nnz = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int()
    pkg._lib.check(L.sgipc_system_bcoo(ls.sys.h, ctypes.byref(v), ctypes.byref(rr), ctypes.byref(cc),
                                       ctypes.byref(nnz)), ctx.h)
    nnz = nnz.value
    n = scene.n
    x = ctx.to_device(np.random.default_rng(9).standard_normal(3 * n))
    avg = ctypes.c_double()
    stored = ctypes.c_int()
    # cold cache: the matrix comes from HBM, as inside PCG (the MAS apply evicts it)
    pkg._lib.check(L.sgipc_spmv_time(ctx.h, pkg.FP64, v, rr, cc, nnz, n, args.spmv_reps, 1, ctypes.byref(avg),
                                     ctypes.byref(stored)), ctx.h)
    spmv_ms = avg.value
    stored = stored.value
    # layout form the solver uses on this matrix (halo: x staged per tile in LDS)
    spm = pkg.SpMat(ctx, _Ptr(v), _Ptr(rr), _Ptr(cc), nnz, n)
    halo_form, halo_max, tl_slots = spm.layout()
    spm.close()
    pkg._lib.check(L.sgipc_spmv_time(ctx.h, pkg.FP64, v, rr, cc, nnz, n, args.spmv_reps, 0, ctypes.byref(avg),
                                     ctypes.byref(ctypes.c_int())), ctx.h)
    spmv_warm_ms = avg.value
    b64 = spmv_bytes(nnz, n)
    gbs64 = b64 / (spmv_ms * 1e-3) / 1e9
    moved64 = layout_bytes(stored, n) / (spmv_ms * 1e-3) / 1e9
    pkg._lib.check(L.sgipc_spmv_time(ctx.h, pkg.FP32_IR, v, rr, cc, nnz, n, args.spmv_reps, 1, ctypes.byref(avg),
                                     ctypes.byref(ctypes.c_int())), ctx.h)
    spmv32_ms = avg.value
    b32 = spmv_bytes(nnz, n, f32_values=True)
    gbs32 = b32 / (spmv32_ms * 1e-3) / 1e9
    moved32 = layout_bytes(stored, n, f32_values=True) / (spmv32_ms * 1e-3) / 1e9
    log(f"SpMV fp64 (cold cache): {spmv_ms * 1e3:.1f} us/launch, {nnz} upper blocks, {b64 / 1e6:.1f} MB algorithmic "
        f"-> {gbs64:.0f} GB/s ({gbs64 / HBM_PEAK_GBS:.1%} of 8 TB/s); layout stores {stored} blocks, streams "
        f"{moved64:.0f} GB/s; back-to-back (matrix partly in the 256 MB Infinity Cache): {spmv_warm_ms * 1e3:.1f} us; "
        f"FP32 values (FP64 vectors) cold: {spmv32_ms * 1e3:.1f} us, {gbs32:.0f} GB/s algorithmic")

    # MAS apply timing
    mas_ms = None
    if not args.only_spmv:
        mh = ls.sys.mas_handle()
        z = ctx.empty(3 * n, np.float64)
        pkg._lib.check(L.sgipc_mas_time_apply(mh, x.vp, z.vp, 20, ctypes.byref(avg)), ctx.h)
        mas_ms = avg.value
        log(f"MAS apply ({args.precision} build): {mas_ms * 1e3:.1f} us")

    if args.only_spmv:
        print(json.dumps({"spmv_us": spmv_ms * 1e3, "spmv_GBs": gbs64, "spmv_warm_us": spmv_warm_ms * 1e3,
                          "spmv32_us": spmv32_ms * 1e3,
                          "spmv32_GBs": gbs32, "nnz": nnz, "stored_blocks": stored, "layout_GBs": moved64,
                          "halo_form": bool(halo_form), "largest_tile_halo": halo_max}),
              flush=True)
        ls.close()
        return

    # ---- plain FP64 additive MAS (B=32) on the same scene, same process
    plain = None
    plain_cfg = (args.precision, args.bank, args.coarse, args.learned) == ("fp64", 32, 0, 0)
    if args.compare_plain > 0 and not plain_cfg:
        pl = pkg.LinearSystem(scene, bank_size=32, coarse_mode=0)
        pl.solve(tol=1e-5, precision="fp64")
        pl.ctx.synchronize()
        t0 = time.perf_counter()
        its = []
        for _ in range(args.compare_plain):
            its.append(pl.solve(tol=1e-5, precision="fp64").iterations)
        pl.ctx.synchronize()
        pms = (time.perf_counter() - t0) * 1e3 / args.compare_plain
        pl.close()
        plain = {"ms": round(pms, 3), "iterations": float(np.mean(its)), "speedup": round(pms / out["ms"], 3)}
        log(f"plain FP64 additive MAS B=32: {pms:.2f} ms/solve, {np.mean(its):.0f} iterations -> "
            f"speedup {pms / out['ms']:.2f}x")

    # ---- CPU baseline: the oracle on the host cores, bounded sample
    cpu = None
    if not args.no_cpu_baseline:
        # the reference's method (additive MAS, B=32, FP64), timed in full
        cpu = cpu_baseline(ge, scene, 32, args.cpu_reps, args.cpu_threads)

    ls.close()
    mixed = args.precision == "fp32_ir"
    kname = "k_spmv_halo" if halo_form else "k_spmv_tsym"
    key = f"{kname}<float,double>" if mixed else f"{kname}<double,double>"
    t_ms, b_alg, gbs, moved = (spmv32_ms, b32, gbs32, moved32) if mixed else (spmv_ms, b64, gbs64, moved64)
    tr = measured_traffic(key)
    roof = {
        "bound": "hbm",
        "kernel": (f"{key} (the PCG loop's SpMV in this configuration: 512-row tiles, in-tile upper blocks once, "
                   + ("x of each tile and its halo staged in LDS, " if halo_form else "x gathered from memory, ")
                   + "emits p.Ap partials)"),
        "layout": {"halo_form": bool(halo_form), "largest_tile_halo": halo_max, "in_tile_slots_max": tl_slots},
        "achieved": round(gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 4),
        "traffic": tr["bytes_per_launch"] if tr else None,
        "traffic_source": tr["source"] if tr else None,
        "bytes_per_launch": b_alg,
        "bytes_basis": ("algorithmic: nnz_upper*(" + ("36" if mixed else "72") +
                        " B values + 8 B row/col) + 24 B/row x + 24 B/row y (FP64 vectors)"),
        "nnz_upper_blocks": nnz,
        "stored_blocks": stored,
        "layout_streamed_GBs": round(moved, 1),
        "avg_launch_us": round(t_ms * 1e3, 3),
        "timing": "HIP event pair recorded by each launch at the kernel's start and end (hipExtLaunchKernelGGL), "
                  "on the library stream; a 512 MiB buffer is streamed before every launch so the matrix comes "
                  "from HBM, as inside PCG",
        "fp64_values_avg_launch_us": round(spmv_ms * 1e3, 3),
        "fp64_values_GBs": round(gbs64, 1),
        "fp64_values_back_to_back_us": round(spmv_warm_ms * 1e3, 3),
    }
    dtype = {"fp64": "f64", "fp32_mas": "f64 (MAS apply f32)",
             "fp32_ir": "f32 SpMV values + MAS apply, f64 vectors/dots/refinement"}[args.precision]
    if args.coarse == 2 and args.level0_half and args.precision != "fp64":
        dtype += "; level-0 domain inverses stored f16 (unit-diagonal scaled), applied in f32"
    line = {
        "metric": "ms per Newton linear solve to rel. residual 1e-5; SpMV HBM GB/s vs peak",
        "value": round(out["ms"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(out["ms"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded 64^3 jittered Kuhn tet grid, per-tet E log-uniform, seeded PT pairs, N(0,1) rhs)",
        "config": {
            "workload": (f"stiff contact scene: {args.grid}^3-vertex grid ({scene.n} verts, {scene.n_tets} tets), "
                         f"stable Neo-Hookean E log-uniform [1e4,1e9] Pa nu=0.45, {scene.n_pairs} point-triangle "
                         f"barrier pairs (d^2 ~ U(0,dhat^2), dhat=1e-3, kappa=1e6), dt=0.01, MAS B={args.bank}, "
                         f"solve to ||b-Ax||/||b|| <= 1e-5"),
            "precision": args.precision,
            "bank_size": args.bank,
            "fused_iteration": args.fuse,
            "pcg_graph_replay": bool(args.graph),
            "mas_coarse_mode": {0: "additive",
                                1: "level-1 V(1,1) cycle, exact dense level-2 solve",
                                2: "two-grid cycle on level 0 (weighted level-0 domain solves before and after) "
                                   "around the level-1 V(1,1) cycle, exact dense level-2 solve"}[args.coarse],
            "level0_smoothing_weight": (args.omega if args.omega is not None else 0.4) if args.coarse == 2 else None,
            "level0_inverse_storage": ("fp16 scaled to a unit diagonal (D C D, D per slot in f32)"
                                       if args.coarse == 2 and args.level0_half and args.precision != "fp64"
                                       else "same as the apply"),
            "learned": learned_info,
            "learned_used_in_timed_steps": out.get("learned_used"),
            "plain_fp64_b32_same_run": plain,
            "n_verts": scene.n,
            "nnz_blocks": out["nnz"],
            "n_triplets": out["n_triplets"],
            "pcg_iterations": out["iters"],
            "true_rel_residual": out["res"],
            "rel_error_vs_fp64_solution": out.get("err_vs_fp64"),
            "phase_ms": {k: round(v, 3) for k, v in out["phases"].items()},
            "mas_apply_us": round(mas_ms * 1e3, 2) if mas_ms else None,
            "spmv_fp32_us": round(spmv32_ms * 1e3, 2),
            "spmv_fp32_GBs": round(gbs32, 1),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


SCENE_WORKLOADS = {
    "c2": "hybrid scene: 48^3-vertex tet grid + 200 affine bodies, 30000 barrier pairs (111392 block rows); "
          "body rows first, MAS over the FEM rows, body block-Jacobi; solve to ||b-Ax||/||b|| <= 1e-5",
    "c4": "cloth scene: 1024^2-vertex sheet (Baraff-Witkin + quadratic bending), 200000 seeded barrier pairs; "
          "MAS B=32; solve to ||b-Ax||/||b|| <= 1e-5",
}


def bench_scene(pkg, args):
    """Whole-solve times of another scene: `--warmup` untimed solves, then
    `--steps` timed ones of the same system; one JSON line.  The SpMV roofline
    and the CPU baseline are reported on the default scene only."""
    t0 = time.perf_counter()
    scene = pkg.scenes.scene_c2() if args.scene == "c2" else pkg.scenes.scene_c4()
    log(f"scene {args.scene}: n={scene.n} ({time.perf_counter() - t0:.1f}s to generate)")
    ls = pkg.LinearSystem(scene, bank_size=args.bank, coarse_mode=args.coarse)
    try:
        if args.omega is not None:
            ls.set_smoothing_weight(args.omega)
        ls.set_level0_half(args.level0_half)
        kw = dict(tol=1e-5, precision=args.precision, fuse=args.fuse, ir_replace_every=args.ir_replace,
                  graph=bool(args.graph))
        for w in range(args.warmup):
            r = ls.solve(**kw)
            log(f"warmup {w}: {r.timing['ms_total']:.2f} ms iters={r.iterations} res={r.true_rel_residual:.2e} "
                f"path={r.coarse_mode} fp64_fallback={r.fp64_fallback}")
        ls.ctx.synchronize()
        times, runs = [], []
        for k in range(args.steps):
            s0 = time.perf_counter()
            r = ls.solve(**kw)
            times.append((time.perf_counter() - s0) * 1e3)
            runs.append(r)
            log(f"step {k}: {times[-1]:.2f} ms iters={r.iterations} res={r.true_rel_residual:.2e} "
                f"path={r.coarse_mode} fp64_fallback={r.fp64_fallback}")
    finally:
        ls.close()
    ms = float(np.mean(times))
    last = runs[-1]
    out = {
        "metric": "ms per Newton linear solve to rel. residual 1e-5",
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (the mixed-precision solve fell back to FP64)" if last.fp64_fallback else
                 {"fp64": "f64", "fp32_mas": "f64 (MAS apply f32)",
                  "fp32_ir": "f32 SpMV values + MAS apply, f64 vectors/dots/refinement"}[args.precision],
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": SCENE_WORKLOADS[args.scene], "precision": args.precision, "bank_size": args.bank,
                   "mas_coarse_mode_requested": args.coarse, "mas_coarse_path": last.coarse_mode,
                   "fp64_fallback": last.fp64_fallback, "n_block_rows": int(scene.n_sys) if hasattr(scene, "n_sys")
                   else int(scene.n),
                   "pcg_iterations": [int(r.iterations) for r in runs],
                   "true_rel_residual": max(float(r.true_rel_residual) for r in runs),
                   "phase_ms": {k: round(float(np.median([r.timing[k] for r in runs])), 3)
                                for k in ("ms_assemble", "ms_convert", "ms_mas_build", "ms_pcg")}},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def train_learned(pkg, ls, args):
    """Train the learned coarse level on a separately seeded 64^3 scene
    (same generator, other seeds) from its plain solver's residuals, on the
    GPU, then attach it to the benchmark system."""
    t0 = time.perf_counter()
    tr_scene = pkg.scenes.tet_grid_scene("train_grid64", args.grid, E_loguniform=(1e4, 1e9), n_pairs=args.pairs,
                                         seed_jitter=10, seed_rhs=11, seed_E=12, seed_pairs=13)
    tr = pkg.LinearSystem(tr_scene, bank_size=args.bank, coarse_mode=args.coarse)
    per = max(args.train_samples // 3, 1)
    data = tr.collect(stride=1, max_samples=per, precision=args.precision)
    for k in range(2):
        b = np.random.default_rng(50 + k).standard_normal(3 * tr_scene.n)
        more = tr.collect(stride=1, max_samples=per, b=b, precision=args.precision)
        X, q = more.samples()
        data.add(X, q)
        more.close()
    t1 = time.perf_counter()
    net = pkg.learned.Mpnn(ls.ctx, layers=3, linear=False, seed=1)
    hist = net.train(data, epochs=args.train_epochs, batch=16, lr=1e-3, seed=0)
    ls.ctx.synchronize()
    t2 = time.perf_counter()
    n_samples = data.size
    data.close()
    tr.close()
    ls.set_learned(net, mode=args.learned)
    # the automatic mode measures both variants on its first two solves
    probe = [ls.solve(tol=1e-5, precision=args.precision, ir_replace_every=args.ir_replace,
                      graph=bool(args.graph)) for _ in range(2)]
    info = {
        "mode": {1: "always", 2: "automatic"}[args.learned],
        "layers": 3, "channels": 64, "samples": n_samples, "epochs": args.train_epochs,
        "collect_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
        "loss_first_epoch": float(hist[0]), "loss_last_epoch": float(hist[-1]),
        "probe_iterations": {("learned" if r.learned else "plain"): r.iterations for r in probe},
        "probe_learned_abandoned": any(r.learned_abandoned for r in probe),
    }
    log(f"learned level: {n_samples} samples ({t1 - t0:.1f}s), trained {args.train_epochs} epochs ({t2 - t1:.1f}s), "
        f"loss {hist[0]:.3e} -> {hist[-1]:.3e}; probe {info['probe_iterations']}")
    return info, net


def cpu_baseline(ge, scene, bank, reps=1, threads=0):
    """The oracle (C restatement of the reference algorithm, OpenMP) on the
    host: one whole solve of the reference's method -- assembly into one
    triplet arena, conversion (parallel radix sort), MAS build (additive, B
    = bank, FP64) and PCG to ||b - Ax|| <= 1e-5 ||b|| -- timed end to end,
    `reps` times (median reported)."""
    import statistics
    orc = ge.load_oracle()
    avail = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except Exception:
        pass
    threads = threads if threads > 0 else avail
    orc.set_threads(threads)
    runs = []
    for _ in range(max(1, reps)):
        t0 = time.perf_counter()
        vals, rows, cols = orc.assemble_arena(scene)
        t1 = time.perf_counter()
        mas = orc.Mas(scene, bank)
        mas.build(vals, rows, cols, 0, scene.pairs)
        t2 = time.perf_counter()
        x, it, res = orc.pcg(vals, rows, cols, scene.n, scene.rhs, mas, tol=1e-5, stop_mode=1)
        t3 = time.perf_counter()
        mas.close()
        del vals, rows, cols, x
        runs.append((t3 - t0, t1 - t0, t2 - t1, t3 - t2, it, res))
        log(f"cpu baseline ({threads} threads): assemble+convert {t1 - t0:.2f}s, MAS build {t2 - t1:.2f}s, "
            f"PCG {it} iterations {t3 - t2:.2f}s (rel. residual {res:.2e}) -> {t3 - t0:.2f}s per solve")
    med = statistics.median(r[0] for r in runs)
    r = min(runs, key=lambda q: abs(q[0] - med))
    return {
        "value": round(med * 1e3, 1),
        "unit": "ms",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle/sgipc_oracle.c (OpenMP, {threads} threads): {len(runs)} whole solve(s) of the "
                   f"reference's method on this scene (additive MAS B={bank}, FP64, to rel. residual 1e-5), "
                   f"median; the median run: assembly+conversion {r[1]:.2f}s (element Hessians with a numerical "
                   f"9x9 eigen projection), MAS build {r[2]:.2f}s, PCG {r[4]} iterations {r[3]:.2f}s"),
        "runs_s": [round(q[0], 3) for q in runs],
    }


if __name__ == "__main__":
    main()
