#!/usr/bin/env python3
"""Benchmark: plain fp32 MLEM iterations on one MI355X.

A step is one MLEM iteration (forward projection + ratio, backprojection +
update) over the whole sinogram. The workload is the 128x128x48-grid,
1,327,104-line system matrix (Siddon lengths, seeds 1001/2001/3001, 1e8
counts); matrix, sinogram and image are resident in HBM before the timed
region starts.

Prints ONE JSON line with the metric, the per-kernel roofline of the dominant
kernel (HIP events around every launch inside the timed region) and the CPU
baseline (the repository's fp32 sequential restatement of the reference
driver, timed on this host on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
METRIC = "MLEM iter/s + HBM GB/s (~200M nnz fp32); batched 64-frame vs plain speedup"


def kernel_bytes(rows: int, cols: int, nnz: int, precision: str = "fp32") -> dict:
    """Algorithmic HBM bytes per launch: every matrix record, index and vector
    element counted once (u32 row/column pointers). Records are 8 bytes
    (u32 index + fp32 value) or 6 bytes with 16-bit values, plus the fp32 row
    scale in the forward pass."""
    rec = 8 if precision == "fp32" else 6
    scale = 0 if precision == "fp32" else 4 * rows
    fwd = rec * nnz + 4 * (rows + 1) + 4 * cols + 4 * rows + 4 * rows + scale
    bwd = rec * nnz + 4 * (cols + 1) + 4 * rows + 4 * cols + 8 * cols
    return {"fwd": fwd, "bwd": bwd, "iteration": fwd + bwd}


def quarter_means(t) -> list:
    """Mean kernel time of the first and of the last quarter of the timed
    launches (clock ramp versus steady state)."""
    t = np.asarray(t, np.float64)
    q = max(1, t.size // 4)
    return [float(np.mean(t[:q])), float(np.mean(t[-q:]))]


def large_run(steps: int, fp64_iters: int) -> dict:
    """The 192x192x64 / 4,194,304-line matrix (about 0.79 G nonzeros) with fp32,
    fp16 and bf16 values (16-bit: per-row fp32 scale): iterations/s, per-kernel
    times and roofline fractions after `steps` timed FAST iterations, and the
    image after `fp64_iters` iterations against FP64 MLEM on the original
    values (relative L2 and L-infinity)."""
    import mlem_amd as M
    from mlem_amd import synth
    from oracle import oracle as orc  # accuracy reference only
    t0 = time.time()
    prob = synth.make_problem("large")
    m = prob.matrix
    out = {"workload": f"large: {'x'.join(map(str, m.grid))} grid, {m.rows} lines, "
                       f"{m.nnz} nonzeros", "rows": m.rows, "cols": m.cols, "nnz": m.nnz,
           "generation_s": time.time() - t0, "steps": steps}
    t0 = time.time()
    x64 = orc.f64_mlem_par(m.rowidx, m.elements, m.cols, prob.y.astype(np.float64),
                           fp64_iters)[0] if fp64_iters > 0 else None
    out["fp64_iterations"] = fp64_iters
    out["fp64_oracle_s"] = time.time() - t0
    for prec in ("fp32", "fp16", "bf16"):
        t0 = time.time()
        A = M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid,
                                     precision=prec)
        open_s = time.time() - t0
        try:
            t0 = time.time()
            A.autotune()
            tune_s = time.time() - t0
            R = M.Recon(A, prob.y, mode="fast", use_graph=False)
            R.run(3)
            M.device_sync()
            t0 = time.perf_counter()
            kf, kb, _ = R.run_timed(steps)
            M.device_sync()
            wall = time.perf_counter() - t0
            R.close()
            B = kernel_bytes(m.rows, m.cols, m.nnz, prec)
            f_ms, b_ms = float(np.mean(kf)), float(np.mean(kb))
            res = {"iterations_per_s": steps / wall, "ms_per_iteration": 1e3 * wall / steps,
                   "launch": A.launch(), "open_s": open_s, "autotune_s": tune_s,
                   "kernel_ms": {"fwd": f_ms, "bwd": b_ms},
                   "roofline_frac": {"fwd": B["fwd"] / (f_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                     "bwd": B["bwd"] / (b_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                     "iteration": B["iteration"] / (wall / steps) / 1e9
                                                  / PEAK_HBM_GBS},
                   "algorithmic_bytes": B}
            if x64 is not None:
                x = M.mlem(A, prob.y, None, fp64_iters, mode="fast").astype(np.float64)
                floor = 1e-6 * float(np.max(np.abs(x64)))
                res["vs_fp64_original"] = {
                    "rel_l2": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
                    "rel_linf": float(np.max(np.abs(x - x64) / np.maximum(np.abs(x64), floor)))}
        finally:
            A.close()
        out[prec] = res
        print(f"[bench] large {prec}: {res}", file=sys.stderr, flush=True)
    f32 = out["fp32"]["iterations_per_s"]
    out["speedup_vs_fp32"] = {p: out[p]["iterations_per_s"] / f32 for p in ("fp16", "bf16")}
    return out


def unrolled_run(A, prob, frames: int, K: int, checkpoint: int, mode: str = "fast") -> dict:
    """Differentiable unrolled MLEM on the same matrix: `frames` sinograms
    with relaxed (real) counts, K iterations from x0 = sum(y)/sum(n), loss
    0.5 ||x_K - x_true||^2; times the forward pass and the reverse pass
    (dL/dy, dL/dx0 with recomputation from checkpoints). Algorithmic bytes of
    one iteration over a block of W frames: the matrix read for both products
    at 8 bytes per element each, and every vector run once (counts, f and r
    per row; x read and written and the column sums per column, W*4 bytes
    each); the reverse pass recomputes K iterations and runs K adjoint
    iterations of the same two products."""
    import mlem_amd as M
    rng = np.random.default_rng(4000)
    Y = np.stack([rng.poisson(prob.ybar * s) + rng.uniform(0, 0.5, prob.ybar.size)
                  for s in rng.uniform(0.5, 2.0, frames)]).astype(np.float32)
    U = M.Unrolled(A, Y, None, K, checkpoint_interval=checkpoint, mode=mode)  # warm-up
    U.backward(U.xK - prob.x_true)
    U.close()
    M.device_sync()
    t0 = time.perf_counter()
    U = M.Unrolled(A, Y, None, K, checkpoint_interval=checkpoint, mode=mode)
    M.device_sync()
    t1 = time.perf_counter()
    gy, gx0 = U.backward(U.xK - prob.x_true)
    M.device_sync()
    t2 = time.perf_counter()
    info = U.info()
    U.close()
    m = prob.matrix
    W = info["block_width"]
    blocks = -(-frames // W)
    per_iter = blocks * (16 * m.nnz + 4 * W * (3 * m.rows + 3 * m.cols))
    fwd_s, bwd_s = t1 - t0, t2 - t1
    fwd_gbs = K * per_iter / fwd_s / 1e9
    bwd_gbs = 2 * K * per_iter / bwd_s / 1e9
    return {"mode": mode, "frames": frames, "K": K, **info,
            "forward_ms": 1e3 * fwd_s, "backward_ms": 1e3 * bwd_s,
            "frame_iterations_per_s": {"forward": frames * K / fwd_s,
                                       "backward": frames * K / bwd_s},
            "roofline": {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                         "algorithmic_bytes_per_iteration": per_iter,
                         "forward": {"achieved": fwd_gbs, "frac": fwd_gbs / PEAK_HBM_GBS},
                         "backward": {"achieved": bwd_gbs, "frac": bwd_gbs / PEAK_HBM_GBS}},
            "note": "host-to-device of y and device-to-host of x_K, dL/dy, dL/dx0 "
                    "included in the two times",
            "grad_finite": bool(np.all(np.isfinite(gy)) and np.all(np.isfinite(gx0)))}


def fp64_check(A, prob, precision: str, iters: int) -> dict:
    """`iters` GPU iterations against FP64 MLEM on the original values (the
    accuracy of the stored precision) and, for 16-bit storage, on the
    dequantized values s_i q_ij (parity of the kernels)."""
    import mlem_amd as M
    from oracle import oracle as orc  # accuracy check only
    m = prob.matrix
    R = M.Recon(A, prob.y, mode="fast")
    R.run(iters)
    x = R.image().astype(np.float64)
    R.close()
    y = prob.y.astype(np.float64)
    t0 = time.time()
    x64 = orc.f64_mlem_par(m.rowidx, m.elements, m.cols, y, iters)[0]
    out = {"iterations": iters, "threads": os.cpu_count()}

    def err(a, b):
        floor = 1e-6 * float(np.max(np.abs(b)))
        return {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                "rel_linf": float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))}

    out["vs_fp64_original"] = err(x, x64)
    print(f"[bench] fp64 oracle on the original values: {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    if precision != "fp32":
        q, s = orc.quantize(m.rowidx, m.elements, precision)
        xq = orc.f64_mlem_par(m.rowidx, q, m.cols, y, iters, rscale=s)[0]
        out["vs_fp64_dequantized"] = err(x, xq)
        out["fp64_dequantized_vs_original"] = err(xq, x64)
    out["oracle_s"] = time.time() - t0
    return out


def parity_check(A, prob, iters: int) -> dict:
    """The image after `iters` iterations in FAST and in EXACT mode against the
    same iterations of the fp32 sequential restatement of seqcsr4mlem."""
    import mlem_amd as M
    from oracle import oracle as orc  # parity check only
    m = prob.matrix
    xo = orc.mlem(m.rowidx, m.elements, m.cols, prob.y, iters)[0]
    floor = 1e-6 * float(np.max(np.abs(xo)))
    out = {"iterations": iters}
    for mode in ("fast", "exact"):
        x = M.mlem(A, prob.y, None, iters, mode=mode)
        d = np.abs(x.astype(np.float64) - xo.astype(np.float64))
        out[mode] = {"rel_linf_vs_seq_fp32": float(np.max(d / np.maximum(np.abs(xo), floor))),
                     "rel_l2_vs_seq_fp32": float(np.linalg.norm(d) / np.linalg.norm(xo)),
                     "bit_identical": bool(np.array_equal(x, xo))}
    return out


def cpu_baseline(problem, iters: int):
    from oracle import oracle as orc  # CPU baseline leg only
    m = problem.matrix
    x, _, _, _, secs = orc.mlem(m.rowidx, m.elements, m.cols, problem.y, iters)
    med = float(np.median(secs))
    out = {"value": 1.0 / med, "unit": "iterations/s", "cores": 1, "kind": "port",
           "sample": f"{iters} iterations of the fp32 sequential restatement of seqcsr4mlem "
                     f"(oracle/mlem_oracle.c) on the full benchmark matrix; median per-iteration "
                     f"time {med:.3f} s"}
    # the OpenMP driver's restatement (openmpcsr4mlem.cpp:69-128) on the host
    # cores this process may use (at most 16)
    threads = min(16, len(os.sched_getaffinity(0)))
    _, osecs = orc.omp_mlem(m.rowidx, m.elements, m.cols, problem.y, iters, threads=threads)
    omed = float(np.median(osecs))
    out["openmp"] = {"value": 1.0 / omed, "unit": "iterations/s", "cores": threads,
                     "sample": f"{iters} iterations of the OpenMP restatement, median "
                               f"{omed:.3f} s per iteration"}
    # thread scaling of the same restatement (2 iterations per thread count)
    scaling = {}
    t = 1
    while t <= threads:
        _, ts = orc.omp_mlem(m.rowidx, m.elements, m.cols, problem.y, 2, threads=t)
        scaling[str(t)] = 1.0 / float(np.median(ts))
        t *= 2
    out["openmp"]["iterations_per_s_by_threads"] = scaling
    return out


def streamed_run(A, prob, block_mib: int, iters: int, plain_value: float,
                 resident_frac: float = 0.0, mode: str = "fast") -> dict:
    """The same reconstruction with the matrix kept in pinned host memory and
    streamed to the device in row blocks of at most `block_mib` MiB per part
    (out-of-core path): iterations/s, host-to-device bytes per iteration
    (every streamed block's CSR part, once: its CSC is built on the device, or
    in ATOMIC mode its rows are scattered while it is there; the CSC parts too
    when the device CSC did not fit) and the PCIe rate they imply, plus the
    image's distance to the in-core FAST run."""
    import mlem_amd as M
    m = prob.matrix
    t0 = time.perf_counter()
    resident = int(resident_frac * 16 * m.nnz)
    S = M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements,
                                 stream_block_bytes=block_mib << 20,
                                 stream_resident_bytes=resident)
    open_s = time.perf_counter() - t0
    nb, nr, passes = S.stream_blocks(), S.resident_blocks(), S.stream_passes()
    R = M.Recon(S, prob.y, mode=mode)
    R.run(1)
    M.device_sync()
    t0 = time.perf_counter()
    R.run(iters)
    M.device_sync()
    wall = time.perf_counter() - t0
    xs = R.image()
    R.close()
    S.close()
    xh = M.mlem(A, prob.y, None, iters + 1, mode="fast")
    d = np.abs(xs.astype(np.float64) - xh.astype(np.float64))
    floor = 1e-6 * float(np.max(np.abs(xh)))
    if mode == "atomic" or passes == 1:
        moved = (8 * m.nnz + 4 * (m.rows + nb)) * (nb - nr) / nb
    else:
        moved = (16 * m.nnz + 4 * (m.rows + nb * (m.cols + 1))) * (nb - nr) / nb
    return {"mode": mode, "block_mib": block_mib, "blocks": nb, "resident_blocks": nr,
            "passes_per_iteration": 1 if mode == "atomic" else passes, "iterations": iters,
            "ms_per_iteration": 1e3 * wall / iters, "iterations_per_s": iters / wall,
            "h2d_bytes_per_iteration": moved, "h2d_gbs": moved * iters / wall / 1e9,
            "slowdown_vs_hbm": plain_value / (iters / wall), "open_s": open_s,
            "rel_linf_vs_in_core": float(np.max(d / np.maximum(np.abs(xh), floor)))}


def batched_run(A, prob, frames: int, width: int, plain_value: float, mode: str = "fast",
                parity: bool = True, Y=None) -> dict:
    """Dynamic-frame workload on the same matrix: per-frame phantom scaling
    U(0.2, 2) per ellipsoid, total counts log-uniform in [1e5, 1e8], iteration
    budgets uniform in 10..100, early stop at a relative log-likelihood change
    below 1e-6. Returns frame-iterations/s and, with `parity`, the frames with
    the fewest and the most iterations against FP64 MLEM (and, in EXACT mode,
    their bit-identity with the fp32 sequential restatement)."""
    import mlem_amd as M
    from mlem_amd import synth
    t0 = time.time()
    budgets = synth.frame_budgets(frames)
    if Y is None:
        Y = synth.dynamic_frames(prob, frames)
    gen = time.time() - t0
    t0 = time.time()
    B = M.Batch(A, Y, budgets, rel_tol=1e-6, block_width=width, mode=mode)
    create_s = time.time() - t0
    B.run(1)  # warm-up: graph capture, first launches
    done0 = B.progress()["frame_iterations"]
    M.device_sync()
    t0 = time.perf_counter()
    n = B.run(1000)
    M.device_sync()
    wall = time.perf_counter() - t0
    fi = B.progress()["frame_iterations"] - done0
    its = [B.frame_info(f)[0] for f in range(frames)]
    out = {"frames": frames, "mode": mode, "block_width": width or "autotuned",
           "block_ms_by_width": B.tuning() if width == 0 else None,
           "frame_iterations": int(fi),
           "batched_iterations": n + 1, "wall_s": wall,
           "frame_iterations_per_s": fi / wall,
           "speedup_vs_plain": (fi / wall) / plain_value,
           "iterations_per_frame": {"min": int(min(its)), "max": int(max(its)),
                                    "mean": float(np.mean(its))},
           "early_stopped_frames": int(sum(1 for f in range(frames) if its[f] < budgets[f])),
           "generation_s": gen, "create_s": create_s,
           "note": "the timed span starts after the first batched iteration (graph "
                   "capture); create_s includes building the matrix's tiling and timing "
                   "the block widths"}
    if parity:
        from oracle import oracle as orc  # parity check only
        m = prob.matrix
        out["parity"] = []
        # the frame that stopped first and the one that ran longest (last)
        for f in (int(np.argmin(its)), int(np.argmax(its))):
            x = B.image(f)
            x64 = orc.f64_mlem_par(m.rowidx, m.elements, m.cols, Y[f].astype(np.float64),
                                   its[f])[0]
            floor = 1e-6 * float(np.max(np.abs(x64)))
            d = np.abs(x.astype(np.float64) - x64)
            rec = {"frame": f, "iterations": its[f],
                   "rel_linf_vs_fp64": float(np.max(d / np.maximum(np.abs(x64), floor))),
                   "rel_l2_vs_fp64": float(np.linalg.norm(d) / np.linalg.norm(x64))}
            if mode == "exact":
                xo = orc.mlem(m.rowidx, m.elements, m.cols, Y[f], its[f])[0]
                rec["bit_identical_to_seq_fp32"] = bool(np.array_equal(x, xo))
            out["parity"].append(rec)
    B.close()
    out["_Y"] = Y
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="pet")
    ap.add_argument("--mode", default="fast", choices=["fast", "atomic", "exact"])
    ap.add_argument("--cpu-iters", type=int, default=4,
                    help="iterations of the CPU baseline sample (0 = skip)")
    ap.add_argument("--parity-iters", type=int, default=60,
                    help="iterations of the FAST / EXACT image checked against the fp32 "
                         "sequential restatement (0 = skip)")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="pick each projection's kernel family and tile size by timing them "
                         "on the matrix (mlem_autotune) before the warm-up")
    ap.add_argument("--keep-order", action="store_true",
                    help="keep file row/voxel order on the device (no locality layout)")
    ap.add_argument("--launch", default="", help="family,param,family,param e.g. tiles,8,groups,32")
    ap.add_argument("--batched-frames", type=int, default=64,
                    help="frames of the batched (dynamic-frame) measurement; 0 = skip")
    ap.add_argument("--batched-width", type=int, default=0,
                    help="widest RHS block (4..32); 0 = autotuned widths")
    ap.add_argument("--batched-mode", default="fast", choices=["fast", "exact"],
                    help="fast: the matrix's two-dimensional tiling (within 1e-4 of FP64); "
                         "exact: every frame bit-identical to the sequential driver")
    ap.add_argument("--batched-exact-too", action=argparse.BooleanOptionalAction, default=True,
                    help="also time the bit-identical (EXACT) batched mode")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="first run the plain copy kernel over this many MiB (calibrates "
                         "hardware byte counters under rocprofv3 --pmc)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp16", "bf16"],
                    help="matrix value storage (16-bit: per-row fp32 scale)")
    ap.add_argument("--unrolled-frames", type=int, default=8,
                    help="frames of the unrolled forward/backward timing (K = 20, "
                         "checkpoints every 5 iterations); 0 = skip")
    ap.add_argument("--streamed-iters", type=int, default=5,
                    help="iterations of the out-of-core (host-streamed) measurement; 0 = skip")
    ap.add_argument("--streamed-block-mib", type=int, default=256,
                    help="largest part of a streamed row block (MiB)")
    ap.add_argument("--large-steps", type=int, default=20,
                    help="timed iterations per value precision on the 192x192x64 matrix "
                         "(fp32, fp16, bf16); 0 = skip")
    ap.add_argument("--large-fp64-iters", type=int, default=5,
                    help="iterations of the 192x192x64 images checked against FP64 MLEM")
    ap.add_argument("--fp64-iters", type=int, default=0,
                    help="also compare this many iterations with FP64 MLEM (oracle, all "
                         "host cores); 0 = skip")
    args = ap.parse_args()
    if args.precision != "fp32" and (args.mode == "atomic" or args.batched_frames):
        args.batched_frames = 0  # the batched path runs on fp32 values only
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU row sharding is a design note, not built)", file=sys.stderr)
        return 2

    import mlem_amd as M
    from mlem_amd import synth

    if args.calibration_copy_mib > 0:
        nbytes = args.calibration_copy_mib << 20
        ms = M.calibration_copy(nbytes, 3)
        print(f"[bench] copy kernel: {nbytes} B read + {nbytes} B written in {ms:.3f} ms "
              f"({2 * nbytes / ms / 1e6:.0f} GB/s)", file=sys.stderr, flush=True)

    t0 = time.time()
    prob = synth.make_problem(args.workload)
    m = prob.matrix
    print(f"[bench] workload {args.workload}: {m.rows} rows, {m.cols} cols, {m.nnz} nnz "
          f"(generated in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    t0 = time.time()
    A = M.Csr4Matrix.from_arrays(m.rows, m.cols, m.rowidx, m.elements, grid=m.grid,
                                 no_csc=(args.mode == "atomic" and args.precision == "fp32"),
                                 keep_order=args.keep_order, precision=args.precision)
    print(f"[bench] matrix in HBM + CSC in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    if args.launch:
        f, fp, b, bp = args.launch.split(",")
        A.set_launch((f, int(fp)), (b, int(bp)))
    tune_s = None
    if args.autotune and not args.launch:
        t0 = time.time()
        A.autotune()
        tune_s = time.time() - t0
        print(f"[bench] autotuned launch {A.launch()} in {tune_s:.1f}s", file=sys.stderr,
              flush=True)
    launch = A.launch()

    R = M.Recon(A, prob.y, mode=args.mode, use_graph=False)
    if args.warmup:
        R.run(args.warmup)
    M.device_sync()
    t_start = time.perf_counter()
    kf, kb, span = R.run_timed(args.steps)
    M.device_sync()
    wall = time.perf_counter() - t_start
    done = R.iterations
    ms_per_step = 1000.0 * wall / args.steps
    value = args.steps / wall

    B = kernel_bytes(m.rows, m.cols, m.nnz, args.precision)
    fwd_ms, bwd_ms = float(np.mean(kf)), float(np.mean(kb))
    dom = "fwd" if fwd_ms >= bwd_ms else "bwd"
    dom_ms = max(fwd_ms, bwd_ms)
    achieved = B[dom] / (dom_ms * 1e-3) / 1e9
    suffix = {"dict512": "_dict", "dict": "_dict", "tiles": "_tiles", "groups": ""}[launch[dom][0]]
    if args.mode == "exact":
        suffix = ""
    name = {"fwd": f"k_fwd{suffix} (CSR forward projection + fused ratio)",
            "bwd": f"k_bwd{suffix} (CSC gather backprojection + fused update)"}[dom]
    if args.mode == "atomic" and dom == "bwd":
        name = "k_scatter + k_update (atomic backprojection)"

    # HBM bytes per launch of the dominant kernel: the L2-to-memory read and
    # write requests of the committed rocprofv3 counter passes over this
    # bench's workload, times their request sizes (mlem_amd/counters.py).
    traffic = None
    ptag = "" if args.precision == "fp32" else f"_{args.precision}"
    tfile = ROOT / "profiles" / f"traffic_{args.workload}{ptag}.json"
    fam, ept = launch[dom]
    if tfile.exists() and fam in ("dict", "dict512") and args.mode == "fast":
        from mlem_amd.counters import kernel_traffic
        vk = {"fp32": 0, "fp16": 1, "bf16": 2}[args.precision]
        nb = 512 if fam == "dict512" else 256
        # the timed launches: FAST, fused ratio (forward, no statistics) or
        # fused update (backward); else the autotuner's launches of the same
        # tile configuration (the same elements streamed, no update)
        summary = json.loads(tfile.read_text())
        tmpls = [f"<{nb}, {ept}, {vk}, false, true, false>("] if dom == "fwd" else \
            [f"<{nb}, {ept}, {vk}, false, true>(", f"<{nb}, {ept}, {vk}, false, false>("]
        for tmpl in tmpls:
            k = kernel_traffic(summary, f"k_{dom}_dict{tmpl}")
            if k and "bytes" in k:
                traffic = k["bytes"]
                break

    cpu = None
    if args.cpu_iters > 0:
        cpu = cpu_baseline(prob, args.cpu_iters)
        cpu["threads_available"] = os.cpu_count()
    parity = None
    if args.parity_iters > 0 and args.precision == "fp32":
        parity = parity_check(A, prob, args.parity_iters)
        print(f"[bench] parity: {parity}", file=sys.stderr, flush=True)

    batched = None
    if args.batched_frames > 0 and args.mode != "atomic":
        try:
            batched = batched_run(A, prob, args.batched_frames, args.batched_width, value,
                                  args.batched_mode)
            Y = batched.pop("_Y")
            if args.batched_exact_too and args.batched_mode != "exact":
                ex = batched_run(A, prob, args.batched_frames, args.batched_width, value,
                                 "exact", parity=False, Y=Y)
                batched["exact_mode"] = {k: ex[k] for k in (
                    "frame_iterations_per_s", "speedup_vs_plain", "frame_iterations", "wall_s",
                    "block_ms_by_width")}
        except Exception as e:  # reported in the line, never hides the plain result
            batched = {"error": f"{type(e).__name__}: {e}"}
        print(f"[bench] batched: {batched}", file=sys.stderr, flush=True)

    unrolled = None
    if args.unrolled_frames > 0 and args.precision == "fp32" and args.mode != "atomic":
        try:
            unrolled = unrolled_run(A, prob, args.unrolled_frames, 20, 5, "fast")
            ex = unrolled_run(A, prob, args.unrolled_frames, 20, 5, "exact")
            unrolled["exact_mode"] = {k: ex[k] for k in ("forward_ms", "backward_ms",
                                                         "frame_iterations_per_s")}
        except Exception as e:  # reported in the line, never hides the plain result
            unrolled = {"error": f"{type(e).__name__}: {e}"}
        print(f"[bench] unrolled: {unrolled}", file=sys.stderr, flush=True)

    streamed = None
    if args.streamed_iters > 0 and args.precision == "fp32" and args.mode == "fast":
        try:
            streamed = streamed_run(A, prob, args.streamed_block_mib, args.streamed_iters, value)
            streamed["half_resident"] = streamed_run(A, prob, args.streamed_block_mib,
                                                     args.streamed_iters, value, 0.5)
            streamed["atomic"] = streamed_run(A, prob, args.streamed_block_mib,
                                              args.streamed_iters, value, 0.0, "atomic")
        except Exception as e:  # reported in the line, never hides the plain result
            streamed = {"error": f"{type(e).__name__}: {e}"}
        print(f"[bench] streamed: {streamed}", file=sys.stderr, flush=True)

    large = None
    if args.large_steps > 0 and args.workload == "pet" and args.mode == "fast" \
            and args.precision == "fp32":
        try:
            large = large_run(args.large_steps, args.large_fp64_iters)
        except Exception as e:  # reported in the line, never hides the plain result
            large = {"error": f"{type(e).__name__}: {e}"}

    accuracy = None
    if args.fp64_iters > 0:
        accuracy = fp64_check(A, prob, args.precision, args.fp64_iters)
        print(f"[bench] fp64 check: {accuracy}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else f"{args.precision} values, f32 accumulate",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: plain MLEM, {args.precision} matrix values, "
                        f"{'x'.join(map(str, m.grid))} grid, {m.rows} lines of response, "
                        f"{m.nnz} nonzeros, {m.cols} voxels",
            "rows": m.rows, "cols": m.cols, "nnz": m.nnz, "mode": args.mode,
            "precision": args.precision, "launch": launch, "iterations_run": done,
        },
        "hbm_gbs_iteration": B["iteration"] / (ms_per_step * 1e-3) / 1e9,
        "kernel_ms": {"fwd": fwd_ms, "bwd": bwd_ms, "span_per_iter": span / args.steps,
                      "fwd_min_median_max": [float(np.min(kf)), float(np.median(kf)),
                                             float(np.max(kf))],
                      "bwd_min_median_max": [float(np.min(kb)), float(np.median(kb)),
                                             float(np.max(kb))],
                      "fwd_first_last_quarter": quarter_means(kf),
                      "bwd_first_last_quarter": quarter_means(kb)},
        "autotune_s": tune_s,
        "roofline": {
            "kernel": name,
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "algorithmic_bytes": B[dom],
            "other_kernel_frac": (B["bwd" if dom == "fwd" else "fwd"] /
                                  (min(fwd_ms, bwd_ms) * 1e-3) / 1e9) / PEAK_HBM_GBS,
        },
        "cpu_baseline": cpu,
        "parity": parity,
        "batched": batched,
        "unrolled": unrolled,
        "streamed": streamed,
        "large": large,
        "fp64_check": accuracy,
    }
    print(json.dumps(line), flush=True)
    R.close()
    A.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
