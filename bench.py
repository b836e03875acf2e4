#!/usr/bin/env python3
"""Benchmark of the libspmv hot path on one MI355X.

Step (the JSON `value`): one pass of the reference's matvec loop
(demos/matvecmain.cpp:70-86) on one rank's share of a 27-point stencil:
160^3 owned rows plus one 160x160 ghost plane fed through L2GMap from a
seeded host halo buffer (FP64 CSR, 109,443,836 stored entries):
    L2GMap::update(x)   device pack + halo unpack
    y = A * x           ordered FP64 CSR SpMV (bit-identical to the reference)
value = algorithmic SpMV bytes per step / measured step time (GB/s), with
    B = nnz*(8+4) + 4*(M+1) + 8*ncols + 8*M   (x read once, y written once).

Also reported: the SpMV kernel's roofline (HIP events on its stream), the
update and product phases timed apart (as demos/matvecmain.cpp:72-80 does),
the streaming-read and copy ceilings measured on the same GPU in the same
run, the CPU oracle timed on the host (1 thread, bounded sample, and all
cores), CG to 1e-10 on the share (GPU, and the oracle on all cores), the
reference's own CPU configuration (5-point 512^2: 100 applications of the
matvec loop and CG to 1e-8, GPU and oracle on 1 thread and all cores), the
CG time-to-1e-10 of FP64 CG versus mixed-precision refinement (FP32 and BF16
values) on the 256^3 7-point Poisson system (with the oracle's CG on all
cores), the batched solver and the adjoint.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "SpMV GB/s (% HBM peak); CG time-to-1e-10 rel. residual, mixed vs FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(nnz, nrows, ncols, val_bytes=8, vec_bytes=8):
    return nnz * (val_bytes + 4) + 4 * (nrows + 1) + vec_bytes * ncols + vec_bytes * nrows


def timed_loop(torch, fn, iters):
    """Average ms of fn() over iters, HIP events on the current stream."""
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(iters):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / iters


def cpu_threads():
    """Host threads for the all-cores CPU figures (the process's CPU share)."""
    n = os.cpu_count() or 1
    try:
        n = min(n, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def run_c1_section(torch, S, P, O, do_cpu, nthreads, n=512, napply=100, rtol=1e-8):
    """The reference's own CPU configuration: 5-point Poisson on n x n (one
    rank, no ghosts), x uniform[-1,1] (seed 11).  The matvec loop of
    demos/matvecmain.cpp:70-86 (update, y = A x, x[:M] = y) napply times,
    then CG (cg.cpp:11-83) with b = A x_true (x_true uniform[-1,1], seed 12),
    x0 = 0, to rtol; on the GPU and on the host (the oracle: 1 thread, and
    all cores), on the same inputs."""
    import numpy as np
    dev = "cuda:0"
    A = P.stencil(5, n, n)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    l2g = S.L2GMap(A.nrows)
    x0 = P.uniform(11, A.ncols, -1, 1)
    Bs = spmv_bytes(A.nnz, A.nrows, A.ncols)
    out = {"workload": f"5-point Poisson {n}^2 ({A.nrows} rows, {A.nnz} nnz), FP64, one rank",
           "matvec_loop": f"{napply} x (L2GMap::update, y = A x, x[:M] = y)", "cg_rtol": rtol}
    x = torch.from_numpy(x0).to(dev)
    y = torch.empty(A.nrows, dtype=torch.float64, device=dev)

    def step():
        l2g.update(x)
        M.mult(x, out=y)
        x[: A.nrows].copy_(y)

    for _ in range(3):
        step()
    x.copy_(torch.from_numpy(x0))
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(napply):
        step()
    torch.cuda.synchronize()
    tg = time.perf_counter() - t
    xg = x.cpu().numpy()
    out["gpu_matvec_loop_seconds"] = tg
    out["gpu_ms_per_apply"] = 1e3 * tg / napply
    out["spmv_bytes"] = Bs
    xt = P.uniform(12, A.ncols, -1, 1)
    oc = O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val)  # the checker and the CPU legs
    bd = M.mult(torch.from_numpy(xt).to(dev))                 # b = A x_true on the GPU
    b = bd.cpu().numpy()
    S.cg(M, bd, 2, rtol)
    torch.cuda.synchronize()
    t = time.perf_counter()
    xs, k = S.cg(M, bd, 100000, rtol)
    torch.cuda.synchronize()
    tcg = time.perf_counter() - t
    xsh = xs.cpu().numpy()
    out["gpu_cg"] = {"seconds": tcg, "iterations": k,
                     "true_rel_residual": float(np.linalg.norm(O.spmv(oc, xsh) - b) / np.linalg.norm(b))}
    log(f"[c1] GPU: {napply} applications {tg * 1e3:.2f} ms, CG {tcg:.4f}s {k} its")
    if do_cpu:
        for th in (1, nthreads):
            xc = x0.copy()
            t = time.perf_counter()
            for _ in range(napply):
                xc[: A.nrows] = O.spmv(oc, xc, threads=th)
            tc = time.perf_counter() - t
            t = time.perf_counter()
            xo, ko, _ = O.cg(oc, b, 100000, rtol, threads=None if th == 1 else th)
            tco = time.perf_counter() - t
            key = "cpu_1_thread" if th == 1 else "cpu_all_cores"
            out[key] = {"threads": th, "matvec_loop_seconds": tc, "ms_per_apply": 1e3 * tc / napply,
                        "matvec_result_bit_identical_to_gpu": bool(np.array_equal(xc, xg)),
                        "cg_seconds": tco, "cg_iterations": ko,
                        "cg_oracle": "sequential dots" if th == 1 else "threaded, 64-chunk dots",
                        "gpu_speedup_matvec": tc / tg, "gpu_speedup_cg": tco / tcg}
            log(f"[c1] CPU {th} thread(s): {napply} applications {tc:.3f}s, CG {tco:.3f}s {ko} its")
    return out


def run_cg_section(torch, S, P, O, n, seed, cpu_iters, nthreads=1):
    """FP64 CG vs mixed-precision refinement to 1e-10 on the n^3 7-point
    Poisson system (b uniform[-1,1])."""
    dev = "cuda:0"
    t0 = time.time()
    A = P.stencil(7, n, n, n)
    log(f"[cg] generated 7-point {n}^3: {A.nnz} nnz in {time.time() - t0:.1f}s")
    M64 = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    M32 = M64.like("f32")
    M16 = M64.like("bf16")
    import numpy as np
    b = torch.from_numpy(P.uniform(seed, A.nrows, -1, 1)).to(dev)
    bnorm = torch.linalg.norm(b).item()

    def true_rel(x):
        r = M64.mult(x)
        return torch.linalg.norm(b - r).item() / bnorm

    out = {"workload": f"7-point Poisson {n}^3, FP64 CSR ({A.nnz} nnz), b uniform[-1,1] seed {seed}",
           "rtol": 1e-10,
           "baseline": "fp64: plain FP64 CG on the CSR arrays (the reference's algorithm)"}
    try:
        info = M64.sliced_info()
        out["sliced_layout"] = {"patterns": info["npatterns"], "longest_row": info["wmax"],
                                "fp64_value_slots": info["slots"],
                                "bf16_value_slots": M16.sliced_info()["slots"]}
    except S.SpmvError as e:
        out["sliced_layout"] = {"unavailable": str(e)}
    # warm up the workspaces and captured graphs once (not timed)
    S.cg(M64, b, 2, 1e-10)
    for Ml in (M32, M16):
        S.ir_cg(M64, Ml, b, rtol=1e-2, inner_rtol=1e-1, max_outer=1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    x64, k = S.cg(M64, b, 100000, 1e-10)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t
    x64n = torch.linalg.norm(x64).item()
    out["fp64"] = {"seconds": t64, "iterations": k, "true_rel_residual": true_rel(x64),
                   "ms_per_iteration": 1e3 * t64 / max(k, 1)}
    log(f"[cg] fp64: {t64:.3f}s {k} its")
    runs = [("fp32_ir", M32, -1), ("bf16_ir", M16, -1)]
    if "patterns" in out["sliced_layout"]:
        runs += [("fp32_ir_sliced", M32, S.SLICED_VARIANT), ("bf16_ir_sliced", M16, S.SLICED_VARIANT)]
        M64.set_kernel(S.SLICED_VARIANT)
        S.cg(M64, b, 2, 1e-10)
        torch.cuda.synchronize()
        t = time.perf_counter()
        xs, ks = S.cg(M64, b, 100000, 1e-10)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t
        out["fp64_sliced"] = {"seconds": ts, "iterations": ks, "true_rel_residual": true_rel(xs),
                              "ms_per_iteration": 1e3 * ts / max(ks, 1),
                              "speedup_vs_fp64": t64 / ts}
        log(f"[cg] fp64_sliced: {ts:.3f}s {ks} its")
        M64.set_kernel(-1)
    for name, Ml, var in runs:
        # the sliced configurations keep both matrices in the sliced layout
        # (the outer FP64 residual too)
        Ml.set_kernel(var)
        M64.set_kernel(var if var == S.SLICED_VARIANT else -1)
        S.ir_cg(M64, Ml, b, rtol=1e-2, inner_rtol=1e-1, max_outer=1)  # layout + graph warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        x, st = S.ir_cg(M64, Ml, b, rtol=1e-10, inner_rtol=1e-4)
        torch.cuda.synchronize()
        tm = time.perf_counter() - t
        out[name] = {"seconds": tm, "outer": st["outer_its"], "inner_iterations": st["inner_its"],
                     "inner_per_outer": st["inner_hist"], "true_rel_residual": true_rel(x),
                     "rel_err_vs_fp64_cg": torch.linalg.norm(x - x64).item() / x64n,
                     "speedup_vs_fp64": t64 / tm}
        if var == S.SLICED_VARIANT and "fp64_sliced" in out:
            out[name]["speedup_vs_fp64_sliced"] = out["fp64_sliced"]["seconds"] / tm
        log(f"[cg] {name}: {tm:.3f}s outer {st['outer_its']} inner {st['inner_its']}")
        # variant beside the configuration: each inner solve asks only for
        # the reduction that brings the outer residual to rtol / 2
        torch.cuda.synchronize()
        t = time.perf_counter()
        x, st = S.ir_cg(M64, Ml, b, rtol=1e-10, inner_rtol=1e-4, adaptive_inner=True)
        torch.cuda.synchronize()
        ta = time.perf_counter() - t
        out[name + "_adaptive_inner"] = {
            "seconds": ta, "outer": st["outer_its"], "inner_iterations": st["inner_its"],
            "inner_per_outer": st["inner_hist"], "true_rel_residual": true_rel(x),
            "speedup_vs_fp64": t64 / ta,
            "variant": "inner rtol max(1e-4, 0.5e-10 / outer relative residual)"}
        if var == S.SLICED_VARIANT and "fp64_sliced" in out:
            out[name + "_adaptive_inner"]["speedup_vs_fp64_sliced"] = out["fp64_sliced"]["seconds"] / ta
        log(f"[cg] {name} adaptive: {ta:.3f}s inner {st['inner_its']}")
        Ml.set_kernel(-1)
        M64.set_kernel(-1)
    if cpu_iters > 0:
        oc = O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val)
        bh = b.cpu().numpy()
        # all cores: the whole solve to 1e-10 (threaded oracle)
        t = time.perf_counter()
        xo, ko, _ = O.cg(oc, bh, 100000, 1e-10, threads=nthreads)
        tc = time.perf_counter() - t
        out["cpu_all_cores"] = {"threads": nthreads, "seconds": tc, "iterations": ko,
                                "rel_diff_vs_gpu_fp64": float(np.linalg.norm(xo - x64.cpu().numpy())
                                                              / np.linalg.norm(xo)),
                                "gpu_fp64_speedup": tc / t64}
        log(f"[cg] CPU {nthreads} threads: {tc:.2f}s {ko} its")
        # one thread: an estimate from a sample of iterations (a whole solve
        # would take minutes)
        t = time.perf_counter()
        O.cg(oc, bh, cpu_iters, 1e-30)
        tc1 = time.perf_counter() - t
        out["cpu_1_thread_estimate"] = {"ms_per_iteration": 1e3 * tc1 / cpu_iters,
                                        "estimated_seconds": tc1 / cpu_iters * ko,
                                        "sample": f"{cpu_iters} oracle CG iterations, 1 thread, "
                                                  "scaled to the all-cores iteration count"}
    del M16, M32, M64
    return out


def run_batch_section(torch, S, P, nsys=256, nrhs=4, rtol=1e-6, seed=17):
    """256 SPD systems (rows log-uniform in [1e3, 2e5], power-law row
    lengths in [3, 64], diagonal 1.1 x |off-diagonal row sum|, FP32), 4 RHS
    each, one persistent hipGraph-captured CG launch to rtol."""
    import numpy as np
    from spmv_amd import autotune as T
    t0 = time.time()
    ns = np.exp(np.log(1e3) + (np.log(2e5) - np.log(1e3)) * P.uniform(seed, nsys)).astype(np.int64)
    systems = [P.random_spd(int(n), 1000 + i) for i, n in enumerate(ns)]
    B = S.Batch(systems, nrhs=nrhs)
    log(f"[batch] {nsys} systems, {B.nrows} rows, {B.nnz} nnz, generated in {time.time() - t0:.1f}s")
    b = torch.from_numpy(P.uniform(23, B.nrows * nrhs, -1, 1).astype(np.float32)
                         .reshape(B.nrows, nrhs)).to("cuda:0")
    best, table = T.tune_batch(B, b, max_its=20000, rtol=rtol, reps=1)
    x, its = B.solve(b, max_its=20000, rtol=rtol)
    torch.cuda.synchronize()
    t = time.perf_counter()
    B.solve(b, x, 20000, rtol, iterations=False)
    torch.cuda.synchronize()
    ts = time.perf_counter() - t
    # true residuals of a sample of systems (CPU, float64)
    import oracle as O
    worst = 0.0
    xh, bh = x.cpu().numpy(), b.cpu().numpy()
    for s_ in list(range(0, nsys, 32)):
        A = systems[s_]
        lo, hi = B.sys_row[s_], B.sys_row[s_ + 1]
        # the system solved: values rounded to FP32
        oc = O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val.astype(np.float32).astype(np.float64))
        for j in range(nrhs):
            r = bh[lo:hi, j].astype(np.float64) - O.spmv(oc, xh[lo:hi, j].astype(np.float64))
            worst = max(worst, float(np.linalg.norm(r) / np.linalg.norm(bh[lo:hi, j])))
    # algorithmic bytes: every right-hand side's iterations x (FP32 SpMV bytes
    # + 9 vector passes), the kernel reading each matrix once for all its RHS
    n_s = np.array([A.nrows for A in systems], np.float64)
    nnz_s = np.array([A.nnz for A in systems], np.float64)
    per_it = nnz_s * 8 + 4 * (n_s + 1) + 8 * n_s + 9 * 4 * n_s
    alg = float(np.sum(its.astype(np.float64) * per_it[:, None]))
    # what one iteration of a system moves when its RHS share the matrix
    shared = float(np.sum(its.max(axis=1) * (nnz_s * 8 + 4 * (n_s + 1)))
                   + np.sum(its.astype(np.float64) * (8 * n_s + 9 * 4 * n_s)[:, None]))
    out = {"workload": f"{nsys} random SPD systems x {nrhs} RHS, {B.nrows} rows, {B.nnz} nnz, FP32, rtol {rtol}",
           "seconds": ts, "systems_per_s": nsys / ts, "rhs_solves_per_s": nsys * nrhs / ts,
           "algorithmic_bytes": alg, "algorithmic_GBps": alg / ts / 1e9,
           "bytes_matrix_shared_by_rhs": shared, "shared_GBps": shared / ts / 1e9,
           "iterations_max": int(its.max()), "iterations_mean": float(its.mean()),
           "config": B.config(), "tuned": best, "worst_true_rel_residual_sampled": worst}
    log(f"[batch] {ts:.3f}s its max {its.max()} mean {its.mean():.1f} config {B.config()} "
        f"{alg / ts / 1e9:.0f} GB/s algorithmic, {shared / ts / 1e9:.0f} GB/s with shared matrix reads")
    del B
    return out


def run_adjoint_section(torch, S, P, fd_entries, n=1024, rtol=1e-13, seed=1):
    """2D anisotropic diffusion 1024^2, theta ~ lognormal(0, 0.5), b ~ U[0,1],
    x_target ~ N(0,1); loss 0.5||x - x_t||^2; forward CG, adjoint CG + SDDMM
    through torch autograd; central differences (eps 1e-6) on seeded
    entries."""
    import numpy as np
    from spmv_amd import adjoint as AD
    dev = "cuda:0"
    ne = P.diffusion_num_edges(n)
    theta = np.exp(0.5 * P.normal(seed, ne))
    b = P.uniform(seed + 1, n * n, 0.0, 1.0)
    xt = P.normal(seed + 2, n * n)
    op = AD.DiffusionOperator(n)
    th = torch.from_numpy(theta).to(dev).requires_grad_(True)
    bd = torch.from_numpy(b).to(dev).requires_grad_(True)
    xtd = torch.from_numpy(xt).to(dev)
    # warm-up: one whole untimed solve (workspaces, captured graphs, the
    # allocator's blocks); a 2-iteration warm-up left 0.04-0.2 s of one-off
    # costs in the timed forward solve
    op.solve(th.detach(), bd.detach(), rtol=rtol)
    torch.cuda.synchronize()
    t = time.perf_counter()
    x = op.solve(th, bd, rtol=rtol)
    loss = 0.5 * ((x - xtd) ** 2).sum()
    torch.cuda.synchronize()
    tf = time.perf_counter() - t
    t = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    tb = time.perf_counter() - t
    # SDDMM kernel alone (grad_A = -lam x^T on the pattern)
    lam = bd.grad.detach()
    g = torch.empty(op.A.nnz, dtype=torch.float64, device=dev)
    sd_ms = timed_loop(torch, lambda: S.sddmm(op.A, lam, x.detach(), alpha=-1.0, out=g), 20)
    sd_bytes = op.A.nnz * (4 + 8) + 4 * (op.A.nrows + 1) + 2 * 8 * op.A.nrows
    its = getattr(op.A, "cg_iterations", {})
    out = {"workload": f"anisotropic 5-point diffusion {n}^2, {op.A.nnz} nnz, {ne} edge conductivities",
           "forward_seconds": tf, "backward_seconds": tb, "rtol": rtol, "loss": loss.item(),
           "forward_cg_iterations": its.get("forward"), "adjoint_cg_iterations": its.get("adjoint"),
           "sddmm_ms": sd_ms, "sddmm_GBps": sd_bytes / sd_ms / 1e6}
    log(f"[adjoint] forward {tf:.3f}s ({its.get('forward')} its) backward {tb:.3f}s "
        f"({its.get('adjoint')} its) sddmm {sd_ms:.3f} ms")
    if fd_entries > 0:
        entries = (P.rand_u64(77, fd_entries) % np.uint64(ne)).astype(np.int64)
        t = time.perf_counter()
        fd = AD.central_difference(op, th.detach(), bd.detach(), xtd, entries, eps=1e-6, rtol=rtol)
        gt = th.grad.cpu().numpy()[entries]
        rel = np.abs(fd - gt) / np.maximum(np.abs(gt), 1e-300)
        out.update({"fd_entries": int(fd_entries), "fd_eps": 1e-6, "fd_max_rel_err": float(rel.max()),
                    "fd_median_rel_err": float(np.median(rel)), "fd_seconds": time.perf_counter() - t})
        log(f"[adjoint] FD on {fd_entries} entries: max rel err {rel.max():.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=160, help="owned grid edge of the 27-point share")
    ap.add_argument("--cg-n", type=int, default=256, help="grid edge of the CG comparison")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--fd-entries", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--mode", choices=["bench", "spmv", "copy", "read", "batch"], default="bench",
                    help="spmv: only the SpMV kernel loop; copy: the calibration copy kernel; "
                         "read: the streaming-read ceiling kernel; batch: the batched solve "
                         "alone (fixed launch configuration)")
    ap.add_argument("--variant", type=int, default=-1,
                    help="--mode spmv: SpMV kernel variant (-1: the library default)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the multi-GPU exchange is not part of this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import spmv_amd as S
    from spmv_amd import problems as P

    dev = "cuda:0"
    torch.cuda.set_device(0)

    if args.mode in ("copy", "read"):
        n = 1 << 28  # 2 GiB of doubles
        src = torch.rand(n, dtype=torch.float64, device=dev)
        if args.mode == "copy":
            dst = torch.empty_like(src)
            fn, written = (lambda: S.copy_f64(dst, src)), 8 * n
        else:
            sink = torch.zeros(1, dtype=torch.float64, device=dev)
            fn, written = (lambda: S.read_f64(src, sink)), 0
        for _ in range(args.warmup):
            fn()
        ms = timed_loop(torch, fn, args.steps)
        print(json.dumps({"mode": args.mode, "bytes_read": 8 * n, "bytes_written": written,
                          "ms": ms, "GBps": (8 * n + written) / ms / 1e6}))
        return

    if args.mode == "batch":
        ns = np.exp(np.log(1e3) + (np.log(2e5) - np.log(1e3)) * P.uniform(17, 256)).astype(np.int64)
        systems = [P.random_spd(int(m), 1000 + i) for i, m in enumerate(ns)]
        Bt = S.Batch(systems, nrhs=4)
        Bt.set_config(1024, 4, 4096)
        b = torch.from_numpy(P.uniform(23, Bt.nrows * 4, -1, 1).astype(np.float32)
                             .reshape(Bt.nrows, 4)).to(dev)
        x, its = Bt.solve(b, max_its=20000, rtol=1e-6)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(max(args.steps, 1)):
            Bt.solve(b, x, 20000, 1e-6, iterations=False)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "batch", "seconds": (time.perf_counter() - t) / max(args.steps, 1),
                          "its_max": int(its.max()), "its_mean": float(its.mean()),
                          "config": Bt.config()}))
        return

    n = args.n
    t0 = time.time()
    A = P.stencil(27, n, n, n, ghost_planes=1)
    log(f"[spmv] generated 27-point {n}^3 + ghost plane: {A.nnz} nnz in {time.time() - t0:.1f}s")
    l2g = S.L2GMap(**P.slab_decomposition(n, n, n))
    halo = P.uniform(2, l2g.num_ghosts(), -1, 1)
    l2g.set_halo(halo)
    M = S.Matrix(A.rowptr, A.colidx, A.val, col_map=l2g, row_map=S.L2GMap(A.nrows))
    x = torch.from_numpy(P.uniform(1, A.ncols, -1, 1)).to(dev)
    y = torch.empty(A.nrows, dtype=torch.float64, device=dev)
    B = spmv_bytes(A.nnz, A.nrows, A.ncols)
    log(f"[spmv] matrix on device: {M.nblocks} row blocks, {B} algorithmic bytes per SpMV")

    def step():
        l2g.update(x)
        M.mult(x, out=y)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    if args.mode == "spmv":
        M.set_kernel(args.variant)
        M.mult(x, out=y)
        ms = timed_loop(torch, lambda: M.mult(x, out=y), args.steps)
        print(json.dumps({"mode": "spmv", "kernel": M.kernel_name(), "ms": ms,
                          "GBps": B / ms / 1e6}))
        return

    # ---- timed steps (the JSON value): update + product as the reference's
    # matvec loop runs them.  Every fifth product is launched with two HIP
    # events that the kernel itself records at its start and end
    # (spmvx_spmv_timed: hipExtLaunchKernel on the stream the product runs on,
    # torch's current stream): the kernel's duration inside the timed steps
    # (the roofline).  Such events lengthen a step by 7-10 us of dispatch, so
    # only that sample of the steps carries them.
    cur = torch.cuda.current_stream()
    sample = {i for i in range(args.steps) if i % 5 == 4} or {args.steps - 1}
    ev = {i: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for i in sample}
    e_all0, e_all1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    e_all0.record(cur)
    for i in range(args.steps):
        l2g.update(x)
        if i in ev:
            M.mult_timed(x, y, ev[i][0], ev[i][1], stream=cur)
        else:
            M.mult(x, out=y)
    e_all1.record(cur)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    step_ms_ev = e_all0.elapsed_time(e_all1) / args.steps
    ms_per_step = 1e3 * wall / args.steps
    value = B / (ms_per_step * 1e-3) / 1e9
    per_step_ms = sorted(a.elapsed_time(b) for a, b in ev.values())
    spmv_ms = sum(per_step_ms) / len(per_step_ms)
    achieved = B / (spmv_ms * 1e-3) / 1e9
    # the same kernel back to back (no update between), for reference
    spmv_alone_ms = timed_loop(torch, lambda: M.mult(x, out=y), max(args.steps, 20))

    def in_step_kernel_ms(nsteps):
        ev2 = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(nsteps)]
        torch.cuda.synchronize()
        for i in range(nsteps):
            l2g.update(x)
            M.mult_timed(x, y2, ev2[i][0], ev2[i][1], stream=cur)
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in ev2) / nsteps

    # the same steps with the 32-bit-index CSR kernel (the plain CSR arrays),
    # when the default reads 16-bit column indices
    variant, kname = M.kernel()
    int32_kernel = None
    if "16-bit" in kname:
        y2 = torch.empty_like(y)
        M.set_kernel(int(os.environ.get("SPMV_INT32_VARIANT", "5")))
        M.mult(x, out=y2)
        for _ in range(args.warmup):
            l2g.update(x)
            M.mult(x, out=y2)
        i32_ms = in_step_kernel_ms(args.steps)
        int32_kernel = {"kernel": M.kernel_name(), "kernel_ms": round(i32_ms, 5),
                        "achieved": round(B / (i32_ms * 1e-3) / 1e9, 1),
                        "frac": round(B / (i32_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "bit_identical": bool(torch.equal(y, y2))}
        log(f"[spmv] 32-bit-index kernel in the steps {i32_ms:.4f} ms")
        M.set_kernel(variant)
    # the two phases of the reference demo's step (matvecmain.cpp:72-80),
    # each timed alone: L2GMap::update, then the product
    update_ms = timed_loop(torch, lambda: l2g.update(x), max(args.steps, 20))
    # PCIe-inclusive step: the halo sent again from the host every step
    # (set_halo: one synchronous host-to-device copy of 204.8 KB)
    npcie = max(args.steps // 2, 10)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(npcie):
        l2g.set_halo(halo)
        l2g.update(x)
        M.mult(x, out=y)
    torch.cuda.synchronize()
    pcie_ms = 1e3 * (time.perf_counter() - t) / npcie
    log(f"[spmv] step {ms_per_step:.4f} ms (events {step_ms_ev:.4f}), SpMV kernel "
        f"{spmv_ms:.4f} ms in the steps ({len(per_step_ms)} sampled; alone "
        f"{spmv_alone_ms:.4f} ms) = {achieved:.0f} GB/s; with the halo from the host each step "
        f"{pcie_ms:.3f} ms")

    # ---- the same product with the sliced row-pattern layout (informational)
    sliced = None
    try:
        info = M.sliced_info()
        M.set_kernel(S.SLICED_VARIANT)
        y2 = torch.empty_like(y)
        M.mult(x, out=y2)
        s_ms = timed_loop(torch, lambda: M.mult(x, out=y2), max(args.steps, 20))
        Bs = 8 * info["slots"] + 2 * A.nrows + 8 * A.ncols + 8 * A.nrows
        sliced = {"kernel": "sliced_spmv (kernel variant %d)" % S.SLICED_VARIANT,
                  "kernel_ms": round(s_ms, 5), "bytes_per_spmv": Bs,
                  "achieved_GBps": round(Bs / (s_ms * 1e-3) / 1e9, 1),
                  "csr_equivalent_GBps": round(B / (s_ms * 1e-3) / 1e9, 1),
                  "speedup_vs_csr": round(spmv_ms / s_ms, 3),
                  "patterns": info["npatterns"],
                  "bit_identical_to_csr": bool(torch.equal(y, y2))}
        log(f"[spmv] sliced layout {s_ms:.4f} ms ({info['npatterns']} patterns)")
        M.set_kernel(-1)
    except S.SpmvError as e:
        sliced = {"unavailable": str(e)}

    # ---- ceilings measured on this GPU in the same run (library kernels over
    # 2 GiB, eight times the 256 MB last-level cache): a streaming read
    # (non-temporal 16-byte loads) and a copy (half the bytes written)
    nc = 1 << 28
    src = torch.rand(nc, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    sink = torch.zeros(1, dtype=torch.float64, device=dev)
    S.read_f64(src, sink)
    S.copy_f64(dst, src)
    read_ms = timed_loop(torch, lambda: S.read_f64(src, sink), 10)
    copy_ms = timed_loop(torch, lambda: S.copy_f64(dst, src), 10)
    ceilings = {"stream_read_GBps": round(8 * nc / (read_ms * 1e-3) / 1e9, 1),
                "copy_GBps": round(16 * nc / (copy_ms * 1e-3) / 1e9, 1),
                "bytes": 8 * nc,
                "spmv_frac_of_stream_read": None}
    ceilings["spmv_frac_of_stream_read"] = round(achieved / ceilings["stream_read_GBps"], 4)
    del src, dst, sink
    log(f"[ceiling] streaming read {ceilings['stream_read_GBps']:.0f} GB/s, "
        f"copy {ceilings['copy_GBps']:.0f} GB/s")

    traffic = None
    tf = os.path.join(ROOT, "profiles", "spmv_27pt_160_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"27-point stencil (diag 26, off -1) on a {n}^3 owned block plus one "
                         f"{n}x{n} ghost plane fed through L2GMap from a seeded host halo; "
                         "step = L2GMap::update + FP64 CSR SpMV"),
            "rows": A.nrows, "cols": A.ncols, "nnz": A.nnz,
            "algorithmic_bytes_per_spmv": B,
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel_ms": round(spmv_ms, 5),
                     "kernel_ms_alone": round(spmv_alone_ms, 5),
                     "kernel_ms_median": round(per_step_ms[len(per_step_ms) // 2], 5),
                     "kernel_ms_min": round(per_step_ms[0], 5),
                     "kernel_ms_samples": len(per_step_ms),
                     "kernel": M.kernel_name(),
                     "hbm_GBps_measured": (round(traffic / (spmv_ms * 1e-3) / 1e9, 1)
                                           if traffic else None),
                     "note": ("achieved = the algorithmic CSR bytes (4-byte column indices) "
                              "per product / kernel time; a kernel reading 16-bit column "
                              "indices moves fewer bytes (traffic), so hbm_GBps_measured is "
                              "the bandwidth it actually draws")},
        "csr_int32_indices": int32_kernel,
        "phases": {"update_ms": round(update_ms, 5), "spmv_ms_in_step": round(spmv_ms, 5),
                   "step_ms": round(step_ms_ev, 5),
                   "note": "demos/matvecmain.cpp:72-80 times update and SpMV separately; "
                           "update alone is timed back to back; spmv_ms_in_step is the "
                           "average of the timed steps' products that record their own start "
                           "and end events (every fifth step)"},
        "pcie_inclusive": {"ms_per_step": round(pcie_ms, 5), "GBps": round(B / (pcie_ms * 1e-3) / 1e9, 1),
                           "step": "set_halo (204.8 KB host-to-device, synchronous) + update + SpMV"},
        "spmv_sliced": sliced,
        "ceilings": ceilings,
    }

    if not args.no_cpu:
        import oracle as O
        oc = O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val)
        xh = x.cpu().numpy()
        reps, tc = 0, 0.0
        while tc < args.cpu_seconds or reps < 2:
            t = time.perf_counter()
            yc = O.spmv(oc, xh)
            tc += time.perf_counter() - t
            reps += 1
        cpu_gbs = B * reps / tc / 1e9
        same = bool(np.array_equal(yc, y.cpu().numpy()))
        ncores = cpu_threads()
        t = time.perf_counter()
        for _ in range(3):
            O.spmv(oc, xh, threads=ncores)
        tmc = (time.perf_counter() - t) / 3
        result["cpu_baseline"] = {"value": round(cpu_gbs, 3), "unit": "GB/s", "cores": 1,
                                  "kind": "port",
                                  "sample": f"{reps} oracle SpMVs of the same matrix and x "
                                            f"({tc:.1f}s), 1 thread; results bit-identical to "
                                            f"the GPU: {same}",
                                  "all_cores": {"cores": ncores,
                                                "value": round(B / tmc / 1e9, 3)}}
        log(f"[cpu] {cpu_gbs:.2f} GB/s (1 thread), {B / tmc / 1e9:.2f} GB/s ({ncores} threads)")

    # CG on the same share, zero halo (homogeneous neighbour), to 1e-10
    l2g.set_halo(None)
    b = torch.from_numpy(P.uniform(3, A.nrows, -1, 1)).to(dev)
    S.cg(M, b, 2, 1e-10)
    torch.cuda.synchronize()
    t = time.perf_counter()
    xs, k = S.cg(M, b, 100000, 1e-10)
    torch.cuda.synchronize()
    tcg = time.perf_counter() - t
    result["cg_27pt"] = {"seconds": round(tcg, 4), "iterations": k, "rtol": 1e-10,
                         "halo": "zero (homogeneous neighbour values for the direction vector)"}
    log(f"[cg27] {tcg:.3f}s {k} its")
    if not args.no_cpu:
        nt = cpu_threads()
        t = time.perf_counter()
        xo, ko, _ = O.cg(oc, b.cpu().numpy(), 100000, 1e-10, threads=nt)
        tco = time.perf_counter() - t
        result["cg_27pt"]["cpu_all_cores"] = {
            "threads": nt, "seconds": round(tco, 3), "iterations": ko,
            "rel_diff_vs_gpu": float(np.linalg.norm(xo - xs.cpu().numpy()) / np.linalg.norm(xo)),
            "gpu_speedup": round(tco / tcg, 1)}
        log(f"[cg27] CPU {nt} threads: {tco:.2f}s {ko} its")
        del xo
    del M, xs

    if not args.no_c1:
        import oracle as O
        result["c1_reference_cpu_config"] = run_c1_section(torch, S, P, O, not args.no_cpu,
                                                           cpu_threads())
    if not args.no_cg:
        import oracle as O
        result["cg"] = run_cg_section(torch, S, P, O, args.cg_n, 5,
                                      0 if args.no_cpu else 3, cpu_threads())
    if not args.no_batch:
        result["batch"] = run_batch_section(torch, S, P)
    if not args.no_adjoint:
        result["adjoint"] = run_adjoint_section(torch, S, P, args.fd_entries)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
