#!/usr/bin/env python3
"""Benchmark: milliseconds to converged eigenvectors on one MI355X.

A step is one full eigensolve of the workload graph's Laplacian -- from the
Laplacian resident in HBM to the n x (n_components+1) eigenvector block plus
its FP64 residual check resident on the device (max residual <= tol).  The
FP32-bulk solver is timed for --steps steps after --warmup untimed ones; the
plain GPU LOBPCG (FP64, block n_components+1, the reference's algorithm) is
timed beside it; the CPU oracle runs a bounded sample on the host.

Prints one JSON line (rank 0, N=1).  Workloads (synthetic, seeded):
  gmm1m   n=1,000,000, 30-d, 40 Zipf(1.2) clusters + 5 % uniform background,
          exact 15-NN fuzzy graph (default)
  gmm100k n=100,000, 50-d, 20 Dirichlet(2) clusters
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "spectral-init_amd"), os.path.join(ROOT, "oracle")]

HBM_PEAK_GBS = 8000.0
BF16_MFMA_PEAK = 2500.0  # TFLOP/s, dense bf16 on the MI355X matrix cores


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def largest_component(indptr, indices, data):
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    n = indptr.size - 1
    a = sp.csr_matrix((data, indices, indptr), shape=(n, n))
    nc, lab = cg.connected_components(a, directed=False)
    if nc == 1:
        return indptr, indices, data, n, 1
    big = np.bincount(lab).argmax()
    keep = np.flatnonzero(lab == big)
    sub = a[keep][:, keep].tocsr()
    sub.sort_indices()
    return (sub.indptr.astype(np.int64), sub.indices.astype(np.int32),
            sub.data.astype(np.float32), n, nc)


def cpu_baseline(lap, si, n_components, plain_iters, budget_s, threads=1):
    """Time the CPU oracle's LOBPCG iteration (k = n_components+1, FP64) on the
    same Laplacian for a bounded number of iterations, on `threads` host cores."""
    import oracle as O
    O.build()
    O.set_threads(threads)
    _, sq, L = lap.export()
    k = n_components + 1
    n = sq.size
    x0 = O.normal_block(42, 1, n * k).reshape(n, k)
    x0[:, 0] = sq / np.linalg.norm(sq)
    t = time.perf_counter()
    O.lobpcg_core(L, x0, tol=0.0, maxiter=2)
    per_it = (time.perf_counter() - t) / 2.0
    iters = int(max(3, min(500, budget_s / max(per_it, 1e-6))))
    t = time.perf_counter()
    r = O.lobpcg_core(L, x0, tol=0.0, maxiter=iters)
    dt = time.perf_counter() - t
    ms_it = 1e3 * dt / max(r["iterations"], 1)
    return dict(value=round(ms_it, 3), unit="ms per LOBPCG iteration (FP64, k=%d)" % k, cores=threads,
                kind="port",
                sample=f"{r['iterations']} LOBPCG iterations of the oracle on the full graph "
                       f"({dt:.1f} s); projected time to eigvecs at the plain GPU solver's "
                       f"{plain_iters} iterations = {ms_it * plain_iters:.0f} ms",
                projected_ms_to_eigvecs=round(ms_it * plain_iters, 1) if plain_iters else None)


def bench_graph_builder(args, torch, graphs, si, ctx):
    """The fuzzy k-NN graph W of the workload's own points built on the device
    (si_fuzzy_graph_build: exact k-NN, smooth_knn_dist, membership, fuzzy
    union, prune), the k-NN step's rate on the FP32 matrix cores that filter
    the candidates, and the CPU oracle's k-NN on a bounded sample of the same
    points."""
    import oracle as O
    x, _ = graphs.workload_points(args.workload, args.seed, args.n)
    n, d = x.shape
    k = 15
    si.fuzzy_graph(ctx, x[:4096], k).close()  # warm the kernels
    torch.cuda.synchronize()
    t = time.perf_counter()
    si.knn_exact(ctx, x, k)
    knn_ms = 1e3 * (time.perf_counter() - t)
    t = time.perf_counter()
    g = si.fuzzy_graph(ctx, x, k)
    graph_ms = 1e3 * (time.perf_counter() - t)
    nnz = g.nnz
    g.close()
    # algorithmic work of the exact k-NN for float32 points: the dot product of
    # each of the n (n - 1) pairs, 2 d flops; the per-point norms, the FP64
    # re-rank of ~k candidates per query and the point and list transfers are
    # < 1 % of it.  The time is the whole si_knn_exact call (upload, filter,
    # re-rank, download).  The filter takes each FP32 product as three bf16
    # part products over DB = 32 ceil((d + 2) / 32) k-slots (coordinates and
    # two norm slots), so the bf16 matrix cores execute 6 DB flops per pair and
    # their dense peak caps the algorithmic rate at peak * 2 d / (6 DB).
    flops = 2.0 * d * n * (n - 1)
    tflops = flops / (knn_ms * 1e-3) / 1e12
    db = 32 * ((d + 2 + 31) // 32)
    executed = 6.0 * db * n * (n - 1) / (knn_ms * 1e-3) / 1e12
    ceiling = BF16_MFMA_PEAK * 2.0 * d / (6.0 * db)
    direct = si.knn_path(-1)
    log(f"graph builder: k-NN {knn_ms:.1f} ms ({tflops:.1f} TFLOP/s algorithmic, "
        f"{executed:.0f} TFLOP/s executed on the bf16 matrix cores, {direct} queries re-run "
        f"by the direct FP64 kernel), whole graph {graph_ms:.1f} ms, nnz {nnz}")
    cpu = None
    if not args.no_cpu:
        O.build()
        O.set_threads(1)
        m = min(n, 12000)  # ~0.4 s of single-thread work: a stable per-pair rate
        t = time.perf_counter()
        O.knn_exact(x[:m].astype(np.float64), k)
        dt = time.perf_counter() - t
        per_pair_ns = 1e9 * dt / (m * (m - 1))
        cpu = dict(unit="ns per k-NN pair (FP64, 1 thread)", value=round(per_pair_ns, 3), cores=1,
                   kind="port",
                   sample=f"oracle exact k-NN of the first {m} points ({dt:.2f} s); projected "
                          f"for all {n} points: {per_pair_ns * n * (n - 1) * 1e-9:.0f} s")
        log(f"graph builder CPU oracle: {cpu['value']} ns/pair -> {cpu['sample']}")
    return dict(points=n, dim=d, k=k, nnz=nnz, ms=round(graph_ms, 1), knn_ms=round(knn_ms, 1),
                knn_direct_queries=direct,
                knn_roofline=dict(bound="mfma", achieved=round(tflops, 2),
                                  peak=round(ceiling, 1), unit="TFLOP/s",
                                  frac=round(tflops / ceiling, 4), flops_per_pair=2 * d,
                                  executed_bf16_tflops=round(executed, 1),
                                  executed_frac_of_bf16_peak=round(executed / BF16_MFMA_PEAK, 4)),
                cpu_baseline=cpu)


def bench_warm_start(args, torch, si, ctx, lap, cfg, fast_ms, plain):
    """The committed warm-start network's block (spectral-init_amd/warmstart_mpnn.npy)
    fed to the FP32-bulk solver and to the plain LOBPCG, against their random
    starts; the network's own cost is timed separately."""
    path = os.path.join(ROOT, "spectral-init_amd", "warmstart_mpnn.npy")
    net = si.WarmStartNet(ctx, params=np.load(path))
    net.warm_start(lap, seed=42)
    ts = []
    for _ in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        blk, used, rn, rr = net.warm_start(lap, seed=42)
        torch.cuda.synchronize()
        ts.append(1e3 * (time.perf_counter() - t))
    net_ms = float(np.median(ts))
    fcfg = si.SpectralInitConfig(tol=cfg.tol, block_size=cfg.block_size, cheb_degree=cfg.cheb_degree,
                                 solver=si.Solver.Fast, warm_start=blk.data_ptr())
    si.solve_eigenproblem(lap, 2, 42, fcfg)
    ft, frep = [], None
    for _ in range(5):
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, _, frep = si.solve_eigenproblem(lap, 2, 42, fcfg)
        torch.cuda.synchronize()
        ft.append(1e3 * (time.perf_counter() - t))
    pcfg = si.SpectralInitConfig(tol=cfg.tol, solver=si.Solver.PlainLobpcg, warm_start=blk.data_ptr())
    si.solve_eigenproblem(lap, 2, 42, pcfg)
    pt, prep = [], None
    for _ in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, _, prep = si.solve_eigenproblem(lap, 2, 42, pcfg)
        torch.cuda.synchronize()
        pt.append(1e3 * (time.perf_counter() - t))
    fast_w, plain_w = float(np.median(ft)), float(np.median(pt))
    plain_r = float(np.median([p["ms"] for p in plain])) if plain else None
    out = dict(network_ms=round(net_ms, 2), network_block_used=bool(used),
               network_block_residual=rn, random_block_residual=rr,
               fast_ms_with_block=round(fast_w, 3), fast_rounds_with_block=frep.iterations,
               fast_ms_with_block_incl_network=round(fast_w + net_ms, 3),
               fast_ms_random_start=round(fast_ms, 3),
               plain_ms_with_block=round(plain_w, 2),
               plain_ms_with_block_incl_network=round(plain_w + net_ms, 2),
               plain_iterations_with_block=prep.iterations,
               plain_ms_random_start=plain_r,
               plain_iterations_random_start=plain[0]["iterations"] if plain else None,
               max_residual_with_block=max(frep.max_residual, prep.max_residual))
    log(f"warm start: network {net_ms:.1f} ms (used {used}); fast {fast_w:.2f} ms "
        f"({frep.iterations} rounds) vs {fast_ms:.2f} random; plain {plain_w:.1f} ms "
        f"({prep.iterations} iterations) vs {plain_r} random")
    return out


def locality_report(si, lap, workload):
    """Where the columns of the stored operator's entries lie relative to their
    rows: the share of entries with |col - row| below a distance, the share
    inside a row block of R rows widened by H rows each side (the x rows one
    workgroup could stage in LDS: 64 bytes per 16-column FP32 row), and the
    share inside the row's own eighth of the rows (one XCD's range)."""
    rp, col, _, _, order = si.laplacian_storage(lap)
    n = rp.size - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp.astype(np.int64)))
    c = col.astype(np.int64)
    d = np.abs(c - rows)
    out = {"workload": workload, "n": n, "stored_entries": int(col.size),
           "order": {si.ORDER_INPUT: "input", si.ORDER_RCM: "rcm",
                     si.ORDER_CLUSTER: "cluster"}.get(order, str(order)),
           "within_distance": {}, "within_row_block_window": {}}
    for w in (256, 1024, 2048, 4096, 16384, 65536, 131072):
        out["within_distance"][str(w)] = round(float(np.mean(d < w)), 4)
    for r, h in ((256, 896), (1024, 512), (1024, 1024), (4096, 0)):
        b0 = (rows // r) * r
        key = f"block {r} rows, halo {h} ({(r + 2 * h) * 64 // 1024} KB of x)"
        out["within_row_block_window"][key] = round(float(np.mean((c >= b0 - h) & (c < b0 + r + h))), 4)
    per = n // 8
    out["within_xcd_range"] = round(float(np.mean((rows // per) == (c // per))), 4)
    return out


def bench_components(args, torch, graphs, si):
    """64 disconnected components (sizes log-uniform in [50, 200 000], each an
    independent 30-d GMM 15-NN graph): spectral_init with all components in one
    batched device solve vs one solve per component (the reference's loop)."""
    t0 = time.perf_counter()
    if args.graph_cache and os.path.exists(args.graph_cache):
        z = np.load(args.graph_cache)
        ip, ix, wt, x, sizes = z["indptr"], z["indices"], z["data"], z["x"], list(z["sizes"])
    else:
        w = graphs.workload("comp64", seed=args.seed, device="cuda")
        ip, ix, wt, x, sizes = w["indptr"], w["indices"], w["data"], w["x"], w["sizes"]
        if args.graph_cache:
            np.savez(args.graph_cache, indptr=ip, indices=ix, data=wt, x=x, sizes=np.array(sizes))
    n, nnz = ip.size - 1, int(ip[-1])
    log(f"graph: n={n} nnz={nnz} components={len(sizes)} sizes {min(sizes)}..{max(sizes)} "
        f"in {time.perf_counter() - t0:.1f}s")
    ctx = si.Context(0)
    res = {}
    # legs: the batched solve; the reference's loop (level 0 = dense EVD below
    # 2000 nodes); the loop with the batched path's levels (dense below 65 nodes)
    legs = ((True, 2000),) if args.batched_only else ((True, 2000), (False, 2000), ("loop65", 65))
    for batch, dense_thr in legs:
        cfg = si.SpectralInitConfig(tol=args.tol, batch_components=batch is True,
                                    dense_threshold=dense_thr)
        runs = []
        for i in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            out, rep = si.spectral_init((ip, ix, wt), args.n_components, 42, data=x, config=cfg,
                                        ctx=ctx, return_report=True)
            dt = 1e3 * (time.perf_counter() - t)
            if i >= args.warmup:
                runs.append(dict(end_to_end_ms=dt, solve_ms=rep.solve_ms, laplacian_ms=rep.laplacian_ms,
                                 rounds=rep.iterations, spmm_calls=rep.spmm_calls,
                                 max_residual=rep.max_residual))
            log(f"batch={batch} run {i}: {dt:.1f} ms end-to-end, solve {rep.solve_ms:.2f} ms, "
                f"laplacian {rep.laplacian_ms:.1f} ms, residual {rep.max_residual:.2e}")
        res[batch] = (runs, out)
    cap, launches = si.batch_graphs()
    med = lambda runs, key: float(np.median([r[key] for r in runs]))
    bs = med(res[True][0], "solve_ms")
    ls = med(res[False][0], "solve_ms") if False in res else None
    loops = False in res
    out = {
        "metric": "ms to eigvecs of 64 components @ residual<=%g (batched device solve)" % args.tol,
        "value": round(bs, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bs, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (64 independent seeded 30-d GMM exact 15-NN fuzzy graphs)",
        "config": {"workload": "comp64", "n": n, "nnz": nnz, "components": len(sizes),
                   "n_components": args.n_components, "tol": args.tol},
        "batched": {"solve_ms": bs, "end_to_end_ms": med(res[True][0], "end_to_end_ms"),
                    "laplacian_ms": med(res[True][0], "laplacian_ms"),
                    "graphs_captured": cap, "graph_launches": launches, "runs": res[True][0]},
        "per_component_loop": {"solve_ms": ls, "end_to_end_ms": med(res[False][0], "end_to_end_ms"),
                               "laplacian_ms": med(res[False][0], "laplacian_ms"),
                               "levels": "reference: dense EVD below 2000 nodes",
                               "runs": res[False][0]} if loops else None,
        "per_component_loop_same_levels": {
            "solve_ms": med(res["loop65"][0], "solve_ms"),
            "end_to_end_ms": med(res["loop65"][0], "end_to_end_ms"),
            "levels": "dense EVD below 65 nodes, as the batched path",
            "runs": res["loop65"][0]} if loops else None,
        "speedup_solve_vs_loop": round(ls / bs, 2) if loops else None,
        "speedup_solve_vs_loop_same_levels":
            round(med(res["loop65"][0], "solve_ms") / bs, 2) if loops else None,
        "max_abs_coord_diff_vs_loop":
            float(np.abs(res[True][1] - res[False][1]).max()) if loops else None,
    }
    print(json.dumps(out), flush=True)


def read_traffic(path, workload, block, order):
    """HBM bytes per launch of the dominant SpMM, from a committed rocprofv3
    --pmc summary (lines "workload: ...", "block: ...", "order: ...",
    "hbm_bytes_per_launch: ...").  None when the summary is absent or measured
    another workload, width or node order."""
    if not path or not os.path.exists(path):
        return None
    kv = {}
    with open(path) as fh:
        for line in fh:
            if ":" in line and not line.startswith(("#", " ", "|")):
                key, _, val = line.partition(":")
                kv[key.strip()] = val.strip()
    if kv.get("workload") != workload or kv.get("block") != str(block) or kv.get("order") != order:
        return None
    try:
        return int(float(kv["hbm_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="gmm1m",
                    choices=["gmm1m", "gmm100k", "blobs2k", "band1m", "comp64"])
    ap.add_argument("--n", type=int, default=None, help="override the node count")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--n-components", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--degree", type=int, default=0, help="max filter degree per round, 0 = automatic")
    ap.add_argument("--plain-runs", type=int, default=10,
                    help="timed plain GPU LOBPCG runs (the median is the speed-up's denominator)")
    ap.add_argument("--plain-warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--e2e-runs", type=int, default=3,
                    help="timed si_spectral_init runs on the host arrays (after one warm-up)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--spmm-only", action="store_true",
                    help="time the SpMM alone for widths 1..64 and exit")
    ap.add_argument("--spmm-sweep", action="store_true",
                    help="also time the SpMM alone for widths 1..64 (FP32 and FP64)")
    ap.add_argument("--rcm", action="store_true",
                    help="store the operator in reverse Cuthill-McKee order (not the clustering order)")
    ap.add_argument("--no-reorder", action="store_true",
                    help="keep the input node order (no locality reordering)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "spmm_wide_f32_gmm1m_traffic.md"),
                    help="rocprofv3 --pmc summary holding the measured HBM bytes per SpMM launch")
    ap.add_argument("--rsvd-runs", type=int, default=1,
                    help="randomized SVD runs beside the solver (oversampling 10, 4 power iterations)")
    ap.add_argument("--graph-cache", default=None,
                    help="load the workload graph from this .npz (written on first use)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="one solve plus a 1 GiB device copy (scale reference for HBM counters), then exit")
    ap.add_argument("--graph-only", action="store_true",
                    help="time the device fuzzy k-NN graph builder alone, print its record and exit")
    ap.add_argument("--no-graph-builder", action="store_true",
                    help="skip timing the device fuzzy k-NN graph builder on the workload's points")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the learned warm start leg (network block fed to both solvers)")
    ap.add_argument("--batched-only", action="store_true",
                    help="comp64: time the batched spectral_init only (no per-component loops)")
    ap.add_argument("--locality", action="store_true",
                    help="print how close each stored entry's column lies to its row in the "
                         "operator's stored node order (the reach of an LDS-staged x block) and exit")
    ap.add_argument("--spmm-shape16", type=int, default=-1,
                    help="pin the FP32 16-column SpMM shape (0-3) instead of autotuning; "
                         "counter runs pin the shape the autotuner picks unprofiled")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(the design notes describe how the solve would shard); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import graphs
    import spectral_mi355x as si

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    si.lib()  # fails loudly if the native library is missing

    if args.workload == "comp64":
        return bench_components(args, torch, graphs, si)
    if args.graph_only:
        ctx = si.Context(0)
        print(json.dumps({"workload": args.workload,
                          "graph_builder": bench_graph_builder(args, torch, graphs, si, ctx)}),
              flush=True)
        return

    t0 = time.perf_counter()
    if args.graph_cache and os.path.exists(args.graph_cache):
        log(f"loading workload graph from {args.graph_cache}")
        z = np.load(args.graph_cache)
        ip, ix, wt = z["indptr"], z["indices"], z["data"]
        n_gen, ncomp = int(z["n_gen"]), int(z["ncomp"])
    else:
        log(f"building workload {args.workload} (n={args.n or 'default'}) ...")
        w = graphs.workload(args.workload, seed=args.seed, n=args.n, device="cuda")
        ip, ix, wt, n_gen, ncomp = largest_component(w["indptr"], w["indices"], w["data"])
        if args.graph_cache:
            np.savez(args.graph_cache, indptr=ip, indices=ix, data=wt, n_gen=n_gen, ncomp=ncomp)
    n = ip.size - 1
    nnz = int(ip[-1])
    log(f"graph: n={n} (generated {n_gen}, components {ncomp}) nnz={nnz} "
        f"in {time.perf_counter() - t0:.1f}s")

    ctx = si.Context(0)
    t0 = time.perf_counter()
    lap = si.Laplacian(ctx, ip, ix, wt, reorder="rcm" if args.rcm else not args.no_reorder)
    torch.cuda.synchronize()
    lap_ms = 1e3 * (time.perf_counter() - t0)
    log(f"laplacian on device: nnz={lap.nnz} hub rows={lap.n_long} ({lap_ms:.1f} ms incl. upload)")
    if args.spmm_shape16 >= 0:
        log(f"FP32 16-column SpMM shape pinned to {si.spmm_shapes(lap, args.spmm_shape16, 0)[0]}")

    if args.locality:
        print(json.dumps(locality_report(si, lap, args.workload)), flush=True)
        return

    sweep = []
    if args.spmm_sweep or args.spmm_only:
        for prec, dt in ((si.SI_F32, torch.float32), (si.SI_F64, torch.float64)):
            for kk in (1, 3, 8, 16, 32, 64):
                x = torch.randn(n, kk, dtype=dt, device="cuda")
                y = torch.empty_like(x)
                lap.spmm(x, y, prec)
                p = si.SpmmProfiler(ctx)
                with p:
                    for _ in range(10):
                        lap.spmm(x, y, prec)
                r = p.result
                gbs = r["bytes"] / (r["ms"] * 1e-3) / 1e9
                log(f"spmm sweep {'f32' if prec == si.SI_F32 else 'f64'} k={kk}: "
                    f"{1e3 * r['ms'] / r['launches']:.1f} us/launch, "
                    f"{r['bytes'] / r['launches'] / 1e6:.1f} MB -> {gbs:.0f} GB/s "
                    f"({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak)")
                sweep.append(dict(precision="f32" if prec == si.SI_F32 else "f64", k=kk,
                                  us=round(1e3 * r["ms"] / r["launches"], 2),
                                  mb=round(r["bytes"] / r["launches"] / 1e6, 2),
                                  gbs=round(gbs, 1)))
                del x, y
        if args.spmm_only:
            probe = {m: round(si.gather_probe(lap, i, 10), 2)
                     for i, m in enumerate(("column_stream_us", "random_rows_us",
                                            "random_rows_2MiB_us"))}
            log(f"gather probe: {probe}")
            print(json.dumps({"workload": args.workload, "n": n, "nnz": nnz,
                              "hub_rows": lap.n_long, "spmm": sweep, "gather_probe": probe}),
                  flush=True)
            return

    cfg = si.SpectralInitConfig(tol=args.tol, block_size=args.block, cheb_degree=args.degree,
                                solver=si.Solver.Fast)
    k = args.n_components + 1
    if args.pmc_probe:
        ev, vec, rep = si.solve_eigenproblem(lap, args.n_components, 42, cfg)
        torch.cuda.synchronize()
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": args.workload, "n": n, "nnz": nnz, "spmm_calls": rep.spmm_calls,
                          "block": rep.block_size, "copy_bytes_read": a.numel() * 4,
                          "copy_bytes_written": b.numel() * 4}), flush=True)
        return
    for i in range(args.warmup):
        ev, vec, rep = si.solve_eigenproblem(lap, args.n_components, 42, cfg)
        log(f"warmup {i}: {rep.solve_ms:.2f} ms level={rep.level} res={rep.max_residual:.2e} "
            f"rounds={rep.iterations} block={rep.block_size} spmm={rep.spmm_calls}")
        del vec

    times, reps = [], []
    prof = si.SpmmProfiler(ctx)
    torch.cuda.synchronize()
    with prof:
        for i in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            ev, vec, rep = si.solve_eigenproblem(lap, args.n_components, 42, cfg)
            torch.cuda.synchronize()
            dt = 1e3 * (time.perf_counter() - t)
            times.append(dt)
            reps.append(rep)
            if rep.level != 1 or not (rep.max_residual <= args.tol):
                log(f"step {i}: WARNING level={rep.level} residual={rep.max_residual:.3e}")
            log(f"step {i}: {dt:.2f} ms (device {rep.solve_ms:.2f} ms) res={rep.max_residual:.2e} "
                f"rounds={rep.iterations} spmm={rep.spmm_calls}")
            if i < args.steps - 1:
                del vec
    sp = prof.result
    fast_ms = float(np.median(times))
    spmm_gbs = sp["bytes_f32"] / (sp["ms_f32"] * 1e-3) / 1e9 if sp["ms_f32"] > 0 else None
    spmm_avg_us = 1e3 * sp["ms_f32"] / max(sp["launches_f32"], 1)
    bytes_per_launch = sp["bytes_f32"] / max(sp["launches_f32"], 1)
    log(f"fast solver median {fast_ms:.2f} ms; FP32 SpMM {sp['launches_f32']} launches avg "
        f"{spmm_avg_us:.1f} us, {bytes_per_launch / 1e6:.1f} MB algorithmic -> {spmm_gbs:.0f} GB/s")
    last = reps[-1]

    plain = []
    plain_iters = None
    for i in range(args.plain_warmup if args.plain_runs else 0):
        try:
            si.lobpcg_solve(lap, args.n_components, 42, tol=args.tol)
        except si.ConvergenceFailure:
            pass
    for i in range(args.plain_runs):
        torch.cuda.synchronize()
        t = time.perf_counter()
        try:
            pev, pvec, prep = si.lobpcg_solve(lap, args.n_components, 42, tol=args.tol)
            ok = prep.max_residual <= args.tol
        except si.ConvergenceFailure:
            prep, ok = None, False
        torch.cuda.synchronize()
        dt = 1e3 * (time.perf_counter() - t)
        plain.append(dict(ms=round(dt, 2), converged=bool(ok),
                          iterations=prep.iterations if prep else None,
                          restarts=prep.restarts if prep else None,
                          max_residual=prep.max_residual if prep else None))
        plain_iters = prep.iterations if prep else plain_iters
        log(f"plain LOBPCG run {i}: {dt:.1f} ms converged={ok} iters={plain[-1]['iterations']} "
            f"res={plain[-1]['max_residual']}")
    plain_ms = float(np.median([p["ms"] for p in plain])) if plain else None

    warm = None
    if not args.no_warm_start and args.n_components == 2:
        warm = bench_warm_start(args, torch, si, ctx, lap, cfg, fast_ms, plain)

    rsvd = []
    for i in range(args.rsvd_runs):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rev, rvec, rrep = si.rsvd_solve(lap, args.n_components, 42, oversampling=10, nbiter=4)
        torch.cuda.synchronize()
        dt = 1e3 * (time.perf_counter() - t)
        rres = si.max_eigenpair_residual(lap, rev, rvec)
        rsvd.append(dict(ms=round(dt, 2), max_residual=rres, eigenvalues=[float(v) for v in rev]))
        log(f"rSVD run {i}: {dt:.1f} ms residual={rres:.3e}")
        del rvec

    # the drop-in end to end: si_spectral_init on host CSR arrays (upload,
    # operator build with the device locality order, components, solve,
    # selection, scaling and noise, coordinates back on the host)
    e2e = []
    ecfg = si.SpectralInitConfig(tol=args.tol)
    for i in range(1 + args.e2e_runs):
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, erep = si.spectral_init((ip, ix, wt), args.n_components, 42, config=ecfg, ctx=ctx,
                                   return_report=True)
        dt = 1e3 * (time.perf_counter() - t)
        if i > 0:
            e2e.append(dict(ms=round(dt, 2), laplacian_ms=round(erep.laplacian_ms, 2),
                            solve_ms=round(erep.solve_ms, 2), max_residual=erep.max_residual))
        log(f"spectral_init end to end run {i}: {dt:.1f} ms (operator {erep.laplacian_ms:.1f} ms, "
            f"solve {erep.solve_ms:.1f} ms)")
    end_to_end = None
    if e2e:
        end_to_end = {"ms": float(np.median([r["ms"] for r in e2e])),
                      "input": "host CSR arrays in, float32 coordinates out", "runs": e2e}

    dense = None
    if n <= 4096:
        # level 0/5 of the chain: the k smallest pairs of the dense Laplacian
        times_d = []
        for _ in range(3):
            torch.cuda.synchronize()
            t = time.perf_counter()
            dev, dvec = si.dense_evd(lap, k)
            torch.cuda.synchronize()
            times_d.append(1e3 * (time.perf_counter() - t))
        dres = si.max_eigenpair_residual(lap, dev, dvec)
        dense = dict(ms=round(float(np.median(times_d)), 2), max_residual=dres,
                     eigenvalues=[float(v) for v in dev])
        log(f"dense EVD: {dense['ms']} ms residual={dres:.2e}")

    cpu = None
    if not args.no_cpu:
        log("CPU oracle sample ...")
        cpu = cpu_baseline(lap, si, args.n_components, plain_iters, args.cpu_seconds)
        log(f"CPU oracle: {cpu['value']} {cpu['unit']}")
        # the same iteration on the host cores this process may use (OpenMP SpMM)
        try:
            ncores = len(os.sched_getaffinity(0))
        except AttributeError:
            ncores = os.cpu_count() or 1
        ncores = max(1, min(ncores, 16))
        cpu_all = cpu_baseline(lap, si, args.n_components, plain_iters, args.cpu_seconds / 3,
                               threads=ncores)
        log(f"CPU oracle on {ncores} cores: {cpu_all['value']} {cpu_all['unit']}")

    order = {0: "input", 1: "rcm", 2: "cluster"}[si.laplacian_order(lap)]
    traffic = read_traffic(args.traffic, args.workload, last.block_size, order)
    # the unit of the roofline: one application of L to an n x b block, FP32
    # value + int32 column per stored entry, int32 row pointers, the block read
    # once and written once (diagonal folded in the epilogue: stored = nnz(W));
    # the Chebyshev recurrence's epilogue also reads y_{j-1} (a third n x b
    # stream), reported beside it
    b = last.block_size
    unit_bytes = nnz * 8.0 + 4.0 * (n + 1) + 2.0 * 4.0 * n * b
    unit_gbs = unit_bytes / (spmm_avg_us * 1e-6) / 1e9 if sp["launches_f32"] else None
    # the ceiling of the gathers on this graph: the SpMM's 64-byte row gathers
    # alone, along its column stream and at random (one 128-byte line each)
    probe = None
    if b == 16:
        probe = {m: round(si.gather_probe(lap, i, 20), 2)
                 for i, m in enumerate(("column_stream_us", "random_rows_us", "random_rows_2MiB_us"))}
        log(f"gather probe: {probe}")

    # the measured device-copy bandwidth beside the 8 TB/s peak: a 1 GiB
    # device-to-device copy, read + write bytes over the kernel time
    a = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    bcp = torch.empty_like(a)
    bcp.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        bcp.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    copy_gbs = 5 * 2 * a.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, bcp
    log(f"device copy: {copy_gbs:.0f} GB/s (1 GiB, read + write)")

    gbuild = None
    if not args.no_graph_builder and args.workload in ("gmm1m", "gmm100k", "blobs2k"):
        gbuild = bench_graph_builder(args, torch, graphs, si, ctx)

    out = {
        "metric": "ms to eigvecs @ residual<=%g (FP32-bulk Chebyshev block solver)" % args.tol,
        "value": round(fast_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fast_ms, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded Gaussian-mixture exact 15-NN fuzzy graph)",
        "config": {"workload": args.workload, "n": n, "nnz": nnz, "n_components": args.n_components,
                   "block": last.block_size, "cheb_degree": args.degree or "auto", "tol": args.tol,
                   "components_generated": ncomp, "node_order": order},
        "plain_gpu_lobpcg": {"ms": plain_ms, "runs": plain},
        "rsvd_oversampling10_4iter": rsvd or None,
        "dense_evd": dense,
        "speedup_vs_plain_gpu_lobpcg": round(plain_ms / fast_ms, 2) if plain_ms else None,
        "fast_solver": {"rounds": last.iterations, "spmm_calls": last.spmm_calls,
                        "max_residual": last.max_residual, "eigenvalues": last.eigenvalues[:k],
                        "step_ms": [round(t, 3) for t in times]},
        "laplacian_build_ms": round(lap_ms, 2),
        "spectral_init_end_to_end": end_to_end,
        "roofline": {"bound": "hbm", "kernel": "k_spmm_wide<float> (FP32 Chebyshev sweeps)",
                     "achieved": round(unit_gbs, 1) if unit_gbs else None,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(unit_gbs / HBM_PEAK_GBS, 4) if unit_gbs else None,
                     "traffic": traffic,
                     "algorithmic_bytes_per_launch": round(unit_bytes),
                     "bytes_formula": "nnz*(4+4) + 4*(n+1) + 2*4*n*b (b = %d)" % b,
                     "avg_launch_us": round(spmm_avg_us, 2),
                     "launches": sp["launches_f32"],
                     "with_recurrence_stream": {
                         "bytes_per_launch": round(bytes_per_launch),
                         "achieved": round(spmm_gbs, 1) if spmm_gbs else None,
                         "frac": round(spmm_gbs / HBM_PEAK_GBS, 4) if spmm_gbs else None},
                     "gather_probe": probe,
                     "measured_device_copy_gbs": round(copy_gbs, 1)},
        "warm_start": warm,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all if not args.no_cpu else None,
        "graph_builder": gbuild,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
