#!/usr/bin/env python3
"""Train the learned warm-start network on the GPU and evaluate what it gains.

Training data are the solver's own converged eigenvectors: seeded Gaussian-
mixture fuzzy 15-NN graphs (20 clusters in 50-d, Dirichlet(2) sizes; the
construction of graphs.workload("gmm100k")) with node counts drawn
log-uniformly from [--n-lo, --n-hi].  Each graph contributes node features for
three probe seeds and the two non-trivial eigenvectors of its Laplacian
(FP32-bulk solver, residual <= 1e-6).  Training minimises the subspace-coverage
loss with the network's own forward / backward kernels and the device Adam
step (spectral_mi355x.WarmStartNet.train_step), cycling through the probe
seeds.

Evaluation on held-out seeds compares, per graph, the plain FP64 LOBPCG and
the FP32-bulk solver started from N(0,1) draws with the same solvers started
from the warm-start block (the Rayleigh-Ritz block of the network's outputs
when its residual beats the random start's, else the random start).  With
--train 0 the script only evaluates the parameters given by --init.

    python spectral-init_amd/train_warmstart.py --out spectral-init_amd/warmstart_mpnn.npy
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)


def log(msg):
    print(f"[train {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def make_graph(graphs, seed, n):
    w = graphs.workload("gmm100k", seed=seed, n=n, device="cuda")
    ip, ix, wt, _ = graphs.largest_connected(w["indptr"], w["indices"], w["data"])
    return ip, ix, wt


def solve(si, lap, cfg):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    ev, vec, rep = si.solve_eigenproblem(lap, 2, 42, cfg)
    torch.cuda.synchronize()
    return vec, dict(ms=1e3 * (time.perf_counter() - t), iterations=rep.iterations,
                     spmm_calls=rep.spmm_calls, residual=rep.max_residual, level=rep.level)


def evaluate(si, graphs, net, seeds, n, tol, on_row=None):
    import torch
    rows = []
    for s in seeds:
        ip, ix, wt = make_graph(graphs, s, n)
        lap = si.Laplacian(net.ctx, ip, ix, wt)
        row = dict(seed=s, size=int(n), n=int(lap.n), nnz=int(lap.nnz))
        for name, solver in (("plain", si.Solver.PlainLobpcg), ("fast", si.Solver.Fast)):
            cfg = si.SpectralInitConfig(tol=tol, solver=solver)
            solve(si, lap, cfg)  # warm the solver's buffers
            _, r = solve(si, lap, cfg)
            row[f"{name}_random"] = r
            torch.cuda.synchronize()
            t = time.perf_counter()
            # same seed as the solver's own random start: the fallback block is
            # that start, so the comparison isolates the network
            blk, used, rn, rr = net.warm_start(lap, seed=42)
            torch.cuda.synchronize()
            ws_ms = 1e3 * (time.perf_counter() - t)
            cfg_w = si.SpectralInitConfig(tol=tol, solver=solver, warm_start=blk.data_ptr())
            _, r = solve(si, lap, cfg_w)
            r["warm_start_ms"] = ws_ms
            r["total_ms"] = ws_ms + r["ms"]
            r["used_net"] = used
            r["res_net_block"] = rn
            r["res_random_block"] = rr
            row[f"{name}_warm"] = r
            del blk
        rows.append(row)
        log(f"eval seed {s} n={row['n']}: plain {row['plain_random']['iterations']} -> "
            f"{row['plain_warm']['iterations']} it ({row['plain_random']['ms']:.1f} -> "
            f"{row['plain_warm']['total_ms']:.1f} ms), fast {row['fast_random']['ms']:.1f} -> "
            f"{row['fast_warm']['total_ms']:.1f} ms, net used {row['plain_warm']['used_net']} "
            f"(block residual {row['plain_warm']['res_net_block']:.3e} vs "
            f"{row['plain_warm']['res_random_block']:.3e})")
        del lap
        if on_row:
            on_row(row)
    return rows


def summarise(rows):
    """Per evaluation size: sums over graphs (divide by "graphs" for means)."""
    summary = {}
    for r in rows:
        d = summary.setdefault(r["size"], dict(graphs=0, plain_it_random=0, plain_it_warm=0,
                                               plain_ms_random=0.0, plain_ms_warm=0.0,
                                               net_used=0, plain_fewer_iterations=0,
                                               plain_max_residual_warm=0.0,
                                               fast_ms_random=0.0, fast_ms_warm=0.0))
        d["graphs"] += 1
        d["plain_it_random"] += r["plain_random"]["iterations"]
        d["plain_it_warm"] += r["plain_warm"]["iterations"]
        d["plain_ms_random"] += r["plain_random"]["ms"]
        d["plain_ms_warm"] += r["plain_warm"]["total_ms"]
        d["net_used"] += int(r["plain_warm"]["used_net"])
        d["plain_fewer_iterations"] += int(r["plain_warm"]["iterations"] <
                                           r["plain_random"]["iterations"])
        d["plain_max_residual_warm"] = max(d["plain_max_residual_warm"],
                                           r["plain_warm"]["residual"])
        d["fast_ms_random"] += r["fast_random"]["ms"]
        d["fast_ms_warm"] += r["fast_warm"]["total_ms"]
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=400, help="training graphs (0: evaluate --init)")
    ap.add_argument("--n-lo", type=int, default=50000)
    ap.add_argument("--n-hi", type=int, default=200000)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--probe-seeds", type=int, default=3, help="feature probe seeds per graph")
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed0", type=int, default=1000, help="first training graph seed")
    ap.add_argument("--eval", default="50000:4,200000:4",
                    help="held-out evaluation graphs as size:count[,size:count...]")
    ap.add_argument("--eval-seed0", type=int, default=5000)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--init", default=None, help="start from these parameters (.npy)")
    ap.add_argument("--out", default=os.path.join(HERE, "warmstart_mpnn.npy"))
    ap.add_argument("--report", default=None, help="write the training / evaluation JSON here")
    args = ap.parse_args()

    import torch
    import graphs
    import spectral_mi355x as si

    ctx = si.Context(0)
    net = si.WarmStartNet(ctx, seed=7, params=np.load(args.init) if args.init else None)
    rng = np.random.default_rng(args.seed0)
    sizes = np.exp(rng.uniform(math.log(args.n_lo), math.log(args.n_hi), args.train)).astype(int)
    data = []
    t0 = time.perf_counter()
    for i, n in enumerate(sizes if args.train > 0 else []):
        ip, ix, wt = make_graph(graphs, args.seed0 + i, int(n))
        lap = si.Laplacian(ctx, ip, ix, wt)
        fs = [si.WarmStartNet.features(lap, seed=p) for p in range(args.probe_seeds)]
        vec, r = solve(si, lap, si.SpectralInitConfig(tol=args.tol))
        if not r["residual"] <= args.tol:
            log(f"graph {i}: solver residual {r['residual']:.2e}, skipped")
            continue
        data.append((lap, fs, vec[:, 1:3].contiguous()))
        if i % 16 == 0:
            log(f"graph {i}: n={lap.n} nnz={lap.nnz} ({time.perf_counter() - t0:.0f} s)")
    log(f"{len(data)} training graphs ready in {time.perf_counter() - t0:.0f} s")

    steps = args.epochs * len(data)
    hist = []
    step = 0
    t0 = time.perf_counter()
    for ep in range(args.epochs if data else 0):
        order = rng.permutation(len(data))
        losses = []
        for j in order:
            lap, fs, v = data[j]
            lr = args.lr * 0.5 * (1.0 + math.cos(math.pi * step / steps))
            losses.append(net.train_step(lap, fs[(step + j) % len(fs)], v, lr))
            step += 1
        hist.append(float(np.mean(losses)))
        log(f"epoch {ep}: mean loss {hist[-1]:.4f}")
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    if data:
        np.save(args.out, net.params())
        log(f"trained {steps} steps in {train_s:.1f} s -> {args.out}")
    del data

    rows = []

    def write_report():
        summary = summarise(rows)
        rep = dict(train_graphs=int(args.train), n_range=[args.n_lo, args.n_hi],
                   epochs=args.epochs, lr=args.lr, probe_seeds=args.probe_seeds, steps=steps,
                   train_seconds=train_s, loss_history=hist, eval=rows,
                   summary={str(k): v for k, v in summary.items()},
                   command="python -u spectral-init_amd/train_warmstart.py " + " ".join(sys.argv[1:]))
        if args.report:  # rewritten after every graph, so a cut-off run keeps its rows
            with open(args.report, "w") as fh:
                json.dump(rep, fh, indent=1)
        return summary

    seed = args.eval_seed0
    for spec in args.eval.split(","):
        size, count = (int(v) for v in spec.split(":"))
        evaluate(si, graphs, net, list(range(seed, seed + count)), size, args.tol,
                 on_row=lambda r: (rows.append(r), write_report()))
        seed += count
    summary = write_report()
    for size, d in sorted(summary.items()):
        g = d["graphs"]
        log(f"n~{size}: {g} graphs, plain LOBPCG mean iterations {d['plain_it_random'] / g:.1f} -> "
            f"{d['plain_it_warm'] / g:.1f}, mean ms {d['plain_ms_random'] / g:.1f} -> "
            f"{d['plain_ms_warm'] / g:.1f} (incl. warm start), network block used {d['net_used']}/{g}; "
            f"fewer iterations on {d['plain_fewer_iterations']}/{g}; FP32-bulk solver mean ms "
            f"{d['fast_ms_random'] / g:.2f} -> {d['fast_ms_warm'] / g:.2f} (incl. warm start)")
    print(json.dumps({"loss_first": hist[0] if hist else None, "loss_last": hist[-1] if hist else None,
                      "eval": [{k: (v if not isinstance(v, dict) else
                                    {kk: v[kk] for kk in ("iterations", "ms", "residual")})
                                for k, v in r.items()} for r in rows]}), flush=True)


if __name__ == "__main__":
    main()
