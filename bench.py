#!/usr/bin/env python
"""Tree-likelihood loglik + gradient benchmark on one MI355X.

Workload (default): batched DNA variational inference, GTR with
Dirichlet(10,10,10,10) frequencies and Gamma(2, rate 0.5) exchangeabilities,
4 Weibull(0.5) rate categories, seeded random 512-taxon tree with branch
lengths ~ Exp(mean 0.02), 100,000 simulated unique patterns with
Poisson(3)+1 weights, K=32 log-normal (sigma 0.1) branch-length draws per
step. A step is one loglik + gradient (branch lengths, GTR exchangeabilities,
frequencies) of all K draws: FP32 partials with power-of-two rescaling, FP64
eigen / exponentials / final sums. Inputs are resident in HBM before timing.

Unit: one site-partial update is one S-vector of partials for one
(node, pattern, category, sample); a loglik+grad step costs
(T-1)*N*C*K post-order + (2T-2)*N*C*K pre-order updates.

Prints one JSON line. ``--gpus`` above 1 exits: this build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "site-partial updates/s + loglik+grad p50/p99 latency, served vs plain GPU"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--taxa", type=int, default=512)
    ap.add_argument("--patterns", type=int, default=100_000)
    ap.add_argument("--samples", type=int, default=32)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 accuracy pass")
    ap.add_argument("--no-fp64-timing", action="store_true", help="skip timing the FP64 plain path")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-patterns", type=int, default=10000)
    ap.add_argument("--cpu-samples", type=int, default=2)
    ap.add_argument("--cpu-reps", type=int, default=3)
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(T, N, C, K, S, rb, nchunks):
    """Bytes each pass must move at minimum with this algorithm (one launch set)."""
    part = K * N * C * S * rb
    scale = K * N * 4
    mats = K * (2 * T - 2) * C * S * S * rb
    tips = T * N * 1
    weights = N * 8
    post = 2 * (T - 2) * part + 2 * (T - 2) * scale + tips + mats + weights
    gpart = (2 * T - 2) * K * nchunks * C * S * S * rb
    pre = 3 * (T - 2) * part + tips + mats + weights + gpart
    return post, pre


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the pattern-sharded multi-GPU design is described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    import torch

    from torchtree_amd import models as hm
    from torchtree_amd import synthetic, timing
    from torchtree_amd.engine import TreeJob
    from torchtree_amd.likelihood import TipData

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t0 = time.time()
    wl = synthetic.config2(seed=args.seed, N=args.patterns, K=args.samples, T=args.taxa)
    log(f"workload {wl.name}: T={wl.T} N={wl.N} C={wl.C} K={wl.K} built in {time.time()-t0:.1f}s")
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    tips = TipData.from_states(list(wl.codes), 4, dev)
    rates = wl.rates.reshape(1, -1).to(dev)
    props = wl.props.reshape(1, -1).to(dev)

    def make_step(job):
        bls = wl.bls.to(dev).clone().requires_grad_(True)
        exch = wl.exch.clone().requires_grad_(True)
        freqs = wl.freqs.clone().requires_grad_(True)

        def step():
            bls.grad = None
            exch.grad = None
            freqs.grad = None
            qn = hm.normalize(hm.gtr_q(exch, freqs), freqs)
            ll = job.loglik(bls, rates, props, qn, freqs)
            ll.sum().backward()
            return ll, bls, exch, freqs

        return step

    def timed(step, steps, warmup):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        lat = []
        start = time.perf_counter()
        for _ in range(steps):
            a = time.perf_counter()
            step()
            torch.cuda.synchronize()
            lat.append((time.perf_counter() - a) * 1e3)
        total = time.perf_counter() - start
        return total, lat

    # ---- FP32 served-precision path (the measured value)
    job32 = TreeJob(post, tips, wl.weights, 4, dtype=torch.float32, device=dev)
    step32 = make_step(job32)
    total, lat = timed(step32, args.steps, args.warmup)
    ms = total / args.steps * 1e3
    upd = wl.updates(grad=True)
    value = upd / (total / args.steps)
    log(f"fp32: {ms:.2f} ms/step, {value:.3e} updates/s")
    with timing.record() as rec:
        ll32, b32, e32, f32 = step32()
    phases = rec.summary()
    log(f"phases: {json.dumps(phases)}")
    ll32 = ll32.detach().cpu()
    g32 = (b32.grad.detach().cpu(), e32.grad.detach().cpu(), f32.grad.detach().cpu())
    nchunks = _nchunks(wl.N, wl.C, wl.K)
    post_b, pre_b = algorithmic_bytes(wl.T, wl.N, wl.C, wl.K, 4, 4, nchunks)
    post_ms = phases["postorder"]["ms"]
    pre_ms = phases["preorder"]["ms"]
    post_gbs = post_b / (post_ms * 1e-3) / 1e9
    pre_gbs = pre_b / (pre_ms * 1e-3) / 1e9
    dom = "preorder" if pre_ms >= post_ms else "postorder"
    dom_b, dom_ms, dom_gbs = (pre_b, pre_ms, pre_gbs) if dom == "preorder" else (post_b, post_ms, post_gbs)
    launches = phases[dom]["launches"]
    del job32
    torch.cuda.empty_cache()

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "site-partial updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched DNA VI loglik+grad: GTR+Weibull4(0.5), random "
                        f"{wl.T}-taxon tree, {wl.N} unique patterns, K={wl.K} draws",
            "taxa": wl.T, "patterns": wl.N, "categories": wl.C, "samples": wl.K, "states": 4,
            "partials": "fp32 + power-of-two rescaling", "seed": args.seed,
        },
        "p50_ms": statistics.median(lat),
        "p99_ms": _pct(lat, 99),
        "updates_per_step": upd,
        "phases_ms": {k: v["ms"] for k, v in phases.items()},
        "roofline": {
            "bound": "hbm", "kernel": f"k_pre (all {launches} depth-level launches)" if dom == "preorder"
            else f"k_post (all {launches} height-level launches)",
            "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dom_gbs / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes": dom_b, "pass_ms": dom_ms,
            "avg_launch_us": dom_ms * 1e3 / max(1, launches - 1),
        },
        "other_pass": {"postorder_gbs": post_gbs, "preorder_gbs": pre_gbs,
                       "postorder_bytes": post_b, "preorder_bytes": pre_b},
    }

    # ---- FP64 plain path: accuracy reference at full size, and its throughput
    if not args.no_check:
        job64 = TreeJob(post, tips, wl.weights, 4, dtype=torch.float64, device=dev)
        step64 = make_step(job64)
        if not args.no_fp64_timing:
            tot64, lat64 = timed(step64, max(2, args.steps // 3), 1)
            result["plain_fp64_updates_per_s"] = upd / (tot64 / max(2, args.steps // 3))
            result["plain_fp64_ms_per_step"] = tot64 / max(2, args.steps // 3) * 1e3
        ll64, b64, e64, f64 = step64()
        ll64 = ll64.detach().cpu()
        g64 = (b64.grad.detach().cpu(), e64.grad.detach().cpu(), f64.grad.detach().cpu())
        result["logl_rel_err_vs_fp64"] = float(((ll32 - ll64).abs() / ll64.abs()).max())
        result["grad_rel_err_vs_fp64"] = {
            n: float((a - b).norm() / b.norm()) for n, a, b in zip(("bls", "exch", "freqs"), g32, g64)}
        del job64
        torch.cuda.empty_cache()
        log(f"fp32 vs fp64: {result['logl_rel_err_vs_fp64']:.2e} {result['grad_rel_err_vs_fp64']}")

    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args, dev)
    print(json.dumps(result), flush=True)


def _nchunks(N, C, K):
    """Chunk count of the S=4 kernels (mirrors make_tiling in csrc/common.hpp)."""
    cp = 1
    while cp < C:
        cp *= 2
    ppb = 256 // cp
    ntiles = (N + ppb - 1) // ppb
    nch = max(1, min(ntiles, 128))
    tpc = (ntiles + nch - 1) // nch
    return (ntiles + tpc - 1) // tpc


def _pct(xs, p):
    s = sorted(xs)
    i = min(len(s) - 1, max(0, int(math.ceil(p / 100 * len(s))) - 1))
    return s[i]


def cpu_baseline(args, dev):
    """FP64 oracle (reference method + autograd) on the host cores, bounded sample."""
    import torch

    from oracle import models as omodels
    from oracle import pruning as opruning
    from torchtree_amd import models as hm
    from torchtree_amd import synthetic
    from torchtree_amd.engine import TreeJob
    from torchtree_amd.likelihood import TipData

    cores = min(16, len(os.sched_getaffinity(0)))
    torch.set_num_threads(cores)
    wl = synthetic.config2(seed=args.seed, N=args.cpu_patterns, K=args.cpu_samples, T=args.taxa)
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    tips = wl.tip_partials() + [None] * (wl.T - 1)

    def run():
        bls = wl.bls.clone().requires_grad_(True)
        exch = wl.exch.clone().requires_grad_(True)
        freqs = wl.freqs.clone().requires_grad_(True)
        q = omodels.gtr_q(exch, freqs)
        mats = omodels.p_t_eigh(q, freqs, bls.unsqueeze(-1) * wl.rates.reshape(1, 1, -1))
        ll = opruning.treelik_discrete_rescaled(tips, wl.weights, post, mats, freqs.reshape(1, -1),
                                                wl.props.reshape(-1, 1, 1))
        ll.sum().backward()
        return ll.detach().reshape(-1), bls.grad

    run()
    times = []
    for _ in range(args.cpu_reps):
        a = time.perf_counter()
        ref_ll, ref_g = run()
        times.append(time.perf_counter() - a)
    med = statistics.median(times)
    upd = wl.updates(grad=True)
    # the same slice through the FP32 HIP path, checked against this oracle result
    job = TreeJob(post, TipData.from_states(list(wl.codes), 4, dev), wl.weights, 4,
                  dtype=torch.float32, device=dev)
    b = wl.bls.to(dev).clone().requires_grad_(True)
    qn = hm.normalize(hm.gtr_q(wl.exch, wl.freqs), wl.freqs)
    got = job.loglik(b, wl.rates.reshape(1, -1), wl.props.reshape(1, -1), qn, wl.freqs)
    got.sum().backward()
    rel = float(((got.detach().cpu() - ref_ll).abs() / ref_ll.abs()).max())
    grel = float((b.grad.cpu() - ref_g).norm() / ref_g.norm())
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    return {
        "value": upd / med, "unit": "site-partial updates/s", "cores": cores, "kind": "port",
        "sample": f"FP64 torch oracle loglik+autograd grad, T={wl.T}, N={wl.N} patterns, "
                  f"C={wl.C}, K={wl.K}; median of {args.cpu_reps} after 1 warm-up "
                  f"({med:.2f} s/eval)",
        "cpu_model": model,
        "fp32_gpu_vs_oracle_on_sample": {"logl_rel_err": rel, "dbls_rel_err": grel},
    }


if __name__ == "__main__":
    main()
