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
import subprocess
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
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--config", default="c2", choices=("c1", "c2", "c3", "c4a", "c4b"),
                    help="workload of the timed step (c2 is the headline configuration)")
    ap.add_argument("--products", default="auto", choices=("auto", "fp32", "bf16x3"),
                    help="dense-state products of the FP32 step: FP32 matrix cores, or split "
                         "BF16 (hi + lo halves, three products, FP32 accumulation); auto: split "
                         "BF16 for the 20- and 61-state workloads, which specify it")
    ap.add_argument("--dense-samples", type=int, default=1, metavar="K",
                    help="samples K of --config c4a / c4b (1 by default; the served dense requests run K up to 32)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-mix measurement")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-plain-limit", type=float, default=150.0,
                    help="seconds of sequential plain FP64 evaluation (stats over completed requests)")
    ap.add_argument("--serve-streams", type=int, default=8)
    ap.add_argument("--serve-dense-products", default="bf16x3", choices=("fp32", "bf16x3"),
                    help="products of the served 20- and 61-state jobs (the four-state jobs "
                         "have no dense products)")
    ap.add_argument("--serve-graph-bytes", type=int, default=64 << 20,
                    help="requests with at most this much workspace replay a captured HIP graph "
                         "(0: every request launches eagerly)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (a known-size kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.products == "auto":
        args.products = "bf16x3" if args.config in ("c4a", "c4b") else "fp32"
    return args


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def walk_need(post, T):
    """Stack entries of the post-order walk (larger-need child first)."""
    need = {}
    for v, a, b in post:
        na, nb = need.get(a, 0), need.get(b, 0)
        need[v] = max(na, nb, 1 + min(na, nb)) if (na and nb) else max(1, na, nb)
    return need[post[-1][0]]


def pre_walk_used(S, C, rb, N, K):
    """The four-state FP32 pre-order runs as one walk (mirrors use_pre_walk, csrc/layout.hpp)."""
    return S == 4 and rb == 4 and C <= 4


def walk_drops(T, post, dev, S, C, N, K, rb):
    """Cherries whose partial the four-state FP32 walks neither store nor read back (the
    post-order walk drops them, the pre-order walk rebuilds them from two tip codes);
    mirrors walk_drop_mode in csrc/layout.hpp."""
    if not (S == 4 and pre_walk_used(S, C, rb, N, K)):
        return 0
    import ctypes

    from torchtree_amd import _lib
    from torchtree_amd.likelihood import get_plan

    plan = get_plan(T, post, dev.index or 0)
    d0, d1 = ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.load().hx_plan_walk_info(plan.handle, None, None, None, ctypes.byref(d0),
                                             ctypes.byref(d1)))
    return d0.value


def algorithmic_bytes(T, N, C, K, S, rb, nchunks, drops=0):
    """Bytes each pass must move at minimum with this algorithm (one launch set).

    P = K N C S rb is one node's partials. Four states: the post-order walk
    writes every internal non-root partial once, except the ``drops`` cherries
    it does not store, and reads none back from HBM (a child that waited on the
    stack is re-read from L2), plus the tip codes. The four-state FP32 pre-order
    walk reads the same partials once (the adjoints live in registers), reads
    each tip's 16-bit code once per sample and each dropped cherry's code pair,
    and writes the per-chunk branch statistics, which the reduction reads
    back. Otherwise the pre-order runs per level: each internal non-root node's
    L and its parent's U are read and its U written (3 P per node). Other state
    counts use level kernels: post 2 P per node (written, read by the parent)
    plus the exponents.
    """
    part = K * N * C * S * rb
    I = T - 1
    mats = K * (2 * T - 2) * C * S * S * rb
    tips = T * N * 1
    weights = N * 8
    gpart = (2 * T - 2) * K * nchunks * C * S * S * rb
    if S == 4:
        post_b = (I - 1 - drops) * part + tips + mats + weights
    else:
        post_b = 2 * (I - 1) * part + 2 * (I - 1) * K * N * 4 + tips + mats + weights
    if pre_walk_used(S, C, rb, N, K):
        pre_b = (I - 1 - drops) * part + K * (T + drops) * N * 2 + mats + weights + 2 * gpart
    else:
        pre_b = 3 * (I - 1) * part + tips + mats + weights + gpart
    return post_b, pre_b


def canonical_bytes(T, N, C, K, S, rb, tip_bytes=1):
    """Bytes per pass of the level-by-level method, per update: the post-order reads two
    children and writes the parent plus a 4-byte scale per pattern and sample,
    N C K S b [(T-1) + (T-2)] + T N tau + (T-1) N K 4; the pre-order reads the parent's
    adjoint and the sibling's partial, writes the child's adjoint and reads its partial for
    the branch term, (2T-2) N C K 4 S b."""
    part = N * C * K * S * rb
    post = part * ((T - 1) + (T - 2)) + T * N * tip_bytes + (T - 1) * N * K * 4
    pre = (2 * T - 2) * part * 4
    return {"postorder": post, "preorder": pre}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "across GPUs the path would shard patterns, one slice and one all-reduce per step",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from torchtree_amd import models as hm
    from torchtree_amd import synthetic, timing
    from torchtree_amd.engine import TreeJob
    from torchtree_amd.likelihood import TipData

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration} bytes")
        del src, dst
    t0 = time.time()
    wl = make_config(args, synthetic)
    log(f"workload {wl.name}: T={wl.T} N={wl.N} C={wl.C} K={wl.K} S={wl.S} built in {time.time()-t0:.1f}s")
    post = [tuple(int(x) for x in r) for r in wl.topo.postorder]
    S = wl.S
    tips = TipData.from_states(list(wl.codes), S, dev)
    rates = wl.rates.reshape(1, -1).to(dev)
    props = wl.props.reshape(1, -1).to(dev)

    q_builder = {4: hm.gtr_q, 20: hm.q_from_exchangeabilities}.get(S)
    if S == 61:
        flags = hm.mg94_flags()
        one = torch.ones(1, dtype=torch.float64)

        def q_builder(ex, fr):  # exch holds (alpha, beta=omega, kappa)
            return hm.mg94_q(ex[0:1], ex[1:2], ex[2:3], fr, flags)

    def make_step(job):
        bls = wl.bls.to(dev).clone().requires_grad_(True)
        exch = wl.exch.clone().requires_grad_(True)
        freqs = wl.freqs.clone().requires_grad_(True)

        def step():
            bls.grad = None
            exch.grad = None
            freqs.grad = None
            qn = hm.normalize(q_builder(exch, freqs), freqs)
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
    job32 = TreeJob(post, tips, wl.weights, S, dtype=torch.float32, device=dev,
                    products=args.products)
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
    drops = walk_drops(wl.T, post, dev, S, wl.C, wl.N, wl.K, 4)
    post_b, pre_b = algorithmic_bytes(wl.T, wl.N, wl.C, wl.K, S, 4, nchunks, drops)
    canon_b = canonical_bytes(wl.T, wl.N, wl.C, wl.K, S, 4)
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
            "workload": DESCRIPTIONS[args.config].format(T=wl.T, N=wl.N, K=wl.K),
            "taxa": wl.T, "patterns": wl.N, "categories": wl.C, "samples": wl.K, "states": S,
            "partials": "fp32 + power-of-two rescaling", "seed": args.seed,
            "dense_products": args.products if S != 4 else "fp32",
        },
        "p50_ms": statistics.median(lat),
        "p99_ms": _pct(lat, 99),
        "updates_per_step": upd,
        "phases_ms": {k: v["ms"] for k, v in phases.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": (("k_pre_walk (one launch, whole tree) + k_reduce_g"
                        if pre_walk_used(S, wl.C, 4, wl.N, wl.K) else
                        f"k_pre ({launches - 1} depth-level launches) + k_reduce_g")
                       if dom == "preorder" else "k_post_walk (one launch, whole tree) + k_reduce_root"),
            "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dom_gbs / HBM_PEAK_GBS,
            "traffic": measured_traffic(args, dom), "algorithmic_bytes": dom_b, "pass_ms": dom_ms,
            # the measured HBM bytes over the pass time, and the per-update bytes of the
            # level-by-level method (every adjoint written and read back) over the same time
            "traffic_frac": (None if measured_traffic(args, dom) is None else
                             measured_traffic(args, dom) / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS),
            "canonical_bytes": canon_b[dom],
            "canonical_frac": canon_b[dom] / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            # the walks are one launch for the whole tree (pass time includes the
            # small reduction that follows); level kernels one launch per level
            "avg_launch_us": dom_ms * 1e3 / (1 if (dom == "postorder" or pre_walk_used(S, wl.C, 4, wl.N, wl.K))
                                             else max(1, launches - 1)),
        },
        "other_pass": {"postorder_gbs": post_gbs, "preorder_gbs": pre_gbs,
                       "postorder_bytes": post_b, "preorder_bytes": pre_b,
                       "walk_stack_entries": walk_need(post, wl.T),
                       "cherries_not_stored": drops},
    }

    # ---- FP64 plain path: accuracy reference at full size, and its throughput
    if not args.no_check:
        job64 = TreeJob(post, tips, wl.weights, S, dtype=torch.float64, device=dev)
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

    if not args.no_cpu_baseline and args.config == "c2":
        result["cpu_baseline"] = cpu_baseline(args, dev)
    if not args.no_serve and args.config == "c2":
        result["serving"] = serving_bench(args, dev)
    print(json.dumps(result), flush=True)


DESCRIPTIONS = {
    "c1": "plain JC69 loglik+grad: random {T}-taxon tree, {N} patterns from 1,000 simulated sites, K={K}",
    "c2": "batched DNA VI loglik+grad: GTR+Weibull4(0.5), random {T}-taxon tree, {N} unique patterns, K={K} draws",
    "c3": "adjoint DNA loglik+grad: GTR+Weibull4(0.5), random {T}-taxon tree, {N} patterns, K={K}",
    "c4a": "amino acid loglik+grad: 20 states, Gamma(1,1) exchangeabilities, gamma4, {T} taxa, {N} patterns, K={K}",
    "c4b": "codon loglik+grad: MG94 kappa=2 omega=0.3, gamma4, {T} taxa, {N} patterns, K={K}",
}


def make_config(args, synthetic):
    if args.config == "c1":
        return synthetic.config1()
    if args.config == "c2":
        return synthetic.config2(seed=args.seed, N=args.patterns, K=args.samples, T=args.taxa)
    if args.config == "c3":
        return synthetic.config3()
    K = args.dense_samples
    if args.config == "c4a":
        return synthetic.make_workload(256, 20_000, K, model="aa", site="gamma", shape=0.5,
                                       bl_mean=0.05, seed=41, name="config4a-aa-256taxa-20000pat")
    return synthetic.make_workload(128, 10_000, K, model="mg94", site="gamma", shape=0.5,
                                   bl_mean=0.05, seed=42, name="config4b-mg94-128taxa-10000pat")


def serving_bench(args, dev):
    """Serving mix: resident jobs, Poisson loglik+grad requests, FP32 served vs plain FP64."""
    import torch

    from torchtree_amd import serving, synthetic

    t0 = time.time()
    specs = synthetic.serving_mix(args.serve_jobs, seed=5)
    log(f"serving mix: {len(specs)} jobs generated in {time.time()-t0:.1f}s")
    budget = 200 << 30
    srv = serving.Server(budget, n_streams=args.serve_streams, pool_bytes=160 << 30,
                         max_request_bytes=40 << 30, graph_max_bytes=args.serve_graph_bytes,
                         device=dev.index or 0)
    jobs32, jobs64 = [], []
    for sp in specs:
        post = [tuple(int(x) for x in r) for r in sp["topo"].postorder]
        T, N, S, C = sp["topo"].T, sp["codes"].shape[1], sp["S"], len(sp["rates"])
        common = (post, sp["codes"], sp["table"], sp["weights"], sp["q_norm"], sp["freqs"],
                  sp["rates"], sp["props"])
        j32 = serving.ResidentJob(sp["name"], *common, dtype=torch.float32,
                                  slice_patterns=srv.slice_for(T, N, C, S, torch.float32),
                                  device=dev.index or 0, kind=sp["kind"],
                                  products=args.serve_dense_products if S != 4 else "fp32")
        srv.register(sp["name"], j32)
        jobs32.append(j32)
        jobs64.append(serving.ResidentJob(sp["name"], *common, dtype=torch.float64,
                                          slice_patterns=srv.slice_for(T, N, C, S, torch.float64),
                                          device=dev.index or 0, kind=sp["kind"]))
    log(f"registered {len(jobs32)} resident jobs ({srv.resident / 2**30:.2f} GiB FP32 tips) in "
        f"{time.time()-t0:.1f}s")

    def requests(jobs, stream):
        return [serving.Request(i, jobs[j], synthetic.draw_bls(specs[j], K, sd).pin_memory(),
                                arrival=t) for i, (j, K, sd, t) in enumerate(stream)]

    n = args.serve_requests
    # capacity: all requests queued at t=0
    burst = synthetic.request_stream(n, len(specs), rate=1e9, seed=7)
    warm = requests(jobs32, burst[: min(20, n)])
    srv.run(warm, realtime=False)
    prog = lambda d, t: log(f"  ... {d}/{t} requests done")
    cap = srv.run(requests(jobs32, burst), realtime=False, progress=prog)
    mean_upd = cap["updates"] / cap["requests"]
    rate = 0.8 * cap["updates_per_s"] / mean_upd
    log(f"serving capacity {cap['updates_per_s']:.3e} updates/s; offered rate {rate:.1f} req/s")
    stream = synthetic.request_stream(n, len(specs), rate=rate, seed=8)
    served = requests(jobs32, stream)
    st = srv.run(served, progress=prog)
    log(f"served: {st}")
    plain = requests(jobs64, stream)
    serving.run_sequential(plain[: min(n, 5)], device=dev.index or 0)  # warm-up
    # the plain path under the same arrivals, one request at a time, first come first served:
    # latencies include its queueing exactly as the server's do
    plain = requests(jobs64, stream)
    pl = serving.run_sequential(plain, device=dev.index or 0, time_limit=args.serve_plain_limit,
                                progress=prog, realtime=True)
    log(f"plain fp64 sequential, same arrivals: {pl}")
    errs, gerrs = [], []
    for a, b in zip(served, plain):
        if b.loglik is None or math.isnan(b.t_done):
            continue
        errs.append(float(((a.loglik - b.loglik).abs() / b.loglik.abs()).max()))
        gerrs.append(float((a.dbls - b.dbls).norm() / b.dbls.norm()))
    kinds = [r.job.kind for r in served]
    return {
        "workload": f"{len(specs)} resident jobs (T~U{{16..1024}}, N log-uniform 100-100000, "
                    f"DNA/AA/codon {kinds.count('dna')}/{kinds.count('aa')}/{kinds.count('codon')} "
                    f"requests), K in {{1,8,32}}, Poisson arrivals at {rate:.1f}/s (80% of capacity), "
                    f"{args.serve_streams} streams, {srv.pool.size >> 30} GiB workspace pool, "
                    f"{args.serve_dense_products} products for the 20- and 61-state jobs",
        "requests": st["requests"],
        "served_updates_per_s": st["updates_per_s"],
        "served_p50_ms": st["p50_ms"], "served_p99_ms": st["p99_ms"],
        "served_counters": {k: st.get(k, 0) for k in ("admitted", "bypassed", "rejected",
                                                     "cancelled", "queued_peak", "graphed",
                                                     "pool_peak_bytes")},
        "plain_same_arrivals": True,
        "capacity_updates_per_s": cap["updates_per_s"],
        "plain_fp64_sequential_updates_per_s": pl["updates_per_s"],
        "plain_fp64_requests_timed": pl["requests"],
        "plain_fp64_p50_ms": pl["p50_ms"], "plain_fp64_p99_ms": pl["p99_ms"],
        "served_vs_plain": st["updates_per_s"] / pl["updates_per_s"],
        "capacity_vs_plain": cap["updates_per_s"] / pl["updates_per_s"],
        "max_logl_rel_err_vs_fp64": max(errs), "max_grad_rel_err_vs_fp64": max(gerrs),
    }


def measured_traffic(args, dom):
    """HBM bytes of the dominant pass per step, from the committed FETCH_SIZE /
    WRITE_SIZE counter runs of this exact workload (rocprofv3, scaled by a
    1 GiB copy; profiles/hbm_traffic_config2*.md), or None for other sizes."""
    if (args.config, args.taxa, args.patterns, args.samples) != ("c2", 512, 100_000, 32):
        return None
    path = Path(__file__).resolve().parent / "profiles" / "hbm_traffic_config2.json"
    try:
        d = json.loads(path.read_text())[dom]
    except (OSError, KeyError, ValueError):
        return None
    return d["read_bytes"] + d["write_bytes"]


def _nchunks(N, C, K, rb=4):
    """Chunk count of the S=4 kernels (mirrors make_tiling in csrc/layout.hpp)."""
    ppb = 64
    ntiles = (N + ppb - 1) // ppb
    if pre_walk_used(4, C, rb, N, K):
        return (ntiles + 3) // 4  # four tiles per pre-order walk workgroup
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

    # every core this process may run on, within the CPU share the host grants it (the
    # OMP_NUM_THREADS it exports, 16 per GPU on the MI355X hosts)
    affinity = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", affinity) or affinity)
    cores = max(1, min(affinity, share))
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
    model = "unknown"
    try:  # lscpu's model name, else /proc/cpuinfo's
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        names = [l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")]
        if not names:
            names = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo")
                     if l.startswith("model name")]
        model = names[0] if names else model
    except Exception:
        pass
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
