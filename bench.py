"""Benchmark of the multilevel k-way partitioner on one MI355X.

Workload (one "step" = one full partition, CSR resident in HBM -> part[] in
HBM): R-MAT graph with 2^22 vertices and 2^26 unique undirected edges
(a,b,c,d = 0.57,0.19,0.19,0.05; self-loops and duplicates removed; seeded
relabelling), int64 xadj, int32 adjncy, unit weights, k = 256, balance bound
1.03, seed 1.

Prints one JSON line: throughput in M edges/s (undirected |E| / time per
step), the HBM roofline of the dominant kernel (per-launch algorithmic bytes /
average launch duration from HIP events on the library's stream), and the
host-CPU baseline: the C restatement of the reference CPU path
(hunyuangraph_rbbisection) timed on one core on a bounded sample of the same
workload (R-MAT scale 19, same parameters and edge factor, k = 256).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "M edges/s partitioned at cut≤1.03×ref, imb≤1.03; batched graphs/s + p99 ms"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cut_and_maxpart(xadj, adjncy, part, k):
    n = len(xadj) - 1
    deg = np.diff(xadj)
    rows = np.repeat(np.arange(n, dtype=np.int32), deg)
    cut = int(np.count_nonzero(part[rows] != part[adjncy])) // 2
    pw = np.bincount(part, minlength=k)
    return cut, int(pw.max())


def cpu_baseline(scale, k, seed, hgp):
    """Reference CPU path restated in C (oracle), single thread, bounded sample."""
    import oracle as O
    x, a, vw, aw = hgp.gen_rmat(scale, 16 << scale, seed=seed)
    E = len(a) // 2
    t0 = time.perf_counter()
    part, cut = O.cpu_rb_partition(x, a, k, vw, aw, seed=seed)
    dt = time.perf_counter() - t0
    return {
        "value": round(E / dt / 1e6, 4),
        "unit": "M edges/s",
        "cores": 1,
        "kind": "port",
        "sample": f"rmat scale {scale} ({E} edges), k={k}, one partition in {dt:.2f} s",
        "cut": int(cut),
        "cut_fraction": round(cut / E, 4),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edges-log2", type=int, default=26)
    ap.add_argument("--k", type=int, default=256)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-scale", type=int, default=19)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ktimer", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; use --gpus 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    log("bench: loading library and torch")
    import hunyuansc25_amd as hgp
    import torch
    log("bench: torch ready")

    if hgp.device_count() < 1:
        raise SystemExit("no HIP device")
    t0 = time.perf_counter()
    x, a, vw, aw = hgp.gen_rmat(args.scale, 1 << args.edges_log2, seed=args.seed)
    n, m = len(x) - 1, len(a)
    E = m // 2
    log(f"generated rmat scale {args.scale}: n={n} |E|={E} in {time.perf_counter() - t0:.1f} s")

    dev = torch.device("cuda:0")
    d_x = torch.from_numpy(x).to(dev)
    d_a = torch.from_numpy(a).to(dev)
    d_vw = torch.from_numpy(vw).to(dev)
    d_aw = torch.from_numpy(aw).to(dev)
    d_part = torch.zeros(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def step():
        return hgp.partition_device(n, d_x.data_ptr(), d_a.data_ptr(), d_vw.data_ptr(), d_aw.data_ptr(), args.k,
                                    d_part.data_ptr(), seed=args.seed)

    for i in range(args.warmup):
        st = step()
        log(f"warmup {i}: {st['ms_total']:.1f} ms device, cut {st['cut']}")

    if not args.no_ktimer:
        hgp.ktimer_enable(True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    stats = []
    for i in range(args.steps):
        stats.append(step())
        log(f"step {i}: {stats[-1]['ms_total']:.1f} ms device")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    kt = hgp.ktimer_report() if not args.no_ktimer else {}
    hgp.ktimer_enable(False)
    ms_per_step = elapsed / args.steps * 1e3
    value = E / (elapsed / args.steps) / 1e6

    part = d_part.cpu().numpy()
    cut, maxp = host_cut_and_maxpart(x, a, part, args.k)
    st = stats[-1]
    tvwgt = int(vw.sum())
    balanced = 100 * args.k * maxp <= 103 * tvwgt
    if cut != st["cut"]:
        raise SystemExit(f"library cut {st['cut']} != host recount {cut}")

    roofline = None
    if kt:
        dom = max(kt.items(), key=lambda kv: kv[1][1])
        name, (launches, ms, byts) = dom
        avg_ms = ms / launches
        achieved = (byts / launches) / (avg_ms * 1e-3) / 1e9
        roofline = {
            "bound": "hbm",
            "kernel": name,
            "launches": int(launches),
            "avg_ms": round(avg_ms, 5),
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
        }
    pipeline_gbs = st["alg_bytes"] / (st["ms_total"] * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: rmat scale {args.cpu_scale} k={args.k} ...")
        cpu = cpu_baseline(args.cpu_scale, args.k, args.seed, hgp)
        log(f"cpu baseline: {cpu['value']} M edges/s ({cpu['sample']})")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "M edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": f"rmat_s{args.scale}_e{args.edges_log2}_k{args.k}",
            "n": n,
            "edges": E,
            "k": args.k,
            "ub": 1.03,
            "seed": args.seed,
            "xadj": "int64",
        },
        "cut": cut,
        "cut_fraction": round(cut / E, 4),
        "imbalance": round(maxp * args.k / tvwgt, 5),
        "balanced_1.03": bool(balanced),
        "device_ms": round(st["ms_total"], 3),
        "phase_ms": {"coarsen": round(st["ms_coarsen"], 3), "init": round(st["ms_init"], 3),
                     "uncoarsen": round(st["ms_uncoarsen"], 3)},
        "levels": st["levels"],
        "refine_iters": st["refine_iters"],
        "pipeline_alg_GBs": round(pipeline_gbs, 2),
        "roofline": roofline,
        "kernels": {k: [v[0], round(v[1], 3), v[2]] for k, v in kt.items()},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
