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
(hunyuangraph_rbbisection) timed on one pinned core, in a child process
started before any GPU call, on the 256x256 grid (k = 8) and the 3-D RGG 2^20
(k = 64) at full size and on a bounded sample of the R-MAT workload (scale 19,
the same generator and edge factor, k = 256); the GPU partitions the same
three graphs for a same-graph speedup and cut ratio.  Also the batched mode
(4096 graphs of 1k..64k vertices) and the serving mode (saturating and
Poisson arrivals).
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


def host_cut_and_maxpart(xadj, adjncy, part, k, vwgt=None, adjwgt=None):
    """Exact cut and heaviest part weight, recounted on the host."""
    n = len(xadj) - 1
    deg = np.diff(xadj)
    rows = np.repeat(np.arange(n, dtype=np.int32), deg)
    diff = part[rows] != part[adjncy]
    cut = (int(np.count_nonzero(diff)) if adjwgt is None else int(adjwgt[diff].astype(np.int64).sum())) // 2
    pw = np.bincount(part, weights=None if vwgt is None else vwgt.astype(np.float64), minlength=k)
    return cut, int(pw.max())


def cpu_graphs(hgp, scale, seed):
    """The CPU baseline's graphs: (name, graph, k)."""
    n2 = 1 << 20
    return [
        ("grid2d_256x256_k8", lambda: hgp.gen_grid2d(256, 256), 8),
        ("rgg3d_2^20_w_k64", lambda: hgp.gen_rgg3d(n2, hgp.rgg_radius(n2, 12.0), seed=seed, vw=(1, 4), ew=(1, 16)), 64),
        (f"rmat_s{scale}_e{scale + 4}_k256", lambda: hgp.gen_rmat(scale, 16 << scale, seed=seed), 256),
    ]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg_child(out_path, core, scale, seed):
    """Child process: pins itself to one core and times the C restatement of
    the reference CPU path (oracle, one thread) on each graph; generation is
    outside the clock.  Touches no GPU."""
    os.sched_setaffinity(0, {core})
    import hunyuansc25_amd as hgp    # host-side generators only
    import oracle as O
    res = {}
    for name, make, k in cpu_graphs(hgp, scale, seed):
        x, a, vw, aw = make()
        t0 = time.perf_counter()
        part, cut = O.cpu_rb_partition(x, a, k, vw, aw, seed=seed)
        dt = time.perf_counter() - t0
        E = len(a) // 2
        tv = int(vw.astype(np.int64).sum())
        pw = np.bincount(part, weights=vw.astype(np.float64), minlength=k)
        res[name] = {"edges": E, "k": k, "seconds": round(dt, 3), "M_edges_per_s": round(E / dt / 1e6, 4),
                     "cut": int(cut), "imbalance": round(float(pw.max()) * k / tv, 5)}
    with open(out_path, "w") as f:
        json.dump(res, f)


def start_cpu_leg(args):
    """Starts the CPU baseline in a child process before this process makes
    any GPU call; the child takes the last allowed core and this process the
    others."""
    import subprocess
    import tempfile
    allowed = sorted(os.sched_getaffinity(0))
    core = allowed[-1]
    out = os.path.join(tempfile.mkdtemp(prefix="hgp_cpu_"), "cpu.json")
    proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-leg-out", out, "--cpu-core", str(core),
                             "--cpu-scale", str(args.cpu_scale), "--seed", str(args.seed)])
    if len(allowed) > 1:
        os.sched_setaffinity(0, set(allowed[:-1]))
    return proc, out, core, len(allowed)


def finish_cpu_leg(leg, gpu_rates):
    """Waits for the child and joins its times with the GPU's on the same graphs."""
    proc, out, core, nallowed = leg
    rc = proc.wait(timeout=900)
    if rc != 0:
        raise SystemExit(f"cpu baseline child failed with status {rc}")
    res = json.load(open(out))
    for name, r in res.items():
        g = gpu_rates.get(name)
        if g:
            r["gpu_M_edges_per_s"] = g["M_edges_per_s"]
            r["gpu_cut"] = g["cut"]
            r["gpu_cut_ratio"] = round(g["cut"] / max(r["cut"], 1), 4)
            r["speedup"] = round(g["M_edges_per_s"] / r["M_edges_per_s"], 1)
    sample = [k for k in res if k.startswith("rmat_")][0]
    s = res[sample]
    return {
        "value": s["M_edges_per_s"],
        "unit": "M edges/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{sample}: R-MAT scale {sample.split('_')[1][1:]} ({s['edges']} edges, edge factor 16 as the "
                  f"2^22/2^26 workload), k=256, one partition in {s['seconds']} s on one pinned core; the 256x256 "
                  f"grid (k=8) and the 3-D RGG 2^20 (k=64) at full size beside it",
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "cores_allowed": nallowed,
        "pinned_core": core,
        "configs": res,
    }


def measured_traffic(workload, kernel):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    of this workload (two rocprofv3 --pmc passes, FETCH_SIZE and WRITE_SIZE,
    scaled by a 1 GiB copy), or None when there is none for it."""
    path = os.path.join(ROOT, "profiles", f"{workload}_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    col = None
    for line in open(path):
        if not line.lstrip().startswith("|"):
            continue
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if col is None:
            if "traffic bytes" in cells:
                col = cells.index("traffic bytes")
            continue
        if len(cells) > col and cells[0] == kernel:
            try:
                return int(float(cells[col]))
            except ValueError:
                return None
    return None


def bench_other(hgp, torch, name, graph, k, seed, steps=3):
    """Throughput of one more configuration, device-resident as the headline."""
    x, a, vw, aw = graph
    n, E = len(x) - 1, len(a) // 2
    dev = torch.device("cuda:0")
    d = [torch.from_numpy(arr).to(dev) for arr in (x, a, vw, aw)]
    d_part = torch.zeros(n, dtype=torch.int32, device=dev)
    run = lambda: hgp.partition_device(n, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), d[3].data_ptr(), k,
                                       d_part.data_ptr(), seed=seed)
    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        st = run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    part = d_part.cpu().numpy()
    cut, maxp = host_cut_and_maxpart(x, a, part, k, vw, aw)
    if cut != st["cut"]:
        raise SystemExit(f"{name}: library cut {st['cut']} != host recount {cut}")
    tv = int(vw.astype(np.int64).sum())
    return {"workload": name, "n": n, "edges": E, "k": k, "M_edges_per_s": round(E / dt / 1e6, 3),
            "ms": round(dt * 1e3, 3), "cut": cut, "imbalance": round(maxp * k / tv, 5),
            "balanced_1.03": bool(100 * k * maxp <= 103 * tv)}


def bench_batch(hgp, count, lanes, seed):
    """Graphs/s of the plain per-graph path (one hgp_partition call after
    another) and of hgp_partition_batch on the same graphs; both from host
    arrays.  The partitions must be identical."""
    from hunyuansc25_amd.workloads import batch_graphs
    gs, ks = batch_graphs(hgp, count, seed)
    E = sum(len(g[1]) // 2 for g in gs)
    hgp.partition_batch(gs[: 2 * lanes], ks[: 2 * lanes], lanes=lanes, seed=seed)   # lanes' arenas warm
    hgp.partition(gs[0][0], gs[0][1], ks[0], seed=seed)
    t0 = time.perf_counter()
    plain = [hgp.partition(g[0], g[1], k, g[2], g[3], seed=seed)[0] for g, k in zip(gs, ks)]
    tp = time.perf_counter() - t0
    t0 = time.perf_counter()
    parts, _ = hgp.partition_batch(gs, ks, lanes=lanes, seed=seed)
    tb = time.perf_counter() - t0
    same = all(np.array_equal(p, q) for p, q in zip(parts, plain))
    return {
        "workload": f"batch_{count}_graphs_1k-64k",
        "graphs": count,
        "edges": int(E),
        "lanes": lanes,
        "plain_graphs_per_s": round(count / tp, 2),
        "batched_graphs_per_s": round(count / tb, 2),
        "speedup": round(tp / tb, 3),
        "identical_to_plain": bool(same),
    }


def bench_serve(hgp, pool_size, count, slots, budget, seed):
    """Serving mode on a pool of pool_size requests (sizes log-uniform in
    [10k, 4M] vertices, R-MAT and 3-D RGG, k in {8..512}); request j is pool
    entry j mod pool_size with that entry's seed, so every result can be
    compared with the entry partitioned alone.
    * plain: the pool one request after another (hgp_partition);
    * saturating: the whole pool submitted at once from host arrays
      (arrivals above any capacity): served graphs/s against plain graphs/s;
    * resident: the pool registered once (graphs resident in HBM), then
      submitted at once twice: the first pass builds and keeps each
      request's coarsening hierarchy, the second reuses them;
    * Poisson: `count` requests on the resident graphs at a fixed rate of
      60 % of the resident capacity: p50/p99 latency (submit to result)."""
    import threading
    from hunyuansc25_amd.workloads import serve_requests
    pool = serve_requests(hgp, pool_size, seed)
    seeds = [seed + i for i in range(pool_size)]
    hgp.partition(pool[0][0][0], pool[0][0][1], pool[0][1], seed=seeds[0])   # warm
    t0 = time.perf_counter()
    alone = [hgp.partition(g[0], g[1], k, g[2], g[3], seed=s)[0] for (g, k), s in zip(pool, seeds)]
    t_plain = time.perf_counter() - t0
    same = True
    with hgp.Server(slots=slots, budget_bytes=budget) as srv:
        # saturating arrivals, host arrays
        start = time.perf_counter()
        tickets = [srv.submit(g[0], g[1], k, g[2], g[3], seed=s) for (g, k), s in zip(pool, seeds)]
        for i, t in enumerate(tickets):
            part, _, _ = srv.wait(t)
            same &= bool(np.array_equal(part, alone[i]))
        t_sat = time.perf_counter() - start
        cap = pool_size / t_sat
        # resident graphs: cold (hierarchies built and kept), then warm
        start = time.perf_counter()
        gids = [srv.register(*g) for g, _ in pool]
        t_reg = time.perf_counter() - start
        passes = []
        for _ in range(2):
            start = time.perf_counter()
            tickets = [srv.submit_graph(gid, k, seed=s) for gid, (_, k), s in zip(gids, pool, seeds)]
            for i, t in enumerate(tickets):
                part, _, _ = srv.wait(t)
                same &= bool(np.array_equal(part, alone[i]))
            passes.append(pool_size / (time.perf_counter() - start))
        kept = srv.kept()
        cap_res = passes[1]
        # Poisson arrivals at a fixed rate on the resident graphs
        rate = 0.6 * cap_res
        rng = np.random.default_rng(seed + 7)
        arrive = np.cumsum(rng.exponential(1.0 / rate, size=count))
        lat = [0.0] * count
        tickets = [None] * count
        ready = threading.Semaphore(0)
        ok = [True]

        def waiter():
            for j in range(count):
                ready.acquire()
                part, st, ms = srv.wait(tickets[j])
                lat[j] = ms
                ok[0] &= bool(np.array_equal(part, alone[j % pool_size]))

        th = threading.Thread(target=waiter)
        th.start()
        start = time.perf_counter()
        for j in range(count):
            delay = start + arrive[j] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            i = j % pool_size
            tickets[j] = srv.submit_graph(gids[i], pool[i][1], seed=seeds[i])
            ready.release()
        th.join()
        span = time.perf_counter() - start
        info = srv.info()
    lat_s = np.sort(np.array(lat))
    plain = pool_size / t_plain
    return {
        "workload": f"serve_pool{pool_size}_10k-4M",
        "slots": slots,
        "budget_GB": round(budget / 2**30, 1),
        "plain_graphs_per_s": round(plain, 3),
        "saturated_graphs_per_s": round(cap, 3),
        "saturated_speedup": round(cap / plain, 3),
        "register_s": round(t_reg, 3),
        "resident_cold_graphs_per_s": round(passes[0], 3),
        "resident_graphs_per_s": round(cap_res, 3),
        "resident_speedup": round(cap_res / plain, 3),
        "hierarchies": kept,
        "poisson_requests": count,
        "poisson_rate_per_s": round(rate, 3),
        "poisson_graphs_per_s": round(count / span, 3),
        "p50_ms": round(float(np.percentile(lat_s, 50)), 2),
        "p99_ms": round(float(np.percentile(lat_s, 99)), 2),
        "peak_reserved_GB": round(info["peak_reserved"] / 2**30, 2),
        "identical_to_alone": bool(same and ok[0]),
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
    ap.add_argument("--no-others", action="store_true", help="skip the grid and RGG configurations")
    ap.add_argument("--no-fm-kinds", action="store_true",
                    help="skip the bisection-FM comparison (scan, queue, reference queue FM) on the headline graph")
    ap.add_argument("--no-refiners", action="store_true",
                    help="skip the older label-propagation refiners (k-way and 2-way) on the headline graph")
    ap.add_argument("--calib-copy", action="store_true",
                    help="one 1 GiB device-to-device copy before the steps (counter calibration)")
    ap.add_argument("--batch-graphs", type=int, default=4096, help="batched-mode graphs (0: skip)")
    ap.add_argument("--lanes", type=int, default=8)
    ap.add_argument("--serve-pool", type=int, default=24, help="serving-mode distinct requests (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=200, help="Poisson-arrival requests")
    ap.add_argument("--cpu-leg-out", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-core", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    args = ap.parse_args()

    if args.cpu_leg_out:
        cpu_leg_child(args.cpu_leg_out, args.cpu_core, args.cpu_scale, args.seed)
        return

    if args.gpus != 1:
        print(json.dumps({"error": "the partitioner runs on one GPU (a graph is not sharded across GPUs); use --gpus 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    leg = None
    if not args.no_cpu:
        leg = start_cpu_leg(args)
        log(f"cpu baseline: child process on core {leg[2]} (this process on the other {leg[3] - 1})")

    log("bench: loading library and torch")
    import torch                      # first: libhgp.so then binds to torch's HIP runtime
    import hunyuansc25_amd as hgp
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

    if args.calib_copy:
        src = torch.ones(1 << 28, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 1073741824 bytes read, 1073741824 bytes written")
        del src, dst

    for i in range(args.warmup):
        st = step()
        log(f"warmup {i}: {st['ms_total']:.1f} ms device, cut {st['cut']}")

    torch.cuda.synchronize()
    t_start = time.perf_counter()
    stats = []
    for i in range(args.steps):
        stats.append(step())
        log(f"step {i}: {stats[-1]['ms_total']:.1f} ms device")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    # one more step, outside the timed region, with the per-kernel HIP-event timer on
    kt = {}
    if not args.no_ktimer:
        hgp.ktimer_enable(True)
        st_k = step()
        kt = hgp.ktimer_report()
        hgp.ktimer_enable(False)
        log(f"kernel-timer step: {st_k['ms_total']:.1f} ms device")
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
        # the dominant kernel: largest total time among kernels with a byte model
        dom = max(((k_, v) for k_, v in kt.items() if v[2] > 0), key=lambda kv: kv[1][1])
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
            "traffic": measured_traffic(f"rmat_s{args.scale}_e{args.edges_log2}_k{args.k}", name),
            "alg_bytes_per_launch": round(byts / launches),
        }
    pipeline_gbs = st["alg_bytes"] / (st["ms_total"] * 1e-3) / 1e9

    fm_kinds = None
    if not args.no_fm_kinds:
        # the initial partition of the headline graph's coarsest level with
        # each FM of the bisection trials (device time of that phase)
        fm_kinds = {}
        for kind, label in ((0, "wave_argmax_scan"), (1, "priority_queue_same_moves"),
                            (2, "reference_priority_queue_fm")):
            d_tmp = torch.zeros(n, dtype=torch.int32, device=dev)
            s_k = hgp.partition_device(n, d_x.data_ptr(), d_a.data_ptr(), d_vw.data_ptr(), d_aw.data_ptr(), args.k,
                                       d_tmp.data_ptr(), seed=args.seed, fm_kind=kind)
            fm_kinds[label] = {"fm_kind": kind, "coarsest_n": s_k["coarsest_n"], "init_ms": round(s_k["ms_init"], 3),
                               "init_cut": s_k["init_cut"], "cut": s_k["cut"], "ms_total": round(s_k["ms_total"], 3)}
            del d_tmp
        log(f"bisection FM kinds: {fm_kinds}")

    refiners = None
    if not args.no_refiners:
        # the older label propagation the reference keeps beside Jet, on the
        # headline graph: the k-way rounds as the level refiner, and the 2-way
        # pass on a random 30 % / 70 % split (host arrays in and out)
        d_tmp = torch.zeros(n, dtype=torch.int32, device=dev)
        s_l = hgp.partition_device(n, d_x.data_ptr(), d_a.data_ptr(), d_vw.data_ptr(), d_aw.data_ptr(), args.k,
                                   d_tmp.data_ptr(), seed=args.seed, refiner=1)
        lp_part = d_tmp.cpu().numpy()
        del d_tmp
        lcut, lmax = host_cut_and_maxpart(x, a, lp_part, args.k)
        w0 = (np.random.default_rng(args.seed).random(n) < 0.3).astype(np.int32)
        t2 = time.perf_counter()
        w1, ran = hgp.refine_2way_lp(x, a, vw, aw, w0)
        t2 = time.perf_counter() - t2
        refiners = {
            "jet": {"ms_total": round(st["ms_total"], 3), "ms_uncoarsen": round(st["ms_uncoarsen"], 3),
                    "cut": cut, "imbalance": round(maxp * args.k / tvwgt, 5), "refine_iters": st["refine_iters"]},
            "older_lp": {"ms_total": round(s_l["ms_total"], 3), "ms_uncoarsen": round(s_l["ms_uncoarsen"], 3),
                         "cut": lcut, "imbalance": round(lmax * args.k / tvwgt, 5),
                         "rounds": s_l["refine_iters"], "cut_check": s_l["cut"] == lcut},
            "lp_2way_pass": {"wall_ms_with_copies": round(t2 * 1e3, 3), "transfer_ran": bool(ran),
                             "cut_before": host_cut_and_maxpart(x, a, w0, 2)[0],
                             "cut_after": host_cut_and_maxpart(x, a, w1, 2)[0]},
        }
        log(f"refiners: {refiners}")

    others = []
    gpu_rates = {}
    if not args.no_others:
        log("other configurations: grid 256x256 k=8, 3-D RGG 2^20 weighted k=64, R-MAT sample ...")
        for name, make, k in cpu_graphs(hgp, args.cpu_scale, args.seed):
            r = bench_other(hgp, torch, name, make(), k, args.seed)
            others.append(r)
            gpu_rates[name] = r
        log(f"other configurations: {others}")

    batched = None
    if args.batch_graphs > 0:
        log(f"batched mode: {args.batch_graphs} graphs, {args.lanes} lanes ...")
        batched = bench_batch(hgp, args.batch_graphs, args.lanes, args.seed)
        log(f"batched mode: {batched}")
    serving = None
    if args.serve_pool > 0:
        log(f"serving mode: pool of {args.serve_pool}, {args.serve_requests} Poisson requests, {args.slots} slots ...")
        serving = bench_serve(hgp, args.serve_pool, args.serve_requests, args.slots, int(args.budget_gb * 2**30),
                              args.seed)
        log(f"serving mode: {serving}")

    cpu = None
    if leg:
        log("cpu baseline: waiting for the child ...")
        cpu = finish_cpu_leg(leg, gpu_rates)
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
        "arena_GB": round(st["arena_bytes"] / 2**30, 2),
        "arena_peak_GB": round(st["arena_peak"] / 2**30, 2),
        "refine_iters": st["refine_iters"],
        "pipeline_alg_GBs": round(pipeline_gbs, 2),
        "roofline": roofline,
        "kernels": {k: [v[0], round(v[1], 3), v[2]] for k, v in kt.items()},
        "bisection_fm": fm_kinds,
        "refiners": refiners,
        "cpu_baseline": cpu,
        "other_configs": others,
        "batched": batched,
        "serving": serving,
    }
    if cpu:
        # the same graph on both sides: the R-MAT sample
        same = cpu["configs"][[k for k in cpu["configs"] if k.startswith("rmat_")][0]]
        if "speedup" in same:
            out["speedup_vs_cpu"] = same["speedup"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
