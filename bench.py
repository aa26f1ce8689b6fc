"""Benchmark of the CS enumeration hot path on one MI355X.

Workload (the dedup-stress configuration): instances 0..M-1 of synthetic
config 2 -- 5 atomic propositions, 40 traces (20 positive / 20 negative),
lengths uniform in [16, 32], bits Bernoulli(0.3), labels from a planted
size-9 formula containing Until, max_formula_size 14, hashed 128-bit keys
(total trace length > 126 bits).  One step = solving all M instances, each
as a plain single-instance solve (the drop-in path).  Inputs are resident:
the instances are generated and their atom rows uploaded to HBM
(vfb_instance_upload) before the timed region.

Prints ONE JSON line: distinct CS/s over the timed steps, plus batched vs
plain instances/s (checked bit-exact), p50/p99 solve latency, the roofline
of the combine+key+dedup kernel (k_expand) from HIP events, and the CPU
baseline (the single-threaded oracle on a bounded sample).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def profiled_traffic():
    """HBM bytes per k_expand launch of this workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by the copy-kernel
    calibration (profiles/k_expand_config2_main.md); None if absent."""
    path = os.path.join(ROOT, "profiles", "k_expand_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)["traffic_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(insts, max_size, budget_s, gpu_results=None):
    """Multi-core CPU build of the same enumeration (oracle/cpu_omp.c,
    OpenMP over this host's cores) on the same instances, timed around each
    solve alone; stops after budget_s seconds of CPU work.  Its results are
    compared with the GPU's."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    done, distinct, t_total, same = 0, 0, 0.0, True
    for k, inst in enumerate(insts):
        names, lens, atoms, P = inst.arrays()
        r, dt = O.enumerate_omp(O.Instance(names, lens, atoms, P), max_size, 3)
        t_total += dt
        distinct += sum(r["distinct"])
        done += 1
        if gpu_results is not None:
            g = gpu_results[k]
            same &= all(r[f] == g[f] for f in ("status", "min_size", "formula", "distinct", "candidates"))
        if t_total > budget_s:
            break
    return {"value": distinct / t_total if t_total else None, "unit": "distinct CS/s",
            "cores": O.omp_threads(), "kind": "port",
            "sample": f"config-2 instances 0..{done - 1} at max_formula_size {max_size} "
                      f"({distinct} distinct CS in {t_total:.2f} s, OpenMP build of the same "
                      f"enumeration, {O.omp_threads()} threads)",
            "bit_exact_vs_gpu": same}


def batched_mix(V, n, max_size, autotune=True):
    """Config 4: n uneven instances, batched vs sequential plain solves.

    The library's autotuner first picks launch / pass parameters on a sample
    of the mix (instances 0..47, batched); both legs then run with them.  Also
    reported: the same comparison without the instances that keep the GPU
    busy alone (plain solve over 10 ms), where batching has idle GPU to fill."""
    insts = [V.Instance.generate(4, k) for k in range(n)]
    tuned = None
    if autotune:
        sample = insts[:min(48, n)]
        tuned = V.autotune(sample, max_size=max_size, batched=True, reps=2)
    V.solve(insts[0], max_size=max_size)
    t0 = time.perf_counter()
    plain = [V.solve(i, max_size=max_size) for i in insts]
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch = V.solve_batch(insts, max_size=max_size)
    t_batch = time.perf_counter() - t0
    fields = ("status", "min_size", "formula", "distinct", "candidates", "level_winners", "otf_level", "otf_stored")
    exact = all(a[f] == b[f] for a, b in zip(plain, batch) for f in fields)
    lat = [r["ms_total"] for r in plain]
    small = [k for k in range(n) if plain[k]["ms_total"] < 10.0]
    t_small_plain = sum(plain[k]["ms_total"] for k in small) / 1e3
    t0 = time.perf_counter()
    sb = V.solve_batch([insts[k] for k in small], max_size=max_size)
    t_small_batch = time.perf_counter() - t0
    exact_small = all(plain[k][f] == b[f] for k, b in zip(small, sb) for f in fields)
    return {"instances": n, "max_formula_size": max_size,
            "instances_per_s_plain": n / t_plain, "instances_per_s_batched": n / t_batch,
            "batched_vs_plain": t_plain / t_batch, "bit_exact": exact,
            "plain_latency_ms_p50": pct(lat, 0.5), "plain_latency_ms_p99": pct(lat, 0.99),
            "distinct_cs": sum(sum(r["level_winners"]) for r in plain),
            "batch_restarts": sum(r["restarts"] for r in batch), "batch_parked": sum(r["parked"] for r in batch),
            "gpu_saturating_instances": n - len(small),
            "gpu_saturating_share_of_plain_time": 1.0 - t_small_plain / t_plain,
            "others": {"instances": len(small), "batched_vs_plain": t_small_plain / t_small_batch,
                       "bit_exact": exact_small},
            "autotuned": tuned}


def served(V, rate, seconds, max_size, max_active, check):
    """Config 5: open-loop Poisson arrivals of config-4-distributed requests."""
    import random
    import threading
    rng = random.Random(5)
    arrivals, t = [], 0.0
    while True:
        t += rng.expovariate(rate)
        if t > seconds:
            break
        arrivals.append(t)
    insts = [V.Instance.generate(5, k) for k in range(len(arrivals))]
    results = [None] * len(insts)
    with V.Server(max_size=max_size, max_active=max_active) as srv:
        tickets = [None] * len(insts)
        submitted = threading.Semaphore(0)

        def collect():
            for k in range(len(insts)):
                submitted.acquire()
                results[k] = srv.wait(tickets[k])
        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()

        def monitor():
            while th.is_alive():
                time.sleep(5)
                done = sum(r is not None for r in results)
                print(f"[bench] served {time.perf_counter() - t0:.0f}s: {done}/{len(insts)} done, "
                      f"{srv.stats()}", file=sys.stderr, flush=True)
        threading.Thread(target=monitor, daemon=True).start()
        for k, a in enumerate(arrivals):
            d = t0 + a - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[k] = srv.submit(insts[k])
            submitted.release()
        th.join()
        elapsed = time.perf_counter() - t0
        stats = srv.stats()
    lat = [r["ms_total"] for r in results]
    sample = list(range(0, len(insts), max(1, len(insts) // check)))[:check]
    plain_t0 = time.perf_counter()
    plain = {k: V.solve(insts[k], max_size=max_size) for k in sample}
    plain_rate = len(sample) / (time.perf_counter() - plain_t0)
    fields = ("status", "min_size", "formula", "distinct", "candidates", "level_winners")
    exact = all(plain[k][f] == results[k][f] for k in sample for f in fields)
    return {"rate_req_s": rate, "seconds": seconds, "requests": len(insts), "max_active": max_active,
            "instances_per_s": len(insts) / elapsed, "latency_ms_p50": pct(lat, 0.5),
            "latency_ms_p99": pct(lat, 0.99), "peak_active": stats["peak_active"],
            "plain_sequential_instances_per_s": plain_rate, "bit_exact_vs_plain_sample": exact,
            "checked": len(sample)}


def limits(V, ks, max_size):
    """Config 3 (10 APs, 63 traces of length 48-64, max size 18): large levels,
    where the combine+dedup kernel runs at its bandwidth roof."""
    out = []
    for k in ks:
        inst = V.Instance.generate(3, k)
        V.solve(inst, max_size=max_size)            # warm
        t0 = time.perf_counter()
        r = V.solve(inst, max_size=max_size, timing=True)
        dt = time.perf_counter() - t0
        ach = r["bytes_expand"] / (r["ms_expand"] * 1e-3) / 1e9 if r["ms_expand"] else None
        out.append({"instance": k, "status": V.status_string(r["status"]), "min_size": r["min_size"],
                    "seconds": dt, "distinct_cs": sum(r["distinct"]),
                    "distinct_per_s": sum(r["distinct"]) / dt,
                    "candidates_per_s": sum(r["level_candidates"]) / dt,
                    "onthefly_level": r["otf_level"],
                    "k_expand_ms": r["ms_expand"], "k_expand_launches": r["expand_launches"],
                    "k_expand_alg_GBps": ach, "k_expand_frac_of_hbm": ach / HBM_PEAK_GBS if ach else None})
    return out


def _stutter(w, n, k):
    """An n-step trace with every step held for k steps (k * n steps)."""
    out = 0
    for t in range(n):
        if (w >> t) & 1:
            out |= ((1 << k) - 1) << (k * t)
    return out


def multiword(V, n, stretch, max_size):
    """Traces over one CS word: config-2 instances 0..n-1 with every step held
    for `stretch` steps (traces of 16*stretch..32*stretch steps, labels kept),
    plain solves.  The combine+dedup kernel is k_expand_mw; its algorithmic
    bytes are (a * 8W + 16) per candidate, W = the row's words."""
    insts, words = [], []
    for k in range(n):
        names, lens, atoms, P = V.Instance.generate(2, k).arrays()
        st = [[_stutter(w, L, stretch) for w, L in zip(row, lens)] for row in atoms]
        inst = V.Instance.from_bits(names, [L * stretch for L in lens], st, P)
        insts.append(inst)
        words.append(V.lib().vfb_instance_words(inst.handle))
    V.solve(insts[0], max_size=max_size)          # warm
    t0 = time.perf_counter()
    rs = [V.solve(i, max_size=max_size, timing=2) for i in insts]
    dt = time.perf_counter() - t0
    ms = sum(r["ms_expand"] for r in rs)
    by = sum(r["bytes_expand"] for r in rs)
    ach = by / (ms * 1e-3) / 1e9 if ms else None
    st = {}
    for r in rs:
        st[V.status_string(r["status"])] = st.get(V.status_string(r["status"]), 0) + 1
    return {"workload": f"config-2 instances 0..{n - 1}, every step held {stretch}x "
                        f"({16 * stretch}-{32 * stretch} steps per trace), max_formula_size {max_size}",
            "words_per_row": [min(words), max(words)], "seconds": dt,
            "distinct_per_s": sum(sum(r["distinct"]) for r in rs) / dt,
            "candidates_per_s": sum(sum(r["level_candidates"]) for r in rs) / dt,
            "statuses": st, "min_sizes": [r["min_size"] for r in rs],
            "k_expand_mw_ms": ms, "k_expand_mw_launches": sum(r["expand_launches"] for r in rs),
            "k_expand_mw_alg_GBps": ach, "k_expand_mw_frac_of_hbm": ach / HBM_PEAK_GBS if ach else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--max-size", type=int, default=14)
    ap.add_argument("--cpu-max-size", type=int, default=14)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mix", type=int, default=512, help="config-4 instances for batched vs plain (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="config-5 served stream length (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-active", type=int, default=128)
    ap.add_argument("--limits", default="4,10", help="config-3 instances to time (empty: skip)")
    ap.add_argument("--multiword", default="8,8,12",
                    help="instances,stretch,max_size of the long-trace workload (empty: skip)")
    ap.add_argument("--only", choices=["main", "limits", "mix", "multiword", "serve"], default=None,
                    help="run one section alone and print its JSON (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(instances shard across GPUs with no exchange)"}))
        sys.exit(2)

    import vfb_amd as V
    dev = V.device_info()
    if dev is None:
        print(json.dumps({"error": "no HIP device"}))
        sys.exit(1)
    print(f"[bench] device {dev}", file=sys.stderr, flush=True)
    if args.only == "limits":
        print(json.dumps({"limits": limits(V, [int(x) for x in args.limits.split(",")], 18)}))
        return
    if args.only == "mix":
        print(json.dumps({"batched_mix": batched_mix(V, args.mix, 16)}))
        return
    if args.only == "serve":
        print(json.dumps({"served": served(V, args.serve_rate, args.serve_seconds, 16, args.serve_active, 256)}))
        return
    if args.only == "multiword":
        n, k, m = (int(x) for x in args.multiword.split(","))
        print(json.dumps({"multiword": multiword(V, n, k, m)}))
        return
    # inputs resident in HBM before the timed region: each instance's atom
    # rows are uploaded once, and the solves read them there
    insts = [V.Instance.generate(2, k).upload() for k in range(args.instances)]

    def step():
        res = []
        for inst in insts:
            # timing=2: HIP events around the expand kernel only (the roofline's)
            res.append(V.solve(inst, max_size=args.max_size, timing=2))
        return res

    for w in range(args.warmup):
        t0 = time.perf_counter()
        step()
        print(f"[bench] warmup {w} {time.perf_counter() - t0:.3f}s", file=sys.stderr, flush=True)

    step_times, all_res = [], []
    for s in range(args.steps):
        t0 = time.perf_counter()
        res = step()
        dt = time.perf_counter() - t0
        step_times.append(dt)
        all_res.append(res)
        print(f"[bench] step {s} {dt:.3f}s", file=sys.stderr, flush=True)
    T = sum(step_times)
    # the reference's count: CS stored per level, and at the level of the
    # returned formula the survivors ordered before it (StoredLTLs)
    distinct = sum(sum(r["distinct"]) for res in all_res for r in res)
    level_distinct = sum(sum(r["level_winners"]) for res in all_res for r in res)
    cands = sum(sum(r["level_candidates"]) for res in all_res for r in res)
    lat = [r["ms_total"] for res in all_res for r in res]
    plain = all_res[-1]
    statuses = {}
    for r in plain:
        statuses[r["status"]] = statuses.get(r["status"], 0) + 1

    if args.only == "main":
        print(json.dumps({"distinct_per_s": distinct / T, "ms_per_step": 1e3 * T / args.steps}))
        return
    # batched: the same instances stepped together; must equal the plain runs
    V.solve_batch(insts, max_size=args.max_size)  # warm
    t0 = time.perf_counter()
    batch = V.solve_batch(insts, max_size=args.max_size)
    t_batch = time.perf_counter() - t0
    exact = all(a[f] == b[f] for a, b in zip(plain, batch)
                for f in ("status", "min_size", "formula", "distinct", "candidates", "level_winners"))
    t_plain = step_times[-1]

    bytes_exp = sum(r["bytes_expand"] for res in all_res for r in res)
    ms_exp = sum(r["ms_expand"] for res in all_res for r in res)
    launches = sum(r["expand_launches"] for res in all_res for r in res)
    achieved = bytes_exp / (ms_exp * 1e-3) / 1e9 if ms_exp else None

    out = {
        "metric": "distinct CS/sec + instances/sec batched vs plain (bit-exact); p50/p99 solve latency",
        "value": distinct / T,
        "unit": "distinct CS/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * T / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"config 2 dedup stress: {args.instances} instances "
                               f"(5 APs, 40 traces 20+/20-, len U[16,32], p=0.3, planted size 9 with U), "
                               f"plain single-instance solves, max_formula_size {args.max_size}, "
                               f"relaxed 128-bit hashed keys",
                   "instances": args.instances, "max_formula_size": args.max_size,
                   "parallelism": "single GPU"},
        "candidates_per_s": cands / T,
        "level_distinct_per_s": level_distinct / T,
        "instances_per_s_plain": len(insts) / t_plain,
        "instances_per_s_batched": len(insts) / t_batch,
        "batched_vs_plain": t_plain / t_batch,
        "batched_bit_exact": exact,
        "latency_ms_p50": pct(lat, 0.5), "latency_ms_p99": pct(lat, 0.99),
        "statuses": {V.status_string(k): v for k, v in statuses.items()},
        "roofline": {"bound": "hbm", "kernel": "k_expand", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": profiled_traffic(),
                     "launches": launches, "ms_per_launch": ms_exp / launches if launches else None,
                     "alg_bytes_per_launch": bytes_exp / launches if launches else None},
        "device": dev,
    }
    if args.limits:
        print("[bench] limits config", file=sys.stderr, flush=True)
        out["limits"] = limits(V, [int(x) for x in args.limits.split(",")], 18)
    if args.multiword:
        print("[bench] multi-word traces", file=sys.stderr, flush=True)
        n, k, m = (int(x) for x in args.multiword.split(","))
        out["multiword"] = multiword(V, n, k, m)
    if args.mix:
        print("[bench] batched mix", file=sys.stderr, flush=True)
        out["batched_mix"] = batched_mix(V, args.mix, 16)
    if args.serve_seconds:
        print("[bench] served stream", file=sys.stderr, flush=True)
        out["served"] = served(V, args.serve_rate, args.serve_seconds, 16, args.serve_active, 256)
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(insts, args.cpu_max_size, args.cpu_seconds,
                                           plain if args.cpu_max_size == args.max_size else None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
