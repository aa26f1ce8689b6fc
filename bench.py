"""Benchmark of the CS enumeration hot path on one MI355X.

Headline workload (the largest single-GPU configuration): the config-4
batched mix -- 512 independent synthetic instances with APs uniform in [2,8],
traces in [8,63] with a random positive/negative split, lengths in [8,64],
bits Bernoulli(p), p uniform in [0.2,0.6], labels from planted random
formulas of size 5-13, max_formula_size 16 -- solved together by
vfb_solve_batch (shared launches, members that fill the GPU alone parked and
run one at a time), with launch parameters from the library's autotuner on a
48-instance sample.  One step = one batched solve of all 512 instances.
Inputs are resident: the instances are generated and their atom rows
uploaded to HBM (vfb_instance_upload) before the timed region.

value = distinct CS per second over the timed steps, where distinct CS is
the reference's StoredLTLs count (`distinct[]`: the CS stored per completed
level, and at the level of the returned formula the survivors ordered before
it), summed over the 512 instances.  Beside it: the same instances solved one
after another by the plain path (instances/s, batched vs plain, checked
bit-exact), the combine+key+dedup kernel's roofline from HIP events over the
timed region, and the CPU baseline (OpenMP over instances, on a bounded
sample of the mix).

Sections: config 2 (dedup stress, 32 plain solves), config 3 (limits,
instances 4 and 10), multi-word traces, and the served stream (config 5:
Poisson arrivals of config-4-distributed requests, 200 req/s for 60 s).

Prints ONE JSON line.
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
MIX_N, MIX_MAX = 512, 16
GIANT_MS = 10.0            # a plain solve this long fills the GPU alone
SERVE_REQUEST_BUDGET = 96 << 30
FIELDS = ("status", "min_size", "formula", "distinct", "candidates", "level_winners", "otf_level", "otf_stored",
          "found_ordinal")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = (len(xs) - 1) * q
    f = int(k)
    c = min(f + 1, len(xs) - 1)
    return xs[f] + (xs[c] - xs[f]) * (k - f)


def profiled_traffic(workload):
    """HBM bytes per k_expand launch of a workload, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs scaled by the copy-kernel
    calibration (profiles/k_expand_traffic.json); None if absent."""
    try:
        with open(os.path.join(ROOT, "profiles", "k_expand_traffic.json")) as f:
            return json.load(f)[workload]
    except (OSError, ValueError, KeyError):
        return None


def roofline(ks, traffic):
    """The combine+key+dedup kernel (k_expand, dedup passes) over the
    counters of vfb_kernel_stats: algorithmic bytes / HIP-event time."""
    ms, by, n = ks["ms_expand"], ks["bytes_expand"], ks["expand_launches"]
    ach = by / (ms * 1e-3) / 1e9 if ms else None
    out = {"bound": "hbm", "kernel": "k_expand", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": ach / HBM_PEAK_GBS if ach else None,
           "traffic": traffic["bytes_per_launch"] if traffic else None,
           "launches": n, "ms_per_launch": ms / n if n else None,
           "alg_bytes_per_launch": by / n if n else None}
    if traffic and ms and n:
        # physical: the profiled HBM bytes of an average launch over the
        # average launch time measured here
        phys = traffic["bytes_per_launch"] / (ms / n * 1e-3) / 1e9
        out.update({"physical_GBps": phys, "physical_frac": phys / HBM_PEAK_GBS,
                    "traffic_source": traffic.get("source")})
    return out


def generate(V, config, n):
    """Instances 0..n-1 of a configuration, generated on 16 host threads
    (the generator's rejection sampling takes ~60 ms per mix instance; the
    library call releases the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    out = [None] * n
    with ThreadPoolExecutor(max_workers=16) as ex:
        futs = {ex.submit(V.Instance.generate, config, k): k for k in range(n)}
        done = 0
        for f in futs:
            out[futs[f]] = f.result()
            done += 1
            if done % 2000 == 0:
                log(f"generated {done}/{n} config-{config} instances")
    return out


def autotune_axes(path, ms_default):
    """Per-axis outcome of an autotuner run, from the '# tried <key> <value>
    <ms>' lines it wrote: the best time over each axis's values (the search
    is coordinate-wise, so each axis starts from the best setting so far)."""
    axes = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 5 and p[:2] == ["#", "tried"] and p[2] != "start":
                ms = float(p[4])
                if p[2] not in axes or ms < axes[p[2]]["best_ms"]:
                    axes[p[2]] = {"best_value": int(p[3]), "best_ms": ms}
    for a in axes.values():
        a["gain_vs_default"] = ms_default / a["best_ms"] if a["best_ms"] else None
    return axes


def _same(a, b):
    return all(a[f] == b[f] for f in FIELDS)


def headline(V, steps, warmup, autotune, batch_only=False, tuning_file=None, save_tuning=None):
    log("config-4 mix: generating and uploading 512 instances")
    insts = [i.upload() for i in generate(V, 4, MIX_N)]
    tuned = None
    if tuning_file:
        tuned = {"tuning": V.load_tuning(tuning_file), "from": tuning_file}
        V.set_tuning(**tuned["tuning"])
        log(f"tuning from {tuning_file}: {tuned['tuning']}")
    elif autotune:
        # a sample of the mix: its first 48 instances and three whose largest
        # levels reach 10^7-10^9 candidates (the large levels set the mix's
        # pace; 104 has the mix's largest, a 2 G-candidate level)
        sample = insts[:48] + [insts[104], insts[219], insts[456]]
        log("autotuning on instances 0..47, 104, 219 and 456 of the mix (batched)")
        cache = save_tuning or os.path.join(ROOT, "gpurun_out", "mix_tuning_autotune.txt")
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        tuned = V.autotune(sample, max_size=MIX_MAX, batched=True, reps=2, cache=cache)
        tuned["per_axis"] = autotune_axes(cache, tuned["ms_default"])
        log(f"tuned {tuned}")
    for w in range(warmup):
        t0 = time.perf_counter()
        V.solve_batch(insts, max_size=MIX_MAX, timing=2)
        log(f"warmup {w} {time.perf_counter() - t0:.3f}s")
    V.kernel_stats(reset=True)
    times, runs = [], []
    for s in range(steps):
        t0 = time.perf_counter()
        res = V.solve_batch(insts, max_size=MIX_MAX, timing=2)
        dt = time.perf_counter() - t0
        times.append(dt)
        runs.append(res)
        log(f"step {s} {dt:.3f}s")
    ks = V.kernel_stats()
    T = sum(times)
    distinct = sum(sum(r["distinct"]) for res in runs for r in res)
    if batch_only:        # the batched steps alone (profiler runs)
        return {"out": {"value": distinct / T, "ms_per_step": 1e3 * T / steps, "kernel_stats": ks}}
    level_distinct = sum(sum(r["level_winners"]) for res in runs for r in res)
    cands = sum(sum(r["level_candidates"]) for res in runs for r in res)
    steady = all(_same(a, b) for res in runs[1:] for a, b in zip(runs[0], res))
    log("plain sequential solves of the same 512")
    t0 = time.perf_counter()
    plain = [V.solve(i, max_size=MIX_MAX) for i in insts]
    t_plain = time.perf_counter() - t0
    exact = all(_same(a, b) for a, b in zip(plain, runs[-1]))
    lat = [r["ms_total"] for r in plain]
    small = [k for k in range(MIX_N) if plain[k]["ms_total"] < GIANT_MS]
    t_small_plain = sum(plain[k]["ms_total"] for k in small) / 1e3
    t0 = time.perf_counter()
    sb = V.solve_batch([insts[k] for k in small], max_size=MIX_MAX)
    t_small_batch = time.perf_counter() - t0
    exact_small = all(_same(plain[k], b) for k, b in zip(small, sb))
    statuses = {}
    for r in runs[-1]:
        statuses[V.status_string(r["status"])] = statuses.get(V.status_string(r["status"]), 0) + 1
    return {
        "insts": insts, "plain": plain, "small": small, "t_small_batch": t_small_batch,
        "out": {
            "value": distinct / T, "ms_per_step": 1e3 * T / steps,
            "level_distinct_per_s": level_distinct / T, "candidates_per_s": cands / T,
            "instances_per_s_batched": MIX_N * steps / T, "instances_per_s_plain": MIX_N / t_plain,
            "batched_vs_plain": t_plain / (T / steps), "bit_exact_vs_plain": exact and steady,
            "plain_latency_ms_p50": pct(lat, 0.5), "plain_latency_ms_p99": pct(lat, 0.99),
            "batch_restarts": sum(r["restarts"] for r in runs[-1]),
            "batch_parked": sum(r["parked"] for r in runs[-1]),
            "statuses": statuses,
            "gpu_saturating_instances": MIX_N - len(small),
            "gpu_saturating_share_of_plain_time": 1.0 - t_small_plain / t_plain,
            "others": {"instances": len(small), "batched_vs_plain": t_small_plain / t_small_batch,
                       "bit_exact": exact_small},
            "kernel_ms": {"expand": ks["ms_expand"]},
            "roofline": roofline(ks, profiled_traffic("config4_mix")),
            "autotuned": tuned,
        },
    }


def cpu_baseline_mix(V, H):
    """OpenMP over instances (oracle/cpu_omp.c, each instance by the
    sequential enumeration, dynamic scheduling) on the mix's instances whose
    plain GPU solve takes under GIANT_MS -- the ones that fill the GPU alone
    take minutes each on the host -- compared bit-exact with the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    small = H["small"]
    oi = []
    for k in small:
        names, lens, atoms, P = H["insts"][k].arrays()
        oi.append(O.Instance(names, lens, atoms, P))
    res, dt = O.enumerate_many_omp(oi, MIX_MAX, 3)
    same = all(all(r[f] == H["plain"][k][f] for f in ("status", "min_size", "formula", "distinct", "candidates",
                                                         "found_ordinal"))
               for r, k in zip(res, small))
    distinct = sum(sum(r["distinct"]) for r in res)
    gpu_distinct = sum(sum(H["plain"][k]["distinct"]) for k in small)
    return {"value": distinct / dt, "unit": "distinct CS/s", "cores": O.omp_threads(), "kind": "port",
            "sample": f"the {len(small)} instances of the 512-instance config-4 mix whose plain GPU solve takes "
                      f"under {GIANT_MS:g} ms ({distinct} distinct CS in {dt:.2f} s; OpenMP over instances, "
                      f"{O.omp_threads()} threads, each instance by the sequential enumeration); the "
                      f"{MIX_N - len(small)} left out fill the GPU alone and take minutes each on the host",
            "gpu_batched_same_sample_distinct_per_s": gpu_distinct / H["t_small_batch"],
            "bit_exact_vs_gpu": same}


def config2(V, n, max_size, steps, cpu_seconds):
    """Config 2 (dedup stress): n plain single-instance solves per step."""
    insts = [V.Instance.generate(2, k).upload() for k in range(n)]
    for i in insts:
        V.solve(i, max_size=max_size)
    V.kernel_stats(reset=True)
    times, runs = [], []
    for s in range(steps):
        t0 = time.perf_counter()
        runs.append([V.solve(i, max_size=max_size, timing=2) for i in insts])
        times.append(time.perf_counter() - t0)
    ks = V.kernel_stats()
    T = sum(times)
    plain = runs[-1]
    V.solve_batch(insts, max_size=max_size)
    t0 = time.perf_counter()
    batch = V.solve_batch(insts, max_size=max_size)
    t_batch = time.perf_counter() - t0
    lat = [r["ms_total"] for res in runs for r in res]
    out = {"workload": f"config 2 dedup stress: instances 0..{n - 1} (5 APs, 40 traces 20+/20-, len U[16,32], "
                       f"p=0.3, planted size 9 with U), plain single-instance solves, max_formula_size {max_size}",
           "distinct_per_s": sum(sum(r["distinct"]) for res in runs for r in res) / T,
           "ms_per_step": 1e3 * T / steps,
           "instances_per_s_plain": n * steps / T, "instances_per_s_batched": n / t_batch,
           "batched_vs_plain": (T / steps) / t_batch,
           "batched_bit_exact": all(_same(a, b) for a, b in zip(plain, batch)),
           "latency_ms_p50": pct(lat, 0.5), "latency_ms_p99": pct(lat, 0.99),
           "roofline": roofline(ks, profiled_traffic("config2_main"))}
    if cpu_seconds:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        done, distinct, t_total, same = 0, 0, 0.0, True
        for k, inst in enumerate(insts):
            names, lens, atoms, P = inst.arrays()
            r, dt = O.enumerate_omp(O.Instance(names, lens, atoms, P), max_size, 3)
            t_total += dt
            distinct += sum(r["distinct"])
            done += 1
            same &= all(r[f] == plain[k][f] for f in ("status", "min_size", "formula", "distinct", "candidates"))
            if t_total > cpu_seconds:
                break
        out["cpu_baseline"] = {"value": distinct / t_total if t_total else None, "unit": "distinct CS/s",
                               "cores": O.omp_threads(), "kind": "port",
                               "sample": f"config-2 instances 0..{done - 1} at max_formula_size {max_size}, "
                                         f"OpenMP within each solve",
                               "bit_exact_vs_gpu": same}
    return out


def limits(V, ks_list, max_size):
    """Config 3 (10 APs, 63 traces of 48-64 steps, max size 18), a fixed set
    of instances: large levels, sets grown toward 2^28 slots."""
    out = []
    for k in ks_list:
        inst = V.Instance.generate(3, k).upload()
        V.solve(inst, max_size=max_size)            # warm
        V.kernel_stats(reset=True)
        t0 = time.perf_counter()
        r = V.solve(inst, max_size=max_size, timing=2)
        dt = time.perf_counter() - t0
        ks = V.kernel_stats()
        rf = roofline(ks, profiled_traffic(f"config3_{k}"))
        out.append({"instance": k, "status": V.status_string(r["status"]), "min_size": r["min_size"],
                    "seconds": dt, "distinct_cs": sum(r["distinct"]),
                    "distinct_per_s": sum(r["distinct"]) / dt,
                    "candidates_per_s": sum(r["level_candidates"]) / dt,
                    "table_slots": r["table_slots"], "onthefly_level": r["otf_level"],
                    "k_expand_ms": ks["ms_expand"], "k_expand_launches": ks["expand_launches"],
                    "k_expand_alg_GBps": rf["achieved"], "k_expand_frac_of_hbm": rf["frac"],
                    "k_expand_physical_frac": rf.get("physical_frac")})
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
        inst = V.Instance.from_bits(names, [L * stretch for L in lens], st, P).upload()
        insts.append(inst)
        words.append(V.lib().vfb_instance_words(inst.handle))
    V.solve(insts[0], max_size=max_size)          # warm
    V.kernel_stats(reset=True)
    t0 = time.perf_counter()
    rs = [V.solve(i, max_size=max_size, timing=2) for i in insts]
    dt = time.perf_counter() - t0
    rf = roofline(V.kernel_stats(), None)
    st = {}
    for r in rs:
        st[V.status_string(r["status"])] = st.get(V.status_string(r["status"]), 0) + 1
    return {"workload": f"config-2 instances 0..{n - 1}, every step held {stretch}x "
                        f"({16 * stretch}-{32 * stretch} steps per trace), max_formula_size {max_size}",
            "words_per_row": [min(words), max(words)], "seconds": dt,
            "distinct_per_s": sum(sum(r["distinct"]) for r in rs) / dt,
            "candidates_per_s": sum(sum(r["level_candidates"]) for r in rs) / dt,
            "statuses": st, "min_sizes": [r["min_size"] for r in rs],
            "k_expand_mw_ms": rf["launches"] and rf["ms_per_launch"] * rf["launches"],
            "k_expand_mw_launches": rf["launches"],
            "k_expand_mw_alg_GBps": rf["achieved"], "k_expand_mw_frac_of_hbm": rf["frac"]}


def served(V, rate, seconds, max_size, max_active, check, request_budget):
    """Config 5: open-loop Poisson arrivals of config-4-distributed requests.
    Every request, and every plain solve it is checked against, runs under
    the same per-request HBM cap (below the arena), so that a request that
    needs more is cut the same way in both and the arena keeps room to admit
    small requests while a giant runs."""
    import random
    import threading
    rng = random.Random(5)
    arrivals, t = [], 0.0
    while True:
        t += rng.expovariate(rate)
        if t > seconds:
            break
        arrivals.append(t)
    log(f"generating {len(arrivals)} requests")
    insts = generate(V, 5, len(arrivals))
    results = [None] * len(insts)
    queued_seen = []
    with V.Server(max_size=max_size, max_active=max_active, request_budget=request_budget) as srv:
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
                time.sleep(1)
                s = srv.stats()
                queued_seen.append(s["queued"])
                if len(queued_seen) % 5 == 0:
                    done = sum(r is not None for r in results)
                    log(f"served {time.perf_counter() - t0:.0f}s: {done}/{len(insts)} done, {s}")
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
    # giants: a level of at least 2^24 candidates (the size at which a
    # batch member fills the device alone)
    giant = [max(r["level_candidates"]) >= (1 << 24) for r in results]
    lat_small = [x for x, g in zip(lat, giant) if not g]
    sample = list(range(0, len(insts), max(1, len(insts) // check)))[:check]
    giants = [k for k, g in enumerate(giant) if g]
    sample = sorted(set(sample) | set(giants[:4]))     # the check includes giants
    plain_t0 = time.perf_counter()
    plain = {k: V.solve(insts[k], max_size=max_size, budget=request_budget) for k in sample}
    plain_rate = len(sample) / (time.perf_counter() - plain_t0)
    exact = all(_same(plain[k], results[k]) for k in sample)
    return {"rate_req_s": rate, "seconds": seconds, "requests": len(insts), "max_active": max_active,
            "request_budget_bytes": request_budget,
            "instances_per_s": len(insts) / elapsed, "latency_ms_p50": pct(lat, 0.5),
            "latency_ms_p99": pct(lat, 0.99), "giant_requests": len(giants),
            "non_giant_latency_ms_p50": pct(lat_small, 0.5), "non_giant_latency_ms_p99": pct(lat_small, 0.99),
            "queued_p50": pct(queued_seen, 0.5), "queued_max": max(queued_seen) if queued_seen else None,
            "peak_active": stats["peak_active"], "restarts": sum(r["restarts"] for r in results),
            "plain_sequential_instances_per_s": plain_rate, "bit_exact_vs_plain_sample": exact,
            "checked": len(sample), "checked_giants": len([k for k in sample if giant[k]])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tuning-file", default=None, help="load the mix's launch tuning instead of autotuning")
    ap.add_argument("--save-tuning", default=None, help="save the autotuned launch tuning here")
    ap.add_argument("--config2", type=int, default=32, help="config-2 instances (0: skip)")
    ap.add_argument("--limits", default="4,10", help="config-3 instances to time (empty: skip)")
    ap.add_argument("--multiword", default="8,8,12",
                    help="instances,stretch,max_size of the long-trace workload (empty: skip)")
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="served stream length (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-active", type=int, default=128)
    ap.add_argument("--only", choices=["mix", "mixbatch", "config2", "limits", "multiword", "serve"], default=None,
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
    log(f"device {dev}")
    if args.only == "limits":
        print(json.dumps({"limits": limits(V, [int(x) for x in args.limits.split(",")], 18)}))
        return
    if args.only == "config2":
        print(json.dumps({"config2": config2(V, args.config2, 14, args.steps, 0)}))
        return
    if args.only == "serve":
        print(json.dumps({"served": served(V, args.serve_rate, args.serve_seconds, MIX_MAX, args.serve_active, 256,
                                           SERVE_REQUEST_BUDGET)}))
        return
    if args.only == "multiword":
        n, k, m = (int(x) for x in args.multiword.split(","))
        print(json.dumps({"multiword": multiword(V, n, k, m)}))
        return

    if args.only == "mixbatch":
        print(json.dumps(headline(V, args.steps, args.warmup, not args.no_autotune, True, args.tuning_file,
                                  args.save_tuning)["out"]))
        return
    H = headline(V, args.steps, args.warmup, not args.no_autotune, False, args.tuning_file, args.save_tuning)
    h = H["out"]
    out = {
        "metric": "distinct CS/sec + instances/sec batched vs plain (bit-exact); p50/p99 solve latency",
        "value": h["value"],
        "unit": "distinct CS/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u64", "data": "synthetic",
        "config": {"workload": "config-4 batched mix: 512 instances (APs U[2,8], traces U[8,63] random +/- split, "
                               "len U[8,64], bits Bernoulli(p) p~U[0.2,0.6], planted formulas size 5-13), "
                               "one vfb_solve_batch per step, max_formula_size 16, autotuned launch parameters, "
                               "relaxed 128-bit hashed keys where total trace length > 126",
                   "instances": MIX_N, "max_formula_size": MIX_MAX, "parallelism": "single GPU"},
        "roofline": h["roofline"],
    }
    out.update({k: v for k, v in h.items() if k not in ("value", "ms_per_step", "roofline")})
    out["latency_ms_p50"] = h["plain_latency_ms_p50"]
    out["latency_ms_p99"] = h["plain_latency_ms_p99"]
    out["device"] = dev
    if args.only == "mix":
        print(json.dumps(out))
        return
    if not args.no_cpu:
        log("cpu baseline (OpenMP over the mix's instances)")
        out["cpu_baseline"] = cpu_baseline_mix(V, H)
    del H
    if args.config2:
        log("config 2")
        out["config2"] = config2(V, args.config2, 14, 3, 0 if args.no_cpu else 20.0)
    if args.limits:
        log("config 3 (limits)")
        out["limits"] = limits(V, [int(x) for x in args.limits.split(",")], 18)
    if args.multiword:
        log("multi-word traces")
        n, k, m = (int(x) for x in args.multiword.split(","))
        out["multiword"] = multiword(V, n, k, m)
    if args.serve_seconds:
        log("served stream")
        out["served"] = served(V, args.serve_rate, args.serve_seconds, MIX_MAX, args.serve_active, 256,
                               SERVE_REQUEST_BUDGET)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
