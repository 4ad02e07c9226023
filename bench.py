#!/usr/bin/env python3
"""Benchmark of the SIFT putative-matching hot path on one MI355X.

Workload: one image pair of 8192 x 8192 SIFT-like uint8 descriptors (40 %
noisy copies, sigma 0.05), rule minScore 0.8 / minRatio 0.8 / mutual check.
A step is one findPutativeMatches job on device-resident descriptors: the
fused int8-MFMA scan kernel plus the finalize kernel (ratio test + mutual
check) writing the per-left match table in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--kernel i8|plain]
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

INT8_PEAK_TOPS = 5000.0   # MI355X dense int8 MFMA, 5.0 POP/s
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector
HBM_PEAK_GBS = 8000.0


def cpu_baseline(L, R, min_score, min_ratio, budget_s=20.0):
    """The oracle's blocked multithreaded rule on the host cores (reported
    beside the GPU number; it is the checker, not the measured product)."""
    from oracle import oracle as O

    threads = len(os.sched_getaffinity(0))
    times = []
    tab = None
    t_start = time.perf_counter()
    while len(times) < 3 and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        _, _, tab, used = O.find_u8_blocked(L, R, min_score, min_ratio, threads=threads)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": L.shape[0] * R.shape[0] / med, "unit": "descriptor pairs scored/s", "cores": int(used),
            "kind": "port",
            "sample": f"the full {L.shape[0]}x{R.shape[0]} pair, oracle blocked rule (exact int32 dots, "
                      f"OpenMP over left rows), median of {len(times)} runs",
            "seconds_per_pair": med}, tab


# committed rocprofv3 memory-counter summaries, one per workload, per scan variant
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_i8_scan_8192x8192_traffic.md")
TRAFFIC_SUMMARY_BATCH = os.path.join(ROOT, "profiles", "k_i8_scan_batched_256pairs_traffic.md")
TRAFFIC_SUMMARY_65536 = os.path.join(ROOT, "profiles", "k_i8_scan_65536x65536_traffic.md")


def measured_traffic(kernel_name, summary=TRAFFIC_SUMMARY):
    """HBM bytes per launch of `kernel_name` on the workload `summary` covers, as
    that committed rocprofv3 counter summary states it (a line
    `traffic_bytes_per_launch <kernel> = <bytes>`), or None."""
    try:
        with open(summary) as f:
            for line in f:
                line = line.strip().strip("`")
                if line.startswith(f"traffic_bytes_per_launch {kernel_name} ="):
                    return float(line.split("=")[1])
    except OSError:
        return None
    return None


def traffic_child(args, v):
    """Counter-pass child of live_traffic: the workload's job with the scan
    variant the parent ran (c2: the 8192^2 pair, 20 jobs; c3: the 256-pair
    batch, 5 jobs; c5: the 65536^2 pair, 5 jobs), then the 256 MiB copy
    kernel whose counter readings scale the scan's."""
    if args.workload == "c3":
        n_pairs = args.pairs or 256
        rng = np.random.default_rng(2024)
        counts = rng.integers(500, 12001, size=(n_pairs, 2))
        images = []
        for i in range(n_pairs):
            d = v.synth_pair(v.synth_seed(3, i), int(counts[i, 0]), int(counts[i, 1]), 0.4, 0.05)
            images += [d["L"], d["R"]]
        m = v.SIFTMatching(0)
        m.enableRoundingWarning(False)
        m.setImages(images)
        m.setPairs([(2 * i, 2 * i + 1) for i in range(n_pairs)])
        m.batchSetGraph(False)
        m.batchSetVariant(args.scan_variant)
        for _ in range(5):
            m.batchMatchAsync(args.min_score, 0.8)
        m.batchFetch()
    else:
        n, seed, rho, sigma, jobs = ((65536, v.synth_seed(5, 0), 0.3, 0.08, 5) if args.workload == "c5"
                                     else (args.n, v.synth_seed(1, 0), 0.4, 0.05, 20))
        d = v.synth_pair(seed, n, n, rho, sigma)
        m = v.SIFTMatching(0)
        m.enableRoundingWarning(False)
        m.allocate(n, n)
        m.setLeftFeatures(n, d["L"])
        m.setRightFeatures(n, d["R"])
        m.setKernel(v.KERNEL_I8)
        m.setScanVariant(args.scan_variant)
        for _ in range(jobs):
            m.findAsync(0.8, 0.8)
        m.synchronize()
    m.close()
    copy_probe(256)


def _short_kernel(name):
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    if name.startswith("v3dm::"):
        name = name[6:]
    return name.split("(")[0]


def scaled_counter_bytes(csv_path, kernel_name):
    """From one rocprofv3 counter pass (counter_collection.csv): the median
    reading of `kernel_name`'s dispatches, in bytes, scaled by the median
    reading of the 256 MiB copy kernel of the same pass (which reads and
    writes exactly 262144 KiB per launch).  None if either is missing."""
    import csv
    scan, copy = [], []
    with open(csv_path) as f:
        for row in csv.DictReader(f):
            k = _short_kernel(row["Kernel_Name"])
            if k == kernel_name:
                scan.append(float(row["Counter_Value"]))
            elif k == "k_copy_probe":
                copy.append(float(row["Counter_Value"]))
    if not scan or not copy or statistics.median(copy) <= 0:
        return None
    return statistics.median(scan) * 262144.0 / statistics.median(copy) * 1024.0


def live_traffic(kernel_name, child_args, timeout_s=150):
    """HBM bytes per launch of `kernel_name` measured now: two rocprofv3
    counter passes (FETCH_SIZE, WRITE_SIZE) over a child process (bench.py
    --traffic-child + child_args) that repeats the workload's job with the
    same scan variant and then runs the 256 MiB copy kernel, each counter
    scaled by what it reads for that copy.  None if rocprofv3 or the child is
    unavailable or fails, or when this process already runs under a profiler."""
    if any(k.startswith("ROCPROF") for k in os.environ):
        return None
    import glob
    import shutil
    import subprocess
    import tempfile

    if shutil.which("rocprofv3") is None:
        return None
    per = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        tmp = tempfile.mkdtemp(prefix="v3dm_traffic_", dir="/tmp")
        try:
            cmd = ["rocprofv3", "--pmc", counter, "--kernel-include-regex", "k_i8_scan|k_copy_probe",
                   "--output-format", "csv", "-d", tmp, "-o", "t", "--", sys.executable,
                   os.path.join(ROOT, "bench.py"), "--traffic-child"] + [str(x) for x in child_args]
            r = subprocess.run(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), timeout=timeout_s,
                               stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            files = glob.glob(os.path.join(tmp, "**", "*counter_collection.csv"), recursive=True)
            if r.returncode != 0 or not files:
                return None
            b = scaled_counter_bytes(files[0], kernel_name)
            if b is None:
                return None
            per[counter] = b
        except (OSError, subprocess.SubprocessError, KeyError, ValueError):
            return None
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    return per["FETCH_SIZE"] + per["WRITE_SIZE"]


def _live_traffic_into(roof, kernel_name, child_args):
    """Replaces roof["traffic"] (the committed summary's figure, kept as
    traffic_summary) with the one measured in this run, when that works."""
    progress("traffic: two rocprofv3 counter passes")
    t = live_traffic(kernel_name, child_args)
    if t is not None:
        roof["traffic_summary"] = roof["traffic"]
        roof["traffic"] = t
        roof["traffic_source"] = ("rocprofv3 FETCH_SIZE and WRITE_SIZE passes in this run (the same job with the "
                                  "same scan variant, medians over its launches), scaled by a 256 MiB copy kernel "
                                  "in the same pass")


def copy_probe(mib, reps=5):
    """A plain device copy of known size (k_copy_probe: reads and writes
    exactly `mib` MiB per launch), run in the same counter pass as the matcher
    so FETCH_SIZE / WRITE_SIZE readings can be scaled to bytes."""
    import v3dsfmtoolkit_amd as v
    v.copy_probe(mib << 20, reps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--kernel", choices=["i8", "plain"], default="i8")
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this size (scales memory counters)")
    ap.add_argument("--workload", choices=["c2", "c3", "c4", "c5", "voctree", "sequence", "voctree-driver",
                                           "overflow", "mfma-ceiling"],
                    default="c2",
                    help="c2: single pair 8192^2 (default line); c3: 256 uneven pairs in one batch (+ float "
                         "variant); c4: learned path on 64 pairs of 16384^2; c5: single pair 65536^2; voctree: "
                         "vocabulary-tree quantisation of 64 views x 7168 features (depth 3, branch factor 50); "
                         "sequence: the match_sequence counterpart on 64 feature files, radius 3; voctree-driver: "
                         "the match_using_voctree counterpart on the same files; overflow: the survivor-overflow "
                         "cases (same-order near-duplicate frames, identical sets, ratio test off) against the "
                         "benchmark pair; mfma-ceiling: the int8 scan's step loop alone (no global loads, barriers "
                         "or survivors) on the benchmark pair's descriptor bytes and on other operand data")
    ap.add_argument("--pairs", type=int, default=0, help="number of pairs for c3 / c4 (0: the workload's own)")
    ap.add_argument("--min-score", type=float, default=0.8,
                    help="c3: the rule's minScore (above 1 no score survives the scan's filter: its time without "
                         "survivors)")
    ap.add_argument("--scan-variant", type=int, default=-1,
                    help="single-pair int8 scan variant (-1: the library measures the fastest)")
    ap.add_argument("--profile-variants", action="store_true",
                    help="after the measurement, run 20 jobs with each scan variant (for profiler passes)")
    ap.add_argument("--traffic", choices=["live", "summary"], default="live",
                    help="c2: HBM traffic of the scan measured in this run by two rocprofv3 counter passes "
                         "(live; the committed summary if they cannot run) or read from the summary only")
    ap.add_argument("--traffic-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-extra", action="store_true",
                    help="c2 only: skip the other workloads' summaries (batch, learned path, 65536^2)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(pairs shard across GPUs without communication; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import v3dsfmtoolkit_amd as v

    if args.traffic_child:
        traffic_child(args, v)
        return
    if args.workload != "c2":
        out = {"c3": run_c3, "c4": run_c4, "c5": run_c5, "voctree": run_voctree,
               "sequence": run_sequence, "voctree-driver": run_voctree_driver,
               "overflow": run_overflow, "mfma-ceiling": run_mfma_ceiling}[args.workload](args, v)
        if args.copy_probe_mib > 0:
            copy_probe(args.copy_probe_mib)
        print(json.dumps(out), flush=True)
        return
    run_c2(args, v)


def run_c2(args, v):
    n = args.n
    min_score, min_ratio = 0.8, 0.8
    d = v.synth_pair(v.synth_seed(1, 0), n, n, 0.4, 0.05)
    L, R = d["L"], d["R"]

    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    m.allocate(n, n)
    m.setLeftFeatures(n, L)
    m.setRightFeatures(n, R)
    kernel = v.KERNEL_I8 if args.kernel == "i8" else v.KERNEL_PLAIN
    m.setKernel(kernel)
    m.setScanVariant(args.scan_variant)

    # the measured kernel must agree with the reference-design kernel
    ref_list = None
    if args.plain_steps > 0:
        m.setKernel(v.KERNEL_PLAIN)
        ref_list = m.findPutativeMatches(min_score, min_ratio)
        m.synchronize()
        m.profileEnable(True)
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            m.findAsync(min_score, min_ratio)
        m.synchronize()
        plain_wall = (time.perf_counter() - t0) / args.plain_steps
        p_scan, p_sel, p_jobs = m.profileRead()
        m.profileEnable(False)
        m.setKernel(kernel)
    got = m.findPutativeMatches(min_score, min_ratio)
    if ref_list is not None and got != ref_list:
        print("bench.py: match lists differ between the measured kernel and the plain kernel", file=sys.stderr)
        sys.exit(1)

    for _ in range(args.warmup):
        m.findAsync(min_score, min_ratio)
    m.synchronize()

    # kernel times: HIP events bound to the scan dispatch of every job
    m.profileEnable(True)
    for _ in range(min(args.steps, 1000)):
        m.findAsync(min_score, min_ratio)
    m.synchronize()
    scan_ms, sel_ms, jobs = m.profileRead()
    m.profileEnable(False)
    # throughput: the same jobs with no events in the stream
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.findAsync(min_score, min_ratio)
    t_enq = time.perf_counter()
    m.synchronize()
    t1 = time.perf_counter()
    final = m.fetchMatches()
    assert final == got
    host_enqueue_us = (t_enq - t0) / args.steps * 1e6
    # what a reference caller pays per pair (match_sequence.cpp:184-203): host
    # descriptors uploaded, the job, the list back, every call blocking
    # (the C-ABI calls themselves, as a C++ caller makes them: no Python list building)
    import ctypes
    lib, ctx = m._lib, m._ctx
    pbuf = np.empty((n, 2), np.int32)
    cnt = ctypes.c_int(0)
    sync = {}
    for kind, (hl, hr) in {"u8": (L, R), "f32": (L.astype(np.float32) / 512, R.astype(np.float32) / 512)}.items():
        ct = ctypes.c_uint8 if kind == "u8" else ctypes.c_float
        setl = lib.v3dm_set_left_u8 if kind == "u8" else lib.v3dm_set_left_f32
        setr = lib.v3dm_set_right_u8 if kind == "u8" else lib.v3dm_set_right_f32
        pl, pr = hl.ctypes.data_as(ctypes.POINTER(ct)), hr.ctypes.data_as(ctypes.POINTER(ct))
        lat = []
        for _ in range(12):
            ts = time.perf_counter()
            rc = setl(ctx, n, pl) or setr(ctx, n, pr) or lib.v3dm_find_putative_matches(
                ctx, min_score, min_ratio, pbuf.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n,
                ctypes.byref(cnt))
            lat.append(time.perf_counter() - ts)
            assert rc == 0
        assert list(map(tuple, pbuf[:cnt.value].tolist())) == got
        sync[kind] = {"median_ms": statistics.median(lat[2:]) * 1e3, "min_ms": min(lat[2:]) * 1e3,
                      "upload_bytes": int(hl.nbytes + hr.nbytes)}
    sync["what"] = ("v3dm_set_left / v3dm_set_right from pageable host arrays + v3dm_find_putative_matches "
                    "(pairs into a host array), blocking, per pair: what a reference caller of "
                    "setLeftFeatures / setRightFeatures / findPutativeMatches pays; f32: the floats u/512 the "
                    "reference callers pass")

    wall = (t1 - t0) / args.steps
    pairs = n * n
    scan_avg_s = scan_ms / 1e3 / max(jobs, 1)
    sel_avg_s = sel_ms / 1e3 / max(jobs, 1)
    if kernel == v.KERNEL_I8:
        achieved = pairs * 256 / scan_avg_s / 1e12
        measured_variant = m.lastScanVariant()
        scan_kernel = _scan_name(v, measured_variant, m.lastScanInline() == 1)
        summary_traffic = measured_traffic(scan_kernel) if n == 8192 else None
        roof = {"bound": "mfma", "kernel": scan_kernel, "achieved": achieved,
                "peak": INT8_PEAK_TOPS,
                "unit": "TOP/s", "frac": achieved / INT8_PEAK_TOPS,
                "traffic": summary_traffic,
                "traffic_source": "committed rocprofv3 summary" if summary_traffic is not None else None,
                "ops_per_pair": 256, "launch_avg_us": scan_avg_s * 1e6, "peak_unit": "int8 dense MFMA, TOP/s",
                # the same work over the whole step (every kernel of the job, launch gaps included)
                "job_frac": pairs * 256 / wall / 1e12 / INT8_PEAK_TOPS}
        if m.lastScanInline() == 1:
            roof["note"] = ("this grid fits the device at once, so the scan kernel also applies its survivors "
                            "(top-2 insertion); no separate insertion kernel runs")
    else:
        achieved = pairs * 256 / scan_avg_s / 1e12
        roof = {"bound": "mfma", "kernel": "k_sgemm_seq", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                "ops_per_pair": 256, "launch_avg_us": scan_avg_s * 1e6}

    out = {
        "metric": "descriptor pairs scored/s",
        "value": pairs / wall,
        "unit": "descriptor pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8" if kernel == v.KERNEL_I8 else "f32",
        "data": "synthetic",
        "config": {"workload": f"single pair {n}x{n} uint8 SIFT-like descriptors, 40% noisy copies (sigma 0.05), "
                               "minScore 0.8, minRatio 0.8, mutual check",
                   "n_left": n, "n_right": n, "kernel": args.kernel, "matches": len(got),
                   "scan_variant": (list(v.scan_variants()[m.lastScanVariant()])
                                    if kernel == v.KERNEL_I8 else None)},
        "image_pairs_per_s": 1.0 / wall,
        "select_avg_us": sel_avg_s * 1e6,
        "host_enqueue_us_per_step": host_enqueue_us,
        "sync_find_putative_matches": sync,
        "roofline": roof,
    }
    if args.plain_steps > 0:
        out["plain_kernel"] = {"ms_per_step": plain_wall * 1e3, "pairs_per_s": pairs / plain_wall,
                               "sgemm_avg_us": p_scan / max(p_jobs, 1) * 1e3,
                               "select_avg_us": p_sel / max(p_jobs, 1) * 1e3,
                               "speedup_of_measured_kernel": plain_wall / wall}
    if not args.no_cpu_baseline:
        cb, tab = cpu_baseline(L, R, min_score, min_ratio)
        from oracle import oracle as O
        cb["agrees_with_gpu"] = O.table_to_pairs(tab) == got
        out["cpu_baseline"] = cb
    if args.profile_variants and kernel == v.KERNEL_I8:
        for var in range(len(v.scan_variants())):
            m.setScanVariant(var)
            for _ in range(20):
                m.findAsync(min_score, min_ratio)
            m.synchronize()
            assert m.fetchMatches() == got
    m.close()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
    if kernel == v.KERNEL_I8 and args.traffic == "live":
        _live_traffic_into(roof, roof["kernel"], ["--workload", "c2", "--n", n, "--scan-variant", measured_variant])
    if not args.no_extra and n == 8192 and kernel == v.KERNEL_I8:
        out["other_workloads"] = run_extra(args, v)
    print(json.dumps(out), flush=True)


def run_extra(args, v):
    """The other workloads of the metric, each summarised (full lines:
    --workload c3 / c4 / c5): the 256-pair batch (image pairs/s, scan roofline,
    float variant), the learned path on 64 pairs of 16384^2 (recall,
    precision, fallback, speed against the plain and the exact int8 path), and
    the 65536^2 pair (scan roofline, top-32 learned recall)."""
    sub = argparse.Namespace(**vars(args))
    sub.pairs = 0
    res = {}
    t = time.perf_counter()
    c3 = run_c3(sub, v)
    res["batch_256_pairs"] = {
        "image_pairs_per_s": c3["image_pairs_per_s"], "descriptor_pairs_per_s": c3["value"],
        "ms_per_job": c3["ms_per_step"], "roofline": c3["roofline"], "config": c3["config"],
        "float_variant": {k: c3["float_variant"][k] for k in ("image_pairs_per_s", "ms_per_job",
                                                               "sample_equals_fp32_oracle",
                                                               "sample_max_abs_score_error_vs_fp64",
                                                               "sample_list_disagreement_vs_fp64_rule",
                                                               "sample_fp64_rule_matches")},
        "sample_equals_oracle": True, "seconds": time.perf_counter() - t}
    progress("batch summary done")
    t = time.perf_counter()
    c4 = run_c4(sub, v)
    res["learned_64_pairs_16384"] = {k: c4[k] for k in ("recall", "precision", "exact_matches",
                                                        "fallback_fraction_left", "fallback_fraction_right",
                                                        "ms_per_pair", "speedup_vs_plain", "training")}
    res["learned_64_pairs_16384"]["speedup_vs_exact_int8"] = (c4["ms_per_pair"]["exact_int8"] /
                                                             c4["ms_per_pair"]["learned"])
    res["learned_64_pairs_16384"]["seconds"] = time.perf_counter() - t
    progress("learned summary done")
    t = time.perf_counter()
    c5 = run_c5(sub, v)
    res["pair_65536"] = {"descriptor_pairs_per_s": c5["value"], "ms_per_job": c5["ms_per_step"],
                         "roofline": c5["roofline"], "matches": c5["config"]["matches"],
                         "learned_top32": {k: c5["learned_top32"][k] for k in ("ms_per_pair", "recall", "precision",
                                                                                "fallback_fraction_left")},
                         "plain": c5["plain"], "seconds": time.perf_counter() - t}
    progress("65536^2 summary done")
    return res


def _scan_name(v, variant, inline=False):
    """Kernel name of a scan variant as rocprofv3 prints it, e.g. k_i8_scan<8, 512, true>
    or k_i8_scan_whole<true> (the last argument: the scan applies its survivors itself)."""
    flag = "true" if inline else "false"
    if v.scan_variant_whole(variant):
        return f"k_i8_scan_whole<{flag}>"
    w, _, tn = v.scan_variants()[variant]
    return f"k_i8_scan<{w}, {tn}, {flag}>"


def _common(metric, value, unit, steps, warmup, wall, dtype, config, higher=True):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": wall * 1e3, "higher_is_better": higher, "scaling": "weak", "vs_baseline": None,
            "dtype": dtype, "data": "synthetic", "config": config}


def run_c3(args, v):
    """256 image pairs, feature counts uniform in [500, 12000], one batched
    job (persistent tile queue, autotuned tile shape, hipGraph replay); then
    the float-descriptor variant (BF16 candidates + FP32 rescoring) on a
    sample of the pairs against the FP64 oracle."""
    from oracle import oracle as O

    n_pairs = args.pairs or 256
    ms = args.min_score
    rng = np.random.default_rng(2024)
    counts = rng.integers(500, 12001, size=(n_pairs, 2))
    images, fimages = [], []
    progress(f"generating {n_pairs} pairs")
    for i in range(n_pairs):
        d = v.synth_pair(v.synth_seed(3, i), int(counts[i, 0]), int(counts[i, 1]), 0.4, 0.05, f32=True)
        images += [d["L"], d["R"]]
        fimages += [d["Lf"], d["Rf"]]
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    c32 = m.setImages(images)
    m.setPairs([(2 * i, 2 * i + 1) for i in range(n_pairs)])
    pairs_scored = int(sum(c32[2 * i] * c32[2 * i + 1] for i in range(n_pairs)))
    best, tune_ms = m.batchAutotune(ms, 0.8, reps=2)
    m.batchSetVariant(best)
    # scan kernel time (events bound to the scan dispatch, eager jobs)
    m.batchSetGraph(False)
    for _ in range(3):
        m.batchMatchAsync(ms, 0.8)
    m.batchFetch()
    # median over jobs profiled one at a time: one slow job (a clock ramp)
    # does not move it
    per_job = []
    for _ in range(15):
        m.profileEnable(True)
        m.batchMatchAsync(ms, 0.8)
        m.batchFetch()
        s_ms, _, nj = m.profileRead()
        per_job.append(s_ms / max(nj, 1))
    m.profileEnable(False)
    scan_ms, jobs = float(np.median(per_job)), 1
    m.batchSetGraph(True)
    tabs = m.batchMatch(ms, 0.8)
    warm = max(args.warmup // 20, 3)
    for _ in range(warm):
        m.batchMatchAsync(ms, 0.8)
    m.batchFetch()
    steps = max(args.steps // 100, 10)
    t0 = time.perf_counter()
    for _ in range(steps):
        m.batchMatchAsync(ms, 0.8)
    tabs2 = m.batchFetch()
    wall = (time.perf_counter() - t0) / steps
    assert all(np.array_equal(a, b) for a, b in zip(tabs, tabs2))
    # parity and CPU baseline on a sample of pairs
    sample = list(range(4)) + list(rng.choice(np.arange(4, n_pairs), 4, replace=False))
    cpu_t, cpu_pairs, agree = 0.0, 0, True
    threads = len(os.sched_getaffinity(0))
    for i in sample:
        L, R = images[2 * i][:c32[2 * i]], images[2 * i + 1][:c32[2 * i + 1]]
        t = time.perf_counter()
        _, _, tab, used = O.find_u8_blocked(L, R, ms, 0.8, threads=threads)
        cpu_t += time.perf_counter() - t
        cpu_pairs += len(L) * len(R)
        agree &= np.array_equal(tab, tabs[i])
    if not agree:
        print("bench.py: batched tables differ from the oracle", file=sys.stderr)
        sys.exit(1)
    scan_avg = scan_ms / 1e3 / max(jobs, 1)
    achieved = pairs_scored * 256 / scan_avg / 1e12
    w, tm, tn = v.scan_variants()[best]
    out = _common("descriptor pairs scored/s", pairs_scored / wall, "descriptor pairs/s", steps, warm, wall, "int8",
                  {"workload": f"{n_pairs} image pairs, feature counts uniform in [500, 12000], 40% noisy copies "
                               f"(sigma 0.05), minScore {ms}, minRatio 0.8, mutual check; one batched job",
                   "pairs": n_pairs, "descriptor_pairs_per_job": pairs_scored, "variant": best,
                   "tile": [tm, tn], "waves": w, "graph": True})
    out["image_pairs_per_s"] = n_pairs / wall
    out["autotune_ms"] = [float(x) for x in tune_ms]
    scan_kernel = _scan_name(v, best, False)
    out["roofline"] = {"bound": "mfma", "kernel": scan_kernel + " (persistent, batched)",
                       "achieved": achieved,
                       "peak": INT8_PEAK_TOPS, "unit": "TOP/s", "frac": achieved / INT8_PEAK_TOPS,
                       "traffic": (measured_traffic(scan_kernel, TRAFFIC_SUMMARY_BATCH)
                                   if n_pairs == 256 else None),
                       "traffic_source": "committed rocprofv3 summary" if n_pairs == 256 else None,
                       "launch_avg_us": scan_avg * 1e6}
    if args.traffic == "live":
        _live_traffic_into(out["roofline"], scan_kernel,
                           ["--workload", "c3", "--pairs", n_pairs, "--scan-variant", best, "--min-score", ms])
    out["cpu_baseline"] = {"value": cpu_pairs / cpu_t, "unit": "descriptor pairs scored/s", "cores": threads,
                           "kind": "port", "sample": f"{len(sample)} of the pairs ({cpu_pairs} descriptor pairs), "
                                                     "oracle blocked rule; their tables equal the GPU's"}
    m.close()
    progress("batched int8 done; float variant")
    # float variant: the same pairs as unquantised unit floats, one batched job
    # (BF16 candidates + FP32 rescoring); sampled pairs against the FP32 and
    # FP64 oracles
    fm = v.SIFTMatching(0)
    fm.enableRoundingWarning(False)
    fc32 = fm.setImages(fimages)
    fm.setPairs([(2 * i, 2 * i + 1) for i in range(n_pairs)])
    fbest, _ = fm.batchAutotune(ms, 0.8, reps=2)
    fm.batchSetVariant(fbest)
    fm.batchSetGraph(True)
    ftabs = fm.batchMatch(ms, 0.8)
    for _ in range(3):
        fm.batchMatchAsync(ms, 0.8)
    fm.batchFetch()
    t0 = time.perf_counter()
    fsteps = max(steps // 2, 5)
    for _ in range(fsteps):
        fm.batchMatchAsync(ms, 0.8)
    fm.batchFetch()
    f_wall = (time.perf_counter() - t0) / fsteps
    fm.close()
    exact_f32, disagree, matches, max_err = True, 0, 0, 0.0
    sm = v.SIFTMatching(0)
    sm.enableRoundingWarning(False)
    sm.allocate(12000, 12000)
    sm.setKernel(v.KERNEL_BF16)
    for i in sample:
        Lf, Rf = fimages[2 * i][:fc32[2 * i]], fimages[2 * i + 1][:fc32[2 * i + 1]]
        exact_f32 &= np.array_equal(ftabs[i], O.find_f32(Lf, Rf, ms, 0.8)[2])
        got = O.table_to_pairs(ftabs[i])
        l64 = set(O.table_to_pairs(O.find_f64(Lf, Rf, ms, 0.8)[2]))
        disagree += len(set(got) ^ l64)
        matches += len(l64)
        # the same rule's rescored scores (single-pair path) against FP64
        sm.setLeftFeatures(len(Lf), Lf)
        sm.setRightFeatures(len(Rf), Rf)
        sm.findPutativeMatches(ms, 0.8)
        bsc, _, arg = sm.leftTop2()
        have = arg >= 0
        if have.any():
            s64 = np.einsum("ij,ij->i", Lf[have].astype(np.float64), Rf[arg[have]].astype(np.float64))
            max_err = max(max_err, float(np.abs(bsc[have] - s64).max()))
    sm.close()
    if not exact_f32:
        print("bench.py: float batch tables differ from the FP32 oracle", file=sys.stderr)
        sys.exit(1)
    fw, _, ftn = v.scan_variants()[fbest]
    out["float_variant"] = {"pairs": n_pairs, "descriptor_pairs_per_s": pairs_scored / f_wall,
                            "image_pairs_per_s": n_pairs / f_wall, "ms_per_job": f_wall * 1e3,
                            "variant": [fw, ftn], "graph": True,
                            "sample_pairs": len(sample), "sample_equals_fp32_oracle": bool(exact_f32),
                            "sample_max_abs_score_error_vs_fp64": max_err,
                            "sample_list_disagreement_vs_fp64_rule": disagree,
                            "sample_fp64_rule_matches": matches,
                            "method": "one batched job: BF16 MFMA candidates within the proven bound, FP32 "
                                      "rescoring, float-key finalize"}
    return out


def progress(*a):
    print("bench.py:", *a, file=sys.stderr, flush=True)


def _train_net(v, pairs, rho, sigma, steps, seed=5):
    from v3dsfmtoolkit_amd import learned as VL
    t0 = time.perf_counter()
    L, R = VL.labeled_pairs(pairs, n=16384, rho=rho, sigma=sigma)
    t1 = time.perf_counter()
    progress(f"labelled {len(L)} pairs in {t1 - t0:.1f} s")
    net = v.EmbeddingNet(0, seed=seed)
    losses = VL.train(net, L, R, steps=steps, batch=1024, lr=2e-3)
    t2 = time.perf_counter()
    progress(f"trained {steps} steps in {t2 - t1:.1f} s, loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    return net, {"pairs": len(L), "steps": steps, "batch": 1024, "label_s": t1 - t0, "train_s": t2 - t1,
                 "loss_first": float(losses[0]), "loss_last": float(np.mean(losses[-20:]))}


def _time_job(fn, reps):
    fn()
    t = time.perf_counter()
    for _ in range(reps):
        r = fn()
    return (time.perf_counter() - t) / reps, r


def _pairset(a):
    return set(map(tuple, np.asarray(a).reshape(-1, 2).tolist()))


def run_c4(args, v):
    """Learned proposal path on 64 pairs of 16384 x 16384 (network trained
    on 2M exact-labelled pairs), beside the exact int8 matcher (the judge of
    recall) and the plain path."""
    n_pairs = args.pairs or 64
    n = 16384
    net, train = _train_net(v, 2_000_000, 0.4, 0.05, steps=2000)
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    m.allocate(n, n)
    tot = {"learned": 0.0, "i8": 0.0, "plain": 0.0, "plain_gpu": 0.0}
    hit = n_exact = n_learned = 0
    flagged = [0, 0]
    fb_pairs = 0
    for i in range(n_pairs):
        d = v.synth_pair(v.synth_seed(4, i), n, n, 0.4, 0.05)
        if i % 8 == 0:
            progress(f"pair {i} of {n_pairs}")
        m.setLeftFeatures(n, d["L"])
        m.setRightFeatures(n, d["R"])
        m.setKernel(v.KERNEL_I8)
        exact = m.findPutativeMatches(0.8, 0.8)
        dt, _ = _time_job(lambda: (m.findAsync(0.8, 0.8), m.synchronize()), 5)
        tot["i8"] += dt
        dt, (got, st) = _time_job(lambda: m.learnedFind(net, 0.05, 0.8, 0.8, 16, asArray=True), 3)
        tot["learned"] += dt
        g, e = _pairset(got), set(exact)
        hit += len(g & e)
        n_exact += len(e)
        n_learned += len(g)
        flagged[0] += st["flagged_left"]
        flagged[1] += st["flagged_right"]
        fb_pairs += st["fallback_pairs"]
        if i < 8:  # the plain path is slow: time it on the first pairs
            m.setKernel(v.KERNEL_PLAIN)
            m.profileEnable(True)
            dt, _ = _time_job(lambda: (m.findAsync(0.8, 0.8), m.synchronize()), 2)
            sm, sl, jobs = m.profileRead()
            m.profileEnable(False)
            tot["plain"] += dt
            tot["plain_gpu"] += (sm + sl) / 1e3 / max(jobs, 1)
            if m.fetchMatches() != exact:
                print("bench.py: plain and int8 lists differ", file=sys.stderr)
                sys.exit(1)
    m.close()
    net.close()
    plain_pp = tot["plain"] / min(n_pairs, 8)
    learned_pp = tot["learned"] / n_pairs
    rec, prec = hit / max(n_exact, 1), hit / max(n_learned, 1)
    out = _common("effective descriptor pairs/s (learned path)", n * n / learned_pp, "descriptor pairs/s",
                  n_pairs, 1, learned_pp, "bf16 retrieval + int8 rescoring",
                  {"workload": f"{n_pairs} pairs of {n}x{n} uint8 descriptors, 40% noisy copies (sigma 0.05), "
                               "minScore 0.8, minRatio 0.8, mutual; top-16 retrieval, margin 0.05 fallback",
                   "pairs": n_pairs, "n": n, "top_k": 16, "margin": 0.05})
    out.update({"recall": rec, "precision": prec, "exact_matches": n_exact,
                "fallback_fraction_left": flagged[0] / (n_pairs * n),
                "fallback_fraction_right": flagged[1] / (n_pairs * n),
                "fallback_pairs_per_pair": fb_pairs / n_pairs,
                "ms_per_pair": {"learned": learned_pp * 1e3, "exact_int8": tot["i8"] / n_pairs * 1e3,
                                "plain": plain_pp * 1e3, "plain_gpu_events": tot["plain_gpu"] / min(n_pairs, 8) * 1e3},
                "speedup_vs_plain": plain_pp / learned_pp, "training": train,
                "timing": "wall clock per job with inputs resident; the learned job includes its one host "
                          "round trip (flag counts)"})
    return out


def run_c5(args, v):
    """One pair of 65536 x 65536 (30% overlap, sigma 0.08): exact int8 beside
    the learned path with top-32 candidates (network trained on pairs of the
    same generator settings)."""
    from oracle import oracle as O
    n = 65536
    d = v.synth_pair(v.synth_seed(5, 0), n, n, 0.3, 0.08)
    m = v.SIFTMatching(0)
    m.enableRoundingWarning(False)
    m.allocate(n, n)
    m.setLeftFeatures(n, d["L"])
    m.setRightFeatures(n, d["R"])
    m.setKernel(v.KERNEL_I8)
    exact = m.findPutativeMatches(0.8, 0.8)
    progress(f"exact int8 list: {len(exact)} matches")
    m.profileEnable(True)
    dt_i8, _ = _time_job(lambda: (m.findAsync(0.8, 0.8), m.synchronize()), 10)
    scan_ms, sel_ms, jobs = m.profileRead()
    m.profileEnable(False)
    scan_var, scan_inline = m.lastScanVariant(), m.lastScanInline() == 1
    scan_avg = scan_ms / 1e3 / max(jobs, 1)
    net, train = _train_net(v, 1_000_000, 0.3, 0.08, steps=1500)
    dt_l, (got, st) = _time_job(lambda: m.learnedFind(net, 0.05, 0.8, 0.8, 32, asArray=True), 3)
    g, e = _pairset(got), set(exact)
    hit = len(g & e)
    m.setKernel(v.KERNEL_PLAIN)
    dt_p, _ = _time_job(lambda: (m.findAsync(0.8, 0.8), m.synchronize()), 1)
    plain_same = m.fetchMatches() == exact
    m.close()
    net.close()
    # CPU baseline: the oracle's blocked rule on a 2048-row slice of the left set
    threads = len(os.sched_getaffinity(0))
    t = time.perf_counter()
    O.find_u8_blocked(d["L"][:2048], d["R"], 0.8, 0.8, threads=threads)
    cpu = 2048 * n / (time.perf_counter() - t)
    achieved = n * n * 256 / scan_avg / 1e12
    out = _common("descriptor pairs scored/s", n * n / dt_i8, "descriptor pairs/s", 10, 1, dt_i8, "int8",
                  {"workload": f"single pair {n}x{n} uint8, 30% noisy copies (sigma 0.08), minScore 0.8, "
                               "minRatio 0.8, mutual", "n_left": n, "n_right": n, "matches": len(exact)})
    scan_kernel = _scan_name(v, scan_var, scan_inline)
    out["roofline"] = {"bound": "mfma", "kernel": scan_kernel, "achieved": achieved,
                       "peak": INT8_PEAK_TOPS,
                       "unit": "TOP/s", "frac": achieved / INT8_PEAK_TOPS,
                       "traffic": measured_traffic(scan_kernel, TRAFFIC_SUMMARY_65536),
                       "traffic_source": "committed rocprofv3 summary",
                       "launch_avg_us": scan_avg * 1e6}
    if args.traffic == "live":
        _live_traffic_into(out["roofline"], scan_kernel, ["--workload", "c5", "--scan-variant", scan_var])
    out["learned_top32"] = {"ms_per_pair": dt_l * 1e3, "effective_pairs_per_s": n * n / dt_l,
                            "recall": hit / max(len(e), 1), "precision": hit / max(len(g), 1),
                            "fallback_fraction_left": st["flagged_left"] / n,
                            "fallback_fraction_right": st["flagged_right"] / n, "training": train}
    out["plain"] = {"ms_per_pair": dt_p * 1e3, "equals_int8": plain_same}
    out["cpu_baseline"] = {"value": cpu, "unit": "descriptor pairs scored/s", "cores": threads, "kind": "port",
                           "sample": "2048 left rows against all 65536 rights, oracle blocked rule"}
    return out


def _survivor_threshold(min_score, min_ratio):
    """The scan's integer survivor threshold (scores in units of 2^-18): a score at or
    below it cannot change a decision; -1 keeps every score (a tiny positive minRatio),
    2^28 none (a NaN threshold accepts nothing)."""
    ms, mr = float(np.float32(min_score)), float(np.float32(min_ratio))
    if mr < 0:
        t = ms
    elif mr > 0:
        t = min(ms, 1.0 - (1.0 - ms) / (mr * (1.0 - 1e-5)))
    elif mr == 0:
        t = max(ms, 1.0)  # only bests above 1 give a negative ratio
    else:
        return 1 << 28  # NaN: nothing is accepted
    if t != t:
        return 1 << 28
    return min(max(-1, int(np.floor(t * 262144.0))), 1 << 28)


def _region_stats(L, R, thr, waves, tile_n):
    """Survivors (exact integer scores > thr, from float32 products of the uint8 rows,
    exact below 2^24) and how many of the scan's (tile, wave) regions -- 64 left rows x
    tile_n right rows, capacity 64 per 512 columns -- overflow."""
    cap = 64 * max(1, tile_n // 512)
    Rt = R.astype(np.float32).T
    surv, over, regions = 0, 0, 0
    for r0 in range(0, len(L), 2048):
        S = L[r0:r0 + 2048].astype(np.float32) @ Rt
        hit = S > thr
        surv += int(hit.sum())
        nb = hit.shape[0] // 64
        for c0 in range(0, hit.shape[1], tile_n):
            cnt = hit[:nb * 64, c0:c0 + tile_n].reshape(nb, 64, -1).sum(axis=(1, 2))
            over += int((cnt > cap).sum())
            regions += nb
    return surv, over, regions, cap


def run_overflow(args, v):
    """The survivor-overflow cliff: pairs whose scores pass the survivor filter far more
    often than the benchmark pair's, against that pair (8192 x 8192, 40% noisy copies).
    A (tile, wave) region holds 64 survivors per 512 right rows; a region that overflows
    is recomputed exactly from the descriptors (v_dot4 over its 64 x tile-width block),
    by the scan wave itself when the grid fits the device, else by k_i8_insert.
    Cases: consecutive frames whose features come in the same order (every right row a
    noisy copy of the left row with its index: one survivor per row, all on the diagonal
    of the tile grid); identical sets (L = R, same order); minRatio 0 (only bests above 1
    can pass, so the filter is max(minScore, 1)); a tiny ratio (minScore 0.9, minRatio
    0.05: no filter is provable, every score survives -- the int8 scan forced, beside the
    plain path the automatic choice takes); and the same-order copies at 16384 x 16384.
    Each case also reports the plain path's time and the automatic kernel choice."""
    def case(name, L, R, min_score, min_ratio, steps):
        m = v.SIFTMatching(0)
        m.enableRoundingWarning(False)
        m.allocate(len(L), len(R))
        m.setLeftFeatures(len(L), L)
        m.setRightFeatures(len(R), R)
        m.setKernel(v.KERNEL_I8)
        got = m.findPutativeMatches(min_score, min_ratio)
        for _ in range(3):
            m.findAsync(min_score, min_ratio)
        m.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            m.findAsync(min_score, min_ratio)
        m.synchronize()
        wall = (time.perf_counter() - t0) / steps
        m.profileEnable(True)  # kernel times: the same jobs again, with events bound to the scan
        for _ in range(steps):
            m.findAsync(min_score, min_ratio)
        m.synchronize()
        scan_ms, sel_ms, jobs = m.profileRead()
        m.profileEnable(False)
        variant, inline = m.lastScanVariant(), m.lastScanInline() == 1
        w, _, tn = v.scan_variants()[variant]
        m.setKernel(v.KERNEL_PLAIN)
        same = m.findPutativeMatches(min_score, min_ratio) == got
        t0 = time.perf_counter()
        for _ in range(3):
            m.findAsync(min_score, min_ratio)
        m.synchronize()
        plain = (time.perf_counter() - t0) / 3
        m.setKernel(v.KERNEL_AUTO)
        auto_same = m.findPutativeMatches(min_score, min_ratio) == got
        auto_kernel = {v.KERNEL_I8: "int8", v.KERNEL_PLAIN: "plain", v.KERNEL_BF16: "bf16"}[m.lastKernel()]
        m.close()
        surv, over, regions, cap = _region_stats(L, R, _survivor_threshold(min_score, min_ratio), w, tn)
        progress(f"overflow case {name}: {wall * 1e3:.4f} ms/job, {over} of {regions} regions overflow")
        return {"n": [len(L), len(R)], "rule": [min_score, min_ratio], "ms_per_job": wall * 1e3,
                "scan_us": scan_ms / max(jobs, 1) * 1e3, "select_us": sel_ms / max(jobs, 1) * 1e3,
                "scan_kernel": _scan_name(v, variant, inline),
                "insertion": "in the scan" if inline else "k_i8_insert", "survivors": surv,
                "survivors_per_left_row": surv / len(L), "regions": regions, "regions_overflowed": over,
                "region_capacity": cap, "matches": len(got), "equals_plain_kernel": same,
                "plain_ms_per_job": plain * 1e3, "automatic_choice": auto_kernel,
                "automatic_choice_equal": auto_same}

    n = 8192
    base = v.synth_pair(v.synth_seed(1, 0), n, n, 0.4, 0.05)
    dup = v.synth_pair(v.synth_seed(7, 0), n, n, 1.0, 0.05)
    ordered = np.empty_like(dup["R"])
    ordered[dup["src"]] = dup["R"]  # right row i: the noisy copy of left row i
    res = {"c2": case("c2", base["L"], base["R"], 0.8, 0.8, args.steps),
           "same_order_copies": case("same_order_copies", dup["L"], ordered, 0.8, 0.8, args.steps),
           "identical": case("identical", base["L"], base["L"], 0.8, 0.8, args.steps),
           "min_ratio_0": case("min_ratio_0", base["L"], base["R"], 0.8, 0.0, args.steps),
           "tiny_ratio_no_filter": case("tiny_ratio_no_filter", base["L"], base["R"], 0.9, 0.05,
                                        max(args.steps // 20, 3))}
    big = v.synth_pair(v.synth_seed(7, 1), 16384, 16384, 1.0, 0.05)
    ordered = np.empty_like(big["R"])
    ordered[big["src"]] = big["R"]
    res["same_order_copies_16384"] = case("same_order_copies_16384", big["L"], ordered, 0.8, 0.8,
                                          max(args.steps // 4, 3))
    big0 = v.synth_pair(v.synth_seed(7, 2), 16384, 16384, 0.4, 0.05)
    res["c2_data_16384"] = case("c2_data_16384", big0["L"], big0["R"], 0.8, 0.8, max(args.steps // 4, 3))
    for k, r in res.items():
        ref = res["c2_data_16384"] if k.endswith("16384") else res["c2"]
        r["slowdown_vs_same_size_benchmark_data"] = r["ms_per_job"] / ref["ms_per_job"]
    c = res["c2"]
    return _common("descriptor pairs scored/s", n * n / (c["ms_per_job"] / 1e3), "descriptor pairs/s",
                   args.steps, 3, c["ms_per_job"] / 1e3, "int8",
                   {"workload": "survivor-overflow cases at 8192 x 8192 and 16384 x 16384 (see cases)"}) | \
        {"cases": res}


def run_mfma_ceiling(args, v):
    """The practical ceiling of the int8 scan's inner loop: its step (8 MFMAs on
    64 rows per wave, the max / compare test, the column term and the four
    right fragments from LDS) run alone, two waves per SIMD, on 128 stored
    rows of the benchmark pair -- and, for comparison, with the fragments kept
    in registers and on other operand bytes.  The int8 MFMA's clock under its
    power limit follows the operand values, so the rate depends on the data."""
    d = v.synth_pair(v.synth_seed(1, 0), 8192, 8192, 0.4, 0.05)
    rng = np.random.default_rng(7)
    u = d["L"][:128]
    data = {
        "benchmark pair rows (stored u ^ 0x80)": u ^ 0x80,
        "same rows stored without offset (u; exact only when every u < 128)": np.minimum(u, 127),
        "uniform random bytes": rng.integers(0, 256, (128, 128), dtype=np.uint8),
        "constant bytes (u = 5, stored 0x85)": np.full((128, 128), 0x85, np.uint8),
    }
    iters = max(2, args.steps) if args.steps != 2000 else 20000
    variants = {}
    for name, rows in data.items():
        for from_lds in (True, False):
            rates = [v.mfma_ceiling_probe(rows, from_lds, iters) for _ in range(3)]
            variants[f"{name}; fragments {'from LDS' if from_lds else 'in registers'}"] = {
                "tops_median": statistics.median(rates), "tops": rates,
                "frac_of_5000": statistics.median(rates) / 5000.0}
    head = variants["benchmark pair rows (stored u ^ 0x80); fragments from LDS"]
    return {"metric": "int8 MFMA rate of the scan's step loop alone", "value": head["tops_median"],
            "unit": "TOP/s", "n_gpus": 1, "steps": iters, "warmup": 1, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int8", "data": "synthetic",
            "config": {"workload": "scan step loop (8 v_mfma_i32_32x32x32_i8 per step, 64 rows per wave in "
                                   "VGPRs, right fragments from a swizzled LDS stage, max / compare test), "
                                   "two 4-wave workgroups per CU, operands: 128 rows of the 8192x8192 "
                                   "benchmark pair"},
            "frac_of_5000": head["frac_of_5000"], "variants": variants}


def run_voctree(args, v):
    """Vocabulary-tree quantisation (the voctree driver's first stage) of 64
    views x 7168 SIFT-like unit float descriptors on a depth-3, branch-factor-50
    tree, keys resident in HBM.  A step is one full quantisation of all keys:
    per level the regrouping kernels and the scoring kernel."""
    from oracle import voctree as OV
    from v3dsfmtoolkit_amd import voctree as VT
    depth, bf, per = 3, 50, 7168
    views = args.pairs or 64
    tree = VT.synthetic_tree(1, depth, bf)
    keys = np.concatenate([v.synth_pair(v.synth_seed(7, i), per, 0, u8=False, f32=True)["Lf"] for i in range(views)])
    keys = np.ascontiguousarray(keys, np.float32)
    n = len(keys)
    t = VT.VocabularyTree(depth, bf, tree)
    t.upload(keys)
    steps, warmup = min(args.steps, 200), min(args.warmup, 20)
    for _ in range(max(warmup, 1)):
        t.run()
    words = t.fetch()
    sample = min(n, 50000)
    same = bool(np.array_equal(words[:sample], OV.quantise(depth, bf, tree, keys[:sample])))
    progress(f"{n} keys quantised; first {sample} equal the oracle: {same}")
    t0 = time.perf_counter()
    for _ in range(steps):
        t.run()
    t.runTimes()  # waits for the stream
    wall = (time.perf_counter() - t0) / steps
    scoring, runs = [], []
    for _ in range(20):
        t.run()
        a, b = t.runTimes()
        scoring.append(a)
        runs.append(b)
    t.setMode(VT.VocabularyTree.EXACT_VALU)  # ablation: every child with the sequential VALU sum
    exact_runs = []
    for _ in range(10):
        t.run()
        exact_runs.append(t.runTimes()[1])
    exact_same = bool(np.array_equal(t.fetch(), words))
    t.setMode(VT.VocabularyTree.AUTO)
    host = []
    for _ in range(3):  # the host-buffer entry point: key upload over PCIe + run + word download
        th = time.perf_counter()
        t.getVisualWords(keys)
        host.append(time.perf_counter() - th)
    t.close()
    level_avg_s = statistics.mean(scoring) / 1e3 / depth
    ops_per_launch = n * bf * 128 * 2  # one multiply and one add per term, per level
    achieved = ops_per_launch / level_avg_s / 1e12
    threads = len(os.sched_getaffinity(0))
    cpu_n = min(n, 100000)
    tc = time.perf_counter()
    OV.quantise(depth, bf, tree, keys[:cpu_n])
    cpu = cpu_n / (time.perf_counter() - tc)
    out = _common("visual words/s", n / wall, "features quantised/s", steps, warmup, wall, "f32",
                  {"workload": f"{views} views x {per} SIFT-like unit float descriptors, vocabulary tree depth "
                               f"{depth}, branch factor {bf} (synthetic unit-norm nodes)", "features": n,
                   "tree_nodes": VT.tree_nodes(depth, bf)})
    out["roofline"] = {"bound": "mfma", "kernel": "k_vt_level_mfma", "achieved": achieved,
                       "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                       "traffic": None, "launch_avg_us": level_avg_s * 1e6,
                       "note": "FP32-MFMA filter over all 50 children (64 x 64 x 128 per wave, 14 children of "
                               "padding counted as waste, not work) plus exact sequential rescoring of the "
                               "children within the proven error bound"}
    out["run_ms"] = statistics.mean(runs)
    out["exact_valu_run_ms"] = statistics.mean(exact_runs)
    out["exact_valu_equals"] = exact_same
    out["pcie_inclusive_ms"] = statistics.median(host) * 1e3
    out["equals_oracle_sample"] = same
    out["cpu_baseline"] = {"value": cpu, "unit": "features quantised/s", "cores": threads, "kind": "port",
                           "sample": f"the first {cpu_n} keys, oracle sequential descent (OpenMP over keys)"}
    return out


def run_sequence(args, v):
    """The putative-matching stage of a whole sequence run through the native
    counterpart driver (v3dm_match_sequence): 64 views of 7168 SIFT-like
    features in V3DBIN files, each view sharing 40 % of its features (noisy
    copies, sigma 0.02) with the previous one, radius 3 (each view against the
    next five), MATCHING_MAX_FEATURES 7168.  A step is one application run:
    read every file, upload every view once, match all pairs as one batched
    job, write the SQLite rows."""
    from v3dsfmtoolkit_amd import drivers as D
    views, per, radius = args.pairs or 64, 7168, 3
    tmp, descs = _sequence_files(v, views, per)
    app = os.path.join(ROOT, "v3dsfmtoolkit_amd", "bin", "v3dm_match_sequence")
    return _run_driver(v, args, tmp, descs, [app, "--quiet", "--db", os.path.join(tmp, "out.db"), "list.txt",
                                             "run.conf", str(radius)], D.sequence_pairs(views, radius).tolist(),
                       f"match_sequence counterpart: {views} views x {per} SIFT-like features in V3DBIN files, 40% "
                       f"shared with the previous view (sigma 0.02), radius {radius}, minScore 0.9, minRatio 0.8")


def _sequence_files(v, views, per):
    """Synthetic run: views of SIFT-like u/512 features, 40 % shared with the previous view."""
    import tempfile
    from v3dsfmtoolkit_amd import features as F
    rng = np.random.default_rng(77)
    tmp = tempfile.mkdtemp(prefix="v3dm_seq_")
    progress(f"writing {views} feature files to {tmp}")
    prev = None
    descs = []
    for i in range(views):
        x = v.synth_pair(v.synth_seed(8, i), per, 0, u8=False, f32=True)["Lf"].astype(np.float64)
        if prev is not None:
            k = int(0.4 * per)
            dst, src = rng.choice(per, k, replace=False), rng.choice(per, k, replace=False)
            w = np.maximum(prev[src] + rng.normal(0, 0.02, (k, 128)), 0)
            w /= np.linalg.norm(w, axis=1, keepdims=True)
            w = np.minimum(w, 0.2)
            x[dst] = w / np.linalg.norm(w, axis=1, keepdims=True)
        prev = x
        q = (np.clip(np.rint(x * 512), 0, 255) / 512).astype(np.float32)
        descs.append(q)
        F.write_features(os.path.join(tmp, f"v{i:04d}.features"), q, ids=np.arange(per, dtype=np.int32),
                         positions=rng.uniform(0, 2000, (per, 2)).astype(np.float32))
    with open(os.path.join(tmp, "list.txt"), "w") as f:
        f.write("".join(f"v{i:04d}\n" for i in range(views)))
    with open(os.path.join(tmp, "run.conf"), "w") as f:
        f.write("MATCHING_MAX_FEATURES 7168\nMATCHING_SCORE_RATIO 0.8\nMINIMUM_RAW_CORRESPONDENCES 100\n"
                "POTENTIAL_VIEWS_COUNT 5\nVISUAL_WORD_MATCH_PROPABILITY 0.1\n")
    return tmp, descs


def _run_driver(v, args, tmp, descs, cmd, pairs, workload):
    """Runs a counterpart driver application steps + 1 times (the first warms caches)
    and reports its matching stage, its whole run and the oracle on four pairs."""
    import re
    import subprocess
    from oracle import oracle as O
    from v3dsfmtoolkit_amd import drivers as D
    views, per = len(descs), len(descs[0])
    steps = max(1, min(args.steps, 5))
    runs = []
    for it in range(steps + 1):
        t0 = time.perf_counter()
        r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True, timeout=300)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(r.stderr)
        m = re.search(r"matched (\d+) view pairs \((\d+) descriptor pairs, (\w+) pool\) in ([\d.]+) ms after a "
                      r"([\d.]+) ms upload; (\d+) putative matches; (\d+) pairs", r.stdout)
        q = re.search(r"quantised (\d+) features of (\d+) views in ([\d.]+) ms; (\d+) candidate pairs", r.stdout)
        if it > 0:  # the first run warms the page cache and the GPU
            runs.append((wall, float(m.group(4)), float(m.group(5)), int(m.group(6)), int(m.group(7)),
                         int(m.group(2)), m.group(3), float(q.group(3)) if q else None, int(q.group(4)) if q else None))
        progress(f"run {it}: {wall:.3f} s wall, matching {m.group(4)} ms")
    wall = statistics.median(x[0] for x in runs)
    match_ms = statistics.median(x[1] for x in runs)
    upload_ms = statistics.median(x[2] for x in runs)
    n_pairs, scored, pool = len(pairs), runs[0][5], runs[0][6]
    if runs[0][8] is not None:
        assert runs[0][8] == n_pairs, (runs[0][8], n_pairs)
    # CPU baseline and parity: the oracle's blocked rule on four of the pairs
    threads = len(os.sched_getaffinity(0))
    sample = [0, 1, len(pairs) // 2, len(pairs) - 1]
    t = time.perf_counter()
    sample_matches = 0
    for p in sample:
        a, b = pairs[p]
        _, _, tab, _ = O.find_u8_blocked(D.descriptors_to_u8(descs[a]), D.descriptors_to_u8(descs[b]), 0.9, 0.8,
                                         threads=threads)
        sample_matches += int((tab >= 0).sum())
    cpu_s = (time.perf_counter() - t) / len(sample)
    out = _common("image pairs matched/s", n_pairs / (match_ms / 1e3), "image pairs/s", len(runs), 1,
                  match_ms / 1e3, "int8",
                  {"workload": workload, "views": views, "view_pairs": n_pairs, "descriptor_pairs": scored,
                   "pool": pool})
    if runs[0][7] is not None:
        out["quantise_ms"] = statistics.median(x[7] for x in runs)
    out["descriptor_pairs_per_s"] = scored / (match_ms / 1e3)
    out["app_wall_s"] = wall
    out["app_image_pairs_per_s"] = n_pairs / wall
    out["upload_ms"] = upload_ms
    out["putative_matches"] = runs[0][3]
    out["stored_pairs"] = runs[0][4]
    out["cpu_baseline"] = {"value": 1.0 / cpu_s, "unit": "image pairs/s", "cores": threads, "kind": "port",
                           "sample": f"{len(sample)} of the view pairs, oracle blocked rule (exact int32 dots)",
                           "sample_matches": sample_matches}
    return out


def run_voctree_driver(args, v):
    """The match_using_voctree counterpart (v3dm_match_using_voctree) on the same
    64 synthetic feature files: GPU quantisation of every feature on a
    depth-3, branch-50 tree (synthetic nodes), binomial candidate choice
    (POTENTIAL_VIEWS_COUNT 5), then the batched matching stage."""
    from oracle import voctree as OV
    from v3dsfmtoolkit_amd import voctree as VT
    views, per = args.pairs or 64, 7168
    tmp, descs = _sequence_files(v, views, per)
    tree = VT.synthetic_tree(1, 3, 50)
    tree.tofile(os.path.join(tmp, "tree.bin"))
    words = OV.quantise(3, 50, tree, np.concatenate(descs))
    pairs = [(a, b) for a, b, _ in OV.candidate_pairs([per] * views, words, 125000, 5, float(np.float32(0.1)))]
    app = os.path.join(ROOT, "v3dsfmtoolkit_amd", "bin", "v3dm_match_using_voctree")
    return _run_driver(v, args, tmp, descs, [app, "--quiet", "--db", os.path.join(tmp, "out.db"), "tree.bin",
                                             "list.txt", "run.conf"], pairs,
                       f"match_using_voctree counterpart: {views} views x {per} SIFT-like features in V3DBIN files, "
                       "vocabulary tree depth 3 x branch 50, POTENTIAL_VIEWS_COUNT 5, minScore 0.9, minRatio 0.8")


if __name__ == "__main__":
    main()
