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


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_i8_scan_8192x8192_traffic.md")


def measured_traffic(kernel_name):
    """HBM bytes per launch of `kernel_name` at the default workload, as the
    committed rocprofv3 counter summary states it (a line
    `traffic_bytes_per_launch <kernel> = <bytes>`), or None."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                line = line.strip().strip("`")
                if line.startswith(f"traffic_bytes_per_launch {kernel_name} ="):
                    return float(line.split("=")[1])
    except OSError:
        return None
    return None


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
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(pairs shard across GPUs without communication; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import v3dsfmtoolkit_amd as v

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

    m.profileEnable(True)
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        m.findAsync(min_score, min_ratio)
    m.synchronize()
    t1 = time.perf_counter()
    scan_ms, sel_ms, jobs = m.profileRead()
    m.profileEnable(False)
    final = m.fetchMatches()
    assert final == got

    wall = (t1 - t0) / args.steps
    pairs = n * n
    scan_avg_s = scan_ms / 1e3 / max(jobs, 1)
    sel_avg_s = sel_ms / 1e3 / max(jobs, 1)
    if kernel == v.KERNEL_I8:
        achieved = pairs * 256 / scan_avg_s / 1e12
        roof = {"bound": "mfma", "kernel": "k_i8_scan", "achieved": achieved, "peak": INT8_PEAK_TOPS,
                "unit": "TFLOP/s", "frac": achieved / INT8_PEAK_TOPS,
                "traffic": measured_traffic("k_i8_scan") if n == 8192 else None,
                "ops_per_pair": 256, "launch_avg_us": scan_avg_s * 1e6}
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
                   "n_left": n, "n_right": n, "kernel": args.kernel, "matches": len(got)},
        "image_pairs_per_s": 1.0 / wall,
        "select_avg_us": sel_avg_s * 1e6,
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
    m.close()
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
