"""The committed rocprofv3 traffic summaries cover every scan variant the
bench may run, so its `roofline.traffic` is never left empty by a change of
the autotuner's pick (CPU only: the variant table is a static library table)."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def lib():
    import v3dsfmtoolkit_amd as v
    return v


def test_single_pair_8192_summary_covers_every_variant(bench, lib):
    # at 8192^2 a variant applies its survivors itself when its grid fits the
    # device, so each variant is listed in the form it runs in there
    for i in range(len(lib.scan_variants())):
        names = [bench._scan_name(lib, i, inline) for inline in (True, False)]
        assert any(bench.measured_traffic(n) is not None for n in names), names


@pytest.mark.parametrize("summary", ["TRAFFIC_SUMMARY_BATCH", "TRAFFIC_SUMMARY_65536"])
def test_large_job_summaries_cover_every_variant(bench, lib, summary):
    # the 256-pair batch and the 65536^2 pair run every variant persistent,
    # with the survivors left to the insertion kernel
    path = getattr(bench, summary)
    for i in range(len(lib.scan_variants())):
        name = bench._scan_name(lib, i, False)
        t = bench.measured_traffic(name, path)
        assert t is not None and t > 0, (summary, name)


def test_traffic_figures_are_at_least_the_algorithmic_bytes(bench):
    # HBM bytes per launch can only exceed what the scan must read: both
    # descriptor sets and their row terms
    assert bench.measured_traffic("k_i8_scan_whole<true>") >= 2 * 8192 * 132
    assert bench.measured_traffic("k_i8_scan_whole<false>", bench.TRAFFIC_SUMMARY_65536) >= 2 * 65536 * 132
    assert bench.measured_traffic("k_i8_scan<4, 4096, false>", bench.TRAFFIC_SUMMARY_BATCH) >= 395726848


def test_counter_pass_scaling(bench, tmp_path):
    # a counter pass as rocprofv3 writes it: the scan's dispatches and the
    # 256 MiB copy kernel's; FETCH_SIZE reads half the copy's bytes, so the
    # scan's median reading is scaled by 2 (and the outlier does not move it)
    header = ("Correlation_Id,Dispatch_Id,Agent_Id,Queue_Id,Process_Id,Thread_Id,Grid_Size,Kernel_Id,Kernel_Name,"
              "Workgroup_Size,LDS_Block_Size,Scratch_Size,VGPR_Count,Accum_VGPR_Count,SGPR_Count,Counter_Name,"
              "Counter_Value,Start_Timestamp,End_Timestamp")
    rows = [("void v3dm::k_i8_scan_whole<true>(v3dm::ScanArgs)", 5000.0),
            ("void v3dm::k_i8_scan_whole<true>(v3dm::ScanArgs)", 5000.0),
            ("void v3dm::k_i8_scan_whole<true>(v3dm::ScanArgs)", 9000.0),
            ("void v3dm::k_i8_scan<4, 512, true>(v3dm::ScanArgs)", 7000.0),
            ("v3dm::k_copy_probe(int4 const*, int4*, long)", 131072.0),
            ("v3dm::k_copy_probe(int4 const*, int4*, long)", 131072.0)]
    p = tmp_path / "t_counter_collection.csv"
    with open(p, "w") as f:
        f.write(header + "\n")
        for i, (name, val) in enumerate(rows):
            f.write(f'{i},{i},"Agent 1",1,1,1,1,1,"{name}",256,0,0,0,0,0,"FETCH_SIZE",{val},0,1\n')
    assert bench.scaled_counter_bytes(str(p), "k_i8_scan_whole<true>") == 5000.0 * 2 * 1024
    assert bench.scaled_counter_bytes(str(p), "k_i8_scan<4, 512, true>") == 7000.0 * 2 * 1024
    assert bench.scaled_counter_bytes(str(p), "k_i8_scan<4, 4096, false>") is None
