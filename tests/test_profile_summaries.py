"""The kernel-trace summaries behind the committed profiles, on synthetic
rocprofv3 kernel-trace CSVs whose answers are worked out by hand (no GPU)."""
import csv

from mun_amd.profile_kernels import summarize_gaps, summarize_trace

FIELDS = ["Kind", "Queue_Id", "Stream_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp", "LDS_Block_Size",
          "VGPR_Count", "Grid_Size_X"]


def _write(tmp_path, rows):
    path = tmp_path / "trace.csv"
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for stream, s, e, name, grid in rows:
            w.writerow({"Kind": "KERNEL_DISPATCH", "Queue_Id": 0, "Stream_Id": stream, "Kernel_Name": name + "(args)",
                        "Start_Timestamp": s, "End_Timestamp": e, "LDS_Block_Size": 0, "VGPR_Count": 8,
                        "Grid_Size_X": grid})
    return str(path)


# times in ns: stream 1 runs three step launches and, after a 100 us gap, a
# fill and a step; stream 2 runs one correlation inside that gap
ROWS = [(1, 0, 50_000, "void sem::step2_kernel<float, 4, 0>", 512),
        (1, 52_000, 100_000, "void sem::step2_kernel<float, 4, 0>", 512),
        (1, 200_000, 210_000, "void sem::fill_kernel<float>", 64),
        (1, 211_000, 260_000, "void sem::step2_kernel<float, 4, 0>", 512),
        (2, 120_000, 150_000, "void sem::corr_kernel<float, 4>", 64)]


def test_trace_summary_groups_by_kernel_and_grid(tmp_path):
    text = summarize_trace(_write(tmp_path, ROWS))
    lines = text.splitlines()
    assert lines[0] == "span 0.000 s, kernel time 0.000 s"
    step = [ln for ln in lines if "step2_kernel" in ln]
    assert len(step) == 1
    # 3 calls, 50 + 48 + 49 = 147 us, 49.0 us each, 147/187 of the kernel time
    assert "| 3 | 49.0 | 0.1 | 78.61 |" in step[0]
    corr = [ln for ln in lines if "corr_kernel" in ln][0]
    assert "| 1 | 30.0 |" in corr
    # rows by total time, largest first
    assert lines.index(step[0]) < lines.index(corr)


def test_gap_summary_attributes_gaps_and_finds_device_idle(tmp_path):
    text = summarize_gaps(_write(tmp_path, ROWS))
    lines = text.splitlines()
    s1 = [ln for ln in lines if ln.startswith("| 1 |")][0]
    # gaps of 2 and 1 us fall under the 5 us threshold; the 100 us gap before
    # the fill is the one counted
    assert "| 4 | 0.2 | 0.3 | 0.10 | 1 |" in s1
    assert "`sem::fill_kernel<float>` (1, 0.10, 100)" in s1
    # 260 us window, kernels cover 50 + 48 + 30 + 10 + 49 = 187 us
    assert lines[-1] == "two-step window 0.3 ms: no kernel running for 0.07 ms (28.08 %)"
