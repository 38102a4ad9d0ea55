"""CPU: the kernel-trace summaries of megadock_amd.profiling on synthetic traces whose
busy time and gaps are known by construction."""
import csv
import os

from megadock_amd import profiling as P


def _write_trace(d, rows):
    with open(os.path.join(d, "run_kernel_trace.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X", "Grid_Size_Y", "Grid_Size_Z"])
        for name, s, e in rows:
            w.writerow([name, s, e, 64, 1, 1])


def test_busy_segments_and_gaps(tmp_path):
    ms = 1000000
    rows, t = [], 0
    # warm-up segment: 5 kernels of 1 ms back to back
    for _ in range(5):
        rows.append(("void (anonymous namespace)::kp_fwd_zy<96, 48, true>(VoxGrid)", t, t + ms))
        t += ms
    t += 300 * ms  # the next process
    # main segment: 20 kernels of 2 ms; after every 4th a 0.7 ms gap, the others overlap by 0.1 ms
    for i in range(20):
        name = "void (anonymous namespace)::kp_x_corr<96, 16>(float2 const*)" if i % 4 == 3 else "k_reduce_top"
        rows.append((name, t, t + 2 * ms))
        t += 2 * ms + (int(0.7 * ms) if i % 4 == 3 else -int(0.1 * ms))
    # a zero-length gap at the end does not count
    rows.append(("k_vox_lds", t - int(0.7 * ms), t - int(0.7 * ms) + ms))
    _write_trace(tmp_path, rows)
    b = P.busy(str(tmp_path))
    assert b["segments"] == 2
    assert b["dispatches"] == 26
    m = b["main"]
    # 20 kernels of 2 ms less 15 overlaps of 0.1 ms, plus the last kernel's tail beyond the 20th
    assert abs(m["idle_ms"] - 4 * 0.7) < 1e-9  # the gaps after kernels 4, 8, 12, 16 (the 20th's follows the end)
    assert m["gaps_by_length"]["0.5-1 ms"] == {"gaps": 4, "ms": 2.8}
    assert m["around_gaps"] == {"kp_x_corr -> k_reduce_top": {"gaps": 4, "ms": 2.8}}
    assert abs(m["busy_ms"] + m["idle_ms"] - m["span_ms"]) < 1e-9
    assert abs(b["idle_ms"] - (300 + 2.8)) < 1e-9
