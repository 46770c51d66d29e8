"""The rocprofv3 CSV summariser (host logic only, synthetic CSV files)."""
import csv
import io
import contextlib

from fv3atm_amd import rocprof_summary as rs


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_short_names():
    assert rs.short("void fv3::k_fused_tp<float, 4, 1, 8, false>(fv3::FusedArgs<float>)") == \
        "fv3::k_fused_tp<float, 4, 1, 8, false>"
    assert rs.short("k_plain(int, float*)") == "k_plain"
    assert rs.short("void fv3::(anonymous namespace)::k_remap(fv3::P)") == "fv3::{anon}::k_remap"


def test_stats_and_counters(tmp_path):
    d = tmp_path / "run" / "host"
    d.mkdir(parents=True)
    k1 = "void fv3::k_a<float>(fv3::Args)"
    k2 = "void at::vectorized_elementwise_kernel<4, bitwise_not_kernel_cuda>(int)"
    _write(d / "1_kernel_trace.csv", ["Kind", "Kernel_Name", "Start_Timestamp", "End_Timestamp"],
           [["K", k1, 100, 300], ["K", k1, 400, 800], ["K", k2, 1000, 1100]])
    agg = rs.kernel_stats(str(tmp_path))
    assert agg["fv3::k_a<float>"] == [2, 600.0]
    assert agg["at::vectorized_elementwise_kernel<4, bitwise_not_kernel_cuda>"] == [1, 100.0]
    # two dispatches of k_a, each reporting its counter in two rows (summed)
    _write(d / "1_counter_collection.csv", ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
           [[1, k1, "FETCH_SIZE", 10], [1, k1, "FETCH_SIZE", 30], [2, k1, "FETCH_SIZE", 60],
            [3, k2, "FETCH_SIZE", 512]])
    c = rs.counter_means(str(tmp_path))
    assert c[("fv3::k_a<float>", "FETCH_SIZE")] == [2, 100.0]
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        rs.main(["counters", str(tmp_path), "--calib-bytes", str(1024 * 1024)])
    text = out.getvalue()
    # 512 KiB read for 1 MiB -> factor 2; k_a: 50 KiB mean -> 102400 bytes
    assert "factor 2.000" in text
    assert "1.0240e+05" in text
