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


def test_traffic_line_and_bench_lookup(tmp_path):
    k1 = "void fv3::k_fused_tp<float, 4, 2, 8, 0, 1>(fv3::FusedArgs<float>)"
    cal = "void at::vectorized_elementwise_kernel<4, bitwise_not_kernel_cuda>(int)"
    for name, counter, vals in (("f", "FETCH_SIZE", (512, 10, 20)), ("w", "WRITE_SIZE", (1024, 3, 5))):
        d = tmp_path / name
        d.mkdir()
        _write(d / "1_counter_collection.csv", ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
               [[1, cal, counter, vals[0]], [2, k1, counter, vals[1]], [3, k1, counter, vals[2]]])
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        rs.main(["traffic", str(tmp_path / "f"), str(tmp_path / "w"), "--kernel", "k_fused_tp",
                 "--launches", "2", "--calib-bytes", str(1024 * 1024), "--tag", "n=8 npz=2 nq=1 cfg=0 rows=60"])
    # fetch: (10 + 20) KiB x 2 = 61440 B; write: (3 + 5) KiB x 1 = 8192 B; per launch 34816 B
    line = [l for l in out.getvalue().splitlines() if l.startswith("bench-traffic:")][0]
    assert line == "bench-traffic: n=8 npz=2 nq=1 cfg=0 rows=60 bytes_per_pass=34816"


def test_spans_of_concurrent_groups(tmp_path):
    d = tmp_path / "t"
    d.mkdir()
    k = "void fv3::k_fused_tp<float, 4, 2, 8, {}, 1>(fv3::FusedArgs<float>)"
    # two launches of two overlapping kernels each: spans 300 and 250
    _write(d / "1_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
           [[k.format(0), 100, 400], [k.format(1), 150, 300], [k.format(0), 500, 700], [k.format(1), 510, 750],
            ["void other(int)", 0, 1000]])
    assert rs.spans(str(tmp_path), "k_fused_tp", 2) == [300.0, 250.0]


def test_valu_rates_and_bench_line(tmp_path):
    k1 = "void fv3::k_fused_tp<float, 4, 2, 8, 0, 1>(fv3::FusedArgs<float>)"
    k2 = "void fv3::k_fused_tp<float, 4, 1, 8, 5, 3>(fv3::FusedArgs<float>)"
    sq, tr = tmp_path / "sq", tmp_path / "tr"
    sq.mkdir()
    tr.mkdir()
    # k1: two dispatches of 6.144e8 VALU each, 1 ms each -> 614.4 G/s = the issue peak
    _write(sq / "1_counter_collection.csv", ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"],
           [[1, k1, "SQ_INSTS_VALU", 6.144e8], [2, k1, "SQ_INSTS_VALU", 6.144e8], [1, k1, "SQ_INSTS_SALU", 1e8],
            [2, k1, "SQ_INSTS_SALU", 1e8], [3, k2, "SQ_INSTS_VALU", 3.072e8]])
    _write(tr / "1_kernel_trace.csv", ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
           [[k1, 0, 1000000], [k1, 2000000, 3000000], [k2, 4000000, 6000000]])
    rows = {r[0]: r for r in rs.valu_table(str(sq), str(tr))}
    n1 = "fv3::k_fused_tp<float, 4, 2, 8, 0, 1>"
    assert abs(rows[n1][1] - 1.0) < 1e-12 and abs(rows[n1][5] - 1.0) < 1e-12
    assert abs(rows["fv3::k_fused_tp<float, 4, 1, 8, 5, 3>"][5] - 0.25) < 1e-12
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        rs.main(["valu", str(sq), str(tr), "--kernel", "k_fused_tp", "--launches", "2", "--tag", "n=8"])
    # (2 x 6.144e8 + 3.072e8) / 2 launches
    line = [l for l in out.getvalue().splitlines() if l.startswith("bench-valu:")][0]
    assert line == "bench-valu: n=8 valu_per_pass=768000000"
