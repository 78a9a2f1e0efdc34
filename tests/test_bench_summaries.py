"""bench.py's rocprofv3 summarizers on synthetic counter files (CPU): the traffic summary scales the
HVP's FETCH_SIZE / WRITE_SIZE by the known-size calibration kernels, and the SQ summary averages the
HVP's dispatches of the last 40% and derives the issue fractions with the stated formulas."""
import csv
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import bench  # noqa: E402


def write_csv(path, rows):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Grid_Size", "Counter_Name", "Counter_Value"])
        for r in rows:
            w.writerow(r)


def footer(path):
    return json.loads(re.search(r"<!-- json: (.*) -->", open(path).read()).group(1))


def test_pmc_summary_scales_by_calibration(tmp_path):
    mib = bench.COPY_BYTES / 1024.0
    for counter, calib, factor in (("FETCH_SIZE", "reduce_kernel", 0.5), ("WRITE_SIZE", "copyBuffer", 1.0)):
        rows = [(0, 0, "at::native::" + calib, 1, counter, factor * mib)]
        rows += [(d, d, "void mpm::k_hvp_pair<true>(x)", 4096, counter, 1000.0 * factor) for d in range(1, 11)]
        write_csv(str(tmp_path / (counter + ".csv")), rows)
    out = str(tmp_path / "traffic.md")
    bench.pmc_summary(str(tmp_path / "FETCH_SIZE.csv"), str(tmp_path / "WRITE_SIZE.csv"), out)
    s = footer(out)
    # 1000 KiB true traffic for each counter after scaling
    assert abs(s["fetch"]["hvp_bytes_per_launch"] - 1000.0 * 1024) < 1e-6
    assert abs(s["traffic_bytes_per_launch"] - 2000.0 * 1024) < 1e-6


def test_sq_summary_fractions(tmp_path):
    paths = []
    vals = {"SQ_BUSY_CYCLES": 32 * 1000.0, "SQ_INSTS_VALU": 1024 * 600 / 4.0, "SQ_LDS_IDX_ACTIVE": 256 * 500.0,
            "SQ_LDS_BANK_CONFLICT": 256 * 100.0, "SQ_LDS_ADDR_CONFLICT": 256 * 50.0,
            "SQ_WAVE_CYCLES": 256 * 1000.0 * 18 / 4.0, "SQ_WAVES": 100.0}
    for k, p in enumerate(bench.SQ_PASSES):
        rows = []
        for d in range(10):
            # the early dispatches (first 60%) carry different values and must be ignored
            scale = 3.0 if d < 6 else 1.0
            for name in p.split():
                rows.append((d, d, "void mpm::k_hvp_pair<true>(x)", 1000, name, scale * vals.get(name, 1.0)))
        path = str(tmp_path / ("p%d.csv" % k))
        write_csv(path, rows)
        paths.append(path)
    out = str(tmp_path / "sq.md")
    bench.sq_summary(paths, out)
    s = footer(out)
    assert s["valu_issue_frac"] == 0.6
    assert s["lds_active_frac"] == 0.5
    assert s["resident_waves_per_cu"] == 18.0


def test_pmc_summary_learned_leg_population(tmp_path):
    """With network dispatches present, the traffic figure is the mean over every HVP dispatch
    between the first and last of them (the learned leg), not the full-grid median."""
    mib = bench.COPY_BYTES / 1024.0
    for counter, calib, factor in (("FETCH_SIZE", "reduce_kernel", 0.5), ("WRITE_SIZE", "copyBuffer", 1.0)):
        rows = [("at::native::" + calib, 1, factor * mib, 1)]
        t = 10
        # warm-up: large launches, no network
        for _ in range(5):
            rows.append(("void mpm::k_hvp_pair<true>(x)", 4096, 5000.0 * factor, t))
            t += 1
        # learned leg: network + HVPs of 1000 and 3000 KiB
        for v in (1000.0, 3000.0, 1000.0, 3000.0):
            rows.append(("void mpm::k_conv<16, 16, 0>(x)", 64, 1.0, t))
            rows.append(("void mpm::k_hvp_pair<true>(x)", 4096 if v > 2000 else 1024, v * factor, t + 1))
            t += 2
        rows.append(("void mpm::k_conv<16, 16, 0>(x)", 64, 1.0, t))
        # plain leg
        for _ in range(5):
            t += 1
            rows.append(("void mpm::k_hvp_pair<true>(x)", 4096, 7000.0 * factor, t))
        with open(str(tmp_path / (counter + ".csv")), "w", newline="") as fh:
            w = csv.writer(fh)
            w.writerow(["Kernel_Name", "Grid_Size", "Counter_Value", "Start_Timestamp"])
            for r in rows:
                w.writerow(r)
    out = str(tmp_path / "traffic.md")
    bench.pmc_summary(str(tmp_path / "FETCH_SIZE.csv"), str(tmp_path / "WRITE_SIZE.csv"), out)
    s = footer(out)
    assert s["fetch"]["leg_launches"] == 4
    assert abs(s["fetch"]["leg_bytes_per_launch"] - 2000.0 * 1024) < 1e-6
    assert abs(s["traffic_bytes_per_launch"] - 4000.0 * 1024) < 1e-6


def test_learned_window_from_markers():
    """The learned leg lies between the first two leg markers (torch kernels bench.py launches
    around each leg's timed steps); counter files repeat a dispatch once per counter, and the
    calibration kernels are not markers. Without markers the network's dispatches bound it."""
    arange = "void (anonymous namespace)::elementwise_kernel_with_index<int, at::native::arange_cuda_out>(x)"
    rows = [("void at::native::vectorized_elementwise_kernel<4, at::native::FillFunctor<float>>(x)", 1),
            ("void at::native::reduce_kernel<512, 1>(x)", 2),
            (arange, 10), (arange, 10), ("void mpm::k_hvp_pair<true>(x)", 11), ("void mpm::k_conv<16, 16, 0>(x)", 12),
            (arange, 20), (arange, 20), (arange, 30), ("void mpm::k_hvp_pair<true>(x)", 31), (arange, 40)]
    assert bench.learned_window(rows) == (10, 20)
    assert bench.learned_window([r for r in rows if "arange" not in r[0]]) == (12, 12)
    assert bench.learned_window([("void mpm::k_hvp_pair<true>(x)", 5)]) is None


def test_stiff_summary_ranges_and_rates(tmp_path):
    """The stiff table splits at the first step whose body count differs from step 0's and computes
    particle-steps/s from the steps' own wall times."""
    def leg(ms, fallbacks):
        return {"value": 0.0, "ms_per_step": ms, "steps": [
            {"ms": ms, "bodies": 8 if i < 6 else 9, "iters_sum": 100 + i, "iters_max": 50, "fallbacks": fallbacks,
             "capped": 1 if i == 7 else 0} for i in range(10)]}
    d = {"plain": leg(500.0, 0), "learned": leg(1000.0, 2),
         "config": "stiff stress case: 1000 particles in 8 seeded boxes and spheres (seed 4)",
         "objects": [{"kind": "box", "center": [0.1, 0.2, 0.3], "size": [0.1, 0.1, 0.1], "count": 1000}],
         "warmup": 0}
    js = tmp_path / "stiff.json"
    js.write_text("[bench] progress\n" + json.dumps(d) + "\n")
    out = str(tmp_path / "stiff.md")
    bench.stiff_summary(str(js), out)
    s = footer(out)
    assert s["split"] == 6
    assert abs(s["ratio_before_split"] - 0.5) < 1e-12
    text = open(out).read()
    # 1000 particles per 0.5 s step = 2000 particle-steps/s over steps 0-5
    assert "| plain | 0-5 | 2e+03 | 500.0 | 102 | 50 | 0.0 | 0 | 8 |" in text
    assert "| learned | 6-9 | 1e+03 | 1000.0 | 108 | 50 | 2.0 | 1 | 9 |" in text
