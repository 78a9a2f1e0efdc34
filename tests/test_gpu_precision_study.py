"""The slope-slab precision study (tests/precision_study.py) at a small size: 6000 particles of the
4M slab's density, materials, h and dt, 12 steps on the GPU and in the FP32 and FP64 oracles. Every
step, the GPU's momentum, position and F errors against FP64 stay within 3x the FP32 oracle's (5x
after the first solve, which runs to the iteration cap; floors of 1e-6 relative, 1e-6 h, 1e-6) at
the maximum and the 99th percentile, and the summary file carries the per-step table."""
import json
import re

import pytest

import precision_study as PS

pytestmark = pytest.mark.gpu


def test_precision_study_small_slab(tmp_path):
    print("[gpu-case] precision study: slope slab 6000 particles, 12 steps", flush=True)
    recs, meta = PS.run_study(6000, 12, str(tmp_path / "steps.csv"), log=lambda s: print(s, flush=True))
    floors = {"mom_rel": 1e-6, "x_max": 1e-6, "x_p99": 1e-6, "F_max": 1e-6, "F_p99": 1e-6}
    for r in recs:
        # the first solve runs to the iteration cap in every precision (the undeformed state's
        # Ginv = I, SnowConstitutiveModel.cpp:123-169) and ends wherever the rounding floor leaves
        # its iterate: there the bar is 5x (measured 3.7x on the momentum)
        k = 5.0 if r["its_fp64"] >= 1000 else 3.0
        for q, fl in floors.items():
            assert r["gpu_" + q] <= k * r["fp32_" + q] + fl, (r["step"], q, r["gpu_" + q], r["fp32_" + q])
    out = tmp_path / "study.md"
    s = PS.summarize(recs, meta, str(out), "python tests/precision_study.py --n 6000 --steps 12")
    foot = json.loads(re.search(r"<!-- json: (.*) -->", out.read_text()).group(1))
    assert foot["steps"] == 12 and foot["particles"] == meta["particles"]
    assert s["max_over_steps"]["gpu"]["x_max"] == max(r["gpu_x_max"] for r in recs)
    assert len((tmp_path / "steps.csv").read_text().splitlines()) == 13


def test_precision_study_gpu64_reference_agrees_with_the_oracle(tmp_path):
    """The --gpu64 mode measures the FP32-bulk step against the FP64 GPU step instead of the FP64
    oracle. On the same scene both references must report the same errors: each quantity of each
    step within a factor 2 of the oracle-based figure, or 5e-6 apart (the two FP64 runs differ by
    up to ~1e-6 h after the capped first solve, which amplifies rounding)."""
    print("[gpu-case] precision study --gpu64: slope slab 6000 particles, 8 steps, against the oracle-based study",
          flush=True)
    ref, _ = PS.run_study(6000, 8, fp32=False, log=lambda s: print(s, flush=True))
    recs, meta = PS.run_study_gpu64(6000, 8, str(tmp_path / "steps64.csv"), log=lambda s: print(s, flush=True),
                                    full_slab=False)
    assert meta["particles"] == 6000 or abs(meta["particles"] - 6000) < 10
    for a, b in zip(ref, recs):
        assert a["its_gpu"] == b["its_gpu"]  # the same FP32-bulk run
        for q in PS.QUANTITIES:
            x, y = a["gpu_" + q], b["gpu_" + q]
            assert (0.5 * x <= y <= 2.0 * x) or abs(x - y) < 5e-6, (a["step"], q, x, y)
    out = tmp_path / "study64.md"
    s = PS.summarize_gpu64(recs, meta, str(out), "python tests/precision_study.py --gpu64 --n 6000 --steps 8")
    foot = json.loads(re.search(r"<!-- json: (.*) -->", out.read_text()).group(1))
    assert foot["steps"] == 8 and s["max_over_steps"]["gpu"]["x_max"] == max(r["gpu_x_max"] for r in recs)
