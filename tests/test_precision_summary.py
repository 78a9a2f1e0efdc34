"""The precision study's summary (tests/precision_study.py) on synthetic per-step records (CPU):
the table's marked steps, the max over steps and the GPU / FP32 ratios."""
import json
import re

import precision_study as PS


def test_summary_table_and_maxima(tmp_path):
    recs = []
    for s in range(1, 31):
        r = {"step": s, "t": s * 5e-4, "its_gpu": 40 + s % 3, "its_fp64": 41, "its_fp32": 42, "bodies_gpu": 1,
             "bodies_fp64": 1}
        for q in PS.QUANTITIES:
            r["gpu_" + q] = 1e-6 * s
            r["fp32_" + q] = 2e-6 * s
        recs.append(r)
    meta = {"particles": 100, "size_m": [0.1, 0.1, 0.1], "h": 1 / 256, "dt": 5e-4, "steps": 30, "seconds": 1.0,
            "oracle_threads": 1}
    out = tmp_path / "s.md"
    s = PS.summarize(recs, meta, str(out), "cmd")
    text = out.read_text()
    rows = [l for l in text.splitlines() if re.match(r"\| \d+ \|", l)]
    assert [int(l.split("|")[1]) for l in rows] == [1, 10, 25, 30]
    assert abs(s["max_over_steps"]["gpu"]["F_max"] - 3e-5) < 1e-18
    assert all(abs(v - 0.5) < 1e-12 for v in s["gpu_over_fp32_max"].values())
    foot = json.loads(re.search(r"<!-- json: (.*) -->", text).group(1))
    assert foot["final"]["its_gpu"] == recs[-1]["its_gpu"]


def test_study_scene_keeps_the_4m_density():
    sc, size = PS.study_scene(60000)
    vol = size[0] * (size[1] - 1.0 / 256) * size[2]
    assert abs(len(sc["m"]) / vol / PS.DENSITY - 1) < 1e-9
    assert sc["h"] == 1.0 / 256 and sc["dt"] == 5e-4
