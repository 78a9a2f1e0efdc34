"""Precision study of the slope-slab configuration: the GPU's FP32-bulk step against a full-FP64 run
of the same scene, per step, over 300 steps (the step compared is Sim::advance, Sim.cpp:57-134).

The full-FP64 run is the FP64 oracle (the reference's algorithm with every scalar in double) on the
CPU, and the reference's own precision is the FP32 oracle beside it; both start from the GPU's
initial state and take the same steps. The FP64 oracle steps about 50k particles per second on
16 cores, so the study runs a slab of the 4M scene's particle density, h, dt, materials (per-particle
E0 log-uniform in [7e4, 2.8e5] Pa, xi in [5, 10]), slope and friction, cut to 0.1 x 0.1 x 0.1 m
(60k particles) instead of 0.8 x 0.2 x 0.4 m.

Per step, for the GPU and for the FP32 oracle against the FP64 oracle: the relative error of the
total momentum sum m v, positions (max and 99th percentile over particles of the largest
per-axis error, in units of h) and F (max and 99th percentile over particles of the largest
elementwise error); and the CR iteration count of the largest body in all three runs.

    python tests/precision_study.py --n 60000 --steps 300 --csv <per-step csv> --out <summary .md>

With --gpu64 the full-FP64 run is the library's FP64 GPU step (mpm_sim64_*, pinned against the FP64
oracle by tests/test_gpu_fp64.py) instead, at the configuration's own size (4M particles):

    python tests/precision_study.py --gpu64 --n 4000000 --steps 300 --csv <per-step csv> --out <summary .md>

This script uses the oracle as the checker; it is not part of the library or of bench.py's timed
path. `tests/test_gpu_precision_study.py` runs both modes at a small size.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

DENSITY = 4.0e6 / (0.8 * 0.2 * 0.4)  # particles per m^3 of the 4M slope slab
SLAB_YZ = (0.1, 0.1)
QUANTITIES = ("mom_rel", "x_max", "x_p99", "F_max", "F_p99")


def study_scene(n, seed=2):
    from mpmsnow_amd import scenes

    sy, sz = SLAB_YZ
    sx = n / (DENSITY * (sy - 1.0 / 256) * sz)
    return scenes.slope_slab(n=n, seed=seed, size=(sx, sy, sz)), (sx, sy, sz)


def _oracle(O, sc, prec):
    mat = sc["material"]
    cols = O.Colliders.create(prec)
    for c, vv, mu, st in sc["planes"]:
        cols.add_plane(c, v=vv, friction=mu, sticky=st)
    f = O.Fields.create(prec)
    f.add_gravity(sc["gravity"])
    model = O.Model.snow(mat["E"], mat["nu"], mat["xi"], mat["theta_c"], mat["theta_s"], prec)
    sim = O.Sim(sc["x"], sc["m"], float(np.float32(sc["h"])), model, cols, f, mu0=sc["mu0"],
                lambda0=sc["lambda0"], xi=sc["xi"], prec=prec)
    sim.pd.set("v", sc["v"])
    sim._keep = (cols, f, model)
    return sim


def _errors(m, h, ref, run):
    """ref, run: dicts of p (n,3), v (n,3), F (n,9)."""
    P64 = (m[:, None].astype(np.float64) * ref["v"]).sum(0)
    P = (m[:, None].astype(np.float64) * run["v"]).sum(0)
    dx = np.abs(run["p"] - ref["p"]).max(1) / h
    dF = np.abs(run["F"] - ref["F"]).max(1)
    return {"mom_rel": float(np.linalg.norm(P - P64) / max(np.linalg.norm(P64), 1e-30)),
            "x_max": float(dx.max()), "x_p99": float(np.percentile(dx, 99)),
            "F_max": float(dF.max()), "F_p99": float(np.percentile(dF, 99))}


def run_study(n, steps, csv_path=None, fp32=True, log=print):
    """Steps the GPU library, the FP64 oracle and (with fp32) the FP32 oracle in lockstep and
    returns the per-step records."""
    import mpmsnow_amd as G
    from oracle import pyoracle as O

    sc, size = study_scene(n)
    n = len(sc["m"])
    h, dt = sc["h"], sc["dt"]
    gpu = G.Sim(sc["x"], sc["m"], h, material=sc["material"], planes=sc["planes"], gravity=[sc["gravity"]],
                v=sc["v"], mu0=sc["mu0"], lambda0=sc["lambda0"], xi=sc["xi"])
    o64 = _oracle(O, sc, "double")
    o32 = _oracle(O, sc, "float") if fp32 else None
    m = sc["m"]
    recs = []
    fh = open(csv_path, "w") if csv_path else None
    cols = ["step", "t"] + ["gpu_" + q for q in QUANTITIES] + (["fp32_" + q for q in QUANTITIES] if fp32 else [])
    cols += ["its_gpu", "its_fp64"] + (["its_fp32"] if fp32 else []) + ["bodies_gpu", "bodies_fp64"]
    if fh:
        fh.write(",".join(cols) + "\n")
    t0 = time.time()
    for step in range(steps):
        st = gpu.advance(dt, sc["max_iters"], sc["tol"])
        o64.advance(dt, sc["max_iters"], sc["tol"])
        if o32 is not None:
            o32.advance(dt, sc["max_iters"], sc["tol"])
        ref = {k: o64.pd.get(k).reshape(n, -1) for k in ("p", "v", "F")}
        rg = {k: gpu.get(k).reshape(n, -1) for k in ("p", "v", "F")}
        rec = {"step": step + 1, "t": (step + 1) * dt}
        for k, v in _errors(m, h, ref, rg).items():
            rec["gpu_" + k] = v
        if o32 is not None:
            r32 = {k: o32.pd.get(k).reshape(n, -1) for k in ("p", "v", "F")}
            for k, v in _errors(m, h, ref, r32).items():
                rec["fp32_" + k] = v
        bg = max((gpu.body_stats(b, hist_cap=0) for b in range(st["num_solved"])), key=lambda r: r["particles"])
        s64 = o64.last_stats()
        rec["its_gpu"] = bg["iterations"]
        rec["its_fp64"] = max(s64, key=lambda r: r["particles"])["iterations"]
        if o32 is not None:
            rec["its_fp32"] = max(o32.last_stats(), key=lambda r: r["particles"])["iterations"]
        rec["bodies_gpu"] = st["num_solved"]
        rec["bodies_fp64"] = len(s64)
        recs.append(rec)
        if fh:
            fh.write(",".join("%.6g" % rec[c] if isinstance(rec[c], float) else str(rec[c]) for c in cols) + "\n")
            fh.flush()
        log("step %d/%d (%.0f s): gpu x_max %.2e h F_max %.2e mom %.2e%s | its gpu %d fp64 %d" % (
            step + 1, steps, time.time() - t0, rec["gpu_x_max"], rec["gpu_F_max"], rec["gpu_mom_rel"],
            (" | fp32 x_max %.2e h F_max %.2e mom %.2e" % (rec["fp32_x_max"], rec["fp32_F_max"], rec["fp32_mom_rel"])
             if o32 is not None else ""), rec["its_gpu"], rec["its_fp64"]))
    if fh:
        fh.close()
    meta = {"particles": n, "size_m": [round(s, 4) for s in size], "h": h, "dt": dt, "steps": steps,
            "seconds": time.time() - t0, "oracle_threads": O.num_threads()}
    return recs, meta


def summarize(recs, meta, out_path, command):
    fp32 = "fp32_mom_rel" in recs[0]
    marks = sorted({s for s in (1, 10, 25, 50, 100, 150, 200, 250, 300) if s <= len(recs)} | {len(recs)})
    lines = [
        "# Precision study, slope slab: FP32-bulk GPU step against a full-FP64 run, %d steps" % meta["steps"],
        "",
        "Command (from the repository root on the GPU host):",
        "",
        "    " + command,
        "",
        "Scene: the 4M slope slab's particle density, h = 1/256, dt = 5e-4, per-particle E0 / xi (seed 2),",
        "30-degree slope with friction 0.2, CR tolerance 1e-5; the slab cut to %s m (%d particles). The"
        % (" x ".join("%.3g" % s for s in meta["size_m"]), meta["particles"]),
        "full-FP64 run is the FP64 oracle on the host's CPU (%d threads); the FP32 oracle is the" % meta["oracle_threads"],
        "reference's own precision. Errors are against the FP64 run: momentum = |sum m v - sum m v64| /",
        "|sum m v64|; x = per particle the largest per-axis position error, in units of h; F = per",
        "particle the largest elementwise error; max and 99th percentile over particles. %.0f s in all." %
        meta["seconds"],
        "",
        "| step | momentum GPU | momentum FP32 | x max GPU | x max FP32 | x p99 GPU | x p99 FP32 | F max GPU "
        "| F max FP32 | F p99 GPU | F p99 FP32 | CR its GPU / FP32 / FP64 |",
        "|---|---|---|---|---|---|---|---|---|---|---|---|",
    ]

    def cell(r, k):
        return "%.2e" % r[k] if k in r else "-"

    for s in marks:
        r = recs[s - 1]
        row = [str(s)]
        for q in QUANTITIES:
            row += [cell(r, "gpu_" + q), cell(r, "fp32_" + q)]
        row.append("%d / %s / %d" % (r["its_gpu"], r.get("its_fp32", "-"), r["its_fp64"]))
        lines.append("| " + " | ".join(row) + " |")
    worst = {}
    for who in ("gpu",) + (("fp32",) if fp32 else ()):
        worst[who] = {q: max(r[who + "_" + q] for r in recs) for q in QUANTITIES}
    row = ["max over steps"]
    for q in QUANTITIES:
        row += ["%.2e" % worst["gpu"][q], "%.2e" % worst["fp32"][q] if fp32 else "-"]
    row.append("")
    lines.append("| " + " | ".join(row) + " |")
    ratio = None
    if fp32:
        ratio = {q: (worst["gpu"][q] / worst["fp32"][q] if worst["fp32"][q] > 0 else None) for q in QUANTITIES}
        lines += ["", "GPU / FP32-oracle, max over steps: " + ", ".join(
            "%s %.3g" % (q, v) for q, v in ratio.items() if v is not None) + "."]
    its = np.array([[r["its_gpu"], r["its_fp64"]] for r in recs], dtype=float)
    lines += ["", "CR iterations of the largest body, mean over steps: GPU %.1f, FP64 %.1f%s." % (
        its[:, 0].mean(), its[:, 1].mean(),
        ", FP32 oracle %.1f" % np.mean([r["its_fp32"] for r in recs]) if fp32 else "")]
    summary = {"workload": "slope_slab_precision", "particles": meta["particles"], "steps": meta["steps"],
               "final": {k: v for k, v in recs[-1].items() if k not in ("step", "t")},
               "max_over_steps": worst, "gpu_over_fp32_max": ratio, "seconds": meta["seconds"]}
    lines += ["", "<!-- json: %s -->" % json.dumps(summary)]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return summary


def read_csv(path, n):
    """Per-step records and the run's description from a csv that run_study wrote (the timing and
    thread count are not in the csv: they read as 0)."""
    import csv

    recs = []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            recs.append({k: (int(v) if k in ("step", "its_gpu", "its_fp64", "its_fp32", "bodies_gpu", "bodies_fp64")
                             else float(v)) for k, v in r.items()})
    sc, size = study_scene(n)
    meta = {"particles": len(sc["m"]), "size_m": [round(s, 4) for s in size], "h": sc["h"], "dt": sc["dt"],
            "steps": len(recs), "seconds": 0.0, "oracle_threads": 0}
    return recs, meta


def run_study_gpu64(n, steps, csv_path=None, log=print, full_slab=True):
    """The same comparison with the full-FP64 run on the GPU (mpm_sim64_*, the reference's algorithm in
    double), which reaches the configuration's own size: the 4M-particle 0.8 x 0.2 x 0.4 m slab
    (full_slab), or the study's cut of the slab at n particles. Per step: the FP32-bulk GPU step's
    momentum, position and F errors against the FP64 GPU step, and both runs' CR iterations."""
    import mpmsnow_amd as G
    from mpmsnow_amd import scenes

    if full_slab:
        sc = scenes.slope_slab(n=n)
        size = (0.8, 0.2, 0.4)
    else:
        sc, size = study_scene(n)
    n = len(sc["m"])
    h, dt = sc["h"], sc["dt"]
    kw = dict(planes=sc["planes"], gravity=[sc["gravity"]], v=sc["v"], mu0=sc["mu0"], lambda0=sc["lambda0"],
              xi=sc["xi"])
    gpu = G.Sim(sc["x"], sc["m"], h, material=sc["material"], **kw)
    g64 = G.Sim64(sc["x"], sc["m"], float(np.float32(h)), sc["material"], **kw)
    m = sc["m"]
    recs = []
    fh = open(csv_path, "w") if csv_path else None
    cols = ["step", "t"] + ["gpu_" + q for q in QUANTITIES] + ["its_gpu", "its_fp64", "bodies_gpu", "bodies_fp64",
                                                              "s_fp32", "s_fp64"]
    if fh:
        fh.write(",".join(cols) + "\n")
    t0 = time.time()
    for step in range(steps):
        ta = time.time()
        st = gpu.advance(dt, sc["max_iters"], sc["tol"])
        tb = time.time()
        s64 = g64.advance(dt, sc["max_iters"], sc["tol"])
        tc = time.time()
        ref = {k: g64.get(k).reshape(n, -1) for k in ("p", "v", "F")}
        rg = {k: gpu.get(k).reshape(n, -1) for k in ("p", "v", "F")}
        rec = {"step": step + 1, "t": (step + 1) * dt}
        for k, v in _errors(m, h, ref, rg).items():
            rec["gpu_" + k] = v
        big = max((gpu.body_stats(b, hist_cap=0) for b in range(st["num_solved"])), key=lambda r: r["particles"])
        big64 = max((g64.body_stats(b, hist_cap=0) for b in range(s64["num_solved"])), key=lambda r: r["particles"])
        rec.update(its_gpu=big["iterations"], its_fp64=big64["iterations"], bodies_gpu=st["num_solved"],
                   bodies_fp64=s64["num_solved"], s_fp32=tb - ta, s_fp64=tc - tb)
        recs.append(rec)
        if fh:
            fh.write(",".join("%.6g" % rec[c] if isinstance(rec[c], float) else str(rec[c]) for c in cols) + "\n")
            fh.flush()
        log("step %d/%d (%.0f s): FP32 vs FP64 x_max %.2e h F_max %.2e mom %.2e | its FP32 %d FP64 %d | "
            "bodies %d / %d | %.3f s / %.3f s" % (step + 1, steps, time.time() - t0, rec["gpu_x_max"],
                                                  rec["gpu_F_max"], rec["gpu_mom_rel"], rec["its_gpu"],
                                                  rec["its_fp64"], rec["bodies_gpu"], rec["bodies_fp64"],
                                                  rec["s_fp32"], rec["s_fp64"]))
    if fh:
        fh.close()
    meta = {"particles": n, "size_m": [round(x, 4) for x in size], "h": h, "dt": dt, "steps": steps,
            "seconds": time.time() - t0, "reference": "gpu64"}
    return recs, meta


def summarize_gpu64(recs, meta, out_path, command):
    marks = sorted({s for s in (1, 2, 5, 10, 25, 50, 100, 150, 200, 250, 300) if s <= len(recs)} | {len(recs)})
    lines = [
        "# Precision study, slope slab: FP32-bulk GPU step against the FP64 GPU step, %d particles, %d steps"
        % (meta["particles"], meta["steps"]),
        "",
        "Command (from the repository root on the GPU host):",
        "",
        "    " + command,
        "",
        "Scene: slope slab %s m, %d particles, h = 1/256, dt = 5e-4, per-particle E0 / xi (seed 2), 30-degree"
        % (" x ".join("%.3g" % x for x in meta["size_m"]), meta["particles"]),
        "slope with friction 0.2, CR tolerance 1e-5. Both runs start from the same particles and step on their",
        "own. The FP64 run is mpm_sim64_advance: the reference's step (Sim.cpp:57-134) with every scalar in",
        "double on the GPU (its parity with the FP64 oracle: tests/test_gpu_fp64.py). Errors of the FP32-bulk",
        "step (mpm_sim_advance: FP32 bulk, FP64 CR reductions) against it: momentum = |sum m v - sum m v64| /",
        "|sum m v64| (the particles' total momentum, which the transfers carry to and from the grid); x = per",
        "particle the largest per-axis position error in units of h; F = per particle the largest elementwise",
        "error; max and 99th percentile over particles. %.0f s in all." % meta["seconds"],
        "",
        "| step | momentum | x max (h) | x p99 (h) | F max | F p99 | CR its FP32 / FP64 | bodies FP32 / FP64 |",
        "|---|---|---|---|---|---|---|---|",
    ]
    for st in marks:
        r = recs[st - 1]
        lines.append("| %d | %.2e | %.2e | %.2e | %.2e | %.2e | %d / %d | %d / %d |" % (
            st, r["gpu_mom_rel"], r["gpu_x_max"], r["gpu_x_p99"], r["gpu_F_max"], r["gpu_F_p99"], r["its_gpu"],
            r["its_fp64"], r["bodies_gpu"], r["bodies_fp64"]))
    worst = {q: max(r["gpu_" + q] for r in recs) for q in QUANTITIES}
    lines.append("| max over steps | %.2e | %.2e | %.2e | %.2e | %.2e | | |" % tuple(worst[q] for q in QUANTITIES))
    its = np.array([[r["its_gpu"], r["its_fp64"]] for r in recs], dtype=float)
    t32 = np.array([r["s_fp32"] for r in recs])
    t64 = np.array([r["s_fp64"] for r in recs])
    lines += ["", "CR iterations of the largest body, mean over steps: FP32-bulk %.1f, FP64 %.1f. Wall time per step"
              " (median): FP32-bulk %.1f ms, FP64 %.0f ms." % (its[:, 0].mean(), its[:, 1].mean(),
                                                               1e3 * np.median(t32), 1e3 * np.median(t64))]
    summary = {"workload": "slope_slab_precision_gpu64", "particles": meta["particles"], "steps": meta["steps"],
               "final": {k: v for k, v in recs[-1].items() if k not in ("step", "t")},
               "max_over_steps": {"gpu": worst}, "its_mean": {"fp32": its[:, 0].mean(), "fp64": its[:, 1].mean()},
               "ms_per_step_median": {"fp32": 1e3 * np.median(t32), "fp64": 1e3 * np.median(t64)},
               "seconds": meta["seconds"]}
    lines += ["", "<!-- json: %s -->" % json.dumps(summary)]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=60000)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--csv", default=None)
    ap.add_argument("--out", required=True)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--from-csv", default=None, help="summarise a per-step csv of an earlier run instead")
    ap.add_argument("--gpu64", action="store_true",
                    help="the full-FP64 run on the GPU (mpm_sim64) instead of the CPU oracle, at the 4M slab "
                         "(--n sets the particle count of the 0.8 x 0.2 x 0.4 m box)")
    a = ap.parse_args()
    if a.gpu64:
        recs, meta = run_study_gpu64(a.n, a.steps, a.csv, log=lambda s: print(s, flush=True))
        cmd = "python tests/precision_study.py --gpu64 --n %d --steps %d --csv <per-step csv> --out <this file>" % (
            a.n, a.steps)
        print(json.dumps(summarize_gpu64(recs, meta, a.out, cmd)))
        return
    if a.from_csv:
        recs, meta = read_csv(a.from_csv, a.n)
        cmd = "python tests/precision_study.py --n %d --steps %d --csv <per-step csv> --out <this file>" % (
            a.n, len(recs))
        print(json.dumps(summarize(recs, meta, a.out, cmd)))
        return
    recs, meta = run_study(a.n, a.steps, a.csv, fp32=not a.no_fp32, log=lambda s: print(s, flush=True))
    cmd = "python tests/precision_study.py --n %d --steps %d --csv <per-step csv> --out <this file>%s" % (
        a.n, a.steps, " --no-fp32" if a.no_fp32 else "")
    print(json.dumps(summarize(recs, meta, a.out, cmd)))


if __name__ == "__main__":
    main()
