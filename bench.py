"""Benchmark: implicit MPM snow step on one MI355X, learned-preconditioned CR against plain CR.

Workload (the metric's configuration): snow slab on a 30-degree slope, 4M particles seeded
uniform in a 0.8 x 0.2 x 0.4 m box, per-particle E0 log-uniform in [7e4, 2.8e5] Pa and xi uniform
in [5, 10] (seed 2, held out from the preconditioner's training scenes), h = 1/256 m, dt = 5e-4 s,
nu = 0.2, theta_c = 2.5e-2, theta_s = 7.5e-3, tilted gravity over the ground plane (mu = 0.2), CR
relative tolerance 1e-5; synthetic seeded particles. A step is one Sim::advance of the whole
scene (body partition, per-body grids, implicit CR solve, G2P, plasticity, advection). After
--warmup plain steps the particle state is kept; --steps steps are timed with the learned
flexible CR (BF16 MFMA network, weights mpmsnow_amd/lpc_weights.npy) and, from the same kept
state, --steps steps with plain diagonal CR. Prints one JSON line.

value = particle-steps/s of the learned-preconditioned step (inputs resident in HBM); `plain`
and `speedup` beside it; CR iterations per step of the largest body for both.
roofline = the dominant kernel (the matrix-free HVP inside CR) in the learned leg: algorithmic
bytes per launch (172 B per particle + 29 B per active node of the bodies still iterating) /
mean launch time, both measured in the timed steps with HIP events on the library's stream
(one HVP launch in 7), against MI355X's 8 TB/s HBM peak. roofline.traffic = measured HBM bytes
per HVP launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload.
network = the preconditioner's forward pass (MFMA BF16): mean time per forward (HIP events), an
upper bound of its FLOPs from the layer shapes, and the MFMA FLOPs that rocprofv3 counted
(committed summary). precision_study = the committed 300-step FP32-vs-FP64 study of the slab (at 60k
particles against the FP64 oracle); precision_study_4M = the same at 4M particles against the
full-FP64 GPU step (mpm_sim64).
cpu_baseline = the reference-faithful CPU oracle (float, OpenMP over the 8-colour voxel
partition like the reference's TBB splats) timed on this host for a bounded sample: a 500k-
particle slab of the same density, materials, h and dt (0.2 m cube) from the GPU's state after
the same warm-up, 2 warm-up steps excluded, the median of 3 timed steps. error_vs_fp64 steps
that sample's state once on the GPU and in the FP64 oracle and compares with the FP32 oracle's
first step.
two_snowballs_1M = the two-snowballs scene (1M particles, +/-3 m/s, h = 1/256, dt = 5e-4, plain
CR), particle-steps/s over --snowball-steps steps.

Other modes:
  --stiff              only the stiff stress case (16M particles in 8 seeded boxes and
                       spheres dropped onto the ground plane and an analytic sphere
                       collider, E0 = 4.8e5 Pa, h = 1/512, dt = 2e-3): plain vs learned CR
                       from the same warmed-up state; one JSON line (--stiff-particles,
                       --stiff-steps, --stiff-warmup)
  --stiff-summary J    the plain vs learned table of a --stiff run's JSON line J
                       (profiles/stiff_drop_16M_plain_vs_learned.md)
  --lpc-weights W, --lpc-fallback R, --lpc-dirs M
                       network parameters (.npy), fallback ratio and flexible-CR
                       direction count of the learned legs (defaults: the committed
                       weights, 0.95, 2)
  --copy-calibration   also runs one 256 MiB device-to-device copy first (the
                       known-size kernel that scales the PMC byte counters)
  --pmc-summary FETCH_CSV WRITE_CSV, --sq-summary P0 P1 P2, --trace-summary STATS_CSV LOG,
  --mfma-summary PMC_CSV TRACE_CSV
                       summarise rocprofv3 counter / kernel-trace runs of bench.py into the
                       profiles/ files of the --workload they ran (slab, or snowballs for the
                       two-snowballs scene run alone with --workload snowballs)
"""
import argparse
import csv
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense BF16 MFMA
HVP_SAMPLE = 7  # HVP launches timed with HIP events in the timed region: one in HVP_SAMPLE
COPY_BYTES = 256 << 20
# rocprofv3 summaries per workload (bench.py --workload W runs W alone)
WORKLOAD_NAMES = {"slab": "slope_slab_4M", "snowballs": "two_snowballs_1M"}
WORKLOAD_TITLES = {"slab": "slope slab (4M particles, learned and plain CR legs)",
                   "snowballs": "two snowballs (1M particles, plain CR)"}


def profile_file(kind, workload):
    name = WORKLOAD_NAMES[workload]
    return os.path.join(ROOT, "profiles", {"traffic": "hvp_%s_traffic.md", "sq": "hvp_%s_sq_counters.md",
                                           "stats": "%s_kernel_stats.md"}[kind] % name)


def progress(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def step_bytes(n_p, n_n, iters):
    """Algorithmic bytes of one step: N_p (692 + 172 (I+3)) + N_n (117 + 29 (I+3) + 120 (I+1))."""
    return n_p * (692 + 172 * (iters + 3)) + n_n * (117 + 29 * (iters + 3) + 120 * (iters + 1))


STATE_FIELDS = ("p", "v", "F", "volume", "Fp", "FinvTrans", "R", "Ginv", "J", "mu", "lambda")


def snapshot(sim):
    return {k: sim.get(k) for k in STATE_FIELDS}


def restore(sim, state):
    for k in STATE_FIELDS:
        sim.ctx.set(k, state[k])


def oracle_sim(sc, state, prec="float"):
    """The CPU oracle's Sim on the scene's colliders, gravity and per-particle material, at a state
    of the GPU run."""
    from oracle import pyoracle as O

    mat = sc["material"]
    cols = O.Colliders.create(prec)
    for c, v, mu, st in sc["planes"]:
        cols.add_plane(c, v=v, friction=mu, sticky=st)
    f = O.Fields.create(prec)
    f.add_gravity(sc["gravity"])
    model = O.Model.snow(mat["E"], mat["nu"], mat["xi"], mat["theta_c"], mat["theta_s"], prec)
    extra = {}
    if "mu0" in sc:
        extra = dict(mu0=sc["mu0"], lambda0=sc["lambda0"], xi=sc["xi"])
    sim = O.Sim(state["p"], sc["m"], sc["h"], model, cols, f, prec=prec, **extra)
    for name in STATE_FIELDS[1:]:
        sim.pd.set(name, state[name])
    sim._keep = (cols, f, model)
    return sim


def gpu_sim(sc):
    import mpmsnow_amd as G

    extra = {}
    if "mu0" in sc:
        extra = dict(mu0=sc["mu0"], lambda0=sc["lambda0"], xi=sc["xi"])
    return G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"], gravity=[sc["gravity"]],
                 v=sc.get("v"), **extra)


def cpu_leg(warmup, n=500000, cpu_warmup=2, cpu_timed=3):
    """cpu_baseline and error_vs_fp64 on a bounded sample of the workload: a slab of the slope-slab
    scene's density, materials, h and dt (n = 500k: box 0.2 x 0.2 x 0.2 m, seed 2), stepped
    `warmup` times on the GPU. From that state the FP32 oracle takes cpu_warmup untimed steps and
    cpu_timed timed ones (the median is the baseline); its first step is compared, with the FP64
    oracle's and the GPU's first step from the same state, against FP64."""
    import numpy as np

    from mpmsnow_amd import scenes
    from oracle import pyoracle as O

    side = 0.2 * (n / 500000.0) ** (1.0 / 3.0)
    sc = scenes.slope_slab(n=n, seed=2, size=(side, side, side))
    n = len(sc["m"])
    sim = gpu_sim(sc)
    for _ in range(warmup):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    state = snapshot(sim)
    o32 = oracle_sim(sc, state)
    progress("cpu baseline: %d-particle slab sample, oracle set up" % n)
    times, its = [], []
    first32 = None
    for k in range(cpu_warmup + cpu_timed):
        t0 = time.perf_counter()
        o32.advance(sc["dt"], sc["max_iters"], sc["tol"])
        dt = time.perf_counter() - t0
        if k == 0:
            first32 = {"p": o32.pd.get("p").reshape(n, 3).copy(), "F": o32.pd.get("F").reshape(n, 9).copy()}
        if k >= cpu_warmup:
            times.append(dt)
            its.append(max(s["iterations"] for s in o32.last_stats()))
        progress("cpu baseline: oracle step %d %.1f s%s" % (k, dt, "" if k >= cpu_warmup else " (warm-up)"))
    med = float(np.median(times))
    line = {"value": n / med, "unit": "particle-steps/s", "cores": O.num_threads(), "kind": "port",
            "sample": "FP32 oracle on a %d-particle slab of the workload's density, materials, h and dt (%.3g m cube, "
                      "seed 2) from the GPU's state after %d steps: %d warm-up steps excluded, median of %d timed "
                      "steps (%s s; %s CR iterations for the largest body)"
                      % (n, side, warmup, cpu_warmup, cpu_timed, ", ".join("%.1f" % t for t in times),
                         ", ".join(str(i) for i in its))}
    progress("error check: FP64 oracle, 1 step")
    o64 = oracle_sim(sc, state, "double")
    o64.advance(sc["dt"], sc["max_iters"], sc["tol"])
    sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    out = {"steps": 1, "particles": n,
           "note": "one step from the CPU sample's state on the GPU, in the FP32 and in the FP64 oracle; "
                   "positions in units of h"}
    ref = {"p": o64.pd.get("p").reshape(n, 3), "F": o64.pd.get("F").reshape(n, 9)}
    for who, got in (("gpu", {"p": sim.get("p").reshape(n, 3), "F": sim.get("F").reshape(n, 9)}),
                     ("fp32_oracle", first32)):
        dx = np.abs(got["p"] - ref["p"]).max(1) / sc["h"]
        dF = np.abs(got["F"] - ref["F"]).max(1)
        out[who] = {"max_dx_over_h": float(dx.max()), "p99_dx_over_h": float(np.percentile(dx, 99)),
                    "max_dF": float(dF.max()), "p99_dF": float(np.percentile(dF, 99))}
    return line, out


def lpc_forward_flops(sim):
    """FLOPs of one network forward over the batch's active tiles: 2 x voxels x 27 x cin x cout per
    convolution (tile levels 0, 1, 2 of 64, 8 and 1 voxels per level-0 tile)."""
    lay = sim.grid_layout()
    t0 = lay["loop_tiles"]
    vox = (64 * t0, 8 * t0, t0)  # upper bounds at levels 1, 2 (their tiles cover 8 x 8 x 8 nodes)
    layers = ((0, 16, 16), (1, 16, 32), (2, 32, 64), (1, 64, 32), (0, 32, 16), (0, 16, 16))
    return sum(2.0 * vox[lv] * 27 * ci * co for lv, ci, co in layers)


def leg_marker(torch):
    """One small torch kernel just outside each leg's timed region (before and after it): nothing
    else in a slab run launches a torch kernel, so the rocprofv3 summaries find the learned leg's
    dispatches between the first two (`learned_window`)."""
    torch.arange(4, device="cuda")


def is_marker(name):
    return "at::" in name and not any(k in name for k in ("FillFunctor", "reduce_kernel", "copy", "Copy"))


def learned_window(rows):
    """(start, end) timestamps of the learned leg from (kernel name, start timestamp) pairs: the
    first two leg markers, else the first and last network dispatch (runs without markers)."""
    marks = sorted({t for n, t in rows if is_marker(n)})  # counter files hold a row per counter
    if len(marks) >= 2:
        return marks[0], marks[1]
    net = [t for n, t in rows if "k_conv" in n]
    return (min(net), max(net)) if net else None


def slab_legs(steps, warmup, n, weights=None, fallback=None, dirs=None, strikes=None):
    """Learned flexible CR and plain diagonal CR on the slope slab, --steps each, from one warmed-up
    state; HVP and network launches timed with HIP events."""
    import numpy as np
    import torch

    from mpmsnow_amd import scenes

    sc = scenes.slope_slab(n=n, seed=2)
    w = np.load(weights or os.path.join(ROOT, "mpmsnow_amd", "lpc_weights.npy"), allow_pickle=False)
    sim = gpu_sim(sc)
    sim.ctx.lpc_set_params(w)
    if fallback is not None:
        sim.ctx.lpc_fallback(fallback)
    if dirs is not None:
        sim.ctx.lpc_directions(dirs)
    if strikes is not None:
        sim.ctx.lpc_strikes(strikes)
    progress("slab: %d particles set up" % n)
    for _ in range(warmup):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    state = snapshot(sim)
    out = {}
    for mode, name in ((1, "learned"), (0, "plain")):
        restore(sim, state)
        sim.ctx.lpc_mode(mode)
        sim.profile(["hvp", "lpc_forward"], period=HVP_SAMPLE)
        its, tot, fbs = [], [], []
        leg_marker(torch)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
            rows = [sim.body_stats(b, hist_cap=0) for b in range(st["num_solved"])]
            big = max(rows, key=lambda r: r["particles"])
            its.append(big["iterations"])
            fbs.append(big["fallbacks"])
            tot.append(st["total_iterations"])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        leg_marker(torch)
        torch.cuda.synchronize()
        prof = sim.profile_query()
        sim.profile(False)
        hv = prof["hvp"]
        hms = hv["ms"] / max(1, hv["launches"])
        hby = hv["bytes"] / max(1, hv["launches"])
        leg = {"value": n * steps / dt, "ms_per_step": dt / steps * 1e3, "cr_iters_per_step": float(np.mean(its)),
               "cr_iters_per_step_all_bodies": float(np.mean(tot)), "fallbacks_per_step": float(np.mean(fbs)),
               "hvp": {"launches_timed": hv["launches"], "ms_per_launch": hms, "bytes_per_launch": hby,
                       "achieved_gbs": hby / (hms * 1e-3) / 1e9 if hms > 0 else 0.0}}
        if mode == 1:
            lp = prof["lpc_forward"]
            if lp["launches"]:
                fms = lp["ms"] / lp["launches"]
                fl = lpc_forward_flops(sim)
                leg["network"] = {"forwards_timed": lp["launches"], "ms_per_forward": fms,
                                  "flops_per_forward_upper_bound": fl,
                                  "tflops_upper_bound": fl / (fms * 1e-3) / 1e12,
                                  "note": "upper bound: 2 x voxels x 27 x cin x cout over every voxel of the "
                                          "loop's tiles at each level, over the whole forward (convolutions, "
                                          "pooling, features); the MFMA work counted by rocprofv3 is in "
                                          "mfma_counters"}
        progress("slab %s: %.2f ms/step, largest body %.1f CR iterations/step, fallbacks %.2f/step"
                 % (name, leg["ms_per_step"], leg["cr_iters_per_step"], leg["fallbacks_per_step"]))
        out[name] = leg
    return out, sc


def snowballs_leg(steps, warmup, n=1000000):
    """Two snowballs colliding, plain CR: particle-steps/s and the HVP's HIP-event rate."""
    import numpy as np
    import torch

    from mpmsnow_amd import scenes

    sc = scenes.two_snowballs(n=n, seed=1)
    sim = gpu_sim(sc)
    progress("two snowballs %d particles set up" % n)
    for _ in range(warmup):
        sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
    sim.profile(["hvp"], period=HVP_SAMPLE)
    stats = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(sim.advance(sc["dt"], sc["max_iters"], sc["tol"]))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    hv = sim.profile_query()["hvp"]
    sim.profile(False)
    hms = hv["ms"] / max(1, hv["launches"])
    hby = hv["bytes"] / max(1, hv["launches"])
    return {"value": n * steps / dt, "ms_per_step": dt / steps * 1e3, "steps": steps, "warmup": warmup,
            "cr_iters_per_step_all_bodies": float(np.mean([s["total_iterations"] for s in stats])),
            "bodies_per_step": float(np.mean([s["num_bodies"] for s in stats])),
            "hvp": {"launches_timed": hv["launches"], "ms_per_launch": hms, "bytes_per_launch": hby,
                    "achieved_gbs": hby / (hms * 1e-3) / 1e9 if hms > 0 else 0.0},
            "config": "two snowballs colliding: 1M particles, +/-3 m/s, h=1/256, dt=5e-4, CR tol 1e-5, plain CR, "
                      "%d steps after %d warm-up" % (steps, warmup)}


def stiff_leg(n, steps, warmup, seed=4, weights=None, fallback=None, dirs=None, plain_only=False):
    """Plain diagonal CR vs learned flexible CR on the stiff stress case (8 seeded boxes and spheres
    dropped onto the ground plane and an analytic sphere collider, E0 = 4.8e5 Pa, h = 1/512,
    dt = 2e-3), both from the state after `warmup` plain steps. Reports per-step CR iterations (sum
    and max over the step's bodies) and the learned solver's fallback triggers (iterations redone
    with the diagonal, summed over bodies)."""
    import numpy as np
    import torch

    import mpmsnow_amd as G
    from mpmsnow_amd import scenes

    sc = scenes.stiff_drop(n=n, seed=seed)
    w = np.load(weights or os.path.join(ROOT, "mpmsnow_amd", "lpc_weights.npy"), allow_pickle=False)
    out = {}
    modes = ((0, "plain"),) if plain_only else ((0, "plain"), (1, "learned"))
    for mode, name in modes:
        sim = G.Sim(sc["x"], sc["m"], sc["h"], material=sc["material"], planes=sc["planes"],
                    spheres=sc["spheres"], gravity=[sc["gravity"]], v=sc["v"])
        sim.ctx.lpc_set_params(w)
        if fallback is not None:
            sim.ctx.lpc_fallback(fallback)
        if dirs is not None:
            sim.ctx.lpc_directions(dirs)
        progress("stiff leg: %s, %d particles set up" % (name, n))
        for k in range(warmup):
            st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
            progress("stiff leg: %s warm-up step %d, %d bodies, CR iterations sum %d max %d"
                     % (name, k, st["num_bodies"], st["total_iterations"], st["max_iterations"]))
        sim.ctx.lpc_mode(mode)
        sim.profile(["hvp"], period=HVP_SAMPLE)
        rows = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            st = sim.advance(sc["dt"], sc["max_iters"], sc["tol"])
            torch.cuda.synchronize()
            te = time.perf_counter()
            bs = [sim.body_stats(b, hist_cap=0) for b in range(st["num_solved"])]
            rows.append({"ms": (te - ts) * 1e3, "bodies": st["num_bodies"], "iters_sum": st["total_iterations"],
                         "iters_max": st["max_iterations"], "fallbacks": int(sum(b["fallbacks"] for b in bs)),
                         "capped": int(sum(b["iterations"] >= sc["max_iters"] for b in bs))})
            progress("stiff leg: %s step %d: %.1f ms, %d bodies, CR iterations sum %d max %d, fallbacks %d"
                     % (name, len(rows), rows[-1]["ms"], st["num_bodies"], st["total_iterations"],
                        st["max_iterations"], rows[-1]["fallbacks"]))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        hv = sim.profile_query()["hvp"]
        sim.profile(False)
        hms = hv["ms"] / max(1, hv["launches"])
        hby = hv["bytes"] / max(1, hv["launches"])
        out[name] = {"value": n * steps / dt, "ms_per_step": dt / steps * 1e3,
                     "hvp": {"launches_timed": hv["launches"], "ms_per_launch": hms, "bytes_per_launch": hby,
                             "achieved_gbs": hby / (hms * 1e-3) / 1e9 if hms > 0 else 0.0},
                     "cr_iters_per_step_sum": float(np.mean([r["iters_sum"] for r in rows])),
                     "cr_iters_per_step_max": float(np.mean([r["iters_max"] for r in rows])),
                     "fallbacks_per_step": float(np.mean([r["fallbacks"] for r in rows])),
                     "capped_solves": int(sum(r["capped"] for r in rows)), "steps": rows}
        del sim
    if not plain_only:
        out["speedup"] = out["learned"]["value"] / out["plain"]["value"]
    out["config"] = ("stiff stress case: %d particles in 8 seeded boxes and spheres (seed %d), plane + analytic "
                     "sphere collider, E0=4.8e5, xi=10, h=1/512, dt=2e-3, CR tol 1e-5, %d steps after %d plain "
                     "warm-up steps" % (n, seed, steps, warmup))
    out["objects"] = sc["objects"]
    out["warmup"] = warmup
    return out


def _pmc_rows(path):
    """(kernel name, grid size, counter value, start timestamp) per dispatch of a rocprofv3
    counter_collection.csv."""
    rows = []
    with open(path) as fh:
        for r in csv.DictReader(fh):
            rows.append((r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"]),
                         int(r.get("Start_Timestamp") or 0)))
    return rows


def bench_command(workload, steps, warmup, extra=""):
    if workload == "slab":
        return "python bench.py --steps %d --warmup %d --no-cpu --no-snowballs%s" % (steps, warmup, extra)
    return "python bench.py --workload snowballs --steps %d --warmup %d%s" % (steps, warmup, extra)


def pmc_summary(fetch_csv, write_csv, out_path, workload="slab"):
    """HVP HBM traffic per launch from FETCH_SIZE / WRITE_SIZE runs, scaled by a known-size copy."""
    res = {}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = _pmc_rows(path)
        # calibration kernel of known traffic: the 256 MiB copy writes COPY_BYTES; the reduction
        # over its destination reads COPY_BYTES
        calib = "reduce_kernel" if counter == "FETCH_SIZE" else "copyBuffer"
        copies = [v for n, g, v, _ in rows if calib in n]
        if not copies:
            raise SystemExit("no %s dispatch in %s (run bench.py with --copy-calibration)" % (calib, path))
        copy_kib = max(copies)
        factor = copy_kib * 1024.0 / COPY_BYTES
        hvp = [(g, v, t) for n, g, v, t in rows if "k_hvp_pair<true>" in n]
        # launches of the steps after the warm-up (the store is body- and tile-sorted from the
        # first step on), full grid (every chunk launched): the median of those
        late = hvp[len(hvp) * 3 // 5:]
        gmax = max(g for g, _, _ in late)
        full = sorted(v for g, v, _ in late if g == gmax)
        kib = full[len(full) // 2]
        res[counter] = {"copy_kib": copy_kib, "factor": factor, "hvp_kib_per_launch": kib,
                        "hvp_bytes_per_launch": kib * 1024.0 / factor, "launches": len(full)}
        # every HVP dispatch of the learned leg (between the markers around its timed steps): the
        # population whose algorithmic bytes the bench line's roofline averages
        win = learned_window([(n, t) for n, _, _, t in rows])
        if win:
            lo, hi = win
            leg = [v for n, _, v, t in rows if "k_hvp" in n and lo <= t <= hi]
            if leg:
                res[counter]["leg_launches"] = len(leg)
                res[counter]["leg_bytes_per_launch"] = sum(leg) / len(leg) * 1024.0 / factor
    full_traffic = res["FETCH_SIZE"]["hvp_bytes_per_launch"] + res["WRITE_SIZE"]["hvp_bytes_per_launch"]
    leg = all("leg_bytes_per_launch" in res[k] for k in res)
    traffic = (res["FETCH_SIZE"]["leg_bytes_per_launch"] + res["WRITE_SIZE"]["leg_bytes_per_launch"] if leg
               else full_traffic)
    # algorithmic bytes per HVP launch of the profiled run itself (its bench line's HIP-event sample)
    algo = None
    log = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(fetch_csv))),
                       os.path.basename(os.path.dirname(os.path.abspath(fetch_csv))) + ".log")
    if os.path.exists(log):
        m = re.search(r'"roofline": (\{[^}]*\})', open(log).read())
        if m:
            algo = json.loads(m.group(1))["bytes_per_launch"]
    summary = {"kernel": "k_hvp_pair<true>",
               "workload": WORKLOAD_NAMES[workload], "traffic_bytes_per_launch": traffic,
               "traffic_full_grid_bytes_per_launch": full_traffic, "algorithmic_bytes_per_launch": algo,
               "fetch": res["FETCH_SIZE"], "write": res["WRITE_SIZE"]}
    cmd = bench_command(workload, 2, 3, " --copy-calibration")
    lines = [
        "# HVP HBM traffic, " + WORKLOAD_TITLES[workload],
        "",
        "Commands (from the repository root on the GPU host):",
        "",
        "    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <F> -o run -- " + cmd,
        "    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <W> -o run -- " + cmd,
        "    python bench.py --workload %s --pmc-summary <F>/run_counter_collection.csv "
        "<W>/run_counter_collection.csv" % workload,
        "",
        "The counters are scaled by what they read for kernels of known traffic on a %d MiB buffer: FETCH_SIZE by"
        % (COPY_BYTES >> 20),
        "torch's sum over it (reads it once), WRITE_SIZE by the device-to-device copy into it",
        "(`__amd_rocclr_copyBuffer`, writes it once): factor = counter KiB x 1024 / buffer bytes.",
        "HVP figures: median over the full-grid `k_hvp_pair<true>` (the CR loop's HVP)",
        "dispatches (every chunk launched) of the last 40% of HVP dispatches (the timed and table steps).",
        "",
        "| counter | calibration KiB | factor | HVP KiB/launch | HVP bytes/launch (scaled) |",
        "|---|---|---|---|---|",
    ]
    for k in ("FETCH_SIZE", "WRITE_SIZE"):
        r = res[k]
        lines.append("| %s | %.0f | %.4f | %.0f | %.4g |" % (k, r["copy_kib"], r["factor"], r["hvp_kib_per_launch"],
                                                          r["hvp_bytes_per_launch"]))
    lines += ["", "traffic (FETCH + WRITE) per full-grid `k_hvp_pair<true>` launch: %.4g bytes" % full_traffic]
    if leg:
        lines += ["", "Every HVP dispatch of the learned leg (between the markers around its timed steps; the "
                  "population the bench line's roofline averages): %d dispatches, FETCH %.4g + WRITE %.4g = "
                  "**%.4g bytes per launch**." % (res["FETCH_SIZE"]["leg_launches"],
                                                  res["FETCH_SIZE"]["leg_bytes_per_launch"],
                                                  res["WRITE_SIZE"]["leg_bytes_per_launch"], traffic)]
    if algo:
        lines += ["", "Algorithmic bytes per HVP launch of the same run (its HIP-event sample of %s): %.4g; "
                  "traffic / algorithmic = %.2f." % ("the learned leg" if workload == "slab" else "the timed steps",
                                                     algo, traffic / algo)]
    lines += ["", "<!-- json: %s -->" % json.dumps(summary)]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(json.dumps(summary))


SQ_PASSES = (
    "SQ_WAVES SQ_BUSY_CYCLES SQ_WAVE_CYCLES SQ_INSTS_VALU SQ_INSTS_LDS SQ_ACTIVE_INST_VALU SQ_ACTIVE_INST_LDS "
    "SQ_LDS_BANK_CONFLICT GRBM_GUI_ACTIVE",
    "SQ_LDS_ADDR_CONFLICT SQ_WAIT_INST_LDS SQ_WAIT_ANY SQ_WAIT_INST_ANY SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR "
    "SQ_ACTIVE_INST_ANY SQ_INSTS_SALU GRBM_GUI_ACTIVE",
    "SQ_LDS_UNALIGNED_STALL SQ_LDS_IDX_ACTIVE SQ_LDS_DATA_FIFO_FULL SQ_LDS_CMD_FIFO_FULL SQ_INST_LEVEL_LDS "
    "SQ_ACCUM_PREV_HIRES GRBM_GUI_ACTIVE",
)


def sq_summary(csv_paths, out_path, kernel="k_hvp_pair<true>", workload="slab"):
    """Mean per-dispatch SQ/LDS counters of the CR loop's HVP from the three `rocprofv3 --pmc` passes
    of SQ_PASSES, and what they imply on MI355X (256 CUs, 1024 SIMDs, 32 shader engines; a wave64
    VALU instruction occupies its SIMD for 4 clocks)."""
    vals = {}
    selection = "the last 40%% of the run's dispatches of that kernel"
    for path in csv_paths:
        per, start, names = {}, {}, []
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r.get("Start_Timestamp"):
                    names.append((r["Kernel_Name"], int(r["Start_Timestamp"])))
                if kernel not in r["Kernel_Name"]:
                    continue
                d = int(r.get("Dispatch_Id") or r["Correlation_Id"])
                per.setdefault(d, {})[r["Counter_Name"]] = float(r["Counter_Value"])
                if r.get("Start_Timestamp"):
                    start[d] = int(r["Start_Timestamp"])
        win = learned_window(names) if start else None
        if win:
            # the learned leg (between the markers around its timed steps), dispatches that found
            # work: a launch after every body of the batch has converged lasts a few percent of a
            # full one (its waves exit at the done flag)
            lo, hi = win
            ids = [d for d in sorted(per) if lo <= start.get(d, -1) <= hi]
            top = max((per[d].get("GRBM_GUI_ACTIVE", 0.0) for d in ids), default=0.0)
            ids = [d for d in ids if per[d].get("GRBM_GUI_ACTIVE", 0.0) >= 0.25 * top]
            selection = ("the learned leg's dispatches (between the markers around its timed steps) that found work "
                         "(GRBM_GUI_ACTIVE, in every pass, at least 25%% of the leg's longest)")
        else:
            ids = sorted(per)[len(per) * 3 // 5:]  # the steps after the warm-up
        if not ids:
            raise SystemExit("no %s dispatch in %s" % (kernel, path))
        for name in per[ids[0]]:
            if name not in vals:
                vals[name] = sum(per[d].get(name, 0.0) for d in ids) / len(ids)
    clocks = vals["SQ_BUSY_CYCLES"] / 32.0
    valu = vals["SQ_INSTS_VALU"] * 4.0 / 1024.0
    lds = vals["SQ_LDS_IDX_ACTIVE"] / 256.0
    bank = vals["SQ_LDS_BANK_CONFLICT"] / 256.0
    addr = vals["SQ_LDS_ADDR_CONFLICT"] / 256.0
    resident = vals["SQ_WAVE_CYCLES"] * 4.0 / 256.0 / clocks
    summary = {"kernel": kernel, "workload": WORKLOAD_NAMES[workload], "valu_issue_frac": round(valu / clocks, 3),
               "lds_active_frac": round(lds / clocks, 3), "resident_waves_per_cu": round(resident, 1)}
    lines = ["# HVP SQ / LDS counters, " + WORKLOAD_TITLES[workload], "",
             "Commands (from the repository root on the GPU host), one counter pass each:", ""]
    for k, p in enumerate(SQ_PASSES):
        lines.append("    rocprofv3 --pmc %s --output-format csv -d <P%d> -o run -- %s"
                     % (p, k, bench_command(workload, 2, 1)))
    lines += ["    python bench.py --workload %s --sq-summary <P0>/run_counter_collection.csv "
              "<P1>/run_counter_collection.csv <P2>/run_counter_collection.csv" % workload, "",
              ("Values: mean per `%s` dispatch over " + selection + ", summed over the device as rocprofv3 reports "
               "them.") % kernel, "",
              "| counter | per dispatch |", "|---|---|"]
    for name in sorted(vals):
        lines.append("| %s | %.4g |" % (name, vals[name]))
    lines += ["", "Derived (256 CUs, 1024 SIMDs, 32 shader engines; a wave64 VALU instruction occupies its SIMD "
              "for 4 clocks; SQ_WAVE_CYCLES read as counting 4-clock units):", "",
              "* kernel length = SQ_BUSY_CYCLES / 32 SEs = %.0f k clocks" % (clocks / 1e3),
              "* VALU issue: SQ_INSTS_VALU x 4 clocks / 1024 SIMDs = %.0f k clocks = %.0f%% of the kernel"
              % (valu / 1e3, 100 * valu / clocks),
              "* LDS active per CU: SQ_LDS_IDX_ACTIVE / 256 = %.0f k clocks = %.0f%%; of it bank conflicts %.0f k "
              "(%.0f%%), address conflicts %.0f k (%.0f%%)" % (lds / 1e3, 100 * lds / clocks, bank / 1e3,
                                                              100 * bank / clocks, addr / 1e3, 100 * addr / clocks),
              "* resident waves per CU: SQ_WAVE_CYCLES x 4 / 256 / kernel clocks = %.1f" % resident,
              "* VALU instructions per wave: %.0f" % (vals["SQ_INSTS_VALU"] / vals["SQ_WAVES"]),
              "", "<!-- json: %s -->" % json.dumps(summary)]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(json.dumps(summary))


MFMA_PASS = ("SQ_INSTS_VALU_MFMA_MOPS_BF16 SQ_INSTS_VALU_MFMA_BF16 SQ_VALU_MFMA_BUSY_CYCLES SQ_BUSY_CYCLES SQ_WAVES "
             "GRBM_GUI_ACTIVE")


def mfma_summary(pmc_csv, trace_csv, out_path):
    """The network's BF16 MFMA convolutions (`k_conv<CIN, COUT, MODE, WAVES>`) on the slope slab's learned
    leg: per kernel, MFMA FLOPs per dispatch from SQ_INSTS_VALU_MFMA_MOPS_BF16 (units of 512 FLOPs)
    and, with each kernel's mean duration from a kernel trace of the same workload, the FLOP rate
    against 2.5 PFLOP/s and the MFMA pipe's busy fraction: SQ_VALU_MFMA_BUSY_CYCLES (summed over the
    1024 SIMDs; a 16x16x32 BF16 MFMA counts 16) / 1024 / (duration x 2.4 GHz). The counter run's
    own kernels run about 2x longer than in the trace (GRBM_GUI_ACTIVE, summed over the 8 XCDs, / 8
    is their length in clocks), so its clocks are not used."""
    per = {}
    with open(pmc_csv) as fh:
        for r in csv.DictReader(fh):
            nm = _short(r["Kernel_Name"])
            if not nm.startswith("k_conv"):
                continue
            d = int(r.get("Dispatch_Id") or r["Correlation_Id"])
            per.setdefault(nm, {}).setdefault(d, {})[r["Counter_Name"]] = float(r["Counter_Value"])
    if not per:
        raise SystemExit("no k_conv dispatch in %s" % pmc_csv)
    dur = {}
    with open(trace_csv) as fh:
        for r in csv.DictReader(fh):
            nm = _short(r["Kernel_Name"])
            if nm.startswith("k_conv"):
                dur.setdefault(nm, []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    rows, tot_fl, tot_ns = [], 0.0, 0.0
    for nm in sorted(per):
        ds = per[nm].values()
        mean = {k: sum(d.get(k, 0.0) for d in ds) / len(ds) for k in next(iter(ds))}
        fl = mean.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0) * 512.0
        ns = sum(dur.get(nm, [0])) / max(1, len(dur.get(nm, [])))
        # MFMA busy cycles per SIMD over the kernel's clocks in the trace run (2.4 GHz)
        busy = mean.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / 1024.0 / max(1.0, ns * 2.4)
        rate = fl / (ns * 1e-9) / 1e12 if ns else 0.0
        rows.append((nm, len(ds), fl, busy, ns, rate))
        tot_fl += fl
        tot_ns += ns
    lines = ["# BF16 MFMA convolutions of the learned preconditioner, " + WORKLOAD_TITLES["slab"], "",
             "Commands (from the repository root on the GPU host):", "",
             "    rocprofv3 --pmc %s --output-format csv -d <M> -o run -- %s" % (MFMA_PASS,
                                                                              bench_command("slab", 2, 3)),
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <D> -o run -- %s" % bench_command("slab", 10, 5),
             "    python bench.py --mfma-summary <M>/run_counter_collection.csv <D>/run_kernel_trace.csv", "",
             "Per kernel: MFMA FLOPs per dispatch = SQ_INSTS_VALU_MFMA_MOPS_BF16 x 512 (mean over the counter run's "
             "dispatches); mean duration from the kernel trace (every dispatch of the kernel); rate = FLOPs / "
             "duration against %.0f TFLOP/s dense BF16; MFMA busy = SQ_VALU_MFMA_BUSY_CYCLES / 1024 SIMDs / "
             "(duration x 2.4 GHz) (the counter sums the SIMDs' busy cycles, 16 per 16x16x32 BF16 MFMA, so it "
             "matches the rate's fraction). The counter run's kernels take about 2x their trace duration (its "
             "GRBM_GUI_ACTIVE, summed over the 8 XCDs, / 8), so its clocks are not used." % MFMA_BF16_PEAK_TFLOPS, "",
             "| kernel | dispatches (counters) | MFMA GFLOP / dispatch | MFMA busy | mean us | TFLOP/s | of peak |",
             "|---|---|---|---|---|---|---|"]
    for nm, nd, fl, busy, ns, rate in rows:
        lines.append("| `%s` | %d | %.2f | %.1f%% | %.1f | %.1f | %.1f%% |" % (
            nm, nd, fl / 1e9, 100 * busy, ns / 1e3, rate, 100 * rate / MFMA_BF16_PEAK_TFLOPS))
    agg = tot_fl / (tot_ns * 1e-9) / 1e12 if tot_ns else 0.0
    lines += ["", "All convolutions of one forward: %.1f GFLOP of MFMA work in %.1f us of kernel time: %.1f TFLOP/s, "
              "%.1f%% of the dense BF16 peak." % (tot_fl / 1e9, tot_ns / 1e3, agg, 100 * agg / MFMA_BF16_PEAK_TFLOPS)]
    summary = {"workload": WORKLOAD_NAMES["slab"], "conv_gflop_per_forward": tot_fl / 1e9,
               "conv_us_per_forward": tot_ns / 1e3, "tflops": agg, "frac_of_bf16_peak": agg / MFMA_BF16_PEAK_TFLOPS,
               "kernels": {nm: {"gflop": fl / 1e9, "mfma_busy": busy, "us": ns / 1e3, "tflops": rate}
                           for nm, nd, fl, busy, ns, rate in rows}}
    lines += ["", "<!-- json: %s -->" % json.dumps(summary)]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(json.dumps(summary))


def _short(name):
    name = name.replace("mpm::(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0]


def trace_summary(stats_csv, bench_log, out_path, workload="slab"):
    """Kernel table of a rocprofv3 --kernel-trace --stats run of the bench, and the HVP check."""
    rows = []
    with open(stats_csv) as fh:
        for r in csv.DictReader(fh):
            rows.append((_short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["Percentage"])))
    rows.sort(key=lambda r: -r[2])
    bench = None
    for line in open(bench_log):
        line = line.strip()
        if line.startswith("{") and '"metric"' in line:
            bench = json.loads(line)
    if bench is None:
        raise SystemExit("no bench JSON line in %s" % bench_log)
    hvp_calls = sum(c for n, c, t, _ in rows if n.startswith("k_hvp"))
    hvp_ns = sum(t for n, c, t, _ in rows if n.startswith("k_hvp"))
    rp_ms = hvp_ns / hvp_calls / 1e6
    ev_ms = bench["roofline"]["ms_per_launch"]
    steps, warmup = bench["steps"], bench["warmup"]
    # the HVP dispatches of the learned leg alone, whose launches bench.py's HIP events sample: from
    # the kernel trace, those between the leg's first and last network dispatch (the warm-up and
    # the plain leg run no network)
    timed = None
    trace_csv = os.path.join(os.path.dirname(stats_csv), os.path.basename(stats_csv).replace("kernel_stats",
                                                                                            "kernel_trace"))
    if os.path.exists(trace_csv):
        hv, names = [], []
        with open(trace_csv) as fh:
            for r in csv.DictReader(fh):
                nm = _short(r["Kernel_Name"])
                be = (int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
                names.append((r["Kernel_Name"], be[0]))
                if nm.startswith("k_hvp"):
                    hv.append(be)
        win = learned_window(names)
        if win:
            hv = [(b, e) for b, e in hv if win[0] <= b <= win[1]]
        else:
            hv.sort()
            hv = hv[-HVP_SAMPLE * bench["roofline"]["launches"]:]
        k = len(hv)
        if k:
            timed = {"launches": k, "ms": sum(e - b for b, e in hv[-k:]) / k / 1e6}
    legs = ("the learned and the plain leg (each --steps steps from the kept state)" if workload == "slab"
            else "the timed steps")
    lines = [
        "# Kernel statistics, " + WORKLOAD_TITLES[workload],
        "",
        "Command (from the repository root on the GPU host):",
        "",
        "    rocprofv3 --kernel-trace --stats --output-format csv -d <D> -o run -- %s > <LOG>"
        % bench_command(workload, steps, warmup),
        "    python bench.py --workload %s --trace-summary <D>/run_kernel_stats.csv <LOG>" % workload,
        "",
        "The profiled run covers %d warm-up steps and %s. bench.py line of that run: %.4g particle-steps/s, "
        "%.3f ms/step (under the profiler)." % (warmup, legs, bench["value"], bench["ms_per_step"]),
        "",
        "HVP (`k_hvp_pair<true>`, the CR loop's, and `k_hvp<false>`, every launch of the run): %d launches, "
        "mean %.4f ms (rocprofv3); bench.py's HIP events over its timed steps (one HVP launch in %d): %d "
        "launches, mean %.4f ms (ratio %.3f)."
        % (hvp_calls, rp_ms, HVP_SAMPLE, bench["roofline"]["launches"], ev_ms,
           rp_ms / ev_ms if ev_ms else float("nan")),
        "",
    ]
    if timed:
        which, rest = (("the learned leg alone", "the warm-up, whose first step solves with the particle store "
                        "still in input order and runs to the iteration cap, and the plain leg")
                       if workload == "slab" else
                       ("the timed steps alone", "the warm-up, whose first step solves with the particle store "
                        "still in input order"))
        lines += ["HVP dispatches of %s (%d, those between the markers bench.py launches around its timed steps, "
                  "in `run_kernel_trace.csv`): mean %.4f ms (rocprofv3) against the HIP events' %.4f ms (ratio %.3f). "
                  "The run's mean above also covers %s."
                  % (which, timed["launches"], timed["ms"], ev_ms, timed["ms"] / ev_ms if ev_ms else float("nan"),
                     rest), ""]
    lines += [
        "| kernel | calls | total ms | mean us | share |",
        "|---|---|---|---|---|",
    ]
    for n, c, t, pct in rows[:20]:
        lines.append("| `%s` | %d | %.2f | %.1f | %.1f%% |" % (n, c, t / 1e6, t / c / 1e3, pct))
    summary = {"workload": WORKLOAD_NAMES[workload], "hvp_launches": hvp_calls, "hvp_ms_rocprof": rp_ms,
               "hvp_ms_events": ev_ms, "steps": steps, "warmup": warmup,
               "hvp_ms_rocprof_timed": timed["ms"] if timed else None}
    lines += ["", "<!-- json: %s -->" % json.dumps(summary)]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(json.dumps(summary))


STIFF_STATS_FILE = os.path.join(ROOT, "profiles", "stiff_drop_16M_kernel_stats.md")
MFMA_FILE = os.path.join(ROOT, "profiles", "lpc_conv_mfma_slope_slab_4M.md")
PRECISION_FILE = os.path.join(ROOT, "profiles", "slope_slab_precision_300_steps.md")
PRECISION_GPU64_FILE = os.path.join(ROOT, "profiles", "slope_slab_4M_precision_fp32_vs_fp64_gpu.md")


def stiff_trace_summary(stats_csv, stiff_json, out_path):
    """Kernel table of a rocprofv3 --kernel-trace --stats run of `bench.py --stiff --stiff-plain-only`, with
    the HVP's algorithmic rate from the run's own HIP-event sample."""
    rows = []
    with open(stats_csv) as fh:
        for r in csv.DictReader(fh):
            rows.append((_short(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]), float(r["Percentage"])))
    rows.sort(key=lambda r: -r[2])
    d = None
    for line in open(stiff_json):
        if line.strip().startswith("{"):
            d = json.loads(line)
    if d is None:
        raise SystemExit("no JSON line in %s" % stiff_json)
    p = d["plain"]
    hvp_calls = sum(c for n, c, t, _ in rows if n.startswith("k_hvp"))
    hvp_ns = sum(t for n, c, t, _ in rows if n.startswith("k_hvp"))
    lines = ["# Kernel statistics, stiff stress case (16M particles, plain CR)", "",
             "Command (from the repository root on the GPU host):", "",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d <D> -o run -- python bench.py --stiff "
             "--stiff-plain-only --stiff-steps %d --stiff-warmup %d > <JSON>" % (len(p["steps"]), d["warmup"]),
             "    python bench.py --stiff-trace-summary <D>/run_kernel_stats.csv <JSON>", "",
             "Workload: " + d["config"] + ".", "",
             "Timed steps under the profiler: %.1f ms/step, %.3g particle-steps/s, CR iterations per step %.0f "
             "(sum over bodies), %.0f (slowest body)." % (p["ms_per_step"], p["value"], p["cr_iters_per_step_sum"],
                                                        p["cr_iters_per_step_max"]), "",
             "HVP (`k_hvp*`, every launch of the run, warm-up included): %d launches, mean %.1f us (rocprofv3). "
             "bench.py's HIP events over the timed steps (one launch in %d): %d launches, mean %.1f us, %.3g "
             "algorithmic bytes per launch (172 B per particle + 29 B per active node of the bodies still "
             "iterating), %.0f GB/s = %.1f%% of 8 TB/s."
             % (hvp_calls, hvp_ns / max(1, hvp_calls) / 1e3, HVP_SAMPLE, p["hvp"]["launches_timed"],
                p["hvp"]["ms_per_launch"] * 1e3, p["hvp"]["bytes_per_launch"], p["hvp"]["achieved_gbs"],
                100 * p["hvp"]["achieved_gbs"] / HBM_PEAK_GBS), "",
             "| kernel | calls | total ms | mean us | share |", "|---|---|---|---|---|"]
    for n, c, t, pct in rows[:20]:
        lines.append("| `%s` | %d | %.2f | %.1f | %.1f%% |" % (n, c, t / 1e6, t / c / 1e3, pct))
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


STIFF_COMPARE_FILE = os.path.join(ROOT, "profiles", "stiff_drop_16M_plain_vs_learned.md")


def stiff_summary(stiff_json, out_path):
    """Plain vs learned table of a `bench.py --stiff` run (its JSON line): per range of steps, before
    and from the first step whose body count differs from step 0's in either run (the onset of
    fragmentation), the throughput, CR iterations, fallback triggers and capped solves, and a
    per-step table (every 4th step, and the last 24)."""
    d = None
    for line in open(stiff_json):
        if line.strip().startswith("{"):
            d = json.loads(line)
    if d is None or "plain" not in d or "learned" not in d:
        raise SystemExit("no plain + learned JSON line in %s" % stiff_json)
    legs = ("plain", "learned")
    steps = {k: d[k]["steps"] for k in legs}
    ns = min(len(steps[k]) for k in legs)
    n = int(re.search(r"(\d+) particles", d["config"]).group(1))
    split = ns
    for i in range(ns):
        if any(steps[k][i]["bodies"] != steps[k][0]["bodies"] for k in legs):
            split = i
            break

    def agg(k, lo, hi):
        rows = steps[k][lo:hi]
        ms = sum(r["ms"] for r in rows)
        m = len(rows)
        return ("| %s | %d-%d | %.3g | %.1f | %.0f | %.0f | %.1f | %d | %d |"
                % (k, lo, hi - 1, n * m / (ms * 1e-3), ms / m, sum(r["iters_sum"] for r in rows) / m,
                   sum(r["iters_max"] for r in rows) / m, sum(r["fallbacks"] for r in rows) / m,
                   sum(r["capped"] for r in rows), max(r["bodies"] for r in rows)), n * m / (ms * 1e-3))

    ranges = [(0, split), (split, ns), (0, ns)] if 0 < split < ns else [(0, ns)]
    table, rate = [], {}
    for lo, hi in ranges:
        for k in legs:
            line, v = agg(k, lo, hi)
            table.append(line)
            rate[(k, lo, hi)] = v
    ratios = ["%.2f over steps %d-%d" % (rate[("learned", lo, hi)] / rate[("plain", lo, hi)], lo, hi - 1)
              for lo, hi in ranges]
    first_cap = {k: next((i for i, r in enumerate(steps[k]) if r["capped"] > 0), None) for k in legs}
    lines = ["# Stiff stress case, 16M particles, %d steps: plain diagonal CR vs learned flexible CR (MI355X)" % ns,
             "", "Commands (repository root):", "",
             "    python -u bench.py --stiff --stiff-steps %d --stiff-warmup %d > <JSON>" % (ns, d["warmup"]),
             "    python bench.py --stiff-summary <JSON>", "",
             "Workload: %s. Each solver runs its own steps from the seeded initial state; ms are wall time per "
             "`Sim::advance` (synchronised)." % d["config"], "",
             "Objects (seed 4): " + "; ".join("%s %.3fx%.3fx%.3f m at (%.3f, %.3f, %.3f), %d particles"
                                              % (o["kind"], *o["size"], *o["center"], o["count"])
                                              for o in d["objects"]) + ".", "",
             "| Solver | steps | particle-steps/s | ms/step | CR iterations/step (sum over bodies) | max over bodies "
             "| fallback triggers/step | capped solves | max bodies |",
             "|---|---|---|---|---|---|---|---|---|"] + table + [
             "", "Learned / plain particle-steps/s: " + ", ".join(ratios) + ".", "",
             "First step with a solve at the iteration cap: plain %s, learned %s. Body count first differs from "
             "step 0's at step %s." % (first_cap["plain"], first_cap["learned"], split if split < ns else "none"),
             "", "Per step (every 4th step, and the last 24):", "",
             "| step | plain ms | bodies | it sum | it max | capped | learned ms | bodies | it sum | it max | "
             "fallbacks | capped |", "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for i in range(ns):
        if i % 4 and i < ns - 24:
            continue
        a, b = steps["plain"][i], steps["learned"][i]
        lines.append("| %d | %.0f | %d | %d | %d | %d | %.0f | %d | %d | %d | %d | %d |"
                     % (i, a["ms"], a["bodies"], a["iters_sum"], a["iters_max"], a["capped"], b["ms"], b["bodies"],
                        b["iters_sum"], b["iters_max"], b["fallbacks"], b["capped"]))
    summary = {"plain": d["plain"]["value"], "learned": d["learned"]["value"], "split": split,
               "ratio_before_split": rate[("learned", 0, split)] / rate[("plain", 0, split)] if 0 < split < ns
               else rate[("learned", 0, ns)] / rate[("plain", 0, ns)]}
    lines += ["", "<!-- json: %s -->" % json.dumps(summary)]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    print(json.dumps(summary))


def read_footer(path):
    """The JSON footer of a committed profiles/ summary, or None."""
    if not os.path.exists(path):
        return None
    m = re.search(r"<!-- json: (.*) -->", open(path).read())
    return json.loads(m.group(1)) if m else None


def read_sq(workload="slab"):
    """VALU-issue and LDS-active fractions of the HVP's clocks from the committed SQ counter summary."""
    path = profile_file("sq", workload)
    if not os.path.exists(path):
        return None
    m = re.search(r"<!-- json: (.*) -->", open(path).read())
    return json.loads(m.group(1)) if m else None


def read_traffic(workload="slab"):
    path = profile_file("traffic", workload)
    if not os.path.exists(path):
        return None
    m = re.search(r"<!-- json: (.*) -->", open(path).read())
    return json.loads(m.group(1))["traffic_bytes_per_launch"] if m else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="timed steps of each leg")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=("slab", "snowballs"), default="slab",
                    help="slab: the headline line (learned vs plain on the 4M slope slab); snowballs: the 1M "
                         "two-snowball scene alone")
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-snowballs", action="store_true", help="skip the two-snowballs line")
    ap.add_argument("--snowball-steps", type=int, default=60)
    ap.add_argument("--stiff", action="store_true",
                    help="run only the stiff stress case (plain vs learned CR) and print its JSON line")
    ap.add_argument("--stiff-particles", type=int, default=16000000)
    ap.add_argument("--stiff-steps", type=int, default=5)
    ap.add_argument("--stiff-warmup", type=int, default=12)
    ap.add_argument("--stiff-plain-only", action="store_true", help="stiff case: the plain CR leg only")
    ap.add_argument("--stiff-trace-summary", nargs=2, metavar=("STATS_CSV", "STIFF_JSON"))
    ap.add_argument("--stiff-summary", metavar="STIFF_JSON",
                    help="write the plain vs learned table of a --stiff run's JSON line to profiles/")
    ap.add_argument("--lpc-weights", default=None, help="network parameters (.npy) for the learned legs")
    ap.add_argument("--lpc-fallback", type=float, default=None, help="fallback ratio of the learned legs")
    ap.add_argument("--lpc-dirs", type=int, default=None, help="directions the learned legs' flexible CR keeps")
    ap.add_argument("--lpc-strikes", type=int, default=None,
                    help="fallbacks after which a body of the learned legs keeps the diagonal (default 1)")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--trace-summary", nargs=2, metavar=("STATS_CSV", "BENCH_LOG"))
    ap.add_argument("--sq-summary", nargs=3, metavar=("PASS0_CSV", "PASS1_CSV", "PASS2_CSV"))
    ap.add_argument("--mfma-summary", nargs=2, metavar=("PMC_CSV", "TRACE_CSV"))
    args = ap.parse_args()
    wl = args.workload
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], profile_file("traffic", wl), wl)
        return
    if args.stiff_trace_summary:
        stiff_trace_summary(args.stiff_trace_summary[0], args.stiff_trace_summary[1], STIFF_STATS_FILE)
        return
    if args.stiff_summary:
        stiff_summary(args.stiff_summary, STIFF_COMPARE_FILE)
        return
    if args.sq_summary:
        sq_summary(args.sq_summary, profile_file("sq", wl), workload=wl)
        return
    if args.mfma_summary:
        mfma_summary(args.mfma_summary[0], args.mfma_summary[1], MFMA_FILE)
        return
    if args.trace_summary:
        trace_summary(args.trace_summary[0], args.trace_summary[1], profile_file("stats", wl), wl)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the step on a single MI355X "
              "(multi-GPU sharding is a design note, not implemented); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.stiff:
        print(json.dumps(stiff_leg(args.stiff_particles, args.stiff_steps, args.stiff_warmup,
                                   weights=args.lpc_weights, fallback=args.lpc_fallback, dirs=args.lpc_dirs,
                                   plain_only=args.stiff_plain_only)))
        return

    import torch

    if args.copy_calibration:
        a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        float(b.sum())  # a read-only pass over the copy's destination
        del a, b

    base = {"metric": "particle-steps/s & CR iters/step at 1e-5 residual, learned-PC vs plain CR",
            "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "data": "synthetic (seeded SplitMix64 particles)"}
    if wl == "snowballs":
        n = args.particles or 1000000
        sb = snowballs_leg(args.steps, args.warmup, n)
        out = dict(base, value=sb["value"], ms_per_step=sb["ms_per_step"], dtype="f32 (FP64 CR reductions)",
                   config={"workload": sb["config"], "particles": n, "preconditioner": "diagonal (plain CR)",
                           "parallelism": "single GPU"})
        h = sb["hvp"]
        out["roofline"] = {"bound": "hbm", "kernel": "hvp", "achieved": round(h["achieved_gbs"], 1),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(h["achieved_gbs"] / HBM_PEAK_GBS, 4),
                           "traffic": read_traffic(wl), "bytes_per_launch": h["bytes_per_launch"],
                           "ms_per_launch": h["ms_per_launch"], "launches": h["launches_timed"]}
        out["two_snowballs_1M"] = sb
        print(json.dumps(out))
        return

    n = args.particles or 4000000
    legs, sc = slab_legs(args.steps, args.warmup, n, args.lpc_weights, args.lpc_fallback, args.lpc_dirs,
                         args.lpc_strikes)
    lr, pl = legs["learned"], legs["plain"]
    h = lr["hvp"]
    out = dict(base, value=lr["value"], ms_per_step=lr["ms_per_step"],
               dtype="f32 (FP64 CR reductions; BF16 MFMA network, FP32 accumulation)",
               config={"workload": "slope slab: %d particles, 0.8 x 0.2 x 0.4 m, per-particle E0 in [7e4, 2.8e5] Pa "
                                   "and xi in [5, 10] (seed 2, held out from training), 30-degree slope, h=1/256, "
                                   "dt=5e-4, CR tol 1e-5; learned flexible CR and plain diagonal CR, %d steps each "
                                   "from the state after %d plain steps" % (n, args.steps, args.warmup),
                       "particles": n, "preconditioner": "learned (flexible CR), against plain diagonal CR",
                       "parallelism": "single GPU"})
    out["value_is"] = "particle-steps/s of the learned-preconditioned step"
    out["cr_iters_per_step"] = lr["cr_iters_per_step"]
    out["learned"] = lr
    out["plain"] = pl
    out["speedup"] = lr["value"] / pl["value"]
    out["iteration_ratio"] = pl["cr_iters_per_step"] / max(lr["cr_iters_per_step"], 1e-9)
    out["roofline"] = {"bound": "hbm", "kernel": "hvp", "achieved": round(h["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(h["achieved_gbs"] / HBM_PEAK_GBS, 4),
                       "traffic": read_traffic(wl), "bytes_per_launch": h["bytes_per_launch"],
                       "ms_per_launch": h["ms_per_launch"], "launches": h["launches_timed"]}
    out["hvp_issue"] = read_sq(wl)
    mf = read_footer(MFMA_FILE)
    if mf and "network" in lr:
        lr["network"]["mfma_counters"] = {
            "profile": os.path.relpath(MFMA_FILE, ROOT), "tflops": round(mf["tflops"], 1),
            "frac_of_bf16_peak": round(mf["frac_of_bf16_peak"], 4),
            "conv_gflop_per_forward": round(mf["conv_gflop_per_forward"], 2)}
    ps = read_footer(PRECISION_FILE)
    if ps:
        out["precision_study"] = {"profile": os.path.relpath(PRECISION_FILE, ROOT), "particles": ps["particles"],
                                  "steps": ps["steps"], "max_over_steps": ps["max_over_steps"],
                                  "note": "FP32-bulk GPU step vs the FP64 oracle (and the FP32 oracle) over the "
                                          "same steps at the slab's density, h and dt"}
    p64 = read_footer(PRECISION_GPU64_FILE)
    if p64:
        out["precision_study_4M"] = {"profile": os.path.relpath(PRECISION_GPU64_FILE, ROOT),
                                     "particles": p64["particles"], "steps": p64["steps"],
                                     "max_over_steps": p64["max_over_steps"]["gpu"], "its_mean": p64["its_mean"],
                                     "ms_per_step_median": p64["ms_per_step_median"],
                                     "note": "FP32-bulk GPU step vs the full-FP64 GPU step (mpm_sim64) on the "
                                             "workload itself, 300 steps"}
    if not args.no_cpu:
        out["cpu_baseline"], out["error_vs_fp64"] = cpu_leg(args.warmup)
        out["cpu_baseline"]["gpu_over_cpu"] = pl["value"] / out["cpu_baseline"]["value"]
    if not args.no_snowballs:
        out["two_snowballs_1M"] = snowballs_leg(args.snowball_steps, 3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
