"""Benchmark of the MI355X sfa/sim step: the 512^3 gauged leapfrog step.

Workload (one GPU): gauged mode, 512^3 periodic, dx=0.5 (L=127.75), dt=0.1,
64 Gaussian Q-balls at seeded uniform positions (amp U[0.6,0.9], width
U[2,4] cells, omega U[0.6,0.95], random phase, alternating charge signs),
theta noise 0.01, links at identity, E = 0 — synthetic data generated on the
host from a fixed seed and resident in HBM before the timed region.

A "step" is one full leapfrog step of the whole grid (one fused launch).
value = site-updates/s over K timed steps (HIP events on the grid's stream)
of the headline precision: FP32C with int8 compensation words on the
positions by default (lo_words 4);
the FP64 path is timed on the same state after it ("fp64", "speedup_vs_fp64").
roofline: algorithmic HBM bytes per site-update = every persistent state
scalar read and written once at its storage width, compensation words not
counted (gauged: 30 scalars; FP64 8 B -> 480 B; FP32C 4 B -> 240 B), times N^3
sites per launch, over the average launch time; peak = 8 TB/s. The bytes the
step actually streams per site (FP32C: 4 B hi words plus the compensation
words, for lo_words 0/1/2/3/4 -> 480/360/240/304/272 B) are reported beside
it as stored_bytes_per_site and stored_frac. traffic: measured HBM bytes per launch
(rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by a plain copy's counters), read
from the counter summary profiles/hbm_traffic.txt when it holds a line for
this kernel and size. cpu_baseline: the reference CPU binary built from source
(oracle/_ref/scp_sim) on every core of this host, on bounded samples of the
CPU-runnable shapes (gauged 128^3 two balls, complex 64^3 one ball).

--mode real runs the same measurement on the real-mode step (6 fields: 12
state scalars, 192 B FP64 / 96 B FP32C per site-update; the state is the real
part of the same 64 seeded balls, theta noise 0.01) with the reference
binary's real-mode 64^3 run as the CPU baseline. It is not the headline line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--prec f64|f32c]
                       [--lo-words 0|1|2|3|4] [--mode gauged|real] [--no-compare] [--no-cpu-baseline]
       python bench.py --depths 128,64,32,16   (march depths on the headline grid, one JSON line)
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

LO_NAMES = {0: "f32", 1: "bf16", 2: "no", 3: "int8", 4: "int8 position"}


NFIELDS = {"gauged": 15, "real": 6}   # fields per site (positions; as many velocities)


def bytes_per_site(prec, mode="gauged"):
    """Algorithmic: the state scalars (gauged 30, real 12), each read once and
    written once per step at the state's width (compensation words are not state)."""
    return 2 * 2 * NFIELDS[mode] * (8 if prec == "f64" else 4)


def stored_bytes_per_site(prec, lo_words, mode="gauged"):
    """What the step streams: the state words plus the compensation words."""
    nf = NFIELDS[mode]
    if prec == "f64":
        return 2 * 2 * nf * 8
    # nf position and nf velocity scalars; int8 words packed four fields to a 4-byte word
    q8 = 4 * ((nf + 3) // 4) / nf
    lo_p, lo_w = {0: (4, 4), 1: (2, 2), 2: (0, 0), 3: (q8, q8), 4: (q8, 0)}[lo_words]
    return 2 * nf * (8 + lo_p + lo_w)


def measured_traffic(kernel, N):
    """HBM bytes per launch for `kernel` at N^3 from profiles/hbm_traffic.txt
    (lines 'traffic <kernel without spaces> N=<N> <bytes>'), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) == 4 and p[0] == "traffic" and p[1] == kernel.replace(" ", "") and p[2] == f"N={N}":
                return float(p[3])
    return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """(threads this process may run on, physical cores of the host, CPU model)."""
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model, phys = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            cur = {}
            for line in f.read().splitlines() + [""]:
                if not line.strip():
                    if "physical id" in cur and "core id" in cur:
                        phys.add((cur["physical id"], cur["core id"]))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
                if k.strip() == "model name":
                    model = v.strip()
    except OSError:
        pass
    return threads, len(phys) or threads, model


def _ref_run(ref, wd, name, cfg, threads):
    """One reference run; returns (ms/step from its 'Wall' line, process wall s)."""
    path = os.path.join(wd, name + ".cfg")
    cfg = dict(cfg, output=os.path.join(wd, name + ".sfa"), diag_file=os.path.join(wd, name + ".tsv"))
    with open(path, "w") as f:
        f.write("".join(f"{k} = {v}\n" for k, v in cfg.items()))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    t0 = time.time()
    res = subprocess.run([ref, path], capture_output=True, text=True, env=env, cwd=wd, timeout=900)
    wall = time.time() - t0
    m = re.search(r"Wall: [\d.]+s \([\d.]+ min\) ([\d.]+)ms/step", res.stdout)
    if res.returncode != 0 or not m:
        log(f"cpu baseline {name} failed rc={res.returncode}: {res.stderr[-300:]}")
        return None, wall
    return float(m.group(1)), wall


def cpu_baseline(mode="gauged", gauged_steps=300):
    """Time the reference CPU path (oracle/_ref/scp_sim, built from the reference
    sources with its Makefile flags -O3 -fopenmp) on every core this process may
    use, on the CPU-runnable measurement shapes: the gauged 128^3 two-ball run
    (periodic, g = 0.05, links N(0, 1e-3), each run's own Gauss-projecting init;
    gauged_steps of its 2000 steps) and the complex 64^3 one-ball run (all 500
    steps), with snap_dt >= T and diag_dt = T. value: the gauged rate (the
    headline's mode); --mode real times the real-mode oscillon 64^3 instead."""
    from scp_amd import sfa
    from scp_amd import seeds
    ref = os.path.join(ROOT, "oracle", "_ref", "scp_sim")
    if not os.path.exists(ref):
        return None
    threads, phys, model = host_cpu()
    base = {"dt_factor": 0.2, "precision": "f64", "bc_type": 2}
    out = {"unit": "Gsite-updates/s", "cores": threads, "kind": "reference"}
    host = f"{model}, {phys} physical cores, OMP_NUM_THREADS={threads}"
    with tempfile.TemporaryDirectory() as wd:
        if mode == "real":
            N, steps = 64, 1500
            L = 0.25 * (N - 1)
            dt = 0.2 * 2 * L / (N - 1)
            ms, wall = _ref_run(ref, wd, "real64", dict(base, N=N, L=L, T=steps * dt, init="oscillon",
                                                        snap_dt=2 * steps * dt, diag_dt=steps * dt), threads)
            if ms is None:
                return None
            out.update(value=N ** 3 / (ms * 1e-3) / 1e9,
                       sample=f"reference scp_sim (gcc -O3 -fopenmp), real 64^3 oscillon, {steps} steps, "
                              f"{host}, 'Wall' {ms:.2f} ms/step, process wall {wall:.1f}s")
            return out
        runs = {}
        for name, N, steps, nb in (("gauged128", 128, gauged_steps, 2), ("complex64", 64, 500, 1)):
            L = 0.25 * (N - 1)
            dt = 0.2 * 2 * L / (N - 1)
            if nb == 2:
                balls = [dict(x0=-6.0, y0=0.0, z0=0.0, amp=0.8, sigma=3.0, omega=0.8, phase=0.0, winding=1),
                         dict(x0=6.0, y0=0.0, z0=0.0, amp=0.8, sigma=3.0, omega=-0.8, phase=0.0, winding=-1)]
                st = seeds.seed_balls(N, L, 2, balls, theta_sigma=0.01, link_sigma=1e-3, seed=1001)
                extra = {"complex_phi": 1, "complex_gauge": 1, "g_gauge": 0.05}
            else:
                balls = [dict(x0=0.0, y0=0.0, z0=0.0, amp=0.8, sigma=3.0, omega=0.8, phase=0.0, winding=1)]
                st = seeds.seed_balls(N, L, 1, balls, theta_sigma=0.01, seed=1000)
                extra = {"complex_phi": 1}
            seed = os.path.join(wd, name + "_seed.sfa")
            sfa.write_sim_sfa(seed, N, L, dt, [(0.0, st)])
            del st
            ms, wall = _ref_run(ref, wd, name, dict(base, N=N, L=L, T=steps * dt, init="sfa", init_sfa=seed,
                                                    snap_dt=2 * steps * dt, diag_dt=steps * dt, **extra), threads)
            if ms is None:
                return None
            runs[name] = (N ** 3 / (ms * 1e-3) / 1e9, ms, steps, wall)
    g, c = runs["gauged128"], runs["complex64"]
    out.update(value=g[0],
               sample=f"reference scp_sim (gcc -O3 -fopenmp) on {host}: gauged 128^3 two opposite balls, "
                      f"periodic, g=0.05, {g[2]} of 2000 steps, 'Wall' {g[1]:.2f} ms/step (process {g[3]:.1f}s "
                      f"incl. its CG init); complex 64^3 one ball, {c[2]} steps, 'Wall' {c[1]:.2f} ms/step "
                      f"= {c[0]:.4f} Gsite-updates/s (process {c[3]:.1f}s)",
               complex_64=c[0])
    return out


def copy_calibration(gib, reps=3):
    """A plain device-to-device copy of known size: rocprof's FETCH_SIZE/WRITE_SIZE
    for it give the scale factor between the counters and bytes moved."""
    import torch
    n = int(gib * (1 << 30)) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.time()
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.time() - t0) / reps
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "reps": reps,
                      "GBps": 2 * n * 4 / dt / 1e9}), flush=True)


def depth_sweep(N, depths, lo_words, reps=3, steps=30):
    """The headline state as a one-grid batch (the batched march runs the same
    tile code as the single-grid launch) timed at each march depth in turn,
    reps rounds of alternating depths, FP32C then FP64: one JSON line of mean
    ms/step per depth."""
    from scp_amd import grid as G
    L = 0.25 * (N - 1)
    st = build_state(N, L)
    out = {"depths": depths, "N": N, "rounds": reps, "steps_per_timing": steps}
    for prec in ("f32c", "f64"):
        params = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05, lo_words=lo_words,
                               precision=G.PREC_F64 if prec == "f64" else G.PREC_F32C)
        with G.Grid(params) as g, G.Batch([g]) as b:
            g.upload(st)
            res = {d: [] for d in depths}
            for r in range(reps):
                for d in depths:
                    b.set_config(di=d)
                    b.step(2)
                    res[d].append(b.time_steps(steps) / steps)
                    log(f"{prec} round {r} depth {d}: {res[d][-1]:.4f} ms/step")
            out[prec] = {str(d): sum(v) / len(v) for d, v in res.items()}
    print(json.dumps(out), flush=True)


REAL_COLS = [0, 1, 2, 3, 4, 5, 12, 13, 14, 15, 16, 17]   # complex columns: Re Phi, Re Theta and their velocities


def build_state(N, L, mode="gauged", seed=1002):
    from scp_amd import seeds
    balls = seeds.draw_balls(N, L, 64, seed)
    if mode == "gauged":
        return seeds.seed_balls(N, L, 2, balls, theta_sigma=0.01, link_sigma=0.0, seed=seed)
    return np.ascontiguousarray(seeds.seed_balls(N, L, 1, balls, theta_sigma=0.01, seed=seed)[REAL_COLS])


def run_one(G, N, L, st, prec, lo_words, steps, warmup, mode="gauged"):
    """Upload the state, warm up, time `steps` launches with HIP events on the
    grid's stream; returns the timing and the roofline figures."""
    params = G.make_params(N, G.MODE_GAUGED if mode == "gauged" else G.MODE_REAL, L, 0.2, g=0.05,
                           lo_words=lo_words, precision=G.PREC_F64 if prec == "f64" else G.PREC_F32C)
    log(f"{prec} lo_words={lo_words}: device bytes {G.grid_bytes(params) / 1e9:.1f} GB")
    with G.Grid(params) as g:
        g.upload(st)
        if warmup > 0:
            g.step(warmup)
        w0 = time.time()
        ms = g.time_steps(steps)
        wall = time.time() - w0
        d = g.diag()
        kernel = g.step_kernel_name
        diag_ms = diag_step_timing(g)
    ms_step = ms / steps
    sites = float(N) ** 3
    log(f"{prec}: {steps} steps {ms:.2f} ms (wall {wall * 1e3:.1f} ms) -> {ms_step:.3f} ms/step; "
        f"E_total={d['E_total']:.9e} Q_total={d['Q_total']:.6e}")
    bps, sbps = bytes_per_site(prec, mode), stored_bytes_per_site(prec, lo_words, mode)
    achieved = bps * sites / (ms_step * 1e-3) / 1e9
    stored = sbps * sites / (ms_step * 1e-3) / 1e9
    log(f"{prec}: diag step (wall, ms): step alone {diag_ms['step']:.2f}, step + stand-alone diagnostics "
        f"{diag_ms['separate']:.2f}, diagnostics fused into the step {diag_ms['fused']:.2f}")
    return {"ms_per_step": ms_step, "value": sites / (ms_step * 1e-3) / 1e9, "kernel": kernel,
            "bytes_per_site": bps, "achieved": achieved, "traffic": measured_traffic(kernel, N),
            "stored_bytes_per_site": sbps, "stored_GBps": stored,
            "diag_step_ms": diag_ms}


def diag_step_timing(g, reps=3):
    """Wall time (synchronised, median of reps) of one step, of one step followed
    by the stand-alone diagnostics pass, and of one step with the diagnostics
    fused into its launch (scp_grid_step_diag): the cost of a diag.tsv row."""
    def med(fn):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))
    return {"step": med(lambda: g.step(1)), "separate": med(lambda: (g.step(1), g.diag())),
            "fused": med(lambda: g.step_diag(1))}


def drift_bench(a):
    """FP32C against FP64 over many steps on the headline state: gauged 512^3,
    64 Q-balls (alternating charge signs), Gauss-projected at t = 0, both grids
    resident and stepped in chunks with the diagnostics fused into the last
    step of each chunk. Reports each path's energy drift and the FP32C-FP64
    differences of E_total and Q_total; the charge difference is relative to
    the sum of the balls' |Q| (sum over sites of |Phi's charge density| x dV of
    the initial state), since the net charge of alternating balls is near 0."""
    from scp_amd import grid as G
    N = a.n
    L = 0.25 * (N - 1)
    dx = 2 * L / (N - 1)
    log(f"seeding {N}^3 gauged state on the host")
    st = build_state(N, L)
    qabs = dx ** 3 * float(np.abs(sum(st[c] * st[18 + c] - st[12 + c] * st[6 + c] for c in range(3))).sum())
    p64 = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05)
    p32 = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05, precision=G.PREC_F32C, lo_words=a.lo_words)
    rows = []
    with G.Grid(p64) as g64, G.Grid(p32) as g32:
        for g in (g64, g32):
            g.upload(st)
            g.gauss_project()
        del st
        d64, d32 = g64.diag(), g32.diag()
        e0, q0 = d64["E_total"], d64["Q_total"]
        log(f"t=0: E {e0:.12e} Q {q0:.6e} sum|Q| {qabs:.6e} gauss_max {d64['gauss_max']:.3e} / {d32['gauss_max']:.3e}")
        done, t0 = 0, time.time()
        while done < a.drift_steps:
            n = min(a.chunk, a.drift_steps - done)
            d64, d32 = g64.step_diag(n), g32.step_diag(n)
            done += n
            row = {"step": done, "dE_f64": (d64["E_total"] - e0) / abs(e0), "dE_f32c": (d32["E_total"] - e0) / abs(e0),
                   "dE_f32c_vs_f64": (d32["E_total"] - d64["E_total"]) / abs(d64["E_total"]),
                   "dQ_f32c_vs_f64": (d32["Q_total"] - d64["Q_total"]) / qabs,
                   "gauss_max_f64": d64["gauss_max"], "gauss_max_f32c": d32["gauss_max"]}
            rows.append(row)
            log(f"step {done}: " + " ".join(f"{k} {v:.3e}" for k, v in row.items() if k != "step")
                + f" ({time.time() - t0:.0f}s)")
    last = rows[-1]
    out = {"metric": "FP32C vs FP64 energy / charge difference over a long run, 512^3 gauged",
           "value": max(abs(last["dE_f32c_vs_f64"]), abs(last["dQ_f32c_vs_f64"])), "unit": "relative",
           "n_gpus": 1, "steps": a.drift_steps, "higher_is_better": False,
           "dtype": f"f32c ({LO_NAMES[a.lo_words]} compensation words) vs f64",
           "data": "synthetic (seeded Gaussian Q-balls, generated on the host, Gauss-projected)",
           "config": {"workload": f"gauged {N}^3 periodic, 64 Q-balls, dx=0.5 dt=0.1, {a.drift_steps} steps",
                      "N": N, "chunk": a.chunk, "lo_words": a.lo_words, "sum_abs_Q": qabs, "E0": e0, "Q0": q0},
           "max_abs": {k: max(abs(r[k]) for r in rows) for k in rows[0] if k != "step"},
           "rows": rows}
    print(json.dumps(out), flush=True)


def sweep_bench(a):
    """Batched parameter sweep: independent gauged grids of edge drawn from
    {32,48,64,80,96} (seeded), one Q-ball each (amp U[0.6,0.9], width U[2,4]
    cells, omega U[0.6,0.95]), FP32C, advanced together by one batch (one
    launch per step over all tiles, chains replayed from HIP graphs, depth and
    launch form autotuned). Compares with stepping each grid on its own."""
    from scp_amd import grid as G
    from scp_amd import seeds
    rng = np.random.default_rng(a.seed)
    if a.sizes == "serve":       # the serving mix: edge log-uniform in [48, 256]
        sizes = np.round(np.exp(rng.uniform(np.log(48), np.log(256), size=a.grids))).astype(int)
    else:
        sizes = rng.choice([32, 48, 64, 80, 96], size=a.grids)
    grids, sites = [], 0
    t0 = time.time()
    for i, N in enumerate(sizes):
        N = int(N)
        L = 0.25 * (N - 1)
        dx = 0.5
        b = dict(x0=float(rng.uniform(-0.2, 0.2) * L), y0=float(rng.uniform(-0.2, 0.2) * L),
                 z0=float(rng.uniform(-0.2, 0.2) * L), amp=float(rng.uniform(0.6, 0.9)),
                 sigma=float(rng.uniform(2.0, 4.0)) * dx, omega=float(rng.uniform(0.6, 0.95)),
                 phase=float(rng.uniform(0, 2 * np.pi)), winding=1)
        st = seeds.seed_balls(N, L, 2, [b], theta_sigma=0.01, link_sigma=0.0, seed=a.seed + i)
        g = G.Grid(G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05, precision=G.PREC_F32C, lo_words=a.lo_words))
        g.upload(st)
        grids.append(g)
        sites += N ** 3
    log(f"{len(grids)} grids, {sites / 1e6:.1f} M sites, set up in {time.time() - t0:.1f}s")
    with G.Batch(grids) as b:
        b.step(2)
        ms_default = b.time_steps(a.steps) / a.steps
        log(f"default config {b.config}: {ms_default:.3f} ms/step")
        taken = b.tune()
        cfg = b.config
        log(f"autotune: {cfg} ({taken} steps)")
        chunk = min(a.steps, 1000)           # chains of at most 1000 launches (one captured graph)
        b.step(chunk)                        # captures the chain graph
        ms, left = 0.0, a.steps
        while left > 0:                      # replays
            n = min(chunk, left)
            ms += b.time_steps(n)
            left -= n
            if a.steps > chunk:
                log(f"sweep: {a.steps - left} of {a.steps} steps timed")
    ms_step = ms / a.steps
    # the same grids stepped one at a time (each its own launch per step)
    t1 = time.time()
    ms_alone = 0.0
    for g in grids:
        ms_alone += g.time_steps(a.steps)
    ms_alone_step = ms_alone / a.steps
    for g in grids:
        g.close()
    out = {"metric": "aggregate Gsite-updates/s, batched sweep of independent gauged grids",
           "value": sites / (ms_step * 1e-3) / 1e9, "unit": "Gsite-updates/s", "n_gpus": 1,
           "steps": a.steps, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": f"f32c ({LO_NAMES[a.lo_words]} compensation words)",
           "data": "synthetic (seeded single Q-ball per grid)",
           "config": {"workload": f"{len(grids)} gauged grids, edges " +
                                  ("from {32,48,64,80,96}" if a.sizes == "sweep" else "log-uniform in [48,256]") +
                                  ", one Q-ball each",
                      "grids": len(grids), "sites": sites, "seed": a.seed, "tuned": cfg},
           "roofline": {"bound": "hbm", "achieved": bytes_per_site("f32c") * sites / (ms_step * 1e-3) / 1e9,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": bytes_per_site("f32c") * sites / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "stored_frac": stored_bytes_per_site("f32c", a.lo_words) * sites / (ms_step * 1e-3) / 1e9
                                       / HBM_PEAK_GBS},
           "unbatched": {"value": sites / (ms_alone_step * 1e-3) / 1e9, "ms_per_step": ms_alone_step},
           "untuned_ms_per_step": ms_default,
           "batch_speedup": ms_alone_step / ms_step}
    print(json.dumps(out), flush=True)


def serve_bench(a, emit=True):
    """Serving: Poisson job arrivals (rate a.rate per second for a.duration
    seconds); each job a gauged periodic grid of edge log-uniform in
    [48, 256], steps uniform in [steps_lo, steps_hi], 1-4 seeded Q-balls;
    admitted by an HBM budget, device-resident, continuously batched. Reports
    p50/p99 completion latency (arrival to completion) and aggregate
    site-updates/s; jobs still running when the horizon ends are counted."""
    from scp_amd import grid as G
    rng = np.random.default_rng(a.seed)
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / a.rate)
        if t > a.duration:
            break
        arrivals.append(t)
    specs = []
    for _ in arrivals:
        N = int(round(np.exp(rng.uniform(np.log(48), np.log(256)))))
        L = 0.25 * (N - 1)
        dx = 0.5
        nb = int(rng.integers(1, 5))
        balls = []
        for b in range(nb):
            sg = 1 if b % 2 == 0 else -1
            balls.append(dict(x0=float(rng.uniform(-0.3, 0.3) * L), y0=float(rng.uniform(-0.3, 0.3) * L),
                              z0=float(rng.uniform(-0.3, 0.3) * L), amp=float(rng.uniform(0.6, 0.9)),
                              sigma=float(rng.uniform(2.0, 4.0)) * dx, omega=sg * float(rng.uniform(0.6, 0.95)),
                              phase=float(rng.uniform(0, 2 * np.pi)), winding=sg))
        specs.append(dict(N=N, L=L, balls=balls, steps=int(rng.integers(a.steps_lo, a.steps_hi + 1))))
    offered = sum(s_["N"] ** 3 * s_["steps"] for s_ in specs)
    log(f"{len(specs)} jobs over {a.duration:.0f}s, offered {offered / 1e9:.1f} G site-updates")
    lat, wait, run, done_sites = [], [], [], 0
    with G.Server(int(a.budget_gb * 1e9), quantum=a.quantum, max_active=a.max_active) as srv:
        pending, i, next_log = {}, 0, 10.0
        while True:
            now = srv.now()
            if now >= next_log:
                st = srv.stats()
                log(f"t={now:6.1f}s submitted {i} queued {st['queued']} active {st['active']} "
                    f"completed {len(lat)} in use {st['bytes_in_use'] / 1e9:.1f} GB")
                next_log = now + 10.0
            while i < len(specs) and arrivals[i] <= now:
                sp = specs[i]
                p = G.make_params(sp["N"], G.MODE_GAUGED, sp["L"], 0.2, g=0.05, precision=G.PREC_F32C,
                                  lo_words=a.lo_words)
                pending[srv.submit(p, sp["balls"], sp["steps"], link_sigma=0.0, seed=a.seed + i)] = i
                i += 1
            if now > a.duration + a.horizon:
                break
            st = srv.stats()
            if st["active"] == 0 and st["queued"] == 0 and st["done"] == 0:
                if i == len(specs):
                    break
                time.sleep(max(0.0, min(arrivals[i] - now, 1.0)))
                continue
            srv.round()
            for jid in list(pending):
                info = srv.job(jid)
                if info["state"] == G.JOB_DONE:
                    k = pending.pop(jid)
                    srv.take(jid, state=False)
                    lat.append(info["t_done"] - arrivals[k])
                    wait.append(info["t_admit"] - arrivals[k])
                    run.append(info["t_done"] - info["t_admit"])
                    done_sites += specs[k]["N"] ** 3 * specs[k]["steps"]
        st = srv.stats()
    elapsed = st["now"]
    completed = len(lat)
    edges = [0, 1, 2, 5, 10, 20, 50, 100, 200, 500, 1000, 2000]
    hist = np.histogram(np.clip(lat, 0, edges[-1]), bins=edges)[0].tolist() if lat else [0] * (len(edges) - 1)
    pct = lambda v, q: float(np.nanpercentile(np.array(v), q)) if v else float("nan")
    lat = np.array(lat) if lat else np.array([float("nan")])
    out = {"metric": "serving: p50/p99 job completion latency and aggregate Gsite-updates/s",
           "value": st["site_updates"] / elapsed / 1e9, "unit": "Gsite-updates/s", "n_gpus": 1,
           "higher_is_better": True, "vs_baseline": None,
           "dtype": f"f32c ({LO_NAMES[a.lo_words]} compensation words)",
           "data": "synthetic (Poisson arrivals, seeded Q-ball jobs)",
           "config": {"workload": "gauged jobs, edge log-uniform [48,256], 1-4 Q-balls", "rate_per_s": a.rate,
                      "duration_s": a.duration, "horizon_s": a.horizon, "steps": [a.steps_lo, a.steps_hi],
                      "hbm_budget_gb": a.budget_gb, "quantum": a.quantum, "max_active": a.max_active,
                      "seed": a.seed},
           "jobs": {"submitted": i, "completed": completed, "unfinished": i - completed},
           "latency_s": {"p50": float(np.nanpercentile(lat, 50)), "p99": float(np.nanpercentile(lat, 99)),
                         "max": float(np.nanmax(lat))},
           # completed jobs: arrival -> admission (queueing) and admission -> completion (on the GPU)
           "wait_s": {"p50": pct(wait, 50), "p99": pct(wait, 99)},
           "run_s": {"p50": pct(run, 50), "p99": pct(run, 99)},
           "latency_histogram": {"edges_s": edges, "counts": hist},
           "offered_gsite_updates_per_s": offered / a.duration / 1e9,
           "busy_fraction": st["busy_seconds"] / elapsed, "elapsed_s": elapsed}
    if emit:
        print(json.dumps(out), flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--prec", default="f32c", choices=["f64", "f32c"],
                    help="precision of the headline value (default: FP32C, the production form)")
    ap.add_argument("--lo-words", type=int, default=4,
                    help="FP32C compensation words: 0 FP32, 1 bfloat16, 2 none (plain FP32), "
                         "3 signed byte in units of ulp/256, 4 such bytes on the positions only (default)")
    ap.add_argument("--mode", default="gauged", choices=["gauged", "real"],
                    help="step workload: gauged (the headline) or real (6-field real-mode step)")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 comparison run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="step workload: skip the short serving window that adds p50/p99 job latency")
    ap.add_argument("--drift-steps", type=int, default=10000, help="drift: steps of the long run")
    ap.add_argument("--chunk", type=int, default=1000, help="drift: steps between diagnostics")
    ap.add_argument("--workload", default="step", choices=["step", "sweep", "serve", "drift"],
                    help="step: the 512^3 gauged step (default); sweep: batched independent grids; "
                         "serve: Poisson job arrivals through the request loop; drift: FP32C against "
                         "FP64 over --drift-steps steps")
    ap.add_argument("--rate", type=float, default=4.0, help="serve: job arrivals per second")
    ap.add_argument("--duration", type=float, default=600.0, help="serve: arrival window (s)")
    ap.add_argument("--horizon", type=float, default=60.0, help="serve: time after the window to finish jobs (s)")
    ap.add_argument("--steps-lo", type=int, default=1000)
    ap.add_argument("--steps-hi", type=int, default=20000)
    ap.add_argument("--budget-gb", type=float, default=250.0, help="serve: HBM budget")
    ap.add_argument("--quantum", type=int, default=200, help="serve: steps per scheduling round")
    ap.add_argument("--max-active", type=int, default=0, help="serve: concurrent jobs limit (0: memory only)")
    ap.add_argument("--grids", type=int, default=256, help="sweep: number of grids")
    ap.add_argument("--sizes", default="sweep", choices=["sweep", "serve"],
                    help="sweep: edges from {32,48,64,80,96}; serve: edges log-uniform in [48,256]")
    ap.add_argument("--seed", type=int, default=3, help="sweep: seed of sizes and balls")
    ap.add_argument("--depths", default="",
                    help="instead of the step: time the headline grid as a one-grid batch at these march "
                         "depths (planes per block, comma-separated), alternating")
    ap.add_argument("--copy-calib", type=float, default=0.0,
                    help="instead of the step: time a plain device copy of this many GiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.copy_calib > 0:
        copy_calibration(a.copy_calib)
        return
    if a.depths:
        depth_sweep(a.n, [int(d) for d in a.depths.split(",")], a.lo_words)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the multi-GPU form (slabs along "
                                   "the march axis with one-plane halo exchange over RCCL) is not built",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "sweep":
        sweep_bench(a)
        return
    if a.workload == "drift":
        drift_bench(a)
        return
    if a.workload == "serve":
        serve_bench(a)
        return
    from scp_amd import grid as G

    N = a.n
    L = 0.25 * (N - 1)            # dx = 0.5
    log(f"seeding {N}^3 {a.mode} state on the host")
    t0 = time.time()
    st = build_state(N, L, a.mode)
    log(f"seeded in {time.time() - t0:.1f}s")
    head = run_one(G, N, L, st, a.prec, a.lo_words, a.steps, a.warmup, a.mode)
    cmp64 = None
    if a.prec == "f32c" and not a.no_compare:
        cmp64 = run_one(G, N, L, st, "f64", 0, a.steps, a.warmup, a.mode)
    del st
    serving = None
    if a.mode == "gauged" and not a.no_serve:
        # the metric's latency half: a short window of the serving workload (Poisson
        # arrivals at 4 jobs/s, the job mix of the full serving run, at most 16 jobs on
        # the GPU so that jobs complete within the window); the full 600 s run is
        # `--workload serve`
        import copy
        sa = copy.copy(a)
        sa.rate, sa.duration, sa.horizon, sa.max_active = 4.0, 15.0, 45.0, 16
        log("serving window: 15 s of arrivals at 4 jobs/s, 45 s horizon, at most 16 active")
        sv = serve_bench(sa, emit=False)
        serving = {"p50_latency_s": sv["latency_s"]["p50"], "p99_latency_s": sv["latency_s"]["p99"],
                   "aggregate_gsite_updates_per_s": sv["value"], "jobs": sv["jobs"],
                   "wait_p50_s": sv["wait_s"]["p50"], "run_p50_s": sv["run_s"]["p50"],
                   "window": "Poisson 4 jobs/s for 15 s + 45 s horizon, at most 16 active, HBM budget 250 GB; "
                             "gauged FP32C jobs, edge log-uniform [48,256], steps U[1000,20000], 1-4 Q-balls"}
    cpu = None if a.no_cpu_baseline else cpu_baseline(a.mode)
    dtype = "f64" if a.prec == "f64" else f"f32c ({LO_NAMES[a.lo_words]} compensation words)"
    out = {
        "metric": ("Gsite-updates/s & %HBM peak, 512^3 gauged, FP32c vs FP64; p50/p99 job latency"
                   if a.mode == "gauged" else f"Gsite-updates/s & %HBM peak, {N}^3 real-mode step"),
        "value": head["value"], "unit": "Gsite-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype, "data": "synthetic (seeded Gaussian Q-balls, generated on the host)",
        "config": {"workload": f"{a.mode} {N}^3 leapfrog step, 64 Q-balls, dx=0.5 dt=0.1",
                   "N": N, "mode": a.mode, "precision": a.prec, "lo_words": a.lo_words,
                   "g": 0.05 if a.mode == "gauged" else None},
        "roofline": {"bound": "hbm", "achieved": head["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": head["achieved"] / HBM_PEAK_GBS, "traffic": head["traffic"],
                     "kernel": head["kernel"], "bytes_per_site": head["bytes_per_site"],
                     "stored_bytes_per_site": head["stored_bytes_per_site"],
                     "stored_frac": head["stored_GBps"] / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "diag_step_ms": head["diag_step_ms"],
        "serving": serving,
    }
    if cmp64:
        out["fp64"] = {"value": cmp64["value"], "ms_per_step": cmp64["ms_per_step"],
                       "roofline_frac": cmp64["achieved"] / HBM_PEAK_GBS, "kernel": cmp64["kernel"],
                       "traffic": cmp64["traffic"], "diag_step_ms": cmp64["diag_step_ms"]}
        out["speedup_vs_fp64"] = head["value"] / cmp64["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
