"""Benchmark of the MI355X sfa/sim step: the 512^3 gauged leapfrog step.

Workload (one GPU): gauged mode, 512^3 periodic, dx=0.5 (L=127.75), dt=0.1,
64 Gaussian Q-balls at seeded uniform positions (amp U[0.6,0.9], width
U[2,4] cells, omega U[0.6,0.95], random phase, alternating charge signs),
theta noise 0.01, links at identity, E = 0 — synthetic data generated on the
host from a fixed seed and resident in HBM before the timed region.

A "step" is one full leapfrog step of the whole grid (one fused launch).
value = site-updates/s over K timed steps (HIP events on the grid's stream).
roofline: algorithmic HBM bytes per site-update = every persistent state
scalar read and written once at its storage width (gauged: 30 scalars ->
480 B in FP64, 240 B in FP32), times N^3 sites per launch, over the average
launch time; peak = 8 TB/s. cpu_baseline: the reference CPU binary built from
source (oracle/_ref/scp_sim) on a bounded 64^3 gauged sample on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--prec f64|f32c]
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
BYTES_PER_SITE = {("gauged", "f64"): 480, ("gauged", "f32c"): 240}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(seconds_hint=15.0):
    """Time the reference CPU path (oracle/_ref/scp_sim, built from the reference
    sources with its Makefile flags -O3 -fopenmp) on a 64^3 gauged sample."""
    ref = os.path.join(ROOT, "oracle", "_ref", "scp_sim")
    if not os.path.exists(ref):
        return None
    cores = min(16, os.cpu_count() or 1)
    N, L, dtf = 64, 15.75, 0.2
    dt = dtf * (2 * L / (N - 1))
    nsteps = 600
    with tempfile.TemporaryDirectory() as wd:
        cfg = os.path.join(wd, "b.cfg")
        with open(cfg, "w") as f:
            f.write(f"N = {N}\nL = {L}\nT = {nsteps * dt}\ndt_factor = {dtf}\ncomplex_phi = 1\n"
                    f"complex_gauge = 1\ng_gauge = 0.05\nbc_type = 2\ninit = oscillon\n"
                    f"precision = f32\nsnap_dt = {10 * nsteps * dt}\ndiag_dt = {nsteps * dt}\n"
                    f"output = {wd}/o.sfa\ndiag_file = {wd}/d.tsv\n")
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        t0 = time.time()
        res = subprocess.run([ref, cfg], capture_output=True, text=True, env=env, cwd=wd, timeout=600)
        wall = time.time() - t0
    m = re.search(r"Wall: [\d.]+s \([\d.]+ min\) ([\d.]+)ms/step", res.stdout)
    if res.returncode != 0 or not m:
        log(f"cpu baseline failed rc={res.returncode}")
        return None
    ms = float(m.group(1))
    return {"value": N ** 3 / (ms * 1e-3) / 1e9, "unit": "Gsite-updates/s", "cores": cores,
            "kind": "reference",
            "sample": f"reference scp_sim (gcc -O3 -fopenmp), gauged 64^3 periodic, init=oscillon, "
                      f"{nsteps} steps, OMP_NUM_THREADS={cores}, 'Wall' ms/step={ms:.2f}, "
                      f"process wall {wall:.1f}s"}


def build_state(N, L, seed=1002):
    from scp_amd import seeds
    balls = seeds.draw_balls(N, L, 64, seed)
    return seeds.seed_balls(N, L, 2, balls, theta_sigma=0.01, link_sigma=0.0, seed=seed)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--prec", default="f64", choices=["f64", "f32c"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU (slab halo exchange "
                                   "over RCCL) is a design, see DESIGN.md", "n_gpus": a.gpus}))
        sys.exit(2)
    from scp_amd import grid as G

    N = a.n
    L = 0.25 * (N - 1)            # dx = 0.5
    prec = G.PREC_F64 if a.prec == "f64" else G.PREC_F32C
    params = G.make_params(N, G.MODE_GAUGED, L, 0.2, precision=prec, g=0.05)
    log(f"seeding {N}^3 gauged state on the host")
    t0 = time.time()
    st = build_state(N, L)
    log(f"seeded in {time.time() - t0:.1f}s; device bytes {G.grid_bytes(params) / 1e9:.1f} GB")
    g = G.Grid(params)
    g.upload(st)
    del st
    log(f"uploaded; warmup {a.warmup} steps")
    if a.warmup > 0:
        g.step(a.warmup)
    log(f"timing {a.steps} steps")
    w0 = time.time()
    ms = g.time_steps(a.steps)
    wall = time.time() - w0
    d = g.diag()
    log(f"timed {ms:.2f} ms (wall {wall * 1e3:.1f} ms); E_total={d['E_total']:.6e} Q_total={d['Q_total']:.6e}")
    ms_step = ms / a.steps
    sites = float(N) ** 3
    value = sites / (ms_step * 1e-3) / 1e9
    bps = BYTES_PER_SITE[("gauged", a.prec)]
    achieved = bps * sites / (ms_step * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"traffic_gauged_{a.prec}_{N}.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    cpu = None if a.no_cpu_baseline else cpu_baseline()
    out = {
        "metric": "Gsite-updates/s & %HBM peak, 512^3 gauged, FP32c vs FP64; p50/p99 job latency",
        "value": value, "unit": "Gsite-updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": a.prec, "data": "synthetic (seeded Gaussian Q-balls, generated on the host)",
        "config": {"workload": f"gauged {N}^3 periodic leapfrog step, 64 Q-balls, dx=0.5 dt=0.1",
                   "N": N, "mode": "gauged", "precision": a.prec, "g": 0.05},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": g.step_kernel_name, "bytes_per_site": bps},
        "cpu_baseline": cpu,
    }
    g.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
