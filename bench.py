"""Benchmark of the MI355X sfa/sim step: the 512^3 gauged leapfrog step.

Workload (one GPU): gauged mode, 512^3 periodic, dx=0.5 (L=127.75), dt=0.1,
64 Gaussian Q-balls at seeded uniform positions (amp U[0.6,0.9], width
U[2,4] cells, omega U[0.6,0.95], random phase, alternating charge signs),
theta noise 0.01, links at identity, E = 0 — synthetic data generated on the
host from a fixed seed and resident in HBM before the timed region.

A "step" is one full leapfrog step of the whole grid (one fused launch).
value = site-updates/s over K timed steps (HIP events on the grid's stream).
roofline: algorithmic HBM bytes per site-update = every persistent state
scalar read and written once at its storage width (gauged: 30 scalars; FP64
8 B -> 480 B; FP32C 4 B hi + 4/2/0 B compensation word for lo_words 0/1/2 ->
480/360/240 B), times N^3 sites per launch, over the average launch time;
peak = 8 TB/s. traffic: measured HBM bytes per launch (rocprofv3 FETCH_SIZE
and WRITE_SIZE scaled by a plain copy's counters), read from the counter
summary profiles/hbm_traffic.txt when it holds a line for this kernel and size. cpu_baseline: the reference CPU binary built from
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

def bytes_per_site(prec, lo_words):
    """30 gauged state scalars, each read once and written once per step."""
    width = 8 if prec == "f64" else 4 + {0: 4, 1: 2, 2: 0}[lo_words]
    return 2 * 30 * width


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
    ap.add_argument("--lo-words", type=int, default=0, help="FP32C compensation words: 0 f32, 1 bf16, 2 none")
    ap.add_argument("--copy-calib", type=float, default=0.0,
                    help="instead of the step: time a plain device copy of this many GiB (HBM counter calibration)")
    a = ap.parse_args()
    if a.copy_calib > 0:
        copy_calibration(a.copy_calib)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only: the multi-GPU form (slabs along "
                                   "the march axis with one-plane halo exchange over RCCL) is not built",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    from scp_amd import grid as G

    N = a.n
    L = 0.25 * (N - 1)            # dx = 0.5
    prec = G.PREC_F64 if a.prec == "f64" else G.PREC_F32C
    params = G.make_params(N, G.MODE_GAUGED, L, 0.2, precision=prec, g=0.05, lo_words=a.lo_words)
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
    bps = bytes_per_site(a.prec, a.lo_words)
    achieved = bps * sites / (ms_step * 1e-3) / 1e9
    traffic = measured_traffic(g.step_kernel_name, N)
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
