"""Generate the golden fixtures under tests/golden/ from the reference CPU binary.

The reference sim (sfa/sim/scp_sim.c) is compiled from its own sources by
`make -C oracle ref` into oracle/_ref/scp_sim (never committed). This script
runs it on small seeded states loaded through init=sfa and stores, per case:
  params   the config keys (JSON string)
  frame0   the reference's frame 0 (post Gauss projection for gauged runs)
  final    the reference's last frame (f64 snapshot)
  times    snapshot times
  diag     the reference diag.tsv text
  stdout   the reference console log (Gauss projection line, Wall line)
Run from the repo root:  python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from scp_amd import sfa  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "scp_sim")
OUT = os.path.join(ROOT, "tests", "golden")


def seed_state(N, L, mode, rng, balls, theta_sigma=0.01, link_sigma=0.0, vel_noise=0.0):
    """Gaussian balls on the periodic box plus noise (fixture inputs only)."""
    dx = 2 * L / (N - 1)
    x = -L + dx * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    ncol = sfa.ncols_for_mode(mode)
    st = np.zeros((ncol, N ** 3))
    box = N * dx

    def mi(d):
        return (d + box / 2) % box - box / 2

    for (x0, y0, z0, A, sig, om, wind) in balls:
        r2 = mi(X - x0) ** 2 + mi(Y - y0) ** 2 + mi(Z - z0) ** 2
        env = (A * np.exp(-r2 / (2 * sig * sig))).ravel()
        for a in range(3):
            ph = wind * 2 * np.pi * a / 3
            re, im = env * np.cos(ph), env * np.sin(ph)
            st[a] += re
            st[6 + a] += -om * im if mode else 0.3 * env
            if mode:
                st[12 + a] += im
                st[18 + a] += om * re
    st[3:6] = theta_sigma * rng.standard_normal((3, N ** 3))
    st[6:12] += vel_noise * rng.standard_normal((6, N ** 3))
    if mode:
        st[15:18] = theta_sigma * rng.standard_normal((3, N ** 3))
    if mode == 2:
        st[24:27] = link_sigma * rng.standard_normal((3, N ** 3))
    return st


CASES = {
    # name: (mode, N, L, keys, balls, noise kwargs, T, init override)
    "real_n8": (0, 8, 3.5, {}, [(0, 0, 0, 0.8, 1.5, 0, 1)], dict(vel_noise=0.05), 2.0, None),
    "real_terms_n8": (0, 8, 3.5, {"mode": 3, "kappa_gamma": 2.0, "kappa_h": 0.3, "alpha_cs": 0.2,
                                  "beta_h": 0.1}, [(0, 0, 0, 0.8, 1.5, 0, 1)], dict(vel_noise=0.05), 2.0, None),
    "real_mode1_n8": (0, 8, 3.5, {"mode": 1}, [(0, 0, 0, 0.8, 1.5, 0, 1)], dict(vel_noise=0.05), 2.0, None),
    "real_oscillon_n10": (0, 10, 4.5, {"bc_type": 0, "damp_width": 2.0, "damp_rate": 0.05},
                          None, {}, 2.0, "oscillon"),
    "complex_n8": (1, 8, 3.5, {}, [(0.3, -0.2, 0.1, 0.8, 1.5, 0.8, 1)], {}, 2.0, None),
    "complex_damp_n10": (1, 10, 4.5, {"bc_type": 0, "damp_width": 2.0, "damp_rate": 0.05},
                         [(0, 0, 0, 0.8, 1.5, 0.8, 1)], {}, 2.0, None),
    "gauged_n8": (2, 8, 3.5, {"complex_gauge": 1, "g_gauge": 0.05},
                  [(-2.0, 0, 0, 0.8, 1.2, 0.8, 1), (2.0, 0, 0, 0.8, 1.2, -0.8, -1)],
                  dict(link_sigma=1e-2), 2.0, None),
    "gauged_g03_n8": (2, 8, 3.5, {"complex_gauge": 1, "g_gauge": 0.3},
                      [(-2.0, 0, 0, 0.8, 1.2, 0.8, 1), (2.0, 0, 0, 0.8, 1.2, -0.8, -1)],
                      dict(link_sigma=0.3), 2.0, None),
    "real_gradient_n10": (0, 10, 4.5, {"bc_type": 1, "gradient_margin": 2},
                          [(0.7, -0.4, 0.3, 0.8, 1.5, 0, 1)], dict(vel_noise=0.05), 2.0, None),
    "complex_gradient_n10": (1, 10, 4.5, {"bc_type": 1, "gradient_margin": 2},
                             [(0.6, 0.5, -0.3, 0.8, 1.5, 0.8, 1)], dict(vel_noise=0.05), 2.0, None),
    "gauged_damp_n10": (2, 10, 4.5, {"complex_gauge": 1, "g_gauge": 0.1, "bc_type": 0,
                                     "damp_width": 2.0, "damp_rate": 0.05},
                        [(0, 0, 0, 0.8, 1.5, 0.8, 1)], dict(link_sigma=1e-2), 2.0, None),
}


def run_case(name, spec, workdir):
    mode, N, L, keys, balls, noise, T, init = spec
    rng = np.random.default_rng(sum(map(ord, name)))
    dtf = 0.1
    dx = 2 * L / (N - 1)
    cfg = {"N": N, "L": L, "T": T, "dt_factor": dtf, "precision": "f64", "bc_type": 2,
           "snap_dt": T, "diag_dt": T / 4, "output": os.path.join(workdir, name + ".sfa"),
           "diag_file": os.path.join(workdir, name + "_diag.tsv")}
    if mode >= 1:
        cfg["complex_phi"] = 1
    cfg.update(keys)
    seed = None
    if init is None:
        st = seed_state(N, L, mode, rng, balls, **noise)
        seed = st
        seed_path = os.path.join(workdir, name + "_seed.sfa")
        sfa.write_sim_sfa(seed_path, N, L, dtf * dx, [(0.0, st)], codec=sfa.CODEC_RAW)
        cfg["init"] = "sfa"
        cfg["init_sfa"] = seed_path
    else:
        cfg["init"] = init
    cfg_path = os.path.join(workdir, name + ".cfg")
    with open(cfg_path, "w") as f:
        for k, v in cfg.items():
            f.write(f"{k} = {v}\n")
    env = dict(os.environ, OMP_NUM_THREADS="4")
    res = subprocess.run([REF, cfg_path], capture_output=True, text=True, env=env, cwd=workdir)
    if res.returncode != 0:
        raise RuntimeError(f"{name}: reference failed\n{res.stdout}\n{res.stderr}")
    with sfa.SFAFile(cfg["output"]) as f:
        nf = f.total_frames
        frame0 = f.read_frame(0)
        final = f.read_frame(nf - 1)
        times = np.array([f.frame_time(i) for i in range(nf)])
        kv = f.kvmd()
    with open(cfg["diag_file"]) as fh:
        diag = fh.read()
    params = {k: v for k, v in cfg.items() if k not in ("output", "diag_file", "init_sfa")}
    extra = {"seed": seed} if (seed is not None and mode == 2) else {}
    np.savez_compressed(os.path.join(OUT, name + ".npz"), params=json.dumps(params),
                        frame0=frame0, final=final, times=times, diag=diag,
                        stdout=res.stdout, kvmd=json.dumps(kv), **extra)
    if name == "complex_n8":
        # a small archive written by the reference itself (COLZSTD f64), kept as a
        # byte-level fixture for the reader and the header layout
        with open(cfg["output"], "rb") as fi, open(os.path.join(OUT, "ref_complex_n8.sfa"), "wb") as fo:
            fo.write(fi.read())
    print(f"{name}: frames={nf} t={times.tolist()} bytes={os.path.getsize(os.path.join(OUT, name + '.npz'))}")


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` first")
    names = sys.argv[1:] or list(CASES)
    with tempfile.TemporaryDirectory() as wd:
        for n in names:
            run_case(n, CASES[n], wd)


if __name__ == "__main__":
    main()
