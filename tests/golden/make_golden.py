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
and, for the vector-frame cases (VEC_CASES), every FRVD payload parsed plus
the f32 fields at the I-frame times.
Run from the repo root:  python tests/golden/make_golden.py [case ...]
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


# Vector frames (FRVD, real mode): name -> (N, L, block size, I-frame interval, delta tol, T).
# Snapshots every vector frame in f32 (the fit reads the fields rounded to f32).
VEC_CASES = {
    "real_vec_n16": (16, 3.5, 8, 4, 0.075, 1.4),
    "real_vec_n12_bs5": (12, 2.75, 5, 3, 0.05, 1.0),
}


def run_vec_case(name, spec, workdir):
    """The reference's vector frames: every FRVD payload, parsed, and the f32
    voxel frame (phi and theta columns) at the first I-frame. The seed is
    rounded to f32 (stored that way) before the reference runs from it."""
    N, L, bs, iint, tol, T = spec
    rng = np.random.default_rng(sum(map(ord, name)))
    dtf = 0.1
    dx = 2 * L / (N - 1)
    dt = dtf * dx
    vdt = 4 * dt
    st = seed_state(N, L, 0, rng, [(0.3, -0.2, 0.1, 0.8, 1.2 * dx * 3, 0, 1)], vel_noise=0.05)
    st = st.astype(np.float32).astype(np.float64)
    seed_path = os.path.join(workdir, name + "_seed.sfa")
    sfa.write_sim_sfa(seed_path, N, L, dt, [(0.0, st)], codec=sfa.CODEC_RAW)
    cfg = {"N": N, "L": L, "T": T, "dt_factor": dtf, "precision": "f32", "bc_type": 2,
           "snap_dt": vdt, "diag_dt": T / 2, "vec_snap_dt": vdt, "vec_block_size": bs,
           "vec_iframe_interval": iint, "vec_delta_tol": tol, "init": "sfa", "init_sfa": seed_path,
           "output": os.path.join(workdir, name + ".sfa"), "diag_file": os.path.join(workdir, name + "_diag.tsv")}
    cfg_path = os.path.join(workdir, name + ".cfg")
    with open(cfg_path, "w") as f:
        for k, v in cfg.items():
            f.write(f"{k} = {v}\n")
    env = dict(os.environ, OMP_NUM_THREADS="4")
    res = subprocess.run([REF, cfg_path], capture_output=True, text=True, env=env, cwd=workdir)
    if res.returncode != 0:
        raise RuntimeError(f"{name}: reference failed\n{res.stdout}\n{res.stderr}")
    out = {}
    vk, last_voxel = 0, None
    with sfa.SFAFile(cfg["output"]) as f:
        types = [f.frame_type(i) for i in range(f.total_frames)]
        for i, t in enumerate(types):
            if t == 0:
                last_voxel = i
                continue
            v = f.read_vec(i)
            pre = f"v{vk}_"
            out[pre + "time"] = v["time"]
            out[pre + "type"] = t
            if t == 1:
                for k in ("origins", "coeffs", "mean", "amp", "phase"):
                    out[pre + k] = np.array(v[k])
                out[pre + "omega"] = v["omega"]
                if vk == 0:   # the voxel frame of the same step (the snapshot precedes the vector frame)
                    fr = f.read_frame(last_voxel, as_f64=False)
                    out["fields0"] = np.stack([fr[c] for c in (0, 1, 2, 3, 4, 5)])
                    out["fields0_time"] = f.frame_time(last_voxel)
            else:
                out[pre + "indices"] = np.array(v["indices"])
                out[pre + "deltas"] = np.array(v["deltas"])
            vk += 1
        times = np.array([f.frame_time(i) for i in range(f.total_frames)])
    params = {k: v for k, v in cfg.items() if k not in ("output", "diag_file", "init_sfa")}
    np.savez_compressed(os.path.join(OUT, name + ".npz"), params=json.dumps(params), frame0=st.astype(np.float32),
                        types=np.array(types), times=times, n_vec=vk, stdout=res.stdout, **out)
    kinds = "".join("IP"[t - 1] for t in types if t)
    nsel = [len(out[f"v{k}_indices"]) for k in range(vk) if f"v{k}_indices" in out]
    print(f"{name}: {len(types)} frames, vector frames {kinds}, P-frame patches {nsel}, "
          f"bytes={os.path.getsize(os.path.join(OUT, name + '.npz'))}")


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle ref` first")
    names = sys.argv[1:] or list(CASES) + list(VEC_CASES)
    with tempfile.TemporaryDirectory() as wd:
        for n in names:
            if n in VEC_CASES:
                run_vec_case(n, VEC_CASES[n], wd)
            else:
                run_case(n, CASES[n], wd)


if __name__ == "__main__":
    main()
