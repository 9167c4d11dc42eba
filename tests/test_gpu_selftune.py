"""Self-tune (kappa at the stability edge, fixed charge) through the C-ABI and the CLI.

The reference runs it in its GPU binary only (scp_sim.cu:5475-5547, real mode):
every st_dt, the diagnostics, a projection phi, phi_vel *= sqrt(Q0/Q) that keeps
the stored acceleration, and a kappa update. Bars:
  * scp_grid_rescale_phi and scp_grid_set_kappa between steps: bitwise equal to
    the oracle stepping with the same scaling / kappa and the stale acceleration
    carried across (FP64 real and complex modes are bitwise steps);
  * the CLI's self-tune run against the oracle's restatement of the loop
    (oracle.self_tune_run): the same update lines (kappa, P_max, Q, branch) at
    the printed precision and the final frame within 1e-12 relative (the scaling
    factor comes from FP64 reductions, whose summation order differs).
"""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import core as oracle
from scp_amd import grid as G
from scp_amd import sfa
from tests.test_oracle import _st_state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "scp_amd", "bin", "scp_sim")


@pytest.mark.parametrize("mode", [0, 1])
def test_rescale_phi_keeps_the_stored_acceleration(mode):
    N = 12
    L = 0.25 * (N - 1)
    st = _st_state(N, L, 3)
    if mode == 1:
        st = np.concatenate([st, 0.5 * st])
    f = 0.93
    p = oracle.make_params(N, mode, L, 0.1)
    a, acc = oracle.step(p, st, 7)
    cols = [0, 1, 2, 6, 7, 8] + ([12, 13, 14, 18, 19, 20] if mode == 1 else [])
    a[cols] *= f
    ref, _ = oracle.step(p, a, 5, acc=acc)
    with G.Grid(G.make_params(N, mode, L, 0.1)) as g:
        g.upload(st)
        g.step(7)
        g.rescale_phi(f)
        g.step(5)
        got = g.download()
    assert np.array_equal(got, ref)


def test_set_kappa_between_steps():
    N = 12
    L = 0.25 * (N - 1)
    st = _st_state(N, L, 4)
    p = oracle.make_params(N, 0, L, 0.1)
    a, acc = oracle.step(p, st, 6)
    p.kappa = 37.5
    ref, _ = oracle.step(p, a, 6, acc=acc)
    with G.Grid(G.make_params(N, 0, L, 0.1)) as g:
        g.upload(st)
        g.step(6)
        g.set_kappa(37.5)
        g.step(6)
        got = g.download()
    assert np.array_equal(got, ref)
    with G.Grid(G.make_params(N, 0, L, 0.1)) as g:
        with pytest.raises(G.ScpError, match="kappa"):
            g.set_kappa(0.0)


@pytest.mark.parametrize("ptarget", [0.0, 0.01])
def test_cli_self_tune_matches_restatement(tmp_path, ptarget):
    N = 12
    L = 0.25 * (N - 1)
    dx = 2 * L / (N - 1)
    dt = 0.1 * dx
    every, n_steps = 6, 36
    st = _st_state(N, L, 7)
    seed = tmp_path / "seed.sfa"
    sfa.write_sim_sfa(seed, N, L, dt, [(0.0, st)])
    cfg = {"N": N, "L": L, "T": n_steps * dt, "dt_factor": 0.1, "bc_type": 2, "precision": "f64",
           "snap_dt": n_steps * dt, "diag_dt": n_steps * dt, "init": "sfa", "init_sfa": str(seed),
           "self_tune": 1, "st_dt": every * dt, "st_pcrit": 0.003, "st_ptarget": ptarget,
           "output": str(tmp_path / "out.sfa"), "diag_file": str(tmp_path / "diag.tsv")}
    path = tmp_path / "run.cfg"
    path.write_text("".join(f"{k} = {v}\n" for k, v in cfg.items()))
    res = subprocess.run([BIN, str(path)], capture_output=True, text=True, cwd=tmp_path, timeout=300)
    print(res.stdout[-2500:], res.stderr[-1500:], flush=True)
    assert res.returncode == 0
    assert f"Self-tune: kappa dynamical, every {every} steps" in res.stdout
    p = oracle.make_params(N, 0, L, 0.1)
    fin, ev = oracle.self_tune_run(p, st, n_steps, every, p.m2, pcrit=0.003, ptarget=ptarget)
    lines = [l for l in res.stdout.splitlines() if l.startswith("[self-tune]")]
    want = []
    for (t, k, Pmax, Q, tag) in ev:
        if ptarget > 0:
            want.append(f"[self-tune] t={t:6.1f} kappa={k:9.4f} Pmax={Pmax:10.3e} target={ptarget:.4f} "
                        f"err={tag:+.3f} Q={Q:9.3e}")
        else:
            want.append(f"[self-tune] t={t:6.1f} kappa={k:9.4f} Pmax={Pmax:10.3e} Q={Q:9.3e} {tag}")
    assert lines == want
    with sfa.SFAFile(tmp_path / "out.sfa") as f:
        got = f.read_frame(f.total_frames - 1)
    assert np.linalg.norm(got - fin) <= 1e-12 * np.linalg.norm(fin)
    kv = re.search(r"kappa=\s*([\d.]+)", lines[-1])
    assert kv and abs(float(kv.group(1)) - p.kappa) < 1e-4
