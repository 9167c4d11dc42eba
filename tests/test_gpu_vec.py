"""Vector frames (FRVD) of the real mode on the GPU, through the C-ABI and the CLI.

Bars:
  * scp_grid_vec_fit against the oracle's restatement of fit_patch
    (scp_sim.c:2567-2640; the oracle is pinned to the reference's I-frames in
    tests/test_oracle.py): at least 99 % of the coefficients bitwise, the rest
    within 2e-7 of the patch's largest coefficient (both cast an FP64 solution
    to FP32; the GPU solves the separable form, the reference the 64 x 64
    normal equations);
  * the CLI's vector frames against the reference binary's (tests/golden/
    real_vec_*.npz): the same I/P sequence, times, origins, omega and P-frame
    patch choice; coefficients within 2e-7 and residuals within 1e-5 of scale.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import core as oracle
from scp_amd import grid as G
from scp_amd import sfa
from tests import fixtures
from tests.test_oracle import VEC_FIXTURES, _compare_vec_frames

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "scp_amd", "bin", "scp_sim")


def _state(N, L, seed):
    rng = np.random.default_rng(seed)
    dx = 2 * L / (N - 1)
    x = -L + dx * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    st = 0.05 * rng.standard_normal((12, N ** 3))
    st[0] += 0.8 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 4.0).ravel()
    return st


@pytest.mark.parametrize("N,bs", [(16, 8), (17, 5), (16, 4), (32, 16), (20, 8)])
def test_vec_fit_matches_oracle(N, bs):
    L = 0.25 * (N - 1)
    st = _state(N, L, N * 31 + bs)
    print(f"case N={N} bs={bs}", flush=True)
    with G.Grid(G.make_params(N, 0, L, 0.1)) as g:
        g.upload(st)
        got = g.vec_fit(bs)
    ref = oracle.vec_fit(st[[0, 1, 2, 3, 4, 5]].astype(np.float32), N, bs)
    assert got.shape == ref.shape == ((N // bs) ** 3, 6, 64)
    exact = np.mean(got == ref)
    scale = np.abs(ref).max(axis=2, keepdims=True)
    err = (np.abs(got.astype(np.float64) - ref) / scale).max()
    print(f"N={N} bs={bs}: {exact:.5f} bitwise, largest difference {err:.2e} of the patch scale")
    assert exact >= 0.99
    assert err <= 2e-7


def test_vec_fit_refusals():
    with G.Grid(G.make_params(12, 1, 2.75, 0.1)) as g:
        with pytest.raises(G.ScpError, match="real-mode"):
            g.vec_fit(4)
    with G.Grid(G.make_params(12, 0, 2.75, 0.1)) as g:
        for bs in (3, 13):
            with pytest.raises(G.ScpError, match="vec_block_size"):
                g.vec_fit(bs)


def _run_fixture(tmp_path, name):
    z = np.load(os.path.join(fixtures.GOLDEN, name + ".npz"))
    p = json.loads(str(z["params"]))
    N, L = int(p["N"]), float(p["L"])
    seed = tmp_path / "seed.sfa"
    sfa.write_sim_sfa(seed, N, L, 0.1 * 2 * L / (N - 1), [(0.0, z["frame0"].astype(np.float64))])
    cfg = dict(p, init_sfa=str(seed), output=str(tmp_path / "out.sfa"), diag_file=str(tmp_path / "diag.tsv"))
    path = tmp_path / "run.cfg"
    path.write_text("".join(f"{k} = {v}\n" for k, v in cfg.items()))
    res = subprocess.run([BIN, str(path)], capture_output=True, text=True, cwd=tmp_path, timeout=300)
    return z, res, tmp_path / "out.sfa"


@pytest.mark.parametrize("name", VEC_FIXTURES)
def test_cli_vector_frames_match_reference(tmp_path, name):
    z, res, out = _run_fixture(tmp_path, name)
    print(res.stdout[-1500:], res.stderr[-1500:], flush=True)
    assert res.returncode == 0
    ref_out = str(z["stdout"])
    for key in ("Vec frames: BS=", "Vec frames: "):
        mine = [l for l in res.stdout.splitlines() if l.startswith(key)]
        theirs = [l for l in ref_out.splitlines() if l.startswith(key)]
        assert [m.split(" written to")[0] for m in mine] == [t.split(" written to")[0] for t in theirs]
    _compare_vec_frames(out, z, first_frame=1)
    with sfa.SFAFile(out) as f:           # the f32 snapshot at the first vector frame (bitwise: real mode)
        types = [f.frame_type(i) for i in range(f.total_frames)]
        assert types == [int(t) for t in z["types"]]
        i = types.index(1) - 1
        assert f.frame_time(i) == float(z["fields0_time"])
        fr = f.read_frame(i, as_f64=False)
        assert np.array_equal(np.stack([fr[c] for c in range(6)]), z["fields0"])


def test_cli_restart_reads_the_last_voxel_frame(tmp_path):
    """A restart from a file that ends with vector frames continues from the last
    voxel frame (vector frames hold fitted patches and no velocities)."""
    z, res, out = _run_fixture(tmp_path, "real_vec_n16")
    assert res.returncode == 0
    with sfa.SFAFile(out) as f:
        types = [f.frame_type(i) for i in range(f.total_frames)]
    last_voxel = max(i for i, t in enumerate(types) if t == 0)
    r = subprocess.run([BIN, str(out), "-output", str(tmp_path / "cont.sfa"), "-diag_file",
                        str(tmp_path / "cont.tsv"), "-vec_snap_dt", "0", "-T", "0.2"],
                       capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"reading frame {last_voxel}" in r.stdout
