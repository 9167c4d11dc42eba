"""Init modes of the drop-in binary against the reference binary's frame 0.

libscp_io's scp_do_init (the reference do_init, scp_init.h:489-499) is run on
the same cfg as oracle/_ref/scp_sim (built from the reference sources), whose
snapshot frame 0 is the initial state (the real and complex modes: bitwise)
or, gauged, the state after its Gauss projection (scp_sim.c:2826-2831, 2940):
there the matter columns are compared bitwise and the oracle's restatement of
that projection, applied to this build's initial state, must land within
1e-12 of the reference's E. Cases: braid (the reference's verify_parity.cfg
values on a smaller grid), oscillon, qball from 2-, 3- and 4-column profile
tables (one ball and two, gauged and not; a 3/4-column table under the
complex mode must give the 2-column seed), template (an SFA stamped by
physical coordinate into a finer grid over the background, plain and with
the gradient background of bc_type 1) and exec (a generator command whose
stdout is the SFA). The profiles are synthetic tables of this test (the
format is what is under test, not a physical solution).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import core as oracle
from scp_amd import config as cfgmod
from scp_amd import sfa

REF_BIN = oracle.REF_BIN
pytestmark = pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference binary not built here")


def write_profile(path, ncol):
    r = np.arange(0.0, 12.0001, 0.05)
    f = 0.8 * np.exp(-r ** 2 / 8.0)
    er = 0.01 * r * np.exp(-r ** 2 / 10.0)
    we = 1.2 + 0.02 * np.exp(-r ** 2 / 5.0)
    cols = [r, f, er, we][:ncol]
    with open(path, "w") as fp:
        fp.write("# synthetic profile: r f [Er [weff]]\n")
        for row in zip(*cols):
            fp.write(" ".join(f"{v:.12e}" for v in row) + "\n")


def run_both(wd, keys, mode):
    """(reference frame 0, this build's initial state, cfg) for the cfg keys."""
    keys = dict(keys)
    N, L = keys["N"], keys["L"]
    dt = keys.get("dt_factor", 0.1) * 2 * L / (N - 1)
    keys.update(T=dt, snap_dt=1000.0, diag_dt=1000.0, output=os.path.join(wd, "o.sfa"),
                diag_file=os.path.join(wd, "d.tsv"), precision="f64")
    path = os.path.join(wd, "run.cfg")
    with open(path, "w") as fp:
        fp.write("".join(f"{k} = {v}\n" for k, v in keys.items()))
    res = subprocess.run([REF_BIN, path], capture_output=True, text=True, cwd=wd,
                         env=dict(os.environ, OMP_NUM_THREADS="4"))
    assert res.returncode == 0, res.stdout[-800:] + res.stderr[-800:]
    with sfa.SFAFile(os.path.join(wd, "o.sfa")) as fo:
        ref0 = fo.read_frame(0)
    c = cfgmod.Config()
    c.load(path)
    assert c.validate() is None
    mine = cfgmod.do_init(c, mode)
    return ref0, mine, c


def compare(ref0, mine, mode, c):
    if mode != 2 or float(c.g_gauge) == 0.0:
        assert np.array_equal(ref0, mine)
        return
    assert np.array_equal(ref0[:27], mine[:27])                  # matter and links untouched
    p = oracle.make_params(c.N, 2, c.L, c.dt_factor, g=c.g_gauge, bc_type=c.bc_type)
    proj, _ = oracle.gauss_project(p, mine)
    scale = max(1.0, float(np.abs(ref0[27:]).max()))
    assert np.abs(proj[27:] - ref0[27:]).max() <= 1e-12 * scale


BRAID = dict(N=24, L=6.0, dt_factor=0.025, m=1.5, m_theta=0.0, eta=0.5, mu=-41.345, kappa=50.0, mode=0,
             damp_width=3.0, damp_rate=0.01, init="braid", A=0.8, A_bg=0.1, R_tube=3.0, ellip=0.3325,
             delta="0,3.0005,4.4325")


@pytest.mark.parametrize("case", ["braid", "oscillon", "qball2_complex", "qball2_two_balls",
                                  "qball3_complex_ignores_er", "qball3_gauged", "qball4_gauged_pair",
                                  "qball2_gauged"])
def test_init_mode_matches_reference_frame0(case):
    with tempfile.TemporaryDirectory() as wd:
        mode = 0
        if case == "braid":
            keys = BRAID
        elif case == "oscillon":
            keys = dict(N=20, L=5.0, init="oscillon", A=0.9, sigma=1.7, delta="0.1,2.0,4.1")
        else:
            ncol = int(case[5])
            prof = os.path.join(wd, "prof.txt")
            write_profile(prof, ncol)
            keys = dict(N=20, L=6.0, complex_phi=1, init="qball", qball_profile=prof, qball_omega=1.35,
                        qball_x0=0.7, qball_y0=-0.4, qball_z0=0.2)
            mode = 1
            if "two_balls" in case or "pair" in case:
                keys.update(qball2_x0=-2.1, qball2_y0=1.0, qball2_z0=-0.3, qball2_sign=-1, qball2_phase=0.9)
            if "gauged" in case:
                keys.update(complex_gauge=1, g_gauge=0.05)
                mode = 2
                if "pair" not in case:
                    keys["bc_type"] = 0          # one ball is net-charged: periodic gauged runs refuse it
        ref0, mine, c = run_both(wd, keys, mode)
        compare(ref0, mine, mode, c)


def test_qball_3col_complex_equals_2col():
    """Extra profile columns are honoured only when gauged (scp_init.h:377-386)."""
    with tempfile.TemporaryDirectory() as wd:
        out = []
        for ncol in (2, 3):
            prof = os.path.join(wd, f"p{ncol}.txt")
            write_profile(prof, ncol)
            c = cfgmod.Config()
            for k, v in dict(N=12, L=4.0, complex_phi=1, init="qball", qball_profile=prof,
                             qball_omega=1.35).items():
                c.set(k, v)
            out.append(cfgmod.do_init(c, 1))
        assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("bc", [2, 1])
def test_init_template_matches_reference_frame0(bc):
    with tempfile.TemporaryDirectory() as wd:
        TN, TL = 12, 3.0
        rng = np.random.default_rng(3)
        tst = 0.2 * rng.standard_normal((12, TN ** 3))
        tpath = os.path.join(wd, "tmpl.sfa")
        sfa.write_sim_sfa(tpath, TN, TL, 0.05, [(0.0, 0.5 * tst), (1.0, tst)])   # the last frame is used
        keys = dict(N=24, L=6.3, init="template", init_sfa=tpath, A_bg=0.1, bc_type=bc)
        if bc == 1:
            keys.update(gradient_A_high=0.15, gradient_A_low=0.05, gradient_margin=2)
        ref0, mine, c = run_both(wd, keys, 0)
        compare(ref0, mine, 0, c)


def test_init_exec_matches_reference_frame0():
    with tempfile.TemporaryDirectory() as wd:
        N, L = 10, 3.0
        st = 0.3 * np.random.default_rng(4).standard_normal((24, N ** 3))
        seed = os.path.join(wd, "gen.sfa")
        sfa.write_sim_sfa(seed, N, L, 0.05, [(0.0, st)])
        ref0, mine, c = run_both(wd, dict(N=N, L=L, complex_phi=1, init="exec", init_exec=f"cat {seed}"), 1)
        compare(ref0, mine, 1, c)
        assert np.array_equal(mine, st)
