"""The fixed-charge gradient flow on the GPU (scp_qflow_*, bin/eta_qflow).

Bars:
  * against the oracle's restatement of sfa/seed/eta_qflow.c (pinned bitwise to
    that program in tests/test_oracle.py), same seed, same Q and step: after
    every iteration count printed, omega, E_Q, s_max and N within 1e-12
    relative, max|F| within 1e-12 of the unit force scale, and the final state within 1e-12 relative field L2 (the only
    difference is the order of the FP64 reductions of N and the energy);
  * the drop-in program bin/eta_qflow against the reference program built from
    source (oracle/_ref/eta_qflow), same command line: every printed row to its
    printed precision, the header's N_init, Q and g_max lines equal, and the
    f32 rotating seed equal (entries at most one f32 ulp apart or, where the
    seed's Theta cancels to rounding noise, within 1e-12 of the column's
    scale: the reductions round differently);
  * the fixed-(Q, J) mode (fixJ = 1) against the oracle's restatement of it
    (rows, multipliers, final state and its Jhat within 1e-12) and the drop-in
    program against the reference program (rows to printed precision, final
    lines equal, the seed with its rotational velocity terms to one f32 ulp).
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import core as oracle
from scp_amd import grid as G
from scp_amd import seeds
from scp_amd import sfa
from tests import qflow_cases as qc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "scp_amd", "bin", "eta_qflow")
REF = os.path.join(ROOT, "oracle", "_ref", "eta_qflow")

CASES = {"ball_bvp": {}, "ring_wind1": dict(wind=1, wtor=2.0), "ring_twist": dict(wind=1, wtor=2.5, twist=1)}


@pytest.mark.parametrize("N", [20, 33])
@pytest.mark.parametrize("case", list(CASES))
def test_qflow_matches_oracle(tmp_path, case, N):
    L, om0, iters, pev = 0.25 * (N - 1), 1.35, 150, 50
    prof = tmp_path / "prof.txt"
    qc.write_profile(prof)
    st, _ = seeds.qflow_seed(N, L, prof, om0, qc.FLOW["eta"], qc.FLOW["mth"], **CASES[case])
    p = oracle.qflow_params(N, L, **qc.FLOW)
    n0 = oracle.qflow_norm(p, st)
    Q, dx = om0 * n0, 2 * L / (N - 1)
    print(f"case qflow {case} N={N}", flush=True)
    ref, orows, om_ref = oracle.qflow(p, st, Q, 0.04 * dx * dx, iters, pev)
    with G.QFlow(N, L, **qc.FLOW) as f:
        f.upload(st)
        nphi, nth = f.norm()
        assert abs(nphi + nth - n0) <= 1e-13 * n0
        rows, om = f.run(Q, 0.04 * dx * dx, iters, pev)
        out = f.download()
    assert [r[0] for r in rows] == [r[0] for r in orows] == [0, 50, 100, 150]
    for a, b in zip(rows, orows):
        # omega, E_Q, s_max, N relative; max|F| against the force scale 1 (the residual
        # force shrinks as the flow converges, its rounding does not)
        for q, (x, y) in enumerate(zip(a[1:], b[1:])):
            assert abs(x - y) <= 1e-12 * (max(abs(y), 1.0) if q == 2 else abs(y)), (case, q, a, b)
    assert abs(om - om_ref) <= 1e-13 * om_ref
    err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    print(f"final state vs oracle: {err:.2e}", flush=True)
    assert err <= 1e-12


def _numbers(rows):
    return [[float(x) for x in r] for r in rows]


@pytest.mark.skipif(not os.path.exists(REF), reason="the reference program (oracle/_ref/eta_qflow) is not built")
@pytest.mark.parametrize("case", ["ball_bvp", "ring_wind1"])
def test_qflow_cli_against_reference_program(tmp_path, case):
    N, L, om0, iters, pev = 32, 7.75, 1.35, 400, 100
    prof = tmp_path / "prof.txt"
    qc.write_profile(prof)
    res = {}
    for name, binary in (("ref", REF), ("gpu", BIN)):
        out = tmp_path / f"{name}.sfa"
        print(f"case qflow cli {case}: {name}", flush=True)
        text, rows = qc.run(qc.argv(binary, N, L, prof, om0, out, iters, pev, **CASES[case]), threads=8)
        with sfa.SFAFile(out) as f:
            res[name] = (text, rows, f.read_frame(0).astype(np.float32))
    (t_ref, r_ref, c_ref), (t_gpu, r_gpu, c_gpu) = res["ref"], res["gpu"]
    head = lambda t: [l.split("dtau=")[0] for l in t.splitlines() if l.startswith("# thinit=")]
    assert head(t_ref) == head(t_gpu)
    assert [r[0] for r in r_ref] == [r[0] for r in r_gpu] == ["0", "100", "200", "300", "400"]
    for a, b in zip(_numbers(r_ref), _numbers(r_gpu)):
        # printed with 6 decimals (omega), 7 significant digits (E_Q), 5 (maxF, s_max), 4 decimals (N)
        tol = [0, 1e-6, 1e-6 * abs(a[2]), 1e-4 * abs(a[3]), 1e-4 * abs(a[4]) + 1e-300, 1e-4]
        for q in range(1, 6):
            assert abs(a[q] - b[q]) <= tol[q], (q, a, b)
    final = lambda t: [l for l in t.splitlines() if l.startswith("# final:")]
    assert final(t_ref) == final(t_gpu)
    # entries equal to one f32 ulp, or (rounding noise of ~1e-21 where the seed's Theta
    # cancels, e.g. on the (1,1,1) diagonal of the ball) within 1e-12 of the column's scale
    same = c_ref == c_gpu
    gap = np.abs(c_ref.astype(np.float64) - c_gpu.astype(np.float64))
    one_ulp = np.spacing(np.maximum(np.abs(c_ref), np.abs(c_gpu))).astype(np.float64)
    scale = np.abs(c_ref).max(axis=1, keepdims=True).astype(np.float64)
    print(f"f32 seed: {same.mean() * 100:.3f} % of entries equal, max gap {np.max(gap / one_ulp):.2f} ulp, "
          f"{np.max(gap / np.maximum(scale, 1e-300)):.2e} of the column scale", flush=True)
    assert np.all((gap <= one_ulp) | (gap <= 1e-12 * scale)) and same.mean() >= 0.98


JCASES = {"ring_wind1_J": (dict(wind=1, wtor=2.0), None), "ball_J": ({}, 0.05),
          "ring_wind1_Jcap": (dict(wind=1, wtor=2.0), 3.0)}


@pytest.mark.parametrize("N", [20, 33])
@pytest.mark.parametrize("case", list(JCASES))
def test_qflow_fixed_q_j_matches_oracle(tmp_path, case, N):
    """The fixed-(Q, J) mode against the oracle's restatement (pinned bitwise to the
    reference program in tests/test_oracle.py): after every printed iteration
    omega, Omega, E_QJ, s_max, N, J_act and cond within 1e-12 relative (cond and
    Omega against their scale 1 and 0.15: the cap and a near-degenerate solve keep
    them O(1)), max|F| within 1e-12 of the unit force scale; the final state and
    its Jhat within 1e-12 relative field L2."""
    L, om0, iters, pev = 0.25 * (N - 1), 1.35, 120, 40
    kw, jscale = JCASES[case]
    prof = tmp_path / "prof.txt"
    qc.write_profile(prof)
    st, _ = seeds.qflow_seed(N, L, prof, om0, qc.FLOW["eta"], qc.FLOW["mth"], **kw)
    p = oracle.qflow_params(N, L, **qc.FLOW)
    n0 = oracle.qflow_norm(p, st)
    Q, dx = om0 * n0, 2 * L / (N - 1)
    Jt = kw.get("wind", 0) * Q if jscale is None else jscale * Q
    print(f"case qflow fixJ {case} N={N}", flush=True)
    ref, orows, olast, ojh = oracle.qflow_j(p, st, Q, Jt, 0.02 * dx * dx, iters, pev)
    with G.QFlow(N, L, **qc.FLOW) as f:
        f.upload(st)
        rows, last = f.run_fixj(Q, Jt, 0.02 * dx * dx, iters, pev)
        out = f.download()
        jh = f.download_jhat()
    assert [r[0] for r in rows] == [r[0] for r in orows] == [0, 40, 80, 120]
    scale = [None, 0, 0.15, 0, 1.0, 0, 0, 0, 1.0]      # absolute floors per column (0: relative only)
    for a, b in zip(rows, orows):
        for q in range(1, 9):
            x, y = a[q], b[q]
            assert abs(x - y) <= 1e-12 * max(abs(y), scale[q]), (case, q, a, b)
    for q, (x, y) in enumerate(zip(last, olast)):
        assert abs(x - y) <= 1e-12 * max(abs(y), 1.0), (case, q, last, olast)
    if case == "ring_wind1_Jcap":
        assert abs(last[1]) == 0.15
    for name, a, b in (("state", out, ref), ("Jhat", jh, ojh)):
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"final {name} vs oracle: {err:.2e}", flush=True)
        assert err <= 1e-12


@pytest.mark.skipif(not os.path.exists(REF), reason="the reference program (oracle/_ref/eta_qflow) is not built")
@pytest.mark.parametrize("case", ["ring_wind1_J", "ball_J"])
def test_qflow_fixed_q_j_cli_against_reference_program(tmp_path, case):
    """bin/eta_qflow fixJ = 1 against the reference program, same command line: the
    header lines (the fixJ target among them) equal, every row to its printed
    precision, the final lines equal, the f32 seed equal as in the fixed-Q test."""
    N, L, om0, iters, pev = 32, 7.75, 1.35, 300, 100
    kw, jscale = JCASES[case]
    prof = tmp_path / "prof.txt"
    qc.write_profile(prof)
    jarg = 0.0
    if jscale is not None:
        st, _ = seeds.qflow_seed(N, L, prof, om0, qc.FLOW["eta"], qc.FLOW["mth"], **kw)
        jarg = jscale * om0 * oracle.qflow_norm(oracle.qflow_params(N, L, **qc.FLOW), st)
    res = {}
    for name, binary in (("ref", REF), ("gpu", BIN)):
        out = tmp_path / f"{name}.sfa"
        print(f"case qflow fixJ cli {case}: {name}", flush=True)
        text, rows = qc.run(qc.argv(binary, N, L, prof, om0, out, iters, pev, fixJ=1, Jtarget=jarg, **kw),
                            threads=8)
        with sfa.SFAFile(out) as f:
            res[name] = (text, rows, f.read_frame(0).astype(np.float32))
    (t_ref, r_ref, c_ref), (t_gpu, r_gpu, c_gpu) = res["ref"], res["gpu"]
    head = lambda t: [l.split("dtau=")[0] for l in t.splitlines() if l.startswith(("# thinit=", "# fixJ:"))]
    assert head(t_ref) == head(t_gpu) and len(head(t_ref)) == 2
    assert [r[0] for r in r_ref] == [r[0] for r in r_gpu] == ["0", "100", "200", "300"]
    for a, b in zip(_numbers(r_ref), _numbers(r_gpu)):
        # printed: omega, Omega 6 decimals; E_QJ 7 significant digits; maxF, s_max 5; N, J_act 4 decimals;
        # cond 4 significant digits: 1.5 units of the last printed digit (a rounding boundary)
        tol = [0, 1.5e-6, 1.5e-6, 1.5e-6 * abs(a[3]), 1.5e-4 * abs(a[4]), 1.5e-4 * abs(a[5]) + 1e-300, 1.5e-4,
               1.5e-4, 1.5e-3 * abs(a[8])]
        for q in range(1, 9):
            assert abs(a[q] - b[q]) <= tol[q], (q, a, b)
    final = lambda t: [l for l in t.splitlines() if l.startswith(("# final:", "# fixJ final:"))]
    assert final(t_ref) == final(t_gpu) and len(final(t_ref)) == 2
    same = c_ref == c_gpu
    gap = np.abs(c_ref.astype(np.float64) - c_gpu.astype(np.float64))
    one_ulp = np.spacing(np.maximum(np.abs(c_ref), np.abs(c_gpu))).astype(np.float64)
    scale = np.abs(c_ref).max(axis=1, keepdims=True).astype(np.float64)
    print(f"f32 seed: {same.mean() * 100:.3f} % of entries equal, max gap {np.max(gap / one_ulp):.2f} ulp", flush=True)
    assert np.all((gap <= one_ulp) | (gap <= 1e-12 * scale)) and same.mean() >= 0.98
