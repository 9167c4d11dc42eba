"""Pins of the oracle (CPU FP64 restatement) before it is trusted as the checker.

Each test is chosen so a plausible restatement mistake (dropped term, wrong
sign or index, transposed operand) fails: the reference binary's own outputs
(golden fixtures, and a live run where the reference builds here), the
reference's documented invariants (Noether charge, real-limit identity, gauge
invariance, Gauss law), and closed forms (lattice plane-wave dispersion,
uniform-field ODE, staple telescoping).
"""
import math
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import core as oracle
from scp_amd import sfa
from tests import fixtures

FIXTURES = fixtures.names()


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_step_bitwise_vs_reference(name):
    fx = fixtures.load(name)
    n, p = fixtures.nsteps(fx)
    st, _ = oracle.step(p, fx["frame0"], n)
    assert np.array_equal(st, fx["final"])


def _g_offset(p, st):
    N = p.N
    s = st.reshape(30, N, N, N)
    divE = sum(s[27 + d] - np.roll(s[27 + d], 1, axis=d) for d in range(3)) / p.dx
    rho = sum(s[a] * s[18 + a] - s[12 + a] * s[6 + a] + s[3 + a] * s[21 + a] - s[15 + a] * s[9 + a]
              for a in range(3))
    return float(np.mean(divE - p.g * rho))


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_diag_vs_reference_tsv(name):
    """every diag.tsv row, to the printed precision (%.6e / %.12e)"""
    fx = fixtures.load(name)
    n, p = fixtures.nsteps(fx)
    head, rows = fixtures.parse_diag(fx["diag"])
    every = int(round(fx["diag_dt"] / p.dt))
    if fx["mode"] == 2:
        p.G_offset = _g_offset(p, fx["frame0"])
    st, acc = fx["frame0"], None
    for r, row in enumerate(rows):
        if r > 0:
            st, acc = oracle.step(p, st, every, acc=acc)
        d = oracle.diag(p, st)
        for col, val in zip(head[1:], row[1:]):
            if col in ("gauss_max", "gauss_l2"):
                assert d[col] <= 1e-12 and val <= 1e-12
                continue
            tol = 5e-7 * abs(val) + 1e-14
            if col in ("Q_phi", "Q_theta", "Q_total", "omega_core", "Q_core", "E_em", "Q_flux",
                       "Q_p0", "Q_p1", "Q_p2"):
                tol = 5e-12 * abs(val) + 1e-12
            assert abs(d[col] - val) <= tol, f"{name} row {r} {col}: {d[col]!r} vs {val!r}"


REF_BIN = oracle.REF_BIN


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference binary not built here")
@pytest.mark.parametrize("mode", [1, 2])
def test_oracle_vs_live_reference_binary(mode):
    """A fresh random case (not one of the fixtures) through oracle/_ref."""
    N, L, steps = 9, 4.0, 15
    rng = np.random.default_rng(100 + mode)
    st = np.zeros((oracle.ncols(mode), N ** 3))
    st[:24] = 0.3 * rng.standard_normal((24, N ** 3))
    if mode == 2:
        st[6:12] = 0.0          # zero velocities: net-neutral by construction (passes C:2264)
        st[18:24] = 0.0
        st[24:27] = 0.4 * rng.standard_normal((3, N ** 3))
    dx = 2 * L / (N - 1)
    with tempfile.TemporaryDirectory() as wd:
        seed = os.path.join(wd, "seed.sfa")
        sfa.write_sim_sfa(seed, N, L, 0.1 * dx, [(0.0, st)])
        cfg = os.path.join(wd, "c.cfg")
        T = steps * 0.1 * dx
        with open(cfg, "w") as f:
            f.write(f"N={N}\nL={L}\nT={T}\ndt_factor=0.1\ncomplex_phi=1\nbc_type=2\nprecision=f64\n"
                    f"init=sfa\ninit_sfa={seed}\nsnap_dt={T}\ndiag_dt={T}\noutput={wd}/o.sfa\n"
                    f"diag_file={wd}/d.tsv\neta=0.7\nm_theta=0.3\n")
            if mode == 2:
                f.write("complex_gauge=1\ng_gauge=0.2\n")
        res = subprocess.run([REF_BIN, cfg], capture_output=True, text=True, cwd=wd,
                             env=dict(os.environ, OMP_NUM_THREADS="2"))
        assert res.returncode == 0, res.stderr
        with sfa.SFAFile(os.path.join(wd, "o.sfa")) as fo:
            frame0, final = fo.read_frame(0), fo.read_frame(fo.total_frames - 1)
    p = oracle.make_params(N, mode, L, 0.1, eta=0.7, mtheta2=0.3 * 0.3, g=0.2)
    if mode == 2:
        proj, _ = oracle.gauss_project(p, st)
        assert np.abs(proj - frame0).max() <= 1e-12       # E after projection
    out, _ = oracle.step(p, frame0, int(round(T / p.dt)))
    assert np.array_equal(out, final)


def _bc_switch_case(wd, extra=""):
    """Complex 10^3 seed with a strong absorbing sphere, run through oracle/_ref
    for 40 steps; returns (params, frame0, final, stdout)."""
    N, L, steps = 10, 4.5, 40
    rng = np.random.default_rng(77)
    st = 0.3 * rng.standard_normal((24, N ** 3))
    dx = 2 * L / (N - 1)
    seed = os.path.join(wd, "seed.sfa")
    sfa.write_sim_sfa(seed, N, L, 0.1 * dx, [(0.0, st)])
    cfg = os.path.join(wd, "c.cfg")
    T = steps * 0.1 * dx
    with open(cfg, "w") as f:
        f.write(f"N={N}\nL={L}\nT={T}\ndt_factor=0.1\ncomplex_phi=1\nbc_type=0\ndamp_width=2.0\n"
                f"damp_rate=0.05\nprecision=f64\ninit=sfa\ninit_sfa={seed}\nsnap_dt={T}\ndiag_dt={T}\n"
                f"output={wd}/o.sfa\ndiag_file={wd}/d.tsv\n{extra}")
    res = subprocess.run([REF_BIN, cfg], capture_output=True, text=True, cwd=wd,
                         env=dict(os.environ, OMP_NUM_THREADS="2"))
    assert res.returncode == 0, res.stderr
    with sfa.SFAFile(os.path.join(wd, "o.sfa")) as fo:
        frame0, final = fo.read_frame(0), fo.read_frame(fo.total_frames - 1)
    p = oracle.make_params(N, 1, L, 0.1, bc_type=0, damp_width=2.0, damp_rate=0.05)
    return p, frame0, final, res.stdout


@pytest.mark.skipif(not os.path.exists(REF_BIN), reason="reference binary not built here")
def test_oracle_bc_switch_vs_live_reference_binary():
    """bc_switch_time (scp_sim.c:3073-3081): the reference sets damp_rate = 0 before
    step lround(bc_switch_time / dt) = 17. The oracle's 16 damped steps followed by
    24 undamped ones (the stored acceleration carried across) are bitwise the
    reference binary's 40 steps; without the switch the two runs differ."""
    with tempfile.TemporaryDirectory() as wd:
        p, frame0, final, out = _bc_switch_case(wd, "bc_switch_time=%r\n" % (17.4 * 0.1 * 2 * 4.5 / 9))
    assert "*** BC SWITCH at t=" in out
    st, a = oracle.step(p, frame0, 16)
    p.damp_rate = 0.0
    st, _ = oracle.step(p, st, 24, acc=a)
    assert np.array_equal(st, final)
    with tempfile.TemporaryDirectory() as wd:
        _, _, final_damped, out = _bc_switch_case(wd)
    assert "BC SWITCH" not in out
    assert np.abs(final_damped - final).max() > 1e-6


def test_q_conservation_eta0_periodic():
    """Noether charge is conserved to round-off at eta=0 (v66 qcons: 7e-14 relative)."""
    N, L = 16, 7.5
    from scp_amd import seeds
    st = seeds.seed_balls(N, L, 1, [dict(x0=0, y0=0, z0=0, amp=0.8, sigma=1.8, omega=1.2, phase=0.2,
                                         winding=1)], theta_sigma=0.01, seed=3)
    p = oracle.make_params(N, 1, L, 0.1, eta=0.0)
    q0 = oracle.diag(p, st)["Q_total"]
    out, _ = oracle.step(p, st, 200)
    q1 = oracle.diag(p, out)["Q_total"]
    assert abs(q1 - q0) / abs(q0) <= 1e-12
    # with eta != 0 the phi and theta charges exchange but their sum is conserved too
    p.eta = 0.5
    out, _ = oracle.step(p, st, 200)
    assert abs(oracle.diag(p, out)["Q_total"] - q0) / abs(q0) <= 1e-12


def test_real_limit_identity():
    """Complex kernel with v = tv = 0 reproduces the real kernel (v66 SPEC §9; ref 2.44e-15)."""
    N, L = 10, 4.5
    rng = np.random.default_rng(5)
    real = 0.5 * rng.standard_normal((12, N ** 3))
    cplx = np.zeros((24, N ** 3))
    cplx[:12] = real
    pr = oracle.make_params(N, 0, L, 0.1)
    pc = oracle.make_params(N, 1, L, 0.1)
    a, _ = oracle.step(pr, real, 20)
    b, _ = oracle.step(pc, cplx, 20)
    assert np.abs(b[12:]).max() == 0.0
    assert np.abs(a - b[:12]).max() <= 1e-13 * max(1.0, np.abs(a).max())


def test_plane_wave_dispersion_closed_form():
    """Free lattice field (mu=0, eta=0): u = A cos(k x) cos(n th) with
    cos th = 1 - Om^2 dt^2/2, Om^2 = m^2 + (4/dx^2) sin^2(k dx/2) — exact for
    the 7-point Laplacian and velocity Verlet. Fails on a wrong Laplacian
    index/sign, a wrong mass term, or a wrong kick/drift order."""
    N, L = 16, 7.5
    p = oracle.make_params(N, 0, L, 0.2, mu=0.0, eta=0.0, m2=1.3, mtheta2=0.7)
    dx, dt = p.dx, p.dt
    x = -L + dx * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    st = np.zeros((12, N ** 3))
    modes = [(0, 2, 0, 0.3, 1.3), (1, 3, 1, 0.2, 1.3), (3, 1, 2, 0.25, 0.7), (5, 4, 1, 0.1, 0.7)]
    for comp, m, axis, A, mass2 in modes:
        k = 2 * np.pi * m / (N * dx)
        coord = [X, Y, Z][axis]
        st[comp] = (A * np.cos(k * coord)).ravel()
    n = 37
    out, _ = oracle.step(p, st, n)
    for comp, m, axis, A, mass2 in modes:
        k = 2 * np.pi * m / (N * dx)
        om2 = mass2 + 4.0 / dx ** 2 * math.sin(k * dx / 2) ** 2
        th = math.acos(1 - om2 * dt * dt / 2)
        coord = [X, Y, Z][axis]
        expect = (A * np.cos(k * coord) * math.cos(n * th)).ravel()
        assert np.abs(out[comp] - expect).max() <= 1e-12


def test_uniform_field_reduces_to_site_ode():
    """k=0 field: every site follows the same 12-dim ODE; a scalar leapfrog of
    u'' = -m^2 u - 2 Vt'(s) u prod_{b!=a}|Phi_b|^2 (theta decoupled) reproduces it."""
    N, L = 6, 2.5
    u0 = np.array([0.7, 0.5, 0.6, 0.1, -0.2, 0.3])        # u, v
    w0 = np.array([0.1, -0.3, 0.2, 0.4, 0.0, -0.1])
    st = np.zeros((24, N ** 3))
    for a in range(3):
        st[a], st[12 + a] = u0[a], u0[3 + a]
        st[6 + a], st[18 + a] = w0[a], w0[3 + a]
    p = oracle.make_params(N, 1, L, 0.1)
    n = 50
    out, _ = oracle.step(p, st, n)

    def acc(z):
        s2 = [z[a] ** 2 + z[3 + a] ** 2 for a in range(3)]
        s = s2[0] * s2[1] * s2[2]
        vp = 0.5 * p.mu / (1 + p.kappa * s) ** 2
        rest = [s2[1] * s2[2], s2[0] * s2[2], s2[0] * s2[1]]
        return np.array([-p.m2 * z[i] - 2 * vp * z[i] * rest[i % 3] for i in range(6)])
    z, v = u0.copy(), w0.copy()
    a = acc(z)
    for _ in range(n):
        v = v + 0.5 * p.dt * a
        z = z + p.dt * v
        a = acc(z)
        v = v + 0.5 * p.dt * a
    for a_ in range(3):
        assert np.abs(out[a_] - z[a_]).max() <= 1e-12
        assert np.abs(out[12 + a_] - z[3 + a_]).max() <= 1e-12


def _rand_gauged(N, seed, link=0.5, amp=0.4):
    rng = np.random.default_rng(seed)
    st = np.zeros((30, N ** 3))
    st[:24] = amp * rng.standard_normal((24, N ** 3))
    st[24:27] = link * rng.standard_normal((3, N ** 3))
    st[27:30] = 0.1 * rng.standard_normal((3, N ** 3))
    return st


def test_gauge_transform_invariance():
    """do_test_gauge_xform (scp_sim.c:2461-2556): |acc|^2 per complex pair, E_acc, G
    unchanged to <= 1e-12 relative under a random lattice gauge transformation."""
    N = 8
    p = oracle.make_params(N, 2, 3.5, 0.1, g=0.4)
    st = _rand_gauged(N, 1)
    a0 = oracle.forces(p, st)
    a1 = oracle.forces(p, oracle.gauge_transform(p, st))
    b0 = [a0[q] ** 2 + a0[6 + q] ** 2 for q in range(3)] + [a0[3 + q] ** 2 + a0[9 + q] ** 2 for q in range(3)]
    b1 = [a1[q] ** 2 + a1[6 + q] ** 2 for q in range(3)] + [a1[3 + q] ** 2 + a1[9 + q] ** 2 for q in range(3)]
    for x, y in zip(b0 + list(a0[12:15]), b1 + list(a1[12:15])):
        assert np.abs(x - y).max() / max(np.abs(x).max(), 1.0) <= 1e-12
    # a non-covariant stencil would fail: rotating matter without moving links changes |acc|
    st2 = st.copy()
    c, s = math.cos(0.3), math.sin(0.3)
    for base_re, base_im in ((0, 12), (3, 15)):
        for a in range(3):
            u, v = st2[base_re + a].copy(), st2[base_im + a].copy()
            idx = np.arange(N ** 3) % 2 == 0       # site-dependent phase
            st2[base_re + a][idx] = c * u[idx] + s * v[idx]
            st2[base_im + a][idx] = c * v[idx] - s * u[idx]
    a2 = oracle.forces(p, st2)
    assert np.abs(a2[0] ** 2 + a2[6] ** 2 - b0[0]).max() > 1e-6


def test_staple_divergence_vanishes():
    """With matter at zero the E-kick is the staple term alone, whose backward
    lattice divergence telescopes to zero (v69 SPEC §3.4)."""
    N = 8
    p = oracle.make_params(N, 2, 3.5, 0.1, g=0.3)
    st = np.zeros((30, N ** 3))
    st[24:27] = np.random.default_rng(4).uniform(-3, 3, (3, N ** 3))
    a = oracle.forces(p, st)
    K = a[12:15].reshape(3, N, N, N)
    div = sum(K[d] - np.roll(K[d], 1, axis=d) for d in range(3))
    assert np.abs(div).max() <= 1e-12 * np.abs(K).max()
    assert np.abs(K).max() > 1e-3


def test_gauss_law_conserved_by_step():
    """After projection the residual stays at round-off for all t (v69 FINDINGS: ~1e-13)."""
    fx = fixtures.load("gauged_g03_n8")
    n, p = fixtures.nsteps(fx)
    p.G_offset = _g_offset(p, fx["frame0"])
    out, _ = oracle.step(p, fx["frame0"], 60)
    d = oracle.diag(p, out)
    assert d["gauss_max"] <= 1e-12


@pytest.mark.parametrize("name", ["gauged_n8", "gauged_g03_n8", "gauged_damp_n10"])
def test_gauss_projection_vs_reference_frame0(name):
    """init_gauss_project (C:2284-2400): the reference's frame 0 is the projected seed."""
    fx = fixtures.load(name)
    n, p = fixtures.nsteps(fx)
    z = np.load(os.path.join(fixtures.GOLDEN, name + ".npz"))
    proj, it = oracle.gauss_project(p, z["seed"])
    assert it > 0
    assert np.abs(proj - fx["frame0"]).max() <= 1e-12 * max(1.0, np.abs(fx["frame0"][27:]).max())
    import re
    m = re.search(r"Gauss projection: (\d+) CG iters", fx["stdout"])
    assert m and abs(int(m.group(1)) - it) <= 2
    d = oracle.diag(p, proj)
    assert d["gauss_max"] <= 1e-12


def test_tiny_g_reduces_to_complex():
    """g -> 1e-8 reproduces the ungauged complex evolution within 1e-9 (v69 SPEC §1.3)."""
    N = 8
    st = _rand_gauged(N, 9, link=0.0, amp=0.3)
    st[27:30] = 0.0
    pg = oracle.make_params(N, 2, 3.5, 0.1, g=1e-8)
    pc = oracle.make_params(N, 1, 3.5, 0.1)
    a, _ = oracle.step(pg, st, 20)
    b, _ = oracle.step(pc, st[:24], 20)
    assert np.abs(a[:24] - b).max() <= 1e-9


def test_g_zero_dispatches_to_complex():
    """complex_gauge with g == 0 runs the complex step bit for bit (C:1586, 1669-1670)."""
    N = 8
    st = _rand_gauged(N, 10)
    pg = oracle.make_params(N, 2, 3.5, 0.1, g=0.0)
    pc = oracle.make_params(N, 1, 3.5, 0.1)
    a, _ = oracle.step(pg, st, 10)
    b, _ = oracle.step(pc, st[:24], 10)
    assert np.array_equal(a[:24], b) and np.array_equal(a[24:], st[24:])


# ---------------------------------------------------------------------------
# Real-mode terms of the reference's GPU step only (scp_sim.cu:1144-1248): no
# executable reference here (the CUDA path), so the restatement is pinned by
# closed forms on fields whose lattice derivatives are known exactly.
# ---------------------------------------------------------------------------

def _real_state(N, th_fn, phi=(0.3, -0.2, 0.5), thv_fn=None):
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    st = np.zeros((12, N ** 3))
    for a in range(3):
        st[a] = phi[a]
    th = th_fn(i, j, k)
    for a in range(3):
        st[3 + a] = np.broadcast_to(th[a], i.shape).ravel()
    if thv_fn is not None:
        thv = thv_fn(i, j, k)
        for a in range(3):
            st[9 + a] = np.broadcast_to(thv[a], i.shape).ravel()
    return st


def _extra(N, st, **kw):
    """acceleration change the GPU-step terms add, from the oracle"""
    L = 2.5
    base = oracle.forces(oracle.make_params(N, 0, L, 0.1), st)
    return oracle.forces(oracle.make_params(N, 0, L, 0.1, **kw), st) - base


def test_gpu_terms_uniform_fields_closed_form():
    """Uniform fields: no gradients or curls, so each term reduces to its local
    form: lambda |th|^2 th, sigma_x |th|^2 phi and |phi|^2 th, sigma_c |th|^2 th,
    gamma (|th|^2 - th_sat^2) th above the threshold and nothing below."""
    N = 6
    th0 = np.array([0.4, 0.1, -0.3])
    phi = np.array([0.3, -0.2, 0.5])
    st = _real_state(N, lambda i, j, k: th0, phi=phi)
    t2, p2 = float(th0 @ th0), float(phi @ phi)
    cases = [
        (dict(lambda_self=0.7), np.zeros(3), -0.7 * t2 * th0),
        (dict(sigma_cross=0.3), -0.3 * t2 * phi, -0.3 * p2 * th0),
        (dict(sigma_cubic=0.9), np.zeros(3), -0.9 * t2 * th0),
        (dict(theta_sat=0.2, gamma_conv=1.3), np.zeros(3), -1.3 * (t2 - 0.04) * th0),
        (dict(theta_sat=0.9, gamma_conv=1.3), np.zeros(3), np.zeros(3)),      # below saturation
        (dict(theta_sat=0.2, gamma_conv=0.0), np.zeros(3), np.zeros(3)),      # both must be > 0
        (dict(sigma_grad=0.5, chi_chiral=2.0), np.zeros(3), np.zeros(3)),     # no gradient
        (dict(sigma_freq=0.5), np.zeros(3), np.zeros(3)),                     # no frequency mismatch
        (dict(lambda_self=-0.7, sigma_cubic=-1.0), np.zeros(3), np.zeros(3)), # negative = off, as there
    ]
    for kw, dphi, dth in cases:
        d = _extra(N, st, **kw)
        for a in range(3):
            assert np.abs(d[a] - dphi[a]).max() <= 1e-15, (kw, a)
            assert np.abs(d[3 + a] - dth[a]).max() <= 1e-15, (kw, a)


def test_gpu_terms_gradient_and_chirality_closed_form():
    """th = (0, A sin(q i), B): the central difference of th_1 along i is
    A sin(q) cos(q i) / dx exactly, so |grad th|^2 = g^2 with g that value, curl
    th = (0, 0, g), helicity h = B g, sigma_eff = sigma (1 + chi h / (|th|^2 +
    0.01)) clamped at 0; th_a gets -sigma_eff g^2 th_a and phi_2 +sigma_eff g^3."""
    N, A, B, sg, chi = 8, 0.35, 0.25, 0.6, 1.7
    q = 2 * np.pi / N
    st = _real_state(N, lambda i, j, k: (0.0 * i, A * np.sin(q * i), B + 0.0 * i))
    dx = 2 * 2.5 / (N - 1)
    i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")[0].ravel()
    g = A * np.sin(q) * np.cos(q * i) / dx
    th = np.stack([0.0 * i, A * np.sin(q * i), B + 0.0 * i])
    t2 = (th ** 2).sum(axis=0)
    se = np.maximum(sg * (1 + chi * B * g / (t2 + 0.01)), 0.0)
    d = _extra(N, st, sigma_grad=sg, chi_chiral=chi)
    for a in range(3):
        assert np.abs(d[3 + a] + se * g * g * th[a]).max() <= 1e-13, a
    assert np.abs(d[0]).max() == 0 and np.abs(d[1]).max() == 0
    assert np.abs(d[2] - se * g * g * g).max() <= 1e-13
    # the curl couples: a cubic term on the same field sources phi_2 with sigma_c |th|^2 g
    d = _extra(N, st, sigma_cubic=0.8)
    assert np.abs(d[2] - 0.8 * t2 * g).max() <= 1e-13


def test_gpu_terms_frequency_mismatch_closed_form():
    """Uniform th of modulus 0.3, theta velocity 0.5 along th at one site only:
    omega there is 0.5 / 0.31 and 0 elsewhere, so that site sees the mean
    mismatch omega0 and each of its six neighbours omega0 / 6; no curl, so phi
    is untouched."""
    N = 6
    th0 = np.array([0.3, 0.0, 0.0])
    s0 = (2, 3, 1)

    def thv(i, j, k):
        m = ((i == s0[0]) & (j == s0[1]) & (k == s0[2])).astype(float)
        return (0.5 * m, 0.0 * m, 0.0 * m)
    st = _real_state(N, lambda i, j, k: th0, thv_fn=thv)
    d = _extra(N, st, sigma_freq=0.4)
    om = 0.5 / (0.3 + 0.01)
    expect = np.zeros((N, N, N))
    expect[s0] = om
    for dd in range(3):
        for sgn in (1, -1):
            n = list(s0)
            n[dd] = (n[dd] + sgn) % N
            expect[tuple(n)] = om / 6
    assert np.abs(d[3] + 0.4 * expect.ravel() * th0[0]).max() <= 1e-15
    assert np.abs(d[:3]).max() == 0 and np.abs(d[4:6]).max() == 0


# ---------------------------------------------------------------------------
# vector frames (real mode, FRVD)
# ---------------------------------------------------------------------------

VEC_FIXTURES = ["real_vec_n16", "real_vec_n12_bs5"]


def _vec_fixture(name):
    import json
    z = np.load(os.path.join(fixtures.GOLDEN, name + ".npz"))
    return z, json.loads(str(z["params"]))


@pytest.mark.parametrize("name", VEC_FIXTURES)
def test_vec_fit_oracle_vs_reference_iframe(name):
    """fit_patch restated (normal equations, Cholesky) on the reference's own f32
    fields reproduces the reference's first I-frame: bitwise but for rare last-bit
    differences of near-zero coefficients."""
    z, p = _vec_fixture(name)
    N, bs = int(p["N"]), int(p["vec_block_size"])
    mine = oracle.vec_fit(z["fields0"], N, bs).reshape(-1, 384)
    ref = z["v0_coeffs"]
    assert np.array_equal(z["v0_origins"][:, 0] % bs, np.zeros(len(ref)))
    exact = np.mean(mine == ref)
    err = np.abs(mine.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-30 + np.abs(ref).max() * 1e-6)
    print(f"{name}: {exact:.5f} of the coefficients bitwise, largest scaled difference {err.max():.2e}")
    assert exact >= 0.999
    assert err.max() <= 2e-7


def test_vec_fit_oracle_closed_form():
    """A field that is a tricubic Chebyshev polynomial on each patch is fitted exactly
    (to FP32 rounding of the samples); the coefficient order is a*16 + b*4 + c with a
    along x (the slowest index)."""
    rng = np.random.default_rng(4)
    for bs in (4, 5, 8):
        N = 2 * bs
        A = oracle.vec_basis(bs)
        truth = rng.standard_normal((8, 64))
        field = np.zeros((N, N, N))
        p = 0
        for bi in range(2):
            for bj in range(2):
                for bk in range(2):
                    field[bi * bs:(bi + 1) * bs, bj * bs:(bj + 1) * bs, bk * bs:(bk + 1) * bs] = \
                        (A @ truth[p]).reshape(bs, bs, bs)
                    p += 1
        f32 = field.astype(np.float32).ravel()
        got = oracle.vec_fit(np.stack([f32] * 6), N, bs)
        assert np.abs(got[:, 0, :] - truth).max() <= 1e-5 * np.abs(truth).max()
        # T_1(x) alone (a linear ramp along i inside a patch) lands in coefficient 16
        ramp = np.zeros((N, N, N), dtype=np.float32)
        s = 2.0 * (np.arange(bs) / (bs - 1)) - 1.0
        ramp[:bs, :bs, :bs] = s[:, None, None]
        c = oracle.vec_fit(np.stack([ramp.ravel()] * 6), N, bs)[0, 0]
        assert abs(c[16] - 1.0) <= 1e-6 and np.abs(np.delete(c, 16)).max() <= 1e-6


@pytest.mark.parametrize("name", VEC_FIXTURES)
def test_vector_frames_host_side_vs_reference(tmp_path, name):
    """The product's host side of the vector frames (scp_vf_write: temporal model,
    I/P choice, FRVD payloads, JMPF types), fed with the oracle's states (the real
    mode is bitwise the reference) and the oracle's fits, reproduces every vector
    frame the reference binary wrote."""
    z, p = _vec_fixture(name)
    N, L, bs = int(p["N"]), float(p["L"]), int(p["vec_block_size"])
    iint, tol, T = int(p["vec_iframe_interval"]), float(p["vec_delta_tol"]), float(p["T"])
    dx = 2 * L / (N - 1)
    dt = 0.1 * dx
    every = max(1, round(float(p["vec_snap_dt"]) / dt))
    n_steps = round(T / dt)
    op = oracle.make_params(N, 0, L, 0.1)
    st = z["frame0"].astype(np.float64)
    frames, acc = [(0.0, st)], None
    for step in range(every, n_steps + 1, every):
        st, acc = oracle.step(op, st, every, acc=acc)
        f32 = st[[0, 1, 2, 3, 4, 5]].astype(np.float32)
        frames.append(("vec", step * dt, oracle.vec_fit(f32, N, bs).reshape(-1, 384)))
    out = tmp_path / "v.sfa"
    sfa.write_sim_sfa(out, N, L, dt, frames, vec=(bs, iint, tol))
    _compare_vec_frames(out, z, first_frame=1)


def _compare_vec_frames(path, z, first_frame):
    """Every vector frame of `path` against the reference's (fixture z)."""
    nv = int(z["n_vec"])
    with sfa.SFAFile(path) as f:
        kinds = [f.frame_type(i) for i in range(f.total_frames)]
        vec_idx = [i for i, t in enumerate(kinds) if t in (1, 2)]
        assert len(vec_idx) == nv
        ref_kinds = [int(t) for t in z["types"]]
        assert [k for k in kinds if k] == [k for k in ref_kinds if k]
        for k, i in enumerate(vec_idx):
            v = f.read_vec(i)
            pre = f"v{k}_"
            assert v["type"] == int(z[pre + "type"])
            assert abs(v["time"] - float(z[pre + "time"])) <= 1e-12
            if v["type"] == 1:
                assert np.array_equal(v["origins"], z[pre + "origins"])
                assert v["omega"] == float(z[pre + "omega"])
                ref = z[pre + "coeffs"]
                scale = np.abs(ref).max()
                assert np.abs(v["coeffs"] - ref).max() <= 2e-7 * scale
                for key in ("mean", "amp"):
                    assert np.abs(v[key] - z[pre + key]).max() <= 1e-6 * scale, key
                # the model's prediction (phase is ill-defined where the amplitude vanishes)
                for t in (v["time"], v["time"] + 1.0):
                    pm = v["mean"] + v["amp"] * np.cos(v["omega"] * t + v["phase"])
                    pr = z[pre + "mean"] + z[pre + "amp"] * np.cos(v["omega"] * t + z[pre + "phase"])
                    assert np.abs(pm - pr).max() <= 1e-5 * scale
            else:
                assert np.array_equal(v["indices"], z[pre + "indices"]), (k, v["indices"], z[pre + "indices"])
                if len(v["indices"]):
                    d = z[pre + "deltas"]
                    assert np.abs(v["deltas"] - d).max() <= 1e-5 * max(np.abs(d).max(), 1e-3)


# ---------------------------------------------------------------------------
# self-tune restatement (scp_sim.cu:5475-5547)
# ---------------------------------------------------------------------------

def _st_state(N, L, seed):
    rng = np.random.default_rng(seed)
    dx = 2 * L / (N - 1)
    x = -L + dx * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    st = 0.05 * rng.standard_normal((12, N ** 3))
    st[0] += 0.8 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / 4.0).ravel()
    return st


@pytest.mark.parametrize("ptarget", [0.0, 0.01])
def test_self_tune_restatement_closed_forms(ptarget):
    """The charge after a projection is the target Q0 (the first update's Q), and
    kappa moves by the stated factor at every update."""
    N = 12
    L = 0.25 * (N - 1)
    p = oracle.make_params(N, 0, L, 0.1)
    k0, every, pcrit, eps, gamma, gain = p.kappa, 6, 0.003, 0.04, 0.12, 0.10
    fin, ev = oracle.self_tune_run(p, _st_state(N, L, 7), 36, every, p.m2, pcrit=pcrit, ptarget=ptarget)
    assert len(ev) == 6
    kappa = k0
    for (t, k, Pmax, Q, tag) in ev:
        if ptarget > 0:
            err = max(-1.0, min(1.0, (Pmax - ptarget) / ptarget))
            assert tag == err
            kappa *= 1.0 + gain * err
        else:
            assert tag == ("COLLAPSE" if Pmax > pcrit else "stable")
            kappa *= (1.0 + gamma) if Pmax > pcrit else (1.0 - eps)
        assert k == max(kappa, 1e-3)
    if ptarget <= 0:
        assert {e[4] for e in ev} == {"COLLAPSE", "stable"}
        d = oracle.diag(p, fin)
        assert abs(2 * d["E_mass"] / p.m2 - ev[0][3]) <= 1e-12 * ev[0][3]


QFLOW_REF = os.path.join(os.path.dirname(REF_BIN), "eta_qflow")


@pytest.mark.skipif(not os.path.exists(QFLOW_REF), reason="reference eta_qflow not built here")
@pytest.mark.parametrize("case", ["ball_bvp", "ring_wind1", "ball_zero_theta"])
def test_oracle_qflow_vs_reference_program(case):
    """The fixed-charge flow: this build's host seed (libscp_io scp_qflow_seed) and
    the oracle's flow against sfa/seed/eta_qflow.c built from source, run on one
    thread (its reductions then run in site order, as the oracle's): every printed
    row (omega, E_Q, maxF, s_max, N) is the same text, and the f32 output seed is
    bitwise the oracle state's columns."""
    from scp_amd import seeds
    from tests import qflow_cases as qc
    N, L, om0, iters, pev = 20, 5.0, 1.35, 90, 30
    kw = dict(ring_wind1=dict(wind=1, wtor=2.0), ball_zero_theta=dict(thinit="zero")).get(case, {})
    with tempfile.TemporaryDirectory() as wd:
        prof = os.path.join(wd, "prof.txt")
        qc.write_profile(prof)
        out = os.path.join(wd, "seed.sfa")
        text, rows = qc.run(qc.argv(QFLOW_REF, N, L, prof, om0, out, iters, pev, **kw))
        with sfa.SFAFile(out) as f:
            ref_cols = f.read_frame(0)
        st, gmax = seeds.qflow_seed(N, L, prof, om0, qc.FLOW["eta"], qc.FLOW["mth"], **kw)
    p = oracle.qflow_params(N, L, **qc.FLOW)
    n0 = oracle.qflow_norm(p, st)
    assert f"N_init={n0:.4f}  Q={om0 * n0:.4f}" in text and f"g_max={gmax:.4e}" in text
    dx = 2 * L / (N - 1)
    fin, orows, om_last = oracle.qflow(p, st, om0 * n0, 0.04 * dx * dx, iters, pev)
    assert [r[0] for r in orows] == [0, 30, 60, 90]
    assert [list(r) for r in rows] == [qc.fmt_row(r) for r in orows]
    assert np.array_equal(ref_cols.astype(np.float32), qc.output_columns(fin, om_last))
    assert np.abs(fin - st).max() > 1e-6          # the flow moved the seed


@pytest.mark.skipif(not os.path.exists(QFLOW_REF), reason="reference eta_qflow not built here")
@pytest.mark.parametrize("case", ["ring_wind1_J", "ball_J", "ring_wind1_Jcap"])
def test_oracle_qflow_fixed_q_j_vs_reference_program(case):
    """The fixed-(Q, J) mode (fixJ = 1): the oracle's restatement against
    sfa/seed/eta_qflow.c built from source, one thread: every printed row (omega,
    Omega, E_QJ, maxF, s_max, N, J_act, cond) is the same text, the final lines
    match, and the f32 output seed (velocities with the -Omega wm Jhat terms) is
    bitwise the oracle's. Cases: a winding-1 ring at its default target J = Q, a
    ball spun up to J = Q/20 (Omega ~ -0.1 from the start, v and Theta_v grow from
    zero), and a ring pushed to J = 3Q, where |Omega| sits at its cap."""
    from scp_amd import seeds
    from tests import qflow_cases as qc
    N, L, om0, iters, pev = 20, 5.0, 1.35, 60, 20
    kw = dict(ring_wind1_J=dict(wind=1, wtor=2.0), ring_wind1_Jcap=dict(wind=1, wtor=2.0),
              ball_J={})[case]
    jarg = 0.0
    with tempfile.TemporaryDirectory() as wd:
        prof = os.path.join(wd, "prof.txt")
        qc.write_profile(prof)
        st, gmax = seeds.qflow_seed(N, L, prof, om0, qc.FLOW["eta"], qc.FLOW["mth"], **kw)
        p = oracle.qflow_params(N, L, **qc.FLOW)
        n0 = oracle.qflow_norm(p, st)
        Q = om0 * n0
        jarg = {"ring_wind1_Jcap": 3.0 * Q, "ball_J": 0.05 * Q}.get(case, 0.0)
        out = os.path.join(wd, "seed.sfa")
        text, rows = qc.run(qc.argv(QFLOW_REF, N, L, prof, om0, out, iters, pev, fixJ=1, Jtarget=jarg, **kw))
        with sfa.SFAFile(out) as f:
            ref_cols = f.read_frame(0)
    Jt = (kw.get("wind", 0) * Q) if jarg == 0.0 else jarg
    assert f"# fixJ: J_target={Jt:.4f}" in text
    dx = 2 * L / (N - 1)
    fin, orows, last, jh = oracle.qflow_j(p, st, Q, Jt, 0.02 * dx * dx, iters, pev)
    assert [r[0] for r in orows] == [0, 20, 40, 60]
    assert [list(r) for r in rows] == [qc.fmt_row_j(r) for r in orows]
    om, Om, cond, ell, K = last
    if case == "ring_wind1_Jcap":
        assert abs(Om) == 0.15
    if case == "ball_J":
        assert 0.05 < -orows[0][2] < 0.15 and np.abs(fin[3:6]).max() > 1e-6   # Omega acts; v grows from 0
    nphi, nth = [float(np.sum(fin[s] ** 2) * dx ** 3) for s in (slice(0, 6), slice(6, 12))]
    fixj_final = [l for l in text.splitlines() if l.startswith("# fixJ final:")]
    assert len(fixj_final) == 1 and f"Omega={Om:.6f}" in fixj_final[0] and f"cond={cond:.3e}" in fixj_final[0]
    assert f"J_actual={om * ell - Om * K:.4f}" in fixj_final[0]
    cols = qc.output_columns_j(fin, om, Om, jh, oracle.qflow_wmask(p))
    assert np.array_equal(ref_cols.astype(np.float32), cols)
    assert np.abs(fin - st).max() > 1e-6
