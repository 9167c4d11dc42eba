"""FP32 compensated step vs the FP64 step (which is itself pinned to the reference).

Stated bars (every FP32C result carries its error against FP64):
  field L2 (relative)            <= 2e-5   after 200 steps of a gauged 32^3 ball
  energy, |dQ|/Q (relative)      <= 1e-6   same run, and over 10^4 steps of the
                                           128^3 two-ball run and of a 48^3 ball
  Gauss residual                 <= 1e-4   absolute (the FP32 current conserves
                                           Gauss's law only to FP32 rounding; the
                                           FP64 reduction reports it): measured
                                           2.1-3.1e-5 on the 32^3 ball (g = 0.1,
                                           links N(0, 0.05^2)), 7.7e-6-1.1e-5 on the
                                           128^3 pair and 7.3e-6-9.8e-6 on the 48^3
                                           ball after 10^4 steps, so the bar is 3x
                                           and 10x those; the 512^3 run is held to
                                           1e-6 (measured 8.3e-8) in
                                           tests/test_gpu_baseline.py
and the compensation words must not do worse than plain FP32 over a long run.
"""
import numpy as np
import pytest

from scp_amd import grid as G
from scp_amd import seeds

pytestmark = pytest.mark.gpu


def pair_state(N, L, link=0.05, seed=21):
    dx = 2 * L / (N - 1)
    balls = [dict(x0=-0.25 * L, y0=0.05 * L, z0=0.0, amp=0.8, sigma=3 * dx, omega=0.8, phase=0.3, winding=1),
             dict(x0=0.25 * L, y0=-0.05 * L, z0=0.1 * L, amp=0.8, sigma=3 * dx, omega=-0.8, phase=1.3,
                  winding=-1)]
    return seeds.seed_balls(N, L, 2, balls, theta_sigma=0.01, link_sigma=link, seed=seed)


def run(params, st, steps, project=True):
    with G.Grid(params) as g:
        g.upload(st)
        if project:
            g.gauss_project()
        d0 = g.diag()
        g.step(steps)
        return g.download(), d0, g.diag()


def ball_state(N, L, link=0.05, seed=22):
    dx = 2 * L / (N - 1)
    balls = [dict(x0=0.1 * L, y0=0.05 * L, z0=-0.1 * L, amp=0.8, sigma=3 * dx, omega=0.8, phase=0.3, winding=1)]
    return seeds.seed_balls(N, L, 2, balls, theta_sigma=0.01, link_sigma=link, seed=seed)


@pytest.mark.parametrize("lo", [0, 1, 3, 4])
def test_f32c_gauged_error_vs_f64(lo):
    N, L, steps = 32, 7.75, 200
    st = ball_state(N, L)
    p64 = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.1)
    p32 = G.make_params(N, G.MODE_GAUGED, L, 0.2, precision=G.PREC_F32C, lo_words=lo, g=0.1)
    a, d0, d64 = run(p64, st, steps)
    b, e0, d32 = run(p32, st, steps)
    field = np.linalg.norm(b - a) / np.linalg.norm(a)
    dE = abs(d32["E_total"] - d64["E_total"]) / abs(d64["E_total"])
    dQ = abs(d32["Q_total"] - d64["Q_total"]) / abs(d64["Q_total"])
    print(f"lo={lo}: field L2 {field:.3e} dE/E {dE:.3e} dQ/|Q| {dQ:.3e} gauss {d32['gauss_max']:.3e}")
    assert field <= 2e-5
    assert dE <= 1e-6
    assert dQ <= 1e-6
    assert d32["gauss_max"] <= 1e-4


def test_compensation_reduces_long_run_drift():
    """Over a long run the compensated words must not do worse than plain FP32
    on the conserved totals (both measured against FP64)."""
    N, L, steps = 16, 3.75, 1500
    st = pair_state(N, L, link=0.02)
    p64 = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.1)
    a, _, d64 = run(p64, st, steps)
    err = {}
    for lo in (0, 1, 3, 4, 2):
        p = G.make_params(N, G.MODE_GAUGED, L, 0.2, precision=G.PREC_F32C, lo_words=lo, g=0.1)
        b, _, d = run(p, st, steps)
        err[lo] = (np.linalg.norm(b - a) / np.linalg.norm(a), abs(d["E_total"] - d64["E_total"]))
    print(f"compensated FP32 {err[0]}, bf16 {err[1]}, byte {err[3]}, byte on positions {err[4]}; plain {err[2]}")
    for lo in (0, 1, 3, 4):
        assert err[lo][0] <= err[2][0] and err[lo][1] <= 1e-6 * abs(d64["E_total"]), lo


def test_f32c_complex_and_real_modes_run():
    for mode, N in ((1, 20), (0, 20)):
        L = 0.25 * (N - 1)
        st = pair_state(N, L)[: {1: 24, 0: 12}[mode]]
        p64 = G.make_params(N, mode, L, 0.2)
        p32 = G.make_params(N, mode, L, 0.2, precision=G.PREC_F32C)
        a, _, d64 = run(p64, st, 50, project=False)
        b, _, d32 = run(p32, st, 50, project=False)
        assert np.linalg.norm(b - a) / np.linalg.norm(a) <= 2e-5
        assert abs(d32["E_total"] - d64["E_total"]) <= 1e-6 * abs(d64["E_total"])


@pytest.mark.parametrize("lo", [1, 3, 4])
def test_f32c_march_and_direct_kernels_agree(lo):
    """The LDS-staged march and the direct-neighbour kernel decode and encode
    the compensation words alike: after 20 steps of a gauged 24^3 ball both
    stay within the FP32C bar of FP64 and of each other."""
    N = 24
    L = 0.25 * (N - 1)
    st = ball_state(N, L)
    a, _, d64 = run(G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.1), st, 20, project=False)
    outs = []
    for sk in (0, 1):
        p = G.make_params(N, G.MODE_GAUGED, L, 0.2, precision=G.PREC_F32C, lo_words=lo, g=0.1, step_kernel=sk)
        b, _, d = run(p, st, 20, project=False)
        outs.append(b)
        assert np.linalg.norm(b - a) / np.linalg.norm(a) <= 2e-5, sk
        assert abs(d["E_total"] - d64["E_total"]) <= 1e-6 * abs(d64["E_total"]), sk
    assert np.linalg.norm(outs[0] - outs[1]) / np.linalg.norm(a) <= 2e-5


# real-mode terms of the pass-1 instance (MODE 3: alpha_cs, beta_h, sigma_freq read from an
# auxiliary array) together with the register-only terms
REAL_PASS1_ALL = dict(alpha_cs=0.1, beta_h=0.2, kappa_h=0.3, sigma_freq=0.3, theta_sat=0.05, gamma_conv=0.8,
                      sigma_grad=0.4, chi_chiral=1.5)


@pytest.mark.parametrize("case", ["base", "pass1_all"])
@pytest.mark.parametrize("lo", [1, 4])
def test_f32c_real_march_and_direct_kernels_agree(lo, case):
    """Real mode, FP32C: the march and the direct-neighbour kernel after 20
    steps of a 24^3 state stay within the FP32C bar of FP64 and of each other;
    with the pass-1 fields the march runs its MODE 3 instance."""
    N = 24
    L = 0.25 * (N - 1)
    st = ball_state(N, L)[:12].copy()
    kw = REAL_PASS1_ALL if case == "pass1_all" else {}
    print(f"case real f32c lo={lo} {case}", flush=True)
    a, _, d64 = run(G.make_params(N, G.MODE_REAL, L, 0.2, **kw), st, 20, project=False)
    outs = []
    for sk in (0, 1):
        p = G.make_params(N, G.MODE_REAL, L, 0.2, precision=G.PREC_F32C, lo_words=lo, step_kernel=sk, **kw)
        with G.Grid(p) as g:
            want = f"k_march<scp::S32<{lo}>, 1, {3 if kw else 0}, false>" if sk == 0 else "k_step32<0"
            assert g.step_kernel_name.startswith(want), g.step_kernel_name
        b, _, d = run(p, st, 20, project=False)
        outs.append(b)
        assert np.linalg.norm(b - a) / np.linalg.norm(a) <= 2e-5, sk
        assert abs(d["E_total"] - d64["E_total"]) <= 1e-6 * abs(d64["E_total"]), sk
    assert np.linalg.norm(outs[0] - outs[1]) / np.linalg.norm(a) <= 2e-5


def q_halves(st, N):
    """Noether charge of Phi in the x < 0 and x >= 0 halves (FP64, host)."""
    q = sum(st[a] * st[18 + a] - st[12 + a] * st[6 + a] for a in range(3)).reshape(N, N, N)
    return q[: N // 2].sum(), q[N // 2:].sum()


@pytest.mark.parametrize("lo", [1, 0, 3, 4, 2])
def test_f32c_ten_thousand_steps_two_balls(lo):
    """128^3 gauged, two opposite Q-balls 24 cells apart (omega = +-0.8), links
    N(0, 1e-3), E = 0, Gauss-projected; 10^4 steps FP32C vs FP64. The charge
    difference is relative to one ball's |Q| (the net charge is near zero).
    Bars: 1e-6 for the production words (int8 on the positions); the other
    storage kinds, kept for comparison, measure up to 4.4e-6 and are held to
    5e-6."""
    N, steps = 128, 10000
    L = 0.25 * (N - 1)
    balls = [dict(x0=-6.0, y0=0.0, z0=0.0, amp=0.8, sigma=3.0, omega=0.8, phase=0.0, winding=1),
             dict(x0=6.0, y0=0.0, z0=0.0, amp=0.8, sigma=3.0, omega=-0.8, phase=0.0, winding=-1)]
    st = seeds.seed_balls(N, L, 2, balls, theta_sigma=0.01, link_sigma=1e-3, seed=7)
    p64 = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05)
    p32 = G.make_params(N, G.MODE_GAUGED, L, 0.2, precision=G.PREC_F32C, lo_words=lo, g=0.05)
    a, _, d64 = run(p64, st, steps)
    b, _, d32 = run(p32, st, steps)
    dx = 2 * L / (N - 1)
    qs = dx ** 3 * sum(abs(q) for q in q_halves(a, N))   # charge scale: |Q| of each ball (density x dV)
    field = np.linalg.norm(b - a) / np.linalg.norm(a)
    dE = abs(d32["E_total"] - d64["E_total"]) / abs(d64["E_total"])
    dQ = abs(d32["Q_total"] - d64["Q_total"]) / qs
    print(f"lo={lo}: field L2 {field:.3e} dE/E {dE:.3e} |dQ|/|Q_ball| {dQ:.3e} gauss {d32['gauss_max']:.3e}")
    assert dE <= 1e-6
    assert dQ <= (1e-6 if lo == 4 else 5e-6)
    assert d32["gauss_max"] <= 1e-4
    assert field <= 1e-4


@pytest.mark.parametrize("lo", [1, 3, 4, 2])
def test_f32c_ten_thousand_steps_charged_ball(lo):
    """One charged ball (net Q != 0, jellium offset from the projection), 48^3,
    10^4 steps: |dQ|/Q and dE/E against FP64."""
    N, steps = 48, 10000
    L = 0.25 * (N - 1)
    st = ball_state(N, L, link=1e-3)
    p64 = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05)
    p32 = G.make_params(N, G.MODE_GAUGED, L, 0.2, precision=G.PREC_F32C, lo_words=lo, g=0.05)
    a, e0, d64 = run(p64, st, steps)
    b, _, d32 = run(p32, st, steps)
    dE = abs(d32["E_total"] - d64["E_total"]) / abs(d64["E_total"])
    dQ = abs(d32["Q_total"] - d64["Q_total"]) / abs(d64["Q_total"])
    print(f"charged ball lo={lo}: Q {d64['Q_total']:.6e} dE/E {dE:.3e} |dQ|/Q {dQ:.3e} gauss {d32['gauss_max']:.3e}")
    assert abs(d64["Q_total"]) > 1.0
    assert dE <= 1e-6
    assert dQ <= 1e-6
    assert d32["gauss_max"] <= 1e-4
