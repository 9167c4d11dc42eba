"""HIP step vs the reference (golden fixtures) and vs the oracle, through the C-ABI.

Bar (FP64 path): bitwise for the real and complex modes (no transcendentals,
same expression trees, no FMA contraction); for the gauged mode the only
difference is device sin/cos vs glibc (ulp level), so the bar there is
<= 1e-12 relative field L2 after N steps.
"""
import numpy as np
import pytest

from oracle import core as oracle
from scp_amd import grid as G
from tests import fixtures

pytestmark = pytest.mark.gpu

FIXTURES = fixtures.names()


def gpu_params(fx, **over):
    kw = dict(fx["kw"])
    kw.update(over)
    return G.make_params(fx["N"], fx["mode"], fx["L"], fx["dt_factor"], **kw)


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.parametrize("name", FIXTURES)
def test_step_matches_reference_fixture(name):
    fx = fixtures.load(name)
    n, _ = fixtures.nsteps(fx)
    print(f"case {name} N={fx['N']} mode={fx['mode']} steps={n}", flush=True)
    with G.Grid(gpu_params(fx)) as g:
        g.upload(fx["frame0"])
        g.step(n)
        out = g.download()
    if fx["mode"] == 2:
        assert rel_l2(out, fx["final"]) <= 1e-12
    else:
        assert np.array_equal(out, fx["final"]), f"max diff {np.abs(out - fx['final']).max():.3e}"


@pytest.mark.parametrize("name", FIXTURES)
def test_forces_match_oracle(name):
    fx = fixtures.load(name)
    n, p = fixtures.nsteps(fx)
    with G.Grid(gpu_params(fx)) as g:
        g.upload(fx["frame0"])
        acc = g.forces()
    ref = oracle.forces(p, fx["frame0"])
    if fx["mode"] == 2:
        assert np.abs(acc - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    else:
        assert np.array_equal(acc, ref)


@pytest.mark.parametrize("prec,lo", [(G.PREC_F64, 0), (G.PREC_F32C, 1), (G.PREC_F32C, 4)])
def test_split_steps_equal_one_call(prec, lo):
    """Emitting on the last launch of a call must not perturb the state (FP64
    and FP32C: the kicks a step takes do not depend on whether it emits)."""
    fx = fixtures.load("gauged_n8")
    with G.Grid(gpu_params(fx, precision=prec, lo_words=lo)) as a, \
            G.Grid(gpu_params(fx, precision=prec, lo_words=lo)) as b:
        a.upload(fx["frame0"])
        b.upload(fx["frame0"])
        a.step(7)
        a.step(13)
        b.step(20)
        assert np.array_equal(a.download(), b.download())
        assert a.steps_done == 20


def seeded(N, mode, L, seed=7, link_sigma=0.0):
    from scp_amd import seeds
    dx = 2 * L / (N - 1)
    balls = [dict(x0=-0.2 * L, y0=0.1 * L, z0=0.0, amp=0.8, sigma=3 * dx, omega=0.8, phase=0.3, winding=1)]
    if mode == 2:
        balls.append(dict(x0=0.25 * L, y0=-0.1 * L, z0=0.05 * L, amp=0.8, sigma=3 * dx, omega=-0.8,
                          phase=1.1, winding=-1))
    return seeds.seed_balls(N, L, mode, balls, theta_sigma=0.01, link_sigma=link_sigma, seed=seed)


@pytest.mark.parametrize("mode,N,margin,steps", [(0, 37, 3, 15), (1, 37, 3, 15), (1, 20, 0, 10), (0, 4, 1, 6),
                                                 (1, 70, 5, 8)])
def test_gradient_bc_matches_oracle(mode, N, margin, steps):
    """bc 1 (pinned x-slabs, extrapolated y edges) against the oracle restatement
    of apply_gradient_bc, bitwise; ragged tiles (N = 37, 70), no slabs (margin 0),
    the smallest grid (N = 4: the edge source rows touch each other)."""
    L = 0.25 * (N - 1)
    st = seeded(N, max(mode, 1), L)
    st = st[:12] if mode == 0 else st[:24]
    print(f"case bc1 mode={mode} N={N} margin={margin} steps={steps}", flush=True)
    p_o = oracle.make_params(N, mode, L, 0.1, bc_type=1, gradient_margin=margin)
    ref, _ = oracle.step(p_o, st, steps)
    with G.Grid(G.make_params(N, mode, L, 0.1, bc_type=1, gradient_margin=margin)) as g:
        g.upload(st)
        g.step(steps - 3)
        g.step(3)
        out = g.download()
        d = g.step_diag(1)
    ref1, _ = oracle.step(p_o, st, steps + 1)
    assert np.array_equal(out, ref.reshape(out.shape)), f"max diff {np.abs(out - ref.reshape(out.shape)).max():.3e}"
    dref = oracle.diag(p_o, ref1)
    for k in ("E_total", "E_grad", "Q_total", "phi_max"):
        assert abs(d[k] - dref[k]) <= 1e-12 * max(1.0, abs(dref[k])), k


@pytest.mark.parametrize("mode,N,steps", [(1, 24, 40), (2, 24, 40), (0, 20, 40), (1, 37, 12), (2, 33, 12)])
def test_step_matches_oracle_larger(mode, N, steps):
    """Larger and odd-sized (ragged tile) grids vs the oracle restatement."""
    L = 0.25 * (N - 1)
    st = seeded(N, max(mode, 1), L, link_sigma=0.05)
    if mode == 0:
        st = np.concatenate([st[:12]])
    print(f"case mode={mode} N={N} steps={steps}", flush=True)
    kw = dict(g=0.1) if mode == 2 else {}
    p_o = oracle.make_params(N, mode, L, 0.1, **kw)
    ref, _ = oracle.step(p_o, st, steps)
    with G.Grid(G.make_params(N, mode, L, 0.1, **kw)) as g:
        g.upload(st)
        g.step(steps)
        out = g.download()
    if mode == 2:
        assert rel_l2(out, ref) <= 1e-12
    else:
        assert np.array_equal(out, ref)


@pytest.mark.parametrize("name", ["complex_n8", "complex_damp_n10", "gauged_n8", "gauged_damp_n10",
                                  "real_n8", "real_terms_n8", "real_mode1_n8", "real_gradient_n10",
                                  "complex_gradient_n10"])
def test_diag_matches_oracle(name):
    fx = fixtures.load(name)
    n, p = fixtures.nsteps(fx)
    with G.Grid(gpu_params(fx)) as g:
        g.upload(fx["frame0"])
        g.step(5)
        d = g.diag()
        st = g.download()
    ref = oracle.diag(p, st)
    for k in G.DIAG_NAMES:
        if k in ("gauss_max", "gauss_l2"):      # rounding-noise level quantities (~1e-14)
            assert abs(d[k] - ref[k]) <= 1e-12, k
            continue
        tol = 1e-12 * max(1.0, abs(ref[k]))
        assert abs(d[k] - ref[k]) <= tol, f"{k}: gpu {d[k]!r} oracle {ref[k]!r}"


@pytest.mark.parametrize("name", ["complex_n8", "gauged_n8", "real_n8", "gauged_damp_n10"])
def test_diag_rows_match_reference_tsv(name):
    """Every diag.tsv row of the reference run, to its printed precision."""
    fx = fixtures.load(name)
    n, p = fixtures.nsteps(fx)
    head, rows = fixtures.parse_diag(fx["diag"])
    every = int(round(fx["diag_dt"] / p.dt))
    with G.Grid(gpu_params(fx)) as g:
        if fx["mode"] == 2:
            g.g_offset = _g_offset(p, fx["frame0"])
        g.upload(fx["frame0"])
        for r, row in enumerate(rows):
            if r > 0:
                g.step(every)
            d = g.diag()
            d["Q_total"] = d["Q_phi"] + d["Q_theta"]
            for col, val in zip(head[1:], row[1:]):
                if col in ("gauss_max", "gauss_l2"):
                    assert d[col] <= 1e-12 and val <= 1e-12
                    continue
                tol = 5e-7 * abs(val) + 1e-14
                if col in ("Q_phi", "Q_theta", "Q_total", "omega_core", "Q_core", "E_em", "Q_flux",
                           "Q_p0", "Q_p1", "Q_p2"):
                    tol = 5e-12 * abs(val) + 1e-12
                assert abs(d[col] - val) <= tol, f"row {r} {col}: gpu {d[col]!r} ref {val!r}"


def _g_offset(p, st):
    """mean Gauss residual of a state (the frozen offset the reference measures after projection)"""
    N = p.N
    s = st.reshape(30, N, N, N)
    E = s[27:30]
    divE = sum(E[d] - np.roll(E[d], 1, axis=d) for d in range(3)) / p.dx
    rho = 0
    for a in range(3):
        rho = rho + s[a] * s[18 + a] - s[12 + a] * s[6 + a] + s[3 + a] * s[21 + a] - s[15 + a] * s[9 + a]
    return float(np.mean(divE - p.g * rho))


def test_gauss_projection_matches_oracle():
    N, L = 16, 7.5
    st = seeded(N, 2, L, link_sigma=0.02)
    st[27:30] = 0.0
    p_o = oracle.make_params(N, 2, L, 0.1, g=0.2)
    ref, it_ref = oracle.gauss_project(p_o, st)
    with G.Grid(G.make_params(N, 2, L, 0.1, g=0.2)) as g:
        g.upload(st)
        it, gmax = g.gauss_project()
        out = g.download()
        off = g.g_offset
    assert it_ref > 0 and abs(it - it_ref) <= 3
    assert gmax <= 1e-12
    assert np.abs(out[27:30] - ref[27:30]).max() <= 1e-11 * max(1.0, np.abs(ref[27:30]).max())
    assert abs(off - p_o.G_offset) <= 1e-12


def test_gauge_transform_invariance():
    """do_test_gauge_xform (scp_sim.c:2461-2556) on the GPU forces."""
    N, L = 12, 5.5
    st = seeded(N, 2, L, link_sigma=0.3)
    p_o = oracle.make_params(N, 2, L, 0.1, g=0.3)
    st, _ = oracle.gauss_project(p_o, st)
    params = G.make_params(N, 2, L, 0.1, g=0.3)
    with G.Grid(params) as g:
        g.upload(st)
        a0 = g.forces()
        g.upload(oracle.gauge_transform(p_o, st))
        a1 = g.forces()

    def bil(a):
        out = [a[q] ** 2 + a[6 + q] ** 2 for q in range(3)] + [a[3 + q] ** 2 + a[9 + q] ** 2 for q in range(3)]
        return out + [a[12 + q] for q in range(3)]
    for x, y in zip(bil(a0), bil(a1)):
        assert np.abs(x - y).max() / max(np.abs(x).max(), 1.0) <= 1e-12


def test_no_cpu_fallback_symbols():
    """The product library is the HIP one, and it exports every ABI entry point."""
    L = G.lib()
    for s in G.ABI_SYMBOLS:
        assert hasattr(L, s)
    assert L.scp_abi_version() == 1


@pytest.mark.parametrize("mode,N", [(1, 10), (2, 10), (1, 70), (2, 70), (1, 4), (2, 4)])
def test_march_kernel_equals_direct_kernel(mode, N):
    """The LDS-staged march (step_kernel 0) and the direct-neighbour march
    (step_kernel 1) run the same per-site arithmetic: FP64 results are bitwise
    equal, also on ragged tiles (N = 70: a 6-wide k tile and a 2-row j tile)
    and on the smallest grid (N = 4: the halo wraps onto the tile itself)."""
    L = 0.25 * (N - 1)
    st = seeded(N, mode, L, link_sigma=0.05)[: 24 if mode == 1 else 30]
    kw = dict(g=0.1) if mode == 2 else {}
    print(f"case mode={mode} N={N}", flush=True)
    outs = []
    for sk in (0, 1):
        with G.Grid(G.make_params(N, mode, L, 0.1, step_kernel=sk, **kw)) as g:
            g.upload(st)
            g.step(9)
            g.step(11)
            outs.append(g.download())
    assert np.array_equal(outs[0], outs[1])


REAL_MARCH_CASES = {
    "base": {},
    "rmode1": dict(rmode=1),
    "rmode3": dict(rmode=3),
    "kappa_h": dict(kappa_h=0.3),
    "gpu_terms": dict(theta_sat=0.05, gamma_conv=0.8, sigma_grad=0.4, chi_chiral=1.5, sigma_cubic=0.7,
                      sigma_cross=0.5, lambda_self=0.9),
    "periodic": dict(bc_type=0),
    "alpha_cs": dict(alpha_cs=0.1),
    "beta_h": dict(beta_h=0.2),
    "sigma_freq": dict(sigma_freq=0.3),
    "pass1_all": dict(alpha_cs=0.1, beta_h=0.2, kappa_h=0.3, sigma_freq=0.3, theta_sat=0.05, gamma_conv=0.8,
                      sigma_grad=0.4, chi_chiral=1.5),
}


@pytest.mark.parametrize("N", [4, 10, 70])
@pytest.mark.parametrize("case", list(REAL_MARCH_CASES))
def test_real_march_equals_direct_kernel(case, N):
    """Real mode on the LDS-staged march against the direct-neighbour kernel:
    bitwise equal in FP64 for every real-mode variant, including the pass-1
    fields (alpha_cs, beta_h, sigma_freq: read from global memory), also on
    ragged tiles (N = 70) and the smallest grid."""
    L = 0.25 * (N - 1)
    st = seeded(N, 1, L)[:12].copy()
    rng = np.random.default_rng(N)
    st[3:6] += 0.1 * rng.standard_normal((3, N ** 3))   # theta large enough for the saturation term
    kw = REAL_MARCH_CASES[case]
    print(f"case {case} N={N}", flush=True)
    outs, names = [], []
    for sk in (0, 1):
        with G.Grid(G.make_params(N, 0, L, 0.1, step_kernel=sk, **kw)) as g:
            names.append(g.step_kernel_name)
            g.upload(st)
            g.step(9)
            g.step(11)
            outs.append(g.download())
    pass1 = any(k in kw for k in ("alpha_cs", "beta_h", "sigma_freq"))
    assert names[0] == f"k_march<scp::S64, 1, {3 if pass1 else 0}, false>", names
    assert names[1].startswith("k_step64<0"), names
    assert np.array_equal(outs[0], outs[1])


GPU_TERMS = {
    "theta_sat": dict(theta_sat=0.05, gamma_conv=0.8),
    "sigma_grad": dict(sigma_grad=0.4, chi_chiral=1.5),
    "sigma_cubic": dict(sigma_cubic=0.7),
    "sigma_freq": dict(sigma_freq=0.3),
    "sigma_cross": dict(sigma_cross=0.5),
    "lambda_self": dict(lambda_self=0.9),
    "all": dict(theta_sat=0.05, gamma_conv=0.8, sigma_grad=0.4, chi_chiral=1.5, sigma_cubic=0.7,
                sigma_freq=0.3, sigma_cross=0.5, lambda_self=0.9),
}


@pytest.mark.parametrize("terms", list(GPU_TERMS))
@pytest.mark.parametrize("N", [12, 13])
def test_real_gpu_terms_match_oracle(terms, N):
    """Real mode with the terms of the reference's GPU step (scp_sim.cu:1144-1248)
    against the oracle's restatement, bitwise in FP64 (theta noise large enough to
    cross the saturation threshold; the frequency term reads the half-kicked
    theta velocities through its own pass)."""
    L = 0.25 * (N - 1)
    st = seeded(N, 1, L)[:12].copy()
    rng = np.random.default_rng(N)
    st[3:6] = 0.2 * rng.standard_normal((3, N ** 3))
    st[9:12] = 0.1 * rng.standard_normal((3, N ** 3))
    kw = GPU_TERMS[terms]
    print(f"case real terms={terms} N={N}", flush=True)
    p_o = oracle.make_params(N, 0, L, 0.1, **kw)
    ref, _ = oracle.step(p_o, st, 12)
    with G.Grid(G.make_params(N, 0, L, 0.1, **kw)) as g:
        g.upload(st)
        g.step(5)
        g.step(7)
        out = g.download()
    assert np.array_equal(out, ref.reshape(out.shape)), f"max diff {np.abs(out - ref.reshape(out.shape)).max():.3e}"
    base, _ = oracle.step(oracle.make_params(N, 0, L, 0.1), st, 12)
    assert np.abs(out - base.reshape(out.shape)).max() > 1e-6      # the terms act
    with G.Grid(G.make_params(N, 0, L, 0.1, precision=G.PREC_F32C, lo_words=0, **kw)) as g:
        g.upload(st)
        g.step(12)
        out32 = g.download()
    assert rel_l2(out32, ref.reshape(out.shape)) <= 1e-5


def test_real_gpu_terms_forces_match_oracle():
    """scp_grid_forces with every term on: the initial velocities feed the frequency term."""
    N = 10
    L = 0.25 * (N - 1)
    st = seeded(N, 1, L)[:12].copy()
    rng = np.random.default_rng(3)
    st[3:6] = 0.2 * rng.standard_normal((3, N ** 3))
    st[9:12] = 0.1 * rng.standard_normal((3, N ** 3))
    kw = GPU_TERMS["all"]
    ref = oracle.forces(oracle.make_params(N, 0, L, 0.1, **kw), st)
    with G.Grid(G.make_params(N, 0, L, 0.1, **kw)) as g:
        g.upload(st)
        acc = g.forces()
    assert np.array_equal(acc, ref)
