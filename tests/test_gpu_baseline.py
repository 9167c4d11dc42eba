"""The measurement workloads at their specified sizes, through the C-ABI.

* gauged 512^3 periodic, 64 seeded Q-balls (alternating charge signs),
  Gauss-projected, 10^4 steps: the FP32C path (int8 position words, the
  production form) against the FP64 path side by side, diagnostics every
  1000 steps. Bars, at every diagnostics row: |E32 - E64| <= 1e-6 |E64| and
  |Q32 - Q64| <= 1e-6 of the balls' summed |Q| (the net charge of alternating
  balls nearly cancels, so it is no scale); Gauss residual max <= 1e-6 in
  FP32C (measured 8e-8 at the end) and <= 1e-12 in FP64 (measured 1e-13).
* 256 independent gauged grids, edges drawn from {32, 48, 64, 80, 96}, one
  Q-ball each, FP32C, one autotuned batch, 5000 steps: sampled members are
  bitwise the same grid stepped alone, and their E_total and Q_total are
  within 1e-6 of an FP64 batch of the same members.
* a sample of the serving job mix (Poisson arrivals at 4 jobs/s, edges
  log-uniform in [48, 256], 1-4 balls, steps uniform in [1000, 20000]) under
  an HBM budget that holds about two of them: the budget holds at every
  round, every job completes, and the sampled jobs are bitwise the same job
  run alone on a grid of its own.
Progress lines go straight to the terminal (they name the case and step).
"""
import math
import time

import numpy as np
import pytest

from scp_amd import grid as G
from scp_amd import seeds

pytestmark = pytest.mark.gpu


def _say(capsys, msg):
    with capsys.disabled():
        print(f"[baseline] {msg}", flush=True)


def ball_charge_scale(balls):
    """Sum over balls of |Q| for Phi_a = A exp(-r^2 / (2 s^2)) e^{i(...)},
    dPhi/dt = i w Phi: |Q| = 3 |w| A^2 pi^(3/2) s^3 (the box holds each ball)."""
    return sum(3 * abs(b["omega"]) * b["amp"] ** 2 * math.pi ** 1.5 * b["sigma"] ** 3 for b in balls)


@pytest.mark.timeout(1200)
def test_cfg2_512_fp32c_against_fp64_over_ten_thousand_steps(capsys):
    N, steps, chunk = 512, 10000, 1000
    L = 0.25 * (N - 1)
    balls = seeds.draw_balls(N, L, 64, 1002)
    qscale = ball_charge_scale(balls)
    p64 = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05)
    p32 = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05, precision=G.PREC_F32C, lo_words=4)
    _say(capsys, f"case cfg2 512^3: seeding on the device, sum|Q| = {qscale:.4e}")
    rows = []
    with G.Grid(p64) as g64, G.Grid(p32) as g32:
        for g in (g64, g32):
            g.seed_balls(balls, theta_sigma=0.01, link_sigma=0.0, seed=1002)
            it, gm = g.gauss_project()
            _say(capsys, f"Gauss projection {it} CG iterations, max|G - Gbar| {gm:.2e}")
        d64, d32 = g64.diag(), g32.diag()
        assert abs(d32["E_total"] - d64["E_total"]) <= 1e-12 * abs(d64["E_total"])
        done, t0 = 0, time.time()
        while done < steps:
            d64, d32 = g64.step_diag(chunk), g32.step_diag(chunk)
            done += chunk
            dE = (d32["E_total"] - d64["E_total"]) / abs(d64["E_total"])
            dQ = (d32["Q_total"] - d64["Q_total"]) / qscale
            rows.append((done, dE, dQ, d32["gauss_max"], d64["gauss_max"]))
            _say(capsys, f"cfg2 step {done}: dE/E {dE:.3e} dQ/sum|Q| {dQ:.3e} gauss f32c {d32['gauss_max']:.2e} "
                         f"f64 {d64['gauss_max']:.2e} E {d64['E_total']:.9e} ({time.time() - t0:.0f}s)")
        s64 = g64.download()
        s32 = g32.download()
    s32 -= s64
    err = math.sqrt(sum(float(np.dot(c, c)) for c in s32)) / math.sqrt(sum(float(np.dot(c, c)) for c in s64))
    del s32, s64
    _say(capsys, f"cfg2 after {steps} steps: FP32C field L2 against FP64 {err:.3e}")
    for done, dE, dQ, g32m, g64m in rows:
        assert abs(dE) <= 1e-6, (done, dE)
        assert abs(dQ) <= 1e-6, (done, dQ)
        assert g32m <= 1e-6, (done, g32m)
        assert g64m <= 1e-12, (done, g64m)
    assert err <= 1e-3


def _sweep_member(N, rng):
    L = 0.25 * (N - 1)
    dx = 0.5
    b = dict(x0=float(rng.uniform(-0.2, 0.2) * L), y0=float(rng.uniform(-0.2, 0.2) * L),
             z0=float(rng.uniform(-0.2, 0.2) * L), amp=float(rng.uniform(0.6, 0.9)),
             sigma=float(rng.uniform(2.0, 4.0)) * dx, omega=float(rng.uniform(0.6, 0.95)),
             phase=float(rng.uniform(0, 2 * np.pi)), winding=1)
    return L, [b]


def _grid(N, L, balls, seed, prec):
    g = G.Grid(G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05, precision=prec,
                             lo_words=4 if prec == G.PREC_F32C else 0))
    g.seed_balls(balls, theta_sigma=0.01, link_sigma=0.0, seed=seed)
    return g


@pytest.mark.timeout(900)
def test_cfg3_batched_sweep_of_256_grids_5000_steps(capsys):
    steps = 5000
    rng = np.random.default_rng(1003)
    sizes = [int(n) for n in rng.choice([32, 48, 64, 80, 96], size=256)]
    specs = [(N,) + _sweep_member(N, rng) for N in sizes]
    _say(capsys, f"case cfg3: 256 grids, {sum(n ** 3 for n in sizes) / 1e6:.1f} M sites")
    grids = [_grid(N, L, balls, 10030000 + k, G.PREC_F32C) for k, (N, L, balls) in enumerate(specs)]
    t0 = time.time()
    with G.Batch(grids) as b:
        taken = b.tune()
        _say(capsys, f"cfg3 tuned {b.config} in {taken} steps")
        left = steps - taken
        while left > 0:
            n = min(1000, left)
            b.step(n)
            left -= n
            _say(capsys, f"cfg3 {steps - left} steps ({time.time() - t0:.0f}s)")
    # one member of every edge, plus the last grid
    sample = sorted({sizes.index(n) for n in set(sizes)} | {len(sizes) - 1})
    f64 = [_grid(*specs[k], 10030000 + k, G.PREC_F64) for k in sample]
    with G.Batch(f64) as b64:
        b64.step(steps)
    for k, h in zip(sample, f64):
        N, L, balls = specs[k]
        g = grids[k]
        assert g.steps_done == steps
        with _grid(N, L, balls, 10030000 + k, G.PREC_F32C) as alone:
            alone.step(steps)
            assert np.array_equal(alone.download(), g.download()), (k, N)
        d32, d64 = g.diag(), h.diag()
        dE = abs(d32["E_total"] - d64["E_total"]) / abs(d64["E_total"])
        dQ = abs(d32["Q_total"] - d64["Q_total"]) / abs(d64["Q_total"])
        _say(capsys, f"cfg3 member {k} (N={N}): batched == alone; FP32C vs FP64 dE/E {dE:.2e} dQ/Q {dQ:.2e}")
        assert dE <= 1e-6 and dQ <= 1e-6, (k, N, dE, dQ)
    for g in grids + f64:
        g.close()


def _serve_mix(rng, n):
    """n jobs of the serving mix with Poisson arrivals at 4 jobs/s."""
    jobs, t = [], 0.0
    for _ in range(n):
        t += rng.exponential(1.0 / 4.0)
        N = int(round(np.exp(rng.uniform(np.log(48), np.log(256)))))
        L = 0.25 * (N - 1)
        dx = 0.5
        balls = []
        for b in range(int(rng.integers(1, 5))):
            sg = 1 if b % 2 == 0 else -1
            balls.append(dict(x0=float(rng.uniform(-0.3, 0.3) * L), y0=float(rng.uniform(-0.3, 0.3) * L),
                              z0=float(rng.uniform(-0.3, 0.3) * L), amp=float(rng.uniform(0.6, 0.9)),
                              sigma=float(rng.uniform(2.0, 4.0)) * dx, omega=sg * float(rng.uniform(0.6, 0.95)),
                              phase=float(rng.uniform(0, 2 * np.pi)), winding=sg))
        p = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05, precision=G.PREC_F32C, lo_words=4)
        jobs.append(dict(t=t, N=N, params=p, balls=balls, nsteps=int(rng.integers(1000, 20001))))
    return jobs


@pytest.mark.timeout(900)
def test_cfg4_serving_job_mix_sample(capsys):
    jobs = _serve_mix(np.random.default_rng(1004), 6)
    sizes = sorted(G.grid_bytes(j["params"]) for j in jobs)
    budget = sizes[-1] + sizes[-2]
    work = sum(j["N"] ** 3 * j["nsteps"] for j in jobs)
    _say(capsys, "case cfg4 sample: " + ", ".join(f"{j['N']}^3 x {j['nsteps']}" for j in jobs)
         + f"; {work / 1e9:.0f} G site-updates, budget {budget / 1e9:.1f} GB")
    results, seen, ids = {}, [], {}
    with G.Server(budget, quantum=200) as s:
        t0 = s.now()
        while len(results) < len(jobs):
            now = s.now() - t0
            for k, j in enumerate(jobs):
                if k not in ids and j["t"] <= now:
                    ids[k] = s.submit(j["params"], j["balls"], j["nsteps"], link_sigma=0.0, seed=10040 + k)
            s.round()
            st = s.stats()
            seen.append(st)
            assert st["bytes_in_use"] <= budget
            for k, jid in ids.items():
                if k not in results and s.job(jid)["state"] == G.JOB_DONE:
                    results[k] = s.take(jid, n3=jobs[k]["N"] ** 3)
                    _say(capsys, f"cfg4 job {k} ({jobs[k]['N']}^3 x {jobs[k]['nsteps']}) done at "
                                 f"{s.now() - t0:.1f}s")
            if len(ids) < len(jobs) and st["active"] == 0 and st["queued"] == 0:
                time.sleep(0.05)
        assert s.stats()["bytes_in_use"] == 0
    assert seen[-1]["site_updates"] == work
    assert max(st["active"] + st["done"] for st in seen) >= 2     # jobs shared the GPU
    # the two jobs of least work, run alone on a grid of their own
    for k in sorted(range(len(jobs)), key=lambda q: jobs[q]["N"] ** 3 * jobs[q]["nsteps"])[:2]:
        j = jobs[k]
        with G.Grid(j["params"]) as g:
            g.seed_balls(j["balls"], link_sigma=0.0, seed=10040 + k)
            g.gauss_project()
            g.step(j["nsteps"])
            assert np.array_equal(g.download(), results[k][0]), k
        _say(capsys, f"cfg4 job {k}: served == alone")
