"""Device seeding (scp_grid_seed_balls) and the serving loop (scp_server_*).

Bars: device seeding equals the host generator (csrc/host/seeds.c) to 1e-13
(the only differences are device exp/log/cos vs glibc, ulp level); a served
job's result is bitwise the result of the same job run alone (a server with
only that job) and of the same grid stepped directly — continuous batching
changes nothing; admission never exceeds the HBM budget."""
import numpy as np
import pytest

from scp_amd import grid as G
from scp_amd import seeds

pytestmark = pytest.mark.gpu


def balls_for(N, L, rng, n):
    dx = 2 * L / (N - 1)
    out = []
    for b in range(n):
        s = 1 if b % 2 == 0 else -1
        out.append(dict(x0=float(rng.uniform(-0.3, 0.3) * L), y0=float(rng.uniform(-0.3, 0.3) * L),
                        z0=float(rng.uniform(-0.3, 0.3) * L), amp=float(rng.uniform(0.6, 0.9)),
                        sigma=float(rng.uniform(2, 4)) * dx, omega=s * float(rng.uniform(0.6, 0.95)),
                        phase=float(rng.uniform(0, 6.28)), winding=s))
    return out


@pytest.mark.parametrize("mode,N", [(2, 24), (1, 20), (2, 33)])
def test_device_seeding_matches_host_generator(mode, N):
    L = 0.25 * (N - 1)
    balls = balls_for(N, L, np.random.default_rng(N), 3)
    host = seeds.seed_balls(N, L, mode, balls, theta_sigma=0.01, link_sigma=0.02, seed=5)
    with G.Grid(G.make_params(N, mode, L, 0.2)) as g:
        g.seed_balls(balls, theta_sigma=0.01, link_sigma=0.02, seed=5)
        dev = g.download()
    scale = np.abs(host).max()
    assert np.abs(dev - host).max() <= 1e-13 * scale


def job_list(seed=9, n=6):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        N = int(rng.choice([16, 24, 33, 40]))
        L = 0.25 * (N - 1)
        prec = G.PREC_F32C if i % 3 else G.PREC_F64
        # FP32C jobs in the production storage form and in bfloat16 words (two batch classes)
        p = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05, precision=prec, lo_words=4 if i % 2 else 1)
        jobs.append(dict(params=p, balls=balls_for(N, L, rng, 1 + i % 2), nsteps=int(rng.integers(5, 60)),
                         seed=100 + i, N=N))
    return jobs


def serve(jobs, budget, quantum, max_active=0, watch=None):
    out = {}
    with G.Server(budget, quantum=quantum, max_active=max_active) as s:
        ids = [s.submit(j["params"], j["balls"], j["nsteps"], link_sigma=1e-3, seed=j["seed"]) for j in jobs]
        for _ in range(10000):
            s.round()
            st = s.stats()
            if watch:
                watch(st)
            for k, jid in enumerate(ids):
                if k not in out and s.job(jid)["state"] == G.JOB_DONE:
                    out[k] = s.take(jid, n3=jobs[k]["N"] ** 3)
            if len(out) == len(jobs):
                break
        assert s.stats()["bytes_in_use"] == 0
    return [out[k] for k in range(len(jobs))]


def test_served_result_is_the_same_alone_and_in_the_mix():
    jobs = job_list()
    mixed = serve(jobs, budget=1 << 40, quantum=7)
    for k, j in enumerate(jobs):
        alone, = serve([j], budget=1 << 40, quantum=13)
        assert np.array_equal(mixed[k][0], alone[0]), k
        # and the same grid stepped directly, without the server
        with G.Grid(j["params"]) as g:
            g.seed_balls(j["balls"], link_sigma=1e-3, seed=j["seed"])
            g.gauss_project()
            g.step(j["nsteps"])
            assert np.array_equal(g.download(), mixed[k][0]), k


def test_admission_respects_the_hbm_budget():
    jobs = job_list(seed=4, n=8)
    sizes = sorted(G.grid_bytes(j["params"]) for j in jobs)
    budget = sizes[-1] + sizes[-2]          # room for about two jobs at a time
    seen = []
    serve(jobs, budget=budget, quantum=5, watch=lambda st: seen.append(st))
    assert max(st["bytes_in_use"] for st in seen) <= budget
    fit = max(k for k in range(len(sizes) + 1) if sum(sizes[:k]) <= budget)   # most jobs that fit at once
    assert max(st["active"] + st["done"] for st in seen) <= fit
    assert seen[-1]["site_updates"] == sum(j["nsteps"] * j["N"] ** 3 for j in jobs)


def test_server_rejects_what_it_cannot_run():
    with G.Server(1 << 20, quantum=5) as s:
        p = G.make_params(64, G.MODE_GAUGED, 15.75, 0.2)
        with pytest.raises(G.ScpError):
            s.submit(p, [], 10)             # more HBM than the budget
    with G.Server(1 << 40, quantum=5) as s:
        with pytest.raises(G.ScpError):
            s.submit(G.make_params(16, G.MODE_REAL, 3.75, 0.2), [], 10)
        with pytest.raises(G.ScpError):
            s.submit(G.make_params(16, G.MODE_GAUGED, 3.75, 0.2), [], 0)
