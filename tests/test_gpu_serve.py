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
    budget = sizes[-1] + sizes[-2]          # room for about two whole grids at a time
    seen = []
    serve(jobs, budget=budget, quantum=5, watch=lambda st: seen.append(st))
    assert max(st["bytes_in_use"] for st in seen) <= budget
    assert max(st["active"] + st["done"] for st in seen) >= 2
    assert seen[-1]["site_updates"] == sum(j["nsteps"] * j["N"] ** 3 for j in jobs)


def emit_bytes(N):
    """The FP64 emit buffers of a gauged grid: 2 x 15 fields x N^3 (64-aligned) doubles."""
    return 2 * 15 * ((N ** 3 + 63) // 64 * 64) * 8


def test_jobs_hold_their_emit_buffers_only_when_they_emit():
    """An active FP32C gauged job holds its step state alone (the grid's bytes less
    the FP64 emit buffers: 212 of 452 bytes per site at edges that are multiples
    of 32, more at this edge, whose step-state rows are padded from 40 to 64);
    a finished one holds the
    emit buffers (its result) and not the step state; taking it frees all."""
    N, L = 40, 9.75
    p = G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.05, precision=G.PREC_F32C, lo_words=4)
    full = G.grid_bytes(p)
    with G.Server(1 << 40, quantum=5) as s:
        jid = s.submit(p, balls_for(N, L, np.random.default_rng(1), 2), 12, link_sigma=1e-3, seed=3)
        s.round()
        info = s.job(jid)
        assert info["state"] == G.JOB_ACTIVE
        # scp_grid_bytes counts the largest diagnostics partials (1024 rows); a 40^3 grid holds 250
        assert full - emit_bytes(N) - (300 << 10) <= info["bytes"] <= full - emit_bytes(N) + (64 << 10), \
            (info["bytes"], full)
        while s.job(jid)["state"] != G.JOB_DONE:
            s.round()
        done = s.job(jid)["bytes"]
        assert emit_bytes(N) <= done <= emit_bytes(N) + (600 << 10), (done, emit_bytes(N))
        assert s.stats()["bytes_in_use"] == done
        s.take(jid, n3=N ** 3)
        assert s.stats()["bytes_in_use"] == 0


@pytest.mark.parametrize("aging", [8, 0])
def test_admission_passes_over_a_blocked_job_until_it_ages(aging):
    """Queue: big, big, small, small under a budget that holds one big job and the
    small ones. Small jobs pass over the blocked second big job (within the aging
    bound); with aging 0 admission keeps strict arrival order."""
    rng = np.random.default_rng(5)
    mk = lambda N, n: dict(params=G.make_params(N, G.MODE_GAUGED, 0.25 * (N - 1), 0.2, g=0.05,
                                                precision=G.PREC_F32C, lo_words=4),
                           balls=balls_for(N, 0.25 * (N - 1), rng, 2), nsteps=n, N=N)
    jobs = [mk(40, 60), mk(40, 10), mk(16, 10), mk(16, 10)]
    admit = lambda N: G.grid_bytes(jobs[0 if N == 40 else 2]["params"]) + 5 * N ** 3 * 8 + (32 << 10)
    resident_big = G.grid_bytes(jobs[0]["params"]) - emit_bytes(40)
    budget = admit(40) + (1 << 20)
    assert resident_big + admit(40) > budget >= resident_big + int(2.5 * admit(16))
    with G.Server(budget, quantum=5) as s:
        s.set_aging(aging)
        ids = [s.submit(j["params"], j["balls"], j["nsteps"], link_sigma=1e-3, seed=40 + k) for k, j in enumerate(jobs)]
        taken = set()
        while len(taken) < len(jobs):
            s.round()
            assert s.stats()["bytes_in_use"] <= budget
            for k, jid in enumerate(ids):
                if k not in taken and s.job(jid)["state"] == G.JOB_DONE:
                    info = s.job(jid)
                    jobs[k]["t_admit"], jobs[k]["t_done"] = info["t_admit"], info["t_done"]
                    s.take(jid, n3=jobs[k]["N"] ** 3)
                    taken.add(k)
    if aging > 0:   # the small jobs ran while the first big job held the memory the second one needed
        assert jobs[2]["t_done"] < jobs[0]["t_done"] and jobs[3]["t_done"] < jobs[0]["t_done"]
        assert jobs[2]["t_admit"] < jobs[1]["t_admit"]
    else:
        assert jobs[2]["t_admit"] >= jobs[1]["t_admit"] and jobs[3]["t_admit"] >= jobs[1]["t_admit"]


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


@pytest.mark.parametrize("mode", [G.MODE_COMPLEX, G.MODE_GAUGED])
def test_served_fp64_jobs_match_the_oracle(mode):
    """FP64 jobs through the request loop (device seeding, Gauss projection when
    gauged, continuous batching in quanta) against the oracle's restatement of
    the reference step from the same seeded and projected state (taken from a
    grid seeded and projected alone, which is bitwise the server's): bitwise in
    complex mode, within 1e-12 relative field L2 gauged."""
    from oracle import core as oracle
    rng = np.random.default_rng(21 + mode)
    jobs = []
    for i, N in enumerate((12, 18, 16)):
        L = 0.25 * (N - 1)
        p = G.make_params(N, mode, L, 0.2, g=0.05)
        jobs.append(dict(params=p, balls=balls_for(N, L, rng, 2), nsteps=int(rng.integers(8, 30)), seed=700 + i, N=N))
    print(f"case served fp64 vs oracle mode={mode}", flush=True)
    served = serve(jobs, budget=1 << 40, quantum=5) if mode == G.MODE_GAUGED else None
    if served is None:   # complex jobs: the server seeds and steps them without projection
        out = {}
        with G.Server(1 << 40, quantum=5) as s:
            ids = [s.submit(j["params"], j["balls"], j["nsteps"], link_sigma=0.0, seed=j["seed"],
                            gauss_project=False) for j in jobs]
            while len(out) < len(jobs):
                s.round()
                for k, jid in enumerate(ids):
                    if k not in out and s.job(jid)["state"] == G.JOB_DONE:
                        out[k] = s.take(jid, mode=mode, n3=jobs[k]["N"] ** 3)
        served = [out[k] for k in range(len(jobs))]
    for k, j in enumerate(jobs):
        N, p = j["N"], j["params"]
        with G.Grid(p) as g:
            g.seed_balls(j["balls"], link_sigma=1e-3 if mode == G.MODE_GAUGED else 0.0, seed=j["seed"])
            goff = 0.0
            if mode == G.MODE_GAUGED:
                g.gauss_project()
                goff = g.g_offset
            st0 = g.download()
        ref, _ = oracle.step(oracle.make_params(N, mode, p.L, 0.2, g=0.05, G_offset=goff), st0, j["nsteps"])
        got = served[k][0]
        if mode == G.MODE_GAUGED:
            assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-12, k
        else:
            assert np.array_equal(got, ref), k
