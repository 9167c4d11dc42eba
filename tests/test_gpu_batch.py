"""Batched stepping (scp_batch_*): many grids of different sizes in one launch
per step, chains replayed from captured HIP graphs. Every result must be
bitwise the result of stepping each grid alone (scp_grid_step), across graph
replays with eager work between them, per-member step counts, the autotuner
and every launch configuration."""
import numpy as np
import pytest

from scp_amd import grid as G
from scp_amd import seeds

pytestmark = pytest.mark.gpu

SIZES = (16, 24, 33, 40, 70)


def ball(N, seed):
    L = 0.25 * (N - 1)
    dx = 2 * L / (N - 1)
    rng = np.random.default_rng(seed)
    b = [dict(x0=float(rng.uniform(-0.3, 0.3) * L), y0=float(rng.uniform(-0.3, 0.3) * L), z0=0.0, amp=0.8,
              sigma=3 * dx, omega=float(rng.uniform(0.6, 0.95)), phase=float(rng.uniform(0, 6.28)), winding=1)]
    return L, seeds.seed_balls(N, L, 2, b, theta_sigma=0.01, link_sigma=0.02, seed=seed)


def make(N, seed, prec, lo):
    L, st = ball(N, seed)
    g = G.Grid(G.make_params(N, G.MODE_GAUGED, L, 0.2, g=0.1, precision=prec, lo_words=lo))
    g.upload(st)
    return g


@pytest.mark.parametrize("prec,lo", [(G.PREC_F64, 0), (G.PREC_F32C, 1), (G.PREC_F32C, 0), (G.PREC_F32C, 3), (G.PREC_F32C, 4)])
def test_batch_equals_each_grid_alone(prec, lo):
    print(f"case prec={prec} lo={lo} sizes={SIZES}", flush=True)
    alone = [make(N, 100 + i, prec, lo) for i, N in enumerate(SIZES)]
    mixed = [make(N, 100 + i, prec, lo) for i, N in enumerate(SIZES)]
    counts = [3, 5, 0, 7, 2]
    with G.Batch(mixed) as b:
        for rep in range(3):                       # the same chain replayed 3 times
            for g in alone:
                g.step(7)
            b.step(7)
            for g, h in zip(alone, mixed):        # eager work between the replays
                h.diag()
                assert np.array_equal(g.download(), h.download()), (rep, g.params.N)
        for g, n in zip(alone, counts):
            if n:
                g.step(n)
        b.step(counts)
        for g, h, n in zip(alone, mixed, counts):
            assert np.array_equal(g.download(), h.download()), ("counts", g.params.N)
            assert h.steps_done == g.steps_done
    for g in alone + mixed:
        g.close()


@pytest.mark.parametrize("N", [62, 126])
def test_single_grid_tile_strips_equal_the_batch(N):
    """The single-grid FP32C launch takes its tiles in strips of four tile rows
    (j fastest in a strip) when the tile rows split into such strips; on such
    grids, ragged ones included (62: one 62-wide k tile and a 2-row last j
    tile; 126: two k tiles, the second 62 wide), its result is bitwise that of
    the same grid in a batch, whose tiles come from a table."""
    print(f"case N={N}", flush=True)
    alone, member = make(N, 300, G.PREC_F32C, 4), make(N, 300, G.PREC_F32C, 4)
    with G.Batch([member]) as b:
        alone.step(9)
        b.step(9)
        assert np.array_equal(alone.download(), member.download())
    alone.close()
    member.close()


@pytest.mark.parametrize("di,persistent", [(16, False), (16, True), (128, True), (32, False)])
def test_batch_launch_configs_agree(di, persistent):
    ref = [make(N, 200 + i, G.PREC_F32C, 1) for i, N in enumerate(SIZES)]
    got = [make(N, 200 + i, G.PREC_F32C, 1) for i, N in enumerate(SIZES)]
    for g in ref:
        g.step(6)
    with G.Batch(got) as b:
        b.set_config(di=di, persistent=persistent)
        b.step(6)
    for g, h in zip(ref, got):
        assert np.array_equal(g.download(), h.download()), g.params.N
        g.close()
        h.close()


def test_batch_autotune_keeps_results():
    ref = [make(N, 300 + i, G.PREC_F32C, 1) for i, N in enumerate(SIZES)]
    got = [make(N, 300 + i, G.PREC_F32C, 1) for i, N in enumerate(SIZES)]
    with G.Batch(got) as b:
        taken = b.tune()
        cfg = b.config
        b.step(4)
    print(f"tuned: {cfg}, steps taken {taken}")
    assert cfg["tuned"] and cfg["di"] in (16, 32, 64, 128) and cfg["tile"] == (64, 4)
    for g, h in zip(ref, got):
        g.step(taken + 4)
        assert np.array_equal(g.download(), h.download()), g.params.N
        g.close()
        h.close()


@pytest.mark.parametrize("prec,lo", [(G.PREC_F64, 0), (G.PREC_F32C, 4)])
def test_real_batch_equals_each_grid_alone(prec, lo):
    def mk(i, N):
        L, st = ball(N, 400 + i)
        g = G.Grid(G.make_params(N, G.MODE_REAL, L, 0.2, precision=prec, lo_words=lo))
        g.upload(st[:12].copy())
        return g
    print(f"case real prec={prec} lo={lo} sizes={SIZES}", flush=True)
    alone = [mk(i, N) for i, N in enumerate(SIZES)]
    mixed = [mk(i, N) for i, N in enumerate(SIZES)]
    with G.Batch(mixed) as b:
        for rep in range(2):
            for g in alone:
                g.step(5)
            b.step(5)
            for g, h in zip(alone, mixed):
                assert np.array_equal(g.download(), h.download()), (rep, g.params.N)
    for g in alone + mixed:
        g.close()


@pytest.mark.parametrize("tile", [None, (32, 8), (0, 0), (64, 4)])
def test_complex_f32c_batch_shapes_equal_each_grid_alone(tile):
    """Complex-mode FP32C grids (int8 position words) batched under every 256-thread
    shape (None: the members' own, here mixed) step bitwise as they step alone."""
    def mk(i, N):
        L, st = ball(N, 450 + i)
        g = G.Grid(G.make_params(N, G.MODE_COMPLEX, L, 0.2, precision=G.PREC_F32C, lo_words=4))
        g.upload(st[:24].copy())
        return g
    print(f"case complex f32c tile={tile}", flush=True)
    alone = [mk(i, N) for i, N in enumerate(SIZES)]
    batched = [mk(i, N) for i, N in enumerate(SIZES)]
    with G.Batch(batched) as b:
        if tile is not None:
            b.set_tile(*tile)
        assert b.config["tile"] == (tile if tile is not None else (0, 0))
        for rep in range(2):
            for g in alone:
                g.step(4)
            b.step(4)
            for g, h in zip(alone, batched):
                assert np.array_equal(g.download(), h.download()), (rep, g.params.N)
    for g in alone + batched:
        g.close()


def test_batch_membership_rules():
    a = make(16, 1, G.PREC_F32C, 1)
    b64 = make(16, 2, G.PREC_F64, 0)
    L = 0.25 * 15
    real = G.Grid(G.make_params(16, G.MODE_REAL, L, 0.2, alpha_cs=0.1))
    with G.Batch([a]) as b:
        with pytest.raises(G.ScpError):
            b.add(b64)                 # other class
        with pytest.raises(G.ScpError, match="cannot join"):
            b.add(real)                # real mode with the pass-1 fields: direct kernel only
        with pytest.raises(G.ScpError):
            b.add(a)                   # already a member
        b.remove(a)
        assert len(b) == 0
        b.add(a)
        a.close()                      # a destroyed grid leaves its batch
        b.members.remove(a)
        assert G.lib().scp_batch_size(b._h) == 0
    b64.close()
    real.close()


@pytest.mark.parametrize("mode", [G.MODE_COMPLEX, G.MODE_GAUGED])
def test_fp64_batch_members_match_the_oracle(mode):
    """FP64 grids of different sizes stepped together by one batch (graph-replayed
    chains, per-member step counts) against the oracle's restatement of the
    reference step on each member's uploaded state: bitwise in complex mode,
    within 1e-12 relative field L2 gauged (this build's link sincos)."""
    from oracle import core as oracle
    sizes, counts = (10, 16, 21), (9, 14, 6)
    states, grids = [], []
    for i, N in enumerate(sizes):
        L, st = ball(N, 500 + i)
        st = st if mode == G.MODE_GAUGED else st[:24].copy()
        g = G.Grid(G.make_params(N, mode, L, 0.2, g=0.1))
        g.upload(st)
        states.append((L, st))
        grids.append(g)
    print(f"case fp64 batch vs oracle mode={mode}", flush=True)
    with G.Batch(grids) as b:
        b.step([4, 4, 4])
        b.step([c - 4 for c in counts])
    for (L, st), g, n in zip(states, grids, counts):
        N = g.params.N
        ref, _ = oracle.step(oracle.make_params(N, mode, L, 0.2, g=0.1), st, n)
        out = g.download()
        if mode == G.MODE_GAUGED:
            assert np.linalg.norm(out - ref) / np.linalg.norm(ref) <= 1e-12, N
        else:
            assert np.array_equal(out, ref), N
        g.close()


@pytest.mark.parametrize("tk,tj,persistent", [(128, 4, False), (64, 8, False), (64, 2, False), (64, 2, True),
                                              (64, 8, True), (32, 8, False), (32, 8, True), (0, 0, False),
                                              (0, 0, True)])
def test_batch_tile_shapes_equal_each_grid_alone(tk, tj, persistent):
    """The other tile shapes of FP32C batches with int8 position words (128 x 4,
    64 x 8, 64 x 2, and 32 x 8 with two 32-site rows per wave, and 0 x 0: each
    grid its own of 64 x 4 and 32 x 8 in one launch; ragged edges 33, 70) step
    every member bitwise as it steps alone, also as persistent blocks and across
    chain replays."""
    print(f"case tile {tk}x{tj} persistent={persistent}", flush=True)
    ref = [make(N, 600 + i, G.PREC_F32C, 4) for i, N in enumerate(SIZES)]
    got = [make(N, 600 + i, G.PREC_F32C, 4) for i, N in enumerate(SIZES)]
    with G.Batch(got) as b:
        b.set_tile(tk, tj)
        b.set_config(di=32, persistent=persistent)
        assert b.config["tile"] == (tk, tj)
        for rep in range(2):
            b.step([5, 3, 6, 5, 4])
            for g, n in zip(ref, [5, 3, 6, 5, 4]):
                g.step(n)
            for g, h in zip(ref, got):
                assert np.array_equal(g.download(), h.download()), (rep, g.params.N)
    for g in ref + got:
        g.close()


def test_batch_autotune_over_tile_shapes_keeps_results():
    """FP32C int8-position-word batches tune over the tile shapes as well (the five
    uniform ones and, members differing in their own shape here, the mixed launch)."""
    ref = [make(N, 700 + i, G.PREC_F32C, 4) for i, N in enumerate(SIZES)]
    got = [make(N, 700 + i, G.PREC_F32C, 4) for i, N in enumerate(SIZES)]
    with G.Batch(got) as b:
        taken = b.tune()
        cfg = b.config
        b.step(3)
    print(f"tuned: {cfg}, steps taken {taken}")
    assert cfg["tuned"] and cfg["tile"] in ((64, 4), (128, 4), (64, 8), (64, 2), (32, 8), (0, 0))
    assert taken == 6 * 2 * 5 * 4 + 2 * 4 * 4     # every candidate, then four finalists twice
    for g, h in zip(ref, got):
        g.step(taken + 3)
        assert np.array_equal(g.download(), h.download()), g.params.N
        g.close()
        h.close()


def test_batch_tile_shapes_are_refused_for_other_classes():
    g = make(16, 1, G.PREC_F64, 0)
    with G.Batch([g]) as b:
        with pytest.raises(G.ScpError, match="int8 position words"):
            b.set_tile(128, 4)
        with pytest.raises(G.ScpError, match="int8 position words"):
            b.set_tile(32, 8)
        with pytest.raises(G.ScpError, match="the shapes are"):
            b.set_tile(64, 3)
        b.set_tile(64, 4)
    g.close()


def test_batch_default_shape_follows_members():
    """An FP32C int8-position-word batch nobody configured gives every member the
    256-thread shape it runs fastest (32 x 8 for edges 16, 24, 40, 70, 64 x 4 for
    33): one mixed launch while both are present, the uniform launch once the
    members agree, re-decided as they leave; bf16-word batches stay at 64 x 4.
    Results stay bitwise each grid's alone across the switches."""
    got = [make(N, 800 + i, G.PREC_F32C, 4) for i, N in enumerate(SIZES)]
    ref = [make(N, 800 + i, G.PREC_F32C, 4) for i, N in enumerate(SIZES)]
    odd = SIZES.index(33)
    with G.Batch(got) as b:
        assert b.config["tile"] == (0, 0)          # per grid
        b.step(3)
        b.remove(got[odd])
        assert b.config["tile"] == (32, 8)
        b.step(2)
    with G.Batch([got[odd]]) as b:
        assert b.config["tile"] == (64, 4)
        b.step(2)
    for i, (g, h) in enumerate(zip(ref, got)):
        g.step(3 + 2)
        assert np.array_equal(g.download(), h.download()), g.params.N
    bf = make(40, 901, G.PREC_F32C, 1)
    with G.Batch([bf]) as b:
        assert b.config["tile"] == (64, 4)
    for g in ref + got + [bf]:
        g.close()
