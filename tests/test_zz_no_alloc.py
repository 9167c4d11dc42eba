"""scp_grid_step_diag_noalloc: the fused step + diagnostics with its work space
in the library's static device scratch, against scp_grid_step_diag.

Two calls in a row on the scratch, the second on a smaller grid (so the
scratch still holds the first call's rows and charge densities), plus a
third on a complex grid: after each, the state is bitwise that of the
allocating entry point on an identical grid and every diagnostic equals its
value (same kernels, same fixed reduction order). Device memory in use does
not grow across a call after the first one on the device (whose symbol lookup
loads the code object that holds the static scratch); a grid beyond the
scratch maximum is refused.
"""
import ctypes as C

import numpy as np
import pytest

from scp_amd import grid as G
from tests.test_gpu_parity import seeded

pytestmark = pytest.mark.gpu


def _free_bytes():
    hip = C.CDLL("libamdhip64.so.7")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _pair(mode, N, prec, steps, project, measure):
    L = 0.25 * (N - 1)
    st = seeded(N, max(mode, 1), L, link_sigma=0.05)[: G.ncols(mode)].copy()
    kw = dict(g=0.1) if mode == 2 else {}
    p = G.make_params(N, mode, L, 0.1, precision=prec, lo_words=4 if prec == G.PREC_F32C else 0, **kw)
    res = []
    for noalloc in (False, True):
        with G.Grid(p) as g:
            g.upload(st)
            if project:
                g.gauss_project()
            g.step(3)
            if noalloc:
                before = _free_bytes()
                d = g.step_diag_noalloc(steps)
                after = _free_bytes()
                if measure:
                    assert after >= before, (before, after)
            else:
                d = g.step_diag(steps)
            res.append((d, g.download(), g.steps_done))
    return res


def test_noalloc_matches_step_diag_on_consecutive_calls():
    cases = [(2, 48, G.PREC_F64, 7, True), (2, 24, G.PREC_F32C, 5, True), (1, 20, G.PREC_F64, 4, False)]
    for c, (mode, N, prec, steps, project) in enumerate(cases):
        print(f"case noalloc mode={mode} N={N} prec={prec} steps={steps}", flush=True)
        (d_a, s_a, n_a), (d_b, s_b, n_b) = _pair(mode, N, prec, steps, project, measure=c > 0)
        assert n_a == n_b == 3 + steps
        assert np.array_equal(s_a, s_b)
        for k in G.DIAG_NAMES:
            assert d_a[k] == d_b[k], (mode, N, k, d_a[k], d_b[k])


def test_noalloc_refuses_grids_above_the_scratch():
    N = 210
    L = 0.25 * (N - 1)
    with G.Grid(G.make_params(N, G.MODE_GAUGED, L, 0.1, precision=G.PREC_F32C, lo_words=4)) as g:
        with pytest.raises(G.ScpError, match="static scratch"):
            g.step_diag_noalloc(1)
        assert g.steps_done == 0
