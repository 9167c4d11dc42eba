"""Python view of the seed generator (scp_amd/csrc/host/seeds.c)."""
import ctypes as C

import numpy as np

from . import _native


class Ball(C.Structure):
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double), ("amp", C.c_double),
                ("sigma", C.c_double), ("omega", C.c_double), ("phase", C.c_double),
                ("winding", C.c_int)]


class QflowSeedArgs(C.Structure):
    _fields_ = [("N", C.c_int), ("L", C.c_double), ("omega0", C.c_double), ("eta", C.c_double),
                ("mth", C.c_double), ("theta_bvp", C.c_int), ("wind", C.c_int), ("twist", C.c_int),
                ("wtor", C.c_double)]


def _lib():
    L = _native.load("libscp_io.so")
    L.scp_seed_balls.argtypes = [C.c_int, C.c_double, C.c_int, C.POINTER(Ball), C.c_int, C.c_double,
                                 C.c_double, C.c_uint64, C.POINTER(C.c_double)]
    L.scp_draw_balls.argtypes = [C.c_int, C.c_double, C.c_int, C.c_uint64, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.POINTER(Ball)]
    L.scp_qflow_seed.argtypes = [C.POINTER(QflowSeedArgs), C.c_char_p, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.c_char_p, C.c_int]
    L.scp_rng_u01.restype = C.c_double
    L.scp_rng_u01.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    L.scp_rng_normal.restype = C.c_double
    L.scp_rng_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
    return L


def seed_balls(N, L, mode, balls, theta_sigma=0.01, link_sigma=0.0, seed=0):
    """State (ncols, N^3) for mode 1 (24 cols) or 2 (30 cols). balls: list of dicts
    with x0,y0,z0,amp,sigma,omega,phase,winding (physical units)."""
    arr = (Ball * max(1, len(balls)))()
    for i, b in enumerate(balls):
        arr[i] = Ball(**b)
    ncol = 30 if mode == 2 else 24
    st = np.zeros((ncol, N ** 3))
    rc = _lib().scp_seed_balls(N, L, mode, arr, len(balls), theta_sigma, link_sigma, seed,
                               st.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise ValueError("scp_seed_balls rejected its arguments")
    return st


def draw_balls(N, L, n, seed, amp=(0.6, 0.9), width_cells=(2.0, 4.0), omega=(0.6, 0.95),
               alternate_signs=True):
    arr = (Ball * n)()
    _lib().scp_draw_balls(N, L, n, seed, amp[0], amp[1], width_cells[0], width_cells[1],
                          omega[0], omega[1], 1 if alternate_signs else 0, arr)
    return [dict(x0=b.x0, y0=b.y0, z0=b.z0, amp=b.amp, sigma=b.sigma, omega=b.omega, phase=b.phase,
                 winding=b.winding) for b in arr]


def qflow_seed(N, L, profile, omega0, eta, mth, thinit="bvp", wind=0, wtor=2.0, twist=0):
    """Seed of the fixed-charge flow (libscp_io scp_qflow_seed): (state (12, N^3), g_max);
    rows u0..2, v0..2, tu0..2, tv0..2."""
    a = QflowSeedArgs(N=N, L=L, omega0=omega0, eta=eta, mth=mth, theta_bvp=1 if thinit == "bvp" else 0,
                      wind=wind, twist=twist, wtor=wtor)
    st = np.zeros((12, N ** 3))
    gmax = C.c_double()
    err = C.create_string_buffer(512)
    if _lib().scp_qflow_seed(C.byref(a), str(profile).encode(), st.ctypes.data_as(C.POINTER(C.c_double)),
                             C.byref(gmax), err, 512):
        raise ValueError(err.value.decode())
    return st, gmax.value
