"""Python view of the configuration module (scp_amd/csrc/host/scp_config.c).

Same keys, defaults, file grammar and validation as the reference
sfa/sim/scp_config.h; fields are read back through a ctypes mirror of the
scp_config struct.
"""
import ctypes as C

from . import _native

_D, _I = C.c_double, C.c_int32


def _s(n):
    return C.c_char * n


class ScpConfig(C.Structure):
    _fields_ = [
        ("N", _I), ("L", _D), ("T", _D), ("dt_factor", _D),
        ("m2", _D), ("mtheta2", _D), ("eta", _D), ("mu", _D), ("kappa", _D), ("kappa_h", _D),
        ("alpha_cs", _D), ("beta_h", _D),
        ("mode", _I), ("inv_alpha", _D), ("inv_beta", _D), ("kappa_gamma", _D), ("gamma_conf", _D),
        ("theta_sat", _D), ("gamma_conv", _D),
        ("sigma_grad", _D), ("chi_chiral", _D), ("sigma_cubic", _D), ("sigma_freq", _D),
        ("sigma_cross", _D), ("lambda_self", _D), ("theta_vev", _D), ("tune_dt", _D),
        ("sweep", _I), ("sweep_T", _D),
        ("self_tune", _I), ("st_dt", _D), ("st_eps", _D), ("st_gamma", _D), ("st_pcrit", _D),
        ("st_charge", _D), ("st_project", _I), ("st_ptarget", _D), ("st_gain", _D),
        ("complex_phi", _I), ("qball_profile", _s(512)), ("qball_omega", _D), ("qball_x0", _D),
        ("qball_y0", _D), ("qball_z0", _D),
        ("qdiag_radius", _D), ("qdiag_probe1", _D), ("qdiag_probe2", _D),
        ("complex_gauge", _I), ("g_gauge", _D), ("qball2_x0", _D), ("qball2_y0", _D), ("qball2_z0", _D),
        ("qball2_sign", _I), ("qball2_phase", _D), ("test_gauge_xform", _I),
        ("higgs_v", _D), ("higgs_lam", _D), ("higgs_kap", _D), ("higgs_rvoid", _D),
        ("g_gauge2", _D), ("gauge2_q0", _D), ("gauge2_q1", _D), ("gauge2_q2", _D),
        ("bc_type", _I), ("damp_width", _D), ("damp_rate", _D), ("bc_switch_time", _D),
        ("gradient_A_high", _D), ("gradient_A_low", _D), ("gradient_margin", _I),
        ("init", _s(32)), ("A", _D), ("sigma", _D), ("A_bg", _D), ("ellip", _D), ("R_tube", _D),
        ("delta", _D * 3),
        ("init_sfa", _s(512)), ("init_frame", _I), ("init_exec", _s(1024)),
        ("output", _s(512)), ("diag_file", _s(512)), ("precision", _I), ("output_split", _I),
        ("snap_dt", _D), ("diag_dt", _D), ("burst_start", _D), ("burst_end", _D), ("burst_every", _D),
        ("vec_snap_dt", _D), ("vec_iframe_interval", _I), ("vec_delta_tol", _D), ("vec_block_size", _I),
        ("n_fabrics", _I), ("mf_stage", _I), ("mf_lock_CQ", _I), ("eps_CQ", _D), ("q_C", _D),
        ("q_Q", _D), ("q_L", _D),
        ("init_sfa_L", _s(512)), ("init_sfa_Q", _s(512)), ("mf_seed_Q", _I),
        ("n_locks", _I), ("locks_file", _s(512)), ("lock0", _s(256)), ("lock1", _s(256)),
        ("lock2", _s(256)), ("lock3", _s(256)),
        ("locks_track", _s(512)), ("lock_soft_r", _D), ("lock_soft_k", _D), ("lock_bag_r", _D),
        ("lock_bag_k", _D),
        ("lock_bag_mode", _I), ("lock_bag_smooth", _I), ("locks_medium_only", _I),
        ("step_precision", _I),
    ]


def _lib():
    L = _native.load("libscp_io.so")
    P = C.POINTER(ScpConfig)
    L.scp_cfg_defaults.argtypes = [P]
    L.scp_cfg_set.argtypes = [P, C.c_char_p, C.c_char_p]
    L.scp_cfg_load.argtypes = [P, C.c_char_p]
    L.scp_cfg_validate.argtypes = [P, C.c_char_p, C.c_int]
    L.scp_cfg_unsupported.argtypes = [P, C.c_char_p, C.c_int]
    L.scp_cfg_kvmd.argtypes = [P, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), C.c_char_p]
    L.scp_f64_to_f16.restype = C.c_ushort
    L.scp_f64_to_f16.argtypes = [C.c_double]
    L.scp_f16_to_f64.restype = C.c_double
    L.scp_f16_to_f64.argtypes = [C.c_ushort]
    return L


class Config:
    def __init__(self):
        self.c = ScpConfig()
        _lib().scp_cfg_defaults(C.byref(self.c))

    def set(self, key, val):
        """Returns True when the key is known (unknown keys warn, as the reference)."""
        return _lib().scp_cfg_set(C.byref(self.c), key.encode(), str(val).encode()) == 0

    def load(self, path):
        if _lib().scp_cfg_load(C.byref(self.c), str(path).encode()) != 0:
            raise OSError(f"Cannot open config: {path}")

    def validate(self):
        buf = C.create_string_buffer(512)
        rc = _lib().scp_cfg_validate(C.byref(self.c), buf, 512)
        return None if rc == 0 else buf.value.decode()

    def unsupported(self):
        buf = C.create_string_buffer(512)
        rc = _lib().scp_cfg_unsupported(C.byref(self.c), buf, 512)
        return None if rc == 0 else buf.value.decode()

    def kvmd(self):
        keys = (C.c_char_p * 32)()
        vals = (C.c_char_p * 32)()
        store = C.create_string_buffer(64 * 32)
        n = _lib().scp_cfg_kvmd(C.byref(self.c), keys, vals, store)
        return [(keys[i].decode(), vals[i].decode()) for i in range(n)]

    def __getattr__(self, k):
        v = getattr(self.c, k)
        return v.decode() if isinstance(v, bytes) else v


def do_init(cfg, mode):
    """The initial state of an init mode (scp_do_init of libscp_io, the reference's
    do_init, scp_init.h:489-499): (12/24/30, N^3) float64 in snapshot column order."""
    import numpy as np
    L = _lib()
    L.scp_do_init.argtypes = [C.POINTER(ScpConfig), C.c_int, C.POINTER(C.c_double), C.c_char_p, C.c_int]
    ncols = {0: 12, 1: 24, 2: 30}[mode]
    st = np.zeros((ncols, cfg.N ** 3))
    err = C.create_string_buffer(512)
    if L.scp_do_init(C.byref(cfg.c), mode, st.ctypes.data_as(C.POINTER(C.c_double)), err, 512):
        raise ValueError(err.value.decode())
    return st


def f64_to_f16(x):
    return _lib().scp_f64_to_f16(float(x))


def f16_to_f64(h):
    return _lib().scp_f16_to_f64(int(h))
