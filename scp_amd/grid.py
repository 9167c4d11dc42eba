"""Host-side mirror of the reference step interface over libscp_amd.so.

The reference drives its step through `verlet_step(Grid*, const Config*)`
(sfa/sim/scp_sim.c:1583) and the GPU path through gpu_alloc / gpu_upload /
gpu_verlet_step / run_gpu_diagnostics (sfa/sim/scp_sim.cu:2659-4041). `Grid`
offers the same life cycle over the C-ABI of include/scp_amd.h:

    g = Grid(params)                # grid_alloc + gpu_alloc
    g.upload(state)                 # gpu_upload + initial compute_forces_*
    g.step(n)                       # n x verlet_step (boundary stage included)
    state = g.download()            # snapshot columns (sfa_snap)
    d = g.diag()                    # one diag.tsv row (compute_energy*, charges, gauss)

Errors raise ScpError with the library's message (the reference exits the
process; a library must not). There is no CPU fallback: without the HIP
library or a GPU, construction raises.
"""
import ctypes as C

import numpy as np

from . import _native

MODE_REAL, MODE_COMPLEX, MODE_GAUGED = 0, 1, 2
PREC_F64, PREC_F32C = 0, 1

DIAG_NAMES = ["E_phi_kin", "E_theta_kin", "E_grad", "E_mass", "E_pot", "E_tgrad", "E_tmass",
              "E_coupling", "E_total", "phi_max", "P_max", "P_int", "theta_rms",
              "Q_phi", "Q_theta", "Q_total", "s_max", "r_core", "omega_core", "Q_core",
              "thp1u", "thp1v", "thp2u", "thp2v", "gauss_max", "gauss_l2", "E_em", "Q_flux",
              "Q_p0", "Q_p1", "Q_p2"]

# every entry point declared in include/scp_amd.h
ABI_SYMBOLS = ["scp_abi_version", "scp_last_error", "scp_grid_bytes", "scp_grid_create",
               "scp_grid_destroy", "scp_grid_upload", "scp_grid_download", "scp_grid_step",
               "scp_grid_diag", "scp_grid_step_diag", "scp_grid_forces", "scp_grid_gauss_project",
               "scp_grid_set_g_offset", "scp_grid_g_offset", "scp_grid_rescale_phi",
               "scp_grid_steps_done", "scp_grid_time_steps", "scp_grid_step_kernel_name",
               "scp_batch_create", "scp_batch_destroy", "scp_batch_add", "scp_batch_remove",
               "scp_batch_size", "scp_batch_step", "scp_batch_step_all", "scp_batch_time_steps",
               "scp_batch_tune", "scp_batch_set_config", "scp_batch_config", "scp_grid_seed_balls",
               "scp_server_create", "scp_server_destroy", "scp_server_submit", "scp_server_round",
               "scp_server_job", "scp_server_take", "scp_server_stats_get", "scp_server_now",
               "scp_grid_vec_fit", "scp_grid_set_kappa", "scp_grid_set_damp_rate",
               "scp_grid_step_diag_noalloc", "scp_qflow_create", "scp_qflow_destroy", "scp_qflow_upload",
               "scp_qflow_download", "scp_qflow_norm", "scp_qflow_run", "scp_qflow_run_fixj",
               "scp_qflow_download_jhat", "scp_server_set_aging",
               "scp_batch_set_tile", "scp_batch_tile"]


class ScpError(RuntimeError):
    pass


class Params(C.Structure):
    """scp_params of include/scp_amd.h."""
    _fields_ = [("N", C.c_int32), ("mode", C.c_int32), ("precision", C.c_int32),
                ("bc_type", C.c_int32), ("rmode", C.c_int32), ("lo_words", C.c_int32),
                ("step_kernel", C.c_int32), ("gradient_margin", C.c_int32),
                ("L", C.c_double), ("dx", C.c_double), ("dt", C.c_double),
                ("m2", C.c_double), ("mtheta2", C.c_double), ("eta", C.c_double),
                ("mu", C.c_double), ("kappa", C.c_double), ("g", C.c_double),
                ("kappa_h", C.c_double), ("alpha_cs", C.c_double), ("beta_h", C.c_double),
                ("inv_alpha", C.c_double), ("inv_beta", C.c_double), ("kappa_gamma", C.c_double),
                ("damp_width", C.c_double), ("damp_rate", C.c_double), ("G_offset", C.c_double),
                ("qdiag_radius", C.c_double), ("qdiag_probe1", C.c_double),
                ("qdiag_probe2", C.c_double),
                ("theta_sat", C.c_double), ("gamma_conv", C.c_double), ("sigma_grad", C.c_double),
                ("chi_chiral", C.c_double), ("sigma_cubic", C.c_double), ("sigma_freq", C.c_double),
                ("sigma_cross", C.c_double), ("lambda_self", C.c_double)]


class QBall(C.Structure):
    """scp_qball of include/scp_amd.h."""
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("z0", C.c_double), ("amp", C.c_double),
                ("sigma", C.c_double), ("omega", C.c_double), ("phase", C.c_double),
                ("winding", C.c_int32), ("reserved", C.c_int32)]


def qballs(balls):
    """ctypes array of QBall from dicts with x0,y0,z0,amp,sigma,omega,phase,winding."""
    arr = (QBall * max(1, len(balls)))()
    for i, b in enumerate(balls):
        arr[i] = QBall(**{k: b[k] for k in ("x0", "y0", "z0", "amp", "sigma", "omega", "phase", "winding")})
    return arr


class Job(C.Structure):
    _fields_ = [("params", Params), ("balls", C.POINTER(QBall)), ("nballs", C.c_int32), ("nsteps", C.c_int32),
                ("theta_sigma", C.c_double), ("link_sigma", C.c_double), ("seed", C.c_uint64),
                ("gauss_project", C.c_int32), ("reserved", C.c_int32)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("nsteps", C.c_int32), ("steps_done", C.c_int64), ("bytes", C.c_int64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("queued", C.c_int64), ("active", C.c_int64), ("done", C.c_int64), ("failed", C.c_int64),
                ("bytes_in_use", C.c_int64), ("site_updates", C.c_int64), ("busy_seconds", C.c_double),
                ("now", C.c_double)]


JOB_FAILED, JOB_QUEUED, JOB_ACTIVE, JOB_DONE = -1, 0, 1, 2


class QflowParams(C.Structure):
    """scp_qflow_params of include/scp_amd.h."""
    _fields_ = [("N", C.c_int32), ("reserved", C.c_int32), ("L", C.c_double), ("eta", C.c_double),
                ("m", C.c_double), ("mth", C.c_double), ("mu", C.c_double), ("kappa", C.c_double)]


class QflowRow(C.Structure):
    _fields_ = [("iter", C.c_int32), ("reserved", C.c_int32), ("omega", C.c_double), ("E_Q", C.c_double),
                ("maxF", C.c_double), ("s_max", C.c_double), ("N", C.c_double)]


class QflowRowJ(C.Structure):
    _fields_ = [("iter", C.c_int32), ("reserved", C.c_int32), ("omega", C.c_double), ("Omega", C.c_double),
                ("E_QJ", C.c_double), ("maxF", C.c_double), ("s_max", C.c_double), ("N", C.c_double),
                ("J_act", C.c_double), ("cond", C.c_double)]


def make_params(N, mode, L, dt_factor, precision=PREC_F64, **kw):
    """Parameters with the reference defaults (scp_config.h:165-215); dx and dt
    are derived exactly as grid_alloc does (scp_sim.c:204-205)."""
    dx = 2.0 * L / (N - 1)
    p = Params(N=N, mode=mode, precision=precision, bc_type=kw.pop("bc_type", 2),
               rmode=kw.pop("rmode", 0), L=L, dx=dx, dt=dt_factor * dx, m2=2.25, mtheta2=0.0,
               eta=0.5, mu=-41.345, kappa=50.0, g=0.05, kappa_h=0.0, alpha_cs=0.0, beta_h=0.0,
               inv_alpha=2.25, inv_beta=5.0, kappa_gamma=2.0, damp_width=3.0, damp_rate=0.01,
               G_offset=0.0, qdiag_radius=8.0, qdiag_probe1=8.0, qdiag_probe2=16.0)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def ncols(mode):
    return {0: 12, 1: 24, 2: 30}[mode]


def nacc(mode):
    return {0: 6, 1: 12, 2: 15}[mode]


_typed = False


def lib():
    global _typed
    L = _native.load("libscp_amd.so")
    if not _typed:
        P = C.POINTER(Params)
        VP = C.c_void_p
        DPP = C.POINTER(C.c_void_p)
        L.scp_last_error.argtypes = [C.c_char_p, C.c_int]
        L.scp_grid_bytes.restype = C.c_int64
        L.scp_grid_bytes.argtypes = [P]
        L.scp_grid_create.argtypes = [P, C.POINTER(VP)]
        L.scp_grid_destroy.argtypes = [VP]
        L.scp_grid_upload.argtypes = [VP, DPP, C.c_int]
        L.scp_grid_download.argtypes = [VP, DPP, C.c_int]
        L.scp_grid_step.argtypes = [VP, C.c_int]
        L.scp_grid_diag.argtypes = [VP, C.POINTER(C.c_double), C.c_int]
        L.scp_grid_step_diag.argtypes = [VP, C.c_int, C.POINTER(C.c_double), C.c_int]
        L.scp_grid_step_diag_noalloc.argtypes = [VP, C.c_int, C.POINTER(C.c_double), C.c_int]
        L.scp_grid_forces.argtypes = [VP, DPP, C.c_int]
        L.scp_grid_vec_fit.argtypes = [VP, C.c_int, C.POINTER(C.c_float), C.c_int64]
        L.scp_grid_gauss_project.argtypes = [VP, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.scp_grid_set_g_offset.argtypes = [VP, C.c_double]
        L.scp_grid_g_offset.restype = C.c_double
        L.scp_grid_g_offset.argtypes = [VP]
        L.scp_grid_rescale_phi.argtypes = [VP, C.c_double]
        L.scp_grid_set_kappa.argtypes = [VP, C.c_double]
        L.scp_grid_set_damp_rate.argtypes = [VP, C.c_double]
        L.scp_grid_steps_done.restype = C.c_int64
        L.scp_grid_steps_done.argtypes = [VP]
        L.scp_grid_time_steps.argtypes = [VP, C.c_int, C.POINTER(C.c_float)]
        L.scp_grid_step_kernel_name.restype = C.c_char_p
        L.scp_grid_step_kernel_name.argtypes = [VP]
        IP = C.POINTER(C.c_int)
        L.scp_batch_create.argtypes = [C.POINTER(VP)]
        L.scp_batch_destroy.argtypes = [VP]
        L.scp_batch_add.argtypes = [VP, VP]
        L.scp_batch_remove.argtypes = [VP, VP]
        L.scp_batch_size.argtypes = [VP]
        L.scp_batch_step.argtypes = [VP, IP, C.c_int]
        L.scp_batch_step_all.argtypes = [VP, C.c_int]
        L.scp_batch_time_steps.argtypes = [VP, C.c_int, C.POINTER(C.c_float)]
        L.scp_batch_tune.argtypes = [VP, IP]
        L.scp_batch_set_config.argtypes = [VP, C.c_int, C.c_int]
        L.scp_batch_config.argtypes = [VP, IP, IP, IP]
        L.scp_batch_set_tile.argtypes = [VP, C.c_int, C.c_int]
        L.scp_batch_tile.argtypes = [VP, IP, IP]
        L.scp_grid_seed_balls.argtypes = [VP, C.POINTER(QBall), C.c_int, C.c_double, C.c_double, C.c_uint64]
        L.scp_server_create.argtypes = [C.c_int64, C.c_int, C.c_int, C.POINTER(VP)]
        L.scp_server_destroy.argtypes = [VP]
        L.scp_server_submit.argtypes = [VP, C.POINTER(Job), C.POINTER(C.c_int64)]
        L.scp_server_round.argtypes = [VP, IP]
        L.scp_server_job.argtypes = [VP, C.c_int64, C.POINTER(JobInfo)]
        L.scp_server_take.argtypes = [VP, C.c_int64, DPP, C.c_int, C.POINTER(C.c_double), C.c_int]
        L.scp_server_stats_get.argtypes = [VP, C.POINTER(ServerStats)]
        L.scp_server_now.restype = C.c_double
        L.scp_server_now.argtypes = [VP]
        L.scp_server_set_aging.argtypes = [VP, C.c_int]
        L.scp_qflow_create.argtypes = [C.POINTER(QflowParams), C.POINTER(VP)]
        L.scp_qflow_destroy.argtypes = [VP]
        L.scp_qflow_upload.argtypes = [VP, DPP, C.c_int]
        L.scp_qflow_download.argtypes = [VP, DPP, C.c_int]
        L.scp_qflow_norm.argtypes = [VP, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.scp_qflow_run.argtypes = [VP, C.c_double, C.c_double, C.c_int, C.c_int, C.POINTER(QflowRow), C.c_int,
                                    IP, C.POINTER(C.c_double)]
        L.scp_qflow_run_fixj.argtypes = [VP, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                         C.POINTER(QflowRowJ), C.c_int, IP, C.POINTER(C.c_double)]
        L.scp_qflow_download_jhat.argtypes = [VP, DPP, C.c_int]
        _typed = True
    return L


def last_error():
    buf = C.create_string_buffer(512)
    lib().scp_last_error(buf, 512)
    return buf.value.decode()


def _check(rc, what):
    if rc != 0:
        raise ScpError(f"{what}: rc={rc}: {last_error()}")


def grid_bytes(params):
    return int(lib().scp_grid_bytes(C.byref(params)))


def _col_ptrs(arr2d):
    return (C.c_void_p * arr2d.shape[0])(*[arr2d[c].ctypes.data for c in range(arr2d.shape[0])])


class Grid:
    def __init__(self, params):
        self.params = params
        self.mode = params.mode
        self.n3 = params.N ** 3
        self._h = C.c_void_p()
        _check(lib().scp_grid_create(C.byref(params), C.byref(self._h)), "scp_grid_create")

    def close(self):
        if self._h:
            lib().scp_grid_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, state):
        st = np.ascontiguousarray(np.asarray(state, dtype=np.float64).reshape(ncols(self.mode), self.n3))
        _check(lib().scp_grid_upload(self._h, _col_ptrs(st), st.shape[0]), "scp_grid_upload")

    def seed_balls(self, balls, theta_sigma=0.01, link_sigma=0.0, seed=0):
        """Seed on the device (scp_grid_seed_balls); balls: list of dicts."""
        arr = qballs(balls)
        _check(lib().scp_grid_seed_balls(self._h, arr, len(balls), float(theta_sigma), float(link_sigma),
                                         int(seed)), "scp_grid_seed_balls")

    def download(self):
        st = np.empty((ncols(self.mode), self.n3))
        _check(lib().scp_grid_download(self._h, _col_ptrs(st), st.shape[0]), "scp_grid_download")
        return st

    def step(self, n=1):
        _check(lib().scp_grid_step(self._h, int(n)), "scp_grid_step")

    def time_steps(self, n):
        ms = C.c_float()
        _check(lib().scp_grid_time_steps(self._h, int(n), C.byref(ms)), "scp_grid_time_steps")
        return ms.value

    def diag(self):
        out = (C.c_double * len(DIAG_NAMES))()
        _check(lib().scp_grid_diag(self._h, out, len(DIAG_NAMES)), "scp_grid_diag")
        return dict(zip(DIAG_NAMES, list(out)))

    def step_diag(self, n=1):
        """n steps, then the diagnostics of the new state, fused into the last
        step launch where the grid runs the march (see scp_grid_step_diag)."""
        out = (C.c_double * len(DIAG_NAMES))()
        _check(lib().scp_grid_step_diag(self._h, int(n), out, len(DIAG_NAMES)), "scp_grid_step_diag")
        return dict(zip(DIAG_NAMES, list(out)))

    def step_diag_noalloc(self, n=1):
        """step_diag with its work space in the library's static device scratch
        (scp_grid_step_diag_noalloc: no device allocation)."""
        out = (C.c_double * len(DIAG_NAMES))()
        _check(lib().scp_grid_step_diag_noalloc(self._h, int(n), out, len(DIAG_NAMES)),
               "scp_grid_step_diag_noalloc")
        return dict(zip(DIAG_NAMES, list(out)))

    def vec_fit(self, block_size=8):
        """Vector-frame patch fits (real mode): ((N // bs)^3, 6, 64) float32."""
        bn = self.params.N // block_size
        out = np.empty((bn ** 3, 6, 64), dtype=np.float32)
        _check(lib().scp_grid_vec_fit(self._h, int(block_size), out.ctypes.data_as(C.POINTER(C.c_float)), out.size),
               "scp_grid_vec_fit")
        return out

    def forces(self):
        acc = np.empty((nacc(self.mode), self.n3))
        _check(lib().scp_grid_forces(self._h, _col_ptrs(acc), acc.shape[0]), "scp_grid_forces")
        return acc

    def gauss_project(self):
        it, gm = C.c_int(), C.c_double()
        _check(lib().scp_grid_gauss_project(self._h, C.byref(it), C.byref(gm)), "scp_grid_gauss_project")
        return it.value, gm.value

    @property
    def g_offset(self):
        return lib().scp_grid_g_offset(self._h)

    @g_offset.setter
    def g_offset(self, v):
        _check(lib().scp_grid_set_g_offset(self._h, float(v)), "scp_grid_set_g_offset")

    def rescale_phi(self, f):
        _check(lib().scp_grid_rescale_phi(self._h, float(f)), "scp_grid_rescale_phi")

    def set_kappa(self, kappa):
        _check(lib().scp_grid_set_kappa(self._h, float(kappa)), "scp_grid_set_kappa")

    def set_damp_rate(self, rate):
        _check(lib().scp_grid_set_damp_rate(self._h, float(rate)), "scp_grid_set_damp_rate")

    @property
    def steps_done(self):
        return int(lib().scp_grid_steps_done(self._h))

    @property
    def step_kernel_name(self):
        return lib().scp_grid_step_kernel_name(self._h).decode()


class Batch:
    """Many grids of one class advanced together (scp_batch_* of include/scp_amd.h).
    Holds references to its member Grid objects (a closed grid leaves the batch)."""

    def __init__(self, grids=()):
        self._h = C.c_void_p()
        self.members = []
        _check(lib().scp_batch_create(C.byref(self._h)), "scp_batch_create")
        for g in grids:
            self.add(g)

    def add(self, grid):
        _check(lib().scp_batch_add(self._h, grid._h), "scp_batch_add")
        self.members.append(grid)

    def remove(self, grid):
        _check(lib().scp_batch_remove(self._h, grid._h), "scp_batch_remove")
        self.members.remove(grid)

    def __len__(self):
        return len(self.members)

    def step(self, nsteps):
        """nsteps: int (every member) or a sequence, one count per member."""
        if isinstance(nsteps, int):
            _check(lib().scp_batch_step_all(self._h, nsteps), "scp_batch_step_all")
        else:
            v = (C.c_int * len(nsteps))(*[int(n) for n in nsteps])
            _check(lib().scp_batch_step(self._h, v, len(nsteps)), "scp_batch_step")

    def time_steps(self, n):
        ms = C.c_float()
        _check(lib().scp_batch_time_steps(self._h, int(n), C.byref(ms)), "scp_batch_time_steps")
        return ms.value

    def tune(self):
        """Autotune depth and launch form on the batch's own steps; returns steps taken."""
        n = C.c_int()
        _check(lib().scp_batch_tune(self._h, C.byref(n)), "scp_batch_tune")
        return n.value

    def set_config(self, di=0, persistent=False):
        _check(lib().scp_batch_set_config(self._h, int(di), 1 if persistent else 0), "scp_batch_set_config")

    def set_tile(self, tk, tj):
        """Tile of the batched launch: tk (k) x tj (j) sites (scp_batch_set_tile)."""
        _check(lib().scp_batch_set_tile(self._h, int(tk), int(tj)), "scp_batch_set_tile")

    @property
    def config(self):
        di, per, tuned = C.c_int(), C.c_int(), C.c_int()
        _check(lib().scp_batch_config(self._h, C.byref(di), C.byref(per), C.byref(tuned)), "scp_batch_config")
        tk, tj = C.c_int(), C.c_int()
        _check(lib().scp_batch_tile(self._h, C.byref(tk), C.byref(tj)), "scp_batch_tile")
        return {"di": di.value, "persistent": bool(per.value), "tuned": bool(tuned.value),
                "tile": (tk.value, tj.value)}

    def close(self):
        if self._h:
            lib().scp_batch_destroy(self._h)
            self._h = C.c_void_p()
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """Request loop over device-resident grids (scp_server_* of include/scp_amd.h)."""

    def __init__(self, hbm_budget, quantum=200, max_active=0):
        self._h = C.c_void_p()
        _check(lib().scp_server_create(int(hbm_budget), int(quantum), int(max_active), C.byref(self._h)),
               "scp_server_create")

    def submit(self, params, balls, nsteps, theta_sigma=0.01, link_sigma=0.0, seed=0, gauss_project=True):
        arr = qballs(balls)
        job = Job(params=params, balls=C.cast(arr, C.POINTER(QBall)), nballs=len(balls), nsteps=int(nsteps),
                  theta_sigma=float(theta_sigma), link_sigma=float(link_sigma), seed=int(seed),
                  gauss_project=1 if gauss_project else 0, reserved=0)
        jid = C.c_int64()
        _check(lib().scp_server_submit(self._h, C.byref(job), C.byref(jid)), "scp_server_submit")
        return jid.value

    def round(self):
        n = C.c_int()
        _check(lib().scp_server_round(self._h, C.byref(n)), "scp_server_round")
        return n.value

    def job(self, jid):
        info = JobInfo()
        _check(lib().scp_server_job(self._h, int(jid), C.byref(info)), "scp_server_job")
        return {k: getattr(info, k) for k, _ in JobInfo._fields_}

    def take(self, jid, mode=MODE_GAUGED, n3=None, state=True):
        """(state or None, diag dict) of a finished job; frees its device memory."""
        out = None
        cols = None
        if state:
            out = np.empty((ncols(mode), n3))
            cols = _col_ptrs(out)
        d = (C.c_double * len(DIAG_NAMES))()
        _check(lib().scp_server_take(self._h, int(jid), cols, out.shape[0] if state else 0, d, len(DIAG_NAMES)),
               "scp_server_take")
        return out, dict(zip(DIAG_NAMES, list(d)))

    def stats(self):
        st = ServerStats()
        _check(lib().scp_server_stats_get(self._h, C.byref(st)), "scp_server_stats_get")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def set_aging(self, rounds):
        _check(lib().scp_server_set_aging(self._h, int(rounds)), "scp_server_set_aging")

    def now(self):
        return lib().scp_server_now(self._h)

    def close(self):
        if self._h:
            lib().scp_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class QFlow:
    """The fixed-charge gradient flow on the device (scp_qflow_* of include/scp_amd.h):
    the relaxer sfa/seed/eta_qflow.c in its fixed-Q and fixed-(Q, J) modes. State: (12, N^3) rows
    u0..2, v0..2, tu0..2, tv0..2."""

    def __init__(self, N, L, eta=0.5, m=1.5, mth=0.0, mu=-41.345, kappa=50.0):
        self.N = N
        self._h = C.c_void_p()
        p = QflowParams(N=N, reserved=0, L=L, eta=eta, m=m, mth=mth, mu=mu, kappa=kappa)
        _check(lib().scp_qflow_create(C.byref(p), C.byref(self._h)), "scp_qflow_create")

    def close(self):
        if self._h:
            lib().scp_qflow_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, state):
        st = np.ascontiguousarray(np.asarray(state, dtype=np.float64).reshape(12, self.N ** 3))
        _check(lib().scp_qflow_upload(self._h, _col_ptrs(st), 12), "scp_qflow_upload")

    def download(self):
        st = np.empty((12, self.N ** 3))
        _check(lib().scp_qflow_download(self._h, _col_ptrs(st), 12), "scp_qflow_download")
        return st

    def norm(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().scp_qflow_norm(self._h, C.byref(a), C.byref(b)), "scp_qflow_norm")
        return a.value, b.value

    def run(self, Q, dtau, iters, printev):
        """Returns (rows [(it, omega, E_Q, maxF, s_max, N)], omega of the last iteration)."""
        maxrows = iters // max(1, printev) + 2
        rows = (QflowRow * maxrows)()
        n, om = C.c_int(), C.c_double()
        _check(lib().scp_qflow_run(self._h, float(Q), float(dtau), int(iters), int(printev), rows, maxrows,
                                   C.byref(n), C.byref(om)), "scp_qflow_run")
        return [(r.iter, r.omega, r.E_Q, r.maxF, r.s_max, r.N) for r in rows[:n.value]], om.value

    def run_fixj(self, Q, J, dtau, iters, printev):
        """The fixed-(Q, J) mode: returns (rows [(it, omega, Omega, E_QJ, maxF, s_max, N,
        J_act, cond)], (omega, Omega, cond) of the last iteration + (ell, K) of the final state)."""
        maxrows = iters // max(1, printev) + 2
        rows = (QflowRowJ * maxrows)()
        n, last = C.c_int(), (C.c_double * 5)()
        _check(lib().scp_qflow_run_fixj(self._h, float(Q), float(J), float(dtau), int(iters), int(printev), rows,
                                        maxrows, C.byref(n), last), "scp_qflow_run_fixj")
        return ([(r.iter, r.omega, r.Omega, r.E_QJ, r.maxF, r.s_max, r.N, r.J_act, r.cond) for r in rows[:n.value]],
                tuple(last))

    def download_jhat(self):
        jh = np.empty((12, self.N ** 3))
        _check(lib().scp_qflow_download_jhat(self._h, _col_ptrs(jh), 12), "scp_qflow_download_jhat")
        return jh
