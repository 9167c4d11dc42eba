"""Python view of the SFA container, backed by libscp_io.so (sfa_io.c).

Column layout of the sim snapshots (reference sfa/sim/scp_sim.c:2862-2897):
12 real columns, +12 imaginary (complex_phi), +6 gauge (complex_gauge).
"""
import ctypes as C

import struct

import numpy as np

from . import _native

F16, F32, F64 = 0, 1, 2
POSITION, ANGLE, VELOCITY = 0, 1, 2
CODEC_RAW, CODEC_ZSTD, CODEC_BSS, CODEC_COLZSTD = 0, 1, 2, 7
_NP = {F16: np.uint16, F32: np.float32, F64: np.float64}

# (name, semantic, component) in snapshot order
SIM_COLUMNS = (
    [("phi_x", POSITION, 0), ("phi_y", POSITION, 1), ("phi_z", POSITION, 2),
     ("theta_x", ANGLE, 0), ("theta_y", ANGLE, 1), ("theta_z", ANGLE, 2),
     ("phi_vx", VELOCITY, 0), ("phi_vy", VELOCITY, 1), ("phi_vz", VELOCITY, 2),
     ("theta_vx", VELOCITY, 3), ("theta_vy", VELOCITY, 4), ("theta_vz", VELOCITY, 5)]
    + [("phiim_x", POSITION, 3), ("phiim_y", POSITION, 4), ("phiim_z", POSITION, 5),
       ("thetaim_x", ANGLE, 3), ("thetaim_y", ANGLE, 4), ("thetaim_z", ANGLE, 5),
       ("phiim_vx", VELOCITY, 6), ("phiim_vy", VELOCITY, 7), ("phiim_vz", VELOCITY, 8),
       ("thetaim_vx", VELOCITY, 9), ("thetaim_vy", VELOCITY, 10), ("thetaim_vz", VELOCITY, 11)]
    + [("th_x", ANGLE, 6), ("th_y", ANGLE, 7), ("th_z", ANGLE, 8),
       ("E_x", VELOCITY, 12), ("E_y", VELOCITY, 13), ("E_z", VELOCITY, 14)])


def ncols_for_mode(mode):
    return {0: 12, 1: 24, 2: 30}[mode]


def _lib():
    lib = _native.load("libscp_io.so")
    if not getattr(lib, "_sfa_typed", False):
        lib.sfaio_create.restype = C.c_void_p
        lib.sfaio_create.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_double, C.c_double, C.c_double, C.c_double]
        lib.sfaio_set_codec.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.sfaio_add_column.argtypes = [C.c_void_p, C.c_char_p, C.c_int, C.c_int, C.c_int]
        lib.sfaio_add_kvmd.argtypes = [C.c_void_p, C.c_uint16, C.c_uint32, C.c_uint32,
                                       C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), C.c_int]
        lib.sfaio_finalize_header.argtypes = [C.c_void_p]
        lib.sfaio_write_frame.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_void_p)]
        lib.sfaio_close.argtypes = [C.c_void_p]
        lib.sfaio_open.restype = C.c_void_p
        lib.sfaio_open.argtypes = [C.c_char_p]
        lib.sfaio_read_frame.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p]
        lib.sfaio_frame_time.restype = C.c_double
        lib.sfaio_frame_time.argtypes = [C.c_void_p, C.c_uint32]
        lib.sfaio_info.argtypes = [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        lib.sfaio_column_info.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.POINTER(C.c_uint8)]
        lib.sfaio_kvmd_count.argtypes = [C.c_void_p]
        lib.sfaio_kvmd_pair.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_char_p),
                                        C.POINTER(C.c_char_p), C.POINTER(C.c_uint32)]
        lib.sfaio_fixup_index.argtypes = [C.c_char_p]
        lib.sfaio_frame_type.argtypes = [C.c_void_p, C.c_uint32]
        lib.scp_vf_init.argtypes = [C.POINTER(VecFrames), C.c_int, C.c_int, C.c_int, C.c_double]
        lib.scp_vf_write.argtypes = [C.POINTER(VecFrames), C.c_void_p, C.c_double]
        lib.scp_vf_free.argtypes = [C.POINTER(VecFrames)]
        lib.sfaio_read_vec.restype = C.c_int64
        lib.sfaio_read_vec.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64]
        lib.sfaio_crc32c.restype = C.c_uint32
        lib.sfaio_crc32c.argtypes = [C.c_void_p, C.c_size_t]
        lib._sfa_typed = True
    return lib


def crc32c(buf):
    b = np.ascontiguousarray(np.frombuffer(bytes(buf), dtype=np.uint8))
    return _lib().sfaio_crc32c(b.ctypes.data, b.size)


class VecFrames(C.Structure):
    """scp_vecframes (vecframes.h): the host side of the real-mode vector frames."""
    _fields_ = [("bs", C.c_int), ("n_patches", C.c_int), ("nc", C.c_int), ("iframe_interval", C.c_int),
                ("refit_interval", C.c_int), ("delta_tol", C.c_double), ("omega", C.c_float),
                ("frame", C.c_int), ("count", C.c_int), ("origins", C.POINTER(C.c_int16)),
                ("coeffs", C.POINTER(C.c_float)), ("mean", C.POINTER(C.c_float)), ("amp", C.POINTER(C.c_float)),
                ("phase", C.POINTER(C.c_float)), ("sum_mean", C.POINTER(C.c_float)),
                ("sum_cos", C.POINTER(C.c_float)), ("sum_sin", C.POINTER(C.c_float)),
                ("delta", C.POINTER(C.c_float)), ("idx", C.POINTER(C.c_uint32))]


def write_sim_sfa(path, N, L, dt, frames, dtype=F64, codec=CODEC_RAW, kvmd=None, ncols=None, vec=None):
    """Write frames [(time, state[ncols, N^3] float64), ...] as a sim snapshot archive.
    With vec = (block_size, iframe_interval, delta_tol), entries ("vec", time,
    coeffs[(N // bs)^3, 384] float32) are vector frames: the temporal model and
    the I/P-frame choice of scp_vf_write (libscp_io) applied to the given fits."""
    lib = _lib()
    ncols = ncols or next(f[1].shape[0] for f in frames if f[0] != "vec")
    vf = None
    if vec:
        vf = VecFrames()
        if lib.scp_vf_init(C.byref(vf), N, vec[0], vec[1], vec[2]):
            raise ValueError("scp_vf_init: bad block size or I-frame interval, or out of memory")
    s = lib.sfaio_create(str(path).encode(), N, N, N, L, L, L, dt)
    if not s:
        raise OSError(f"cannot create {path}")
    rc = lib.sfaio_set_codec(s, codec, 1)
    if rc:
        lib.sfaio_close(s)
        raise OSError(f"codec {codec} unavailable (rc={rc})")
    if kvmd:
        keys = (C.c_char_p * len(kvmd))(*[k.encode() for k, _ in kvmd])
        vals = (C.c_char_p * len(kvmd))(*[str(v).encode() for _, v in kvmd])
        lib.sfaio_add_kvmd(s, 0, 0xFFFFFFFF, 0xFFFFFFFF, keys, vals, len(kvmd))
    for name, sem, comp in SIM_COLUMNS[:ncols]:
        lib.sfaio_add_column(s, name.encode(), dtype, sem, comp)
    lib.sfaio_finalize_header(s)
    for fr in frames:
        if fr[0] == "vec":
            c = np.ascontiguousarray(fr[2], dtype=np.float32).ravel()
            if c.size != vf.n_patches * vf.nc:
                raise ValueError(f"vector frame needs {vf.n_patches} x {vf.nc} coefficients")
            C.memmove(vf.coeffs, c.ctypes.data, c.nbytes)
            if lib.scp_vf_write(C.byref(vf), s, float(fr[1])):
                lib.sfaio_close(s)
                raise OSError("vector frame write failed")
            continue
        t, st = fr
        st = np.asarray(st, dtype=np.float64).reshape(ncols, -1)
        if dtype == F64:
            cols = [np.ascontiguousarray(st[c]) for c in range(ncols)]
        elif dtype == F32:
            cols = [np.ascontiguousarray(st[c].astype(np.float32)) for c in range(ncols)]
        else:
            cols = [f64_to_f16(st[c]) for c in range(ncols)]
        ptrs = (C.c_void_p * ncols)(*[c.ctypes.data for c in cols])
        if lib.sfaio_write_frame(s, float(t), ptrs):
            lib.sfaio_close(s)
            raise OSError("frame write failed")
    if vf is not None:
        lib.scp_vf_free(C.byref(vf))
    if lib.sfaio_close(s):
        raise OSError("close failed")


class SFAFile:
    """Read-only handle: header fields, columns, KVMD sets and frames."""

    def __init__(self, path):
        self._lib = _lib()
        self._s = self._lib.sfaio_open(str(path).encode())
        if not self._s:
            raise OSError(f"cannot open SFA {path}")
        u = (C.c_uint32 * 6)()
        d = (C.c_double * 4)()
        self._lib.sfaio_info(self._s, u, d)
        self.Nx, self.Ny, self.Nz, self.n_columns, self.total_frames, self.flags = list(u)
        self.Lx, self.Ly, self.Lz, self.dt = list(d)
        self.columns = []
        for c in range(self.n_columns):
            name = C.create_string_buffer(13)
            meta = (C.c_uint8 * 4)()
            self._lib.sfaio_column_info(self._s, c, name, meta)
            self.columns.append((name.value.decode(), meta[0], meta[1], meta[2]))
        self.n_total = self.Nx * self.Ny * self.Nz

    def kvmd(self):
        out = []
        for sidx in range(self._lib.sfaio_kvmd_count(self._s)):
            k, v, ff = C.c_char_p(), C.c_char_p(), C.c_uint32()
            n = self._lib.sfaio_kvmd_pair(self._s, sidx, -1, C.byref(k), C.byref(v), C.byref(ff))
            pairs = []
            for p in range(n):
                self._lib.sfaio_kvmd_pair(self._s, sidx, p, C.byref(k), C.byref(v), C.byref(ff))
                pairs.append((k.value.decode(), v.value.decode()))
            out.append((ff.value, pairs))
        return out

    def frame_time(self, f):
        return self._lib.sfaio_frame_time(self._s, f)

    def read_frame(self, f, as_f64=True):
        nbytes = sum(self.n_total * [2, 4, 8][dt] for _, dt, _, _ in self.columns)
        buf = np.empty(nbytes, dtype=np.uint8)
        rc = self._lib.sfaio_read_frame(self._s, f, buf.ctypes.data)
        if rc:
            raise OSError(f"read_frame({f}) failed rc={rc}")
        cols, off = [], 0
        for _, dt, _, _ in self.columns:
            nb = self.n_total * [2, 4, 8][dt]
            a = buf[off:off + nb].view(_NP[dt])
            if as_f64:
                a = f16_to_f64(a) if dt == F16 else a.astype(np.float64)
            cols.append(a)
            off += nb
        return np.stack(cols) if as_f64 else cols

    def frame_type(self, f):
        """0 voxel frame (FRMD), 1 vector I-frame, 2 vector P-frame (FRVD)."""
        return self._lib.sfaio_frame_type(self._s, f)

    def read_vec(self, f):
        """A vector frame's payload, parsed (layouts of sfa.h:1643-1720)."""
        cap = 100 * self.n_total + 4096       # an I-frame with its model, patches of 4^3 or larger
        buf = np.empty(cap, dtype=np.uint8)
        n = self._lib.sfaio_read_vec(self._s, f, buf.ctypes.data, cap)
        if n < 0:
            raise OSError(f"read_vec({f}) failed rc={n}")
        b = buf[:n].tobytes()
        t = self.frame_type(f)
        out = {"type": t, "time": self.frame_time(f)}
        if t == 1:
            npch, bs, nc, flags = struct.unpack_from("<IBHB", b, 0)
            o = 8
            out.update(n_patches=npch, block_size=bs, n_coeffs=nc, flags=flags)
            out["origins"] = np.frombuffer(b, np.int16, npch * 3, o).reshape(npch, 3); o += npch * 6
            out["coeffs"] = np.frombuffer(b, np.float32, npch * nc, o).reshape(npch, nc); o += npch * nc * 4
            if flags & 1:
                out["omega"] = struct.unpack_from("<f", b, o)[0]; o += 4
                for k in ("mean", "amp", "phase"):
                    out[k] = np.frombuffer(b, np.float32, npch * nc, o).reshape(npch, nc); o += npch * nc * 4
        else:
            nd, nc = struct.unpack_from("<IH", b, 0)
            out.update(n_deltas=nd, n_coeffs=nc)
            out["indices"] = np.frombuffer(b, np.uint32, nd, 8)
            out["deltas"] = np.frombuffer(b, np.float32, nd * nc, 8 + 4 * nd).reshape(nd, nc)
        return out

    def close(self):
        if self._s:
            self._lib.sfaio_close(self._s)
            self._s = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def fixup_index(path):
    return _lib().sfaio_fixup_index(str(path).encode())


def f64_to_f16(x):
    """Truncating float->half of the reference (scp_config.h:548-558), vectorised."""
    f = np.asarray(x, dtype=np.float64).astype(np.float32).view(np.uint32)
    sign = ((f >> 16) & 0x8000).astype(np.uint16)
    e = ((f >> 23) & 0xFF).astype(np.int32) - 127 + 15
    mant = ((f >> 13) & 0x3FF).astype(np.uint16)
    out = np.where(e <= 0, sign, np.where(e >= 31, sign | 0x7C00,
                                            sign | (np.clip(e, 0, 31).astype(np.uint16) << 10) | mant))
    return np.ascontiguousarray(out.astype(np.uint16))


def f16_to_f64(h):
    h = np.asarray(h, dtype=np.uint16).astype(np.uint32)
    sign = h & 0x8000
    e = (h >> 10) & 0x1F
    mant = h & 0x3FF
    bits = (sign << 16) | ((e - 15 + 127) << 23) | (mant << 13)
    f = bits.astype(np.uint32).view(np.float32).astype(np.float64)
    f = np.where(e == 0, np.where(sign != 0, -0.0, 0.0), f)
    f = np.where(e == 31, np.where(sign != 0, -np.inf, np.inf), f)
    return f
