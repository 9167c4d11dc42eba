"""CPU tests of the host side of the boundary: SFA container, config keys,
KVMD, f16 casts, and the exported C-ABI of the HIP library (no GPU calls)."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

from scp_amd import config, sfa
from scp_amd import grid as G
from tests import fixtures

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SFA = os.path.join(ROOT, "tests", "golden", "ref_complex_n8.sfa")


def test_crc32c_known_answers():
    # CRC-32C check value (RFC 3720 / iSCSI): "123456789" -> 0xE3069283
    assert sfa.crc32c(b"123456789") == 0xE3069283
    assert sfa.crc32c(b"") == 0
    assert sfa.crc32c(b"\x00" * 32) == 0x8A9136AA          # RFC 3720 B.4 vector
    assert sfa.crc32c(b"\xff" * 32) == 0x62A8AB43


@pytest.mark.parametrize("codec", [sfa.CODEC_RAW, sfa.CODEC_ZSTD, sfa.CODEC_BSS, sfa.CODEC_COLZSTD])
@pytest.mark.parametrize("dtype", [sfa.F64, sfa.F32, sfa.F16])
def test_sfa_roundtrip(tmp_path, codec, dtype):
    """sfa_test.c round trip (sfa/format/sfa_test.c): write frames, read back."""
    N, nf = 6, 5
    rng = np.random.default_rng(3)
    frames = [(t * 10.0, rng.standard_normal((12, N ** 3))) for t in range(nf)]
    p = tmp_path / "t.sfa"
    kv = [("N", "6"), ("L", "2.500000")]
    sfa.write_sim_sfa(p, N, 2.5, 0.05, frames, dtype=dtype, codec=codec, kvmd=kv)
    with sfa.SFAFile(p) as f:
        assert (f.Nx, f.Ny, f.Nz, f.n_columns, f.total_frames) == (N, N, N, 12, nf)
        assert f.flags == codec                          # streaming flag cleared on close
        assert f.kvmd() == [(0xFFFFFFFF, kv)]
        assert [c[0] for c in f.columns] == [c[0] for c in sfa.SIM_COLUMNS[:12]]
        for i, (t, st) in enumerate(frames):
            assert f.frame_time(i) == t
            got = f.read_frame(i)
            if dtype == sfa.F64:
                assert np.array_equal(got, st)
            elif dtype == sfa.F32:
                assert np.array_equal(got, st.astype(np.float32).astype(np.float64))
            else:
                assert np.array_equal(got, sfa.f16_to_f64(sfa.f64_to_f16(st)))


def test_sfa_many_frames_second_jump_table(tmp_path):
    """More than 1024 frames opens a second JMPF table (sfa.h:872-905)."""
    N = 4
    frames = [(0.5 * t, np.full((12, N ** 3), float(t))) for t in range(1030)]
    p = tmp_path / "m.sfa"
    sfa.write_sim_sfa(p, N, 1.5, 0.5, frames)
    with sfa.SFAFile(p) as f:
        assert f.total_frames == 1030
        for i in (0, 1023, 1024, 1029):
            assert f.frame_time(i) == 0.5 * i
            assert np.all(f.read_frame(i) == i)


def test_reference_written_archive_reads():
    """An archive written by the reference binary (COLZSTD f64) decodes with a valid CRC."""
    fx = fixtures.load("complex_n8")
    with sfa.SFAFile(REF_SFA) as f:
        assert f.flags == sfa.CODEC_COLZSTD
        assert f.n_columns == 24 and f.total_frames == 2
        assert np.array_equal(f.read_frame(0), fx["frame0"])
        assert np.array_equal(f.read_frame(1), fx["final"])
        (ff, pairs), = f.kvmd()
        assert ff == 0xFFFFFFFF and len(pairs) == 27


def test_header_bytes_match_reference_writer(tmp_path):
    """SFAH, CDEF, KVMD and the JTOP/JMPF skeleton are byte-identical to what
    the reference writer produced for the same grid, columns and KVMD."""
    with open(REF_SFA, "rb") as fh:
        ref = fh.read()
    with sfa.SFAFile(REF_SFA) as f:
        (ff, pairs), = f.kvmd()
        frames = [(f.frame_time(i), f.read_frame(i)) for i in range(f.total_frames)]
        N, L, dt = f.Nx, f.Lx, f.dt
    p = tmp_path / "mine.sfa"
    sfa.write_sim_sfa(p, N, L, dt, frames, codec=sfa.CODEC_COLZSTD, kvmd=pairs)
    mine = open(p, "rb").read()
    jtop = struct.unpack_from("<Q", ref, 72)[0]
    assert mine[:jtop] == ref[:jtop]
    # JTOP + JMPF: identical except compressed sizes/offsets when zstd versions differ
    jm = struct.unpack_from("<Q", ref, jtop + 28)[0]
    assert struct.unpack_from("<Q", mine, jtop + 28)[0] == jm
    for i in range(2):
        o = jm + 20 + 32 * i
        t_r, off_r, sz_r, crc_r, ty_r = struct.unpack_from("<dQQII", ref, o)
        t_m, off_m, sz_m, crc_m, ty_m = struct.unpack_from("<dQQII", mine, o)
        assert (t_r, crc_r, ty_r) == (t_m, crc_m, ty_m)


def test_fixup_index_recovers_unindexed_frames(tmp_path):
    """sfa_fixup_index (sfa.h:2425-2631): zeroed JMPF slots and a truncated tail."""
    N = 5
    rng = np.random.default_rng(1)
    frames = [(float(t), rng.standard_normal((12, N ** 3))) for t in range(4)]
    p = tmp_path / "s.sfa"
    sfa.write_sim_sfa(p, N, 2.0, 0.1, frames, codec=sfa.CODEC_COLZSTD)
    b = bytearray(open(p, "rb").read())
    jtop = struct.unpack_from("<Q", b, 72)[0]
    jm = struct.unpack_from("<Q", b, jtop + 28)[0]
    # a writer that died: streaming flag set, total_frames 0, index slots 1..3 empty, tail cut
    struct.pack_into("<I", b, 16, struct.unpack_from("<I", b, 16)[0] | 0x20)
    struct.pack_into("<I", b, 68, 0)
    for i in range(1, 4):
        b[jm + 20 + 32 * i + 8: jm + 20 + 32 * i + 32] = bytes(24)
    b = b[:-100]
    open(p, "wb").write(bytes(b))
    assert sfa.fixup_index(p) == 3
    with sfa.SFAFile(p) as f:
        assert f.total_frames == 3 and f.flags == sfa.CODEC_COLZSTD
        for i in range(3):
            assert np.array_equal(f.read_frame(i), frames[i][1])


def test_corrupted_frame_reports_checksum_error(tmp_path):
    N = 4
    p = tmp_path / "c.sfa"
    sfa.write_sim_sfa(p, N, 1.5, 0.1, [(0.0, np.ones((12, N ** 3)))], codec=sfa.CODEC_RAW)
    b = bytearray(open(p, "rb").read())
    b[-5] ^= 0xFF
    open(p, "wb").write(bytes(b))
    with sfa.SFAFile(p) as f:
        with pytest.raises(OSError, match="rc=-2"):
            f.read_frame(0)


def test_config_defaults_and_keys(tmp_path):
    c = config.Config()
    assert (c.N, c.L, c.T, c.dt_factor) == (128, 10.0, 200.0, 0.025)
    assert (c.m2, c.eta, c.mu, c.kappa, c.g_gauge) == (2.25, 0.5, -41.345, 50.0, 0.05)
    assert c.precision == 1 and c.init == "oscillon" and c.output == "output.sfa"
    cfg = tmp_path / "a.cfg"
    cfg.write_text("# comment\n  N = 32   \nm=2.0 # inline comment\n\tprecision = f64\n"
                   "delta = 0.1,0.2,0.3\ninit = sfa\nno_equals_line\nmystery_key = 3\n")
    c.load(cfg)
    assert c.N == 32 and c.m2 == 4.0 and c.precision == 2 and c.init == "sfa"
    assert list(c.delta) == [0.1, 0.2, 0.3]
    assert not c.set("mystery_key", "1")
    c.set("precision", "f99")                            # ignored, as the reference
    assert c.precision == 2


@pytest.mark.parametrize("keys,msg", [
    ({"complex_phi": 1, "alpha_cs": 0.1}, "complex_phi=1 is incompatible with alpha_cs"),
    ({"complex_phi": 1, "init": "braid"}, "braid init not defined for complex mode"),
    ({"complex_gauge": 1}, "complex_gauge=1 requires complex_phi=1"),
    ({"complex_phi": 1, "complex_gauge": 1, "bc_type": 1}, "incompatible with bc_type=1"),
    ({"complex_phi": 1, "complex_gauge": 1, "g_gauge": -1}, "requires g_gauge >= 0"),
    ({"init": "qball"}, "init=qball requires complex_phi=1"),
    ({"n_fabrics": 2}, "n_fabrics must be 1 or 3"),
])
def test_config_validation_errors(keys, msg):
    c = config.Config()
    for k, v in keys.items():
        c.set(k, v)
    err = c.validate()
    assert err is not None and msg in err


@pytest.mark.parametrize("name", ["real_n8", "complex_n8", "gauged_n8"])
def test_kvmd_matches_reference(name):
    """KVMD set 0 written by the reference (scp_init.h:505-556) for the fixture's config."""
    fx = fixtures.load(name)
    c = config.Config()
    for k, v in fx["params"].items():
        if k not in ("init",):
            c.set(k, v)
    (ff, pairs), = fx["kvmd"]
    assert ff == 0xFFFFFFFF
    assert [tuple(p) for p in pairs] == c.kvmd()


def test_f16_cast_matches_reference_semantics():
    # truncating mantissa (not rounding), flush below 2^-14, saturate to inf
    for x, h in [(1.0, 0x3C00), (-2.0, 0xC000), (1.0009765625, 0x3C01), (1.0014, 0x3C01),
                 (2.0 ** -14, 0x0400), (2.0 ** -15, 0x0000), (70000.0, 0x7C00), (1e-8, 0x0000)]:
        assert config.f64_to_f16(x) == h
        assert sfa.f64_to_f16(np.array([x]))[0] == h
    assert config.f16_to_f64(0x3C01) == 1.0009765625
    assert config.f16_to_f64(0x7C00) == float("inf")


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "scp_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(scp_[a-z_0-9]+)\s*\(", txt)))


def test_abi_library_exports_every_declared_symbol():
    lib = os.path.join(ROOT, "scp_amd", "lib", "libscp_amd.so")
    assert os.path.exists(lib), "libscp_amd.so not built"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (scp_\w+)", out))
    declared = _declared_symbols()
    assert declared and set(declared) <= exported, set(declared) - exported
    assert set(declared) == set(G.ABI_SYMBOLS)


def test_abi_library_loads_and_refuses_bad_params():
    """Loadable without a GPU; argument errors come back as status codes, not exits."""
    p = G.make_params(8, 2, 3.5, 0.1)
    p.bc_type = 1
    assert G.grid_bytes(p) == -1                         # gauged: pinning is gauge-dependent
    assert "incompatible" in G.last_error()
    q = G.make_params(8, 1, 3.5, 0.1, bc_type=1, gradient_margin=2)
    assert G.grid_bytes(q) > G.grid_bytes(G.make_params(8, 1, 3.5, 0.1))   # pins and edge rows
    q.precision = G.PREC_F32C
    assert G.grid_bytes(q) == -1 and "FP64" in G.last_error()
    q = G.make_params(8, 1, 3.5, 0.1, gradient_margin=2)
    assert G.grid_bytes(q) == -1 and "bc_type=1 only" in G.last_error()
    q = G.make_params(8, 1, 3.5, 0.1, sigma_cubic=0.5)          # real-mode term on a complex grid
    assert G.grid_bytes(q) == -1 and "real-mode terms" in G.last_error()
    q = G.make_params(8, 0, 3.5, 0.1, sigma_freq=0.5)           # frequency term: a 7-field pass-1 array
    assert G.grid_bytes(q) == G.grid_bytes(G.make_params(8, 0, 3.5, 0.1)) + 7 * 512 * 8
    p.bc_type = 2
    assert G.grid_bytes(p) > 0
    p.N = 3
    assert G.grid_bytes(p) == -1
    assert "out of range" in G.last_error()


def test_config_runs_the_gpu_step_real_terms():
    """The real-mode keys of the reference's GPU step are run, not refused; theta_vev
    and gamma_conf are parsed and never used by the reference, so they are accepted
    and change nothing; the terms stay refused with complex_phi, as there."""
    c = config.Config()
    for k, v in (("theta_sat", "0.1"), ("gamma_conv", "0.5"), ("sigma_grad", "0.2"), ("chi_chiral", "1"),
                 ("sigma_cubic", "0.3"), ("sigma_freq", "0.1"), ("sigma_cross", "0.2"), ("lambda_self", "0.4"),
                 ("theta_vev", "0.5"), ("gamma_conf", "0.2")):
        c.set(k, v)
    assert c.validate() is None and c.unsupported() is None
    c.set("complex_phi", "1")
    assert "incompatible with sigma_grad" in c.validate()


def test_vector_frames_in_the_container(tmp_path):
    """FRVD chunks (sfa.h:1563-1720): indexed with their I/P type, read back as
    payloads, refused as voxel data, and re-indexed by the streaming fixup (type
    I where the slot was lost, time from the chunk, as sfa_fixup_index does)."""
    N, bs = 8, 4
    rng = np.random.default_rng(3)
    st = rng.standard_normal((12, N ** 3))
    fits = [rng.standard_normal((8, 384)).astype(np.float32) for _ in range(3)]
    frames = [(0.0, st), ("vec", 0.5, fits[0]), ("vec", 1.0, fits[1]), ("vec", 1.5, fits[2])]
    p = tmp_path / "v.sfa"
    sfa.write_sim_sfa(p, N, 2.0, 0.1, frames, codec=sfa.CODEC_COLZSTD, vec=(bs, 2, 0.0))
    with sfa.SFAFile(p) as f:
        assert [f.frame_type(i) for i in range(4)] == [0, 1, 2, 1]
        assert np.array_equal(f.read_frame(0), st)
        with pytest.raises(OSError, match="rc=-4"):
            f.read_frame(1)
        v = f.read_vec(1)
        assert v["n_patches"] == 8 and v["block_size"] == bs and v["n_coeffs"] == 384 and v["flags"] == 1
        assert np.array_equal(v["coeffs"], fits[0]) and np.array_equal(v["mean"], fits[0])
        assert np.array_equal(v["origins"][:, 2], [0, 4] * 4)
        w = f.read_vec(2)                      # tolerance 0: every patch differs from the bootstrap model
        assert w["n_deltas"] == 8 and np.allclose(w["deltas"], fits[1] - fits[0], atol=1e-6)
        assert [f.frame_time(i) for i in range(4)] == [0.0, 0.5, 1.0, 1.5]
    b = bytearray(open(p, "rb").read())
    jtop = struct.unpack_from("<Q", b, 72)[0]
    jm = struct.unpack_from("<Q", b, jtop + 28)[0]
    struct.pack_into("<I", b, 16, struct.unpack_from("<I", b, 16)[0] | 0x20)
    struct.pack_into("<I", b, 68, 0)
    for i in range(1, 4):
        b[jm + 20 + 32 * i: jm + 20 + 32 * i + 32] = bytes(32)
    open(p, "wb").write(bytes(b))
    assert sfa.fixup_index(p) == 4
    with sfa.SFAFile(p) as f:
        assert [f.frame_type(i) for i in range(4)] == [0, 1, 1, 1]
        assert [f.frame_time(i) for i in range(4)] == [0.0, 0.5, 1.0, 1.5]
        assert np.array_equal(f.read_vec(3)["coeffs"], fits[2])


def test_config_vector_frames():
    """vec_snap_dt runs in the real mode; complex_phi refuses it (scp_config.h:458);
    this build needs 4 <= vec_block_size <= 16 and vec_iframe_interval >= 1."""
    c = config.Config()
    c.set("N", "32")
    c.set("vec_snap_dt", "1.0")
    assert c.validate() is None and c.unsupported() is None
    c.set("vec_block_size", "3")
    assert "vec_block_size=3" in c.unsupported()
    c.set("vec_block_size", "0")               # 0: the default 8
    assert c.unsupported() is None
    c.set("vec_iframe_interval", "0")
    assert "vec_iframe_interval" in c.unsupported()
    c.set("vec_iframe_interval", "5")
    c.set("complex_phi", "1")
    assert "incompatible with vec_snap_dt" in c.validate()
