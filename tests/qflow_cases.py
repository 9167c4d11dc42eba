"""Shared cases of the fixed-charge flow tests: a synthetic radial profile (a
table in the reference's (r f) format, uniform spacing from r = 0) and the
reference program's command line and printed rows (sfa/seed/eta_qflow.c)."""
import os
import re
import subprocess

import numpy as np

FLOW = dict(eta=0.5, m=1.5, mth=1.6, mu=-41.345, kappa=50.0)


def write_profile(path, amp=0.9, width=2.2, rmax=12.0, dr=0.02):
    r = np.arange(0.0, rmax + 0.5 * dr, dr)
    with open(path, "w") as fp:
        fp.write("# synthetic Q-ball profile: r f\n")
        for a, b in zip(r, amp * np.exp(-r ** 2 / (2 * width ** 2))):
            fp.write(f"{a:.6f} {b:.12e}\n")


def argv(binary, N, L, profile, omega0, out, iters, printev, thinit="bvp", wind=0, wtor=2.0, twist=0, Q=0.0,
         dtau=0.0, fixJ=None, Jtarget=0.0):
    f = FLOW
    cmd = [binary, str(N), repr(L), str(profile), repr(omega0), repr(f["eta"]), repr(f["m"]), repr(f["mth"]),
           repr(f["mu"]), repr(f["kappa"]), str(out), repr(Q), str(iters), repr(dtau), thinit, str(printev),
           str(wind), repr(wtor), str(twist)]
    if fixJ is not None:
        cmd += [str(fixJ), repr(Jtarget)]
    return cmd


ROW = re.compile(r"^\s+(\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s*$")
ROW_J = re.compile(r"^\s+(\d+)" + r"\s+(\S+)" * 8 + r"\s*$")


def run(cmd, threads=1, cwd=None):
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, timeout=600,
                         env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
    assert res.returncode == 0, res.stdout[-1000:] + res.stderr[-1000:]
    pat = ROW_J if "# fixJ:" in res.stdout else ROW
    rows = [m.groups() for m in map(pat.match, res.stdout.splitlines()) if m]
    return res.stdout, rows


def fmt_row(r):
    """A row as the reference prints it (eta_qflow.c:370)."""
    it, om, eq, mf, sm, nn = r
    return f"  {int(it):<8d} {om:<10.6f} {eq:<12.6e} {mf:<12.4e} {sm:<12.4e} {nn:<12.4f}".split()


def fmt_row_j(r):
    """A fixed-(Q, J) row as the reference prints it (eta_qflow.c:356-357)."""
    it, om, Om, eq, mf, sm, nn, ja, cd = r
    return (f"  {int(it):<8d} {om:<10.6f} {Om:<10.6f} {eq:<12.6e} {mf:<12.4e} {sm:<12.4e} {nn:<12.4f} "
            f"{ja:<10.4f} {cd:<10.3e}").split()


def output_columns_j(st, omega, Omega, jhat, wmask):
    """The 24 float32 columns of the fixed-(Q, J) seed (eta_qflow.c:447-460): the
    velocities carry -Omega wm Jhat of each field; wmask is the N x N table."""
    n3 = st.shape[1]
    N = wmask.shape[0]
    wm = np.repeat(wmask.reshape(-1), N)                 # site (i, j, k) -> wmask[i, j]
    assert wm.size == n3
    u, v, tu, tv = st[0:3], st[3:6], st[6:9], st[9:12]
    ju, jv, jtu, jtv = jhat[0:3], jhat[3:6], jhat[6:9], jhat[9:12]
    cols = np.zeros((24, n3), dtype=np.float32)
    for a in range(3):
        cols[a] = u[a]
        cols[12 + a] = v[a]
        cols[6 + a] = -omega * v[a] - Omega * (wm * ju[a])
        cols[18 + a] = omega * u[a] - Omega * (wm * jv[a])
        cols[3 + a] = tu[a]
        cols[15 + a] = tv[a]
        cols[9 + a] = -omega * tv[a] - Omega * (wm * jtu[a])
        cols[21 + a] = omega * tu[a] - Omega * (wm * jtv[a])
    return cols


def output_columns(st, omega):
    """The 24 float32 columns eta_qflow writes (gen_qball_boost order, eta_qflow.c:445-456)."""
    n3 = st.shape[1]
    u, v, tu, tv = st[0:3], st[3:6], st[6:9], st[9:12]
    cols = np.zeros((24, n3), dtype=np.float32)
    for a in range(3):
        cols[a] = u[a]
        cols[12 + a] = v[a]
        cols[6 + a] = -omega * v[a]
        cols[18 + a] = omega * u[a]
        cols[3 + a] = tu[a]
        cols[15 + a] = tv[a]
        cols[9 + a] = -omega * tv[a]
        cols[21 + a] = omega * tu[a]
    return cols
