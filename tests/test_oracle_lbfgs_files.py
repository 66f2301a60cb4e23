"""The file forms of progs/inversion/L_BFGS.f90's out-of-core protocol, as the
reference's own LBFGSOC and the program's Fortran I/O write them
(oracle/ref_lbfgs_x.f90 built with flang beside libs/libnlinv/nlinv_lbfgs.f).
These pin what sfd_lbfgsoc_step_files reads and writes:

- WorkArrayOut.dat and DiagOut.dat are one sequential unformatted record each,
  framed by 4-byte little-endian byte counts (8 (n(2m+1)+2m) and 8n bytes);
- IflagOut.dat and InfoOut.dat hold one `(I2)` record;
- LbfgsDatOut.dat holds LBFGSOC's 33 scalars in 5 list-directed records
  (7 reals, 8 reals, 9 integers, 8 integers, 1 logical) and McsrchDatOut.dat
  MCSRCHOC's 25 (2 integers, 2 logicals, 8, 9 and 4 reals); the runtime wraps
  long records onto continuation lines, so a reader must let one READ
  statement run over several lines;
- ISPT = N+2M, IYPT = ISPT+NM, NPT = POINT*N and the loop indexes I = J = N+1
  at every write; INMC, IYCN and ISCN are 1-based W indexes.
A run of the program on its files follows the reference's LBFGSOC with W and
DIAG in memory bitwise, so the files carry the whole state exactly.
"""
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.skipif(not O.nlref_available(), reason="reference LBFGSOC not built (oracle/_ref)")

NAMES = ("Iflag", "Info", "Diag", "WorkArray", "LbfgsDat", "McsrchDat")
LBFGS_FIELDS = [("GTOL", "r"), ("ONE", "r"), ("ZERO", "r"), ("GNORM", "r"), ("STP1", "r"), ("FTOL", "r"),
                ("STPMIN", "r"), None,
                ("STPMAX", "r"), ("STP", "r"), ("YS", "r"), ("YY", "r"), ("SQ", "r"), ("YR", "r"), ("BETA", "r"),
                ("XNORM", "r"), None,
                ("MP", "i"), ("LP", "i"), ("ITER", "i"), ("NFUN", "i"), ("POINT", "i"), ("ISPT", "i"),
                ("IYPT", "i"), ("MAXFEV", "i"), ("INFO", "i"), None,
                ("BOUND", "i"), ("NPT", "i"), ("CP", "i"), ("I", "i"), ("NFEV", "i"), ("INMC", "i"), ("IYCN", "i"),
                ("ISCN", "i"), None,
                ("FINISH", "l"), None]
MCSRCH_FIELDS = [("INFOC", "i"), ("J", "i"), None, ("BRACKT", "l"), ("STAGE1", "l"), None,
                 ("DG", "r"), ("DGM", "r"), ("DGINIT", "r"), ("DGTEST", "r"), ("DGX", "r"), ("DGXM", "r"),
                 ("DGY", "r"), ("DGYM", "r"), None,
                 ("FINIT", "r"), ("FTEST1", "r"), ("FM", "r"), ("FX", "r"), ("FXM", "r"), ("FY", "r"), ("FYM", "r"),
                 ("STX", "r"), ("STY", "r"), None,
                 ("STMIN", "r"), ("STMAX", "r"), ("WIDTH", "r"), ("WIDTH1", "r"), None]


def read_list(path, fields):
    """Reads list-directed records as the READ statements of the reference do:
    each statement (ended by None in `fields`) takes its items from the next
    line on, continuing over lines, and skips the rest of its last line."""
    lines = open(path).read().split("\n")
    out, li, stmt = {}, 0, []
    for f in fields:
        if f is not None:
            stmt.append(f)
            continue
        toks = []
        while len(toks) < len(stmt):
            toks += lines[li].replace(",", " ").split()
            li += 1
        for (name, kind), t in zip(stmt, toks):
            if kind == "r":
                out[name] = float(t.upper().replace("D", "E"))
            elif kind == "i":
                out[name] = int(t)
            else:
                out[name] = t.lstrip(".")[0].upper() == "T"
        stmt = []
    return out


def read_record(path):
    """One sequential unformatted record of doubles with 4-byte markers."""
    b = open(path, "rb").read()
    head = int(np.frombuffer(b[:4], "<i4")[0])
    tail = int(np.frombuffer(b[-4:], "<i4")[0])
    assert head == tail == len(b) - 8, (head, tail, len(b))
    return np.frombuffer(b[4:-4], "<f8")


def rotate(d):
    for name in NAMES:
        src = os.path.join(d, name + "Out.dat")
        if os.path.exists(src):
            os.replace(src, os.path.join(d, name + ".dat"))


def bounded_quad(n=40, seed=5):
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T / n + np.eye(n)
    xs = np.concatenate([2.0 + rng.standard_normal(n // 2), -1.0 + rng.standard_normal(n - n // 2)])

    def fg(x):
        r = x - xs
        return 0.5 * r @ A @ r, A @ r
    return fg


KW = dict(m=5, xmin=(0.5, -2.0), xmax=(3.0, 0.3), stp1_opt=1)


def test_reference_program_file_forms(tmp_path):
    fg, n, m = bounded_quad(), 40, 5
    d = str(tmp_path)
    x = np.full(n, 1.0)
    for call in range(12):
        f, g = fg(x)
        iflag, info = O.ref_lbfgs_x(d, x, f, g, boundt=2, **KW)
        w = read_record(os.path.join(d, "WorkArrayOut.dat"))
        assert w.size == n * (2 * m + 1) + 2 * m
        assert read_record(os.path.join(d, "DiagOut.dat")).size == n
        assert open(os.path.join(d, "IflagOut.dat")).read() == "%2d\n" % iflag
        assert open(os.path.join(d, "InfoOut.dat")).read() == "%2d\n" % info
        lb = read_list(os.path.join(d, "LbfgsDatOut.dat"), LBFGS_FIELDS)
        assert (lb["ONE"], lb["ZERO"], lb["ISPT"], lb["IYPT"], lb["I"]) == (1.0, 0.0, n + 2 * m, n + 2 * m + n * m,
                                                                         n + 1)
        assert lb["NPT"] % n == 0 and 0 <= lb["POINT"] < m and (lb["MP"], lb["LP"]) == (1020, 0)
        assert lb["INFO"] == -1 and iflag == 1
        # the scratch indexes: W(INMC) is alpha(cp), W(IYCN+1:) Y(cp), W(ISCN+1:) S(cp)
        if lb["ITER"] > 1:
            assert n + m + 1 <= lb["INMC"] <= n + 2 * m
            assert (lb["IYCN"] - (n + 2 * m + n * m)) % n == 0 and (lb["ISCN"] - (n + 2 * m)) % n == 0
        mc = read_list(os.path.join(d, "McsrchDatOut.dat"), MCSRCH_FIELDS)
        assert mc["J"] == n + 1 and 0 <= mc["INFOC"] <= 5   # MCSTEP's INFO
        rotate(d)
    raw = open(os.path.join(d, "LbfgsDat.dat")).read().split("\n")
    assert len([ln for ln in raw if ln.strip()]) > 5, "records were expected to wrap onto continuation lines"


def test_reference_program_follows_in_memory_lbfgsoc(tmp_path):
    """The program's files (text scalars, unformatted W and DIAG) carry the
    whole state: run after run it follows one in-memory LBFGSOC bitwise."""
    fg, n = bounded_quad(), 40
    a, b = tmp_path / "prog", tmp_path / "mem"
    a.mkdir()
    b.mkdir()
    ref = O.RefLBFGSOC(n, str(b), boundt=2, **KW)
    xa, xb = np.full(n, 1.0), np.full(n, 1.0)
    for call in range(60):
        fa, ga = fg(xa)
        fb, gb = fg(xb)
        ra = O.ref_lbfgs_x(str(a), xa, fa, ga, boundt=2, **KW)
        rotate(str(a))
        rb = ref.step(xb, fb, gb)
        assert ra == rb and np.array_equal(xa, xb), "call %d" % call
        if ra[0] <= 0:
            break
    assert call >= 10
