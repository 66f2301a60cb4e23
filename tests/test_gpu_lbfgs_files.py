"""sfd_lbfgsoc_step_files: one run of progs/inversion/L_BFGS.f90 (lbfgs_type 1)
on the reference's own state files, against that program run by the
reference's LBFGSOC with its Fortran file handling (oracle/ref_lbfgs_x.f90,
pinned in test_oracle_lbfgs_files.py).  A job that alternates between the two
programs, run by run, must follow the reference alone: the same IFLAG/MYINFO
sequence and iterates within 1e-9 (the device sums its FP64 dot products in
another order).  Each file either program writes must match the other's field
by field, and the reader must take every legal list-directed form of the
scalars (D exponents, commas, repeat counts, one value per line) and 8-byte
record markers to the same result bitwise."""
import os
import shutil

import numpy as np
import pytest

import oracle_lib as O
from test_oracle_lbfgs_files import LBFGS_FIELDS, MCSRCH_FIELDS, bounded_quad, read_list, read_record, rotate

sa = pytest.importorskip("seismic_amd")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.nlref_available(), reason="reference LBFGSOC not built (oracle/_ref)")]

KW = dict(m=5, xmin=(0.5, -2.0), xmax=(3.0, 0.3), stp1_opt=1)


def rosen(x):
    f = np.sum(100.0 * (x[1::2] - x[::2] ** 2) ** 2 + (1 - x[::2]) ** 2)
    g = np.zeros_like(x)
    g[::2] = -400.0 * x[::2] * (x[1::2] - x[::2] ** 2) - 2 * (1 - x[::2])
    g[1::2] = 200.0 * (x[1::2] - x[::2] ** 2)
    return f, g


def run_ref(d, x, f, g, clip, kw):
    return O.ref_lbfgs_x(d, x, f, g, boundt=clip, **kw)


def run_dev(d, x, f, g, clip, kw):
    return sa.lbfgsoc_step_files(d, x, f, g, clip_type=clip, **kw)


def drive(d, fg, x0, who, clip, kw, fp32, nmax=60):
    """Runs the program once per evaluation, rotating the files in between;
    who(call) picks the reference or the device run."""
    x = x0.copy()
    rec = []
    for call in range(nmax):
        f, g = fg(x)
        if fp32:   # L_BFGS.f90 reads fct, gdt and mod from FP32 SEP files
            f, g = float(np.float32(f)), g.astype(np.float32).astype(np.float64)
        r = who(call)(d, x, f, g, clip, kw)
        if fp32:
            x[:] = x.astype(np.float32)
        rotate(d)
        rec.append(r + (x.copy(),))
        if r[0] <= 0:
            break
    return rec


PROBLEMS = {
    "quad": (bounded_quad, 40, lambda n: np.full(n, 1.0), KW, 2),
    "quad_clip1": (bounded_quad, 40, lambda n: np.full(n, 1.0), dict(KW, stp1_opt=0), 1),
    "rosen": (lambda: rosen, 20, lambda n: np.tile([-1.2, 1.0], n // 2), dict(m=5, xmin=(-1.5,), xmax=(1.5,),
                                                                            stp1_opt=1), 2),
}


@pytest.mark.parametrize("start", ["device", "reference"])
@pytest.mark.parametrize("prob,fp32", [("quad", True), ("quad_clip1", False), ("rosen", False)])
def test_alternating_with_reference_program(prob, fp32, start, tmp_path):
    mk, n, x0f, kw, clip = PROBLEMS[prob]
    print("case: L_BFGS.x files, %s fp32=%s, first run on the %s" % (prob, fp32, start), flush=True)
    fg, x0 = mk(), x0f(n)
    (tmp_path / "ref").mkdir()
    (tmp_path / "mix").mkdir()
    ref = drive(str(tmp_path / "ref"), fg, x0, lambda c: run_ref, clip, kw, fp32)
    first, second = (run_dev, run_ref) if start == "device" else (run_ref, run_dev)
    mix = drive(str(tmp_path / "mix"), fg, x0, lambda c: first if c % 2 == 0 else second, clip, kw, fp32)
    k = min(len(ref), len(mix), 40)
    assert k >= 10
    for j in range(k):
        assert ref[j][:2] == mix[j][:2], "run %d: reference %s, alternating %s" % (j, ref[j][:2], mix[j][:2])
        x1, x2 = ref[j][2], mix[j][2]
        assert np.linalg.norm(x1 - x2) <= 1e-9 * max(1.0, np.linalg.norm(x1)), "run %d" % j
    print("runs compared: %d (reference alone %d, alternating %d), final iflag %d / %d" %
          (k, len(ref), len(mix), ref[-1][0], mix[-1][0]), flush=True)


# Scratch variables the files carry although every use assigns them first.
# Until their first assignment in a run they hold whatever the process left in
# the SAVE'd locals (the reference program starts with static zeros, the
# harness process with leftovers of earlier runs), so they are compared only
# from the point on where both runs have assigned them: the two-loop's values
# from ITER 2, DG and FTEST1 after the first evaluation of a line search.  The
# modified-function values of MCSRCHOC (FM ... DGYM) are assigned only in one
# branch of stage 1 and are not compared.
AFTER_ITER1 = {"YS", "YY", "SQ", "YR", "BETA", "XNORM", "NPT", "CP", "INMC", "IYCN", "ISCN"}
AFTER_EVAL1 = {"DG", "FTEST1"}
NEVER = {"FM", "FXM", "FYM", "DGM", "DGXM", "DGYM"}


def close(a, b, what, skip=()):
    for key in a:
        if key in skip:
            continue
        va, vb = a[key], b[key]
        if isinstance(va, float):
            assert abs(va - vb) <= 1e-9 * max(1.0, abs(va), abs(vb)), "%s %s: %r vs %r" % (what, key, va, vb)
        else:
            assert va == vb, "%s %s: %r vs %r" % (what, key, va, vb)


def test_device_files_match_reference_files(tmp_path):
    """From the same start, 14 device runs and 14 reference runs write the
    same LbfgsDatOut / McsrchDatOut fields (integers and logicals equal,
    reals within 1e-9; scratch values once assigned), W, DIAG, IFLAG and
    MYINFO."""
    print("case: L_BFGS.x files, field by field", flush=True)
    fg, n = bounded_quad(), 40
    dd, dr = str(tmp_path / "dev"), str(tmp_path / "ref")
    os.mkdir(dd)
    os.mkdir(dr)
    xd, xr = np.full(n, 1.0), np.full(n, 1.0)
    for call in range(14):
        fd_, gd = fg(xd)
        fr, gr = fg(xr)
        a = run_dev(dd, xd, fd_, gd, 2, KW)
        b = run_ref(dr, xr, fr, gr, 2, KW)
        assert a == b, "run %d: %s vs %s" % (call, a, b)
        # a run that ends in an error writes neither scalar file (or only the
        # one of a completed iteration): the same files must exist
        for name in ("LbfgsDatOut.dat", "McsrchDatOut.dat"):
            assert os.path.exists(os.path.join(dd, name)) == os.path.exists(os.path.join(dr, name)), name
        if os.path.exists(os.path.join(dr, "LbfgsDatOut.dat")):
            la = read_list(os.path.join(dd, "LbfgsDatOut.dat"), LBFGS_FIELDS)
            lb = read_list(os.path.join(dr, "LbfgsDatOut.dat"), LBFGS_FIELDS)
            assert la["ITER"] == lb["ITER"]
            close(la, lb, "run %d LbfgsDatOut" % call, skip=AFTER_ITER1 if lb["ITER"] < 2 else ())
        if os.path.exists(os.path.join(dr, "McsrchDatOut.dat")):
            close(read_list(os.path.join(dd, "McsrchDatOut.dat"), MCSRCH_FIELDS),
                  read_list(os.path.join(dr, "McsrchDatOut.dat"), MCSRCH_FIELDS), "run %d McsrchDatOut" % call,
                  skip=NEVER | (AFTER_EVAL1 if call == 0 else set()))
        for name in ("WorkArrayOut.dat", "DiagOut.dat"):
            u, v = read_record(os.path.join(dd, name)), read_record(os.path.join(dr, name))
            assert u.size == v.size and np.linalg.norm(u - v) <= 1e-9 * max(1.0, np.linalg.norm(v)), name
        for name in ("IflagOut.dat", "InfoOut.dat"):
            assert open(os.path.join(dd, name)).read() == open(os.path.join(dr, name)).read()
        rotate(dd)
        rotate(dr)
        if a[0] <= 0:
            break
    assert call >= 10


def rewrite_list(path, fields, style):
    """Writes the values of a state file again in another legal
    list-directed form."""
    vals = read_list(path, fields)
    recs, cur = [], []
    for f in fields:
        if f is None:
            recs.append(cur)
            cur = []
            continue
        name, kind = f
        v = vals[name]
        if kind == "r":
            t = repr(float(v))
            if style == "dexp":
                t = "%.17E" % v
                t = t.replace("E", "D")
            cur.append(t)
        elif kind == "i":
            cur.append(str(v))
        else:
            cur.append(".TRUE." if v else ".FALSE.")
    lines = []
    for r in recs:
        if style == "commas":
            lines.append(", ".join(r) + ",  trailing text the READ skips")
        elif style == "split":
            lines += r
        elif style == "repeat":
            out, k = [], 0
            while k < len(r):
                j = k
                while j + 1 < len(r) and r[j + 1] == r[k]:
                    j += 1
                out.append(r[k] if j == k else "%d*%s" % (j - k + 1, r[k]))
                k = j + 1
            lines.append(" ".join(out))
        else:
            lines.append(" " + " ".join(r))
    open(path, "w").write("\n".join(lines) + "\n")


def to_8byte_markers(path):
    b = open(path, "rb").read()
    n = int(np.frombuffer(b[:4], "<i4")[0])
    m8 = np.array([n], "<i8").tobytes()
    open(path, "wb").write(m8 + b[4:-4] + m8)


def prepared_state(tmp_path, runs=7):
    """The reference program's files after `runs` runs, with the next
    evaluation (x, f, g)."""
    fg, n = bounded_quad(), 40
    d = str(tmp_path / "base")
    os.mkdir(d)
    x = np.full(n, 1.0)
    for _ in range(runs):
        f, g = fg(x)
        run_ref(d, x, f, g, 2, KW)
        rotate(d)
    f, g = fg(x)
    return d, x, f, g


@pytest.mark.parametrize("style", ["dexp", "commas", "split", "repeat", "marker8"])
def test_reader_takes_every_legal_form(style, tmp_path):
    print("case: L_BFGS.x files, %s form" % style, flush=True)
    base, x, f, g = prepared_state(tmp_path)
    alt = str(tmp_path / "alt")
    shutil.copytree(base, alt)
    if style == "marker8":
        to_8byte_markers(os.path.join(alt, "WorkArray.dat"))
        to_8byte_markers(os.path.join(alt, "Diag.dat"))
    else:
        rewrite_list(os.path.join(alt, "LbfgsDat.dat"), LBFGS_FIELDS, style)
        rewrite_list(os.path.join(alt, "McsrchDat.dat"), MCSRCH_FIELDS, style)
    xa, xb = x.copy(), x.copy()
    ra = run_dev(base, xa, f, g, 2, KW)
    rb = run_dev(alt, xb, f, g, 2, KW)
    assert ra == rb and np.array_equal(xa, xb)
    for name in ("LbfgsDatOut.dat", "McsrchDatOut.dat"):
        if os.path.exists(os.path.join(base, name)):
            assert open(os.path.join(base, name)).read() == open(os.path.join(alt, name)).read()


def expect_error(d, x, f, g, code):
    with pytest.raises(sa.SeisFDError) as e:
        run_dev(d, x.copy(), f, g, 2, KW)
    assert e.value.code == code, e.value


def test_bad_state_files_are_refused(tmp_path):
    """IFLAG 2 (DIAGCO re-entry) is SFD_EINVAL; a missing, short or damaged
    record, and scalars written for another N (ISPT) or with ONE not 1, are
    SFD_EIO; nothing is written then."""
    print("case: L_BFGS.x files, refused states", flush=True)
    base, x, f, g = prepared_state(tmp_path)

    def variant(tag):
        d = str(tmp_path / tag)
        shutil.copytree(base, d)
        return d

    d = variant("iflag2")
    open(os.path.join(d, "Iflag.dat"), "w").write(" 2\n")
    expect_error(d, x, f, g, -1)
    assert not os.path.exists(os.path.join(d, "IflagOut.dat"))
    d = variant("nowork")
    os.remove(os.path.join(d, "WorkArray.dat"))
    expect_error(d, x, f, g, -8)
    d = variant("short")
    w = read_record(os.path.join(d, "WorkArray.dat"))[:-3]
    open(os.path.join(d, "WorkArray.dat"), "wb").write(
        np.array([w.nbytes], "<i4").tobytes() + w.tobytes() + np.array([w.nbytes], "<i4").tobytes())
    expect_error(d, x, f, g, -8)
    d = variant("marker")
    b = bytearray(open(os.path.join(d, "Diag.dat"), "rb").read())
    b[-1] ^= 1
    open(os.path.join(d, "Diag.dat"), "wb").write(bytes(b))
    expect_error(d, x, f, g, -8)
    for tag, key, val in (("ispt", "ISPT", 51), ("one", "ONE", 2.0)):
        d = variant(tag)
        p = os.path.join(d, "LbfgsDat.dat")
        vals = read_list(p, LBFGS_FIELDS)
        vals[key] = val
        recs, cur = [], []
        for fl in LBFGS_FIELDS:
            if fl is None:
                recs.append(" " + " ".join(cur))
                cur = []
                continue
            v = vals[fl[0]]
            cur.append(("T" if v else "F") if fl[1] == "l" else repr(v))
        open(p, "w").write("\n".join(recs) + "\n")
        expect_error(d, x, f, g, -8)
