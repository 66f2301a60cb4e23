"""Out-of-core L-BFGS (sfd_lbfgsoc_step, LBFGSOC of libs/libnlinv/nlinv_lbfgs.f)
against the oracle's restatement orc_lbfgsoc, which is pinned bitwise to the
reference's compiled LBFGSOC (test_oracle_lbfgsoc.py).  The device sums its
FP64 dot products in a different order from the serial oracle, so iterates
agree to ~1e-10 rather than bitwise; the IFLAG/MYINFO sequence must be equal.
Every device call reads the optimiser state back from its file."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O

sa = pytest.importorskip("seismic_amd")
pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rosen(x):
    f = np.sum(100.0 * (x[1::2] - x[::2] ** 2) ** 2 + (1 - x[::2]) ** 2)
    g = np.zeros_like(x)
    g[::2] = -400.0 * x[::2] * (x[1::2] - x[::2] ** 2) - 2 * (1 - x[::2])
    g[1::2] = 200.0 * (x[1::2] - x[::2] ** 2)
    return f, g


def bounded_quad(n=40, seed=5):
    rng = np.random.default_rng(seed)
    Q = rng.standard_normal((n, n))
    A = Q @ Q.T / n + np.eye(n)
    xs = np.concatenate([2.0 + rng.standard_normal(n // 2), -1.0 + rng.standard_normal(n - n // 2)])

    def fg(x):
        r = x - xs
        return 0.5 * r @ A @ r, A @ r
    return fg


def drive(opt, fg, x0, nmax=200):
    x = x0.copy()
    rec = []
    for _ in range(nmax):
        f, g = fg(x)
        iflag, info = opt.step(x, f, g)
        rec.append((iflag, info, x.copy()))
        if iflag <= 0:
            break
    return rec


def compare(ro, rd, kmin):
    k = min(len(ro), len(rd), 40)
    assert k >= kmin
    for j in range(k):
        assert ro[j][:2] == rd[j][:2], "call %d: oracle %s device %s" % (j, ro[j][:2], rd[j][:2])
        x1, x2 = ro[j][2], rd[j][2]
        assert np.linalg.norm(x1 - x2) <= 1e-9 * max(1.0, np.linalg.norm(x1)), "call %d" % j
    return k


CASES = [
    ("quad", ((0.5, -2.0), (3.0, 0.3)), 1, 999999.0, 2),
    ("quad", ((0.5, -2.0), (3.0, 0.3)), 0, 999999.0, 1),
    ("quad", ((-1e30,), (1e30,)), 1, 0.05, 0),
    ("rosen", ((-1.5,), (1.5,)), 1, 999999.0, 2),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "%s-stp1opt%d-clip%d" % (c[0], c[2], c[4]))
def test_device_lbfgsoc_matches_oracle(case, tmp_path):
    prob, (xmin, xmax), stp1_opt, step_init, clip = case
    print("case: lbfgsoc", case, flush=True)
    if prob == "quad":
        fg, n = bounded_quad(), 40
        x0 = np.full(n, 1.0)
    else:
        fg, n = rosen, 20
        x0 = np.tile([-1.2, 1.0], n // 2)
    ro = drive(O.LBFGSOC(n, m=5, xmin=xmin, xmax=xmax, stp1_opt=stp1_opt, step_init=step_init, boundt=clip),
               fg, x0)
    rd = drive(sa.LBFGSOC(n, str(tmp_path / "state.bin"), m=5, xmin=xmin, xmax=xmax, stp1_opt=stp1_opt,
                          step_init=step_init, clip_type=clip), fg, x0)
    k = compare(ro, rd, 5)
    print("calls compared: %d of %d / %d, final iflag oracle %d device %d" % (k, len(ro), len(rd), ro[-1][0],
                                                                              rd[-1][0]))


def test_lbfgsoc_resume_from_file(tmp_path):
    """A run interrupted after 7 calls and resumed by a new object from the
    state file follows the uninterrupted run bitwise."""
    print("case: lbfgsoc resume", flush=True)
    fg, n = bounded_quad(), 40
    kw = dict(m=5, xmin=(0.5, -2.0), xmax=(3.0, 0.3), stp1_opt=1, clip_type=2)
    full = drive(sa.LBFGSOC(n, str(tmp_path / "a.bin"), **kw), fg, np.full(n, 1.0))
    x = np.full(n, 1.0)
    first = sa.LBFGSOC(n, str(tmp_path / "b.bin"), **kw)
    part = []
    for _ in range(7):
        f, g = fg(x)
        part.append(first.step(x, f, g) + (x.copy(),))
    again = sa.LBFGSOC(n, str(tmp_path / "b.bin"), resume=True, info=first.info.value, **kw)
    part += drive(again, fg, x)
    assert len(part) == len(full)
    for (a, b) in zip(full, part):
        assert a[:2] == b[:2] and np.array_equal(a[2], b[2])


def test_lbfgsoc_state_checks(tmp_path):
    """A missing state file, or one written for another n or other bounds,
    is SFD_EIO; IFLAG 2 (DIAGCO) is not accepted."""
    print("case: lbfgsoc state checks", flush=True)
    fg, n = bounded_quad(), 40
    x = np.full(n, 1.0)
    f, g = fg(x)
    with pytest.raises(sa.SeisFDError) as e:
        sa.LBFGSOC(n, str(tmp_path / "none.bin"), resume=True).step(x.copy(), f, g)
    assert e.value.code == -8
    lb = sa.LBFGSOC(n, str(tmp_path / "s.bin"), xmin=(0.5,), xmax=(3.0,))
    lb.step(x.copy(), f, g)
    with pytest.raises(sa.SeisFDError) as e:
        sa.LBFGSOC(n, str(tmp_path / "s.bin"), xmin=(0.4,), xmax=(3.0,), resume=True).step(x.copy(), f, g)
    assert e.value.code == -8
    x2 = np.full(n + 2, 1.0)
    with pytest.raises(sa.SeisFDError) as e:
        sa.LBFGSOC(n + 2, str(tmp_path / "s.bin"), xmin=(0.5,), xmax=(3.0,), resume=True).step(
            x2, 1.0, np.ones(n + 2))
    assert e.value.code == -8
    bad = sa.LBFGSOC(n, str(tmp_path / "t.bin"))
    bad.iflag.value = 2
    with pytest.raises(sa.SeisFDError) as e:
        bad.step(x.copy(), f, g)
    assert e.value.code == -1


def test_fortran_lbfgsoc_one_process_per_update(tmp_path):
    """progs/inversion/L_BFGS.f90's protocol: every model update is a new
    process (the Fortran example's --lbfgsoc mode) that reads the state file,
    makes one LBFGSOC call and writes the state and the next model back."""
    exe = os.path.join(ROOT, "seismic_amd", "lib", "seisfd_example")
    if not os.path.exists(exe):
        pytest.skip("Fortran example not built")
    print("case: fortran lbfgsoc processes", flush=True)
    fg, n = bounded_quad(), 40
    bmin, bmax = (0.5, -2.0), (3.0, 0.3)
    d = str(tmp_path)
    with open(os.path.join(d, "opts.txt"), "w") as fo:
        fo.write("%d 5 2 1 999999.0 2\n%r %r\n%r %r\n" % (n, bmin[0], bmin[1], bmax[0], bmax[1]))
    x = np.full(n, 1.0)
    x.tofile(os.path.join(d, "model.f64"))
    rec = []
    for call in range(12):
        f, g = fg(x)
        np.concatenate([[f], g]).tofile(os.path.join(d, "fg.f64"))
        r = subprocess.run([exe, "--lbfgsoc", d], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        iflag, info = (int(v) for v in open(os.path.join(d, "ctl.txt")).read().split())
        x = np.fromfile(os.path.join(d, "model.f64"))
        rec.append((iflag, info, x.copy()))
        if iflag <= 0:
            break
    ro = drive(O.LBFGSOC(n, m=5, xmin=bmin, xmax=bmax, stp1_opt=1, boundt=2), fg, np.full(n, 1.0))
    compare(ro, rec, len(rec))
    print("process-per-update calls: %d" % len(rec))


def test_lbfgsoc_damaged_state_file(tmp_path):
    """The state file ends with a checksum of its contents: a truncated file,
    one flipped byte in the stored vectors or in the checksum itself is
    SFD_EIO; the undamaged file resumes."""
    print("case: lbfgsoc damaged state", flush=True)
    fg, n = bounded_quad(), 40
    kw = dict(m=5, xmin=(0.5, -2.0), xmax=(3.0, 0.3), stp1_opt=1, clip_type=2)
    path = str(tmp_path / "s.bin")
    lb = sa.LBFGSOC(n, path, **kw)
    x = np.full(n, 1.0)
    for _ in range(4):
        f, g = fg(x)
        lb.step(x, f, g)
    good = open(path, "rb").read()
    flip = lambda b, k: b[:k] + bytes([b[k] ^ 0x10]) + b[k + 1:]
    for name, bad in (("truncated", good[:-16]), ("vector byte", flip(good, len(good) // 2)),
                      ("checksum byte", flip(good, len(good) - 3))):
        open(path, "wb").write(bad)
        f, g = fg(x)
        with pytest.raises(sa.SeisFDError) as e:
            sa.LBFGSOC(n, path, resume=True, info=lb.info.value, **kw).step(x.copy(), f, g)
        assert e.value.code == -8, name
    open(path, "wb").write(good)
    f, g = fg(x)
    iflag, _ = sa.LBFGSOC(n, path, resume=True, info=lb.info.value, **kw).step(x.copy(), f, g)
    assert iflag in (0, 1)
