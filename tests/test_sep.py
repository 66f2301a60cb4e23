"""SEP files through the C-ABI (sfd_sep_*, host code: no GPU needed).

The format is SEPlib's (the reference reads and writes every array through
it: from_aux / sreed / srite, libs/libwave/IO.f90:14-51): history files
written by a chain of programs, each appending its own lines, the last
assignment of a name winning; quoted values; `in` relative to the working
directory or to the header; big-endian data by default ("xdr_float").  The
reference holds no SEP files, so the fixtures are written here in that
format; the coordinate-file grouping follows readgathercoords /
number_of_traces (ReadSouVelRho.f90:161-241, :828-890): a shot is a run of
consecutive traces with equal (sx, sy), its source keys come from its last
trace."""
import os

import numpy as np
import pytest

from seismic_amd import sep


def test_round_trip(tmp_path):
    a = np.random.default_rng(1).standard_normal((7, 5, 3)).astype(np.float32)
    p = str(tmp_path / "a.H")
    sep.write(p, a, o=(1.5, 0.0, -2.0), d=(0.004, 12.5, 1.0))
    b, h = sep.read(p)
    assert np.array_equal(a, b)
    assert h["n"][:3] == [3, 5, 7] and h["ndim"] == 3 and h["xdr"] == 0
    assert h["o"][0] == np.float32(1.5) and h["d"][1] == np.float32(12.5)


def test_history_written_by_a_chain_of_programs(tmp_path):
    d = tmp_path / "data"
    d.mkdir()
    vals = np.arange(12, dtype=np.float32).reshape(4, 3) * np.float32(0.25) - 1
    vals.astype(">f4").tofile(str(d / "v.H@"))          # big-endian, as SEPlib writes by default
    (tmp_path / "v.H").write_text(
        "\nSpike:   sergey@host   Mon Dec  1 10:00:00 2025\n"
        "\t\tsets next: in=\"/nonexistent/v.H@\"\n"
        "\tn1=5 n2=4 o1=0 d1=10 label1=\"depth (m)\"\n"
        "# a comment with n1=99 in it\n"
        "\nWindow:   sergey@host   Mon Dec  1 10:01:00 2025\n"
        "\tn1=3 esize=4 in='data/v.H@'\n")
    cwd = os.getcwd()
    os.chdir(str(tmp_path / "data"))       # `in` resolves next to the header when not found from here
    try:
        b, h = sep.read(str(tmp_path / "v.H"))
    finally:
        os.chdir(cwd)
    assert h["n"][:2] == [3, 4] and h["d"][0] == 10.0 and h["xdr"] == 1
    assert np.array_equal(b, vals)


def test_native_format_and_trailing_eot(tmp_path):
    v = np.linspace(-1, 1, 6).astype(np.float32)
    v.tofile(str(tmp_path / "w.bin"))
    (tmp_path / "w.H").write_bytes(b"\tn1=6 in=\"" + str(tmp_path / "w.bin").encode() +
                                   b"\" data_format=\"native_float\"\n\x04\x04\x04garbage n1=1")
    b, _ = sep.read(str(tmp_path / "w.H"))
    assert np.array_equal(b, v)


@pytest.mark.parametrize("text,code", [
    ("n1=4 esize=8 in=x.H@", -4),          # complex data: not supported
    ("n1=4", -8),                           # no in=
    ("n1=4 in=missing.H@", -8),             # no data file
    ("n1=4 in=short.H@", -8),               # fewer bytes than the header says
    ("n1=0 in=short.H@", -8),               # bad axis length
])
def test_bad_files_are_refused(tmp_path, text, code):
    np.zeros(2, np.float32).tofile(str(tmp_path / "short.H@"))
    (tmp_path / "b.H").write_text(text + "\n")
    with pytest.raises(sep.SepError) as e:
        sep.read(str(tmp_path / "b.H"))
    assert e.value.code == code


def test_gathers_group_consecutive_traces_by_source(tmp_path):
    # keys as the reference's test recipe (test/Makefile testgeo): sx=1 sy=2
    # selev=gelev=3 gx=4 gy=5
    rows = []
    shots = [(100.0, 5), (350.0, 3), (100.0, 4), (600.0, 1)]   # the 3rd repeats the 1st's x: a new run
    for k, (sx, ntr) in enumerate(shots):
        for j in range(ntr):
            elev = 10.0 * k + j                                    # differs per trace: the source's is the last's
            rows.append([sx, 0.0, elev, sx - 50.0 + 25.0 * j, 0.0])
    c = np.array(rows, np.float32)
    p = str(tmp_path / "coord.H")
    sep.write(p, c)
    g = sep.gathers(p, keygx=4, keygy=5, keysx=1, keysy=2, keyselev=3, keygelev=3)
    assert [len(s[2]) for s in g] == [n for _, n in shots]
    off = 0
    for (sz, sx, rz, rx), (sx0, ntr) in zip(g, shots):
        assert sx == sx0
        assert sz == c[off + ntr - 1, 2]
        assert np.array_equal(rz, c[off:off + ntr, 2]) and np.array_equal(rx, c[off:off + ntr, 3])
        off += ntr
    with pytest.raises(sep.SepError):
        sep.gathers(p, keygx=6, keygy=5, keysx=1, keysy=2, keyselev=3, keygelev=3)
