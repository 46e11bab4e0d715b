"""The library's parallel xyz reader (mdpd_read_xyz) against the reference's
format and abort cases (io_util::ReadXYZForm, io_util.hpp:96-129;
FPDPD::readAscii, ptcl_class.hpp:60-65).  Host code only: runs without a GPU.

Values are pinned by C's own conversions: a file written with %.15g and read
back must give the records Python's correctly rounded float() gives for the
same text (fpdpd.read_xyz), bit for bit, and the columns of %.17g text must
round-trip exactly."""
import numpy as np
import pytest

from implicit_dpd_amd import capi
from implicit_dpd_amd.capi import MdpdError
from implicit_dpd_amd.engine import read_xyz as read_xyz_native
from implicit_dpd_amd.fpdpd import ATOM_TYPE, FPDPD_DTYPE, read_xyz, write_xyz


def records(n, seed=3):
    g = np.random.default_rng(seed)
    r = np.zeros(n, FPDPD_DTYPE)
    r["id"] = g.permutation(n)
    r["prop"] = g.integers(0, 4, n)
    r["amp_id"] = g.integers(0, 2**32 - 1, n, dtype=np.uint64)
    r["unit"] = g.integers(0, 4, n)
    for k in ("pos", "vel", "vel_buf", "acc"):
        r[k] = g.standard_normal((n, 3)) * 10.0 ** g.integers(-300, 300, (n, 3))
    special = [[0.0, -0.0, 1e-320], [5e-324, 1.7976931348623157e308, -2.5e-308],
               [1.0, 0.1, 3.0], [np.pi, -np.e, 1e15]]
    if n:
        r["pos"][:4] = special[: min(4, n)]
    return r


def same(a, b):
    for k in ("id", "prop", "amp_id", "unit", "pos", "vel", "vel_buf", "acc", "delta_sumr",
              "press", "nei_cm_pos", "density"):
        assert np.ascontiguousarray(a[k]).tobytes() == np.ascontiguousarray(b[k]).tobytes(), k


@pytest.mark.parametrize("n", [0, 1, 7, 5000, 60000])
def test_native_reader_matches_python_reader(tmp_path, n):
    rec = records(n)
    f = tmp_path / "c.xyz"
    write_xyz(f, rec, 4242)
    a, ta = read_xyz(f)
    b, tb = read_xyz_native(f)
    assert ta == tb == 4242 and len(b) == n
    same(a, b)


def test_seventeen_digit_columns_round_trip(tmp_path):
    rec = records(3000, seed=9)
    f = tmp_path / "c.xyz"
    with open(f, "w") as fp:
        fp.write(f"{len(rec)}\ntime 0\n")
        for r in rec:
            cols = [ATOM_TYPE[int(r["prop"])]] + ["%.17g" % x for x in r["pos"]]
            cols += ["%u" % r[k] for k in ("id", "prop", "amp_id", "unit")]
            cols += ["%.17g" % x for k in ("vel", "vel_buf", "acc") for x in r[k]]
            fp.write("\t ".join(cols) + "\r\n")
    b, _ = read_xyz_native(f)
    for k in ("pos", "vel", "vel_buf", "acc", "id", "prop", "amp_id", "unit"):
        assert np.ascontiguousarray(rec[k]).tobytes() == np.ascontiguousarray(b[k]).tobytes(), k


def test_parallel_chunks_keep_file_order(tmp_path):
    # > 1 MiB per chunk boundary: several threads, lines cut at chunk edges
    rec = records(40000, seed=5)
    rec["id"] = np.arange(len(rec))
    f = tmp_path / "c.xyz"
    write_xyz(f, rec, 1)
    with open(f, "a") as fp:
        fp.write("\n  \n")  # trailing blank lines are no records
    assert f.stat().st_size > 8 << 20
    b, _ = read_xyz_native(f)
    assert np.array_equal(b["id"], np.arange(len(rec)))
    same(read_xyz(f)[0], b)


def _text(tmp_path, body):
    f = tmp_path / "bad.xyz"
    f.write_text(body)
    return f


LINE = "O 1 2 3 0 0 0 0 0 0 0 0 0 0 0 0 0\n"


@pytest.mark.parametrize("body,msg", [
    ("2\nstep 5\n" + LINE * 2, "may be old"),                 # io_util.hpp:105-113
    ("2\n" + LINE * 2, "may be old"),                           # no time line
    ("3\ntime 5\n" + LINE * 2, "# of lines"),                   # fewer lines than the header
    ("1\ntime 5\n" + LINE * 2, "# of lines"),                   # more (io_util.hpp:124-127)
    ("2\ntime 5\n" + LINE + "O 1 2 3 0 0 0 0 0 0 0 0 0 0 0 0\n", "17 columns"),  # 16 columns
    ("1\ntime 5\nO 1 2 x 0 0 0 0 0 0 0 0 0 0 0 0 0\n", "17 columns"),
])
def test_reference_abort_cases_are_errors(tmp_path, body, msg):
    f = _text(tmp_path, body)
    with pytest.raises(MdpdError) as e:
        read_xyz_native(f)
    assert e.value.code == capi.MDPD_ERR_PARSE and msg in str(e.value)


def test_capacity_and_missing_file(tmp_path):
    import ctypes as C
    f = _text(tmp_path, "3\ntime 5\n" + LINE * 3)
    buf = np.zeros(2, FPDPD_DTYPE)
    n, t = C.c_uint32(), C.c_uint32()
    rc = capi.lib().mdpd_read_xyz(str(f).encode(), buf.ctypes.data_as(C.c_void_p), 2, 232,
                                  C.byref(n), C.byref(t))
    assert rc == capi.MDPD_ERR_CAPACITY
    assert b"exceeds maximum buffer" in capi.lib().mdpd_global_error()
    rc = capi.lib().mdpd_read_xyz(str(f).encode(), buf.ctypes.data_as(C.c_void_p), 2, 200,
                                  C.byref(n), C.byref(t))
    assert rc == capi.MDPD_ERR_ARG  # stride below sizeof(FPDPD)
    with pytest.raises(MdpdError) as e:
        read_xyz_native(tmp_path / "absent.xyz")
    assert e.value.code == capi.MDPD_ERR_PARSE


@pytest.mark.parametrize("text,ok", [
    (b"", False),
    (b"   \n", False),
    (b"7", False),                      # no time line
    (b"7\n", False),
    (b"7\ntime", False),
    (b"7\ntime 12", True),              # header only, no trailing newline
    (b"7\ntime 12\n", True),
    (b"7\ntime " + b"9" * 300, False),  # longer than the header buffer
    (b"7" * 300, False),
])
def test_header_only_and_malformed_headers(tmp_path, text, ok):
    """mdpd_read_xyz_header reads at most the first 256 bytes and parses
    them as a terminated string: empty, short and overlong headers are parse
    errors, never reads past the buffer."""
    import ctypes as C
    path = tmp_path / "h.xyz"
    path.write_bytes(text)
    n, t = C.c_uint32(), C.c_uint32()
    rc = capi.lib().mdpd_read_xyz_header(str(path).encode(), C.byref(n), C.byref(t))
    if ok:
        assert rc == 0 and (n.value, t.value) == (7, 12)
    else:
        assert rc == capi.MDPD_ERR_PARSE
