"""The include/utils.hpp helpers (reference source/utils.cu:16-208) on the CPU,
through tests/cpp/utils_main.cpp built with g++ against libncorr.so.

* logspace against numpy.logspace and exact powers of ten;
* generateRandomPosition against an independent Mersenne Twister stream
  (numpy's legacy MT19937, itself checked against the C++ standard's
  known answer) through libstdc++'s generate_canonical;
* the seeded catalogue: reproducible, prefix-stable, uniform (KS test);
* the writers' byte format against Python's own "%.6f" formatting;
* readCatalogTxt's column map, bounds error and line numbers;
* readCatalogFits on binary tables written byte by byte here from the FITS
  standard (no FITS library or reference fixture exists in this image).
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest
from scipy import stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "cudancorr_amd", "lib")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("utils") / "utils_main")
    src = os.path.join(ROOT, "tests", "cpp", "utils_main.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"), src, "-L" + LIBDIR,
                    "-lncorr", "-lncorr_hip", "-Wl,-rpath," + LIBDIR, "-o", out], check=True)
    return out


def run(exe, *args, ok=True):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=120)
    if ok:
        assert r.returncode == 0, r.stdout + r.stderr
    return r


def points(text):
    rows = [list(map(float, ln.split())) for ln in text.splitlines() if ln.strip()]
    return np.array(rows).reshape(-1, 4)


def error_of(r):
    assert r.returncode == 3, r.stdout + r.stderr
    assert r.stdout.startswith("EXCEPTION ")
    return r.stdout[len("EXCEPTION "):].strip()


# ------------------------------------------------------------------ logspace


def test_logspace_matches_numpy(exe):
    # the bins of source/main.cu:21
    got = np.array([float(x) for x in run(exe, "logspace", 0.1, 20.0, 13).stdout.split()])
    want = np.logspace(math.log10(0.1), math.log10(20.0), 13)
    assert got.size == 13
    assert np.allclose(got, want, rtol=4e-15, atol=0)
    assert got[0] == 10.0 ** math.log10(0.1)
    assert np.all(np.diff(np.log10(got)) > 0)


def test_logspace_exact_decades(exe):
    got = [float(x) for x in run(exe, "logspace", 1, 1e6, 7).stdout.split()]
    assert got == [1.0, 10.0, 100.0, 1e3, 1e4, 1e5, 1e6]


def test_logspace_single_edge_is_nan(exe):
    # step = (b - a)/0 and 0 * inf: the reference's arithmetic gives NaN
    got = run(exe, "logspace", 1.0, 10.0, 1).stdout.split()
    assert len(got) == 1 and math.isnan(float(got[0]))
    assert run(exe, "logspace", 1.0, 10.0, 0).stdout == ""


# --------------------------------------------------------------- random points


def _mt_raw(seed, n):
    return np.random.RandomState(seed)._bit_generator.random_raw(n).astype(np.float64)


def test_numpy_mt19937_known_answer():
    # C++ [rand.predef]: the 10000th draw of a default mt19937 (seed 5489)
    raw = np.random.RandomState(5489)._bit_generator.random_raw(10000)
    assert raw[-1] == 4123659995


@pytest.mark.parametrize("seed", [5489, 12345])
def test_random_position_draw_order(exe, seed):
    """uniform_real_distribution(0, b) over mt19937: two 32-bit draws g1, g2
    make (g1 + g2 * 2^32) / 2^64, scaled by b; x, y, z in that order, w = 1."""
    n = 50
    box = (2000.0, 1500.0, 750.0)
    got = points(run(exe, "randpos", seed, n, *box).stdout)
    raw = _mt_raw(seed, 6 * n).reshape(n, 3, 2)
    u = (raw[..., 0] + raw[..., 1] * 2.0 ** 32) / 2.0 ** 64
    u = np.minimum(u, np.nextafter(1.0, 0.0))
    want = u * np.array(box)
    assert np.array_equal(got[:, :3], want)
    assert np.all(got[:, 3] == 1.0)


def _cat(exe, tmp_path, name, np_, box, seed=None):
    f = tmp_path / name
    run(exe, "randcat", f, np_, *box, *([] if seed is None else [seed]))
    return np.fromfile(f, dtype=np.float64).reshape(-1, 4)


def test_seeded_catalogue_is_reproducible_and_uniform(exe, tmp_path):
    box = (2000.0, 1000.0, 500.0)
    n = 300_000  # several blocks of 65536 points
    a = _cat(exe, tmp_path, "a.bin", n, box, seed=7)
    b = _cat(exe, tmp_path, "b.bin", n, box, seed=7)
    c = _cat(exe, tmp_path, "c.bin", n, box, seed=8)
    short = _cat(exe, tmp_path, "s.bin", 70_000, box, seed=7)
    assert a.shape == (n, 4)
    assert np.array_equal(a, b)
    assert not np.array_equal(a[:, :3], c[:, :3])
    assert np.array_equal(short, a[:70_000])
    assert np.all(a[:, 3] == 1.0)
    for k in range(3):
        assert np.all((a[:, k] >= 0) & (a[:, k] < box[k]))
        assert stats.kstest(a[:, k] / box[k], "uniform").pvalue > 1e-4
    # the blocks are not copies of one another
    assert not np.array_equal(a[:65536, :3], a[65536:131072, :3])
    assert abs(np.corrcoef(a[:65536, 0], a[65536:131072, 0])[0, 1]) < 0.02


def test_unseeded_catalogue_differs_per_call(exe, tmp_path):
    box = (10.0, 10.0, 10.0)
    a = _cat(exe, tmp_path, "u1.bin", 1000, box)
    b = _cat(exe, tmp_path, "u2.bin", 1000, box)
    assert a.shape == (1000, 4) and not np.array_equal(a, b)
    assert np.all((a[:, :3] >= 0) & (a[:, :3] < 10.0))


def test_negative_catalogue_size(exe, tmp_path):
    assert "negative" in error_of(run(exe, "randcat", tmp_path / "n.bin", -1, 1, 1, 1, 3, ok=False))


# ------------------------------------------------------------------- writers


def _rows(k, n=40, seed=3):
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1e4, 1e7, (n, k))
    v[0] = [0.0000005, -0.0000004, 1e15 + 0.25, 2.5e-7, 0.1234565, -0.0][:k]  # rounding and sign edges
    v[1, -1] = 123456789012.0
    return v


@pytest.mark.parametrize("mode,k", [("write3d", 4), ("writerppi", 6)])
def test_writer_format(exe, tmp_path, mode, k):
    v = _rows(k)
    src, out = tmp_path / "in.bin", tmp_path / "out.txt"
    v.tofile(src)
    run(exe, mode, out, src)
    want = "".join(" ".join("%.6f" % x for x in row) + "\n" for row in v)
    assert out.read_text() == want


@pytest.mark.parametrize("mode,k", [("write3d", 4), ("writerppi", 6)])
def test_writer_errors(exe, tmp_path, mode, k):
    src = tmp_path / "in.bin"
    _rows(k, n=3).tofile(src)
    assert error_of(run(exe, mode, tmp_path / "o.txt", src, "extra", ok=False)) == "Vectors have different sizes."
    bad = tmp_path / "no_such_dir" / "o.txt"
    assert error_of(run(exe, mode, bad, src, ok=False)) == f"Unable to open file: {bad}"


def test_written_counts_read_back(exe, tmp_path):
    """writeToFileThreeD output is a readCatalogTxt input: r1 r2 r3 count with w = count."""
    rng = np.random.default_rng(11)
    v = np.c_[rng.uniform(0, 20, (25, 3)), rng.integers(0, 10 ** 6, 25)]
    src, out = tmp_path / "in.bin", tmp_path / "out.txt"
    v.tofile(src)
    run(exe, "write3d", out, src)
    got = points(run(exe, "readtxt", out, 20, 20, 20, 0, 1, 2, 3).stdout)
    assert np.allclose(got, np.round(v, 6), rtol=0, atol=5e-7)
    assert np.array_equal(got[:, 3], v[:, 3])


# ---------------------------------------------------------------- text input


def test_txt_column_map(exe, tmp_path):
    f = tmp_path / "cat.txt"
    f.write_text("9 1.5 2.5 3.5 0.75\n"
                 "\n"
                 "# a comment line\n"
                 "8 4 5 6 2 trailing words 7 7\n"
                 "  7\t10 0 0 1e0   \n")
    got = points(run(exe, "readtxt", f, 10, 10, 10, 1, 2, 3, 4).stdout)
    assert np.array_equal(got, [[1.5, 2.5, 3.5, 0.75], [4, 5, 6, 2], [10, 0, 0, 1]])
    # permuted columns, x from column 0 (10 is on the box face: inside)
    got = points(run(exe, "readtxt", f, 10, 10, 10, 3, 0, 1, 2).stdout)
    assert np.array_equal(got, [[3.5, 9, 1.5, 2.5], [6, 8, 4, 5], [0, 7, 10, 0]])


def test_txt_weight_defaults_to_one(exe, tmp_path):
    f = tmp_path / "cat.txt"
    f.write_text("1 2 3\n4 5 6 0.5\n")
    got = points(run(exe, "readtxt", f, 10, 10, 10, 0, 1, 2, 3).stdout)
    assert np.array_equal(got, [[1, 2, 3, 1], [4, 5, 6, 0.5]])
    got = points(run(exe, "readtxt", f, 10, 10, 10, 0, 1, 2).stdout)
    assert np.array_equal(got[:, 3], [1, 1])


@pytest.mark.parametrize("bad,line", [("-0.5 1 1", 3), ("1 10.000001 1", 3), ("1 1 nan_text 4", 3)])
def test_txt_errors_name_the_line(exe, tmp_path, bad, line):
    f = tmp_path / "cat.txt"
    f.write_text("1 1 1\n\n" + bad + "\n2 2 2\n")
    msg = error_of(run(exe, "readtxt", f, 10, 10, 10, 0, 1, 2, 3, ok=False))
    assert msg.startswith(f"Error at line {line}:")
    if "text" not in bad:
        assert msg == f"Error at line {line}: double3 value is out of bounds."


def test_txt_missing_file(exe, tmp_path):
    f = tmp_path / "none.txt"
    assert error_of(run(exe, "readtxt", f, 1, 1, 1, 0, 1, 2, 3, ok=False)) == f"Unable to open file: {f}"


# ---------------------------------------------------------------- FITS input


class Raw(str):
    """A header value written as given (not quoted)."""


def _card(key, value=None, comment=""):
    if value is None:
        s = key.ljust(8)
    else:
        if isinstance(value, bool):
            v = ("T" if value else "F").rjust(20)
        elif isinstance(value, Raw):
            v = str(value).rjust(20)
        elif isinstance(value, str):
            v = ("'" + value.replace("'", "''").ljust(8) + "'").ljust(20)
        else:
            v = repr(value).rjust(20)
        s = key.ljust(8) + "= " + v + (" / " + comment if comment else "")
    assert len(s) <= 80
    return s.ljust(80).encode("ascii")


def _header(cards):
    b = b"".join(cards) + _card("END")
    return b + b" " * (-len(b) % 2880)


def _pad(data):
    return data + b"\0" * (-len(data) % 2880)


def _primary():
    return _header([_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0), _card("EXTEND", True)])


def _image(shape=(10, 7)):
    cards = [_card("XTENSION", "IMAGE"), _card("BITPIX", -32), _card("NAXIS", len(shape))]
    cards += [_card(f"NAXIS{i + 1}", n) for i, n in enumerate(shape)]
    cards += [_card("PCOUNT", 0), _card("GCOUNT", 1)]
    data = np.arange(np.prod(shape), dtype=">f4").tobytes()
    return _header(cards) + _pad(data)


_DT = {"D": ">f8", "E": ">f4", "J": ">i4", "I": ">i2", "K": ">i8", "B": "u1"}


def _bintable(cols, nrows=None, truncate=0, extra=()):
    """cols: (ttype, tform, values, {keyword: value}); tform "rT" with T in
    B I J K E D A."""
    fields, arrays = [], []
    for i, (name, form, vals, kw) in enumerate(cols):
        rep = int(form[:-1] or 1)
        t = form[-1]
        dt = f"S{rep}" if t == "A" else (_DT[t], (rep,)) if rep != 1 else _DT[t]
        fields.append((f"f{i}",) + ((dt,) if isinstance(dt, str) else dt))
        arrays.append(vals)
    n = len(arrays[0]) if nrows is None else nrows
    rec = np.zeros(n, dtype=fields)
    for i, a in enumerate(arrays):
        rec[f"f{i}"] = a[:n]
    cards = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
             _card("NAXIS1", rec.dtype.itemsize), _card("NAXIS2", n), _card("PCOUNT", 0), _card("GCOUNT", 1),
             _card("TFIELDS", len(cols))]
    for i, (name, form, _, kw) in enumerate(cols, start=1):
        cards += [_card(f"TTYPE{i}", name), _card(f"TFORM{i}", form, "data format")]
        cards += [_card(f"{k}{i}", v) for k, v in kw.items()]
    cards += [_card("COMMENT", None)] + [_card(k, v) for k, v in extra]
    data = _pad(rec.tobytes())
    return _header(cards) + (data[: len(data) - truncate] if truncate else data)


def _catalogue(n=200, seed=5):
    rng = np.random.default_rng(seed)
    ra = rng.uniform(0, 360, n)
    dec = rng.uniform(-90, 90, n).astype(np.float32)
    z = np.round(rng.uniform(0.0, 1.2, n), 2)  # ties with zmin/zmax
    wraw = rng.integers(-1000, 1000, n).astype(np.int32)
    u16 = rng.integers(0, 65536, n)
    vec = rng.uniform(0, 1, (n, 3))
    names = np.array([b"gal%05d" % i for i in range(n)])
    cols = [("ID", "K", np.arange(n, dtype=np.int64) + 2 ** 40, {}),
            ("ra", "D", ra, {"TUNIT": "deg"}),
            ("NAME", "8A", names, {}),
            ("Dec", "E", dec, {}),
            ("FLAG", "B", (np.arange(n) % 256).astype(np.uint8), {}),
            ("z", "D", z, {}),
            ("weight", "J", wraw, {"TSCAL": Raw("5.0D-1"), "TZERO": 1.0}),
            ("U16", "I", (u16 - 32768).astype(np.int16), {"TZERO": 32768}),
            ("VEC", "3D", vec, {})]
    truth = dict(ra=ra, dec=dec.astype(np.float64), z=z, w=wraw * 0.5 + 1.0, u16=u16, vec0=vec[:, 0])
    return cols, truth


def _write(tmp_path, name, *hdus):
    f = tmp_path / name
    f.write_bytes(b"".join(hdus))
    return f


NAMES = ["RA", "DEC", "Z", "WEIGHT"]


def test_fits_reads_scaled_columns(exe, tmp_path):
    cols, t = _catalogue()
    f = _write(tmp_path, "cat.fits", _primary(), _image(), _bintable(cols, extra=[("OBSERVER", "O'Brien")]))
    got = points(run(exe, "readfits", f, 3, 0.2, 1.0, *NAMES).stdout)
    keep = (t["z"] > 0.2) & (t["z"] < 1.0)
    assert 0 < keep.sum() < len(keep) and np.any(t["z"] == 0.2) and np.any(t["z"] == 1.0)
    want = np.c_[t["ra"], t["dec"], t["z"], t["w"]][keep]
    assert np.array_equal(got, want)


def test_fits_weight_defaults_to_one(exe, tmp_path):
    cols, t = _catalogue(n=40)
    f = _write(tmp_path, "cat.fits", _primary(), _bintable(cols))
    got = points(run(exe, "readfits", f, 2, -1, 2, "Z", "RA", "DEC").stdout)
    assert np.array_equal(got, np.c_[t["ra"], t["dec"], t["z"], np.ones(40)])


def test_fits_other_types_and_first_vector_element(exe, tmp_path):
    """Only RA, DEC, Z, WEIGHT are kept (utils.cu:103-106); the other names
    must still be found and be readable, and the first vector element is the
    one read: checked by naming them in WEIGHT's place in a table whose columns
    are renamed."""
    cols, t = _catalogue(n=30)
    for src, key in (("U16", "u16"), ("VEC", "vec0")):
        ren = [(("WEIGHT" if c[0] == src else ("W_OLD" if c[0] == "weight" else c[0])),) + c[1:] for c in cols]
        f = _write(tmp_path, f"{src}.fits", _primary(), _bintable(ren))
        got = points(run(exe, "readfits", f, 2, -1, 2, *NAMES).stdout)
        assert np.array_equal(got[:, 3], t[key].astype(np.float64))
    # K column with 2^40 offsets and a B column read through the same path
    ren = [(("WEIGHT" if c[0] == "ID" else ("W_OLD" if c[0] == "weight" else c[0])),) + c[1:] for c in cols]
    f = _write(tmp_path, "id.fits", _primary(), _bintable(ren))
    got = points(run(exe, "readfits", f, 2, -1, 2, *NAMES, "flag").stdout)
    assert np.array_equal(got[:, 3], np.arange(30) + 2.0 ** 40)


def test_fits_hdu_errors(exe, tmp_path):
    cols, _ = _catalogue(n=10)
    ascii_tab = _header([_card("XTENSION", "TABLE"), _card("BITPIX", 8), _card("NAXIS", 2), _card("NAXIS1", 8),
                         _card("NAXIS2", 1), _card("PCOUNT", 0), _card("GCOUNT", 1), _card("TFIELDS", 1),
                         _card("TFORM1", "F8.3"), _card("TBCOL1", 1)]) + _pad(b"   1.000")
    f = _write(tmp_path, "multi.fits", _primary(), _image(), ascii_tab, _bintable(cols))
    assert len(points(run(exe, "readfits", f, 4, -1, 2, *NAMES).stdout)) == 10
    for hdu in (1, 2, 3):
        assert error_of(run(exe, "readfits", f, hdu, -1, 2, *NAMES, ok=False)) == \
            "The specified HDU is not a Binary Table."
    for hdu in (0, 5, 9):
        assert error_of(run(exe, "readfits", f, hdu, -1, 2, *NAMES, ok=False)) == "Error moving to the specified HDU."


def test_fits_column_errors(exe, tmp_path):
    cols, _ = _catalogue(n=10)
    f = _write(tmp_path, "cat.fits", _primary(), _bintable(cols))
    assert error_of(run(exe, "readfits", f, 2, -1, 2, "RA", "MAG", ok=False)) == "Error finding column named MAG"
    # a string column cannot be read as a number
    assert error_of(run(exe, "readfits", f, 2, -1, 2, "NAME", "MAG", ok=False)) == "Error reading FITS file data."
    # a template matching several columns is not unique; '#' and '?' templates
    assert error_of(run(exe, "readfits", f, 2, -1, 2, "*A*", ok=False)) == "Error finding column named *A*"
    assert len(points(run(exe, "readfits", f, 2, -1, 2, "U#", "R?", "*EIGHT", *NAMES).stdout)) == 10
    dup = cols + [("Z", "D", np.zeros(10), {})]
    g = _write(tmp_path, "dup.fits", _primary(), _bintable(dup))
    assert error_of(run(exe, "readfits", g, 2, -1, 2, *NAMES, ok=False)) == "Error finding column named Z"
    # no rows: no column is looked up (utils.cu:88)
    e = _write(tmp_path, "empty.fits", _primary(), _bintable(cols, nrows=0))
    assert run(exe, "readfits", e, 2, -1, 2, "MAG").stdout == ""


def test_fits_file_errors(exe, tmp_path):
    cols, _ = _catalogue(n=100)
    short = _write(tmp_path, "short.fits", _primary(), _bintable(cols, truncate=2880))
    assert error_of(run(exe, "readfits", short, 2, -1, 2, *NAMES, ok=False)) == "Error reading FITS file data."
    for name, data in (("text.fits", b"1 2 3\n" * 1000), ("empty.fits", b"")):
        f = _write(tmp_path, name, data)
        assert error_of(run(exe, "readfits", f, 2, -1, 2, *NAMES, ok=False)) == f"Unable to open FITS file: {f}"
    missing = tmp_path / "none.fits"
    assert error_of(run(exe, "readfits", missing, 2, -1, 2, *NAMES, ok=False)) == \
        f"Unable to open FITS file: {missing}"


def test_fits_long_header_and_wide_rows(exe, tmp_path):
    """A header over several 2880-byte blocks, and rows of 8.8 MB so that the
    table is read in several chunks."""
    n, rep = 9, 1_100_000
    rng = np.random.default_rng(2)
    big = rng.uniform(0, 1, (n, rep))
    z = np.linspace(0.05, 0.95, n)
    cols = [("RA", f"{rep}D", big, {}), ("DEC", "D", -big[:, 1], {}), ("Z", "E", z.astype(np.float32), {})]
    extra = [(f"HIST{i:04d}", i) for i in range(120)]
    f = _write(tmp_path, "wide.fits", _primary(), _bintable(cols, extra=extra))
    got = points(run(exe, "readfits", f, 2, 0.0, 1.0, "RA", "DEC", "Z").stdout)
    assert np.array_equal(got, np.c_[big[:, 0], -big[:, 1], z.astype(np.float32), np.ones(n)])


def test_fits_header_value_parsing():
    """The test's own card writer follows the fixed format (FITS 4.0 sec. 4.2)."""
    c = _card("TTYPE1", "O'Brien", "x")
    assert c[:10] == b"TTYPE1  = " and c[10:21] == b"'O''Brien' " and len(c) == 80
    assert _card("NAXIS1", 12)[10:30] == b" " * 18 + b"12"
    assert struct.unpack(">d", np.array([1.5], ">f8").tobytes())[0] == 1.5
