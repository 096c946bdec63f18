"""The C++ drop-in class on the GPU: ncorr::calculateD1D2D3 followed by
getTriCountsOut, called exactly as source/main.cu:31-45 and 72-95 do, checked
against the oracle's compacted histogram (cudaNcorr.cu:72-86)."""
import os
import subprocess

import numpy as np
import pytest

from cudancorr_amd.gen import edges_linear, prep_cat_survey, survey_cap, uniform_box
from oracle import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "cudancorr_amd", "lib")


@pytest.fixture(scope="module")
def dropin(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dropin") / "dropin_main")
    src = os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), src, "-L" + LIBDIR,
                    "-lncorr", "-lncorr_hip", "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


def _run(exe, tmp, cats, names, survey, precision, box, edges):
    paths = {}
    for n, c in cats.items():
        p = str(tmp / f"{n}.bin")
        np.ascontiguousarray(c, dtype=np.float64).tofile(p)
        paths[n] = p
    args = [exe, "count", paths[names[0]], paths[names[1]], paths[names[2]], str(int(survey)), str(precision)]
    args += [repr(float(b)) for b in box] + [repr(float(e)) for e in edges]
    print("dropin", names, "survey", survey, "precision", precision, flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    return np.array([float(x) for x in out.stdout.split()])


@pytest.mark.parametrize("precision", [0, 1])
def test_dropin_box_ddd(dropin, tmp_path, precision):
    p = uniform_box(8000, 150.0, seed=31)
    edges = edges_linear(2.0, 20.0, 10)
    box = [150.0] * 3
    got = _run(dropin, tmp_path, {"D": p}, ("D", "D", "D"), False, precision, box, edges)
    c, _ = oracle.count(p, p, p, edges, box)
    want = oracle.compact(edges, c.astype(np.float64))
    assert got.size == 150
    assert np.array_equal(got, want)


@pytest.mark.parametrize("term", ["DDD", "DDR", "DRR", "RRR"])
def test_dropin_survey_terms(dropin, tmp_path, term):
    d, r = survey_cap(3000, 4500, seed=12)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 40.0, 20)
    got = _run(dropin, tmp_path, {"D": d, "R": r}, tuple(term), True, 1, box, edges)
    cats = {"D": d, "R": r}
    _, ws = oracle.count(cats[term[0]], cats[term[1]], cats[term[2]], edges, box, survey=True)
    want = oracle.compact(edges, ws)
    nz = want != 0
    assert np.array_equal(got[~nz], want[~nz])
    assert np.max(np.abs(got[nz] - want[nz]) / want[nz]) <= 1e-9


@pytest.mark.parametrize("precision", [0, 1])
def test_dropin_non_diagonal_bin_position(dropin, tmp_path, precision):
    """Known answer for the output order (cudaNcorr.cu:72-86, 145-154): one
    triangle with sides 1.5, 2, 2.5 (a 3-4-5 triangle halved) and edges
    1, 2, 3, 4 lies in the sorted bins (0, 1, 1) -- 2 is an edge and goes to
    the upper bin.  Its 6 ordered triplets must appear at the one
    getTriCountsOut() index whose getTriBinsOut() entry holds the mids
    (1.5, 2.5, 2.5), and nowhere else."""
    tri = np.array([[5.0, 5.0, 5.0, 1.0], [6.5, 5.0, 5.0, 1.0], [5.0, 7.0, 5.0, 1.0]])
    edges = [1.0, 2.0, 3.0, 4.0]
    p = str(tmp_path / "T.bin")
    tri.tofile(p)
    args = [dropin, "countbins", p, p, p, "1", str(precision), "20.0", "20.0", "20.0"] + [repr(e) for e in edges]
    print("dropin countbins precision", precision, flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = np.array([[float(x) for x in line.split()] for line in out.stdout.split("\n") if line.strip()])
    # mids 1.5, 2.5, 3.5: the 9 sorted triples passing mid_k <= mid_i + mid_j
    # (only bins (0, 0, 2), mids 1.5, 1.5, 3.5, fail), in (i, j, k) order
    assert rows[:, 1:].tolist() == [[1.5, 1.5, 1.5], [1.5, 1.5, 2.5], [1.5, 2.5, 2.5], [1.5, 2.5, 3.5],
                                    [1.5, 3.5, 3.5], [2.5, 2.5, 2.5], [2.5, 2.5, 3.5], [2.5, 3.5, 3.5],
                                    [3.5, 3.5, 3.5]]
    assert rows[:, 0].tolist() == [0, 0, 6, 0, 0, 0, 0, 0, 0]


def test_dropin_allocation_failure_is_runtime_error(dropin, tmp_path):
    """A device allocation failing inside calculateD1D2D3 (the library's test
    knob NC_FAIL_DEVICE_ALLOC=0 fails the context's first one) reaches the
    caller as std::runtime_error naming the failed call, as the reference's
    gpuErrchk does (gpuerrchk.h:9-14)."""
    p = uniform_box(500, 50.0, seed=3)
    f = str(tmp_path / "P.bin")
    p.tofile(f)
    args = [dropin, "count", f, f, f, "0", "0", "50.0", "50.0", "50.0"] + [repr(float(e)) for e in edges_linear(0.0, 5.0, 5)]
    print("dropin allocation failure", flush=True)
    out = subprocess.run(args, capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, NC_FAIL_DEVICE_ALLOC="0"))
    assert out.returncode == 3, out.stdout + out.stderr
    assert out.stdout.startswith("EXCEPTION runtime_error calculateD1D2D3:"), out.stdout
    assert "out of memory" in out.stdout


@pytest.mark.parametrize("nth", [0, 3, 9])
def test_allocation_failure_leaves_context_usable(monkeypatch, nth):
    """NC_FAIL_DEVICE_ALLOC=n fails the n-th device allocation of a fresh
    context (catalogue upload or a buffer of the count): the call returns
    NC_ERR_NOMEM, and the next calls on the same context count correctly."""
    from cudancorr_amd import Context, NcError

    monkeypatch.setenv("NC_FAIL_DEVICE_ALLOC", str(nth))
    p = uniform_box(3000, 80.0, seed=5)
    edges = edges_linear(0.0, 8.0, 8)
    box = [80.0] * 3
    want, _ = oracle.count(p, p, p, edges, box)
    print("allocation failure at", nth, flush=True)
    with Context(0) as c:
        failed = False
        for _ in range(12):  # the countdown passes the failing allocation within these calls
            try:
                _, got = c.count(p, p, p, edges, box)
            except NcError as e:
                assert e.code == 4 and "out of memory" in str(e)
                failed = True
                continue
            assert np.array_equal(got, want)
        assert failed
        _, got = c.count(p, p, p, edges, box)
        assert np.array_equal(got, want)
