"""rppi binning on the GPU (k_rppi through the C-ABI and the drop-in class)
against the rppi oracle (device.cu:203-255, 317-379, 434-505).  Unweighted
counts bit-exact; weighted sums within 1e-9 relative."""
import os
import subprocess

import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN, Context, NcError
from cudancorr_amd.gen import neyman_scott_box, prep_cat_survey, survey_cap, uniform_box
from oracle import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "cudancorr_amd", "lib")
WEIGHT_RTOL = 1e-9


def _log(msg):
    print(msg, flush=True)


def _wcheck(got, want):
    nz = want != 0
    assert np.array_equal(got[~nz], want[~nz])
    assert np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz]), initial=0.0) <= WEIGHT_RTOL


def _box_case(seed=4):
    # box: 12 x 12 x 10 cells of (rpmax, rpmax, npi) = (8, 8, 6)
    p = neyman_scott_box(300, 20, 2.0, 96.0, seed=seed)
    p[:, 2] = np.mod(p[:, 2], 60.0)
    edges = np.array([0.5, 1.0, 2.0, 3.0, 4.5, 6.0, 8.0])
    return p, edges, 6, np.array([96.0, 96.0, 60.0])


@pytest.mark.parametrize("path", [NC_PATH_PLAIN, NC_PATH_MIXED])
def test_box_ddd(ctx, path):
    _log(f"rppi box ddd path={path}")
    p, edges, npi, box = _box_case()
    want, _, lost = oracle.count_rppi(p, p, p, edges, npi, box)
    hist, got = ctx.count(p, p, p, edges, box, path=path, npi=npi)
    assert want.sum() > 100_000 and lost == 0
    assert got.shape == (6, 6, 6, npi, npi)
    assert np.array_equal(got, want)
    assert np.array_equal(hist, want.astype(np.float64))


def test_box_asymmetric_weighted(ctx):
    _log("rppi box weighted ddr")
    p, edges, npi, box = _box_case()
    rng = np.random.default_rng(9)
    d = p.copy()
    d[:, 3] = rng.lognormal(0.0, 0.3, d.shape[0])
    r = uniform_box(5000, 96.0, seed=17)
    r[:, 2] = np.mod(r[:, 2], 60.0)
    for (a, b, c) in ((d, d, r), (d, r, r), (r, d, d)):
        cnt, ws, _ = oracle.count_rppi(a, b, c, edges, npi, box)
        hist, _ = ctx.count(a, b, c, edges, box, npi=npi)
        assert cnt.sum() > 10_000
        _wcheck(hist, ws)


def test_box_ties_on_a_lattice(ctx):
    """Integer coordinates: many rp^2 ties between the sides of a triangle, where
    the two orders of (p2, p3) may fall in different pi bins."""
    _log("rppi box lattice ties")
    g = np.arange(0, 24, 1.5)
    X, Y, Z = np.meshgrid(g, g, g[:12], indexing="ij")
    p = np.c_[X.ravel(), Y.ravel(), Z.ravel(), np.ones(X.size)]
    edges = np.array([0.5, 1.6, 2.2, 3.1, 4.0])
    box = np.array([24.0, 24.0, 18.0])  # cells (4, 4, 3): 6 x 6 x 6
    want, _, _ = oracle.count_rppi(p, p, p, edges, 3, box)
    _, got = ctx.count(p, p, p, edges, box, npi=3)
    assert want.sum() > 100_000
    assert np.array_equal(got, want)


@pytest.mark.parametrize("term", ["DDD", "DDR", "DRR", "RRR"])
def test_survey_terms(ctx, term):
    _log(f"rppi survey {term}")
    d, r = survey_cap(2500, 4000, seed=13)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.1
    box, ori = prep_cat_survey(d, r)
    edges = np.array([0.0, 2.0, 4.0, 6.0, 9.0])
    npi = 12  # cells sqrtf(9^2 + 12^2) = 15
    cats = {"D": d, "R": r}
    a, b, c = (cats[ch] for ch in term)
    cnt, ws, lost = oracle.count_rppi(a, b, c, edges, npi, box, survey=True, ori=ori)
    hist, got = ctx.count(a, b, c, edges, box, survey=True, npi=npi, ori=ori)
    assert cnt.sum() > 1000 and lost == 0
    if term == "RRR":
        assert np.array_equal(got, cnt)
    else:
        _wcheck(hist, ws)


def test_batched_replays(ctx):
    _log("rppi batched")
    p, edges, npi, box = _box_case(seed=6)
    q = uniform_box(4000, 96.0, seed=23)
    q[:, 2] = np.mod(q[:, 2], 60.0)
    P, Q = ctx.catalog(p), ctx.catalog(q)
    terms = [(P, P, P), (P, P, Q), (Q, Q, Q)]
    want = [oracle.count_rppi(a, b, c, edges, npi, box)[0] for a, b, c in ((p, p, p), (p, p, q), (q, q, q))]
    for rep in range(3):
        _, cnt = ctx.count_batched(terms, edges, box, path=NC_PATH_PLAIN, npi=npi)
        for i in range(3):
            assert np.array_equal(cnt[i], want[i]), f"replay {rep} term {i}"
    P.close()
    Q.close()


@pytest.fixture(scope="module")
def dropin(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dropin_rppi") / "dropin_main")
    src = os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), src, "-L" + LIBDIR,
                    "-lncorr", "-lncorr_hip", "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


@pytest.mark.parametrize("survey", [False, True])
def test_dropin_rppi(dropin, tmp_path, survey):
    _log(f"rppi dropin survey={survey}")
    if survey:
        d, r = survey_cap(2000, 10, seed=14)
        d[:, :3] = 1000.0 + (d[:, :3] - 1000.0) * 0.1
        box, ori = prep_cat_survey(d, r)
        p, edges, npi = d, np.array([0.0, 2.0, 4.0, 6.0, 9.0]), 12
    else:
        p, edges, npi, box = _box_case(seed=8)
        ori = (0.0, 0.0, 0.0)
    path = str(tmp_path / "p.bin")
    np.ascontiguousarray(p, dtype=np.float64).tofile(path)
    args = [dropin, "rppi", path, path, path, str(int(survey)), str(npi), repr(float(ori[0]))]
    args += [repr(float(b)) for b in box] + [repr(float(e)) for e in edges]
    out = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = np.array([float(x) for x in out.stdout.split()])
    cnt, ws, _ = oracle.count_rppi(p, p, p, edges, npi, box, survey=survey, ori=ori)
    want = oracle.compact_rppi(edges, npi, ws)
    assert got.size == want.size
    nz = want != 0
    assert np.array_equal(got[~nz], want[~nz])
    assert np.max(np.abs(got[nz] - want[nz]) / want[nz], initial=0.0) <= WEIGHT_RTOL


def test_key_stream_layout_and_direct_atomics(ctx, monkeypatch):
    """The key stream (binned per bucket in LDS) and the direct global atomics
    give the same counts; weighted sums agree to rounding."""
    _log("rppi stream vs direct")
    p, edges, npi, box = _box_case(seed=12)
    rng = np.random.default_rng(3)
    w = p.copy()
    w[:, 3] = rng.lognormal(0.0, 0.3, w.shape[0])
    h_s, c_s = ctx.count(p, p, p, edges, box, npi=npi)
    assert ctx.timing().hist_layout == 4
    hw_s, _ = ctx.count(w, w, p, edges, box, npi=npi)
    monkeypatch.setenv("NC_RPPI_STREAM", "0")
    with Context(0) as c2:
        h_d, c_d = c2.count(p, p, p, edges, box, npi=npi)
        assert c2.timing().hist_layout == 3
        hw_d, _ = c2.count(w, w, p, edges, box, npi=npi)
    assert c_s.sum() > 100_000
    assert np.array_equal(c_s, c_d)
    _wcheck(hw_s, hw_d)


@pytest.mark.parametrize("survey", [False, True])
def test_key_stream_region_overflow(monkeypatch, survey):
    """Regions sized at 5% of the sampled need: the keys past them are added
    directly, on every call, and the results stay exact."""
    _log(f"rppi region overflow survey={survey}")
    monkeypatch.setenv("NC_RPPI_REGION_SCALE", "0.05")
    monkeypatch.setenv("NC_RPPI_STREAM", "1")  # surveys default to direct atomics
    if survey:
        d, r = survey_cap(2500, 10, seed=13)
        d[:, :3] = 1000.0 + (d[:, :3] - 1000.0) * 0.1
        box, ori = prep_cat_survey(d, r)
        p, edges, npi = d, np.array([0.0, 2.0, 4.0, 6.0, 9.0]), 12
    else:
        p, edges, npi, box = _box_case(seed=5)
        ori = (0.0, 0.0, 0.0)
    cnt, ws, _ = oracle.count_rppi(p, p, p, edges, npi, box, survey=survey, ori=ori)
    with Context(0) as c2:
        for _ in range(2):
            hist, got = c2.count(p, p, p, edges, box, survey=survey, npi=npi, ori=ori)
            assert c2.timing().hist_layout == 4
            if survey:
                _wcheck(hist, ws)
            else:
                assert np.array_equal(got, cnt)


def test_key_stream_batched_mixed_weights(ctx):
    """One launch over weighted and unit-weight terms: the unit-weight terms
    stage their keys in the weighted layout; every term against the oracle,
    over replays of the captured graph."""
    _log("rppi stream batched mixed weights")
    p, edges, npi, box = _box_case(seed=15)
    rng = np.random.default_rng(4)
    w = p.copy()
    w[:, 3] = rng.lognormal(0.0, 0.3, w.shape[0])
    q = uniform_box(3000, 96.0, seed=29)
    q[:, 2] = np.mod(q[:, 2], 60.0)
    arrs = {"p": p, "w": w, "q": q}
    cats = {k: ctx.catalog(v) for k, v in arrs.items()}
    specs = ["ppp", "wwq", "qqq", "wqq", "pqp"]
    terms = [tuple(cats[ch] for ch in sp) for sp in specs]
    want = [oracle.count_rppi(*(arrs[ch] for ch in sp), edges, npi, box) for sp in specs]
    for rep in range(3):
        hist, cnt = ctx.count_batched(terms, edges, box, path=NC_PATH_PLAIN, npi=npi)
        assert ctx.timing().hist_layout == 4
        for i, sp in enumerate(specs):
            c, ws, _ = want[i]
            if "w" in sp:
                _wcheck(hist[i], ws)
            else:
                assert np.array_equal(cnt[i], c), f"replay {rep} term {sp}"
    for c in cats.values():
        c.close()


@pytest.mark.parametrize("survey", [False, True])
def test_key_stream_rounds(monkeypatch, survey):
    """The work queue cut into many rounds (NC_RPPI_ROUND_KEYS), each with its
    own regions: the same results, eagerly and over graph replays."""
    _log(f"rppi rounds survey={survey}")
    monkeypatch.setenv("NC_RPPI_ROUND_KEYS", "5000")
    monkeypatch.setenv("NC_RPPI_STREAM", "1")  # surveys default to direct atomics
    if survey:
        d, r = survey_cap(2500, 2000, seed=19)
        for c in (d, r):
            c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.1
        box, ori = prep_cat_survey(d, r)
        edges, npi = np.array([0.0, 2.0, 4.0, 6.0, 9.0]), 12
        specs = {"D": d, "R": r}
        terms = ["DDD", "DRR", "RRR"]
    else:
        p, edges, npi, box = _box_case(seed=21)
        q = uniform_box(3000, 96.0, seed=31)
        q[:, 2] = np.mod(q[:, 2], 60.0)
        ori = (0.0, 0.0, 0.0)
        specs = {"D": p, "R": q}
        terms = ["DDD", "DDR", "RRR"]
    want = [oracle.count_rppi(*(specs[ch] for ch in t), edges, npi, box, survey=survey, ori=ori) for t in terms]
    with Context(0) as c2:
        cats = {k: c2.catalog(v) for k, v in specs.items()}
        tl = [tuple(cats[ch] for ch in t) for t in terms]
        _, got = c2.count(*tl[-1], edges, box, survey=survey, npi=npi, ori=ori)
        assert c2.timing().n_kernel >= 6  # at least three rounds of two kernels
        assert np.array_equal(got, want[-1][0])
        for rep in range(3):
            hist, cnt = c2.count_batched(tl, edges, box, survey=survey, npi=npi, ori=ori)
            for i, t in enumerate(terms):
                c, ws, _ = want[i]
                if "D" in t and survey:
                    _wcheck(hist[i], ws)
                else:
                    assert np.array_equal(cnt[i], c), f"replay {rep} term {t}"
        for c in cats.values():
            c.close()


def test_survey_stream_workgroup_limit(monkeypatch):
    """The survey key-stream kernel is built for at most 8 waves per workgroup:
    16 is refused with NC_ERR_LIMIT before any launch, 8 counts as usual, and
    the autotuner never picks 16 there."""
    _log("rppi survey stream workgroup limit")
    monkeypatch.setenv("NC_RPPI_STREAM", "1")  # surveys default to direct atomics
    d, r = survey_cap(1500, 2000, seed=31)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.1
    box, ori = prep_cat_survey(d, r)
    edges = np.array([0.0, 2.0, 4.0, 6.0, 9.0])
    npi = 12
    cnt, _, _ = oracle.count_rppi(r, r, r, edges, npi, box, survey=True, ori=ori)
    with Context(0) as c2:
        c2.set_tuning(waves_per_group=16)
        with pytest.raises(NcError) as e:
            c2.count(r, r, r, edges, box, survey=True, npi=npi, ori=ori)
        assert e.value.code == 6
        c2.set_tuning(waves_per_group=8)
        _, got = c2.count(r, r, r, edges, box, survey=True, npi=npi, ori=ori)
        assert np.array_equal(got, cnt)
        R = c2.catalog(r)
        best = c2.autotune([(R, R, R)], edges, box, survey=True, npi=npi, ori=ori)
        assert best["waves_per_group"] <= 8
        _, got = c2.count_batched([(R, R, R)], edges, box, survey=True, npi=npi, ori=ori)
        assert np.array_equal(got[0], cnt)
        R.close()


def test_default_output_path(monkeypatch):
    """Without NC_RPPI_STREAM, boxes take the key stream (layout 4) and surveys
    the direct atomics (layout 3); NC_RPPI_STREAM=1 streams a survey too."""
    _log("rppi default output path")
    p, edges, npi, box = _box_case(seed=7)
    cnt_b, _, _ = oracle.count_rppi(p, p, p, edges, npi, box)
    d, r = survey_cap(1500, 10, seed=11)
    d[:, :3] = 1000.0 + (d[:, :3] - 1000.0) * 0.1
    sbox, ori = prep_cat_survey(d, r)
    sedges, snpi = np.array([0.0, 2.0, 4.0, 6.0, 9.0]), 12
    d[:, 3] = 1.0
    cnt_s, _, _ = oracle.count_rppi(d, d, d, sedges, snpi, sbox, survey=True, ori=ori)
    monkeypatch.delenv("NC_RPPI_STREAM", raising=False)
    with Context(0) as c2:
        _, got = c2.count(p, p, p, edges, box, npi=npi)
        assert c2.timing().hist_layout == 4 and np.array_equal(got, cnt_b)
        _, got = c2.count(d, d, d, sedges, sbox, survey=True, npi=snpi, ori=ori)
        assert c2.timing().hist_layout == 3 and np.array_equal(got, cnt_s)
    monkeypatch.setenv("NC_RPPI_STREAM", "1")
    with Context(0) as c2:
        _, got = c2.count(d, d, d, sedges, sbox, survey=True, npi=snpi, ori=ori)
        assert c2.timing().hist_layout == 4 and np.array_equal(got, cnt_s)
