"""C-ABI and drop-in host logic -- CPU only (no compute call needs a GPU).

* libncorr_hip.so loads and exports every function include/ncorr_hip.h declares;
* size helpers;
* the C++ drop-in classes (include/cudaNcorr.hpp, include/survey.hpp) compile
  with g++ against libncorr.so, construct without a GPU and keep the
  reference's output bin list, errors and survey preparation
  (source/cudaNcorr.cu:118-157, source/survey.cu:60-128);
* the comoving distance (survey.cu:26-37) matches scipy.integrate.quad, which
  wraps the same QUADPACK QAGS algorithm as GSL, to 1e-7 relative.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from cudancorr_amd import _lib
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ncorr_hip.h")
LIBDIR = os.path.join(ROOT, "cudancorr_amd", "lib")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(nc_\w+)\s*\(", text, re.M)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(_lib.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_layouts_match_binding(tmp_path):
    """Every field offset and size of nc_params, nc_timing and nc_tuning, as a
    C compiler lays them out from include/ncorr_hip.h, equals the ctypes
    binding's."""
    structs = {"nc_params": _lib.NcParams, "nc_timing": _lib.NcTiming, "nc_tuning": _lib.NcTuning}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "ncorr_hip.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'    printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "layout")
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", exe], check=True)
    got = {}
    for line in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines():
        cname, f, v = line.split()
        got[(cname, f)] = int(v)
    for cname, cls in structs.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


def test_size_helpers():
    L = _lib.lib()
    for ne, nh, nc in ((11, 1000, 220), (16, 3375, 680), (21, 8000, 1540), (26, 15625, 2925), (1, 0, 0)):
        assert L.nc_hist_size(ne) == nh
        assert L.nc_compact_size(ne) == nc


def test_hist_len():
    L = _lib.lib()
    e = np.arange(0.0, 31.0, 2.0)
    p, _ = _lib.Context.params(e, (100.0, 100.0, 100.0))
    assert L.nc_hist_len(ctypes.byref(p)) == 15 ** 3
    p, _ = _lib.Context.params(e, (100.0, 100.0, 100.0), npi=20)
    assert p.bin_type == _lib.NC_BIN_RPPI
    assert L.nc_hist_len(ctypes.byref(p)) == 15 ** 3 * 20 * 20
    p.num_pi_bins = 0
    assert L.nc_hist_len(ctypes.byref(p)) == 0
    assert L.nc_hist_len(None) == 0


def test_null_arguments_are_rejected():
    L = _lib.lib()
    assert L.nc_create(None, 0) != 0
    assert L.nc_destroy(None) != 0
    assert L.nc_catalog_destroy(None) != 0
    assert L.nc_catalog_size(None) == -1
    assert L.nc_last_timing(None, None) != 0


@pytest.fixture(scope="module")
def dropin(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("dropin") / "dropin_main")
    src = os.path.join(ROOT, "tests", "cpp", "dropin_main.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), src, "-L" + LIBDIR,
                    "-lncorr", "-lncorr_hip", "-Wl,-rpath," + LIBDIR, "-o", exe], check=True)
    return exe


def test_dropin_constructor(dropin):
    out = subprocess.run([dropin, "ctor"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("OK")


def test_dropin_survey_prep_and_distance(dropin):
    out = subprocess.run([dropin, "survey"], capture_output=True, text=True, check=True).stdout
    rows = {}
    for line in out.splitlines():
        k, *v = line.split()
        rows.setdefault(k, []).append([float(x) for x in v])
    # joint min (9,-5,3), max (12,2,20): box = extent + 1, ori = -min, shift by -min
    assert rows["box"][0] == [4.0, 8.0, 18.0]
    assert rows["ori"][0] == [-9.0, 5.0, -3.0]
    assert rows["d0"][0] == [1.0, 0.0, 0.0]
    assert rows["r1"][0] == [2.0, 7.0, 17.0]
    from scipy.integrate import quad

    om, ol = 0.313772, 0.686228
    for z, dc in rows["dc"]:
        val, _ = quad(lambda t: 1.0 / np.sqrt(om * (1 + t) ** 3 + ol), 0.0, z, epsabs=0.0, epsrel=1e-10, limit=1000)
        want = 299792.458 * val / 100.0
        assert abs(dc - want) <= 1e-7 * want
    x, y, zc, w = rows["xyz"][0]
    d = rows["dc"][1][1]  # z = 0.5
    ra, dec = np.deg2rad(30.0), np.deg2rad(20.0)
    np.testing.assert_allclose([x, y, zc], [d * np.cos(dec) * np.cos(ra), d * np.cos(dec) * np.sin(ra),
                                            d * np.sin(dec)], rtol=1e-14)
    assert w == 1.5


def test_output_bins_match_oracle():
    from cudancorr_amd.gen import edges_linear

    for e in (edges_linear(2.0, 20.0, 10), edges_linear(0.0, 50.0, 25)):
        assert oracle.output_bins(e).shape[0] == oracle.compact(e, np.zeros((len(e) - 1) ** 3)).size
