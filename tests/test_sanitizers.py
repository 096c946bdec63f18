"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

* The CPU oracle (oracle/oracle.c) built with the sanitizers, driven through
  every entry point by tests/cpp/oracle_main.c (edge cases included); its
  output must equal the same driver built without them.
* The host libraries of the drop-in classes (ncorr.cpp, survey.cpp,
  utils.cpp) built with the sanitizers against a C-ABI stand-in
  (tests/cpp/nc_stub.c, no GPU runtime): the constructor, survey and helper
  paths run clean, and a device allocation failure reported by the C-ABI
  (NC_ERR_NOMEM) reaches the caller of calculateD1D2D3 as std::runtime_error
  with the library's message, as the reference's gpuErrchk throws
  (gpuerrchk.h:9-14); a bad argument (NC_ERR_ARG) as std::invalid_argument.
Host code only: the sanitizers are never applied to GPU code.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(ROOT, "cudancorr_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="needs gcc and g++")


def _run(exe, *args, env=None):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300, env=env or ENV)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    return r


@pytest.fixture(scope="module")
def oracle_bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("san_oracle")
    srcs = [os.path.join(ROOT, "oracle", "oracle.c"), os.path.join(ROOT, "tests", "cpp", "oracle_main.c")]
    common = ["-std=c99", "-fopenmp", "-ffp-contract=off", "-fno-fast-math"]
    san, plain = str(d / "oracle_san"), str(d / "oracle_plain")
    subprocess.run(["gcc", *common, *SAN, *srcs, "-lm", "-o", san], check=True)
    subprocess.run(["gcc", *common, "-O2", *srcs, "-lm", "-o", plain], check=True)
    return san, plain


def test_oracle_under_sanitizers(oracle_bins):
    san, plain = oracle_bins
    out = _run(san)
    assert out.returncode == 0, out.stdout + out.stderr
    ref = subprocess.run([plain], capture_output=True, text=True, check=True).stdout
    assert out.stdout == ref
    rows = {line.split()[0]: line.split()[1:] for line in out.stdout.splitlines()}
    assert rows["box_ppp"][0] == "0" and int(rows["box_ppp"][1]) > 0
    assert rows["empty"][:2] == ["0", "0"]
    assert rows["bad_edges"][0] != "0"  # rejected, no access past the edges
    assert int(rows["compact"][0]) > 0


@pytest.fixture(scope="module")
def host_bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("san_host")
    stub_o = str(d / "nc_stub.o")
    subprocess.run(["gcc", "-std=c99", *SAN, "-I" + INC, "-c", os.path.join(ROOT, "tests", "cpp", "nc_stub.c"),
                    "-o", stub_o], check=True)
    libs = [os.path.join(CSRC, f) for f in ("ncorr.cpp", "survey.cpp", "utils.cpp")]
    exes = {}
    for name in ("dropin_main", "utils_main"):
        exe = str(d / (name + "_san"))
        subprocess.run(["g++", "-std=c++17", *SAN, "-fopenmp", "-I" + INC,
                        os.path.join(ROOT, "tests", "cpp", name + ".cpp"), *libs, stub_o, "-o", exe], check=True)
        exes[name] = exe
    return exes


def test_dropin_host_paths_under_sanitizers(host_bins):
    exe = host_bins["dropin_main"]
    r = _run(exe, "ctor")
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
    r = _run(exe, "survey")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "box 4 8 18" in r.stdout


def test_allocation_failure_reaches_caller_as_runtime_error(host_bins, tmp_path):
    exe = host_bins["dropin_main"]
    p = tmp_path / "P.bin"
    np.array([[1.0, 1.0, 1.0, 1.0], [2.0, 1.0, 1.0, 1.0]]).tofile(p)
    args = ["count", p, p, p, "0", "0", "10", "10", "10", "0", "1", "2"]
    r = _run(exe, *args, env=dict(ENV, NC_STUB_STATUS="4"))  # NC_ERR_NOMEM
    assert r.returncode == 3
    assert r.stdout.startswith("EXCEPTION runtime_error calculateD1D2D3: stub: device allocation failed")
    r = _run(exe, *args, env=dict(ENV, NC_STUB_STATUS="5"))  # NC_ERR_HIP
    assert r.stdout.startswith("EXCEPTION runtime_error")
    r = _run(exe, *args, env=dict(ENV, NC_STUB_STATUS="1"))  # NC_ERR_ARG
    assert r.stdout.startswith("EXCEPTION invalid_argument")


def test_utils_under_sanitizers(host_bins, tmp_path):
    exe = host_bins["utils_main"]
    r = _run(exe, "logspace", 0.1, 20, 13)
    assert r.returncode == 0 and len(r.stdout.split()) == 13
    r = _run(exe, "randpos", 7, 5, 10, 20, 30)
    assert r.returncode == 0 and len(r.stdout.splitlines()) == 5
    cat = tmp_path / "c.bin"
    r = _run(exe, "randcat", cat, 70000, 50, 50, 50, 3)  # more than one 65536-point block
    assert r.returncode == 0 and cat.stat().st_size == 70000 * 32
    rows = tmp_path / "rows.bin"
    np.array([[1.0, 2.0, 3.0, 4.0], [0.5, 0.5, 0.75, 7.0]]).tofile(rows)
    out = tmp_path / "o.txt"
    r = _run(exe, "write3d", out, rows)
    assert r.returncode == 0 and out.read_text().count("\n") == 2
    txt = tmp_path / "in.txt"
    txt.write_text("# comment\n1 2 3 0.5\n4 5 6 1.5\n")
    r = _run(exe, "readtxt", txt, 10, 10, 10, 0, 1, 2, 3)
    assert r.returncode == 0 and len(r.stdout.splitlines()) == 2
    r = _run(exe, "readtxt", tmp_path / "missing.txt", 10, 10, 10, 0, 1, 2)
    assert r.returncode == 3 and r.stdout.startswith("EXCEPTION")
