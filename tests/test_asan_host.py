"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/asan/host_check.cpp runs, built with -fsanitize=address,undefined and
libstdc++'s container assertions, everything the C interface computes on the
host before it launches a kernel: the recursion, seed, quadrature and
polar-skip tables (shtools_amd/csrc/tables.hpp), the ring FFT plans
(fftplan.hpp), the batched plans' geometry, field offsets and work-item
lists (layout.hpp) and the argument checks of the plain transforms
(checks.hpp, including extreme integer arguments), and checks the
invariants the kernels index by: offsets inside their arrays, each unit of
work covered exactly once, LDS classes large enough for their members.
tests/fft_host/fft_host_check.cpp, the device ring-FFT stage code run on the
host, is built with the same sanitizers and run on a sample of ring lengths
of each plan kind (direct, Bluestein, pruned, split).  Any sanitizer report
fails the test.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FLAGS = ["-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-Wno-attributes",
         "-fno-omit-frame-pointer", "-D_GLIBCXX_ASSERTIONS", "-Wno-unknown-pragmas"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _build(tmp_path_factory, name, src, extra=()):
    exe = str(tmp_path_factory.mktemp("asan") / name)
    r = subprocess.run(["g++", *FLAGS, *extra, "-o", exe, src], capture_output=True, text=True)
    if r.returncode != 0 and "asan" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("g++ without the sanitizer runtimes")
    assert r.returncode == 0, r.stderr
    return exe


@pytest.fixture(scope="module")
def host_check(tmp_path_factory):
    return _build(tmp_path_factory, "host_check", os.path.join(HERE, "asan", "host_check.cpp"))


@pytest.fixture(scope="module")
def fft_check(tmp_path_factory):
    return _build(tmp_path_factory, "fft_host_check", os.path.join(HERE, "fft_host", "fft_host_check.cpp"),
                  ["-I", os.path.join(HERE, "fft_host")])


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    return r.stdout


@pytest.mark.parametrize("section", ["tables", "plans", "layout", "checks"])
def test_host_code(host_check, section):
    out = _run([host_check, section])
    assert out.startswith(f"ok {section} "), out


def test_device_fft_code_on_host(fft_check):
    # direct lengths, Bluestein GLQ rings, pruned DH rings and the split plans
    # of the largest DH rings (lmax 1023 and 1100: n = 4096, 4404)
    lengths = [1, 2, 8, 49, 360, 2 * 7 ** 3, 11, 97, 1009, 2 * 63 + 1, 2 * 511 + 1, 4 * 301, 4 * 700, 4096, 4404,
               4 * 1290]
    out = _run([fft_check] + [str(n) for n in lengths])
    rows = [ln.split() for ln in out.strip().splitlines()]
    assert len(rows) == len(lengths)
    bad = [(row[0], row[-1]) for row in rows if not float(row[-1]) <= 1e-13]
    assert not bad, bad
