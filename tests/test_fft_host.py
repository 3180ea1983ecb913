"""CPU run of the library's device ring-FFT code against direct DFT sums.

tests/fft_host/fft_host_check.cpp compiles shtools_amd/csrc/sht_device.hpp
for the host (one thread runs every butterfly; barriers are no-ops) with the
plans of shtools_amd/csrc/fftplan.hpp, so the stage code of every radix,
the Bluestein turn and the pruned plans are checked here
on every ring length the transforms use, without a GPU: DH rings
n = 4(lmax+1) (sampling 2) and 2(lmax+1) (sampling 1, complex grids), GLQ
rings n = 2 lmax + 1, plus lengths that exercise each radix and the prime
fallbacks.  Outputs are compared with long-double direct sums (48 sampled
orders per transform; all of them for short rings).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "fft_host", "fft_host_check.cpp")
TOL = 1e-13  # max |X - X_direct| / max |X_direct|, FP64 FFT


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("fft_host") / "fft_host_check")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wno-attributes", "-I", os.path.join(HERE, "fft_host"),
                        "-Wno-unknown-pragmas", "-o", exe, SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, lengths):
    r = subprocess.run([exe] + [str(n) for n in lengths], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [ln.split() for ln in r.stdout.strip().splitlines()]
    assert len(rows) == len(lengths)
    return rows


def _check(rows):
    bad = [(row[0], row[-1]) for row in rows if not float(row[-1]) <= TOL]
    assert not bad, bad


def test_dh2_rings(checker):
    """Every 4th DH2 ring length up to lmax 1100 and the bench extremes."""
    Ls = sorted(set(list(range(0, 1101, 4)) + [1, 2, 3, 63, 64, 1022, 1023, 1024]))
    _check(_run(checker, [4 * (L + 1) for L in Ls]))


def test_dh1_and_glq_rings(checker):
    Ls = sorted(set(list(range(1, 700, 9)) + [2, 3, 4, 5, 6, 7, 127, 128, 511, 512]))
    _check(_run(checker, [2 * (L + 1) for L in Ls] + [2 * L + 1 for L in Ls]))


def test_radix_coverage(checker):
    """Lengths that are products of each register radix, radix 7, and primes
    that force Bluestein."""
    lengths = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 15, 16, 49, 343, 2401, 2 * 7 ** 3, 4096, 3 ** 7, 5 ** 5,
               2 ** 6 * 3 ** 2 * 5, 11, 13, 17, 97, 1009, 4099, 8191, 6 * 1021]
    rows = _run(checker, lengths)
    _check(rows)
    radices = {int(r) for row in rows for r in row[4:-1]}
    assert {2, 3, 4, 5, 7, 8} <= radices, radices
    assert any(row[2] == "1" for row in rows)  # Bluestein plans present


def test_high_degree_rings(checker):
    """Ring lengths of degrees 2047..2800 (the high-degree workload)."""
    _check(_run(checker, [4 * (L + 1) for L in (2047, 2048, 2399, 2400, 2799)]))


def test_prime_radix_plans(checker):
    """Direct plans ending with a prime radix (11 .. 23, the last stage, run
    as symmetric sums over compile-time roots): ring lengths p * k for
    smooth k, DH and GLQ alike, and their Bluestein fallbacks (two such
    primes, or rings beyond the LDS)."""
    primes = [11, 13, 17, 19, 23]
    lengths = [p * k for p in primes for k in (1, 2, 3, 4, 6, 8, 12, 20, 64, 100, 420)]
    rows = _run(checker, lengths + [11 * 13 * 4, 23 * 512 * 2])
    _check(rows)
    for row in rows[:len(lengths)]:
        n, blue, last = int(row[0]), int(row[2]), int(row[-2])
        if n <= 10240:
            assert blue == 0 and last in primes, row
    assert int(rows[-2][2]) == 1 and int(rows[-1][2]) == 1
