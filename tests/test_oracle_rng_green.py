"""Pins the oracle's RNG and Green CDF.

Philox4x32-10: Random123 known-answer vectors (kat_vectors) and rocRAND.
Green CDF: closed-form invariants of the square's harmonic measure and the
reference's own sample of 10 000 panel indices (randomWalk/rand.data).
"""
import math
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest
from scipy import stats

import oracle_ffi as O
from conftest import GOLDEN

# Random123 kat_vectors, philox4x32_10 (counter, key, expected output)
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_kat(oracle_lib, ctr, key, want):
    assert list(O.philox(ctr, key)) == want


def test_walk_stream_layout(oracle_lib):
    """Hop h of walk w reads word h%4 of Philox((w, job, h/4, 'WALK'), seed)."""
    L = O.oracle()
    seed = (7 << 32) | 12345
    for w, job, h in [(0, 0, 0), (5, 3, 9), (123456, 0, 1023)]:
        block = O.philox([w, job, h >> 2, 0x57414C4B], [seed & 0xFFFFFFFF, seed >> 32])
        assert L.orc_walk_u32(seed, w, job, h) == block[h & 3]


def test_philox_matches_rocrand(oracle_lib):
    """rocRAND's philox4x32_10 engine (open source, AMD) gives the same blocks."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    hdr = "/opt/rocm/include/rocrand/rocrand_philox4x32_10.h"
    if not (os.path.exists(hipcc) and os.path.exists(hdr)):
        pytest.skip("rocRAND headers not available")
    src = r"""
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
  for (unsigned long long seed : {0ull, 1ull, 0x0000000700003039ull}) {
    rocrand_device::philox4x32_10_engine e(seed, 0, 0);
    for (int i = 0; i < 8; ++i) { uint4 v = e.next4(); printf("%u %u %u %u\n", v.x, v.y, v.z, v.w); }
  }
}
"""
    with tempfile.TemporaryDirectory() as d:
        cpp = os.path.join(d, "k.cpp")
        exe = os.path.join(d, "k")
        open(cpp, "w").write(src)
        r = subprocess.run([hipcc, "-O1", "-x", "hip", "--offload-arch=gfx950", cpp, "-o", exe],
                           capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("could not compile the rocRAND probe: " + r.stderr[-300:])
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = np.array(out, dtype=np.uint64).reshape(3, 8, 4)
    for s, seed in enumerate([0, 1, 0x0000000700003039]):
        for i in range(8):
            # rocRAND engine: counter = (i, 0, 0, 0), key = (seed lo, seed hi)
            want = O.philox([i, 0, 0, 0], [seed & 0xFFFFFFFF, seed >> 32])
            assert list(got[s, i]) == list(want)


def _side_density(s, terms=10):
    n = np.arange(1, 2 * terms, 2)
    sgn = np.where(n % 4 == 1, 1.0, -1.0)
    return np.sum(sgn[:, None] * np.sin(np.outer(n, np.pi * np.atleast_1d(s))) /
                  np.cosh(n * np.pi / 2)[:, None], axis=0)


def test_green_cdf_invariants(oracle_lib):
    intg = O.green_cdf()
    assert intg[0] == 0.0 and intg[200] == 1.0
    assert np.all(np.diff(intg) > 0)
    for k in range(5):
        assert abs(intg[50 * k] - k / 4) < 1e-15
    w = np.diff(intg)
    side = w[:50]
    assert np.allclose(side, side[::-1], rtol=0, atol=1e-16)          # mirror symmetry
    assert np.allclose(w.reshape(4, 50), np.tile(side, (4, 1)), rtol=0, atol=1e-15)  # equal sides
    # the exact side integral of the truncated series (1/4 for the full series)
    n = np.arange(1, 20, 2)
    exact = np.sum(np.where(n % 4 == 1, 1.0, -1.0) * 2 / (n * np.pi * np.cosh(n * np.pi / 2)))
    assert abs(exact - 0.25) < 1e-12
    # panel weights are midpoint-rule samples of that density
    mids = (np.arange(50) + 0.5) / 50
    dens = _side_density(mids)
    assert np.allclose(side / side.sum(), dens / dens.sum(), rtol=1e-12, atol=0)
    # density at the side centre: sum over odd n of 1/cosh(n pi/2)
    assert abs(_side_density(0.5)[0] - 0.41731) < 5e-5


def test_green_cdf_matches_reference_sample(oracle_lib):
    """randomWalk/rand.data: 10 000 panel indices drawn by the reference."""
    d = np.array(open(os.path.join(GOLDEN, "rand.data")).read().split(), dtype=int)
    assert len(d) == 10000 and d.min() >= 0 and d.max() < 200
    p = np.diff(O.green_cdf())
    cnt = np.bincount(d, minlength=200)

    def chi2(q):
        e = q * len(d)
        return float(((cnt - e) ** 2 / e).sum())

    assert stats.chi2.sf(chi2(p), 199) > 0.01
    # plausible mistakes are rejected by the same sample
    assert stats.chi2.sf(chi2(np.roll(p, 1)), 199) < 1e-4       # off-by-one panel
    assert stats.chi2.sf(chi2(np.full(200, 1 / 200)), 199) < 1e-10  # flat kernel


def test_thresholds_and_sampling(oracle_lib):
    L = O.oracle()
    intg = O.green_cdf()
    T = O.green_thresholds()
    assert T[0] == 0 and np.all(np.diff(T.astype(np.int64)) > 0)
    assert np.array_equal(T, np.array([round(x * 2 ** 32) for x in intg[:200]], dtype=np.uint32))
    rng = np.random.default_rng(3)
    us = rng.integers(0, 2 ** 32, size=20000, dtype=np.uint64)
    tp = T.ctypes.data_as(O._U32)
    ip = intg.ctypes.data_as(O._D)
    mism = 0
    for u in us:
        a = L.orc_sample_index(tp, int(u))
        b = L.orc_get_index_fp(ip, float(u) / 2 ** 32)   # reference getIndex on r = u/2^32
        mism += a != b
        assert T[a] <= u and (a == 199 or T[a + 1] > u)
    assert mism == 0
    # boundaries: u = T[i] selects panel i, u = T[i]-1 selects i-1
    for i in range(1, 200):
        assert L.orc_sample_index(tp, int(T[i])) == i
        assert L.orc_sample_index(tp, int(T[i]) - 1) == i - 1
    assert L.orc_sample_index(tp, 2 ** 32 - 1) == 199
