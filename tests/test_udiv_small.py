"""The intra kernels divide small sums by block sizes with `udiv_small`
(compute-vp9_amd/csrc/kernels.hip): truncate x * rcp(n) in float, then one
correction step each way. The GPU's v_rcp_f32 is within 1 ulp of 1/n; this
checks, over every x < 2^16 and n in [1, 128], that the result equals integer
division for every reciprocal within 2 ulp of the correctly rounded one."""
import numpy as np


def _emulate(x, n, r):
    q = np.trunc((x.astype(np.float32) * np.float32(r)).astype(np.float32)).astype(np.int64)
    rem = x - q * n
    return q + (rem >= n) - (rem < 0)


def test_udiv_small_exact_over_its_range():
    x = np.arange(0, 1 << 16, dtype=np.int64)
    for n in range(1, 129):
        r0 = np.float32(1.0) / np.float32(n)
        for d in (-2, -1, 0, 1, 2):
            r = r0
            for _ in range(abs(d)):
                r = np.nextafter(r, np.float32(np.inf) if d > 0 else np.float32(0))
            assert np.array_equal(_emulate(x, n, r), x // n), (n, d)


def test_udiv_small_needs_its_correction_steps():
    # the plain truncated product is off by one somewhere in the range: the
    # correction steps are not dead code
    x = np.arange(0, 1 << 16, dtype=np.int64)
    off = 0
    for n in range(1, 129):
        r = np.nextafter(np.float32(1.0) / np.float32(n), np.float32(0))
        q = np.trunc((x.astype(np.float32) * r).astype(np.float32)).astype(np.int64)
        off += int(np.count_nonzero(q != x // n))
    assert off > 0
