"""The traversal kernel replaces `float(size) / tChildExit > maxFootprint` (raytracing.cpp:276) by
an equivalent comparison when the footprint test is disabled (maxFootprint = -1). For a size
c = 2^k > 0 and IEEE round-to-nearest division, RN(c / t) > -1 exactly when t > 0, t = +0 or
t < -c; the kernel tests the first two cases as one unsigned comparison of t's bit pattern with
that of +inf. Checked here against numpy's IEEE float32 division on boundary, special and random values.
"""
import numpy as np


def fast(c, t):
    # mirrors traverse_core.h footprint_large() for max_footprint == -1: the bit pattern of t is at
    # most that of +inf (sign clear, not NaN), or t < -c
    return (t.view(np.uint32) <= np.uint32(0x7F800000)) | (t < -c)


def slow(c, t):
    # the three cases spelled out
    return (t > 0) | ((t == 0) & ~np.signbit(t)) | (t < -c)


def test_disabled_footprint_predicate_matches_division():
    rng = np.random.Generator(np.random.PCG64(0))
    for k in range(0, 31):
        c = np.float32(2.0 ** k)
        near = [c, -c]
        for base in (c, -c):
            v = np.float32(base)
            up = v
            dn = v
            for _ in range(64):
                up = np.nextafter(up, np.float32(np.inf))
                dn = np.nextafter(dn, np.float32(-np.inf))
                near += [up, dn]
        special = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 1.17549435e-38, -1.17549435e-38,
                   3.4028235e38, -3.4028235e38, 1.0, -1.0, 0.5, -0.5]
        rnd = np.concatenate([rng.uniform(-4 * float(c), 4 * float(c), 20000),
                              -np.exp(rng.uniform(-80, 80, 20000)), np.exp(rng.uniform(-80, 80, 20000))])
        t = np.asarray(near + special + list(rnd), dtype=np.float32)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            exact = (c / t) > np.float32(-1.0)
        assert exact.dtype == bool
        assert np.array_equal(exact, fast(c, t)), k
        assert np.array_equal(exact, slow(c, t)), k
        nans = np.array([0x7F800001, 0x7FC00000, 0x7FFFFFFF, 0xFF800001, 0xFFC00000], np.uint32).view(np.float32)
        assert not fast(c, nans).any()
