"""The host-side N(0,1) transform of the seeded u64 stream (Box-Muller in f64,
rounded once): closed forms at chosen words and the moments of a sample."""
import math

import numpy as np

from massively_amd.random import normal_from_u64


def test_box_muller_closed_forms():
    # lo = 2^32 - 1 gives u1 = 1, so z = 0 whatever hi is.
    assert normal_from_u64(np.array([0xFFFFFFFF, 0x12345678FFFFFFFF], np.uint64)).tolist() == [0.0, 0.0]
    # hi = 0 gives cos(0) = 1: z = sqrt(-2 ln((lo + 1) / 2^32)).
    lo = 0x7FFFFFFF  # u1 = 1/2
    z = normal_from_u64(np.array([lo], np.uint64))[0]
    assert z == np.float32(math.sqrt(2 * math.log(2)))
    # hi = 2^31 gives cos(pi) = -1.
    z = normal_from_u64(np.array([(1 << 63) | lo], np.uint64))[0]
    assert z == np.float32(-math.sqrt(2 * math.log(2)))
    # hi = 2^30 gives cos(pi/2) = 0 (to f64 rounding).
    z = normal_from_u64(np.array([(1 << 62) | lo], np.uint64))[0]
    assert abs(float(z)) < 1e-15


def test_box_muller_moments():
    rng = np.random.default_rng(3)
    z = normal_from_u64(rng.integers(0, 2**64, 1 << 20, dtype=np.uint64)).astype(np.float64)
    assert abs(z.mean()) < 5e-3
    assert abs(z.std() - 1.0) < 5e-3
    assert abs(np.abs(z).mean() - math.sqrt(2 / math.pi)) < 5e-3
