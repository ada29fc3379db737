"""The tile-coordinate division of the move kernels (udiv_small in
mc-polymer-simulation_amd/csrc/elp_device.h): q = trunc(float(n) * rcp(float(d)))
corrected once by the remainder.  Its claim is exactness for n < 2^22 and
1 <= d < 2^22 when the hardware reciprocal (v_rcp_f32) is within one ulp of
1/d.  Checked here in float32 arithmetic for every tile count d up to 2^11
(a box edge over a tile edge), with the reciprocal taken as the correctly
rounded 1/d and as each float neighbour of it, against integer division, on
n at and around every multiple of d near the top of the range and on random
n below 2^22."""
import numpy as np


def _udiv_small(n, d, rcp):
    q = (n.astype(np.float32) * rcp).astype(np.float32)
    q = np.trunc(q).astype(np.int64)
    r = n.astype(np.int64) - q * d
    return np.where(r < 0, q - 1, np.where(r >= d, q + 1, q))


def test_udiv_small_exact_for_tile_ranges():
    rng = np.random.default_rng(7)
    top = 1 << 22
    for d in range(1, 1 << 11):
        base = np.float32(1.0) / np.float32(d)
        # multiples of d near 2^22 and their neighbours, plus random n
        k = np.arange(top // d - 64, top // d + 1, dtype=np.int64)
        n = np.concatenate([k * d - 1, k * d, k * d + 1, rng.integers(0, top, 256), [0, 1, top - 1]])
        n = n[(n >= 0) & (n < top)].astype(np.uint32)
        for rcp in (np.nextafter(base, np.float32(0)), base, np.nextafter(base, np.float32(1))):
            q = _udiv_small(n, d, np.float32(rcp))
            assert np.array_equal(q, n.astype(np.int64) // d), d
