"""elp_ring_frame_summary_noalloc, whose device buffers come from the
library's static scratch array instead of hipMalloc, against
elp_ring_frame_summary: identical results on two calls in a row, the second
with a smaller input (fewer, shorter rings in a smaller box), a repeated call
on the scratch the previous call left behind, and the size limit."""
import ctypes as C
import sys

import numpy as np
import pytest

import elp

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(got, want):
    for k in want:
        assert np.array_equal(got[k], want[k]), k


def test_noalloc_frame_summary_matches_allocating_entry_point():
    tile = (16, 16, 16)
    for L, N, seed in ((64, 200, 3), (32, 100, 5)):
        _say(f"[no-alloc frame summary] ring L={L} N={N} seed={seed}")
        h = elp.RingHost(N=N, L=L, density=1.2, seed=seed)
        eng = elp.Engine("ring", h.L, tile, elp.SUBLAT_DENSE, 13)
        eng.upload(h.lattice())
        h.load_shift(tile)
        eng.run(h.origins(tile, 50), tau0=0, seed=h.seed)
        want = eng.ring_frame_summary(h)
        assert want["contacts"].sum() > 0 and np.all(want["rg2"] > 0)
        _same(eng.ring_frame_summary(h, noalloc=True), want)
        _same(eng.ring_frame_summary(h, noalloc=True), want)
        eng.close()
        h.close()


def test_noalloc_frame_summary_rejects_sizes_beyond_its_scratch():
    h = elp.RingHost(N=100, L=32, density=1.2, seed=1)
    eng = elp.Engine("ring", h.L, (16, 16, 16), elp.SUBLAT_DENSE, 13)
    eng.upload(h.lattice())
    n_pol, max_len = 40000, 1000                      # about 44 MB of buffers, over the 32 MiB scratch
    l2i = np.full(1 << 18, -1, np.int32)
    i32 = np.zeros(3 * n_pol, np.int32)
    i64 = np.zeros(3 * n_pol, np.int64)
    lib = eng.lib
    rc = lib.elp_ring_frame_summary_noalloc(eng.h, l2i.ctypes.data, n_pol, max_len, i32.ctypes.data, i64.ctypes.data,
                                            i32.ctypes.data, None, None, None)
    assert rc == -1                                   # ELP_E_ARG
    assert b"scratch" in lib.elp_last_error()
    eng.close()
    h.close()
