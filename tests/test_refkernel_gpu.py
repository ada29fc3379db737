"""The oracle's move rules against the reference kernels themselves.

oracle/_ref/refkernel_{ring,lin} run gpupol2/ringpol3.cu and gpupol/linpol.cu,
compiled unmodified from the reference tree, with the reference launch
sequence and seed table.  The oracle's reference-mode replays (same move
functions as the replays that check the gfx950 kernels, driven by the
reference's per-lane xorshift128 / LFSR113 states) must reproduce their
lattice and final RNG states bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

import elp
from ringcheck import ROOT, oracle

pytestmark = pytest.mark.gpu
REF = os.path.join(ROOT, "oracle", "_ref")


def _need(exe):
    path = os.path.join(REF, exe)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (needs the reference sources at build time)")
    return path


def _run(exe, args, tmp):
    print(f"[reference kernel] {exe} {args[:3]}", flush=True)
    subprocess.run([exe] + [str(a) for a in args], check=True, timeout=120, cwd=tmp)


@pytest.mark.parametrize("L,N,seed,ntime", [(32, 100, 1, 6), (64, 200, 4, 3)])
def test_ringpol3_reference_kernel_matches_oracle(tmp_path, L, N, seed, ntime):
    import ctypes as C
    exe = _need("refkernel_ring")
    h = elp.RingHost(N=N, L=L, density=1.2, seed=seed)
    lat0 = h.lattice()
    seeds0 = h.seed_table()
    h.load_shift((32, 32, 32))
    org = h.origins((32, 32, 32), ntime + 1)          # frame origins; org[0] = load shift
    d = np.empty_like(org)
    d[0] = org[0]
    d[1:] = (org[1:] - org[:-1]) % 32                 # shifts drawn per time unit
    lat0.tofile(tmp_path / "lat.in")
    seeds0.tofile(tmp_path / "seeds.in")
    d.astype(np.int32).tofile(tmp_path / "shift.in")
    _run(exe, [L, 51, ntime, "lat.in", "seeds.in", "shift.in", "lat.out", "seeds.out"], tmp_path)
    got = np.fromfile(tmp_path / "lat.out", dtype=np.uint8)
    got_seeds = np.fromfile(tmp_path / "seeds.out", dtype=np.uint32)

    lib = oracle()
    lib.orc_ring_run_ref.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    want = lat0.copy()
    want_seeds = seeds0.copy()
    # work-group w owns the tile at (accumulated shift mod L) + 32*w: the
    # binding of seed streams to tiles needs the shift mod L, not mod 32
    o = np.ascontiguousarray(np.cumsum(d, axis=0)[:ntime] % L, dtype=np.int32)
    assert lib.orc_ring_run_ref(want.ctypes.data, L, want_seeds.ctypes.data, 51, ntime, o.ctypes.data) == 0
    assert np.count_nonzero(want != lat0) > 1000
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} sites differ"
    assert np.array_equal(got_seeds, want_seeds)


@pytest.mark.parametrize("L,N,seed,ntime", [(32, 200, 2, 5), (64, 100, 6, 2)])
def test_linpol_reference_kernel_matches_oracle(tmp_path, L, N, seed, ntime):
    import ctypes as C
    exe = _need("refkernel_lin")
    h = elp.LinHost(N=N, L=L, density=1.2, seed=seed)
    lat0 = h.lattice()
    seeds0 = h.seed_table()
    org = h.origins((32, 32, 32), ntime)
    lat0.tofile(tmp_path / "lat.in")
    seeds0.tofile(tmp_path / "seeds.in")
    org.astype(np.int32).tofile(tmp_path / "off.in")
    _run(exe, [L, 64, ntime, "lat.in", "seeds.in", "off.in", "lat.out", "seeds.out"], tmp_path)
    got = np.fromfile(tmp_path / "lat.out", dtype=np.uint16)
    got_seeds = np.fromfile(tmp_path / "seeds.out", dtype=np.uint32)

    lib = oracle()
    lib.orc_lin_run_ref.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    want = lat0.copy()
    want_seeds = seeds0.copy()
    o = np.ascontiguousarray(org, dtype=np.int32)
    assert lib.orc_lin_run_ref(want.ctypes.data, L, want_seeds.ctypes.data, 64, ntime, o.ctypes.data) == 0
    assert np.count_nonzero(want != lat0) > 1000
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} sites differ"
    assert np.array_equal(got_seeds, want_seeds)
