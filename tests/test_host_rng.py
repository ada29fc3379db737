"""The host random stream of the drop-in (start-up draws, RedistribSL
shuffle, tile-grid shifts) is the LFSR113 step of gpupol2/gpupol.h:84-95,
which is util/rng.h:10-23, the generator of efpol.  The efpol restatement is
pinned byte-exact against the reference efpol's golden tuv sequences, so its
generator is the reference's; the host generator must give the same stream
from the same state."""
import ctypes as C
import os
import subprocess

import numpy as np

import elp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_lfsr113_equals_efpol_generator():
    lib = C.CDLL(os.path.join(elp.PKG_DIR, "lib", "libgpupol3_ring.so"))
    lib.elp_host_lfsr113.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    exe = os.path.join(ROOT, "oracle", "bin", "efpol_port_ring")
    rng = np.random.default_rng(7)
    # states as the seedings make them (every word above the generator's minimum of 2, 8, 16, 128)
    states = [np.array([s ^ 85127698, s ^ 12369123, s ^ 81261212, s ^ 51203016], np.uint32) for s in (0, 1, 239864012)]
    states += [rng.integers(1 << 8, 1 << 32, size=4, dtype=np.uint64).astype(np.uint32) for _ in range(5)]
    for st in states:
        n = 2000
        out = np.zeros(n, np.uint32)
        s = st.copy()
        assert lib.elp_host_lfsr113(s.ctypes.data, out.ctypes.data, n) == 0
        ref = subprocess.run([exe, "rng", *(str(int(x)) for x in st), str(n)], capture_output=True, text=True,
                             check=True).stdout.split()
        assert np.array_equal(out, np.array(ref, dtype=np.uint64).astype(np.uint32)), st
