"""ctypes binding of libelp_stats.so (polymer statistics of ELP configurations)."""
import ctypes as C
import os

import numpy as np

from .core import LIB_DIR, PKG_DIR, ElpError

_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(LIB_DIR, "libelp_stats.so")
        if not os.path.exists(path):
            raise ElpError(f"{path} is not built: run `make -C {PKG_DIR}`")
        lib = C.CDLL(path)
        P, I = C.c_void_p, C.c_int
        lib.elp_series_new.restype = P
        lib.elp_series_new.argtypes = [I, I, I, I, I, I, I]
        lib.elp_series_free.argtypes = [P]
        lib.elp_series_add.argtypes = [P, P, P, I]
        lib.elp_series_add_tuv.argtypes = [P, C.c_char_p]
        lib.elp_series_frames.argtypes = [P]
        lib.elp_series_positions.argtypes = [P, I, I, P]
        lib.elp_series_rg2.argtypes = [P, I, P]
        lib.elp_series_statics.argtypes = [P, I, C.POINTER(C.c_double), C.POINTER(C.c_double), P]
        lib.elp_series_msd.argtypes = [P, I, I, P, P]
        lib.elp_series_msd_reference.argtypes = [P, I, I, P, P]
        lib.elp_series_new_summary.restype = P
        lib.elp_series_new_summary.argtypes = [I, I, I, I, I, I, I]
        lib.elp_series_add_summary.argtypes = [P, P, P, P, P, P, P]
        lib.elp_series_summary.argtypes = [P, I, P, P, P, P, P, P]
        _lib = lib
    return _lib


class Series:
    """Frames of one system (n_chain chains of n_mono monomers) and their statistics."""

    def __init__(self, ring, L, n_chain, n_mono, max_frames, summary=False):
        """summary: keep per-chain reductions of each frame (add_summary) instead of all positions."""
        self.lib = _load()
        self.ring, self.L = bool(ring), tuple(L)
        self.n_chain, self.n_mono = int(n_chain), int(n_mono)
        new = self.lib.elp_series_new_summary if summary else self.lib.elp_series_new
        h = new(int(ring), *self.L, self.n_chain, self.n_mono, int(max_frames))
        if not h:
            raise ElpError("elp_series_new failed")
        self.h = C.c_void_p(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.elp_series_free(self.h)
            self.h = None

    __del__ = close

    @property
    def frames(self):
        return self.lib.elp_series_frames(self.h)

    def add(self, starts, bonds):
        """starts: (n_chain, 3) int; bonds: (n_chain, >= n_mono-1) bond codes."""
        s = np.ascontiguousarray(starts, dtype=np.int32)
        b = np.ascontiguousarray(bonds, dtype=np.uint8)
        if self.lib.elp_series_add(self.h, s.ctypes.data, b.ctypes.data, b.shape[1]):
            raise ElpError("elp_series_add failed")

    def add_host(self, host):
        """Add the chains recovered by a RingHost / LinHost (after set_lattice)."""
        starts = np.zeros((self.n_chain, 3), np.int32)
        bonds = np.zeros((self.n_chain, self.n_mono), np.uint8)
        for c in range(self.n_chain):
            st, b = host.polymer(c)
            starts[c] = st
            bonds[c, :len(b)] = b[:self.n_mono]
        self.add(starts, bonds)

    def add_summary(self, sm):
        """Add one frame given as per-chain reductions (a dict as Engine.ring_frame_summary returns)."""
        a = [np.ascontiguousarray(sm["start3"], np.int32), np.ascontiguousarray(sm["sum3"], np.int64),
             np.ascontiguousarray(sm["mid3"], np.int32), np.ascontiguousarray(sm["rg2"], np.float64),
             np.ascontiguousarray(sm["nsl"], np.int32), np.ascontiguousarray(sm["contacts"], np.uint64)]
        if a[0].shape != (self.n_chain, 3) or a[5].shape[0] < self.n_mono:
            raise ElpError("summary shape does not match the series")
        if self.lib.elp_series_add_summary(self.h, *(x.ctypes.data for x in a)):
            raise ElpError("elp_series_add_summary failed")

    def summary(self, f):
        """Per-chain reductions of frame f of a full series (what the device returns for that frame)."""
        n = self.n_chain
        out = {"start3": np.zeros((n, 3), np.int32), "sum3": np.zeros((n, 3), np.int64),
               "mid3": np.zeros((n, 3), np.int32), "rg2": np.zeros(n), "nsl": np.zeros(n, np.int32),
               "contacts": np.zeros(self.n_mono, np.uint64)}
        if self.lib.elp_series_summary(self.h, int(f), *(out[k].ctypes.data for k in ("start3", "sum3", "mid3", "rg2",
                                                                                       "nsl", "contacts"))):
            raise ElpError("elp_series_summary failed")
        return out

    def add_tuv(self, path):
        rc = self.lib.elp_series_add_tuv(self.h, os.fsencode(path))
        if rc:
            raise ElpError(f"elp_series_add_tuv({path}) = {rc}")

    def positions(self, f, c):
        out = np.empty((self.n_mono, 3), np.int32)
        if self.lib.elp_series_positions(self.h, int(f), int(c), out.ctypes.data):
            raise ElpError("bad frame/chain")
        return out

    def rg2(self, f):
        out = np.empty(self.n_chain)
        if self.lib.elp_series_rg2(self.h, int(f), out.ctypes.data):
            raise ElpError("bad frame")
        return out

    def statics(self, n_therm):
        rg2, sl = C.c_double(), C.c_double()
        pc = np.zeros(self.n_mono)
        if self.lib.elp_series_statics(self.h, int(n_therm), C.byref(rg2), C.byref(sl), pc.ctypes.data):
            raise ElpError("elp_series_statics failed")
        return {"rg2": rg2.value, "slrat": sl.value, "pc": pc}

    def msd(self, n_therm, n_out, convention="drift"):
        """g1, g3 at lags 0..n_out-1 frames from frame n_therm on.  convention
        "drift" removes the whole displacement of the system mean (this
        library's default); "reference" removes it as the reference analysis
        does (g1 half of it, g3 1/sqrt(2): elp_series_msd_reference)."""
        fn = {"drift": self.lib.elp_series_msd, "reference": self.lib.elp_series_msd_reference}[convention]
        g1, g3 = np.zeros(n_out), np.zeros(n_out)
        if fn(self.h, int(n_therm), int(n_out), g1.ctypes.data, g3.ctypes.data):
            raise ElpError("elp_series_msd failed")
        return g1, g3
