"""ctypes bindings of libelp.so and the drop-in host libraries."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
REPO_DIR = os.path.dirname(PKG_DIR)
HEADER = os.path.join(REPO_DIR, "include", "elp_gpu.h")

SUBLAT_CUBIC = 0
SUBLAT_DENSE = 1
RULES_REFERENCE = 0      # ringpol3's move rates
RULES_BALANCED = 1       # forced TransBack at rate 1/4: uniform measure stationary


class ElpError(RuntimeError):
    pass


_libs = {}


class Schedule(C.Structure):
    """elp_schedule of include/elp_gpu.h."""
    _fields_ = [("tt", C.c_int), ("tu", C.c_int), ("tv", C.c_int), ("sublattice", C.c_int), ("n_iter", C.c_int),
                ("threads", C.c_int), ("units_per_graph", C.c_int), ("attempts_per_s", C.c_double),
                ("rules", C.c_int), ("accepted_per_s", C.c_double)]

    @property
    def tile(self):
        return (self.tt, self.tu, self.tv)

    def record(self):
        buf = C.create_string_buffer(256)
        _check(load_elp().elp_schedule_format(C.byref(self), buf, len(buf)))
        return buf.value.decode()

    @classmethod
    def parse(cls, text):
        s = cls()
        _check(load_elp().elp_schedule_parse(os.fsencode(text), C.byref(s)))
        return s


def autotune(kind, lattices, boxes, time_units=20, device=0, rules=None, sublattice=None):
    """Compiled schedule with the most accepted moves per second for these
    replicas, among schedules with the same statics (elp_autotune): under the
    reference rates (default) only `sublattice` (default dense) competes;
    under RULES_BALANCED sublattice=None lets both compete.  Returns
    (Schedule, log text).  The lattices are only read."""
    rules = RULES_REFERENCE if rules is None else rules
    if sublattice is None:
        sublattice = SUBLAT_DENSE if rules == RULES_REFERENCE else -1
    lib = load_elp()
    arrs = [np.ascontiguousarray(a, dtype=np.uint8 if kind == "ring" else np.uint16) for a in lattices]
    ptrs = (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    bx = np.ascontiguousarray(boxes, dtype=np.int32).reshape(-1)
    best = Schedule()
    log = C.create_string_buffer(8192)
    _check(lib.elp_autotune(device, 0 if kind == "ring" else 1, len(arrs), bx.ctypes.data, ptrs, int(time_units),
                            int(rules), int(sublattice), C.byref(best), log, len(log)), lib)
    return best, log.value.decode()


def _load(name):
    if name in _libs:
        return _libs[name]
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise ElpError(f"{path} is not built: run `make -C {PKG_DIR}` (or __graft_entry__.build())")
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL if name == "libelp.so" else C.RTLD_LOCAL)
    _libs[name] = lib
    return lib


def load_elp():
    lib = _load("libelp.so")
    if not getattr(lib, "_elp_typed", False):
        P, I, U32, U64 = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64
        lib.elp_last_error.restype = C.c_char_p
        lib.elp_device_count.argtypes = [C.POINTER(C.c_int)]
        for fn in ("elp_ring_create", "elp_lin_create"):
            getattr(lib, fn).argtypes = [C.POINTER(P), I, I, I, I, I, I, I, I, I]
        lib.elp_destroy.argtypes = [P]
        lib.elp_upload.argtypes = [P, P, U64]
        lib.elp_download.argtypes = [P, P, U64]
        lib.elp_run.argtypes = [P, I, P, U64, U32, U32]
        lib.elp_sync.argtypes = [P]
        lib.elp_last_run_ms.argtypes = [P, C.POINTER(C.c_float)]
        lib.elp_launch_info.argtypes = [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)]
        lib.elp_bench_copy.argtypes = [I, U64, I, C.POINTER(C.c_float)]
        for fn in ("elp_ring_batch_create", "elp_lin_batch_create"):
            getattr(lib, fn).argtypes = [C.POINTER(P), I, I, P, I, I, I, I, I]
        lib.elp_batch_destroy.argtypes = [P]
        lib.elp_batch_upload.argtypes = [P, I, P, U64]
        lib.elp_batch_download.argtypes = [P, I, P, U64]
        lib.elp_batch_run.argtypes = [P, I, P, P, P, I]
        lib.elp_batch_sync.argtypes = [P]
        lib.elp_batch_last_run_ms.argtypes = [P, C.POINTER(C.c_float)]
        lib.elp_batch_info.argtypes = [P, C.POINTER(I), C.POINTER(I), C.POINTER(U64)]
        lib.elp_batch_set_threads.argtypes = [P, I]
        lib.elp_set_threads.argtypes = [P, I]
        lib.elp_batch_set_rules.argtypes = [P, I]
        lib.elp_set_rules.argtypes = [P, I]
        lib.elp_ring_chains.argtypes = [P, P, I, I, P, P, P]
        lib.elp_ring_tuv.argtypes = [P, P, I, I, I, C.c_char_p, U64, C.POINTER(U64)]
        lib.elp_ring_chain_stats.argtypes = [P, P, I, I, P, P]
        lib.elp_ring_frame_summary.argtypes = [P, P, I, I, P, P, P, P, P, P]
        lib.elp_ring_frame_summary_noalloc.argtypes = [P, P, I, I, P, P, P, P, P, P]
        lib.elp_lin_frame_summary.argtypes = [P, P, I, I, P, P, P, P, P, P]
        lib.elp_batch_frame_summary.argtypes = [P, I, P, I, I, P, P, P, P, P, P]
        lib.elp_lin_chains.argtypes = [P, P, I, I, P, P, P]
        lib.elp_lin_tuv.argtypes = [P, P, I, I, I, C.c_char_p, U64, C.POINTER(U64)]
        lib.elp_net_create.argtypes = [C.POINTER(P), I, U32]
        lib.elp_net_destroy.argtypes = [P]
        lib.elp_net_params.argtypes = [P, P, U64]
        lib.elp_net_set_params.argtypes = [P, P, U64]
        lib.elp_net_forward.argtypes = [P, P, I, P]
        lib.elp_net_gradient.argtypes = [P, P, P, I, P, C.POINTER(C.c_float)]
        lib.elp_net_harvest.argtypes = [P, P, I, U32]
        lib.elp_net_samples.argtypes = [P, C.POINTER(I), P, P]
        lib.elp_net_train.argtypes = [P, I, I, C.c_float, C.POINTER(C.c_float)]
        lib.elp_run_learned.argtypes = [P, P, I, P, U64, U32, U32, C.POINTER(U64), C.POINTER(U64)]
        lib.elp_net_patch_moves.argtypes = [I, P, P, I, P, P, P]
        lib.elp_net_patch_features.argtypes = [I, P, I, P]
        lib.elp_net_tile_check.argtypes = [I, P, I, P, I, P]
        lib.elp_autotune.argtypes = [I, I, I, P, P, I, I, I, C.POINTER(Schedule), C.c_char_p, U64]
        lib.elp_count_moves.argtypes = [P, I]
        lib.elp_move_counts.argtypes = [P, P]
        lib.elp_batch_count_moves.argtypes = [P, I]
        lib.elp_batch_move_counts.argtypes = [P, P]
        lib.elp_schedule_format.argtypes = [C.POINTER(Schedule), C.c_char_p, U64]
        lib.elp_schedule_parse.argtypes = [C.c_char_p, C.POINTER(Schedule)]
        lib._elp_typed = True
    return lib


def _check(rc, lib=None):
    if rc != 0:
        lib = lib or load_elp()
        raise ElpError(f"elp error {rc}: {lib.elp_last_error().decode()}")


def device_count():
    n = C.c_int(0)
    _check(load_elp().elp_device_count(C.byref(n)))
    return n.value


def declared_symbols():
    """extern "C" entry points declared in include/elp_gpu.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(elp_\w+)\s*\(", text, re.M)))


def exported_symbols(path=None):
    path = path or os.path.join(LIB_DIR, "libelp.so")
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return sorted({ln.split()[-1] for ln in out.splitlines() if ln.strip()})


def bench_copy(nbytes, reps=20, device=0):
    """Average ms of one device-to-device copy of nbytes (plain streaming kernel)."""
    ms = C.c_float(0)
    _check(load_elp().elp_bench_copy(device, int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def _frame_summary(lib, fn, lead, host):
    """Call one of the elp_*_frame_summary entry points (leading arguments `lead`) for the chains of `host`."""
    l2i = host.label2id()
    n = host.n_pol
    out = {"start3": np.zeros((n, 3), np.int32), "sum3": np.zeros((n, 3), np.int64),
           "mid3": np.zeros((n, 3), np.int32), "rg2": np.zeros(n), "nsl": np.zeros(n, np.int32),
           "contacts": np.zeros(host.N, np.uint64)}
    _check(fn(*lead, l2i.ctypes.data, n, host.N,
              *(out[k].ctypes.data for k in ("start3", "sum3", "mid3", "rg2", "nsl", "contacts"))), lib)
    return out


class Engine:
    """One device lattice + stream (ring: uint8 per site, linear: uint16)."""

    def __init__(self, kind, L, tile=(32, 32, 32), sublattice=SUBLAT_CUBIC, n_iter=51, device=0):
        self.lib = load_elp()
        self.kind = kind
        self.L = tuple(int(x) for x in (L if hasattr(L, "__len__") else (L, L, L)))
        self.tile = tuple(int(x) for x in tile)
        self.dtype = np.uint8 if kind == "ring" else np.uint16
        h = C.c_void_p()
        fn = self.lib.elp_ring_create if kind == "ring" else self.lib.elp_lin_create
        _check(fn(C.byref(h), device, *self.L, *self.tile, int(sublattice), int(n_iter)), self.lib)
        self.h = h

    def close(self):
        if self.h:
            self.lib.elp_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def shape_tuv(self):
        return self.L

    def upload(self, lat):
        a = np.ascontiguousarray(lat, dtype=self.dtype)
        _check(self.lib.elp_upload(self.h, a.ctypes.data, a.nbytes), self.lib)

    def download(self):
        a = np.empty(self.L[0] * self.L[1] * self.L[2], dtype=self.dtype)
        _check(self.lib.elp_download(self.h, a.ctypes.data, a.nbytes), self.lib)
        return a

    def run(self, origins, tau0, seed, stream=0, sync=True):
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
        _check(self.lib.elp_run(self.h, len(o), o.ctypes.data, int(tau0), int(seed) & 0xFFFFFFFF,
                                int(stream) & 0xFFFFFFFF), self.lib)
        if sync:
            self.sync()

    def sync(self):
        _check(self.lib.elp_sync(self.h), self.lib)

    def last_run_ms(self):
        ms = C.c_float(0)
        _check(self.lib.elp_last_run_ms(self.h, C.byref(ms)), self.lib)
        return ms.value

    def launch_info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(self.lib.elp_launch_info(self.h, C.byref(a), C.byref(b), C.byref(c)), self.lib)
        return {"threads": a.value, "blocks": b.value, "lds_bytes": c.value}

    def set_threads(self, threads):
        _check(self.lib.elp_set_threads(self.h, int(threads)), self.lib)

    def set_rules(self, rules):
        """ring move rates: RULES_REFERENCE (default) or RULES_BALANCED"""
        _check(self.lib.elp_set_rules(self.h, int(rules)), self.lib)

    def count_moves(self, enable=True):
        """Launch the counting build of the kernel in the following runs (same moves); enable clears the counts."""
        _check(self.lib.elp_count_moves(self.h, int(bool(enable))), self.lib)

    def move_counts(self):
        """{"attempts", "empty", "face", "accepted"} since count_moves(True)."""
        c = np.zeros(4, np.uint64)
        _check(self.lib.elp_move_counts(self.h, c.ctypes.data), self.lib)
        return dict(zip(("attempts", "empty", "face", "accepted"), (int(x) for x in c)))

    def ring_chains(self, host):
        """Rings of the device lattice, recovered on the device: (bonds[n_pol][N], lengths, starts[n_pol][3])
        in polymer-id order, with the host's label table."""
        l2i = host.label2id()
        bonds = np.zeros((host.n_pol, host.N), np.uint8)
        lens = np.zeros(host.n_pol, np.int32)
        starts = np.zeros((host.n_pol, 3), np.int32)
        _check(self.lib.elp_ring_chains(self.h, l2i.ctypes.data, host.n_pol, host.N, bonds.ctypes.data,
                                        lens.ctypes.data, starts.ctypes.data), self.lib)
        return bonds, lens, starts

    def ring_stats(self, host):
        """Per polymer id, from the device-side walk: (Rg^2 in bond units, stored-length bond count)."""
        l2i = host.label2id()
        rg2 = np.zeros(host.n_pol)
        nsl = np.zeros(host.n_pol, np.int32)
        _check(self.lib.elp_ring_chain_stats(self.h, l2i.ctypes.data, host.n_pol, host.N, rg2.ctypes.data,
                                             nsl.ctypes.data), self.lib)
        return rg2, nsl

    def ring_frame_summary(self, host, noalloc=False):
        """One frame of statistics reduced on the device: per polymer id the start site, the sum of its
        monomer positions and its monomer N/2 relative to the start, Rg^2 and the stored-length bond count;
        contacts[g] = monomer pairs g apart that touch, summed over rings.  Feed to stats.Series(summary=True).
        noalloc: use elp_ring_frame_summary_noalloc (device buffers from the library's static scratch)."""
        fn = self.lib.elp_ring_frame_summary_noalloc if noalloc else self.lib.elp_ring_frame_summary
        return _frame_summary(self.lib, fn, (self.h,), host)

    def lin_frame_summary(self, host):
        """ring_frame_summary for a linear-chain engine (middle monomer (N-1)/2)."""
        return _frame_summary(self.lib, self.lib.elp_lin_frame_summary, (self.h,), host)

    def ring_tuv(self, host):
        """The tuv file text of the device lattice (bytes), from the device-side recovery."""
        return self._tuv(self.lib.elp_ring_tuv, host)

    def lin_chains(self, host):
        """Linear chains of the device lattice, recovered on the device: (bonds[n_pol][N], lengths,
        starts[n_pol][3]) in polymer-id order; the last bond of a chain is 0xf."""
        l2i = host.label2id()
        bonds = np.zeros((host.n_pol, host.N), np.uint8)
        lens = np.zeros(host.n_pol, np.int32)
        starts = np.zeros((host.n_pol, 3), np.int32)
        _check(self.lib.elp_lin_chains(self.h, l2i.ctypes.data, host.n_pol, host.N, bonds.ctypes.data,
                                       lens.ctypes.data, starts.ctypes.data), self.lib)
        return bonds, lens, starts

    def lin_tuv(self, host):
        """The tuv file text of a linear-chain device lattice (bytes), from the device-side recovery."""
        return self._tuv(self.lib.elp_lin_tuv, host)

    def _tuv(self, fn, host):
        l2i = host.label2id()
        n = C.c_uint64()
        _check(fn(self.h, l2i.ctypes.data, host.n_pol, host.N, host.max_pol_length, None, 0, C.byref(n)), self.lib)
        buf = C.create_string_buffer(n.value)
        _check(fn(self.h, l2i.ctypes.data, host.n_pol, host.N, host.max_pol_length, buf, n.value, C.byref(n)),
               self.lib)
        return buf.raw[:n.value]


class Batch:
    """Independent replicas of any boxes in one device buffer; one launch per time
    unit covers every tile of every replica (elp_batch_* in include/elp_gpu.h)."""

    def __init__(self, kind, boxes, tile=(16, 16, 16), sublattice=SUBLAT_DENSE, n_iter=13, device=0):
        self.lib = load_elp()
        self.kind = kind
        self.boxes = [tuple(int(x) for x in (b if hasattr(b, "__len__") else (b, b, b))) for b in boxes]
        self.n_rep = len(self.boxes)
        self.tile = tuple(int(x) for x in tile)
        self.dtype = np.uint8 if kind == "ring" else np.uint16
        bx = np.ascontiguousarray(self.boxes, dtype=np.int32).reshape(-1)
        h = C.c_void_p()
        fn = self.lib.elp_ring_batch_create if kind == "ring" else self.lib.elp_lin_batch_create
        _check(fn(C.byref(h), device, self.n_rep, bx.ctypes.data, *self.tile, int(sublattice), int(n_iter)), self.lib)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.elp_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def upload(self, rep, lat):
        a = np.ascontiguousarray(lat, dtype=self.dtype)
        _check(self.lib.elp_batch_upload(self.h, int(rep), a.ctypes.data, a.nbytes), self.lib)

    def download(self, rep):
        L = self.boxes[rep]
        a = np.empty(L[0] * L[1] * L[2], dtype=self.dtype)
        _check(self.lib.elp_batch_download(self.h, int(rep), a.ctypes.data, a.nbytes), self.lib)
        return a

    def run(self, origins, tau0, seeds, streams=None, units_per_graph=0, sync=True):
        """origins: (n_time, n_rep, 3); tau0, seeds, streams: one per replica."""
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, self.n_rep, 3)
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(tau0, dtype=np.uint64), (self.n_rep,)))
        streams = np.zeros(self.n_rep, dtype=np.uint32) if streams is None else streams
        k = np.empty((self.n_rep, 2), dtype=np.uint32)
        k[:, 0] = np.asarray(seeds, dtype=np.uint64) & 0xFFFFFFFF
        k[:, 1] = np.asarray(streams, dtype=np.uint64) & 0xFFFFFFFF
        _check(self.lib.elp_batch_run(self.h, len(o), o.ctypes.data, t.ctypes.data, k.ctypes.data,
                                      int(units_per_graph)), self.lib)
        if sync:
            self.sync()

    def sync(self):
        _check(self.lib.elp_batch_sync(self.h), self.lib)

    def frame_summary(self, rep, host):
        """Frame summary (as Engine.ring_frame_summary) of replica `rep`, read in place on the device;
        `host` holds that replica's label table and sizes."""
        return _frame_summary(self.lib, self.lib.elp_batch_frame_summary, (self.h, int(rep)), host)

    def last_run_ms(self):
        ms = C.c_float(0)
        _check(self.lib.elp_batch_last_run_ms(self.h, C.byref(ms)), self.lib)
        return ms.value

    def launch_info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_uint64()
        _check(self.lib.elp_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c)), self.lib)
        return {"threads": a.value, "blocks": b.value, "sites": c.value}

    def set_threads(self, threads):
        _check(self.lib.elp_batch_set_threads(self.h, int(threads)), self.lib)

    def set_rules(self, rules):
        _check(self.lib.elp_batch_set_rules(self.h, int(rules)), self.lib)

    def count_moves(self, enable=True):
        _check(self.lib.elp_batch_count_moves(self.h, int(bool(enable))), self.lib)

    def move_counts(self):
        """{"attempts", "empty", "face", "accepted"} over all replicas since count_moves(True)."""
        c = np.zeros(4, np.uint64)
        _check(self.lib.elp_batch_move_counts(self.h, c.ctypes.data), self.lib)
        return dict(zip(("attempts", "empty", "face", "accepted"), (int(x) for x in c)))


NET_NPARAM = 14416
NET_NIN = 144
NET_NMOVE = 11


class Net:
    """Learned move proposal (elp_net_* in include/elp_gpu.h)."""

    def __init__(self, seed=0, device=0):
        self.lib = load_elp()
        h = C.c_void_p()
        _check(self.lib.elp_net_create(C.byref(h), device, int(seed) & 0xFFFFFFFF), self.lib)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.elp_net_destroy(self.h)
            self.h = None

    __del__ = close

    def params(self):
        p = np.empty(NET_NPARAM, np.float32)
        _check(self.lib.elp_net_params(self.h, p.ctypes.data, p.size), self.lib)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(self.lib.elp_net_set_params(self.h, p.ctypes.data, p.size), self.lib)

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.int8).reshape(-1, NET_NIN)
        prob = np.empty((len(x), NET_NMOVE), np.float32)
        _check(self.lib.elp_net_forward(self.h, x.ctypes.data, len(x), prob.ctypes.data), self.lib)
        return prob

    def gradient(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.int8).reshape(-1, NET_NIN)
        y = np.ascontiguousarray(y, dtype=np.int32)
        g = np.empty(NET_NPARAM, np.float32)
        loss = C.c_float()
        _check(self.lib.elp_net_gradient(self.h, x.ctypes.data, y.ctypes.data, len(x), g.ctypes.data,
                                         C.byref(loss)), self.lib)
        return g, loss.value

    def harvest(self, engine, count, seed=0):
        _check(self.lib.elp_net_harvest(self.h, engine.h, int(count), int(seed) & 0xFFFFFFFF), self.lib)

    def samples(self):
        n = C.c_int()
        _check(self.lib.elp_net_samples(self.h, C.byref(n), None, None), self.lib)
        x = np.empty((n.value, NET_NIN), np.int8)
        y = np.empty(n.value, np.int32)
        _check(self.lib.elp_net_samples(self.h, C.byref(n), x.ctypes.data, y.ctypes.data), self.lib)
        return x, y

    def train(self, epochs=1, batch=4096, lr=1e-3):
        loss = C.c_float()
        _check(self.lib.elp_net_train(self.h, int(epochs), int(batch), float(lr), C.byref(loss)), self.lib)
        return loss.value

    def run(self, engine, origins, tau0, seed, stream=0):
        """Learned-proposal time units on a ring Engine (16^3 tiles, cubic);
        returns (proposals, accepted)."""
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
        npr, nac = C.c_uint64(), C.c_uint64()
        _check(self.lib.elp_run_learned(engine.h, self.h, len(o), o.ctypes.data, int(tau0), int(seed) & 0xFFFFFFFF,
                                        int(stream) & 0xFFFFFFFF, C.byref(npr), C.byref(nac)), self.lib)
        return npr.value, nac.value


def patch_moves(patches, moves, device=0):
    """Move m at the centre of each 5^3 patch: (patches after, ok, reverse move)."""
    lib = load_elp()
    p = np.ascontiguousarray(patches, dtype=np.uint8).reshape(-1, 125)
    m = np.ascontiguousarray(moves, dtype=np.int32)
    out = np.empty_like(p)
    ok = np.empty(len(p), np.int32)
    rev = np.empty(len(p), np.int32)
    _check(lib.elp_net_patch_moves(device, p.ctypes.data, m.ctypes.data, len(p), out.ctypes.data, ok.ctypes.data,
                                   rev.ctypes.data), lib)
    return out, ok, rev


def patch_features(patches, device=0):
    lib = load_elp()
    p = np.ascontiguousarray(patches, dtype=np.uint8).reshape(-1, 125)
    x = np.empty((len(p), NET_NIN), np.int8)
    _check(lib.elp_net_patch_features(device, p.ctypes.data, len(p), x.ctypes.data), lib)
    return x


def tile_check(tiles, cases, device=0):
    """Learned engine's tile path against the patch functions: per case
    (tile, t, u, v, move) the failed-check bits and whether the move is
    possible (elp_net_tile_check)."""
    lib = load_elp()
    t = np.ascontiguousarray(tiles, dtype=np.uint8).reshape(-1, 4096)
    c = np.ascontiguousarray(cases, dtype=np.int32).reshape(-1, 5)
    r = np.empty((len(c), 2), np.int32)
    _check(lib.elp_net_tile_check(device, t.ctypes.data, len(t), c.ctypes.data, len(c), r.ctypes.data), lib)
    return r[:, 0], r[:, 1]


class _Host:
    """Handle API of a drop-in host library (fresh start, shift stream, tuv)."""

    prefix = None
    libname = None
    dtype = None

    def __init__(self, N=None, L=None, density=1.2, seed=0, path=None, box_unit=0, init="reference",
                 site_fraction=0.0):
        """box_unit: box edges are truncated to multiples of it (0 = 32, the
        reference's rule; 16 allows boxes such as 48 for 16^3-tile engines);
        with a path, the edges of the file's box must be multiples of it (0 =
        the largest of 32, 16, 8, 4 that divides them).
        init: "reference" (the reference's compact chains) or, for linear
        chains, "saw" (random self-avoiding walks covering site_fraction of
        the box; 0 takes the library default)."""
        self.lib = _load(self.libname)
        p = self.prefix
        lib = self.lib
        if not getattr(lib, "_host_typed", False):
            getattr(lib, p + "new").restype = C.c_void_p
            getattr(lib, p + "new").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint32]
            getattr(lib, p + "new_unit").restype = C.c_void_p
            getattr(lib, p + "new_unit").argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint32,
                                                     C.c_int]
            getattr(lib, p + "from_file").restype = C.c_void_p
            getattr(lib, p + "from_file").argtypes = [C.c_char_p, C.c_uint32]
            getattr(lib, p + "free").argtypes = [C.c_void_p]
            getattr(lib, p + "info").argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 6
            getattr(lib, p + "get_lattice").argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
            getattr(lib, p + "set_lattice").argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
            getattr(lib, p + "write_tuv").argtypes = [C.c_void_p, C.c_char_p]
            getattr(lib, p + "polymer").argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
            getattr(lib, p + "load_shift").argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
            getattr(lib, p + "origins").argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
            getattr(lib, p + "seeds").argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
            getattr(lib, p + "seed").restype = C.c_uint32
            getattr(lib, p + "seed").argtypes = [C.c_void_p]
            lib._host_typed = True
        if init not in ("reference", "saw"):
            raise ValueError(f"unknown init {init!r}")
        if path is not None:
            if box_unit:
                f = getattr(lib, p + "from_file_unit")
                f.restype = C.c_void_p
                f.argtypes = [C.c_char_p, C.c_uint32, C.c_int]
                h = f(os.fsencode(path), int(seed), int(box_unit))
            else:
                h = getattr(lib, p + "from_file")(os.fsencode(path), int(seed))
        elif init == "saw":
            if not hasattr(lib, p + "new_saw"):
                raise ElpError("the self-avoiding-walk start exists for linear chains only")
            f = getattr(lib, p + "new_saw")
            f.restype = C.c_void_p
            f.argtypes = [C.c_int] * 4 + [C.c_double, C.c_uint32, C.c_int, C.c_double]
            Ls = L if hasattr(L, "__len__") else (L, L, L)
            h = f(int(N), int(Ls[0]), int(Ls[1]), int(Ls[2]), float(density), int(seed), int(box_unit),
                  float(site_fraction))
        else:
            Ls = L if hasattr(L, "__len__") else (L, L, L)
            h = getattr(lib, p + "new_unit")(int(N), int(Ls[0]), int(Ls[1]), int(Ls[2]), float(density), int(seed),
                                             int(box_unit))
        if not h:
            raise ElpError(f"{p}new/from_file failed")
        self.h = C.c_void_p(h)
        vals = [C.c_int() for _ in range(6)]
        getattr(lib, p + "info")(self.h, *[C.byref(v) for v in vals])
        self.L = (vals[0].value, vals[1].value, vals[2].value)
        self.n_pol, self.N, self.max_pol_length = vals[3].value, vals[4].value, vals[5].value

    def close(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.prefix + "free")(self.h)
            self.h = None

    __del__ = close

    @property
    def nsite(self):
        return self.L[0] * self.L[1] * self.L[2]

    def lattice(self):
        a = np.empty(self.nsite, dtype=self.dtype)
        if getattr(self.lib, self.prefix + "get_lattice")(self.h, a.ctypes.data, a.nbytes):
            raise ElpError("get_lattice failed")
        return a

    def set_lattice(self, lat):
        a = np.ascontiguousarray(lat, dtype=self.dtype)
        rc = getattr(self.lib, self.prefix + "set_lattice")(self.h, a.ctypes.data, a.nbytes)
        if rc:
            raise ElpError("set_lattice: chain recovery failed")

    def write_tuv(self, path):
        if getattr(self.lib, self.prefix + "write_tuv")(self.h, os.fsencode(path)):
            raise ElpError("write_tuv failed")

    def polymer(self, pid):
        bonds = np.zeros(self.max_pol_length + 2, dtype=np.uint32)
        start = np.zeros(3, dtype=np.uint32)
        n = getattr(self.lib, self.prefix + "polymer")(self.h, int(pid), bonds.ctypes.data, len(bonds),
                                                       start.ctypes.data)
        if n < 0:
            raise ElpError(f"polymer {pid} not found")
        return tuple(int(x) for x in start), bonds[:n].copy()

    def load_shift(self, tile):
        getattr(self.lib, self.prefix + "load_shift")(self.h, *tile)

    def origins(self, tile, n):
        o = np.zeros((n, 3), dtype=np.int32)
        getattr(self.lib, self.prefix + "origins")(self.h, *tile, int(n), o.ctypes.data)
        return o

    def seed_table(self):
        """Per-lane seed table of the reference start-up (uint32[nwg*512*8])."""
        n = (self.L[0] // 32) * (self.L[1] // 32) * (self.L[2] // 32) * 512 * 8
        a = np.zeros(n, dtype=np.uint32)
        if getattr(self.lib, self.prefix + "seeds")(self.h, a.ctypes.data, n):
            raise ElpError("seed table size mismatch")
        return a

    @property
    def seed(self):
        return getattr(self.lib, self.prefix + "seed")(self.h)

    def label2id(self):
        """the 2^18-entry table from a chain's 18 id label bits to its polymer id (-1 unused)"""
        f = getattr(self.lib, self.prefix + "label2id")
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        out = np.empty(1 << 18, np.int32)
        if f(self.h, out.ctypes.data, out.size):
            raise ElpError("label table not available")
        return out


class RingHost(_Host):
    prefix = "elp_ring_host_"
    libname = "libgpupol3_ring.so"
    dtype = np.uint8


class LinHost(_Host):
    prefix = "elp_lin_host_"
    libname = "libgpupol_lin.so"
    dtype = np.uint16
