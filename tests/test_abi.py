"""The C-ABI libraries load and export every declared entry point (no GPU calls)."""
import ctypes as C
import os
import subprocess

import elp

LIB = elp.LIB_DIR
GPULIB_RING = ["GPULibInit", "GPULibLoadBuffers", "GPULibRun", "GPULibEqRun", "GPULoadLattices", "GPULatticeToCPU",
               "GPULibRelease"]
GPULIB_LIN = ["GPULibInit", "GPULibLoadBuffers", "GPULibRun", "GPULibRelease", "GPULatticeToCPU"]
HOST = ["new", "from_file", "free", "info", "get_lattice", "set_lattice", "write_tuv", "polymer", "load_shift",
        "origins", "seed", "seeds"]


def test_libelp_exports_every_declared_symbol():
    declared = elp.declared_symbols()
    assert "elp_run" in declared and "elp_ring_create" in declared and len(declared) >= 12
    exported = set(elp.exported_symbols())
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    lib = elp.load_elp()
    for s in declared:
        getattr(lib, s)


def test_host_libraries_export_gpulib_boundary():
    ring = set(elp.exported_symbols(os.path.join(LIB, "libgpupol3_ring.so")))
    lin = set(elp.exported_symbols(os.path.join(LIB, "libgpupol_lin.so")))
    assert all(s in ring for s in GPULIB_RING)
    assert all(s in lin for s in GPULIB_LIN)
    assert all("elp_ring_host_" + s in ring for s in HOST)
    assert all("elp_lin_host_" + s in lin for s in HOST)


def _c_functions(header):
    """Names of the functions a C header declares (prototypes ending in ');')."""
    import re
    text = re.sub(r"/\*.*?\*/", "", open(header).read(), flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"#[^\n]*", "", text)
    protos = re.findall(r"[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(([^;{}()]|\([^()]*\))*\)\s*;", text)
    return sorted({name for name, _ in protos} - {"sizeof"})


def test_host_libraries_export_every_declared_function():
    """Every function include/gpupol_ring.h and include/gpupol_lin.h declare is
    exported by libgpupol3_ring.so / libgpupol_lin.so."""
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    for header, lib in (("gpupol_ring.h", "libgpupol3_ring.so"), ("gpupol_lin.h", "libgpupol_lin.so")):
        declared = _c_functions(os.path.join(inc, header))
        assert "GPULibRun" in declared and "LoadPolymers" in declared and len(declared) > 30, declared
        exported = set(elp.exported_symbols(os.path.join(LIB, lib)))
        missing = [f for f in declared if f not in exported]
        assert not missing, (header, missing)


def test_error_paths_without_gpu_calls():
    """Argument validation happens before any device call."""
    lib = elp.load_elp()
    h = C.c_void_p()
    assert lib.elp_ring_create(C.byref(h), 0, 96, 96, 96, 24, 24, 24, 0, 51) == -1   # schedule not compiled in
    assert b"" != lib.elp_last_error()
    assert lib.elp_ring_create(C.byref(h), 0, 100, 96, 96, 16, 16, 16, 0, 51) == -1  # box not divisible
    assert b"divisible" in lib.elp_last_error()
    assert lib.elp_run(None, 1, None, 0, 0, 0) == -1


def test_batch_error_paths_without_gpu_calls():
    lib = elp.load_elp()
    h = C.c_void_p()
    boxes = (C.c_int32 * 6)(32, 32, 32, 40, 32, 32)
    assert lib.elp_ring_batch_create(C.byref(h), 0, 2, boxes, 16, 16, 16, 1, 13) == -1
    assert b"replica 1" in lib.elp_last_error()
    boxes = (C.c_int32 * 3)(32, 32, 32)
    assert lib.elp_lin_batch_create(C.byref(h), 0, 0, boxes, 16, 16, 16, 1, 16) == -1
    assert lib.elp_batch_run(None, 1, None, None, None, 0) == -1


def test_schedule_record_round_trip():
    s = elp.Schedule(16, 16, 16, elp.SUBLAT_DENSE, 13, 256, 16, 5.25e11, elp.RULES_BALANCED, 1.5e11)
    rec = s.record()
    assert rec == ("tile=16x16x16 sublattice=dense n_iter=13 threads=256 units_per_graph=16 attempts_per_s=5.250000e+11"
                   " rules=balanced accepted_per_s=1.500000e+11")
    t = elp.Schedule.parse(rec)
    assert (t.tile, t.sublattice, t.n_iter, t.threads, t.units_per_graph) == ((16, 16, 16), 1, 13, 256, 16)
    assert (t.attempts_per_s, t.rules, t.accepted_per_s) == (5.25e11, elp.RULES_BALANCED, 1.5e11)
    # records written before the rules and accepted-move fields: reference rates
    u = elp.Schedule.parse("tile=32x32x32 sublattice=cubic n_iter=51 threads=256 units_per_graph=0")
    assert (u.tile, u.sublattice, u.n_iter, u.rules, u.accepted_per_s) == ((32, 32, 32), 0, 51, 0, 0.0)
    v = elp.Schedule.parse("tile=16x16x16 sublattice=dense n_iter=13 threads=256 units_per_graph=16 "
                           "attempts_per_s=5.25e11")
    assert (v.attempts_per_s, v.rules) == (5.25e11, elp.RULES_REFERENCE)
    for bad in ("tile=16x16 sublattice=dense", "tile=16x16x16 sublattice=sparse n_iter=1 threads=64 units_per_graph=0",
                "tile=16x16x16 sublattice=dense n_iter=13 threads=256 units_per_graph=16 attempts_per_s=1 rules=fast"):
        try:
            elp.Schedule.parse(bad)
        except elp.ElpError:
            continue
        raise AssertionError(bad)


def test_drop_in_executables_print_usage():
    for exe in ("gpupol3_cuda_ring", "gpupol_cuda_lin"):
        r = subprocess.run([os.path.join(elp.PKG_DIR, "bin", exe)], capture_output=True, text=True)
        assert r.returncode == 2 and "usage" in r.stderr
