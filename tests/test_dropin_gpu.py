"""End-to-end runs of the drop-in executables (argv of do_run.sh:355) on the GPU."""
import filecmp
import os
import subprocess

import numpy as np
import pytest

import elp
from ringcheck import replay_lin, replay_ring

pytestmark = pytest.mark.gpu
BIN = os.path.join(elp.PKG_DIR, "bin")


def _run(exe, args, env=None):
    print(f"[drop-in] {exe} {' '.join(map(str, args))}", flush=True)
    r = subprocess.run([os.path.join(BIN, exe)] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def _settings(d):
    return dict(ln.split(" = ", 1) for ln in open(os.path.join(d, "simulation_settings.txt")).read().splitlines())


def test_ring_dropin_matches_replay_and_resumes(tmp_path):
    d = tmp_path / "ring"
    out = _run("gpupol3_cuda_ring", [100, 200, 7, d, 1.2, 0, 100, 32, 32, 32])
    assert "Simulation complete" in out
    files = sorted(os.listdir(d))
    assert {"t=0_dev=0.res", "t=100_dev=0.res", "t=200_dev=0.res", "simulation_settings.txt"} <= set(files)
    s = _settings(d)
    assert s["Start_seed"] == "7" and s["Npol"] == "393" and s["Executable"] == "gpupol3"

    # the same start-up through the handle API writes the same t=0 file
    h = elp.RingHost(N=100, L=32, density=1.2, seed=7)
    h.write_tuv(tmp_path / "t0.res")
    assert filecmp.cmp(d / "t=0_dev=0.res", tmp_path / "t0.res", shallow=False)

    # t=100 == replay of 100 time units with the drop-in's host shift stream
    lat0 = h.lattice()
    h.load_shift((32, 32, 32))
    org = h.origins((32, 32, 32), 100)
    want = replay_ring(lat0, h.L, (32, 32, 32), elp.SUBLAT_CUBIC, 51, org, 0, 7)
    h.set_lattice(want)
    h.write_tuv(tmp_path / "t100.res")
    assert filecmp.cmp(d / "t=100_dev=0.res", tmp_path / "t100.res", shallow=False)

    # deterministic, and the host writer (ELP_GPU_TUV=0) gives the same files
    # as the device-side recovery the executable uses by default
    d2 = tmp_path / "ring2"
    _run("gpupol3_cuda_ring", [100, 200, 7, d2, 1.2, 0, 100, 32, 32, 32], env=dict(os.environ, ELP_GPU_TUV="0"))
    assert filecmp.cmp(d / "t=100_dev=0.res", d2 / "t=100_dev=0.res", shallow=False)
    assert filecmp.cmp(d / "t=200_dev=0.res", d2 / "t=200_dev=0.res", shallow=False)

    # resume from t=200: seed += 200/100, two more intervals
    _run("gpupol3_cuda_ring", [100, 200, 7, d, 1.2, 0, 100, 32, 32, 32])
    assert {"t=300_dev=0.res", "t=400_dev=0.res"} <= set(os.listdir(d))
    assert _settings(d)["Start_seed"] == "9"
    r = elp.RingHost(path=str(d / "t=400_dev=0.res"))
    assert r.n_pol == 393 and r.N == 100


def test_ring_dropin_fast_eq_and_dense_schedule(tmp_path):
    d = tmp_path / "feq"
    env = dict(os.environ, ELP_TILE="16", ELP_SUBLATTICE="dense", ELP_NITER="13")
    _run("gpupol3_cuda_ring", [200, 100, 3, d, 1.2, 25, 100, 64, 64, 64], env=env)
    s = _settings(d)
    assert s["Fast_equilibration"] == "1" and s["FEQ_timestep"] == "25"
    r = elp.RingHost(path=str(d / "t=100_dev=0.res"))
    assert r.L == (64, 64, 64) and r.N == 200


def test_linear_dropin(tmp_path):
    d = tmp_path / "lin"
    out = _run("gpupol_cuda_lin", [200, 200, 5, d, 1.2, 0, 100, 32, 32, 32])
    assert "Simulation complete" in out
    assert _settings(d)["Executable"] == "gpupol"
    text = open(d / "t=200_dev=0.res").read().split("\n")
    assert text[3] == "np= 197"
    chains = [ln.split("\t")[1] for ln in text[5:] if "\t" in ln]
    assert len(chains) == 197 and all(c.endswith("f") and len(c) == 200 for c in chains)
    r = elp.LinHost(path=str(d / "t=200_dev=0.res"))
    assert r.n_pol == 197
    # the files come from the device-side recovery; the host path (ELP_GPU_TUV=0) writes the same bytes
    d2 = tmp_path / "lin2"
    _run("gpupol_cuda_lin", [200, 200, 5, d2, 1.2, 0, 100, 32, 32, 32], env=dict(os.environ, ELP_GPU_TUV="0"))
    for t in (0, 100, 200):
        assert filecmp.cmp(d / f"t={t}_dev=0.res", d2 / f"t={t}_dev=0.res", shallow=False), t


def test_linear_dropin_saw_start(tmp_path):
    """ELP_LIN_INIT=saw: the executable's fresh start is the handle API's
    self-avoiding-walk start, and t=100 is its replay."""
    d = tmp_path / "lsaw"
    _run("gpupol_cuda_lin", [100, 100, 6, d, 1.2, 0, 100, 32, 32, 32], env=dict(os.environ, ELP_LIN_INIT="saw"))
    h = elp.LinHost(N=100, L=32, density=1.2, seed=6, init="saw")
    h.write_tuv(tmp_path / "t0.res")
    assert filecmp.cmp(d / "t=0_dev=0.res", tmp_path / "t0.res", shallow=False)
    lat0 = h.lattice()
    org = h.origins((32, 32, 32), 100)
    h.set_lattice(replay_lin(lat0, h.L, (32, 32, 32), elp.SUBLAT_CUBIC, 64, org, 0, 6))
    h.write_tuv(tmp_path / "t100.res")
    assert filecmp.cmp(d / "t=100_dev=0.res", tmp_path / "t100.res", shallow=False)
