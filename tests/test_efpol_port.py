"""The efpol restatement (oracle/efpol_port.c) against golden tuv files
written by the reference efpol binaries (tests/golden/make_efpol_golden.sh)."""
import filecmp
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "efpol")
CASES = sorted(os.listdir(GOLD)) if os.path.isdir(GOLD) else []


@pytest.mark.parametrize("case", CASES)
def test_efpol_port_byte_exact(case, tmp_path):
    argv = open(os.path.join(GOLD, case, "ARGV")).read().split()
    exe, seed, _dir, dens, time_, intv, nmono, dbl, L, model = argv
    port = os.path.join(ROOT, "oracle", "bin", "efpol_port_" + ("ring" if exe.endswith("ring") else "lin"))
    out = tmp_path / "run"
    subprocess.run([port, seed, str(out), dens, time_, intv, nmono, dbl, L, model], check=True,
                   stdout=subprocess.DEVNULL)
    golden = sorted(f for f in os.listdir(os.path.join(GOLD, case)) if f.startswith("t="))
    assert len(golden) >= 2
    for f in golden:
        assert filecmp.cmp(os.path.join(GOLD, case, f), out / f"N_{nmono}" / f, shallow=False), f
    assert filecmp.cmp(os.path.join(GOLD, case, "simulation_settings.txt"),
                       out / f"N_{nmono}" / "simulation_settings.txt", shallow=False)


def test_efpol_port_bench_mode():
    port = os.path.join(ROOT, "oracle", "bin", "efpol_port_ring")
    out = subprocess.run([port, "bench", "16", "32", "1.2", "1", "20", "5"], check=True, capture_output=True,
                         text=True).stdout
    kv = dict(x.split("=") for x in out.split())
    assert int(kv["attempts"]) == 20 * 32 * 154 and float(kv["seconds"]) > 0


def test_bench_cpu_baseline_record():
    """bench.py's CPU leg on a tiny melt: a rate, the cores and CPU model used, and
    the monomer-normalised rate (efpol's time unit is one attempt per monomer)."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    rec = bench.cpu_baseline(32, 100, 1.2, cores=1, warm=20, timed=10)
    assert rec is not None and rec["unit"] == "attempts/s" and rec["cores"] == 1 and rec["value"] > 0
    assert rec["kind"] in ("reference", "port")
    assert rec["monomer_time_units_per_s"] == rec["value"]
    assert "cpu_model" in rec
