"""Compile the replay engine's burst-size rules (gnx_replay_plan.h) for the
host and check them, including a simulation of the HMC's solve pattern
(CPU only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_replay_plan(tmp_path):
    exe = tmp_path / "replay_plan_check"
    src = os.path.join(ROOT, "tests", "host", "replay_plan_check.hip")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", src, "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
