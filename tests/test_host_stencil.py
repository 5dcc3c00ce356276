"""Compile the kernels' inline geometry/stencil functions for the host and
check them bit-exactly against the oracle on 11 lattice shapes (CPU only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_stencil(tmp_path, P):
    P.lib()  # builds oracle/liboracle.so if needed
    exe = tmp_path / "stencil_check"
    src = os.path.join(ROOT, "tests", "host", "stencil_check.hip")
    oracle = os.path.join(ROOT, "oracle")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off",
                    src, f"-L{oracle}", "-loracle", f"-Wl,-rpath,{oracle}", "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
