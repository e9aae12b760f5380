"""Host code of the solver set-up, compiled for the CPU with g++: the SELL
layouts of the renumbered multigrid levels (32-bit columns, 16-bit
offsets + Dirichlet lists, and the same with slots packed by four) hold exactly the rows of the permuted CSR pattern,
and every index a SELL kernel forms from them stays inside its array
(slot < number of slots, column < rows) on Q1/Q2 quadrilaterals, P1
triangles and Q2 hexahedra."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "mymoonmd_amd", "csrc", "host")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_sell_layouts_hold_the_rows_and_stay_in_bounds(tmp_path):
    exe = str(tmp_path / "check_sell_layouts")
    src = [os.path.join(ROOT, "tests", "host", "check_sell_layouts.cpp")] + [
        os.path.join(HOST, f) for f in ("space.cpp", "mesh.cpp", "fe.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-fopenmp", "-I", HOST, *src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "all checks passed" in out.stdout
