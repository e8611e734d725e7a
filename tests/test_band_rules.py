"""The byte-parallel rules of the band walks on the CPU (no GPU): Life and
Generations four cells a dword (life_rule, gen_rule and the interval constants
of csrc/kernels_life.hpp), the Checkerboard pattern and the open-cell select of
csrc/gate.hpp, over every Int8 cell value and every live count up to 120,
against the rules stated one cell at a time (src/rules.jl:385-389, 446-456;
src/schemes.jl:84-89).  The kernels on the GPU are checked against the oracle
in tests/test_gpu_parity.py and tests/test_gpu_gated.py."""
import os
import shutil
import subprocess

import pytest


@pytest.mark.skipif(shutil.which("/opt/rocm/bin/hipcc") is None, reason="needs hipcc")
def test_swar_rules_match_the_cell_rules(tmp_path):
    here = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(os.path.dirname(here), "cellularautomatacore.jl_amd", "csrc")
    exe = str(tmp_path / "swar_rules_selftest")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-std=c++17", "-O1", "-I", csrc,
                    os.path.join(here, "native", "swar_rules_selftest.hip"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
