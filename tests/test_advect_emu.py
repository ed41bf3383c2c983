"""The LDS-staged advection kernels, run on the CPU through a serial
emulation of their launches (tests/advect_emu), against the same kernels
reading global memory: bit for bit on random fields over grid sizes whose
tiles include the interior margin exactly, at CFL numbers below and above the
reach limit.  The emulation also aborts on any LDS read outside its window
and on any clamp-free sampling axis (interior tiles) that differs from the
clamped one, so it shows on the CPU that the unchecked reads stay in bounds
and that dropping the clamp changes no index or weight.  The one-pass
MacCormack dye kernel (its staged windows and its forward-pass ring, each
read checked against its own extent) runs against the two global-gather
passes too, at CFL numbers on both sides of its reach test."""
import os
import re
import subprocess
import sys

import pytest

HERE = os.path.join(os.path.dirname(__file__), "advect_emu")
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANG), reason="clang++ (ROCm LLVM) not installed")
def test_lds_advection_matches_global_gathers(tmp_path):
    sys.path.insert(0, HERE)
    import make_emu
    src = tmp_path / "src"
    make_emu.make(str(src))
    exe = tmp_path / "emu"
    subprocess.run([CLANG, "-std=c++17", "-O2", "-I", HERE, "-I", str(src),
                    "-Wno-unused-command-line-argument", os.path.join(HERE, "main.cpp"),
                    str(src / "fh_advect_emu.cpp"), "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL BIT-IDENTICAL" in r.stdout
    m = re.search(r"staged blocks (\d+), global blocks (\d+), checked LDS reads (\d+), "
                  r"clamp-free axes (\d+)", r.stdout)
    assert m and all(int(g) > 0 for g in m.groups()), r.stdout
    # some one-pass dye launches went over the reach (the two-pass kernels
    # redid them), the others did not
    m = re.search(r"one-pass dye launches marked over the reach (\d+)", r.stdout)
    assert m and 0 < int(m.group(1)) < 35, r.stdout
