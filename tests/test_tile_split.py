"""The panel kernel's matvec tile split (weighted toward the non-owner ranks) checked on the host: the
producer (panel_map: a rank's range and row slot) and the consumer (tall_slots: how many row slots a
strip has) must agree exactly, or a row pass would wait for a slot nobody writes."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="no hipcc")
def test_tile_split_ranges_and_slots(tmp_path):
    exe = str(tmp_path / "tile_split_check")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                    "-I", os.path.join(ROOT, "skewlinearalgebra.jl_amd", "csrc"), "-I", os.path.join(ROOT, "include"),
                    os.path.join(HERE, "host", "tile_split_check.cpp"), "-o", exe], check=True,
                   stderr=subprocess.DEVNULL)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    assert " bad 0" in r.stdout
