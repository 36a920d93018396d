"""bench.py's contract pieces that run without a GPU: the one-GPU refusal and
the bytes-per-update arithmetic of the roofline."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bytes_per_lattice_update():
    # 2 * Q * sizeof(storage) + 1 flag byte, D3Q19
    assert bench.bytes_per_update("f64") == 305
    assert bench.bytes_per_update("f32s") == 153
    assert bench.bytes_per_update("f32") == 153
    assert bench.bytes_per_update("f16s") == 77
    assert bench.algorithmic_bytes("f32s") == 256 ** 3 * 153


def test_fluid_cells_of_the_headline_mask():
    mask = bench.solid_mask(0.1)
    n = bench.fluid_cells(mask)
    # the wall layers are not fluid; about 10% of the rest are solid
    interior = 256 * 256 * 254
    assert 0.88 * interior < n < 0.92 * interior
    assert bench.fluid_cells(None) == interior
