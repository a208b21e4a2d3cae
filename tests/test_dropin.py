"""The C++ drop-in API: the test program links (CPU) and passes the reference's known answers (GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "gpu-voxels_amd", "lib", "test_dropin")


def test_dropin_program_links():
    assert os.path.exists(BIN), "run gpu-voxels_amd/build.py"
    out = subprocess.run(["ldd", BIN], capture_output=True, text=True).stdout
    assert "libgpu_voxels.so" in out and "libgvx.so" in out and "not found" not in out


@pytest.mark.gpu
def test_dropin_reference_known_answers():
    env = dict(os.environ, GPU_VOXELS_MODEL_PATH=os.path.join(ROOT, "tests", "golden", "models"))
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300, env=env)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-3000:]
    assert "0 failed" in r.stdout
