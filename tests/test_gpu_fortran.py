"""The reference's own Fortran checks (fft_c2c_x / fft_r2c_x / their _z
copies, fft_multiple_grids, test2d, the dtt and halo examples) run through the Fortran decomp_2d /
decomp_2d_fft modules over the C ABI (2decomp-fft_amd/fortran)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "2decomp-fft_amd", "fortran", "fft_hip_tests")


def test_fortran_reference_checks(torch_gpu):
    assert os.path.exists(EXE), "build() did not produce the Fortran test program"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0
    assert "fft_hip_tests completed" in r.stdout
    assert "transposes exact" in r.stdout
    assert r.stdout.count("error / mesh point") == 12
    assert r.stdout.count("dtt checks passed: 512 type triples") == 2
    assert "halo divergence checks passed" in r.stdout
    assert "periodic halo checks passed" in r.stdout
