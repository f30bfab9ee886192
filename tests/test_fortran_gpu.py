"""The Fortran (iso_c_binding) side of the C ABI: the host-array wrapper of
filt_da on a single Fourier mode and 5 whole steps through les_step, run as
a Fortran executable (les_amd/fortran/les_fortran_check)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "les_amd", "fortran", "les_fortran_check")


@pytest.mark.gpu
def test_fortran_binding_runs_on_gpu():
    if not os.path.exists(EXE):
        pytest.fail("les_fortran_check not built: run __graft_entry__.build()")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "LES_FORTRAN_OK" in r.stdout


def test_fortran_binding_compiles():
    """Interfaces compile with flang (no GPU needed)."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "les_amd", "fortran")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(EXE)
