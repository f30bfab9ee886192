"""The Fortran (iso_c_binding) side of the C ABI: a generated interface for
every function the header declares, and a Fortran executable
(les_amd/fortran/les_fortran_check) that runs the host-array wrappers against
closed forms and whole steps with every SGS model."""
import os
import re
import subprocess
import sys

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


def test_capi_module_is_current_and_complete():
    """les_amd_capi.f90 is what gen_capi.py writes from the header, with one
    bind(C) interface per declared function."""
    r = subprocess.run([sys.executable, os.path.join(ROOT, "les_amd", "fortran", "gen_capi.py"), "--check"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    from les_amd._lib import declared_symbols
    txt = open(os.path.join(ROOT, "les_amd", "fortran", "les_amd_capi.f90")).read()
    bound = set(re.findall(r"bind\(C, *&?\s*name='(les_\w+)'\)", txt))
    assert set(declared_symbols()) == bound, set(declared_symbols()) ^ bound
