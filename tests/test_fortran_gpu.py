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


@pytest.fixture(scope="module")
def fortran_run(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail("les_fortran_check not built: run __graft_entry__.build()")
    out = tmp_path_factory.mktemp("fortran_states")
    r = subprocess.run([EXE, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    return r, out


@pytest.mark.gpu
def test_fortran_binding_runs_on_gpu(torch_mod, fortran_run):
    r, _ = fortran_run
    assert r.returncode == 0, r.stdout + r.stderr
    assert "LES_FORTRAN_OK" in r.stdout


def _record_fields(path, n3, nfields):
    import numpy as np
    from les_amd.api import record_read
    raw = record_read(str(path))
    assert len(raw) >= 8 * n3 * nfields
    return [np.frombuffer(raw[8 * n3 * i:8 * n3 * (i + 1)], dtype="<f8") for i in range(nfields)]


@pytest.mark.gpu
@pytest.mark.parametrize("model", [1, 2, 3, 4, 5, 6, "convective"])
def test_fortran_whole_steps_vs_oracle(torch_mod, fortran_run, model):
    """The states the Fortran program reaches through les_step (32x32x9, seed
    42, 6 steps with SGS model 1..6, model 1 with the Coriolis force and the
    sponge of damping_method 1; 5 convective steps with theta, q and a uniform
    surface heat flux) against the oracle run with the same set-up: u, v, w
    and RHSx..RHSz (and theta) to 1e-9 relative, the bound of the 10-step
    FP64 whole-step tests."""
    import numpy as np
    from oracle.oracle import Oracle, OracleSim, scalar_params
    from conftest import relerr
    r, out = fortran_run
    assert r.returncode == 0, r.stdout + r.stderr
    print("fortran whole steps", model, flush=True)
    nx, ny, nz = 32, 32, 9
    o = Oracle(nx, ny, nz, dt=2e-4, z0=1e-4, force=1.0, inilag=1, initu=1, cs_count=5)
    if model == "convective":
        o.set_scalars(scalar_params(1, 7, 0, 0, 0.55, 1000.0, 1.0, 400.0, 100.0, 9.81, 0.01, -0.004),
                      np.full((ny, nx), 0.05), None)
        so = OracleSim(o, seed=42, scalar_ic=dict(T_init=283.0, z_inv=0.8, inv_strength=0.01, amp_K=0.1,
                                                  q_init=1.0))
        so.step(5)
        names = ("u", "v", "w", "theta", "qmix", "RHSx", "RHSy", "RHSz")
        path = out / "convective.out"
    else:
        o.set_sgs_model(model)
        if model == 1:
            o.set_forcing(coriolis=1, ug0=10.0, vg0=0.0, u_star=0.45, z_i=1000.0, ubc=1, damping_method=1)
        so = OracleSim(o, seed=42)
        so.step(6)
        names = ("u", "v", "w", "RHSx", "RHSy", "RHSz")
        path = out / f"model_{model}.out"
    ld = 2 * (nx // 2 + 1)
    segs = _record_fields(path, ld * ny * nz, len(names))
    for name, seg in zip(names, segs):
        a = seg.reshape(nz, ny, ld)[:nz - 1, :, :nx]
        b = so[name][:nz - 1, :, :nx]
        e = relerr(a, b)
        print(f"  {name}: {e:.2e}", flush=True)
        assert e < 1e-9, (name, e)


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
