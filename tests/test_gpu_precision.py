"""FP32-bulk precision mode against the FP64 GPU path.

The FP32-bulk step keeps the pressure tridiagonal, the wall stress and all
reductions in FP64 and rounds everything else to single precision.  These
tests report the per-field error of that policy after 1, 10 and 100 steps
from the same seeded initial condition and bound it.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "dudz", "txz", "Nu_t", "RHSx", "RHSz")


def _rel(a, b):
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d


def _interior(name, x, nx, nz):
    # physical columns 1..nx, levels 1..nz-1 (u/p-nodes) or 2..nz-1 (w)
    if name == "p":
        return x[1:nz, :, :nx]
    if name in ("w", "RHSz"):
        return x[1:nz - 1, :, :nx]
    return x[:nz - 1, :, :nx]


@pytest.fixture(scope="module")
def runs(torch_mod):
    from les_amd.api import LesSim
    nx, ny, nz = 64, 64, 33
    out = {}
    sims = {prec: LesSim(nx, ny, nz, precision=prec, dt=2e-4, graph=True) for prec in ("fp64", "fp32")}
    for s in sims.values():
        s.init_ic(seed=11)
    done = 0
    for nsteps in (1, 10, 100):
        print("fp32-bulk vs fp64 at step", nsteps, flush=True)
        for s in sims.values():
            s.step(nsteps - done)
        done = nsteps
        out[nsteps] = {name: (sims["fp32"].get(name), sims["fp64"].get(name)) for name in FIELDS}
    return nx, ny, nz, out


# Bounds sit about 10x above the measured 64x64x32 errors.  The relative
# error of v and w is large against their own (small) norm because the FP32
# rounding scales with the streamwise velocity u ~ 20 u*.
@pytest.mark.parametrize("nsteps,bound", [(1, 1e-4), (10, 5e-4), (100, 2e-3)])
def test_fp32_bulk_field_error(runs, nsteps, bound):
    nx, ny, nz, out = runs
    table = []
    for name in FIELDS:
        a, b = out[nsteps][name]
        a, b = _interior(name, a, nx, nz), _interior(name, b, nx, nz)
        assert np.all(np.isfinite(a)), name
        table.append((name, _rel(a, b)))
    print(f"step {nsteps}: " + "  ".join(f"{n}={e:.2e}" for n, e in table), flush=True)
    for name, e in table:
        assert e < bound, (name, e, bound)


@pytest.mark.parametrize("nsteps", [1, 10, 100])
def test_fp32_bulk_mean_profiles(runs, nsteps):
    """Plane-mean profiles of u, v and the resolved + SGS stress: the
    quantities an ABL study reports."""
    nx, ny, nz, out = runs
    for name in ("u", "v", "txz"):
        a, b = out[nsteps][name]
        pa = a[:nz - 1, :, :nx].mean(axis=(1, 2))
        pb = b[:nz - 1, :, :nx].mean(axis=(1, 2))
        scale = np.abs(pb).max()
        err = np.abs(pa - pb).max() / scale
        print(f"step {nsteps} mean {name}: max rel {err:.2e}", flush=True)
        assert err < 1e-4, (name, err)
