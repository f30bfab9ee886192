"""Statistics as FP64 reductions on the GPU (les_spectrum, les_avg_stats,
les_surface_stats and the state-level les_state_stats*) against the oracle's
restatement of io.f90 spectrum / avg_stats and obukhov's plane averages:
routine level on the same inputs (1e-12), and the state's statistics after
several steps against the same statistics of the oracle's state."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

GRIDS = [(32, 32, 9), (48, 32, 7)]


def G(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def H(t):
    return t.double().cpu().numpy()


def blocks_err(a, b, nz, nx):
    """relative max error of each named part of two statistics vectors; the
    plane-mean w, zero up to round-off in a divergence-free flow, is measured
    against the mean streamwise velocity"""
    from les_amd.api import stats_fields
    fa, fb = stats_fields(a, nz, nx), stats_fields(b, nz, nx)
    out = {k: relerr(fa[k], fb[k]) if np.max(np.abs(fb[k])) > 0 else float(np.max(np.abs(fa[k])))
           for k in fb if k != "wbar"}
    out["wbar"] = float(np.max(np.abs(fa["wbar"] - fb["wbar"])) / np.max(np.abs(fb["ubar"])))
    return out


@pytest.mark.parametrize("grid", GRIDS)
def test_routine_stats_vs_oracle(torch_mod, grid):
    print("statistics routines", grid, flush=True)
    from oracle.oracle import Oracle, scalar_params
    from les_amd.api import Les
    nx, ny, nz = grid
    o, g = Oracle(nx, ny, nz), Les(nx, ny, nz)
    p = scalar_params()
    o.set_scalars(p, None, None)
    g.set_scalars(**{f[0]: getattr(p, f[0]) for f in p._fields_})
    rng = np.random.default_rng(3)
    u, v, w, txz = (o.field() for _ in range(4))
    for a, off in ((u, 5.0), (v, 0.0), (w, 0.0), (txz, -0.5)):
        a[..., :nx] = off + rng.standard_normal(a[..., :nx].shape)
    spec_o = o.spectrum(u)
    spec_g = H(g.spectrum(G(u)))
    assert relerr(spec_g, spec_o) < 1e-12
    st_o = o.avg_stats(u, v, w, txz)
    st_g = H(g.avg_stats(G(u), G(v), G(w), G(txz)))
    errs = blocks_err(st_g, st_o, nz, nx)
    print("  avg_stats:", {k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12, errs
    n2 = (ny, nx)
    psi, psi0, us = rng.uniform(-0.2, 0.2, n2), rng.uniform(-0.1, 0.1, n2), rng.uniform(0.2, 0.5, n2)
    th, q = 0.71 + 1e-3 * rng.standard_normal((ny, o.ld)), 0.01 + 1e-3 * rng.standard_normal((ny, o.ld))
    t3, q3 = 1e-4 * rng.standard_normal((ny, o.ld)) + 2e-4, 1e-5 * rng.standard_normal((ny, o.ld))
    planes = [u[0], v[0], txz[0], w[0]]
    so = o.surface_stats(*planes, psi, psi0, us, th, q, t3, q3)
    sg = H(g.surface_stats(*[G(a) for a in planes], G(psi), G(psi0), G(us), G(th), G(q), G(t3), G(q3)))
    print("  surface:", so, sg, flush=True)
    assert np.allclose(sg, so, rtol=1e-12, atol=0)
    sn = H(g.surface_stats(*[G(a) for a in planes], G(psi), G(psi0), G(us)))
    assert np.allclose(sn[:3], so[:3], rtol=1e-12, atol=0) and np.all(sn[3:] == 0)


def _oracle_state_stats(o, so, scal):
    st = o.avg_stats(so["u"], so["v"], so["w"], so["txz"])
    extra = (so["theta"][0], so["qmix"][0], so["sgs_t3"], so["sgs_q3"]) if scal else ()
    sf = o.surface_stats(so["u"][0], so["v"][0], so["txz"][0], so["tyz"][0], so["psi_m"], so["psi_m0"],
                         so["ustar_avg"], *extra)
    return st, sf


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("conv", [False, True])
def test_state_stats_vs_oracle_state(torch_mod, prec, conv):
    """The state's statistics after 10 steps (two dynamic updates) against
    the oracle's statistics of its own state after the same steps: FP64 at
    1e-10 (the states themselves agree to ~1e-13), FP32-bulk at 5e-4 (the
    variances and <u'w'> are differences of O(u^2) terms)."""
    print("state statistics", prec, conv, flush=True)
    from oracle.oracle import Oracle, OracleSim, scalar_params
    from les_amd.api import LesSim, checkerboard_flux
    nx, ny, nz = 32, 32, 17
    o = Oracle(nx, ny, nz)
    ic = dict(T_init=283.0, z_inv=0.8, inv_strength=0.01, amp_K=0.1, q_init=1.0)
    if conv:
        p, wt = scalar_params(), checkerboard_flux(nx, ny)
        o.set_scalars(p, wt, None)
        so = OracleSim(o, scalar_ic=ic)
        sg = LesSim(nx, ny, nz, precision=prec, scalars=dict({f[0]: getattr(p, f[0]) for f in p._fields_}, wt_s=wt))
        sg.init_ic()
        sg.init_scalar_ic(**ic)
    else:
        so = OracleSim(o)
        sg = LesSim(nx, ny, nz, precision=prec)
        sg.init_ic()
    so.step(10)
    sg.step(10)
    st_o, sf_o = _oracle_state_stats(o, so, conv)
    st_g, sf_g = sg.stats()
    tol = 1e-10 if prec == "fp64" else 5e-4
    errs = blocks_err(st_g, st_o, nz, nx)
    print("  avg_stats:", {k: f"{e:.1e}" for k, e in errs.items()}, flush=True)
    assert max(errs.values()) < tol, errs
    n_sf = 8 if conv else 3
    se = np.abs(sf_g[:n_sf] - sf_o[:n_sf]) / np.abs(sf_o[:n_sf])
    print("  surface:", sf_g[:n_sf], "rel", se, flush=True)
    assert np.all(se < (1e-10 if prec == "fp64" else 1e-4)), se
    if not conv:
        assert np.all(sf_g[3:] == 0)


def test_state_stats_running_means(torch_mod):
    """les_state_stats_accumulate keeps (n avg + cur)/(n+1) on the device:
    after three samples it equals the mean of the three snapshots; reset
    starts a new average."""
    print("running statistics", flush=True)
    from les_amd.api import LesSim
    s = LesSim(32, 32, 17, precision="fp64")
    s.init_ic(seed=9)
    snaps, surf = [], []
    for _ in range(3):
        s.step(2)
        a, b = s.stats()
        snaps.append(a)
        surf.append(b)
        s.stats_accumulate()
    avg, savg, n = s.stats_average(reset=True)
    assert n == 3
    assert np.allclose(avg, np.mean(snaps, axis=0), rtol=1e-14, atol=1e-300)
    assert np.allclose(savg, np.mean(surf, axis=0), rtol=1e-14, atol=0)
    s.step(1)
    s.stats_accumulate()
    avg2, _, n2 = s.stats_average()
    assert n2 == 1 and np.array_equal(avg2, s.stats()[0])
