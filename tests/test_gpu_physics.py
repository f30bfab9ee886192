"""SGS models 1, 2 and 4 and the Coriolis / sponge forcing on the GPU against
the oracle: routine level (std_dynamic, lagrange_Ssim, interpolag_Ssim, the
sponge profile) and whole steps with each switch, FP64 and FP32-bulk."""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu

GRIDS = [(32, 32, 9), (48, 32, 7)]
TOL = 1e-12


def mk(nx, ny, nz, **kw):
    from oracle.oracle import Oracle
    from les_amd.api import Les
    return Oracle(nx, ny, nz, **kw), Les(nx, ny, nz, **kw)


def rnd(o, seed):
    rng = np.random.default_rng(seed)
    a = o.field()
    a[...] = rng.standard_normal(a.shape)
    return a


def G(t):
    import torch
    return torch.from_numpy(np.ascontiguousarray(t)).cuda()


def H(t):
    return t.double().cpu().numpy()


def _inputs(o, seed):
    u, v, w = rnd(o, seed) + 3.0, rnd(o, seed + 1), rnd(o, seed + 2)
    S = [0.5 * rnd(o, seed + 3 + i) for i in range(6)]
    for a in S:
        a[..., o.nx:] = 0.0
    return S, u, v, w


def _sgs_pair(o, g, seed):
    st, arr = o.sgs_state()
    rng = np.random.default_rng(seed)
    for n in ("F_LM", "F_MM", "F_QN", "F_NN"):
        arr[n][...] = 0.1 + rng.random(arr[n].shape)
    for n in ("u_lag", "v_lag", "w_lag"):
        arr[n][...] = 5.0 * rng.standard_normal(arr[n].shape)
    gd = g.sgs_fields()
    for n in gd:
        gd[n].copy_(G(arr[n]))
    return st, arr, gd


@pytest.mark.parametrize("grid", GRIDS)
def test_std_dynamic(torch_mod, grid):
    """per-level Cs_opt2 = max(0, sum L:M / sum M:M): the sums differ only in
    their order (a tree of partial sums on the GPU, compensated in the oracle)"""
    print("std_dynamic", grid, flush=True)
    o, g = mk(*grid)
    S, u, v, w = _inputs(o, 10)
    cs_o = o.std_dynamic(S, u, v, w)
    cs_g = H(g.std_dynamic([G(a) for a in S], G(u), G(v), G(w)))
    print("  Cs:", cs_o, flush=True)
    assert np.any(cs_o > 0)
    assert np.max(np.abs(cs_g - cs_o)) <= 1e-10 * np.max(np.abs(cs_o))


@pytest.mark.parametrize("grid", GRIDS)
def test_interpolag_ssim(torch_mod, grid):
    print("interpolag_ssim", grid, flush=True)
    o, g = mk(*grid, dt=2e-3)
    st, arr, gd = _sgs_pair(o, g, 20)
    o.interpolag_ssim(st)
    g.interpolag_ssim(gd)
    for n in ("F_LM", "F_MM", "u_lag", "v_lag", "w_lag"):
        assert relerr(H(gd[n])[..., :o.nx], arr[n][..., :o.nx]) < TOL, n
    for n in ("F_QN", "F_NN"):   # not touched by the scale-invariant model
        assert np.array_equal(H(gd[n]), arr[n]), n


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("init", [True, False])
def test_lagrange_ssim(torch_mod, grid, init):
    print("lagrange_ssim", grid, init, flush=True)
    o, g = mk(*grid, dt=2e-3)
    st, arr, gd = _sgs_pair(o, g, 30)
    S, u, v, w = _inputs(o, 31)
    o.set_time(5, 600, flm_init=not init, fqn_init=True)
    g.set_time(5, 600, flm_init=not init, fqn_init=True)
    o.lagrange_ssim(S, u, v, w, st)
    g.lagrange_ssim([G(a) for a in S], G(u), G(v), G(w), gd)
    for n in ("F_LM", "F_MM", "Cs_opt2"):
        e = relerr(H(gd[n])[..., :o.nx], arr[n][..., :o.nx])
        print(f"  {n}: {e:.1e}", flush=True)
        assert e < 1e-11, n
    for n in ("Beta", "Cs_opt2"):
        assert np.array_equal(H(gd[n])[..., o.nx:], arr[n][..., o.nx:]), n
    assert np.all(H(gd["Beta"]) == 1.0)
    assert np.array_equal(H(gd["Beta_avg"]), arr["Beta_avg"])
    assert np.array_equal(H(gd["Betaclip_avg"]), arr["Betaclip_avg"])


@pytest.mark.parametrize("method", [1, 2])
def test_sponge_profile(torch_mod, method):
    """setsponge runs on the host in both: the same doubles"""
    o, g = mk(32, 32, 33, Lz=3.0)
    kw = dict(ubc=1, damping_method=method, u_star=0.45, z_i=1000.0)
    o.set_forcing(**kw)
    g.set_forcing(**kw)
    assert np.array_equal(g.sponge(), o.sponge())
    assert o.sponge().max() > 0


def test_switch_checks(torch_mod):
    from les_amd.api import Les, LesError
    g = Les(32, 32, 9)
    for m in (0, 6, 7):
        with pytest.raises(LesError):
            g.set_sgs_model(m)
    with pytest.raises(LesError):
        g.set_forcing(ubc=2)
    with pytest.raises(LesError):
        g.set_forcing(ubc=1, damping_method=3)
    with pytest.raises(LesError):
        g.set_forcing(coriolis=1, u_star=0.0)
    g.set_sgs_model(1)
    with pytest.raises(LesError):
        g.set_scalars(enable=1, model=7)      # the scalar Lagrangian model needs lagrange_Sdep
    g.set_scalars(enable=1, model=5)          # constant Pr with any momentum model


def _pair(prec, model=5, forcing=None, conv=False, Lz=1.0, seed=42):
    from oracle.oracle import Oracle, OracleSim, scalar_params
    from les_amd.api import LesSim, checkerboard_flux
    nx, ny, nz = 32, 32, 17
    o = Oracle(nx, ny, nz, Lz=Lz)
    o.set_sgs_model(model)
    if forcing:
        o.set_forcing(**forcing)
    scal = None
    ic = dict(T_init=283.0, z_inv=0.8 * Lz, inv_strength=0.01, amp_K=0.1, q_init=1.0)
    if conv:
        p = scalar_params(model=5)
        wt = checkerboard_flux(nx, ny)
        o.set_scalars(p, wt, None)
        scal = dict({f[0]: getattr(p, f[0]) for f in p._fields_}, wt_s=wt)
    so = OracleSim(o, seed=seed, scalar_ic=ic if conv else None)
    sg = LesSim(nx, ny, nz, precision=prec, sgs_model=model, scalars=scal, forcing=forcing, Lz=Lz)
    sg.init_ic(seed=seed)
    if conv:
        sg.init_scalar_ic(seed=seed, **ic)
    return so, sg


def _compare(so, sg, prec, names, tol):
    nx = sg.nx
    for name in names:
        a, b = sg.get(name), so[name]
        nl = b.shape[0] - 1
        e = relerr(a[:nl, :, :nx], b[:nl, :, :nx])
        print(f"  {name}: {e:.2e}", flush=True)
        assert e < tol, (name, e)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("model", [1, 2, 4])
def test_step_sgs_models(torch_mod, prec, model):
    """10 steps (dynamic updates at jt = 5, 10) with SGS model 1, 2 or 4"""
    print("step model", model, prec, flush=True)
    so, sg = _pair(prec, model=model)
    so.step(10)
    sg.step(10)
    tol = 1e-9 if prec == "fp64" else 1e-4
    _compare(so, sg, prec, ("u", "v", "w", "p"), tol)
    _compare(so, sg, prec, ("Nu_t", "Cs_opt2"), 1e-9 if prec == "fp64" else 5e-3)
    if model == 4:
        _compare(so, sg, prec, ("F_LM", "F_MM"), 1e-9 if prec == "fp64" else 5e-3)


FORCINGS = {
    "coriolis": dict(coriolis=1, ug0=8.0, vg0=-1.0, u_star=0.45, z_i=1000.0),
    "sponge1": dict(ubc=1, damping_method=1, u_star=0.45, z_i=1000.0),
    "sponge2+coriolis": dict(coriolis=1, ug0=8.0, vg0=0.5, ubc=1, damping_method=2, u_star=0.45, z_i=1000.0),
}


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", sorted(FORCINGS))
def test_step_forcing(torch_mod, prec, case):
    """10 steps of a 3 z_i deep domain with the Coriolis force and / or the
    sponge of the top quarter"""
    print("step forcing", case, prec, flush=True)
    so, sg = _pair(prec, forcing=FORCINGS[case], Lz=3.0)
    so.step(10)
    sg.step(10)
    _compare(so, sg, prec, ("u", "v", "w", "p"), 1e-9 if prec == "fp64" else 1e-4)
    # FP32-bulk RHS: differences of O(u^2) advection terms plus the O(coriol ug)
    # Coriolis terms, each rounded in single precision
    _compare(so, sg, prec, ("RHSx", "RHSy", "RHSz"), 1e-9 if prec == "fp64" else 2e-3)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_step_convective_sponge(torch_mod, prec):
    """the convective step with damping_method 2: theta and q relax to their
    plane means in the sponge, momentum as above, with SGS model 1"""
    print("step convective sponge", prec, flush=True)
    f = dict(coriolis=1, ug0=5.0, vg0=0.0, ubc=1, damping_method=2, u_star=1.0, z_i=1000.0)
    so, sg = _pair(prec, model=1, forcing=f, conv=True, Lz=3.0)
    # the same humidity perturbation in both (the initial q is uniform, so the
    # q sponge would otherwise act on nothing)
    import torch
    nx, nz = sg.nx, sg.nz
    pert = 1e-3 * np.random.default_rng(5).standard_normal((nz - 1, sg.ny, nx))
    so["qmix"][:nz - 1, :, :nx] += pert
    qv = sg.view("qmix")
    qv[:nz - 1, :, :nx] += torch.from_numpy(pert).to(device=qv.device, dtype=qv.dtype)
    so.step(10)
    sg.step(10)
    tol = 1e-9 if prec == "fp64" else 1e-4
    _compare(so, sg, prec, ("u", "v", "w", "p"), tol)
    for name in ("theta", "qmix"):
        a, b = sg.get(name), so[name]
        nl = b.shape[0] - 1
        da, db = a[:nl, :, :sg.nx] - b[:nl, :, :sg.nx].mean(), b[:nl, :, :sg.nx] - b[:nl, :, :sg.nx].mean()
        e = np.max(np.abs(da - db)) / np.max(np.abs(db))
        print(f"  {name} deviations: {e:.2e}", flush=True)
        assert e < (1e-9 if prec == "fp64" else 1e-3), name


@pytest.mark.parametrize("model", [1, 2, 4])
def test_graph_replay_with_switches(torch_mod, model):
    """captured step windows replay the model and forcing branches bit for bit"""
    print("graph replay model", model, flush=True)
    from les_amd.api import LesSim
    f = FORCINGS["sponge2+coriolis"]
    a = LesSim(32, 32, 17, sgs_model=model, forcing=f, Lz=3.0)
    b = LesSim(32, 32, 17, sgs_model=model, forcing=f, Lz=3.0, graph=True)
    for s in (a, b):
        s.init_ic(seed=7)
    a.step(30)
    for _ in range(3):
        b.step(10)
    for name in ("u", "v", "w", "p", "Cs_opt2", "Nu_t", "RHSx"):
        assert np.array_equal(a.get(name), b.get(name)), name
