"""SGS models 1, 2, 4 and 6 and the Coriolis / sponge forcing on the GPU
against the oracle: routine level (std_dynamic, lagrange_Ssim,
interpolag_Ssim, new_scaledep_dynamic, scalar_scaledep_dynamic, the sponge
profile) and whole steps with each switch, FP64 and FP32-bulk."""
import numpy as np
import pytest

from conftest import FLIP_ULPS, MULTI_STEP_L2, flip_check, float_ulps, relerr

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


@pytest.mark.parametrize("grid", GRIDS)
def test_new_scaledep_dynamic(torch_mod, grid):
    """model 6: per-level Cs_opt2 from four plane sums (a tree of partial sums
    on the GPU, compensated in the oracle) through single-precision clips of
    Cs2_2d, Cs2_4d and Beta: equal up to a flip of one of those roundings,
    which moves Cs_opt2 by a few float ulps"""
    print("new_scaledep_dynamic", grid, flush=True)
    o, g = mk(*grid)
    S, u, v, w = _inputs(o, 40)
    cs_o = o.new_scaledep_dynamic(S, u, v, w)
    cs_g = H(g.new_scaledep_dynamic([G(a) for a in S], G(u), G(v), G(w)))
    ul = float_ulps(cs_g, cs_o)
    print("  Cs:", cs_o, " float ulps:", ul, flush=True)
    assert np.any(cs_o > 0)
    assert np.all(ul <= FLIP_ULPS["Cs_opt2"]) and np.mean(ul > 0) <= 0.5


@pytest.mark.parametrize("grid", GRIDS)
def test_scalar_scaledep_dynamic(torch_mod, grid):
    """model 6 with scalars: theta filtered in place (1e-12), Cs2_Pr per
    level in the first nz-1 values of the Cs2_Pr array and Pr_t = Cs_opt2 /
    Cs2_Pr on levels 1..nz-1 (a flip of the clipped ratios moves them by a
    few float ulps); the lid level of Pr_t is left as it was"""
    print("scalar_scaledep_dynamic", grid, flush=True)
    from oracle.oracle import scalar_params
    from les_amd.api import checkerboard_flux
    o, g = mk(*grid)
    nx, ny, nz = o.nx, o.ny, o.nz
    wt = checkerboard_flux(nx, ny)
    o.set_sgs_model(6)
    o.set_scalars(scalar_params(model=6), wt, None)
    g.set_sgs_model(6)
    g.set_scalars(enable=1, model=6, wt_s=wt)
    S, u, v, w = _inputs(o, 50)
    rng = np.random.default_rng(51)
    th = o.field()
    th[..., :nx] = 0.7 + 0.01 * rng.standard_normal(th[..., :nx].shape)
    cs = o.field()
    cs[..., :nx] = 0.01 + 0.02 * rng.random(cs[..., :nx].shape)
    psi_m = 0.1 * rng.standard_normal((ny, nx))
    phi_h = 1.0 + 0.1 * rng.random((ny, nx))
    pr_o = o.field()
    pr_o[nz - 1] = 0.33
    th_g = G(th)
    ss = g.scalar_sgs_fields(phi_h=G(phi_h), psi_m=G(psi_m))
    ss["Pr_t"].copy_(G(pr_o))
    cpr_o = o.scalar_scaledep_dynamic(S, u, v, w, th, cs, psi_m, phi_h, pr_o)
    g.scalar_scaledep_dynamic([G(a) for a in S], G(u), G(v), G(w), th_g, G(cs), ss)
    assert relerr(H(th_g)[..., :nx], th[..., :nx]) < 1e-12
    cpr_g = H(ss["Cs2_Pr"]).ravel()[:nz - 1]
    ul = float_ulps(cpr_g, cpr_o)
    print("  Cs2_Pr:", cpr_o, " float ulps:", ul, flush=True)
    assert np.all(cpr_o > 0) and np.all(ul <= FLIP_ULPS["Cs2_Pr"])
    pr_g = H(ss["Pr_t"])
    flip_check(pr_g[:nz - 1, :, :nx], pr_o[:nz - 1, :, :nx], "Pr_t", rel=1e-12, frac=0.5,
               max_ulps=FLIP_ULPS["Pr_t"], l2=1e-6)
    assert np.all(pr_g[nz - 1] == 0.33)


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
    for m in (0, 7):
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
    g.set_scalars(enable=1, model=5)          # constant Pr with models 1-4
    with pytest.raises(LesError):
        g.set_scalars(enable=1, model=6)      # scalar model 6 goes with SGS model 6 only
    h = Les(32, 32, 9)
    h.set_sgs_model(6)
    with pytest.raises(LesError):
        h.set_scalars(enable=1, model=5)      # SGS model 6 takes Pr_t from scalar_scaledep_dynamic
    h.set_scalars(enable=1, model=6)


def _pair(prec, model=5, forcing=None, conv=False, Lz=1.0, seed=42, smodel=5):
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
        p = scalar_params(model=smodel)
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
@pytest.mark.parametrize("model", [1, 2, 4, 6])
def test_step_sgs_models(torch_mod, prec, model):
    """10 steps (dynamic updates at jt = 5, 10) with SGS model 1, 2, 4 or 6
    (model 6 still inside the constant period of jt_total < 500)"""
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


@pytest.mark.parametrize("model", [1, 2, 4, 6])
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


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("conv", [False, True])
def test_step_model6_crossing_dynamic_switch(torch_mod, prec, conv):
    """SGS model 6 from jt_total = 490 across 500, where Cs_opt2 (and with
    scalars Pr_t, scalar model 6) switch from the constants to
    new_scaledep_dynamic / scalar_scaledep_dynamic (sgs_stag.f90:143-173):
    state after 10 and 16 steps (updates at jt = 5, 10, 15)"""
    print("model 6 across jt_total = 500", prec, "convective" if conv else "neutral", flush=True)
    so, sg = _pair(prec, model=6, conv=conv, smodel=6)
    so.set_time(0, 490)
    sg.set_time(0, 490)
    nx, nz = sg.nx, sg.nz
    names = ("u", "v", "w", "p") + (("theta", "qmix") if conv else ())
    for n, tol in ((10, 1e-10 if prec == "fp64" else 1e-4), (6, 1e-9 if prec == "fp64" else 2e-4)):
        so.step(n)
        sg.step(n)
        _compare(so, sg, prec, names, tol)
        for name in ("Cs_opt2", "Nu_t") + (("Pr_t",) if conv else ()):
            a, b = sg.get(name)[:nz - 1, :, :nx], so[name][:nz - 1, :, :nx]
            if prec == "fp64":
                flip_check(a, b, name, rel=1e-9, max_ulps=None, l2=MULTI_STEP_L2)
                continue
            # FP32-bulk: a level's plane sum L.M can cancel to near zero, where
            # the 1E-24 clip makes Cs_opt2 jump between ~1e-9 and ~0 on the
            # fields' single-precision rounding (up to a few 1e-3 of the
            # largest level); Pr_t spans 1e-3..1e37 with Cs2_Pr, so it is
            # compared through the scalar diffusivity Nu_t / Pr_t it sets
            if name == "Pr_t":
                a = sg.get("Nu_t")[:nz - 1, :, :nx] / a
                b = so["Nu_t"][:nz - 1, :, :nx] / b
            e = relerr(a, b)
            print(f"  {name}: {e:.2e}", flush=True)
            assert e < 2e-2, name
    cs = so["Cs_opt2"][:nz - 1, :, :nx]
    assert np.ptp(cs) > 0 and np.all(np.ptp(cs, axis=(1, 2)) == 0)   # one dynamic value per level
    if conv:
        assert not np.all(so["Pr_t"][:nz - 1, :, :nx] == 0.55)
