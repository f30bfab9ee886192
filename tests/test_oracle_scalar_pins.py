"""Pins for the scalar part of the CPU oracle (obukhov, calcbeta,
scalar_RHS_calc, step_scalar, lagrange_scalar_Sdep, the model-7 step):
  * Monin-Obukhov functions against their defining relation
    d(psi)/d(zeta) = (1 - phi(zeta))/zeta, which ties the two independent
    outputs psi and phi together, and against the neutral limit;
  * the Obukhov length against its definition L = -u*^3 theta / (kappa g w'theta');
  * buoyancy against single Fourier modes (exact plane means);
  * scalar advection and SGS diffusion against single-mode and linear-in-z
    closed forms (dealiased products of band-limited fields are exact);
  * the scalar Lagrangian model against invariants (Galilean, a constant
    added to theta, x<->y symmetry) and the relaxation of the
    Meneveau-Lund-Cabot time scale;
  * the whole convective step against horizontal translation invariance.
"""
import math

import numpy as np
import pytest

from conftest import relerr

VONK = 0.41


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def grid_xy(o):
    x = np.arange(o.nx) * o.dx
    y = np.arange(o.ny) * o.dy
    return np.meshgrid(x, y)


def checker(nx, ny, hi=0.1, lo=0.02, patches=4):
    px = (np.arange(nx) * patches) // nx
    py = (np.arange(ny) * patches) // ny
    return np.where((px[None, :] + py[:, None]) % 2 == 0, hi, lo)


def mk(O, nx=16, ny=16, nz=9, wt=None, **sp):
    o = O.Oracle(nx, ny, nz, dt=2e-4)
    o.set_scalars(O.scalar_params(**sp), wt, None)
    return o


def _obukhov(o, U, theta0, wt, jt=2):
    """uniform u = U, theta = theta0 (nondim), sgs_t3 = wt: returns outputs."""
    u = np.full((o.nz, o.ny, o.ld), U)
    v = o.field()
    th = np.full((o.nz, o.ny, o.ld), theta0)
    q = o.field()
    t3 = np.full((o.ny, o.ld), wt)
    q3 = np.zeros((o.ny, o.ld))
    out = {n: np.zeros((o.ny, o.nx)) for n in ("psi_m", "psi_m0", "psi_h", "psi_h0", "phi_m", "phi_h", "ustar")}
    o.set_time(jt, jt)
    o.obukhov(u, v, th, q, t3, q3, out["psi_m"], out["psi_m0"], out["psi_h"], out["psi_h0"], out["phi_m"],
              out["phi_h"], out["ustar"])
    return out


def test_obukhov_neutral_limit(O):
    o = mk(O)
    r = _obukhov(o, 3.0, 0.7, 0.0)
    assert np.all(r["psi_m"] == 0) and np.all(r["phi_m"] == 1) and np.all(r["phi_h"] == 1)
    # patch_flag=1 log law with psi = 0 (old values): u* = U kappa / ln(dz/2/z0)
    assert np.allclose(r["ustar"], 3.0 * VONK / math.log(0.5 * o.dz / 1e-4), rtol=1e-14)


def _zeta_of(o, U, theta0, wt, sp):
    """Obukhov length from its definition with the dimensionless gravity
    g z_i/u_star^2, evaluated at z = dz/2."""
    us = U * VONK / math.log(0.5 * o.dz / 1e-4)
    g_ = sp.g * sp.z_i / sp.u_star ** 2
    L = -us ** 3 * theta0 / (VONK * g_ * wt)
    return 0.5 * o.dz / L, 1e-4 / L, 1e-4 / 7.39 / L


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_obukhov_psi_phi_integral_relation(O, sign):
    """Sweep the surface flux so zeta covers (-5, 0) (unstable) or (0, 2)
    (stable); psi and phi must satisfy psi(zeta) = int_0^zeta (1-phi)/z dz,
    for momentum and heat, and psi_m0 / psi_h0 are the same functions at
    zeta0 = z0/L and zeta0s = z0/7.39/L."""
    o = mk(O)
    sp = O.scalar_params()
    U, th = 3.0, 0.7
    z1, _, _ = _zeta_of(o, U, th, 1.0, sp)          # zeta is linear in the flux
    targets = -sign * np.geomspace(1e-5, 5.0 if sign > 0 else 1.5, 200)
    wts = targets / z1
    Z, pm, ph, fm, fh, pm0, ph0, z0s, z0hs = [], [], [], [], [], [], [], [], []
    for wt in wts:
        r = _obukhov(o, U, th, wt, jt=2)
        z, z0, z0h = _zeta_of(o, U, th, wt, sp)
        Z.append(z); z0s.append(z0); z0hs.append(z0h)
        pm.append(r["psi_m"][0, 0]); ph.append(r["psi_h"][0, 0])
        fm.append(r["phi_m"][0, 0]); fh.append(r["phi_h"][0, 0])
        pm0.append(r["psi_m0"][0, 0]); ph0.append(r["psi_h0"][0, 0])
    Z, pm, ph, fm, fh = map(np.asarray, (Z, pm, ph, fm, fh))
    keep = (Z > -30) & (Z < 2)
    Z, pm, ph, fm, fh = Z[keep], pm[keep], ph[keep], fm[keep], fh[keep]
    order = np.argsort(Z)
    Z, pm, ph, fm, fh = Z[order], pm[order], ph[order], fm[order], fh[order]
    # the integrand (1 - phi)/zeta has a finite limit at 0; integrate from
    # the smallest |zeta| point with the trapezoid rule on a log-spaced grid
    for psi, phi in ((pm, fm), (ph, fh)):
        integrand = (1.0 - phi) / Z
        if sign > 0:   # unstable: zeta < 0, integrate from the point nearest 0 (last) downward
            ref = psi[-1]
            cum = np.concatenate([[0.0], np.cumsum(0.5 * (integrand[1:] + integrand[:-1]) * np.diff(Z))])
            pred = ref - (cum[-1] - cum)
        else:
            ref = psi[0]
            cum = np.concatenate([[0.0], np.cumsum(0.5 * (integrand[1:] + integrand[:-1]) * np.diff(Z))])
            pred = ref + cum
        # psi(0) = 0: the reference point is within O(zeta_min) of zero
        assert abs(ref) < 5e-3
        assert np.max(np.abs(pred - psi)) < 5e-3 * max(1.0, np.max(np.abs(psi)))
    if sign > 0:
        # Businger-Dyer: phi_h = phi_m^2 in unstable conditions
        assert np.allclose(fh, fm ** 2, rtol=1e-12)
    else:
        assert np.allclose(fh, fm, rtol=0) and np.allclose(pm, ph, rtol=0)
    # psi_m0 is psi_m evaluated at zeta0 (interpolate the sweep)
    z0s, pm0 = np.asarray(z0s)[keep][order], np.asarray(pm0)[keep][order]
    inside = (z0s > Z.min()) & (z0s < Z.max())
    if inside.any():
        assert np.allclose(np.interp(z0s[inside], Z, pm), pm0[inside], atol=2e-3)


def test_obukhov_clamps_zeta(O):
    """zeta is clamped to [-30, 2] (scalars_module.f90:583-605): a stronger
    flux than the one that reaches the clamp changes nothing."""
    o = mk(O)
    sp = O.scalar_params()
    a = _obukhov(o, 0.5, 0.7, 5.0)
    b = _obukhov(o, 0.5, 0.7, 50.0)
    z, _, _ = _zeta_of(o, 0.5, 0.7, 5.0, sp)
    assert z < -30
    for k in ("psi_m", "phi_m", "psi_h", "phi_h"):
        assert np.array_equal(a[k], b[k]), k
    # stable side: zeta clamped at 2 -> phi = 1 + 5*2
    c = _obukhov(o, 0.3, 0.7, -5.0)
    assert np.allclose(c["phi_m"], 11.0) and np.allclose(c["psi_m"], -10.0)


def test_obukhov_first_step_sets_surface_flux(O):
    """jt == SCAL_init without initsc: sgs_t3(1) = wt_s/(u* T_scale) and the
    psi arrays restart from 0 (scalars_module.f90:521-538)."""
    wt = checker(16, 16)
    o = mk(O, wt=wt)
    u = np.full((o.nz, o.ny, o.ld), 2.0)
    th = np.full((o.nz, o.ny, o.ld), 0.7)
    t3, q3 = np.zeros((o.ny, o.ld)), np.zeros((o.ny, o.ld))
    arrs = [np.full((o.ny, o.nx), 0.3) for _ in range(7)]
    o.set_time(1, 1)
    o.obukhov(u, o.field(), th, o.field(), t3, q3, *arrs)
    assert np.allclose(t3[:, :o.nx], wt / 400.0, rtol=1e-15)
    assert np.all(q3 == 0)
    # every point is unstable (wt > 0): phi_m < 1
    assert np.all(arrs[4] < 1.0)


def test_calcbeta_single_mode(O):
    """theta = t0 + a cos(kx x) + b(z), q = q0: plane means are exact, so
    beta = g_hat a cos(kx x) ((1/tb_k) + (1/tb_{k-1}))/2 on k=2..nz-1, 0 on
    levels 1 and nz (virtual factor 1 + 0.61 q)."""
    o = mk(O, nx=16, ny=8, nz=7)
    sp = O.scalar_params()
    X, Y = grid_xy(o)
    a, q0 = 0.01, 0.012
    th = o.field(); q = o.field()
    tb = [0.7 + 0.01 * k for k in range(o.nz)]
    for k in range(o.nz):
        th[k, :, :o.nx] = tb[k] + a * np.cos(2 * X)
        q[k, :, :o.nx] = q0
    beta = o.field()
    o.calcbeta(th, q, beta)
    gh = sp.g * sp.z_i / sp.u_star ** 2
    fv = 1 + 0.61e-3 * q0 * sp.q_scale
    for k in range(1, o.nz - 1):
        ex = gh * (a * fv * np.cos(2 * X) / (tb[k] * fv) + a * fv * np.cos(2 * X) / (tb[k - 1] * fv)) / 2
        assert np.max(np.abs(beta[k, :, :o.nx] - ex)) < 1e-12 * gh
    assert np.all(beta[0] == 0) and np.all(beta[-1] == 0)


def _rhs(o, which, th, dsdx, dsdy, dsdz, u, v, w, nu, pr, ustar=None, phi_h=None):
    R = o.field()
    t3 = np.zeros((o.ny, o.ld))
    us = np.full((o.ny, o.nx), 1.0) if ustar is None else ustar
    ph = np.ones((o.ny, o.nx)) if phi_h is None else phi_h
    o.scalar_rhs_calc(which, dsdx, dsdy, dsdz, u, v, w, nu, pr, us, ph, R, t3)
    return R, t3


def test_scalar_advection_single_mode(O):
    """u = U, v = V, w = 0, theta = sin(2x + 3y), no SGS: RHS = -(U d/dx + V d/dy) theta."""
    o = mk(O, nx=16, ny=16, nz=6)
    X, Y = grid_xy(o)
    U, V = 1.3, -0.4
    th_xy = np.sin(2 * X + 3 * Y)
    dsdx = o.field(); dsdy = o.field(); dsdz = o.field()
    u = o.field(); v = o.field(); w = o.field()
    for k in range(o.nz):
        dsdx[k, :, :o.nx] = 2 * np.cos(2 * X + 3 * Y)
        dsdy[k, :, :o.nx] = 3 * np.cos(2 * X + 3 * Y)
        u[k, :, :o.nx] = U
        v[k, :, :o.nx] = V
    nu = o.field()
    pr = np.full_like(nu, 0.55)
    R, _ = _rhs(o, 0, None, dsdx, dsdy, dsdz, u, v, w, nu, pr)
    ex = -(U * 2 + V * 3) * np.cos(2 * X + 3 * Y)
    for k in range(o.nz - 1):
        assert np.max(np.abs(R[k, :, :o.nx] - ex)) < 1e-12, k
    assert np.all(R[-1] == 0)


def test_scalar_vertical_advection_staggering(O):
    """Plane-uniform w(z) and dsdz(z): RHS(k) = -(w_k g_k + w_{k+1} g_{k+1})/2,
    and at the first level -(w_2/2) g_2 (scalars_module.f90:322-338)."""
    o = mk(O, nx=8, ny=8, nz=7)
    W = [0.0, 0.3, -0.2, 0.5, 0.1, 0.25, 0.0]
    Gz = [9.0, 1.5, 2.0, -1.0, 0.5, 3.0, 7.0]   # level 1 replaced by the surface gradient
    u = o.field(); v = o.field(); w = o.field(); dsdz = o.field()
    for k in range(o.nz):
        w[k, :, :o.nx] = W[k]
        dsdz[k, :, :o.nx] = Gz[k]
    nu = o.field(); pr = np.full_like(nu, 0.55)
    R, _ = _rhs(o, 0, None, o.field(), o.field(), dsdz, u, v, w, nu, pr)
    # wt_s = 0 -> surface gradient 0 at level 1
    assert np.allclose(R[0, :, :o.nx], -(0.5 * W[1]) * Gz[1], rtol=1e-13)
    for k in range(1, o.nz - 1):
        assert np.allclose(R[k, :, :o.nx], -(W[k] * Gz[k] + W[k + 1] * Gz[k + 1]) / 2, rtol=1e-12, atol=1e-14), k


def test_scalar_sgs_diffusion_closed_form(O):
    """Constant Nu_t and Pr_, theta = sin(3x) + G z: horizontal SGS term
    -(Nu/Pr) 9 sin(3x); vertical flux constant above the wall, so only the
    first level feels the surface flux: ((Nu/Pr) G + wt/(T_scale u*))/dz."""
    wt = 0.05
    o = mk(O, nx=16, ny=8, nz=6, wt=np.full((8, 16), wt))
    X, Y = grid_xy(o)
    nu_c, pr_c, G = 0.02, 0.55, 1.7
    dsdx = o.field(); dsdy = o.field(); dsdz = o.field()
    for k in range(o.nz):
        dsdx[k, :, :o.nx] = 3 * np.cos(3 * X)
        dsdz[k, :, :o.nx] = G
    nu = np.full((o.nz, o.ny, o.ld), nu_c)
    pr = np.full_like(nu, pr_c)
    z = o.field()
    us = np.full((o.ny, o.nx), 0.8)
    phi = np.full((o.ny, o.nx), 0.9)
    R, t3 = _rhs(o, 0, None, dsdx, dsdy, dsdz, z, z, z, nu, pr, ustar=us, phi_h=phi)
    sf = wt / 400.0
    hor = -(nu_c / pr_c) * 9 * np.sin(3 * X)
    assert np.max(np.abs(R[0, :, :o.nx] - (hor + ((nu_c / pr_c) * G + sf) / o.dz))) < 1e-11
    for k in range(1, o.nz - 1):
        assert np.max(np.abs(R[k, :, :o.nx] - hor)) < 1e-11, k
    assert np.allclose(t3[:, :o.nx], sf)
    # surface gradient written into dsdz(1): -phi_h flux / (u* kappa dz/2)
    assert np.allclose(dsdz[0, :, :o.nx], -0.9 * sf / (0.8 * VONK * o.dz / 2), rtol=1e-14)


def test_humidity_surface_flux_scaling(O):
    """q uses wq_s/q_scale/u_star (scalars_module_q.f90:139-145)."""
    o = O.Oracle(8, 8, 5, dt=2e-4)
    o.set_scalars(O.scalar_params(), None, np.full((8, 8), 0.3))
    z = o.field(); nu = o.field(); pr = np.full_like(nu, 0.55)
    R, t3 = _rhs(o, 1, None, o.field(), o.field(), o.field(), z, z, z, nu, pr)
    assert np.allclose(t3[:, :8], 0.3 / 100.0, rtol=1e-15)


@pytest.mark.parametrize("Lz", [1.0, 3.0])
def test_step_scalar_ab2_and_lid(O, Lz):
    """AB2 on levels 1..nz-1 over the physical columns; the lid level takes the
    inversion gradient only when the domain is taller than z_i."""
    o = O.Oracle(8, 8, 5, dt=2e-4, Lz=Lz)
    o.set_scalars(O.scalar_params(dTdz_top=0.01), None, None)
    rng = np.random.default_rng(0)
    s = rng.standard_normal((o.nz, o.ny, o.ld))
    R = rng.standard_normal(s.shape)
    Rf = rng.standard_normal(s.shape)
    s0 = s.copy()
    o.step_scalar(0, s, R, Rf)
    assert np.allclose(s[:-1, :, :8], s0[:-1, :, :8] + 2e-4 * (1.5 * R[:-1, :, :8] - 0.5 * Rf[:-1, :, :8]), rtol=1e-15)
    assert np.array_equal(s[:-1, :, 8:], s0[:-1, :, 8:])
    lid = 0.01 / 400 * 1000 * o.dz if Lz > 1 else 0.0
    assert np.allclose(s[-1], s[-2] + lid, rtol=1e-15)


# ------------------------------------------------ scalar Lagrangian model
def _lag_scalar_inputs(o, seed):
    rng = np.random.default_rng(seed)
    X, Y = grid_xy(o)
    f = lambda: o.field()
    u, v, w, th = f(), f(), f(), f()
    S = [f() for _ in range(6)]
    for k in range(o.nz):
        for a in (u, v, w, th):
            ph = rng.uniform(0, 2 * np.pi, 3)
            a[k, :, :o.nx] = (np.sin(X + ph[0]) + 0.5 * np.cos(2 * Y + ph[1]) + 0.3 * np.sin(3 * X + 2 * Y + ph[2]))
        for a in S:
            ph = rng.uniform(0, 2 * np.pi, 2)
            a[k, :, :o.nx] = 0.4 * np.sin(2 * X + ph[0]) + 0.2 * np.cos(X + 3 * Y + ph[1])
    th += 0.7
    return u, v, w, th, S


def _lag_scalar_run(O, o, u, v, w, th, S, seed, Tn=None):
    st, arr = o.sgs_state()
    rng = np.random.default_rng(seed)
    arr["TnLMMM"][...] = Tn if Tn is not None else rng.uniform(0.5, 2.0, arr["TnLMMM"].shape)
    arr["TnQNNN"][...] = rng.uniform(0.5, 2.0, arr["TnQNNN"].shape)
    arr["Cs_opt2"][...] = rng.uniform(0.01, 0.05, arr["Cs_opt2"].shape)
    sst, sarr = o.scalar_sgs_state()
    for n in ("G_LM", "G_MM", "G_QN", "G_NN"):
        sarr[n][...] = rng.uniform(0.1, 1.0, sarr[n].shape)
    o.set_time(5, 600)
    o.set_scalar_lag_init(1, 1)
    o.lagrange_scalar_sdep(S, u, v, w, th, st, sst)
    return arr, sarr


def test_lagrange_scalar_galilean_and_theta_offset(O):
    """L_i = filt(u theta) - filt(u) filt(theta) is invariant under u -> u + c
    and theta -> theta + c; M, N depend on gradients only."""
    o = mk(O, nx=16, ny=16, nz=5)
    outs = []
    for du, dth in ((0.0, 0.0), (2.5, 0.0), (0.0, 0.3)):
        u, v, w, th, S = _lag_scalar_inputs(o, 3)
        u[..., :o.nx] += du
        th[..., :o.nx] += dth
        arr, sarr = _lag_scalar_run(O, o, u, v, w, th, S, 5)
        outs.append({k: sarr[k][..., :o.nx].copy() for k in ("G_LM", "G_MM", "G_QN", "G_NN", "Cs2_Pr", "Pr_t")})
    for o2 in outs[1:]:
        for k in ("G_MM", "G_NN"):
            assert relerr(o2[k], outs[0][k]) < 1e-9, k
        for k in ("G_LM", "G_QN", "Cs2_Pr", "Pr_t"):   # FP32 rounding points
            bad = np.abs(o2[k] - outs[0][k]) > 1e-9 * np.abs(outs[0][k]) + 2e-7 * np.abs(outs[0][k])
            assert bad.mean() < 0.02, k


def test_lagrange_scalar_xy_symmetry(O):
    o = mk(O, nx=16, ny=16, nz=5)
    u, v, w, th, S = _lag_scalar_inputs(o, 9)
    T = lambda a: np.ascontiguousarray(np.pad(np.transpose(a[..., :o.nx], (0, 2, 1)), ((0, 0), (0, 0), (0, 2))))
    rng = np.random.default_rng(2)
    Tn = rng.uniform(0.5, 2.0, (o.nz, o.ny, o.ld))
    a0, s0 = _lag_scalar_run(O, o, u, v, w, th, S, 4, Tn=Tn)
    # transposed inputs with transposed state (seeded identically, then transposed)
    st, arr = o.sgs_state()
    rng = np.random.default_rng(4)
    arr["TnLMMM"][...] = T(Tn)
    arr["TnQNNN"][...] = T(rng.uniform(0.5, 2.0, arr["TnQNNN"].shape))
    arr["Cs_opt2"][...] = T(rng.uniform(0.01, 0.05, arr["Cs_opt2"].shape))
    sst, sarr = o.scalar_sgs_state()
    for n in ("G_LM", "G_MM", "G_QN", "G_NN"):
        sarr[n][...] = T(rng.uniform(0.1, 1.0, sarr[n].shape))
    o.set_time(5, 600)
    o.set_scalar_lag_init(1, 1)
    SS = [T(S[3]), T(S[1]), T(S[4]), T(S[0]), T(S[2]), T(S[5])]
    o.lagrange_scalar_sdep(SS, T(v), T(u), T(w), T(th), st, sst)
    for k in ("G_MM", "G_NN"):
        assert relerr(T(sarr[k])[..., :o.nx], s0[k][..., :o.nx]) < 1e-10, k
    for k in ("G_LM", "Cs2_Pr"):
        a, b = T(sarr[k])[..., :o.nx], s0[k][..., :o.nx]
        assert (np.abs(a - b) > 1e-7 * np.abs(b)).mean() < 0.02, k


def test_lagrange_scalar_relaxation_time_scale(O):
    """theta uniform and u = const: L = Q = 0 so G_LM relaxes by (1 - eps),
    eps = (dt cs/T)/(1 + dt cs/T), T = 1.5 Delta TnLMMM^(-1/8) with the
    momentum model's TnLMMM (lagrange_scalar_Sdep.f90:318-331)."""
    o = mk(O, nx=8, ny=8, nz=4)
    u = np.full((o.nz, o.ny, o.ld), 2.0)
    th = np.full((o.nz, o.ny, o.ld), 0.7)
    z = o.field()
    S = [o.field() for _ in range(6)]
    st, arr = o.sgs_state()
    rng = np.random.default_rng(1)
    arr["TnLMMM"][...] = rng.uniform(0.5, 2.0, arr["TnLMMM"].shape)
    arr["TnQNNN"][...] = rng.uniform(0.5, 2.0, arr["TnLMMM"].shape)
    arr["Cs_opt2"][...] = 0.02
    sst, sarr = o.scalar_sgs_state()
    GLM = rng.uniform(0.1, 1.0, sarr["G_LM"].shape)
    GMM = rng.uniform(0.1, 1.0, sarr["G_MM"].shape)
    sarr["G_LM"][...] = GLM
    sarr["G_MM"][...] = GMM
    o.set_time(5, 600)
    o.set_scalar_lag_init(1, 1)
    o.lagrange_scalar_sdep(S, u, z, z, th, st, sst)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    e24 = float(np.float32(1e-24))
    delta = math.sqrt(o.dx * o.dy)
    T = np.maximum(e24, f32(1.5 * delta * arr["TnLMMM"] ** (-1.0 / 8)))
    eps = (1e-3 / T) / (1 + 1e-3 / T)
    sl = (slice(None), slice(None), slice(0, o.nx))
    # S = 0 but the surface gradient makes M != 0 at levels 1, 2: check the
    # levels where L = M = 0 exactly
    for k in range(2, o.nz - 1):
        exLM = np.maximum(e24, f32((1 - eps[k]) * GLM[k]))
        assert np.allclose(sarr["G_LM"][k][:, :o.nx], exLM[:, :o.nx], rtol=1e-15)
        assert np.allclose(sarr["G_MM"][k][:, :o.nx], ((1 - eps[k]) * GMM[k])[:, :o.nx], rtol=1e-15)
        c2 = np.maximum(e24, f32(exLM / ((1 - eps[k]) * GMM[k])))
        assert np.allclose(sarr["Pr_t"][k][:, :o.nx] * 0 + 0.02 / sarr["Cs2_Pr"][k][:, :o.nx],
                           sarr["Pr_t"][k][:, :o.nx], rtol=1e-15)
        assert np.all(sarr["Cs2_Pr"][k][:, :o.nx] <= c2[:, :o.nx] * 8.0000001)
    assert np.all(sarr["Cs2_Pr"][:, :, o.nx:] == e24)
    assert np.all(arr["Beta"][-1] == 1.0)


# ------------------------------------------------------------ whole step
def test_convective_step_translation_invariance(O):
    """Rolling the initial condition and the surface-flux pattern by whole
    cells rolls the state after 6 steps (incl. both dynamic updates)."""
    nx, ny, nz = 16, 16, 9
    res = []
    for shift in ((0, 0), (4, 8)):
        wt = np.roll(checker(nx, ny), shift, axis=(0, 1))
        o = mk(O, nx, ny, nz, wt=wt)
        s = O.OracleSim(o, seed=7, scalar_ic={})
        if shift != (0, 0):
            for name in ("u", "v", "w", "theta"):
                a = s[name]
                a[..., :nx] = np.roll(a[..., :nx], shift, axis=(1, 2))
        s.step(6)
        res.append(s)
    for name in ("u", "w", "theta", "qmix", "p", "G_LM", "G_MM", "Pr_t", "beta_scal"):
        a = res[0][name][:-1, :, :nx]
        b = np.roll(res[1][name][:-1, :, :nx], (-4, -8), axis=(1, 2))
        if name in ("G_LM",):
            bad = np.abs(a - b) > 1e-10 * np.max(np.abs(a)) + 2e-7 * np.abs(a)
            assert bad.mean() < 1e-2, name
        else:
            assert relerr(b, a) < 1e-9, name


def test_convective_step_heats_and_conserves_q(O):
    """With heating only at the surface, no humidity flux and no lid
    gradients, the domain-mean theta rises by the surface flux integrated
    over time, and q stays uniform (no sources, no gradients)."""
    nx, ny, nz = 16, 16, 9
    o = mk(O, nx, ny, nz, wt=np.full((ny, nx), 0.1), dTdz_top=0.0, dqdz_top=0.0)
    s = O.OracleSim(o, seed=1, scalar_ic=dict(amp_K=0.0, inv_strength=0.0))
    th0 = s["theta"][:-1, :, :nx].mean()
    s.step(5)
    th1 = s["theta"][:-1, :, :nx].mean()
    # column heat content gain = flux * time / height: nondim flux wt/(T_scale u*), dt per step
    gain = 5 * 2e-4 * (0.1 / 400.0) / (o.dz * (nz - 1))
    assert abs((th1 - th0) - gain) < 1e-3 * gain
    q = s["qmix"][:-1, :, :nx]
    assert np.allclose(q, 1.0 / 100.0, rtol=1e-12)
