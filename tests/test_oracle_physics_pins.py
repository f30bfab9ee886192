"""Closed-form pins of the oracle's physics switches (oracle/les_oracle.c):
setsponge (topbc.f90:14-74), the Coriolis and sponge terms of the step
(main.f90:319-348, scalars_module.f90:143-148), Mason's mixing length of
model 1 (sgs_stag.f90:73-104), std_dynamic (std_dynamic.f90) against an
independent numpy restatement with numpy's FFT, and lagrange_Ssim /
interpolag_Ssim (model 4) through limits and exact invariants."""
import math

import numpy as np
import pytest

CORIOL = float(np.float32(1.3962634e-04))   # the default-real literal of param.f90:220


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


# ---------------------------------------------------------------- sponge
def test_sponge_cosine_profile(O):
    """damping_method 2 on nz = 13, L_z = 3000 m: uv-node heights (k - 1/2)
    250 m; zero below z_d = 2250 m, 1.95 (1 - cos(pi (z - z_d) / 750 m)) above:
    levels 10, 11, 12 sit at pi/6, pi/2, 5 pi/6; sponge(13) = sponge(12)."""
    o = O.Oracle(16, 16, 13, Lz=3.0)
    o.set_forcing(ubc=1, damping_method=2, u_star=0.45, z_i=1000.0)
    sp = o.sponge()
    r3 = math.sqrt(3.0) / 2
    ex = np.zeros(13)
    ex[9], ex[10], ex[11] = 1.95 * (1 - r3), 1.95, 1.95 * (1 + r3)
    ex[12] = ex[11]
    assert np.allclose(sp, ex, rtol=1e-14, atol=1e-15), (sp, ex)


@pytest.mark.parametrize("nz,kend,factor", [(32, 25, 1.0), (17, 13, 9.0 / 5.25)])
def test_sponge_relaxation_profile(O, nz, kend, factor):
    """damping_method 1: sponge(nz) = z_i / (50 u*), divided by 5^factor per
    level down to level int(0.75 nz + 1), factor = 9 / (nz - 0.75 nz + 1);
    zero below."""
    o = O.Oracle(16, 16, nz, Lz=1.0)
    o.set_forcing(ubc=1, damping_method=1, u_star=0.4, z_i=800.0)
    sp = o.sponge()
    top = 800.0 / (50 * 0.4)
    ex = np.zeros(nz)
    for k in range(kend, nz + 1):
        ex[k - 1] = top / 5.0 ** ((nz - k) * factor)
    assert np.allclose(sp, ex, rtol=1e-13, atol=0), (sp, ex)


# ---------------------------------------------------------------- forcing terms
def _forcing_delta(O, forcing, conv=False):
    """One step of two identical oracle states, with and without the forcing;
    returns the pre-step fields and the difference of the pre-pressure RHS
    (RHS + grad p, the step subtracts the pressure gradient last) and of the
    scalar RHS."""
    nx, ny, nz = 16, 16, 17
    sims, pre = [], None
    for f in (None, forcing):
        o = O.Oracle(nx, ny, nz, Lz=3.0, initu=1)
        if conv:
            p = O.scalar_params(model=5)
            o.set_scalars(p, np.full((ny, nx), 2e-4), None)
        s = O.OracleSim(o, seed=11, scalar_ic={} if conv else None)
        s.step(2)
        if f is not None:        # the same history, the forcing in the last step only
            o.set_forcing(**f)
        if pre is None:
            # the step reads u, v, w (and theta) after filt_da has removed their
            # Nyquist modes in place (main.f90:228-233, scalars_module.f90:87)
            pre = {n: s[n].copy() for n in ("u", "v", "w") + (("theta",) if conv else ())}
            for a in pre.values():
                o.filt_da(a, o.field(), o.field())
        s.step(1)
        sims.append((o, s))
    (o0, s0), (o1, s1) = sims
    d = {}
    for n, g in (("RHSx", "dpdx"), ("RHSy", "dpdy"), ("RHSz", "dpdz")):
        d[n] = (s1[n] + s1[g]) - (s0[n] + s0[g])
    if conv:
        d["RHS_T"] = s1["RHS_T"] - s0["RHS_T"]
    return o1, pre, d


def _mean(a, nx):
    return a[..., :nx].mean(axis=(-1, -2))


def test_coriolis_term(O):
    """RHSx gains coriol (v - vg), RHSy -coriol (u - ug), RHSz nothing, with
    coriol = 1.3962634e-4 z_i / u*, ug = ug0/u*, vg = vg0/u*, on whole rows."""
    f = dict(coriolis=1, ug0=8.0, vg0=-2.0, u_star=0.45, z_i=1000.0)
    o, pre, d = _forcing_delta(O, f)
    cor = CORIOL * 1000.0 / 0.45
    nl = o.nz - 1
    ex_x = cor * pre["v"][:nl] - cor * (-2.0 / 0.45)
    ex_y = -cor * pre["u"][:nl] + cor * (8.0 / 0.45)
    scale = max(np.abs(ex_x).max(), np.abs(ex_y).max())
    assert np.abs(d["RHSx"][:nl] - ex_x).max() < 1e-11 * scale
    assert np.abs(d["RHSy"][:nl] - ex_y).max() < 1e-11 * scale
    assert np.abs(d["RHSz"][:nl, :, :o.nx]).max() < 1e-11 * scale


@pytest.mark.parametrize("method", [1, 2])
def test_sponge_terms(O, method):
    """damping_method 2: RHSx -= (sponge(jz) + sponge(jz+1))/2 (u - <u>),
    RHSy likewise, RHSz -= sponge(jz)/2 (w - <w>); method 1: RHSz -=
    sponge(jz) w; on the nx*ny points of levels 1..nz-1."""
    f = dict(ubc=1, damping_method=method, u_star=0.45, z_i=1000.0)
    o, pre, d = _forcing_delta(O, f)
    sp = o.sponge()
    nl, nx = o.nz - 1, o.nx
    assert sp[:nl].max() > 0.1
    cxy = 0.5 * (sp[:nl] + sp[1:nl + 1])
    if method == 2:
        ex = {"RHSx": -cxy[:, None, None] * (pre["u"][:nl, :, :nx] - _mean(pre["u"][:nl], nx)[:, None, None]),
              "RHSy": -cxy[:, None, None] * (pre["v"][:nl, :, :nx] - _mean(pre["v"][:nl], nx)[:, None, None]),
              "RHSz": -0.5 * sp[:nl, None, None] * (pre["w"][:nl, :, :nx] - _mean(pre["w"][:nl], nx)[:, None, None])}
    else:
        ex = {"RHSx": np.zeros((nl, o.ny, nx)), "RHSy": np.zeros((nl, o.ny, nx)),
              "RHSz": -sp[:nl, None, None] * pre["w"][:nl, :, :nx]}
    scale = max(np.abs(e).max() for e in ex.values())
    assert scale > 0
    for n in ex:
        err = np.abs(d[n][:nl, :, :nx] - ex[n]).max()
        assert err < 1e-10 * scale, (n, err, scale)


def test_scalar_sponge_term(O):
    """damping_method 2 with scalars: RHS_T -= (sponge(jz) + sponge(jz+1))/2
    (theta - <theta>) before the scalar step."""
    f = dict(ubc=1, damping_method=2, u_star=1.0, z_i=1000.0)
    o, pre, d = _forcing_delta(O, f, conv=True)
    sp = o.sponge()
    nl, nx = o.nz - 1, o.nx
    cxy = 0.5 * (sp[:nl] + sp[1:nl + 1])
    th = pre["theta"][:nl, :, :nx]
    ex = -cxy[:, None, None] * (th - _mean(pre["theta"][:nl], nx)[:, None, None])
    assert np.abs(ex).max() > 0
    assert np.abs(d["RHS_T"][:nl, :, :nx] - ex).max() < 1e-9 * np.abs(ex).max()


# ---------------------------------------------------------------- model 1
def _mason_l(O, Lx, nz=9, Lz=1.0):
    o = O.Oracle(16, 16, nz, Lx=Lx, Ly=Lx, Lz=Lz)
    rng = np.random.default_rng(2)
    grads = [o.field() for _ in range(9)]
    for a in grads:
        a[..., :o.nx] = rng.standard_normal(a[..., :o.nx].shape)
    u, v, w = o.field(), o.field(), o.field()
    taus = [o.field() for _ in range(6)]
    st, arr = o.sgs_state()
    o.set_time(3, 3)
    o.sgs_stag_model(grads, u, v, w, taus, st, 1)
    s11, s12, s13, s22, s23, s33 = o.calc_sij(grads)
    Sm = np.sqrt(2 * (s11 ** 2 + s22 ** 2 + s33 ** 2 + 2 * (s12 ** 2 + s13 ** 2 + s23 ** 2)))
    l2 = arr["Nu_t"][..., :o.nx] / (Sm[..., :o.nx] * 0.16 ** 2)
    assert np.all(arr["Cs_opt2"] == 0.16 * 0.16)
    return o, np.sqrt(l2.mean(axis=(1, 2)))


def test_mason_length_wall_limit(O):
    """Delta >> kappa z / Co (a wide horizontal grid): l(jz) = kappa zz / Co
    with zz = dz/2 on the first uv-node and (jz - 1) dz on the w-nodes above
    (relative correction (kappa zz / (Co Delta))^2 / 2 < 3e-8)."""
    o, l = _mason_l(O, Lx=2e5)
    zz = np.array([0.5 * o.dz] + [k * o.dz for k in range(1, o.nz)])
    ex = 0.41 * zz / 0.16
    assert np.allclose(l, ex, rtol=1e-7), (l, ex)


def test_mason_length_outer_limit(O):
    """Delta << kappa z / Co (a fine horizontal grid, a tall domain): l = Delta."""
    o, l = _mason_l(O, Lx=1e-3, Lz=1e3)
    delta = math.sqrt(o.dx * o.dy)
    assert np.allclose(l[1:], delta, rtol=1e-6), (l, delta)


# ---------------------------------------------------------------- model 2
def _np_filter(f, o, alpha):
    """sharp spectral cutoff at pi/(alpha Delta), Nyquist modes removed
    (test_filtermodule.f90:32-63), with numpy's FFT"""
    nx, ny = o.nx, o.ny
    kx = 2 * np.pi / (o.dx * nx) * np.arange(nx // 2 + 1)
    ky = 2 * np.pi / (o.dy * ny) * np.fft.fftfreq(ny, 1.0 / ny)
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    delta = alpha * math.sqrt(o.dx * o.dy)
    mask = (k2 < (np.pi / delta) ** 2).astype(float)
    mask[:, nx // 2] = 0.0
    mask[ny // 2, :] = 0.0
    return np.fft.irfft2(np.fft.rfft2(f) * mask, s=(ny, nx))


def _np_lm_mm(o, S, u, v, w, k, delta):
    nx = o.nx
    if k == 0:
        ub, vb, wb = u[0, :, :nx], v[0, :, :nx], 0.25 * w[1, :, :nx]
    else:
        ub = 0.5 * (u[k, :, :nx] + u[k - 1, :, :nx])
        vb = 0.5 * (v[k, :, :nx] + v[k - 1, :, :nx])
        wb = w[k, :, :nx]
    F = lambda a: _np_filter(a, o, 2.0)
    vel = (ub, vb, wb)
    pairs = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))
    fb = [F(a) for a in vel]
    L = [F(vel[a] * vel[b]) - fb[a] * fb[b] for a, b in pairs]
    s = [S[q][k, :, :nx] for q in range(6)]
    wts = np.array([1, 2, 2, 1, 2, 1], dtype=float)
    Sm = np.sqrt(2 * sum(wq * sq ** 2 for wq, sq in zip(wts, s)))
    sb = [F(a) for a in s]
    Sb = np.sqrt(2 * sum(wq * sq ** 2 for wq, sq in zip(wts, sb)))
    ssb = [F(Sm * a) for a in s]
    M = [2 * delta ** 2 * (ssb[q] - 4 * Sb * sb[q]) for q in range(6)]
    LM = sum(wts[q] * L[q] * M[q] for q in range(6))
    MM = sum(wts[q] * M[q] * M[q] for q in range(6))
    return LM, MM


def _rand_fields(o, seed):
    rng = np.random.default_rng(seed)
    fs = []
    for _ in range(9):
        a = o.field()
        # smooth enough that the 2-Delta filter keeps part of every field
        a[..., :o.nx] = np.cumsum(rng.standard_normal(a[..., :o.nx].shape), axis=-1) * 0.3
        a[..., :o.nx] -= a[..., :o.nx].mean(axis=-1, keepdims=True)
        fs.append(a)
    u, v, w = fs[:3]
    u[..., :o.nx] += 5.0
    return fs[3:], u, v, w


def test_std_dynamic_numpy(O):
    """Cs_opt2(jz) = max(0, sum L:M / sum M:M) per level, Delta = (dx dy
    dz)^(1/3), against the same sums formed with numpy's FFT filter."""
    o = O.Oracle(16, 16, 7)
    S, u, v, w = _rand_fields(o, 3)
    cs = o.std_dynamic(S, u, v, w)
    delta = (o.dx * o.dy * o.dz) ** (1.0 / 3.0)
    ex = []
    for k in range(o.nz - 1):
        LM, MM = _np_lm_mm(o, S, u, v, w, k, delta)
        ex.append(max(0.0, LM.sum() / MM.sum()))
    assert np.any(np.array(ex) > 0)
    assert np.allclose(cs[:o.nz - 1], ex, rtol=1e-10, atol=1e-14), (cs, ex)


def test_std_dynamic_invariances(O):
    """Galilean (u, v + constants) and scale (u_i, S_ij times alpha)
    invariance of the dynamic coefficient."""
    o = O.Oracle(16, 16, 7)
    S, u, v, w = _rand_fields(o, 4)
    cs = o.std_dynamic(S, u, v, w)
    u2, v2 = u.copy(), v.copy()
    u2[..., :o.nx] += 3.0
    v2[..., :o.nx] -= 1.5
    assert np.allclose(o.std_dynamic(S, u2, v2, w), cs, rtol=1e-9, atol=1e-15)
    a = 2.5
    assert np.allclose(o.std_dynamic([a * s for s in S], a * u, a * v, a * w), cs, rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- model 4
def test_interpolag_ssim_vertical_shift(O):
    """F linear in the level index, w_lag = cs_count W (W < 0 uniform): the
    back-trajectory displacement z2 = -W cs_count dt / dz on w-nodes, doubled at level 2
    only when negative, 0.25 z2 at level 1 doubled when positive, no clamp:
    F(k) + zp(k) dF/dk, the top level unchanged (its neighbour clamps to
    itself)."""
    o = O.Oracle(16, 16, 9, dt=0.05, cs_count=5)
    st, arr = o.sgs_state()
    nz = o.nz
    lev = np.arange(nz, dtype=float)[:, None, None]
    arr["F_LM"][...] = 2.0 + 0.5 * lev
    arr["F_MM"][...] = 1.0 - 0.1 * lev
    W = -0.3
    arr["w_lag"][...] = 5 * W
    z2 = -W * 5 * 0.05 / o.dz
    assert 0.5 < z2 < 1.0
    o.interpolag_ssim(st)
    zp = np.full(nz, z2)
    zp[0] = 2 * 0.25 * z2
    zp[nz - 1] = 0.0
    for n, slope in (("F_LM", 0.5), ("F_MM", -0.1)):
        ex = (2.0 if n == "F_LM" else 1.0) + slope * (np.arange(nz) + zp)
        got = arr[n][:, :, :o.nx].mean(axis=(1, 2))
        assert np.allclose(got, ex, rtol=1e-13), (n, got, ex)
        assert np.ptp(arr[n][:, :, :o.nx], axis=(1, 2)).max() < 1e-13


def test_lagrange_ssim_limits(O):
    """Initialisation step with u_lag = 0: F_MM = MM exactly (a convex
    combination of MM with itself); F_LM between LM and 0.025 MM with the
    weight eps = (dt cs/T)/(1 + dt cs/T), T = 1.5 Delta (F_LM F_MM)^(-1/8) of
    Meneveau, Lund & Cabot (1996); L, M against numpy; Beta = 1 and
    Beta_avg = nx*ny."""
    o = O.Oracle(16, 16, 7, dt=3e-3, cs_count=5)
    S, u, v, w = _rand_fields(o, 6)
    st, arr = o.sgs_state()
    o.set_time(5, 5, flm_init=0, fqn_init=0)
    o.lagrange_ssim(S, u, v, w, st)
    delta = math.sqrt(o.dx * o.dy)
    nx = o.nx
    assert np.all(arr["Beta"] == 1.0) and np.all(arr["Beta_avg"] == nx * o.ny)
    for k in (0, 2, 5):
        LM, MM = _np_lm_mm(o, S, u, v, w, k, delta)
        fmm, flm = arr["F_MM"][k, :, :nx], arr["F_LM"][k, :, :nx]
        assert np.allclose(fmm, MM, rtol=1e-10, atol=1e-12 * np.abs(MM).max())
        T = 1.5 * delta * np.maximum(0.025 * MM * MM, 1e-32) ** (-1.0 / 8)
        dum = 3e-3 * 5 / T
        eps = dum / (1 + dum)
        ex = np.maximum(1e-32, eps * LM + (1 - eps) * 0.025 * MM)
        assert np.allclose(flm, ex, rtol=1e-9, atol=1e-12 * np.abs(ex).max())
        assert np.allclose(arr["Cs_opt2"][k, :, :nx], np.maximum(1e-32, flm / fmm), rtol=1e-14)
        assert np.all(arr["Cs_opt2"][k, :, nx:] == 1e-32)


# ------------------------------------------------ model 6 (new_scaledep)
F24 = float(np.float32(1e-24))


def _f32(x):
    return float(np.float32(x))


def _clipped(c2, c4):
    """new_scaledep_dynamic.f90:213-220 / scalar_scaledep_dynamic.f90:233-240"""
    c2, c4 = max(F24, _f32(c2)), max(F24, _f32(c4))
    beta = c4 / c2
    return max(0.0, c2 / max(_f32(beta), _f32(0.125))), beta


def _strain(o, a):
    """S11 = a cos x, S12 = a sin x, S22 = -S11: trace-free, |S| = 2a at every
    point, and the single x mode passes both test filters unchanged."""
    x = np.arange(o.nx) * o.dx
    S = [o.field() for _ in range(6)]
    for k in range(o.nz):
        S[0][k, :, :o.nx] = a * np.cos(x)[None, :]
        S[1][k, :, :o.nx] = a * np.sin(x)[None, :]
        S[3][k, :, :o.nx] = -a * np.cos(x)[None, :]
    return S


def _xfield(o, modes):
    """sum of amp sin(m x) on every level"""
    x = np.arange(o.nx) * o.dx
    f = o.field()
    f[:, :, :o.nx] = sum(amp * np.sin(m * x) for m, amp in modes)[None, None, :]
    return f


def test_new_scaledep_closed_form(O):
    """Model 6 on 32^2 (2-Delta filter keeps kx < 8, 4-Delta kx < 4) with the
    constant-norm strain and u = A sin 10x + C sin 11x + D sin 5x + E sin 6x,
    v = w = 0.  The 10/11 pair lies beyond both filters, the 5/6 pair between
    them, so the only cos x parts are L11 = ... + AC cos x and Q11 = ... + DE
    cos x; with M_ij = 2 Delta^2 (2a - 4 2a) S_ij and N_ij = 2 Delta^2 (2a -
    16 2a) S_ij the plane sums are LM = -6 a^2 Delta^2 AC nxy, MM = 288 a^4
    Delta^4 nxy, QN = -30 a^2 Delta^2 DE nxy, NN = 7200 a^4 Delta^4 nxy:
    Cs2_2d = -AC / (48 a^2 Delta^2), Cs2_4d = -DE / (240 a^2 Delta^2), Beta =
    DE / (5 AC) = 0.5, and Cs_opt2 = Cs2_2d / Beta on every level."""
    o = O.Oracle(32, 32, 6)
    a, A, C, D, E = 1.0, 0.3, -0.2, 0.25, -0.6
    S = _strain(o, a)
    u = _xfield(o, [(10, A), (11, C), (5, D), (6, E)])
    v, w = o.field(), o.field()
    cs = o.new_scaledep_dynamic(S, u, v, w)
    d2 = o.dx * o.dy
    want, beta = _clipped(-A * C / (48 * a * a * d2), -D * E / (240 * a * a * d2))
    assert abs(beta - 0.5) < 1e-9
    print("Cs_opt2", cs, "closed form", want)
    assert np.allclose(cs, want, rtol=2.5e-7, atol=0)


def test_scalar_scaledep_closed_form(O):
    """scalar_scaledep_dynamic on the same strain, zero surface flux (so dT/dz
    = 0 on the first level), u = A sin 11x + D sin 5x, v = w = 0 and theta =
    T0 + Gamma z + b sin x + g sin 12x + h sin 6x.  Only the pair (11, 12)
    puts a cos x part into L1 = Ag/2 cos x (the (5, 6) pair's passes both
    filtered sides and cancels); Q1 = Dh/2 cos x (the 2-Delta bars' (5, 6)
    product through the 4-Delta filter).  dT/dx bar = b cos x + 6h cos 6x,
    dT/dx hat = b cos x, dT/dz = Gamma above the first level, so with const
    = Delta^2: M1 = -6 a Delta^2 (b cos x + 6h cos 6x), M3 = -6 a Gamma
    Delta^2, N1 = -30 a b Delta^2 cos x, N3 = -30 a Gamma Delta^2, and per
    level LM = -1.5 A g a b Delta^2 nxy, MM = 36 a^2 Delta^4 ((b^2 + 36 h^2)/2
    + Gamma^2) nxy, QN = -7.5 D h a b Delta^2 nxy, NN = 900 a^2 Delta^4 (b^2/2
    + Gamma^2) nxy (Gamma -> 0 on the first level); Cs2_Pr from the clipped
    ratios and Pr_t = Cs_opt2 / Cs2_Pr."""
    o = O.Oracle(32, 32, 6)
    o.set_scalars(O.scalar_params(), np.zeros((o.ny, o.nx)), None)
    a, A, D = 1.0, 0.3, 0.25
    b, g, h, G, T0 = 0.5, -0.2, -0.4, 0.3, 0.7
    S = _strain(o, a)
    u = _xfield(o, [(11, A), (5, D)])
    v, w = o.field(), o.field()
    th = _xfield(o, [(1, b), (12, g), (6, h)])
    for k in range(o.nz):
        th[k, :, :o.nx] += T0 + G * (k + 0.5) * o.dz
    cs = o.field()
    for k in range(o.nz):
        cs[k] = 0.02 + 0.001 * k
    pr = o.field()
    cpr = o.scalar_scaledep_dynamic(S, u, v, w, th, cs, np.zeros((o.ny, o.nx)), np.ones((o.ny, o.nx)), pr)
    d2 = o.dx * o.dy
    for k in range(o.nz - 1):
        Gk = 0.0 if k == 0 else G
        mm = 36 * a * a * d2 * d2 * ((b * b + 36 * h * h) / 2 + Gk * Gk)
        nn = 900 * a * a * d2 * d2 * (b * b / 2 + Gk * Gk)
        want, beta = _clipped(-1.5 * A * g * a * b * d2 / mm, -7.5 * D * h * a * b * d2 / nn)
        assert beta > 1.5                     # unclipped and far from 1: an inverted ratio cannot pass
        print(k, "Cs2_Pr", cpr[k], "closed form", want)
        assert abs(cpr[k] - want) <= 2.5e-7 * want
        assert np.allclose(pr[k, :, :], cs[k] / cpr[k], rtol=1e-15)
    assert np.all(pr[o.nz - 1] == 0.0)        # the lid level is not written
