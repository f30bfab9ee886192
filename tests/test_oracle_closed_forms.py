"""Closed-form pins of the oracle functions whose formulas the invariance
tests cannot see: the test-filter factors of M_ij / N_ij and the orientation
of Beta in lagrange_Sdep (lagrange_Sdep.f90:268-421) and in the scalar model
(lagrange_scalar_Sdep.f90:247-395), the time-step assembly (AB2 weights,
forcing on u only, the Euler first step; main.f90:357-378) and the boundary
rows of the pressure system with a non-zero divtz
(press_stag_array.f90:142-165,238-247).

Every expected value is computed here from the inputs' closed forms (Fourier
modes whose filtered values and products are known exactly, a plane-uniform
laminar column whose eddy viscosity and wall stress are textbook, the
discrete exponential solution of a constant-coefficient three-term
recurrence); the oracle's own outputs are never fed back into an expected
value.  tests/test_oracle_mutations.py rebuilds the oracle with each mistake
these pins target planted and checks that the pin then fails.
"""
import math

import numpy as np
import pytest

from conftest import relerr


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def f32(a):
    return np.asarray(a, dtype=np.float64).astype(np.float32).astype(np.float64)


E24 = float(np.float32(1e-24))
CS2_CONST = float(np.float32(0.16) ** 2)           # sgs_stag.f90:169, real(0.16_4**2) in single precision
VONK = 0.41                                        # param.f90:98


def grid_xy(o):
    x = np.arange(o.nx) * o.dx
    y = np.arange(o.ny) * o.dy
    return np.meshgrid(x, y)  # (ny, nx)


def put(o, planes):
    f = o.field(len(planes))
    for k, a in enumerate(planes):
        f[k, :, :o.nx] = a
    return f


def _relax(X, F0, Tn, delta, lagran_dt):
    """One Lagrangian relaxation with the Meneveau-Lund-Cabot time scale
    T = 1.5 Delta Tn^(-1/8): F = eps X + (1 - eps) F0 (FP64 value, before any
    rounding of F)."""
    T = np.maximum(E24, f32(1.5 * delta * Tn ** (-1.0 / 8)))
    eps = (lagran_dt / T) / (1.0 + lagran_dt / T)
    return eps * X + (1.0 - eps) * F0, eps


def _close(a, b, rtol, what):
    err = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))
    assert err <= rtol, (what, err, rtol)


# one single-precision ulp is at most 2^-23 of a value: "rounded" outputs are
# compared at 2.5e-7 (two ulps), quotients of rounded values at 5e-7
R1, R2 = 2.5e-7, 5e-7


def test_lagrange_constant_norm_strain_closed_form(O):
    """S11 = a cos x, S12 = a sin x, S22 = -S11 (trace-free, |S| = 2a
    everywhere) passes both test filters unchanged, so
      M_ij = 2 Delta^2 (2a S_ij - 4 (2a) S_ij)  = -12 a Delta^2 S_ij,
      N_ij = 2 Delta^2 (2a S_ij - 16 (2a) S_ij) = -60 a Delta^2 S_ij,
      MM = 288 a^4 Delta^4,  NN = 7200 a^4 Delta^4.
    u = A sin(5x), v = B sin(5y) (5 inside the 2-Delta cutoff 8, 10 outside;
    both outside the 4-Delta cutoff 4): L11 = A^2 cos(10x)/2, L22 =
    B^2 cos(10y)/2, L12 = 0 (the (5,5) product passes the 2-Delta filter),
    Q11 = A^2/2, Q22 = B^2/2, Q12 = 0; hence
      LM = -6 a^2 Delta^2 cos x (A^2 cos 10x - B^2 cos 10y),
      QN = 30 a^2 Delta^2 cos x (B^2 - A^2).
    With inilag at jt = cs_count, F_MM = MM and F_LM = 0.03 MM before the
    relaxation, so TnLMMM = real(0.03 MM^2) and every output follows."""
    o = O.Oracle(32, 32, 5, dt=0.1)
    X, Y = grid_xy(o)
    a, A, B, m = 1.0, 0.16, 0.12, 5
    delta = math.sqrt(o.dx * o.dy)
    u = put(o, [A * np.sin(m * X)] * o.nz)
    v = put(o, [B * np.sin(m * Y)] * o.nz)
    w = o.field()
    S = [o.field() for _ in range(6)]
    for k in range(o.nz):
        S[0][k, :, :o.nx] = a * np.cos(X)
        S[1][k, :, :o.nx] = a * np.sin(X)
        S[3][k, :, :o.nx] = -a * np.cos(X)
    st, arr = o.sgs_state()
    o.set_time(5, 600, flm_init=False, fqn_init=False)
    o.lagrange_sdep(S, u, v, w, st)

    d2 = delta * delta
    MM = 288.0 * a ** 4 * d2 * d2 * np.ones_like(X)
    NN = 7200.0 * a ** 4 * d2 * d2 * np.ones_like(X)
    LM = -6.0 * a * a * d2 * np.cos(X) * (A * A * np.cos(2 * m * X) - B * B * np.cos(2 * m * Y))
    QN = 30.0 * a * a * d2 * np.cos(X) * (B * B - A * A)
    lagran_dt = 0.1 * 5
    TnL = np.maximum(f32(0.03 * MM * MM), E24)
    TnQ = np.maximum(f32(0.03 * NN * NN), E24)
    FLM, _ = _relax(LM, 0.03 * MM, TnL, delta, lagran_dt)
    FMM, _ = _relax(MM, MM, TnL, delta, lagran_dt)
    FQN, _ = _relax(QN, 0.03 * NN, TnQ, delta, lagran_dt)
    FNN, _ = _relax(NN, NN, TnQ, delta, lagran_dt)
    FLM, FQN = np.maximum(E24, f32(FLM)), np.maximum(E24, f32(FQN))
    assert FLM.min() > 1e-3 and FQN.min() > 1e-3     # the relaxation stays clear of the clip
    cs2d = np.maximum(E24, f32(FLM / FMM))
    cs4d = np.maximum(E24, f32(FQN / FNN))
    beta = cs4d / cs2d                               # exponent ln2/(ln4 - ln2) = 1
    assert np.mean(np.abs(beta - 1.0) > 0.05) > 0.5  # beta far from 1: an inverted ratio cannot pass
    for k in range(o.nz):
        nx = o.nx
        _close(arr["TnLMMM"][k, :, :nx], TnL, R1, f"TnLMMM k={k}")
        _close(arr["TnQNNN"][k, :, :nx], TnQ, R1, f"TnQNNN k={k}")
        _close(arr["F_MM"][k, :, :nx], FMM, 1e-12, f"F_MM k={k}")
        _close(arr["F_NN"][k, :, :nx], FNN, 1e-12, f"F_NN k={k}")
        _close(arr["F_LM"][k, :, :nx], FLM, R1, f"F_LM k={k}")
        _close(arr["F_QN"][k, :, :nx], FQN, R1, f"F_QN k={k}")
        bk = np.ones_like(beta) if k == o.nz - 1 else beta   # Beta = 1 at the lid (:402-405)
        _close(arr["Beta"][k, :, :nx], bk, R2, f"Beta k={k}")
        cso = np.maximum(E24, f32(cs2d / np.maximum(f32(bk), f32(0.125))))
        _close(arr["Cs_opt2"][k, :, :nx], cso, 2 * R2, f"Cs_opt2 k={k}")
    # plane means of lagrange_Sdep.f90:425-426
    cs_last = np.maximum(E24, f32(cs2d / np.maximum(f32(beta), f32(0.125))))
    _close(arr["Beta_avg"][1], cs2d.sum() / cs_last.sum(), 1e-6, "Beta_avg")
    _close(arr["Betaclip_avg"][1], cs4d.sum() / cs2d.sum(), 1e-6, "Betaclip_avg")


def test_lagrange_scalar_constant_norm_strain_closed_form(O):
    """The scalar model with the same constant-norm strain (|S| = 2a, passes
    both filters) and theta = T0 + Gamma z + b sin x (the x mode passes both
    filters; surface flux 0, so dT/dz at the first node is 0):
      M_i = Delta^2 (2a dT_i - 4 (2a) dT_i)  = -6 a Delta^2 dT_i,
      N_i = Delta^2 (2a dT_i - 16 (2a) dT_i) = -30 a Delta^2 dT_i,
    with dT = (b cos x, 0, Gamma_k), Gamma_k = Gamma (Gamma/2 on the first
    node, the mean of the surface gradient 0 and Gamma);
      MM = 36 a^2 Delta^4 |dT|^2, NN = 900 a^2 Delta^4 |dT|^2.
    u = A sin(7x): u theta holds modes 7, 6 and 8, of which the 2-Delta
    filter drops 8, so L1 = (A b / 2) cos 8x, and the 4-Delta filter drops
    all of them and u, so Q = 0:
      LM = -3 A a b^2 Delta^2 cos 8x cos x,  QN = 0.
    The time scales are the momentum model's Tn (inputs), lagrange_scalar_Sdep.f90:321-352."""
    o = O.Oracle(32, 32, 6, dt=0.1)
    o.set_scalars(O.scalar_params(dTdz_top=0.0), np.zeros((o.ny, o.nx)), None)
    X, Y = grid_xy(o)
    a, A, b, Gam, T0 = 1.0, 0.01, 0.5, 0.5, 0.7
    delta = math.sqrt(o.dx * o.dy)
    u = put(o, [A * np.sin(7 * X)] * o.nz)
    v, w = o.field(), o.field()
    th = put(o, [T0 + Gam * (k + 0.5) * o.dz + b * np.sin(X) for k in range(o.nz)])
    S = [o.field() for _ in range(6)]
    for k in range(o.nz):
        S[0][k, :, :o.nx] = a * np.cos(X)
        S[1][k, :, :o.nx] = a * np.sin(X)
        S[3][k, :, :o.nx] = -a * np.cos(X)
    st, arr = o.sgs_state()
    TnL, TnQ, cs0 = 0.004, 2.5, 0.02
    arr["TnLMMM"][...] = TnL
    arr["TnQNNN"][...] = TnQ
    arr["Cs_opt2"][...] = cs0
    sst, sarr = o.scalar_sgs_state()
    o.set_time(5, 600)
    o.set_scalar_lag_init(0, 0)
    o.lagrange_scalar_sdep(S, u, v, w, th, st, sst)

    d2 = delta * delta
    lagran_dt = 0.1 * 5
    for k in range(o.nz - 1):            # the lid level takes dTdz_top
        g = Gam / 2 if k == 0 else Gam
        dT2 = (b * np.cos(X)) ** 2 + g * g
        MM = 36.0 * a * a * d2 * d2 * dT2
        NN = 900.0 * a * a * d2 * d2 * dT2
        LM = -3.0 * A * a * b * b * d2 * np.cos(8 * X) * np.cos(X)
        GLM, _ = _relax(LM, 0.03 * MM, TnL, delta, lagran_dt)
        GMM, _ = _relax(MM, MM, TnL, delta, lagran_dt)
        GQN, _ = _relax(0.0, 0.03 * NN, TnQ, delta, lagran_dt)
        GNN, _ = _relax(NN, NN, TnQ, delta, lagran_dt)
        GLM, GQN = np.maximum(E24, f32(GLM)), np.maximum(E24, f32(GQN))
        assert GLM.min() > 0.0
        c2 = np.maximum(E24, f32(GLM / GMM))
        c4 = np.maximum(E24, f32(GQN / GNN))
        beta = c4 / c2
        assert np.mean(np.abs(beta - 1.0) > 0.05) > 0.5   # an inverted ratio cannot pass
        cpr = np.maximum(E24, f32(c2 / np.maximum(f32(beta), f32(0.125))))
        nx = o.nx
        _close(sarr["G_MM"][k, :, :nx], GMM, 1e-12, f"G_MM k={k}")
        _close(sarr["G_NN"][k, :, :nx], GNN, 1e-12, f"G_NN k={k}")
        _close(sarr["G_LM"][k, :, :nx], GLM, R1, f"G_LM k={k}")
        _close(sarr["G_QN"][k, :, :nx], GQN, R1, f"G_QN k={k}")
        _close(arr["Beta"][k, :, :nx], beta, R2, f"Beta k={k}")
        _close(sarr["Cs2_Pr"][k, :, :nx], cpr, 2 * R2, f"Cs2_Pr k={k}")
        _close(sarr["Pr_t"][k, :, :nx], cs0 / cpr, 2 * R2, f"Pr_t k={k}")


def _column_rhs(u, v, dz, delta, z0, nzl):
    """Right-hand sides of a plane-uniform laminar column (w = 0, no
    horizontal variation): only the stress divergence d(tau_xz)/dz remains,
    with the Smagorinsky eddy viscosity Nu_t = Cs^2 Delta^2 |du/dz|
    (|S| = sqrt(2 S_ij S_ij) = sqrt(u_z^2 + v_z^2) on the w-nodes), the
    neutral log-law wall stress tau_xz(1) = -u*^2 u/|u|,
    u* = |u(1)| kappa / ln(dz / 2 z0), and tau = 0 at the lid.
    u, v: the nzl = nz - 1 uvp-node values.  Returns RHSx, RHSy (nzl)."""
    U = math.hypot(u[0], v[0])
    us = U * VONK / math.log(0.5 * dz / z0)
    txz, tyz = np.zeros(nzl + 1), np.zeros(nzl + 1)
    txz[0], tyz[0] = -us * us * u[0] / U, -us * us * v[0] / U
    for k in range(1, nzl):              # w-nodes 2 .. nz-1
        uz, vz = (u[k] - u[k - 1]) / dz, (v[k] - v[k - 1]) / dz
        nu = CS2_CONST * delta * delta * math.sqrt(uz * uz + vz * vz)
        txz[k], tyz[k] = -nu * uz, -nu * vz
    return -(txz[1:] - txz[:-1]) / dz, -(tyz[1:] - tyz[:-1]) / dz


@pytest.mark.parametrize("initu", [1, 0])
def test_step_laminar_column_closed_form(O, initu):
    """A plane-uniform log-law column u(z), with v(z) = 0.3 u(z) and w = 0:
    advection vanishes (w = 0 and no horizontal variation), the pressure
    only cancels the vertical part, so each step is
      u <- u + dt (1.5 R - 0.5 R_f + force),  v <- v + dt (1.5 R_v - 0.5 R_v,f)
    with R the stress divergence of the column (main.f90:357-378).  With
    initu the history R_f starts at 0 (AB2 from the first step); without it
    the first step is Euler (R_f = R, :363-370).  Two steps in closed form."""
    nx = ny = 8
    nz, dt, z0 = 17, 2e-4, 1e-4
    o = O.Oracle(nx, ny, nz, dt=dt, z0=z0, initu=initu)
    so = O.OracleSim(o, amp=0.0)
    nzl = nz - 1
    z = (np.arange(nzl) + 0.5) * o.dz
    u0 = np.log(z / z0) / 0.4
    v0 = 0.3 * u0
    so["v"][:nzl, :, :nx] = v0[:, None, None]
    so["v"][nzl, :, :nx] = v0[-1]
    assert np.allclose(so["u"][:nzl, 0, 0], u0, rtol=1e-14)
    delta = math.sqrt(o.dx * o.dy)
    u, v = u0.copy(), v0.copy()
    Rf = Rfv = None
    for step in (1, 2):
        R, Rv = _column_rhs(u, v, o.dz, delta, z0, nzl)
        if step == 1:
            Rf, Rfv = (R, Rv) if not initu else (np.zeros(nzl), np.zeros(nzl))
        u = u + dt * (1.5 * R - 0.5 * Rf + 1.0)
        v = v + dt * (1.5 * Rv - 0.5 * Rfv)
        Rf, Rfv = R, Rv
        so.step(1)
        gu, gv, gw = so["u"][:nzl, :, :nx], so["v"][:nzl, :, :nx], so["w"][:nzl, :, :nx]
        assert np.max(np.abs(gu - gu[:, :1, :1])) < 1e-12       # stays plane-uniform
        assert relerr(gu[:, 0, 0], u) < 1e-12, ("u", step)
        assert relerr(gv[:, 0, 0], v) < 1e-12, ("v", step)
        assert np.max(np.abs(gw)) < 1e-12
        assert np.all(so["u"][nzl, :, :nx] == so["u"][nzl - 1, :, :nx])   # lid u(nz) = u(nz-1)


def _recurrence_solution(K2, dz, nz, Db, Dt):
    """Exact solution of the pressure rows for one Fourier mode with H = 0:
    P_1 - P_0 = -dz Db (wall row), P_{j-1} - (2 + K^2 dz^2) P_j + P_{j+1} = 0
    (interior), P_nz - P_{nz-1} = -dz Dt (lid row).  The interior recurrence
    has the solutions r^j and r^-j with r + 1/r = 2 + K^2 dz^2."""
    lam = 2.0 + K2 * dz * dz
    r = 0.5 * (lam + math.sqrt(lam * lam - 4.0))
    j = np.arange(nz + 1, dtype=np.float64)
    # P_j = alpha r^(j - nz) + beta r^-j, scaled to avoid overflow
    A = np.array([[r ** (1 - nz) - r ** (-nz), r ** -1 - 1.0],
                  [1.0 - r ** -1, r ** (-nz) - r ** (1 - nz)]])
    rhs = np.array([-dz * Db, -dz * Dt])
    al, be = np.linalg.solve(A, rhs)
    return al * r ** (j - nz) + be * r ** (-j)


def test_press_stag_boundary_rows_closed_form(O):
    """u = v = w = 0 and divtz non-zero on the wall and lid planes only: the
    pressure is the discrete harmonic function whose wall and lid rows carry
    -dz divtz (press_stag_array.f90:142-165).  A mode cos(2x + y) with
    amplitudes Db (wall) and Dt (lid) gives p = P_j cos(2x + y) with P the
    exact recurrence solution; a plane-uniform divtz = Dc on the wall gives
    the zero mode p = 0 at level 0 and -dz Dc above (:238-247)."""
    nx = ny = 16
    nz = 9
    o = O.Oracle(nx, ny, nz)
    X, Y = grid_xy(o)
    Db, Dt, Dc = 0.7, -0.4, 0.25
    z = o.field()
    divtz = o.field()
    divtz[0, :, :nx] = Db * np.cos(2 * X + Y) + Dc
    divtz[nz - 1, :, :nx] = Dt * np.cos(2 * X + Y)
    p, dpdx, dpdy = o.field(nz + 1), o.field(), o.field()
    o.press_stag_array(z, z, z, divtz, p, dpdx, dpdy)
    P = _recurrence_solution(5.0, o.dz, nz, Db, Dt)
    for j in range(nz):
        zero_mode = 0.0 if j == 0 else -o.dz * Dc
        ex = P[j] * np.cos(2 * X + Y) + zero_mode
        assert np.max(np.abs(p[j, :, :nx] - ex)) < 1e-13, j
        if 1 <= j <= nz - 1:
            assert np.max(np.abs(dpdx[j - 1, :, :nx] + 2 * P[j] * np.sin(2 * X + Y))) < 1e-12, j
            assert np.max(np.abs(dpdy[j - 1, :, :nx] + P[j] * np.sin(2 * X + Y))) < 1e-12, j
    assert abs(P[1] - P[0] + o.dz * Db) < 1e-14 and abs(P[nz] - P[nz - 1] + o.dz * Dt) < 1e-14
