"""Closed-form pins of the oracle's statistics (oracle/les_oracle.c):
spectrum (io.f90:1613-1646), the current statistics of avg_stats
(io.f90:1722-1755) and obukhov's plane averages (scalars_module.f90:545-561,
655-661), on fields made of Fourier modes whose plane means, variances,
covariances and spectral energies are known exactly."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def grid_xy(o):
    x = np.arange(o.nx) * o.dx
    y = np.arange(o.ny) * o.dy
    return np.meshgrid(x, y)


def test_spectrum_modes(O):
    """f = C + B sin(3y) + A cos(m x + 2y): every row's mean is C + B sin(3y)
    and its x mode m has |c_m|^2 = A^2/4, so spec(0) = (C^2 + B^2/2)/2 (the
    mean counted with 1/2), spec(m) = A^2/4 and nothing else."""
    o = O.Oracle(32, 24, 5)
    X, Y = grid_xy(o)
    f = o.field()
    Cs, As, Bs, ms = (1.5, -0.5, 0.0, 2.0), (0.3, 0.7, 1.1, 0.0), (0.2, 0.0, 0.4, 0.9), (3, 7, 15, 1)
    for k in range(4):
        f[k, :, :o.nx] = Cs[k] + Bs[k] * np.sin(3 * Y) + As[k] * np.cos(ms[k] * X + 2 * Y)
    spec = o.spectrum(f[:4])
    for k in range(4):
        ex = np.zeros(o.nx // 2)
        ex[0] = 0.5 * (Cs[k] ** 2 + 0.5 * Bs[k] ** 2)
        ex[ms[k]] += As[k] ** 2 / 4
        assert np.max(np.abs(spec[k] - ex)) < 1e-14, (k, spec[k], ex)


def test_avg_stats_modes(O):
    """u_k = U_k + A_k cos(mx), v_k = V_k + Av sin(mx + y), w_k = W_k +
    Aw_k cos(mx), txz_k = T_k + At cos(x) on nl = nz-1 levels.  The w-node
    fluctuations of io.f90:1729-1742 are u' = (A_k + A_(k-1))/2 cos(mx),
    v' = Av sin(mx + y), w' = Aw_k cos(mx) (zero on the first level), so
      <u'^2> = (A_k + A_(k-1))^2 / 8, <v'^2> = Av^2/2, <w'^2> = Aw_k^2/2,
      <u'w'> = (A_k + A_(k-1)) Aw_k / 4, stress = <u'w'> + T_k,
    and E11/z = [U_k^2/2 at k1 = 0, A_k^2/4 at k1 = m] / ((k + 1/2) dz).
    A level-dependent A_k pins the averaging with the level below."""
    o = O.Oracle(32, 32, 7)
    X, Y = grid_xy(o)
    nl, m = o.nz - 1, 4
    U = 1.0 + 0.5 * np.arange(nl)
    V = -0.2 + 0.1 * np.arange(nl)
    W = 0.05 * np.arange(nl)
    T = -0.3 + 0.04 * np.arange(nl)
    A = 0.2 + 0.1 * np.arange(nl)
    Aw = 0.05 + 0.02 * np.arange(nl)
    Av, At = 0.15, 0.6
    u, v, w, txz = o.field(), o.field(), o.field(), o.field()
    for k in range(nl):
        u[k, :, :o.nx] = U[k] + A[k] * np.cos(m * X)
        v[k, :, :o.nx] = V[k] + Av * np.sin(m * X + Y)
        w[k, :, :o.nx] = W[k] + Aw[k] * np.cos(m * X)
        txz[k, :, :o.nx] = T[k] + At * np.cos(X)
    st = o.avg_stats(u, v, w, txz)
    ex = {"ubar": U, "vbar": V, "wbar": W, "txz": T}
    Am = np.concatenate([[0.0], 0.5 * (A[1:] + A[:-1])])
    Awz = np.concatenate([[0.0], Aw[1:]])
    ex["u2"] = Am ** 2 / 2
    ex["v2"] = np.concatenate([[0.0], np.full(nl - 1, Av * Av / 2)])
    ex["w2"] = Awz ** 2 / 2
    ex["uw"] = Am * Awz / 2
    ex["stress"] = ex["uw"] + T
    names = ("ubar", "vbar", "wbar", "u2", "v2", "w2", "uw", "txz", "stress")
    for i, n in enumerate(names):
        assert np.max(np.abs(st[i * nl:(i + 1) * nl] - ex[n])) < 1e-14, (n, st[i * nl:(i + 1) * nl], ex[n])
    eoz = st[9 * nl:].reshape(nl, o.nx // 2)
    for k in range(nl):
        e = np.zeros(o.nx // 2)
        e[0] = 0.5 * U[k] ** 2
        e[m] = A[k] ** 2 / 4
        z = (k + 0.5) * o.dz
        assert np.max(np.abs(eoz[k] - e / z)) < 1e-13 * max(1.0, np.max(e / z)), k


def test_surface_stats_uniform(O):
    """Uniform first-level fields: u* from the log law over the mean speed,
    the mean local u*, u* from the mean stress (txz^2 + tyz^2)^(1/4), the
    first-level theta, theta_v = theta (1 + 0.61e-3 q q_scale), q, the
    surface flux and the Obukhov length -u*^3 / (kappa (g z_i/u*^2 / theta)
    (wt + 0.61 theta 1e-3 q_scale wq))."""
    o = O.Oracle(16, 16, 9, z0=2e-4)
    sp = O.scalar_params()
    o.set_scalars(sp, None, None)
    U, V, T1, T2, P1, P0, S = 3.0, -1.2, -0.8, 0.3, 0.25, 0.05, 0.9
    Th, Q, Wt, Wq = 0.71, 0.012, 2e-4, -1e-5
    pl = lambda a: np.full((o.ny, o.ld), a)
    p2 = lambda a: np.full((o.ny, o.nx), a)
    out = o.surface_stats(pl(U), pl(V), pl(T1), pl(T2), p2(P1), p2(P0), p2(S), pl(Th), pl(Q), pl(Wt), pl(Wq))
    kappa = 0.41
    us = math.hypot(U, V) * kappa / (math.log(0.5 * o.dz / 2e-4) - P1 + P0)
    thv = Th * (1 + 0.61e-3 * Q * sp.q_scale)
    wtv = Wt + 0.61 * Th * (1e-3 * sp.q_scale * Wq)
    g_hat = sp.g * sp.z_i / sp.u_star ** 2
    ex = [us, S, (T1 * T1 + T2 * T2) ** 0.25, Th, thv, Q, Wt, -us ** 3 / (kappa * (g_hat / Th) * wtv)]
    assert np.allclose(out, ex, rtol=1e-13, atol=0), (out, ex)
    neutral = o.surface_stats(pl(U), pl(V), pl(T1), pl(T2), p2(P1), p2(P0), p2(S))
    assert np.allclose(neutral[:3], ex[:3], rtol=1e-13) and np.all(neutral[3:] == 0.0)
