"""ctypes wrapper of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the les_amd product path.

Fields are numpy float64 arrays of shape (nlev, ny, ld) in C order, which is
the reference's Fortran (ld, ny, nlev) column-major layout byte for byte.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# LES_ORACLE_LIB: another build of the same source (the mutation tests load
# oracles with planted mistakes this way)
_LIB = os.environ.get("LES_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None

D = C.POINTER(C.c_double)


class Params(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double),
                ("dt", C.c_double), ("z0", C.c_double), ("force", C.c_double),
                ("inilag", C.c_int), ("initu", C.c_int), ("cs_count", C.c_int)]


class SgsState(C.Structure):
    _fields_ = [(n, D) for n in ("F_LM", "F_MM", "F_QN", "F_NN", "Beta", "Cs_opt2", "Nu_t",
                                 "u_lag", "v_lag", "w_lag", "TnLMMM", "TnQNNN",
                                 "L11t", "L22t", "L33t", "Q11t", "Q22t", "Q33t",
                                 "Beta_avg", "Betaclip_avg")]


SgsState._fields_names = staticmethod(lambda: [f[0] for f in SgsState._fields_])


class ScalarParams(C.Structure):
    _fields_ = [("enable", C.c_int), ("model", C.c_int), ("initsc", C.c_int), ("passive", C.c_int),
                ("Pr", C.c_double), ("z_i", C.c_double), ("u_star", C.c_double), ("T_scale", C.c_double),
                ("q_scale", C.c_double), ("g", C.c_double), ("dTdz_top", C.c_double), ("dqdz_top", C.c_double)]


SCALAR_SGS_NAMES = ("G_LM", "G_MM", "G_QN", "G_NN", "u_lag_s", "v_lag_s", "w_lag_s", "Pr_t", "Cs2_Pr",
                    "phi_h", "psi_m")


class Forcing(C.Structure):
    _fields_ = [("coriolis", C.c_int), ("ug0", C.c_double), ("vg0", C.c_double), ("u_star", C.c_double),
                ("z_i", C.c_double), ("ubc", C.c_int), ("damping_method", C.c_int)]


class ScalarSgs(C.Structure):
    _fields_ = [(n, D) for n in SCALAR_SGS_NAMES]


class State(C.Structure):
    _fields_ = ([("u", D), ("v", D), ("w", D), ("grad", D * 9),
                 ("RHSx", D), ("RHSy", D), ("RHSz", D), ("RHSx_f", D), ("RHSy_f", D), ("RHSz_f", D),
                 ("tau", D * 6), ("divtx", D), ("divty", D), ("divtz", D),
                 ("p", D), ("dpdx", D), ("dpdy", D), ("dpdz", D),
                 ("psi_m", D), ("psi_m0", D), ("phi_m", D), ("ustar_avg", D),
                 ("sgs", SgsState), ("jt", C.c_int), ("jt_total", C.c_int),
                 ("theta", D), ("qmix", D), ("dTdx", D), ("dTdy", D), ("dTdz", D),
                 ("dqdx", D), ("dqdy", D), ("dqdz", D),
                 ("RHS_T", D), ("RHS_Tf", D), ("RHS_q", D), ("RHS_qf", D), ("beta_scal", D),
                 ("sgs_t3", D), ("sgs_q3", D), ("psi_h", D), ("psi_h0", D), ("phi_h", D),
                 ("ssgs", ScalarSgs)])


def scalar_params(enable=1, model=7, initsc=0, passive=0, Pr=0.55, z_i=1000.0, u_star=1.0, T_scale=400.0,
                  q_scale=100.0, g=9.81, dTdz_top=0.01, dqdz_top=-0.004):
    """param.f90:53,72-90,139-146,174 defaults; dTdz_top = the inversion of the
    convective configuration (0.01 K/m)."""
    return ScalarParams(enable, model, initsc, passive, Pr, z_i, u_star, T_scale, q_scale, g, dTdz_top, dqdz_top)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        vp = C.c_void_p
        L.or_create.restype = vp
        L.or_create.argtypes = [C.POINTER(Params)]
        L.or_destroy.argtypes = [vp]
        L.or_state_create.restype = C.POINTER(State)
        L.or_state_create.argtypes = [vp]
        L.or_state_destroy.argtypes = [C.POINTER(State)]
        L.or_state_init_ic.argtypes = [vp, C.POINTER(State), C.c_uint64, C.c_double, C.c_double, C.c_double]
        L.or_state_init_scalar_ic.argtypes = [vp, C.POINTER(State), C.c_uint64, C.c_double, C.c_double,
                                              C.c_double, C.c_double, C.c_double]
        L.or_set_scalars.argtypes = [vp, C.POINTER(ScalarParams), D, D]
        L.or_set_forcing.argtypes = [vp, C.POINTER(Forcing)]
        L.or_step.argtypes = [vp, C.POINTER(State)]
        L.or_uniform.restype = C.c_double
        L.or_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.or_rmsdiv.restype = C.c_double
        L.or_tridag_array.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous float64"
    return a.ctypes.data_as(D)


class Oracle:
    """One grid: owns the C context (wavenumber/filter tables, FFT plans)."""

    def __init__(self, nx, ny, nz, Lx=2 * math.pi, Ly=2 * math.pi, Lz=1.0, dt=2e-4,
                 z0=1e-4, force=1.0, inilag=1, initu=1, cs_count=5):
        self.L = lib()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.lh = nx // 2 + 1
        self.ld = 2 * self.lh
        self.nx2, self.ny2 = 3 * nx // 2, 3 * ny // 2
        self.ld_big = 2 * (self.nx2 // 2 + 1)
        self.params = Params(nx, ny, nz, Lx, Ly, Lz, dt, z0, force, inilag, initu, cs_count)
        self.ctx = self.L.or_create(C.byref(self.params))
        self.dx, self.dy, self.dz = Lx / nx, Ly / ny, Lz / (nz - 1)

    def __del__(self):
        try:
            self.L.or_destroy(self.ctx)
        except Exception:
            pass

    # ---- helpers ----
    def field(self, nlev=None, big=False):
        nlev = self.nz if nlev is None else nlev
        if big:
            return np.zeros((nlev, self.ny2, self.ld_big))
        return np.zeros((nlev, self.ny, self.ld))

    def tables(self):
        n = self.lh * self.ny
        out = [np.zeros((self.ny, self.lh)) for _ in range(5)]
        self.L.or_get_tables(C.c_void_p(self.ctx), *[_p(o) for o in out])
        return dict(zip(("kx", "ky", "k2", "G_test", "G_test_test"), out))

    def set_sgs_model(self, model):
        self.L.or_set_sgs_model(C.c_void_p(self.ctx), int(model))

    def set_time(self, jt, jt_total, flm_init=None, fqn_init=None):
        """jt/jt_total as the routine-level calls see them (param.f90:147-148)."""
        self.L.or_set_time(C.c_void_p(self.ctx), jt, jt_total)
        if flm_init is not None:
            self.L.or_set_lag_init(C.c_void_p(self.ctx), int(flm_init), int(fqn_init))

    def sgs_state(self):
        """Numpy-backed or_sgs_state; returns (struct, dict of arrays)."""
        arrs = {n: self.field() for n in SgsState._fields_names() if n not in ("Beta_avg", "Betaclip_avg")}
        arrs["Beta_avg"] = np.zeros(self.nz)
        arrs["Betaclip_avg"] = np.zeros(self.nz)
        st = SgsState(*[_p(arrs[n]) for n in SgsState._fields_names()])
        return st, arrs

    # ---- scalars ----
    def set_scalars(self, params, wt_s=None, wq_s=None):
        """params: ScalarParams; wt_s, wq_s: (ny, nx) kinematic surface fluxes."""
        self.sp = params
        self._wt = np.ascontiguousarray(np.zeros((self.ny, self.nx)) if wt_s is None else wt_s, dtype=np.float64)
        self._wq = np.ascontiguousarray(np.zeros((self.ny, self.nx)) if wq_s is None else wq_s, dtype=np.float64)
        self.L.or_set_scalars(C.c_void_p(self.ctx), C.byref(params), _p(self._wt), _p(self._wq))

    def set_scalar_lag_init(self, glm_init, gqn_init):
        self.L.or_set_scalar_lag_init(C.c_void_p(self.ctx), int(glm_init), int(gqn_init))

    def scalar_sgs_state(self, phi_h=None, psi_m=None):
        arrs = {n: self.field() for n in SCALAR_SGS_NAMES[:9]}
        arrs["phi_h"] = np.ones((self.ny, self.nx)) if phi_h is None else phi_h
        arrs["psi_m"] = np.zeros((self.ny, self.nx)) if psi_m is None else psi_m
        st = ScalarSgs(*[_p(arrs[n]) for n in SCALAR_SGS_NAMES])
        return st, arrs

    def obukhov(self, u, v, theta, qmix, sgs_t3, sgs_q3, psi_m, psi_m0, psi_h, psi_h0, phi_m, phi_h, ustar_avg):
        self.L.or_obukhov(C.c_void_p(self.ctx), *[_p(a) for a in (u, v, theta, qmix, sgs_t3, sgs_q3, psi_m, psi_m0,
                                                                   psi_h, psi_h0, phi_m, phi_h, ustar_avg)])

    def calcbeta(self, theta, qmix, beta):
        self.L.or_calcbeta(C.c_void_p(self.ctx), _p(theta), _p(qmix), _p(beta))

    def scalar_rhs_calc(self, which, dsdx, dsdy, dsdz, u, v, w, nu_t, pr, ustar_avg, phi_h, rhs, sgs_vert1):
        self.L.or_scalar_rhs_calc(C.c_void_p(self.ctx), int(which),
                                  *[_p(a) for a in (dsdx, dsdy, dsdz, u, v, w, nu_t, pr, ustar_avg, phi_h, rhs,
                                                    sgs_vert1)])

    def step_scalar(self, which, scalar, rhs, rhs_f):
        self.L.or_step_scalar(C.c_void_p(self.ctx), int(which), _p(scalar), _p(rhs), _p(rhs_f))

    def interpolag_scalar_sdep(self, sst):
        self.L.or_interpolag_scalar_sdep(C.c_void_p(self.ctx), C.byref(sst))

    def lagrange_scalar_sdep(self, S, u, v, w, theta, st, sst):
        s6 = (D * 6)(*[_p(a) for a in S])
        self.L.or_lagrange_scalar_sdep(C.c_void_p(self.ctx), s6, _p(u), _p(v), _p(w), _p(theta), C.byref(st),
                                       C.byref(sst))

    def sgs_stag_scalar(self, grads, u, v, w, theta, taus, st, sst):
        g = (D * 9)(*[_p(a) for a in grads])
        t = (D * 6)(*[_p(a) for a in taus])
        self.L.or_sgs_stag_scalar(C.c_void_p(self.ctx), g, _p(u), _p(v), _p(w), _p(theta), t, C.byref(st),
                                  C.byref(sst))

    def rtnewt(self, A):
        a = np.ascontiguousarray(A, dtype=np.float64)
        self.L.or_rtnewt.restype = C.c_double
        return self.L.or_rtnewt(_p(a))

    def scaledep_dynamic(self, S, u, v, w):
        """S: the six dummy-argument fields S11..S33; returns (Cs_opt2, beta) per level."""
        cs, beta = np.zeros(self.nz), np.zeros(self.nz)
        s6 = (D * 6)(*[_p(a) for a in S])
        self.L.or_scaledep_dynamic(C.c_void_p(self.ctx), s6, _p(u), _p(v), _p(w), _p(cs), _p(beta))
        return cs, beta

    def sgs_stag_model3(self, grads, u, v, w, taus, st):
        g = (D * 9)(*[_p(a) for a in grads])
        t = (D * 6)(*[_p(a) for a in taus])
        self.L.or_sgs_stag_model3(C.c_void_p(self.ctx), g, _p(u), _p(v), _p(w), t, C.byref(st))

    def interpolag_sdep(self, st):
        self.L.or_interpolag_sdep(C.c_void_p(self.ctx), C.byref(st))

    def lagrange_sdep(self, S, u, v, w, st):
        s6 = (D * 6)(*[_p(a) for a in S])
        self.L.or_lagrange_sdep(C.c_void_p(self.ctx), s6, _p(u), _p(v), _p(w), C.byref(st))

    def sgs_stag(self, grads, u, v, w, taus, st):
        g = (D * 9)(*[_p(a) for a in grads])
        t = (D * 6)(*[_p(a) for a in taus])
        self.L.or_sgs_stag(C.c_void_p(self.ctx), g, _p(u), _p(v), _p(w), t, C.byref(st))

    def sgs_stag_model(self, grads, u, v, w, taus, st, model):
        g = (D * 9)(*[_p(a) for a in grads])
        t = (D * 6)(*[_p(a) for a in taus])
        self.L.or_sgs_stag_model(C.c_void_p(self.ctx), g, _p(u), _p(v), _p(w), t, C.byref(st), C.c_int(model))

    def std_dynamic(self, S, u, v, w):
        cs = np.zeros(self.nz)
        s6 = (D * 6)(*[_p(a) for a in S])
        self.L.or_std_dynamic(C.c_void_p(self.ctx), s6, _p(u), _p(v), _p(w), _p(cs))
        return cs

    def new_scaledep_dynamic(self, S, u, v, w):
        """new_scaledep_dynamic.f90 (model 6): Cs_opt2 per level."""
        cs = np.zeros(self.nz)
        s6 = (D * 6)(*[_p(a) for a in S])
        self.L.or_new_scaledep_dynamic(C.c_void_p(self.ctx), s6, _p(u), _p(v), _p(w), _p(cs))
        return cs

    def scalar_scaledep_dynamic(self, S, u, v, w, theta, Cs_opt2, psi_m, phi_h, Pr_t):
        """scalar_scaledep_dynamic.f90 (model 6 with scalars): filters theta in
        place, writes Pr_t levels 1..nz-1; returns Cs2_Pr per level (nz-1)."""
        cpr = np.zeros(self.nz - 1)
        s6 = (D * 6)(*[_p(a) for a in S])
        pm = np.ascontiguousarray(psi_m, dtype=np.float64)
        ph = np.ascontiguousarray(phi_h, dtype=np.float64)
        self.L.or_scalar_scaledep_dynamic(C.c_void_p(self.ctx), s6, _p(u), _p(v), _p(w), _p(theta), _p(Cs_opt2),
                                          _p(pm), _p(ph), _p(cpr), _p(Pr_t))
        return cpr

    def interpolag_ssim(self, st):
        self.L.or_interpolag_ssim(C.c_void_p(self.ctx), C.byref(st))

    def lagrange_ssim(self, S, u, v, w, st):
        s6 = (D * 6)(*[_p(a) for a in S])
        self.L.or_lagrange_ssim(C.c_void_p(self.ctx), s6, _p(u), _p(v), _p(w), C.byref(st))

    def set_forcing(self, coriolis=0, ug0=0.0, vg0=0.0, u_star=0.45, z_i=1000.0, ubc=0, damping_method=1):
        self.forcing = Forcing(coriolis, ug0, vg0, u_star, z_i, ubc, damping_method)
        self.L.or_set_forcing(C.c_void_p(self.ctx), C.byref(self.forcing))

    def sponge(self):
        sp = np.zeros(self.nz)
        self.L.or_get_sponge(C.c_void_p(self.ctx), _p(sp))
        return sp

    # ---- transforms ----
    def rfft2_fwd(self, plane, big=False):
        self.L.or_rfft2_fwd(C.c_void_p(self.ctx), _p(plane), int(big))

    def rfft2_inv(self, plane, big=False):
        self.L.or_rfft2_inv(C.c_void_p(self.ctx), _p(plane), int(big))

    # ---- routines ----
    def filt_da(self, f, dfdx, dfdy):
        self.L.or_filt_da(C.c_void_p(self.ctx), _p(f), _p(dfdx), _p(dfdy), f.shape[0])

    def ddx(self, dfdx, f):
        self.L.or_ddx(C.c_void_p(self.ctx), _p(dfdx), _p(f), f.shape[0])

    def ddy(self, dfdy, f):
        self.L.or_ddy(C.c_void_p(self.ctx), _p(dfdy), _p(f), f.shape[0])

    def ddxy(self, dfdx, dfdy, f):
        self.L.or_ddxy(C.c_void_p(self.ctx), _p(dfdx), _p(dfdy), _p(f), f.shape[0])

    def test_filter(self, plane, G):
        self.L.or_test_filter(C.c_void_p(self.ctx), _p(plane), _p(G))

    def test_filter_init(self, alpha):
        G = np.zeros((self.ny, self.lh))
        self.L.or_test_filter_init(C.c_void_p(self.ctx), C.c_double(alpha), _p(G))
        return G

    def padd(self, big_c, small_c):
        self.L.or_padd(C.c_void_p(self.ctx), _p(big_c), _p(small_c))

    def unpadd(self, small_c, big_c):
        self.L.or_unpadd(C.c_void_p(self.ctx), _p(small_c), _p(big_c))

    def dealias1(self, u, u_big):
        self.L.or_dealias1(C.c_void_p(self.ctx), _p(u), _p(u_big), u.shape[0])

    def dealias2(self, u, u_big):
        self.L.or_dealias2(C.c_void_p(self.ctx), _p(u), _p(u_big), u.shape[0])

    def ddz_uv(self, dfdz, f):
        self.L.or_ddz_uv(C.c_void_p(self.ctx), _p(dfdz), _p(f))

    def ddz_w(self, dfdz, f):
        self.L.or_ddz_w(C.c_void_p(self.ctx), _p(dfdz), _p(f))

    def convec(self, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz):
        self.L.or_convec(C.c_void_p(self.ctx), *[_p(a) for a in (u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz)])

    def wallstress(self, u, v, dudz, dvdz, txz, tyz, psi_m=None, psi_m0=None, phi_m=None):
        n2 = (self.ny, self.nx)
        psi_m = np.zeros(n2) if psi_m is None else psi_m
        psi_m0 = np.zeros(n2) if psi_m0 is None else psi_m0
        phi_m = np.ones(n2) if phi_m is None else phi_m
        ustar = np.zeros(n2)
        self.L.or_wallstress(C.c_void_p(self.ctx), *[_p(a) for a in (u, v, dudz, dvdz, txz, tyz, psi_m, psi_m0, phi_m, ustar)])
        return ustar

    def calc_sij(self, grads):
        S = [self.field() for _ in range(6)]
        g = (D * 9)(*[_p(a) for a in grads])
        s = (D * 6)(*[_p(a) for a in S])
        self.L.or_calc_sij(C.c_void_p(self.ctx), g, s)
        return S

    def divstress_uv(self, divt, tx, ty, tz):
        self.L.or_divstress_uv(C.c_void_p(self.ctx), _p(divt), _p(tx), _p(ty), _p(tz))

    def divstress_w(self, divt, tx, ty, tz):
        self.L.or_divstress_w(C.c_void_p(self.ctx), _p(divt), _p(tx), _p(ty), _p(tz))

    def press_stag_array(self, u, v, w, divtz, p, dpdx, dpdy):
        self.L.or_press_stag_array(C.c_void_p(self.ctx), *[_p(a) for a in (u, v, w, divtz, p, dpdx, dpdy)])

    def tridag_array(self, a, b, c, r, u):
        return self.L.or_tridag_array(C.c_void_p(self.ctx), _p(a), _p(b), _p(c), _p(r), _p(u))

    def project(self, u, v, w, dpdx, dpdy, dpdz):
        self.L.or_project(C.c_void_p(self.ctx), *[_p(a) for a in (u, v, w, dpdx, dpdy, dpdz)])

    def rmsdiv(self, dudx, dvdy, dwdz):
        return self.L.or_rmsdiv(C.c_void_p(self.ctx), _p(dudx), _p(dvdy), _p(dwdz))

    def spectrum(self, f):
        spec = np.zeros((f.shape[0], self.nx // 2))
        self.L.or_spectrum(C.c_void_p(self.ctx), _p(f), f.shape[0], _p(spec))
        return spec

    def avg_stats(self, u, v, w, txz):
        out = np.zeros(9 * (self.nz - 1) + (self.nz - 1) * (self.nx // 2))
        self.L.or_avg_stats(C.c_void_p(self.ctx), _p(u), _p(v), _p(w), _p(txz), _p(out))
        return out

    def surface_stats(self, u, v, txz, tyz, psi_m, psi_m0, ustar_avg, theta=None, qmix=None, sgs_t3=None,
                      sgs_q3=None):
        """First-level planes (ny, ld) of u, v, txz, tyz (and theta, qmix,
        sgs_t3, sgs_q3), (ny, nx) psi_m, psi_m0, ustar_avg."""
        out = np.zeros(8)
        sc = [None if a is None else _p(np.ascontiguousarray(a)) for a in (theta, qmix, sgs_t3, sgs_q3)]
        self.L.or_surface_stats(C.c_void_p(self.ctx), *[_p(np.ascontiguousarray(a)) for a in
                                                       (u, v, txz, tyz, psi_m, psi_m0, ustar_avg)], *sc, _p(out))
        return out

    def timestep_conditions(self, u, v, w, nu_t):
        out = np.zeros(4)
        self.L.or_timestep_conditions(C.c_void_p(self.ctx), _p(u), _p(v), _p(w), _p(nu_t), _p(out))
        return out


class OracleSim:
    """Whole-step oracle state (main.f90:202-415, neutral model=5)."""

    FIELDS3 = ("u", "v", "w", "RHSx", "RHSy", "RHSz", "RHSx_f", "RHSy_f", "RHSz_f",
               "divtx", "divty", "divtz", "dpdx", "dpdy", "dpdz")
    SGS3 = ("F_LM", "F_MM", "F_QN", "F_NN", "Beta", "Cs_opt2", "Nu_t", "u_lag", "v_lag", "w_lag",
            "TnLMMM", "TnQNNN", "L11t", "L22t", "L33t", "Q11t", "Q22t", "Q33t")
    GRADS = ("dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz")
    TAUS = ("txx", "txy", "txz", "tyy", "tyz", "tzz")

    SCAL3 = ("theta", "qmix", "dTdx", "dTdy", "dTdz", "dqdx", "dqdy", "dqdz", "RHS_T", "RHS_Tf", "RHS_q",
             "RHS_qf", "beta_scal")
    SSGS3 = ("G_LM", "G_MM", "G_QN", "G_NN", "u_lag_s", "v_lag_s", "w_lag_s", "Pr_t", "Cs2_Pr")

    def __init__(self, orc: Oracle, seed=42, amp=0.1, ustar=1.0, kappa_ic=0.4, scalar_ic=None):
        """scalar_ic: dict(T_init, z_inv, inv_strength, amp_K, q_init) when the
        oracle's scalars are enabled (Oracle.set_scalars)."""
        self.o = orc
        self.L = orc.L
        self.s = self.L.or_state_create(C.c_void_p(orc.ctx))
        self.L.or_state_init_ic(C.c_void_p(orc.ctx), self.s, seed, amp, ustar, kappa_ic)
        if scalar_ic is not None:
            k = dict(T_init=283.0, z_inv=0.8, inv_strength=0.01, amp_K=0.1, q_init=1.0)
            k.update(scalar_ic)
            self.L.or_state_init_scalar_ic(C.c_void_p(orc.ctx), self.s, seed, k["T_init"], k["z_inv"],
                                           k["inv_strength"], k["amp_K"], k["q_init"])

    def __del__(self):
        try:
            self.L.or_state_destroy(self.s)
        except Exception:
            pass

    def _view(self, ptr, nlev=None):
        o = self.o
        nlev = o.nz if nlev is None else nlev
        return np.ctypeslib.as_array(ptr, shape=(nlev, o.ny, o.ld))

    def __getitem__(self, name):
        st = self.s.contents
        if name in self.FIELDS3:
            return self._view(getattr(st, name))
        if name == "p":
            return self._view(st.p, self.o.nz + 1)
        if name in self.GRADS:
            return self._view(st.grad[self.GRADS.index(name)])
        if name in self.TAUS:
            return self._view(st.tau[self.TAUS.index(name)])
        if name in self.SGS3:
            return self._view(getattr(st.sgs, name))
        if name in ("Beta_avg", "Betaclip_avg"):
            return np.ctypeslib.as_array(getattr(st.sgs, name), shape=(self.o.nz,))
        if name in ("psi_m", "psi_m0", "phi_m", "ustar_avg", "psi_h", "psi_h0", "phi_h"):
            return np.ctypeslib.as_array(getattr(st, name), shape=(self.o.ny, self.o.nx))
        if name in self.SCAL3:
            return self._view(getattr(st, name))
        if name in ("sgs_t3", "sgs_q3"):
            return self._view(getattr(st, name), 1)[0]
        if name in self.SSGS3:
            return self._view(getattr(st.ssgs, name))
        raise KeyError(name)

    @property
    def jt(self):
        return self.s.contents.jt

    def set_time(self, jt, jt_total):
        """Step counters of the state (param.f90:147-148); the next step is
        jt + 1, jt_total + 1."""
        st = self.s.contents
        st.jt = int(jt)
        st.jt_total = int(jt_total)

    def step(self, n=1):
        for _ in range(n):
            self.L.or_step(C.c_void_p(self.o.ctx), self.s)


MLP_NIN, MLP_NH, MLP_NOUT = 8, 64, 2


def mlp_nparam():
    return lib().or_mlp_nparam()


def mlp_forward(theta, x, bf16=True):
    """Learned-closure MLP on (P, 8) inputs -> (P, 2); theta flat float64."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    P = x.shape[0]
    y = np.zeros((P, MLP_NOUT))
    lib().or_mlp_forward(_p(theta), _p(x), C.c_long(P), _p(y), int(bf16))
    return y


def mlp_loss_grad(theta, x, t, bf16=True):
    """Loss 1/(2P) sum (y - t)^2 and its gradient with respect to theta."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.zeros_like(theta)
    L = lib()
    L.or_mlp_loss_grad.restype = C.c_double
    loss = L.or_mlp_loss_grad(_p(theta), _p(x), _p(t), C.c_long(x.shape[0]), _p(g), int(bf16))
    return loss, g


def mlp_slices():
    """Offsets of W1 b1 W2 b2 W3 b3 W4 b4 in the flat parameter vector."""
    shapes = [(MLP_NH, MLP_NIN), (MLP_NH,), (MLP_NH, MLP_NH), (MLP_NH,), (MLP_NH, MLP_NH), (MLP_NH,),
              (MLP_NOUT, MLP_NH), (MLP_NOUT,)]
    out, o = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append((o, sh))
        o += n
    return out


def uniform(seed, index):
    return lib().or_uniform(seed, index)


def threads():
    """OpenMP threads of the oracle's transforms (its results do not depend on them)."""
    return lib().or_threads()


def set_threads(n):
    lib().or_set_threads(int(n))
