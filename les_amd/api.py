"""Host-side mirror of the reference's routine interface over the HIP C ABI.

Fields are torch tensors on the GPU with shape (nlev, ny, ld), C-contiguous,
float64 (or float32 for the FP32-bulk variants): byte-for-byte the
reference's Fortran (ld, ny, nlev) column-major arrays.  Method names and
argument order follow the reference subroutines (filt_da(f, dfdx, dfdy),
ddx(dfdx, f), convec(cx, cy, cz) ...); module globals they read (u, v, w,
gradients, F_LM ...) are passed explicitly.  Errors the reference turned into
`stop` raise LesError.
"""
import ctypes as C
import math

import numpy as np

import torch

from . import _lib
from ._lib import Grid, SgsFields, SGS_NAMES, ScalarParams, ScalarSgsFields, SCALAR_SGS_NAMES


class LesError(RuntimeError):
    pass


def _ptr(t, dtype=None, numel=None, what="array"):
    """Device address of `t` after checking what the C ABI assumes: a
    contiguous GPU tensor, of `dtype` when given, holding at least `numel`
    elements when given (the kernels index the full reference layout)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise LesError(f"{what}: expected a torch tensor on the GPU, got {type(t).__name__}")
    if not t.is_cuda:
        raise LesError(f"{what}: les_amd works on GPU tensors only (no CPU fallback)")
    if not t.is_contiguous():
        raise LesError(f"{what}: les_amd needs contiguous tensors in the reference layout")
    if dtype is not None and t.dtype != dtype:
        raise LesError(f"{what}: expected {dtype}, got {t.dtype}")
    if numel is not None and t.numel() < numel:
        raise LesError(f"{what}: needs at least {numel} elements, got {t.numel()} (shape {tuple(t.shape)})")
    return C.c_void_p(t.data_ptr())


F64, F32 = torch.float64, torch.float32


class Les:
    """One grid: the les_ctx (wavenumbers, test filters, FFT twiddles, scratch)."""

    def __init__(self, nx, ny, nz, Lx=2 * math.pi, Ly=2 * math.pi, Lz=1.0, dt=2e-4, z0=1e-4,
                 force=1.0, inilag=1, initu=1, cs_count=5, stream=None):
        if not torch.cuda.is_available():
            raise LesError("les_amd needs a GPU (HIP); none is visible")
        self.L = _lib.load()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.lh = nx // 2 + 1
        self.ld = 2 * self.lh
        self.nx2, self.ny2 = 3 * nx // 2, 3 * ny // 2
        self.lh_big = self.nx2 // 2 + 1
        self.ld_big = 2 * self.lh_big
        self.dz = Lz / (nz - 1)
        self.grid = Grid(nx, ny, nz, Lx, Ly, Lz, dt, z0, force, inilag, initu, cs_count)
        h = C.c_void_p()
        rc = self.L.les_create(C.byref(self.grid), C.byref(h))
        if rc != 0:
            raise LesError(f"les_create failed ({rc})")
        self.ctx = h
        s = stream if stream is not None else torch.cuda.current_stream()
        self._chk(self.L.les_set_stream(self.ctx, C.c_void_p(s.cuda_stream)))

    def __del__(self):
        try:
            if getattr(self, "ctx", None):
                self.L.les_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            msg = self.L.les_last_error(self.ctx)
            raise LesError((msg or b"les_amd error").decode())

    @staticmethod
    def _sfx(t):
        if not isinstance(t, torch.Tensor):
            raise LesError(f"expected a torch tensor on the GPU, got {type(t).__name__}")
        if t.dtype not in (F64, F32):
            raise LesError(f"les_amd routines take float64 (or float32 for the FP32-bulk variants), got {t.dtype}")
        return "_f32" if t.dtype == F32 else ""

    # ---- argument checks: sizes of the reference layout ----
    def _f(self, t, what, nlev=None, dtype=F64):
        """A (nlev, ny, ld) field (nlev defaults to nz)."""
        nlev = self.nz if nlev is None else nlev
        return _ptr(t, dtype, nlev * self.ny * self.ld, what)

    def _fbig(self, t, what, nlev, dtype=F64):
        """A (nlev, ny2, ld_big) padded field."""
        return _ptr(t, dtype, nlev * self.ny2 * self.ld_big, what)

    def _pl(self, t, what, dtype=F64):
        """One (ny, ld) plane."""
        return _ptr(t, dtype, self.ny * self.ld, what)

    def _p2(self, t, what):
        """One (ny, nx) surface array (always float64)."""
        return _ptr(t, F64, self.ny * self.nx, what)

    def _nlev(self, f, what="f"):
        if not isinstance(f, torch.Tensor) or f.dim() < 1:
            raise LesError(f"{what}: expected a (nlev, ny, ld) tensor")
        return f.shape[0]

    # ---- allocation helpers (reference layout) ----
    def field(self, nlev=None, dtype=torch.float64, big=False):
        nlev = self.nz if nlev is None else nlev
        if big:
            return torch.zeros((nlev, self.ny2, self.ld_big), dtype=dtype, device="cuda")
        return torch.zeros((nlev, self.ny, self.ld), dtype=dtype, device="cuda")

    def set_time(self, jt, jt_total, flm_init=None, fqn_init=None):
        self._chk(self.L.les_set_time(self.ctx, jt, jt_total))
        if flm_init is not None:
            self._chk(self.L.les_set_lag_init(self.ctx, int(flm_init), int(fqn_init)))

    def tables(self):
        import numpy as np
        out = [np.zeros((self.ny, self.lh)) for _ in range(5)]
        p = [o.ctypes.data_as(C.c_void_p) for o in out]
        self._chk(self.L.les_get_tables(self.ctx, *p))
        return dict(zip(("kx", "ky", "k2", "G_test", "G_test_test"), out))

    def sync(self):
        self._chk(self.L.les_sync(self.ctx))

    # ---- routines (reference argument order) ----
    def filt_da(self, f, dfdx, dfdy):
        sfx, n = self._sfx(f), self._nlev(f)
        dt = F32 if sfx else F64
        self._chk(getattr(self.L, "les_filt_da" + sfx)(self.ctx, self._f(f, "f", n, dt), self._f(dfdx, "dfdx", n, dt),
                                                       self._f(dfdy, "dfdy", n, dt), n))

    def ddx(self, dfdx, f):
        sfx, n = self._sfx(f), self._nlev(f)
        dt = F32 if sfx else F64
        self._chk(getattr(self.L, "les_ddx" + sfx)(self.ctx, self._f(dfdx, "dfdx", n, dt), self._f(f, "f", n, dt), n))

    def ddy(self, dfdy, f):
        sfx, n = self._sfx(f), self._nlev(f)
        dt = F32 if sfx else F64
        self._chk(getattr(self.L, "les_ddy" + sfx)(self.ctx, self._f(dfdy, "dfdy", n, dt), self._f(f, "f", n, dt), n))

    def ddxy(self, dfdx, dfdy, f):
        n = self._nlev(f)
        self._chk(self.L.les_ddxy(self.ctx, self._f(dfdx, "dfdx", n), self._f(dfdy, "dfdy", n), self._f(f, "f", n), n))

    def test_filter_init(self, alpha):
        G = torch.zeros((self.ny, self.lh), dtype=torch.float64, device="cuda")
        self._chk(self.L.les_test_filter_init(self.ctx, C.c_double(alpha), _ptr(G)))
        return G

    def test_filter(self, plane, G):
        """G must have the plane's dtype (the FP32 variant reads float G)."""
        sfx = self._sfx(plane)
        dt = F32 if sfx else F64
        self._chk(getattr(self.L, "les_test_filter" + sfx)(self.ctx, self._pl(plane, "plane", dt),
                                                           _ptr(G, dt, self.ny * self.lh, "G")))

    def padd(self, u_big, u):
        self._chk(self.L.les_padd(self.ctx, self._fbig(u_big, "u_big", 1), self._pl(u, "u")))

    def unpadd(self, cc, cc_big):
        self._chk(self.L.les_unpadd(self.ctx, self._pl(cc, "cc"), self._fbig(cc_big, "cc_big", 1)))

    def dealias1(self, u, u_big):
        sfx, n = self._sfx(u), self._nlev(u, "u")
        dt = F32 if sfx else F64
        self._chk(getattr(self.L, "les_dealias1" + sfx)(self.ctx, self._f(u, "u", n, dt),
                                                        self._fbig(u_big, "u_big", n, dt), n))

    def dealias2(self, u, u_big):
        sfx, n = self._sfx(u), self._nlev(u, "u")
        dt = F32 if sfx else F64
        self._chk(getattr(self.L, "les_dealias2" + sfx)(self.ctx, self._f(u, "u", n, dt),
                                                        self._fbig(u_big, "u_big", n, dt), n))

    def ddz_uv(self, dfdz, f):
        self._chk(self.L.les_ddz_uv(self.ctx, self._f(dfdz, "dfdz"), self._f(f, "f")))

    def ddz_w(self, dfdz, f):
        self._chk(self.L.les_ddz_w(self.ctx, self._f(dfdz, "dfdz"), self._f(f, "f")))

    def convec(self, u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz):
        sfx = self._sfx(u)
        dt = F32 if sfx else F64
        names = ("u", "v", "w", "dudy", "dudz", "dvdx", "dvdz", "dwdx", "dwdy", "cx", "cy", "cz")
        args = (u, v, w, dudy, dudz, dvdx, dvdz, dwdx, dwdy, cx, cy, cz)
        self._chk(getattr(self.L, "les_convec" + sfx)(self.ctx, *[self._f(t, n, None, dt) for t, n in zip(args, names)]))

    def wallstress(self, u, v, dudz, dvdz, txz, tyz, psi_m=None, psi_m0=None, phi_m=None):
        n2 = (self.ny, self.nx)
        dev = dict(dtype=torch.float64, device="cuda")
        psi_m = torch.zeros(n2, **dev) if psi_m is None else psi_m
        psi_m0 = torch.zeros(n2, **dev) if psi_m0 is None else psi_m0
        phi_m = torch.ones(n2, **dev) if phi_m is None else phi_m
        ustar = torch.zeros(n2, **dev)
        f = [self._f(t, n) for t, n in ((u, "u"), (v, "v"), (dudz, "dudz"), (dvdz, "dvdz"), (txz, "txz"),
                                         (tyz, "tyz"))]
        p = [self._p2(t, n) for t, n in ((psi_m, "psi_m"), (psi_m0, "psi_m0"), (phi_m, "phi_m"), (ustar, "ustar"))]
        self._chk(self.L.les_wallstress(self.ctx, *f, *p))
        return ustar

    def sgs_fields(self):
        """The Lagrangian SGS arrays of sgsmodule.f90 for the routine-level
        entry points, which are FP64 (the FP32-bulk state keeps its own)."""
        d = {n: self.field() for n in SGS_NAMES}
        d["Beta_avg"] = torch.zeros(self.nz, dtype=torch.float64, device="cuda")
        d["Betaclip_avg"] = torch.zeros(self.nz, dtype=torch.float64, device="cuda")
        return d

    def _sgs_struct(self, d):
        return SgsFields(*[self._f(d[n], n) for n in SGS_NAMES],
                         _ptr(d["Beta_avg"], F64, self.nz, "Beta_avg"),
                         _ptr(d["Betaclip_avg"], F64, self.nz, "Betaclip_avg"))

    def _arr(self, ts, names):
        if len(ts) != len(names):
            raise LesError(f"expected {len(names)} arrays ({', '.join(names)}), got {len(ts)}")
        return (C.c_void_p * len(names))(*[self._f(t, n).value for t, n in zip(ts, names)])

    GRAD_NAMES = ("dudx", "dudy", "dudz", "dvdx", "dvdy", "dvdz", "dwdx", "dwdy", "dwdz")
    TAU_NAMES = ("txx", "txy", "txz", "tyy", "tyz", "tzz")
    S_NAMES = ("S11", "S12", "S13", "S22", "S23", "S33")

    def calc_sij(self, grads):
        S = [self.field() for _ in range(6)]
        self._chk(self.L.les_calc_sij(self.ctx, self._arr(grads, self.GRAD_NAMES), self._arr(S, self.S_NAMES)))
        return S

    def sgs_stag(self, grads, u, v, w, taus, sgs):
        g = self._arr(grads, self.GRAD_NAMES)
        t = self._arr(taus, self.TAU_NAMES)
        st = self._sgs_struct(sgs)
        self._chk(self.L.les_sgs_stag(self.ctx, g, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"), t,
                                      C.byref(st)))

    def lagrange_sdep(self, S, u, v, w, sgs):
        s = self._arr(S, self.S_NAMES)
        st = self._sgs_struct(sgs)
        self._chk(self.L.les_lagrange_sdep(self.ctx, s, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"),
                                           C.byref(st)))

    def interpolag_sdep(self, sgs):
        st = self._sgs_struct(sgs)
        self._chk(self.L.les_interpolag_sdep(self.ctx, C.byref(st)))

    def set_sgs_model(self, model):
        """param.f90 'model': 1 Smagorinsky with Mason's wall damping, 2 standard
        dynamic, 3 plane-averaged scale-dependent (quintic for beta), 4
        Lagrangian scale-invariant, 5 Lagrangian scale-dependent, 6
        plane-averaged scale-dependent (new_scaledep_dynamic; with scalars,
        scalar model 6: Pr_t of scalar_scaledep_dynamic)."""
        self._chk(self.L.les_set_sgs_model(self.ctx, int(model)))

    def std_dynamic(self, S, u, v, w):
        """std_dynamic.f90: the per-level Cs_opt2 as an (nz,) float64 GPU tensor."""
        s = self._arr(S, self.S_NAMES)
        cs = torch.zeros(self.nz, dtype=torch.float64, device="cuda")
        self._chk(self.L.les_std_dynamic(self.ctx, s, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"), _ptr(cs)))
        return cs

    def new_scaledep_dynamic(self, S, u, v, w):
        """new_scaledep_dynamic.f90 (model 6): the per-level Cs_opt2 as an (nz,)
        float64 GPU tensor."""
        s = self._arr(S, self.S_NAMES)
        cs = torch.zeros(self.nz, dtype=torch.float64, device="cuda")
        self._chk(self.L.les_new_scaledep_dynamic(self.ctx, s, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"),
                                                  _ptr(cs)))
        return cs

    def lagrange_ssim(self, S, u, v, w, sgs):
        s = self._arr(S, self.S_NAMES)
        st = self._sgs_struct(sgs)
        self._chk(self.L.les_lagrange_ssim(self.ctx, s, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"),
                                           C.byref(st)))

    def interpolag_ssim(self, sgs):
        st = self._sgs_struct(sgs)
        self._chk(self.L.les_interpolag_ssim(self.ctx, C.byref(st)))

    def set_forcing(self, coriolis=0, ug0=0.0, vg0=0.0, u_star=0.45, z_i=1000.0, ubc=0, damping_method=1):
        """Coriolis forcing toward the geostrophic wind (ug0, vg0) m/s and the
        top sponge (ubc=1, damping_method 1 or 2), main.f90:319-348."""
        self._forcing = _lib.Forcing(int(coriolis), float(ug0), float(vg0), float(u_star), float(z_i), int(ubc),
                                     int(damping_method))
        self._chk(self.L.les_set_forcing(self.ctx, C.byref(self._forcing)))

    def sponge(self):
        """setsponge's sponge(1..nz) as an (nz,) numpy array."""
        sp = np.zeros(self.nz)
        self._chk(self.L.les_get_sponge(self.ctx, sp.ctypes.data_as(C.c_void_p)))
        return sp

    def scaledep_dynamic(self, S, u, v, w):
        """scaledep_dynamic.f90 on the dummies S (S11..S33 in that order):
        returns the per-level Cs_opt2 and beta as (nz,) float64 GPU tensors."""
        s = self._arr(S, self.S_NAMES)
        cs = torch.zeros(self.nz, dtype=torch.float64, device="cuda")
        beta = torch.zeros(self.nz, dtype=torch.float64, device="cuda")
        self._chk(self.L.les_scaledep_dynamic(self.ctx, s, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"),
                                              _ptr(cs), _ptr(beta)))
        return cs, beta

    # ---- autotuner (launch shapes and FFT plans) ----
    def autotune(self, what=7):
        """Time column-FFT tile widths and plans (what & 1), tridiagonal
        schedules (what & 2) and the padded-product variant (what & 4) on this
        context's state and keep the fastest; returns the measured ms per
        candidate and the choice.  Tile widths, the Thomas-sweep storage and the
        padded-product variant leave the results unchanged; the FFT plan
        (radix-4 or radix-8 passes) and the parallel-cyclic-reduction schedule
        of the tridiagonal solve change the order of operations, so results
        after tuning agree with the untuned ones to rounding, not bit for bit."""
        import json
        buf = C.create_string_buffer(4096)
        self._chk(self.L.les_autotune(self.ctx, int(what), buf, len(buf)))
        return json.loads(buf.value.decode())

    def get_tune(self):
        import numpy as np
        a = np.zeros(8, dtype=np.int32)
        self._chk(self.L.les_get_tune(self.ctx, a.ctypes.data_as(C.c_void_p)))
        return dict(col_tx_fp32=int(a[0]), col_tx_fp64=int(a[1]), tri_mode=int(a[2]), tri_block=int(a[3]),
                    col_plan_fp32=int(a[4]), col_plan_fp64=int(a[5]), pad_mode_fp32=int(a[6]),
                    pad_mode_fp64=int(a[7]))

    def set_tune(self, col_tx_fp32=16, col_tx_fp64=8, tri_mode=0, tri_block=128, col_plan_fp32=0, col_plan_fp64=0,
                 pad_mode_fp32=0, pad_mode_fp64=0):
        import numpy as np
        a = np.array([col_tx_fp32, col_tx_fp64, tri_mode, tri_block, col_plan_fp32, col_plan_fp64, pad_mode_fp32,
                      pad_mode_fp64], dtype=np.int32)
        self._chk(self.L.les_set_tune(self.ctx, a.ctypes.data_as(C.c_void_p)))

    def set_layout(self, xs_tiled):
        """x-spectral intermediate layout: 0 row-major, 1 8-column tiles
        (nx a multiple of 16); results are bit-identical in both."""
        self._chk(self.L.les_set_layout(self.ctx, int(xs_tiled)))

    def get_layout(self):
        return int(self.L.les_get_layout(self.ctx))

    def set_tune_choice(self, chosen):
        """Apply the 'chosen' entry of an autotune report."""
        self.set_tune(col_tx_fp32=chosen["col_tx"][0], col_tx_fp64=chosen["col_tx"][1],
                      tri_mode=chosen["tri_mode"], tri_block=chosen["tri_block"],
                      col_plan_fp32=chosen["col_plan"][0], col_plan_fp64=chosen["col_plan"][1],
                      pad_mode_fp32=chosen["pad_mode"][0], pad_mode_fp64=chosen["pad_mode"][1])

    # ---- learned closure (no counterpart in the reference code) ----
    MLP_NIN, MLP_NOUT = 8, 2

    def mlp_init(self, seed=0):
        self._chk(self.L.les_mlp_init(self.ctx, int(seed)))
        self.mlp_np = self.L.les_mlp_nparam()

    def mlp_params(self):
        import numpy as np
        th = np.zeros(self.L.les_mlp_nparam(), dtype=np.float32)
        self._chk(self.L.les_mlp_get_params(self.ctx, th.ctypes.data_as(C.c_void_p)))
        return th

    def mlp_set_params(self, theta):
        import numpy as np
        th = np.ascontiguousarray(theta, dtype=np.float32)
        assert th.size == self.L.les_mlp_nparam()
        self._chk(self.L.les_mlp_set_params(self.ctx, th.ctypes.data_as(C.c_void_p)))

    def mlp_set_norm(self, mu, sd, ymu, ysd):
        import numpy as np
        a = [np.ascontiguousarray(x, dtype=np.float32) for x in (mu, sd, ymu, ysd)]
        self._chk(self.L.les_mlp_set_norm(self.ctx, *[x.ctypes.data_as(C.c_void_p) for x in a]))

    def mlp_norm(self):
        import numpy as np
        a = [np.zeros(n, dtype=np.float32) for n in (8, 8, 2, 2)]
        self._chk(self.L.les_mlp_get_norm(self.ctx, *[x.ctypes.data_as(C.c_void_p) for x in a]))
        return a

    @staticmethod
    def _xy(X, T, what="T"):
        P = X.shape[0] if isinstance(X, torch.Tensor) and X.dim() == 2 else -1
        if P < 0:
            raise LesError("X: expected a (P, 8) float32 tensor")
        px = _ptr(X, F32, P * 8, "X")
        pt = None if T is None else _ptr(T, F32, P * 2, what)
        return P, px, pt

    def mlp_fit_norm(self, X, Y):
        P, px, py = self._xy(X, Y, "Y")
        self._chk(self.L.les_mlp_fit_norm(self.ctx, px, py, P))

    def mlp_forward(self, X):
        """(P, 8) float32 raw features on the GPU -> (P, 2) (Cs_opt2, Beta)."""
        P, px, _ = self._xy(X, None)
        Y = torch.empty((P, 2), dtype=torch.float32, device=X.device)
        self._chk(self.L.les_mlp_forward(self.ctx, px, P, _ptr(Y)))
        self.sync()
        return Y

    def mlp_grad(self, X, T):
        P, px, pt = self._xy(X, T)
        g = torch.empty(self.L.les_mlp_nparam(), dtype=torch.float32, device=X.device)
        loss = torch.empty(1, dtype=torch.float32, device=X.device)
        self._chk(self.L.les_mlp_grad(self.ctx, px, pt, P, _ptr(g), _ptr(loss)))
        self.sync()
        return float(loss.item()), g

    def mlp_train(self, X, T, batch=4096, steps=1000, lr=1e-3, seed=0):
        P, px, pt = self._xy(X, T)
        hist = torch.empty(max(steps, 1), dtype=torch.float32, device=X.device)
        self._chk(self.L.les_mlp_train(self.ctx, px, pt, P, int(batch), int(steps),
                                       float(lr), int(seed), _ptr(hist)))
        self.sync()
        return hist[:steps]

    def set_closure(self, mode, period=10, threshold=0.2):
        self._chk(self.L.les_set_closure(self.ctx, int(mode), int(period), float(threshold)))

    def closure_judge(self):
        import numpy as np
        out = np.zeros(4)
        self._chk(self.L.les_closure_judge(self.ctx, out.ctypes.data_as(C.c_void_p)))
        return out

    # ---- scalars (scalars_module.f90, scalars_module_q.f90, lagrange_scalar_Sdep.f90) ----
    def set_scalars(self, enable=1, model=7, initsc=0, passive=0, Pr=0.55, z_i=1000.0, u_star=1.0,
                    T_scale=400.0, q_scale=100.0, g=9.81, dTdz_top=0.01, dqdz_top=-0.004, wt_s=None, wq_s=None):
        """Scalar switches/constants and the (ny, nx) kinematic surface fluxes
        wt_s (K m/s), wq_s ((g/kg) m/s) as host arrays; call before the state."""
        import numpy as np
        self.scalar_params = ScalarParams(enable, model, initsc, passive, Pr, z_i, u_star, T_scale, q_scale, g,
                                          dTdz_top, dqdz_top)
        self._wt = None if wt_s is None else np.ascontiguousarray(wt_s, dtype=np.float64)
        self._wq = None if wq_s is None else np.ascontiguousarray(wq_s, dtype=np.float64)
        pw = None if self._wt is None else self._wt.ctypes.data_as(C.c_void_p)
        pq = None if self._wq is None else self._wq.ctypes.data_as(C.c_void_p)
        self._chk(self.L.les_set_scalars(self.ctx, C.byref(self.scalar_params), pw, pq))

    def set_scalar_lag_init(self, glm_init, gqn_init):
        self._chk(self.L.les_set_scalar_lag_init(self.ctx, int(glm_init), int(gqn_init)))

    def obukhov(self, u, v, theta, qmix, sgs_t3, sgs_q3, psi_m, psi_m0, psi_h, psi_h0, phi_m, phi_h, ustar_avg):
        f = [self._f(t, n) for t, n in ((u, "u"), (v, "v"), (theta, "theta"), (qmix, "qmix"))]
        pl = [self._pl(sgs_t3, "sgs_t3"), self._pl(sgs_q3, "sgs_q3")]
        p2 = [self._p2(t, n) for t, n in ((psi_m, "psi_m"), (psi_m0, "psi_m0"), (psi_h, "psi_h"), (psi_h0, "psi_h0"),
                                          (phi_m, "phi_m"), (phi_h, "phi_h"), (ustar_avg, "ustar_avg"))]
        self._chk(self.L.les_obukhov(self.ctx, *f, *pl, *p2))

    def calcbeta(self, theta, qmix, beta):
        self._chk(self.L.les_calcbeta(self.ctx, self._f(theta, "theta"), self._f(qmix, "qmix"),
                                      self._f(beta, "beta_scal")))

    def scalar_rhs_calc(self, which, dsdx, dsdy, dsdz, u, v, w, nu_t, pr, ustar_avg, phi_h, rhs, sgs_vert1):
        f = [self._f(t, n) for t, n in ((dsdx, "dsdx"), (dsdy, "dsdy"), (dsdz, "dsdz"), (u, "u"), (v, "v"),
                                         (w, "w"), (nu_t, "Nu_t"), (pr, "Pr_"))]
        self._chk(self.L.les_scalar_rhs_calc(self.ctx, int(which), *f, self._p2(ustar_avg, "ustar_avg"),
                                             self._p2(phi_h, "phi_h"), self._f(rhs, "RHS"),
                                             self._pl(sgs_vert1, "sgs_vert1")))

    def step_scalar(self, which, scalar, rhs, rhs_f):
        self._chk(self.L.les_step_scalar(self.ctx, int(which), self._f(scalar, "scalar"), self._f(rhs, "RHS"),
                                         self._f(rhs_f, "RHS_f")))

    def scalar_sgs_fields(self, phi_h=None, psi_m=None):
        """The scalar SGS arrays of sgsmodule.f90 (FP64, as the routine-level
        entry points take them)."""
        d = {n: self.field() for n in SCALAR_SGS_NAMES}
        n2 = (self.ny, self.nx)
        d["phi_h"] = torch.ones(n2, dtype=torch.float64, device="cuda") if phi_h is None else phi_h
        d["psi_m"] = torch.zeros(n2, dtype=torch.float64, device="cuda") if psi_m is None else psi_m
        return d

    def _ssgs_struct(self, d):
        return ScalarSgsFields(*[self._f(d[n], n) for n in SCALAR_SGS_NAMES],
                               self._p2(d["phi_h"], "phi_h"), self._p2(d["psi_m"], "psi_m"))

    def interpolag_scalar_sdep(self, ssgs):
        st = self._ssgs_struct(ssgs)
        self._chk(self.L.les_interpolag_scalar_sdep(self.ctx, C.byref(st)))

    def lagrange_scalar_sdep(self, S, u, v, w, theta, sgs, ssgs):
        s = self._arr(S, self.S_NAMES)
        m, st = self._sgs_struct(sgs), self._ssgs_struct(ssgs)
        self._chk(self.L.les_lagrange_scalar_sdep(self.ctx, s, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"),
                                                  self._f(theta, "theta"), C.byref(m), C.byref(st)))

    def scalar_scaledep_dynamic(self, S, u, v, w, theta, Cs_opt2, ssgs):
        """scalar_scaledep_dynamic.f90 (model 6 with scalars): filters theta in
        place, writes ssgs["Pr_t"] on levels 1..nz-1 and the per-level Cs2_Pr
        into the first nz-1 values of ssgs["Cs2_Pr"]."""
        s = self._arr(S, self.S_NAMES)
        st = self._ssgs_struct(ssgs)
        self._chk(self.L.les_scalar_scaledep_dynamic(self.ctx, s, self._f(u, "u"), self._f(v, "v"),
                                                     self._f(w, "w"), self._f(theta, "theta"),
                                                     self._f(Cs_opt2, "Cs_opt2"), C.byref(st)))

    def sgs_stag_scalar(self, grads, u, v, w, theta, taus, sgs, ssgs):
        g = self._arr(grads, self.GRAD_NAMES)
        t = self._arr(taus, self.TAU_NAMES)
        m, st = self._sgs_struct(sgs), self._ssgs_struct(ssgs)
        self._chk(self.L.les_sgs_stag_scalar(self.ctx, g, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"),
                                             self._f(theta, "theta"), t, C.byref(m), C.byref(st)))

    def _divstress(self, name, divt, tx, ty, tz):
        sfx = self._sfx(divt)
        dt = F32 if sfx else F64
        self._chk(getattr(self.L, name + sfx)(self.ctx, self._f(divt, "divt", None, dt), self._f(tx, "tx", None, dt),
                                              self._f(ty, "ty", None, dt), self._f(tz, "tz", None, dt)))

    def divstress_uv(self, divt, tx, ty, tz):
        self._divstress("les_divstress_uv", divt, tx, ty, tz)

    def divstress_w(self, divt, tx, ty, tz):
        self._divstress("les_divstress_w", divt, tx, ty, tz)

    def press_stag_array(self, u, v, w, divtz, p, dpdx, dpdy):
        sfx = self._sfx(u)
        dt = F32 if sfx else F64
        f = [self._f(t, n, None, dt) for t, n in ((u, "u"), (v, "v"), (w, "w"), (divtz, "divtz"))]
        self._chk(getattr(self.L, "les_press_stag_array" + sfx)(
            self.ctx, *f, self._f(p, "p", self.nz + 1, dt), self._f(dpdx, "dpdx", None, dt),
            self._f(dpdy, "dpdy", None, dt)))

    def tridag_array(self, a, b, c, r, u):
        """a, b, c real (nz+1, ny, lh); r, u complex as (nz+1, ny, 2 lh) float64."""
        nr = (self.nz + 1) * self.ny * self.lh
        co = [_ptr(t, F64, nr, n) for t, n in ((a, "a"), (b, "b"), (c, "c"))]
        self._chk(self.L.les_tridag_array(self.ctx, *co, _ptr(r, F64, 2 * nr, "r"), _ptr(u, F64, 2 * nr, "u")))

    def tridag_array_noalloc(self, a, b, c, r, u):
        """tridag_array without device allocation (the Thomas factors in the
        library's per-GPU scratch array); results identical to tridag_array."""
        nr = (self.nz + 1) * self.ny * self.lh
        co = [_ptr(t, F64, nr, n) for t, n in ((a, "a"), (b, "b"), (c, "c"))]
        self._chk(self.L.les_tridag_array_noalloc(self.ctx, *co, _ptr(r, F64, 2 * nr, "r"),
                                                  _ptr(u, F64, 2 * nr, "u")))

    def project(self, u, v, w, dpdx, dpdy, dpdz):
        self._chk(self.L.les_project(self.ctx, *[self._f(t, n) for t, n in ((u, "u"), (v, "v"), (w, "w"),
                                                                            (dpdx, "dpdx"), (dpdy, "dpdy"),
                                                                            (dpdz, "dpdz"))]))

    def rmsdiv(self, dudx, dvdy, dwdz):
        r = C.c_double()
        self._chk(self.L.les_rmsdiv(self.ctx, self._f(dudx, "dudx"), self._f(dvdy, "dvdy"), self._f(dwdz, "dwdz"),
                                    C.byref(r)))
        return r.value

    # ---- statistics (io.f90 spectrum / avg_stats, obukhov's plane averages) ----
    def stats_len(self):
        return stats_len(self.nz, self.nx)

    def spectrum(self, f):
        """spectrum (io.f90:1613-1646) of every level of f: (nlev, nx/2) float64 GPU tensor."""
        n = self._nlev(f)
        spec = torch.zeros((n, self.nx // 2), dtype=F64, device="cuda")
        self._chk(self.L.les_spectrum(self.ctx, self._f(f, "f", n), n, _ptr(spec)))
        return spec

    def avg_stats(self, u, v, w, txz):
        """Current statistics of avg_stats (io.f90:1722-1755) as a float64 GPU
        tensor in the layout of stats_fields."""
        out = torch.zeros(self.stats_len(), dtype=F64, device="cuda")
        self._chk(self.L.les_avg_stats(self.ctx, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"),
                                       self._f(txz, "txz"), _ptr(out)))
        return out

    def surface_stats(self, u, v, txz, tyz, psi_m, psi_m0, ustar_avg, theta=None, qmix=None, sgs_t3=None,
                      sgs_q3=None):
        """obukhov's plane averages (scalars_module.f90:545-561,655-661) from
        first-level fields: float64 GPU tensor (8,) in the order of SURFACE_NAMES."""
        out = torch.zeros(8, dtype=F64, device="cuda")
        sc = [theta, qmix, sgs_t3, sgs_q3]
        if any(x is None for x in sc) and not all(x is None for x in sc):
            raise LesError("surface_stats: theta, qmix, sgs_t3, sgs_q3 go together")
        scp = [None] * 4 if sc[0] is None else [self._pl(theta, "theta"), self._pl(qmix, "qmix"),
                                                 self._pl(sgs_t3, "sgs_t3"), self._pl(sgs_q3, "sgs_q3")]
        self._chk(self.L.les_surface_stats(self.ctx, self._pl(u, "u"), self._pl(v, "v"), self._pl(txz, "txz"),
                                           self._pl(tyz, "tyz"), self._p2(psi_m, "psi_m"),
                                           self._p2(psi_m0, "psi_m0"), self._p2(ustar_avg, "ustar_avg"), *scp,
                                           _ptr(out)))
        return out

    def timestep_conditions(self, u, v, w, nu_t):
        out = (C.c_double * 4)()
        self._chk(self.L.les_timestep_conditions(self.ctx, self._f(u, "u"), self._f(v, "v"), self._f(w, "w"),
                                                 self._f(nu_t, "Nu_t"), out))
        return list(out)


class LesSim(Les):
    """Whole-step driver on context-owned device state (main.f90:202-415)."""

    def __init__(self, nx, ny, nz, precision="fp64", graph=False, scalars=None, sgs_model=5, forcing=None, **kw):
        """scalars: None (neutral) or a dict of set_scalars arguments (the
        convective step with theta, q, buoyancy and the scalar SGS model).
        sgs_model: param.f90 'model', 1..6 (set_sgs_model).  forcing: None or a
        dict of set_forcing arguments (Coriolis, sponge)."""
        # a stream of its own (default priority): graph capture needs a
        # non-default stream, and the state is only touched through the library
        self.stream = torch.cuda.Stream()
        super().__init__(nx, ny, nz, stream=self.stream, **kw)
        if sgs_model != 5:
            self.set_sgs_model(sgs_model)
        if scalars is not None:
            self.set_scalars(**scalars)
        if forcing is not None:
            self.set_forcing(**forcing)
        self.precision = precision
        self._chk(self.L.les_state_alloc(self.ctx, 1 if precision == "fp32" else 0))
        self._chk(self.L.les_set_graph(self.ctx, 1 if graph else 0))

    def set_graph(self, enable):
        self._chk(self.L.les_set_graph(self.ctx, 1 if enable else 0))

    def mlp_harvest(self, labels=True):
        """Features (P, 8) and labels (P, 2) of the last step, which must have
        run the Lagrangian update; P = nx*ny*(nz-1)."""
        P = self.nx * self.ny * (self.nz - 1)
        X = torch.empty((P, 8), dtype=torch.float32, device="cuda")
        Y = torch.empty((P, 2), dtype=torch.float32, device="cuda") if labels else None
        self._chk(self.L.les_mlp_harvest(self.ctx, _ptr(X, F32, P * 8, "X"), _ptr(Y, F32, P * 2, "Y")))
        self.sync()
        return X, Y

    def graph_prepare(self):
        self._chk(self.L.les_graph_prepare(self.ctx))

    def init_ic(self, seed=42, amp=0.1, ustar=1.0, kappa_ic=0.4):
        self._chk(self.L.les_state_init_ic(self.ctx, seed, amp, ustar, kappa_ic))

    def init_scalar_ic(self, seed=42, T_init=283.0, z_inv=0.8, inv_strength=0.01, amp_K=0.1, q_init=1.0):
        """Convective scalar IC (after init_ic; needs set_scalars before the state)."""
        self._chk(self.L.les_state_init_scalar_ic(self.ctx, seed, T_init, z_inv, inv_strength, amp_K, q_init))

    def step(self, n=1):
        self._chk(self.L.les_step(self.ctx, n))

    def ptr(self, name):
        nlev, es = C.c_int(), C.c_int()
        p = self.L.les_state_field(self.ctx, name.encode(), C.byref(nlev), C.byref(es))
        if not p:
            raise KeyError(name)
        return p, nlev.value, es.value

    def view(self, name):
        """Zero-copy torch view of a state field (reference layout)."""
        p, nlev, es = self.ptr(name)
        if name in ("psi_m", "psi_m0", "phi_m", "ustar_avg", "psi_h", "psi_h0", "phi_h"):
            shape, typestr = (self.ny, self.nx), "<f8"
        elif name in ("sgs_t3", "sgs_q3"):
            shape, typestr = (self.ny, self.ld), "<f8"
        elif name in ("Beta_avg", "Betaclip_avg"):
            shape, typestr = (self.nz,), "<f8"
        else:
            shape, typestr = (nlev, self.ny, self.ld), ("<f4" if es == 4 else "<f8")

        class _Dev:
            __cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (p, False),
                                        "version": 2, "strides": None}
        return torch.as_tensor(_Dev(), device="cuda")

    def get(self, name):
        """Copy a state field to a host numpy array (float64).  theta and qmix
        are returned in the reference convention (the FP32-bulk state stores
        deviations from a reference value)."""
        self.sync()
        a = self.view(name).double().cpu().numpy()
        if name in ("theta", "qmix") and self.precision == "fp32":
            tr, qr = C.c_double(), C.c_double()
            self._chk(self.L.les_scalar_ref(self.ctx, C.byref(tr), C.byref(qr)))
            a[:, :, :self.nx] += tr.value if name == "theta" else qr.value
        return a

    @property
    def jt(self):
        a, b = C.c_int(), C.c_int()
        self._chk(self.L.les_get_time(self.ctx, C.byref(a), C.byref(b)))
        return a.value

    # ---- statistics of the state (FP64 reductions on the device) ----
    def stats(self):
        """Current statistics of the state: (stats, surface) host float64
        arrays; stats_fields() names the parts."""
        import numpy as np
        st, sf = np.zeros(self.stats_len()), np.zeros(8)
        self._chk(self.L.les_state_stats(self.ctx, st.ctypes.data_as(C.c_void_p), sf.ctypes.data_as(C.c_void_p)))
        return st, sf

    def checkpoint(self, path):
        """Write the state as io.f90's checkpoint record (vel_sc.out layout
        with scalars, vel.out without)."""
        self._chk(self.L.les_state_checkpoint(self.ctx, _b(path)))

    def read_initial(self, path, model=5, initsc=0, inilag=0):
        """Read an initial-condition record with initial.f90's list for
        model / initsc / inilag into the state."""
        self._chk(self.L.les_state_read_initial(self.ctx, _b(path), int(model), int(initsc), int(inilag)))

    def stats_accumulate(self):
        """Fold the current statistics into running means kept on the device."""
        self._chk(self.L.les_state_stats_accumulate(self.ctx))

    def stats_average(self, reset=False):
        """Running means since the last reset: (stats, surface, samples)."""
        import numpy as np
        st, sf, n = np.zeros(self.stats_len()), np.zeros(8), C.c_int()
        self._chk(self.L.les_state_stats_average(self.ctx, st.ctypes.data_as(C.c_void_p),
                                                 sf.ctypes.data_as(C.c_void_p), C.byref(n), int(bool(reset))))
        return st, sf, n.value


# ---- run files (host only: no GPU needed) ----
def _b(path):
    return str(path).encode()


def record_write(path, arrays, append=False):
    """One Fortran sequential unformatted record holding the arrays (numpy,
    in order), as a Fortran WRITE of that list would."""
    import numpy as np
    arrs = [np.ascontiguousarray(a) for a in arrays]
    n = len(arrs)
    ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
    sizes = (C.c_uint64 * n)(*[a.nbytes for a in arrs])
    if _lib.load().les_record_write(_b(path), int(append), n, ptrs, sizes) != 0:
        raise LesError(f"record_write failed: {path}")


def record_read(path, index=0):
    """Bytes of record `index` of a Fortran sequential unformatted file."""
    L = _lib.load()
    nb = C.c_uint64()
    if L.les_record_read(_b(path), int(index), None, 0, C.byref(nb)) != 0:
        raise LesError(f"record_read: no record {index} in {path}")
    buf = C.create_string_buffer(max(nb.value, 1))
    L.les_record_read(_b(path), int(index), buf, nb.value, C.byref(nb))
    return buf.raw[:nb.value]


def read_total_time(path):
    """jt_total of total_time.dat, or 0 if the file is missing (param.f90:200-210)."""
    v = C.c_int()
    rc = _lib.load().les_read_total_time(_b(path), C.byref(v))
    if rc == 1:
        return 0
    if rc != 0:
        raise LesError(f"read_total_time: malformed {path}")
    return v.value


def write_total_time(path, jt_total):
    if _lib.load().les_write_total_time(_b(path), int(jt_total)) != 0:
        raise LesError(f"write_total_time failed: {path}")


def read_surface(directory, nx, ny, z_i=1000.0):
    """wts.dat, wqs.dat, zom.dat of a run directory (bottombc.f90:23-57):
    (wt_s K m/s, wq_s (g/kg) m/s, zo in units of z_i), each (ny, nx)."""
    import numpy as np
    out = [np.zeros((ny, nx)) for _ in range(3)]
    rc = _lib.load().les_read_surface(_b(directory), nx, ny, float(z_i), *[a.ctypes.data for a in out])
    if rc != 0:
        raise LesError(f"read_surface failed ({rc}) in {directory}")
    return tuple(out)


def read_param_nml(nml_path, total_time_path=None):
    """param_nml of input.nml with the reference's defaults and restart rule
    (param.f90:174-230) as a dict; status 3 (an unknown item) is returned
    beside it."""
    from ._lib import ParamNml
    p = ParamNml()
    rc = _lib.load().les_read_param_nml(None if nml_path is None else _b(nml_path),
                                        None if total_time_path is None else _b(total_time_path), C.byref(p))
    if rc not in (0, 3):
        raise LesError(f"read_param_nml failed ({rc}): {nml_path}")
    return {f[0]: getattr(p, f[0]) for f in p._fields_}, rc


def read_nml_value(path, group, key):
    v = C.c_double()
    rc = _lib.load().les_read_nml_value(_b(path), group.encode(), key.encode(), C.byref(v))
    if rc != 0:
        raise LesError(f"read_nml_value: {group}/{key} not found in {path} ({rc})")
    return v.value


SURFACE_NAMES = ("ustar_", "ustar_avg_time", "ustar_time", "theta_", "thetav_", "q_time", "wt_", "L_")
STATS_NAMES = ("ubar", "vbar", "wbar", "u2", "v2", "w2", "uw", "txz", "stress")


def stats_len(nz, nx):
    """Length of the statistics vector of a grid: 9 (nz-1) + (nz-1) nx/2."""
    return 9 * (nz - 1) + (nz - 1) * (nx // 2)


def stats_fields(a, nz, nx):
    """Named views of a statistics vector: the nine plane profiles (nz-1
    levels each; u2, v2, w2, uw on w-nodes, zero on the first level; stress =
    uw + txz) and Eoz, E11(k1)/z as (nz-1, nx/2)."""
    nl = nz - 1
    d = {n: a[i * nl:(i + 1) * nl] for i, n in enumerate(STATS_NAMES)}
    d["Eoz"] = a[9 * nl:].reshape(nl, nx // 2)
    return d


def checkerboard_flux(nx, ny, hi=0.1, lo=0.02, patches=4):
    """(ny, nx) surface heat flux of patches x patches alternating tiles
    (kinematic, K m/s): the heterogeneous surface of the convective case."""
    import numpy as np
    px = (np.arange(nx) * patches) // nx
    py = (np.arange(ny) * patches) // ny
    return np.where((px[None, :] + py[:, None]) % 2 == 0, hi, lo).astype(np.float64)


class LesEnsemble:
    """Several LesSim instances (any grids, surface parameters, precisions)
    advanced together: one HIP graph per window of steps for all of them.
    Each member's results are bit-identical to stepping it alone."""

    def __init__(self, members):
        self.members = list(members)
        self.L = _lib.load()
        arr = (C.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        h = C.c_void_p()
        rc = self.L.les_ensemble_create(arr, len(self.members), C.byref(h))
        if rc != 0:
            raise LesError(f"les_ensemble_create failed ({rc}): members need a state and one device")
        self.h = h

    @property
    def window(self):
        return self.L.les_ensemble_window(self.h)

    def step(self, n=1):
        if self.L.les_ensemble_step(self.h, int(n)) != 0:
            raise LesError(self.L.les_ensemble_last_error(self.h).decode())

    def sync(self):
        if self.L.les_ensemble_sync(self.h) != 0:
            raise LesError("les_ensemble_sync failed")

    def close(self):
        if getattr(self, "h", None):
            self.L.les_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
