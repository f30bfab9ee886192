import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) GPU")


def relerr(a, b):
    """max|a-b| / max|b| (interior values)."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    if den == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)


def float_ulps(a, b):
    """|a-b| in units of one single-precision ulp of the larger magnitude."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    m = np.maximum(np.abs(a), np.abs(b)).astype(np.float32)
    ulp = np.spacing(m).astype(np.float64)
    return np.abs(a - b) / ulp


# Largest float-ulp difference of an element whose value passes through the
# single-precision rounding points of lagrange_Sdep.f90:321-421 and
# lagrange_scalar_Sdep.f90:324-388, where an FP64 input of a rounding differs
# in its last bits.  A flip moves a rounded value by one ulp, at most 2^-23 of
# it; a result's own ulp can be as small as 2^-24 of it, so a relative change
# of m 2^-23 counts up to 2m ulps, plus one for the result's own rounding:
#   1  a value rounded from FP64 inputs that carry no flip (F_*, Tn*, G_*);
#   2  an FP64 product with one rounded value (Nu_t, tau: Cs_opt2 |S| Delta^2);
#   4  an FP64 quotient of two rounded values (Beta = Cs2_4d / Cs2_2d: 2 2^-23);
#   8  a rounded quotient of a rounded value by the rounded Beta
#      (Cs_opt2 = real(Cs2_2d / max(real(Beta), 1/8)), Cs2_Pr likewise:
#      2^-23 + 2 2^-23 + 2^-24 before the last rounding);
#  16  the FP64 quotient of two such values (Pr_t = Cs_opt2 / Cs2_Pr).
FLIP_ULPS = dict(F_LM=1, F_MM=1, F_QN=1, F_NN=1, TnLMMM=1, TnQNNN=1, G_LM=1, G_MM=1, G_QN=1, G_NN=1,
                 Nu_t=2, tau=2, Beta=4, Cs_opt2=8, Cs2_Pr=8, Pr_t=16)

# After several steps the FP64 states differ in their last bits everywhere,
# and more single-precision roundings of the Lagrangian update flip than in a
# one-call test with identical inputs (values near a zero crossing differ in
# relative terms by far more than the fields' 1e-13): up to 0.1% of elements
# at a few float ulps (~1e-7) each make a relative L2 of a few 1e-9.
MULTI_STEP_L2 = 5e-9


def buoyancy_tol(beta, g_hat, ulps=4):
    """Bound for beta_scal (calcbeta, scalars_module.f90:175-228) relative to
    its largest magnitude: beta = g_hat (theta_v - <theta_v>) / <theta_v>
    with g_hat = g z_i / u*^2 ~ 1e4, and theta_v has passed through the FFT
    round trip of filt_da, which leaves a few ulps; `ulps` of theta_v move
    beta by ulps * eps * g_hat.  Never below 1e-12."""
    return max(1e-12, ulps * np.finfo(np.float64).eps * g_hat / float(np.max(np.abs(beta))))


def flip_check(a, b, what="", rel=1e-12, frac=1e-3, max_ulps=1.0, l2=1e-10):
    """FP64 parity of a field that passes through single-precision rounding
    points (lagrange_Sdep.f90:321-421, lagrange_scalar_Sdep.f90:324-388):
    where the two FP64 inputs of such a rounding differ in their last bits the
    float result can flip by one ulp.  Every element must match to `rel`
    relative to the field's largest magnitude, except at most a fraction
    `frac`, each of which may differ by no more than `max_ulps` float ulps of
    its value; the relative L2 error over all elements must be <= `l2`.
    Returns (fraction beyond rel, max float ulps among them, relative L2)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    d = np.abs(a - b)
    bad = d > rel * scale
    nb = np.linalg.norm(b.ravel())
    rl2 = float(np.linalg.norm((a - b).ravel()) / nb) if nb > 0 else float(np.linalg.norm(d.ravel()))
    worst = float(np.max(float_ulps(a[bad], b[bad]))) if bad.any() else 0.0
    print(f"  {what}: {int(bad.sum())}/{bad.size} beyond {rel:g}, max {worst:.2f} float ulps, "
          f"rel L2 {rl2:.2e}", flush=True)
    assert bad.mean() <= frac, (what, "fraction beyond rel", float(bad.mean()))
    if max_ulps is not None:   # None: an FP64 field downstream of flipped coefficients
        assert worst <= max_ulps * (1 + 1e-9), (what, "float ulps", worst)
    assert rl2 <= l2, (what, "relative L2", rl2)
    return float(bad.mean()), worst, rl2


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch
