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
# in its last bits (a "flip" moves a rounded value by one ulp, at most 2^-23
# of it; measured in ulps of a result whose own ulp may be as small as 2^-24
# of it, each flip can count twice):
#   1  a value rounded from FP64 inputs that carry no flip (F_*, Tn*, G_*);
#   2  an FP64 product with one rounded value (Nu_t, tau: Cs_opt2 |S| Delta^2);
#   4  an FP64 quotient of two rounded values (Beta = Cs2_4d / Cs2_2d), or a
#      rounded quotient of a rounded value by a rounded quotient of rounded
#      values (Cs_opt2 = Cs2_2d / max(real(Beta), 1/8), Cs2_Pr likewise);
#   8  an FP64 quotient of two such values (Pr_t = Cs_opt2 / Cs2_Pr).
FLIP_ULPS = dict(F_LM=1, F_MM=1, F_QN=1, F_NN=1, TnLMMM=1, TnQNNN=1, G_LM=1, G_MM=1, G_QN=1, G_NN=1,
                 Nu_t=2, tau=2, Beta=4, Cs_opt2=4, Cs2_Pr=4, Pr_t=8)


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
