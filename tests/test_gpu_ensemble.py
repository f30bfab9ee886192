"""Ensembles: independent instances with uneven grids, surface parameters,
precisions and physics advanced together (one HIP graph per window for all
members).  The pin: every member's state is bit-identical to the same
instance stepped alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "RHSx", "Cs_opt2", "F_LM", "Nu_t", "txz")


def make(spec):
    from les_amd.api import LesSim
    nx, ny, nz, prec, z0, seed, kind, cs = spec
    kw = dict(precision=prec, z0=z0, dt=2e-4, cs_count=cs)
    if kind == "stable":  # surface cooling: a stably stratified instance
        kw["scalars"] = dict(wt_s=np.full((ny, nx), -0.01), dTdz_top=0.003)
    s = LesSim(nx, ny, nz, **kw)
    s.init_ic(seed=seed)
    if kind == "stable":
        s.init_scalar_ic(seed=seed, amp_K=0.05)
    return s


SPECS = [
    (32, 32, 17, "fp64", 1e-4, 1, "neutral", 5),
    (48, 32, 13, "fp32", 1e-3, 2, "neutral", 5),
    (32, 48, 17, "fp32", 3e-5, 3, "stable", 5),
    (64, 64, 9, "fp64", 1e-5, 4, "neutral", 5),
]


def compare(ens_members, specs, nsteps, extra=()):
    singles = [make(sp) for sp in specs]
    for s, sp in zip(singles, specs):
        for (i, jt, jtt) in extra:
            if specs[i] is sp:
                s.set_time(jt, jtt)
        s.step(nsteps)
    for i, (a, b) in enumerate(zip(ens_members, singles)):
        assert a.jt == b.jt, i
        names = FIELDS + (("theta",) if specs[i][6] == "stable" else ())
        for n in names:
            assert np.array_equal(a.get(n), b.get(n)), (i, n)


def test_ensemble_bit_identical_to_single_runs(torch_mod):
    print("ensemble uneven grids", flush=True)
    from les_amd.api import LesEnsemble
    members = [make(sp) for sp in SPECS]
    ens = LesEnsemble(members)
    assert ens.window == 10
    ens.step(7)      # eager (members not yet warmed up)
    ens.step(33)     # then captured windows of 10
    ens.sync()
    compare(members, SPECS, 40)


def test_ensemble_with_one_off_window_and_mixed_periods(torch_mod):
    """One member crosses jt_total = 500 (the constant-Cs switch) and one has
    cs_count = 3 (window 6): the ensemble window is 30 and the window holding
    the switch runs eagerly."""
    print("ensemble one-off", flush=True)
    from les_amd.api import LesEnsemble
    specs = [SPECS[0], (32, 32, 9, "fp32", 2e-4, 5, "neutral", 3), SPECS[1]]
    members = [make(sp) for sp in specs]
    members[2].set_time(0, 440)
    ens = LesEnsemble(members)
    assert ens.window == 30
    ens.step(95)
    ens.sync()
    compare(members, specs, 95, extra=[(2, 0, 440)])
