"""The autotuner's launch shapes: every column-FFT tile width and tridiagonal
schedule gives bit-identical states (the per-element arithmetic does not
depend on them); the radix-8 column plan and the streaming padded products
are other orders of the same DFTs (FP64 states agree to rounding);
les_autotune picks a measured candidate."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "p", "RHSx", "Cs_opt2", "Nu_t", "txz")


def run(prec, tune=None, steps=12, conv=False, autotune=False, n=32, ny=None, nz=17):
    from les_amd.api import LesSim, checkerboard_flux
    ny = ny or n
    kw = dict(scalars=dict(wt_s=checkerboard_flux(n, ny), dTdz_top=0.01)) if conv else {}
    s = LesSim(n, ny, nz, precision=prec, dt=2e-4, **kw)
    s.init_ic(seed=3)
    if conv:
        s.init_scalar_ic(seed=3)
    rep = None
    if tune:
        s.set_tune(**tune)
    if autotune:
        rep = s.autotune(7)
    s.step(steps)
    return s, rep


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_launch_shapes_do_not_change_results(torch_mod, prec):
    print("tune shapes", prec, flush=True)
    base, _ = run(prec, conv=True)
    ref = {n: base.get(n) for n in FIELDS + ("theta",)}
    txs = (8, 32) if prec == "fp32" else (4, 16)
    shapes = [dict(col_tx_fp32=tx if prec == "fp32" else 16, col_tx_fp64=tx if prec == "fp64" else 8)
              for tx in txs]
    shapes += [dict(tri_mode=m, tri_block=b) for (m, b) in ((0, 64), (0, 256), (1, 128), (2, 32))]
    for sh in shapes:
        s, _ = run(prec, tune=sh, conv=True)
        assert s.get_tune()["tri_mode"] == sh.get("tri_mode", 0)
        for n, a in ref.items():
            assert np.array_equal(s.get(n), a), (sh, n)


def test_autotune_chooses_a_measured_candidate(torch_mod):
    print("autotune", flush=True)
    base, _ = run("fp32", steps=15)
    s, rep = run("fp32", steps=15, autotune=True)
    print(rep)
    ch = rep["chosen"]
    t = s.get_tune()
    assert f"{t['col_tx_fp32']}/{t['col_plan_fp32']}" in rep["col"]
    assert f"{t['tri_mode']}/{t['tri_block']}" in rep["tridiag"]
    assert ch["tri_mode"] == t["tri_mode"] and ch["tri_block"] == t["tri_block"]
    assert str(t["pad_mode_fp32"]) in rep["pad"] and ch["pad_mode"][0] == t["pad_mode_fp32"]
    assert all(v > 0 for k in ("col", "tridiag", "pad") for v in rep[k].values())
    if t["col_plan_fp32"] == 0 and t["pad_mode_fp32"] == 0 and t["tri_mode"] not in (3, 4):
        for n in FIELDS:
            assert np.array_equal(s.get(n), base.get(n)), n


@pytest.mark.parametrize("prec,tx", [("fp64", 4), ("fp64", 8), ("fp32", 8), ("fp32", 16)])
def test_radix8_column_plan(torch_mod, prec, tx):
    """The radix-8 plan changes only the rounding of the column transforms:
    after 12 convective steps the FP64 state agrees with the radix-4 plan to
    1e-12 relative; the FP32 state to FP32 round-off, which for v and w is
    set by u ~ 20 u* (the FP32-bulk error of v and w against FP64 after 10
    steps is 3e-5 .. 7e-5 of their own norms), so 2e-4."""
    print("radix-8 plan", prec, tx, flush=True)
    base, _ = run(prec, conv=True)
    sh = dict(col_tx_fp32=tx, col_plan_fp32=2) if prec == "fp32" else dict(col_tx_fp64=tx, col_plan_fp64=2)
    s, _ = run(prec, tune=sh, conv=True)
    tol = 1e-12 if prec == "fp64" else 2e-4
    for n in ("u", "v", "w", "p", "theta"):
        a, b = s.get(n)[..., :32], base.get(n)[..., :32]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err < tol, (n, err)


@pytest.mark.parametrize("prec,conv,n,ny,nz,steps,mode", [
    ("fp64", True, 64, 64, 17, 8, 1), ("fp32", True, 64, 64, 17, 8, 1), ("fp64", False, 64, 64, 17, 8, 1),
    ("fp32", False, 64, 64, 17, 8, 1), ("fp64", True, 256, 256, 17, 3, 1), ("fp32", False, 256, 256, 17, 3, 1),
    ("fp64", True, 96, 64, 17, 6, 1), ("fp32", True, 96, 64, 17, 6, 1),
    ("fp32", True, 256, 256, 17, 3, 2), ("fp64", True, 256, 256, 17, 3, 2), ("fp32", True, 256, 128, 13, 3, 2),
    ("fp64", True, 256, 128, 14, 2, 2), ("fp32", True, 256, 128, 14, 2, 2), ("fp32", True, 256, 128, 6, 3, 2)])
def test_streaming_padded_products(torch_mod, prec, conv, n, ny, nz, steps, mode):
    """Streaming level runs (radix-8 plan; mode 2: one transform per
    wavefront, cz of two levels in one transform) against level groups
    (two-pass plan) for the padded products, on 64^2, 256^2, 256x128 and
    96x64 grids (padded lengths 96, 384 and 144; the 384 FP64 ring is the
    largest LDS tile): momentum with both scalars, and momentum alone.
    nz = 13, 14 and 6 give level runs of odd length (3), of one level and of
    two levels.
    Tolerances as for the radix-8 column plan."""
    print("streaming padded products", prec, conv, n, ny, nz, mode, flush=True)
    key = "pad_mode_" + prec
    base, _ = run(prec, conv=conv, n=n, ny=ny, nz=nz, steps=steps)
    s, _ = run(prec, tune={key: mode}, conv=conv, n=n, ny=ny, nz=nz, steps=steps)
    assert s.get_tune()[key] == mode
    tol = 1e-12 if prec == "fp64" else 2e-4
    names = ("u", "v", "w", "p") + (("theta",) if conv else ())
    for nm in names:
        a, b = s.get(nm)[..., :n], base.get(nm)[..., :n]
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(nm, err)
        assert err < tol, (nm, err)
    assert any(not np.array_equal(s.get(nm), base.get(nm)) for nm in names)  # the streaming kernel ran
    if mode == 2:  # the wavefront kernel ran, not the streaming kernel of mode 1
        s1, _ = run(prec, tune={key: 1}, conv=conv, n=n, ny=ny, nz=nz, steps=steps)
        assert any(not np.array_equal(s.get(nm), s1.get(nm)) for nm in names)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("mode,block", [(3, 64), (4, 32), (4, 64)])
def test_other_elimination_orders_pressure_solve(torch_mod, prec, mode, block):
    """Tridiagonal schedules 3 (parallel cyclic reduction, one wavefront per
    mode) and 4 (two-ended Thomas sweep, four threads per mode) against the
    one-ended sweep: after 12 convective steps the FP64 state agrees to 1e-12
    relative (the z systems are diagonally dominant except for the boundary
    rows), the FP32-bulk state to FP32 round-off (2e-4, as for the transform
    plans); on 32^2 and 64^2 grids (nz = 17: the two ends meet at row 8)."""
    print("elimination order", mode, block, prec, flush=True)
    tol = 1e-12 if prec == "fp64" else 2e-4
    for n in (32, 64):
        base, _ = run(prec, conv=True, n=n)
        s, _ = run(prec, tune=dict(tri_mode=mode, tri_block=block), conv=True, n=n)
        assert s.get_tune()["tri_mode"] == mode
        for nm in ("u", "v", "w", "p", "theta"):
            a, b = s.get(nm)[..., :n], base.get(nm)[..., :n]
            err = np.max(np.abs(a - b)) / np.max(np.abs(b))
            print(n, nm, err)
            assert err < tol, (n, nm, err)


def test_set_tune_rejects_bad_shapes(torch_mod):
    from les_amd.api import Les, LesError
    g = Les(16, 16, 9)
    with pytest.raises(LesError):
        g.set_tune(col_tx_fp32=12)
    with pytest.raises(LesError):
        g.set_tune(tri_mode=2, tri_block=1024)   # (nz+1) x 1024 x 24 B does not fit in LDS
    with pytest.raises(LesError):
        g.set_tune(tri_mode=5, tri_block=64)
    with pytest.raises(LesError):
        g.set_tune(tri_mode=4, tri_block=512)   # four threads per mode: at most 256 modes per block
    with pytest.raises(LesError):
        g.set_tune(pad_mode_fp32=3)


def run_layout(prec, layout, tune=None, n=32, ny=None, nz=17, steps=12, smodel=5):
    from les_amd.api import LesSim, checkerboard_flux
    ny = ny or n
    s = LesSim(n, ny, nz, precision=prec, dt=2e-4, sgs_model=6 if smodel == 6 else 5,
               scalars=dict(wt_s=checkerboard_flux(n, ny), dTdz_top=0.01, model=smodel))
    s.init_ic(seed=3)
    s.init_scalar_ic(seed=3)
    if tune:
        s.set_tune(**tune)
    s.set_layout(layout)
    assert s.get_layout() == layout
    s.step(steps)
    return s


@pytest.mark.parametrize("prec,n,ny,tune,smodel", [
    ("fp32", 32, 32, None, 7), ("fp64", 32, 32, None, 6),
    ("fp32", 48, 32, dict(col_tx_fp32=8), 5), ("fp32", 32, 48, dict(col_tx_fp32=32), 7),
    ("fp32", 32, 48, dict(col_tx_fp32=16, col_plan_fp32=2), 6),
    ("fp64", 48, 32, dict(col_tx_fp64=4), 7), ("fp64", 32, 32, dict(col_tx_fp64=16), 5),
    ("fp64", 48, 48, dict(col_tx_fp64=8, col_plan_fp64=2), 7),
    ("fp32", 256, 256, dict(col_tx_fp32=8, col_plan_fp32=2, pad_mode_fp32=2), 7),
    ("fp64", 256, 128, dict(col_tx_fp64=8, pad_mode_fp64=1), 7),
])
def test_tiled_spectral_layout_is_bit_identical(torch_mod, prec, n, ny, tune, smodel):
    """The tiled x-spectral layout moves only addresses: twelve convective
    steps (two dynamic updates; filt_da's kept spectra feeding the zero-pad
    pass of the convective products, divstress; every column tile width and
    plan) give bit-identical states."""
    print("tiled layout", prec, n, ny, tune, smodel, flush=True)
    a = run_layout(prec, 0, tune, n, ny, smodel=smodel)
    b = run_layout(prec, 1, tune, n, ny, smodel=smodel)
    for name in FIELDS + ("theta", "qmix", "RHS_T"):
        assert np.array_equal(a.get(name), b.get(name)), name


def test_tiled_layout_needs_nx_multiple_of_16(torch_mod):
    from les_amd.api import LesSim
    s = LesSim(24, 24, 9, precision="fp32", dt=2e-4)
    with pytest.raises(RuntimeError):
        s.set_layout(1)
    with pytest.raises(RuntimeError):
        s.set_layout(2)
    assert s.get_layout() == 0
