"""Checkpoint and initial-condition records of the device state
(io.f90:1314-1338, initial.f90:46-175): the record's layout field by field
against the state, a restart from it that continues bit-identically, the
model-7 read list that does not match the checkpoint's layout, and FP32-bulk
states widened to real(8)."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _segments(raw, sizes):
    out, off = [], 0
    for n in sizes:
        out.append(np.frombuffer(raw[off:off + 8 * n], dtype="<f8"))
        off += 8 * n
    return out, off


def _conv_sim(prec, nx=32, ny=32, nz=17, **kw):
    from les_amd.api import LesSim, checkerboard_flux
    s = LesSim(nx, ny, nz, precision=prec, scalars=dict(wt_s=checkerboard_flux(nx, ny), dTdz_top=0.01), **kw)
    s.init_ic(seed=4)
    s.init_scalar_ic(seed=4)
    return s


NEUTRAL = ("u", "v", "w", "RHSx", "RHSy", "RHSz", "Cs_opt2", "F_LM", "F_MM", "F_QN", "F_NN")


def test_checkpoint_layout_and_restart_neutral(torch_mod, tmp_path):
    """vel.out layout (11 fields of (ld, ny, nz) real(8)); a second instance
    that reads it with model 5, no scalars, inilag false (the reference's
    continued run) and continues at jt_total = 10 is bit-identical to the
    first one continuing, because the record holds everything the step
    carries over at a Lagrangian update (u_lag is zero there)."""
    print("checkpoint neutral", flush=True)
    from les_amd.api import LesSim
    nx, ny, nz = 32, 32, 17
    a = LesSim(nx, ny, nz, precision="fp64")
    a.init_ic(seed=6)
    a.step(10)
    path = tmp_path / "vel.out"
    a.checkpoint(path)
    raw = path.read_bytes()
    n3 = a.ld * ny * nz
    body = 11 * n3 * 8
    assert len(raw) == body + 8
    assert struct.unpack("<i", raw[:4])[0] == body and struct.unpack("<i", raw[-4:])[0] == body
    segs, _ = _segments(raw[4:-4], [n3] * 11)
    for name, seg in zip(NEUTRAL, segs):
        assert np.array_equal(seg.reshape(nz, ny, a.ld), a.get(name)), name
    b = LesSim(nx, ny, nz, precision="fp64", inilag=0)
    b.read_initial(path, model=5, initsc=0, inilag=0)
    b.set_time(0, 10)
    for name in NEUTRAL:
        assert np.array_equal(b.get(name), a.get(name)), name
    a.step(10)
    b.step(10)
    for name in ("u", "v", "w", "p", "RHSx", "F_LM", "F_QN", "Cs_opt2", "Nu_t"):
        assert np.array_equal(a.get(name), b.get(name)), name


def test_checkpoint_scalars_and_model7_read(torch_mod, tmp_path):
    """vel_sc.out layout (20 fields, the two surface-flux planes, four (nx, ny)
    stability arrays); read back with the model-5 list it restores every
    field, with the model-7 list (initial.f90:104-114, no qmix, RHS_q, sgs_q3,
    psi_m0, psi_h, psi_h0) the fields after theta come out shifted, as in the
    reference."""
    print("checkpoint scalars", flush=True)
    a = _conv_sim("fp64")
    a.step(5)
    path = tmp_path / "vel_sc.out"
    a.checkpoint(path)
    nx, ny, nz, ld = a.nx, a.ny, a.nz, a.ld
    n3, pl, n2 = ld * ny * nz, ld * ny, nx * ny
    names = ["u", "v", "w", "theta", "qmix", "RHSx", "RHSy", "RHSz", "RHS_T", "RHS_q", "sgs_t3", "sgs_q3",
             "psi_m", "psi_m0", "psi_h", "psi_h0", "Cs_opt2", "F_LM", "F_MM", "F_QN", "F_NN", "G_LM", "G_MM",
             "G_QN", "G_NN", "Pr_t"]
    sizes = [n3] * 10 + [pl] * 2 + [n2] * 4 + [n3] * 10
    raw = path.read_bytes()
    assert len(raw) == 8 * sum(sizes) + 8
    segs, _ = _segments(raw[4:-4], sizes)
    for name, seg, n in zip(names, segs, sizes):
        ref = a.get(name).reshape(-1)
        assert ref.size == n and np.array_equal(seg, ref), name
    b = _conv_sim("fp64")
    b.read_initial(path, model=5, initsc=1, inilag=0)
    for name in names:
        assert np.array_equal(b.get(name), a.get(name)), name
    c = _conv_sim("fp64")
    c.read_initial(path, model=7, initsc=1, inilag=0)
    for name in ("u", "v", "w", "theta"):
        assert np.array_equal(c.get(name), a.get(name)), name
    assert np.array_equal(c.get("RHSx"), a.get("qmix"))         # the 5th item of the record
    assert np.array_equal(c.get("RHS_T"), a.get("RHSz"))
    from les_amd.api import LesError
    with pytest.raises(LesError):
        c.read_initial(path, model=9, initsc=1, inilag=0)


def test_checkpoint_fp32_state(torch_mod, tmp_path):
    """An FP32-bulk state is written in real(8): each value widened exactly,
    theta and q in the reference convention; reading it back restores the
    state bit for bit, except that theta and q, stored as deviations from a
    reference value, come back from (deviation + reference) - reference: one
    rounding of the reference value (below 1e-16) per value."""
    print("checkpoint fp32", flush=True)
    a = _conv_sim("fp32")
    a.step(3)
    path = tmp_path / "vel_sc.out"
    a.checkpoint(path)
    n3 = a.ld * a.ny * a.nz
    raw = path.read_bytes()
    segs, _ = _segments(raw[4:], [n3] * 5)
    for name, seg in zip(("u", "v", "w", "theta", "qmix"), segs):
        assert np.array_equal(seg.reshape(a.nz, a.ny, a.ld), a.get(name)), name
    b = _conv_sim("fp32")
    b.read_initial(path, model=5, initsc=1, inilag=0)
    for name in ("u", "v", "w", "RHSx", "F_LM", "G_NN", "psi_m", "sgs_t3"):
        assert np.array_equal(b.view(name).cpu().numpy(), a.view(name).cpu().numpy()), name
    for name in ("theta", "qmix"):
        x, y = b.view(name).double().cpu().numpy(), a.view(name).double().cpu().numpy()
        assert np.max(np.abs(x - y)) <= 1e-16, name


def test_read_initial_short_record(torch_mod, tmp_path):
    from les_amd.api import LesError, LesSim, record_write
    s = LesSim(32, 32, 9)
    path = tmp_path / "short.out"
    record_write(path, [np.zeros(s.ld * s.ny * s.nz * 3)])   # u, v, w only
    with pytest.raises(LesError, match="record holds"):
        s.read_initial(path, model=5, initsc=0, inilag=1)
