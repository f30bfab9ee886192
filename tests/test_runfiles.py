"""The reference's run files, host side (no GPU): Fortran sequential
unformatted records (checked against the Fortran runtime of the ROCm flang
compiler and against hand-built subrecords), total_time.dat, the surface
files wts.dat / wqs.dat / zom.dat with bottombc.f90's conversions, and the
param_nml namelist with param.f90's defaults and restart rule; the namelist
reader also reads the reference's own namelist dump (tests/golden/fort.984)."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = os.environ.get("LES_FC", "/opt/rocm/lib/llvm/bin/flang")


@pytest.fixture(scope="module")
def api():
    from les_amd import api as a
    return a


def test_record_framing(api, tmp_path):
    p = tmp_path / "r.bin"
    a, b = np.arange(5.0), np.arange(3, dtype=np.int32)
    api.record_write(p, [a, b])
    api.record_write(p, [b], append=True)
    raw = p.read_bytes()
    n0 = a.nbytes + b.nbytes
    assert struct.unpack("<i", raw[:4])[0] == n0 and struct.unpack("<i", raw[4 + n0:8 + n0])[0] == n0
    assert raw[4:4 + n0] == a.tobytes() + b.tobytes()
    assert api.record_read(p, 0) == a.tobytes() + b.tobytes()
    assert api.record_read(p, 1) == b.tobytes()
    with pytest.raises(api.LesError):
        api.record_read(p, 2)


def test_subrecords(api, tmp_path):
    """A logical record split into three subrecords: leading counts negative
    while more follow, trailing counts negative after the first."""
    parts = [b"A" * 8, b"B" * 5, b"C" * 3]
    heads = [-8, -5, 3]
    tails = [8, -5, -3]
    raw = b"".join(struct.pack("<i", h) + d + struct.pack("<i", t) for h, d, t in zip(heads, parts, tails))
    p = tmp_path / "sub.bin"
    p.write_bytes(raw + struct.pack("<i", 2) + b"ZZ" + struct.pack("<i", 2))
    assert api.record_read(p, 0) == b"".join(parts)
    assert api.record_read(p, 1) == b"ZZ"
    bad = tmp_path / "bad.bin"
    bad.write_bytes(struct.pack("<i", 4) + b"abcd" + struct.pack("<i", 5))
    with pytest.raises(api.LesError):
        api.record_read(bad, 0)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="no flang")
def test_records_match_fortran_runtime(api, tmp_path):
    exe = tmp_path / "records"
    subprocess.run([FLANG, "-O1", os.path.join(ROOT, "tests", "fortran", "records.f90"), "-o", str(exe)],
                   check=True, cwd=tmp_path)
    # Fortran writes, the library reads
    f = tmp_path / "fortran.bin"
    subprocess.run([str(exe), "write", str(f)], check=True)
    a = 0.5 * np.arange(1, 8)
    b = np.array([[10 * i + j for j in (1, 2)] for i in (1, 2, 3)], dtype=np.float64)  # b(i, j)
    assert api.record_read(f, 0) == a.tobytes() + b.tobytes(order="F")
    assert api.record_read(f, 1) == b[:, 0].tobytes()
    # the library writes a longer record, Fortran reads a prefix of it
    g = tmp_path / "library.bin"
    c, d, extra = np.linspace(-1, 1, 7), np.array([3.5, -2.25, 1e-300]), np.array([7.0, 8.0])
    api.record_write(g, [c, d, extra])
    out = subprocess.run([str(exe), "read", str(g)], check=True, capture_output=True, text=True).stdout.split()
    vals = np.array([float(x) for x in out])
    assert np.array_equal(vals, np.concatenate([c, d]))


def test_total_time(api, tmp_path):
    p = tmp_path / "total_time.dat"
    assert api.read_total_time(p) == 0        # missing: the reference assumes 0
    api.write_total_time(p, 12500)
    assert api.read_total_time(p) == 12500
    p.write_text("      700\n")               # any list-directed spacing
    assert api.read_total_time(p) == 700


def test_surface_files(api, tmp_path):
    """ny lines of nx values each (one list-directed READ per line, repeat
    counts and D exponents allowed); wt_s = wts / (rho_d cp), wq_s = 1000 wqs
    / (rho_d Le), zo = zom / z_i (bottombc.f90:50-57)."""
    nx, ny = 4, 3
    wts = np.array([[1.0, 2.0, 3.0, 4.0], [5.0, 6.0, 7.0, 8.0], [9.0, 10.0, 11.0, 12.0]])
    (tmp_path / "wts.dat").write_text("1 2 3 4\n5,6,7,8\n9 10 11 12 trailing\n")
    (tmp_path / "wqs.dat").write_text("4*0.5\n0.5D0 0.5 2*0.5\n4*0.5\n")
    (tmp_path / "zom.dat").write_text("\n".join(" ".join(["0.1"] * nx) for _ in range(ny)) + "\n")
    wt, wq, zo = api.read_surface(tmp_path, nx, ny, z_i=1000.0)
    rho_d, cp, Le = 1.2923, 1005.0, float(np.float32(2.5008e6))
    assert np.allclose(wt, wts / (rho_d * cp), rtol=1e-15)
    assert np.allclose(wq, 1000.0 * 0.5 / (rho_d * Le), rtol=1e-15)
    assert np.allclose(zo, 1e-4, rtol=1e-15)
    (tmp_path / "zom.dat").write_text("0.1 0.1\n")
    with pytest.raises(api.LesError):
        api.read_surface(tmp_path, nx, ny)


def test_namelist_reads_reference_dump(api):
    """tests/golden/fort.984 is the reference's own namelist dump (upper-case
    names, aligned '=', E+000 exponents)."""
    p = os.path.join(ROOT, "tests", "golden", "fort.984")
    assert api.read_nml_value(p, "scales_nml", "q_scale") == 10.66
    assert api.read_nml_value(p, "SCALES_NML", "UG_DIM") == 2.0
    assert api.read_nml_value(p, "scales_nml", "vg_dim") == 0.0
    with pytest.raises(api.LesError):
        api.read_nml_value(p, "scales_nml", "nope")


def test_param_nml_defaults_values_and_restart_rule(api, tmp_path):
    nml = tmp_path / "input.nml"
    nml.write_text("! run input\n&PARAM_NML\n  u_star = 0.45, dt_dim = 0.2D0,\n  nsteps = 500, c_count = 100\n"
                   "  initu = .false., initsc = F, inilag = .TRUE.\n  T_init = 290.5 /\n")
    p, rc = api.read_param_nml(nml)
    assert rc == 0
    assert p["u_star"] == 0.45 and p["dt_dim"] == 0.2 and p["nsteps"] == 500 and p["c_count"] == 100
    assert p["initu"] == 0 and p["initsc"] == 0 and p["inilag"] == 1 and p["T_init"] == 290.5
    assert p["T_scale"] == 400.0 and p["q_scale"] == 100.0 and p["dTdz_top"] == 0.005   # param.f90 defaults
    assert p["dt"] == 0.2 * 0.45 / 1000.0                                               # param.f90:219
    tt = tmp_path / "total_time.dat"
    api.write_total_time(tt, 3000)
    p2, _ = api.read_param_nml(nml, tt)
    assert p2["jt_total"] == 3000 and p2["initu"] == 1 and p2["initsc"] == 1 and p2["inilag"] == 0
    bad = tmp_path / "bad.nml"
    bad.write_text("&param_nml u_star = 0.3, bogus_key = 1, nsteps = 9 /\n")
    p3, rc3 = api.read_param_nml(bad)
    assert rc3 == 3 and p3["u_star"] == 0.3 and p3["nsteps"] == 0   # stops at the unknown item
