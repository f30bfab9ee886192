"""The closed-form pins catch the mistakes they target.

Each case builds the oracle from its committed source with one planted
mistake (les_oracle.c, OR_MUTANT = n) into a temporary directory, runs the
pin that should catch it against that build in a child process
(LES_ORACLE_LIB points oracle/oracle.py at it), and requires the pin to fail.
The unmodified oracle passes the same pins in the main suite
(tests/test_oracle_closed_forms.py, tests/test_oracle_physics_pins.py).
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PINS = "tests/test_oracle_closed_forms.py::"
PHYS = "tests/test_oracle_physics_pins.py::"
MUTANTS = {
    1: ("lagrange_Sdep: 4 <-> 16 in M_ij / N_ij", PINS + "test_lagrange_constant_norm_strain_closed_form"),
    2: ("lagrange_Sdep: Beta inverted", PINS + "test_lagrange_constant_norm_strain_closed_form"),
    3: ("step: tadv1 <-> tadv2", PINS + "test_step_laminar_column_closed_form[1]"),
    4: ("step: forcing on v", PINS + "test_step_laminar_column_closed_form[1]"),
    5: ("press_stag_array: divtz rows with the wrong sign", PINS + "test_press_stag_boundary_rows_closed_form"),
    6: ("lagrange_scalar_Sdep: 4 <-> 16", PINS + "test_lagrange_scalar_constant_norm_strain_closed_form"),
    7: ("lagrange_scalar_Sdep: Beta inverted", PINS + "test_lagrange_scalar_constant_norm_strain_closed_form"),
    8: ("step: no Euler first step without initu", PINS + "test_step_laminar_column_closed_form[0]"),
    9: ("step: Coriolis term of RHSy with the wrong sign", PHYS + "test_coriolis_term"),
    10: ("step: sponge of RHSx at sponge(jz) only", PHYS + "test_sponge_terms[2]"),
    11: ("interpolag_Ssim: interpolag_Sdep's w_lag clamp", PHYS + "test_interpolag_ssim_vertical_shift"),
    12: ("model 1: wall-level length at z = dz", PHYS + "test_mason_length_wall_limit"),
    13: ("new_scaledep_dynamic: 4 <-> 16 in M_ij / N_ij", PHYS + "test_new_scaledep_closed_form"),
    14: ("scalar_scaledep_dynamic: Beta inverted", PHYS + "test_scalar_scaledep_closed_form"),
}


def _cflags():
    """CFLAGS of oracle/Makefile, so the mutants differ from the oracle's own
    build only in OR_MUTANT."""
    for line in open(os.path.join(ROOT, "oracle", "Makefile")):
        if line.startswith("CFLAGS"):
            return shlex.split(line.split("=", 1)[1])
    raise RuntimeError("no CFLAGS in oracle/Makefile")


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("oracle_mutants")
    src = [os.path.join(ROOT, "oracle", f) for f in ("les_oracle.c", "les_mlp_oracle.c")]

    def build(n):
        out = str(d / f"liboracle_mutant{n}.so")
        subprocess.run(["gcc", *_cflags(), f"-DOR_MUTANT={n}", "-shared", "-o", out, *src, "-lm"], check=True)
        return n, out

    with ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build, MUTANTS))


@pytest.mark.parametrize("n", sorted(MUTANTS))
def test_pin_fails_on_planted_mistake(mutant_libs, n):
    what, pin = MUTANTS[n]
    env = dict(os.environ, LES_ORACLE_LIB=mutant_libs[n])
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", pin],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    assert r.returncode == 1, (what, r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "1 failed" in r.stdout, (what, r.stdout[-2000:])
