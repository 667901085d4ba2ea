"""Mutation check of the oracle's pins (CPU only).

Each build `oracle/_build/liboracle_mut<N>.so` (`make -C oracle mutants`, the ORACLE_MUTANT
switch of oracle/mpulse_oracle.c) plants one plausible mistake in a part of the path whose
pin is not a golden vector: the incident H-sheet plane, the CPML index layout, fdtd-nl and
fdtd-disp.  For each, the physics checks of tests/test_oracle_pins.py named below must fail
on the mutant -- the same checks pass on the oracle itself in that module.
"""
import os
import subprocess
from functools import partial

import pytest

import oracle_lib as ol
import test_oracle_pins as pins
from mpulse_amd import abi

ORACLE = os.path.join(ol.ROOT, "oracle")

# mutant -> (what it plants, the pins that must catch it)
MUTANTS = {
    1: ("H sheet of the low face on plane d, not d - 1",
        [pins.test_incident_sheet_planes, pins.test_focused_pulse_injects_one_way_and_focuses]),
    2: ("no half-cell offset of the H-side CPML grading",
        [pins.test_cpml_coefficient_profiles, partial(pins.test_cpml_3d_reflects_little_on_every_face, 2),
         pins.test_cpml_absorbs]),
    3: ("outermost H-side CPML node clamped to x = 1",
        [pins.test_cpml_coefficient_profiles]),
    4: ("CPML b/c tables one cell toward the interior",
        [pins.test_cpml_coefficient_profiles, partial(pins.test_cpml_3d_reflects_little_on_every_face, 2)]),
    5: ("sign of the E-side CPML auxiliary update flipped",
        [partial(pins.test_cpml_3d_reflects_little_on_every_face, 2), pins.test_cpml_absorbs]),
    6: ("low-side kappa_E line one cell toward the interior",
        [partial(pins.test_cpml_3d_reflects_little_on_every_face, 2)]),
    7: ("fdtd-nl without the chi E^2 term of D",
        [pins.test_nl_vector_cubic_with_uniform_curl]),
    8: ("fdtd-nl with the curl's sign flipped",
        [pins.test_nl_vector_cubic_with_uniform_curl, partial(pins.test_vacuum_yee_energy_invariant, abi.SOLVER_NL)]),
    9: ("fdtd-disp pole strength from eps instead of eps2",
        [pins.test_disp_zero_dimensional_oscillators_converge]),
    10: ("fdtd-disp without the previous polarisation in E",
         [pins.test_disp_zero_dimensional_oscillators_converge]),
}


@pytest.fixture(scope="module", autouse=True)
def mutant_builds():
    subprocess.run(["make", "-s", "-C", ORACLE, "-j8", "mutants"], check=True)


@pytest.mark.parametrize("m", sorted(MUTANTS))
def test_every_planted_mistake_fails_its_pins(m, monkeypatch):
    what, checks = MUTANTS[m]
    path = ol.mutant_path(m)
    monkeypatch.setattr(ol, "LIB", path)
    monkeypatch.setattr(ol, "LIB_OMP", path)
    for check in checks:
        with pytest.raises(AssertionError):
            check()
        print(f"mutant {m} ({what}): caught by {getattr(check, 'func', check).__name__}")
