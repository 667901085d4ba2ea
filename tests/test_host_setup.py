"""The C++ host layer: setup-file grammar (src/rebuild.cpp:15-116) and registries
(src/fieldsim.cpp:114-194, src/fdtd_plrc.t:411-473) produce the descriptor of the workload
(CPU only), and a setup file runs through FieldSolver on the GPU."""
import math

import numpy as np
import pytest

from mpulse_amd import abi, configs

SETUP_SMALL = """
// 64^3, one Lorentz pole + Kerr, 10-cell CPML, short pulse from z = 12
globals { grid-x 64 grid-y 64 grid-z 64 grid-dx 0.3141592653589793 grid-dy 0.3141592653589793
          grid-dz 0.3141592653589793 dt 0.09068996821171089 T-total 400 }
fieldsim {
  fdtd-plrc-nonlin { eps 1 chi 2e-4  L1eps 1.1 L1Om2 4 L1delta 0.005
                     L2eps 0 L2Om2 1 L3eps 0 L3Om2 1
                     cpml_border { d 10 kappaMax 15 aMax 0.25 sigmaMax 8.795 }
                     short_pulse_inject { d 12 length 20.0 width 4.71238898038469 om0 1 amp 1
                                          eps 2.4666666666666668 TShift 40 ZShift -6.440264939859076 } }
  single_periodic
  energy { interval 10 }
}
"""


def test_setup_file_to_descriptor():
    d, total = abi.parse_setup(SETUP_SMALL)
    assert total == 400
    assert (d.nx, d.ny, d.nz) == (64, 64, 64)
    assert d.solver == abi.SOLVER_PLRC_NONLIN and d.boundary == abi.BOUNDARY_PERIODIC
    assert d.chi == 2e-4 and d.l_eps[0] == 1.1 and d.l_om2[0] == 4 and d.l_delta[0] == 0.005
    assert d.l_om2[1] == 1 and d.l_eps[1] == 0
    assert d.n_factories == 2
    cp, sp = d.factories[0], d.factories[1]
    assert cp.kind == abi.FACTORY_CPML and cp.d == 10 and cp.sigma_max == 8.795
    assert sp.kind == abi.FACTORY_SHORT_PULSE and sp.d == 12 and sp.tshift == 40 and sp.length == 20.0
    ref, _ = configs.config_small()
    assert abs(sp.zshift - ref.factories[1].zshift) < 1e-12


def test_defaults_and_optional_braces():
    """Keys left out take the reference's MakeParamMap defaults; an object key without braces
    builds the object with defaults (src/parameter.t:14-27)."""
    d, total = abi.parse_setup("globals { } fieldsim { fdtd-plrc-lin { cpml_border } single_xy_periodic }")
    assert (d.nx, d.dx, d.dt, total) == (128, 0.1, 1e-2, 100000)
    assert d.solver == abi.SOLVER_PLRC_LIN and d.boundary == abi.BOUNDARY_XY_PERIODIC
    assert d.eps == 1.0 and list(d.l_om2) == [1.0, 1.0, 1.0]
    f = d.factories[0]
    assert d.n_factories == 1 and (f.d, f.kappa_max, f.a_max, f.sigma_max) == (8, 15, 0.25, 3)
    d, _ = abi.parse_setup("globals { } fieldsim { fdtd-plrc-nonlin { short_pulse_inject { } } }")
    f = d.factories[0]
    assert (f.d, f.length, f.width, f.amp, f.eps) == (15, 1.0, 1.0, 1.0, 1.0)
    assert abs(f.om0 - 2 * math.pi) < 1e-15
    assert d.chi == 0.1


def test_setup_errors_follow_the_reference():
    with pytest.raises(abi.MpxError, match="No globals specified"):
        abi.parse_setup("fieldsim { fdtd-plain { } }")
    with pytest.raises(abi.MpxError, match="No field simulation specified"):
        abi.parse_setup("globals { grid-x 8 }")
    with pytest.raises(abi.MpxError, match="side_inject is not provided"):
        abi.parse_setup("globals { } fieldsim { fdtd-plrc-nonlin { side_inject { } } }")


def test_gauss_inject_keys():
    """gauss_inject takes the keys of src/gaussinject.cpp:33-49 with their defaults and becomes
    one GaussInject factory."""
    d, _ = abi.parse_setup("globals { } fieldsim { fdtd-plrc-nonlin { gauss_inject { } } }")
    f = d.factories[0]
    assert d.n_factories == 1 and f.kind == abi.FACTORY_GAUSS_INJECT
    assert (f.d, f.gsize_t, f.length, f.width, f.gshift, f.gtime, f.amp, f.eps, f.toff) == \
        (15, 4096, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0, 0)
    assert abs(f.k0 - 2 * math.pi) < 1e-15
    d, _ = abi.parse_setup("globals { } fieldsim { fdtd-plrc-nonlin { gauss_inject { d 9 SizeT 512 length 2.5 "
                           "width 3.5 k0 6 Shift 1.25 Time 0.5 amp 0.7 eps 2.25 toff 4 } } }")
    f = d.factories[0]
    assert (f.d, f.gsize_t, f.length, f.width, f.k0, f.gshift, f.gtime, f.amp, f.eps, f.toff) == \
        (9, 512, 2.5, 3.5, 6.0, 1.25, 0.5, 0.7, 2.25, 4)


def test_focused_pulse_keys():
    """focused_pulse_inject takes the keys of src/focusedpulseinject.cpp:38-56 (oversampling
    keys accepted) and becomes one focused-pulse factory."""
    d, _ = abi.parse_setup("globals { } fieldsim { fdtd-plrc-lin { focused_pulse_inject { d 11 "
                           "length 30 width 20 om0 1 TShift 60 ZShift 4 Phase 0.5 OversampleX 2 "
                           "OversampleY 2 amp 3 eps 2.25 } } }")
    assert d.n_factories == 1
    f = d.factories[0]
    assert f.kind == abi.FACTORY_FOCUSED_PULSE
    assert (f.d, f.length, f.width, f.om0, f.tshift, f.zshift, f.phase, f.amp, f.eps) == \
        (11, 30, 20, 1, 60, 4, 0.5, 3, 2.25)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,path", [(abi.PREC_FP64, abi.PATH_FUSED), (abi.PREC_FP32_BULK, abi.PATH_FUSED),
                                       (abi.PREC_FP32_BULK, abi.PATH_SWEEP), (abi.PREC_FP64, abi.PATH_REFERENCE)])
def test_setup_file_runs_like_the_descriptor_path(prec, path):
    """The same workload through the C++ FieldSolver layer and through mpx_create directly, for
    the precision and kernel path options of mpulse_hip_run (--fp32, --sweep, --reference)."""
    print(f"case: host driver prec {prec} path {path}", flush=True)
    text = SETUP_SMALL.replace("T-total 400", "T-total 40")
    sim = abi.HostSimulation(text, precision=prec, path=path)
    sim.run()
    d, _ = abi.parse_setup(text)
    d.precision, d.path = prec, path
    with abi.Solver(d) as s:
        s.step(40)
        for n in ("Ex", "By", "Ez"):
            a, b = sim.get(n), s.get(n)
            assert np.array_equal(a, b), n
    sim.close()
