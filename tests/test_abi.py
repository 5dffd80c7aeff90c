"""The C ABI library loads and exports every symbol include/fvmr.h declares; host-side
entry points return status codes (never exit) and report the reason."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cfd_solver_2d_gpu_amd as F
from cfd_solver_2d_gpu_amd import _abi
from conftest import ROOT, gpu_available


def declared_functions():
    src = open(os.path.join(ROOT, "include", "fvmr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fvmr_\w+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    lib = C.CDLL(_abi.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds exactly the declared set
    assert sorted(_abi.SIGNATURES) == names


def test_struct_sizes_match_header_layout():
    # fvmr_params: 7 doubles, 14 ints, 3x4 doubles, 5 doubles
    assert C.sizeof(F.Params) == 7 * 8 + 14 * 4 + 12 * 8 + 5 * 8
    assert C.sizeof(F.StepInfo) == 6 * 8 + 4 * 4
    assert C.sizeof(F.MeshSoA) == 16 + 8 * (2 + 4 * 3 + 2 + 1 + 2 + 5 + 2 + 2 + 1)


def test_params_default_is_the_reference_globalpara():
    p = F.default_params()
    assert (p.gamma, p.R, p.Pr, p.T0, p.mu0) == (1.4, 287.06, 0.73, 288.16, 17.9e-6)
    assert p.flux_scheme == F.SOL_ROE and p.equation == F.EQ_EULER
    assert p.time_advance == F.EVO_EXPLICIT and p.Pr_viscous == 0.72 and p.dsr_shock_speed == 20.0
    assert (p.renumber, p.kernels, p.tile_cells) == (F.RENUMBER_HILBERT, F.KERNELS_AUTO, 0)


def test_ruvp_from_mach_closed_form():
    r = F.ruvp_from_mach(0.5, 0.0)
    T = 288.16 / (1 + 0.2 * 0.25)
    p = 101325.0 * (T / 288.16) ** 3.5
    assert np.allclose(r, [p / (287.06 * T), 0.5 * np.sqrt(1.4 * 287.06 * T), 0.0, p], rtol=1e-14)


def test_host_errors_are_status_codes(tmp_path):
    with pytest.raises(F.FvmrError) as e:
        F.Mesh.read_su2(tmp_path / "missing.su2")
    assert e.value.code == 5 and "cannot open file" in str(e.value)
    bad = tmp_path / "bad.su2"
    bad.write_text("NDIME= 3\n")
    with pytest.raises(F.FvmrError) as e:
        F.Mesh.read_su2(bad)
    assert e.value.code == 3 and "not 2D" in str(e.value)
    with pytest.raises(F.FvmrError):
        F.Mesh.generate(99, 4, 4)
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, 4, 4, b=1.0)
    with pytest.raises(F.FvmrError) as e:
        F.field_init(m, F.default_params(), 9, (0, 0, 0))
    assert "initialCondition" in str(e.value)


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_device_fails_cleanly():
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, 4, 4, b=1.0)
    with pytest.raises(F.FvmrError) as e:
        F.GPUSolver2(m, F.default_params())
    assert e.value.code == 2


def test_invalid_params_rejected_before_device():
    m = F.Mesh.generate(F.GEN_PERIODIC_SQUARE, 4, 4, b=1.0)
    for field, val, msg in (("flux_scheme", 2005, "scheme"), ("time_advance", 3105, "temporal"),
                            ("gradient", 7, "graident"), ("limiter", 3, "limiter")):
        p = F.default_params(**{field: val})
        with pytest.raises(F.FvmrError) as e:
            F.GPUSolver2(m, p)
        assert msg in str(e.value)


def test_state_arrays_must_be_four_rows_of_the_cell_count(tmp_path):
    """The library copies num_element doubles from each of the four rows: a state of another
    shape is refused before any copy."""
    import cfd_solver_2d_gpu_amd as FF
    with pytest.raises(ValueError):
        FF._rows(np.zeros((4, 5)), 6)
    with pytest.raises(ValueError):
        FF._rows(np.zeros((3, 6)), 6)
    U, rows = FF._rows(np.zeros((4, 6)), 6)
    assert U.shape == (4, 6)
    m = FF.Mesh.generate(FF.GEN_PERIODIC_SQUARE, 4, 4, b=1.0)
    with pytest.raises(ValueError):
        m.write_continue(tmp_path / "pause_x[1].dat", np.ones((4, m.num_element - 1)), 0.0, 1, 0.5)
