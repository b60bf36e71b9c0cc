"""The C-ABI library: it loads, exports every entry point include/fot_capi.h
declares, and reports errors as statuses (the reference's exceptions and
exits become fot_status).  No GPU compute."""
import ctypes as C

import numpy as np
import pytest

from cuda_fot_amd import ColourOverflow, FotError, InvalidInputFile, capi, plan


def test_exports_every_declared_symbol():
    names = capi.declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(capi.lib, n)]
    assert not missing, missing
    # and the Python binding covers them all
    assert not set(names) - set(capi._PROTOS)


def test_exceptions_become_statuses():
    # every entry point runs through the library's exception guard: a C++
    # exception (out of host memory or any other) returns a status and a
    # message instead of crossing the C ABI (tests/cpp/guard_test.cpp)
    import os
    import subprocess
    binary = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin", "guard_test")
    r = subprocess.run([binary], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def test_status_strings():
    for s in range(9):
        assert capi.lib.fot_status_string(s)
    assert capi.lib.fot_version() >= 1


def test_device_count_without_gpu_does_not_crash():
    n = C.c_int(-1)
    st = capi.lib.fot_device_count(C.byref(n))
    assert st in (capi.OK, capi.ERR_NO_DEVICE) and n.value >= 0


def test_invalid_arguments_return_status():
    assert capi.lib.fot_colour_cells(None, 10, 2, 10, None, None) == capi.ERR_INVALID_ARGUMENT
    assert capi.lib.fot_hplan_build(None, None, 0, 0, 0, None, 0, None) == \
        capi.ERR_INVALID_ARGUMENT
    assert capi.lib.fot_loop_create(None, None, None, None) == capi.ERR_INVALID_ARGUMENT
    f = capi.Functor()
    assert capi.lib.fot_functor_init(C.byref(f), capi.KERNEL_MINE, 3, 4, 4) == \
        capi.ERR_INVALID_ARGUMENT
    assert capi.lib.fot_functor_init(C.byref(f), capi.KERNEL_MINE, 2, 4, 2) == \
        capi.ERR_INVALID_ARGUMENT
    assert capi.lib.fot_functor_init(C.byref(f), 99, 0, 0, 0) == capi.ERR_INVALID_ARGUMENT


def test_missing_point_or_cell_data_is_refused():
    # a null array the functor reads would fault on the device: refused on
    # the host before any plan or launch
    d = capi.ProblemDesc()
    edges = np.array([[0, 1], [1, 2]], dtype=np.uint32)
    d.num_cells, d.num_mappings = 2, 1
    d.cell_to_node[0] = edges.ctypes.data
    d.mesh_dim[0], d.num_points[0], d.point_dim[0] = 2, 3, 4
    d.num_cell_data, d.cell_dim[0] = 1, 4
    d.point_data_out = 1  # never dereferenced: refused before
    f = capi.functor(capi.KERNEL_MINE, 2, 4, 4)
    cfg = capi.LoopConfig(capi.STRATEGY_ATOMIC, 64, None, 0, 0)
    loop = C.c_void_p()
    assert capi.lib.fot_loop_create(C.byref(f), C.byref(d), C.byref(cfg), C.byref(loop)) == \
        capi.ERR_INVALID_ARGUMENT
    assert b"point data 0" in capi.lib.fot_last_error_message()
    d.point_data[0] = 1
    assert capi.lib.fot_loop_create(C.byref(f), C.byref(d), C.byref(cfg), C.byref(loop)) == \
        capi.ERR_INVALID_ARGUMENT
    assert b"cell data 0" in capi.lib.fot_last_error_message()


def test_out_of_range_mesh_is_invalid_input():
    # the reference throws InvalidInputFile on a mapping beyond its point set
    bad = np.array([[0, 1], [1, 7]], dtype=np.uint32)
    with pytest.raises(InvalidInputFile):
        plan.colour_cells(bad, 4)
    with pytest.raises(InvalidInputFile):
        plan.hplan([bad], 4, 32)
    with pytest.raises(FotError):
        plan.point_renumbering_first_touch(bad, 4)


def test_colour_overflow_is_reported():
    # 65 cells on one point need 65 colours; the reference has 64 (mesh.hpp:128)
    star = np.array([[0, i + 1] for i in range(65)], dtype=np.uint32)
    with pytest.raises(ColourOverflow):
        plan.colour_cells(star, 66)
    # and more than 128 block colours (colouring.hpp:125-136)
    hub = np.array([[0, i + 1] for i in range(200)], dtype=np.uint32)
    with pytest.raises(ColourOverflow):
        plan.hplan([hub], 201, 1)


def test_last_error_message_names_the_problem():
    bad = np.array([[0, 9]], dtype=np.uint32)
    with pytest.raises(InvalidInputFile) as e:
        plan.colour_cells(bad, 3)
    assert "out of range" in str(e.value)
