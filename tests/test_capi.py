"""C-ABI surface checks that need no GPU: the library loads and exports every
entry point ``include/nufft3d.h`` declares; host-only calls behave."""
import ctypes as C
import os
import re

import pytest

from nufft_3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "nufft3d.h")).read()
    return sorted(set(re.findall(r"\b(nufft3d_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    names = _declared()
    for n in ("nufft3d_plan_create", "nufft3d_forward", "nufft3d_adjoint", "nufft3d_inverse",
              "nufft3d_density", "nufft3d_deap", "nufft3d_plan_destroy", "nufft3d_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for n in _declared():
        assert hasattr(lib, n), n
    assert set(_declared()) <= set(_lib.SIGNATURES)


def test_opts_default_matches_reference_defaults():
    lib = _lib.load()
    o = _lib.Opts()
    assert lib.nufft3d_opts_default(C.byref(o)) == 0
    # nufft_3d.m:27-31: J=4, u=2, alpha=0 (Beatty), radial=1; gpu=1 single
    assert (o.J, o.u, o.alpha, o.radial, o.precision) == (4.0, 2.0, 0.0, 1, 1)
    assert lib.nufft3d_opts_default(None) == _lib.NUFFT3D_ERR_INVALID
    assert b"NULL" in lib.nufft3d_last_error()


def test_null_plan_is_rejected_without_touching_the_device():
    lib = _lib.load()
    assert lib.nufft3d_forward(None, None, 1, None) == _lib.NUFFT3D_ERR_INVALID
    assert lib.nufft3d_nnz(None) == -1
    assert lib.nufft3d_plan_destroy(None) == 0


def test_plan_argument_errors_before_device_use():
    lib = _lib.load()
    o = _lib.Opts()
    lib.nufft3d_opts_default(C.byref(o))
    h = C.c_void_p()
    N = (C.c_int32 * 3)(8, 8, 8)
    om = (C.c_double * 3)(0, 0, 0)
    o.u = 7.0   # nufft_3d.m:82-84
    assert lib.nufft3d_plan_create(om, 1, N, C.byref(o), C.byref(h)) == _lib.NUFFT3D_ERR_INVALID
    assert b"u=7" in lib.nufft3d_last_error()
    o.u = 2.0
    Nodd = (C.c_int32 * 3)(7, 8, 8)   # nufft_3d.m:101-103
    assert lib.nufft3d_plan_create(om, 1, Nodd, C.byref(o), C.byref(h)) == _lib.NUFFT3D_ERR_INVALID
    assert b"even" in lib.nufft3d_last_error()
    assert lib.nufft3d_plan_create(None, 0, N, C.byref(o), C.byref(h)) == _lib.NUFFT3D_ERR_INVALID


def test_operator_refuses_cpu_mode():
    from nufft_3d_amd import nufft_3d
    with pytest.raises(ValueError):
        nufft_3d([[0.0], [0.0], [0.0]], 8, gpu=0)
