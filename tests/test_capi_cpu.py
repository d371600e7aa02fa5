"""The C-ABI library without a GPU: exports, settings parsing, error behaviour.

Nothing here launches a kernel; program build and kernel-name validation are host
code.  Error messages mirror the reference's compile-time #error diagnostics
(induction_Eq_volume.cl:22,35,48,68,81,94,114,127,140; kernel_div_cleaning.cl:52;
*OrderExtended.h NS_D0 #error).
"""
import ctypes
import os
import re

import pytest

from inductioneq_amd import _capi
from inductioneq_amd.settings import ProgramConfig, generate_settings

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_capi.LIB_PATH):
        from inductioneq_amd import build

        build.build(verbose=False)
    return _capi.load()


def test_every_declared_symbol_is_exported(lib):
    declared = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            declared |= set(re.findall(r"\b(ieq_\w+)\s*\(", open(os.path.join(ROOT, "include", h)).read()))
    assert declared == set(_capi.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def _build(lib, settings, modules):
    h = ctypes.c_void_p()
    mods = (ctypes.c_char_p * len(modules))(*[m.encode() for m in modules])
    rc = lib.ieq_program_build(ctypes.byref(h), settings.encode(), mods, len(modules))
    return rc, h


def test_build_and_info_roundtrip(lib):
    cfg = ProgramConfig(real="double", opset="6thOrder", nx=40, ny=36, nz=12, periodic=True, hall=True,
                        dissipation="HIGH_ORDER", ho_factor=0.5, testcase="hall_periodic", w_size=16,
                        xmin=0, xmax=1, ymin=0, ymax=1, zmin=0, zmax=1, dt=1e-3)
    rc, h = _build(lib, cfg.settings_string(), cfg.modules())
    assert rc == 0, lib.ieq_last_error(None)
    info = _capi.ProgramInfo()
    assert lib.ieq_program_get_info(h, ctypes.byref(info)) == 0
    assert (info.nx, info.ny, info.nz) == (40, 36, 12)
    assert info.real_bytes == 8 and info.opset == 2 and info.order == 6 and info.num_bounds == 6
    assert info.periodic == 1 and info.hall == 1 and info.dissipation == 1 and info.ho_factor == 0.5
    assert info.testcase == 2 and info.w_size == 16
    assert info.dx == cfg.dx and info.dt == 1e-3
    assert info.num_nodes == 40 * 36 * 12 and info.num_nodes_pad == cfg.num_nodes_pad
    assert info.stride % 64 == 0 and info.stride >= info.num_nodes_pad
    assert lib.ieq_vector_field_bytes(h) == 3 * info.stride * 8
    lib.ieq_program_release(h)


def test_settings_string_roundtrip():
    cfg = ProgramConfig(real="float", opset="4thOrderExtended", nx=20, ny=20, nz=10, hall=False,
                        dissipation="ADAPTIVE", form_uibj="SPLIT", form_source="ZERO", form_ujbi="PRODUCT",
                        testcase="rotation_3D", laplace="WIDE_STENCIL_LNS", dt=0.01)
    back = ProgramConfig.from_settings(cfg.settings_string(), cfg.modules())
    for f in ("real", "opset", "nx", "ny", "nz", "dx", "dy", "dz", "dt", "periodic", "hall", "dissipation", "form_uibj",
              "form_source", "form_ujbi", "testcase", "laplace", "bnodes"):
        assert getattr(back, f) == getattr(cfg, f), f


def test_generate_settings_formats():
    # matlab/generate_settings.m: USE strings, %d integers, %.16e floats, optimizations verbatim
    s = generate_settings({"a": "USE_HALL", "N": 40, "DX": 0.05, "optimizations": " -cl-mad-enable", "REAL": "double"},
                          ["a", "N", "DX", "optimizations", "REAL"])
    assert s == " -DUSE_HALL=1 -DN=40 -DDX=5.0000000000000003e-02 -cl-mad-enable -DREAL=double"


@pytest.mark.parametrize("drop,msg", [
    ("-DUSE_UIBJ_PRODUCT=1", "No discretization of 'D_j(u_i B_j)' specified"),
    ("-DUSE_SOURCE_CENTRAL=1", "No discretization of 'u_i D_j B_j' specified"),
    ("-DUSE_UJBI_CENTRAL=1", "No discretization of 'D_j(u_j B_i)' specified"),
    ("-DREAL=double", "REAL must be defined"),
])
def test_missing_defines_fail_like_the_reference(lib, drop, msg):
    cfg = ProgramConfig(testcase="rotation_2D")
    rc, _ = _build(lib, cfg.settings_string().replace(" " + drop, ""), cfg.modules())
    assert rc == _capi.IEQ_ERR_SETTINGS
    assert msg in lib.ieq_last_error(None).decode()


def test_narrow_dirichlet_laplacian_rejected_for_extended(lib):
    cfg = ProgramConfig(opset="4thOrderExtended", laplace="NARROW_STENCIL_DIRICHLET", testcase="rotation_2D")
    rc, _ = _build(lib, cfg.settings_string(), cfg.modules())
    assert rc == _capi.IEQ_ERR_SETTINGS
    assert "narrow stencil operator" in lib.ieq_last_error(None).decode()


def test_divergence_cleaning_needs_a_laplacian(lib):
    cfg = ProgramConfig(testcase="rotation_2D", laplace="WIDE_STENCIL_DIRICHLET")
    s = cfg.settings_string().replace(" -DUSE_LAPLACE_WIDE_STENCIL_DIRICHLET=1", "")
    rc, _ = _build(lib, s, cfg.modules())
    assert rc == _capi.IEQ_ERR_SETTINGS
    assert "Define a discretization of the Laplace operator" in lib.ieq_last_error(None).decode()


def test_unknown_and_missing_kernels(lib):
    cfg = ProgramConfig(testcase="rotation_2D")  # no divergence-cleaning module
    rc, h = _build(lib, cfg.settings_string(), cfg.modules())
    assert rc == 0
    g = (ctypes.c_uint32 * 3)(4, 4, 4)
    args = (ctypes.c_void_p * 2)(1, 2)
    for name in (b"no_such_kernel", b"calc_laplace_divcleaning"):
        names = (ctypes.c_char_p * 1)(name)
        rc = lib.ieq_run(h, names, 1, g, g, args, 2, None)
        assert rc == _capi.IEQ_ERR_KERNEL
        assert "kernel not in program" in lib.ieq_last_error(h).decode()
    names = (ctypes.c_char_p * 1)(b"SSPRK33_1")
    assert lib.ieq_run(h, names, 1, g, g, args, 2, None) == _capi.IEQ_ERR_ARGS  # needs 7 arguments
    lib.ieq_program_release(h)


def test_typed_entry_points_cover_every_reference_kernel():
    """One ieq_<name> per kernel of the reference program (95), with the reference buffer count:
    7 for the three-register integrators and their helpers, 6 for the 2N ones
    (kernel_time_integrator.cl:224-227, 630-633, 1023-1026)."""
    assert len(_capi.KERNEL_BUFFERS) == 95
    assert _capi.KERNEL_BUFFERS["SSPRK33_1"] == 7 and _capi.KERNEL_BUFFERS["KennedyCarpenterLewis2R54C_5"] == 7
    assert _capi.KERNEL_BUFFERS["CarpenterKennedy2N54_1a"] == 6 and _capi.KERNEL_BUFFERS["calc_curlB_rho_1_2args"] == 6
    assert _capi.KERNEL_BUFFERS["calc_axpy"] == 3 and _capi.KERNEL_BUFFERS["init_b"] == 1


def test_typed_entry_points_check_like_ieq_run(lib):
    """ieq_<name> goes through the list interface: a kernel missing from the program, a NULL buffer and an
    NDRange outside the grid give ieq_run's codes and messages, before anything is launched."""
    cfg = ProgramConfig(testcase="rotation_2D", nx=8, ny=8, nz=8)  # no divergence-cleaning module
    rc, h = _build(lib, cfg.settings_string(), cfg.modules())
    assert rc == 0
    g = (ctypes.c_uint32 * 3)(8, 8, 8)
    big = (ctypes.c_uint32 * 3)(9, 8, 8)
    p = [ctypes.c_void_p(64 * (k + 1)) for k in range(7)]  # never dereferenced: every call fails its checks first
    assert lib.ieq_calc_laplace_divcleaning(h, g, None, p[0], p[1], None) == _capi.IEQ_ERR_KERNEL
    assert "kernel not in program: calc_laplace_divcleaning" in lib.ieq_last_error(h).decode()
    assert lib.ieq_SSPRK33_1(h, g, None, p[0], None, *p[2:], None) == _capi.IEQ_ERR_ARGS
    assert "SSPRK33_1: argument 1 is NULL" in lib.ieq_last_error(h).decode()
    assert lib.ieq_SSPRK33_1(h, big, None, *p, None) == _capi.IEQ_ERR_ARGS
    assert "NDRange exceeds the grid" in lib.ieq_last_error(h).decode()
    lib.ieq_program_release(h)


def test_run_host_checks_arguments_before_copying(lib):
    """ieq_run_host (MatCL's copy-in / run / write-back call) rejects a NULL host array and one larger than a
    field before it allocates or copies anything."""
    import numpy as np

    cfg = ProgramConfig(testcase="rotation_2D", nx=8, ny=8, nz=8)
    rc, h = _build(lib, cfg.settings_string(), cfg.modules())
    assert rc == 0
    g = (ctypes.c_uint32 * 3)(8, 8, 8)
    names = (ctypes.c_char_p * 1)(b"init_b")
    a = np.zeros((cfg.num_nodes_pad, 4))
    args = (_capi.HostArg * 1)(_capi.HostArg(None, cfg.num_nodes_pad, 1))
    assert lib.ieq_run_host(h, names, 1, g, g, args, 1, None) == _capi.IEQ_ERR_ARGS
    assert "host argument 0 is NULL" in lib.ieq_last_error(h).decode()
    args = (_capi.HostArg * 1)(_capi.HostArg(a.ctypes.data, 1 << 40, 1))
    assert lib.ieq_run_host(h, names, 1, g, g, args, 1, None) == _capi.IEQ_ERR_ARGS
    assert "larger than a field" in lib.ieq_last_error(h).decode()
    lib.ieq_program_release(h)


def test_member_stride(lib):
    assert lib.ieq_member_stride(10, 10, 10) == 1024
    assert lib.ieq_member_stride(8, 8, 1) == 64
    assert lib.ieq_member_stride(0, 4, 4) == 0


def test_ensemble_rejects_before_touching_the_gpu(lib):
    """Invalid ensembles fail at argument validation (no HIP call is reached)."""
    def create(cfg, integrator, members):
        rc, h = _build(lib, cfg.settings_string(), cfg.modules())
        assert rc == 0
        arr = (_capi.Member * len(members))(*members)
        e = ctypes.c_void_p()
        rc = lib.ieq_ensemble_create(ctypes.byref(e), h, arr, len(members), integrator)
        msg = lib.ieq_last_error(h).decode()
        lib.ieq_program_release(h)
        return rc, msg

    m = _capi.Member(8, 8, 8, 0.1, 0.1, 0.1, 1e-3, 1.0, 0, 0, 0, None, None, None, None)
    periodic = ProgramConfig(opset="4thOrder", nx=8, ny=8, nz=8, periodic=True, hall=True, testcase="uniform_flow")
    bounded = ProgramConfig(opset="4thOrder", nx=8, ny=8, nz=8, periodic=False, testcase="rotation_3D")
    rc, msg = create(bounded, b"SSPRK33", [m])
    assert rc == _capi.IEQ_ERR_UNSUPPORTED and "periodic" in msg
    rc, msg = create(periodic, b"SSPRK104", [m])
    assert rc == _capi.IEQ_ERR_UNSUPPORTED and "SSPRK33" in msg
    rc, msg = create(periodic, b"SSPRK33", [m])  # null fields
    assert rc == _capi.IEQ_ERR_ARGS and "invalid member 0" in msg
