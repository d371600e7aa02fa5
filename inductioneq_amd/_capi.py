"""ctypes binding of ``libinductioneq_hip.so`` (declared in ``include/inductioneq.h``).

There is no fallback: if the library has not been built, loading raises.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("IEQ_LIB_PATH") or os.path.join(_PKG, "libinductioneq_hip.so")

IEQ_OK = 0
IEQ_ERR_SETTINGS = -1
IEQ_ERR_KERNEL = -2
IEQ_ERR_ARGS = -3
IEQ_ERR_HIP = -4
IEQ_ERR_UNSUPPORTED = -5

# every entry point declared in include/inductioneq.h
EXPORTS = [
    "ieq_program_build", "ieq_program_release", "ieq_last_error", "ieq_program_get_info", "ieq_set_stream",
    "ieq_synchronize", "ieq_alloc", "ieq_free", "ieq_vector_field_bytes", "ieq_scalar_field_bytes",
    "ieq_upload_real4", "ieq_download_real4", "ieq_upload_real", "ieq_download_real", "ieq_run", "ieq_run_host",
    "ieq_step", "ieq_step_uniform", "ieq_last_step_path", "ieq_diagnostics", "ieq_diagnostics_noalloc",
    "ieq_clean_div_cg", "ieq_member_stride", "ieq_ensemble_create", "ieq_ensemble_step", "ieq_ensemble_config",
    "ieq_ensemble_release", "ieq_adjoint_create", "ieq_adjoint_forward", "ieq_adjoint_backward", "ieq_adjoint_counts",
    "ieq_adjoint_release",
]

# typed per-kernel entry points declared in include/inductioneq_kernels.h, by buffer count
# (ieq_<name>(prog, g, l, buffers..., kernel_seconds))
KERNEL_BUFFERS = {}
for _n in (
    "init_b", "init_u", "init_rho",
):
    KERNEL_BUFFERS[_n] = 1
for _n in (
    "calc_div_divcleaning", "calc_laplace_divcleaning", "projector_divcleaning", "analytical_b", "norm2",
    "norm2_S", "calc_div", "calc_curl", "calc_grad",
):
    KERNEL_BUFFERS[_n] = 2
for _n in (
    "calc_axpy", "calc_xpay", "dot_product", "norm2_diff",
):
    KERNEL_BUFFERS[_n] = 3
for _n in (
    "calc_time_2_args", "calc_u_2_args", "calc_curlB_rho_1_2args", "calc_curlB_rho_2_2args",
    "CarpenterKennedy2N54_1a", "CarpenterKennedy2N54_1b", "CarpenterKennedy2N54_2a", "CarpenterKennedy2N54_2b",
    "CarpenterKennedy2N54_3a", "CarpenterKennedy2N54_3b", "CarpenterKennedy2N54_4a", "CarpenterKennedy2N54_4b",
    "CarpenterKennedy2N54_5a", "CarpenterKennedy2N54_5b", "ToulorgeDesmet2N84F_1a", "ToulorgeDesmet2N84F_1b",
    "ToulorgeDesmet2N84F_2a", "ToulorgeDesmet2N84F_2b", "ToulorgeDesmet2N84F_3a", "ToulorgeDesmet2N84F_3b",
    "ToulorgeDesmet2N84F_4a", "ToulorgeDesmet2N84F_4b", "ToulorgeDesmet2N84F_5a", "ToulorgeDesmet2N84F_5b",
    "ToulorgeDesmet2N84F_6a", "ToulorgeDesmet2N84F_6b", "ToulorgeDesmet2N84F_7a", "ToulorgeDesmet2N84F_7b",
    "ToulorgeDesmet2N84F_8a", "ToulorgeDesmet2N84F_8b",
):
    KERNEL_BUFFERS[_n] = 6
for _n in (
    "calc_time_3_args", "calc_u_3_args", "calc_curlB_rho_1_3args", "calc_curlB_rho_2_3args",
    "calc_curlB_rho_3_3args", "SSPRK33_1", "SSPRK33_2a", "SSPRK33_2b", "SSPRK33_3a", "SSPRK33_3b", "SSPRK104_01",
    "SSPRK104_02a", "SSPRK104_02b", "SSPRK104_03a", "SSPRK104_03b", "SSPRK104_04a", "SSPRK104_04b", "SSPRK104_05a",
    "SSPRK104_05b", "SSPRK104_06", "SSPRK104_07", "SSPRK104_08a", "SSPRK104_08b", "SSPRK104_09a", "SSPRK104_09b",
    "SSPRK104_10a", "SSPRK104_10b", "SSPRK104_11a", "SSPRK104_11b", "KennedyCarpenterLewis2R54C_1a",
    "KennedyCarpenterLewis2R54C_1b", "KennedyCarpenterLewis2R54C_2a", "KennedyCarpenterLewis2R54C_2b",
    "KennedyCarpenterLewis2R54C_3a", "KennedyCarpenterLewis2R54C_3b", "KennedyCarpenterLewis2R54C_4a",
    "KennedyCarpenterLewis2R54C_4b", "KennedyCarpenterLewis2R54C_5", "CalvoFrancoRandez2R64_1a",
    "CalvoFrancoRandez2R64_1b", "CalvoFrancoRandez2R64_2a", "CalvoFrancoRandez2R64_2b", "CalvoFrancoRandez2R64_3a",
    "CalvoFrancoRandez2R64_3b", "CalvoFrancoRandez2R64_4a", "CalvoFrancoRandez2R64_4b", "CalvoFrancoRandez2R64_5a",
    "CalvoFrancoRandez2R64_5b", "CalvoFrancoRandez2R64_6",
):
    KERNEL_BUFFERS[_n] = 7
EXPORTS += ["ieq_" + _n for _n in KERNEL_BUFFERS]

_NOT_INT = ("ieq_program_release", "ieq_last_error", "ieq_vector_field_bytes", "ieq_scalar_field_bytes",
            "ieq_member_stride", "ieq_ensemble_release", "ieq_adjoint_release")


class HostArg(ctypes.Structure):
    _fields_ = [("host", ctypes.c_void_p), ("n", ctypes.c_size_t), ("vector", ctypes.c_int)]


class CgResult(ctypes.Structure):
    _fields_ = [("divergence_norm_initial", ctypes.c_double), ("divergence_norm_final", ctypes.c_double),
                ("residual_norm", ctypes.c_double), ("iterations", ctypes.c_int), ("seconds", ctypes.c_double)]


class Member(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double), ("rho", ctypes.c_double),
                ("ux", ctypes.c_double), ("uy", ctypes.c_double), ("uz", ctypes.c_double), ("b1", ctypes.c_void_p),
                ("b2", ctypes.c_void_p), ("b3", ctypes.c_void_p), ("emf", ctypes.c_void_p)]


class ProgramInfo(ctypes.Structure):
    _fields_ = [
        ("real_bytes", ctypes.c_int), ("opset", ctypes.c_int), ("order", ctypes.c_int), ("num_bounds", ctypes.c_int),
        ("stencil_width", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double), ("xmin", ctypes.c_double),
        ("ymin", ctypes.c_double), ("zmin", ctypes.c_double), ("dt", ctypes.c_double),
        ("periodic", ctypes.c_int), ("hall", ctypes.c_int), ("dissipation", ctypes.c_int),
        ("form_uibj", ctypes.c_int), ("form_source", ctypes.c_int), ("form_ujbi", ctypes.c_int),
        ("testcase", ctypes.c_int), ("laplace", ctypes.c_int), ("bnodes", ctypes.c_int), ("w_size", ctypes.c_int),
        ("ho_factor", ctypes.c_double), ("mp2min", ctypes.c_double), ("mp2max", ctypes.c_double),
        ("cmin", ctypes.c_double), ("cmax", ctypes.c_double), ("ux", ctypes.c_double), ("uy", ctypes.c_double),
        ("uz", ctypes.c_double), ("num_nodes", ctypes.c_longlong), ("num_nodes_pad", ctypes.c_longlong),
        ("stride", ctypes.c_longlong), ("num_groups", ctypes.c_longlong),
    ]


_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libinductioneq_hip.so is not built (%s); run `python -m inductioneq_amd.build` "
                          "or __graft_entry__.build()" % LIB_PATH)
    try:
        # torch ships its own HIP runtime; with torch loaded first the library binds to that same runtime, so
        # device memory from torch tensors and the library's own calls share one runtime (loaded the other way
        # round, the library's first allocation found no device once torch had started)
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    lib.ieq_program_build.argtypes = [ctypes.POINTER(P), ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]
    lib.ieq_program_build.restype = ctypes.c_int
    lib.ieq_program_release.argtypes = [P]
    lib.ieq_program_release.restype = None
    lib.ieq_last_error.argtypes = [P]
    lib.ieq_last_error.restype = ctypes.c_char_p
    lib.ieq_program_get_info.argtypes = [P, ctypes.POINTER(ProgramInfo)]
    lib.ieq_set_stream.argtypes = [P, P]
    lib.ieq_synchronize.argtypes = [P]
    lib.ieq_alloc.argtypes = [P, ctypes.c_size_t, ctypes.POINTER(P)]
    lib.ieq_free.argtypes = [P, P]
    lib.ieq_vector_field_bytes.argtypes = [P]
    lib.ieq_vector_field_bytes.restype = ctypes.c_size_t
    lib.ieq_scalar_field_bytes.argtypes = [P]
    lib.ieq_scalar_field_bytes.restype = ctypes.c_size_t
    for fn in ("ieq_upload_real4", "ieq_download_real4", "ieq_upload_real", "ieq_download_real"):
        getattr(lib, fn).argtypes = [P, P, P, ctypes.c_size_t]
    u32x3 = ctypes.POINTER(ctypes.c_uint32)
    lib.ieq_run.argtypes = [P, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, u32x3, u32x3, ctypes.POINTER(P),
                            ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.ieq_run_host.argtypes = [P, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, u32x3, u32x3, ctypes.POINTER(HostArg),
                                 ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.ieq_step.argtypes = [P, ctypes.c_char_p, ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(P), ctypes.c_int,
                             ctypes.POINTER(ctypes.c_double)]
    lib.ieq_step_uniform.argtypes = [P, ctypes.c_char_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                     ctypes.POINTER(P), ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.ieq_last_step_path.argtypes = [P]
    lib.ieq_diagnostics.argtypes = [P, P, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
    lib.ieq_diagnostics_noalloc.argtypes = [P, P, ctypes.c_double, P]
    lib.ieq_clean_div_cg.argtypes = [P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(CgResult)]
    lib.ieq_member_stride.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.ieq_member_stride.restype = ctypes.c_longlong
    lib.ieq_ensemble_create.argtypes = [ctypes.POINTER(P), P, ctypes.POINTER(Member), ctypes.c_int, ctypes.c_char_p]
    lib.ieq_ensemble_step.argtypes = [P, ctypes.c_longlong, ctypes.POINTER(ctypes.c_double)]
    ip = ctypes.POINTER(ctypes.c_int)
    lib.ieq_ensemble_config.argtypes = [P, ip, ip, ip]
    lib.ieq_ensemble_release.argtypes = [P]
    lib.ieq_ensemble_release.restype = None
    dp = ctypes.POINTER(ctypes.c_double)
    lib.ieq_adjoint_create.argtypes = [ctypes.POINTER(P), P, ctypes.c_longlong, ctypes.c_int]
    lib.ieq_adjoint_forward.argtypes = [P, P, P, ctypes.c_double, dp, dp]
    lib.ieq_adjoint_backward.argtypes = [P, P, P, dp, dp]
    llp = ctypes.POINTER(ctypes.c_longlong)
    lib.ieq_adjoint_counts.argtypes = [P, llp, llp]
    lib.ieq_adjoint_release.argtypes = [P]
    lib.ieq_adjoint_release.restype = None
    for name, nbuf in KERNEL_BUFFERS.items():
        getattr(lib, "ieq_" + name).argtypes = [P, u32x3, u32x3] + [P] * nbuf + [ctypes.POINTER(ctypes.c_double)]
    for fn in EXPORTS:
        if fn not in _NOT_INT:
            getattr(lib, fn).restype = ctypes.c_int
    _lib = lib
    return lib


class IEQError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


def check(rc: int, handle=None):
    if rc != IEQ_OK:
        msg = load().ieq_last_error(handle)
        raise IEQError(rc, msg.decode() if msg else "error")
