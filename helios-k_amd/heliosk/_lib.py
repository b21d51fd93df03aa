"""ctypes binding of the C ABI in include/heliosk.h (lib/libheliosk.so).

The library must be built (helios-k_amd/Makefile); loading fails loudly when
it is missing -- there is no fallback path.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libheliosk.so")

HK_OK = 0
HK_ERR_ARG, HK_ERR_HIP, HK_ERR_WORKSPACE, HK_ERR_UNSUPPORTED = 1, 2, 3, 4
HK_FORM_HITRAN, HK_FORM_EXOMOL, HK_FORM_SUPERLINE = 0, 1, 2
HK_VOIGT_FP64 = 0
HK_VOIGT_FP32C = 1
HK_VOIGT_COMPAT = 2
HK_PROFILE_VOIGT, HK_PROFILE_LORENTZ, HK_PROFILE_DOPPLER, HK_PROFILE_BINNED_GAUSS = 1, 2, 3, 4
HK_MAX_ISO = 16

c_d = ctypes.c_double
c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_size = ctypes.c_size_t
c_vp = ctypes.c_void_p
P_d = ctypes.POINTER(ctypes.c_double)


class Lines(ctypes.Structure):
    _fields_ = [("nlines", c_i64),
                ("nu", c_vp), ("S", c_vp), ("EL", c_vp), ("A", c_vp), ("delta", c_vp),
                ("vy", c_vp), ("n", c_vp), ("iso", c_vp), ("niso", c_int), ("form", c_int),
                ("iso_mass", c_d * HK_MAX_ISO), ("iso_ab", c_d * HK_MAX_ISO),
                ("iso_sscale", c_d * HK_MAX_ISO), ("max_abs_delta", c_d)]


class TP(ctypes.Structure):
    _fields_ = [("T", c_d), ("P", c_d), ("Q", c_d * HK_MAX_ISO)]


class Cut(ctypes.Structure):
    _fields_ = [("cut_mode", c_int), ("cut", c_d)]


class Chi(ctypes.Structure):
    _fields_ = [("sigma", ctypes.c_double * 3), ("B", ctypes.c_double * 3),
                ("range", ctypes.c_double * 2)]


class Grid(ctypes.Structure):
    _fields_ = [("numin", c_d), ("dnu", c_d), ("nx", c_i64)]


class LineState(ctypes.Structure):
    _fields_ = [("nlines", c_i64), ("nu", c_vp), ("S", c_vp), ("ialphaD", c_vp), ("y", c_vp),
                ("hw", c_vp), ("ey", c_vp), ("blk_max", c_vp), ("radius", c_vp), ("key", c_vp)]


class Tune(ctypes.Structure):
    _fields_ = [("tile_points", c_int), ("exact_pairs", c_int), ("stage_lines", c_int)]


class Spectrum(ctypes.Structure):
    _fields_ = [("st", LineState), ("grid", Grid), ("kmin", c_d), ("accumulate", c_int),
                ("K", c_vp)]


class VoigtBatch(ctypes.Structure):
    _fields_ = [("workspace", c_vp), ("nspec", c_int), ("mode", c_int), ("tile_points", c_int),
                ("stage_lines", c_int), ("items_main", c_i64), ("items_core", c_i64),
                ("off_args", c_size), ("off_cargs", c_size), ("off_items", c_size),
                ("off_citems", c_size), ("off_counter", c_size)]


_lib = None


class HKError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HKError("libheliosk.so is not built (%s); run __graft_entry__.build() "
                          "or make -C helios-k_amd" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.hk_last_error.restype = ctypes.c_char_p
        L.hk_version.restype = ctypes.c_char_p
        L.hk_line_state_bytes.restype = c_size
        L.hk_line_state_bytes.argtypes = [c_i64]
        L.hk_line_state_init.argtypes = [P(LineState), c_i64, c_vp, c_size]
        L.hk_line_params.argtypes = [P(Lines), P(TP), P(Cut), P(LineState), c_vp]
        L.hk_lines_sort_workspace_bytes.restype = c_size
        L.hk_lines_sort_workspace_bytes.argtypes = [c_i64]
        L.hk_lines_sort.argtypes = [P(Lines), P(Lines), c_vp, c_size, c_vp]
        L.hk_profile_workspace_bytes.restype = c_size
        L.hk_profile_workspace_bytes.argtypes = [P(Grid), c_i64]
        L.hk_profile_accumulate.argtypes = [P(LineState), P(Grid), c_d, c_int, c_int, c_int, c_int,
                                            P(Chi), c_vp, c_vp, c_size, c_vp]
        L.hk_means_x.argtypes = [c_vp, c_vp, c_i64, c_d, c_vp, c_vp, c_size, c_vp]
        L.hk_sort_bins_workspace_bytes.restype = c_size
        L.hk_sort_bins_workspace_bytes.argtypes = [c_i64, c_int]
        L.hk_sort_bins.argtypes = [c_vp, c_i64, ctypes.POINTER(c_i64), c_int, c_vp, c_vp, c_size,
                                   c_vp]
        L.hk_means_workspace_bytes.restype = c_size
        L.hk_means_workspace_bytes.argtypes = []
        L.hk_means.argtypes = [c_vp, P(Grid), c_d, c_vp, c_vp, c_size, c_vp]
        L.hk_transmission.argtypes = [c_vp, c_int, c_int, c_vp, c_d, c_int, c_d, c_vp, c_vp]
        L.hk_voigt_workspace_bytes.restype = c_size
        L.hk_voigt_workspace_bytes.argtypes = [P(Grid), P(Tune)]
        L.hk_voigt_accumulate.argtypes = [P(LineState), P(Grid), c_d, c_int, c_int, P(Tune),
                                          c_vp, c_vp, c_size, c_vp]
        L.hk_voigt_accumulate_noalloc.argtypes = [P(LineState), P(Grid), c_d, c_int, c_int,
                                                  P(Tune), c_vp, c_vp]
        L.hk_voigt_pair_count.argtypes = [P(LineState), P(Grid), c_vp, c_vp]
        L.hk_voigt_work_count.argtypes = [P(LineState), P(Grid), c_int, P(Tune), c_vp, c_vp]
        L.hk_voigt_autotune.argtypes = [P(LineState), P(Grid), c_int, c_vp, c_vp, c_size, P(Tune),
                                        ctypes.POINTER(ctypes.c_float), c_vp]
        L.hk_voigt_autotune_reset.restype = None
        L.hk_voigt_autotune_reset.argtypes = []
        L.hk_voigt_batched_workspace_bytes.restype = c_size
        L.hk_voigt_batched_workspace_bytes.argtypes = [c_int, P(Spectrum), P(Tune)]
        L.hk_voigt_batched.argtypes = [c_int, P(Spectrum), c_int, P(Tune), c_vp, c_size, c_vp]
        L.hk_voigt_batched_prepare.argtypes = [c_int, P(Spectrum), c_int, P(Tune), c_vp, c_size,
                                               P(VoigtBatch), c_vp]
        L.hk_voigt_batched_launch.argtypes = [P(VoigtBatch), c_vp]
        L.hk_ktable_workspace_bytes.restype = c_size
        L.hk_ktable_workspace_bytes.argtypes = [c_i64, c_int, c_int, c_int]
        L.hk_ktable_prepare.argtypes = [P(Grid), c_d, c_int, c_int, c_vp, c_size]
        L.hk_ktable.argtypes = [c_vp, c_int, P(Grid), c_d, c_int, c_int, c_d, c_d, c_vp, c_vp,
                                c_vp, c_vp, c_size, c_vp]
        L.hk_ktable_backward_workspace_bytes.restype = c_size
        L.hk_ktable_backward_workspace_bytes.argtypes = [c_i64, c_int, c_int, c_int]
        L.hk_ktable_backward.argtypes = [c_vp, c_vp, c_vp, c_int, P(Grid), c_d, c_int, c_int, c_vp,
                                         c_vp, c_vp, c_size, c_vp]
        L.hk_ktable_reconstruct.argtypes = [c_vp, c_vp, c_int, c_int, c_int, c_int, c_d, c_d,
                                            c_vp, c_vp, c_vp]
        L.hk_ktable_plan_host.argtypes = [c_int, c_int, P_d]
        L.hk_voigt_adjoint_workspace_bytes.restype = c_size
        L.hk_voigt_adjoint_workspace_bytes.argtypes = [c_i64]
        L.hk_voigt_adjoint.argtypes = [P(Lines), P(TP), P(Cut), P(Grid), c_vp, c_vp, c_vp, c_vp,
                                       c_vp, c_size, c_vp]
        L.hk_bin_segments.argtypes = [P(Grid), c_d, c_int, ctypes.POINTER(c_i64)]
        L.hk_synth_lines.argtypes = [ctypes.c_uint64, c_int, c_i64, c_int, P_d, P_d, P_d, P_d,
                                     P_d]
        _lib = L
    return _lib


def check(rc):
    if rc != HK_OK:
        raise HKError("heliosk error %d: %s" % (rc, lib().hk_last_error().decode()))


# Every symbol include/heliosk.h declares.
EXPORTED = ["hk_last_error", "hk_version", "hk_line_state_bytes", "hk_line_state_init",
            "hk_lines_sort_workspace_bytes", "hk_lines_sort", "hk_means_workspace_bytes",
            "hk_means", "hk_transmission", "hk_profile_workspace_bytes", "hk_profile_accumulate",
            "hk_means_x", "hk_sort_bins_workspace_bytes", "hk_sort_bins",
            "hk_line_params", "hk_voigt_workspace_bytes", "hk_voigt_accumulate",
            "hk_voigt_accumulate_noalloc", "hk_voigt_pair_count",
            "hk_voigt_work_count",
            "hk_voigt_autotune", "hk_voigt_autotune_reset",
            "hk_voigt_batched_workspace_bytes", "hk_voigt_batched", "hk_voigt_batched_prepare",
            "hk_voigt_batched_launch", "hk_ktable_workspace_bytes",
            "hk_ktable_prepare", "hk_ktable", "hk_ktable_reconstruct", "hk_ktable_plan_host",
            "hk_ktable_backward_workspace_bytes", "hk_ktable_backward",
            "hk_voigt_adjoint_workspace_bytes", "hk_voigt_adjoint", "hk_bin_segments",
            "hk_synth_lines"]
