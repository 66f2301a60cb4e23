"""ctypes mirror of include/seisfd.h (the engine's C-ABI).

The structures below are byte-for-byte the C structs; the loader finds the
in-tree built library ``seismic_amd/lib/libseisfd.so`` and fails loudly when
it is missing (there is no CPU fallback for any engine entry point).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libseisfd.so")

SFD_OK = 0
SFD_EINVAL = -1
SFD_ENOMEM = -2
SFD_EHIP = -3
SFD_EUNSUPPORTED = -4
SFD_ESTATE = -5
SFD_EUNSTABLE = -6
SFD_ENOTFINITE = -7
SFD_EIO = -8
SFD_EALIASED = -9

SFD_NORM_L2 = 2
SFD_NORM_L1 = 1
SFD_NORM_HUBER = 12
SFD_NORM_CAUCHY = 3
SFD_NORM_HYPERBOLIC = 4
SFD_SNAP_FP32 = 0
SFD_SNAP_FP16S = 1
SFD_SNAP_FP16G = 2

SFD_EXTRACT_SINC_V2 = 0
SFD_EXTRACT_NEAREST = 1
SFD_EXTRACT_SINC_LSRTM = 2
SFD_EXTRACT_SINC_UNIT = 3

SFD_OPT_GRAPHS = 1
SFD_OPT_GRAPH_CHUNK = 2
SFD_OPT_XCD_REMAP = 3
SFD_OPT_KERNEL_EVENTS = 4
SFD_OPT_STRIP_PX = 5
SFD_OPT_TIME_BLOCK = 6
SFD_OPT_PLAN_REUSE = 7
SFD_OPT_BATCH_MIB = 8


class sfd_grid(C.Structure):
    _fields_ = [
        ("nz", C.c_int32), ("nx", C.c_int32),
        ("oz", C.c_float), ("ox", C.c_float),
        ("dz", C.c_float), ("dx", C.c_float),
        ("dt", C.c_float),
        ("ntaper", C.c_int32), ("lsinc", C.c_int32), ("snapi", C.c_int32),
        ("aperture_x", C.c_float),
        ("rec_type", C.c_int32), ("shot_type", C.c_int32), ("surf_type", C.c_int32),
        ("fmax", C.c_float),
    ]


class sfd_shot(C.Structure):
    _fields_ = [
        ("sz", C.c_float), ("sx", C.c_float),
        ("nt", C.c_int32), ("ntr", C.c_int32),
        ("wavelet", C.POINTER(C.c_float)),
        ("rz", C.POINTER(C.c_float)),
        ("rx", C.POINTER(C.c_float)),
    ]


class sfd_grid3d(C.Structure):
    _fields_ = [("nz", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("oz", C.c_float), ("ox", C.c_float), ("oy", C.c_float),
                ("dz", C.c_float), ("dx", C.c_float), ("dy", C.c_float), ("dt", C.c_float),
                ("ntaper", C.c_int32), ("lsinc", C.c_int32),
                ("aperture_x", C.c_float), ("aperture_y", C.c_float), ("fmax", C.c_float), ("snapi", C.c_int32)]


class sfd_shot3d(C.Structure):
    _fields_ = [("sz", C.c_float), ("sx", C.c_float), ("sy", C.c_float), ("nt", C.c_int32), ("ntr", C.c_int32),
                ("wavelet", C.POINTER(C.c_float)), ("rz", C.POINTER(C.c_float)), ("rx", C.POINTER(C.c_float)),
                ("ry", C.POINTER(C.c_float))]


class sfd_fwi_opts(C.Structure):
    _fields_ = [
        ("illupow", C.c_float),
        ("rect_z", C.c_int32), ("rect_x", C.c_int32),
        ("do_z", C.c_int32), ("do_x", C.c_int32),
        ("snap_mode", C.c_int32),
        ("norm", C.c_int32), ("norm_thresh", C.c_float),
    ]


class sfd_afwi_opts(C.Structure):
    _fields_ = [("snap_mode", C.c_int32), ("norm", C.c_int32), ("norm_thresh", C.c_float),
                ("vprho_param", C.c_int32)]


class sfd_bandpass_opts(C.Structure):
    _fields_ = [("flo", C.c_float), ("fhi", C.c_float), ("nplo", C.c_int32), ("nphi", C.c_int32),
                ("phase", C.c_int32)]


class sfd_mute_opts(C.Structure):
    _fields_ = [("v0", C.c_float), ("t0", C.c_float), ("v1", C.c_float), ("t1", C.c_float), ("tramp", C.c_float)]


class sfd_sep_header(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("n", C.c_int64 * 6), ("o", C.c_float * 6), ("d", C.c_float * 6),
                ("esize", C.c_int32), ("xdr", C.c_int32), ("in_", C.c_char * 1024)]


class sfd_lbfgs_opts(C.Structure):
    _fields_ = [
        ("niter", C.c_int), ("neval", C.c_int), ("m", C.c_int), ("const_type", C.c_int),
        ("freeze_soft", C.c_int), ("eval0", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double),
    ]


class sfd_lbfgsoc_opts(C.Structure):
    _fields_ = [
        ("m", C.c_int), ("nparam", C.c_int),
        ("xmin", C.c_void_p), ("xmax", C.c_void_p),
        ("stp1_opt", C.c_int), ("step_init", C.c_double), ("clip_type", C.c_int),
    ]


class sfd_lbfgs_report(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("evaluations", C.c_int), ("iflag", C.c_int), ("info", C.c_int),
        ("nhist", C.c_int),
        ("hist_cap", C.c_int), ("seconds", C.c_double),
        ("f_eval", C.POINTER(C.c_double)), ("t_eval", C.POINTER(C.c_double)),
        ("accepted", C.POINTER(C.c_int)),
        ("pc_iterations", C.c_int), ("pc_fallbacks", C.c_int), ("pc_rejects", C.c_int),
        ("pc_train_seconds", C.c_double), ("pc_loss_first", C.c_double), ("pc_loss_last", C.c_double),
    ]


class sfd_precond_opts(C.Structure):
    _fields_ = [
        ("enable", C.c_int), ("train_after", C.c_int), ("train_steps", C.c_int), ("lr", C.c_float),
        ("seed", C.c_int), ("max_evals", C.c_int), ("max_fallbacks", C.c_int), ("target", C.c_int),
        ("use_net", C.c_int), ("illum_h0", C.c_int), ("coarsen", C.c_int),
    ]


SFD_CNN_NPARAM = 9280


SFD_FCTGDT = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                         C.POINTER(C.c_double))

# Every symbol declared in include/seisfd.h, in declaration order.
EXPORTS = [
    "sfd_create", "sfd_destroy", "sfd_last_error", "sfd_device_info",
    "sfd_set_model", "sfd_set_density", "sfd_set_elevation", "sfd_forward", "sfd_forward_noalloc", "sfd_rtm", "sfd_set_traces", "sfd_born", "sfd_adjoint_exact",
    "sfd_born_adjoint_exact", "sfd_lsrtm", "sfd_bandpass", "sfd_mute_mask",
    "sfd_fwi_gradient", "sfd_afwi_gradient", "sfd_last_timing", "sfd_set_tile", "sfd_get_tile", "sfd_set_option", "sfd_last_waves", "sfd_kernel_stats",
    "sfd_kernel_totals", "sfd_copy_calibration", "sfd_stream_calibration", "sfd_lbfgs_create", "sfd_lbfgs_step", "sfd_lbfgs_destroy",
    "sfd_lbfgsoc_step", "sfd_lbfgsoc_step_files", "sfd_lbfgs_run", "sfd_fwi_lbfgs",
    "sfd_cnn_create", "sfd_cnn_destroy", "sfd_cnn_get_params", "sfd_cnn_set_params", "sfd_cnn_forward",
    "sfd_cnn_loss_grad", "sfd_cnn_train",
    "sfd_sep_read_header", "sfd_sep_read", "sfd_sep_write", "sfd_sep_gathers",
    "sfd_create3d", "sfd_destroy3d", "sfd_last_error3d", "sfd_set_model3d", "sfd_set_density3d", "sfd_forward3d", "sfd_rtm3d", "sfd_lsrtm3d", "sfd_born3d",
    "sfd_afwi_gradient3d",
]

_lib = None


def load():
    """Load libseisfd.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libseisfd.so not built at %s; run __graft_entry__.build()" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    lib.sfd_create.argtypes = [P(C.c_void_p), P(sfd_grid)]
    lib.sfd_create3d.argtypes = [P(C.c_void_p), P(sfd_grid3d)]
    lib.sfd_destroy3d.argtypes = [C.c_void_p]
    lib.sfd_last_error3d.argtypes = [C.c_void_p]
    lib.sfd_last_error3d.restype = C.c_char_p
    lib.sfd_set_model3d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.sfd_rtm3d.argtypes = [C.c_void_p, P(sfd_shot3d), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.sfd_lsrtm3d.argtypes = [C.c_void_p, P(sfd_shot3d), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.sfd_born3d.argtypes = [C.c_void_p, P(sfd_shot3d), C.c_int, C.c_void_p, C.c_void_p]
    lib.sfd_afwi_gradient3d.argtypes = [C.c_void_p, P(sfd_shot3d), C.c_int, C.c_void_p, C.c_void_p,
                                        P(sfd_afwi_opts), P(C.c_double), C.c_void_p, C.c_void_p]
    lib.sfd_set_density3d.argtypes = [C.c_void_p, C.c_void_p]
    lib.sfd_forward3d.argtypes = [C.c_void_p, P(sfd_shot3d), C.c_int, C.c_void_p, C.c_int]
    lib.sfd_destroy.argtypes = [C.c_void_p]
    lib.sfd_last_error.argtypes = [C.c_void_p]
    lib.sfd_last_error.restype = C.c_char_p
    lib.sfd_device_info.argtypes = [C.c_char_p, C.c_int, P(C.c_int), P(C.c_size_t)]
    lib.sfd_set_model.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.sfd_set_elevation.argtypes = [C.c_void_p, C.c_void_p]
    lib.sfd_set_density.argtypes = [C.c_void_p, C.c_void_p]
    lib.sfd_forward.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_int]
    lib.sfd_forward_noalloc.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_int]
    lib.sfd_rtm.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_void_p,
                            C.c_void_p, C.c_int]
    lib.sfd_set_traces.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    lib.sfd_born.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_void_p]
    lib.sfd_adjoint_exact.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_void_p, C.c_int]
    lib.sfd_born_adjoint_exact.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_void_p]
    lib.sfd_lsrtm.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_void_p,
                              C.c_void_p]
    lib.sfd_fwi_gradient.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, P(sfd_fwi_opts), P(C.c_double), C.c_void_p,
                                     C.c_void_p, C.c_void_p]
    lib.sfd_afwi_gradient.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_void_p, P(sfd_afwi_opts),
                                      P(C.c_double), C.c_void_p, C.c_void_p]
    lib.sfd_last_timing.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
    lib.sfd_set_tile.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.sfd_get_tile.argtypes = [C.c_void_p, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), C.c_char_p, C.c_int]
    lib.sfd_set_option.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.sfd_kernel_stats.argtypes = [C.c_void_p, P(C.c_double), P(C.c_int)]
    lib.sfd_last_waves.argtypes = [C.c_void_p, P(C.c_int)]
    lib.sfd_kernel_totals.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_int)]
    lib.sfd_copy_calibration.argtypes = [C.c_size_t, C.c_int, P(C.c_double)]
    lib.sfd_stream_calibration.argtypes = [C.c_size_t, C.c_int, P(C.c_double)]
    lib.sfd_lbfgs_create.argtypes = [P(C.c_void_p), C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.sfd_lbfgs_step.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p,
                                   P(C.c_int), P(C.c_int)]
    lib.sfd_lbfgs_destroy.argtypes = [C.c_void_p]
    lib.sfd_lbfgsoc_step.argtypes = [C.c_char_p, C.c_char_p, P(sfd_lbfgsoc_opts), C.c_int, C.c_void_p,
                                     C.c_double, C.c_void_p, P(C.c_int), P(C.c_int)]
    lib.sfd_sep_read_header.argtypes = [C.c_char_p, P(sfd_sep_header)]
    lib.sfd_sep_read.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t, C.c_void_p]
    lib.sfd_sep_write.argtypes = [C.c_char_p, P(sfd_sep_header), C.c_void_p, C.c_size_t]
    lib.sfd_sep_gathers.argtypes = [C.c_char_p, P(C.c_int), P(C.c_int), C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    lib.sfd_bandpass.argtypes = [C.c_void_p, P(sfd_bandpass_opts), C.c_int, C.c_int, C.c_void_p]
    lib.sfd_mute_mask.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_float, P(sfd_mute_opts), C.c_void_p]
    lib.sfd_lbfgsoc_step_files.argtypes = [C.c_char_p, P(sfd_lbfgsoc_opts), C.c_int, C.c_void_p, C.c_double,
                                           C.c_void_p, P(C.c_int), P(C.c_int)]
    lib.sfd_lbfgs_run.argtypes = [P(sfd_lbfgs_opts), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  SFD_FCTGDT, C.c_void_p, P(sfd_lbfgs_report)]
    lib.sfd_fwi_lbfgs.argtypes = [C.c_void_p, P(sfd_shot), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                  P(sfd_fwi_opts), P(sfd_lbfgs_opts), P(sfd_precond_opts), C.c_void_p,
                                  P(sfd_lbfgs_report)]
    lib.sfd_cnn_create.argtypes = [P(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int]
    lib.sfd_cnn_destroy.argtypes = [C.c_void_p]
    lib.sfd_cnn_get_params.argtypes = [C.c_void_p, C.c_void_p]
    lib.sfd_cnn_set_params.argtypes = [C.c_void_p, C.c_void_p]
    lib.sfd_cnn_forward.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.sfd_cnn_loss_grad.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, P(C.c_double), C.c_void_p]
    lib.sfd_cnn_train.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_float,
                                  P(C.c_double), P(C.c_double)]
    _lib = lib
    return lib
