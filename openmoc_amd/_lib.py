"""ctypes binding of libopenmoc_amd.so (the C ABI declared in
include/openmoc_amd.h).  The library is built in-tree by
``python -m openmoc_amd.build``; importing this module without it raises."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libopenmoc_amd.so"

MOC_OK = 0
SCALAR_FLUX, FISSION_SOURCE, TOTAL_SOURCE = 0, 1, 2
EXP_RATIONAL, EXP_HW = 0, 1
# scattering-source arithmetic (moc_solver_set_scatter_precision)
SCATTER_AUTO, SCATTER_FP64, SCATTER_MFMA_F16, SCATTER_FP32 = 0, 1, 2, 3
# stabilizationType (src/Solver.h:103-113)
DIAGONAL, YAMAMOTO, GLOBAL = 0, 1, 2
# solverMode (src/constants.h)
FORWARD, ADJOINT = 0, 1


class MOCError(RuntimeError):
    """Raised when a C ABI call returns a non-zero status (the reference
    raises RuntimeError from log_printf(ERROR), openmoc/swig/openmoc.i:106-112)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


class TracksInfo(C.Structure):
    _fields_ = [("num_groups", C.c_int32), ("num_materials", C.c_int32),
                ("num_azim", C.c_int32), ("num_polar", C.c_int32),
                ("num_tracks", C.c_int64), ("num_segments", C.c_int64),
                ("num_fsrs", C.c_int64), ("max_optical_length", C.c_double),
                ("tolerance", C.c_double), ("max_iters", C.c_int32),
                ("res_type", C.c_int32), ("is_3d", C.c_int32), ("num_layers", C.c_int32)]


class TrackData(C.Structure):
    _fields_ = [("num_azim", C.c_int32), ("num_polar", C.c_int32),
                ("num_tracks", C.c_int64), ("num_segments", C.c_int64),
                ("seg_offset", C.c_void_p), ("seg_length", C.c_void_p),
                ("seg_fsr", C.c_void_p), ("azim", C.c_void_p),
                ("next_fwd", C.c_void_p), ("next_bwd", C.c_void_p),
                ("next_fwd_is_fwd", C.c_void_p), ("next_bwd_is_fwd", C.c_void_p),
                ("bc_fwd", C.c_void_p), ("bc_bwd", C.c_void_p),
                ("total_weight", C.c_void_p), ("sin_theta", C.c_void_p),
                ("azim_phi", C.c_void_p), ("azim_weight", C.c_void_p),
                ("polar_weight", C.c_void_p), ("seg_start", C.c_void_p),
                ("polar", C.c_void_p), ("xy", C.c_void_p)]


class RequestResult(C.Structure):
    _fields_ = [("keff", C.c_double), ("num_iterations", C.c_int32),
                ("latency_s", C.c_double), ("queue_s", C.c_double), ("solve_s", C.c_double)]


# (name, restype, argtypes) of every exported function
_P = C.c_void_p
_SIGS = [
    ("moc_last_error", C.c_char_p, []),
    ("moc_version", C.c_char_p, []),
    ("moc_tracks_generate", C.c_int, [C.c_char_p, C.c_int, C.POINTER(_P)]),
    ("moc_tracks_destroy", None, [_P]),
    ("moc_tracks_dump_segments", C.c_int, [_P, C.c_char_p]),
    ("moc_tracks_read_segments", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(_P)]),
    ("moc_tracks_info", C.c_int, [_P, C.POINTER(TracksInfo)]),
    ("moc_tracks_export_tracks", C.c_int, [_P] + [_P] * 11),
    ("moc_tracks_export_segments", C.c_int, [_P, _P, _P]),
    ("moc_tracks_export_segment_starts", C.c_int, [_P, _P]),
    ("moc_tracks_export_fsrs", C.c_int, [_P, _P, _P, _P, _P]),
    ("moc_tracks_export_quadrature", C.c_int, [_P] * 7),
    ("moc_tracks_export_tracks_3d", C.c_int, [_P] * 7),
    ("moc_tracks_export_materials", C.c_int, [_P] * 5),
    ("moc_solver_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("moc_solver_destroy", None, [_P]),
    ("moc_solver_set_tracks", C.c_int, [_P, C.POINTER(TrackData)]),
    ("moc_solver_set_materials", C.c_int, [_P, C.c_int32, C.c_int32, _P, _P, _P, _P]),
    ("moc_solver_set_fsrs", C.c_int, [_P, C.c_int64, _P, _P]),
    ("moc_solver_set_convergence_threshold", C.c_int, [_P, C.c_double]),
    ("moc_solver_set_exponential", C.c_int, [_P, C.c_int]),
    ("moc_solver_allow_negative_fluxes", C.c_int, [_P, C.c_int]),
    ("moc_solver_set_scatter_precision", C.c_int, [_P, C.c_int, C.POINTER(C.c_int)]),
    ("moc_solver_compute_fsr_sources", C.c_int, [_P, C.c_int32, C.c_void_p, C.c_int64]),
    ("moc_solver_set_fixed_source_by_fsr", C.c_int, [_P, C.c_int64, C.c_int32, C.c_double]),
    ("moc_device_copy_calibration", C.c_int, [C.c_int, C.c_int64, C.c_int32,
                                              C.POINTER(C.c_double)]),
    ("moc_engine_create", C.c_int, [C.c_int, C.c_int32, C.c_int64, C.POINTER(_P)]),
    ("moc_engine_destroy", None, [_P]),
    ("moc_engine_add_geometry", C.c_int, [_P, C.POINTER(TrackData), C.c_int64, _P, _P,
                                          C.c_int32, C.c_int32, _P, _P, _P, _P,
                                          C.POINTER(C.c_int32)]),
    ("moc_engine_submit", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _P, _P, _P, _P,
                                    C.c_int32, C.c_int32, C.c_double, C.c_int,
                                    C.POINTER(C.c_int64)]),
    ("moc_engine_wait", C.c_int, [_P, C.c_int64, C.c_double, C.POINTER(RequestResult), _P,
                                  C.c_int64]),
    ("moc_engine_stats", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    ("moc_engine_memory", C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64)]),
    ("moc_solver_stabilize_transport", C.c_int, [_P, C.c_double, C.c_int32]),
    ("moc_solver_set_keff_from_neutron_balance", C.c_int, [_P, C.c_int]),
    ("moc_solver_fission_transport_sweep", C.c_int, [_P]),
    ("moc_solver_set_linear_source", C.c_int, [_P, C.c_int]),
    ("moc_solver_get_flux_moments", C.c_int, [_P, _P, C.c_int64]),
    ("moc_solver_get_fsr_centroids", C.c_int, [_P, _P, C.c_int64]),
    ("moc_solver_scatter_transport_sweep", C.c_int, [_P]),
    ("moc_solver_compute_eigenvalue", C.c_int, [_P, C.c_int32, C.c_int32]),
    ("moc_solver_compute_flux", C.c_int, [_P, C.c_int32, C.c_int32]),
    ("moc_solver_compute_source", C.c_int, [_P, C.c_int32, C.c_double, C.c_int32]),
    ("moc_solver_get_keff", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("moc_solver_get_num_iterations", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("moc_solver_get_flux", C.c_int, [_P, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    ("moc_solver_get_fluxes", C.c_int, [_P, _P, C.c_int64]),
    ("moc_solver_set_fluxes", C.c_int, [_P, _P, C.c_int64]),
    ("moc_solver_set_fsr_centroids", C.c_int, [_P, _P, C.c_int64]),
    ("moc_solver_dump_fsr_fluxes", C.c_int, [_P, C.c_char_p]),
    ("moc_solver_load_fsr_fluxes", C.c_int, [_P, C.c_char_p, C.c_int32, C.c_double, _P]),
    ("moc_solver_load_initial_fsr_fluxes", C.c_int, [_P, C.c_char_p]),
    ("moc_solver_set_residual_by_reference", C.c_int, [_P, C.c_char_p]),
    ("moc_solver_set_restart_status", C.c_int, [_P, C.c_int32]),
    ("moc_solver_set_solver_mode", C.c_int, [_P, C.c_int32]),
    ("moc_solver_set_fixed_source_by_material", C.c_int, [_P, C.c_int32, C.c_int32, C.c_double]),
    ("moc_solver_set_fsr_cells", C.c_int, [_P, _P, C.c_int64]),
    ("moc_solver_set_fixed_source_by_cell", C.c_int, [_P, C.c_int32, C.c_int32, C.c_double]),
    ("moc_solver_reset_fixed_sources", C.c_int, [_P]),
    ("moc_solver_set_chi_spectrum_material", C.c_int, [_P, C.c_int32]),
    ("moc_flux_file_write", C.c_int, [C.c_char_p, C.c_double, C.c_int32, C.c_int64, _P, _P]),
    ("moc_flux_file_read", C.c_int, [C.c_char_p, _P, _P, _P, _P, C.c_int64, _P, C.c_int64]),
    ("moc_fsr_match_centroids", C.c_int, [_P, C.c_int64, _P, C.c_int64, C.c_double, _P, _P]),
    ("moc_solver_compute_fsr_fission_rates", C.c_int, [_P, _P, C.c_int64]),
    ("moc_solver_get_timing", C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int64)]),
    ("moc_solver_transport_sweep", C.c_int, [_P, _P, _P, _P]),
    ("moc_solver_transport_sweep_noalloc", C.c_int, [_P, _P, _P, _P, _P, _P]),
    ("moc_solver_noalloc_scratch_bytes", C.c_int, [_P, C.POINTER(C.c_int64)]),
    ("moc_solver_set_launch_shape", C.c_int, [_P, _P]),
    ("moc_solver_get_launch_shape", C.c_int, [_P, _P]),
    ("moc_solver_run_iterations", C.c_int, [_P, C.c_int32, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]),
    ("moc_solver_time_sweeps", C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
]

EXPORTED = [n for n, _, _ in _SIGS]

_lib = None


def lib():
    """Load (once) and return the library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} not found: build it with `python -m openmoc_amd.build` "
                "(the HIP extension is required; there is no CPU fallback)")
        h = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIGS:
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(code: int) -> None:
    if code != MOC_OK:
        raise MOCError(code, lib().moc_last_error().decode(errors="replace"))
