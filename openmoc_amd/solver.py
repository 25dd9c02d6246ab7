"""Host-side mirror of the reference OpenMOC Python interface for the hot path.

``TrackGenerator`` is the counterpart of ``openmoc.TrackGenerator`` (CPU
producer: laydown + ray tracing) and ``HipSolver`` (alias ``GPUSolver``) the
counterpart of ``openmoc.CPUSolver`` / ``openmoc.cuda.GPUSolver``: same method
names, argument meaning and error behaviour (errors raise ``RuntimeError``
subclasses, as SWIG maps ``log_printf(ERROR)``, openmoc/swig/openmoc.i:106-112).

Usage, as in the reference's sample inputs
(sample-input/benchmarks/c5g7/c5g7-2d.py:88-95)::

    tg = TrackGenerator(problem)        # problem: workloads.Problem or text
    tg.generateTracks()
    solver = HipSolver(tg)
    solver.setConvergenceThreshold(1e-5)
    solver.computeEigenvalue(1000)
    keff = solver.getKeff()
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Union

import numpy as np

from . import _lib
from ._lib import (ADJOINT, DIAGONAL, EXP_HW, EXP_RATIONAL, FISSION_SOURCE, FORWARD, GLOBAL,
                   SCALAR_FLUX,
                   SCATTER_AUTO, SCATTER_FP32, SCATTER_FP64, SCATTER_MFMA_F16, TOTAL_SOURCE,
                   YAMAMOTO, MOCError, check, lib)
from .workloads import Problem

# FSR keys of the tracer: lattice cell * KEY_STRIDE + zone * 64 + sector
# (3D: + axial layer * lattice cells * KEY_STRIDE)
KEY_STRIDE = 4096

__all__ = ["TrackGenerator", "HipSolver", "GPUSolver", "HipLSSolver", "SCALAR_FLUX",
           "FISSION_SOURCE",
           "TOTAL_SOURCE", "EXP_RATIONAL", "EXP_HW", "DIAGONAL", "YAMAMOTO", "GLOBAL",
           "SCATTER_AUTO", "SCATTER_FP64", "SCATTER_MFMA_F16", "SCATTER_FP32", "MOCError"]


def _track_data(tg: "TrackGenerator"):
    """moc_track_data_t of a generator's tracks, and the arrays it points
    into (keep them alive while the struct is in use)."""
    i = tg.info
    td = _lib.TrackData()
    td.num_azim, td.num_polar = i.num_azim, i.num_polar
    td.num_tracks, td.num_segments = i.num_tracks, i.num_segments
    keep = [tg.seg_offset, tg.seg_length, tg.seg_fsr, tg.azim, tg.next_fwd, tg.next_bwd,
            tg.next_fwd_is_fwd, tg.next_bwd_is_fwd, tg.bc_fwd, tg.bc_bwd,
            np.ascontiguousarray(tg.total_weight), tg.sin_theta, tg.phi,
            np.ascontiguousarray(tg.azim_spacing * tg.azim_weight), tg.polar_weight,
            np.ascontiguousarray(tg.seg_start)]
    (td.seg_offset, td.seg_length, td.seg_fsr, td.azim, td.next_fwd, td.next_bwd,
     td.next_fwd_is_fwd, td.next_bwd_is_fwd, td.bc_fwd, td.bc_bwd, td.total_weight,
     td.sin_theta, td.azim_phi, td.azim_weight, td.polar_weight,
     td.seg_start) = [a.ctypes.data for a in keep]
    if getattr(tg, "polar", None) is not None:
        keep += [tg.polar, tg.xy]
        td.polar = tg.polar.ctypes.data
        td.xy = tg.xy.ctypes.data
        td.seg_start = None
    return td, keep


def _path(fname) -> bytes:
    return os.fsencode(fname)


def fsr_centroids(tg: "TrackGenerator") -> np.ndarray:
    """FSR centroids [N, 3] of a TrackGenerator's tracks, as the reference's
    Geometry::getFSRCentroid holds them (TrackGenerator::generateFSRCentroids,
    src/TrackGenerator.cpp:1681-1735); z is 0 for 2D tracks."""
    n = tg.getNumFSRs()
    xyz = np.zeros((n, 3))
    xyz[:, :2] = tg.fsr_centroid.reshape(n, 2)
    if getattr(tg, "fsr_centroid_z", None) is not None:
        xyz[:, 2] = tg.fsr_centroid_z
    return xyz


def write_flux_file(fname, keff: float, xyz: np.ndarray, flux: np.ndarray) -> None:
    """Writes a flux file in the layout of Solver::dumpFSRFluxes
    (src/Solver.cpp:2066-2112) from centroids [N, 3] and fluxes [N, G]."""
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(flux, np.float64).reshape(len(x), -1)
    check(lib().moc_flux_file_write(_path(fname), float(keff), f.shape[1], len(x), _ptr(x),
                                    _ptr(f)))


def read_flux_file(fname):
    """(k_eff, centroids [N, 3], fluxes [N, G]) of a flux file."""
    k, g, n = C.c_double(), C.c_int32(), C.c_int64()
    check(lib().moc_flux_file_read(_path(fname), C.byref(k), C.byref(g), C.byref(n), None, 0,
                                   None, 0))
    xyz = np.zeros((n.value, 3))
    flux = np.zeros((n.value, g.value))
    check(lib().moc_flux_file_read(_path(fname), None, None, None, _ptr(xyz), xyz.size,
                                   _ptr(flux), flux.size))
    return k.value, xyz, flux


def match_centroids(file_xyz: np.ndarray, own_xyz: np.ndarray, tolerance: float = 0.01):
    """(index, max_error): for every own centroid the matched file centroid
    (the rule of Solver::loadFSRFluxes, src/Solver.cpp:2198-2306)."""
    a = np.ascontiguousarray(file_xyz, np.float64).reshape(-1, 3)
    b = np.ascontiguousarray(own_xyz, np.float64).reshape(-1, 3)
    idx = np.zeros(len(b), np.int64)
    err = C.c_double()
    check(lib().moc_fsr_match_centroids(_ptr(a), len(a), _ptr(b), len(b), float(tolerance),
                                        _ptr(idx), C.byref(err)))
    return idx, err.value


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class TrackGenerator:
    """Lays down and ray-traces 2D tracks over a Cartesian pin lattice."""

    def __init__(self, problem: Union[Problem, str], num_threads: int = 0):
        if isinstance(problem, Problem):
            self.problem, self._text = problem, problem.to_text()
        else:
            # the C++ parser is authoritative and reports malformed input
            self._text = problem
            try:
                self.problem = Problem.from_text(problem)
            except (ValueError, KeyError, IndexError):
                self.problem = None
        self._threads = int(num_threads)
        self._h = None
        self.info = None

    def setNumThreads(self, num_threads: int) -> None:
        self._threads = int(num_threads)

    def generateTracks(self) -> None:
        h = C.c_void_p()
        check(lib().moc_tracks_generate(self._text.encode(), self._threads, C.byref(h)))
        self._free()
        self._h = h
        info = _lib.TracksInfo()
        check(lib().moc_tracks_info(h, C.byref(info)))
        self.info = info
        self._export()

    def _export(self) -> None:
        i = self.info
        T, S, N, G, M = i.num_tracks, i.num_segments, i.num_fsrs, i.num_groups, i.num_materials
        na2, npol = i.num_azim // 2, i.num_polar
        self.seg_offset = np.zeros(T + 1, np.int64)
        self.azim = np.zeros(T, np.int32)
        self.xy = np.zeros(T, np.int32)
        self.start = np.zeros((T, 2))
        self.end = np.zeros((T, 2))
        self.next_fwd = np.zeros(T, np.int64)
        self.next_bwd = np.zeros(T, np.int64)
        self.next_fwd_is_fwd = np.zeros(T, np.uint8)
        self.next_bwd_is_fwd = np.zeros(T, np.uint8)
        self.bc_fwd = np.zeros(T, np.uint8)
        self.bc_bwd = np.zeros(T, np.uint8)
        check(lib().moc_tracks_export_tracks(
            self._h, _ptr(self.seg_offset), _ptr(self.azim), _ptr(self.xy), _ptr(self.start),
            _ptr(self.end), _ptr(self.next_fwd), _ptr(self.next_bwd), _ptr(self.next_fwd_is_fwd),
            _ptr(self.next_bwd_is_fwd), _ptr(self.bc_fwd), _ptr(self.bc_bwd)))
        self.seg_length = np.zeros(S)
        self.seg_fsr = np.zeros(S, np.int32)
        check(lib().moc_tracks_export_segments(self._h, _ptr(self.seg_length), _ptr(self.seg_fsr)))
        if i.is_3d:
            self.seg_start = np.zeros((0, 2))
        else:
            self.seg_start = np.zeros((S, 2))
            check(lib().moc_tracks_export_segment_starts(self._h, _ptr(self.seg_start)))
        self.fsr_material = np.zeros(N, np.int32)
        self.fsr_volume = np.zeros(N)
        self.fsr_centroid = np.zeros((N, 2))
        self.fsr_key = np.zeros(N, np.int64)
        check(lib().moc_tracks_export_fsrs(self._h, _ptr(self.fsr_material), _ptr(self.fsr_volume),
                                           _ptr(self.fsr_centroid), _ptr(self.fsr_key)))
        self.phi = np.zeros(na2)
        self.azim_spacing = np.zeros(na2)
        self.azim_weight = np.zeros(na2)
        self.sin_theta = np.zeros(npol)
        self.polar_weight = np.zeros(npol)
        self.total_weight = np.zeros((na2, npol))
        check(lib().moc_tracks_export_quadrature(
            self._h, _ptr(self.phi), _ptr(self.azim_spacing), _ptr(self.azim_weight),
            _ptr(self.sin_theta), _ptr(self.polar_weight), _ptr(self.total_weight)))
        self.sigma_t = np.zeros((M, G))
        self.sigma_s = np.zeros((M, G * G))
        self.nu_sigma_f = np.zeros((M, G))
        self.chi = np.zeros((M, G))
        check(lib().moc_tracks_export_materials(self._h, _ptr(self.sigma_t), _ptr(self.sigma_s),
                                                _ptr(self.nu_sigma_f), _ptr(self.chi)))
        self.polar = None
        if i.is_3d:
            # explicit 3D tracks (TrackGenerator3D, EXPLICIT_3D)
            self.polar = np.zeros(T, np.int32)
            self.start_z = np.zeros(T)
            self.end_z = np.zeros(T)
            self.theta = np.zeros((na2, npol))
            self.polar_spacing = np.zeros((na2, npol))
            self.fsr_centroid_z = np.zeros(N)
            check(lib().moc_tracks_export_tracks_3d(
                self._h, _ptr(self.polar), _ptr(self.start_z), _ptr(self.end_z),
                _ptr(self.theta), _ptr(self.polar_spacing), _ptr(self.fsr_centroid_z)))

    # reference TrackGenerator getters
    def getNumTracks(self) -> int:
        return int(self.info.num_tracks)

    getNum2DTracks = getNumTracks

    def getNumSegments(self) -> int:
        return int(self.info.num_segments)

    def getNumFSRs(self) -> int:
        return int(self.info.num_fsrs)

    def getNumEnergyGroups(self) -> int:
        return int(self.info.num_groups)

    # --- segment files (TrackGenerator.cpp:1388-1638) ----------------------
    def setDumpSegments(self, dump_segments: bool) -> None:
        """TrackGenerator::setDumpSegments: whether dumpSegmentsToFile writes."""
        self._dump_segments = bool(dump_segments)

    def getTestFilename(self, directory: str = "tracks") -> str:
        """Segment-file name for these tracks in `directory`, in the pattern
        of TrackGenerator(3D)::getTestFilename (TrackGenerator.cpp:1353-1371,
        TrackGenerator3D.cpp:1480-1532)."""
        p = self.problem
        if p is None:
            raise MOCError(-2, "the problem text could not be parsed for a file name")
        if p.is_3d:
            return (f"{directory}/3D_{p.num_azim}_azim_{p.num_polar}_polar_"
                    f"{p.azim_spacing:g}x{p.z_spacing:g}_cm_spacing_quad_GAUSS_LEGENDRE_"
                    f"track_EXPLICIT_3D.data")
        return f"{directory}/2D_{p.num_azim}_angles_{p.azim_spacing:g}_cm_spacing.data"

    def dumpSegmentsToFile(self, fname: Optional[str] = None) -> None:
        """TrackGenerator::dumpSegmentsToFile (TrackGenerator.cpp:1388-1501):
        nothing if setDumpSegments(False); otherwise the tracks and segments go
        to `fname` (default getTestFilename(), its directory created)."""
        if not getattr(self, "_dump_segments", True):
            return
        if not self.containsSegments():
            raise MOCError(-2, "Unable to dump Segments to a file since no Segments have "
                               "been generated")
        path = fname or self.getTestFilename()
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        check(lib().moc_tracks_dump_segments(self._h, _path(path)))

    def readSegmentsFromFile(self, fname: Optional[str] = None) -> bool:
        """TrackGenerator::readSegmentsFromFile (TrackGenerator.cpp:1518-1638):
        takes the tracks and segments from `fname` (default getTestFilename())
        instead of generating them: a file dumpSegmentsToFile wrote, or a 2D
        file the reference's own dumpSegmentsToFile wrote (read on this
        problem's track laydown).  False, with nothing changed, if the file
        was written for another problem."""
        h = C.c_void_p()
        check(lib().moc_tracks_read_segments(self._text.encode(),
                                             _path(fname or self.getTestFilename()), C.byref(h)))
        if not h.value:
            return False
        self._free()
        self._h = h
        info = _lib.TracksInfo()
        check(lib().moc_tracks_info(h, C.byref(info)))
        self.info = info
        self._export()
        return True

    def getFSRVolumes(self) -> np.ndarray:
        return self.fsr_volume

    def getFSRCells(self) -> np.ndarray:
        """Geometry::findCellContainingFSR as a table [N]: the cell of each
        FSR as pin * 64 + index, with `pin` the pin universe's index in the
        problem and `index` its cell (cylinders innermost first, the outer
        cell last).  A cylinder's rings and sectors are all of its cell."""
        p = self.problem
        if p is None:
            raise MOCError(-2, "the problem text could not be parsed for the cells")
        key = np.asarray(self.fsr_key, np.int64)
        lat = (key // KEY_STRIDE) % (p.nx * p.ny)          # iy * nx + ix, iy from ymin
        zone = (key % KEY_STRIDE) // 64
        iy, ix = lat // p.nx, lat % p.nx
        pin = np.asarray(p.cells).reshape(p.ny, p.nx)[p.ny - 1 - iy, ix]
        # zone -> cell index: each cylinder covers its rings' zones
        out = np.empty(len(key), np.int32)
        for q, t in enumerate(p.pins):
            first = np.cumsum([0] + [max(1, c.rings) for c in t.cylinders])
            sel = pin == q
            out[sel] = q * 64 + np.searchsorted(first, zone[sel], side="right") - 1
        return out

    def getMaxOpticalLength(self) -> float:
        return float(self.info.max_optical_length)

    def containsSegments(self) -> bool:
        return self._h is not None

    def _free(self) -> None:
        if self._h is not None:
            lib().moc_tracks_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


class HipSolver:
    """MOC source iteration on one MI355X; Solver API of the reference."""

    def __init__(self, track_generator: Optional[TrackGenerator] = None, device: int = -1):
        h = C.c_void_p()
        check(lib().moc_solver_create(int(device), C.byref(h)))
        self._h = h
        self._tg = None
        self._num_fsrs = 0
        self._num_groups = 0
        if track_generator is not None:
            self.setTrackGenerator(track_generator)

    # --- setup -----------------------------------------------------------
    def setTrackGenerator(self, tg: TrackGenerator) -> None:
        if not tg.containsSegments():
            raise MOCError(-2, "Unable to set the TrackGenerator for the Solver since the "
                               "TrackGenerator has not yet generated tracks")
        i = tg.info
        td, keep = _track_data(tg)
        check(lib().moc_solver_set_materials(
            self._h, i.num_materials, i.num_groups, _ptr(tg.sigma_t), _ptr(tg.sigma_s),
            _ptr(tg.nu_sigma_f), _ptr(tg.chi)))
        check(lib().moc_solver_set_fsrs(self._h, i.num_fsrs, _ptr(tg.fsr_material),
                                        _ptr(tg.fsr_volume)))
        check(lib().moc_solver_set_tracks(self._h, C.byref(td)))
        del keep
        self._tg = tg
        self._num_fsrs = int(i.num_fsrs)
        self._num_groups = int(i.num_groups)
        self.setConvergenceThreshold(i.tolerance)
        self.setFSRCentroids(fsr_centroids(tg))

    def setArrays(self, *, num_azim, num_polar, seg_offset, seg_length, seg_fsr, azim,
                  next_fwd, next_bwd, next_fwd_is_fwd, next_bwd_is_fwd, bc_fwd, bc_bwd,
                  total_weight, sin_theta, sigma_t, sigma_s, nu_sigma_f, chi, fsr_material,
                  fsr_volume) -> None:
        """Set tracks/materials/FSRs from raw arrays (e.g. tracks produced by
        another tracer).  sigma_s is dest-major [M, G*G]."""
        c = np.ascontiguousarray
        arrs = dict(seg_offset=c(seg_offset, np.int64), seg_length=c(seg_length, np.float64),
                    seg_fsr=c(seg_fsr, np.int32), azim=c(azim, np.int32),
                    next_fwd=c(next_fwd, np.int64), next_bwd=c(next_bwd, np.int64),
                    next_fwd_is_fwd=c(next_fwd_is_fwd, np.uint8),
                    next_bwd_is_fwd=c(next_bwd_is_fwd, np.uint8), bc_fwd=c(bc_fwd, np.uint8),
                    bc_bwd=c(bc_bwd, np.uint8), total_weight=c(total_weight, np.float64),
                    sin_theta=c(sin_theta, np.float64))
        td = _lib.TrackData()
        td.num_azim, td.num_polar = int(num_azim), int(num_polar)
        td.num_tracks, td.num_segments = len(arrs["azim"]), len(arrs["seg_length"])
        for k, v in arrs.items():
            setattr(td, k, v.ctypes.data)
        st, ss, nsf, ch = (c(x, np.float64) for x in (sigma_t, sigma_s, nu_sigma_f, chi))
        M, G = st.shape
        check(lib().moc_solver_set_materials(self._h, M, G, _ptr(st), _ptr(ss), _ptr(nsf),
                                             _ptr(ch)))
        fm, fv = c(fsr_material, np.int32), c(fsr_volume, np.float64)
        check(lib().moc_solver_set_fsrs(self._h, len(fm), _ptr(fm), _ptr(fv)))
        check(lib().moc_solver_set_tracks(self._h, C.byref(td)))
        self._num_fsrs, self._num_groups = len(fm), G
        self._tg = None
        self._psi_size = 2 * len(arrs["azim"]) * (int(num_polar) // 2) * G

    def setFSRCentroids(self, xyz: np.ndarray) -> None:
        """FSR centroids [N, 3] that name the FSRs in flux files (set from the
        TrackGenerator by setTrackGenerator)."""
        a = np.ascontiguousarray(xyz, np.float64)
        check(lib().moc_solver_set_fsr_centroids(self._h, _ptr(a), a.size))

    # --- flux files (Solver.cpp:2056-2318) --------------------------------
    def dumpFSRFluxes(self, fname: str) -> None:
        """Solver::dumpFSRFluxes (src/Solver.cpp:2066-2112): k_eff, FSR
        centroids and the current scalar fluxes, binary."""
        check(lib().moc_solver_dump_fsr_fluxes(self._h, _path(fname)))

    def loadFSRFluxes(self, fname: str, assign_k_eff: bool = False,
                      tolerance: float = 0.01) -> float:
        """Solver::loadFSRFluxes (src/Solver.cpp:2135-2318): the fluxes of
        the file's FSRs nearest to this geometry's centroids become the current
        fluxes.  Returns the largest centroid mismatch (cm)."""
        err = C.c_double()
        check(lib().moc_solver_load_fsr_fluxes(self._h, _path(fname), int(bool(assign_k_eff)),
                                               float(tolerance), C.byref(err)))
        return err.value

    def loadInitialFSRFluxes(self, fname: Optional[str]) -> None:
        """Solver::loadInitialFSRFluxes (src/Solver.cpp:2120-2123): eigenvalue
        and source solves restart from the file's fluxes and k_eff after 30
        sweeps on their sources.  None switches it off."""
        check(lib().moc_solver_load_initial_fsr_fluxes(self._h, _path(fname or "")))

    def setResidualByReference(self, fname: Optional[str]) -> None:
        """Solver::setResidualByReference (src/Solver.cpp:2056-2059):
        residuals against the file's fluxes.  None switches it off."""
        check(lib().moc_solver_set_residual_by_reference(self._h, _path(fname or "")))

    def setSolverMode(self, solver_mode: int) -> None:
        """Solver::setSolverMode (src/Solver.cpp:597-602): FORWARD or ADJOINT
        (transposed scattering and fission matrices for the sources)."""
        check(lib().moc_solver_set_solver_mode(self._h, int(solver_mode)))

    def setRestartStatus(self, is_restart: bool) -> None:
        """Solver::setRestartStatus (src/Solver.cpp:610-615): following solves
        continue from this solver's k_eff, angular and scalar fluxes."""
        check(lib().moc_solver_set_restart_status(self._h, int(bool(is_restart))))

    def setMaterials(self, sigma_t, sigma_s, nu_sigma_f, chi) -> None:
        """Replace the cross sections (sigma_t/nu_sigma_f/chi [M, G], sigma_s
        [M, G*G] dest-major), e.g. a perturbed copy of the geometry's materials
        (Material::setSigma*, src/Material.cpp:628-979)."""
        c = np.ascontiguousarray
        st, ss, nsf, ch = (c(x, np.float64) for x in (sigma_t, sigma_s, nu_sigma_f, chi))
        M, G = st.shape
        check(lib().moc_solver_set_materials(self._h, M, G, _ptr(st), _ptr(ss), _ptr(nsf),
                                             _ptr(ch)))
        self._num_groups = int(G)

    def setConvergenceThreshold(self, threshold: float) -> None:
        check(lib().moc_solver_set_convergence_threshold(self._h, float(threshold)))
        self._threshold = float(threshold)

    # --- reference getters (src/Solver.h:443-462) ---------------------------
    def getConvergenceThreshold(self) -> float:
        return getattr(self, "_threshold", 1e-5)

    def getTrackGenerator(self):
        return self._tg

    def _need_tracks(self):
        if self._tg is None:
            raise MOCError(-2, "the solver has no TrackGenerator")
        return self._tg

    def getNumPolarAngles(self) -> int:
        return int(self._need_tracks().info.num_polar)

    def getMaxOpticalLength(self) -> float:
        return self._need_tracks().getMaxOpticalLength()

    def is3D(self) -> bool:
        return bool(self._need_tracks().info.is_3d)

    def getFSRVolume(self, fsr_id: int) -> float:
        """Solver::getFSRVolume, with its range check (src/Solver.cpp:192-203)."""
        tg = self._need_tracks()
        if not 0 <= int(fsr_id) < tg.getNumFSRs():
            raise MOCError(-1, f"Unable to get the volume for FSR {fsr_id} since the max "
                               f"FSR ID = {tg.getNumFSRs() - 1}")
        return float(tg.fsr_volume[int(fsr_id)])

    def getTotalTime(self) -> float:
        return self.getTiming()["total_seconds"]

    def isUsingDoublePrecision(self) -> bool:
        """Angular fluxes and the sweep are FP32 (per-FSR arithmetic FP64)."""
        return False

    def isUsingExponentialInterpolation(self) -> bool:
        return False

    def useExponentialIntrinsic(self) -> None:
        """Solver::useExponentialIntrinsic: the hardware exponential."""
        self.setExponential(EXP_HW)

    def setNumThreads(self, num_threads: int) -> None:
        """Accepted for API compatibility; the sweep runs on the GPU."""
        if int(num_threads) <= 0:
            raise MOCError(-1, f"Unable to set the number of threads to {num_threads} "
                               "since it is less than or equal to 0")

    def setExponential(self, mode: int) -> None:
        check(lib().moc_solver_set_exponential(self._h, int(mode)))

    def allowNegativeFluxes(self, allow: bool) -> None:
        check(lib().moc_solver_allow_negative_fluxes(self._h, int(bool(allow))))

    def setScatterPrecision(self, mode: int = SCATTER_AUTO) -> int:
        """Arithmetic of the scattering-source contraction: SCATTER_AUTO
        (matrix cores from 32 groups on, FP64 below), SCATTER_FP64,
        SCATTER_MFMA_F16 (fp16 MFMA on hi/lo split operands, FP32
        accumulation) or SCATTER_FP32.  Returns the path in use."""
        used = C.c_int()
        check(lib().moc_solver_set_scatter_precision(self._h, int(mode), C.byref(used)))
        return used.value

    def computeFSRSources(self, iteration: int = 0) -> np.ndarray:
        """CPUSolver::computeFSRSources (src/CPUSolver.cpp:1939-2023) from the
        current flux and keff; returns the sources per steradian [N, G]."""
        n = self._num_fsrs * self._num_groups
        out = np.empty(n)
        check(lib().moc_solver_compute_fsr_sources(self._h, int(iteration), _ptr(out), n))
        return out.reshape(self._num_fsrs, self._num_groups)

    def stabilizeTransport(self, stabilization_factor: float,
                           stabilization_type: int = DIAGONAL) -> None:
        """Solver::stabilizeTransport (src/Solver.cpp:676-681)."""
        check(lib().moc_solver_stabilize_transport(self._h, float(stabilization_factor),
                                                   int(stabilization_type)))

    def setKeffFromNeutronBalance(self) -> None:
        """Solver::setKeffFromNeutronBalance (src/Solver.cpp:2047-2049)."""
        check(lib().moc_solver_set_keff_from_neutron_balance(self._h, 1))

    def useLinearSource(self, on: bool = True) -> None:
        """Linear source in every FSR (the reference's CPULSSolver)."""
        check(lib().moc_solver_set_linear_source(self._h, int(bool(on))))

    def getFluxMoments(self) -> np.ndarray:
        """Normalised scalar-flux moments [N, G, 2] of the last linear-source solve."""
        out = np.zeros(self._num_fsrs * self._num_groups * 2)
        check(lib().moc_solver_get_flux_moments(self._h, _ptr(out), out.size))
        return out.reshape(self._num_fsrs, self._num_groups, 2)

    def getFSRCentroids(self) -> np.ndarray:
        out = np.zeros(self._num_fsrs * 2)
        check(lib().moc_solver_get_fsr_centroids(self._h, _ptr(out), out.size))
        return out.reshape(self._num_fsrs, 2)

    def fissionTransportSweep(self) -> None:
        """Solver::fissionTransportSweep (src/Solver.cpp:1282-1286)."""
        check(lib().moc_solver_fission_transport_sweep(self._h))

    def scatterTransportSweep(self) -> None:
        """Solver::scatterTransportSweep (src/Solver.cpp:1293-1297)."""
        check(lib().moc_solver_scatter_transport_sweep(self._h))

    def setFixedSourceByFSR(self, fsr_id: int, group: int, source: float) -> None:
        check(lib().moc_solver_set_fixed_source_by_fsr(self._h, int(fsr_id), int(group),
                                                       float(source)))

    def setFixedSourceByMaterial(self, material, group: int, source: float) -> None:
        """Solver::setFixedSourceByMaterial: `material` is the index of the
        material (or its name in the problem description)."""
        if isinstance(material, str):
            names = [m.name for m in self._tg.problem.materials] if self._tg is not None and \
                self._tg.problem is not None else []
            if material not in names:
                raise MOCError(-1, f"unknown material {material!r}")
            material = names.index(material)
        check(lib().moc_solver_set_fixed_source_by_material(self._h, int(material), int(group),
                                                            float(source)))

    def setFixedSourceByCell(self, cell, group: int, source: float) -> None:
        """Solver::setFixedSourceByCell: `cell` is (pin, index), the pin
        universe's index in the problem and the cell's index in it
        (cylinders innermost first, then the outer cell; a cylinder's rings
        and sectors are all of its cell).  Applied when a solve starts, before
        the material sources, as Solver::initializeFixedSources does."""
        if self._tg is None:
            raise MOCError(-2, "setFixedSourceByCell needs the solver's TrackGenerator")
        if not getattr(self, "_cells_set", False):
            cells = np.ascontiguousarray(self._tg.getFSRCells(), np.int32)
            check(lib().moc_solver_set_fsr_cells(self._h, _ptr(cells), len(cells)))
            self._cells_set = True
        pin, index = cell
        check(lib().moc_solver_set_fixed_source_by_cell(self._h, int(pin) * 64 + int(index),
                                                        int(group), float(source)))

    def setChiSpectrumMaterial(self, material) -> None:
        """Solver::setChiSpectrumMaterial: solves start from that material's
        chi (index or name; None for the flat guess)."""
        if material is None:
            material = -1
        elif isinstance(material, str):
            names = [m.name for m in self._tg.problem.materials] if self._tg is not None and \
                self._tg.problem is not None else []
            if material not in names:
                raise MOCError(-1, f"unknown material {material!r}")
            material = names.index(material)
        check(lib().moc_solver_set_chi_spectrum_material(self._h, int(material)))

    def resetFixedSources(self) -> None:
        """CPUSolver::resetFixedSources: every fixed source zero."""
        check(lib().moc_solver_reset_fixed_sources(self._h))

    # --- solves ----------------------------------------------------------
    def computeEigenvalue(self, max_iters: int = 1000, res_type: int = FISSION_SOURCE) -> None:
        check(lib().moc_solver_compute_eigenvalue(self._h, int(max_iters), int(res_type)))

    def computeFlux(self, max_iters: int = 1000, only_fixed_source: bool = True) -> None:
        check(lib().moc_solver_compute_flux(self._h, int(max_iters), int(bool(only_fixed_source))))

    def computeSource(self, max_iters: int = 1000, k_eff: float = 1.0,
                      res_type: int = TOTAL_SOURCE) -> None:
        check(lib().moc_solver_compute_source(self._h, int(max_iters), float(k_eff),
                                              int(res_type)))

    # --- results ---------------------------------------------------------
    def getKeff(self) -> float:
        v = C.c_double()
        check(lib().moc_solver_get_keff(self._h, C.byref(v)))
        return v.value

    def getNumIterations(self) -> int:
        v = C.c_int32()
        check(lib().moc_solver_get_num_iterations(self._h, C.byref(v)))
        return v.value

    def getFlux(self, fsr_id: int, group: int) -> float:
        v = C.c_double()
        check(lib().moc_solver_get_flux(self._h, int(fsr_id), int(group), C.byref(v)))
        return v.value

    def getFluxes(self, num_fluxes: Optional[int] = None) -> np.ndarray:
        n = self._num_fsrs * self._num_groups if num_fluxes is None else int(num_fluxes)
        out = np.zeros(n)
        check(lib().moc_solver_get_fluxes(self._h, _ptr(out), n))
        return out

    def setFluxes(self, fluxes: np.ndarray) -> None:
        a = np.ascontiguousarray(fluxes, np.float64).ravel()
        check(lib().moc_solver_set_fluxes(self._h, _ptr(a), a.size))

    def computeFSRFissionRates(self, num_fsrs: Optional[int] = None, nu: bool = True) -> np.ndarray:
        """CPUSolver::computeFSRFissionRates (src/CPUSolver.cpp:2825-2886):
        volume-weighted nu-sigma-f rates.  Problems carry nu-sigma-f only, so
        nu=False (sigma-f rates) is refused."""
        if not nu:
            raise MOCError(-1, "sigma-f is not part of the problem description; use nu=True")
        n = self._num_fsrs if num_fsrs is None else int(num_fsrs)
        out = np.zeros(n)
        check(lib().moc_solver_compute_fsr_fission_rates(self._h, _ptr(out), n))
        return out

    def getTiming(self):
        sw, tot, ni = C.c_double(), C.c_double(), C.c_int64()
        check(lib().moc_solver_get_timing(self._h, C.byref(sw), C.byref(tot), C.byref(ni)))
        return {"iteration_seconds": sw.value, "total_seconds": tot.value,
                "num_integrations": ni.value}

    def printTimerReport(self) -> None:
        t = self.getTiming()
        n = max(t["num_integrations"], 1)
        print(f"Total time to solution ...................... {t['total_seconds']:.4e} sec")
        print(f"Device time of source iterations ............. {t['iteration_seconds']:.4e} sec")
        print(f"Integrations ................................. {t['num_integrations']}")
        print(f"Integration time (device) .................... {t['iteration_seconds'] / n:.4e} sec")

    # --- kernel-level entry points --------------------------------------
    def transportSweep(self, Q: np.ndarray, psi: np.ndarray):
        """One sweep on given reduced sources; returns (tally, psi_out)."""
        q = np.ascontiguousarray(Q, np.float64).ravel()
        p = np.array(psi, np.float32, copy=True).ravel()
        if self._tg is None and not hasattr(self, "_psi_size"):
            raise MOCError(-2, "tracks have not been set")
        pol = 1 if (self._tg is not None and self._tg.info.is_3d) else (
            self._tg.info.num_polar // 2 if self._tg is not None else 0)
        n_psi = (2 * self._tg.getNumTracks() * pol * self._num_groups
                 if self._tg is not None else self._psi_size)
        if q.size != self._num_fsrs * self._num_groups or p.size != n_psi:
            raise MOCError(-1, f"transportSweep expects Q of {self._num_fsrs * self._num_groups} "
                               f"and psi of {n_psi} values")
        tally = np.zeros(q.size)
        check(lib().moc_solver_transport_sweep(self._h, _ptr(q), _ptr(p), _ptr(tally)))
        return tally, p

    SHAPE_KEYS = ("groups_per_lane", "lanes_per_track", "tracks_per_wave", "block",
                  "tally_copies", "lds_groups", "lds_block", "lds_blocks_per_sub", "lds_q",
                  "lds_flush", "seg_store")

    def setLaunchShape(self, **shape) -> None:
        """Override the sweep launch shape (keys of getLaunchShape())."""
        cur = self.getLaunchShape()
        unknown = set(shape) - set(self.SHAPE_KEYS)
        if unknown:
            raise MOCError(-1, f"unknown launch shape keys {sorted(unknown)}")
        cur.update(shape)
        a = np.array([cur[k] for k in self.SHAPE_KEYS], np.int32)
        check(lib().moc_solver_set_launch_shape(self._h, _ptr(a)))

    def getLaunchShape(self) -> dict:
        a = np.zeros(len(self.SHAPE_KEYS), np.int32)
        check(lib().moc_solver_get_launch_shape(self._h, _ptr(a)))
        return dict(zip(self.SHAPE_KEYS, (int(x) for x in a)))

    def runIterations(self, num_iterations: int, time_sweep: bool = True):
        """Benchmark form: num_iterations full source iterations from the
        current state, no stopping rule.  Returns (mean device seconds per
        iteration, mean seconds of the sweep kernel alone or None), the latter
        from up to 8 further iterations with events around the sweep."""
        v, w = C.c_double(), C.c_double()
        check(lib().moc_solver_run_iterations(self._h, int(num_iterations), C.byref(v),
                                              C.byref(w) if time_sweep else None))
        return v.value, (w.value if time_sweep else None)

    def timeSweeps(self, repeats: int = 10) -> float:
        v = C.c_double()
        check(lib().moc_solver_time_sweeps(self._h, int(repeats), C.byref(v)))
        return v.value

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None:
                lib().moc_solver_destroy(self._h)
                self._h = None
        except Exception:
            pass


GPUSolver = HipSolver


class HipLSSolver(HipSolver):
    """Linear-source solver (the reference's CPULSSolver) on one MI355X."""

    def __init__(self, track_generator: Optional[TrackGenerator] = None, device: int = -1):
        super().__init__(track_generator, device)
        if track_generator is not None:
            self.useLinearSource(True)

    def setTrackGenerator(self, tg: TrackGenerator) -> None:
        super().setTrackGenerator(tg)
        self.useLinearSource(True)
