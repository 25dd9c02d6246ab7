"""Continuous-batching serving of eigenvalue solves on one MI355X.

A geometry is uploaded once (``add_geometry``); each request brings its own
cross sections and is admitted as batch rows and device memory free up.  The
solves in flight of one geometry advance together, one launch per iteration
step for all of them (captured graphs of iteration steps); batches of
different geometries overlap on the device.  A result is the same alone and
in the mix (bitwise on the 7-group path: integer fixed-point tallies,
fixed-order reductions).
There is no counterpart in the reference (its GPUSolver runs one blocking
solve); each request follows Solver::computeEigenvalue
(src/Solver.cpp:1542-1689).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._lib import FISSION_SOURCE, check, lib
from .solver import TrackGenerator, _ptr, _track_data
from .workloads import Problem

ZERO_SIGMA_T = 1e-6      # reference ZERO_SIGMA_T (src/Material.cpp:628-650)
FLT_EPSILON = 1e-12


def material_arrays(problem: Problem) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """(sigma_t [M,G], sigma_s [M,G*G] dest-major, nu_sigma_f [M,G], chi [M,G])
    as the solver consumes them: sigma_t floored at ZERO_SIGMA_T, scattering
    transposed from the origin-major input, chi normalised to sum 1 (the same
    steps as the problem parser, src/Material.cpp:628-650, 721-732, 894-910)."""
    G = problem.num_groups
    st, ss, nsf, chi = [], [], [], []
    for m in problem.materials:
        t = np.array(m.sigma_t, np.float64)
        t = np.where(np.abs(t) < ZERO_SIGMA_T, ZERO_SIGMA_T, t)
        st.append(t)
        ss.append(np.asarray(m.sigma_s, np.float64).reshape(G, G).T.ravel())
        nsf.append(np.asarray(m.nu_sigma_f, np.float64))
        c = np.array(m.chi, np.float64)
        cs = float(np.sum(c))
        if abs(cs) >= FLT_EPSILON:
            c = c / cs
        chi.append(c)
    c = np.ascontiguousarray
    return c(st), c(ss), c(nsf), c(chi)


@dataclass
class Result:
    keff: float
    num_iterations: int
    latency_s: float
    queue_s: float
    solve_s: float
    flux: Optional[np.ndarray] = None   # [N, G] when requested


class ServeEngine:
    """Request loop on one GPU (moc_engine_* C ABI)."""

    def __init__(self, device: int = 0, max_active: int = 16, mem_budget_bytes: int = 0):
        h = C.c_void_p()
        check(lib().moc_engine_create(int(device), int(max_active), int(mem_budget_bytes),
                                      C.byref(h)))
        self._h = h
        self._geo = {}

    def close(self) -> None:
        if self._h:
            lib().moc_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_geometry(self, tg: TrackGenerator) -> int:
        """Uploads the generator's tracks and FSRs (its materials fix the group
        and material counts).  Returns the geometry id."""
        i = tg.info
        td, keep = _track_data(tg)
        gid = C.c_int32()
        check(lib().moc_engine_add_geometry(
            self._h, C.byref(td), i.num_fsrs, _ptr(tg.fsr_material), _ptr(tg.fsr_volume),
            i.num_materials, i.num_groups, _ptr(tg.sigma_t), _ptr(tg.sigma_s),
            _ptr(tg.nu_sigma_f), _ptr(tg.chi), C.byref(gid)))
        del keep
        self._geo[gid.value] = (int(i.num_fsrs), int(i.num_groups), int(i.num_materials))
        return gid.value

    def submit(self, geo: int, materials, max_iters: int = 1000,
               res_type: int = FISSION_SOURCE, tolerance: float = 1e-5,
               want_flux: bool = False) -> int:
        """materials: a Problem (its materials are used) or the tuple returned
        by material_arrays."""
        if isinstance(materials, Problem):
            materials = material_arrays(materials)
        st, ss, nsf, chi = (np.ascontiguousarray(a, np.float64) for a in materials)
        M, G = st.shape
        rid = C.c_int64()
        check(lib().moc_engine_submit(self._h, int(geo), M, G, _ptr(st), _ptr(ss), _ptr(nsf),
                                      _ptr(chi), int(max_iters), int(res_type),
                                      float(tolerance), int(bool(want_flux)), C.byref(rid)))
        self._want = getattr(self, "_want", {})
        self._want[rid.value] = (geo, want_flux)
        return rid.value

    def wait(self, request: int, timeout: float = -1.0) -> Result:
        geo, want = self._want.pop(request, (None, False))
        r = _lib.RequestResult()
        flux = None
        n = 0
        if want and geo is not None:
            N, G, _ = self._geo[geo]
            flux = np.zeros(N * G)
            n = N * G
        rc = lib().moc_engine_wait(self._h, int(request), float(timeout), C.byref(r), _ptr(flux), n)
        if rc != 0:
            self._want[request] = (geo, want)  # still owned by the engine
        check(rc)
        return Result(r.keff, r.num_iterations, r.latency_s, r.queue_s, r.solve_s,
                      None if flux is None else flux.reshape(-1, self._geo[geo][1]))

    def stats(self) -> dict:
        c, f, it, p = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        check(lib().moc_engine_stats(self._h, C.byref(c), C.byref(f), C.byref(it), C.byref(p)))
        return {"completed": c.value, "failed": f.value, "iterations": it.value,
                "pending": p.value}

    def memory(self) -> dict:
        """Device bytes of the solves: budget, held now (live or idle
        solvers of all batches) and the most held at once."""
        b, h, m = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().moc_engine_memory(self._h, C.byref(b), C.byref(h), C.byref(m)))
        return {"budget": b.value, "held": h.value, "held_max": m.value}
