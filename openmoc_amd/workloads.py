"""Problem definitions for the MOC source iteration.

A *problem* is everything the transport sweep and the per-FSR steps need:
multigroup materials, a 2D Cartesian grid of pin cells (each cell a pin
template of concentric cylinders with ring/sector subdivisions, or a
homogeneous block), boundary conditions, the track laydown parameters and the
solver settings.

The same text serialisation ("moc-problem 1") is read by this package's C++
tracer and by the reference driver under ``oracle/``, so both sides start from
identical bytes of cross-section data.

Geometry conventions follow the reference CSG:
  * the domain is centred on the origin, ``x in [-nx*px/2, nx*px/2]``;
  * lattice cells are listed row-major from the *upper-left* corner (largest y
    first), as ``Lattice::setUniverses`` expects (src/Universe.cpp:1604-1619);
  * fuel rings are equal-area subdivisions of the annulus they split
    (src/Cell.cpp:1483-1517);
  * sector planes have normals at ``i*2pi/n + pi/4`` (src/Cell.cpp:1335-1350);
  * scattering matrices are given origin-major, ``s[orig*G + dest]``, the
    layout ``Material::setSigmaS`` takes (src/Material.cpp:721-732).

Synthetic cross sections are drawn from ``numpy.random.default_rng(seed)``:
  * sigma_t = 0.5 * exp(0.3 * N(0,1))  (median 0.5 /cm, log-sd 0.3);
  * scattering is downscatter-dominated, upscatter only inside the thermal
    block (groups 6-7 of 7; a band in the thermal block at 70 groups), with a
    per-origin scattering ratio <= 0.95;
  * fuel nu*sigma_f ~ U(0.005, 0.3), chi on the first three groups;
  * moderator is non-fissile.
"""

from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence

import numpy as np

BC_CODES = {"V": 0, "R": 1, "P": 2}


@dataclasses.dataclass
class MaterialXS:
    name: str
    sigma_t: np.ndarray          # [G]
    sigma_s: np.ndarray          # [G*G], origin-major s[orig*G + dest]
    nu_sigma_f: np.ndarray       # [G]
    chi: np.ndarray              # [G]

    @property
    def num_groups(self) -> int:
        return int(self.sigma_t.shape[0])

    @property
    def fissionable(self) -> bool:
        return bool(np.any(self.nu_sigma_f != 0.0))


@dataclasses.dataclass
class Cylinder:
    radius: float
    material: int
    rings: int = 1
    sectors: int = 0


@dataclasses.dataclass
class PinTemplate:
    """Concentric cylinders (innermost first) inside a square lattice cell."""
    cylinders: List[Cylinder]
    outer_material: int
    outer_sectors: int = 0


@dataclasses.dataclass
class Problem:
    name: str
    materials: List[MaterialXS]
    pins: List[PinTemplate]
    nx: int
    ny: int
    pitch_x: float
    pitch_y: float
    cells: np.ndarray            # [ny, nx] pin ids, row 0 = top (max y)
    bc: Sequence[str] = ("R", "R", "R", "R")   # xmin xmax ymin ymax
    num_azim: int = 4
    azim_spacing: float = 0.1
    num_polar: int = 6
    quadrature: str = "TY"
    tolerance: float = 1e-5
    max_iters: int = 1000
    res_type: str = "FISSION_SOURCE"
    # axially extruded 3D problem (height > 0): explicit 3D tracks with the
    # Gauss-Legendre polar set, `layers` equal axial FSR layers
    height: float = 0.0
    layers: int = 1
    z_spacing: float = 0.0
    bc_z: Sequence[str] = ("R", "R")             # zmin zmax

    @property
    def is_3d(self) -> bool:
        return self.height > 0

    @property
    def num_groups(self) -> int:
        return self.materials[0].num_groups

    @property
    def width_x(self) -> float:
        return self.nx * self.pitch_x

    @property
    def width_y(self) -> float:
        return self.ny * self.pitch_y

    # ------------------------------------------------------------------
    def to_text(self) -> str:
        g = self.num_groups
        out = ["moc-problem 1", f"name {self.name}", f"groups {g}"]
        fmt = lambda arr: " ".join(repr(float(v)) for v in np.asarray(arr).ravel())
        for mid, m in enumerate(self.materials):
            out.append(f"material {mid} {m.name}")
            out.append("sigma_t " + fmt(m.sigma_t))
            out.append("sigma_s " + fmt(m.sigma_s))
            out.append("nu_sigma_f " + fmt(m.nu_sigma_f))
            out.append("chi " + fmt(m.chi))
            out.append("endmaterial")
        for pid, p in enumerate(self.pins):
            parts = [f"pin {pid} {len(p.cylinders)}"]
            for c in p.cylinders:
                parts.append(f"{c.radius!r} {c.material} {c.rings} {c.sectors}")
            parts.append(f"{p.outer_material} {p.outer_sectors}")
            out.append(" ".join(parts))
        out.append(f"grid {self.nx} {self.ny} {self.pitch_x!r} {self.pitch_y!r}")
        out.append("cells " + " ".join(str(int(v)) for v in np.asarray(self.cells).ravel()))
        out.append("bc " + " ".join(self.bc))
        out.append(f"tracks {self.num_azim} {self.azim_spacing!r} {self.num_polar} {self.quadrature}")
        if self.is_3d:
            out.append(f"axial {self.height!r} {self.layers} {self.z_spacing!r} "
                       f"{self.bc_z[0]} {self.bc_z[1]}")
        out.append(f"solver {self.tolerance!r} {self.max_iters} {self.res_type}")
        out.append("end")
        return "\n".join(out) + "\n"

    @staticmethod
    def from_text(text: str) -> "Problem":
        lines = [ln.split() for ln in text.splitlines() if ln.strip() and not ln.startswith("#")]
        if lines[0][:2] != ["moc-problem", "1"]:
            raise ValueError("not a moc-problem 1 file")
        kw: Dict[str, object] = {}
        mats: List[MaterialXS] = []
        pins: List[PinTemplate] = []
        cur: Dict[str, object] = {}
        for tok in lines[1:]:
            key = tok[0]
            if key == "name":
                kw["name"] = tok[1]
            elif key == "groups":
                pass
            elif key == "material":
                cur = {"name": tok[2]}
            elif key in ("sigma_t", "sigma_s", "nu_sigma_f", "chi"):
                cur[key] = np.array([float(v) for v in tok[1:]], dtype=np.float64)
            elif key == "endmaterial":
                mats.append(MaterialXS(**cur))  # type: ignore[arg-type]
            elif key == "pin":
                n = int(tok[2])
                cyl = []
                for i in range(n):
                    r, m, rings, sec = tok[3 + 4 * i: 7 + 4 * i]
                    cyl.append(Cylinder(float(r), int(m), int(rings), int(sec)))
                om, osec = tok[3 + 4 * n: 5 + 4 * n]
                pins.append(PinTemplate(cyl, int(om), int(osec)))
            elif key == "grid":
                kw.update(nx=int(tok[1]), ny=int(tok[2]), pitch_x=float(tok[3]), pitch_y=float(tok[4]))
            elif key == "cells":
                kw["cells"] = np.array([int(v) for v in tok[1:]], dtype=np.int32)
            elif key == "bc":
                kw["bc"] = tuple(tok[1:5])
            elif key == "tracks":
                kw.update(num_azim=int(tok[1]), azim_spacing=float(tok[2]),
                          num_polar=int(tok[3]), quadrature=tok[4])
            elif key == "axial":
                kw.update(height=float(tok[1]), layers=int(tok[2]), z_spacing=float(tok[3]),
                          bc_z=(tok[4], tok[5]))
            elif key == "solver":
                kw.update(tolerance=float(tok[1]), max_iters=int(tok[2]), res_type=tok[3])
            elif key == "end":
                break
            else:
                raise ValueError(f"unknown key {key!r}")
        kw["cells"] = np.asarray(kw["cells"]).reshape(int(kw["ny"]), int(kw["nx"]))  # type: ignore[arg-type]
        return Problem(materials=mats, pins=pins, **kw)  # type: ignore[arg-type]


# ----------------------------------------------------------------------
# Synthetic cross sections
# ----------------------------------------------------------------------

def _scatter_matrix(rng: np.random.Generator, sigma_t: np.ndarray,
                    upscatter_from: int, band: Optional[int]) -> np.ndarray:
    """Origin-major scattering matrix with row sums <= 0.95 sigma_t.

    Origins below ``upscatter_from`` only scatter to themselves and to higher
    (slower) groups; origins inside the thermal block may also scatter up by
    at most ``band`` groups (``None``: anywhere inside the block).
    """
    g_n = sigma_t.shape[0]
    s = np.zeros((g_n, g_n))
    for g in range(g_n):
        ratio = rng.uniform(0.55, 0.95)
        total = ratio * sigma_t[g]
        dest = list(range(g, g_n))
        if g >= upscatter_from:
            lo = upscatter_from if band is None else max(upscatter_from, g - band)
            dest = list(range(lo, g)) + dest
        dest = np.array(dest)
        w = np.zeros(dest.shape[0])
        for i, d in enumerate(dest):
            if d == g:
                w[i] = rng.uniform(4.0, 12.0)          # self-scatter dominates
            elif d > g:
                w[i] = math.exp(-0.7 * (d - g)) * rng.uniform(0.5, 1.5)
            else:
                w[i] = 0.15 * math.exp(-0.7 * (g - d)) * rng.uniform(0.5, 1.5)
        w /= w.sum()
        s[g, dest] = total * w
    return s.ravel()


def synthetic_material(rng: np.random.Generator, num_groups: int, name: str,
                       fissile: bool) -> MaterialXS:
    sigma_t = 0.5 * np.exp(0.3 * rng.standard_normal(num_groups))
    if num_groups == 7:
        upscatter_from, band = 5, None                  # groups 6-7 (1-based)
    else:
        upscatter_from, band = int(round(0.65 * num_groups)), 8
    sigma_s = _scatter_matrix(rng, sigma_t, upscatter_from, band)
    if fissile:
        nu_sigma_f = rng.uniform(0.005, 0.3, num_groups)
        chi = np.zeros(num_groups)
        chi[:3] = rng.uniform(0.2, 1.0, 3)
        chi /= chi.sum()
    else:
        nu_sigma_f = np.zeros(num_groups)
        chi = np.zeros(num_groups)
    return MaterialXS(name, sigma_t, sigma_s, nu_sigma_f, chi)


def perturb_material(m: MaterialXS, rng: np.random.Generator, frac: float = 0.10) -> MaterialXS:
    """Per-group multiplicative perturbation (+-frac) of every reaction.

    Total and scattering move together so the scattering ratio stays <= 0.95.
    """
    f = 1.0 + rng.uniform(-frac, frac, m.num_groups)
    g_n = m.num_groups
    s = m.sigma_s.reshape(g_n, g_n) * f[:, None]
    return MaterialXS(m.name, m.sigma_t * f, s.ravel(),
                      m.nu_sigma_f * (1.0 + rng.uniform(-frac, frac, g_n)), m.chi.copy())


# C5G7 7-group UO2 and water cross sections (OECD/NEA C5G7 benchmark data, as
# tabulated in the reference's sample-input/c5g7-mgxs-hdf5.py:30-55, 236-252).
C5G7_UO2 = MaterialXS(
    "UO2",
    np.array([1.779490E-01, 3.298050E-01, 4.803880E-01, 5.543670E-01,
              3.118010E-01, 3.951680E-01, 5.644060E-01]),
    np.array([1.275370E-01, 4.237800E-02, 9.437400E-06, 5.516300E-09, 0., 0., 0.,
              0., 3.244560E-01, 1.631400E-03, 3.142700E-09, 0., 0., 0.,
              0., 0., 4.509400E-01, 2.679200E-03, 0., 0., 0.,
              0., 0., 0., 4.525650E-01, 5.566400E-03, 0., 0.,
              0., 0., 0., 1.252500E-04, 2.714010E-01, 1.025500E-02, 1.002100E-08,
              0., 0., 0., 0., 1.296800E-03, 2.658020E-01, 1.680900E-02,
              0., 0., 0., 0., 0., 8.545800E-03, 2.730800E-01]),
    np.array([2.005998E-02, 2.027303E-03, 1.570599E-02, 4.518301E-02,
              4.334208E-02, 2.020901E-01, 5.257105E-01]),
    np.array([5.87910E-01, 4.11760E-01, 3.39060E-04, 1.17610E-07, 0., 0., 0.]),
)
C5G7_WATER = MaterialXS(
    "Water",
    np.array([1.592060E-01, 4.129700E-01, 5.903100E-01, 5.843500E-01,
              7.180000E-01, 1.254450E+00, 2.650380E+00]),
    np.array([4.447770E-02, 1.134000E-01, 7.234700E-04, 3.749900E-06, 5.318400E-08, 0., 0.,
              0., 2.823340E-01, 1.299400E-01, 6.234000E-04, 4.800200E-05, 7.448600E-06, 1.045500E-06,
              0., 0., 3.452560E-01, 2.245700E-01, 1.699900E-02, 2.644300E-03, 5.034400E-04,
              0., 0., 0., 9.102840E-02, 4.155100E-01, 6.373200E-02, 1.213900E-02,
              0., 0., 0., 7.143700E-05, 1.391380E-01, 5.118200E-01, 6.122900E-02,
              0., 0., 0., 0., 2.215700E-03, 6.999130E-01, 5.373200E-01,
              0., 0., 0., 0., 0., 1.324400E-01, 2.480700E+00]),
    np.zeros(7), np.zeros(7),
)

# 17x17 guide-tube / instrument-tube positions (row, col), row 0 at the top.
GUIDE_POSITIONS_17 = [
    (2, 5), (2, 8), (2, 11), (3, 3), (3, 13),
    (5, 2), (5, 5), (5, 8), (5, 11), (5, 14),
    (8, 2), (8, 5), (8, 8), (8, 11), (8, 14),
    (11, 2), (11, 5), (11, 8), (11, 11), (11, 14),
    (13, 3), (13, 13), (14, 5), (14, 8), (14, 11),
]

PITCH = 1.26
FUEL_RADIUS = 0.54


def _pin(fuel_mat: int, mod_mat: int) -> PinTemplate:
    return PinTemplate([Cylinder(FUEL_RADIUS, fuel_mat, 3, 8)], mod_mat, 8)


# ----------------------------------------------------------------------
# Workloads
# ----------------------------------------------------------------------

def hom_inf_medium() -> Problem:
    """2-group homogeneous infinite medium (reference tests/input_set.py:32-92)."""
    m = MaterialXS("infinite_medium",
                   np.array([0.2208, 1.604]),
                   np.array([0.1, 0.117, 0.0, 1.42]),
                   np.array([0.0015, 0.325]), np.array([1.0, 0.0]))
    return Problem("hom_inf_medium", [m], [PinTemplate([], 0, 0)], 10, 10, 0.25, 0.25,
                   np.zeros((10, 10), np.int32), num_azim=4, azim_spacing=0.1,
                   tolerance=1e-5, max_iters=500)


def c5g7_pin_cell() -> Problem:
    """C5G7 UO2 pin cell, 4x4 cm, r=1 (reference tests/input_set.py:95-135)."""
    return Problem("c5g7_pin_cell", [C5G7_UO2, C5G7_WATER],
                   [PinTemplate([Cylinder(1.0, 0, 1, 0)], 1, 0)], 1, 1, 4.0, 4.0,
                   np.zeros((1, 1), np.int32), num_azim=4, azim_spacing=0.1,
                   tolerance=1e-5, max_iters=500)


def water_box_source() -> Problem:
    """Fixed-source box of the reference's tests/test_compute_flux and
    tests/test_compute_source (test_compute_flux.py:13-91): the 2.5 cm
    HomInfMedInput grid of 10x10 cells (tests/input_set.py:54-92) filled with
    C5G7 water, vacuum on all four sides, and the cell at lattice row 3,
    column 3 (from the top left, the universes[0][3][3] of the test) a
    separate universe that carries the fixed source.  4 azimuthal angles at
    0.1 cm (the test harness defaults, tests/testing_harness.py:33-37)."""
    cells = np.zeros((10, 10), np.int32)
    cells[3, 3] = 1
    return Problem("water_box_source", [C5G7_WATER], [PinTemplate([], 0, 0), PinTemplate([], 0, 0)],
                   10, 10, 0.25, 0.25, cells, bc=("V", "V", "V", "V"), num_azim=4,
                   azim_spacing=0.1, tolerance=1e-5, max_iters=500, res_type="TOTAL_SOURCE")


def pin_cell_7g(seed: int = 1, num_azim: int = 32) -> Problem:
    """Single 1.26 cm pin cell, r=0.54 (3 rings x 8 sectors), moderator 8
    sectors, reflective; 7 synthetic groups; 32 azimuthal (default), 0.05 cm,
    TY 6."""
    rng = np.random.default_rng(seed)
    fuel = synthetic_material(rng, 7, "fuel", True)
    mod = synthetic_material(rng, 7, "moderator", False)
    return Problem("pin_cell_7g", [fuel, mod], [_pin(0, 1)], 1, 1, PITCH, PITCH,
                   np.zeros((1, 1), np.int32), num_azim=num_azim, azim_spacing=0.05,
                   tolerance=1e-5, max_iters=1000)


def lattice_17x17_7g(seed: int = 2, num_azim: int = 64, spacing: float = 0.03) -> Problem:
    """17x17 pin lattice (264 fuel pins, 25 moderator-filled guide positions
    with the same ring/sector scheme), reflective, 7 synthetic groups."""
    rng = np.random.default_rng(seed)
    fuel = synthetic_material(rng, 7, "fuel", True)
    mod = synthetic_material(rng, 7, "moderator", False)
    cells = np.zeros((17, 17), np.int32)
    for r, c in GUIDE_POSITIONS_17:
        cells[r, c] = 1
    return Problem("lattice_17x17_7g", [fuel, mod], [_pin(0, 1), _pin(1, 1)], 17, 17,
                   PITCH, PITCH, cells, num_azim=num_azim, azim_spacing=spacing,
                   tolerance=1e-5, max_iters=1000)


def minicore_70g(seed: int = 3, num_azim: int = 32, spacing: float = 0.05) -> Problem:
    """3x3 assembly mini-core: 2x2 fuel assemblies (17x17 pins each) in the
    upper-left, 21.42 cm water-like reflector in the right column and bottom
    row (each reflector assembly a 17x17 grid of homogeneous 1.26 cm blocks),
    reflective on xmin/ymax, vacuum on xmax/ymin; 70 synthetic groups."""
    rng = np.random.default_rng(seed)
    fuel = synthetic_material(rng, 70, "fuel", True)
    mod = synthetic_material(rng, 70, "moderator", False)
    refl = synthetic_material(rng, 70, "reflector", False)
    pins = [_pin(0, 1), _pin(1, 1), PinTemplate([], 2, 0)]
    n = 17 * 3
    cells = np.full((n, n), 2, np.int32)
    for ar in range(2):
        for ac in range(2):
            blk = np.zeros((17, 17), np.int32)
            for r, c in GUIDE_POSITIONS_17:
                blk[r, c] = 1
            cells[ar * 17:(ar + 1) * 17, ac * 17:(ac + 1) * 17] = blk
    return Problem("minicore_70g", [fuel, mod, refl], pins, n, n, PITCH, PITCH, cells,
                   bc=("R", "V", "V", "R"), num_azim=num_azim, azim_spacing=spacing,
                   tolerance=1e-5, max_iters=1000)


def minicore_7g(seed: int = 7, num_azim: int = 32, spacing: float = 0.4) -> Problem:
    """The minicore_70g geometry (2x2 fuel assemblies, reflector column and
    row, vacuum on xmax/ymin) with 7 synthetic groups; 32 azimuthal angles at
    0.4 cm by default: the 3x3 mini-core kind of the serving stream."""
    p = minicore_70g(num_azim=num_azim, spacing=spacing)
    rng = np.random.default_rng(seed)
    p.name = "minicore_7g"
    p.materials = [synthetic_material(rng, 7, "fuel", True), synthetic_material(rng, 7, "moderator", False),
                   synthetic_material(rng, 7, "reflector", False)]
    return p


def pin_cell_70g(seed: int = 6, num_azim: int = 16) -> Problem:
    """pin_cell_7g geometry with 70 synthetic groups (dense lower-triangular
    plus band scattering): the 70-group parity case."""
    rng = np.random.default_rng(seed)
    fuel = synthetic_material(rng, 70, "fuel", True)
    mod = synthetic_material(rng, 70, "moderator", False)
    return Problem("pin_cell_70g", [fuel, mod], [_pin(0, 1)], 1, 1, PITCH, PITCH,
                   np.zeros((1, 1), np.int32), num_azim=num_azim, azim_spacing=0.05,
                   tolerance=1e-5, max_iters=1000)


def lattice_3x3_vac_7g(seed: int = 4, num_azim: int = 16) -> Problem:
    """3x3 pin lattice (centre position moderator-filled), vacuum on all four
    sides, 7 synthetic groups; 16 azimuthal angles, 0.05 cm.  Leaks
    neutrons, so it exercises the neutron-balance keff and the Krylov
    operator sweeps (which the reference requires to have vacuum
    boundaries, openmoc/krylov.py:98-108)."""
    rng = np.random.default_rng(seed)
    fuel = synthetic_material(rng, 7, "fuel", True)
    mod = synthetic_material(rng, 7, "moderator", False)
    cells = np.zeros((3, 3), np.int32)
    cells[1, 1] = 1
    return Problem("lattice_3x3_vac_7g", [fuel, mod], [_pin(0, 1), _pin(1, 1)], 3, 3,
                   PITCH, PITCH, cells, bc=("V", "V", "V", "V"), num_azim=num_azim,
                   azim_spacing=0.05, tolerance=1e-5, max_iters=1000)


def pin_cell_7g_tc(seed: int = 5) -> Problem:
    """pin_cell_7g with transport-corrected-style moderator data: the
    in-group scattering of the two thermal groups is negative (what the
    reference's DIAGONAL transport stabilisation is for)."""
    rng = np.random.default_rng(seed)
    fuel = synthetic_material(rng, 7, "fuel", True)
    mod = synthetic_material(rng, 7, "moderator", False)
    ss = mod.sigma_s.reshape(7, 7).copy()
    for g in (5, 6):
        ss[g, g] = -0.1 * mod.sigma_t[g]
    mod = MaterialXS(mod.name, mod.sigma_t, ss.ravel(), mod.nu_sigma_f, mod.chi)
    return Problem("pin_cell_7g_tc", [fuel, mod], [_pin(0, 1)], 1, 1, PITCH, PITCH,
                   np.zeros((1, 1), np.int32), num_azim=32, azim_spacing=0.05,
                   tolerance=1e-5, max_iters=1000)


def pin_cell_3d_7g(seed: int = 1, top: str = "R") -> Problem:
    """pin_cell_7g extruded to 2 cm in 4 axial FSR layers: explicit 3D tracks,
    8 azimuthal angles at 0.1 cm, 6 Gauss-Legendre polar angles, 0.5 cm
    axial spacing, reflective bottom, `top` boundary (R or V)."""
    p = pin_cell_7g(seed)
    p.name = "pin_cell_3d_7g" if top == "R" else "pin_cell_3d_vac_7g"
    p.num_azim, p.azim_spacing, p.num_polar, p.quadrature = 8, 0.1, 6, "GL"
    p.height, p.layers, p.z_spacing, p.bc_z = 2.0, 4, 0.5, ("R", top)
    return p


def pin_cell_3d_vac_7g(seed: int = 1) -> Problem:
    """pin_cell_3d_7g with a vacuum top (axial leakage)."""
    return pin_cell_3d_7g(seed, top="V")


def lattice_3x3_3d_7g(seed: int = 4) -> Problem:
    """lattice_3x3_vac_7g (vacuum on the four sides) extruded to 3 cm in 3
    axial layers, vacuum bottom and reflective top: explicit 3D tracks, 8
    azimuthal angles at 0.15 cm, 6 Gauss-Legendre polar angles, 0.6 cm axial
    spacing."""
    p = lattice_3x3_vac_7g(seed)
    p.name = "lattice_3x3_3d_7g"
    p.num_azim, p.azim_spacing, p.num_polar, p.quadrature = 8, 0.15, 6, "GL"
    p.height, p.layers, p.z_spacing, p.bc_z = 3.0, 3, 0.6, ("V", "R")
    return p


def lattice_17x17_3d_7g(seed: int = 2, num_azim: int = 16, spacing: float = 0.1,
                        z_spacing: float = 0.5) -> Problem:
    """lattice_17x17_7g extruded to 20 cm in 10 axial FSR layers, reflective
    top and bottom: explicit 3D tracks, 16 azimuthal angles at 0.1 cm, 6
    Gauss-Legendre polar angles, 0.5 cm axial spacing."""
    p = lattice_17x17_7g(seed)
    p.name = "lattice_17x17_3d_7g"
    p.num_azim, p.azim_spacing, p.num_polar, p.quadrature = num_azim, spacing, 6, "GL"
    p.height, p.layers, p.z_spacing, p.bc_z = 20.0, 10, z_spacing, ("R", "R")
    return p


WORKLOADS = {
    "pin_cell_3d_7g": pin_cell_3d_7g,
    "pin_cell_3d_vac_7g": pin_cell_3d_vac_7g,
    "lattice_3x3_3d_7g": lattice_3x3_3d_7g,
    "lattice_17x17_3d_7g": lattice_17x17_3d_7g,
    "pin_cell_70g": pin_cell_70g,
    "lattice_3x3_vac_7g": lattice_3x3_vac_7g,
    "pin_cell_7g_tc": pin_cell_7g_tc,
    "hom_inf_medium": hom_inf_medium,
    "c5g7_pin_cell": c5g7_pin_cell,
    "water_box_source": water_box_source,
    "pin_cell_7g": pin_cell_7g,
    "lattice_17x17_7g": lattice_17x17_7g,
    "minicore_70g": minicore_70g,
    "minicore_7g": minicore_7g,
}


def get(name: str, **kw) -> Problem:
    return WORKLOADS[name](**kw)
