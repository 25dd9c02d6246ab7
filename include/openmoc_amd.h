/*
 * openmoc_amd.h -- C ABI of the MI355X (gfx950) method-of-characteristics
 * source-iteration library.
 *
 * The library replaces the reference OpenMOC Solver back-end (CPUSolver /
 * GPUSolver, src/Solver.h:121-585) on the hot path: transport sweep, FSR
 * sources, scalar-flux update, keff, normalisation and residual all run in
 * hand-written HIP kernels on one GPU.  The CPU producer (problem parsing,
 * quadrature, track laydown and ray tracing) is exposed too, as the
 * counterpart of the reference TrackGenerator.
 *
 * Conventions
 *   - Every function returns MOC_OK (0) or a negative MOC_ERR_* code; the
 *     message of the last failure on the calling thread is moc_last_error().
 *     Nothing throws across this boundary (reference: log_printf(ERROR)
 *     throws std::logic_error, src/log.cpp:588-599).
 *   - Groups are 1-based where the reference is 1-based (getFlux), 0-based in
 *     flat arrays ([fsr*G + g]).
 *   - Inputs are copied at call time; no caller buffer is retained
 *     (deviation: reference CPUSolver::setFluxes aliases the caller's buffer,
 *     src/CPUSolver.cpp:180-192).
 *   - A solver handle may be used from one host thread at a time.
 */
#ifndef OPENMOC_AMD_H_
#define OPENMOC_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOC_OK 0
#define MOC_ERR_INVALID (-1)   /* bad argument / id / size */
#define MOC_ERR_STATE (-2)     /* call out of order (e.g. no tracks set) */
#define MOC_ERR_HIP (-3)       /* HIP runtime failure */
#define MOC_ERR_MEMORY (-4)    /* allocation / admission refused */
#define MOC_ERR_NO_DEVICE (-5) /* no GPU visible */
#define MOC_ERR_IO (-6)        /* file missing, short or malformed */

/* residualType (src/Solver.h / constants) */
#define MOC_SCALAR_FLUX 0
#define MOC_FISSION_SOURCE 1
#define MOC_TOTAL_SOURCE 2

/* solverMode (src/constants.h) */
#define MOC_FORWARD 0
#define MOC_ADJOINT 1

/* exponential evaluation in the sweep */
#define MOC_EXP_RATIONAL 0   /* the reference's 5/6 rational fit (default) */
#define MOC_EXP_HW 1         /* hardware exp2: 1 - 2^(-x log2 e) */

/* Arithmetic of the scattering-source contraction sum_g sigma_s(g->g') phi(g)
 * (and the fission rate sum_g nu-sigma-f(g) phi(g)) in computeFSRSources. */
#define MOC_SCATTER_AUTO 0     /* matrix cores from 32 groups on, FP64 below */
#define MOC_SCATTER_FP64 1     /* FP64 loops (the reference's arithmetic) */
#define MOC_SCATTER_MFMA_F16 2 /* fp16 MFMA on hi/lo split operands, FP32 accumulation */
#define MOC_SCATTER_FP32 3     /* FP32 vector fallback */

/* Thread-local text of the last error ("" if none). */
const char* moc_last_error(void);

/* Library version string. */
const char* moc_version(void);

/* ========================================================================
 * CPU producer: counterpart of TrackGenerator::generateTracks
 * (src/TrackGenerator.cpp:796-880) for a Cartesian pin lattice described by
 * a "moc-problem 1" text.
 * ====================================================================== */
typedef struct moc_tracks moc_tracks_t;

typedef struct {
  int32_t num_groups;
  int32_t num_materials;
  int32_t num_azim;        /* full circle */
  int32_t num_polar;       /* full sphere */
  int64_t num_tracks;
  int64_t num_segments;
  int64_t num_fsrs;
  double max_optical_length;
  double tolerance;        /* from the problem's solver line */
  int32_t max_iters;
  int32_t res_type;
  int32_t is_3d;           /* 1: explicit 3D tracks (TrackGenerator3D, EXPLICIT_3D) */
  int32_t num_layers;      /* axial FSR layers (3D) */
} moc_tracks_info_t;

/* Parses the problem, lays down and traces the tracks (threads<=0: all
 * cores).  Replaces Geometry::initializeFlatSourceRegions +
 * TrackGenerator::generateTracks. */
int moc_tracks_generate(const char* problem_text, int threads, moc_tracks_t** out);
/* TrackGenerator::dumpSegmentsToFile (src/TrackGenerator.cpp:1388-1501;
 * records of DumpSegments::onTrack, src/TrackTraversingAlgorithms.cpp:
 * 1101-1144): the problem text, the laydown, every track's segment records
 * (f64 length, i32 material, i64 region, i32 track index, f64 start x, y, z)
 * and the FSR table.  MOC_ERR_IO if the file cannot be written. */
int moc_tracks_dump_segments(const moc_tracks_t* tracks, const char* path);
/* TrackGenerator::readSegmentsFromFile (src/TrackGenerator.cpp:1518-1638):
 * tracks and segments from a file written by moc_tracks_dump_segments for the
 * same problem text, without laydown or ray tracing; or from a 2D file the
 * reference's own dumpSegmentsToFile wrote: the tracks are then this
 * problem's laydown (as the reference lays them down before reading), the
 * segments the file's, and each FSR's material and key come from the point
 * the file gives for it.  *out is NULL (and the call returns MOC_OK) if the
 * file was written for another problem (for a reference-written file: its
 * records do not fit this problem's laydown), as the reference returns false;
 * MOC_ERR_IO for a missing or malformed file. */
int moc_tracks_read_segments(const char* problem_text, const char* path, moc_tracks_t** out);
void moc_tracks_destroy(moc_tracks_t* tracks);
int moc_tracks_info(const moc_tracks_t* tracks, moc_tracks_info_t* info);

/* Copies track arrays out (any pointer may be NULL to skip). Sizes follow
 * moc_tracks_info. */
int moc_tracks_export_tracks(const moc_tracks_t* tracks, int64_t* seg_offset /*T+1*/,
                             int32_t* azim /*T*/, int32_t* xy /*T*/,
                             double* start_xy /*2T*/, double* end_xy /*2T*/,
                             int64_t* next_fwd, int64_t* next_bwd,
                             uint8_t* next_fwd_is_fwd, uint8_t* next_bwd_is_fwd,
                             uint8_t* bc_fwd, uint8_t* bc_bwd);
int moc_tracks_export_segments(const moc_tracks_t* tracks, double* length /*S*/,
                               int32_t* fsr /*S*/);
/* Start point (x, y) of every segment [2S] (Geometry::segmentize2D records
 * it, src/Geometry.cpp:2196-2210; the linear source re-centres it on the
 * FSR centroid). */
int moc_tracks_export_segment_starts(const moc_tracks_t* tracks, double* start_xy /*2S*/);
int moc_tracks_export_fsrs(const moc_tracks_t* tracks, int32_t* material /*N*/,
                           double* volume /*N*/, double* centroid /*2N*/, int64_t* key /*N*/);
/* Quadrature after laydown: phi/spacing/azim_weight [num_azim/2],
 * sin_theta/polar_weight [num_polar], total_weight [num_azim/2 * num_polar]. */
int moc_tracks_export_quadrature(const moc_tracks_t* tracks, double* phi, double* spacing,
                                 double* azim_weight, double* sin_theta,
                                 double* polar_weight, double* total_weight);
/* 3D tracks (TrackGenerator3D::get3DTracks, src/TrackGenerator3D.cpp:
 * 1404-1460; uid order = stacks by azimuthal, xy and polar index, then z):
 * polar index [T] and z of the start and end points [T]; corrected polar
 * angles and polar spacings [num_azim/2 * num_polar]
 * (src/TrackGenerator3D.cpp:750-808); FSR centroid z [N].  Errors on 2D
 * track sets. */
int moc_tracks_export_tracks_3d(const moc_tracks_t* tracks, int32_t* polar, double* start_z,
                                double* end_z, double* theta, double* polar_spacing,
                                double* fsr_centroid_z);
/* Materials: sigma_t/nu_sigma_f/chi [M*G], sigma_s [M*G*G] dest-major. */
int moc_tracks_export_materials(const moc_tracks_t* tracks, double* sigma_t,
                                double* sigma_s, double* nu_sigma_f, double* chi);

/* ========================================================================
 * Solver: counterpart of the reference Solver API implemented on one GPU.
 * ====================================================================== */
typedef struct moc_solver moc_solver_t;

/* Track data handed to the solver (Solver::setTrackGenerator,
 * src/Solver.cpp:412-453; device upload as GPUSolver::initializeTracks,
 * src/accel/cuda/GPUSolver.cu:1312-1353). 2D tracks, uid order.  Segment
 * lengths are stored on the device as float; FSR ids as int32. */
typedef struct {
  int32_t num_azim;                 /* full circle */
  int32_t num_polar;                /* full sphere; the 2D sweep uses num_polar/2 */
  int64_t num_tracks;               /* T */
  int64_t num_segments;             /* S */
  const int64_t* seg_offset;        /* [T+1] */
  const double* seg_length;         /* [S] cm */
  const int32_t* seg_fsr;           /* [S] */
  const int32_t* azim;              /* [T] azimuthal index in [0, num_azim/2) */
  const int64_t* next_fwd;          /* [T] Track::getTrackNextFwd */
  const int64_t* next_bwd;          /* [T] Track::getTrackNextBwd */
  const uint8_t* next_fwd_is_fwd;   /* [T] Track::getNextFwdFwd */
  const uint8_t* next_bwd_is_fwd;   /* [T] Track::getNextBwdFwd */
  const uint8_t* bc_fwd;            /* [T] 0 vacuum, 1 reflective, 2 periodic */
  const uint8_t* bc_bwd;            /* [T] */
  const double* total_weight;       /* [num_azim/2 * num_polar] Quadrature w(a,p) */
  const double* sin_theta;          /* [num_polar] */
  /* Needed only for the linear source (may be NULL otherwise): */
  const double* azim_phi;           /* [num_azim/2] track angle of each azimuthal index */
  const double* azim_weight;        /* [num_azim/2] azimuthal spacing * azimuthal weight */
  const double* polar_weight;       /* [num_polar] Quadrature::getPolarWeight */
  const double* seg_start;          /* [2S] start point (x, y) of each segment */
  /* 3D tracks (Track3D, explicit segments): polar index of each track in
   * [0, num_polar), or NULL for 2D tracks.  With it the sweep is the 3D one
   * (CPUSolver::tallyScalarFlux 3D branch, src/CPUSolver.cpp:2411-2453): G
   * fluxes per track end, segment lengths are 3D lengths, the exponential
   * takes tau itself, and total_weight[azim * num_polar + polar] weights the
   * whole track (src/TrackTraversingAlgorithms.cpp:902-910). */
  const int32_t* polar;             /* [T] */
  /* 3D only, optional: Track::getXYIndex of each track (the 2D track it lies
   * over).  With it the tracks over one 2D track (its z-stacks) are swept by
   * one block with their FSRs' tally in LDS. */
  const int32_t* xy;                /* [T] */
} moc_track_data_t;

/* device < 0: current device. */
int moc_solver_create(int device, moc_solver_t** out);
void moc_solver_destroy(moc_solver_t* solver);

/* Solver::setTrackGenerator */
int moc_solver_set_tracks(moc_solver_t* solver, const moc_track_data_t* tracks);
/* Geometry materials (Material::setSigma*, src/Material.cpp:628-979):
 * sigma_t/nu_sigma_f/chi [M*G], sigma_s [M*G*G] dest-major s[dest*G+orig]. */
int moc_solver_set_materials(moc_solver_t* solver, int32_t num_materials, int32_t num_groups,
                             const double* sigma_t, const double* sigma_s,
                             const double* nu_sigma_f, const double* chi);
/* Solver::initializeFSRs (volumes and materials, src/Solver.cpp:820-873). */
int moc_solver_set_fsrs(moc_solver_t* solver, int64_t num_fsrs, const int32_t* material,
                        const double* volume);

int moc_solver_set_convergence_threshold(moc_solver_t* solver, double threshold);
int moc_solver_set_exponential(moc_solver_t* solver, int mode /* MOC_EXP_* */);
int moc_solver_allow_negative_fluxes(moc_solver_t* solver, int allow);
/* Scattering-source arithmetic (MOC_SCATTER_*).  Replaces the FP64 loops of
 * CPUSolver::computeFSRSources (src/CPUSolver.cpp:1939-2023) and the moment
 * sources of CPULSSolver::computeFSRSources (src/CPULSSolver.cpp:386-524);
 * the reference has no such switch (its GPUSolver uses FP32 loops,
 * src/accel/cuda/GPUSolver.cu:288-338).  *in_use (nullable) receives the
 * path MOC_SCATTER_AUTO resolves to for the current group count. */
int moc_solver_set_scatter_precision(moc_solver_t* solver, int mode, int* in_use);
/* CPUSolver::computeFSRSources(iteration) (src/CPUSolver.cpp:1939-2023; a
 * protected method of Solver, src/Solver.h:396) from the current flux and
 * keff: sources [N*G] per steradian, the reference's _reduced_sources, with
 * negative sources set to 1e-25 below iteration 30 unless negative fluxes
 * are allowed. */
int moc_solver_compute_fsr_sources(moc_solver_t* solver, int32_t iteration, double* sources,
                                   int64_t n);
/* Solver::setFixedSourceByFSR (1-based group). */
int moc_solver_set_fixed_source_by_fsr(moc_solver_t* solver, int64_t fsr, int32_t group,
                                       double source);
/* Solver::setFixedSourceByMaterial(material, group, source) (src/Solver.cpp:
 * 538-543), material = index into the materials of moc_solver_set_materials;
 * applied to every FSR of the material when the next solve starts, after the
 * per-FSR and cell sources (src/Solver.cpp:1119-1137). */
int moc_solver_set_fixed_source_by_material(moc_solver_t* solver, int32_t material, int32_t group,
                                            double source);
/* Geometry::findCellContainingFSR (src/Geometry.cpp) as a table: the id of
 * the material cell each FSR lies in [num_fsrs], the cell ids that
 * moc_solver_set_fixed_source_by_cell names (any numbering the caller keeps;
 * the Python mirror numbers a pin's cells pin * 64 + index, cylinders
 * innermost first, then the outer cell). */
int moc_solver_set_fsr_cells(moc_solver_t* solver, const int32_t* cell_of_fsr, int64_t num_fsrs);
/* Solver::setFixedSourceByCell(cell, group, source) (src/Solver.cpp:510-527):
 * every FSR of the cell.  Fixed sources are applied when a solve starts, as
 * Solver::initializeFixedSources does (src/Solver.cpp:1089-1137): cell
 * sources, then material sources, overwrite the per-FSR ones of their FSRs,
 * whatever the call order.  MOC_ERR_STATE without the FSR cell table. */
int moc_solver_set_fixed_source_by_cell(moc_solver_t* solver, int32_t cell, int32_t group,
                                        double source);
/* CPUSolver::resetFixedSources (src/CPUSolver.cpp:228-250): all fixed
 * sources zero. */
int moc_solver_reset_fixed_sources(moc_solver_t* solver);

/* stabilizationType (src/Solver.h:103-113) */
#define MOC_STAB_NONE (-1)
#define MOC_STAB_DIAGONAL 0
#define MOC_STAB_YAMAMOTO 1
#define MOC_STAB_GLOBAL 2
/* Solver::stabilizeTransport(factor, type) (src/Solver.cpp:676-681); applied
 * by computeEigenvalue from its second iteration on (src/Solver.cpp:1616-1635,
 * CPUSolver::computeStabilizingFlux / stabilizeFlux, src/CPUSolver.cpp:
 * 2665-2805).  type MOC_STAB_NONE switches it off. */
int moc_solver_stabilize_transport(moc_solver_t* solver, double factor, int32_t type);
/* Solver::setKeffFromNeutronBalance (src/Solver.cpp:2047-2049): keff =
 * fission / (absorption + vacuum leakage) instead of the fission-rate ratio
 * (CPUSolver::computeKeff, src/CPUSolver.cpp:2258-2328).  on = 0 restores the
 * default. */
int moc_solver_set_keff_from_neutron_balance(moc_solver_t* solver, int on);
/* Krylov operator sweeps (Solver::fissionTransportSweep / scatterTransportSweep,
 * src/Solver.cpp:1282-1297): sources from the current flux (the last solve's,
 * or moc_solver_set_fluxes'), fission-only (CPUSolver::computeFSRFissionSources,
 * src/CPUSolver.cpp:2030-2066) or scatter-only (computeFSRScatterSources,
 * :2072-2104); one sweep from the start fluxes the previous operator sweep
 * left (zero at first); addSourceToScalarFlux into the flux. */
int moc_solver_fission_transport_sweep(moc_solver_t* solver);

/* Linear source (CPULSSolver, src/CPULSSolver.cpp; 2D): on = 1 makes
 * compute_eigenvalue solve with linear sources in every FSR.  Needs the
 * track data's azim_phi, azim_weight, polar_weight and seg_start.  The FSR
 * centroids (TrackGenerator::generateFSRCentroids, src/TrackGenerator.cpp:
 * 1681-1735) and the linear expansion coefficients
 * (LinearExpansionGenerator, src/TrackTraversingAlgorithms.cpp:470-831) are
 * computed at the first solve after tracks, FSRs or materials change.
 * Stabilisation, the neutron-balance keff and the Krylov sweeps are
 * flat-source only. */
int moc_solver_set_linear_source(moc_solver_t* solver, int on);
/* Normalised scalar-flux moments [N*G*2], (x, y) per FSR and group, layout
 * [(fsr*G + g)*2 + m] (CPULSSolver::_scalar_flux_xyz). */
int moc_solver_get_flux_moments(moc_solver_t* solver, double* out, int64_t num_values);
/* FSR centroids [2N] (x, y) the linear source expands about. */
int moc_solver_get_fsr_centroids(moc_solver_t* solver, double* out, int64_t num_values);
int moc_solver_scatter_transport_sweep(moc_solver_t* solver);

/* Solver::computeEigenvalue(max_iters, res_type) (src/Solver.cpp:1542-1689). */
int moc_solver_compute_eigenvalue(moc_solver_t* solver, int32_t max_iters, int32_t res_type);
/* Solver::computeFlux(max_iters, only_fixed_source) (src/Solver.cpp:1352-1420):
 * k_eff = 1; only_fixed_source = 0 after an earlier solve starts from that
 * solve's scalar and angular fluxes, otherwise from zero (:1379-1386). */
int moc_solver_compute_flux(moc_solver_t* solver, int32_t max_iters, int32_t only_fixed_source);
/* Solver::computeSource(max_iters, k_eff, res_type) (src/Solver.cpp:1459-1516). */
int moc_solver_compute_source(moc_solver_t* solver, int32_t max_iters, double k_eff,
                              int32_t res_type);

int moc_solver_get_keff(const moc_solver_t* solver, double* keff);
int moc_solver_get_num_iterations(const moc_solver_t* solver, int32_t* iters);
/* Solver::getFlux(fsr_id, group) with 1-based group (src/Solver.cpp:298-321). */
int moc_solver_get_flux(moc_solver_t* solver, int64_t fsr, int32_t group, double* flux);
/* CPUSolver::getFluxes(out, N*G) (src/CPUSolver.cpp:66-130). */
int moc_solver_get_fluxes(moc_solver_t* solver, double* out, int64_t num_fluxes);
/* CPUSolver::setFluxes (src/CPUSolver.cpp:166-192): copied in; the operand
 * of the next Krylov operator sweep and what get_fluxes returns until then.
 * A compute_* call starts from its own flat guess, as the reference does
 * without a restart (Solver::computeInitialFluxGuess, src/Solver.cpp:1710-1731). */
int moc_solver_set_fluxes(moc_solver_t* solver, const double* in, int64_t num_fluxes);
/* CPUSolver::computeFSRFissionRates(rates, N, nu = true) (src/CPUSolver.cpp:
 * 2825-2886): volume-weighted nu-sigma-f rates.  The problem description
 * carries nu-sigma-f only, so the reference's default nu = false (sigma-f
 * rates) has no counterpart. */
int moc_solver_compute_fsr_fission_rates(moc_solver_t* solver, double* rates, int64_t num_fsrs);

/* ---- FSR flux files (checkpoint / restart) ----
 * Layout of Solver::dumpFSRFluxes (src/Solver.cpp:2066-2112), native byte
 * order: f64 k_eff | i32 num_groups | i64 num_fsrs | per FSR f64 x, y, z
 * (centroid) and f64 flux[num_groups].  FSRs of a file are matched to the
 * solver's by nearest centroid among the 3x3x3 cells of width `tolerance` cm
 * around each centroid (src/Solver.cpp:2198-2306). */

/* FSR centroids [3N] naming the solver's FSRs in flux files: the reference's
 * Geometry::getFSRCentroid values, from TrackGenerator::generateFSRCentroids
 * (src/TrackGenerator.cpp:1681-1735; z = 0 for 2D tracks). */
int moc_solver_set_fsr_centroids(moc_solver_t* solver, const double* xyz, int64_t num_values);
/* Solver::dumpFSRFluxes(fname) (src/Solver.cpp:2066-2112): k_eff and the
 * current scalar fluxes. */
int moc_solver_dump_fsr_fluxes(moc_solver_t* solver, const char* path);
/* Solver::loadFSRFluxes(fname, assign_k_eff, tolerance) (src/Solver.cpp:
 * 2135-2318): the matched fluxes become the current fluxes (as set_fluxes);
 * assign_k_eff sets k_eff to the file's.  *max_centroid_error (nullable)
 * receives the largest centroid mismatch (cm).  MOC_ERR_IO for an unreadable
 * file, MOC_ERR_INVALID for another FSR or group count or an unmatched FSR. */
int moc_solver_load_fsr_fluxes(moc_solver_t* solver, const char* path, int32_t assign_k_eff,
                               double tolerance, double* max_centroid_error);
/* Solver::loadInitialFSRFluxes(fname) (src/Solver.cpp:2120-2123; used in
 * computeInitialFluxGuess, src/Solver.cpp:1733-1751): eigenvalue and source
 * solves start from the file's fluxes and k_eff, normalised to the FSR count,
 * then run 30 sweeps on their sources before the source iteration.  The file
 * is read when each solve starts (tolerance 0.01 cm).  NULL or "" switches it
 * off. */
int moc_solver_load_initial_fsr_fluxes(moc_solver_t* solver, const char* path);
/* Solver::setResidualByReference(fname) (src/Solver.cpp:2056-2059; used in
 * CPUSolver::computeResidual, src/CPUSolver.cpp:2113-2252): residuals compare
 * the normalised flux with the file's fluxes instead of the previous
 * iteration's.  Read when each solve starts; NULL or "" switches it off. */
int moc_solver_set_residual_by_reference(moc_solver_t* solver, const char* path);
/* Solver::setRestartStatus(is_restart) (src/Solver.cpp:610-615): while on,
 * compute_eigenvalue keeps k_eff, and compute_eigenvalue / compute_source keep
 * the angular fluxes and the scalar flux of the solver's previous solve
 * instead of a flat guess (src/Solver.cpp:1554-1572, 1721-1731); the iteration
 * count starts at 1 (flux moments carry over with the linear source).
 * MOC_ERR_STATE without an earlier solve on this solver. */
int moc_solver_set_restart_status(moc_solver_t* solver, int32_t is_restart);
/* Solver::setSolverMode(solver_mode) (src/Solver.cpp:597-602; applied in
 * initializeMaterials, :792-812): MOC_ADJOINT solves with the scattering and
 * fission matrices transposed (Material::transposeProductionMatrices,
 * src/Material.cpp:1091-1103); keff, normalisation and the fission-source
 * residual keep nu-sigma-f, as in the reference.  Stays in effect for every
 * following solve. */
int moc_solver_set_solver_mode(moc_solver_t* solver, int32_t mode);
/* Solver::setChiSpectrumMaterial(material) (src/Solver.cpp:708-710; used by
 * computeInitialFluxGuess, :1722-1725, and flattenFSRFluxesChiSpectrum,
 * src/CPUSolver.cpp:1829-1840): eigenvalue and source solves start from the
 * material's chi in every FSR.  material = index into the materials; -1 for
 * the flat guess. */
int moc_solver_set_chi_spectrum_material(moc_solver_t* solver, int32_t material);

/* Host-only flux-file helpers (no GPU needed), the steps the solver calls
 * above use.  moc_flux_file_read with null xyz / flux returns the header only;
 * otherwise xyz_len must be 3*num_fsrs and flux_len num_fsrs*num_groups. */
int moc_flux_file_write(const char* path, double keff, int32_t num_groups, int64_t num_fsrs,
                        const double* xyz, const double* flux);
int moc_flux_file_read(const char* path, double* keff, int32_t* num_groups, int64_t* num_fsrs,
                       double* xyz, int64_t xyz_len, double* flux, int64_t flux_len);
/* index[i] = the file centroid (of num_file, [3*num_file]) matched to own
 * centroid i (of num_own) by the rule above; *max_error (nullable) the largest
 * matched distance. */
int moc_fsr_match_centroids(const double* file_xyz, int64_t num_file, const double* own_xyz,
                            int64_t num_own, double tolerance, int64_t* index, double* max_error);

/* Timing of the last solve: device time of all sweeps (s), total wall time
 * (s), number of integrations (2 * fluxes_per_track * segments * iters,
 * src/Solver.cpp:1901-1905). */
int moc_solver_get_timing(const moc_solver_t* solver, double* sweep_seconds,
                          double* total_seconds, int64_t* num_integrations);

/* One transport sweep on caller-given reduced sources Q [N*G] (per
 * steradian, CPUSolver::_reduced_sources) and start angular fluxes
 * psi [T*2*(P*G)] in the reference layout ((t*2+d)*fpt + p*G + g), P =
 * num_polar/2.  Writes the raw FSR tally [N*G] (CPUSolver::_scalar_flux right
 * after transportSweep, before addSourceToScalarFlux) and replaces psi with
 * the transferred outgoing fluxes (CPUSolver::_start_flux after the sweep).
 * Host pointers; for parity checks of the sweep kernel in isolation. */
int moc_solver_transport_sweep(moc_solver_t* solver, const double* Q, float* psi,
                               double* tally);

/* Bytes of device scratch the allocation-free sweep holds per GPU. */
#define MOC_NOALLOC_SCRATCH_BYTES (64 * 1024 * 1024)

/* Allocation-free form of moc_solver_transport_sweep.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Q [N*G], psi_in and psi_out [T*2*(P*G)] (reference layout, as above) and
 * tally [N*G] are DEVICE pointers; psi_out may equal psi_in.  The work is
 * queued on `stream` (a hipStream_t; NULL: the solver's own stream) with no
 * host synchronisation, and the results equal moc_solver_transport_sweep's
 * bit for bit.  The call makes no hipMalloc/hipFree and no memset: the
 * sweep's own buffers (start-flux slots, Q/sigma_t table, tally, iteration
 * state) live in one static device array of MOC_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race
 * and must be serialised by the caller.  MOC_ERR_INVALID when the problem
 * needs more scratch than that (2*psi slots + tables + tally; e.g. the
 * 17x17 lattice at 7 groups needs about 11 MB), or for the linear source. */
int moc_solver_transport_sweep_noalloc(moc_solver_t* solver, const double* Q,
                                       const float* psi_in, float* psi_out, double* tally,
                                       void* stream);
/* Scratch bytes (<= MOC_NOALLOC_SCRATCH_BYTES when it fits) that
 * moc_solver_transport_sweep_noalloc needs for this solver's problem. */
int moc_solver_noalloc_scratch_bytes(const moc_solver_t* solver, int64_t* bytes);

/* Sweep launch shape, shape11 = {groups_per_lane, lanes_per_track,
 * tracks_per_wave, block, tally_copies, lds_groups, lds_block,
 * lds_blocks_per_sub, lds_q, lds_flush, seg_store} (chosen by default from
 * the device's LDS size, or by the autotuner, openmoc_amd/autotune.py).
 *   lds_groups > 0: LDS-privatised tally; each block holds the [N][lds_groups]
 *     int64 fixed-point tally slice of one group subset, 1/V per FSR and
 *     sigma_t by material in LDS, plus the subset's FP32 Q/sigma_t slice when
 *     lds_q = 1 (8*N*lds_groups + 4*N + 4*M*lds_groups
 *     + 4*N*lds_groups*lds_q bytes must fit the device's LDS); lds_block
 *     threads (multiple of 64, <= 1024), lds_blocks_per_sub blocks per
 *     subset.  lds_flush = 0: each block writes its slice to a partial buffer
 *     that the per-FSR step sums; lds_flush = 1: each block adds the
 *     slice's non-zero entries to one int64 tally with atomics.  Both sums
 *     are exact, so the results are the same bits.
 *   lds_groups = 0: the global-tally sweep; groups_per_lane in
 *     {1,2,4,9,16}, lanes_per_track = ceil(G/groups_per_lane),
 *     tracks_per_wave * lanes_per_track <= 64, block a multiple of 64
 *     (<= 256).  seg_store = 1 (default): each segment's weighted
 *     delta-psi per group is stored (no atomics) and the per-FSR step sums
 *     an FSR's segments in a fixed order in FP64, so repeated solves give the
 *     same bits (device memory: 4 * segments * 2 * G bytes, three times that
 *     with the linear source); seg_store = 0: device-scope float
 *     atomics into tally_copies private tallies (block b uses copy
 *     b % copies), whose order varies from run to run. */
int moc_solver_set_launch_shape(moc_solver_t* solver, const int32_t* shape11);
int moc_solver_get_launch_shape(const moc_solver_t* solver, int32_t* shape11);

/* Benchmark form of the eigenvalue iteration (no counterpart in the
 * reference): runs num_iterations source iterations (sweep + all per-FSR
 * steps, in the graph batches a solve uses) from the current state with the
 * stopping rule disabled, synchronises, and returns the mean device time per
 * iteration; sweep_seconds_avg (nullable) receives the mean time of the
 * transport-sweep kernel alone, from HIP events around it on the solver's
 * stream in up to 8 further iterations. */
int moc_solver_run_iterations(moc_solver_t* solver, int32_t num_iterations,
                              double* step_seconds_avg, double* sweep_seconds_avg);

/* Device-side sweep timing: runs `repeats` sweeps back to back on the
 * solver's current state and returns the mean kernel time (s) measured with
 * HIP events on the sweep's stream. */
int moc_solver_time_sweeps(moc_solver_t* solver, int32_t repeats, double* seconds_per_sweep);

/* ========================================================================
 * Serving: continuous batching of eigenvalue solves on one GPU.  No
 * counterpart in the reference (its GPUSolver runs one synchronous solve,
 * src/accel/cuda/GPUSolver.cu:1664); each solve follows
 * Solver::computeEigenvalue (src/Solver.cpp:1542-1689).
 * A geometry (tracks, FSRs; materials fix G, M and the launch shape) is
 * uploaded once; requests bring their own cross sections.  Requests are
 * admitted as solve slots and device memory (mem_budget_bytes <= 0: 80% of
 * the free memory) allow; their iterations overlap on the device.  A
 * result does not depend on what else ran (bitwise the same as the solve
 * alone).
 * ====================================================================== */
typedef struct moc_engine moc_engine_t;

typedef struct {
  double keff;
  int32_t num_iterations;
  double latency_s;   /* submit -> result */
  double queue_s;     /* submit -> admitted */
  double solve_s;     /* admitted -> result */
} moc_request_result_t;

int moc_engine_create(int device, int32_t max_active, int64_t mem_budget_bytes,
                      moc_engine_t** out);
/* Finishes the solves in flight, then frees everything. */
void moc_engine_destroy(moc_engine_t* engine);
int moc_engine_add_geometry(moc_engine_t* engine, const moc_track_data_t* tracks,
                            int64_t num_fsrs, const int32_t* fsr_material,
                            const double* fsr_volume, int32_t num_materials, int32_t num_groups,
                            const double* sigma_t, const double* sigma_s,
                            const double* nu_sigma_f, const double* chi, int32_t* geo_id);
/* Queues computeEigenvalue(max_iters, res_type) with convergence threshold
 * `tolerance` on geometry geo_id with these cross sections (layout as in
 * moc_solver_set_materials; counts must match the geometry's). */
int moc_engine_submit(moc_engine_t* engine, int32_t geo_id, int32_t num_materials,
                      int32_t num_groups, const double* sigma_t, const double* sigma_s,
                      const double* nu_sigma_f, const double* chi, int32_t max_iters,
                      int32_t res_type, double tolerance, int want_flux, int64_t* request_id);
/* Waits up to timeout_s (< 0: no limit) for a request; MOC_ERR_STATE if it
 * is still running.  flux (nullable, N*G) receives the normalised flux when
 * want_flux was set.  A returned request is forgotten. */
int moc_engine_wait(moc_engine_t* engine, int64_t request_id, double timeout_s,
                    moc_request_result_t* result, double* flux, int64_t num_fluxes);
int moc_engine_stats(moc_engine_t* engine, int64_t* completed, int64_t* failed,
                     int64_t* iterations, int32_t* pending);
/* Device memory of the engine's solves: the budget, the bytes the solvers of
 * all batches hold now (live or idle; an idle solver is freed when another
 * geometry needs the memory), and the most they have held at once.  A new
 * solver is created only while held + its bytes <= budget, except when no
 * solve is live at all (the queue always moves). */
int moc_engine_memory(moc_engine_t* engine, int64_t* budget, int64_t* held, int64_t* held_max);

/* Diagnostic (no counterpart in the reference): a plain device-to-device copy
 * of `bytes` bytes (a multiple of 16) run `repeats` times on `device`; the mean
 * kernel time goes to *seconds.  The copy kernel reads and writes exactly
 * `bytes` each time, which calibrates profiler HBM byte counters. */
int moc_device_copy_calibration(int device, int64_t bytes, int32_t repeats, double* seconds);

#ifdef __cplusplus
}
#endif

#endif /* OPENMOC_AMD_H_ */
