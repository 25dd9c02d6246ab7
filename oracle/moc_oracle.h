/*
 * moc_oracle: CPU restatement (FP64, OpenMP) of the reference OpenMOC
 * source iteration, used ONLY as a checker by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg.  It is never linked into the product.
 *
 * Every entry point takes plain arrays (the same SoA track layout the product
 * consumes), so the oracle can run either on this package's tracks or on the
 * tracks dumped from the compiled reference (oracle/ref_driver.cpp).
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * comment of moc_oracle.cpp.
 */
#ifndef MOC_ORACLE_H_
#define MOC_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int num_groups;            /* G */
  int num_polar_half;        /* P = num_polar/2 (2D) */
  int num_azim_half;         /* num_azim/2 */
  int64_t num_tracks;        /* T */
  int64_t num_fsrs;          /* N */
  int num_materials;         /* M */
  /* tracks, uid order */
  const int64_t* seg_offset;       /* [T+1] */
  const double* seg_length;        /* [S] */
  const int32_t* seg_fsr;          /* [S] */
  const int32_t* trk_azim;         /* [T] */
  const int64_t* next_fwd;         /* [T] */
  const int64_t* next_bwd;         /* [T] */
  const uint8_t* next_fwd_is_fwd;  /* [T] */
  const uint8_t* next_bwd_is_fwd;  /* [T] */
  const uint8_t* bc_fwd;           /* [T] 0 vacuum, 1 reflective, 2 periodic */
  const uint8_t* bc_bwd;           /* [T] */
  /* quadrature */
  const double* total_weight;      /* [num_azim_half][P] w(a,p) */
  const double* inv_sin_theta;     /* [P] */
  /* materials, per material m: */
  const double* sigma_t;           /* [M][G] */
  const double* sigma_s;           /* [M][G*G] dest-major s[dest*G+orig] */
  const double* nu_sigma_f;        /* [M][G] */
  const double* chi;               /* [M][G] */
  /* FSRs */
  const int32_t* fsr_material;     /* [N] */
  const double* fsr_volume;        /* [N] */
  const double* fixed_source;      /* [N*G] or NULL */
  /* linear source only (NULL otherwise) */
  const double* azim_phi;          /* [num_azim_half] track angle */
  const double* azim_weight;       /* [num_azim_half] azimuthal spacing * weight */
  const double* polar_weight;      /* [P] polar weight of the quadrature */
  const double* seg_start;         /* [2S] segment start points (x, y) */
  /* 3D tracks (NULL for 2D): polar index per track; total_weight is then
   * [num_azim_half][num_polar_3d] and one polar angle travels per track */
  const int32_t* trk_polar;        /* [T] */
  int num_polar_3d;
} OracleProblem;

enum { ORACLE_SCALAR_FLUX = 0, ORACLE_FISSION_SOURCE = 1, ORACLE_TOTAL_SOURCE = 2 };

typedef struct {
  int max_iters;
  double tolerance;
  int res_type;               /* ORACLE_* */
  int psi_fp64;               /* 0: float angular fluxes (reference), 1: double */
  int threads;                /* <=0: OpenMP default */
  int keff_from_fission_rates;/* 1: reference default; 0: neutron balance */
  int allow_negative_fluxes;  /* 0: reference default */
  int stab_type;              /* -1 off; 0 DIAGONAL, 1 YAMAMOTO, 2 GLOBAL */
  double stab_factor;
} OracleOptions;

typedef struct {
  double keff;
  int num_iterations;
  double residual;
  double sweep_seconds;       /* wall time inside transport sweeps */
  int64_t num_integrations;   /* 2 * G * P * S * iterations */
} OracleResult;

/* Solver::computeEigenvalue (src/Solver.cpp:1542-1689) with CPUSolver steps.
 * flux_out: [N*G] converged, normalised scalar flux. Returns 0 on success. */
int oracle_compute_eigenvalue(const OracleProblem* p, const OracleOptions* o,
                              OracleResult* r, double* flux_out);

/* CPULSSolver::computeEigenvalue (linear source, 2D): src/CPULSSolver.cpp
 * with the LinearExpansionGenerator (src/TrackTraversingAlgorithms.cpp:
 * 470-831) and the FSR centroids of TrackGenerator::generateFSRCentroids
 * (src/TrackGenerator.cpp:1681-1735).  flux_xy_out (nullable): [N*G*2]
 * scalar-flux moments (x, y) per FSR and group.  Stabilisation and the
 * neutron-balance keff are not restated for the linear source (-3). */
int oracle_compute_eigenvalue_ls(const OracleProblem* p, const OracleOptions* o,
                                 OracleResult* r, double* flux_out, double* flux_xy_out,
                                 double* centroid_out);

/* Solver::computeFlux(max_iters, only_fixed_source=true):
 * src/Solver.cpp:1352-1420 (sources computed once from phi = 0). */
int oracle_compute_flux(const OracleProblem* p, const OracleOptions* o,
                        OracleResult* r, double* flux_out);

/* Solver::computeSource(max_iters, k_eff, res_type): src/Solver.cpp:1459-1516. */
int oracle_compute_source(const OracleProblem* p, const OracleOptions* o, double k_eff,
                          OracleResult* r, double* flux_out);
/* mode 0 eigenvalue, 1 flux, 2 source (k_eff); phi0 [N*G] / psi0 [2*T*fpt]
 * (nullable, flux mode only): the state a previous solve left, for
 * computeFlux(max_iters, only_fixed_source = false); psi_out (nullable)
 * receives the start fluxes at the end.  FP32 angular fluxes only. */
int oracle_solve_state(const OracleProblem* p, const OracleOptions* o, int mode, double k_eff,
                       const double* phi0, const float* psi0, OracleResult* r, double* flux_out,
                       float* psi_out);

/* One CPUSolver::transportSweep (src/CPUSolver.cpp:2338-2389) on given
 * reduced sources Q [N*G] and start fluxes psi_start [T*2*P*G] (float, the
 * reference layout (t*2+d)*fpt + p*G + g).  Writes the raw scalar-flux tally
 * [N*G] (before addSourceToScalarFlux) and updates psi_start in place with the
 * transferred outgoing fluxes. */
int oracle_transport_sweep(const OracleProblem* p, const double* Q, float* psi_start,
                           double* tally_out, int threads);

/* Solver::fissionTransportSweep (kind 1) / scatterTransportSweep (kind 2),
 * src/Solver.cpp:1282-1297: sources from phi (computeFSRFissionSources /
 * computeFSRScatterSources, src/CPUSolver.cpp:2030-2104), one sweep from
 * psi_start (updated in place), addSourceToScalarFlux into phi (in place). */
int oracle_krylov_sweep(const OracleProblem* p, int kind, int allow_negative, double* phi,
                        float* psi_start, int threads);

/* CPUSolver::computeFSRSources (src/CPUSolver.cpp:1939-2023). */
int oracle_compute_fsr_sources(const OracleProblem* p, const double* phi, double keff,
                               int iteration, int allow_negative, double* Q_out);

/* CPUSolver::addSourceToScalarFlux (src/CPUSolver.cpp:2608-2659), in place. */
int oracle_add_source_to_scalar_flux(const OracleProblem* p, const double* Q,
                                     int allow_negative, double* phi_inout);

/* Rational (1 - exp(-x))/x fit used by the reference (src/exponentials.h:156-192). */
double oracle_exp_f1(double x);

/* pairwise_sum (src/pairwise_sum.h:19-39) with VEC_LENGTH = 4 (FP64 build). */
double oracle_pairwise_sum(const double* v, int64_t n);

#ifdef __cplusplus
}
#endif

#endif
