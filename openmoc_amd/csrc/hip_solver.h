// HipSolver: the MOC source iteration on one MI355X (gfx950).
//
// Host-side counterpart of the reference Solver API (src/Solver.h:121-585),
// with every per-iteration step in hand-written HIP kernels:
//   transportSweep            -> sweep2d_kernel  (FP32, LDS-free, atomics)
//   addSourceToScalarFlux +
//   fission-rate partials     -> post_sweep_kernel (FP64)
//   computeKeff + normalise   -> keff_kernel (FP64, 1 block, fixed order)
//   normalise + residual +
//   storeFSRFluxes +
//   computeFSRSources (next)  -> update_kernel (FP64)
//   residual / convergence    -> residual_kernel (FP64, 1 block)
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <map>
#include <vector>

#include "../../include/openmoc_amd.h"

namespace moc {

// Iteration state kept on the device; the host reads it once per batch.
struct IterState {
  double k;             // current eigenvalue
  double k_prev;        // eigenvalue of the previous iteration
  double norm_factor;   // last normalisation factor (N / total fission)
  double psi_scale;     // factor applied to start fluxes on load
  double residual;
  double fission_total;
  int iteration;        // completed iterations of the current solve
  int converged;        // 1 once the stopping rule is met (2: iteration limit, batched solves)
  int dk;               // int(1e5 * (k - k_prev)) of the last iteration
  int reported_iters;   // Solver::_num_iterations semantics
  int64_t negative_fluxes;
  int64_t negative_sources;
  // Fixed-point LDS tally (see sweep2d_lds_kernel): contributions / V are
  // accumulated as int64 multiples of 1/tally_scale, with tally_scale set from
  // a bound on |psi| and |Q/sigma_t| before every sweep.
  double tally_scale;
  uint32_t psi_max_bits;    // max |outgoing psi| of the running sweep (float bits)
  uint32_t qmax_bits;       // max |Q/sigma_t| of the sources being computed (float bits)
  float qmax;               // max |Q/sigma_t| of the sources in use
  float psi_max;            // max |start psi| of the sweep in use (before psi_scale)
  double scale_cap;         // upper limit of the FP32 scale s32 (1e38 / max 1/V)
  int old_sel;              // which old-flux buffer holds the current normalised flux
  // iteration limit enforced on the device (batched solves; 0: none): the
  // solve stops with converged = 2 once `iteration` reaches it
  int max_iters;
};

struct LaunchShape {
  int groups_per_lane = 1;   // template parameter of the sweep
  int lanes_per_track = 1;   // ceil(G / groups_per_lane)
  int tracks_per_wave = 1;   // 64 / lanes_per_track
  int block = 256;           // threads per block (multiple of 64)
  int tally_copies = 8;      // private FP32 tallies (one per XCD)
  // LDS-privatised tally (used when lds_groups > 0): groups per subset
  // (= lanes per track), threads per block, blocks per group subset.
  int lds_groups = 0;
  int lds_block = 512;
  int lds_blocks_per_sub = 0;
  int lds_q = 0;             // 1: the subset's Q/sigma_t slice is staged in LDS too
  // 1: each block adds the non-zero entries of its LDS tally slice to one
  // global int64 tally with exact atomics; 0: it writes its whole slice to
  // a per-block partial buffer that the per-FSR step sums
  int lds_flush = 0;
  // global-tally sweep (lds_groups = 0): 1 stores each segment's weighted
  // delta-psi per group (no atomics) and the per-FSR step sums an FSR's
  // segments in a fixed order (deterministic); 0 adds them with float atomics
  // into tally_copies private tallies
  int seg_store = 0;
};

class DeviceError : public std::exception {
 public:
  DeviceError(int code, std::string msg) : code_(code), msg_(std::move(msg)) {}
  const char* what() const noexcept override { return msg_.c_str(); }
  int code() const { return code_; }

 private:
  int code_;
  std::string msg_;
};

// Plain device-to-device copy of `bytes` bytes, `repeats` times; returns the
// mean kernel time (s).  Calibrates HBM byte counters against a known size.
double deviceCopyCalibration(int device, int64_t bytes, int repeats);

struct FsrArgs;      // arguments of the per-FSR kernels
struct SweepArgs;    // arguments of the 2D sweep kernels
struct LdsParams;    // LDS-privatised tally of the 2D sweep
struct ScatArgs;     // scattering contraction on the matrix cores / FP32
struct BatchSlot;    // one solve's row in a batched launch
struct SweepBuffers; // per-sweep device buffers that replace the solver's own
class SolveBatch;

class HipSolver {
 public:
  explicit HipSolver(int device);
  ~HipSolver();
  HipSolver(const HipSolver&) = delete;
  HipSolver& operator=(const HipSolver&) = delete;

  void setTracks(const moc_track_data_t& tracks);
  void setMaterials(int num_materials, int num_groups, const double* sigma_t,
                    const double* sigma_s, const double* nu_sigma_f, const double* chi);
  void setFSRs(int64_t num_fsrs, const int32_t* material, const double* volume);
  void setConvergenceThreshold(double t);
  void setExponential(int mode);
  void allowNegativeFluxes(bool allow) {
    allow_negative_ = allow;
    graph_dirty_ = true;
  }
  // Solver::stabilizeTransport(factor, type) (src/Solver.cpp:676-681);
  // type < 0 switches stabilisation off.
  void stabilizeTransport(double factor, int type);
  // Solver::setKeffFromNeutronBalance (src/Solver.cpp:2047-2049).
  void setKeffFromNeutronBalance(bool on) {
    balance_ = on;
    graph_dirty_ = true;
  }
  // Solver::fissionTransportSweep / scatterTransportSweep (src/Solver.cpp:
  // 1282-1297): sources from the current flux, one sweep from the current
  // start fluxes, addSourceToScalarFlux.
  void krylovSweep(int source_kind);
  // Linear source (CPULSSolver): on/off; needs the track data's azimuthal
  // angles and weights, polar weights and segment start points.
  void setLinearSource(bool on);
  bool linearSource() const { return linear_; }
  // Normalised scalar-flux moments [N*G*2] (x, y) of the last linear-source solve.
  void getFluxMoments(double* out, int64_t n);
  // FSR centroids [2N] used by the linear source.
  void getCentroids(double* out, int64_t n);

  // Uses the tracks, FSRs and launch shape of `src` (device buffers shared,
  // not copied) in this fresh solver; materials stay per solver.  For many
  // solves of one geometry with different cross sections (serving).
  void shareGeometry(HipSolver& src);
  // Asynchronous solve: startSolve launches the first batch of iterations;
  // pollSolve returns true once the solve has finished (results ready),
  // launching the next batch when the previous one completed.  mode 0
  // eigenvalue, 1 flux, 2 source (k_fixed).
  void startSolve(int mode, int max_iters, int res_type, double k_fixed);
  bool pollSolve();
  bool solveActive() const { return async_active_; }
  // Device bytes of one solve's own arrays (shared geometry excluded).
  size_t solveBytes() const;
  void useGraphs(bool on) { use_graphs_ = on; }
  int device() const { return device_; }
  void setFixedSourceByFSR(int64_t fsr, int group, double source);
  void setFixedSourceByMaterial(int material, int group, double source);
  void setFSRCells(const int32_t* cell, int64_t n);
  void setFixedSourceByCell(int cell, int group, double source);
  void resetFixedSources();
  // Arithmetic of the scattering-source contraction (MOC_SCATTER_*);
  // scatterPrecisionInUse resolves MOC_SCATTER_AUTO for the current groups.
  void setScatterPrecision(int mode);
  // computeFSRSources(iteration) from the current flux and keff; out [N*G]
  // receives the sources per steradian (the reference's _reduced_sources).
  void computeFSRSources(int iteration, double* out, int64_t n);
  int scatterPrecisionInUse() const;

  void computeEigenvalue(int max_iters, int res_type);
  void computeFlux(int max_iters, bool only_fixed_source);
  void computeSource(int max_iters, double k_eff, int res_type);

  double getKeff() const { return keff_; }
  int getNumIterations() const { return num_iterations_; }
  double getFlux(int64_t fsr, int group);
  void getFluxes(double* out, int64_t n);
  void setFluxes(const double* in, int64_t n);
  void computeFSRFissionRates(double* out, int64_t n);
  void timing(double* sweep_s, double* total_s, int64_t* integrations) const;

  void transportSweepOnce(const double* Q, float* psi, double* tally);
  // transportSweepOnce on device buffers (Q [N*G], psi_in/psi_out in the
  // reference layout, tally [N*G]), asynchronous on `s` (null: the solver's
  // stream), allocating nothing: the sweep's own buffers (start-flux slots,
  // Q/sigma_t table, fixed-point tally, iteration state) are carved from a
  // static device array of kNoAllocScratchBytes per GPU, shared by every call
  // on that GPU (overlapping calls on one GPU race).
  void transportSweepNoAlloc(const double* Q, const float* psi_in, float* psi_out,
                             double* tally, hipStream_t s);
  // Bytes of that device array a call on this solver needs.
  size_t noAllocScratchBytes() const;
  // Runs n eigenvalue source iterations with the stopping rule disabled
  // (benchmark form), as a solve runs them (graph batches).  Returns the mean
  // device time per iteration; *sweep_seconds (nullable) receives the mean
  // time of the sweep kernel alone, from events around it in up to 8 more
  // iterations.
  double runIterations(int n, double* sweep_seconds);
  double timeSweeps(int repeats);

  const LaunchShape& shape() const { return shape_; }
  void setShape(const LaunchShape& s);

  // FSR centroids [3N] (x, y, z) that identify FSRs in flux files.
  void setFSRCentroids(const double* xyz, int64_t n);
  // Solver::dumpFSRFluxes (src/Solver.cpp:2066-2112): k_eff, centroids and
  // the current scalar fluxes.
  void dumpFSRFluxes(const std::string& path);
  // Solver::loadFSRFluxes (src/Solver.cpp:2135-2318): scalar fluxes of the
  // file FSRs nearest to this solver's centroids become the current fluxes;
  // returns the largest centroid mismatch.
  double loadFSRFluxes(const std::string& path, bool assign_k_eff, double tolerance);
  // Solver::loadInitialFSRFluxes (src/Solver.cpp:2120-2123): eigenvalue and
  // source solves start from the file's fluxes and k_eff, normalised, after 30
  // sweeps on their sources (src/Solver.cpp:1733-1751).  "" switches it off.
  void loadInitialFSRFluxes(const std::string& path);
  // Solver::setResidualByReference (src/Solver.cpp:2056-2059): residuals
  // compare the flux with the file's fluxes instead of the previous
  // iteration's.  "" switches it off.
  void setResidualByReference(const std::string& path);
  // Solver::setRestartStatus (src/Solver.cpp:610-615): the next eigenvalue
  // and source solves keep k_eff (eigenvalue), the angular fluxes and the
  // scalar flux of the previous solve instead of starting from a flat guess
  // (src/Solver.cpp:1554-1572, 1721-1731).
  void setRestartStatus(bool on) { restart_status_ = on; }
  // Solver::setSolverMode (src/Solver.cpp:597-602): MOC_FORWARD or
  // MOC_ADJOINT (transposed scattering and fission matrices).
  void setSolverMode(int mode);
  // Solver::setChiSpectrumMaterial (src/Solver.cpp:708-710): the initial flux
  // guess of eigenvalue and source solves is that material's chi in every FSR
  // (src/Solver.cpp:1722-1725); -1 restores the flat guess.
  void setChiSpectrumMaterial(int material);

 private:
  enum Mode { MODE_EIGEN = 0, MODE_FLUX = 1, MODE_SOURCE = 2 };
  void requireReady() const;
  void allocateIterationArrays();
  void runSolve(Mode mode, int max_iters, int res_type, double k_fixed);
  void launchSweep(const float* psi_in, float* psi_out, hipStream_t s,
                   hipEvent_t ev_after_sweep = nullptr, float* leak = nullptr,
                   const SweepBuffers* bufs = nullptr);
  void launchIteration(Mode mode, int res_type, bool recompute_sources, double tol,
                       hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
  void launchInit(Mode mode, double k0);
  void readFluxFiles(Mode mode);
  void launchStartupSweeps();
  void launchRestart(Mode mode, double k0);
  void uploadProductionMatrices(int M, int G);
  void zeroTallies(hipStream_t s);
  std::vector<double> matchedFlux(const std::string& path, double tolerance, double* keff,
                                  double* max_error);
  void ensureGraph(int mode, int res_type);
  void launchIterations(int mode, int res_type, int n);
  void launchAsyncBatch();
  void finishSolve(int max_iters);
  void prepareLinearSource();
  void prepareScatter();
  void buildScatterTiles();
  void launchScatter(hipStream_t s);
  void buildPositionStream();
  FsrArgs fsrArgs();
  void sweepArgs(SweepArgs& a, const float* psi_in, float* psi_out, float* leak,
                 const SweepBuffers* bufs) const;
  void ldsParams(LdsParams& L, const SweepBuffers* bufs) const;
  bool scatArgs(ScatArgs& a) const;
  friend class SolveBatch;
  void refreshStabilization();
  const double* currentFlux() const;
  void pickShape();
  void countFissionable();
  void buildTileLayout(int tracks_per_wave);
  int sweepTracksPerWave() const;
  size_t ldsBytes(int groups, int with_q) const;

  int device_ = 0;
  hipStream_t stream_ = nullptr;

  // sizes
  int G_ = 0, M_ = 0, P_ = 0, num_azim_ = 0, num_polar_ = 0;
  bool is3d_ = false;                  // 3D tracks: P_ = 1, weight class per track
  // 3D z-stack groups (tracks over a pair of neighbouring 2D tracks), for the
  // LDS stack sweep
  int32_t ngrp_ = 0;                   // 0: no groups (2D, or no xy given)
  int32_t max_grp_slots_ = 0;          // most FSRs touched by one group
  std::vector<int32_t> h_trk_grp_;     // [T] group of each track
  std::vector<int32_t> h_slot_;        // [S] segment -> slot among its group's FSRs
  std::vector<int32_t> h_grp_slot0_, h_grp_nslots_, h_slot_fsr_;
  int32_t* d_grp_ = nullptr;           // [ngrp][4] first tile, tiles, first slot, slots
  int32_t* d_slot_fsr_ = nullptr;      // [total slots] global FSR of each slot
  unsigned long long* d_tally_i64_ = nullptr;  // [N*G] fixed-point tally (stack sweep)
  bool stackMode() const { return is3d_ && ngrp_ > 0 && shape_.lds_groups > 0; }
  // z-stack per-slot sums: entries per slot row (G rounded up to 8)
  int stackRow() const { return (G_ + 7) / 8 * 8; }
  // the sweep tallies into the global int64 fixed-point tally d_tally_i64_
  // (3D z-stack sweep, or the 2D LDS sweep with the atomic flush)
  bool int64Tally() const { return stackMode() || (shape_.lds_groups > 0 && shape_.lds_flush); }
  int64_t T_ = 0, S_ = 0, N_ = 0;
  int64_t num_fissionable_ = 0;
  bool tracks_set_ = false, materials_set_ = false, fsrs_set_ = false;
  double threshold_ = 1e-5;
  int exp_mode_ = MOC_EXP_RATIONAL;
  bool allow_negative_ = false;
  LaunchShape shape_;
  bool shape_forced_ = false;

  // device arrays (raw pointers owned by this object)
  float* d_seg_len_ = nullptr;
  int32_t* d_seg_key_ = nullptr;       // tile-interleaved FSR | material << 24
  int32_t* d_work_ = nullptr;          // [ntiles*tpw] directed tracks, longest first (-1 pad)
  int32_t* d_tile_ = nullptr;          // [ntiles][2] first row, rows
  int32_t* d_trk_ = nullptr;           // [T][4] seg_begin, seg_count, out_fwd, out_bwd
  int32_t* d_trk_azim_ = nullptr;
  float* d_weights_ = nullptr;         // [num_azim/2][P]
  float inv_sin_[4] = {0, 0, 0, 0};
  float* d_psi_[2] = {nullptr, nullptr};
  float* d_tally_ = nullptr;           // [N*G]
  float2* d_qs_ = nullptr;             // [N*G] (Q/sigma_t, sigma_t)
  double* d_phi_ = nullptr;            // [N*G]
  double* d_phi_old_[2] = {nullptr, nullptr};  // [N*G] normalised fluxes, ping-pong
  double* d_Q_ = nullptr;              // [N*G]
  double* d_fixed_ = nullptr;          // [N*G]
  double* d_sigma_t_ = nullptr;        // [M*G]
  double* d_sigma_s_ = nullptr;        // [M*G*G]
  double* d_nsf_ = nullptr;            // [M*G]
  double* d_chi_ = nullptr;            // [M*G] fission-source spectrum (nu-sigma-f when adjoint)
  double* d_fis_nsf_ = nullptr;        // [M*G] fission-source production (chi when adjoint)
  int32_t* d_mat_fissionable_ = nullptr;  // [M]
  int32_t* d_fsr_mat_ = nullptr;       // [N]
  double* d_fsr_vol_ = nullptr;        // [N]
  float* d_inv_vol_ = nullptr;         // [N] 1/V (0 for zero volume)
  double sigma_t_max_ = 0.0;
  double scale_cap_ = 1e38;       // 1e38 / max(1/V): keeps (1/V) * s32 finite in FP32
  double* d_partials_ = nullptr;       // [max blocks]
  unsigned long long* d_lds_partials_ = nullptr;  // [nsub][blocks_per_sub][N*GB] fixed point
  float* d_qsub_ = nullptr;            // [nsub][N][GB] Q/sigma_t for the LDS sweep
  double* d_sigma_a_ = nullptr;        // [M*G] absorption
  float* d_leak_ = nullptr;            // [2T*G] vacuum leakage (neutron-balance keff)
  double* d_stab_ratio_ = nullptr;     // [G] YAMAMOTO ratios
  std::vector<double> h_sigma_t_, h_sigma_s_;  // host copies (dest-major scattering in use)
  std::vector<double> h_sigma_s_fwd_, h_nsf_fwd_, h_chi_fwd_;  // the materials as given
  int solver_mode_ = MOC_FORWARD;
  int chi_material_ = -1;              // setChiSpectrumMaterial
  std::vector<double> h_chi_guess_;    // [N*G] staging of that guess
  int stab_type_ = -1;
  double stab_factor_ = 0.0;
  bool balance_ = false;
  bool phi_is_old_ = false;            // current flux lives in d_phi_old_[old_sel] (after a solve)
  int old_sel_ = 0;                    // host copy of IterState::old_sel after a solve
  bool krylov_ready_ = false;          // start fluxes initialised for Krylov sweeps
  int pstride_ = 0;                    // stride of the per-block partial arrays
  // linear source
  bool linear_ = false;
  bool ls_ready_ = false;              // constants and position stream are current
  std::vector<double> h_seg_start_;    // [2S]
  std::vector<double> h_azim_phi_, h_azim_w_, h_polar_w_;
  std::vector<double> h_centroid_;     // [2N]
  std::vector<double> h_fsr_xyz_;      // [3N] centroids naming FSRs in flux files
  std::string init_flux_file_;         // loadInitialFSRFluxes
  std::string ref_flux_file_;          // setResidualByReference
  std::vector<double> h_init_flux_;    // [N*G] loaded at solve start
  double init_keff_ = 1.0;
  double* d_ref_flux_ = nullptr;       // [N*G] reference flux of the residual
  size_t ref_flux_n_ = 0;              // values read into d_ref_flux_
  bool restart_status_ = false;        // setRestartStatus
  bool restart_active_ = false;        // the running solve continues the previous one
  bool continue_flux_ = false;         // computeFlux(only_fixed_source = false) after a solve
  bool solved_ = false;                // a solve finished since the arrays were laid out
  int psi_cur0_ = 0;                   // start-flux buffer at the running solve's start
  double graph_tol_ = -1.0;            // stopping threshold the iteration graphs were captured with
  std::vector<int32_t> h_work_, h_tile_;  // host copies of the stream layout
  std::vector<int32_t> h_wpos_;        // [2T] work position of each directed track
  int32_t* d_out_wp_ = nullptr;        // [2T] work position of the linked start slot
  int64_t psi_w_ = 0;                  // start-flux slots per group (ntiles * tpw)
  size_t psiCount() const { return (size_t)G_ * psi_w_ * P_; }
  float2* d_seg_pos_ = nullptr;        // stream positions (LS)
  float* d_azim_cos_ = nullptr;
  float* d_azim_sin_ = nullptr;
  double* d_ilem_ = nullptr;           // [N*3]
  double* d_src_c_ = nullptr;          // [N*3*G]
  double* d_phi_xy_ = nullptr;         // [N*G*2]
  double* d_phi_xy_n_ = nullptr;       // [N*G*2]
  double* d_Qxy_ = nullptr;            // [N*G*2]
  float4* d_qls_ = nullptr;            // [N*G]
  float* d_tally_xy_ = nullptr;        // [copies][2][N*G] x and y planes
  // scattering contraction on the matrix cores / FP32 fallback
  int scatter_mode_ = MOC_SCATTER_AUTO;
  float* d_scat_ = nullptr;            // [C][N][G+1]
  int32_t* d_scat_rows_ = nullptr;     // [tiles*16] FSRs grouped by material (-1 pad)
  int32_t* d_scat_tile_mat_ = nullptr; // [tiles]
  uint32_t* d_scat_b_ = nullptr;       // [M][ks][nt][64] x 4 words of split fp16 fragments
  float* d_scat_colmul_ = nullptr;     // [M][nt*16]
  double* d_ss_colsum_ = nullptr;      // [M*G] sum over destinations of sigma_s
  int64_t scat_ntiles_ = 0;
  bool scat_tiles_ok_ = false, scat_frag_ok_ = false;
  struct GeoHold;
  std::shared_ptr<GeoHold> geo_hold_;  // set when the geometry buffers are shared
  // captured iteration batch
  // iteration graphs, one per start-flux buffer the batch begins at (the
  // sweeps alternate the two buffers, so each parity needs its own capture)
  hipGraphExec_t graph_exec_[2] = {nullptr, nullptr};
  void dropGraphs();
  int graph_mode_ = -1, graph_res_type_ = -1;
  bool graph_dirty_ = true;
  bool use_graphs_ = true;
  // asynchronous solve
  bool async_active_ = false;
  hipEvent_t async_ev_ = nullptr, async_ev0_ = nullptr;
  int async_mode_ = 0, async_res_type_ = 0, async_max_ = 0, async_done_ = 0;
  double async_ms_ = 0.0;
  std::chrono::steady_clock::time_point async_t0_;
  double* d_tally64_ = nullptr;        // [N*G] reduced LDS-path tally
  // per-segment store of the global-tally sweep (shape_.seg_store)
  bool segStore() const { return shape_.seg_store && shape_.lds_groups == 0 && !is3d_; }
  unsigned long long* d_stack_partial_ = nullptr;  // z-stack sweep: [slots][G] per-slot sums
  int64_t* d_stack_off_ = nullptr;                 // [N+1] first entry of each FSR's slot list
  int32_t* d_stack_slots_ = nullptr;               // slots ordered by FSR
  float* d_segout_ = nullptr;          // [planes][seg_rows_][seg_stride_]
  int64_t* d_csr_off_ = nullptr;       // [N+1] each FSR's rows of d_segout_ (geometry)
  int32_t* d_seg_slot_ = nullptr;      // tile-interleaved: each segment's row (geometry)
  bool csr_ready_ = false;             // d_csr_off_, d_seg_slot_ match the stream layout
  int seg_stride_ = 0;                 // floats per row: G
  int64_t seg_rows_ = 0;               // stored segment passes + the padding rows
  void buildSegmentIndex();
  int num_cus_ = 0;
  size_t lds_max_ = 0;                 // max dynamic LDS per block (bytes)
  size_t lds_per_cu_ = 0;
  IterState* d_state_ = nullptr;
  IterState* h_state_ = nullptr;       // pinned
  int num_partials_ = 0;
  int psi_cur_ = 0;                    // which psi buffer holds the start fluxes
  bool state_valid_ = false;           // an iteration state exists on the device

  int32_t max_seg_fsr_ = -1;
  std::vector<float> h_len_;             // host copies for the stream layout
  std::vector<int32_t> h_fsr_, h_trk_, h_order_;
  int layout_tpw_ = 0, layout_ntiles_ = 0;
  bool layout_stack_ = false;            // streams laid out by z-stack group (3D)
  bool layout_shared_ = false;           // one stream for both directions of a track
  std::vector<int32_t> layout_grp_ntiles_;
  int64_t layout_rows_ = 0;
  std::vector<int32_t> mat_fissionable_;  // host copy, per material
  std::vector<int32_t> h_fsr_mat_;        // host copy, per FSR
  std::vector<double> h_fsr_vol_;         // host copy, per FSR
  std::vector<double> h_fixed_;        // host copy of fixed sources
  bool fixed_on_ = false;
  // Fixed sources as the reference keeps them (Solver::_fix_src_FSR_map and
  // _fix_src_material_map, src/Solver.cpp:479-543): folded into the per-FSR
  // array when a solve starts (Solver::initializeFixedSources,
  // src/Solver.cpp:1089-1137, CPUSolver::initializeFixedSources,
  // src/CPUSolver.cpp:425-456), so a material's source wins over an
  // overlapping per-FSR one whatever the call order.
  std::map<std::pair<int64_t, int>, double> fix_fsr_map_;
  std::map<std::pair<int, int>, double> fix_mat_map_;
  std::map<std::pair<int, int>, double> fix_cell_map_;  // Solver::_fix_src_cell_map
  std::vector<int32_t> h_fsr_cell_;                      // cell of each FSR (findCellContainingFSR)
  bool fixed_dirty_ = false;
  void applyFixedSources();

  // results
  double keff_ = 1.0;
  int num_iterations_ = 0;
  double sweep_seconds_ = 0, total_seconds_ = 0;
  int64_t integrations_ = 0;

  // batched solves (SolveBatch)
  // New cross sections of the same material and group counts, with nothing
  // allocated and nothing synchronous: the tables are staged in pinned host
  // memory and copied on the solver's stream (a re-armed row solver).
  void rearmMaterials(const double* sigma_t, const double* sigma_s, const double* nu_sigma_f,
                      const double* chi);
  double* h_xs_stage_ = nullptr;       // pinned staging of rearmMaterials
  size_t h_xs_stage_n_ = 0;
  bool stream_borrowed_ = false;       // stream_ belongs to a SolveBatch
  int device_max_iters_ = 0;           // IterState::max_iters of the next launchInit
  void adoptStream(hipStream_t s);
};

// A solve of a SolveBatch that has stopped.
struct BatchResult {
  void* tag = nullptr;
  double keff = 0.0;
  int num_iterations = 0;
  std::vector<double> flux;  // normalised FSR scalar flux [N*G] when requested
};

// Eigenvalue solves advanced together (serving).  Every launch of an
// iteration step carries all live solves, one row each of a device table of
// per-solve kernel arguments (blockIdx.y = row); the steps of kGraphIters
// iterations are one captured graph per live-row count.  The rows may belong
// to different geometries whose kernels are the same (batchKey): each row
// brings its own tracks, FSRs and cross sections, launches are sized for the
// largest member, and a row's blocks beyond its own work return at entry.
// Each solve stops on the device (stopping rule or iteration limit), after
// which its rows return at entry, and a finished row is refilled with the
// next request between graph launches (continuous batching).  Every row keeps
// its angular and scalar fluxes resident; its arithmetic is that of a plain
// HipSolver solve of the same inputs.
class SolveBatch {
 public:
  SolveBatch(int device, int capacity, int res_type);
  ~SolveBatch();
  SolveBatch(const SolveBatch&) = delete;
  SolveBatch& operator=(const SolveBatch&) = delete;

  // The kernels (and their template arguments) a geometry's solves run; only
  // geometries with equal keys can share a batch.  Throws for geometries a
  // batch cannot run (3D, linear source).
  static std::string batchKey(const HipSolver& geometry);
  // Whether a geometry's solves are small enough to share launches with
  // other geometries' (an LDS sweep of at most 64 blocks and 64 KB of LDS per
  // solve); larger ones fill launches on their own.
  static bool shareable(const HipSolver& geometry);
  // Lets this batch run solves on `geometry` (tracks, FSRs and materials set;
  // its rows share its geometry buffers and launch shape).
  void addGeometry(HipSolver& geometry);
  bool hasGeometry(const HipSolver* geometry) const;

  int capacity() const { return capacity_; }
  int live() const { return (int)live_.size(); }
  bool busy() const { return !live_.empty() || !inflight_.empty(); }
  bool hasIdleSolver(const HipSolver* geometry) const;
  size_t bytesHeld() const;
  // Starts an eigenvalue solve on `geometry` with these cross sections in a
  // free row (its initial state is queued behind the launches in flight);
  // `tag` names it in the results.  Creates a row solver when none is idle.
  void start(HipSolver* geometry, void* tag, int M, int G, const double* sigma_t,
             const double* sigma_s, const double* nu_sigma_f, const double* chi, int max_iters,
             double tol, bool want_flux);
  // Collects the solves that stopped in completed launches and keeps up to
  // two graph launches in flight.  Never waits for the device.
  void progress(std::vector<BatchResult>& done);
  // Waits for the launches in flight and collects them (a batch without
  // live solves only; their rows have all stopped).
  void drain();
  // Frees the idle row solvers that no launch in flight refers to; returns
  // the bytes freed.
  size_t releaseIdle();

 private:
  struct Row {
    void* tag;
    uint64_t serial;  // this batch's number of the solve (tags may be reused)
    int si;           // row solver: solvers_[si], its table entry slots_[si]
    int max_iters;
    bool want_flux;
  };
  struct Launch {
    hipEvent_t ev;
    int buf;
    std::vector<std::pair<uint64_t, int>> rows;  // (serial, solver) of row j of this launch
  };
  // Launch sizes: the largest over the member geometries.
  struct Dims {
    int sweep_blocks = 0, sweep_threads = 0;
    size_t lds_bytes = 0;
    int post_blocks = 0, upd_blocks = 0;
    int mfma_blocks = 0, fp32_blocks = 0, fsr_blocks = 0;  // 0: no member needs the kernel
  };
  void fillRow(BatchSlot& b, HipSolver& s, double tol) const;
  void setRow(int row, const BatchSlot& b);
  void launchStep(hipStream_t s, int n, int parity);
  hipGraphExec_t graphFor(int n);
  void launchGraph();
  void finish(size_t idx, const IterState& st, std::vector<BatchResult>& done);
  int newSolver(HipSolver* geometry);

  int device_ = 0;
  int capacity_ = 0;
  int res_type_ = MOC_FISSION_SOURCE;
  std::string key_;
  std::vector<HipSolver*> geos_;       // member geometries (not owned)
  Dims dims_;
  hipStream_t stream_ = nullptr;
  hipStream_t capture_stream_ = nullptr;
  BatchSlot* d_table_ = nullptr;       // [capacity]
  IterState* d_snap_ = nullptr;        // [2][capacity]
  IterState* h_snap_ = nullptr;        // [2][capacity] pinned
  int next_buf_ = 0;
  uint64_t next_serial_ = 1;
  std::vector<std::unique_ptr<HipSolver>> solvers_;  // null where one was released
  std::vector<HipSolver*> solver_geo_;               // geometry of each solver
  std::vector<std::unique_ptr<BatchSlot>> slots_;    // host copy of each solver's row
  std::vector<int> idle_;              // solvers without a solve
  std::vector<Row> live_;              // row j of the table = live_[j]
  std::vector<Launch> inflight_;       // oldest first
  std::vector<hipEvent_t> free_events_;
  std::map<int, hipGraphExec_t> graphs_;    // by live-row count, for dims_
  std::vector<hipGraphExec_t> retired_;     // graphs of smaller dims, destroyed when idle
};

}  // namespace moc
