// extern "C" boundary of libopenmoc_amd: status codes + thread-local error
// text; no exception crosses this file.
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/openmoc_amd.h"
#include "fluxfile.h"
#include "segfile.h"
#include "hip_solver.h"
#include "serve.h"
#include "tracks.h"

struct moc_tracks {
  moc::Problem problem;
  moc::TrackSet ts;
  std::string text;  // the problem text the tracks were made from
};

struct moc_solver {
  moc::HipSolver* impl;
};

struct moc_engine {
  moc::ServeEngine* impl;
};

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return MOC_OK;
  } catch (const moc::DeviceError& e) {
    return fail(e.code(), e.what());
  } catch (const std::bad_alloc&) {
    return fail(MOC_ERR_MEMORY, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(MOC_ERR_INVALID, e.what());
  } catch (...) {
    return fail(MOC_ERR_INVALID, "unknown error");
  }
}

int residual_code(const std::string& s) {
  if (s == "SCALAR_FLUX") return MOC_SCALAR_FLUX;
  if (s == "TOTAL_SOURCE") return MOC_TOTAL_SOURCE;
  return MOC_FISSION_SOURCE;
}

template <typename T>
void copy_out(T* dst, const std::vector<T>& src) {
  if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(T));
}

}  // namespace

extern "C" {

const char* moc_last_error(void) { return g_last_error.c_str(); }

const char* moc_version(void) { return "openmoc_amd 0.1 (gfx950)"; }

// ---------------------------------------------------------------- tracks
int moc_tracks_generate(const char* problem_text, int threads, moc_tracks_t** out) {
  if (!problem_text || !out) return fail(MOC_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    auto* t = new moc_tracks;
    try {
      t->text = problem_text;
      t->problem = moc::parse_problem(problem_text);
      t->ts = moc::generate_tracks(t->problem, threads);
      // the exponential evaluators support optical lengths up to 100
      // (MAX_OPTICAL_LENGTH, src/constants.h; Solver::initializeExpEvaluators)
      moc::split_segments(t->ts, t->problem, 100.0);
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

void moc_tracks_destroy(moc_tracks_t* tracks) { delete tracks; }

int moc_tracks_dump_segments(const moc_tracks_t* t, const char* path) {
  if (!t || !path || !*path) return fail(MOC_ERR_INVALID, "null argument");
  return guarded([&] { moc::dump_segment_file(path, t->text, t->ts); });
}

int moc_tracks_read_segments(const char* problem_text, const char* path, moc_tracks_t** out) {
  if (!problem_text || !path || !*path || !out) return fail(MOC_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    auto* t = new moc_tracks;
    try {
      t->text = problem_text;
      t->problem = moc::parse_problem(problem_text);
      if (!moc::read_segment_file(path, t->text, t->ts)) {
        delete t;
        return;  // written for another problem
      }
      if (t->ts.num_groups != t->problem.num_groups ||
          t->ts.fsr_material.size() != t->ts.fsr_volume.size())
        throw moc::DeviceError(MOC_ERR_IO, std::string("segment file ") + path +
                                               " does not match its problem text");
      for (int32_t m : t->ts.fsr_material)
        if (m < 0 || m >= (int32_t)t->problem.materials.size())
          throw moc::DeviceError(MOC_ERR_IO, std::string("segment file ") + path +
                                                 ": material id out of range");
    } catch (...) {
      delete t;
      throw;
    }
    *out = t;
  });
}

int moc_tracks_info(const moc_tracks_t* t, moc_tracks_info_t* info) {
  if (!t || !info) return fail(MOC_ERR_INVALID, "null argument");
  info->num_groups = t->problem.num_groups;
  info->num_materials = (int32_t)t->problem.materials.size();
  info->num_azim = t->problem.num_azim;
  info->num_polar = t->problem.num_polar;
  info->num_tracks = t->ts.num_tracks;
  info->num_segments = (int64_t)t->ts.seg_length.size();
  info->num_fsrs = t->ts.num_fsrs;
  info->max_optical_length = t->ts.max_optical_length;
  info->tolerance = t->problem.tolerance;
  info->max_iters = t->problem.max_iters;
  info->res_type = residual_code(t->problem.res_type);
  info->is_3d = t->ts.is_3d ? 1 : 0;
  info->num_layers = t->ts.num_layers;
  return MOC_OK;
}

int moc_tracks_export_tracks_3d(const moc_tracks_t* t, int32_t* polar, double* start_z,
                                double* end_z, double* theta, double* polar_spacing,
                                double* fsr_centroid_z) {
  if (!t) return fail(MOC_ERR_INVALID, "null argument");
  if (!t->ts.is_3d) return fail(MOC_ERR_INVALID, "not a 3D track set");
  copy_out(polar, t->ts.polar);
  copy_out(start_z, t->ts.start_z);
  copy_out(end_z, t->ts.end_z);
  copy_out(theta, t->ts.quad.theta);
  copy_out(polar_spacing, t->ts.quad.polar_spacing);
  copy_out(fsr_centroid_z, t->ts.fsr_centroid_z);
  return MOC_OK;
}

int moc_tracks_export_tracks(const moc_tracks_t* t, int64_t* seg_offset, int32_t* azim,
                             int32_t* xy, double* start_xy, double* end_xy, int64_t* next_fwd,
                             int64_t* next_bwd, uint8_t* next_fwd_is_fwd,
                             uint8_t* next_bwd_is_fwd, uint8_t* bc_fwd, uint8_t* bc_bwd) {
  if (!t) return fail(MOC_ERR_INVALID, "null argument");
  const moc::TrackSet& s = t->ts;
  copy_out(seg_offset, s.seg_offset);
  copy_out(azim, s.azim);
  copy_out(xy, s.xy);
  if (start_xy)
    for (long i = 0; i < s.num_tracks; i++) {
      start_xy[2 * i] = s.start_x[i];
      start_xy[2 * i + 1] = s.start_y[i];
    }
  if (end_xy)
    for (long i = 0; i < s.num_tracks; i++) {
      end_xy[2 * i] = s.end_x[i];
      end_xy[2 * i + 1] = s.end_y[i];
    }
  copy_out(next_fwd, s.next_fwd);
  copy_out(next_bwd, s.next_bwd);
  copy_out(next_fwd_is_fwd, s.next_fwd_is_fwd);
  copy_out(next_bwd_is_fwd, s.next_bwd_is_fwd);
  copy_out(bc_fwd, s.bc_fwd);
  copy_out(bc_bwd, s.bc_bwd);
  return MOC_OK;
}

int moc_tracks_export_segments(const moc_tracks_t* t, double* length, int32_t* fsr) {
  if (!t) return fail(MOC_ERR_INVALID, "null argument");
  copy_out(length, t->ts.seg_length);
  copy_out(fsr, t->ts.seg_fsr);
  return MOC_OK;
}

int moc_tracks_export_segment_starts(const moc_tracks_t* t, double* start_xy) {
  if (!t) return fail(MOC_ERR_INVALID, "null argument");
  copy_out(start_xy, t->ts.seg_start);
  return MOC_OK;
}

int moc_tracks_export_fsrs(const moc_tracks_t* t, int32_t* material, double* volume,
                           double* centroid, int64_t* key) {
  if (!t) return fail(MOC_ERR_INVALID, "null argument");
  copy_out(material, t->ts.fsr_material);
  copy_out(volume, t->ts.fsr_volume);
  copy_out(centroid, t->ts.fsr_centroid);
  copy_out(key, t->ts.fsr_key);
  return MOC_OK;
}

int moc_tracks_export_quadrature(const moc_tracks_t* t, double* phi, double* spacing,
                                 double* azim_weight, double* sin_theta, double* polar_weight,
                                 double* total_weight) {
  if (!t) return fail(MOC_ERR_INVALID, "null argument");
  const moc::Quadrature& q = t->ts.quad;
  copy_out(phi, q.phi);
  copy_out(spacing, q.azim_spacing);
  copy_out(azim_weight, q.azim_weight);
  copy_out(sin_theta, q.sin_theta);
  copy_out(polar_weight, q.polar_weight);
  copy_out(total_weight, q.total_weight);
  return MOC_OK;
}

int moc_tracks_export_materials(const moc_tracks_t* t, double* sigma_t, double* sigma_s,
                                double* nu_sigma_f, double* chi) {
  if (!t) return fail(MOC_ERR_INVALID, "null argument");
  const int G = t->problem.num_groups;
  size_t m = 0;
  for (const moc::MaterialData& md : t->problem.materials) {
    if (sigma_t) std::memcpy(sigma_t + m * G, md.sigma_t.data(), G * sizeof(double));
    if (sigma_s) std::memcpy(sigma_s + m * G * G, md.sigma_s.data(), (size_t)G * G * sizeof(double));
    if (nu_sigma_f) std::memcpy(nu_sigma_f + m * G, md.nu_sigma_f.data(), G * sizeof(double));
    if (chi) std::memcpy(chi + m * G, md.chi.data(), G * sizeof(double));
    m++;
  }
  return MOC_OK;
}

// ---------------------------------------------------------------- solver
int moc_solver_create(int device, moc_solver_t** out) {
  if (!out) return fail(MOC_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    auto* s = new moc_solver;
    s->impl = nullptr;
    try {
      s->impl = new moc::HipSolver(device);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
  });
}

void moc_solver_destroy(moc_solver_t* s) {
  if (!s) return;
  delete s->impl;
  delete s;
}

#define MOC_SOLVER_CALL(solver, body)                                   \
  do {                                                                  \
    if (!(solver) || !(solver)->impl) return fail(MOC_ERR_INVALID, "null solver"); \
    return guarded([&] { body; });                                      \
  } while (0)

int moc_solver_set_tracks(moc_solver_t* s, const moc_track_data_t* tracks) {
  if (!tracks) return fail(MOC_ERR_INVALID, "null track data");
  MOC_SOLVER_CALL(s, s->impl->setTracks(*tracks));
}

int moc_solver_set_materials(moc_solver_t* s, int32_t num_materials, int32_t num_groups,
                             const double* sigma_t, const double* sigma_s,
                             const double* nu_sigma_f, const double* chi) {
  MOC_SOLVER_CALL(s, s->impl->setMaterials(num_materials, num_groups, sigma_t, sigma_s,
                                           nu_sigma_f, chi));
}

int moc_solver_set_fsrs(moc_solver_t* s, int64_t num_fsrs, const int32_t* material,
                        const double* volume) {
  MOC_SOLVER_CALL(s, s->impl->setFSRs(num_fsrs, material, volume));
}

int moc_solver_set_convergence_threshold(moc_solver_t* s, double threshold) {
  MOC_SOLVER_CALL(s, s->impl->setConvergenceThreshold(threshold));
}

int moc_solver_set_exponential(moc_solver_t* s, int mode) {
  MOC_SOLVER_CALL(s, s->impl->setExponential(mode));
}

int moc_solver_allow_negative_fluxes(moc_solver_t* s, int allow) {
  MOC_SOLVER_CALL(s, s->impl->allowNegativeFluxes(allow != 0));
}

int moc_solver_compute_fsr_sources(moc_solver_t* s, int32_t iteration, double* sources, int64_t n) {
  MOC_SOLVER_CALL(s, s->impl->computeFSRSources(iteration, sources, n));
}

int moc_solver_set_scatter_precision(moc_solver_t* s, int mode, int* in_use) {
  MOC_SOLVER_CALL(s, {
    s->impl->setScatterPrecision(mode);
    if (in_use) *in_use = s->impl->scatterPrecisionInUse();
  });
}

int moc_solver_set_fixed_source_by_fsr(moc_solver_t* s, int64_t fsr, int32_t group,
                                       double source) {
  MOC_SOLVER_CALL(s, s->impl->setFixedSourceByFSR(fsr, group, source));
}

int moc_engine_create(int device, int32_t max_active, int64_t mem_budget_bytes,
                      moc_engine_t** out) {
  if (!out) return fail(MOC_ERR_INVALID, "null argument");
  *out = nullptr;
  return guarded([&] {
    auto* e = new moc_engine;
    e->impl = nullptr;
    try {
      e->impl = new moc::ServeEngine(device, max_active, mem_budget_bytes);
    } catch (...) {
      delete e;
      throw;
    }
    *out = e;
  });
}

void moc_engine_destroy(moc_engine_t* e) {
  if (!e) return;
  try {
    delete e->impl;
  } catch (...) {
  }
  delete e;
}

#define MOC_ENGINE_CALL(e, body)                                            \
  do {                                                                      \
    if (!(e) || !(e)->impl) return fail(MOC_ERR_INVALID, "null engine");    \
    return guarded([&] { body; });                                          \
  } while (0)

int moc_engine_add_geometry(moc_engine_t* e, const moc_track_data_t* tracks, int64_t num_fsrs,
                            const int32_t* fsr_material, const double* fsr_volume,
                            int32_t num_materials, int32_t num_groups, const double* sigma_t,
                            const double* sigma_s, const double* nu_sigma_f, const double* chi,
                            int32_t* geo_id) {
  if (!tracks || !geo_id) return fail(MOC_ERR_INVALID, "null argument");
  MOC_ENGINE_CALL(e, *geo_id = e->impl->addGeometry(*tracks, num_fsrs, fsr_material, fsr_volume,
                                                    num_materials, num_groups, sigma_t, sigma_s,
                                                    nu_sigma_f, chi));
}

int moc_engine_submit(moc_engine_t* e, int32_t geo_id, int32_t num_materials, int32_t num_groups,
                      const double* sigma_t, const double* sigma_s, const double* nu_sigma_f,
                      const double* chi, int32_t max_iters, int32_t res_type, double tolerance,
                      int want_flux, int64_t* request_id) {
  if (!request_id) return fail(MOC_ERR_INVALID, "null argument");
  MOC_ENGINE_CALL(e, *request_id = e->impl->submit(geo_id, num_materials, num_groups, sigma_t,
                                                   sigma_s, nu_sigma_f, chi, max_iters, res_type,
                                                   tolerance, want_flux != 0));
}

int moc_engine_wait(moc_engine_t* e, int64_t request_id, double timeout_s,
                    moc_request_result_t* result, double* flux, int64_t num_fluxes) {
  if (!e || !e->impl) return fail(MOC_ERR_INVALID, "null engine");
  bool done = false;
  moc::RequestResult r;
  const int rc = guarded([&] { done = e->impl->wait(request_id, timeout_s, &r, flux, num_fluxes); });
  if (rc != MOC_OK) return rc;
  if (!done) return fail(MOC_ERR_STATE, "request not finished within the timeout");
  if (result) {
    result->keff = r.keff;
    result->num_iterations = r.num_iterations;
    result->latency_s = r.latency_s;
    result->queue_s = r.queue_s;
    result->solve_s = r.solve_s;
  }
  return MOC_OK;
}

int moc_engine_stats(moc_engine_t* e, int64_t* completed, int64_t* failed, int64_t* iterations,
                     int32_t* pending) {
  MOC_ENGINE_CALL(e, {
    const moc::ServeStats s = e->impl->stats();
    if (completed) *completed = s.completed;
    if (failed) *failed = s.failed;
    if (iterations) *iterations = s.iterations;
    if (pending) *pending = s.pending;
  });
}

int moc_engine_memory(moc_engine_t* e, int64_t* budget, int64_t* held, int64_t* held_max) {
  MOC_ENGINE_CALL(e, {
    const moc::ServeStats s = e->impl->stats();
    if (budget) *budget = s.budget;
    if (held) *held = s.held;
    if (held_max) *held_max = s.held_max;
  });
}

int moc_device_copy_calibration(int device, int64_t bytes, int32_t repeats, double* seconds) {
  return guarded([&] {
    const double t = moc::deviceCopyCalibration(device, bytes, repeats);
    if (seconds) *seconds = t;
  });
}

int moc_solver_stabilize_transport(moc_solver_t* s, double factor, int32_t type) {
  MOC_SOLVER_CALL(s, s->impl->stabilizeTransport(factor, type));
}

int moc_solver_set_keff_from_neutron_balance(moc_solver_t* s, int on) {
  MOC_SOLVER_CALL(s, s->impl->setKeffFromNeutronBalance(on != 0));
}

int moc_solver_set_linear_source(moc_solver_t* s, int on) {
  MOC_SOLVER_CALL(s, s->impl->setLinearSource(on != 0));
}

int moc_solver_get_flux_moments(moc_solver_t* s, double* out, int64_t num_values) {
  MOC_SOLVER_CALL(s, s->impl->getFluxMoments(out, num_values));
}

int moc_solver_get_fsr_centroids(moc_solver_t* s, double* out, int64_t num_values) {
  MOC_SOLVER_CALL(s, s->impl->getCentroids(out, num_values));
}

int moc_solver_fission_transport_sweep(moc_solver_t* s) {
  MOC_SOLVER_CALL(s, s->impl->krylovSweep(1));
}

int moc_solver_scatter_transport_sweep(moc_solver_t* s) {
  MOC_SOLVER_CALL(s, s->impl->krylovSweep(2));
}

int moc_solver_compute_eigenvalue(moc_solver_t* s, int32_t max_iters, int32_t res_type) {
  MOC_SOLVER_CALL(s, s->impl->computeEigenvalue(max_iters, res_type));
}

int moc_solver_compute_flux(moc_solver_t* s, int32_t max_iters, int32_t only_fixed_source) {
  MOC_SOLVER_CALL(s, s->impl->computeFlux(max_iters, only_fixed_source != 0));
}

int moc_solver_compute_source(moc_solver_t* s, int32_t max_iters, double k_eff,
                              int32_t res_type) {
  MOC_SOLVER_CALL(s, s->impl->computeSource(max_iters, k_eff, res_type));
}

int moc_solver_get_keff(const moc_solver_t* s, double* keff) {
  if (!s || !s->impl || !keff) return fail(MOC_ERR_INVALID, "null argument");
  *keff = s->impl->getKeff();
  return MOC_OK;
}

int moc_solver_get_num_iterations(const moc_solver_t* s, int32_t* iters) {
  if (!s || !s->impl || !iters) return fail(MOC_ERR_INVALID, "null argument");
  *iters = s->impl->getNumIterations();
  return MOC_OK;
}

int moc_solver_get_flux(moc_solver_t* s, int64_t fsr, int32_t group, double* flux) {
  if (!flux) return fail(MOC_ERR_INVALID, "null argument");
  MOC_SOLVER_CALL(s, *flux = s->impl->getFlux(fsr, group));
}

int moc_solver_get_fluxes(moc_solver_t* s, double* out, int64_t num_fluxes) {
  MOC_SOLVER_CALL(s, s->impl->getFluxes(out, num_fluxes));
}

int moc_solver_set_fluxes(moc_solver_t* s, const double* in, int64_t num_fluxes) {
  MOC_SOLVER_CALL(s, s->impl->setFluxes(in, num_fluxes));
}

int moc_solver_compute_fsr_fission_rates(moc_solver_t* s, double* rates, int64_t num_fsrs) {
  MOC_SOLVER_CALL(s, s->impl->computeFSRFissionRates(rates, num_fsrs));
}

int moc_solver_set_fsr_centroids(moc_solver_t* s, const double* xyz, int64_t num_values) {
  MOC_SOLVER_CALL(s, s->impl->setFSRCentroids(xyz, num_values));
}

int moc_solver_dump_fsr_fluxes(moc_solver_t* s, const char* path) {
  if (!path || !*path) return fail(MOC_ERR_INVALID, "empty file name");
  MOC_SOLVER_CALL(s, s->impl->dumpFSRFluxes(path));
}

int moc_solver_load_fsr_fluxes(moc_solver_t* s, const char* path, int32_t assign_k_eff,
                               double tolerance, double* max_centroid_error) {
  if (!path || !*path) return fail(MOC_ERR_INVALID, "empty file name");
  MOC_SOLVER_CALL(s, {
    const double e = s->impl->loadFSRFluxes(path, assign_k_eff != 0, tolerance);
    if (max_centroid_error) *max_centroid_error = e;
  });
}

int moc_solver_load_initial_fsr_fluxes(moc_solver_t* s, const char* path) {
  MOC_SOLVER_CALL(s, s->impl->loadInitialFSRFluxes(path ? path : ""));
}

int moc_solver_set_residual_by_reference(moc_solver_t* s, const char* path) {
  MOC_SOLVER_CALL(s, s->impl->setResidualByReference(path ? path : ""));
}

int moc_solver_set_restart_status(moc_solver_t* s, int32_t is_restart) {
  MOC_SOLVER_CALL(s, s->impl->setRestartStatus(is_restart != 0));
}

int moc_solver_set_fixed_source_by_material(moc_solver_t* s, int32_t material, int32_t group,
                                            double source) {
  MOC_SOLVER_CALL(s, s->impl->setFixedSourceByMaterial(material, group, source));
}

int moc_solver_set_fsr_cells(moc_solver_t* s, const int32_t* cell_of_fsr, int64_t num_fsrs) {
  MOC_SOLVER_CALL(s, s->impl->setFSRCells(cell_of_fsr, num_fsrs));
}

int moc_solver_set_fixed_source_by_cell(moc_solver_t* s, int32_t cell, int32_t group,
                                        double source) {
  MOC_SOLVER_CALL(s, s->impl->setFixedSourceByCell(cell, group, source));
}

int moc_solver_reset_fixed_sources(moc_solver_t* s) {
  MOC_SOLVER_CALL(s, s->impl->resetFixedSources());
}

int moc_solver_set_chi_spectrum_material(moc_solver_t* s, int32_t material) {
  MOC_SOLVER_CALL(s, s->impl->setChiSpectrumMaterial(material));
}

int moc_solver_set_solver_mode(moc_solver_t* s, int32_t mode) {
  MOC_SOLVER_CALL(s, s->impl->setSolverMode(mode));
}

int moc_flux_file_write(const char* path, double keff, int32_t num_groups, int64_t num_fsrs,
                        const double* xyz, const double* flux) {
  if (!path || !*path) return fail(MOC_ERR_INVALID, "empty file name");
  return guarded([&] { moc::write_flux_file(path, keff, num_groups, num_fsrs, xyz, flux); });
}

int moc_flux_file_read(const char* path, double* keff, int32_t* num_groups, int64_t* num_fsrs,
                       double* xyz, int64_t xyz_len, double* flux, int64_t flux_len) {
  if (!path || !*path) return fail(MOC_ERR_INVALID, "empty file name");
  return guarded([&] {
    const moc::FluxFile f = moc::read_flux_file(path);
    if (xyz && xyz_len != (int64_t)f.xyz.size())
      throw moc::DeviceError(MOC_ERR_INVALID, "xyz_len must be 3*num_fsrs");
    if (flux && flux_len != (int64_t)f.flux.size())
      throw moc::DeviceError(MOC_ERR_INVALID, "flux_len must be num_fsrs*num_groups");
    if (keff) *keff = f.keff;
    if (num_groups) *num_groups = f.num_groups;
    if (num_fsrs) *num_fsrs = f.num_fsrs;
    if (xyz) copy_out(xyz, f.xyz);
    if (flux) copy_out(flux, f.flux);
  });
}

int moc_fsr_match_centroids(const double* file_xyz, int64_t num_file, const double* own_xyz,
                            int64_t num_own, double tolerance, int64_t* index, double* max_error) {
  if (!index && num_own > 0) return fail(MOC_ERR_INVALID, "null index array");
  return guarded([&] {
    const std::vector<int64_t> idx =
        moc::match_centroids(file_xyz, num_file, own_xyz, num_own, tolerance, max_error);
    copy_out(index, idx);
  });
}

int moc_solver_get_timing(const moc_solver_t* s, double* sweep_seconds, double* total_seconds,
                          int64_t* num_integrations) {
  if (!s || !s->impl) return fail(MOC_ERR_INVALID, "null solver");
  s->impl->timing(sweep_seconds, total_seconds, num_integrations);
  return MOC_OK;
}

int moc_solver_transport_sweep(moc_solver_t* s, const double* Q, float* psi, double* tally) {
  MOC_SOLVER_CALL(s, s->impl->transportSweepOnce(Q, psi, tally));
}

int moc_solver_transport_sweep_noalloc(moc_solver_t* s, const double* Q, const float* psi_in,
                                       float* psi_out, double* tally, void* stream) {
  MOC_SOLVER_CALL(s, s->impl->transportSweepNoAlloc(Q, psi_in, psi_out, tally,
                                                    static_cast<hipStream_t>(stream)));
}

int moc_solver_noalloc_scratch_bytes(const moc_solver_t* s, int64_t* bytes) {
  if (!s || !s->impl || !bytes) return fail(MOC_ERR_INVALID, "null argument");
  *bytes = (int64_t)s->impl->noAllocScratchBytes();
  return MOC_OK;
}

int moc_solver_set_launch_shape(moc_solver_t* s, const int32_t* shape11) {
  if (!shape11) return fail(MOC_ERR_INVALID, "null argument");
  MOC_SOLVER_CALL(s, {
    moc::LaunchShape sh;
    sh.groups_per_lane = shape11[0];
    sh.lanes_per_track = shape11[1];
    sh.tracks_per_wave = shape11[2];
    sh.block = shape11[3];
    sh.tally_copies = shape11[4];
    sh.lds_groups = shape11[5];
    sh.lds_block = shape11[6];
    sh.lds_blocks_per_sub = shape11[7];
    sh.lds_q = shape11[8];
    sh.lds_flush = shape11[9];
    sh.seg_store = shape11[10];
    s->impl->setShape(sh);
  });
}

int moc_solver_get_launch_shape(const moc_solver_t* s, int32_t* shape11) {
  if (!s || !s->impl || !shape11) return fail(MOC_ERR_INVALID, "null argument");
  const moc::LaunchShape& sh = s->impl->shape();
  shape11[0] = sh.groups_per_lane;
  shape11[1] = sh.lanes_per_track;
  shape11[2] = sh.tracks_per_wave;
  shape11[3] = sh.block;
  shape11[4] = sh.tally_copies;
  shape11[5] = sh.lds_groups;
  shape11[6] = sh.lds_block;
  shape11[7] = sh.lds_blocks_per_sub;
  shape11[8] = sh.lds_q;
  shape11[9] = sh.lds_flush;
  shape11[10] = sh.seg_store;
  return MOC_OK;
}

int moc_solver_run_iterations(moc_solver_t* s, int32_t num_iterations,
                              double* step_seconds_avg, double* sweep_seconds_avg) {
  if (!step_seconds_avg) return fail(MOC_ERR_INVALID, "null argument");
  MOC_SOLVER_CALL(s, *step_seconds_avg = s->impl->runIterations(num_iterations, sweep_seconds_avg));
}

int moc_solver_time_sweeps(moc_solver_t* s, int32_t repeats, double* seconds_per_sweep) {
  if (!seconds_per_sweep) return fail(MOC_ERR_INVALID, "null argument");
  MOC_SOLVER_CALL(s, *seconds_per_sweep = s->impl->timeSweeps(repeats));
}

}  // extern "C"
