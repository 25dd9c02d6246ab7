// Continuous-batching solve engine: many eigenvalue solves in flight on one
// GPU, each on its own stream with device-resident state, admitted as slots
// and device memory free up.
//
// A geometry (tracks, FSRs and the launch shape) is uploaded once and shared
// by every solve on it; a request brings only its cross sections.  Each
// admitted request is bound to a solver from the geometry's pool (re-armed
// in place, so serving makes no hipMalloc/hipFree), started with
// HipSolver::startSolve and advanced by pollSolve, whose iteration batches
// are captured HIP graphs.  Solves on different streams overlap on the
// device; a solve's arithmetic does not depend on what else runs (integer
// fixed-point tallies and fixed-order reductions), so its result is the same
// alone and in the mix.
#include "serve.h"

#include <algorithm>
#include <cstring>

namespace moc {

namespace {
double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

ServeEngine::ServeEngine(int device, int max_active, int64_t mem_budget)
    : device_(device), max_active_(std::max(1, max_active)) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    throw DeviceError(MOC_ERR_NO_DEVICE, "no HIP device visible");
  if (device_ < 0) (void)hipGetDevice(&device_);
  if (device_ >= count) throw DeviceError(MOC_ERR_INVALID, "device index out of range");
  if (hipSetDevice(device_) != hipSuccess) throw DeviceError(MOC_ERR_HIP, "hipSetDevice failed");
  if (mem_budget > 0) {
    budget_ = (size_t)mem_budget;
  } else {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
      throw DeviceError(MOC_ERR_HIP, "hipMemGetInfo failed");
    budget_ = free_b / 10 * 8;  // 80% of what is free now
  }
  worker_ = std::thread([this] { loop(); });
}

ServeEngine::~ServeEngine() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_all();
  if (worker_.joinable()) worker_.join();
}

int ServeEngine::addGeometry(const moc_track_data_t& tracks, int64_t num_fsrs,
                             const int32_t* fsr_material, const double* fsr_volume, int M, int G,
                             const double* sigma_t, const double* sigma_s, const double* nu_sigma_f,
                             const double* chi) {
  auto geo = std::make_unique<Geometry>();
  // the template solver owns the shared buffers; it is also a pooled solver
  geo->tmpl = std::make_unique<HipSolver>(device_);
  geo->tmpl->setTracks(tracks);
  geo->tmpl->setMaterials(M, G, sigma_t, sigma_s, nu_sigma_f, chi);
  geo->tmpl->setFSRs(num_fsrs, fsr_material, fsr_volume);
  geo->M = M;
  geo->G = G;
  geo->N = num_fsrs;
  geo->bytes = geo->tmpl->solveBytes();
  geo->work = 2.0 * (double)tracks.num_segments * G * (tracks.polar ? 1 : tracks.num_polar / 2);
  std::lock_guard<std::mutex> lk(geo_mu_);
  geos_.push_back(std::move(geo));
  return (int)geos_.size() - 1;
}

int64_t ServeEngine::submit(int geo, int M, int G, const double* sigma_t, const double* sigma_s,
                            const double* nu_sigma_f, const double* chi, int max_iters, int res_type,
                            double tol, bool want_flux) {
  Geometry* g = nullptr;
  {
    std::lock_guard<std::mutex> lk(geo_mu_);
    if (geo < 0 || geo >= (int)geos_.size()) throw DeviceError(MOC_ERR_INVALID, "unknown geometry id");
    g = geos_[geo].get();
  }
  if (M != g->M || G != g->G)
    throw DeviceError(MOC_ERR_INVALID, "materials must have the geometry's material and group counts");
  if (!sigma_t || !sigma_s || !nu_sigma_f || !chi) throw DeviceError(MOC_ERR_INVALID, "null cross sections");
  if (max_iters <= 0 || !(tol > 0.0) || res_type < MOC_SCALAR_FLUX || res_type > MOC_TOTAL_SOURCE)
    throw DeviceError(MOC_ERR_INVALID, "invalid solve options");
  auto r = std::make_shared<Request>();
  const size_t mg = (size_t)M * G;
  r->geo = geo;
  r->st.assign(sigma_t, sigma_t + mg);
  r->ss.assign(sigma_s, sigma_s + mg * G);
  r->nsf.assign(nu_sigma_f, nu_sigma_f + mg);
  r->chi.assign(chi, chi + mg);
  r->max_iters = max_iters;
  r->res_type = res_type;
  r->tol = tol;
  r->want_flux = want_flux;
  r->t_submit = now_s();
  std::lock_guard<std::mutex> lk(mu_);
  r->id = next_id_++;
  pending_.push_back(r);
  all_[r->id] = r;
  cv_work_.notify_all();
  return r->id;
}

bool ServeEngine::wait(int64_t id, double timeout_s, RequestResult* out, double* flux,
                       int64_t flux_n) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = all_.find(id);
  if (it == all_.end()) throw DeviceError(MOC_ERR_INVALID, "unknown request id");
  std::shared_ptr<Request> r = it->second;
  auto done = [&] { return r->state >= kDone; };
  if (timeout_s < 0) {
    cv_done_.wait(lk, done);
  } else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
    return false;
  }
  if (r->state == kFailed) {
    all_.erase(id);
    throw DeviceError(MOC_ERR_INVALID, "request failed: " + r->error);
  }
  if (out) {
    out->keff = r->keff;
    out->num_iterations = r->iters;
    out->latency_s = r->t_done - r->t_submit;
    out->queue_s = r->t_start - r->t_submit;
    out->solve_s = r->t_done - r->t_start;
  }
  if (flux) {
    if ((int64_t)r->flux.size() != flux_n)
      throw DeviceError(MOC_ERR_INVALID, "flux buffer size must be N*G (and want_flux set at submit)");
    std::memcpy(flux, r->flux.data(), r->flux.size() * sizeof(double));
  }
  all_.erase(id);
  return true;
}

ServeStats ServeEngine::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  ServeStats s = stats_;
  s.pending = (int)pending_.size();
  return s;
}

// Binds a request to a pooled solver of its geometry and starts it.
void ServeEngine::admit(const std::shared_ptr<Request>& r) {
  Geometry* g;
  {
    std::lock_guard<std::mutex> lk(geo_mu_);
    g = geos_[r->geo].get();
  }
  HipSolver* s = nullptr;
  if (!g->tmpl_busy) {
    s = g->tmpl.get();
    g->tmpl_busy = true;
  } else if (!g->idle.empty()) {
    s = g->idle.back();
    g->idle.pop_back();
  } else {
    g->pool.push_back(std::make_unique<HipSolver>(device_));
    s = g->pool.back().get();
    s->shareGeometry(*g->tmpl);
  }
  Active a{r, s, g};
  try {
    s->setMaterials(g->M, g->G, r->st.data(), r->ss.data(), r->nsf.data(), r->chi.data());
    s->setConvergenceThreshold(r->tol);
    r->t_start = now_s();
    s->startSolve(0, r->max_iters, r->res_type, 1.0);
  } catch (const std::exception& e) {
    release(a);
    std::lock_guard<std::mutex> lk(mu_);
    r->state = kFailed;
    r->error = e.what();
    stats_.failed++;
    cv_done_.notify_all();
    return;
  }
  mem_used_ += g->bytes;
  active_.push_back(a);
}

void ServeEngine::release(const Active& a) {
  if (a.solver == a.geo->tmpl.get())
    a.geo->tmpl_busy = false;
  else
    a.geo->idle.push_back(a.solver);
}

void ServeEngine::loop() {
  (void)hipSetDevice(device_);
  while (true) {
    // admission: free slots and memory.  Longest first: the pending request
    // with the most work per sweep starts next, so long solves overlap the
    // short ones instead of trailing them; a request that has been passed
    // over max_active_ times is admitted before anything younger.
    std::vector<std::shared_ptr<Request>> admit_now;
    {
      std::unique_lock<std::mutex> lk(mu_);
      if (active_.empty() && pending_.empty()) {
        cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
      }
      if (stop_ && pending_.empty() && active_.empty()) break;
      size_t mem = mem_used_;
      size_t slots = active_.size();
      while (!pending_.empty() && slots < (size_t)max_active_) {
        size_t pick = 0;
        size_t need;
        {
          std::lock_guard<std::mutex> gl(geo_mu_);
          if (pending_.front()->passed_over < max_active_) {
            double best = -1.0;
            for (size_t i = 0; i < pending_.size(); i++) {
              const double w = geos_[pending_[i]->geo]->work;
              if (w > best) {
                best = w;
                pick = i;
              }
            }
          }
          need = geos_[pending_[pick]->geo]->bytes;
        }
        if (slots > 0 && mem + need > budget_) break;  // wait for memory
        mem += need;
        slots++;
        for (size_t i = 0; i < pick; i++) pending_[i]->passed_over++;
        admit_now.push_back(pending_[pick]);
        pending_.erase(pending_.begin() + (long)pick);
      }
      if (stop_ && !admit_now.empty()) {
        // shutting down: fail what was not started
        for (auto& r : admit_now) {
          r->state = kFailed;
          r->error = "engine stopped";
        }
        admit_now.clear();
        cv_done_.notify_all();
      }
    }
    for (auto& r : admit_now) admit(r);
    // progress every active solve
    bool finished_any = false;
    for (size_t i = 0; i < active_.size();) {
      Active& a = active_[i];
      bool done = false;
      std::string err;
      try {
        done = a.solver->pollSolve();
      } catch (const std::exception& e) {
        done = true;
        err = e.what();
      }
      if (!done) {
        i++;
        continue;
      }
      std::shared_ptr<Request> r = a.req;
      if (err.empty()) {
        r->keff = a.solver->getKeff();
        r->iters = a.solver->getNumIterations();
        if (r->want_flux) {
          r->flux.resize((size_t)a.geo->N * a.geo->G);
          try {
            a.solver->getFluxes(r->flux.data(), (int64_t)r->flux.size());
          } catch (const std::exception& e) {
            err = e.what();
          }
        }
      }
      const double t = now_s();
      mem_used_ -= a.geo->bytes;
      release(a);
      active_.erase(active_.begin() + (long)i);
      {
        std::lock_guard<std::mutex> lk(mu_);
        r->t_done = t;
        if (err.empty()) {
          r->state = kDone;
          stats_.completed++;
          stats_.iterations += r->iters;
        } else {
          r->state = kFailed;
          r->error = err;
          stats_.failed++;
        }
      }
      finished_any = true;
    }
    if (finished_any) cv_done_.notify_all();
    if (!active_.empty() && !finished_any) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

}  // namespace moc
