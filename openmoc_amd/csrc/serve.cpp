// Continuous-batching solve engine: many eigenvalue solves in flight on one
// GPU with device-resident state, admitted as rows and device memory free up.
//
// A geometry (tracks, FSRs and the launch shape) is uploaded once and shared
// by every solve on it; a request brings only its cross sections.  Solves form
// SolveBatches (by residual type): the solves of all small geometries that run
// the same kernels share one batch, a large geometry has a batch of its own.
// Every launch of an iteration step carries all solves of a batch, one row
// each, in captured graphs of iteration steps, and a row whose solve stopped
// is refilled with the next request between graph launches.  Batches run on
// their own streams and overlap on the device.  A solve's arithmetic does not
// depend on what else runs (integer fixed-point tallies and fixed-order
// reductions on the 7-group path), so its result is the same alone and in the
// mix.
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
  // the template solver owns the shared buffers; the batches' row solvers
  // share them
  geo->tmpl = std::make_unique<HipSolver>(device_);
  geo->tmpl->setTracks(tracks);
  geo->tmpl->setMaterials(M, G, sigma_t, sigma_s, nu_sigma_f, chi);
  geo->tmpl->setFSRs(num_fsrs, fsr_material, fsr_volume);
  geo->M = M;
  geo->G = G;
  geo->N = num_fsrs;
  geo->bytes = geo->tmpl->solveBytes();
  geo->work = 2.0 * (double)tracks.num_segments * G * (tracks.polar ? 1 : tracks.num_polar / 2);
  const std::string key = SolveBatch::batchKey(*geo->tmpl);  // throws for what batches cannot run
  std::lock_guard<std::mutex> lk(geo_mu_);
  geo->batch_key = SolveBatch::shareable(*geo->tmpl) ? key : "geometry " + std::to_string(geos_.size());
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
  s.budget = (int64_t)budget_;
  s.held = held_.load();
  s.held_max = held_max_.load();
  return s;
}

void ServeEngine::noteHeld() {
  const int64_t h = (int64_t)bytesHeld();
  held_.store(h);
  if (h > held_max_.load()) held_max_.store(h);
}

size_t ServeEngine::bytesHeld() {
  size_t b = 0;
  for (auto& kv : batches_) b += kv.second->bytesHeld();
  return b;
}

// Starts r in the batch of its geometry's key and residual type.  A row solver that
// has to be created counts against the memory budget with every row solver
// already held (idle ones included); idle solvers not referenced by a launch
// in flight are freed first when that helps.  With no solve live anywhere a
// request is started whatever the budget, so the queue always moves.
bool ServeEngine::admit(const std::shared_ptr<Request>& r, int live) {
  Geometry* g;
  {
    std::lock_guard<std::mutex> lk(geo_mu_);
    g = geos_[r->geo].get();
  }
  const auto key = std::make_pair(g->batch_key, r->res_type);
  std::unique_ptr<SolveBatch>& b = batches_[key];
  try {
    if (!b) b = std::make_unique<SolveBatch>(device_, max_active_, r->res_type);
    if (!b->hasGeometry(g->tmpl.get())) b->addGeometry(*g->tmpl);
  } catch (const std::exception& e) {
    if (b && !b->busy() && b->bytesHeld() == 0) batches_.erase(key);
    finishRequest(r, nullptr, e.what());
    return true;
  }
  if (!b->hasIdleSolver(g->tmpl.get())) {
    size_t held = bytesHeld();
    if (held + g->bytes > budget_) {
      for (auto& kv : batches_) {
        // with nothing live, the launches still in flight only finish
        // stopped rows: let them end so every idle solver can go
        if (live == 0) kv.second->drain();
        if (held + g->bytes > budget_) held -= kv.second->releaseIdle();
      }
      held_.store((int64_t)held);
    }
    if (held + g->bytes > budget_ && live > 0) return false;
  }
  r->t_start = now_s();
  try {
    b->start(g->tmpl.get(), r.get(), g->M, g->G, r->st.data(), r->ss.data(), r->nsf.data(),
             r->chi.data(), r->max_iters, r->tol, r->want_flux);
  } catch (const std::exception& e) {
    finishRequest(r, nullptr, e.what());
    return true;
  }
  running_[r.get()] = r;
  noteHeld();
  return true;
}

void ServeEngine::finishRequest(const std::shared_ptr<Request>& r, const BatchResult* res,
                                const std::string& err) {
  const double t = now_s();
  std::lock_guard<std::mutex> lk(mu_);
  r->t_done = t;
  if (res && err.empty()) {
    r->keff = res->keff;
    r->iters = res->num_iterations;
    r->flux = res->flux;
    r->state = kDone;
    stats_.completed++;
    stats_.iterations += r->iters;
  } else {
    r->state = kFailed;
    r->error = err;
    stats_.failed++;
  }
}

void ServeEngine::loop() {
  (void)hipSetDevice(device_);
  std::vector<BatchResult> done;
  while (true) {
    // admission: free rows and memory.  Longest first: the pending request
    // with the most work per sweep starts next, so long solves overlap the
    // short ones instead of trailing them; a request that has been passed
    // over max_active_ times is admitted before anything younger.
    std::vector<std::shared_ptr<Request>> admit_now;
    int live = (int)running_.size();
    {
      std::unique_lock<std::mutex> lk(mu_);
      if (running_.empty() && pending_.empty()) {
        cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
      }
      if (stop_ && pending_.empty() && running_.empty()) break;
      int slots = live;
      while (!pending_.empty() && slots < max_active_) {
        size_t pick = 0;
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
        }
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
    bool changed = false;
    for (size_t i = 0; i < admit_now.size(); i++) {
      if (!admit(admit_now[i], live)) {
        // short of memory: the rest go back to the front of the queue, in order
        std::lock_guard<std::mutex> lk(mu_);
        pending_.insert(pending_.begin(), admit_now.begin() + (long)i, admit_now.end());
        break;
      }
      live = (int)running_.size();
      changed = true;
    }
    // progress every batch: collect stopped solves, keep launches in flight
    done.clear();
    for (auto& kv : batches_) {
      try {
        kv.second->progress(done);
      } catch (const std::exception& e) {
        // a device error: every solve of the engine is lost
        for (auto& kv : running_) finishRequest(kv.second, nullptr, e.what());
        running_.clear();
        changed = true;
        break;
      }
    }
    for (const BatchResult& res : done) {
      auto it = running_.find(res.tag);
      if (it == running_.end()) continue;
      finishRequest(it->second, &res, "");
      running_.erase(it);
      changed = true;
    }
    if (changed) cv_done_.notify_all();
    if (!running_.empty() && !changed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

}  // namespace moc
