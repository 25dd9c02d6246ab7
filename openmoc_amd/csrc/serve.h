// Continuous-batching solve engine (see serve.cpp).
#ifndef OPENMOC_AMD_SERVE_H_
#define OPENMOC_AMD_SERVE_H_

#include <atomic>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hip_solver.h"

namespace moc {

struct RequestResult {
  double keff = 0.0;
  int num_iterations = 0;
  double latency_s = 0.0;  // submit -> result
  double queue_s = 0.0;    // submit -> admitted
  double solve_s = 0.0;    // admitted -> result
};

struct ServeStats {
  int64_t completed = 0;
  int64_t failed = 0;
  int64_t iterations = 0;
  int pending = 0;
  int64_t budget = 0;     // device bytes for the solves' own arrays
  int64_t held = 0;       // bytes of the row solvers held now (live or idle)
  int64_t held_max = 0;   // the largest `held` so far
};

class ServeEngine {
 public:
  // max_active: solves in flight at once; mem_budget: device bytes for the
  // solves' own arrays (<= 0: 80% of the free memory at creation).
  ServeEngine(int device, int max_active, int64_t mem_budget);
  ~ServeEngine();
  ServeEngine(const ServeEngine&) = delete;
  ServeEngine& operator=(const ServeEngine&) = delete;

  // Uploads a geometry once (tracks, FSRs; the materials given fix the group
  // and material counts and the launch shape).  Returns its id.
  int addGeometry(const moc_track_data_t& tracks, int64_t num_fsrs, const int32_t* fsr_material,
                  const double* fsr_volume, int M, int G, const double* sigma_t,
                  const double* sigma_s, const double* nu_sigma_f, const double* chi);
  // Queues an eigenvalue solve of geometry `geo` with these cross sections.
  int64_t submit(int geo, int M, int G, const double* sigma_t, const double* sigma_s,
                 const double* nu_sigma_f, const double* chi, int max_iters, int res_type,
                 double tol, bool want_flux);
  // Waits for a request (timeout_s < 0: forever); false on timeout.  The
  // request is forgotten once its result has been returned.
  bool wait(int64_t id, double timeout_s, RequestResult* out, double* flux, int64_t flux_n);
  ServeStats stats();

 private:
  enum { kPending = 0, kDone = 1, kFailed = 2 };
  struct Request {
    int64_t id = 0;
    int geo = 0;
    std::vector<double> st, ss, nsf, chi;
    int max_iters = 0, res_type = 0;
    double tol = 0.0;
    bool want_flux = false;
    int state = kPending;
    std::string error;
    double keff = 0.0;
    int iters = 0;
    std::vector<double> flux;
    double t_submit = 0, t_start = 0, t_done = 0;
    int passed_over = 0;  // admissions that went to a later request
  };
  struct Geometry {
    std::unique_ptr<HipSolver> tmpl;  // owns the shared geometry buffers
    // batch of its solves (with the residual type): small geometries that
    // run the same kernels share one, a large geometry has its own
    std::string batch_key;
    int M = 0, G = 0;
    int64_t N = 0;
    size_t bytes = 0;                 // device bytes of one solve
    double work = 0.0;                // integrations per sweep
  };
  void loop();
  // Starts r in its geometry's batch; false when device memory is short.
  bool admit(const std::shared_ptr<Request>& r, int live);
  size_t bytesHeld();                 // all row solvers of all batches
  void finishRequest(const std::shared_ptr<Request>& r, const BatchResult* res,
                     const std::string& err);

  int device_;
  int max_active_;
  size_t budget_ = 0;
  std::mutex geo_mu_;
  std::vector<std::unique_ptr<Geometry>> geos_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<std::shared_ptr<Request>> pending_;
  std::map<int64_t, std::shared_ptr<Request>> all_;
  std::map<void*, std::shared_ptr<Request>> running_;  // worker thread only: by batch tag
  // worker thread only: batches by (batch key, residual type)
  std::map<std::pair<std::string, int>, std::unique_ptr<SolveBatch>> batches_;
  int64_t next_id_ = 1;
  bool stop_ = false;
  ServeStats stats_;
  std::atomic<int64_t> held_{0}, held_max_{0};  // written by the worker thread
  void noteHeld();
  std::thread worker_;
};

}  // namespace moc

#endif
