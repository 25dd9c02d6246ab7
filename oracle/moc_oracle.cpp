/*
 * moc_oracle.cpp -- FP64 CPU restatement of the reference OpenMOC source
 * iteration.  TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, as the checker.
 *
 * What each function restates, and how it is pinned:
 *
 *   oracle_exp_f1                 src/exponentials.h:156-192
 *       pinned: closed form (1-e^-x)/x to the fit's stated 2e-6 over
 *       [0, 100], plus the reference unit-test golden values
 *       (tests/unit_tests/test_exponentials.py:55-106).
 *   oracle_pairwise_sum           src/pairwise_sum.h:19-39
 *       pinned: exact sums of integers / brute force.
 *   oracle_compute_fsr_sources    src/CPUSolver.cpp:1939-2023
 *   oracle_add_source_to_scalar_flux  src/CPUSolver.cpp:2608-2659
 *   oracle_transport_sweep        src/CPUSolver.cpp:2338-2389, 2402-2497,
 *                                 2507-2527, 2560-2601;
 *                                 src/TrackTraversingAlgorithms.cpp:890-1052
 *       pinned: single-segment analytic attenuation, particle balance, and
 *       bit-for-bit-class agreement with the compiled reference CPUSolver on
 *       the reference's own tracks (keff / flux fixtures).
 *   oracle_compute_eigenvalue     src/Solver.cpp:1542-1689 with
 *                                 CPUSolver::computeKeff :2258-2328,
 *                                 normalizeFluxes :1860-1931,
 *                                 computeResidual :2113-2252
 *       pinned: analytic infinite-medium k (1- and 2-group), the reference
 *       regression fixtures (tests/test_forward_pin_cell/results_true.dat,
 *       tests/test_forward_hom_inf_medium/results_true.dat) and the compiled
 *       reference on the synthetic workloads.
 *   oracle_compute_flux           src/Solver.cpp:1352-1420
 *       pinned: infinite-medium fixed-source closed form phi = 4pi q / sigma_a
 *       (q isotropic source per steradian, Q = fixed/4pi).
 *   neutron-balance keff          src/CPUSolver.cpp:2258-2328 (rates of
 *                                 fission, absorption and the vacuum leakage
 *                                 tallied by transferBoundaryFlux :2593-2600)
 *       pinned: the compiled reference with setKeffFromNeutronBalance, and
 *       the reflective infinite medium where it equals the fission-rate k.
 *   transport stabilisation       src/Solver.cpp:1616-1635,
 *                                 src/CPUSolver.cpp:2665-2805
 *       pinned: the compiled reference with stabilizeTransport (each type),
 *       and the fixed point: a converged flux is unchanged by it.
 *   oracle_krylov_sweep           src/Solver.cpp:1282-1297,
 *                                 src/CPUSolver.cpp:2030-2104
 *       pinned: the compiled reference's fissionTransportSweep /
 *       scatterTransportSweep outputs, and linearity (the two operator
 *       sweeps from zero start fluxes add up to one sweep on the sum of
 *       their sources).
 */
#include "moc_oracle.h"

#include <omp.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr double kFourPi = 12.566370614359172;
constexpr double kOneOverFourPi = 0.07957747154594767;
constexpr double kFluxEpsilon = 1.0e-25;
constexpr double kFltEpsilon = 1.0e-12;

template <typename T>
double pairwise(const T* v, int64_t n) {
  double sum = 0;
  if (n < 4) {
    for (int64_t i = 0; i < n; i++) sum += v[i];
  } else {
    int64_t off = n % 2;
    int64_t half = n / 2;
    sum = pairwise(v, half) + pairwise(v + half, half + off);
  }
  return sum;
}

inline double exp_f1(double x) {
  // Josey 5/6 rational fit of (1 - exp(-x)) / x (coefficients published with
  // the reference, src/exponentials.h:167-181).
  const double p1 = 2.4172687328033081e-1, p2 = 6.2804790965268531e-2,
               p3 = 1.0567595009016521e-2, p4 = 1.0059468082903561e-3,
               p5 = 1.9309063097411041e-4;
  const double d1 = 7.4169266112320541e-1, d2 = 2.6722515319494311e-1,
               d3 = 6.1643725066901411e-2, d4 = 1.0590759992367811e-2,
               d5 = 1.0057980007137651e-3, d6 = 1.9309063097411041e-4;
  double den = ((((((d6 * x + d5) * x + d4) * x + d3) * x + d2) * x + d1) * x + 1.0);
  double num = (((((p5 * x + p4) * x + p3) * x + p2) * x + p1) * x + 1.0);
  return num / den;
}

struct View {
  const OracleProblem* p;
  int G, P, fpt;
  int64_t T, N;
  const double* sig_t(int64_t r) const { return p->sigma_t + (int64_t)p->fsr_material[r] * G; }
  const double* sig_s(int64_t r) const { return p->sigma_s + (int64_t)p->fsr_material[r] * G * G; }
  const double* nsf(int64_t r) const { return p->nu_sigma_f + (int64_t)p->fsr_material[r] * G; }
  const double* chi(int64_t r) const { return p->chi + (int64_t)p->fsr_material[r] * G; }
  bool fissionable(int64_t r) const {
    const double* v = nsf(r);
    for (int g = 0; g < G; g++)
      if (v[g] != 0.0) return true;
    return false;
  }
};

View make_view(const OracleProblem* p) {
  View v;
  v.p = p;
  v.G = p->num_groups;
  v.P = p->num_polar_half;
  v.fpt = v.G * v.P;
  v.T = p->num_tracks;
  v.N = p->num_fsrs;
  return v;
}

// Q[r,G] = (sum_g s[G,g] phi[g] + sum_g chi[G] nsf[g] phi[g] / k + fixed) / 4pi
void fsr_sources(const View& v, const double* phi, double k, int it, bool allow_neg,
                 double* Q, int mode /*0 total, 1 fission only, 2 scatter only*/) {
  const int G = v.G;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < v.N; r++) {
    const double* s = v.sig_s(r);
    const double* chi = v.chi(r);
    const double* nsf = v.nsf(r);
    bool fis = v.fissionable(r);
    std::vector<double> sc(G), fi(G);
    for (int Gd = 0; Gd < G; Gd++) {
      for (int g = 0; g < G; g++) {
        sc[g] = s[Gd * G + g] * phi[r * G + g];
        fi[g] = phi[r * G + g] * (fis ? chi[Gd] * nsf[g] : 0.0);
      }
      double scatter = pairwise(sc.data(), G);
      double fission = pairwise(fi.data(), G);
      double q;
      if (mode == 1) q = fission;
      else if (mode == 2) q = scatter;
      else {
        q = fission / k + scatter;
        if (v.p->fixed_source) q += v.p->fixed_source[r * G + Gd];
      }
      q *= kOneOverFourPi;
      if (mode == 0 && q < 0.0 && it < 30 && !allow_neg) q = kFluxEpsilon;
      Q[r * G + Gd] = q;
    }
  }
}

template <typename PsiT>
void sweep(const View& v, const double* Q, PsiT* start, double* tally, int threads,
           float* leakage = nullptr /* [T] per track, or null */) {
  const OracleProblem* p = v.p;
  const int G = v.G, P = v.P, fpt = v.fpt;
  const int64_t T = v.T;
  std::memset(tally, 0, sizeof(double) * v.N * G);
  // boundary <- start (CPUSolver::copyBoundaryFluxes, :483-490)
  std::vector<PsiT> boundary(start, start + 2 * T * fpt);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    std::vector<double> buf(G, 0.0), dpsi(fpt);
#pragma omp for schedule(dynamic)
    for (int64_t t = 0; t < T; t++) {
      const int a = p->trk_azim[t];
      const double* w = p->total_weight + (int64_t)a * P;
      const int64_t s0 = p->seg_offset[t], s1 = p->seg_offset[t + 1];
      const int64_t ns = s1 - s0;
      for (int d = 0; d < 2; d++) {
        PsiT* psi = &boundary[(t * 2 + d) * fpt];
        for (int64_t kk = 0; kk < ns; kk++) {
          int64_t k = d == 0 ? s0 + kk : s1 - 1 - kk;
          const int32_t r = p->seg_fsr[k];
          const double len = p->seg_length[k];
          const double* st = v.sig_t(r);
          // tallyScalarFlux (2D, :2457-2496)
          for (int pp = 0; pp < P; pp++) {
            const double isin = p->inv_sin_theta[pp];
            for (int g = 0; g < G; g++) {
              double tau = st[g] * len;
              double e = isin * exp_f1(tau * isin);
              double dl = (tau * (double)psi[pp * G + g] - len * Q[r * G + g]) * e;
              psi[pp * G + g] = (PsiT)((double)psi[pp * G + g] - dl);
              dpsi[pp * G + g] = dl * w[pp];
            }
          }
          for (int pp = 0; pp < P; pp++)
            for (int g = 0; g < G; g++) buf[g] += dpsi[pp * G + g];
          // flush when the next segment (in this direction) changes FSR
          bool last = kk == ns - 1;
          int64_t knext = d == 0 ? k + 1 : k - 1;
          bool flush = (d == 0) ? (!last && p->seg_fsr[knext] != r) : (last || p->seg_fsr[knext] != r);
          if (flush) {
            for (int g = 0; g < G; g++) {
#pragma omp atomic
              tally[r * G + g] += buf[g];
              buf[g] = 0.0;
            }
          }
        }
        // transferBoundaryFlux (:2560-2601)
        uint8_t bc = d == 0 ? p->bc_fwd[t] : p->bc_bwd[t];
        if (bc == 1 || bc == 2) {
          int64_t nt = d == 0 ? p->next_fwd[t] : p->next_bwd[t];
          int nd = d == 0 ? !p->next_fwd_is_fwd[t] : !p->next_bwd_is_fwd[t];
          PsiT* out = start + (nt * 2 + nd) * fpt;
          for (int e = 0; e < fpt; e++) out[e] = psi[e];
        } else if (leakage) {
          // vacuum leakage with the weight of polar index 0 (the 2D sweep
          // passes polar_index = 0, src/TrackTraversingAlgorithms.cpp:902)
          const double w0 = w[0];
          for (int e = 0; e < fpt; e++) leakage[t] += (float)(w0 * (double)psi[e]);
        }
      }
    }
  }
}

void add_source(const View& v, const double* Q, bool allow_neg, double* phi) {
  const int G = v.G;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < v.N; r++) {
    double vol = v.p->fsr_volume[r];
    if (vol < kFltEpsilon) vol = 1e30;
    const double* st = v.sig_t(r);
    for (int g = 0; g < G; g++) {
      double f = phi[r * G + g] / (st[g] * vol);
      f += kFourPi * Q[r * G + g] / st[g];
      if (f < 0.0 && !allow_neg) f = kFluxEpsilon;
      phi[r * G + g] = f;
    }
  }
}

double fission_rate_total(const View& v, const double* phi, std::vector<double>& scratch) {
  const int G = v.G;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < v.N; r++) {
    std::vector<double> gr(G);
    const double* nsf = v.nsf(r);
    for (int g = 0; g < G; g++) gr[g] = nsf[g] * phi[r * G + g] * v.p->fsr_volume[r];
    scratch[r] = pairwise(gr.data(), G);
  }
  return pairwise(scratch.data(), v.N);
}

double residual(const View& v, const double* phi, const double* old, int res_type, double k,
                std::vector<double>& res) {
  const int G = v.G;
  int64_t norm = 0;
  std::fill(res.begin(), res.end(), 0.0);
  if (res_type == ORACLE_SCALAR_FLUX) {
    norm = v.N;
    for (int64_t r = 0; r < v.N; r++)
      for (int g = 0; g < G; g++)
        if (old[r * G + g] > 0.0) {
          double d = (phi[r * G + g] - old[r * G + g]) / old[r * G + g];
          res[r] += d * d;
        }
  } else if (res_type == ORACLE_FISSION_SOURCE) {
    for (int64_t r = 0; r < v.N; r++) {
      if (!v.fissionable(r)) continue;
      norm++;
      const double* nsf = v.nsf(r);
      double nw = 0, od = 0;
      for (int g = 0; g < G; g++) {
        nw += phi[r * G + g] * nsf[g];
        od += old[r * G + g] * nsf[g];
      }
      if (od > 0.0) res[r] = ((nw - od) / od) * ((nw - od) / od);
    }
  } else {
    norm = v.N;
    for (int64_t r = 0; r < v.N; r++) {
      double nw = 0, od = 0;
      if (v.fissionable(r)) {
        const double* nsf = v.nsf(r);
        for (int g = 0; g < G; g++) {
          nw += phi[r * G + g] * nsf[g];
          od += old[r * G + g] * nsf[g];
        }
        nw /= k;
        od /= k;
      }
      const double* s = v.sig_s(r);
      for (int Gd = 0; Gd < G; Gd++)
        for (int g = 0; g < G; g++) {
          nw += s[Gd * G + g] * phi[r * G + g];
          od += s[Gd * G + g] * old[r * G + g];
        }
      if (od > 0.0) res[r] = ((nw - od) / od) * ((nw - od) / od);
    }
  }
  double sum = pairwise(res.data(), v.N);
  if (sum < 0) sum = 0;
  if (norm <= 0) norm = 1;
  return std::sqrt(sum / (double)norm);
}

enum Mode { MODE_EIGEN = 0, MODE_FLUX = 1, MODE_SOURCE = 2 };

// Solver::computeEigenvalue (:1542-1689), computeFlux (:1352-1420) and
// computeSource (:1459-1516) share this loop.
template <typename PsiT>
int solve(const OracleProblem* p, const OracleOptions* o, OracleResult* out, double* flux_out,
          Mode mode, double k_fixed) {
  View v = make_view(p);
  const int G = v.G;
  const int64_t N = v.N, T = v.T;
  if (o->threads > 0) omp_set_num_threads(o->threads);
  std::vector<double> phi(N * G, mode == MODE_FLUX ? 0.0 : 1.0), old(N * G), Q(N * G),
      scratch(N), tally(N * G);
  std::vector<PsiT> start(2 * T * v.fpt, (PsiT)0);
  bool allow_neg = o->allow_negative_fluxes != 0;
  double k = mode == MODE_SOURCE ? k_fixed : 1.0;
  double sweep_s = 0;
  int iters = 0;
  double res = 0;

  auto normalize = [&]() {
    double tot = fission_rate_total(v, phi.data(), scratch);
    double f = (double)N / tot;
    for (auto& x : phi) x *= f;
    for (auto& x : start) x = (PsiT)(x * f);
  };
  if (mode == MODE_EIGEN) normalize();  // computeInitialFluxGuess (:1710-1731)
  old = phi;
  if (mode == MODE_FLUX) fsr_sources(v, phi.data(), k, 0, allow_neg, Q.data(), 0);

  double k_prev = k;
  bool converged = false;
  const bool balance = mode == MODE_EIGEN && o->keff_from_fission_rates == 0;
  const bool stab = mode == MODE_EIGEN && o->stab_type >= 0;
  std::vector<float> leakage(balance ? T : 0);
  std::vector<double> stab_flux(stab ? N * G : 0), stab_ratio(stab ? G : 0);
  if (stab && o->stab_type == 1) {
    // YAMAMOTO: largest |sigma_s(g->g) / sigma_t(g)| over the FSRs, times the factor
    for (int g = 0; g < G; g++) {
      double mx = 0.0;
      for (int64_t r = 0; r < N; r++) {
        double ratio = std::abs(v.sig_s(r)[g * G + g] / v.sig_t(r)[g]);
        if (ratio > mx) mx = ratio;
      }
      stab_ratio[g] = mx * o->stab_factor;
    }
  }
  for (int i = 0; i < o->max_iters; i++) {
    if (stab && i > 0) {
      // computeStabilizingFlux (src/CPUSolver.cpp:2665-2723) from the
      // normalised flux of the previous iteration
      for (int64_t r = 0; r < N; r++)
        for (int g = 0; g < G; g++) {
          const int64_t q = r * G + g;
          if (o->stab_type == 0) {
            const double ss = v.sig_s(r)[g * G + g];
            if (ss < 0.0) stab_flux[q] = -phi[q] * o->stab_factor * ss / v.sig_t(r)[g];
          } else if (o->stab_type == 1) {
            stab_flux[q] = phi[q] * stab_ratio[g];
          } else {
            stab_flux[q] = (1.0 / o->stab_factor - 1.0) * phi[q];
          }
        }
    }
    if (mode != MODE_FLUX) fsr_sources(v, phi.data(), k, i, allow_neg, Q.data(), 0);
    if (balance) std::fill(leakage.begin(), leakage.end(), 0.f);
    auto t0 = std::chrono::steady_clock::now();
    sweep<PsiT>(v, Q.data(), start.data(), tally.data(), o->threads,
                balance ? leakage.data() : nullptr);
    sweep_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    phi = tally;
    add_source(v, Q.data(), allow_neg, phi.data());
    if (mode == MODE_EIGEN) {
      double rate = fission_rate_total(v, phi.data(), scratch);
      if (balance) {
        // k = fission / (absorption + leakage) (:2264-2325), sigma_a from
        // Material::getSigmaA (src/Material.cpp:228-257)
        for (int64_t r = 0; r < N; r++) {
          std::vector<double> gr(G);
          const double* st = v.sig_t(r);
          const double* ss = v.sig_s(r);
          for (int g = 0; g < G; g++) {
            double sa = st[g];
            for (int d = 0; d < G; d++) sa -= ss[d * G + g];
            gr[g] = sa * phi[r * G + g];
          }
          scratch[r] = pairwise(gr.data(), G) * p->fsr_volume[r];
        }
        double absorption = pairwise(scratch.data(), N);
        double leak = pairwise(leakage.data(), T);
        k = rate / (absorption + leak);
      } else {
        k *= rate / (double)N;  // computeKeff with fission rates (:2258-2328)
      }
      if (stab && i > 0) {
        // stabilizeFlux (src/CPUSolver.cpp:2730-2805)
        for (int64_t r = 0; r < N; r++)
          for (int g = 0; g < G; g++) {
            const int64_t q = r * G + g;
            if (o->stab_type == 0) {
              const double ss = v.sig_s(r)[g * G + g];
              if (ss < 0.0) {
                phi[q] += stab_flux[q];
                phi[q] /= (1.0 - o->stab_factor * ss / v.sig_t(r)[g]);
              }
            } else if (o->stab_type == 1) {
              phi[q] += stab_flux[q];
              phi[q] /= (1 + stab_ratio[g]);
            } else {
              phi[q] += stab_flux[q];
              phi[q] *= o->stab_factor;
            }
          }
      }
      normalize();
      res = residual(v, phi.data(), old.data(), o->res_type, k, scratch);
      int dk = (int)(1e5 * (k - k_prev));
      k_prev = k;
      old = phi;
      iters++;
      if (res < o->tolerance && std::abs(dk) < 1) { converged = true; break; }
    } else {
      int rt = mode == MODE_FLUX ? ORACLE_SCALAR_FLUX : o->res_type;
      res = residual(v, phi.data(), old.data(), rt, k, scratch);
      old = phi;
      if (i > 1 && res < o->tolerance) { iters = i; converged = true; break; }
      iters = i + 1;
    }
  }
  if (!converged && mode != MODE_EIGEN) iters = o->max_iters;
  std::memcpy(flux_out, phi.data(), sizeof(double) * N * G);
  out->keff = k;
  out->num_iterations = iters;
  out->residual = res;
  out->sweep_seconds = sweep_s;
  out->num_integrations = 2LL * G * v.P * p->seg_offset[T] * iters;
  return 0;
}

}  // namespace

extern "C" {

double oracle_exp_f1(double x) { return exp_f1(x); }

double oracle_pairwise_sum(const double* v, int64_t n) { return pairwise(v, n); }

int oracle_compute_fsr_sources(const OracleProblem* p, const double* phi, double keff,
                               int iteration, int allow_negative, double* Q_out) {
  View v = make_view(p);
  fsr_sources(v, phi, keff, iteration, allow_negative != 0, Q_out, 0);
  return 0;
}

int oracle_add_source_to_scalar_flux(const OracleProblem* p, const double* Q,
                                     int allow_negative, double* phi_inout) {
  View v = make_view(p);
  add_source(v, Q, allow_negative != 0, phi_inout);
  return 0;
}

int oracle_transport_sweep(const OracleProblem* p, const double* Q, float* psi_start,
                           double* tally_out, int threads) {
  View v = make_view(p);
  sweep<float>(v, Q, psi_start, tally_out, threads);
  return 0;
}

int oracle_compute_eigenvalue(const OracleProblem* p, const OracleOptions* o,
                              OracleResult* r, double* flux_out) {
  if (!p || !o || !r || !flux_out) return -1;
  return o->psi_fp64 ? solve<double>(p, o, r, flux_out, MODE_EIGEN, 1.0)
                     : solve<float>(p, o, r, flux_out, MODE_EIGEN, 1.0);
}

int oracle_krylov_sweep(const OracleProblem* p, int kind, int allow_negative, double* phi,
                        float* psi_start, int threads) {
  if (!p || !phi || !psi_start || (kind != 1 && kind != 2)) return -1;
  View v = make_view(p);
  std::vector<double> Q(v.N * v.G), tally(v.N * v.G);
  fsr_sources(v, phi, 1.0, 0, true, Q.data(), kind);
  sweep<float>(v, Q.data(), psi_start, tally.data(), threads);
  std::memcpy(phi, tally.data(), sizeof(double) * v.N * v.G);
  add_source(v, Q.data(), allow_negative != 0, phi);
  return 0;
}

int oracle_compute_flux(const OracleProblem* p, const OracleOptions* o, OracleResult* r,
                        double* flux_out) {
  if (!p || !o || !r || !flux_out) return -1;
  return o->psi_fp64 ? solve<double>(p, o, r, flux_out, MODE_FLUX, 1.0)
                     : solve<float>(p, o, r, flux_out, MODE_FLUX, 1.0);
}

int oracle_compute_source(const OracleProblem* p, const OracleOptions* o, double k_eff,
                          OracleResult* r, double* flux_out) {
  if (!p || !o || !r || !flux_out || !(k_eff > 0)) return -1;
  return o->psi_fp64 ? solve<double>(p, o, r, flux_out, MODE_SOURCE, k_eff)
                     : solve<float>(p, o, r, flux_out, MODE_SOURCE, k_eff);
}

}  // extern "C"
