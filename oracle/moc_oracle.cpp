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
 *   oracle_compute_source         src/Solver.cpp:1459-1516
 *   oracle_solve_state            the same loops from a previous solve's
 *                                 scalar and start fluxes (computeFlux with
 *                                 only_fixed_source = false, :1379-1386)
 *       pinned: infinite-medium fixed-source closed form phi = 4pi q / sigma_a
 *       (q isotropic source per steradian, Q = fixed/4pi); the compiled
 *       reference's computeFlux / computeSource with fixed sources by cell,
 *       FSR and material, including its own tests/test_compute_flux and
 *       tests/test_compute_source regression files, and the continuation
 *       of an eigenvalue and a source solve.
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
  v.P = p->trk_polar ? 1 : p->num_polar_half;
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

// 3D tracks (TransportSweep::onTrack with a Track3D, src/TrackTraversingAlgorithms.cpp:
// 890-1052; CPUSolver::tallyScalarFlux 3D branch, src/CPUSolver.cpp:2411-2453):
// G fluxes per track end, delta = (tau psi - l Q) F1(tau) with the segment's 3D
// length, accumulated unweighted per FSR run and weighted by the track's
// w(a, p) when flushed (accumulateScalarFluxContribution, :2507-2527).
template <typename PsiT>
void sweep3d(const View& v, const double* Q, PsiT* start, double* tally, int threads) {
  const OracleProblem* p = v.p;
  const int G = v.G;
  const int64_t T = v.T;
  std::memset(tally, 0, sizeof(double) * v.N * G);
  std::vector<PsiT> boundary(start, start + 2 * T * G);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    std::vector<double> buf(G, 0.0);
#pragma omp for schedule(dynamic)
    for (int64_t t = 0; t < T; t++) {
      const double w = p->total_weight[(int64_t)p->trk_azim[t] * p->num_polar_3d + p->trk_polar[t]];
      const int64_t s0 = p->seg_offset[t], s1 = p->seg_offset[t + 1];
      const int64_t ns = s1 - s0;
      for (int d = 0; d < 2; d++) {
        PsiT* psi = &boundary[(t * 2 + d) * G];
        for (int64_t kk = 0; kk < ns; kk++) {
          const int64_t k = d == 0 ? s0 + kk : s1 - 1 - kk;
          const int32_t r = p->seg_fsr[k];
          const double len = p->seg_length[k];
          const double* st = v.sig_t(r);
          for (int g = 0; g < G; g++) {
            const double tau = st[g] * len;
            const double dl = (tau * (double)psi[g] - len * Q[r * G + g]) * exp_f1(tau);
            psi[g] = (PsiT)((double)psi[g] - dl);
            buf[g] += dl;
          }
          const bool last = kk == ns - 1;
          const int64_t knext = d == 0 ? k + 1 : k - 1;
          const bool flush = (d == 0) ? (!last && p->seg_fsr[knext] != r) : (last || p->seg_fsr[knext] != r);
          if (flush) {
            for (int g = 0; g < G; g++) {
#pragma omp atomic
              tally[r * G + g] += w * buf[g];
              buf[g] = 0.0;
            }
          }
        }
        const uint8_t bc = d == 0 ? p->bc_fwd[t] : p->bc_bwd[t];
        if (bc == 1 || bc == 2) {
          const int64_t nt = d == 0 ? p->next_fwd[t] : p->next_bwd[t];
          const int nd = d == 0 ? !p->next_fwd_is_fwd[t] : !p->next_bwd_is_fwd[t];
          PsiT* out = start + (nt * 2 + nd) * G;
          for (int g = 0; g < G; g++) out[g] = psi[g];
        }
      }
    }
  }
}

template <typename PsiT>
void sweep(const View& v, const double* Q, PsiT* start, double* tally, int threads,
           float* leakage = nullptr /* [T] per track, or null */) {
  const OracleProblem* p = v.p;
  if (p->trk_polar) {
    sweep3d(v, Q, start, tally, threads);
    return;
  }
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
// phi0 / psi0 (nullable, MODE_FLUX only): the scalar flux [N*G] and start
// fluxes [2*T*fpt] a previous solve left; computeFlux(only_fixed_source =
// false) after a solve keeps them instead of the zero start (:1379-1386), with
// k = 1 for the sources of computeFSRSources(0) (:1370, :1389).  psi_out
// (nullable) receives the start fluxes at the end.
template <typename PsiT>
int solve(const OracleProblem* p, const OracleOptions* o, OracleResult* out, double* flux_out,
          Mode mode, double k_fixed, const double* phi0 = nullptr, const PsiT* psi0 = nullptr,
          PsiT* psi_out = nullptr) {
  View v = make_view(p);
  const int G = v.G;
  const int64_t N = v.N, T = v.T;
  if (o->threads > 0) omp_set_num_threads(o->threads);
  std::vector<double> phi(N * G, mode == MODE_FLUX ? 0.0 : 1.0), old(N * G), Q(N * G),
      scratch(N), tally(N * G);
  std::vector<PsiT> start(2 * T * v.fpt, (PsiT)0);
  if (mode == MODE_FLUX && phi0) std::memcpy(phi.data(), phi0, sizeof(double) * N * G);
  if (mode == MODE_FLUX && psi0) std::memcpy(start.data(), psi0, sizeof(PsiT) * start.size());
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
  if (psi_out) std::memcpy(psi_out, start.data(), sizeof(PsiT) * start.size());
  out->keff = k;
  out->num_iterations = iters;
  out->residual = res;
  out->sweep_seconds = sweep_s;
  out->num_integrations = 2LL * G * v.P * p->seg_offset[T] * iters;
  return 0;
}

// ---------------------------------------------------------------------------
// Linear source (2D), CPULSSolver
// ---------------------------------------------------------------------------

// 5/6 rational fit of G(x) = 1/x - (1 - e^-x)/x^2 (src/exponentials.h:100-145)
inline double exp_g(double x) {
  const double p0 = 0.5, p1 = 1.76558112351595e-1, p2 = 4.041584305811143e-2,
               p3 = 6.178333902037397e-3, p4 = 6.429894635552992e-4,
               p5 = 6.064409107557148e-5;
  const double d1 = 6.864462055546078e-1, d2 = 2.263358514260129e-1,
               d3 = 4.721469893686252e-2, d4 = 6.883236664917246e-3,
               d5 = 7.036272419147752e-4, d6 = 6.064409107557148e-5;
  double den = d6 * x + d5;
  den = den * x + d4;
  den = den * x + d3;
  den = den * x + d2;
  den = den * x + d1;
  den = den * x + 1.0;
  den = 1.0 / den;
  double num = p5 * x + p4;
  num = num * x + p3;
  num = num * x + p2;
  num = num * x + p1;
  num = num * x + p0;
  return num * den;
}

// 5/5 rational fit of G2(x) (src/exponentials.h:280-323)
inline double exp_g2(double x) {
  const double a1 = -8.335775885589858e-2, a2 = -3.603942303847604e-3,
               a3 = 3.7673183263550827e-3, a4 = 1.124183494990467e-5,
               a5 = 1.6837426505799449e-4;
  const double b1 = 7.454048371823628e-1, b2 = 2.3794300531408347e-1,
               b3 = 5.367250964303789e-2, b4 = 6.125197988351906e-3,
               b5 = 1.0102514456857377e-3;
  double num = a5 * x + a4;
  num = num * x + a3;
  num = num * x + a2;
  num = num * x + a1;
  num *= x;
  double den = b5 * x + b4;
  den = den * x + b3;
  den = den * x + b2;
  den = den * x + b1;
  den = den * x + 1.0;
  return num / den;
}

constexpr double kMinDet = 1e-10;  // MIN_DET (src/constants.h:70)

struct LinearSource {
  std::vector<double> centroid;  // [2N]
  std::vector<double> pos;       // [2S] segment start relative to its FSR centroid
  std::vector<double> ilem;      // [3N] inverse linear-expansion matrix
  std::vector<double> src_c;     // [N][3][G] source constants
};

// FSR centroids (CentroidGenerator, src/TrackTraversingAlgorithms.cpp:377-460,
// from the track start and the segment lengths), the segments re-centred on
// them (RecenterSegments, :1322-1332), and the LinearExpansionGenerator
// (:470-831): moments of the segments about the centroid, their inverse and
// the source constants.
LinearSource ls_setup(const View& v) {
  const OracleProblem* p = v.p;
  const int G = v.G, P = v.P;
  const int64_t N = v.N, T = v.T;
  LinearSource ls;
  ls.centroid.assign(2 * N, 0.0);
  for (int64_t t = 0; t < T; t++) {
    const int a = p->trk_azim[t];
    const double w = p->azim_weight[a];
    const double cp = std::cos(p->azim_phi[a]), sp = std::sin(p->azim_phi[a]);
    const int64_t s0 = p->seg_offset[t], s1 = p->seg_offset[t + 1];
    if (s1 == s0) continue;
    double x = p->seg_start[2 * s0], y = p->seg_start[2 * s0 + 1];
    for (int64_t k = s0; k < s1; k++) {
      const int32_t r = p->seg_fsr[k];
      const double l = p->seg_length[k];
      ls.centroid[2 * r] += w * (x + cp * l / 2.0) * l / p->fsr_volume[r];
      ls.centroid[2 * r + 1] += w * (y + sp * l / 2.0) * l / p->fsr_volume[r];
      x += cp * l;
      y += sp * l;
    }
  }
  const int64_t S = p->seg_offset[T];
  ls.pos.resize(2 * S);
  for (int64_t k = 0; k < S; k++) {
    const int32_t r = p->seg_fsr[k];
    ls.pos[2 * k] = p->seg_start[2 * k] - ls.centroid[2 * r];
    ls.pos[2 * k + 1] = p->seg_start[2 * k + 1] - ls.centroid[2 * r + 1];
  }
  std::vector<double> lem(3 * N, 0.0);
  ls.src_c.assign(N * 3 * G, 0.0);
  for (int64_t t = 0; t < T; t++) {
    const int a = p->trk_azim[t];
    const double wgt = p->azim_weight[a];
    const double cp = std::cos(p->azim_phi[a]), sp = std::sin(p->azim_phi[a]);
    for (int64_t k = p->seg_offset[t]; k < p->seg_offset[t + 1]; k++) {
      const int32_t r = p->seg_fsr[k];
      const double l = p->seg_length[k];
      const double vol = p->fsr_volume[r];
      const double xc = ls.pos[2 * k] + l * 0.5 * cp;
      const double yc = ls.pos[2 * k + 1] + l * 0.5 * sp;
      const double vol_impact = wgt * l / vol;
      const double src_constant = vol_impact * l / 2.0;
      const double* st = v.sig_t(r);
      double* c = &ls.src_c[r * 3 * G];
      for (int g = 0; g < G; g++) {
        double c0 = vol_impact * xc * xc, c1 = vol_impact * yc * yc, c2 = vol_impact * xc * yc;
        const double tau = l * st[g];
        for (int q = 0; q < P; q++) {
          const double sin_t = 1.0 / p->inv_sin_theta[q];
          const double g2 = l * exp_g2(tau / sin_t) * src_constant * 2 * p->polar_weight[q] * sin_t;
          c0 += cp * cp * g2;
          c1 += sp * sp * g2;
          c2 += sp * cp * g2;
        }
        c[g] += c0;
        c[G + g] += c1;
        c[2 * G + g] += c2;
      }
      lem[3 * r] += wgt * l / vol * (xc * xc + std::pow(cp * l, 2) / 12.0);
      lem[3 * r + 1] += wgt * l / vol * (yc * yc + std::pow(sp * l, 2) / 12.0);
      lem[3 * r + 2] += wgt * l / vol * (xc * yc + sp * cp * std::pow(l, 2) / 12.0);
    }
  }
  ls.ilem.assign(3 * N, 0.0);
  for (int64_t r = 0; r < N; r++) {
    const double det = lem[3 * r] * lem[3 * r + 1] - lem[3 * r + 2] * lem[3 * r + 2];
    if (std::abs(det) >= kMinDet) {
      ls.ilem[3 * r] = lem[3 * r + 1] / det;
      ls.ilem[3 * r + 1] = lem[3 * r] / det;
      ls.ilem[3 * r + 2] = -lem[3 * r + 2] / det;
    }
  }
  return ls;
}

// Source moments (CPULSSolver::computeFSRSources, src/CPULSSolver.cpp:386-524),
// after the flat sources Q of the same iteration.
void ls_sources(const View& v, const LinearSource& ls, const double* phi_xy, const double* Q,
                double k, int it, bool allow_neg, double* Qxy) {
  const int G = v.G;
  for (int64_t r = 0; r < v.N; r++) {
    const double* s = v.sig_s(r);
    const double* chi = v.chi(r);
    const double* nsf = v.nsf(r);
    const bool fis = v.fissionable(r);
    std::vector<double> tmp(G);
    for (int g = 0; g < G; g++) {
      double src[2];
      for (int m = 0; m < 2; m++) {
        double fission = 0.0;
        if (fis) {
          for (int gp = 0; gp < G; gp++) tmp[gp] = chi[g] * nsf[gp] * phi_xy[(r * G + gp) * 2 + m];
          fission = pairwise(tmp.data(), G) / k;
        }
        for (int gp = 0; gp < G; gp++) tmp[gp] = s[g * G + gp] * phi_xy[(r * G + gp) * 2 + m];
        src[m] = pairwise(tmp.data(), G) + fission;
      }
      double* q = &Qxy[(r * G + g) * 2];
      if (allow_neg || Q[r * G + g] > 10 * kFluxEpsilon || it > 29) {
        q[0] = kOneOverFourPi / 2 * (ls.ilem[3 * r] * src[0] + ls.ilem[3 * r + 2] * src[1]);
        q[1] = kOneOverFourPi / 2 * (ls.ilem[3 * r + 2] * src[0] + ls.ilem[3 * r + 1] * src[1]);
      } else {
        q[0] = 0.0;
        q[1] = 0.0;
      }
    }
  }
}

// Linear-source transport sweep (tallyLSScalarFlux 2D, src/CPULSSolver.cpp:
// 542-739, accumulateLinearFluxContribution :749-780, TransportSweep::onTrack).
template <typename PsiT>
void ls_sweep(const View& v, const LinearSource& ls, const double* Q, const double* Qxy,
              PsiT* start, double* tally, double* tally_xy) {
  const OracleProblem* p = v.p;
  const int G = v.G, P = v.P, fpt = v.fpt;
  const int64_t T = v.T;
  std::memset(tally, 0, sizeof(double) * v.N * G);
  std::memset(tally_xy, 0, sizeof(double) * v.N * G * 2);
  std::vector<PsiT> boundary(start, start + 2 * T * fpt);
#pragma omp parallel
  {
    std::vector<double> buf(G), bx(G), by(G);
#pragma omp for schedule(dynamic)
    for (int64_t t = 0; t < T; t++) {
      const int a = p->trk_azim[t];
      const double* w = p->total_weight + (int64_t)a * P;
      const double cp = std::cos(p->azim_phi[a]), sp = std::sin(p->azim_phi[a]);
      const int64_t s0 = p->seg_offset[t], s1 = p->seg_offset[t + 1];
      const int64_t ns = s1 - s0;
      // the FSR buffers carry over from the forward to the backward pass, as
      // the reference flushes only when the FSR changes (and at the very end)
      std::fill(buf.begin(), buf.end(), 0.0);
      std::fill(bx.begin(), bx.end(), 0.0);
      std::fill(by.begin(), by.end(), 0.0);
      for (int d = 0; d < 2; d++) {
        PsiT* psi = &boundary[(t * 2 + d) * fpt];
        const double dx = d == 0 ? cp : -cp, dy = d == 0 ? sp : -sp;
        for (int64_t kk = 0; kk < ns; kk++) {
          const int64_t k = d == 0 ? s0 + kk : s1 - 1 - kk;
          const int32_t r = p->seg_fsr[k];
          const double l = p->seg_length[k];
          const double* st = v.sig_t(r);
          // forward: the segment start; backward: its end
          double px = ls.pos[2 * k], py = ls.pos[2 * k + 1];
          if (d == 1) {
            px += cp * l;
            py += sp * l;
          }
          const double cx = 2 * px + l * dx, cy = 2 * py + l * dy;
          for (int g = 0; g < G; g++) {
            const double* qxy = &Qxy[(r * G + g) * 2];
            const double src_flat = Q[r * G + g] + qxy[0] * cx + qxy[1] * cy;
            const double src_lin = dx * qxy[0] + dy * qxy[1];
            const double tau = st[g] * l;
            for (int q = 0; q < P; q++) {
              const double isin = p->inv_sin_theta[q];
              const double tq = std::max(1e-8, tau * isin);
              const double eg = exp_g(tq);
              const double f1 = (1.0 - tq * eg) * isin;
              const double gq = eg * isin;
              const double f2 = 2.0 * gq - f1;
              double h = f1 - gq;
              const double wgt = w[q];
              const double ps = (double)psi[q * G + g];
              h *= wgt * tau * l * ps;
              double dl = (tau * ps - l * src_flat) * f1 - l * l * src_lin * f2;
              psi[q * G + g] = (PsiT)(ps - dl);
              dl *= wgt;
              buf[g] += dl;
              bx[g] += h * dx + dl * px;
              by[g] += h * dy + dl * py;
            }
          }
          const bool last = kk == ns - 1;
          const int64_t knext = d == 0 ? k + 1 : k - 1;
          const bool flush = (d == 0) ? (!last && p->seg_fsr[knext] != r)
                                      : (last || p->seg_fsr[knext] != r);
          if (flush) {
            for (int g = 0; g < G; g++) {
#pragma omp atomic
              tally[r * G + g] += buf[g];
#pragma omp atomic
              tally_xy[(r * G + g) * 2] += bx[g];
#pragma omp atomic
              tally_xy[(r * G + g) * 2 + 1] += by[g];
              buf[g] = bx[g] = by[g] = 0.0;
            }
          }
        }
        const uint8_t bc = d == 0 ? p->bc_fwd[t] : p->bc_bwd[t];
        if (bc == 1 || bc == 2) {
          const int64_t nt = d == 0 ? p->next_fwd[t] : p->next_bwd[t];
          const int nd = d == 0 ? !p->next_fwd_is_fwd[t] : !p->next_bwd_is_fwd[t];
          PsiT* out = start + (nt * 2 + nd) * fpt;
          for (int e = 0; e < fpt; e++) out[e] = psi[e];
        }
      }
    }
  }
}

// CPULSSolver::addSourceToScalarFlux (src/CPULSSolver.cpp:787-882).
void ls_add_source(const View& v, const LinearSource& ls, const double* Q, const double* Qxy,
                   const double* old, bool allow_neg, double* phi, double* phi_xy) {
  const int G = v.G;
  for (int64_t r = 0; r < v.N; r++) {
    double vol = v.p->fsr_volume[r];
    if (vol < kFltEpsilon) vol = 1e30;
    const double* st = v.sig_t(r);
    const double* c = &ls.src_c[r * 3 * G];
    for (int g = 0; g < G; g++) {
      const int64_t i = r * G + g;
      const double fc = kFourPi * 2;
      phi[i] /= vol;
      phi[i] += kFourPi * Q[i];
      phi[i] /= st[g];
      double* m = &phi_xy[i * 2];
      const double* q = &Qxy[i * 2];
      m[0] /= vol;
      m[0] += fc * q[0] * c[g];
      m[0] += fc * q[1] * c[2 * G + g];
      m[1] /= vol;
      m[1] += fc * q[0] * c[2 * G + g];
      m[1] += fc * q[1] * c[G + g];
      m[0] /= st[g];
      m[1] /= st[g];
      if (phi[i] < 0.0 && !allow_neg) {
        phi[i] = std::max(old[i], kFluxEpsilon);
        m[0] = 0.0;
        m[1] = 0.0;
      }
    }
  }
}

template <typename PsiT>
int solve_ls(const OracleProblem* p, const OracleOptions* o, OracleResult* out,
             double* flux_out, double* flux_xy_out, double* centroid_out) {
  if (!p->azim_phi || !p->azim_weight || !p->polar_weight || !p->seg_start) return -1;
  if (o->keff_from_fission_rates == 0 || o->stab_type >= 0) return -3;
  View v = make_view(p);
  const int G = v.G;
  const int64_t N = v.N, T = v.T;
  if (o->threads > 0) omp_set_num_threads(o->threads);
  const LinearSource ls = ls_setup(v);
  std::vector<double> phi(N * G, 1.0), old(N * G), Q(N * G), scratch(N), tally(N * G);
  std::vector<double> phi_xy(N * G * 2, 0.0), Qxy(N * G * 2, 0.0), tally_xy(N * G * 2);
  std::vector<PsiT> start(2 * T * v.fpt, (PsiT)0);
  const bool allow_neg = o->allow_negative_fluxes != 0;
  double k = 1.0, res = 0.0, sweep_s = 0.0;
  int iters = 0;
  auto normalize = [&]() {
    const double tot = fission_rate_total(v, phi.data(), scratch);
    const double f = (double)N / tot;
    for (auto& x : phi) x *= f;
    for (auto& x : start) x = (PsiT)(x * f);
    for (auto& x : phi_xy) x *= f;
  };
  normalize();
  old = phi;
  double k_prev = k;
  for (int i = 0; i < o->max_iters; i++) {
    fsr_sources(v, phi.data(), k, i, allow_neg, Q.data(), 0);
    ls_sources(v, ls, phi_xy.data(), Q.data(), k, i, allow_neg, Qxy.data());
    auto t0 = std::chrono::steady_clock::now();
    ls_sweep<PsiT>(v, ls, Q.data(), Qxy.data(), start.data(), tally.data(), tally_xy.data());
    sweep_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    phi = tally;
    phi_xy = tally_xy;
    ls_add_source(v, ls, Q.data(), Qxy.data(), old.data(), allow_neg, phi.data(), phi_xy.data());
    const double rate = fission_rate_total(v, phi.data(), scratch);
    k *= rate / (double)N;
    normalize();
    res = residual(v, phi.data(), old.data(), o->res_type, k, scratch);
    const int dk = (int)(1e5 * (k - k_prev));
    k_prev = k;
    old = phi;
    iters++;
    if (res < o->tolerance && std::abs(dk) < 1) break;
  }
  std::memcpy(flux_out, phi.data(), sizeof(double) * N * G);
  if (flux_xy_out) std::memcpy(flux_xy_out, phi_xy.data(), sizeof(double) * N * G * 2);
  if (centroid_out) std::memcpy(centroid_out, ls.centroid.data(), sizeof(double) * 2 * N);
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

int oracle_compute_eigenvalue_ls(const OracleProblem* p, const OracleOptions* o,
                                 OracleResult* r, double* flux_out, double* flux_xy_out,
                                 double* centroid_out) {
  if (!p || !o || !r || !flux_out) return -1;
  return o->psi_fp64 ? solve_ls<double>(p, o, r, flux_out, flux_xy_out, centroid_out)
                     : solve_ls<float>(p, o, r, flux_out, flux_xy_out, centroid_out);
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

int oracle_solve_state(const OracleProblem* p, const OracleOptions* o, int mode, double k_eff,
                       const double* phi0, const float* psi0, OracleResult* r, double* flux_out,
                       float* psi_out) {
  if (!p || !o || !r || !flux_out || mode < MODE_EIGEN || mode > MODE_SOURCE || o->psi_fp64)
    return -1;
  if (mode == MODE_SOURCE && !(k_eff > 0)) return -1;
  if ((phi0 || psi0) && mode != MODE_FLUX) return -1;
  return solve<float>(p, o, r, flux_out, (Mode)mode, mode == MODE_SOURCE ? k_eff : 1.0, phi0,
                      psi0, psi_out);
}

}  // extern "C"
