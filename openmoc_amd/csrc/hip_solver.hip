// HIP kernels and host driver of the MOC source iteration on gfx950.
//
// Per source iteration (Solver::computeEigenvalue, src/Solver.cpp:1614-1682):
//   sweep2d_kernel     transportSweep: FP32 angular fluxes in registers, one
//                      wave per tile of directed tracks, lanes over groups,
//                      polar angles unrolled in-lane, per-segment tally with
//                      hardware float atomics into an L2-resident FP32 tally;
//                      the outgoing flux is written straight into the linked
//                      track's start slot of the other ping-pong buffer
//                      (transferBoundaryFlux fused, src/CPUSolver.cpp:2560-2601)
//   post_sweep_kernel  addSourceToScalarFlux (FP64) + per-FSR fission rate
//                      partials + zeroing of the tally for the next sweep
//   keff_kernel        computeKeff + normalisation factor (FP64, fixed order)
//   update_kernel      normalizeFluxes (phi), computeResidual partials,
//                      storeFSRFluxes and computeFSRSources for the next
//                      iteration (FP64), and the FP32 (Q/sigma_t, sigma_t)
//                      table the next sweep gathers
//   residual_kernel    RMS residual, delta-k and the stopping rule on device
// The normalisation of the start fluxes (CPUSolver::normalizeFluxes scales
// _start_flux, src/CPUSolver.cpp:1920-1928) is applied when the next sweep
// loads them (psi_scale), so no extra pass over the angular fluxes is made.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>

#include "hip_solver.h"

namespace moc {

#define HIP_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw DeviceError(MOC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

constexpr double kFourPi = 12.566370614359172;
constexpr double kOneOverFourPi = 0.07957747154594767;
constexpr double kFluxEpsilon = 1.0e-25;
constexpr double kFltEpsilon = 1.0e-12;
constexpr int kMaxPolarHalf = 3;
constexpr int kReduceBlock = 256;

template <typename T>
void dev_alloc(T** p, size_t n) {
  if (*p) {
    HIP_CHECK(hipFree(*p));
    *p = nullptr;
  }
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    throw DeviceError(MOC_ERR_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
}

template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// ---------------------------------------------------------------------------
// Exponential attenuation 1 - exp(-x)
// ---------------------------------------------------------------------------
// x * F1(x) with F1 the 5/6 rational fit of (1-exp(-x))/x the reference
// evaluates in its CPU and GPU sweeps (src/exponentials.h:156-192,
// src/accel/cuda/dev_exponential.h:32-56).  Horner in FMA, one v_rcp_f32.
__device__ __forceinline__ float attenuation_rational(float x) {
  float den = fmaf(1.9309063097411041e-4f, x, 1.0057980007137651e-3f);
  den = fmaf(den, x, 1.0590759992367811e-2f);
  den = fmaf(den, x, 6.1643725066901411e-2f);
  den = fmaf(den, x, 2.6722515319494311e-1f);
  den = fmaf(den, x, 7.4169266112320541e-1f);
  den = fmaf(den, x, 1.0f);
  float num = fmaf(1.9309063097411041e-4f, x, 1.0059468082903561e-3f);
  num = fmaf(num, x, 1.0567595009016521e-2f);
  num = fmaf(num, x, 6.2804790965268531e-2f);
  num = fmaf(num, x, 2.4172687328033081e-1f);
  num = fmaf(num, x, 1.0f);
  return x * num * __builtin_amdgcn_rcpf(den);
}

// 1 - 2^(-y) with y = x log2(e) folded into the polar factor by the caller.
__device__ __forceinline__ float attenuation_hw(float y) {
  return 1.0f - __builtin_amdgcn_exp2f(-y);
}

// ---------------------------------------------------------------------------
// Transport sweep
// ---------------------------------------------------------------------------
struct SweepArgs {
  const int32_t* work;      // directed track ids (t*2+d)
  int num_work;
  const int4* trk;          // seg_begin, seg_count, out slot fwd, out slot bwd
  const int32_t* trk_azim;
  const float* seg_len;
  const int32_t* seg_fsr;
  const float2* qs;         // [N*G] (Q/sigma_t, sigma_t)
  const float* psi_in;      // [2T*fpt]
  float* psi_out;           // [2T*fpt]
  float* tally;             // [N*G]
  const float* weights;     // [num_azim/2][P]
  float pf0, pf1, pf2;      // polar factors: 1/sin(theta_p) (times log2 e for HW exp)
  const IterState* state;
  int G, P, fpt, lanes_per_track, tracks_per_wave;
};

template <int GPL, int EXPM>
__global__ __launch_bounds__(256) void sweep2d_kernel(SweepArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int slot = lane / a.lanes_per_track;
  const int gl = lane - slot * a.lanes_per_track;
  if (slot >= a.tracks_per_wave) return;
  const int widx = wave * a.tracks_per_wave + slot;
  if (widx >= a.num_work) return;
  if (a.state->converged) return;

  const int dt = a.work[widx];
  const int t = dt >> 1;
  const int dir = dt & 1;
  const int4 tr = a.trk[t];
  const int azim = a.trk_azim[t];
  const int G = a.G, P = a.P, fpt = a.fpt;
  const float scale = (float)a.state->psi_scale;
  const float pf[kMaxPolarHalf] = {a.pf0, a.pf1, a.pf2};
  float w[kMaxPolarHalf];
#pragma unroll
  for (int p = 0; p < kMaxPolarHalf; p++) w[p] = p < P ? a.weights[azim * P + p] : 0.f;

  int g[GPL];
  bool ok[GPL];
  float psi[GPL][kMaxPolarHalf];
  const float* pin = a.psi_in + (size_t)dt * fpt;
#pragma unroll
  for (int k = 0; k < GPL; k++) {
    g[k] = gl + k * a.lanes_per_track;
    ok[k] = g[k] < G;
#pragma unroll
    for (int p = 0; p < kMaxPolarHalf; p++)
      psi[k][p] = (ok[k] && p < P) ? pin[p * G + g[k]] * scale : 0.f;
  }

  const int n = tr.y;
  const int step = dir ? -1 : 1;
  int s = dir ? tr.x + n - 1 : tr.x;
  float len = 0.f;
  int fsr = 0;
  if (n > 0) {
    len = a.seg_len[s];
    fsr = a.seg_fsr[s];
  }
  for (int i = 0; i < n; i++) {
    const int sn = s + step;
    float len_n = 0.f;
    int fsr_n = 0;
    if (i + 1 < n) {  // software prefetch of the next segment
      len_n = a.seg_len[sn];
      fsr_n = a.seg_fsr[sn];
    }
    const size_t base = (size_t)fsr * G;
#pragma unroll
    for (int k = 0; k < GPL; k++) {
      if (!ok[k]) continue;
      const float2 q = a.qs[base + g[k]];
      const float tau = q.y * len;
      float acc = 0.f;
#pragma unroll
      for (int p = 0; p < kMaxPolarHalf; p++) {
        if (p < P) {
          const float x = tau * pf[p];
          const float A = EXPM == MOC_EXP_RATIONAL ? attenuation_rational(x) : attenuation_hw(x);
          const float d = (psi[k][p] - q.x) * A;
          psi[k][p] -= d;
          acc = fmaf(w[p], d, acc);
        }
      }
      unsafeAtomicAdd(&a.tally[base + g[k]], acc);
    }
    len = len_n;
    fsr = fsr_n;
    s = sn;
  }

  const int out = dir ? tr.w : tr.z;
  if (out >= 0) {
    float* po = a.psi_out + (size_t)out * fpt;
#pragma unroll
    for (int k = 0; k < GPL; k++)
#pragma unroll
      for (int p = 0; p < kMaxPolarHalf; p++)
        if (ok[k] && p < P) po[p * G + g[k]] = psi[k][p];
  }
}

// ---------------------------------------------------------------------------
// FP64 per-FSR kernels
// ---------------------------------------------------------------------------
struct FsrArgs {
  int64_t N;
  int G;
  const int32_t* fsr_mat;
  const double* fsr_vol;
  const double* sigma_t;
  const double* sigma_s;      // dest-major
  const double* nsf;
  const double* chi;
  const int32_t* fissionable;
  const double* fixed;        // nullable
  double* phi;
  double* phi_old;
  double* Q;
  float2* qs;
  float* tally;
  double* partials;
  IterState* state;
  int allow_negative;
};

// Fixed-order block sum (deterministic): wave shuffles, then LDS.
__device__ double block_sum(double v) {
  __shared__ double sh[kReduceBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sh[i];
  __syncthreads();
  return r;  // valid in thread 0
}

// from_tally=1: phi = tally/(sigma_t V) + 4 pi Q / sigma_t (addSourceToScalarFlux,
// src/CPUSolver.cpp:2608-2659) and the tally is cleared for the next sweep.
// Always: partials[block] = sum over the block's FSRs of V * sum_g nsf*phi.
__global__ __launch_bounds__(kReduceBlock) void post_sweep_kernel(FsrArgs a, int from_tally) {
  if (a.state->converged) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rate = 0.0;
  int64_t neg = 0;
  if (r < a.N) {
    const int G = a.G;
    const int m = a.fsr_mat[r];
    const double vol_true = a.fsr_vol[r];
    const double vol = vol_true < kFltEpsilon ? 1e30 : vol_true;
    const double* st = a.sigma_t + (size_t)m * G;
    const double* nsf = a.nsf + (size_t)m * G;
    for (int g = 0; g < G; g++) {
      const size_t i = (size_t)r * G + g;
      double f;
      if (from_tally) {
        f = (double)a.tally[i] / (st[g] * vol) + kFourPi * a.Q[i] / st[g];
        if (f < 0.0 && !a.allow_negative) {
          f = kFluxEpsilon;
          neg++;
        }
        a.phi[i] = f;
        a.tally[i] = 0.f;
      } else {
        f = a.phi[i];
      }
      rate += nsf[g] * f;
    }
    rate *= vol_true;
  }
  const double s = block_sum(rate);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
  if (neg) atomicAdd((unsigned long long*)&a.state->negative_fluxes, (unsigned long long)neg);
}

// phase 0: initial guess normalisation (k unchanged); phase 1: computeKeff
// (k *= F/N, src/CPUSolver.cpp:2258-2328) then normalizeFluxes' factor N/F
// (src/CPUSolver.cpp:1860-1931).
__global__ __launch_bounds__(kReduceBlock) void keff_kernel(const double* partials, int nparts,
                                                           IterState* state, double num_fsrs,
                                                           int phase) {
  if (state->converged) return;
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += partials[i];
  const double total = block_sum(v);
  if (threadIdx.x == 0) {
    if (phase == 1) state->k = state->k * total / num_fsrs;
    state->fission_total = total;
    state->norm_factor = num_fsrs / total;
    state->psi_scale = num_fsrs / total;
  }
}

// normalise (optional), residual partials, store old flux, next sources.
__global__ __launch_bounds__(kReduceBlock) void update_kernel(FsrArgs a, int normalize,
                                                             int res_type, int do_residual,
                                                             int compute_sources,
                                                             int source_iteration_offset) {
  if (a.state->converged) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double res = 0.0;
  int64_t negsrc = 0;
  if (r < a.N) {
    const int G = a.G;
    const int m = a.fsr_mat[r];
    const double f = normalize ? a.state->norm_factor : 1.0;
    const double k = a.state->k;
    const double* st = a.sigma_t + (size_t)m * G;
    const double* ss = a.sigma_s + (size_t)m * G * G;
    const double* nsf = a.nsf + (size_t)m * G;
    const double* chi = a.chi + (size_t)m * G;
    const bool fis = a.fissionable[m] != 0;
    double* phi = a.phi + (size_t)r * G;
    double* old = a.phi_old + (size_t)r * G;
    double new_fis = 0.0, old_fis = 0.0;
    if (normalize)
      for (int g = 0; g < G; g++) phi[g] *= f;
    if (do_residual) {
      if (res_type == MOC_SCALAR_FLUX) {
        for (int g = 0; g < G; g++)
          if (old[g] > 0.0) {
            const double d = (phi[g] - old[g]) / old[g];
            res += d * d;
          }
      } else {
        double nw = 0.0, od = 0.0;
        if (fis) {
          for (int g = 0; g < G; g++) {
            nw += phi[g] * nsf[g];
            od += old[g] * nsf[g];
          }
          if (res_type == MOC_TOTAL_SOURCE) {
            nw /= k;
            od /= k;
          }
        }
        if (res_type == MOC_TOTAL_SOURCE) {
          for (int gd = 0; gd < G; gd++)
            for (int g = 0; g < G; g++) {
              nw += ss[gd * G + g] * phi[g];
              od += ss[gd * G + g] * old[g];
            }
        }
        if ((res_type == MOC_TOTAL_SOURCE || fis) && od > 0.0) {
          const double d = (nw - od) / od;
          res = d * d;
        }
      }
    }
    for (int g = 0; g < G; g++) old[g] = phi[g];
    if (compute_sources) {
      // computeFSRSources (src/CPUSolver.cpp:1939-2023)
      const int it = a.state->iteration + source_iteration_offset;
      if (fis)
        for (int g = 0; g < G; g++) new_fis += nsf[g] * phi[g];
      for (int gd = 0; gd < G; gd++) {
        double sc = 0.0;
        const double* row = ss + (size_t)gd * G;
        for (int g = 0; g < G; g++) sc += row[g] * phi[g];
        double q = (fis ? chi[gd] * new_fis : 0.0) / k + sc;
        if (a.fixed) q += a.fixed[(size_t)r * G + gd];
        q *= kOneOverFourPi;
        if (q < 0.0) {
          negsrc++;
          if (it < 30 && !a.allow_negative) q = kFluxEpsilon;
        }
        const size_t i = (size_t)r * G + gd;
        a.Q[i] = q;
        a.qs[i] = make_float2((float)(q / st[gd]), (float)st[gd]);
      }
    }
    (void)old_fis;
  }
  const double s = block_sum(res);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
  if (negsrc) atomicAdd((unsigned long long*)&a.state->negative_sources, (unsigned long long)negsrc);
}

// mode 0 eigen: stop when res < tol && |int(1e5 dk)| < 1 (src/Solver.cpp:1643-1681)
// mode 1/2 flux/source: stop when i > 1 && res < tol, _num_iterations = i.
__global__ __launch_bounds__(kReduceBlock) void residual_kernel(const double* partials, int nparts,
                                                               IterState* state, double norm,
                                                               double tol, int mode) {
  if (state->converged) return;
  double v = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += partials[i];
  const double sum = block_sum(v);
  if (threadIdx.x == 0) {
    double s = sum < 0.0 ? 0.0 : sum;
    const double res = sqrt(s / (norm <= 0.0 ? 1.0 : norm));
    state->residual = res;
    const int i = state->iteration;
    if (mode == 0) {
      const int dk = (int)(1e5 * (state->k - state->k_prev));
      state->dk = dk;
      state->k_prev = state->k;
      state->iteration = i + 1;
      state->reported_iters = i + 1;
      if (res < tol && abs(dk) < 1) state->converged = 1;
    } else {
      state->iteration = i + 1;
      state->reported_iters = i + 1;
      if (i > 1 && res < tol) {
        state->converged = 1;
        state->reported_iters = i;
      }
    }
  }
}

__global__ void fill_kernel(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void zero_f32_kernel(float* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

__global__ void make_qs_kernel(const double* Q, const int32_t* fsr_mat, const double* sigma_t,
                               float2* qs, int64_t N, int G) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N * G) {
    const int64_t r = i / G;
    const int g = (int)(i - r * G);
    const double st = sigma_t[(size_t)fsr_mat[r] * G + g];
    qs[i] = make_float2((float)(Q[i] / st), (float)st);
  }
}

__global__ void fission_rates_kernel(const double* phi, const int32_t* fsr_mat,
                                     const double* nsf, const double* vol, double* out,
                                     int64_t N, int G) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < N) {
    const double* v = nsf + (size_t)fsr_mat[r] * G;
    double s = 0.0;
    for (int g = 0; g < G; g++) s += v[g] * phi[(size_t)r * G + g] * vol[r];
    out[r] = s;
  }
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ===========================================================================
// Host driver
// ===========================================================================
HipSolver::HipSolver(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    throw DeviceError(MOC_ERR_NO_DEVICE, "no HIP device visible");
  if (device < 0) HIP_CHECK(hipGetDevice(&device));
  if (device >= count) throw DeviceError(MOC_ERR_INVALID, "device index out of range");
  device_ = device;
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HIP_CHECK(hipMalloc((void**)&d_state_, sizeof(IterState)));
  HIP_CHECK(hipHostMalloc((void**)&h_state_, sizeof(IterState), hipHostMallocDefault));
  std::memset(h_state_, 0, sizeof(IterState));
}

HipSolver::~HipSolver() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  dev_free(d_seg_len_);
  dev_free(d_seg_fsr_);
  dev_free(d_work_);
  dev_free(d_trk_);
  dev_free(d_trk_azim_);
  dev_free(d_weights_);
  dev_free(d_psi_[0]);
  dev_free(d_psi_[1]);
  dev_free(d_tally_);
  dev_free(d_qs_);
  dev_free(d_phi_);
  dev_free(d_phi_old_);
  dev_free(d_Q_);
  dev_free(d_fixed_);
  dev_free(d_sigma_t_);
  dev_free(d_sigma_s_);
  dev_free(d_nsf_);
  dev_free(d_chi_);
  dev_free(d_mat_fissionable_);
  dev_free(d_fsr_mat_);
  dev_free(d_fsr_vol_);
  dev_free(d_partials_);
  dev_free(d_state_);
  if (h_state_) (void)hipHostFree(h_state_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void HipSolver::pickShape() {
  if (shape_forced_ || G_ <= 0) return;
  LaunchShape s;
  static const int kGpl[] = {1, 2, 4, 9, 16};
  s.groups_per_lane = 0;
  for (int gpl : kGpl) {
    if ((G_ + gpl - 1) / gpl <= 64) {
      s.groups_per_lane = gpl;
      break;
    }
  }
  if (s.groups_per_lane == 0) throw DeviceError(MOC_ERR_INVALID, "too many energy groups (max 1024)");
  s.lanes_per_track = (G_ + s.groups_per_lane - 1) / s.groups_per_lane;
  s.tracks_per_wave = 64 / s.lanes_per_track;
  s.block = 256;
  shape_ = s;
}

void HipSolver::setShape(const LaunchShape& s) {
  static const int kGpl[] = {1, 2, 4, 9, 16};
  bool okg = false;
  for (int gpl : kGpl) okg |= gpl == s.groups_per_lane;
  if (!okg || s.lanes_per_track <= 0 || s.lanes_per_track > 64 ||
      s.lanes_per_track * s.groups_per_lane < G_ ||
      s.tracks_per_wave <= 0 || s.tracks_per_wave * s.lanes_per_track > 64 ||
      s.block <= 0 || s.block % 64 != 0 || s.block > 1024)
    throw DeviceError(MOC_ERR_INVALID, "invalid sweep launch shape");
  shape_ = s;
  shape_forced_ = true;
}

void HipSolver::setTracks(const moc_track_data_t& td) {
  HIP_CHECK(hipSetDevice(device_));
  if (td.num_tracks <= 0 || td.num_segments <= 0 || !td.seg_offset || !td.seg_length ||
      !td.seg_fsr || !td.azim || !td.next_fwd || !td.next_bwd || !td.next_fwd_is_fwd ||
      !td.next_bwd_is_fwd || !td.bc_fwd || !td.bc_bwd || !td.total_weight || !td.sin_theta)
    throw DeviceError(MOC_ERR_INVALID, "incomplete track data");
  if (td.num_azim < 4 || td.num_azim % 4 || td.num_polar < 2 || td.num_polar % 2)
    throw DeviceError(MOC_ERR_INVALID, "invalid quadrature sizes");
  if (td.num_polar / 2 > kMaxPolarHalf)
    throw DeviceError(MOC_ERR_INVALID, "2D sweep supports at most 6 polar angles");
  if (td.num_segments >= (int64_t)1 << 31 || td.num_tracks >= (int64_t)1 << 29)
    throw DeviceError(MOC_ERR_INVALID, "track set too large for 32-bit indexing");
  const int64_t T = td.num_tracks, S = td.num_segments;
  if (td.seg_offset[0] != 0 || td.seg_offset[T] != S)
    throw DeviceError(MOC_ERR_INVALID, "seg_offset must start at 0 and end at num_segments");
  const int na2 = td.num_azim / 2;
  const int P = td.num_polar / 2;

  std::vector<int32_t> trk(4 * T), azim(T);
  std::vector<int> in_count(2 * T, 0);
  for (int64_t t = 0; t < T; t++) {
    const int64_t b = td.seg_offset[t], e = td.seg_offset[t + 1];
    if (e < b) throw DeviceError(MOC_ERR_INVALID, "seg_offset must be non-decreasing");
    if (td.azim[t] < 0 || td.azim[t] >= na2) throw DeviceError(MOC_ERR_INVALID, "azimuthal index out of range");
    azim[t] = td.azim[t];
    trk[4 * t] = (int32_t)b;
    trk[4 * t + 1] = (int32_t)(e - b);
    for (int d = 0; d < 2; d++) {
      const uint8_t bc = d == 0 ? td.bc_fwd[t] : td.bc_bwd[t];
      int32_t slot = -1;
      if (bc == 1 || bc == 2) {
        const int64_t nt = d == 0 ? td.next_fwd[t] : td.next_bwd[t];
        const int nd = d == 0 ? !td.next_fwd_is_fwd[t] : !td.next_bwd_is_fwd[t];
        if (nt < 0 || nt >= T) throw DeviceError(MOC_ERR_INVALID, "linked track out of range");
        slot = (int32_t)(nt * 2 + nd);
        in_count[slot]++;
      } else if (bc != 0) {
        throw DeviceError(MOC_ERR_INVALID, "unknown boundary condition code");
      }
      trk[4 * t + 2 + d] = slot;
    }
  }
  // Every start slot is written by at most one outgoing track-direction each
  // sweep (the reflective/periodic links form a permutation); this is what
  // lets the sweep write the ping-pong buffer without races.
  for (int64_t i = 0; i < 2 * T; i++)
    if (in_count[i] > 1) throw DeviceError(MOC_ERR_INVALID, "track links are not one-to-one");

  std::vector<float> len(S);
  std::vector<int32_t> fsr(S);
  int32_t max_fsr = -1;
  for (int64_t s = 0; s < S; s++) {
    if (!(td.seg_length[s] >= 0.0)) throw DeviceError(MOC_ERR_INVALID, "negative segment length");
    if (td.seg_fsr[s] < 0) throw DeviceError(MOC_ERR_INVALID, "negative FSR id");
    len[s] = (float)td.seg_length[s];
    fsr[s] = td.seg_fsr[s];
    max_fsr = std::max(max_fsr, fsr[s]);
  }
  // directed tracks, longest first; fwd/bwd of a track adjacent
  std::vector<int32_t> order(T);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return trk[4 * x + 1] > trk[4 * y + 1]; });
  std::vector<int32_t> work(2 * T);
  for (int64_t i = 0; i < T; i++) {
    work[2 * i] = 2 * order[i];
    work[2 * i + 1] = 2 * order[i] + 1;
  }
  std::vector<float> w((size_t)na2 * P);
  for (int a = 0; a < na2; a++)
    for (int p = 0; p < P; p++) w[(size_t)a * P + p] = (float)td.total_weight[(size_t)a * td.num_polar + p];
  for (int p = 0; p < kMaxPolarHalf; p++)
    inv_sin_[p] = p < P ? (float)(1.0 / td.sin_theta[p]) : 0.f;

  dev_alloc(&d_seg_len_, S);
  dev_alloc(&d_seg_fsr_, S);
  dev_alloc(&d_trk_, 4 * T);
  dev_alloc(&d_trk_azim_, T);
  dev_alloc(&d_work_, 2 * T);
  dev_alloc(&d_weights_, w.size());
  HIP_CHECK(hipMemcpy(d_seg_len_, len.data(), S * sizeof(float), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_seg_fsr_, fsr.data(), S * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_trk_, trk.data(), trk.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_trk_azim_, azim.data(), T * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_work_, work.data(), 2 * T * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_weights_, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));

  T_ = T;
  S_ = S;
  P_ = P;
  num_azim_ = td.num_azim;
  num_polar_ = td.num_polar;
  tracks_set_ = true;
  if (fsrs_set_ && max_fsr >= N_) {
    tracks_set_ = false;
    throw DeviceError(MOC_ERR_INVALID, "segment FSR id exceeds the number of FSRs");
  }
  max_seg_fsr_ = max_fsr;
  allocateIterationArrays();
}

void HipSolver::setMaterials(int num_materials, int num_groups, const double* sigma_t,
                             const double* sigma_s, const double* nu_sigma_f, const double* chi) {
  HIP_CHECK(hipSetDevice(device_));
  if (num_materials <= 0 || num_groups <= 0 || !sigma_t || !sigma_s || !nu_sigma_f || !chi)
    throw DeviceError(MOC_ERR_INVALID, "invalid material arrays");
  if (num_groups > 1024) throw DeviceError(MOC_ERR_INVALID, "at most 1024 energy groups");
  const int M = num_materials, G = num_groups;
  std::vector<int32_t> fis(M, 0);
  for (int m = 0; m < M; m++) {
    for (int g = 0; g < G; g++) {
      if (!(sigma_t[(size_t)m * G + g] > 0.0))
        throw DeviceError(MOC_ERR_INVALID, "sigma_t must be positive");
      if (nu_sigma_f[(size_t)m * G + g] != 0.0) fis[m] = 1;
    }
  }
  dev_alloc(&d_sigma_t_, (size_t)M * G);
  dev_alloc(&d_sigma_s_, (size_t)M * G * G);
  dev_alloc(&d_nsf_, (size_t)M * G);
  dev_alloc(&d_chi_, (size_t)M * G);
  dev_alloc(&d_mat_fissionable_, M);
  HIP_CHECK(hipMemcpy(d_sigma_t_, sigma_t, (size_t)M * G * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_sigma_s_, sigma_s, (size_t)M * G * G * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_nsf_, nu_sigma_f, (size_t)M * G * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_chi_, chi, (size_t)M * G * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_mat_fissionable_, fis.data(), M * sizeof(int32_t), hipMemcpyHostToDevice));
  mat_fissionable_ = fis;
  if (G != G_) {
    G_ = G;
    shape_forced_ = false;
    fixed_on_ = false;
    h_fixed_.clear();
  }
  M_ = M;
  materials_set_ = true;
  pickShape();
  if (fsrs_set_) {
    for (int64_t r = 0; r < N_; r++)
      if (h_fsr_mat_[r] >= M_) {
        materials_set_ = false;
        throw DeviceError(MOC_ERR_INVALID, "FSR material id exceeds the number of materials");
      }
    countFissionable();
  }
  allocateIterationArrays();
}

void HipSolver::countFissionable() {
  num_fissionable_ = 0;
  for (int64_t r = 0; r < N_; r++)
    if (h_fsr_mat_[r] < (int)mat_fissionable_.size() && mat_fissionable_[h_fsr_mat_[r]])
      num_fissionable_++;
}

void HipSolver::setFSRs(int64_t num_fsrs, const int32_t* material, const double* volume) {
  HIP_CHECK(hipSetDevice(device_));
  if (num_fsrs <= 0 || !material || !volume) throw DeviceError(MOC_ERR_INVALID, "invalid FSR arrays");
  if (num_fsrs >= (int64_t)1 << 31) throw DeviceError(MOC_ERR_INVALID, "too many FSRs");
  for (int64_t r = 0; r < num_fsrs; r++) {
    if (material[r] < 0 || (materials_set_ && material[r] >= M_))
      throw DeviceError(MOC_ERR_INVALID, "FSR material id out of range");
    if (!(volume[r] >= 0.0)) throw DeviceError(MOC_ERR_INVALID, "FSR volume must be >= 0");
  }
  if (tracks_set_ && max_seg_fsr_ >= num_fsrs)
    throw DeviceError(MOC_ERR_INVALID, "segment FSR id exceeds the number of FSRs");
  dev_alloc(&d_fsr_mat_, num_fsrs);
  dev_alloc(&d_fsr_vol_, num_fsrs);
  HIP_CHECK(hipMemcpy(d_fsr_mat_, material, num_fsrs * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_fsr_vol_, volume, num_fsrs * sizeof(double), hipMemcpyHostToDevice));
  h_fsr_mat_.assign(material, material + num_fsrs);
  if (num_fsrs != N_) {
    fixed_on_ = false;
    h_fixed_.clear();
  }
  N_ = num_fsrs;
  fsrs_set_ = true;
  countFissionable();
  allocateIterationArrays();
}

void HipSolver::allocateIterationArrays() {
  if (!(tracks_set_ && materials_set_ && fsrs_set_)) return;
  const size_t NG = (size_t)N_ * G_;
  const size_t psi_n = (size_t)2 * T_ * P_ * G_;
  dev_alloc(&d_psi_[0], psi_n);
  dev_alloc(&d_psi_[1], psi_n);
  dev_alloc(&d_tally_, NG);
  dev_alloc(&d_qs_, NG);
  dev_alloc(&d_phi_, NG);
  dev_alloc(&d_phi_old_, NG);
  dev_alloc(&d_Q_, NG);
  num_partials_ = (int)blocks_for(N_, kReduceBlock);
  dev_alloc(&d_partials_, num_partials_);
  HIP_CHECK(hipMemset(d_psi_[0], 0, psi_n * sizeof(float)));
  HIP_CHECK(hipMemset(d_psi_[1], 0, psi_n * sizeof(float)));
  HIP_CHECK(hipMemset(d_tally_, 0, NG * sizeof(float)));
  HIP_CHECK(hipMemset(d_phi_, 0, NG * sizeof(double)));
  HIP_CHECK(hipMemset(d_phi_old_, 0, NG * sizeof(double)));
  HIP_CHECK(hipMemset(d_Q_, 0, NG * sizeof(double)));
  HIP_CHECK(hipMemset(d_qs_, 0, NG * sizeof(float2)));
  if (fixed_on_ && h_fixed_.size() == NG) {
    dev_alloc(&d_fixed_, NG);
    HIP_CHECK(hipMemcpy(d_fixed_, h_fixed_.data(), NG * sizeof(double), hipMemcpyHostToDevice));
  }
  psi_cur_ = 0;
}

void HipSolver::requireReady() const {
  if (!tracks_set_) throw DeviceError(MOC_ERR_STATE, "tracks have not been set");
  if (!materials_set_) throw DeviceError(MOC_ERR_STATE, "materials have not been set");
  if (!fsrs_set_) throw DeviceError(MOC_ERR_STATE, "FSRs have not been set");
}

void HipSolver::setConvergenceThreshold(double t) {
  if (!(t > 0.0)) throw DeviceError(MOC_ERR_INVALID, "convergence threshold must be > 0");
  threshold_ = t;
}

void HipSolver::setExponential(int mode) {
  if (mode != MOC_EXP_RATIONAL && mode != MOC_EXP_HW)
    throw DeviceError(MOC_ERR_INVALID, "unknown exponential mode");
  exp_mode_ = mode;
}

void HipSolver::setFixedSourceByFSR(int64_t fsr, int group, double source) {
  requireReady();
  if (fsr < 0 || fsr >= N_) throw DeviceError(MOC_ERR_INVALID, "FSR id out of range");
  if (group < 1 || group > G_) throw DeviceError(MOC_ERR_INVALID, "group out of range (1-based)");
  const size_t NG = (size_t)N_ * G_;
  if (h_fixed_.size() != NG) h_fixed_.assign(NG, 0.0);
  h_fixed_[(size_t)fsr * G_ + group - 1] = source;
  fixed_on_ = true;
  HIP_CHECK(hipSetDevice(device_));
  if (!d_fixed_) dev_alloc(&d_fixed_, NG);
  HIP_CHECK(hipMemcpy(d_fixed_, h_fixed_.data(), NG * sizeof(double), hipMemcpyHostToDevice));
}

void HipSolver::launchSweep(const float* psi_in, float* psi_out, hipStream_t s) {
  SweepArgs a;
  a.work = d_work_;
  a.num_work = (int)(2 * T_);
  a.trk = reinterpret_cast<const int4*>(d_trk_);
  a.trk_azim = d_trk_azim_;
  a.seg_len = d_seg_len_;
  a.seg_fsr = d_seg_fsr_;
  a.qs = d_qs_;
  a.psi_in = psi_in;
  a.psi_out = psi_out;
  a.tally = d_tally_;
  a.weights = d_weights_;
  const float l2e = exp_mode_ == MOC_EXP_HW ? 1.4426950408889634f : 1.0f;
  a.pf0 = inv_sin_[0] * l2e;
  a.pf1 = inv_sin_[1] * l2e;
  a.pf2 = inv_sin_[2] * l2e;
  a.state = d_state_;
  a.G = G_;
  a.P = P_;
  a.fpt = P_ * G_;
  a.lanes_per_track = shape_.lanes_per_track;
  a.tracks_per_wave = shape_.tracks_per_wave;
  const int64_t waves = (a.num_work + a.tracks_per_wave - 1) / a.tracks_per_wave;
  const int wpb = shape_.block / 64;
  const unsigned grid = (unsigned)((waves + wpb - 1) / wpb);
  dim3 blk(shape_.block);
#define MOC_SWEEP_CASE(GPL)                                                                  \
  case GPL:                                                                                  \
    if (exp_mode_ == MOC_EXP_HW)                                                             \
      hipLaunchKernelGGL((sweep2d_kernel<GPL, MOC_EXP_HW>), dim3(grid), blk, 0, s, a);       \
    else                                                                                     \
      hipLaunchKernelGGL((sweep2d_kernel<GPL, MOC_EXP_RATIONAL>), dim3(grid), blk, 0, s, a); \
    break;
  switch (shape_.groups_per_lane) {
    MOC_SWEEP_CASE(1)
    MOC_SWEEP_CASE(2)
    MOC_SWEEP_CASE(4)
    MOC_SWEEP_CASE(9)
    MOC_SWEEP_CASE(16)
    default:
      throw DeviceError(MOC_ERR_INVALID, "unsupported groups per lane");
  }
#undef MOC_SWEEP_CASE
  HIP_CHECK(hipGetLastError());
}

namespace {
FsrArgs make_fsr_args(int64_t N, int G, int32_t* fsr_mat, double* vol, double* st, double* ss,
                      double* nsf, double* chi, int32_t* fis, double* fixed, double* phi,
                      double* old, double* Q, float2* qs, float* tally, double* partials,
                      IterState* state, bool allow_neg) {
  FsrArgs a;
  a.N = N;
  a.G = G;
  a.fsr_mat = fsr_mat;
  a.fsr_vol = vol;
  a.sigma_t = st;
  a.sigma_s = ss;
  a.nsf = nsf;
  a.chi = chi;
  a.fissionable = fis;
  a.fixed = fixed;
  a.phi = phi;
  a.phi_old = old;
  a.Q = Q;
  a.qs = qs;
  a.tally = tally;
  a.partials = partials;
  a.state = state;
  a.allow_negative = allow_neg ? 1 : 0;
  return a;
}
}  // namespace

// One source iteration; all launches on stream_, no host synchronisation.
void HipSolver::launchIteration(Mode mode, int res_type, bool recompute_sources, double tol,
                                hipEvent_t ev0, hipEvent_t ev1) {
  FsrArgs fa = make_fsr_args(N_, G_, d_fsr_mat_, d_fsr_vol_, d_sigma_t_, d_sigma_s_, d_nsf_,
                             d_chi_, d_mat_fissionable_, fixed_on_ ? d_fixed_ : nullptr, d_phi_,
                             d_phi_old_, d_Q_, d_qs_, d_tally_, d_partials_, d_state_,
                             allow_negative_);
  const unsigned nb = blocks_for(N_, kReduceBlock);
  if (ev0) HIP_CHECK(hipEventRecord(ev0, stream_));
  launchSweep(d_psi_[psi_cur_], d_psi_[psi_cur_ ^ 1], stream_);
  if (ev1) HIP_CHECK(hipEventRecord(ev1, stream_));
  psi_cur_ ^= 1;
  hipLaunchKernelGGL(post_sweep_kernel, dim3(nb), dim3(kReduceBlock), 0, stream_, fa, 1);
  if (mode == MODE_EIGEN)
    hipLaunchKernelGGL(keff_kernel, dim3(1), dim3(kReduceBlock), 0, stream_, d_partials_,
                       num_partials_, d_state_, (double)N_, 1);
  const double norm = (mode == MODE_EIGEN && res_type == MOC_FISSION_SOURCE)
                          ? (double)num_fissionable_
                          : (double)N_;
  const int rt = mode == MODE_FLUX ? MOC_SCALAR_FLUX : res_type;
  hipLaunchKernelGGL(update_kernel, dim3(nb), dim3(kReduceBlock), 0, stream_, fa,
                     mode == MODE_EIGEN ? 1 : 0, rt, 1, recompute_sources ? 1 : 0, 1);
  hipLaunchKernelGGL(residual_kernel, dim3(1), dim3(kReduceBlock), 0, stream_, d_partials_,
                     num_partials_, d_state_, rt == MOC_FISSION_SOURCE ? (double)num_fissionable_ : norm,
                     tol, mode == MODE_EIGEN ? 0 : 1);
  HIP_CHECK(hipGetLastError());
}

// Fresh state for a solve: zero angular fluxes and tally, initial guess,
// its normalisation (eigen) and the sources of iteration 0.
void HipSolver::launchInit(Mode mode, double k0) {
  const size_t NG = (size_t)N_ * G_;
  const size_t psi_n = (size_t)2 * T_ * P_ * G_;
  IterState st;
  std::memset(&st, 0, sizeof(st));
  st.k = k0;
  st.k_prev = st.k;
  st.norm_factor = 1.0;
  st.psi_scale = 1.0;
  HIP_CHECK(hipMemcpyAsync(d_state_, &st, sizeof(st), hipMemcpyHostToDevice, stream_));
  // fresh angular fluxes and tally (initializeFluxArrays, src/CPUSolver.cpp:281-380)
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(psi_n, 256)), dim3(256), 0, stream_, d_psi_[0], (int64_t)psi_n);
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(psi_n, 256)), dim3(256), 0, stream_, d_psi_[1], (int64_t)psi_n);
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)NG);
  psi_cur_ = 0;

  FsrArgs fa = make_fsr_args(N_, G_, d_fsr_mat_, d_fsr_vol_, d_sigma_t_, d_sigma_s_, d_nsf_,
                             d_chi_, d_mat_fissionable_, fixed_on_ ? d_fixed_ : nullptr, d_phi_,
                             d_phi_old_, d_Q_, d_qs_, d_tally_, d_partials_, d_state_,
                             allow_negative_);
  const unsigned nb = blocks_for(N_, kReduceBlock);
  // initial guess (computeInitialFluxGuess, src/Solver.cpp:1710-1731; computeFlux
  // starts from phi = 0, src/Solver.cpp:1383-1390)
  hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_phi_,
                     (int64_t)NG, mode == MODE_FLUX ? 0.0 : 1.0);
  if (mode == MODE_EIGEN) {
    hipLaunchKernelGGL(post_sweep_kernel, dim3(nb), dim3(kReduceBlock), 0, stream_, fa, 0);
    hipLaunchKernelGGL(keff_kernel, dim3(1), dim3(kReduceBlock), 0, stream_, d_partials_,
                       num_partials_, d_state_, (double)N_, 0);
  }
  // normalise (eigen), store old flux, sources of iteration 0
  hipLaunchKernelGGL(update_kernel, dim3(nb), dim3(kReduceBlock), 0, stream_, fa,
                     mode == MODE_EIGEN ? 1 : 0, 0, 0, 1, 0);
  // start fluxes are zero: the scale of the first sweep is irrelevant
  HIP_CHECK(hipGetLastError());

  state_valid_ = true;
}

void HipSolver::runSolve(Mode mode, int max_iters, int res_type, double k_fixed) {
  requireReady();
  HIP_CHECK(hipSetDevice(device_));
  if (max_iters <= 0) throw DeviceError(MOC_ERR_INVALID, "max_iters must be > 0");
  if (res_type < MOC_SCALAR_FLUX || res_type > MOC_TOTAL_SOURCE)
    throw DeviceError(MOC_ERR_INVALID, "unknown residual type");
  if (mode != MODE_FLUX && res_type == MOC_FISSION_SOURCE && num_fissionable_ == 0)
    throw DeviceError(MOC_ERR_INVALID,
                      "cannot compute a FISSION_SOURCE residual without fissionable FSRs");
  if (mode == MODE_EIGEN && num_fissionable_ == 0)
    throw DeviceError(MOC_ERR_INVALID, "eigenvalue problem without fissionable material");
  const auto t0 = std::chrono::steady_clock::now();
  launchInit(mode, mode == MODE_SOURCE ? k_fixed : 1.0);

  const bool recompute = mode != MODE_FLUX;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  int done = 0;
  int batch = 4;
  float sweep_ms_total = 0.f;
  while (done < max_iters) {
    const int nb_it = std::min(batch, max_iters - done);
    HIP_CHECK(hipEventRecord(e0, stream_));
    for (int i = 0; i < nb_it; i++) launchIteration(mode, res_type, recompute, threshold_);
    HIP_CHECK(hipEventRecord(e1, stream_));
    HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    sweep_ms_total += ms;
    done += nb_it;
    if (h_state_->converged) break;
    batch = std::min(batch * 2, 32);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  keff_ = h_state_->k;
  num_iterations_ = h_state_->converged ? h_state_->reported_iters : max_iters;
  total_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // device time of the iteration batches (sweeps plus the per-FSR steps)
  sweep_seconds_ = sweep_ms_total * 1e-3;
  integrations_ = 2LL * G_ * P_ * S_ * (int64_t)h_state_->iteration;
}

void HipSolver::computeEigenvalue(int max_iters, int res_type) {
  runSolve(MODE_EIGEN, max_iters, res_type, 1.0);
}

void HipSolver::computeFlux(int max_iters, bool only_fixed_source) {
  (void)only_fixed_source;  // the flux arrays are always re-initialised
  if (!fixed_on_) throw DeviceError(MOC_ERR_STATE, "computeFlux requires fixed sources");
  runSolve(MODE_FLUX, max_iters, MOC_SCALAR_FLUX, 1.0);
}

void HipSolver::computeSource(int max_iters, double k_eff, int res_type) {
  if (!(k_eff > 0.0)) throw DeviceError(MOC_ERR_INVALID, "k_eff must be positive");
  runSolve(MODE_SOURCE, max_iters, res_type, k_eff);
}

double HipSolver::getFlux(int64_t fsr, int group) {
  requireReady();
  if (fsr < 0 || fsr >= N_) throw DeviceError(MOC_ERR_INVALID, "FSR id out of range");
  if (group < 1 || group > G_) throw DeviceError(MOC_ERR_INVALID, "group out of range (1-based)");
  double v = 0;
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipMemcpyAsync(&v, d_phi_ + (size_t)fsr * G_ + group - 1, sizeof(double),
                           hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  return v;
}

void HipSolver::getFluxes(double* out, int64_t n) {
  requireReady();
  if (!out || n != N_ * G_) throw DeviceError(MOC_ERR_INVALID, "getFluxes size must be N*G");
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipMemcpyAsync(out, d_phi_, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

void HipSolver::setFluxes(const double* in, int64_t n) {
  requireReady();
  if (!in || n != N_ * G_) throw DeviceError(MOC_ERR_INVALID, "setFluxes size must be N*G");
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipMemcpyAsync(d_phi_, in, n * sizeof(double), hipMemcpyHostToDevice, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

void HipSolver::computeFSRFissionRates(double* out, int64_t n) {
  requireReady();
  if (!out || n != N_) throw DeviceError(MOC_ERR_INVALID, "fission rates size must be N");
  HIP_CHECK(hipSetDevice(device_));
  double* d_out = nullptr;
  dev_alloc(&d_out, N_);
  hipLaunchKernelGGL(fission_rates_kernel, dim3(blocks_for(N_, 256)), dim3(256), 0, stream_,
                     d_phi_, d_fsr_mat_, d_nsf_, d_fsr_vol_, d_out, N_, G_);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(out, d_out, N_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  dev_free(d_out);
}

void HipSolver::timing(double* sweep_s, double* total_s, int64_t* integrations) const {
  if (sweep_s) *sweep_s = sweep_seconds_;
  if (total_s) *total_s = total_seconds_;
  if (integrations) *integrations = integrations_;
}

void HipSolver::transportSweepOnce(const double* Q, float* psi, double* tally) {
  requireReady();
  if (!Q || !psi || !tally) throw DeviceError(MOC_ERR_INVALID, "null argument");
  HIP_CHECK(hipSetDevice(device_));
  const size_t NG = (size_t)N_ * G_;
  const size_t psi_n = (size_t)2 * T_ * P_ * G_;
  IterState st;
  std::memset(&st, 0, sizeof(st));
  st.psi_scale = 1.0;
  st.k = 1.0;
  HIP_CHECK(hipMemcpyAsync(d_state_, &st, sizeof(st), hipMemcpyHostToDevice, stream_));
  HIP_CHECK(hipMemcpyAsync(d_Q_, Q, NG * sizeof(double), hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(make_qs_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_Q_,
                     d_fsr_mat_, d_sigma_t_, d_qs_, N_, G_);
  HIP_CHECK(hipMemcpyAsync(d_psi_[0], psi, psi_n * sizeof(float), hipMemcpyHostToDevice, stream_));
  // slots that no track writes keep their value (vacuum inflow), as in
  // CPUSolver::_start_flux
  HIP_CHECK(hipMemcpyAsync(d_psi_[1], psi, psi_n * sizeof(float), hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)NG);
  launchSweep(d_psi_[0], d_psi_[1], stream_);
  std::vector<float> t32(NG);
  HIP_CHECK(hipMemcpyAsync(t32.data(), d_tally_, NG * sizeof(float), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipMemcpyAsync(psi, d_psi_[1], psi_n * sizeof(float), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  for (size_t i = 0; i < NG; i++) tally[i] = t32[i];
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)NG);
  HIP_CHECK(hipStreamSynchronize(stream_));
  psi_cur_ = 0;
}

double HipSolver::runIterations(int n) {
  requireReady();
  if (n <= 0) throw DeviceError(MOC_ERR_INVALID, "number of iterations must be > 0");
  if (num_fissionable_ == 0) throw DeviceError(MOC_ERR_INVALID, "no fissionable material");
  HIP_CHECK(hipSetDevice(device_));
  if (!state_valid_) launchInit(MODE_EIGEN, 1.0);
  // keep iterating from the current state; the stopping rule is disabled
  // (tolerance 0) and a converged flag from an earlier solve is cleared
  HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  h_state_->converged = 0;
  HIP_CHECK(hipMemcpyAsync(d_state_, h_state_, sizeof(IterState), hipMemcpyHostToDevice, stream_));
  std::vector<hipEvent_t> ev(2 * n);
  for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
  for (int i = 0; i < n; i++)
    launchIteration(MODE_EIGEN, MOC_FISSION_SOURCE, true, 0.0, ev[2 * i], ev[2 * i + 1]);
  HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  double total_ms = 0;
  for (int i = 0; i < n; i++) {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
    total_ms += ms;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  keff_ = h_state_->k;
  return total_ms * 1e-3 / n;
}

double HipSolver::timeSweeps(int repeats) {
  requireReady();
  if (repeats <= 0) throw DeviceError(MOC_ERR_INVALID, "repeats must be > 0");
  HIP_CHECK(hipSetDevice(device_));
  IterState st;
  std::memset(&st, 0, sizeof(st));
  st.psi_scale = 1.0;
  st.k = keff_;
  HIP_CHECK(hipMemcpyAsync(d_state_, &st, sizeof(st), hipMemcpyHostToDevice, stream_));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  launchSweep(d_psi_[psi_cur_], d_psi_[psi_cur_ ^ 1], stream_);  // warm-up
  HIP_CHECK(hipEventRecord(e0, stream_));
  for (int i = 0; i < repeats; i++) {
    launchSweep(d_psi_[psi_cur_], d_psi_[psi_cur_ ^ 1], stream_);
    psi_cur_ ^= 1;
  }
  HIP_CHECK(hipEventRecord(e1, stream_));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const size_t NG = (size_t)N_ * G_;
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)NG);
  HIP_CHECK(hipStreamSynchronize(stream_));
  return ms * 1e-3 / repeats;
}

}  // namespace moc
