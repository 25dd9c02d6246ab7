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
#include <type_traits>

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
constexpr int kSweepBlockMax = 256;  // __launch_bounds__ of the sweep kernel

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
  const int32_t* work;      // [ntiles*tpw] directed track ids (t*2+d) or -1, longest first
  int ntiles;               // wave tiles of tpw directed tracks
  int tpw;                  // directed tracks per tile (= per wave)
  const int2* tile;         // [ntiles] (first row, rows) of the interleaved streams
  const int4* trk;          // seg_begin, seg_count, out slot fwd, out slot bwd
  const int32_t* trk_azim;
  const float* seg_len;     // tile-interleaved: [row][tpw], traversal order, 0-padded
  const int32_t* seg_key;   // tile-interleaved: [row][tpw], FSR | material << 24
  const float2* qs;         // [N*G] (Q/sigma_t, sigma_t)
  const float* psi_in;      // [2T*fpt]
  float* psi_out;           // [2T*fpt]
  float* tally;             // global path: [tally_copies][N*G]
  int tally_copies;         // global path: private copies (block b -> copy b % copies)
  size_t tally_stride;      // N*G
  const float* weights;     // [num_azim/2][P]
  float pf0, pf1, pf2;      // polar factors: 1/sin(theta_p) (times log2 e for HW exp)
  const IterState* state;
  unsigned* psi_max_bits;   // &state->psi_max_bits
  int G, P, fpt, lanes_per_track, tracks_per_wave;
  int N;
};

// Segment streams carry a packed key: FSR id in the low 24 bits, material in
// the high 8 bits (N < 2^24, M <= 256 are checked on the host).
constexpr int kFsrBits = 24;
constexpr int kFsrMask = (1 << kFsrBits) - 1;

// Source policies: (Q/sigma_t, sigma_t) of group g of the segment's FSR.
struct GatherSource {  // global FP32 table qs[N*G]
  static constexpr bool kLocal = false;
  const float2* qs;
  int G;
  __device__ float2 load(int key, int g, int /*loc*/) const {
    return qs[(size_t)(key & kFsrMask) * G + g];
  }
};
struct LdsSource {     // LDS slices: Q/sigma_t [N][GB] and sigma_t by material [M][GB]
  static constexpr bool kLocal = true;
  const float* q;
  const float* sig;
  int GB;
  __device__ float2 load(int key, int /*g*/, int loc) const {
    return make_float2(q[(key & kFsrMask) * GB + loc],
                       sig[(int)((unsigned)key >> kFsrBits) * GB + loc]);
  }
};
struct SubsetSource {  // global subset-major Q/sigma_t [N][GB] + LDS sigma_t by material
  static constexpr bool kLocal = false;
  const float* q;
  const float* sig;
  int GB;
  __device__ float2 load(int key, int /*g*/, int loc) const {
    return make_float2(q[(size_t)(key & kFsrMask) * GB + loc],
                       sig[(int)((unsigned)key >> kFsrBits) * GB + loc]);
  }
};

// Tally policies: where a lane's weighted delta-psi of one segment goes.
struct GlobalTally {   // device-scope float atomics into this block's copy
  float* t;
  int G;
  __device__ void add(int fsr, int g, int /*loc*/, float v) const {
    unsafeAtomicAdd(&t[(size_t)fsr * G + g], v);
  }
};
// Fixed-point LDS tally: the block's group slice [N][GB] of int64 sums of
// contribution / V_fsr in units of 1/scale.  Integer LDS atomics run at full
// rate (float LDS atomics do not), and integer sums are exact and
// order-independent, so the tally is bitwise reproducible.
struct LdsFixedTally {
  unsigned long long* t;
  const float* inv_vol;   // [N] 1/V (0 for zero-volume FSRs)
  double scale;
  int GB;
  __device__ void add(int fsr, int /*g*/, int loc, float v) const {
    const double x = (double)(v * inv_vol[fsr]) * scale;
    atomicAdd(&t[fsr * GB + loc], (unsigned long long)(long long)x);
  }
};

// Integrates directed track dt (tile slot `slot`) for the lane's groups
// g_begin + gl + k*lpt (k < GPL, valid while < g_begin + g_count), taking
// sources from `src`, tallying through `tally`, and writes the outgoing
// angular flux to the linked track's start slot.
template <int GPL, int EXPM, int NP, class Src, class Tally>
__device__ __forceinline__ void sweep_track(const SweepArgs& a, int tile, int slot, int dt, int gl,
                                            int lpt, int g_begin, int g_count, const Src& src,
                                            const Tally& tally) {
  const int t = dt >> 1;
  const int dir = dt & 1;
  const int4 tr = a.trk[t];
  const int2 ti = a.tile[tile];
  const int azim = a.trk_azim[t];
  const int G = a.G, fpt = a.fpt;
  constexpr int P = NP;  // polar half-angles, compile time: the segment body is straight-line
  const float scale = (float)a.state->psi_scale;
  const float pf[kMaxPolarHalf] = {a.pf0, a.pf1, a.pf2};
  float w[P];
#pragma unroll
  for (int p = 0; p < P; p++) w[p] = a.weights[azim * P + p];

  // Lanes past the last group work on a clamped group and add exactly zero,
  // so the loop below has no per-lane branches.
  int g[GPL], loc[GPL];
  bool ok[GPL];
  float psi[GPL][P];
  const float* pin = a.psi_in + (size_t)dt * fpt;
#pragma unroll
  for (int k = 0; k < GPL; k++) {
    const int l = gl + k * lpt;
    ok[k] = l < g_count;
    loc[k] = ok[k] ? l : g_count - 1;
    g[k] = g_begin + loc[k];
#pragma unroll
    for (int p = 0; p < P; p++) psi[k][p] = ok[k] ? pin[p * G + g[k]] * scale : 0.f;
  }

  // Deep software pipeline.  The tile's rows are processed in chunks of D
  // rows held in three rotating register sets, so the stream loads of a chunk
  // are issued two chunks (2*D segments) before it is integrated.  Sources
  // that live in LDS are read one segment ahead; sources gathered from global
  // memory are fetched for the whole next chunk before the current one is
  // integrated.  The waits the compiler inserts therefore always leave the
  // most recent loads (and the tally's memory operations) in flight.  The
  // streams are tile-interleaved ([row][tpw], one coalesced line per row and
  // wave) and padded to the tile's longest track with length-0 segments,
  // which integrate as exact no-ops (tau = 0 -> attenuation 0 -> delta 0);
  // rows past the tile are clamped to its last row, so every lane runs the
  // same trip count with unconditional loads.
  constexpr int D = GPL >= 4 ? 2 : 4;
  const int n = ti.y;
  const int last = n - 1;
  const size_t row0 = (size_t)ti.x * a.tpw + slot;
  auto load_chunk = [&](int c, float* L, int* K) {
#pragma unroll
    for (int j = 0; j < D; j++) {
      const int r = c * D + j;
      const size_t idx = row0 + (size_t)min(r, last) * a.tpw;
      L[j] = a.seg_len[idx] * (r < n ? 1.f : 0.f);
      K[j] = a.seg_key[idx];
    }
  };
  auto fetch = [&](int key, float2* q) {
#pragma unroll
    for (int k = 0; k < GPL; k++) q[k] = src.load(key, g[k], loc[k]);
  };
  auto integrate = [&](float len, int key, const float2* q) {
    const int fsr = key & kFsrMask;
#pragma unroll
    for (int k = 0; k < GPL; k++) {
      const float tau = q[k].y * len;
      float acc = 0.f;
#pragma unroll
      for (int p = 0; p < P; p++) {
        const float x = tau * pf[p];
        const float A = EXPM == MOC_EXP_RATIONAL ? attenuation_rational(x) : attenuation_hw(x);
        const float d = (psi[k][p] - q[k].x) * A;
        psi[k][p] -= d;
        acc = fmaf(w[p], d, acc);
      }
      if (!ok[k]) acc = 0.f;
      tally.add(fsr, g[k], loc[k], acc);
    }
  };
  // Integrates one chunk.  Local sources: read one segment ahead.  Global
  // sources: `cur` holds this chunk's sources (fetched one chunk earlier) and
  // the next chunk's sources are fetched into `nxt` first.
  auto process = [&](const float* L, const int* K, const int* Knext, float2 (*cur)[GPL],
                     float2 (*nxt)[GPL]) {
    if constexpr (Src::kLocal) {
      float2 qa[GPL], qb[GPL];
      fetch(K[0], qa);
#pragma unroll
      for (int j = 0; j < D; j++) {
        if (j + 1 < D) fetch(K[j + 1], (j & 1) ? qa : qb);
        integrate(L[j], K[j], (j & 1) ? qb : qa);
      }
      (void)Knext;
      (void)cur;
      (void)nxt;
    } else {
#pragma unroll
      for (int j = 0; j < D; j++) fetch(Knext[j], nxt[j]);
#pragma unroll
      for (int j = 0; j < D; j++) integrate(L[j], K[j], cur[j]);
    }
  };

  if (n > 0) {
    const int nch = (n + D - 1) / D;
    // three rotating sets: chunk c lives in set c % 3 (streams and sources)
    float L0[D], L1[D], L2[D];
    int K0[D], K1[D], K2[D];
    float2 Q0[Src::kLocal ? 1 : D][GPL], Q1[Src::kLocal ? 1 : D][GPL], Q2[Src::kLocal ? 1 : D][GPL];
    load_chunk(0, L0, K0);
    load_chunk(1, L1, K1);
    load_chunk(2, L2, K2);
    if constexpr (!Src::kLocal) {
#pragma unroll
      for (int j = 0; j < D; j++) fetch(K0[j], Q0[j]);
    }
    for (int c = 0; c < nch; c += 3) {
      process(L0, K0, K1, Q0, Q1);
      load_chunk(c + 3, L0, K0);
      if (c + 1 < nch) {
        process(L1, K1, K2, Q1, Q2);
        load_chunk(c + 4, L1, K1);
      }
      if (c + 2 < nch) {
        process(L2, K2, K0, Q2, Q0);
        load_chunk(c + 5, L2, K2);
      }
    }
  }

  const int out = dir ? tr.w : tr.z;
  float amax = 0.f;
  if (out >= 0) {
    float* po = a.psi_out + (size_t)out * fpt;
#pragma unroll
    for (int k = 0; k < GPL; k++)
#pragma unroll
      for (int p = 0; p < P; p++)
        if (ok[k]) {
          po[p * G + g[k]] = psi[k][p];
          amax = fmaxf(amax, fabsf(psi[k][p]));
        }
  }
  // running max of |start psi| for the next sweep's fixed-point scale
  // (non-negative floats order like their bit patterns)
  if (amax > 0.f) atomicMax(&a.psi_max_bits[0], __float_as_uint(amax));
}

// Global-atomic sweep: one wave per tile of tpw directed tracks,
// lanes_per_track lanes per track, GPL groups per lane.
template <int GPL, int EXPM, int NP>
__global__ __launch_bounds__(kSweepBlockMax) void sweep2d_kernel(SweepArgs a) {
  const int lane = threadIdx.x & 63;
  const int tile = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int slot = lane / a.lanes_per_track;
  const int gl = lane - slot * a.lanes_per_track;
  if (slot >= a.tpw || tile >= a.ntiles) return;
  if (a.state->converged) return;
  const int dt = a.work[tile * a.tpw + slot];
  if (dt < 0) return;
  GatherSource src{a.qs, a.G};
  GlobalTally tally{a.tally + (size_t)(blockIdx.x % a.tally_copies) * a.tally_stride, a.G};
  sweep_track<GPL, EXPM, NP>(a, tile, slot, dt, gl, a.lanes_per_track, 0, a.G, src, tally);
}

// LDS-privatised sweep.  Block b owns group subset b % nsub (GB groups, one
// lane per group) and block slot b / nsub.  LDS holds the FP32 tally slice
// [N][GB] of the subset, sigma_t of the subset by material [M][GB] and, when
// QLDS, the subset's Q/sigma_t slice [N][GB] (otherwise gathered from the
// subset-major global table).  The block processes the tiles
// slot*wpb + wave, + blocks_per_sub*wpb, ... (longest first, so the strided
// assignment balances) and finally writes its tally slice to its own partial
// buffer; reduce_partials_kernel sums the partials in a fixed order.
struct LdsParams {
  unsigned long long* partials;  // [nsub][blocks_per_sub][N*GB] fixed-point sums
  const float* qsub;      // [nsub][N][GB] Q/sigma_t, subset-major
  const float* inv_vol;   // [N]
  const double* sigma_t;  // [M][G]
  int M;
  int GB;                 // groups per subset (= lanes per track; tpw = 64 / GB)
  int nsub;               // ceil(G / GB)
  int blocks_per_sub;
};

template <int EXPM, int QLDS, int NP>
__global__ __launch_bounds__(1024) void sweep2d_lds_kernel(SweepArgs a, LdsParams L) {
  extern __shared__ unsigned long long lds64[];
  if (a.state->converged) return;  // uniform: read by every thread of the block
  const int sub = blockIdx.x % L.nsub;
  const int bslot = blockIdx.x / L.nsub;
  const int g0 = sub * L.GB;
  const int gb = min(L.GB, a.G - g0);
  const int nl = a.N * L.GB;
  // LDS: int64 tally [N*GB] | float 1/V [N] | sigma_t [M*GB] | Q/sigma_t [N*GB] (QLDS)
  unsigned long long* lds_tally = lds64;
  float* lds_ivol = reinterpret_cast<float*>(lds64 + nl);
  float* lds_sig = lds_ivol + a.N;
  float* lds_q = lds_sig + L.M * L.GB;
  for (int i = threadIdx.x; i < nl; i += blockDim.x) lds_tally[i] = 0ull;
  for (int i = threadIdx.x; i < a.N; i += blockDim.x) lds_ivol[i] = L.inv_vol[i];
  for (int i = threadIdx.x; i < L.M * L.GB; i += blockDim.x) {
    const int m = i / L.GB, l = i - m * L.GB;
    lds_sig[i] = (float)L.sigma_t[(size_t)m * a.G + g0 + min(l, gb - 1)];
  }
  const float* qg = L.qsub + (size_t)sub * nl;
  if (QLDS)
    for (int i = threadIdx.x; i < nl; i += blockDim.x) lds_q[i] = qg[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int slot = lane / L.GB;
  const int gl = lane - slot * L.GB;
  if (slot < a.tpw) {
    LdsFixedTally tally{lds_tally, lds_ivol, a.state->tally_scale, L.GB};
    for (int tile = bslot * wpb + wib; tile < a.ntiles; tile += L.blocks_per_sub * wpb) {
      const int dt = a.work[tile * a.tpw + slot];
      if (dt < 0) continue;
      if (QLDS) {
        LdsSource src{lds_q, lds_sig, L.GB};
        sweep_track<1, EXPM, NP>(a, tile, slot, dt, gl, L.GB, g0, gb, src, tally);
      } else {
        SubsetSource src{qg, lds_sig, L.GB};
        sweep_track<1, EXPM, NP>(a, tile, slot, dt, gl, L.GB, g0, gb, src, tally);
      }
    }
  }
  __syncthreads();
  unsigned long long* out = L.partials + ((size_t)sub * L.blocks_per_sub + bslot) * nl;
  for (int i = threadIdx.x; i < nl; i += blockDim.x) out[i] = lds_tally[i];
}

// tally64[r*G + g] = sum over the subset's blocks of partials (fixed order).
__global__ void reduce_partials_kernel(const unsigned long long* partials, double* tally64,
                                       int64_t N, int G, int GB, int nsub, int blocks_per_sub,
                                       const IterState* state) {
  if (state->converged) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // over nsub * N * GB
  const int64_t nl = N * GB;
  if (i >= nl * nsub) return;
  const int sub = (int)(i / nl);
  const int64_t j = i - sub * nl;
  const int64_t r = j / GB;
  const int l = (int)(j - r * GB);
  const int g = sub * GB + l;
  if (g >= G) return;
  const unsigned long long* p = partials + (size_t)sub * blocks_per_sub * nl + j;
  unsigned long long t = 0ull;
  for (int b = 0; b < blocks_per_sub; b++) t += p[(size_t)b * nl];  // exact (two's complement)
  // tally64 = sum of contribution / V (the volume division of
  // addSourceToScalarFlux is folded into the fixed-point tally)
  tally64[r * G + g] = (double)(long long)t / state->tally_scale;
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
  float2* qs;                 // [N*G] (global-atomic sweep), or null
  float* qsub;                // [nsub][N][GB] Q/sigma_t (LDS sweep), or null
  int qsub_groups;            // GB
  float* tally;               // [tally_copies][N*G] (global-atomic sweep)
  int tally_copies;
  const double* tally64;      // [N*G] (LDS sweep), or null
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
        if (a.tally64) {
          // fixed-point tally already holds sum(contribution) / V
          f = a.tally64[i] / st[g] + kFourPi * a.Q[i] / st[g];
        } else {
          // sum of the private tally copies in a fixed order
          double t = 0.0;
          const size_t stride = (size_t)a.N * G;
          for (int c = 0; c < a.tally_copies; c++) {
            t += (double)a.tally[c * stride + i];
            a.tally[c * stride + i] = 0.f;
          }
          f = t / (st[g] * vol) + kFourPi * a.Q[i] / st[g];
        }
        if (f < 0.0 && !a.allow_negative) {
          f = kFluxEpsilon;
          neg++;
        }
        a.phi[i] = f;
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
  double qmax = 0.0;
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
        qmax = fmax(qmax, fabs(q / st[gd]));
        if (a.qs) a.qs[i] = make_float2((float)(q / st[gd]), (float)st[gd]);
        if (a.qsub) {
          const int sb = gd / a.qsub_groups, l = gd - sb * a.qsub_groups;
          a.qsub[((size_t)sb * a.N + r) * a.qsub_groups + l] = (float)(q / st[gd]);
        }
      }
    }
    (void)old_fis;
  }
  const double s = block_sum(res);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = s;
  if (negsrc) atomicAdd((unsigned long long*)&a.state->negative_sources, (unsigned long long)negsrc);
  if (qmax > 0.0) {
    // round up so the float bound is not below the double value
    const float qf = __uint_as_float(__float_as_uint((float)qmax) + 1u);
    atomicMax(&a.state->qmax_bits, __float_as_uint(qf));
  }
}

// Fixed-point scale of the next sweep's LDS tally.  Over one sweep,
// sum |w delta| / V over an FSR's segments <= 4 pi sigma_t (|psi|max +
// |Q/sigma_t|max), because sum over its segments (both directions, all polar
// angles) of w l / sin(theta) = 4 pi V and |delta| <= (1 - e^-tau)|psi - q|
// <= tau / sin(theta) |psi - q|, with psi a convex mix of its inflow and the
// sources.  The scale keeps that bound at 2^59 (int64 range 2^63).
__device__ void set_tally_scale(IterState* st, double sigma_t_max, int take_qmax) {
  if (take_qmax) {
    st->qmax = __uint_as_float(st->qmax_bits);
    st->qmax_bits = 0u;
  }
  st->psi_max = __uint_as_float(st->psi_max_bits);
  st->psi_max_bits = 0u;
  const double bound =
      2.0 * kFourPi * sigma_t_max * ((double)st->psi_max * st->psi_scale + (double)st->qmax);
  st->tally_scale = bound > 1e-200 ? 0x1p60 / bound : 0x1p60;
}

__global__ void tally_scale_kernel(IterState* state, double sigma_t_max, int take_qmax) {
  if (threadIdx.x == 0 && blockIdx.x == 0) set_tally_scale(state, sigma_t_max, take_qmax);
}

// mode 0 eigen: stop when res < tol && |int(1e5 dk)| < 1 (src/Solver.cpp:1643-1681)
// mode 1/2 flux/source: stop when i > 1 && res < tol, _num_iterations = i.
__global__ __launch_bounds__(kReduceBlock) void residual_kernel(const double* partials, int nparts,
                                                               IterState* state, double norm,
                                                               double tol, int mode,
                                                               double sigma_t_max, int take_qmax) {
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
    set_tally_scale(state, sigma_t_max, take_qmax);
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
                               float2* qs, float* qsub, int GB, int64_t N, int G) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N * G) {
    const int64_t r = i / G;
    const int g = (int)(i - r * G);
    const double st = sigma_t[(size_t)fsr_mat[r] * G + g];
    if (qs) qs[i] = make_float2((float)(Q[i] / st), (float)st);
    if (qsub) {
      const int sb = g / GB, l = g - sb * GB;
      qsub[((size_t)sb * N + r) * GB + l] = (float)(Q[i] / st);
    }
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
  int v = 0;
  HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device_));
  num_cus_ = v;
  HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, device_));
  lds_max_ = (size_t)v;
  if (lds_max_ == 0) {
    HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device_));
    lds_max_ = (size_t)v;
  }
  HIP_CHECK(hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerMultiprocessor, device_));
  lds_per_cu_ = v > 0 ? (size_t)v : lds_max_;
}

HipSolver::~HipSolver() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  dev_free(d_seg_len_);
  dev_free(d_seg_key_);
  dev_free(d_work_);
  dev_free(d_tile_);
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
  dev_free(d_inv_vol_);
  dev_free(d_partials_);
  dev_free(d_lds_partials_);
  dev_free(d_qsub_);
  dev_free(d_tally64_);
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
  // LDS-privatised tally when at least one group's slice fits in LDS: the
  // largest subset that fits, evened out over the subsets.
  // LDS-privatised tally when a group's slice fits in LDS: with the source
  // slice resident too (lds_q = 1) if that fits, else gathered.  The largest
  // subset that fits, evened out over the subsets.
  s.lds_groups = 0;
  s.lds_q = 0;
  if (N_ > 0 && lds_max_ > 0) {
    for (int q = 1; q >= 0 && s.lds_groups == 0; q--) {
      int gb = 0;
      while (gb < std::min(G_, 64) && ldsBytes(gb + 1, q) <= lds_max_) gb++;
      if (gb < 1) continue;
      const int nsub = (G_ + gb - 1) / gb;
      gb = (G_ + nsub - 1) / nsub;
      s.lds_groups = gb;
      s.lds_q = q;
      s.lds_block = 1024;
      // about two blocks per CU: the strided tile assignment then balances
      // the longest-first tiles across blocks (measured on the 17x17 lattice)
      s.lds_blocks_per_sub = std::max(1, (2 * num_cus_ + nsub - 1) / nsub);
    }
  }
  shape_ = s;
}

// LDS of the LDS sweep: int64 tally [N*gb] | 1/V [N] | sigma_t [M*gb] | Q/sigma_t [N*gb]
size_t HipSolver::ldsBytes(int gb, int with_q) const {
  return (size_t)N_ * gb * sizeof(uint64_t) + (size_t)N_ * sizeof(float) +
         (size_t)std::max(M_, 1) * gb * sizeof(float) +
         (with_q ? (size_t)N_ * gb * sizeof(float) : 0);
}

void HipSolver::setShape(const LaunchShape& s) {
  static const int kGpl[] = {1, 2, 4, 9, 16};
  bool okg = false;
  for (int gpl : kGpl) okg |= gpl == s.groups_per_lane;
  if (!okg || s.lanes_per_track <= 0 || s.lanes_per_track > 64 ||
      s.lanes_per_track * s.groups_per_lane < G_ ||
      s.tracks_per_wave <= 0 || s.tracks_per_wave * s.lanes_per_track > 64 ||
      s.block <= 0 || s.block % 64 != 0 || s.block > kSweepBlockMax || s.tally_copies < 1 ||
      s.tally_copies > 64)
    throw DeviceError(MOC_ERR_INVALID, "invalid sweep launch shape");
  if (s.lds_groups < 0 || s.lds_groups > 64 || s.lds_groups > G_ || s.lds_q < 0 || s.lds_q > 1 ||
      (s.lds_groups > 0 &&
       (ldsBytes(s.lds_groups, s.lds_q) > lds_max_ || s.lds_block <= 0 ||
        s.lds_block % 64 != 0 || s.lds_block > 1024 || s.lds_blocks_per_sub <= 0)))
    throw DeviceError(MOC_ERR_INVALID, "invalid LDS sweep shape");
  const bool realloc = s.tally_copies != shape_.tally_copies || s.lds_groups != shape_.lds_groups ||
                       s.lds_blocks_per_sub != shape_.lds_blocks_per_sub;
  shape_ = s;
  shape_forced_ = true;
  if (realloc) allocateIterationArrays();
  buildTileLayout(sweepTracksPerWave());
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
  h_order_.resize(2 * T);
  for (int64_t i = 0; i < T; i++) {
    h_order_[2 * i] = 2 * order[i];
    h_order_[2 * i + 1] = 2 * order[i] + 1;
  }
  std::vector<float> w((size_t)na2 * P);
  for (int a = 0; a < na2; a++)
    for (int p = 0; p < P; p++) w[(size_t)a * P + p] = (float)td.total_weight[(size_t)a * td.num_polar + p];
  for (int p = 0; p < kMaxPolarHalf; p++)
    inv_sin_[p] = p < P ? (float)(1.0 / td.sin_theta[p]) : 0.f;

  dev_alloc(&d_trk_, 4 * T);
  dev_alloc(&d_trk_azim_, T);
  dev_alloc(&d_weights_, w.size());
  HIP_CHECK(hipMemcpy(d_trk_, trk.data(), trk.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_trk_azim_, azim.data(), T * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_weights_, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
  h_len_.swap(len);
  h_fsr_.swap(fsr);
  h_trk_.swap(trk);
  layout_tpw_ = 0;  // streams are laid out for the launch shape once it is known

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

// Tile-interleaved segment streams for tiles of `tpw` directed tracks taken
// in work order (longest first): row i of a tile holds the i-th segment, in
// traversal order, of each of its tracks; shorter tracks are padded with
// length-0 segments that repeat their last FSR.
void HipSolver::buildTileLayout(int tpw) {
  if (tpw == layout_tpw_) return;
  if (!(tracks_set_ && fsrs_set_)) return;
  if (N_ >= ((int64_t)1 << kFsrBits) || M_ > 256)
    throw DeviceError(MOC_ERR_INVALID, "at most 2^24 FSRs and 256 materials");
  const int64_t W = 2 * T_;
  const int64_t ntiles = (W + tpw - 1) / tpw;
  if (ntiles >= (int64_t)1 << 31) throw DeviceError(MOC_ERR_INVALID, "too many tiles");
  std::vector<int32_t> work(ntiles * tpw, -1);
  std::vector<int32_t> tile(2 * ntiles);
  int64_t rows = 0;
  for (int64_t tl = 0; tl < ntiles; tl++) {
    int32_t nmax = 0;
    for (int sl = 0; sl < tpw; sl++) {
      const int64_t w = tl * tpw + sl;
      if (w >= W) break;
      const int32_t dt = h_order_[w];
      work[tl * tpw + sl] = dt;
      nmax = std::max(nmax, h_trk_[4 * (dt >> 1) + 1]);
    }
    tile[2 * tl] = (int32_t)rows;
    tile[2 * tl + 1] = nmax;
    rows += nmax;
    if (rows >= (int64_t)1 << 31) throw DeviceError(MOC_ERR_INVALID, "segment streams too large");
  }
  const size_t total = (size_t)rows * tpw;
  std::vector<float> ilen(total, 0.f);
  std::vector<int32_t> ifsr(total, 0);
  auto key = [&](int32_t f) {
    return (int32_t)((uint32_t)f | ((uint32_t)h_fsr_mat_[f] << kFsrBits));
  };
  for (int64_t tl = 0; tl < ntiles; tl++) {
    const int64_t r0 = tile[2 * tl];
    const int32_t nr = tile[2 * tl + 1];
    for (int sl = 0; sl < tpw; sl++) {
      const int32_t dt = work[tl * tpw + sl];
      if (dt < 0) continue;
      const int32_t t = dt >> 1, dir = dt & 1;
      const int32_t b = h_trk_[4 * t], n = h_trk_[4 * t + 1];
      int32_t lastf = n > 0 ? h_fsr_[dir ? b : b + n - 1] : 0;
      for (int32_t i = 0; i < nr; i++) {
        const size_t o = (size_t)(r0 + i) * tpw + sl;
        if (i < n) {
          const int32_t sidx = dir ? b + n - 1 - i : b + i;
          ilen[o] = h_len_[sidx];
          ifsr[o] = key(h_fsr_[sidx]);
        } else {
          ilen[o] = 0.f;
          ifsr[o] = key(lastf);
        }
      }
    }
  }
  dev_alloc(&d_seg_len_, total);
  dev_alloc(&d_seg_key_, total);
  dev_alloc(&d_work_, work.size());
  dev_alloc(&d_tile_, tile.size());
  HIP_CHECK(hipMemcpy(d_seg_len_, ilen.data(), total * sizeof(float), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_seg_key_, ifsr.data(), total * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_work_, work.data(), work.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_tile_, tile.data(), tile.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  layout_tpw_ = tpw;
  layout_ntiles_ = (int)ntiles;
  layout_rows_ = rows;
}

int HipSolver::sweepTracksPerWave() const {
  return shape_.lds_groups > 0 ? 64 / shape_.lds_groups : shape_.tracks_per_wave;
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
  sigma_t_max_ = 0.0;
  for (size_t i = 0; i < (size_t)M * G; i++) sigma_t_max_ = std::max(sigma_t_max_, sigma_t[i]);
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
  {
    // 1/V for the fixed-point tally; zero-volume FSRs get 0, as the reference
    // divides their tally by 1e30 (src/CPUSolver.cpp:2621-2622)
    std::vector<float> iv(num_fsrs);
    for (int64_t r = 0; r < num_fsrs; r++)
      iv[r] = volume[r] < kFltEpsilon ? 0.f : (float)(1.0 / volume[r]);
    dev_alloc(&d_inv_vol_, num_fsrs);
    HIP_CHECK(hipMemcpy(d_inv_vol_, iv.data(), num_fsrs * sizeof(float), hipMemcpyHostToDevice));
  }
  h_fsr_mat_.assign(material, material + num_fsrs);
  h_fsr_vol_.assign(volume, volume + num_fsrs);
  layout_tpw_ = 0;
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
  dev_alloc(&d_tally_, NG * shape_.tally_copies);
  dev_alloc(&d_qs_, NG);
  dev_alloc(&d_phi_, NG);
  dev_alloc(&d_phi_old_, NG);
  dev_alloc(&d_Q_, NG);
  if (!shape_forced_) pickShape();
  if (shape_.lds_groups > 0) {
    const int nsub = (G_ + shape_.lds_groups - 1) / shape_.lds_groups;
    dev_alloc(&d_lds_partials_, (size_t)nsub * shape_.lds_blocks_per_sub * N_ * shape_.lds_groups);
    dev_alloc(&d_qsub_, (size_t)nsub * N_ * shape_.lds_groups);
    HIP_CHECK(hipMemset(d_qsub_, 0, (size_t)nsub * N_ * shape_.lds_groups * sizeof(float)));
    dev_alloc(&d_tally64_, NG);
    HIP_CHECK(hipMemset(d_tally64_, 0, NG * sizeof(double)));
  }
  num_partials_ = (int)blocks_for(N_, kReduceBlock);
  dev_alloc(&d_partials_, num_partials_);
  buildTileLayout(sweepTracksPerWave());
  HIP_CHECK(hipMemset(d_psi_[0], 0, psi_n * sizeof(float)));
  HIP_CHECK(hipMemset(d_psi_[1], 0, psi_n * sizeof(float)));
  HIP_CHECK(hipMemset(d_tally_, 0, NG * shape_.tally_copies * sizeof(float)));
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

void HipSolver::launchSweep(const float* psi_in, float* psi_out, hipStream_t s,
                            hipEvent_t ev_after_sweep) {
  SweepArgs a;
  if (layout_tpw_ != sweepTracksPerWave()) buildTileLayout(sweepTracksPerWave());
  a.work = d_work_;
  a.ntiles = layout_ntiles_;
  a.tpw = layout_tpw_;
  a.tile = reinterpret_cast<const int2*>(d_tile_);
  a.trk = reinterpret_cast<const int4*>(d_trk_);
  a.trk_azim = d_trk_azim_;
  a.seg_len = d_seg_len_;
  a.seg_key = d_seg_key_;
  a.qs = d_qs_;
  a.psi_in = psi_in;
  a.psi_out = psi_out;
  a.tally = d_tally_;
  a.tally_copies = shape_.tally_copies;
  a.tally_stride = (size_t)N_ * G_;
  a.weights = d_weights_;
  const float l2e = exp_mode_ == MOC_EXP_HW ? 1.4426950408889634f : 1.0f;
  a.pf0 = inv_sin_[0] * l2e;
  a.pf1 = inv_sin_[1] * l2e;
  a.pf2 = inv_sin_[2] * l2e;
  a.state = d_state_;
  a.psi_max_bits = reinterpret_cast<unsigned*>(reinterpret_cast<char*>(d_state_) +
                                               offsetof(IterState, psi_max_bits));
  a.G = G_;
  a.P = P_;
  a.fpt = P_ * G_;
  a.N = (int)N_;
  a.lanes_per_track = shape_.lanes_per_track;
  a.tracks_per_wave = shape_.tracks_per_wave;
  if (shape_.lds_groups > 0) {
    LdsParams L;
    L.GB = shape_.lds_groups;
    L.nsub = (G_ + L.GB - 1) / L.GB;
    L.blocks_per_sub = shape_.lds_blocks_per_sub;
    L.partials = d_lds_partials_;
    L.qsub = d_qsub_;
    L.inv_vol = d_inv_vol_;
    L.sigma_t = d_sigma_t_;
    L.M = M_;
    const size_t lds_bytes = ldsBytes(L.GB, shape_.lds_q);
    if (lds_bytes > lds_max_) throw DeviceError(MOC_ERR_INVALID, "LDS slice exceeds the device limit");
    const unsigned grid = (unsigned)(L.nsub * L.blocks_per_sub);
    auto launch = [&](auto kernel) {
      HIP_CHECK(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds_bytes));
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(shape_.lds_block), lds_bytes, s, a, L);
    };
    auto by_np = [&](auto e, auto q) {
      constexpr int E = decltype(e)::value, Q = decltype(q)::value;
      switch (P_) {
        case 1: launch(sweep2d_lds_kernel<E, Q, 1>); break;
        case 2: launch(sweep2d_lds_kernel<E, Q, 2>); break;
        default: launch(sweep2d_lds_kernel<E, Q, 3>); break;
      }
    };
    using EH = std::integral_constant<int, MOC_EXP_HW>;
    using ER = std::integral_constant<int, MOC_EXP_RATIONAL>;
    using Q1 = std::integral_constant<int, 1>;
    using Q0 = std::integral_constant<int, 0>;
    if (exp_mode_ == MOC_EXP_HW)
      shape_.lds_q ? by_np(EH{}, Q1{}) : by_np(EH{}, Q0{});
    else
      shape_.lds_q ? by_np(ER{}, Q1{}) : by_np(ER{}, Q0{});
    HIP_CHECK(hipGetLastError());
    if (ev_after_sweep) HIP_CHECK(hipEventRecord(ev_after_sweep, s));
    const int64_t nred = (int64_t)L.nsub * N_ * L.GB;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(blocks_for(nred, 256)), dim3(256), 0, s,
                       (const unsigned long long*)d_lds_partials_, d_tally64_, (int64_t)N_, G_,
                       L.GB, L.nsub,
                       L.blocks_per_sub, (const IterState*)d_state_);
    HIP_CHECK(hipGetLastError());
    return;
  }
  const int64_t waves = a.ntiles;
  const int wpb = shape_.block / 64;
  const unsigned grid = (unsigned)((waves + wpb - 1) / wpb);
  dim3 blk(shape_.block);
  auto go = [&](auto gpl) {
    constexpr int GPL = decltype(gpl)::value;
    auto by_np = [&](auto kernel1, auto kernel2, auto kernel3) {
      switch (P_) {
        case 1: hipLaunchKernelGGL(kernel1, dim3(grid), blk, 0, s, a); break;
        case 2: hipLaunchKernelGGL(kernel2, dim3(grid), blk, 0, s, a); break;
        default: hipLaunchKernelGGL(kernel3, dim3(grid), blk, 0, s, a); break;
      }
    };
    if (exp_mode_ == MOC_EXP_HW)
      by_np(sweep2d_kernel<GPL, MOC_EXP_HW, 1>, sweep2d_kernel<GPL, MOC_EXP_HW, 2>,
            sweep2d_kernel<GPL, MOC_EXP_HW, 3>);
    else
      by_np(sweep2d_kernel<GPL, MOC_EXP_RATIONAL, 1>, sweep2d_kernel<GPL, MOC_EXP_RATIONAL, 2>,
            sweep2d_kernel<GPL, MOC_EXP_RATIONAL, 3>);
  };
  switch (shape_.groups_per_lane) {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 9: go(std::integral_constant<int, 9>{}); break;
    case 16: go(std::integral_constant<int, 16>{}); break;
    default:
      throw DeviceError(MOC_ERR_INVALID, "unsupported groups per lane");
  }
  HIP_CHECK(hipGetLastError());
  if (ev_after_sweep) HIP_CHECK(hipEventRecord(ev_after_sweep, s));
}

namespace {
// Points the per-FSR kernels at the source/tally buffers of the LDS sweep.
void apply_lds_args(FsrArgs& fa, const LaunchShape& sh, double* tally64, float* qsub) {
  if (sh.lds_groups > 0) {
    fa.tally64 = tally64;
    fa.qsub = qsub;
    fa.qsub_groups = sh.lds_groups;
    fa.qs = nullptr;
  }
}

FsrArgs make_fsr_args(int64_t N, int G, int32_t* fsr_mat, double* vol, double* st, double* ss,
                      double* nsf, double* chi, int32_t* fis, double* fixed, double* phi,
                      double* old, double* Q, float2* qs, float* tally, int copies,
                      double* partials, IterState* state, bool allow_neg) {
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
  a.tally_copies = copies;
  a.tally64 = nullptr;
  a.qsub = nullptr;
  a.qsub_groups = 1;
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
                             d_phi_old_, d_Q_, d_qs_, d_tally_, shape_.tally_copies, d_partials_, d_state_,
                             allow_negative_);
  apply_lds_args(fa, shape_, d_tally64_, d_qsub_);
  const unsigned nb = blocks_for(N_, kReduceBlock);
  if (ev0) HIP_CHECK(hipEventRecord(ev0, stream_));
  launchSweep(d_psi_[psi_cur_], d_psi_[psi_cur_ ^ 1], stream_, ev1);
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
                     tol, mode == MODE_EIGEN ? 0 : 1, sigma_t_max_, recompute_sources ? 1 : 0);
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
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG * shape_.tally_copies, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)(NG * shape_.tally_copies));
  psi_cur_ = 0;

  FsrArgs fa = make_fsr_args(N_, G_, d_fsr_mat_, d_fsr_vol_, d_sigma_t_, d_sigma_s_, d_nsf_,
                             d_chi_, d_mat_fissionable_, fixed_on_ ? d_fixed_ : nullptr, d_phi_,
                             d_phi_old_, d_Q_, d_qs_, d_tally_, shape_.tally_copies, d_partials_, d_state_,
                             allow_negative_);
  apply_lds_args(fa, shape_, d_tally64_, d_qsub_);
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
  // fixed-point scale of the first sweep (start fluxes are zero)
  hipLaunchKernelGGL(tally_scale_kernel, dim3(1), dim3(1), 0, stream_, d_state_, sigma_t_max_, 1);
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
  {
    // fixed-point scale from the inputs' bounds (see set_tally_scale)
    std::vector<double> st_h((size_t)M_ * G_);
    HIP_CHECK(hipMemcpy(st_h.data(), d_sigma_t_, st_h.size() * sizeof(double), hipMemcpyDeviceToHost));
    double qmax = 0.0, pmax = 0.0;
    for (size_t i = 0; i < NG; i++)
      qmax = std::max(qmax, std::fabs(Q[i] / st_h[(size_t)h_fsr_mat_[i / G_] * G_ + i % G_]));
    for (size_t i = 0; i < psi_n; i++) pmax = std::max(pmax, (double)std::fabs(psi[i]));
    st.qmax = (float)(qmax * (1 + 1e-6));
    st.psi_max = (float)(pmax * (1 + 1e-6));
    const double bound = 2.0 * kFourPi * sigma_t_max_ * ((double)st.psi_max + (double)st.qmax);
    st.tally_scale = bound > 1e-200 ? 0x1p60 / bound : 0x1p60;
  }
  HIP_CHECK(hipMemcpyAsync(d_state_, &st, sizeof(st), hipMemcpyHostToDevice, stream_));
  HIP_CHECK(hipMemcpyAsync(d_Q_, Q, NG * sizeof(double), hipMemcpyHostToDevice, stream_));
  const bool lds = shape_.lds_groups > 0;
  hipLaunchKernelGGL(make_qs_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_Q_,
                     d_fsr_mat_, d_sigma_t_, lds ? nullptr : d_qs_, lds ? d_qsub_ : nullptr,
                     std::max(1, shape_.lds_groups), N_, G_);
  HIP_CHECK(hipMemcpyAsync(d_psi_[0], psi, psi_n * sizeof(float), hipMemcpyHostToDevice, stream_));
  // slots that no track writes keep their value (vacuum inflow), as in
  // CPUSolver::_start_flux
  HIP_CHECK(hipMemcpyAsync(d_psi_[1], psi, psi_n * sizeof(float), hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG * shape_.tally_copies, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)(NG * shape_.tally_copies));
  launchSweep(d_psi_[0], d_psi_[1], stream_);
  std::vector<float> t32;
  if (shape_.lds_groups > 0) {
    // the fixed-point tally holds sum / V; the raw tally is sum
    HIP_CHECK(hipMemcpyAsync(tally, d_tally64_, NG * sizeof(double), hipMemcpyDeviceToHost, stream_));
  } else {
    t32.resize(NG * shape_.tally_copies);
    HIP_CHECK(hipMemcpyAsync(t32.data(), d_tally_, NG * shape_.tally_copies * sizeof(float),
                             hipMemcpyDeviceToHost, stream_));
  }
  HIP_CHECK(hipMemcpyAsync(psi, d_psi_[1], psi_n * sizeof(float), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  if (shape_.lds_groups == 0) {
    for (size_t i = 0; i < NG; i++) {
      double t = 0.0;
      for (int c = 0; c < shape_.tally_copies; c++) t += t32[c * NG + i];
      tally[i] = t;
    }
  } else {
    for (size_t i = 0; i < NG; i++) {
      const double v = h_fsr_vol_[i / G_];
      tally[i] *= v < kFltEpsilon ? 0.0 : v;
    }
  }
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG * shape_.tally_copies, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)(NG * shape_.tally_copies));
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
  if (!state_valid_) launchInit(MODE_EIGEN, 1.0);
  // repeated sweeps on the current sources and start fluxes (psi_scale 1, so
  // |psi| stays within the bound the fixed-point scale is computed from)
  HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  h_state_->converged = 0;
  h_state_->psi_scale = 1.0;
  h_state_->psi_max_bits = __builtin_bit_cast(uint32_t, std::max(h_state_->psi_max,
                                                   __builtin_bit_cast(float, h_state_->psi_max_bits)));
  HIP_CHECK(hipMemcpyAsync(d_state_, h_state_, sizeof(IterState), hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(tally_scale_kernel, dim3(1), dim3(1), 0, stream_, d_state_, sigma_t_max_, 0);
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
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG * shape_.tally_copies, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)(NG * shape_.tally_copies));
  HIP_CHECK(hipStreamSynchronize(stream_));
  return ms * 1e-3 / repeats;
}

}  // namespace moc
