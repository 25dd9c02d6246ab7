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
//   keff_kernel        computeKeff + normalisation factor (FP64, fixed order);
//                      in an eigen iteration every update_kernel block forms
//                      the same sums itself and block 0 stores them
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
#include <map>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>

#include "hip_solver.h"
#include "fluxfile.h"

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
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, workgroups dispatched round-robin
// Segment-stream rows: every tile's row count is a multiple of kRowPad (the
// pipeline chunk sizes divide it), and kTailRows rows precede the first tile
// and follow the last, so the pipeline's look-ahead loads (at most 3 chunks
// of 4 rows past a tile's end, or before its start in a backward pass) stay
// inside the stream without clamping.
constexpr int kRowPad = 4;
constexpr int kTailRows = 12;
// Tile table entries are (first row, rows | kTileBackward).  Where both
// directions of a track read one stream, a tile has two work tiles: the
// forward pass reads its rows first to last, the backward pass last to first.
constexpr int kTileBackward = 1 << 30;
constexpr int kTileRowsMask = kTileBackward - 1;
// Source iterations captured once into a graph and relaunched as a batch:
// one launch call per batch instead of six per iteration.  The batch length
// is even, so the ping-pong start-flux buffers return to their places.
constexpr int kGraphIters = 8;

// Fixed-point tally scale for a bound B on |sum of contribution / V| of any
// FSR over one sweep: contributions / V are multiplied by the float
// s32 = 2^28 / B and split exactly into an int64 in units of 2^-32, so B maps
// to 2^60 (int64 range 2^63).  Returns the effective scale s32 * 2^32.
// `cap` keeps (1/V) * s32 finite in FP32 for the smallest FSR volume.
__host__ __device__ inline double tally_scale_for(double bound, double cap) {
  const double s = 0x1p28 / (bound > 1e-300 ? bound : 1e-300);
  const float s32 = (float)(s < cap ? s : cap);
  return (double)s32 * 0x1p32;
}

// Device allocations remember their size, so a buffer that is big enough is
// kept when it is "allocated" again (a solver re-armed with new cross
// sections then makes no hipMalloc/hipFree, which would stall other streams).
std::mutex g_alloc_mu;
std::map<void*, size_t> g_alloc_bytes;

template <typename T>
void dev_alloc(T** p, size_t n) {
  if (n == 0) n = 1;
  const size_t bytes = n * sizeof(T);
  if (*p) {
    {
      std::lock_guard<std::mutex> lk(g_alloc_mu);
      auto it = g_alloc_bytes.find((void*)*p);
      if (it != g_alloc_bytes.end() && it->second >= bytes) return;
      if (it != g_alloc_bytes.end()) g_alloc_bytes.erase(it);
    }
    HIP_CHECK(hipFree(*p));
    *p = nullptr;
  }
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    throw DeviceError(MOC_ERR_MEMORY, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  }
  std::lock_guard<std::mutex> lk(g_alloc_mu);
  g_alloc_bytes[(void*)*p] = bytes;
}

template <typename T>
void dev_free(T*& p) {
  if (p) {
    {
      std::lock_guard<std::mutex> lk(g_alloc_mu);
      g_alloc_bytes.erase((void*)p);
    }
    (void)hipFree(p);
  }
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
}  // namespace

struct SweepArgs {
  const int32_t* work;      // [ntiles*tpw] directed track ids (t*2+d) or -1, longest first
  int ntiles;               // work tiles of tpw directed tracks (one wave each)
  int tpw;                  // directed tracks per tile (= per wave)
  const int2* tile;         // [ntiles] (first row, rows | kTileBackward) of the interleaved streams
  const int4* trk;          // seg_begin, seg_count, out slot fwd, out slot bwd
  const int32_t* trk_azim;
  const float* seg_len;     // tile-interleaved: [row][tpw], forward order (or traversal order), 0-padded
  const int32_t* seg_key;   // tile-interleaved: [row][tpw], FSR | material << 24
  unsigned stream_bytes;    // bytes of each stream (rows + tail) * tpw * 4
  const float2* qs;         // [N*G] (Q/sigma_t, sigma_t)
  const float* psi_in;      // [G][W][P], W = ntiles * tpw: group-major by work position
  float* psi_out;           // [G][W][P]
  const int32_t* out_wp;    // [2T] work position of the linked start slot, -1 vacuum
  int psi_w;                // W
  float* leak;              // [2T*G] vacuum leakage w(a,0) * sum_p psi_out, or null
  float* tally;             // global path: [tally_copies][N*G]
  int tally_copies;         // global path: private copies (block b -> copy b % copies)
  float* segout;            // global path, per-segment store: [slots][seg_stride], or null
  const int32_t* seg_slot;  // tile-interleaved like seg_len: each segment's store row
  int seg_stride;           // floats per store row (G)
  size_t tally_stride;      // N*G
  const float* weights;     // [num_azim/2][P]
  float pf0, pf1, pf2;      // polar factors: 1/sin(theta_p) (times log2 e for HW exp)
  const IterState* state;
  unsigned* psi_max_bits;   // &state->psi_max_bits
  int G, P, fpt, lanes_per_track, tracks_per_wave;
  int N;
};

namespace {

// Segment streams carry a packed key: FSR id in the low 24 bits, material in
// the high 8 bits (N < 2^24, M <= 256 are checked on the host).
constexpr int kFsrBits = 24;
constexpr int kFsrMask = (1 << kFsrBits) - 1;

// Source policies: Q/sigma_t and sigma_t of group g of the segment's FSR,
// plus the FSR's fixed-point tally factor (1/V times the tally scale) where
// the tally needs it.
struct SrcVal {
  float q;    // Q / sigma_t
  float sig;  // sigma_t
  float iv;   // (1/V) * s32 (fixed-point tally only)
};
struct GatherSource {  // global FP32 table qs[N*G]
  static constexpr bool kLocal = false;
  const float2* qs;
  int G;
  __device__ SrcVal load(int key, int g, int /*loc*/) const {
    const float2 v = qs[(size_t)(key & kFsrMask) * G + g];
    return {v.x, v.y, 0.f};
  }
};
struct LdsSource {     // LDS: {Q/sigma_t, iv} [N][GB] and sigma_t by material [M][GB]
  static constexpr bool kLocal = true;
  const float2* qi;
  const float* sig;
  int GB;
  __device__ SrcVal load(int key, int /*g*/, int loc) const {
    const float2 v = qi[__umul24(key & kFsrMask, GB) + loc];
    return {v.x, sig[__umul24((unsigned)key >> kFsrBits, GB) + loc], v.y};
  }
};
struct SubsetSource {  // global subset-major Q/sigma_t [N][GB]; LDS iv [N], sigma_t [M][GB]
  static constexpr bool kLocal = false;
  const float* q;
  const float* iv;
  const float* sig;
  int GB;
  __device__ SrcVal load(int key, int /*g*/, int loc) const {
    const int f = key & kFsrMask;
    return {q[__umul24(f, GB) + loc], sig[__umul24((unsigned)key >> kFsrBits, GB) + loc], iv[f]};
  }
};

// Tally policies: where a lane's weighted delta-psi of one segment goes.
struct GlobalTally {   // device-scope float atomics into this block's copy
  static constexpr bool kStore = false;
  float* t;
  int G;
  __device__ void add(int fsr, int g, int /*loc*/, float v, float /*iv*/) const {
    unsafeAtomicAdd(&t[(size_t)fsr * G + g], v);
  }
};
// Fixed-point LDS tally: the block's group slice [N][GB] of int64 sums of
// contribution / V_fsr in units of 2^-32 / s32 (see tally_scale_for).
// Integer LDS atomics run at full rate (float LDS atomics do not), and
// integer sums are exact and order-independent, so the tally is bitwise
// reproducible.
struct LdsFixedTally {
  static constexpr bool kStore = false;
  unsigned long long* t;
  int GB;
  __device__ void add(int fsr, int /*g*/, int loc, float v, float iv) const {
    // y = contribution * (1/V) * s32, |y| <= 2^28; the int64 floor(y * 2^32)
    // is built from floor(y) (high word) and (y - floor(y)) * 2^32 (low
    // word), both exact in FP32 down to the last bit of y
    const float y = v * iv;
    const float hi = floorf(y);
    const unsigned lo = (unsigned)((y - hi) * 4294967296.f);
    const unsigned long long x =
        ((unsigned long long)(unsigned)(int)hi << 32) | (unsigned long long)lo;
    atomicAdd(&t[__umul24(fsr, GB) + loc], x);
  }
};

// Per-segment store: the weighted delta-psi of a segment and group g goes to
// row `slot` of out, out[slot * stride + g], no atomics.  Slots are laid out
// FSR by FSR (each FSR's segments in a fixed order), so the per-FSR step sums
// a contiguous run of rows and the tally is deterministic.
struct SegmentStore {
  static constexpr bool kStore = true;
  float* out;
  int stride;
  __device__ void store(int slot, int g, float v) const { out[(size_t)slot * stride + g] = v; }
};

// Integrates directed track dt (tile slot `slot`) for the lane's groups
// g_begin + gl + k*lpt (k < GPL, valid while < g_begin + g_count), taking
// sources from `src`, tallying through `tally`, and writes the outgoing
// angular flux to the linked track's start slot.
// Running maximum of non-negative floats by their bit patterns (which order
// like the values).  The current maximum is read first and the atomic is
// issued only to raise it, so after the first waves most lanes issue none:
// same-address atomics from every lane of a launch serialise in L2.
__device__ __forceinline__ void raise_max_bits(unsigned* p, float v) {
  const unsigned b = __float_as_uint(v);
  if (b > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(p, b);
}

template <int GPL, int EXPM, int NP, class Src, class Tally>
__device__ __forceinline__ void sweep_track(const SweepArgs& a, int tile, int slot, int dt, int gl,
                                            int lpt, int g_begin, int g_count, const Src& src,
                                            const Tally& tally) {
  const int t = dt >> 1;
  const int2 ti = a.tile[tile];
  const int azim = a.trk_azim[t];
  constexpr int P = NP;  // polar half-angles, compile time: the segment body is straight-line
  // start fluxes: group-major, the P polar values of a (group, directed
  // track) contiguous, tracks in work order, so a wave's loads are one run
  const size_t wp = (size_t)tile * a.tpw + slot;
  const float scale = (float)a.state->psi_scale;
  const float pf[kMaxPolarHalf] = {a.pf0, a.pf1, a.pf2};
  float w[P];
#pragma unroll
  for (int p = 0; p < P; p++) w[p] = a.weights[azim * P + p];
  if (GPL == 1 && gl >= g_count) {
    // a lane past the last group tallies exactly zero
#pragma unroll
    for (int p = 0; p < P; p++) w[p] = 0.f;
  }

  // Lanes past the last group work on a clamped group and add exactly zero,
  // so the loop below has no per-lane branches.
  int g[GPL], loc[GPL];
  bool ok[GPL];
  float psi[GPL][P];
#pragma unroll
  for (int k = 0; k < GPL; k++) {
    const int l = gl + k * lpt;
    ok[k] = l < g_count;
    loc[k] = ok[k] ? l : g_count - 1;
    g[k] = g_begin + loc[k];
    const float* pin = a.psi_in + ((size_t)g[k] * a.psi_w + wp) * P;
#pragma unroll
    for (int p = 0; p < P; p++) psi[k][p] = ok[k] ? pin[p] * scale : 0.f;
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
  static_assert(kRowPad % D == 0 && kTailRows >= 3 * D, "stream padding");
  const int n = ti.y & kTileRowsMask;  // multiple of kRowPad; tile is wave-uniform
  // backward pass over a stream shared by both directions: chunk c is rows
  // [n - (c+1) D, n - c D) of the tile, last row first; the tile's padding
  // rows (length 0, exact no-ops) then come before its segments
  const bool rev = (ti.y & kTileBackward) != 0;
  // Buffer loads: 32-bit lane offset plus a scalar row offset, bounds-checked
  // by the hardware against the stream size (out-of-range reads give 0).
  const __amdgpu_buffer_rsrc_t rs_len =
      __builtin_amdgcn_make_buffer_rsrc((void*)a.seg_len, (short)0, (int)a.stream_bytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rs_key =
      __builtin_amdgcn_make_buffer_rsrc((void*)a.seg_key, (short)0, (int)a.stream_bytes, 0x00020000);
  const unsigned row_bytes = (unsigned)a.tpw * 4u;
  const unsigned lane_off = (unsigned)ti.x * row_bytes + (unsigned)slot * 4u;
  const __amdgpu_buffer_rsrc_t rs_slot = __builtin_amdgcn_make_buffer_rsrc(
      (void*)a.seg_slot, (short)0, Tally::kStore ? (int)a.stream_bytes : 0, 0x00020000);
  auto fetch = [&](int key, SrcVal* q) {
#pragma unroll
    for (int k = 0; k < GPL; k++) q[k] = src.load(key, g[k], loc[k]);
  };
  auto integrate = [&](float len, int key, const SrcVal* q, int sidx) {
    const int fsr = key & kFsrMask;
#pragma unroll
    for (int k = 0; k < GPL; k++) {
      const float tau = q[k].sig * len;
      float acc = 0.f;
#pragma unroll
      for (int p = 0; p < P; p++) {
        const float x = tau * pf[p];
        const float A = EXPM == MOC_EXP_RATIONAL ? attenuation_rational(x) : attenuation_hw(x);
        // delta = (psi - Q/sigma_t) A; the flux update is one fused
        // multiply-add on the same difference and the tally takes the
        // rounded product.  Written out so that every kernel built on this
        // loop rounds alike, whatever the compiler would fuse on its own.
        const float dq = psi[k][p] - q[k].q;
        const float d = dq * A;
        psi[k][p] = fmaf(-dq, A, psi[k][p]);
        acc = fmaf(w[p], d, acc);
      }
      if constexpr (Tally::kStore) {
        if (ok[k]) tally.store(sidx, g[k], acc);
      } else {
        if (GPL > 1 && !ok[k]) acc = 0.f;
        tally.add(fsr, g[k], loc[k], acc, q[k].iv);
      }
    }
  };
  // Integrates one chunk.  Local sources: read one segment ahead.  Global
  // sources: `cur` holds this chunk's sources (fetched one chunk earlier) and
  // the next chunk's sources are fetched into `nxt` first.
  auto process = [&](const float* L, const int* K, const int* S, const int* Knext,
                     SrcVal (*cur)[GPL], SrcVal (*nxt)[GPL]) {
    if constexpr (Src::kLocal) {
      SrcVal qa[GPL], qb[GPL];
      fetch(K[0], qa);
#pragma unroll
      for (int j = 0; j < D; j++) {
        if (j + 1 < D) fetch(K[j + 1], (j & 1) ? qa : qb);
        integrate(L[j], K[j], (j & 1) ? qb : qa, Tally::kStore ? S[j] : 0);
      }
      (void)Knext;
      (void)cur;
      (void)nxt;
    } else {
#pragma unroll
      for (int j = 0; j < D; j++) fetch(Knext[j], nxt[j]);
#pragma unroll
      for (int j = 0; j < D; j++) integrate(L[j], K[j], cur[j], Tally::kStore ? S[j] : 0);
    }
  };

  // The pipeline, with the row order a compile-time choice (a run-time one
  // makes the compiler split the steady-state loop into branches).  A
  // backward pass's first row may lie before the tile: the wrap of the
  // unsigned row offset cancels in the sum, which stays inside the stream.
  auto pipeline = [&](auto rev_tag) {
  constexpr bool REV = decltype(rev_tag)::value;
  auto load_chunk = [&](int c, float* L, int* K, int* S) {
    const int r = REV ? n - (c + 1) * D : c * D;
    const unsigned vo = lane_off + (unsigned)r * row_bytes;
#pragma unroll
    for (int j = 0; j < D; j++) {
      const int so = (REV ? D - 1 - j : j) * (int)row_bytes;
      L[j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs_len, vo, so, 0));
      K[j] = (int)__builtin_amdgcn_raw_buffer_load_b32(rs_key, vo, so, 0);
      // a shared stream's backward pass stores to the row after the forward one
      if constexpr (Tally::kStore)
        S[j] = (int)__builtin_amdgcn_raw_buffer_load_b32(rs_slot, vo, so, 0) + (REV ? 1 : 0);
    }
  };
  if (n > 0) {
    const int nch = n / D;
    // three rotating sets: chunk c lives in set c % 3 (streams and sources)
    float L0[D], L1[D], L2[D];
    int K0[D], K1[D], K2[D];
    int S0[D], S1[D], S2[D];  // store rows (per-segment store only)
    SrcVal Q0[Src::kLocal ? 1 : D][GPL], Q1[Src::kLocal ? 1 : D][GPL], Q2[Src::kLocal ? 1 : D][GPL];
    load_chunk(0, L0, K0, S0);
    load_chunk(1, L1, K1, S1);
    load_chunk(2, L2, K2, S2);
    if constexpr (!Src::kLocal) {
#pragma unroll
      for (int j = 0; j < D; j++) fetch(K0[j], Q0[j]);
    }
    // The steady-state loop has no conditional sections, so the compiler's
    // wait counts at its head see all three sets in flight; the last one or
    // two chunks are integrated after it.
    // Scheduling barriers keep each chunk's work in its own section: without
    // them the scheduler hoists later chunks' address arithmetic to the loop
    // head, where it has to wait for the newest loads.
    int c = 0;
    for (; c + 3 <= nch; c += 3) {
      process(L0, K0, S0, K1, Q0, Q1);
      load_chunk(c + 3, L0, K0, S0);
      __builtin_amdgcn_sched_barrier(0);
      process(L1, K1, S1, K2, Q1, Q2);
      load_chunk(c + 4, L1, K1, S1);
      __builtin_amdgcn_sched_barrier(0);
      process(L2, K2, S2, K0, Q2, Q0);
      load_chunk(c + 5, L2, K2, S2);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (c < nch) process(L0, K0, S0, K1, Q0, Q1);
    if (c + 1 < nch) process(L1, K1, S1, K2, Q1, Q2);
  }
  };
  if (rev)
    pipeline(std::true_type{});
  else
    pipeline(std::false_type{});

  const int out = a.out_wp[dt];
  float amax = 0.f;
  if (out >= 0) {
#pragma unroll
    for (int k = 0; k < GPL; k++) {
      float* po = a.psi_out + ((size_t)g[k] * a.psi_w + out) * P;
#pragma unroll
      for (int p = 0; p < P; p++)
        if (ok[k]) {
          po[p] = psi[k][p];
          amax = fmaxf(amax, fabsf(psi[k][p]));
        }
    }
  }
  if (a.leak) {
    // vacuum leakage of the neutron-balance keff (transferBoundaryFlux,
    // src/CPUSolver.cpp:2593-2600: the 2D sweep passes polar index 0, so
    // every polar angle of the track is weighted by w(a, 0)); every directed
    // track writes its groups each sweep
    const float w0 = a.weights[azim * P];
#pragma unroll
    for (int k = 0; k < GPL; k++)
      if (ok[k]) {
        float v = 0.f;
#pragma unroll
        for (int p = 0; p < P; p++) v += psi[k][p];
        a.leak[(size_t)dt * a.G + g[k]] = out < 0 ? w0 * v : 0.f;
      }
  }
  // running max of |start psi| for the next sweep's fixed-point scale
  // (non-negative floats order like their bit patterns)
  if (amax > 0.f) raise_max_bits(&a.psi_max_bits[0], amax);
}

// Global-atomic sweep: one wave per tile of tpw directed tracks,
// lanes_per_track lanes per track, GPL groups per lane.
template <int GPL, int EXPM, int NP, int STORE>
__device__ __forceinline__ void sweep2d_kernel_body(const SweepArgs& a) {
  const int lane = threadIdx.x & 63;
  const int tile = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int slot = lane / a.lanes_per_track;
  const int gl = lane - slot * a.lanes_per_track;
  if (slot >= a.tpw || tile >= a.ntiles) return;
  if (a.state->converged) return;
  const int dt = a.work[tile * a.tpw + slot];
  if (dt < 0) return;
  GatherSource src{a.qs, a.G};
  if constexpr (STORE) {
    SegmentStore tally{a.segout, a.seg_stride};
    sweep_track<GPL, EXPM, NP>(a, tile, slot, dt, gl, a.lanes_per_track, 0, a.G, src, tally);
  } else {
    GlobalTally tally{a.tally + (size_t)(blockIdx.x % a.tally_copies) * a.tally_stride, a.G};
    sweep_track<GPL, EXPM, NP>(a, tile, slot, dt, gl, a.lanes_per_track, 0, a.G, src, tally);
  }
}

template <int GPL, int EXPM, int NP, int STORE>
__global__ __launch_bounds__(kSweepBlockMax) void sweep2d_kernel(SweepArgs a) {
  sweep2d_kernel_body<GPL, EXPM, NP, STORE>(a);
}

// ---------------------------------------------------------------------------
// Linear-source sweep (CPULSSolver::tallyLSScalarFlux 2D,
// src/CPULSSolver.cpp:542-739)
// ---------------------------------------------------------------------------
// 5/6 rational fit of G(x) = 1/x - (1 - e^-x)/x^2 the reference uses to form
// F1, F2 and H (src/exponentials.h:100-145, src/ExpEvaluator.h:349-379).
__device__ __forceinline__ float exp_g_rational(float x) {
  float den = fmaf(6.064409107557148e-5f, x, 7.036272419147752e-4f);
  den = fmaf(den, x, 6.883236664917246e-3f);
  den = fmaf(den, x, 4.721469893686252e-2f);
  den = fmaf(den, x, 2.263358514260129e-1f);
  den = fmaf(den, x, 6.864462055546078e-1f);
  den = fmaf(den, x, 1.0f);
  float num = fmaf(6.064409107557148e-5f, x, 6.429894635552992e-4f);
  num = fmaf(num, x, 6.178333902037397e-3f);
  num = fmaf(num, x, 4.041584305811143e-2f);
  num = fmaf(num, x, 1.76558112351595e-1f);
  num = fmaf(num, x, 0.5f);
  return num * __builtin_amdgcn_rcpf(den);
}

struct LsArgs {
  const float2* seg_pos;    // tile-interleaved [row][tpw]: the segment's entry point
                            // in traversal order, relative to its FSR centroid
  const float4* qls;        // [N*G] (Q, Q_x, Q_y, sigma_t)
  float* tally_xy;          // [tally_copies][2][N*G] (x plane, y plane)
  const float* azim_cos;    // [num_azim/2]
  const float* azim_sin;
  // per-segment store instead of the atomics, or null: planes (flat, x, y)
  // of [slots][seg_stride] floats, seg_plane floats apart
  float* segout;
  int seg_stride;
  size_t seg_plane;
};

// One lane per (directed track, group subset of GPL groups); float atomics
// into the block's private tally copy (flat tally and the x/y moments).
template <int GPL, int NP>
__global__ __launch_bounds__(kSweepBlockMax) void sweep2d_ls_kernel(SweepArgs a, LsArgs L) {
  const int lane = threadIdx.x & 63;
  const int tile = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  const int slot = lane / a.lanes_per_track;
  const int gl = lane - slot * a.lanes_per_track;
  if (slot >= a.tpw || tile >= a.ntiles) return;
  if (a.state->converged) return;
  const int dt = a.work[tile * a.tpw + slot];
  if (dt < 0) return;
  const int t = dt >> 1;
  const int dir = dt & 1;
  const int2 ti = a.tile[tile];
  const int azim = a.trk_azim[t];
  const int G = a.G, lpt = a.lanes_per_track;
  const size_t wp = (size_t)tile * a.tpw + slot;
  constexpr int P = NP;
  const float scale = (float)a.state->psi_scale;
  const float pf[kMaxPolarHalf] = {a.pf0, a.pf1, a.pf2};
  float w[P];
#pragma unroll
  for (int p = 0; p < P; p++) w[p] = a.weights[azim * P + p];
  const float dx = dir ? -L.azim_cos[azim] : L.azim_cos[azim];
  const float dy = dir ? -L.azim_sin[azim] : L.azim_sin[azim];
  const size_t copy = (size_t)(blockIdx.x % a.tally_copies);
  float* tf = a.tally + copy * a.tally_stride;
  float* tx = L.tally_xy + copy * a.tally_stride * 2;  // x and y planes of this copy
  float* ty = tx + a.tally_stride;

  int g[GPL];
  bool ok[GPL];
  float psi[GPL][P];
#pragma unroll
  for (int k = 0; k < GPL; k++) {
    const int l = gl + k * lpt;
    ok[k] = l < G;
    g[k] = ok[k] ? l : G - 1;
    const float* pin = a.psi_in + ((size_t)g[k] * a.psi_w + wp) * P;
#pragma unroll
    for (int p = 0; p < P; p++) psi[k][p] = ok[k] ? pin[p] * scale : 0.f;
  }
  const int n = ti.y & kTileRowsMask;  // one stream per direction: never a backward pass
  // Two-deep pipeline: the streams of segment r + 2 and the sources of
  // segment r + 1 are loaded before segment r is integrated.  Rows past the
  // tile (the next tile's or the zero-length tail rows keyed to valid FSRs)
  // are read but never integrated; at most two of them are read
  // (kTailRows >= 2).
  const float* lp = a.seg_len + (size_t)ti.x * a.tpw + slot;
  const int* kp = a.seg_key + (size_t)ti.x * a.tpw + slot;
  const float2* pp = L.seg_pos + (size_t)ti.x * a.tpw + slot;
  // store rows of the per-segment store (the key stream stands in without it)
  const int* sp = (L.segout ? a.seg_slot : a.seg_key) + (size_t)ti.x * a.tpw + slot;
  const size_t tpw = (size_t)a.tpw;
  float len_a = lp[0], len_b = lp[tpw];
  int key_a = kp[0], key_b = kp[tpw];
  int sl_a = sp[0], sl_b = sp[tpw];
  float2 pos_a = pp[0], pos_b = pp[tpw];
  float4 q_a[GPL], q_b[GPL];
#pragma unroll
  for (int k = 0; k < GPL; k++) q_a[k] = L.qls[(size_t)(key_a & kFsrMask) * G + g[k]];
  for (int r = 0; r < n; r++) {
    const size_t o2 = (size_t)(r + 2) * tpw;
    const float len_c = lp[o2];
    const int key_c = kp[o2];
    const int sl_c = sp[o2];
    const float2 pos_c = pp[o2];
#pragma unroll
    for (int k = 0; k < GPL; k++) q_b[k] = L.qls[(size_t)(key_b & kFsrMask) * G + g[k]];
    const float len = len_a;
    const int key = key_a;
    const float2 pos = pos_a;
    const int fsr = key & kFsrMask;
    const float cx = 2.f * pos.x + len * dx, cy = 2.f * pos.y + len * dy;
#pragma unroll
    for (int k = 0; k < GPL; k++) {
      const float4 q = q_a[k];
      const float src_flat = q.x + q.y * cx + q.z * cy;
      const float src_lin = dx * q.y + dy * q.z;
      const float tau = q.w * len;
      float acc = 0.f, ax = 0.f, ay = 0.f;
#pragma unroll
      for (int p = 0; p < P; p++) {
        const float isin = pf[p];
        const float tq = fmaxf(1e-8f, tau * isin);
        const float eg = exp_g_rational(tq);
        const float f1 = (1.f - tq * eg) * isin;
        const float gq = eg * isin;
        const float f2 = 2.f * gq - f1;
        const float h = (f1 - gq) * w[p] * tau * len * psi[k][p];
        float d = (tau * psi[k][p] - len * src_flat) * f1 - len * len * src_lin * f2;
        psi[k][p] -= d;
        d *= w[p];
        acc += d;
        ax += h * dx + d * pos.x;
        ay += h * dy + d * pos.y;
      }
      if (L.segout) {
        if (ok[k]) {
          float* o = L.segout + (size_t)sl_a * L.seg_stride + g[k];
          o[0] = acc;
          o[L.seg_plane] = ax;
          o[2 * L.seg_plane] = ay;
        }
      } else if (ok[k] && len > 0.f) {
        const size_t i = (size_t)fsr * G + g[k];
        unsafeAtomicAdd(&tf[i], acc);
        unsafeAtomicAdd(&tx[i], ax);
        unsafeAtomicAdd(&ty[i], ay);
      }
    }
    len_a = len_b;
    key_a = key_b;
    sl_a = sl_b;
    pos_a = pos_b;
    len_b = len_c;
    key_b = key_c;
    sl_b = sl_c;
    pos_b = pos_c;
#pragma unroll
    for (int k = 0; k < GPL; k++) q_a[k] = q_b[k];
  }
  const int out = a.out_wp[dt];
  float amax = 0.f;
  if (out >= 0) {
#pragma unroll
    for (int k = 0; k < GPL; k++) {
      float* po = a.psi_out + ((size_t)g[k] * a.psi_w + out) * P;
#pragma unroll
      for (int p = 0; p < P; p++)
        if (ok[k]) {
          po[p] = psi[k][p];
          amax = fmaxf(amax, fabsf(psi[k][p]));
        }
    }
  }
  if (amax > 0.f) raise_max_bits(&a.psi_max_bits[0], amax);
}

// LDS-privatised sweep.  Block b owns group subset b % nsub (GB groups, one
// lane per group) and block slot b / nsub.  LDS holds the int64 fixed-point
// tally slice [N][GB] of the subset, 1/V per FSR, sigma_t of the subset by material [M][GB] and, when
// QLDS, the subset's Q/sigma_t slice [N][GB] (otherwise gathered from the
// subset-major global table).  The block processes the tiles
// slot*wpb + wave, + blocks_per_sub*wpb, ... (longest first, so the strided
// assignment balances) and finally writes its tally slice to its own partial
// buffer, which the per-FSR step sums, or adds the slice's non-zero entries to
// one global int64 tally with atomics.  Both sums are exact integer sums, so
// the two flushes give the same bits.
}  // namespace

struct LdsParams {
  unsigned long long* partials;  // [nsub][blocks_per_sub][N*GB] fixed-point sums, or null
  unsigned long long* tally64;   // [N*G] fixed-point sums (atomic flush), or null
  const float* qsub;      // [nsub][N][GB] Q/sigma_t, subset-major
  const float* inv_vol;   // [N]
  const double* sigma_t;  // [M][G]
  int M;
  int GB;                 // groups per subset (= lanes per track; tpw = 64 / GB)
  int nsub;               // ceil(G / GB)
  int blocks_per_sub;
};

namespace {

template <int EXPM, int QLDS, int NP>
__device__ __forceinline__ void sweep2d_lds_kernel_body(const SweepArgs& a, const LdsParams& L) {
  extern __shared__ unsigned long long lds64[];
  if (a.state->converged) return;  // uniform: read by every thread of the block
  // Workgroups are dispatched to the 8 XCDs round-robin by block id.  With
  // blocks_per_sub a multiple of 8, the nsub subset blocks of one slot (the
  // same tiles, different groups) are placed on the same XCD one after the
  // other, so the segment streams they all read come from that XCD's L2.
  int sub, bslot;
  if (L.blocks_per_sub % kNumXcd == 0) {
    const int x = blockIdx.x % kNumXcd, j = blockIdx.x / kNumXcd;
    sub = j % L.nsub;
    bslot = x + kNumXcd * (j / L.nsub);
  } else {
    sub = blockIdx.x % L.nsub;
    bslot = blockIdx.x / L.nsub;
  }
  const int g0 = sub * L.GB;
  const int gb = min(L.GB, a.G - g0);
  const int nl = a.N * L.GB;
  // LDS: int64 tally [N*GB] | then QLDS: {Q/sigma_t, iv} [N*GB], else iv [N] |
  // sigma_t [M*GB]; iv = (1/V) * s32 is the fixed-point factor of the FSR
  const float s32 = (float)(a.state->tally_scale * 0x1p-32);
  unsigned long long* lds_tally = lds64;
  float2* lds_qi = reinterpret_cast<float2*>(lds64 + nl);
  float* lds_iv = reinterpret_cast<float*>(lds64 + nl);
  float* lds_sig = QLDS ? reinterpret_cast<float*>(lds_qi + nl) : lds_iv + a.N;
  const float* qg = L.qsub + (size_t)sub * nl;
  for (int i = threadIdx.x; i < nl; i += blockDim.x) lds_tally[i] = 0ull;
  if (QLDS) {
    for (int i = threadIdx.x; i < nl; i += blockDim.x)
      lds_qi[i] = make_float2(qg[i], L.inv_vol[i / L.GB] * s32);
  } else {
    for (int i = threadIdx.x; i < a.N; i += blockDim.x) lds_iv[i] = L.inv_vol[i] * s32;
  }
  for (int i = threadIdx.x; i < L.M * L.GB; i += blockDim.x) {
    const int m = i / L.GB, l = i - m * L.GB;
    lds_sig[i] = (float)L.sigma_t[(size_t)m * a.G + g0 + min(l, gb - 1)];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int slot = lane / L.GB;
  const int gl = lane - slot * L.GB;
  if (slot < a.tpw) {
    LdsFixedTally tally{lds_tally, L.GB};
    for (int tile = __builtin_amdgcn_readfirstlane(bslot * wpb + wib); tile < a.ntiles;
         tile += L.blocks_per_sub * wpb) {
      const int dt = a.work[tile * a.tpw + slot];
      if (dt < 0) continue;
      if (QLDS) {
        LdsSource src{lds_qi, lds_sig, L.GB};
        sweep_track<1, EXPM, NP>(a, tile, slot, dt, gl, L.GB, g0, gb, src, tally);
      } else {
        SubsetSource src{qg, lds_iv, lds_sig, L.GB};
        sweep_track<1, EXPM, NP>(a, tile, slot, dt, gl, L.GB, g0, gb, src, tally);
      }
    }
  }
  __syncthreads();
  if (L.tally64) {
    for (int i = threadIdx.x; i < nl; i += blockDim.x) {
      const unsigned long long v = lds_tally[i];
      const int r = i / L.GB, l = i - r * L.GB;
      if (l < gb && v != 0ull) atomicAdd(&L.tally64[(size_t)r * a.G + g0 + l], v);
    }
    return;
  }
  unsigned long long* out = L.partials + ((size_t)sub * L.blocks_per_sub + bslot) * nl;
  for (int i = threadIdx.x; i < nl; i += blockDim.x) out[i] = lds_tally[i];
}

template <int EXPM, int QLDS, int NP>
__global__ __launch_bounds__(1024) void sweep2d_lds_kernel(SweepArgs a, LdsParams L) {
  sweep2d_lds_kernel_body<EXPM, QLDS, NP>(a, L);
}

// 3D z-stack sweep.  A group is the set of 3D tracks over a pair of
// neighbouring parallel 2D tracks (their z-stacks, every polar angle); its
// segments touch only the FSRs along those two 2D lines, one per (2D region,
// axial layer), which get local slots 0..n-1.
// Block (group, group subset) keeps the fixed-point tally and the
// {Q/sigma_t, 1/V} table of its slots in LDS, sweeps the group's tiles, and
// adds its tally to the global int64 tally (exact, order-independent).  The
// segment keys of the stream are local slots | material << 24.  Sources are
// read from LDS only: the pipeline's look-ahead reads keys of rows past the
// tile, which may be slots of another group, so a key is only ever used as
// an index into the block's own LDS slice (never to address global memory).
struct StackParams {
  const int4* grp;             // [ngrp] first tile, tiles, first slot, slots
  const int32_t* slot_fsr;     // global FSR of each slot
  unsigned long long* tally;   // [N*G] fixed-point sums of contribution / V (atomic flush)
  unsigned long long* partial; // [slots][prow] each slot's sums, written once per sweep, or null
  int prow;                    // row stride: G rounded up to 8 (64-byte rows)
  const float* qsub;           // [nsub][N][GB] Q/sigma_t, subset-major
  const float* inv_vol;        // [N]
  const double* sigma_t;       // [M][G]
  int M;
  int GB;
  int nsub;
  int ngrp;
  int max_slots;               // LDS layout stride (largest group)
};

template <int EXPM>
__global__ __launch_bounds__(1024) void sweep3d_stack_kernel(SweepArgs a, StackParams L) {
  extern __shared__ unsigned long long lds64[];
  if (a.state->converged) return;
  // the nsub subset blocks of one group are dispatched to the same XCD
  // (blocks are dealt to the 8 XCDs round-robin), so they share its L2
  const int x = blockIdx.x % kNumXcd, j = blockIdx.x / kNumXcd;
  const int grp = x + kNumXcd * (j / L.nsub);
  const int sub = j % L.nsub;
  if (grp >= L.ngrp) return;  // uniform per block
  const int4 gi = L.grp[grp];
  const int g0 = sub * L.GB;
  const int gb = min(L.GB, a.G - g0);
  const int nl = L.max_slots * L.GB;
  const int ns = gi.w * L.GB;
  const float s32 = (float)(a.state->tally_scale * 0x1p-32);
  unsigned long long* lds_tally = lds64;
  float2* lds_qi = reinterpret_cast<float2*>(lds64 + nl);
  float* lds_sig = reinterpret_cast<float*>(lds_qi + nl);
  const float* qg = L.qsub + (size_t)sub * a.N * L.GB;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    lds_tally[i] = 0ull;
    const int s = i / L.GB, l = i - s * L.GB;
    const int f = L.slot_fsr[gi.z + s];
    lds_qi[i] = make_float2(qg[(size_t)f * L.GB + l], L.inv_vol[f] * s32);
  }
  for (int i = threadIdx.x; i < L.M * L.GB; i += blockDim.x) {
    const int m = i / L.GB, l = i - m * L.GB;
    lds_sig[i] = (float)L.sigma_t[(size_t)m * a.G + g0 + min(l, gb - 1)];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int slot = lane / L.GB;
  const int gl = lane - slot * L.GB;
  if (slot < a.tpw) {
    LdsFixedTally tally{lds_tally, L.GB};
    LdsSource src{lds_qi, lds_sig, L.GB};
    for (int tile = __builtin_amdgcn_readfirstlane(gi.x + wib); tile < gi.x + gi.y; tile += wpb) {
      const int dt = a.work[tile * a.tpw + slot];
      if (dt < 0) continue;
      sweep_track<1, EXPM, 1>(a, tile, slot, dt, gl, L.GB, g0, gb, src, tally);
    }
  }
  __syncthreads();
  if (L.partial) {
    // every (slot, group) of the sweep has exactly one block: its sums are
    // written, zero or not, and stack_reduce_kernel adds an FSR's slots.
    // (Rows of G sums per slot: the reduction then reads one row per slot.
    // Writing each block's slice contiguously instead, [subset][slot][GB],
    // cut these writes by a third but scattered the reduction's reads: 1.2 GB
    // fetched against 0.24 GB.)
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
      const int s = i / L.GB, l = i - s * L.GB;
      if (l < gb) L.partial[(size_t)(gi.z + s) * L.prow + g0 + l] = lds_tally[i];
    }
    return;
  }
  for (int i = threadIdx.x; i < ns; i += blockDim.x) {
    const int s = i / L.GB, l = i - s * L.GB;
    const unsigned long long v = lds_tally[i];
    if (l < gb && v != 0ull)
      atomicAdd(&L.tally[(size_t)L.slot_fsr[gi.z + s] * a.G + g0 + l], v);
  }
}

// tally[r][g] = sum over FSR r's z-stack slots of their sums (an exact
// integer sum, so the bits equal those of the atomic flush).  off [N+1] and
// slots list each FSR's slots; reads are one slot row of G sums per step.
// Rows of prow sums (G rounded up to 8, 64 bytes at 7 groups) keep the 7
// blocks' pieces of a row in one 64-byte line.
__global__ void stack_reduce_kernel(const unsigned long long* __restrict__ partial,
                                    const int64_t* __restrict__ off, const int32_t* __restrict__ slots,
                                    unsigned long long* __restrict__ tally, int64_t N, int G,
                                    int prow, const IterState* state) {
  if (state->converged) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N * G) return;
  const int64_t r = i / G;
  const int g = (int)(i - r * G);
  const unsigned long long* pl = partial + g;
  // four independent loads in flight (integer sums: the order is free)
  unsigned long long t0 = 0ull, t1 = 0ull, t2 = 0ull, t3 = 0ull;
  int64_t q = off[r];
  const int64_t e = off[r + 1];
  for (; q + 4 <= e; q += 4) {
    const int s0 = slots[q], s1 = slots[q + 1], s2 = slots[q + 2], s3 = slots[q + 3];
    t0 += pl[(size_t)s0 * prow];
    t1 += pl[(size_t)s1 * prow];
    t2 += pl[(size_t)s2 * prow];
    t3 += pl[(size_t)s3 * prow];
  }
  for (; q < e; q++) t0 += pl[(size_t)slots[q] * prow];
  tally[i] = (t0 + t1) + (t2 + t3);
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
}  // namespace

// Arguments of the per-FSR kernels (declared in hip_solver.h).
struct FsrArgs {
  int64_t N;
  int G;
  const int32_t* fsr_mat;
  const double* fsr_vol;
  const double* sigma_t;
  const double* sigma_s;      // dest-major
  const double* nsf;          // nu-sigma-f: keff, normalisation, residual
  const double* chi;          // fission-source spectrum (nu-sigma-f when adjoint)
  const double* fis_nsf;      // fission-source production (chi when adjoint)
  const double* sigma_a;      // [M*G] absorption (neutron-balance keff)
  const int32_t* fissionable;
  const double* fixed;        // nullable
  double* phi;                // [N*G] scalar flux of the current sweep
  double* phi_old[2];         // normalised fluxes, ping-pong (state->old_sel is current)
  double* Q;
  float2* qs;                 // [N*G] (global-atomic sweep), or null
  float* qsub;                // [nsub][N][GB] Q/sigma_t (LDS sweep), or null
  int qsub_groups;            // GB
  float* tally;               // [tally_copies][N*G] (global-atomic sweep)
  int tally_copies;
  const unsigned long long* lds_partials;  // [nsub][bps][N*GB] fixed point (LDS sweep), or null
  unsigned long long* tally_i64;           // [N*G] fixed point (3D z-stack sweep, LDS flush), or null
  int lds_bps;
  // per-segment store of the global-tally sweep: [planes][slots][seg_stride]
  // floats (planes flat, x, y); FSR r's segments are rows
  // csr_off[r] .. csr_off[r+1] - 1, summed here in that order; or null
  const float* segout;
  const int64_t* csr_off;
  int seg_stride;
  size_t seg_plane;
  // post_sweep_kernel's threads run in the LDS sweep's subset-major order
  // (sub, r, l) (both of its flushes), so its FP64 partial sums are added in
  // the same order whichever flush ran
  int lds_order;
  const float* leak;          // [2T*G] vacuum leakage per directed track and group, or null
  int64_t leak_n;
  double* partials;           // [kNumPartials][pstride] per-block sums
  int pstride;
  IterState* state;
  int allow_negative;
  int stab_type;              // -1 off, else DIAGONAL / YAMAMOTO / GLOBAL
  double stab_factor;
  const double* stab_ratio;   // [G] YAMAMOTO: factor * max |sigma_s(g->g) / sigma_t(g)|
  // linear source (null / 0 for the flat source)
  int linear;
  float* tally_xy;            // [tally_copies][2][N*G] sweep moment tallies (x, y planes)
  double* phi_xy;             // [N*G*2] flux moments of the current sweep
  double* phi_xy_n;           // [N*G*2] normalised flux moments
  double* Qxy;                // [N*G*2] source moments
  float4* qls;                // [N*G] (Q, Q_x, Q_y, sigma_t) for the sweep
  const double* ilem;         // [N*3] inverse linear expansion matrix
  const double* src_c;        // [N][3][G] source constants
  // scattering contraction computed beforehand (scatter_mfma_kernel /
  // scatter_fp32_kernel), or null for the FP64 loops of update_kernel:
  // [C][N][G+1], column gd = sum_g sigma_s(g->gd) x(g), column G =
  // sum_g nu-sigma-f(g) x(g), of x = phi (c = 0) and the moments (c = 1, 2)
  const float* scat;
  const double* ss_colsum;    // [M*G] sum over destinations of sigma_s(g->d)
  // residual against a fixed reference flux [N*G] instead of the old flux
  // (setResidualByReference), or null
  const double* ref_flux;
};

// Device buffers of one sweep that replace the solver's own (the
// allocation-free sweep): Q/sigma_t tables, tallies and the iteration state.
struct SweepBuffers {
  const float2* qs;               // [N*G] (Q/sigma_t, sigma_t): global-tally sweep
  float* qsub;                    // [nsub][N][GB] Q/sigma_t: LDS and z-stack sweeps
  float* tally;                   // [copies][N*G]: global-tally sweep
  unsigned long long* tally_i64;  // [N*G] fixed point: LDS (atomic flush) and z-stack sweeps
  IterState* state;
};

namespace {

// per-block partial sums written by post_sweep_kernel / update_kernel
constexpr int kPartFission = 0;     // sum V nu-sigma-f phi (unstabilised)
constexpr int kPartAbsorption = 1;  // sum V sigma_a phi
constexpr int kPartNormFission = 2; // sum V nu-sigma-f phi after stabilisation
constexpr int kPartLeakage = 3;     // sum of vacuum leakage
constexpr int kPartResidual = 4;    // residual terms (update_kernel, fsr_residual_kernel)
constexpr int kNumPartials = 5;

// stabilizationType (src/Solver.h:103-113)
constexpr int kStabDiagonal = 0;
constexpr int kStabYamamoto = 1;
constexpr int kStabGlobal = 2;

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

// The sum of x[0..n) formed by one wave alone, in every lane with the same
// bits: lane l adds x[l], x[l + 64], ... in order, then six butterfly steps
// add the lanes pairwise (each step adds the same two values in every lane of
// a pair, and IEEE addition commutes).  No LDS and no barrier, so the other
// loads of the calling block stay in flight.
__device__ __forceinline__ double wave_sum_all(const double* x, int n) {
  const int lane = threadIdx.x & 63;
  double v = 0.0;
#pragma unroll 4
  for (int i = lane; i < n; i += 64) v += x[i];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// block_sum of two values in one pass (one pair of barriers): each result
// has block_sum's bits and is valid in thread 0.  MX: the second is a
// maximum, taken over the same tree.
template <bool MX>
__device__ void block_reduce2(double& a, double& b) {
  __shared__ double sh[2][kReduceBlock / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    const double o = __shfl_down(b, off, 64);
    b = MX ? fmax(b, o) : b + o;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][wid] = a;
    sh[1][wid] = b;
  }
  __syncthreads();
  double ra = 0, rb = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
      ra += sh[0][i];
      rb = MX ? fmax(rb, sh[1][i]) : rb + sh[1][i];
    }
  __syncthreads();
  a = ra;
  b = rb;
}

// Tally of FSR r, group g from the per-segment store: the sum over the FSR's
// rows, in FP64, in a fixed order (eight interleaved partial sums keep eight
// loads in flight), so it does not depend on scheduling.  The threads of one
// FSR read its contiguous rows of G floats one after the other.
__device__ __forceinline__ double segment_sum(const FsrArgs& a, int64_t r, int g, int plane) {
  const float* base = a.segout + (size_t)plane * a.seg_plane + g;
  const int64_t e0 = a.csr_off[r], e1 = a.csr_off[r + 1];
  constexpr int kAcc = 8;
  double acc[kAcc];
#pragma unroll
  for (int u = 0; u < kAcc; u++) acc[u] = 0.0;
  int64_t e = e0;
  for (; e + kAcc <= e1; e += kAcc) {
#pragma unroll
    for (int u = 0; u < kAcc; u++) acc[u] += (double)base[(size_t)(e + u) * a.seg_stride];
  }
  for (; e < e1; e++) acc[0] += (double)base[(size_t)e * a.seg_stride];
#pragma unroll
  for (int w = kAcc / 2; w > 0; w /= 2)
#pragma unroll
    for (int u = 0; u < w; u++) acc[u] += acc[u + w];
  return acc[0];
}

// One thread per (FSR, group), in the LDS sweep's subset-major order
// (sub, r, l) so the fixed-point partials are read coalesced.
//   from_tally=1: the tally of the sweep becomes the scalar flux
//     phi = tally/(sigma_t V) + 4 pi Q / sigma_t with the negative-flux clamp
//     (addSourceToScalarFlux, src/CPUSolver.cpp:2608-2659); the fixed-point
//     partials are summed exactly here (the global-atomic path sums and clears
//     its FP32 tally copies), then transport stabilisation
//     (computeStabilizingFlux / stabilizeFlux, src/CPUSolver.cpp:2665-2805)
//     is applied from iteration 1 on, from the previous normalised flux.
//   from_tally=0: phi is the current array (initial guess).
// Per-block partials: fission rate before and after stabilisation,
// absorption rate and vacuum leakage (computeKeff, src/CPUSolver.cpp:2258-2328).
// lds_bps: blocks per subset whose partial slices are summed (the LDS
// sweep's launch; a.lds_bps unless a batched launch used fewer)
__device__ __forceinline__ void post_sweep_kernel_body(const FsrArgs& a, int from_tally, int balance,
                                                       int lds_bps) {
  if (a.state->converged) return;
  const int G = a.G;
  const int GB = a.qsub_groups;
  const int64_t NG = a.N * G;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double rate = 0.0, rate_s = 0.0, absr = 0.0, leak = 0.0;
  int64_t neg = 0;
  // LDS sweep: (sub, r, l) -> (r, g), the order of its partials; otherwise
  // FSR-major (r, g), the order of the tally and flux arrays
  const int64_t nl = a.N * GB;
  int64_t sub = 0, j = 0, r;
  int g;
  bool valid;
  if (a.lds_order) {
    sub = i / nl;
    j = i - sub * nl;
    r = j / GB;
    g = (int)(sub * GB + (j - r * GB));
    valid = i < (int64_t)((G + GB - 1) / GB) * nl && g < G;
  } else {
    r = i / G;
    g = (int)(i - r * G);
    valid = i < NG;
  }
  if (valid) {
    const int m = a.fsr_mat[r];
    const double vol_true = a.fsr_vol[r];
    const double vol = vol_true < kFltEpsilon ? 1e30 : vol_true;
    const double st = a.sigma_t[(size_t)m * G + g];
    const size_t k = (size_t)r * G + g;
    double f;
    if (from_tally) {
      if (a.lds_partials) {
        const unsigned long long* p = a.lds_partials + (size_t)sub * lds_bps * nl + j;
        // exact (two's complement) sums, so any order gives the same bits;
        // sixteen independent accumulators keep sixteen loads in flight
        constexpr int kAcc = 16;
        unsigned long long acc[kAcc];
#pragma unroll
        for (int u = 0; u < kAcc; u++) acc[u] = 0ull;
        int b = 0;
        for (; b + kAcc <= lds_bps; b += kAcc) {
#pragma unroll
          for (int u = 0; u < kAcc; u++) acc[u] += p[(size_t)(b + u) * nl];
        }
        for (; b < lds_bps; b++) acc[0] += p[(size_t)b * nl];
#pragma unroll
        for (int w = kAcc / 2; w > 0; w /= 2)
#pragma unroll
          for (int u = 0; u < w; u++) acc[u] += acc[u + w];
        const unsigned long long t = acc[0];
        // the fixed-point tally holds sum(contribution) / V
        f = ((double)(long long)t / a.state->tally_scale) / st + kFourPi * a.Q[k] / st;
      } else if (a.tally_i64) {
        const unsigned long long t = a.tally_i64[k];
        a.tally_i64[k] = 0ull;
        f = ((double)(long long)t / a.state->tally_scale) / st + kFourPi * a.Q[k] / st;
      } else if (a.linear) {
        // CPULSSolver::addSourceToScalarFlux (src/CPULSSolver.cpp:787-882)
        double t = 0.0, tx = 0.0, ty = 0.0;
        if (a.segout) {
          t = segment_sum(a, r, g, 0);
          tx = segment_sum(a, r, g, 1);
          ty = segment_sum(a, r, g, 2);
        } else {
          for (int c = 0; c < a.tally_copies; c++) {
            t += (double)a.tally[c * NG + k];
            tx += (double)a.tally_xy[2 * c * NG + k];
            ty += (double)a.tally_xy[(2 * c + 1) * NG + k];
            a.tally[c * NG + k] = 0.f;
            a.tally_xy[2 * c * NG + k] = 0.f;
            a.tally_xy[(2 * c + 1) * NG + k] = 0.f;
          }
        }
        f = t / vol;
        f += kFourPi * a.Q[k];
        f /= st;
        const double* c3 = a.src_c + (size_t)r * 3 * G;
        const double fc = kFourPi * 2;
        const double qx = a.Qxy[2 * k], qy = a.Qxy[2 * k + 1];
        double mx = tx / vol;
        mx += fc * qx * c3[g];
        mx += fc * qy * c3[2 * G + g];
        double my = ty / vol;
        my += fc * qx * c3[2 * G + g];
        my += fc * qy * c3[G + g];
        mx /= st;
        my /= st;
        if (f < 0.0 && !a.allow_negative) {
          f = fmax(a.phi_old[a.state->old_sel][k], kFluxEpsilon);
          mx = 0.0;
          my = 0.0;
          neg++;
        }
        a.phi_xy[2 * k] = mx;
        a.phi_xy[2 * k + 1] = my;
      } else if (a.segout) {
        f = segment_sum(a, r, g, 0) / (st * vol) + kFourPi * a.Q[k] / st;
      } else {
        double t = 0.0;
        for (int c = 0; c < a.tally_copies; c++) {
          t += (double)a.tally[c * NG + k];
          a.tally[c * NG + k] = 0.f;
        }
        f = t / (st * vol) + kFourPi * a.Q[k] / st;
      }
      if (f < 0.0 && !a.allow_negative) {
        f = kFluxEpsilon;
        neg++;
      }
    } else {
      f = a.phi[k];
    }
    const double nsf = a.nsf[(size_t)m * G + g];
    rate = nsf * f * vol_true;
    if (balance) absr = a.sigma_a[(size_t)m * G + g] * f * vol_true;
    if (from_tally && a.stab_type >= 0 && a.state->iteration > 0) {
      const double old = a.phi_old[a.state->old_sel][k];
      const double rho = a.stab_factor;
      if (a.stab_type == kStabDiagonal) {
        const double ss = a.sigma_s[((size_t)m * G + g) * G + g];
        if (ss < 0.0) {
          const double s = -old * rho * ss / st;
          f += s;
          f /= (1.0 - rho * ss / st);
        }
      } else if (a.stab_type == kStabYamamoto) {
        const double mr = a.stab_ratio[g];
        const double s = old * mr;
        f += s;
        f /= (1 + mr);
      } else if (a.stab_type == kStabGlobal) {
        const double s = (1.0 / rho - 1.0) * old;
        f += s;
        f *= rho;
      }
    }
    rate_s = nsf * f * vol_true;
    if (from_tally) a.phi[k] = f;
  }
  if (balance && a.leak)
    for (int64_t q = i; q < a.leak_n; q += (int64_t)gridDim.x * blockDim.x) leak += (double)a.leak[q];
  double s0 = rate, s2 = rate_s;
  block_reduce2<false>(s0, s2);
  double s1 = 0.0, s3 = 0.0;
  if (balance) {
    s1 = absr;
    s3 = leak;
    block_reduce2<false>(s1, s3);
  }
  if (threadIdx.x == 0) {
    a.partials[kPartFission * a.pstride + blockIdx.x] = s0;
    a.partials[kPartNormFission * a.pstride + blockIdx.x] = s2;
    a.partials[kPartAbsorption * a.pstride + blockIdx.x] = s1;
    a.partials[kPartLeakage * a.pstride + blockIdx.x] = s3;
  }
  if (neg) atomicAdd((unsigned long long*)&a.state->negative_fluxes, (unsigned long long)neg);
}

__global__ __launch_bounds__(kReduceBlock) void post_sweep_kernel(FsrArgs a, int from_tally,
                                                                 int balance) {
  post_sweep_kernel_body(a, from_tally, balance, a.lds_bps);
}

// phase 0: initial guess normalisation (k unchanged); phase 1: computeKeff
// (k *= F/N, or k = F / (A + L) from the neutron balance; src/CPUSolver.cpp:
// 2258-2328) then normalizeFluxes' factor N/F (src/CPUSolver.cpp:1860-1931)
// from the (stabilised) fluxes.
// The sums of post_sweep_kernel's per-block partials for the keff step inside
// update_kernel: formed by every wave on its own, so every thread of every
// block gets the same bits (keff_kernel, one block, sums the same partials
// with block_sum; the two orders differ in the last bits, and a solve always
// takes the same one): fission before and after stabilisation, and with the
// neutron balance absorption and leakage.  One order for every caller, so a
// block of update_kernel that takes keff_kernel's place gets its bits.
struct KeffSums {
  double fiss, fiss_norm, absr, leak;
};
// The keff step runs inside update_kernel only up to this many post-sweep
// partials: every update block reads all of them, so the reads grow with the
// square of the block count (at 254 partials, the 17x17 lattice, the fused
// step saves 2.1 us an iteration; at 1 052, the 7-group mini-core, it breaks
// even; far above, 70 groups, it would read gigabytes).
constexpr int kFuseKeffMaxParts = 512;
__device__ __forceinline__ KeffSums keff_sums(const double* partials, int nparts, int pstride,
                                              int balance) {
  KeffSums r;
  r.fiss = wave_sum_all(partials + (size_t)kPartFission * pstride, nparts);
  r.fiss_norm = wave_sum_all(partials + (size_t)kPartNormFission * pstride, nparts);
  r.absr = 0.0;
  r.leak = 0.0;
  if (balance) {
    r.absr = wave_sum_all(partials + (size_t)kPartAbsorption * pstride, nparts);
    r.leak = wave_sum_all(partials + (size_t)kPartLeakage * pstride, nparts);
  }
  return r;
}

__device__ __forceinline__ void keff_kernel_body(const double* partials, int nparts,
                                                           int pstride, IterState* state,
                                                           double num_fsrs, int phase,
                                                           int balance) {
  if (state->converged) return;
  double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
    v0 += partials[kPartFission * pstride + i];
    v2 += partials[kPartNormFission * pstride + i];
    if (balance) {
      v1 += partials[kPartAbsorption * pstride + i];
      v3 += partials[kPartLeakage * pstride + i];
    }
  }
  const double fiss = block_sum(v0);
  const double fiss_norm = block_sum(v2);
  double absr = 0.0, leak = 0.0;
  if (balance) {
    absr = block_sum(v1);
    leak = block_sum(v3);
  }
  if (threadIdx.x == 0) {
    if (phase == 1) {
      if (balance)
        state->k = fiss / (absr + leak);
      else
        state->k = state->k * fiss / num_fsrs;
    }
    state->fission_total = fiss_norm;
    state->norm_factor = num_fsrs / fiss_norm;
    state->psi_scale = num_fsrs / fiss_norm;
  }
}

__global__ __launch_bounds__(kReduceBlock) void keff_kernel(const double* partials, int nparts,
                                                           int pstride, IterState* state,
                                                           double num_fsrs, int phase,
                                                           int balance) {
  keff_kernel_body(partials, nparts, pstride, state, num_fsrs, phase, balance);
}

// From this many groups on, the per-FSR source residual terms are summed by
// fsr_residual_kernel (one lane per FSR) instead of one lane of every FSR's
// G lanes in update_kernel.
constexpr int kFsrResidualGroups = 32;

// FISSION_SOURCE / TOTAL_SOURCE residual term of FSR r (computeResidual,
// src/CPUSolver.cpp:2113-2252), from the new flux phi * f and the old flux.
__device__ double source_residual_term(const FsrArgs& a, int64_t r, int res_type, int normalize,
                                       double f, double k) {
  const int G = a.G;
  const int m = a.fsr_mat[r];
  const bool fis = a.fissionable[m] != 0;
  const double* nsf = a.nsf + (size_t)m * G;
  const double* phi = a.phi + (size_t)r * G;
  const double* old = (a.ref_flux ? a.ref_flux : a.phi_old[a.state->old_sel]) + (size_t)r * G;
  double nw = 0.0, od = 0.0;
  if (fis) {
    for (int g = 0; g < G; g++) {
      nw += (normalize ? phi[g] * f : phi[g]) * nsf[g];
      od += old[g] * nsf[g];
    }
    if (res_type == MOC_TOTAL_SOURCE) {
      nw /= k;
      od /= k;
    }
  }
  if (res_type == MOC_TOTAL_SOURCE) {
    // sum over d and g of sigma_s(g->d) x(g) = sum_g colsum(g) x(g)
    const double* cs = a.ss_colsum + (size_t)m * G;
    for (int g = 0; g < G; g++) {
      nw += cs[g] * (normalize ? phi[g] * f : phi[g]);
      od += cs[g] * old[g];
    }
  }
  if ((res_type == MOC_TOTAL_SOURCE || fis) && od > 0.0) {
    const double d = (nw - od) / od;
    return d * d;
  }
  return 0.0;
}

// Source residual terms for G >= kFsrResidualGroups, one thread per FSR,
// after update_kernel (which leaves phi and the current old flux as they
// were and writes zero partials); per-block sums replace its partials.
__device__ __forceinline__ void fsr_residual_kernel_body(const FsrArgs& a, int normalize,
                                                                   int res_type) {
  if (a.state->converged) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double res = 0.0;
  if (r < a.N)
    res = source_residual_term(a, r, res_type, normalize, normalize ? a.state->norm_factor : 1.0,
                               a.state->k);
  const double s = block_sum(res);
  if (threadIdx.x == 0) a.partials[kPartResidual * a.pstride + blockIdx.x] = s;
}

__global__ __launch_bounds__(kReduceBlock) void fsr_residual_kernel(FsrArgs a, int normalize,
                                                                   int res_type) {
  fsr_residual_kernel_body(a, normalize, res_type);
}

// Source kinds of update_kernel
constexpr int kSrcTotal = 0;    // computeFSRSources (src/CPUSolver.cpp:1939-2023)
constexpr int kSrcFission = 1;  // computeFSRFissionSources (src/CPUSolver.cpp:2030-2066)
constexpr int kSrcScatter = 2;  // computeFSRScatterSources (src/CPUSolver.cpp:2072-2104)

// One thread per (FSR, destination group): normalise (optional), residual
// partials, store the normalised flux as the next "old" flux, sources of the
// next sweep.  Reads phi and the current old flux only, writes the other
// old buffer, so threads of one FSR never see each other's writes.
// keff_parts > 0: this launch also does keff_kernel's phase-1 step (eigen
// iteration) from the first keff_parts post-sweep partials.  Every block
// forms the sums with keff_sums (the same bits in every block) and block 0
// stores k and the normalisation factor.  k_prev is read for k: the previous
// residual step (or the solve's set-up) made them equal, and unlike k it is
// not written during the launch.  FUSE = 0 compiles the step out (keff_parts
// is then ignored), so launches without it keep the plain kernel.
template <int FUSE>
__device__ __forceinline__ void update_kernel_body(const FsrArgs& a, int normalize,
                                                             int res_type, int do_residual,
                                                             int compute_sources,
                                                             int source_iteration_offset,
                                                             int source_kind, int keff_parts = 0,
                                                             int balance = 0,
                                                             double num_fsrs = 0.0) {
  if (a.state->converged) return;
  const int G = a.G;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double res = 0.0;
  double qmax = 0.0;
  int64_t negsrc = 0;
  double k_new = 0.0, f_new = 0.0;
  const bool fused = FUSE && keff_parts > 0;
  if (fused) {
    const KeffSums ks = keff_sums(a.partials, keff_parts, a.pstride, balance);
    k_new = balance ? ks.fiss / (ks.absr + ks.leak) : a.state->k_prev * ks.fiss / num_fsrs;
    f_new = num_fsrs / ks.fiss_norm;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.state->k = k_new;
      a.state->fission_total = ks.fiss_norm;
      a.state->norm_factor = f_new;
      a.state->psi_scale = f_new;
    }
  }
  if (i < a.N * G) {
    const int64_t r = i / G;
    const int gd = (int)(i - r * G);
    const int m = a.fsr_mat[r];
    const double f = normalize ? (fused ? f_new : a.state->norm_factor) : 1.0;
    const double k = fused ? k_new : a.state->k;
    const double st = a.sigma_t[(size_t)m * G + gd];
    const double* ss = a.sigma_s + (size_t)m * G * G;
    // fission source: chi(gd) sum_g fnsf(g) x(g) (transposed roles when adjoint)
    const double* fnsf = a.fis_nsf + (size_t)m * G;
    const bool fis = a.fissionable[m] != 0;
    const double* phi = a.phi + (size_t)r * G;
    const int sel = a.state->old_sel;
    const double* old = a.phi_old[sel] + (size_t)r * G;
    double* nxt = a.phi_old[sel ^ 1] + (size_t)r * G;
    const double phin = normalize ? phi[gd] * f : phi[gd];
    nxt[gd] = phin;
    if (do_residual) {
      // computeResidual (src/CPUSolver.cpp:2113-2252)
      if (res_type == MOC_SCALAR_FLUX) {
        const double ref = a.ref_flux ? a.ref_flux[i] : old[gd];
        if (ref > 0.0) {
          const double d = (phin - ref) / ref;
          res = d * d;
        }
      } else if (gd == 0 && G < kFsrResidualGroups) {
        res = source_residual_term(a, r, res_type, normalize, f, k);
      }
    }
    if (compute_sources) {
      const int it = a.state->iteration + source_iteration_offset;
      double q;
      const float* sct = a.scat ? a.scat + (size_t)r * (G + 1) : nullptr;
      if (source_kind == kSrcFission) {
        // chi(gd) nu-sigma-f(g) phi(g), without 1/k or fixed sources
        double fsrc = 0.0;
        if (fis) {
          if (sct)
            fsrc = a.chi[(size_t)m * G + gd] * (double)sct[G];
          else
            for (int g = 0; g < G; g++) fsrc += a.chi[(size_t)m * G + gd] * fnsf[g] * phi[g];
        }
        q = fsrc * kOneOverFourPi;
      } else if (source_kind == kSrcScatter) {
        double sc = 0.0;
        if (sct)
          sc = (double)sct[gd];
        else
          for (int g = 0; g < G; g++) sc += ss[(size_t)gd * G + g] * phi[g];
        q = sc * kOneOverFourPi;
      } else {
        double new_fis = 0.0, sc = 0.0;
        if (sct) {
          // the contraction is linear: normalising after it is the same
          if (fis) new_fis = (double)sct[G] * f;
          sc = (double)sct[gd] * f;
        } else {
          if (fis)
            for (int g = 0; g < G; g++) new_fis += fnsf[g] * (normalize ? phi[g] * f : phi[g]);
          const double* row = ss + (size_t)gd * G;
          for (int g = 0; g < G; g++) sc += row[g] * (normalize ? phi[g] * f : phi[g]);
        }
        q = (fis ? a.chi[(size_t)m * G + gd] * new_fis : 0.0) / k + sc;
        if (a.fixed) q += a.fixed[(size_t)r * G + gd];
        q *= kOneOverFourPi;
        if (q < 0.0) {
          negsrc++;
          if (it < 30 && !a.allow_negative) q = kFluxEpsilon;
        }
      }
      a.Q[i] = q;
      qmax = fabs(q / st);
      if (a.qs) a.qs[i] = make_float2((float)(q / st), (float)st);
      if (a.linear) {
        // source moments (CPULSSolver::computeFSRSources, src/CPULSSolver.cpp:
        // 386-524) from the normalised flux moments
        const double* mom = a.phi_xy + (size_t)r * G * 2;
        const double fm = normalize ? f : 1.0;
        double src[2];
        for (int c = 0; c < 2; c++) {
          double fis_c = 0.0, sc_c = 0.0;
          if (sct) {
            const float* sm = a.scat + ((size_t)(c + 1) * a.N + r) * (G + 1);
            if (fis) fis_c = a.chi[(size_t)m * G + gd] * ((double)sm[G] * fm) / k;
            sc_c = (double)sm[gd] * fm;
          } else {
            if (fis) {
              const double cg = a.chi[(size_t)m * G + gd];
              for (int g = 0; g < G; g++) fis_c += cg * fnsf[g] * (mom[2 * g + c] * fm);
              fis_c /= k;
            }
            const double* row = ss + (size_t)gd * G;
            for (int g = 0; g < G; g++) sc_c += row[g] * (mom[2 * g + c] * fm);
          }
          src[c] = sc_c + fis_c;
        }
        a.phi_xy_n[2 * i] = mom[2 * gd] * fm;
        a.phi_xy_n[2 * i + 1] = mom[2 * gd + 1] * fm;
        double qx = 0.0, qy = 0.0;
        if (a.allow_negative || q > 10 * kFluxEpsilon || it > 29) {
          const double* il = a.ilem + (size_t)r * 3;
          qx = kOneOverFourPi / 2 * (il[0] * src[0] + il[2] * src[1]);
          qy = kOneOverFourPi / 2 * (il[2] * src[0] + il[1] * src[1]);
        }
        a.Qxy[2 * i] = qx;
        a.Qxy[2 * i + 1] = qy;
        a.qls[i] = make_float4((float)q, (float)qx, (float)qy, (float)st);
      }
      if (a.qsub) {
        const int sb = gd / a.qsub_groups, l = gd - sb * a.qsub_groups;
        a.qsub[((size_t)sb * a.N + r) * a.qsub_groups + l] = (float)(q / st);
      }
    }
  }
  double s = res, qm = qmax;
  block_reduce2<true>(s, qm);
  if (threadIdx.x == 0) {
    a.partials[kPartResidual * a.pstride + blockIdx.x] = s;
    if (qm > 0.0) {
      // round up so the float bound is not below the double value
      const float qf = __uint_as_float(__float_as_uint((float)qm) + 1u);
      raise_max_bits(&a.state->qmax_bits, qf);
    }
  }
  if (negsrc) atomicAdd((unsigned long long*)&a.state->negative_sources, (unsigned long long)negsrc);
}

template <int FUSE>
__global__ __launch_bounds__(kReduceBlock) void update_kernel(FsrArgs a, int normalize,
                                                             int res_type, int do_residual,
                                                             int compute_sources,
                                                             int source_iteration_offset,
                                                             int source_kind, int keff_parts,
                                                             int balance, double num_fsrs) {
  update_kernel_body<FUSE>(a, normalize, res_type, do_residual, compute_sources,
                           source_iteration_offset, source_kind, keff_parts, balance, num_fsrs);
}

// ---------------------------------------------------------------------------
// Scattering-source contraction on the matrix cores
// ---------------------------------------------------------------------------
// For every FSR r of material m and component c (the scalar flux, and for the
// linear source the two flux moments):
//   y[c][r][gd] = sum_g sigma_s[m](g->gd) x_c[r][g]    gd < G
//   y[c][r][G]  = sum_g nu-sigma-f[m](g)  x_c[r][g]    (fission rate)
// (computeFSRSources, src/CPUSolver.cpp:1939-2023; the moment sources of
// src/CPULSSolver.cpp:386-524).  FSRs are grouped by material into tiles of
// 16 rows; one wave multiplies a tile's 16 x K block of x by the material's
// K x (G+1) matrix with v_mfma_f32_16x16x16f16.  Operands are split into
// fp16 high and low parts (x = hi + lo, each row of x scaled by a power of
// two so its largest entry lies in [0.5, 1); each matrix column likewise),
// and hi*hi + hi*lo + lo*hi is accumulated in FP32: about 22 significant
// bits, against 11 for a plain fp16 product.
constexpr int kScatRows = 16;
constexpr int kMfmaMinGroups = 32;   // automatic choice: matrix cores from 32 groups on

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

}  // namespace

struct ScatArgs {
  int64_t N;
  int G;
  int ks_n;                    // K steps of 16 (ceil(G / 16))
  int nt_n;                    // column tiles of 16 (ceil((G + 1) / 16))
  int64_t ntiles;
  const int32_t* rows;         // [ntiles * 16] FSR ids, -1 pads a tile
  const int32_t* tile_mat;     // [ntiles] material of each tile
  const uint4* bfrag;          // [M][ks_n][nt_n][64] split fp16 operand fragments
  const float* colmul;         // [M][nt_n * 16] 2^e undoing each column's scaling
  const double* sigma_s;       // [M][G][G] dest-major (FP32 path)
  const double* nsf;           // [M][G] (FP32 path)
  const int32_t* fsr_mat;      // (FP32 path)
  const double* x[3];          // components; element g of row r at x[c][(r*G + g)*xs[c]]
  int xs[3];
  int C;
  float* out;                  // [C][N][G + 1]
  const IterState* state;      // nothing to do once converged (null: always run)
};

namespace {

__device__ __forceinline__ void split_f16(double v, _Float16& hi, _Float16& lo) {
  const float f = (float)v;
  hi = (_Float16)f;
  lo = (_Float16)(f - (float)hi);
}

// 2^-e for the largest magnitude mx = f 2^e, f in [0.5, 1) (1 for mx = 0),
// with e kept where both 2^e and 2^-e are normal floats.
__device__ __forceinline__ int scale_exponent(double mx) {
  if (!(mx > 0.0)) return 0;
  int e;
  (void)frexp(mx, &e);
  return e < -120 ? -120 : (e > 120 ? 120 : e);
}

// Operand fragments of the matrix side, one thread per (m, ks, nt, lane):
// lane l holds rows k = 16 ks + 4 (l / 16) + j (j = 0..3) of column
// n = 16 nt + l % 16, as v_mfma_f32_16x16x16f16 reads its B operand.
__global__ void scatter_frag_kernel(const double* sigma_s, const double* nsf, int M, int G,
                                    int ks_n, int nt_n, uint4* bfrag, float* colmul) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)M * ks_n * nt_n * 64;
  if (t >= total) return;
  const int lane = (int)(t & 63);
  int64_t u = t >> 6;
  const int nt = (int)(u % nt_n);
  u /= nt_n;
  const int ks = (int)(u % ks_n);
  const int m = (int)(u / ks_n);
  const int n = nt * 16 + (lane & 15);
  auto coef = [&](int k) -> double {
    if (k >= G || n > G) return 0.0;
    return n < G ? sigma_s[((size_t)m * G + n) * G + k] : nsf[(size_t)m * G + k];
  };
  double mx = 0.0;
  for (int k = 0; k < G; k++) mx = fmax(mx, fabs(coef(k)));
  const int e = scale_exponent(mx);
  const double s = ldexp(1.0, -e);
  if (ks == 0 && lane < 16) colmul[(size_t)m * nt_n * 16 + n] = ldexpf(1.f, e);
  _Float16 hi[4], lo[4];
  for (int j = 0; j < 4; j++) split_f16(coef(ks * 16 + 4 * (lane >> 4) + j) * s, hi[j], lo[j]);
  auto pack = [](_Float16 a, _Float16 b) {
    return (unsigned)__builtin_bit_cast(unsigned short, a) |
           ((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
  };
  bfrag[t] = make_uint4(pack(hi[0], hi[1]), pack(hi[2], hi[3]), pack(lo[0], lo[1]),
                        pack(lo[2], lo[3]));
}

// One wave per tile of 16 FSRs (4 waves per block).  Lane l loads row
// i = l % 16 of the tile at k = 16 ks + 4 (l / 16) + j, the A operand layout;
// the accumulator holds rows 4 (l / 16) + j of column 16 nt + l % 16.
__device__ __forceinline__ void scatter_mfma_kernel_body(const ScatArgs& a) {
  if (a.state && a.state->converged) return;
  const int lane = threadIdx.x & 63;
  const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= a.ntiles) return;  // whole waves
  const int G = a.G;
  const int i = lane & 15, q = lane >> 4;
  const int m = a.tile_mat[tile];
  const int r = a.rows[tile * kScatRows + i];
  const uint4* bf = a.bfrag + (size_t)m * a.ks_n * a.nt_n * 64 + lane;
  const float* cm = a.colmul + (size_t)m * a.nt_n * 16;
  constexpr int kChunk = 5;  // column tiles per pass (80 columns)
  for (int c = 0; c < a.C; c++) {
    const double* x = a.x[c];
    const int xs = a.xs[c];
    auto xval = [&](int k) -> double {
      return (r >= 0 && k < G) ? x[((size_t)r * G + k) * xs] : 0.0;
    };
    // row scale: largest |x| of the row over the 4 lanes holding it
    double mx = 0.0;
    for (int ks = 0; ks < a.ks_n; ks++)
      for (int j = 0; j < 4; j++) mx = fmax(mx, fabs(xval(ks * 16 + 4 * q + j)));
    mx = fmax(mx, __shfl_xor(mx, 16, 64));
    mx = fmax(mx, __shfl_xor(mx, 32, 64));
    const int e = scale_exponent(mx);
    const double s = ldexp(1.0, -e);
    float rmul[4];
    int rrow[4];
    for (int j = 0; j < 4; j++) {
      rmul[j] = ldexpf(1.f, __shfl(e, 4 * q + j, 64));
      rrow[j] = __shfl(r, 4 * q + j, 64);
    }
    float* out = a.out + (size_t)c * a.N * (G + 1);
    for (int nt0 = 0; nt0 < a.nt_n; nt0 += kChunk) {
      const int ntc = min(kChunk, a.nt_n - nt0);
      float4_t acc[kChunk];
      for (int t = 0; t < kChunk; t++) acc[t] = float4_t{0.f, 0.f, 0.f, 0.f};
      for (int ks = 0; ks < a.ks_n; ks++) {
        half4_t ah, al;
        for (int j = 0; j < 4; j++) {
          _Float16 h, l;
          split_f16(xval(ks * 16 + 4 * q + j) * s, h, l);
          ah[j] = h;
          al[j] = l;
        }
#pragma unroll
        for (int t = 0; t < kChunk; t++) {
          if (t < ntc) {
            const uint4 b = bf[((size_t)ks * a.nt_n + nt0 + t) * 64];
            const half4_t bh = __builtin_bit_cast(half4_t, make_uint2(b.x, b.y));
            const half4_t bl = __builtin_bit_cast(half4_t, make_uint2(b.z, b.w));
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(al, bh, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bl, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, bh, acc[t], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < kChunk; t++) {
        const int col = (nt0 + t) * 16 + i;
        if (t < ntc && col <= G) {
          const float cmul = cm[col];
          for (int j = 0; j < 4; j++)
            if (rrow[j] >= 0) out[(size_t)rrow[j] * (G + 1) + col] = acc[t][j] * rmul[j] * cmul;
        }
      }
    }
  }
}

__global__ __launch_bounds__(256) void scatter_mfma_kernel(ScatArgs a) {
  scatter_mfma_kernel_body(a);
}

// FP32 fallback of the same contraction on the vector ALUs: one thread per
// (component, FSR, column), FP32 products and sums.
__device__ __forceinline__ void scatter_fp32_kernel_body(const ScatArgs& a) {
  if (a.state && a.state->converged) return;
  const int G = a.G;
  const int64_t per = a.N * (G + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= per * a.C) return;
  const int c = (int)(t / per);
  const int64_t u = t - (int64_t)c * per;
  const int64_t r = u / (G + 1);
  const int col = (int)(u - r * (G + 1));
  const int m = a.fsr_mat[r];
  const double* coef = col < G ? a.sigma_s + ((size_t)m * G + col) * G : a.nsf + (size_t)m * G;
  const double* x = a.x[c] + (size_t)r * G * a.xs[c];
  float acc = 0.f;
  for (int g = 0; g < G; g++) acc = fmaf((float)coef[g], (float)x[(size_t)g * a.xs[c]], acc);
  a.out[t] = acc;
}

__global__ __launch_bounds__(256) void scatter_fp32_kernel(ScatArgs a) {
  scatter_fp32_kernel_body(a);
}

// Fixed-point scale of the next sweep's LDS tally.  Over one sweep,
// sum |w delta| / V over an FSR's segments <= 4 pi sigma_t (|psi|max +
// |Q/sigma_t|max), because sum over its segments (both directions, all polar
// angles) of w l / sin(theta) = 4 pi V and |delta| <= (1 - e^-tau)|psi - q|
// <= tau / sin(theta) |psi - q|, with psi a convex mix of its inflow and the
// sources.  Twice that bound is mapped to 2^60 (int64 range 2^63).
__device__ void set_tally_scale(IterState* st, double sigma_t_max, int take_qmax) {
  if (take_qmax) {
    st->qmax = __uint_as_float(st->qmax_bits);
    st->qmax_bits = 0u;
  }
  st->psi_max = __uint_as_float(st->psi_max_bits);
  st->psi_max_bits = 0u;
  const double bound =
      2.0 * kFourPi * sigma_t_max * ((double)st->psi_max * st->psi_scale + (double)st->qmax);
  st->tally_scale = tally_scale_for(bound, st->scale_cap);
}

// flip: 1 makes the old-flux buffer written by update_kernel the current one
__global__ void tally_scale_kernel(IterState* state, double sigma_t_max, int take_qmax, int flip) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    set_tally_scale(state, sigma_t_max, take_qmax);
    if (flip) state->old_sel ^= 1;
  }
}

// mode 0 eigen: stop when res < tol && |int(1e5 dk)| < 1 (src/Solver.cpp:1643-1681)
// mode 1/2 flux/source: stop when i > 1 && res < tol, _num_iterations = i.
__device__ __forceinline__ void residual_kernel_body(const double* partials, int nparts,
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
    state->old_sel ^= 1;
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
    // a batched solve stops at its iteration limit (converged = 2: stopped
    // without meeting the stopping rule)
    if (!state->converged && state->max_iters > 0 && state->iteration >= state->max_iters)
      state->converged = 2;
    set_tally_scale(state, sigma_t_max, take_qmax);
  }
}

__global__ __launch_bounds__(kReduceBlock) void residual_kernel(const double* partials, int nparts,
                                                               IterState* state, double norm,
                                                               double tol, int mode,
                                                               double sigma_t_max, int take_qmax) {
  residual_kernel_body(partials, nparts, state, norm, tol, mode, sigma_t_max, take_qmax);
}

// ---------------------------------------------------------------------------
// Batched iteration: the solves of one geometry advance together, one launch
// per step for all of them.  Solve j of a launch is row j of a device table
// of per-solve arguments (blockIdx.y = j); each kernel returns at entry for a
// solve that has stopped, so a row's work depends only on its own arrays.
// ---------------------------------------------------------------------------
}  // namespace

struct BatchSlot {
  SweepArgs sw[2];      // start fluxes in psi[0] / psi[1] (parity of the iteration)
  LdsParams lds;        // LDS sweep (lds_groups > 0)
  ScatArgs scat;        // scattering contraction on the matrix cores / FP32 (from 32 groups)
  FsrArgs fsr;
  int nparts_post;      // post_sweep_kernel blocks: partials summed by keff_kernel
  int nparts_upd;       // update_kernel blocks: partials summed by residual_kernel
  int nparts_fsr;       // fsr_residual_kernel blocks (0 below kFsrResidualGroups groups)
  int scat_path;        // MOC_SCATTER_* of the contraction (FP64: no contraction launch)
  double num_fsrs;      // N (keff and normalisation)
  double norm;          // residual normaliser (FSRs, or fissionable FSRs)
  double tol;           // convergence threshold of this solve
  double sigma_t_max;   // largest sigma_t of this solve's materials
};

namespace {

// a launch argument of set_batch_slot_kernel
static_assert(sizeof(BatchSlot) <= 2048, "batch row arguments exceed the kernel-argument budget");

template <int GPL, int EXPM, int NP, int STORE>
__global__ __launch_bounds__(kSweepBlockMax) void sweep2d_batch_kernel(const BatchSlot* __restrict__ t,
                                                                       int parity) {
  sweep2d_kernel_body<GPL, EXPM, NP, STORE>(t[blockIdx.y].sw[parity]);
}

// bps > 0: every row's subsets get bps blocks instead of the row's own count
// (a launch of many rows of one geometry needs fewer blocks per row to fill
// the device; the tally is an exact integer sum, so the bits are the same)
template <int EXPM, int QLDS, int NP>
__global__ __launch_bounds__(1024) void sweep2d_lds_batch_kernel(const BatchSlot* __restrict__ t,
                                                                 int parity, int bps) {
  const BatchSlot& b = t[blockIdx.y];
  LdsParams L = b.lds;
  if (bps > 0) L.blocks_per_sub = bps;
  // rows of smaller geometries than the largest in the batch use fewer blocks
  if ((int)blockIdx.x >= L.nsub * L.blocks_per_sub) return;
  sweep2d_lds_kernel_body<EXPM, QLDS, NP>(b.sw[parity], L);
}

// bps > 0: the LDS sweep of this launch used bps blocks per subset for every
// row (see sweep2d_lds_batch_kernel)
__global__ __launch_bounds__(kReduceBlock) void post_sweep_batch_kernel(const BatchSlot* __restrict__ t,
                                                                       int bps) {
  const BatchSlot& b = t[blockIdx.y];
  if ((int)blockIdx.x >= b.nparts_post) return;  // its partials end there
  post_sweep_kernel_body(b.fsr, 1, 0, bps > 0 ? bps : b.fsr.lds_bps);
}

// rows with few partials take the keff step in update_batch_kernel instead
__global__ __launch_bounds__(kReduceBlock) void keff_batch_kernel(const BatchSlot* __restrict__ t) {
  const BatchSlot& b = t[blockIdx.y];
  if (b.nparts_post <= kFuseKeffMaxParts) return;
  keff_kernel_body(b.fsr.partials, b.nparts_post, b.fsr.pstride, b.fsr.state, b.num_fsrs, 1, 0);
}

__global__ __launch_bounds__(256) void scatter_mfma_batch_kernel(const BatchSlot* __restrict__ t) {
  const BatchSlot& b = t[blockIdx.y];
  if (b.scat_path != MOC_SCATTER_MFMA_F16) return;
  scatter_mfma_kernel_body(b.scat);
}

__global__ __launch_bounds__(256) void scatter_fp32_batch_kernel(const BatchSlot* __restrict__ t) {
  const BatchSlot& b = t[blockIdx.y];
  if (b.scat_path != MOC_SCATTER_FP32) return;
  scatter_fp32_kernel_body(b.scat);
}

// FUSE: with keff_kernel's step for rows with few post-sweep partials, as in
// the row's single solve (launchIteration)
template <int FUSE>
__global__ __launch_bounds__(kReduceBlock) void update_batch_kernel(const BatchSlot* __restrict__ t,
                                                                   int res_type) {
  const BatchSlot& b = t[blockIdx.y];
  if ((int)blockIdx.x >= b.nparts_upd) return;
  update_kernel_body<FUSE>(b.fsr, 1, res_type, 1, 1, 1, kSrcTotal,
                           b.nparts_post <= kFuseKeffMaxParts ? b.nparts_post : 0, 0, b.num_fsrs);
}

__global__ __launch_bounds__(kReduceBlock) void fsr_residual_batch_kernel(const BatchSlot* __restrict__ t,
                                                                         int res_type) {
  const BatchSlot& b = t[blockIdx.y];
  if ((int)blockIdx.x >= b.nparts_fsr) return;  // none below kFsrResidualGroups groups
  fsr_residual_kernel_body(b.fsr, 1, res_type);
}

__global__ __launch_bounds__(kReduceBlock) void residual_batch_kernel(const BatchSlot* __restrict__ t) {
  const BatchSlot& b = t[blockIdx.y];
  residual_kernel_body(b.fsr.partials + (size_t)kPartResidual * b.fsr.pstride, b.nparts_upd, b.fsr.state,
                       b.norm, b.tol, 0, b.sigma_t_max, 1);
}

// Row `row` of the table := v (a launch argument, so the host's copy of v is
// free as soon as the launch is queued).
__global__ void set_batch_slot_kernel(BatchSlot* t, int row, BatchSlot v) {
  if (threadIdx.x == 0) t[row] = v;
}

// Iteration states of rows 0..n-1, gathered for one copy to the host.
__global__ void gather_states_kernel(const BatchSlot* __restrict__ t, int n, IterState* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = *t[j].fsr.state;
}

// Plain device copy (16-byte vector loads and stores), the reference kernel
// for calibrating HBM byte counters.
__global__ void copy_kernel(const int4* __restrict__ src, int4* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
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

// Raw FSR tally [N*G] from the per-segment store (single-sweep entry point).
__global__ void segment_tally_kernel(FsrArgs a, double* out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.N * a.G) return;
  const int64_t r = i / a.G;
  out[i] = segment_sum(a, r, (int)(i - r * a.G), 0);
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

// Dynamic LDS above the default needs the opt-in attribute; it is set once
// per kernel and device (before any graph capture that launches the kernel).
void ensure_lds_attribute(const void* kernel, int device, size_t bytes) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> done;
  std::lock_guard<std::mutex> lk(mu);
  size_t& cur = done[{kernel, device}];
  if (cur < bytes) {
    HIP_CHECK(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    cur = bytes;
  }
}

// ---------------------------------------------------------------------------
// Allocation-free single sweep: kernels over caller buffers and the static
// scratch array
// ---------------------------------------------------------------------------
// Fresh iteration state of a single sweep (psi_scale 1, nothing converged)
// and cleared bound words: bounds[0] max |Q/sigma_t| (double bits), bounds[1]
// max |psi| (float bits).
__global__ void noalloc_begin_kernel(IterState* st, unsigned long long* bounds, double scale_cap) {
  IterState z = {};
  z.k = 1.0;
  z.psi_scale = 1.0;
  z.scale_cap = scale_cap;
  *st = z;
  bounds[0] = 0ull;
  bounds[1] = 0ull;
}

// Maxima of |Q / sigma_t| (FP64, as the host computes them) and |psi| over
// the inputs: per-thread maxima, a wave reduction, one atomic per wave.
__global__ __launch_bounds__(256) void noalloc_bounds_kernel(const double* Q, const int32_t* fsr_mat,
                                                             const double* sigma_t, int64_t N, int G,
                                                             const float* psi, int64_t psi_n,
                                                             unsigned long long* bounds) {
  double qm = 0.0;
  float pm = 0.f;
  const int64_t NG = N * G;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < NG; i += stride) {
    const int64_t r = i / G;
    const int g = (int)(i - r * G);
    qm = fmax(qm, fabs(Q[i] / sigma_t[(size_t)fsr_mat[r] * G + g]));
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < psi_n; i += stride)
    pm = fmaxf(pm, fabsf(psi[i]));
  for (int off = 32; off > 0; off >>= 1) {
    qm = fmax(qm, __shfl_xor(qm, off, 64));
    pm = fmaxf(pm, __shfl_xor(pm, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    // non-negative floating-point values order like their bit patterns
    atomicMax(&bounds[0], (unsigned long long)__double_as_longlong(qm));
    atomicMax(&bounds[1], (unsigned long long)__float_as_uint(pm));
  }
}

// Fixed-point tally scale from the two maxima, with the host's arithmetic
// (transportSweepOnce).
__global__ void noalloc_scale_kernel(IterState* st, const unsigned long long* bounds,
                                     double sigma_t_max) {
  const double qmax = __longlong_as_double((long long)bounds[0]);
  const double pmax = (double)__uint_as_float((unsigned)bounds[1]);
  st->qmax = (float)(qmax * (1 + 1e-6));
  st->psi_max = (float)(pmax * (1 + 1e-6));
  const double bound = 2.0 * kFourPi * sigma_t_max * ((double)st->psi_max + (double)st->qmax);
  st->tally_scale = tally_scale_for(bound, st->scale_cap);
}

// Reference layout (dt*P*G + p*G + g) <-> start-flux slots [G][W][P] by work
// position (work[i] = directed track at position i, -1 padding).
__global__ void psi_to_slots_kernel(const float* psi, const int32_t* work, int64_t W, int P, int G,
                                    float* s0, float* s1) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)G * W * P) return;
  const int p = (int)(e % P);
  const int64_t i = (e / P) % W;
  const int g = (int)(e / ((int64_t)P * W));
  const int dt = work[i];
  if (dt < 0) return;
  const float v = psi[(size_t)dt * P * G + (size_t)p * G + g];
  s0[e] = v;
  s1[e] = v;
}

__global__ void slots_to_psi_kernel(const float* s, const int32_t* work, int64_t W, int P, int G,
                                    float* psi) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)G * W * P) return;
  const int p = (int)(e % P);
  const int64_t i = (e / P) % W;
  const int g = (int)(e / ((int64_t)P * W));
  const int dt = work[i];
  if (dt >= 0) psi[(size_t)dt * P * G + (size_t)p * G + g] = s[e];
}

// The raw FSR tally (sum of contributions, CPUSolver::_scalar_flux after the
// sweep) from the fixed-point tally (sum / V in units of 1/tally_scale) or
// from the FP32 tally copies summed in copy order.
__global__ void noalloc_tally_kernel(const unsigned long long* t64, const float* t32, int copies,
                                     const double* vol, const IterState* st, int64_t N, int G,
                                     double* out) {
  const int64_t NG = N * G;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NG) return;
  if (t64) {
    const double v = vol[i / G];
    out[i] = (double)(long long)t64[i] / st->tally_scale * (v < kFltEpsilon ? 0.0 : v);
  } else {
    double t = 0.0;
    for (int c = 0; c < copies; c++) t += t32[(size_t)c * NG + i];
    out[i] = t;
  }
}

}  // namespace

// Scratch memory of the allocation-free sweep (transportSweepNoAlloc): one
// copy per GPU, held until the process ends.  Host code never uses this
// name as an address; it uses the per-device pointer from
// hipGetSymbolAddress (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[MOC_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current), looked up
// once per device.
unsigned char* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) throw DeviceError(MOC_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw DeviceError(MOC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[device];
}

// Byte offsets of the allocation-free sweep's buffers within the scratch
// array (each 256-byte aligned).
struct NoAllocLayout {
  size_t state, bounds, tally, table, psi0, psi1, total;
};
inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

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
  dropGraphs();
  if (async_ev_) (void)hipEventDestroy(async_ev_);
  if (async_ev0_) (void)hipEventDestroy(async_ev0_);
  if (geo_hold_) {
    // geometry buffers belong to the shared holder
    d_seg_len_ = nullptr;
    d_seg_key_ = nullptr;
    d_work_ = nullptr;
    d_tile_ = nullptr;
    d_trk_ = nullptr;
    d_trk_azim_ = nullptr;
    d_weights_ = nullptr;
    d_fsr_mat_ = nullptr;
    d_fsr_vol_ = nullptr;
    d_inv_vol_ = nullptr;
    d_out_wp_ = nullptr;
    d_csr_off_ = nullptr;
    d_seg_slot_ = nullptr;
  }
  dev_free(d_seg_len_);
  dev_free(d_csr_off_);
  dev_free(d_seg_slot_);
  dev_free(d_segout_);
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
  dev_free(d_phi_old_[0]);
  dev_free(d_phi_old_[1]);
  dev_free(d_sigma_a_);
  dev_free(d_leak_);
  dev_free(d_seg_pos_);
  dev_free(d_azim_cos_);
  dev_free(d_azim_sin_);
  dev_free(d_ilem_);
  dev_free(d_src_c_);
  dev_free(d_phi_xy_);
  dev_free(d_phi_xy_n_);
  dev_free(d_Qxy_);
  dev_free(d_qls_);
  dev_free(d_tally_xy_);
  dev_free(d_scat_);
  dev_free(d_grp_);
  dev_free(d_slot_fsr_);
  dev_free(d_tally_i64_);
  dev_free(d_stack_partial_);
  dev_free(d_stack_off_);
  dev_free(d_stack_slots_);
  dev_free(d_scat_rows_);
  dev_free(d_scat_tile_mat_);
  dev_free(d_scat_b_);
  dev_free(d_scat_colmul_);
  dev_free(d_ss_colsum_);
  dev_free(d_stab_ratio_);
  dev_free(d_ref_flux_);
  dev_free(d_Q_);
  dev_free(d_fixed_);
  dev_free(d_sigma_t_);
  dev_free(d_sigma_s_);
  dev_free(d_nsf_);
  dev_free(d_chi_);
  dev_free(d_fis_nsf_);
  dev_free(d_mat_fissionable_);
  dev_free(d_fsr_mat_);
  dev_free(d_fsr_vol_);
  dev_free(d_inv_vol_);
  dev_free(d_out_wp_);
  dev_free(d_partials_);
  dev_free(d_lds_partials_);
  dev_free(d_qsub_);
  dev_free(d_tally64_);
  dev_free(d_state_);
  if (h_state_) (void)hipHostFree(h_state_);
  if (h_xs_stage_) (void)hipHostFree(h_xs_stage_);
  if (stream_ && !stream_borrowed_) (void)hipStreamDestroy(stream_);
}

// setMaterials for a solver that already has materials of these counts:
// the same tables (production matrices, absorption, column sums, fissionable
// flags), built in a pinned staging buffer and copied asynchronously on the
// solver's stream; the matrix-core operands are rebuilt there too.  The
// iteration arrays are kept (launchInit starts the next solve afresh).
void HipSolver::rearmMaterials(const double* sigma_t, const double* sigma_s, const double* nu_sigma_f,
                               const double* chi) {
  if (!materials_set_) throw DeviceError(MOC_ERR_STATE, "materials have not been set");
  if (!sigma_t || !sigma_s || !nu_sigma_f || !chi) throw DeviceError(MOC_ERR_INVALID, "invalid material arrays");
  const int M = M_, G = G_;
  const size_t MG = (size_t)M * G;
  std::vector<int32_t> fis(M, 0);
  double st_max = 0.0;
  for (int m = 0; m < M; m++)
    for (int g = 0; g < G; g++) {
      if (!(sigma_t[(size_t)m * G + g] > 0.0)) throw DeviceError(MOC_ERR_INVALID, "sigma_t must be positive");
      if (nu_sigma_f[(size_t)m * G + g] != 0.0) fis[m] = 1;
      st_max = std::max(st_max, sigma_t[(size_t)m * G + g]);
    }
  h_sigma_t_.assign(sigma_t, sigma_t + MG);
  h_sigma_s_fwd_.assign(sigma_s, sigma_s + MG * G);
  h_nsf_fwd_.assign(nu_sigma_f, nu_sigma_f + MG);
  h_chi_fwd_.assign(chi, chi + MG);
  const bool adj = solver_mode_ == MOC_ADJOINT;
  // staging: sigma_t | nu-sigma-f | sigma_s (dest-major) | chi | fission nsf |
  // sigma_a | column sums | fissionable flags (as doubles' worth of int32)
  const size_t need = MG * 7 + MG * G + ((size_t)M + 1) / 2;
  HIP_CHECK(hipSetDevice(device_));
  if (h_xs_stage_n_ < need) {
    if (h_xs_stage_) HIP_CHECK(hipHostFree(h_xs_stage_));
    h_xs_stage_ = nullptr;
    HIP_CHECK(hipHostMalloc((void**)&h_xs_stage_, need * sizeof(double), hipHostMallocDefault));
    h_xs_stage_n_ = need;
  }
  double* st = h_xs_stage_;
  double* nsf = st + MG;
  double* ss = nsf + MG;
  double* ch = ss + MG * G;
  double* fn = ch + MG;
  double* sa = fn + MG;
  double* cs = sa + MG;
  int32_t* fl = reinterpret_cast<int32_t*>(cs + MG);
  std::memcpy(st, sigma_t, MG * sizeof(double));
  std::memcpy(nsf, nu_sigma_f, MG * sizeof(double));
  for (int m = 0; m < M; m++)
    for (int d = 0; d < G; d++)
      for (int g = 0; g < G; g++)
        ss[((size_t)m * G + d) * G + g] =
            adj ? sigma_s[((size_t)m * G + g) * G + d] : sigma_s[((size_t)m * G + d) * G + g];
  std::memcpy(ch, adj ? nu_sigma_f : chi, MG * sizeof(double));
  std::memcpy(fn, adj ? chi : nu_sigma_f, MG * sizeof(double));
  for (int m = 0; m < M; m++)
    for (int g = 0; g < G; g++) {
      double v = sigma_t[(size_t)m * G + g], c = 0.0;
      for (int d = 0; d < G; d++) {
        v -= ss[((size_t)m * G + d) * G + g];
        c += ss[((size_t)m * G + d) * G + g];
      }
      sa[(size_t)m * G + g] = v;
      cs[(size_t)m * G + g] = c;
    }
  std::memcpy(fl, fis.data(), M * sizeof(int32_t));
  h_sigma_s_.assign(ss, ss + MG * G);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
  };
  up(d_sigma_t_, st, MG * sizeof(double));
  up(d_nsf_, nsf, MG * sizeof(double));
  up(d_sigma_s_, ss, MG * G * sizeof(double));
  up(d_chi_, ch, MG * sizeof(double));
  up(d_fis_nsf_, fn, MG * sizeof(double));
  up(d_sigma_a_, sa, MG * sizeof(double));
  up(d_ss_colsum_, cs, MG * sizeof(double));
  up(d_mat_fissionable_, fl, M * sizeof(int32_t));
  mat_fissionable_ = fis;
  sigma_t_max_ = st_max;
  countFissionable();
  scat_frag_ok_ = false;
  prepareScatter();  // matrix-core operands of the new matrices, on stream_
  graph_dirty_ = true;
  state_valid_ = false;
  solved_ = false;
}

// Runs this solver's launches on `s` (owned by the caller) from now on.
void HipSolver::adoptStream(hipStream_t s) {
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  if (!stream_borrowed_) HIP_CHECK(hipStreamDestroy(stream_));
  stream_ = s;
  stream_borrowed_ = true;
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
  // Up to 4 groups per lane, the count that keeps the most lanes busy: at 70
  // groups 2 per lane fill 35 lanes of a wave (one track), 4 per lane 54 of
  // them (three tracks of 18 lanes); measured on the 70-group mini-core the
  // flat sweep takes 1.52 ms instead of 1.72 ms, the linear-source one 4.38
  // instead of 4.47.  More groups per lane hold more registers and measured
  // slower (9: 1.82 ms).
  {
    auto busy = [&](int gpl) {
      const int lpt = (G_ + gpl - 1) / gpl;
      return lpt > 64 ? 0.0 : (double)G_ / (lpt * gpl) * (64 / lpt) * lpt / 64.0;
    };
    for (int gpl : {2, 4})
      if (gpl > s.groups_per_lane && busy(gpl) > busy(s.groups_per_lane) + 0.05) s.groups_per_lane = gpl;
  }
  s.lanes_per_track = (G_ + s.groups_per_lane - 1) / s.groups_per_lane;
  s.tracks_per_wave = 64 / s.lanes_per_track;
  s.block = 256;
  // the global-tally path stores per segment by default (deterministic)
  s.seg_store = 1;
  // LDS-privatised tally when at least one group's slice fits in LDS: the
  // largest subset that fits, evened out over the subsets.
  // LDS-privatised tally when a group's slice fits in LDS: with the source
  // slice resident too (lds_q = 1) if that fits, else gathered.  The largest
  // subset that fits, evened out over the subsets.
  s.lds_groups = 0;
  s.lds_q = 0;
  if (N_ > 0 && lds_max_ > 0 && !linear_) {
    for (int q = 1; q >= 0 && s.lds_groups == 0; q--) {
      int gb = 0;
      while (gb < std::min(G_, 64) && ldsBytes(gb + 1, q) <= lds_max_) gb++;
      if (gb < 1) continue;
      const int nsub = (G_ + gb - 1) / gb;
      gb = (G_ + nsub - 1) / nsub;
      s.lds_groups = gb;
      s.lds_q = q;
      s.lds_block = 1024;
      // Blocks run in rounds of (CUs x blocks resident per CU).  About 4/3
      // of a round, rounded down to a multiple of the XCD count (which
      // places the subsets of one track slot on one XCD, where they share its
      // L2 for the segment stream): on the 17x17 lattice (one block per CU)
      // 7 x 48 to 7 x 80 blocks all take 146-150 us per source iteration,
      // 7 x 40 153-163 us and one round (7 x 32 or 36) 163-168 us; the
      // fewest blocks that keep the speed write the fewest partial slices.
      // 7 x 73 without the XCD placement took 145 us but fetched 357 MB per
      // sweep against 47 MB.
      const int64_t resident =
          std::max<int64_t>(1, std::min<int64_t>((int64_t)(lds_max_ / ldsBytes(gb, q)), 2048 / s.lds_block));
      s.lds_blocks_per_sub =
          (int)std::max<int64_t>(kNumXcd, 4 * num_cus_ * resident / (3 * nsub) / kNumXcd * kNumXcd);
      // no more blocks than the subset has waves of tiles: a small problem
      // (a pin cell) then launches a few blocks instead of hundreds of idle
      // ones, each of which would clear and write a partial tally slice
      if (T_ > 0) {
        const int64_t tiles = (2 * T_ + 64 / gb - 1) / (64 / gb);
        const int64_t waves = std::max<int64_t>(1, (tiles + s.lds_block / 64 - 1) / (s.lds_block / 64));
        s.lds_blocks_per_sub = (int)std::min<int64_t>(s.lds_blocks_per_sub, waves);
        s.lds_blocks_per_sub = (s.lds_blocks_per_sub + kNumXcd - 1) / kNumXcd * kNumXcd;
      }
    }
    // z-stack sweep: one group per block.  Narrow slices keep several blocks
    // on a CU: on the 3D 17x17 lattice 1 group per block sweeps in 1.10 ms,
    // 2 or 4 (one block per CU) in 1.43-1.45 ms.
    if (is3d_ && ngrp_ > 0 && s.lds_groups > 0) {
      s.lds_groups = 1;
      s.lds_q = 1;
    }
  }
  shape_ = s;
}

// LDS of the LDS sweep: int64 tally [N*gb] | 1/V [N] | sigma_t [M*gb] | Q/sigma_t [N*gb]
size_t HipSolver::ldsBytes(int gb, int with_q) const {
  // z-stack sweep (sources always in LDS): tally [slots*gb] i64 |
  // {Q/sigma_t, iv} [slots*gb] | sigma_t [M*gb]
  if (is3d_ && ngrp_ > 0)
    return (size_t)max_grp_slots_ * gb * (sizeof(uint64_t) + 2 * sizeof(float)) +
           (size_t)std::max(M_, 1) * gb * sizeof(float);
  return (size_t)N_ * gb * sizeof(uint64_t) +
         (with_q ? (size_t)N_ * gb * 2 * sizeof(float) : (size_t)N_ * sizeof(float)) +
         (size_t)std::max(M_, 1) * gb * sizeof(float);
}

void HipSolver::setShape(const LaunchShape& s) {
  graph_dirty_ = true;
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
      s.lds_flush < 0 || s.lds_flush > 1 || s.seg_store < 0 || s.seg_store > 1 ||
      (s.lds_groups > 0 &&
       (ldsBytes(s.lds_groups, s.lds_q) > lds_max_ || s.lds_block <= 0 ||
        s.lds_block % 64 != 0 || s.lds_block > 1024 || s.lds_blocks_per_sub <= 0)))
    throw DeviceError(MOC_ERR_INVALID, "invalid LDS sweep shape");
  const bool realloc = s.tally_copies != shape_.tally_copies || s.lds_groups != shape_.lds_groups ||
                       s.lds_blocks_per_sub != shape_.lds_blocks_per_sub;
  (void)realloc;
  shape_ = s;
  shape_forced_ = true;
  // streams, start-flux slots and tallies follow the shape (buffers that are
  // big enough are kept)
  allocateIterationArrays();
}

void HipSolver::setTracks(const moc_track_data_t& td) {
  HIP_CHECK(hipSetDevice(device_));
  if (geo_hold_) throw DeviceError(MOC_ERR_STATE, "this solver shares its geometry; use a new solver");
  if (td.num_tracks <= 0 || td.num_segments <= 0 || !td.seg_offset || !td.seg_length ||
      !td.seg_fsr || !td.azim || !td.next_fwd || !td.next_bwd || !td.next_fwd_is_fwd ||
      !td.next_bwd_is_fwd || !td.bc_fwd || !td.bc_bwd || !td.total_weight || !td.sin_theta)
    throw DeviceError(MOC_ERR_INVALID, "incomplete track data");
  if (td.num_azim < 4 || td.num_azim % 4 || td.num_polar < 2 || td.num_polar % 2)
    throw DeviceError(MOC_ERR_INVALID, "invalid quadrature sizes");
  const bool is3d = td.polar != nullptr;
  if (!is3d && td.num_polar / 2 > kMaxPolarHalf)
    throw DeviceError(MOC_ERR_INVALID, "2D sweep supports at most 6 polar angles");
  if (td.num_segments >= (int64_t)1 << 31 || td.num_tracks >= (int64_t)1 << 29)
    throw DeviceError(MOC_ERR_INVALID, "track set too large for 32-bit indexing");
  const int64_t T = td.num_tracks, S = td.num_segments;
  if (td.seg_offset[0] != 0 || td.seg_offset[T] != S)
    throw DeviceError(MOC_ERR_INVALID, "seg_offset must start at 0 and end at num_segments");
  const int na2 = td.num_azim / 2;
  // 3D: one polar angle per track; the weight class azim * num_polar + polar
  // takes the place of the azimuthal index
  const int P = is3d ? 1 : td.num_polar / 2;

  std::vector<int32_t> trk(4 * T), azim(T);
  std::vector<int> in_count(2 * T, 0);
  for (int64_t t = 0; t < T; t++) {
    const int64_t b = td.seg_offset[t], e = td.seg_offset[t + 1];
    if (e < b) throw DeviceError(MOC_ERR_INVALID, "seg_offset must be non-decreasing");
    if (td.azim[t] < 0 || td.azim[t] >= na2) throw DeviceError(MOC_ERR_INVALID, "azimuthal index out of range");
    azim[t] = td.azim[t];
    if (is3d) {
      if (td.polar[t] < 0 || td.polar[t] >= td.num_polar)
        throw DeviceError(MOC_ERR_INVALID, "polar index out of range");
      azim[t] = td.azim[t] * td.num_polar + td.polar[t];
    }
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
  std::vector<float> w;
  if (is3d) {
    w.resize((size_t)na2 * td.num_polar);
    for (size_t c = 0; c < w.size(); c++) w[c] = (float)td.total_weight[c];
    inv_sin_[0] = 1.f;  // lengths are 3D lengths
    for (int p = 1; p < kMaxPolarHalf; p++) inv_sin_[p] = 0.f;
  } else {
    w.resize((size_t)na2 * P);
    for (int a = 0; a < na2; a++)
      for (int p = 0; p < P; p++) w[(size_t)a * P + p] = (float)td.total_weight[(size_t)a * td.num_polar + p];
    for (int p = 0; p < kMaxPolarHalf; p++)
      inv_sin_[p] = p < P ? (float)(1.0 / td.sin_theta[p]) : 0.f;
  }

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
  // linear-source inputs (optional)
  h_seg_start_.clear();
  h_azim_phi_.clear();
  h_azim_w_.clear();
  h_polar_w_.clear();
  if (!is3d && td.seg_start && td.azim_phi && td.azim_weight && td.polar_weight) {
    h_seg_start_.assign(td.seg_start, td.seg_start + 2 * S);
    h_azim_phi_.assign(td.azim_phi, td.azim_phi + na2);
    h_azim_w_.assign(td.azim_weight, td.azim_weight + na2);
    h_polar_w_.assign(td.polar_weight, td.polar_weight + P);
  }
  ls_ready_ = false;

  // 3D: group the tracks by the 2D track they lie over (azimuthal index and
  // xy index / kStackPairing: neighbouring parallel 2D tracks, which cross
  // nearly the same FSRs, share a group, so a group's tally flush covers the
  // work of both); each group's FSRs get slots 0..n-1 in order of first
  // appearance.  On the 3D 17x17 lattice pairs sweep 12% faster than single
  // 2D tracks; groups of 4 and 8 gain less or lose (larger LDS slices).
  constexpr int32_t kStackPairing = 2;
  ngrp_ = 0;
  max_grp_slots_ = 0;
  h_trk_grp_.clear();
  h_slot_.clear();
  h_grp_slot0_.clear();
  h_grp_nslots_.clear();
  h_slot_fsr_.clear();
  if (is3d && td.xy) {
    std::map<std::pair<int32_t, int32_t>, int32_t> gid;
    h_trk_grp_.resize(T);
    for (int64_t t = 0; t < T; t++) {
      const std::pair<int32_t, int32_t> key(td.azim[t], td.xy[t] / kStackPairing);
      auto it = gid.find(key);
      if (it == gid.end()) it = gid.emplace(key, (int32_t)gid.size()).first;
      h_trk_grp_[t] = it->second;
    }
    ngrp_ = (int32_t)gid.size();
    // groups numbered by work, most segments first: blocks are dispatched in
    // group order, so the largest groups start first and the last blocks of
    // the launch are short ones
    {
      std::vector<int64_t> work(ngrp_, 0);
      for (int64_t t = 0; t < T; t++) work[h_trk_grp_[t]] += td.seg_offset[t + 1] - td.seg_offset[t];
      std::vector<int32_t> by_work(ngrp_);
      std::iota(by_work.begin(), by_work.end(), 0);
      std::stable_sort(by_work.begin(), by_work.end(),
                       [&](int32_t x, int32_t y) { return work[x] > work[y]; });
      std::vector<int32_t> rank(ngrp_);
      for (int32_t i = 0; i < ngrp_; i++) rank[by_work[i]] = i;
      for (int64_t t = 0; t < T; t++) h_trk_grp_[t] = rank[h_trk_grp_[t]];
    }
    std::vector<std::vector<int64_t>> members(ngrp_);
    for (int64_t t = 0; t < T; t++) members[h_trk_grp_[t]].push_back(t);
    h_slot_.assign(S, 0);
    int32_t max_f = 0;
    for (int64_t k = 0; k < S; k++) max_f = std::max(max_f, td.seg_fsr[k]);
    std::vector<int32_t> slot_of((size_t)max_f + 1, -1);  // reset per group
    for (int32_t g = 0; g < ngrp_; g++) {
      const int32_t s0 = (int32_t)h_slot_fsr_.size();
      h_grp_slot0_.push_back(s0);
      for (int64_t t : members[g])
        for (int64_t k = td.seg_offset[t]; k < td.seg_offset[t + 1]; k++) {
          int32_t& sl = slot_of[td.seg_fsr[k]];
          if (sl < 0) {
            sl = (int32_t)h_slot_fsr_.size() - s0;
            h_slot_fsr_.push_back(td.seg_fsr[k]);
          }
          h_slot_[k] = sl;
        }
      const int32_t n = (int32_t)h_slot_fsr_.size() - s0;
      {
        // slots in FSR order within the group: lanes on neighbouring tracks
        // then meet on fewer LDS banks than with first-encounter numbering
        std::vector<int32_t> perm(n), newpos(n), sorted(n);
        std::iota(perm.begin(), perm.end(), 0);
        std::sort(perm.begin(), perm.end(),
                  [&](int32_t x, int32_t y) { return h_slot_fsr_[s0 + x] < h_slot_fsr_[s0 + y]; });
        for (int32_t i = 0; i < n; i++) {
          newpos[perm[i]] = i;
          sorted[i] = h_slot_fsr_[s0 + perm[i]];
        }
        std::copy(sorted.begin(), sorted.end(), h_slot_fsr_.begin() + s0);
        for (int64_t t : members[g])
          for (int64_t k = td.seg_offset[t]; k < td.seg_offset[t + 1]; k++) h_slot_[k] = newpos[h_slot_[k]];
      }
      for (int32_t i = 0; i < n; i++) slot_of[h_slot_fsr_[s0 + i]] = -1;
      h_grp_nslots_.push_back(n);
      max_grp_slots_ = std::max(max_grp_slots_, n);
    }
    dev_alloc(&d_slot_fsr_, h_slot_fsr_.size());
    HIP_CHECK(hipMemcpy(d_slot_fsr_, h_slot_fsr_.data(), h_slot_fsr_.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
  }

  T_ = T;
  S_ = S;
  P_ = P;
  is3d_ = is3d;
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

// Tile-interleaved segment streams for tiles of `tpw` tracks taken in work
// order (longest first): row i of a tile holds the i-th segment of each of
// its tracks; shorter tracks are padded with length-0 segments that repeat
// their last FSR.  Flat-source sweeps read one stream for both directions of
// a track: each stream tile has a forward and a backward work tile (the
// backward pass reads the rows last to first), which halves the stream
// traffic.  The linear-source sweep keeps one stream per directed track (its
// segment entry points differ by direction).
void HipSolver::buildTileLayout(int tpw) {
  const bool stack = stackMode();
  const bool shared = !linear_;
  if (tpw == layout_tpw_ && stack == layout_stack_ && shared == layout_shared_) return;
  if (geo_hold_)
    throw DeviceError(MOC_ERR_INVALID,
                      "shared geometry is laid out for another launch shape (same group count needed)");
  if (!(tracks_set_ && fsrs_set_)) return;
  if (N_ >= ((int64_t)1 << kFsrBits) || M_ > 256)
    throw DeviceError(MOC_ERR_INVALID, "at most 2^24 FSRs and 256 materials");
  // stream units in work order: directed tracks longest first, or with a
  // shared stream each track once (as its forward direction); for the
  // z-stack sweep, group by group (each group's units longest first, its
  // last tile padded)
  const int reps = shared ? 2 : 1;  // work tiles per stream tile
  std::vector<int32_t> units;
  units.reserve(h_order_.size() / reps);
  for (int32_t dt : h_order_)
    if (!shared || (dt & 1) == 0) units.push_back(dt);
  std::vector<int32_t> order_grp;
  std::vector<int32_t> grp_tile0, grp_ntiles;  // in work tiles
  const std::vector<int32_t>* order = &units;
  if (stack) {
    std::vector<std::vector<int32_t>> mem(ngrp_);
    for (int32_t u : units) mem[h_trk_grp_[u >> 1]].push_back(u);
    for (int32_t g = 0; g < ngrp_; g++) {
      const int32_t t0 = (int32_t)(order_grp.size() / tpw);
      order_grp.insert(order_grp.end(), mem[g].begin(), mem[g].end());
      while (order_grp.size() % tpw) order_grp.push_back(-1);
      grp_tile0.push_back(t0 * reps);
      grp_ntiles.push_back(((int32_t)(order_grp.size() / tpw) - t0) * reps);
    }
    order = &order_grp;
  }
  const int64_t U = (int64_t)order->size();
  const int64_t nst = (U + tpw - 1) / tpw;  // stream tiles
  const int64_t ntiles = nst * reps;        // work tiles
  if (ntiles >= (int64_t)1 << 30) throw DeviceError(MOC_ERR_INVALID, "too many tiles");
  std::vector<int32_t> work(ntiles * tpw, -1);
  std::vector<int32_t> tile(2 * ntiles);
  std::vector<int64_t> st_row0(nst);
  int64_t rows = kTailRows;  // head rows: a backward pass's look-ahead before the first tile
  for (int64_t ts = 0; ts < nst; ts++) {
    int32_t nmax = 0;
    for (int sl = 0; sl < tpw; sl++) {
      const int64_t w = ts * tpw + sl;
      if (w >= U) break;
      const int32_t u = (*order)[w];
      if (u < 0) continue;
      for (int d = 0; d < reps; d++) work[(ts * reps + d) * tpw + sl] = u + d;
      nmax = std::max(nmax, h_trk_[4 * (u >> 1) + 1]);
    }
    nmax = (nmax + kRowPad - 1) / kRowPad * kRowPad;
    st_row0[ts] = rows;
    for (int d = 0; d < reps; d++) {
      tile[2 * (ts * reps + d)] = (int32_t)rows;
      tile[2 * (ts * reps + d) + 1] = nmax | (d ? kTileBackward : 0);
    }
    rows += nmax;
    if (rows >= (int64_t)1 << 30) throw DeviceError(MOC_ERR_INVALID, "segment streams too large");
  }
  // zero-length head and tail rows keyed to FSR 0 (read by the look-ahead,
  // never integrated)
  const size_t total = (size_t)(rows + kTailRows) * tpw;
  if (total * sizeof(float) >= ((size_t)1 << 31))
    throw DeviceError(MOC_ERR_INVALID, "segment streams exceed 2 GiB");
  std::vector<float> ilen(total, 0.f);
  std::vector<int32_t> ifsr(total, 0);
  auto key = [&](int32_t f) {
    return (int32_t)((uint32_t)f | ((uint32_t)h_fsr_mat_[f] << kFsrBits));
  };
  // key of segment k: its FSR, or (z-stack sweep) its slot in the group
  auto seg_key = [&](int64_t k) {
    if (!stack) return key(h_fsr_[k]);
    return (int32_t)((uint32_t)h_slot_[k] | ((uint32_t)h_fsr_mat_[h_fsr_[k]] << kFsrBits));
  };
  for (int64_t ts = 0; ts < nst; ts++) {
    const int64_t r0 = st_row0[ts];
    const int32_t nr = tile[2 * ts * reps + 1] & kTileRowsMask;
    for (int sl = 0; sl < tpw; sl++) {
      const int32_t u = work[ts * reps * tpw + sl];
      if (u < 0) continue;
      const int32_t t = u >> 1, dir = u & 1;  // dir 0 for a shared stream
      const int32_t b = h_trk_[4 * t], n = h_trk_[4 * t + 1];
      // padding repeats the key of the stream's last segment (the last
      // segment in traversal order, which is segment b for a backward track)
      const int32_t lastk = n > 0 ? (dir ? b : b + n - 1) : -1;
      const int32_t lastkey = lastk >= 0 ? seg_key(lastk) : 0;
      for (int32_t i = 0; i < nr; i++) {
        const size_t o = (size_t)(r0 + i) * tpw + sl;
        if (i < n) {
          const int32_t sidx = dir ? b + n - 1 - i : b + i;
          ilen[o] = h_len_[sidx];
          ifsr[o] = seg_key(sidx);
        } else {
          ilen[o] = 0.f;
          ifsr[o] = lastkey;
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
  layout_stack_ = stack;
  layout_shared_ = shared;
  layout_ntiles_ = (int)ntiles;
  layout_rows_ = rows;
  layout_grp_ntiles_.clear();
  if (stack) {
    std::vector<int32_t> g4((size_t)ngrp_ * 4);
    for (int32_t g = 0; g < ngrp_; g++) {
      g4[4 * g] = grp_tile0[g];
      g4[4 * g + 1] = grp_ntiles[g];
      g4[4 * g + 2] = h_grp_slot0_[g];
      g4[4 * g + 3] = h_grp_nslots_[g];
    }
    dev_alloc(&d_grp_, g4.size());
    HIP_CHECK(hipMemcpy(d_grp_, g4.data(), g4.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    layout_grp_ntiles_ = grp_ntiles;
  }
  // start-flux slots live at work positions: position of each directed
  // track, and the position its outgoing flux is written to
  h_wpos_.assign(2 * T_, -1);
  for (size_t i = 0; i < work.size(); i++)
    if (work[i] >= 0) h_wpos_[work[i]] = (int32_t)i;
  std::vector<int32_t> out_wp(2 * T_, -1);
  for (int64_t dt = 0; dt < 2 * T_; dt++) {
    const int32_t link = h_trk_[4 * (dt >> 1) + 2 + (dt & 1)];
    out_wp[dt] = link >= 0 ? h_wpos_[link] : -1;
  }
  dev_alloc(&d_out_wp_, out_wp.size());
  HIP_CHECK(hipMemcpy(d_out_wp_, out_wp.data(), out_wp.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  psi_w_ = (int64_t)ntiles * tpw;
  h_work_.swap(work);
  h_tile_.swap(tile);
  ls_ready_ = false;
  csr_ready_ = false;  // the per-segment index follows the stream layout
  state_valid_ = false;  // start fluxes are laid out by work position
}

// Rows of the per-segment store: FSR by FSR (a counting sort of the
// segments by FSR, each FSR's segments in stream order), then rows at the
// end for the padding segments; csr_off[r] is FSR r's first row and the
// seg_slot stream gives each stream position its row.  With a stream shared
// by both directions each position has two adjacent rows, the forward pass's
// and, after it, the backward pass's.  Geometry data, shared with solvers
// that share the streams.
void HipSolver::buildSegmentIndex() {
  // rows packed back to back: the per-FSR sums read an FSR's rows, and the
  // FSRs, in row order, so rows that share a cache line are read together
  seg_stride_ = G_;
  if (csr_ready_) return;
  if (geo_hold_)
    throw DeviceError(MOC_ERR_STATE, "the shared geometry has no per-segment index for this launch shape");
  const int tpw = layout_tpw_;
  const int per = layout_shared_ ? 2 : 1;  // rows per stream position
  auto each = [&](auto fn) {
    for (int64_t tl = 0; tl < layout_ntiles_; tl++) {
      // a shared stream's positions are visited once, by the forward pass
      if (h_tile_[2 * tl + 1] & kTileBackward) continue;
      const int64_t r0 = h_tile_[2 * tl];
      const int32_t nr = h_tile_[2 * tl + 1] & kTileRowsMask;
      for (int32_t i = 0; i < nr; i++)
        for (int sl = 0; sl < tpw; sl++) {
          const int32_t dt = h_work_[tl * tpw + sl];
          if (dt < 0) continue;
          const int32_t t = dt >> 1, dir = dt & 1;
          const int32_t b = h_trk_[4 * t], n = h_trk_[4 * t + 1];
          if (i >= n) continue;
          fn(h_fsr_[dir ? b + n - 1 - i : b + i], (size_t)(r0 + i) * tpw + sl);
        }
    }
  };
  std::vector<int64_t> off((size_t)N_ + 1, 0);
  each([&](int32_t f, size_t) { off[(size_t)f + 1] += per; });
  for (int64_t r = 0; r < N_; r++) off[r + 1] += off[r];
  if (off[N_] + per >= ((int64_t)1 << 31)) throw DeviceError(MOC_ERR_INVALID, "too many segments");
  seg_rows_ = off[N_] + per;
  const size_t total = (size_t)(layout_rows_ + kTailRows) * tpw;
  std::vector<int32_t> slot(total, (int32_t)off[N_]);  // padding: the last rows
  std::vector<int64_t> cur(off.begin(), off.end() - 1);
  each([&](int32_t f, size_t q) {
    slot[q] = (int32_t)cur[f];
    cur[f] += per;
  });
  dev_alloc(&d_csr_off_, off.size());
  dev_alloc(&d_seg_slot_, total);
  HIP_CHECK(hipMemcpy(d_csr_off_, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_seg_slot_, slot.data(), total * sizeof(int32_t), hipMemcpyHostToDevice));
  csr_ready_ = true;
}

int HipSolver::sweepTracksPerWave() const {
  return shape_.lds_groups > 0 ? 64 / shape_.lds_groups : shape_.tracks_per_wave;
}

void HipSolver::setMaterials(int num_materials, int num_groups, const double* sigma_t,
                             const double* sigma_s, const double* nu_sigma_f, const double* chi) {
  graph_dirty_ = true;
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
  dev_alloc(&d_nsf_, (size_t)M * G);
  dev_alloc(&d_mat_fissionable_, M);
  HIP_CHECK(hipMemcpy(d_sigma_t_, sigma_t, (size_t)M * G * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_nsf_, nu_sigma_f, (size_t)M * G * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_mat_fissionable_, fis.data(), M * sizeof(int32_t), hipMemcpyHostToDevice));
  mat_fissionable_ = fis;
  h_sigma_t_.assign(sigma_t, sigma_t + (size_t)M * G);
  h_sigma_s_fwd_.assign(sigma_s, sigma_s + (size_t)M * G * G);
  h_nsf_fwd_.assign(nu_sigma_f, nu_sigma_f + (size_t)M * G);
  h_chi_fwd_.assign(chi, chi + (size_t)M * G);
  uploadProductionMatrices(M, G);
  sigma_t_max_ = 0.0;
  for (size_t i = 0; i < (size_t)M * G; i++) sigma_t_max_ = std::max(sigma_t_max_, sigma_t[i]);
  if (G != G_) {
    G_ = G;
    shape_forced_ = false;
    fixed_on_ = false;
    h_fixed_.clear();
    fix_fsr_map_.clear();
    fix_mat_map_.clear();
    fix_cell_map_.clear();
    fixed_dirty_ = false;
  }
  M_ = M;
  materials_set_ = true;
  ls_ready_ = false;
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

// The production matrices the sources use (Solver::initializeMaterials,
// src/Solver.cpp:792-812): forward, sigma_s and the fission matrix
// chi (x) nu-sigma-f; adjoint, both transposed (Material::
// transposeProductionMatrices, src/Material.cpp:1091-1103), i.e. sigma_s^T and
// nu-sigma-f (x) chi.  keff, the normalisation and the fission-source residual
// keep nu-sigma-f (src/CPUSolver.cpp:1860-1931, 2113-2328); the absorption and
// the total-source residual follow the matrix in use, as the reference's read
// its transposed sigma_s.
void HipSolver::uploadProductionMatrices(int M, int G) {
  const bool adj = solver_mode_ == MOC_ADJOINT;
  const size_t MG = (size_t)M * G;
  h_sigma_s_.resize(MG * G);
  for (int m = 0; m < M; m++)
    for (int d = 0; d < G; d++)
      for (int g = 0; g < G; g++)
        h_sigma_s_[((size_t)m * G + d) * G + g] =
            adj ? h_sigma_s_fwd_[((size_t)m * G + g) * G + d] : h_sigma_s_fwd_[((size_t)m * G + d) * G + g];
  const std::vector<double>& src_chi = adj ? h_nsf_fwd_ : h_chi_fwd_;
  const std::vector<double>& src_nsf = adj ? h_chi_fwd_ : h_nsf_fwd_;
  dev_alloc(&d_sigma_s_, MG * G);
  dev_alloc(&d_chi_, MG);
  dev_alloc(&d_fis_nsf_, MG);
  HIP_CHECK(hipMemcpy(d_sigma_s_, h_sigma_s_.data(), MG * G * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_chi_, src_chi.data(), MG * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_fis_nsf_, src_nsf.data(), MG * sizeof(double), hipMemcpyHostToDevice));
  // Material::getSigmaA: sigma_t minus the out-scattering of the group
  // (src/Material.cpp:228-257); column sums for the total-source residual
  std::vector<double> sa(MG), cs(MG, 0.0);
  for (int m = 0; m < M; m++)
    for (int g = 0; g < G; g++) {
      double v = h_sigma_t_[(size_t)m * G + g];
      for (int d = 0; d < G; d++) {
        v -= h_sigma_s_[((size_t)m * G + d) * G + g];
        cs[(size_t)m * G + g] += h_sigma_s_[((size_t)m * G + d) * G + g];
      }
      sa[(size_t)m * G + g] = v;
    }
  dev_alloc(&d_sigma_a_, MG);
  HIP_CHECK(hipMemcpy(d_sigma_a_, sa.data(), sa.size() * sizeof(double), hipMemcpyHostToDevice));
  dev_alloc(&d_ss_colsum_, MG);
  HIP_CHECK(hipMemcpy(d_ss_colsum_, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice));
  scat_frag_ok_ = false;
  graph_dirty_ = true;
}

void HipSolver::setChiSpectrumMaterial(int material) {
  if (material >= 0 && (!materials_set_ || material >= M_))
    throw DeviceError(MOC_ERR_INVALID, "material id out of range");
  chi_material_ = material < 0 ? -1 : material;
}

void HipSolver::setSolverMode(int mode) {
  if (mode != MOC_FORWARD && mode != MOC_ADJOINT) throw DeviceError(MOC_ERR_INVALID, "unknown solver mode");
  if (mode == solver_mode_) return;
  solver_mode_ = mode;
  if (materials_set_) {
    HIP_CHECK(hipSetDevice(device_));
    uploadProductionMatrices(M_, G_);
    // the matrix-core operands follow the matrices in use
    if (tracks_set_ && fsrs_set_) prepareScatter();
  }
  // the fluxes of the last solve stay (Solver::setSolverMode only records the
  // mode, src/Solver.cpp:597-602): a restart in the other mode starts from
  // them, with sources from the matrices of the new mode
  krylov_ready_ = false;
}

void HipSolver::countFissionable() {
  num_fissionable_ = 0;
  for (int64_t r = 0; r < N_; r++)
    if (h_fsr_mat_[r] < (int)mat_fissionable_.size() && mat_fissionable_[h_fsr_mat_[r]])
      num_fissionable_++;
}

void HipSolver::setFSRs(int64_t num_fsrs, const int32_t* material, const double* volume) {
  HIP_CHECK(hipSetDevice(device_));
  if (geo_hold_) throw DeviceError(MOC_ERR_STATE, "this solver shares its geometry; use a new solver");
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
    float iv_max = 0.f;
    for (float v : iv) iv_max = std::max(iv_max, v);
    scale_cap_ = iv_max > 0.f ? 1e38 / iv_max : 1e38;
    dev_alloc(&d_inv_vol_, num_fsrs);
    HIP_CHECK(hipMemcpy(d_inv_vol_, iv.data(), num_fsrs * sizeof(float), hipMemcpyHostToDevice));
  }
  h_fsr_mat_.assign(material, material + num_fsrs);
  h_fsr_vol_.assign(volume, volume + num_fsrs);
  scat_tiles_ok_ = false;
  layout_tpw_ = 0;
  if (num_fsrs != N_) {
    fixed_on_ = false;
    h_fixed_.clear();
    fix_fsr_map_.clear();
    fix_mat_map_.clear();
    fix_cell_map_.clear();
    fixed_dirty_ = false;
  }
  N_ = num_fsrs;
  fsrs_set_ = true;
  ls_ready_ = false;
  countFissionable();
  allocateIterationArrays();
}

void HipSolver::allocateIterationArrays() {
  graph_dirty_ = true;
  state_valid_ = false;
  solved_ = false;
  if (!(tracks_set_ && materials_set_ && fsrs_set_)) return;
  const size_t NG = (size_t)N_ * G_;
  if (!shape_forced_) pickShape();
  buildTileLayout(sweepTracksPerWave());
  const size_t psi_n = psiCount();
  dev_alloc(&d_psi_[0], psi_n);
  dev_alloc(&d_psi_[1], psi_n);
  dev_alloc(&d_tally_, NG * shape_.tally_copies);
  dev_alloc(&d_qs_, NG);
  dev_alloc(&d_phi_, NG);
  dev_alloc(&d_phi_old_[0], NG);
  dev_alloc(&d_phi_old_[1], NG);
  dev_alloc(&d_leak_, (size_t)2 * T_ * G_);
  if (linear_) {
    dev_alloc(&d_phi_xy_, NG * 2);
    dev_alloc(&d_phi_xy_n_, NG * 2);
    dev_alloc(&d_Qxy_, NG * 2);
    dev_alloc(&d_qls_, NG);
    dev_alloc(&d_tally_xy_, NG * 2 * shape_.tally_copies);
    HIP_CHECK(hipMemsetAsync(d_tally_xy_, 0, NG * 2 * shape_.tally_copies * sizeof(float), stream_));
    HIP_CHECK(hipMemsetAsync(d_phi_xy_n_, 0, NG * 2 * sizeof(double), stream_));
    HIP_CHECK(hipMemsetAsync(d_phi_xy_, 0, NG * 2 * sizeof(double), stream_));
    ls_ready_ = false;
  }
  dev_alloc(&d_Q_, NG);
  if (!shape_forced_) pickShape();
  if (stackMode()) {
    const int nsub = (G_ + shape_.lds_groups - 1) / shape_.lds_groups;
    dev_alloc(&d_qsub_, (size_t)nsub * N_ * shape_.lds_groups);
    HIP_CHECK(hipMemsetAsync(d_qsub_, 0, (size_t)nsub * N_ * shape_.lds_groups * sizeof(float), stream_));
    dev_alloc(&d_tally_i64_, NG);
    HIP_CHECK(hipMemsetAsync(d_tally_i64_, 0, NG * sizeof(unsigned long long), stream_));
    // per-slot sums of the z-stack sweep and each FSR's slots (a counting
    // sort of the slots by FSR)
    const size_t nslots = h_slot_fsr_.size();
    std::vector<int64_t> off((size_t)N_ + 1, 0);
    for (int32_t f : h_slot_fsr_) {
      if (f < 0 || f >= N_) throw DeviceError(MOC_ERR_INVALID, "segment FSR id exceeds the number of FSRs");
      off[(size_t)f + 1]++;
    }
    for (int64_t r = 0; r < N_; r++) off[r + 1] += off[r];
    std::vector<int32_t> slots(std::max<size_t>(nslots, 1), 0);
    {
      std::vector<int64_t> cur(off.begin(), off.end() - 1);
      for (size_t q = 0; q < nslots; q++) slots[cur[h_slot_fsr_[q]]++] = (int32_t)q;
    }
    dev_alloc(&d_stack_partial_, std::max<size_t>(nslots, 1) * stackRow());
    dev_alloc(&d_stack_off_, off.size());
    dev_alloc(&d_stack_slots_, slots.size());
    HIP_CHECK(hipMemcpy(d_stack_off_, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(d_stack_slots_, slots.data(), slots.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  } else if (shape_.lds_groups > 0) {
    const int nsub = (G_ + shape_.lds_groups - 1) / shape_.lds_groups;
    if (shape_.lds_flush) {
      dev_alloc(&d_tally_i64_, NG);
      HIP_CHECK(hipMemsetAsync(d_tally_i64_, 0, NG * sizeof(unsigned long long), stream_));
    } else {
      dev_alloc(&d_lds_partials_, (size_t)nsub * shape_.lds_blocks_per_sub * N_ * shape_.lds_groups);
    }
    dev_alloc(&d_qsub_, (size_t)nsub * N_ * shape_.lds_groups);
    HIP_CHECK(hipMemsetAsync(d_qsub_, 0, (size_t)nsub * N_ * shape_.lds_groups * sizeof(float), stream_));
    dev_alloc(&d_tally64_, NG);
    HIP_CHECK(hipMemsetAsync(d_tally64_, 0, NG * sizeof(double), stream_));
  }
  if (segStore()) {
    buildTileLayout(sweepTracksPerWave());  // positions follow the final shape
    buildSegmentIndex();
    const int planes = linear_ ? 3 : 1;
    dev_alloc(&d_segout_, (size_t)planes * seg_rows_ * seg_stride_);
    dev_alloc(&d_tally64_, NG);
  }
  {
    // per-block partials: post_sweep_kernel runs over nsub*N*GB threads,
    // update_kernel over N*G
    const int gb = std::max(1, shape_.lds_groups);
    const int64_t npost = (int64_t)((G_ + gb - 1) / gb) * N_ * gb;
    num_partials_ = (int)blocks_for(npost, kReduceBlock);
    pstride_ = std::max(num_partials_, (int)blocks_for(NG, kReduceBlock));
    dev_alloc(&d_partials_, (size_t)kNumPartials * pstride_);
  }
  buildTileLayout(sweepTracksPerWave());
  HIP_CHECK(hipMemsetAsync(d_psi_[0], 0, psi_n * sizeof(float), stream_));
  HIP_CHECK(hipMemsetAsync(d_psi_[1], 0, psi_n * sizeof(float), stream_));
  HIP_CHECK(hipMemsetAsync(d_tally_, 0, NG * shape_.tally_copies * sizeof(float), stream_));
  HIP_CHECK(hipMemsetAsync(d_phi_, 0, NG * sizeof(double), stream_));
  HIP_CHECK(hipMemsetAsync(d_phi_old_[0], 0, NG * sizeof(double), stream_));
  HIP_CHECK(hipMemsetAsync(d_phi_old_[1], 0, NG * sizeof(double), stream_));
  HIP_CHECK(hipMemsetAsync(d_leak_, 0, (size_t)2 * T_ * G_ * sizeof(float), stream_));
  phi_is_old_ = false;
  krylov_ready_ = false;
  HIP_CHECK(hipMemsetAsync(d_Q_, 0, NG * sizeof(double), stream_));
  HIP_CHECK(hipMemsetAsync(d_qs_, 0, NG * sizeof(float2), stream_));
  if (fixed_on_ && h_fixed_.size() == NG) {
    dev_alloc(&d_fixed_, NG);
    HIP_CHECK(hipMemcpy(d_fixed_, h_fixed_.data(), NG * sizeof(double), hipMemcpyHostToDevice));
  }
  fixed_dirty_ = fixed_on_;
  psi_cur_ = 0;
  // the clears run on the solver's stream, which does not wait for the
  // default stream: they are complete before any later upload or launch
  HIP_CHECK(hipStreamSynchronize(stream_));
  prepareScatter();
}

int HipSolver::scatterPrecisionInUse() const {
  if (scatter_mode_ != MOC_SCATTER_AUTO) return scatter_mode_;
  return G_ >= kMfmaMinGroups ? MOC_SCATTER_MFMA_F16 : MOC_SCATTER_FP64;
}

void HipSolver::setScatterPrecision(int mode) {
  if (mode < MOC_SCATTER_AUTO || mode > MOC_SCATTER_FP32)
    throw DeviceError(MOC_ERR_INVALID, "unknown scattering precision");
  if (async_active_) throw DeviceError(MOC_ERR_STATE, "a solve is still running on this solver");
  HIP_CHECK(hipSetDevice(device_));
  scatter_mode_ = mode;
  graph_dirty_ = true;
  prepareScatter();
}

// FSRs grouped by material in tiles of kScatRows (ascending FSR ids within a
// material; the last tile of a material is padded with -1).
void HipSolver::buildScatterTiles() {
  std::vector<std::vector<int32_t>> by(M_);
  for (int64_t r = 0; r < N_; r++) by[h_fsr_mat_[r]].push_back((int32_t)r);
  std::vector<int32_t> rows, tmat;
  for (int m = 0; m < M_; m++)
    for (size_t s = 0; s < by[m].size(); s += kScatRows) {
      tmat.push_back(m);
      for (size_t j = 0; j < (size_t)kScatRows; j++)
        rows.push_back(s + j < by[m].size() ? by[m][s + j] : -1);
    }
  scat_ntiles_ = (int64_t)tmat.size();
  dev_alloc(&d_scat_rows_, rows.size());
  dev_alloc(&d_scat_tile_mat_, tmat.size());
  HIP_CHECK(hipMemcpy(d_scat_rows_, rows.data(), rows.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_scat_tile_mat_, tmat.data(), tmat.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  scat_tiles_ok_ = true;
}

// Buffers of the contraction path in use; the matrix-core operand fragments
// are rebuilt whenever the cross sections change.
void HipSolver::prepareScatter() {
  if (!(materials_set_ && fsrs_set_ && tracks_set_)) return;
  const int path = scatterPrecisionInUse();
  if (path == MOC_SCATTER_FP64) return;
  const int C = linear_ ? 3 : 1;
  dev_alloc(&d_scat_, (size_t)C * N_ * (G_ + 1));
  if (path != MOC_SCATTER_MFMA_F16) return;
  if (!scat_tiles_ok_) buildScatterTiles();
  if (!scat_frag_ok_) {
    const int ks_n = (G_ + 15) / 16, nt_n = (G_ + 1 + 15) / 16;
    const int64_t nfrag = (int64_t)M_ * ks_n * nt_n * 64;
    dev_alloc(&d_scat_b_, (size_t)nfrag * 4);
    dev_alloc(&d_scat_colmul_, (size_t)M_ * nt_n * 16);
    hipLaunchKernelGGL(scatter_frag_kernel, dim3(blocks_for(nfrag, 256)), dim3(256), 0, stream_,
                       d_sigma_s_, d_fis_nsf_, M_, G_, ks_n, nt_n,
                       reinterpret_cast<uint4*>(d_scat_b_), d_scat_colmul_);
    HIP_CHECK(hipGetLastError());
    scat_frag_ok_ = true;
  }
}

// The contraction of the current flux (and moments) the next update_kernel
// reads; nothing on the FP64 path, whose update_kernel loops over groups.
// Arguments of the scattering contraction; false on the FP64 path (no
// contraction launch).
bool HipSolver::scatArgs(ScatArgs& a) const {
  const int path = scatterPrecisionInUse();
  if (path == MOC_SCATTER_FP64) return false;
  if (!d_scat_ || (path == MOC_SCATTER_MFMA_F16 && !(scat_tiles_ok_ && scat_frag_ok_)))
    throw DeviceError(MOC_ERR_STATE, "scattering contraction buffers are not prepared");
  a.N = N_;
  a.G = G_;
  a.ks_n = (G_ + 15) / 16;
  a.nt_n = (G_ + 1 + 15) / 16;
  a.ntiles = scat_ntiles_;
  a.rows = d_scat_rows_;
  a.tile_mat = d_scat_tile_mat_;
  a.bfrag = reinterpret_cast<const uint4*>(d_scat_b_);
  a.colmul = d_scat_colmul_;
  a.sigma_s = d_sigma_s_;
  a.nsf = d_fis_nsf_;
  a.fsr_mat = d_fsr_mat_;
  a.C = linear_ ? 3 : 1;
  a.x[0] = d_phi_;
  a.xs[0] = 1;
  a.x[1] = linear_ ? d_phi_xy_ : nullptr;
  a.xs[1] = 2;
  a.x[2] = linear_ ? d_phi_xy_ + 1 : nullptr;
  a.xs[2] = 2;
  a.out = d_scat_;
  a.state = d_state_;
  return true;
}

void HipSolver::launchScatter(hipStream_t s) {
  ScatArgs a;
  if (!scatArgs(a)) return;
  const int path = scatterPrecisionInUse();
  if (path == MOC_SCATTER_MFMA_F16)
    hipLaunchKernelGGL(scatter_mfma_kernel, dim3(blocks_for(scat_ntiles_, 4)), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(scatter_fp32_kernel, dim3(blocks_for((int64_t)a.C * N_ * (G_ + 1), 256)),
                       dim3(256), 0, s, a);
  HIP_CHECK(hipGetLastError());
}

void HipSolver::requireReady() const {
  if (!tracks_set_) throw DeviceError(MOC_ERR_STATE, "tracks have not been set");
  if (!materials_set_) throw DeviceError(MOC_ERR_STATE, "materials have not been set");
  if (!fsrs_set_) throw DeviceError(MOC_ERR_STATE, "FSRs have not been set");
  if (async_active_) throw DeviceError(MOC_ERR_STATE, "a solve is still running on this solver");
}

void HipSolver::setConvergenceThreshold(double t) {
  graph_dirty_ = true;
  if (!(t > 0.0)) throw DeviceError(MOC_ERR_INVALID, "convergence threshold must be > 0");
  threshold_ = t;
}

void HipSolver::setExponential(int mode) {
  graph_dirty_ = true;
  if (mode != MOC_EXP_RATIONAL && mode != MOC_EXP_HW)
    throw DeviceError(MOC_ERR_INVALID, "unknown exponential mode");
  exp_mode_ = mode;
}

// Solver::setFixedSourceByFSR (src/Solver.cpp:479-499): recorded in the
// per-FSR map, applied when the next solve starts.
void HipSolver::setFixedSourceByFSR(int64_t fsr, int group, double source) {
  graph_dirty_ = true;
  requireReady();
  if (fsr < 0 || fsr >= N_) throw DeviceError(MOC_ERR_INVALID, "FSR id out of range");
  if (group < 1 || group > G_) throw DeviceError(MOC_ERR_INVALID, "group out of range (1-based)");
  fix_fsr_map_[{fsr, group}] = source;
  fixed_on_ = true;
  fixed_dirty_ = true;
}

// Solver::setFixedSourceByMaterial (src/Solver.cpp:538-543): every FSR of the
// material, applied when the next solve starts after the per-FSR sources.
void HipSolver::setFixedSourceByMaterial(int material, int group, double source) {
  graph_dirty_ = true;
  requireReady();
  if (material < 0 || material >= M_) throw DeviceError(MOC_ERR_INVALID, "material id out of range");
  if (group < 1 || group > G_) throw DeviceError(MOC_ERR_INVALID, "group out of range (1-based)");
  fix_mat_map_[{material, group}] = source;
  fixed_on_ = true;
  fixed_dirty_ = true;
}

// Geometry::findCellContainingFSR (src/Geometry.cpp) as a table: the id of
// the material cell holding each FSR, for setFixedSourceByCell.
void HipSolver::setFSRCells(const int32_t* cell, int64_t n) {
  requireReady();
  if (!cell || n != N_) throw DeviceError(MOC_ERR_INVALID, "FSR cell table must hold N values");
  h_fsr_cell_.assign(cell, cell + n);
  fixed_dirty_ = fixed_on_;
}

// Solver::setFixedSourceByCell (src/Solver.cpp:510-527): every FSR of the
// cell, applied when the next solve starts, before the material sources.
void HipSolver::setFixedSourceByCell(int cell, int group, double source) {
  graph_dirty_ = true;
  requireReady();
  if ((int64_t)h_fsr_cell_.size() != N_)
    throw DeviceError(MOC_ERR_STATE, "the FSR cell table has not been set");
  if (group < 1 || group > G_) throw DeviceError(MOC_ERR_INVALID, "group out of range (1-based)");
  if (std::find(h_fsr_cell_.begin(), h_fsr_cell_.end(), cell) == h_fsr_cell_.end())
    throw DeviceError(MOC_ERR_INVALID, "no FSR lies in this cell");
  fix_cell_map_[{cell, group}] = source;
  fixed_on_ = true;
  fixed_dirty_ = true;
}

// Solver::initializeFixedSources + CPUSolver::initializeFixedSources
// (src/Solver.cpp:1089-1137, src/CPUSolver.cpp:425-456): each cell entry, then
// each material entry, is written into the per-FSR map for every FSR of the
// cell or material (where it stays, as in the reference), then the per-FSR
// map fills the array.
void HipSolver::applyFixedSources() {
  if (!fixed_on_ || !fixed_dirty_) return;
  if ((int64_t)h_fsr_cell_.size() == N_)
    for (const auto& e : fix_cell_map_)
      for (int64_t r = 0; r < N_; r++)
        if (h_fsr_cell_[r] == e.first.first) fix_fsr_map_[{r, e.first.second}] = e.second;
  for (const auto& e : fix_mat_map_)
    for (int64_t r = 0; r < N_; r++)
      if (h_fsr_mat_[r] == e.first.first) fix_fsr_map_[{r, e.first.second}] = e.second;
  const size_t NG = (size_t)N_ * G_;
  h_fixed_.assign(NG, 0.0);
  for (const auto& e : fix_fsr_map_) h_fixed_[(size_t)e.first.first * G_ + e.first.second - 1] = e.second;
  HIP_CHECK(hipSetDevice(device_));
  dev_alloc(&d_fixed_, NG);
  HIP_CHECK(hipMemcpy(d_fixed_, h_fixed_.data(), NG * sizeof(double), hipMemcpyHostToDevice));
  fixed_dirty_ = false;
}

// CPUSolver::resetFixedSources (src/CPUSolver.cpp:228-250): every recorded
// fixed source and the array become zero (the solver still counts as having
// fixed sources).
void HipSolver::resetFixedSources() {
  graph_dirty_ = true;
  requireReady();
  for (auto& e : fix_fsr_map_) e.second = 0.0;
  for (auto& e : fix_mat_map_) e.second = 0.0;
  for (auto& e : fix_cell_map_) e.second = 0.0;
  if (h_fixed_.empty()) return;
  std::fill(h_fixed_.begin(), h_fixed_.end(), 0.0);
  HIP_CHECK(hipSetDevice(device_));
  if (d_fixed_)
    HIP_CHECK(hipMemcpy(d_fixed_, h_fixed_.data(), h_fixed_.size() * sizeof(double),
                        hipMemcpyHostToDevice));
}

// Arguments of the 2D / 3D sweep kernels over this solver's arrays, or over
// caller-provided buffers (the allocation-free sweep).
void HipSolver::sweepArgs(SweepArgs& a, const float* psi_in, float* psi_out, float* leak,
                          const SweepBuffers* bufs) const {
  if (layout_tpw_ != sweepTracksPerWave())
    throw DeviceError(MOC_ERR_STATE, "segment streams are not laid out for the launch shape");
  a.work = d_work_;
  a.ntiles = layout_ntiles_;
  a.tpw = layout_tpw_;
  a.tile = reinterpret_cast<const int2*>(d_tile_);
  a.trk = reinterpret_cast<const int4*>(d_trk_);
  a.trk_azim = d_trk_azim_;
  a.seg_len = d_seg_len_;
  a.seg_key = d_seg_key_;
  a.stream_bytes = (unsigned)(((size_t)layout_rows_ + kTailRows) * layout_tpw_ * sizeof(float));
  a.qs = d_qs_;
  a.psi_in = psi_in;
  a.psi_out = psi_out;
  a.out_wp = d_out_wp_;
  a.psi_w = (int)psi_w_;
  a.leak = leak;
  a.tally = d_tally_;
  a.tally_copies = shape_.tally_copies;
  a.tally_stride = (size_t)N_ * G_;
  a.segout = (!bufs && segStore()) ? d_segout_ : nullptr;
  a.seg_slot = d_seg_slot_;
  a.seg_stride = seg_stride_;
  a.weights = d_weights_;
  const float l2e = exp_mode_ == MOC_EXP_HW ? 1.4426950408889634f : 1.0f;
  a.pf0 = inv_sin_[0] * l2e;
  a.pf1 = inv_sin_[1] * l2e;
  a.pf2 = inv_sin_[2] * l2e;
  a.state = d_state_;
  a.G = G_;
  a.P = P_;
  a.fpt = P_ * G_;
  a.N = (int)N_;
  a.lanes_per_track = shape_.lanes_per_track;
  a.tracks_per_wave = shape_.tracks_per_wave;
  if (bufs) {
    // caller-provided buffers (the allocation-free sweep): the LDS sweep then
    // always flushes into the int64 tally
    a.qs = bufs->qs;
    a.tally = bufs->tally;
    a.state = bufs->state;
  }
  a.psi_max_bits = &const_cast<IterState*>(a.state)->psi_max_bits;
}

// Parameters of the 2D LDS sweep (shape_.lds_groups > 0).
void HipSolver::ldsParams(LdsParams& L, const SweepBuffers* bufs) const {
  L.GB = shape_.lds_groups;
  L.nsub = (G_ + L.GB - 1) / L.GB;
  L.blocks_per_sub = shape_.lds_blocks_per_sub;
  L.partials = (shape_.lds_flush || bufs) ? nullptr : d_lds_partials_;
  L.tally64 = bufs ? bufs->tally_i64 : shape_.lds_flush ? d_tally_i64_ : nullptr;
  L.qsub = bufs ? bufs->qsub : d_qsub_;
  L.inv_vol = d_inv_vol_;
  L.sigma_t = d_sigma_t_;
  L.M = M_;
}

void HipSolver::launchSweep(const float* psi_in, float* psi_out, hipStream_t s,
                            hipEvent_t ev_after_sweep, float* leak, const SweepBuffers* bufs) {
  SweepArgs a;
  sweepArgs(a, psi_in, psi_out, leak, bufs);
  if (linear_) {
    if (bufs) throw DeviceError(MOC_ERR_INVALID, "the linear source has no caller-buffer sweep");
    if (!ls_ready_) throw DeviceError(MOC_ERR_STATE, "linear-source data not prepared");
    LsArgs L{d_seg_pos_, d_qls_, d_tally_xy_, d_azim_cos_, d_azim_sin_,
             a.segout, seg_stride_, (size_t)seg_rows_ * seg_stride_};
    // the reference's LS exponentials take plain 1/sin(theta)
    a.pf0 = inv_sin_[0];
    a.pf1 = inv_sin_[1];
    a.pf2 = inv_sin_[2];
    const unsigned grid = (unsigned)(((int64_t)layout_ntiles_ * 64 + shape_.block - 1) / shape_.block);
    auto by_np = [&](auto k1, auto k2, auto k3) {
      switch (P_) {
        case 1: hipLaunchKernelGGL(k1, dim3(grid), dim3(shape_.block), 0, s, a, L); break;
        case 2: hipLaunchKernelGGL(k2, dim3(grid), dim3(shape_.block), 0, s, a, L); break;
        default: hipLaunchKernelGGL(k3, dim3(grid), dim3(shape_.block), 0, s, a, L); break;
      }
    };
    switch (shape_.groups_per_lane) {
      case 1: by_np(sweep2d_ls_kernel<1, 1>, sweep2d_ls_kernel<1, 2>, sweep2d_ls_kernel<1, 3>); break;
      case 2: by_np(sweep2d_ls_kernel<2, 1>, sweep2d_ls_kernel<2, 2>, sweep2d_ls_kernel<2, 3>); break;
      case 4: by_np(sweep2d_ls_kernel<4, 1>, sweep2d_ls_kernel<4, 2>, sweep2d_ls_kernel<4, 3>); break;
      case 9: by_np(sweep2d_ls_kernel<9, 1>, sweep2d_ls_kernel<9, 2>, sweep2d_ls_kernel<9, 3>); break;
      default: by_np(sweep2d_ls_kernel<16, 1>, sweep2d_ls_kernel<16, 2>, sweep2d_ls_kernel<16, 3>); break;
    }
    HIP_CHECK(hipGetLastError());
    if (ev_after_sweep) HIP_CHECK(hipEventRecord(ev_after_sweep, s));
    return;
  }
  if (stackMode()) {
    StackParams L;
    L.grp = reinterpret_cast<const int4*>(d_grp_);
    L.slot_fsr = d_slot_fsr_;
    L.tally = bufs ? bufs->tally_i64 : d_tally_i64_;
    // the solver's own sweep writes per-slot sums and reduces them per FSR;
    // on caller buffers (no room for the slot sums) the blocks add theirs
    // to the tally with atomics
    L.partial = bufs ? nullptr : d_stack_partial_;
    L.prow = stackRow();
    L.qsub = bufs ? bufs->qsub : d_qsub_;
    L.inv_vol = d_inv_vol_;
    L.sigma_t = d_sigma_t_;
    L.M = M_;
    L.GB = shape_.lds_groups;
    L.nsub = (G_ + L.GB - 1) / L.GB;
    L.ngrp = ngrp_;
    L.max_slots = max_grp_slots_;
    const size_t lds_bytes = ldsBytes(L.GB, 1);
    if (lds_bytes > lds_max_) throw DeviceError(MOC_ERR_INVALID, "LDS slice exceeds the device limit");
    if (layout_grp_ntiles_.size() != (size_t)ngrp_)
      throw DeviceError(MOC_ERR_STATE, "segment streams are not laid out by z-stack group");
    // groups padded to a multiple of the XCD count; block b -> group, subset
    const unsigned grid = (unsigned)((ngrp_ + kNumXcd - 1) / kNumXcd * kNumXcd * L.nsub);
    auto launch = [&](auto kernel) {
      ensure_lds_attribute((const void*)kernel, device_, lds_max_);
      hipLaunchKernelGGL(kernel, dim3(grid), dim3(shape_.lds_block), lds_bytes, s, a, L);
    };
    if (exp_mode_ == MOC_EXP_HW)
      launch(sweep3d_stack_kernel<MOC_EXP_HW>);
    else
      launch(sweep3d_stack_kernel<MOC_EXP_RATIONAL>);
    if (L.partial)
      hipLaunchKernelGGL(stack_reduce_kernel, dim3(blocks_for((int64_t)N_ * G_, 256)), dim3(256), 0, s,
                         L.partial, d_stack_off_, d_stack_slots_, d_tally_i64_, N_, G_, L.prow, a.state);
    HIP_CHECK(hipGetLastError());
    if (ev_after_sweep) HIP_CHECK(hipEventRecord(ev_after_sweep, s));
    return;
  }
  if (shape_.lds_groups > 0) {
    LdsParams L;
    ldsParams(L, bufs);
    const size_t lds_bytes = ldsBytes(L.GB, shape_.lds_q);
    if (lds_bytes > lds_max_) throw DeviceError(MOC_ERR_INVALID, "LDS slice exceeds the device limit");
    const unsigned grid = (unsigned)(L.nsub * L.blocks_per_sub);
    auto launch = [&](auto kernel) {
      ensure_lds_attribute((const void*)kernel, device_, lds_max_);
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
    auto by_exp = [&](auto st) {
      constexpr int ST = decltype(st)::value;
      if (exp_mode_ == MOC_EXP_HW)
        by_np(sweep2d_kernel<GPL, MOC_EXP_HW, 1, ST>, sweep2d_kernel<GPL, MOC_EXP_HW, 2, ST>,
              sweep2d_kernel<GPL, MOC_EXP_HW, 3, ST>);
      else
        by_np(sweep2d_kernel<GPL, MOC_EXP_RATIONAL, 1, ST>, sweep2d_kernel<GPL, MOC_EXP_RATIONAL, 2, ST>,
              sweep2d_kernel<GPL, MOC_EXP_RATIONAL, 3, ST>);
    };
    if (a.segout)
      by_exp(std::integral_constant<int, 1>{});
    else
      by_exp(std::integral_constant<int, 0>{});
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

FsrArgs HipSolver::fsrArgs() {
  if (linear_ && !ls_ready_)
    throw DeviceError(MOC_ERR_STATE, "linear-source data not prepared");
  FsrArgs a;
  a.N = N_;
  a.G = G_;
  a.fsr_mat = d_fsr_mat_;
  a.fsr_vol = d_fsr_vol_;
  a.sigma_t = d_sigma_t_;
  a.sigma_s = d_sigma_s_;
  a.nsf = d_nsf_;
  a.chi = d_chi_;
  a.fis_nsf = d_fis_nsf_;
  a.sigma_a = d_sigma_a_;
  a.fissionable = d_mat_fissionable_;
  a.fixed = fixed_on_ ? d_fixed_ : nullptr;
  a.phi = d_phi_;
  a.phi_old[0] = d_phi_old_[0];
  a.phi_old[1] = d_phi_old_[1];
  a.Q = d_Q_;
  a.tally = d_tally_;
  a.tally_copies = shape_.tally_copies;
  a.leak = balance_ ? d_leak_ : nullptr;
  a.leak_n = (int64_t)2 * T_ * G_;
  a.partials = d_partials_;
  a.pstride = pstride_;
  a.state = d_state_;
  a.allow_negative = allow_negative_ ? 1 : 0;
  a.stab_type = stab_type_;
  a.stab_factor = stab_factor_;
  a.stab_ratio = d_stab_ratio_;
  a.linear = linear_ ? 1 : 0;
  a.tally_xy = d_tally_xy_;
  a.phi_xy = d_phi_xy_;
  a.phi_xy_n = d_phi_xy_n_;
  a.Qxy = d_Qxy_;
  a.qls = d_qls_;
  a.ilem = d_ilem_;
  a.src_c = d_src_c_;
  a.scat = scatterPrecisionInUse() == MOC_SCATTER_FP64 ? nullptr : d_scat_;
  a.ss_colsum = d_ss_colsum_;
  a.ref_flux = (!ref_flux_file_.empty() && ref_flux_n_ == (size_t)N_ * G_) ? d_ref_flux_ : nullptr;
  a.tally_i64 = nullptr;
  a.lds_order = 0;
  a.segout = segStore() ? d_segout_ : nullptr;
  a.csr_off = d_csr_off_;
  a.seg_stride = seg_stride_;
  a.seg_plane = (size_t)seg_rows_ * seg_stride_;
  if (stackMode()) {
    a.qs = nullptr;
    a.qsub = d_qsub_;
    a.qsub_groups = shape_.lds_groups;
    a.lds_partials = nullptr;
    a.lds_bps = 0;
    a.tally_i64 = d_tally_i64_;
  } else if (shape_.lds_groups > 0) {
    a.qs = nullptr;
    a.qsub = d_qsub_;
    a.qsub_groups = shape_.lds_groups;
    a.lds_partials = shape_.lds_flush ? nullptr : d_lds_partials_;
    a.lds_bps = shape_.lds_flush ? 0 : shape_.lds_blocks_per_sub;
    a.tally_i64 = shape_.lds_flush ? d_tally_i64_ : nullptr;
    a.lds_order = 1;
  } else {
    a.qs = d_qs_;
    a.qsub = nullptr;
    a.qsub_groups = 1;
    a.lds_partials = nullptr;
    a.lds_bps = 0;
  }
  return a;
}

// One source iteration; all launches on stream_, no host synchronisation.
void HipSolver::launchIteration(Mode mode, int res_type, bool recompute_sources, double tol,
                                hipEvent_t ev0, hipEvent_t ev1) {
  const FsrArgs fa = fsrArgs();
  const bool balance = balance_ && mode == MODE_EIGEN;
  if (ev0) HIP_CHECK(hipEventRecord(ev0, stream_));
  launchSweep(d_psi_[psi_cur_], d_psi_[psi_cur_ ^ 1], stream_, ev1, balance ? d_leak_ : nullptr);
  psi_cur_ ^= 1;
  hipLaunchKernelGGL(post_sweep_kernel, dim3(num_partials_), dim3(kReduceBlock), 0, stream_, fa, 1,
                     balance ? 1 : 0);
  // eigen: keff_kernel's step runs inside update_kernel (one launch fewer)
  // when the post-sweep partials are few enough
  const int keff_parts = mode == MODE_EIGEN && num_partials_ <= kFuseKeffMaxParts ? num_partials_ : 0;
  if (mode == MODE_EIGEN && keff_parts == 0)
    hipLaunchKernelGGL(keff_kernel, dim3(1), dim3(kReduceBlock), 0, stream_, d_partials_,
                       num_partials_, pstride_, d_state_, (double)N_, 1, balance ? 1 : 0);
  const double norm = (mode == MODE_EIGEN && res_type == MOC_FISSION_SOURCE)
                          ? (double)num_fissionable_
                          : (double)N_;
  const int rt = mode == MODE_FLUX ? MOC_SCALAR_FLUX : res_type;
  const unsigned nb_upd = blocks_for((int64_t)N_ * G_, kReduceBlock);
  if (recompute_sources) launchScatter(stream_);
  auto* const upd = keff_parts > 0 ? update_kernel<1> : update_kernel<0>;
  hipLaunchKernelGGL(upd, dim3(nb_upd), dim3(kReduceBlock), 0, stream_, fa,
                     mode == MODE_EIGEN ? 1 : 0, rt, 1, recompute_sources ? 1 : 0, 1, kSrcTotal,
                     keff_parts, balance ? 1 : 0, (double)N_);
  if (rt != MOC_SCALAR_FLUX && G_ >= kFsrResidualGroups)
    hipLaunchKernelGGL(fsr_residual_kernel, dim3(blocks_for(N_, kReduceBlock)), dim3(kReduceBlock), 0,
                       stream_, fa, mode == MODE_EIGEN ? 1 : 0, rt);
  hipLaunchKernelGGL(residual_kernel, dim3(1), dim3(kReduceBlock), 0, stream_,
                     d_partials_ + (size_t)kPartResidual * pstride_,
                     (int)nb_upd, d_state_, rt == MOC_FISSION_SOURCE ? (double)num_fissionable_ : norm,
                     tol, mode == MODE_EIGEN ? 0 : 1, sigma_t_max_, recompute_sources ? 1 : 0);
  HIP_CHECK(hipGetLastError());
}

// Fresh state for a solve: zero angular fluxes and tally, initial guess,
// its normalisation (eigen) and the sources of iteration 0.
void HipSolver::launchInit(Mode mode, double k0) {
  // the linear source's per-FSR data (expansion matrices, source constants,
  // position stream) must exist before its kernels read them
  if (linear_ && !ls_ready_) prepareLinearSource();
  restart_active_ = false;
  const bool cont = mode == MODE_FLUX ? continue_flux_ : restart_status_;
  continue_flux_ = false;
  if (cont) {
    launchRestart(mode, k0);
    return;
  }
  const size_t NG = (size_t)N_ * G_;
  const size_t psi_n = psiCount();
  // a flux file's fluxes and k_eff replace the flat guess of eigenvalue and
  // source solves (loadFSRFluxes(file, true), src/Solver.cpp:1733-1736)
  const bool restart = mode != MODE_FLUX && h_init_flux_.size() == NG;
  if (restart) k0 = init_keff_;
  IterState st;
  std::memset(&st, 0, sizeof(st));
  st.k = k0;
  st.k_prev = st.k;
  st.norm_factor = 1.0;
  st.psi_scale = 1.0;
  st.scale_cap = scale_cap_;
  st.max_iters = device_max_iters_;
  HIP_CHECK(hipMemcpyAsync(d_state_, &st, sizeof(st), hipMemcpyHostToDevice, stream_));
  // fresh angular fluxes and tally (initializeFluxArrays, src/CPUSolver.cpp:281-380)
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(psi_n, 256)), dim3(256), 0, stream_, d_psi_[0], (int64_t)psi_n);
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(psi_n, 256)), dim3(256), 0, stream_, d_psi_[1], (int64_t)psi_n);
  zeroTallies(stream_);
  psi_cur_ = 0;

  refreshStabilization();
  const FsrArgs fa = fsrArgs();
  // initial guess (computeInitialFluxGuess, src/Solver.cpp:1710-1731; computeFlux
  // starts from phi = 0, src/Solver.cpp:1383-1390)
  if (restart) {
    HIP_CHECK(hipMemcpyAsync(d_phi_, h_init_flux_.data(), NG * sizeof(double),
                             hipMemcpyHostToDevice, stream_));
  } else if (mode != MODE_FLUX && chi_material_ >= 0) {
    // flattenFSRFluxesChiSpectrum (src/CPUSolver.cpp:1829-1840): every FSR
    // starts from the chi of the chosen material
    if (chi_material_ >= M_ || h_chi_fwd_.size() != (size_t)M_ * G_)
      throw DeviceError(MOC_ERR_STATE, "the chi-spectrum material is not among the materials");
    h_chi_guess_.resize(NG);
    for (int64_t r = 0; r < N_; r++)
      for (int g = 0; g < G_; g++)
        h_chi_guess_[(size_t)r * G_ + g] = h_chi_fwd_[(size_t)chi_material_ * G_ + g];
    HIP_CHECK(hipMemcpyAsync(d_phi_, h_chi_guess_.data(), NG * sizeof(double),
                             hipMemcpyHostToDevice, stream_));
  } else {
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_phi_,
                       (int64_t)NG, mode == MODE_FLUX ? 0.0 : 1.0);
  }
  if (linear_)
    // flattenFSRFluxes zeroes the moments (src/CPULSSolver.cpp:343-356)
    hipLaunchKernelGGL(fill_kernel, dim3(blocks_for(2 * NG, 256)), dim3(256), 0, stream_,
                       d_phi_xy_, (int64_t)(2 * NG), 0.0);
  // the loaded fluxes are normalised in source solves too (src/Solver.cpp:1736)
  const bool normalize = mode == MODE_EIGEN || restart;
  if (normalize) {
    hipLaunchKernelGGL(post_sweep_kernel, dim3(num_partials_), dim3(kReduceBlock), 0, stream_, fa, 0, 0);
    hipLaunchKernelGGL(keff_kernel, dim3(1), dim3(kReduceBlock), 0, stream_, d_partials_,
                       num_partials_, pstride_, d_state_, (double)N_, 0, 0);
  }
  // normalise, store old flux, sources of iteration 0
  launchScatter(stream_);
  hipLaunchKernelGGL(update_kernel<0>, dim3(blocks_for(NG, kReduceBlock)), dim3(kReduceBlock), 0,
                     stream_, fa, normalize ? 1 : 0, 0, 0, 1, 0, kSrcTotal, 0, 0, 0.0);
  // fixed-point scale of the first sweep (start fluxes are zero); the old
  // flux just written becomes the current one
  hipLaunchKernelGGL(tally_scale_kernel, dim3(1), dim3(1), 0, stream_, d_state_, sigma_t_max_, 1, 1);
  HIP_CHECK(hipGetLastError());
  if (restart) {
    launchStartupSweeps();
    // the copy above is complete once the startup sweeps synchronised
    h_init_flux_.clear();
  }

  state_valid_ = true;
  phi_is_old_ = true;
  krylov_ready_ = false;
}

// A solve that continues the previous one (setRestartStatus, or computeFlux
// without only_fixed_source): k_eff (kept for eigenvalue solves,
// src/Solver.cpp:1563-1564; set to k0 by computeSource and computeFlux,
// :1474, :1366), the start fluxes with their pending normalisation and
// the current scalar flux carry over; the flux is stored as the old flux and
// gives the sources of iteration 0, without the flat guess or its
// normalisation (src/Solver.cpp:1721-1731).
void HipSolver::launchRestart(Mode mode, double k0) {
  if (!solved_ || !state_valid_)
    throw DeviceError(MOC_ERR_STATE, "a restart needs an earlier solve on this solver");
  if (h_init_flux_.size() == (size_t)N_ * G_)
    throw DeviceError(MOC_ERR_INVALID, "a restart and initial fluxes from a file are exclusive");
  const size_t NG = (size_t)N_ * G_;
  HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  IterState st = *h_state_;
  // computeSource sets k_eff (src/Solver.cpp:1474), computeFlux sets 1
  // (src/Solver.cpp:1366-1367)
  if (mode != MODE_EIGEN) st.k = k0;
  st.k_prev = st.k;
  st.residual = 0.0;
  st.iteration = 0;
  st.converged = 0;
  st.dk = 0;
  st.reported_iters = 0;
  st.negative_fluxes = 0;
  st.negative_sources = 0;
  // the next sweep's start fluxes are the last sweep's outgoing ones, bounded
  // by its running maximum (psi_scale still applies to them)
  st.psi_max_bits = __builtin_bit_cast(uint32_t, st.psi_max);
  st.qmax_bits = 0u;
  *h_state_ = st;
  HIP_CHECK(hipMemcpyAsync(d_state_, h_state_, sizeof(IterState), hipMemcpyHostToDevice, stream_));
  if (phi_is_old_)
    HIP_CHECK(hipMemcpyAsync(d_phi_, d_phi_old_[st.old_sel & 1], NG * sizeof(double),
                             hipMemcpyDeviceToDevice, stream_));
  // linear source: the normalised flux moments of the last iteration too
  if (linear_)
    HIP_CHECK(hipMemcpyAsync(d_phi_xy_, d_phi_xy_n_, 2 * NG * sizeof(double),
                             hipMemcpyDeviceToDevice, stream_));
  zeroTallies(stream_);
  refreshStabilization();
  const FsrArgs fa = fsrArgs();
  // storeFSRFluxes and the sources of iteration 0, no normalisation
  launchScatter(stream_);
  hipLaunchKernelGGL(update_kernel<0>, dim3(blocks_for(NG, kReduceBlock)), dim3(kReduceBlock), 0,
                     stream_, fa, 0, 0, 0, 1, 0, kSrcTotal, 0, 0, 0.0);
  hipLaunchKernelGGL(tally_scale_kernel, dim3(1), dim3(1), 0, stream_, d_state_, sigma_t_max_, 1, 1);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(stream_));
  // only computeEigenvalue counts a restart as an iteration (src/Solver.cpp:1554)
  restart_active_ = mode == MODE_EIGEN;
  state_valid_ = true;
  phi_is_old_ = true;
  krylov_ready_ = false;
}

// Zero tallies that sweeps accumulate into (the LDS sweep's per-block
// partials are overwritten instead).
void HipSolver::zeroTallies(hipStream_t s) {
  const size_t NG = (size_t)N_ * G_;
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG * shape_.tally_copies, 256)), dim3(256), 0, s,
                     d_tally_, (int64_t)(NG * shape_.tally_copies));
  if (int64Tally())
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(2 * NG, 256)), dim3(256), 0, s,
                       reinterpret_cast<float*>(d_tally_i64_), (int64_t)(2 * NG));
  if (linear_)
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(2 * NG * shape_.tally_copies, 256)),
                       dim3(256), 0, s, d_tally_xy_, (int64_t)(2 * NG * shape_.tally_copies));
}

// After a flux file's fluxes were normalised, stored and turned into the
// sources of iteration 0: 30 sweeps on those sources converge the angular
// fluxes, then addSourceToScalarFlux (src/Solver.cpp:1744-1751).  The source
// iteration then starts from computeFSRSources(0) of that flux, the old flux
// staying the loaded one.
void HipSolver::launchStartupSweeps() {
  constexpr int kStartupSweeps = 30;
  const size_t NG = (size_t)N_ * G_;
  // the start fluxes are zero, so the normalisation has nothing to scale in
  // them and the 30 sweeps run unscaled
  HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  h_state_->psi_scale = 1.0;
  HIP_CHECK(hipMemcpyAsync(d_state_, h_state_, sizeof(IterState), hipMemcpyHostToDevice, stream_));
  const FsrArgs fa = fsrArgs();
  for (int i = 0; i < kStartupSweeps; i++) {
    // transportSweep starts from a zero scalar-flux tally (src/CPUSolver.cpp:2338-2389)
    zeroTallies(stream_);
    // fixed-point scale from the previous sweep's outgoing angular fluxes
    if (i > 0)
      hipLaunchKernelGGL(tally_scale_kernel, dim3(1), dim3(1), 0, stream_, d_state_, sigma_t_max_, 0, 0);
    launchSweep(d_psi_[psi_cur_], d_psi_[psi_cur_ ^ 1], stream_);
    psi_cur_ ^= 1;
  }
  hipLaunchKernelGGL(post_sweep_kernel, dim3(num_partials_), dim3(kReduceBlock), 0, stream_, fa, 1, 0);
  // computeFSRSources(0) of this flux, unnormalised; the old flux is kept
  launchScatter(stream_);
  hipLaunchKernelGGL(update_kernel<0>, dim3(blocks_for(NG, kReduceBlock)), dim3(kReduceBlock), 0,
                     stream_, fa, 0, 0, 0, 1, 0, kSrcTotal, 0, 0, 0.0);
  hipLaunchKernelGGL(tally_scale_kernel, dim3(1), dim3(1), 0, stream_, d_state_, sigma_t_max_, 1, 0);
  HIP_CHECK(hipGetLastError());
}

// YAMAMOTO: per group, the factor times the largest |sigma_s(g->g) / sigma_t(g)|
// over the FSRs' materials (src/CPUSolver.cpp:2692-2710).
void HipSolver::refreshStabilization() {
  if (stab_type_ != 1) return;
  std::vector<int> used(M_, 0);
  for (int64_t r = 0; r < N_; r++) used[h_fsr_mat_[r]] = 1;
  std::vector<double> ratio(G_, 0.0);
  for (int g = 0; g < G_; g++) {
    double mx = 0.0;
    for (int m = 0; m < M_; m++)
      if (used[m]) {
        const double v = std::fabs(h_sigma_s_[((size_t)m * G_ + g) * G_ + g] / h_sigma_t_[(size_t)m * G_ + g]);
        if (v > mx) mx = v;
      }
    ratio[g] = mx * stab_factor_;
  }
  dev_alloc(&d_stab_ratio_, G_);
  HIP_CHECK(hipMemcpy(d_stab_ratio_, ratio.data(), G_ * sizeof(double), hipMemcpyHostToDevice));
}

void HipSolver::stabilizeTransport(double factor, int type) {
  graph_dirty_ = true;
  if (type < 0) {
    stab_type_ = -1;
    return;
  }
  if (type > 2) throw DeviceError(MOC_ERR_INVALID, "unknown stabilization type");
  if (!(factor > 0.0) && type != 1) throw DeviceError(MOC_ERR_INVALID, "stabilization factor must be positive");
  stab_type_ = type;
  stab_factor_ = factor;
}

const double* HipSolver::currentFlux() const {
  return phi_is_old_ ? d_phi_old_[old_sel_] : d_phi_;
}

// One Krylov operator sweep (Solver::fissionTransportSweep /
// scatterTransportSweep): sources from the current flux, the sweep from the
// start fluxes left by the previous operator sweep (zero at first, as after
// initializeFluxArrays), then addSourceToScalarFlux into the flux.
void HipSolver::krylovSweep(int source_kind) {
  requireReady();
  if (linear_) throw DeviceError(MOC_ERR_INVALID, "Krylov operator sweeps are flat-source only");
  if (source_kind != kSrcFission && source_kind != kSrcScatter)
    throw DeviceError(MOC_ERR_INVALID, "unknown Krylov source kind");
  HIP_CHECK(hipSetDevice(device_));
  const size_t NG = (size_t)N_ * G_;
  const size_t psi_n = psiCount();
  if (phi_is_old_) {
    // the flux of the last solve becomes the operand
    HIP_CHECK(hipMemcpyAsync(d_phi_, d_phi_old_[old_sel_], NG * sizeof(double),
                             hipMemcpyDeviceToDevice, stream_));
    phi_is_old_ = false;
  }
  if (!krylov_ready_) {
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(psi_n, 256)), dim3(256), 0, stream_, d_psi_[0], (int64_t)psi_n);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(psi_n, 256)), dim3(256), 0, stream_, d_psi_[1], (int64_t)psi_n);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG * shape_.tally_copies, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)(NG * shape_.tally_copies));
    psi_cur_ = 0;
    krylov_ready_ = true;
  }
  IterState st;
  std::memset(&st, 0, sizeof(st));
  st.k = 1.0;
  st.k_prev = 1.0;
  st.norm_factor = 1.0;
  st.psi_scale = 1.0;
  st.scale_cap = scale_cap_;
  st.iteration = 30;  // no negative-source clamp
  // the start fluxes' bound for the fixed-point scale: the largest outgoing
  // flux of the previous operator sweep
  st.psi_max_bits = h_state_ ? h_state_->psi_max_bits : 0u;
  HIP_CHECK(hipStreamSynchronize(stream_));
  *h_state_ = st;
  HIP_CHECK(hipMemcpyAsync(d_state_, h_state_, sizeof(IterState), hipMemcpyHostToDevice, stream_));
  FsrArgs fa = fsrArgs();
  fa.leak = nullptr;
  fa.stab_type = -1;
  launchScatter(stream_);
  hipLaunchKernelGGL(update_kernel<0>, dim3(blocks_for(NG, kReduceBlock)), dim3(kReduceBlock), 0,
                     stream_, fa, 0, 0, 0, 1, 0, source_kind, 0, 0, 0.0);
  hipLaunchKernelGGL(tally_scale_kernel, dim3(1), dim3(1), 0, stream_, d_state_, sigma_t_max_, 1, 0);
  launchSweep(d_psi_[psi_cur_], d_psi_[psi_cur_ ^ 1], stream_);
  psi_cur_ ^= 1;
  hipLaunchKernelGGL(post_sweep_kernel, dim3(num_partials_), dim3(kReduceBlock), 0, stream_, fa, 1, 0);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  state_valid_ = false;
}

// CPUSolver::computeFSRSources (src/CPUSolver.cpp:1939-2023) on its own:
// sources of the current flux with the current keff, the negative-source
// clamp for iterations below 30.
void HipSolver::computeFSRSources(int iteration, double* out, int64_t n) {
  requireReady();
  const size_t NG = (size_t)N_ * G_;
  if (!out || n != (int64_t)NG) throw DeviceError(MOC_ERR_INVALID, "source buffer must hold N*G values");
  if (iteration < 0) throw DeviceError(MOC_ERR_INVALID, "iteration must be >= 0");
  HIP_CHECK(hipSetDevice(device_));
  applyFixedSources();
  if (linear_ && !ls_ready_) prepareLinearSource();
  if (phi_is_old_) {
    HIP_CHECK(hipMemcpyAsync(d_phi_, d_phi_old_[old_sel_], NG * sizeof(double),
                             hipMemcpyDeviceToDevice, stream_));
    phi_is_old_ = false;
  }
  HIP_CHECK(hipStreamSynchronize(stream_));
  IterState st;
  std::memset(&st, 0, sizeof(st));
  st.k = keff_;
  st.k_prev = keff_;
  st.norm_factor = 1.0;
  st.psi_scale = 1.0;
  st.scale_cap = scale_cap_;
  st.iteration = iteration;
  *h_state_ = st;
  HIP_CHECK(hipMemcpyAsync(d_state_, h_state_, sizeof(IterState), hipMemcpyHostToDevice, stream_));
  FsrArgs fa = fsrArgs();
  fa.leak = nullptr;
  fa.stab_type = -1;
  launchScatter(stream_);
  hipLaunchKernelGGL(update_kernel<0>, dim3(blocks_for(NG, kReduceBlock)), dim3(kReduceBlock), 0,
                     stream_, fa, 0, 0, 0, 1, 0, kSrcTotal, 0, 0, 0.0);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(out, d_Q_, NG * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  state_valid_ = false;
  krylov_ready_ = false;
}

void HipSolver::runSolve(Mode mode, int max_iters, int res_type, double k_fixed) {
  startSolve(mode, max_iters, res_type, k_fixed);
  while (true) {
    HIP_CHECK(hipEventSynchronize(async_ev_));
    if (pollSolve()) break;
  }
}

void HipSolver::computeEigenvalue(int max_iters, int res_type) {
  runSolve(MODE_EIGEN, max_iters, res_type, 1.0);
}

void HipSolver::computeFlux(int max_iters, bool only_fixed_source) {
  // only_fixed_source = false after an earlier solve: its scalar and angular
  // fluxes are the starting point instead of zero (src/Solver.cpp:1379-1386)
  continue_flux_ = !only_fixed_source && solved_ && state_valid_;
  if (!fixed_on_ && !continue_flux_)
    throw DeviceError(MOC_ERR_STATE, "computeFlux from zero requires fixed sources");
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
  HIP_CHECK(hipMemcpyAsync(&v, currentFlux() + (size_t)fsr * G_ + group - 1, sizeof(double),
                           hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  return v;
}

void HipSolver::getFluxes(double* out, int64_t n) {
  requireReady();
  if (!out || n != N_ * G_) throw DeviceError(MOC_ERR_INVALID, "getFluxes size must be N*G");
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipMemcpyAsync(out, currentFlux(), n * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

void HipSolver::setFluxes(const double* in, int64_t n) {
  requireReady();
  if (!in || n != N_ * G_) throw DeviceError(MOC_ERR_INVALID, "setFluxes size must be N*G");
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipMemcpyAsync(d_phi_, in, n * sizeof(double), hipMemcpyHostToDevice, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  phi_is_old_ = false;
}

void HipSolver::setFSRCentroids(const double* xyz, int64_t n) {
  if (!fsrs_set_) throw DeviceError(MOC_ERR_STATE, "FSRs have not been set");
  if (!xyz || n != 3 * N_) throw DeviceError(MOC_ERR_INVALID, "centroid array must hold 3*N values");
  h_fsr_xyz_.assign(xyz, xyz + n);
}

void HipSolver::dumpFSRFluxes(const std::string& path) {
  requireReady();
  if (h_fsr_xyz_.size() != (size_t)(3 * N_))
    throw DeviceError(MOC_ERR_STATE, "FSR centroids have not been set");
  std::vector<double> flux((size_t)N_ * G_);
  getFluxes(flux.data(), (int64_t)flux.size());
  write_flux_file(path, keff_, G_, N_, h_fsr_xyz_.data(), flux.data());
}

// The file's fluxes reordered to this solver's FSRs by nearest centroid
// (src/Solver.cpp:2135-2318).
std::vector<double> HipSolver::matchedFlux(const std::string& path, double tolerance, double* keff,
                                           double* max_error) {
  if (h_fsr_xyz_.size() != (size_t)(3 * N_))
    throw DeviceError(MOC_ERR_STATE, "FSR centroids have not been set");
  const FluxFile f = read_flux_file(path);
  if (f.num_fsrs != N_)
    throw DeviceError(MOC_ERR_INVALID, "the flux file holds " + std::to_string(f.num_fsrs) +
                                           " FSRs, the geometry " + std::to_string(N_));
  if (f.num_groups != G_)
    throw DeviceError(MOC_ERR_INVALID, "the flux file holds " + std::to_string(f.num_groups) +
                                           " groups, the materials " + std::to_string(G_));
  const std::vector<int64_t> idx =
      match_centroids(f.xyz.data(), f.num_fsrs, h_fsr_xyz_.data(), N_, tolerance, max_error);
  std::vector<double> flux((size_t)N_ * G_);
  for (int64_t r = 0; r < N_; r++)
    std::memcpy(flux.data() + r * G_, f.flux.data() + idx[(size_t)r] * G_, G_ * sizeof(double));
  if (keff) *keff = f.keff;
  return flux;
}

double HipSolver::loadFSRFluxes(const std::string& path, bool assign_k_eff, double tolerance) {
  requireReady();
  double k = 0.0, err = 0.0;
  const std::vector<double> flux = matchedFlux(path, tolerance, &k, &err);
  setFluxes(flux.data(), (int64_t)flux.size());
  if (assign_k_eff) keff_ = k;
  return err;
}

void HipSolver::loadInitialFSRFluxes(const std::string& path) {
  init_flux_file_ = path;
  h_init_flux_.clear();
}

void HipSolver::setResidualByReference(const std::string& path) {
  ref_flux_file_ = path;
  graph_dirty_ = true;
}

// Files named by loadInitialFSRFluxes / setResidualByReference are read when
// a solve starts, as the reference reads them in computeInitialFluxGuess
// (src/Solver.cpp:1710-1736); centroid tolerance 0.01 cm (src/Solver.h:470).
void HipSolver::readFluxFiles(Mode mode) {
  constexpr double kTolerance = 0.01;
  h_init_flux_.clear();
  if (!init_flux_file_.empty() && mode != MODE_FLUX) {
    if (num_fissionable_ == 0)
      throw DeviceError(MOC_ERR_INVALID, "initial fluxes from a file are normalised by the fission "
                                         "rate, which needs fissionable material");
    h_init_flux_ = matchedFlux(init_flux_file_, kTolerance, &init_keff_, nullptr);
  }
  if (!ref_flux_file_.empty()) {
    const std::vector<double> ref = matchedFlux(ref_flux_file_, kTolerance, nullptr, nullptr);
    dev_alloc(&d_ref_flux_, ref.size());
    ref_flux_n_ = ref.size();
    HIP_CHECK(hipMemcpyAsync(d_ref_flux_, ref.data(), ref.size() * sizeof(double),
                             hipMemcpyHostToDevice, stream_));
    HIP_CHECK(hipStreamSynchronize(stream_));
    graph_dirty_ = true;
  }
}

void HipSolver::computeFSRFissionRates(double* out, int64_t n) {
  requireReady();
  if (!out || n != N_) throw DeviceError(MOC_ERR_INVALID, "fission rates size must be N");
  HIP_CHECK(hipSetDevice(device_));
  double* d_out = nullptr;
  dev_alloc(&d_out, N_);
  hipLaunchKernelGGL(fission_rates_kernel, dim3(blocks_for(N_, 256)), dim3(256), 0, stream_,
                     currentFlux(), d_fsr_mat_, d_nsf_, d_fsr_vol_, d_out, N_, G_);
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
  if (linear_) throw DeviceError(MOC_ERR_INVALID, "the single-sweep entry point is flat-source only");
  if (!Q || !psi || !tally) throw DeviceError(MOC_ERR_INVALID, "null argument");
  HIP_CHECK(hipSetDevice(device_));
  const size_t NG = (size_t)N_ * G_;
  const size_t psi_n = psiCount();
  const size_t ref_n = (size_t)2 * T_ * P_ * G_;  // reference layout (t*2+d)*fpt + p*G + g
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
    for (size_t i = 0; i < ref_n; i++) pmax = std::max(pmax, (double)std::fabs(psi[i]));
    st.qmax = (float)(qmax * (1 + 1e-6));
    st.psi_max = (float)(pmax * (1 + 1e-6));
    const double bound = 2.0 * kFourPi * sigma_t_max_ * ((double)st.psi_max + (double)st.qmax);
    st.scale_cap = scale_cap_;
    st.tally_scale = tally_scale_for(bound, scale_cap_);
  }
  HIP_CHECK(hipMemcpyAsync(d_state_, &st, sizeof(st), hipMemcpyHostToDevice, stream_));
  HIP_CHECK(hipMemcpyAsync(d_Q_, Q, NG * sizeof(double), hipMemcpyHostToDevice, stream_));
  const bool lds = shape_.lds_groups > 0;
  hipLaunchKernelGGL(make_qs_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, d_Q_,
                     d_fsr_mat_, d_sigma_t_, lds ? nullptr : d_qs_, lds ? d_qsub_ : nullptr,
                     std::max(1, shape_.lds_groups), N_, G_);
  // reference layout -> start-flux slots [G][W][P] by work position
  const int fpt = P_ * G_;
  std::vector<float> slots(psi_n, 0.f);
  for (int64_t dt = 0; dt < 2 * T_; dt++)
    for (int p = 0; p < P_; p++)
      for (int g = 0; g < G_; g++)
        slots[((size_t)g * psi_w_ + h_wpos_[dt]) * P_ + p] = psi[(size_t)dt * fpt + p * G_ + g];
  HIP_CHECK(hipMemcpyAsync(d_psi_[0], slots.data(), psi_n * sizeof(float), hipMemcpyHostToDevice, stream_));
  // slots that no track writes keep their value (vacuum inflow), as in
  // CPUSolver::_start_flux
  HIP_CHECK(hipMemcpyAsync(d_psi_[1], slots.data(), psi_n * sizeof(float), hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(NG * shape_.tally_copies, 256)), dim3(256), 0, stream_, d_tally_, (int64_t)(NG * shape_.tally_copies));
  if (int64Tally())
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(2 * NG, 256)), dim3(256), 0, stream_,
                       reinterpret_cast<float*>(d_tally_i64_), (int64_t)(2 * NG));
  launchSweep(d_psi_[0], d_psi_[1], stream_);
  std::vector<float> t32;
  std::vector<unsigned long long> t64;
  if (int64Tally()) {
    t64.resize(NG);
    HIP_CHECK(hipMemcpyAsync(t64.data(), d_tally_i64_, NG * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, stream_));
  } else if (shape_.lds_groups > 0) {
    const int gb = shape_.lds_groups, nsub = (G_ + gb - 1) / gb;
    const int64_t nred = (int64_t)nsub * N_ * gb;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(blocks_for(nred, 256)), dim3(256), 0, stream_,
                       (const unsigned long long*)d_lds_partials_, d_tally64_, (int64_t)N_, G_, gb,
                       nsub, shape_.lds_blocks_per_sub, (const IterState*)d_state_);
    HIP_CHECK(hipGetLastError());
    // the fixed-point tally holds sum / V; the raw tally is sum
    HIP_CHECK(hipMemcpyAsync(tally, d_tally64_, NG * sizeof(double), hipMemcpyDeviceToHost, stream_));
  } else if (segStore()) {
    hipLaunchKernelGGL(segment_tally_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, stream_, fsrArgs(),
                       d_tally64_);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(tally, d_tally64_, NG * sizeof(double), hipMemcpyDeviceToHost, stream_));
  } else {
    t32.resize(NG * shape_.tally_copies);
    HIP_CHECK(hipMemcpyAsync(t32.data(), d_tally_, NG * shape_.tally_copies * sizeof(float),
                             hipMemcpyDeviceToHost, stream_));
  }
  HIP_CHECK(hipMemcpyAsync(slots.data(), d_psi_[1], psi_n * sizeof(float), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  if (int64Tally()) {
    // the fixed-point tally holds sum / V
    for (size_t i = 0; i < NG; i++) {
      const double v = h_fsr_vol_[i / G_];
      tally[i] = (double)(long long)t64[i] / st.tally_scale * (v < kFltEpsilon ? 0.0 : v);
    }
    hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for(2 * NG, 256)), dim3(256), 0, stream_,
                       reinterpret_cast<float*>(d_tally_i64_), (int64_t)(2 * NG));
  } else if (segStore()) {
    // the raw tally came back as is
  } else if (shape_.lds_groups == 0) {
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
  for (int64_t dt = 0; dt < 2 * T_; dt++)
    for (int p = 0; p < P_; p++)
      for (int g = 0; g < G_; g++)
        psi[(size_t)dt * fpt + p * G_ + g] = slots[((size_t)g * psi_w_ + h_wpos_[dt]) * P_ + p];
  psi_cur_ = 0;
}

double HipSolver::runIterations(int n, double* sweep_seconds) {
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
  const double tol = threshold_;
  threshold_ = 0.0;
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  double step_ms = 0.0;
  try {
    // the iterations as a solve runs them (graph batches); the graphs of
    // both start-flux parities are captured here, before the timed launches,
    // and kept for the next call (the threshold is part of their key)
    if (use_graphs_) {
      const int p0 = psi_cur_;
      for (int q = 0; q < 2; q++) {
        psi_cur_ = q;
        ensureGraph(MODE_EIGEN, MOC_FISSION_SOURCE);
      }
      psi_cur_ = p0;
    }
    HIP_CHECK(hipEventRecord(e0, stream_));
    launchIterations(MODE_EIGEN, MOC_FISSION_SOURCE, n);
    HIP_CHECK(hipEventRecord(e1, stream_));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    step_ms = ms / n;
    if (sweep_seconds) {
      // the sweep kernel alone: events around it in a few more iterations
      const int m = std::min(n, 8);
      std::vector<hipEvent_t> ev(2 * m);
      for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
      for (int i = 0; i < m; i++)
        launchIteration(MODE_EIGEN, MOC_FISSION_SOURCE, true, 0.0, ev[2 * i], ev[2 * i + 1]);
      HIP_CHECK(hipStreamSynchronize(stream_));
      double sw = 0.0;
      for (int i = 0; i < m; i++) {
        float x = 0.f;
        HIP_CHECK(hipEventElapsedTime(&x, ev[2 * i], ev[2 * i + 1]));
        sw += x;
      }
      for (auto& e : ev) (void)hipEventDestroy(e);
      *sweep_seconds = sw * 1e-3 / m;
    }
  } catch (...) {
    threshold_ = tol;
    graph_dirty_ = true;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  threshold_ = tol;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  keff_ = h_state_->k;
  return step_ms * 1e-3;
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
  hipLaunchKernelGGL(tally_scale_kernel, dim3(1), dim3(1), 0, stream_, d_state_, sigma_t_max_, 0, 0);
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

// Geometry buffers shared by the solvers of one track set (serving): freed
// with the last solver that uses them.
struct HipSolver::GeoHold {
  int device = 0;
  std::vector<void*> bufs;
  ~GeoHold() {
    (void)hipSetDevice(device);
    for (void* p : bufs)
      if (p) {
        {
          std::lock_guard<std::mutex> lk(g_alloc_mu);
          g_alloc_bytes.erase(p);
        }
        (void)hipFree(p);
      }
  }
};

void HipSolver::shareGeometry(HipSolver& src) {
  if (&src == this) return;
  if (!(src.tracks_set_ && src.fsrs_set_ && src.layout_tpw_ > 0))
    throw DeviceError(MOC_ERR_STATE, "the source solver has no complete geometry (tracks, FSRs, materials)");
  if (src.device_ != device_) throw DeviceError(MOC_ERR_INVALID, "solvers on different devices");
  if (src.is3d_) throw DeviceError(MOC_ERR_INVALID, "3D track sets are not shared between solvers");
  if (tracks_set_ || fsrs_set_ || geo_hold_)
    throw DeviceError(MOC_ERR_STATE, "geometry can only be shared into a fresh solver");
  if (!src.geo_hold_) {
    src.geo_hold_ = std::make_shared<GeoHold>();
    src.geo_hold_->device = src.device_;
    src.geo_hold_->bufs = {src.d_seg_len_, src.d_seg_key_, src.d_work_,  src.d_tile_,
                           src.d_trk_,     src.d_trk_azim_, src.d_weights_, src.d_fsr_mat_,
                           src.d_fsr_vol_, src.d_inv_vol_, src.d_out_wp_, src.d_csr_off_,
                           src.d_seg_slot_};
  }
  geo_hold_ = src.geo_hold_;
  d_seg_len_ = src.d_seg_len_;
  d_seg_key_ = src.d_seg_key_;
  d_work_ = src.d_work_;
  d_tile_ = src.d_tile_;
  d_trk_ = src.d_trk_;
  d_trk_azim_ = src.d_trk_azim_;
  d_weights_ = src.d_weights_;
  d_fsr_mat_ = src.d_fsr_mat_;
  d_fsr_vol_ = src.d_fsr_vol_;
  d_inv_vol_ = src.d_inv_vol_;
  d_out_wp_ = src.d_out_wp_;
  d_csr_off_ = src.d_csr_off_;
  d_seg_slot_ = src.d_seg_slot_;
  csr_ready_ = src.csr_ready_;
  seg_stride_ = src.seg_stride_;
  seg_rows_ = src.seg_rows_;
  psi_w_ = src.psi_w_;
  h_wpos_ = src.h_wpos_;
  T_ = src.T_;
  S_ = src.S_;
  P_ = src.P_;
  is3d_ = src.is3d_;
  N_ = src.N_;
  G_ = src.G_;
  num_azim_ = src.num_azim_;
  num_polar_ = src.num_polar_;
  for (int p = 0; p < kMaxPolarHalf; p++) inv_sin_[p] = src.inv_sin_[p];
  max_seg_fsr_ = src.max_seg_fsr_;
  layout_tpw_ = src.layout_tpw_;
  layout_ntiles_ = src.layout_ntiles_;
  layout_rows_ = src.layout_rows_;
  layout_shared_ = src.layout_shared_;
  layout_stack_ = src.layout_stack_;
  h_fsr_mat_ = src.h_fsr_mat_;
  h_fsr_vol_ = src.h_fsr_vol_;
  scale_cap_ = src.scale_cap_;
  shape_ = src.shape_;
  shape_forced_ = true;
  exp_mode_ = src.exp_mode_;
  tracks_set_ = true;
  fsrs_set_ = true;
}

size_t HipSolver::solveBytes() const {
  const size_t NG = (size_t)N_ * G_;
  const size_t psi = (size_t)2 * T_ * P_ * G_ * sizeof(float);
  const int gb = std::max(1, shape_.lds_groups);
  const size_t nsub = (size_t)(G_ + gb - 1) / gb;
  size_t b = 2 * psi + NG * (6 * sizeof(double) + sizeof(float2)) +
             NG * shape_.tally_copies * sizeof(float) + (size_t)2 * T_ * G_ * sizeof(float) +
             (size_t)M_ * G_ * (G_ + 4) * sizeof(double);
  if (stackMode())
    b += h_slot_fsr_.size() * (stackRow() * sizeof(uint64_t) + sizeof(int32_t)) + NG * sizeof(uint64_t) +
         nsub * N_ * gb * sizeof(float);
  else if (shape_.lds_groups > 0)
    b += nsub * shape_.lds_blocks_per_sub * N_ * gb * sizeof(uint64_t) + nsub * N_ * gb * sizeof(float);
  if (segStore())
    b += (size_t)(linear_ ? 3 : 1) * seg_rows_ * seg_stride_ * sizeof(float) + NG * sizeof(double);
  if (scatterPrecisionInUse() != MOC_SCATTER_FP64)
    b += (size_t)N_ * (G_ + 1) * sizeof(float) +
         (size_t)M_ * ((G_ + 15) / 16) * ((G_ + 16) / 16) * 64 * 16;
  return b;
}


void HipSolver::dropGraphs() {
  for (auto& g : graph_exec_)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// The graph of kGraphIters iterations starting at start-flux buffer psi_cur_
// (its sweeps alternate the two buffers from there).  Both parities' graphs
// are kept while mode, residual type and threshold stay the same.
void HipSolver::ensureGraph(int mode, int res_type) {
  const bool recompute = mode != MODE_FLUX;
  if (graph_dirty_ || graph_mode_ != mode || graph_res_type_ != res_type ||
      graph_tol_ != threshold_) {
    dropGraphs();
    graph_dirty_ = false;
    graph_mode_ = mode;
    graph_res_type_ = res_type;
    graph_tol_ = threshold_;
  }
  if (graph_exec_[psi_cur_]) return;
  if (layout_tpw_ != sweepTracksPerWave())
    throw DeviceError(MOC_ERR_STATE, "segment streams are not laid out for the launch shape");
  // kernels that need the LDS opt-in get it before the capture
  const int psi0 = psi_cur_;
  {
    const void* ks[] = {
        (const void*)sweep2d_lds_kernel<MOC_EXP_RATIONAL, 0, 1>, (const void*)sweep2d_lds_kernel<MOC_EXP_RATIONAL, 0, 2>,
        (const void*)sweep2d_lds_kernel<MOC_EXP_RATIONAL, 0, 3>, (const void*)sweep2d_lds_kernel<MOC_EXP_RATIONAL, 1, 1>,
        (const void*)sweep2d_lds_kernel<MOC_EXP_RATIONAL, 1, 2>, (const void*)sweep2d_lds_kernel<MOC_EXP_RATIONAL, 1, 3>,
        (const void*)sweep2d_lds_kernel<MOC_EXP_HW, 0, 1>, (const void*)sweep2d_lds_kernel<MOC_EXP_HW, 0, 2>,
        (const void*)sweep2d_lds_kernel<MOC_EXP_HW, 0, 3>, (const void*)sweep2d_lds_kernel<MOC_EXP_HW, 1, 1>,
        (const void*)sweep2d_lds_kernel<MOC_EXP_HW, 1, 2>, (const void*)sweep2d_lds_kernel<MOC_EXP_HW, 1, 3>,
        (const void*)sweep3d_stack_kernel<MOC_EXP_RATIONAL>, (const void*)sweep3d_stack_kernel<MOC_EXP_HW>};
    for (const void* k : ks) ensure_lds_attribute(k, device_, lds_max_);
  }
  hipGraph_t g = nullptr;
  HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  try {
    for (int i = 0; i < kGraphIters; i++)
      launchIteration((Mode)mode, res_type, recompute, threshold_);
  } catch (...) {
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    psi_cur_ = psi0;
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(stream_, &g));
  hipError_t e = hipGraphInstantiate(&graph_exec_[psi0], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  psi_cur_ = psi0;  // an even number of iterations was recorded
  if (e != hipSuccess) {
    graph_exec_[psi0] = nullptr;
    throw DeviceError(MOC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
}

// Launches `n` iterations: whole graphs while at least kGraphIters remain and
// psi_cur_ is where the graph expects it, single iterations otherwise.
void HipSolver::launchIterations(int mode, int res_type, int n) {
  const bool recompute = mode != MODE_FLUX;
  while (n > 0) {
    if (use_graphs_ && n >= kGraphIters) {
      ensureGraph(mode, res_type);
      HIP_CHECK(hipGraphLaunch(graph_exec_[psi_cur_], stream_));
      n -= kGraphIters;
    } else {
      launchIteration((Mode)mode, res_type, recompute, threshold_);
      n--;
    }
  }
}

void HipSolver::startSolve(int mode, int max_iters, int res_type, double k_fixed) {
  requireReady();
  HIP_CHECK(hipSetDevice(device_));
  if (mode < MODE_EIGEN || mode > MODE_SOURCE) throw DeviceError(MOC_ERR_INVALID, "unknown solve mode");
  if (max_iters <= 0) throw DeviceError(MOC_ERR_INVALID, "max_iters must be > 0");
  if (res_type < MOC_SCALAR_FLUX || res_type > MOC_TOTAL_SOURCE)
    throw DeviceError(MOC_ERR_INVALID, "unknown residual type");
  if (mode != MODE_FLUX && res_type == MOC_FISSION_SOURCE && num_fissionable_ == 0)
    throw DeviceError(MOC_ERR_INVALID,
                      "cannot compute a FISSION_SOURCE residual without fissionable FSRs");
  if (mode == MODE_EIGEN && num_fissionable_ == 0)
    throw DeviceError(MOC_ERR_INVALID, "eigenvalue problem without fissionable material");
  if (mode == MODE_FLUX && !fixed_on_ && !continue_flux_)
    throw DeviceError(MOC_ERR_STATE, "computeFlux from zero requires fixed sources");
  applyFixedSources();
  if (linear_) {
    if (mode != MODE_EIGEN) throw DeviceError(MOC_ERR_INVALID, "the linear source supports computeEigenvalue");
    if (stab_type_ >= 0 || balance_)
      throw DeviceError(MOC_ERR_INVALID,
                        "stabilisation and the neutron-balance keff are not available with the linear source");
    if (!ls_ready_) prepareLinearSource();
  }
  if (!async_ev_) HIP_CHECK(hipEventCreate(&async_ev_));
  if (!async_ev0_) HIP_CHECK(hipEventCreate(&async_ev0_));
  async_ms_ = 0.0;
  async_t0_ = std::chrono::steady_clock::now();
  readFluxFiles((Mode)mode);
  launchInit((Mode)mode, mode == MODE_SOURCE ? k_fixed : 1.0);
  psi_cur0_ = psi_cur_;
  if (use_graphs_ && max_iters >= kGraphIters) ensureGraph(mode, res_type);
  async_mode_ = mode;
  async_res_type_ = res_type;
  async_max_ = max_iters;
  async_done_ = 0;
  async_active_ = true;
  launchAsyncBatch();
}

void HipSolver::launchAsyncBatch() {
  const int n = std::min(kGraphIters, async_max_ - async_done_);
  HIP_CHECK(hipEventRecord(async_ev0_, stream_));
  launchIterations(async_mode_, async_res_type_, n);
  async_done_ += n;
  HIP_CHECK(hipMemcpyAsync(h_state_, d_state_, sizeof(IterState), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipEventRecord(async_ev_, stream_));
}

bool HipSolver::pollSolve() {
  if (!async_active_) return true;
  const hipError_t q = hipEventQuery(async_ev_);
  if (q == hipErrorNotReady) return false;
  HIP_CHECK(q);
  {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, async_ev0_, async_ev_) == hipSuccess) async_ms_ += ms;
  }
  if (!(h_state_->converged || async_done_ >= async_max_)) {
    launchAsyncBatch();
    return false;
  }
  finishSolve(async_max_);
  async_active_ = false;
  return true;
}

void HipSolver::finishSolve(int max_iters) {
  keff_ = h_state_->k;
  old_sel_ = h_state_->old_sel;
  phi_is_old_ = true;
  // a restart counts from 1 (src/Solver.cpp:1554)
  num_iterations_ = (h_state_->converged == 1 ? h_state_->reported_iters : max_iters) +
                    (restart_active_ ? 1 : 0);
  // iterations launched after convergence return at entry, so the start
  // fluxes of a following sweep are where the last live iteration put them
  psi_cur_ = psi_cur0_ ^ (h_state_->iteration & 1);
  solved_ = true;
  total_seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - async_t0_).count();
  // device time of the iteration batches (sweeps plus the per-FSR steps)
  sweep_seconds_ = async_ms_ * 1e-3;
  integrations_ = 2LL * G_ * P_ * S_ * (int64_t)h_state_->iteration;
}

// ---------------------------------------------------------------------------
// Linear source: host set-up
// ---------------------------------------------------------------------------
void HipSolver::setLinearSource(bool on) {
  if (on == linear_) return;
  if (on && is3d_) throw DeviceError(MOC_ERR_INVALID, "the linear source is 2D only");
  if (on && geo_hold_) throw DeviceError(MOC_ERR_INVALID, "linear source on a shared geometry");
  linear_ = on;
  graph_dirty_ = true;
  ls_ready_ = false;
  // the linear source tallies through device-scope atomics
  shape_forced_ = false;
  pickShape();
  layout_tpw_ = 0;
  allocateIterationArrays();
}

namespace {
// 5/5 rational fit of G2 (src/exponentials.h:280-323), for the source constants.
double exp_g2_host(double x) {
  double num = 1.6837426505799449e-4 * x + 1.124183494990467e-5;
  num = num * x + 3.7673183263550827e-3;
  num = num * x + -3.603942303847604e-3;
  num = num * x + -8.335775885589858e-2;
  num *= x;
  double den = 1.0102514456857377e-3 * x + 6.125197988351906e-3;
  den = den * x + 5.367250964303789e-2;
  den = den * x + 2.3794300531408347e-1;
  den = den * x + 7.454048371823628e-1;
  den = den * x + 1.0;
  return num / den;
}
}  // namespace

// FSR centroids (TrackGenerator::generateFSRCentroids, src/TrackGenerator.cpp:
// 1681-1735), the inverse linear expansion matrices and the source constants
// (LinearExpansionGenerator, src/TrackTraversingAlgorithms.cpp:470-831), then
// the stream of segment entry points relative to the centroids.  Host work,
// once per geometry and cross-section set; tracks are split over threads and
// the per-thread sums are added in thread order (deterministic).
void HipSolver::prepareLinearSource() {
  if (h_seg_start_.empty())
    throw DeviceError(MOC_ERR_INVALID,
                      "linear source needs azimuthal angles/weights, polar weights and segment "
                      "start points in the track data");
  const int G = G_, P = P_;
  const int64_t N = N_, T = T_;
  std::vector<double> cosa(h_azim_phi_.size()), sina(h_azim_phi_.size());
  for (size_t i = 0; i < h_azim_phi_.size(); i++) {
    cosa[i] = std::cos(h_azim_phi_[i]);
    sina[i] = std::sin(h_azim_phi_[i]);
  }
  // track azimuthal indices and offsets from the host copies
  std::vector<int64_t> off(T + 1);
  std::vector<int32_t> az(T);
  {
    for (int64_t t = 0; t < T; t++) off[t] = h_trk_[4 * t];
    off[T] = S_;
    HIP_CHECK(hipMemcpy(az.data(), d_trk_azim_, T * sizeof(int32_t), hipMemcpyDeviceToHost));
  }
  const int nth = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  auto parallel = [&](auto&& fn) {
    std::vector<std::thread> th;
    for (int i = 0; i < nth; i++) th.emplace_back(fn, i);
    for (auto& x : th) x.join();
  };
  // centroids: sum of w l (x_start + l/2 dir) / V with x accumulated from the
  // track's first segment
  std::vector<std::vector<double>> cpart(nth, std::vector<double>(2 * N, 0.0));
  parallel([&](int id) {
    for (int64_t t = id; t < T; t += nth) {
      const int a = az[t];
      const double w = h_azim_w_[a];
      if (off[t + 1] == off[t]) continue;
      double x = h_seg_start_[2 * off[t]], y = h_seg_start_[2 * off[t] + 1];
      for (int64_t k = off[t]; k < off[t + 1]; k++) {
        const int32_t r = h_fsr_[k];
        const double l = h_len_[k];
        cpart[id][2 * r] += w * (x + cosa[a] * l / 2.0) * l / h_fsr_vol_[r];
        cpart[id][2 * r + 1] += w * (y + sina[a] * l / 2.0) * l / h_fsr_vol_[r];
        x += cosa[a] * l;
        y += sina[a] * l;
      }
    }
  });
  h_centroid_.assign(2 * N, 0.0);
  for (int id = 0; id < nth; id++)
    for (int64_t i = 0; i < 2 * N; i++) h_centroid_[i] += cpart[id][i];
  cpart.clear();
  // expansion matrices and source constants
  std::vector<double> st((size_t)M_ * G);
  HIP_CHECK(hipMemcpy(st.data(), d_sigma_t_, st.size() * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<std::vector<double>> lpart(nth, std::vector<double>(3 * N, 0.0));
  std::vector<std::vector<double>> spart(nth, std::vector<double>((size_t)N * 3 * G, 0.0));
  parallel([&](int id) {
    std::vector<double>& lem = lpart[id];
    std::vector<double>& sc = spart[id];
    for (int64_t t = id; t < T; t += nth) {
      const int a = az[t];
      const double wgt = h_azim_w_[a], cp = cosa[a], sp = sina[a];
      for (int64_t k = off[t]; k < off[t + 1]; k++) {
        const int32_t r = h_fsr_[k];
        const double l = h_len_[k];
        const double vol = h_fsr_vol_[r];
        const double xc = (h_seg_start_[2 * k] - h_centroid_[2 * r]) + l * 0.5 * cp;
        const double yc = (h_seg_start_[2 * k + 1] - h_centroid_[2 * r + 1]) + l * 0.5 * sp;
        const double vol_impact = wgt * l / vol;
        const double src_constant = vol_impact * l / 2.0;
        const double* s = st.data() + (size_t)h_fsr_mat_[r] * G;
        double* c = sc.data() + (size_t)r * 3 * G;
        for (int g = 0; g < G; g++) {
          double c0 = vol_impact * xc * xc, c1 = vol_impact * yc * yc, c2 = vol_impact * xc * yc;
          const double tau = l * s[g];
          for (int p = 0; p < P; p++) {
            const double sin_t = 1.0 / (double)inv_sin_[p];
            const double g2 =
                l * exp_g2_host(tau / sin_t) * src_constant * 2 * h_polar_w_[p] * sin_t;
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
  });
  std::vector<double> lem(3 * N, 0.0), src_c((size_t)N * 3 * G, 0.0);
  for (int id = 0; id < nth; id++) {
    for (int64_t i = 0; i < 3 * N; i++) lem[i] += lpart[id][i];
    for (size_t i = 0; i < src_c.size(); i++) src_c[i] += spart[id][i];
  }
  lpart.clear();
  spart.clear();
  std::vector<double> ilem(3 * N, 0.0);
  for (int64_t r = 0; r < N; r++) {
    const double det = lem[3 * r] * lem[3 * r + 1] - lem[3 * r + 2] * lem[3 * r + 2];
    if (std::fabs(det) >= 1e-10) {  // MIN_DET (src/constants.h:70)
      ilem[3 * r] = lem[3 * r + 1] / det;
      ilem[3 * r + 1] = lem[3 * r] / det;
      ilem[3 * r + 2] = -lem[3 * r + 2] / det;
    }
  }
  dev_alloc(&d_ilem_, 3 * N);
  dev_alloc(&d_src_c_, src_c.size());
  HIP_CHECK(hipMemcpy(d_ilem_, ilem.data(), ilem.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_src_c_, src_c.data(), src_c.size() * sizeof(double), hipMemcpyHostToDevice));
  std::vector<float> fc(cosa.begin(), cosa.end()), fs(sina.begin(), sina.end());
  dev_alloc(&d_azim_cos_, fc.size());
  dev_alloc(&d_azim_sin_, fs.size());
  HIP_CHECK(hipMemcpy(d_azim_cos_, fc.data(), fc.size() * sizeof(float), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(d_azim_sin_, fs.data(), fs.size() * sizeof(float), hipMemcpyHostToDevice));
  buildPositionStream();
  ls_ready_ = true;
}

// Entry point of each segment in traversal order, relative to its FSR
// centroid, in the tile-interleaved layout of the length/key streams
// (forward: the segment start; backward: its end, as the reference moves
// the starting position to the end after the forward pass,
// src/CPULSSolver.cpp:736-738).
void HipSolver::buildPositionStream() {
  const int tpw = layout_tpw_;
  const int64_t ntiles = layout_ntiles_;
  const size_t total = (size_t)(layout_rows_ + kTailRows) * tpw;
  std::vector<float2> pos(total, make_float2(0.f, 0.f));
  std::vector<double> cosa(h_azim_phi_.size()), sina(h_azim_phi_.size());
  for (size_t i = 0; i < h_azim_phi_.size(); i++) {
    cosa[i] = std::cos(h_azim_phi_[i]);
    sina[i] = std::sin(h_azim_phi_[i]);
  }
  std::vector<int32_t> az(T_);
  HIP_CHECK(hipMemcpy(az.data(), d_trk_azim_, T_ * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int64_t tl = 0; tl < ntiles; tl++) {
    const int64_t r0 = h_tile_[2 * tl];
    for (int sl = 0; sl < tpw; sl++) {
      const int32_t dt = h_work_[tl * tpw + sl];
      if (dt < 0) continue;
      const int32_t t = dt >> 1, dir = dt & 1;
      const int32_t b = h_trk_[4 * t], n = h_trk_[4 * t + 1];
      for (int32_t i = 0; i < n; i++) {
        const int32_t k = dir ? b + n - 1 - i : b + i;
        const int32_t r = h_fsr_[k];
        double x = h_seg_start_[2 * k] - h_centroid_[2 * r];
        double y = h_seg_start_[2 * k + 1] - h_centroid_[2 * r + 1];
        if (dir) {
          x += cosa[az[t]] * h_len_[k];
          y += sina[az[t]] * h_len_[k];
        }
        pos[(size_t)(r0 + i) * tpw + sl] = make_float2((float)x, (float)y);
      }
    }
  }
  dev_alloc(&d_seg_pos_, total);
  HIP_CHECK(hipMemcpy(d_seg_pos_, pos.data(), total * sizeof(float2), hipMemcpyHostToDevice));
}

void HipSolver::getFluxMoments(double* out, int64_t n) {
  requireReady();
  if (!linear_) throw DeviceError(MOC_ERR_STATE, "the linear source is not enabled");
  if (!out || n != 2 * N_ * G_) throw DeviceError(MOC_ERR_INVALID, "moments size must be 2*N*G");
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipMemcpyAsync(out, d_phi_xy_n_, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIP_CHECK(hipStreamSynchronize(stream_));
}

void HipSolver::getCentroids(double* out, int64_t n) {
  requireReady();
  if (!out || n != 2 * N_) throw DeviceError(MOC_ERR_INVALID, "centroid size must be 2*N");
  if (linear_ && !ls_ready_) prepareLinearSource();
  if (h_centroid_.size() != (size_t)(2 * N_)) throw DeviceError(MOC_ERR_STATE, "no centroids yet");
  std::memcpy(out, h_centroid_.data(), 2 * N_ * sizeof(double));
}

// ===========================================================================
// SolveBatch
// ===========================================================================
SolveBatch::SolveBatch(int device, int capacity, int res_type)
    : device_(device), capacity_(capacity), res_type_(res_type) {
  if (capacity <= 0 || capacity > 65535) throw DeviceError(MOC_ERR_INVALID, "batch capacity out of range");
  if (res_type < MOC_SCALAR_FLUX || res_type > MOC_TOTAL_SOURCE)
    throw DeviceError(MOC_ERR_INVALID, "unknown residual type");
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&capture_stream_, hipStreamNonBlocking));
  dev_alloc(&d_table_, (size_t)capacity);
  dev_alloc(&d_snap_, (size_t)2 * capacity);
  HIP_CHECK(hipHostMalloc((void**)&h_snap_, (size_t)2 * capacity * sizeof(IterState), hipHostMallocDefault));
  std::memset(h_snap_, 0, (size_t)2 * capacity * sizeof(IterState));
}

SolveBatch::~SolveBatch() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
  for (hipGraphExec_t g : retired_) (void)hipGraphExecDestroy(g);
  for (auto& l : inflight_) (void)hipEventDestroy(l.ev);
  for (hipEvent_t e : free_events_) (void)hipEventDestroy(e);
  solvers_.clear();  // their stream is stream_: destroyed below
  dev_free(d_table_);
  dev_free(d_snap_);
  if (h_snap_) (void)hipHostFree(h_snap_);
  if (capture_stream_) (void)hipStreamDestroy(capture_stream_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

std::string SolveBatch::batchKey(const HipSolver& g) {
  g.requireReady();
  if (g.is3d_ || g.linear_) throw DeviceError(MOC_ERR_INVALID, "batched solves are 2D flat-source solves");
  if (g.layout_tpw_ != g.sweepTracksPerWave())
    throw DeviceError(MOC_ERR_STATE, "segment streams are not laid out for the launch shape");
  const LaunchShape& sh = g.shape_;
  std::string k = "P" + std::to_string(g.P_) + ".e" + std::to_string(g.exp_mode_);
  if (sh.lds_groups > 0)
    k += ".lds.q" + std::to_string(sh.lds_q) + ".b" + std::to_string(sh.lds_block);
  else
    k += ".glob.g" + std::to_string(sh.groups_per_lane) + ".b" + std::to_string(sh.block) + ".s" +
         std::to_string(g.segStore() ? 1 : 0);
  return k;
}

bool SolveBatch::shareable(const HipSolver& g) {
  const LaunchShape& sh = g.shape_;
  if (sh.lds_groups <= 0) return false;
  const int nsub = (g.G_ + sh.lds_groups - 1) / sh.lds_groups;
  return nsub * sh.lds_blocks_per_sub <= 64 && g.ldsBytes(sh.lds_groups, sh.lds_q) <= ((size_t)64 << 10);
}

bool SolveBatch::hasGeometry(const HipSolver* g) const {
  return std::find(geos_.begin(), geos_.end(), g) != geos_.end();
}

void SolveBatch::addGeometry(HipSolver& g) {
  if (hasGeometry(&g)) return;
  const std::string k = batchKey(g);
  if (!geos_.empty() && k != key_)
    throw DeviceError(MOC_ERR_INVALID, "the geometry runs other kernels than the batch's");
  if (g.device() != device_) throw DeviceError(MOC_ERR_INVALID, "geometry on another device");
  key_ = k;
  geos_.push_back(&g);
  // launch sizes: the largest over the members
  Dims d;
  for (const HipSolver* m : geos_) {
    const LaunchShape& sh = m->shape_;
    if (sh.lds_groups > 0) {
      const int nsub = (m->G_ + sh.lds_groups - 1) / sh.lds_groups;
      d.sweep_blocks = std::max(d.sweep_blocks, nsub * sh.lds_blocks_per_sub);
      d.sweep_threads = sh.lds_block;
      d.lds_bytes = std::max(d.lds_bytes, m->ldsBytes(sh.lds_groups, sh.lds_q));
      if (d.lds_bytes > m->lds_max_) throw DeviceError(MOC_ERR_INVALID, "LDS slice exceeds the device limit");
    } else {
      const int wpb = sh.block / 64;
      d.sweep_blocks = std::max(d.sweep_blocks, (m->layout_ntiles_ + wpb - 1) / wpb);
      d.sweep_threads = sh.block;
    }
    d.post_blocks = std::max(d.post_blocks, m->num_partials_);
    d.upd_blocks = std::max(d.upd_blocks, (int)blocks_for(m->N_ * m->G_, kReduceBlock));
    const int path = m->scatterPrecisionInUse();
    if (path == MOC_SCATTER_MFMA_F16)
      d.mfma_blocks = std::max(d.mfma_blocks, (int)blocks_for(m->scat_ntiles_, 4));
    else if (path == MOC_SCATTER_FP32)
      d.fp32_blocks = std::max(d.fp32_blocks, (int)blocks_for(m->N_ * (m->G_ + 1), 256));
    if (res_type_ != MOC_SCALAR_FLUX && m->G_ >= kFsrResidualGroups)
      d.fsr_blocks = std::max(d.fsr_blocks, (int)blocks_for(m->N_, kReduceBlock));
  }
  const bool grew = d.sweep_blocks != dims_.sweep_blocks || d.lds_bytes != dims_.lds_bytes ||
                    d.post_blocks != dims_.post_blocks || d.upd_blocks != dims_.upd_blocks ||
                    d.mfma_blocks != dims_.mfma_blocks || d.fp32_blocks != dims_.fp32_blocks ||
                    d.fsr_blocks != dims_.fsr_blocks || d.sweep_threads != dims_.sweep_threads;
  dims_ = d;
  if (grew) {
    // graphs of the old sizes may still be in flight: kept until the batch
    // is idle or destroyed
    for (auto& gr : graphs_) retired_.push_back(gr.second);
    graphs_.clear();
    // the graphs of every row count are captured now, so a change of the
    // number of live solves never waits for a capture
    for (int n = 1; n <= capacity_; n++) (void)graphFor(n);
  }
}

size_t SolveBatch::bytesHeld() const {
  size_t b = 0;
  for (size_t i = 0; i < solvers_.size(); i++)
    if (solvers_[i]) b += solver_geo_[i]->solveBytes();
  return b;
}

bool SolveBatch::hasIdleSolver(const HipSolver* g) const {
  for (int si : idle_)
    if (solver_geo_[si] == g) return true;
  return false;
}

int SolveBatch::newSolver(HipSolver* g) {
  auto s = std::make_unique<HipSolver>(device_);
  s->shareGeometry(*g);
  s->adoptStream(stream_);
  for (size_t i = 0; i < solvers_.size(); i++)
    if (!solvers_[i]) {
      solvers_[i] = std::move(s);
      solver_geo_[i] = g;
      return (int)i;
    }
  solvers_.push_back(std::move(s));
  solver_geo_.push_back(g);
  slots_.push_back(std::make_unique<BatchSlot>());
  return (int)solvers_.size() - 1;
}

// Row arguments of solver s: both start-flux parities of the sweep, the LDS
// tally, the scattering contraction and the per-FSR steps, as launchIteration
// passes them for an eigenvalue iteration, with the row's own block counts.
void SolveBatch::fillRow(BatchSlot& b, HipSolver& s, double tol) const {
  std::memset(&b, 0, sizeof(b));
  s.sweepArgs(b.sw[0], s.d_psi_[0], s.d_psi_[1], nullptr, nullptr);
  s.sweepArgs(b.sw[1], s.d_psi_[1], s.d_psi_[0], nullptr, nullptr);
  if (s.shape_.lds_groups > 0) s.ldsParams(b.lds, nullptr);
  b.scat_path = s.scatArgs(b.scat) ? s.scatterPrecisionInUse() : MOC_SCATTER_FP64;
  b.fsr = s.fsrArgs();
  b.nparts_post = s.num_partials_;
  b.nparts_upd = (int)blocks_for(s.N_ * s.G_, kReduceBlock);
  b.nparts_fsr = (res_type_ != MOC_SCALAR_FLUX && s.G_ >= kFsrResidualGroups)
                     ? (int)blocks_for(s.N_, kReduceBlock)
                     : 0;
  b.num_fsrs = (double)s.N_;
  b.norm = res_type_ == MOC_FISSION_SOURCE ? (double)s.num_fissionable_ : (double)s.N_;
  b.tol = tol;
  b.sigma_t_max = s.sigma_t_max_;
}

void SolveBatch::setRow(int row, const BatchSlot& b) {
  hipLaunchKernelGGL(set_batch_slot_kernel, dim3(1), dim3(64), 0, stream_, d_table_, row, b);
  HIP_CHECK(hipGetLastError());
}

void SolveBatch::start(HipSolver* g, void* tag, int M, int G, const double* sigma_t,
                       const double* sigma_s, const double* nu_sigma_f, const double* chi,
                       int max_iters, double tol, bool want_flux) {
  if (!hasGeometry(g)) throw DeviceError(MOC_ERR_INVALID, "geometry not in the batch");
  if ((int)live_.size() >= capacity_) throw DeviceError(MOC_ERR_STATE, "no free row in the batch");
  if (max_iters <= 0 || !(tol > 0.0)) throw DeviceError(MOC_ERR_INVALID, "invalid solve options");
  HIP_CHECK(hipSetDevice(device_));
  int si = -1;
  for (size_t i = 0; i < idle_.size(); i++)
    if (solver_geo_[idle_[i]] == g) {
      si = idle_[i];
      idle_.erase(idle_.begin() + (long)i);
      break;
    }
  if (si < 0) si = newSolver(g);
  HipSolver& s = *solvers_[si];
  try {
    // the solver's previous solve (if any) has stopped: launches still in
    // flight return at entry for its row, so its arrays can be re-armed
    if (s.materials_set_ && s.M_ == M && s.G_ == G)
      s.rearmMaterials(sigma_t, sigma_s, nu_sigma_f, chi);
    else
      s.setMaterials(M, G, sigma_t, sigma_s, nu_sigma_f, chi);
    s.setConvergenceThreshold(tol);
    if (s.num_fissionable_ == 0)
      throw DeviceError(MOC_ERR_INVALID, "eigenvalue problem without fissionable material");
    s.device_max_iters_ = max_iters;
    s.launchInit(HipSolver::MODE_EIGEN, 1.0);
    s.device_max_iters_ = 0;
    fillRow(*slots_[si], s, tol);
  } catch (...) {
    s.device_max_iters_ = 0;
    idle_.push_back(si);
    throw;
  }
  setRow((int)live_.size(), *slots_[si]);
  live_.push_back(Row{tag, next_serial_++, si, max_iters, want_flux});
}

// One eigenvalue iteration of rows 0..n-1 (launchIteration's sequence),
// sized for the largest member geometry.
void SolveBatch::launchStep(hipStream_t st, int n, int parity) {
  const HipSolver& g = *geos_.front();  // kernels and template arguments are the members' own
  const LaunchShape& sh = g.shape_;
  // one geometry's LDS sweep: n rows share the blocks one solve would use
  // (about two per CU), a multiple of the XCD count per subset, at least 8
  int bps = 0;
  dim3 sgrid((unsigned)dims_.sweep_blocks, (unsigned)n);
  if (sh.lds_groups > 0 && geos_.size() == 1 && n > 1) {
    // n rows of one geometry read the same segment stream: each row gets
    // about a 1/n share of one solve's blocks per subset, rounded up to a
    // multiple of the XCD count, so the blocks of one track slot (every
    // subset of every row) run on one XCD and share its L2 for the stream
    // (measured on five 17x17 solves: 16 blocks per row and subset 55.7 ms,
    // 14 without that placement 58.8 ms)
    const int nsub = (g.G_ + sh.lds_groups - 1) / sh.lds_groups;
    bps = std::max(kNumXcd, (sh.lds_blocks_per_sub / n + kNumXcd - 1) / kNumXcd * kNumXcd);
    bps = std::min(bps, sh.lds_blocks_per_sub);
    sgrid.x = (unsigned)(nsub * bps);
  }
  const dim3 sblock((unsigned)dims_.sweep_threads);
  if (sh.lds_groups > 0) {
    auto launch = [&](auto kernel) {
      ensure_lds_attribute((const void*)kernel, device_, g.lds_max_);
      hipLaunchKernelGGL(kernel, sgrid, sblock, dims_.lds_bytes, st, d_table_, parity, bps);
    };
    auto by_np = [&](auto e, auto q) {
      constexpr int E = decltype(e)::value, Q = decltype(q)::value;
      switch (g.P_) {
        case 1: launch(sweep2d_lds_batch_kernel<E, Q, 1>); break;
        case 2: launch(sweep2d_lds_batch_kernel<E, Q, 2>); break;
        default: launch(sweep2d_lds_batch_kernel<E, Q, 3>); break;
      }
    };
    using EH = std::integral_constant<int, MOC_EXP_HW>;
    using ER = std::integral_constant<int, MOC_EXP_RATIONAL>;
    using Q1 = std::integral_constant<int, 1>;
    using Q0 = std::integral_constant<int, 0>;
    if (g.exp_mode_ == MOC_EXP_HW)
      sh.lds_q ? by_np(EH{}, Q1{}) : by_np(EH{}, Q0{});
    else
      sh.lds_q ? by_np(ER{}, Q1{}) : by_np(ER{}, Q0{});
  } else {
    auto go = [&](auto gpl) {
      constexpr int GPL = decltype(gpl)::value;
      auto by_np = [&](auto k1, auto k2, auto k3) {
        switch (g.P_) {
          case 1: hipLaunchKernelGGL(k1, sgrid, sblock, 0, st, d_table_, parity); break;
          case 2: hipLaunchKernelGGL(k2, sgrid, sblock, 0, st, d_table_, parity); break;
          default: hipLaunchKernelGGL(k3, sgrid, sblock, 0, st, d_table_, parity); break;
        }
      };
      auto by_exp = [&](auto stv) {
        constexpr int ST = decltype(stv)::value;
        if (g.exp_mode_ == MOC_EXP_HW)
          by_np(sweep2d_batch_kernel<GPL, MOC_EXP_HW, 1, ST>, sweep2d_batch_kernel<GPL, MOC_EXP_HW, 2, ST>,
                sweep2d_batch_kernel<GPL, MOC_EXP_HW, 3, ST>);
        else
          by_np(sweep2d_batch_kernel<GPL, MOC_EXP_RATIONAL, 1, ST>,
                sweep2d_batch_kernel<GPL, MOC_EXP_RATIONAL, 2, ST>,
                sweep2d_batch_kernel<GPL, MOC_EXP_RATIONAL, 3, ST>);
      };
      if (g.segStore())
        by_exp(std::integral_constant<int, 1>{});
      else
        by_exp(std::integral_constant<int, 0>{});
    };
    switch (sh.groups_per_lane) {
      case 1: go(std::integral_constant<int, 1>{}); break;
      case 2: go(std::integral_constant<int, 2>{}); break;
      case 4: go(std::integral_constant<int, 4>{}); break;
      case 9: go(std::integral_constant<int, 9>{}); break;
      case 16: go(std::integral_constant<int, 16>{}); break;
      default: throw DeviceError(MOC_ERR_INVALID, "unsupported groups per lane");
    }
  }
  const unsigned ny = (unsigned)n;
  hipLaunchKernelGGL(post_sweep_batch_kernel, dim3((unsigned)dims_.post_blocks, ny), dim3(kReduceBlock), 0,
                     st, d_table_, bps);
  // rows with few partials take the keff step inside update_kernel
  if (dims_.post_blocks > kFuseKeffMaxParts)
    hipLaunchKernelGGL(keff_batch_kernel, dim3(1, ny), dim3(kReduceBlock), 0, st, d_table_);
  if (dims_.mfma_blocks > 0)
    hipLaunchKernelGGL(scatter_mfma_batch_kernel, dim3((unsigned)dims_.mfma_blocks, ny), dim3(256), 0, st,
                       d_table_);
  if (dims_.fp32_blocks > 0)
    hipLaunchKernelGGL(scatter_fp32_batch_kernel, dim3((unsigned)dims_.fp32_blocks, ny), dim3(256), 0, st,
                       d_table_);
  bool fuse_any = false;  // some row takes the keff step in update_kernel
  for (const HipSolver* m : geos_) fuse_any = fuse_any || m->num_partials_ <= kFuseKeffMaxParts;
  auto* const upd = fuse_any ? update_batch_kernel<1> : update_batch_kernel<0>;
  hipLaunchKernelGGL(upd, dim3((unsigned)dims_.upd_blocks, ny), dim3(kReduceBlock), 0, st, d_table_, res_type_);
  if (dims_.fsr_blocks > 0)
    hipLaunchKernelGGL(fsr_residual_batch_kernel, dim3((unsigned)dims_.fsr_blocks, ny), dim3(kReduceBlock),
                       0, st, d_table_, res_type_);
  hipLaunchKernelGGL(residual_batch_kernel, dim3(1, ny), dim3(kReduceBlock), 0, st, d_table_);
  HIP_CHECK(hipGetLastError());
}

// The graph of kGraphIters iterations of n rows (an even count, so every
// row's start fluxes are back in psi[0] after it).
hipGraphExec_t SolveBatch::graphFor(int n) {
  auto it = graphs_.find(n);
  if (it != graphs_.end()) return it->second;
  {
    // the LDS opt-in is set before the capture
    const void* ks[] = {
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_RATIONAL, 0, 1>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_RATIONAL, 0, 2>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_RATIONAL, 0, 3>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_RATIONAL, 1, 1>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_RATIONAL, 1, 2>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_RATIONAL, 1, 3>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_HW, 0, 1>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_HW, 0, 2>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_HW, 0, 3>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_HW, 1, 1>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_HW, 1, 2>,
        (const void*)sweep2d_lds_batch_kernel<MOC_EXP_HW, 1, 3>};
    for (const void* k : ks) ensure_lds_attribute(k, device_, geos_.front()->lds_max_);
  }
  hipGraph_t g = nullptr;
  HIP_CHECK(hipStreamBeginCapture(capture_stream_, hipStreamCaptureModeThreadLocal));
  try {
    for (int i = 0; i < kGraphIters; i++) launchStep(capture_stream_, n, i & 1);
  } catch (...) {
    (void)hipStreamEndCapture(capture_stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(capture_stream_, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) throw DeviceError(MOC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  graphs_[n] = ex;
  return ex;
}

void SolveBatch::launchGraph() {
  const int n = (int)live_.size();
  hipGraphExec_t ex = graphFor(n);
  HIP_CHECK(hipGraphLaunch(ex, stream_));
  const int buf = next_buf_;
  next_buf_ ^= 1;
  IterState* d = d_snap_ + (size_t)buf * capacity_;
  hipLaunchKernelGGL(gather_states_kernel, dim3(blocks_for(n, 64)), dim3(64), 0, stream_,
                     (const BatchSlot*)d_table_, n, d);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(h_snap_ + (size_t)buf * capacity_, d, (size_t)n * sizeof(IterState),
                           hipMemcpyDeviceToHost, stream_));
  hipEvent_t ev;
  if (!free_events_.empty()) {
    ev = free_events_.back();
    free_events_.pop_back();
  } else {
    HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  HIP_CHECK(hipEventRecord(ev, stream_));
  Launch l{ev, buf, {}};
  l.rows.reserve(n);
  for (const Row& r : live_) l.rows.emplace_back(r.serial, r.si);
  inflight_.push_back(std::move(l));
}

// Result of row idx, whose solve stopped with state st; the last row moves
// into its place.
void SolveBatch::finish(size_t idx, const IterState& st, std::vector<BatchResult>& done) {
  const Row r = live_[idx];
  HipSolver& s = *solvers_[r.si];
  BatchResult res;
  res.tag = r.tag;
  res.keff = st.k;
  res.num_iterations = st.converged == 1 ? st.reported_iters : r.max_iters;
  if (r.want_flux) {
    // the normalised flux of the last live iteration; no launch writes this
    // solver's arrays any more (its rows return at entry)
    res.flux.resize((size_t)s.N_ * s.G_);
    HIP_CHECK(hipMemcpy(res.flux.data(), s.d_phi_old_[st.old_sel], res.flux.size() * sizeof(double),
                        hipMemcpyDeviceToHost));
  }
  s.keff_ = st.k;
  s.old_sel_ = st.old_sel;
  s.phi_is_old_ = true;
  s.num_iterations_ = res.num_iterations;
  const size_t last = live_.size() - 1;
  if (idx != last) {
    live_[idx] = live_[last];
    setRow((int)idx, *slots_[live_[idx].si]);
  }
  live_.pop_back();
  idle_.push_back(r.si);
  done.push_back(std::move(res));
}

void SolveBatch::progress(std::vector<BatchResult>& done) {
  HIP_CHECK(hipSetDevice(device_));
  if (!retired_.empty() && inflight_.empty()) {
    for (hipGraphExec_t g : retired_) (void)hipGraphExecDestroy(g);
    retired_.clear();
  }
  while (!inflight_.empty()) {
    Launch& l = inflight_.front();
    const hipError_t q = hipEventQuery(l.ev);
    if (q == hipErrorNotReady) break;
    HIP_CHECK(q);
    const IterState* snap = h_snap_ + (size_t)l.buf * capacity_;
    for (size_t j = 0; j < l.rows.size(); j++) {
      if (!snap[j].converged) continue;
      // still live (a finished row's solver may hold a newer solve by now)
      for (size_t i = 0; i < live_.size(); i++)
        if (live_[i].serial == l.rows[j].first) {
          finish(i, snap[j], done);
          break;
        }
    }
    free_events_.push_back(l.ev);
    inflight_.erase(inflight_.begin());
  }
  while (inflight_.size() < 2 && !live_.empty()) launchGraph();
}

void SolveBatch::drain() {
  if (!live_.empty() || inflight_.empty()) return;
  HIP_CHECK(hipSetDevice(device_));
  HIP_CHECK(hipStreamSynchronize(stream_));
  std::vector<BatchResult> none;
  progress(none);  // pops the finished launches; no row is live, none is launched
}

size_t SolveBatch::releaseIdle() {
  HIP_CHECK(hipSetDevice(device_));
  size_t freed = 0;
  std::vector<int> keep;
  for (int si : idle_) {
    // a launch in flight may still read the row's state
    bool referenced = false;
    for (const Launch& l : inflight_)
      for (const auto& row : l.rows) referenced |= row.second == si;
    if (referenced) {
      keep.push_back(si);
      continue;
    }
    freed += solver_geo_[si]->solveBytes();
    solvers_[si].reset();
  }
  idle_.swap(keep);
  return freed;
}

size_t HipSolver::noAllocScratchBytes() const {
  const size_t NG = (size_t)N_ * G_;
  const bool i64 = shape_.lds_groups > 0;  // LDS and z-stack sweeps: int64 tally, qsub table
  const int gb = std::max(1, shape_.lds_groups);
  const size_t nsub = (size_t)(G_ + gb - 1) / gb;
  size_t b = align256(sizeof(IterState)) + 256;
  b += align256(i64 ? NG * sizeof(unsigned long long) : NG * shape_.tally_copies * sizeof(float));
  b += align256(i64 ? nsub * N_ * gb * sizeof(float) : NG * sizeof(float2));
  b += 2 * align256(psiCount() * sizeof(float));
  return b;
}

// The sweep of transportSweepOnce on device buffers, with every buffer of
// its own in the static scratch array: no allocation, no host
// synchronisation, no memset (fill kernels only), so it can be captured into
// a HIP graph.  The fixed-point scale comes from the same bounds, computed
// with the same arithmetic on the device, so the results are the bits of
// transportSweepOnce (the LDS sweep always flushes its slice into the int64
// tally here; both flushes are exact integer sums).
void HipSolver::transportSweepNoAlloc(const double* Q, const float* psi_in, float* psi_out,
                                      double* tally, hipStream_t s) {
  requireReady();
  if (linear_) throw DeviceError(MOC_ERR_INVALID, "the single-sweep entry point is flat-source only");
  if (!Q || !psi_in || !psi_out || !tally) throw DeviceError(MOC_ERR_INVALID, "null argument");
  const size_t need = noAllocScratchBytes();
  if (need > (size_t)MOC_NOALLOC_SCRATCH_BYTES)
    throw DeviceError(MOC_ERR_INVALID,
                      "problem too large for the allocation-free sweep: needs " + std::to_string(need) +
                          " bytes of scratch, the maximum is " +
                          std::to_string((size_t)MOC_NOALLOC_SCRATCH_BYTES));
  if (layout_tpw_ != sweepTracksPerWave())
    throw DeviceError(MOC_ERR_STATE, "segment streams are not laid out for the launch shape");
  HIP_CHECK(hipSetDevice(device_));
  unsigned char* base = noalloc_scratch(device_);
  if (!s) s = stream_;
  const size_t NG = (size_t)N_ * G_;
  const bool i64 = shape_.lds_groups > 0;
  const int gb = std::max(1, shape_.lds_groups);
  const size_t nsub = (size_t)(G_ + gb - 1) / gb;
  const size_t psi_n = psiCount();
  size_t off = 0;
  IterState* st = reinterpret_cast<IterState*>(base + off);
  off += align256(sizeof(IterState));
  unsigned long long* bounds = reinterpret_cast<unsigned long long*>(base + off);
  off += 256;
  unsigned char* tally_mem = base + off;
  const size_t tally_bytes = i64 ? NG * sizeof(unsigned long long) : NG * shape_.tally_copies * sizeof(float);
  off += align256(tally_bytes);
  unsigned char* table_mem = base + off;
  off += align256(i64 ? nsub * N_ * gb * sizeof(float) : NG * sizeof(float2));
  float* slot0 = reinterpret_cast<float*>(base + off);
  off += align256(psi_n * sizeof(float));
  float* slot1 = reinterpret_cast<float*>(base + off);

  SweepBuffers B;
  B.qs = i64 ? nullptr : reinterpret_cast<const float2*>(table_mem);
  B.qsub = i64 ? reinterpret_cast<float*>(table_mem) : nullptr;
  B.tally = i64 ? nullptr : reinterpret_cast<float*>(tally_mem);
  B.tally_i64 = i64 ? reinterpret_cast<unsigned long long*>(tally_mem) : nullptr;
  B.state = st;

  const int64_t ref_n = (int64_t)2 * T_ * P_ * G_;
  hipLaunchKernelGGL(noalloc_begin_kernel, dim3(1), dim3(1), 0, s, st, bounds, scale_cap_);
  const unsigned nb = (unsigned)std::min<int64_t>(
      std::max<int64_t>(1, (std::max<int64_t>((int64_t)NG, ref_n) + 255) / 256), 2048);
  hipLaunchKernelGGL(noalloc_bounds_kernel, dim3(nb), dim3(256), 0, s, Q, d_fsr_mat_, d_sigma_t_,
                     N_, G_, psi_in, ref_n, bounds);
  hipLaunchKernelGGL(noalloc_scale_kernel, dim3(1), dim3(1), 0, s, st, (const unsigned long long*)bounds,
                     sigma_t_max_);
  hipLaunchKernelGGL(make_qs_kernel, dim3(blocks_for(NG, 256)), dim3(256), 0, s, Q, d_fsr_mat_,
                     d_sigma_t_, i64 ? nullptr : reinterpret_cast<float2*>(table_mem), B.qsub, gb, N_, G_);
  // the tally starts at zero (what the last call left there is cleared)
  hipLaunchKernelGGL(zero_f32_kernel, dim3(blocks_for((int64_t)(tally_bytes / 4), 256)), dim3(256), 0, s,
                     reinterpret_cast<float*>(tally_mem), (int64_t)(tally_bytes / 4));
  // both slot buffers get the inputs: slots no track writes keep their value
  // (vacuum inflow), as CPUSolver::_start_flux does
  hipLaunchKernelGGL(psi_to_slots_kernel, dim3(blocks_for((int64_t)psi_n, 256)), dim3(256), 0, s, psi_in,
                     d_work_, psi_w_, P_, G_, slot0, slot1);
  HIP_CHECK(hipGetLastError());
  launchSweep(slot0, slot1, s, nullptr, nullptr, &B);
  hipLaunchKernelGGL(noalloc_tally_kernel, dim3(blocks_for((int64_t)NG, 256)), dim3(256), 0, s,
                     B.tally_i64, B.tally, shape_.tally_copies, d_fsr_vol_, (const IterState*)st, N_, G_,
                     tally);
  hipLaunchKernelGGL(slots_to_psi_kernel, dim3(blocks_for((int64_t)psi_n, 256)), dim3(256), 0, s, slot1,
                     d_work_, psi_w_, P_, G_, psi_out);
  HIP_CHECK(hipGetLastError());
}

double deviceCopyCalibration(int device, int64_t bytes, int repeats) {
  if (bytes <= 0 || bytes % 16 != 0 || repeats <= 0)
    throw DeviceError(MOC_ERR_INVALID, "copy size must be a positive multiple of 16 bytes");
  HIP_CHECK(hipSetDevice(device));
  int4 *a = nullptr, *b = nullptr;
  const int64_t n = bytes / 16;
  dev_alloc(&a, (size_t)n);
  dev_alloc(&b, (size_t)n);
  HIP_CHECK(hipMemset(a, 1, (size_t)bytes));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  int cus = 0;
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)cus * 16));
  HIP_CHECK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < repeats; r++) hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, nullptr, a, b, n);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(e1, nullptr));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(a);
  dev_free(b);
  return ms * 1e-3 / repeats;
}

}  // namespace moc
