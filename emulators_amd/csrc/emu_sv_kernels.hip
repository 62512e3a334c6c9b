// emu_sv_kernels.hip -- gfx950 kernels and C ABI for the state-vector
// Krylov time evolution (H.psi inside Lanczos exponential steps).
//
// Layout: psi is a contiguous complex vector of 2^N elements (interleaved
// re/im, float2 for c64, double2 for c128); qubit q lives in bit N-1-q.
//
// H.psi is computed in one or more "passes".  A pass owns LDS tiles whose
// local index bits are a contiguous low range [0,c) plus one high range
// [s,s+m) of the global index.  All 2^(c+m) amplitudes of a tile are staged
// in LDS, so a flip X_b of any tile bit b is a read of a neighbouring LDS
// slot.  Pass 1 covers the low bits and adds the diagonal; later passes
// cover the remaining high bits with a few low bits kept for coalesced
// segments, and accumulate into the output.  The last pass also produces the
// per-workgroup partial sums of one Lanczos iteration (||w||^2, <v|w>,
// <prev|w>, <v|prev>).
//
// The diagonal  -sum_b d_b x_b + sum_{b<b'} U_bb' x_b x_b'  is never stored:
// per tile, the outer (fixed) bits give a constant E_out and per-local-bit
// offsets a_p; per thread the low 8 local bits are fixed, which leaves a
// per-thread base and one offset G_j per remaining local bit, plus a tiny
// per-tile table B(r) of the pair terms among those bits.  Cost: a few FP64
// adds per amplitude, all in FP64 even for c64 vectors.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <cmath>
#include <complex>
#include <atomic>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "emu_sv_hip.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    return fail(EMU_SV_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return EMU_SV_OK;
}

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kMaxTileLog = 14;   // LDS tile <= 2^14 elements
constexpr int kThreadBits = 8;    // log2(kThreads)
constexpr int kNumDots = 8;
constexpr int kDiagRow = 16;  // E_out, a_0 .. a_{t-1} of one pass-1 tile, stamp in the last slot
constexpr int kDiagStamp = kDiagRow - 1;
static_assert(kDiagStamp >= kMaxTileLog + 1,
              "a diagonal row holds E_out, one a_p per tile bit and the stamp");
constexpr int kStreamBlocks = 2048;  // grid of the streaming kernels
constexpr int kDefaultTileLogC64 = 13;
constexpr int kDefaultTileLogC128 = 12;
constexpr int kDefaultContigC64 = 3;
constexpr int kDefaultWaveLogC64 = 11;
constexpr int kDefaultWaveLogC128 = 10;
constexpr int kKernelLds = 1;
constexpr int kKernelWave = 2;
constexpr int kDefaultContigC128 = 2;

// Offsets into the packed parameter block (see emu_sv_hip.h).
constexpr int kCre = 0, kCim = 32, kDelta = 64, kU = 96, kFlagComplex = 1120;
constexpr int kFlagNoise = 1121, kNoise = 1122;  // K00, K01, K10, K11 (re, im)
constexpr int kGeneration = 1135;  // pack generation (process-unique, > 0; 0 = unstamped)

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Block reduction of K doubles held per thread; thread 0..K-1 writes
// out[k].  Uses a caller-provided LDS scratch of kWaves*K doubles.
template <int K>
__device__ __forceinline__ void block_reduce_store(double (&v)[K], double *red,
                                                   double *out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) red[wave * K + k] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s += red[w * K + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

struct PassArgs {
  const void *v;      // input vector (stored form); actual v = v_scale * v
  const void *prev;   // previous Lanczos vector or nullptr (last pass only)
  void *out;          // output (read-modify-write when accumulate)
  const double *ham;  // packed parameter block (device)
  double *partials;   // per-workgroup partial sums (kNumDots each) or nullptr
  const double *dtab; // per-tile diagonal constants of pass 1 (kDiagRow doubles per tile)
  // Orthogonalising load (pass 1 of a fused Lanczos matvec): the tile input
  // is z = v_scale*v - ca*pa - cb*pb, written back to v (zout) and reduced to
  // ||z||^2 in partials2.
  int ortho;
  void *zout;
  const void *pa, *pb;
  double ca_re, ca_im, cb_re, cb_im;
  double pb_scale;
  double *partials2;
  // Orthogonalising pass of a pipelined Lanczos iteration: when set, ca_re,
  // ca_im, cb_re, cb_im, pb_scale and v_scale are read from these 6 device
  // doubles instead (formed on the device by emu_sv_lanczos_next_coef).
  const double *dcoef;
  double v_scale, prev_scale;
  double sig_re, sig_im;
  int n;               // qubits
  int c, s, m;         // tile bits: [0,c) U [s,s+m)
  uint32_t flip_mask;  // global bits whose flip terms this pass applies
  int diag;            // add the diagonal term
  int accumulate;      // out += ...
  int xcd;             // XCD-aware tile order
};

// Load one input amplitude of a pass: v_scale*v[x], or in orthogonalising
// mode z = v_scale*v[x] - ca*pa[x] - cb*pb[x] (FP64 arithmetic, stored back
// to v and accumulated into nrm).
// Reductions of the orthogonalising load: ||z||^2, <z|pb_s>, <z|w_s> with
// pb_s = pb_scale*pb the current Lanczos vector and w_s = v_scale*v the
// previous matvec (5 doubles).
constexpr int kOrthoSums = 5;

// The arithmetic of load_input on already loaded v, pa[x], pb[x].
template <typename T>
__device__ __forceinline__ typename Vec2<T>::type finish_input(const PassArgs &a, uint64_t x,
                                                               typename Vec2<T>::type v,
                                                               typename Vec2<T>::type pa,
                                                               typename Vec2<T>::type pb,
                                                               double (&osum)[kOrthoSums]) {
  using T2 = typename Vec2<T>::type;
  T2 z;
  if (!a.ortho) {
    const T vs = static_cast<T>(a.v_scale);
    z.x = v.x * vs;
    z.y = v.y * vs;
    return z;
  }
  const double wr = a.v_scale * static_cast<double>(v.x), wi = a.v_scale * static_cast<double>(v.y);
  double zr = wr, zi = wi;
  if (a.pa) {
    zr -= a.ca_re * pa.x - a.ca_im * pa.y;
    zi -= a.ca_re * pa.y + a.ca_im * pa.x;
  }
  double br = 0.0, bi = 0.0;
  if (a.pb) {
    br = a.pb_scale * static_cast<double>(pb.x);
    bi = a.pb_scale * static_cast<double>(pb.y);
    zr -= a.cb_re * pb.x - a.cb_im * pb.y;
    zi -= a.cb_re * pb.y + a.cb_im * pb.x;
  }
  z.x = static_cast<T>(zr);
  z.y = static_cast<T>(zi);
  reinterpret_cast<T2 *>(a.zout)[x] = z;
  const double qr = z.x, qi = z.y;
  osum[0] += qr * qr + qi * qi;
  osum[1] += qr * br + qi * bi;
  osum[2] += qr * bi - qi * br;
  osum[3] += qr * wr + qi * wi;
  osum[4] += qr * wi - qi * wr;
  return z;
}

template <typename T>
__device__ __forceinline__ typename Vec2<T>::type load_input(const PassArgs &a, uint64_t x,
                                                             double (&osum)[kOrthoSums]) {
  using T2 = typename Vec2<T>::type;
  const T2 v = reinterpret_cast<const T2 *>(a.v)[x];
  T2 z;
  if (!a.ortho) {
    const T vs = static_cast<T>(a.v_scale);
    z.x = v.x * vs;
    z.y = v.y * vs;
    return z;
  }
  const double wr = a.v_scale * static_cast<double>(v.x), wi = a.v_scale * static_cast<double>(v.y);
  double zr = wr, zi = wi;
  if (a.pa) {
    const T2 p = reinterpret_cast<const T2 *>(a.pa)[x];
    zr -= a.ca_re * p.x - a.ca_im * p.y;
    zi -= a.ca_re * p.y + a.ca_im * p.x;
  }
  double br = 0.0, bi = 0.0;
  if (a.pb) {
    const T2 p = reinterpret_cast<const T2 *>(a.pb)[x];
    br = a.pb_scale * static_cast<double>(p.x);
    bi = a.pb_scale * static_cast<double>(p.y);
    zr -= a.cb_re * p.x - a.cb_im * p.y;
    zi -= a.cb_re * p.y + a.cb_im * p.x;
  }
  z.x = static_cast<T>(zr);
  z.y = static_cast<T>(zi);
  reinterpret_cast<T2 *>(a.zout)[x] = z;
  const double qr = z.x, qi = z.y;
  osum[0] += qr * qr + qi * qi;
  osum[1] += qr * br + qi * bi;
  osum[2] += qr * bi - qi * br;
  osum[3] += qr * wr + qi * wi;
  osum[4] += qr * wi - qi * wr;
  return z;
}

// ---------------------------------------------------------------------------
// H.psi pass kernel.  CPLX: complex flip coefficients (some phi != 0).
// DOTS: fuse the Lanczos partial sums (last pass).
// ---------------------------------------------------------------------------
// One tile of a pass: tile index g, diagonal constants row[0] = E_out,
// row[1 + p] = a_p (when a.diag), partial sums into row prow of the partials.
// NOISE (with CPLX): the block's single-qubit term K is added, K10 / K01 to
// the flip coefficients and (n - popcount(x)) K00 + popcount(x) K11 to the
// diagonal of amplitude x.
// The static LDS of one pass_tile, declared once by the calling kernel: a
// kernel that inlines several pass_tile variants (the batch engine's) then
// holds one copy instead of one per variant.
struct PassLds {
  double ap[kMaxTileLog];
  double U[kMaxTileLog][kMaxTileLog];
  double B[1 << (kMaxTileLog - kThreadBits)];
  double eout;
  double red[kWaves * kNumDots];
};

template <typename T, bool CPLX, bool DOTS, bool NOISE = false>
__device__ __forceinline__ void pass_tile(const PassArgs &a, uint64_t g, const double *row,
                                          uint64_t prow, PassLds &L) {
  static_assert(!NOISE || CPLX, "the noise term needs complex coefficients");
  using T2 = typename Vec2<T>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T2 *tile = reinterpret_cast<T2 *>(smem_raw);
  double *s_ap = L.ap;
  auto &s_U = L.U;
  double *s_B = L.B;
  double &s_eout = L.eout;
  double *s_red = L.red;

  const int tid = threadIdx.x;
  const int c = a.c, s = a.s, m = a.m, n = a.n;
  const int t = c + m;
  const uint32_t tsize = 1u << t;
  const int nlo = s - c;
  const uint64_t base = ((g & ((1ull << nlo) - 1)) << c) | ((g >> nlo) << (s + m));
  const uint32_t cmask = (1u << c) - 1;
  auto gidx = [&](uint32_t l) -> uint64_t {
    return base | (uint64_t)(l & cmask) | ((uint64_t)(l >> c) << s);
  };
  auto gbit = [&](int i) -> int { return i < c ? i : s + (i - c); };

  const double *__restrict__ H = a.ham;

  // Stage the tile.  Whole groups of four elements per thread issue all
  // their global loads before any arithmetic (several loads in flight).
  double osum[kOrthoSums] = {0, 0, 0, 0, 0};
  constexpr int SG = 4;
  if (tsize % (SG * kThreads) == 0) {
    const T2 *__restrict__ vin = reinterpret_cast<const T2 *>(a.v);
    const T2 *__restrict__ pav = reinterpret_cast<const T2 *>(a.pa);
    const T2 *__restrict__ pbv = reinterpret_cast<const T2 *>(a.pb);
    for (uint32_t l0 = tid; l0 < tsize; l0 += SG * kThreads) {
      T2 rv[SG], ra[SG], rb[SG];
      uint64_t xs[SG];
#pragma unroll
      for (int u = 0; u < SG; ++u) {
        xs[u] = gidx(l0 + u * kThreads);
        rv[u] = vin[xs[u]];
      }
      if (a.ortho && pav) {
#pragma unroll
        for (int u = 0; u < SG; ++u) ra[u] = pav[xs[u]];
      }
      if (a.ortho && pbv) {
#pragma unroll
        for (int u = 0; u < SG; ++u) rb[u] = pbv[xs[u]];
      }
#pragma unroll
      for (int u = 0; u < SG; ++u)
        tile[l0 + u * kThreads] = finish_input<T>(a, xs[u], rv[u], ra[u], rb[u], osum);
    }
  } else {
    for (uint32_t l = tid; l < tsize; l += kThreads) tile[l] = load_input<T>(a, gidx(l), osum);
  }
  if (a.diag) {
    // Outer-bit constants of this tile.
    if (tid < t) s_ap[tid] = row[1 + tid];
    if (tid == 63) s_eout = row[0];
    for (int e = tid; e < t * t; e += kThreads) {
      const int i = e / t, j = e % t;
      s_U[i][j] = H[kU + gbit(i) * 32 + gbit(j)];
    }
  }
  __syncthreads();

  const int tb = t < kThreadBits ? t : kThreadBits;  // thread-fixed local bits
  const int k = t - tb;                              // loop local bits
  double dbase = 0.0;
  double Gj[kMaxTileLog - kThreadBits];
  if (a.diag) {
    for (int r = tid; r < (1 << k); r += kThreads) {
      double e = 0.0;
      for (int j = 0; j < k; ++j)
        if ((r >> j) & 1)
          for (int j2 = 0; j2 < j; ++j2)
            if ((r >> j2) & 1) e += s_U[tb + j][tb + j2];
      s_B[r] = e;
    }
    dbase = s_eout;
    for (int p = 0; p < tb; ++p) {
      if ((tid >> p) & 1) {
        dbase += s_ap[p];
        for (int p2 = 0; p2 < p; ++p2)
          if ((tid >> p2) & 1) dbase += s_U[p][p2];
      }
    }
#pragma unroll
    for (int j = 0; j < kMaxTileLog - kThreadBits; ++j) {
      double gj = 0.0;
      if (j < k) {
        gj = s_ap[tb + j];
        for (int p = 0; p < tb; ++p)
          if ((tid >> p) & 1) gj += s_U[tb + j][p];
      }
      Gj[j] = gj;
    }
    __syncthreads();
  }

  // Flip coefficients of the local bits (uniform values).  With NOISE the
  // two directions differ: bit set <1|H|0> = c + K10 (cre, cim), bit clear
  // <0|H|1> = conj(c) + K01 (dre, dim).
  uint32_t fm_local = 0;
  T cre[kMaxTileLog], cim[kMaxTileLog];
  T dre[NOISE ? kMaxTileLog : 1], dim[NOISE ? kMaxTileLog : 1];
#pragma unroll
  for (int i = 0; i < kMaxTileLog; ++i) {
    cre[i] = 0;
    cim[i] = 0;
    if (NOISE) dre[i] = dim[i] = 0;
    if (i < t) {
      const int b = gbit(i);
      if ((a.flip_mask >> b) & 1u) {
        fm_local |= 1u << i;
        if (NOISE) {
          cre[i] = static_cast<T>(H[kCre + b] + H[kNoise + 4]);
          cim[i] = static_cast<T>(H[kCim + b] + H[kNoise + 5]);
          dre[i] = static_cast<T>(H[kCre + b] + H[kNoise + 2]);
          dim[i] = static_cast<T>(-H[kCim + b] + H[kNoise + 3]);
        } else {
          cre[i] = static_cast<T>(H[kCre + b]);
          cim[i] = static_cast<T>(H[kCim + b]);
        }
      }
    }
  }
  // diagonal noise: n K00 + popcount(x) (K11 - K00)
  double k0r = 0.0, k0i = 0.0, k1r = 0.0, k1i = 0.0;
  if (NOISE) {
    k0r = n * H[kNoise + 0];
    k0i = n * H[kNoise + 1];
    k1r = H[kNoise + 6] - H[kNoise + 0];
    k1i = H[kNoise + 7] - H[kNoise + 1];
  }
  const T sr = static_cast<T>(a.sig_re), si = static_cast<T>(a.sig_im);
  T2 *__restrict__ out = reinterpret_cast<T2 *>(a.out);
  const T2 *__restrict__ prevp = reinterpret_cast<const T2 *>(a.prev);
  const T ps = static_cast<T>(a.prev_scale);
  double dots[7] = {0, 0, 0, 0, 0, 0, 0};

  // One output amplitude: flips from LDS, the diagonal, sigma, the partial
  // sum o (when accumulating) and the dots with p = prev[x].
  auto emit = [&](uint32_t r, const T2 o, const T2 pv) {
    const uint32_t l = (uint32_t)tid | (r << tb);
    const T2 psi = tile[l];
    T hr = 0, hi = 0;
#pragma unroll
    for (int i = 0; i < kMaxTileLog; ++i) {
      if ((fm_local >> i) & 1u) {
        const T2 nb = tile[l ^ (1u << i)];
        if (NOISE) {
          const bool up = (l >> i) & 1u;
          const T fr = up ? cre[i] : dre[i], fi = up ? cim[i] : dim[i];
          hr += fr * nb.x - fi * nb.y;
          hi += fr * nb.y + fi * nb.x;
        } else if (CPLX) {
          // bit set: <1|H|0> = c ; bit clear: <0|H|1> = conj(c)
          const T ci = ((l >> i) & 1u) ? cim[i] : -cim[i];
          hr += cre[i] * nb.x - ci * nb.y;
          hi += cre[i] * nb.y + ci * nb.x;
        } else {
          hr += cre[i] * nb.x;
          hi += cre[i] * nb.y;
        }
      }
    }
    const uint64_t x = gidx(l);
    if (a.diag) {
      double e = dbase + s_B[r];
#pragma unroll
      for (int j = 0; j < kMaxTileLog - kThreadBits; ++j)
        if ((r >> j) & 1u) e += Gj[j];
      if (NOISE) {
        const double pc = (double)__popcll(x);
        const double er = e + k0r + pc * k1r, ei = k0i + pc * k1i;
        const double pr = static_cast<double>(psi.x), pi = static_cast<double>(psi.y);
        hr += static_cast<T>(er * pr - ei * pi);
        hi += static_cast<T>(er * pi + ei * pr);
      } else {
        hr += static_cast<T>(e * static_cast<double>(psi.x));
        hi += static_cast<T>(e * static_cast<double>(psi.y));
      }
    }
    T2 w;
    w.x = sr * hr - si * hi;
    w.y = sr * hi + si * hr;
    if (a.accumulate) {
      w.x += o.x;
      w.y += o.y;
    }
    out[x] = w;
    if (DOTS) {
      const double wr = w.x, wi = w.y, vr = psi.x, vi = psi.y;
      dots[0] += wr * wr + wi * wi;
      dots[1] += vr * wr + vi * wi;
      dots[2] += vr * wi - vi * wr;
      if (prevp) {
        const double pr = static_cast<double>(pv.x * ps), pi = static_cast<double>(pv.y * ps);
        dots[3] += pr * wr + pi * wi;
        dots[4] += pr * wi - pi * wr;
        dots[5] += vr * pr + vi * pi;
        dots[6] += vr * pi - vi * pr;
      }
    }
  };

  if (tid < (1 << tb)) {
    const uint32_t nr = 1u << k;
    T2 zero;
    zero.x = zero.y = 0;
    if (nr % SG == 0) {
      // groups of SG amplitudes: the partial-sum and prev loads first
      for (uint32_t r0 = 0; r0 < nr; r0 += SG) {
        T2 o[SG], pv[SG];
        uint64_t xs[SG];
#pragma unroll
        for (int u = 0; u < SG; ++u) {
          xs[u] = gidx((uint32_t)tid | ((r0 + u) << tb));
          o[u] = zero;
          pv[u] = zero;
        }
        if (a.accumulate) {
#pragma unroll
          for (int u = 0; u < SG; ++u) o[u] = out[xs[u]];
        }
        if (DOTS && prevp) {
#pragma unroll
          for (int u = 0; u < SG; ++u) pv[u] = prevp[xs[u]];
        }
#pragma unroll
        for (int u = 0; u < SG; ++u) emit(r0 + u, o[u], pv[u]);
      }
    } else {
      for (uint32_t r = 0; r < nr; ++r) {
        const uint64_t x = gidx((uint32_t)tid | (r << tb));
        const T2 o = a.accumulate ? out[x] : zero;
        const T2 pv = (DOTS && prevp) ? prevp[x] : zero;
        emit(r, o, pv);
      }
    }
  }
  if (DOTS) block_reduce_store<7>(dots, s_red, a.partials + prow * kNumDots);
  if (a.ortho) {
    __syncthreads();
    block_reduce_store<kOrthoSums>(osum, s_red, a.partials2 + prow * kNumDots);
  }
}

template <typename T, bool CPLX, bool DOTS, bool NOISE = false>
__global__ __launch_bounds__(kThreads) void hpsi_pass_kernel(const PassArgs args) {
  PassArgs a = args;
  if (args.dcoef) {
    a.ca_re = args.dcoef[0];
    a.ca_im = args.dcoef[1];
    a.cb_re = args.dcoef[2];
    a.cb_im = args.dcoef[3];
    a.pb_scale = args.dcoef[4];
    a.v_scale = args.dcoef[5];
  }
  const int t = a.c + a.m;
  const uint64_t ntiles = 1ull << (a.n - t);
  uint64_t g = blockIdx.x;
  if (a.xcd && ntiles >= 8) g = (g & 7) * (ntiles >> 3) + (g >> 3);
  __shared__ PassLds lds;
  pass_tile<T, CPLX, DOTS, NOISE>(a, g, a.diag ? a.dtab + g * kDiagRow : nullptr, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// Wave-tile H.psi pass kernel.  One wavefront owns a tile of 2^(6+KREG)
// amplitudes: local bits 0..5 are the lane index, local bits 6..6+KREG-1
// index KREG-bit register slots.  A flip of a register bit reads another
// register of the same lane; a flip of a lane bit is a cross-lane exchange
// (DPP quad_perm / row mirrors for xor 1,2,4,8, v_permlane16/32_swap for xor
// 16, 32).  No LDS, no barrier: occupancy is set by registers alone.
// ---------------------------------------------------------------------------
template <int MASK>
__device__ __forceinline__ int lane_xor(int v, int lane) {
  if constexpr (MASK == 1) {
    return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
  } else if constexpr (MASK == 2) {
    return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
  } else if constexpr (MASK == 4) {
    const int t = __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    return __builtin_amdgcn_update_dpp(0, t, 0x1B, 0xF, 0xF, false);         // quad_perm 3,2,1,0
  } else if constexpr (MASK == 8) {
    const int t = __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);  // row_mirror
    return __builtin_amdgcn_update_dpp(0, t, 0x141, 0xF, 0xF, false);        // row_half_mirror
  } else if constexpr (MASK == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false);
    return (lane & 16) ? (int)p[0] : (int)p[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    return (lane & 32) ? (int)p[0] : (int)p[1];
  }
}

template <int MASK>
__device__ __forceinline__ float lane_xor(float v, int lane) {
  return __int_as_float(lane_xor<MASK>(__float_as_int(v), lane));
}

template <int MASK>
__device__ __forceinline__ double lane_xor(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_xor<MASK>((int)(b & 0xffffffffll), lane);
  const int hi = lane_xor<MASK>((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <typename T, int P>
__device__ __forceinline__ void lane_flip_add(const typename Vec2<T>::type psi, int lane,
                                              uint32_t fm, T cr, T cis, T &hr, T &hi) {
  if ((fm >> P) & 1u) {
    const T nx = lane_xor<1 << P>(psi.x, lane);
    const T ny = lane_xor<1 << P>(psi.y, lane);
    hr += cr * nx - cis * ny;
    hi += cr * ny + cis * nx;
  }
}

template <typename T, int KREG, bool CPLX, bool DOTS>
__global__ __launch_bounds__(kThreads) void hpsi_wave_kernel(const PassArgs a) {
  using T2 = typename Vec2<T>::type;
  constexpr int R = 1 << KREG;
  constexpr int t = 6 + KREG;
  const int lane = threadIdx.x & 63;
  const int c = a.c, s = a.s, m = a.m, n = a.n;
  const uint64_t ntiles = 1ull << (n - t);
  uint64_t blk = blockIdx.x;
  const uint64_t nblk = ntiles >= (uint64_t)kWaves ? ntiles / kWaves : 1;
  if (a.xcd && nblk >= 8) blk = (blk & 7) * (nblk >> 3) + (blk >> 3);
  const uint64_t g = blk * kWaves + (threadIdx.x >> 6);
  if (g >= ntiles) return;  // whole wave exits (small problems only)
  const int nlo = s - c;
  const uint64_t base = ((g & ((1ull << nlo) - 1)) << c) | ((g >> nlo) << (s + m));
  const uint32_t cmask = (1u << c) - 1;
  auto gidx = [&](uint32_t l) -> uint64_t {
    return base | (uint64_t)(l & cmask) | ((uint64_t)(l >> c) << s);
  };
  auto gbit = [&](int i) -> int { return i < c ? i : s + (i - c); };
  const double *__restrict__ H = a.ham;
  T2 *__restrict__ out = reinterpret_cast<T2 *>(a.out);
  const T2 *__restrict__ prevp = reinterpret_cast<const T2 *>(a.prev);

  // Load the tile: slot r of lane holds local index lane | r << 6.
  T2 psi[R];
  double osum[kOrthoSums] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int r = 0; r < R; ++r) psi[r] = load_input<T>(a, gidx((uint32_t)lane | ((uint32_t)r << 6)), osum);

  // Flip coefficients per local bit (uniform), conj for a clear bit.
  uint32_t fm = 0;
  T cre[t], cim[t];
#pragma unroll
  for (int i = 0; i < t; ++i) {
    const int b = gbit(i);
    const bool on = (a.flip_mask >> b) & 1u;
    fm |= on ? (1u << i) : 0u;
    cre[i] = on ? static_cast<T>(H[kCre + b]) : T(0);
    cim[i] = on ? static_cast<T>(H[kCim + b]) : T(0);
  }

  // Diagonal: E(lane, r) = dbase(lane) + sum_{j in r} G_j(lane) + sum_{j<j' in r} U_jj'.
  double dbase = 0.0, Gj[KREG], Ur[KREG][KREG];
  if (a.diag) {
    // Outer-bit constants of this tile, tabulated by diag_prep_kernel.
    const double *row = a.dtab + g * kDiagRow;
    const double ai = lane < t ? row[1 + lane] : 0.0;
    const double eout = row[0];
    dbase = eout;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      const double ap = __shfl(ai, p, 64);
      if ((lane >> p) & 1) {
        dbase += ap;
#pragma unroll
        for (int p2 = 0; p2 < p; ++p2)
          if ((lane >> p2) & 1) dbase += H[kU + gbit(p) * 32 + gbit(p2)];
      }
    }
#pragma unroll
    for (int j = 0; j < KREG; ++j) {
      double gj = __shfl(ai, 6 + j, 64);
#pragma unroll
      for (int p = 0; p < 6; ++p)
        if ((lane >> p) & 1) gj += H[kU + gbit(6 + j) * 32 + gbit(p)];
      Gj[j] = gj;
#pragma unroll
      for (int j2 = 0; j2 < KREG; ++j2) Ur[j][j2] = H[kU + gbit(6 + j) * 32 + gbit(6 + j2)];
    }
  }

  // Per-lane signed imaginary parts for the lane-bit flips.
  T cis[6];
#pragma unroll
  for (int p = 0; p < 6; ++p) cis[p] = ((lane >> p) & 1) ? cim[p] : -cim[p];

  const T sr = static_cast<T>(a.sig_re), si = static_cast<T>(a.sig_im);
  const T ps = static_cast<T>(a.prev_scale);
  double dots[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const T2 p0 = psi[r];
    T hr = 0, hi = 0;
    // register-bit flips (compile-time slots)
#pragma unroll
    for (int j = 0; j < KREG; ++j) {
      if ((fm >> (6 + j)) & 1u) {
        const T2 nb = psi[r ^ (1 << j)];
        const T ci = CPLX ? (((r >> j) & 1) ? cim[6 + j] : -cim[6 + j]) : T(0);
        hr += cre[6 + j] * nb.x - ci * nb.y;
        hi += cre[6 + j] * nb.y + ci * nb.x;
      }
    }
    // lane-bit flips
    lane_flip_add<T, 0>(p0, lane, fm, cre[0], CPLX ? cis[0] : T(0), hr, hi);
    lane_flip_add<T, 1>(p0, lane, fm, cre[1], CPLX ? cis[1] : T(0), hr, hi);
    lane_flip_add<T, 2>(p0, lane, fm, cre[2], CPLX ? cis[2] : T(0), hr, hi);
    lane_flip_add<T, 3>(p0, lane, fm, cre[3], CPLX ? cis[3] : T(0), hr, hi);
    lane_flip_add<T, 4>(p0, lane, fm, cre[4], CPLX ? cis[4] : T(0), hr, hi);
    lane_flip_add<T, 5>(p0, lane, fm, cre[5], CPLX ? cis[5] : T(0), hr, hi);
    if (a.diag) {
      double e = dbase;
#pragma unroll
      for (int j = 0; j < KREG; ++j) {
        if ((r >> j) & 1) {
          e += Gj[j];
#pragma unroll
          for (int j2 = 0; j2 < j; ++j2)
            if ((r >> j2) & 1) e += Ur[j][j2];
        }
      }
      hr += static_cast<T>(e * static_cast<double>(p0.x));
      hi += static_cast<T>(e * static_cast<double>(p0.y));
    }
    T2 w;
    w.x = sr * hr - si * hi;
    w.y = sr * hi + si * hr;
    const uint64_t x = gidx((uint32_t)lane | ((uint32_t)r << 6));
    if (a.accumulate) {
      const T2 o = out[x];
      w.x += o.x;
      w.y += o.y;
    }
    out[x] = w;
    if (DOTS) {
      const double wr = w.x, wi = w.y, vr = p0.x, vi = p0.y;
      dots[0] += wr * wr + wi * wi;
      dots[1] += vr * wr + vi * wi;
      dots[2] += vr * wi - vi * wr;
      if (prevp) {
        const T2 q = prevp[x];
        const double pr = static_cast<double>(q.x * ps), pi = static_cast<double>(q.y * ps);
        dots[3] += pr * wr + pi * wi;
        dots[4] += pr * wi - pi * wr;
        dots[5] += vr * pr + vi * pi;
        dots[6] += vr * pi - vi * pr;
      }
    }
  }
  if (DOTS) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const double sm = wave_sum(dots[k]);
      if (lane == 0) a.partials[g * kNumDots + k] = sm;
    }
  }
  if (a.ortho) {
#pragma unroll
    for (int k = 0; k < kOrthoSums; ++k) {
      const double sm = wave_sum(osum[k]);
      if (lane == 0) a.partials2[g * kNumDots + k] = sm;
    }
  }
}

// Per-tile constants of the diagonal for a pass-1 tile [0, t): with h the
// outer bits (x >> t), E_out(h) = sum_{q in h} -d_q + sum_{q<q' in h} U_qq'
// and a_p(h) = -d_p + sum_{q in h} U_pq for the t tile bits.  One thread per
// table entry (row h, entry p): entry 0 is E_out(h), entry 1 + p is a_p(h),
// entries beyond t are zero, and the last entry is the row's stamp
// (diag_stamp: the block's pack generation, n and t).  The table is written
// coalesced.  With keep, a row whose stamp matches is left as it is; rows of
// an unstamped block (generation 0) are always rebuilt.
__device__ __forceinline__ double diag_stamp(const double *H, int n, int t) {
  // exact in FP64 while the generation stays below 2^41
  return H[kGeneration] * 4096.0 + (double)(n * 64 + t);
}

__global__ __launch_bounds__(kThreads) void diag_prep_kernel(int n, int t, const double *H,
                                                             double *table, int keep) {
  const uint64_t ntiles = 1ull << (n - t);
  const uint64_t idx = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  const uint64_t h = idx / kDiagRow;
  const int p = (int)(idx % kDiagRow);
  const double stamp = diag_stamp(H, n, t);
  // every row lies inside one block (kThreads is a multiple of kDiagRow):
  // all its entries read the old stamp before any entry is rewritten
  const bool fresh = h < ntiles && keep && H[kGeneration] > 0.0 &&
                     table[h * kDiagRow + kDiagStamp] == stamp;
  __syncthreads();
  if (h >= ntiles || fresh) return;
  const uint64_t outer = h << t;
  double e = 0.0;
  if (p == kDiagStamp) {
    e = stamp;
  } else if (p == 0) {
    for (int q = t; q < n; ++q) {
      if (!((outer >> q) & 1)) continue;
      e -= H[kDelta + q];
      for (int q2 = q + 1; q2 < n; ++q2)
        if ((outer >> q2) & 1) e += H[kU + q * 32 + q2];
    }
  } else if (p <= t) {
    e = -H[kDelta + p - 1];
    for (int q = t; q < n; ++q)
      if ((outer >> q) & 1) e += H[kU + (p - 1) * 32 + q];
  }
  table[idx] = e;
}
static_assert(kThreads % kDiagRow == 0, "diag_prep_kernel keeps each row inside one block");

// Two-stage deterministic reduction of nrow rows of kNumDots partial sums:
// every block sums a fixed contiguous chunk, the last block to finish (ticket
// counter) adds the block sums in block order, writes K results and re-arms
// the counter.
constexpr int kFinalizeBlocks = 128;
__global__ __launch_bounds__(kThreads) void finalize2_kernel(const double *rows, int nrow, int K,
                                                             double *stage, unsigned *ticket,
                                                             double *out) {
  __shared__ double red[kThreads];
  __shared__ bool last;
  const int per = (nrow + gridDim.x - 1) / gridDim.x;
  const int r0 = blockIdx.x * per;
  const int r1 = min(nrow, r0 + per);
  // thread = (row lane, value) with 32 row lanes x 8 values
  const int v = threadIdx.x & 7, rl = threadIdx.x >> 3;
  double acc = 0.0;
  for (int r = r0 + rl; r < r1; r += kThreads / 8) acc += rows[(uint64_t)r * kNumDots + v];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int hlf = kThreads / 2; hlf >= 8; hlf >>= 1) {
    if (threadIdx.x < hlf) red[threadIdx.x] += red[threadIdx.x + hlf];
    __syncthreads();
  }
  if (threadIdx.x < kNumDots) stage[blockIdx.x * kNumDots + threadIdx.x] = red[threadIdx.x];
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  // same (row lane, value) layout over the stage rows, then the same tree
  double sm = 0.0;
  for (unsigned b = rl; b < gridDim.x; b += kThreads / 8)
    sm += __hip_atomic_load(&stage[b * kNumDots + v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  red[threadIdx.x] = sm;
  __syncthreads();
  for (int hlf = kThreads / 2; hlf >= 8; hlf >>= 1) {
    if (threadIdx.x < hlf) red[threadIdx.x] += red[threadIdx.x + hlf];
    __syncthreads();
  }
  if (threadIdx.x < K) out[threadIdx.x] = red[threadIdx.x];
  if (threadIdx.x == 0) atomicExch(ticket, 0u);
}

// Sum nblk rows of K partials in a fixed order (deterministic).
__global__ __launch_bounds__(kThreads) void finalize_kernel(const double *partials, int nblk,
                                                            int stride, int K, double *out) {
  __shared__ double red[kThreads];
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads) s += partials[(uint64_t)b * stride + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kThreads / 2; h > 0; h >>= 1) {
      if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
    __syncthreads();
  }
}

// Streaming kernels work on 16-byte packets: one complex128 or two complex64
// amplitudes per lane and load.
template <typename T> struct Packet;
template <> struct Packet<float> {
  using type = float4;
  static constexpr int amps = 2;
};
template <> struct Packet<double> {
  using type = double2;
  static constexpr int amps = 1;
};

template <typename P>
__device__ __forceinline__ double packet_re(const P &p, int k);
template <>
__device__ __forceinline__ double packet_re<float4>(const float4 &p, int k) { return k ? p.z : p.x; }
template <>
__device__ __forceinline__ double packet_re<double2>(const double2 &p, int) { return p.x; }
template <typename P>
__device__ __forceinline__ double packet_im(const P &p, int k);
template <>
__device__ __forceinline__ double packet_im<float4>(const float4 &p, int k) { return k ? p.w : p.y; }
template <>
__device__ __forceinline__ double packet_im<double2>(const double2 &p, int) { return p.y; }
__device__ __forceinline__ void packet_set(float4 &p, int k, double re, double im) {
  if (k) {
    p.z = (float)re;
    p.w = (float)im;
  } else {
    p.x = (float)re;
    p.y = (float)im;
  }
}
__device__ __forceinline__ void packet_set(double2 &p, int, double re, double im) {
  p.x = re;
  p.y = im;
}

// w <- w - ca*pa - cb*pb ; partial ||w||^2
template <typename T>
__global__ __launch_bounds__(kThreads) void lanczos_update_kernel(uint64_t npk, void *wv,
                                                                  const void *pav, double car,
                                                                  double cai, const void *pbv,
                                                                  double cbr, double cbi,
                                                                  double *partials) {
  using P = typename Packet<T>::type;
  P *w = reinterpret_cast<P *>(wv);
  const P *pa = reinterpret_cast<const P *>(pav);
  const P *pb = reinterpret_cast<const P *>(pbv);
  double nrm[1] = {0.0};
  __shared__ double red[kWaves];
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < npk; x += stride) {
    P z = w[x];
    P qa, qb;
    if (pa) qa = pa[x];
    if (pb) qb = pb[x];
#pragma unroll
    for (int k = 0; k < Packet<T>::amps; ++k) {
      double zr = packet_re(z, k), zi = packet_im(z, k);
      if (pa) {
        const double ar = packet_re(qa, k), ai = packet_im(qa, k);
        zr -= car * ar - cai * ai;
        zi -= car * ai + cai * ar;
      }
      if (pb) {
        const double br = packet_re(qb, k), bi = packet_im(qb, k);
        zr -= cbr * br - cbi * bi;
        zi -= cbr * bi + cbi * br;
      }
      packet_set(z, k, zr, zi);
      const double qr = packet_re(z, k), qi = packet_im(z, k);
      nrm[0] += qr * qr + qi * qi;
    }
    w[x] = z;
  }
  block_reduce_store<1>(nrm, red, partials + (uint64_t)blockIdx.x * kNumDots);
}

// ||a||^2 and <a|b>
template <typename T>
__global__ __launch_bounds__(kThreads) void dot_kernel(uint64_t npk, const void *av, const void *bv,
                                                       double *partials) {
  using P = typename Packet<T>::type;
  const P *a = reinterpret_cast<const P *>(av);
  const P *b = reinterpret_cast<const P *>(bv);
  double acc[3] = {0.0, 0.0, 0.0};
  __shared__ double red[kWaves * 3];
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < npk; x += stride) {
    const P p = a[x];
    P q;
    if (b) q = b[x];
#pragma unroll
    for (int k = 0; k < Packet<T>::amps; ++k) {
      const double ar = packet_re(p, k), ai = packet_im(p, k);
      acc[0] += ar * ar + ai * ai;
      if (b) {
        const double br = packet_re(q, k), bi = packet_im(q, k);
        acc[1] += ar * br + ai * bi;
        acc[2] += ar * bi - ai * br;
      }
    }
  }
  block_reduce_store<3>(acc, red, partials + (uint64_t)blockIdx.x * kNumDots);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void scale_kernel(uint64_t npk, void *vv, T s) {
  using P = typename Packet<T>::type;
  P *v = reinterpret_cast<P *>(vv);
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < npk; x += stride) {
    P z = v[x];
#pragma unroll
    for (int k = 0; k < Packet<T>::amps; ++k)
      packet_set(z, k, packet_re(z, k) * s, packet_im(z, k) * s);
    v[x] = z;
  }
}

// One launch of the Arnoldi modified Gram-Schmidt chain: w <- w - h u_prev
// with h = pscale2 * (rec[1], rec[2]) (the previous launch's reduced
// overlap, read from device memory), then the partial sums ||w||^2 and
// <u_next|w> of the updated w.  u_prev / u_next may be null.
template <typename T>
__global__ __launch_bounds__(kThreads) void arnoldi_mgs_kernel(uint64_t npk, void *wv,
                                                               const void *upv, const double *rec,
                                                               double pscale2, const void *unv,
                                                               double *partials) {
  using P = typename Packet<T>::type;
  P *w = reinterpret_cast<P *>(wv);
  const P *up = reinterpret_cast<const P *>(upv);
  const P *un = reinterpret_cast<const P *>(unv);
  double hr = 0.0, hi = 0.0;
  if (up) {
    hr = pscale2 * rec[1];
    hi = pscale2 * rec[2];
  }
  double acc[3] = {0.0, 0.0, 0.0};
  __shared__ double red[kWaves * 3];
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < npk; x += stride) {
    P z = w[x];
    P a, b;
    if (up) a = up[x];
    if (un) b = un[x];
#pragma unroll
    for (int k = 0; k < Packet<T>::amps; ++k) {
      if (up) {
        const double ar = packet_re(a, k), ai = packet_im(a, k);
        const double zr = packet_re(z, k) - (hr * ar - hi * ai);
        const double zi = packet_im(z, k) - (hr * ai + hi * ar);
        packet_set(z, k, zr, zi);
      }
      const double qr = packet_re(z, k), qi = packet_im(z, k);
      acc[0] += qr * qr + qi * qi;
      if (un) {
        const double br = packet_re(b, k), bi = packet_im(b, k);
        acc[1] += br * qr + bi * qi;
        acc[2] += br * qi - bi * qr;
      }
    }
    if (up) w[x] = z;
  }
  block_reduce_store<3>(acc, red, partials + (uint64_t)blockIdx.x * kNumDots);
}

// The whole modified Gram-Schmidt of a small vector (D <= 2^13) in one
// workgroup: w stays in registers (NT threads, EPT = D / NT elements each),
// every step reduces (||w||^2, <u_k|w>) over the workgroup, subtracts
// s_k^2 <u_k|w> u_k and rounds w to the storage type, as the launch chain
// does.  One launch instead of 2 (count + 1).  The steps are sequential, so
// the cost is the per-step reduction: four waves (eight at D = 2^13) beat
// sixteen (n = 10: 15 us against 33 us; n = 12: 27 us against 41 us).
constexpr int kMgsSmallLog = 13;
constexpr int kArnoldiMaxDim = 256;  // emu_sv_arnoldi_exp: Krylov dimension limit
constexpr int kMgsSmallCount = 128;
struct MgsSmallArgs {
  const void *u[kMgsSmallCount];
  double s2[kMgsSmallCount];
};

// Workgroup sums with one barrier: consecutive calls alternate between two
// buffers (red = the call's own one).  A thread writing buffer b in call k+2
// has passed call k+1's barrier, which every thread reaches only after its
// reads of buffer b in call k, so no second barrier is needed.
template <int NT, int K>
__device__ __forceinline__ void mgs_block_sum(double (&v)[K], double *red) {
  constexpr int NW = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double sm = wave_sum(v[k]);
    if (lane == 0) red[wave * K + k] = sm;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double sm = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) sm += red[w * K + k];  // same order in every thread
    v[k] = sm;
  }
}

template <typename T, int EPT, int NT>
__global__ __launch_bounds__(NT) void arnoldi_mgs_small_kernel(int count, void *wv,
                                                               const MgsSmallArgs args,
                                                               double *records, int normalize) {
  using T2 = typename Vec2<T>::type;
  __shared__ double red[2][(NT / 64) * 3];  // alternating per workgroup sum
  T2 *w = reinterpret_cast<T2 *>(wv);
  double wr[EPT], wi[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const T2 v = w[threadIdx.x + e * NT];
    wr[e] = v.x;
    wi[e] = v.y;
  }
  for (int k = 0; k < count; ++k) {
    const T2 *u = reinterpret_cast<const T2 *>(args.u[k]);
    double ur[EPT], ui[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const T2 q = u[threadIdx.x + e * NT];
      ur[e] = q.x;
      ui[e] = q.y;
    }
    double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      acc[0] += wr[e] * wr[e] + wi[e] * wi[e];
      acc[1] += ur[e] * wr[e] + ui[e] * wi[e];
      acc[2] += ur[e] * wi[e] - ui[e] * wr[e];
    }
    mgs_block_sum<NT, 3>(acc, red[k & 1]);
    if (threadIdx.x == 0) {
      records[4 * k] = acc[0];
      records[4 * k + 1] = acc[1];
      records[4 * k + 2] = acc[2];
    }
    const double hr = args.s2[k] * acc[1], hi = args.s2[k] * acc[2];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const double zr = wr[e] - (hr * ur[e] - hi * ui[e]);
      const double zi = wi[e] - (hr * ui[e] + hi * ur[e]);
      wr[e] = static_cast<double>(static_cast<T>(zr));
      wi[e] = static_cast<double>(static_cast<T>(zi));
    }
  }
  double nrm[1] = {0.0};
#pragma unroll
  for (int e = 0; e < EPT; ++e) nrm[0] += wr[e] * wr[e] + wi[e] * wi[e];
  mgs_block_sum<NT, 1>(nrm, red[count & 1]);
  if (threadIdx.x == 0) records[4 * count] = nrm[0];
  // normalize: store the unit vector (the next basis vector), rounded once
  const double sc = (normalize && nrm[0] > 0.0) ? 1.0 / sqrt(nrm[0]) : 1.0;
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    T2 o;
    o.x = static_cast<T>(wr[e] * sc);
    o.y = static_cast<T>(wi[e] * sc);
    w[threadIdx.x + e * NT] = o;
  }
}

// w <- w / sqrt(*norm2) in place (left as it is when the norm is zero): the
// normalisation of the launch-chain Arnoldi step, from its device record.
template <typename T>
__global__ __launch_bounds__(kThreads) void scale_by_norm_kernel(uint64_t npk, void *wv, const double *norm2) {
  using P = typename Packet<T>::type;
  P *v = reinterpret_cast<P *>(wv);
  const double nn = *norm2;
  if (!(nn > 0.0)) return;
  const double s = 1.0 / sqrt(nn);
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < npk; x += stride) {
    P z = v[x];
#pragma unroll
    for (int k = 0; k < Packet<T>::amps; ++k)
      packet_set(z, k, static_cast<double>(packet_re(z, k)) * s, static_cast<double>(packet_im(z, k)) * s);
    v[x] = z;
  }
}

template <typename T>
int launch_mgs_small(int n, int count, const void *const *vecs, const double *scales, void *w,
                     double *records, int normalize, hipStream_t st) {
  MgsSmallArgs args;
  for (int k = 0; k < count; ++k) {
    args.u[k] = vecs[k];
    args.s2[k] = scales[k] * scales[k];
  }
  switch (n) {
    case 10: hipLaunchKernelGGL((arnoldi_mgs_small_kernel<T, 4, 256>), dim3(1), dim3(256), 0, st, count, w, args, records, normalize); break;
    case 11: hipLaunchKernelGGL((arnoldi_mgs_small_kernel<T, 8, 256>), dim3(1), dim3(256), 0, st, count, w, args, records, normalize); break;
    case 12: hipLaunchKernelGGL((arnoldi_mgs_small_kernel<T, 16, 256>), dim3(1), dim3(256), 0, st, count, w, args, records, normalize); break;
    case 13: hipLaunchKernelGGL((arnoldi_mgs_small_kernel<T, 16, 512>), dim3(1), dim3(512), 0, st, count, w, args, records, normalize); break;
    default: return fail(EMU_SV_EINVAL, "arnoldi_orthogonalize: small path needs 10 <= n <= 13");
  }
  return hip_check(hipGetLastError(), "arnoldi_mgs_small launch");
}

// ---------------------------------------------------------------------------
// Arnoldi Gram-Schmidt with the Gram matrix of the basis (emu_sv_arnoldi_gram).
// The reference's modified Gram-Schmidt (krylov_exp.py:54-58 with
// is_hermitian=False) subtracts the basis vectors a_k one at a time,
// h_k = <a_k|w_k>, w_{k+1} = w_k - h_k a_k.  Expanding w_k gives
//   h_k = c_k - sum_{i<k} h_i G(k, i),  c_k = <a_k|w>,  G(k, i) = <a_k|a_i>,
// so the same coefficients follow from one sweep of overlaps and the Gram
// matrix of the basis: the new row G(j, .) of a_j is reduced in the same
// sweep (it reads a_j anyway), earlier rows come from earlier calls.  Per
// call: K1 the overlaps (count + 1 vectors read), K2 the reduction and the
// forward substitution (one workgroup), K3 w' = a-scaled w - sum h_k a_k
// with ||w'||^2 (count + 1 read, 1 written), K4 the norm and the next
// basis vector's scale 1 / ||w'||.  The basis is stored unnormalised
// (u_k, a_k = s_k u_k) with the scales on the device, so no normalisation
// pass is needed and the caller can enqueue the next matvec (on u_{j+1},
// scale 1) before reading this call's records.
// State (doubles, ld >= count + 1): [0, ld) scales s_k; G(k, i) (i < k) as
// (re, im) at ld + 2 (k (k - 1) / 2 + i) (packed lower triangle); 2 ld K3
// coefficients; 8 misc; then the partial sums, kGramBlocks rows of 4 ld + 4.
// ---------------------------------------------------------------------------
constexpr int kGramMaxCount = 256;
constexpr int kGramBlocks = 256;
constexpr int kGramEpt = 2;  // packets per thread and chunk of the overlap sweep
struct GramVecs {
  const void *u[kGramMaxCount];
};
__host__ __device__ constexpr size_t gram_off_g(int ld) { return (size_t)ld; }
__host__ __device__ constexpr size_t gram_off_coef(int ld) { return (size_t)ld + (size_t)ld * (ld - 1); }
__host__ __device__ constexpr size_t gram_tri(int k, int i) { return 2 * ((size_t)k * (k - 1) / 2 + i); }
__host__ __device__ constexpr size_t gram_off_misc(int ld) { return gram_off_coef(ld) + 2 * (size_t)ld; }
__host__ __device__ constexpr size_t gram_off_part(int ld) { return gram_off_misc(ld) + 8; }
__host__ __device__ constexpr size_t gram_part_stride(int ld) { return 4 * (size_t)ld + 4; }
__host__ __device__ constexpr size_t gram_state_doubles(int ld) {
  return gram_off_part(ld) + (size_t)kGramBlocks * gram_part_stride(ld);
}

// K1: workgroup (x, y) sums, over its chunks of the vectors, <u_i|w> and
// <u_i|u_j> (j = count - 1; w and u_j stay in registers per chunk) for the
// kGramGroup basis vectors i of group y, and (y = 0) ||w||^2.  Partial row x
// layout: 4 i + (0, 1) = Re, Im <u_i|w>; 4 i + (2, 3) = Re, Im <u_i|u_j>;
// 4 count = ||w||^2.  The second grid dimension fills the GPU on the small
// vectors of the noisy paths (2^14 .. 2^18 elements) without more chunks.
constexpr int kGramGroup = 8;
template <typename T>
__global__ __launch_bounds__(kThreads) void arnoldi_gram_dots_kernel(uint64_t npk, int count,
                                                                     const GramVecs vecs,
                                                                     const void *wv, double *part,
                                                                     uint64_t stride) {
  using P = typename Packet<T>::type;
  constexpr int E = kGramEpt, A = Packet<T>::amps, GI = kGramGroup;
  __shared__ double s_acc[kWaves][GI][4];
  __shared__ double s_w[kWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i0 = blockIdx.y * GI;
  const int ni = min(GI, count - i0);
  for (int v = threadIdx.x; v < kWaves * GI * 4; v += kThreads) (&s_acc[0][0][0])[v] = 0.0;
  __syncthreads();
  const P *w = reinterpret_cast<const P *>(wv);
  const P *uj = reinterpret_cast<const P *>(vecs.u[count - 1]);
  const uint64_t chunk = (uint64_t)kThreads * E;
  const uint64_t nchunks = (npk + chunk - 1) / chunk;
  P zero;
  for (int k = 0; k < A; ++k) packet_set(zero, k, 0.0, 0.0);
  double wn = 0.0;
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    P wp[E], jp[E];
    uint64_t xs[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      xs[e] = c * chunk + (uint64_t)e * kThreads + threadIdx.x;
      wp[e] = xs[e] < npk ? w[xs[e]] : zero;
      jp[e] = xs[e] < npk ? uj[xs[e]] : zero;
    }
    if (blockIdx.y == 0) {
#pragma unroll
      for (int e = 0; e < E; ++e)
#pragma unroll
        for (int k = 0; k < A; ++k) {
          const double r = packet_re(wp[e], k), m = packet_im(wp[e], k);
          wn += r * r + m * m;
        }
    }
    // the group's loads first, then its sums
    P q[GI][E];
#pragma unroll
    for (int g = 0; g < GI; ++g) {
      const int i = i0 + g;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        q[g][e] = jp[e];  // i = j (u_j is in registers) and the group's unused slots
        if (g < ni && i != count - 1 && xs[e] < npk) q[g][e] = reinterpret_cast<const P *>(vecs.u[i])[xs[e]];
      }
    }
#pragma unroll
    for (int g = 0; g < GI; ++g) {
      if (g >= ni) break;
      double a4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int e = 0; e < E; ++e)
#pragma unroll
        for (int k = 0; k < A; ++k) {
          const double qr = packet_re(q[g][e], k), qi = packet_im(q[g][e], k);
          const double wr = packet_re(wp[e], k), wi = packet_im(wp[e], k);
          const double jr = packet_re(jp[e], k), ji = packet_im(jp[e], k);
          a4[0] += qr * wr + qi * wi;
          a4[1] += qr * wi - qi * wr;
          a4[2] += qr * jr + qi * ji;
          a4[3] += qr * ji - qi * jr;
        }
#pragma unroll
      for (int v = 0; v < 4; ++v) a4[v] = wave_sum(a4[v]);
      if (lane == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) s_acc[wave][g][v] += a4[v];
      }
    }
  }
  wn = wave_sum(wn);
  if (lane == 0) s_w[wave] = wn;
  __syncthreads();
  double *row = part + (uint64_t)blockIdx.x * stride;
  for (int v = threadIdx.x; v < 4 * ni; v += kThreads) {
    double sm = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) sm += s_acc[q][v >> 2][v & 3];
    row[4 * i0 + v] = sm;
  }
  if (blockIdx.y == 0 && threadIdx.x == 0) {
    double sm = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) sm += s_w[q];
    row[4 * count] = sm;
  }
}

// K2 (one workgroup of 1024 threads): the partial rows summed in a fixed
// order (four row slices of every value, each with its loads in flight
// together, then the slices in order), c_k and the new Gram row, then the
// forward substitution L h = c, L unit lower triangular with L(k, i) =
// G(k, i).  The basis is orthonormal to rounding, so |L - I| is at rounding
// level and Jacobi sweeps h <- c - (L - I) h, one thread per row, reach the
// substitution's solution within rounding in two sweeps (the error shrinks
// by |L - I| per sweep); three are run, on the Gram triangle staged in LDS
// (up to 64 basis vectors).  A basis whose Gram matrix is off by more than
// 1e-3 anywhere is solved by sequential substitution instead (thread 0).
// Writes the records (row k: [0, Re h_k, Im h_k, 0], row 0 [0] = ||a-scaled
// w||^2), the K3 coefficients h_k s_k, s_j (misc[0]) and G(j, .).
constexpr int kGramSweeps = 3;
constexpr int kGramSolveThreads = 1024;
constexpr int kGramLdsCount = 64;  // Gram triangles up to this count staged in LDS
__global__ __launch_bounds__(kGramSolveThreads) void arnoldi_gram_solve_kernel(int count, int nb, int ld,
                                                                               double s0, double *state,
                                                                               double *records) {
  constexpr int NT = kGramSolveThreads, S = 4, V = NT / S;
  __shared__ double sums[4 * kGramMaxCount + 1];
  __shared__ double s_part[S][V];
  __shared__ double s_c[kGramMaxCount][2];
  __shared__ double s_h[2][kGramMaxCount][2];
  __shared__ double s_s[kGramMaxCount];
  __shared__ double s_G[kGramLdsCount * (kGramLdsCount - 1)];  // packed triangle, (re, im)
  __shared__ int s_far;
  double *sc = state;
  double *G = state + gram_off_g(ld);
  double *coef = state + gram_off_coef(ld);
  double *misc = state + gram_off_misc(ld);
  const double *part = state + gram_off_part(ld);
  const uint64_t stride = gram_part_stride(ld);
  const int tid = threadIdx.x;
  const int nvals = 4 * count + 1;
  const int j = count - 1;
  const bool lds_g = count <= kGramLdsCount;
  if (tid == 0) s_far = 0;
  // earlier rows of the Gram triangle (written by earlier calls), staged
  if (lds_g)
    for (int e = tid; e < j * (j - 1); e += NT) s_G[e] = G[e];
  for (int k = tid; k < count; k += NT) s_s[k] = k == 0 ? s0 : sc[k];
  for (int v0 = 0; v0 < nvals; v0 += V) {
    const int v = v0 + tid % V, sl = tid / V;
    double acc = 0.0;
    if (v < nvals) {
#pragma unroll 8
      for (int b = sl; b < nb; b += S) acc += part[(uint64_t)b * stride + v];
    }
    s_part[sl][tid % V] = acc;
    __syncthreads();
    if (sl == 0 && v < nvals) sums[v] = ((s_part[0][tid] + s_part[1][tid]) + s_part[2][tid]) + s_part[3][tid];
    __syncthreads();
  }
  const double sj = s_s[j];
  for (int k = tid; k < count; k += NT) {
    const double f = s_s[k] * sj;
    s_c[k][0] = f * sums[4 * k];
    s_c[k][1] = f * sums[4 * k + 1];
    s_h[0][k][0] = s_c[k][0];
    s_h[0][k][1] = s_c[k][1];
    if (k < j) {  // G(j, k) = <a_j|a_k> = conj(<a_k|a_j>)
      const double gr = f * sums[4 * k + 2], gi = -f * sums[4 * k + 3];
      G[gram_tri(j, k)] = gr;
      G[gram_tri(j, k) + 1] = gi;
      if (lds_g) {
        s_G[gram_tri(j, k)] = gr;
        s_G[gram_tri(j, k) + 1] = gi;
      }
    }
  }
  if (tid == 0) {
    sc[0] = s0;
    misc[0] = sj;
  }
  __syncthreads();
  // G(k, i) from LDS when staged; row j from the sums otherwise
  auto gram_at = [&](int k, int i, double &gr, double &gi) {
    if (lds_g) {
      gr = s_G[gram_tri(k, i)];
      gi = s_G[gram_tri(k, i) + 1];
    } else if (k == j) {
      const double f = s_s[i] * sj;
      gr = f * sums[4 * i + 2];
      gi = -f * sums[4 * i + 3];
    } else {
      gr = G[gram_tri(k, i)];
      gi = G[gram_tri(k, i) + 1];
    }
  };
  int cur = 0;
  for (int sw = 0; sw < kGramSweeps; ++sw) {
    for (int k = tid; k < count; k += NT) {
      double ar = 0.0, ai = 0.0;
      for (int i = 0; i < k; ++i) {
        double gr, gi;
        gram_at(k, i, gr, gi);
        if (sw == 0 && fabs(gr) + fabs(gi) > 1e-3) s_far = 1;
        const double hr = s_h[cur][i][0], hi = s_h[cur][i][1];
        ar += hr * gr - hi * gi;
        ai += hr * gi + hi * gr;
      }
      s_h[cur ^ 1][k][0] = s_c[k][0] - ar;
      s_h[cur ^ 1][k][1] = s_c[k][1] - ai;
    }
    __syncthreads();
    cur ^= 1;
  }
  if (s_far) {  // far from orthonormal: sequential substitution
    if (tid == 0) {
      for (int k = 0; k < count; ++k) {
        double ar = 0.0, ai = 0.0;
        for (int i = 0; i < k; ++i) {
          double gr, gi;
          gram_at(k, i, gr, gi);
          ar += s_h[cur][i][0] * gr - s_h[cur][i][1] * gi;
          ai += s_h[cur][i][0] * gi + s_h[cur][i][1] * gr;
        }
        s_h[cur][k][0] = s_c[k][0] - ar;
        s_h[cur][k][1] = s_c[k][1] - ai;
      }
    }
    __syncthreads();
  }
  for (int k = tid; k < count; k += NT) {
    const double hr = s_h[cur][k][0], hi = s_h[cur][k][1];
    records[4 * k] = k == 0 ? sj * sj * sums[4 * count] : 0.0;
    records[4 * k + 1] = hr;
    records[4 * k + 2] = hi;
    records[4 * k + 3] = 0.0;
    coef[2 * k] = hr * s_s[k];
    coef[2 * k + 1] = hi * s_s[k];
  }
}

// K3: w' = s_j w - sum_k coef_k u_k (FP64, rounded once on store) and the
// per-workgroup ||w'||^2 of the stored value.
template <typename T>
__global__ __launch_bounds__(kThreads) void arnoldi_gram_update_kernel(uint64_t npk, int count,
                                                                       const GramVecs vecs,
                                                                       void *wv, const double *state,
                                                                       int ld, double *part) {
  using P = typename Packet<T>::type;
  constexpr int A = Packet<T>::amps;
  __shared__ double s_cf[kGramMaxCount][2];
  __shared__ double red[kWaves];
  const double *coef = state + gram_off_coef(ld);
  for (int k = threadIdx.x; k < count; k += kThreads) {
    s_cf[k][0] = coef[2 * k];
    s_cf[k][1] = coef[2 * k + 1];
  }
  const double sj = state[gram_off_misc(ld)];
  __syncthreads();
  P *w = reinterpret_cast<P *>(wv);
  double nrm = 0.0;
  const uint64_t step = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < npk; x += step) {
    const P z = w[x];
    double zr[A], zi[A];
#pragma unroll
    for (int k = 0; k < A; ++k) {
      zr[k] = sj * packet_re(z, k);
      zi[k] = sj * packet_im(z, k);
    }
#pragma unroll 4
    for (int i = 0; i < count; ++i) {
      const P q = reinterpret_cast<const P *>(vecs.u[i])[x];
      const double cr = s_cf[i][0], ci = s_cf[i][1];
#pragma unroll
      for (int k = 0; k < A; ++k) {
        const double qr = packet_re(q, k), qi = packet_im(q, k);
        zr[k] -= cr * qr - ci * qi;
        zi[k] -= cr * qi + ci * qr;
      }
    }
    P o;
#pragma unroll
    for (int k = 0; k < A; ++k) packet_set(o, k, zr[k], zi[k]);
    w[x] = o;
#pragma unroll
    for (int k = 0; k < A; ++k) {
      const double r = packet_re(o, k), m = packet_im(o, k);
      nrm += r * r + m * m;
    }
  }
  nrm = wave_sum(nrm);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nrm;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sm = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) sm += red[q];
    part[blockIdx.x] = sm;
  }
}

// K4 (one workgroup): ||w'||^2 summed in a fixed order into records[0] and
// the next basis vector's scale s_count = 1 / ||w'|| (0 for a zero w').
__global__ __launch_bounds__(kThreads) void arnoldi_gram_norm_kernel(int nb, const double *part,
                                                                     double *rec, double *scale_out) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += kThreads) acc += part[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n2 = red[0];
    rec[0] = n2;
    rec[1] = rec[2] = rec[3] = 0.0;
    *scale_out = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
  }
}

// A 2x2 operator on one bit: thread p owns the pair (x0, x0 | 2^bit).
struct Op1q {
  double a[8];  // A00, A01, A10, A11 as (re, im)
};
template <typename T>
__global__ __launch_bounds__(kThreads) void apply_1q_kernel(uint64_t npairs, void *psiv, int bit,
                                                            const Op1q op) {
  using T2 = typename Vec2<T>::type;
  T2 *psi = reinterpret_cast<T2 *>(psiv);
  const uint64_t lomask = (1ull << bit) - 1;
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t p = (uint64_t)blockIdx.x * kThreads + threadIdx.x; p < npairs; p += stride) {
    const uint64_t x0 = ((p & ~lomask) << 1) | (p & lomask), x1 = x0 | (1ull << bit);
    const T2 u = psi[x0], v = psi[x1];
    const double ur = u.x, ui = u.y, vr = v.x, vi = v.y;
    const double *A = op.a;
    T2 o0, o1;
    o0.x = static_cast<T>(A[0] * ur - A[1] * ui + A[2] * vr - A[3] * vi);
    o0.y = static_cast<T>(A[0] * ui + A[1] * ur + A[2] * vi + A[3] * vr);
    o1.x = static_cast<T>(A[4] * ur - A[5] * ui + A[6] * vr - A[7] * vi);
    o1.y = static_cast<T>(A[4] * ui + A[5] * ur + A[6] * vi + A[7] * vr);
    psi[x0] = o0;
    psi[x1] = o1;
  }
}

// A 4x4 operator on the bit pair (a, b): thread g owns the four amplitudes
// whose bits a and b take every value, local index 2*bit_a + bit_b.
struct Op2q {
  double s[32];  // row-major 4x4, (re, im)
};
__device__ __forceinline__ uint64_t insert_zero(uint64_t x, int bit) {
  const uint64_t lo = x & ((1ull << bit) - 1);
  return ((x >> bit) << (bit + 1)) | lo;
}
template <typename T>
__global__ __launch_bounds__(kThreads) void pair_op_kernel(uint64_t ngroups, const void *inv,
                                                           void *outv, int a, int b,
                                                           const Op2q op, int accumulate) {
  using T2 = typename Vec2<T>::type;
  const T2 *in = reinterpret_cast<const T2 *>(inv);
  T2 *out = reinterpret_cast<T2 *>(outv);
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t g = (uint64_t)blockIdx.x * kThreads + threadIdx.x; g < ngroups; g += stride) {
    const uint64_t x = insert_zero(insert_zero(g, lo), hi);
    uint64_t idx[4];
    idx[0] = x;
    idx[1] = x | (1ull << b);
    idx[2] = x | (1ull << a);
    idx[3] = x | (1ull << a) | (1ull << b);
    double vr[4], vi[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const T2 v = in[idx[k]];
      vr[k] = v.x;
      vi[k] = v.y;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double yr = 0.0, yi = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double sr = op.s[2 * (4 * r + c)], si = op.s[2 * (4 * r + c) + 1];
        yr += sr * vr[c] - si * vi[c];
        yi += sr * vi[c] + si * vr[c];
      }
      if (accumulate) {
        const T2 o = out[idx[r]];
        yr += o.x;
        yi += o.y;
      }
      T2 w;
      w.x = static_cast<T>(yr);
      w.y = static_cast<T>(yi);
      out[idx[r]] = w;
    }
  }
}

// out[x] += sum_j S (row bit j+N, column bit j) of in, for every qubit j of
// an N-qubit density matrix stored as a 2N-bit vector: the gathered form of
// N pair operators with one S, one launch (small matrices, cache resident).
template <typename T>
__global__ __launch_bounds__(kThreads) void pairs_all_kernel(int N, const void *inv, void *outv,
                                                             const Op2q op) {
  using T2 = typename Vec2<T>::type;
  const T2 *in = reinterpret_cast<const T2 *>(inv);
  T2 *out = reinterpret_cast<T2 *>(outv);
  const uint64_t D = 1ull << (2 * N);
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < D; x += stride) {
    double ar = 0.0, ai = 0.0;
    for (int j = 0; j < N; ++j) {
      const int rb = j + N, cb = j;
      const int row = 2 * (int)((x >> rb) & 1) + (int)((x >> cb) & 1);
      const uint64_t base = x & ~((1ull << rb) | (1ull << cb));
#pragma unroll
      for (int ab = 0; ab < 4; ++ab) {
        const uint64_t y = base | ((uint64_t)(ab >> 1) << rb) | ((uint64_t)(ab & 1) << cb);
        const T2 v = in[y];
        const double sr = op.s[2 * (4 * row + ab)], si = op.s[2 * (4 * row + ab) + 1];
        ar += sr * (double)v.x - si * (double)v.y;
        ai += sr * (double)v.y + si * (double)v.x;
      }
    }
    T2 o = out[x];
    o.x = static_cast<T>((double)o.x + ar);
    o.y = static_cast<T>((double)o.y + ai);
    out[x] = o;
  }
}

// out = s (M - M^H) for a 2^k x 2^k row-major matrix M.  Block (bx, by) with
// by <= bx owns the tile pair (I, J) = (by, bx) and (J, I): both tiles are
// staged in LDS before either is written, so out may alias M.
constexpr int kAhTile = 32;
template <typename T>
__global__ __launch_bounds__(kThreads) void antiherm_kernel(int side_log, const void *mv, void *outv,
                                                            double s_re, double s_im) {
  using T2 = typename Vec2<T>::type;
  __shared__ T2 ta[kAhTile][kAhTile + 1];
  __shared__ T2 tb[kAhTile][kAhTile + 1];
  if (blockIdx.y > blockIdx.x) return;
  const uint64_t side = 1ull << side_log;
  const int ts = side < (uint64_t)kAhTile ? (int)side : kAhTile;
  const uint64_t I = (uint64_t)blockIdx.y * ts, J = (uint64_t)blockIdx.x * ts;
  const T2 *m = reinterpret_cast<const T2 *>(mv);
  T2 *out = reinterpret_cast<T2 *>(outv);
  const int tx = threadIdx.x % kAhTile, ty = threadIdx.x / kAhTile;  // 32 x 8
  for (int r = ty; r < ts; r += kThreads / kAhTile) {
    if (tx < ts) {
      ta[r][tx] = m[(I + r) * side + J + tx];  // tile (I, J)
      tb[r][tx] = m[(J + r) * side + I + tx];  // tile (J, I)
    }
  }
  __syncthreads();
  for (int r = ty; r < ts; r += kThreads / kAhTile) {
    if (tx < ts) {
      // out[I+r][J+tx] = s (M[I+r][J+tx] - conj(M[J+tx][I+r]))
      const T2 p = ta[r][tx], q = tb[tx][r];
      double dr = (double)p.x - (double)q.x, di = (double)p.y + (double)q.y;
      T2 w;
      w.x = static_cast<T>(s_re * dr - s_im * di);
      w.y = static_cast<T>(s_re * di + s_im * dr);
      out[(I + r) * side + J + tx] = w;
      // out[J+r][I+tx] = s (M[J+r][I+tx] - conj(M[I+tx][J+r]))
      const T2 p2 = tb[r][tx], q2 = ta[tx][r];
      dr = (double)p2.x - (double)q2.x;
      di = (double)p2.y + (double)q2.y;
      w.x = static_cast<T>(s_re * dr - s_im * di);
      w.y = static_cast<T>(s_re * di + s_im * dr);
      out[(J + r) * side + I + tx] = w;
    }
  }
}

// ---------------------------------------------------------------------------
// Inverse-CDF sampling of basis indices.  p(x) = |src[x]|^2 (mode 0) or
// |src[x * stride]| (mode 1, the diagonal of a density matrix).  The vector
// is cut into chunks of kSampleChunk elements; a wavefront owns a chunk with
// lane l summing its kSampleChunk/64 consecutive elements, and the same
// decomposition is used when a sample descends into a chunk, so the chunk
// sums and the in-chunk search add in one order.
// ---------------------------------------------------------------------------
constexpr int kSampleChunkLog = 12;

template <typename T>
__device__ __forceinline__ double sample_prob(const void *src, uint64_t stride, int mode,
                                              uint64_t x) {
  using T2 = typename Vec2<T>::type;
  const T2 v = reinterpret_cast<const T2 *>(src)[x * stride];
  const double re = v.x, im = v.y;
  return mode == 0 ? re * re + im * im : sqrt(re * re + im * im);
}

// Lane l's element range of chunk c (chunk of min(D, 2^12) elements).
__device__ __forceinline__ void sample_lane_range(uint64_t D, uint64_t c, int lane, uint64_t &x0,
                                                  uint64_t &x1) {
  const uint64_t chunk = D < (1ull << kSampleChunkLog) ? D : (1ull << kSampleChunkLog);
  const uint64_t per = chunk >= 64 ? chunk / 64 : 1;
  const uint64_t base = c * chunk;
  x0 = base + (uint64_t)lane * per;
  x1 = x0 + per;
  if (x0 > base + chunk) x0 = base + chunk;
  if (x1 > base + chunk) x1 = base + chunk;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void sample_chunk_kernel(uint64_t D, uint64_t nchunks,
                                                                const void *src, uint64_t stride,
                                                                int mode, double *sums) {
  const int lane = threadIdx.x & 63;
  const uint64_t c = (uint64_t)blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (c >= nchunks) return;
  uint64_t x0, x1;
  sample_lane_range(D, c, lane, x0, x1);
  double s = 0.0;
  for (uint64_t x = x0; x < x1; ++x) s += sample_prob<T>(src, stride, mode, x);
  s = wave_sum(s);
  if (lane == 0) sums[c] = s;
}

// Inclusive prefix sums of the chunk sums in place (one workgroup; thread t
// owns a run of consecutive chunks).
__global__ __launch_bounds__(kThreads) void sample_scan_kernel(uint64_t nchunks, double *sums) {
  __shared__ double part[kThreads];
  const uint64_t per = (nchunks + kThreads - 1) / kThreads;
  const uint64_t c0 = threadIdx.x * per;
  const uint64_t c1 = c0 + per < nchunks ? c0 + per : nchunks;
  double s = 0.0;
  for (uint64_t c = c0; c < c1; ++c) s += sums[c];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int t = 0; t < kThreads; ++t) {
      const double v = part[t];
      part[t] = acc;  // exclusive offset of thread t's run
      acc += v;
    }
  }
  __syncthreads();
  double acc = part[threadIdx.x];
  for (uint64_t c = c0; c < c1; ++c) {
    acc += sums[c];
    sums[c] = acc;
  }
}

// One wavefront per draw u in [0, 1): the first index x whose cumulative
// probability exceeds u * total (searchsorted(cdf, u, right=True)).
template <typename T>
__global__ __launch_bounds__(kThreads) void sample_draw_kernel(uint64_t D, uint64_t nchunks,
                                                               const void *src, uint64_t stride,
                                                               int mode, const double *ends,
                                                               const double *uniforms, int count,
                                                               int64_t *out) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * kWaves + (threadIdx.x >> 6);
  if (k >= count) return;
  // The reference's test, searchsorted(cdf / total, u, right=True): the first
  // index whose normalised cumulative weight exceeds u.  Normalising each
  // cumulative sum (a correctly rounded division, monotone in the sum)
  // rather than scaling u makes every decision equal to the reference's
  // whenever the cumulative sums themselves are exact.
  const double total = ends[nchunks - 1];
  const double u = uniforms[k];
  // first chunk whose normalised inclusive end exceeds u
  uint64_t lo = 0, hi = nchunks;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (ends[mid] / total > u) hi = mid;
    else lo = mid + 1;
  }
  if (lo >= nchunks) lo = nchunks - 1;
  const uint64_t c = lo;
  const double base = c > 0 ? ends[c - 1] : 0.0;
  uint64_t x0, x1;
  sample_lane_range(D, c, lane, x0, x1);
  double s = 0.0;
  for (uint64_t x = x0; x < x1; ++x) s += sample_prob<T>(src, stride, mode, x);
  // inclusive prefix over lanes
  double pre = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double o = __shfl_up(pre, off, 64);
    if (lane >= off) pre += o;
  }
  const unsigned long long hit = __ballot((base + pre) / total > u);
  int64_t found = -1;
  if (hit) {
    const int l = __ffsll((long long)hit) - 1;
    if (lane == l) {
      double run = base + (pre - s);
      for (uint64_t x = x0; x < x1; ++x) {
        run += sample_prob<T>(src, stride, mode, x);
        if (run / total > u) {
          found = (int64_t)x;
          break;
        }
      }
      if (found < 0) found = (int64_t)(x1 > x0 ? x1 - 1 : x0);
    }
    found = __shfl((long long)found, l, 64);
  } else {
    // rounding kept every normalised sum of the chunk at or below u: the
    // chunk's last element with non-zero probability
    uint64_t cx0, cx1;
    sample_lane_range(D, c, 0, cx0, cx1);
    const uint64_t chunk = D < (1ull << kSampleChunkLog) ? D : (1ull << kSampleChunkLog);
    int64_t last = (int64_t)(c * chunk + chunk - 1);
    for (int64_t x = last; x >= (int64_t)(c * chunk); --x)
      if (sample_prob<T>(src, stride, mode, (uint64_t)x) > 0.0) {
        last = x;
        break;
      }
    found = last;
  }
  if (lane == 0) out[k] = found < (int64_t)D ? found : (int64_t)D - 1;
}

template <typename S, typename D2>
__global__ __launch_bounds__(kThreads) void convert_kernel(uint64_t D,
                                                           const typename Vec2<S>::type *src,
                                                           typename Vec2<D2>::type *dst) {
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < D; x += stride) {
    const auto z = src[x];
    typename Vec2<D2>::type o;
    o.x = static_cast<D2>(z.x);
    o.y = static_cast<D2>(z.y);
    dst[x] = o;
  }
}

// Same-type copy in 16-byte packets: each thread issues its four loads
// (a block-stride apart, so every load instruction is coalesced) before its
// four stores; one pass, no grid-stride loop.  npk is a multiple of
// 4 * kThreads or the tail is guarded.
__global__ __launch_bounds__(kThreads) void copy_packets_kernel(uint64_t npk, const uint4 *src,
                                                                uint4 *dst) {
  const uint64_t base = (uint64_t)blockIdx.x * (4 * kThreads) + threadIdx.x;
  uint4 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint64_t i = base + (uint64_t)k * kThreads;
    if (i < npk) v[k] = src[i];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint64_t i = base + (uint64_t)k * kThreads;
    if (i < npk) dst[i] = v[k];
  }
}

constexpr int kLincombBatch = 112;  // vectors per launch: 2.7 KiB of kernel arguments, so a
                                    // whole Krylov basis (<= 101 vectors) is one launch
constexpr int kMaxLincomb = 1024;
struct LincombArgs {
  const void *vec[kLincombBatch];
  double cr[kLincombBatch];
  double ci[kLincombBatch];
  int count;
  int accumulate;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void lincomb_kernel(uint64_t npk, const LincombArgs args,
                                                           void *outv) {
  using P = typename Packet<T>::type;
  constexpr int A = Packet<T>::amps;
  P *out = reinterpret_cast<P *>(outv);
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < npk; x += stride) {
    double zr[A], zi[A];
    P o;
    if (args.accumulate) o = out[x];
#pragma unroll
    for (int k = 0; k < A; ++k) {
      zr[k] = args.accumulate ? packet_re(o, k) : 0.0;
      zi[k] = args.accumulate ? packet_im(o, k) : 0.0;
    }
    for (int j = 0; j < args.count; ++j) {
      const P p = reinterpret_cast<const P *>(args.vec[j])[x];
      const double cr = args.cr[j], ci = args.ci[j];
#pragma unroll
      for (int k = 0; k < A; ++k) {
        const double pr = packet_re(p, k), pi = packet_im(p, k);
        zr[k] += cr * pr - ci * pi;
        zi[k] += cr * pi + ci * pr;
      }
    }
#pragma unroll
    for (int k = 0; k < A; ++k) packet_set(o, k, zr[k], zi[k]);
    out[x] = o;
  }
}

// Correlation <n_b n_b'> over bit positions.  Each block walks chunks of 2^K
// consecutive amplitudes (K = min(n, 14)); thread tid owns the amplitudes
// chunk + tid + 256 r.  Per thread: P (sum |psi|^2), P_j (r bit j set),
// P_jj' (r bits j, j' set).  The block folds these into the chunk-local
// K x K matrix and then into an n x n partial held in LDS.
constexpr int kCorrRBits = 6;
constexpr int kCorrVals = 1 + kCorrRBits + kCorrRBits * (kCorrRBits - 1) / 2;  // 22

template <typename T>
__global__ __launch_bounds__(kThreads) void correlation_kernel(int n, const typename Vec2<T>::type *psi,
                                                               double *partials) {
  using T2 = typename Vec2<T>::type;
  __shared__ double vals[kCorrVals][kThreads];
  __shared__ double loc[kMaxTileLog][kMaxTileLog];
  __shared__ double ltot;
  __shared__ double acc[EMU_SV_MAX_QUBITS * EMU_SV_MAX_QUBITS];
  const int tid = threadIdx.x;
  const int K = n < kMaxTileLog ? n : kMaxTileLog;
  const int tb = K < kThreadBits ? K : kThreadBits;
  const int rb = K - tb;
  const uint64_t nchunks = 1ull << (n - K);
  for (int e = tid; e < n * n; e += kThreads) acc[e] = 0.0;
  for (uint64_t ch = blockIdx.x; ch < nchunks; ch += gridDim.x) {
    const uint64_t cbase = ch << K;
    double P = 0.0, Pj[kCorrRBits], Pjj[kCorrVals - 1 - kCorrRBits];
#pragma unroll
    for (int j = 0; j < kCorrRBits; ++j) Pj[j] = 0.0;
#pragma unroll
    for (int j = 0; j < kCorrVals - 1 - kCorrRBits; ++j) Pjj[j] = 0.0;
    if (tid < (1 << tb)) {
      for (uint32_t r = 0; r < (1u << rb); ++r) {
        const T2 z = psi[cbase + tid + ((uint64_t)r << tb)];
        const double p = (double)z.x * z.x + (double)z.y * z.y;
        P += p;
        int idx = 0;
#pragma unroll
        for (int j = 0; j < kCorrRBits; ++j) {
          if ((r >> j) & 1u) Pj[j] += p;
#pragma unroll
          for (int j2 = 0; j2 < j; ++j2) {
            if (((r >> j) & 1u) && ((r >> j2) & 1u)) Pjj[idx] += p;
            ++idx;
          }
        }
      }
    }
    __syncthreads();  // previous chunk's readers of vals are done
    vals[0][tid] = P;
#pragma unroll
    for (int j = 0; j < kCorrRBits; ++j) vals[1 + j][tid] = Pj[j];
#pragma unroll
    for (int j = 0; j < kCorrVals - 1 - kCorrRBits; ++j) vals[1 + kCorrRBits + j][tid] = Pjj[j];
    __syncthreads();
    // chunk-local K x K (upper triangle incl. diagonal) and total
    const int nent = K * (K + 1) / 2;
    for (int e = tid; e <= nent; e += kThreads) {
      if (e == nent) {
        double sm = 0.0;
        for (int u = 0; u < (1 << tb); ++u) sm += vals[0][u];
        ltot = sm;
        continue;
      }
      int i = 0, rem = e;
      while (rem >= K - i) {
        rem -= K - i;
        ++i;
      }
      const int j = i + rem;  // i <= j
      double sm = 0.0;
      if (j < tb) {
        for (int u = 0; u < (1 << tb); ++u)
          if (((u >> i) & 1) && ((u >> j) & 1)) sm += vals[0][u];
      } else if (i < tb) {
        const int jr = j - tb;
        for (int u = 0; u < (1 << tb); ++u)
          if ((u >> i) & 1) sm += vals[1 + jr][u];
      } else if (i == j) {
        const int jr = j - tb;
        for (int u = 0; u < (1 << tb); ++u) sm += vals[1 + jr][u];
      } else {
        const int ir = i - tb, jr = j - tb;  // ir < jr
        const int idx = jr * (jr - 1) / 2 + ir;
        for (int u = 0; u < (1 << tb); ++u) sm += vals[1 + kCorrRBits + idx][u];
      }
      loc[i][j] = sm;
      loc[j][i] = sm;
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += kThreads) {
      const int b = e / n, b2 = e % n;
      const bool ob = b >= K, ob2 = b2 >= K;
      double add;
      if (ob && ob2)
        add = (((cbase >> b) & 1) && ((cbase >> b2) & 1)) ? ltot : 0.0;
      else if (ob)
        add = ((cbase >> b) & 1) ? loc[b2][b2] : 0.0;
      else if (ob2)
        add = ((cbase >> b2) & 1) ? loc[b][b] : 0.0;
      else
        add = loc[b][b2];
      acc[e] += add;
    }
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += kThreads)
    partials[(uint64_t)blockIdx.x * n * n + e] = acc[e];
}

// Reorder the summed bit-indexed matrix into qubit order.
__global__ void corr_finalize_kernel(int n, const double *partials, int nblk, double *out) {
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    double sm = 0.0;
    for (int b = 0; b < nblk; ++b) sm += partials[(uint64_t)b * n * n + e];
    const int bi = e / n, bj = e % n;
    out[(n - 1 - bi) * n + (n - 1 - bj)] = sm;
  }
}


// ---------------------------------------------------------------------------
// Parameter traces: per-bit sums of conj(W_l) Z_l over flipped / masked
// index sets (the per-qubit traces of every Hamiltonian-parameter gradient).
// Grid (gx, n): block row b owns bit b; its blocks stride over chunks of
// kTraceChunk elements.  The flip partner Z_l(x ^ 2^b) is a second read of
// Z, served from L2/MALL since the n block rows of a chunk run together.
// ---------------------------------------------------------------------------
constexpr int kTraceChunk = 1024;
constexpr int kTraceBlocks = 2048;

template <int K>
__device__ __forceinline__ void block_reduce_partial(double (&v)[K], int nk, double *red, double *out) {
  // red: kWaves*K doubles of LDS; sums the first nk entries of v over the block
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k < nk) {
      const double sm = wave_sum(v[k]);
      if (lane == 0) red[wave * K + k] = sm;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nk; k += kThreads) {
    double sm = 0.0;
    for (int w = 0; w < kWaves; ++w) sm += red[w * K + k];
    out[k] = sm;
  }
}

template <typename T, bool PAIRS>
__global__ __launch_bounds__(kThreads) void traces_kernel(int n, int count, const typename Vec2<T>::type *W,
                                                          const typename Vec2<T>::type *Z, uint64_t stride,
                                                          double *partials) {
  using T2 = typename Vec2<T>::type;
  constexpr int K = PAIRS ? 6 + 2 * EMU_SV_MAX_QUBITS : 6;
  __shared__ double red[kWaves * K];
  const int b = blockIdx.y;
  const uint64_t D = 1ull << n;
  const uint64_t ch = D < (uint64_t)kTraceChunk ? D : (uint64_t)kTraceChunk;
  const uint64_t nchunks = D / ch;
  const uint64_t fb = 1ull << b;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    for (uint64_t off = threadIdx.x; off < ch; off += kThreads) {
      const uint64_t x = c * ch + off;
      const bool set = (x >> b) & 1;
      for (int l = 0; l < count; ++l) {
        const uint64_t row = (uint64_t)l * stride;
        const T2 w = W[row + x];
        const T2 zz = Z[row + x];
        const T2 zp = Z[row + (x ^ fb)];
        const double wr = w.x, wi = w.y;
        // conj(w) * zp
        const double fr = wr * (double)zp.x + wi * (double)zp.y;
        const double fi = wr * (double)zp.y - wi * (double)zp.x;
        if (set) {
          const double pr = wr * (double)zz.x + wi * (double)zz.y;
          const double pi = wr * (double)zz.y - wi * (double)zz.x;
          acc[0] += fr;
          acc[1] += fi;
          acc[4] += pr;
          acc[5] += pi;
          if (PAIRS) {
#pragma unroll
            for (int b2 = 0; b2 < EMU_SV_MAX_QUBITS; ++b2) {
              if (b2 > b && b2 < n && ((x >> b2) & 1)) {
                acc[6 + 2 * b2] += pr;
                acc[7 + 2 * b2] += pi;
              }
            }
          }
        } else {
          acc[2] += fr;
          acc[3] += fi;
        }
      }
    }
  }
  const int nk = PAIRS ? 6 + 2 * n : 6;
  block_reduce_partial<K>(acc, nk, red,
                          partials + ((uint64_t)b * gridDim.x + blockIdx.x) * (6 + 2 * EMU_SV_MAX_QUBITS));
}

// out: up[n], dn[n], occ[n], pair[n*n] complex; partials [b][gx][6 + 2*32].
__global__ __launch_bounds__(kThreads) void traces_finalize_kernel(int n, int gx, int pairs,
                                                                   const double *partials, double *out) {
  const int per = 6 + 2 * EMU_SV_MAX_QUBITS;
  const int nout = pairs ? 6 * n + 2 * n * n : 6 * n;
  for (int e = blockIdx.x * kThreads + threadIdx.x; e < nout; e += gridDim.x * kThreads) {
    int b, k;
    if (e < 6 * n) {
      const int q = e / (2 * n);          // 0 up, 1 dn, 2 occ
      const int r = e % (2 * n);
      b = r / 2;
      k = 2 * q + (r & 1);
    } else {
      const int r = e - 6 * n;            // pair: (b * n + b2) * 2 + re/im
      b = r / (2 * n);
      const int b2 = (r / 2) % n;
      if (b2 <= b) {
        out[e] = 0.0;
        continue;
      }
      k = 6 + 2 * b2 + (r & 1);
    }
    double sm = 0.0;
    for (int g = 0; g < gx; ++g) sm += partials[((uint64_t)b * gx + g) * per + k];
    out[e] = sm;
  }
}

#include "emu_sv_tile.inc"
#include "emu_sv_btile.inc"
#include "emu_sv_batch.inc"

// ---------------------------------------------------------------------------
// Planner
// ---------------------------------------------------------------------------
struct Pass {
  int c, s, m;
  uint32_t fmask;
  int diag;
};

uint32_t bits_mask(int lo, int hi) {  // bits [lo, hi)
  uint64_t m = ((1ull << hi) - 1) ^ ((1ull << lo) - 1);
  return (uint32_t)m;
}

int make_plan(int n, int tl, int cmin, Pass *out) {
  if (n <= tl) {
    out[0] = {n, n, 0, bits_mask(0, n), 1};
    return 1;
  }
  int np = 0;
  out[np++] = {tl, tl, 0, bits_mask(0, tl), 1};
  int next = tl;
  while (next < n) {
    const int rem = n - next;
    int m = rem < tl - cmin ? rem : tl - cmin;
    int c = tl - m;
    if (c > next) c = next;
    out[np++] = {c, next, m, bits_mask(next, next + m), 0};
    next += m;
  }
  return np;
}

int kernel_kind(int flags) {
  const int k = (flags >> 8) & 0xF;
  return k == 0 ? kKernelWave : k;
}

void default_tiles(int dtype, int kind, int &tl, int &cmin) {
  if (tl <= 0) {
    if (kind == kKernelWave)
      tl = dtype == EMU_SV_C64 ? kDefaultWaveLogC64 : kDefaultWaveLogC128;
    else
      tl = dtype == EMU_SV_C64 ? kDefaultTileLogC64 : kDefaultTileLogC128;
  }
  if (cmin <= 0) cmin = dtype == EMU_SV_C64 ? kDefaultContigC64 : kDefaultContigC128;
}

template <typename T, bool CPLX, bool DOTS, bool NOISE>
int set_lds_attr(size_t lds) {
  // Dynamic LDS above 64 KiB must be enabled per kernel.
  static std::mutex mu;
  static size_t enabled = 65536;
  std::lock_guard<std::mutex> lock(mu);
  if (lds <= enabled) return EMU_SV_OK;
  hipError_t err =
      hipFuncSetAttribute(reinterpret_cast<const void *>(&hpsi_pass_kernel<T, CPLX, DOTS, NOISE>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err == hipSuccess) enabled = lds;
  return hip_check(err, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

template <typename T, bool CPLX, bool DOTS, bool NOISE = false>
int launch_pass(const PassArgs &a, hipStream_t st) {
  const int t = a.c + a.m;
  const uint64_t nt = 1ull << (a.n - t);
  const size_t lds = sizeof(typename Vec2<T>::type) << t;
  int rc = set_lds_attr<T, CPLX, DOTS, NOISE>(lds);
  if (rc) return rc;
  hipLaunchKernelGGL((hpsi_pass_kernel<T, CPLX, DOTS, NOISE>), dim3((unsigned)nt), dim3(kThreads),
                     lds, st, a);
  return hip_check(hipGetLastError(), "hpsi_pass_kernel launch");
}

double *partials2_ptr(void *ws, int n);
double *partials_ptr(void *ws, int n);
int finalize(const double *rows, int nrow, int K, void *ws, int n, double *out, hipStream_t st);
// Outputs of the next-iteration coefficients formed by the last reduction
// of a pipelined Lanczos iteration (next == nullptr: none).
struct NextCoefOut {
  const double *prev = nullptr;  // the coefficient record this iteration read
  double *next = nullptr, *tout = nullptr;
  double sig_re = 0.0, sig_im = 0.0;
};
int finalize_pair(const double *rows_a, int nrow_a, int ka, double *out_a, const double *rows_b,
                  int nrow_b, int kb, double *out_b, void *ws, int n, const NextCoefOut &nc,
                  hipStream_t st);
constexpr int kFinFoldMaxRows = 8192;  // rows of the folded one-block reduction
int finalize_folded(const double *rows, int nrow, double *out_a, double *out_b,
                    const NextCoefOut &nc, hipStream_t st);
const double *dtab_ptr(void *ws, int n);

template <typename T, int KREG, bool CPLX, bool DOTS>
int launch_wave(const PassArgs &a, hipStream_t st) {
  const uint64_t nt = 1ull << (a.n - (6 + KREG));
  const uint64_t nblk = nt >= (uint64_t)kWaves ? nt / kWaves : 1;
  hipLaunchKernelGGL((hpsi_wave_kernel<T, KREG, CPLX, DOTS>), dim3((unsigned)nblk), dim3(kThreads),
                     0, st, a);
  return hip_check(hipGetLastError(), "hpsi_wave_kernel launch");
}

template <typename T, int KREG>
int launch_wave_any(const PassArgs &a, bool cplx, bool dots, hipStream_t st) {
  if (cplx)
    return dots ? launch_wave<T, KREG, true, true>(a, st) : launch_wave<T, KREG, true, false>(a, st);
  return dots ? launch_wave<T, KREG, false, true>(a, st) : launch_wave<T, KREG, false, false>(a, st);
}

struct OrthoIn {
  const void *pa, *pb;
  double ca_re, ca_im, cb_re, cb_im, pb_scale;
  double *norm2_out;  // kOrthoSums doubles
  const double *dcoef = nullptr;  // device coefficients (LDS and tile kernels)
  NextCoefOut nc;                 // next coefficients formed on the device (optional)
};

// The diagonal table of the first pass lives in the workspace.  Its rows
// carry a stamp of the parameter block's pack generation, n and the
// first-pass tile log; with keep (EMU_SV_FLAG_KEEP_DIAG) the prep kernel
// leaves rows whose stamp matches and rebuilds the others, so a table is
// never reused for a different block, size or tile, whatever address the
// block has.  Without keep every row is rebuilt.
// Set by emu_sv_arnoldi_exp around its matvecs after the first, and by
// emu_sv_krylov_step around its iterations: the table was built by the
// step's first matvec, on this workspace and stream, with nothing in
// between, so the stamp check (a launch on the iteration's critical path)
// is skipped.  Calling-thread state, never set across a return to the caller.
thread_local bool t_diag_known_fresh = false;

int prepare_diag(int n, int t1, const double *ham, void *ws, bool keep, hipStream_t st) {
  if (keep && t_diag_known_fresh) return EMU_SV_OK;
  const uint64_t ntiles = 1ull << (n - t1);
  hipLaunchKernelGGL(diag_prep_kernel, dim3((unsigned)((ntiles * kDiagRow + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, n, t1, ham, const_cast<double *>(dtab_ptr(ws, n)),
                     keep ? 1 : 0);
  return hip_check(hipGetLastError(), "diag_prep launch");
}

template <typename T>
int run_hpsi(int n, const double *ham, double sr, double si, const void *v, double vsc,
             const void *prev, double psc, void *w, double *dots, void *ws, int kind, int tl,
             int cmin, bool cplx, hipStream_t st, const OrthoIn *ortho = nullptr,
             bool noise = false, bool keep_diag = false) {
  if (kind == kKernelWave && n < tl) kind = kKernelLds;  // one small tile: LDS kernel
  if (noise && kind != kKernelLds) return fail(EMU_SV_EINVAL, "hpsi: the noise term needs the LDS kernel");
  if (ortho && ortho->dcoef && (kind == kKernelWave || noise))
    return fail(EMU_SV_EINVAL, "hpsi: device coefficients need the LDS or tile kernel");
  Pass plan[EMU_SV_MAX_QUBITS];
  const int np = make_plan(n, tl, cmin, plan);
  // pass 1 is the contiguous tile [0, t1)
  if (int rc = prepare_diag(n, plan[0].c + plan[0].m, ham, ws, keep_diag, st)) return rc;
  int nrow_a = 0;
  for (int p = 0; p < np; ++p) {
    PassArgs a{};
    a.v = v;
    a.prev = prev;
    a.out = w;
    a.ham = ham;
    a.partials = partials_ptr(ws, n);
    a.dtab = dtab_ptr(ws, n);
    a.ortho = (ortho != nullptr && p == 0) ? 1 : 0;
    a.zout = const_cast<void *>(v);
    a.pa = ortho ? ortho->pa : nullptr;
    a.pb = ortho ? ortho->pb : nullptr;
    a.ca_re = ortho ? ortho->ca_re : 0.0;
    a.ca_im = ortho ? ortho->ca_im : 0.0;
    a.cb_re = ortho ? ortho->cb_re : 0.0;
    a.cb_im = ortho ? ortho->cb_im : 0.0;
    a.pb_scale = ortho ? ortho->pb_scale : 0.0;
    a.partials2 = partials2_ptr(ws, n);
    a.dcoef = (ortho != nullptr && p == 0) ? ortho->dcoef : nullptr;
    a.v_scale = (ortho != nullptr && p > 0) ? 1.0 : vsc;
    a.prev_scale = psc;
    a.sig_re = sr;
    a.sig_im = si;
    a.n = n;
    a.c = plan[p].c;
    a.s = plan[p].s;
    a.m = plan[p].m;
    a.flip_mask = plan[p].fmask;
    a.diag = plan[p].diag;
    a.accumulate = p > 0;
    a.xcd = 1;
    const bool last = p == np - 1;
    const bool fuse = last && dots != nullptr;
    int rc;
    if (noise) {
      rc = fuse ? launch_pass<T, true, true, true>(a, st) : launch_pass<T, true, false, true>(a, st);
    } else if (kind == kKernelWave) {
      switch (tl - 6) {
        case 3: rc = launch_wave_any<T, 3>(a, cplx, fuse, st); break;
        case 4: rc = launch_wave_any<T, 4>(a, cplx, fuse, st); break;
        case 5: rc = launch_wave_any<T, 5>(a, cplx, fuse, st); break;
        default: return fail(EMU_SV_EINVAL, "hpsi: wave tiles need tile_log in [9, 11]");
      }
    } else if (cplx) {
      rc = fuse ? launch_pass<T, true, true>(a, st) : launch_pass<T, true, false>(a, st);
    } else {
      rc = fuse ? launch_pass<T, false, true>(a, st) : launch_pass<T, false, false>(a, st);
    }
    if (rc) return rc;
    const int t = a.c + a.m;
    const int nrow = (int)(1ull << (n - t));
    // a Lanczos iteration reduces its first-pass sums together with the last
    // pass's dots (one launch); the first-pass rows stay in place until then
    if (a.ortho) {
      if (dots) {
        nrow_a = nrow;
      } else {
        rc = finalize(a.partials2, nrow, kOrthoSums, ws, n, ortho->norm2_out, st);
        if (rc) return rc;
      }
    }
    if (fuse) {
      rc = ortho ? finalize_pair(a.partials2, nrow_a, kOrthoSums, ortho->norm2_out, a.partials, nrow,
                                 7, dots, ws, n, ortho->nc, st)
                 : finalize(a.partials, nrow, 7, ws, n, dots, st);
      if (rc) return rc;
    }
  }
  return EMU_SV_OK;
}


// ---------------------------------------------------------------------------
// Tile-kernel passes: A = contiguous [0, TL) (+ diagonal, + orthogonalisation),
// then B passes of c + m = TL tile bits; the last B pass takes up to
// max_partner leftover top bits as partner reads.
// ---------------------------------------------------------------------------
struct TilePass {
  int tl, c, s, m;
  uint32_t fmask;
  int pbits;
  bool modeb;
};

// Partner counts with a compiled kernel variant.
bool partner_count_ok(int r) { return r == 0 || r == 2 || r == 4 || r == 6; }

// A: [0, tla); B passes of tlb tile bits with >= cmin contiguous low bits;
// the last B pass may read up to max_partner leftover top bits directly.
// Requires n >= max(tla, tlb).
int make_tile_plan(int n, int tla, int tlb, int cmin, int max_partner, TilePass *out) {
  int np = 0;
  out[np++] = {tla, tla, tla, 0, bits_mask(0, tla), 0, false};
  int next = tla;
  while (next < n) {
    const int rem = n - next;
    int m = rem < tlb - cmin ? rem : tlb - cmin;
    if (tlb - m > next) m = tlb - next;  // keep c <= s (small leftovers)
    const int left = rem - m;
    if (left > 0) {
      int r = left;
      while (r <= max_partner && !partner_count_ok(r)) ++r;
      if (r <= max_partner && r <= rem && rem - r >= 1 && tlb - (rem - r) <= next) {
        const int mm = rem - r;
        out[np++] = {tlb, tlb - mm, next, mm, bits_mask(next, next + mm), r, true};
        break;
      }
    }
    out[np++] = {tlb, tlb - m, next, m, bits_mask(next, next + m), 0, true};
    next += m;
  }
  return np;
}

int tile_max_log(int dtype) { return dtype == EMU_SV_C64 ? 14 : 13; }

template <typename T, int TL, int NT, bool CPLX, bool MODEB, int PR>
int launch_tile_one(const TileArgs &a, hipStream_t st) {
  static std::once_flag once;
  static hipError_t attr = hipSuccess;
  const size_t lds = sizeof(typename Vec2<T>::type) << TL;
  if (lds > 65536)
    std::call_once(once, [&] {
      attr = hipFuncSetAttribute(
          reinterpret_cast<const void *>(&tile_kernel<T, TL, NT, CPLX, MODEB, PR>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    });
  if (int rc = hip_check(attr, "hipFuncSetAttribute(tile kernel LDS)")) return rc;
  const uint64_t nt = 1ull << (a.n - TL);
  hipLaunchKernelGGL((tile_kernel<T, TL, NT, CPLX, MODEB, PR>), dim3((unsigned)nt), dim3(NT), lds,
                     st, a);
  return hip_check(hipGetLastError(), "tile kernel launch");
}

template <typename T, int TL, bool CPLX>
int launch_tile_b(const TileArgs &a, hipStream_t st) {
  constexpr int NT = TL >= 13 ? 512 : 256;
  switch (a.partner_bits) {
    case 0: return launch_tile_one<T, TL, NT, CPLX, true, 0>(a, st);
    case 2: return launch_tile_one<T, TL, NT, CPLX, true, 2>(a, st);
    case 4: return launch_tile_one<T, TL, NT, CPLX, true, 4>(a, st);
    case 6: return launch_tile_one<T, TL, NT, CPLX, true, 6>(a, st);
    default: return fail(EMU_SV_EINVAL, "tile kernel: unsupported partner count");
  }
}

template <typename T, int TL>
int launch_tile_tl(const TileArgs &a, bool cplx, bool modeb, hipStream_t st) {
  constexpr int NT = TL >= 13 ? 512 : 256;
  if (modeb) return cplx ? launch_tile_b<T, TL, true>(a, st) : launch_tile_b<T, TL, false>(a, st);
  return cplx ? launch_tile_one<T, TL, NT, true, false, 0>(a, st)
              : launch_tile_one<T, TL, NT, false, false, 0>(a, st);
}

template <typename T>
int launch_tile(const TileArgs &a, int tl, bool cplx, bool modeb, hipStream_t st) {
  switch (tl) {
    case 10: return launch_tile_tl<T, 10>(a, cplx, modeb, st);
    case 11: return launch_tile_tl<T, 11>(a, cplx, modeb, st);
    case 12: return launch_tile_tl<T, 12>(a, cplx, modeb, st);
    case 13: return launch_tile_tl<T, 13>(a, cplx, modeb, st);
    case 14:
      if constexpr (sizeof(T) == 4) return launch_tile_tl<T, 14>(a, cplx, modeb, st);
      break;
    default: break;
  }
  return fail(EMU_SV_EINVAL, "tile kernel: tile_log out of range");
}

template <typename T>
int run_tile(int n, const double *ham, double sr, double si, const void *v, double vsc,
             const void *prev, double psc, void *w, double *dots, void *ws, int tla, int tlb,
             int cmin, int max_partner, bool cplx, hipStream_t st, const OrthoIn *ortho,
             bool keep_diag = false) {
  TilePass plan[EMU_SV_MAX_QUBITS];
  const int np = make_tile_plan(n, tla, tlb, cmin, max_partner, plan);
  if (int rc = prepare_diag(n, tla, ham, ws, keep_diag, st)) return rc;
  int nrow_a = 0;
  for (int p = 0; p < np; ++p) {
    const bool last = p == np - 1;
    TileArgs a{};
    a.v = v;  // with orthogonalisation, pass A has overwritten v with z
    a.acc = w;
    a.out = w;
    a.prev = last ? prev : nullptr;
    a.ham = ham;
    a.dtab = dtab_ptr(ws, n);
    a.partials = partials_ptr(ws, n);
    a.partials2 = partials2_ptr(ws, n);
    a.ortho = (ortho != nullptr && p == 0) ? 1 : 0;
    a.zout = const_cast<void *>(v);
    a.pa = ortho ? ortho->pa : nullptr;
    a.pb = ortho ? ortho->pb : nullptr;
    a.ca_re = ortho ? ortho->ca_re : 0.0;
    a.ca_im = ortho ? ortho->ca_im : 0.0;
    a.cb_re = ortho ? ortho->cb_re : 0.0;
    a.cb_im = ortho ? ortho->cb_im : 0.0;
    a.pb_scale = ortho ? ortho->pb_scale : 0.0;
    a.dcoef = (ortho != nullptr && p == 0) ? ortho->dcoef : nullptr;
    a.v_scale = (ortho != nullptr && p > 0) ? 1.0 : vsc;  // later passes read z = v_scale*w already
    a.prev_scale = psc;
    a.sig_re = sr;
    a.sig_im = si;
    a.n = n;
    a.c = plan[p].c;
    a.s = plan[p].s;
    a.m = plan[p].m;
    a.flip_mask = plan[p].fmask;
    a.partner_bits = plan[p].pbits;
    a.dots = (last && dots != nullptr) ? 1 : 0;
    int rc = launch_tile<T>(a, plan[p].tl, cplx, plan[p].modeb, st);
    if (rc) return rc;
    const int nrow = (int)(1ull << (n - plan[p].tl));
    if (a.ortho) {
      if (dots) {
        nrow_a = nrow;  // reduced with the last pass's dots
      } else {
        rc = finalize(a.partials2, nrow, kOrthoSums, ws, n, ortho->norm2_out, st);
        if (rc) return rc;
      }
    }
    if (a.dots) {
      rc = ortho ? finalize_pair(a.partials2, nrow_a, kOrthoSums, ortho->norm2_out, a.partials, nrow,
                                 7, dots, ws, n, ortho->nc, st)
                 : finalize(a.partials, nrow, 7, ws, n, dots, st);
      if (rc) return rc;
    }
  }
  return EMU_SV_OK;
}


// ---------------------------------------------------------------------------
// Big-tile passes (btile_kernel): A = contiguous [0, TB) with the diagonal
// (and the orthogonalisation), then B passes of c + m = TB tile bits with
// c >= cmin contiguous low bits; at n <= 2 TB - cmin one B pass covers the
// rest.  TB = 14 (c128) or 15 (c64); needs n > TB.
// ---------------------------------------------------------------------------
int bt_tile_bits_of(int dtype) { return dtype == EMU_SV_C64 ? 15 : 14; }

int make_btile_plan(int n, int tb, int cmin, Pass *out) {
  int np = 0;
  out[np++] = {tb, tb, 0, bits_mask(0, tb), 1};
  int next = tb;
  while (next < n) {
    const int rem = n - next;
    const int m = rem < tb - cmin ? rem : tb - cmin;
    out[np++] = {tb - m, next, m, bits_mask(next, next + m), 0};
    next += m;
  }
  return np;
}

// 0: valid; 1: n not above one tile (use the LDS kernel); < 0: error.
int check_btile_schedule(int n, int dtype, int &cmin, const char *what) {
  const int tb = bt_tile_bits_of(dtype);
  if (cmin <= 0) cmin = dtype == EMU_SV_C64 ? 4 : 2;
  if (cmin < 1 || cmin >= tb)
    return fail(EMU_SV_EINVAL, std::string(what) + ": btile min_contig must be in [1, " +
                                   std::to_string(tb) + ")");
  return n <= tb ? 1 : 0;
}

template <typename T, bool CPLX, bool MODEB, bool DOTS, bool NOISE>
int launch_btile_one(const BtArgs &a, hipStream_t st) {
  static std::once_flag once;
  static hipError_t attr = hipSuccess;
  const size_t lds = (size_t)kBtRoundSlots * kBtThreads * 16;
  std::call_once(once, [&] {
    attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&btile_kernel<T, CPLX, MODEB, DOTS, NOISE>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  });
  if (int rc = hip_check(attr, "hipFuncSetAttribute(btile kernel LDS)")) return rc;
  const uint64_t nt = 1ull << (a.n - bt_tile_bits<T>());
  hipLaunchKernelGGL((btile_kernel<T, CPLX, MODEB, DOTS, NOISE>), dim3((unsigned)nt),
                     dim3(kBtThreads), lds, st, a);
  return hip_check(hipGetLastError(), "btile kernel launch");
}

template <typename T, bool CPLX, bool NOISE>
int launch_btile_c(const BtArgs &a, bool modeb, bool dots, hipStream_t st) {
  if (!modeb) return launch_btile_one<T, CPLX, false, false, NOISE>(a, st);
  return dots ? launch_btile_one<T, CPLX, true, true, NOISE>(a, st)
              : launch_btile_one<T, CPLX, true, false, NOISE>(a, st);
}

template <typename T>
int launch_btile(const BtArgs &a, bool cplx, bool noise, bool modeb, bool dots, hipStream_t st) {
  if (noise) return launch_btile_c<T, true, true>(a, modeb, dots, st);
  return cplx ? launch_btile_c<T, true, false>(a, modeb, dots, st)
              : launch_btile_c<T, false, false>(a, modeb, dots, st);
}

// noise: the Monte-Carlo term K of the parameter block (bare H.psi only:
// no orthogonalisation, which needs a Hermitian H)
template <typename T>
int run_btile(int n, const double *ham, double sr, double si, const void *v, double vsc, void *w,
              double *dots, void *ws, int cmin, bool cplx, hipStream_t st, const OrthoIn *ortho,
              bool noise = false) {
  if (noise && ortho)
    return fail(EMU_SV_EINVAL, "btile: the noise term is for bare H.psi (no orthogonalisation)");
  constexpr int TB = bt_tile_bits<T>();
  Pass plan[EMU_SV_MAX_QUBITS];
  const int np = make_btile_plan(n, TB, cmin, plan);
  int nrow_a = 0;
  for (int p = 0; p < np; ++p) {
    const bool last = p == np - 1;
    BtArgs a{};
    a.v = v;  // with orthogonalisation, pass A has overwritten v with z
    a.acc = w;
    a.out = w;
    a.ham = ham;
    a.partials = partials_ptr(ws, n);
    a.partials2 = partials2_ptr(ws, n);
    a.ortho = (ortho != nullptr && p == 0) ? 1 : 0;
    a.zout = const_cast<void *>(v);
    a.pa = ortho ? ortho->pa : nullptr;
    a.pb = ortho ? ortho->pb : nullptr;
    a.ca_re = ortho ? ortho->ca_re : 0.0;
    a.ca_im = ortho ? ortho->ca_im : 0.0;
    a.cb_re = ortho ? ortho->cb_re : 0.0;
    a.cb_im = ortho ? ortho->cb_im : 0.0;
    a.pb_scale = ortho ? ortho->pb_scale : 0.0;
    a.dcoef = (ortho != nullptr && p == 0) ? ortho->dcoef : nullptr;
    a.v_scale = (ortho != nullptr && p > 0) ? 1.0 : vsc;  // later passes read z = v_scale*w already
    a.sig_re = sr;
    a.sig_im = si;
    a.n = n;
    a.c = plan[p].c;
    a.s = plan[p].s;
    a.m = plan[p].m;
    a.flip_mask = plan[p].fmask;
    const bool want_dots = last && dots != nullptr;
    if (want_dots && p == 0)
      return fail(EMU_SV_EINVAL, "btile: the reductions need a second pass (n above one tile)");
    const int nrow = (int)(1ull << (n - TB));
    // the last pass folds pass A's rows into its own (one set to reduce)
    const bool fold = want_dots && ortho != nullptr && nrow <= kFinFoldMaxRows;
    a.fold = fold ? a.partials2 : nullptr;
    int rc = launch_btile<T>(a, cplx || noise, noise, p > 0, want_dots, st);
    if (rc) return rc;
    if (a.ortho) {
      if (dots) {
        nrow_a = nrow;  // reduced with the last pass's dots
      } else {
        rc = finalize(a.partials2, nrow, kOrthoSums, ws, n, ortho->norm2_out, st);
        if (rc) return rc;
      }
    }
    if (fold) {
      rc = finalize_folded(a.partials, nrow, ortho->norm2_out, dots, ortho->nc, st);
      if (rc) return rc;
    } else if (want_dots) {
      // the fused dots of a btile pass: [0] ||w||^2, [1,2] <v|w>; [3..6] zero
      rc = ortho ? finalize_pair(a.partials2, nrow_a, kOrthoSums, ortho->norm2_out, a.partials, nrow,
                                 7, dots, ws, n, ortho->nc, st)
                 : finalize(a.partials, nrow, 7, ws, n, dots, st);
      if (rc) return rc;
    }
  }
  return EMU_SV_OK;
}


// ---------------------------------------------------------------------------
// Batch engine host side.
// ---------------------------------------------------------------------------
struct BatchHandle {
  int count = 0, dtype = 0, max_dim = 0, npass_max = 0;
  double tol = 0.0;
  BatchReg *regs = nullptr;
  BatchEntry *tab[4] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t ntab[4] = {0, 0, 0, 0};
  BatchEntry *comb = nullptr;
  uint64_t ncomb = 0;
  int *counter = nullptr;
  size_t lds = 0;
  hipStream_t capture_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<BatchReg> host_regs;
};

struct BatchLayout {
  size_t regs, arrays, partials, tables, counter, total;
  uint64_t ntab[4], ncomb;
};

int batch_plan(const emu_sv_batch_register &r, int dtype, int tl, int cmin, BatchReg &out) {
  if (r.n < 1 || r.n > 30 || r.steps < 1 || !r.ham || !r.dt || !r.psi || !r.psi_alt || !r.rows)
    return fail(EMU_SV_EINVAL, "batch: bad register description");
  if (r.flags & EMU_SV_FLAG_NOISE)
    return fail(EMU_SV_EINVAL, "batch: the Lanczos engine needs a Hermitian H (no noise term)");
  if ((reinterpret_cast<uintptr_t>(r.psi) | reinterpret_cast<uintptr_t>(r.psi_alt) |
       reinterpret_cast<uintptr_t>(r.rows)) & 15)
    return fail(EMU_SV_EINVAL, "batch: psi, psi_alt and rows must be 16-byte aligned");
  // a register may carry its own schedule in its flags (0: the batch's)
  const int rtl = (r.flags >> 16) & 0xF, rcmin = (r.flags >> 20) & 0xF;
  if (rtl) tl = rtl;
  if (rcmin) cmin = rcmin;
  const int tmax = dtype == EMU_SV_C64 ? 14 : 13;  // 128 KiB of LDS
  if (tl < kThreadBits || tl > tmax || cmin < 1 || cmin >= tl)
    return fail(EMU_SV_EINVAL, "batch: register schedule needs tile_log in [8, 13] (complex64: 14) "
                               "and min_contig in [1, tile_log)");
  const int t = r.n < tl ? r.n : tl;
  Pass plan[EMU_SV_MAX_QUBITS];
  const int np = make_plan(r.n, t, cmin, plan);
  if (np > 4) return fail(EMU_SV_EINVAL, "batch: register needs more than 4 passes");
  out = BatchReg{};
  out.n = r.n;
  out.steps = r.steps;
  out.cplx = (r.flags & EMU_SV_FLAG_COMPLEX) ? 1 : 0;
  out.npass = np;
  out.pause = r.pause_each_step ? 1 : 0;
  out.ntiles_max = 0;
  for (int p = 0; p < np; ++p) {
    out.plan[p] = {plan[p].c, plan[p].s, plan[p].m, plan[p].diag, plan[p].fmask};
    out.tiles[p] = 1ull << (r.n - (plan[p].c + plan[p].m));
    if (out.tiles[p] > out.ntiles_max) out.ntiles_max = out.tiles[p];
  }
  return EMU_SV_OK;
}

int batch_layout(int count, const emu_sv_batch_register *regs, int dtype, int tl, int cmin,
                 BatchLayout &L, std::vector<BatchReg> *built) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  L = BatchLayout{};
  size_t off = 0;
  L.regs = off;
  off = al(off + sizeof(BatchReg) * (size_t)count);
  L.arrays = off;
  off = al(off + sizeof(double) * BatchArr::kTotal * (size_t)count);
  L.partials = off;
  uint64_t part = 0;
  for (int i = 0; i < count; ++i) {
    BatchReg br;
    if (int rc = batch_plan(regs[i], dtype, tl, cmin, br)) return rc;
    part += 2 * br.ntiles_max * kNumDots;
    for (int p = 0; p < br.npass; ++p) L.ntab[p] += br.tiles[p];
    L.ncomb += ((1ull << br.n) + kBatchChunk - 1) / kBatchChunk;
    if (built) built->push_back(br);
  }
  off = al(off + sizeof(double) * part);
  L.tables = off;
  uint64_t ent = L.ncomb;
  for (int p = 0; p < 4; ++p) ent += L.ntab[p];
  off = al(off + sizeof(BatchEntry) * ent);
  L.counter = off;
  off = al(off + 256);
  L.total = off;
  return EMU_SV_OK;
}

int batch_tile_logs(int dtype, int &tl, int &cmin) {
  // defaults measured on the 256-register workload (N 8..20): 2^12 / 4 for
  // complex64, 2^11 / 2 for complex128 (32 KiB tiles)
  if (tl <= 0) tl = dtype == EMU_SV_C64 ? 12 : 11;
  if (cmin <= 0) cmin = dtype == EMU_SV_C64 ? 4 : 2;
  const int tmax = dtype == EMU_SV_C64 ? 14 : 13;  // 128 KiB of LDS
  if (tl < kThreadBits || tl > tmax || cmin < 1 || cmin >= tl)
    return fail(EMU_SV_EINVAL, "batch: tile_log must be in [8, 13] (complex64: 14), "
                               "min_contig in [1, tile_log)");
  return EMU_SV_OK;
}

template <typename T>
int batch_enqueue_iteration(BatchHandle &h, hipStream_t st) {
  for (int p = 0; p < h.npass_max; ++p) {
    if (h.ntab[p] == 0) continue;
    hipLaunchKernelGGL(batch_pass_kernel<T>, dim3((unsigned)h.ntab[p]), dim3(kThreads), h.lds, st,
                       h.regs, h.tab[p], p);
    if (int rc = hip_check(hipGetLastError(), "batch pass launch")) return rc;
  }
  hipLaunchKernelGGL(batch_decide_kernel, dim3((unsigned)h.count), dim3(64), 0, st, h.regs, h.tol,
                     h.max_dim);
  if (int rc = hip_check(hipGetLastError(), "batch decide launch")) return rc;
  hipLaunchKernelGGL(batch_combine_kernel<T>, dim3((unsigned)h.ncomb), dim3(kThreads), 0, st, h.regs,
                     h.comb);
  if (int rc = hip_check(hipGetLastError(), "batch combine launch")) return rc;
  hipLaunchKernelGGL(batch_advance_kernel, dim3((unsigned)((h.count + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                     st, h.regs, h.count);
  return hip_check(hipGetLastError(), "batch advance launch");
}

// Workspace layout (bytes): [partial rows | pass-1 diagonal table |
// finalize stage | ticket counter].  The caller zero-fills it once; the
// ticket is re-armed by every reduction.
constexpr int kCorrBlocks = 1024;
struct WsLayout {
  size_t partials, partials2, dtab, stage, ticket, total;
};

WsLayout ws_layout(int n) {
  // The ticket and the finalize stage sit at fixed offsets, so a workspace
  // sized for a larger n can serve any smaller n without stale counters.
  const uint64_t tiles = n > kThreadBits ? (1ull << (n - kThreadBits)) : 1ull;
  const uint64_t rows = tiles > (uint64_t)kStreamBlocks ? tiles : (uint64_t)kStreamBlocks;
  uint64_t part = rows * kNumDots * sizeof(double);
  const uint64_t corr = (uint64_t)kCorrBlocks * EMU_SV_MAX_QUBITS * EMU_SV_MAX_QUBITS * sizeof(double);
  if (corr > part) part = corr;
  WsLayout l;
  l.ticket = 0;
  l.stage = 256;
  l.partials = l.stage + kFinalizeBlocks * kNumDots * sizeof(double);
  l.partials2 = l.partials + part;
  l.dtab = l.partials2 + tiles * kNumDots * sizeof(double);
  l.total = l.dtab + tiles * kDiagRow * sizeof(double);
  return l;
}

double *partials_ptr(void *ws, int n) {
  return reinterpret_cast<double *>(reinterpret_cast<char *>(ws) + ws_layout(n).partials);
}

double *partials2_ptr(void *ws, int n) {
  return reinterpret_cast<double *>(reinterpret_cast<char *>(ws) + ws_layout(n).partials2);
}

const double *dtab_ptr(void *ws, int n) {
  return reinterpret_cast<const double *>(reinterpret_cast<char *>(ws) + ws_layout(n).dtab);
}

// The reductions that end a fused Lanczos iteration, in one launch: set A
// (the orthogonalisation sums of the first pass, KA values per row) and set
// B (the dots of the last pass), each summed in a fixed order as by
// finalize2_kernel (blocks [0, nbA) take A, the rest B; the last block adds
// the block sums).  With next != nullptr the last block also forms the next
// iteration's device coefficients (lanczos_next_coef_body).
__global__ __launch_bounds__(kThreads) void finalize_pair_kernel(
    const double *rows_a, int nrow_a, int ka, double *out_a, int nb_a, const double *rows_b,
    int nrow_b, int kb, double *out_b, double *stage, unsigned *ticket, NextCoefOut nc) {
  __shared__ double red[kThreads];
  __shared__ double res[2][kNumDots];
  __shared__ bool last;
  const bool is_a = (int)blockIdx.x < nb_a;
  const int blk = is_a ? (int)blockIdx.x : (int)blockIdx.x - nb_a;
  const int nblk = is_a ? nb_a : (int)gridDim.x - nb_a;
  const double *rows = is_a ? rows_a : rows_b;
  const int nrow = is_a ? nrow_a : nrow_b;
  const int per = (nrow + nblk - 1) / nblk;
  const int r0 = blk * per;
  const int r1 = min(nrow, r0 + per);
  const int v = threadIdx.x & 7, rl = threadIdx.x >> 3;  // 32 row lanes x 8 values
  double acc = 0.0;
  for (int r = r0 + rl; r < r1; r += kThreads / 8) acc += rows[(uint64_t)r * kNumDots + v];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int hlf = kThreads / 2; hlf >= 8; hlf >>= 1) {
    if (threadIdx.x < hlf) red[threadIdx.x] += red[threadIdx.x + hlf];
    __syncthreads();
  }
  if (threadIdx.x < kNumDots) stage[blockIdx.x * kNumDots + threadIdx.x] = red[threadIdx.x];
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int set = 0; set < 2; ++set) {
    const unsigned b0 = set == 0 ? 0u : (unsigned)nb_a;
    const unsigned b1 = set == 0 ? (unsigned)nb_a : gridDim.x;
    double sm = 0.0;
    for (unsigned b = b0 + rl; b < b1; b += kThreads / 8)
      sm += __hip_atomic_load(&stage[b * kNumDots + v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    red[threadIdx.x] = sm;
    __syncthreads();
    for (int hlf = kThreads / 2; hlf >= 8; hlf >>= 1) {
      if (threadIdx.x < hlf) red[threadIdx.x] += red[threadIdx.x + hlf];
      __syncthreads();
    }
    if (threadIdx.x < kNumDots) res[set][threadIdx.x] = red[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x < ka) out_a[threadIdx.x] = res[0][threadIdx.x];
  if (threadIdx.x < kb) out_b[threadIdx.x] = res[1][threadIdx.x];
  if (threadIdx.x == 0) {
    if (nc.next) lanczos_next_coef_body(res[0], res[1], nc.sig_re, nc.sig_im, nc.prev, nc.next, nc.tout);
    atomicExch(ticket, 0u);
  }
}

// The same two reductions in one workgroup, for the row counts of the
// two-pass (btile) schedules: one CU reads the rows with 16-byte loads, up
// to sixteen in flight per thread, and no block sums are handed over
// through a ticket (11 against 18 us for 2 x 4096 rows).  Larger row counts
// take the grid form.  Fixed summation order.  FOLD: one set of rows
// holding both reductions, [0, 3) the dots of set B and [3, 8) the KA = 5
// sums of set A (btile pass B copies pass A's row of the same block into
// its own), so half the bytes are read.
constexpr int kFinOneThreads = 1024;
constexpr int kFinOneMaxRows = 4096;  // rows per set of the two-set form
template <bool FOLD>
__global__ __launch_bounds__(kFinOneThreads) void finalize_one_kernel(
    const double *rows_a, int nrow_a, int ka, double *out_a, const double *rows_b, int nrow_b,
    int kb, double *out_b, NextCoefOut nc) {
  static_assert(kNumDots == 8, "a row is four 16-byte packets");
  constexpr int NW = kFinOneThreads / 64;
  constexpr int NL = kFinOneThreads / 4;  // row lanes
  constexpr int CH = 16;                  // independent load chains per thread
  constexpr int NSET = FOLD ? 1 : 2;
  __shared__ double red[2][NW][kNumDots];
  __shared__ double res[2][kNumDots];
  const int q = threadIdx.x & 3;  // packet of the row: values 2q, 2q+1
  const int rl = threadIdx.x >> 2;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int set = 0; set < NSET; ++set) {
    const double2 *rows = reinterpret_cast<const double2 *>(set == 0 && !FOLD ? rows_a : rows_b);
    const int nrow = set == 0 && !FOLD ? nrow_a : nrow_b;
    double ax[CH], ay[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) ax[c] = ay[c] = 0.0;
    int r = rl;
    for (; r + (CH - 1) * NL < nrow; r += CH * NL) {
      double2 x[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c) x[c] = rows[(uint64_t)(r + c * NL) * 4 + q];
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        ax[c] += x[c].x;
        ay[c] += x[c].y;
      }
    }
    // the rest: fewer than CH rows per lane, each into its own chain
#pragma unroll
    for (int c = 0; c < CH - 1; ++c) {
      const int rr = r + c * NL;
      if (rr < nrow) {
        const double2 x = rows[(uint64_t)rr * 4 + q];
        ax[c] += x.x;
        ay[c] += x.y;
      }
    }
    double sx = 0.0, sy = 0.0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      sx += ax[c];
      sy += ay[c];
    }
    // the lanes of one packet index q: lane bits 2..5
#pragma unroll
    for (int off = 4; off < 64; off <<= 1) {
      sx += __shfl_xor(sx, off, 64);
      sy += __shfl_xor(sy, off, 64);
    }
    if (lane < 4) {
      red[set][wave][2 * lane] = sx;
      red[set][wave][2 * lane + 1] = sy;
    }
  }
  __syncthreads();
  if (threadIdx.x < NSET * kNumDots) {
    const int set = threadIdx.x / kNumDots, v = threadIdx.x % kNumDots;
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[set][w][v];
    if (!FOLD)
      res[set][v] = s;
    else if (v < 3)
      res[1][v] = s;
    else
      res[0][v - 3] = s;
  }
  if (FOLD && threadIdx.x >= 3 && threadIdx.x < kNumDots) res[1][threadIdx.x] = 0.0;
  __syncthreads();
  if ((int)threadIdx.x < ka) out_a[threadIdx.x] = res[0][threadIdx.x];
  if ((int)threadIdx.x < kb) out_b[threadIdx.x] = res[1][threadIdx.x];
  if (threadIdx.x == 0 && nc.next)
    lanczos_next_coef_body(res[0], res[1], nc.sig_re, nc.sig_im, nc.prev, nc.next, nc.tout);
}

// The reductions of a btile iteration whose last pass folded pass A's rows
// into its own (FOLD above); nrow <= kFinFoldMaxRows.
int finalize_folded(const double *rows, int nrow, double *out_a, double *out_b,
                    const NextCoefOut &nc, hipStream_t st) {
  hipLaunchKernelGGL(finalize_one_kernel<true>, dim3(1), dim3(kFinOneThreads), 0, st, nullptr, 0,
                     kOrthoSums, out_a, rows, nrow, 7, out_b, nc);
  return hip_check(hipGetLastError(), "finalize launch");
}

int finalize_pair(const double *rows_a, int nrow_a, int ka, double *out_a, const double *rows_b,
                  int nrow_b, int kb, double *out_b, void *ws, int n, const NextCoefOut &nc,
                  hipStream_t st) {
  if (nrow_a <= kFinOneMaxRows && nrow_b <= kFinOneMaxRows) {
    hipLaunchKernelGGL(finalize_one_kernel<false>, dim3(1), dim3(kFinOneThreads), 0, st, rows_a,
                       nrow_a, ka, out_a, rows_b, nrow_b, kb, out_b, nc);
    return hip_check(hipGetLastError(), "finalize launch");
  }
  const WsLayout l = ws_layout(n);
  char *base = reinterpret_cast<char *>(ws);
  auto blocks = [](int nrow) {
    int nb = (nrow + 31) / 32;
    if (nb > kFinalizeBlocks / 2) nb = kFinalizeBlocks / 2;
    return nb < 1 ? 1 : nb;
  };
  const int nb_a = blocks(nrow_a), nb_b = blocks(nrow_b);
  hipLaunchKernelGGL(finalize_pair_kernel, dim3(nb_a + nb_b), dim3(kThreads), 0, st, rows_a, nrow_a,
                     ka, out_a, nb_a, rows_b, nrow_b, kb, out_b,
                     reinterpret_cast<double *>(base + l.stage),
                     reinterpret_cast<unsigned *>(base + l.ticket), nc);
  return hip_check(hipGetLastError(), "finalize launch");
}

int finalize(const double *rows, int nrow, int K, void *ws, int n, double *out, hipStream_t st) {
  const WsLayout l = ws_layout(n);
  char *base = reinterpret_cast<char *>(ws);
  int nb = (nrow + 31) / 32;
  if (nb > kFinalizeBlocks) nb = kFinalizeBlocks;
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(finalize2_kernel, dim3(nb), dim3(kThreads), 0, st, rows, nrow, K,
                     reinterpret_cast<double *>(base + l.stage),
                     reinterpret_cast<unsigned *>(base + l.ticket), out);
  return hip_check(hipGetLastError(), "finalize launch");
}

// 16-byte packets in a 2^n vector (c64: two amplitudes per packet; n >= 1).
uint64_t packets(int n, int dtype) { return dtype == EMU_SV_C64 ? (1ull << (n - 1)) : (1ull << n); }

int stream_blocks(uint64_t D) {
  uint64_t b = (D + kThreads - 1) / kThreads;
  return (int)(b < (uint64_t)kStreamBlocks ? b : (uint64_t)kStreamBlocks);
}

int check_n(int n) {
  if (n < 1 || n > EMU_SV_MAX_QUBITS)
    return fail(EMU_SV_EINVAL, "number of qubits must be in [1, 32], got " + std::to_string(n));
  return EMU_SV_OK;
}

int check_dtype(int dtype) {
  if (dtype != EMU_SV_C64 && dtype != EMU_SV_C128)
    return fail(EMU_SV_EINVAL, "dtype must be EMU_SV_C128 (0) or EMU_SV_C64 (1)");
  return EMU_SV_OK;
}

int tile_partners(int flags) { return (flags >> 12) & 0xF; }

// First-pass tile log of the tile kernel: flags bits 16-19 (0 = tile_log).
int tile_first_log(int flags, int tl) {
  const int t = (flags >> 16) & 0xF;
  return t ? t : tl;
}

// 0: valid tile schedule; 1: n below one tile (use the LDS kernel); <0: error.
int check_tile_schedule(int n, int dtype, int &tl, int &cmin, int &tla, const char *what) {
  if (tl <= 0) tl = dtype == EMU_SV_C64 ? 13 : 12;
  if (cmin <= 0) cmin = dtype == EMU_SV_C64 ? 5 : 4;
  if (tla <= 0) tla = tl;
  const int mx = tile_max_log(dtype);
  if (tl < 10 || tl > mx || tla < 10 || tla > mx)
    return fail(EMU_SV_EINVAL, std::string(what) + ": tile kernel needs tile logs in [10, " +
                                   std::to_string(mx) + "]");
  if (cmin < 1 || cmin >= tl)
    return fail(EMU_SV_EINVAL, std::string(what) + ": min_contig must be in [1, tile_log)");
  return (n < tl || n < tla) ? 1 : 0;
}

int lanczos_matvec_dev_impl(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                            void *w, const void *pa, const void *pb, const double *coef, void *r,
                            double *norm2_out, double *dots_out, void *workspace, int tile_log,
                            int min_contig, int flags, hipStream_t stream, const NextCoefOut &nc) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!ham || !w || !r || !coef || !norm2_out || !dots_out || !workspace)
    return fail(EMU_SV_EINVAL, "lanczos_matvec_dev: null pointer");
  if (w == r || w == pa || w == pb || r == pa || r == pb)
    return fail(EMU_SV_EINVAL, "lanczos_matvec_dev: w, r, pa and pb must be distinct buffers");
  const int kind = kernel_kind(flags);
  if (kind != kKernelTile && kind != kKernelLds && kind != kKernelBtile)
    return fail(EMU_SV_EINVAL, "lanczos_matvec_dev: needs the tile, btile or LDS kernel kind");
  if (flags & EMU_SV_FLAG_NOISE)
    return fail(EMU_SV_EINVAL, "lanczos_matvec_dev: the Lanczos recurrence needs a Hermitian H (no noise term)");
  if (kind == kKernelBtile) {
    const int rc = check_btile_schedule(n, dtype, min_contig, "lanczos_matvec_dev");
    if (rc < 0) return rc;
    if (rc > 0) return fail(EMU_SV_EINVAL, "lanczos_matvec_dev: btile needs more qubits than one tile");
    OrthoIn o;
    o.pa = pa;
    o.pb = pb;
    o.ca_re = o.ca_im = o.cb_re = o.cb_im = 0.0;
    o.pb_scale = 0.0;
    o.norm2_out = norm2_out;
    o.dcoef = coef;
    o.nc = nc;
    const bool bc = (flags & EMU_SV_FLAG_COMPLEX) != 0;
    if (dtype == EMU_SV_C128)
      return run_btile<double>(n, ham, sigma_re, sigma_im, w, 1.0, r, dots_out, workspace, min_contig,
                               bc, stream, &o);
    return run_btile<float>(n, ham, sigma_re, sigma_im, w, 1.0, r, dots_out, workspace, min_contig, bc,
                            stream, &o);
  }
  int tla = 0;
  if (kind == kKernelTile) {
    tla = tile_first_log(flags, 0);
    const int rc = check_tile_schedule(n, dtype, tile_log, min_contig, tla, "lanczos_matvec_dev");
    if (rc < 0) return rc;
    if (rc > 0) return fail(EMU_SV_EINVAL, "lanczos_matvec_dev: fewer qubits than one tile");
  } else {
    default_tiles(dtype, kind, tile_log, min_contig);
    if (tile_log < kThreadBits || tile_log > kMaxTileLog || (dtype == EMU_SV_C128 && tile_log > 13) ||
        min_contig < 0 || min_contig >= tile_log)
      return fail(EMU_SV_EINVAL, "lanczos_matvec_dev: bad tile_log / min_contig");
  }
  OrthoIn o;
  o.pa = pa;
  o.pb = pb;
  o.ca_re = o.ca_im = o.cb_re = o.cb_im = 0.0;
  o.pb_scale = 0.0;
  o.norm2_out = norm2_out;
  o.dcoef = coef;
  o.nc = nc;
  const bool cplx = (flags & EMU_SV_FLAG_COMPLEX) != 0;
  const bool keep = (flags & EMU_SV_FLAG_KEEP_DIAG) != 0;
  if (kind == kKernelLds) {
    if (dtype == EMU_SV_C128)
      return run_hpsi<double>(n, ham, sigma_re, sigma_im, w, 1.0, nullptr, 0.0, r, dots_out, workspace,
                              kind, tile_log, min_contig, cplx, stream, &o, false, keep);
    return run_hpsi<float>(n, ham, sigma_re, sigma_im, w, 1.0, nullptr, 0.0, r, dots_out, workspace,
                           kind, tile_log, min_contig, cplx, stream, &o, false, keep);
  }
  if (dtype == EMU_SV_C128)
    return run_tile<double>(n, ham, sigma_re, sigma_im, w, 1.0, nullptr, 0.0, r, dots_out, workspace,
                            tla, tile_log, min_contig, tile_partners(flags), cplx, stream, &o, keep);
  return run_tile<float>(n, ham, sigma_re, sigma_im, w, 1.0, nullptr, 0.0, r, dots_out, workspace, tla,
                         tile_log, min_contig, tile_partners(flags), cplx, stream, &o, keep);
}

int hpsi_impl(int n, int dtype, const double *ham, double sigma_re, double sigma_im, const void *v,
              double v_scale, const void *prev, double prev_scale, void *w, double *dots_out,
              void *workspace, int tile_log, int min_contig, int flags, hipStream_t stream);

}  // namespace

// ---------------------------------------------------------------------------
// Static scratch of emu_sv_hpsi_static: the workspace of callers that cannot
// allocate device memory (code under stream capture, tight serving loops).
// Each GPU holds its own copy of the array until the process ends.  In host
// code the array's name is only a stand-in for the symbol: its device
// address is looked up once per device (hipGetSymbolAddress, with that
// device current) and only that address is handed to kernels.  Every call
// on a GPU shares the array: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kStaticScratchBytes = 24ull << 20;  // the workspace of n <= 24
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
void *g_static_ptr[kMaxDevices] = {};

// Clears the reduction ticket and the finalize stage (the only parts of a
// workspace that must start at zero; the ticket is re-armed by every
// reduction).  A fill kernel rather than a memset, so the first call may
// also be recorded into a graph.
__global__ void zero_words_kernel(uint32_t *p, size_t words) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

int static_scratch(hipStream_t st, void **out) {
  int dev = 0;
  if (int rc = hip_check(hipGetDevice(&dev), "hipGetDevice")) return rc;
  if (dev < 0 || dev >= kMaxDevices)
    return fail(EMU_SV_EINVAL, "hpsi_static: device index above " + std::to_string(kMaxDevices));
  std::lock_guard<std::mutex> lock(g_static_mu);
  if (!g_static_ptr[dev]) {
    void *p = nullptr;
    if (int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                           "hipGetSymbolAddress(static scratch)"))
      return rc;
    const size_t words = ws_layout(1).partials / sizeof(uint32_t);  // ticket + stage
    hipLaunchKernelGGL(zero_words_kernel, dim3(8), dim3(kThreads), 0, st,
                       reinterpret_cast<uint32_t *>(p), words);
    if (int rc = hip_check(hipGetLastError(), "static scratch clear")) return rc;
    g_static_ptr[dev] = p;
  }
  *out = g_static_ptr[dev];
  return EMU_SV_OK;
}
}  // namespace

namespace {
int lanczos_iterate_impl(int n, int dtype, const double *ham, double sigma_re, double sigma_im, void *w,
                         const void *pa, const void *pb, const double *coef, void *r, double *norm2_out,
                         double *dots_out, double *next_coef, double *t_out, void *workspace,
                         int tile_log, int min_contig, int flags, hipStream_t stream);

// The device address of pinned (page-locked) host memory, or nullptr for
// any other pointer (pageable memory keeps explicit copies).
double *device_view_of_pinned(double *host) {
  hipPointerAttribute_t attr{};
  if (hipPointerGetAttributes(&attr, host) != hipSuccess) {
    (void)hipGetLastError();  // a pageable pointer is not an error here
    return nullptr;
  }
  if (attr.type != hipMemoryTypeHost || !attr.devicePointer || !attr.hostPointer) return nullptr;
  const ptrdiff_t off = reinterpret_cast<char *>(host) - static_cast<char *>(attr.hostPointer);
  return reinterpret_cast<double *>(static_cast<char *>(attr.devicePointer) + off);
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *emu_sv_last_error(void) { return g_last_error.c_str(); }

const char *emu_sv_version(void) { return "emulators_amd emu_sv_hip 0.1.0 (gfx950)"; }

int emu_sv_pack_hamiltonian(int n, const double *omega, const double *delta, const double *phi,
                            const double *interaction, double *block) {
  if (int rc = check_n(n)) return rc;
  if (!omega || !delta || !phi || !interaction || !block)
    return fail(EMU_SV_EINVAL, "pack_hamiltonian: null pointer");
  memset(block, 0, sizeof(double) * EMU_SV_HAM_DOUBLES);
  bool cplx = false;
  for (int q = 0; q < n; ++q) {
    const int b = n - 1 - q;
    const double half = 0.5 * omega[q];
    const double cr = half * cos(phi[q]), ci = half * sin(phi[q]);
    block[kCre + b] = phi[q] != 0.0 ? cr : half;
    block[kCim + b] = phi[q] != 0.0 ? ci : 0.0;
    if (phi[q] != 0.0) cplx = true;
    block[kDelta + b] = delta[q];
  }
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      const int bi = n - 1 - i, bj = n - 1 - j;
      block[kU + bi * 32 + bj] = interaction[i * n + j];
      block[kU + bj * 32 + bi] = interaction[i * n + j];
    }
  block[kFlagComplex] = cplx ? 1.0 : 0.0;
  static std::atomic<uint64_t> generation{1};
  block[kGeneration] = (double)generation.fetch_add(1, std::memory_order_relaxed);
  return EMU_SV_OK;
}

int emu_sv_pack_noise(const double *noise, double *block) {
  if (!noise || !block) return fail(EMU_SV_EINVAL, "pack_noise: null pointer");
  for (int k = 0; k < 8; ++k) {
    if (!std::isfinite(noise[k])) return fail(EMU_SV_EINVAL, "pack_noise: non-finite entry");
    block[kNoise + k] = noise[k];
  }
  block[kFlagNoise] = 1.0;
  return EMU_SV_OK;
}

int emu_sv_plan_passes(int n, int dtype, int tile_log, int min_contig, int flags) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (kernel_kind(flags) == kKernelBtile) {
    const int rc = check_btile_schedule(n, dtype, min_contig, "plan_passes");
    if (rc < 0) return rc;
    if (rc == 0) {
      Pass bplan[EMU_SV_MAX_QUBITS];
      return make_btile_plan(n, bt_tile_bits_of(dtype), min_contig, bplan);
    }
    tile_log = 0;
    min_contig = 0;
    flags = EMU_SV_KERNEL_LDS;
  }
  if (kernel_kind(flags) == kKernelTile) {
    int tla = tile_first_log(flags, 0);
    const int rc = check_tile_schedule(n, dtype, tile_log, min_contig, tla, "plan_passes");
    if (rc < 0) return rc;
    if (rc == 0) {
      TilePass tplan[EMU_SV_MAX_QUBITS];
      return make_tile_plan(n, tla, tile_log, min_contig, tile_partners(flags), tplan);
    }
    tile_log = 0;
    min_contig = 0;
    flags = EMU_SV_KERNEL_LDS;
  }
  default_tiles(dtype, kernel_kind(flags), tile_log, min_contig);
  Pass plan[EMU_SV_MAX_QUBITS];
  return make_plan(n, tile_log, min_contig, plan);
}

size_t emu_sv_workspace_bytes(int n, int dtype) {
  (void)dtype;
  if (n < 1 || n > EMU_SV_MAX_QUBITS) return 0;
  return ws_layout(n).total;
}

int emu_sv_hpsi(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                const void *v, double v_scale, const void *prev, double prev_scale, void *w,
                double *dots_out, void *workspace, int tile_log, int min_contig, int flags,
                hipStream_t stream) {
  if (!workspace) return fail(EMU_SV_EINVAL, "hpsi: null workspace (it holds the diagonal table)");
  return hpsi_impl(n, dtype, ham, sigma_re, sigma_im, v, v_scale, prev, prev_scale, w, dots_out,
                   workspace, tile_log, min_contig, flags, stream);
}

int emu_sv_hpsi_static(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                       const void *v, double v_scale, const void *prev, double prev_scale, void *w,
                       double *dots_out, int tile_log, int min_contig, int flags,
                       hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  const size_t need = ws_layout(n).total;
  if (need > kStaticScratchBytes)
    return fail(EMU_SV_EINVAL, "hpsi_static: n = " + std::to_string(n) + " needs " +
                                   std::to_string(need) + " bytes of scratch, above the library's " +
                                   std::to_string(kStaticScratchBytes) +
                                   "-byte static array (n <= 24); use emu_sv_hpsi with a workspace");
  void *ws = nullptr;
  if (int rc = static_scratch(stream, &ws)) return rc;
  return hpsi_impl(n, dtype, ham, sigma_re, sigma_im, v, v_scale, prev, prev_scale, w, dots_out, ws,
                   tile_log, min_contig, flags, stream);
}

}  // extern "C"

namespace {
int hpsi_impl(int n, int dtype, const double *ham, double sigma_re, double sigma_im, const void *v,
              double v_scale, const void *prev, double prev_scale, void *w, double *dots_out,
              void *workspace, int tile_log, int min_contig, int flags, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!ham || !v || !w) return fail(EMU_SV_EINVAL, "hpsi: null pointer");
  if (v == w) return fail(EMU_SV_EINVAL, "hpsi: input and output must not alias");
  int kind = kernel_kind(flags);
  if (kind != kKernelLds && kind != kKernelWave && kind != kKernelTile && kind != kKernelBtile)
    return fail(EMU_SV_EINVAL, "hpsi: unknown kernel kind in flags");
  const bool noise = (flags & EMU_SV_FLAG_NOISE) != 0;
  const bool cplx = noise || (flags & EMU_SV_FLAG_COMPLEX) != 0;
  const bool keep = (flags & EMU_SV_FLAG_KEEP_DIAG) != 0;
  if (kind == kKernelBtile) {
    const int rc = check_btile_schedule(n, dtype, min_contig, "hpsi");
    if (rc < 0) return rc;
    if (rc == 0 && !prev) {  // the btile passes do not reduce the prev dots
      if (!workspace) return fail(EMU_SV_EINVAL, "hpsi: null workspace");
      if (dtype == EMU_SV_C128)
        return run_btile<double>(n, ham, sigma_re, sigma_im, v, v_scale, w, dots_out, workspace,
                                 min_contig, cplx, stream, nullptr, noise);
      return run_btile<float>(n, ham, sigma_re, sigma_im, v, v_scale, w, dots_out, workspace,
                              min_contig, cplx, stream, nullptr, noise);
    }
    kind = kKernelLds;  // n within one tile, or prev dots wanted: the LDS kernel
    tile_log = 0;
    min_contig = 0;
  }
  if (noise && kind != kKernelLds) {  // the LDS kernel carries the noise term below one btile
    kind = kKernelLds;
    tile_log = 0;
    min_contig = 0;
  }
  int tla = kind == kKernelTile ? tile_first_log(flags, 0) : 0;
  if (kind == kKernelTile) {
    int rc = check_tile_schedule(n, dtype, tile_log, min_contig, tla, "hpsi");
    if (rc > 0) {
      kind = kKernelLds;  // fewer qubits than one tile: the LDS kernel
      tile_log = 0;
      min_contig = 0;
    } else if (rc < 0) {
      return rc;
    } else if (dtype == EMU_SV_C128) {
      return run_tile<double>(n, ham, sigma_re, sigma_im, v, v_scale, prev, prev_scale, w, dots_out,
                              workspace, tla, tile_log, min_contig, tile_partners(flags), cplx,
                              stream, nullptr, keep);
    } else {
      return run_tile<float>(n, ham, sigma_re, sigma_im, v, v_scale, prev, prev_scale, w, dots_out,
                             workspace, tla, tile_log, min_contig, tile_partners(flags), cplx,
                             stream, nullptr, keep);
    }
  }
  default_tiles(dtype, kind, tile_log, min_contig);
  if (kind == kKernelWave && (tile_log < 9 || tile_log > 11))
    return fail(EMU_SV_EINVAL, "hpsi: wave tiles need tile_log in [9, 11]");
  if (tile_log < kThreadBits || tile_log > kMaxTileLog)
    return fail(EMU_SV_EINVAL, "hpsi: tile_log must be in [8, 14]");
  if (dtype == EMU_SV_C128 && tile_log > 13)
    return fail(EMU_SV_EINVAL, "hpsi: complex128 tiles are limited to 2^13 elements");
  if (min_contig < 0 || min_contig >= tile_log)
    return fail(EMU_SV_EINVAL, "hpsi: min_contig must be in [0, tile_log)");
  if (dtype == EMU_SV_C128)
    return run_hpsi<double>(n, ham, sigma_re, sigma_im, v, v_scale, prev, prev_scale, w, dots_out,
                            workspace, kind, tile_log, min_contig, cplx, stream, nullptr, noise, keep);
  return run_hpsi<float>(n, ham, sigma_re, sigma_im, v, v_scale, prev, prev_scale, w, dots_out,
                         workspace, kind, tile_log, min_contig, cplx, stream, nullptr, noise, keep);
}
}  // namespace

extern "C" {

int emu_sv_lanczos_matvec(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                          void *w, double w_scale, const void *pa, double pa_scale, double a_re,
                          double a_im, const void *pb, double pb_scale, double b_re, double b_im,
                          void *r, double *norm2_out, double *dots_out, void *workspace,
                          int tile_log, int min_contig, int flags, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!ham || !w || !r || !norm2_out || !dots_out || !workspace)
    return fail(EMU_SV_EINVAL, "lanczos_matvec: null pointer");
  if (w == r || w == pa || w == pb || r == pa || r == pb)
    return fail(EMU_SV_EINVAL, "lanczos_matvec: w, r, pa and pb must be distinct buffers");
  int kind = kernel_kind(flags);
  if (kind != kKernelLds && kind != kKernelWave && kind != kKernelTile && kind != kKernelBtile)
    return fail(EMU_SV_EINVAL, "lanczos_matvec: unknown kernel kind in flags");
  if (flags & EMU_SV_FLAG_NOISE)
    return fail(EMU_SV_EINVAL, "lanczos_matvec: the Lanczos recurrence needs a Hermitian H (no noise term)");
  if (kind == kKernelBtile) {
    const int rc = check_btile_schedule(n, dtype, min_contig, "lanczos_matvec");
    if (rc < 0) return rc;
    if (rc == 0) {
      OrthoIn o;
      o.pa = pa;
      o.pb = pb;
      o.ca_re = a_re * pa_scale;
      o.ca_im = a_im * pa_scale;
      o.cb_re = b_re * pb_scale;
      o.cb_im = b_im * pb_scale;
      o.pb_scale = pb_scale;
      o.norm2_out = norm2_out;
      const bool bc = (flags & EMU_SV_FLAG_COMPLEX) != 0;
      if (dtype == EMU_SV_C128)
        return run_btile<double>(n, ham, sigma_re, sigma_im, w, w_scale, r, dots_out, workspace,
                                 min_contig, bc, stream, &o);
      return run_btile<float>(n, ham, sigma_re, sigma_im, w, w_scale, r, dots_out, workspace,
                              min_contig, bc, stream, &o);
    }
    kind = kKernelLds;
    tile_log = 0;
    min_contig = 0;
  }
  bool use_tile = false;
  int tla = kind == kKernelTile ? tile_first_log(flags, 0) : 0;
  if (kind == kKernelTile) {
    const int rc = check_tile_schedule(n, dtype, tile_log, min_contig, tla, "lanczos_matvec");
    if (rc < 0) return rc;
    if (rc > 0) {
      kind = kKernelLds;
      tile_log = 0;
      min_contig = 0;
    } else {
      use_tile = true;
    }
  }
  if (!use_tile) default_tiles(dtype, kind, tile_log, min_contig);
  if (kind == kKernelWave && (tile_log < 9 || tile_log > 11))
    return fail(EMU_SV_EINVAL, "lanczos_matvec: wave tiles need tile_log in [9, 11]");
  if (!use_tile && (tile_log < kThreadBits || tile_log > kMaxTileLog ||
                    (dtype == EMU_SV_C128 && tile_log > 13) || min_contig < 0 ||
                    min_contig >= tile_log))
    return fail(EMU_SV_EINVAL, "lanczos_matvec: bad tile_log / min_contig");
  OrthoIn o;
  o.pa = pa;
  o.pb = pb;
  o.ca_re = a_re * pa_scale;
  o.ca_im = a_im * pa_scale;
  o.cb_re = b_re * pb_scale;
  o.cb_im = b_im * pb_scale;
  o.pb_scale = pb_scale;
  o.norm2_out = norm2_out;
  const bool cplx = (flags & EMU_SV_FLAG_COMPLEX) != 0;
  const bool keep = (flags & EMU_SV_FLAG_KEEP_DIAG) != 0;
  if (use_tile) {
    if (dtype == EMU_SV_C128)
      return run_tile<double>(n, ham, sigma_re, sigma_im, w, w_scale, nullptr, 0.0, r, dots_out,
                              workspace, tla, tile_log, min_contig, tile_partners(flags), cplx,
                              stream, &o, keep);
    return run_tile<float>(n, ham, sigma_re, sigma_im, w, w_scale, nullptr, 0.0, r, dots_out,
                           workspace, tla, tile_log, min_contig, tile_partners(flags), cplx, stream,
                           &o, keep);
  }
  if (dtype == EMU_SV_C128)
    return run_hpsi<double>(n, ham, sigma_re, sigma_im, w, w_scale, nullptr, 0.0, r, dots_out,
                            workspace, kind, tile_log, min_contig, cplx, stream, &o, false, keep);
  return run_hpsi<float>(n, ham, sigma_re, sigma_im, w, w_scale, nullptr, 0.0, r, dots_out,
                         workspace, kind, tile_log, min_contig, cplx, stream, &o, false, keep);
}

int emu_sv_lanczos_matvec_dev(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                              void *w, const void *pa, const void *pb, const double *coef, void *r,
                              double *norm2_out, double *dots_out, void *workspace, int tile_log,
                              int min_contig, int flags, hipStream_t stream) {
  return lanczos_matvec_dev_impl(n, dtype, ham, sigma_re, sigma_im, w, pa, pb, coef, r, norm2_out,
                                 dots_out, workspace, tile_log, min_contig, flags, stream,
                                 NextCoefOut{});
}

int emu_sv_lanczos_iterate_dev(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                               void *w, const void *pa, const void *pb, const double *coef, void *r,
                               double *norm2_out, double *dots_out, double *next_coef,
                               double *t_out, void *workspace, int tile_log, int min_contig,
                               int flags, hipStream_t stream) {
  // the reductions and side outputs may sit in pinned host memory: the
  // kernels then write them there (no copy command behind the iteration)
  auto at = [](double *p) {
    double *d = p ? device_view_of_pinned(p) : nullptr;
    return d ? d : p;
  };
  return lanczos_iterate_impl(n, dtype, ham, sigma_re, sigma_im, w, pa, pb, coef, r, at(norm2_out),
                              at(dots_out), next_coef, at(t_out), workspace, tile_log, min_contig, flags,
                              stream);
}
}  // extern "C"

namespace {
int lanczos_iterate_impl(int n, int dtype, const double *ham, double sigma_re, double sigma_im, void *w,
                         const void *pa, const void *pb, const double *coef, void *r, double *norm2_out,
                         double *dots_out, double *next_coef, double *t_out, void *workspace,
                         int tile_log, int min_contig, int flags, hipStream_t stream) {
  if (!next_coef || !t_out) return fail(EMU_SV_EINVAL, "lanczos_iterate_dev: null pointer");
  if (next_coef == coef) return fail(EMU_SV_EINVAL, "lanczos_iterate_dev: next_coef must not alias coef");
  NextCoefOut nc;
  nc.prev = coef;
  nc.next = next_coef;
  nc.tout = t_out;
  nc.sig_re = sigma_re;
  nc.sig_im = sigma_im;
  return lanczos_matvec_dev_impl(n, dtype, ham, sigma_re, sigma_im, w, pa, pb, coef, r, norm2_out,
                                 dots_out, workspace, tile_log, min_contig, flags, stream, nc);
}
}  // namespace

extern "C" {

int emu_sv_lanczos_next_coef(const double *norm2, const double *dots, double sigma_re, double sigma_im,
                             const double *prev, double *next, double *t_out, hipStream_t stream) {
  if (!norm2 || !dots || !prev || !next || !t_out) return fail(EMU_SV_EINVAL, "lanczos_next_coef: null");
  hipLaunchKernelGGL(lanczos_next_coef_kernel, dim3(1), dim3(64), 0, stream, norm2, dots, sigma_re,
                     sigma_im, prev, next, t_out);
  return hip_check(hipGetLastError(), "lanczos_next_coef launch");
}

int emu_sv_lanczos_update(int n, int dtype, void *w, const void *pa, double pa_scale, double a_re,
                          double a_im, const void *pb, double pb_scale, double b_re, double b_im,
                          double *norm2_out, void *workspace, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!w || !norm2_out || !workspace) return fail(EMU_SV_EINVAL, "lanczos_update: null pointer");
  const uint64_t npk = packets(n, dtype);
  const int nb = stream_blocks(npk);
  double *part = partials_ptr(workspace, n);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(lanczos_update_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, npk, w,
                       pa, a_re * pa_scale, a_im * pa_scale, pb, b_re * pb_scale,
                       b_im * pb_scale, part);
  else
    hipLaunchKernelGGL(lanczos_update_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, npk, w,
                       pa, a_re * pa_scale, a_im * pa_scale, pb, b_re * pb_scale,
                       b_im * pb_scale, part);
  if (int rc = hip_check(hipGetLastError(), "lanczos_update launch")) return rc;
  return finalize(part, nb, 1, workspace, n, norm2_out, stream);
}

int emu_sv_lincomb(int n, int dtype, int count, const void *const *vecs, const double *coef,
                   void *out, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (count < 1 || count > kMaxLincomb)
    return fail(EMU_SV_EINVAL, "lincomb: count must be in [1, 1024]");
  if (!vecs || !coef || !out) return fail(EMU_SV_EINVAL, "lincomb: null pointer");
  for (int k = 0; k < count; ++k) {
    if (!vecs[k]) return fail(EMU_SV_EINVAL, "lincomb: null vector");
    if (k > 0 && vecs[k] == out)
      return fail(EMU_SV_EINVAL, "lincomb: only the first vector may alias the output");
  }
  const uint64_t npk = packets(n, dtype);
  const int nb = stream_blocks(npk);
  for (int k0 = 0; k0 < count; k0 += kLincombBatch) {
    LincombArgs args;
    memset(&args, 0, sizeof(args));
    args.count = count - k0 < kLincombBatch ? count - k0 : kLincombBatch;
    args.accumulate = k0 > 0;
    for (int k = 0; k < args.count; ++k) {
      args.vec[k] = vecs[k0 + k];
      args.cr[k] = coef[2 * (k0 + k)];
      args.ci[k] = coef[2 * (k0 + k) + 1];
    }
    if (dtype == EMU_SV_C128)
      hipLaunchKernelGGL(lincomb_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, npk, args,
                         out);
    else
      hipLaunchKernelGGL(lincomb_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, npk, args,
                         out);
    if (int rc = hip_check(hipGetLastError(), "lincomb launch")) return rc;
  }
  return EMU_SV_OK;
}

int emu_sv_dot(int n, int dtype, const void *a, const void *b, double *dots_out, void *workspace,
               hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!a || !dots_out || !workspace) return fail(EMU_SV_EINVAL, "dot: null pointer");
  const uint64_t npk = packets(n, dtype);
  const int nb = stream_blocks(npk);
  double *part = partials_ptr(workspace, n);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(dot_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, npk, a, b, part);
  else
    hipLaunchKernelGGL(dot_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, npk, a, b, part);
  if (int rc = hip_check(hipGetLastError(), "dot launch")) return rc;
  return finalize(part, nb, 3, workspace, n, dots_out, stream);
}

int arnoldi_mgs(int n, int dtype, int count, const void *const *vecs, const double *scales, void *w,
                double *records, void *workspace, int normalize, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (count < 1 || count > kMaxLincomb)
    return fail(EMU_SV_EINVAL, "arnoldi_orthogonalize: count must be in [1, 1024]");
  if (!vecs || !scales || !w || !records || !workspace)
    return fail(EMU_SV_EINVAL, "arnoldi_orthogonalize: null pointer");
  for (int k = 0; k < count; ++k) {
    if (!vecs[k]) return fail(EMU_SV_EINVAL, "arnoldi_orthogonalize: null basis vector");
    if (vecs[k] == w) return fail(EMU_SV_EINVAL, "arnoldi_orthogonalize: w aliases a basis vector");
  }
  if (n >= 10 && n <= kMgsSmallLog && count <= kMgsSmallCount)  // one workgroup, one launch
    return dtype == EMU_SV_C128
               ? launch_mgs_small<double>(n, count, vecs, scales, w, records, normalize, stream)
               : launch_mgs_small<float>(n, count, vecs, scales, w, records, normalize, stream);
  const uint64_t npk = packets(n, dtype);
  const int nb = stream_blocks(npk);
  double *part = partials_ptr(workspace, n);
  for (int k = 0; k <= count; ++k) {
    const void *up = k > 0 ? vecs[k - 1] : nullptr;
    const void *un = k < count ? vecs[k] : nullptr;
    const double *rec = k > 0 ? records + 4 * (k - 1) : records;
    const double ps2 = k > 0 ? scales[k - 1] * scales[k - 1] : 0.0;
    if (dtype == EMU_SV_C128)
      hipLaunchKernelGGL(arnoldi_mgs_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, npk, w,
                         up, rec, ps2, un, part);
    else
      hipLaunchKernelGGL(arnoldi_mgs_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, npk, w,
                         up, rec, ps2, un, part);
    if (int rc = hip_check(hipGetLastError(), "arnoldi_mgs launch")) return rc;
    if (int rc = finalize(part, nb, 3, workspace, n, records + 4 * k, stream)) return rc;
  }
  if (!normalize) return EMU_SV_OK;
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(scale_by_norm_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, npk, w,
                       records + 4 * count);
  else
    hipLaunchKernelGGL(scale_by_norm_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, npk, w,
                       records + 4 * count);
  return hip_check(hipGetLastError(), "arnoldi normalisation launch");
}

size_t emu_sv_arnoldi_gram_state_doubles(int ld) {
  return ld < 2 ? 0 : gram_state_doubles(ld);
}

int emu_sv_arnoldi_gram(int n, int dtype, int count, const void *const *vecs, void *w,
                        double v0_scale, double *state, int ld, double *records_out,
                        hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (count < 1 || count > kGramMaxCount)
    return fail(EMU_SV_EINVAL, "arnoldi_gram: count must be in [1, 256]");
  if (ld <= count) return fail(EMU_SV_EINVAL, "arnoldi_gram: ld must exceed count");
  if (!vecs || !w || !state || !records_out) return fail(EMU_SV_EINVAL, "arnoldi_gram: null pointer");
  GramVecs gv;
  for (int k = 0; k < count; ++k) {
    if (!vecs[k]) return fail(EMU_SV_EINVAL, "arnoldi_gram: null basis vector");
    if (vecs[k] == w) return fail(EMU_SV_EINVAL, "arnoldi_gram: w aliases a basis vector");
    gv.u[k] = vecs[k];
  }
  const uint64_t npk = packets(n, dtype);
  const uint64_t chunk = (uint64_t)kThreads * kGramEpt;
  const uint64_t nchunks = (npk + chunk - 1) / chunk;
  const int nb1 = (int)(nchunks < (uint64_t)kGramBlocks ? nchunks : (uint64_t)kGramBlocks);
  const dim3 grid1((unsigned)nb1, (unsigned)((count + kGramGroup - 1) / kGramGroup));
  double *part = state + gram_off_part(ld);
  const uint64_t stride = gram_part_stride(ld);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(arnoldi_gram_dots_kernel<double>, grid1, dim3(kThreads), 0, stream, npk,
                       count, gv, w, part, stride);
  else
    hipLaunchKernelGGL(arnoldi_gram_dots_kernel<float>, grid1, dim3(kThreads), 0, stream, npk,
                       count, gv, w, part, stride);
  if (int rc = hip_check(hipGetLastError(), "arnoldi_gram overlaps launch")) return rc;
  hipLaunchKernelGGL(arnoldi_gram_solve_kernel, dim3(1), dim3(kGramSolveThreads), 0, stream, count,
                     nb1, ld, v0_scale, state, records_out);
  if (int rc = hip_check(hipGetLastError(), "arnoldi_gram solve launch")) return rc;
  const int nb3 = stream_blocks(npk);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(arnoldi_gram_update_kernel<double>, dim3(nb3), dim3(kThreads), 0, stream,
                       npk, count, gv, w, state, ld, part);
  else
    hipLaunchKernelGGL(arnoldi_gram_update_kernel<float>, dim3(nb3), dim3(kThreads), 0, stream,
                       npk, count, gv, w, state, ld, part);
  if (int rc = hip_check(hipGetLastError(), "arnoldi_gram update launch")) return rc;
  hipLaunchKernelGGL(arnoldi_gram_norm_kernel, dim3(1), dim3(kThreads), 0, stream, nb3, part,
                     records_out + 4 * count, state + count);
  return hip_check(hipGetLastError(), "arnoldi_gram norm launch");
}

int emu_sv_arnoldi_orthogonalize(int n, int dtype, int count, const void *const *vecs,
                                 const double *scales, void *w, double *records, void *workspace,
                                 hipStream_t stream) {
  return arnoldi_mgs(n, dtype, count, vecs, scales, w, records, workspace, 0, stream);
}

int emu_sv_arnoldi_orthonormalize(int n, int dtype, int count, const void *const *vecs,
                                  const double *scales, void *w, double *records, void *workspace,
                                  hipStream_t stream) {
  return arnoldi_mgs(n, dtype, count, vecs, scales, w, records, workspace, 1, stream);
}

// The whole Arnoldi exponential of one step on the library's host side: the
// Arnoldi branch of krylov_exp (emu_base/math/krylov_exp.py:44-107 with
// is_hermitian=False), pipelined.  Iteration j enqueues H.psi of the unit
// vector v_j, the Gram-Schmidt of the result with its normalisation on the
// device and an asynchronous copy of its records; iteration j+1 is enqueued
// before the host waits for iteration j (unless half the last error estimate
// is already below the tolerance), so the host's small exponential and test
// overlap the device's next matvec.  No per-iteration interpreter work.
namespace {
// Wait for an event by polling it: the library's drivers wait once per
// Krylov iteration for a few tens of microseconds of GPU work, where a
// blocking wait's wake-up costs more than the poll.
int wait_event(hipEvent_t e, const char *what) {
  for (;;) {
    const hipError_t q = hipEventQuery(e);
    if (q == hipSuccess) return EMU_SV_OK;
    if (q != hipErrorNotReady) return hip_check(q, what);
  }
}

struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  ~EventPair() {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
  }
};
}  // namespace

size_t emu_sv_arnoldi_exp_records_doubles(int max_dim) {
  if (max_dim < 1 || max_dim > kArnoldiMaxDim) return 0;
  return ((size_t)max_dim * (max_dim + 2) + 1) * 4 + gram_state_doubles(max_dim + 1);
}

int emu_sv_arnoldi_exp(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                       void *v, void *out, void *basis, int max_dim, double exp_tolerance,
                       double norm_tolerance, double *records, double *host_records,
                       void *workspace, int tile_log, int min_contig, int flags,
                       hipStream_t stream, int *result_out) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!ham || !v || !out || !basis || !records || !host_records || !workspace || !result_out)
    return fail(EMU_SV_EINVAL, "arnoldi_exp: null pointer");
  if (max_dim < 1 || max_dim > kArnoldiMaxDim)
    return fail(EMU_SV_EINVAL, "arnoldi_exp: max_dim must be in [1, 256]");
  if (out == basis) return fail(EMU_SV_EINVAL, "arnoldi_exp: out must not alias the basis");
  const size_t vbytes = (dtype == EMU_SV_C128 ? sizeof(double2) : sizeof(float2)) << n;
  const size_t row = (size_t)(max_dim + 2) * 4;  // doubles per iteration's records
  double *dot_dev = records + (size_t)max_dim * row;
  double *dot_host = host_records + (size_t)max_dim * row;
  // Gram-Schmidt: the one-workgroup kernel for 2^10..2^13 elements, else the
  // Gram-matrix form (unnormalised basis, scales on the device)
  const bool use_gram = !(n >= 10 && n <= kMgsSmallLog && max_dim <= kMgsSmallCount);
  const int gld = max_dim + 1;
  double *gstate = dot_dev + 4;
  const int hflags = flags & ~EMU_SV_FLAG_KEEP_DIAG;
  EventPair ev;
  for (hipEvent_t &x : ev.e)
    if (int rc = hip_check(hipEventCreateWithFlags(&x, hipEventDisableTiming), "arnoldi_exp: event")) return rc;

  // initial norm: the only synchronisation outside the iterations
  if (int rc = emu_sv_dot(n, dtype, v, nullptr, dot_dev, workspace, stream)) return rc;
  if (int rc = hip_check(hipMemcpyAsync(dot_host, dot_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, stream),
                         "arnoldi_exp: copy")) return rc;
  if (int rc = hip_check(hipStreamSynchronize(stream), "arnoldi_exp: synchronize")) return rc;
  const double initial_norm = std::sqrt(std::max(dot_host[0], 0.0));
  if (!(initial_norm > 0.0)) return fail(EMU_SV_EINVAL, "arnoldi_exp: zero input vector");

  std::vector<const void *> vec((size_t)max_dim + 1);
  std::vector<double> sc((size_t)max_dim + 1, 1.0);  // later basis vectors are unit vectors
  vec[0] = v;
  sc[0] = 1.0 / initial_norm;
  int enqueued = 0, matvecs = 0;
  auto enqueue = [&](int j) -> int {
    void *w = static_cast<char *>(basis) + (size_t)j * vbytes;
    t_diag_known_fresh = j > 0;
    const int hrc = hpsi_impl(n, dtype, ham, sigma_re, sigma_im, vec[j], use_gram ? 1.0 : sc[j], nullptr,
                              0.0, w, nullptr, workspace, tile_log, min_contig,
                              hflags | (j > 0 ? EMU_SV_FLAG_KEEP_DIAG : 0), stream);
    t_diag_known_fresh = false;
    if (hrc) return hrc;
    double *rec = records + (size_t)j * row;
    if (use_gram) {
      if (int rc = emu_sv_arnoldi_gram(n, dtype, j + 1, vec.data(), w, sc[0], gstate, gld, rec, stream))
        return rc;
    } else if (int rc = arnoldi_mgs(n, dtype, j + 1, vec.data(), sc.data(), w, rec, workspace, 1, stream)) {
      return rc;
    }
    if (int rc = hip_check(hipMemcpyAsync(host_records + (size_t)j * row, rec, (size_t)(j + 2) * 4 * sizeof(double),
                                          hipMemcpyDeviceToHost, stream), "arnoldi_exp: copy"))
      return rc;
    if (int rc = hip_check(hipEventRecord(ev.e[j & 1], stream), "arnoldi_exp: event")) return rc;
    vec[j + 1] = w;
    enqueued = j + 1;
    ++matvecs;
    return EMU_SV_OK;
  };
  using cd = std::complex<double>;
  const int dim = max_dim + 2;
  std::vector<cd> T((size_t)dim * dim, cd(0.0, 0.0));
  std::vector<double> sub, col;
  // col = expm(T[:m, :m]) e_0
  auto expm_col = [&](int m) -> int {
    sub.assign((size_t)2 * m * m, 0.0);
    for (int r = 0; r < m; ++r)
      for (int c = 0; c < m; ++c) {
        sub[2 * ((size_t)r * m + c)] = T[(size_t)r * dim + c].real();
        sub[2 * ((size_t)r * m + c) + 1] = T[(size_t)r * dim + c].imag();
      }
    col.assign((size_t)2 * m, 0.0);
    if (emu_sv_expm_column(m, sub.data(), col.data()) != EMU_SV_OK)
      return fail(EMU_SV_EINVAL, "arnoldi_exp: small exponential failed");
    return EMU_SV_OK;
  };
  // out = initial_norm sum_k col_k sc_k V_k over the first count vectors
  auto combine = [&](int count) -> int {
    std::vector<double> coef((size_t)2 * count);
    for (int k = 0; k < count; ++k) {
      coef[2 * k] = col[2 * k] * sc[k] * initial_norm;
      coef[2 * k + 1] = col[2 * k + 1] * sc[k] * initial_norm;
    }
    return emu_sv_lincomb(n, dtype, count, vec.data(), coef.data(), out, stream);
  };
  auto finish = [&](int count, int iterations, int converged, int happy) -> int {
    if (int rc = combine(count)) return rc;
    result_out[0] = iterations;
    result_out[1] = converged;
    result_out[2] = happy;
    result_out[3] = matvecs;
    return EMU_SV_OK;
  };

  if (int rc = enqueue(0)) return rc;
  double last_err = -1.0;
  for (int j = 0; j < max_dim; ++j) {
    const bool likely_done = last_err >= 0.0 && 0.5 * last_err < exp_tolerance;
    if (j + 1 < max_dim && !likely_done)
      if (int rc = enqueue(j + 1)) return rc;
    if (int rc = wait_event(ev.e[j & 1], "arnoldi_exp: wait")) return rc;
    const double *rec = host_records + (size_t)j * row;
    const double nrm = std::sqrt(std::max(rec[0], 0.0));  // ||op(v_j)||, before orthogonalisation
    // Gram form: the records hold T[k, j] itself, and the next (unnormalised)
    // basis vector's scale is 1 / n2
    for (int k = 0; k <= j; ++k)
      T[(size_t)k * dim + j] = (use_gram ? 1.0 : sc[k]) * cd(rec[4 * k + 1], rec[4 * k + 2]);
    const double n2 = std::sqrt(std::max(rec[4 * (j + 1)], 0.0));
    T[(size_t)(j + 1) * dim + j] = n2;
    if (use_gram) sc[j + 1] = n2 > 0.0 ? 1.0 / n2 : 0.0;
    if (n2 < norm_tolerance) {  // happy breakdown
      if (int rc = expm_col(j + 1)) return rc;
      return finish(j + 1, j + 1, 1, 1);
    }
    T[(size_t)(j + 2) * dim + j + 1] = 1.0;
    if (int rc = expm_col(j + 3)) return rc;
    const double err1 = std::abs(cd(col[2 * (j + 1)], col[2 * (j + 1) + 1]));
    const double err2 = std::abs(cd(col[2 * (j + 2)], col[2 * (j + 2) + 1])) * nrm;
    const double err = err1 < err2 ? err1 : err1 * err2 / (err1 - err2);
    if (err < exp_tolerance) return finish(j + 2, j + 1, 1, 0);
    last_err = err;
    if (j + 1 < max_dim && enqueued == j + 1)
      if (int rc = enqueue(j + 1)) return rc;  // the speculation was skipped and j did not converge
  }
  return finish(max_dim + 1, max_dim, 0, 0);
}

// ---------------------------------------------------------------------------
// One Krylov time step of the Hermitian (Lanczos) branch with its loop in the
// library: the pipelined engine of the Python package in native code.
// Records: rows of kStepRow doubles; row j (iteration j) holds [0, 5) the
// orthogonalisation sums, [5, 12) the last pass's dots, [16, 22) the device
// coefficients the iteration reads, [24, 30) the side outputs (n2, b, a,
// ||r||^2 / n2^2 of iteration j+1); row max_dim + 2 the opening dot and
// H.psi reductions.
// ---------------------------------------------------------------------------
constexpr int kStepRow = 32;

size_t emu_sv_krylov_step_records_doubles(int max_dim) {
  if (max_dim < 1 || max_dim > kArnoldiMaxDim) return 0;
  return (size_t)kStepRow * (max_dim + 3);
}

size_t emu_sv_krylov_step_decomp_doubles(int max_dim) {
  if (max_dim < 1 || max_dim > kArnoldiMaxDim) return 0;
  return 4 + 8 * (size_t)(max_dim + 2);
}

int emu_sv_krylov_step(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                       const void *v, void *out, void *basis, int max_dim, double exp_tolerance,
                       double norm_tolerance, double *records, double *host_records,
                       void *workspace, int tile_log, int min_contig, int flags,
                       hipStream_t stream, int *result_out, double *decomp_out) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!ham || !v || !out || !basis || !records || !host_records || !workspace || !result_out)
    return fail(EMU_SV_EINVAL, "krylov_step: null pointer");
  if (max_dim < 1 || max_dim > kArnoldiMaxDim)
    return fail(EMU_SV_EINVAL, "krylov_step: max_dim must be in [1, 256]");
  if (flags & EMU_SV_FLAG_NOISE)
    return fail(EMU_SV_EINVAL, "krylov_step: the Lanczos recurrence needs a Hermitian H (no noise term)");
  const int kind = kernel_kind(flags);
  if (kind != kKernelTile && kind != kKernelLds && kind != kKernelBtile)
    return fail(EMU_SV_EINVAL, "krylov_step: needs the tile, btile or LDS kernel kind");
  const size_t vbytes = (dtype == EMU_SV_C128 ? sizeof(double2) : sizeof(float2)) << n;
  // basis row k (k >= 1) holds w_{k-1} and then, orthogonalised in place,
  // basis vector k; rows 1 .. max_dim + 1
  auto row_ptr = [&](int k) -> void * { return static_cast<char *>(basis) + (size_t)(k - 1) * vbytes; };
  {
    const char *o = static_cast<const char *>(out), *b = static_cast<const char *>(basis);
    if (o >= b && o < b + (size_t)(max_dim + 1) * vbytes)
      return fail(EMU_SV_EINVAL, "krylov_step: out must not alias the basis");
  }
  double *open_dev = records + (size_t)(max_dim + 2) * kStepRow;
  double *open_host = host_records + (size_t)(max_dim + 2) * kStepRow;
  const int base_flags = flags & ~EMU_SV_FLAG_KEEP_DIAG;
  const std::complex<double> sig(sigma_re, sigma_im);

  // initial norm and the opening H.psi with its dots: the two synchronisations
  // outside the iterations
  if (int rc = emu_sv_dot(n, dtype, v, nullptr, open_dev, workspace, stream)) return rc;
  if (int rc = hip_check(hipMemcpyAsync(open_host, open_dev, 3 * sizeof(double), hipMemcpyDeviceToHost, stream),
                         "krylov_step: copy")) return rc;
  if (int rc = hip_check(hipStreamSynchronize(stream), "krylov_step: synchronize")) return rc;
  const double initial_norm = std::sqrt(std::max(open_host[0], 0.0));
  if (!(initial_norm > 0.0)) return fail(EMU_SV_EINVAL, "krylov_step: zero input vector");
  const double s0 = 1.0 / initial_norm;
  if (int rc = hpsi_impl(n, dtype, ham, sigma_re, sigma_im, v, s0, nullptr, 0.0, row_ptr(1), open_dev + 8,
                         workspace, tile_log, min_contig, base_flags, stream)) return rc;
  if (int rc = hip_check(hipMemcpyAsync(open_host + 8, open_dev + 8, 7 * sizeof(double),
                                        hipMemcpyDeviceToHost, stream), "krylov_step: copy")) return rc;
  if (int rc = hip_check(hipStreamSynchronize(stream), "krylov_step: synchronize")) return rc;
  const std::complex<double> b0(open_host[9], open_host[10]);
  double ww = open_host[8];  // ||op(v_0)||^2
  // iteration 0's coefficients: a = 0, b = b0 (pb = v with scale s0), w_scale = 1
  double *c0 = host_records + 16;
  c0[0] = 0.0;
  c0[1] = 0.0;
  c0[2] = b0.real() * s0;
  c0[3] = b0.imag() * s0;
  c0[4] = s0;
  c0[5] = 1.0;
  if (int rc = hip_check(hipMemcpyAsync(records + 16, c0, 6 * sizeof(double), hipMemcpyHostToDevice, stream),
                         "krylov_step: copy")) return rc;

  std::vector<double> tdiag((size_t)2 * (max_dim + 2), 0.0), tsub((size_t)max_dim + 2, 0.0);
  std::vector<double> tsup((size_t)2 * (max_dim + 2), 0.0);  // T[k, k+1]
  tdiag[0] = b0.real();
  tdiag[1] = b0.imag();
  std::vector<double> sc((size_t)max_dim + 2, 0.0);
  sc[0] = s0;
  std::vector<const void *> vec((size_t)max_dim + 2, nullptr);
  vec[0] = v;
  EventPair ev;
  for (hipEvent_t &x : ev.e)
    if (int rc = hip_check(hipEventCreateWithFlags(&x, hipEventDisableTiming), "krylov_step: event")) return rc;
  // Pinned host records take the iterations' reductions straight from the
  // kernel that forms them, so no copy command sits between two iterations
  // (at 12-20 qubits an iteration is 40-80 us of GPU work, and each copy
  // and the launch gap behind it cost about 10 us).  Pageable records keep
  // the asynchronous copy.
  double *host_rec_dev = device_view_of_pinned(host_records);
  int enqueued = 0, matvecs = 1;
  auto enqueue = [&](int j) -> int {
    double *rec = records + (size_t)j * kStepRow;
    // where the reductions land: the host row itself, or the device row
    double *out = host_rec_dev ? host_rec_dev + (size_t)j * kStepRow : rec;
    // iteration j reads w_j (row j + 1), orthogonalises it against basis
    // vectors j - 1 and j and writes r_j = w_{j+1} into row j + 2
    const void *pa = j > 0 ? (j == 1 ? v : row_ptr(j - 1)) : nullptr;
    const void *pb = j == 0 ? v : row_ptr(j);
    // the opening H.psi built the first-pass diagonal table on this
    // workspace and stream: the iterations skip its stamp check
    t_diag_known_fresh = true;
    const int irc = lanczos_iterate_impl(n, dtype, ham, sigma_re, sigma_im, row_ptr(j + 1), pa, pb,
                                               rec + 16, row_ptr(j + 2), out, out + 5,
                                               records + (size_t)(j + 1) * kStepRow + 16, out + 24,
                                               workspace, tile_log, min_contig,
                                               base_flags | EMU_SV_FLAG_KEEP_DIAG, stream);
    t_diag_known_fresh = false;
    if (irc) return irc;
    if (!host_rec_dev)
      if (int rc = hip_check(hipMemcpyAsync(host_records + (size_t)j * kStepRow, rec, kStepRow * sizeof(double),
                                            hipMemcpyDeviceToHost, stream), "krylov_step: copy"))
        return rc;
    if (int rc = hip_check(hipEventRecord(ev.e[j & 1], stream), "krylov_step: event")) return rc;
    enqueued = j + 1;
    ++matvecs;
    return EMU_SV_OK;
  };
  std::vector<double> ex;
  // exp(T) of the leading m x m block; with coefficients: the combination
  auto texp = [&](int m, double n2, int want) -> int {
    ex.assign(want ? 2 + 2 * (size_t)(m + 1) : 2, 0.0);
    if (emu_sv_lanczos_exp(m, tdiag.data(), tsub.data(), sigma_re, sigma_im, n2, want, ex.data()) != EMU_SV_OK)
      return fail(EMU_SV_EINVAL, "krylov_step: small exponential failed");
    return EMU_SV_OK;
  };
  auto finish = [&](int m, double n2, int happy, int converged) -> int {
    if (int rc = texp(m, n2, 1)) return rc;
    const int count = happy ? m : m + 1;
    std::vector<double> coef((size_t)2 * count);
    for (int k = 0; k < count; ++k) {
      coef[2 * k] = ex[2 + 2 * k] * sc[k] * initial_norm;
      coef[2 * k + 1] = ex[3 + 2 * k] * sc[k] * initial_norm;
    }
    if (int rc = emu_sv_lincomb(n, dtype, count, vec.data(), coef.data(), out, stream)) return rc;
    result_out[0] = m;
    result_out[1] = converged;
    result_out[2] = happy;
    result_out[3] = matvecs;
    if (decomp_out) {  // the Lanczos decomposition (layout: emu_sv_krylov_step_decomp_doubles)
      const size_t M = (size_t)max_dim + 2;
      double *d = decomp_out;
      d[0] = m;
      d[1] = happy ? 0.0 : n2;
      d[2] = happy;
      d[3] = converged;
      for (int k = 0; k < 2 * m; ++k) d[4 + k] = tdiag[k];
      for (int k = 0; k < m; ++k) d[4 + 2 * M + k] = tsub[k];
      for (int k = 0; k < 2 * m; ++k) d[4 + 3 * M + k] = tsup[k];
      for (int k = 0; k < count; ++k) d[4 + 5 * M + k] = sc[k];
      for (int k = 0; k < 2 * count; ++k) d[4 + 6 * M + k] = ex[2 + k];
    }
    return EMU_SV_OK;
  };

  if (int rc = enqueue(0)) return rc;
  double last_err = -1.0;
  for (int j = 0; j < max_dim; ++j) {
    // speculate on iteration j+1 unless iteration j is likely to converge
    // (half the last error estimate below the tolerance)
    const bool likely_done = last_err >= 0.0 && 0.5 * last_err < exp_tolerance;
    if (j + 1 < max_dim && !likely_done)
      if (int rc = enqueue(j + 1)) return rc;
    if (int rc = wait_event(ev.e[j & 1], "krylov_step: wait")) return rc;
    const double *rec = host_records + (size_t)j * kStepRow;
    const double nrm = std::sqrt(std::max(ww, 0.0));  // ||op(v_j)||, before orthogonalisation
    const double n2 = rec[24];
    tsub[j] = n2;
    if (n2 < norm_tolerance) return finish(j + 1, 0.0, 1, 1);  // happy breakdown
    vec[j + 1] = row_ptr(j + 1);
    sc[j + 1] = 1.0 / n2;
    if (int rc = texp(j + 1, n2, 0)) return rc;
    const double e1 = ex[0], err2 = ex[1] * nrm;
    const double err = e1 < err2 ? e1 : e1 * err2 / (e1 - err2);
    if (err < exp_tolerance) return finish(j + 1, n2, 0, 1);
    last_err = err;
    tdiag[2 * (j + 1)] = rec[25];
    tdiag[2 * (j + 1) + 1] = rec[26];
    tsup[2 * j] = rec[27];
    tsup[2 * j + 1] = rec[28];
    ww = rec[29];
    if (j + 1 < max_dim && enqueued == j + 1)
      if (int rc = enqueue(j + 1)) return rc;  // the speculation was skipped and j did not converge
  }
  return finish(max_dim, tsub[max_dim - 1], 0, 0);
}

int emu_sv_apply_1q(int n, int dtype, void *psi, int bit, const double *a, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!psi || !a) return fail(EMU_SV_EINVAL, "apply_1q: null pointer");
  if (bit < 0 || bit >= n) return fail(EMU_SV_EINVAL, "apply_1q: bit must be in [0, n)");
  Op1q op;
  for (int k = 0; k < 8; ++k) op.a[k] = a[k];
  const uint64_t npairs = 1ull << (n - 1);
  const int nb = stream_blocks(npairs);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(apply_1q_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, npairs, psi,
                       bit, op);
  else
    hipLaunchKernelGGL(apply_1q_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, npairs, psi,
                       bit, op);
  return hip_check(hipGetLastError(), "apply_1q launch");
}

int emu_sv_apply_pair(int n, int dtype, const void *in, void *out, int bit_a, int bit_b,
                      const double *op, int accumulate, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!in || !out || !op) return fail(EMU_SV_EINVAL, "apply_pair: null pointer");
  if (n < 2 || bit_a < 0 || bit_a >= n || bit_b < 0 || bit_b >= n || bit_a == bit_b)
    return fail(EMU_SV_EINVAL, "apply_pair: bits must be distinct and in [0, n)");
  Op2q o;
  for (int k = 0; k < 32; ++k) o.s[k] = op[k];
  const uint64_t ngroups = 1ull << (n - 2);
  const int nb = stream_blocks(ngroups);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(pair_op_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, ngroups, in,
                       out, bit_a, bit_b, o, accumulate ? 1 : 0);
  else
    hipLaunchKernelGGL(pair_op_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, ngroups, in,
                       out, bit_a, bit_b, o, accumulate ? 1 : 0);
  return hip_check(hipGetLastError(), "apply_pair launch");
}

int emu_sv_apply_pairs_all(int n_side, int dtype, const void *in, void *out, const double *op,
                           hipStream_t stream) {
  if (n_side < 1 || n_side > 8) return fail(EMU_SV_EINVAL, "apply_pairs_all: n_side must be in [1, 8]");
  if (int rc = check_dtype(dtype)) return rc;
  if (!in || !out || !op) return fail(EMU_SV_EINVAL, "apply_pairs_all: null pointer");
  if (in == out) return fail(EMU_SV_EINVAL, "apply_pairs_all: in and out must not alias");
  Op2q o;
  for (int k = 0; k < 32; ++k) o.s[k] = op[k];
  const uint64_t D = 1ull << (2 * n_side);
  const int nb = stream_blocks(D);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(pairs_all_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, n_side, in,
                       out, o);
  else
    hipLaunchKernelGGL(pairs_all_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, n_side, in,
                       out, o);
  return hip_check(hipGetLastError(), "apply_pairs_all launch");
}

int emu_sv_antihermitian(int side_log, int dtype, const void *m, void *out, double s_re,
                         double s_im, hipStream_t stream) {
  if (side_log < 0 || side_log > EMU_SV_MAX_QUBITS / 2)
    return fail(EMU_SV_EINVAL, "antihermitian: side_log must be in [0, 16]");
  if (int rc = check_dtype(dtype)) return rc;
  if (!m || !out) return fail(EMU_SV_EINVAL, "antihermitian: null pointer");
  const uint64_t side = 1ull << side_log;
  const uint64_t ts = side < (uint64_t)kAhTile ? side : (uint64_t)kAhTile;
  const unsigned nt = (unsigned)(side / ts);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(antiherm_kernel<double>, dim3(nt, nt), dim3(kThreads), 0, stream, side_log,
                       m, out, s_re, s_im);
  else
    hipLaunchKernelGGL(antiherm_kernel<float>, dim3(nt, nt), dim3(kThreads), 0, stream, side_log,
                       m, out, s_re, s_im);
  return hip_check(hipGetLastError(), "antihermitian launch");
}

size_t emu_sv_sample_workspace_bytes(int n) {
  if (n < 1 || n > EMU_SV_MAX_QUBITS) return 0;
  const uint64_t D = 1ull << n;
  const uint64_t chunk = D < (1ull << kSampleChunkLog) ? D : (1ull << kSampleChunkLog);
  return (size_t)(D / chunk) * sizeof(double);
}

int emu_sv_sample(int n, int dtype, const void *src, uint64_t stride, int mode,
                  const double *uniforms, int count, int64_t *indices_out, void *workspace,
                  size_t workspace_bytes, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!src || !uniforms || !indices_out || !workspace)
    return fail(EMU_SV_EINVAL, "sample: null pointer");
  if (mode != 0 && mode != 1) return fail(EMU_SV_EINVAL, "sample: mode must be 0 or 1");
  if (stride < 1) return fail(EMU_SV_EINVAL, "sample: stride must be >= 1");
  if (count < 0) return fail(EMU_SV_EINVAL, "sample: count must be >= 0");
  if (workspace_bytes < emu_sv_sample_workspace_bytes(n))
    return fail(EMU_SV_EINVAL, "sample: workspace too small");
  if (count == 0) return EMU_SV_OK;
  const uint64_t D = 1ull << n;
  const uint64_t chunk = D < (1ull << kSampleChunkLog) ? D : (1ull << kSampleChunkLog);
  const uint64_t nchunks = D / chunk;
  double *sums = reinterpret_cast<double *>(workspace);
  const unsigned g1 = (unsigned)((nchunks + kWaves - 1) / kWaves);
  const unsigned g3 = (unsigned)((count + kWaves - 1) / kWaves);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(sample_chunk_kernel<double>, dim3(g1), dim3(kThreads), 0, stream, D,
                       nchunks, src, stride, mode, sums);
  else
    hipLaunchKernelGGL(sample_chunk_kernel<float>, dim3(g1), dim3(kThreads), 0, stream, D,
                       nchunks, src, stride, mode, sums);
  if (int rc = hip_check(hipGetLastError(), "sample chunk launch")) return rc;
  hipLaunchKernelGGL(sample_scan_kernel, dim3(1), dim3(kThreads), 0, stream, nchunks, sums);
  if (int rc = hip_check(hipGetLastError(), "sample scan launch")) return rc;
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(sample_draw_kernel<double>, dim3(g3), dim3(kThreads), 0, stream, D,
                       nchunks, src, stride, mode, sums, uniforms, count, indices_out);
  else
    hipLaunchKernelGGL(sample_draw_kernel<float>, dim3(g3), dim3(kThreads), 0, stream, D,
                       nchunks, src, stride, mode, sums, uniforms, count, indices_out);
  return hip_check(hipGetLastError(), "sample draw launch");
}

int emu_sv_scale(int n, int dtype, void *v, double s, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!v) return fail(EMU_SV_EINVAL, "scale: null pointer");
  const uint64_t npk = packets(n, dtype);
  const int nb = stream_blocks(npk);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(scale_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, npk, v, s);
  else
    hipLaunchKernelGGL(scale_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, npk, v, (float)s);
  return hip_check(hipGetLastError(), "scale launch");
}

int emu_sv_convert(int n, int src_dtype, const void *src, int dst_dtype, void *dst,
                   hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(src_dtype)) return rc;
  if (int rc = check_dtype(dst_dtype)) return rc;
  if (!src || !dst) return fail(EMU_SV_EINVAL, "convert: null pointer");
  const uint64_t D = 1ull << n;
  const int nb = stream_blocks(D);
  if (src_dtype == EMU_SV_C128 && dst_dtype == EMU_SV_C64)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(nb), dim3(kThreads), 0, stream, D,
                       (const double2 *)src, (float2 *)dst);
  else if (src_dtype == EMU_SV_C64 && dst_dtype == EMU_SV_C128)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(nb), dim3(kThreads), 0, stream, D,
                       (const float2 *)src, (double2 *)dst);
  else {
    const uint64_t npk = packets(n, src_dtype);
    const uint64_t per = 4ull * kThreads;
    hipLaunchKernelGGL(copy_packets_kernel, dim3((unsigned)((npk + per - 1) / per)), dim3(kThreads), 0,
                       stream, npk, (const uint4 *)src, (uint4 *)dst);
  }
  return hip_check(hipGetLastError(), "convert launch");
}

int emu_sv_correlation(int n, int dtype, const void *psi, double *corr_out, void *workspace,
                       hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!psi || !corr_out || !workspace) return fail(EMU_SV_EINVAL, "correlation: null pointer");
  const int K = n < kMaxTileLog ? n : kMaxTileLog;
  const uint64_t nchunks = 1ull << (n - K);
  const int nb = (int)(nchunks < (uint64_t)kCorrBlocks ? nchunks : (uint64_t)kCorrBlocks);
  double *part = partials_ptr(workspace, n);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(correlation_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, n,
                       (const double2 *)psi, part);
  else
    hipLaunchKernelGGL(correlation_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, n,
                       (const float2 *)psi, part);
  if (int rc = hip_check(hipGetLastError(), "correlation launch")) return rc;
  hipLaunchKernelGGL(corr_finalize_kernel, dim3(1), dim3(kThreads), 0, stream, n, part, nb,
                     corr_out);
  return hip_check(hipGetLastError(), "correlation finalize launch");
}

int emu_sv_parameter_traces(int n, int dtype, int count, const void *w, const void *z,
                            uint64_t stride, int want_pairs, double *out, void *workspace,
                            hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (count < 1) return fail(EMU_SV_EINVAL, "parameter_traces: count must be >= 1");
  if (!w || !z || !out || !workspace) return fail(EMU_SV_EINVAL, "parameter_traces: null pointer");
  if (count > 1 && stride < (1ull << n))
    return fail(EMU_SV_EINVAL, "parameter_traces: rows overlap (stride < 2^n)");
  const uint64_t D = 1ull << n;
  const uint64_t ch = D < (uint64_t)kTraceChunk ? D : (uint64_t)kTraceChunk;
  const uint64_t nchunks = D / ch;
  uint64_t gx = (uint64_t)(kTraceBlocks / n);
  if (gx > nchunks) gx = nchunks;
  if (gx < 1) gx = 1;
  double *part = partials_ptr(workspace, n);
  const dim3 grid((unsigned)gx, (unsigned)n);
  if (dtype == EMU_SV_C128) {
    if (want_pairs)
      hipLaunchKernelGGL((traces_kernel<double, true>), grid, dim3(kThreads), 0, stream, n, count,
                         (const double2 *)w, (const double2 *)z, stride, part);
    else
      hipLaunchKernelGGL((traces_kernel<double, false>), grid, dim3(kThreads), 0, stream, n, count,
                         (const double2 *)w, (const double2 *)z, stride, part);
  } else {
    if (want_pairs)
      hipLaunchKernelGGL((traces_kernel<float, true>), grid, dim3(kThreads), 0, stream, n, count,
                         (const float2 *)w, (const float2 *)z, stride, part);
    else
      hipLaunchKernelGGL((traces_kernel<float, false>), grid, dim3(kThreads), 0, stream, n, count,
                         (const float2 *)w, (const float2 *)z, stride, part);
  }
  if (int rc = hip_check(hipGetLastError(), "parameter_traces launch")) return rc;
  const int nout = want_pairs ? 6 * n + 2 * n * n : 6 * n;
  hipLaunchKernelGGL(traces_finalize_kernel, dim3((nout + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     stream, n, (int)gx, want_pairs, (const double *)part, out);
  return hip_check(hipGetLastError(), "parameter_traces finalize launch");
}

size_t emu_sv_batch_workspace_bytes(int count, int dtype, const emu_sv_batch_register *regs,
                                    int tile_log, int min_contig) {
  if (count < 1 || !regs || check_dtype(dtype)) return 0;
  if (batch_tile_logs(dtype, tile_log, min_contig)) return 0;
  BatchLayout L;
  if (batch_layout(count, regs, dtype, tile_log, min_contig, L, nullptr)) return 0;
  return L.total;
}

int emu_sv_batch_create(int count, int dtype, const emu_sv_batch_register *regs, int max_dim,
                        double tolerance, int tile_log, int min_contig, void *workspace,
                        size_t workspace_bytes, hipStream_t stream, void **handle_out) {
  if (!handle_out) return fail(EMU_SV_EINVAL, "batch_create: null handle pointer");
  *handle_out = nullptr;
  if (count < 1 || !regs || !workspace) return fail(EMU_SV_EINVAL, "batch_create: bad arguments");
  if (reinterpret_cast<uintptr_t>(workspace) & 255)
    return fail(EMU_SV_EINVAL, "batch_create: workspace must be 256-byte aligned");
  if (int rc = check_dtype(dtype)) return rc;
  if (max_dim < 2 || max_dim > kBatchMaxDim)
    return fail(EMU_SV_EINVAL, "batch_create: max_dim must be in [2, 62]");
  if (!(tolerance > 0.0)) return fail(EMU_SV_EINVAL, "batch_create: tolerance must be > 0");
  if (int rc = batch_tile_logs(dtype, tile_log, min_contig)) return rc;
  auto *h = new BatchHandle();
  BatchLayout L;
  if (int rc = batch_layout(count, regs, dtype, tile_log, min_contig, L, &h->host_regs)) {
    delete h;
    return rc;
  }
  if (workspace_bytes < L.total) {
    delete h;
    return fail(EMU_SV_EINVAL, "batch_create: workspace too small");
  }
  char *base = reinterpret_cast<char *>(workspace);
  h->count = count;
  h->dtype = dtype;
  h->max_dim = max_dim;
  h->tol = tolerance;
  h->regs = reinterpret_cast<BatchReg *>(base + L.regs);
  double *arrays = reinterpret_cast<double *>(base + L.arrays);
  double *partials = reinterpret_cast<double *>(base + L.partials);
  BatchEntry *tables = reinterpret_cast<BatchEntry *>(base + L.tables);
  h->counter = reinterpret_cast<int *>(base + L.counter);
  const size_t esz = dtype == EMU_SV_C64 ? sizeof(float2) : sizeof(double2);
  // host copies of everything the device needs
  std::vector<BatchEntry> tab_host;
  uint64_t toff[4], coff;
  {
    uint64_t o = 0;
    for (int p = 0; p < 4; ++p) {
      toff[p] = o;
      o += L.ntab[p];
    }
    coff = o;
    tab_host.resize(o + L.ncomb);
  }
  uint64_t tfill[4] = {0, 0, 0, 0}, cfill = 0, poff = 0;
  int max_t = 0;
  for (int i = 0; i < count; ++i) {
    BatchReg &R = h->host_regs[i];
    const emu_sv_batch_register &U = regs[i];
    R.ham = U.ham;
    R.dt = U.dt;
    R.psi[0] = U.psi;
    R.psi[1] = U.psi_alt;
    R.user_psi = U.psi;
    R.rows = U.rows;
    R.partials = partials + poff;
    poff += 2 * R.ntiles_max * kNumDots;
    R.arr = arrays + (size_t)i * BatchArr::kTotal;
    R.cur = 0;
    R.step = 0;
    R.j = 0;
    R.phase = kPhStart;
    R.m = 0;
    R.matvecs = 0;
    double dt0 = 0.0;
    if (int rc = hip_check(hipMemcpy(&dt0, U.dt, sizeof(double), hipMemcpyDeviceToHost),
                           "batch_create: read dt")) {
      delete h;
      return rc;
    }
    R.sig_re = 0.0;
    R.sig_im = -dt0;
    for (int p = 0; p < R.npass; ++p) {
      const int t = R.plan[p].c + R.plan[p].m;
      if (t > max_t) max_t = t;
      for (uint64_t g = 0; g < R.tiles[p]; ++g) tab_host[toff[p] + tfill[p]++] = {(uint32_t)i, (uint32_t)g};
      if (p + 1 > h->npass_max) h->npass_max = p + 1;
    }
    const uint64_t nch = ((1ull << R.n) + kBatchChunk - 1) / kBatchChunk;
    for (uint64_t c = 0; c < nch; ++c) tab_host[coff + cfill++] = {(uint32_t)i, (uint32_t)c};
  }
  for (int p = 0; p < 4; ++p) {
    h->ntab[p] = L.ntab[p];
    h->tab[p] = tables + toff[p];
  }
  h->ncomb = L.ncomb;
  h->comb = tables + coff;
  // the first pass stages U in the same dynamic area before its tile
  h->lds = std::max(esz << max_t, kBatchUStageBytes);
  int rc = EMU_SV_OK;
  if (h->lds > 65536) {  // dynamic LDS above 64 KiB is enabled per kernel
    rc = hip_check(hipFuncSetAttribute(dtype == EMU_SV_C128
                                           ? reinterpret_cast<const void *>(&batch_pass_kernel<double>)
                                           : reinterpret_cast<const void *>(&batch_pass_kernel<float>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)h->lds),
                   "batch_create: LDS attribute");
    if (rc) {
      delete h;
      return rc;
    }
  }
  std::vector<double> arr_zero((size_t)count * BatchArr::kTotal, 0.0);
  rc = hip_check(hipMemcpyAsync(h->regs, h->host_regs.data(), sizeof(BatchReg) * count,
                                    hipMemcpyHostToDevice, stream), "batch_create: upload");
  if (!rc) rc = hip_check(hipMemcpyAsync(arrays, arr_zero.data(), sizeof(double) * arr_zero.size(),
                                         hipMemcpyHostToDevice, stream), "batch_create: upload");
  if (!rc) rc = hip_check(hipMemcpyAsync(tables, tab_host.data(), sizeof(BatchEntry) * tab_host.size(),
                                         hipMemcpyHostToDevice, stream), "batch_create: upload");
  if (!rc) rc = hip_check(hipStreamSynchronize(stream), "batch_create: sync");
  if (!rc) rc = hip_check(hipStreamCreateWithFlags(&h->capture_stream, hipStreamNonBlocking),
                          "batch_create: stream");
  if (rc) {
    delete h;
    return rc;
  }
  *handle_out = h;
  return EMU_SV_OK;
}

int emu_sv_batch_run(void *handle, int max_iterations, int check_every, hipStream_t stream,
                     int *remaining_out) {
  auto *h = reinterpret_cast<BatchHandle *>(handle);
  if (!h || max_iterations < 0 || check_every < 1) return fail(EMU_SV_EINVAL, "batch_run: bad arguments");
  if (!h->exec) {
    // capture one batched iteration (passes, decide, combine) once
    hipStream_t cs = h->capture_stream;
    int rc = hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "batch_run: capture");
    if (!rc) rc = h->dtype == EMU_SV_C64 ? batch_enqueue_iteration<float>(*h, cs)
                                         : batch_enqueue_iteration<double>(*h, cs);
    hipGraph_t g = nullptr;
    const int rc2 = hip_check(hipStreamEndCapture(cs, &g), "batch_run: end capture");
    if (rc || rc2) {
      if (g) hipGraphDestroy(g);
      return rc ? rc : rc2;
    }
    h->graph = g;
    if (int rc3 = hip_check(hipGraphInstantiate(&h->exec, g, nullptr, nullptr, 0), "batch_run: instantiate"))
      return rc3;
  }
  int remaining = 1;
  for (int it = 0; it < max_iterations; ++it) {
    if (int rc = hip_check(hipGraphLaunch(h->exec, stream), "batch_run: graph launch")) return rc;
    if ((it + 1) % check_every == 0 || it + 1 == max_iterations) {
      hipLaunchKernelGGL(batch_count_kernel, dim3(1), dim3(kThreads), 0, stream, h->regs, h->count,
                         h->counter);
      if (int rc = hip_check(hipGetLastError(), "batch count launch")) return rc;
      if (int rc = hip_check(hipMemcpyAsync(&remaining, h->counter, sizeof(int), hipMemcpyDeviceToHost,
                                            stream), "batch_run: read counter")) return rc;
      if (int rc = hip_check(hipStreamSynchronize(stream), "batch_run: sync")) return rc;
      if (remaining == 0) break;
    }
  }
  if (remaining_out) *remaining_out = remaining;
  return EMU_SV_OK;
}

int emu_sv_batch_status(void *handle, int *phase_out, int *steps_out, int *matvecs_out,
                        hipStream_t stream) {
  auto *h = reinterpret_cast<BatchHandle *>(handle);
  if (!h) return fail(EMU_SV_EINVAL, "batch_status: null handle");
  std::vector<BatchReg> r(h->count);
  if (int rc = hip_check(hipMemcpyAsync(r.data(), h->regs, sizeof(BatchReg) * h->count,
                                        hipMemcpyDeviceToHost, stream), "batch_status: read")) return rc;
  if (int rc = hip_check(hipStreamSynchronize(stream), "batch_status: sync")) return rc;
  for (int i = 0; i < h->count; ++i) {
    if (phase_out) phase_out[i] = r[i].phase;
    if (steps_out) steps_out[i] = r[i].step;
    if (matvecs_out) matvecs_out[i] = r[i].matvecs;
  }
  return EMU_SV_OK;
}

int emu_sv_batch_publish(void *handle, hipStream_t stream) {
  auto *h = reinterpret_cast<BatchHandle *>(handle);
  if (!h) return fail(EMU_SV_EINVAL, "batch_publish: null handle");
  if (h->dtype == EMU_SV_C64)
    hipLaunchKernelGGL(batch_publish_kernel<float>, dim3((unsigned)h->ncomb), dim3(kThreads), 0, stream,
                       h->regs, h->comb);
  else
    hipLaunchKernelGGL(batch_publish_kernel<double>, dim3((unsigned)h->ncomb), dim3(kThreads), 0, stream,
                       h->regs, h->comb);
  return hip_check(hipGetLastError(), "batch publish launch");
}

int emu_sv_batch_resume(void *handle, hipStream_t stream) {
  auto *h = reinterpret_cast<BatchHandle *>(handle);
  if (!h) return fail(EMU_SV_EINVAL, "batch_resume: null handle");
  hipLaunchKernelGGL(batch_resume_kernel, dim3((h->count + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     stream, h->regs, h->count);
  return hip_check(hipGetLastError(), "batch_resume launch");
}

int emu_sv_batch_step_info(void *handle, int reg, int *info_out, double *scalars_out, double *arrays_out,
                           void **state_out, void **prev_state_out, void **rows_out, hipStream_t stream) {
  auto *h = reinterpret_cast<BatchHandle *>(handle);
  if (!h || reg < 0 || reg >= h->count || !info_out || !scalars_out || !arrays_out)
    return fail(EMU_SV_EINVAL, "batch_step_info: bad arguments");
  BatchReg r;
  if (int rc = hip_check(hipMemcpyAsync(&r, h->regs + reg, sizeof(BatchReg), hipMemcpyDeviceToHost, stream),
                         "batch_step_info: read")) return rc;
  if (int rc = hip_check(hipStreamSynchronize(stream), "batch_step_info: sync")) return rc;
  if (int rc = hip_check(hipMemcpyAsync(arrays_out, r.arr, sizeof(double) * BatchArr::kTotal,
                                        hipMemcpyDeviceToHost, stream), "batch_step_info: read")) return rc;
  if (int rc = hip_check(hipStreamSynchronize(stream), "batch_step_info: sync")) return rc;
  info_out[0] = r.phase;
  info_out[1] = r.step;
  info_out[2] = r.m;
  info_out[3] = r.happy;
  info_out[4] = r.matvecs;
  scalars_out[0] = r.initial_norm;
  if (state_out) *state_out = r.psi[r.cur];
  if (prev_state_out) *prev_state_out = r.psi[r.cur ^ 1];
  if (rows_out) *rows_out = r.rows;
  return EMU_SV_OK;
}

int emu_sv_batch_array_layout(int *offsets_out) {
  if (!offsets_out) return fail(EMU_SV_EINVAL, "batch_array_layout: null pointer");
  offsets_out[0] = BatchArr::kDiag;
  offsets_out[1] = BatchArr::kSub;
  offsets_out[2] = BatchArr::kScale;
  offsets_out[3] = BatchArr::kCoef;
  offsets_out[4] = BatchArr::kSup;
  offsets_out[5] = BatchArr::kTotal;
  return EMU_SV_OK;
}

int emu_sv_batch_destroy(void *handle) {
  auto *h = reinterpret_cast<BatchHandle *>(handle);
  if (!h) return EMU_SV_OK;
  if (h->exec) hipGraphExecDestroy(h->exec);
  if (h->graph) hipGraphDestroy(h->graph);
  if (h->capture_stream) hipStreamDestroy(h->capture_stream);
  delete h;
  return EMU_SV_OK;
}

int emu_sv_pack_hamiltonian_steps(int n, int steps, const double *omega, const double *delta,
                                  const double *phi, const double *interaction, double *blocks_out) {
  if (steps < 1 || !blocks_out) return fail(EMU_SV_EINVAL, "pack_hamiltonian_steps: bad arguments");
  for (int k = 0; k < steps; ++k)
    if (int rc = emu_sv_pack_hamiltonian(n, omega + (size_t)k * n, delta + (size_t)k * n,
                                         phi + (size_t)k * n, interaction,
                                         blocks_out + (size_t)k * EMU_SV_HAM_DOUBLES))
      return rc;
  return EMU_SV_OK;
}

}  // extern "C"
