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
#include <complex>
#include <mutex>
#include <string>
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
constexpr int kStreamBlocks = 1024;  // grid of the streaming kernels
constexpr int kDefaultTileLogC64 = 13;
constexpr int kDefaultTileLogC128 = 12;
constexpr int kDefaultContigC64 = 3;
constexpr int kDefaultContigC128 = 2;

// Offsets into the packed parameter block (see emu_sv_hip.h).
constexpr int kCre = 0, kCim = 32, kDelta = 64, kU = 96, kFlagComplex = 1120;

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
  double v_scale, prev_scale;
  double sig_re, sig_im;
  int n;               // qubits
  int c, s, m;         // tile bits: [0,c) U [s,s+m)
  uint32_t flip_mask;  // global bits whose flip terms this pass applies
  int diag;            // add the diagonal term
  int accumulate;      // out += ...
  int xcd;             // XCD-aware tile order
};

// ---------------------------------------------------------------------------
// H.psi pass kernel.  CPLX: complex flip coefficients (some phi != 0).
// DOTS: fuse the Lanczos partial sums (last pass).
// ---------------------------------------------------------------------------
template <typename T, bool CPLX, bool DOTS>
__global__ __launch_bounds__(kThreads) void hpsi_pass_kernel(const PassArgs a) {
  using T2 = typename Vec2<T>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T2 *tile = reinterpret_cast<T2 *>(smem_raw);
  __shared__ double s_ap[kMaxTileLog];
  __shared__ double s_U[kMaxTileLog][kMaxTileLog];
  __shared__ double s_B[1 << (kMaxTileLog - kThreadBits)];
  __shared__ double s_eout;
  __shared__ double s_red[kWaves * kNumDots];

  const int tid = threadIdx.x;
  const int c = a.c, s = a.s, m = a.m, n = a.n;
  const int t = c + m;
  const uint32_t tsize = 1u << t;
  const uint64_t ntiles = 1ull << (n - t);
  uint64_t g = blockIdx.x;
  if (a.xcd && ntiles >= 8) g = (g & 7) * (ntiles >> 3) + (g >> 3);
  const int nlo = s - c;
  const uint64_t base = ((g & ((1ull << nlo) - 1)) << c) | ((g >> nlo) << (s + m));
  const uint32_t cmask = (1u << c) - 1;
  auto gidx = [&](uint32_t l) -> uint64_t {
    return base | (uint64_t)(l & cmask) | ((uint64_t)(l >> c) << s);
  };
  auto gbit = [&](int i) -> int { return i < c ? i : s + (i - c); };

  const double *__restrict__ H = a.ham;
  const T2 *__restrict__ vin = reinterpret_cast<const T2 *>(a.v);
  const T vs = static_cast<T>(a.v_scale);

  // Stage the tile.
  for (uint32_t l = tid; l < tsize; l += kThreads) {
    T2 z = vin[gidx(l)];
    z.x *= vs;
    z.y *= vs;
    tile[l] = z;
  }
  if (a.diag) {
    if (tid < t) {
      const int b = gbit(tid);
      double acc = -H[kDelta + b];
      uint64_t bm = base;
      while (bm) {
        const int q = __ffsll((unsigned long long)bm) - 1;
        bm &= bm - 1;
        acc += H[kU + b * 32 + q];
      }
      s_ap[tid] = acc;
    }
    if (tid == 63) {
      double e = 0.0;
      uint64_t bm = base;
      while (bm) {
        const int q = __ffsll((unsigned long long)bm) - 1;
        bm &= bm - 1;
        e -= H[kDelta + q];
        uint64_t bm2 = bm;
        while (bm2) {
          const int q2 = __ffsll((unsigned long long)bm2) - 1;
          bm2 &= bm2 - 1;
          e += H[kU + q * 32 + q2];
        }
      }
      s_eout = e;
    }
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

  // Flip coefficients of the local bits (uniform values).
  uint32_t fm_local = 0;
  T cre[kMaxTileLog], cim[kMaxTileLog];
#pragma unroll
  for (int i = 0; i < kMaxTileLog; ++i) {
    cre[i] = 0;
    cim[i] = 0;
    if (i < t) {
      const int b = gbit(i);
      if ((a.flip_mask >> b) & 1u) {
        fm_local |= 1u << i;
        cre[i] = static_cast<T>(H[kCre + b]);
        cim[i] = static_cast<T>(H[kCim + b]);
      }
    }
  }
  const T sr = static_cast<T>(a.sig_re), si = static_cast<T>(a.sig_im);
  T2 *__restrict__ out = reinterpret_cast<T2 *>(a.out);
  const T2 *__restrict__ prevp = reinterpret_cast<const T2 *>(a.prev);
  const T ps = static_cast<T>(a.prev_scale);
  double dots[7] = {0, 0, 0, 0, 0, 0, 0};

  if (tid < (1 << tb)) {
    for (uint32_t r = 0; r < (1u << k); ++r) {
      const uint32_t l = (uint32_t)tid | (r << tb);
      const T2 psi = tile[l];
      T hr = 0, hi = 0;
#pragma unroll
      for (int i = 0; i < kMaxTileLog; ++i) {
        if ((fm_local >> i) & 1u) {
          const T2 nb = tile[l ^ (1u << i)];
          if (CPLX) {
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
      if (a.diag) {
        double e = dbase + s_B[r];
#pragma unroll
        for (int j = 0; j < kMaxTileLog - kThreadBits; ++j)
          if ((r >> j) & 1u) e += Gj[j];
        hr += static_cast<T>(e * static_cast<double>(psi.x));
        hi += static_cast<T>(e * static_cast<double>(psi.y));
      }
      T2 w;
      w.x = sr * hr - si * hi;
      w.y = sr * hi + si * hr;
      const uint64_t x = gidx(l);
      if (a.accumulate) {
        const T2 o = out[x];
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
          T2 p = prevp[x];
          const double pr = static_cast<double>(p.x * ps), pi = static_cast<double>(p.y * ps);
          dots[3] += pr * wr + pi * wi;
          dots[4] += pr * wi - pi * wr;
          dots[5] += vr * pr + vi * pi;
          dots[6] += vr * pi - vi * pr;
        }
      }
    }
  }
  if (DOTS) block_reduce_store<7>(dots, s_red, a.partials + (uint64_t)blockIdx.x * kNumDots);
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

// w <- w - ca*pa - cb*pb ; partial ||w||^2
template <typename T>
__global__ __launch_bounds__(kThreads) void lanczos_update_kernel(
    uint64_t D, typename Vec2<T>::type *w, const typename Vec2<T>::type *pa, double car,
    double cai, const typename Vec2<T>::type *pb, double cbr, double cbi, double *partials) {
  using T2 = typename Vec2<T>::type;
  double nrm[1] = {0.0};
  __shared__ double red[kWaves];
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < D; x += stride) {
    T2 z = w[x];
    double zr = z.x, zi = z.y;
    if (pa) {
      const T2 p = pa[x];
      zr -= car * p.x - cai * p.y;
      zi -= car * p.y + cai * p.x;
    }
    if (pb) {
      const T2 p = pb[x];
      zr -= cbr * p.x - cbi * p.y;
      zi -= cbr * p.y + cbi * p.x;
    }
    z.x = static_cast<T>(zr);
    z.y = static_cast<T>(zi);
    w[x] = z;
    const double qr = z.x, qi = z.y;
    nrm[0] += qr * qr + qi * qi;
  }
  block_reduce_store<1>(nrm, red, partials + (uint64_t)blockIdx.x * kNumDots);
}

// ||a||^2 and <a|b>
template <typename T>
__global__ __launch_bounds__(kThreads) void dot_kernel(uint64_t D, const typename Vec2<T>::type *a,
                                                       const typename Vec2<T>::type *b,
                                                       double *partials) {
  using T2 = typename Vec2<T>::type;
  double acc[3] = {0.0, 0.0, 0.0};
  __shared__ double red[kWaves * 3];
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < D; x += stride) {
    const T2 p = a[x];
    const double ar = p.x, ai = p.y;
    acc[0] += ar * ar + ai * ai;
    if (b) {
      const T2 q = b[x];
      const double br = q.x, bi = q.y;
      acc[1] += ar * br + ai * bi;
      acc[2] += ar * bi - ai * br;
    }
  }
  block_reduce_store<3>(acc, red, partials + (uint64_t)blockIdx.x * kNumDots);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void scale_kernel(uint64_t D, typename Vec2<T>::type *v,
                                                         T s) {
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < D; x += stride) {
    auto z = v[x];
    z.x *= s;
    z.y *= s;
    v[x] = z;
  }
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

constexpr int kLincombBatch = 48;  // vectors per launch (kernel-argument size)
constexpr int kMaxLincomb = 1024;
struct LincombArgs {
  const void *vec[kLincombBatch];
  double cr[kLincombBatch];
  double ci[kLincombBatch];
  int count;
  int accumulate;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void lincomb_kernel(uint64_t D, const LincombArgs args,
                                                           typename Vec2<T>::type *out) {
  using T2 = typename Vec2<T>::type;
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < D; x += stride) {
    double zr = 0.0, zi = 0.0;
    if (args.accumulate) {
      const T2 o = out[x];
      zr = o.x;
      zi = o.y;
    }
    for (int k = 0; k < args.count; ++k) {
      const T2 p = reinterpret_cast<const T2 *>(args.vec[k])[x];
      const double pr = p.x, pi = p.y, cr = args.cr[k], ci = args.ci[k];
      zr += cr * pr - ci * pi;
      zi += cr * pi + ci * pr;
    }
    T2 o;
    o.x = static_cast<T>(zr);
    o.y = static_cast<T>(zi);
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

void default_tiles(int dtype, int &tl, int &cmin) {
  if (tl <= 0) tl = dtype == EMU_SV_C64 ? kDefaultTileLogC64 : kDefaultTileLogC128;
  if (cmin <= 0) cmin = dtype == EMU_SV_C64 ? kDefaultContigC64 : kDefaultContigC128;
}

template <typename T, bool CPLX, bool DOTS>
int set_lds_attr(size_t lds) {
  // Dynamic LDS above 64 KiB must be enabled per kernel.
  static std::mutex mu;
  static size_t enabled = 65536;
  std::lock_guard<std::mutex> lock(mu);
  if (lds <= enabled) return EMU_SV_OK;
  hipError_t err =
      hipFuncSetAttribute(reinterpret_cast<const void *>(&hpsi_pass_kernel<T, CPLX, DOTS>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (err == hipSuccess) enabled = lds;
  return hip_check(err, "hipFuncSetAttribute(MaxDynamicSharedMemorySize)");
}

template <typename T, bool CPLX, bool DOTS>
int launch_pass(const PassArgs &a, hipStream_t st) {
  const int t = a.c + a.m;
  const uint64_t nt = 1ull << (a.n - t);
  const size_t lds = sizeof(typename Vec2<T>::type) << t;
  int rc = set_lds_attr<T, CPLX, DOTS>(lds);
  if (rc) return rc;
  hipLaunchKernelGGL((hpsi_pass_kernel<T, CPLX, DOTS>), dim3((unsigned)nt), dim3(kThreads), lds,
                     st, a);
  return hip_check(hipGetLastError(), "hpsi_pass_kernel launch");
}

template <typename T>
int run_hpsi(int n, const double *ham, double sr, double si, const void *v, double vsc,
             const void *prev, double psc, void *w, double *dots, void *ws, int tl, int cmin,
             bool cplx, hipStream_t st) {
  Pass plan[EMU_SV_MAX_QUBITS];
  const int np = make_plan(n, tl, cmin, plan);
  for (int p = 0; p < np; ++p) {
    PassArgs a;
    a.v = v;
    a.prev = prev;
    a.out = w;
    a.ham = ham;
    a.partials = reinterpret_cast<double *>(ws);
    a.v_scale = vsc;
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
    if (cplx)
      rc = fuse ? launch_pass<T, true, true>(a, st) : launch_pass<T, true, false>(a, st);
    else
      rc = fuse ? launch_pass<T, false, true>(a, st) : launch_pass<T, false, false>(a, st);
    if (rc) return rc;
    if (fuse) {
      const int t = a.c + a.m;
      const int nblk = (int)(1ull << (n - t));
      hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, st,
                         reinterpret_cast<const double *>(ws), nblk, kNumDots, 7, dots);
      rc = hip_check(hipGetLastError(), "finalize launch");
      if (rc) return rc;
    }
  }
  return EMU_SV_OK;
}

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
  return EMU_SV_OK;
}

int emu_sv_plan_passes(int n, int dtype, int tile_log, int min_contig) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  default_tiles(dtype, tile_log, min_contig);
  Pass plan[EMU_SV_MAX_QUBITS];
  return make_plan(n, tile_log, min_contig, plan);
}

size_t emu_sv_workspace_bytes(int n, int dtype) {
  (void)dtype;
  if (n < 1 || n > EMU_SV_MAX_QUBITS) return 0;
  const uint64_t tiles = n > kThreadBits ? (1ull << (n - kThreadBits)) : 1ull;
  uint64_t rows = tiles > (uint64_t)kStreamBlocks ? tiles : (uint64_t)kStreamBlocks;
  uint64_t bytes = rows * kNumDots * sizeof(double);
  const uint64_t corr = (uint64_t)kStreamBlocks * EMU_SV_MAX_QUBITS * EMU_SV_MAX_QUBITS * sizeof(double);
  return (size_t)(bytes > corr ? bytes : corr);
}

int emu_sv_hpsi(int n, int dtype, const double *ham, double sigma_re, double sigma_im,
                const void *v, double v_scale, const void *prev, double prev_scale, void *w,
                double *dots_out, void *workspace, int tile_log, int min_contig, int flags,
                hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!ham || !v || !w) return fail(EMU_SV_EINVAL, "hpsi: null pointer");
  if (v == w) return fail(EMU_SV_EINVAL, "hpsi: input and output must not alias");
  if (dots_out && !workspace) return fail(EMU_SV_EINVAL, "hpsi: reductions need a workspace");
  default_tiles(dtype, tile_log, min_contig);
  if (tile_log < kThreadBits || tile_log > kMaxTileLog)
    return fail(EMU_SV_EINVAL, "hpsi: tile_log must be in [8, 14]");
  if (dtype == EMU_SV_C128 && tile_log > 13)
    return fail(EMU_SV_EINVAL, "hpsi: complex128 tiles are limited to 2^13 elements");
  if (min_contig < 0 || min_contig >= tile_log)
    return fail(EMU_SV_EINVAL, "hpsi: min_contig must be in [0, tile_log)");
  const bool cplx = (flags & EMU_SV_FLAG_COMPLEX) != 0;
  if (dtype == EMU_SV_C128)
    return run_hpsi<double>(n, ham, sigma_re, sigma_im, v, v_scale, prev, prev_scale, w, dots_out,
                            workspace, tile_log, min_contig, cplx, stream);
  return run_hpsi<float>(n, ham, sigma_re, sigma_im, v, v_scale, prev, prev_scale, w, dots_out,
                         workspace, tile_log, min_contig, cplx, stream);
}

int emu_sv_lanczos_update(int n, int dtype, void *w, const void *pa, double pa_scale, double a_re,
                          double a_im, const void *pb, double pb_scale, double b_re, double b_im,
                          double *norm2_out, void *workspace, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!w || !norm2_out || !workspace) return fail(EMU_SV_EINVAL, "lanczos_update: null pointer");
  const uint64_t D = 1ull << n;
  const int nb = stream_blocks(D);
  double *part = reinterpret_cast<double *>(workspace);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(lanczos_update_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, D,
                       (double2 *)w, (const double2 *)pa, a_re * pa_scale, a_im * pa_scale,
                       (const double2 *)pb, b_re * pb_scale, b_im * pb_scale, part);
  else
    hipLaunchKernelGGL(lanczos_update_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, D,
                       (float2 *)w, (const float2 *)pa, a_re * pa_scale, a_im * pa_scale,
                       (const float2 *)pb, b_re * pb_scale, b_im * pb_scale, part);
  if (int rc = hip_check(hipGetLastError(), "lanczos_update launch")) return rc;
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, stream, part, nb, kNumDots, 1,
                     norm2_out);
  return hip_check(hipGetLastError(), "finalize launch");
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
  const uint64_t D = 1ull << n;
  const int nb = stream_blocks(D);
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
      hipLaunchKernelGGL(lincomb_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, D, args,
                         (double2 *)out);
    else
      hipLaunchKernelGGL(lincomb_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, D, args,
                         (float2 *)out);
    if (int rc = hip_check(hipGetLastError(), "lincomb launch")) return rc;
  }
  return EMU_SV_OK;
}

int emu_sv_dot(int n, int dtype, const void *a, const void *b, double *dots_out, void *workspace,
               hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!a || !dots_out || !workspace) return fail(EMU_SV_EINVAL, "dot: null pointer");
  const uint64_t D = 1ull << n;
  const int nb = stream_blocks(D);
  double *part = reinterpret_cast<double *>(workspace);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(dot_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, D,
                       (const double2 *)a, (const double2 *)b, part);
  else
    hipLaunchKernelGGL(dot_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, D,
                       (const float2 *)a, (const float2 *)b, part);
  if (int rc = hip_check(hipGetLastError(), "dot launch")) return rc;
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(kThreads), 0, stream, part, nb, kNumDots, 3,
                     dots_out);
  return hip_check(hipGetLastError(), "finalize launch");
}

int emu_sv_scale(int n, int dtype, void *v, double s, hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!v) return fail(EMU_SV_EINVAL, "scale: null pointer");
  const uint64_t D = 1ull << n;
  const int nb = stream_blocks(D);
  if (dtype == EMU_SV_C128)
    hipLaunchKernelGGL(scale_kernel<double>, dim3(nb), dim3(kThreads), 0, stream, D, (double2 *)v, s);
  else
    hipLaunchKernelGGL(scale_kernel<float>, dim3(nb), dim3(kThreads), 0, stream, D, (float2 *)v,
                       (float)s);
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
  else if (src_dtype == EMU_SV_C128)
    hipLaunchKernelGGL((convert_kernel<double, double>), dim3(nb), dim3(kThreads), 0, stream, D,
                       (const double2 *)src, (double2 *)dst);
  else
    hipLaunchKernelGGL((convert_kernel<float, float>), dim3(nb), dim3(kThreads), 0, stream, D,
                       (const float2 *)src, (float2 *)dst);
  return hip_check(hipGetLastError(), "convert launch");
}

int emu_sv_correlation(int n, int dtype, const void *psi, double *corr_out, void *workspace,
                       hipStream_t stream) {
  if (int rc = check_n(n)) return rc;
  if (int rc = check_dtype(dtype)) return rc;
  if (!psi || !corr_out || !workspace) return fail(EMU_SV_EINVAL, "correlation: null pointer");
  const int K = n < kMaxTileLog ? n : kMaxTileLog;
  const uint64_t nchunks = 1ull << (n - K);
  const int nb = (int)(nchunks < (uint64_t)kStreamBlocks ? nchunks : (uint64_t)kStreamBlocks);
  double *part = reinterpret_cast<double *>(workspace);
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

}  // extern "C"
