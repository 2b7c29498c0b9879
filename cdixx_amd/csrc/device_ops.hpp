// device_ops.hpp -- device helpers shared by the CDI kernels: line
// configurations of the FFT, the reference's pointwise operations, workgroup
// reductions, grid-stride helpers and the Philox generator.
#pragma once
#include "fft_line.hpp"
#include "kernels.hpp"
#include <math.h>

namespace cdx {

// ---------------------------------------------------------------------------
// line configurations
// ---------------------------------------------------------------------------
// lanes per line (TL), exchange-buffer padding (PAD, one element per 2^PAD;
// 0: none) and Stockham radices (forward order) per line length; every radix
// divides the E = N/TL points a lane holds.  PAD is measured per length: the
// 384- and 768-point lines (last radix 6 / 12) run faster unpadded
// (per iteration: 384^2 17.06 -> 15.98 us, 768^2 20.59 -> 19.89 us), 2048
// needs the padding (57.3 us padded, 67.1 unpadded), 1024 is indifferent
// (padding every 8, 16 or 32 elements or an XOR swizzle: 17.8-17.9 us;
// unpadded 19.6).
template <int N>
struct LineCfg;
#define CDX_LINE(n, tl, pad, ...)            \
  template <>                                \
  struct LineCfg<n> {                        \
    static constexpr int TL = tl;            \
    static constexpr int PAD = pad;          \
    using RL = Radices<__VA_ARGS__>;         \
  };
CDX_LINE(16, 4, 5, 4, 4)
CDX_LINE(32, 4, 5, 8, 4)
CDX_LINE(64, 8, 5, 8, 8)
CDX_LINE(128, 16, 5, 8, 8, 2)
CDX_LINE(256, 16, 5, 16, 16)
CDX_LINE(384, 16, 0, 8, 8, 6)
CDX_LINE(512, 32, 5, 16, 16, 2)
CDX_LINE(768, 32, 0, 8, 8, 12)
CDX_LINE(1024, 128, 5, 8, 8, 8, 2)
CDX_LINE(2048, 128, 5, 16, 16, 8)
#undef CDX_LINE

// Per precision: FP64 lines holding 16 or 24 complex doubles per lane (512
// and 2048 on 32 and 128 lanes; 384 and 768 on 16 and 32) exceed, with the
// row pass's g, S and twiddles, the 256 VGPRs of a wave (up to 462 spilled);
// they run on more lanes with 8 or 12 points each.  The 256-point FP64 line
// keeps 16 points per lane (66 spilled VGPRs in the update pass).
template <class T, int N, int V = 0>
struct LineCfgT : LineCfg<N> {};
template <>
struct LineCfgT<double, 512> {
  static constexpr int TL = 64;
  static constexpr int PAD = 5;
  using RL = Radices<8, 8, 8>;
};
template <>
struct LineCfgT<double, 384> {
  static constexpr int TL = 32;
  static constexpr int PAD = 0;
  using RL = Radices<4, 4, 4, 6>;
};
template <>
struct LineCfgT<double, 768> {
  static constexpr int TL = 64;
  static constexpr int PAD = 0;
  using RL = Radices<4, 4, 4, 12>;
};
template <>
struct LineCfgT<double, 2048> {
  static constexpr int TL = 256;
  static constexpr int PAD = 5;
  using RL = Radices<8, 8, 8, 4>;
};

// Variant 1 of the FP32 lines of the batched passes (128-512 points): twice
// the lanes with half the points each, the schedules of the FP64 lines (for
// 128 points, 4 points per lane); the batch autotuner picks a variant per
// frame size.  Other lengths and FP64 have variant 0 only.
template <>
struct LineCfgT<float, 128, 1> {
  static constexpr int TL = 32;
  static constexpr int PAD = 5;
  using RL = Radices<4, 4, 4, 2>;
};
template <>
struct LineCfgT<float, 256, 1> {
  static constexpr int TL = 32;
  static constexpr int PAD = 5;
  using RL = Radices<8, 8, 4>;
};
template <>
struct LineCfgT<float, 384, 1> {
  static constexpr int TL = 32;
  static constexpr int PAD = 0;
  using RL = Radices<4, 4, 4, 6>;
};
template <>
struct LineCfgT<float, 512, 1> {
  static constexpr int TL = 64;
  static constexpr int PAD = 5;
  using RL = Radices<8, 8, 8>;
};

template <class T, int N, int V = 0>
using LineOf = Line<T, N, LineCfgT<T, N, V>::TL, typename LineCfgT<T, N, V>::RL, LineCfgT<T, N, V>::PAD>;

// LDS layout of a workgroup of 2^lsh lines: two exchange buffers (the
// transforms alternate between them, one barrier per exchange); buffer b of
// line c starts at (b * 2^lsh + c) * pitch.  The pitch is the exchange
// footprint rounded to a 128-byte bank row (EPB complex elements) plus
// EPB / 2^lsh, so the lines of a column tile (lanes (t, c) with c fastest)
// start on different banks.
template <class T, int N, int V = 0>
__host__ __device__ constexpr int lds_pitch(int lsh) {
  constexpr int EPB = 128 / (int)sizeof(cplx<T>);
  constexpr int e = LineOf<T, N, V>::LDS_ELEMS;
  constexpr int base = (e + EPB - 1) / EPB * EPB;
  return base + (lsh < 31 ? (EPB >> lsh) & (EPB - 1) : 0);
}
template <class T, int N, int V = 0>
__host__ __device__ constexpr size_t lds_bytes(int lsh) {
  return (size_t)2 * ((size_t)1 << lsh) * lds_pitch<T, N, V>(lsh) * sizeof(cplx<T>);
}

// index of element (r, c) of a rows-row image in a layout of column tiles W
// wide: the loop state T1 (t1_index) and the packed measurement (i_index)
template <int W>
__device__ __forceinline__ size_t tile_index(int r, int c, int rows) {
  return (size_t)(c / W) * ((size_t)rows * W) + (size_t)r * W + (c % W);
}
__device__ __forceinline__ size_t t1_index(int r, int c, int rows) { return tile_index<kT1Tile>(r, c, rows); }
__device__ __forceinline__ size_t i_index(int r, int c, int rows) { return tile_index<kITile>(r, c, rows); }

// index of element (r, c) of a cols-column image in the pair-of-rows layout
// of the single-frame loop's T2: rows 2k and 2k+1 interleaved in groups of
// four columns, so the column pass writes 64-byte pieces (two rows of its
// four-column tile) and a row-pass workgroup of two rows reads one
// contiguous block (rows even, cols a multiple of 4)
__device__ __forceinline__ size_t t2_index(int r, int c, int cols) {
  return (size_t)(r >> 1) * (2 * (size_t)cols) + (size_t)((c >> 2) << 3) + (size_t)((r & 1) << 2) + (c & 3);
}

// ---------------------------------------------------------------------------
// pointwise device helpers (reference semantics)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dhypot(float a, float b) { return hypotf(a, b); }
__device__ __forceinline__ double dhypot(double a, double b) { return hypot(a, b); }

// applyModAbs (cudaConfig.cu:827-841) with state == NULL
template <class T>
__device__ __forceinline__ cplx<T> mod_abs(cplx<T> f, T I) {
  T mod = dhypot(f.x, f.y);
  T rat = I > (T)0 ? dsqrt(I) : (T)0;
  if (mod == (T)0) return {rat, (T)0};
  rat /= mod;
  return {f.x * rat, f.y * rat};
}

// applyMod, 5-argument form (cdi.cu:438-485) with bsnorm = 1; beamstop pixels
// arrive as NaN intensities and are left unchanged.
template <class T>
__device__ __forceinline__ cplx<T> mod_noise(cplx<T> f, T I, int noise, T scale, T rcolor) {
  if (isnan(I)) return f;
  T mod2 = I > (T)0 ? I : (T)0;
  T src2 = f.x * f.x + f.y * f.y;
  T maximum = scale * (T)0.95;
  if (mod2 >= maximum) mod2 = mod2 > src2 ? mod2 : src2;
  T tol = dsqrt((T)noise + mod2 * rcolor / scale) * scale / rcolor;
  T diff = mod2 - src2;
  T val = mod2;
  if (diff > tol) val -= tol;
  else if (diff < -tol) val += tol;
  if (src2 == (T)0) return {val, (T)0};
  T s = dsqrt((val > (T)0 ? val : (T)0) / src2);
  return {s * f.x, s * f.y};
}

// applyModAccurate (cdi.cu:351-388) on r = bsnorm * F.  I: the intensity,
// NaN on beamstop pixels (where the reference writes r back).  Where
// |r|^2 == 0 the reference leaves F unwritten; r is zero there as well.
template <class T>
__device__ __forceinline__ cplx<T> mod_accurate(cplx<T> r, T I, T scale) {
  if (isnan(I)) return r;
  const T mod2 = I > (T)0 ? I : (T)0;
  const T src2 = r.x * r.x + r.y * r.y;
  if (src2 == (T)0) return r;
  if (mod2 >= scale * (T)0.95 && src2 >= mod2) return r;
  const T s = dsqrt(mod2 / src2);
  return {s * r.x, s * r.y};
}

template <class T>
__device__ __forceinline__ cplx<T> project(cplx<T> f, T I, const ProjArgs<T>& pa) {
  if (pa.mode == PROJ_MOD_NOISE) return mod_noise(f, I, pa.noiseLevel, pa.scale, pa.rcolor);
  if (pa.mode == PROJ_MOD_ABS_BS && isnan(I)) return f;
  return mod_abs(f, I);
}

// Apply*Support (cdi.cu:518-616) behind the applySupport dispatcher
// (cdi.cc:25-36).  g = current object estimate, p = projected estimate g'.
template <class T>
__device__ __forceinline__ cplx<T> support_update(cplx<T> g, cplx<T> p, bool in, int algo, T beta, T relax) {
  switch (algo) {
    case 0:  // RAAR
      if (in) return p;
      return {beta * g.x + ((T)1 - (T)2 * beta) * p.x, beta * g.y + ((T)1 - (T)2 * beta) * p.y};
    case 1:  // ER
      if (in) return {g.x + relax * (p.x - g.x), g.y + relax * (p.y - g.y)};
      return {(T)0, (T)0};
    case 2:  // POSER
      if (in && p.x > (T)0) return p;
      return {(T)0, (T)0};
    case 3:  // HIO
      if (in) return p;
      return {g.x - beta * p.x, g.y - beta * p.y};
    case 4: {  // POSHIO
      cplx<T> r = g;
      if (p.x > (T)0 && in) r.x = g.x + relax * (p.x - g.x);
      else r.x = g.x - beta * p.x;
      r.y = g.y - beta * p.y;
      return r;
    }
    default:  // FHIO, KKT, XCORRELATION, ...
      if (in) return {g.x + relax * (p.x - g.x), g.y + relax * (p.y - g.y)};
      return {g.x - (T)1.2 * p.x, g.y - (T)1.2 * p.y};
  }
}

// ---------------------------------------------------------------------------
// workgroup reductions
// ---------------------------------------------------------------------------
template <class V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <class V>
__device__ __forceinline__ V wave_max(V v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    V w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
// sum of a per-thread double over the workgroup, added atomically to *dst
__device__ __forceinline__ void block_atomic_sum(double v, double* dst, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < nw; i++) s += sh[i];
    atomicAdd(dst, s);
  }
}
// Workgroup barrier ordering LDS only: unlike __syncthreads it does not wait
// for the wave's outstanding global stores, so a reduction after a pass's
// stores does not stall on their write latency.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// workgroup sums of two per-thread values (reduced together in the working
// precision, one shuffle chain), added atomically in double to dst[0], dst[1]
template <class V>
__device__ __forceinline__ void block_atomic_sum2(V a, V b, double* dst, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
  lds_barrier();
  if (l == 0) {
    sh[2 * w] = (double)a;
    sh[2 * w + 1] = (double)b;
  }
  lds_barrier();
  if (threadIdx.x == 0) {
    double sa = 0, sb = 0;
    for (int i = 0; i < nw; i++) {
      sa += sh[2 * i];
      sb += sh[2 * i + 1];
    }
    atomicAdd(dst, sa);
    atomicAdd(dst + 1, sb);
  }
}
// workgroup sum of a per-thread double stored at *dst by thread 0
__device__ __forceinline__ void block_store_sum(double v, double* dst, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int i = 0; i < nw; i++) s += sh[i];
    *dst = s;
  }
}
// workgroup sum stored at rs.partials[blockIdx.x] (fixed-order reduction)
__device__ __forceinline__ void block_partial_sum(double v, const ReduceSlots& rs, double* sh) {
  block_store_sum(v, rs.partials + blockIdx.x, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) *rs.nparts = (int)gridDim.x;
}
__device__ __forceinline__ void atomic_max_pos(void* bits, float v) {
  atomicMax((unsigned int*)bits, __float_as_uint(v));
}
__device__ __forceinline__ void atomic_max_pos(void* bits, double v) {
  atomicMax((unsigned long long*)bits, (unsigned long long)__double_as_longlong(v));
}
template <class T>
__device__ __forceinline__ void block_atomic_max(T v, void* dst, double* sh) {
  v = wave_max(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (l == 0) sh[w] = (double)v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0;
    for (int i = 0; i < nw; i++) m = sh[i] > m ? sh[i] : m;
    atomic_max_pos(dst, (T)m);
  }
}

// ---------------------------------------------------------------------------
// elementwise kernels (grid-stride, 256 threads)
// ---------------------------------------------------------------------------
inline dim3 ew_grid(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}
// grids of kernels that end in one atomic per workgroup: one workgroup per CU
inline dim3 red_grid(int64_t n) {
  dim3 g = ew_grid(n);
  if (g.x > 256) g.x = 256;
  return g;
}
#define EW_LOOP(n) for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// ---------------------------------------------------------------------------
// Philox4x64-10 (Salmon et al., SC'11) for random starts and the forward model
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint64_t lo0 = 0xD2E7470EE14C6C93ULL * c[0], hi0 = __umul64hi(0xD2E7470EE14C6C93ULL, c[0]);
    const uint64_t lo1 = 0xCA5A826395121157ULL * c[2], hi1 = __umul64hi(0xCA5A826395121157ULL, c[2]);
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B97F4A7C15ULL;
    k1 += 0xBB67AE8584CAA73BULL;
  }
}
__device__ __forceinline__ double philox_uniform(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t c[4] = {idx, 0, 0, 0};
  philox4x64(c, seed, stream);
  return (double)(c[0] >> 11) * 0x1.0p-53;
}

}  // namespace cdx
