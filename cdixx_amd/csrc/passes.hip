// passes.hip -- the fused FFT passes of the CDI phase-retrieval loop.
//
// One iteration of CDI::phaseRetrieve (CDIxx src/post_gpu/cdi.cc:349-354,381)
// is projection -> IFFT -> 1/N -> support update -> FFT.  Here it is two
// kernels that each stream the image once:
//   column pass  (k_col, COL_FWD_PROJ_INV): forward column FFT of T1 = rowFFT(g),
//                 modulus projection (+R-factor sums), inverse column FFT -> T2
//   row pass     (k_row, ROW_INV_UPDATE_FWD): inverse row FFT of T2, 1/N,
//                 HIO/RAAR/ER/... update against g and the support, store g,
//                 forward row FFT -> T1
// so F = colFFT(T1) is never stored between iterations.  The same kernels
// with other OP values serve the standalone 2D FFT and shrink-wrap.
//
// Lines per workgroup (columns for k_col, rows for k_row) are a launch
// parameter (power of two, 64..256 threads), chosen per size by the tuner.
#include "device_ops.hpp"

namespace cdx {

int line_supported(int n) {
  switch (n) {
    case 16: case 32: case 64: case 128: case 256: case 384: case 512: case 768: case 1024: case 2048:
      return 1;
    default:
      return 0;
  }
}
void line_factors(int n, int* P, int* Q) {
  switch (n) {
#define F(n_) case n_: *P = LineCfg<n_>::P; *Q = LineCfg<n_>::Q; return;
    F(16) F(32) F(64) F(128) F(256) F(384) F(512) F(768) F(1024) F(2048)
#undef F
    default: *P = *Q = 0;
  }
}
static int line_lanes(int n) {
  switch (n) {
#define F(n_) case n_: return LineCfg<n_>::TL;
    F(16) F(32) F(64) F(128) F(256) F(384) F(512) F(768) F(1024) F(2048)
#undef F
    default: return 0;
  }
}
static size_t line_pitch_bytes(int n, size_t csz) {
  switch (n) {
#define F(n_) case n_: return (size_t)(csz == 8 ? lds_pitch<float, n_>() : lds_pitch<double, n_>()) * csz;
    F(16) F(32) F(64) F(128) F(256) F(384) F(512) F(768) F(1024) F(2048)
#undef F
    default: return 0;
  }
}

constexpr int kMaxThreads = 256;
constexpr size_t kMaxLds = 160 * 1024;

int tiling_valid(int n, size_t csz, int lines) {
  const int tl = line_lanes(n);
  if (tl == 0 || lines <= 0 || (lines & (lines - 1))) return 0;
  const int threads = lines * tl;
  if (threads < 64 || threads > kMaxThreads) return 0;
  return (size_t)lines * line_pitch_bytes(n, csz) <= kMaxLds;
}
int tiling_default(int n, size_t csz, int column_pass) {
  const int tl = line_lanes(n);
  if (!tl) return 0;
  int lines = column_pass ? (csz == 8 ? 8 : 4) : 128 / tl;
  while (lines * tl < 64) lines *= 2;
  while (!tiling_valid(n, csz, lines) && lines > 1) lines /= 2;
  return tiling_valid(n, csz, lines) ? lines : 0;
}

// ---------------------------------------------------------------------------
// column pass
// ---------------------------------------------------------------------------
template <class T, int N, int OP>
__global__ void __launch_bounds__(kMaxThreads)
k_col(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int cols, const cplx<T>* __restrict__ tw,
      const T* __restrict__ I, ProjArgs<T> pa, T gsig, int lsh) {
  constexpr int P = LineCfg<N>::P, Q = LineCfg<N>::Q, TL = LineCfg<N>::TL;
  constexpr int PITCH = lds_pitch<T, N>();
  using L = Line<T, P, Q, TL>;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
  const int c = threadIdx.x & ((1 << lsh) - 1);
  const int t = threadIdx.x >> lsh;
  const int col = (blockIdx.x << lsh) + c;
  const bool ok = col < cols;
  cplx<T>* my = lds + c * PITCH;
  const cplx<T>* twf = tw;
  const cplx<T>* twi = tw + N;
  cplx<T> v[L::E];

  constexpr bool from_f = (OP == COL_INV || OP == COL_PROJ_INV || OP == COL_GAUSS_INV);
  if constexpr (from_f) {
#pragma unroll
    for (int i = 0; i < L::E; i++) {
      const int r = L::dout_pos(i, t);
      v[i] = ok ? in[(size_t)r * cols + col] : cplx<T>{(T)0, (T)0};
    }
  } else {
#pragma unroll
    for (int i = 0; i < L::E; i++) {
      const int r = L::din_pos(i, t);
      v[i] = ok ? in[(size_t)r * cols + col] : cplx<T>{(T)0, (T)0};
    }
    L::forward(v, my, twf, t);
  }

  if constexpr (OP == COL_FWD) {
    if (ok) {
#pragma unroll
      for (int i = 0; i < L::E; i++) out[(size_t)L::dout_pos(i, t) * cols + col] = v[i];
    }
    return;
  }

  if constexpr (OP == COL_FWD_PROJ_INV || OP == COL_PROJ_INV) {
    double rnum = 0, rden = 0;
#pragma unroll
    for (int i = 0; i < L::E; i++) {
      const int r = L::dout_pos(i, t);
      const T Iv = ok ? I[(size_t)r * cols + col] : (T)NAN;
      if (pa.rf && !isnan(Iv)) {
        const T s = Iv > (T)0 ? dsqrt(Iv) : (T)0;
        rnum += (double)fabs(dhypot(v[i].x, v[i].y) - s);
        rden += (double)s;
      }
      if (ok) v[i] = project(v[i], Iv, pa);
    }
    if (pa.rf) {
      __shared__ double sh[16];
      block_atomic_sum(rnum, pa.rf, sh);
      block_atomic_sum(rden, pa.rf + 1, sh);
    }
  }
  if constexpr (OP == COL_FWD_GAUSS_INV || OP == COL_GAUSS_INV) {
    // applyGaussMult(isFreq=1), cudaConfig.cu:380-395, wrapped coordinates
    const int rows = N;
    const T yrel = (T)(col >= cols / 2 ? col - cols : col);
#pragma unroll
    for (int i = 0; i < L::E; i++) {
      const int r = L::dout_pos(i, t);
      const T xrel = (T)(r >= rows / 2 ? r - rows : r);
      const T f = (T)exp(-(xrel * xrel + yrel * yrel) / ((T)2 * gsig * gsig));
      v[i].x *= f;
      v[i].y *= f;
    }
  }

  __syncthreads();  // LDS region reuse
  L::inverse(v, my, twi, t);
  if (ok) {
#pragma unroll
    for (int i = 0; i < L::E; i++) out[(size_t)L::din_pos(i, t) * cols + col] = v[i];
  }
}

// ---------------------------------------------------------------------------
// row pass
// ---------------------------------------------------------------------------
template <class T, int N, int OP>
__global__ void __launch_bounds__(kMaxThreads)
k_row(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int rows, const cplx<T>* __restrict__ tw,
      cplx<T>* __restrict__ g, const T* __restrict__ spt, UpdArgs<T> ua, T scale, T* __restrict__ sout,
      ReduceSlots rs, int lsh) {
  constexpr int P = LineCfg<N>::P, Q = LineCfg<N>::Q, TL = LineCfg<N>::TL;
  constexpr int PITCH = lds_pitch<T, N>();
  using L = Line<T, P, Q, TL>;
  extern __shared__ __align__(16) unsigned char smem[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem);
  const int t = threadIdx.x % TL;
  const int rr = threadIdx.x / TL;
  const int row = (blockIdx.x << lsh) + rr;
  const bool ok = row < rows;
  cplx<T>* my = lds + rr * PITCH;
  const cplx<T>* twf = tw;
  const cplx<T>* twi = tw + N;
  const size_t base = (size_t)row * N;
  cplx<T> v[L::E];

  if constexpr (OP == ROW_FWD) {
#pragma unroll
    for (int i = 0; i < L::E; i++) v[i] = ok ? in[base + L::din_pos(i, t)] : cplx<T>{(T)0, (T)0};
    L::forward(v, my, twf, t);
    if (ok) {
#pragma unroll
      for (int i = 0; i < L::E; i++) out[base + L::dout_pos(i, t)] = v[i];
    }
    return;
  } else {
#pragma unroll
    for (int i = 0; i < L::E; i++) v[i] = ok ? in[base + L::dout_pos(i, t)] : cplx<T>{(T)0, (T)0};
    L::inverse(v, my, twi, t);
  }

  if constexpr (OP == ROW_INV) {
    if (ok) {
#pragma unroll
      for (int i = 0; i < L::E; i++) out[base + L::din_pos(i, t)] = {v[i].x * scale, v[i].y * scale};
    }
    return;
  }

  if constexpr (OP == ROW_INV_MOD2) {
    T m = 0;
    double s = 0;
#pragma unroll
    for (int i = 0; i < L::E; i++) {
      const T a = v[i].x * v[i].x + v[i].y * v[i].y;
      if (ok) {
        sout[base + L::din_pos(i, t)] = a;
        m = a > m ? a : m;
        s += (double)a;
      }
    }
    __shared__ double sh[16];
    block_atomic_max(m, rs.maxbits, sh);
    block_partial_sum(s, rs, sh);
    return;
  }

  if constexpr (OP == ROW_INV_UPDATE_FWD) {
    const T thr = (T)ua.scal[0];
#pragma unroll
    for (int i = 0; i < L::E; i++) {
      const size_t a = base + L::din_pos(i, t);
      const cplx<T> p = {v[i].x * ua.invN, v[i].y * ua.invN};
      const cplx<T> go = ok ? g[a] : cplx<T>{(T)0, (T)0};
      const T sv = ok ? spt[a] : (T)0;
      v[i] = support_update(go, p, sv > thr, ua.algo, ua.beta, ua.relax);
      if (ok) g[a] = v[i];
    }
    __syncthreads();  // LDS region reuse
    L::forward(v, my, twf, t);
    if (ok) {
#pragma unroll
      for (int i = 0; i < L::E; i++) out[base + L::dout_pos(i, t)] = v[i];
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// Each kernel instantiation raises its own dynamic-LDS limit (above the
// 64 KiB default) once, to the largest tiling it is launched with.
template <class K>
static hipError_t lds_limit(K kern, size_t bytes, size_t& done) {
  if (bytes > 64 * 1024 && bytes > done) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    done = bytes;
  }
  return hipSuccess;
}
static int log2i(int v) {
  int s = 0;
  while ((1 << s) < v) s++;
  return s;
}

template <class T, int N, int OP>
static hipError_t col_launch_n(int cols, const void* in, void* out, const void* tw, const T* I,
                               const ProjArgs<T>& pa, double gs, hipStream_t st, int lines) {
  static size_t lds_done = 0;
  constexpr int TL = LineCfg<N>::TL;
  if (lines == 0) lines = tiling_default(N, sizeof(cplx<T>), 1);
  if (!tiling_valid(N, sizeof(cplx<T>), lines)) return hipErrorInvalidConfiguration;
  const size_t lds = (size_t)lines * lds_pitch<T, N>() * sizeof(cplx<T>);
  auto kern = k_col<T, N, OP>;
  hipError_t e = lds_limit(kern, lds, lds_done);
  if (e != hipSuccess) return e;
  dim3 grid((cols + lines - 1) / lines), block(lines * TL);
  hipLaunchKernelGGL(kern, grid, block, lds, st, (const cplx<T>*)in, (cplx<T>*)out, cols, (const cplx<T>*)tw, I,
                     pa, (T)gs, log2i(lines));
  return hipGetLastError();
}

template <class T, int OP>
static hipError_t col_launch_op(int rows, int cols, const void* in, void* out, const void* tw, const T* I,
                                const ProjArgs<T>& pa, double gs, hipStream_t st, int lines) {
  switch (rows) {
#define CASE(n) case n: return col_launch_n<T, n, OP>(cols, in, out, tw, I, pa, gs, st, lines);
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(384) CASE(512) CASE(768) CASE(1024) CASE(2048)
#undef CASE
    default: return hipErrorInvalidValue;
  }
}

template <class T>
hipError_t launch_col(int op, int rows, int cols, const void* in, void* out, const void* tw, const T* I,
                      const ProjArgs<T>& pa, double gs, hipStream_t st, int lines) {
  switch (op) {
    case COL_FWD: return col_launch_op<T, COL_FWD>(rows, cols, in, out, tw, I, pa, gs, st, lines);
    case COL_INV: return col_launch_op<T, COL_INV>(rows, cols, in, out, tw, I, pa, gs, st, lines);
    case COL_FWD_PROJ_INV: return col_launch_op<T, COL_FWD_PROJ_INV>(rows, cols, in, out, tw, I, pa, gs, st, lines);
    case COL_PROJ_INV: return col_launch_op<T, COL_PROJ_INV>(rows, cols, in, out, tw, I, pa, gs, st, lines);
    case COL_FWD_GAUSS_INV:
      return col_launch_op<T, COL_FWD_GAUSS_INV>(rows, cols, in, out, tw, I, pa, gs, st, lines);
    case COL_GAUSS_INV: return col_launch_op<T, COL_GAUSS_INV>(rows, cols, in, out, tw, I, pa, gs, st, lines);
    default: return hipErrorInvalidValue;
  }
}

template <class T, int N, int OP>
static hipError_t row_launch_n(int rows, const void* in, void* out, const void* tw, void* g, const T* spt,
                               const UpdArgs<T>& ua, T scale, T* sout, const ReduceSlots& rs, hipStream_t st,
                               int lines) {
  static size_t lds_done = 0;
  constexpr int TL = LineCfg<N>::TL;
  if (lines == 0) lines = tiling_default(N, sizeof(cplx<T>), 0);
  if (!tiling_valid(N, sizeof(cplx<T>), lines)) return hipErrorInvalidConfiguration;
  const size_t lds = (size_t)lines * lds_pitch<T, N>() * sizeof(cplx<T>);
  auto kern = k_row<T, N, OP>;
  hipError_t e = lds_limit(kern, lds, lds_done);
  if (e != hipSuccess) return e;
  dim3 grid((rows + lines - 1) / lines), block(lines * TL);
  hipLaunchKernelGGL(kern, grid, block, lds, st, (const cplx<T>*)in, (cplx<T>*)out, rows, (const cplx<T>*)tw,
                     (cplx<T>*)g, spt, ua, scale, sout, rs, log2i(lines));
  return hipGetLastError();
}

template <class T, int OP>
static hipError_t row_launch_op(int rows, int cols, const void* in, void* out, const void* tw, void* g,
                                const T* spt, const UpdArgs<T>& ua, T scale, T* sout, const ReduceSlots& rs,
                                hipStream_t st, int lines) {
  switch (cols) {
#define CASE(n) case n: return row_launch_n<T, n, OP>(rows, in, out, tw, g, spt, ua, scale, sout, rs, st, lines);
    CASE(16) CASE(32) CASE(64) CASE(128) CASE(256) CASE(384) CASE(512) CASE(768) CASE(1024) CASE(2048)
#undef CASE
    default: return hipErrorInvalidValue;
  }
}

template <class T>
hipError_t launch_row(int op, int rows, int cols, const void* in, void* out, const void* tw, void* g,
                      const T* spt, const UpdArgs<T>& ua, T scale, T* sout, const ReduceSlots& rs,
                      hipStream_t st, int lines) {
  switch (op) {
    case ROW_FWD: return row_launch_op<T, ROW_FWD>(rows, cols, in, out, tw, g, spt, ua, scale, sout, rs, st, lines);
    case ROW_INV: return row_launch_op<T, ROW_INV>(rows, cols, in, out, tw, g, spt, ua, scale, sout, rs, st, lines);
    case ROW_INV_UPDATE_FWD:
      return row_launch_op<T, ROW_INV_UPDATE_FWD>(rows, cols, in, out, tw, g, spt, ua, scale, sout, rs, st, lines);
    case ROW_INV_MOD2:
      return row_launch_op<T, ROW_INV_MOD2>(rows, cols, in, out, tw, g, spt, ua, scale, sout, rs, st, lines);
    default: return hipErrorInvalidValue;
  }
}

template hipError_t launch_col<float>(int, int, int, const void*, void*, const void*, const float*,
                                      const ProjArgs<float>&, double, hipStream_t, int);
template hipError_t launch_col<double>(int, int, int, const void*, void*, const void*, const double*,
                                       const ProjArgs<double>&, double, hipStream_t, int);
template hipError_t launch_row<float>(int, int, int, const void*, void*, const void*, void*, const float*,
                                      const UpdArgs<float>&, float, float*, const ReduceSlots&, hipStream_t, int);
template hipError_t launch_row<double>(int, int, int, const void*, void*, const void*, void*, const double*,
                                       const UpdArgs<double>&, double, double*, const ReduceSlots&, hipStream_t,
                                       int);

}  // namespace cdx
