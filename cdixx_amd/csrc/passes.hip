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
// The projection mode, the R-factor switch and the update rule of the two
// hot kernels are template parameters (no per-pixel branching on them), and
// every operand a pass reads besides the line itself (I; g and S) is loaded
// at kernel start so its latency overlaps the first transform.
//
// Lines per workgroup are a launch parameter (power of two, 64..512
// threads), chosen per size by the tuner.
//
// This file is compiled four times (CDX_PASS_T = float/double, CDX_PASS_ROW =
// 0/1) so the instantiations build in parallel; CDX_PASS_HOST marks the
// translation unit that also defines the host-side helpers.
#include <algorithm>

#include "pass_bodies.hpp"

#ifndef CDX_PASS_T
#define CDX_PASS_T float
#endif
#ifndef CDX_PASS_ROW
#define CDX_PASS_ROW 0
#endif

namespace cdx {

constexpr int kMaxThreads = 512;
constexpr size_t kMaxLds = 160 * 1024;

#define CDX_SIZES(F) F(16) F(32) F(64) F(128) F(256) F(384) F(512) F(768) F(1024) F(2048)

static int line_lanes(int n, size_t csz) {
  switch (n) {
#define F(n_) case n_: return csz == 8 ? LineCfgT<float, n_>::TL : LineCfgT<double, n_>::TL;
    CDX_SIZES(F)
#undef F
    default: return 0;
  }
}
static int log2i(int v) {
  int s = 0;
  while ((1 << s) < v) s++;
  return s;
}
// LDS bytes of a workgroup of `lines` lines (both exchange buffers)
static size_t wg_lds_bytes(int n, size_t csz, int lines) {
  const int lsh = log2i(lines);
  switch (n) {
#define F(n_) case n_: return csz == 8 ? lds_bytes<float, n_>(lsh) : lds_bytes<double, n_>(lsh);
    CDX_SIZES(F)
#undef F
    default: return 0;
  }
}

#ifdef CDX_PASS_HOST
int line_supported(int n) { return line_lanes(n, 8) != 0; }

size_t support_mask_bytes(int rows, int cols, size_t csz) {
  return (size_t)rows * (size_t)line_lanes(cols, csz) * sizeof(unsigned);
}

int tiling_valid(int n, size_t csz, int lines) {
  const int tl = line_lanes(n, csz);
  if (tl == 0 || lines <= 0 || (lines & (lines - 1))) return 0;
  const int threads = lines * tl;
  if (threads < 64 || threads > kMaxThreads) return 0;
  return wg_lds_bytes(n, csz, lines) <= kMaxLds;
}
int tiling_default(int n, size_t csz, int column_pass) {
  const int tl = line_lanes(n, csz);
  if (!tl) return 0;
  // columns: 64-byte row segments; rows: 256 threads
  int lines = column_pass ? (int)(64 / csz) : 256 / tl;
  if (lines < 1) lines = 1;
  while (lines * tl < 64) lines *= 2;
  while (!tiling_valid(n, csz, lines) && lines > 1) lines /= 2;
  return tiling_valid(n, csz, lines) ? lines : 0;
}
#endif

// ---------------------------------------------------------------------------
// single-frame kernels
// ---------------------------------------------------------------------------
template <class T, int N, int OP, int MODE, int RF>
__global__ void __launch_bounds__(kMaxThreads)
k_col(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int cols, const cplx<T>* __restrict__ tw,
      const T* __restrict__ I, ProjArgs<T> pa, T gsig, int lsh) {
  extern __shared__ __align__(16) unsigned char smem[];
  // the single-frame loop keeps T1 and the packed measurement in the column-tile layout
  constexpr bool tiled = OP == COL_FWD_PROJ_INV || OP == COL_FWD_GAUSS_INV || OP == COL_FWD_T1;
  // the loop's T2 (the inverse column transforms read by the update and
  // shrink-wrap row passes) is in the pair-of-rows layout
  constexpr bool t2l = OP == COL_FWD_PROJ_INV || OP == COL_PROJ_INV || OP == COL_FWD_GAUSS_INV || OP == COL_GAUSS_INV;
  col_body<T, N, OP, MODE, RF, tiled, true, 0, t2l>(in, out, cols, tw, I, pa, gsig, xcd_tile(blockIdx.x, gridDim.x),
                                                     lsh, smem);
}
template <class T, int N, int OP, int ALGO>
__global__ void __launch_bounds__(kMaxThreads)
k_row(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int rows, const cplx<T>* __restrict__ tw,
      cplx<T>* __restrict__ g, const T* __restrict__ spt, UpdArgs<T> ua, T scale, T* __restrict__ sout,
      ReduceSlots rs, int lsh) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr bool tiled = OP == ROW_INV_UPDATE_FWD || OP == ROW_INV_UPDATE_FWD_PRTF;
  constexpr bool t2l = OP == ROW_INV_UPDATE_FWD || OP == ROW_INV_UPDATE_FWD_PRTF || OP == ROW_INV_MOD2;
  // Update passes of 1024 and 2048 points with one row per workgroup:
  // neighbouring row workgroups run on one XCD, so the two rows of each
  // 128-byte T1 tile line leave from the same L2 (2048^2: 18.4k -> 18.7k
  // iterations/s, the tuner then taking one row per workgroup).  A two-row
  // workgroup fills whole lines itself; remapping those cost 0.3% at 1024^2.
  const bool xcd_rows = tiled && N >= 1024 && lsh == 0;
  // the loop's update pass tests the support from its bit mask
  constexpr bool smask = OP == ROW_INV_UPDATE_FWD;
  row_body<T, N, OP, ALGO, tiled, 0, t2l, tiled && N <= 1024, smask>(
      in, out, rows, tw, g, spt, ua, scale, sout, rs, xcd_rows ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x,
      lsh, smem);
}
// support mask of the update pass: word w = row * TL + t, bit i =
// spt[row][din_pos(i, t)] > (T)scal[0], the comparison row_body makes
template <class T, int N>
__global__ void __launch_bounds__(256) k_support_mask(int rows, const T* __restrict__ spt,
                                                      const double* __restrict__ scal, unsigned* __restrict__ mask) {
  using L = LineOf<T, N, 0>;
  constexpr int TL = LineCfgT<T, N, 0>::TL;
  static_assert(L::E <= 32, "one mask word per lane");
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= rows * TL) return;
  const int row = w / TL, t = w % TL;
  const T thr = (T)scal[0];
  const T* s = spt + (size_t)row * N;
  unsigned m = 0;
#pragma unroll
  for (int i = 0; i < L::E; i++) m |= (s[L::din_pos(i, t)] > thr ? 1u : 0u) << i;
  mask[w] = m;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// A kernel instantiation whose tilings can exceed the default 64 KiB of
// dynamic LDS gets its cap raised once to the largest valid tiling's LDS
// (lds_cap).  prepare_col/prepare_row
// do that for every instantiation of a line length, so a context raises the
// caps at creation and no attribute call happens inside a stream capture.
template <class T, int N>
static size_t max_lds() {
  size_t m = 0;
  for (int lines = 1; lines <= kMaxThreads; lines *= 2)
    if (tiling_valid(N, sizeof(cplx<T>), lines)) m = std::max(m, lds_bytes<T, N>(log2i(lines)));
  return m;
}
// prepare: raise the cap to the largest valid tiling's LDS, best effort (a
// size whose largest tiling the device refuses still launches its smaller
// tilings: the launch raises the cap to what it uses)
template <class K, class T, int N>
static void lds_prepare(K kern, bool& prepared) {
  if (prepared) return;
  (void)lds_cap((const void*)kern, max_lds<T, N>());
  prepared = true;
}
#if CDX_PASS_ROW == 0
template <class T, int N, int OP, int MODE, int RF>
static hipError_t col_launch_n(int cols, const void* in, void* out, const void* tw, const T* I,
                               const ProjArgs<T>& pa, double gs, hipStream_t st, int lines, bool prep) {
  static bool prepared = false;
  constexpr int TL = LineCfgT<T, N>::TL;
  auto kern = k_col<T, N, OP, MODE, RF>;
  if (prep) {
    lds_prepare<decltype(kern), T, N>(kern, prepared);
    return hipSuccess;
  }
  if (!tiling_valid(N, sizeof(cplx<T>), lines)) return hipErrorInvalidConfiguration;
  const size_t lds = lds_bytes<T, N>(log2i(lines));
  hipError_t e = lds_cap((const void*)kern, lds);
  if (e != hipSuccess) return e;
  dim3 grid((cols + lines - 1) / lines), block(lines * TL);
  hipLaunchKernelGGL(kern, grid, block, lds, st, (const cplx<T>*)in, (cplx<T>*)out, cols, (const cplx<T>*)tw, I, pa,
                     (T)gs, log2i(lines));
  return hipGetLastError();
}

template <class T, int N>
static hipError_t col_launch_sz(int op, int cols, const void* in, void* out, const void* tw, const T* I,
                                const ProjArgs<T>& pa, double gs, hipStream_t st, int lines, bool prep) {
  const bool rf = pa.rf != nullptr;
  switch (op) {
    case COL_FWD: return col_launch_n<T, N, COL_FWD, 0, 0>(cols, in, out, tw, I, pa, gs, st, lines, prep);
    case COL_FWD_T1: return col_launch_n<T, N, COL_FWD_T1, 0, 0>(cols, in, out, tw, I, pa, gs, st, lines, prep);
    case COL_INV: return col_launch_n<T, N, COL_INV, 0, 0>(cols, in, out, tw, I, pa, gs, st, lines, prep);
    case COL_FWD_PROJ_INV:
      if (pa.fsave)
        return col_launch_n<T, N, COL_FWD_PROJ_INV, -1, -1>(cols, in, out, tw, I, pa, gs, st, lines, prep);
#define M(mode)                                                                                           \
  return rf ? col_launch_n<T, N, COL_FWD_PROJ_INV, mode, 1>(cols, in, out, tw, I, pa, gs, st, lines, prep) \
            : col_launch_n<T, N, COL_FWD_PROJ_INV, mode, 0>(cols, in, out, tw, I, pa, gs, st, lines, prep);
      switch (pa.mode) {
        case PROJ_MOD_ABS: M(PROJ_MOD_ABS)
        case PROJ_MOD_NOISE: M(PROJ_MOD_NOISE)
        case PROJ_MOD_ABS_BS: M(PROJ_MOD_ABS_BS)
        default: return hipErrorInvalidValue;
      }
#undef M
    case COL_PROJ_INV:
      return col_launch_n<T, N, COL_PROJ_INV, -1, -1>(cols, in, out, tw, I, pa, gs, st, lines, prep);
    case COL_FWD_GAUSS_INV:
      return col_launch_n<T, N, COL_FWD_GAUSS_INV, 0, 0>(cols, in, out, tw, I, pa, gs, st, lines, prep);
    case COL_GAUSS_INV:
      return col_launch_n<T, N, COL_GAUSS_INV, 0, 0>(cols, in, out, tw, I, pa, gs, st, lines, prep);
    default: return hipErrorInvalidValue;
  }
}

template <class T>
static hipError_t col_dispatch(int op, int rows, int cols, const void* in, void* out, const void* tw, const T* I,
                               const ProjArgs<T>& pa, double gs, hipStream_t st, int lines, bool prep) {
  if (lines == 0) lines = tiling_default(rows, sizeof(cplx<T>), 1);
  switch (rows) {
#define F(n) case n: return col_launch_sz<T, n>(op, cols, in, out, tw, I, pa, gs, st, lines, prep);
    CDX_SIZES(F)
#undef F
    default: return hipErrorInvalidValue;
  }
}
template <class T>
hipError_t launch_col(int op, int rows, int cols, const void* in, void* out, const void* tw, const T* I,
                      const ProjArgs<T>& pa, double gs, hipStream_t st, int lines) {
  return col_dispatch<T>(op, rows, cols, in, out, tw, I, pa, gs, st, lines, false);
}
template <class T>
hipError_t prepare_col(int rows) {
  // every instantiation col_launch_sz can pick: each op, and for the
  // projection ops each mode with and without R-factor sums and the
  // run-time-mode (PRTF) variant
  double dummy = 0;
  for (int op : {COL_FWD, COL_INV, COL_FWD_PROJ_INV, COL_PROJ_INV, COL_FWD_GAUSS_INV, COL_GAUSS_INV, COL_FWD_T1})
    for (int mode : {PROJ_MOD_ABS, PROJ_MOD_NOISE, PROJ_MOD_ABS_BS})
      for (int v = 0; v < 3; v++) {
        ProjArgs<T> pa{};
        pa.mode = mode;
        pa.rf = v == 1 ? &dummy : nullptr;
        pa.fsave = v == 2 ? (void*)&dummy : nullptr;
        hipError_t e = col_dispatch<T>(op, rows, 0, nullptr, nullptr, nullptr, nullptr, pa, 0, nullptr, 0, true);
        if (e != hipSuccess) return e;
      }
  return hipSuccess;
}
template hipError_t launch_col<CDX_PASS_T>(int, int, int, const void*, void*, const void*, const CDX_PASS_T*,
                                           const ProjArgs<CDX_PASS_T>&, double, hipStream_t, int);
template hipError_t prepare_col<CDX_PASS_T>(int);
#else
template <class T, int N, int OP, int ALGO>
static hipError_t row_launch_n(int rows, const void* in, void* out, const void* tw, void* g, const T* spt,
                               const UpdArgs<T>& ua, T scale, T* sout, const ReduceSlots& rs, hipStream_t st,
                               int lines, bool prep) {
  static bool prepared = false;
  constexpr int TL = LineCfgT<T, N>::TL;
  auto kern = k_row<T, N, OP, ALGO>;
  if (prep) {
    lds_prepare<decltype(kern), T, N>(kern, prepared);
    return hipSuccess;
  }
  if (!tiling_valid(N, sizeof(cplx<T>), lines)) return hipErrorInvalidConfiguration;
  const size_t lds = lds_bytes<T, N>(log2i(lines));
  hipError_t e = lds_cap((const void*)kern, lds);
  if (e != hipSuccess) return e;
  dim3 grid((rows + lines - 1) / lines), block(lines * TL);
  hipLaunchKernelGGL(kern, grid, block, lds, st, (const cplx<T>*)in, (cplx<T>*)out, rows, (const cplx<T>*)tw,
                     (cplx<T>*)g, spt, ua, scale, sout, rs, log2i(lines));
  return hipGetLastError();
}

template <class T, int N>
static hipError_t row_launch_sz(int op, int rows, const void* in, void* out, const void* tw, void* g,
                                const T* spt, const UpdArgs<T>& ua, T scale, T* sout, const ReduceSlots& rs,
                                hipStream_t st, int lines, bool prep) {
#define R(op_, a_) \
  return row_launch_n<T, N, op_, a_>(rows, in, out, tw, g, spt, ua, scale, sout, rs, st, lines, prep);
  switch (op) {
    case ROW_FWD: R(ROW_FWD, 0)
    case ROW_INV: R(ROW_INV, 0)
    case ROW_INV_MOD2: R(ROW_INV_MOD2, 0)
    case ROW_INV_UPDATE_FWD:
      switch (ua.algo) {
        case 0: R(ROW_INV_UPDATE_FWD, 0)
        case 1: R(ROW_INV_UPDATE_FWD, 1)
        case 2: R(ROW_INV_UPDATE_FWD, 2)
        case 3: R(ROW_INV_UPDATE_FWD, 3)
        case 4: R(ROW_INV_UPDATE_FWD, 4)
        default: R(ROW_INV_UPDATE_FWD, 5)  // FHIO and every other token (the dispatcher's default)
      }
    case ROW_INV_UPDATE_FWD_PRTF:
      if (!prep && 2 * (size_t)((rows + lines - 1) / lines) > (size_t)kMaxPartials)
        return hipErrorInvalidConfiguration;
      R(ROW_INV_UPDATE_FWD_PRTF, -1)
    default: return hipErrorInvalidValue;
  }
#undef R
}

template <class T>
static hipError_t row_dispatch(int op, int rows, int cols, const void* in, void* out, const void* tw, void* g,
                               const T* spt, const UpdArgs<T>& ua, T scale, T* sout, const ReduceSlots& rs,
                               hipStream_t st, int lines, bool prep) {
  if (lines == 0) lines = tiling_default(cols, sizeof(cplx<T>), 0);
  switch (cols) {
#define F(n) \
  case n: return row_launch_sz<T, n>(op, rows, in, out, tw, g, spt, ua, scale, sout, rs, st, lines, prep);
    CDX_SIZES(F)
#undef F
    default: return hipErrorInvalidValue;
  }
}
template <class T>
hipError_t launch_row(int op, int rows, int cols, const void* in, void* out, const void* tw, void* g,
                      const T* spt, const UpdArgs<T>& ua, T scale, T* sout, const ReduceSlots& rs,
                      hipStream_t st, int lines) {
  return row_dispatch<T>(op, rows, cols, in, out, tw, g, spt, ua, scale, sout, rs, st, lines, false);
}
template <class T>
hipError_t prepare_row(int cols) {
  for (int op : {ROW_FWD, ROW_INV, ROW_INV_MOD2, ROW_INV_UPDATE_FWD, ROW_INV_UPDATE_FWD_PRTF})
    for (int algo = 0; algo <= 5; algo++) {
      UpdArgs<T> ua{};
      ua.algo = algo;
      hipError_t e = row_dispatch<T>(op, 0, cols, nullptr, nullptr, nullptr, nullptr, nullptr, ua, (T)1, nullptr,
                                     ReduceSlots{}, nullptr, 0, true);
      if (e != hipSuccess) return e;
    }
  return hipSuccess;
}
template <class T>
hipError_t launch_support_mask(int rows, int cols, const T* spt, const double* scal, unsigned* mask, hipStream_t st) {
  switch (cols) {
#define F(n_)                                                                                                    \
  case n_: {                                                                                                     \
    const int words = rows * LineCfgT<T, n_>::TL;                                                               \
    hipLaunchKernelGGL((k_support_mask<T, n_>), dim3((words + 255) / 256), dim3(256), 0, st, rows, spt, scal, mask); \
    return hipGetLastError();                                                                                    \
  }
    CDX_SIZES(F)
#undef F
    default: return hipErrorInvalidValue;
  }
}
template hipError_t launch_support_mask<CDX_PASS_T>(int, int, const CDX_PASS_T*, const double*, unsigned*,
                                                    hipStream_t);
template hipError_t launch_row<CDX_PASS_T>(int, int, int, const void*, void*, const void*, void*,
                                           const CDX_PASS_T*, const UpdArgs<CDX_PASS_T>&, CDX_PASS_T, CDX_PASS_T*,
                                           const ReduceSlots&, hipStream_t, int);
template hipError_t prepare_row<CDX_PASS_T>(int);
#endif

}  // namespace cdx
