// kernels.hip -- setup, shrink-wrap, diagnostic and forward-model kernels of
// the CDI path (the fused FFT passes are in passes.hip).
#include "device_ops.hpp"

#include <map>
#include <mutex>
#include <utility>

namespace cdx {

hipError_t lds_cap(const void* kern, size_t bytes) {
  if (bytes <= 64 * 1024) return hipSuccess;
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, size_t> raised;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(mu);
  size_t& have = raised[{dev, kern}];
  if (bytes <= have) return hipSuccess;
  e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // leave no sticky error for the next launch check
    return e;
  }
  have = bytes;
  return hipSuccess;
}

// beamstop folded into the intensities as a NaN sentinel (bs > 0.5)
template <class T>
__global__ void k_pack(int64_t n, const T* __restrict__ I, const T* __restrict__ bs, T* __restrict__ out, int amp) {
  EW_LOOP(n) {
    const T v = I[i];
    out[i] = (bs && bs[i] > (T)0.5) ? (T)NAN : amp ? (v > (T)0 ? dsqrt(v) : (T)0) : v;
  }
}
template <class T>
hipError_t launch_pack_intensity(int64_t n, const T* I, const T* bs, T* out, int amp, hipStream_t st) {
  hipLaunchKernelGGL(k_pack<T>, ew_grid(n), dim3(256), 0, st, n, I, bs, out, amp);
  return hipGetLastError();
}
// the same, written in the column-tile layout of the measurement (i_index)
template <class T>
__global__ void k_pack_tiled(int rows, int cols, const T* __restrict__ I, const T* __restrict__ bs,
                             T* __restrict__ out, int amp) {
  const int64_t n = (int64_t)rows * cols;
  EW_LOOP(n) {
    const T v = I[i];
    const int r = (int)(i / cols), c = (int)(i % cols);
    out[i_index(r, c, rows)] = (bs && bs[i] > (T)0.5) ? (T)NAN : amp ? (v > (T)0 ? dsqrt(v) : (T)0) : v;
  }
}
template <class T>
hipError_t launch_pack_intensity_tiled(int rows, int cols, const T* I, const T* bs, T* out, int amp,
                                       hipStream_t st) {
  hipLaunchKernelGGL(k_pack_tiled<T>, ew_grid((int64_t)rows * cols), dim3(256), 0, st, rows, cols, I, bs, out, amp);
  return hipGetLastError();
}

template <class T>
__global__ void k_mod_abs(int64_t n, cplx<T>* F, const T* I) {
  EW_LOOP(n) F[i] = mod_abs(F[i], I[i]);
}
template <class T>
hipError_t launch_mod_abs(int64_t n, void* F, const T* I, hipStream_t st) {
  hipLaunchKernelGGL(k_mod_abs<T>, ew_grid(n), dim3(256), 0, st, n, (cplx<T>*)F, I);
  return hipGetLastError();
}

// applyMod with the explicit beamstop array and bsnorm (cdi.cu:438-485)
template <class T>
__global__ void k_mod_noise(int64_t n, cplx<T>* F, const T* I, const T* bs, int noise, T scale, T rcolor, T bsnorm) {
  EW_LOOP(n) {
    cplx<T> f = F[i];
    if (bs && bs[i] > (T)0.5) {
      if (bsnorm != (T)1) F[i] = {f.x * bsnorm, f.y * bsnorm};
      continue;
    }
    if (bsnorm != (T)1) f = {f.x * bsnorm, f.y * bsnorm};
    const T Iv = I[i];
    F[i] = mod_noise(f, isnan(Iv) ? (T)0 : Iv, noise, scale, rcolor);
  }
}
template <class T>
hipError_t launch_mod_noise(int64_t n, void* F, const T* I, const T* bs, int noise, T scale, T rcolor, T bsnorm,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_mod_noise<T>, ew_grid(n), dim3(256), 0, st, n, (cplx<T>*)F, I, bs, noise, scale, rcolor,
                     bsnorm);
  return hipGetLastError();
}

template <class T>
__global__ void k_support(int64_t n, cplx<T>* g, const cplx<T>* gp, const T* spt, UpdArgs<T> ua) {
  const T thr = (T)ua.scal[0];
  EW_LOOP(n) g[i] = support_update(g[i], gp[i], spt[i] > thr, ua.algo, ua.beta, ua.relax);
}
template <class T>
hipError_t launch_support(int64_t n, void* g, const void* gp, const T* spt, const UpdArgs<T>& ua, hipStream_t st) {
  hipLaunchKernelGGL(k_support<T>, ew_grid(n), dim3(256), 0, st, n, (cplx<T>*)g, (const cplx<T>*)gp, spt, ua);
  return hipGetLastError();
}

template <class T>
__global__ void k_norm_c(int64_t n, cplx<T>* d, T f) {
  EW_LOOP(n) d[i] = {d[i].x * f, d[i].y * f};
}
template <class T>
hipError_t launch_norm_c(int64_t n, void* d, T f, hipStream_t st) {
  hipLaunchKernelGGL(k_norm_c<T>, ew_grid(n), dim3(256), 0, st, n, (cplx<T>*)d, f);
  return hipGetLastError();
}
template <class T>
__global__ void k_norm_r(int64_t n, T* d, T f) {
  EW_LOOP(n) d[i] *= f;
}
template <class T>
hipError_t launch_norm_r(int64_t n, T* d, T f, hipStream_t st) {
  hipLaunchKernelGGL(k_norm_r<T>, ew_grid(n), dim3(256), 0, st, n, d, f);
  return hipGetLastError();
}

template <class T>
__global__ void k_mod2(int64_t n, T* out, const cplx<T>* in, T norm) {
  EW_LOOP(n) {
    cplx<T> a = in[i];
    out[i] = (a.x * a.x + a.y * a.y) * norm;
  }
}
template <class T>
hipError_t launch_mod2(int64_t n, T* out, const void* in, T norm, hipStream_t st) {
  hipLaunchKernelGGL(k_mod2<T>, ew_grid(n), dim3(256), 0, st, n, out, (const cplx<T>*)in, norm);
  return hipGetLastError();
}

template <class T>
__global__ void k_gauss_freq(int rows, int cols, const cplx<T>* in, cplx<T>* out, T sigma) {
  const int64_t n = (int64_t)rows * cols;
  EW_LOOP(n) {
    const int x = (int)(i / cols), y = (int)(i % cols);
    const T xr = (T)(x >= rows / 2 ? x - rows : x);
    const T yr = (T)(y >= cols / 2 ? y - cols : y);
    const T f = (T)exp(-(xr * xr + yr * yr) / ((T)2 * sigma * sigma));
    out[i] = {f * in[i].x, f * in[i].y};
  }
}
template <class T>
hipError_t launch_gauss_mult_freq(int rows, int cols, const void* in, void* out, T sigma, hipStream_t st) {
  hipLaunchKernelGGL(k_gauss_freq<T>, ew_grid((int64_t)rows * cols), dim3(256), 0, st, rows, cols,
                     (const cplx<T>*)in, (cplx<T>*)out, sigma);
  return hipGetLastError();
}

// cudaConvertFO (cudaConfig.cu:202-216): quadrant swap; out may equal in
template <class E>
__global__ void k_convert_fo(int rows, int cols, const E* in, E* out) {
  const int64_t half = (int64_t)rows * cols / 2;
  EW_LOOP(half) {
    const int x = (int)(i / cols), y = (int)(i % cols);
    const int64_t ip = (int64_t)(x + (rows >> 1)) * cols + (y >= (cols >> 1) ? y - (cols >> 1) : y + (cols >> 1));
    E a = in[i], b = in[ip];
    out[i] = b;
    out[ip] = a;
  }
}
template <class T>
hipError_t launch_convert_fo(int rows, int cols, const void* in, void* out, int cpx, hipStream_t st) {
  const int64_t half = (int64_t)rows * cols / 2;
  if (cpx)
    hipLaunchKernelGGL(k_convert_fo<cplx<T>>, ew_grid(half), dim3(256), 0, st, rows, cols, (const cplx<T>*)in,
                       (cplx<T>*)out);
  else
    hipLaunchKernelGGL(k_convert_fo<T>, ew_grid(half), dim3(256), 0, st, rows, cols, (const T*)in, (T*)out);
  return hipGetLastError();
}

// createMask<rect> (cudaConfig.cu:1291-1303, rect::isInside :89-92)
template <class T>
__global__ void k_rect(int rows, int cols, int sx, int sy, int ex, int ey, T* out) {
  const int64_t n = (int64_t)rows * cols;
  EW_LOOP(n) {
    const int x = (int)(i / cols), y = (int)(i % cols);
    out[i] = (x >= sx && x < ex && y >= sy && y < ey) ? (T)1 : (T)0;
  }
}
template <class T>
hipError_t launch_rect_mask(int rows, int cols, int sx, int sy, int ex, int ey, T* out, hipStream_t st) {
  hipLaunchKernelGGL(k_rect<T>, ew_grid((int64_t)rows * cols), dim3(256), 0, st, rows, cols, sx, sy, ex, ey, out);
  return hipGetLastError();
}
// createMask<C_circle> centred at (rows>>1, cols>>1) (cudaConfig.cu:93-98),
// freq=1 applies the quadrant shift first (createBeamStop, misc.cc:191-195)
template <class T>
__global__ void k_circle(int rows, int cols, float radius, int freq, T* out) {
  const int64_t n = (int64_t)rows * cols;
  EW_LOOP(n) {
    int x = (int)(i / cols), y = (int)(i % cols);
    if (freq) {
      x = x >= rows / 2 ? x - rows / 2 : x + rows / 2;
      y = y >= cols / 2 ? y - cols / 2 : y + cols / 2;
    }
    const float dx = (float)(x - (rows >> 1)), dy = (float)(y - (cols >> 1));
    out[i] = sqrtf(dx * dx + dy * dy) < radius ? (T)1 : (T)0;
  }
}
template <class T>
hipError_t launch_circle_mask(int rows, int cols, float radius, int freq, T* out, hipStream_t st) {
  hipLaunchKernelGGL(k_circle<T>, ew_grid((int64_t)rows * cols), dim3(256), 0, st, rows, cols, radius, freq, out);
  return hipGetLastError();
}

// max (init 0, as max_op with Real(0) in cub_wrap.cu:88-92) and sum
template <class T>
__global__ void k_reduce(int64_t n, const T* x, ReduceSlots rs) {
  T m = 0;
  double s = 0;
  EW_LOOP(n) {
    const T v = x[i];
    m = v > m ? v : m;
    s += (double)v;
  }
  __shared__ double sh[16];
  block_atomic_max(m, rs.maxbits, sh);
  block_partial_sum(s, rs, sh);
}
template <class T>
hipError_t launch_reduce(int64_t n, const T* x, const ReduceSlots& rs, hipStream_t st) {
  dim3 grid = ew_grid(n);
  if (grid.x > 1024) grid.x = 1024;  // partial-sum capacity
  hipLaunchKernelGGL(k_reduce<T>, grid, dim3(256), 0, st, n, x, rs);
  return hipGetLastError();
}

// Shrink-wrap threshold from the reduced max/sum, then clear the slots.
// mode 0: cdi.cc:330-333   threshold = min(max, 5*avg)*shrinkThreshold,
//         avg = 4*sum/N taken on the first call only (scal[1] holds it)
// mode 1: holo.cc:141      threshold = max*shrinkThreshold
// one wavefront; the partial sums are added in a fixed order
template <class T>
__global__ void __launch_bounds__(64) k_threshold(double* scal, ReduceSlots rs, double shrinkThreshold, int rows,
                                                  int cols, int mode) {
  const int np = *rs.nparts;
  double s = 0;
  for (int i = threadIdx.x; i < np; i += 64) s += rs.partials[i];
  s = wave_sum(s);
  if (threadIdx.x != 0) return;
  T maxs;
  if (sizeof(T) == 4) maxs = (T)__uint_as_float(*(unsigned int*)rs.maxbits);
  else maxs = (T)__longlong_as_double(*(long long*)rs.maxbits);
  if (mode == 0) {
    T avg = (T)scal[1];
    if (avg == (T)0) {
      avg = (T)s * (T)4 / (T)rows / (T)cols;
      scal[1] = (double)avg;
    }
    if (avg * (T)5 < maxs) maxs = avg * (T)5;
  }
  scal[0] = (double)(T)((double)maxs * shrinkThreshold);
  scal[2] = (double)maxs;
  if (sizeof(T) == 4) *(unsigned int*)rs.maxbits = 0u;
  else *(unsigned long long*)rs.maxbits = 0ull;
}
template <class T>
hipError_t launch_threshold(double* scal, const ReduceSlots& rs, double shrinkThreshold, int rows, int cols,
                            int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_threshold<T>, dim3(1), dim3(64), 0, st, scal, rs, shrinkThreshold, rows, cols, mode);
  return hipGetLastError();
}

// Real-space shrink-wrap blur (applyGaussConv, misc.cc:80-86 with createGauss
// cudaConfig.cu:447-457 and applyConvolution :475-507): the 2D kernel
// exp(-(dx^2+dy^2)/(2s^2))/(2 pi s^2) on a (2r+1)^2 window, r = floor(6s)>>1,
// zero padding, is the product of two 1D kernels, so it runs as a row pass
// and a column pass through LDS.
template <class T, bool CPXIN>
__global__ void k_blur_rows(int rows, int cols, const void* __restrict__ src, T* __restrict__ tmp, T sigma, int r) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* line = reinterpret_cast<T*>(smem);  // cols + 2r
  T* w = line + cols + 2 * r;            // 2r+1 weights
  const int x = blockIdx.x;
  const double norm = 1.0 / (2.5066282746310002 * (double)sigma);  // 1/(sqrt(2 pi) s)
  for (int j = threadIdx.x; j <= 2 * r; j += blockDim.x) {
    const int d = j - r;
    w[j] = (T)(norm * exp(-(double)(d * d) / (2.0 * (double)sigma * (double)sigma)));
  }
  for (int j = threadIdx.x; j < cols + 2 * r; j += blockDim.x) {
    const int y = j - r;
    T v = 0;
    if (y >= 0 && y < cols) {
      if constexpr (CPXIN) {
        const cplx<T> a = ((const cplx<T>*)src)[(size_t)x * cols + y];
        v = a.x * a.x + a.y * a.y;
      } else {
        v = ((const T*)src)[(size_t)x * cols + y];
      }
    }
    line[j] = v;
  }
  __syncthreads();
  for (int y = threadIdx.x; y < cols; y += blockDim.x) {
    T s = 0;
    for (int k = 0; k <= 2 * r; k++) s += line[y + k] * w[k];
    tmp[(size_t)x * cols + y] = s;
  }
}
// column pass: a workgroup owns CW adjacent columns (64 B per row), staged
// through LDS: (rows + 2r) * 64 B <= 133 KB for 2048 rows
template <class T>
__global__ void k_blur_cols(int rows, int cols, const T* __restrict__ tmp, T* __restrict__ spt, T sigma, int r,
                            ReduceSlots rs) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int CW = 64 / sizeof(T);
  T* tile = reinterpret_cast<T*>(smem);  // (rows + 2r) x CW
  T* w = tile + (rows + 2 * r) * CW;
  const int c0 = blockIdx.x * CW;
  const double norm = 1.0 / (2.5066282746310002 * (double)sigma);
  for (int j = threadIdx.x; j <= 2 * r; j += blockDim.x) {
    const int d = j - r;
    w[j] = (T)(norm * exp(-(double)(d * d) / (2.0 * (double)sigma * (double)sigma)));
  }
  for (int j = threadIdx.x; j < (rows + 2 * r) * CW; j += blockDim.x) {
    const int xx = j / CW - r, cc = j % CW;
    const int y = c0 + cc;
    tile[j] = (xx >= 0 && xx < rows && y < cols) ? tmp[(size_t)xx * cols + y] : (T)0;
  }
  __syncthreads();
  T m = 0;
  double sum = 0;
  for (int j = threadIdx.x; j < rows * CW; j += blockDim.x) {
    const int x = j / CW, cc = j % CW;
    const int y = c0 + cc;
    T s = 0;
    for (int k = 0; k <= 2 * r; k++) s += tile[(x + k) * CW + cc] * w[k];
    if (y < cols) {
      spt[(size_t)x * cols + y] = s;
      m = s > m ? s : m;
      sum += (double)s;
    }
  }
  __shared__ double sh[16];
  block_atomic_max(m, rs.maxbits, sh);
  block_partial_sum(sum, rs, sh);
}
// g complex (blur of |g|^2) when cpx_in, else a real image
template <class T>
hipError_t launch_blur_real(int rows, int cols, const void* g, int cpx_in, T* tmp, T* spt, T sigma,
                            const ReduceSlots& rs, hipStream_t st) {
  const int r = ((int)floor((double)sigma * 6)) >> 1;
  size_t l1 = (size_t)(cols + 2 * r + 2 * r + 1) * sizeof(T);
  if (cpx_in)
    hipLaunchKernelGGL((k_blur_rows<T, true>), dim3(rows), dim3(256), l1, st, rows, cols, g, tmp, sigma, r);
  else
    hipLaunchKernelGGL((k_blur_rows<T, false>), dim3(rows), dim3(256), l1, st, rows, cols, g, tmp, sigma, r);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  constexpr int CW = 64 / sizeof(T);
  size_t l2 = (size_t)((rows + 2 * r) * CW + 2 * r + 1) * sizeof(T);
  if (l2 > 160 * 1024) return hipErrorInvalidValue;
  auto kern = k_blur_cols<T>;
  if ((e = lds_cap((const void*)kern, l2)) != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3((cols + CW - 1) / CW), dim3(256), l2, st, rows, cols, (const T*)tmp, spt, sigma, r,
                     rs);
  return hipGetLastError();
}

// LDS of the blur's column kernel for `rows`-point columns at sigma
template <class T>
static size_t blur_col_lds(int rows, T sigma) {
  const int r = ((int)floor((double)sigma * 6)) >> 1;
  constexpr int CW = 64 / sizeof(T);
  return (size_t)((rows + 2 * r) * CW + 2 * r + 1) * sizeof(T);
}
template <class T>
hipError_t prepare_blur(int rows, T sigma) {
  const size_t l2 = blur_col_lds<T>(rows, sigma);
  if (l2 > 160 * 1024) return hipErrorInvalidValue;
  return lds_cap((const void*)k_blur_cols<T>, l2);
}

// Start state (prepareIter, cdi.cc:222-223): createWaveFront(sqrt(I)) then
// applyRandomPhase (cdi.cu:486-502), zero where bs > threshold.
// mode 0: random sign (the reference's (U>0.5)*pi phase, evaluated as
//         sincosf(pi) in float); mode 1: uniform phase in [-pi, pi).
template <class T>
__global__ void k_start(int64_t n, const T* I, const T* bs, double thr, uint64_t seed, int mode, cplx<T>* F) {
  EW_LOOP(n) {
    if (bs && (double)bs[i] > thr) {
      F[i] = {(T)0, (T)0};
      continue;
    }
    const T Iv = I[i];
    const T m = Iv > (T)0 ? dsqrt(Iv) : (T)0;
    const double u = philox_uniform(seed, 3, (uint64_t)i);
    if (mode == 0) {
      if (u > 0.5) F[i] = {-m, m * (T)(-8.742277657347586e-08)};
      else F[i] = {m, (T)0};
    } else {
      const double ph = -kPi + 2.0 * kPi * u;
      F[i] = {(T)(m * cos(ph)), (T)(m * sin(ph))};
    }
  }
}
template <class T>
hipError_t launch_start(int64_t n, const T* I, const T* bs, double thr, uint64_t seed, int mode, void* F,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_start<T>, ew_grid(n), dim3(256), 0, st, n, I, bs, thr, seed, mode, (cplx<T>*)F);
  return hipGetLastError();
}

// plain device copy in 16-byte words (tail bytes by the last thread)
__global__ void k_copy(int64_t n16, const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t tail,
                       const unsigned char* __restrict__ s8, unsigned char* __restrict__ d8) {
  EW_LOOP(n16) dst[i] = src[i];
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int64_t k = 0; k < tail; k++) d8[n16 * 16 + k] = s8[n16 * 16 + k];
}
hipError_t launch_copy(size_t bytes, const void* src, void* dst, hipStream_t st) {
  const int64_t n16 = (int64_t)(bytes / 16);
  hipLaunchKernelGGL(k_copy, ew_grid(n16 > 0 ? n16 : 1), dim3(256), 0, st, n16, (const uint4*)src, (uint4*)dst,
                     (int64_t)(bytes % 16), (const unsigned char*)src, (unsigned char*)dst);
  return hipGetLastError();
}

__global__ void k_fill_f64(int64_t n, double* p, double v) {
  EW_LOOP(n) p[i] = v;
}
hipError_t launch_fill_f64(int64_t n, double* p, double v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_f64, ew_grid(n), dim3(256), 0, st, n, p, v);
  return hipGetLastError();
}

// Final residual of phaseRetrieve (cdi.cc:413-421): sum over I < 0.99 of
// ((|F|^2 - I)*(1-bs))^2
template <class T>
__global__ void k_residual(int64_t n, const cplx<T>* F, const T* I, const T* bs, double* acc) {
  double s = 0;
  EW_LOOP(n) {
    const cplx<T> f = F[i];
    T d = I[i] < (T)0.99 ? (f.x * f.x + f.y * f.y) - I[i] : (T)0;
    if (bs) d *= (T)1 - bs[i];
    s += (double)(d * d);
  }
  __shared__ double sh[16];
  block_atomic_sum(s, acc, sh);
}
template <class T>
hipError_t launch_residual(int64_t n, const void* F, const T* I, const T* bs, double* acc, hipStream_t st) {
  hipLaunchKernelGGL(k_residual<T>, red_grid(n), dim3(256), 0, st, n, (const cplx<T>*)F, I, bs, acc);
  return hipGetLastError();
}

// R-factor sums of an explicit F against the packed intensities
template <class T>
__global__ void k_rfactor(int64_t n, const cplx<T>* F, const T* Ipk, double* acc) {
  double a = 0, b = 0;
  EW_LOOP(n) {
    const T Iv = Ipk[i];
    if (isnan(Iv)) continue;
    const T s = Iv > (T)0 ? dsqrt(Iv) : (T)0;
    a += (double)fabs(dhypot(F[i].x, F[i].y) - s);
    b += (double)s;
  }
  __shared__ double sh[16];
  block_atomic_sum(a, acc, sh);
  block_atomic_sum(b, acc + 1, sh);
}
template <class T>
hipError_t launch_rfactor(int64_t n, const void* F, const T* Ipk, double* acc, hipStream_t st) {
  hipLaunchKernelGGL(k_rfactor<T>, red_grid(n), dim3(256), 0, st, n, (const cplx<T>*)F, Ipk, acc);
  return hipGetLastError();
}

template <class T>
__global__ void k_cast_c(int64_t n, const void* in, int in_f64, cplx<T>* out) {
  EW_LOOP(n) {
    if (in_f64) {
      const cplx<double> a = ((const cplx<double>*)in)[i];
      out[i] = {(T)a.x, (T)a.y};
    } else {
      const cplx<float> a = ((const cplx<float>*)in)[i];
      out[i] = {(T)a.x, (T)a.y};
    }
  }
}
template <class T>
hipError_t launch_cast_c(int64_t n, const void* in, int in_f64, void* out, hipStream_t st) {
  hipLaunchKernelGGL(k_cast_c<T>, ew_grid(n), dim3(256), 0, st, n, in, in_f64, (cplx<T>*)out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// synthetic forward model (same streams and arithmetic as the CPU oracle)
//
// The kernels take a stack of B frames of rows x cols side by side: element
// (r, c) of frame b at r * B * cols + b * cols + c (B = 1: one frame in its
// usual layout).  Per frame they use frame-local pixel indices for the
// Philox counters and the one-frame summation orders, so a frame of a stack
// equals the frame made alone.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void stack_pos(int64_t k, int cols, int64_t W, int& r, int& b, int& c) {
  r = (int)(k / W);
  const int rem = (int)(k - (int64_t)r * W);
  b = rem / cols;
  c = rem - b * cols;
}
// both fields of the object: U(0,1) from Philox streams 0 (amplitude) and 1 (phase)
__global__ void k_uniform2_stack(int n, SynthFrames f, double* __restrict__ a, double* __restrict__ ph) {
  const int64_t W = (int64_t)f.B * n;
  EW_LOOP(W * n) {
    int r, b, c;
    stack_pos(i, n, W, r, b, c);
    const uint64_t il = (uint64_t)r * n + c;
    a[i] = philox_uniform(f.seed[b], 0, il);
    ph[i] = philox_uniform(f.seed[b], 1, il);
  }
}
// periodic Gaussian smoothing along one axis, radius ceil(4 sigma), weights
// normalised to sum 1 (accumulated in the same order as the oracle).  Each
// workgroup forms the 2r+1 weights once in LDS, with the expression and
// summation order of the per-pixel form, so the table holds the same values.
constexpr int kSmoothMaxTaps = 257;
__global__ void __launch_bounds__(256) k_smooth(int n, const double* in, double* out, double sigma, int axis) {
  __shared__ double wt[kSmoothMaxTaps];
  const int r = (int)ceil(4 * sigma);
  const int taps = 2 * r + 1;
  const bool table = taps <= kSmoothMaxTaps;
  if (table) {
    if (threadIdx.x == 0) {
      double ws = 0;
      for (int k = -r; k <= r; k++) ws += exp(-(double)k * k / (2 * sigma * sigma));
      wt[0] = ws;
    }
    __syncthreads();
    const double ws = wt[0];
    __syncthreads();
    for (int t = threadIdx.x; t < taps; t += blockDim.x) {
      const int k = t - r;
      wt[t] = exp(-(double)k * k / (2 * sigma * sigma)) / ws;
    }
    __syncthreads();
  }
  const int64_t N = (int64_t)n * n;
  EW_LOOP(N) {
    const int x = (int)(i / n), y = (int)(i % n);
    double ws = 0;
    if (!table)
      for (int k = -r; k <= r; k++) ws += exp(-(double)k * k / (2 * sigma * sigma));
    double s = 0;
    // unrolled so that several taps' loads are in flight (same summation order)
#pragma unroll 8
    for (int k = -r; k <= r; k++) {
      const double w = table ? wt[k + r] : exp(-(double)k * k / (2 * sigma * sigma)) / ws;
      const int64_t j = axis ? (int64_t)(((x + k) % n + n) % n) * n + y : (int64_t)x * n + ((y + k) % n + n) % n;
      s += w * in[j];
    }
    out[i] = s;
  }
}
// The same smoothing with one workgroup per line: the line (row for axis 0,
// column for axis 1) and the weights are read into LDS once, then every
// output sums its taps from LDS in the order of k_smooth, so the results are
// identical.  Lines of up to kSmoothMaxLine points, radius below n.
constexpr int kSmoothMaxLine = 2048;
// Every line of a stack of B frames, two fields (blockIdx.y) per launch.
__global__ void __launch_bounds__(256) k_smooth_line_stack(int n, int B, const double* in0, double* out0,
                                                           const double* in1, double* out1, double sigma, int axis) {
  __shared__ double wt[kSmoothMaxTaps];
  __shared__ double line[kSmoothMaxLine];
  const double* in = blockIdx.y ? in1 : in0;
  double* out = blockIdx.y ? out1 : out0;
  const int r = (int)ceil(4 * sigma);
  const int taps = 2 * r + 1;
  const int b = blockIdx.x / n, l = blockIdx.x % n;
  const int64_t W = (int64_t)B * n;
  if (threadIdx.x == 0) {
    double ws = 0;
    for (int k = -r; k <= r; k++) ws += exp(-(double)k * k / (2 * sigma * sigma));
    wt[0] = ws;
  }
  __syncthreads();
  const double ws = wt[0];
  __syncthreads();
  for (int t = threadIdx.x; t < taps; t += blockDim.x) {
    const int k = t - r;
    wt[t] = exp(-(double)k * k / (2 * sigma * sigma)) / ws;
  }
  // element p of line l of frame b
  auto at = [&](int p) { return axis ? (int64_t)p * W + (int64_t)b * n + l : (int64_t)l * W + (int64_t)b * n + p; };
  for (int p = threadIdx.x; p < n; p += blockDim.x) line[p] = in[at(p)];
  __syncthreads();
  for (int p = threadIdx.x; p < n; p += blockDim.x) {
    double s = 0;
    for (int k = -r; k <= r; k++) {
      int q = p + k;
      q += q < 0 ? n : (q >= n ? -n : 0);
      s += wt[k + r] * line[q];
    }
    out[at(p)] = s;
  }
}
// Min and max of each frame of two fields as order-preserving integer keys
// (doubles of any sign): mm[2 * (field * B + b) + {0, 1}], seeded with the
// keys of (+inf, -inf).  Waves reduce by shuffles, the workgroup in LDS, and
// one atomic pair per workgroup combines them.
__global__ void k_seed_minmax_stack(int count, unsigned long long* mm) {
  for (int k = threadIdx.x; k < count; k += blockDim.x) {
    mm[2 * k] = ~0ull;
    mm[2 * k + 1] = 0ull;
  }
}
__global__ void __launch_bounds__(256) k_minmax_stack(int n, int B, const double* x0, const double* x1, double* mm) {
  __shared__ double slo[4], shi[4];
  const double* x = blockIdx.z ? x1 : x0;
  const int b = blockIdx.y;
  const int64_t W = (int64_t)B * n, N = (int64_t)n * n;
  double lo = INFINITY, hi = -INFINITY;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[(i / n) * W + (int64_t)b * n + i % n];
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    slo[wave] = lo;
    shi[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      lo = fmin(lo, slo[w]);
      hi = fmax(hi, shi[w]);
    }
    long long bl = __double_as_longlong(lo), bh = __double_as_longlong(hi);
    unsigned long long kl = bl < 0 ? ~(unsigned long long)bl : ((unsigned long long)bl | 0x8000000000000000ULL);
    unsigned long long kh = bh < 0 ? ~(unsigned long long)bh : ((unsigned long long)bh | 0x8000000000000000ULL);
    unsigned long long* dst = (unsigned long long*)mm + 2 * ((int)blockIdx.z * B + b);
    atomicMin(dst, kl);
    atomicMax(dst + 1, kh);
  }
}
__device__ __forceinline__ double unkey(unsigned long long k) {
  const unsigned long long b = (k & 0x8000000000000000ULL) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
  return __longlong_as_double((long long)b);
}
// inside test of a support shape centred at (c, c): disk |d| < a (strict, as
// createMask<C_circle>), ellipse with semi-axes a, b rotated by rot, or a
// polygon star-shaped about the centre (vertices at increasing angles)
__device__ __forceinline__ bool shape_inside(const ShapeParams& sp, float dx, float dy) {
  if (sp.kind == 0) return sqrtf(dx * dx + dy * dy) < sp.a;
  if (sp.kind == 1) {
    const float c = cosf(sp.rot), s = sinf(sp.rot);
    const float u = (dx * c + dy * s) / sp.a, v = (-dx * s + dy * c) / sp.b;
    return u * u + v * v < 1.f;
  }
  float phi = atan2f(dy, dx);
  if (phi < 0.f) phi += 2.f * (float)kPi;
  int i = sp.nv - 1;  // sector [ang[i], ang[i+1]), wrapping after the last vertex
  for (int k = 0; k + 1 < sp.nv; k++)
    if (phi >= sp.ang[k] && phi < sp.ang[k + 1]) i = k;
  const int j = i + 1 < sp.nv ? i + 1 : 0;
  const float x0 = sp.rad[i] * cosf(sp.ang[i]), y0 = sp.rad[i] * sinf(sp.ang[i]);
  const float x1 = sp.rad[j] * cosf(sp.ang[j]), y1 = sp.rad[j] * sinf(sp.ang[j]);
  return (x1 - x0) * (dy - y0) - (y1 - y0) * (dx - x0) >= 0.f;
}
template <class T>
__global__ void k_compose_stack(int n, SynthFrames f, const double* __restrict__ a, const double* __restrict__ ph,
                                const double* __restrict__ mm, cplx<T>* __restrict__ obj) {
  const int64_t W = (int64_t)f.B * n;
  const int c0 = n >> 1;
  EW_LOOP(W * n) {
    int x, b, y;
    stack_pos(i, n, W, x, b, y);
    const unsigned long long* ka = (const unsigned long long*)mm + 2 * b;
    const unsigned long long* kp = (const unsigned long long*)mm + 2 * (f.B + b);
    const double amin = unkey(ka[0]), amax = unkey(ka[1]);
    const double pmin = unkey(kp[0]), pmax = unkey(kp[1]);
    if (!shape_inside(f.sp[b], (float)(x - c0), (float)(y - c0))) {
      obj[i] = {(T)0, (T)0};
      continue;
    }
    const double amp = 0.2 + 0.8 * (a[i] - amin) / (amax - amin);
    const double phs = -kPi + 2 * kPi * (ph[i] - pmin) / (pmax - pmin);
    obj[i] = {(T)(amp * cos(phs)), (T)(amp * sin(phs))};
  }
}
// sum |F|^2 per frame of rows x cols (one atomic per workgroup into acc[b])
template <class T>
__global__ void k_sum_mod2_stack(int rows, int cols, int B, const cplx<T>* __restrict__ F, double* acc) {
  const int b = blockIdx.y;
  const int64_t W = (int64_t)B * cols, N = (int64_t)rows * cols;
  double s = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
    const cplx<T> z = F[(i / cols) * W + (int64_t)b * cols + i % cols];
    s += (double)z.x * z.x + (double)z.y * z.y;
  }
  __shared__ double sh[16];
  block_atomic_sum(s, acc + b, sh);
}
// Poisson(lam): multiplication method below 10, PTRS (Hoermann 1993) above;
// uniforms from Philox stream 2, block counter (pixel, j).
struct UStream {
  uint64_t seed, idx, blk;
  uint64_t buf[4];
  int pos;
  __device__ double next() {
    if (pos == 4) {
      buf[0] = idx; buf[1] = blk++; buf[2] = 0; buf[3] = 0;
      philox4x64(buf, seed, 2);
      pos = 0;
    }
    return (double)(buf[pos++] >> 11) * 0x1.0p-53;
  }
};
__device__ double poisson_draw(double lam, uint64_t seed, uint64_t idx) {
  UStream u{seed, idx, 0, {0, 0, 0, 0}, 4};
  if (!(lam > 0) || !(lam < 1e300)) return 0;  // also NaN / inf
  if (lam < 10) {
    const double enlam = exp(-lam);
    double prod = 1.0, X = 0;
    for (int it = 0; it < 4096; it++) {
      prod *= u.next();
      if (prod > enlam) X += 1;
      else return X;
    }
    return X;
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2);
  // PTRS accepts ~90% of proposals; the cap only bounds the loop
  for (int it = 0; it < 4096; it++) {
    const double U = u.next() - 0.5;
    const double V = u.next();
    const double us = 0.5 - fabs(U);
    const double k = floor((2 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1)) return k;
  }
  return floor(lam + 0.5);
}
// I = |F|^2 photons / sum |F|^2 per frame, drawn Poisson (counter: the
// frame-local pixel index) or kept as the mean
template <class T>
__global__ void k_poisson_stack(int rows, int cols, SynthFrames f, const cplx<T>* __restrict__ F,
                                const double* __restrict__ tot, int poisson, T* __restrict__ I) {
  const int64_t W = (int64_t)f.B * cols;
  EW_LOOP(W * rows) {
    int r, b, c;
    stack_pos(i, cols, W, r, b, c);
    const double sc = tot[b] > 0 ? f.photons[b] / tot[b] : 0;
    const double lam = ((double)F[i].x * F[i].x + (double)F[i].y * F[i].y) * sc;
    I[i] = poisson ? (T)poisson_draw(lam, f.seed[b], (uint64_t)r * cols + c) : (T)lam;
  }
}
bool synth_stack_supported(int n, double sigma) {
  const int r = (int)ceil(4 * sigma);
  return n <= kSmoothMaxLine && 2 * r + 1 <= kSmoothMaxTaps && r < n;
}
template <class T>
hipError_t launch_synth_fields(int n, const SynthFrames& f, double sigma, double* a, double* ph, double* ta,
                               double* tph, double* mm, void* obj, hipStream_t st) {
  const bool lines = synth_stack_supported(n, sigma);
  if (f.B < 1 || f.B > kSynthMax || (!lines && f.B != 1)) return hipErrorInvalidValue;
  const int64_t tot = (int64_t)f.B * n * n;
  hipLaunchKernelGGL(k_uniform2_stack, ew_grid(tot), dim3(256), 0, st, n, f, a, ph);
  if (lines) {
    hipLaunchKernelGGL(k_smooth_line_stack, dim3(f.B * n, 2), dim3(256), 0, st, n, f.B, (const double*)a, ta,
                       (const double*)ph, tph, sigma, 0);
    hipLaunchKernelGGL(k_smooth_line_stack, dim3(f.B * n, 2), dim3(256), 0, st, n, f.B, (const double*)ta, a,
                       (const double*)tph, ph, sigma, 1);
  } else {  // one frame, lines beyond the line kernel's limits
    const dim3 g = ew_grid((int64_t)n * n);
    hipLaunchKernelGGL(k_smooth, g, dim3(256), 0, st, n, (const double*)a, ta, sigma, 0);
    hipLaunchKernelGGL(k_smooth, g, dim3(256), 0, st, n, (const double*)ta, a, sigma, 1);
    hipLaunchKernelGGL(k_smooth, g, dim3(256), 0, st, n, (const double*)ph, tph, sigma, 0);
    hipLaunchKernelGGL(k_smooth, g, dim3(256), 0, st, n, (const double*)tph, ph, sigma, 1);
  }
  hipLaunchKernelGGL(k_seed_minmax_stack, dim3(1), dim3(64), 0, st, 2 * f.B, (unsigned long long*)mm);
  // 4 elements per thread at least, at most 256 workgroups per frame
  int64_t nb = ((int64_t)n * n + 256 * 4 - 1) / (256 * 4);
  nb = nb < 1 ? 1 : (nb > 256 ? 256 : nb);
  hipLaunchKernelGGL(k_minmax_stack, dim3((unsigned)nb, f.B, 2), dim3(256), 0, st, n, f.B, (const double*)a,
                     (const double*)ph, mm);
  hipLaunchKernelGGL(k_compose_stack<T>, ew_grid(tot), dim3(256), 0, st, n, f, (const double*)a, (const double*)ph,
                     (const double*)mm, (cplx<T>*)obj);
  return hipGetLastError();
}
template <class T>
hipError_t launch_synth_intensity(int rows, int cols, const SynthFrames& f, const void* F, double* tot, int poisson,
                                  T* I, hipStream_t st) {
  if (f.B < 1 || f.B > kSynthMax) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(256), 0, st, (int64_t)f.B, tot, 0.0);
  const dim3 g1 = red_grid((int64_t)rows * cols);  // per frame
  hipLaunchKernelGGL(k_sum_mod2_stack<T>, dim3(g1.x, f.B), dim3(256), 0, st, rows, cols, f.B, (const cplx<T>*)F, tot);
  hipLaunchKernelGGL(k_poisson_stack<T>, ew_grid((int64_t)f.B * rows * cols), dim3(256), 0, st, rows, cols, f,
                     (const cplx<T>*)F, (const double*)tot, poisson, I);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// setup kernels of CDI::prepareIter / readObjectWave / checkAutoCorrelation
// ---------------------------------------------------------------------------
// convertFOPhase (cudaConfig.cu:861-870): multiply by (-1)^(x+y) * norm
template <class T>
__global__ void k_fo_phase(int rows, int cols, cplx<T>* d, T norm) {
  const int64_t n = (int64_t)rows * cols;
  EW_LOOP(n) {
    const int x = (int)(i / cols), y = (int)(i % cols);
    const T f = ((x + y) & 1) ? -norm : norm;
    d[i] = {d[i].x * f, d[i].y * f};
  }
}
template <class T>
hipError_t launch_fo_phase(int rows, int cols, void* d, T norm, hipStream_t st) {
  hipLaunchKernelGGL(k_fo_phase<T>, ew_grid((int64_t)rows * cols), dim3(256), 0, st, rows, cols, (cplx<T>*)d, norm);
  return hipGetLastError();
}
// out = in * (invert ? 1 - m : m)   (linearConst + multiply, cdi.cc:185-187)
template <class T>
__global__ void k_mask_mul(int64_t n, T* d, const T* m, int invert) {
  EW_LOOP(n) d[i] *= invert ? (T)1 - m[i] : m[i];
}
template <class T>
hipError_t launch_mask_mul(int64_t n, T* d, const T* m, int invert, hipStream_t st) {
  hipLaunchKernelGGL(k_mask_mul<T>, ew_grid(n), dim3(256), 0, st, n, d, m, invert);
  return hipGetLastError();
}
// applyPoissonNoise_WO (cdi.cu:291-296): quantise to 1/rcolor steps of scale
// and add zero-mean dark noise Poisson(noiseLevel) - noiseLevel
template <class T>
__global__ void k_poisson_wo(int64_t n, T* w, double noise, T scale, T rcolor, uint64_t seed) {
  EW_LOOP(n) {
    const double dark = poisson_draw(noise, seed, (uint64_t)i);
    w[i] = scale * (T)((double)(int)(w[i] * rcolor / scale) + dark - noise) / rcolor;
  }
}
template <class T>
hipError_t launch_poisson_wo(int64_t n, T* w, double noise, T scale, T rcolor, uint64_t seed, hipStream_t st) {
  hipLaunchKernelGGL(k_poisson_wo<T>, ew_grid(n), dim3(256), 0, st, n, w, noise, scale, rcolor, seed);
  return hipGetLastError();
}
// extendToComplex (cudaConfig.cu:682-686): (a, 0)
template <class T>
__global__ void k_extend(int64_t n, const T* a, cplx<T>* b) {
  EW_LOOP(n) b[i] = {a[i], (T)0};
}
template <class T>
hipError_t launch_extend(int64_t n, const T* a, void* b, hipStream_t st) {
  hipLaunchKernelGGL(k_extend<T>, ew_grid(n), dim3(256), 0, st, n, a, (cplx<T>*)b);
  return hipGetLastError();
}
// createWaveFront, integer-size form (cudaConfig.cu:688-709): an objrow x
// objcol image centred in the frame, amplitude sqrt(max(I,0)) (1 without I),
// phase 2*pi*(P - 0.5) when a phase image is given
template <class T>
__global__ void k_wavefront(int rows, int cols, const T* I, const T* P, int orow, int ocol, cplx<T>* out) {
  const int64_t n = (int64_t)rows * cols;
  const int mx = (rows - orow) / 2, my = (cols - ocol) / 2;
  EW_LOOP(n) {
    const int x = (int)(i / cols), y = (int)(i % cols);
    if (x < mx || x >= mx + orow || y < my || y >= my + ocol) {
      out[i] = {(T)0, (T)0};
      continue;
    }
    const int64_t t = (int64_t)(x - mx) * ocol + (y - my);
    const T mod = I ? dsqrt(I[t] > (T)0 ? I[t] : (T)0) : (T)1;
    if (P) {
      const T ph = (P[t] - (T)0.5) * (T)2 * (T)kPi;
      out[i] = {mod * (T)cos(ph), mod * (T)sin(ph)};
    } else {
      out[i] = {mod, (T)0};
    }
  }
}
template <class T>
hipError_t launch_wavefront(int rows, int cols, const T* I, const T* P, int orow, int ocol, void* out,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_wavefront<T>, ew_grid((int64_t)rows * cols), dim3(256), 0, st, rows, cols, I, P, orow, ocol,
                     (cplx<T>*)out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// PRTF (cdi.cc:355-380 accumulation, 423-440 final)
// ---------------------------------------------------------------------------
// One wavefront: lane-strided partial sums in a fixed order, then a wave sum.
__global__ void k_prtf_phase(double* phi0, double* theta, const double* partials, const int* nparts, int first) {
  const int np = *nparts;
  double re = 0, im = 0;
  for (int i = threadIdx.x; i < np; i += 64) {
    re += partials[2 * i];
    im += partials[2 * i + 1];
  }
  re = wave_sum(re);
  im = wave_sum(im);
  if (threadIdx.x == 0) {
    if (first) {
      *phi0 = -atan2(im, re) / 2;
      *theta = *phi0;
    } else {
      *theta = *phi0 + atan2(im, re);
    }
  }
}
hipError_t launch_prtf_phase(double* phi0, double* theta, const ReduceSlots& rs, int first, hipStream_t st) {
  hipLaunchKernelGGL(k_prtf_phase, dim3(1), dim3(64), 0, st, phi0, theta, (const double*)rs.partials, rs.nparts,
                     first);
  return hipGetLastError();
}

// addPhase (cudaConfig.cu:876-887)
template <class T>
__global__ void k_prtf_add(int64_t n, cplx<T>* __restrict__ map, const cplx<T>* __restrict__ F,
                           const double* theta) {
  const T th = (T)*theta;
  const T c = cos(th), s = sin(th);
  EW_LOOP(n) {
    const cplx<T> b = F[i];
    const T m = dhypot(b.x, b.y);
    if (m > (T)1e-6) {
      const T r = b.x / m, q = b.y / m;
      cplx<T> a = map[i];
      a.x += r * c - q * s;
      a.y += r * s + q * c;
      map[i] = a;
    }
  }
}
template <class T>
hipError_t launch_prtf_add(int64_t n, void* map, const void* F, const double* theta, hipStream_t st) {
  hipLaunchKernelGGL(k_prtf_add<T>, ew_grid(n), dim3(256), 0, st, n, (cplx<T>*)map, (const cplx<T>*)F, theta);
  return hipGetLastError();
}

// cart2polar_kernel + edgeReduce (cudaConfig.cu:1027-1063) over the fftshifted
// image (cudaConvertFO, even sizes: a half-size shift on both axes), one
// workgroup per radius.  Bilinear weights as bilinear(), cudaConfig.cu:1027-1047.
template <class T>
__device__ __forceinline__ T radial_sample(const cplx<T>* src, int rows, int cols, int x, int y, int kind, T norm) {
  const int xs = x + (rows >> 1) < rows ? x + (rows >> 1) : x - (rows >> 1);
  const int ys = y + (cols >> 1) < cols ? y + (cols >> 1) : y - (cols >> 1);
  const cplx<T> v = src[(size_t)xs * cols + ys];
  const T m = dhypot(v.x, v.y);
  return kind == 0 ? (m > (T)1e-6 ? (T)1 : (T)0) : m * norm;
}
template <class T>
__global__ void k_radial(const cplx<T>* __restrict__ src, int rows, int cols, int kind, T norm, double* out) {
  const int r = blockIdx.x;
  const int nang = 2 * cols;
  double acc = 0;
  for (int a = threadIdx.x; a < nang; a += blockDim.x) {
    const T th = (T)(a * (2.0 * M_PI / nang));
    const T X = (T)r * cos(th) + (T)(rows >> 1);
    const T Y = (T)r * sin(th) + (T)(cols >> 1);
    const int x0 = (int)floor(X), y0 = (int)floor(Y);
    if (x0 < 0 || x0 + 1 >= rows || y0 < 0 || y0 + 1 >= cols) continue;
    const T fx = X - (T)x0, fy = Y - (T)y0;
    const T v00 = radial_sample(src, rows, cols, x0, y0, kind, norm);
    const T v01 = radial_sample(src, rows, cols, x0, y0 + 1, kind, norm);
    const T v10 = radial_sample(src, rows, cols, x0 + 1, y0, kind, norm);
    const T v11 = radial_sample(src, rows, cols, x0 + 1, y0 + 1, kind, norm);
    acc += (double)((1 - fx) * (1 - fy) * v00 + (1 - fx) * fy * v01 + fx * (1 - fy) * v10 + fx * fy * v11);
  }
  __shared__ double sh[16];
  block_store_sum(acc, out + r, sh);
}
template <class T>
hipError_t launch_radial(const void* src, int rows, int cols, int kind, T norm, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_radial<T>, dim3(rows / 2), dim3(256), 0, st, (const cplx<T>*)src, rows, cols, kind, norm,
                     out);
  return hipGetLastError();
}

// explicit instantiations
#define INST(T)                                                                                                  \
  template hipError_t launch_fo_phase<T>(int, int, void*, T, hipStream_t);                                       \
  template hipError_t launch_mask_mul<T>(int64_t, T*, const T*, int, hipStream_t);                               \
  template hipError_t launch_poisson_wo<T>(int64_t, T*, double, T, T, uint64_t, hipStream_t);                    \
  template hipError_t launch_extend<T>(int64_t, const T*, void*, hipStream_t);                                   \
  template hipError_t launch_wavefront<T>(int, int, const T*, const T*, int, int, void*, hipStream_t);           \
  template hipError_t launch_pack_intensity<T>(int64_t, const T*, const T*, T*, int, hipStream_t);                    \
  template hipError_t launch_pack_intensity_tiled<T>(int, int, const T*, const T*, T*, int, hipStream_t);             \
  template hipError_t launch_mod_abs<T>(int64_t, void*, const T*, hipStream_t);                                  \
  template hipError_t launch_mod_noise<T>(int64_t, void*, const T*, const T*, int, T, T, T, hipStream_t);        \
  template hipError_t launch_support<T>(int64_t, void*, const void*, const T*, const UpdArgs<T>&, hipStream_t);  \
  template hipError_t launch_norm_c<T>(int64_t, void*, T, hipStream_t);                                          \
  template hipError_t launch_norm_r<T>(int64_t, T*, T, hipStream_t);                                             \
  template hipError_t launch_mod2<T>(int64_t, T*, const void*, T, hipStream_t);                                  \
  template hipError_t launch_gauss_mult_freq<T>(int, int, const void*, void*, T, hipStream_t);                   \
  template hipError_t launch_convert_fo<T>(int, int, const void*, void*, int, hipStream_t);                      \
  template hipError_t launch_rect_mask<T>(int, int, int, int, int, int, T*, hipStream_t);                        \
  template hipError_t launch_circle_mask<T>(int, int, float, int, T*, hipStream_t);                              \
  template hipError_t launch_reduce<T>(int64_t, const T*, const ReduceSlots&, hipStream_t);                      \
  template hipError_t launch_threshold<T>(double*, const ReduceSlots&, double, int, int, int, hipStream_t);      \
  template hipError_t launch_blur_real<T>(int, int, const void*, int, T*, T*, T, const ReduceSlots&, hipStream_t);    \
  template hipError_t prepare_blur<T>(int, T);                                                                       \
  template hipError_t launch_start<T>(int64_t, const T*, const T*, double, uint64_t, int, void*, hipStream_t);   \
  template hipError_t launch_residual<T>(int64_t, const void*, const T*, const T*, double*, hipStream_t);        \
  template hipError_t launch_rfactor<T>(int64_t, const void*, const T*, double*, hipStream_t);                   \
  template hipError_t launch_cast_c<T>(int64_t, const void*, int, void*, hipStream_t);                           \
  template hipError_t launch_synth_fields<T>(int, const SynthFrames&, double, double*, double*, double*, double*, \
                                             double*, void*, hipStream_t);                                         \
  template hipError_t launch_synth_intensity<T>(int, int, const SynthFrames&, const void*, double*, int, T*,       \
                                                hipStream_t);                                                      \
  template hipError_t launch_prtf_add<T>(int64_t, void*, const void*, const double*, hipStream_t);               \
  template hipError_t launch_radial<T>(const void*, int, int, int, T, double*, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace cdx
