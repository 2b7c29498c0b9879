// preprocess.hip -- pattern preprocessing before reconstruction, the steps of
// CDIxx util/processimg.cpp:11-109: background subtraction, optional centring
// on the thresholded centroid (findMiddle, src/post_gpu/misc.cc:27-42) with a
// sub-pixel FFT shift (shiftWave, misc.cc:97-104) or the modulus re-centring
// (shiftMiddle, misc.cc:88-95), cropping to a multiple of 4*nmerge, and
// nmerge x nmerge pixel binning (mergePixel, src/gpu/cudaConfig.cu:949-961).
#include "../../include/cdx_abi.h"
#include "device_ops.hpp"
#include "errors.hpp"
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <vector>

using namespace cdx;

namespace {

int pfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  cdx::set_last_error(buf);
  return code;
}
#define PCHK(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return pfail(CDX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
#define PRC(expr)      \
  do {                 \
    int r_ = (expr);   \
    if (r_) return r_; \
  } while (0)

int64_t blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

// crop (cudaConfig.cu:348-358): dest[x][y] = src[(x + (R - r)/2 + sx) * C + y + (C - c)/2 + sy]
__global__ void k_crop(const float* __restrict__ src, int R, int C, float* __restrict__ dst, int r, int c, int sx,
                       int sy) {
  const int64_t n = (int64_t)r * c;
  EW_LOOP(n) {
    const int x = (int)(i / c), y = (int)(i % c);
    dst[i] = src[(int64_t)(x + (R - r) / 2 + sx) * C + y + (C - c) / 2 + sy];
  }
}
// add(d_sig, d_bkg, -1)
__global__ void k_sub(int64_t n, float* __restrict__ a, const float* __restrict__ b) {
  EW_LOOP(n) a[i] = a[i] - b[i];
}
// one workgroup: max (pass 0) or the centroid sums of applyThreshold(sig, max/2)
// weighted by ((x + 0.5)/rows - 0.5, (y + 0.5)/cols - 0.5) (multiplyx/y,
// cudaConfig.cu:1427-1437)
__global__ void k_centroid(const float* __restrict__ s, int rows, int cols, double* out) {
  const int64_t n = (int64_t)rows * cols;
  float mx = -INFINITY;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) mx = fmaxf(mx, s[i]);
  __shared__ float smx[1024];
  smx[threadIdx.x] = mx;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) smx[threadIdx.x] = fmaxf(smx[threadIdx.x], smx[threadIdx.x + o]);
    __syncthreads();
  }
  const float thr = smx[0] / 2;
  double sw = 0, sx = 0, sy = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const float v = s[i] > thr ? s[i] : 0.f;
    const int x = (int)(i / cols), y = (int)(i % cols);
    sw += v;
    sx += (double)(v * (((float)x + 0.5f) / (float)rows - 0.5f));
    sy += (double)(v * (((float)y + 0.5f) / (float)cols - 0.5f));
  }
  __shared__ double sh[16];
  block_store_sum(sw, out + 2, sh);
  block_store_sum(sx, out, sh);
  block_store_sum(sy, out + 1, sh);
  if (threadIdx.x == 0) out[3] = smx[0];
}
// multiplyShift (cudaConfig.cu:373-378) on the centred spectrum
__global__ void k_shift_phase(int r, int c, cplx<float>* __restrict__ z, float shx, float shy) {
  const int64_t n = (int64_t)r * c;
  EW_LOOP(n) {
    const int x = (int)(i / c), y = (int)(i % c);
    const float ph = -2.f * (float)M_PI * (shx * (float)(x - r / 2) / (float)r + shy * (float)(y - c / 2) / (float)c);
    float s, co;
    sincosf(ph, &s, &co);
    const cplx<float> v = z[i];
    z[i] = {v.x * co - v.y * s, v.x * s + v.y * co};
  }
}
// rotateToReal (cudaConfig.cu:441-445): z = |z|
__global__ void k_modulus(int64_t n, cplx<float>* z) {
  EW_LOOP(n) z[i] = {hypotf(z[i].x, z[i].y), 0.f};
}
// getReal then mergePixel: out[x][y] = mean of the nm x nm block of Re(z)
__global__ void k_merge_real(const cplx<float>* __restrict__ z, int c, int nm, float* __restrict__ out, int fr, int fc) {
  const int64_t n = (int64_t)fr * fc;
  EW_LOOP(n) {
    const int x = (int)(i / fc), y = (int)(i % fc);
    float s = 0.f;
    for (int dx = 0; dx < nm; dx++)
      for (int dy = 0; dy < nm; dy++) s += z[(int64_t)(x * nm + dx) * c + y * nm + dy].x;
    out[i] = s / (float)(nm * nm);
  }
}

}  // namespace

extern "C" int cdx_preprocess_pattern(const float* sig, int rows, int cols, const float* bkg, int brows, int bcols,
                                      int nmerge, int centre, double shift_row, double shift_col, int shift_mode,
                                      float* out, int cap, int* out_size, double* mid) {
  if (!sig || rows <= 0 || cols <= 0 || nmerge < 1 || !out || !out_size)
    return pfail(CDX_ERR_ARG, "cdx_preprocess_pattern: bad argument");
  if (bkg && (brows <= 0 || bcols <= 0)) return pfail(CDX_ERR_ARG, "cdx_preprocess_pattern: bad background size");
  // device copies; the larger image is cropped to the smaller (processimg.cpp:27-44,
  // which compares the row counts only)
  int R = rows, C = cols;
  float *ds = nullptr, *db = nullptr, *tmp = nullptr;
  double* sums = nullptr;
  void* z = nullptr;
  cdx_ctx* ctx = nullptr;
  std::vector<float> host;
  int rc = CDX_OK;
  auto cleanup = [&]() {
    (void)hipFree(ds);
    (void)hipFree(db);
    (void)hipFree(tmp);
    (void)hipFree(sums);
    (void)hipFree(z);
    if (ctx) cdx_destroy(ctx);
  };
  do {
    if (hipMalloc(&ds, (size_t)rows * cols * 4) || hipMalloc(&sums, 8 * sizeof(double)) ||
        hipMemcpy(ds, sig, (size_t)rows * cols * 4, hipMemcpyHostToDevice)) {
      rc = pfail(CDX_ERR_HIP, "cdx_preprocess_pattern: device allocation/copy failed");
      break;
    }
    if (bkg) {
      if (hipMalloc(&db, (size_t)brows * bcols * 4) ||
          hipMemcpy(db, bkg, (size_t)brows * bcols * 4, hipMemcpyHostToDevice)) {
        rc = pfail(CDX_ERR_HIP, "cdx_preprocess_pattern: device allocation/copy failed");
        break;
      }
      if (brows != rows) {
        const bool crop_bkg = brows > rows;
        const int r = crop_bkg ? rows : brows, c = crop_bkg ? cols : bcols;
        if (hipMalloc(&tmp, (size_t)r * c * 4)) {
          rc = pfail(CDX_ERR_HIP, "cdx_preprocess_pattern: device allocation failed");
          break;
        }
        const float* src = crop_bkg ? db : ds;
        const int sR = crop_bkg ? brows : rows, sC = crop_bkg ? bcols : cols;
        if (r > sR || c > sC) {
          rc = pfail(CDX_ERR_ARG, "cdx_preprocess_pattern: images cannot be cropped to a common size");
          break;
        }
        hipLaunchKernelGGL(k_crop, dim3(blocks((int64_t)r * c)), dim3(256), 0, 0, src, sR, sC, tmp, r, c, 0, 0);
        if (crop_bkg) {
          (void)hipFree(db);
          db = tmp;
        } else {
          (void)hipFree(ds);
          ds = tmp;
          R = r;
          C = c;
        }
        tmp = nullptr;
      } else if (bcols != cols) {
        rc = pfail(CDX_ERR_ARG, "cdx_preprocess_pattern: background has %d columns, pattern %d", bcols, cols);
        break;
      }
      hipLaunchKernelGGL(k_sub, dim3(blocks((int64_t)R * C)), dim3(256), 0, 0, (int64_t)R * C, ds, (const float*)db);
    }
    // centroid of the pattern thresholded at half its maximum (processimg.cpp:60-64)
    hipLaunchKernelGGL(k_centroid, dim3(1), dim3(1024), 0, 0, (const float*)ds, R, C, sums);
    double hs[4];
    if (hipMemcpy(hs, sums, sizeof hs, hipMemcpyDeviceToHost)) {
      rc = pfail(CDX_ERR_HIP, "cdx_preprocess_pattern: reduction failed");
      break;
    }
    float midx = (float)(hs[2] != 0 ? hs[0] / hs[2] : 0) * (float)R;
    float midy = (float)(hs[2] != 0 ? hs[1] / hs[2] : 0) * (float)C;
    if (centre) {  // processimg.cpp:69-71: argv[7] shifts rows, argv[6] columns
      midx += (float)shift_row;
      midy -= (float)shift_col;
    } else {
      midx = midy = 0.f;
    }
    if (mid) {
      mid[0] = midx;
      mid[1] = midy;
    }
    const int step = nmerge * 4;
    int outrow = (R - (int)fabsf(midx) * 2) / step * step;
    const int outcol = (C - (int)fabsf(midy) * 2) / step * step;
    outrow = outrow < outcol ? outrow : outcol;
    const int n = outrow;
    if (n < step) {
      rc = pfail(CDX_ERR_ARG, "cdx_preprocess_pattern: nothing left after centring and cropping");
      break;
    }
    if (!cdx_size_supported(n)) {
      rc = pfail(CDX_ERR_UNSUPPORTED, "cdx_preprocess_pattern: cropped size %d is not an FFT length of this build", n);
      break;
    }
    const int fin = n / nmerge;
    *out_size = fin;
    if ((int64_t)fin * fin > cap) {
      rc = pfail(CDX_ERR_ARG, "cdx_preprocess_pattern: output needs %d x %d values", fin, fin);
      break;
    }
    const float shx = (float)(int)midx - midx, shy = (float)(int)midy - midy;
    // crop around the centre (float arithmetic of crop(), cudaConfig.cu:348-358,
    // including its use of the row count in the column clamp)
    const float mxn = midx / (float)R, myn = midy / (float)C;
    int sx = (int)((float)R * mxn);
    if (sx + n / 2 > R / 2) sx = (R - n) / 2;
    else if (sx - n / 2 < -R / 2) sx = (n - R) / 2;
    int sy = (int)((float)C * myn);
    if (sy + n / 2 > C / 2) sy = (R - n) / 2;
    else if (sy - n / 2 < -C / 2) sy = (n - C) / 2;
    if ((n - 1) + (R - n) / 2 + sx >= R || (R - n) / 2 + sx < 0 || (n - 1) + (C - n) / 2 + sy >= C ||
        (C - n) / 2 + sy < 0) {
      rc = pfail(CDX_ERR_ARG, "cdx_preprocess_pattern: crop window leaves the image");
      break;
    }
    if ((rc = cdx_create(&ctx, n, n, CDX_F32, nullptr))) break;
    hipStream_t st = (hipStream_t)cdx_stream(ctx);
    if (hipDeviceSynchronize() || hipMalloc(&tmp, (size_t)n * n * 4) || hipMalloc(&z, (size_t)n * n * 8)) {
      rc = pfail(CDX_ERR_HIP, "cdx_preprocess_pattern: device allocation failed");
      break;
    }
    hipLaunchKernelGGL(k_crop, dim3(blocks((int64_t)n * n)), dim3(256), 0, st, (const float*)ds, R, C, tmp, n, n, sx,
                       sy);
    if ((rc = cdx_extend_to_complex(ctx, tmp, z))) break;
    const int64_t N = (int64_t)n * n;
    if (shift_mode == 1) {  // shiftMiddle
      if ((rc = cdx_convert_fo(ctx, z, nullptr, 1)) || (rc = cdx_fft(ctx, z, z, CDX_FORWARD))) break;
      hipLaunchKernelGGL(k_modulus, dim3(blocks(N)), dim3(256), 0, st, N, (cplx<float>*)z);
      if ((rc = cdx_apply_norm(ctx, z, 1.0 / (double)N)) || (rc = cdx_fft(ctx, z, z, CDX_INVERSE)) ||
          (rc = cdx_convert_fo(ctx, z, nullptr, 1)))
        break;
    } else {  // shiftWave
      if ((rc = cdx_fft(ctx, z, z, CDX_FORWARD)) || (rc = cdx_convert_fo(ctx, z, nullptr, 1))) break;
      hipLaunchKernelGGL(k_shift_phase, dim3(blocks(N)), dim3(256), 0, st, n, n, (cplx<float>*)z, shx, shy);
      if ((rc = cdx_convert_fo(ctx, z, nullptr, 1)) || (rc = cdx_apply_norm(ctx, z, 1.0 / (double)N)) ||
          (rc = cdx_fft(ctx, z, z, CDX_INVERSE)))
        break;
    }
    hipLaunchKernelGGL(k_merge_real, dim3(blocks((int64_t)fin * fin)), dim3(256), 0, st, (const cplx<float>*)z, n,
                       nmerge, tmp, fin, fin);
    if (hipStreamSynchronize(st) || hipMemcpy(out, tmp, (size_t)fin * fin * 4, hipMemcpyDeviceToHost)) {
      rc = pfail(CDX_ERR_HIP, "cdx_preprocess_pattern: %s", hipGetErrorString(hipGetLastError()));
      break;
    }
  } while (0);
  cleanup();
  return rc;
}
