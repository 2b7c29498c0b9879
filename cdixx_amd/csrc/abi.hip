// abi.hip -- C ABI of the CDI hot path (include/cdx_abi.h) on top of kernels.hip.
#include "../../include/cdx_abi.h"
#include "errors.hpp"
#include "kernels.hpp"
#include <algorithm>
#include <hip/hip_runtime.h>
#include <math.h>
#include <mutex>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>

using namespace cdx;

static thread_local std::string g_err;
static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
#define HIPCHK(expr)                                                                          \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) return fail(CDX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

extern "C" const char* cdx_last_error(void) { return g_err.c_str(); }
namespace cdx {
void set_last_error(const char* msg) { g_err = msg ? msg : ""; }
}
extern "C" const char* cdx_version(void) { return "cdixx_amd 0.1 (gfx950)"; }
extern "C" int cdx_size_supported(int n) { return line_supported(n); }

// scal[] layout (doubles on the device)
enum {
  S_THR = 0,     // support threshold (cudaVars.threshold)
  S_AVG = 1,     // shrink-wrap avg from the first call
  S_LASTMAX = 2,
  S_MAXBITS = 4, // reduce slot: max as bits
  S_SUM = 5,     // reduce slot: sum
  S_ACC = 6,     // [6],[7] generic accumulators
  S_MMA = 8,     // [8],[9] min/max keys
  S_MMP = 10,    // [10],[11]
  S_TOT = 12,
  S_PHI0 = 13,   // PRTF phi0 (cdi.cc:360)
  S_THETA = 14,  // PRTF phase of the current step
  S_COUNT = 16
};

struct cdx_ctx {
  int rows, cols, prec;
  size_t rsz, csz;  // bytes per real / complex element
  int64_t n;
  hipStream_t st;
  bool own_stream;
  void* tw_row;  // twiddles for line length cols
  void* tw_col;  // twiddles for line length rows
  void* T1;
  void* T2;
  void* Ipk;
  void* tmpr;
  double* scal;
  double* partials;  // kMaxPartials per-workgroup sums
  double* ring;
  int ring_cap;
  unsigned* smask = nullptr;  // support bits of the update pass (launch_support_mask)
  int rcolor;
  double scale, beta, relax;
  int col_lines = 0, row_lines = 0;  // lines per workgroup of the passes (0: default)
  int timing_warm = 0;  // cdx_time_iterations: untimed iterations before the timed ones (at least)
  // PRTF of the last cdx_phase_retrieve (allocated on first use)
  void* prtf_map = nullptr;   // summed unit phasors, complex N
  void* gamma0 = nullptr;     // reference object P_S(g'), complex N
  double* prtf_dev = nullptr; // [0, rows/2): prtf sums; [1024, 1024 + rows/2): weights
  std::vector<double> prtf_h, weights_h;
  int prtf_count = 0;
};

// twiddle table of a line length n: tw[m] = W_n^m = exp(-2 pi i m/n), m < n,
// evaluated in double with exact integer reduction of the angle.  An FP32
// table (n float pairs) is followed by the same n values as double pairs,
// from which the FP32 passes form the powers of a twiddle before rounding.
int make_twiddles(int n, int prec, void** dptr) {
  std::vector<double> re((size_t)n), im((size_t)n);
  for (int m = 0; m < n; m++) {
    // reduce to the first octant so cos/sin see |angle| <= pi/4
    const long e8 = 8L * m;
    const int quad = (int)((e8 / n + 1) / 2);
    const double r = -2.0 * M_PI * (double)(4L * m - (long)quad * n) / (4.0 * n);
    const double c = cos(r), s = sin(r);
    // multiply (c + i s) by (-i)^quad
    switch (quad & 3) {
      case 0: re[m] = c; im[m] = s; break;
      case 1: re[m] = s; im[m] = -c; break;
      case 2: re[m] = -c; im[m] = -s; break;
      default: re[m] = -s; im[m] = c; break;
    }
  }
  const size_t dbl_off = prec == CDX_F64 ? 0 : 8 * (size_t)n;
  std::vector<unsigned char> host(dbl_off + 16 * (size_t)n);
  for (size_t i = 0; i < (size_t)n; i++) {
    const double v[2] = {re[i], im[i]};
    memcpy(&host[dbl_off + i * 16], v, 16);
    if (prec != CDX_F64) {
      const float f[2] = {(float)re[i], (float)im[i]};
      memcpy(&host[i * 8], f, 8);
    }
  }
  HIPCHK(hipMalloc(dptr, host.size()));
  HIPCHK(hipMemcpy(*dptr, host.data(), host.size(), hipMemcpyHostToDevice));
  return CDX_OK;
}

extern "C" int cdx_malloc(void** ptr, size_t bytes) {
  if (!ptr) return fail(CDX_ERR_ARG, "cdx_malloc: null out pointer");
  HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
  return CDX_OK;
}
extern "C" int cdx_free(void* ptr) {
  if (ptr) HIPCHK(hipFree(ptr));
  return CDX_OK;
}
extern "C" int cdx_memcpy_h2d(void* d, const void* s, size_t b) {
  HIPCHK(hipMemcpy(d, s, b, hipMemcpyHostToDevice));
  return CDX_OK;
}
extern "C" int cdx_memcpy_d2h(void* d, const void* s, size_t b) {
  HIPCHK(hipMemcpy(d, s, b, hipMemcpyDeviceToHost));
  return CDX_OK;
}
extern "C" int cdx_memcpy_d2d(void* d, const void* s, size_t b) {
  HIPCHK(hipMemcpy(d, s, b, hipMemcpyDeviceToDevice));
  return CDX_OK;
}
extern "C" int cdx_memset0(void* p, size_t b) {
  HIPCHK(hipMemset(p, 0, b));
  return CDX_OK;
}
extern "C" int cdx_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return CDX_OK;
}

extern "C" int cdx_create(cdx_ctx** out, int rows, int cols, int precision, void* stream) {
  if (!out) return fail(CDX_ERR_ARG, "cdx_create: null out pointer");
  *out = nullptr;
  if (!line_supported(rows) || !line_supported(cols))
    return fail(CDX_ERR_ARG, "cdx_create: unsupported size %dx%d (lines of 16..2048, 2^k or 3*2^k)", rows, cols);
  if (precision != CDX_F32 && precision != CDX_F64) return fail(CDX_ERR_ARG, "cdx_create: bad precision %d", precision);
  cdx_ctx* c = new cdx_ctx();
  c->rows = rows;
  c->cols = cols;
  c->prec = precision;
  c->rsz = precision == CDX_F64 ? 8 : 4;
  c->csz = 2 * c->rsz;
  c->n = (int64_t)rows * cols;
  c->rcolor = 65535;
  c->scale = 1;
  c->beta = 1;
  c->relax = 1.9;
  int rc;
  if (stream) {
    c->st = (hipStream_t)stream;
    c->own_stream = false;
  } else {
    // a blocking stream: ordered with null-stream copies, like the reference's
    // single legacy stream (default priority)
    if (hipStreamCreate(&c->st) != hipSuccess) {
      delete c;
      return fail(CDX_ERR_HIP, "cdx_create: hipStreamCreate failed");
    }
    c->own_stream = true;
  }
  if ((rc = make_twiddles(cols, precision, &c->tw_row)) || (rc = make_twiddles(rows, precision, &c->tw_col))) {
    cdx_destroy(c);
    return rc;
  }
  c->ring_cap = 8192;
  if (hipMalloc(&c->T1, c->n * c->csz) != hipSuccess || hipMalloc(&c->T2, c->n * c->csz) != hipSuccess ||
      hipMalloc(&c->Ipk, c->n * c->rsz) != hipSuccess || hipMalloc(&c->tmpr, c->n * c->rsz) != hipSuccess ||
      hipMalloc((void**)&c->scal, S_COUNT * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->partials, kMaxPartials * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->ring, 2 * kRfSlots * (size_t)c->ring_cap * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&c->smask, support_mask_bytes(rows, cols, c->csz)) != hipSuccess) {
    cdx_destroy(c);
    return fail(CDX_ERR_HIP, "cdx_create: out of device memory for %dx%d", rows, cols);
  }
  // raise the LDS caps of every pass this context can launch now, outside
  // any stream capture
  hipError_t pe = precision == CDX_F64 ? prepare_col<double>(rows) : prepare_col<float>(rows);
  if (pe == hipSuccess) pe = precision == CDX_F64 ? prepare_row<double>(cols) : prepare_row<float>(cols);
  if (pe != hipSuccess) {
    cdx_destroy(c);
    return fail(CDX_ERR_HIP, "cdx_create: raising the LDS limits of the passes failed: %s", hipGetErrorString(pe));
  }
  double init[S_COUNT] = {0};
  init[S_THR] = 0.5;  // init_cuda_image, src/gpu/cdi.cu:20
  if (hipMemcpy(c->scal, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess) {
    cdx_destroy(c);
    return fail(CDX_ERR_HIP, "cdx_create: scalar upload failed");
  }
  *out = c;
  return CDX_OK;
}

extern "C" int cdx_destroy(cdx_ctx* c) {
  if (!c) return CDX_OK;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (void* p : {c->tw_row, c->tw_col, c->T1, c->T2, c->Ipk, c->tmpr, (void*)c->scal, (void*)c->partials,
                  (void*)c->ring, (void*)c->smask, c->prtf_map, c->gamma0, (void*)c->prtf_dev})
    (void)hipFree(p);
  if (c->own_stream && c->st) (void)hipStreamDestroy(c->st);
  delete c;
  return CDX_OK;
}

#define CTX_CHECK(c)                                          \
  do {                                                        \
    if (!(c)) return fail(CDX_ERR_ARG, "%s: null context", __func__); \
  } while (0)

extern "C" int cdx_context_info(cdx_ctx* c, int* rows, int* cols, int* precision) {
  CTX_CHECK(c);
  if (rows) *rows = c->rows;
  if (cols) *cols = c->cols;
  if (precision) *precision = c->prec;
  return CDX_OK;
}

extern "C" int cdx_synchronize(cdx_ctx* c) {
  CTX_CHECK(c);
  HIPCHK(hipStreamSynchronize(c->st));
  return CDX_OK;
}
extern "C" void* cdx_stream(cdx_ctx* c) { return c ? (void*)c->st : nullptr; }

extern "C" int cdx_init_vars(cdx_ctx* c, int rcolor, double scale) {
  CTX_CHECK(c);
  if (rcolor != 0) c->rcolor = rcolor;
  if (scale == scale) c->scale = scale;
  return CDX_OK;
}
extern "C" int cdx_set_threshold(cdx_ctx* c, double thr) {
  CTX_CHECK(c);
  // stored at the precision of the support comparison
  double v = c->prec == CDX_F64 ? thr : (double)(float)thr;
  HIPCHK(hipMemcpyAsync(c->scal + S_THR, &v, sizeof v, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return CDX_OK;
}
extern "C" int cdx_get_threshold(cdx_ctx* c, double* thr) {
  CTX_CHECK(c);
  if (!thr) return fail(CDX_ERR_ARG, "cdx_get_threshold: null out");
  HIPCHK(hipMemcpyAsync(thr, c->scal + S_THR, sizeof(double), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return CDX_OK;
}
extern "C" int cdx_set_beta(cdx_ctx* c, double b) {
  CTX_CHECK(c);
  c->beta = b;
  return CDX_OK;
}
extern "C" int cdx_set_relax(cdx_ctx* c, double r) {
  CTX_CHECK(c);
  c->relax = r;
  return CDX_OK;
}

static ReduceSlots slots(cdx_ctx* c) {
  return ReduceSlots{(void*)(c->scal + S_MAXBITS), c->partials, (int*)(c->scal + S_SUM)};
}

// ---------------------------------------------------------------------------
// precision dispatch
// ---------------------------------------------------------------------------
template <class T>
static int fft_t(cdx_ctx* c, const void* in, void* out, int dir) {
  ProjArgs<T> pa{};
  UpdArgs<T> ua{};
  ReduceSlots rs{};
  if (dir == CDX_FORWARD) {
    HIPCHK(launch_row<T>(ROW_FWD, c->rows, c->cols, in, out, c->tw_row, nullptr, nullptr, ua, (T)1, nullptr, rs, c->st));
    HIPCHK(launch_col<T>(COL_FWD, c->rows, c->cols, out, out, c->tw_col, nullptr, pa, 0, c->st));
  } else {
    HIPCHK(launch_col<T>(COL_INV, c->rows, c->cols, in, out, c->tw_col, nullptr, pa, 0, c->st));
    HIPCHK(launch_row<T>(ROW_INV, c->rows, c->cols, out, out, c->tw_row, nullptr, nullptr, ua, (T)1, nullptr, rs, c->st));
  }
  return CDX_OK;
}
#define DISPATCH(c, fn, ...) ((c)->prec == CDX_F64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

extern "C" int cdx_fft(cdx_ctx* c, const void* in, void* out, int dir) {
  CTX_CHECK(c);
  if (!in || !out) return fail(CDX_ERR_ARG, "cdx_fft: null buffer");
  if (dir != CDX_FORWARD && dir != CDX_INVERSE) return fail(CDX_ERR_ARG, "cdx_fft: bad direction %d", dir);
  return DISPATCH(c, fft_t, c, in, out, dir);
}

extern "C" int cdx_apply_mod_abs(cdx_ctx* c, void* F, const void* I) {
  CTX_CHECK(c);
  if (!F || !I) return fail(CDX_ERR_ARG, "cdx_apply_mod_abs: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_mod_abs<double>(c->n, F, (const double*)I, c->st));
  else HIPCHK(launch_mod_abs<float>(c->n, F, (const float*)I, c->st));
  return CDX_OK;
}
extern "C" int cdx_apply_mod(cdx_ctx* c, void* F, const void* I, const void* bs, int noise, double bsnorm) {
  CTX_CHECK(c);
  if (!F || !I) return fail(CDX_ERR_ARG, "cdx_apply_mod: null buffer");
  if (c->prec == CDX_F64)
    HIPCHK(launch_mod_noise<double>(c->n, F, (const double*)I, (const double*)bs, noise, c->scale, c->rcolor, bsnorm, c->st));
  else
    HIPCHK(launch_mod_noise<float>(c->n, F, (const float*)I, (const float*)bs, noise, (float)c->scale,
                                   (float)c->rcolor, (float)bsnorm, c->st));
  return CDX_OK;
}
template <class T>
static UpdArgs<T> upd_args(cdx_ctx* c, int algo, double beta, double relax) {
  UpdArgs<T> ua;
  ua.algo = algo;
  ua.beta = (T)beta;
  ua.relax = (T)relax;
  ua.invN = (T)(1.0 / (double)c->n);
  ua.scal = c->scal;
  ua.gamma0 = nullptr;
  ua.prtf = 0;
  ua.smask = c->smask;
  return ua;
}
extern "C" int cdx_apply_support(cdx_ctx* c, void* g, const void* gp, int algo, const void* spt) {
  CTX_CHECK(c);
  if (!g || !gp || !spt) return fail(CDX_ERR_ARG, "cdx_apply_support: null buffer");
  if (c->prec == CDX_F64)
    HIPCHK(launch_support<double>(c->n, g, gp, (const double*)spt, upd_args<double>(c, algo, c->beta, c->relax), c->st));
  else
    HIPCHK(launch_support<float>(c->n, g, gp, (const float*)spt, upd_args<float>(c, algo, c->beta, c->relax), c->st));
  return CDX_OK;
}
extern "C" int cdx_apply_norm(cdx_ctx* c, void* d, double f) {
  CTX_CHECK(c);
  if (!d) return fail(CDX_ERR_ARG, "cdx_apply_norm: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_norm_c<double>(c->n, d, f, c->st));
  else HIPCHK(launch_norm_c<float>(c->n, d, (float)f, c->st));
  return CDX_OK;
}
extern "C" int cdx_apply_norm_real(cdx_ctx* c, void* d, double f) {
  CTX_CHECK(c);
  if (!d) return fail(CDX_ERR_ARG, "cdx_apply_norm_real: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_norm_r<double>(c->n, (double*)d, f, c->st));
  else HIPCHK(launch_norm_r<float>(c->n, (float*)d, (float)f, c->st));
  return CDX_OK;
}
extern "C" int cdx_get_mod2(cdx_ctx* c, void* out, const void* in, double norm) {
  CTX_CHECK(c);
  if (!out || !in) return fail(CDX_ERR_ARG, "cdx_get_mod2: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_mod2<double>(c->n, (double*)out, in, norm, c->st));
  else HIPCHK(launch_mod2<float>(c->n, (float*)out, in, (float)norm, c->st));
  return CDX_OK;
}
extern "C" int cdx_gauss_mult_freq(cdx_ctx* c, const void* in, void* out, double sigma) {
  CTX_CHECK(c);
  if (!in || !out) return fail(CDX_ERR_ARG, "cdx_gauss_mult_freq: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_gauss_mult_freq<double>(c->rows, c->cols, in, out, sigma, c->st));
  else HIPCHK(launch_gauss_mult_freq<float>(c->rows, c->cols, in, out, (float)sigma, c->st));
  return CDX_OK;
}
extern "C" int cdx_convert_fo(cdx_ctx* c, void* d, void* out, int is_complex) {
  CTX_CHECK(c);
  if (!d) return fail(CDX_ERR_ARG, "cdx_convert_fo: null buffer");
  if (!out) out = d;
  if (c->prec == CDX_F64) HIPCHK(launch_convert_fo<double>(c->rows, c->cols, d, out, is_complex, c->st));
  else HIPCHK(launch_convert_fo<float>(c->rows, c->cols, d, out, is_complex, c->st));
  return CDX_OK;
}
extern "C" int cdx_create_support(cdx_ctx* c, void* spt, double os) {
  CTX_CHECK(c);
  if (!spt) return fail(CDX_ERR_ARG, "cdx_create_support: null buffer");
  if (!(os > 0)) return fail(CDX_ERR_ARG, "cdx_create_support: oversampling_spt must be > 0");
  // rect re; re.startx = (oversampling_spt-1)/2*row/oversampling_spt-1 (cdi.cc:242-245)
  int sx = (int)((os - 1) / 2 * c->rows / os - 1);
  int sy = (int)((os - 1) / 2 * c->cols / os - 1);
  int ex = c->rows - sx, ey = c->cols - sy;
  if (c->prec == CDX_F64) HIPCHK(launch_rect_mask<double>(c->rows, c->cols, sx, sy, ex, ey, (double*)spt, c->st));
  else HIPCHK(launch_rect_mask<float>(c->rows, c->cols, sx, sy, ex, ey, (float*)spt, c->st));
  return CDX_OK;
}
extern "C" int cdx_create_beamstop(cdx_ctx* c, void* bs, double r) {
  CTX_CHECK(c);
  if (!bs) return fail(CDX_ERR_ARG, "cdx_create_beamstop: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_circle_mask<double>(c->rows, c->cols, (float)r, 1, (double*)bs, c->st));
  else HIPCHK(launch_circle_mask<float>(c->rows, c->cols, (float)r, 1, (float*)bs, c->st));
  return CDX_OK;
}
extern "C" int cdx_create_circle(cdx_ctx* c, void* m, double r) {
  CTX_CHECK(c);
  if (!m) return fail(CDX_ERR_ARG, "cdx_create_circle: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_circle_mask<double>(c->rows, c->cols, (float)r, 0, (double*)m, c->st));
  else HIPCHK(launch_circle_mask<float>(c->rows, c->cols, (float)r, 0, (float*)m, c->st));
  return CDX_OK;
}

static int read_reduce(cdx_ctx* c, double* mx, double* sm) {
  double h[S_COUNT];
  HIPCHK(hipMemcpyAsync(h, c->scal, sizeof h, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (mx) {
    if (c->prec == CDX_F64) {
      double v;
      memcpy(&v, &h[S_MAXBITS], 8);
      *mx = v;
    } else {
      float v;
      memcpy(&v, &h[S_MAXBITS], 4);
      *mx = v;
    }
  }
  if (sm) {
    int np;
    memcpy(&np, &h[S_SUM], sizeof np);
    if (np < 0 || np > kMaxPartials) return fail(CDX_ERR_HIP, "reduction: bad partial count %d", np);
    std::vector<double> part(np);
    HIPCHK(hipMemcpy(part.data(), c->partials, np * sizeof(double), hipMemcpyDeviceToHost));
    double s = 0;
    for (double v : part) s += v;
    *sm = s;
  }
  // clear the max slot
  double z = 0;
  HIPCHK(hipMemcpyAsync(c->scal + S_MAXBITS, &z, sizeof z, hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return CDX_OK;
}
extern "C" int cdx_find_max(cdx_ctx* c, const void* x, int64_t n, double* out) {
  CTX_CHECK(c);
  if (!x || !out) return fail(CDX_ERR_ARG, "cdx_find_max: null buffer");
  if (n <= 0) n = c->n;
  if (c->prec == CDX_F64) HIPCHK(launch_reduce<double>(n, (const double*)x, slots(c), c->st));
  else HIPCHK(launch_reduce<float>(n, (const float*)x, slots(c), c->st));
  return read_reduce(c, out, nullptr);
}
extern "C" int cdx_find_sum(cdx_ctx* c, const void* x, int64_t n, double* out) {
  CTX_CHECK(c);
  if (!x || !out) return fail(CDX_ERR_ARG, "cdx_find_sum: null buffer");
  if (n <= 0) n = c->n;
  if (c->prec == CDX_F64) HIPCHK(launch_reduce<double>(n, (const double*)x, slots(c), c->st));
  else HIPCHK(launch_reduce<float>(n, (const float*)x, slots(c), c->st));
  return read_reduce(c, nullptr, out);
}

extern "C" int cdx_gauss_conv(cdx_ctx* c, const void* in, void* out, double sigma) {
  CTX_CHECK(c);
  if (!in || !out) return fail(CDX_ERR_ARG, "cdx_gauss_conv: null buffer");
  if (!(sigma > 0)) return fail(CDX_ERR_ARG, "cdx_gauss_conv: sigma must be > 0");
  if (in == out) return fail(CDX_ERR_ARG, "cdx_gauss_conv: in and out must differ (as applyConvolution's)");
  if (c->prec == CDX_F64)
    HIPCHK(launch_blur_real<double>(c->rows, c->cols, in, 0, (double*)c->tmpr, (double*)out, sigma, slots(c), c->st));
  else
    HIPCHK(launch_blur_real<float>(c->rows, c->cols, in, 0, (float*)c->tmpr, (float*)out, (float)sigma, slots(c), c->st));
  return read_reduce(c, nullptr, nullptr);  // clears the max/sum slots the blur filled
}

extern "C" int cdx_random_start(cdx_ctx* c, const void* I, const void* bs, uint64_t seed, int mode, void* F) {
  CTX_CHECK(c);
  if (!I || !F) return fail(CDX_ERR_ARG, "cdx_random_start: null buffer");
  double thr;
  int rc = cdx_get_threshold(c, &thr);
  if (rc) return rc;
  if (c->prec == CDX_F64) HIPCHK(launch_start<double>(c->n, (const double*)I, (const double*)bs, thr, seed, mode, F, c->st));
  else HIPCHK(launch_start<float>(c->n, (const float*)I, (const float*)bs, thr, seed, mode, F, c->st));
  return CDX_OK;
}

static int read_acc(cdx_ctx* c, double* a, double* b) {
  double h[2];
  HIPCHK(hipMemcpyAsync(h, c->scal + S_ACC, sizeof h, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (a) *a = h[0];
  if (b) *b = h[1];
  return CDX_OK;
}
extern "C" int cdx_residual(cdx_ctx* c, const void* F, const void* I, const void* bs, double* out) {
  CTX_CHECK(c);
  if (!F || !I || !out) return fail(CDX_ERR_ARG, "cdx_residual: null buffer");
  HIPCHK(launch_fill_f64(2, c->scal + S_ACC, 0.0, c->st));
  if (c->prec == CDX_F64) HIPCHK(launch_residual<double>(c->n, F, (const double*)I, (const double*)bs, c->scal + S_ACC, c->st));
  else HIPCHK(launch_residual<float>(c->n, F, (const float*)I, (const float*)bs, c->scal + S_ACC, c->st));
  return read_acc(c, out, nullptr);
}
extern "C" int cdx_rfactor(cdx_ctx* c, const void* F, const void* I, const void* bs, double* out) {
  CTX_CHECK(c);
  if (!F || !I || !out) return fail(CDX_ERR_ARG, "cdx_rfactor: null buffer");
  HIPCHK(launch_fill_f64(2, c->scal + S_ACC, 0.0, c->st));
  if (c->prec == CDX_F64) {
    HIPCHK(launch_pack_intensity<double>(c->n, (const double*)I, (const double*)bs, (double*)c->Ipk, 0, c->st));
    HIPCHK(launch_rfactor<double>(c->n, F, (const double*)c->Ipk, c->scal + S_ACC, c->st));
  } else {
    HIPCHK(launch_pack_intensity<float>(c->n, (const float*)I, (const float*)bs, (float*)c->Ipk, 0, c->st));
    HIPCHK(launch_rfactor<float>(c->n, F, (const float*)c->Ipk, c->scal + S_ACC, c->st));
  }
  double a, b;
  int rc = read_acc(c, &a, &b);
  if (rc) return rc;
  *out = b > 0 ? a / b : 0;
  return CDX_OK;
}

extern "C" void cdx_params_default(cdx_params* p) {
  if (!p) return;
  memset(p, 0, sizeof *p);
  p->rcolor = 65535;
  p->exposure = 1;
  p->beta = 1;
  p->relax = 1.9;
  p->shrinkThreshold = 0;
  p->shrinkMode = 0;
  p->sigma0 = 3;
  p->sigmaMin = 2;
  p->sigmaDecay = 0.99;
  p->prtfIter = -1;
}

// ---------------------------------------------------------------------------
// the loop
// ---------------------------------------------------------------------------
namespace {
template <class T>
struct Loop {
  cdx_ctx* c;
  const cdx_params* p;
  ProjArgs<T> pa;
  bool f_pending;  // F holds the current Fourier estimate; T1 not yet valid
  void* F;
  void* g;
  T* spt;

  int setup(const void* I, const void* bs) {
    const bool pack_bs = p->useBS && bs && (p->simCCDbit || p->freeBeamstop);
    // packed in the column-tile layout the single-frame column pass reads
    HIPCHK(launch_pack_intensity_tiled<T>(c->rows, c->cols, (const T*)I, pack_bs ? (const T*)bs : nullptr,
                                          (T*)c->Ipk, !p->simCCDbit, c->st));
    pa.mode = p->simCCDbit ? PROJ_MOD_NOISE : (pack_bs ? PROJ_MOD_ABS_BS : PROJ_MOD_ABS);
    pa.noiseLevel = p->noiseLevel;
    pa.scale = (T)(1.0 / p->exposure);
    pa.rcolor = (T)p->rcolor;
    pa.rf = nullptr;
    pa.rfs = 2 * c->ring_cap;
    // the update pass's support bits from the caller's support and threshold
    HIPCHK(launch_support_mask<T>(c->rows, c->cols, spt, c->scal, c->smask, c->st));
    return CDX_OK;
  }
  int iterate(int algo, double* rf) {
    pa.rf = rf;
    if (f_pending)
      HIPCHK(launch_col<T>(COL_PROJ_INV, c->rows, c->cols, F, c->T2, c->tw_col, (const T*)c->Ipk, pa, 0, c->st,
                           c->col_lines));
    else
      HIPCHK(launch_col<T>(COL_FWD_PROJ_INV, c->rows, c->cols, c->T1, c->T2, c->tw_col, (const T*)c->Ipk, pa, 0,
                           c->st, c->col_lines));
    UpdArgs<T> ua = upd_args<T>(c, algo, p->beta, p->relax);
    HIPCHK(launch_row<T>(ROW_INV_UPDATE_FWD, c->rows, c->cols, c->T2, c->T1, c->tw_row, g, spt, ua, (T)1, nullptr,
                         ReduceSlots{}, c->st, c->row_lines));
    f_pending = false;
    return CDX_OK;
  }
  // one iteration with the PRTF accumulation of cdi.cc:355-380: the column
  // pass also stores the projected F (patternWave), the row pass forms the
  // phase-alignment sums, then map += phasor(F) e^{i theta}
  int iterate_prtf(int algo, double* rf, bool first) {
    pa.rf = rf;
    pa.fsave = F;
    hipError_t e = f_pending ? launch_col<T>(COL_PROJ_INV, c->rows, c->cols, F, c->T2, c->tw_col, (const T*)c->Ipk,
                                             pa, 0, c->st, c->col_lines)
                             : launch_col<T>(COL_FWD_PROJ_INV, c->rows, c->cols, c->T1, c->T2, c->tw_col,
                                             (const T*)c->Ipk, pa, 0, c->st, c->col_lines);
    pa.fsave = nullptr;
    HIPCHK(e);
    UpdArgs<T> ua = upd_args<T>(c, algo, p->beta, p->relax);
    ua.gamma0 = c->gamma0;
    ua.prtf = first ? 1 : 2;
    HIPCHK(launch_row<T>(ROW_INV_UPDATE_FWD_PRTF, c->rows, c->cols, c->T2, c->T1, c->tw_row, g, spt, ua, (T)1,
                         nullptr, slots(c), c->st, c->row_lines));
    HIPCHK(launch_prtf_phase(c->scal + S_PHI0, c->scal + S_THETA, slots(c), first, c->st));
    HIPCHK(launch_prtf_add<T>(c->n, c->prtf_map, F, c->scal + S_THETA, c->st));
    if (first) HIPCHK(launch_radial<T>(F, c->rows, c->cols, 0, (T)1, c->prtf_dev + 1024, c->st));
    f_pending = false;
    return CDX_OK;
  }
  int shrink_wrap(double sigma) {
    ProjArgs<T> none{};
    if (p->shrinkMode == 0) {
      // applyGaussMult(patternWave, ., Real(row>>2)/gaussianSigma, 1) (cdi.cc:327)
      T sf = (T)(c->rows >> 2) / (T)sigma;
      if (f_pending)
        HIPCHK(launch_col<T>(COL_GAUSS_INV, c->rows, c->cols, F, c->T2, c->tw_col, nullptr, none, (double)sf, c->st,
                             c->col_lines));
      else
        HIPCHK(launch_col<T>(COL_FWD_GAUSS_INV, c->rows, c->cols, c->T1, c->T2, c->tw_col, nullptr, none, (double)sf,
                             c->st, c->col_lines));
      HIPCHK(launch_row<T>(ROW_INV_MOD2, c->rows, c->cols, c->T2, nullptr, c->tw_row, nullptr, nullptr, UpdArgs<T>{},
                           (T)1, spt, slots(c), c->st, c->row_lines));
      HIPCHK(launch_threshold<T>(c->scal, slots(c), p->shrinkThreshold, c->rows, c->cols, 0, c->st));
    } else {
      HIPCHK(launch_blur_real<T>(c->rows, c->cols, g, 1, (T*)c->tmpr, spt, (T)sigma, slots(c), c->st));
      HIPCHK(launch_threshold<T>(c->scal, slots(c), p->shrinkThreshold, c->rows, c->cols, 1, c->st));
    }
    // new support and threshold: new support bits
    HIPCHK(launch_support_mask<T>(c->rows, c->cols, spt, c->scal, c->smask, c->st));
    return CDX_OK;
  }
  int finish() {
    if (!f_pending) {
      ProjArgs<T> none{};
      HIPCHK(launch_col<T>(COL_FWD_T1, c->rows, c->cols, c->T1, F, c->tw_col, nullptr, none, 0, c->st));
      f_pending = true;
    }
    return CDX_OK;
  }
};
}  // namespace

extern "C" int cdx_expand_schedule(const char* algorithm, int* out, int cap);

template <class T>
static int phase_retrieve_t(cdx_ctx* c, const char* algorithm, const cdx_params* p, void* F, void* g, const void* I,
                            const void* bs, void* spt, double* rtrace, int rcap, int* n_iter) {
  int nt = cdx_expand_schedule(algorithm, nullptr, 0);
  if (nt < 0) return fail(CDX_ERR_SCHEDULE, "cdx_phase_retrieve: bad algorithm string \"%s\"", algorithm);
  std::vector<int> toks(nt);
  cdx_expand_schedule(algorithm, toks.data(), nt);
  for (int t : toks)
    if (t == CDX_TV) return fail(CDX_ERR_UNSUPPORTED, "cdx_phase_retrieve: TV steps are outside this path");
  Loop<T> L{c, p, {}, true, F, g, (T*)spt};
  int rc = L.setup(I, bs);
  if (rc) return rc;
  const T rootN = (T)sqrt((double)c->n);
  HIPCHK(launch_norm_c<T>(c->n, g, (T)1 / rootN, c->st));  // cdi.cc:313
  // the shrink-wrap average is taken on the first shrink-wrap of each call
  // (a local of CDI::phaseRetrieve, cdi.cc:315,331)
  HIPCHK(launch_fill_f64(1, c->scal + S_AVG, 0.0, c->st));
  const int nrec = p->recordR ? (rcap < c->ring_cap ? rcap : c->ring_cap) : 0;
  if (nrec > 0) HIPCHK(launch_fill_f64(2 * kRfSlots * (int64_t)c->ring_cap, c->ring, 0.0, c->st));

  // PRTF buffers (before any capture); the map starts at zero
  const bool want_prtf = p->prtfIter >= 0;
  if (want_prtf) {
    if (!c->prtf_map) HIPCHK(hipMalloc(&c->prtf_map, c->n * c->csz));
    if (!c->gamma0) HIPCHK(hipMalloc(&c->gamma0, c->n * c->csz));
    if (!c->prtf_dev) HIPCHK(hipMalloc((void**)&c->prtf_dev, 2048 * sizeof(double)));
    HIPCHK(launch_fill_f64(c->n * (int64_t)c->csz / 8, (double*)c->prtf_map, 0.0, c->st));
    HIPCHK(launch_fill_f64(2048, c->prtf_dev, 0.0, c->st));
  }

  // the real-space blur's LDS grows with sigma, largest at the start
  if (p->shrinkMode == 1) HIPCHK(prepare_blur<T>(c->rows, (T)p->sigma0));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (p->useGraph) HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  double sigma = p->sigma0;
  int iter = 0, nproj = 0;
  bool prtf_started = false;
  for (int tok : toks) {
    if (tok == CDX_SHRINKWRAP) {
      if ((rc = L.shrink_wrap(sigma))) break;
      if (sigma > p->sigmaMin) sigma *= p->sigmaDecay;
    } else {
      double* rf = nproj < nrec ? c->ring + 2 * (size_t)nproj : nullptr;
      if (want_prtf && iter >= p->prtfIter) {
        if ((rc = L.iterate_prtf(tok, rf, !prtf_started))) break;
        prtf_started = true;
      } else if ((rc = L.iterate(tok, rf))) {
        break;
      }
      nproj++;
    }
    iter++;
  }
  // final PRTF: radial sums of |map| / (iter - prtfIter) (cdi.cc:423-433)
  if (!rc && prtf_started) {
    hipError_t e = launch_radial<T>(c->prtf_map, c->rows, c->cols, 1, (T)(1. / (iter - p->prtfIter)), c->prtf_dev,
                                    c->st);
    if (e != hipSuccess) rc = fail(CDX_ERR_HIP, "launch_radial: %s", hipGetErrorString(e));
  }
  if (!rc) rc = L.finish();
  if (p->useGraph) {
    hipError_t e = hipStreamEndCapture(c->st, &graph);
    if (!rc && e != hipSuccess) rc = fail(CDX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (!rc) {
      e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = fail(CDX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    if (!rc) {
      e = hipGraphLaunch(exec, c->st);
      if (e != hipSuccess) rc = fail(CDX_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
    }
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
  }
  if (rc) return rc;
  HIPCHK(launch_norm_c<T>(c->n, g, rootN, c->st));  // cdi.cc:406
  if (n_iter) *n_iter = iter;
  c->prtf_count = prtf_started ? iter - p->prtfIter : 0;
  c->prtf_h.clear();
  c->weights_h.clear();
  if (prtf_started) {
    c->prtf_h.assign((size_t)c->rows / 2, 0.0);
    c->weights_h.assign((size_t)c->rows / 2, 0.0);
    HIPCHK(hipMemcpyAsync(c->prtf_h.data(), c->prtf_dev, c->prtf_h.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->st));
    HIPCHK(hipMemcpyAsync(c->weights_h.data(), c->prtf_dev + 1024, c->weights_h.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  if (nrec > 0 && rtrace) {
    // the first nrec entries of every slot row
    std::vector<double> h(2 * kRfSlots * (size_t)nrec);
    HIPCHK(hipMemcpy2DAsync(h.data(), 2 * nrec * sizeof(double), c->ring, 2 * c->ring_cap * sizeof(double),
                            2 * nrec * sizeof(double), kRfSlots, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipStreamSynchronize(c->st));
    int m = nproj < nrec ? nproj : nrec;
    for (int i = 0; i < m; i++) {
      double num = 0, den = 0;
      for (int k = 0; k < kRfSlots; k++) {
        num += h[2 * ((size_t)k * nrec + i)];
        den += h[2 * ((size_t)k * nrec + i) + 1];
      }
      rtrace[i] = den > 0 ? num / den : 0;
    }
  }
  return CDX_OK;
}

extern "C" int cdx_phase_retrieve(cdx_ctx* c, const char* algorithm, const cdx_params* p, void* F, void* g,
                                  const void* I, const void* bs, void* spt, double* rtrace, int rcap, int* n_iter) {
  CTX_CHECK(c);
  if (!algorithm || !p || !F || !g || !I || !spt) return fail(CDX_ERR_ARG, "cdx_phase_retrieve: null argument");
  if (!(p->exposure != 0)) return fail(CDX_ERR_ARG, "cdx_phase_retrieve: exposure must be non-zero");
  return DISPATCH(c, phase_retrieve_t, c, algorithm, p, F, g, I, bs, spt, rtrace, rcap, n_iter);
}

extern "C" int cdx_prtf(cdx_ctx* c, double* prtf, double* weights, int cap, int* count) {
  CTX_CHECK(c);
  // rows/2 values; zeros when the last call accumulated nothing
  const int m = c->rows / 2 < cap ? c->rows / 2 : cap;
  const bool have = (int)c->prtf_h.size() >= m;
  for (int i = 0; i < m; i++) {
    if (prtf) prtf[i] = have ? c->prtf_h[i] : 0.0;
    if (weights) weights[i] = have ? c->weights_h[i] : 0.0;
  }
  if (count) *count = c->prtf_count;
  return CDX_OK;
}

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
// Events and graphs of a timing run, released on every exit path.
struct TimingRes {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~TimingRes() {
    drop_graph();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// a list of events released on every exit path
struct EventList {
  std::vector<hipEvent_t> e;
  int create(int k) {
    e.assign((size_t)k, nullptr);
    for (auto& x : e)
      if (hipEventCreate(&x) != hipSuccess) return fail(CDX_ERR_HIP, "hipEventCreate failed");
    return CDX_OK;
  }
  ~EventList() {
    for (auto x : e)
      if (x) (void)hipEventDestroy(x);
  }
};

// Captures `body` (a sequence of launches on st) into res.graph and
// instantiates it.  A capture that fails part-way is still ended, so the
// stream is never left capturing.
template <class Body>
static int capture(hipStream_t st, TimingRes& res, Body body) {
  res.drop_graph();
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = body();
  hipError_t e = hipStreamEndCapture(st, &res.graph);
  if (rc) return rc;
  HIPCHK(e);
  HIPCHK(hipGraphInstantiate(&res.exec, res.graph, nullptr, nullptr, 0));
  return CDX_OK;
}

template <class T>
static int time_t_(cdx_ctx* c, const cdx_params* p, int algo, void* F, void* g, const void* I, const void* bs,
                   const void* spt, int n, int use_graph, double* ms_total, double* col_ms, double* row_ms,
                   int per_launch = 0, int n_pass = 0) {
  Loop<T> L{c, p, {}, true, F, g, (T*)spt};
  int rc = L.setup(I, bs);
  if (rc) return rc;
  TimingRes res;
  HIPCHK(hipEventCreate(&res.e0));
  HIPCHK(hipEventCreate(&res.e1));
  // g as CDI::prepareIter leaves it: the loop carries IFFT(F)/N (cdi.cc:313)
  const T rootN = (T)sqrt((double)c->n);
  HIPCHK(launch_norm_c<T>(c->n, g, (T)1 / rootN, c->st));
  // R-factor sums when p->recordR (accumulated into the first ring entry)
  double* rf = p->recordR ? c->ring : nullptr;
  // first iteration leaves the F-pending state; it is not timed
  if ((rc = L.iterate(algo, rf))) return rc;
  if (use_graph) {
    rc = capture(c->st, res, [&]() {
      int r = CDX_OK;
      for (int i = 0; i < n && !r; i++) r = L.iterate(algo, rf);
      return r;
    });
    if (rc) return rc;
    // warm the graph: once, or until at least timing_warm iterations have run
    const int reps = c->timing_warm > n ? (c->timing_warm + n - 1) / n : 1;
    for (int w = 0; w < reps; w++) HIPCHK(hipGraphLaunch(res.exec, c->st));
  } else {
    for (int w = 0; w < c->timing_warm && !rc; w++) rc = L.iterate(algo, rf);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipEventRecord(res.e0, c->st));
  if (use_graph) {
    HIPCHK(hipGraphLaunch(res.exec, c->st));
  } else {
    for (int i = 0; i < n && !rc; i++) rc = L.iterate(algo, rf);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(res.e1, c->st));
  HIPCHK(hipEventSynchronize(res.e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, res.e0, res.e1));
  if (ms_total) *ms_total = ms;
  res.drop_graph();
  if (col_ms || row_ms) {
    // average launch duration of each pass: 200 launches of that pass
    // alone (whatever the number of timed iterations: the first launches of
    // a graph run slower), back to back (one captured graph when use_graph,
    // as the loop runs), between two events; the median of three such runs.  Every
    // column-pass launch reads the same T1 and every row-pass launch the same
    // T2, so each moves the bytes of its pass in the loop.  The row-pass
    // launches re-apply the update to g, so the state afterwards is no
    // iterate of the loop.
    n = n_pass > 0 ? n_pass : 200;
    UpdArgs<T> ua = upd_args<T>(c, algo, p->beta, p->relax);
    L.pa.rf = rf;
    for (int pass = 0; pass < 2; pass++) {
      double* dst = pass == 0 ? col_ms : row_ms;
      if (!dst) continue;
      auto launch = [&]() -> hipError_t {
        if (pass == 0)
          return launch_col<T>(COL_FWD_PROJ_INV, c->rows, c->cols, c->T1, c->T2, c->tw_col, (const T*)c->Ipk, L.pa,
                               0, c->st, c->col_lines);
        return launch_row<T>(ROW_INV_UPDATE_FWD, c->rows, c->cols, c->T2, c->T1, c->tw_row, g, (const T*)spt, ua,
                             (T)1, nullptr, ReduceSlots{}, c->st, c->row_lines);
      };
      if (per_launch) {
        // every launch between its own pair of events: launches are then
        // serialized as under a kernel-trace profiler (mean over launches)
        EventList ev;
        rc = ev.create(2 * n);
        if (rc) return rc;
        HIPCHK(launch());  // warm
        HIPCHK(hipStreamSynchronize(c->st));
        for (int i = 0; i < n; i++) {
          HIPCHK(hipEventRecord(ev.e[2 * i], c->st));
          HIPCHK(launch());
          HIPCHK(hipEventRecord(ev.e[2 * i + 1], c->st));
        }
        HIPCHK(hipEventSynchronize(ev.e[2 * n - 1]));
        double sum = 0;
        for (int i = 0; i < n; i++) {
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, ev.e[2 * i], ev.e[2 * i + 1]));
          sum += ms;
        }
        *dst = sum / n;
        continue;
      }
      if (use_graph) {
        rc = capture(c->st, res, [&]() {
          for (int i = 0; i < n; i++) HIPCHK(launch());
          return (int)CDX_OK;
        });
        if (rc) return rc;
        HIPCHK(hipGraphLaunch(res.exec, c->st));  // warm the graph once
      }
      // median of three timed runs (single runs vary by up to 20% between
      // processes on one GPU)
      float pms[3];
      for (int k = 0; k < 3; k++) {
        HIPCHK(hipStreamSynchronize(c->st));
        HIPCHK(hipEventRecord(res.e0, c->st));
        if (use_graph) {
          HIPCHK(hipGraphLaunch(res.exec, c->st));
        } else {
          for (int i = 0; i < n; i++) HIPCHK(launch());
        }
        HIPCHK(hipEventRecord(res.e1, c->st));
        HIPCHK(hipEventSynchronize(res.e1));
        HIPCHK(hipEventElapsedTime(&pms[k], res.e0, res.e1));
      }
      std::sort(pms, pms + 3);
      *dst = pms[1] / n;
      res.drop_graph();
    }
  }
  // leave F = FFT(g) and g in the caller's scaling (cdi.cc:406)
  if ((rc = L.finish())) return rc;
  HIPCHK(launch_norm_c<T>(c->n, g, rootN, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return CDX_OK;
}

extern "C" int cdx_time_passes(cdx_ctx* c, const cdx_params* p, int algo, void* F, void* g, const void* I,
                               const void* bs, const void* spt, int n, int mode, double* col_ms, double* row_ms) {
  CTX_CHECK(c);
  if (!p || !F || !g || !I || !spt || n <= 0 || (mode != 0 && mode != 1) || !col_ms || !row_ms)
    return fail(CDX_ERR_ARG, "cdx_time_passes: bad argument");
  return DISPATCH(c, time_t_, c, p, algo, F, g, I, bs, spt, 1, mode == 0, nullptr, col_ms, row_ms, mode, n);
}

extern "C" int cdx_time_iterations(cdx_ctx* c, const cdx_params* p, int algo, void* F, void* g, const void* I,
                                   const void* bs, const void* spt, int n, int use_graph, double* ms_total,
                                   double* col_ms, double* row_ms) {
  CTX_CHECK(c);
  if (!p || !F || !g || !I || !spt || n <= 0) return fail(CDX_ERR_ARG, "cdx_time_iterations: bad argument");
  return DISPATCH(c, time_t_, c, p, algo, F, g, I, bs, spt, n, use_graph, ms_total, col_ms, row_ms);
}

// ---------------------------------------------------------------------------
// synthetic forward model
// ---------------------------------------------------------------------------

static int sim_object_sp(cdx_ctx* c, const ShapeParams& sp, uint64_t seed, double sigma, void* obj) {
  if (!obj) return fail(CDX_ERR_ARG, "cdx_sim_object: null buffer");
  if (c->rows != c->cols) return fail(CDX_ERR_ARG, "cdx_sim_object: square contexts only");
  const int n = c->rows;
  SynthFrames f{};
  f.B = 1;
  f.seed[0] = seed;
  f.sp[0] = sp;
  double* fld = nullptr;  // a, ph and their smoothing temporaries
  HIPCHK(hipMallocAsync((void**)&fld, 4 * (size_t)c->n * sizeof(double), c->st));
  double *a = fld, *ph = fld + c->n, *ta = fld + 2 * c->n, *tph = fld + 3 * c->n;
  // min/max keys of a and ph in scal[S_MMA .. S_MMA + 3]
  static_assert(S_MMP == S_MMA + 2, "min/max keys of both fields are contiguous");
  if (c->prec == CDX_F64) HIPCHK(launch_synth_fields<double>(n, f, sigma, a, ph, ta, tph, c->scal + S_MMA, obj, c->st));
  else HIPCHK(launch_synth_fields<float>(n, f, sigma, a, ph, ta, tph, c->scal + S_MMA, obj, c->st));
  HIPCHK(hipFreeAsync(fld, c->st));
  return CDX_OK;  // ordered on the context's stream like the other context calls
}

// The forward model for the frames of f side by side (launch_synth_fields):
// objects into obj, their 2D transforms into F (the same row and column
// passes as cdx_fft, over B*n rows of n points and n rows of B*n columns:
// each frame's transform is the one cdx_fft makes of it), Poisson
// intensities into I.  Square FP32 contexts; ordered on the context's stream.
int sim_stack(cdx_ctx* c, const SynthFrames& f, double sigma, double* a, double* ph, double* ta, double* tph,
              double* mm, double* tot, void* obj, void* F, float* I) {
  CTX_CHECK(c);
  if (c->prec != CDX_F32 || c->rows != c->cols) return fail(CDX_ERR_ARG, "sim_stack: square FP32 contexts only");
  if (f.B < 1 || f.B > kSynthMax || !synth_stack_supported(c->rows, sigma))
    return fail(CDX_ERR_ARG, "sim_stack: %d frames of %d points, smoothing %g not supported", f.B, c->rows, sigma);
  const int n = c->rows;
  HIPCHK(launch_synth_fields<float>(n, f, sigma, a, ph, ta, tph, mm, obj, c->st));
  ProjArgs<float> pa{};
  UpdArgs<float> ua{};
  ReduceSlots rs{};
  HIPCHK(launch_row<float>(ROW_FWD, f.B * n, n, obj, F, c->tw_row, nullptr, nullptr, ua, 1.f, nullptr, rs, c->st));
  HIPCHK(launch_col<float>(COL_FWD, n, f.B * n, F, F, c->tw_col, nullptr, pa, 0, c->st));
  HIPCHK(launch_synth_intensity<float>(n, n, f, F, tot, 1, I, c->st));
  return CDX_OK;
}

extern "C" int cdx_sim_object(cdx_ctx* c, double radius, uint64_t seed, double sigma, void* obj) {
  CTX_CHECK(c);
  ShapeParams sp{};
  sp.kind = 0;
  sp.a = (float)radius;
  return sim_object_sp(c, sp, seed, sigma, obj);
}

extern "C" int cdx_sim_object_shape(cdx_ctx* c, int kind, const double* prm, int np, uint64_t seed, double sigma,
                                    void* obj) {
  CTX_CHECK(c);
  ShapeParams sp{};
  sp.kind = kind;
  if (!prm) return fail(CDX_ERR_ARG, "cdx_sim_object_shape: null parameters");
  if (kind == 0 && np >= 1) {
    sp.a = (float)prm[0];
  } else if (kind == 1 && np >= 3) {
    sp.a = (float)prm[0];
    sp.b = (float)prm[1];
    sp.rot = (float)prm[2];
  } else if (kind == 2 && np >= 6 && np <= 16 && np % 2 == 0) {
    sp.nv = np / 2;
    for (int i = 0; i < sp.nv; i++) {
      sp.ang[i] = (float)prm[i];
      sp.rad[i] = (float)prm[sp.nv + i];
      if (i && !(sp.ang[i] > sp.ang[i - 1])) return fail(CDX_ERR_ARG, "cdx_sim_object_shape: angles must increase");
    }
    if (!(sp.ang[0] >= 0) || !(sp.ang[sp.nv - 1] < 2 * M_PI))
      return fail(CDX_ERR_ARG, "cdx_sim_object_shape: angles outside [0, 2 pi)");
  } else {
    return fail(CDX_ERR_ARG, "cdx_sim_object_shape: bad shape kind %d / parameter count %d", kind, np);
  }
  return sim_object_sp(c, sp, seed, sigma, obj);
}

extern "C" int cdx_sim_intensity(cdx_ctx* c, const void* obj, double photons, uint64_t seed, int poisson, void* I) {
  CTX_CHECK(c);
  if (!obj || !I) return fail(CDX_ERR_ARG, "cdx_sim_intensity: null buffer");
  int rc = cdx_fft(c, obj, c->T2, CDX_FORWARD);
  if (rc) return rc;
  SynthFrames f{};
  f.B = 1;
  f.seed[0] = seed;
  f.photons[0] = photons;
  if (c->prec == CDX_F64)
    HIPCHK(launch_synth_intensity<double>(c->rows, c->cols, f, c->T2, c->scal + S_TOT, poisson, (double*)I, c->st));
  else
    HIPCHK(launch_synth_intensity<float>(c->rows, c->cols, f, c->T2, c->scal + S_TOT, poisson, (float*)I, c->st));
  return CDX_OK;  // ordered on the context's stream
}

// ---------------------------------------------------------------------------
// setup operations of CDI::prepareIter / readObjectWave / checkAutoCorrelation
// ---------------------------------------------------------------------------
extern "C" int cdx_convert_fo_phase(cdx_ctx* c, void* d, double norm) {
  CTX_CHECK(c);
  if (!d) return fail(CDX_ERR_ARG, "cdx_convert_fo_phase: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_fo_phase<double>(c->rows, c->cols, d, norm, c->st));
  else HIPCHK(launch_fo_phase<float>(c->rows, c->cols, d, (float)norm, c->st));
  return CDX_OK;
}
extern "C" int cdx_mask_multiply(cdx_ctx* c, void* d, const void* m, int invert) {
  CTX_CHECK(c);
  if (!d || !m) return fail(CDX_ERR_ARG, "cdx_mask_multiply: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_mask_mul<double>(c->n, (double*)d, (const double*)m, invert, c->st));
  else HIPCHK(launch_mask_mul<float>(c->n, (float*)d, (const float*)m, invert, c->st));
  return CDX_OK;
}
extern "C" int cdx_poisson_noise_wo(cdx_ctx* c, void* d, double noiseLevel, uint64_t seed) {
  CTX_CHECK(c);
  if (!d) return fail(CDX_ERR_ARG, "cdx_poisson_noise_wo: null buffer");
  if (c->prec == CDX_F64)
    HIPCHK(launch_poisson_wo<double>(c->n, (double*)d, noiseLevel, c->scale, c->rcolor, seed, c->st));
  else
    HIPCHK(launch_poisson_wo<float>(c->n, (float*)d, noiseLevel, (float)c->scale, (float)c->rcolor, seed, c->st));
  return CDX_OK;
}
extern "C" int cdx_extend_to_complex(cdx_ctx* c, const void* a, void* b) {
  CTX_CHECK(c);
  if (!a || !b) return fail(CDX_ERR_ARG, "cdx_extend_to_complex: null buffer");
  if (c->prec == CDX_F64) HIPCHK(launch_extend<double>(c->n, (const double*)a, b, c->st));
  else HIPCHK(launch_extend<float>(c->n, (const float*)a, b, c->st));
  return CDX_OK;
}
extern "C" int cdx_create_wavefront(cdx_ctx* c, const void* I, const void* P, int orow, int ocol, void* out) {
  CTX_CHECK(c);
  if (!out) return fail(CDX_ERR_ARG, "cdx_create_wavefront: null buffer");
  if (orow <= 0 || ocol <= 0 || orow > c->rows || ocol > c->cols)
    return fail(CDX_ERR_ARG, "cdx_create_wavefront: object %dx%d does not fit the %dx%d frame", orow, ocol, c->rows,
                c->cols);
  if (c->prec == CDX_F64)
    HIPCHK(launch_wavefront<double>(c->rows, c->cols, (const double*)I, (const double*)P, orow, ocol, out, c->st));
  else
    HIPCHK(launch_wavefront<float>(c->rows, c->cols, (const float*)I, (const float*)P, orow, ocol, out, c->st));
  return CDX_OK;
}

// ---------------------------------------------------------------------------
// tiling of the passes and its tuner
// ---------------------------------------------------------------------------
extern "C" int cdx_set_timing_warmup(cdx_ctx* c, int iters) {
  CTX_CHECK(c);
  if (iters < 0) return fail(CDX_ERR_ARG, "cdx_set_timing_warmup: negative iteration count %d", iters);
  c->timing_warm = iters;
  return CDX_OK;
}

extern "C" int cdx_set_tiling(cdx_ctx* c, int col_lines, int row_lines) {
  CTX_CHECK(c);
  if (col_lines && !tiling_valid(c->rows, c->csz, col_lines))
    return fail(CDX_ERR_ARG, "cdx_set_tiling: %d columns per workgroup is not valid for %d-point columns", col_lines,
                c->rows);
  if (row_lines && !tiling_valid(c->cols, c->csz, row_lines))
    return fail(CDX_ERR_ARG, "cdx_set_tiling: %d rows per workgroup is not valid for %d-point rows", row_lines,
                c->cols);
  c->col_lines = col_lines;
  c->row_lines = row_lines;
  return CDX_OK;
}
extern "C" int cdx_get_tiling(cdx_ctx* c, int* col_lines, int* row_lines) {
  CTX_CHECK(c);
  if (col_lines) *col_lines = c->col_lines ? c->col_lines : tiling_default(c->rows, c->csz, 1);
  if (row_lines) *row_lines = c->row_lines ? c->row_lines : tiling_default(c->cols, c->csz, 0);
  return CDX_OK;
}

template <class T>
static int autotune_t(cdx_ctx* c, int reps, double* report, int cap) {
  void* g = nullptr;
  void* spt = nullptr;
  HIPCHK(hipMalloc(&g, c->n * c->csz));
  HIPCHK(hipMalloc(&spt, c->n * c->rsz));
  HIPCHK(hipMemsetAsync(g, 0, c->n * c->csz, c->st));
  HIPCHK(hipMemsetAsync(spt, 0, c->n * c->rsz, c->st));
  HIPCHK(hipMemsetAsync(c->T1, 0, c->n * c->csz, c->st));
  HIPCHK(hipMemsetAsync(c->Ipk, 0, c->n * c->rsz, c->st));
  ProjArgs<T> pa{};
  pa.scale = 1;
  pa.rcolor = 65535;
  UpdArgs<T> ua = upd_args<T>(c, CDX_HIO, 1.0, 1.9);
  HIPCHK(launch_support_mask<T>(c->rows, c->cols, (const T*)spt, c->scal, c->smask, c->st));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int nrep = 0;
  double best_col = 1e30, best_row = 1e30;
  int bc = 0, br = 0;
  for (int pass = 0; pass < 2; pass++) {
    const int line = pass == 0 ? c->rows : c->cols;
    for (int lines = 1; lines <= 256; lines *= 2) {
      if (!tiling_valid(line, c->csz, lines)) continue;
      float ms = 0;
      for (int trial = 0; trial < 2; trial++) {  // the first trial warms caches and code
        HIPCHK(hipEventRecord(e0, c->st));
        for (int i = 0; i < reps; i++) {
          if (pass == 0)
            HIPCHK(launch_col<T>(COL_FWD_PROJ_INV, c->rows, c->cols, c->T1, c->T2, c->tw_col, (const T*)c->Ipk, pa,
                                 0, c->st, lines));
          else
            HIPCHK(launch_row<T>(ROW_INV_UPDATE_FWD, c->rows, c->cols, c->T2, c->T1, c->tw_row, g, (const T*)spt, ua,
                                 (T)1, nullptr, ReduceSlots{}, c->st, lines));
        }
        HIPCHK(hipEventRecord(e1, c->st));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      }
      const double us = 1000.0 * ms / reps;
      if (report && nrep + 3 <= cap) {
        report[nrep++] = pass;
        report[nrep++] = lines;
        report[nrep++] = us;
      }
      if (pass == 0 && us < best_col) best_col = us, bc = lines;
      if (pass == 1 && us < best_row) best_row = us, br = lines;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(g);
  hipFree(spt);
  c->col_lines = bc;
  c->row_lines = br;
  return CDX_OK;
}

extern "C" int cdx_autotune(cdx_ctx* c, int reps, double* report, int cap) {
  CTX_CHECK(c);
  if (reps <= 0) reps = 50;
  return DISPATCH(c, autotune_t, c, reps, report, cap);
}

// ---------------------------------------------------------------------------
// cdx_phase_retrieve_static: the loop for callers that cannot allocate
// device memory.  Its workspace (T1, T2, the packed measurement, the blur
// buffer, twiddle tables, scalars, partial sums, support bits) is the file-scope array
// `scratch`, sized for frames of up to kStaticMaxSide points per side in
// FP32; the kernels are those of cdx_phase_retrieve.
// ---------------------------------------------------------------------------
namespace {
constexpr int kStaticMaxSide = 512;
constexpr size_t kStaticMaxN = (size_t)kStaticMaxSide * kStaticMaxSide;
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kOffT1 = 0;
constexpr size_t kOffT2 = kOffT1 + al256(kStaticMaxN * 8);
constexpr size_t kOffIpk = kOffT2 + al256(kStaticMaxN * 8);
constexpr size_t kOffTmp = kOffIpk + al256(kStaticMaxN * 4);
// FP32 twiddle tables: float pairs, then the same values as double pairs
constexpr size_t kOffTwR = kOffTmp + al256(kStaticMaxN * 4);
constexpr size_t kOffTwC = kOffTwR + al256(kStaticMaxSide * 24);
constexpr size_t kOffScal = kOffTwC + al256(kStaticMaxSide * 24);
constexpr size_t kOffPart = kOffScal + al256(S_COUNT * sizeof(double));
constexpr size_t kOffMask = kOffPart + al256(kMaxPartials * sizeof(double));
// support bits: rows x lanes words, at most 512 x 32 for these line lengths
constexpr size_t kMaskBytes = (size_t)kStaticMaxSide * 32 * sizeof(unsigned);
constexpr size_t kScratchBytes = kOffMask + al256(kMaskBytes);  // 6 MiB + 120 KiB
constexpr int kMaxDevices = 64;
}  // namespace

// One copy per GPU, held until the process ends.  Host code never uses this
// name as an address: the per-device pointer from hipGetSymbolAddress is.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// device address of the current device's copy of `scratch`, looked up on
// the first call on that device
static int scratch_address(void** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(CDX_ERR_ARG, "cdx_phase_retrieve_static: device %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CDX_ERR_HIP, "cdx_phase_retrieve_static: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return CDX_OK;
}

// twiddle table tw[m] = W_n^m = exp(-2 pi i m/n) in float pairs followed by
// double pairs, the values make_twiddles uploads (double evaluation after
// exact octant reduction)
__global__ void k_twiddles_f32(int n, float* tw) {
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < n; m += gridDim.x * blockDim.x) {
    const long e8 = 8L * m;
    const int quad = (int)((e8 / n + 1) / 2);
    const double r = -2.0 * M_PI * (double)(4L * m - (long)quad * n) / (4.0 * n);
    const double c = cos(r), s = sin(r);
    double re, im;
    switch (quad & 3) {
      case 0: re = c; im = s; break;
      case 1: re = s; im = -c; break;
      case 2: re = -c; im = -s; break;
      default: re = -s; im = c; break;
    }
    tw[2 * m] = (float)re;
    tw[2 * m + 1] = (float)im;
    double* twd = reinterpret_cast<double*>(tw + 2 * n);
    twd[2 * m] = re;
    twd[2 * m + 1] = im;
  }
}

extern "C" int cdx_phase_retrieve_static(int rows, int cols, const char* algorithm, const cdx_params* p, void* F,
                                         void* g, const void* I, const void* bs, void* spt, double* threshold,
                                         void* stream, int* n_iter) {
  if (!algorithm || !p || !F || !g || !I || !spt || !threshold)
    return fail(CDX_ERR_ARG, "cdx_phase_retrieve_static: null argument");
  if (!line_supported(rows) || !line_supported(cols) || rows > kStaticMaxSide || cols > kStaticMaxSide)
    return fail(CDX_ERR_ARG,
                "cdx_phase_retrieve_static: %dx%d does not fit the static workspace (FFT lengths up to %dx%d)", rows,
                cols, kStaticMaxSide, kStaticMaxSide);
  if (!(p->exposure != 0)) return fail(CDX_ERR_ARG, "cdx_phase_retrieve_static: exposure must be non-zero");
  if (p->prtfIter >= 0) return fail(CDX_ERR_ARG, "cdx_phase_retrieve_static: no PRTF (prtfIter must be < 0)");
  unsigned char* base = nullptr;
  int rc = scratch_address((void**)&base);
  if (rc) return rc;
  cdx_params q = *p;
  q.useGraph = 0;  // graph instantiation allocates
  q.recordR = 0;
  cdx_ctx c{};  // a context on the stack whose buffers are slices of scratch
  c.rows = rows;
  c.cols = cols;
  c.prec = CDX_F32;
  c.rsz = 4;
  c.csz = 8;
  c.n = (int64_t)rows * cols;
  c.st = (hipStream_t)stream;
  c.own_stream = false;
  c.T1 = base + kOffT1;
  c.T2 = base + kOffT2;
  c.Ipk = base + kOffIpk;
  c.tmpr = base + kOffTmp;
  c.tw_row = base + kOffTwR;
  c.tw_col = base + kOffTwC;
  c.scal = (double*)(base + kOffScal);
  c.partials = (double*)(base + kOffPart);
  c.smask = (unsigned*)(base + kOffMask);
  if (support_mask_bytes(rows, cols, c.csz) > kMaskBytes)
    return fail(CDX_ERR_ARG, "cdx_phase_retrieve_static: %dx%d support mask exceeds the static workspace", rows, cols);
  c.ring = nullptr;
  c.ring_cap = 0;
  c.rcolor = 65535;
  c.scale = 1;
  c.beta = 1;
  c.relax = 1.9;
  hipLaunchKernelGGL(k_twiddles_f32, dim3((cols + 255) / 256), dim3(256), 0, c.st, cols, (float*)c.tw_row);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_twiddles_f32, dim3((rows + 255) / 256), dim3(256), 0, c.st, rows, (float*)c.tw_col);
  HIPCHK(hipGetLastError());
  // scalars from zero (a previous call's are still there), then the threshold
  HIPCHK(launch_fill_f64(S_COUNT, c.scal, 0.0, c.st));
  HIPCHK(launch_fill_f64(1, c.scal + S_THR, (double)(float)*threshold, c.st));
  rc = phase_retrieve_t<float>(&c, algorithm, &q, F, g, I, bs, spt, nullptr, 0, n_iter);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(threshold, c.scal + S_THR, sizeof(double), hipMemcpyDeviceToHost, c.st));
  HIPCHK(hipStreamSynchronize(c.st));
  return CDX_OK;
}
