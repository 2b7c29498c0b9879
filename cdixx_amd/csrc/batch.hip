// batch.hip -- many independent reconstructions of mixed sizes advanced
// together: every pass of every frame in one launch.
//
// Frames are packed back to back in shared buffers (F, g, I, beamstop,
// support and the T1/T2 intermediates).  A workgroup map lists, for every
// workgroup of a pass, its (frame, tile); the kernels switch on the frame's
// size and run the same pass bodies as the single-frame path, so a frame's
// arithmetic is identical in both.  Each frame keeps its own support
// threshold, shrink-wrap state and R-factor ring.  Workgroups have 256
// threads for every size (16 columns/rows of 256 or 384 points, 8 of 512),
// so column tiles are 128 bytes wide.
#include "../../include/cdx_abi.h"
#include "errors.hpp"
#include "pass_bodies.hpp"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <string>
#include <vector>

namespace cdx {

constexpr int kBatchThreads = 256;
constexpr int kBatchSizes[] = {128, 256, 384, 512};
constexpr int kNumBatchSizes = 4;
constexpr int kScal = 16;  // scalars per frame (same layout as a context's)

__host__ __device__ constexpr int bsize_index(int n) {
  return n == 128 ? 0 : n == 256 ? 1 : n == 384 ? 2 : n == 512 ? 3 : -1;
}
// default threads per workgroup: 256 (FP32) / 128 (FP64, whose two exchange
// buffers per line would otherwise exceed the LDS at 384); cdx_batch_autotune
// times 64, 128 and 256
template <class T>
constexpr int bthreads() { return sizeof(T) == 8 ? kBatchThreads / 2 : kBatchThreads; }

struct BFrame {
  int n;
  int wg_row0, nwg_row;  // this frame's workgroups in the row pass
  int pad;
  int64_t off;           // element offset of the frame in the packed buffers
  double* scal;          // kScal scalars: [0] threshold, [1] avg, [4] max bits
  double* rf;            // R-factor ring (2 doubles per projection)
};

template <class T>
struct BArgs {
  const BFrame* frames;
  const cplx<T>* src;
  cplx<T>* dst;
  cplx<T>* g;
  const T* Ipk;
  T* spt;
  const cplx<T>* tw[kNumBatchSizes];
  ProjArgs<T> pa;
  T beta, relax;
  double sigma;      // shrink-wrap blur sigma (the frame's sigma_f = (n>>2)/sigma)
  double* partials;
  int* nparts;
  int rslot;         // R-factor ring slot of this projection, -1: none
  int lsh[kNumBatchSizes];  // log2 lines per workgroup, per frame size
  int var[kNumBatchSizes];  // line variant (radix schedule), per frame size
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <class T, int OP, int MODE, int RF>
__global__ void __launch_bounds__(kBatchThreads) kb_col(BArgs<T> a, const int2* __restrict__ map) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int2 m = map[blockIdx.x];
  if (m.x < 0) return;  // padding entry (whole workgroup)
  // the batch's T2 (read by the update and shrink-wrap row passes) is in the
  // pair-of-rows layout, per frame
  constexpr bool t2l = OP == COL_FWD_PROJ_INV || OP == COL_PROJ_INV || OP == COL_FWD_GAUSS_INV || OP == COL_GAUSS_INV;
  // T1 (written by the update row pass) is in column tiles per frame
  constexpr bool t1l = OP == COL_FWD_PROJ_INV || OP == COL_FWD_GAUSS_INV || OP == COL_FWD;
  const BFrame f = a.frames[m.x];
  ProjArgs<T> pa = a.pa;
  pa.rf = (RF && a.rslot >= 0) ? f.rf + 2 * (size_t)a.rslot : nullptr;
  const T gs = (T)(f.n >> 2) / (T)a.sigma;
  const cplx<T>* in = a.src + f.off;
  cplx<T>* out = a.dst + f.off;
  const T* I = a.Ipk + f.off;
  switch (f.n) {
#define CASE(n_)                                                                                              \
  case n_:                                                                                                    \
    if (sizeof(T) == 4 && a.var[bsize_index(n_)])                                                             \
      col_body<T, n_, OP, MODE, RF, t1l, true, sizeof(T) == 4 ? 1 : 0, t2l>(                                  \
          in, out, n_, a.tw[bsize_index(n_)], I, pa, gs, m.y, a.lsh[bsize_index(n_)], smem);                  \
    else                                                                                                      \
      col_body<T, n_, OP, MODE, RF, t1l, true, 0, t2l>(in, out, n_, a.tw[bsize_index(n_)], I, pa, gs, m.y,     \
                                                          a.lsh[bsize_index(n_)], smem);                      \
    break;
    CASE(128) CASE(256) CASE(384) CASE(512)
#undef CASE
    default: break;
  }
}

template <class T, int OP, int ALGO>
__global__ void __launch_bounds__(kBatchThreads) kb_row(BArgs<T> a, const int2* __restrict__ map) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int2 m = map[blockIdx.x];
  if (m.x < 0) return;
  constexpr bool t2l = OP == ROW_INV_UPDATE_FWD || OP == ROW_INV_MOD2;
  constexpr bool t1l = OP == ROW_INV_UPDATE_FWD;
  const BFrame f = a.frames[m.x];
  UpdArgs<T> ua;
  ua.algo = ALGO;
  ua.beta = a.beta;
  ua.relax = a.relax;
  ua.invN = (T)(1.0 / ((double)f.n * (double)f.n));
  ua.scal = f.scal;
  ReduceSlots rs{(void*)(f.scal + 4), a.partials, a.nparts};
  const cplx<T>* in = a.src + f.off;
  cplx<T>* out = a.dst + f.off;
  switch (f.n) {
#define CASE(n_)                                                                                              \
  case n_:                                                                                                    \
    if (sizeof(T) == 4 && a.var[bsize_index(n_)])                                                             \
      row_body<T, n_, OP, ALGO, t1l, sizeof(T) == 4 ? 1 : 0, t2l>(in, out, n_, a.tw[bsize_index(n_)],          \
                                                                    a.g + f.off, a.spt + f.off, ua, (T)1,       \
                                                                    a.spt + f.off, rs, m.y,                     \
                                                                    a.lsh[bsize_index(n_)], smem);              \
    else                                                                                                      \
      row_body<T, n_, OP, ALGO, t1l, 0, t2l>(in, out, n_, a.tw[bsize_index(n_)], a.g + f.off, a.spt + f.off,   \
                                               ua, (T)1, a.spt + f.off, rs, m.y, a.lsh[bsize_index(n_)], smem); \
    break;
    CASE(128) CASE(256) CASE(384) CASE(512)
#undef CASE
    default: break;
  }
}

__global__ void kb_reset_avg(double* scal, int nf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nf) scal[(size_t)i * kScal + 1] = 0.0;
}

// per-frame shrink-wrap threshold (frequency form, cdi.cc:330-333; real form
// holo.cc:141): one wavefront per frame adds its partials in a fixed order
template <class T>
__global__ void __launch_bounds__(64) kb_threshold(const BFrame* frames, const double* partials,
                                                   double shrinkThreshold, int mode) {
  const BFrame f = frames[blockIdx.x];
  double s = 0;
  for (int i = threadIdx.x; i < f.nwg_row; i += 64) s += partials[f.wg_row0 + i];
  s = wave_sum(s);
  if (threadIdx.x != 0) return;
  double* scal = f.scal;
  T maxs;
  if (sizeof(T) == 4) maxs = (T)__uint_as_float(*(unsigned int*)(scal + 4));
  else maxs = (T)__longlong_as_double(*(long long*)(scal + 4));
  if (mode == 0) {
    T avg = (T)scal[1];
    if (avg == (T)0) {
      avg = (T)s * (T)4 / (T)f.n / (T)f.n;
      scal[1] = (double)avg;
    }
    if (avg * (T)5 < maxs) maxs = avg * (T)5;
  }
  scal[0] = (double)(T)((double)maxs * shrinkThreshold);
  scal[2] = (double)maxs;
  scal[4] = 0;
}

// elementwise over all frames: g *= s_f with s_f = 1/n_f (dir 0) or n_f (dir 1)
template <class T>
__global__ void kb_scale_g(const BFrame* frames, int nf, cplx<T>* g, int dir) {
  const BFrame f = frames[blockIdx.y];
  const int64_t n = (int64_t)f.n * f.n;
  const T s = dir == 0 ? (T)(1.0 / f.n) : (T)f.n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    cplx<T> v = g[f.off + i];
    g[f.off + i] = {v.x * s, v.y * s};
  }
}

}  // namespace cdx

using namespace cdx;

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct cdx_batch {
  int nf = 0, prec = 0;
  size_t rsz = 4, csz = 8;
  std::vector<int> n;
  std::vector<int64_t> off;
  int64_t total = 0;
  hipStream_t st = nullptr;
  bool own_stream = false;
  void *F = nullptr, *g = nullptr, *T1 = nullptr, *T2 = nullptr, *Ipk = nullptr, *I = nullptr, *bs = nullptr,
       *spt = nullptr;
  void* tw[kNumBatchSizes] = {nullptr, nullptr, nullptr, nullptr};
  BFrame* dframes = nullptr;
  int2* col_map = nullptr;
  int2* row_map = nullptr;
  int ncol = 0, nrow = 0;
  double* scal = nullptr;
  double* ring = nullptr;
  int ring_cap = 0;
  double* partials = nullptr;
  int* nparts = nullptr;
  size_t lds_col = 0, lds_row = 0;
  int threads = 0;                       // threads per workgroup of both passes
  int lsh[kNumBatchSizes] = {0, 0, 0, 0};  // log2 lines per workgroup, per frame size
  int var[kNumBatchSizes] = {0, 0, 0, 0};  // line variant (radix schedule) per frame size (FP32)
  int partials_cap = 0;
  std::vector<cdx_ctx*> ctx;  // a single-frame context per frame size (setup operations on slices)
  std::vector<int> ctx_n;
};

static thread_local char g_berr[512];
static int bfail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_berr, sizeof g_berr, fmt, ap);
  va_end(ap);
  set_last_error(g_berr);
  return code;
}
#define BCHK(expr)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return bfail(CDX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

// LDS bytes of a batched workgroup of frame size n with 2^lsh lines of
// variant v (both exchange buffers); lanes per line of size n, variant v
static size_t wg_lds_bytes(int n, size_t csz, int lsh, int v) {
  switch (n) {
#define F(n_)                                                                                     \
  case n_:                                                                                        \
    return csz == 8 ? (v ? lds_bytes<float, n_, 1>(lsh) : lds_bytes<float, n_>(lsh)) : lds_bytes<double, n_>(lsh);
    F(128) F(256) F(384) F(512)
#undef F
    default: return 0;
  }
}
static int line_lanes_b(int n, size_t csz, int v) {
  switch (n) {
#define F(n_) \
  case n_: return csz == 8 ? (v ? LineCfgT<float, n_, 1>::TL : LineCfgT<float, n_>::TL) : LineCfgT<double, n_>::TL;
    F(128) F(256) F(384) F(512)
#undef F
    default: return 0;
  }
}
// radix schedule of frame size n, variant v (forward order), e.g. "16x16"
static std::string radix_name(int n, size_t csz, int v) {
  auto fmt = [](const int* r, int m) {
    std::string s;
    for (int i = 0; i < m; i++) s += (i ? "x" : "") + std::to_string(r[i]);
    return s;
  };
  switch (n) {
#define F(n_)                                                                                            \
  case n_:                                                                                               \
    if (csz == 8 && v) return fmt(LineCfgT<float, n_, 1>::RL::R, LineCfgT<float, n_, 1>::RL::M);         \
    if (csz == 8) return fmt(LineCfgT<float, n_>::RL::R, LineCfgT<float, n_>::RL::M);                    \
    return fmt(LineCfgT<double, n_>::RL::R, LineCfgT<double, n_>::RL::M);
    F(128) F(256) F(384) F(512)
#undef F
    default: return "";
  }
}

// raises the dynamic-LDS cap of every batched pass instantiation the loop
// can launch to `lds` (both passes use the same), when the tiling is set and
// so before any stream capture
template <class T>
static hipError_t batch_prepare(size_t lds) {
  const void* ks[] = {
      (const void*)kb_col<T, COL_PROJ_INV, PROJ_MOD_ABS, 0>,    (const void*)kb_col<T, COL_PROJ_INV, PROJ_MOD_ABS, 1>,
      (const void*)kb_col<T, COL_PROJ_INV, PROJ_MOD_NOISE, 0>,  (const void*)kb_col<T, COL_PROJ_INV, PROJ_MOD_NOISE, 1>,
      (const void*)kb_col<T, COL_PROJ_INV, PROJ_MOD_ABS_BS, 0>, (const void*)kb_col<T, COL_PROJ_INV, PROJ_MOD_ABS_BS, 1>,
      (const void*)kb_col<T, COL_FWD_PROJ_INV, PROJ_MOD_ABS, 0>,
      (const void*)kb_col<T, COL_FWD_PROJ_INV, PROJ_MOD_ABS, 1>,
      (const void*)kb_col<T, COL_FWD_PROJ_INV, PROJ_MOD_NOISE, 0>,
      (const void*)kb_col<T, COL_FWD_PROJ_INV, PROJ_MOD_NOISE, 1>,
      (const void*)kb_col<T, COL_FWD_PROJ_INV, PROJ_MOD_ABS_BS, 0>,
      (const void*)kb_col<T, COL_FWD_PROJ_INV, PROJ_MOD_ABS_BS, 1>,
      (const void*)kb_col<T, COL_GAUSS_INV, 0, 0>,              (const void*)kb_col<T, COL_FWD_GAUSS_INV, 0, 0>,
      (const void*)kb_col<T, COL_FWD, 0, 0>,                    (const void*)kb_row<T, ROW_INV_UPDATE_FWD, 0>,
      (const void*)kb_row<T, ROW_INV_UPDATE_FWD, 1>,            (const void*)kb_row<T, ROW_INV_UPDATE_FWD, 2>,
      (const void*)kb_row<T, ROW_INV_UPDATE_FWD, 3>,            (const void*)kb_row<T, ROW_INV_UPDATE_FWD, 4>,
      (const void*)kb_row<T, ROW_INV_UPDATE_FWD, 5>,            (const void*)kb_row<T, ROW_INV_MOD2, 0>};
  for (const void* k : ks) {
    hipError_t e = lds_cap(k, lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Workgroup maps of both passes for `threads` threads per workgroup (64,
// 128 or 256): frame size n runs threads / TL(n) lines per workgroup.
// Returns CDX_ERR_ARG (maps unchanged) when a size present cannot take it.
static int batch_tiling(cdx_batch* b, int threads, const int* var) {
  if (threads != 64 && threads != 128 && threads != 256)
    return bfail(CDX_ERR_ARG, "cdx_batch_set_tiling: %d threads per workgroup (64, 128 or 256)", threads);
  int lsh[kNumBatchSizes] = {0, 0, 0, 0};
  size_t lds = 0;
  for (int k = 0; k < kNumBatchSizes; k++) {
    if (var[k] != 0 && (var[k] != 1 || b->prec != CDX_F32))
      return bfail(CDX_ERR_ARG, "cdx_batch_set_schedule: line variant %d for size %d (0, or 1 in FP32)", var[k],
                   kBatchSizes[k]);
    bool used = false;
    for (int s : b->n) used |= s == kBatchSizes[k];
    if (!used) continue;
    const int tl = line_lanes_b(kBatchSizes[k], b->csz, var[k]), lines = threads / tl;
    if (lines < 1 || lines * tl != threads || kBatchSizes[k] % lines)
      return bfail(CDX_ERR_ARG, "cdx_batch_set_tiling: %d threads do not tile %d-point lines", threads, kBatchSizes[k]);
    while ((1 << lsh[k]) < lines) lsh[k]++;
    lds = std::max(lds, wg_lds_bytes(kBatchSizes[k], b->csz, lsh[k], var[k]));
  }
  if (lds > 160 * 1024)
    return bfail(CDX_ERR_ARG, "cdx_batch_set_tiling: %d threads need %zu bytes of LDS", threads, lds);
  // best effort here; a launch the device refuses reports the error itself
  (void)(b->prec == CDX_F64 ? batch_prepare<double>(lds) : batch_prepare<float>(lds));
  std::vector<int2> cols, rows;
  std::vector<BFrame> fr(b->nf);
  for (int i = 0; i < b->nf; i++) {
    const int n = b->n[i], lines = 1 << lsh[bsize_index(n)];
    for (int t = 0; t < n / lines; t++) cols.push_back(make_int2(i, t));
    fr[i].n = n;
    fr[i].wg_row0 = (int)rows.size();
    fr[i].nwg_row = n / lines;
    for (int t = 0; t < n / lines; t++) rows.push_back(make_int2(i, t));
    fr[i].pad = 0;
    fr[i].off = b->off[i];
    fr[i].scal = b->scal + (size_t)i * kScal;
    fr[i].rf = b->ring + (size_t)i * b->ring_cap * 2 * kRfSlots;
  }
  if ((int)rows.size() > b->partials_cap)
    return bfail(CDX_ERR_ARG, "cdx_batch_set_tiling: too many row workgroups (%zu)", rows.size());
  // column tiles sharing cache lines go to the same XCD: workgroup id b runs
  // tile xcd(b) of the (padded) tile list
  while (cols.size() % 8) cols.push_back(make_int2(-1, 0));
  std::vector<int2> cmap(cols.size());
  const int nb = (int)cols.size();
  for (int bid = 0; bid < nb; bid++) cmap[bid] = cols[(bid & 7) * (nb >> 3) + (bid >> 3)];
  int2 *cm = nullptr, *rm = nullptr;
  BFrame* df = nullptr;
  if (hipMalloc((void**)&cm, cmap.size() * sizeof(int2)) || hipMalloc((void**)&rm, rows.size() * sizeof(int2)) ||
      hipMalloc((void**)&df, fr.size() * sizeof(BFrame)) ||
      hipMemcpy(cm, cmap.data(), cmap.size() * sizeof(int2), hipMemcpyHostToDevice) ||
      hipMemcpy(rm, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice) ||
      hipMemcpy(df, fr.data(), fr.size() * sizeof(BFrame), hipMemcpyHostToDevice)) {
    hipFree(cm);
    hipFree(rm);
    hipFree(df);
    return bfail(CDX_ERR_HIP, "cdx_batch_set_tiling: out of device memory");
  }
  if (b->st) hipStreamSynchronize(b->st);
  hipFree(b->col_map);
  hipFree(b->row_map);
  hipFree(b->dframes);
  b->col_map = cm;
  b->row_map = rm;
  b->dframes = df;
  b->ncol = nb;
  b->nrow = (int)rows.size();
  b->threads = threads;
  for (int k = 0; k < kNumBatchSizes; k++) {
    b->lsh[k] = lsh[k];
    b->var[k] = var[k];
  }
  b->lds_col = b->lds_row = lds;
  return CDX_OK;
}

extern "C" int cdx_batch_destroy(cdx_batch* b) {
  if (!b) return CDX_OK;
  if (b->st) hipStreamSynchronize(b->st);
  for (void* p : {b->F, b->g, b->T1, b->T2, b->Ipk, b->I, b->bs, b->spt}) hipFree(p);
  for (void* p : b->tw) hipFree(p);
  hipFree(b->dframes);
  hipFree(b->col_map);
  hipFree(b->row_map);
  hipFree(b->scal);
  hipFree(b->ring);
  hipFree(b->partials);
  hipFree(b->nparts);
  for (auto* c : b->ctx) cdx_destroy(c);
  if (b->own_stream && b->st) hipStreamDestroy(b->st);
  delete b;
  return CDX_OK;
}

extern "C" int cdx_batch_create(cdx_batch** out, int nframes, const int* sizes, int precision, void* stream) {
  if (!out || nframes <= 0 || !sizes) return bfail(CDX_ERR_ARG, "cdx_batch_create: bad arguments");
  if (precision != CDX_F32 && precision != CDX_F64) return bfail(CDX_ERR_ARG, "cdx_batch_create: bad precision");
  *out = nullptr;
  for (int i = 0; i < nframes; i++)
    if (bsize_index(sizes[i]) < 0)
      return bfail(CDX_ERR_ARG, "cdx_batch_create: frame %d has size %d (batched sizes: 128, 256, 384, 512)", i,
                   sizes[i]);
  cdx_batch* b = new cdx_batch();
  b->nf = nframes;
  b->prec = precision;
  b->rsz = precision == CDX_F64 ? 8 : 4;
  b->csz = 2 * b->rsz;
  b->n.assign(sizes, sizes + nframes);
  b->off.resize(nframes);
  for (int i = 0; i < nframes; i++) {
    b->off[i] = b->total;
    b->total += (int64_t)sizes[i] * sizes[i];
  }
  int rc;
  if (stream) {
    b->st = (hipStream_t)stream;
  } else {
    if (hipStreamCreate(&b->st) != hipSuccess) {
      delete b;
      return bfail(CDX_ERR_HIP, "cdx_batch_create: hipStreamCreate failed");
    }
    b->own_stream = true;
  }
  auto bad = [&](int code) {
    cdx_batch_destroy(b);
    return code;
  };
  const size_t tc = (size_t)b->total * b->csz, tr = (size_t)b->total * b->rsz;
  if (hipMalloc(&b->F, tc) || hipMalloc(&b->g, tc) || hipMalloc(&b->T1, tc) || hipMalloc(&b->T2, tc) ||
      hipMalloc(&b->Ipk, tr) || hipMalloc(&b->I, tr) || hipMalloc(&b->bs, tr) || hipMalloc(&b->spt, tr))
    return bad(bfail(CDX_ERR_HIP, "cdx_batch_create: out of device memory"));
  hipMemset(b->bs, 0, tr);
  // twiddles and single-frame contexts for the sizes present
  for (int k = 0; k < kNumBatchSizes; k++) {
    bool used = false;
    for (int s : b->n) used |= s == kBatchSizes[k];
    if (!used) continue;
    if ((rc = make_twiddles(kBatchSizes[k], precision, &b->tw[k]))) return bad(rc);
    cdx_ctx* c;
    if ((rc = cdx_create(&c, kBatchSizes[k], kBatchSizes[k], precision, b->st))) return bad(rc);
    b->ctx.push_back(c);
    b->ctx_n.push_back(kBatchSizes[k]);
  }
  b->ring_cap = 4096;
  if (hipMalloc((void**)&b->scal, (size_t)nframes * kScal * sizeof(double)) ||
      hipMalloc((void**)&b->ring, (size_t)nframes * b->ring_cap * 2 * kRfSlots * sizeof(double)))
    return bad(bfail(CDX_ERR_HIP, "cdx_batch_create: out of device memory"));
  // one partial sum per row workgroup: at most one workgroup per row
  for (int i = 0; i < nframes; i++) b->partials_cap += b->n[i];
  if (hipMalloc((void**)&b->partials, (size_t)b->partials_cap * sizeof(double)) ||
      hipMalloc((void**)&b->nparts, sizeof(int)))
    return bad(bfail(CDX_ERR_HIP, "cdx_batch_create: out of device memory"));
  if ((rc = batch_tiling(b, precision == CDX_F64 ? bthreads<double>() : bthreads<float>(), b->var))) return bad(rc);
  std::vector<double> sc((size_t)nframes * kScal, 0.0);
  for (int i = 0; i < nframes; i++) sc[(size_t)i * kScal] = 0.5;  // init_cuda_image threshold
  if (hipMemcpy(b->scal, sc.data(), sc.size() * sizeof(double), hipMemcpyHostToDevice))
    return bad(bfail(CDX_ERR_HIP, "cdx_batch_create: upload failed"));
  *out = b;
  return CDX_OK;
}

extern "C" int cdx_batch_frame(cdx_batch* b, int i, void** F, void** g, void** I, void** bs, void** spt, int* n) {
  if (!b || i < 0 || i >= b->nf) return bfail(CDX_ERR_ARG, "cdx_batch_frame: bad frame index");
  const int64_t o = b->off[i];
  if (F) *F = (char*)b->F + o * b->csz;
  if (g) *g = (char*)b->g + o * b->csz;
  if (I) *I = (char*)b->I + o * b->rsz;
  if (bs) *bs = (char*)b->bs + o * b->rsz;
  if (spt) *spt = (char*)b->spt + o * b->rsz;
  if (n) *n = b->n[i];
  return CDX_OK;
}

extern "C" cdx_ctx* cdx_batch_context(cdx_batch* b, int i) {
  if (!b || i < 0 || i >= b->nf) return nullptr;
  for (size_t k = 0; k < b->ctx.size(); k++)
    if (b->ctx_n[k] == b->n[i]) return b->ctx[k];
  return nullptr;
}

extern "C" int cdx_batch_set_threshold(cdx_batch* b, double thr) {
  if (!b) return bfail(CDX_ERR_ARG, "cdx_batch_set_threshold: null batch");
  std::vector<double> h((size_t)b->nf * kScal);
  BCHK(hipMemcpy(h.data(), b->scal, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  // the shrink-wrap average is left alone: each cdx_batch_phase_retrieve
  // takes it again on its first shrink-wrap (kb_reset_avg), as the
  // single-frame cdx_set_threshold leaves the context's
  for (int i = 0; i < b->nf; i++) h[(size_t)i * kScal] = b->prec == CDX_F64 ? thr : (double)(float)thr;
  BCHK(hipMemcpy(b->scal, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  return CDX_OK;
}

extern "C" int cdx_batch_get_thresholds(cdx_batch* b, double* out) {
  if (!b || !out) return bfail(CDX_ERR_ARG, "cdx_batch_get_thresholds: bad arguments");
  std::vector<double> h((size_t)b->nf * kScal);
  BCHK(hipStreamSynchronize(b->st));
  BCHK(hipMemcpy(h.data(), b->scal, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i < b->nf; i++) out[i] = h[(size_t)i * kScal];
  return CDX_OK;
}

namespace {
template <class T>
struct BLoop {
  cdx_batch* b;
  const cdx_params* p;
  ProjArgs<T> pa;
  bool f_pending = true;

  BArgs<T> args(const void* src, void* dst) {
    BArgs<T> a{};
    a.frames = b->dframes;
    a.src = (const cplx<T>*)src;
    a.dst = (cplx<T>*)dst;
    a.g = (cplx<T>*)b->g;
    a.Ipk = (const T*)b->Ipk;
    a.spt = (T*)b->spt;
    for (int k = 0; k < kNumBatchSizes; k++) a.tw[k] = (const cplx<T>*)b->tw[k];
    a.pa = pa;
    a.beta = (T)p->beta;
    a.relax = (T)p->relax;
    a.sigma = 1;
    a.partials = b->partials;
    a.nparts = b->nparts;
    a.rslot = -1;
    for (int k = 0; k < kNumBatchSizes; k++) {
      a.lsh[k] = b->lsh[k];
      a.var[k] = b->var[k];
    }
    return a;
  }
  template <class K>
  hipError_t launch(K kern, int nwg, const int2* map, const BArgs<T>& a, size_t lds) {
    // the caps were raised when the tiling was set (batch_prepare); lds_cap
    // only looks them up here
    hipError_t e = lds_cap((const void*)kern, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(b->threads), lds, b->st, a, map);
    return hipGetLastError();
  }
  int setup() {
    const bool pack_bs = p->useBS && (p->simCCDbit || p->freeBeamstop);
    // packed per frame in the column-tile layout the batched column pass reads
    for (int i = 0; i < b->nf; i++) {
      const int64_t o = b->off[i];
      BCHK(launch_pack_intensity_tiled<T>(b->n[i], b->n[i], (const T*)b->I + o, pack_bs ? (const T*)b->bs + o : nullptr,
                                          (T*)b->Ipk + o, !p->simCCDbit, b->st));
    }
    pa = ProjArgs<T>{};
    pa.mode = p->simCCDbit ? PROJ_MOD_NOISE : (pack_bs ? PROJ_MOD_ABS_BS : PROJ_MOD_ABS);
    pa.noiseLevel = p->noiseLevel;
    pa.scale = (T)(1.0 / p->exposure);
    pa.rcolor = (T)p->rcolor;
    pa.rfs = 2 * b->ring_cap;
    return CDX_OK;
  }
  int iterate(int algo, int rslot) {
    BArgs<T> a = args(f_pending ? b->F : b->T1, b->T2);
    a.rslot = rslot;
    const bool rf = rslot >= 0;
    if (f_pending) {
      ProjArgs<T> rpa = pa;
      a.pa = rpa;
#define M(mode)                                                                                                  \
  BCHK(rf ? launch(kb_col<T, COL_PROJ_INV, mode, 1>, b->ncol, b->col_map, a, b->lds_col)                         \
          : launch(kb_col<T, COL_PROJ_INV, mode, 0>, b->ncol, b->col_map, a, b->lds_col));
      switch (pa.mode) {
        case PROJ_MOD_ABS: M(PROJ_MOD_ABS) break;
        case PROJ_MOD_NOISE: M(PROJ_MOD_NOISE) break;
        default: M(PROJ_MOD_ABS_BS) break;
      }
#undef M
    } else {
#define M(mode)                                                                                                  \
  BCHK(rf ? launch(kb_col<T, COL_FWD_PROJ_INV, mode, 1>, b->ncol, b->col_map, a, b->lds_col)                     \
          : launch(kb_col<T, COL_FWD_PROJ_INV, mode, 0>, b->ncol, b->col_map, a, b->lds_col));
      switch (pa.mode) {
        case PROJ_MOD_ABS: M(PROJ_MOD_ABS) break;
        case PROJ_MOD_NOISE: M(PROJ_MOD_NOISE) break;
        default: M(PROJ_MOD_ABS_BS) break;
      }
#undef M
    }
    BArgs<T> r = args(b->T2, b->T1);
    switch (algo) {
#define A(k) \
  case k: BCHK(launch(kb_row<T, ROW_INV_UPDATE_FWD, k>, b->nrow, b->row_map, r, b->lds_row)); break;
      A(0) A(1) A(2) A(3) A(4)
      default: BCHK(launch(kb_row<T, ROW_INV_UPDATE_FWD, 5>, b->nrow, b->row_map, r, b->lds_row)); break;
#undef A
    }
    f_pending = false;
    return CDX_OK;
  }
  int shrink_wrap(double sigma) {
    if (p->shrinkMode != 0) return bfail(CDX_ERR_UNSUPPORTED, "batched shrink-wrap: frequency form only");
    BArgs<T> a = args(f_pending ? b->F : b->T1, b->T2);
    a.sigma = sigma;
    if (f_pending) BCHK(launch(kb_col<T, COL_GAUSS_INV, 0, 0>, b->ncol, b->col_map, a, b->lds_col));
    else BCHK(launch(kb_col<T, COL_FWD_GAUSS_INV, 0, 0>, b->ncol, b->col_map, a, b->lds_col));
    BArgs<T> r = args(b->T2, nullptr);
    BCHK(launch(kb_row<T, ROW_INV_MOD2, 0>, b->nrow, b->row_map, r, b->lds_row));
    hipLaunchKernelGGL(kb_threshold<T>, dim3(b->nf), dim3(64), 0, b->st, (const BFrame*)b->dframes,
                       (const double*)b->partials, p->shrinkThreshold, 0);
    BCHK(hipGetLastError());
    return CDX_OK;
  }
  int finish() {
    if (!f_pending) {
      BArgs<T> a = args(b->T1, b->F);
      BCHK(launch(kb_col<T, COL_FWD, 0, 0>, b->ncol, b->col_map, a, b->lds_col));
      f_pending = true;
    }
    return CDX_OK;
  }
  int scale_g(int dir) {
    int maxn = 0;
    for (int s : b->n) maxn = std::max(maxn, s);
    dim3 grid((unsigned)std::min<int64_t>(((int64_t)maxn * maxn + 255) / 256, 1024), (unsigned)b->nf);
    hipLaunchKernelGGL(kb_scale_g<T>, grid, dim3(256), 0, b->st, (const BFrame*)b->dframes, b->nf, (cplx<T>*)b->g,
                       dir);
    BCHK(hipGetLastError());
    return CDX_OK;
  }
};

template <class T>
int batch_retrieve_t(cdx_batch* b, const char* algorithm, const cdx_params* p, double* rtrace, int rcap,
                     int* n_iter) {
  int nt = cdx_expand_schedule(algorithm, nullptr, 0);
  if (nt < 0) return bfail(CDX_ERR_SCHEDULE, "cdx_batch_phase_retrieve: bad algorithm string \"%s\"", algorithm);
  std::vector<int> toks(nt);
  cdx_expand_schedule(algorithm, toks.data(), nt);
  for (int t : toks)
    if (t == CDX_TV) return bfail(CDX_ERR_UNSUPPORTED, "cdx_batch_phase_retrieve: TV steps are outside this path");
  BLoop<T> L{b, p};
  int rc = L.setup();
  if (rc) return rc;
  // each frame's shrink-wrap average is taken on the first shrink-wrap of
  // each call (a local of CDI::phaseRetrieve, cdi.cc:315,331)
  hipLaunchKernelGGL(kb_reset_avg, dim3((b->nf + 63) / 64), dim3(64), 0, b->st, b->scal, b->nf);
  BCHK(hipGetLastError());
  const int nrec = p->recordR ? std::min(rcap, b->ring_cap) : 0;
  if (nrec > 0) BCHK(launch_fill_f64((int64_t)b->nf * b->ring_cap * 2 * kRfSlots, b->ring, 0.0, b->st));
  if ((rc = L.scale_g(0))) return rc;  // cdi.cc:313
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (p->useGraph) BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  double sigma = p->sigma0;
  int iter = 0, nproj = 0;
  for (int tok : toks) {
    if (tok == CDX_SHRINKWRAP) {
      if ((rc = L.shrink_wrap(sigma))) break;
      if (sigma > p->sigmaMin) sigma *= p->sigmaDecay;
    } else {
      if ((rc = L.iterate(tok, nproj < nrec ? nproj : -1))) break;
      nproj++;
    }
    iter++;
  }
  if (!rc) rc = L.finish();
  if (p->useGraph) {
    hipError_t e = hipStreamEndCapture(b->st, &graph);
    if (!rc && e != hipSuccess) rc = bfail(CDX_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (!rc && (e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0)) != hipSuccess)
      rc = bfail(CDX_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    if (!rc && (e = hipGraphLaunch(exec, b->st)) != hipSuccess)
      rc = bfail(CDX_ERR_HIP, "hipGraphLaunch: %s", hipGetErrorString(e));
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
  }
  if (rc) return rc;
  if ((rc = L.scale_g(1))) return rc;  // cdi.cc:406
  BCHK(hipStreamSynchronize(b->st));
  if (n_iter) *n_iter = iter;
  if (rtrace && nrec > 0) {
    std::vector<double> h((size_t)b->nf * b->ring_cap * 2 * kRfSlots);
    BCHK(hipMemcpy(h.data(), b->ring, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    const int m = std::min(nproj, nrec);
    for (int f = 0; f < b->nf; f++)
      for (int i = 0; i < m; i++) {
        // frame f's ring: kRfSlots rows of ring_cap (num, den) pairs
        const double* s = &h[(size_t)f * b->ring_cap * 2 * kRfSlots + 2 * (size_t)i];
        double num = 0, den = 0;
        for (int k = 0; k < kRfSlots; k++) {
          num += s[2 * (size_t)k * b->ring_cap];
          den += s[2 * (size_t)k * b->ring_cap + 1];
        }
        rtrace[(size_t)f * rcap + i] = den > 0 ? num / den : 0;
      }
  }
  return CDX_OK;
}

// events and graph of a batched timing run, released on every exit path
struct BTimingRes {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~BTimingRes() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// one untimed iteration, with use_graph one untimed launch of the captured
// n iterations, then n timed iterations; g enters and leaves in the scaling
// of cdx_batch_phase_retrieve (cdi.cc:313,406)
template <class T>
int batch_time_t(cdx_batch* b, const cdx_params* p, int algo, int n, int use_graph, double* ms_out) {
  BLoop<T> L{b, p};
  int rc = L.setup();
  if (rc) return rc;
  BTimingRes res;
  BCHK(hipEventCreate(&res.e0));
  BCHK(hipEventCreate(&res.e1));
  if ((rc = L.scale_g(0))) return rc;
  if ((rc = L.iterate(algo, -1))) return rc;
  if (use_graph) {
    BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < n && !rc; i++) rc = L.iterate(algo, -1);
    hipError_t e = hipStreamEndCapture(b->st, &res.graph);
    if (rc) return rc;
    BCHK(e);
    BCHK(hipGraphInstantiate(&res.exec, res.graph, nullptr, nullptr, 0));
    BCHK(hipGraphLaunch(res.exec, b->st));
  }
  BCHK(hipStreamSynchronize(b->st));
  BCHK(hipEventRecord(res.e0, b->st));
  if (use_graph) {
    BCHK(hipGraphLaunch(res.exec, b->st));
  } else {
    for (int i = 0; i < n && !rc; i++) rc = L.iterate(algo, -1);
    if (rc) return rc;
  }
  BCHK(hipEventRecord(res.e1, b->st));
  BCHK(hipEventSynchronize(res.e1));
  float ms = 0;
  BCHK(hipEventElapsedTime(&ms, res.e0, res.e1));
  *ms_out = ms;
  if ((rc = L.finish())) return rc;
  if ((rc = L.scale_g(1))) return rc;
  BCHK(hipStreamSynchronize(b->st));
  return CDX_OK;
}
}  // namespace

extern "C" int cdx_batch_phase_retrieve(cdx_batch* b, const char* algorithm, const cdx_params* p, double* rtrace,
                                        int rcap, int* n_iter) {
  if (!b || !algorithm || !p) return bfail(CDX_ERR_ARG, "cdx_batch_phase_retrieve: null argument");
  if (!(p->exposure != 0)) return bfail(CDX_ERR_ARG, "cdx_batch_phase_retrieve: exposure must be non-zero");
  if (p->prtfIter >= 0)
    return bfail(CDX_ERR_UNSUPPORTED, "cdx_batch_phase_retrieve: PRTF is computed by single-frame contexts only");
  return b->prec == CDX_F64 ? batch_retrieve_t<double>(b, algorithm, p, rtrace, rcap, n_iter)
                            : batch_retrieve_t<float>(b, algorithm, p, rtrace, rcap, n_iter);
}

extern "C" int cdx_batch_time_iterations(cdx_batch* b, const cdx_params* p, int algo, int n, int use_graph,
                                         double* ms) {
  if (!b || !p || n <= 0 || !ms) return bfail(CDX_ERR_ARG, "cdx_batch_time_iterations: bad arguments");
  return b->prec == CDX_F64 ? batch_time_t<double>(b, p, algo, n, use_graph, ms)
                            : batch_time_t<float>(b, p, algo, n, use_graph, ms);
}

extern "C" int cdx_batch_set_tiling(cdx_batch* b, int threads) {
  if (!b) return bfail(CDX_ERR_ARG, "cdx_batch_set_tiling: null batch");
  return batch_tiling(b, threads, b->var);
}

extern "C" int cdx_batch_get_tiling(cdx_batch* b, int* threads) {
  if (!b || !threads) return bfail(CDX_ERR_ARG, "cdx_batch_get_tiling: bad arguments");
  *threads = b->threads;
  return CDX_OK;
}

extern "C" int cdx_batch_set_schedule(cdx_batch* b, int threads, const int* variants) {
  if (!b || !variants) return bfail(CDX_ERR_ARG, "cdx_batch_set_schedule: bad arguments");
  return batch_tiling(b, threads, variants);
}

extern "C" int cdx_batch_get_schedule(cdx_batch* b, int* threads, int* variants, char* radices, int cap) {
  if (!b || !threads || !variants) return bfail(CDX_ERR_ARG, "cdx_batch_get_schedule: bad arguments");
  *threads = b->threads;
  std::string r;
  for (int k = 0; k < kNumBatchSizes; k++) {
    variants[k] = b->var[k];
    bool used = false;
    for (int s : b->n) used |= s == kBatchSizes[k];
    if (used)
      r += (r.empty() ? "" : " ") + std::to_string(kBatchSizes[k]) + ":" + radix_name(kBatchSizes[k], b->csz, b->var[k]);
  }
  if (radices && cap > 0) snprintf(radices, (size_t)cap, "%s", r.c_str());
  return CDX_OK;
}

// Times `reps` HIO iterations (one captured graph) for every schedule: each
// workgroup size (64, 128, 256 threads) with each combination of line
// variants over the frame sizes present (FP32; FP64 has one variant), and
// keeps the fastest.  F and g are restored afterwards, so the frames' state
// is as before the call.  us[k] = microseconds per iteration of the best
// schedule at 64 << k threads (0 where no schedule fits the batch).
extern "C" int cdx_batch_autotune(cdx_batch* b, int reps, double* us, int nus) {
  if (!b || reps <= 0) return bfail(CDX_ERR_ARG, "cdx_batch_autotune: bad arguments");
  const size_t bytes = (size_t)b->total * b->csz;
  void* save = nullptr;
  BCHK(hipMalloc(&save, 2 * bytes));
  BCHK(hipStreamSynchronize(b->st));
  BCHK(hipMemcpy(save, b->F, bytes, hipMemcpyDeviceToDevice));
  BCHK(hipMemcpy((char*)save + bytes, b->g, bytes, hipMemcpyDeviceToDevice));
  cdx_params p;
  cdx_params_default(&p);
  // sizes present get a variant bit each (FP32)
  int present[kNumBatchSizes], np = 0;
  for (int k = 0; k < kNumBatchSizes; k++) {
    bool used = false;
    for (int s : b->n) used |= s == kBatchSizes[k];
    if (used && b->prec == CDX_F32) present[np++] = k;
  }
  const int start_th = b->threads;
  int start_var[kNumBatchSizes], best_var[kNumBatchSizes];
  for (int k = 0; k < kNumBatchSizes; k++) start_var[k] = best_var[k] = b->var[k];
  int best_th = start_th, rc = CDX_OK;
  double best_ms = 0;
  for (int k = 0; k < 3 && !rc; k++) {
    const int th = 64 << k;
    if (k < nus && us) us[k] = 0;
    for (int mask = 0; mask < (1 << np) && !rc; mask++) {
      int var[kNumBatchSizes] = {0, 0, 0, 0};
      for (int q = 0; q < np; q++) var[present[q]] = (mask >> q) & 1;
      if (batch_tiling(b, th, var) != CDX_OK) continue;  // does not fit the batch: skipped
      double ms = 0;
      rc = cdx_batch_time_iterations(b, &p, CDX_HIO, reps, 1, &ms);
      if (rc) break;
      const double u = 1e3 * ms / reps;
      if (k < nus && us && (us[k] == 0 || u < us[k])) us[k] = u;
      if (best_ms == 0 || ms < best_ms) {
        best_ms = ms;
        best_th = th;
        for (int q = 0; q < kNumBatchSizes; q++) best_var[q] = var[q];
      }
    }
  }
  const int rc2 = rc ? batch_tiling(b, start_th, start_var) : batch_tiling(b, best_th, best_var);
  hipError_t e1 = hipMemcpy(b->F, save, bytes, hipMemcpyDeviceToDevice);
  hipError_t e2 = hipMemcpy(b->g, (char*)save + bytes, bytes, hipMemcpyDeviceToDevice);
  hipFree(save);
  if (rc) return rc;
  if (rc2) return rc2;
  if (e1 != hipSuccess || e2 != hipSuccess) return bfail(CDX_ERR_HIP, "cdx_batch_autotune: restore failed");
  set_last_error("");
  return CDX_OK;
}

extern "C" int cdx_batch_synchronize(cdx_batch* b) {
  if (!b) return bfail(CDX_ERR_ARG, "cdx_batch_synchronize: null batch");
  BCHK(hipStreamSynchronize(b->st));
  return CDX_OK;
}
