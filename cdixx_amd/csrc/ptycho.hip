// ptycho.hip -- the scan-position loop of CDIxx's ptychography program
// (util/ptycho.cpp:514-561) on the FFT-pair machinery of the CDI loop.
//
// For every scan position i, in the caller's order, the reference does
//   O_w  = getWindow(O, sx_i, sy_i)                            ptycho.cu:17-23
//   psi  = P * O_w                                              ptycho.cpp:526
//   F    = FFT(psi), projected with bsnorm = 1/sqrt(N) by
//          applyModAccurate (cdi.cu:351-388) or applyMod (cdi.cu:438-485)
//   psi' = IFFT(F) / sqrt(N),  U = psi - psi'                   ptycho.cpp:549-552
//   O_w  = rPIE(O_w, P, U)  [+ probeStep += rPIE_step(P, O_w, U) * probeStepSize]
//                                                               ptycho.cu:50-108
//   updateWindow(O, sx_i, sy_i, O_w)                            ptycho.cu:25-30
// one position after another.  Here the positions are grouped into levels: a
// position's level is one more than the highest level of the earlier
// positions whose windows (clipped to the object) overlap its own.  The
// positions of one level touch disjoint object pixels, and each reads the
// object as the sequential loop leaves it, so the object after a sweep is the
// sequential loop's bit for bit.  Only the probe-step sum is added in level
// order instead of scan order.
//
// A level is three launches over (tile, position) grids:
//   kp_fwd   gather O_w * P, forward row FFT                   -> slot A (column tiles)
//   kp_col   the CDI column pass body (col_body, COL_FWD_PROJ_INV) with the
//            ptychographic projection                          A -> slot B
//   kp_upd   inverse row FFT, 1/sqrt(N), U, rPIE, scatter into O
//            [per-slot probe-step terms, added in slot order by kp_pstep]
#include "../../include/cdx_abi.h"
#include "errors.hpp"
#include "pass_bodies.hpp"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <vector>

namespace cdx {

constexpr int kPtSlots = 64;  // positions in flight per launch (slot buffers)
constexpr float kAlpha = 0.1f;  // ALPHA, ptycho.cu:5 (object step)
constexpr float kBeta = 0.5f;   // BETA, ptycho.cu:6 (probe step)

// lines per workgroup (log2) for NT threads per workgroup
template <int N, int NT>
constexpr int plsh() {
  int s = 0;
  while ((1 << s) < NT / LineCfg<N>::TL) s++;
  return s;
}

// one scan position: pattern index and window offsets (row, column)
struct PtPos {
  int scan, sx, sy, pad;
};

struct PtArgs {
  const PtPos* pos;  // positions of this launch, one per blockIdx.y
  cplx<float>* obj;
  int orow, ocol;
  const cplx<float>* probe;
  cplx<float>* slotA;  // kPtSlots buffers of N*N
  cplx<float>* slotB;
  const float* Ipk;    // packed patterns (NaN on the beamstop), nscan * N*N
  const cplx<float>* tw;
  cplx<float>* pterm;  // per-slot probe-step terms (updateProbe) or null
  ProjArgs<float> pa;
  float norm;          // 1/sqrt(N*N)
  float probeMax, objMax, objStep, probeStepSize;
  const float* probeMaxDev;  // if set: max |P|^2 computed on the device
  int updateProbe;
};

// rPIE (ptycho.cu:58-64): target -= conj(source) diff * step / ((1-a)|source|^2 + a max(maxi, 1e-5))
__device__ __forceinline__ void rpie(cplx<float>& target, cplx<float> s, cplx<float> d, float maxi, float param,
                                     float step) {
  float denom = s.x * s.x + s.y * s.y;
  denom = step / ((1 - param) * denom + param * fmaxf(maxi, 1e-5f));
  target.x -= (s.x * d.x + s.y * d.y) * denom;
  target.y -= (s.x * d.y - s.y * d.x) * denom;
}
// rPIE_step (ptycho.cu:65-71): target = -conj(source) diff / ((1-a)|source|^2 + a max(maxi, 1e-5))
__device__ __forceinline__ cplx<float> rpie_step(cplx<float> s, cplx<float> d, float maxi, float param) {
  float denom = s.x * s.x + s.y * s.y;
  denom = 1.f / ((1 - param) * denom + param * fmaxf(maxi, 1e-5f));
  return {-(s.x * d.x + s.y * d.y) * denom, -(s.x * d.y - s.y * d.x) * denom};
}

// gather the window, multiply by the probe, forward row FFT -> slot A
template <int N, int NT>
__global__ void __launch_bounds__(NT) kp_fwd(PtArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  using L = LineOf<float, N>;
  constexpr int TL = LineCfg<N>::TL, E = L::E, LSH = plsh<N, NT>();
  const int pitch = lds_pitch<float, N>(LSH), alt = pitch << LSH;
  int ph = 0;
  const int t = threadIdx.x % TL, rr = threadIdx.x / TL;
  const int row = (blockIdx.x << LSH) + rr;
  const PtPos p = a.pos[blockIdx.y];
  cplx<float>* my = reinterpret_cast<cplx<float>*>(smem) + rr * pitch;
  const int x = row + p.sx;
  const bool xin = x >= 0 && x < a.orow;
  cplx<float> v[E];
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int c = L::din_pos(i, t), y = c + p.sy;
    const cplx<float> o = (xin && y >= 0 && y < a.ocol) ? a.obj[(size_t)x * a.ocol + y] : cplx<float>{0.f, 0.f};
    v[i] = cmul(a.probe[(size_t)row * N + c], o);
  }
  const typename L::TwF twf = L::template tw_load<-1>(a.tw, t);
  L::forward(v, my, alt, ph, twf, t);
  // slot A in column tiles (the layout of the CDI loop state): the column
  // kernel reads each tile contiguously
  cplx<float>* out = a.slotA + (size_t)blockIdx.y * N * N;
#pragma unroll
  for (int i = 0; i < E; i++) out[t1_index(row, L::dout_pos(i, t), N)] = v[i];
}

// column FFT, projection, inverse column FFT: slot A -> slot B
template <int N, int NT, int MODE>
__global__ void __launch_bounds__(NT) kp_col(PtArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const PtPos p = a.pos[blockIdx.y];
  const size_t so = (size_t)blockIdx.y * N * N;
  col_body<float, N, COL_FWD_PROJ_INV, MODE, 0, true>(a.slotA + so, a.slotB + so, N, a.tw,
                                                      a.Ipk + (size_t)p.scan * N * N, a.pa, 0.f, blockIdx.x,
                                                      plsh<N, NT>(), smem);
}

// forward column FFT only (forward model): slot A -> slot B
template <int N, int NT>
__global__ void __launch_bounds__(NT) kp_colfwd(PtArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const size_t so = (size_t)blockIdx.y * N * N;
  col_body<float, N, COL_FWD, 0, 0, true>(a.slotA + so, a.slotB + so, N, a.tw, nullptr, a.pa, 0.f, blockIdx.x,
                                          plsh<N, NT>(), smem);
}

// inverse row FFT of slot B, exit-wave difference, rPIE update, scatter
template <int N, int NT>
__global__ void __launch_bounds__(NT) kp_upd(PtArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  using L = LineOf<float, N>;
  constexpr int TL = LineCfg<N>::TL, E = L::E, LSH = plsh<N, NT>();
  const int pitch = lds_pitch<float, N>(LSH), alt = pitch << LSH;
  int ph = 0;
  const int t = threadIdx.x % TL, rr = threadIdx.x / TL;
  const int row = (blockIdx.x << LSH) + rr;
  const PtPos p = a.pos[blockIdx.y];
  cplx<float>* my = reinterpret_cast<cplx<float>*>(smem) + rr * pitch;
  const int x = row + p.sx;
  const bool xin = x >= 0 && x < a.orow;
  const cplx<float>* in = a.slotB + (size_t)blockIdx.y * N * N + (size_t)row * N;
  cplx<float> v[E], ov[E], pv[E];
#pragma unroll
  for (int i = 0; i < E; i++) v[i] = in[L::dout_pos(i, t)];
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int c = L::din_pos(i, t), y = c + p.sy;
    ov[i] = (xin && y >= 0 && y < a.ocol) ? a.obj[(size_t)x * a.ocol + y] : cplx<float>{0.f, 0.f};
    pv[i] = a.probe[(size_t)row * N + c];
  }
  const typename L::TwI twi = L::template tw_load<+1>(a.tw, t);
  L::inverse(v, my, alt, ph, twi, t);
  const float pmax = a.probeMaxDev ? *a.probeMaxDev : a.probeMax;
  cplx<float>* pt = a.updateProbe ? a.pterm + (size_t)blockIdx.y * N * N + (size_t)row * N : nullptr;
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int c = L::din_pos(i, t), y = c + p.sy;
    const cplx<float> psi = cmul(pv[i], ov[i]);
    const cplx<float> d = {psi.x - v[i].x * a.norm, psi.y - v[i].y * a.norm};
    cplx<float> o = ov[i];
    rpie(o, pv[i], d, pmax, kAlpha, a.objStep);
    if (a.updateProbe) {
      // updateObjectAndProbeStep (ptycho.cu:91-101): the probe term uses the updated object
      const cplx<float> q = rpie_step(o, d, a.objMax, kBeta);
      pt[c] = {q.x * a.probeStepSize, q.y * a.probeStepSize};
    }
    if (xin && y >= 0 && y < a.ocol) a.obj[(size_t)x * a.ocol + y] = o;
  }
}

// probeStep += term[0] + term[1] + ... in slot order
__global__ void kp_pstep(int64_t n, int nslots, const cplx<float>* __restrict__ term, cplx<float>* __restrict__ step) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    cplx<float> s = step[j];
    for (int k = 0; k < nslots; k++) {
      const cplx<float> q = term[(size_t)k * n + j];
      s = {s.x + q.x, s.y + q.y};
    }
    step[j] = s;
  }
}

// patterns[scan] = |F|^2 / N (getMod2 with norm 1/(row*column), ptycho.cpp:343)
template <int N>
__global__ void kp_mod2(const PtPos* __restrict__ pos, const cplx<float>* __restrict__ F, float* __restrict__ pat) {
  constexpr int64_t n = (int64_t)N * N;
  const PtPos p = pos[blockIdx.y];
  const cplx<float>* f = F + (size_t)blockIdx.y * n;
  float* o = pat + (size_t)p.scan * n;
  const float inv = 1.f / (float)n;
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x)
    o[j] = (f[j].x * f[j].x + f[j].y * f[j].y) * inv;
}

// packed pattern: NaN where bs > 0.5, the intensity elsewhere
__global__ void kp_pack(int64_t n, int64_t total, const float* __restrict__ I, const float* __restrict__ bs,
                        float* __restrict__ out) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x)
    out[j] = (bs && bs[j % n] > 0.5f) ? NAN : I[j];
}

// max |P|^2 (getMod2 + findMax, ptycho.cpp:479-480), one workgroup
__global__ void __launch_bounds__(1024) kp_probe_max(int64_t n, const cplx<float>* __restrict__ P, float* out) {
  float m = 0.f;
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) m = fmaxf(m, P[j].x * P[j].x + P[j].y * P[j].y);
  __shared__ float sh[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmaxf(m, sh[w]);
    *out = fmaxf(m, sh[0]);
  }
}

// probe += probeStep; probeStep = 0 (ptycho.cpp:563-566)
__global__ void kp_apply_step(int64_t n, cplx<float>* __restrict__ P, cplx<float>* __restrict__ step) {
  for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const cplx<float> s = step[j];
    P[j] = {P[j].x + s.x, P[j].y + s.y};
    step[j] = {0.f, 0.f};
  }
}

}  // namespace cdx

using namespace cdx;

struct cdx_ptycho {
  int n = 0, orow = 0, ocol = 0, nscan = 0;
  hipStream_t st = nullptr;
  bool own_stream = false;
  std::vector<int> sx, sy;
  void *slotA = nullptr, *slotB = nullptr, *pterm = nullptr, *Ipk = nullptr, *tw = nullptr;
  float* dmax = nullptr;  // device max |P|^2
  PtPos* dpos[2] = {nullptr, nullptr};
  PtPos* hpos[2] = {nullptr, nullptr};  // pinned staging of the level lists
  hipEvent_t staged[2] = {nullptr, nullptr};
  int cur = 0;
  bool have_patterns = false;
};

static thread_local char g_perr[512];
static int pfail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_perr, sizeof g_perr, fmt, ap);
  va_end(ap);
  set_last_error(g_perr);
  return code;
}
#define PCHK(expr)                                                                             \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return pfail(CDX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

static bool pt_size_ok(int n) { return n == 64 || n == 128 || n == 256 || n == 384 || n == 512; }

template <class K>
static hipError_t pt_launch(K kern, dim3 grid, int nt, size_t lds, hipStream_t st, const PtArgs& a) {
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, grid, dim3(nt), lds, st, a);
  return hipGetLastError();
}

// The three launches of one group of independent positions.  A group holds
// a few positions, so small workgroups fill the chip: the row kernels run one
// wave per workgroup (rows are contiguous at any width), the column kernel two
// waves (8 complex64 columns = 64-byte row segments for N = 128..384).
template <int N>
static hipError_t pt_group(const PtArgs& a, int k, int mode, hipStream_t st) {
  constexpr int RT = 64, CT = 128;
  const size_t lr = lds_bytes<float, N>(plsh<N, RT>()), lc = lds_bytes<float, N>(plsh<N, CT>());
  const dim3 gr(N >> plsh<N, RT>(), k), gc(N >> plsh<N, CT>(), k);
  hipError_t e = pt_launch(kp_fwd<N, RT>, gr, RT, lr, st, a);
  if (e == hipSuccess)
    e = mode == CDX_PT_NOISE ? pt_launch(kp_col<N, CT, PROJ_PT_NOISE>, gc, CT, lc, st, a)
                             : pt_launch(kp_col<N, CT, PROJ_PT_ACCURATE>, gc, CT, lc, st, a);
  if (e == hipSuccess) e = pt_launch(kp_upd<N, RT>, gr, RT, lr, st, a);
  return e;
}

template <int N>
static hipError_t pt_sim_group(const PtArgs& a, int k, float* pat, hipStream_t st) {
  constexpr int NT = 256;
  const size_t lds = lds_bytes<float, N>(plsh<N, NT>());
  const dim3 grid(N >> plsh<N, NT>(), k);
  hipError_t e = pt_launch(kp_fwd<N, NT>, grid, NT, lds, st, a);
  if (e == hipSuccess) e = pt_launch(kp_colfwd<N, NT>, grid, NT, lds, st, a);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(kp_mod2<N>, dim3(std::max(1, N * N / 4096), k), dim3(256), 0, st, a.pos,
                       (const cplx<float>*)a.slotB, pat);
    e = hipGetLastError();
  }
  return e;
}

static hipError_t pt_dispatch_group(int n, const PtArgs& a, int k, int mode, hipStream_t st) {
  switch (n) {
    case 64: return pt_group<64>(a, k, mode, st);
    case 128: return pt_group<128>(a, k, mode, st);
    case 256: return pt_group<256>(a, k, mode, st);
    case 384: return pt_group<384>(a, k, mode, st);
    case 512: return pt_group<512>(a, k, mode, st);
    default: return hipErrorInvalidValue;
  }
}
static hipError_t pt_dispatch_sim(int n, const PtArgs& a, int k, float* pat, hipStream_t st) {
  switch (n) {
    case 64: return pt_sim_group<64>(a, k, pat, st);
    case 128: return pt_sim_group<128>(a, k, pat, st);
    case 256: return pt_sim_group<256>(a, k, pat, st);
    case 384: return pt_sim_group<384>(a, k, pat, st);
    case 512: return pt_sim_group<512>(a, k, pat, st);
    default: return hipErrorInvalidValue;
  }
}

extern "C" int cdx_ptycho_destroy(cdx_ptycho* p) {
  if (!p) return CDX_OK;
  if (p->st) (void)hipStreamSynchronize(p->st);
  for (void* q : {p->slotA, p->slotB, p->pterm, p->Ipk, p->tw, (void*)p->dmax, (void*)p->dpos[0], (void*)p->dpos[1]})
    (void)hipFree(q);
  for (int b = 0; b < 2; b++) {
    if (p->hpos[b]) (void)hipHostFree(p->hpos[b]);
    if (p->staged[b]) (void)hipEventDestroy(p->staged[b]);
  }
  if (p->own_stream && p->st) (void)hipStreamDestroy(p->st);
  delete p;
  return CDX_OK;
}

extern "C" int cdx_ptycho_create(cdx_ptycho** out, int n, int obj_rows, int obj_cols, int nscan, const int* sx,
                                 const int* sy, void* stream) {
  if (!out) return pfail(CDX_ERR_ARG, "cdx_ptycho_create: out is NULL");
  *out = nullptr;
  if (!pt_size_ok(n)) return pfail(CDX_ERR_ARG, "cdx_ptycho_create: probe size %d (supported: 64, 128, 256, 384, 512)", n);
  if (obj_rows <= 0 || obj_cols <= 0 || nscan <= 0 || !sx || !sy)
    return pfail(CDX_ERR_ARG, "cdx_ptycho_create: bad object size, scan count or positions");
  cdx_ptycho* p = new cdx_ptycho();
  p->n = n;
  p->orow = obj_rows;
  p->ocol = obj_cols;
  p->nscan = nscan;
  p->sx.assign(sx, sx + nscan);
  p->sy.assign(sy, sy + nscan);
  if (stream) {
    p->st = (hipStream_t)stream;
  } else {
    if (hipStreamCreate(&p->st) != hipSuccess) {
      delete p;
      return pfail(CDX_ERR_HIP, "cdx_ptycho_create: hipStreamCreate failed");
    }
    p->own_stream = true;
  }
  const size_t nn = (size_t)n * n;
  const int slots = std::min(nscan, kPtSlots);
  int rc;
  if (hipMalloc(&p->slotA, slots * nn * 8) || hipMalloc(&p->slotB, slots * nn * 8) ||
      hipMalloc(&p->pterm, slots * nn * 8) || hipMalloc(&p->Ipk, (size_t)nscan * nn * 4) ||
      hipMalloc((void**)&p->dmax, 256) || hipMalloc((void**)&p->dpos[0], nscan * sizeof(PtPos)) ||
      hipMalloc((void**)&p->dpos[1], nscan * sizeof(PtPos)) ||
      hipHostMalloc((void**)&p->hpos[0], nscan * sizeof(PtPos)) ||
      hipHostMalloc((void**)&p->hpos[1], nscan * sizeof(PtPos)) || hipEventCreate(&p->staged[0]) ||
      hipEventCreate(&p->staged[1])) {
    cdx_ptycho_destroy(p);
    return pfail(CDX_ERR_HIP, "cdx_ptycho_create: out of memory (%d positions of %dx%d)", nscan, n, n);
  }
  if ((rc = make_twiddles(n, CDX_F32, &p->tw))) {
    cdx_ptycho_destroy(p);
    return rc;
  }
  *out = p;
  return CDX_OK;
}

extern "C" int cdx_ptycho_set_positions(cdx_ptycho* p, const int* sx, const int* sy) {
  if (!p || !sx || !sy) return pfail(CDX_ERR_ARG, "cdx_ptycho_set_positions: bad arguments");
  p->sx.assign(sx, sx + p->nscan);
  p->sy.assign(sy, sy + p->nscan);
  return CDX_OK;
}

extern "C" int cdx_ptycho_set_patterns(cdx_ptycho* p, const void* patterns, const void* bs) {
  if (!p || !patterns) return pfail(CDX_ERR_ARG, "cdx_ptycho_set_patterns: bad arguments");
  const int64_t nn = (int64_t)p->n * p->n, total = nn * p->nscan;
  hipLaunchKernelGGL(kp_pack, dim3((unsigned)std::min<int64_t>((total + 255) / 256, 8192)), dim3(256), 0, p->st, nn,
                     total, (const float*)patterns, (const float*)bs, (float*)p->Ipk);
  PCHK(hipGetLastError());
  PCHK(hipStreamSynchronize(p->st));  // the caller may release patterns and bs on return
  p->have_patterns = true;
  return CDX_OK;
}

extern "C" int cdx_ptycho_params_default(cdx_ptycho_params* q) {
  if (!q) return pfail(CDX_ERR_ARG, "cdx_ptycho_params_default: NULL");
  memset(q, 0, sizeof *q);
  q->mode = CDX_PT_ACCURATE;
  q->noiseLevel = 0;
  q->scale = 1.0;       // 1/exposure
  q->rcolor = 65535.0;  // format.hpp:21
  q->probeMax = -1;     // computed on the device
  q->objMax = 0;
  q->objStep = 1.0;
  q->probeStepSize = 0;
  q->updateProbe = 0;
  q->maxConcurrent = 0;
  return CDX_OK;
}

// Level lists of a sweep: positions grouped by level (stable within a level),
// each level split into groups of at most cap positions.  Returns the groups'
// sizes; fills list[] with the positions in launch order.
static std::vector<int> pt_levels(int n, int orow, int ocol, const int* sx, const int* sy, const int* order, int m,
                                  int cap, PtPos* list) {
  struct Rect {
    int x0, x1, y0, y1;
  };
  std::vector<Rect> r(m);
  std::vector<int> lvl(m, 0);
  int nl = 0;
  // Clipped windows are at most n wide and start inside the object, so two
  // that overlap start less than n apart in both directions: bucket the
  // earlier positions by start / n and look only at the 3 x 3 neighbouring
  // cells.
  const int gx = orow / n + 1, gy = ocol / n + 1;
  std::vector<std::vector<int>> cell((size_t)gx * gy);
  for (int j = 0; j < m; j++) {
    const int s = order[j];
    r[j] = {std::max(sx[s], 0), std::min(sx[s] + n, orow), std::max(sy[s], 0),
            std::min(sy[s] + n, ocol)};
    int l = 0;
    if (r[j].x0 < r[j].x1 && r[j].y0 < r[j].y1) {
      const int cx = r[j].x0 / n, cy = r[j].y0 / n;
      for (int ax = std::max(cx - 1, 0); ax <= std::min(cx + 1, gx - 1); ax++)
        for (int ay = std::max(cy - 1, 0); ay <= std::min(cy + 1, gy - 1); ay++)
          for (int i : cell[(size_t)ax * gy + ay]) {
            if (lvl[i] + 1 > l && r[i].x0 < r[j].x1 && r[j].x0 < r[i].x1 && r[i].y0 < r[j].y1 &&
                r[j].y0 < r[i].y1)
              l = lvl[i] + 1;
          }
      cell[(size_t)cx * gy + cy].push_back(j);
    }
    lvl[j] = l;
    nl = std::max(nl, l + 1);
  }
  std::vector<std::vector<int>> by(nl);
  for (int j = 0; j < m; j++) by[lvl[j]].push_back(j);
  std::vector<int> groups;
  int w = 0;
  for (auto& v : by) {
    for (size_t a = 0; a < v.size(); a += cap) {
      const int k = (int)std::min<size_t>(cap, v.size() - a);
      for (int q = 0; q < k; q++) {
        const int s = order[v[a + q]];
        list[w++] = {s, sx[s], sy[s], 0};
      }
      groups.push_back(k);
    }
  }
  return groups;
}

// next of the two (pinned host, device) position-list buffers: waits for the
// copy that last read the host buffer; the caller fills it, copies it and
// records staged[b] after the copy
static int pt_stage(cdx_ptycho* p, int* b) {
  *b = p->cur;
  p->cur ^= 1;
  PCHK(hipEventSynchronize(p->staged[*b]));
  return CDX_OK;
}

static PtArgs pt_args(cdx_ptycho* p, void* obj, const void* probe) {
  PtArgs a{};
  a.obj = (cplx<float>*)obj;
  a.orow = p->orow;
  a.ocol = p->ocol;
  a.probe = (const cplx<float>*)probe;
  a.slotA = (cplx<float>*)p->slotA;
  a.slotB = (cplx<float>*)p->slotB;
  a.Ipk = (const float*)p->Ipk;
  a.tw = (const cplx<float>*)p->tw;
  a.pterm = (cplx<float>*)p->pterm;
  a.norm = (float)(1.0 / sqrt((double)p->n * p->n));
  return a;
}

extern "C" int cdx_ptycho_sweep(cdx_ptycho* p, const int* order, int norder, const cdx_ptycho_params* q,
                                void* object, const void* probe, void* probe_step, int* n_groups) {
  if (!p || !q || !object || !probe) return pfail(CDX_ERR_ARG, "cdx_ptycho_sweep: bad arguments");
  if (!p->have_patterns) return pfail(CDX_ERR_ARG, "cdx_ptycho_sweep: patterns not set");
  if (q->mode != CDX_PT_ACCURATE && q->mode != CDX_PT_NOISE)
    return pfail(CDX_ERR_ARG, "cdx_ptycho_sweep: mode %d (CDX_PT_ACCURATE or CDX_PT_NOISE)", q->mode);
  if (q->updateProbe && !probe_step)
    return pfail(CDX_ERR_ARG, "cdx_ptycho_sweep: updateProbe needs a probe_step buffer");
  const int m = order ? norder : p->nscan;
  if (m < 0 || m > p->nscan) return pfail(CDX_ERR_ARG, "cdx_ptycho_sweep: %d positions (at most %d)", m, p->nscan);
  std::vector<int> ident;
  if (!order) {
    ident.resize(m);
    for (int i = 0; i < m; i++) ident[i] = i;
    order = ident.data();
  }
  for (int i = 0; i < m; i++)
    if (order[i] < 0 || order[i] >= p->nscan)
      return pfail(CDX_ERR_ARG, "cdx_ptycho_sweep: order[%d] = %d out of range", i, order[i]);
  const int slots = std::min(p->nscan, kPtSlots);
  const int cap = q->maxConcurrent > 0 ? std::min(q->maxConcurrent, slots) : slots;
  int b;
  int rc = pt_stage(p, &b);
  if (rc) return rc;
  PtPos* dev = p->dpos[b];
  const std::vector<int> groups = pt_levels(p->n, p->orow, p->ocol, p->sx.data(), p->sy.data(), order, m, cap,
                                            p->hpos[b]);
  PCHK(hipMemcpyAsync(dev, p->hpos[b], m * sizeof(PtPos), hipMemcpyHostToDevice, p->st));
  PCHK(hipEventRecord(p->staged[b], p->st));
  PtArgs a = pt_args(p, object, probe);
  a.pa.mode = q->mode == CDX_PT_NOISE ? PROJ_PT_NOISE : PROJ_PT_ACCURATE;
  a.pa.noiseLevel = q->noiseLevel;
  a.pa.scale = (float)q->scale;
  a.pa.rcolor = (float)q->rcolor;
  a.pa.bsnorm = a.norm;
  a.probeMax = (float)q->probeMax;
  a.objMax = (float)q->objMax;
  a.objStep = (float)q->objStep;
  a.probeStepSize = (float)q->probeStepSize;
  a.updateProbe = q->updateProbe ? 1 : 0;
  const int64_t nn = (int64_t)p->n * p->n;
  if (q->probeMax < 0) {
    hipLaunchKernelGGL(kp_probe_max, dim3(1), dim3(1024), 0, p->st, nn, (const cplx<float>*)probe, p->dmax);
    PCHK(hipGetLastError());
    a.probeMaxDev = p->dmax;
  }
  int off = 0;
  for (int k : groups) {
    a.pos = dev + off;
    PCHK(pt_dispatch_group(p->n, a, k, q->mode, p->st));
    if (a.updateProbe) {
      hipLaunchKernelGGL(kp_pstep, dim3((unsigned)std::min<int64_t>((nn + 255) / 256, 2048)), dim3(256), 0, p->st,
                         nn, k, (const cplx<float>*)p->pterm, (cplx<float>*)probe_step);
      PCHK(hipGetLastError());
    }
    off += k;
  }
  if (n_groups) *n_groups = (int)groups.size();
  return CDX_OK;
}

extern "C" int cdx_ptycho_levels(int n, int obj_rows, int obj_cols, int nscan, const int* sx, const int* sy,
                                 const int* order, int norder, int max_concurrent, int* groups, int* list,
                                 int* n_groups) {
  if (n <= 0 || obj_rows <= 0 || obj_cols <= 0 || nscan <= 0 || !sx || !sy || !groups || !list || !n_groups)
    return pfail(CDX_ERR_ARG, "cdx_ptycho_levels: bad arguments");
  const int m = order ? norder : nscan;
  if (m < 0 || m > nscan) return pfail(CDX_ERR_ARG, "cdx_ptycho_levels: %d positions (at most %d)", m, nscan);
  std::vector<int> ident;
  if (!order) {
    ident.resize(m);
    for (int i = 0; i < m; i++) ident[i] = i;
    order = ident.data();
  }
  for (int i = 0; i < m; i++)
    if (order[i] < 0 || order[i] >= nscan)
      return pfail(CDX_ERR_ARG, "cdx_ptycho_levels: order[%d] = %d out of range", i, order[i]);
  const int cap = max_concurrent > 0 ? max_concurrent : kPtSlots;
  std::vector<PtPos> pl(m > 0 ? m : 1);
  const std::vector<int> g = pt_levels(n, obj_rows, obj_cols, sx, sy, order, m, cap, pl.data());
  for (size_t i = 0; i < g.size(); i++) groups[i] = g[i];
  for (int i = 0; i < m; i++) list[i] = pl[i].scan;
  *n_groups = (int)g.size();
  return CDX_OK;
}

extern "C" int cdx_ptycho_simulate(cdx_ptycho* p, const void* object, const void* probe, void* patterns) {
  if (!p || !object || !probe || !patterns) return pfail(CDX_ERR_ARG, "cdx_ptycho_simulate: bad arguments");
  int b;
  int rc = pt_stage(p, &b);
  if (rc) return rc;
  PtPos* dev = p->dpos[b];
  for (int s = 0; s < p->nscan; s++) p->hpos[b][s] = {s, p->sx[s], p->sy[s], 0};
  PCHK(hipMemcpyAsync(dev, p->hpos[b], p->nscan * sizeof(PtPos), hipMemcpyHostToDevice, p->st));
  PCHK(hipEventRecord(p->staged[b], p->st));
  PtArgs a = pt_args(p, (void*)object, probe);
  const int slots = std::min(p->nscan, kPtSlots);
  for (int off = 0; off < p->nscan; off += slots) {
    const int k = std::min(slots, p->nscan - off);
    a.pos = dev + off;
    PCHK(pt_dispatch_sim(p->n, a, k, (float*)patterns, p->st));
  }
  return CDX_OK;
}

extern "C" int cdx_ptycho_apply_probe_step(cdx_ptycho* p, void* probe, void* probe_step) {
  if (!p || !probe || !probe_step) return pfail(CDX_ERR_ARG, "cdx_ptycho_apply_probe_step: bad arguments");
  const int64_t nn = (int64_t)p->n * p->n;
  hipLaunchKernelGGL(kp_apply_step, dim3((unsigned)std::min<int64_t>((nn + 255) / 256, 2048)), dim3(256), 0, p->st,
                     nn, (cplx<float>*)probe, (cplx<float>*)probe_step);
  PCHK(hipGetLastError());
  return CDX_OK;
}

extern "C" int cdx_ptycho_synchronize(cdx_ptycho* p) {
  if (!p) return pfail(CDX_ERR_ARG, "cdx_ptycho_synchronize: NULL");
  PCHK(hipStreamSynchronize(p->st));
  return CDX_OK;
}

extern "C" int cdx_ptycho_time_sweeps(cdx_ptycho* p, const int* order, int norder, const cdx_ptycho_params* q,
                                      void* object, const void* probe, void* probe_step, int n_sweeps,
                                      double* ms_total, int* n_groups) {
  if (!p || !ms_total || n_sweeps <= 0) return pfail(CDX_ERR_ARG, "cdx_ptycho_time_sweeps: bad arguments");
  int rc = cdx_ptycho_sweep(p, order, norder, q, object, probe, probe_step, n_groups);  // untimed first sweep
  if (rc) return rc;
  hipEvent_t e0, e1;
  PCHK(hipEventCreate(&e0));
  PCHK(hipEventCreate(&e1));
  PCHK(hipStreamSynchronize(p->st));
  PCHK(hipEventRecord(e0, p->st));
  for (int i = 0; i < n_sweeps && !rc; i++) rc = cdx_ptycho_sweep(p, order, norder, q, object, probe, probe_step, nullptr);
  PCHK(hipEventRecord(e1, p->st));
  PCHK(hipEventSynchronize(e1));
  float ms = 0;
  PCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms_total = ms;
  return rc;
}
