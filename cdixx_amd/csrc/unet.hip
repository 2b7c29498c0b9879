// unet.hip -- learned warm start for the phase-retrieval loop (no counterpart
// in the reference, whose loop always starts from random signs,
// src/post_gpu/cdi.cc:222-223).
//
// A 2D U-Net on the centred log-intensity predicts the object amplitude, its
// phase and a support logit.  Everything runs on hand-written kernels:
//   * 3x3 / 1x1 convolutions as implicit GEMMs on BF16 MFMA
//     (v_mfma_f32_32x32x16_bf16, FP32 accumulation), NHWC activations;
//   * input gradients as the same convolution with flipped, transposed
//     weights; weight gradients as a split-K GEMM over pixels with FP32
//     atomics; bias gradients, ReLU/pool/upsample/concat backward;
//   * the loss, Adam (FP32 master weights, BF16 copies for the MFMA);
//   * training pairs from the project's own GPU forward model (random
//     ellipse / polygon supports, smoothed-noise amplitude and phase, Poisson
//     intensities).
// Layout of a conv layer's weights: w[co][k], k = tap * Cin + ci (tap = 3 dy
// + dx for 3x3), zero-padded to Kp = roundup(taps * Cin, 32) per row.
#include "../../include/cdx_abi.h"
#include "device_ops.hpp"
#include "errors.hpp"
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

using namespace cdx;

namespace {

int ufail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  cdx::set_last_error(buf);
  return code;
}
#define UCHK(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return ufail(CDX_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
#define RCHK(expr)          \
  do {                      \
    int r_ = (expr);        \
    if (r_) return r_;      \
  } while (0)

using u16 = unsigned short;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__host__ __device__ inline float bf2f(u16 h) {
  const unsigned u = (unsigned)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
// round to nearest even
__host__ __device__ inline u16 f2bf(float f) {
  unsigned u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}

constexpr int kBK = 32;  // K padding of the weight rows (multiple of the conv's 16-wide MFMA steps)
// workspace of the weight gradient's pixel splits, in floats
constexpr int64_t kWgradPartFloats = (int64_t)6 << 20;

// ---------------------------------------------------------------------------
// convolution: out[m][co] = act(bias[co] + sum_k A[m][k] W[co][k]),
// A[m][k] = x[b, h + dy, w + dx, ci] (zero outside the image)
// ---------------------------------------------------------------------------
struct ConvArgs {
  const u16* x;       // [B][H][W][Cin]
  const u16* w;       // [Cout][Kp]
  const float* bias;  // [Cout] or null
  u16* y;             // [B][H][W][Cout] or null
  float* yf;          // FP32 copy of the output or null
  int B, H, W, Cin, Cout, Kp, taps, relu;
  const u16* mask = nullptr;  // y is zeroed where mask <= 0 (a ReLU output's gradient), or null
};

__device__ __forceinline__ void pixel_of(int m, int HW, int W, int& b, int& h, int& w) {
  b = m / HW;
  const int r = m - b * HW;
  h = r / W;
  w = r - h * W;
}

// 8 consecutive k (one tap, 8 consecutive input channels) of pixel (b,h,w)
__device__ __forceinline__ uint4 load_patch8(const u16* x, int H, int W, int Cin, int taps, int b, int h, int w,
                                             int k) {
  const int tap = k / Cin, ci = k - tap * Cin;
  int dy = 0, dx = 0;
  if (taps == 9) {
    dy = tap / 3 - 1;
    dx = tap - (tap / 3) * 3 - 1;
  }
  const int hh = h + dy, ww = w + dx;
  if (hh < 0 || hh >= H || ww < 0 || ww >= W) return make_uint4(0, 0, 0, 0);
  return *(const uint4*)(x + ((size_t)(b * H + hh) * W + ww) * Cin + ci);
}

// 8 BF16 values of a 16-byte piece <-> floats
__device__ __forceinline__ void unpack8(const uint4& v, float* f) {
  const unsigned* w = (const unsigned*)&v;
#pragma unroll
  for (int e = 0; e < 8; e++) f[e] = bf2f((u16)((w[e >> 1] >> ((e & 1) * 16)) & 0xFFFFu));
}
__device__ __forceinline__ uint4 pack8(const float* f) {
  uint4 v;
  unsigned* w = (unsigned*)&v;
#pragma unroll
  for (int d = 0; d < 4; d++) w[d] = (unsigned)f2bf(f[2 * d]) | ((unsigned)f2bf(f[2 * d + 1]) << 16);
  return v;
}
// zero where the ReLU output y is not positive
__device__ __forceinline__ void relu_mask8(float* f, const uint4& y) {
  float m[8];
  unpack8(y, m);
#pragma unroll
  for (int e = 0; e < 8; e++)
    if (!(m[e] > 0.f)) f[e] = 0.f;
}
// row of accumulator element r in the 32x32 block (column = lane & 31)
__device__ __forceinline__ int acc_row(int r, int lane) { return (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3); }

// Implicit-GEMM convolution on a TM x TN output tile: four waves in a WM x WN
// grid, each computing BM x BN blocks of 32 x 32 (TM = 32 WM BM pixels, TN =
// 32 WN BN channels), so a wave reads BM + BN LDS operands per BM BN MFMAs.
// K in steps of BK through two LDS buffers, the next step's global loads in
// flight during the current step's MFMAs, one barrier per step.
template <int WM, int WN, int BM, int BN, int BK>
__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  constexpr int TM = 32 * WM * BM, TN = 32 * WN * BN, LP = BK + 8;
  constexpr int KC = BK / 8;                       // 16-byte pieces per row and step
  constexpr int AP = TM * KC, BP = TN * KC;        // pieces per step
  constexpr int AV = (AP + 255) / 256, BV = (BP + 255) / 256;
  static_assert(WM * WN == 4 && BK % 16 == 0, "four waves, 16-wide MFMA steps");
  // the two operand buffers, and after the K loop the epilogue's staging
  // area (a 32 x 33 float block per wave) in the same LDS
  constexpr size_t kOps = (size_t)2 * (TM + TN) * LP * sizeof(u16), kStage = (size_t)4 * 32 * 33 * sizeof(float);
  __shared__ __align__(16) unsigned char smem[kOps > kStage ? kOps : kStage];
  u16(*As)[TM][LP] = reinterpret_cast<u16(*)[TM][LP]>(smem);
  u16(*Bs)[TN][LP] = reinterpret_cast<u16(*)[TN][LP]>(smem + (size_t)2 * TM * LP * sizeof(u16));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int HW = a.H * a.W, M = a.B * HW, K = a.taps * a.Cin;
  const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
  // this thread's A pieces: row (pixel) and k offset within the step; the
  // pixel's element offset in x
  int arow[AV], akc[AV], ab[AV], ah[AV], aw[AV];
  size_t apix[AV];
  bool aok[AV], aon[AV];
#pragma unroll
  for (int j = 0; j < AV; j++) {
    const int idx = tid + 256 * j;
    aon[j] = idx < AP;
    arow[j] = aon[j] ? idx / KC : 0;
    akc[j] = (idx % KC) * 8;
    const int m = m0 + arow[j];
    aok[j] = aon[j] && m < M;
    ab[j] = ah[j] = aw[j] = 0;
    if (aok[j]) pixel_of(m, HW, a.W, ab[j], ah[j], aw[j]);
    apix[j] = ((size_t)(ab[j] * a.H + ah[j]) * a.W + aw[j]) * a.Cin;
  }
  // When BK divides Cin every step lies in one tap: the tap and its pixel
  // offset are the same for the whole workgroup (no per-piece division)
  const bool one_tap = a.Cin % BK == 0;
  int brow[BV], bkc[BV];
  bool bok[BV], bon[BV];
#pragma unroll
  for (int j = 0; j < BV; j++) {
    const int idx = tid + 256 * j;
    bon[j] = idx < BP;
    brow[j] = bon[j] ? idx / KC : 0;
    bkc[j] = (idx % KC) * 8;
    bok[j] = bon[j] && n0 + brow[j] < a.Cout;
  }
  uint4 ra[AV], rb[BV];
  auto load = [&](int k0) {
    if (one_tap) {
      const int tap = k0 / a.Cin, ci0 = k0 - tap * a.Cin;
      int dy = 0, dx = 0;
      if (a.taps == 9) {
        dy = tap / 3 - 1;
        dx = tap - (tap / 3) * 3 - 1;
      }
      const long long off = ((long long)dy * a.W + dx) * a.Cin + ci0;
      const bool kin = k0 < K;
#pragma unroll
      for (int j = 0; j < AV; j++) {
        const int hh = ah[j] + dy, ww = aw[j] + dx;
        const bool in = aok[j] && kin && hh >= 0 && hh < a.H && ww >= 0 && ww < a.W;
        ra[j] = in ? *(const uint4*)(a.x + (long long)apix[j] + off + akc[j]) : make_uint4(0, 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int j = 0; j < AV; j++) {
        const int k = k0 + akc[j];
        ra[j] = (aok[j] && k < K) ? load_patch8(a.x, a.H, a.W, a.Cin, a.taps, ab[j], ah[j], aw[j], k)
                                  : make_uint4(0, 0, 0, 0);
      }
    }
#pragma unroll
    for (int j = 0; j < BV; j++) {
      const int k = k0 + bkc[j];
      rb[j] = (bok[j] && k < a.Kp) ? *(const uint4*)(a.w + (size_t)(n0 + brow[j]) * a.Kp + k) : make_uint4(0, 0, 0, 0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < AV; j++)
      if (aon[j]) *(uint4*)&As[buf][arow[j]][akc[j]] = ra[j];
#pragma unroll
    for (int j = 0; j < BV; j++)
      if (bon[j]) *(uint4*)&Bs[buf][brow[j]][bkc[j]] = rb[j];
  };
  f32x16 acc[BM][BN];
#pragma unroll
  for (int i = 0; i < BM; i++)
#pragma unroll
    for (int j = 0; j < BN; j++) acc[i][j] = f32x16{};
  load(0);
  store(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < a.Kp; k0 += BK) {
    const bool more = k0 + BK < a.Kp;
    if (more) load(k0 + BK);
#pragma unroll
    for (int s4 = 0; s4 < BK / 16; s4++) {
      const int kk = s4 * 16 + (lane >> 5) * 8;
      bf16x8 av[BM], bv[BN];
#pragma unroll
      for (int i = 0; i < BM; i++) av[i] = *(const bf16x8*)&As[buf][(wm * BM + i) * 32 + (lane & 31)][kk];
#pragma unroll
      for (int j = 0; j < BN; j++) bv[j] = *(const bf16x8*)&Bs[buf][(wn * BN + j) * 32 + (lane & 31)][kk];
#pragma unroll
      for (int i = 0; i < BM; i++)
#pragma unroll
        for (int j = 0; j < BN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
    // the other buffer was last read in the previous step, before its barrier
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // Epilogue, one 32 x 32 block at a time: the wave stages its accumulators
  // in LDS (element (row, col) at row * 33 + col), then lane l takes pixel
  // row l / 2 and channels 16 (l % 2) .. +15 as two 8-channel segments, so
  // bias, ReLU, mask and rounding run per 8 channels and y moves in 16-byte
  // stores (channel counts are multiples of 8: a segment is all in or out).
  float* stg = reinterpret_cast<float*>(smem) + wave * 32 * 33;
  const int srow = lane >> 1, sc0 = (lane & 1) * 16;
#pragma unroll
  for (int i = 0; i < BM; i++) {
#pragma unroll
    for (int j = 0; j < BN; j++) {
      __syncthreads();  // the K loop's or the previous block's LDS reads are done
#pragma unroll
      for (int r = 0; r < 16; r++) stg[acc_row(r, lane) * 33 + (lane & 31)] = acc[i][j][r];
      __syncthreads();
      const int mm = m0 + (wm * BM + i) * 32 + srow;
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        const int c = sc0 + 8 * sg, coo = n0 + (wn * BN + j) * 32 + c;
        if (mm >= M || coo >= a.Cout) continue;
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = stg[srow * 33 + c + e] + (a.bias ? a.bias[coo + e] : 0.f);
        if (a.relu)
#pragma unroll
          for (int e = 0; e < 8; e++) v[e] = fmaxf(v[e], 0.f);
        const size_t o = (size_t)mm * a.Cout + coo;
        if (a.mask) relu_mask8(v, *(const uint4*)(a.mask + o));
        if (a.y) *(uint4*)(a.y + o) = pack8(v);
        if (a.yf) {
          *(float4*)(a.yf + o) = make_float4(v[0], v[1], v[2], v[3]);
          *(float4*)(a.yf + o + 4) = make_float4(v[4], v[5], v[6], v[7]);
        }
      }
    }
  }
}
// 3x3 convolution on a block of TH x TW pixels of one image (TM = TH TW =
// 32 WM BM pixels) x TN = 32 WN BN output channels, for Cin a multiple of 32
// and sides divisible by the block.  The input channels go in chunks of 32:
// per chunk the (TH + 2) x (TW + 2) halo of the block (zero outside the
// image) and the chunk's 9 x 32 weight columns of the TN channels are staged
// in LDS once, and the 9 taps x 2 MFMA k-steps read shifted windows of the
// halo at compile-time offsets (k_conv instead gathers every tap's im2col
// rows from global memory, one K step of 32 at a time).  The next chunk's
// pieces are loaded into registers during the current chunk's MFMAs.  LDS
// pitches: 40 bf16 per halo pixel and 9 x 32 + 8 per weight row, so the 32
// rows of an MFMA operand read spread over all banks.  Same epilogue and
// arguments as k_conv (taps = 9, Kp = 9 Cin).
template <int WM, int WN, int BM, int BN, int TW>
__global__ void __launch_bounds__(256) k_conv3_halo(ConvArgs a) {
  constexpr int TM = 32 * WM * BM, TN = 32 * WN * BN, TH = TM / TW;
  constexpr int HW2 = TW + 2, HPIX = (TH + 2) * HW2;  // halo pixels
  constexpr int HP = 40, WP = 9 * 32 + 8;            // LDS pitches in bf16
  constexpr int HPC = HPIX * 4, WPC = TN * 36;        // 16-byte pieces per chunk
  constexpr int HV = (HPC + 255) / 256, WV = (WPC + 255) / 256;
  static_assert(WM * WN == 4 && TM % TW == 0, "four waves, whole pixel rows");
  constexpr size_t kOps = ((size_t)HPIX * HP + (size_t)TN * WP) * sizeof(u16),
                   kStage = (size_t)4 * 32 * 33 * sizeof(float);
  __shared__ __align__(16) unsigned char smem[kOps > kStage ? kOps : kStage];
  u16(*Hs)[HP] = reinterpret_cast<u16(*)[HP]>(smem);
  u16(*Ws)[WP] = reinterpret_cast<u16(*)[WP]>(smem + (size_t)HPIX * HP * sizeof(u16));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int nwb = a.W / TW, nhb = a.H / TH;
  const int bx = blockIdx.x, wb = bx % nwb, hb = (bx / nwb) % nhb, b = bx / (nwb * nhb);
  const int h0 = hb * TH, w0 = wb * TW, n0 = blockIdx.y * TN;

  uint4 rh[HV], rw[WV];
  auto load = [&](int ci0) {
#pragma unroll
    for (int j = 0; j < HV; j++) {
      const int q = tid + 256 * j;
      rh[j] = make_uint4(0, 0, 0, 0);
      if (q < HPC) {
        const int hp = q >> 2, part = q & 3;
        const int hh = h0 + hp / HW2 - 1, ww = w0 + hp % HW2 - 1;
        if (hh >= 0 && hh < a.H && ww >= 0 && ww < a.W)
          rh[j] = *(const uint4*)(a.x + ((size_t)(b * a.H + hh) * a.W + ww) * a.Cin + ci0 + part * 8);
      }
    }
#pragma unroll
    for (int j = 0; j < WV; j++) {
      const int q = tid + 256 * j;
      rw[j] = make_uint4(0, 0, 0, 0);
      if (q < WPC) {
        const int co = q / 36, rem = q - co * 36, tap = rem >> 2, part = rem & 3;
        if (n0 + co < a.Cout)
          rw[j] = *(const uint4*)(a.w + (size_t)(n0 + co) * a.Kp + tap * a.Cin + ci0 + part * 8);
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int j = 0; j < HV; j++) {
      const int q = tid + 256 * j;
      if (q < HPC) *(uint4*)&Hs[q >> 2][(q & 3) * 8] = rh[j];
    }
#pragma unroll
    for (int j = 0; j < WV; j++) {
      const int q = tid + 256 * j;
      if (q < WPC) {
        const int co = q / 36, rem = q - co * 36;
        *(uint4*)&Ws[co][(rem >> 2) * 32 + (rem & 3) * 8] = rw[j];
      }
    }
  };
  // halo pixel of this lane's pixel row in block i (tap (0, 0))
  int hbase[BM];
#pragma unroll
  for (int i = 0; i < BM; i++) {
    const int q = (wm * BM + i) * 32 + (lane & 31);
    hbase[i] = (q / TW) * HW2 + q % TW;
  }
  const int kl = (lane >> 5) * 8;
  f32x16 acc[BM][BN];
#pragma unroll
  for (int i = 0; i < BM; i++)
#pragma unroll
    for (int j = 0; j < BN; j++) acc[i][j] = f32x16{};
  load(0);
  store();
  __syncthreads();
  for (int ci0 = 0; ci0 < a.Cin; ci0 += 32) {
    const bool more = ci0 + 32 < a.Cin;
    if (more) load(ci0 + 32);
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
      const int toff = (tap / 3) * HW2 + tap % 3;
#pragma unroll
      for (int k16 = 0; k16 < 2; k16++) {
        const int kk = k16 * 16 + kl;
        bf16x8 av[BM], bv[BN];
#pragma unroll
        for (int i = 0; i < BM; i++) av[i] = *(const bf16x8*)&Hs[hbase[i] + toff][kk];
#pragma unroll
        for (int j = 0; j < BN; j++) bv[j] = *(const bf16x8*)&Ws[(wn * BN + j) * 32 + (lane & 31)][tap * 32 + kk];
#pragma unroll
        for (int i = 0; i < BM; i++)
#pragma unroll
          for (int j = 0; j < BN; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
      }
    }
    if (more) {
      __syncthreads();  // this chunk's operands are read
      store();
      __syncthreads();
    }
  }
  // epilogue as k_conv's, with the block's pixel rows mapped to the image
  float* stg = reinterpret_cast<float*>(smem) + wave * 32 * 33;
  const int srow = lane >> 1, sc0 = (lane & 1) * 16;
#pragma unroll
  for (int i = 0; i < BM; i++) {
#pragma unroll
    for (int j = 0; j < BN; j++) {
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 16; r++) stg[acc_row(r, lane) * 33 + (lane & 31)] = acc[i][j][r];
      __syncthreads();
      const int q = (wm * BM + i) * 32 + srow;
      const int mm = (b * a.H + h0 + q / TW) * a.W + w0 + q % TW;
#pragma unroll
      for (int sg = 0; sg < 2; sg++) {
        const int c = sc0 + 8 * sg, coo = n0 + (wn * BN + j) * 32 + c;
        if (coo >= a.Cout) continue;
        float v[8];
#pragma unroll
        for (int e = 0; e < 8; e++) v[e] = stg[srow * 33 + c + e] + (a.bias ? a.bias[coo + e] : 0.f);
        if (a.relu)
#pragma unroll
          for (int e = 0; e < 8; e++) v[e] = fmaxf(v[e], 0.f);
        const size_t o = (size_t)mm * a.Cout + coo;
        if (a.mask) relu_mask8(v, *(const uint4*)(a.mask + o));
        if (a.y) *(uint4*)(a.y + o) = pack8(v);
        if (a.yf) {
          *(float4*)(a.yf + o) = make_float4(v[0], v[1], v[2], v[3]);
          *(float4*)(a.yf + o + 4) = make_float4(v[4], v[5], v[6], v[7]);
        }
      }
    }
  }
}

// The halo kernel's block for a layer, or false when it does not apply
// (1x1 taps, Cin not a multiple of 32, sides not divisible by the block):
// TN = 64 when the channels divide by 64, else 32; the largest block of
// TM = 256 / 128 / 64 pixels (16 wide) that still gives about 1024
// workgroups.
bool launch_conv_halo(const ConvArgs& a, int cout, hipStream_t st) {
  if (a.taps != 9 || a.Cin % 32 || a.Kp != 9 * a.Cin || a.W % 16 || cout % 32) return false;
  const long long px = (long long)a.B * a.H * a.W;
  const dim3 blk(256);
  auto blocks = [&](int tm) { return (unsigned)(px / tm); };
  if (cout % 64 == 0) {
    const unsigned ny = (unsigned)(cout / 64);
    if (a.H % 8 == 0 && blocks(128) * ny >= 1024) {
      hipLaunchKernelGGL((k_conv3_halo<2, 2, 2, 1, 16>), dim3(blocks(128), ny), blk, 0, st, a);
      return true;
    }
    if (a.H % 4 == 0) {
      hipLaunchKernelGGL((k_conv3_halo<2, 2, 1, 1, 16>), dim3(blocks(64), ny), blk, 0, st, a);
      return true;
    }
    return false;
  }
  const unsigned ny = (unsigned)(cout / 32);
  if (a.H % 16 == 0 && blocks(256) * ny >= 1024) {
    hipLaunchKernelGGL((k_conv3_halo<4, 1, 2, 1, 16>), dim3(blocks(256), ny), blk, 0, st, a);
    return true;
  }
  if (a.H % 8 == 0) {
    hipLaunchKernelGGL((k_conv3_halo<4, 1, 1, 1, 16>), dim3(blocks(128), ny), blk, 0, st, a);
    return true;
  }
  return false;
}

// The conv tile for `cout` output channels: the largest of 256 x 32 / 128 x 32
// (at most 32 channels: no idle MFMA columns), 128 x 128, 128 x 64 and 64 x 64
// that still gives about 1024 workgroups (four per CU).  3x3 layers with
// 32-channel chunks run the halo kernel.
void launch_conv(const ConvArgs& a, int cout, hipStream_t st) {
  if (launch_conv_halo(a, cout, st)) return;
  const long long M = (long long)a.B * a.H * a.W;
  auto wgs = [&](int tm, int tn) { return ((M + tm - 1) / tm) * ((cout + tn - 1) / tn); };
  const dim3 blk(256);
  if (cout <= 32) {
    if (wgs(256, 32) >= 1024)
      hipLaunchKernelGGL((k_conv<4, 1, 2, 1, 32>), dim3((unsigned)((M + 255) / 256), 1), blk, 0, st, a);
    else
      hipLaunchKernelGGL((k_conv<4, 1, 1, 1, 32>), dim3((unsigned)((M + 127) / 128), 1), blk, 0, st, a);
  } else if (cout % 128 == 0 && wgs(128, 128) >= 1024) {
    hipLaunchKernelGGL((k_conv<2, 2, 2, 2, 32>), dim3((unsigned)((M + 127) / 128), cout / 128), blk, 0, st, a);
  } else if (wgs(128, 64) >= 1024) {
    hipLaunchKernelGGL((k_conv<2, 2, 2, 1, 32>), dim3((unsigned)((M + 127) / 128), (cout + 63) / 64), blk, 0, st, a);
  } else {
    hipLaunchKernelGGL((k_conv<2, 2, 1, 1, 32>), dim3((unsigned)((M + 63) / 64), (cout + 63) / 64), blk, 0, st, a);
  }
}

// ---------------------------------------------------------------------------
// weight gradient: dw[co][n] = sum_m dz[m][co] A[m][n] (A the im2col rows of
// the layer input) and db[co] = sum_m dz[m][co].  grid.z splits the pixels;
// each split writes its partial tile to a workspace and k_wgrad_reduce adds
// the splits in a fixed order, so the gradient is the same on every run.
// ---------------------------------------------------------------------------
struct WgradArgs {
  const u16* dz;  // [M][Cout]
  const u16* x;   // [B][H][W][Cin]
  float* part;  // [nsplit][Cout*K + tn*Cout]: each split's weight then bias partials
  int B, H, W, Cin, Cout, taps, chunk;
};

// TC x TN tile (TC = 32 WM output channels, TN = 32 WN weight columns, four
// waves of 32 x 32), BK pixels per step through two LDS buffers, the next
// step's global loads in flight during the current step's MFMAs.  Both MFMA
// operands run along the pixels, which are the rows of dz and of the input:
// each thread loads a block of 4 pixels x 8 channels (four 16-byte rows),
// transposes it in registers (byte permutes) and stores 8 rows of 4 pixels.
// A thread's im2col column (tap, channel) is fixed for the whole kernel and
// its pixel advances by the step, so no index division is left in the loop.
// The bias gradient rides on the dz blocks: the workgroup of n-tile y adds
// the blocks of the steps s with s % tn == y, so the tn workgroups of a
// (channel tile, split) share that work and each stores its own partial.
// BK is chosen so that the blocks of a step, (TC + TN) / 8 x BK / 4, are at
// most one per thread: a second block on some threads would set the pace of
// every step.  Requires W % 4 == 0 (a 4-pixel block lies in one image row).
template <int WM, int WN, int BK>
__global__ void __launch_bounds__(256) k_wgrad(WgradArgs a) {
  constexpr int TC = 32 * WM, TN = 32 * WN, LP = BK + 8;
  static_assert(BK % 16 == 0, "16-wide MFMA steps");
  constexpr int NA = TC / 8 * (BK / 4), NB = TN / 8 * (BK / 4);  // 4x8 blocks per step
  constexpr int NJ = (NA + NB + 255) / 256;
  static_assert(WM * WN == 4, "four waves");
  __shared__ __align__(16) u16 As[2][TC][LP];  // [co][pixel]
  __shared__ __align__(16) u16 Bs[2][TN][LP];  // [n][pixel]
  __shared__ float bsum[TC];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int HW = a.H * a.W, M = a.B * HW, K = a.taps * a.Cin;
  const int co0 = blockIdx.x * TC, n0 = blockIdx.y * TN, tn = gridDim.y;
  const int p_lo = blockIdx.z * a.chunk, p_hi = min(M, p_lo + a.chunk);
  // this thread's blocks: A (dz) blocks first, then B (im2col) blocks
  bool isA[NJ], act[NJ], cok[NJ];
  int pg[NJ], col[NJ];
  // B blocks: the column's tap offset and channel, the current pixel
  int bdy[NJ], bdx[NJ], bci[NJ], pb[NJ], ph[NJ], pw[NJ];
#pragma unroll
  for (int jj = 0; jj < NJ; jj++) {
    const int j = tid + 256 * jj;
    act[jj] = j < NA + NB;
    isA[jj] = j < NA;
    const int q = isA[jj] ? j : j - NA;
    pg[jj] = q % (BK / 4);
    col[jj] = (q / (BK / 4)) * 8;
    cok[jj] = act[jj] && (isA[jj] ? co0 + col[jj] < a.Cout : n0 + col[jj] < K);
    bdy[jj] = bdx[jj] = bci[jj] = 0;
    if (!isA[jj]) {
      const int n = n0 + col[jj], tap = n / a.Cin;
      bci[jj] = n - tap * a.Cin;
      if (a.taps == 9) {
        bdy[jj] = tap / 3 - 1;
        bdx[jj] = tap - (tap / 3) * 3 - 1;
      }
    }
    pixel_of(min(p_lo + 4 * pg[jj], M - 1), HW, a.W, pb[jj], ph[jj], pw[jj]);
  }
  uint4 r[NJ][4];
  float bs[NJ][8];
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
#pragma unroll
    for (int e = 0; e < 8; e++) bs[jj][e] = 0.f;
  // loads of the step at p0; then every B block's pixel advances by BK
  auto load = [&](int p0) {
#pragma unroll
    for (int jj = 0; jj < NJ; jj++) {
      const int m = p0 + 4 * pg[jj];
#pragma unroll
      for (int k = 0; k < 4; k++) r[jj][k] = make_uint4(0, 0, 0, 0);
      if (cok[jj] && m < p_hi) {
        if (isA[jj]) {
#pragma unroll
          for (int k = 0; k < 4; k++)
            if (m + k < p_hi) r[jj][k] = *(const uint4*)(a.dz + (size_t)(m + k) * a.Cout + co0 + col[jj]);
        } else {
          const int hh = ph[jj] + bdy[jj];
          if (hh >= 0 && hh < a.H) {
            const u16* row = a.x + ((size_t)(pb[jj] * a.H + hh) * a.W) * a.Cin + bci[jj];
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const int ww = pw[jj] + k + bdx[jj];
              if (m + k < p_hi && ww >= 0 && ww < a.W) r[jj][k] = *(const uint4*)(row + (size_t)ww * a.Cin);
            }
          }
        }
      }
      if (!isA[jj]) {
        pw[jj] += BK;
        while (pw[jj] >= a.W) {
          pw[jj] -= a.W;
          if (++ph[jj] == a.H) {
            ph[jj] = 0;
            ++pb[jj];
          }
        }
      }
    }
  };
  // transpose to [channel][pixel] rows: channel e of the four pixels is
  // half (e & 1) of dword e >> 1 of each 16-byte row
  auto store = [&](int buf, bool bias) {
#pragma unroll
    for (int jj = 0; jj < NJ; jj++) {
      if (!act[jj]) continue;
      u16(*T)[LP] = isA[jj] ? As[buf] : Bs[buf];
      const unsigned* w0 = (const unsigned*)&r[jj][0];
      const unsigned* w1 = (const unsigned*)&r[jj][1];
      const unsigned* w2 = (const unsigned*)&r[jj][2];
      const unsigned* w3 = (const unsigned*)&r[jj][3];
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int d = e >> 1;
        const unsigned sel = (e & 1) ? 0x07060302u : 0x05040100u;
        const unsigned lo = __builtin_amdgcn_perm(w1[d], w0[d], sel);
        const unsigned hi = __builtin_amdgcn_perm(w3[d], w2[d], sel);
        *(uint2*)&T[col[jj] + e][4 * pg[jj]] = make_uint2(lo, hi);
        if (bias && isA[jj])
          bs[jj][e] += (bf2f((u16)(lo & 0xFFFFu)) + bf2f((u16)(lo >> 16))) +
                       (bf2f((u16)(hi & 0xFFFFu)) + bf2f((u16)(hi >> 16)));
      }
    }
  };
  if (tid < TC) bsum[tid] = 0.f;
  f32x16 acc = {};
  load(p_lo);
  store(0, blockIdx.y == 0);
  __syncthreads();
  int buf = 0, step = 0;
  for (int p0 = p_lo; p0 < p_hi; p0 += BK, step++) {
    const bool more = p0 + BK < p_hi;
    if (more) load(p0 + BK);
#pragma unroll
    for (int s4 = 0; s4 < BK / 16; s4++) {
      const int kk = s4 * 16 + (lane >> 5) * 8;
      const bf16x8 av = *(const bf16x8*)&As[buf][wm * 32 + (lane & 31)][kk];
      const bf16x8 bv = *(const bf16x8*)&Bs[buf][wn * 32 + (lane & 31)][kk];
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    // the other buffer was last read in the previous step, before its barrier
    if (more) store(buf ^ 1, (step + 1) % tn == (int)blockIdx.y);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int jj = 0; jj < NJ; jj++)
    if (cok[jj] && isA[jj])
#pragma unroll
      for (int e = 0; e < 8; e++) atomicAdd(&bsum[col[jj] + e], bs[jj][e]);
  __syncthreads();
  const size_t nout = (size_t)a.Cout * (K + tn);
  float* part = a.part + (size_t)blockIdx.z * nout;
  if (tid < TC && co0 + tid < a.Cout) part[(size_t)a.Cout * K + (size_t)blockIdx.y * a.Cout + co0 + tid] = bsum[tid];
  const int n = n0 + wn * 32 + (lane & 31);
  if (n < K) {
#pragma unroll
    for (int r2 = 0; r2 < 16; r2++) {
      const int co = co0 + wm * 32 + acc_row(r2, lane);
      if (co < a.Cout) part[(size_t)co * K + n] = acc[r2];
    }
  }
}

// Weight gradient of a 3x3 layer by pixel blocks with the input halo:
// workgroup (x, y, z) computes dw for output channels [32x, 32x + 32), all 9
// taps and input channels [32y, 32y + 32) over split z of the pixel blocks
// (TH = 4 RW rows x 16 columns of one image).  Per block, dz [pixels][32
// channels] and the (TH + 2) x 18 halo of the input [pixels][32 channels]
// go to LDS as loaded (16-byte pieces), and wave w takes rows RW w ..
// RW w + RW - 1 of the block, 16 pixels per MFMA k-step: its dz operand (32
// channels x 16 pixels) is read once and each tap's input operand (16
// pixels x 32 channels) is the halo window shifted by the tap, both with LDS
// transpose reads (ds_read_b64_tr_b16: within each group of 16 lanes, lane i
// receives element i % 4 of the 8 bytes addressed by lane 4e + i / 4 as its
// element e, so lanes 4e' + q addressing pixel k0 + e', channels c + 4q ..
// c + 4q + 3 give lane i pixels k0 .. k0 + 3 of channel c + i; two such
// reads make one operand of v_mfma_f32_32x32x16_bf16).  So dz is read once
// per input-channel chunk and the input once per output-channel tile (plus
// the halo), where k_wgrad re-reads dz for every n-tile and the input for
// every tap.  Nine 32 x 32 accumulators per wave (the waves' sums over their
// rows added in wave order at the end); the next block's pieces load during
// the current block's MFMAs.  Bias gradient (chunk 0 only):
// per-thread sums of the dz pieces, then a fixed-order sum over the threads
// of each 8-channel group.  Partials as k_wgrad with tn = 1.  Requires Cin
// and Cout multiples of 32, H % TH == 0, W % 16 == 0.
template <int RW>
__global__ void __launch_bounds__(256, 2) k_wgrad_halo(WgradArgs a) {
  constexpr int TH = 4 * RW, HPIX = (TH + 2) * 18, PIX = TH * 16;
  constexpr int ND = PIX * 4, NH = HPIX * 4;  // 16-byte pieces per block
  constexpr int JD = (ND + 255) / 256, JH = (NH + 255) / 256;
  // dz [pixel][channel] and the input halo [pixel][channel], two buffers
  // each; after the pixel loop the same LDS adds the waves' accumulators
  constexpr size_t kOps = (size_t)2 * (PIX + HPIX) * 32 * sizeof(u16), kRed = (size_t)4 * 1024 * sizeof(float);
  __shared__ __align__(16) unsigned char smem[kOps > kRed ? kOps : kRed];
  u16(*Ds)[PIX][32] = reinterpret_cast<u16(*)[PIX][32]>(smem);
  u16(*Xs)[HPIX][32] = reinterpret_cast<u16(*)[HPIX][32]>(smem + (size_t)2 * PIX * 32 * sizeof(u16));
  __shared__ float bpart[256][9];
  typedef short v4s __attribute__((ext_vector_type(4)));
  typedef v4s __attribute__((address_space(3))) * lds4;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int co0 = blockIdx.x * 32, ci0 = blockIdx.y * 32;
  const int K = 9 * a.Cin;
  const int nwb = a.W / 16, nhb = a.H / TH, nblk = a.B * nhb * nwb;
  const int b_lo = blockIdx.z * a.chunk, b_hi = min(nblk, b_lo + a.chunk);
  const bool bias = blockIdx.y == 0;
  uint4 rd[JD], rx[JH];
  float bs[8];
#pragma unroll
  for (int e = 0; e < 8; e++) bs[e] = 0.f;
  auto load = [&](int blk) {
    const int wb = blk % nwb, hb = (blk / nwb) % nhb, bb = blk / (nwb * nhb);
    const int h0 = hb * TH, w0 = wb * 16;
#pragma unroll
    for (int j = 0; j < JD; j++) {
      const int q = tid + 256 * j, px = q >> 2;
      rd[j] = make_uint4(0, 0, 0, 0);
      if (q < ND) {
        const size_t m = (size_t)(bb * a.H + h0 + px / 16) * a.W + w0 + px % 16;
        rd[j] = *(const uint4*)(a.dz + m * a.Cout + co0 + 8 * (q & 3));
      }
    }
#pragma unroll
    for (int j = 0; j < JH; j++) {
      const int q = tid + 256 * j, hp = q >> 2;
      rx[j] = make_uint4(0, 0, 0, 0);
      if (q < NH) {
        const int hh = h0 + hp / 18 - 1, ww = w0 + hp % 18 - 1;
        if (hh >= 0 && hh < a.H && ww >= 0 && ww < a.W)
          rx[j] = *(const uint4*)(a.x + ((size_t)(bb * a.H + hh) * a.W + ww) * a.Cin + ci0 + 8 * (q & 3));
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int j = 0; j < JD; j++) {
      const int q = tid + 256 * j;
      if (q < ND) {
        *(uint4*)&Ds[buf][q >> 2][8 * (q & 3)] = rd[j];
        if (bias) {
          float f[8];
          unpack8(rd[j], f);
#pragma unroll
          for (int e = 0; e < 8; e++) bs[e] += f[e];
        }
      }
    }
#pragma unroll
    for (int j = 0; j < JH; j++) {
      const int q = tid + 256 * j;
      if (q < NH) *(uint4*)&Xs[buf][q >> 2][8 * (q & 3)] = rx[j];
    }
  };
  // transpose-read rows and columns of this lane: pixel offset within a
  // 16-pixel k-step and channel, for operand half h = 0 (h = 1: +4 pixels)
  const int g = lane >> 4, s = lane & 15;
  const int tk = 8 * (g >> 1) + (s >> 2), tc = 16 * (g & 1) + 4 * (s & 3);
  f32x16 acc[9];
#pragma unroll
  for (int t = 0; t < 9; t++) acc[t] = f32x16{};
  if (b_lo < b_hi) {
    load(b_lo);
    store(0);
  }
  __syncthreads();
  int buf = 0;
  for (int blk = b_lo; blk < b_hi; blk++) {
    const bool more = blk + 1 < b_hi;
    if (more) load(blk + 1);
#pragma unroll
    for (int rr = 0; rr < RW; rr++) {
      const int row = RW * wave + rr;  // block row of this k-step
      bf16x8 av;
      {
        const v4s lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds4)&Ds[buf][16 * row + tk][tc]);
        const v4s hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds4)&Ds[buf][16 * row + tk + 4][tc]);
        const short tmp[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        __builtin_memcpy(&av, tmp, 16);
      }
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const int hp = (row + t / 3) * 18 + t % 3 + tk;  // halo pixel of k = tk, tap t
        const v4s lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds4)&Xs[buf][hp][tc]);
        const v4s hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds4)&Xs[buf][hp + 4][tc]);
        const short tmp[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        bf16x8 bv;
        __builtin_memcpy(&bv, tmp, 16);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[t], 0, 0, 0);
      }
    }
    if (more) {
      __syncthreads();  // this block's operands are read
      store(buf ^ 1);
    }
    __syncthreads();
    buf ^= 1;
  }
  const size_t nout = (size_t)a.Cout * (K + 1);
  float* part = a.part + (size_t)blockIdx.z * nout;
  if (bias) {
#pragma unroll
    for (int e = 0; e < 8; e++) bpart[tid][e] = bs[e];
    __syncthreads();
    if (tid < 32) {
      // threads of channel group tid / 8 hold pieces q = tid' with q % 4 == tid / 8
      float sum = 0.f;
      for (int t2 = tid >> 3; t2 < 256; t2 += 4) sum += bpart[t2][tid & 7];
      part[(size_t)a.Cout * K + co0 + tid] = sum;
    }
  }
  // The four waves hold the sums over their own block rows: per tap, their
  // 32 x 32 blocks (acc[t] element r: output channel acc_row(r, lane), input
  // channel lane % 32) are added in wave order through LDS, then stored
  // channel-row by channel-row
  float* red = reinterpret_cast<float*>(smem);
#pragma unroll
  for (int t = 0; t < 9; t++) {
    __syncthreads();  // the pixel loop's (or the previous tap's) LDS reads are done
#pragma unroll
    for (int r2 = 0; r2 < 16; r2++) red[wave * 1024 + acc_row(r2, lane) * 32 + (lane & 31)] = acc[t][r2];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = tid + 256 * k, row = e >> 5, col = e & 31;
      const float v = ((red[e] + red[1024 + e]) + red[2048 + e]) + red[3072 + e];
      part[(size_t)(co0 + row) * K + t * a.Cin + ci0 + col] = v;
    }
  }
}

// The splits' partials are added in two fixed-order stages.  Stage 1:
// workgroup (x, g) adds splits [g*zc, (g+1)*zc) of 64 consecutive partial
// values (4 lanes of 64 threads, each lane every 4th split, then the lanes in
// order) into part2[g][.].  nout = Cout*K + tn*Cout values: the weight
// partials [Cout][K] followed by the bias partials [tn][Cout] of each split.
__global__ void __launch_bounds__(256) k_wgrad_reduce1(int nsplit, int zc, int64_t nout, const float* __restrict__ part,
                                                       float* __restrict__ part2) {
  __shared__ float red[4][64];
  const int il = threadIdx.x & 63, zl = threadIdx.x >> 6;
  const int64_t i = blockIdx.x * (int64_t)64 + il;
  const int z0 = blockIdx.y * zc, z1 = min(nsplit, z0 + zc);
  float s = 0.f;
  if (i < nout)
    for (int z = z0 + zl; z < z1; z += 4) s += part[z * nout + i];
  red[zl][il] = s;
  __syncthreads();
  if (zl == 0 && i < nout) part2[blockIdx.y * nout + i] = (red[0][il] + red[1][il]) + (red[2][il] + red[3][il]);
}
// Stage 2: dw[co][n] += sum_g part2[g][co*K + n], and (threads past the
// weights) db[c] += sum over g < ng, y < tn of part2[g][Cout*K + y*Cout + c],
// both in that fixed order
__global__ void k_wgrad_reduce2(int ng, int Cout, int K, int Kp, int tn, const float* __restrict__ part2,
                                float* __restrict__ dw, float* __restrict__ db) {
  const int64_t nw = (int64_t)Cout * K, nout = nw + (int64_t)tn * Cout;
  EW_LOOP(nw + Cout) {
    float s = 0.f;
    if (i < nw) {
      for (int g = 0; g < ng; g++) s += part2[g * nout + i];
      const int co = (int)(i / K), n = (int)(i - (int64_t)co * K);
      dw[(size_t)co * Kp + n] += s;
    } else {
      const int c = (int)(i - nw);
      for (int g = 0; g < ng; g++)
        for (int y = 0; y < tn; y++) s += part2[g * nout + nw + (int64_t)y * Cout + c];
      db[c] += s;
    }
  }
}

// The layout kernels below move 8 channels (16 bytes) per thread: every
// channel count of the network is a multiple of 8, and the element counts
// they take (pixels x channels / 8) fit 32 bits (checked by the launchers).
#define EW_LOOP32(n) for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

// 2x2 average pool, NHWC; n8 = B * H/2 * W/2 * C/8
__global__ void k_pool(int n8, int H, int W, int C, const u16* __restrict__ x, u16* __restrict__ y) {
  const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
  EW_LOOP32(n8) {
    const int g = i % C8, p = i / C8;
    const int wo = p % Wo, t = p / Wo, ho = t % Ho, b = t / Ho;
    const uint4* x0 = (const uint4*)(x + ((size_t)(b * H + 2 * ho) * W + 2 * wo) * C) + g;
    const size_t row = (size_t)W * C8;  // uint4 per image row
    float f0[8], f1[8], f2[8], f3[8], o[8];
    unpack8(x0[0], f0);
    unpack8(x0[C8], f1);
    unpack8(x0[row], f2);
    unpack8(x0[row + C8], f3);
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = 0.25f * (f0[e] + f1[e] + f2[e] + f3[e]);
    ((uint4*)y)[i] = pack8(o);
  }
}
// dx[b,h,w,c] = (dx + dy[b,h/2,w/2,c] / 4) * (y > 0): the pooled input's
// gradient completed and passed back through its ReLU; n8 = B*H*W*C/8
__global__ void k_pool_bwd_add(int n8, int H, int W, int C, const u16* __restrict__ dy, u16* __restrict__ dx,
                               const u16* __restrict__ y) {
  const int C8 = C / 8;
  EW_LOOP32(n8) {
    const int g = i % C8, p = i / C8;
    const int w = p % W, t = p / W, h = t % H, b = t / H;
    float a[8], d[8];
    unpack8(((const uint4*)dx)[i], a);
    unpack8(((const uint4*)(dy + ((size_t)(b * (H / 2) + h / 2) * (W / 2) + w / 2) * C))[g], d);
#pragma unroll
    for (int e = 0; e < 8; e++) a[e] = bf2f(f2bf(a[e] + 0.25f * d[e]));
    relu_mask8(a, ((const uint4*)y)[i]);
    ((uint4*)dx)[i] = pack8(a);
  }
}
// cat[b,h,w,:] = (low[b,h/2,w/2,0:Cu], skip[b,h,w,0:Cs]); n8 = B*H*W*(Cu+Cs)/8
__global__ void k_upcat(int n8, int H, int W, int Cu, int Cs, const u16* __restrict__ low,
                        const u16* __restrict__ skip, u16* __restrict__ cat) {
  const int C8 = (Cu + Cs) / 8, Cu8 = Cu / 8, Cs8 = Cs / 8;
  EW_LOOP32(n8) {
    const int g = i % C8, p = i / C8;
    uint4 v;
    if (g < Cu8) {
      const int w = p % W, t = p / W, h = t % H, b = t / H;
      v = ((const uint4*)low)[((size_t)(b * (H / 2) + h / 2) * (W / 2) + w / 2) * Cu8 + g];
    } else {
      v = ((const uint4*)skip)[(size_t)p * Cs8 + g - Cu8];
    }
    ((uint4*)cat)[i] = v;
  }
}
// dskip = dcat[..., Cu:]; n8 = P*Cs/8
__global__ void k_upcat_bwd_skip(int n8, int Cu, int Cs, const u16* __restrict__ dcat, u16* __restrict__ dskip) {
  const int Cs8 = Cs / 8, C8 = (Cu + Cs) / 8;
  EW_LOOP32(n8) {
    const int g = i % Cs8, p = i / Cs8;
    ((uint4*)dskip)[i] = ((const uint4*)dcat)[(size_t)p * C8 + Cu / 8 + g];
  }
}
// dlow[b,h,w,c] = (sum over the 2x2 block of dcat[..., c]) * (ylow > 0);
// n8 = B*Hl*Wl*Cu/8
__global__ void k_upcat_bwd_low(int n8, int Hl, int Wl, int Cu, int Cs, const u16* __restrict__ dcat,
                                u16* __restrict__ dlow, const u16* __restrict__ ylow) {
  const int C8 = (Cu + Cs) / 8, Cu8 = Cu / 8, H = 2 * Hl, W = 2 * Wl;
  EW_LOOP32(n8) {
    const int g = i % Cu8, p = i / Cu8;
    const int w = p % Wl, t = p / Wl, h = t % Hl, b = t / Hl;
    const uint4* d0 = (const uint4*)dcat + ((size_t)(b * H + 2 * h) * W + 2 * w) * C8 + g;
    const size_t row = (size_t)W * C8;
    float f0[8], f1[8], f2[8], f3[8], o[8];
    unpack8(d0[0], f0);
    unpack8(d0[C8], f1);
    unpack8(d0[row], f2);
    unpack8(d0[row + C8], f3);
#pragma unroll
    for (int e = 0; e < 8; e++) o[e] = bf2f(f2bf(f0[e] + f1[e] + f2[e] + f3[e]));
    relu_mask8(o, ((const uint4*)ylow)[i]);
    ((uint4*)dlow)[i] = pack8(o);
  }
}

// wf[ci][t' * Cout + co] = w[co][t * Cin + ci], t' = taps - 1 - t
// Every layer in one launch: layer l covers elements [start[l], start[l+1])
// of the concatenated flips (at most kMaxLayers = 4 * 6 - 1 convolutions).
constexpr int kMaxLayers = 23;
struct FlipTable {
  int n;
  int64_t start[kMaxLayers + 1];
  int cout[kMaxLayers], cin[kMaxLayers], taps[kMaxLayers], kp[kMaxLayers], kp2[kMaxLayers];
  const u16* w[kMaxLayers];
  u16* wf[kMaxLayers];
};
__global__ void k_flip_all(FlipTable f) {
  EW_LOOP(f.start[f.n]) {
    int l = 0;
    while (i >= f.start[l + 1]) l++;
    const int64_t j = i - f.start[l];
    const int Cout = f.cout[l], Cin = f.cin[l], taps = f.taps[l];
    const int co = (int)(j % Cout);
    const int t2 = (int)((j / Cout) % taps);
    const int ci = (int)(j / ((int64_t)Cout * taps));
    f.wf[l][(size_t)ci * f.kp2[l] + t2 * Cout + co] = f.w[l][(size_t)co * f.kp[l] + (taps - 1 - t2) * Cin + ci];
  }
}

// Loss per pixel over the head output (amplitude, phase, support logit) and
// the target (amplitude, phase, mask):
//   (a - a*)^2 + lp * m* (1 - cos(p - p*)) + ls * BCE(s, m*), averaged over pixels;
// dz = dLoss/dout as BF16 (channels 3..7 zero).
__global__ void k_loss(int M, const float* __restrict__ yf, const float* __restrict__ tg, float lp, float ls,
                       u16* __restrict__ dz, double* loss) {
  double acc = 0;
  const float inv = 1.f / (float)M;
  EW_LOOP(M) {
    const float a = yf[i * 8], p = yf[i * 8 + 1], s = yf[i * 8 + 2];
    const float at = tg[i * 4], pt = tg[i * 4 + 1], mt = tg[i * 4 + 2];
    const float sg = 1.f / (1.f + expf(-s));
    const float d = p - pt;
    acc += (double)((a - at) * (a - at) + lp * mt * (1.f - cosf(d)) +
                    ls * (fmaxf(s, 0.f) - s * mt + log1pf(expf(-fabsf(s)))));
    u16* o = dz + (size_t)i * 8;
    o[0] = f2bf(2.f * (a - at) * inv);
    o[1] = f2bf(lp * mt * sinf(d) * inv);
    o[2] = f2bf(ls * (sg - mt) * inv);
#pragma unroll
    for (int c = 3; c < 8; c++) o[c] = 0;
  }
  __shared__ double sh[16];
  block_store_sum(acc, loss + blockIdx.x, sh);
}

// Adam on the flat FP32 parameters; refreshes the BF16 copy, clears the gradient
__global__ void k_adam(int64_t n, float* __restrict__ w, float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, u16* __restrict__ wb, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  EW_LOOP(n) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    const float wi = w[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
    g[i] = 0.f;
  }
}
__global__ void k_to_bf16(int64_t n, const float* __restrict__ w, u16* __restrict__ wb) {
  EW_LOOP(n) wb[i] = f2bf(w[i]);
}
__global__ void k_zero_f32(int64_t n, float* p) {
  EW_LOOP(n) p[i] = 0.f;
}
__global__ void k_zero_u16(int64_t n, u16* p) {
  EW_LOOP(n) p[i] = 0;
}

// network input of sample b: channel 0 = log(1 + I) / 16 of the centred
// (fftshifted) pattern, channels 1..7 zero
__global__ void k_pack_input(int n, const float* __restrict__ I, u16* __restrict__ x) {
  const int64_t N = (int64_t)n * n;
  EW_LOOP(N) {
    const int h = (int)(i / n), w = (int)(i % n);
    const int hs = h + n / 2 < n ? h + n / 2 : h - n / 2, ws = w + n / 2 < n ? w + n / 2 : w - n / 2;
    const float v = I[(size_t)hs * n + ws];
    u16* o = x + (size_t)i * 8;
    o[0] = f2bf(log1pf(fmaxf(v, 0.f)) * (1.f / 16.f));
#pragma unroll
    for (int c = 1; c < 8; c++) o[c] = 0;
  }
}
// one workgroup: sum of a complex float image (fixed order)
__global__ void k_csum(int64_t n, const cplx<float>* __restrict__ z, double* out) {
  double re = 0, im = 0;
  // unrolled so that several loads are in flight (same summation order)
#pragma unroll 8
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    re += z[i].x;
    im += z[i].y;
  }
  __shared__ double sh[16];
  block_store_sum(re, out, sh);
  block_store_sum(im, out + 1, sh);
}
// targets: amplitude, phase relative to the global phase arg(sum obj), mask
__global__ void k_pack_target(int64_t n, const cplx<float>* __restrict__ obj, const double* csum,
                              float* __restrict__ tg) {
  const float th = (float)atan2(csum[1], csum[0]);
  const float c = cosf(th), s = sinf(th);
  EW_LOOP(n) {
    const cplx<float> z = obj[i];
    const float re = z.x * c + z.y * s, im = z.y * c - z.x * s;  // z e^{-i th}
    const float a = sqrtf(z.x * z.x + z.y * z.y);
    float* o = tg + (size_t)i * 4;
    o[0] = a;
    o[1] = a > 0.f ? atan2f(im, re) : 0.f;
    o[2] = a > 0.f ? 1.f : 0.f;
    o[3] = 0.f;
  }
}
// The same packing for the frames of a forward-model stack (frame b's
// element (r, c) at r * B * n + b * n + c) into batch entries b0 + b.
__global__ void k_pack_input_stack(int n, int B, int b0, const float* __restrict__ I, u16* __restrict__ x) {
  const int64_t N = (int64_t)n * n, W = (int64_t)B * n;
  EW_LOOP((int64_t)B * N) {
    const int b = (int)(i / N);
    const int64_t k = i - (int64_t)b * N;
    const int h = (int)(k / n), w = (int)(k % n);
    const int hs = h + n / 2 < n ? h + n / 2 : h - n / 2, ws = w + n / 2 < n ? w + n / 2 : w - n / 2;
    const float v = I[(int64_t)hs * W + (int64_t)b * n + ws];
    u16* o = x + ((size_t)(b0 + b) * N + k) * 8;
    o[0] = f2bf(log1pf(fmaxf(v, 0.f)) * (1.f / 16.f));
#pragma unroll
    for (int c = 1; c < 8; c++) o[c] = 0;
  }
}
// k_csum per frame of a stack: one workgroup per frame, the same order
__global__ void k_csum_stack(int n, int B, const cplx<float>* __restrict__ z, double* out) {
  const int b = blockIdx.x;
  const int64_t N = (int64_t)n * n, W = (int64_t)B * n;
  double re = 0, im = 0;
#pragma unroll 8
  for (int64_t i = threadIdx.x; i < N; i += blockDim.x) {
    const cplx<float> v = z[(i / n) * W + (int64_t)b * n + i % n];
    re += v.x;
    im += v.y;
  }
  __shared__ double sh[16];
  block_store_sum(re, out + 2 * b, sh);
  block_store_sum(im, out + 2 * b + 1, sh);
}
__global__ void k_pack_target_stack(int n, int B, int b0, const cplx<float>* __restrict__ obj,
                                    const double* __restrict__ csum, float* __restrict__ tg) {
  const int64_t N = (int64_t)n * n, W = (int64_t)B * n;
  EW_LOOP((int64_t)B * N) {
    const int b = (int)(i / N);
    const int64_t k = i - (int64_t)b * N;
    const float th = (float)atan2(csum[2 * b + 1], csum[2 * b]);
    const float c = cosf(th), s = sinf(th);
    const cplx<float> z = obj[(k / n) * W + (int64_t)b * n + k % n];
    const float re = z.x * c + z.y * s, im = z.y * c - z.x * s;  // z e^{-i th}
    const float a = sqrtf(z.x * z.x + z.y * z.y);
    float* o = tg + ((size_t)(b0 + b) * N + k) * 4;
    o[0] = a;
    o[1] = a > 0.f ? atan2f(im, re) : 0.f;
    o[2] = a > 0.f ? 1.f : 0.f;
    o[3] = 0.f;
  }
}
// warm start from the head output of sample 0: support = sigmoid(logit) > thr,
// g0 = amplitude e^{i phase} inside the support
__global__ void k_ws_compose(int64_t n, const float* __restrict__ yf, float thr, cplx<float>* __restrict__ g0,
                             float* __restrict__ spt) {
  EW_LOOP(n) {
    const float a = fmaxf(yf[i * 8], 0.f), p = yf[i * 8 + 1], s = yf[i * 8 + 2];
    const bool in = 1.f / (1.f + expf(-s)) > thr;
    g0[i] = in ? cplx<float>{a * cosf(p), a * sinf(p)} : cplx<float>{0.f, 0.f};
    spt[i] = in ? 1.f : 0.f;
  }
}
__global__ void k_sum_sq(int64_t n, const cplx<float>* __restrict__ z, const float* __restrict__ I, double* out) {
  double a = 0, b = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    a += (double)z[i].x * z[i].x + (double)z[i].y * z[i].y;
    b += (double)fmaxf(I[i], 0.f);
  }
  __shared__ double sh[16];
  block_store_sum(a, out, sh);
  block_store_sum(b, out + 1, sh);
}

// host Philox4x64-10 (Salmon et al., SC'11), same streams as the device one
void philox_host(uint64_t c[4], uint64_t k0, uint64_t k1) {
  for (int r = 0; r < 10; r++) {
    const __uint128_t p0 = (__uint128_t)0xD2E7470EE14C6C93ULL * c[0];
    const __uint128_t p1 = (__uint128_t)0xCA5A826395121157ULL * c[2];
    const uint64_t n0 = (uint64_t)(p1 >> 64) ^ c[1] ^ k0, n2 = (uint64_t)(p0 >> 64) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint64_t)p1;
    c[2] = n2;
    c[3] = (uint64_t)p0;
    k0 += 0x9E3779B97F4A7C15ULL;
    k1 += 0xBB67AE8584CAA73BULL;
  }
}
double uniform_host(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t c[4] = {idx, 0, 0, 0};
  philox_host(c, seed, stream);
  return (double)(c[0] >> 11) * 0x1.0p-53;
}

int64_t ew_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

}  // namespace

// ---------------------------------------------------------------------------
// the network
// ---------------------------------------------------------------------------
struct UConv {
  int cin, cout, taps, kp, kp2, relu;
  int64_t w_off, b_off;  // offsets into the flat parameter arrays
  u16* wflip = nullptr;  // [cin][kp2] for the input gradient
};
struct UTensor {
  u16* y = nullptr;
  u16* dy = nullptr;
  int h = 0, c = 0;
};

struct cdx_unet {
  int size, levels, base, max_batch;
  hipStream_t st = nullptr;
  cdx_ctx* ctx = nullptr;  // size x size FP32 context for the forward model
  std::vector<UConv> conv;  // enc (a,b) per level, dec (a,b) per level < L-1, head
  int64_t nparam = 0;
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  u16* wb = nullptr;
  // activations
  UTensor in0;
  std::vector<UTensor> ea, e, p, cat, da, d;
  float* yf = nullptr;  // head output [B][H][W][8]
  u16* dz_head = nullptr;
  float* tg = nullptr;  // targets [B][H][W][4]
  double* dscal = nullptr;  // reductions (loss partials, sums)
  float* wg_part = nullptr;  // weight-gradient partials of the pixel splits
  void *obj = nullptr, *I = nullptr;  // forward-model scratch (one frame)
  // forward model for up to synth_chunk frames side by side: fields a, ph
  // and their smoothing temporaries, min/max keys, |F|^2 totals and
  // complex sums (syn); objects, transforms and intensities
  int synth_chunk = 0;
  double* syn = nullptr;
  void *sobj = nullptr, *sF = nullptr;
  float* sI = nullptr;
  int64_t adam_t = 0;
};

namespace {

int enc_a(int l) { return 2 * l; }
int enc_b(int l) { return 2 * l + 1; }
int dec_a(const cdx_unet* u, int l) { return 2 * u->levels + 2 * l; }
int dec_b(const cdx_unet* u, int l) { return 2 * u->levels + 2 * l + 1; }
int head(const cdx_unet* u) { return 2 * u->levels + 2 * (u->levels - 1); }

int alloc_tensor(UTensor& t, int B, int h, int c) {
  t.h = h;
  t.c = c;
  const size_t bytes = (size_t)B * h * h * c * sizeof(u16);
  UCHK(hipMalloc(&t.y, bytes));
  UCHK(hipMalloc(&t.dy, bytes));
  return CDX_OK;
}
void free_tensor(UTensor& t) {
  hipFree(t.y);
  hipFree(t.dy);
  t.y = t.dy = nullptr;
}

int conv_fwd(cdx_unet* u, const UConv& L, int B, int h, const u16* x, u16* y, float* yf, bool relu) {
  ConvArgs a{x, u->wb + L.w_off, u->w + L.b_off, y, yf, B, h, h, L.cin, L.cout, L.kp, L.taps, relu ? 1 : 0};
  launch_conv(a, L.cout, u->st);
  UCHK(hipGetLastError());
  return CDX_OK;
}
// input gradient: dx = conv(dz, flipped weights), no bias, passed back
// through the ReLU whose output was x (x_relu, or null when x is no ReLU output)
int conv_dgrad(cdx_unet* u, const UConv& L, int B, int h, const u16* dz, u16* dx, const u16* x_relu) {
  ConvArgs a{dz, L.wflip, nullptr, dx, nullptr, B, h, h, L.cout, L.cin, L.kp2, L.taps, 0, x_relu};
  launch_conv(a, L.cin, u->st);
  UCHK(hipGetLastError());
  return CDX_OK;
}
// weight and bias gradients of layer L: dz [B*h*h][cout], x its input
int conv_wgrad(cdx_unet* u, const UConv& L, int B, int h, const u16* dz, const u16* x) {
  if (h % 4 || L.cin % 8 || L.cout % 8)
    return ufail(CDX_ERR_ARG, "unet: weight gradient needs sides divisible by 4 and channels by 8");
  const int M = B * h * h, K = L.taps * L.cin;
  // 3x3 layers with 32-channel chunks: the halo kernel (blocks of 4 x 16
  // pixels), each workgroup all 9 taps of a 32 x 32 channel pair
  const bool halo = L.taps == 9 && L.cin % 32 == 0 && L.cout % 32 == 0 && h % 16 == 0;
  // otherwise tiles of 32 x 128 with 48-pixel steps for at most 32 channels,
  // else 64 x 64 with 64
  const bool narrow = L.cout <= 32;
  const int TC = halo ? 32 : narrow ? 32 : 64, TN = halo ? 0 : narrow ? 128 : 64, BK = narrow ? 48 : 64;
  const int tm = (L.cout + TC - 1) / TC, tn = halo ? 1 : (K + TN - 1) / TN;
  const int ty = halo ? L.cin / 32 : tn;              // grid y: input-channel chunks or n-tiles
  const int units = halo ? B * (h / 4) * (h / 16) : M;  // what the splits share: pixel blocks or pixels
  const int64_t nout = (int64_t)L.cout * (K + tn);
  // about 1024 workgroups (512 for the halo kernel: two per CU, each with
  // nine partial tiles to write), at least 256 pixels each, within the
  // workspace (the splits' partials and the reduction's first stage, 1/16 of
  // them)
  int nsplit = (halo ? 512 : 1024) / (tm * ty);
  nsplit = std::max(1, std::min(nsplit, halo ? units / 2 : M / 256));
  nsplit = (int)std::min<int64_t>(nsplit, kWgradPartFloats / (nout + nout / 16 + 1) - 1);
  if (nsplit < 1) return ufail(CDX_ERR_ARG, "unet: layer too large for the weight-gradient workspace");
  int chunk = (units + nsplit - 1) / nsplit;
  if (!halo) chunk = (chunk + BK - 1) / BK * BK;
  nsplit = (units + chunk - 1) / chunk;
  WgradArgs a{dz, x, u->wg_part, B, h, h, L.cin, L.cout, L.taps, chunk};
  if (halo)
    hipLaunchKernelGGL((k_wgrad_halo<1>), dim3(tm, ty, nsplit), dim3(256), 0, u->st, a);
  else if (narrow)
    hipLaunchKernelGGL((k_wgrad<1, 4, 48>), dim3(tm, tn, nsplit), dim3(256), 0, u->st, a);
  else
    hipLaunchKernelGGL((k_wgrad<2, 2, 64>), dim3(tm, tn, nsplit), dim3(256), 0, u->st, a);
  UCHK(hipGetLastError());
  const int zc = 16, ng = (nsplit + zc - 1) / zc;
  float* part2 = u->wg_part + (int64_t)nsplit * nout;
  hipLaunchKernelGGL(k_wgrad_reduce1, dim3((unsigned)((nout + 63) / 64), ng), dim3(256), 0, u->st, nsplit, zc, nout,
                     (const float*)u->wg_part, part2);
  UCHK(hipGetLastError());
  hipLaunchKernelGGL(k_wgrad_reduce2, dim3(ew_blocks((int64_t)L.cout * K + L.cout)), dim3(256), 0, u->st, ng,
                     L.cout, K, L.kp, tn, (const float*)part2, u->g + L.w_off, u->g + L.b_off);
  UCHK(hipGetLastError());
  return CDX_OK;
}
// threads for a layout kernel over n8 16-byte groups
int ew8_grid(int64_t n8, dim3& grid) {
  if (n8 >= ((int64_t)1 << 31)) return ufail(CDX_ERR_ARG, "unet: activation too large");
  grid = dim3((unsigned)ew_blocks(n8));
  return CDX_OK;
}
// the flipped, transposed BF16 weights of the input gradients, all layers
int flip_weights(cdx_unet* u) {
  FlipTable f{};
  if ((int)u->conv.size() > kMaxLayers) return ufail(CDX_ERR_ARG, "unet: more layers than the flip table holds");
  f.n = (int)u->conv.size();
  for (int l = 0; l < f.n; l++) {
    const UConv& L = u->conv[l];
    f.start[l + 1] = f.start[l] + (int64_t)L.cin * L.taps * L.cout;
    f.cout[l] = L.cout;
    f.cin[l] = L.cin;
    f.taps[l] = L.taps;
    f.kp[l] = L.kp;
    f.kp2[l] = L.kp2;
    f.w[l] = (const u16*)(u->wb + L.w_off);
    f.wf[l] = L.wflip;
  }
  hipLaunchKernelGGL(k_flip_all, dim3(ew_blocks(f.start[f.n])), dim3(256), 0, u->st, f);
  UCHK(hipGetLastError());
  return CDX_OK;
}
int refresh_weights(cdx_unet* u) {
  hipLaunchKernelGGL(k_to_bf16, dim3(ew_blocks(u->nparam)), dim3(256), 0, u->st, u->nparam, u->w, u->wb);
  UCHK(hipGetLastError());
  return flip_weights(u);
}

int forward(cdx_unet* u, int B) {
  const int Lv = u->levels;
  for (int l = 0; l < Lv; l++) {
    const int h = u->size >> l;
    const u16* x = l == 0 ? u->in0.y : u->p[l].y;
    RCHK(conv_fwd(u, u->conv[enc_a(l)], B, h, x, u->ea[l].y, nullptr, true));
    RCHK(conv_fwd(u, u->conv[enc_b(l)], B, h, u->ea[l].y, u->e[l].y, nullptr, true));
    if (l + 1 < Lv) {
      const int64_t n8 = (int64_t)B * (h / 2) * (h / 2) * u->e[l].c / 8;
      dim3 g;
      RCHK(ew8_grid(n8, g));
      hipLaunchKernelGGL(k_pool, g, dim3(256), 0, u->st, (int)n8, h, h, u->e[l].c, u->e[l].y, u->p[l + 1].y);
      UCHK(hipGetLastError());
    }
  }
  for (int l = Lv - 2; l >= 0; l--) {
    const int h = u->size >> l;
    const UTensor& low = l == Lv - 2 ? u->e[Lv - 1] : u->d[l + 1];
    const int64_t n8 = (int64_t)B * h * h * u->cat[l].c / 8;
    dim3 g;
    RCHK(ew8_grid(n8, g));
    hipLaunchKernelGGL(k_upcat, g, dim3(256), 0, u->st, (int)n8, h, h, low.c, u->e[l].c, low.y, u->e[l].y,
                       u->cat[l].y);
    UCHK(hipGetLastError());
    RCHK(conv_fwd(u, u->conv[dec_a(u, l)], B, h, u->cat[l].y, u->da[l].y, nullptr, true));
    RCHK(conv_fwd(u, u->conv[dec_b(u, l)], B, h, u->da[l].y, u->d[l].y, nullptr, true));
  }
  const UTensor& top = Lv >= 2 ? u->d[0] : u->e[0];
  RCHK(conv_fwd(u, u->conv[head(u)], B, u->size, top.y, nullptr, u->yf, false));
  return CDX_OK;
}

// Each ReLU's backward is folded into the kernel that completes the
// gradient of its output: the head's and the b-convs' input gradients
// (conv_dgrad's mask), the upsampling's low branch (k_upcat_bwd_low) and the
// pooling's backward, which adds the last term of a skip tensor's gradient.
int backward(cdx_unet* u, int B) {
  const int Lv = u->levels;
  const int H0 = u->size;
  UTensor& top = Lv >= 2 ? u->d[0] : u->e[0];
  const UConv& hd = u->conv[head(u)];
  RCHK(conv_wgrad(u, hd, B, H0, u->dz_head, top.y));
  RCHK(conv_dgrad(u, hd, B, H0, u->dz_head, top.dy, top.y));
  for (int l = 0; l + 1 < Lv; l++) {
    const int h = H0 >> l;
    const int64_t np = (int64_t)B * h * h;
    RCHK(conv_wgrad(u, u->conv[dec_b(u, l)], B, h, u->d[l].dy, u->da[l].y));
    RCHK(conv_dgrad(u, u->conv[dec_b(u, l)], B, h, u->d[l].dy, u->da[l].dy, u->da[l].y));
    RCHK(conv_wgrad(u, u->conv[dec_a(u, l)], B, h, u->da[l].dy, u->cat[l].y));
    RCHK(conv_dgrad(u, u->conv[dec_a(u, l)], B, h, u->da[l].dy, u->cat[l].dy, nullptr));
    UTensor& low = l == Lv - 2 ? u->e[Lv - 1] : u->d[l + 1];
    const int Cs = u->e[l].c;
    dim3 g;
    RCHK(ew8_grid(np * Cs / 8, g));
    hipLaunchKernelGGL(k_upcat_bwd_skip, g, dim3(256), 0, u->st, (int)(np * Cs / 8), low.c, Cs, u->cat[l].dy,
                       u->e[l].dy);
    UCHK(hipGetLastError());
    const int64_t nl8 = (int64_t)B * (h / 2) * (h / 2) * low.c / 8;
    RCHK(ew8_grid(nl8, g));
    hipLaunchKernelGGL(k_upcat_bwd_low, g, dim3(256), 0, u->st, (int)nl8, h / 2, h / 2, low.c, Cs, u->cat[l].dy,
                       low.dy, low.y);
    UCHK(hipGetLastError());
  }
  for (int l = Lv - 1; l >= 0; l--) {
    const int h = H0 >> l;
    RCHK(conv_wgrad(u, u->conv[enc_b(l)], B, h, u->e[l].dy, u->ea[l].y));
    RCHK(conv_dgrad(u, u->conv[enc_b(l)], B, h, u->e[l].dy, u->ea[l].dy, u->ea[l].y));
    const u16* x = l == 0 ? u->in0.y : u->p[l].y;
    RCHK(conv_wgrad(u, u->conv[enc_a(l)], B, h, u->ea[l].dy, x));
    if (l > 0) {
      RCHK(conv_dgrad(u, u->conv[enc_a(l)], B, h, u->ea[l].dy, u->p[l].dy, nullptr));
      const int hp = h * 2;
      const int64_t n8 = (int64_t)B * hp * hp * u->e[l - 1].c / 8;
      dim3 g;
      RCHK(ew8_grid(n8, g));
      hipLaunchKernelGGL(k_pool_bwd_add, g, dim3(256), 0, u->st, (int)n8, hp, hp, u->e[l - 1].c, u->p[l].dy,
                         u->e[l - 1].dy, u->e[l - 1].y);
      UCHK(hipGetLastError());
    }
  }
  return CDX_OK;
}

int loss_grad(cdx_unet* u, int B, double* loss) {
  const int M = B * u->size * u->size;
  const int nb = 256;  // fixed grid: partial sums in dscal[0..nb)
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, u->st, M, u->yf, u->tg, 0.5f, 1.0f, u->dz_head, u->dscal);
  UCHK(hipGetLastError());
  if (loss) {
    std::vector<double> h(nb);
    UCHK(hipMemcpyAsync(h.data(), u->dscal, nb * sizeof(double), hipMemcpyDeviceToHost, u->st));
    UCHK(hipStreamSynchronize(u->st));
    double s = 0;
    for (double v : h) s += v;
    *loss = s / M;
  }
  return CDX_OK;
}

int adam(cdx_unet* u, double lr) {
  u->adam_t++;
  const double b1 = 0.9, b2 = 0.999;
  const float c1 = (float)(1 - pow(b1, (double)u->adam_t)), c2 = (float)(1 - pow(b2, (double)u->adam_t));
  hipLaunchKernelGGL(k_adam, dim3(ew_blocks(u->nparam)), dim3(256), 0, u->st, u->nparam, u->w, u->g, u->m1, u->m2,
                     u->wb, (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2);
  UCHK(hipGetLastError());
  return flip_weights(u);
}

// smoothing of the synthetic objects' fields, in pixels
constexpr double kSynthSigma = 3.0;
// forward-model sample: shape, photons and seeds drawn from `seed`
struct SampleSpec {
  int kind, np;
  double prm[16];
  double photons;
};
SampleSpec draw_sample(uint64_t seed, double lo, double hi) {
  SampleSpec s{};
  auto U = [&](int k) { return uniform_host(seed, 10, (uint64_t)k); };
  s.kind = U(0) < 0.5 ? 1 : 2;
  if (s.kind == 1) {
    s.np = 3;
    s.prm[0] = 20 + 40 * U(1);
    s.prm[1] = 20 + 40 * U(2);
    s.prm[2] = M_PI * U(3);
  } else {
    const int nv = 3 + (int)(6 * U(1));
    s.np = 2 * nv;
    for (int i = 0; i < nv; i++) {
      s.prm[i] = 2 * M_PI * (i + 0.4 + 0.2 * U(2 + i)) / nv;
      s.prm[nv + i] = 20 + 40 * U(10 + i);
    }
  }
  s.photons = pow(10.0, log10(lo) + (log10(hi) - log10(lo)) * U(20));
  return s;
}

int synth_frame(cdx_unet* u, uint64_t seed, double lo, double hi, void* obj, void* I) {
  const SampleSpec s = draw_sample(seed, lo, hi);
  RCHK(cdx_sim_object_shape(u->ctx, s.kind, s.prm, s.np, seed, kSynthSigma, obj));
  RCHK(cdx_sim_intensity(u->ctx, obj, s.photons, seed, 1, I));
  return CDX_OK;
}

// the support shape of a drawn sample (the parameters cdx_sim_object_shape takes)
ShapeParams sample_shape(const SampleSpec& s) {
  ShapeParams sp{};
  sp.kind = s.kind;
  if (s.kind == 1) {
    sp.a = (float)s.prm[0];
    sp.b = (float)s.prm[1];
    sp.rot = (float)s.prm[2];
  } else {
    sp.nv = s.np / 2;
    for (int i = 0; i < sp.nv; i++) {
      sp.ang[i] = (float)s.prm[i];
      sp.rad[i] = (float)s.prm[sp.nv + i];
    }
  }
  return sp;
}

// Batch entries b of seeds seed0 + b.  Frames go through the forward model
// synth_chunk at a time side by side (the same values as one frame at a
// time, in a few launches per chunk instead of ~17 per frame).
int synth_batch(cdx_unet* u, int B, uint64_t seed0, double lo, double hi) {
  const int n = u->size;
  const int64_t N = (int64_t)n * n;
  if (u->synth_chunk > 0) {
    for (int b0 = 0; b0 < B; b0 += u->synth_chunk) {
      SynthFrames f{};
      f.B = B - b0 < u->synth_chunk ? B - b0 : u->synth_chunk;
      for (int j = 0; j < f.B; j++) {
        const uint64_t seed = seed0 + (uint64_t)(b0 + j);
        const SampleSpec spec = draw_sample(seed, lo, hi);
        f.seed[j] = seed;
        f.photons[j] = spec.photons;
        f.sp[j] = sample_shape(spec);
      }
      const size_t fN = (size_t)u->synth_chunk * N;
      double* a = u->syn;
      double* mm = u->syn + 4 * fN;
      double* tot = mm + 4 * (size_t)u->synth_chunk;
      double* cs = tot + u->synth_chunk;
      RCHK(sim_stack(u->ctx, f, kSynthSigma, a, a + fN, a + 2 * fN, a + 3 * fN, mm, tot, u->sobj, u->sF, u->sI));
      const int64_t tot_px = (int64_t)f.B * N;
      hipLaunchKernelGGL(k_pack_input_stack, dim3(ew_blocks(tot_px)), dim3(256), 0, u->st, n, f.B, b0,
                         (const float*)u->sI, u->in0.y);
      UCHK(hipGetLastError());
      hipLaunchKernelGGL(k_csum_stack, dim3(f.B), dim3(1024), 0, u->st, n, f.B, (const cplx<float>*)u->sobj, cs);
      UCHK(hipGetLastError());
      hipLaunchKernelGGL(k_pack_target_stack, dim3(ew_blocks(tot_px)), dim3(256), 0, u->st, n, f.B, b0,
                         (const cplx<float>*)u->sobj, (const double*)cs, u->tg);
      UCHK(hipGetLastError());
    }
    return CDX_OK;
  }
  for (int b = 0; b < B; b++) {
    RCHK(synth_frame(u, seed0 + (uint64_t)b, lo, hi, u->obj, u->I));
    hipLaunchKernelGGL(k_pack_input, dim3(ew_blocks(N)), dim3(256), 0, u->st, n, (const float*)u->I,
                       u->in0.y + (size_t)b * N * 8);
    UCHK(hipGetLastError());
    hipLaunchKernelGGL(k_csum, dim3(1), dim3(1024), 0, u->st, N, (const cplx<float>*)u->obj, u->dscal);
    UCHK(hipGetLastError());
    hipLaunchKernelGGL(k_pack_target, dim3(ew_blocks(N)), dim3(256), 0, u->st, N, (const cplx<float>*)u->obj,
                       (const double*)u->dscal, u->tg + (size_t)b * N * 4);
    UCHK(hipGetLastError());
  }
  return CDX_OK;
}

void destroy(cdx_unet* u) {
  if (!u) return;
  if (u->st) hipStreamSynchronize(u->st);
  for (auto& L : u->conv) hipFree(L.wflip);
  for (void* p : {(void*)u->w, (void*)u->g, (void*)u->m1, (void*)u->m2, (void*)u->wb, (void*)u->yf,
                  (void*)u->dz_head, (void*)u->tg, (void*)u->dscal, (void*)u->wg_part, u->obj, u->I, (void*)u->syn,
                  u->sobj, u->sF, (void*)u->sI})
    hipFree(p);
  free_tensor(u->in0);
  for (auto* v : {&u->ea, &u->e, &u->p, &u->cat, &u->da, &u->d})
    for (auto& t : *v) free_tensor(t);
  if (u->ctx) cdx_destroy(u->ctx);
  delete u;
}

}  // namespace

#define UNET_CHECK(u) \
  if (!(u)) return ufail(CDX_ERR_ARG, "%s: null network", __func__)

extern "C" int cdx_unet_create(cdx_unet** out, int size, int levels, int base, int max_batch, uint64_t seed) {
  if (!out) return ufail(CDX_ERR_ARG, "cdx_unet_create: null out pointer");
  *out = nullptr;
  if (levels < 1 || levels > 6 || base < 8 || base % 8 || max_batch < 1 || size < 16 || (size >> (levels - 1)) < 2 ||
      (size >> (levels - 1)) << (levels - 1) != size || !cdx_size_supported(size))
    return ufail(CDX_ERR_ARG, "cdx_unet_create: unsupported geometry size %d levels %d base %d batch %d", size,
                 levels, base, max_batch);
  cdx_unet* u = new cdx_unet();
  u->size = size;
  u->levels = levels;
  u->base = base;
  u->max_batch = max_batch;
  int rc = cdx_create(&u->ctx, size, size, CDX_F32, nullptr);
  if (rc) {
    delete u;
    return rc;
  }
  u->st = (hipStream_t)cdx_stream(u->ctx);
  // layer table
  auto add = [&](int cin, int cout, int taps, int relu) {
    UConv L{};
    L.cin = cin;
    L.cout = cout;
    L.taps = taps;
    L.relu = relu;
    L.kp = (taps * cin + kBK - 1) / kBK * kBK;
    L.kp2 = (taps * cout + kBK - 1) / kBK * kBK;
    L.w_off = u->nparam;
    u->nparam += (int64_t)cout * L.kp;
    L.b_off = u->nparam;
    u->nparam += (cout + 7) / 8 * 8;
    u->conv.push_back(L);
  };
  for (int l = 0; l < levels; l++) {
    const int c = base << l;
    add(l == 0 ? 8 : (base << (l - 1)), c, 9, 1);
    add(c, c, 9, 1);
  }
  for (int l = 0; l + 1 < levels; l++) {
    const int c = base << l;
    add((base << (l + 1)) + c, c, 9, 1);
    add(c, c, 9, 1);
  }
  add(base, 8, 1, 0);
  const size_t pb = (size_t)u->nparam * sizeof(float);
  const int B = max_batch;
  const int64_t M = (int64_t)B * size * size;
  bool ok = hipMalloc(&u->w, pb) == hipSuccess && hipMalloc(&u->g, pb) == hipSuccess &&
            hipMalloc(&u->m1, pb) == hipSuccess && hipMalloc(&u->m2, pb) == hipSuccess &&
            hipMalloc(&u->wb, (size_t)u->nparam * sizeof(u16)) == hipSuccess &&
            hipMalloc(&u->yf, (size_t)M * 8 * sizeof(float)) == hipSuccess &&
            hipMalloc(&u->dz_head, (size_t)M * 8 * sizeof(u16)) == hipSuccess &&
            hipMalloc(&u->tg, (size_t)M * 4 * sizeof(float)) == hipSuccess &&
            hipMalloc(&u->dscal, 4096 * sizeof(double)) == hipSuccess &&
            hipMalloc(&u->wg_part, kWgradPartFloats * sizeof(float)) == hipSuccess &&
            hipMalloc(&u->obj, (size_t)size * size * 8) == hipSuccess &&
            hipMalloc(&u->I, (size_t)size * size * 4) == hipSuccess;
  if (ok && synth_stack_supported(size, kSynthSigma)) {
    const int ch = max_batch < kSynthMax ? max_batch : kSynthMax;
    const size_t fN = (size_t)ch * size * size;
    ok = hipMalloc(&u->syn, (4 * fN + 7 * (size_t)ch) * sizeof(double)) == hipSuccess &&
         hipMalloc(&u->sobj, fN * 8) == hipSuccess && hipMalloc(&u->sF, fN * 8) == hipSuccess &&
         hipMalloc((void**)&u->sI, fN * 4) == hipSuccess;
    u->synth_chunk = ch;
  }
  for (auto& L : u->conv) {
    ok = ok && hipMalloc(&L.wflip, (size_t)L.cin * L.kp2 * sizeof(u16)) == hipSuccess;
    if (ok) {  // the K padding of the flipped weights stays zero
      const int64_t nf = (int64_t)L.cin * L.kp2;
      hipLaunchKernelGGL(k_zero_u16, dim3(ew_blocks(nf)), dim3(256), 0, u->st, nf, L.wflip);
    }
  }
  u->ea.resize(levels);
  u->e.resize(levels);
  u->p.resize(levels);
  u->cat.resize(levels);
  u->da.resize(levels);
  u->d.resize(levels);
  ok = ok && alloc_tensor(u->in0, B, size, 8) == CDX_OK;
  for (int l = 0; l < levels && ok; l++) {
    const int h = size >> l, c = base << l;
    ok = alloc_tensor(u->ea[l], B, h, c) == CDX_OK && alloc_tensor(u->e[l], B, h, c) == CDX_OK;
    if (ok && l > 0) ok = alloc_tensor(u->p[l], B, h, base << (l - 1)) == CDX_OK;
    if (ok && l + 1 < levels)
      ok = alloc_tensor(u->cat[l], B, h, (base << (l + 1)) + c) == CDX_OK &&
           alloc_tensor(u->da[l], B, h, c) == CDX_OK && alloc_tensor(u->d[l], B, h, c) == CDX_OK;
  }
  if (!ok) {
    destroy(u);
    return ufail(CDX_ERR_HIP, "cdx_unet_create: out of device memory");
  }
  // He-normal weights (fan-in of the real input channels), zero biases and
  // padding; the head starts small
  std::vector<float> h((size_t)u->nparam, 0.f);
  uint64_t ctr = 0;
  auto normal = [&]() {
    const double u1 = uniform_host(seed, 11, ctr++), u2 = uniform_host(seed, 11, ctr++);
    return sqrt(-2 * log(u1 + 1e-300)) * cos(2 * M_PI * u2);
  };
  for (size_t i = 0; i < u->conv.size(); i++) {
    const UConv& L = u->conv[i];
    const int cin_real = i == 0 ? 1 : L.cin;
    const bool is_head = i + 1 == u->conv.size();
    const double sd = is_head ? 0.1 / sqrt((double)L.cin) : sqrt(2.0 / (L.taps * cin_real));
    for (int co = 0; co < L.cout; co++)
      for (int t = 0; t < L.taps; t++)
        for (int ci = 0; ci < L.cin; ci++)
          h[L.w_off + (size_t)co * L.kp + t * L.cin + ci] = (i == 0 && ci > 0) ? 0.f : (float)(sd * normal());
  }
  rc = cdx_unet_set_params(u, h.data(), u->nparam);
  if (!rc) {
    hipLaunchKernelGGL(k_zero_f32, dim3(ew_blocks(u->nparam)), dim3(256), 0, u->st, u->nparam, u->g);
    hipLaunchKernelGGL(k_zero_f32, dim3(ew_blocks(u->nparam)), dim3(256), 0, u->st, u->nparam, u->m1);
    hipLaunchKernelGGL(k_zero_f32, dim3(ew_blocks(u->nparam)), dim3(256), 0, u->st, u->nparam, u->m2);
    hipLaunchKernelGGL(k_zero_u16, dim3(ew_blocks(M * 8)), dim3(256), 0, u->st, M * 8, u->in0.y);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(u->st) != hipSuccess)
      rc = ufail(CDX_ERR_HIP, "cdx_unet_create: initialisation failed");
  }
  if (rc) {
    destroy(u);
    return rc;
  }
  *out = u;
  return CDX_OK;
}

extern "C" int cdx_unet_destroy(cdx_unet* u) {
  destroy(u);
  return CDX_OK;
}

extern "C" int cdx_unet_num_params(cdx_unet* u, int64_t* n) {
  UNET_CHECK(u);
  if (n) *n = u->nparam;
  return CDX_OK;
}

extern "C" int cdx_unet_layers(cdx_unet* u, int64_t* info, int cap) {
  UNET_CHECK(u);
  const int nl = (int)u->conv.size();
  for (int i = 0; i < nl && 6 * (i + 1) <= cap; i++) {
    const UConv& L = u->conv[i];
    int64_t* o = info + 6 * i;
    o[0] = L.cin;
    o[1] = L.cout;
    o[2] = L.taps;
    o[3] = L.kp;
    o[4] = L.w_off;
    o[5] = L.b_off;
  }
  return nl;
}

extern "C" int cdx_unet_set_params(cdx_unet* u, const float* host, int64_t n) {
  UNET_CHECK(u);
  if (!host || n != u->nparam) return ufail(CDX_ERR_ARG, "cdx_unet_set_params: expected %lld values", (long long)u->nparam);
  UCHK(hipMemcpyAsync(u->w, host, (size_t)n * sizeof(float), hipMemcpyHostToDevice, u->st));
  RCHK(refresh_weights(u));
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}
extern "C" int cdx_unet_get_params(cdx_unet* u, float* host, int64_t n) {
  UNET_CHECK(u);
  if (!host || n != u->nparam) return ufail(CDX_ERR_ARG, "cdx_unet_get_params: expected %lld values", (long long)u->nparam);
  UCHK(hipMemcpyAsync(host, u->w, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}
extern "C" int cdx_unet_get_grads(cdx_unet* u, float* host, int64_t n) {
  UNET_CHECK(u);
  if (!host || n != u->nparam) return ufail(CDX_ERR_ARG, "cdx_unet_get_grads: expected %lld values", (long long)u->nparam);
  UCHK(hipMemcpyAsync(host, u->g, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}
extern "C" int cdx_unet_zero_grads(cdx_unet* u) {
  UNET_CHECK(u);
  hipLaunchKernelGGL(k_zero_f32, dim3(ew_blocks(u->nparam)), dim3(256), 0, u->st, u->nparam, u->g);
  UCHK(hipGetLastError());
  return CDX_OK;
}

// host batch: x [B][size][size] (network input channel 0, already scaled),
// targets [B][size][size][3] (amplitude, phase, mask) or NULL
extern "C" int cdx_unet_set_batch(cdx_unet* u, int B, const float* x, const float* targets) {
  UNET_CHECK(u);
  if (B < 1 || B > u->max_batch || !x) return ufail(CDX_ERR_ARG, "cdx_unet_set_batch: bad batch");
  const size_t P = (size_t)B * u->size * u->size;
  std::vector<u16> xi(P * 8, 0);
  for (size_t i = 0; i < P; i++) xi[i * 8] = f2bf(x[i]);
  UCHK(hipMemcpyAsync(u->in0.y, xi.data(), xi.size() * sizeof(u16), hipMemcpyHostToDevice, u->st));
  if (targets) {
    std::vector<float> t(P * 4, 0.f);
    for (size_t i = 0; i < P; i++)
      for (int c = 0; c < 3; c++) t[i * 4 + c] = targets[i * 3 + c];
    UCHK(hipMemcpyAsync(u->tg, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, u->st));
  }
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}
// current input / targets back to the host (same layouts as cdx_unet_set_batch)
extern "C" int cdx_unet_get_batch(cdx_unet* u, int B, float* x, float* targets) {
  UNET_CHECK(u);
  if (B < 1 || B > u->max_batch) return ufail(CDX_ERR_ARG, "cdx_unet_get_batch: bad batch");
  const size_t P = (size_t)B * u->size * u->size;
  if (x) {
    std::vector<u16> xi(P * 8);
    UCHK(hipMemcpyAsync(xi.data(), u->in0.y, xi.size() * sizeof(u16), hipMemcpyDeviceToHost, u->st));
    UCHK(hipStreamSynchronize(u->st));
    for (size_t i = 0; i < P; i++) x[i] = bf2f(xi[i * 8]);
  }
  if (targets) {
    std::vector<float> t(P * 4);
    UCHK(hipMemcpyAsync(t.data(), u->tg, t.size() * sizeof(float), hipMemcpyDeviceToHost, u->st));
    UCHK(hipStreamSynchronize(u->st));
    for (size_t i = 0; i < P; i++)
      for (int c = 0; c < 3; c++) targets[i * 3 + c] = t[i * 4 + c];
  }
  return CDX_OK;
}

extern "C" int cdx_unet_forward(cdx_unet* u, int B) {
  UNET_CHECK(u);
  if (B < 1 || B > u->max_batch) return ufail(CDX_ERR_ARG, "cdx_unet_forward: bad batch");
  RCHK(forward(u, B));
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}
// head output [B][size][size][3]: amplitude, phase, support logit
extern "C" int cdx_unet_output(cdx_unet* u, int B, float* host) {
  UNET_CHECK(u);
  if (B < 1 || B > u->max_batch || !host) return ufail(CDX_ERR_ARG, "cdx_unet_output: bad batch");
  const size_t P = (size_t)B * u->size * u->size;
  std::vector<float> y(P * 8);
  UCHK(hipMemcpyAsync(y.data(), u->yf, y.size() * sizeof(float), hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  for (size_t i = 0; i < P; i++)
    for (int c = 0; c < 3; c++) host[i * 3 + c] = y[i * 8 + c];
  return CDX_OK;
}
// loss of the last forward against the targets, gradients added to the
// parameter gradients (cleared by cdx_unet_zero_grads and by each Adam step)
extern "C" int cdx_unet_backward(cdx_unet* u, int B, double* loss) {
  UNET_CHECK(u);
  if (B < 1 || B > u->max_batch) return ufail(CDX_ERR_ARG, "cdx_unet_backward: bad batch");
  RCHK(loss_grad(u, B, loss));
  RCHK(backward(u, B));
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}
extern "C" int cdx_unet_adam(cdx_unet* u, double lr) {
  UNET_CHECK(u);
  RCHK(adam(u, lr));
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}

extern "C" int cdx_unet_synth_batch(cdx_unet* u, int B, uint64_t seed0, double photons_lo, double photons_hi) {
  UNET_CHECK(u);
  if (B < 1 || B > u->max_batch || !(photons_lo > 0) || !(photons_hi >= photons_lo))
    return ufail(CDX_ERR_ARG, "cdx_unet_synth_batch: bad arguments");
  RCHK(synth_batch(u, B, seed0, photons_lo, photons_hi));
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}
extern "C" int cdx_unet_synth_frame(cdx_unet* u, uint64_t seed, double photons_lo, double photons_hi, void* obj,
                                    void* I) {
  UNET_CHECK(u);
  if (!obj || !I) return ufail(CDX_ERR_ARG, "cdx_unet_synth_frame: null buffer");
  return synth_frame(u, seed, photons_lo, photons_hi, obj, I);
}

extern "C" int cdx_unet_train(cdx_unet* u, int steps, int B, uint64_t seed0, double lr, double photons_lo,
                              double photons_hi, double* losses) {
  UNET_CHECK(u);
  if (steps < 0 || B < 1 || B > u->max_batch) return ufail(CDX_ERR_ARG, "cdx_unet_train: bad arguments");
  for (int s = 0; s < steps; s++) {
    RCHK(synth_batch(u, B, seed0 + (uint64_t)s * B, photons_lo, photons_hi));
    RCHK(forward(u, B));
    double l = 0;
    RCHK(loss_grad(u, B, losses ? &l : nullptr));
    RCHK(backward(u, B));
    RCHK(adam(u, lr));
    if (losses) losses[s] = l;
  }
  UCHK(hipStreamSynchronize(u->st));
  return CDX_OK;
}

// warm start of a reconstruction on `ctx` (size x size, FP32) from the
// measured pattern I (FFT order): F = FFT(c g0), g = IFFT(F)/sqrt(N) (the
// state CDI::prepareIter leaves), spt = predicted support (0/1, centred),
// with c scaling g0 to the measured total intensity (Parseval)
extern "C" int cdx_unet_warm_start(cdx_unet* u, cdx_ctx* ctx, const void* I, double support_threshold, void* F,
                                   void* g, void* spt) {
  UNET_CHECK(u);
  if (!ctx || !I || !F || !g || !spt) return ufail(CDX_ERR_ARG, "cdx_unet_warm_start: null argument");
  int rows = 0, cols = 0, prec = 0;
  RCHK(cdx_context_info(ctx, &rows, &cols, &prec));
  if (rows != u->size || cols != u->size || prec != CDX_F32)
    return ufail(CDX_ERR_ARG, "cdx_unet_warm_start: context must be %dx%d FP32", u->size, u->size);
  const int n = u->size;
  const int64_t N = (int64_t)n * n;
  RCHK(cdx_synchronize(ctx));
  hipLaunchKernelGGL(k_pack_input, dim3(ew_blocks(N)), dim3(256), 0, u->st, n, (const float*)I, u->in0.y);
  UCHK(hipGetLastError());
  RCHK(forward(u, 1));
  hipLaunchKernelGGL(k_ws_compose, dim3(ew_blocks(N)), dim3(256), 0, u->st, N, (const float*)u->yf,
                     (float)support_threshold, (cplx<float>*)g, (float*)spt);
  UCHK(hipGetLastError());
  hipLaunchKernelGGL(k_sum_sq, dim3(1), dim3(1024), 0, u->st, N, (const cplx<float>*)g, (const float*)I, u->dscal);
  UCHK(hipGetLastError());
  double sums[2];
  UCHK(hipMemcpyAsync(sums, u->dscal, sizeof sums, hipMemcpyDeviceToHost, u->st));
  UCHK(hipStreamSynchronize(u->st));
  const double c = sums[0] > 0 ? sqrt(sums[1] / ((double)N * sums[0])) : 0.0;
  RCHK(cdx_apply_norm(ctx, g, c));
  RCHK(cdx_fft(ctx, g, F, CDX_FORWARD));
  RCHK(cdx_apply_norm(ctx, g, sqrt((double)N)));
  RCHK(cdx_synchronize(ctx));
  return CDX_OK;
}

// ---------------------------------------------------------------------------
// warm start of a frame larger than the network: the network sees the
// pattern's central ns x ns frequencies (FFT order), i.e. the object at
// f = nl/ns times coarser pixels over the same field of view; its phases seed
// those frequencies of the large frame, the rest keep a random start, and
// its support is upsampled (nearest) about the centre.
// ---------------------------------------------------------------------------
// signed frequency of FFT-order index u on an n-point axis, wrapped onto m points
__device__ __forceinline__ int band_index(int u, int n, int m) { return u < n / 2 ? u : u - n + m; }

__global__ void k_crop_band(int nl, int ns, const float* __restrict__ Il, float* __restrict__ Is) {
  const int64_t N = (int64_t)ns * ns;
  EW_LOOP(N) {
    const int u = (int)(i / ns), v = (int)(i % ns);
    Is[i] = Il[(size_t)band_index(u, ns, nl) * nl + band_index(v, ns, nl)];
  }
}
// F_l (random start: sqrt(I) e^{i phi}, 0 in the beamstop) takes the phase of
// F_s at the band's frequencies where F_s != 0; its modulus is kept
__global__ void k_band_phase(int nl, int ns, const cplx<float>* __restrict__ Fs, cplx<float>* __restrict__ Fl) {
  const int64_t N = (int64_t)ns * ns;
  EW_LOOP(N) {
    const int u = (int)(i / ns), v = (int)(i % ns);
    const cplx<float> s = Fs[i];
    const float m2 = s.x * s.x + s.y * s.y;
    if (!(m2 > 0.f)) continue;
    cplx<float>* d = Fl + (size_t)band_index(u, ns, nl) * nl + band_index(v, ns, nl);
    const cplx<float> r = *d;
    const float a = sqrtf(r.x * r.x + r.y * r.y) * rsqrtf(m2);
    *d = {s.x * a, s.y * a};
  }
}
// centred nearest upsampling by f = nl/ns: (X, Y) -> (x, y) = (X/f, Y/f)
// keeps the centre nl/2 on ns/2; with dil > 0 the small support is first
// dilated by a disk of radius dil (small-frame pixels: (x, y) is in when some
// (x + dx, y + dy) with dx^2 + dy^2 <= dil^2 inside the frame is)
__global__ void k_upsample_support(int nl, int ns, int dil, const float* __restrict__ ss, float* __restrict__ sl) {
  const int64_t N = (int64_t)nl * nl;
  const int f = nl / ns;
  EW_LOOP(N) {
    const int x = (int)(i / nl) / f, y = (int)(i % nl) / f;
    float v = ss[(size_t)x * ns + y];
    for (int dx = -dil; dx <= dil && v == 0.f; dx++)
      for (int dy = -dil; dy <= dil; dy++) {
        const int xx = x + dx, yy = y + dy;
        if (dx * dx + dy * dy > dil * dil || xx < 0 || xx >= ns || yy < 0 || yy >= ns) continue;
        if (ss[(size_t)xx * ns + yy] != 0.f) {
          v = 1.f;
          break;
        }
      }
    sl[i] = v;
  }
}

extern "C" int cdx_unet_warm_start_scaled(cdx_unet* u, cdx_ctx* ctx_small, cdx_ctx* ctx, const void* I,
                                          const void* bs, double support_threshold, int dilate, uint64_t seed,
                                          void* F, void* g, void* spt) {
  UNET_CHECK(u);
  if (!ctx_small || !ctx || !I || !F || !g || !spt)
    return ufail(CDX_ERR_ARG, "cdx_unet_warm_start_scaled: null argument");
  int rows = 0, cols = 0, prec = 0, rs = 0, cs = 0, ps = 0;
  RCHK(cdx_context_info(ctx, &rows, &cols, &prec));
  RCHK(cdx_context_info(ctx_small, &rs, &cs, &ps));
  if (rs != u->size || cs != u->size || ps != CDX_F32)
    return ufail(CDX_ERR_ARG, "cdx_unet_warm_start_scaled: small context must be %dx%d FP32", u->size, u->size);
  if (dilate < 0 || dilate > 16) return ufail(CDX_ERR_ARG, "cdx_unet_warm_start_scaled: dilate %d outside 0..16", dilate);
  if (rows != cols || prec != CDX_F32 || rows < rs || rows % rs)
    return ufail(CDX_ERR_ARG, "cdx_unet_warm_start_scaled: frame %dx%d must be square FP32, a multiple of %d", rows,
                 cols, rs);
  const int nl = rows, ns = rs;
  const int64_t Ns = (int64_t)ns * ns, Nl = (int64_t)nl * nl;
  // small-frame pattern and support (float), prediction F and g (complex)
  void* buf = nullptr;
  UCHK(hipMalloc(&buf, (size_t)Ns * (2 * sizeof(float) + 2 * sizeof(cplx<float>))));
  float* Is = (float*)buf;
  float* ss = Is + Ns;
  cplx<float>* Fs = (cplx<float>*)(ss + Ns);
  cplx<float>* gs = Fs + Ns;
  int rc = CDX_OK;
  do {
    if ((rc = cdx_synchronize(ctx))) break;
    hipLaunchKernelGGL(k_crop_band, dim3(ew_blocks(Ns)), dim3(256), 0, u->st, nl, ns, (const float*)I, Is);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(u->st) != hipSuccess) {
      rc = ufail(CDX_ERR_HIP, "cdx_unet_warm_start_scaled: crop failed");
      break;
    }
    if ((rc = cdx_unet_warm_start(u, ctx_small, Is, support_threshold, Fs, gs, ss))) break;
    if ((rc = cdx_random_start(ctx, I, bs, seed, 1, F))) break;
    if ((rc = cdx_synchronize(ctx))) break;
    hipLaunchKernelGGL(k_band_phase, dim3(ew_blocks(Ns)), dim3(256), 0, u->st, nl, ns, (const cplx<float>*)Fs,
                       (cplx<float>*)F);
    hipLaunchKernelGGL(k_upsample_support, dim3(ew_blocks(Nl)), dim3(256), 0, u->st, nl, ns, dilate,
                       (const float*)ss, (float*)spt);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(u->st) != hipSuccess) {
      rc = ufail(CDX_ERR_HIP, "cdx_unet_warm_start_scaled: band transfer failed");
      break;
    }
    // g = IFFT(F)/sqrt(N), the state CDI::prepareIter leaves
    if ((rc = cdx_fft(ctx, F, g, CDX_INVERSE))) break;
    if ((rc = cdx_apply_norm(ctx, g, 1.0 / sqrt((double)Nl)))) break;
    rc = cdx_synchronize(ctx);
  } while (0);
  hipFree(buf);
  return rc;
}

// parameters as a small binary file: "CDXUNET1", size, levels, base, count, FP32 values
extern "C" int cdx_unet_save(cdx_unet* u, const char* path) {
  UNET_CHECK(u);
  std::vector<float> h((size_t)u->nparam);
  RCHK(cdx_unet_get_params(u, h.data(), u->nparam));
  FILE* f = path ? fopen(path, "wb") : nullptr;
  if (!f) return ufail(CDX_ERR_IO, "cdx_unet_save: cannot open %s", path ? path : "(null)");
  const int32_t hdr[3] = {u->size, u->levels, u->base};
  const int64_t n = u->nparam;
  bool ok = fwrite("CDXUNET1", 1, 8, f) == 8 && fwrite(hdr, sizeof hdr, 1, f) == 1 && fwrite(&n, 8, 1, f) == 1 &&
            fwrite(h.data(), sizeof(float), h.size(), f) == h.size();
  ok = (fclose(f) == 0) && ok;
  return ok ? CDX_OK : ufail(CDX_ERR_IO, "cdx_unet_save: write failed on %s", path);
}
extern "C" int cdx_unet_load(cdx_unet* u, const char* path) {
  UNET_CHECK(u);
  FILE* f = path ? fopen(path, "rb") : nullptr;
  if (!f) return ufail(CDX_ERR_IO, "cdx_unet_load: cannot open %s", path ? path : "(null)");
  char magic[8];
  int32_t hdr[3];
  int64_t n = 0;
  bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "CDXUNET1", 8) == 0 && fread(hdr, sizeof hdr, 1, f) == 1 &&
            fread(&n, 8, 1, f) == 1;
  if (!ok || hdr[0] != u->size || hdr[1] != u->levels || hdr[2] != u->base || n != u->nparam) {
    fclose(f);
    return ufail(CDX_ERR_IO, "cdx_unet_load: %s does not hold a %d/%d/%d network", path, u->size, u->levels, u->base);
  }
  std::vector<float> h((size_t)n);
  ok = fread(h.data(), sizeof(float), h.size(), f) == h.size();
  fclose(f);
  if (!ok) return ufail(CDX_ERR_IO, "cdx_unet_load: short read on %s", path);
  return cdx_unet_set_params(u, h.data(), n);
}
