// kernels.hpp -- launch interface of the CDI hot-path kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cdx {

// Projection modes of the fused column pass.
enum ProjMode : int {
  PROJ_MOD_ABS = 0,     // applyModAbs, src/gpu/cudaConfig.cu:827-841
  PROJ_MOD_NOISE = 1,   // applyMod, src/gpu/cdi.cu:438-485 (beamstop as NaN sentinel)
  PROJ_MOD_ABS_BS = 2,  // applyModAbs, beamstop pixels left unconstrained
  // ptychography (util/ptycho.cpp:539-548): F *= bsnorm first, intensity packed, NaN on the beamstop
  PROJ_PT_ACCURATE = 3,  // applyModAccurate, src/gpu/cdi.cu:351-388
  PROJ_PT_NOISE = 4,     // applyMod, src/gpu/cdi.cu:438-485
};

// Column-pass operations (one workgroup owns C whole columns).
enum ColOp : int {
  COL_FWD = 0,            // Din load -> forward -> Dout store
  COL_INV = 1,            // Dout load -> inverse -> Din store
  COL_FWD_PROJ_INV = 2,   // forward -> [R-factor] -> projection -> inverse
  COL_PROJ_INV = 3,       // input already in Fourier space
  COL_FWD_GAUSS_INV = 4,  // forward -> Gaussian multiply -> inverse (shrink-wrap)
  COL_GAUSS_INV = 5,
  COL_FWD_T1 = 6,         // COL_FWD reading the loop state T1 in its column-tile layout
};

// The loop state T1 = rowFFT(g) of a single-frame context is stored in
// column tiles: element (r, c) of a rows x cols image at
// (c / kT1Tile) * rows * kT1Tile + r * kT1Tile + c % kT1Tile.  The row pass
// writes 8-element (64-byte) pieces, so a workgroup's two rows fill whole
// 128-byte lines; a 4-column workgroup of the column pass reads half of every
// piece of its tile, and the workgroup holding the other half runs on the same
// XCD (xcd_tile).  The packed measurement, written once per loop, is kept in
// tiles of kITile columns, each a 4-column workgroup's contiguous block.
// 1024^2 loop, against T1 in 4-column tiles (three A/B pairs): 61.3k against
// 59.6-59.8k iterations/s (8-column measurement tiles: 60.7k; 16-column T1
// tiles: 58.4k).  Every supported line length is a multiple of both widths.
constexpr int kT1Tile = 8;
constexpr int kITile = 4;

// Row-pass operations (one workgroup owns R whole rows).
enum RowOp : int {
  ROW_FWD = 0,
  ROW_INV = 1,            // optional scale
  ROW_INV_UPDATE_FWD = 2, // inverse -> 1/N -> support update -> store g -> forward
  ROW_INV_MOD2 = 3,       // inverse -> |.|^2 -> S, max and sum (shrink-wrap)
  ROW_INV_UPDATE_FWD_PRTF = 4,  // ROW_INV_UPDATE_FWD with the PRTF phase-alignment sums
};

// R-factor sums of one projection: kRfSlots (num, den) pairs; workgroup b
// adds into pair b mod kRfSlots, readers add the pairs.  Rings of these sums
// are slot-major (pair k of entry i at ring[k * rfs + 2 i], rfs = 2 x entries),
// so the pairs of one projection lie on distinct cache lines: double atomics
// to one line serialize, and with all pairs of a projection on one or two
// lines they cost about as much as the rest of the pass's tail.
constexpr int kRfSlots = 16;

template <class T>
struct ProjArgs {
  int mode;
  int noiseLevel;
  T scale;   // cudaVars.scale = 1/exposure
  T rcolor;  // cudaVars.rcolor
  double* rf;  // R-factor accumulators (kRfSlots (num, den) pairs, rfs doubles apart) or nullptr
  int rfs;
  void* fsave;  // if set (run-time-mode kernels only): projected F stored here
  T bsnorm;     // PROJ_PT_* modes: F *= bsnorm before the projection
};

template <class T>
struct UpdArgs {
  int algo;
  T beta, relax, invN;
  const double* scal;  // scal[0] = support threshold
  void* gamma0;        // PRTF reference object P_S(g') (ROW_INV_UPDATE_FWD_PRTF)
  int prtf;            // 1: store gamma0, sum gamma0^2; 2: sum P_S(g') conj(gamma0)
  // support test S > threshold of the single-frame update pass, one word per
  // (row, lane): bit i is the lane's i-th pixel (launch_support_mask)
  const unsigned* smask;
};

// Deterministic reductions: every workgroup of the producing kernel stores
// its partial sum at partials[blockIdx.x] (workgroup 0 stores the count);
// the consumer adds them in a fixed order.  The max is an atomic on the bit
// pattern (order-independent for non-negative values).
struct ReduceSlots {
  void* maxbits;     // uint32 (float) or uint64 (double) bit pattern of the max
  double* partials;  // one entry per workgroup (capacity kMaxPartials)
  int* nparts;
};
constexpr int kMaxPartials = 4096;

// line lengths of the FFT
int line_supported(int n);
}  // namespace cdx
// twiddle table tw[m] = exp(-2 pi i m/n), m < n, uploaded to *dptr (C ABI status)
int make_twiddles(int n, int prec, void** dptr);
namespace cdx {
// lines per workgroup of a pass: valid (power of two, 64..256 threads, LDS
// within 160 KiB) and the default choice; csz = bytes per complex element
int tiling_valid(int n, size_t csz, int lines);
int tiling_default(int n, size_t csz, int column_pass);

// The passes; tw = twiddle table of the transformed line length (2n
// entries); lines = lines per workgroup (0: default).
template <class T>
hipError_t launch_col(int op, int rows, int cols, const void* in, void* out, const void* tw,
                      const T* I, const ProjArgs<T>& pa, double gauss_sigma, hipStream_t st, int lines = 0);
// The support mask of the single-frame update pass (ROW_INV_UPDATE_FWD):
// word row * lanes + t holds, in bit i, spt[row][x_i] > (T)scal[0] for the
// i-th pixel x_i the row pass gives lane t (its E <= 32 pixels of the row).
// Rebuilt whenever the support or the threshold changes; the test is the
// one the pass made on the float support, so the iterates are unchanged.
template <class T>
hipError_t launch_support_mask(int rows, int cols, const T* spt, const double* scal, unsigned* mask, hipStream_t st);
// bytes of that mask for a rows x cols frame of element size csz (0: unsupported length)
size_t support_mask_bytes(int rows, int cols, size_t csz);
template <class T>
hipError_t launch_row(int op, int rows, int cols, const void* in, void* out, const void* tw,
                      void* g, const T* spt, const UpdArgs<T>& ua, T scale, T* sout,
                      const ReduceSlots& rs, hipStream_t st, int lines = 0);
// Raise the dynamic-LDS cap of every column-pass instantiation for rows-point
// columns / row-pass instantiation for cols-point rows to its largest valid
// tiling, so no attribute call is left for a launch inside a stream capture.
template <class T> hipError_t prepare_col(int rows);
template <class T> hipError_t prepare_row(int cols);
// Lets `kern` launch with `bytes` of dynamic LDS (above the 64 KiB default):
// hipFuncSetAttribute(MaxDynamicSharedMemorySize) when bytes exceed what was
// already raised for that kernel on this device (thread-safe; a lookup
// otherwise)
hipError_t lds_cap(const void* kern, size_t bytes);
// the real-space shrink-wrap blur's column kernel (k_blur_cols<T>) for
// rows-point columns at blur width sigma (its LDS grows with sigma)
template <class T> hipError_t prepare_blur(int rows, T sigma);

// PRTF (cdi.cc:355-380,423-440).  prtf_phase sums the complex partials of
// ROW_INV_UPDATE_FWD_PRTF: first step phi0 = -arg(sum)/2, theta = phi0;
// later steps theta = phi0 + arg(sum).  prtf_add: map += (F/|F|) e^{i theta}
// where |F| > 1e-6 (addPhase, cudaConfig.cu:876-887).  radial: out[r], r <
// rows/2 = sum over 2 cols angles of the bilinear sample of the fftshifted
// image at radius r (cart2polar_kernel + edgeReduce, cudaConfig.cu:1027-1063);
// kind 0 samples the map |src| > 1e-6, kind 1 samples |src| * norm.
hipError_t launch_prtf_phase(double* phi0, double* theta, const ReduceSlots& rs, int first, hipStream_t st);
template <class T> hipError_t launch_prtf_add(int64_t n, void* map, const void* F, const double* theta, hipStream_t st);
template <class T> hipError_t launch_radial(const void* src, int rows, int cols, int kind, T norm, double* out,
                                            hipStream_t st);

// elementwise / small kernels
// packed measurement: NaN where bs > 0.5 (unmeasured), else I, or the
// amplitude sqrt(max(I, 0)) when amp != 0 (what the modulus projection uses)
template <class T> hipError_t launch_pack_intensity(int64_t n, const T* I, const T* bs, T* out, int amp, hipStream_t st);
// packed measurement in the column-tile layout (single-frame loop)
template <class T>
hipError_t launch_pack_intensity_tiled(int rows, int cols, const T* I, const T* bs, T* out, int amp, hipStream_t st);
template <class T> hipError_t launch_mod_abs(int64_t n, void* F, const T* I, hipStream_t st);
template <class T> hipError_t launch_mod_noise(int64_t n, void* F, const T* I, const T* bs, int noise, T scale,
                                               T rcolor, T bsnorm, hipStream_t st);
template <class T> hipError_t launch_support(int64_t n, void* g, const void* gp, const T* spt,
                                             const UpdArgs<T>& ua, hipStream_t st);
template <class T> hipError_t launch_norm_c(int64_t n, void* d, T f, hipStream_t st);
template <class T> hipError_t launch_norm_r(int64_t n, T* d, T f, hipStream_t st);
template <class T> hipError_t launch_mod2(int64_t n, T* out, const void* in, T norm, hipStream_t st);
template <class T> hipError_t launch_gauss_mult_freq(int rows, int cols, const void* in, void* out, T sigma,
                                                     hipStream_t st);
template <class T> hipError_t launch_convert_fo(int rows, int cols, const void* in, void* out, int cpx,
                                                hipStream_t st);
template <class T> hipError_t launch_rect_mask(int rows, int cols, int sx, int sy, int ex, int ey, T* out,
                                               hipStream_t st);
template <class T> hipError_t launch_circle_mask(int rows, int cols, float radius, int freq, T* out,
                                                 hipStream_t st);
template <class T> hipError_t launch_reduce(int64_t n, const T* x, const ReduceSlots& rs, hipStream_t st);
template <class T> hipError_t launch_threshold(double* scal, const ReduceSlots& rs, double shrinkThreshold,
                                               int rows, int cols, int mode, hipStream_t st);
template <class T> hipError_t launch_blur_real(int rows, int cols, const void* g, int cpx_in, T* tmp, T* spt,
                                               T sigma, const ReduceSlots& rs, hipStream_t st);
template <class T> hipError_t launch_start(int64_t n, const T* I, const T* bs, double thr, uint64_t seed,
                                           int mode, void* F, hipStream_t st);
hipError_t launch_fill_f64(int64_t n, double* p, double v, hipStream_t st);
hipError_t launch_copy(size_t bytes, const void* src, void* dst, hipStream_t st);
template <class T> hipError_t launch_residual(int64_t n, const void* F, const T* I, const T* bs, double* acc,
                                              hipStream_t st);
template <class T> hipError_t launch_rfactor(int64_t n, const void* F, const T* Ipk, double* acc,
                                             hipStream_t st);
template <class T> hipError_t launch_cast_c(int64_t n, const void* in, int in_f64, void* out, hipStream_t st);

// setup kernels (prepareIter / readObjectWave / checkAutoCorrelation)
template <class T> hipError_t launch_fo_phase(int rows, int cols, void* d, T norm, hipStream_t st);
template <class T> hipError_t launch_mask_mul(int64_t n, T* d, const T* m, int invert, hipStream_t st);
template <class T> hipError_t launch_poisson_wo(int64_t n, T* w, double noise, T scale, T rcolor, uint64_t seed,
                                                hipStream_t st);
template <class T> hipError_t launch_extend(int64_t n, const T* a, void* b, hipStream_t st);
template <class T> hipError_t launch_wavefront(int rows, int cols, const T* I, const T* P, int orow, int ocol,
                                               void* out, hipStream_t st);

// synthetic forward model
// support shape of a synthetic object: kind 0 disk (radius a), 1 ellipse
// (semi-axes a, b, rotation rot), 2 polygon (nv <= 8 vertices at increasing
// angles ang[] in [0, 2 pi) and radii rad[], star-shaped about the centre)
struct ShapeParams {
  int kind, nv;
  float a, b, rot;
  float ang[8], rad[8];
};
// B <= kSynthMax frames side by side (element (r, c) of frame b at
// r * B * cols + b * cols + c; B = 1: one frame): per frame the seed, support
// shape and photon count.  launch_synth_fields makes the objects of n x n
// frames (field buffers a, ph, ta, tph: B * n * n doubles; mm: 4 * B
// doubles, the min/max keys); launch_synth_intensity the intensities of
// their transforms F (tot: B doubles), Poisson draws or (poisson = 0) the
// means.  Each frame of a stack equals the frame made alone from the same
// seed, shape and photons.
constexpr int kSynthMax = 16;
struct SynthFrames {
  int B;
  uint64_t seed[kSynthMax];
  double photons[kSynthMax];
  ShapeParams sp[kSynthMax];
};
// whether the line smoothing serves n x n frames (otherwise one frame at a time)
bool synth_stack_supported(int n, double sigma);
template <class T>
hipError_t launch_synth_fields(int n, const SynthFrames& f, double sigma, double* a, double* ph, double* ta,
                               double* tph, double* mm, void* obj, hipStream_t st);
template <class T>
hipError_t launch_synth_intensity(int rows, int cols, const SynthFrames& f, const void* F, double* tot, int poisson,
                                  T* I, hipStream_t st);

}  // namespace cdx

// the forward model for a stack of frames on a context's stream (abi.hip)
struct cdx_ctx;
int sim_stack(cdx_ctx* c, const cdx::SynthFrames& f, double sigma, double* a, double* ph, double* ta, double* tph,
              double* mm, double* tot, void* obj, void* F, float* I);
