/*
 * cdx_abi.h -- C ABI of the MI355X CDI phase-retrieval path (libcdixx_amd.so).
 *
 * Every entry point returns CDX_OK (0) or a CDX_ERR_* status; the message of
 * the last failure on the calling thread is cdx_last_error().  Nothing here
 * aborts or exits (the reference aborts in gpuAssert, src/gpu/cudaConfig.cu:27-34,
 * and exits on bad data, src/post_gpu/cdi.cc:97-101).
 *
 * The hidden globals of the reference (image size cuda_imgsz, the cuFFT
 * plan, __constant__ cudaVars{rcolor, beta_HIO, scale, threshold}:
 * src/gpu/cudaConfig.cu:14-18, src/gpu/cdi.cu:8-30) live in an explicit
 * context.  One context per host thread; calls on one context are ordered
 * on its stream.
 *
 * Device buffers are caller-owned raw pointers.  In a CDX_F32 context real
 * arrays are float and complex arrays interleaved float pairs (the layout of
 * the reference's complexFormat = _Complex float, include/format.hpp:131-132);
 * in a CDX_F64 context they are double and double pairs.  Arrays are
 * row-major, index = x*cols + y; Fourier-space arrays are in FFT order.
 */
#ifndef CDX_ABI_H
#define CDX_ABI_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
  CDX_OK = 0,
  CDX_ERR_ARG = 1,         /* bad argument (null pointer, unsupported size, ...) */
  CDX_ERR_HIP = 2,         /* HIP runtime error */
  CDX_ERR_SCHEDULE = 3,    /* malformed algorithm string / unknown token */
  CDX_ERR_UNSUPPORTED = 4, /* token outside this path (TV) */
  CDX_ERR_DATA = 5,        /* NaN or all-zero pattern */
  CDX_ERR_IO = 6,          /* file read/write failure */
  CDX_ERR_CONFIG = 7       /* unreadable configuration file */
};
enum { CDX_F32 = 0, CDX_F64 = 1 };
/* algorithm tokens: the values of enum Algorithm, include/format.hpp:22 */
enum { CDX_RAAR = 0, CDX_ER, CDX_POSER, CDX_HIO, CDX_POSHIO, CDX_FHIO, CDX_SHRINKWRAP,
       CDX_XCORRELATION, CDX_KKT, CDX_TV };
enum { CDX_FORWARD = -1, CDX_INVERSE = 1 };

typedef struct cdx_ctx cdx_ctx;

const char* cdx_last_error(void);
const char* cdx_version(void);
/* 1 if a line length is supported by the FFT (16..2048, 2^k and 3*2^k) */
int cdx_size_supported(int n);

/* ---- device memory: replaces memMngr borrow/return and myMemcpy* ----
 * include/cudaConfig.hpp:16-21,32-38 */
int cdx_malloc(void** ptr, size_t bytes);
int cdx_free(void* ptr);
int cdx_memcpy_h2d(void* dst, const void* src, size_t bytes);
int cdx_memcpy_d2h(void* dst, const void* src, size_t bytes);
int cdx_memcpy_d2d(void* dst, const void* src, size_t bytes);
int cdx_memset0(void* ptr, size_t bytes);
int cdx_device_synchronize(void);
/* myMemcpyD2D as a kernel on `stream` (NULL: legacy stream); blocking */
int cdx_copy(void* stream, void* dst, const void* src, size_t bytes);

/* ---- context: replaces resize_cuda_image + init_fft + init_cuda_image ----
 * include/cudaConfig.hpp:22-23,45; src/gpu/cudaConfig.cu:69-74,137-160;
 * src/gpu/cdi.cu:16-30.  stream may be NULL (the context creates one at
 * default priority). */
int cdx_create(cdx_ctx** out, int rows, int cols, int precision, void* stream);
int cdx_destroy(cdx_ctx* ctx);
int cdx_synchronize(cdx_ctx* ctx);
/* size and precision of a context */
int cdx_context_info(cdx_ctx* ctx, int* rows, int* cols, int* precision);
void* cdx_stream(cdx_ctx* ctx);
/* init_cuda_image(rcolor, scale), include/cudaConfig.hpp:23 */
int cdx_init_vars(cdx_ctx* ctx, int rcolor, double scale);
/* setThreshold, include/cudaConfig.hpp:28 (src/gpu/cdi.cu:242-244) */
int cdx_set_threshold(cdx_ctx* ctx, double threshold);
int cdx_get_threshold(cdx_ctx* ctx, double* threshold);
/* vars.beta_HIO (fixed to 1 in the reference, src/gpu/cdi.cu:22) and the
 * 1.9 relaxation literal of ER/FHIO/POSHIO (src/gpu/cdi.cu:534-616) */
int cdx_set_beta(cdx_ctx* ctx, double beta);
int cdx_set_relax(cdx_ctx* ctx, double relax);

/* ---- free functions on the path ---- */
/* myFFT / myIFFT, include/cudaConfig.hpp:46-47: unnormalised 2D C2C, in==out allowed */
int cdx_fft(cdx_ctx* ctx, const void* in, void* out, int direction);
/* applyModAbs(source, target), include/cudaConfig.hpp:111 */
int cdx_apply_mod_abs(cdx_ctx* ctx, void* F, const void* I);
/* applyMod(source, target, bs, noiseLevel, norm), include/cudaConfig.hpp:107 */
int cdx_apply_mod(cdx_ctx* ctx, void* F, const void* I, const void* bs, int noiseLevel, double bsnorm);
/* applySupport(gkp1, gkprime, algo, spt), include/cdi.hpp:43 */
int cdx_apply_support(cdx_ctx* ctx, void* g, const void* gprime, int algo, const void* spt);
/* applyNorm(complexFormat*, Real) and applyNorm(Real*, Real), include/cudaConfig.hpp:71,73 */
int cdx_apply_norm(cdx_ctx* ctx, void* data, double factor);
int cdx_apply_norm_real(cdx_ctx* ctx, void* data, double factor);
/* getMod2(Real*, complexFormat*, norm), include/cudaConfig.hpp:93 */
int cdx_get_mod2(cdx_ctx* ctx, void* out, const void* in, double norm);
/* applyGaussMult(complex in, out, sigma, isFreq=1), include/cudaConfig.hpp:85 */
int cdx_gauss_mult_freq(cdx_ctx* ctx, const void* in, void* out, double sigma);
/* applyGaussConv(input, output, gaussMem, sigma), include/misc.hpp; real input */
int cdx_gauss_conv(cdx_ctx* ctx, const void* in, void* out, double sigma);
/* cudaConvertFO(data, out), include/cudaConfig.hpp:185; out may be NULL (in place) */
int cdx_convert_fo(cdx_ctx* ctx, void* data, void* out, int is_complex);
/* CDI::createSupport rectangle, src/post_gpu/cdi.cc:240-251 */
int cdx_create_support(cdx_ctx* ctx, void* spt, double oversampling_spt);
/* createBeamStop(row, column, beamStopSize), src/post_gpu/misc.cc:191-195 */
int cdx_create_beamstop(cdx_ctx* ctx, void* bs, double radius);
/* centred disk support of the given radius (createMask<C_circle>, isFreq=0) */
int cdx_create_circle(cdx_ctx* ctx, void* mask, double radius);
/* findMax / findSum, include/cub_wrap.hpp:3,6; blocking, result on the host */
int cdx_find_max(cdx_ctx* ctx, const void* x, int64_t n, double* out);
int cdx_find_sum(cdx_ctx* ctx, const void* x, int64_t n, double* out);
/* createWaveFront(I) + applyRandomPhase, src/post_gpu/cdi.cc:222-223:
 * mode 0 = random sign (the reference's start), 1 = uniform random phase.
 * Pixels with bs > threshold start at 0. */
int cdx_random_start(cdx_ctx* ctx, const void* I, const void* bs, uint64_t seed, int mode, void* F);
/* residual of phaseRetrieve (src/post_gpu/cdi.cc:413-421); blocking */
int cdx_residual(cdx_ctx* ctx, const void* F, const void* I, const void* bs, double* out);
/* R = sum | |F| - sqrt(I) | / sum sqrt(I) over pixels outside the beamstop; blocking */
int cdx_rfactor(cdx_ctx* ctx, const void* F, const void* I, const void* bs, double* out);

/* ---- setup operations used by CDI::prepareIter / readObjectWave ---- */
/* convertFOPhase(data, norm), include/cudaConfig.hpp:125: times (-1)^(x+y)*norm */
int cdx_convert_fo_phase(cdx_ctx* ctx, void* data, double norm);
/* multiply(data, data, mask) or with 1-mask (linearConst + multiply,
 * src/post_gpu/cdi.cc:185-187); include/cudaConfig.hpp:118 */
int cdx_mask_multiply(cdx_ctx* ctx, void* data, const void* mask, int invert);
/* applyPoissonNoise_WO(wave, noiseLevel, state), include/cudaConfig.hpp:100:
 * quantise and add dark Poisson noise (Philox stream 2 with this seed) */
int cdx_poisson_noise_wo(cdx_ctx* ctx, void* data, double noiseLevel, uint64_t seed);
/* extendToComplex(a, b), include/cudaConfig.hpp:70 */
int cdx_extend_to_complex(cdx_ctx* ctx, const void* a, void* b);
/* createWaveFront(I, phase, objectWave, row, col), include/cudaConfig.hpp:82
 * (I and phase may be NULL) */
int cdx_create_wavefront(cdx_ctx* ctx, const void* I, const void* phase, int obj_rows, int obj_cols, void* out);

/* ---- the .bin image format: struct imageFile, include/imageFile.hpp:3-17;
 * readImage_c .bin branch, src/independant/imgio.c:23-43; writeComplexImage /
 * writeFloatImage, imgio.c:223-243.  Host memory; type 3 = FLOAT, 4 = FLOAT2,
 * 5 = DOUBLE, 6 = DOUBLE2. ---- */
int cdx_bin_read_header(const char* path, int* type, int* rows, int* cols);
int cdx_bin_read(const char* path, void* host, size_t capacity);
int cdx_bin_write(const char* path, int type, int rows, int cols, const void* host);

/* ---- the loop: CDI::phaseRetrieve, src/post_gpu/cdi.cc:297-421 ---- */
typedef struct cdx_params {
  int simCCDbit;        /* 1: applyMod (noise-tolerant, beamstop-aware); 0: applyModAbs */
  int useBS;            /* beamstop in use */
  int freeBeamstop;     /* applyModAbs path leaves beamstop pixels unconstrained (new; default 0) */
  int noiseLevel;
  int rcolor;           /* 65535 in the reference (format.hpp:127) */
  double exposure;      /* vars.scale = 1/exposure */
  double beta;          /* vars.beta_HIO; reference 1 */
  double relax;         /* ER/FHIO/POSHIO relaxation; reference 1.9 */
  double shrinkThreshold;
  int shrinkMode;       /* 0: frequency-domain (cdi.cc:326-341); 1: real-space (holo.cc:138-146) */
  double sigma0;        /* 3 */
  double sigmaMin;      /* 2 (freq) / 1.5 (real) */
  double sigmaDecay;    /* 0.99 */
  int recordR;          /* record the R-factor before every projection */
  int useGraph;         /* capture the schedule in a hipGraph and launch it once */
  int prtfIter;         /* PRTF accumulation from this iteration counter on (cdi.cc:355-380); -1 off */
} cdx_params;
void cdx_params_default(cdx_params* p);

/* F: patternWave (in: start wave; out: FFT of the final object).
 * g: objectWave (in: as prepareIter leaves it, IFFT(F)/sqrt(N); out: final
 *    object times sqrt(N), as cdi.cc:313,406 leave it).
 * I: patternData (FFT order); bs: beamstop or NULL; spt: support (in/out,
 *    rewritten by shrink-wrap).  The context threshold is used and updated.
 * rtrace (host, optional): R-factor before each projection, up to rtrace_cap.
 * n_iter (optional): value of the reference's iteration counter at exit. */
int cdx_phase_retrieve(cdx_ctx* ctx, const char* algorithm, const cdx_params* p, void* F, void* g,
                       const void* I, const void* bs, void* spt, double* rtrace, int rtrace_cap, int* n_iter);

/* No counterpart in the reference: cdx_phase_retrieve for callers that
 * cannot allocate device memory (code that must not call hipMalloc: tight
 * serving loops, harnesses that forbid it).  FP32 frames of rows x cols <=
 * 512 x 512 (FFT lengths); F, g, I, bs, spt as in cdx_phase_retrieve; the
 * support threshold comes in and goes out through *threshold (host); stream
 * may be NULL.  Its workspace is one static device array of 6 MiB + 120 KiB
 * per GPU, held until the process ends: no device memory is allocated or
 * freed (the schedule string is parsed on the host).  Launches go straight
 * to the stream (no graph); PRTF accumulation is not available (prtfIter
 * must be < 0) and recordR is ignored.  Calls that overlap on one GPU share
 * that workspace and race.  Larger frames: CDX_ERR_ARG. */
int cdx_phase_retrieve_static(int rows, int cols, const char* algorithm, const cdx_params* p, void* F, void* g,
                              const void* I, const void* bs, void* spt, double* threshold, void* stream,
                              int* n_iter);

/* PRTF of the last cdx_phase_retrieve with p->prtfIter >= 0 (cdi.cc:423-440):
 * prtf[r] and weights[r], r < rows/2, are the radial sums whose ratio the
 * reference writes to prtf.txt; *count = iter - prtfIter (0: no step
 * accumulated, both arrays zero).  Host arrays, up to cap values each. */
int cdx_prtf(cdx_ctx* ctx, double* prtf, double* weights, int cap, int* count);

/* ---- batched reconstructions (no counterpart in the reference, which runs
 * one frame per process through CDI::phaseRetrieve): frames of sizes 128,
 * 256, 384 or 512 (square) advance together, one launch per pass for all
 * frames.  The batch owns packed buffers; cdx_batch_frame returns the device
 * slices of frame i (same layouts as the single-frame entry points), and
 * cdx_batch_context a single-frame context of that frame's size on the
 * batch's stream for setup operations on the slices. ---- */
typedef struct cdx_batch cdx_batch;
int cdx_batch_create(cdx_batch** out, int nframes, const int* sizes, int precision, void* stream);
int cdx_batch_destroy(cdx_batch* b);
int cdx_batch_frame(cdx_batch* b, int i, void** F, void** g, void** I, void** bs, void** spt, int* n);
cdx_ctx* cdx_batch_context(cdx_batch* b, int i);
int cdx_batch_set_threshold(cdx_batch* b, double threshold);
int cdx_batch_get_thresholds(cdx_batch* b, double* thresholds);
int cdx_batch_synchronize(cdx_batch* b);
/* CDI::phaseRetrieve on every frame with one schedule (frequency-form
 * shrink-wrap, per-frame thresholds).  rtrace (optional, host):
 * nframes x rtrace_cap R-factors, frame-major. */
int cdx_batch_phase_retrieve(cdx_batch* b, const char* algorithm, const cdx_params* p, double* rtrace,
                             int rtrace_cap, int* n_iter);
/* times n iterations of `algo` for the whole batch (graph or stream launches):
 * one untimed iteration, with use_graph one untimed launch of the captured n
 * iterations, then n timed ones; F and g enter and leave as
 * cdx_batch_phase_retrieve takes and leaves them */
int cdx_batch_time_iterations(cdx_batch* b, const cdx_params* p, int algo, int n, int use_graph, double* ms);
/* threads per workgroup of both batched passes (64, 128 or 256; frame size n
 * runs threads / lanes(n) lines per workgroup); CDX_ERR_ARG, tiling
 * unchanged, when a frame size present cannot take it.  Default 256 (FP32),
 * 128 (FP64). */
int cdx_batch_set_tiling(cdx_batch* b, int threads);
int cdx_batch_get_tiling(cdx_batch* b, int* threads);
/* the whole schedule: threads per workgroup and, per frame size 128, 256,
 * 384, 512 (variants[0..3]), the line variant: 0 = the default radix
 * schedule, 1 (FP32 only) = twice the lanes with half the points each
 * (128: 4x4x4x2, 256: 8x8x4, 384: 4x4x4x6, 512: 8x8x8).  get_schedule also
 * writes the radices of the sizes present ("256:16x16 384:8x8x6 ...") into
 * radices[cap] (optional). */
int cdx_batch_set_schedule(cdx_batch* b, int threads, const int* variants);
int cdx_batch_get_schedule(cdx_batch* b, int* threads, int* variants, char* radices, int cap);
/* autotuner: times `reps` HIO iterations (one graph) for every schedule (64,
 * 128 and 256 threads per workgroup, each with every combination of line
 * variants over the sizes present) and keeps the fastest; us[k] (k < nus,
 * optional) = microseconds per iteration of the best schedule at 64 << k
 * threads, 0 where none fits.  F and g are restored, so the frames' state is
 * unchanged. */
int cdx_batch_autotune(cdx_batch* b, int reps, double* us, int nus);

/* ---- the schedule grammar: AlgoParser, src/independant/readConfig.cc:82-163 ---- */
/* expands the schedule into tokens; returns the token count or -CDX_ERR_SCHEDULE */
int cdx_expand_schedule(const char* algorithm, int* out, int cap);

/* ---- configuration: readConfig::readConfig + print, include/readConfig.hpp:36-49 ----
 * Parses a libconfig file and writes every readConfig field as one JSON
 * object into buf (host only, no GPU needed). */
int cdx_read_config_json(const char* path, char* buf, int cap);

/* ---- measurement helpers (timing with HIP events on the context stream) ---- */
/* Runs iterations of `algo` from the state F, g (g as cdx_phase_retrieve
 * takes it, IFFT(F)/sqrt(N)) without shrink-wrap: one untimed iteration, with
 * use_graph one untimed launch of the captured n_iter iterations (repeated
 * until at least the cdx_set_timing_warmup count has run), without it that
 * many untimed iterations, then n_iter timed iterations.  ms_total: wall time
 * of the n_iter timed iterations between two events.  F and g are left as
 * cdx_phase_retrieve leaves them.  If col_ms/row_ms are given, each pass is
 * then launched 200 times on its own, back to back (one graph when
 * use_graph), between two events, three times, and the median mean time per
 * launch of the column pass and of the row pass is returned; those launches
 * re-apply the update to the same input, so F and g are then no iterate of
 * the loop.  With p->recordR the column pass also forms the R-factor sums
 * (the cost of recording R every iteration). */
int cdx_time_iterations(cdx_ctx* ctx, const cdx_params* p, int algo, void* F, void* g, const void* I,
                        const void* bs, const void* spt, int n_iter, int use_graph, double* ms_total,
                        double* col_ms, double* row_ms);

/* Untimed iterations cdx_time_iterations runs right before its timed ones
 * (at least; default 0: one untimed launch of the captured iterations).  No
 * counterpart in the reference; for callers timing short runs: after a
 * fraction of a millisecond of work an MI355X runs the loop ~3% slower than
 * after tens of milliseconds. */
int cdx_set_timing_warmup(cdx_ctx* ctx, int iters);

/* Per-pass launch durations of the loop (ms per launch) from the state F, g
 * after one untimed iteration: each pass launched n times on its own.
 * mode 0: back to back in one captured graph between two events, the median
 * of three runs (the figure cdx_time_iterations returns); mode 1: every
 * launch between its own pair of events on the context stream, mean over
 * the launches -- launches serialized the way a kernel-trace profiler
 * serializes them.  F and g are afterwards no iterate of the loop (the row
 * pass re-applies the update to the same input). */
int cdx_time_passes(cdx_ctx* ctx, const cdx_params* p, int algo, void* F, void* g, const void* I, const void* bs,
                    const void* spt, int n, int mode, double* col_ms, double* row_ms);

/* ---- tiling of the FFT passes (no counterpart in the reference, whose
 * cuFFT plan chooses internally; init_fft, src/gpu/cudaConfig.cu:137-160) ----
 * lines = columns (column pass) / rows (row pass) per workgroup; 0 = default. */
int cdx_set_tiling(cdx_ctx* ctx, int col_lines, int row_lines);
int cdx_get_tiling(cdx_ctx* ctx, int* col_lines, int* row_lines);
/* Times every valid tiling of both passes (reps launches each, HIP events)
 * and keeps the fastest.  report (optional) receives triples
 * (pass 0=column/1=row, lines, microseconds per launch), up to cap doubles. */
int cdx_autotune(cdx_ctx* ctx, int reps, double* report, int cap);

/* ---- synthetic forward model (Philox4x64-10 streams) ----
 * Asynchronous on the context's stream, like the other context calls. */
/* n x n object: smoothed-noise amplitude in [0.2,1] and phase in [-pi,pi]
 * inside a centred disk of the given radius.  Context must be n x n. */
int cdx_sim_object(cdx_ctx* ctx, double radius, uint64_t seed, double smooth_sigma, void* obj);
/* the same object inside another support shape centred at (n/2, n/2):
 * kind 0 disk {radius}; 1 ellipse {a, b, rotation}, |rotated (dx/a, dy/b)| < 1;
 * 2 polygon {ang_0..ang_{k-1}, rad_0..rad_{k-1}}, 3 <= k <= 8 vertices at
 * increasing angles in [0, 2 pi), star-shaped about the centre (np = 2k) */
int cdx_sim_object_shape(cdx_ctx* ctx, int kind, const double* params, int np, uint64_t seed, double smooth_sigma,
                         void* obj);
/* I = |FFT(obj)|^2 scaled to `photons` in total, Poisson sampled if `poisson` */
int cdx_sim_intensity(cdx_ctx* ctx, const void* obj, double photons, uint64_t seed, int poisson, void* I);

/* ---- pattern preprocessing: util/processimg.cpp:11-109 (host arrays) ----
 * sig (rows x cols) minus bkg (optional; the larger image is first cropped
 * centrally to the smaller, comparing row counts as the reference does);
 * the centroid of the result thresholded at half its maximum (findMiddle) is
 * used when `centre` is set (then midx += shift_row, midy -= shift_col:
 * argv[7], argv[6] of the reference), else the centre is (0, 0); the image is
 * cropped around it to the largest multiple of 4*nmerge that fits, shifted by
 * the sub-pixel remainder through the FFT (shift_mode 0, shiftWave) or
 * re-centred by the modulus of its spectrum (shift_mode 1, shiftMiddle), and
 * binned nmerge x nmerge.  out receives out_size^2 values (cap), mid the
 * centre (midx, midy) in pixels.  The cropped size must be an FFT length of
 * this build (cdx_size_supported). */
int cdx_preprocess_pattern(const float* sig, int rows, int cols, const float* bkg, int brows, int bcols,
                           int nmerge, int centre, double shift_row, double shift_col, int shift_mode, float* out,
                           int cap, int* out_size, double* mid);

/* ---- learned warm start (no counterpart in the reference, whose loop
 * starts from random signs, src/post_gpu/cdi.cc:222-223): a U-Net on the
 * centred log-intensity log(1 + I)/16 predicting the object amplitude, its
 * phase (relative to the global phase arg(sum obj)) and a support logit.
 * BF16 MFMA convolutions with FP32 accumulation; Adam on FP32 master weights.
 * Geometry: size x size input (size a supported FFT length divisible by
 * 2^(levels-1)), `levels` resolution levels of base*2^l channels, two 3x3
 * convolutions + ReLU per level, 2x2 average pooling, nearest upsampling and
 * skip concatenation, 1x1 head.  One network per host thread. ---- */
typedef struct cdx_unet cdx_unet;
int cdx_unet_create(cdx_unet** out, int size, int levels, int base, int max_batch, uint64_t seed);
int cdx_unet_destroy(cdx_unet* net);
int cdx_unet_num_params(cdx_unet* net, int64_t* n);
/* per conv layer (encoder a,b per level; decoder a,b per level below the
 * bottom; head): cin, cout, taps, Kp, weight offset, bias offset into the
 * flat parameter array (weights [cout][Kp], k = tap*cin + ci); returns the
 * layer count, fills up to cap values */
int cdx_unet_layers(cdx_unet* net, int64_t* info, int cap);
int cdx_unet_set_params(cdx_unet* net, const float* host, int64_t n);
int cdx_unet_get_params(cdx_unet* net, float* host, int64_t n);
int cdx_unet_get_grads(cdx_unet* net, float* host, int64_t n);
int cdx_unet_zero_grads(cdx_unet* net);
/* host batch in: x [B][size][size] network input, targets [B][size][size][3]
 * (amplitude, phase, mask) or NULL; and back out */
int cdx_unet_set_batch(cdx_unet* net, int batch, const float* x, const float* targets);
int cdx_unet_get_batch(cdx_unet* net, int batch, float* x, float* targets);
int cdx_unet_forward(cdx_unet* net, int batch);
/* head output of the last forward: [B][size][size][3] amplitude, phase, logit */
int cdx_unet_output(cdx_unet* net, int batch, float* host);
/* loss of the last forward (mean over pixels of (a-a*)^2 + 0.5 m*(1-cos(p-p*))
 * + BCE(logit, m*)) and its gradients added to the parameter gradients */
int cdx_unet_backward(cdx_unet* net, int batch, double* loss);
int cdx_unet_adam(cdx_unet* net, double lr);
/* forward-model training pairs on the GPU: sample b uses seed seed0 + b
 * (ellipse or polygon support of radius 20-60 px, photons log-uniform in
 * [lo, hi], Poisson noise) */
int cdx_unet_synth_batch(cdx_unet* net, int batch, uint64_t seed0, double photons_lo, double photons_hi);
/* one frame of the same generator into caller buffers (complex obj, real I,
 * FFT order; FP32, size x size) */
int cdx_unet_synth_frame(cdx_unet* net, uint64_t seed, double photons_lo, double photons_hi, void* obj, void* I);
/* steps x (synthesise batch, forward, backward, Adam); losses[steps] optional */
int cdx_unet_train(cdx_unet* net, int steps, int batch, uint64_t seed0, double lr, double photons_lo,
                   double photons_hi, double* losses);
/* warm start on ctx (size x size, FP32) from the measured I (FFT order):
 * spt = sigmoid(logit) > support_threshold (0/1, centred), g0 = amplitude
 * e^{i phase} on the support scaled to the measured total intensity,
 * F = FFT(g0), g = IFFT(F)/sqrt(N) (the state CDI::prepareIter leaves) */
int cdx_unet_warm_start(cdx_unet* net, cdx_ctx* ctx, const void* I, double support_threshold, void* F, void* g,
                        void* spt);
/* warm start of a square FP32 frame of n = f * size pixels per side (ctx; f
 * a positive integer) with ctx_small (size x size, FP32) as the network's
 * work context: the network sees the pattern's central size x size
 * frequencies (FFT order), i.e. the object at f-times coarser pixels over the
 * same field of view.  F = cdx_random_start(I, bs, seed, mode 1), then at
 * those frequencies F takes the phase of the prediction's FFT (modulus
 * kept); spt = the predicted support, dilated by a disk of `dilate` (0..16)
 * small-frame pixels, upsampled by f about the centre (nearest); g =
 * IFFT(F)/sqrt(N).  Configuration: a 2048^2 frame of a 320 px object becomes
 * a 256^2 pattern of a 40 px object, inside the trained range. */
int cdx_unet_warm_start_scaled(cdx_unet* net, cdx_ctx* ctx_small, cdx_ctx* ctx, const void* I, const void* bs,
                               double support_threshold, int dilate, uint64_t seed, void* F, void* g, void* spt);
int cdx_unet_save(cdx_unet* net, const char* path);
int cdx_unet_load(cdx_unet* net, const char* path);

/* ---- ptychography: the scan-position loop of util/ptycho.cpp:514-561 ----
 * The reference keeps the object, probe, patterns and scan positions in
 * `class ptycho` (util/ptycho.cpp:40-69) and runs the loop over the free
 * functions of include/ptycho.hpp (getWindow, updateWindow, updateObject,
 * updateObjectAndProbeStep), myFFT/myIFFT and applyModAccurate / applyMod
 * (include/cudaConfig.hpp:107-110).  A cdx_ptycho holds the positions, the
 * packed patterns and the work buffers; object, probe and probe step are the
 * caller's device arrays (complex64, row-major).  FP32 only, as the
 * reference (Real = float). */
typedef struct cdx_ptycho cdx_ptycho;
enum { CDX_PT_ACCURATE = 0, CDX_PT_NOISE = 1 };
typedef struct cdx_ptycho_params {
  int mode;              /* CDX_PT_ACCURATE: applyModAccurate (ptycho.cpp:540,545, iter < 2000);
                            CDX_PT_NOISE: applyMod with noiseLevel (:542,547) */
  int noiseLevel;
  double scale;          /* cudaVars.scale = 1/exposure (init_cuda_image, ptycho.cpp:249) */
  double rcolor;         /* cudaVars.rcolor (65535) */
  double probeMax;       /* max |P|^2 (ptycho.cpp:479-480); < 0: computed on the device */
  double objMax;         /* max |O|^2 used by the probe step (ptycho.cpp:481-488) */
  double objStep;        /* object step: 1 (updateObject, :555), 0.5 (updateObjectAndProbeStep, :557) */
  double probeStepSize;  /* probe-step weight (1/nscan, ptycho.cpp:453) */
  int updateProbe;       /* 0: updateObject; 1: updateObjectAndProbeStep (iter >= update_probe_iter) */
  int maxConcurrent;     /* positions per launch (0: up to 64); 1 runs the positions one by one */
} cdx_ptycho_params;
int cdx_ptycho_params_default(cdx_ptycho_params* p);
/* probe n x n (64, 128, 256, 384, 512); object obj_rows x obj_cols; nscan
 * window offsets sx (rows), sy (columns): round(scanpos + shift) of
 * ptycho.cpp:517-523 (the sub-pixel shift there truncates to zero) */
int cdx_ptycho_create(cdx_ptycho** out, int n, int obj_rows, int obj_cols, int nscan, const int* sx, const int* sy,
                      void* stream);
int cdx_ptycho_destroy(cdx_ptycho* pt);
int cdx_ptycho_set_positions(cdx_ptycho* pt, const int* sx, const int* sy);
/* patterns: device f32 [nscan][n][n], FFT order, divided by the exposure
 * (ptycho.cpp:734-736); bs: device f32 [n][n] (> 0.5 = beamstop) or NULL.
 * Both are copied (packed) before the call returns. */
int cdx_ptycho_set_patterns(cdx_ptycho* pt, const void* patterns, const void* bs);
/* one pass over the positions order[0..norder) (order NULL: 0..nscan-1):
 * the object (and probe_step when updateProbe) as the reference's loop leaves
 * them.  n_groups: launch groups used (optional). */
int cdx_ptycho_sweep(cdx_ptycho* pt, const int* order, int norder, const cdx_ptycho_params* p, void* object,
                     const void* probe, void* probe_step, int* n_groups);
/* the launch schedule of a sweep (host only, no device needed): groups of
 * mutually independent positions in launch order.  A position's level is one
 * more than the highest level of the earlier positions in `order` whose
 * windows (clipped to the object) overlap its own; levels are split into
 * groups of at most max_concurrent (0: 64).  groups[n_groups] holds the group
 * sizes (at most norder entries), list[norder] the scan indices in launch
 * order.  Has no counterpart in the reference, whose loop runs one position
 * at a time (ptycho.cpp:514). */
int cdx_ptycho_levels(int n, int obj_rows, int obj_cols, int nscan, const int* sx, const int* sy, const int* order,
                      int norder, int max_concurrent, int* groups, int* list, int* n_groups);
/* probe += probe_step; probe_step = 0 (ptycho.cpp:563-566) */
int cdx_ptycho_apply_probe_step(cdx_ptycho* pt, void* probe, void* probe_step);
/* forward model without noise: patterns[i] = |FFT(P * window_i(O))|^2 / (n n)
 * (createPattern, ptycho.cpp:328-347) */
int cdx_ptycho_simulate(cdx_ptycho* pt, const void* object, const void* probe, void* patterns);
int cdx_ptycho_synchronize(cdx_ptycho* pt);
/* n_sweeps sweeps with the same order after one untimed sweep; ms_total
 * between two events on the context's stream */
int cdx_ptycho_time_sweeps(cdx_ptycho* pt, const int* order, int norder, const cdx_ptycho_params* p, void* object,
                           const void* probe, void* probe_step, int n_sweeps, double* ms_total, int* n_groups);

#ifdef __cplusplus
}
#endif
#endif
