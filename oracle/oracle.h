/*
 * oracle.h -- CPU restatement of CDIxx's CDI phase-retrieval path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP path in
 * cdixx_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * The same source is compiled twice: with REAL=double (liboracle_f64.so)
 * and REAL=float (liboracle_f32.so).  Exported names are identical in both.
 * Complex data is interleaved (re, im) pairs of REAL, i.e. the layout of
 * the reference's complexFormat (_Complex float, format.hpp:131-132).
 *
 * Every array is row-major, index = x*cols + y (cudaDefs_h.cu:61-65).
 * Fourier-space arrays are in FFT order (DC at [0,0]).
 */
#ifndef CDX_ORACLE_H
#define CDX_ORACLE_H
#include <stdint.h>
#ifndef REAL
#define REAL double
#endif
typedef struct { REAL re, im; } ocpx;

/* Algorithm tokens: same values as enum Algorithm, format.hpp:22. */
enum { O_RAAR = 0, O_ER, O_POSER, O_HIO, O_POSHIO, O_FHIO, O_SHRINKWRAP,
       O_XCORRELATION, O_KKT, O_TV, O_CNT };

/* ---- FFT (a5) ---- */
void orc_fft1d(int n, ocpx* x, int sign);                 /* in place, unnormalised */
void orc_fft2d(int rows, int cols, const ocpx* in, ocpx* out, int sign);
void orc_dft2d_brute(int rows, int cols, const ocpx* in, ocpx* out, int sign);

/* ---- elementwise ops on the path ---- */
void orc_apply_mod_abs(int64_t n, ocpx* F, const REAL* I);
void orc_apply_mod(int64_t n, ocpx* F, const REAL* I, const REAL* bs,
                   int noiseLevel, REAL scale, int rcolor, REAL bsnorm);
void orc_apply_support(int64_t n, ocpx* g, const ocpx* gp, int algo, const REAL* spt,
                       REAL threshold, REAL beta, REAL relax);
void orc_apply_norm(int64_t n, ocpx* x, REAL f);
void orc_get_mod2(int64_t n, REAL* out, const ocpx* x, REAL norm);
void orc_gauss_mult_freq(int rows, int cols, const ocpx* in, ocpx* out, REAL sigma);
void orc_gauss_conv_real(int rows, int cols, const REAL* in, REAL* out, REAL sigma);
void orc_convert_fo_real(int rows, int cols, REAL* data);
void orc_convert_fo_cpx(int rows, int cols, ocpx* data);
void orc_create_support_rect(int rows, int cols, double oversampling_spt, REAL* out);
void orc_create_beamstop(int rows, int cols, REAL radius, REAL* out);
void orc_create_circle(int rows, int cols, REAL radius, REAL* out);
void orc_random_sign_start(int rows, int cols, const REAL* I, const REAL* bs,
                           REAL threshold, const uint8_t* flip, ocpx* F);
REAL orc_find_max(int64_t n, const REAL* x);
REAL orc_find_sum(int64_t n, const REAL* x);

/* ---- shrink-wrap (a8, a8') ---- */
typedef struct {
  REAL sigma;       /* current blur sigma */
  REAL avg;         /* freq mode: 4*sum/N from the first call, 0 until then */
  REAL threshold;   /* global support threshold (cudaVars.threshold) */
  REAL decay;       /* sigma *= decay per call while sigma > the mode minimum (0.99) */
} orc_sw_state;
void orc_shrink_wrap_freq(int rows, int cols, const ocpx* F, REAL* spt, REAL shrinkThreshold,
                          orc_sw_state* st, ocpx* scratch);
void orc_shrink_wrap_real(int rows, int cols, const ocpx* g, REAL* spt, REAL shrinkThreshold,
                          REAL sigmaMin, orc_sw_state* st, REAL* scratch);

/* ---- schedule (a2) ---- */
typedef struct orc_parser orc_parser;
orc_parser* orc_parser_new(const char* formula, int* err);
void orc_parser_free(orc_parser* p);
void orc_parser_restart(orc_parser* p);
int  orc_parser_next(orc_parser* p);
int  orc_parser_expand(const char* formula, int* out, int cap);  /* -> count, or -1 */

/* ---- reductions for diagnostics ---- */
void orc_rfactor(int64_t n, const ocpx* F, const REAL* I, const REAL* bs, double* num, double* den);
double orc_residual(int64_t n, const ocpx* F, const REAL* I, const REAL* bs);

/* ---- one iteration and the full loop (a1) ---- */
typedef struct {
  int simCCDbit, useBS, noiseLevel, rcolor, shrinkMode, startSign;
  REAL scale, beta, relax, shrinkThreshold, sigmaStart, sigmaMin, sigmaDecay;
  /* applyModAbs path only: pixels with bs > 0.5 keep F (the beamstop is left
   * unconstrained); a mode of this build, the reference has no such switch */
  int freeBeamstop;
} orc_params;
void orc_iteration(int rows, int cols, ocpx* F, ocpx* g, const REAL* I, const REAL* bs,
                   const REAL* spt, int algo, const orc_params* p, REAL threshold, ocpx* gp);
int  orc_phase_retrieve(int rows, int cols, const char* algorithm, ocpx* F, ocpx* g,
                        const REAL* I, const REAL* bs, REAL* spt, const orc_params* p,
                        REAL* threshold_inout, double* rtrace, int rtrace_cap, int* n_iter_out);

/* ---- PRTF (a11) ---- */
void orc_add_phase(int64_t n, ocpx* a, const ocpx* b, REAL phi);
void orc_radial_sum(int rows, int cols, const REAL* img, REAL* out);
int  orc_phase_retrieve_prtf(int rows, int cols, const char* algorithm, ocpx* F, ocpx* g, const REAL* I,
                             const REAL* bs, REAL* spt, const orc_params* p, REAL* threshold_inout,
                             double* rtrace, int rtrace_cap, int* n_iter_out, int prtfIter, REAL* prtf_out,
                             REAL* weights_out, int* prtf_count);

/* ---- setup steps of prepareIter / readObjectWave (a14) ---- */
void orc_create_wavefront(int rows, int cols, const REAL* I, const REAL* P, int orow, int ocol, ocpx* out);
void orc_convert_fo_phase(int rows, int cols, ocpx* d, REAL norm);
void orc_poisson_noise_wo(int64_t n, REAL* w, double noise, REAL scale, REAL rcolor, uint64_t seed);
void orc_extend_to_complex(int64_t n, const REAL* a, ocpx* b);

/* ---- synthetic forward model (Philox4x64-10 streams) ---- */
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t idx);
void orc_sim_object(int n, REAL radius, uint64_t seed, REAL smooth_sigma, ocpx* obj);
void orc_sim_intensity(int rows, int cols, const ocpx* obj, double photons, uint64_t seed,
                       int poisson, REAL* I);
int64_t orc_poisson(double lam, uint64_t seed, uint64_t stream, uint64_t idx);
#endif
