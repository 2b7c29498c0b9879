/*
 * oracle.c -- CPU restatement of CDIxx's CDI phase-retrieval path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  Compiled twice (REAL=double, REAL=float) by oracle/build.sh.
 *
 * Each function names the reference file:line it restates (paths relative
 * to the CDIxx tree).  How each function is pinned (closed forms, brute
 * force, library cross-checks, known-answer vectors) is listed in the
 * docstrings of tests/test_oracle.py.
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* FFT: own mixed-radix Stockham autosort transform.                          */
/* Replaces the cuFFT plan used by init_fft/myFFT/myIFFT,                     */
/* src/gpu/cudaConfig.cu:137-187: unnormalised C2C, forward kernel            */
/* exp(-2*pi*i*k*n/N) (sign=-1), inverse exp(+...) (sign=+1).                 */
/* ------------------------------------------------------------------------ */
static int factorize(int n, int* f) {
  int k = 0;
  while (n % 4 == 0) { f[k++] = 4; n /= 4; }
  while (n % 2 == 0) { f[k++] = 2; n /= 2; }
  for (int p = 3; n > 1; p += 2)
    while (n % p == 0) { f[k++] = p; n /= p; }
  return k;
}

void orc_fft1d(int n, ocpx* x, int sign) {
  if (n <= 1) return;
  int fac[64];
  int nf = factorize(n, fac);
  ocpx* work = (ocpx*)malloc(sizeof(ocpx) * (size_t)n);
  ocpx* src = x;
  ocpx* dst = work;
  int ns = 1;  /* length of the sub-transforms already done */
  ocpx v[64];
  double wr[64], wi[64];
  for (int s = 0; s < nf; s++) {
    int r = fac[s];
    int m = n / r;
    /* roots of the radix-r DFT */
    for (int q = 0; q < r; q++) {
      double a = sign * 2.0 * M_PI * q / r;
      wr[q] = cos(a); wi[q] = sin(a);
    }
    for (int j = 0; j < m; j++) {
      int k = j % ns;
      double ang = sign * 2.0 * M_PI * (double)k / (double)(ns * r);
      for (int q = 0; q < r; q++) {
        double c = cos(ang * q), sn = sin(ang * q);
        double ar = src[j + q * m].re, ai = src[j + q * m].im;
        v[q].re = (REAL)(ar * c - ai * sn);
        v[q].im = (REAL)(ar * sn + ai * c);
      }
      int base = (j / ns) * ns * r + k;
      for (int u = 0; u < r; u++) {
        double sr = 0, si = 0;
        for (int q = 0; q < r; q++) {
          int e = (u * q) % r;
          sr += (double)v[q].re * wr[e] - (double)v[q].im * wi[e];
          si += (double)v[q].re * wi[e] + (double)v[q].im * wr[e];
        }
        dst[base + u * ns].re = (REAL)sr;
        dst[base + u * ns].im = (REAL)si;
      }
    }
    ns *= r;
    ocpx* t = src; src = dst; dst = t;
  }
  if (src != x) memcpy(x, src, sizeof(ocpx) * (size_t)n);
  free(work);
}

/* 2D transform, rows then columns (separable; cudaConfig.cu:175-180). */
void orc_fft2d(int rows, int cols, const ocpx* in, ocpx* out, int sign) {
  if (out != in) memcpy(out, in, sizeof(ocpx) * (size_t)rows * cols);
#pragma omp parallel for schedule(static)
  for (int x = 0; x < rows; x++) orc_fft1d(cols, out + (size_t)x * cols, sign);
#pragma omp parallel
  {
    ocpx* col = (ocpx*)malloc(sizeof(ocpx) * (size_t)rows);
#pragma omp for schedule(static)
    for (int y = 0; y < cols; y++) {
      for (int x = 0; x < rows; x++) col[x] = out[(size_t)x * cols + y];
      orc_fft1d(rows, col, sign);
      for (int x = 0; x < rows; x++) out[(size_t)x * cols + y] = col[x];
    }
    free(col);
  }
}

/* O(N^2) direct DFT straight from the definition, for pinning the FFT. */
void orc_dft2d_brute(int rows, int cols, const ocpx* in, ocpx* out, int sign) {
  for (int kx = 0; kx < rows; kx++)
    for (int ky = 0; ky < cols; ky++) {
      double sr = 0, si = 0;
      for (int x = 0; x < rows; x++)
        for (int y = 0; y < cols; y++) {
          long ex = ((long)kx * x) % rows, ey = ((long)ky * y) % cols;
          double a = sign * 2.0 * M_PI * ((double)ex / rows + (double)ey / cols);
          double c = cos(a), s = sin(a);
          double re = in[x * cols + y].re, im = in[x * cols + y].im;
          sr += re * c - im * s;
          si += re * s + im * c;
        }
      out[kx * cols + ky].re = (REAL)sr;
      out[kx * cols + ky].im = (REAL)si;
    }
}

/* ------------------------------------------------------------------------ */
/* Modulus projections                                                        */
/* ------------------------------------------------------------------------ */
/* applyModAbs, src/gpu/cudaConfig.cu:827-841, with state==NULL as CDI passes
 * it (src/post_gpu/cdi.cc:350): F <- F*sqrt(max(I,0))/|F|; |F|==0 -> (r,0). */
void orc_apply_mod_abs(int64_t n, ocpx* F, const REAL* I) {
  for (int64_t i = 0; i < n; i++) {
    REAL mod = (REAL)hypot((double)F[i].re, (double)F[i].im);
    REAL rat = I[i];
    rat = rat > 0 ? (REAL)sqrt((double)rat) : (REAL)0;
    if (mod == 0) { F[i].re = rat; F[i].im = 0; continue; }
    rat /= mod;
    F[i].re *= rat;
    F[i].im *= rat;
  }
}

/* applyMod (5-argument form), src/gpu/cdi.cu:438-485, with the globals
 * vars.scale = 1/exposure and vars.rcolor (cdi.cu:8-30, format.hpp:127). */
void orc_apply_mod(int64_t n, ocpx* F, const REAL* I, const REAL* bs, int noiseLevel,
                   REAL scale, int rcolor, REAL bsnorm) {
  for (int64_t i = 0; i < n; i++) {
    REAL rx = F[i].re, ry = F[i].im;
    REAL mod2 = I[i] > 0 ? I[i] : (REAL)0;
    if (bs && bs[i] > (REAL)0.5) {
      if (bsnorm != 1) { F[i].re = rx * bsnorm; F[i].im = ry * bsnorm; }
      continue;
    }
    if (bsnorm != 1) { rx *= bsnorm; ry *= bsnorm; }
    REAL src2 = rx * rx + ry * ry;
    REAL maximum = scale * (REAL)0.95;
    if (mod2 >= maximum) mod2 = mod2 > src2 ? mod2 : src2;
    REAL tol = (REAL)sqrt((double)(noiseLevel + mod2 * rcolor / scale)) * scale / rcolor;
    REAL diff = mod2 - src2;
    REAL val = mod2;
    if (diff > tol) val -= tol;
    else if (diff < -tol) val += tol;
    if (src2 == 0) { F[i].re = val; F[i].im = 0; continue; }
    REAL s = (REAL)sqrt((double)((val > 0 ? val : 0) / src2));
    F[i].re = s * rx;
    F[i].im = s * ry;
  }
}

/* ------------------------------------------------------------------------ */
/* Real-space updates: dispatcher src/post_gpu/cdi.cc:25-36 and kernels       */
/* src/gpu/cdi.cu:518-616.  inside = spt > threshold (cdi.cu:525).  beta is   */
/* vars.beta_HIO, hard-wired to 1 in the reference (cdi.cu:22); relax is the  */
/* reference's literal 1.9.                                                   */
/* ------------------------------------------------------------------------ */
void orc_apply_support(int64_t n, ocpx* g, const ocpx* gp, int algo, const REAL* spt,
                       REAL threshold, REAL beta, REAL relax) {
  for (int64_t i = 0; i < n; i++) {
    int in = spt[i] > threshold;
    ocpx a = g[i], b = gp[i];
    switch (algo) {
      case O_RAAR:  /* cdi.cu:550-568 */
        if (in) a = b;
        else { a.re = beta * a.re + (1 - 2 * beta) * b.re; a.im = beta * a.im + (1 - 2 * beta) * b.im; }
        break;
      case O_ER:    /* cdi.cu:585-600 */
        if (in) { a.re += relax * (b.re - a.re); a.im += relax * (b.im - a.im); }
        else { a.re = 0; a.im = 0; }
        break;
      case O_POSER: /* cdi.cu:569-584 */
        if (in && b.re > 0) a = b;
        else { a.re = 0; a.im = 0; }
        break;
      case O_POSHIO: /* cdi.cu:601-616 */
        if (b.re > 0 && in) a.re += relax * (b.re - a.re);
        else a.re -= beta * b.re;
        a.im -= beta * b.im;
        break;
      case O_HIO:   /* cdi.cu:518-533 */
        if (in) a = b;
        else { a.re -= beta * b.re; a.im -= beta * b.im; }
        break;
      default:      /* FHIO and every other token, cdi.cu:534-549 */
        if (in) { a.re += relax * (b.re - a.re); a.im += relax * (b.im - a.im); }
        else { a.re -= (REAL)1.2 * b.re; a.im -= (REAL)1.2 * b.im; }
        break;
    }
    g[i] = a;
  }
}

/* applyNorm(complex*), cudaConfig.cu:425-429 */
void orc_apply_norm(int64_t n, ocpx* x, REAL f) {
  for (int64_t i = 0; i < n; i++) { x[i].re *= f; x[i].im *= f; }
}

/* getMod2(Real*, complex*, norm), cudaConfig.cu:773-777 */
void orc_get_mod2(int64_t n, REAL* out, const ocpx* x, REAL norm) {
  for (int64_t i = 0; i < n; i++) out[i] = (x[i].re * x[i].re + x[i].im * x[i].im) * norm;
}

/* applyGaussMult(complex, isFreq=1), cudaConfig.cu:380-395: wrapped
 * frequency coordinates, factor exp(-(kx^2+ky^2)/(2 sigma^2)). */
void orc_gauss_mult_freq(int rows, int cols, const ocpx* in, ocpx* out, REAL sigma) {
  for (int x = 0; x < rows; x++)
    for (int y = 0; y < cols; y++) {
      REAL xr = x >= rows / 2 ? (REAL)(x - rows) : (REAL)x;
      REAL yr = y >= cols / 2 ? (REAL)(y - cols) : (REAL)y;
      REAL f = (REAL)exp(-(double)(xr * xr + yr * yr) / (2 * (double)sigma * sigma));
      size_t i = (size_t)x * cols + y;
      out[i].re = f * in[i].re;
      out[i].im = f * in[i].im;
    }
}

/* applyGaussConv, src/post_gpu/misc.cc:80-86 -> createGauss
 * (cudaConfig.cu:447-457) + applyConvolution (cudaConfig.cu:475-507):
 * radius floor(6 sigma)>>1, weights exp(-(dx^2+dy^2)/(2s^2))/(2 pi s^2),
 * zero padding outside the image. */
void orc_gauss_conv_real(int rows, int cols, const REAL* in, REAL* out, REAL sigma) {
  int r = ((int)floor((double)sigma * 6)) >> 1;
  int w = 2 * r + 1;
  REAL* k = (REAL*)malloc(sizeof(REAL) * (size_t)w * w);
  for (int i = 0; i < w * w; i++) {
    int dx = i / w - r, dy = i % w - r;
    k[i] = (REAL)(1.0 / (2 * M_PI * (double)sigma * sigma) *
                  exp(-(double)(dx * dx + dy * dy) / (2 * (double)sigma * sigma)));
  }
#pragma omp parallel for schedule(static)
  for (int x = 0; x < rows; x++)
    for (int y = 0; y < cols; y++) {
      double s = 0;
      for (int a = -r; a <= r; a++) {
        int xx = x + a;
        if (xx < 0 || xx >= rows) continue;
        for (int b = -r; b <= r; b++) {
          int yy = y + b;
          if (yy < 0 || yy >= cols) continue;
          s += (double)in[(size_t)xx * cols + yy] * k[(a + r) * w + (b + r)];
        }
      }
      out[(size_t)x * cols + y] = (REAL)s;
    }
  free(k);
}

/* cudaConvertFO, cudaConfig.cu:202-216: swap quadrants (fftshift for even
 * sizes). */
void orc_convert_fo_real(int rows, int cols, REAL* d) {
  int half = rows * cols / 2;
  for (int idx = 0; idx < half; idx++) {
    int x = idx / cols, y = idx % cols;
    int ip = (x + (rows >> 1)) * cols + (y >= (cols >> 1) ? y - (cols >> 1) : y + (cols >> 1));
    REAL t = d[idx]; d[idx] = d[ip]; d[ip] = t;
  }
}
void orc_convert_fo_cpx(int rows, int cols, ocpx* d) {
  int half = rows * cols / 2;
  for (int idx = 0; idx < half; idx++) {
    int x = idx / cols, y = idx % cols;
    int ip = (x + (rows >> 1)) * cols + (y >= (cols >> 1) ? y - (cols >> 1) : y + (cols >> 1));
    ocpx t = d[idx]; d[idx] = d[ip]; d[ip] = t;
  }
}

/* CDI::createSupport, src/post_gpu/cdi.cc:240-251 -> createMask<rect>
 * (cudaConfig.cu:1291-1303) and rect::isInside (cudaConfig.cu:89-92). */
void orc_create_support_rect(int rows, int cols, double os, REAL* out) {
  int sx = (int)((os - 1) / 2 * rows / os - 1);
  int sy = (int)((os - 1) / 2 * cols / os - 1);
  int ex = rows - sx, ey = cols - sy;
  for (int x = 0; x < rows; x++)
    for (int y = 0; y < cols; y++)
      out[(size_t)x * cols + y] = (x >= sx && x < ex && y >= sy && y < ey) ? 1 : 0;
}

/* createBeamStop, src/post_gpu/misc.cc:191-195 -> createCircleMask(isFreq=1)
 * (misc.cc:161-170) -> createMask<C_circle> with the quadrant shift
 * (cudaConfig.cu:1291-1301) and C_circle::isInside (cudaConfig.cu:93-98):
 * a disk of radius r around DC in FFT order, strict '<'. */
void orc_create_beamstop(int rows, int cols, REAL radius, REAL* out) {
  int x0 = rows >> 1, y0 = cols >> 1;
  for (int x = 0; x < rows; x++)
    for (int y = 0; y < cols; y++) {
      int xs = x >= rows / 2 ? x - rows / 2 : x + rows / 2;
      int ys = y >= cols / 2 ? y - cols / 2 : y + cols / 2;
      float dx = (float)(xs - x0), dy = (float)(ys - y0);
      float dr = sqrtf(dx * dx + dy * dy);
      out[(size_t)x * cols + y] = dr < (float)radius ? 1 : 0;
    }
}

/* Centred disk (createMask<C_circle>, isFreq=0); the object support of the
 * synthetic workloads. */
void orc_create_circle(int rows, int cols, REAL radius, REAL* out) {
  int x0 = rows >> 1, y0 = cols >> 1;
  for (int x = 0; x < rows; x++)
    for (int y = 0; y < cols; y++) {
      float dx = (float)(x - x0), dy = (float)(y - y0);
      out[(size_t)x * cols + y] = sqrtf(dx * dx + dy * dy) < (float)radius ? 1 : 0;
    }
}

/* createWaveFront(patternData, 0, patternWave, 1) (cudaConfig.cu:710-732)
 * followed by applyRandomPhase (src/gpu/cdi.cu:486-502): F = sqrt(max(I,0))
 * times +1 or -1 (flip[i]), zero where bs > threshold.  The reference draws
 * the flips from cuRAND; here the caller supplies them. */
void orc_random_sign_start(int rows, int cols, const REAL* I, const REAL* bs, REAL threshold,
                           const uint8_t* flip, ocpx* F) {
  int64_t n = (int64_t)rows * cols;
  for (int64_t i = 0; i < n; i++) {
    if (bs && bs[i] > threshold) { F[i].re = 0; F[i].im = 0; continue; }
    REAL m = I[i] > 0 ? (REAL)sqrt((double)I[i]) : (REAL)0;
    /* sincosf(pi) as the reference evaluates it in float */
    F[i].re = flip[i] ? -m : m;
    F[i].im = flip[i] ? m * (REAL)sin((double)(float)M_PI) : 0;
  }
}

/* findMax / findSum, src/gpu/cub_wrap.cu:88-92,154-158 (max_op with init 0). */
REAL orc_find_max(int64_t n, const REAL* x) {
  REAL m = 0;
  for (int64_t i = 0; i < n; i++) if (x[i] > m) m = x[i];
  return m;
}
REAL orc_find_sum(int64_t n, const REAL* x) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += x[i];
  return (REAL)s;
}

/* ------------------------------------------------------------------------ */
/* Shrink-wrap                                                                */
/* ------------------------------------------------------------------------ */
/* Frequency-domain form, src/post_gpu/cdi.cc:326-341. */
void orc_shrink_wrap_freq(int rows, int cols, const ocpx* F, REAL* spt, REAL shrinkThreshold,
                          orc_sw_state* st, ocpx* scratch) {
  /* sigma schedule: *= decay while > 2 (cdi.cc:338-340) */
  int64_t n = (int64_t)rows * cols;
  orc_gauss_mult_freq(rows, cols, F, scratch, (REAL)(rows >> 2) / st->sigma);
  orc_fft2d(rows, cols, scratch, scratch, +1);
  orc_get_mod2(n, spt, scratch, 1);
  REAL maxs = orc_find_max(n, spt);
  if (st->avg == 0) st->avg = orc_find_sum(n, spt) * 4 / rows / cols;
  if (st->avg * 5 < maxs) maxs = st->avg * 5;
  st->threshold = maxs * shrinkThreshold;
  if (st->sigma > 2) st->sigma *= st->decay;
}

/* Real-space form, src/post_gpu/holo.cc:138-146: S = blur(|g|^2),
 * threshold = max(S) * shrinkThreshold, sigma *= 0.99 while > sigmaMin. */
void orc_shrink_wrap_real(int rows, int cols, const ocpx* g, REAL* spt, REAL shrinkThreshold,
                          REAL sigmaMin, orc_sw_state* st, REAL* scratch) {
  int64_t n = (int64_t)rows * cols;
  orc_get_mod2(n, scratch, g, 1);
  orc_gauss_conv_real(rows, cols, scratch, spt, st->sigma);
  st->threshold = orc_find_max(n, spt) * shrinkThreshold;
  if (st->sigma > sigmaMin) st->sigma *= st->decay;
}

/* ------------------------------------------------------------------------ */
/* AlgoParser, src/independant/readConfig.cc:82-163.  The grammar is         */
/* "count*TOKEN+count*TOKEN+..." where TOKEN may be a parenthesised group.    */
/* Groups are parsed recursively (the reference handles one nesting level;    */
/* both agree there).  next() is the reference's state machine.               */
/* ------------------------------------------------------------------------ */
struct orc_parser {
  int n;
  int* count;
  int* algo;                 /* token, or O_CNT + sub index */
  struct orc_parser** subs;
  int nsub;
  int cur;
  int curCount;
};

static int token_id(const char* s, int len) {
  static const struct { const char* name; int id; } tab[] = {
    {"RAAR", O_RAAR}, {"HIO", O_HIO}, {"ER", O_ER}, {"POSER", O_POSER},
    {"POSHIO", O_POSHIO}, {"FHIO", O_FHIO}, {"KKT", O_KKT},
    {"shrinkWrap", O_SHRINKWRAP}, {"TV", O_TV}, {"XCORRELATION", O_XCORRELATION}};
  for (unsigned i = 0; i < sizeof(tab) / sizeof(tab[0]); i++)
    if ((int)strlen(tab[i].name) == len && !strncmp(tab[i].name, s, len)) return tab[i].id;
  return -1;
}

static orc_parser* parse_range(const char* f, int len, int* err) {
  orc_parser* p = (orc_parser*)calloc(1, sizeof(orc_parser));
  int cap = 8;
  p->count = (int*)malloc(sizeof(int) * cap);
  p->algo = (int*)malloc(sizeof(int) * cap);
  p->subs = (orc_parser**)malloc(sizeof(orc_parser*) * cap);
  int i = 0;
  while (i < len) {
    /* count: atoi semantics up to '*' */
    int j = i;
    while (j < len && f[j] != '*') j++;
    if (j >= len) { *err = 1; return p; }
    char buf[32];
    int cl = j - i < 31 ? j - i : 31;
    memcpy(buf, f + i, cl); buf[cl] = 0;
    int num = atoi(buf);
    i = j + 1;
    int id;
    if (i < len && f[i] == '(') {
      int depth = 0, k = i;
      for (; k < len; k++) {
        if (f[k] == '(') depth++;
        else if (f[k] == ')') { if (--depth == 0) break; }
      }
      if (k >= len) { *err = 1; return p; }
      orc_parser* sub = parse_range(f + i + 1, k - i - 1, err);
      if (*err) { orc_parser_free(sub); return p; }
      p->subs[p->nsub] = sub;
      id = O_CNT + p->nsub++;
      i = k + 1;
      if (i < len && f[i] != '+') { *err = 1; return p; }
    } else {
      int k = i;
      while (k < len && f[k] != '+') k++;
      id = token_id(f + i, k - i);
      if (id < 0) { *err = 2; return p; }
      i = k;
    }
    if (p->n == cap) {
      cap *= 2;
      p->count = (int*)realloc(p->count, sizeof(int) * cap);
      p->algo = (int*)realloc(p->algo, sizeof(int) * cap);
      p->subs = (orc_parser**)realloc(p->subs, sizeof(orc_parser*) * cap);
    }
    p->count[p->n] = num;
    p->algo[p->n] = id;
    p->n++;
    if (i < len && f[i] == '+') i++;
  }
  if (p->n == 0) *err = 1;
  return p;
}

orc_parser* orc_parser_new(const char* formula, int* err) {
  size_t L = strlen(formula);
  char* s = (char*)malloc(L + 1);
  int j = 0;
  for (size_t i = 0; i < L; i++) if (formula[i] != ' ') s[j++] = formula[i];
  s[j] = 0;
  *err = 0;
  orc_parser* p = parse_range(s, j, err);
  free(s);
  if (*err) { orc_parser_free(p); return NULL; }
  orc_parser_restart(p);
  return p;
}

void orc_parser_free(orc_parser* p) {
  if (!p) return;
  for (int i = 0; i < p->nsub; i++) orc_parser_free(p->subs[i]);
  free(p->count); free(p->algo); free(p->subs); free(p);
}

void orc_parser_restart(orc_parser* p) {
  p->cur = 0;
  p->curCount = p->count[0];
  for (int i = 0; i < p->nsub; i++) orc_parser_restart(p->subs[i]);
}

int orc_parser_next(orc_parser* p) {
  if (p->curCount == 0) {
    if (p->cur == p->n - 1) return -1;
    p->curCount = p->count[++p->cur];
  }
  if (p->algo[p->cur] >= O_CNT) {
    orc_parser* sub = p->subs[p->algo[p->cur] - O_CNT];
    int r = orc_parser_next(sub);
    if (r == -1) {
      p->curCount--;
      orc_parser_restart(sub);
      return orc_parser_next(p);
    }
    return r;
  }
  p->curCount--;
  return p->algo[p->cur];
}

int orc_parser_expand(const char* formula, int* out, int cap) {
  int err;
  orc_parser* p = orc_parser_new(formula, &err);
  if (!p) return -1;
  int n = 0, t;
  while ((t = orc_parser_next(p)) >= 0) {
    if (n < cap) out[n] = t;
    n++;
    if (n > 10000000) { n = -1; break; }  /* a zero count after the first element never ends */
  }
  orc_parser_free(p);
  return n;
}

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                                */
/* ------------------------------------------------------------------------ */
/* R = sum_{k not in bs} | |F_k| - sqrt(I_k) |  /  sum_{k not in bs} sqrt(I_k)
 * (the reference has no R-factor; this is the build's definition). */
void orc_rfactor(int64_t n, const ocpx* F, const REAL* I, const REAL* bs, double* num, double* den) {
  double a = 0, b = 0;
  for (int64_t i = 0; i < n; i++) {
    if (bs && bs[i] > 0.5) continue;
    double s = I[i] > 0 ? sqrt((double)I[i]) : 0;
    a += fabs(hypot((double)F[i].re, (double)F[i].im) - s);
    b += s;
  }
  *num = a; *den = b;
}

/* Final residual of CDI::phaseRetrieve, src/post_gpu/cdi.cc:413-421:
 * sum over pixels with I < 0.99 of ((|F|^2 - I)*(1-bs))^2. */
double orc_residual(int64_t n, const ocpx* F, const REAL* I, const REAL* bs) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) {
    double m2 = (double)F[i].re * F[i].re + (double)F[i].im * F[i].im;
    double d = I[i] < 0.99 ? m2 - I[i] : 0;
    if (bs) d *= 1 - bs[i];
    s += d * d;
  }
  return s;
}

/* ------------------------------------------------------------------------ */
/* One iteration of the loop body, src/post_gpu/cdi.cc:349-354,381:           */
/* projection -> IFFT -> 1/N -> support update -> FFT.  gp is scratch.        */
/* ------------------------------------------------------------------------ */
/* The projection of cdi.cc:349-350: applyMod (simCCDbit) or applyModAbs;
 * with freeBeamstop (this build's mode, applyModAbs path) pixels with
 * bs > 0.5 keep F. */
static void orc_project(int64_t n, ocpx* F, const REAL* I, const REAL* bs, const orc_params* p) {
  if (p->simCCDbit) {
    orc_apply_mod(n, F, I, p->useBS ? bs : NULL, p->noiseLevel, p->scale, p->rcolor, 1);
  } else if (p->freeBeamstop && p->useBS && bs) {
    for (int64_t i = 0; i < n; i++)
      if (!(bs[i] > (REAL)0.5)) orc_apply_mod_abs(1, F + i, I + i);
  } else {
    orc_apply_mod_abs(n, F, I);
  }
}

void orc_iteration(int rows, int cols, ocpx* F, ocpx* g, const REAL* I, const REAL* bs,
                   const REAL* spt, int algo, const orc_params* p, REAL threshold, ocpx* gp) {
  int64_t n = (int64_t)rows * cols;
  orc_project(n, F, I, bs, p);
  orc_fft2d(rows, cols, F, gp, +1);
  orc_apply_norm(n, gp, (REAL)(1.0 / (double)n));
  orc_apply_support(n, g, gp, algo, spt, threshold, p->beta, p->relax);
  orc_fft2d(rows, cols, g, F, -1);
}

/* ---- PRTF (cdi.cc:355-380 accumulation, 423-440 final) ---- */

/* addPhase, cudaConfig.cu:876-887: where |b| > 1e-6, a += (b/|b|) e^{i phi} */
void orc_add_phase(int64_t n, ocpx* a, const ocpx* b, REAL phi) {
  const REAL c = (REAL)cos(phi), s = (REAL)sin(phi);
  for (int64_t i = 0; i < n; i++) {
    const REAL m = (REAL)hypot(b[i].re, b[i].im);
    if (m > (REAL)1e-6) {
      const REAL r = b[i].re / m, q = b[i].im / m;
      a[i].re += r * c - q * s;
      a[i].im += r * s + q * c;
    }
  }
}

/* bilinear sample at centred coordinates (X, Y) of a w x h row-major image,
 * 0 where the 2x2 stencil leaves the image (cudaConfig.cu:1027-1047) */
static REAL bilinear_centred(const REAL* img, int w, int h, REAL X, REAL Y) {
  X += (REAL)(w >> 1);
  Y += (REAL)(h >> 1);
  const int x0 = (int)floor(X), y0 = (int)floor(Y);
  if (x0 < 0 || x0 + 1 >= w || y0 < 0 || y0 + 1 >= h) return 0;
  const REAL fx = X - (REAL)x0, fy = Y - (REAL)y0;
  const size_t i0 = (size_t)x0 * h + y0;
  return (1 - fx) * (1 - fy) * img[i0] + (1 - fx) * fy * img[i0 + 1] + fx * (1 - fy) * img[i0 + h] +
         fx * fy * img[i0 + h + 1];
}

/* cart2polar_kernel (cudaConfig.cu:1050-1055) on a (2 cols) x (rows/2) polar
 * grid, then edgeReduce (:1057-1063): out[r] = sum over the 2 cols angles
 * theta_a = a * 2pi / (2 cols) of the sample at (r cos theta_a, r sin theta_a). */
void orc_radial_sum(int rows, int cols, const REAL* img, REAL* out) {
  const int nang = 2 * cols, nrad = rows / 2;
  for (int r = 0; r < nrad; r++) out[r] = 0;
  for (int a = 0; a < nang; a++) {
    const REAL th = (REAL)(a * (2.0 * M_PI / nang));
    const REAL c = (REAL)cos(th), s = (REAL)sin(th);
    for (int r = 0; r < nrad; r++) out[r] += bilinear_centred(img, rows, cols, (REAL)r * c, (REAL)r * s);
  }
}

/* CDI::phaseRetrieve, src/post_gpu/cdi.cc:297-440.  g enters as prepareIter
 * leaves it (IFFT(F)/sqrt(N)) and leaves scaled by sqrt(N) (cdi.cc:313,406).
 * rtrace (optional) receives the R-factor of F before each projection.
 *
 * PRTF (prtfIter >= 0): from the first projection step whose counter is
 * >= prtfIter, the projected F's unit phasors are summed into a map with the
 * global-phase alignment of cdi.cc:358-378: gamma0 = P_S(g') at the first such
 * step, phi0 = -arg(sum gamma0^2)/2; later steps add phi0 + arg(sum P_S(g')
 * conj(gamma0)).  P_S keeps g' inside the support and zeroes it outside (the
 * reference applies its relaxed ER kernel to a scratch buffer it never
 * initialises).  weights = radial sums of the fftshifted map |F| > 1e-6 at
 * the first step; prtf = radial sums of the fftshifted |map| / (iter -
 * prtfIter).  prtf_out/weights_out hold rows/2 values; *prtf_count receives
 * iter - prtfIter, or 0 when no step accumulated.
 * Returns 0, or -1 for a bad schedule / unsupported token. */
int orc_phase_retrieve_prtf(int rows, int cols, const char* algorithm, ocpx* F, ocpx* g, const REAL* I,
                            const REAL* bs, REAL* spt, const orc_params* p, REAL* threshold_inout,
                            double* rtrace, int rtrace_cap, int* n_iter_out, int prtfIter, REAL* prtf_out,
                            REAL* weights_out, int* prtf_count) {
  int64_t n = (int64_t)rows * cols;
  int err;
  orc_parser* ap = orc_parser_new(algorithm, &err);
  if (!ap) return -1;
  ocpx* gp = (ocpx*)malloc(sizeof(ocpx) * n);
  REAL* rs = (REAL*)malloc(sizeof(REAL) * n);
  ocpx *map = NULL, *gamma0 = NULL;
  REAL phi0 = 0;
  int started = 0;
  if (prtfIter >= 0) {
    map = (ocpx*)calloc(n, sizeof(ocpx));
    gamma0 = (ocpx*)malloc(sizeof(ocpx) * n);
  }
  orc_apply_norm(n, g, (REAL)(1.0 / sqrt((double)n)));
  orc_sw_state st = {p->sigmaStart, 0, *threshold_inout, p->sigmaDecay};
  int iter = 0, nr = 0, rc = 0;
  for (;; iter++) {
    int a = orc_parser_next(ap);
    if (a < 0) break;
    if (a == O_SHRINKWRAP) {
      if (p->shrinkMode == 0) orc_shrink_wrap_freq(rows, cols, F, spt, p->shrinkThreshold, &st, gp);
      else {
        /* the object estimate in the units the loop carries it */
        orc_shrink_wrap_real(rows, cols, g, spt, p->shrinkThreshold, p->sigmaMin, &st, rs);
      }
      continue;
    }
    if (a == O_TV) { rc = -1; break; }
    if (rtrace && nr < rtrace_cap) {
      /* beamstop pixels are left out where the projection leaves them
       * unconstrained (applyMod with useBS, or the free-beamstop mode) */
      const int bs_free = p->useBS && (p->simCCDbit || p->freeBeamstop);
      double num, den;
      orc_rfactor(n, F, I, bs_free ? bs : NULL, &num, &den);
      rtrace[nr++] = den > 0 ? num / den : 0;
    }
    if (prtfIter < 0 || iter < prtfIter) {
      orc_iteration(rows, cols, F, g, I, bs, spt, a, p, st.threshold, gp);
      continue;
    }
    /* orc_iteration split around the PRTF accumulation */
    orc_project(n, F, I, bs, p);
    orc_fft2d(rows, cols, F, gp, +1);
    orc_apply_norm(n, gp, (REAL)(1.0 / (double)n));
    orc_apply_support(n, g, gp, a, spt, st.threshold, p->beta, p->relax);
    double sre = 0, sim = 0;
    if (!started) {
      for (int64_t i = 0; i < n; i++) {
        const int in = spt[i] > st.threshold;
        gamma0[i].re = in ? gp[i].re : 0;
        gamma0[i].im = in ? gp[i].im : 0;
        sre += (double)gamma0[i].re * gamma0[i].re - (double)gamma0[i].im * gamma0[i].im;
        sim += 2.0 * gamma0[i].re * gamma0[i].im;
      }
      phi0 = (REAL)(-atan2(sim, sre) / 2);
      orc_add_phase(n, map, F, phi0);
      for (int64_t i = 0; i < n; i++) rs[i] = (REAL)hypot(F[i].re, F[i].im) > (REAL)1e-6 ? 1 : 0;
      orc_convert_fo_real(rows, cols, rs);
      orc_radial_sum(rows, cols, rs, weights_out);
      started = 1;
    } else {
      for (int64_t i = 0; i < n; i++) {
        if (!(spt[i] > st.threshold)) continue;
        sre += (double)gp[i].re * gamma0[i].re + (double)gp[i].im * gamma0[i].im;
        sim += (double)gp[i].im * gamma0[i].re - (double)gp[i].re * gamma0[i].im;
      }
      orc_add_phase(n, map, F, (REAL)(phi0 + atan2(sim, sre)));
    }
    orc_fft2d(rows, cols, g, F, -1);
  }
  orc_apply_norm(n, g, (REAL)sqrt((double)n));
  *threshold_inout = st.threshold;
  if (n_iter_out) *n_iter_out = iter;
  if (prtf_count) *prtf_count = started ? iter - prtfIter : 0;
  if (started) {
    const REAL f = (REAL)(1. / (iter - prtfIter));
    for (int64_t i = 0; i < n; i++) rs[i] = (REAL)hypot(map[i].re, map[i].im) * f;
    orc_convert_fo_real(rows, cols, rs);
    orc_radial_sum(rows, cols, rs, prtf_out);
  } else if (prtfIter >= 0) {
    for (int r = 0; r < rows / 2; r++) prtf_out[r] = weights_out[r] = 0;
  }
  orc_parser_free(ap);
  free(gp); free(rs); free(map); free(gamma0);
  return rc;
}

int orc_phase_retrieve(int rows, int cols, const char* algorithm, ocpx* F, ocpx* g,
                       const REAL* I, const REAL* bs, REAL* spt, const orc_params* p,
                       REAL* threshold_inout, double* rtrace, int rtrace_cap, int* n_iter_out) {
  return orc_phase_retrieve_prtf(rows, cols, algorithm, F, g, I, bs, spt, p, threshold_inout, rtrace, rtrace_cap,
                                 n_iter_out, -1, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* Synthetic forward model.                                                   */
/* Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), the published           */
/* counter-based generator.  Streams: 0 = amplitude field, 1 = phase field,   */
/* 2 = Poisson draws, 3 = start signs/phases.                                 */
/* ------------------------------------------------------------------------ */
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
  __uint128_t p = (__uint128_t)a * b;
  *hi = (uint64_t)(p >> 64);
  *lo = (uint64_t)p;
}

void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
    mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B97F4A7C15ULL;
    k1 += 0xBB67AE8584CAA73BULL;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* U[0,1) with 53 random bits: word 0 of block (idx,0,0,0) under key (seed,stream). */
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t c[4] = {idx, 0, 0, 0}, k[2] = {seed, stream}, o[4];
  orc_philox4x64(c, k, o);
  return (double)(o[0] >> 11) * 0x1.0p-53;
}

/* periodic separable Gaussian smoothing of an n x n double field */
static void smooth_periodic(int n, double* f, double sigma) {
  int r = (int)ceil(4 * sigma);
  double* w = (double*)malloc(sizeof(double) * (2 * r + 1));
  double ws = 0;
  for (int i = -r; i <= r; i++) { w[i + r] = exp(-(double)i * i / (2 * sigma * sigma)); ws += w[i + r]; }
  for (int i = 0; i <= 2 * r; i++) w[i] /= ws;
  double* t = (double*)malloc(sizeof(double) * (size_t)n * n);
  for (int x = 0; x < n; x++)
    for (int y = 0; y < n; y++) {
      double s = 0;
      for (int i = -r; i <= r; i++) s += w[i + r] * f[(size_t)x * n + ((y + i) % n + n) % n];
      t[(size_t)x * n + y] = s;
    }
  for (int x = 0; x < n; x++)
    for (int y = 0; y < n; y++) {
      double s = 0;
      for (int i = -r; i <= r; i++) s += w[i + r] * t[(size_t)(((x + i) % n + n) % n) * n + y];
      f[(size_t)x * n + y] = s;
    }
  free(t); free(w);
}

/* Object: amplitude = smoothed U(0,1) rescaled to [0.2,1], phase = second
 * smoothed field rescaled to [-pi,pi], inside a centred disk of the given
 * radius (strict '<', the C_circle rule of cudaConfig.cu:93-98). */
void orc_sim_object(int n, REAL radius, uint64_t seed, REAL smooth_sigma, ocpx* obj) {
  size_t N = (size_t)n * n;
  double* a = (double*)malloc(sizeof(double) * N);
  double* ph = (double*)malloc(sizeof(double) * N);
  for (size_t i = 0; i < N; i++) { a[i] = orc_uniform(seed, 0, i); ph[i] = orc_uniform(seed, 1, i); }
  smooth_periodic(n, a, smooth_sigma);
  smooth_periodic(n, ph, smooth_sigma);
  double amin = a[0], amax = a[0], pmin = ph[0], pmax = ph[0];
  for (size_t i = 1; i < N; i++) {
    if (a[i] < amin) amin = a[i];
    if (a[i] > amax) amax = a[i];
    if (ph[i] < pmin) pmin = ph[i];
    if (ph[i] > pmax) pmax = ph[i];
  }
  int c = n >> 1;
  for (int x = 0; x < n; x++)
    for (int y = 0; y < n; y++) {
      size_t i = (size_t)x * n + y;
      float dx = (float)(x - c), dy = (float)(y - c);
      if (!(sqrtf(dx * dx + dy * dy) < (float)radius)) { obj[i].re = 0; obj[i].im = 0; continue; }
      double amp = 0.2 + 0.8 * (a[i] - amin) / (amax - amin);
      double phs = -M_PI + 2 * M_PI * (ph[i] - pmin) / (pmax - pmin);
      obj[i].re = (REAL)(amp * cos(phs));
      obj[i].im = (REAL)(amp * sin(phs));
    }
  free(a); free(ph);
}

/* per-pixel uniform stream for the Poisson sampler */
typedef struct { uint64_t seed, stream, idx, blk; uint64_t buf[4]; int pos; } ustream;
static double us_next(ustream* u) {
  if (u->pos == 4) {
    uint64_t c[4] = {u->idx, u->blk++, 0, 0}, k[2] = {u->seed, u->stream};
    orc_philox4x64(c, k, u->buf);
    u->pos = 0;
  }
  return (double)(u->buf[u->pos++] >> 11) * 0x1.0p-53;
}

/* Poisson(lam): multiplication (inversion) method for lam < 10, PTRS
 * transformed rejection (Hoermann 1993) above. */
int64_t orc_poisson(double lam, uint64_t seed, uint64_t stream, uint64_t idx) {
  ustream u = {seed, stream, idx, 0, {0, 0, 0, 0}, 4};
  if (!(lam > 0) || !(lam < 1e300)) return 0;  /* also NaN / inf */
  if (lam < 10) {
    double enlam = exp(-lam), prod = 1.0;
    int64_t X = 0;
    for (;;) {
      prod *= us_next(&u);
      if (prod > enlam) X++;
      else return X;
    }
  }
  double slam = sqrt(lam), loglam = log(lam);
  double b = 0.931 + 2.53 * slam;
  double a = -0.059 + 0.02483 * b;
  double invalpha = 1.1239 + 1.1328 / (b - 3.4);
  double vr = 0.9277 - 3.6224 / (b - 2);
  for (;;) {
    double U = us_next(&u) - 0.5;
    double V = us_next(&u);
    double us = 0.5 - fabs(U);
    int64_t k = (int64_t)floor((2 * a / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(invalpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma((double)k + 1))
      return k;
  }
}

/* ---- setup steps of CDI::prepareIter / readObjectWave (cdi.cc:62-82,161-200) ---- */

/* createWaveFront, integer-size form (src/gpu/cudaConfig.cu:688-709) with
 * shift 0 and phaseFactor 0, as readObjectWave calls it (cdi.cc:78): an
 * orow x ocol image at margins ((rows-orow)/2, (cols-ocol)/2), amplitude
 * sqrt(max(I,0)) (1 without I), phase (P-0.5)*2*pi with a phase image, zero
 * elsewhere. */
void orc_create_wavefront(int rows, int cols, const REAL* I, const REAL* P, int orow, int ocol, ocpx* out) {
  const int mx = (rows - orow) / 2, my = (cols - ocol) / 2;
  for (int x = 0; x < rows; x++)
    for (int y = 0; y < cols; y++) {
      ocpx* o = &out[(size_t)x * cols + y];
      if (x < mx || x >= mx + orow || y < my || y >= my + ocol) {
        o->re = o->im = 0;
        continue;
      }
      const size_t t = (size_t)(x - mx) * ocol + (y - my);
      const REAL mod = I ? (REAL)sqrt(I[t] > 0 ? (double)I[t] : 0.0) : (REAL)1;
      const REAL ph = P ? (REAL)(((double)P[t] - 0.5) * 2 * M_PI) : (REAL)0;
      if (ph != 0) {
        o->re = mod * (REAL)cos((double)ph);
        o->im = mod * (REAL)sin((double)ph);
      } else {
        o->re = mod;
        o->im = 0;
      }
    }
}

/* convertFOPhase (cudaConfig.cu:861-870): data *= norm, negated where x+y is odd */
void orc_convert_fo_phase(int rows, int cols, ocpx* d, REAL norm) {
  for (int x = 0; x < rows; x++)
    for (int y = 0; y < cols; y++) {
      const REAL f = ((x + y) & 1) ? -norm : norm;
      d[(size_t)x * cols + y].re *= f;
      d[(size_t)x * cols + y].im *= f;
    }
}

/* applyPoissonNoise_WO (src/gpu/cdi.cu:291-296) with scale = vars.scale:
 * w = scale * (int(w*rcolor/scale) + Poisson(noiseLevel) - noiseLevel) / rcolor.
 * The reference draws from cuRAND; here from the build's Philox Poisson
 * stream 2 at counter i. */
void orc_poisson_noise_wo(int64_t n, REAL* w, double noise, REAL scale, REAL rcolor, uint64_t seed) {
  for (int64_t i = 0; i < n; i++) {
    const double dark = (double)orc_poisson(noise, seed, 2, (uint64_t)i);
    const REAL q = (REAL)((double)(int)(w[i] * rcolor / scale) + dark - noise);
    w[i] = scale * q / rcolor;
  }
}

/* extendToComplex (cudaConfig.cu:682-686): b = (a, 0) */
void orc_extend_to_complex(int64_t n, const REAL* a, ocpx* b) {
  for (int64_t i = 0; i < n; i++) {
    b[i].re = a[i];
    b[i].im = 0;
  }
}

/* I = |FFT(obj)|^2 (FFT order) scaled to `photons` in total, then Poisson
 * sampled (stream 2) when `poisson` is set. */
void orc_sim_intensity(int rows, int cols, const ocpx* obj, double photons, uint64_t seed,
                       int poisson, REAL* I) {
  size_t N = (size_t)rows * cols;
  ocpx* F = (ocpx*)malloc(sizeof(ocpx) * N);
  orc_fft2d(rows, cols, obj, F, -1);
  double tot = 0;
  for (size_t i = 0; i < N; i++) tot += (double)F[i].re * F[i].re + (double)F[i].im * F[i].im;
  double sc = tot > 0 ? photons / tot : 0;
  for (size_t i = 0; i < N; i++) {
    double lam = ((double)F[i].re * F[i].re + (double)F[i].im * F[i].im) * sc;
    I[i] = poisson ? (REAL)orc_poisson(lam, seed, 2, i) : (REAL)lam;
  }
  free(F);
}
