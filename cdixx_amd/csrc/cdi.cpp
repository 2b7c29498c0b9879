// cdi.cpp -- the CDI class (include/cdi.hpp) over the C ABI.
// Reference: src/post_gpu/cdi.cc (class CDI) and util/cdi.cpp (driver).
#include "../../include/cdi.hpp"
#include "../../include/cdx_abi.h"
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <chrono>
#include <string>
#include <vector>

namespace {
CDI* g_last = nullptr;  // context of applySupport(), as the reference's globals

void chk(int rc, const char* what) {
  if (rc != CDX_OK) throw cdx::Error(rc, std::string(what) + ": " + cdx_last_error());
}

// readImage (src/independant/imgio.cc:6) for the .bin format; pixel codecs
// (PNG/TIFF/JPEG) are outside this path
std::vector<float> read_real_image(const char* path, int& rows, int& cols) {
  if (!path) throw cdx::Error(CDX_ERR_IO, "no image file given");
  const char* ext = strrchr(path, '.');
  if (!ext || strcmp(ext, ".bin") != 0)
    throw cdx::Error(CDX_ERR_IO, std::string(path) + ": only the .bin image format is supported");
  int type;
  chk(cdx_bin_read_header(path, &type, &rows, &cols), "readImage");
  size_t n = (size_t)rows * cols;
  if (type == 3) {
    std::vector<float> v(n);
    chk(cdx_bin_read(path, v.data(), n * 4), "readImage");
    return v;
  }
  if (type == 5) {
    std::vector<double> d(n);
    chk(cdx_bin_read(path, d.data(), n * 8), "readImage");
    return std::vector<float>(d.begin(), d.end());
  }
  throw cdx::Error(CDX_ERR_IO, std::string(path) + ": expected a real (FLOAT) image");
}

// centre crop (crop with midx = midy = 0, src/gpu/cudaConfig.cu:348-358)
std::vector<float> centre_crop(const std::vector<float>& src, int row, int col, int r, int c) {
  std::vector<float> out((size_t)r * c);
  for (int x = 0; x < r; x++)
    for (int y = 0; y < c; y++)
      out[(size_t)x * c + y] = src[(size_t)(x + (row - r) / 2) * col + y + (col - c) / 2];
  return out;
}
}  // namespace

// CDI::CDI, src/post_gpu/cdi.cc:37-40
CDI::CDI(const char* configfile) : readConfig(configfile) {
  if (verbose >= 4) print();
  if (runSim) d = oversampling_spt * pixelsize * beamspotsize / lambda;
  g_last = this;
}

CDI::~CDI() {
  if (ctx) {
    cdx_free(objectWave);
    cdx_free(patternWave);
    cdx_free(autoCorrelation);
    cdx_free(patternData);
    cdx_free(support);
    cdx_free(beamstop);
    cdx_destroy(ctx);
  }
  if (g_last == this) g_last = nullptr;
}

// CDI::allocateMem, src/post_gpu/cdi.cc:42-61
void CDI::allocateMem() {
  if (objectWave) return;
  if (exposure == 0) throw cdx::Error(CDX_ERR_CONFIG, "exposure must be set (vars.scale = 1/exposure)");
  chk(cdx_create(&ctx, row, column, CDX_F32, nullptr), "allocateMem");
  size_t sz = (size_t)row * column;
  void* p;
  chk(cdx_malloc(&p, sz * 8), "allocateMem");
  objectWave = (complexFormat*)p;
  chk(cdx_malloc(&p, sz * 8), "allocateMem");
  patternWave = (complexFormat*)p;
  chk(cdx_malloc(&p, sz * 8), "allocateMem");
  autoCorrelation = (complexFormat*)p;
  chk(cdx_malloc(&p, sz * 4), "allocateMem");
  patternData = (Real*)p;
  chk(cdx_init_vars(ctx, 65535, 1. / exposure), "init_cuda_image");
  propagate.row = row;
  propagate.column = column;
  propagate.pixelsize = pixelsize;
  propagate.lambda = lambda;
  propagate.distance = d;
}

// CDI::readObjectWave, src/post_gpu/cdi.cc:62-82 (.bin intensity/phase images)
void CDI::readObjectWave() {
  if (domnist) throw cdx::Error(CDX_ERR_CONFIG, "domnist: the MNIST loader is outside this path");
  int orow = 0, ocol = 0, prow = 0, pcol = 0;
  std::vector<float> inten, phase;
  if (intensityModulation) inten = read_real_image(common.Intensity, orow, ocol);
  if (phaseModulation != 0) {
    phase = read_real_image(common.Phase, prow, pcol);
    if (!intensityModulation) {
      orow = prow;
      ocol = pcol;
    } else if (prow != orow || pcol != ocol) {
      phase = centre_crop(phase, prow, pcol, std::min(prow, orow), std::min(pcol, ocol));
      if (prow < orow || pcol < ocol) throw cdx::Error(CDX_ERR_IO, "phase image smaller than the intensity image");
    }
    for (auto& v : phase) v *= (float)phaseModulation;  // applyNorm(d_phase, phaseModulation)
  }
  if (orow == 0) throw cdx::Error(CDX_ERR_CONFIG, "runSim needs an Intensity or Phase image");
  row = (int)(orow * oversampling);
  column = (int)(ocol * oversampling);
  allocateMem();
  void *dI = nullptr, *dP = nullptr;
  size_t n = (size_t)orow * ocol * 4;
  if (!inten.empty()) {
    chk(cdx_malloc(&dI, n), "readObjectWave");
    chk(cdx_memcpy_h2d(dI, inten.data(), n), "readObjectWave");
  }
  if (!phase.empty()) {
    chk(cdx_malloc(&dP, n), "readObjectWave");
    chk(cdx_memcpy_h2d(dP, phase.data(), n), "readObjectWave");
  }
  int rc = cdx_create_wavefront(ctx, dI, dP, orow, ocol, objectWave);
  cdx_synchronize(ctx);
  cdx_free(dI);
  cdx_free(dP);
  chk(rc, "createWaveFront");
}

// CDI::readPattern, src/post_gpu/cdi.cc:83-108
void CDI::readPattern() {
  int r, c;
  std::vector<float> pat = read_real_image(common.Pattern, r, c);
  if (cropPattern) {
    pat = centre_crop(pat, r, c, cropPattern, cropPattern);
    r = c = cropPattern;
  }
  row = r;
  column = c;
  allocateMem();
  chk(cdx_memcpy_h2d(patternData, pat.data(), pat.size() * 4), "readPattern");
  if (!cropPattern) {
    double mx;
    chk(cdx_find_max(ctx, patternData, 0, &mx), "findMax");
    if (mx < 1e-7 || mx != mx) throw cdx::Error(CDX_ERR_DATA, "pattern max = " + std::to_string(mx));
  }
  chk(cdx_convert_fo(ctx, patternData, nullptr, 0), "cudaConvertFO");
  chk(cdx_apply_norm_real(ctx, patternData, 1. / exposure), "applyNorm");
}

// CDI::readFiles, src/post_gpu/cdi.cc:109-117
void CDI::readFiles() {
  if (runSim) readObjectWave();
  else readPattern();
}

// CDI::setPattern_c, src/post_gpu/cdi.cc:118-122
void CDI::setPattern_c(complexFormat* pattern) {
  chk(cdx_convert_fo(ctx, pattern, patternWave, 1), "cudaConvertFO");
  chk(cdx_get_mod2(ctx, patternData, patternWave, 1), "getMod2");
  chk(cdx_random_start(ctx, patternData, useBS ? beamstop : nullptr, seed_ + draws_++, startMode, patternWave),
      "applyRandomPhase");
}

// CDI::setPattern, src/post_gpu/cdi.cc:124-128
void CDI::setPattern(Real* pattern) {
  chk(cdx_convert_fo(ctx, pattern, patternData, 0), "cudaConvertFO");
  chk(cdx_random_start(ctx, patternData, useBS ? beamstop : nullptr, seed_ + draws_++, startMode, patternWave),
      "createWaveFront + applyRandomPhase");
}

// CDI::init, src/post_gpu/cdi.cc:129-159
void CDI::init() {
  allocateMem();
  if (useBS) {
    void* p;
    chk(cdx_malloc(&p, (size_t)row * column * 4), "init");
    beamstop = (Real*)p;
    if (BSimage && BSimage[0] != '\0') {
      int r, c;
      std::vector<float> bs = read_real_image(BSimage, r, c);
      if (r != row || c != column)
        throw cdx::Error(CDX_ERR_ARG, "Beam stop mask (" + std::to_string(r) + ", " + std::to_string(c) +
                                          ") is not of the same size as diffraction pattern (" + std::to_string(row) +
                                          ", " + std::to_string(column) + ")!");
      chk(cdx_memcpy_h2d(beamstop, bs.data(), bs.size() * 4), "init");
      chk(cdx_convert_fo(ctx, beamstop, nullptr, 0), "cudaConvertFO");
      // invert(beamstop): 1 - bs
      std::vector<float> ones(bs.size(), 1.0f);
      void* t;
      chk(cdx_malloc(&t, bs.size() * 4), "init");
      chk(cdx_memcpy_h2d(t, ones.data(), bs.size() * 4), "init");
      chk(cdx_mask_multiply(ctx, t, beamstop, 1), "invert");
      chk(cdx_memcpy_d2d(beamstop, t, bs.size() * 4), "init");
      cdx_free(t);
    } else {
      chk(cdx_create_beamstop(ctx, beamstop, beamStopSize), "createBeamStop");
    }
  }
  if (dopupil) {
    double k = row * pixelsize * pixelsize / (lambda * d);
    dpupil = d * k / (k + 1);
    resolution = (Real)(lambda * dpupil / (row * pixelsize));
  }
  createSupport();
  seed_ = seed >= 0 ? (unsigned long long)seed
                    : (unsigned long long)std::chrono::system_clock::now().time_since_epoch().count();
}

// CDI::prepareIter, src/post_gpu/cdi.cc:160-230
void CDI::prepareIter() {
  const double N = (double)row * column;
  if (runSim) {
    if (domnist) throw cdx::Error(CDX_ERR_CONFIG, "domnist: the MNIST loader is outside this path");
    if (isFresnel) throw cdx::Error(CDX_ERR_CONFIG, "isFresnel: the Fresnel propagator is outside this path");
    chk(cdx_fft(ctx, objectWave, patternWave, CDX_FORWARD), "myFFT");
    chk(cdx_convert_fo_phase(ctx, patternWave, 1. / sqrt(N)), "convertFOPhase");
    chk(cdx_get_mod2(ctx, patternData, patternWave, 1), "getMod2");
    if (useBS) chk(cdx_mask_multiply(ctx, patternData, beamstop, 1), "multiply");
    chk(cdx_convert_fo(ctx, patternData, nullptr, 0), "cudaConvertFO");
    if (simCCDbit) chk(cdx_poisson_noise_wo(ctx, patternData, noiseLevel, seed_ + draws_++), "applyPoissonNoise_WO");
    chk(cdx_convert_fo(ctx, patternData, nullptr, 0), "cudaConvertFO");
  }
  bool restarted = false;
  if (restart && common.restart) {
    int type, r, c;
    if (cdx_bin_read_header(common.restart, &type, &r, &c) == CDX_OK && type == 4 && r == row && c == column) {
      std::vector<float> wf((size_t)r * c * 2);
      chk(cdx_bin_read(common.restart, wf.data(), wf.size() * 4), "restart");
      chk(cdx_memcpy_h2d(patternWave, wf.data(), wf.size() * 4), "restart");
      restarted = true;
    } else {
      fprintf(stderr, "Restart file %s unusable or size mismatch; random start\n", common.restart);
      restart = 0;
    }
  }
  if (!restarted)
    chk(cdx_random_start(ctx, patternData, useBS ? beamstop : nullptr, seed_ + draws_++, startMode, patternWave),
        "applyRandomPhase");
  chk(cdx_fft(ctx, patternWave, objectWave, CDX_INVERSE), "myIFFT");
  chk(cdx_apply_norm(ctx, objectWave, 1. / sqrt(N)), "applyNorm");
  initSupport();
}

// CDI::checkAutoCorrelation, src/post_gpu/cdi.cc:231-239 (R2C + redundant fill
// == C2C of the real pattern)
void CDI::checkAutoCorrelation() {
  chk(cdx_extend_to_complex(ctx, patternData, autoCorrelation), "extendToComplex");
  chk(cdx_fft(ctx, autoCorrelation, autoCorrelation, CDX_FORWARD), "myFFT");
  chk(cdx_apply_norm(ctx, autoCorrelation, 1. / sqrt((double)row * column)), "applyNorm");
}

// CDI::createSupport / initSupport, src/post_gpu/cdi.cc:240-254
void CDI::createSupport() {
  if (!support) {
    void* p;
    chk(cdx_malloc(&p, (size_t)row * column * 4), "createSupport");
    support = (Real*)p;
  }
  initSupport();
}
void CDI::initSupport() { chk(cdx_create_support(ctx, support, oversampling_spt), "createMask"); }

// CDI::saveState, src/post_gpu/cdi.cc:255-259 (the restart file; PNG plots are
// outside this path)
void CDI::saveState() {
  if (!common.restart) return;
  std::vector<float> h((size_t)row * column * 2);
  chk(cdx_memcpy_d2h(h.data(), patternWave, h.size() * 4), "saveState");
  chk(cdx_bin_write(common.restart, 4, row, column, h.data()), "writeComplexImage");
}

// CDI::phaseRetrieve, src/post_gpu/cdi.cc:297-421
complexFormat* CDI::phaseRetrieve() {
  cdx_params p;
  cdx_params_default(&p);
  p.simCCDbit = simCCDbit;
  p.useBS = useBS;
  p.freeBeamstop = freeBeamstop;
  p.noiseLevel = noiseLevel;
  p.exposure = exposure;
  p.beta = betaHIO;
  p.relax = relaxER;
  p.shrinkThreshold = shrinkThreshold;
  p.shrinkMode = shrinkWrapMode;
  p.sigma0 = shrinkSigma0;
  p.sigmaMin = shrinkSigmaMin >= 0 ? shrinkSigmaMin : (shrinkWrapMode == 0 ? 2.0 : 1.5);
  p.sigmaDecay = shrinkSigmaDecay;
  p.useGraph = useGraph;
  p.prtfIter = prtfIter;
  int n_iter = 0;
  chk(cdx_phase_retrieve(ctx, algorithm, &p, patternWave, objectWave, patternData, useBS ? beamstop : nullptr,
                         support, nullptr, 0, &n_iter),
      "phaseRetrieve");
  double res = 0;
  chk(cdx_residual(ctx, patternWave, patternData, useBS ? beamstop : nullptr, &res), "residual");
  residual = (Real)res;
  if (verbose >= 1) printf("residual= %f\n", residual);
  // prtf.txt: "i, prtf/weight" for i < row/2 (cdi.cc:435-439); written only
  // when a step accumulated (the reference divides uninitialised weights
  // otherwise)
  std::vector<double> pr((size_t)row / 2), wt((size_t)row / 2);
  int count = 0;
  chk(cdx_prtf(ctx, pr.data(), wt.data(), row / 2, &count), "prtf");
  if (count > 0) {
    FILE* f = fopen("prtf.txt", "w");
    if (!f) throw cdx::Error(CDX_ERR_IO, "phaseRetrieve: cannot write prtf.txt");
    for (int i = 0; i < row / 2; i++) fprintf(f, "%d, %f\n", i, (float)(pr[i] / wt[i]));
    fclose(f);
  }
  return patternWave;
}

// applySupport dispatcher, src/post_gpu/cdi.cc:25-36
void applySupport(void* gkp1, void* gkprime, int algo, Real* spt) {
  if (!g_last || !g_last->ctx) throw cdx::Error(CDX_ERR_ARG, "applySupport: no CDI context");
  chk(cdx_apply_support(g_last->ctx, gkp1, gkprime, algo, spt), "applySupport");
}
