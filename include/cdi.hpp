/*
 * cdi.hpp -- drop-in for CDIxx include/cdi.hpp:1-43 (class CDI).
 *
 * Same public members and methods, so existing callers (util/cdi.cpp,
 * html/runCDI.sh through cdi_run) recompile unchanged.  Device buffers are
 * raw pointers owned by the object, as in the reference; the hidden GPU
 * state of the reference lives in `ctx` (a cdx_ctx, include/cdx_abi.h).
 * Real = float, complexFormat = interleaved float pairs (_Complex float in
 * the reference, include/format.hpp:131-132).
 *
 * Failures throw cdx::Error (carrying a CDX_ERR_* code) where the reference
 * calls exit()/abort().
 */
#ifndef CDX_CDI_HPP
#define CDX_CDI_HPP
#include <string>
#include "readConfig.hpp"

#ifndef Real
#define Real float
#endif
#ifndef complexFormat
#define complexFormat _Complex float
#endif

struct cdx_ctx;

/* Fresnel propagator parameters (include/propagator.hpp); the propagation
 * itself (isFresnel) is outside this path. */
struct propagator {
  int row = 0, column = 0;
  double pixelsize = 0, lambda = 0, distance = 0;
};

class CDI : public readConfig {
 public:
  CDI(const char* configfile);
  ~CDI();
  Real* patternData = 0;
  Real* beamstop = 0;
  propagator propagate;
  complexFormat* patternWave = 0;
  complexFormat* autoCorrelation = 0;
  int row = 0, column = 0;
  complexFormat* objectWave = 0;
  Real* support = 0;
  Real residual = 0;
  Real resolution = 0;
  void* cuda_spt = 0;
  void* mnist_dat = 0;
  std::string save_suffix = "";
  void* devstates = 0;
  cdx_ctx* ctx = 0;   /* explicit replacement of the reference's global GPU state */
  void allocateMem();
  void readObjectWave();
  void readPattern();
  void readFiles();
  void setPattern(Real* pattern);
  void setPattern_c(complexFormat* pattern);
  void init();
  void prepareIter();
  void checkAutoCorrelation();
  void createSupport();
  void initSupport();
  complexFormat* phaseRetrieve();
  void saveState();

 private:
  unsigned long long seed_ = 0;
  unsigned long long draws_ = 0;
  int threshold_set_ = 0;
};

/* applySupport(gkp1, gkprime, algo, spt), include/cdi.hpp:43 (src/post_gpu/cdi.cc:25-36),
 * on the context of the last constructed CDI object. */
void applySupport(void* gkp1, void* gkprime, int algo, Real* spt);
#endif
