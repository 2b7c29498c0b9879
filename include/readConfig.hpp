/*
 * readConfig.hpp -- drop-in for CDIxx include/readConfig.hpp:1-63.
 *
 * Same field names, types and defaults (the X-macro lists of
 * include/readConfig.hpp:25-35), same groups (InputImages.default/pupil and
 * Job, src/independant/readConfig.cc:38-60) and the same lookup rules: a key
 * of the wrong type leaves the default in place, a float field also accepts
 * an integer, unknown keys are ignored.  libconfig is replaced by a parser
 * of the subset CDIxx's .cfg files use.  Keys marked "new" do not exist in
 * the reference; their defaults reproduce the reference's fixed behaviour.
 *
 * Errors throw cdx::Error instead of exit() (readConfig.cc:33-35,132-135).
 */
#ifndef CDX_READCONFIG_HPP
#define CDX_READCONFIG_HPP
#include <deque>
#include <stdexcept>
#include <string>
#include <vector>

namespace cdx {
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
}  // namespace cdx

struct CDIfiles {
  const char* Pattern = nullptr;
  const char* Intensity = nullptr;
  const char* Phase = nullptr;
  const char* restart = nullptr;
};

#define DECINT(x, y) int x = y;
#define DECBOOL(x, y) bool x = y;
#define DECREAL(x, y) double x = y;
#define DECSTR(x, y) const char* x = y;

#define INTVAR(F)                                                                                           \
  F(beamStopSize, 5) F(nIter, 1) F(nIterpupil, 1) F(noiseLevel, 0) F(noiseLevel_pupil, 0) F(verbose, 2)   \
  F(mnistN, 3) F(cropPattern, 0) F(spectrumSamplingStep, 10) F(saveVideoEveryIter, 0) F(prtfIter, 1000)   \
  F(zernikeIter, 0) F(scanx, 3) F(scany, 3) F(computeErrorEveryIter, 0)                                    \
  /* new */ F(shrinkWrapMode, 0) F(startMode, 0) F(seed, -1)

#define BOOLVAR(F)                                                                                           \
  F(isFlip, 0) F(runSim, 0) F(simCCDbit, 0) F(isFresnel, 0) F(doIteration, 0) F(useGaussionLumination, 0)  \
  F(useGaussionHERALDO, 0) F(doCentral, 0) F(useRectHERALDO, 0) F(dopupil, 0) F(useDM, 0) F(useBS, 0)      \
  F(reconAC, 0) F(phaseModulation_pupil, 0) F(intensityModulation, 1) F(restart, 0) F(saveIter, 0)         \
  F(domnist, 0) F(saveLMDB, 0) F(solveSpectrum, 0) F(mPIE, 1) F(online, 0)                                 \
  /* new */ F(freeBeamstop, 0) F(useGraph, 1)

#define REALVAR(F)                                                                                           \
  F(exposure, 0) F(exposurepupil, 0) F(oversampling, 0) F(oversampling_spt, 0) F(lambda, 0) F(d, 0)        \
  F(dpupil, 0) F(pixelsize, 0) F(beamspotsize, 0) F(shrinkThreshold, 0) F(positionUncertainty, 0)          \
  F(costheta, 1) F(photoncount, 1e9) F(phaseModulation, 0) F(stepSize, 64) F(pupilSize, 500)               \
  /* new */ F(betaHIO, 1) F(relaxER, 1.9) F(shrinkSigma0, 3) F(shrinkSigmaMin, -1) F(shrinkSigmaDecay, 0.99)

#define STRVAR(F)                                                                                            \
  F(mnistData, "data") F(algorithm, "1000*HIO")                                                              \
  F(ccd_response, "/home/boyang/SharedNTFS/images/ccd_response.txt")                                        \
  F(spectrum, "/home/boyang/SharedNTFS/images/spectrum/spectrum_after_sample.txt") F(BSimage, "")           \
  F(outputDir, "./") F(distFile, "./scan.txt")

class readConfig {
 public:
  BOOLVAR(DECBOOL)
  INTVAR(DECINT)
  REALVAR(DECREAL)
  STRVAR(DECSTR)
  CDIfiles pupil;
  CDIfiles common;
  readConfig(const char* configfile);
  void print();
  ~readConfig() {}

 private:
  std::deque<std::string> strings_;  // storage behind the const char* fields (stable addresses)
};

/* AlgoParser, include/readConfig.hpp:51-62 and readConfig.cc:82-163:
 * "count*TOKEN+count*TOKEN+..." where TOKEN may be a parenthesised group.
 * next() returns the enum Algorithm value of the next step, -1 at the end. */
class AlgoParser {
 public:
  void* subParsersp;
  void* countp;
  void* algoListp;
  int nAlgo;
  unsigned int currentAlgo;
  int currentCount;
  AlgoParser(const char* formula);
  ~AlgoParser();
  void restart();
  int next();
};
#endif
