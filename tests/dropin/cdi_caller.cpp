// cdi_caller.cpp -- a caller of the drop-in CDI class (include/cdi.hpp)
// written the way the reference's own drivers use it: the util/cdi.cpp:8-50
// sequence (construct from a config, readFiles, init, prepareIter,
// phaseRetrieve, saveState, nIter restarts keeping the lowest residual,
// checkAutoCorrelation) and direct member access (util/mono.cpp:205-219
// reads and writes cdi.patternData / cdi.patternWave).  Test program: it
// also writes the simulated pattern (pattern.bin, FFT order) for a
// forward-model check, checks the saveState -> restart round trip, and
// prints one JSON line.  Exit status 0 when every check holds.
#include "cdi.hpp"
#include "cdx_abi.h"

#include <stdio.h>
#include <string.h>
#include <vector>

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s config.cfg\n", argv[0]);
    return 2;
  }
  try {
    CDI setups(argv[1]);
    setups.readFiles();
    setups.init();
    setups.prepareIter();
    const size_t N = (size_t)setups.row * setups.column;
    // member access: the simulated pattern as prepareIter leaves it
    std::vector<float> pat(N);
    if (cdx_memcpy_d2h(pat.data(), setups.patternData, N * sizeof(float)) ||
        cdx_bin_write("pattern.bin", 3, setups.row, setups.column, pat.data()))
      throw cdx::Error(CDX_ERR_IO, cdx_last_error());
    double smallresidual = 0;
    if (setups.doIteration) {
      setups.phaseRetrieve();
      setups.saveState();
      smallresidual = setups.residual;
    }
    for (int i = 0; i < setups.nIter; i++) {
      setups.prepareIter();
      if (setups.doIteration) {
        setups.phaseRetrieve();
        if (smallresidual > setups.residual) {
          smallresidual = setups.residual;
          setups.saveState();
        }
      }
    }
    // saveState -> restart: the saved patternWave is the start of the next
    // prepareIter, bit for bit (cdi.cc:201-221,255-259)
    std::vector<float> saved(2 * N), back(2 * N);
    int type = 0, r = 0, c = 0;
    if (cdx_bin_read_header(setups.common.restart, &type, &r, &c) || type != 4 || r != setups.row ||
        c != setups.column || cdx_bin_read(setups.common.restart, saved.data(), saved.size() * sizeof(float)))
      throw cdx::Error(CDX_ERR_IO, "restart file unreadable");
    setups.restart = 1;
    setups.runSim = 0;  // keep the measured pattern; only the start comes from the file
    setups.prepareIter();
    if (cdx_memcpy_d2h(back.data(), setups.patternWave, back.size() * sizeof(float)))
      throw cdx::Error(CDX_ERR_HIP, cdx_last_error());
    const bool restart_ok = memcmp(saved.data(), back.data(), saved.size() * sizeof(float)) == 0 && setups.restart;
    setups.checkAutoCorrelation();
    cdx_synchronize(setups.ctx);
    printf("{\"rows\": %d, \"cols\": %d, \"residual\": %.9g, \"smallest_residual\": %.9g, \"restart_bitexact\": %s}\n",
           setups.row, setups.column, (double)setups.residual, smallresidual, restart_ok ? "true" : "false");
    return restart_ok ? 0 : 1;
  } catch (const cdx::Error& e) {
    fprintf(stderr, "cdi_caller: %s\n", e.what());
    return e.code ? e.code : 1;
  }
}
