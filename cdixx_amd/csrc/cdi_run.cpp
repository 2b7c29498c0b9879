// cdi_run.cpp -- command-line driver, the equivalent of CDIxx util/cdi.cpp:8-50:
// construct -> readFiles -> init -> prepareIter -> phaseRetrieve -> saveState,
// then nIter restarts keeping the lowest residual.
#include "../../include/cdi.hpp"
#include "../../include/cdx_abi.h"
#include <stdio.h>

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s config.cfg\n", argv[0]);
    return 2;
  }
  try {
    CDI setups(argv[1]);
    setups.readFiles();
    setups.init();
    printf("Imaging distance = %4.2fcm\n", setups.d * 1e-4);
    setups.prepareIter();
    if (setups.doIteration) {
      setups.phaseRetrieve();
      setups.saveState();
    }
    double smallresidual = setups.residual;
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
    setups.checkAutoCorrelation();
    cdx_synchronize(setups.ctx);
  } catch (const cdx::Error& e) {
    fprintf(stderr, "cdi_run: %s\n", e.what());
    return e.code ? e.code : 1;
  }
  return 0;
}
