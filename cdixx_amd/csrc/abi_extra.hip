// abi_extra.hip -- small C-ABI entry points that need no context state.
#include "../../include/cdx_abi.h"
#include "errors.hpp"
#include "kernels.hpp"
#include <hip/hip_runtime.h>
#include <string>

// device-to-device copy as a kernel on the given stream (capturable in a
// hipGraph as a kernel node, unlike a memcpy node)
extern "C" int cdx_copy(void* stream, void* dst, const void* src, size_t bytes) {
  if (!dst || !src) {
    cdx::set_last_error("cdx_copy: null buffer");
    return CDX_ERR_ARG;
  }
  hipError_t e = cdx::launch_copy(bytes, src, dst, (hipStream_t)stream);
  if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) {
    cdx::set_last_error((std::string("cdx_copy: ") + hipGetErrorString(e)).c_str());
    return CDX_ERR_HIP;
  }
  return CDX_OK;
}
