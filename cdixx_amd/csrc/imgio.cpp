// imgio.cpp -- the .bin image format of CDIxx (host side).
// Header: struct imageFile {char type; int rows; int cols; int nchann; int
// typesize;} (include/imageFile.hpp:3-17), 20 bytes with natural alignment,
// then rows*cols elements.  Reader: readImage_c .bin branch
// (src/independant/imgio.c:23-43, FLOAT and FLOAT2 only there; DOUBLE and
// DOUBLE2 are accepted here as well).  Writers: writeComplexImage /
// writeFloatImage (imgio.c:223-243).
#include "../../include/cdx_abi.h"
#include "errors.hpp"
#include <stdio.h>
#include <string.h>
#include <string>

namespace {
struct imageFile {
  char type;
  int rows;
  int cols;
  int nchann;
  int typesize;
};
static_assert(sizeof(imageFile) == 20, "imageFile header must be 20 bytes");
const int kTypeSizes[] = {0, 1, 2, 4, 8, 8, 16};

int io_fail(int code, const std::string& m) {
  cdx::set_last_error(m.c_str());
  return code;
}
}  // namespace

extern "C" int cdx_bin_read_header(const char* path, int* type, int* rows, int* cols) {
  FILE* f = path ? fopen(path, "rb") : nullptr;
  if (!f) return io_fail(CDX_ERR_IO, std::string("cannot open ") + (path ? path : "(null)"));
  imageFile h;
  size_t got = fread(&h, sizeof h, 1, f);
  fclose(f);
  if (got != 1) return io_fail(CDX_ERR_IO, std::string(path) + ": empty or truncated header");
  if (h.type < 3 || h.type > 6 || h.rows <= 0 || h.cols <= 0)
    return io_fail(CDX_ERR_IO, std::string(path) + ": unsupported .bin type " + std::to_string((int)h.type));
  if (type) *type = h.type;
  if (rows) *rows = h.rows;
  if (cols) *cols = h.cols;
  return CDX_OK;
}

extern "C" int cdx_bin_read(const char* path, void* host, size_t cap) {
  int type, rows, cols;
  int rc = cdx_bin_read_header(path, &type, &rows, &cols);
  if (rc) return rc;
  size_t bytes = (size_t)rows * cols * kTypeSizes[type];
  if (!host || cap < bytes) return io_fail(CDX_ERR_ARG, std::string(path) + ": destination buffer too small");
  FILE* f = fopen(path, "rb");
  if (!f) return io_fail(CDX_ERR_IO, std::string("cannot open ") + path);
  fseek(f, (long)sizeof(imageFile), SEEK_SET);
  size_t got = fread(host, 1, bytes, f);
  fclose(f);
  if (got != bytes) return io_fail(CDX_ERR_IO, std::string(path) + ": truncated data");
  return CDX_OK;
}

extern "C" int cdx_bin_write(const char* path, int type, int rows, int cols, const void* host) {
  if (type < 3 || type > 6 || rows <= 0 || cols <= 0 || !host)
    return io_fail(CDX_ERR_ARG, "cdx_bin_write: bad arguments");
  FILE* f = path ? fopen(path, "wb") : nullptr;
  if (!f) return io_fail(CDX_ERR_IO, std::string("cannot write ") + (path ? path : "(null)"));
  imageFile h;
  memset(&h, 0, sizeof h);
  h.type = (char)type;
  h.rows = rows;
  h.cols = cols;
  h.nchann = 1;
  h.typesize = kTypeSizes[type];
  size_t bytes = (size_t)rows * cols * kTypeSizes[type];
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(host, 1, bytes, f) == bytes;
  ok = (fclose(f) == 0) && ok;
  if (!ok) return io_fail(CDX_ERR_IO, std::string("short write to ") + path);
  return CDX_OK;
}
