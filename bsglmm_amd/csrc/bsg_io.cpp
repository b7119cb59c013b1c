// bsg_io.cpp -- NIfTI-1 single-file images and the design file (bsglmm_io.h).
//
// Header fields are read and written at their NIfTI-1 byte offsets; only the
// fields the reference program reads or writes are interpreted.
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/bsglmm.h"
#include "../../include/bsglmm_io.h"
#include "bsg_internal.h"

using bsg::set_error;

namespace {

constexpr int HDR = 348;            // sizeof_hdr of NIfTI-1
constexpr int OFF_DIM = 40;         // short dim[8]
constexpr int OFF_DATATYPE = 70;    // short
constexpr int OFF_BITPIX = 72;      // short
constexpr int OFF_PIXDIM = 76;      // float pixdim[8]
constexpr int OFF_VOX_OFFSET = 108; // float
constexpr int OFF_XYZT_UNITS = 123; // char
constexpr int OFF_MAGIC = 344;      // char[4]

bool ends_with(const std::string &s, const char *suf)
{
  const size_t n = strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

bool exists(const std::string &p)
{
  FILE *f = fopen(p.c_str(), "rb");
  if (!f) return false;
  fclose(f);
  return true;
}

// nifti_expand's name rule (read_data.cpp:81-116)
int resolve(const char *name, std::string &out)
{
  if (!name) {
    set_error("NIfTI: NULL file name");
    return BSG_ERR_ARG;
  }
  std::string s(name);
  const bool gz = ends_with(s, ".nii.gz"), nii = ends_with(s, ".nii");
  if (!gz && !nii) {
    set_error("Image ('" + s + "') is not single file NIFTI (.nii or .nii.gz)");
    return BSG_ERR_ARG;
  }
  if (exists(s)) {
    out = s;
    return BSG_OK;
  }
  const std::string alt = gz ? s.substr(0, s.size() - 3) : s + ".gz";
  if (exists(alt)) {
    out = alt;
    return BSG_OK;
  }
  set_error("Cannot open '" + s + "'");
  return BSG_ERR_ARG;
}

template <typename T>
T get(const unsigned char *h, int off, bool swap)
{
  T v;
  memcpy(&v, h + off, sizeof(T));
  if (swap) {
    unsigned char b[sizeof(T)];
    memcpy(b, &v, sizeof(T));
    for (size_t i = 0; i < sizeof(T) / 2; i++) std::swap(b[i], b[sizeof(T) - 1 - i]);
    memcpy(&v, b, sizeof(T));
  }
  return v;
}

template <typename T>
void put(unsigned char *h, int off, T v)
{
  memcpy(h + off, &v, sizeof(T));
}

struct Header {
  int dims[3];
  int datatype, bitpix;
  long vox_offset;
  bool swap;
};

int read_header(gzFile f, const std::string &path, Header &hd)
{
  unsigned char h[HDR];
  if (gzread(f, h, HDR) != HDR) {
    set_error("NIfTI: cannot read the header of '" + path + "'");
    return BSG_ERR_ARG;
  }
  int32_t sz;
  memcpy(&sz, h, 4);
  if (sz == HDR) {
    hd.swap = false;
  } else if (__builtin_bswap32((uint32_t)sz) == (uint32_t)HDR) {
    hd.swap = true;
  } else {
    set_error("NIfTI: '" + path + "' is not a NIfTI-1 file (sizeof_hdr " + std::to_string(sz) + ")");
    return BSG_ERR_ARG;
  }
  const short ndim = get<short>(h, OFF_DIM, hd.swap);
  for (int k = 0; k < 3; k++) {
    const short d = get<short>(h, OFF_DIM + 2 * (k + 1), hd.swap);
    hd.dims[k] = (k < ndim) ? d : 1;
  }
  if (ndim < 1 || ndim > 7 || hd.dims[0] < 1 || hd.dims[1] < 1 || hd.dims[2] < 1) {
    set_error("NIfTI: bad dimensions in '" + path + "'");
    return BSG_ERR_ARG;
  }
  hd.datatype = get<short>(h, OFF_DATATYPE, hd.swap);
  hd.bitpix = get<short>(h, OFF_BITPIX, hd.swap);
  const float vo = get<float>(h, OFF_VOX_OFFSET, hd.swap);
  hd.vox_offset = vo < 352.0f ? 352 : (long)vo;  // single-file images start at 352 or later
  return BSG_OK;
}

size_t type_size(int dt)
{
  switch (dt) {
    case BSG_NIFTI_UINT8: return 1;
    case BSG_NIFTI_INT16: return 2;
    case BSG_NIFTI_INT32: return 4;
    case BSG_NIFTI_FLOAT32: return 4;
    case BSG_NIFTI_FLOAT64: return 8;
    default: return 0;
  }
}

bool gz_read_all(gzFile f, void *dst, size_t bytes)
{
  unsigned char *p = static_cast<unsigned char *>(dst);
  while (bytes > 0) {
    const unsigned chunk = bytes > (1u << 30) ? (1u << 30) : (unsigned)bytes;
    const int got = gzread(f, p, chunk);
    if (got <= 0) return false;
    p += got;
    bytes -= (size_t)got;
  }
  return true;
}

bool gz_write_all(gzFile f, const void *src, size_t bytes)
{
  const unsigned char *p = static_cast<const unsigned char *>(src);
  while (bytes > 0) {
    const unsigned chunk = bytes > (1u << 30) ? (1u << 30) : (unsigned)bytes;
    const int put_n = gzwrite(f, p, chunk);
    if (put_n <= 0) return false;
    p += put_n;
    bytes -= (size_t)put_n;
  }
  return true;
}

}  // namespace

extern "C" {

int bsg_nifti_resolve(const char *name, char *resolved, size_t cap)
{
  std::string p;
  int rc = resolve(name, p);
  if (rc != BSG_OK) return rc;
  if (resolved) {
    if (p.size() + 1 > cap) {
      set_error("bsg_nifti_resolve: output buffer too small");
      return BSG_ERR_ARG;
    }
    memcpy(resolved, p.c_str(), p.size() + 1);
  }
  return BSG_OK;
}

int bsg_nifti_info(const char *name, int dims[3], int *datatype)
{
  std::string p;
  int rc = resolve(name, p);
  if (rc != BSG_OK) return rc;
  gzFile f = gzopen(p.c_str(), "rb");
  if (!f) {
    set_error("Cannot open '" + p + "'");
    return BSG_ERR_ARG;
  }
  Header hd;
  rc = read_header(f, p, hd);
  gzclose(f);
  if (rc != BSG_OK) return rc;
  if (dims)
    for (int k = 0; k < 3; k++) dims[k] = hd.dims[k];
  if (datatype) *datatype = hd.datatype;
  return BSG_OK;
}

int bsg_nifti_read(const char *name, double *out, size_t n)
{
  std::string p;
  int rc = resolve(name, p);
  if (rc != BSG_OK) return rc;
  gzFile f = gzopen(p.c_str(), "rb");
  if (!f) {
    set_error("Cannot open '" + p + "'");
    return BSG_ERR_ARG;
  }
  Header hd;
  if ((rc = read_header(f, p, hd)) != BSG_OK) {
    gzclose(f);
    return rc;
  }
  const size_t nv = (size_t)hd.dims[0] * hd.dims[1] * hd.dims[2];
  const size_t ts = type_size(hd.datatype);
  if (!ts || hd.bitpix != (int)(8 * ts)) {
    gzclose(f);
    set_error("NIfTI: unsupported datatype " + std::to_string(hd.datatype) + " in '" + p + "'");
    return BSG_ERR_ARG;
  }
  if (n != nv || !out) {
    gzclose(f);
    set_error("bsg_nifti_read: '" + p + "' holds " + std::to_string(nv) + " voxels, caller asked for " +
              std::to_string(n));
    return BSG_ERR_ARG;
  }
  if (gzseek(f, hd.vox_offset, SEEK_SET) != hd.vox_offset) {
    gzclose(f);
    set_error("NIfTI: cannot seek to the data of '" + p + "'");
    return BSG_ERR_ARG;
  }
  std::vector<unsigned char> raw(nv * ts);
  const bool ok = gz_read_all(f, raw.data(), raw.size());
  gzclose(f);
  if (!ok) {
    set_error("Error reading volume 1 from " + p);
    return BSG_ERR_ARG;
  }
  const unsigned char *r = raw.data();
  for (size_t i = 0; i < nv; i++) {
    const unsigned char *e = r + i * ts;
    switch (hd.datatype) {
      case BSG_NIFTI_UINT8: out[i] = (double)e[0]; break;
      case BSG_NIFTI_INT16: out[i] = (double)get<int16_t>(e, 0, hd.swap); break;
      case BSG_NIFTI_INT32: out[i] = (double)get<int32_t>(e, 0, hd.swap); break;
      case BSG_NIFTI_FLOAT32: out[i] = (double)get<float>(e, 0, hd.swap); break;
      default: out[i] = get<double>(e, 0, hd.swap); break;
    }
  }
  return BSG_OK;
}

int bsg_nifti_write_f64(const char *name, int nx, int ny, int nz, const double *data)
{
  if (!name || !data || nx < 1 || ny < 1 || nz < 1 || nx > 32767 || ny > 32767 || nz > 32767) {
    set_error("bsg_nifti_write_f64: bad argument");
    return BSG_ERR_ARG;
  }
  std::string p(name);
  const bool gz = ends_with(p, ".gz");
  if (!ends_with(p, gz ? ".nii.gz" : ".nii")) {
    set_error("bsg_nifti_write_f64: file to be written must end with .nii or .nii.gz");
    return BSG_ERR_ARG;
  }
  unsigned char h[352];
  memset(h, 0, sizeof(h));
  put<int32_t>(h, 0, HDR);
  const short dim[5] = {3, (short)nx, (short)ny, (short)nz, 1};
  for (int k = 0; k < 5; k++) put<short>(h, OFF_DIM + 2 * k, dim[k]);
  put<short>(h, OFF_DATATYPE, (short)BSG_NIFTI_FLOAT64);
  put<short>(h, OFF_BITPIX, (short)64);
  for (int k = 1; k <= 3; k++) put<float>(h, OFF_PIXDIM + 4 * k, 2.0f);
  put<float>(h, OFF_VOX_OFFSET, 352.0f);
  h[OFF_XYZT_UNITS] = 2;  // NIFTI_UNITS_MM
  memcpy(h + OFF_MAGIC, "n+1\0", 4);
  const size_t bytes = (size_t)nx * ny * nz * sizeof(double);
  bool ok;
  if (gz) {
    gzFile f = gzopen(p.c_str(), "wb6");
    if (!f) {
      set_error("Error opening header file " + p + " for write");
      return BSG_ERR_ARG;
    }
    ok = gz_write_all(f, h, sizeof(h)) && gz_write_all(f, data, bytes);
    ok = (gzclose(f) == Z_OK) && ok;
  } else {
    FILE *f = fopen(p.c_str(), "wb");
    if (!f) {
      set_error("Error opening header file " + p + " for write");
      return BSG_ERR_ARG;
    }
    ok = fwrite(h, 1, sizeof(h), f) == sizeof(h) && fwrite(data, 1, bytes, f) == bytes;
    ok = (fclose(f) == 0) && ok;
  }
  if (!ok) {
    set_error("Error writing data to " + p);
    return BSG_ERR_ARG;
  }
  return BSG_OK;
}

int bsg_design_read(const char *path, int ncovar, int *nsubs, float *covar, char *names, char *images)
{
  if (!path || ncovar < 1 || !nsubs) {
    set_error("bsg_design_read: bad argument");
    return BSG_ERR_ARG;
  }
  std::ifstream in(path);
  if (!in) {
    set_error(std::string("Cannot open design file '") + path + "'");
    return BSG_ERR_ARG;
  }
  std::vector<std::string> tok;
  std::string t;
  while (in >> t) tok.push_back(t);
  const size_t row = (size_t)ncovar + 2;
  if (tok.size() < row) {
    set_error("design file: fewer than NCov + 2 header tokens");
    return BSG_ERR_ARG;
  }
  const size_t body = tok.size() - row;
  if (body % row != 0 || body == 0) {
    set_error("design file: " + std::to_string(body) + " tokens after the header are not whole rows of " +
              std::to_string(row) + " (subject, " + std::to_string(ncovar) + " covariates, image)");
    return BSG_ERR_ARG;
  }
  const int n = (int)(body / row);
  *nsubs = n;
  if (names)
    for (int c = 0; c < ncovar; c++) {
      const std::string &s = tok[1 + c];
      const size_t k = s.size() < 63 ? s.size() : 63;
      memcpy(names + 64 * c, s.data(), k);
      names[64 * c + k] = 0;
    }
  for (int i = 0; i < n; i++) {
    const size_t b = row * (i + 1);
    if (covar)
      for (int c = 0; c < ncovar; c++) covar[(size_t)c * n + i] = (float)atof(tok[b + 1 + c].c_str());
    if (images) {
      const std::string img = "./images/" + tok[b + row - 1];
      if (img.size() > 511) {
        set_error("design file: image name too long: " + tok[b + row - 1]);
        return BSG_ERR_ARG;
      }
      memcpy(images + 512 * (size_t)i, img.c_str(), img.size() + 1);
    }
  }
  return BSG_OK;
}

}  // extern "C"
