// EM file format reader/writer (host code; the on-disk boundary of the path).
//
// Header (512 bytes, Oscar3/c++Picker/include/tom/core/io.h:49-63,
// IOfun/tom_emread.m:153-176): int8 machine, byte2, byte3, type;
// uint32 dims[3]; char comment[80]; int32 emdata[40]; uint8 userdata[256].
// machine 0 (OS-9), 3 (SGI), 5 (Mac) are big-endian, the rest (6 = PC)
// little-endian (tom_emread.m:160-164). type: 1 int8, 2 int16, 4 int32,
// 5 float32, 8 complex float32, 9 float64, 10 complex float64. Data follow
// x fastest.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/tomhip.h"

namespace {

thread_local std::string em_err;

bool host_big_endian() {
  const uint16_t v = 1;
  return *reinterpret_cast<const uint8_t*>(&v) == 0;
}

void swap_bytes(void* p, size_t elsize, size_t n) {
  uint8_t* b = static_cast<uint8_t*>(p);
  for (size_t i = 0; i < n; ++i, b += elsize)
    for (size_t k = 0; k < elsize / 2; ++k) std::swap(b[k], b[elsize - 1 - k]);
}

bool file_big_endian(int8_t machine) { return machine == 0 || machine == 3 || machine == 5; }

int elsize_of(int8_t type) {
  switch (type) {
    case 1: return 1;
    case 2: return 2;
    case 4: return 4;
    case 5: return 4;
    case 8: return 8;
    case 9: return 8;
    case 10: return 16;
    default: return 0;
  }
}

int fail(int code, const std::string& msg) {
  em_err = msg;
  return code;
}

int read_header(FILE* f, tomhip_em_header* h) {
  uint8_t raw[512];
  if (fread(raw, 1, 512, f) != 512) return fail(-1, "file shorter than the 512-byte EM header");
  std::memcpy(h, raw, 512);
  const bool swap = file_big_endian(h->machine) != host_big_endian();
  if (swap) {
    swap_bytes(h->dims, 4, 3);
    swap_bytes(h->emdata, 4, 40);
  }
  return 0;
}

}  // namespace

static_assert(sizeof(tomhip_em_header) == 512, "EM header must be 512 bytes");

extern "C" {

const char* tomhip_em_last_error(void) { return em_err.c_str(); }

int tomhip_em_read_header(const char* path, tomhip_em_header* hdr) {
  if (!path || !hdr) return fail(-1, "null pointer");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(-1, std::string("cannot open ") + path);
  const int rc = read_header(f, hdr);
  fclose(f);
  return rc;
}

static int em_read_f32(const char* path, float* out, int64_t capacity, tomhip_em_header* hdr_out);

int tomhip_em_read_f32(const char* path, float* out, int64_t capacity, tomhip_em_header* hdr_out) {
  // no C++ exception crosses the C ABI
  try {
    return em_read_f32(path, out, capacity, hdr_out);
  } catch (const std::bad_alloc&) {
    return fail(-2, "out of host memory");
  } catch (const std::exception& e) {
    return fail(-1, e.what());
  }
}

static int em_read_f32(const char* path, float* out, int64_t capacity, tomhip_em_header* hdr_out) {
  if (!path || !out) return fail(-1, "null pointer");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(-1, std::string("cannot open ") + path);
  tomhip_em_header h;
  int rc = read_header(f, &h);
  if (rc) {
    fclose(f);
    return rc;
  }
  const int es = elsize_of(h.type);
  if (es == 0 || h.type == 8 || h.type == 10) {
    fclose(f);
    return fail(-3, "EM data type " + std::to_string((int)h.type) + " is not a real type");
  }
  // element count without overflow (three 32-bit dims), checked against
  // the caller's buffer and against the bytes the file holds
  const unsigned __int128 n128 = (unsigned __int128)h.dims[0] * h.dims[1] * h.dims[2];
  if (n128 > (unsigned __int128)(capacity < 0 ? 0 : capacity)) {
    fclose(f);
    return fail(-1, "output buffer too small for the header's dims");
  }
  const int64_t n = (int64_t)n128;
  long here = ftell(f);
  if (here < 0 || fseek(f, 0, SEEK_END) != 0) {
    fclose(f);
    return fail(-1, "cannot size the EM file");
  }
  const long end = ftell(f);
  if (end < 0 || fseek(f, here, SEEK_SET) != 0) {
    fclose(f);
    return fail(-1, "cannot size the EM file");
  }
  if ((unsigned __int128)(end - here) < n128 * (unsigned)es) {
    fclose(f);
    return fail(-1, "EM file shorter than its header says");
  }
  std::vector<uint8_t> raw((size_t)n * es);
  if (fread(raw.data(), es, (size_t)n, f) != (size_t)n) {
    fclose(f);
    return fail(-1, "EM file shorter than its header says");
  }
  fclose(f);
  if (es > 1 && file_big_endian(h.machine) != host_big_endian()) swap_bytes(raw.data(), es, (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const uint8_t* p = raw.data() + (size_t)i * es;
    switch (h.type) {
      case 1: out[i] = (float)*reinterpret_cast<const int8_t*>(p); break;
      case 2: { int16_t v; std::memcpy(&v, p, 2); out[i] = (float)v; break; }
      case 4: { int32_t v; std::memcpy(&v, p, 4); out[i] = (float)v; break; }
      case 5: { float v; std::memcpy(&v, p, 4); out[i] = v; break; }
      case 9: { double v; std::memcpy(&v, p, 8); out[i] = (float)v; break; }
    }
  }
  if (hdr_out) *hdr_out = h;
  return 0;
}

static int em_write_f32(const char* path, const float* data, const int64_t dims[3], const tomhip_em_header* hdr);

int tomhip_em_write_f32(const char* path, const float* data, const int64_t dims[3], const tomhip_em_header* hdr) {
  try {
    return em_write_f32(path, data, dims, hdr);
  } catch (const std::bad_alloc&) {
    return fail(-2, "out of host memory");
  } catch (const std::exception& e) {
    return fail(-1, e.what());
  }
}

static int em_write_f32(const char* path, const float* data, const int64_t dims[3], const tomhip_em_header* hdr) {
  if (!path || !data || !dims) return fail(-1, "null pointer");
  tomhip_em_header h;
  std::memset(&h, 0, sizeof(h));
  if (hdr) h = *hdr;
  h.machine = 6;  // PC, little-endian
  h.type = 5;
  for (int d = 0; d < 3; ++d) {
    if (dims[d] < 1 || dims[d] > 0xffffffffLL) return fail(-1, "bad dims");
    h.dims[d] = (uint32_t)dims[d];
  }
  if ((unsigned __int128)dims[0] * dims[1] * dims[2] > (unsigned __int128)(INT64_MAX / 8))
    return fail(-1, "bad dims");
  const int64_t n = dims[0] * dims[1] * dims[2];
  FILE* f = fopen(path, "wb");
  if (!f) return fail(-1, std::string("cannot create ") + path);
  tomhip_em_header hw = h;
  std::vector<float> tmp;
  const float* src = data;
  if (host_big_endian()) {
    swap_bytes(hw.dims, 4, 3);
    swap_bytes(hw.emdata, 4, 40);
    tmp.assign(data, data + n);
    swap_bytes(tmp.data(), 4, (size_t)n);
    src = tmp.data();
  }
  const bool ok = fwrite(&hw, 1, 512, f) == 512 && fwrite(src, 4, (size_t)n, f) == (size_t)n;
  fclose(f);
  return ok ? 0 : fail(-1, "write failed");
}

}  // extern "C"
