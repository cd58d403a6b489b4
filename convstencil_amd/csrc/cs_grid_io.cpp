// cs_grid_io.cpp -- host-side grid files: the reference's text dumps and a binary FP64 format.
//
// Text: the format of the reference's save_to_txt (src/1d/main.cu:21-36,
// src/2d/main.cu:36-55, src/3d/main.cu:22-40), which its drivers write to
// output_gpu.txt when built with WRITE_OUTPUT (src/2d/main.cu:410-415): the whole padded
// array, each value printed with "%.0f" (rounded to an integer, as the reference does for
// its integer-valued fills), 1-D one "index value" line per cell, 2-D one line of
// tab-terminated values per row, 3-D the rows of each plane followed by "\nNext\n".
//
// Binary (.f64): a 96-byte header -- the 8 bytes "CSGRID64", uint32 format version 1,
// uint32 ndim, int64 ext[3], int64 org[3], int64 n[3], 8 reserved zero bytes -- then the
// padded array as little-endian IEEE doubles, row-major, slowest axis first.  Values
// round-trip exactly, so fixtures and dumps can be compared bit for bit.
#include <cstdio>
#include <cstring>

#include "../../include/convstencil.h"

namespace {

constexpr char kMagic[8] = {'C', 'S', 'G', 'R', 'I', 'D', '6', '4'};

struct Header {
  char magic[8];
  uint32_t version, ndim;
  int64_t ext[3], org[3], n[3];
  uint64_t reserved;
};
static_assert(sizeof(Header) == 96, "fixed header size");

// Extents come from file headers too, so every product is checked: the cell count must
// fit, in bytes, in an int64 (a wrapped product could pass a capacity check while the
// extents describe far more cells than were read).
bool layout_ok(const cs_layout *L) {
  if (!L || L->ndim < 1 || L->ndim > 3) return false;
  for (int d = 0; d < 3; ++d) {
    if (L->ext[d] < 1 || L->n[d] < 0 || L->org[d] < 0 || L->org[d] > L->ext[d] || L->n[d] > L->ext[d] - L->org[d])
      return false;
    if (d < 3 - L->ndim && (L->ext[d] != 1 || L->n[d] != 1 || L->org[d] != 0)) return false;
  }
  int64_t c = 0, bytes = 0;
  if (__builtin_mul_overflow(L->ext[0], L->ext[1], &c) || __builtin_mul_overflow(c, L->ext[2], &c) ||
      __builtin_mul_overflow(c, (int64_t)sizeof(double), &bytes))
    return false;
  return c > 0;
}

int64_t cells(const cs_layout *L) { return L->ext[0] * L->ext[1] * L->ext[2]; }

}  // namespace

extern "C" {

cs_status cs_grid_save_txt(const char *path, const cs_layout *layout, const double *host) {
  if (!path || !host || !layout_ok(layout)) return CS_ERR_INVALID;
  FILE *f = fopen(path, "w");
  if (!f) return CS_ERR_INVALID;
  const int64_t Z = layout->ext[0], Y = layout->ext[1], X = layout->ext[2];
  bool ok = true;
  if (layout->ndim == 1) {
    for (int64_t i = 0; i < X && ok; ++i) ok = fprintf(f, "%lld %.0f\n", (long long)i, host[i]) > 0;
  } else {
    for (int64_t z = 0; z < Z && ok; ++z) {
      for (int64_t y = 0; y < Y && ok; ++y) {
        const double *row = host + (z * Y + y) * X;
        for (int64_t x = 0; x < X && ok; ++x) ok = fprintf(f, "%.0f\t", row[x]) > 0;
        ok = ok && fputc('\n', f) != EOF;
      }
      if (layout->ndim == 3) ok = ok && fputs("\nNext\n", f) != EOF;
    }
  }
  ok = (fclose(f) == 0) && ok;
  return ok ? CS_OK : CS_ERR_INVALID;
}

cs_status cs_grid_save_f64(const char *path, const cs_layout *layout, const double *host) {
  if (!path || !host || !layout_ok(layout)) return CS_ERR_INVALID;
  Header h{};
  memcpy(h.magic, kMagic, 8);
  h.version = 1;
  h.ndim = (uint32_t)layout->ndim;
  for (int d = 0; d < 3; ++d) {
    h.ext[d] = layout->ext[d];
    h.org[d] = layout->org[d];
    h.n[d] = layout->n[d];
  }
  FILE *f = fopen(path, "wb");
  if (!f) return CS_ERR_INVALID;
  const size_t nc = (size_t)cells(layout);
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(host, sizeof(double), nc, f) == nc;
  ok = (fclose(f) == 0) && ok;
  return ok ? CS_OK : CS_ERR_INVALID;
}

cs_status cs_grid_load_f64(const char *path, cs_layout *layout, double *host, int64_t capacity) {
  if (!path || !layout) return CS_ERR_INVALID;
  FILE *f = fopen(path, "rb");
  if (!f) return CS_ERR_INVALID;
  Header h{};
  cs_status s = CS_OK;
  if (fread(&h, sizeof h, 1, f) != 1 || memcmp(h.magic, kMagic, 8) != 0 || h.version != 1) {
    s = CS_ERR_INVALID;
  } else {
    cs_layout L{};
    L.ndim = (int32_t)h.ndim;
    for (int d = 0; d < 3; ++d) {
      L.ext[d] = h.ext[d];
      L.org[d] = h.org[d];
      L.n[d] = h.n[d];
    }
    if (!layout_ok(&L)) {
      s = CS_ERR_INVALID;
    } else {
      *layout = L;
      if (host) {
        const int64_t nc = cells(&L);
        if (capacity < nc)
          s = CS_ERR_INVALID;
        else if (fread(host, sizeof(double), (size_t)nc, f) != (size_t)nc)
          s = CS_ERR_INVALID;  // truncated file
      }
    }
  }
  fclose(f);
  return s;
}

}  // extern "C"
