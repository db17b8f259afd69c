// draw_png.cpp — edge drawing of a layout into an RGBA PNG (part of
// libphde_graph.so, host only): the step after the layout path, as the
// reference's draw_graph.c does it.
//
// Followed from draw_graph.c:
//  - image size (:228-248): aspect = (xmax-xmin)/(ymax-ymin) over the two
//    chosen coordinates, with the extremes starting at xmax = ymax = -1 and
//    xmin = ymin = 1; width = sqrt(100 n / aspect), height = aspect * width,
//    both truncated, then divided by 10 when either exceeds 100,000 or by 5
//    when either exceeds 20,000;
//  - pixel positions (:252-263): (v - min) / (max - min) * size, clamped to
//    [0, size - 1], floored when drawn;
//  - background RGBA (4, 4, 4, 255) (:266-276);
//  - every directed edge (u ascending, then adjacency order) drawn with
//    Bresenham's line algorithm (:108-126), coloured by its length bin
//    floor(len / max_len * 7) in single precision, where len is the
//    Euclidean distance of the floored end points (:278-316): bins 0, 1, 5, 6
//    take their own colour while that bin's running count is within 5% of
//    the undirected edge count (bin 0 compares an integer quotient, so it
//    always qualifies; bin 0 falls through to bin 1's test, bin 6 to bin
//    5's), otherwise they move two bins inward; bins 2-4 have fixed colours;
//    bin 7 (the longest edges) keeps black (:45-107).
// The reference draws in parallel with unsynchronised counters; here the
// edges are drawn in order, which is what its single-thread run does.
#include <zlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "phde_graph.h"

namespace {

struct Rgba {
  unsigned char r, g, b, a;
};

struct Painter {
  std::vector<unsigned char> &img;
  unsigned width;
  long m;  // undirected edges
  long lw1 = 0, lw2 = 0, hg1 = 0, hg2 = 0;

  Rgba colour(int bin) {
    Rgba c{0, 0, 0, 255};
    bool set = false;
    switch (bin) {
      case 0:
        if ((double)(lw1 / m) <= 0.05) {
          c = {251, 83, 73, 255};
          lw1++;
          set = true;
          break;
        }
        [[fallthrough]];
      case 1:
        if ((double)lw2 / (double)m <= 0.05) {
          c = {251, 165, 4, 255};
          lw2++;
          set = true;
        }
        break;
      case 6:
        if ((double)hg2 / (double)m <= 0.05) {
          c = {13, 152, 186, 255};
          hg2++;
          set = true;
          break;
        }
        [[fallthrough]];
      case 5:
        if ((double)hg1 / (double)m <= 0.05) {
          c = {4, 251, 4, 255};
          hg1++;
          set = true;
        }
        break;
      default:
        break;
    }
    if ((bin == 0 || bin == 1) && !set) bin += 2;
    if ((bin == 5 || bin == 6) && !set) bin -= 2;
    switch (bin) {
      case 2:
        c = {251, 174, 66, 255};
        break;
      case 3:
        c = {251, 251, 4, 255};
        break;
      case 4:
        c = {154, 251, 50, 255};
        break;
      default:
        break;
    }
    return c;
  }

  void line(int x0, int y0, int x1, int y1, Rgba c) {
    const int dx = std::abs(x1 - x0), sx = x0 < x1 ? 1 : -1;
    const int dy = std::abs(y1 - y0), sy = y0 < y1 ? 1 : -1;
    int err = (dx > dy ? dx : -dy) / 2;
    for (;;) {
      unsigned char *p = &img[4 * ((size_t)width * y0 + x0)];
      p[0] = c.r;
      p[1] = c.g;
      p[2] = c.b;
      p[3] = c.a;
      if (x0 == x1 && y0 == y1) break;
      const int e2 = err;
      if (e2 > -dx) {
        err -= dy;
        x0 += sx;
      }
      if (e2 < dy) {
        err += dx;
        y0 += sy;
      }
    }
  }
};

double length(double x1, double y1, double x2, double y2) {
  return std::sqrt(std::pow(x2 - x1, 2) + std::pow(y2 - y1, 2));
}

bool put_chunk(FILE *f, const char *type, const unsigned char *data, size_t len) {
  unsigned char hdr[8] = {(unsigned char)(len >> 24), (unsigned char)(len >> 16), (unsigned char)(len >> 8),
                          (unsigned char)len, (unsigned char)type[0], (unsigned char)type[1],
                          (unsigned char)type[2], (unsigned char)type[3]};
  uLong crc = crc32(0L, Z_NULL, 0);
  crc = crc32(crc, hdr + 4, 4);
  if (len) crc = crc32(crc, data, (uInt)len);
  const unsigned char tail[4] = {(unsigned char)(crc >> 24), (unsigned char)(crc >> 16),
                                 (unsigned char)(crc >> 8), (unsigned char)crc};
  return fwrite(hdr, 1, 8, f) == 8 && (len == 0 || fwrite(data, 1, len, f) == len) &&
         fwrite(tail, 1, 4, f) == 4;
}

// 8-bit RGBA, no interlace, filter type 0 on every row, zlib level 6.
int write_png(const char *path, const std::vector<unsigned char> &img, unsigned w, unsigned h) {
  std::vector<unsigned char> raw((size_t)h * (4 * (size_t)w + 1));
  for (unsigned y = 0; y < h; y++) {
    raw[(size_t)y * (4 * w + 1)] = 0;
    std::memcpy(&raw[(size_t)y * (4 * w + 1) + 1], &img[(size_t)y * 4 * w], 4 * (size_t)w);
  }
  uLongf zlen = compressBound((uLong)raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return PG_ERR_NOMEM;
  FILE *f = fopen(path, "wb");
  if (!f) return PG_ERR_IO;
  static const unsigned char sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
  const unsigned char ihdr[13] = {(unsigned char)(w >> 24), (unsigned char)(w >> 16), (unsigned char)(w >> 8),
                                  (unsigned char)w,         (unsigned char)(h >> 24), (unsigned char)(h >> 16),
                                  (unsigned char)(h >> 8),  (unsigned char)h,         8, 6, 0, 0, 0};
  bool ok = fwrite(sig, 1, 8, f) == 8 && put_chunk(f, "IHDR", ihdr, 13) && put_chunk(f, "IDAT", z.data(), zlen) &&
            put_chunk(f, "IEND", nullptr, 0);
  ok = (fclose(f) == 0) && ok;
  return ok ? PG_OK : PG_ERR_IO;
}

}  // namespace

extern "C" {

int pg_draw_png(const pg_csr *g, const double *coords, int num_coord, int first, int second, const char *path,
                unsigned *width_out, unsigned *height_out) {
  if (!g || !coords || !path || g->n < 1 || num_coord < 1 || first < 0 || second < 0 || first >= num_coord ||
      second >= num_coord)
    return PG_ERR_ARG;
  const int64_t n = g->n;
  std::vector<double> vx(n), vy(n);
  double xmax = -1, xmin = 1, ymax = -1, ymin = 1;
  for (int64_t i = 0; i < n; i++) {
    vx[i] = coords[(size_t)first * n + i];
    vy[i] = coords[(size_t)second * n + i];
    if (vx[i] > xmax) xmax = vx[i];
    if (vx[i] < xmin) xmin = vx[i];
    if (vy[i] > ymax) ymax = vy[i];
    if (vy[i] < ymin) ymin = vy[i];
  }
  const double aspect = (xmax - xmin) / (ymax - ymin);
  unsigned width = (unsigned)std::sqrt(100 * n / aspect);
  unsigned height = (unsigned)(aspect * width);
  if (width > 100000 || height > 100000) {
    width /= 10;
    height /= 10;
  } else if (width > 20000 || height > 20000) {
    width /= 5;
    height /= 5;
  }
  if (width < 1 || height < 1) return PG_ERR_RANGE;
  const double xri = 1.0 / (xmax - xmin), yri = 1.0 / (ymax - ymin);
  for (int64_t i = 0; i < n; i++) {
    vx[i] = (-xmin + vx[i]) * xri * width;
    vy[i] = (-ymin + vy[i]) * yri * height;
    if (vx[i] < 0) vx[i] = 0;
    if (vy[i] < 0) vy[i] = 0;
    if (vx[i] > (width - 1)) vx[i] = width - 1;
    if (vy[i] > (height - 1)) vy[i] = height - 1;
  }
  std::vector<unsigned char> img((size_t)width * height * 4);
  for (size_t p = 0; p < (size_t)width * height; p++) {
    img[4 * p + 0] = 4;
    img[4 * p + 1] = 4;
    img[4 * p + 2] = 4;
    img[4 * p + 3] = 255;
  }
  float max_len = 0;
  for (int64_t u = 0; u < n; u++)
    for (uint32_t j = g->rowptr[u]; j < g->rowptr[u + 1]; j++) {
      const int64_t v = g->adj[j];
      const float l = (float)length(std::floor(vx[u]), std::floor(vy[u]), std::floor(vx[v]), std::floor(vy[v]));
      if (l > max_len) max_len = l;
    }
  Painter pt{img, width, (long)(g->m / 2)};
  for (int64_t u = 0; u < n; u++)
    for (uint32_t j = g->rowptr[u]; j < g->rowptr[u + 1]; j++) {
      const int64_t v = g->adj[j];
      const float len = (float)length(std::floor(vx[u]), std::floor(vy[u]), std::floor(vx[v]), std::floor(vy[v]));
      const float t = (len / max_len) * 7;
      const int bin = (int)std::floor(t);
      pt.line((int)std::floor(vx[u]), (int)std::floor(vy[u]), (int)std::floor(vx[v]), (int)std::floor(vy[v]),
              pt.colour(bin));
    }
  const int rc = write_png(path, img, width, height);
  if (rc == PG_OK) {
    if (width_out) *width_out = width;
    if (height_out) *height_out = height;
  }
  return rc;
}

}  // extern "C"
