// draw — PNG picture of a layout's edges; command-line drop-in for the
// reference's draw_graph.c (same six arguments: binary CSR, .nxyz
// coordinates, output PNG, number of coordinates per line, and the 1-based
// coordinates for x and y).  The rendering is pg_draw_png.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "phde_graph.h"

int main(int argc, char **argv) {
  if (argc != 7) {
    std::fprintf(stderr,
                 "Usage: %s <binary csr filename> <vertex coords filename> <output png filename> "
                 "<num coordinates> <first coord> <second coord>\n",
                 argv[0]);
    return 1;
  }
  const int nc = std::atoi(argv[4]), first = std::atoi(argv[5]) - 1, second = std::atoi(argv[6]) - 1;
  pg_csr g{};
  if (pg_read_csr(argv[1], 0, &g)) {
    std::fprintf(stderr, "Error: Could not open CSR file. Exiting ...\n");
    return 1;
  }
  std::fprintf(stderr, "Num edges: %lld, num vertices: %lld\n", (long long)g.m / 2, (long long)g.n);
  FILE *f = std::fopen(argv[2], "r");
  if (!f || nc < 1) {
    std::fprintf(stderr, "Error: Could not open vertex coords file. Exiting ...\n");
    pg_free(&g);
    return 1;
  }
  // .nxyz: one vertex per line, nc comma-separated values (column-major here)
  std::vector<double> xy((size_t)nc * g.n);
  for (int64_t i = 0; i < g.n; i++)
    for (int j = 0; j < nc; j++) {
      double x = 0;
      if (std::fscanf(f, j + 1 < nc ? "%lf," : "%lf", &x) != 1) {
        std::fprintf(stderr, "Error: vertex coords file ends early (vertex %lld). Exiting ...\n", (long long)i);
        std::fclose(f);
        pg_free(&g);
        return 1;
      }
      xy[(size_t)j * g.n + i] = x;
    }
  std::fclose(f);
  unsigned w = 0, h = 0;
  const int rc = pg_draw_png(&g, xy.data(), nc, first, second, argv[3], &w, &h);
  pg_free(&g);
  if (rc) {
    std::fprintf(stderr, "error %d writing %s\n", rc, argv[3]);
    return 1;
  }
  std::fprintf(stderr, "Image %u x %u written to %s\n", w, h, argv[3]);
  return 0;
}
