// mtx2csr — Matrix Market coordinate file to the binary CSR that `embed`
// reads; command-line drop-in for the reference's mtx2csr.cpp (same two
// arguments, same output file).  The conversion itself is pg_mtx2csr.
#include <cstdio>

#include "phde_graph.h"

int main(int argc, char **argv) {
  if (argc != 3) {
    std::printf("Usage: %s <MM filename> <binary CSR filename>\n", argv[0]);
    return 1;
  }
  pg_csr full{};
  int kind = 0;
  int rc = pg_read_mtx(argv[1], &full, &kind);
  if (rc) {
    std::printf("Error: could not read Matrix Market coordinate file %s (code %d)! Exiting\n", argv[1], rc);
    return 1;
  }
  if (kind == PG_MTX_REAL) std::printf("Matrix has real weights, ignoring them.\n");
  if (kind == PG_MTX_INTEGER) std::printf("Matrix has integer weights, using them\n");
  std::printf("After deduplication and self loop removal, n: %lld, m: %lld\n", (long long)full.n,
              (long long)full.m / 2);
  pg_csr lcc{};
  int64_t ncomp = 0;
  rc = pg_lcc(&full, &lcc, &ncomp);
  pg_free(&full);
  if (rc) {
    std::printf("Error: component extraction failed (code %d)\n", rc);
    return 1;
  }
  if (ncomp == 1)
    std::printf("There is just one connected component.\n");
  else
    std::printf("There are %lld components and the largest component size is %lld\n", (long long)ncomp,
                (long long)lcc.n);
  std::printf("Writing binary file to disk.\n");
  rc = pg_write_csr(argv[2], &lcc);
  pg_free(&lcc);
  if (rc) {
    std::printf("Error: could not write output CSR file %s! Exiting\n", argv[2]);
    return 1;
  }
  return 0;
}
