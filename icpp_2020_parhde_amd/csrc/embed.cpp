// embed.cpp — the `embed` command, drop-in for spectralDrawing.cpp:1-202.
//
//   embed <csr> [-w x] [-c n] [-a algo] [-p strat] [-r s] [-h]
//         [-f fp64|fp32] [-e power|jacobi|generalized] [-E]
//
// Same argument parsing (getopt "w:hc:a:p:r:" plus the GPU options), same
// binary CSR input (spectralDrawing.cpp:136-164), same `<csr>_hde.nxyz`
// output written with default ostream formatting (:41-65), same progress
// lines.  The layout runs on the GPU through the DrawingBase classes.
#include <getopt.h>
#include <unistd.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>

#include "PHDE.h"
#include "ParHDE.h"
#include "PivotMDS.h"
#include "WParHDE.h"

static int writeCoords(long n, double *coordinates, int numCoord, const char *inputFilename) {
  std::ofstream fout;
  std::string coordFilename(inputFilename);
  coordFilename += "_hde.nxyz";
  std::cout << "Writing coordinates to file " << coordFilename << std::endl;
  fout.open(coordFilename);
  for (long i = 0; i < n; i++) {
    for (int j = 0; j < numCoord - 1; j++) fout << coordinates[i + j * n] << ",";
    fout << coordinates[i + (numCoord - 1) * n] << std::endl;
  }
  fout.close();
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::cerr << "Usage: " << argv[0] << " [csr filename] [options]" << std::endl;
    return 1;
  }
  char *inputFilename = argv[1];
  int weighted = 0, numCoord = 2, doRand = 2, drawingAlgorithm = 3, numSources = 10;
  int precision = PHDE_PREC_FP64, eig = PHDE_EIG_POWER, eager = 0;
  int opt;
  int judge = 0;
  const char *netPath = nullptr;
  while ((opt = getopt(argc, argv, "w:hc:a:p:r:f:e:EJn:")) != -1) {
    switch (opt) {
      case 'w':
        weighted = 1;
        break;
      case 'c':
        numCoord = atoi(optarg);
        break;
      case 'p':
        if (!strcmp(optarg, "RandFine"))
          doRand = 1;
        else if (!strcmp(optarg, "RandCoarse"))
          doRand = 0;
        else if (!strcmp(optarg, "Kcenters"))
          doRand = 2;
        else if (!strcmp(optarg, "Learned"))
          doRand = 3;
        else {
          std::cerr << "Unrecognized Strategy" << std::endl;
          abort();
        }
        break;
      case 'a':
        if (!strcmp(optarg, "PivotMDS"))
          drawingAlgorithm = 0;
        else if (!strcmp(optarg, "PHDE"))
          drawingAlgorithm = 1;
        else if (!strcmp(optarg, "wParHDE"))
          drawingAlgorithm = 2;
        else if (!strcmp(optarg, "ParHDE"))
          drawingAlgorithm = 3;
        else {
          std::cerr << "Unrecognized algorithm" << std::endl;
          abort();
        }
        break;
      case 'r':
        numSources = atoi(optarg);
        break;
      case 'f':
        precision = !strcmp(optarg, "fp32") ? PHDE_PREC_FP32 : PHDE_PREC_FP64;
        break;
      case 'e':
        eig = !strcmp(optarg, "jacobi") ? PHDE_EIG_JACOBI
                                        : (!strcmp(optarg, "generalized") ? PHDE_EIG_GENERALIZED
                                                                          : PHDE_EIG_POWER);
        break;
      case 'E':
        eager = 1;
        break;
      case 'J':
        judge = 1;
        break;
      case 'n':
        netPath = optarg;
        break;
      case 'h':
        std::cout << "Usage: " << argv[0]
                  << " [csr filename]\n"
                     "-w         : Indicates if CSR file is weighted or unweighted (default)\n"
                     "-c <value> : number of coordinates to generate( default=2)\n"
                     "-a <algo>  : Indicate the algorithm to run (PivotMDS,PHDE, wParHDE, ParHDE(default))\n"
                     "-p <strat> : Pivot picking strategy to use (RandFine,RandCoarse,Kcenters(default),Learned)\n"
                     "-r <value> : Number of pivots (default=10)"
                     "-f <prec>  : fp64 (default) or fp32 orthogonalisation / SpMM\n"
                     "-e <eig>   : power (default), jacobi or generalized eigensolver\n"
                     "-E         : eager traversals (3s, as the reference)\n"
                     "-n <file>  : pivot network weights for -p Learned (default: weights/pivot_net.bin\n"
                     "             next to the library)\n"
                     "-J         : with -p Learned, also run k-centers and keep it when its Hall\n"
                     "             energy is more than 2% lower\n"
                     "-h         : help"
                  << std::endl;
        exit(1);
      default:
        abort();
    }
  }
  auto startTimer = std::chrono::high_resolution_clock::now();
  FILE *infp = fopen(inputFilename, "rb");
  if (infp == NULL) {
    std::cout << "Error: Could not open input file. Exiting ..." << std::endl;
    return 1;
  }
  long n, m, rest[4];
  if (fread(&n, sizeof(long), 1, infp) != 1 || fread(&m, sizeof(long), 1, infp) != 1 ||
      fread(rest, sizeof(long), 4, infp) != 4 || n < 1 || m < 0) {
    std::cout << "Error: malformed CSR header" << std::endl;
    return 1;
  }
  unsigned int *rowOffsets = (unsigned int *)malloc(sizeof(unsigned int) * (n + 1));
  unsigned int *adj = (unsigned int *)malloc(sizeof(unsigned int) * (m > 0 ? m : 1));
  int32_t *weights = NULL;
  bool ok = rowOffsets && adj && fread(rowOffsets, sizeof(unsigned int), n + 1, infp) == (size_t)(n + 1) &&
            fread(adj, sizeof(unsigned int), m, infp) == (size_t)m;
  if (ok && weighted) {
    weights = (int32_t *)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    ok = weights && fread(weights, sizeof(int32_t), m, infp) == (size_t)m;
  }
  fclose(infp);
  if (!ok) {
    std::cout << "Error: truncated CSR file" << std::endl;
    return 1;
  }
  std::chrono::duration<double> elt = std::chrono::high_resolution_clock::now() - startTimer;
  std::cout << "CSR read time: " << elt.count() << " s." << std::endl;
  std::cout << "Num edges: " << m / 2 << ", vertices: " << n << ", nnz: " << m << std::endl;

  DrawingBase *draw;
  switch (drawingAlgorithm) {
    case 0:
      draw = new PivotMDS(rowOffsets, adj, NULL, n, m, numSources, doRand, numCoord);
      break;
    case 1:
      draw = new PHDE(rowOffsets, adj, NULL, n, m, numSources, doRand, numCoord);
      break;
    case 2:
      draw = new WParHDE(rowOffsets, adj, weights, n, m, numSources, numCoord);
      break;
    default:
      draw = new ParHDE(rowOffsets, adj, NULL, n, m, numSources, doRand, numCoord);
  }
  draw->options().precision = precision;
  draw->options().eig_mode = eig;
  draw->options().eager = eager;
  draw->options().verbose = 1;
  phde_mpnn *net = nullptr;
  if (doRand == 3) {
    std::string path;
    if (netPath) {
      path = netPath;
    } else {
      // bin/embed -> ../weights/pivot_net.bin
      char exe[4096];
      const ssize_t len = readlink("/proc/self/exe", exe, sizeof(exe) - 1);
      exe[len > 0 ? len : 0] = 0;
      path = std::string(exe);
      path = path.substr(0, path.find_last_of('/')) + "/../weights/pivot_net.bin";
    }
    int rc = phde_net_create(0, 1, &net);
    if (rc == PHDE_OK) rc = phde_net_load(net, path.c_str());
    if (rc != PHDE_OK) {
      std::cerr << "Cannot load the pivot network (" << path << "): " << phde_last_error() << std::endl;
      return 1;
    }
    draw->setPivotNet(net);
    draw->options().judge = judge;
  }
  draw->run(1);
  double *coordinates = draw->getCoordinates();
  writeCoords(n, coordinates, numCoord, inputFilename);
  free(rowOffsets);
  free(adj);
  if (weighted) free(weights);
  elt = std::chrono::high_resolution_clock::now() - startTimer;
  std::cout << "Overall time: " << elt.count() << " s." << std::endl;
  if (net) {
    // the context that borrows the network goes first; deleted through its
    // own class, which is the object's dynamic type (the destructor is
    // non-virtual, as in the reference)
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wdelete-non-virtual-dtor"
    switch (drawingAlgorithm) {
      case 0: delete static_cast<PivotMDS *>(draw); break;
      case 1: delete static_cast<PHDE *>(draw); break;
      case 2: delete static_cast<WParHDE *>(draw); break;
      default: delete static_cast<ParHDE *>(draw);
    }
#pragma GCC diagnostic pop
    phde_net_destroy(net);
  }
  return 0;
}
