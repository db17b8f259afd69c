// spmm.h — internal interface of the Laplacian / Gram / coordinates stage.
#pragma once
#include "phde_common.h"

namespace phde {

// Row-major work copies of S and L*S plus the s x s product.
struct Lap {
  phde_ctx *ctx = nullptr;
  int s = 0;
  bool fp32 = false;
  int nblk_gram = 0;
  void *Srow = nullptr, *LSrow = nullptr;
  double *part = nullptr, *SLS = nullptr;  // SLS: s*s device, reference storage order
  double *hub_part = nullptr;                // SpMM partials of the hub rows' chunks
  int setup(phde_ctx *c, int64_t n, int s, bool fp32);
  void release();
  ~Lap() { release(); }
  int run(const void *Scol);        // SLS = S^T (L S)
  int gram_self(const void *Acol);  // SLS = A^T A
};

// The SpMM's rows longer than one chunk (ctx sp_* fields), once per context.
int spmm_hubs(phde_ctx *c);
int coords_compute(phde_ctx *c, const void *Scol, bool fp32, int s, int nc, const double *V_dev,
                   double *X_dev, double *part, double *inv);
int coords_normalise(phde_ctx *c, double *X_dev, int nc, const double *part, int nblk, double *inv);
int hall_energy_dev(phde_ctx *c, const double *Y_dev, int nc, double *part_dev, double *h_out2);

}  // namespace phde
