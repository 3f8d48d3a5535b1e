// Column-wise PCG passes (blas.hip) used by the resident state's global solve.
#pragma once
#include "common.h"

namespace dlp {

struct PcgBufs {
  int n, C;
  int64_t ld;
  float* Z;
  float* R;
  float* P;
  float* AP;
  const float* diag;
  const float* eps;
  double* rz;
  double* rz_new;
  double* pAp;
  double* alpha;
  double* beta;
  float* viol;
};

// P = R/diag, rz = sum R*R/diag, pAp = 0.
int pcg_start(const PcgBufs& b, hipStream_t st);
// After AP = A P has been formed: pAp, alpha, Z/R update (rz_new, viol), beta, new P.
int pcg_finish_iteration(const PcgBufs& b, hipStream_t st);

}  // namespace dlp
