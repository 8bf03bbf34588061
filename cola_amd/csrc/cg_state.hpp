// Layout of the CG state buffer (ck_cg_ctrl followed by per-column fp32 arrays), shared by
// the chained CG kernels (cg.hip) and the persistent fused solve (fused_cg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/cola_krylov.h"

namespace ck {

struct CgView {
  ck_cg_ctrl* ctrl;
  float* mult;
  float* tol;
  float* gamma;
  float* alpha;
  float* beta;
  float* rnorm;
  float* errors;
};

// floats per per-column array: nrhs rounded up to 64
__host__ __device__ inline int cg_stride_of(int nrhs) { return (nrhs + 63) / 64 * 64; }

__host__ __device__ inline CgView cg_view(void* state, int nrhs) {
  CgView v;
  v.ctrl = static_cast<ck_cg_ctrl*>(state);
  float* f = reinterpret_cast<float*>(static_cast<char*>(state) + sizeof(ck_cg_ctrl));
  const int s = cg_stride_of(nrhs);
  v.mult = f + CK_CG_MULT * s;
  v.tol = f + CK_CG_TOL * s;
  v.gamma = f + CK_CG_GAMMA * s;
  v.alpha = f + CK_CG_ALPHA * s;
  v.beta = f + CK_CG_BETA * s;
  v.rnorm = f + CK_CG_RNORM * s;
  v.errors = f + CK_CG_ERRORS * s;
  return v;
}

}  // namespace ck
