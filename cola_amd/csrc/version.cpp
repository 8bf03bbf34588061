// Version and status strings of the C ABI (host-only translation unit).
#include "../../include/cola_krylov.h"

extern "C" const char* ck_version(void) { return "cola_krylov 0.1.0 gfx950"; }

extern "C" const char* ck_status_string(int status) {
  switch (status) {
    case CK_OK:
      return "ok";
    case CK_ERR_BAD_ARG:
      return "bad argument (shape, size or null pointer)";
    case CK_ERR_HIP:
      return "HIP runtime or kernel launch error";
    case CK_ERR_UNSUPPORTED:
      return "operator structure not supported by this entry point";
    case CK_ERR_WORKSPACE:
      return "workspace too small";
    case CK_ERR_NOCONV:
      return "tridiagonal QL iteration did not converge";
    case CK_ERR_MAXITER:
      return "iteration cap reached before the tolerance";
    case CK_ERR_BREAKDOWN:
      return "breakdown: p.Ap <= 0 (operator or preconditioner not positive definite)";
    case CK_ERR_NAN:
      return "non-finite residual";
    case CK_ERR_TIMEOUT:
      return "grid-wide wait timed out";
    default:
      return "unknown status";
  }
}
