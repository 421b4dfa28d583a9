// decomp.h -- launchers of the decomposition epilogue kernels (decomp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace mmpbsa {

void launch_gb_weight(const int *sel, int n, const double *qa, const double *Rc,
                      const double *Ri, double f, double *w, hipStream_t s);
void launch_gb_share(const double *w, int n, const double *scal, double *out, hipStream_t s);
void launch_sa_share(const int *sel, int n, const double *sc, const double *si, double gamma,
                     double *out, hipStream_t s);
void launch_dgb_scalar(const double *e3, const double *tw, double *scal, hipStream_t s);

}  // namespace mmpbsa
