// decomp.hip -- per-atom polar / SA shares of the per-residue decomposition
// (decomposition.rs:302-375).  Residue sums use launch_segment_sum.
#include <hip/hip_runtime.h>

#include "decomp.h"

namespace mmpbsa {

// w_k = | f q^2 / R_complex - f q^2 / R_isolated |, f = -1/2 (1/eps_in - 1/eps_out)
// (decomposition.rs:311-337; GB self energy without salt, as the reference).
__global__ void gb_weight_kernel(const int *__restrict__ sel, int n, const double *__restrict__ qa,
                                 const double *__restrict__ Rc, const double *__restrict__ Ri,
                                 double f, double *__restrict__ w)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int g = sel[k];
    const double q = qa[g];
    const double sc = f * q * q / Rc[g];
    const double si = f * q * q / Ri[k];
    w[k] = fabs(sc - si);
}

// polar share: dGB * (w / sum w) when sum w > 1e-20 (decomposition.rs:340-352)
__global__ void gb_share_kernel(const double *__restrict__ w, int n,
                                const double *__restrict__ scal /* [dgb, tw] */,
                                double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double dgb = scal[0], tw = scal[1];
    out[k] = tw > 1e-20 ? dgb * (w[k] / tw) : 0.0;
}

// gamma * (SASA_complex[sel[k]] - SASA_isolated[k]) (decomposition.rs:365-375)
__global__ void sa_share_kernel(const int *__restrict__ sel, int n, const double *__restrict__ sc,
                                const double *__restrict__ si, double gamma,
                                double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = gamma * (sc[sel[k]] - si[k]);
}

__global__ void dgb_scalar_kernel(const double *__restrict__ e3, const double *__restrict__ tw,
                                  double *__restrict__ scal)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        scal[0] = e3[0] - e3[1] - e3[2];
        scal[1] = tw[0];
    }
}

void launch_gb_weight(const int *sel, int n, const double *qa, const double *Rc,
                      const double *Ri, double f, double *w, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(gb_weight_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, sel, n, qa, Rc,
                       Ri, f, w);
    LAUNCH_OK();
}

void launch_gb_share(const double *w, int n, const double *scal, double *out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(gb_share_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, w, n, scal, out);
    LAUNCH_OK();
}

void launch_sa_share(const int *sel, int n, const double *sc, const double *si, double gamma,
                     double *out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(sa_share_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, sel, n, sc, si,
                       gamma, out);
    LAUNCH_OK();
}

void launch_dgb_scalar(const double *e3, const double *tw, double *scal, hipStream_t s)
{
    hipLaunchKernelGGL(dgb_scalar_kernel, dim3(1), dim3(64), 0, s, e3, tw, scal);
    LAUNCH_OK();
}

}  // namespace mmpbsa
