// Quasi-harmonic entropy of a subsystem over frames (entropy.rs:97-217):
// the mass-weighted covariance of the atomic fluctuations, its eigenvalues,
// and the quantum harmonic oscillator entropy of each mode.
//
// On the device: the mass-weighted means (one thread per coordinate, frames
// in order, as the reference sums them), the covariance as a rank-F update
// tiled through LDS (every element accumulates d_i d_j over the frames in
// frame order and is divided by F, so the matrix equals the reference's bit
// for bit), the eigenvalues by rocSOLVER's divide-and-conquer symmetric
// eigensolver (the reference uses cyclic Jacobi to 1e-10: the two agree to
// rounding of the significant modes), and the mode entropies summed in
// descending eigenvalue order by one lane.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <shared_mutex>

#include "engine.h"
#include "kernels.h"

namespace mmpbsa {

namespace {

constexpr int kQhTile = 64;   // covariance tile edge
constexpr int kQhChunk = 16;  // frames staged in LDS per step

// mean[k] = (sum over frames, in order, of x[f][k] sqrt(m)) / F
__global__ __launch_bounds__(256) void qh_mean_kernel(const double *__restrict__ x,
                                                      const double *__restrict__ sqm, int nf,
                                                      int dim, double *__restrict__ mean)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dim) return;
    const double w = sqm[k / 3];
    double s = 0.0;
    for (int f = 0; f < nf; ++f) s += x[(int64_t)f * dim + k] * w;
    mean[k] = s / (double)nf;
}

// d[f][k] = x[f][k] sqrt(m) - mean[k]
__global__ __launch_bounds__(256) void qh_delta_kernel(const double *__restrict__ x,
                                                       const double *__restrict__ sqm,
                                                       const double *__restrict__ mean, int nf,
                                                       int dim, double *__restrict__ d)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nf * dim) return;
    const int k = (int)(i % dim);
    d[i] = x[i] * sqm[k / 3] - mean[k];
}

// cov (column-major, lda = dim, both triangles) = sum over frames of d_i d_j
// in frame order, / F.  Workgroup: a 64 x 64 tile of the upper block triangle
// (tile column >= tile row), 256 threads of 4 x 4 elements; frames staged 16
// at a time.  The mirrored element is the same product sum (multiplication
// commutes), so both triangles are written from it.
__global__ __launch_bounds__(256) void qh_cov_kernel(const double *__restrict__ d, int nf, int dim,
                                                     int ntile, double *__restrict__ cov)
{
    __shared__ double di[kQhChunk][kQhTile], dj[kQhChunk][kQhTile];
    // blockIdx.x -> (tile row ti <= tile column tj) in row-major order of the
    // upper block triangle
    int b = blockIdx.x, ti = 0;
    while (b >= ntile - ti) {
        b -= ntile - ti;
        ++ti;
    }
    const int tj = ti + b;
    const int i0 = ti * kQhTile, j0 = tj * kQhTile;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;
    for (int f0 = 0; f0 < nf; f0 += kQhChunk) {
        for (int t = threadIdx.x; t < kQhChunk * kQhTile; t += 256) {
            const int r = t / kQhTile, c = t - r * kQhTile;
            const int f = f0 + r;
            const bool fin = f < nf;
            di[r][c] = fin && i0 + c < dim ? d[(int64_t)f * dim + i0 + c] : 0.0;
            dj[r][c] = fin && j0 + c < dim ? d[(int64_t)f * dim + j0 + c] : 0.0;
        }
        __syncthreads();
        const int nr = min(kQhChunk, nf - f0);
        for (int r = 0; r < nr; ++r) {
            double a[4], c[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = di[r][ty + 16 * q];
                c[q] = dj[r][tx + 16 * q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * c[q];
        }
        __syncthreads();
    }
    const double fn = (double)nf;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = i0 + ty + 16 * p, j = j0 + tx + 16 * q;
            if (i < dim && j < dim && i <= j) {
                const double v = acc[p][q] / fn;
                cov[(int64_t)j * dim + i] = v;
                cov[(int64_t)i * dim + j] = v;
            }
        }
}

// S/kB of each mode (entropy.rs:199-212): eigenvalue lambda in A^2 amu;
// modes below 1e-10 or with alpha above 500 contribute nothing
__global__ __launch_bounds__(256) void qh_modes_kernel(const double *__restrict__ lam, int dim,
                                                       double kbt_si, double conv,
                                                       double *__restrict__ s)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dim) return;
    const double hbar = 1.054571817e-34;
    const double l = lam[k];
    double v = 0.0;
    if (!(l < 1e-10)) {
        const double alpha = hbar / sqrt(kbt_si * (l * conv));
        if (!(alpha > 500.0)) {
            const double ea = exp(alpha);
            v = alpha / (ea - 1.0) - log(1.0 - exp(-alpha));
        }
    }
    s[k] = v;
}

// sum in descending eigenvalue order (the solver returns ascending order)
__global__ void qh_sum_kernel(const double *__restrict__ s, int dim, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double t = 0.0;
    for (int k = dim - 1; k >= 0; --k) t += s[k];
    *out = t;
}

void rb_ok(rocblas_status st, const char *what)
{
    if (st != rocblas_status_success)
        throw DeviceError(MMPBSA_ERR_DEVICE, std::string(what) + ": rocBLAS status " +
                                                 std::to_string((int)st));
}

}  // namespace

int64_t qh_max_dim() { return 60000; }

double qh_entropy(Worker &w, const double *h_xyz, int nf, int na, const double *h_masses,
                  double temperature)
{
    const int64_t dim64 = 3 * (int64_t)na;
    if (dim64 > qh_max_dim())
        throw DeviceError(MMPBSA_ERR_TOO_LARGE,
                          "quasi-harmonic entropy: more than 20,000 atoms (a 3N x 3N covariance)");
    const int dim = (int)dim64;
    // a negative or non-finite mass, or a non-finite coordinate, makes the
    // reference's covariance NaN, and with it its eigenvalues and result
    for (int i = 0; i < na; ++i)
        if (!(h_masses[i] >= 0.0) || !std::isfinite(h_masses[i])) return std::nan("");
    for (int64_t i = 0; i < (int64_t)nf * dim; ++i)
        if (!std::isfinite(h_xyz[i])) return std::nan("");
    std::vector<double> sqm(na);
    for (int i = 0; i < na; ++i) sqm[i] = std::sqrt(h_masses[i]);
    const size_t nx = (size_t)nf * dim;
    DBuf<double> x(w.pool, nx), dm(w.pool, na), mean(w.pool, dim), d(w.pool, nx);
    DBuf<double> cov(w.pool, (size_t)dim * dim), lam(w.pool, dim), e(w.pool, dim),
        s(w.pool, dim), tot(w.pool, 1);
    DBuf<int> info(w.pool, 1);
    HIP_OK(hipMemcpyAsync(x.get(), h_xyz, sizeof(double) * nx, hipMemcpyHostToDevice, w.stream));
    HIP_OK(hipMemcpyAsync(dm.get(), sqm.data(), sizeof(double) * na, hipMemcpyHostToDevice,
                          w.stream));
    hipLaunchKernelGGL(qh_mean_kernel, dim3(ceil_div(dim, 256)), dim3(256), 0, w.stream, x.get(),
                       dm.get(), nf, dim, mean.get());
    LAUNCH_OK();
    hipLaunchKernelGGL(qh_delta_kernel, dim3(ceil_div((int64_t)nx, 256)), dim3(256), 0, w.stream,
                       x.get(), dm.get(), mean.get(), nf, dim, d.get());
    LAUNCH_OK();
    const int ntile = ceil_div(dim, kQhTile);
    const int64_t nblk = (int64_t)ntile * (ntile + 1) / 2;
    hipLaunchKernelGGL(qh_cov_kernel, dim3((unsigned)nblk), dim3(256), 0, w.stream, d.get(), nf,
                       dim, ntile, cov.get());
    LAUNCH_OK();
    {
        // the solver allocates its workspace: no graph capture may be open
        std::unique_lock<std::shared_mutex> g(capture_gate());
        rocblas_handle h = nullptr;
        rb_ok(rocblas_create_handle(&h), "rocblas_create_handle");
        struct Destroy {
            rocblas_handle h;
            ~Destroy() { (void)rocblas_destroy_handle(h); }
        } destroy{h};
        rb_ok(rocblas_set_stream(h, w.stream), "rocblas_set_stream");
        rb_ok(rocsolver_dsyevd(h, rocblas_evect_none, rocblas_fill_upper, dim, cov.get(), dim,
                               lam.get(), e.get(), info.get()),
              "rocsolver_dsyevd");
        // the handle's workspace is released only after the solver has run
        HIP_OK(hipStreamSynchronize(w.stream));
    }
    const double kbt_si = 1.380649e-23 * temperature;
    const double conv = 1e-20 * 1.66054e-27;  // A^2 amu -> m^2 kg
    hipLaunchKernelGGL(qh_modes_kernel, dim3(ceil_div(dim, 256)), dim3(256), 0, w.stream,
                       lam.get(), dim, kbt_si, conv, s.get());
    LAUNCH_OK();
    hipLaunchKernelGGL(qh_sum_kernel, dim3(1), dim3(64), 0, w.stream, s.get(), dim, tot.get());
    LAUNCH_OK();
    HIP_OK(hipMemcpyAsync(w.pin, tot.get(), sizeof(double), hipMemcpyDeviceToHost, w.stream));
    HIP_OK(hipMemcpyAsync(w.pin_i, info.get(), sizeof(int), hipMemcpyDeviceToHost, w.stream));
    w.sync();
    if (w.pin_i[0] != 0)
        throw DeviceError(MMPBSA_ERR_DEVICE, "quasi-harmonic entropy: eigensolver did not converge");
    // -T dS = -T S/kB R, R = 0.00198688 kcal/(mol K)
    return -temperature * 0.00198688 * w.pin[0];
}

}  // namespace mmpbsa
