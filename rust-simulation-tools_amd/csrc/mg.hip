// mg.hip -- linearized-PB multigrid on gfx950.
//
// The reference V-cycle (pb_solver.rs:842-939) is affine in (phi, rhs): one
// cycle maps phi to phi + M (b - A phi).  It is run here in correction form:
//   r_k  = b - A phi_k             FP64, fused with phi_k = phi_{k-1} + e_{k-1}
//   e_k  = V(0, r_k)               FP32, same smoother, omega per level, 0.25
//                                  restriction scale and prolongation onto every
//                                  node (even-sized levels extrapolate, as the
//                                  reference does) -- so the iterate, its
//                                  boundary drift and the cycle count follow the
//                                  FP64 reference to FP32 precision of e_k
//   convergence: RMS of r_k (FP64), the reference's test at the same cycle.
// A device status word stops every later kernel once the solve has converged,
// diverged or run out of cycles, so the host can enqueue cycles in batches.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "kernels.h"
#include "mg.h"

namespace mmpbsa {

__device__ __forceinline__ bool stopped(const MgStatus *st) { return st->state != 0; }

// ---- FP32 red-black half-sweep (sor_smooth, pb_solver.rs:214-327) -----------
__global__ __launch_bounds__(256) void rb_sweep_kernel(MgLevelDev L, int color, float omega,
                                                       const MgStatus *__restrict__ st)
{
    if (stopped(st)) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int half = (nx - 1) / 2;  // slots per row
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t rows = (int64_t)(ny - 2) * (nz - 2);
    if (t >= rows * half) return;
    const int s = (int)(t % half);
    const int64_t row = t / half;
    const int iy = 1 + (int)(row % (ny - 2));
    const int iz = 1 + (int)(row / (ny - 2));
    const int ix = 1 + ((1 + iy + iz + color) & 1) + 2 * s;
    if (ix > nx - 2) return;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int64_t idx = ix + sy * iy + sz * iz;
    float *__restrict__ e = L.e;
    float nb, dg;
    if (L.uni) {
        const float u = L.eps_u;
        nb = u * ((e[idx + 1] + e[idx - 1]) * L.ihx + (e[idx + sy] + e[idx - sy]) * L.ihy +
                  (e[idx + sz] + e[idx - sz]) * L.ihz);
        dg = u * (2.0f * L.ihx + 2.0f * L.ihy + 2.0f * L.ihz);
    } else {
        const float exp_ = L.ex[idx], exm = L.ex[idx - 1];
        const float eyp = L.ey[idx], eym = L.ey[idx - sy];
        const float ezp = L.ez[idx], ezm = L.ez[idx - sz];
        nb = fmaf(exp_, e[idx + 1], exm * e[idx - 1]) * L.ihx +
             fmaf(eyp, e[idx + sy], eym * e[idx - sy]) * L.ihy +
             fmaf(ezp, e[idx + sz], ezm * e[idx - sz]) * L.ihz;
        dg = (exp_ + exm) * L.ihx + (eyp + eym) * L.ihy + (ezp + ezm) * L.ihz + L.k2[idx];
    }
    const float old = e[idx];
    const float gs = (nb + L.f[idx]) / dg;
    e[idx] = fmaf(omega, gs - old, old);
}

// residual res = f + nb - dg * e on the interior, 0 on the boundary
__global__ __launch_bounds__(256) void residual32_kernel(MgLevelDev L,
                                                         const MgStatus *__restrict__ st)
{
    if (stopped(st)) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int iz = (int)(idx / sz);
    const int iy = (int)((idx - (int64_t)iz * sz) / nx);
    const int ix = (int)(idx - (int64_t)iz * sz - (int64_t)iy * nx);
    if (ix == 0 || iy == 0 || iz == 0 || ix == nx - 1 || iy == ny - 1 || iz == nz - 1) {
        L.res[idx] = 0.0f;
        return;
    }
    const float *__restrict__ e = L.e;
    float nb, dg;
    if (L.uni) {
        const float u = L.eps_u;
        nb = u * ((e[idx + 1] + e[idx - 1]) * L.ihx + (e[idx + sy] + e[idx - sy]) * L.ihy +
                  (e[idx + sz] + e[idx - sz]) * L.ihz);
        dg = u * (2.0f * L.ihx + 2.0f * L.ihy + 2.0f * L.ihz);
    } else {
        const float exp_ = L.ex[idx], exm = L.ex[idx - 1];
        const float eyp = L.ey[idx], eym = L.ey[idx - sy];
        const float ezp = L.ez[idx], ezm = L.ez[idx - sz];
        nb = fmaf(exp_, e[idx + 1], exm * e[idx - 1]) * L.ihx +
             fmaf(eyp, e[idx + sy], eym * e[idx - sy]) * L.ihy +
             fmaf(ezp, e[idx + sz], ezm * e[idx - sz]) * L.ihz;
        dg = (exp_ + exm) * L.ihx + (eyp + eym) * L.ihy + (ezp + ezm) * L.ihz + L.k2[idx];
    }
    L.res[idx] = L.f[idx] + nb - dg * e[idx];
}

// restrict_into (pb_solver.rs:578-624): normalised full weighting, x scale.
// Also clears the coarse correction (the reference zeroes it before recursing).
__global__ __launch_bounds__(256) void restrict_kernel(MgLevelDev F, MgLevelDev C, float scale,
                                                       const MgStatus *__restrict__ st)
{
    if (stopped(st)) return;
    const int64_t N = (int64_t)C.nx * C.ny * C.nz;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int64_t csz = (int64_t)C.nx * C.ny;
    const int cz = (int)(idx / csz);
    const int cy = (int)((idx - (int64_t)cz * csz) / C.nx);
    const int cx = (int)(idx - (int64_t)cz * csz - (int64_t)cy * C.nx);
    float sum = 0.0f, ws = 0.0f;
    for (int dz = -1; dz <= 1; ++dz) {
        const int z = 2 * cz + dz;
        if (z < 0 || z >= F.nz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int y = 2 * cy + dy;
            if (y < 0 || y >= F.ny) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = 2 * cx + dx;
                if (x < 0 || x >= F.nx) continue;
                const float w = (dx == 0 ? 2.0f : 1.0f) * (dy == 0 ? 2.0f : 1.0f) *
                                (dz == 0 ? 2.0f : 1.0f);
                sum = fmaf(w, F.res[x + (int64_t)F.nx * (y + (int64_t)F.ny * z)], sum);
                ws += w;
            }
        }
    }
    C.f[idx] = scale * sum / ws;
    C.e[idx] = 0.0f;
}

// prolongate_add (pb_solver.rs:628-681): trilinear with the coarse cell
// clamped to n_c - 2, added to every fine node.
__global__ __launch_bounds__(256) void prolong_kernel(MgLevelDev C, MgLevelDev F,
                                                      const MgStatus *__restrict__ st)
{
    if (stopped(st)) return;
    const int64_t N = (int64_t)F.nx * F.ny * F.nz;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int64_t fsz = (int64_t)F.nx * F.ny;
    const int fz = (int)(idx / fsz);
    const int fy = (int)((idx - (int64_t)fz * fsz) / F.nx);
    const int fx = (int)(idx - (int64_t)fz * fsz - (int64_t)fy * F.nx);
    const int x0 = min(fx >> 1, max(C.nx - 2, 0));
    const int y0 = min(fy >> 1, max(C.ny - 2, 0));
    const int z0 = min(fz >> 1, max(C.nz - 2, 0));
    const float dx = 0.5f * fx - x0, dy = 0.5f * fy - y0, dz = 0.5f * fz - z0;
    const int64_t cs = C.nx, csz = (int64_t)C.nx * C.ny;
    const float *__restrict__ c = C.e + x0 + cs * y0 + csz * z0;
    const float c00 = c[0] * (1.0f - dx) + c[1] * dx;
    const float c10 = c[cs] * (1.0f - dx) + c[cs + 1] * dx;
    const float c01 = c[csz] * (1.0f - dx) + c[csz + 1] * dx;
    const float c11 = c[csz + cs] * (1.0f - dx) + c[csz + cs + 1] * dx;
    const float c0 = c00 * (1.0f - dy) + c10 * dy;
    const float c1 = c01 * (1.0f - dy) + c11 * dy;
    F.e[idx] += c0 * (1.0f - dz) + c1 * dz;
}

// coarsen_dielectrics / coarsen_kappa by injection (pb_solver.rs:684-748)
__global__ void inject_kernel(MgLevelDev F, MgLevelDev C)
{
    const int64_t N = (int64_t)C.nx * C.ny * C.nz;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int64_t csz = (int64_t)C.nx * C.ny;
    const int cz = (int)(idx / csz);
    const int cy = (int)((idx - (int64_t)cz * csz) / C.nx);
    const int cx = (int)(idx - (int64_t)cz * csz - (int64_t)cy * C.nx);
    const int x = min(2 * cx, F.nx - 1), y = min(2 * cy, F.ny - 1), z = min(2 * cz, F.nz - 1);
    const int64_t f = x + (int64_t)F.nx * (y + (int64_t)F.ny * z);
    // the coarse level owns these arrays; they are const only for the smoother
    const_cast<float *>(C.ex)[idx] = F.ex[f];
    const_cast<float *>(C.ey)[idx] = F.ey[f];
    const_cast<float *>(C.ez)[idx] = F.ez[f];
    const_cast<float *>(C.k2)[idx] = F.k2[f];
}

__global__ void zero_kernel(float *p, int64_t n, const MgStatus *__restrict__ st)
{
    if (st && stopped(st)) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ---- FP64 outer step: phi_out = phi_in + e, f0 = b - A phi_out (stencil_at,
// pb_solver.rs:159-194), block partial sums of r^2.
__global__ __launch_bounds__(256) void outer64_kernel(MgOuterDev O, const MgStatus *__restrict__ st)
{
    __shared__ double red[4];
    if (stopped(st)) return;
    const int nx = O.nx, ny = O.ny, nz = O.nz;
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    double r2acc = 0.0;
    if (idx < N) {
        const float *__restrict__ e = O.e;
        const double *__restrict__ p = O.phi_in;
        const double pc = e ? p[idx] + (double)e[idx] : p[idx];
        if (e) O.phi_out[idx] = pc;
        const int iz = (int)(idx / sz);
        const int iy = (int)((idx - (int64_t)iz * sz) / nx);
        const int ix = (int)(idx - (int64_t)iz * sz - (int64_t)iy * nx);
        if (ix == 0 || iy == 0 || iz == 0 || ix == nx - 1 || iy == ny - 1 || iz == nz - 1) {
            O.f0[idx] = 0.0f;
        } else {
            double pxp, pxm, pyp, pym, pzp, pzm;
            if (e) {
                pxp = p[idx + 1] + (double)e[idx + 1];
                pxm = p[idx - 1] + (double)e[idx - 1];
                pyp = p[idx + sy] + (double)e[idx + sy];
                pym = p[idx - sy] + (double)e[idx - sy];
                pzp = p[idx + sz] + (double)e[idx + sz];
                pzm = p[idx - sz] + (double)e[idx - sz];
            } else {
                pxp = p[idx + 1];
                pxm = p[idx - 1];
                pyp = p[idx + sy];
                pym = p[idx - sy];
                pzp = p[idx + sz];
                pzm = p[idx - sz];
            }
            double exp_, exm, eyp, eym, ezp, ezm, k2;
            if (O.uni) {
                exp_ = exm = eyp = eym = ezp = ezm = O.eps_u;
                k2 = 0.0;
            } else {
                exp_ = O.ex[idx];
                exm = O.ex[idx - 1];
                eyp = O.ey[idx];
                eym = O.ey[idx - sy];
                ezp = O.ez[idx];
                ezm = O.ez[idx - sz];
                k2 = O.k2[idx];
            }
            const double nb = (exp_ * pxp + exm * pxm) * O.ihx + (eyp * pyp + eym * pym) * O.ihy +
                              (ezp * pzp + ezm * pzm) * O.ihz;
            const double dg = (exp_ + exm) * O.ihx + (eyp + eym) * O.ihy + (ezp + ezm) * O.ihz + k2;
            const double r = O.scale * O.rho[idx] + nb - dg * pc;
            O.f0[idx] = (float)r;
            r2acc = r * r;
        }
    }
    r2acc = wave_sum(r2acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r2acc;
    __syncthreads();
    if (threadIdx.x == 0) O.partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Convergence / divergence test of solve_lpbe_multigrid (pb_solver.rs:1085-1147).
__global__ __launch_bounds__(256) void mg_check_kernel(const double *__restrict__ partials, int nbk,
                                                       double count, double tol, int cycle,
                                                       int max_cycles, MgStatus *st)
{
    __shared__ double red[256];
    if (st->state != 0) return;
    double a = 0.0;
    for (int b = threadIdx.x; b < nbk; b += 256) a += partials[b];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double rms = count > 0.0 ? sqrt(red[0] / count) : 0.0;
    st->rms = rms;
    if (cycle == 0) return;  // residual of the initial potential: no test
    if (rms < st->best) {
        st->best = rms;
        st->growing = 0;
    } else {
        st->growing += 1;
    }
    const bool div = !isfinite(rms) || rms > st->best * 1e4 || (st->growing >= 5 && rms > st->prev);
    st->cycles = cycle;
    if (div) {
        st->state = 2;
        return;
    }
    st->prev = rms;
    if (rms < tol) {
        st->state = 1;
        return;
    }
    if (cycle >= max_cycles) st->state = 3;
}

// ---- FP64 Chebyshev SOR (solve_lpbe_initial, pb_solver.rs:373-447) ----------
__global__ __launch_bounds__(256) void sor64_kernel(MgOuterDev O, int color, double omega,
                                                    double *__restrict__ partials,
                                                    const MgStatus *__restrict__ st)
{
    __shared__ double red[4];
    if (stopped(st)) return;
    const int nx = O.nx, ny = O.ny, nz = O.nz;
    const int half = (nx - 1) / 2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t rows = (int64_t)(ny - 2) * (nz - 2);
    double acc = 0.0;
    if (t < rows * half) {
        const int s = (int)(t % half);
        const int64_t row = t / half;
        const int iy = 1 + (int)(row % (ny - 2));
        const int iz = 1 + (int)(row / (ny - 2));
        const int ix = 1 + ((1 + iy + iz + color) & 1) + 2 * s;
        if (ix <= nx - 2) {
            const int64_t sy = nx, sz = (int64_t)nx * ny;
            const int64_t idx = ix + sy * iy + sz * iz;
            double *__restrict__ p = O.phi_out;
            double exp_, exm, eyp, eym, ezp, ezm, k2;
            if (O.uni) {
                exp_ = exm = eyp = eym = ezp = ezm = O.eps_u;
                k2 = 0.0;
            } else {
                exp_ = O.ex[idx];
                exm = O.ex[idx - 1];
                eyp = O.ey[idx];
                eym = O.ey[idx - sy];
                ezp = O.ez[idx];
                ezm = O.ez[idx - sz];
                k2 = O.k2[idx];
            }
            const double nb = (exp_ * p[idx + 1] + exm * p[idx - 1]) * O.ihx +
                              (eyp * p[idx + sy] + eym * p[idx - sy]) * O.ihy +
                              (ezp * p[idx + sz] + ezm * p[idx - sz]) * O.ihz;
            const double dg = (exp_ + exm) * O.ihx + (eyp + eym) * O.ihy + (ezp + ezm) * O.ihz + k2;
            const double rhs = O.scale * O.rho[idx];
            const double gs = (nb + rhs) / dg;
            const double pn = p[idx] + omega * (gs - p[idx]);
            const double r = gs - p[idx];
            acc = r * r;
            p[idx] = pn;
        }
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void sor_check_kernel(const double *__restrict__ partials,
                                                        int nbk, double count, double tol,
                                                        int iter, int max_iter, MgStatus *st)
{
    __shared__ double red[256];
    if (st->state != 0) return;
    double a = 0.0;
    for (int b = threadIdx.x; b < nbk; b += 256) a += partials[b];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double rms = count > 0.0 ? sqrt(red[0] / count) : 0.0;
    st->rms = rms;
    st->cycles = iter;
    if (rms < tol) {
        st->state = 1;
        return;
    }
    if (iter >= max_iter) st->state = 3;
}

// ---- launchers ---------------------------------------------------------------
static inline dim3 grid1(int64_t n) { return dim3((unsigned)ceil_div(n, 256)); }

void mg_rb_sweep(const MgLevelDev &L, int color, float omega, const MgStatus *st, hipStream_t s)
{
    const int64_t work = (int64_t)(L.ny - 2) * (L.nz - 2) * ((L.nx - 1) / 2);
    if (work <= 0) return;
    hipLaunchKernelGGL(rb_sweep_kernel, grid1(work), dim3(256), 0, s, L, color, omega, st);
    LAUNCH_OK();
}

void mg_residual32(const MgLevelDev &L, const MgStatus *st, hipStream_t s)
{
    hipLaunchKernelGGL(residual32_kernel, grid1((int64_t)L.nx * L.ny * L.nz), dim3(256), 0, s, L,
                       st);
    LAUNCH_OK();
}

void mg_restrict(const MgLevelDev &F, const MgLevelDev &C, const MgStatus *st, hipStream_t s)
{
    hipLaunchKernelGGL(restrict_kernel, grid1((int64_t)C.nx * C.ny * C.nz), dim3(256), 0, s, F, C,
                       0.25f, st);
    LAUNCH_OK();
}

void mg_prolong(const MgLevelDev &C, const MgLevelDev &F, const MgStatus *st, hipStream_t s)
{
    hipLaunchKernelGGL(prolong_kernel, grid1((int64_t)F.nx * F.ny * F.nz), dim3(256), 0, s, C, F,
                       st);
    LAUNCH_OK();
}

void mg_inject(const MgLevelDev &F, const MgLevelDev &C, hipStream_t s)
{
    hipLaunchKernelGGL(inject_kernel, grid1((int64_t)C.nx * C.ny * C.nz), dim3(256), 0, s, F, C);
    LAUNCH_OK();
}

void mg_zero(float *p, int64_t n, const MgStatus *st, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_kernel, grid1(n), dim3(256), 0, s, p, n, st);
    LAUNCH_OK();
}

int mg_outer_blocks(const MgOuterDev &O) { return ceil_div((int64_t)O.nx * O.ny * O.nz, 256); }

void mg_outer64(const MgOuterDev &O, const MgStatus *st, hipStream_t s)
{
    hipLaunchKernelGGL(outer64_kernel, dim3(mg_outer_blocks(O)), dim3(256), 0, s, O, st);
    LAUNCH_OK();
}

void mg_check(const double *partials, int nbk, double count, double tol, int cycle,
              int max_cycles, MgStatus *st, hipStream_t s)
{
    hipLaunchKernelGGL(mg_check_kernel, dim3(1), dim3(256), 0, s, partials, nbk, count, tol, cycle,
                       max_cycles, st);
    LAUNCH_OK();
}

int sor64_blocks(const MgOuterDev &O)
{
    const int64_t work = (int64_t)(O.ny - 2) * (O.nz - 2) * ((O.nx - 1) / 2);
    return std::max(1, ceil_div(work, 256));
}

void mg_sor64(const MgOuterDev &O, int color, double omega, double *partials, const MgStatus *st,
              hipStream_t s)
{
    hipLaunchKernelGGL(sor64_kernel, dim3(sor64_blocks(O)), dim3(256), 0, s, O, color, omega,
                       partials, st);
    LAUNCH_OK();
}

void mg_sor_check(const double *partials, int nbk, double count, double tol, int iter,
                  int max_iter, MgStatus *st, hipStream_t s)
{
    hipLaunchKernelGGL(sor_check_kernel, dim3(1), dim3(256), 0, s, partials, nbk, count, tol, iter,
                       max_iter, st);
    LAUNCH_OK();
}

}  // namespace mmpbsa
