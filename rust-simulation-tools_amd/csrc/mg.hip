// mg.hip -- linearized-PB multigrid on gfx950, batched over independent solves.
//
// The reference V-cycle (pb_solver.rs:842-939) is affine in (phi, rhs): one
// cycle maps phi to phi + M (b - A phi).  It runs here in correction form:
//   r_k  = b - A phi_k             FP64, fused with phi_k = phi_{k-1} + e_{k-1}
//   e_k  = V(0, r_k)               FP32, same smoother and omega per level, the
//                                  0.25 restriction scale, and prolongation onto
//                                  every node (even-sized levels extrapolate, as
//                                  the reference does) -- so the iterate, its
//                                  boundary drift and the cycle count follow the
//                                  FP64 reference to FP32 precision of e_k
//   test: RMS of r_k (FP64), the reference's convergence test at that cycle.
// Each solve has a status word; kernels of a finished solve return at once,
// so cycles can be enqueued (or replayed from a graph) in batches.
// Levels below ~16k nodes run inside one workgroup (mg_tail), which replaces
// roughly a hundred tiny launches per cycle.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.h"
#include "mg.h"

namespace mmpbsa {

__device__ __forceinline__ bool stopped(const MgStatus *st) { return st->state != 0; }

// ---- node operations shared by the batched kernels and the tail --------------
template <typename T>
__device__ __forceinline__ T face_of(uint32_t code, T cin, T cout, const T *__restrict__ a, int64_t i)
{
    return code == 0u ? cout : (code == 1u ? cin : a[i]);
}

// The six face dielectrics and kappa-bar^2 of interior node idx.
template <typename T, typename V>
__device__ __forceinline__ void coeffs(const V &L, int64_t idx, T &exp_, T &exm, T &eyp, T &eym,
                                       T &ezp, T &ezm, T &k2)
{
    const int64_t sy = L.nx, sz = (int64_t)L.nx * L.ny;
    if (L.cw) {
        const uint32_t w = L.cw[idx];
        exp_ = face_of<T>(w & 3u, L.c_in, L.c_out, L.ex, idx);
        exm = face_of<T>((w >> 2) & 3u, L.c_in, L.c_out, L.ex, idx - 1);
        eyp = face_of<T>((w >> 4) & 3u, L.c_in, L.c_out, L.ey, idx);
        eym = face_of<T>((w >> 6) & 3u, L.c_in, L.c_out, L.ey, idx - sy);
        ezp = face_of<T>((w >> 8) & 3u, L.c_in, L.c_out, L.ez, idx);
        ezm = face_of<T>((w >> 10) & 3u, L.c_in, L.c_out, L.ez, idx - sz);
        k2 = ((w >> 12) & 1u) ? (T)L.c_k2 : (T)0;
    } else {
        exp_ = L.ex[idx];
        exm = L.ex[idx - 1];
        eyp = L.ey[idx];
        eym = L.ey[idx - sy];
        ezp = L.ez[idx];
        ezm = L.ez[idx - sz];
        k2 = L.k2[idx];
    }
}

// mode 0: ordinary update; 1: first red half-sweep of a zero correction (no
// neighbour reads); 2: first black half-sweep (own old value taken as zero).
__device__ __forceinline__ void rb_node(const MgLevelDev &L, int64_t idx, int mode)
{
    const int64_t sy = L.nx, sz = (int64_t)L.nx * L.ny;
    float *__restrict__ e = L.e;
    float nb = 0.0f, dg;
    if (L.uni) {
        const float u = L.eps_u;
        if (mode != 1)
            nb = u * ((e[idx + 1] + e[idx - 1]) * L.ihx + (e[idx + sy] + e[idx - sy]) * L.ihy +
                      (e[idx + sz] + e[idx - sz]) * L.ihz);
        dg = u * (2.0f * L.ihx + 2.0f * L.ihy + 2.0f * L.ihz);
    } else {
        float exp_, exm, eyp, eym, ezp, ezm, k2;
        coeffs<float>(L, idx, exp_, exm, eyp, eym, ezp, ezm, k2);
        if (mode != 1)
            nb = fmaf(exp_, e[idx + 1], exm * e[idx - 1]) * L.ihx +
                 fmaf(eyp, e[idx + sy], eym * e[idx - sy]) * L.ihy +
                 fmaf(ezp, e[idx + sz], ezm * e[idx - sz]) * L.ihz;
        dg = (exp_ + exm) * L.ihx + (eyp + eym) * L.ihy + (ezp + ezm) * L.ihz + k2;
    }
    const float old = mode == 0 ? e[idx] : 0.0f;
    const float gs = (nb + L.f[idx]) / dg;
    e[idx] = fmaf(L.omega, gs - old, old);
}

// work item t of a half-sweep -> node index (or -1)
__device__ __forceinline__ int64_t rb_index(const MgLevelDev &L, int64_t t, int color)
{
    const uint32_t half = L.fd_half.d;
    const uint32_t rows = (uint32_t)(L.ny - 2) * (uint32_t)(L.nz - 2);
    if (t >= (int64_t)rows * half) return -1;
    const uint32_t ut = (uint32_t)t;
    const uint32_t row = L.fd_half.div(ut);
    const uint32_t s = ut - row * half;
    const uint32_t zq = L.fd_row.div(row);
    const int iy = 1 + (int)(row - zq * (uint32_t)(L.ny - 2));
    const int iz = 1 + (int)zq;
    const int ix = 1 + ((1 + iy + iz + color) & 1) + 2 * (int)s;
    if (ix > L.nx - 2) return -1;
    return (int64_t)ix + (int64_t)L.nx * (iy + (int64_t)L.ny * iz);
}

// node index -> (ix, iy, iz) with the level's FastDiv helpers
template <typename V>
__device__ __forceinline__ void split3(const V &L, int64_t idx, int &ix, int &iy, int &iz)
{
    const uint32_t u = (uint32_t)idx;
    const uint32_t z = L.fd_nxy.div(u);
    const uint32_t r = u - z * L.fd_nxy.d;
    const uint32_t y = L.fd_nx.div(r);
    iz = (int)z;
    iy = (int)y;
    ix = (int)(r - y * L.fd_nx.d);
}

void mg_level_index(MgLevelDev &L)
{
    L.fd_half = FastDiv((uint32_t)std::max(1, (L.nx - 1) / 2));
    L.fd_row = FastDiv((uint32_t)std::max(1, L.ny - 2));
    L.fd_nx = FastDiv((uint32_t)std::max(1, L.nx));
    L.fd_nxy = FastDiv((uint32_t)std::max(1, L.nx * L.ny));
}

__device__ __forceinline__ int64_t rb_work(const MgLevelDev &L)
{
    if (L.nx < 3 || L.ny < 3 || L.nz < 3) return 0;
    return (int64_t)(L.ny - 2) * (L.nz - 2) * ((L.nx - 1) / 2);
}

__device__ __forceinline__ void residual_node(const MgLevelDev &L, int64_t idx)
{
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    int ix, iy, iz;
    split3(L, idx, ix, iy, iz);
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
        float exp_, exm, eyp, eym, ezp, ezm, k2;
        coeffs<float>(L, idx, exp_, exm, eyp, eym, ezp, ezm, k2);
        nb = fmaf(exp_, e[idx + 1], exm * e[idx - 1]) * L.ihx +
             fmaf(eyp, e[idx + sy], eym * e[idx - sy]) * L.ihy +
             fmaf(ezp, e[idx + sz], ezm * e[idx - sz]) * L.ihz;
        dg = (exp_ + exm) * L.ihx + (eyp + eym) * L.ihy + (ezp + ezm) * L.ihz + k2;
    }
    L.res[idx] = L.f[idx] + nb - dg * e[idx];
}

// restrict_into (pb_solver.rs:578-624) for coarse node idx; clears its correction.
__device__ __forceinline__ void restrict_node(const MgLevelDev &F, const MgLevelDev &C, int64_t idx)
{
    int cx, cy, cz;
    split3(C, idx, cx, cy, cz);
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
    C.f[idx] = 0.25f * sum / ws;
    C.e[idx] = 0.0f;
}

// prolongate_add (pb_solver.rs:628-681) for fine node idx
__device__ __forceinline__ void prolong_node(const MgLevelDev &C, const MgLevelDev &F, int64_t idx)
{
    int fx, fy, fz;
    split3(F, idx, fx, fy, fz);
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

__device__ __forceinline__ int64_t nodes(const MgLevelDev &L) { return (int64_t)L.nx * L.ny * L.nz; }

// ---- batched kernels (blockIdx.y = solve) -----------------------------------
__global__ __launch_bounds__(256) void rb_sweep_kernel(const MgLevelDev *__restrict__ Ls, int color,
                                                       int first, const MgStatus *__restrict__ st)
{
    const int s = blockIdx.y;
    if (stopped(st + s)) return;
    const MgLevelDev L = Ls[s];
    if (L.nx == 0) return;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t idx = rb_index(L, t, color);
    if (idx < 0) return;
    rb_node(L, idx, first ? (color == 0 ? 1 : 2) : 0);
}

__global__ __launch_bounds__(256) void residual32_kernel(const MgLevelDev *__restrict__ Ls,
                                                         const MgStatus *__restrict__ st)
{
    const int s = blockIdx.y;
    if (stopped(st + s)) return;
    const MgLevelDev L = Ls[s];
    if (L.nx == 0) return;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes(L)) return;
    residual_node(L, idx);
}

__global__ __launch_bounds__(256) void restrict_kernel(const MgLevelDev *__restrict__ Fs,
                                                       const MgLevelDev *__restrict__ Cs,
                                                       const MgStatus *__restrict__ st)
{
    const int s = blockIdx.y;
    if (stopped(st + s)) return;
    const MgLevelDev C = Cs[s];
    if (C.nx == 0) return;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes(C)) return;
    restrict_node(Fs[s], C, idx);
}

__global__ __launch_bounds__(256) void prolong_kernel(const MgLevelDev *__restrict__ Cs,
                                                      const MgLevelDev *__restrict__ Fs,
                                                      const MgStatus *__restrict__ st)
{
    const int s = blockIdx.y;
    if (stopped(st + s)) return;
    const MgLevelDev F = Fs[s];
    if (F.nx == 0) return;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes(F)) return;
    prolong_node(Cs[s], F, idx);
}

// One workgroup per solve runs the V-cycle over its tail levels.
__global__ __launch_bounds__(1024) void tail_kernel(const MgTailDev *__restrict__ Ts,
                                                    const MgStatus *__restrict__ st)
{
    constexpr int kLds = 16384;  // floats: the coarsest levels live in LDS
    __shared__ float lds[kLds];
    const int s = blockIdx.x;
    if (stopped(st + s)) return;
    const int n = Ts[s].n;
    if (n == 0) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    MgLevelDev lv[kMaxTail];
    for (int l = 0; l < n; ++l) lv[l] = Ts[s].lv[l];
    {
        // from the coarsest level upwards, move levels (except the entry level,
        // whose correction is read back by the caller) into LDS while they fit
        int used = 0;
        for (int l = n - 1; l >= 1; --l) {
            const int N = (int)nodes(lv[l]);
            const int need = (lv[l].uni ? 3 : 7) * N;
            if (used + need > kLds) break;
            float *base = lds + used;
            if (!lv[l].uni) {
                float *cx = base + 3 * N, *cy = base + 4 * N, *cz = base + 5 * N, *ck = base + 6 * N;
                for (int i = tid; i < N; i += nt) {
                    cx[i] = lv[l].ex[i];
                    cy[i] = lv[l].ey[i];
                    cz[i] = lv[l].ez[i];
                    ck[i] = lv[l].k2[i];
                }
                lv[l].ex = cx;
                lv[l].ey = cy;
                lv[l].ez = cz;
                lv[l].k2 = ck;
            }
            lv[l].e = base;
            lv[l].f = base + N;
            lv[l].res = base + 2 * N;
            used += need;
        }
        __syncthreads();
    }
    auto sweeps = [&](const MgLevelDev &L, int count) {
        const int64_t w = rb_work(L);
        for (int k = 0; k < count; ++k)
            for (int color = 0; color < 2; ++color) {
                for (int64_t t = tid; t < w; t += nt) {
                    const int64_t idx = rb_index(L, t, color);
                    if (idx >= 0) rb_node(L, idx, 0);
                }
                __syncthreads();
            }
    };
    {
        const MgLevelDev L0 = lv[0];
        for (int64_t i = tid; i < nodes(L0); i += nt) L0.e[i] = 0.0f;
        __syncthreads();
    }
    for (int l = 0; l + 1 < n; ++l) {
        const MgLevelDev L = lv[l];
        const MgLevelDev C = lv[l + 1];
        sweeps(L, 3);
        for (int64_t i = tid; i < nodes(L); i += nt) residual_node(L, i);
        __syncthreads();
        for (int64_t i = tid; i < nodes(C); i += nt) restrict_node(L, C, i);
        __syncthreads();
    }
    sweeps(lv[n - 1], 50);
    for (int l = n - 2; l >= 0; --l) {
        const MgLevelDev L = lv[l];
        const MgLevelDev C = lv[l + 1];
        for (int64_t i = tid; i < nodes(L); i += nt) prolong_node(C, L, i);
        __syncthreads();
        sweeps(L, 3);
    }
}

// ---- FP64 outer step: phi_out = phi_in + e, f0 = b - A phi_out (stencil_at,
// pb_solver.rs:159-194), block partial sums of r^2; clears the boundary of the
// next correction buffer (ezero) for the next V-cycle.
__global__ __launch_bounds__(256) void outer64_kernel(const MgOuterDev *__restrict__ Os,
                                                      float *const *__restrict__ ezero,
                                                      const MgStatus *__restrict__ st)
{
    __shared__ double red[4];
    const int s = blockIdx.y;
    if (stopped(st + s)) return;
    const MgOuterDev O = Os[s];
    if (blockIdx.x >= (unsigned)O.nbk) return;
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
        int ix, iy, iz;
        split3(O, idx, ix, iy, iz);
        if (ix == 0 || iy == 0 || iz == 0 || ix == nx - 1 || iy == ny - 1 || iz == nz - 1) {
            O.f0[idx] = 0.0f;
            if (ezero && ezero[s]) ezero[s][idx] = 0.0f;
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
                coeffs<double>(O, idx, exp_, exm, eyp, eym, ezp, ezm, k2);
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

// Convergence / divergence test of solve_lpbe_multigrid (pb_solver.rs:1085-1147),
// one block per solve.  count_cycle = 0 records the initial residual only.
__global__ __launch_bounds__(256) void mg_check_kernel(const MgOuterDev *__restrict__ Os,
                                                       int count_cycle, MgStatus *__restrict__ sts)
{
    __shared__ double red[256];
    const int s = blockIdx.x;
    MgStatus *st = sts + s;
    if (st->state != 0) return;
    const double *__restrict__ partials = Os[s].partials;
    const int nbk = Os[s].nbk;
    double a = 0.0;
    for (int b = threadIdx.x; b < nbk; b += 256) a += partials[b];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double rms = st->count > 0.0 ? sqrt(red[0] / st->count) : 0.0;
    st->rms = rms;
    if (!count_cycle) return;
    const int cycle = st->cycles + 1;
    st->cycles = cycle;
    if (rms < st->best) {
        st->best = rms;
        st->growing = 0;
    } else {
        st->growing += 1;
    }
    const bool div = !isfinite(rms) || rms > st->best * 1e4 || (st->growing >= 5 && rms > st->prev);
    if (div) {
        st->state = 2;
        return;
    }
    st->prev = rms;
    if (rms < st->tol) {
        st->state = 1;
        return;
    }
    if (cycle >= st->max_cycles) st->state = 3;
}

// coarsen_dielectrics / coarsen_kappa by injection (pb_solver.rs:684-748)
__global__ void inject_kernel(MgLevelDev F, MgLevelDev C)
{
    const int64_t N = nodes(C);
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

// Packed finest-level coefficient words (layout in MgLevelDev).  Face values
// equal to eps_in / eps_out get codes 1 / 0 (exact FP64 equality), any other
// value code 2; kappa-bar^2 is either 0 or c_k2 by construction.
__global__ void pack_codes_kernel(int nx, int ny, int nz, const double *__restrict__ ex,
                                  const double *__restrict__ ey, const double *__restrict__ ez,
                                  const double *__restrict__ k2, double cin, double cout,
                                  double ck2, uint32_t *__restrict__ cw)
{
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    auto code = [&](const double *a, int64_t i) -> uint32_t {
        if (i < 0) return 0u;
        const double v = a[i];
        return v == cout ? 0u : (v == cin ? 1u : 2u);
    };
    uint32_t w = code(ex, idx) | (code(ex, idx - 1) << 2) | (code(ey, idx) << 4) |
                 (code(ey, idx - sy) << 6) | (code(ez, idx) << 8) | (code(ez, idx - sz) << 10);
    if (k2[idx] != 0.0) w |= 1u << 12;
    cw[idx] = w;
}

void mg_pack_codes(int nx, int ny, int nz, const double *ex, const double *ey, const double *ez,
                   const double *k2, double c_in, double c_out, double c_k2, uint32_t *cw,
                   hipStream_t s)
{
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(pack_codes_kernel, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, nx, ny,
                       nz, ex, ey, ez, k2, c_in, c_out, c_k2, cw);
    LAUNCH_OK();
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
                coeffs<double>(O, idx, exp_, exm, eyp, eym, ezp, ezm, k2);
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
static inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, ceil_div(n, 256)); }

void mg_rb_sweep(const MgLevelDev *L, int n_solves, int64_t max_work, int color, int first,
                 const MgStatus *st, hipStream_t s)
{
    if (max_work <= 0 || n_solves <= 0) return;
    hipLaunchKernelGGL(rb_sweep_kernel, dim3(nblk(max_work), n_solves), dim3(256), 0, s, L, color,
                       first, st);
    LAUNCH_OK();
}

void mg_residual32(const MgLevelDev *L, int n_solves, int64_t max_nodes, const MgStatus *st,
                   hipStream_t s)
{
    if (max_nodes <= 0 || n_solves <= 0) return;
    hipLaunchKernelGGL(residual32_kernel, dim3(nblk(max_nodes), n_solves), dim3(256), 0, s, L, st);
    LAUNCH_OK();
}

void mg_restrict(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int64_t max_coarse,
                 const MgStatus *st, hipStream_t s)
{
    if (max_coarse <= 0 || n_solves <= 0) return;
    hipLaunchKernelGGL(restrict_kernel, dim3(nblk(max_coarse), n_solves), dim3(256), 0, s, F, C, st);
    LAUNCH_OK();
}

void mg_prolong(const MgLevelDev *C, const MgLevelDev *F, int n_solves, int64_t max_fine,
                const MgStatus *st, hipStream_t s)
{
    if (max_fine <= 0 || n_solves <= 0) return;
    hipLaunchKernelGGL(prolong_kernel, dim3(nblk(max_fine), n_solves), dim3(256), 0, s, C, F, st);
    LAUNCH_OK();
}

void mg_tail(const MgTailDev *T, int n_solves, const MgStatus *st, hipStream_t s)
{
    if (n_solves <= 0) return;
    hipLaunchKernelGGL(tail_kernel, dim3(n_solves), dim3(1024), 0, s, T, st);
    LAUNCH_OK();
}

void mg_outer64(const MgOuterDev *O, int n_solves, int64_t max_nodes, const MgStatus *st,
                hipStream_t s)
{
    mg_outer64_z(O, nullptr, n_solves, max_nodes, st, s);
}

void mg_outer64_z(const MgOuterDev *O, float *const *ezero, int n_solves, int64_t max_nodes,
                  const MgStatus *st, hipStream_t s)
{
    if (n_solves <= 0) return;
    hipLaunchKernelGGL(outer64_kernel, dim3(nblk(max_nodes), n_solves), dim3(256), 0, s, O, ezero,
                       st);
    LAUNCH_OK();
}

void mg_check(const MgOuterDev *O, int n_solves, int count_cycle, MgStatus *st, hipStream_t s)
{
    if (n_solves <= 0) return;
    hipLaunchKernelGGL(mg_check_kernel, dim3(n_solves), dim3(256), 0, s, O, count_cycle, st);
    LAUNCH_OK();
}

void mg_status_init(MgStatus *st, const MgStatus *init_host, int n, hipStream_t s)
{
    HIP_OK(hipMemcpyAsync(st, init_host, sizeof(MgStatus) * n, hipMemcpyHostToDevice, s));
}

void mg_inject(const MgLevelDev &F, const MgLevelDev &C, hipStream_t s)
{
    hipLaunchKernelGGL(inject_kernel, dim3(nblk((int64_t)C.nx * C.ny * C.nz)), dim3(256), 0, s, F, C);
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
