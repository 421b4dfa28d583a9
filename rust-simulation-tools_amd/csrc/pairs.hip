// pairs.hip -- all-pairs kernels of the MM-PBSA path on gfx950.
//
// Atoms are staged through LDS in tiles of kTile; each lane owns one atom of
// the row tile and walks the column tile.  Every reduction writes per-block
// partials that a fixed-order kernel sums, so results are bit-reproducible.
// Built with -ffp-contract=off: an FMA appears only where fma() is written.
#include <hip/hip_runtime.h>

#include <atomic>

#include <algorithm>

#include "kernels.h"

namespace mmpbsa {

// ---------------------------------------------------------------------------
// HCT descreening integral for one ordered pair (gb_energy.rs:164-215).
// For j's sphere wholly outside i's (rho_i <= r - s) the closed form reduces to
// (1/r) * sum_{k>=1} k/(2k+1) x^(2k+1), x = s/r, which has no cancellation;
// eight terms are exact to 3e-10 relative for x < 0.25.  Other pairs use the
// reference expression.
__device__ __forceinline__ double hct_term(double r2, double rho_i, double sj)
{
    if (r2 == 0.0) return 0.0;
    const double rinv = rsqrt(r2);
    const double x = sj * rinv;
    const double r_est = r2 * rinv;
    if (x < 0.25 && rho_i <= r_est - sj) {
        const double x2 = x * x;
        double p = 8.0 / 17.0;
        p = fma(p, x2, 7.0 / 15.0);
        p = fma(p, x2, 6.0 / 13.0);
        p = fma(p, x2, 5.0 / 11.0);
        p = fma(p, x2, 4.0 / 9.0);
        p = fma(p, x2, 3.0 / 7.0);
        p = fma(p, x2, 2.0 / 5.0);
        p = fma(p, x2, 1.0 / 3.0);
        return x * x2 * p * rinv;
    }
    const double r = sqrt(r2);
    if (rho_i >= r + sj) return 0.0;
    const double u = r + sj;
    const double ad = fabs(r - sj);
    const double l = rho_i > ad ? rho_i : ad;
    const double li = 1.0 / l, ui = 1.0 / u;
    double c = 0.5 * (li - ui + 0.25 * (ui * ui - li * li) * (r - sj * sj / r) +
                      log(l / u) / (2.0 * r));
    if (rho_i < sj - r) c += 2.0 * (1.0 / rho_i - li);
    return c;
}

__global__ __launch_bounds__(kTile) void gb_psi_kernel(const double *__restrict__ xyz,
                                                        const double *__restrict__ radii,
                                                        const double *__restrict__ screen,
                                                        int n, double offset, double cut2,
                                                        int chunk, double *__restrict__ part)
{
    __shared__ double sx[kTile], sy[kTile], sz[kTile], ss[kTile];
    const int t = threadIdx.x;
    const int i = blockIdx.x * kTile + t;
    const int j0 = blockIdx.y * chunk;
    const int j1 = min(n, j0 + chunk);
    const bool act = i < n;
    double xi = 0, yi = 0, zi = 0, rho_i = 1;
    if (act) {
        xi = xyz[3 * i];
        yi = xyz[3 * i + 1];
        zi = xyz[3 * i + 2];
        rho_i = radii[i] - offset;
    }
    double acc = 0.0;
    for (int jt = j0; jt < j1; jt += kTile) {
        const int j = jt + t;
        if (j < j1) {
            sx[t] = xyz[3 * j];
            sy[t] = xyz[3 * j + 1];
            sz[t] = xyz[3 * j + 2];
            ss[t] = screen[j] * (radii[j] - offset);
        }
        __syncthreads();
        const int m = min(kTile, j1 - jt);
        if (act) {
            for (int k = 0; k < m; ++k) {
                const double dx = xi - sx[k], dy = yi - sy[k], dz = zi - sz[k];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (jt + k == i || r2 > cut2) continue;
                acc += hct_term(r2, rho_i, ss[k]);
            }
        }
        __syncthreads();
    }
    if (act) part[(size_t)blockIdx.y * n + i] = acc;
}

// OBC / HCT radius map, gb_energy.rs:219-255.
__device__ __forceinline__ double born_radius(double psi, double radius, double offset, int model,
                                              double rgbmax)
{
    const double rho = radius - offset;
    double inv;
    if (model == 0) {
        inv = 1.0 / rho - psi;
    } else {
        const double a = model == 1 ? 0.8 : 1.0;
        const double b = model == 1 ? 0.0 : 0.8;
        const double g = model == 1 ? 2.909125 : 4.85;
        const double pr = psi * rho;
        const double p2 = pr * pr;
        const double p3 = p2 * pr;
        const double th = tanh(a * pr - b * p2 + g * p3);
        inv = 1.0 / rho - th / radius;
    }
    double R = inv > 0.0 ? 1.0 / inv : rgbmax;
    if (R > rgbmax) R = rgbmax;
    return R;
}

__global__ void gb_finalize_kernel(const double *__restrict__ part, int chunks, int n,
                                   const double *__restrict__ radii, double offset, int model,
                                   double rgbmax, double *__restrict__ born)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double psi = 0.0;
    for (int c = 0; c < chunks; ++c) psi += part[(size_t)c * n + i];
    born[i] = born_radius(psi, radii[i], offset, model, rgbmax);
}

// psi of the atoms of A from the atoms of B (complex indices; the partners of
// a split are disjoint, so no pair is an atom with itself), j-chunked as
// gb_psi_kernel: part[chunk][ia]
__global__ __launch_bounds__(kTile) void gb_psi_cross_kernel(
    const double *__restrict__ xyz, const double *__restrict__ radii,
    const double *__restrict__ screen, const int *__restrict__ A, int nA,
    const int *__restrict__ B, int nB, double offset, double cut2, int chunk,
    double *__restrict__ part)
{
    __shared__ double sx[kTile], sy[kTile], sz[kTile], ss[kTile];
    const int t = threadIdx.x;
    const int ia = blockIdx.x * kTile + t;
    const int j0 = blockIdx.y * chunk;
    const int j1 = min(nB, j0 + chunk);
    const bool act = ia < nA;
    double xi = 0, yi = 0, zi = 0, rho_i = 1;
    if (act) {
        const int i = A[ia];
        xi = xyz[3 * i];
        yi = xyz[3 * i + 1];
        zi = xyz[3 * i + 2];
        rho_i = radii[i] - offset;
    }
    double acc = 0.0;
    for (int jt = j0; jt < j1; jt += kTile) {
        const int jj = jt + t;
        if (jj < j1) {
            const int j = B[jj];
            sx[t] = xyz[3 * j];
            sy[t] = xyz[3 * j + 1];
            sz[t] = xyz[3 * j + 2];
            ss[t] = screen[j] * (radii[j] - offset);
        }
        __syncthreads();
        const int m = min(kTile, j1 - jt);
        if (act) {
            for (int k = 0; k < m; ++k) {
                const double dx = xi - sx[k], dy = yi - sy[k], dz = zi - sz[k];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > cut2) continue;
                acc += hct_term(r2, rho_i, ss[k]);
            }
        }
        __syncthreads();
    }
    if (act) part[(size_t)blockIdx.y * nA + ia] = acc;
}

// Born radii of one partner's atoms in the complex: psi from the partner's
// own partials plus the cross partials, atom map[p] of the complex
__global__ void gb_finalize_parts_kernel(const double *__restrict__ own, int own_chunks,
                                         const double *__restrict__ cross, int cross_chunks,
                                         const int *__restrict__ map, int n,
                                         const double *__restrict__ radii, double offset,
                                         int model, double rgbmax, double *__restrict__ born)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double psi = 0.0;
    for (int c = 0; c < own_chunks; ++c) psi += own[(size_t)c * n + p];
    double x = 0.0;
    for (int c = 0; c < cross_chunks; ++c) x += cross[(size_t)c * n + p];
    const int a = map[p];
    born[a] = born_radius(psi + x, radii[a], offset, model, rgbmax);
}

// j-chunks of the Born-radius kernel: enough workgroups to fill the GPU
// (about 4096) unless the tuning table fixes the count
static std::atomic<int> g_gb_chunks_override{0};
void set_gb_chunks_override(int c) { g_gb_chunks_override.store(c > 0 ? c : 0); }

int gb_chunks(int n)
{
    const int nb = std::max(1, ceil_div(n, kTile));
    const int o = g_gb_chunks_override.load();
    int c = o > 0 ? std::min(o, 64) : std::max(1, std::min(64, ceil_div(4096, nb)));
    c = std::min(c, std::max(1, ceil_div(n, kTile)));
    return c;
}

int gb_psi_chunks(int n)
{
    if (n <= 0) return 0;
    const int chunk = ceil_div(ceil_div(n, gb_chunks(n)), kTile) * kTile;
    return ceil_div(n, chunk);
}

void launch_gb_radii(const double *xyz, const PairTopo &t, int model, double offset,
                     double rgbmax, double cutoff, double *psi_work, double *born_out,
                     hipStream_t s)
{
    if (t.n == 0) return;
    const int chunks = gb_chunks(t.n);
    const int chunk = ceil_div(ceil_div(t.n, chunks), kTile) * kTile;
    const int used = ceil_div(t.n, chunk);
    dim3 grid(ceil_div(t.n, kTile), used);
    hipLaunchKernelGGL(gb_psi_kernel, grid, dim3(kTile), 0, s, xyz, t.radii, t.screen, t.n,
                       offset, cutoff * cutoff, chunk, psi_work);
    LAUNCH_OK();
    hipLaunchKernelGGL(gb_finalize_kernel, dim3(ceil_div(t.n, 256)), dim3(256), 0, s, psi_work,
                       used, t.n, t.radii, offset, model, rgbmax, born_out);
    LAUNCH_OK();
}

// ---------------------------------------------------------------------------
// Block reduction helpers (4 waves of 64).
template <int NT>
__device__ inline void block_store(double v, double *red, double *out)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < NT / 64; ++k) s += red[k];
        *out = s;
    }
    __syncthreads();
}

template <int NT>
__device__ inline void block_store_dd(dd v, dd *red, double *out)
{
    v = wave_sum_dd(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        dd s = {0.0, 0.0};
        for (int k = 0; k < NT / 64; ++k) dd_add(s, red[k]);
        out[0] = s.s;
        out[1] = s.c;
    }
    __syncthreads();
}

// GB energy: self terms on diagonal tiles plus 2 * sum_{i<j} pair terms,
// gb_energy.rs:283-355.  Grid (nb, nb); blocks below the diagonal exit.
__global__ __launch_bounds__(kTile) void gb_energy_kernel(
    const double *__restrict__ xyz, const double *__restrict__ q, const double *__restrict__ R,
    int n, int nb, double kappa, double inv_ein, double inv_eout, double cut2,
    double *__restrict__ partials)
{
    __shared__ double sx[kTile], sy[kTile], sz[kTile], sR[kTile], sq[kTile];
    __shared__ double red[kTile / 64];
    const int ib = blockIdx.x, jb = blockIdx.y;
    double *outp = partials + (size_t)ib * nb + jb;
    if (jb < ib) {
        if (threadIdx.x == 0) *outp = 0.0;
        return;
    }
    const int t = threadIdx.x;
    const int i = ib * kTile + t;
    const int j = jb * kTile + t;
    if (j < n) {
        sx[t] = xyz[3 * j];
        sy[t] = xyz[3 * j + 1];
        sz[t] = xyz[3 * j + 2];
        sR[t] = R[j];
        sq[t] = q[j];
    }
    __syncthreads();
    const bool no_salt = !(kappa > 0.0);
    const double dns = -0.5 * (inv_ein - inv_eout);
    const double KS = 0.73;
    double acc = 0.0;
    if (i < n) {
        const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
        const double qi = q[i], Ri = R[i];
        if (ib == jb) {
            if (no_salt)
                acc += dns * qi * qi / Ri;
            else
                acc += -0.5 * qi * qi * (inv_ein - exp(-KS * kappa * Ri) * inv_eout) / Ri;
        }
        const int m = min(kTile, n - jb * kTile);
        const int k0 = (ib == jb) ? t + 1 : 0;
        for (int k = k0; k < m; ++k) {
            const double dx = xi - sx[k], dy = yi - sy[k], dz = zi - sz[k];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > cut2) continue;
            const double rr = Ri * sR[k];
            const double f = sqrt(r2 + rr * exp(-r2 / (4.0 * rr)));
            double pe;
            if (no_salt)
                pe = dns * qi * sq[k] / f;
            else
                pe = -0.5 * qi * sq[k] * (inv_ein - exp(-KS * kappa * f) * inv_eout) / f;
            acc += 2.0 * pe;
        }
    }
    block_store<kTile>(acc, red, outp);
}

// j-chunks of the cross kernel: about 4096 workgroups, at most one per tile of B
static int cross_chunk_len(int nA, int nB)
{
    const int nb = std::max(1, ceil_div(nA, kTile));
    const int c = std::max(1, std::min(ceil_div(4096, nb), ceil_div(nB, kTile)));
    return ceil_div(ceil_div(nB, c), kTile) * kTile;
}

int gb_psi_cross_chunks(int nA, int nB)
{
    return nA > 0 && nB > 0 ? ceil_div(nB, cross_chunk_len(nA, nB)) : 0;
}

void launch_gb_psi_cross(const double *xyz, const PairTopo &t, const int *A, int nA, const int *B,
                         int nB, double offset, double cutoff, double *part, hipStream_t s)
{
    if (nA <= 0 || nB <= 0) return;
    const int chunk = cross_chunk_len(nA, nB);
    hipLaunchKernelGGL(gb_psi_cross_kernel, dim3(ceil_div(nA, kTile), ceil_div(nB, chunk)),
                       dim3(kTile), 0, s, xyz, t.radii, t.screen, A, nA, B, nB, offset,
                       cutoff * cutoff, chunk, part);
    LAUNCH_OK();
}

void launch_gb_finalize_parts(const double *own, int own_chunks, const double *cross,
                              int cross_chunks, const int *map, int n, const PairTopo &t,
                              int model, double offset, double rgbmax, double *born,
                              hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(gb_finalize_parts_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, own,
                       own_chunks, cross, cross_chunks, map, n, t.radii, offset, model, rgbmax,
                       born);
    LAUNCH_OK();
}

int gb_energy_blocks(int n)
{
    const int nb = std::max(1, ceil_div(n, kTile));
    return nb * nb;
}

void launch_gb_energy(const double *xyz, const PairTopo &t, const double *born, double kappa,
                      double eps_in, double eps_out, double cutoff, double *partials,
                      hipStream_t s)
{
    const int nb = std::max(1, ceil_div(t.n, kTile));
    hipLaunchKernelGGL(gb_energy_kernel, dim3(nb, nb), dim3(kTile), 0, s, xyz, t.q_amber, born,
                       t.n, nb, kappa, 1.0 / eps_in, 1.0 / eps_out, cutoff * cutoff, partials);
    LAUNCH_OK();
}

// ---------------------------------------------------------------------------
// Non-bonded MM, mm_energy.rs:217-259.  1-4 membership is tested before the
// exclusion test, as the reference does.  Sums are double-double.
__global__ __launch_bounds__(kTile) void mm_pair_kernel(const double *__restrict__ xyz,
                                                         PairTopo T, double *__restrict__ partials)
{
    __shared__ double sx[kTile], sy[kTile], sz[kTile], sq[kTile];
    __shared__ int st[kTile];
    __shared__ dd red[kTile / 64];
    const int ib = blockIdx.x, jb = blockIdx.y, nb = T.nb, n = T.n;
    double *outp = partials + ((size_t)ib * nb + jb) * 8;
    if (jb < ib) {
        if (threadIdx.x < 8) outp[threadIdx.x] = 0.0;
        return;
    }
    const int t = threadIdx.x;
    const int i = ib * kTile + t;
    const int j0 = jb * kTile;
    if (j0 + t < n) {
        const int j = j0 + t;
        sx[t] = xyz[3 * j];
        sy[t] = xyz[3 * j + 1];
        sz[t] = xyz[3 * j + 2];
        sq[t] = T.q_amber[j];
        st[t] = T.type[j];
    }
    __syncthreads();
    dd vdw = {0, 0}, elec = {0, 0}, v14 = {0, 0}, e14 = {0, 0};
    if (i < n) {
        const bool flagged = T.tile_flag[(size_t)ib * nb + jb] != 0;
        uint64_t exm[4] = {0, 0, 0, 0}, p4m[4] = {0, 0, 0, 0};
        if (flagged) {
            for (int k = T.ex_ptr[i]; k < T.ex_ptr[i + 1]; ++k) {
                const int d = T.ex_col[k] - j0;
                if (d >= 0 && d < kTile) exm[d >> 6] |= 1ull << (d & 63);
            }
            for (int k = T.p14_ptr[i]; k < T.p14_ptr[i + 1]; ++k) {
                const int d = T.p14_col[k] - j0;
                if (d >= 0 && d < kTile) p4m[d >> 6] |= 1ull << (d & 63);
            }
        }
        const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
        const double qi = T.q_amber[i];
        const double *rowA = T.ljA + (size_t)T.type[i] * T.n_types;
        const double *rowB = T.ljB + (size_t)T.type[i] * T.n_types;
        const int m = min(kTile, n - j0);
        const int k0 = (ib == jb) ? t + 1 : 0;
        const double iscnb = 1.0 / T.scnb, iscee = 1.0 / T.scee;
        for (int k = k0; k < m; ++k) {
            const double dx = xi - sx[k], dy = yi - sy[k], dz = zi - sz[k];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool is14 = (p4m[k >> 6] >> (k & 63)) & 1ull;
            const bool isex = (exm[k >> 6] >> (k & 63)) & 1ull;
            if (!is14 && isex) continue;
            // one division per pair: A / r^12 - B / r^6 and q q / r from 1 / r
            // (terms within a few ulp of the divided forms)
            const double ir = 1.0 / sqrt(r2);
            const double ir2 = ir * ir;
            const double ir6 = ir2 * ir2 * ir2;
            const double A = rowA[st[k]], B = rowB[st[k]];
            const double lj = A * ir6 * ir6 - B * ir6;
            const double cq = qi * sq[k] * ir;
            if (is14) {
                dd_add(v14, lj * iscnb);
                dd_add(e14, cq * iscee);
            } else {
                dd_add(vdw, lj);
                dd_add(elec, cq);
            }
        }
    }
    block_store_dd<kTile>(vdw, red, outp + 0);
    block_store_dd<kTile>(elec, red, outp + 2);
    block_store_dd<kTile>(v14, red, outp + 4);
    block_store_dd<kTile>(e14, red, outp + 6);
}

// The receptor x ligand pairs of a complex whose partners share no excluded
// or 1-4 pair: the vdW and Coulomb terms of mm_pair_kernel for exactly those
// pairs (same expressions, the LJ table read in the complex's index order), in
// its partial layout (vdw, elec, vdw_14 = 0, elec_14 = 0 as double-doubles).
// Grid (ceil(nA / kTile), ceil(nB / kTile)); A and B hold complex atom indices.
__global__ __launch_bounds__(kTile) void mm_cross_kernel(const double *__restrict__ xyz,
                                                         PairTopo T, const int *__restrict__ A,
                                                         int nA, const int *__restrict__ Bv,
                                                         int nB, double *__restrict__ partials)
{
    __shared__ double sx[kTile], sy[kTile], sz[kTile], sq[kTile];
    __shared__ int st[kTile], sid[kTile];
    __shared__ dd red[kTile / 64];
    const int t = threadIdx.x;
    const int ia = blockIdx.x * kTile + t;
    const int j0 = blockIdx.y * kTile;
    double *outp = partials + ((size_t)blockIdx.x * gridDim.y + blockIdx.y) * 8;
    if (j0 + t < nB) {
        const int j = Bv[j0 + t];
        sid[t] = j;
        sx[t] = xyz[3 * j];
        sy[t] = xyz[3 * j + 1];
        sz[t] = xyz[3 * j + 2];
        sq[t] = T.q_amber[j];
        st[t] = T.type[j];
    }
    __syncthreads();
    dd vdw = {0, 0}, elec = {0, 0};
    if (ia < nA) {
        const int i = A[ia];
        const double xi = xyz[3 * i], yi = xyz[3 * i + 1], zi = xyz[3 * i + 2];
        const double qi = T.q_amber[i];
        const int ti = T.type[i];
        const int m = min(kTile, nB - j0);
        for (int k = 0; k < m; ++k) {
            const double dx = xi - sx[k], dy = yi - sy[k], dz = zi - sz[k];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double ir = 1.0 / sqrt(r2);
            const double ir2 = ir * ir;
            const double ir6 = ir2 * ir2 * ir2;
            // row of the lower complex index, as mm_pair_kernel
            const size_t cell = i < sid[k] ? (size_t)ti * T.n_types + st[k]
                                           : (size_t)st[k] * T.n_types + ti;
            const double LA = T.ljA[cell], LB = T.ljB[cell];
            const double lj = LA * ir6 * ir6 - LB * ir6;
            const double cq = qi * sq[k] * ir;
            dd_add(vdw, lj);
            dd_add(elec, cq);
        }
    }
    block_store_dd<kTile>(vdw, red, outp + 0);
    block_store_dd<kTile>(elec, red, outp + 2);
    if (t < 4) outp[4 + t] = 0.0;
}

int mm_cross_blocks(int nA, int nB)
{
    return nA > 0 && nB > 0 ? ceil_div(nA, kTile) * ceil_div(nB, kTile) : 0;
}

void launch_mm_cross(const double *xyz, const PairTopo &t, const int *A, int nA, const int *B,
                     int nB, double *partials, hipStream_t s)
{
    if (nA <= 0 || nB <= 0) return;
    hipLaunchKernelGGL(mm_cross_kernel, dim3(ceil_div(nA, kTile), ceil_div(nB, kTile)),
                       dim3(kTile), 0, s, xyz, t, A, nA, B, nB, partials);
    LAUNCH_OK();
}

int mm_blocks(int n)
{
    const int nb = std::max(1, ceil_div(n, kTile));
    return nb * nb;
}

void launch_mm_pairs(const double *xyz, const PairTopo &t, double *partials, hipStream_t s)
{
    hipLaunchKernelGGL(mm_pair_kernel, dim3(t.nb, t.nb), dim3(kTile), 0, s, xyz, t, partials);
    LAUNCH_OK();
}

// Bonded terms, mm_energy.rs:187-212 (bond / angle / dihedral).
__device__ inline void dcross(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

__global__ __launch_bounds__(256) void bonded_kernel(const double *__restrict__ xyz, PairTopo T,
                                                     double *__restrict__ partials)
{
    __shared__ double red[4];
    const int g = blockIdx.x * 256 + threadIdx.x;
    double eb = 0.0, ea = 0.0, ed = 0.0;
    if (g < T.n_bonds) {
        const int *b = T.bonds + 3 * g;
        const double *p = xyz + 3 * b[0], *q = xyz + 3 * b[1];
        const double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double k = T.bond_k[b[2]], r0 = T.bond_r0[b[2]];
        eb = k * (r - r0) * (r - r0);
    }
    const int ga = g - T.n_bonds;
    if (ga >= 0 && ga < T.n_angles) {
        const int *a = T.angles + 4 * ga;
        const double *A = xyz + 3 * a[0], *B = xyz + 3 * a[1], *Cc = xyz + 3 * a[2];
        const double ba[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
        const double bc[3] = {Cc[0] - B[0], Cc[1] - B[1], Cc[2] - B[2]};
        const double dot = ba[0] * bc[0] + ba[1] * bc[1] + ba[2] * bc[2];
        const double m1 = sqrt(ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2]);
        const double m2 = sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
        double ct = dot / (m1 * m2);
        ct = ct < -1.0 ? -1.0 : (ct > 1.0 ? 1.0 : ct);
        const double th = acos(ct);
        const double k = T.angle_k[a[3]], t0 = T.angle_t0[a[3]];
        ea = k * (th - t0) * (th - t0);
    }
    const int gd = g - T.n_bonds - T.n_angles;
    if (gd >= 0 && gd < T.n_dihs) {
        const int *d = T.dihs + 6 * gd;
        const double *a = xyz + 3 * d[0], *b = xyz + 3 * d[1], *c = xyz + 3 * d[2],
                     *e = xyz + 3 * d[3];
        const double b1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double b2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
        const double b3[3] = {e[0] - c[0], e[1] - c[1], e[2] - c[2]};
        double n1[3], n2[3], m1[3], u[3];
        dcross(b1, b2, n1);
        dcross(b2, b3, n2);
        const double len = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
        if (len < 1e-20) {
            u[0] = u[1] = u[2] = 0.0;
        } else {
            u[0] = b2[0] / len;
            u[1] = b2[1] / len;
            u[2] = b2[2] / len;
        }
        dcross(n1, u, m1);
        const double x = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
        const double y = m1[0] * n2[0] + m1[1] * n2[1] + m1[2] * n2[2];
        const double phi = atan2(-y, x);
        ed = T.dih_pk[d[4]] * (1.0 + cos(T.dih_pn[d[4]] * phi - T.dih_ph[d[4]]));
    }
    block_store<256>(eb, red, partials + (size_t)blockIdx.x * 3 + 0);
    block_store<256>(ea, red, partials + (size_t)blockIdx.x * 3 + 1);
    block_store<256>(ed, red, partials + (size_t)blockIdx.x * 3 + 2);
}

int bonded_blocks(const PairTopo &t)
{
    return std::max(1, ceil_div((int64_t)t.n_bonds + t.n_angles + t.n_dihs, 256));
}

void launch_bonded(const double *xyz, const PairTopo &t, double *partials, hipStream_t s)
{
    hipLaunchKernelGGL(bonded_kernel, dim3(bonded_blocks(t)), dim3(256), 0, s, xyz, t, partials);
    LAUNCH_OK();
}

// Fixed-order sums of block partials (one block, 256 lanes).
__global__ __launch_bounds__(256) void sum_partials_kernel(const double *__restrict__ p, int nbk,
                                                           int stride, int nterm,
                                                           double *__restrict__ out)
{
    __shared__ double red[4];
    for (int term = 0; term < nterm; ++term) {
        double a = 0.0;
        for (int b = threadIdx.x; b < nbk; b += 256) a += p[(size_t)b * stride + term];
        block_store<256>(a, red, out + term);
    }
}

__global__ __launch_bounds__(256) void sum_partials_dd_kernel(const double *__restrict__ p,
                                                              int nbk, int nterm,
                                                              double *__restrict__ out)
{
    __shared__ dd red[4];
    __shared__ double tmp[2];
    for (int term = 0; term < nterm; ++term) {
        dd a = {0.0, 0.0};
        for (int b = threadIdx.x; b < nbk; b += 256) {
            dd_add(a, p[(size_t)b * 2 * nterm + 2 * term]);
            dd_add(a, p[(size_t)b * 2 * nterm + 2 * term + 1]);
        }
        block_store_dd<256>(a, red, tmp);
        if (threadIdx.x == 0) out[term] = tmp[0] + tmp[1];
        __syncthreads();
    }
}

void launch_sum_partials(const double *partials, int n_blocks, int stride, int n_terms,
                         double *out, hipStream_t s)
{
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(256), 0, s, partials, n_blocks, stride,
                       n_terms, out);
    LAUNCH_OK();
}

void launch_sum_partials_dd(const double *partials, int n_blocks, int n_terms, double *out,
                            hipStream_t s)
{
    hipLaunchKernelGGL(sum_partials_dd_kernel, dim3(1), dim3(256), 0, s, partials, n_blocks,
                       n_terms, out);
    LAUNCH_OK();
}

// ---------------------------------------------------------------------------
// Boundary node enumeration in the reference's face order (pb_solver.rs:31-60).
__host__ __device__ inline int64_t bnd_count(int nx, int ny, int nz)
{
    const int64_t zi = nz > 2 ? nz - 2 : 0, yi = ny > 2 ? ny - 2 : 0;
    return 2LL * nx * ny + 2LL * nx * zi + 2LL * yi * zi;
}

__device__ inline void bnd_node(int nx, int ny, int nz, int64_t b, int &ix, int &iy, int &iz)
{
    const int64_t f0 = 2LL * nx * ny;
    if (b < f0) {
        const int s = (int)(b / ((int64_t)nx * ny));
        const int64_t r = b % ((int64_t)nx * ny);
        iz = s ? nz - 1 : 0;
        iy = (int)(r / nx);
        ix = (int)(r % nx);
        return;
    }
    b -= f0;
    const int64_t zi = nz > 2 ? nz - 2 : 0;
    const int64_t f1 = 2LL * nx * zi;
    if (b < f1) {
        iz = 1 + (int)(b / (2LL * nx));
        const int64_t r = b % (2LL * nx);
        iy = (r / nx) ? ny - 1 : 0;
        ix = (int)(r % nx);
        return;
    }
    b -= f1;
    const int64_t yi = ny > 2 ? ny - 2 : 0;
    iz = 1 + (int)(b / (2 * yi));
    const int64_t r = b % (2 * yi);
    iy = 1 + (int)(r / 2);
    ix = (r % 2) ? nx - 1 : 0;
}

int64_t boundary_count(const GridDesc &g) { return bnd_count(g.nx, g.ny, g.nz); }

// set_dh_boundary, pb_solver.rs:68-93, for one or two media at once: sum A
// = sum_i q_i exp(-kappa r) / r (kappa > 0) and sum B = sum_i q_i / r share
// each pair's distance.  Positions relative to the grid origin in FP32 (|dx|
// error ~1e-5 A at r >= the grid buffer), terms FP32 (one rsqrt per pair),
// sums over 256 atoms FP32 then FP64.  Grid (boundary blocks, atom chunks):
// each block sums one atom chunk into work[chunk][node] (A) and
// work[chunks + chunk][node] (B); dh_reduce_kernel adds the chunks in order
// and applies 332.0522 / eps_out in FP64.
template <bool SALT, bool PAIR>
__global__ __launch_bounds__(256) void dh_boundary_kernel(GridDesc g, int n, int chunk, int chunks,
                                                          const double *__restrict__ xyz,
                                                          const double *__restrict__ q,
                                                          float kappa, double *__restrict__ work)
{
    __shared__ float4 sa[256];
    const int64_t nbnd = bnd_count(g.nx, g.ny, g.nz);
    const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int ix = 0, iy = 0, iz = 0;
    const bool act = b < nbnd;
    if (act) bnd_node(g.nx, g.ny, g.nz, b, ix, iy, iz);
    const float px = (float)((double)ix * g.h), py = (float)((double)iy * g.h),
                pz = (float)((double)iz * g.h);
    const int a_lo = blockIdx.y * chunk, a_hi = min(n, a_lo + chunk);
    double acc_a = 0.0, acc_b = 0.0;
    for (int a0 = a_lo; a0 < a_hi; a0 += 256) {
        const int a = a0 + threadIdx.x;
        if (a < a_hi)
            sa[threadIdx.x] = make_float4((float)(xyz[3 * a] - g.ox), (float)(xyz[3 * a + 1] - g.oy),
                                          (float)(xyz[3 * a + 2] - g.oz), (float)q[a]);
        __syncthreads();
        const int m = min(256, a_hi - a0);
        float pa = 0.0f, pb = 0.0f;
        if (act) {
            for (int k = 0; k < m; ++k) {
                const float4 A = sa[k];
                const float dx = px - A.x, dy = py - A.y, dz = pz - A.z;
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const bool ok = r2 > 1e-20f;
                const float ir = __builtin_amdgcn_rsqf(r2);
                const float t = A.w * ir;  // q / r
                if (SALT) {
                    const float e = __expf(-kappa * (r2 * ir));
                    pa = ok ? fmaf(t, e, pa) : pa;
                }
                if (!SALT || PAIR) pb = ok ? pb + t : pb;
            }
            acc_a += (double)pa;
            acc_b += (double)pb;
        }
        __syncthreads();
    }
    if (act) {
        if (SALT) work[(size_t)blockIdx.y * nbnd + b] = acc_a;
        if (!SALT || PAIR) work[(size_t)(blockIdx.y + (SALT ? chunks : 0)) * nbnd + b] = acc_b;
    }
}

// phi_x[node] = scale_x * (sum of the chunks of sum x), x = A (work rows
// 0 .. chunks - 1) and, with phi_b, B (rows chunks .. 2 chunks - 1)
__global__ void dh_reduce_kernel(GridDesc g, int chunks, const double *__restrict__ work,
                                 double scale_a, double *__restrict__ phi_a, double scale_b,
                                 double *__restrict__ phi_b)
{
    const int64_t nbnd = bnd_count(g.nx, g.ny, g.nz);
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbnd) return;
    int ix, iy, iz;
    bnd_node(g.nx, g.ny, g.nz, b, ix, iy, iz);
    const int64_t idx = (int64_t)ix + (int64_t)g.nx * (iy + (int64_t)g.ny * iz);
    double s = 0.0;
    for (int c = 0; c < chunks; ++c) s += work[(size_t)c * nbnd + b];
    phi_a[idx] = scale_a * s;
    if (phi_b) {
        double t = 0.0;
        for (int c = 0; c < chunks; ++c) t += work[(size_t)(chunks + c) * nbnd + b];
        phi_b[idx] = scale_b * t;
    }
}

static int dh_chunks(const GridDesc &g, int n)
{
    const int64_t blocks = std::max<int64_t>(1, ceil_div(boundary_count(g), 256));
    const int want = (int)std::max<int64_t>(1, ceil_div(4096, blocks));
    return std::max(1, std::min({want, 64, ceil_div(std::max(n, 1), 256)}));
}

size_t dh_work_size(const GridDesc &g, int n)
{
    return 2 * (size_t)dh_chunks(g, n) * (size_t)std::max<int64_t>(1, boundary_count(g));
}

void launch_dh_boundary(const GridDesc &g, int n, const double *xyz, const double *q,
                        double kappa, double eps_a, double *phi_a, double eps_b, double *phi_b,
                        double *work, hipStream_t s)
{
    const int64_t nb = boundary_count(g);
    if (nb == 0) return;
    const int chunks = dh_chunks(g, n);
    const int chunk = ceil_div(ceil_div(std::max(n, 1), chunks), 256) * 256;
    const int used = std::max(1, ceil_div(std::max(n, 1), chunk));
    const dim3 grid(ceil_div(nb, 256), used);
    const double sa = kCoulomb / eps_a, sb = kCoulomb / eps_b;
    if (kappa > 0.0) {
        if (phi_b)
            hipLaunchKernelGGL((dh_boundary_kernel<true, true>), grid, dim3(256), 0, s, g, n, chunk,
                               used, xyz, q, (float)kappa, work);
        else
            hipLaunchKernelGGL((dh_boundary_kernel<true, false>), grid, dim3(256), 0, s, g, n,
                               chunk, used, xyz, q, (float)kappa, work);
        LAUNCH_OK();
        hipLaunchKernelGGL(dh_reduce_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, g, used,
                           work, sa, phi_a, sb, phi_b);
    } else {
        // without salt both media take the same sum
        hipLaunchKernelGGL((dh_boundary_kernel<false, false>), grid, dim3(256), 0, s, g, n, chunk,
                           used, xyz, q, 0.0f, work);
        LAUNCH_OK();
        hipLaunchKernelGGL(dh_reduce_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, g, used,
                           work, sa, phi_a, 0.0, nullptr);
        LAUNCH_OK();
        if (phi_b)
            hipLaunchKernelGGL(dh_reduce_kernel, dim3(ceil_div(nb, 256)), dim3(256), 0, s, g, used,
                               work, sb, phi_b, 0.0, nullptr);
    }
    LAUNCH_OK();
}

// ---------------------------------------------------------------------------
// Decomposition cross pairs, decomposition.rs:240-283.  Lane = atom of set A,
// column tiles over set B (chunked, partials per chunk, fixed-order reduce).
// Pair (ri, li) = (A, B) when a_is_rec, else (B, A).
__global__ __launch_bounds__(kTile) void decomp_kernel(const double *__restrict__ xyz, PairTopo T,
                                                       const int *__restrict__ A, int nA,
                                                       const int *__restrict__ Bv, int nB,
                                                       int chunk, int a_is_rec, int cross_excl,
                                                       double *__restrict__ work)
{
    __shared__ double sx[kTile], sy[kTile], sz[kTile], sq[kTile];
    __shared__ int st[kTile], sid[kTile];
    const int t = threadIdx.x;
    const int ia = blockIdx.x * kTile + t;
    const int b0 = blockIdx.y * chunk, b1 = min(nB, b0 + chunk);
    const bool act = ia < nA;
    int a = 0;
    double xa = 0, ya = 0, za = 0, qa = 0;
    int ta = 0;
    if (act) {
        a = A[ia];
        xa = xyz[3 * a];
        ya = xyz[3 * a + 1];
        za = xyz[3 * a + 2];
        qa = T.q_amber[a];
        ta = T.type[a];
    }
    dd vv = {0, 0}, ee = {0, 0};
    for (int bt = b0; bt < b1; bt += kTile) {
        if (bt + t < b1) {
            const int b = Bv[bt + t];
            sid[t] = b;
            sx[t] = xyz[3 * b];
            sy[t] = xyz[3 * b + 1];
            sz[t] = xyz[3 * b + 2];
            sq[t] = T.q_amber[b];
            st[t] = T.type[b];
        }
        __syncthreads();
        const int m = min(kTile, b1 - bt);
        if (act) {
            for (int k = 0; k < m; ++k) {
                const double dx = a_is_rec ? xa - sx[k] : sx[k] - xa;
                const double dy = a_is_rec ? ya - sy[k] : sy[k] - ya;
                const double dz = a_is_rec ? za - sz[k] : sz[k] - za;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r < 1e-10) continue;
                bool is14 = false, isex = false;
                if (cross_excl) {
                    const int b = sid[k];
                    const int lo = a < b ? a : b, hi = a < b ? b : a;
                    if (lo != hi) {
                        for (int p = T.p14_ptr[lo]; p < T.p14_ptr[lo + 1]; ++p)
                            if (T.p14_col[p] == hi) is14 = true;
                        for (int p = T.ex_ptr[lo]; p < T.ex_ptr[lo + 1]; ++p)
                            if (T.ex_col[p] == hi) isex = true;
                    }
                }
                if (!is14 && isex) continue;
                const double r2 = r * r;
                const double r6 = r2 * r2 * r2;
                const double r12 = r6 * r6;
                const int tr = a_is_rec ? ta : st[k], tl = a_is_rec ? st[k] : ta;
                const double LA = T.ljA[(size_t)tr * T.n_types + tl];
                const double LB = T.ljB[(size_t)tr * T.n_types + tl];
                const double qr = a_is_rec ? qa : sq[k], ql = a_is_rec ? sq[k] : qa;
                double v, e;
                if (is14) {
                    v = (LA / r12 - LB / r6) / T.scnb;
                    e = qr * ql / r / T.scee;
                } else {
                    v = LA / r12 - LB / r6;
                    e = qr * ql / r;
                }
                dd_add(vv, v * 0.5);
                dd_add(ee, e * 0.5);
            }
        }
        __syncthreads();
    }
    if (act) {
        double *w = work + ((size_t)blockIdx.y * nA + ia) * 4;
        w[0] = vv.s;
        w[1] = vv.c;
        w[2] = ee.s;
        w[3] = ee.c;
    }
}

__global__ void decomp_reduce_kernel(const double *__restrict__ work, int chunks, int nA,
                                     double *__restrict__ ov, double *__restrict__ oe)
{
    const int ia = blockIdx.x * blockDim.x + threadIdx.x;
    if (ia >= nA) return;
    dd v = {0, 0}, e = {0, 0};
    for (int c = 0; c < chunks; ++c) {
        const double *w = work + ((size_t)c * nA + ia) * 4;
        dd_add(v, w[0]);
        dd_add(v, w[1]);
        dd_add(e, w[2]);
        dd_add(e, w[3]);
    }
    ov[ia] = dd_val(v);
    oe[ia] = dd_val(e);
}

int decomp_chunks(int nB) { return std::max(1, std::min(64, ceil_div(nB, 2048))); }

void launch_decomp_pairs(const double *xyz, const PairTopo &t, const int *A, int nA,
                         const int *B, int nB, int a_is_rec, int cross_excl, double *work,
                         double *out_vdw, double *out_elec, hipStream_t s)
{
    if (nA == 0) return;
    const int chunks = decomp_chunks(nB);
    const int chunk = ceil_div(ceil_div(std::max(nB, 1), chunks), kTile) * kTile;
    const int used = std::max(1, ceil_div(std::max(nB, 1), chunk));
    hipLaunchKernelGGL(decomp_kernel, dim3(ceil_div(nA, kTile), used), dim3(kTile), 0, s, xyz, t,
                       A, nA, B, nB, chunk, a_is_rec, cross_excl, work);
    LAUNCH_OK();
    hipLaunchKernelGGL(decomp_reduce_kernel, dim3(ceil_div(nA, 256)), dim3(256), 0, s, work, used,
                       nA, out_vdw, out_elec);
    LAUNCH_OK();
}

// ---------------------------------------------------------------------------
// Shrake-Rupley SASA, core/src/sasa.rs:225-276: one wavefront per atom.  A
// neighbour is kept when it lies within ext_i + max_ext (the reference's
// KD-tree radius); a test point is accessible when d^2 > (r_j + p)^2 for every
// neighbour.  All FP64, no contraction: the count is bit-exact.
constexpr int kSasaCap = 192;

__global__ __launch_bounds__(256) void sasa_kernel(const double *__restrict__ xyz,
                                                   const double *__restrict__ radii, int n,
                                                   double probe, const double *__restrict__ pts,
                                                   int npts, CellList cl, double max_ext,
                                                   double *__restrict__ out, int *overflow)
{
    __shared__ double nx_[4][kSasaCap], ny_[4][kSasaCap], nz_[4][kSasaCap], no_[4][kSasaCap];
    __shared__ int cnt[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + w;
    if (lane == 0) cnt[w] = 0;
    __syncthreads();
    double xi = 0, yi = 0, zi = 0, ext = 0;
    if (i < n) {
        xi = xyz[3 * i];
        yi = xyz[3 * i + 1];
        zi = xyz[3 * i + 2];
        ext = radii[i] + probe;
        const double search = ext + max_ext;
        const double s2 = search * search;
        const int cx0 = max(0, (int)floor((xi - search - cl.ox) * cl.inv));
        const int cx1 = min(cl.dx - 1, (int)floor((xi + search - cl.ox) * cl.inv));
        const int cy0 = max(0, (int)floor((yi - search - cl.oy) * cl.inv));
        const int cy1 = min(cl.dy - 1, (int)floor((yi + search - cl.oy) * cl.inv));
        const int cz0 = max(0, (int)floor((zi - search - cl.oz) * cl.inv));
        const int cz1 = min(cl.dz - 1, (int)floor((zi + search - cl.oz) * cl.inv));
        const int ncx = cx1 - cx0 + 1, ncy = cy1 - cy0 + 1, ncz = cz1 - cz0 + 1;
        const int ncell = (ncx > 0 && ncy > 0 && ncz > 0) ? ncx * ncy * ncz : 0;
        for (int c = lane; c < ncell; c += 64) {
            const int cx = cx0 + c % ncx, cy = cy0 + (c / ncx) % ncy, cz = cz0 + c / (ncx * ncy);
            const int cid = cx + cl.dx * (cy + cl.dy * cz);
            for (int k = cl.start[cid]; k < cl.start[cid + 1]; ++k) {
                const int j = cl.atoms[k];
                if (j == i) continue;
                const double dx = xyz[3 * j] - xi, dy = xyz[3 * j + 1] - yi, dz = xyz[3 * j + 2] - zi;
                if (dx * dx + dy * dy + dz * dz > s2) continue;
                const int slot = atomicAdd(&cnt[w], 1);
                if (slot < kSasaCap) {
                    nx_[w][slot] = xyz[3 * j];
                    ny_[w][slot] = xyz[3 * j + 1];
                    nz_[w][slot] = xyz[3 * j + 2];
                    const double oe = radii[j] + probe;
                    no_[w][slot] = oe * oe;
                }
            }
        }
    }
    __syncthreads();
    int m = cnt[w];
    if (m > kSasaCap) {
        if (lane == 0) atomicExch(overflow, 1);
        m = kSasaCap;
    }
    int acc = 0;
    if (i < n) {
        for (int p = lane; p < npts; p += 64) {
            const double tx = xi + pts[3 * p] * ext, ty = yi + pts[3 * p + 1] * ext,
                         tz = zi + pts[3 * p + 2] * ext;
            bool ok = true;
            for (int k = 0; k < m && ok; ++k) {
                const double dx = tx - nx_[w][k], dy = ty - ny_[w][k], dz = tz - nz_[w][k];
                const double d2 = dx * dx + dy * dy + dz * dz;
                if (!(d2 > no_[w][k])) ok = false;
            }
            acc += ok ? 1 : 0;
        }
    }
    acc = wave_sum(acc);
    if (i < n && lane == 0) {
        const double frac = (double)acc / (double)npts;
        const double area = 4.0 * kPi * (ext * ext);
        out[i] = frac * area;
    }
}

void launch_sasa(const double *xyz, const double *radii, int n, double probe, const double *pts,
                 int n_pts, const CellList &cl, double max_ext, double *per_atom, int *overflow,
                 hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL(sasa_kernel, dim3(ceil_div(n, 4)), dim3(256), 0, s, xyz, radii, n, probe,
                       pts, n_pts, cl, max_ext, per_atom, overflow);
    LAUNCH_OK();
}

// ---- statistics of a served request (binding.rs:704-743, entropy.rs:53-78) ----
// One workgroup over n frames' (dMM, dPolar, dSA, dTotal): means, sample
// standard deviations about them (two passes, as the reference), the SEM of
// dTotal and the interaction entropy -T dS = kT (m + ln(<exp(x - m)>)) of x =
// (dMM - <dMM>) / kT, m = max x (log-sum-exp).  Fixed-order tree reductions, so
// the result is deterministic; it equals the sequential host sums to rounding.
// out: mean mm, polar, sa, total; std total, mm, polar, sa; sem total;
// -T dS (NaN below two frames).
__device__ void block_sum4(double (&v)[4], double (*red)[4])
{
    const int t = threadIdx.x;
    for (int k = 0; k < 4; ++k) red[t][k] = v[k];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (t < off)
            for (int k = 0; k < 4; ++k) red[t][k] += red[t + off][k];
        __syncthreads();
    }
    for (int k = 0; k < 4; ++k) v[k] = red[0][k];
    __syncthreads();
}

__global__ __launch_bounds__(256) void frame_stats_kernel(const double *__restrict__ d, int n,
                                                          double kt, double *__restrict__ out)
{
    __shared__ double red[256][4];
    const int t = threadIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int f = t; f < n; f += blockDim.x)
        for (int k = 0; k < 4; ++k) s[k] += d[4 * f + k];
    block_sum4(s, red);
    double mu[4];
    for (int k = 0; k < 4; ++k) mu[k] = s[k] / (double)n;
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    for (int f = t; f < n; f += blockDim.x)
        for (int k = 0; k < 4; ++k) {
            const double e = d[4 * f + k] - mu[k];
            q[k] += e * e;
        }
    block_sum4(q, red);
    // entropy: the largest x, then the sum of exp(x - m)
    double m[4] = {-INFINITY, 0.0, 0.0, 0.0};
    for (int f = t; f < n; f += blockDim.x) m[0] = fmax(m[0], (d[4 * f] - mu[0]) / kt);
    red[t][0] = m[0];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (t < off) red[t][0] = fmax(red[t][0], red[t + off][0]);
        __syncthreads();
    }
    const double mx = red[0][0];
    __syncthreads();
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int f = t; f < n; f += blockDim.x) a[0] += exp((d[4 * f] - mu[0]) / kt - mx);
    block_sum4(a, red);
    if (t != 0) return;
    const double sd[4] = {n > 1 ? sqrt(q[0] / (double)(n - 1)) : 0.0,
                          n > 1 ? sqrt(q[1] / (double)(n - 1)) : 0.0,
                          n > 1 ? sqrt(q[2] / (double)(n - 1)) : 0.0,
                          n > 1 ? sqrt(q[3] / (double)(n - 1)) : 0.0};
    out[0] = mu[0];
    out[1] = mu[1];
    out[2] = mu[2];
    out[3] = mu[3];
    out[4] = sd[3];
    out[5] = sd[0];
    out[6] = sd[1];
    out[7] = sd[2];
    out[8] = sd[3] / sqrt((double)n);
    out[9] = n >= 2 ? kt * (mx + log(a[0] / (double)n)) : NAN;
}

void launch_frame_stats(const double *terms, int n, double kt, double *out10, hipStream_t s)
{
    hipLaunchKernelGGL(frame_stats_kernel, dim3(1), dim3(256), 0, s, terms, n, kt, out10);
    LAUNCH_OK();
}

}  // namespace mmpbsa
