// grid.hip -- Poisson-Boltzmann grid assembly on gfx950: cell lists, charge
// spreading, SAS / ion-exclusion maps, SES classification, face dielectrics,
// boundary interpolation and the energy gather.
//
// Geometry tests are FP64 with no contraction, in the reference's operation
// order, so boolean maps and face values match the CPU bit for bit.
#include <hip/hip_runtime.h>

#include <cfloat>

#include <algorithm>
#include <cmath>

#include "kernels.h"

namespace mmpbsa {

__device__ __forceinline__ int64_t gi(const GridDesc &g, int x, int y, int z)
{
    return (int64_t)x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
}

__device__ __forceinline__ void node_of(const GridDesc &g, int64_t idx, int &x, int &y, int &z)
{
    const int64_t nxy = (int64_t)g.nx * g.ny;
    z = (int)(idx / nxy);
    const int64_t r = idx - (int64_t)z * nxy;
    y = (int)(r / g.nx);
    x = (int)(r - (int64_t)y * g.nx);
}

// ---- cell list -------------------------------------------------------------
__global__ void cell_count_kernel(const double *__restrict__ xyz, int n, double ox, double oy,
                                  double oz, double inv, int dx, int dy, int dz,
                                  int *__restrict__ cell_of, int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cx = (int)floor((xyz[3 * i] - ox) * inv);
    int cy = (int)floor((xyz[3 * i + 1] - oy) * inv);
    int cz = (int)floor((xyz[3 * i + 2] - oz) * inv);
    cx = min(max(cx, 0), dx - 1);
    cy = min(max(cy, 0), dy - 1);
    cz = min(max(cz, 0), dz - 1);
    const int c = cx + dx * (cy + dy * cz);
    cell_of[i] = c;
    atomicAdd(&count[c], 1);
}

// Exclusive scan of count[0..nc) into start[0..nc]; one block of 1024.
__global__ __launch_bounds__(1024) void scan_kernel(const int *__restrict__ count, int nc,
                                                    int *__restrict__ start)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nc + 1023) / 1024;
    const int a = t * per, b = min(nc, a + per);
    int s = 0;
    for (int k = a; k < b; ++k) s += count[k];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? part[t - 1] : 0;
    for (int k = a; k < b; ++k) {
        start[k] = run;
        run += count[k];
    }
    if (t == 1023) start[nc] = part[1023];
}

__global__ void cell_fill_kernel(const int *__restrict__ cell_of, int n,
                                 const int *__restrict__ start, int *__restrict__ fill,
                                 int *__restrict__ atoms)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell_of[i];
    const int p = atomicAdd(&fill[c], 1);
    atoms[start[c] + p] = i;
}

void launch_cell_build(const double *xyz, int n, double ox, double oy, double oz, double inv,
                       int dx, int dy, int dz, int *cell_of, int *start, int *fill, int *atoms,
                       hipStream_t s)
{
    const int64_t nc = (int64_t)dx * dy * dz;
    launch_fill_i32(fill, 0, nc, s);
    if (n > 0) {
        hipLaunchKernelGGL(cell_count_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, xyz, n, ox,
                           oy, oz, inv, dx, dy, dz, cell_of, fill);
        LAUNCH_OK();
    }
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, s, fill, (int)nc, start);
    LAUNCH_OK();
    launch_fill_i32(fill, 0, nc, s);
    if (n > 0) {
        hipLaunchKernelGGL(cell_fill_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, cell_of, n,
                           start, fill, atoms);
        LAUNCH_OK();
    }
}

// ---- charge spreading, map_charges (pb_grid.rs:220-292) ---------------------
// Contributions are rounded to 2^-40 fixed point and summed with integer
// atomics: the map is independent of atom scheduling.
constexpr double kFix = 1099511627776.0;  // 2^40

__global__ void spread_kernel(GridDesc g, int n, const double *__restrict__ xyz,
                              const double *__restrict__ q, double vol,
                              unsigned long long *__restrict__ acc)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const double fx = (xyz[3 * a] - g.ox) / g.h;
    const double fy = (xyz[3 * a + 1] - g.oy) / g.h;
    const double fz = (xyz[3 * a + 2] - g.oz) / g.h;
    const double flx = floor(fx), fly = floor(fy), flz = floor(fz);
    if (flx < 0.0 || flx >= (double)(g.nx - 1) || fly < 0.0 || fly >= (double)(g.ny - 1) ||
        flz < 0.0 || flz >= (double)(g.nz - 1))
        return;
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    const double dx = fx - (double)ix, dy = fy - (double)iy, dz = fz - (double)iz;
    const double qd = q[a] / vol;
    const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const double w = qd * wx[i] * wy[j] * wz[k];
                const long long f = llrint(w * kFix);
                atomicAdd(&acc[gi(g, ix + i, iy + j, iz + k)], (unsigned long long)f);
            }
}

__global__ void fixed_to_f64_kernel(const unsigned long long *__restrict__ acc, int64_t n,
                                    double *__restrict__ rho)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rho[i] = (double)(long long)acc[i] * (1.0 / kFix);
}

void launch_spread(const GridDesc &g, int n, const double *xyz, const double *q,
                   unsigned long long *acc, double *rho, hipStream_t s)
{
    const int64_t N = g.size();
    launch_fill_u64(acc, 0ull, N, s);
    if (n > 0) {
        hipLaunchKernelGGL(spread_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, g, n, xyz, q,
                           g.h * g.h * g.h, acc);
        LAUNCH_OK();
    }
    hipLaunchKernelGGL(fixed_to_f64_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, acc, N, rho);
    LAUNCH_OK();
}

// ---- inside tests (AtomSpatialHash::point_inside, pb_grid.rs:400-442) -------
// inside[i] = exists atom a with |p_i - x_a|^2 < (r_a + probe)^2, built by
// atom: a workgroup per atom marks the nodes of its inflated sphere.  Each
// node-atom test is the per-node search's own expression (p_i from the node
// index, d^2 summed x, y, z, threshold (r + probe)^2), and marking is an
// idempotent store of 1, so the map does not depend on the order.  Most
// nodes lie farther than r + probe from every atom; by atom, no work is spent
// on them (a search by node visits every cell within reach).
__global__ __launch_bounds__(256) void inside_scatter_kernel(GridDesc g, const double *__restrict__ xyz,
                                                             const double *__restrict__ radii,
                                                             double pa, double pb,
                                                             uint8_t *__restrict__ out_a,
                                                             uint8_t *__restrict__ out_b)
{
    const int a = blockIdx.x;
    const double ax = xyz[3 * a], ay = xyz[3 * a + 1], az = xyz[3 * a + 2];
    const double ta = radii[a] + pa, tb = radii[a] + pb;
    const double R = out_b ? fmax(ta, tb) : ta;
    // node ranges covering the sphere with a node to spare on each side (the
    // exact test decides), clamped before the integer conversion
    auto lo = [&](double v, int n) { return (int)fmin(fmax(floor(v) - 1.0, 0.0), (double)n); };
    auto hi = [&](double v, int n) { return (int)fmax(fmin(ceil(v) + 1.0, (double)(n - 1)), -1.0); };
    const int x0 = lo((ax - R - g.ox) / g.h, g.nx), x1 = hi((ax + R - g.ox) / g.h, g.nx);
    const int y0 = lo((ay - R - g.oy) / g.h, g.ny), y1 = hi((ay + R - g.oy) / g.h, g.ny);
    const int z0 = lo((az - R - g.oz) / g.h, g.nz), z1 = hi((az + R - g.oz) / g.h, g.nz);
    if (x0 > x1 || y0 > y1 || z0 > z1) return;
    const int wx = x1 - x0 + 1, wy = y1 - y0 + 1;
    const int cnt = wx * wy * (z1 - z0 + 1);
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const int dz_ = k / (wx * wy), r_ = k - dz_ * (wx * wy);
        const int dy_ = r_ / wx, x = x0 + (r_ - dy_ * wx), y = y0 + dy_, z = z0 + dz_;
        const double px = g.ox + (double)x * g.h, py = g.oy + (double)y * g.h,
                     pz = g.oz + (double)z * g.h;
        const double dx = px - ax, dy = py - ay, dz = pz - az;
        const double d2 = dx * dx + dy * dy + dz * dz;
        const int64_t idx = (int64_t)x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
        if (d2 < ta * ta) out_a[idx] = 1;
        if (out_b && d2 < tb * tb) out_b[idx] = 1;
    }
}

void launch_inside(const GridDesc &g, int n, const double *xyz, const double *radii, double probe,
                   uint8_t *inside, hipStream_t s)
{
    launch_fill_u8(inside, 0, g.size(), s);
    if (n <= 0) return;
    hipLaunchKernelGGL(inside_scatter_kernel, dim3(n), dim3(256), 0, s, g, xyz, radii, probe,
                       probe, inside, (uint8_t *)nullptr);
    LAUNCH_OK();
}

void launch_inside2(const GridDesc &g, int n, const double *xyz, const double *radii, double pa,
                    double pb, uint8_t *out_a, uint8_t *out_b, hipStream_t s)
{
    launch_fill_u8(out_a, 0, g.size(), s);
    launch_fill_u8(out_b, 0, g.size(), s);
    if (n <= 0) return;
    hipLaunchKernelGGL(inside_scatter_kernel, dim3(n), dim3(256), 0, s, g, xyz, radii, pa, pb,
                       out_a, out_b);
    LAUNCH_OK();
}

// ---- SES: flood fill of the exterior + bounded squared distance -----------
// compute_inside_ses, pb_grid.rs:576-708.  Connected exterior = outside-SAS
// nodes reachable from the box faces through outside-SAS nodes.  Line sweeps
// propagate along maximal runs; rounds repeat until nothing changes.
__global__ void ext_init_kernel(GridDesc g, const uint8_t *__restrict__ sas, uint8_t *__restrict__ ext)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.size()) return;
    int x, y, z;
    node_of(g, idx, x, y, z);
    const bool bnd = x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 || z == g.nz - 1;
    ext[idx] = (bnd && !sas[idx]) ? 1 : 0;
}

// Bounded separable squared EDT (values capped at T+1) and the SES test
// inside <=> min squared index distance to the exterior > T, T = floor((p/h)^2).
__global__ void dil_x_kernel(GridDesc g, int R, int T, const uint8_t *__restrict__ ext,
                             uint8_t *__restrict__ out)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.size()) return;
    int x, y, z;
    node_of(g, idx, x, y, z);
    int best = T + 1;
    for (int d = -R; d <= R; ++d) {
        const int xx = x + d;
        if (xx < 0 || xx >= g.nx) continue;
        if (ext[idx + d]) best = min(best, d * d);
    }
    out[idx] = (uint8_t)best;
}

__global__ void dil_yz_kernel(GridDesc g, int R, int T, int axis, const uint8_t *__restrict__ in,
                              uint8_t *__restrict__ out, int final_test)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.size()) return;
    int x, y, z;
    node_of(g, idx, x, y, z);
    const int pos = axis == 1 ? y : z;
    const int len = axis == 1 ? g.ny : g.nz;
    const int64_t st = axis == 1 ? (int64_t)g.nx : (int64_t)g.nx * g.ny;
    int best = T + 1;
    for (int d = -R; d <= R; ++d) {
        const int p = pos + d;
        if (p < 0 || p >= len) continue;
        best = min(best, (int)in[idx + d * st] + d * d);
    }
    best = min(best, T + 1);
    out[idx] = final_test ? (uint8_t)(best > T ? 1 : 0) : (uint8_t)best;
}

// Forward fill on a 32-node chunk of a line: bits of F (exterior, or the
// carry-in at bit 0) spread upwards through runs of open (outside-SAS)
// nodes, as the node-by-node sweep would.  Log-step prefix: after the step
// with shift s, F holds every open node at most 2 s - 1 open nodes above a
// seed, and M the nodes whose s lower neighbours are all open.
__device__ __forceinline__ uint32_t fill_up(uint32_t F, uint32_t open)
{
    uint32_t M = open;
#pragma unroll
    for (int s = 1; s < 32; s <<= 1) {
        F |= M & (F << s);
        M &= M << s;
    }
    return F;
}
__device__ __forceinline__ uint32_t fill_down(uint32_t F, uint32_t open)
{
    uint32_t M = open;
#pragma unroll
    for (int s = 1; s < 32; s <<= 1) {
        F |= M & (F >> s);
        M &= M >> s;
    }
    return F;
}

// One thread per line along `axis` of map blockIdx.y of the table: forward
// then backward propagation through outside-SAS nodes, 32 nodes at a time
// (the chunk's loads are independent; the sweep itself runs on bit masks).
// After both passes every open run of the line that holds an exterior node
// is exterior, as after the node-by-node sweeps.
__global__ void flood_batch_kernel(const FloodJob *__restrict__ jobs, int axis)
{
    const FloodJob J = jobs[blockIdx.y];
    if (*J.active == 0) return;  // this map reached its fixpoint in an earlier round
    const GridDesc g = J.g;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t base, stride;
    int len;
    if (axis == 0) {
        if (t >= (int64_t)g.ny * g.nz) return;
        base = t * g.nx;
        stride = 1;
        len = g.nx;
    } else if (axis == 1) {
        if (t >= (int64_t)g.nx * g.nz) return;
        const int64_t x = t % g.nx, z = t / g.nx;
        base = x + (int64_t)g.nx * g.ny * z;
        stride = g.nx;
        len = g.ny;
    } else {
        if (t >= (int64_t)g.nx * g.ny) return;
        base = t;
        stride = (int64_t)g.nx * g.ny;
        len = g.nz;
    }
    const auto sas = to_global<const uint8_t>(J.sas);
    const auto ext = to_global<uint8_t>(J.ext);
    const int nch = (len + 31) / 32;
    // all 64 loads of a chunk are issued before the first use (nodes past the
    // line's end re-read its last node and are masked off)
    auto load = [&](int c, uint32_t &open, uint32_t &E) {
        const int k0 = 32 * c, n = min(32, len - k0);
        uint8_t sv[32], ev[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int64_t i = base + (int64_t)(k0 + min(k, n - 1)) * stride;
            sv[k] = sas[i];
            ev[k] = ext[i];
        }
        open = 0u;
        E = 0u;
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            open |= (uint32_t)(sv[k] == 0) << k;
            E |= (uint32_t)(ev[k] != 0) << k;
        }
        const uint32_t live = n == 32 ? ~0u : (1u << n) - 1u;
        open &= live;
        E &= live;
    };
    auto store = [&](int c, uint32_t nw) {
        const int k0 = 32 * c;
        while (nw) {
            const int k = __ffs(nw) - 1;
            ext[base + (int64_t)(k0 + k) * stride] = 1;
            nw &= nw - 1u;
        }
    };
    uint32_t ch = 0u;
    uint32_t carry = 0u;  // the previous chunk's last node is open and exterior
    for (int c = 0; c < nch; ++c) {
        uint32_t open, E;
        load(c, open, E);
        const int n = min(32, len - 32 * c);
        const uint32_t F = fill_up((E | carry) & open, open);
        store(c, F & ~E);
        ch |= F & ~E;
        carry = (F >> (n - 1)) & 1u;
    }
    carry = 0u;
    for (int c = nch - 1; c >= 0; --c) {
        uint32_t open, E;
        load(c, open, E);
        const int n = min(32, len - 32 * c);
        const uint32_t F = fill_down((E | (carry << (n - 1))) & open, open);
        store(c, F & ~E);
        ch |= F & ~E;
        carry = F & 1u;
    }
    if (ch) *J.changed = 1;
}

// The same sweeps with the chunks of a line side by side: G lanes per line
// (G a power of two >= the line's 32-node chunks), all chunk loads at once.
// A chunk maps a carry-in to a carry-out (filled from 0 and from 1), and a
// log-step scan of those maps across the group gives each chunk the carry
// the node-by-node sweep brings to it; the backward pass then does the same
// on the forward result.  Lanes past the line's end carry nothing.
__device__ __forceinline__ uint32_t carry_compose(uint32_t later, uint32_t earlier)
{
    // maps as 2 bits: bit 0 = image of 0, bit 1 = image of 1
    const uint32_t a = (later >> (earlier & 1u)) & 1u;
    const uint32_t b = (later >> ((earlier >> 1) & 1u)) & 1u;
    return a | (b << 1);
}

__global__ void flood_par_kernel(const FloodJob *__restrict__ jobs, int axis, int G)
{
    const FloodJob J = jobs[blockIdx.y];
    if (*J.active == 0) return;
    const GridDesc g = J.g;
    const int64_t tg = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t t = tg / G;  // line
    const int c = (int)(tg - t * G);  // chunk
    int64_t base = 0, stride = 1, nlines;
    int len;
    if (axis == 0) {
        nlines = (int64_t)g.ny * g.nz;
        base = t * g.nx;
        stride = 1;
        len = g.nx;
    } else if (axis == 1) {
        nlines = (int64_t)g.nx * g.nz;
        const int64_t x = t % g.nx, z = t / g.nx;
        base = x + (int64_t)g.nx * g.ny * z;
        stride = g.nx;
        len = g.ny;
    } else {
        nlines = (int64_t)g.nx * g.ny;
        base = t;
        stride = (int64_t)g.nx * g.ny;
        len = g.nz;
    }
    // whole groups stay together for the shuffles (a line past the end is idle)
    const bool line_ok = t < nlines;
    const int k0 = 32 * c, n = line_ok ? max(0, min(32, len - k0)) : 0;
    const auto sas = to_global<const uint8_t>(J.sas);
    const auto ext = to_global<uint8_t>(J.ext);
    uint32_t open = 0u, E = 0u;
    if (n > 0) {
        uint8_t sv[32], ev[32];
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            const int64_t i = base + (int64_t)(k0 + min(k, n - 1)) * stride;
            sv[k] = sas[i];
            ev[k] = ext[i];
        }
#pragma unroll
        for (int k = 0; k < 32; ++k) {
            open |= (uint32_t)(sv[k] == 0) << k;
            E |= (uint32_t)(ev[k] != 0) << k;
        }
        const uint32_t live = n == 32 ? ~0u : (1u << n) - 1u;
        open &= live;
        E &= live;
    }
    const uint32_t top = n > 0 ? 1u << (n - 1) : 0u;
    // forward: carry maps, inclusive scan from chunk 0, carry-in = the scan up
    // to the previous chunk applied to 0
    uint32_t fm = 0u;
    if (n > 0) {
        const uint32_t F0 = fill_up(E & open, open), F1 = fill_up((E | 1u) & open, open);
        fm = ((F0 & top) ? 1u : 0u) | ((F1 & top) ? 2u : 0u);
    }
    uint32_t h = fm;
    for (int d = 1; d < G; d <<= 1) {
        const uint32_t hp = __shfl_up(h, d, G);
        if (c >= d) h = carry_compose(h, hp);
    }
    const uint32_t hprev = __shfl_up(h, 1, G);
    const uint32_t cin = c > 0 ? (hprev & 1u) : 0u;
    const uint32_t Ff = n > 0 ? fill_up((E | cin) & open, open) : 0u;
    // backward on the forward result: carry maps from the line's end
    uint32_t bm = 0u;
    if (n > 0) {
        const uint32_t B0 = fill_down(Ff & open, open), B1 = fill_down((Ff | top) & open, open);
        bm = (B0 & 1u) | ((B1 & 1u) << 1);
    }
    uint32_t hb = bm;
    for (int d = 1; d < G; d <<= 1) {
        const uint32_t hn = __shfl_down(hb, d, G);
        if (c + d < G) hb = carry_compose(hb, hn);
    }
    const uint32_t hnext = __shfl_down(hb, 1, G);
    const uint32_t cinb = (c + 1 < G) ? (hnext & 1u) : 0u;
    const uint32_t Fb = n > 0 ? fill_down((Ff | (cinb ? top : 0u)) & open, open) : 0u;
    uint32_t nw = Fb & ~E;
    const bool ch = nw != 0u;
    while (nw) {
        const int k = __ffs(nw) - 1;
        ext[base + (int64_t)(k0 + k) * stride] = 1;
        nw &= nw - 1u;
    }
    if (ch) *J.changed = 1;
}

// End of a flood round of one map: active = changed, changed = 0.
__global__ void flood_round_end_kernel(const FloodJob *__restrict__ job)
{
    if (threadIdx.x == 0) {
        *job->active = *job->changed;
        *job->changed = 0;
    }
}

// ---- the same flood on bits ------------------------------------------------
// Rows of W = ceil(nx / 32) words along x (padding bits closed).  A sweep along
// x runs fill_up / fill_down word by word with the carry between words; a
// sweep along y or z gives a wavefront to each column of words: lane l holds
// K = ceil(len / 64) consecutive words, and the line's forward fill
// F_j = e_j | (o_j & F_(j-1)) is a scan of the maps c -> A | (P & c) (A, P bit
// vectors: a word's map is (e_j, o_j)), composed within the lane and then
// across lanes in log2(64) shuffle steps; the backward fill is the same scan
// from the line's end.  32 lines of the grid go through each word at once.
// The fixpoint (open nodes connected to an open boundary node) is the byte
// flood's.
constexpr int kFloodMaxK = 8;  // y and z lines of up to 512 nodes

int64_t ses_bits_words(const std::vector<SesJob> &jobs)
{
    int64_t t = 0;
    for (const SesJob &J : jobs) {
        if (J.g.ny > 64 * kFloodMaxK || J.g.nz > 64 * kFloodMaxK) return 0;
        t += 2 * (int64_t)((J.g.nx + 31) / 32) * J.g.ny * J.g.nz;
    }
    return t;
}

// open and exterior-seed bits of map blockIdx.y: one thread per bit of the
// padded rows, a wavefront's 64 bits are two whole words
__global__ __launch_bounds__(256) void flood_pack_kernel(const FloodJob *__restrict__ jobs)
{
    const FloodJob J = jobs[blockIdx.y];
    const GridDesc g = J.g;
    const int64_t rowbits = (int64_t)J.W * 32;
    const int64_t total = rowbits * g.ny * g.nz;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ((int64_t)blockIdx.x * blockDim.x >= total) return;  // whole workgroup past the map
    bool o = false, e = false;
    if (v < total) {
        const int64_t row = v / rowbits;
        const int x = (int)(v - row * rowbits);
        if (x < g.nx) {
            const int y = (int)(row % g.ny), z = (int)(row / g.ny);
            o = J.sas[row * g.nx + x] == 0;
            e = o && (x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1 || z == 0 ||
                      z == g.nz - 1);
        }
    }
    const uint64_t bo = __ballot(o), be = __ballot(e);
    const int lane = threadIdx.x & 63;
    if (v < total && (lane & 31) == 0) {
        J.open[v >> 5] = (uint32_t)(bo >> lane);
        J.extb[v >> 5] = (uint32_t)(be >> lane);
    }
}

// x sweeps: one thread per row of map blockIdx.y
__global__ __launch_bounds__(256) void flood_bits_x_kernel(const FloodJob *__restrict__ jobs)
{
    const FloodJob J = jobs[blockIdx.y];
    if (*J.active == 0) return;
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (int64_t)J.g.ny * J.g.nz) return;
    const int W = J.W;
    const uint32_t *__restrict__ op = J.open + r * W;
    uint32_t *__restrict__ ex = J.extb + r * W;
    bool ch = false;
    uint32_t c = 0u;
    for (int w = 0; w < W; ++w) {
        const uint32_t o = op[w], e = ex[w];
        const uint32_t F = fill_up(e | (c & o), o);
        if (F != e) {
            ex[w] = F;
            ch = true;
        }
        c = F >> 31;
    }
    c = 0u;
    for (int w = W - 1; w >= 0; --w) {
        const uint32_t o = op[w], e = ex[w];
        const uint32_t B = fill_down(e | ((c << 31) & o), o);
        if (B != e) {
            ex[w] = B;
            ch = true;
        }
        c = B & 1u;
    }
    if (ch) *J.changed = 1;
}

// y (axis 1) or z (axis 2) sweeps: one wavefront per column of words of map
// blockIdx.y
__global__ __launch_bounds__(256) void flood_bits_yz_kernel(const FloodJob *__restrict__ jobs,
                                                            int axis)
{
    const FloodJob J = jobs[blockIdx.y];
    if (*J.active == 0) return;
    const GridDesc g = J.g;
    const int W = J.W;
    const int lane = threadIdx.x & 63;
    const int64_t col = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    int64_t base, stride;
    int len;
    if (axis == 1) {  // column (w, z)
        if (col >= (int64_t)W * g.nz) return;
        const int64_t z = col / W, w = col - z * W;
        base = z * g.ny * W + w;
        stride = W;
        len = g.ny;
    } else {  // column (w, y)
        if (col >= (int64_t)W * g.ny) return;
        base = col;
        stride = (int64_t)W * g.ny;
        len = g.nz;
    }
    const int K = (len + 63) / 64;
    uint32_t o[kFloodMaxK], e[kFloodMaxK];
#pragma unroll
    for (int k = 0; k < kFloodMaxK; ++k) {
        const int j = lane * K + k;
        const bool ok = k < K && j < len;
        o[k] = ok ? J.open[base + j * stride] : 0u;
        e[k] = ok ? J.extb[base + j * stride] : 0u;
    }
    // forward: the lane's map, then the scan across lanes
    uint32_t A = 0u, P = ~0u;
#pragma unroll
    for (int k = 0; k < kFloodMaxK; ++k) {
        if (k < K) {
            A = e[k] | (o[k] & A);
            P &= o[k];
        }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t Ap = __shfl_up(A, d, 64), Pp = __shfl_up(P, d, 64);
        if (lane >= d) {
            A |= P & Ap;
            P &= Pp;
        }
    }
    uint32_t c = __shfl_up(A, 1, 64);
    if (lane == 0) c = 0u;
    uint32_t F[kFloodMaxK];
#pragma unroll
    for (int k = 0; k < kFloodMaxK; ++k) {
        F[k] = e[k] | (o[k] & c);
        if (k < K) c = F[k];
    }
    // backward on the forward result, from the line's end
    A = 0u;
    P = ~0u;
#pragma unroll
    for (int k = kFloodMaxK - 1; k >= 0; --k) {
        if (k < K) {
            A = F[k] | (o[k] & A);
            P &= o[k];
        }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t An = __shfl_down(A, d, 64), Pn = __shfl_down(P, d, 64);
        if (lane + d < 64) {
            A |= P & An;
            P &= Pn;
        }
    }
    c = __shfl_down(A, 1, 64);
    if (lane == 63) c = 0u;
    bool ch = false;
#pragma unroll
    for (int k = kFloodMaxK - 1; k >= 0; --k) {
        if (k < K) {
            const uint32_t B = F[k] | (o[k] & c);
            c = B;
            const int j = lane * K + k;
            if (j < len && B != e[k]) {
                J.extb[base + j * stride] = B;
                ch = true;
            }
        }
    }
    if (__ballot(ch) != 0 && lane == 0) *J.changed = 1;
}

// End of a flood round of every map: active = changed, changed = 0.
__global__ void flood_rounds_end_kernel(const FloodJob *__restrict__ jobs, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    *jobs[k].active = *jobs[k].changed;
    *jobs[k].changed = 0;
}

// exterior bytes of map blockIdx.y from its bits
__global__ __launch_bounds__(256) void flood_unpack_kernel(const FloodJob *__restrict__ jobs)
{
    const FloodJob J = jobs[blockIdx.y];
    const GridDesc g = J.g;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.size()) return;
    const int64_t row = idx / g.nx;
    const int x = (int)(idx - row * g.nx);
    J.ext[idx] = (uint8_t)((J.extb[row * J.W + (x >> 5)] >> (x & 31)) & 1u);
}

void launch_ses_batch(const std::vector<SesJob> &jobs, FloodJob *d_table, int *d_flags,
                      uint32_t *d_bits, hipStream_t s)
{
    const int n = (int)jobs.size();
    if (n == 0) return;
    std::vector<FloodJob> ht(n);
    int64_t off = 0, max_words = 0, max_rows = 0, max_cols = 0, max_n = 0;
    for (int k = 0; k < n; ++k) {
        const SesJob &J = jobs[k];
        const int64_t N = J.g.size();
        const int W = (J.g.nx + 31) / 32;
        ht[k] = FloodJob{J.g, J.sas, J.ext, d_flags + k, d_flags + n + k, W, nullptr, nullptr};
        if (d_bits) {
            const int64_t words = (int64_t)W * J.g.ny * J.g.nz;
            ht[k].open = d_bits + off;
            ht[k].extb = d_bits + off + words;
            off += 2 * words;
            max_words = std::max(max_words, words);
            max_rows = std::max(max_rows, (int64_t)J.g.ny * J.g.nz);
            max_cols = std::max(max_cols, (int64_t)W * std::max(J.g.ny, J.g.nz));
            max_n = std::max(max_n, N);
        } else {
            hipLaunchKernelGGL(ext_init_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, J.g,
                               J.sas, J.ext);
            LAUNCH_OK();
        }
    }
    HIP_OK(hipMemcpyAsync(d_table, ht.data(), sizeof(FloodJob) * n, hipMemcpyHostToDevice, s));
    launch_fill_i32(d_flags, 0, n, s);
    launch_fill_i32(d_flags + n, 1, n, s);
    if (d_bits) {
        // every map in the same launches; kRoundsBits rounds per host check
        constexpr int kRoundsBits = 6;
        hipLaunchKernelGGL(flood_pack_kernel, dim3(ceil_div(32 * max_words, 256), n), dim3(256), 0,
                           s, d_table);
        LAUNCH_OK();
        std::vector<int> act(n, 1);
        for (int round = 0; round < 100000; round += kRoundsBits) {
            for (int r = 0; r < kRoundsBits; ++r) {
                hipLaunchKernelGGL(flood_bits_x_kernel, dim3(ceil_div(max_rows, 256), n),
                                   dim3(256), 0, s, d_table);
                LAUNCH_OK();
                for (int ax = 1; ax < 3; ++ax) {
                    hipLaunchKernelGGL(flood_bits_yz_kernel, dim3(ceil_div(max_cols, 4), n),
                                       dim3(256), 0, s, d_table, ax);
                    LAUNCH_OK();
                }
                hipLaunchKernelGGL(flood_rounds_end_kernel, dim3(ceil_div(n, 64)), dim3(64), 0, s,
                                   d_table, n);
                LAUNCH_OK();
            }
            HIP_OK(hipMemcpyAsync(act.data(), d_flags + n, sizeof(int) * n, hipMemcpyDeviceToHost,
                                  s));
            HIP_OK(hipStreamSynchronize(s));
            if (std::none_of(act.begin(), act.end(), [](int a) { return a != 0; })) break;
        }
        hipLaunchKernelGGL(flood_unpack_kernel, dim3(ceil_div(max_n, 256), n), dim3(256), 0, s,
                           d_table);
        LAUNCH_OK();
    }
    // rounds of each map until a round changes nothing.  kRounds rounds of
    // every map still flooding are queued per host check (one check for all
    // of them); a round after a map's fixpoint finds active = 0 and its
    // launches exit at once.  Each map keeps launches of its own: sharing
    // launches between maps measured slower, as each map's x sweeps then
    // compete for the caches.
    constexpr int kRounds = 4;
    std::vector<int> live(n), act(n, 1);
    for (int k = 0; k < n; ++k) live[k] = k;
    if (d_bits) live.clear();
    for (int round = 0; !live.empty() && round < 100000; round += kRounds) {
        for (int k : live) {
            const GridDesc &g = jobs[k].g;
            const int64_t ln[3] = {(int64_t)g.ny * g.nz, (int64_t)g.nx * g.nz,
                                   (int64_t)g.nx * g.ny};
            for (int r = 0; r < kRounds; ++r) {
                for (int ax = 0; ax < 3; ++ax) {
                    const int len = ax == 0 ? g.nx : ax == 1 ? g.ny : g.nz;
                    const int nch = (len + 31) / 32;
                    if (nch <= 64) {
                        int G = 1;
                        while (G < nch) G <<= 1;
                        hipLaunchKernelGGL(flood_par_kernel, dim3(ceil_div(ln[ax] * G, 256), 1),
                                           dim3(256), 0, s, d_table + k, ax, G);
                    } else {
                        hipLaunchKernelGGL(flood_batch_kernel, dim3(ceil_div(ln[ax], 128), 1),
                                           dim3(128), 0, s, d_table + k, ax);
                    }
                    LAUNCH_OK();
                }
                hipLaunchKernelGGL(flood_round_end_kernel, dim3(1), dim3(64), 0, s, d_table + k);
                LAUNCH_OK();
            }
        }
        HIP_OK(hipMemcpyAsync(act.data(), d_flags + n, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        live.erase(std::remove_if(live.begin(), live.end(), [&](int k) { return act[k] == 0; }),
                   live.end());
    }

    for (const SesJob &J : jobs) {
        const int64_t N = J.g.size();
        const int R = (int)std::floor(std::sqrt((double)J.thresh));
        hipLaunchKernelGGL(dil_x_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, J.g, R,
                           J.thresh, J.ext, J.work_a);
        LAUNCH_OK();
        hipLaunchKernelGGL(dil_yz_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, J.g, R,
                           J.thresh, 1, J.work_a, J.work_b, 0);
        LAUNCH_OK();
        hipLaunchKernelGGL(dil_yz_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, J.g, R,
                           J.thresh, 2, J.work_b, J.inside_out, 1);
        LAUNCH_OK();
    }
}

// ---- face dielectrics, assign_dielectrics (pb_grid.rs:827-940) -------------
// Mixed edges take the VDW-sphere fraction of fraction_inside_along_axis
// (pb_grid.rs:717-807) from atoms in nearby cells (the same interval set as the
// reference's all-atom scan), sorted and merged in FP64.
constexpr int kMaxIv = 48;

__device__ double edge_fraction(double lo, double hi, double p1, double p2, int axis,
                                const double *__restrict__ xyz, const double *__restrict__ radii,
                                double max_r, const CellList &cl, int *overflow)
{
    const double len = hi - lo;
    if (len <= 0.0) return 0.0;
    const int a1 = axis == 0 ? 1 : 0;
    const int a2 = axis == 2 ? 1 : 2;
    double mid[3];
    mid[axis] = 0.5 * (lo + hi);
    mid[a1] = p1;
    mid[a2] = p2;
    const double w = max_r + 0.5 * len + 1e-9;
    const int cx0 = max(0, (int)floor((mid[0] - w - cl.ox) * cl.inv));
    const int cx1 = min(cl.dx - 1, (int)floor((mid[0] + w - cl.ox) * cl.inv));
    const int cy0 = max(0, (int)floor((mid[1] - w - cl.oy) * cl.inv));
    const int cy1 = min(cl.dy - 1, (int)floor((mid[1] + w - cl.oy) * cl.inv));
    const int cz0 = max(0, (int)floor((mid[2] - w - cl.oz) * cl.inv));
    const int cz1 = min(cl.dz - 1, (int)floor((mid[2] + w - cl.oz) * cl.inv));
    double ia[kMaxIv], ib[kMaxIv];
    int m = 0;
    for (int cz = cz0; cz <= cz1; ++cz)
        for (int cy = cy0; cy <= cy1; ++cy)
            for (int cx = cx0; cx <= cx1; ++cx) {
                const int c = cx + cl.dx * (cy + cl.dy * cz);
                for (int k = cl.start[c]; k < cl.start[c + 1]; ++k) {
                    const int a = cl.atoms[k];
                    const double R = radii[a] + 0.0;
                    const double d1 = p1 - xyz[3 * a + a1];
                    const double d2 = p2 - xyz[3 * a + a2];
                    const double rs = R * R - d1 * d1 - d2 * d2;
                    if (rs <= 0.0) continue;
                    const double rho = sqrt(rs);
                    const double ca = xyz[3 * a + axis];
                    const double ilo = fmax(ca - rho, lo);
                    const double ihi = fmin(ca + rho, hi);
                    if (ilo < ihi) {
                        if (m < kMaxIv) {
                            // insertion by start (ties in any order merge identically)
                            int p = m++;
                            while (p > 0 && ia[p - 1] > ilo) {
                                ia[p] = ia[p - 1];
                                ib[p] = ib[p - 1];
                                --p;
                            }
                            ia[p] = ilo;
                            ib[p] = ihi;
                        } else {
                            atomicExch(overflow, 1);
                        }
                    }
                }
            }
    if (m == 0) return 0.0;
    double tot = 0.0, cl_ = ia[0], ch = ib[0];
    for (int k = 1; k < m; ++k) {
        if (ia[k] <= ch) {
            ch = fmax(ch, ib[k]);
        } else {
            tot += ch - cl_;
            cl_ = ia[k];
            ch = ib[k];
        }
    }
    tot += ch - cl_;
    double f = tot / len;
    f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
    return f;
}

__global__ void faces_kernel(GridDesc g, const uint8_t *__restrict__ inside,
                             const double *__restrict__ xyz, const double *__restrict__ radii,
                             double max_r, CellList cl, double ein, double eout,
                             double *__restrict__ ex, double *__restrict__ ey,
                             double *__restrict__ ez, float *__restrict__ fx,
                             float *__restrict__ fy, float *__restrict__ fz, int *overflow)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.size()) return;
    int x, y, z;
    node_of(g, idx, x, y, z);
    const uint8_t a = inside[idx];
    const double pt[3] = {g.ox + (double)x * g.h, g.oy + (double)y * g.h, g.oz + (double)z * g.h};
    const int pos[3] = {x, y, z};
    const int lim[3] = {g.nx, g.ny, g.nz};
    const int64_t step[3] = {1, (int64_t)g.nx, (int64_t)g.nx * g.ny};
    double *dst[3] = {ex, ey, ez};
    float *dstf[3] = {fx, fy, fz};
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        double e = eout;
        if (pos[axis] + 1 < lim[axis]) {
            const uint8_t b = inside[idx + step[axis]];
            if (a == b) {
                e = a ? ein : eout;
            } else {
                const double lo = pt[axis], hi = pt[axis] + g.h;
                const double p1 = axis == 0 ? pt[1] : pt[0];
                const double p2 = axis == 2 ? pt[1] : pt[2];
                const double f = edge_fraction(lo, hi, p1, p2, axis, xyz, radii, max_r, cl, overflow);
                if (f <= 0.0)
                    e = eout;
                else if (f >= 1.0)
                    e = ein;
                else
                    e = 1.0 / (f / ein + (1.0 - f) / eout);
            }
        }
        dst[axis][idx] = e;
        dstf[axis][idx] = (float)e;
    }
}

// The same faces in two launches.  faces_split_kernel writes the faces whose
// two nodes lie on one side and lists the mixed edges ((node << 2) | axis; one
// atomic per workgroup, so the list's order varies, but each entry's result
// does not).  faces_mixed_kernel gives each listed edge a wavefront: lanes test
// the candidate atoms side by side (a row of cells at a time: the atoms of
// consecutive x cells are contiguous in the cell list) and append the clipped
// intervals to a wave-local LDS list in the atom scan order of edge_fraction;
// a stable rank by start puts them in edge_fraction's insertion order, and the
// merge runs in that order on wave-uniform values, so every face is bitwise
// the one-thread-per-node kernel's.
constexpr int kSplitBlock = 1024;

// LDS written by some lanes of a wavefront, then read by others
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// lane l's value (l wave-uniform)
__device__ inline double readlane_f64(double v, int l)
{
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__global__ __launch_bounds__(kSplitBlock) void faces_split_kernel(
    GridDesc g, const uint8_t *__restrict__ inside, double ein, double eout,
    double *__restrict__ ex, double *__restrict__ ey, double *__restrict__ ez,
    float *__restrict__ fx, float *__restrict__ fy, float *__restrict__ fz,
    uint64_t *__restrict__ list, int cap, int *__restrict__ count)
{
    __shared__ int wcnt[kSplitBlock / 64];
    __shared__ int bbase;
    const int64_t N = g.size();
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = idx < N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int pos[3] = {0, 0, 0};
    if (in) node_of(g, idx, pos[0], pos[1], pos[2]);
    const uint8_t a = in ? inside[idx] : 0;
    const int lim[3] = {g.nx, g.ny, g.nz};
    const int64_t step[3] = {1, (int64_t)g.nx, (int64_t)g.nx * g.ny};
    double *dst[3] = {ex, ey, ez};
    float *dstf[3] = {fx, fy, fz};
    uint64_t bm[3];
    int wtot = 0;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        bool mixed = false;
        double e = eout;
        if (in && pos[axis] + 1 < lim[axis]) {
            const uint8_t b = inside[idx + step[axis]];
            if (a == b)
                e = a ? ein : eout;
            else
                mixed = true;
        }
        if (in && !mixed) {
            dst[axis][idx] = e;
            dstf[axis][idx] = (float)e;
        }
        bm[axis] = __ballot(mixed);
        wtot += __popcll(bm[axis]);
    }
    if (lane == 0) wcnt[wave] = wtot;
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < kSplitBlock / 64; ++k) t += wcnt[k];
        bbase = t ? atomicAdd(count, t) : 0;
    }
    __syncthreads();
    int base = bbase;
    for (int k = 0; k < wave; ++k) base += wcnt[k];
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        const uint64_t m = bm[axis];
        if ((m >> lane) & 1u) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            const int slot = base + (int)below;
            if (slot < cap) list[slot] = ((uint64_t)idx << 2) | (uint64_t)axis;
        }
        base += __popcll(m);
    }
}

__global__ __launch_bounds__(256) void faces_mixed_kernel(
    GridDesc g, const double *__restrict__ xyz, const double *__restrict__ radii, double max_r,
    CellList cl, double ein, double eout, const uint64_t *__restrict__ list,
    const int *__restrict__ count, int cap, double *__restrict__ ex, double *__restrict__ ey,
    double *__restrict__ ez, float *__restrict__ fx, float *__restrict__ fy,
    float *__restrict__ fz, int *overflow)
{
    __shared__ double ia_s[4][kMaxIv], ib_s[4][kMaxIv];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    double *ia = ia_s[wave], *ib = ib_s[wave];
    const int n = min(*count, cap);
    double *dst[3] = {ex, ey, ez};
    float *dstf[3] = {fx, fy, fz};
    for (int q = blockIdx.x * 4 + wave; q < n; q += gridDim.x * 4) {
        const uint64_t ent = list[q];
        const int64_t idx = (int64_t)(ent >> 2);
        const int axis = (int)(ent & 3u);
        int x, y, z;
        node_of(g, idx, x, y, z);
        const double pt[3] = {g.ox + (double)x * g.h, g.oy + (double)y * g.h,
                              g.oz + (double)z * g.h};
        // edge_fraction (lo, hi, p1, p2) of this edge, atoms tested by lanes
        const double lo = pt[axis], hi = pt[axis] + g.h;
        const double p1 = axis == 0 ? pt[1] : pt[0];
        const double p2 = axis == 2 ? pt[1] : pt[2];
        const double len = hi - lo;
        double f = 0.0;
        if (len > 0.0) {
            const int a1 = axis == 0 ? 1 : 0;
            const int a2 = axis == 2 ? 1 : 2;
            double mid[3];
            mid[axis] = 0.5 * (lo + hi);
            mid[a1] = p1;
            mid[a2] = p2;
            const double w = max_r + 0.5 * len + 1e-9;
            const int cx0 = max(0, (int)floor((mid[0] - w - cl.ox) * cl.inv));
            const int cx1 = min(cl.dx - 1, (int)floor((mid[0] + w - cl.ox) * cl.inv));
            const int cy0 = max(0, (int)floor((mid[1] - w - cl.oy) * cl.inv));
            const int cy1 = min(cl.dy - 1, (int)floor((mid[1] + w - cl.oy) * cl.inv));
            const int cz0 = max(0, (int)floor((mid[2] - w - cl.oz) * cl.inv));
            const int cz1 = min(cl.dz - 1, (int)floor((mid[2] + w - cl.oz) * cl.inv));
            // the candidate cells form rows of consecutive x cells, and a
            // row's atoms are one contiguous range of the cell list: lanes
            // take the rows' ranges (cz outer, cy inner, as edge_fraction
            // scans them), then the atoms of the ranges laid end to end
            const int ncy = cy1 - cy0 + 1, ncz = cz1 - cz0 + 1;
            const int nr = (cx0 <= cx1 && ncy > 0 && ncz > 0) ? ncy * ncz : 0;
            int m = 0;  // intervals found (wave-uniform)
            for (int r0 = 0; r0 < nr; r0 += 64) {
                const int r = r0 + lane;
                int kb = 0, kn = 0;
                if (r < nr) {
                    const int cy = cy0 + r % ncy, cz = cz0 + r / ncy;
                    const int crow = cl.dx * (cy + cl.dy * cz);
                    kb = cl.start[crow + cx0];
                    kn = cl.start[crow + cx1 + 1] - kb;
                }
                int incl = kn;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int t = __shfl_up(incl, d, 64);
                    if (lane >= d) incl += t;
                }
                const int excl = incl - kn;
                const int total = __builtin_amdgcn_readlane(incl, 63);
                const int rows = min(64, nr - r0);
                for (int j0 = 0; j0 < total; j0 += 64) {
                    const int j = j0 + lane;
                    // the last row starting at or before j holds it
                    int k = 0;
                    for (int q = 0; q < rows; ++q) {
                        const int eq = __builtin_amdgcn_readlane(excl, q);
                        const int bq = __builtin_amdgcn_readlane(kb, q);
                        if (j >= eq) k = bq + (j - eq);
                    }
                    bool hit = false;
                    double ilo = 0.0, ihi = 0.0;
                    if (j < total) {
                        const int at = cl.atoms[k];
                        const double R = radii[at] + 0.0;
                        const double d1 = p1 - xyz[3 * at + a1];
                        const double d2 = p2 - xyz[3 * at + a2];
                        const double rs = R * R - d1 * d1 - d2 * d2;
                        if (rs > 0.0) {
                            const double rho = sqrt(rs);
                            const double ca = xyz[3 * at + axis];
                            ilo = fmax(ca - rho, lo);
                            ihi = fmin(ca + rho, hi);
                            hit = ilo < ihi;
                        }
                    }
                    const uint64_t hm = __ballot(hit);
                    if (hit) {
                        const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                            (uint32_t)(hm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0u));
                        const int slot = m + (int)below;
                        if (slot < kMaxIv) {
                            ia[slot] = ilo;
                            ib[slot] = ihi;
                        }
                    }
                    m += __popcll(hm);
                }
            }
            if (m > kMaxIv) {
                if (lane == 0) atomicExch(overflow, 1);
                m = kMaxIv;
            }
            wave_lds_sync();
            // lane p holds interval p; its stable rank by start is its place
            // in edge_fraction's insertion order (the first kMaxIv of the scan)
            double sa = 0.0, sb = 0.0;
            if (lane < m) {
                sa = ia[lane];
                sb = ib[lane];
            }
            int rank = 0;
            for (int q = 0; q < m; ++q) {
                const double aq = readlane_f64(sa, q);
                rank += (aq < sa || (aq == sa && q < lane)) ? 1 : 0;
            }
            wave_lds_sync();
            if (lane < m) {
                ia[rank] = sa;
                ib[rank] = sb;
            }
            wave_lds_sync();
            if (lane < m) {
                sa = ia[lane];
                sb = ib[lane];
            }
            if (m > 0) {
                // the merge of edge_fraction, in the same order, on uniform values
                double tot = 0.0, c_lo = readlane_f64(sa, 0), c_hi = readlane_f64(sb, 0);
                for (int k = 1; k < m; ++k) {
                    const double ak = readlane_f64(sa, k), bk = readlane_f64(sb, k);
                    if (ak <= c_hi) {
                        c_hi = fmax(c_hi, bk);
                    } else {
                        tot += c_hi - c_lo;
                        c_lo = ak;
                        c_hi = bk;
                    }
                }
                tot += c_hi - c_lo;
                f = tot / len;
                f = f < 0.0 ? 0.0 : (f > 1.0 ? 1.0 : f);
            }
        }
        if (lane == 0) {
            double e;
            if (f <= 0.0)
                e = eout;
            else if (f >= 1.0)
                e = ein;
            else
                e = 1.0 / (f / ein + (1.0 - f) / eout);
            dst[axis][idx] = e;
            dstf[axis][idx] = (float)e;
        }
        wave_lds_sync();
    }
}

int faces_list_cap(int64_t n) { return (int)std::min<int64_t>(n / 4 + 1024, (int64_t)1 << 30); }

void launch_faces_split(const GridDesc &g, const uint8_t *inside, const double *xyz,
                        const double *radii, double max_r, const CellList &cl, double eps_in,
                        double eps_out, double *ex, double *ey, double *ez, float *fx, float *fy,
                        float *fz, uint64_t *list, int cap, int *count, int *overflow,
                        hipStream_t s)
{
    const int64_t N = g.size();
    hipLaunchKernelGGL(faces_split_kernel, dim3(ceil_div(N, kSplitBlock)), dim3(kSplitBlock), 0, s, g, inside,
                       eps_in, eps_out, ex, ey, ez, fx, fy, fz, list, cap, count);
    LAUNCH_OK();
    hipLaunchKernelGGL(faces_mixed_kernel, dim3(2048), dim3(256), 0, s, g, xyz, radii, max_r, cl,
                       eps_in, eps_out, list, count, cap, ex, ey, ez, fx, fy, fz, overflow);
    LAUNCH_OK();
}

void launch_faces(const GridDesc &g, const uint8_t *inside, const double *xyz,
                  const double *radii, double max_r, const CellList &cl, double eps_in,
                  double eps_out, double *ex, double *ey, double *ez, float *fx, float *fy,
                  float *fz, int *overflow, hipStream_t s)
{
    const int64_t N = g.size();
    hipLaunchKernelGGL(faces_kernel, dim3(ceil_div(N, 128)), dim3(128), 0, s, g, inside, xyz,
                       radii, max_r, cl, eps_in, eps_out, ex, ey, ez, fx, fy, fz, overflow);
    LAUNCH_OK();
}

// assign_kappa, pb_grid.rs:951-985
__global__ void kappa_map_kernel(int64_t n, const uint8_t *__restrict__ ion_inside, double kb2,
                                 double *__restrict__ k2, float *__restrict__ k2f)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = ion_inside[i] ? 0.0 : kb2;
    k2[i] = v;
    k2f[i] = (float)v;
}

void launch_kappa_map(const GridDesc &g, const uint8_t *ion_inside, double kbar2, double *k2,
                      float *k2f, hipStream_t s)
{
    const int64_t N = g.size();
    hipLaunchKernelGGL(kappa_map_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, N, ion_inside,
                       kbar2, k2, k2f);
    LAUNCH_OK();
}

// ---- trilinear interpolation, PbGrid::interpolate_with_data (pb_grid.rs:140-182)
__device__ double interp3(const GridDesc &g, const double *__restrict__ d, double px, double py,
                          double pz)
{
    const double fx = (px - g.ox) / g.h, fy = (py - g.oy) / g.h, fz = (pz - g.oz) / g.h;
    const double flx = floor(fx), fly = floor(fy), flz = floor(fz);
    if (flx < 0.0 || flx >= (double)(g.nx - 1) || fly < 0.0 || fly >= (double)(g.ny - 1) ||
        flz < 0.0 || flz >= (double)(g.nz - 1))
        return 0.0;
    const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
    const double dx = fx - (double)ix, dy = fy - (double)iy, dz = fz - (double)iz;
    const double c000 = d[gi(g, ix, iy, iz)], c100 = d[gi(g, ix + 1, iy, iz)];
    const double c010 = d[gi(g, ix, iy + 1, iz)], c110 = d[gi(g, ix + 1, iy + 1, iz)];
    const double c001 = d[gi(g, ix, iy, iz + 1)], c101 = d[gi(g, ix + 1, iy, iz + 1)];
    const double c011 = d[gi(g, ix, iy + 1, iz + 1)], c111 = d[gi(g, ix + 1, iy + 1, iz + 1)];
    const double c00 = c000 * (1.0 - dx) + c100 * dx;
    const double c10 = c010 * (1.0 - dx) + c110 * dx;
    const double c01 = c001 * (1.0 - dx) + c101 * dx;
    const double c11 = c011 * (1.0 - dx) + c111 * dx;
    const double c0 = c00 * (1.0 - dy) + c10 * dy;
    const double c1 = c01 * (1.0 - dy) + c11 * dy;
    return c0 * (1.0 - dz) + c1 * dz;
}

// interpolated_boundary, pb_solver.rs:100-121
__global__ void interp_boundary_kernel(GridDesc f, GridDesc c, const double *__restrict__ cphi,
                                       double *__restrict__ phi)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= f.size()) return;
    int x, y, z;
    node_of(f, idx, x, y, z);
    const bool bnd = x == 0 || x == f.nx - 1 || y == 0 || y == f.ny - 1 || z == 0 || z == f.nz - 1;
    if (!bnd) return;
    phi[idx] = interp3(c, cphi, f.ox + (double)x * f.h, f.oy + (double)y * f.h,
                       f.oz + (double)z * f.h);
}

void launch_interp_boundary(const GridDesc &fine, const GridDesc &coarse, const double *cphi,
                            double *phi, hipStream_t s)
{
    const int64_t N = fine.size();
    hipLaunchKernelGGL(interp_boundary_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, fine,
                       coarse, cphi, phi);
    LAUNCH_OK();
}

// compute_elec_energy, pb_solver.rs:1174-1186 (block partials of q * phi(r))
__global__ __launch_bounds__(256) void gather_kernel(GridDesc g, const double *__restrict__ phi,
                                                     int n, const double *__restrict__ xyz,
                                                     const double *__restrict__ q,
                                                     double *__restrict__ partials)
{
    __shared__ double red[4];
    const int a = blockIdx.x * 256 + threadIdx.x;
    double v = 0.0;
    if (a < n) v = q[a] * interp3(g, phi, xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]);
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

void launch_energy_gather(const GridDesc &g, const double *phi, int n, const double *xyz,
                          const double *q, double *partials, int *n_blocks, hipStream_t s)
{
    *n_blocks = std::max(1, ceil_div(n, 256));
    hipLaunchKernelGGL(gather_kernel, dim3(*n_blocks), dim3(256), 0, s, g, phi, n, xyz, q,
                       partials);
    LAUNCH_OK();
}

// Per-atom PB terms 0.5 q_i (phi_solv - phi_ref)(r_i); their sum over the
// atoms is the subsystem's total of compute_pb_energy (pb_energy.rs:423-430)
__global__ __launch_bounds__(256) void atom_terms_kernel(GridDesc g,
                                                         const double *__restrict__ phi_s,
                                                         const double *__restrict__ phi_r, int n,
                                                         const double *__restrict__ xyz,
                                                         const double *__restrict__ q,
                                                         double *__restrict__ out)
{
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= n) return;
    const double x = xyz[3 * a], y = xyz[3 * a + 1], z = xyz[3 * a + 2];
    out[a] = 0.5 * q[a] * (interp3(g, phi_s, x, y, z) - interp3(g, phi_r, x, y, z));
}

void launch_pb_atom_terms(const GridDesc &g, const double *phi_solv, const double *phi_ref, int n,
                          const double *xyz, const double *q, double *out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(atom_terms_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, g, phi_solv,
                       phi_ref, n, xyz, q, out);
    LAUNCH_OK();
}

// max |v| ignoring NaN (as f64::max does), pb_energy.rs:188-194
__global__ __launch_bounds__(256) void maxabs_kernel(const double *__restrict__ v, int64_t n,
                                                     double *__restrict__ partials)
{
    __shared__ double red[256];
    double m = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        m = fmax(m, fabs(v[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// one wavefront: lane maxima, then a shuffle tree (a maximum is exact, and
// fmax from +0 drops NaNs in any order, so the result is order-independent)
__global__ __launch_bounds__(64) void max_partials_kernel(const double *__restrict__ p, int nb,
                                                          double *out)
{
    double m = 0.0;
    for (int k = threadIdx.x; k < nb; k += 64) m = fmax(m, p[k]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if (threadIdx.x == 0) *out = m;
}

void launch_maxabs(const double *v, int64_t n, double *partials, int *n_blocks, hipStream_t s)
{
    *n_blocks = (int)std::min<int64_t>(1024, std::max<int64_t>(1, ceil_div(n, 256)));
    hipLaunchKernelGGL(maxabs_kernel, dim3(*n_blocks), dim3(256), 0, s, v, n, partials);
    LAUNCH_OK();
}

void launch_max_partials(const double *partials, int n_blocks, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(max_partials_kernel, dim3(1), dim3(64), 0, s, partials, n_blocks, out);
    LAUNCH_OK();
}

// ---- small helpers ----------------------------------------------------------
template <typename T>
__global__ void fill_kernel(T *p, T v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void launch_fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel<uint8_t>, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, v, n);
    LAUNCH_OK();
}
void launch_fill_f64(double *p, double v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel<double>, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, v, n);
    LAUNCH_OK();
}
void launch_fill_i32(int *p, int v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel<int>, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, v, n);
    LAUNCH_OK();
}
void launch_fill_u64(unsigned long long *p, unsigned long long v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel<unsigned long long>, dim3(ceil_div(n, 256)), dim3(256), 0, s, p,
                       v, n);
    LAUNCH_OK();
}
void launch_fill_f32(float *p, float v, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(fill_kernel<float>, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, v, n);
    LAUNCH_OK();
}

__global__ void gather_xyz_kernel(const double *__restrict__ src, const int *__restrict__ idx,
                                  int n, double *__restrict__ dst)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int a = idx[k];
    dst[3 * k] = src[3 * a];
    dst[3 * k + 1] = src[3 * a + 1];
    dst[3 * k + 2] = src[3 * a + 2];
}

void launch_gather_xyz(const double *src, const int *idx, int n, double *dst, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(gather_xyz_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, src, idx, n, dst);
    LAUNCH_OK();
}

// Bounding box of n AoS points (min then max per axis) by one workgroup.
// Min and max are exact, so the result does not depend on the reduction order.
__global__ __launch_bounds__(1024) void bbox_kernel(const double *__restrict__ xyz, int n,
                                                    double *__restrict__ out)
{
    __shared__ double red[6][1024];
    double v[6] = {DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        for (int k = 0; k < 3; ++k) {
            const double c = xyz[3 * (int64_t)i + k];
            if (c < v[k]) v[k] = c;
            if (c > v[3 + k]) v[3 + k] = c;
        }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int k = 0; k < 6; ++k) {
                const double o = red[k][threadIdx.x + off];
                double &m = red[k][threadIdx.x];
                if (k < 3 ? o < m : o > m) m = o;
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = red[threadIdx.x][0];
}

// Plain copy of n 16-byte words (a reference point for traffic counters).
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

void launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t s)
{
    if (n16 <= 0) return;
    const int blocks = (int)std::min<int64_t>(ceil_div(n16, 256), 8192);
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s, (const uint4 *)src, (uint4 *)dst,
                       n16);
    LAUNCH_OK();
}

void launch_bbox(const double *xyz, int n, double *out6, hipStream_t s)
{
    hipLaunchKernelGGL(bbox_kernel, dim3(1), dim3(1024), 0, s, xyz, n, out6);
    LAUNCH_OK();
}

// Per-segment sums in segment order (one lane per segment; segments are
// residues of at most a few hundred atoms).
__global__ void segment_sum_kernel(const double *__restrict__ v, const int *__restrict__ seg_atoms,
                                   const int *__restrict__ seg_ptr, int n_seg,
                                   double *__restrict__ out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seg) return;
    double acc = 0.0;
    for (int k = seg_ptr[s]; k < seg_ptr[s + 1]; ++k) acc += v[seg_atoms[k]];
    out[s] = acc;
}

void launch_segment_sum(const double *v, const int *seg_atoms, const int *seg_ptr, int n_seg,
                        double *out, hipStream_t s)
{
    if (n_seg <= 0) return;
    hipLaunchKernelGGL(segment_sum_kernel, dim3(ceil_div(n_seg, 256)), dim3(256), 0, s, v,
                       seg_atoms, seg_ptr, n_seg, out);
    LAUNCH_OK();
}

}  // namespace mmpbsa
