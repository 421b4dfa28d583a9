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

#include <algorithm>
#include <cfloat>

#include "common.h"
#include "mg.h"

#include <type_traits>

namespace mmpbsa {

__device__ __forceinline__ bool stopped(const MgStatus *st) { return st->state != 0; }

// (solve, block) of a batched launch, XCD-ordered so that each XCD walks
// contiguous blocks of one solve and z-neighbour planes are still in its L2.
// Without a work list the grid is blocks x solves (every solve gets the
// largest solve's block count; the surplus returns at once).  With a work
// list wl = {n, pre[0..n], id[0..n-1]} the grid is one-dimensional with
// pre[n] blocks: blocks pre[k] .. pre[k+1]-1 belong to solve id[k], found by a
// binary search that is uniform across the workgroup (scalar loads).
struct BatchBlock {
    int s, b;
};
__device__ __forceinline__ BatchBlock batch_block(const uint32_t *__restrict__ wl)
{
    if (!wl) {
        const uint32_t flat =
            xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
        const uint32_t s = flat / gridDim.x;
        return {(int)s, (int)(flat - s * gridDim.x)};
    }
    const uint32_t flat = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t n = wl[0];
    uint32_t lo = 0, hi = n;  // wl[1 + lo] <= flat < wl[1 + hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (wl[1 + mid] <= flat)
            lo = mid;
        else
            hi = mid;
    }
    return {(int)wl[n + 2 + lo], (int)(flat - wl[1 + lo])};
}

// ---- node operations shared by the batched kernels and the tail --------------
template <typename T>
__device__ __forceinline__ T face_bit(uint32_t w, int bit, T cin, T cout)
{
    return (w >> bit) & 1u ? cin : cout;
}

// The six face dielectrics and kappa-bar^2 of interior node idx.
template <typename T, typename V>
__device__ __forceinline__ void coeffs(const V &L, int64_t idx, T &exp_, T &exm, T &eyp, T &eym,
                                       T &ezp, T &ezm, T &k2)
{
    const int64_t sy = L.nx, sz = (int64_t)L.nx * L.ny;
    if (L.cw) {
        const uint32_t w = L.cw[idx];
        k2 = (w & kCodeKappa) ? (T)L.c_k2 : (T)0;
        if (!(w & kCodeSpecial)) {
            exp_ = face_bit<T>(w, 0, L.c_in, L.c_out);
            exm = face_bit<T>(w, 1, L.c_in, L.c_out);
            eyp = face_bit<T>(w, 2, L.c_in, L.c_out);
            eym = face_bit<T>(w, 3, L.c_in, L.c_out);
            ezp = face_bit<T>(w, 4, L.c_in, L.c_out);
            ezm = face_bit<T>(w, 5, L.c_in, L.c_out);
            return;
        }
    } else {
        k2 = L.k2[idx];
    }
    exp_ = L.ex[idx];
    exm = L.ex[idx - 1];
    eyp = L.ey[idx];
    eym = L.ey[idx - sy];
    ezp = L.ez[idx];
    ezm = L.ez[idx - sz];
}

// In-place Gauss-Seidel / SOR update of interior node idx (the tail's smoother).
__device__ __forceinline__ void rb_node(const MgLevelDev &L, int64_t idx)
{
    const int64_t sy = L.nx, sz = (int64_t)L.nx * L.ny;
    float *__restrict__ e = L.e;
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
    const float old = e[idx];
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

// restrict_into (pb_solver.rs:578-624) for coarse node idx; with clear, also
// zeroes its correction (the tail smooths in place from zero).
__device__ __forceinline__ void restrict_node(const MgLevelDev &F, const MgLevelDev &C, int64_t idx,
                                              bool clear)
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
                // generic access: the tail kernel keeps small levels in LDS
                sum = fmaf(w, F.res[x + (int64_t)F.nx * (y + (int64_t)F.ny * z)], sum);
                ws += w;
            }
        }
    }
    C.f[idx] = 0.25f * sum / ws;
    if (clear) C.e[idx] = 0.0f;
}

// prolongate_add (pb_solver.rs:628-681): the coarse correction interpolated
// at fine node (fx, fy, fz); cells beyond the last coarse interval (even fine
// sizes) extrapolate from it
template <typename P>
__device__ __forceinline__ float prolong_value(const MgLevelDev &C, P ce, int fx, int fy, int fz)
{
    const int x0 = min(fx >> 1, max(C.nx - 2, 0));
    const int y0 = min(fy >> 1, max(C.ny - 2, 0));
    const int z0 = min(fz >> 1, max(C.nz - 2, 0));
    const float dx = 0.5f * fx - x0, dy = 0.5f * fy - y0, dz = 0.5f * fz - z0;
    const uint32_t cs = (uint32_t)C.nx, csz = (uint32_t)C.nx * (uint32_t)C.ny;
    const uint32_t o = (uint32_t)x0 + cs * (uint32_t)y0 + csz * (uint32_t)z0;
    const float c00 = ce[o] * (1.0f - dx) + ce[o + 1] * dx;
    const float c10 = ce[o + cs] * (1.0f - dx) + ce[o + cs + 1] * dx;
    const float c01 = ce[o + csz] * (1.0f - dx) + ce[o + csz + 1] * dx;
    const float c11 = ce[o + csz + cs] * (1.0f - dx) + ce[o + csz + cs + 1] * dx;
    const float c0 = c00 * (1.0f - dy) + c10 * dy;
    const float c1 = c01 * (1.0f - dy) + c11 * dy;
    return c0 * (1.0f - dz) + c1 * dz;
}

__device__ __forceinline__ void prolong_node(const MgLevelDev &C, const MgLevelDev &F, int64_t idx)
{
    int fx, fy, fz;
    split3(F, idx, fx, fy, fz);
    F.e[idx] += prolong_value(C, (const float *)C.e, fx, fy, fz);
}

__device__ __forceinline__ int64_t nodes(const MgLevelDev &L) { return (int64_t)L.nx * L.ny * L.nz; }

// ---- batched kernels (blockIdx.y = solve) -----------------------------------
// ---- tiled z-streaming smoother -------------------------------------------------
// One pass = one red-black sweep over a level, out of place.  Half-sweep k
// (k = 1 red, k = 2 black) of step t updates plane t - k + 1, so within a step
// both half-sweeps touch the nodes with x + y + t even, each reads only values
// the reference order has already produced, and plane z is final after step
// z + 1.  The red half-sweep runs on the region shrunk by one node, which is
// exact on the tile core widened by one; the black one is then exact on the
// core.  A thread owns column pairs (x, x+1) of the region; in every step
// exactly one column of each pair is active, alternating with t.
namespace sm {
constexpr int K = 2;                  // half-sweeps per step
constexpr int H = 2;                  // halo
constexpr int TILE = 28;              // core edge
constexpr int RW = TILE + 2 * H;      // region row: 32 words
constexpr int RN = RW * RW;
constexpr int RING = 4;               // planes t-2 .. t+1
constexpr int PF = RN / 256;          // plane values per thread
constexpr int NP = RN / 2;            // column pairs
constexpr int PP = NP / 256;          // column pairs per thread
constexpr int WB = (TILE * TILE + 255) / 256;  // core values written back per thread
static_assert(RN % 256 == 0 && NP % 256 == 0, "region must fill the workgroup");
}  // namespace sm

int smooth_tiles(int nx, int ny) { return ceil_div(nx, sm::TILE) * ceil_div(ny, sm::TILE); }

// About 48 planes per workgroup: long enough that the two extra planes a
// chunk loads stay cheap, short enough that large grids split into many
// workgroups whatever else shares the launch.
int smooth_zc(int nz)
{
    const int nch = std::max(1, (nz + 24) / 48);
    return ceil_div(nz, nch);
}

// face value: bit set -> cin, clear -> cout (a bitfield select on the bits)
__device__ __forceinline__ float face_sel(uint32_t w, int bit, uint32_t cin, uint32_t cout)
{
    const uint32_t m = (uint32_t)((int32_t)(w << (31 - bit)) >> 31);
    return __uint_as_float((m & cin) | (~m & cout));
}

// Loads are unconditional (lanes without work load from a harmless address and
// discard the value): loads under control flow make the compiler copy their
// results at the join and wait for every older load there.  The right-hand
// side and code bytes are loaded once per plane for both columns of a pair,
// coalesced, and kept in a three-plane register ring: plane z serves the red
// half-sweep of step z and the black half-sweep of step z + 1.
template <bool FIRST, bool PROLONG>
__global__ __launch_bounds__(256) void smooth_kernel(const MgLevelDev *__restrict__ Ls,
                                                     const MgLevelDev *__restrict__ Cs, int dir,
                                                     const MgStatus *__restrict__ st,
                                                     const uint32_t *__restrict__ wl)
{
    using namespace sm;
    __shared__ float ring[RING * RN];
    const BatchBlock bb = batch_block(wl);
    const int s = bb.s, b = bb.b;
    if (stopped(st + s)) return;
    const MgLevelDev L = Ls[s];
    if (L.nx == 0) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz, zc = L.zc;
    const int ntx = (nx + TILE - 1) / TILE, nty = (ny + TILE - 1) / TILE;
    const int nch = (nz + zc - 1) / zc;
    if (b >= ntx * nty * nch) return;
    const int tile = b % (ntx * nty), ch = b / (ntx * nty);
    const int cx = (nx + ntx - 1) / ntx, cy = (ny + nty - 1) / nty;
    const int x0 = (tile % ntx) * cx, x1 = min(nx, x0 + cx);
    const int y0 = (tile / ntx) * cy, y1 = min(ny, y0 + cy);
    const int z0 = ch * zc, z1 = min(nz, z0 + zc);
    const int xl = x0 - H, yl = y0 - H;
    // element offsets fit 32 bits (mg_smooth checks), so every load is a
    // 32-bit vector offset from a scalar base
    const uint32_t sxy = (uint32_t)nx * (uint32_t)ny;
    const bool uni = L.uni != 0;
    // non-uniform levels always carry code bytes; uniform ones read none, and
    // their dummy loads go to f
    const auto src = to_global<const float>(dir ? L.e : L.e2);
    const auto dst = to_global<float>((FIRST || dir) ? L.e2 : L.e);
    const auto fg = to_global<const float>(L.f);
    MgLevelDev C{};
    if (PROLONG) C = Cs[s];
    const auto ceg = to_global<const float>(PROLONG ? C.e : L.f);
    const auto cwg = to_global<const uint8_t>(uni ? (const uint8_t *)L.f : L.cw);
    const auto exg = to_global<const float>(uni ? L.f : L.ex);
    const auto eyg = to_global<const float>(uni ? L.f : L.ey);
    const auto ezg = to_global<const float>(uni ? L.f : L.ez);
    const float ih = L.ihx;  // the grids are cubic: ihx == ihy == ihz
    const float omega = L.omega, ck2 = L.c_k2;
    const uint32_t cin = __float_as_uint(L.c_in), cout = __float_as_uint(L.c_out);
    const float uih = L.eps_u * ih, udg_inv = 1.0f / (6.0f * L.eps_u * ih);
    const int tid = threadIdx.x;
    const int t0 = z0 - K + 1;
    const int t1 = z1 + K - 1;  // last step: writes plane z1 - 1

    // Column pair m: region row and first column, in-plane offset of the pair,
    // which columns exist in the grid, and per step parity q = (t - t0) & 1 the
    // active column (0 / 1) and whether it is interior to grid and region.
    int pri[PP], pgo[PP];
    bool pin0[PP], pin1[PP], act1[2][PP], ok[2][PP];
#pragma unroll
    for (int m = 0; m < PP; ++m) {
        const int P = tid + 256 * m;
        const int ry = P / (RW / 2);
        const int rx0 = 2 * (P - ry * (RW / 2));
        const int y = yl + ry, x = xl + rx0;
        pri[m] = ry * RW + rx0;
        pgo[m] = x + nx * y;
        const bool yin = y >= 0 && y < ny;
        pin0[m] = yin && x >= 0 && x < nx;
        pin1[m] = yin && x + 1 >= 0 && x + 1 < nx;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int c = (y + t0 + q - xl) & 1;
            const int rx = rx0 + c, xa = x + c;
            act1[q][m] = c != 0;
            ok[q][m] = ry >= 1 && ry <= RW - 2 && rx >= 1 && rx <= RW - 2 && xa >= 1 &&
                       xa <= nx - 2 && y >= 1 && y <= ny - 2;
        }
    }
    // plane loads: region value n = tid + 256 j sits in region row
    // (tid >> 5) + 8 j, column tid & 31
    static_assert(RW == 32, "plane load layout assumes 32-word rows");
    const int lx = xl + (tid & 31), ly = yl + (tid >> 5);
    const bool lxok = lx >= 0 && lx < nx;
    const int lgo0 = lx + nx * ly;
    // write-back: core value n = tid + 256 j at core row n / TILE
    const int wx = x1 - x0, wy = y1 - y0;

    auto slot = [&](int z) { return ring + ((z + 2 * RING) & (RING - 1)) * RN; };
    float pf[PF];
    auto load_plane = [&](int z) {
        if (FIRST) {
#pragma unroll
            for (int j = 0; j < PF; ++j) pf[j] = 0.0f;
            return;
        }
        const bool zin = z >= 0 && z < nz;
        const uint32_t zo = sxy * (uint32_t)(zin ? z : 0);
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            const int y = ly + 8 * j;
            const bool v = zin && lxok && y >= 0 && y < ny;
            float x = src[v ? zo + (uint32_t)(lgo0 + 8 * j * nx) : 0u];
            if (PROLONG)  // in-range indices for lanes without a node
                x += prolong_value(C, ceg, v ? lx : 0, v ? y : 0, v ? z : 0);
            pf[j] = v ? x : 0.0f;
        }
    };
    auto store_plane = [&](int z) {
        float *p = slot(z);
#pragma unroll
        for (int j = 0; j < PF; ++j) p[tid + 256 * j] = pf[j];
    };
    auto write_back = [&](int z) {
        if (z < z0 || z >= z1) return;
        const float *p = slot(z);
        const uint32_t zo = sxy * (uint32_t)z;
#pragma unroll
        for (int j = 0; j < WB; ++j) {
            const int n = tid + 256 * j;
            const int r = n / TILE, c = n - r * TILE;
            if (r < wy && c < wx)
                dst[zo + (uint32_t)((x0 + c) + nx * (y0 + r))] = p[(y0 + r - yl) * RW + (x0 + c - xl)];
        }
    };
    auto zok = [&](int k, int z) {
        return z >= 1 && z <= nz - 2 && z >= z0 - (K - k) && z <= z1 - 1 + (K - k);
    };
    // f and code bytes of both columns of every pair at plane z
    struct PlaneCoef {
        float f0[PP], f1[PP];
        uint32_t w0[PP], w1[PP];
    };
    auto load_coef = [&](int z, PlaneCoef &Q) {
        const bool zin = z >= 1 && z <= nz - 2;
        const uint32_t zo = sxy * (uint32_t)(zin ? z : 0);
#pragma unroll
        for (int m = 0; m < PP; ++m) {
            const uint32_t g0 = (zin && pin0[m]) ? zo + (uint32_t)pgo[m] : 0u;
            const uint32_t g1 = (zin && pin1[m]) ? zo + (uint32_t)pgo[m] + 1u : 0u;
            Q.f0[m] = fg[g0];
            Q.f1[m] = fg[g1];
            Q.w0[m] = cwg[g0];
            Q.w1[m] = cwg[g1];
        }
    };
    // half-sweep k of step t (parity q) on plane z = t - k + 1 with that
    // plane's coefficients Q
    auto stage = [&](int k, int t, int q, const PlaneCoef &Q) {
        const int z = t - k + 1;
        if (!zok(k, z)) return;
        const float *__restrict__ pm = slot(z - 1);
        float *__restrict__ pc = slot(z);
        const float *__restrict__ pp = slot(z + 1);
#pragma unroll
        for (int m = 0; m < PP; ++m) {
            if (!ok[q][m]) continue;
            const bool c1 = act1[q][m];
            const int i = pri[m] + (c1 ? 1 : 0);
            const float fv = c1 ? Q.f1[m] : Q.f0[m];
            const float exx = pc[i + 1], emx = pc[i - 1], eyy = pc[i + RW], emy = pc[i - RW];
            const float ezz = pp[i], emz = pm[i], old = pc[i];
            float gs;
            if (uni) {
                const float sum = ((exx + emx) + (eyy + emy)) + (ezz + emz);
                gs = fmaf(sum, uih, fv) * udg_inv;
            } else {
                const uint32_t w = c1 ? Q.w1[m] : Q.w0[m];
                float a0, a1, a2, a3, a4, a5;
                if (!(w & kCodeSpecial)) {
                    a0 = face_sel(w, 0, cin, cout);
                    a1 = face_sel(w, 1, cin, cout);
                    a2 = face_sel(w, 2, cin, cout);
                    a3 = face_sel(w, 3, cin, cout);
                    a4 = face_sel(w, 4, cin, cout);
                    a5 = face_sel(w, 5, cin, cout);
                } else {  // rare: faces cut by the dielectric boundary
                    const uint32_t g = sxy * (uint32_t)z + (uint32_t)(pgo[m] + (c1 ? 1 : 0));
                    a0 = exg[g];
                    a1 = exg[g - 1];
                    a2 = eyg[g];
                    a3 = eyg[g - nx];
                    a4 = ezg[g];
                    a5 = ezg[g - sxy];
                }
                const float k2 = (w & kCodeKappa) ? ck2 : 0.0f;
                const float sum =
                    fmaf(a0, exx, fmaf(a1, emx, fmaf(a2, eyy, fmaf(a3, emy, fmaf(a4, ezz, a5 * emz)))));
                const float dg = fmaf(((a0 + a1) + (a2 + a3)) + (a4 + a5), ih, k2);
                gs = fmaf(sum, ih, fv) * __builtin_amdgcn_rcpf(dg);
            }
            pc[i] = fmaf(omega, gs - old, old);
        }
    };
    // one step: stage plane t+1 of e, write plane t-2 back, start the loads of
    // the coefficients of plane t+1 and of e at plane t+2, then the red
    // half-sweep on plane t (coefficients Qb) and the black one on plane t-1
    // (coefficients Qa)
    PlaneCoef Qa, Qb, Qn;
    auto step = [&](int t, int q) {
        store_plane(t + 1);
        write_back(t - K);
        load_coef(t + 1, Qn);
        load_plane(t + 2);
        __syncthreads();
        stage(1, t, q, Qb);
        __syncthreads();
        stage(2, t, q, Qa);
        __syncthreads();
        Qa = Qb;
        Qb = Qn;
    };
    load_plane(t0 - 1);
    store_plane(t0 - 1);
    load_plane(t0);
    store_plane(t0);
    load_coef(t0 - 1, Qa);
    load_coef(t0, Qb);
    load_plane(t0 + 1);
    for (int t = t0; t <= t1; t += 2) {
        step(t, 0);
        if (t + 1 <= t1) step(t + 1, 1);
    }
}

// Batched launch grid: blocks x solves, or the work list's total as a 1-D grid.
static inline dim3 batch_grid(int64_t blocks, int n_solves, const uint32_t *wl)
{
    return wl ? dim3((unsigned)blocks) : dim3((unsigned)blocks, (unsigned)n_solves);
}

void mg_smooth(const MgLevelDev *L, const MgLevelDev *C, int n_solves, int first, int dir,
               int max_blocks, const MgStatus *st, hipStream_t s, const uint32_t *wl)
{
    if (n_solves <= 0 || max_blocks <= 0) return;
    const dim3 grid = batch_grid(max_blocks, n_solves, wl);
    if (first)
        hipLaunchKernelGGL((smooth_kernel<true, false>), grid, dim3(256), 0, s, L, C, dir, st, wl);
    else if (C)
        hipLaunchKernelGGL((smooth_kernel<false, true>), grid, dim3(256), 0, s, L, C, dir, st, wl);
    else
        hipLaunchKernelGGL((smooth_kernel<false, false>), grid, dim3(256), 0, s, L, C, dir, st,
                           wl);
    LAUNCH_OK();
}

// ---- temporally blocked smoother: three red-black sweeps per pass -------------
// The schedule of smooth_kernel with K = 6 half-sweeps: half-sweep k of step t
// updates plane t - k + 1 (k odd red, k even black), so a pass reads the
// correction, the right-hand side and the code bytes once and writes the
// correction once for three sweeps.  With a halo of 6 the tile core is 20 x 20
// inside the 32 x 32 region; half-sweep k only has to be exact on the region
// shrunk by k nodes, so it skips the outer k - 1 rings.  In every step the
// active column of a pair (x, x + 1) is the one with x + y + t even, the same
// for all six half-sweeps.
//
// Coefficients: the right-hand side of both columns of a pair and their code
// bytes stay in registers for the six planes a step touches.  Faces whose
// seven code bits (six faces, kappa) say everything come from a 128-entry LDS
// table {a0..a5, 1/diag} built with the update's own expressions.  Nodes with a
// fractional face are compacted per wave at the start of each step; one lane
// per such node loads its six faces from the FP32 face arrays before the
// step's prefetches are issued, so waiting for them never waits for the
// prefetches, and writes {a0..a5, 1/diag} to a per-wave LDS slot.  A wave with
// more such nodes than slots loads the rest inside the half-sweep.
//
// Memory access is through buffer resources: the plane offset is a scalar
// offset, per-lane offsets are fixed for the whole pass, and lanes without a
// node use an offset past every array, for which buffer loads return zero and
// buffer stores do nothing.
namespace s3 {
constexpr int K = 6;                   // half-sweeps per pass
constexpr int H = K;                   // halo
constexpr int RING = 8;                // correction planes t-6 .. t+1
constexpr int TAB = 128;               // code-table entries
constexpr uint32_t kOob = 0x80000000u; // offset outside every array
constexpr int kRsrcWord3 = 0x00020000; // raw buffer, 32-bit data format
static_assert(RING >= K + 2 && (RING & (RING - 1)) == 0, "ring holds planes t-K .. t+1");
// Tiles and z-chunks cover the interior nodes 1 .. n - 2 of each axis (on
// 161^3, 8 x 8 tiles instead of 9 x 9); a tile on a face of the grid also
// writes back the boundary nodes beside its core.
__host__ __device__ constexpr int inner(int n) { return n > 2 ? n - 2 : 1; }

// Region geometry: RW columns x RH rows, a core of (RW - 2 H) x (RH - 2 H).
// Thread t < NP owns column pairs t and t + NP (rows t / PPR and t / PPR + PR,
// columns 2 (t % PPR) and the next); NT rounds NP up to whole waves, and the
// threads past NP own no pair (they load, store and write back only their
// share of the plane and write-back rectangle).
template <int RW_, int RH_, int SL_>
struct Geo {
    static constexpr int RW = RW_, RH = RH_, RN = RW * RH;
    static constexpr int SLOTS = SL_;                       // fractional-face slots per wave
    static constexpr int TX = RW - 2 * H, TY = RH - 2 * H;  // core
    static constexpr int PPR = RW / 2;                      // column pairs per row
    static constexpr int PR = RH / 2;                       // rows between a thread's pairs
    static constexpr int NP = RN / 4;                       // threads owning pairs
    static constexpr int NT = (NP + 63) / 64 * 64;          // threads
    static constexpr int NW = NT / 64;                      // waves
    static constexpr int PF = 4, PP = 2;                    // plane values, pairs per thread
    static constexpr int WX = TX + 2, WY = TY + 2;          // written nodes: core plus faces
    static constexpr int WB = (WX * WY + NT - 1) / NT;      // written back per thread
    // LDS bytes: the correction ring and the face entries (table + slots)
    static constexpr int LDS = RING * RN * 4 + 2 * (TAB + NW * SLOTS) * 16;
    // both pairs of a thread share the red-black parity of their rows
    static_assert(RW % 2 == 0 && PR % 2 == 0 && TAB <= NT, "region layout");
};
// 32 x 32 (20 x 20 core): 256 threads, 40 KB, 4 workgroups per CU.
// 44 x 44 (32 x 32 core): 512 threads (484 own pairs), 79.5 KB, 2 per CU.
// Both give 16 waves per CU.  (A 32 x 40 region with 320 threads measured 22%
// slower than 32 x 32 on 161^3: its 5-wave workgroups lowered the waves per CU.)
// The slots per wave fill what the LDS budget leaves: 32 in the small region
// (4 workgroups per CU, 40 KB each), 60 in the large one (2 per CU, 79.5 KB).
using Small = Geo<32, 32, 32>;
using Large = Geo<44, 44, 60>;

// Thread -> pair of the 44 x 44 region, row << 8 | pair column (0xffff: no
// pair).  The LDS has 32 banks of 4 bytes; a half-sweep's reads by one wave
// hit bank (44 r + 2 i + c_r) mod 32, where the active column c_r alternates
// with the row.  A wave spanning whole 22-pair rows puts up to three lanes on
// one bank (36% of LDS cycles were conflicts); 4 rows x 16 pairs cover every
// bank exactly twice.  Waves 0-4 take such blocks of rows 0-19; the other 164
// pairs go, one by one, to whichever of waves 5-7 has the lower count on the
// two banks the pair's active column meets (one per step parity).
struct PairMap {
    uint16_t at[512];
};
constexpr PairMap large_pair_map()
{
    PairMap m{};
    for (int t = 0; t < 512; ++t) m.at[t] = 0xffff;
    for (int w = 0; w < 5; ++w)
        for (int l = 0; l < 64; ++l)
            m.at[64 * w + l] = (uint16_t)(((4 * w + l / 16) << 8) | (l % 16));
    int cnt[3][2][32] = {};
    int fill[3] = {0, 0, 0};
    const int cap[3] = {64, 64, 36};
    auto place = [&](int r, int i) {
        int best = -1, bm = 0;
        for (int k = 0; k < 3; ++k) {
            if (fill[k] == cap[k]) continue;
            int mk = 0;
            for (int c0 = 0; c0 < 2; ++c0) {
                const int b = (44 * r + 2 * i + ((r + c0) & 1)) % 32;
                mk = cnt[k][c0][b] > mk ? cnt[k][c0][b] : mk;
            }
            if (best < 0 || mk < bm) {
                best = k;
                bm = mk;
            }
        }
        for (int c0 = 0; c0 < 2; ++c0) ++cnt[best][c0][(44 * r + 2 * i + ((r + c0) & 1)) % 32];
        m.at[64 * (5 + best) + fill[best]++] = (uint16_t)((r << 8) | i);
    };
    for (int r = 20; r < 22; ++r)
        for (int i = 0; i < 16; ++i) place(r, i);
    for (int r = 0; r < 22; ++r)
        for (int i = 16; i < 22; ++i) place(r, i);
    return m;
}
constexpr bool pair_map_is_bijection(const PairMap &m)
{
    bool seen[22][22] = {};
    int n = 0;
    for (int t = 0; t < 512; ++t) {
        if (m.at[t] == 0xffff) continue;
        const int r = m.at[t] >> 8, i = m.at[t] & 0xff;
        if (r >= 22 || i >= 22 || seen[r][i]) return false;
        seen[r][i] = true;
        ++n;
    }
    return n == 484;
}
static_assert(pair_map_is_bijection(large_pair_map()), "every pair of the large region owned once");
__device__ constexpr PairMap kLargePairs = large_pair_map();
static_assert(Small::NT == 256 && Large::NT == 512, "thread counts");
static_assert(4 * Small::LDS <= 160 * 1024 && 2 * Large::LDS <= 160 * 1024, "LDS per CU");
template <class G>
int tiles(int nx, int ny)
{
    return ceil_div(inner(nx), G::TX) * ceil_div(inner(ny), G::TY);
}
}  // namespace s3

int smooth3_threads(int rg) { return rg ? s3::Large::NT : s3::Small::NT; }

int smooth3_tiles(int nx, int ny, int rg)
{
    return rg ? s3::tiles<s3::Large>(nx, ny) : s3::tiles<s3::Small>(nx, ny);
}

int smooth3_blocks(int nx, int ny, int nz, int zc, int rg)
{
    return smooth3_tiles(nx, ny, rg) * ceil_div(s3::inner(nz), zc);
}

// Region of a level: the 44 x 44 region when its lanes per plane (tiles x
// threads, each thread updating two pairs per half-sweep) come to at most 90%
// of the 32 x 32 region's.  On 161^3 that is 25 x 512 against 64 x 256 lanes;
// 81^3 (9 x 512 against 16 x 256) and 41^3 keep the small region.
int smooth3_region(int nx, int ny)
{
    const int64_t small = (int64_t)s3::tiles<s3::Small>(nx, ny) * s3::Small::NT;
    const int64_t large = (int64_t)s3::tiles<s3::Large>(nx, ny) * s3::Large::NT;
    return 10 * large <= 9 * small ? 1 : 0;
}

// Planes per workgroup: about 96 (a chunk also streams 2 K - 2 = 10 planes it
// does not write), shorter when a launch of n_solves such levels would
// otherwise have fewer than 2048 workgroups (8 per CU) of the small region's
// size (1024 of the large one's), but at least 8.
// Measured on the 161^3 batch: chunks of ~96 planes (two per column) beat
// 64, 40 and 32; whole columns were no faster (the halo reads of
// neighbouring tiles then meet less often in L2).
int smooth3_zc(int nx, int ny, int nz, int n_solves, int rg)
{
    const int64_t tiles = (int64_t)smooth3_tiles(nx, ny, rg) * std::max(n_solves, 1);
    const int64_t floor = rg ? 1024 : 2048;
    int nch = std::max(1, (nz + 48) / 96);
    nch = std::max<int>(nch, (int)((floor + tiles - 1) / tiles));
    nch = std::min(nch, std::max(1, nz / 8));
    return ceil_div(nz, nch);
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             s3::kRsrcWord3);
}
__device__ __forceinline__ float bload(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so)
{
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ uint32_t bload8(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so)
{
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ double bload64(__amdgpu_buffer_rsrc_t r, uint32_t vo, uint32_t so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ void bstore64(__amdgpu_buffer_rsrc_t r, double v, uint32_t vo,
                                         uint32_t so)
{
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)vo, (int)so, 0);
}

// four waves per SIMD (128 VGPRs) for both geometries
template <bool FIRST, bool PROLONG, class G>
__global__ __launch_bounds__(G::NT, 4) void smooth3_kernel(const MgLevelDev *__restrict__ Ls,
                                                          const MgLevelDev *__restrict__ Cs,
                                                          int dir,
                                                          const MgStatus *__restrict__ st,
                                                          const uint32_t *__restrict__ wl)
{
    using namespace s3;
    constexpr int RW = G::RW, RH = G::RH, RN = G::RN, NT = G::NT, NW = G::NW, TX = G::TX,
                  TY = G::TY, PF = G::PF, PP = G::PP, PR = G::PR, PPR = G::PPR, NP = G::NP,
                  WX = G::WX, WB = G::WB, SLOTS = G::SLOTS;
    __shared__ float ring[RING * RN];
    // face entries {a0..a3}, {a4, a5, 1/diag, -}: rows 0..127 the code table,
    // then SLOTS per wave for the step's fractional-face nodes
    __shared__ float4 fac[2 * (TAB + NW * SLOTS)];
    float4 *const tab = fac;
    float4 *const spc = fac + 2 * TAB;
    const BatchBlock bb = batch_block(wl);
    const int s = bb.s, b = bb.b;
    if (stopped(st + s)) return;
    const MgLevelDev L = Ls[s];
    if (L.nx == 0) return;
    const int nx = L.nx, ny = L.ny, nz = L.nz, zc = L.zc;
    const int ix = inner(nx), iy = inner(ny), iz = inner(nz);
    const int ntx = (ix + TX - 1) / TX, nty = (iy + TY - 1) / TY;
    const int nch = (iz + zc - 1) / zc;
    if (b >= ntx * nty * nch) return;
    const int tile = b % (ntx * nty), ch = b / (ntx * nty);
    const int cx = (ix + ntx - 1) / ntx, cy = (iy + nty - 1) / nty;
    const int x0 = 1 + (tile % ntx) * cx, x1 = min(nx - 1, x0 + cx);
    const int y0 = 1 + (tile / ntx) * cy, y1 = min(ny - 1, y0 + cy);
    const int z0 = 1 + ch * zc, z1 = min(nz - 1, z0 + zc);
    // nodes written back: the core plus the boundary nodes beside it.  Those
    // are never updated but carry the source's values, which need not be zero:
    // below a level of even size the prolongation extrapolates onto the
    // boundary (prolongate_add, pb_solver.rs:628-681).
    const int wx0 = x0 == 1 ? 0 : x0, wx1 = x1 == nx - 1 ? nx : x1;
    const int wy0 = y0 == 1 ? 0 : y0, wy1 = y1 == ny - 1 ? ny : y1;
    const int wz0 = z0 == 1 ? 0 : z0, wz1 = z1 == nz - 1 ? nz : z1;
    const int xl = x0 - H, yl = y0 - H;
    const uint32_t sxy = (uint32_t)nx * (uint32_t)ny;
    const uint32_t N = sxy * (uint32_t)nz;  // mg_smooth3 checks 4 N < 2^31
    const bool uni = L.uni != 0;
    const auto src = buf_rsrc(dir ? L.e : L.e2, 4u * N);
    const auto dst = buf_rsrc((FIRST || dir) ? L.e2 : L.e, 4u * N);
    const auto frs = buf_rsrc(L.f, 4u * N);
    const auto crs = buf_rsrc(uni ? nullptr : L.cw, uni ? 0u : N);
    // fractional-face entries (device memory, 64-bit addressing; loaded as
    // native 4-vectors)
    typedef float f4 __attribute__((ext_vector_type(4)));
    const auto etab = to_global<const f4>(reinterpret_cast<const f4 *>(L.ent));
    MgLevelDev C{};
    if (PROLONG) C = Cs[s];
    const auto ceg = to_global<const float>(PROLONG ? C.e : L.f);
    const float ih = L.ihx;  // the grids are cubic: ihx == ihy == ihz
    const float omega = L.omega, ck2 = L.c_k2;
    const uint32_t cin = __float_as_uint(L.c_in), cout = __float_as_uint(L.c_out);
    const float uih = L.eps_u * ih, udg_inv = 1.0f / (6.0f * L.eps_u * ih);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int t0 = z0 - K + 1, t1 = z1 + K - 1;

    // code table: entry w (face bits 0-5, kappa bit 6) -> {a0..a5, 1/diag}
    if (!uni && tid < TAB) {
        const uint32_t w = (uint32_t)tid;
        const float a0 = face_sel(w, 0, cin, cout), a1 = face_sel(w, 1, cin, cout);
        const float a2 = face_sel(w, 2, cin, cout), a3 = face_sel(w, 3, cin, cout);
        const float a4 = face_sel(w, 4, cin, cout), a5 = face_sel(w, 5, cin, cout);
        const float k2 = (w & kCodeKappa) ? ck2 : 0.0f;
        const float dg = fmaf(((a0 + a1) + (a2 + a3)) + (a4 + a5), ih, k2);
        tab[2 * tid] = make_float4(a0, a1, a2, a3);
        tab[2 * tid + 1] = make_float4(a4, a5, __builtin_amdgcn_rcpf(dg), 0.0f);
    }

    // column pairs m: region row ry = (tid >> 4) + PR m, columns rx0, rx0 + 1
    // (threads without a pair: every offset out of range, ring distance 0)
    bool own;
    int ry0, rx0;
    if constexpr (RW == 44) {
        const uint32_t pm = kLargePairs.at[tid];
        own = pm != 0xffffu;
        ry0 = own ? (int)(pm >> 8) : 0;
        rx0 = own ? 2 * (int)(pm & 0xffu) : 0;
    } else {
        own = tid < NP;
        ry0 = own ? tid / PPR : 0;
        rx0 = own ? 2 * (tid % PPR) : 0;
    }
    const int li0 = ry0 * RW + rx0;  // LDS index of column 0 of pair 0 (pair 1: + PR RW)
    // plane loads: a thread loads and stores the four nodes of its own column
    // pairs, value j = 2 m + c at region row ry0 + PR m, column rx0 + c, so
    // that within a step it alone reads the plane it stores.  One offset per
    // pair, that of its second column (the first is 4 bytes below).  Region
    // nodes outside the grid are never read by an update nor written back, so
    // their values do not matter: only in-grid nodes need exact offsets, and
    // theirs are never negative (x >= 0, y >= 0).  Any other offset either
    // wraps past the array (the load returns zero) or reads another node.
    const int lx = xl + rx0, ly = yl + ry0;
    uint32_t pvo[PP];
#pragma unroll
    for (int m = 0; m < PP; ++m)
        pvo[m] = own ? 4u * (uint32_t)(lx + 1 + nx * (ly + PR * m)) : kOob;
    const int par = (xl + rx0 + yl + ry0) & 1;  // active column at step t: par ^ (t & 1)
    uint32_t cvo[PP];   // node index of column 0 (kOob when the pair leaves the grid)
    uint32_t rdpk = 0;  // byte 2m + c: ring distance of column c if it is updated, else 0
    // spread[c]: bit 2 (k - 1) + m set when half-sweep k updates column c of pair m
#pragma unroll
    for (int m = 0; m < PP; ++m) {
        const int ry = ry0 + PR * m, y = yl + ry, x = xl + rx0;
        const bool pin = own && y >= 0 && y < ny && x >= 0 && x + 1 < nx;
        cvo[m] = pin ? (uint32_t)(x + nx * y) : kOob;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int rx = rx0 + c, xa = x + c;
            const int rd = min(min(rx, RW - 1 - rx), min(ry, RH - 1 - ry));
            const bool in = own && xa >= 1 && xa <= nx - 2 && y >= 1 && y <= ny - 2;
            rdpk |= (uint32_t)(in ? rd : 0) << (8 * (2 * m + c));
        }
    }
    uint32_t spread[2] = {0u, 0u};
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 1; k <= K; ++k)
#pragma unroll
            for (int m = 0; m < PP; ++m)
                if (((rdpk >> (8 * (2 * m + c))) & 0xffu) >= (uint32_t)k)
                    spread[c] |= 1u << (2 * (k - 1) + m);
    // write-back: value n = tid + NT j at row n / WX of the written rectangle
    uint32_t wvo[WB];
    int wli[WB];
#pragma unroll
    for (int j = 0; j < WB; ++j) {
        const int n = tid + NT * j, r = n / WX, c = n - r * WX;
        const bool ok = r < wy1 - wy0 && c < wx1 - wx0;
        wvo[j] = ok ? 4u * (uint32_t)((wx0 + c) + nx * (wy0 + r)) : kOob;
        wli[j] = ok ? (wy0 + r - yl) * RW + (wx0 + c - xl) : 0;
    }

    auto slot = [&](int z) { return ring + ((z + 2 * RING) & (RING - 1)) * RN; };
    float pf[PF];
    auto load_plane = [&](int z) {
        if (FIRST || z < 0 || z >= nz) {
#pragma unroll
            for (int j = 0; j < PF; ++j) pf[j] = 0.0f;
            return;
        }
        const uint32_t zo = 4u * sxy * (uint32_t)z;
#pragma unroll
        for (int j = 0; j < PF; ++j) {
            const uint32_t o = pvo[j >> 1];
            float x = bload(src, (j & 1) ? o : o - 4u, zo);
            if (PROLONG) {
                const int gx = lx + (j & 1), gy = ly + PR * (j >> 1);
                const bool v = own && gx >= 0 && gx < nx && gy >= 0 && gy < ny;
                x += prolong_value(C, ceg, v ? gx : 0, v ? gy : 0, z);
                x = v ? x : 0.0f;
            }
            pf[j] = x;
        }
    };
    auto store_plane = [&](int z) {
        float *p = slot(z);
        if (own)
#pragma unroll
            for (int j = 0; j < PF; ++j) p[li0 + PR * RW * (j >> 1) + (j & 1)] = pf[j];
    };
    // finest level, post-smoothing: phi += e in FP64 instead of storing e (the
    // outer step's own addition).  At the write-back point of plane z the loads
    // of phi are issued; after the step's six half-sweeps, before its barrier,
    // the final values of plane z (unchanged by the step) are read from the
    // ring, added and stored (fold_finish), so the loads are not waited for and
    // no ring value is held across the half-sweeps.  The slot is refilled only
    // after the barrier.  Lanes outside the core are skipped by predicate (8 N
    // may exceed kOob).
    const bool fold = !FIRST && L.phi != nullptr;
    const auto phr = buf_rsrc(fold ? L.phi : nullptr, fold ? 8u * N : 0u);
    double phv[WB];
    int fold_z = -1;
    auto write_back = [&](int z) {
        if (z < wz0 || z >= wz1) return;
        const uint32_t zo = 4u * sxy * (uint32_t)z;
        if (fold) {
#pragma unroll
            for (int j = 0; j < WB; ++j)
                phv[j] = wvo[j] != kOob ? bload64(phr, 2u * wvo[j], 2u * zo) : 0.0;
            fold_z = z;
            return;
        }
        const float *p = slot(z);
#pragma unroll
        for (int j = 0; j < WB; ++j)
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(p[wli[j]]), dst, (int)wvo[j],
                                                  (int)zo, 0);
    };
    auto fold_finish = [&]() {
        if (!fold || fold_z < 0) return;
        const uint32_t zo = 4u * sxy * (uint32_t)fold_z;
        const float *p = slot(fold_z);
#pragma unroll
        for (int j = 0; j < WB; ++j) {
            const float ev = p[wli[j]];
            if (wvo[j] != kOob) bstore64(phr, phv[j] + (double)ev, 2u * wvo[j], 2u * zo);
        }
        fold_z = -1;
    };
    // coefficients of a plane: f of both columns of both pairs, code bytes packed
    struct Coef {
        float f[PP][2];
        uint32_t w;
    };
    struct Next {
        float f[PP][2];
        uint32_t b[PP][2];
    };
    auto load_coef = [&](int z, Next &Q, bool UNI) {
        // planes outside 1 .. nz - 2 are never updated: every offset out of range
        const bool zin = z >= 1 && z <= nz - 2;
        const uint32_t zo = zin ? sxy * (uint32_t)z : 0u;
#pragma unroll
        for (int m = 0; m < PP; ++m) {
            const uint32_t o = zin ? cvo[m] : kOob, fo = o == kOob ? kOob : 4u * o;
            Q.f[m][0] = bload(frs, fo, 4u * zo);
            Q.f[m][1] = bload(frs, fo + 4u, 4u * zo);
            Q.b[m][0] = UNI ? 0u : bload8(crs, o, zo);
            Q.b[m][1] = UNI ? 0u : bload8(crs, o + 1u, zo);
        }
    };
    auto pack = [](const Next &Q) {
        Coef c;
#pragma unroll
        for (int m = 0; m < PP; ++m) {
            c.f[m][0] = Q.f[m][0];
            c.f[m][1] = Q.f[m][1];
        }
        c.w = Q.b[0][0] | (Q.b[0][1] << 8) | (Q.b[1][0] << 16) | (Q.b[1][1] << 24);
        return c;
    };
    auto zok = [&](int k, int z) {
        return z >= 1 && z <= nz - 2 && z >= z0 - (K - k) && z <= z1 - 1 + (K - k);
    };

    using Fast = std::integral_constant<int, 0>;
    using Slow = std::integral_constant<int, 1>;
    using NoSpec = std::integral_constant<int, 2>;
    uint32_t *req = reinterpret_cast<uint32_t *>(spc) + wv * SLOTS * 8;

    // the step loop, compiled separately for uniform and non-uniform levels
    auto run = [&](auto uni_tag) {
        constexpr bool UNI = decltype(uni_tag)::value;
        // entry indices packed in bytes: only when every entry (table rows,
        // then SLOTS per wave) is below 256, i.e. the 32 x 32 region
        constexpr bool ENT = FIRST && TAB + NW * SLOTS <= 256;
        Coef Q[K];  // Q[i]: plane t - i
        Next nxt;

        // Fractional-face nodes of step t.  Update j = 2 (k - 1) + m (half-
        // sweep k, pair m) of a lane is one when bit j of spm is set; the wave
        // numbers them lane by lane (exclusive prefix sum of the lane counts,
        // spx) and in bit order within a lane, so update j uses slot spx +
        // popcount of the lower bits.  Slots past SLOTS are read in the
        // half-sweep itself.
        uint32_t spm = 0;
        int spx = 0, total = 0;
        bool nospec = false;  // no code word of the wave's six planes has the fractional bit
        f4 sa, sb;  // the entry a lane loads for its slot
        uint32_t sreq = 0xffffffffu;
        // Rolling mask of the window's fractional-face bits, column 0 in bits
        // 0-11 and column 1 in bits 16-27: bit 2 (k - 1) + m is bit 7 of the
        // code byte of pair m's column at plane t - k + 1.  Each step shifts it
        // by one plane and inserts the new plane's bits (roll), so the
        // numbering reads it instead of reassembling twelve bits from the six
        // code words.
        uint32_t fmask = 0u;
        auto roll = [&](uint32_t w) {
            // code bytes of w: pair 0 column 0, pair 0 column 1, pair 1 column 0,
            // pair 1 column 1
            fmask = ((fmask << 2) & 0x0ffc0ffcu) | ((w >> 7) & 1u) | ((w >> 22) & 2u) |
                    ((w << 1) & 0x10000u) | ((w >> 14) & 0x20000u);
        };
        // face entries of the step's updates (ENT): byte j & 3 of ent[j >> 2]
        // is the entry of update j = 2 (k - 1) + m, a table row or the wave's
        // slot TAB + wave * SLOTS + slot, valid outside the SLOW mode
        uint32_t ent[3] = {0u, 0u, 0u};
        auto specials_issue = [&](int t, bool all) {
            // bytes c and 2 + c (the active column of both pairs) of the code
            // words of two consecutive planes, kappa and face bits only
            const uint32_t sel = 0x06040200u + 0x01010101u * (uint32_t)(par ^ (t & 1));
            if (ENT)
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    ent[i] = __builtin_amdgcn_perm(Q[2 * i + 1].w, Q[2 * i].w, sel) & 0x7f7f7f7fu;
            nospec = __ballot(fmask != 0u) == 0;
            if (nospec) {
                spm = 0;
                spx = 0;
                total = 0;
                sreq = 0xffffffffu;
                return;
            }
            const int c = par ^ (t & 1);
            uint32_t zmask = all ? (1u << (2 * K)) - 1u : 0u;  // half-sweeps with a plane
            if (!all)
#pragma unroll
                for (int k = 1; k <= K; ++k)
                    zmask |= zok(k, t - k + 1) ? 3u << (2 * (k - 1)) : 0u;
            // bit j: code bit 7 (fractional faces) of update j
            const uint32_t hb = (c ? fmask >> 16 : fmask) & 0xfffu;
            spm = hb & zmask & (c ? spread[1] : spread[0]);
            // exclusive prefix sum of the lane counts (0..12) and the wave
            // total from one ballot per count bit: mbcnt gives the set bits
            // below the lane, so no cross-lane shuffles (LDS round trips)
            const uint32_t cnt = (uint32_t)__popc(spm);
            int pre = 0, tot = 0;
#pragma unroll
            for (int bit = 0; bit < 4; ++bit) {
                const uint64_t bm = __ballot((cnt >> bit) & 1u);
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(bm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bm, 0u));
                pre += (int)below << bit;
                tot += __popcll(bm) << bit;
            }
            spx = pre;
            total = __builtin_amdgcn_readfirstlane(tot);
            if (total == 0) return;
            // requests: node index of each numbered update
            const uint32_t a0 = cvo[0] + (uint32_t)c + sxy * (uint32_t)(t + 1);
            const uint32_t a1 = cvo[1] + (uint32_t)c + sxy * (uint32_t)(t + 1);
            uint32_t rem = spm;
            int sl = spx;
            while (__ballot(rem != 0u)) {
                if (rem) {
                    const int j = __ffs(rem) - 1;
                    rem &= rem - 1u;
                    const uint32_t g = ((j & 1) ? a1 : a0) - sxy * (uint32_t)((j >> 1) + 1);
                    if (sl < SLOTS) {
                        req[sl] = g;
                        if (ENT) {
                            const uint32_t sh = 8u * (uint32_t)(j & 3);
                            const uint32_t v = (uint32_t)(TAB + wv * SLOTS + sl) << sh;
                            const uint32_t keep = ~(0xffu << sh);
                            const int wi = j >> 2;
                            ent[0] = wi == 0 ? (ent[0] & keep) | v : ent[0];
                            ent[1] = wi == 1 ? (ent[1] & keep) | v : ent[1];
                            ent[2] = wi == 2 ? (ent[2] & keep) | v : ent[2];
                        }
                    }
                    ++sl;
                }
            }
            // one lane per numbered update loads its precomputed entry
            sreq = lane < min(total, SLOTS) ? req[lane] : 0xffffffffu;
            if (sreq != 0xffffffffu) {
                sa = etab[2 * (size_t)sreq];
                sb = etab[2 * (size_t)sreq + 1];
            }
        };
        auto specials_store = [&]() {
            if (total == 0 || sreq == 0xffffffffu) return;
            spc[2 * (wv * SLOTS + lane)] = make_float4(sa.x, sa.y, sa.z, sa.w);
            spc[2 * (wv * SLOTS + lane) + 1] = make_float4(sb.x, sb.y, sb.z, sb.w);
        };

        // per step: the active column's ring distance and the byte offset,
        // within a plane, of the node one row above it (so every in-plane
        // neighbour is a non-negative offset from one address)
        uint32_t rdm[PP], ibm[PP];
        auto step_cols = [&](int t) {
            const int c = par ^ (t & 1);
#pragma unroll
            for (int m = 0; m < PP; ++m) {
                rdm[m] = (rdpk >> (8 * (2 * m + c))) & 0xffu;
                // (clamped for region rows 0 and RH - 1, whose lanes never
                // update, so that their rows -1 .. +1 stay inside the slot)
                ibm[m] = 4u * (uint32_t)min(max(li0 + PR * RW * m + c - RW, 0),
                                            (RH - 3) * RW + RW - 1);
            }
        };
        // half-sweep k of step t on plane t - k + 1.  MODE 0: fast path; 1:
        // SLOW (the wave has more fractional-face updates this step than
        // slots); 2: the wave has no fractional-face node in its six planes
        // (table only)
        // all: every half-sweep of step t has its plane in range (no test;
        // uniform, so the scalar unit branches over the tests)
        auto stage = [&](int k, int t, auto mode_tag, bool all) {
            constexpr int MODE = decltype(mode_tag)::value;
            constexpr bool SLOW = MODE == 1, NOSPEC = MODE == 2;
            const int z = t - k + 1;
            if (!all && !zok(k, z)) return;
            const int c = par ^ (t & 1);
            const char *ringb = reinterpret_cast<const char *>(ring);
            const uint32_t sc = 4u * (uint32_t)(slot(z) - ring);
            const uint32_t sm = 4u * (uint32_t)(slot(z - 1) - ring) + 4u * RW;
            const uint32_t sp = 4u * (uint32_t)(slot(z + 1) - ring) + 4u * RW;
            const Coef &q = Q[k - 1];
            // both updates of the thread are computed unconditionally (reads
            // of in-region addresses only) and only the stores are predicated,
            // so their loads and fma chains overlap
            float nv[PP];
            float *dp[PP];
            bool act[PP];
#pragma unroll
            for (int m = 0; m < PP; ++m) {
                const int jb = 2 * m + c;
                act[m] = rdm[m] >= (uint32_t)k;
                const float *r = reinterpret_cast<const float *>(ringb + sc + ibm[m]);
                const float fv = c ? q.f[m][1] : q.f[m][0];
                // pre-smoothing starts from zero: at its first half-sweep a
                // node and all six neighbours are still +0, so the neighbour
                // sum is +0 (the same bits as summing them)
                constexpr bool ZERO = FIRST;
                const bool zs = ZERO && k == 1;
                const float old = r[RW];
                float emy = 0.0f, emx = 0.0f, exx = 0.0f, eyy = 0.0f, ezz = 0.0f, emz = 0.0f;
                if (!zs) {
                    emy = r[0];
                    emx = r[RW - 1];
                    exx = r[RW + 1];
                    eyy = r[2 * RW];
                    ezz = *reinterpret_cast<const float *>(ringb + sp + ibm[m]);
                    emz = *reinterpret_cast<const float *>(ringb + sm + ibm[m]);
                }
                float *pc_i = const_cast<float *>(r) + RW;
                float gs;
                if (UNI) {
                    const float sum = zs ? 0.0f : ((exx + emx) + (eyy + emy)) + (ezz + emz);
                    gs = fmaf(sum, uih, fv) * udg_inv;
                } else {
                    const uint32_t w = (q.w >> (8 * jb)) & 0xffu;
                    const int j = 2 * (k - 1) + m;
                    // SLOW: a lane that does not update reads its table row
                    // (its slot number is not one of the wave's); slots past
                    // SLOTS load their faces below
                    constexpr bool OLD = SLOW || !ENT;
                    const int sl = OLD && !NOSPEC ? spx + __popc(spm & ((1u << j) - 1u)) : 0;
                    const bool sp = OLD && !NOSPEC && act[m] && w >= kCodeSpecial;
                    const uint32_t se = (uint32_t)(TAB + wv * SLOTS) +
                                        (uint32_t)(SLOW ? min(sl, SLOTS - 1) : sl);
                    // otherwise the entry numbered at the start of the step
                    const uint32_t ei =
                        OLD ? (sp ? se : (w & 0x7fu)) : (ent[j >> 2] >> (8 * (j & 3))) & 0xffu;
                    const float4 *e = fac + 2u * ei;
                    float4 A = e[0], B = e[1];
                    if (SLOW && sp && sl >= SLOTS) {
                        const size_t g = cvo[m] + (uint32_t)c + sxy * (uint32_t)z;
                        const f4 a = etab[2 * g], b = etab[2 * g + 1];
                        A = make_float4(a.x, a.y, a.z, a.w);
                        B = make_float4(b.x, b.y, b.z, b.w);
                    }
                    const float sum = zs ? 0.0f : fmaf(A.x, exx, fmaf(A.y, emx, fmaf(A.z, eyy,
                                      fmaf(A.w, emy, fmaf(B.x, ezz, B.y * emz)))));
                    gs = fmaf(sum, ih, fv) * B.z;
                }
                dp[m] = pc_i;
                nv[m] = fmaf(omega, gs - old, old);
            }
#pragma unroll
            for (int m = 0; m < PP; ++m)
                if (act[m]) *dp[m] = nv[m];
        };

        // prologue: planes t0 - 1 and t0 in the ring, coefficients of t0 and
        // plane t0 + 1 in flight
        load_plane(t0 - 1);
        store_plane(t0 - 1);
        load_plane(t0);
        store_plane(t0);
        load_coef(t0, nxt, UNI);
        load_plane(t0 + 1);
        // the code table and the first two planes are read by every wave
        __syncthreads();
        // steps ts .. te have all six half-sweeps in range: those run without
        // the per-half-sweep range tests
        const int ts = max(t0, max(K, z0 + K - 1)), te = min(t1, min(nz - 2, z1 + K - 2));
        auto steady = [&](int t) { return t >= ts && t <= te; };
        // the coefficient window of step t (planes t .. t - 5), its active
        // columns, and its fractional faces numbered with their loads issued
        auto advance = [&](int t) {
#pragma unroll
            for (int i = K - 1; i > 0; --i) Q[i] = Q[i - 1];
            Q[0] = pack(nxt);
            step_cols(t);
            if (!UNI) {
                roll(Q[0].w);
                specials_issue(t, steady(t));
            }
        };
        advance(t0);
        auto step = [&](int t, bool all_tag) {
            store_plane(t + 1);
            // order of issue: write-back, prefetches; the step's fractional
            // faces were issued before them, at the end of the previous step,
            // so waiting for the faces waits for nothing younger
            write_back(t - K);
            if (t < t1) {
                load_coef(t + 1, nxt, UNI);
                load_plane(t + 2);
            }
            // One barrier per step.  Every half-sweep of step t updates only
            // nodes with x + y + t even: their in-plane neighbours have the
            // other parity and are not written during the step, and their z
            // neighbours lie in the thread's own column, written by its own
            // earlier half-sweeps or by its own store of plane t + 1.  The
            // barrier follows the last half-sweep: the next step's in-plane
            // neighbours and write-back read its values, and the next plane
            // store refills the slot of plane t - 6, which this step's
            // write-back and last half-sweep read.  Before it, the next
            // step's window moves in and its fractional faces are requested,
            // so their loads overlap the potential update, the barrier and the
            // next step's stores and prefetches.
            if (UNI) {
#pragma unroll
                for (int k = 1; k <= K; ++k) stage(k, t, Fast{}, all_tag);
                if (t < t1) advance(t + 1);
                fold_finish();
                __syncthreads();
                return;
            }
            specials_store();
            // the branches are uniform per wave: a wave with more fractional-
            // face updates than slots this step takes the SLOW path (its
            // slots are its own, written and read by its own lanes)
            if (total > SLOTS) {
#pragma unroll
                for (int k = 1; k <= K; ++k) stage(k, t, Slow{}, all_tag);
            } else if (nospec) {
#pragma unroll
                for (int k = 1; k <= K; ++k) stage(k, t, NoSpec{}, all_tag);
            } else {
#pragma unroll
                for (int k = 1; k <= K; ++k) stage(k, t, Fast{}, all_tag);
            }
            if (t < t1) advance(t + 1);
            fold_finish();
            __syncthreads();
        };
        for (int t = t0; t <= t1; ++t) step(t, steady(t));
        // the grid's last plane (never updated, still in the ring) follows
        // the last step
        if (wz1 > z1) {
            write_back(z1);
            fold_finish();
        }
    };
    if (uni)
        run(std::integral_constant<bool, true>{});
    else
        run(std::integral_constant<bool, false>{});
}

// e2 += prolongation of the coarse correction (the start of post-smoothing)
__global__ __launch_bounds__(256) void prolong_e2_kernel(const MgLevelDev *__restrict__ Fs,
                                                         const MgLevelDev *__restrict__ Cs,
                                                         const MgStatus *__restrict__ st,
                                                         const uint32_t *__restrict__ wl)
{
    const BatchBlock bb = batch_block(wl);
    const int s = bb.s;
    if (stopped(st + s)) return;
    MgLevelDev F = Fs[s];
    if (F.nx == 0) return;
    const int64_t idx = (int64_t)bb.b * blockDim.x + threadIdx.x;
    if (idx >= nodes(F)) return;
    const MgLevelDev C = Cs[s];
    F.e = F.e2;
    prolong_node(C, F, idx);
}

void mg_prolong_e2(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int64_t max_nodes,
                   const MgStatus *st, hipStream_t s, const uint32_t *wl)
{
    if (n_solves <= 0 || max_nodes <= 0) return;
    const dim3 grid = batch_grid(wl ? max_nodes : ceil_div(max_nodes, 256), n_solves, wl);
    hipLaunchKernelGGL(prolong_e2_kernel, grid, dim3(256), 0, s, F, C, st, wl);
    LAUNCH_OK();
}

void mg_smooth3(const MgLevelDev *L, const MgLevelDev *C, int n_solves, int first, int dir,
                int max_blocks, int rg, const MgStatus *st, hipStream_t s, const uint32_t *wl)
{
    if (n_solves <= 0 || max_blocks <= 0) return;
    const dim3 grid = batch_grid(max_blocks, n_solves, wl);
    // the prolongation runs as its own kernel (mg_prolong_e2) before this pass
    if (C) throw DeviceError(1, "mg_smooth3: no fused prolongation");  // invalid argument
    auto launch = [&](auto geo) {
        using G = decltype(geo);
        if (first)
            hipLaunchKernelGGL((smooth3_kernel<true, false, G>), grid, dim3(G::NT), 0, s, L, C,
                               dir, st, wl);
        else
            hipLaunchKernelGGL((smooth3_kernel<false, false, G>), grid, dim3(G::NT), 0, s, L, C,
                               dir, st, wl);
    };
    if (rg)
        launch(s3::Large{});
    else
        launch(s3::Small{});
    LAUNCH_OK();
}

// ---- z-streaming per-node kernels ---------------------------------------------
// One thread per in-plane index p = x + nx y (lanes contiguous in memory) over a
// range of L.zc planes.  Its z neighbours come from a register window; the
// in-plane neighbours, the right-hand side and the code byte of plane z + 1 are
// loaded one plane ahead, and loads at fractional-face nodes are issued before
// each plane's prefetches, so waiting for them never waits for the prefetches.
// Each kernel evaluates the same expressions as its per-node counterpart
// (residual_node, coeffs, prolong_value), so results are equal bit for bit.
__device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, float v, uint32_t vo, uint32_t so)
{
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)vo, (int)so, 0);
}

// workgroups of one solve: ceil(nx ny / 256) in-plane blocks times ceil(nz / zc) chunks
__host__ __device__ inline int stream_blocks(int nx, int ny, int nz, int zc)
{
    return ((nx * ny + 255) / 256) * ((nz + zc - 1) / zc);
}

struct StreamPos {
    int p, x, y, z0, z1;
};
// false: this thread has no column (past the plane or past the chunks)
__device__ __forceinline__ bool stream_pos(int nx, int ny, int nz, int zc, int b, StreamPos &q)
{
    const int sxy = nx * ny, pb = (sxy + 255) / 256;
    const int ch = b / pb;
    q.p = (b - ch * pb) * 256 + (int)threadIdx.x;
    q.z0 = ch * zc;
    q.z1 = min(nz, q.z0 + zc);
    if (q.p >= sxy || q.z0 >= nz) return false;
    q.y = q.p / nx;
    q.x = q.p - q.y * nx;
    return true;
}

// prolong_e2_kernel: e2 += prolongation of the coarse correction
__global__ __launch_bounds__(256) void prolong_e2s_kernel(const MgLevelDev *__restrict__ Fs,
                                                          const MgLevelDev *__restrict__ Cs,
                                                          const MgStatus *__restrict__ st,
                                                          const uint32_t *__restrict__ wl)
{
    const BatchBlock bb = batch_block(wl);
    if (stopped(st + bb.s)) return;
    const MgLevelDev F = Fs[bb.s];
    if (F.nx == 0) return;
    const MgLevelDev C = Cs[bb.s];
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    StreamPos q;
    if (bb.b >= stream_blocks(nx, ny, nz, F.zc) || !stream_pos(nx, ny, nz, F.zc, bb.b, q)) return;
    const uint32_t sxy = (uint32_t)nx * (uint32_t)ny, N = sxy * (uint32_t)nz;
    const uint32_t cs = (uint32_t)C.nx, csz = (uint32_t)C.nx * (uint32_t)C.ny;
    const auto er = buf_rsrc(F.e2, 4u * N);
    const auto cr = buf_rsrc(C.e, 4u * csz * (uint32_t)C.nz);
    // prolong_value's cell and weights in x and y are fixed for the column
    const int x0 = min(q.x >> 1, max(C.nx - 2, 0)), y0 = min(q.y >> 1, max(C.ny - 2, 0));
    const float dx = 0.5f * q.x - x0, dy = 0.5f * q.y - y0;
    const uint32_t co = 4u * ((uint32_t)x0 + cs * (uint32_t)y0);
    // coarse plane cz interpolated in x and y (c0 / c1 of prolong_value), from
    // its four corner values
    auto craw = [&](int cz, float (&v)[4]) {
        const bool ok = cz < C.nz;
        const uint32_t zo = ok ? 4u * csz * (uint32_t)cz : 0u;
        v[0] = bload(cr, ok ? co : s3::kOob, zo);
        v[1] = bload(cr, ok ? co + 4u : s3::kOob, zo);
        v[2] = bload(cr, ok ? co + 4u * cs : s3::kOob, zo);
        v[3] = bload(cr, ok ? co + 4u * cs + 4u : s3::kOob, zo);
    };
    auto cmix = [&](const float (&v)[4]) {
        const float c00 = v[0] * (1.0f - dx) + v[1] * dx;
        const float c10 = v[2] * (1.0f - dx) + v[3] * dx;
        return c00 * (1.0f - dy) + c10 * dy;
    };
    auto cplane = [&](int cz) {
        float v[4];
        craw(cz, v);
        return cmix(v);
    };
    const uint32_t po = 4u * (uint32_t)q.p, zs = 4u * sxy;
    // A thread owns its column, so its loads and stores never alias: each fine
    // plane is loaded D planes ahead of its store (D loads in flight instead
    // of one), and the corners of coarse plane z0c + 1 + CR when c_hi moves to
    // plane z0c + 1, so every wait is for a load issued about D fine planes
    // earlier (vector memory loads retire in order: a wait for a recent load
    // would also wait for the prefetches before it).  The arithmetic is
    // unchanged.
    constexpr int D = 8, CR = D / 2;
    float fb[D];
#pragma unroll
    for (int i = 0; i < D; ++i) fb[i] = q.z0 + i < q.z1 ? bload(er, po, zs * (uint32_t)(q.z0 + i)) : 0.0f;
    int zc_lo = -2;  // the first plane computes both coarse planes directly
    float c_lo = 0.0f, c_hi = 0.0f;
    float nv[CR][4];  // corners of coarse planes zc_lo + 2 .. zc_lo + 1 + CR
    for (int zb = q.z0; zb < q.z1; zb += D) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            const int z = zb + i;
            if (z >= q.z1) break;
            const float cur = fb[i];
            fb[i] = z + D < q.z1 ? bload(er, po, zs * (uint32_t)(z + D)) : 0.0f;
            const int z0c = min(z >> 1, max(C.nz - 2, 0));
            if (z0c != zc_lo) {
                if (z0c == zc_lo + 1) {
                    c_lo = c_hi;
                    c_hi = cmix(nv[0]);  // plane z0c + 1
#pragma unroll
                    for (int r = 0; r + 1 < CR; ++r)
#pragma unroll
                        for (int k = 0; k < 4; ++k) nv[r][k] = nv[r + 1][k];
                    craw(z0c + 1 + CR, nv[CR - 1]);
                } else {
                    c_lo = cplane(z0c);
                    c_hi = cplane(z0c + 1);
#pragma unroll
                    for (int r = 0; r < CR; ++r) craw(z0c + 2 + r, nv[r]);
                }
                zc_lo = z0c;
            }
            const float dz = 0.5f * z - z0c;
            bstore(er, cur + (c_lo * (1.0f - dz) + c_hi * dz), po, zs * (uint32_t)z);
        }
    }
}

// Residual and restriction in one pass (residual32_kernel then
// restrict_kernel without the residual array).  A workgroup owns a 16 x 16
// tile of interior coarse columns and a chunk of C.zc interior coarse planes
// (a boundary node's right-hand side is never read), and streams the
// fine planes 2 cz0 - 1 .. 2 cz1 - 1: the pre-smoothed correction e2 goes
// through a 4-plane LDS ring (35 x 35 nodes, loaded two planes ahead), each
// fine plane's residual on the 33 x 33 nodes the tile restricts from is
// written to LDS (double-buffered, one barrier per plane), and every coarse
// column extends its fma chain in restrict_node's order and skips, so both
// the residuals and the interior coarse right-hand sides equal the two
// kernels' bit for bit.
namespace rr {
constexpr int CT = 16;               // coarse tile edge
constexpr int FR = 2 * CT + 1;       // fine residual region edge: 33
constexpr int ER = FR + 2;           // correction region edge: 35
constexpr int NE = (ER * ER + 255) / 256;  // correction values per thread: 5
constexpr int NR = (FR * FR + 255) / 256;  // residual nodes per thread: 5
// residual plane in LDS: row stride RS, even columns at 0 .. 16 and odd ones at
// 17 .. 32, so a coarse column's reads are consecutive across lanes and rows
// two apart land 16 banks apart
constexpr int RS = 40, ODD = (FR + 1) / 2;
__host__ __device__ constexpr int rpos(int ry, int rx) { return ry * RS + ((rx & 1) ? ODD : 0) + (rx >> 1); }
}  // namespace rr

// Tiles and chunks cover the interior coarse nodes only: nothing reads a
// right-hand side on a level's boundary.
int mg_rr_blocks(int cnx, int cny, int cnz, int czc)
{
    return ceil_div(s3::inner(cnx), rr::CT) * ceil_div(s3::inner(cny), rr::CT) *
           ceil_div(s3::inner(cnz), czc);
}

__global__ __launch_bounds__(256, 4) void resrestrict_kernel(const MgLevelDev *__restrict__ Fs,
                                                          const MgLevelDev *__restrict__ Cs,
                                                          const MgStatus *__restrict__ st,
                                                          const uint32_t *__restrict__ wl)
{
    using namespace rr;
    using s3::kOob;
    __shared__ float ering[4 * ER * ER];
    __shared__ float rpl[2 * FR * RS];
    __shared__ float4 tabr[2 * 128];  // code byte bits 0-6 -> {a0..a3}, {a4, a5, diag, -}
    const BatchBlock bb = batch_block(wl);
    if (stopped(st + bb.s)) return;
    const MgLevelDev C = Cs[bb.s];
    if (C.nx == 0) return;
    const MgLevelDev F = Fs[bb.s];
    const int ntx = (s3::inner(C.nx) + CT - 1) / CT, nty = (s3::inner(C.ny) + CT - 1) / CT;
    const int nch = (s3::inner(C.nz) + C.zc - 1) / C.zc;
    if (bb.b >= ntx * nty * nch) return;
    const int tile = bb.b % (ntx * nty), ch = bb.b / (ntx * nty);
    const int cx0 = 1 + (tile % ntx) * CT, cy0 = 1 + (tile / ntx) * CT;
    const int cz0 = 1 + ch * C.zc, cz1 = min(C.nz - 1, cz0 + C.zc);
    const int nx = F.nx, ny = F.ny, nz = F.nz;
    const int fx0 = 2 * cx0 - 1, fy0 = 2 * cy0 - 1;  // residual region origin
    const uint32_t sxy = (uint32_t)nx * (uint32_t)ny, N = sxy * (uint32_t)nz;
    const uint32_t csxy = (uint32_t)C.nx * (uint32_t)C.ny;
    const bool uni = F.uni != 0, hasw = !uni && F.cw;
    const auto er = buf_rsrc(F.e2, 4u * N), fr = buf_rsrc(F.f, 4u * N);
    const auto cr = buf_rsrc(hasw ? F.cw : nullptr, hasw ? N : 0u);
    const auto exr = buf_rsrc(uni ? nullptr : F.ex, uni ? 0u : 4u * N);
    const auto eyr = buf_rsrc(uni ? nullptr : F.ey, uni ? 0u : 4u * N);
    const auto ezr = buf_rsrc(uni ? nullptr : F.ez, uni ? 0u : 4u * N);
    const auto k2r = buf_rsrc(uni || hasw ? nullptr : F.k2, uni || hasw ? 0u : 4u * N);
    const auto cfr = buf_rsrc(C.f, 4u * csxy * (uint32_t)C.nz);
    const int tid = threadIdx.x;
    const float u = F.eps_u, ihx = F.ihx, ihy = F.ihy, ihz = F.ihz;
    // faces and diagonal of every code without the fractional bit, from the
    // residual's own expressions (the same bits); visible after the
    // prologue's barrier
    if (hasw && tid < 128) {
        const uint32_t w = (uint32_t)tid;
        const float a0 = face_bit<float>(w, 0, F.c_in, F.c_out);
        const float a1 = face_bit<float>(w, 1, F.c_in, F.c_out);
        const float a2 = face_bit<float>(w, 2, F.c_in, F.c_out);
        const float a3 = face_bit<float>(w, 3, F.c_in, F.c_out);
        const float a4 = face_bit<float>(w, 4, F.c_in, F.c_out);
        const float a5 = face_bit<float>(w, 5, F.c_in, F.c_out);
        const float k2 = (w & kCodeKappa) ? F.c_k2 : 0.0f;
        const float dg = (a0 + a1) * ihx + (a2 + a3) * ihy + (a4 + a5) * ihz + k2;
        tabr[2 * tid] = make_float4(a0, a1, a2, a3);
        tabr[2 * tid + 1] = make_float4(a4, a5, dg, 0.0f);
    }

    // correction region value n = tid + 256 j: node (fx0 - 1 + n % ER, fy0 - 1 + n / ER)
    uint32_t eo[NE];
#pragma unroll
    for (int j = 0; j < NE; ++j) {
        const int n = tid + 256 * j, ry = n / ER, rx = n - ry * ER;
        const int x = fx0 - 1 + rx, y = fy0 - 1 + ry;
        eo[j] = (n < ER * ER && x >= 0 && x < nx && y >= 0 && y < ny) ? 4u * (uint32_t)(x + nx * y)
                                                                        : kOob;
    }
    // residual node n = tid + 256 j: node (fx0 + n % FR, fy0 + n / FR); bit j of
    // rin: interior column (in-plane neighbours exist), the rest give residual 0
    uint32_t ro[NR];
    int rli[NR], rpo[NR];
    uint32_t rin = 0u;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const int n = tid + 256 * j, ry = n / FR, rx = n - ry * FR;
        const int x = fx0 + rx, y = fy0 + ry;
        const bool in = n < FR * FR && x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2;
        ro[j] = in ? (uint32_t)(x + nx * y) : kOob;
        rli[j] = (ry + 1) * ER + (rx + 1);
        rpo[j] = rpos(ry, rx);
        if (in) rin |= 1u << j;
    }
    auto eslot = [&](int z) { return ering + ((z + 8) & 3) * (ER * ER); };
    float ev[NE];
    auto load_e = [&](int z) {
        const bool zi = z >= 0 && z < nz;
        const uint32_t zo = zi ? 4u * sxy * (uint32_t)z : 0u;
#pragma unroll
        for (int j = 0; j < NE; ++j) ev[j] = bload(er, zi ? eo[j] : kOob, zo);
    };
    auto store_e = [&](int z) {
        float *sl = eslot(z);
#pragma unroll
        for (int j = 0; j < NE; ++j)
            if (tid + 256 * j < ER * ER) sl[tid + 256 * j] = ev[j];
    };
    float fv[NR];
    uint32_t wv[NR];
    auto load_fw = [&](int z) {
        const bool zi = z >= 1 && z <= nz - 2;
        const uint32_t zo = zi ? 4u * sxy * (uint32_t)z : 0u;
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const uint32_t o = zi ? ro[j] : kOob;
            fv[j] = bload(fr, o == kOob ? kOob : 4u * o, zo);
            wv[j] = hasw ? bload8(cr, o, zo >> 2) : 0u;
        }
    };

    const int zs = 2 * cz0 - 1, ze = 2 * cz1 - 1;  // fine planes streamed: zs .. ze
    for (int z = zs - 1; z <= zs + 1; ++z) {
        load_e(z);
        store_e(z);
    }
    load_e(zs + 2);
    load_fw(zs);
    __syncthreads();

    // coarse column of this thread and its fine 3 x 3 neighbourhood in rpl:
    // restrict_node's weights (2 on the axis, 1 off it) with 0 where it skips
    // a node outside the grid.  Such nodes hold residual 0 and the chain
    // starts at +0 and never reaches -0, so fma(0, 0, sum) leaves its bits
    // unchanged; the weight sums are small integers, exact in any order.
    const int ccx = cx0 + (tid & (CT - 1)), ccy = cy0 + (tid / CT);
    const bool cown = ccx < C.nx - 1 && ccy < C.ny - 1;
    float wxy[9];
    float wsxy = 0.0f;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = 2 * ccx + dx, y = 2 * ccy + dy;
            const bool in = x >= 0 && x < nx && y >= 0 && y < ny;
            const float w = in ? (dx == 0 ? 2.0f : 1.0f) * (dy == 0 ? 2.0f : 1.0f) : 0.0f;
            wxy[3 * (dy + 1) + (dx + 1)] = w;
            wsxy += w;
        }
    // fine (x, y) = (2 ccx + dx, 2 ccy + dy) sits at region column 2 (ccx - cx0)
    // + dx + 1 and row 2 (ccy - cy0) + dy + 1
    const int cb = rpos(2 * (ccy - cy0) + 1, 2 * (ccx - cx0) + 1) - ODD;  // even column of dx = -1
    float sum = 0.0f, ws = 0.0f;
    for (int z = zs; z <= ze; ++z) {
        store_e(z + 2);
        const bool zin = z >= 1 && z <= nz - 2;
        // faces of fractional-face nodes (or of every node without code bytes)
        // (loads only for the lanes that need them: a wave without such a
        // node skips them)
        float fa[NR][7];
        const uint32_t zo = 4u * sxy * (uint32_t)(zin ? z : 0);
#pragma unroll
        for (int j = 0; j < NR; ++j) {
#pragma unroll
            for (int k = 0; k < 7; ++k) fa[j][k] = 0.0f;
            const bool arr = !uni && zin && ((rin >> j) & 1u) && (!hasw || (wv[j] & kCodeSpecial));
            if (arr) {
                const uint32_t g = 4u * ro[j];
                fa[j][0] = bload(exr, g, zo);
                fa[j][1] = bload(exr, g - 4u, zo);
                fa[j][2] = bload(eyr, g, zo);
                fa[j][3] = bload(eyr, g - 4u * (uint32_t)nx, zo);
                fa[j][4] = bload(ezr, g, zo);
                fa[j][5] = bload(ezr, g, zo - 4u * sxy);  // plane z - 1
                if (!hasw) fa[j][6] = bload(k2r, g, zo);
            }
        }
        const float *em = eslot(z - 1), *ec = eslot(z), *ep = eslot(z + 1);
        float *rp = rpl + (z & 1) * (FR * RS);
        float fcur[NR];
        uint32_t wcur[NR];
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            fcur[j] = fv[j];
            wcur[j] = wv[j];
        }
        if (z + 1 <= ze) {
            load_e(z + 3);
            load_fw(z + 1);
        }
#pragma unroll
        for (int j = 0; j < NR; ++j) {
            const int n = tid + 256 * j;
            if (n >= FR * FR) continue;
            float r = 0.0f;
            if (zin && ((rin >> j) & 1u)) {
                const int li = rli[j];
                const float xp = ec[li + 1], xm = ec[li - 1], yp = ec[li + ER], ym = ec[li - ER];
                const float zp = ep[li], zm = em[li], c = ec[li];
                float nb, dg;
                if (uni) {
                    nb = u * ((xp + xm) * ihx + (yp + ym) * ihy + (zp + zm) * ihz);
                    dg = u * (2.0f * ihx + 2.0f * ihy + 2.0f * ihz);
                } else {
                    float exp_ = fa[j][0], exm = fa[j][1], eyp = fa[j][2], eym = fa[j][3];
                    float ezp = fa[j][4], ezm = fa[j][5], k2 = fa[j][6];
                    const uint32_t w = wcur[j];
                    if (hasw && !(w & kCodeSpecial)) {
                        const float4 A = tabr[2 * (w & 0x7fu)], B = tabr[2 * (w & 0x7fu) + 1];
                        exp_ = A.x;
                        exm = A.y;
                        eyp = A.z;
                        eym = A.w;
                        ezp = B.x;
                        ezm = B.y;
                        dg = B.z;
                    } else {
                        if (hasw) k2 = (w & kCodeKappa) ? F.c_k2 : 0.0f;
                        dg = (exp_ + exm) * ihx + (eyp + eym) * ihy + (ezp + ezm) * ihz + k2;
                    }
                    nb = fmaf(exp_, xp, exm * xm) * ihx + fmaf(eyp, yp, eym * ym) * ihy +
                         fmaf(ezp, zp, ezm * zm) * ihz;
                }
                r = fcur[j] + nb - dg * c;
            }
            rp[rpo[j]] = r;
        }
        __syncthreads();
        // restriction: plane z is dz = 0 of coarse plane z / 2 (z even), or
        // dz = +1 of (z - 1) / 2, which is then complete, and dz = -1 of
        // (z + 1) / 2 (z odd)
        auto terms = [&](int dz) {
            if (z < 0 || z >= nz) return;
            const float wz = dz == 0 ? 2.0f : 1.0f;
            const float *q = rp + cb;
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
                const int ro_ = dy * RS;
                sum = fmaf(wxy[3 * (dy + 1) + 0] * wz, q[ro_], sum);            // dx = -1
                sum = fmaf(wxy[3 * (dy + 1) + 1] * wz, q[ro_ + ODD], sum);      // dx = 0
                sum = fmaf(wxy[3 * (dy + 1) + 2] * wz, q[ro_ + 1], sum);        // dx = +1
            }
            ws += wsxy * wz;
        };
        if (cown) {
            if ((z & 1) == 0) {
                terms(0);
            } else {
                terms(1);
                const int cz = (z - 1) >> 1;
                if (cz >= cz0)
                    bstore(cfr, 0.25f * sum / ws, 4u * (uint32_t)(ccx + C.nx * ccy),
                           4u * csxy * (uint32_t)cz);
                sum = 0.0f;
                ws = 0.0f;
                terms(-1);
            }
        }
    }
}

void mg_residual_restrict(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int max_blocks,
                          const MgStatus *st, hipStream_t s, const uint32_t *wl)
{
    if (n_solves <= 0 || max_blocks <= 0) return;
    hipLaunchKernelGGL(resrestrict_kernel, batch_grid(max_blocks, n_solves, wl), dim3(256), 0, s,
                       F, C, st, wl);
    LAUNCH_OK();
}

void mg_prolong_e2s(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int max_blocks,
                    const MgStatus *st, hipStream_t s, const uint32_t *wl)
{
    if (n_solves <= 0 || max_blocks <= 0) return;
    hipLaunchKernelGGL(prolong_e2s_kernel, batch_grid(max_blocks, n_solves, wl), dim3(256), 0, s,
                       F, C, st, wl);
    LAUNCH_OK();
}
int mg_stream_blocks(int nx, int ny, int nz, int zc) { return stream_blocks(nx, ny, nz, zc); }

__global__ __launch_bounds__(256) void residual32_kernel(const MgLevelDev *__restrict__ Ls,
                                                         const MgStatus *__restrict__ st,
                                                         const uint32_t *__restrict__ wl)
{
    const BatchBlock bb = batch_block(wl);
    const int s = bb.s;
    if (stopped(st + s)) return;
    MgLevelDev L = Ls[s];
    if (L.nx == 0) return;
    const int64_t idx = (int64_t)bb.b * blockDim.x + threadIdx.x;
    if (idx >= nodes(L)) return;
    L.e = L.e2;  // the pre-smoothed correction
    residual_node(L, idx);
}

__global__ __launch_bounds__(256) void restrict_kernel(const MgLevelDev *__restrict__ Fs,
                                                       const MgLevelDev *__restrict__ Cs,
                                                       const MgStatus *__restrict__ st,
                                                       const uint32_t *__restrict__ wl)
{
    const BatchBlock bb = batch_block(wl);
    const int s = bb.s;
    if (stopped(st + s)) return;
    const MgLevelDev C = Cs[s];
    if (C.nx == 0) return;
    const int64_t idx = (int64_t)bb.b * blockDim.x + threadIdx.x;
    if (idx >= nodes(C)) return;
    // batched levels start smoothing from zero without reading e, and the
    // tail clears its entry level itself
    restrict_node(Fs[s], C, idx, false);
}

// One workgroup per solve runs the V-cycle over its tail levels.  Level
// descriptors are read from the table with uniform (scalar) loads wherever
// they are used, never copied into per-lane arrays.
__global__ __launch_bounds__(1024) void tail_kernel(const MgTailDev *__restrict__ Ts,
                                                    const MgStatus *__restrict__ st)
{
    // floats: 158 KB of gfx950's 160 KB of LDS per workgroup hold the
    // coarsest levels and, on the 161^3 chain, the 21^3 entry level's
    // correction, right-hand side and residual
    constexpr int kLds = 40448;
    __shared__ float lds[kLds];
    const int s = blockIdx.x;
    if (stopped(st + s)) return;
    const MgTailDev &T = Ts[s];
    const int n = T.n;
    if (n == 0) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    // LDS placement, from the coarsest level upwards while levels fit: e, f
    // and the residual, plus the FP32 faces when those fit too (uniform)
    struct Place {
        int off;    // float offset of e (f at + N, res at + 2 N), or -1: global memory
        bool full;  // faces at + 3 N .. + 6 N
    };
    auto place = [&](int l) {
        int used = 0;
        for (int q = n - 1; q >= l; --q) {
            const MgLevelDev &G = T.lv[q];
            const int N = G.nx * G.ny * G.nz;
            const bool full = used + (G.uni ? 3 : 7) * N <= kLds;
            if (!full && used + 3 * N > kLds) break;
            if (q == l) return Place{used, full && !G.uni};
            used += (full && !G.uni) ? 7 * N : 3 * N;
        }
        return Place{-1, false};
    };
    auto level = [&](int l) {
        MgLevelDev L = T.lv[l];
        const Place p = place(l);
        if (p.off >= 0) {
            const int N = L.nx * L.ny * L.nz;
            float *b = lds + p.off;
            L.e = b;
            L.f = b + N;
            L.res = b + 2 * N;
            if (p.full) {
                L.ex = b + 3 * N;
                L.ey = b + 4 * N;
                L.ez = b + 5 * N;
                L.k2 = b + 6 * N;
            }
        }
        return L;
    };
    // copy in: the entry level's right-hand side (from the batched
    // restriction) and the faces of levels placed whole; the entry level's
    // correction goes back to global memory at the end, where the
    // prolongation that follows reads it
    for (int l = 0; l < n; ++l) {
        const Place p = place(l);
        if (p.off < 0) continue;
        const MgLevelDev &G = T.lv[l];
        const int N = G.nx * G.ny * G.nz;
        float *b = lds + p.off;
        if (l == 0)
            for (int i = tid; i < N; i += nt) b[N + i] = G.f[i];
        if (p.full)
            for (int i = tid; i < N; i += nt) {
                b[3 * N + i] = G.ex[i];
                b[4 * N + i] = G.ey[i];
                b[5 * N + i] = G.ez[i];
                b[6 * N + i] = G.k2[i];
            }
    }
    __syncthreads();
    auto sweeps = [&](const MgLevelDev &L, int count) {
        const int64_t w = rb_work(L);
        for (int k = 0; k < count; ++k)
            for (int color = 0; color < 2; ++color) {
                for (int64_t t = tid; t < w; t += nt) {
                    const int64_t idx = rb_index(L, t, color);
                    if (idx >= 0) rb_node(L, idx);
                }
                __syncthreads();
            }
    };
    {
        const MgLevelDev L0 = level(0);
        for (int64_t i = tid; i < nodes(L0); i += nt) L0.e[i] = 0.0f;
        __syncthreads();
    }
    for (int l = 0; l + 1 < n; ++l) {
        const MgLevelDev L = level(l);
        const MgLevelDev C = level(l + 1);
        sweeps(L, 3);
        for (int64_t i = tid; i < nodes(L); i += nt) residual_node(L, i);
        __syncthreads();
        for (int64_t i = tid; i < nodes(C); i += nt) restrict_node(L, C, i, true);
        __syncthreads();
    }
    {
        const MgLevelDev L = level(n - 1);
        if (L.nx == 3 && L.ny == 3 && L.nz == 3) {
            // one interior node, whose neighbours are boundary nodes that no
            // sweep changes: its 50 updates (one per sweep, in the odd colour)
            // need no other thread
            if (tid == 0)
                for (int k = 0; k < 50; ++k) rb_node(L, 13);
            __syncthreads();
        } else {
            sweeps(L, 50);
        }
    }
    for (int l = n - 2; l >= 0; --l) {
        const MgLevelDev L = level(l);
        const MgLevelDev C = level(l + 1);
        for (int64_t i = tid; i < nodes(L); i += nt) prolong_node(C, L, i);
        __syncthreads();
        sweeps(L, 3);
    }
    if (place(0).off >= 0) {
        const MgLevelDev L0 = level(0);
        float *g = T.lv[0].e;
        for (int64_t i = tid; i < nodes(L0); i += nt) g[i] = L0.e[i];
    }
}

// ---- FP64 outer step: phi_out = phi_in + e, f0 = b - A phi_out (stencil_at,
// pb_solver.rs:159-194), block partial sums of r^2.  The charge density rho is
// read by the caller (interior nodes only) and passed in.
__device__ __forceinline__ double outer_node(const MgOuterDev &O, int64_t idx, bool interior,
                                             double rho)
{
    const int nx = O.nx, ny = O.ny;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    // descriptor pointers are device memory: global (not flat) accesses
    const auto e = to_global<const float>(O.e);
    const auto p = to_global<const double>(O.phi_in);
    const auto po = to_global<double>(O.phi_out);
    const auto f0 = to_global<float>(O.f0);
    const double pc = e ? p[idx] + (double)e[idx] : p[idx];
    if (e) po[idx] = pc;
    if (!interior) {
        f0[idx] = 0.0f;
        return 0.0;
    }
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
        // coeffs<double> with global accesses
        const auto cw = to_global<const uint8_t>(O.cw);
        bool arrays = true;
        if (O.cw) {
            const uint32_t w = cw[idx];
            k2 = (w & kCodeKappa) ? O.c_k2 : 0.0;
            if (!(w & kCodeSpecial)) {
                exp_ = face_bit<double>(w, 0, O.c_in, O.c_out);
                exm = face_bit<double>(w, 1, O.c_in, O.c_out);
                eyp = face_bit<double>(w, 2, O.c_in, O.c_out);
                eym = face_bit<double>(w, 3, O.c_in, O.c_out);
                ezp = face_bit<double>(w, 4, O.c_in, O.c_out);
                ezm = face_bit<double>(w, 5, O.c_in, O.c_out);
                arrays = false;
            }
        } else {
            k2 = to_global<const double>(O.k2)[idx];
        }
        if (arrays) {
            const auto ex = to_global<const double>(O.ex);
            const auto ey = to_global<const double>(O.ey);
            const auto ez = to_global<const double>(O.ez);
            exp_ = ex[idx];
            exm = ex[idx - 1];
            eyp = ey[idx];
            eym = ey[idx - sy];
            ezp = ez[idx];
            ezm = ez[idx - sz];
        }
    }
    const double nb = (exp_ * pxp + exm * pxm) * O.ihx + (eyp * pyp + eym * pym) * O.ihy +
                      (ezp * pzp + ezm * pzm) * O.ihz;
    const double dg = (exp_ + exm) * O.ihx + (eyp + eym) * O.ihy + (ezp + ezm) * O.ihz + k2;
    const double r = O.scale * rho + nb - dg * pc;
    f0[idx] = (float)r;
    return r * r;
}

// One block: 256 nodes of solve s and, when it has a partner (same grid, same
// charge density: the solvated and reference solves of one grid), the same
// nodes of the partner, so rho is read once for both.  Either side may have
// stopped; the work list holds the lead solve while either side runs.
__global__ __launch_bounds__(256) void outer64_kernel(const MgOuterDev *__restrict__ Os,
                                                      const MgStatus *__restrict__ st,
                                                      const uint32_t *__restrict__ wl)
{
    __shared__ double red[2][4];
    const BatchBlock bb = batch_block(wl);
    const int s = bb.s;
    const MgOuterDev O = Os[s];
    const int s2 = O.partner;
    const bool ra = !stopped(st + s), rb = s2 >= 0 && !stopped(st + s2);
    if (!ra && !rb) return;
    if (bb.b >= O.nbk) return;
    const int nx = O.nx, ny = O.ny, nz = O.nz;
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t idx = (int64_t)bb.b * blockDim.x + threadIdx.x;
    double a2 = 0.0, b2 = 0.0;
    if (idx < N) {
        int ix, iy, iz;
        split3(O, idx, ix, iy, iz);
        const bool interior =
            !(ix == 0 || iy == 0 || iz == 0 || ix == nx - 1 || iy == ny - 1 || iz == nz - 1);
        const double rho = interior ? to_global<const double>(O.rho)[idx] : 0.0;
        if (ra) a2 = outer_node(O, idx, interior, rho);
        if (rb) b2 = outer_node(Os[s2], idx, interior, rho);
    }
    a2 = wave_sum(a2);
    b2 = wave_sum(b2);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a2;
        red[1][threadIdx.x >> 6] = b2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (ra) O.partials[bb.b] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        if (rb) Os[s2].partials[bb.b] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
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
    const bool forced = st->diverge_at > 0 && cycle == st->diverge_at;
    const bool div = forced || !isfinite(rms) || rms > st->best * 1e4 ||
                     (st->growing >= 5 && rms > st->prev);
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

// Code bytes (layout at kCodeSpecial).  Face values equal to eps_in / eps_out
// (exact equality) set / clear their bit; any other value marks the node
// special; kappa-bar^2 is either 0 or c_k2 by construction.
template <typename T>
__global__ void pack_codes_kernel(int nx, int ny, int nz, const T *__restrict__ ex,
                                  const T *__restrict__ ey, const T *__restrict__ ez,
                                  const T *__restrict__ k2, T cin, T cout, uint8_t *__restrict__ cw)
{
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    uint32_t w = 0u;
    auto face = [&](const T *a, int64_t i, int bit) {
        if (i < 0) return;
        const T v = a[i];
        if (v == cin)
            w |= 1u << bit;
        else if (v != cout)
            w |= kCodeSpecial;
    };
    face(ex, idx, 0);
    face(ex, idx - 1, 1);
    face(ey, idx, 2);
    face(ey, idx - sy, 3);
    face(ez, idx, 4);
    face(ez, idx - sz, 5);
    if (k2[idx] != (T)0) w |= kCodeKappa;
    cw[idx] = (uint8_t)w;
}

template <typename T>
static void pack_codes(int nx, int ny, int nz, const T *ex, const T *ey, const T *ez, const T *k2,
                       T c_in, T c_out, uint8_t *cw, hipStream_t s)
{
    const int64_t N = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(pack_codes_kernel<T>, dim3((unsigned)ceil_div(N, 256)), dim3(256), 0, s, nx,
                       ny, nz, ex, ey, ez, k2, c_in, c_out, cw);
    LAUNCH_OK();
}

// coarsen_dielectrics / coarsen_kappa for level C of every listed solve in one
// launch (grid: blocks x solves; a solve whose C.nx is 0 has no such level)
__global__ void inject_batch_kernel(const MgLevelDev *__restrict__ Fs,
                                    const MgLevelDev *__restrict__ Cs)
{
    const MgLevelDev C = Cs[blockIdx.y];
    if (C.nx == 0) return;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes(C)) return;
    const MgLevelDev F = Fs[blockIdx.y];
    const int64_t csz = (int64_t)C.nx * C.ny;
    const int cz = (int)(idx / csz);
    const int cy = (int)((idx - (int64_t)cz * csz) / C.nx);
    const int cx = (int)(idx - (int64_t)cz * csz - (int64_t)cy * C.nx);
    const int x = min(2 * cx, F.nx - 1), y = min(2 * cy, F.ny - 1), z = min(2 * cz, F.nz - 1);
    const int64_t f = x + (int64_t)F.nx * (y + (int64_t)F.ny * z);
    const_cast<float *>(C.ex)[idx] = F.ex[f];
    const_cast<float *>(C.ey)[idx] = F.ey[f];
    const_cast<float *>(C.ez)[idx] = F.ez[f];
    const_cast<float *>(C.k2)[idx] = F.k2[f];
}

// code bytes of level L of every listed solve (pack_codes_kernel's rule)
__global__ void pack_codes_batch_kernel(const MgLevelDev *__restrict__ Ls)
{
    const MgLevelDev L = Ls[blockIdx.y];
    if (L.nx == 0) return;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes(L)) return;
    const int64_t sy = L.nx, sz = (int64_t)L.nx * L.ny;
    const float cin = L.c_in, cout = L.c_out;
    uint32_t w = 0u;
    auto face = [&](const float *a, int64_t i, int bit) {
        if (i < 0) return;
        const float v = a[i];
        if (v == cin)
            w |= 1u << bit;
        else if (v != cout)
            w |= kCodeSpecial;
    };
    face(L.ex, idx, 0);
    face(L.ex, idx - 1, 1);
    face(L.ey, idx, 2);
    face(L.ey, idx - sy, 3);
    face(L.ez, idx, 4);
    face(L.ez, idx - sz, 5);
    if (L.k2[idx] != 0.0f) w |= kCodeKappa;
    const_cast<uint8_t *>(L.cw)[idx] = (uint8_t)w;
}

// The entry of a fractional-face node with the phase smoother's own
// expressions (a sum of the six faces times ih plus kappa-bar^2, its
// reciprocal), so that reading it is bitwise the same as computing it there.
__global__ void fill_entries_kernel(const MgLevelDev *__restrict__ Ls)
{
    const MgLevelDev L = Ls[blockIdx.y];
    if (L.nx == 0 || !L.ent) return;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nodes(L)) return;
    const uint32_t w = L.cw[idx];
    if (!(w & kCodeSpecial)) return;
    int ix, iy, iz;
    split3(L, idx, ix, iy, iz);
    if (ix == 0 || iy == 0 || iz == 0 || ix == L.nx - 1 || iy == L.ny - 1 || iz == L.nz - 1)
        return;
    const int64_t sy = L.nx, sz = (int64_t)L.nx * L.ny;
    const float a0 = L.ex[idx], a1 = L.ex[idx - 1], a2 = L.ey[idx], a3 = L.ey[idx - sy];
    const float a4 = L.ez[idx], a5 = L.ez[idx - sz];
    const float k2 = (w & kCodeKappa) ? L.c_k2 : 0.0f;
    const float dg = fmaf(((a0 + a1) + (a2 + a3)) + (a4 + a5), L.ihx, k2);
    float4 *e = const_cast<float4 *>(L.ent);
    e[2 * idx] = make_float4(a0, a1, a2, a3);
    e[2 * idx + 1] = make_float4(a4, a5, __builtin_amdgcn_rcpf(dg), 0.0f);
}

void mg_fill_entries(const MgLevelDev *L, int n, int64_t max_nodes, hipStream_t s)
{
    if (n <= 0 || max_nodes <= 0) return;
    const dim3 grid((unsigned)std::max<int64_t>(1, ceil_div(max_nodes, 256)), (unsigned)n);
    hipLaunchKernelGGL(fill_entries_kernel, grid, dim3(256), 0, s, L);
    LAUNCH_OK();
}

void mg_inject_pack_batch(const MgLevelDev *F, const MgLevelDev *C, int n, int64_t max_nodes,
                          hipStream_t s)
{
    if (n <= 0 || max_nodes <= 0) return;
    const dim3 grid((unsigned)std::max<int64_t>(1, ceil_div(max_nodes, 256)), (unsigned)n);
    hipLaunchKernelGGL(inject_batch_kernel, grid, dim3(256), 0, s, F, C);
    LAUNCH_OK();
    hipLaunchKernelGGL(pack_codes_batch_kernel, grid, dim3(256), 0, s, C);
    LAUNCH_OK();
}

void mg_pack_codes(int nx, int ny, int nz, const double *ex, const double *ey, const double *ez,
                   const double *k2, double c_in, double c_out, uint8_t *cw, hipStream_t s)
{
    pack_codes<double>(nx, ny, nz, ex, ey, ez, k2, c_in, c_out, cw, s);
}

void mg_pack_codes(int nx, int ny, int nz, const float *ex, const float *ey, const float *ez,
                   const float *k2, float c_in, float c_out, uint8_t *cw, hipStream_t s)
{
    pack_codes<float>(nx, ny, nz, ex, ey, ez, k2, c_in, c_out, cw, s);
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

void mg_residual32(const MgLevelDev *L, int n_solves, int64_t max_nodes, const MgStatus *st,
                   hipStream_t s, const uint32_t *wl)
{
    if (max_nodes <= 0 || n_solves <= 0) return;
    hipLaunchKernelGGL(residual32_kernel, batch_grid(wl ? max_nodes : nblk(max_nodes), n_solves, wl),
                       dim3(256), 0, s, L, st, wl);
    LAUNCH_OK();
}

void mg_restrict(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int64_t max_coarse,
                 const MgStatus *st, hipStream_t s, const uint32_t *wl)
{
    if (max_coarse <= 0 || n_solves <= 0) return;
    hipLaunchKernelGGL(restrict_kernel, batch_grid(wl ? max_coarse : nblk(max_coarse), n_solves, wl),
                       dim3(256), 0, s, F, C, st, wl);
    LAUNCH_OK();
}

void mg_tail(const MgTailDev *T, int n_solves, const MgStatus *st, hipStream_t s)
{
    if (n_solves <= 0) return;
    hipLaunchKernelGGL(tail_kernel, dim3(n_solves), dim3(1024), 0, s, T, st);
    LAUNCH_OK();
}

void mg_outer64(const MgOuterDev *O, int n_solves, int64_t max_nodes, const MgStatus *st,
                hipStream_t s, const uint32_t *wl)
{
    if (n_solves <= 0 || (wl && max_nodes <= 0)) return;
    hipLaunchKernelGGL(outer64_kernel, batch_grid(wl ? max_nodes : nblk(max_nodes), n_solves, wl),
                       dim3(256), 0, s, O, st, wl);
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
