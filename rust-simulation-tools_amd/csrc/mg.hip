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

namespace mmpbsa {

__device__ __forceinline__ bool stopped(const MgStatus *st) { return st->state != 0; }

// (solve, block) of a batched launch (grid = blocks x solves), XCD-ordered so
// that each XCD walks contiguous blocks of one solve and z-neighbour planes
// are still in its L2.
struct BatchBlock {
    int s, b;
};
__device__ __forceinline__ BatchBlock batch_block()
{
    const uint32_t flat = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const uint32_t s = flat / gridDim.x;
    return {(int)s, (int)(flat - s * gridDim.x)};
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
                                                     const MgStatus *__restrict__ st)
{
    using namespace sm;
    __shared__ float ring[RING * RN];
    const uint32_t flat = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int s = (int)(flat / gridDim.x);
    const int b = (int)(flat - (uint32_t)s * gridDim.x);
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

void mg_smooth(const MgLevelDev *L, const MgLevelDev *C, int n_solves, int first, int dir,
               int max_blocks, const MgStatus *st, hipStream_t s)
{
    if (n_solves <= 0 || max_blocks <= 0) return;
    const dim3 grid(max_blocks, n_solves);
    if (first)
        hipLaunchKernelGGL((smooth_kernel<true, false>), grid, dim3(256), 0, s, L, C, dir, st);
    else if (C)
        hipLaunchKernelGGL((smooth_kernel<false, true>), grid, dim3(256), 0, s, L, C, dir, st);
    else
        hipLaunchKernelGGL((smooth_kernel<false, false>), grid, dim3(256), 0, s, L, C, dir, st);
    LAUNCH_OK();
}

__global__ __launch_bounds__(256) void residual32_kernel(const MgLevelDev *__restrict__ Ls,
                                                         const MgStatus *__restrict__ st)
{
    const BatchBlock bb = batch_block();
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
                                                       const MgStatus *__restrict__ st)
{
    const BatchBlock bb = batch_block();
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
                    if (idx >= 0) rb_node(L, idx);
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
        for (int64_t i = tid; i < nodes(C); i += nt) restrict_node(L, C, i, true);
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
// pb_solver.rs:159-194), block partial sums of r^2.
__global__ __launch_bounds__(256) void outer64_kernel(const MgOuterDev *__restrict__ Os,
                                                      const MgStatus *__restrict__ st)
{
    __shared__ double red[4];
    const BatchBlock bb = batch_block();
    const int s = bb.s;
    if (stopped(st + s)) return;
    const MgOuterDev O = Os[s];
    if (bb.b >= O.nbk) return;
    const int nx = O.nx, ny = O.ny, nz = O.nz;
    const int64_t N = (int64_t)nx * ny * nz;
    const int64_t idx = (int64_t)bb.b * blockDim.x + threadIdx.x;
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
    if (threadIdx.x == 0) O.partials[bb.b] = red[0] + red[1] + red[2] + red[3];
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

void mg_tail(const MgTailDev *T, int n_solves, const MgStatus *st, hipStream_t s)
{
    if (n_solves <= 0) return;
    hipLaunchKernelGGL(tail_kernel, dim3(n_solves), dim3(1024), 0, s, T, st);
    LAUNCH_OK();
}

void mg_outer64(const MgOuterDev *O, int n_solves, int64_t max_nodes, const MgStatus *st,
                hipStream_t s)
{
    if (n_solves <= 0) return;
    hipLaunchKernelGGL(outer64_kernel, dim3(nblk(max_nodes), n_solves), dim3(256), 0, s, O, st);
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
