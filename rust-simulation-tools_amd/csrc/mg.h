// mg.h -- device views and launchers of the multigrid kernels (mg.hip).
// Every kernel serves a batch of independent solves: blockIdx.y (or
// blockIdx.x for one-block kernels) selects the solve, and each solve has its
// own status word, so converged solves drop out while the rest continue.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace mmpbsa {

// Solver status word of one solve.
// state: 0 running, 1 converged, 2 diverged, 3 out of cycles.
struct MgStatus {
    int state;
    int cycles;      // V-cycles completed (the reference's iteration count)
    int growing;
    int max_cycles;
    double best;
    double prev;
    double rms;
    double tol;
    double count;    // interior nodes (RMS denominator)
    int diverge_at;  // > 0: the divergence test fires at this cycle (set from the host)
    int pad_;
};

// One FP32 level of the correction V-cycle.  nx == 0 marks "no work for this
// solve at this level" in a batched launch.
struct MgLevelDev {
    int nx, ny, nz;
    int uni;           // uniform dielectric, zero kappa
    float eps_u;
    float ihx, ihy, ihz;
    float omega;
    const float *ex, *ey, *ez, *k2;  // coefficients (unused when uni)
    // Packed coefficients: when cw is set, cw[idx] is the code byte of node
    // idx (layout at kCodeSpecial), so a sweep reads one byte per node
    // instead of four coefficient arrays.
    const uint8_t *cw;
    float c_in, c_out, c_k2;
    float *e;          // correction (result of post-smoothing; in-place for the tail)
    float *e2;         // correction after pre-smoothing (batched levels)
    float *f;          // right-hand side
    float *res;        // residual scratch
    int zc;            // smoother: planes per workgroup (z-chunk length)
    int rg;            // smoother region (host bookkeeping): 0 32 x 32, 1 44 x 44
    // Fractional-face entries of a level the phase smoother sweeps (non-
    // uniform maps): at every interior node whose code byte has kCodeSpecial,
    // ent[2 idx] = {a0, a1, a2, a3} and ent[2 idx + 1] = {a4, a5, 1/diag, 0}
    // (faces x+, x-, y+, y-, z+, z-); other entries are never read.  Null
    // elsewhere.
    const float4 *ent;
    // finest level only (else null): the FP64 potential the phase smoother's
    // post-smoothing pass adds its final correction to, phi += e, in place of
    // storing e (the outer step then reads phi alone)
    double *phi;
    // index helpers: (nx-1)/2 slots per row, ny-2 interior rows, nx, nx*ny
    FastDiv fd_half, fd_row, fd_nx, fd_nxy;
};
void mg_level_index(MgLevelDev &L);  // fills the FastDiv helpers from nx, ny

constexpr int kMaxTail = 8;
// Levels handled by the one-workgroup tail kernel (coarse end of the V-cycle).
struct MgTailDev {
    int n;                     // number of levels, n == 0: no tail
    MgLevelDev lv[kMaxTail];   // lv[0] is entered with f set and e zero
};

// FP64 outer step / SOR view of the finest grid.
struct MgOuterDev {
    int nx, ny, nz;
    int uni;
    double eps_u;
    double ihx, ihy, ihz;
    double scale;                       // 4 pi * 332.0522
    const double *ex, *ey, *ez, *k2;    // FP64 coefficients (unused when uni)
    const uint8_t *cw;                  // code bytes as in MgLevelDev (or null)
    double c_in, c_out, c_k2;
    const double *rho;                  // charge density
    const double *phi_in;
    double *phi_out;
    const float *e;                     // correction to add (nullptr: none)
    float *f0;                          // FP32 residual for the V-cycle
    double *partials;                   // per-block sums of r^2
    int nbk;                            // number of partials (blocks per solve)
    FastDiv fd_nx, fd_nxy;
    // outer64: index of a solve with the same grid and charge density whose
    // outer step runs in the same threads (one read of rho for both), or -1
    int partner;
};

// Batched launchers.  L / C / F / O / T are device arrays with one entry per
// solve; `max_nodes` is the largest node count among the entries.
// Tiled smoother: one red-black Gauss-Seidel sweep over the level, out of
// place.  dir 0 reads e2 and writes e, dir 1 reads e and writes e2; first = 1
// starts from a zero correction (nothing is read) and writes e2.  Each
// workgroup owns an (x, y) tile and a range of L.zc planes, streams it along z
// through an LDS ring and recomputes a halo, so the result equals that of two
// in-place half-sweeps over the whole level.
int smooth_tiles(int nx, int ny);  // (x, y) tiles of one plane
int smooth_zc(int nz);             // z-chunk length for a level of nz planes
// max_blocks: the largest smooth_tiles * ceil(nz / zc) among the solves
// C (one coarse level per solve, or null): add the coarse correction's
// prolongation to the source while loading it (first post-smoothing pass).
// Every batched launcher takes an optional work list wl (device memory,
// {n, pre[0..n], id[0..n-1]}: solve id[k] owns blocks pre[k] .. pre[k+1]-1);
// with a list, the block-count argument is the list's total pre[n] and the
// grid is one-dimensional, without one it is that count times n_solves.
void mg_smooth(const MgLevelDev *L, const MgLevelDev *C, int n_solves, int first, int dir,
               int max_blocks, const MgStatus *st, hipStream_t s, const uint32_t *wl = nullptr);
// Temporally blocked smoother: one pass = three red-black sweeps (a whole
// smoothing phase of the V-cycle), same buffers and flags as mg_smooth, with
// smooth3_blocks workgroups per solve, L.zc from smooth3_zc.  The result
// equals three mg_smooth passes bit for bit.  Needs 4 * nodes < 2^31.
// rg selects the (x, y) region of a workgroup: 0 = 32 x 32 nodes (a 20 x 20
// core), 1 = 44 x 44 (32 x 32 core); every solve of a launch uses the same.
int smooth3_region(int nx, int ny);  // the built-in choice for a level
int smooth3_threads(int rg);         // workgroup size of a region
int smooth3_tiles(int nx, int ny, int rg);
int smooth3_blocks(int nx, int ny, int nz, int zc, int rg);
int smooth3_zc(int nx, int ny, int nz, int n_solves, int rg);
void mg_smooth3(const MgLevelDev *L, const MgLevelDev *C, int n_solves, int first, int dir,
                int max_blocks, int rg, const MgStatus *st, hipStream_t s,
                const uint32_t *wl = nullptr);
// z-streaming version of prolong_e2 (same results bit for bit): one thread
// per in-plane node over L.zc planes, mg_stream_blocks workgroups per solve.
int mg_stream_blocks(int nx, int ny, int nz, int zc);
void mg_prolong_e2s(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int max_blocks,
                    const MgStatus *st, hipStream_t s, const uint32_t *wl = nullptr);
// Residual of e2 and its restriction into C.f in one pass (no residual
// array), equal bit for bit to mg_residual32 then mg_restrict at the interior
// coarse nodes (the boundary's right-hand sides are never read, nor written
// here); mg_rr_blocks(coarse dims, C.zc) workgroups per solve.
int mg_rr_blocks(int cnx, int cny, int cnz, int czc);
void mg_residual_restrict(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int max_blocks,
                          const MgStatus *st, hipStream_t s, const uint32_t *wl = nullptr);
// e2 += prolongation of C's correction (post-smoothing input of smooth3)
void mg_prolong_e2(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int64_t max_nodes,
                   const MgStatus *st, hipStream_t s, const uint32_t *wl = nullptr);
void mg_residual32(const MgLevelDev *L, int n_solves, int64_t max_nodes, const MgStatus *st,
                   hipStream_t s, const uint32_t *wl = nullptr);
void mg_restrict(const MgLevelDev *F, const MgLevelDev *C, int n_solves, int64_t max_coarse,
                 const MgStatus *st, hipStream_t s, const uint32_t *wl = nullptr);
void mg_tail(const MgTailDev *T, int n_solves, const MgStatus *st, hipStream_t s);
void mg_outer64(const MgOuterDev *O, int n_solves, int64_t max_nodes, const MgStatus *st,
                hipStream_t s, const uint32_t *wl = nullptr);
void mg_check(const MgOuterDev *O, int n_solves, int count_cycle, MgStatus *st, hipStream_t s);
void mg_status_init(MgStatus *st, const MgStatus *init_host, int n, hipStream_t s);
void mg_inject(const MgLevelDev &F, const MgLevelDev &C, hipStream_t s);  // host-side structs
// mg_inject then the code bytes (mg_pack_codes) of level C for n solves at once:
// F and C are device arrays of n descriptors (C[k].nx == 0: skip solve k),
// max_nodes the largest C level.
void mg_inject_pack_batch(const MgLevelDev *F, const MgLevelDev *C, int n, int64_t max_nodes,
                          hipStream_t s);
// fill L[k].ent (where set) from the level's FP32 faces, code bytes, ihx and
// c_k2, for n descriptors at once (device array; max_nodes the largest level)
void mg_fill_entries(const MgLevelDev *L, int n, int64_t max_nodes, hipStream_t s);

// Code byte of a node: bits 0-5 = face x+, x-, y+, y-, z+, z- has the solute
// dielectric (else the solvent one), bit 6 = kappa-bar^2 is c_k2 (else 0),
// bit 7 = some face has neither value: read all six from ex / ey / ez.
constexpr uint32_t kCodeKappa = 0x40u, kCodeSpecial = 0x80u;
void mg_pack_codes(int nx, int ny, int nz, const double *ex, const double *ey, const double *ez,
                   const double *k2, double c_in, double c_out, uint8_t *cw, hipStream_t s);
void mg_pack_codes(int nx, int ny, int nz, const float *ex, const float *ey, const float *ez,
                   const float *k2, float c_in, float c_out, uint8_t *cw, hipStream_t s);

// Single-solve FP64 Chebyshev SOR (fallback path).
int sor64_blocks(const MgOuterDev &O);
void mg_sor64(const MgOuterDev &O, int color, double omega, double *partials, const MgStatus *st,
              hipStream_t s);
void mg_sor_check(const double *partials, int nbk, double count, double tol, int iter,
                  int max_iter, MgStatus *st, hipStream_t s);

}  // namespace mmpbsa
