// mg.h -- device views and launchers of the multigrid kernels (mg.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mmpbsa {

// Solver status word shared by every kernel of one solve.
// state: 0 running, 1 converged, 2 diverged, 3 out of cycles.
struct MgStatus {
    int state;
    int cycles;
    int growing;
    int pad;
    double best;
    double prev;
    double rms;
};

// One FP32 level of the correction V-cycle.
struct MgLevelDev {
    int nx, ny, nz;
    int uni;           // uniform dielectric, zero kappa
    float eps_u;
    float ihx, ihy, ihz;
    const float *ex, *ey, *ez, *k2;  // coefficients (unused when uni)
    float *e;          // correction
    float *f;          // right-hand side
    float *res;        // residual scratch
};

// FP64 outer step / SOR view of the finest grid.
struct MgOuterDev {
    int nx, ny, nz;
    int uni;
    double eps_u;
    double ihx, ihy, ihz;
    double scale;                       // 4 pi * 332.0522
    const double *ex, *ey, *ez, *k2;    // FP64 coefficients (unused when uni)
    const double *rho;                  // charge density
    const double *phi_in;
    double *phi_out;
    const float *e;                     // correction to add (nullptr: none)
    float *f0;                          // FP32 residual for the V-cycle
    double *partials;                   // per-block sums of r^2
};

void mg_rb_sweep(const MgLevelDev &L, int color, float omega, const MgStatus *st, hipStream_t s);
void mg_residual32(const MgLevelDev &L, const MgStatus *st, hipStream_t s);
void mg_restrict(const MgLevelDev &F, const MgLevelDev &C, const MgStatus *st, hipStream_t s);
void mg_prolong(const MgLevelDev &C, const MgLevelDev &F, const MgStatus *st, hipStream_t s);
void mg_inject(const MgLevelDev &F, const MgLevelDev &C, hipStream_t s);
void mg_zero(float *p, int64_t n, const MgStatus *st, hipStream_t s);
int mg_outer_blocks(const MgOuterDev &O);
void mg_outer64(const MgOuterDev &O, const MgStatus *st, hipStream_t s);
void mg_check(const double *partials, int nbk, double count, double tol, int cycle,
              int max_cycles, MgStatus *st, hipStream_t s);
int sor64_blocks(const MgOuterDev &O);
void mg_sor64(const MgOuterDev &O, int color, double omega, double *partials, const MgStatus *st,
              hipStream_t s);
void mg_sor_check(const double *partials, int nbk, double count, double tol, int iter,
                  int max_iter, MgStatus *st, hipStream_t s);

}  // namespace mmpbsa
