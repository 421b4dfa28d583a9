/*
 * mmpbsa_oracle.h -- FP64 CPU restatement of the reference MM-PBSA path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it; the product library never links it.
 *
 * Each function follows the reference source cited beside it (paths relative
 * to the rust-simulation-tools 0.2.4 checkout).  Arithmetic is IEEE binary64
 * in the reference's operation order, compiled with -ffp-contract=off and no
 * fast-math, with glibc libm (the libm Rust's f64 methods call on Linux).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   kappa, single-ion GB, bond/angle/dihedral, single-atom SASA, PB Born ion,
 *   odd auto_grid dims, charge conservation, node interpolation, SOR/MG
 *   convergence, restrict/prolong constants: the reference's own unit tests.
 *   Born radii: two-atom closed form (analytic HCT integral, far-field series)
 *   and the isolated-atom limit.  GB / MM pair sums: brute-force long-double
 *   sums on small systems and two-atom analytic energies.  SES / EDT / flood
 *   fill: brute-force O(G^2) search and naive BFS on tiny grids.  Face
 *   fractions: dense sampling.  LPBE: scipy sparse direct solve of the same
 *   discrete operator on small grids.  SASA: isolated atom = 4 pi (r+p)^2.
 *   Decomposition: sum over residues == receptor x ligand cross-pair sums.
 *   Quasi-harmonic entropy: Jacobi eigenvalues on the reference's unit-test
 *   matrices and against LAPACK; one oscillator's closed form and classical
 *   limit; invariance under rigid motions; additivity over uncorrelated parts.
 */
#ifndef MMPBSA_ORACLE_H
#define MMPBSA_ORACLE_H

#include <stdint.h>
#include "../include/mmpbsa_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
    int64_t n[3];
    double h[3];
    double o[3];
} orc_grid;

typedef struct orc_solve {
    int32_t iterations;
    int32_t converged;
    double final_residual;
} orc_solve;

/* gb_energy.rs:126-139 */
double orc_compute_kappa(double salt, double eps, double temperature);
/* gb_energy.rs:144-258 */
void orc_gb_born_radii(const mmpbsa_topology_view *t, const double *xyz,
                       const mmpbsa_gb_params *p, double *born_out);
/* gb_energy.rs:283-355 */
double orc_gb_energy_from_radii(const mmpbsa_topology_view *t, const double *xyz,
                                const mmpbsa_gb_params *p, const double *born);
/* mm_energy.rs:168-267 */
void orc_mm_energy(const mmpbsa_topology_view *t, const double *xyz, mmpbsa_mm_terms *out);
/* core/src/sasa.rs:147-289 + sa_energy.rs:51-75 */
void orc_sasa_per_atom(int32_t n, const double *xyz, const double *radii,
                       double probe, int32_t n_points, double *out);

/* ---- PB pieces (pb_grid.rs, pb_solver.rs, pb_energy.rs) ---- */
void orc_auto_grid(int32_t n, const double *xyz, double spacing, double buffer, orc_grid *g);
void orc_coarse_grid(int32_t n, const double *xyz, double fine_spacing, double fillratio,
                     int64_t fscale, orc_grid *g);
void orc_map_charges(const orc_grid *g, int32_t n, const double *xyz, const double *q,
                     double *out);
void orc_inside_ses(const orc_grid *g, int32_t n, const double *xyz, const double *radii,
                    double probe, uint8_t *inside_out, uint8_t *sas_out, uint8_t *ext_out);
double orc_fraction_inside(double lo, double hi, double p1, double p2, int axis, int32_t n,
                           const double *xyz, const double *radii, double probe);
void orc_assign_dielectrics(const orc_grid *g, int32_t n, const double *xyz,
                            const double *radii, double probe, double eps_in, double eps_out,
                            double *ex, double *ey, double *ez);
void orc_assign_kappa(const orc_grid *g, int32_t n, const double *xyz, const double *radii,
                      double ion_radius, double kappa, double eps_solvent, double *out);
void orc_dh_boundary(const orc_grid *g, int32_t n, const double *xyz, const double *q,
                     double kappa, double eps_out, double *phi);
void orc_interp_boundary(const orc_grid *fine, const orc_grid *coarse, const double *coarse_phi,
                         double *phi);
/* bc: 0 = zero, 1 = Debye-Hueckel, 2 = values already in phi's boundary */
void orc_solve_multigrid(const orc_grid *g, const double *charge, const double *ex,
                         const double *ey, const double *ez, const double *k2, int bc,
                         int32_t n, const double *xyz, const double *q, double kappa_bulk,
                         double eps_out, double tol, int64_t max_it, double *phi,
                         orc_solve *res);
void orc_solve_sor(const orc_grid *g, const double *charge, const double *ex, const double *ey,
                   const double *ez, const double *k2, int bc, int32_t n, const double *xyz,
                   const double *q, double kappa_bulk, double eps_out, double tol,
                   int64_t max_it, double *phi, orc_solve *res);
double orc_rms_residual(const orc_grid *g, const double *phi, const double *charge,
                        const double *ex, const double *ey, const double *ez, const double *k2,
                        double rhs_scale);
double orc_elec_energy(const orc_grid *g, const double *phi, int32_t n, const double *xyz,
                       const double *q);
void orc_restrict(const double *fine, const int64_t *fdims, double *coarse,
                  const int64_t *cdims, double scale);
void orc_prolongate_add(const double *coarse, const int64_t *cdims, double *fine,
                        const int64_t *fdims);
/* pb_energy.rs:332-438 (with the fine_dims and explicit focusing-grid
 * extensions) */
void orc_pb_energy(const mmpbsa_topology_view *t, const double *xyz, const mmpbsa_pb_params *p,
                   mmpbsa_pb_result *out);
/* the same, and per atom 0.5 q_i (phi_solv - phi_ref)(r_i) on the fine grid
 * (the per-residue PB terms are sums of these) */
void orc_pb_atom_terms(const mmpbsa_topology_view *t, const double *xyz,
                       const mmpbsa_pb_params *p, mmpbsa_pb_result *out, double *atom_terms);
/* decomposition.rs:240-283: receptor x ligand pairs into per-residue sums */
void orc_decomp_pairs(const mmpbsa_topology_view *t, const double *xyz, int32_t n_rec,
                      const int32_t *rec_atoms, int32_t n_lig, const int32_t *lig_atoms,
                      const int32_t *atom_res, double *res_vdw, double *res_elec);
/* Face-fraction strategy: 0 = scan all atoms per mixed edge (pb_grid.rs:742),
 * 1 = restrict to atoms in nearby cells (same interval set, same result). */
void orc_set_face_scan(int mode);
/* entropy.rs:227-275: eigenvalues of a symmetric dim x dim matrix (row-major,
 * overwritten) by cyclic Jacobi rotations (off-diagonal threshold 1e-10, at
 * most 100 sweeps), in descending order */
void orc_jacobi_eigenvalues(double *a, int32_t dim, double *eig);
/* entropy.rs:97-217: -T dS (kcal/mol) of frames xyz[nf][na][3] (A) with
 * masses (amu); returns 0 (the reference's None) with fewer than two frames
 * or no atoms, else 1 */
int orc_quasi_harmonic_entropy(const double *xyz, int32_t nf, int32_t na, const double *masses,
                               double temperature, double *minus_tds);
void orc_set_threads(int n);
void orc_set_force_divergence(int cycle);

#ifdef __cplusplus
}
#endif
#endif
