// kernels.h -- host-side launchers of the MM-PBSA HIP kernels (gfx950).
// Every launcher enqueues on `s` and never synchronises.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.h"

namespace mmpbsa {

// ---- atom-pair kernels (pairs.hip) ------------------------------------------
struct PairTopo {          // device views of one subsystem
    int n;
    int n_types;
    const double *q_amber;   // [n]
    const double *radii;     // [n]
    const double *screen;    // [n]
    const int *type;         // [n]
    const double *ljA;       // [n_types^2] (0 where nb_parm_index < 1)
    const double *ljB;
    const int *ex_ptr;       // [n+1] CSR of excluded partners j > i
    const int *ex_col;
    const int *p14_ptr;      // [n+1] CSR of 1-4 partners j > i
    const int *p14_col;
    const uint8_t *tile_flag;  // [nb*nb] tile (ib, jb) holds an excluded / 1-4 pair
    int nb;                    // tiles per side (256 atoms per tile)
    const int *bonds; int n_bonds; const double *bond_k; const double *bond_r0;
    const int *angles; int n_angles; const double *angle_k; const double *angle_t0;
    const int *dihs; int n_dihs; const double *dih_pk; const double *dih_pn; const double *dih_ph;
    double scee, scnb;
};

constexpr int kTile = 256;

// Born radii: psi partials then the OBC/HCT map.  work: [chunks * n] doubles.
int gb_chunks(int n);
void set_gb_chunks_override(int c);  // 0 restores the heuristic
void launch_gb_radii(const double *xyz, const PairTopo &t, int model, double offset,
                     double rgbmax, double cutoff, double *psi_work, double *born_out,
                     hipStream_t s);
// The chunks of psi partials launch_gb_radii writes for n atoms.
int gb_psi_chunks(int n);
// Complex Born radii from its partners: psi partials of the atoms of A (complex
// indices) from the atoms of B, [gb_psi_cross_chunks(nA, nB)][nA]; then the
// radii of one partner's atoms (map: their complex indices) from the partner's
// own psi partials plus those, written at the complex indices.
int gb_psi_cross_chunks(int nA, int nB);
void launch_gb_psi_cross(const double *xyz, const PairTopo &t, const int *A, int nA, const int *B,
                         int nB, double offset, double cutoff, double *part, hipStream_t s);
void launch_gb_finalize_parts(const double *own, int own_chunks, const double *cross,
                              int cross_chunks, const int *map, int n, const PairTopo &t,
                              int model, double offset, double rgbmax, double *born,
                              hipStream_t s);
// GB energy: writes partials; final sum by launch_sum_partials.
int gb_energy_blocks(int n);
void launch_gb_energy(const double *xyz, const PairTopo &t, const double *born, double kappa,
                      double eps_in, double eps_out, double cutoff, double *partials,
                      hipStream_t s);
// MM: 8 doubles (4 dd terms) per tile block, plus bonded partials.
int mm_blocks(int n);
void launch_mm_pairs(const double *xyz, const PairTopo &t, double *partials, hipStream_t s);
// Receptor x ligand pairs of a complex (A, B: complex atom indices; no
// excluded or 1-4 pair across them): mm_pairs' vdW and Coulomb terms of those
// pairs, 8 doubles per block as launch_mm_pairs.
int mm_cross_blocks(int nA, int nB);
void launch_mm_cross(const double *xyz, const PairTopo &t, const int *A, int nA, const int *B,
                     int nB, double *partials, hipStream_t s);
int bonded_blocks(const PairTopo &t);
void launch_bonded(const double *xyz, const PairTopo &t, double *partials, hipStream_t s);
// Deterministic reductions of block partials.
void launch_sum_partials(const double *partials, int n_blocks, int stride, int n_terms,
                         double *out, hipStream_t s);
void launch_sum_partials_dd(const double *partials, int n_blocks, int n_terms, double *out,
                            hipStream_t s);

// Debye-Hueckel boundary on the boundary nodes of grid g (FP32 terms, FP64 sums).
struct GridDesc {
    int nx, ny, nz;
    double ox, oy, oz;
    double h;
    __host__ __device__ int64_t size() const { return (int64_t)nx * ny * nz; }
};
int64_t boundary_count(const GridDesc &g);
size_t dh_work_size(const GridDesc &g, int n);  // doubles of scratch
// phi_a: the medium with kappa and eps_a; phi_b (or null): a salt-free medium
// with eps_b, from the same pass over the atoms.
void launch_dh_boundary(const GridDesc &g, int n, const double *xyz, const double *q,
                        double kappa, double eps_a, double *phi_a, double eps_b, double *phi_b,
                        double *work, hipStream_t s);

// Decomposition atom sums: for each atom a in A, sum over b in B of the
// half pair vdW / elec (dd) into work, then reduce into out_vdw / out_elec.
int decomp_chunks(int nB);
void launch_decomp_pairs(const double *xyz, const PairTopo &t, const int *A, int nA,
                         const int *B, int nB, int a_is_rec, int cross_excl, double *work,
                         double *out_vdw, double *out_elec, hipStream_t s);

// SASA per atom (Shrake-Rupley) using a cell list.
struct CellList {
    double ox, oy, oz, inv;
    int dx, dy, dz;
    const int *start;   // [ncell + 1]
    const int *atoms;   // [n]
};
// statistics of n frames' (dMM, dPolar, dSA, dTotal) at kT (10 doubles out:
// four means, std of total, mm, polar, sa, SEM of total, -T dS)
void launch_frame_stats(const double *terms, int n, double kt, double *out10, hipStream_t s);
void launch_sasa(const double *xyz, const double *radii, int n, double probe,
                 const double *pts, int n_pts, const CellList &cl, double max_ext,
                 double *per_atom, int *overflow, hipStream_t s);

// ---- grid kernels (grid.hip) ------------------------------------------------
void launch_cell_build(const double *xyz, int n, double ox, double oy, double oz, double inv,
                       int dx, int dy, int dz, int *cell_of, int *start, int *fill, int *atoms,
                       hipStream_t s);
void launch_spread(const GridDesc &g, int n, const double *xyz, const double *q,
                   unsigned long long *acc, double *rho, hipStream_t s);
// inside[i] = exists atom a (of n) with |p_i - x_a|^2 < (r_a + probe)^2
void launch_inside(const GridDesc &g, int n, const double *xyz, const double *radii, double probe,
                   uint8_t *inside, hipStream_t s);
// the maps of probes pa and pb from one pass (equal to two launch_inside)
void launch_inside2(const GridDesc &g, int n, const double *xyz, const double *radii, double pa,
                    double pb, uint8_t *out_a, uint8_t *out_b, hipStream_t s);
// SES classification from SAS maps (flood fill of the exterior + bounded
// distance transform) for many maps: flood rounds of every map (one host check
// per few rounds), then each map's distance test.
struct SesJob {
    GridDesc g;
    const uint8_t *sas;
    uint8_t *ext, *work_a, *work_b, *inside_out;
    int thresh;
};
struct FloodJob {
    GridDesc g;
    const uint8_t *sas;
    uint8_t *ext;
    int *changed;
    int *active;  // 0 once a round changed nothing in this map
    // bit-packed flood: rows of W words along x, word (w, y, z) at
    // (z ny + y) W + w; open = outside-SAS nodes, extb = exterior nodes
    int W;
    uint32_t *open, *extb;
};
// words the bit-packed flood of these jobs needs (open and exterior bits of
// every map); 0 when a y or z line is longer than it takes (512 nodes)
int64_t ses_bits_words(const std::vector<SesJob> &jobs);
// d_table: device room for jobs.size() FloodJobs; d_flags: 2 jobs.size() ints;
// d_bits: ses_bits_words(jobs) words for the bit-packed flood, or null for
// the byte flood (one thread group per line of bytes, launches per map)
void launch_ses_batch(const std::vector<SesJob> &jobs, FloodJob *d_table, int *d_flags,
                      uint32_t *d_bits, hipStream_t s);
// Face dielectrics (FP64 + FP32 copies) from the SES map.
// The same faces in two launches (trivial faces per node, then a wavefront per
// mixed edge): list: cap entries (faces_list_cap(N) is ample); count: one int,
// zeroed by the caller, the mixed edges found (above cap: the faces are
// incomplete, run launch_faces).
int faces_list_cap(int64_t n);
void launch_faces_split(const GridDesc &g, const uint8_t *inside, const double *xyz,
                        const double *radii, double max_r, const CellList &cl, double eps_in,
                        double eps_out, double *ex, double *ey, double *ez, float *fx, float *fy,
                        float *fz, uint64_t *list, int cap, int *count, int *overflow,
                        hipStream_t s);
void launch_faces(const GridDesc &g, const uint8_t *inside, const double *xyz,
                  const double *radii, double max_r, const CellList &cl, double eps_in,
                  double eps_out, double *ex, double *ey, double *ez, float *fx, float *fy,
                  float *fz, int *overflow, hipStream_t s);
void launch_kappa_map(const GridDesc &g, const uint8_t *ion_inside, double kbar2, double *k2,
                      float *k2f, hipStream_t s);
void launch_interp_boundary(const GridDesc &fine, const GridDesc &coarse, const double *cphi,
                            double *phi, hipStream_t s);
void launch_energy_gather(const GridDesc &g, const double *phi, int n, const double *xyz,
                          const double *q, double *partials, int *n_blocks, hipStream_t s);
// out[a] = 0.5 q_a (phi_solv - phi_ref)(r_a)
void launch_pb_atom_terms(const GridDesc &g, const double *phi_solv, const double *phi_ref, int n,
                          const double *xyz, const double *q, double *out, hipStream_t s);
void launch_maxabs(const double *v, int64_t n, double *partials, int *n_blocks, hipStream_t s);
void launch_max_partials(const double *partials, int n_blocks, double *out, hipStream_t s);
void launch_fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s);
void launch_fill_f64(double *p, double v, int64_t n, hipStream_t s);
void launch_fill_f32(float *p, float v, int64_t n, hipStream_t s);
void launch_fill_i32(int *p, int v, int64_t n, hipStream_t s);
void launch_fill_u64(unsigned long long *p, unsigned long long v, int64_t n, hipStream_t s);
void launch_gather_xyz(const double *src, const int *idx, int n, double *dst, hipStream_t s);
// mn[3] then mx[3] of n AoS points into out6 (device)
void launch_bbox(const double *xyz, int n, double *out6, hipStream_t s);
// plain device copy of n16 16-byte words
void launch_copy16(const void *src, void *dst, int64_t n16, hipStream_t s);
void launch_segment_sum(const double *v, const int *seg_atoms, const int *seg_ptr, int n_seg,
                        double *out, hipStream_t s);

}  // namespace mmpbsa
