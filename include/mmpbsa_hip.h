/*
 * mmpbsa_hip.h -- C ABI of the MI355X (gfx950) MM-PBSA solvation-energy engine.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes and
 * returns an int status (MMPBSA_OK == 0).  No C++ exception crosses the
 * boundary; the message of the last failure on a context is available from
 * mmpbsa_last_error().  Coordinates are Angstrom, AoS [n_atoms][3] float64,
 * exactly as the reference crate's `&[[f64; 3]]` arguments.
 *
 * Reference interface replaced (rust-simulation-tools 0.2.4, crate rst-mmpbsa):
 *   crates/mmpbsa/src/gb_energy.rs   GbModel :13, GbParams :24-56, compute_gb_energy :357
 *   crates/mmpbsa/src/mm_energy.rs   MmEnergy :43-58, compute_mm_energy :168-172
 *   crates/mmpbsa/src/pb_energy.rs   PbParams :21-66, compute_pb_energy :332-438
 *   crates/mmpbsa/src/sa_energy.rs   SaParams :12-32, compute_sa_energy :51-75
 *   crates/mmpbsa/src/binding.rs     FrameEnergy :34-56, compute_binding_energy_single_frame :749
 *   crates/mmpbsa/src/decomposition.rs  decompose_binding_energy :75, ResidueContribution :29-48
 *   crates/core/src/amber/prmtop.rs  AmberTopology :34-105 (the topology view below)
 * Threading: a context is externally synchronised (one host thread at a time);
 * several contexts may share the device.
 */
#ifndef MMPBSA_HIP_H
#define MMPBSA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MMPBSA_OK 0
#define MMPBSA_ERR_INVALID_ARG 1   /* null pointer, bad size, bad parameter   */
#define MMPBSA_ERR_BAD_RESIDUE 2   /* residue index out of range (prmtop.rs:167-175) */
#define MMPBSA_ERR_DEVICE 3        /* HIP runtime / kernel launch failure      */
#define MMPBSA_ERR_OOM 4           /* device allocation failed / over budget   */
#define MMPBSA_ERR_NO_DEVICE 5     /* no gfx950 device visible                 */
#define MMPBSA_ERR_NOT_READY 6     /* server ticket not finished yet           */
#define MMPBSA_ERR_TOO_LARGE 7     /* request exceeds a fixed capacity         */
#define MMPBSA_ERR_IO 8            /* file missing, unreadable or malformed    */

/* ---- topology view: caller-owned, read-only; copied at context creation --
 * Mirrors AmberTopology (crates/core/src/amber/prmtop.rs:34-105).
 * All indices are 0-based.  excluded_atoms uses -1 for AMBER's placeholder
 * entry (the reference's usize::MAX, prmtop.rs:674-683).                   */
typedef struct mmpbsa_topology_view {
    int32_t n_atoms;
    int32_t n_residues;
    int32_t n_types;
    const double *charges;        /* [n_atoms] elementary charges (PB)         */
    const double *charges_amber;  /* [n_atoms] charge * 18.2223 (GB, MM)       */
    const double *radii;          /* [n_atoms] intrinsic radii, Angstrom       */
    const double *screen;         /* [n_atoms] HCT screening factors           */
    const int32_t *atom_type;     /* [n_atoms] 0-based LJ type                 */
    const int64_t *nb_parm_index; /* [n_types*n_types] 1-based, <1 = no LJ     */
    const double *lj_acoef;       /* [n_lj]                                    */
    const double *lj_bcoef;       /* [n_lj]                                    */
    int32_t n_lj;
    const int32_t *residue_pointers; /* [n_residues] first atom of residue     */
    int32_t n_bonds;
    const int32_t *bonds;         /* [n_bonds][3]: i, j, type                  */
    const double *bond_k;         /* per bond type */
    const double *bond_r0;
    int32_t n_bond_types;
    int32_t n_angles;
    const int32_t *angles;        /* [n_angles][4]: i, j, k, type              */
    const double *angle_k;
    const double *angle_theta0;
    int32_t n_angle_types;
    int32_t n_dihedrals;
    const int32_t *dihedrals;     /* [n_dihedrals][6]: i, j, k, l, type, ignore_14 */
    const double *dihedral_pk;
    const double *dihedral_pn;
    const double *dihedral_phase;
    int32_t n_dihedral_types;
    const int32_t *num_excluded;  /* [n_atoms] NUMBER_EXCLUDED_ATOMS           */
    const int32_t *excluded_atoms;/* flat EXCLUDED_ATOMS_LIST, -1 = placeholder */
    int64_t n_excluded;           /* length of excluded_atoms                  */
    double scee;                  /* SCEE_SCALE_FACTOR (default 1.2)           */
    double scnb;                  /* SCNB_SCALE_FACTOR (default 2.0)           */
} mmpbsa_topology_view;

/* ---- parameter blocks (defaults = the reference's Default impls) -------- */
enum { MMPBSA_GB_HCT = 0, MMPBSA_GB_OBC1 = 1, MMPBSA_GB_OBC2 = 2 }; /* gb_energy.rs:13-21 */

typedef struct mmpbsa_gb_params {   /* gb_energy.rs:24-56 */
    int32_t model;                  /* default OBC-I */
    double solute_dielectric;       /* 1.0  */
    double solvent_dielectric;      /* 78.3 */
    double salt_concentration;      /* 0.15 */
    double temperature;             /* 300  */
    double offset;                  /* 0.09 */
    double rgbmax;                  /* 25.0 */
    double cutoff;                  /* 9999 */
} mmpbsa_gb_params;

typedef struct mmpbsa_pb_params {   /* pb_energy.rs:21-66 */
    double grid_spacing;            /* 0.5   */
    double grid_buffer;             /* 10.0  */
    double solute_dielectric;       /* 1.0   */
    double solvent_dielectric;      /* 80.0  */
    double salt_concentration;      /* 0.15  */
    double temperature;             /* 300   */
    double probe_radius;            /* 1.4   */
    double ion_radius;              /* 2.0   */
    double tolerance;               /* 1e-6  */
    int64_t max_iterations;         /* 10000 */
    double fillratio;               /* 4.0   */
    int64_t fscale;                 /* 8     */
    /* Extension (no counterpart in PbParams): when fine_dims[0] > 0 the fine
     * grid uses these dimensions instead of auto_grid's, centred on the
     * subsystem's bounding-box centre exactly as auto_grid centres
     * (pb_grid.rs:188-215).  Zero keeps the reference behaviour.           */
    int32_t fine_dims[3];
    /* Extension: explicit focusing grids shared by all subsystems (a fixed
     * frame, as a coarse-to-fine setup around an interface needs).  When
     * coarse_dims[0] > 0 the coarse grid of pb_solve_with_focusing
     * (pb_energy.rs:134-322) has these dimensions and coarse_spacing and is
     * centred at coarse_center (origin = centre - (n-1) h / 2 per axis), in
     * place of build_coarse_grid's; when fine_centered != 0 the fine grid
     * (auto_grid's or fine_dims) is centred at fine_center in place of the
     * subsystem's bounding-box centre.  Zeros keep the reference behaviour.  */
    int32_t coarse_dims[3];
    double coarse_spacing;
    double coarse_center[3];
    int32_t fine_centered;
    double fine_center[3];
} mmpbsa_pb_params;

typedef struct mmpbsa_sa_params {   /* sa_energy.rs:12-32 */
    double surface_tension;         /* 0.0072 */
    double offset;                  /* 0.0    */
    double probe_radius;            /* 1.4    */
    int32_t n_sphere_points;        /* 480    */
} mmpbsa_sa_params;

void mmpbsa_gb_params_default(mmpbsa_gb_params *p);
void mmpbsa_pb_params_default(mmpbsa_pb_params *p);
void mmpbsa_sa_params_default(mmpbsa_sa_params *p);

/* ---- result blocks ------------------------------------------------------ */
typedef struct mmpbsa_mm_terms {    /* MmEnergy, mm_energy.rs:43-58 */
    double bond, angle, dihedral, vdw, elec, vdw_14, elec_14;
} mmpbsa_mm_terms;

typedef struct mmpbsa_pb_solve_info {   /* PbSolveResult, pb_solver.rs:143-152 */
    double energy;          /* 1/2 sum q phi on the fine grid, kcal/mol     */
    double final_residual;  /* RMS residual of the fine solve               */
    int32_t iterations;     /* V-cycles (or SOR sweeps) of the fine solve  */
    int32_t converged;      /* 1 if the fine solve met the tolerance        */
    int32_t focused;        /* 1 if the coarse focusing solve was accepted  */
    int32_t coarse_iterations;
} mmpbsa_pb_solve_info;

typedef struct mmpbsa_pb_result {   /* PbEnergy, pb_energy.rs:68-73 + diagnostics */
    double total;           /* E_solv - E_ref */
    mmpbsa_pb_solve_info solv;
    mmpbsa_pb_solve_info ref;
    int32_t fine_dims[3];
    int32_t coarse_dims[3];
} mmpbsa_pb_result;

typedef struct mmpbsa_subsystem_energy { double mm, polar, sa; } mmpbsa_subsystem_energy; /* binding.rs:34-45 */

typedef struct mmpbsa_frame_terms {    /* FrameEnergy, binding.rs:47-56 */
    mmpbsa_subsystem_energy complex_, receptor, ligand;
    double delta_mm, delta_polar, delta_sa, delta_total;
} mmpbsa_frame_terms;

typedef struct mmpbsa_residue_contrib { /* ResidueContribution, decomposition.rs:29-42 */
    int32_t residue_index;  /* 0-based index in the caller's topology */
    double vdw, elec, polar, sa;
} mmpbsa_residue_contrib;

enum { MMPBSA_METHOD_GB = 0, MMPBSA_METHOD_PB = 1 };  /* SolvationMethod, binding.rs:24-31 */

/* ---- library ------------------------------------------------------------ */
/* Returns MMPBSA_OK when a gfx950 device is usable; *n_devices may be NULL. */
int mmpbsa_init(int32_t *n_devices);
const char *mmpbsa_version(void);

/* ---- context: one topology resident on the device ----------------------- */
typedef struct mmpbsa_ctx mmpbsa_ctx;

/* Copies the view (host and device); the caller may free it afterwards. */
int mmpbsa_ctx_create(const mmpbsa_topology_view *top, int32_t device, mmpbsa_ctx **out);
int mmpbsa_ctx_destroy(mmpbsa_ctx *ctx);
const char *mmpbsa_last_error(const mmpbsa_ctx *ctx);
int32_t mmpbsa_ctx_n_atoms(const mmpbsa_ctx *ctx);

/* compute_mm_energy (mm_energy.rs:168) */
int mmpbsa_mm_energy(mmpbsa_ctx *ctx, const double *xyz, mmpbsa_mm_terms *out);
/* compute_gb_energy (gb_energy.rs:357); born_radii_out [n_atoms] may be NULL */
int mmpbsa_gb_energy(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_gb_params *p,
                     double *born_radii_out, double *total_out);
/* compute_sa_energy (sa_energy.rs:51); per_atom_sasa_out may be NULL */
int mmpbsa_sa_energy(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_sa_params *p,
                     double *per_atom_sasa_out, double *total_sasa_out, double *total_out);
/* compute_pb_energy (pb_energy.rs:332) */
int mmpbsa_pb_energy(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_pb_params *p,
                     mmpbsa_pb_result *out);

/* compute_pb_energy_solvated (pb_energy.rs:442-482): with solute_res (0-based
 * residue indices, n_solute of them) the topology is stripped to those
 * residues (strip_to_solute: build_selection + extract_subtopology, the
 * coordinates sliced alike) before compute_pb_energy; solute_res == NULL is
 * mmpbsa_pb_energy.  A residue index >= n_residues is MMPBSA_ERR_BAD_RESIDUE
 * (prmtop.rs:167-175).  The stripped topology stays resident on the context
 * while later calls name the same residues. */
int mmpbsa_pb_energy_solvated(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_pb_params *p,
                              const int32_t *solute_res, int32_t n_solute, mmpbsa_pb_result *out);

/* compute_gb_energy_from_radii (gb_energy.rs:283-355): the GB energy of the
 * given Born radii [n_atoms] without computing radii (the step that
 * compute_gb_energy_cached, gb_energy.rs:372-387, takes on a cache hit). */
int mmpbsa_gb_energy_from_radii(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_gb_params *p,
                                const double *born_radii, double *total_out);

/* GbCache (gb_energy.rs:66-120): Born radii of the last frame computed
 * through mmpbsa_gb_energy_cached, with its coordinates (host memory).  A
 * cache may serve any context; it is externally synchronised. */
typedef struct mmpbsa_gb_cache mmpbsa_gb_cache;
int mmpbsa_gb_cache_create(double rmsd_threshold, mmpbsa_gb_cache **out);
int mmpbsa_gb_cache_destroy(mmpbsa_gb_cache *cache);
/* compute_gb_energy_cached (gb_energy.rs:372-387): when the cache holds as
 * many atoms as the context and the RMSD to its coordinates (summed in atom
 * order, GbCache::try_reuse) is below the threshold, its radii are reused
 * unchanged; otherwise radii are computed and the cache takes them and xyz.
 * *reused (nullable) is 1 on reuse; born_radii_out [n_atoms] may be NULL. */
int mmpbsa_gb_energy_cached(mmpbsa_ctx *ctx, mmpbsa_gb_cache *cache, const double *xyz,
                            const mmpbsa_gb_params *p, double *born_radii_out, double *total_out,
                            int32_t *reused);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mmpbsa_gb_energy with every temporary (Born-radius
 * partial sums, the radii unless d_born_radii_out is given, energy partials)
 * in one fixed 16 MiB static device array per GPU instead of the library's
 * pool; the call allocates nothing.  Inputs and outputs are device pointers
 * on the context's device: d_xyz [n_atoms][3] Angstrom, d_born_radii_out
 * [n_atoms] (nullable), d_total_out [1].  Results equal mmpbsa_gb_energy's.
 * MMPBSA_ERR_TOO_LARGE when the temporaries exceed the array (about 200,000
 * atoms).  All calls on one GPU share the array: calls that overlap on one
 * GPU race. */
int mmpbsa_gb_energy_noalloc(mmpbsa_ctx *ctx, const double *d_xyz, const mmpbsa_gb_params *p,
                             double *d_born_radii_out, double *d_total_out);

/* ---- receptor / ligand split (binding.rs:142-211, subsystem.rs:22-169) -- */
typedef struct mmpbsa_split mmpbsa_split;

/* Residue lists are 0-based indices into ctx's topology.  Returns
 * MMPBSA_ERR_BAD_RESIDUE for an index >= n_residues (prmtop.rs:167-175). */
int mmpbsa_split_create(mmpbsa_ctx *ctx, const int32_t *receptor_res, int32_t n_rec,
                        const int32_t *ligand_res, int32_t n_lig, mmpbsa_split **out);
int mmpbsa_split_destroy(mmpbsa_split *s);
/* which: 0 = complex atoms (in the caller's topology), 1 = receptor atoms,
 * 2 = ligand atoms (both in the effective complex).  Copies up to cap ints. */
int mmpbsa_split_atoms(const mmpbsa_split *s, int32_t which, int32_t *out, int32_t cap,
                       int32_t *n_out);

/* compute_frame_energy (binding.rs:333-415).  xyz holds all atoms of the
 * caller's topology (the split slices the complex out, binding.rs:213-222). */
int mmpbsa_frame_energy(mmpbsa_split *s, const double *xyz, int32_t method,
                        const mmpbsa_gb_params *gb, const mmpbsa_pb_params *pb,
                        const mmpbsa_sa_params *sa, mmpbsa_frame_terms *out);

/* Many frames of one split: the per-frame body of compute_binding_energy
 * (binding.rs:563-689) over frames already read.  mode 0 = plain (one frame at
 * a time, host-synchronous, no graphs), mode 1 = batched (the PB solves of
 * groups of frames advance together in shared, graph-replayed launches).
 * xyz_frames is [n_frames][n_atoms][3] in host memory. */
int mmpbsa_frames_energy(mmpbsa_split *s, const double *xyz_frames, int32_t n_frames,
                         int32_t method, const mmpbsa_gb_params *gb,
                         const mmpbsa_pb_params *pb, const mmpbsa_sa_params *sa,
                         int32_t mode, mmpbsa_frame_terms *out);

/* As mmpbsa_frames_energy with the frames already in device memory (HBM) of
 * the context's device: d_xyz_frames is a device pointer (for example from
 * mmpbsa_device_alloc), read in place on the library's stream, and must be
 * fully written before the call.  Results are identical to
 * mmpbsa_frames_energy's.  No counterpart in the reference (its frames always
 * come from files). */
int mmpbsa_frames_energy_device(mmpbsa_split *s, const double *d_xyz_frames, int32_t n_frames,
                                int32_t method, const mmpbsa_gb_params *gb,
                                const mmpbsa_pb_params *pb, const mmpbsa_sa_params *sa,
                                int32_t mode, mmpbsa_frame_terms *out);

/* decompose_binding_energy (decomposition.rs:75).  rec_out/lig_out have one
 * entry per requested residue, in the order given at split creation. */
int mmpbsa_decompose(mmpbsa_split *s, const double *xyz, const mmpbsa_gb_params *gb,
                     const mmpbsa_sa_params *sa, mmpbsa_residue_contrib *rec_out,
                     mmpbsa_residue_contrib *lig_out);

/* decompose_binding_energy (decomposition.rs:75) with a PB polar term in place
 * of the GB-weighted share (decomposition.rs:311-352).  vdw, elec and sa are
 * as in mmpbsa_decompose; polar is, per residue, the sum over its atoms of
 * 0.5 q_i (dphi_cplx - dphi_iso)(r_i), dphi = phi_solv - phi_ref on each
 * subsystem's fine grid (compute_pb_energy, pb_energy.rs:332-438).  The sum
 * of polar over all residues is dG_PB, returned in *delta_pb (nullable).
 * Extension: the reference has no PB decomposition. */
int mmpbsa_decompose_pb(mmpbsa_split *s, const double *xyz, const mmpbsa_pb_params *pb,
                        const mmpbsa_sa_params *sa, mmpbsa_residue_contrib *rec_out,
                        mmpbsa_residue_contrib *lig_out, double *delta_pb);

/* ---- topology ingest (host; no device needed) ----------------------------
 * parse_prmtop (crates/core/src/amber/prmtop.rs:436-733): %FLAG sections,
 * indices made 0-based, bond / angle / dihedral entries divided by 3,
 * charges = CHARGE / 18.2223 (charges_amber keeps CHARGE), ignore_14 from a
 * negative third or fourth dihedral index, EXCLUDED_ATOMS_LIST 0 -> -1,
 * SCEE / SCNB_SCALE_FACTOR first value (defaults 1.2 / 2.0), missing MASS /
 * RADII / SCREEN read as zeros.  Errors are MMPBSA_ERR_IO with the
 * reference's messages (mmpbsa_prmtop_last_error); a per-atom section shorter
 * than n_atoms is also an error here.                                        */
typedef struct mmpbsa_prmtop mmpbsa_prmtop;

typedef struct mmpbsa_prmtop_info {
    int32_t n_atoms, n_residues, n_types;
    const double *masses;          /* [n_atoms] amu                              */
    const char *atom_names;        /* [n_atoms][5]: up to 4 characters, NUL-ended */
    const char *residue_labels;    /* [n_residues][5]                            */
} mmpbsa_prmtop_info;

int mmpbsa_prmtop_read(const char *path, mmpbsa_prmtop **out);
const char *mmpbsa_prmtop_last_error(void);
int mmpbsa_prmtop_destroy(mmpbsa_prmtop *t);
/* A topology view of the parsed file (pointers valid until destroy), ready
 * for mmpbsa_ctx_create. */
int mmpbsa_prmtop_view(const mmpbsa_prmtop *t, mmpbsa_topology_view *view);
int mmpbsa_prmtop_info_get(const mmpbsa_prmtop *t, mmpbsa_prmtop_info *info);

/* ---- trajectory ingest (host; no device needed) ---------------------------
 * DCD: crates/core/src/trajectory/dcd.rs (header :197-348, frames :351-416);
 * positions are returned in nm as there.  mdcrd: crates/mmpbsa/src/mdcrd.rs
 * (positions in A).  Errors are MMPBSA_ERR_IO with the reference's messages
 * (mmpbsa_traj_last_error).                                                 */
typedef struct mmpbsa_dcd_header {
    int64_t n_frames, n_atoms;
    int32_t start_timestep, timestep_interval;
    float timestep;
    int32_t has_unit_cell, is_charmm, is_64bit, big_endian;
    int64_t first_frame_offset, frame_size;
} mmpbsa_dcd_header;

typedef struct mmpbsa_dcd mmpbsa_dcd;
typedef struct mmpbsa_mdcrd mmpbsa_mdcrd;

const char *mmpbsa_traj_last_error(void);
int mmpbsa_dcd_open(const char *path, mmpbsa_dcd **out);
int mmpbsa_dcd_close(mmpbsa_dcd *d);
/* header, the number of title lines, and title k (NUL-terminated into buf) */
int mmpbsa_dcd_get_header(const mmpbsa_dcd *d, mmpbsa_dcd_header *out, int32_t *n_titles);
int mmpbsa_dcd_title(const mmpbsa_dcd *d, int32_t k, char *buf, int32_t cap);
int mmpbsa_dcd_seek(mmpbsa_dcd *d, int64_t frame);
int64_t mmpbsa_dcd_current(const mmpbsa_dcd *d);
/* next frame: xyz_nm [n_atoms][3]; box6 (nullable) a, b, c in nm and alpha,
 * beta, gamma in degrees, set when has_unit_cell; *got = 0 at the end. */
int mmpbsa_dcd_read_frame(mmpbsa_dcd *d, double *xyz_nm, double *box6, int32_t *got);

int mmpbsa_mdcrd_open(const char *path, int64_t n_atoms, int32_t has_box, mmpbsa_mdcrd **out);
int mmpbsa_mdcrd_close(mmpbsa_mdcrd *m);
int64_t mmpbsa_mdcrd_current(const mmpbsa_mdcrd *m);
/* next frame: xyz_A [n_atoms][3]; box3 (nullable) X, Y, Z in A; *got = 0 at the end */
int mmpbsa_mdcrd_read_frame(mmpbsa_mdcrd *m, double *xyz_A, double *box3, int32_t *got);

/* reimage_complex (binding.rs:234-308): unwrap each bonded component by
 * breadth-first minimum-image steps, then move components next to the first
 * one by the minimum image of their geometric centres.  bonds: [n_bonds][2]. */
int mmpbsa_reimage(double *xyz, int32_t n_atoms, const int32_t *bonds, int32_t n_bonds,
                   const double box[3]);

/* compute_binding_energy (binding.rs:461-744) over a trajectory file: frames
 * start_frame, start_frame + stride, ... below end_frame (end_frame < 0: all)
 * are read (format 0 = mdcrd in A, has_box; 1 = DCD in nm, converted),
 * sliced to the split's complex, reimaged when the file carries a box, and
 * processed in batches of 64 frames through the batched path.  The results
 * are kept in the split until mmpbsa_trajectory_frames copies them out.
 * No frames is MMPBSA_ERR_INVALID_ARG ("No frames found in trajectory"). */
enum { MMPBSA_TRAJ_MDCRD = 0, MMPBSA_TRAJ_DCD = 1 };
int mmpbsa_binding_energy_trajectory(mmpbsa_split *s, const char *path, int32_t format,
                                     int32_t has_box, int64_t stride, int64_t start_frame,
                                     int64_t end_frame, int32_t method,
                                     const mmpbsa_gb_params *gb, const mmpbsa_pb_params *pb,
                                     const mmpbsa_sa_params *sa, int32_t *n_frames);
/* Copies the last trajectory call's n frames (and the complex coordinates of
 * the last frame processed, [n_complex_atoms][3] A, nullable). */
int mmpbsa_trajectory_frames(const mmpbsa_split *s, mmpbsa_frame_terms *out, int32_t n,
                             double *last_frame_xyz);

/* ---- statistics and entropy over frames (host; no device needed) ---------- */
typedef struct mmpbsa_binding_stats {    /* BindingResult (binding.rs:704-743) */
    double mean_delta_mm, mean_delta_polar, mean_delta_sa, mean_delta_total;
    double std_delta_total, std_delta_mm, std_delta_polar, std_delta_sa;  /* sample std */
    double sem_delta_total;                                                /* std / sqrt(n) */
} mmpbsa_binding_stats;

/* binding.rs:704-743: means in frame order, sample standard deviations (0 for
 * one frame), SEM of delta_total.  n == 0 is MMPBSA_ERR_INVALID_ARG ("No
 * frames found in trajectory", binding.rs:700-702). */
int mmpbsa_binding_stats_compute(const mmpbsa_frame_terms *frames, int32_t n,
                                 mmpbsa_binding_stats *out);

/* interaction_entropy (entropy.rs:53-78): -T dS = kT ln <exp((dMM - <dMM>)/kT)>
 * by log-sum-exp, kT = 0.00198688 T.  Fewer than two frames (the reference's
 * None) is MMPBSA_ERR_INVALID_ARG. */
int mmpbsa_interaction_entropy(const mmpbsa_frame_terms *frames, int32_t n, double temperature,
                               double *minus_tds);

/* quasi_harmonic_entropy (entropy.rs:97-217; lib.rs:2226-2237): the
 * mass-weighted covariance of the atomic fluctuations over n_frames frames
 * xyz_frames[n_frames][n_atoms][3] (A, host memory), its eigenvalues, and the
 * quantum harmonic oscillator entropy of the modes, as -T dS in kcal/mol, on
 * GPU `device`.  Fewer than two frames or no atoms (the reference's None) is
 * MMPBSA_ERR_INVALID_ARG; more than 20,000 atoms is MMPBSA_ERR_TOO_LARGE (the
 * 3N x 3N covariance).  A negative or non-finite mass or a non-finite
 * coordinate gives NaN, as the reference's does. */
int mmpbsa_quasi_harmonic_entropy(int32_t device, const double *xyz_frames, int32_t n_frames,
                                  int32_t n_atoms, const double *masses, double temperature,
                                  double *minus_tds);

/* ---- serving (no counterpart in the reference) ---------------------------
 * A server owns one worker thread and HIP stream.  Requests (frames of one
 * split, copied at submission) wait in a queue; each scheduling round takes
 * frames from the waiting requests oldest first (continuous batching: new
 * requests join the next round) while the round holds at most
 * max_round_frames frames and the sum of their device-memory estimates fits
 * hbm_budget_bytes, and runs them through the batched path of
 * mmpbsa_frames_energy, frames of different complexes sharing solver
 * launches.  A frame's result depends only on its own inputs: it is identical
 * alone and in any mix.  Splits (and their contexts) must outlive their
 * requests.                                                                 */
typedef struct mmpbsa_server mmpbsa_server;

typedef struct mmpbsa_server_params {
    int64_t hbm_budget_bytes;   /* admission limit; 0 = 3/4 of free device memory */
    int32_t max_round_frames;   /* frames per scheduling round (default 16)      */
} mmpbsa_server_params;

typedef struct mmpbsa_request_stats {
    double queue_ms;            /* submission -> first frame scheduled           */
    double total_ms;            /* submission -> last frame finished             */
    int32_t n_frames;
    int32_t status;             /* MMPBSA_OK or the error of the request         */
} mmpbsa_request_stats;

void mmpbsa_server_params_default(mmpbsa_server_params *p);
int mmpbsa_server_create(const mmpbsa_server_params *p, int32_t device, mmpbsa_server **out);
/* Finishes the requests already submitted, then stops. */
int mmpbsa_server_destroy(mmpbsa_server *srv);
/* MMPBSA_ERR_TOO_LARGE when one frame alone exceeds the budget. */
int mmpbsa_server_submit(mmpbsa_server *srv, mmpbsa_split *split, const double *xyz_frames,
                         int32_t n_frames, int32_t method, const mmpbsa_gb_params *gb,
                         const mmpbsa_pb_params *pb, const mmpbsa_sa_params *sa,
                         uint64_t *ticket);
/* MMPBSA_ERR_NOT_READY until every frame is done; then copies n_frames
 * results to out (may be null), fills stats (may be null) and releases the
 * ticket.  wait blocks instead of returning NOT_READY. */
int mmpbsa_server_poll(mmpbsa_server *srv, uint64_t ticket, mmpbsa_frame_terms *out,
                       mmpbsa_request_stats *stats);
int mmpbsa_server_wait(mmpbsa_server *srv, uint64_t ticket, mmpbsa_frame_terms *out,
                       mmpbsa_request_stats *stats);

/* Statistics of a served request, computed on the device by the server thread
 * when the request's last frame finishes (before its waiters wake): the
 * BindingResult statistics (binding.rs:704-743) and interaction_entropy
 * (entropy.rs:53-78) of its frames at the request's temperature (the PB
 * parameters' for MMPBSA_METHOD_PB, else the GB parameters').  Fixed-order
 * FP64 tree reductions: equal to mmpbsa_binding_stats_compute and
 * mmpbsa_interaction_entropy over the returned frames to rounding. */
typedef struct mmpbsa_request_summary {
    mmpbsa_binding_stats stats;
    double minus_tds;     /* NaN for a single frame */
    double temperature;   /* K */
} mmpbsa_request_summary;
/* poll / wait that also fill summary (may be null). */
int mmpbsa_server_poll_summary(mmpbsa_server *srv, uint64_t ticket, mmpbsa_frame_terms *out,
                               mmpbsa_request_stats *stats, mmpbsa_request_summary *summary);
int mmpbsa_server_wait_summary(mmpbsa_server *srv, uint64_t ticket, mmpbsa_frame_terms *out,
                               mmpbsa_request_stats *stats, mmpbsa_request_summary *summary);

/* ---- device memory --------------------------------------------------------
 * For callers that stage inputs in HBM without another GPU runtime (no
 * counterpart in the reference).  Allocation is on `device`; copies are
 * synchronous.  mmpbsa_device_copy runs a plain copy kernel (bytes a multiple
 * of 16), whose hardware counters calibrate traffic measurements.            */
int mmpbsa_device_alloc(int32_t device, int64_t bytes, void **out);
int mmpbsa_device_free(void *ptr);
int mmpbsa_device_upload(void *dst, const void *src, int64_t bytes);
int mmpbsa_device_download(void *dst, const void *src, int64_t bytes);
int mmpbsa_device_copy(void *dst, const void *src, int64_t bytes);

/* ---- launch tuning (no counterpart in the reference) -----------------------
 * Launch shapes of the batched solver and the pair kernels, as chosen by an
 * autotuner (rust_simulation_tools.autotune) for the workloads at hand.  The
 * reference fixes its equivalents (8^3 SOR tiles and thread chunking,
 * pb_solver.rs:214-327, pb_grid.rs:233).  Smoother z-chunking and the check
 * interval never change a result bit; the Born-radius chunk count changes
 * the order of its partial sums (results within rounding).  A level of a
 * solve group uses the first entry whose node range holds its size and whose
 * solve range holds the group's solve count; otherwise the built-in choice. */
typedef struct mmpbsa_tune_entry {
    int64_t nodes_min, nodes_max;   /* level size range, nodes                     */
    int32_t solves_min, solves_max; /* solves sharing the launch                   */
    int32_t smoother_chunks;        /* z-chunks per column of the phase smoother   */
    int32_t smoother_region;        /* workgroup region: 0 built-in, 1 32 x 32
                                       nodes (20 x 20 core), 2 44 x 44 (32 x 32)  */
} mmpbsa_tune_entry;

typedef struct mmpbsa_tuning {
    int32_t check_cycles;  /* V-cycles between host status checks (default 8) */
    int32_t gb_chunks;     /* Born-radius j-chunks per atom block (0: built-in) */
    int32_t n_entries;     /* entries[0..n_entries) copied by mmpbsa_tuning_set */
    const mmpbsa_tune_entry *entries;
} mmpbsa_tuning;

/* NULL restores the built-in choices.  At most 256 entries. */
int mmpbsa_tuning_set(const mmpbsa_tuning *t);
/* Current values; up to cap entries copied to entries (nullable), their
 * total count in t->n_entries (t->entries is left NULL). */
int mmpbsa_tuning_get(mmpbsa_tuning *t, mmpbsa_tune_entry *entries, int32_t cap);
/* The smoother z-chunk length the solver would use for a level of nx*ny*nz
 * nodes in a group of n_solves solves (for reports). */
int32_t mmpbsa_tuning_smoother_zc(int32_t nx, int32_t ny, int32_t nz, int32_t n_solves);
/* The smoother region (1: 32 x 32, 2: 44 x 44) the solver would use for such
 * a level (for reports). */
int32_t mmpbsa_tuning_smoother_region(int32_t nx, int32_t ny, int32_t nz, int32_t n_solves);
/* Observation for the autotuner: while on, every batched multigrid level the
 * solver lays out is recorded as (level nodes, solves in its group) with a
 * count.  Turning it on clears the record.  mmpbsa_tuning_observed copies up
 * to cap records and returns how many there are. */
int mmpbsa_tuning_observe(int32_t on);
int32_t mmpbsa_tuning_observed(int64_t *nodes, int32_t *solves, int64_t *count, int32_t cap);

/* ---- measurement (no counterpart in the reference) ----------------------
 * While enabled, each level-0 smoother pass of the fine focusing stage (the
 * dominant kernel: one red-black sweep per launch) is bracketed by HIP events
 * on its own stream, and solver cycles are launched directly rather than
 * replayed from a graph.  mmpbsa_profile_read returns the summed event time,
 * the number of launches, their algorithmic bytes (per node: correction read
 * 4 unless the pass starts from zero, right-hand side 4, code byte 1 on
 * non-uniform maps, correction write 4, plus 0.5 when the pass prolongates)
 * and the size in threads of the last timed launch. */
int mmpbsa_profile_enable(int32_t on);
int mmpbsa_profile_read(double *total_ms, int64_t *launches, double *alg_bytes,
                        int64_t *grid_threads);
/* While profiling is enabled, these launches are also timed with HIP events
 * on the library's stream and booked by name with their algorithmic work:
 *   "mm_pair_kernel"     20 FLOP per unordered atom pair (FP64)
 *   "gb_radii"           32 FLOP per ordered pair (FP64; Born-radius integral)
 *   "gb_energy_kernel"   26 FLOP per unordered pair (FP64)
 *   "dh_boundary_kernel" 15 FLOP per boundary node and atom (FP32 terms)
 *   "faces_kernel"       37 bytes per grid node (SES byte read, 3 FP64 and
 *                        3 FP32 face values written)
 * mmpbsa_profile_kernel returns the summed time, launch count and work of one
 * name since profiling was last enabled (zeros for a name never timed). */
int mmpbsa_profile_kernel(const char *name, double *total_ms, int64_t *launches, double *work);

#ifdef __cplusplus
}
#endif
#endif /* MMPBSA_HIP_H */
