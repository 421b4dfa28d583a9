// engine.h -- host runtime of the MM-PBSA engine: topology, device
// residency, memory pool, per-term drivers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

#include "../../include/mmpbsa_hip.h"
#include "kernels.h"
#include "mg.h"

namespace mmpbsa {

// ---- host topology: AmberTopology fields the path reads (prmtop.rs:34-105)
struct HostTopo {
    int n = 0, n_res = 0, n_types = 0;
    std::vector<double> q, qa, radii, screen;
    std::vector<int> type;
    std::vector<int64_t> nbidx;
    std::vector<double> ljA, ljB;
    std::vector<int> res_ptr;
    std::vector<int> bonds, angles, dihs;  // 3 / 4 / 6 ints per term
    std::vector<double> bond_k, bond_r0, angle_k, angle_t0, dih_pk, dih_pn, dih_ph;
    std::vector<int> nex, exl;
    double scee = 1.2, scnb = 2.0;

    static HostTopo from_view(const mmpbsa_topology_view &v);
    std::vector<int> atom_residue_indices() const;                    // prmtop.rs:120-132
    std::vector<int> build_selection(const std::vector<int> &res) const;  // prmtop.rs:165-204
    HostTopo extract(const std::vector<int> &atoms) const;            // subsystem.rs:22-169
};

// ---- device memory pool (per worker; reuse follows stream order) ----------
class Pool {
public:
    ~Pool();
    void *get(size_t bytes);
    void put(void *p);
    void trim();
    size_t cached() const { return cached_; }

private:
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> live_;
    size_t cached_ = 0;
};

template <typename T>
struct DBuf {
    Pool *pool = nullptr;
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(Pool &pl, size_t count) : pool(&pl), n(count)
    {
        p = static_cast<T *>(pl.get(sizeof(T) * (count ? count : 1)));
    }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; }
    DBuf &operator=(DBuf &&o) noexcept
    {
        release();
        pool = o.pool;
        p = o.p;
        n = o.n;
        o.p = nullptr;
        return *this;
    }
    ~DBuf() { release(); }
    void release()
    {
        if (p && pool) pool->put(p);
        p = nullptr;
    }
    T *get() const { return p; }
};

// A graph capture in one thread and a call that synchronises with the legacy
// stream or the device in another (a synchronous copy, a free, a launch on
// the null stream) must not overlap: HIP rejects the call while a capture is
// open.  Captures hold this gate shared, such calls hold it exclusively.
std::shared_mutex &capture_gate();

// Owned device array (not pooled): topology residency.
template <typename T>
struct DArr {
    T *p = nullptr;
    size_t n = 0;
    DArr() = default;
    DArr(const DArr &) = delete;
    DArr &operator=(const DArr &) = delete;
    ~DArr()
    {
        if (!p) return;
        std::unique_lock<std::shared_mutex> g(capture_gate());
        (void)hipFree(p);
    }
    void upload(const std::vector<T> &v)
    {
        std::unique_lock<std::shared_mutex> g(capture_gate());
        if (p) (void)hipFree(p);
        p = nullptr;
        n = v.size();
        HIP_OK(hipMalloc(&p, sizeof(T) * (n ? n : 1)));
        if (n) HIP_OK(hipMemcpy(p, v.data(), sizeof(T) * n, hipMemcpyHostToDevice));
    }
};

// ---- one topology resident on the device ----------------------------------
struct DevTopo {
    HostTopo h;
    double max_radius = 0.0;
    int cross_dummy = 0;
    DArr<double> q, qa, radii, screen, ljA, ljB;
    DArr<int> type, ex_ptr, ex_col, p14_ptr, p14_col;
    DArr<uint8_t> tile_flag;
    DArr<int> bonds, angles, dihs;
    DArr<double> bond_k, bond_r0, angle_k, angle_t0, dih_pk, dih_pn, dih_ph;
    std::vector<int> h_ex_ptr, h_ex_col, h_p14_ptr, h_p14_col;
    int nb = 1;
    void build(const HostTopo &t);
    PairTopo view() const;
};

// A worker: one stream, its pool and small pinned staging areas.
struct Worker {
    hipStream_t stream = nullptr;
    Pool pool;
    double *pin = nullptr;   // pinned host doubles
    int *pin_i = nullptr;    // pinned host ints
    Worker();
    ~Worker();
    void sync();
    // further workers (own stream and pool) for work that runs beside this
    // one's in other host threads, made on first use: 0 the frames' MM, GB
    // and SA terms, 1 the fine dielectric maps during the coarse PB stage
    Worker &aux(int i);

private:
    std::unique_ptr<Worker> aux_[2];
};

// Results of one PB solve on one grid (solve_lpbe_multigrid semantics).
struct SolveInfo {
    int iterations = 0;
    int converged = 0;
    double final_residual = 0.0;
};

// quasi_harmonic_entropy (entropy.rs:97-217) of nf frames [nf][na][3] (A)
// with masses (amu): -T dS in kcal/mol (NaN where the reference's is); at
// most qh_max_dim() coordinates (3 na)
int64_t qh_max_dim();
double qh_entropy(Worker &w, const double *h_xyz, int nf, int na, const double *h_masses,
                  double temperature);

// ---- drivers (device coordinates already resident on w.stream) ------------
void run_mm(Worker &w, const DevTopo &t, const double *d_xyz, mmpbsa_mm_terms *out);
// complex, receptor and ligand MM terms of one frame from the partners' pairs
// and the receptor x ligand pairs (a split whose complex is exactly the two
// partners, with no excluded, 1-4 or bonded term across them)
void run_mm_parts(Worker &w, const DevTopo &ct, const DevTopo &rt, const DevTopo &lt,
                  const double *dc, const double *dr, const double *dl, const int *d_ra, int nr,
                  const int *d_la, int nl, mmpbsa_mm_terms out[3]);
// complex, receptor and ligand GB energies of one frame, the complex's Born
// radii from the partners' psi sums plus the cross sums (a split whose complex
// is exactly the two partners)
void run_gb_parts(Worker &w, const DevTopo &ct, const DevTopo &rt, const DevTopo &lt,
                  const double *dc, const double *dr, const double *dl, const int *d_ra, int nr,
                  const int *d_la, int nl, const mmpbsa_gb_params &p, double out[3]);
double run_gb(Worker &w, const DevTopo &t, const double *d_xyz, const mmpbsa_gb_params &p,
              double *d_born /* nullable: pooled scratch used */, std::vector<double> *born_out);
double run_gb_from_radii(Worker &w, const DevTopo &t, const double *d_xyz,
                         const mmpbsa_gb_params &p, const double *d_born);
// h_xyz (host copy of d_xyz) may be null: the bounding box is then taken on
// the device.
void run_sa(Worker &w, const DevTopo &t, const double *d_xyz, const double *h_xyz,
            const mmpbsa_sa_params &p, double *d_per_atom_out, std::vector<double> *per_atom,
            double *total_sasa, double *total);
void run_pb(Worker &w, const DevTopo &t, const double *d_xyz, const double *h_xyz,
            const mmpbsa_pb_params &p, mmpbsa_pb_result *out);

// One PB energy (compute_pb_energy) of one subsystem in one frame.  Jobs run
// together through run_pb_jobs: every multigrid solve of a stage (all coarse
// focusing solves, then all fine solves) advances in shared launches.
struct PbJob {
    const DevTopo *t = nullptr;
    const double *d_xyz = nullptr;
    const double *h_xyz = nullptr;  // host copy, or null (bounding box on the device)
    mmpbsa_pb_params p{};
    mmpbsa_pb_result *out = nullptr;
    double *d_atom_terms = nullptr;  // device [n] 0.5 q (phi_solv - phi_ref)(r), or null
    struct State;
    std::shared_ptr<State> st;
};
// batched: replay solver cycles from a hipGraph (else launch them directly).
void run_pb_jobs(Worker &w, std::vector<PbJob> &jobs, bool batched);

// FP64 kappa (gb_energy.rs:126-139)
double compute_kappa(double salt, double eps, double temperature);

// ---- geometry shared by SA and PB (engine.cpp) -------------------------------
struct DevCells {
    DBuf<int> cell_of, start, fill, atoms;
    CellList cl;
};
void bbox(int n, const double *xyz, double *mn, double *mx);
// the same from device coordinates (one small readback)
void device_bbox(Worker &w, int n, const double *d_xyz, double *mn, double *mx);
GridDesc auto_grid(const double *mn, const double *mx, double h, double buffer);
GridDesc coarse_grid(const double *mn, const double *mx, double h, double fillratio,
                     int64_t fscale);
void build_cells(Worker &w, const double *d_xyz, int n, const double *mn, const double *mx,
                 double cs, DevCells &c);

// Launch-shape choices of the batched solver.  None of them changes a result
// bit: the phase smoother, the streaming restriction and prolongation are
// exact for any z-chunking, and the status checks only decide when the host
// looks.  Entries come from the autotuner (mmpbsa_tuning_set).
struct TuneEntry {
    int64_t nodes_min = 0, nodes_max = 0;  // level size range
    int solves_min = 0, solves_max = 0;    // solves sharing the launch
    int smoother_chunks = 0;               // z-chunks per column (0: heuristic)
    int smoother_region = 0;               // 1: 32 x 32, 2: 44 x 44 (0: heuristic)
};
struct Tuning {
    int64_t tail_nodes = 16384;
    int batch_solves = 1 << 20;
    int check_cycles = 8;   // V-cycles between host status checks
    int gb_chunks = 0;      // Born-radius j-chunks per atom block (0: heuristic)
    std::vector<TuneEntry> entries;
};
// A snapshot of the current tuning (the table may be replaced concurrently).
Tuning tuning();
void set_tuning(const Tuning &t);
// the smoother's z-chunk length for a level of these dimensions in a launch of
// n_solves solves: the tuned chunk count when an entry matches, else the
// built-in heuristic (smooth3_zc)
int tuned_smooth3_zc(const Tuning &t, int nx, int ny, int nz, int n_solves, int rg);
// the smoother region (0: 32 x 32, 1: 44 x 44) for such a level: the tuned
// one when an entry with a region matches, else smooth3_region
int tuned_smooth3_region(const Tuning &t, int nx, int ny, int nz, int n_solves);
// record the (level nodes, group solves) of batched smoother levels
void tuning_observe(bool on);
int tuning_observed(int64_t *nodes, int32_t *solves, int64_t *count, int cap);

// Kernel timing of the finest-grid smoother (no counterpart in the reference).
struct ProfileStats {
    bool enabled = false;
    double total_ms = 0.0;
    int64_t launches = 0;
    double alg_bytes = 0.0;
    int64_t grid_threads = 0;  // launch size (threads) of the last timed pass
};
ProfileStats &profile_stats();
std::mutex &profile_mutex();

// Named-kernel timing (measurement only, no counterpart in the reference):
// while profiling is enabled, a KernelTimer brackets the launches made in its
// scope with HIP events on the worker's stream and books them under `name`
// with `work` algorithmic units (FLOP or bytes, per the kernel's own rule).
class KernelTimer {
public:
    KernelTimer(Worker &w, const char *name, double work);
    ~KernelTimer();
    KernelTimer(const KernelTimer &) = delete;
    KernelTimer &operator=(const KernelTimer &) = delete;

private:
    Worker &w_;
    const char *name_;
    double work_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};
// Totals of one name (events of finished launches are read and released).
void kernel_profile_read(const char *name, double *ms, int64_t *launches, double *work);
void kernel_profile_reset();

}  // namespace mmpbsa
