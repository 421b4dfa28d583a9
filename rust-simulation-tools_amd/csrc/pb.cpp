// pb.cpp -- the Poisson-Boltzmann pipeline (pb_energy.rs:134-438): grid
// assembly, batched multigrid solves (pb_solver.rs:1015-1167) and the energy
// gather, for one or many (subsystem, frame) jobs at a time.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <future>
#include <map>
#include <string>
#include <thread>

#include "engine.h"

namespace mmpbsa {

static const double kFourPiEc = 4.0 * kPi * kCoulomb;
// Launch-shape knobs of the batched solver (tuned per device, see bench.py):
// levels with at most tail_nodes nodes run inside the one-workgroup tail;
// at most batch_solves solves share a launch.
namespace {
std::mutex &tuning_mutex()
{
    static std::mutex m;
    return m;
}
Tuning &tuning_store()
{
    static Tuning t = [] {
        Tuning v;
        if (const char *e = std::getenv("MMPBSA_TAIL_NODES")) v.tail_nodes = std::atoll(e);
        if (const char *e = std::getenv("MMPBSA_BATCH_SOLVES")) v.batch_solves = std::atoi(e);
        return v;
    }();
    return t;
}
}  // namespace

Tuning tuning()
{
    std::lock_guard<std::mutex> lk(tuning_mutex());
    return tuning_store();
}

void set_tuning(const Tuning &t)
{
    std::lock_guard<std::mutex> lk(tuning_mutex());
    tuning_store() = t;
}

namespace {
// (level nodes, solves in the group) -> level-solve occurrences, recorded while
// observation is on (the autotuner's view of a workload)
std::mutex &observe_mutex()
{
    static std::mutex m;
    return m;
}
bool g_observe = false;
std::map<std::pair<int64_t, int>, int64_t> &observed()
{
    static std::map<std::pair<int64_t, int>, int64_t> m;
    return m;
}
}  // namespace

void tuning_observe(bool on)
{
    std::lock_guard<std::mutex> lk(observe_mutex());
    g_observe = on;
    observed().clear();
}

int tuning_observed(int64_t *nodes, int32_t *solves, int64_t *count, int cap)
{
    std::lock_guard<std::mutex> lk(observe_mutex());
    int k = 0;
    for (const auto &kv : observed()) {
        if (k < cap) {
            if (nodes) nodes[k] = kv.first.first;
            if (solves) solves[k] = kv.first.second;
            if (count) count[k] = kv.second;
        }
        ++k;
    }
    return k;
}

static void observe_level(int64_t nodes, int n_solves)
{
    std::lock_guard<std::mutex> lk(observe_mutex());
    if (g_observe) observed()[{nodes, n_solves}] += 1;
}

static bool entry_holds(const TuneEntry &e, int64_t nodes, int n_solves)
{
    return nodes >= e.nodes_min && nodes <= e.nodes_max && n_solves >= e.solves_min &&
           n_solves <= e.solves_max;
}

int tuned_smooth3_zc(const Tuning &t, int nx, int ny, int nz, int n_solves, int rg)
{
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (const TuneEntry &e : t.entries) {
        if (!entry_holds(e, nodes, n_solves) || e.smoother_chunks <= 0) continue;
        const int nch = std::min(e.smoother_chunks, std::max(1, nz - 2));
        return ceil_div(std::max(1, nz - 2), nch);
    }
    return smooth3_zc(nx, ny, nz, n_solves, rg);
}

int tuned_smooth3_region(const Tuning &t, int nx, int ny, int nz, int n_solves)
{
    const int64_t nodes = (int64_t)nx * ny * nz;
    for (const TuneEntry &e : t.entries)
        if (entry_holds(e, nodes, n_solves) && e.smoother_region > 0) return e.smoother_region - 1;
    return smooth3_region(nx, ny);
}

ProfileStats &profile_stats()
{
    static ProfileStats s;
    return s;
}

std::mutex &profile_mutex()
{
    static std::mutex m;
    return m;
}

namespace {
struct KernelBook {
    double ms = 0.0, work = 0.0;
    int64_t launches = 0;
};
struct PendingTimer {
    std::string name;
    double work;
    hipEvent_t a, b;
};
std::map<std::string, KernelBook> &kernel_books()
{
    static std::map<std::string, KernelBook> m;
    return m;
}
std::vector<PendingTimer> &pending_timers()
{
    static std::vector<PendingTimer> v;
    return v;
}
// with profile_mutex held
void drain_timers()
{
    for (PendingTimer &t : pending_timers()) {
        float ms = 0.0f;
        if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            KernelBook &k = kernel_books()[t.name];
            k.ms += ms;
            k.work += t.work;
            k.launches += 1;
        } else {
            (void)hipGetLastError();
        }
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    pending_timers().clear();
}
}  // namespace

KernelTimer::KernelTimer(Worker &w, const char *name, double work)
    : w_(w), name_(name), work_(work)
{
    if (!profile_stats().enabled) return;
    if (hipEventCreate(&a_) != hipSuccess || hipEventCreate(&b_) != hipSuccess ||
        hipEventRecord(a_, w_.stream) != hipSuccess) {
        (void)hipGetLastError();
        if (a_) (void)hipEventDestroy(a_);
        if (b_) (void)hipEventDestroy(b_);
        a_ = b_ = nullptr;
    }
}

KernelTimer::~KernelTimer()
{
    if (!a_) return;
    if (hipEventRecord(b_, w_.stream) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipEventDestroy(a_);
        (void)hipEventDestroy(b_);
        return;
    }
    std::lock_guard<std::mutex> lk(profile_mutex());
    pending_timers().push_back(PendingTimer{name_, work_, a_, b_});
}

void kernel_profile_read(const char *name, double *ms, int64_t *launches, double *work)
{
    std::lock_guard<std::mutex> lk(profile_mutex());
    drain_timers();
    const auto it = kernel_books().find(name);
    const KernelBook k = it == kernel_books().end() ? KernelBook{} : it->second;
    if (ms) *ms = k.ms;
    if (launches) *launches = k.launches;
    if (work) *work = k.work;
}

void kernel_profile_reset()
{
    std::lock_guard<std::mutex> lk(profile_mutex());
    drain_timers();
    kernel_books().clear();
}

struct MapSet {
    bool uni = false;  // uniform dielectric and zero kappa
    double eps_u = 1.0;
    DBuf<double> ex, ey, ez, k2;
    DBuf<float> fx, fy, fz, k2f;
    DBuf<uint8_t> cw;   // code bytes of the finest level
    double c_in = 1.0, c_out = 1.0, c_k2 = 0.0;
};

static double spectral_radius(int nx, int ny, int nz)
{
    return (std::cos(kPi / (double)nx) + std::cos(kPi / (double)ny) + std::cos(kPi / (double)nz)) /
           3.0;
}

static double estimate_omega(int nx, int ny, int nz)
{
    const double r = spectral_radius(nx, ny, nz);
    return 2.0 / (1.0 + std::sqrt(1.0 - r * r));
}

// Dielectric and kappa maps (pb_grid.rs:296-313, :827-985) for many maps at
// once (uniform media need none): the SAS tests, the flood fills (shared
// launches, one host check per round), the distance tests and the faces of
// every map are issued together, with one host check of the face overflow
// flags, so small maps do not leave the GPU idle between host round trips.
struct MapReq {
    const DevTopo *t;
    const double *d_xyz;
    const DevCells *c;
    GridDesc g;
    const mmpbsa_pb_params *p;
    double ein, eout;
    bool ionic;
    double kappa;
    MapSet *m;
};

static void build_maps_batch(Worker &w, std::vector<MapReq> &rq)
{
    struct Tmp {
        DBuf<uint8_t> sas, inside, ext, wa, wb, ion;
    };
    std::vector<Tmp> tmp(rq.size());
    std::vector<SesJob> ses;
    std::vector<size_t> faces;  // maps that need faces from an SES map
    for (size_t k = 0; k < rq.size(); ++k) {
        MapReq &r = rq[k];
        MapSet &m = *r.m;
        const GridDesc &g = r.g;
        const int64_t N = g.size();
        const bool uni_eps = std::fabs(r.ein - r.eout) < DBL_EPSILON;
        if (uni_eps && !r.ionic) {
            m.uni = true;
            m.eps_u = r.ein;
            continue;
        }
        m.uni = false;
        m.ex = DBuf<double>(w.pool, N);
        m.ey = DBuf<double>(w.pool, N);
        m.ez = DBuf<double>(w.pool, N);
        m.k2 = DBuf<double>(w.pool, N);
        m.fx = DBuf<float>(w.pool, N);
        m.fy = DBuf<float>(w.pool, N);
        m.fz = DBuf<float>(w.pool, N);
        m.k2f = DBuf<float>(w.pool, N);
        if (uni_eps) {
            launch_fill_f64(m.ex.get(), r.ein, N, w.stream);
            launch_fill_f64(m.ey.get(), r.ein, N, w.stream);
            launch_fill_f64(m.ez.get(), r.ein, N, w.stream);
            launch_fill_f32(m.fx.get(), (float)r.ein, N, w.stream);
            launch_fill_f32(m.fy.get(), (float)r.ein, N, w.stream);
            launch_fill_f32(m.fz.get(), (float)r.ein, N, w.stream);
            continue;
        }
        Tmp &tm = tmp[k];
        tm.inside = DBuf<uint8_t>(w.pool, N);
        const double probe = r.p->probe_radius;
        if (probe <= 0.0) {
            launch_inside(g, r.t->h.n, r.d_xyz, r.t->radii.p, 0.0, tm.inside.get(), w.stream);
        } else {
            const double th2 = (probe / g.h) * (probe / g.h);
            const double T = std::floor(th2);
            if (T > 250.0)
                throw DeviceError(MMPBSA_ERR_TOO_LARGE, "probe_radius / grid_spacing above 15.8");
            tm.sas = DBuf<uint8_t>(w.pool, N);
            tm.ext = DBuf<uint8_t>(w.pool, N);
            tm.wa = DBuf<uint8_t>(w.pool, N);
            tm.wb = DBuf<uint8_t>(w.pool, N);
            if (r.ionic) {
                // the ion-accessibility map of the kappa map from the same pass
                tm.ion = DBuf<uint8_t>(w.pool, N);
                launch_inside2(g, r.t->h.n, r.d_xyz, r.t->radii.p, probe, r.p->ion_radius,
                               tm.sas.get(), tm.ion.get(), w.stream);
            } else {
                launch_inside(g, r.t->h.n, r.d_xyz, r.t->radii.p, probe, tm.sas.get(), w.stream);
            }
            ses.push_back(SesJob{g, tm.sas.get(), tm.ext.get(), tm.wa.get(), tm.wb.get(),
                                 tm.inside.get(), (int)T});
        }
        faces.push_back(k);
    }
    if (!ses.empty()) {
        DBuf<FloodJob> table(w.pool, ses.size());
        DBuf<int> flags(w.pool, 2 * ses.size());
        // MMPBSA_FLOOD=bytes: the byte flood (a test of one against the other)
        const char *fe = std::getenv("MMPBSA_FLOOD");
        const int64_t words =
            (fe && std::strcmp(fe, "bytes") == 0) ? 0 : ses_bits_words(ses);
        DBuf<uint32_t> bits;
        if (words > 0) bits = DBuf<uint32_t>(w.pool, (size_t)words);
        launch_ses_batch(ses, table.get(), flags.get(), bits.get(), w.stream);
    }
    if (!faces.empty()) {
        // per map: the interval-overflow flag, then the mixed-edge count; one
        // mixed-edge list serves the maps in turn (stream order).
        // MMPBSA_FACES_LIST_CAP lowers the list's capacity (a test of the
        // fallback below).
        const size_t nf = faces.size();
        int64_t nmax = 0;
        for (size_t f = 0; f < nf; ++f) nmax = std::max(nmax, rq[faces[f]].g.size());
        int cap_env = -1;
        if (const char *e = std::getenv("MMPBSA_FACES_LIST_CAP")) cap_env = std::max(0, std::atoi(e));
        auto cap_of = [&](int64_t n) {
            const int c = faces_list_cap(n);
            return cap_env >= 0 ? std::min(c, cap_env) : c;
        };
        DBuf<int> ovf(w.pool, 2 * nf);
        DBuf<uint64_t> list(w.pool, (size_t)std::max(1, cap_of(nmax)));
        launch_fill_i32(ovf.get(), 0, 2 * (int64_t)nf, w.stream);
        auto issue = [&](size_t f, bool split) {
            MapReq &r = rq[faces[f]];
            MapSet &m = *r.m;
            // algorithmic bytes: the SES byte read, three FP64 and three FP32
            // face values written per node
            KernelTimer kt(w, "face_dielectrics", 37.0 * (double)r.g.size());
            if (split)
                launch_faces_split(r.g, tmp[faces[f]].inside.get(), r.d_xyz, r.t->radii.p,
                                   r.t->max_radius, r.c->cl, r.ein, r.eout, m.ex.get(),
                                   m.ey.get(), m.ez.get(), m.fx.get(), m.fy.get(), m.fz.get(),
                                   list.get(), cap_of(r.g.size()), ovf.get() + nf + f,
                                   ovf.get() + f, w.stream);
            else
                launch_faces(r.g, tmp[faces[f]].inside.get(), r.d_xyz, r.t->radii.p,
                             r.t->max_radius, r.c->cl, r.ein, r.eout, m.ex.get(), m.ey.get(),
                             m.ez.get(), m.fx.get(), m.fy.get(), m.fz.get(), ovf.get() + f,
                             w.stream);
        };
        for (size_t f = 0; f < nf; ++f) issue(f, true);
        std::vector<int> ho(2 * nf);
        auto fetch = [&]() {
            HIP_OK(hipMemcpyAsync(ho.data(), ovf.get(), sizeof(int) * ho.size(),
                                  hipMemcpyDeviceToHost, w.stream));
            w.sync();
        };
        fetch();
        // a map with more mixed edges than its list holds takes the
        // one-thread-per-node pass over every face
        bool again = false;
        for (size_t f = 0; f < nf; ++f)
            if (ho[nf + f] > cap_of(rq[faces[f]].g.size())) {
                launch_fill_i32(ovf.get() + f, 0, 1, w.stream);
                issue(f, false);
                again = true;
            }
        if (again) fetch();
        for (size_t f = 0; f < nf; ++f)
            if (ho[f]) throw DeviceError(MMPBSA_ERR_TOO_LARGE, "face interval list overflow");
    }
    for (size_t k = 0; k < rq.size(); ++k) {
        MapReq &r = rq[k];
        MapSet &m = *r.m;
        if (m.uni) continue;
        const int64_t N = r.g.size();
        if (r.ionic) {
            DBuf<uint8_t> ion;
            if (tmp[k].ion.get()) {
                ion = std::move(tmp[k].ion);
            } else {
                ion = DBuf<uint8_t>(w.pool, N);
                launch_inside(r.g, r.t->h.n, r.d_xyz, r.t->radii.p, r.p->ion_radius, ion.get(),
                              w.stream);
            }
            const double kb2 = r.p->solvent_dielectric * r.kappa * r.kappa;
            launch_kappa_map(r.g, ion.get(), kb2, m.k2.get(), m.k2f.get(), w.stream);
        } else {
            launch_fill_f64(m.k2.get(), 0.0, N, w.stream);
            launch_fill_f32(m.k2f.get(), 0.0f, N, w.stream);
        }
        m.c_in = r.ein;
        m.c_out = r.eout;
        m.c_k2 = r.ionic ? r.p->solvent_dielectric * r.kappa * r.kappa : 0.0;
        m.cw = DBuf<uint8_t>(w.pool, N);
        mg_pack_codes(r.g.nx, r.g.ny, r.g.nz, m.ex.get(), m.ey.get(), m.ez.get(), m.k2.get(),
                      m.c_in, m.c_out, m.cw.get(), w.stream);
    }
}

static MgOuterDev outer_view(const GridDesc &g, const MapSet &m, const double *rho)
{
    MgOuterDev O{};
    O.nx = g.nx;
    O.ny = g.ny;
    O.nz = g.nz;
    O.uni = m.uni ? 1 : 0;
    O.eps_u = m.eps_u;
    const double ih = 1.0 / (g.h * g.h);
    O.ihx = O.ihy = O.ihz = ih;
    O.scale = kFourPiEc;
    O.ex = m.ex.get();
    O.ey = m.ey.get();
    O.ez = m.ez.get();
    O.k2 = m.k2.get();
    O.cw = m.uni ? nullptr : m.cw.get();
    O.c_in = m.c_in;
    O.c_out = m.c_out;
    O.c_k2 = m.c_k2;
    O.rho = rho;
    O.fd_nx = FastDiv((uint32_t)std::max(1, g.nx));
    O.fd_nxy = FastDiv((uint32_t)std::max(1, g.nx * g.ny));
    O.partner = -1;
    return O;
}

__global__ void copy_boundary_kernel(GridDesc g, const double *__restrict__ src,
                                     double *__restrict__ dst)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.size()) return;
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const int z = (int)(idx / nxy);
    const int y = (int)((idx - z * nxy) / g.nx);
    const int x = (int)(idx - z * nxy - (int64_t)y * g.nx);
    if (x == 0 || y == 0 || z == 0 || x == g.nx - 1 || y == g.ny - 1 || z == g.nz - 1)
        dst[idx] = src[idx];
}

static MgStatus status_template(double tol, double count, int max_cycles)
{
    MgStatus s{};
    s.state = 0;
    s.cycles = 0;
    s.growing = 0;
    s.max_cycles = max_cycles;
    s.best = DBL_MAX;
    s.prev = DBL_MAX;
    s.rms = 0.0;
    s.tol = tol;
    s.count = count;
    s.diverge_at = 0;
    s.pad_ = 0;
    return s;
}

// MMPBSA_FORCE_DIVERGENCE_CYCLE=k (k > 0) makes the divergence test of every
// multigrid solve fire at cycle k, so that the fallback of
// pb_solver.rs:1110-1138 (Chebyshev SOR from the partial potential) can be
// exercised from the host; unset or 0 leaves the reference's test alone.
static int forced_divergence_cycle()
{
    const char *e = std::getenv("MMPBSA_FORCE_DIVERGENCE_CYCLE");
    return e ? std::max(0, std::atoi(e)) : 0;
}

// solve_lpbe_initial (pb_solver.rs:373-447): Chebyshev-ramped FP64 SOR from phi.
static SolveInfo sor_solve(Worker &w, const GridDesc &g, const MapSet &m, const double *rho,
                           double *phi, double tol, int64_t max_it)
{
    SolveInfo info;
    info.final_residual = DBL_MAX;
    if (max_it <= 0) return info;
    MgOuterDev O = outer_view(g, m, rho);
    O.phi_out = phi;
    const int nbk = sor64_blocks(O);
    DBuf<double> part(w.pool, 2 * (size_t)nbk);
    DBuf<MgStatus> st(w.pool, 1);
    const MgStatus init = status_template(tol, 0.0, 0);
    mg_status_init(st.get(), &init, 1, w.stream);
    const double nint = (g.nx > 2 && g.ny > 2 && g.nz > 2)
                            ? (double)(g.nx - 2) * (g.ny - 2) * (g.nz - 2)
                            : 0.0;
    const double rj = spectral_radius(g.nx, g.ny, g.nz);
    const double r2 = rj * rj;
    double omega = 1.0;
    int64_t k = 0;
    MgStatus hs{};
    while (k < max_it) {
        const int64_t chunk_end = std::min<int64_t>(max_it, k + 64);
        for (; k < chunk_end; ++k) {
            mg_sor64(O, 0, omega, part.get(), st.get(), w.stream);
            mg_sor64(O, 1, omega, part.get() + nbk, st.get(), w.stream);
            mg_sor_check(part.get(), 2 * nbk, nint, tol, (int)(k + 1), (int)max_it, st.get(),
                         w.stream);
            omega = (k == 0) ? 1.0 / (1.0 - r2 / 2.0) : 1.0 / (1.0 - r2 * omega / 4.0);
        }
        HIP_OK(hipMemcpyAsync(w.pin, st.get(), sizeof(MgStatus), hipMemcpyDeviceToHost, w.stream));
        w.sync();
        std::memcpy(&hs, w.pin, sizeof(MgStatus));
        if (hs.state != 0) break;
    }
    info.iterations = hs.cycles;
    info.converged = hs.state == 1;
    info.final_residual = hs.rms;
    return info;
}

// ---- batched multigrid -------------------------------------------------------
struct MgTask {
    GridDesc g{};
    const MapSet *m = nullptr;
    const double *rho = nullptr;
    double *phi = nullptr;  // boundary values in, solution out
    double tol = 1e-6;
    int64_t max_it = 10000;
    SolveInfo info;
};

namespace {

struct TaskLevel {
    GridDesc g;
    DBuf<float> ex, ey, ez, k2, e, e2, f, res;
    DBuf<uint8_t> cw;
    DBuf<float4> ent;  // fractional-face entries (levels the phase pass sweeps)
    MgLevelDev d;
};

struct TaskWork {
    MgTask *t = nullptr;
    std::vector<TaskLevel> L;
    int tail_start = 0;
    int maxv = 1;
    DBuf<double> buf1, bc, partials;
    bool fold = false;  // the phase pass adds the finest correction to phi
    MgOuterDev O[2];        // per cycle parity
    MgOuterDev O0;          // cycle-0 residual
};

int64_t lv_nodes(const GridDesc &g) { return g.size(); }

int64_t lv_work(const GridDesc &g)
{
    if (g.nx < 3 || g.ny < 3 || g.nz < 3) return 0;
    return (int64_t)(g.ny - 2) * (g.nz - 2) * ((g.nx - 1) / 2);
}

MgLevelDev level_with_e(const MgLevelDev &d, float *e, float *e2, double *phi)
{
    MgLevelDev r = d;
    r.e = e;
    r.e2 = e2;
    r.phi = phi;
    return r;
}

MgLevelDev null_level()
{
    MgLevelDev z{};
    return z;
}

}  // namespace

static void mg_solve_group(Worker &w, std::vector<MgTask *> &tasks, bool use_graph,
                           bool profile_level0);

// profile_level0: when profiling is enabled, time this batch's level-0
// smoother passes (the fine focusing stage: the dominant kernel launches).
static void mg_solve_batch(Worker &w, std::vector<MgTask *> &tasks, bool use_graph,
                           bool profile_level0)
{
    const size_t g = (size_t)std::max(1, tuning().batch_solves);  // one snapshot per batch
    for (size_t a = 0; a < tasks.size(); a += g) {
        std::vector<MgTask *> part(tasks.begin() + a, tasks.begin() + std::min(tasks.size(), a + g));
        mg_solve_group(w, part, use_graph, profile_level0);
    }
}

static void mg_solve_group(Worker &w, std::vector<MgTask *> &tasks, bool use_graph,
                           bool profile_level0)
{
    const Tuning tun = tuning();
    std::vector<MgTask *> mgt;
    for (MgTask *t : tasks) {
        const int md = std::min(t->g.nx, std::min(t->g.ny, t->g.nz));
        if (md <= 8)
            t->info = sor_solve(w, t->g, *t->m, t->rho, t->phi, t->tol, t->max_it);
        else
            mgt.push_back(t);
    }
    const int S = (int)mgt.size();
    if (S == 0) return;
    ProfileStats &prof = profile_stats();
    const bool profiling = prof.enabled && profile_level0;
    if (profiling) use_graph = false;

    // one pass per smoothing phase (three sweeps, smooth3_kernel) unless
    // MMPBSA_SMOOTHER=sweep asks for one pass per sweep (smooth_kernel)
    const char *sm_env = std::getenv("MMPBSA_SMOOTHER");
    const bool phase_pass = !(sm_env && std::strcmp(sm_env, "sweep") == 0);
    // z-streaming residual + restriction and prolongation unless MMPBSA_STREAM=0
    const char *st_env = std::getenv("MMPBSA_STREAM");
    const bool stream = !(st_env && std::strcmp(st_env, "0") == 0);
    // the phase pass adds the finest level's final correction to phi itself
    // (one FP64 potential per solve, no ping-pong copy; per solve: TaskWork::fold)
    const bool fold = phase_pass;

    // smoother region per level index: the solves of a launch share one, that
    // of the group's largest grid at the level (levels as built below)
    std::vector<int> level_rg;
    {
        std::vector<int64_t> largest;
        for (MgTask *t : mgt) {
            int d[3] = {t->g.nx, t->g.ny, t->g.nz};
            for (size_t l = 0;; ++l) {
                const int64_t N = (int64_t)d[0] * d[1] * d[2];
                if (l == largest.size()) {
                    largest.push_back(-1);
                    level_rg.push_back(0);
                }
                if (N > largest[l]) {
                    largest[l] = N;
                    level_rg[l] = tuned_smooth3_region(tun, d[0], d[1], d[2], S);
                }
                if (std::min(d[0], std::min(d[1], d[2])) <= 5) break;
                for (int k = 0; k < 3; ++k) d[k] = (d[k] + 1) / 2;
            }
        }
    }

    std::vector<TaskWork> W(S);
    int maxT = 0;
    for (int s = 0; s < S; ++s) {
        TaskWork &tw = W[s];
        MgTask *t = mgt[s];
        tw.t = t;
        tw.maxv = (int)std::min<int64_t>(200, std::max<int64_t>(1, t->max_it / 6));
        int d[3] = {t->g.nx, t->g.ny, t->g.nz};
        double h = t->g.h;
        for (;;) {  // build_multigrid_levels (pb_solver.rs:776-836), min size 5
            TaskLevel lv;
            lv.g = t->g;
            lv.g.nx = d[0];
            lv.g.ny = d[1];
            lv.g.nz = d[2];
            lv.g.h = h;
            tw.L.push_back(std::move(lv));
            if (std::min(d[0], std::min(d[1], d[2])) <= 5) break;
            for (int k = 0; k < 3; ++k) d[k] = (d[k] + 1) / 2;
            h *= 2.0;
        }
        const int nl = (int)tw.L.size();
        tw.tail_start = nl - 1;
        for (int l = 0; l < nl; ++l)
            if (lv_nodes(tw.L[l].g) <= tun.tail_nodes) {
                tw.tail_start = l;
                break;
            }
        maxT = std::max(maxT, tw.tail_start);
        const MapSet &m = *t->m;
        for (int l = 0; l < nl; ++l) {
            TaskLevel &lv = tw.L[l];
            const int64_t N = lv_nodes(lv.g);
            lv.e = DBuf<float>(w.pool, N);
            if (l < tw.tail_start || l == 0) lv.e2 = DBuf<float>(w.pool, N);
            lv.f = DBuf<float>(w.pool, N);
            if (l + 1 < nl) lv.res = DBuf<float>(w.pool, N);
            MgLevelDev &dv = lv.d;
            dv = MgLevelDev{};
            dv.nx = lv.g.nx;
            dv.ny = lv.g.ny;
            dv.nz = lv.g.nz;
            dv.uni = m.uni ? 1 : 0;
            dv.eps_u = (float)m.eps_u;
            const float ih = (float)(1.0 / (lv.g.h * lv.g.h));
            dv.ihx = dv.ihy = dv.ihz = ih;
            dv.omega = (float)estimate_omega(lv.g.nx, lv.g.ny, lv.g.nz);
            dv.rg = level_rg[l];
            dv.zc = phase_pass ? tuned_smooth3_zc(tun, lv.g.nx, lv.g.ny, lv.g.nz, S, dv.rg)
                               : smooth_zc(lv.g.nz);
            if (l < tw.tail_start) observe_level(N, S);
            dv.e = lv.e.get();
            dv.e2 = lv.e2.get();
            dv.f = lv.f.get();
            dv.res = lv.res.get();
            mg_level_index(dv);
            if (!m.uni) {
                if (l == 0) {
                    dv.ex = m.fx.get();
                    dv.ey = m.fy.get();
                    dv.ez = m.fz.get();
                    dv.k2 = m.k2f.get();
                    dv.cw = m.cw.get();
                    dv.c_in = (float)m.c_in;
                    dv.c_out = (float)m.c_out;
                    dv.c_k2 = (float)m.c_k2;
                } else {
                    lv.ex = DBuf<float>(w.pool, N);
                    lv.ey = DBuf<float>(w.pool, N);
                    lv.ez = DBuf<float>(w.pool, N);
                    lv.k2 = DBuf<float>(w.pool, N);
                    dv.ex = lv.ex.get();
                    dv.ey = lv.ey.get();
                    dv.ez = lv.ez.get();
                    dv.k2 = lv.k2.get();
                    lv.cw = DBuf<uint8_t>(w.pool, N);
                    dv.cw = lv.cw.get();
                    dv.c_in = (float)m.c_in;
                    dv.c_out = (float)m.c_out;
                    dv.c_k2 = (float)m.c_k2;
                    // injected faces and code bytes: one launch per level for
                    // every solve of the group, after this loop
                }
                // fractional-face entries of the levels the phase smoother
                // sweeps, filled once the faces and code bytes are there
                if (phase_pass && l < tw.tail_start) {
                    lv.ent = DBuf<float4>(w.pool, 2 * (size_t)N);
                    dv.ent = lv.ent.get();
                }
            }
        }
        const int64_t N0 = lv_nodes(t->g);
        // only where the phase pass smooths the finest level (not a grid the
        // tail kernel solves whole)
        tw.fold = fold && tw.tail_start > 0;
        if (!tw.fold) tw.buf1 = DBuf<double>(w.pool, N0);
        tw.bc = DBuf<double>(w.pool, N0);
        HIP_OK(hipMemcpyAsync(tw.bc.get(), t->phi, sizeof(double) * N0, hipMemcpyDeviceToDevice,
                              w.stream));
        MgOuterDev O = outer_view(t->g, m, t->rho);
        O.f0 = tw.L[0].f.get();
        O.nbk = ceil_div(N0, 256);
        tw.partials = DBuf<double>(w.pool, O.nbk);
        O.partials = tw.partials.get();
        double *bufs[2] = {t->phi, tw.buf1.get()};
        float *E[2] = {tw.L[0].e.get(), tw.L[0].e2.get()};
        tw.O0 = O;
        tw.O0.phi_in = tw.O0.phi_out = bufs[0];
        tw.O0.e = nullptr;
        for (int p = 0; p < 2; ++p) {
            tw.O[p] = O;
            if (tw.fold) {
                // the post-smoothing pass has already added e to phi in place
                tw.O[p].phi_in = tw.O[p].phi_out = t->phi;
                tw.O[p].e = nullptr;
            } else {
                tw.O[p].phi_in = bufs[1 - p];
                tw.O[p].phi_out = bufs[p];
                tw.O[p].e = E[p];
            }
        }
    }

    // Solves with the same grid and charge density (the solvated and reference
    // solves of one grid) pair up for the FP64 outer step: the lead's blocks
    // also do the partner's nodes, so rho is read once for both.
    std::vector<int> partner(S, -1), lead_of(S, -1);
    for (int s = 0; s < S; ++s) {
        if (lead_of[s] >= 0) continue;
        for (int s2 = s + 1; s2 < S; ++s2) {
            const GridDesc &a = W[s].t->g, &b = W[s2].t->g;
            if (lead_of[s2] >= 0 || partner[s2] >= 0 || W[s2].t->rho != W[s].t->rho ||
                a.nx != b.nx || a.ny != b.ny || a.nz != b.nz)
                continue;
            partner[s] = s2;
            lead_of[s2] = s;
            break;
        }
        W[s].O0.partner = W[s].O[0].partner = W[s].O[1].partner = partner[s];
    }

    // coarsened dielectrics and kappa (coarsen_dielectrics / coarsen_kappa,
    // pb_solver.rs:684-748) and their code bytes, level by level for all
    // solves with non-uniform maps at once
    std::vector<std::vector<MgLevelDev>> inj_host;
    std::vector<DBuf<MgLevelDev>> inj_dev;
    {
        size_t maxL = 0;
        for (const TaskWork &tw : W) maxL = std::max(maxL, tw.L.size());
        for (size_t l = 1; l < maxL; ++l) {
            std::vector<MgLevelDev> h(2 * (size_t)S, null_level());
            int64_t mx = 0;
            for (int s = 0; s < S; ++s) {
                const TaskWork &tw = W[s];
                if (tw.t->m->uni || l >= tw.L.size()) continue;
                h[s] = tw.L[l - 1].d;
                h[S + s] = tw.L[l].d;
                mx = std::max(mx, lv_nodes(tw.L[l].g));
            }
            if (mx == 0) continue;
            inj_dev.emplace_back(w.pool, 2 * (size_t)S);
            HIP_OK(hipMemcpyAsync(inj_dev.back().get(), h.data(), sizeof(MgLevelDev) * h.size(),
                                  hipMemcpyHostToDevice, w.stream));
            mg_inject_pack_batch(inj_dev.back().get(), inj_dev.back().get() + S, S, mx, w.stream);
            mg_fill_entries(inj_dev.back().get() + S, S, mx, w.stream);
            inj_host.push_back(std::move(h));
        }
        // level 0: the maps' own FP32 faces and code bytes
        std::vector<MgLevelDev> h0(S, null_level());
        int64_t mx0 = 0;
        for (int s = 0; s < S; ++s)
            if (W[s].L[0].d.ent) {
                h0[s] = W[s].L[0].d;
                mx0 = std::max(mx0, lv_nodes(W[s].L[0].g));
            }
        if (mx0 > 0) {
            inj_dev.emplace_back(w.pool, (size_t)S);
            HIP_OK(hipMemcpyAsync(inj_dev.back().get(), h0.data(), sizeof(MgLevelDev) * h0.size(),
                                  hipMemcpyHostToDevice, w.stream));
            mg_fill_entries(inj_dev.back().get(), S, mx0, w.stream);
            inj_host.push_back(std::move(h0));
        }
    }

    // ---- device descriptor tables (two parities) ----
    // per parity: F[maxT][S], C[maxT][S], T[S], O[S], ezero[S]; plus O0[S], ezero0[S]
    const size_t nF = (size_t)maxT * S;
    struct Host {
        std::vector<MgLevelDev> F, C;
        std::vector<MgTailDev> T;
        std::vector<MgOuterDev> O;
    } H[2];
    std::vector<int64_t> maxnodes(maxT, 0), maxcoarse(maxT, 0);
    std::vector<int> sblocks(maxT, 0);  // smoother launch width per level
    std::vector<int> strblk(maxT, 0);   // streaming-kernel launch width per level
    std::vector<int> rrblk(maxT, 0);    // residual + restriction width
    int64_t maxN0 = 0;
    for (int p = 0; p < 2; ++p) {
        H[p].F.assign(nF, null_level());
        H[p].C.assign(nF, null_level());
        H[p].T.assign(S, MgTailDev{});
        H[p].O.resize(S);
        for (int s = 0; s < S; ++s) {
            TaskWork &tw = W[s];
            // level 0 alternates its two correction buffers with the cycle
            // parity: the post-smoothed correction of parity p is read by that
            // cycle's outer step while the other buffer serves as e2
            float *E[2] = {tw.L[0].e.get(), tw.L[0].e2.get()};
            auto lev = [&](int l) {
                return l == 0 ? level_with_e(tw.L[0].d, E[p], E[1 - p], tw.fold ? tw.t->phi : nullptr)
                              : tw.L[l].d;
            };
            for (int l = 0; l < tw.tail_start; ++l) {
                H[p].F[(size_t)l * S + s] = lev(l);
                H[p].C[(size_t)l * S + s] = lev(l + 1);
                const GridDesc &g = tw.L[l].g;
                sblocks[l] = std::max(sblocks[l],
                                      phase_pass ? smooth3_blocks(g.nx, g.ny, g.nz, tw.L[l].d.zc,
                                                                  tw.L[l].d.rg)
                                                 : smooth_tiles(g.nx, g.ny) *
                                                       ceil_div(g.nz, smooth_zc(g.nz)));
                if (phase_pass && 4 * lv_nodes(g) >= ((int64_t)1 << 31))
                    throw DeviceError(MMPBSA_ERR_TOO_LARGE, "multigrid level above 2^29 nodes");
                maxnodes[l] = std::max(maxnodes[l], lv_nodes(tw.L[l].g));
                strblk[l] = std::max(strblk[l], mg_stream_blocks(g.nx, g.ny, g.nz, tw.L[l].d.zc));
                const GridDesc &gc = tw.L[l + 1].g;
                rrblk[l] = std::max(rrblk[l], mg_rr_blocks(gc.nx, gc.ny, gc.nz, tw.L[l + 1].d.zc));
                maxcoarse[l] = std::max(maxcoarse[l], lv_nodes(tw.L[l + 1].g));
            }
            MgTailDev &T = H[p].T[s];
            const int nl = (int)tw.L.size();
            T.n = nl - tw.tail_start;
            if (T.n > kMaxTail) throw DeviceError(MMPBSA_ERR_TOO_LARGE, "too many multigrid levels");
            for (int l = tw.tail_start; l < nl; ++l) T.lv[l - tw.tail_start] = lev(l);
            H[p].O[s] = tw.O[p];
            maxN0 = std::max(maxN0, lv_nodes(tw.t->g));
        }
    }
    std::vector<MgOuterDev> O0(S);
    for (int s = 0; s < S; ++s) O0[s] = W[s].O0;
    // one device blob holding every table
    auto bytes_of = [](size_t n, size_t sz) { return (n * sz + 255) & ~(size_t)255; };
    const size_t szF = bytes_of(nF ? nF : 1, sizeof(MgLevelDev));
    const size_t szT = bytes_of(S, sizeof(MgTailDev));
    const size_t szO = bytes_of(S, sizeof(MgOuterDev));
    const size_t per = 2 * szF + szT + szO;
    const size_t total = 2 * per + szO + bytes_of(S, sizeof(MgStatus));
    DBuf<char> blob(w.pool, total);
    std::vector<char> hb(total, 0);
    char *dbase = blob.get();
    size_t off = 0;
    struct Dev {
        MgLevelDev *F, *C;
        MgTailDev *T;
        MgOuterDev *O;
    } D[2];
    for (int p = 0; p < 2; ++p) {
        std::memcpy(hb.data() + off, H[p].F.data(), nF * sizeof(MgLevelDev));
        D[p].F = (MgLevelDev *)(dbase + off);
        off += szF;
        std::memcpy(hb.data() + off, H[p].C.data(), nF * sizeof(MgLevelDev));
        D[p].C = (MgLevelDev *)(dbase + off);
        off += szF;
        std::memcpy(hb.data() + off, H[p].T.data(), S * sizeof(MgTailDev));
        D[p].T = (MgTailDev *)(dbase + off);
        off += szT;
        std::memcpy(hb.data() + off, H[p].O.data(), S * sizeof(MgOuterDev));
        D[p].O = (MgOuterDev *)(dbase + off);
        off += szO;
    }
    std::memcpy(hb.data() + off, O0.data(), S * sizeof(MgOuterDev));
    MgOuterDev *dO0 = (MgOuterDev *)(dbase + off);
    off += szO;
    std::vector<MgStatus> init(S);
    const int div_at = forced_divergence_cycle();
    for (int s = 0; s < S; ++s) {
        const GridDesc &g = W[s].t->g;
        init[s] = status_template(W[s].t->tol, (double)(g.nx - 2) * (g.ny - 2) * (g.nz - 2),
                                  W[s].maxv);
        init[s].diverge_at = div_at;
    }
    std::memcpy(hb.data() + off, init.data(), S * sizeof(MgStatus));
    MgStatus *dst = (MgStatus *)(dbase + off);
    HIP_OK(hipMemcpyAsync(dbase, hb.data(), total, hipMemcpyHostToDevice, w.stream));

    // ---- work lists: per launch kind and level, the solves that still run
    // and their block counts (a launch's grid is their sum, not the largest
    // solve's count times S).  kinds: 0 smoothing pass, 1 prolongation,
    // 2 residual + restriction (per-node residual without streaming), 3
    // per-node restriction; list WL_OUT: the FP64 outer step.
    constexpr int kKinds = 4;
    const int WL_OUT = kKinds * maxT;
    auto kind_blocks = [&](int kind, int l, int s) -> int64_t {
        const TaskWork &tw = W[s];
        if (kind < 0) return ceil_div(lv_nodes(tw.t->g), 256);
        if (l >= tw.tail_start) return 0;
        const MgLevelDev &d = tw.L[l].d;
        const GridDesc &g = tw.L[l].g;
        switch (kind) {
        case 0:
            return phase_pass ? smooth3_blocks(g.nx, g.ny, g.nz, d.zc, d.rg)
                              : (int64_t)smooth_tiles(g.nx, g.ny) * ceil_div(g.nz, d.zc);
        case 1:
            return stream ? mg_stream_blocks(g.nx, g.ny, g.nz, d.zc) : ceil_div(lv_nodes(g), 256);
        case 2: {
            const GridDesc &gc = tw.L[l + 1].g;
            return stream ? mg_rr_blocks(gc.nx, gc.ny, gc.nz, tw.L[l + 1].d.zc)
                          : ceil_div(lv_nodes(g), 256);
        }
        default:
            return stream ? 0 : ceil_div(lv_nodes(tw.L[l + 1].g), 256);
        }
    };
    struct WlRef {
        size_t off = 0;
        int64_t total = 0;
    };
    std::vector<WlRef> wref(WL_OUT + 1);
    std::vector<uint32_t> hwl;
    DBuf<uint32_t> dwl(w.pool, (size_t)(WL_OUT + 1) * (2 * (size_t)S + 2));
    std::vector<char> active(S, 1);
    auto build_lists = [&]() {
        hwl.clear();
        for (int li = 0; li <= WL_OUT; ++li) {
            const int kind = li == WL_OUT ? -1 : li / std::max(1, maxT);
            const int l = li == WL_OUT ? 0 : li % std::max(1, maxT);
            std::vector<uint32_t> ids, pre(1, 0);
            int64_t acc = 0;
            for (int s = 0; s < S; ++s) {
                // the outer step lists a lead while either side of its pair runs
                const bool on = kind >= 0 ? (bool)active[s]
                                          : lead_of[s] < 0 &&
                                                (active[s] || (partner[s] >= 0 && active[partner[s]]));
                const int64_t b = on ? kind_blocks(kind, l, s) : 0;
                if (b <= 0) continue;
                ids.push_back((uint32_t)s);
                acc += b;
                pre.push_back((uint32_t)acc);
            }
            if (acc >= ((int64_t)1 << 31)) throw DeviceError(MMPBSA_ERR_TOO_LARGE, "launch above 2^31 blocks");
            wref[li].off = hwl.size();
            wref[li].total = acc;
            hwl.push_back((uint32_t)ids.size());
            hwl.insert(hwl.end(), pre.begin(), pre.end());
            hwl.insert(hwl.end(), ids.begin(), ids.end());
        }
        HIP_OK(hipMemcpyAsync(dwl.get(), hwl.data(), sizeof(uint32_t) * hwl.size(),
                              hipMemcpyHostToDevice, w.stream));
    };
    build_lists();
    auto wl_of = [&](int li) { return dwl.get() + wref[li].off; };
    auto tot_of = [&](int li) { return wref[li].total; };

    // cycle 0: residual of the boundary-only potential (no test)
    mg_outer64(dO0, S, tot_of(WL_OUT), dst, w.stream, wl_of(WL_OUT));
    mg_check(dO0, S, 0, dst, w.stream);

    // profiled level-0 smoother passes: events, cycle number, pass reads e,
    // pass adds the prolongated coarse correction
    struct Ev {
        hipEvent_t a, b;
        int cycle;
        bool reads_e, prolong;
    };
    std::vector<Ev> evs;
    int cycle_counter = 0;
    // three sweeps per smoothing phase (pb_solver.rs:842-939), one pass each:
    // pre-smoothing ends in e2 and post-smoothing in e, alternating in between
    constexpr int npass = 3;
    // The first post-smoothing pass also applies prolongate_add from C.
    auto smooth_phase = [&](const MgLevelDev *F, const MgLevelDev *C, int l, bool pre) {
        if (phase_pass) {
            // pre: from zero into e2; post: e2 plus the prolongated coarse
            // correction (added by its own kernel) into e
            if (!pre) {
                const int li = 1 * maxT + l;
                if (stream)
                    mg_prolong_e2s(F, C, S, (int)tot_of(li), dst, w.stream, wl_of(li));
                else
                    mg_prolong_e2(F, C, S, tot_of(li), dst, w.stream, wl_of(li));
            }
            // the events bracket the smoother launch alone
            Ev ev{nullptr, nullptr, cycle_counter, !pre, false};
            if (profiling && l == 0) {
                HIP_OK(hipEventCreate(&ev.a));
                HIP_OK(hipEventCreate(&ev.b));
                HIP_OK(hipEventRecord(ev.a, w.stream));
            }
            mg_smooth3(F, nullptr, S, pre ? 1 : 0, pre ? 1 : 0, (int)tot_of(l), level_rg[l], dst,
                       w.stream, wl_of(l));
            if (profiling && l == 0) {
                HIP_OK(hipEventRecord(ev.b, w.stream));
                evs.push_back(ev);
            }
            return;
        }
        for (int k = 0; k < npass; ++k) {
            const bool first = pre && k == 0;
            const bool prolong = !pre && k == 0;
            const int dir = ((npass - 1 - k) & 1) ^ (pre ? 1 : 0);
            Ev ev{nullptr, nullptr, cycle_counter, !first, prolong};
            if (profiling && l == 0) {
                HIP_OK(hipEventCreate(&ev.a));
                HIP_OK(hipEventCreate(&ev.b));
                HIP_OK(hipEventRecord(ev.a, w.stream));
            }
            mg_smooth(F, prolong ? C : nullptr, S, first ? 1 : 0, dir, (int)tot_of(l), dst,
                      w.stream, wl_of(l));
            if (profiling && l == 0) {
                HIP_OK(hipEventRecord(ev.b, w.stream));
                evs.push_back(ev);
            }
        }
    };
    auto enqueue_cycle = [&](int p) {
        ++cycle_counter;
        for (int l = 0; l < maxT; ++l) {
            const MgLevelDev *F = D[p].F + (size_t)l * S;
            const MgLevelDev *C = D[p].C + (size_t)l * S;
            smooth_phase(F, C, l, true);
            const int li = 2 * maxT + l;
            if (stream) {
                mg_residual_restrict(F, C, S, (int)tot_of(li), dst, w.stream, wl_of(li));
            } else {
                mg_residual32(F, S, tot_of(li), dst, w.stream, wl_of(li));
                mg_restrict(F, C, S, tot_of(3 * maxT + l), dst, w.stream, wl_of(3 * maxT + l));
            }
        }
        mg_tail(D[p].T, S, dst, w.stream);
        for (int l = maxT - 1; l >= 0; --l) {
            const MgLevelDev *F = D[p].F + (size_t)l * S;
            const MgLevelDev *C = D[p].C + (size_t)l * S;
            smooth_phase(F, C, l, false);
        }
        mg_outer64(D[p].O, S, tot_of(WL_OUT), dst, w.stream, wl_of(WL_OUT));
        mg_check(D[p].O, S, 1, dst, w.stream);
    };

    int maxv = 0;
    for (auto &tw : W) maxv = std::max(maxv, tw.maxv);
    std::vector<MgStatus> hs(S);
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    auto drop_graph = [&]() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    };
    auto capture = [&]() {
        std::shared_lock<std::shared_mutex> gate(capture_gate());
        HIP_OK(hipStreamBeginCapture(w.stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_cycle(1);
            enqueue_cycle(0);
        } catch (...) {
            hipGraph_t g2 = nullptr;
            (void)hipStreamEndCapture(w.stream, &g2);
            if (g2) (void)hipGraphDestroy(g2);
            throw;
        }
        HIP_OK(hipStreamEndCapture(w.stream, &graph));
        HIP_OK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    };
    struct GraphGuard {
        std::function<void()> f;
        ~GraphGuard() { f(); }
    } graph_guard{drop_graph};
    if (use_graph) capture();
    // blocks of the FP64 outer step of the running solves when the lists were
    // last built (a proxy for the level-0 work of every launch), counted per
    // solve
    auto active_work = [&]() {
        int64_t a = 0;
        for (int s = 0; s < S; ++s)
            if (active[s]) a += kind_blocks(-1, 0, s);
        return a;
    };
    int64_t listed = active_work();
    const int chunk_cycles = std::max(2, tun.check_cycles);
    int done_cycles = 0;
    while (done_cycles < maxv) {
        const int n = std::min(chunk_cycles, maxv - done_cycles + (maxv - done_cycles) % 2);
        if (use_graph) {
            for (int k = 0; k < n; k += 2) HIP_OK(hipGraphLaunch(exec, w.stream));
        } else {
            for (int k = 0; k < n; ++k) enqueue_cycle((done_cycles + k + 1) & 1);
        }
        done_cycles += n;
        HIP_OK(hipMemcpyAsync(hs.data(), dst, sizeof(MgStatus) * S, hipMemcpyDeviceToHost,
                              w.stream));
        w.sync();
        bool all = true;
        for (int s = 0; s < S; ++s) all = all && hs[s].state != 0;
        if (all) break;
        // Once finished solves leave a quarter of the listed level-0 work
        // idle, rebuild the lists without them (and the graph with the
        // smaller grids).  The stream is idle here.  Profiled runs keep one
        // launch size throughout.
        if (!profiling) {
            int64_t running = 0;
            for (int s = 0; s < S; ++s)
                if (active[s] && hs[s].state == 0) running += kind_blocks(-1, 0, s);
            if (4 * running <= 3 * listed) {
                for (int s = 0; s < S; ++s) active[s] = active[s] && hs[s].state == 0;
                build_lists();
                listed = active_work();
                if (use_graph) {
                    drop_graph();
                    capture();
                }
            }
        }
    }
    drop_graph();
    if (profiling && !evs.empty()) {
        // algorithmic bytes of one pass over a level-0 solve, per node: the
        // correction read (unless starting from zero) and written (the
        // post-smoothing pass that adds it to phi reads and writes the FP64
        // potential instead), f, the code byte (non-uniform maps) and, when
        // prolongating, the coarse correction (one eighth of a float per fine
        // node)
        double ms_tot = 0.0, bytes = 0.0;
        for (const Ev &ev : evs) {
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, ev.a, ev.b));
            ms_tot += ms;
            for (int s = 0; s < S; ++s) {
                if (W[s].tail_start == 0 || ev.cycle > hs[s].cycles) continue;
                const double n = (double)lv_nodes(W[s].t->g);
                const double out = (ev.reads_e && W[s].fold) ? 16.0 : 4.0;
                bytes += n * ((ev.reads_e ? 4.0 : 0.0) + out + 4.0 + (W[s].t->m->uni ? 0.0 : 1.0) +
                              (ev.prolong ? 0.5 : 0.0));
            }
            (void)hipEventDestroy(ev.a);
            (void)hipEventDestroy(ev.b);
        }
        std::lock_guard<std::mutex> lk(profile_mutex());
        prof.total_ms += ms_tot;
        prof.launches += (int64_t)evs.size();
        prof.alg_bytes += bytes;
        prof.grid_threads = tot_of(0) * (phase_pass ? smooth3_threads(level_rg[0]) : 256);
    }
    for (int s = 0; s < S; ++s) {
        TaskWork &tw = W[s];
        MgTask *t = tw.t;
        const int fin = hs[s].cycles;
        const int64_t N0 = lv_nodes(t->g);
        if (!tw.fold && (fin & 1))
            HIP_OK(hipMemcpyAsync(t->phi, tw.buf1.get(), sizeof(double) * N0,
                                  hipMemcpyDeviceToDevice, w.stream));
        if (hs[s].state == 2) {
            // divergence (pb_solver.rs:1110-1138): Chebyshev SOR from the partial
            // potential with the boundary condition re-applied
            hipLaunchKernelGGL(copy_boundary_kernel, dim3(ceil_div(N0, 256)), dim3(256), 0,
                               w.stream, t->g, tw.bc.get(), t->phi);
            LAUNCH_OK();
            t->info = sor_solve(w, t->g, *t->m, t->rho, t->phi, t->tol, t->max_it);
            continue;
        }
        t->info.iterations = hs[s].state == 1 ? fin : tw.maxv;
        t->info.converged = hs[s].state == 1;
        t->info.final_residual = hs[s].rms;
    }
    w.sync();
}

// ---- PB jobs -------------------------------------------------------------------
struct PbJob::State {
    GridDesc fine{}, cg{};
    bool focus = false;
    double kappa = 0.0;
    double mn[3], mx[3];
    DevCells cells;
    DBuf<double> rho_f, rho_c;
    struct Side {
        double ein = 1.0, eout = 1.0;
        bool ionic = false;
        MapSet cm, fm;
        DBuf<double> cphi, phi;
        MgTask ct, ft;
        bool have_bc = false;
        mmpbsa_pb_solve_info *info = nullptr;
    } side[2];
};

// Debye-Hueckel boundary values of grid g for the job's solvated medium (side
// 0, with salt) and / or its reference medium (side 1, salt-free, eps_in):
// one pass over the atoms when both are wanted.
static void dh_sides(Worker &w, PbJob *j, const GridDesc &g, bool want0, bool want1,
                     double *phi0, double *phi1)
{
    const PbJob::State &S = *j->st;
    const auto &s0 = S.side[0], &s1 = S.side[1];
    const int n = j->t->h.n;
    DBuf<double> dhw(w.pool, dh_work_size(g, n));
    // 15 FP32 FLOP per boundary node and atom (one pass serves both media
    // when the reference medium is salt-free)
    const double bnodes = (double)g.size() - (double)std::max(0, g.nx - 2) *
                                                 std::max(0, g.ny - 2) * std::max(0, g.nz - 2);
    KernelTimer kt(w, "dh_boundary_kernel", 15.0 * bnodes * n);
    const double k0 = s0.ionic ? S.kappa : 0.0, k1 = s1.ionic ? S.kappa : 0.0;
    if (want0 && want1 && k1 == 0.0) {
        launch_dh_boundary(g, n, j->d_xyz, j->t->q.p, k0, s0.eout, phi0, s1.eout, phi1, dhw.get(),
                           w.stream);
        return;
    }
    if (want0)
        launch_dh_boundary(g, n, j->d_xyz, j->t->q.p, k0, s0.eout, phi0, 0.0, nullptr, dhw.get(),
                           w.stream);
    if (want1)
        launch_dh_boundary(g, n, j->d_xyz, j->t->q.p, k1, s1.eout, phi1, 0.0, nullptr, dhw.get(),
                           w.stream);
}

void run_pb_jobs(Worker &w, std::vector<PbJob> &jobs, bool batched)
{
    std::vector<PbJob *> act;
    for (PbJob &j : jobs) {
        std::memset(j.out, 0, sizeof(*j.out));
        if (!(j.p.grid_spacing > 0.0)) throw DeviceError(MMPBSA_ERR_INVALID_ARG, "grid_spacing <= 0");
        if (j.t->h.n == 0) continue;
        act.push_back(&j);
    }
    // bounding boxes of jobs without a host copy of their coordinates: one
    // reduction per job, one readback for all
    std::vector<double> dev_box;
    {
        std::vector<PbJob *> need;
        for (PbJob *j : act)
            if (!j->h_xyz) need.push_back(j);
        if (!need.empty()) {
            DBuf<double> d(w.pool, 6 * need.size());
            for (size_t k = 0; k < need.size(); ++k)
                launch_bbox(need[k]->d_xyz, need[k]->t->h.n, d.get() + 6 * k, w.stream);
            dev_box.resize(6 * need.size());
            HIP_OK(hipMemcpyAsync(dev_box.data(), d.get(), sizeof(double) * dev_box.size(),
                                  hipMemcpyDeviceToHost, w.stream));
            w.sync();
        }
    }
    size_t dev_box_k = 0;
    // geometry, cell lists, charge maps (pb_energy.rs:332-372)
    for (PbJob *j : act) {
        j->st = std::make_shared<PbJob::State>();
        PbJob::State &S = *j->st;
        const mmpbsa_pb_params &p = j->p;
        const int n = j->t->h.n;
        S.kappa = compute_kappa(p.salt_concentration, p.solvent_dielectric, p.temperature);
        if (j->h_xyz) {
            bbox(n, j->h_xyz, S.mn, S.mx);
        } else {
            for (int k = 0; k < 3; ++k) {
                S.mn[k] = dev_box[6 * dev_box_k + k];
                S.mx[k] = dev_box[6 * dev_box_k + 3 + k];
            }
            ++dev_box_k;
        }
        S.fine = auto_grid(S.mn, S.mx, p.grid_spacing, p.grid_buffer);
        if (p.fine_dims[0] > 0) {
            S.fine.nx = p.fine_dims[0];
            S.fine.ny = p.fine_dims[1] > 0 ? p.fine_dims[1] : p.fine_dims[0];
            S.fine.nz = p.fine_dims[2] > 0 ? p.fine_dims[2] : p.fine_dims[0];
            const int d[3] = {S.fine.nx, S.fine.ny, S.fine.nz};
            double o[3];
            for (int k = 0; k < 3; ++k)
                o[k] = 0.5 * (S.mn[k] + S.mx[k]) - 0.5 * ((double)(d[k] - 1) * p.grid_spacing);
            S.fine.ox = o[0];
            S.fine.oy = o[1];
            S.fine.oz = o[2];
        }
        // explicit frames (fine_center, coarse_dims): origin = centre - (n-1) h / 2
        if (p.fine_centered) {
            S.fine.ox = p.fine_center[0] - 0.5 * ((double)(S.fine.nx - 1) * p.grid_spacing);
            S.fine.oy = p.fine_center[1] - 0.5 * ((double)(S.fine.ny - 1) * p.grid_spacing);
            S.fine.oz = p.fine_center[2] - 0.5 * ((double)(S.fine.nz - 1) * p.grid_spacing);
        }
        const bool explicit_cg = p.coarse_dims[0] > 0;
        S.focus = explicit_cg || (p.fillratio > 0.0 && p.fscale > 1);
        if (explicit_cg) {
            if (p.coarse_dims[1] <= 0 || p.coarse_dims[2] <= 0 || !(p.coarse_spacing > 0.0))
                throw DeviceError(MMPBSA_ERR_INVALID_ARG,
                                  "coarse_dims and coarse_spacing must be positive");
            S.cg.nx = p.coarse_dims[0];
            S.cg.ny = p.coarse_dims[1];
            S.cg.nz = p.coarse_dims[2];
            S.cg.h = p.coarse_spacing;
            S.cg.ox = p.coarse_center[0] - 0.5 * ((double)(S.cg.nx - 1) * p.coarse_spacing);
            S.cg.oy = p.coarse_center[1] - 0.5 * ((double)(S.cg.ny - 1) * p.coarse_spacing);
            S.cg.oz = p.coarse_center[2] - 0.5 * ((double)(S.cg.nz - 1) * p.coarse_spacing);
        } else if (S.focus) {
            S.cg = coarse_grid(S.mn, S.mx, p.grid_spacing, p.fillratio, p.fscale);
        }
        if (S.fine.size() > (int64_t)1 << 31 || (S.focus && S.cg.size() > (int64_t)1 << 31))
            throw DeviceError(MMPBSA_ERR_TOO_LARGE, "PB grid above 2^31 nodes");
        j->out->fine_dims[0] = S.fine.nx;
        j->out->fine_dims[1] = S.fine.ny;
        j->out->fine_dims[2] = S.fine.nz;
        if (S.focus) {
            j->out->coarse_dims[0] = S.cg.nx;
            j->out->coarse_dims[1] = S.cg.ny;
            j->out->coarse_dims[2] = S.cg.nz;
        }
        build_cells(w, j->d_xyz, n, S.mn, S.mx,
                    j->t->max_radius + std::max(p.probe_radius, std::max(p.ion_radius, 0.0)),
                    S.cells);
        {
            DBuf<unsigned long long> acc(w.pool, S.fine.size());
            S.rho_f = DBuf<double>(w.pool, S.fine.size());
            launch_spread(S.fine, n, j->d_xyz, j->t->q.p, acc.get(), S.rho_f.get(), w.stream);
        }
        if (S.focus) {
            DBuf<unsigned long long> acc(w.pool, S.cg.size());
            S.rho_c = DBuf<double>(w.pool, S.cg.size());
            launch_spread(S.cg, n, j->d_xyz, j->t->q.p, acc.get(), S.rho_c.get(), w.stream);
        }
        // pb_energy.rs:375-406: solvated (eps_in, eps_out, ions) and reference (eps_in)
        S.side[0].ein = p.solute_dielectric;
        S.side[0].eout = p.solvent_dielectric;
        S.side[0].ionic = true;
        S.side[0].info = &j->out->solv;
        S.side[1].ein = p.solute_dielectric;
        S.side[1].eout = p.solute_dielectric;
        S.side[1].ionic = false;
        S.side[1].info = &j->out->ref;
    }
    // The dielectric and kappa maps depend only on the coordinates, the cell
    // lists and the grids.  With batched solves (and no profiling) they are
    // built on an auxiliary worker in a second host thread: the coarse maps
    // while the main thread computes the coarse Debye-Hueckel boundaries, then
    // the fine maps while it runs the coarse stage, so the flood fills' short
    // launches and host checks overlap that work.
    std::vector<MapReq> rq_coarse, rq_fine;
    for (PbJob *j : act) {
        PbJob::State &S = *j->st;
        for (auto &sd : S.side) {
            if (S.focus)
                rq_coarse.push_back(MapReq{j->t, j->d_xyz, &S.cells, S.cg, &j->p, sd.ein, sd.eout,
                                           sd.ionic, S.kappa, &sd.cm});
            rq_fine.push_back(MapReq{j->t, j->d_xyz, &S.cells, S.fine, &j->p, sd.ein, sd.eout,
                                     sd.ionic, S.kappa, &sd.fm});
        }
    }
    // (the thread's promise and error slot are declared before the thread, so
    // that on unwinding its join runs before they are destroyed)
    std::promise<void> coarse_ready;
    std::future<void> coarse_maps = coarse_ready.get_future();
    std::exception_ptr maps_err;
    struct Joiner {
        std::thread t;
        ~Joiner()
        {
            if (t.joinable()) t.join();
        }
    } maps_thread;
    const bool maps_beside = batched && !profile_stats().enabled && !rq_fine.empty();
    if (maps_beside) {
        w.sync();  // cell lists and coordinates complete before the other stream reads them
        Worker &mw = w.aux(1);
        maps_thread.t = std::thread([&rq_coarse, &rq_fine, &mw, &coarse_ready, &maps_err] {
            try {
                build_maps_batch(mw, rq_coarse);
                mw.sync();
            } catch (...) {
                coarse_ready.set_exception(std::current_exception());
                return;
            }
            coarse_ready.set_value();
            try {
                build_maps_batch(mw, rq_fine);
                mw.sync();
            } catch (...) {
                maps_err = std::current_exception();
            }
        });
    } else {
        build_maps_batch(w, rq_coarse);
    }
    // coarse focusing solves with Debye-Hueckel boundaries (pb_energy.rs:150-186)
    std::vector<MgTask *> tasks;
    for (PbJob *j : act) {
        PbJob::State &S = *j->st;
        if (!S.focus) continue;
        for (auto &sd : S.side) {
            sd.cphi = DBuf<double>(w.pool, S.cg.size());
            launch_fill_f64(sd.cphi.get(), 0.0, S.cg.size(), w.stream);
        }
        dh_sides(w, j, S.cg, true, true, S.side[0].cphi.get(), S.side[1].cphi.get());
        for (auto &sd : S.side) {
            sd.ct.g = S.cg;
            sd.ct.m = &sd.cm;
            sd.ct.rho = S.rho_c.get();
            sd.ct.phi = sd.cphi.get();
            sd.ct.tol = j->p.tolerance * 10.0;
            sd.ct.max_it = j->p.max_iterations;
            tasks.push_back(&sd.ct);
        }
    }
    if (maps_beside) coarse_maps.get();  // rethrows a failure of the coarse maps
    if (!tasks.empty()) mg_solve_batch(w, tasks, batched, false);
    // coarse validity (pb_energy.rs:188-194), then fine boundary values
    {
        std::vector<DBuf<double>> parts;
        DBuf<double> res(w.pool, 2 * act.size() + 1);
        int k = 0;
        for (PbJob *j : act) {
            PbJob::State &S = *j->st;
            for (auto &sd : S.side) {
                if (S.focus && sd.ct.info.converged) {
                    parts.emplace_back(w.pool, 1024);
                    int nbk = 0;
                    launch_maxabs(sd.cphi.get(), S.cg.size(), parts.back().get(), &nbk, w.stream);
                    launch_max_partials(parts.back().get(), nbk, res.get() + k, w.stream);
                }
                ++k;
            }
        }
        std::vector<double> mx(2 * act.size() + 1, 0.0);
        HIP_OK(hipMemcpyAsync(mx.data(), res.get(), sizeof(double) * mx.size(),
                              hipMemcpyDeviceToHost, w.stream));
        w.sync();
        k = 0;
        for (PbJob *j : act) {
            PbJob::State &S = *j->st;
            bool need[2] = {false, false};
            for (int si = 0; si < 2; ++si) {
                auto &sd = S.side[si];
                sd.info->focused = 0;
                sd.info->coarse_iterations = S.focus ? sd.ct.info.iterations : 0;
                const bool ok = S.focus && sd.ct.info.converged && std::isfinite(mx[k]) &&
                                mx[k] < 1e10;
                sd.phi = DBuf<double>(w.pool, S.fine.size());
                launch_fill_f64(sd.phi.get(), 0.0, S.fine.size(), w.stream);
                if (ok) {
                    launch_interp_boundary(S.fine, S.cg, sd.cphi.get(), sd.phi.get(), w.stream);
                    sd.info->focused = 1;
                } else {
                    need[si] = true;
                }
                // (the coarse maps are released once the map thread is done:
                // they live in its worker's pool)
                sd.cphi.release();
                ++k;
            }
            dh_sides(w, j, S.fine, need[0], need[1], S.side[0].phi.get(), S.side[1].phi.get());
        }
    }
    // fine solves
    tasks.clear();
    if (maps_beside) {
        maps_thread.t.join();
        if (maps_err) std::rethrow_exception(maps_err);
    } else {
        build_maps_batch(w, rq_fine);
    }
    for (PbJob *j : act)
        for (auto &sd : j->st->side) sd.cm = MapSet();
    for (PbJob *j : act) {
        PbJob::State &S = *j->st;
        for (auto &sd : S.side) {
            sd.ft.g = S.fine;
            sd.ft.m = &sd.fm;
            sd.ft.rho = S.rho_f.get();
            sd.ft.phi = sd.phi.get();
            sd.ft.tol = j->p.tolerance;
            sd.ft.max_it = j->p.max_iterations;
            tasks.push_back(&sd.ft);
        }
    }
    if (!tasks.empty()) mg_solve_batch(w, tasks, batched, true);
    // energies, compute_elec_energy (pb_solver.rs:1174-1186)
    {
        const size_t nres = 2 * act.size() + 1;
        DBuf<double> res(w.pool, nres);
        std::vector<DBuf<double>> parts;
        int k = 0;
        for (PbJob *j : act) {
            PbJob::State &S = *j->st;
            if (j->d_atom_terms)
                launch_pb_atom_terms(S.fine, S.side[0].phi.get(), S.side[1].phi.get(), j->t->h.n,
                                     j->d_xyz, j->t->q.p, j->d_atom_terms, w.stream);
            for (auto &sd : S.side) {
                sd.fm = MapSet();
                parts.emplace_back(w.pool, std::max(1, ceil_div(j->t->h.n, 256)));
                int nbk = 0;
                launch_energy_gather(S.fine, sd.phi.get(), j->t->h.n, j->d_xyz, j->t->q.p,
                                     parts.back().get(), &nbk, w.stream);
                launch_sum_partials(parts.back().get(), nbk, 1, 1, res.get() + k, w.stream);
                ++k;
            }
        }
        std::vector<double> e(nres, 0.0);
        HIP_OK(hipMemcpyAsync(e.data(), res.get(), sizeof(double) * nres, hipMemcpyDeviceToHost,
                              w.stream));
        w.sync();
        k = 0;
        for (PbJob *j : act) {
            PbJob::State &S = *j->st;
            for (auto &sd : S.side) {
                sd.info->energy = 0.5 * e[k];
                sd.info->iterations = sd.ft.info.iterations;
                sd.info->converged = sd.ft.info.converged;
                sd.info->final_residual = sd.ft.info.final_residual;
                ++k;
            }
            j->out->total = j->out->solv.energy - j->out->ref.energy;
            j->st.reset();
        }
    }
}

// compute_pb_energy (pb_energy.rs:332-438) for one subsystem, unbatched
void run_pb(Worker &w, const DevTopo &t, const double *d_xyz, const double *h_xyz,
            const mmpbsa_pb_params &p, mmpbsa_pb_result *out)
{
    std::vector<PbJob> jobs(1);
    jobs[0].t = &t;
    jobs[0].d_xyz = d_xyz;
    jobs[0].h_xyz = h_xyz;
    jobs[0].p = p;
    jobs[0].out = out;
    run_pb_jobs(w, jobs, false);
}

}  // namespace mmpbsa
