// capi.cpp -- the extern "C" boundary (include/mmpbsa_hip.h): contexts,
// receptor/ligand splits, frame energies, batched frames, decomposition.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <map>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <set>
#include <thread>
#include <unordered_map>

#include "decomp.h"
#include "engine.h"
#include "trajectory.h"

using namespace mmpbsa;

struct mmpbsa_ctx {
    DevTopo top;
    std::unique_ptr<Worker> w;
    std::string err;
    int device = 0;
    // the solute of the last mmpbsa_pb_energy_solvated call, kept resident
    std::vector<int> solute_res, solute_sel;
    std::unique_ptr<DevTopo> solute;
};

// GbCache (gb_energy.rs:66-120): the coordinates and Born radii of the last
// frame whose radii were computed, on the host as in the reference.
struct mmpbsa_gb_cache {
    double threshold = 0.5;
    std::vector<double> xyz, radii;
};

struct mmpbsa_split {
    mmpbsa_ctx *ctx = nullptr;
    bool identity = true;
    std::vector<int> csel, ra, la;          // complex atoms; rec / lig atoms in the complex
    std::vector<int> rec_res, lig_res;      // residues as requested (caller numbering)
    std::vector<int> rec_loc, lig_loc;      // the same residues in the complex numbering
    DevTopo ctop, rtop, ltop;
    DArr<int> d_csel, d_ra, d_la;
    // decomposition: complex residue -> positions in [rec atoms..., lig atoms...]
    DArr<int> d_seg_pos, d_seg_ptr;
    int cross_excl = 0;
    // the complex is exactly the two partners: its GB radii come from the
    // partners' psi sums plus the cross sums; when also no excluded, 1-4 or
    // bonded term spans them, its MM terms from the partners' and the cross pairs
    bool gb_parts = false, mm_parts = false;
    std::vector<std::unique_ptr<Worker>> workers;
    // results of the last trajectory call
    std::vector<mmpbsa_frame_terms> traj_out;
    std::vector<double> traj_last;
};

namespace {

thread_local std::string g_err;

int fail(mmpbsa_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    g_err = msg;
    return code;
}

template <typename F>
int guarded(mmpbsa_ctx *ctx, F &&f)
{
    try {
        f();
        return MMPBSA_OK;
    } catch (const DeviceError &e) {
        return fail(ctx, e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(ctx, MMPBSA_ERR_OOM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(ctx, MMPBSA_ERR_DEVICE, e.what());
    }
}

bool g_device_ok = false;

void require_device()
{
    if (g_device_ok) return;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        throw DeviceError(MMPBSA_ERR_NO_DEVICE, "no HIP device visible");
    }
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, 0));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        throw DeviceError(MMPBSA_ERR_NO_DEVICE,
                          std::string("device is ") + prop.gcnArchName + ", library built for gfx950");
    g_device_ok = true;
}

// atoms of the split's complex (an identity split keeps no selection list)
size_t complex_atoms(const mmpbsa_split *s)
{
    return s->identity ? (size_t)s->ctx->top.h.n : s->csel.size();
}

std::vector<double> gather_host(const double *xyz, const std::vector<int> &idx)
{
    std::vector<double> out(idx.size() * 3);
    for (size_t k = 0; k < idx.size(); ++k)
        for (int c = 0; c < 3; ++c) out[3 * k + c] = xyz[3 * (size_t)idx[k] + c];
    return out;
}

struct FrameCoords {
    DBuf<double> full, c, r, l;
    const double *dc = nullptr;
    std::vector<double> hc, hr, hl;
    const double *hcp = nullptr;
};

// Slice the complex, receptor and ligand coordinates of one frame
// (binding.rs:213-222, :345-346) on the device.  Host frames are uploaded
// first; device frames (on_device) are read in place and leave no host copy
// (bounding boxes are then taken on the device).
void stage_frame(Worker &w, mmpbsa_split *s, const double *xyz, FrameCoords &f, bool on_device,
                 bool complex_local = false)
{
    if (complex_local) {  // host coordinates of the complex atoms only
        const size_t nc = complex_atoms(s);
        f.c = DBuf<double>(w.pool, 3 * nc + 3);
        HIP_OK(hipMemcpyAsync(f.c.get(), xyz, sizeof(double) * 3 * nc, hipMemcpyHostToDevice,
                              w.stream));
        f.dc = f.c.get();
        f.hcp = xyz;
        f.r = DBuf<double>(w.pool, 3 * s->ra.size() + 3);
        f.l = DBuf<double>(w.pool, 3 * s->la.size() + 3);
        launch_gather_xyz(f.dc, s->d_ra.p, (int)s->ra.size(), f.r.get(), w.stream);
        launch_gather_xyz(f.dc, s->d_la.p, (int)s->la.size(), f.l.get(), w.stream);
        f.hr = gather_host(f.hcp, s->ra);
        f.hl = gather_host(f.hcp, s->la);
        return;
    }
    const int n_all = s->ctx->top.h.n;
    const double *full = xyz;
    if (!on_device) {
        f.full = DBuf<double>(w.pool, 3 * (size_t)n_all);
        HIP_OK(hipMemcpyAsync(f.full.get(), xyz, sizeof(double) * 3 * n_all,
                              hipMemcpyHostToDevice, w.stream));
        full = f.full.get();
    }
    if (s->identity) {
        f.dc = full;
        f.hcp = on_device ? nullptr : xyz;
    } else {
        f.c = DBuf<double>(w.pool, 3 * s->csel.size());
        launch_gather_xyz(full, s->d_csel.p, (int)s->csel.size(), f.c.get(), w.stream);
        f.dc = f.c.get();
        if (!on_device) {
            f.hc = gather_host(xyz, s->csel);
            f.hcp = f.hc.data();
        }
    }
    f.r = DBuf<double>(w.pool, 3 * s->ra.size() + 3);
    f.l = DBuf<double>(w.pool, 3 * s->la.size() + 3);
    launch_gather_xyz(f.dc, s->d_ra.p, (int)s->ra.size(), f.r.get(), w.stream);
    launch_gather_xyz(f.dc, s->d_la.p, (int)s->la.size(), f.l.get(), w.stream);
    if (!on_device) {
        f.hr = gather_host(f.hcp, s->ra);
        f.hl = gather_host(f.hcp, s->la);
    }
}

double mm_total(const mmpbsa_mm_terms &m)
{
    return m.bond + m.angle + m.dihedral + m.vdw + m.elec + m.vdw_14 + m.elec_14;
}

// One frame of one split with its method and parameters.
struct FrameItem {
    mmpbsa_split *s = nullptr;
    const double *xyz = nullptr;  // all atoms of the caller's topology (host or device)
    int method = MMPBSA_METHOD_GB;
    mmpbsa_gb_params gb{};
    mmpbsa_pb_params pb{};
    mmpbsa_sa_params sa{};
    mmpbsa_frame_terms *out = nullptr;
    bool complex_local = false;   // xyz holds only the split's complex atoms
};

// compute_frame_energy (binding.rs:333-415) for a list of frames, possibly of
// different splits, methods and parameters.  With batched, the PB solves of
// all frames and subsystems advance together (shared launches, graph-replayed
// cycles); otherwise each solve runs on its own.  Every frame's result depends
// only on its own inputs, so it is the same alone and in any batch.
void run_items(Worker &w, const std::vector<FrameItem> &items, bool batched, bool on_device)
{
    const size_t nf = items.size();
    std::vector<FrameCoords> fc(nf);
    for (size_t f = 0; f < nf; ++f)
        stage_frame(w, items[f].s, items[f].xyz, fc[f], on_device, items[f].complex_local);
    auto tops = [&](size_t f, int k) -> const DevTopo * {
        mmpbsa_split *s = items[f].s;
        return k == 0 ? &s->ctop : (k == 1 ? &s->rtop : &s->ltop);
    };
    auto dptr = [&](size_t f, int k) -> const double * {
        return k == 0 ? fc[f].dc : (k == 1 ? fc[f].r.get() : fc[f].l.get());
    };
    auto hptr = [&](size_t f, int k) -> const double * {
        if (on_device) return nullptr;
        return k == 0 ? fc[f].hcp : (k == 1 ? fc[f].hr.data() : fc[f].hl.data());
    };
    std::vector<mmpbsa_pb_result> pr(3 * nf);
    std::vector<PbJob> jobs;
    for (size_t f = 0; f < nf; ++f) {
        if (items[f].method != MMPBSA_METHOD_PB) continue;
        for (int k = 0; k < 3; ++k) {
            PbJob j;
            j.t = tops(f, k);
            j.d_xyz = dptr(f, k);
            j.h_xyz = hptr(f, k);
            j.p = items[f].pb;
            j.out = &pr[3 * f + k];
            jobs.push_back(j);
        }
    }
    // MM, GB and SA of every frame: their own kernels on their own inputs,
    // independent of the PB solves.
    std::vector<double> mmv(3 * nf), gbv(3 * nf, 0.0), sav(3 * nf);
    auto other_terms = [&](Worker &x) {
        for (size_t f = 0; f < nf; ++f) {
            const FrameItem &it = items[f];
            const mmpbsa_split *sp = it.s;
            mmpbsa_mm_terms mm3[3];
            if (sp->mm_parts)
                run_mm_parts(x, sp->ctop, sp->rtop, sp->ltop, dptr(f, 0), dptr(f, 1), dptr(f, 2),
                             sp->d_ra.p, (int)sp->ra.size(), sp->d_la.p, (int)sp->la.size(), mm3);
            const bool gb = it.method != MMPBSA_METHOD_PB;
            if (gb && sp->gb_parts)
                run_gb_parts(x, sp->ctop, sp->rtop, sp->ltop, dptr(f, 0), dptr(f, 1), dptr(f, 2),
                             sp->d_ra.p, (int)sp->ra.size(), sp->d_la.p, (int)sp->la.size(),
                             it.gb, &gbv[3 * f]);
            for (int k = 0; k < 3; ++k) {
                if (!sp->mm_parts) run_mm(x, *tops(f, k), dptr(f, k), &mm3[k]);
                mmv[3 * f + k] = mm_total(mm3[k]);
                if (gb && !sp->gb_parts)
                    gbv[3 * f + k] = run_gb(x, *tops(f, k), dptr(f, k), it.gb, nullptr, nullptr);
                double tsasa = 0.0, tsa = 0.0;
                run_sa(x, *tops(f, k), dptr(f, k), hptr(f, k), it.sa, nullptr, nullptr, &tsasa,
                       &tsa);
                sav[3 * f + k] = tsa;
            }
        }
    };
    if (!jobs.empty() && batched && !profile_stats().enabled) {
        // Beside the batched solver: on an auxiliary worker (its own stream)
        // in a second host thread, so these kernels fill the GPU while the
        // solver's narrow launches (the one-workgroup tail, small levels) and
        // host checks leave it idle.  The staged coordinates are complete
        // before the other stream reads them; profiled runs keep them apart so
        // that kernel timers see each kernel alone.
        w.sync();
        Worker &side = w.aux(0);
        std::exception_ptr err;
        std::thread th([&] {
            try {
                other_terms(side);
            } catch (...) {
                err = std::current_exception();
            }
        });
        try {
            run_pb_jobs(w, jobs, true);
        } catch (...) {
            th.join();
            throw;
        }
        th.join();
        if (err) std::rethrow_exception(err);
    } else {
        if (!jobs.empty()) {
            if (batched) {
                run_pb_jobs(w, jobs, true);
            } else {
                for (auto &j : jobs) {
                    std::vector<PbJob> one(1, j);
                    run_pb_jobs(w, one, false);
                }
            }
        }
        other_terms(w);
    }
    for (size_t f = 0; f < nf; ++f) {
        const FrameItem &it = items[f];
        mmpbsa_frame_terms *o = it.out;
        mmpbsa_subsystem_energy *e[3] = {&o->complex_, &o->receptor, &o->ligand};
        for (int k = 0; k < 3; ++k) {
            e[k]->mm = mmv[3 * f + k];
            e[k]->polar = it.method == MMPBSA_METHOD_PB ? pr[3 * f + k].total : gbv[3 * f + k];
            e[k]->sa = sav[3 * f + k];
        }
        o->delta_mm = o->complex_.mm - o->receptor.mm - o->ligand.mm;
        o->delta_polar = o->complex_.polar - o->receptor.polar - o->ligand.polar;
        o->delta_sa = o->complex_.sa - o->receptor.sa - o->ligand.sa;
        o->delta_total = o->delta_mm + o->delta_polar + o->delta_sa;
    }
}

// nf frames of one split at xyz + stride * f
void frames_group(Worker &w, mmpbsa_split *s, const double *xyz, int nf, size_t stride,
                  int method, const mmpbsa_gb_params &gb, const mmpbsa_pb_params &pb,
                  const mmpbsa_sa_params &sa, mmpbsa_frame_terms *out, bool batched,
                  bool on_device = false)
{
    std::vector<FrameItem> items(nf);
    for (int f = 0; f < nf; ++f) {
        FrameItem &it = items[f];
        it.s = s;
        it.xyz = xyz + stride * f;
        it.method = method;
        it.gb = gb;
        it.pb = pb;
        it.sa = sa;
        it.out = out + f;
    }
    run_items(w, items, batched, on_device);
}

void frame_energy(Worker &w, mmpbsa_split *s, const double *xyz, int method,
                  const mmpbsa_gb_params &gb, const mmpbsa_pb_params &pb,
                  const mmpbsa_sa_params &sa, mmpbsa_frame_terms *out)
{
    frames_group(w, s, xyz, 1, 0, method, gb, pb, sa, out, false);
}

// ---- serving ------------------------------------------------------------------
using Clock = std::chrono::steady_clock;

struct ServerRequest {
    uint64_t ticket = 0;
    mmpbsa_split *split = nullptr;
    std::vector<double> xyz;          // the caller's frames, copied at submission
    int n_frames = 0;
    int method = MMPBSA_METHOD_GB;
    mmpbsa_gb_params gb{};
    mmpbsa_pb_params pb{};
    mmpbsa_sa_params sa{};
    std::vector<mmpbsa_frame_terms> out;
    int next = 0;                      // first frame not yet scheduled
    int done = 0;                      // frames finished
    double frame_bytes = 0.0;          // admission estimate of one frame in flight
    Clock::time_point t_submit, t_start, t_end;
    bool started = false;
    int status = MMPBSA_OK;
    std::string err;
    mmpbsa_request_summary summary{};  // device epilogue, set before the last frame counts
};

// Device bytes one frame holds while its PB solves run: per grid node of each
// subsystem's fine and coarse grids, the FP64 maps, charge density, two
// potentials per medium, the FP32 multigrid arrays of both solves and the
// geometry scratch (an upper estimate used for admission only).
constexpr double kPbBytesPerNode = 320.0;
// Fractional-face entry tables of the solvated solve (the reference solve's
// map is uniform and has none): 32 bytes per node of every level the phase
// smoother sweeps, at most 8/7 of the finest level's nodes.
constexpr double kEntryBytesPerNode = 32.0 * 8.0 / 7.0;

double frame_bytes_estimate(const mmpbsa_split *s, const double *xyz, int method,
                            const mmpbsa_pb_params &p)
{
    const size_t n = complex_atoms(s);
    if (method != MMPBSA_METHOD_PB || n == 0) return 600.0 * (double)n * 3.0 + 4096.0;
    std::vector<double> hc = s->identity ? std::vector<double>(xyz, xyz + 3 * n)
                                         : gather_host(xyz, s->csel);
    double mn[3], mx[3];
    bbox((int)n, hc.data(), mn, mx);
    GridDesc f = auto_grid(mn, mx, p.grid_spacing, p.grid_buffer);
    double nodes = (double)f.size();
    if (p.fine_dims[0] > 0) {
        const double a = p.fine_dims[0], b = p.fine_dims[1] > 0 ? p.fine_dims[1] : a,
                     c = p.fine_dims[2] > 0 ? p.fine_dims[2] : a;
        nodes = a * b * c;
    }
    if (p.coarse_dims[0] > 0)
        nodes += (double)p.coarse_dims[0] * std::max(p.coarse_dims[1], 1) *
                 std::max(p.coarse_dims[2], 1);
    else if (p.fillratio > 0.0 && p.fscale > 1)
        nodes += (double)coarse_grid(mn, mx, p.grid_spacing, p.fillratio, p.fscale).size();
    // the receptor and ligand grids are no larger than the complex's
    return 3.0 * (kPbBytesPerNode + kEntryBytesPerNode) * nodes;
}

}  // namespace

struct mmpbsa_server {
    int device = 0;
    double budget = 0.0;
    int max_round = 16;
    std::unique_ptr<Worker> w;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerRequest *> pending;   // requests with frames not yet scheduled
    std::unordered_map<uint64_t, std::unique_ptr<ServerRequest>> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    double in_flight = 0.0;
    std::thread th;
    void loop();
};

// The device epilogue of a finished request (frame_stats_kernel).
void summarize(Worker &w, ServerRequest &r)
{
    const int n = r.n_frames;
    std::vector<double> h(4 * (size_t)n);
    for (int f = 0; f < n; ++f) {
        h[4 * f] = r.out[f].delta_mm;
        h[4 * f + 1] = r.out[f].delta_polar;
        h[4 * f + 2] = r.out[f].delta_sa;
        h[4 * f + 3] = r.out[f].delta_total;
    }
    const double temp = r.method == MMPBSA_METHOD_PB ? r.pb.temperature : r.gb.temperature;
    DBuf<double> d(w.pool, h.size() + 10);
    HIP_OK(hipMemcpyAsync(d.get(), h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                          w.stream));
    launch_frame_stats(d.get(), n, 0.00198688 * temp, d.get() + h.size(), w.stream);
    double o[10];
    HIP_OK(hipMemcpyAsync(o, d.get() + h.size(), sizeof(o), hipMemcpyDeviceToHost, w.stream));
    HIP_OK(hipStreamSynchronize(w.stream));
    mmpbsa_binding_stats &s = r.summary.stats;
    s.mean_delta_mm = o[0];
    s.mean_delta_polar = o[1];
    s.mean_delta_sa = o[2];
    s.mean_delta_total = o[3];
    s.std_delta_total = o[4];
    s.std_delta_mm = o[5];
    s.std_delta_polar = o[6];
    s.std_delta_sa = o[7];
    s.sem_delta_total = o[8];
    r.summary.minus_tds = o[9];
    r.summary.temperature = temp;
}

// Continuous batching: each round takes frames from the waiting requests,
// oldest first (a request's remaining frames, then the next request's), while
// the round holds at most max_round frames and their admission estimate fits
// the HBM budget; requests submitted meanwhile join the next round.  Oldest
// first finishes requests in arrival order, so a request's latency is its
// queueing plus its own frames rather than the whole backlog.
void mmpbsa_server::loop()
{
    (void)hipSetDevice(device);
    for (;;) {
        std::vector<std::pair<ServerRequest *, int>> picks;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !pending.empty(); });
            if (pending.empty()) return;  // stop requested and nothing left
            double bytes = 0.0;
            while (!pending.empty() && (int)picks.size() < max_round) {
                ServerRequest *r = pending.front();
                if (!picks.empty() && bytes + r->frame_bytes > budget) break;
                if (!r->started) {
                    r->started = true;
                    r->t_start = Clock::now();
                }
                picks.push_back({r, r->next++});
                bytes += r->frame_bytes;
                if (r->next >= r->n_frames) pending.pop_front();
            }
            in_flight = bytes;
        }
        std::vector<FrameItem> items(picks.size());
        for (size_t k = 0; k < picks.size(); ++k) {
            ServerRequest *r = picks[k].first;
            const int f = picks[k].second;
            const size_t stride = 3 * (size_t)r->split->ctx->top.h.n;
            FrameItem &it = items[k];
            it.s = r->split;
            it.xyz = r->xyz.data() + stride * f;
            it.method = r->method;
            it.gb = r->gb;
            it.pb = r->pb;
            it.sa = r->sa;
            it.out = &r->out[f];
        }
        std::string err;
        int st = guarded(nullptr, [&] { run_items(*w, items, true, false); });
        if (st != MMPBSA_OK) err = g_err;
        // requests whose last frames ran in this round: their statistics on
        // the device before the waiters wake (done and status change only on
        // this thread, so they are read here without the lock)
        if (st == MMPBSA_OK) {
            std::map<ServerRequest *, int> ran;
            for (auto &pk : picks) ++ran[pk.first];
            for (auto &kv : ran) {
                ServerRequest *r = kv.first;
                if (r->status != MMPBSA_OK || r->done + kv.second != r->n_frames) continue;
                st = guarded(nullptr, [&] { summarize(*w, *r); });
                if (st != MMPBSA_OK) {
                    err = g_err;
                    break;
                }
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            in_flight = 0.0;
            for (auto &pk : picks) {
                ServerRequest *r = pk.first;
                if (st != MMPBSA_OK && r->status == MMPBSA_OK) {
                    r->status = st;
                    r->err = err;
                }
                if (++r->done == r->n_frames) r->t_end = Clock::now();
            }
        }
        cv_done.notify_all();
    }
}

namespace {

unsigned char *scratch_on(int device);  // the allocation-free GB entry point's array

void check_xyz(const double *xyz) {
    if (!xyz) throw DeviceError(MMPBSA_ERR_INVALID_ARG, "null coordinates");
}

}  // namespace

extern "C" {

const char *mmpbsa_version(void) { return "mmpbsa-hip 0.1.0 (gfx950)"; }

int mmpbsa_init(int32_t *n_devices)
{
    return guarded(nullptr, [&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        if (n_devices) *n_devices = n;
        require_device();
        // set-up of mmpbsa_gb_energy_noalloc's scratch on every gfx950 device
        for (int d = 0; d < n; ++d) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, d) != hipSuccess) {
                (void)hipGetLastError();
                continue;
            }
            if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) scratch_on(d);
        }
    });
}

void mmpbsa_gb_params_default(mmpbsa_gb_params *p)
{
    if (!p) return;
    p->model = MMPBSA_GB_OBC1;
    p->solute_dielectric = 1.0;
    p->solvent_dielectric = 78.3;
    p->salt_concentration = 0.15;
    p->temperature = 300.0;
    p->offset = 0.09;
    p->rgbmax = 25.0;
    p->cutoff = 9999.0;
}

void mmpbsa_pb_params_default(mmpbsa_pb_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->grid_spacing = 0.5;
    p->grid_buffer = 10.0;
    p->solute_dielectric = 1.0;
    p->solvent_dielectric = 80.0;
    p->salt_concentration = 0.15;
    p->temperature = 300.0;
    p->probe_radius = 1.4;
    p->ion_radius = 2.0;
    p->tolerance = 1e-6;
    p->max_iterations = 10000;
    p->fillratio = 4.0;
    p->fscale = 8;
}

void mmpbsa_sa_params_default(mmpbsa_sa_params *p)
{
    if (!p) return;
    p->surface_tension = 0.0072;
    p->offset = 0.0;
    p->probe_radius = 1.4;
    p->n_sphere_points = 480;
}

int mmpbsa_ctx_create(const mmpbsa_topology_view *top, int32_t device, mmpbsa_ctx **out)
{
    if (!top || !out) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<mmpbsa_ctx> c(new mmpbsa_ctx());
    const int rc = guarded(nullptr, [&] {
        require_device();
        HIP_OK(hipSetDevice(device));
        c->device = device;
        c->top.build(HostTopo::from_view(*top));
        c->w.reset(new Worker());
    });
    if (rc == MMPBSA_OK) *out = c.release();
    return rc;
}

int mmpbsa_ctx_destroy(mmpbsa_ctx *ctx)
{
    delete ctx;
    return MMPBSA_OK;
}

const char *mmpbsa_last_error(const mmpbsa_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_err.c_str();
}

int32_t mmpbsa_ctx_n_atoms(const mmpbsa_ctx *ctx) { return ctx ? ctx->top.h.n : -1; }

int mmpbsa_mm_energy(mmpbsa_ctx *ctx, const double *xyz, mmpbsa_mm_terms *out)
{
    if (!ctx || !out) return fail(ctx, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        check_xyz(xyz);
        Worker &w = *ctx->w;
        DBuf<double> d(w.pool, 3 * (size_t)ctx->top.h.n + 3);
        HIP_OK(hipMemcpyAsync(d.get(), xyz, sizeof(double) * 3 * ctx->top.h.n,
                              hipMemcpyHostToDevice, w.stream));
        run_mm(w, ctx->top, d.get(), out);
    });
}

int mmpbsa_gb_energy(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_gb_params *p,
                     double *born_radii_out, double *total_out)
{
    if (!ctx || !p || !total_out) return fail(ctx, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        check_xyz(xyz);
        Worker &w = *ctx->w;
        const int n = ctx->top.h.n;
        DBuf<double> d(w.pool, 3 * (size_t)n + 3);
        HIP_OK(hipMemcpyAsync(d.get(), xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, w.stream));
        std::vector<double> born;
        *total_out = run_gb(w, ctx->top, d.get(), *p, nullptr, born_radii_out ? &born : nullptr);
        if (born_radii_out && n) std::memcpy(born_radii_out, born.data(), sizeof(double) * n);
    });
}

int mmpbsa_sa_energy(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_sa_params *p,
                     double *per_atom_sasa_out, double *total_sasa_out, double *total_out)
{
    if (!ctx || !p) return fail(ctx, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        check_xyz(xyz);
        Worker &w = *ctx->w;
        const int n = ctx->top.h.n;
        DBuf<double> d(w.pool, 3 * (size_t)n + 3);
        HIP_OK(hipMemcpyAsync(d.get(), xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, w.stream));
        std::vector<double> per;
        double ts = 0.0, t = 0.0;
        run_sa(w, ctx->top, d.get(), xyz, *p, nullptr, per_atom_sasa_out ? &per : nullptr, &ts, &t);
        if (per_atom_sasa_out && n) std::memcpy(per_atom_sasa_out, per.data(), sizeof(double) * n);
        if (total_sasa_out) *total_sasa_out = ts;
        if (total_out) *total_out = t;
    });
}

int mmpbsa_pb_energy(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_pb_params *p,
                     mmpbsa_pb_result *out)
{
    if (!ctx || !p || !out) return fail(ctx, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        check_xyz(xyz);
        Worker &w = *ctx->w;
        const int n = ctx->top.h.n;
        DBuf<double> d(w.pool, 3 * (size_t)n + 3);
        HIP_OK(hipMemcpyAsync(d.get(), xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, w.stream));
        run_pb(w, ctx->top, d.get(), xyz, *p, out);
    });
}

// compute_pb_energy_solvated (pb_energy.rs:442-482): strip_to_solute, then
// compute_pb_energy of the solute.  The solute's topology stays resident on
// the context until a call names other residues.
int mmpbsa_pb_energy_solvated(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_pb_params *p,
                              const int32_t *solute_res, int32_t n_solute, mmpbsa_pb_result *out)
{
    if (!ctx || !p || !out || n_solute < 0 || (n_solute > 0 && !solute_res))
        return fail(ctx, MMPBSA_ERR_INVALID_ARG, "invalid argument");
    if (!solute_res) return mmpbsa_pb_energy(ctx, xyz, p, out);
    return guarded(ctx, [&] {
        check_xyz(xyz);
        const std::vector<int> res(solute_res, solute_res + n_solute);
        if (!ctx->solute || res != ctx->solute_res) {
            std::vector<int> sel = ctx->top.h.build_selection(res);  // prmtop.rs:165-204 errors
            std::unique_ptr<DevTopo> t(new DevTopo());
            t->build(ctx->top.h.extract(sel));
            ctx->solute = std::move(t);
            ctx->solute_sel = std::move(sel);
            ctx->solute_res = res;
        }
        Worker &w = *ctx->w;
        const std::vector<double> hx = gather_host(xyz, ctx->solute_sel);
        const int n = (int)ctx->solute_sel.size();
        DBuf<double> d(w.pool, 3 * (size_t)n + 3);
        if (n)
            HIP_OK(hipMemcpyAsync(d.get(), hx.data(), sizeof(double) * 3 * n,
                                  hipMemcpyHostToDevice, w.stream));
        run_pb(w, *ctx->solute, d.get(), hx.data(), *p, out);
    });
}

int mmpbsa_gb_energy_from_radii(mmpbsa_ctx *ctx, const double *xyz, const mmpbsa_gb_params *p,
                                const double *born_radii, double *total_out)
{
    if (!ctx || !p || !born_radii || !total_out)
        return fail(ctx, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        check_xyz(xyz);
        Worker &w = *ctx->w;
        const int n = ctx->top.h.n;
        DBuf<double> d(w.pool, 3 * (size_t)n + 3), R(w.pool, (size_t)n + 1);
        HIP_OK(hipMemcpyAsync(d.get(), xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, w.stream));
        HIP_OK(hipMemcpyAsync(R.get(), born_radii, sizeof(double) * n, hipMemcpyHostToDevice,
                              w.stream));
        *total_out = run_gb_from_radii(w, ctx->top, d.get(), *p, R.get());
    });
}

int mmpbsa_gb_cache_create(double rmsd_threshold, mmpbsa_gb_cache **out)
{
    if (!out) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    *out = new (std::nothrow) mmpbsa_gb_cache();
    if (!*out) return fail(nullptr, MMPBSA_ERR_OOM, "host allocation failed");
    (*out)->threshold = rmsd_threshold;
    return MMPBSA_OK;
}

int mmpbsa_gb_cache_destroy(mmpbsa_gb_cache *c)
{
    delete c;
    return MMPBSA_OK;
}

// compute_gb_energy_cached (gb_energy.rs:372-387): radii are reused when the
// cache holds as many atoms and the RMSD (GbCache::try_reuse, the reference's
// summation order) is below the threshold; otherwise they are computed and
// stored together with these coordinates.
int mmpbsa_gb_energy_cached(mmpbsa_ctx *ctx, mmpbsa_gb_cache *cache, const double *xyz,
                            const mmpbsa_gb_params *p, double *born_radii_out, double *total_out,
                            int32_t *reused)
{
    if (!ctx || !cache || !p || !total_out) return fail(ctx, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        check_xyz(xyz);
        const int n = ctx->top.h.n;
        bool hit = false;
        if (n > 0 && cache->xyz.size() == 3 * (size_t)n) {
            double ss = 0.0;
            for (int i = 0; i < n; ++i) {
                const double dx = cache->xyz[3 * i] - xyz[3 * i];
                const double dy = cache->xyz[3 * i + 1] - xyz[3 * i + 1];
                const double dz = cache->xyz[3 * i + 2] - xyz[3 * i + 2];
                ss += dx * dx + dy * dy + dz * dz;
            }
            hit = std::sqrt(ss / (double)n) < cache->threshold;
        }
        Worker &w = *ctx->w;
        DBuf<double> d(w.pool, 3 * (size_t)n + 3);
        HIP_OK(hipMemcpyAsync(d.get(), xyz, sizeof(double) * 3 * n, hipMemcpyHostToDevice, w.stream));
        if (hit) {
            DBuf<double> R(w.pool, (size_t)n + 1);
            HIP_OK(hipMemcpyAsync(R.get(), cache->radii.data(), sizeof(double) * n,
                                  hipMemcpyHostToDevice, w.stream));
            *total_out = run_gb_from_radii(w, ctx->top, d.get(), *p, R.get());
            if (born_radii_out) std::memcpy(born_radii_out, cache->radii.data(), sizeof(double) * n);
        } else {
            std::vector<double> born;
            *total_out = run_gb(w, ctx->top, d.get(), *p, nullptr, &born);
            cache->xyz.assign(xyz, xyz + 3 * (size_t)n);
            cache->radii = born;
            if (born_radii_out && n) std::memcpy(born_radii_out, born.data(), sizeof(double) * n);
        }
        if (reused) *reused = hit ? 1 : 0;
    });
}

// build_complex_subsystem (binding.rs:142-211) + extract_subtopology x2
int mmpbsa_split_create(mmpbsa_ctx *ctx, const int32_t *receptor_res, int32_t n_rec,
                        const int32_t *ligand_res, int32_t n_lig, mmpbsa_split **out)
{
    if (!ctx || !out || n_rec < 0 || n_lig < 0 || (n_rec && !receptor_res) || (n_lig && !ligand_res))
        return fail(ctx, MMPBSA_ERR_INVALID_ARG, "invalid argument");
    *out = nullptr;
    std::unique_ptr<mmpbsa_split> s(new mmpbsa_split());
    const int rc = guarded(ctx, [&] {
        const HostTopo &T = ctx->top.h;
        s->ctx = ctx;
        s->rec_res.assign(receptor_res, receptor_res + n_rec);
        s->lig_res.assign(ligand_res, ligand_res + n_lig);
        std::set<int> cres(s->rec_res.begin(), s->rec_res.end());
        cres.insert(s->lig_res.begin(), s->lig_res.end());
        const std::vector<int> cresv(cres.begin(), cres.end());
        const std::vector<int> csel = T.build_selection(cresv);
        HostTopo ctop;
        if ((int)csel.size() == T.n) {
            s->identity = true;
            ctop = T;
            s->rec_loc = s->rec_res;
            s->lig_loc = s->lig_res;
        } else {
            s->identity = false;
            s->csel = csel;
            ctop = T.extract(csel);
            for (int r : s->rec_res)
                s->rec_loc.push_back((int)(std::lower_bound(cresv.begin(), cresv.end(), r) - cresv.begin()));
            for (int r : s->lig_res)
                s->lig_loc.push_back((int)(std::lower_bound(cresv.begin(), cresv.end(), r) - cresv.begin()));
        }
        s->ra = ctop.build_selection(s->rec_loc);
        s->la = ctop.build_selection(s->lig_loc);
        s->ctop.build(ctop);
        s->rtop.build(ctop.extract(s->ra));
        s->ltop.build(ctop.extract(s->la));
        s->d_csel.upload(s->csel);
        s->d_ra.upload(s->ra);
        s->d_la.upload(s->la);
        // residue -> positions in [ra..., la...] for the decomposition sums
        const std::vector<int> ares = ctop.atom_residue_indices();
        std::vector<std::vector<int>> pos(ctop.n_res);
        for (size_t k = 0; k < s->ra.size(); ++k) pos[ares[s->ra[k]]].push_back((int)k);
        for (size_t k = 0; k < s->la.size(); ++k)
            pos[ares[s->la[k]]].push_back((int)(s->ra.size() + k));
        std::vector<int> sp, sptr(1, 0);
        for (auto &v : pos) {
            sp.insert(sp.end(), v.begin(), v.end());
            sptr.push_back((int)sp.size());
        }
        s->d_seg_pos.upload(sp);
        s->d_seg_ptr.upload(sptr);
        // any excluded / 1-4 pair between a receptor and a ligand atom?
        std::vector<char> is_r(ctop.n, 0), is_l(ctop.n, 0);
        for (int a : s->ra) is_r[a] = 1;
        for (int a : s->la) is_l[a] = 1;
        auto scan = [&](const std::vector<int> &ptr, const std::vector<int> &col) {
            for (int i = 0; i < ctop.n; ++i)
                for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
                    const int j = col[k];
                    if ((is_r[i] && is_l[j]) || (is_l[i] && is_r[j])) s->cross_excl = 1;
                }
        };
        scan(s->ctop.h_ex_ptr, s->ctop.h_ex_col);
        scan(s->ctop.h_p14_ptr, s->ctop.h_p14_col);
        bool cover = true;
        for (int a = 0; a < ctop.n && cover; ++a) cover = is_r[a] != is_l[a];
        s->gb_parts = cover;
        bool parts = cover && !s->cross_excl;
        auto one_side = [&](const std::vector<int> &terms, int stride, int atoms) {
            for (size_t b = 0; b + stride <= terms.size() && parts; b += stride) {
                const bool r = is_r[terms[b]];
                for (int k = 1; k < atoms; ++k)
                    if ((bool)is_r[terms[b + k]] != r) parts = false;
            }
        };
        one_side(ctop.bonds, 3, 2);
        one_side(ctop.angles, 4, 3);
        one_side(ctop.dihs, 6, 4);
        s->mm_parts = parts;
    });
    if (rc == MMPBSA_OK) *out = s.release();
    return rc;
}

int mmpbsa_split_destroy(mmpbsa_split *s)
{
    delete s;
    return MMPBSA_OK;
}

int mmpbsa_split_atoms(const mmpbsa_split *s, int32_t which, int32_t *out, int32_t cap,
                       int32_t *n_out)
{
    if (!s || which < 0 || which > 2) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "invalid argument");
    std::vector<int> v;
    if (which == 0) {
        if (s->identity) {
            v.resize(s->ctx->top.h.n);
            for (int i = 0; i < (int)v.size(); ++i) v[i] = i;
        } else {
            v = s->csel;
        }
    } else {
        v = which == 1 ? s->ra : s->la;
    }
    if (n_out) *n_out = (int32_t)v.size();
    if (out) std::memcpy(out, v.data(), sizeof(int32_t) * std::min<size_t>(v.size(), cap > 0 ? cap : 0));
    return MMPBSA_OK;
}

int mmpbsa_frame_energy(mmpbsa_split *s, const double *xyz, int32_t method,
                        const mmpbsa_gb_params *gb, const mmpbsa_pb_params *pb,
                        const mmpbsa_sa_params *sa, mmpbsa_frame_terms *out)
{
    if (!s || !out) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    mmpbsa_ctx *ctx = s->ctx;
    return guarded(ctx, [&] {
        check_xyz(xyz);
        mmpbsa_gb_params g;
        mmpbsa_pb_params p;
        mmpbsa_sa_params a;
        mmpbsa_gb_params_default(&g);
        mmpbsa_pb_params_default(&p);
        mmpbsa_sa_params_default(&a);
        if (gb) g = *gb;
        if (pb) p = *pb;
        if (sa) a = *sa;
        frame_energy(*ctx->w, s, xyz, method, g, p, a, out);
    });
}

}  // extern "C"

namespace {
int frames_energy_impl(mmpbsa_split *s, const double *xyz_frames, int32_t n_frames,
                       int32_t method, const mmpbsa_gb_params *gb, const mmpbsa_pb_params *pb,
                       const mmpbsa_sa_params *sa, int32_t mode, mmpbsa_frame_terms *out,
                       bool on_device)
{
    if (!s || !out || n_frames < 0) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "invalid argument");
    mmpbsa_ctx *ctx = s->ctx;
    return guarded(ctx, [&] {
        if (n_frames == 0) return;
        check_xyz(xyz_frames);
        mmpbsa_gb_params g;
        mmpbsa_pb_params p;
        mmpbsa_sa_params a;
        mmpbsa_gb_params_default(&g);
        mmpbsa_pb_params_default(&p);
        mmpbsa_sa_params_default(&a);
        if (gb) g = *gb;
        if (pb) p = *pb;
        if (sa) a = *sa;
        const size_t stride = 3 * (size_t)ctx->top.h.n;
        if (mode == 0) {
            for (int f = 0; f < n_frames; ++f)
                frames_group(*ctx->w, s, xyz_frames + stride * f, 1, 0, method, g, p, a, out + f,
                             false, on_device);
            return;
        }
        // batched: groups of frames whose PB solves advance together
        const int group = 8;
        for (int f0 = 0; f0 < n_frames; f0 += group) {
            const int nf = std::min(group, n_frames - f0);
            frames_group(*ctx->w, s, xyz_frames + stride * f0, nf, stride, method, g, p, a,
                         out + f0, true, on_device);
        }
    });
}
}  // namespace

extern "C" {

int mmpbsa_frames_energy(mmpbsa_split *s, const double *xyz_frames, int32_t n_frames,
                         int32_t method, const mmpbsa_gb_params *gb, const mmpbsa_pb_params *pb,
                         const mmpbsa_sa_params *sa, int32_t mode, mmpbsa_frame_terms *out)
{
    return frames_energy_impl(s, xyz_frames, n_frames, method, gb, pb, sa, mode, out, false);
}

int mmpbsa_frames_energy_device(mmpbsa_split *s, const double *d_xyz_frames, int32_t n_frames,
                                int32_t method, const mmpbsa_gb_params *gb,
                                const mmpbsa_pb_params *pb, const mmpbsa_sa_params *sa,
                                int32_t mode, mmpbsa_frame_terms *out)
{
    return frames_energy_impl(s, d_xyz_frames, n_frames, method, gb, pb, sa, mode, out, true);
}

// decompose_binding_energy_impl (decomposition.rs:210-410)
namespace {

// decompose_binding_energy_impl (decomposition.rs:210-410).  The polar share
// comes from GB weights (gbp != null, the reference's rule) or, as an
// extension, from per-atom PB terms (pbp != null): 0.5 q_i (dphi_cplx -
// dphi_iso)(r_i) with dphi = phi_solv - phi_ref, whose sum over all residues
// is dG_PB (returned in *delta_pb).
void decompose_impl(mmpbsa_split *s, const double *xyz, const mmpbsa_gb_params *gbp,
                    const mmpbsa_pb_params *pbp, const mmpbsa_sa_params &sa,
                    mmpbsa_residue_contrib *rec_out, mmpbsa_residue_contrib *lig_out,
                    double *delta_pb)
{
    check_xyz(xyz);
    Worker &w = *s->ctx->w;
    FrameCoords f;
    stage_frame(w, s, xyz, f, false);
    const int nr = (int)s->ra.size(), nl = (int)s->la.size(), na = nr + nl;
    const PairTopo cv = s->ctop.view();
    // per-atom vdW / elec halves (receptor block, then ligand block)
    DBuf<double> vdw(w.pool, na + 1), elec(w.pool, na + 1);
    {
        const size_t wr = (size_t)decomp_chunks(nl) * std::max(nr, 1) * 4;
        const size_t wl = (size_t)decomp_chunks(nr) * std::max(nl, 1) * 4;
        DBuf<double> work(w.pool, std::max(wr, wl));
        launch_decomp_pairs(f.dc, cv, s->d_ra.p, nr, s->d_la.p, nl, 1, s->cross_excl,
                            work.get(), vdw.get(), elec.get(), w.stream);
        launch_decomp_pairs(f.dc, cv, s->d_la.p, nl, s->d_ra.p, nr, 0, s->cross_excl,
                            work.get(), vdw.get() + nr, elec.get() + nr, w.stream);
    }
    DBuf<double> pol(w.pool, na + 1);
    if (gbp) {
        const mmpbsa_gb_params &gb = *gbp;
        DBuf<double> Rc(w.pool, s->ctop.h.n + 1), Rr(w.pool, nr + 1), Rl(w.pool, nl + 1);
        const double ec = run_gb(w, s->ctop, f.dc, gb, Rc.get(), nullptr);
        const double er = run_gb(w, s->rtop, f.r.get(), gb, Rr.get(), nullptr);
        const double el = run_gb(w, s->ltop, f.l.get(), gb, Rl.get(), nullptr);
        const double fac = -0.5 * (1.0 / gb.solute_dielectric - 1.0 / gb.solvent_dielectric);
        DBuf<double> wts(w.pool, na + 1), scal(w.pool, 2);
        launch_gb_weight(s->d_ra.p, nr, s->ctop.qa.p, Rc.get(), Rr.get(), fac, wts.get(), w.stream);
        launch_gb_weight(s->d_la.p, nl, s->ctop.qa.p, Rc.get(), Rl.get(), fac, wts.get() + nr,
                         w.stream);
        launch_fill_f64(scal.get(), ec - er - el, 1, w.stream);
        launch_sum_partials(wts.get(), na, 1, 1, scal.get() + 1, w.stream);
        launch_gb_share(wts.get(), na, scal.get(), pol.get(), w.stream);
    } else {
        // the three subsystems' PB solves in one batch, with per-atom terms
        DBuf<double> Tc(w.pool, s->ctop.h.n + 1), Tr(w.pool, nr + 1), Tl(w.pool, nl + 1);
        mmpbsa_pb_result res[3];
        std::vector<PbJob> jobs(3);
        const DevTopo *tops[3] = {&s->ctop, &s->rtop, &s->ltop};
        const double *dx[3] = {f.dc, f.r.get(), f.l.get()};
        const double *hx[3] = {f.hcp, f.hr.data(), f.hl.data()};
        double *tt[3] = {Tc.get(), Tr.get(), Tl.get()};
        for (int k = 0; k < 3; ++k) {
            jobs[k].t = tops[k];
            jobs[k].d_xyz = dx[k];
            jobs[k].h_xyz = hx[k];
            jobs[k].p = *pbp;
            jobs[k].out = &res[k];
            jobs[k].d_atom_terms = tt[k];
        }
        run_pb_jobs(w, jobs, true);
        launch_sa_share(s->d_ra.p, nr, Tc.get(), Tr.get(), 1.0, pol.get(), w.stream);
        launch_sa_share(s->d_la.p, nl, Tc.get(), Tl.get(), 1.0, pol.get() + nr, w.stream);
        if (delta_pb) *delta_pb = res[0].total - res[1].total - res[2].total;
    }
    // SA shares
    DBuf<double> Sc(w.pool, s->ctop.h.n + 1), Sr(w.pool, nr + 1), Sl(w.pool, nl + 1);
    double ts, tt;
    run_sa(w, s->ctop, f.dc, f.hcp, sa, Sc.get(), nullptr, &ts, &tt);
    run_sa(w, s->rtop, f.r.get(), f.hr.data(), sa, Sr.get(), nullptr, &ts, &tt);
    run_sa(w, s->ltop, f.l.get(), f.hl.data(), sa, Sl.get(), nullptr, &ts, &tt);
    DBuf<double> sav(w.pool, na + 1);
    launch_sa_share(s->d_ra.p, nr, Sc.get(), Sr.get(), sa.surface_tension, sav.get(), w.stream);
    launch_sa_share(s->d_la.p, nl, Sc.get(), Sl.get(), sa.surface_tension, sav.get() + nr,
                    w.stream);
    // per-residue sums (complex numbering)
    const int ncr = s->ctop.h.n_res;
    DBuf<double> res(w.pool, 4 * (size_t)std::max(ncr, 1));
    launch_segment_sum(vdw.get(), s->d_seg_pos.p, s->d_seg_ptr.p, ncr, res.get(), w.stream);
    launch_segment_sum(elec.get(), s->d_seg_pos.p, s->d_seg_ptr.p, ncr, res.get() + ncr, w.stream);
    launch_segment_sum(pol.get(), s->d_seg_pos.p, s->d_seg_ptr.p, ncr, res.get() + 2 * ncr, w.stream);
    launch_segment_sum(sav.get(), s->d_seg_pos.p, s->d_seg_ptr.p, ncr, res.get() + 3 * ncr, w.stream);
    std::vector<double> h(4 * (size_t)std::max(ncr, 1));
    HIP_OK(hipMemcpyAsync(h.data(), res.get(), sizeof(double) * 4 * ncr, hipMemcpyDeviceToHost,
                          w.stream));
    w.sync();
    auto fill = [&](const std::vector<int> &orig, const std::vector<int> &loc,
                    mmpbsa_residue_contrib *o) {
        for (size_t k = 0; k < orig.size(); ++k) {
            const int r = loc[k];
            o[k].residue_index = orig[k];
            o[k].vdw = h[r];
            o[k].elec = h[ncr + r];
            o[k].polar = h[2 * ncr + r];
            o[k].sa = h[3 * ncr + r];
        }
    };
    fill(s->rec_res, s->rec_loc, rec_out);
    fill(s->lig_res, s->lig_loc, lig_out);
}

}  // namespace

int mmpbsa_decompose(mmpbsa_split *s, const double *xyz, const mmpbsa_gb_params *gbp,
                     const mmpbsa_sa_params *sap, mmpbsa_residue_contrib *rec_out,
                     mmpbsa_residue_contrib *lig_out)
{
    if (!s || (!rec_out && !s->rec_res.empty()) || (!lig_out && !s->lig_res.empty()))
        return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(s->ctx, [&] {
        mmpbsa_gb_params gb;
        mmpbsa_sa_params sa;
        mmpbsa_gb_params_default(&gb);
        mmpbsa_sa_params_default(&sa);
        if (gbp) gb = *gbp;
        if (sap) sa = *sap;
        decompose_impl(s, xyz, &gb, nullptr, sa, rec_out, lig_out, nullptr);
    });
}

int mmpbsa_decompose_pb(mmpbsa_split *s, const double *xyz, const mmpbsa_pb_params *pbp,
                        const mmpbsa_sa_params *sap, mmpbsa_residue_contrib *rec_out,
                        mmpbsa_residue_contrib *lig_out, double *delta_pb)
{
    if (!s || (!rec_out && !s->rec_res.empty()) || (!lig_out && !s->lig_res.empty()))
        return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(s->ctx, [&] {
        mmpbsa_pb_params pb;
        mmpbsa_sa_params sa;
        mmpbsa_pb_params_default(&pb);
        mmpbsa_sa_params_default(&sa);
        if (pbp) pb = *pbp;
        if (sap) sa = *sap;
        decompose_impl(s, xyz, nullptr, &pb, sa, rec_out, lig_out, delta_pb);
    });
}

int mmpbsa_binding_energy_trajectory(mmpbsa_split *s, const char *path, int32_t format,
                                     int32_t has_box, int64_t stride, int64_t start_frame,
                                     int64_t end_frame, int32_t method,
                                     const mmpbsa_gb_params *gb, const mmpbsa_pb_params *pb,
                                     const mmpbsa_sa_params *sa, int32_t *n_frames)
{
    if (!s || !path || !n_frames) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    *n_frames = 0;
    mmpbsa_ctx *ctx = s->ctx;
    return guarded(ctx, [&] {
        mmpbsa_gb_params g;
        mmpbsa_pb_params p;
        mmpbsa_sa_params a;
        mmpbsa_gb_params_default(&g);
        mmpbsa_pb_params_default(&p);
        mmpbsa_sa_params_default(&a);
        if (gb) g = *gb;
        if (pb) p = *pb;
        if (sa) a = *sa;
        const int64_t st = std::max<int64_t>(1, stride);  // binding.rs:525
        const int64_t end = end_frame < 0 ? INT64_MAX : end_frame;
        const int n_all = ctx->top.h.n;
        const size_t nc = complex_atoms(s);
        const std::vector<int> &cb = s->ctop.h.bonds;  // (i, j, type), complex-local
        std::vector<double> full(3 * (size_t)n_all), box(6, 0.0);
        std::unique_ptr<DcdFile> dcd;
        std::unique_ptr<MdcrdFile> mdcrd;
        try {
            if (format == MMPBSA_TRAJ_DCD) {
                dcd = std::make_unique<DcdFile>(path);
                if (dcd->hd.n_atoms < n_all)
                    throw TrajError("DCD has " + std::to_string(dcd->hd.n_atoms) +
                                    " atoms, the topology " + std::to_string(n_all));
                full.resize(3 * (size_t)dcd->hd.n_atoms);
            } else {
                mdcrd = std::make_unique<MdcrdFile>(path, n_all, has_box != 0);
            }
        } catch (const TrajError &e) {
            throw DeviceError(MMPBSA_ERR_IO, e.what());
        }
        const int kBatch = 64;  // binding.rs:527
        std::vector<std::vector<double>> batch;
        std::vector<mmpbsa_frame_terms> out;
        auto flush = [&] {
            if (batch.empty()) return;
            const size_t base = out.size();
            out.resize(base + batch.size());
            std::vector<FrameItem> items(batch.size());
            for (size_t k = 0; k < batch.size(); ++k) {
                FrameItem &it = items[k];
                it.s = s;
                it.xyz = batch[k].data();
                it.method = method;
                it.gb = g;
                it.pb = p;
                it.sa = a;
                it.out = &out[base + k];
                it.complex_local = true;
            }
            run_items(*ctx->w, items, true, false);
            batch.clear();
        };
        std::vector<double> last;
        int64_t idx = 0;
        for (;;) {
            bool got;
            try {
                got = dcd ? dcd->read_frame(full.data(), box.data())
                          : mdcrd->read_frame(full.data(), box.data());
            } catch (const TrajError &e) {
                throw DeviceError(MMPBSA_ERR_IO, e.what());
            }
            if (!got || idx >= end) break;
            if (idx >= start_frame && (idx - start_frame) % st == 0) {
                std::vector<double> c(3 * nc);
                const bool nm = dcd != nullptr;  // DCD positions are nm: back to A
                for (size_t k = 0; k < nc; ++k) {
                    const size_t src = s->identity ? k : (size_t)s->csel[k];
                    for (int d = 0; d < 3; ++d)
                        c[3 * k + d] = nm ? full[3 * src + d] * 10.0 : full[3 * src + d];
                }
                const bool boxed = dcd ? dcd->hd.has_unit_cell != 0 : has_box != 0;
                if (boxed) {
                    const double bx[3] = {nm ? box[0] * 10.0 : box[0], nm ? box[1] * 10.0 : box[1],
                                          nm ? box[2] * 10.0 : box[2]};
                    reimage_complex(c.data(), (int)nc, cb.data(), (int)(cb.size() / 3), 3, bx);
                }
                last = c;
                batch.push_back(std::move(c));
                if ((int)batch.size() >= kBatch) flush();
            }
            ++idx;
        }
        flush();
        if (out.empty()) throw DeviceError(MMPBSA_ERR_INVALID_ARG, "No frames found in trajectory");
        s->traj_out = std::move(out);
        s->traj_last = std::move(last);
        *n_frames = (int32_t)s->traj_out.size();
    });
}

int mmpbsa_trajectory_frames(const mmpbsa_split *s, mmpbsa_frame_terms *out, int32_t n,
                             double *last_frame_xyz)
{
    if (!s || (!out && n > 0)) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    if (n > (int32_t)s->traj_out.size())
        return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "more frames requested than computed");
    if (n > 0) std::memcpy(out, s->traj_out.data(), sizeof(mmpbsa_frame_terms) * n);
    if (last_frame_xyz && !s->traj_last.empty())
        std::memcpy(last_frame_xyz, s->traj_last.data(), sizeof(double) * s->traj_last.size());
    return MMPBSA_OK;
}

int mmpbsa_binding_stats_compute(const mmpbsa_frame_terms *frames, int32_t n,
                                 mmpbsa_binding_stats *out)
{
    if (!out || (n > 0 && !frames)) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    if (n <= 0) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "No frames found in trajectory");
    auto mean = [&](double mmpbsa_frame_terms::*m) {
        double acc = 0.0;
        for (int32_t k = 0; k < n; ++k) acc += frames[k].*m;
        return acc / (double)n;
    };
    auto sample_std = [&](double mmpbsa_frame_terms::*m, double mu) {
        if (n <= 1) return 0.0;
        double acc = 0.0;
        for (int32_t k = 0; k < n; ++k) {
            const double d = frames[k].*m - mu;
            acc += d * d;
        }
        return std::sqrt(acc / (double)(n - 1));
    };
    out->mean_delta_mm = mean(&mmpbsa_frame_terms::delta_mm);
    out->mean_delta_polar = mean(&mmpbsa_frame_terms::delta_polar);
    out->mean_delta_sa = mean(&mmpbsa_frame_terms::delta_sa);
    out->mean_delta_total = mean(&mmpbsa_frame_terms::delta_total);
    out->std_delta_total = sample_std(&mmpbsa_frame_terms::delta_total, out->mean_delta_total);
    out->std_delta_mm = sample_std(&mmpbsa_frame_terms::delta_mm, out->mean_delta_mm);
    out->std_delta_polar = sample_std(&mmpbsa_frame_terms::delta_polar, out->mean_delta_polar);
    out->std_delta_sa = sample_std(&mmpbsa_frame_terms::delta_sa, out->mean_delta_sa);
    out->sem_delta_total = out->std_delta_total / std::sqrt((double)n);
    return MMPBSA_OK;
}

int mmpbsa_interaction_entropy(const mmpbsa_frame_terms *frames, int32_t n, double temperature,
                               double *minus_tds)
{
    if (!minus_tds || (n > 0 && !frames))
        return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    if (n < 2) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "interaction entropy needs >= 2 frames");
    const double kt = 0.00198688 * temperature;
    double mean_de = 0.0;
    for (int32_t k = 0; k < n; ++k) mean_de += frames[k].delta_mm;
    mean_de /= (double)n;
    std::vector<double> ex(n);
    double mx = -std::numeric_limits<double>::infinity();
    for (int32_t k = 0; k < n; ++k) {
        ex[k] = (frames[k].delta_mm - mean_de) / kt;
        mx = std::max(mx, ex[k]);
    }
    double acc = 0.0;
    for (int32_t k = 0; k < n; ++k) acc += std::exp(ex[k] - mx);
    *minus_tds = kt * (mx + std::log(acc / (double)n));
    return MMPBSA_OK;
}

int mmpbsa_quasi_harmonic_entropy(int32_t device, const double *xyz_frames, int32_t n_frames,
                                  int32_t n_atoms, const double *masses, double temperature,
                                  double *minus_tds)
{
    if (!minus_tds || n_frames < 0 || n_atoms < 0 ||
        (n_frames > 0 && n_atoms > 0 && (!xyz_frames || !masses)))
        return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "invalid argument");
    if (n_frames < 2 || n_atoms == 0)
        return fail(nullptr, MMPBSA_ERR_INVALID_ARG,
                    "quasi-harmonic entropy needs >= 2 frames and >= 1 atom");
    return guarded(nullptr, [&] {
        require_device();
        HIP_OK(hipSetDevice(device));
        Worker w;
        *minus_tds = qh_entropy(w, xyz_frames, n_frames, n_atoms, masses, temperature);
    });
}

void mmpbsa_server_params_default(mmpbsa_server_params *p)
{
    if (!p) return;
    p->hbm_budget_bytes = 0;
    p->max_round_frames = 16;
}

int mmpbsa_server_create(const mmpbsa_server_params *params, int32_t device, mmpbsa_server **out)
{
    if (!out) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    return guarded(nullptr, [&] {
        require_device();
        mmpbsa_server_params p;
        mmpbsa_server_params_default(&p);
        if (params) p = *params;
        if (p.max_round_frames <= 0 || p.hbm_budget_bytes < 0)
            throw DeviceError(MMPBSA_ERR_INVALID_ARG, "bad server parameters");
        HIP_OK(hipSetDevice(device));
        auto srv = std::make_unique<mmpbsa_server>();
        srv->device = device;
        srv->max_round = p.max_round_frames;
        if (p.hbm_budget_bytes > 0) {
            srv->budget = (double)p.hbm_budget_bytes;
        } else {
            size_t free_b = 0, total_b = 0;
            HIP_OK(hipMemGetInfo(&free_b, &total_b));
            srv->budget = 0.75 * (double)free_b;
        }
        srv->w = std::make_unique<Worker>();
        srv->th = std::thread([s = srv.get()] { s->loop(); });
        *out = srv.release();
    });
}

int mmpbsa_server_destroy(mmpbsa_server *srv)
{
    if (!srv) return MMPBSA_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->th.joinable()) srv->th.join();
    delete srv;
    return MMPBSA_OK;
}

int mmpbsa_server_submit(mmpbsa_server *srv, mmpbsa_split *split, const double *xyz_frames,
                         int32_t n_frames, int32_t method, const mmpbsa_gb_params *gb,
                         const mmpbsa_pb_params *pb, const mmpbsa_sa_params *sa, uint64_t *ticket)
{
    if (!srv || !split || !ticket || n_frames <= 0)
        return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "invalid argument");
    return guarded(nullptr, [&] {
        check_xyz(xyz_frames);
        auto r = std::make_unique<ServerRequest>();
        r->split = split;
        r->n_frames = n_frames;
        r->method = method;
        mmpbsa_gb_params_default(&r->gb);
        mmpbsa_pb_params_default(&r->pb);
        mmpbsa_sa_params_default(&r->sa);
        if (gb) r->gb = *gb;
        if (pb) r->pb = *pb;
        if (sa) r->sa = *sa;
        const size_t stride = 3 * (size_t)split->ctx->top.h.n;
        r->xyz.assign(xyz_frames, xyz_frames + stride * (size_t)n_frames);
        r->out.assign(n_frames, mmpbsa_frame_terms{});
        r->frame_bytes = frame_bytes_estimate(split, xyz_frames, method, r->pb);
        if (r->frame_bytes > srv->budget)
            throw DeviceError(MMPBSA_ERR_TOO_LARGE, "one frame needs more device memory (" +
                                                        std::to_string((int64_t)r->frame_bytes) +
                                                        " B) than the server's budget");
        r->t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stop) throw DeviceError(MMPBSA_ERR_INVALID_ARG, "server is shutting down");
        r->ticket = srv->next_ticket++;
        *ticket = r->ticket;
        srv->pending.push_back(r.get());
        srv->reqs[r->ticket] = std::move(r);
        srv->cv_work.notify_one();
    });
}

}  // extern "C"

namespace {
// Copy out a finished request and release its ticket (caller holds srv->mu).
int finish_request(mmpbsa_server *srv, std::unordered_map<uint64_t,
                   std::unique_ptr<ServerRequest>>::iterator it, mmpbsa_frame_terms *out,
                   mmpbsa_request_stats *stats, mmpbsa_request_summary *summary = nullptr)
{
    ServerRequest &r = *it->second;
    if (summary && r.status == MMPBSA_OK) *summary = r.summary;
    auto ms = [](Clock::time_point a, Clock::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    if (stats) {
        stats->queue_ms = ms(r.t_submit, r.t_start);
        stats->total_ms = ms(r.t_submit, r.t_end);
        stats->n_frames = r.n_frames;
        stats->status = r.status;
    }
    const int st = r.status;
    const std::string err = r.err;
    if (out && st == MMPBSA_OK) std::memcpy(out, r.out.data(), sizeof(mmpbsa_frame_terms) * r.n_frames);
    srv->reqs.erase(it);
    if (st != MMPBSA_OK) return fail(nullptr, st, err);
    return MMPBSA_OK;
}
}  // namespace

extern "C" {

int mmpbsa_server_poll(mmpbsa_server *srv, uint64_t ticket, mmpbsa_frame_terms *out,
                       mmpbsa_request_stats *stats)
{
    if (!srv) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "unknown ticket");
    if (it->second->done < it->second->n_frames) return MMPBSA_ERR_NOT_READY;
    return finish_request(srv, it, out, stats);
}

int mmpbsa_server_wait(mmpbsa_server *srv, uint64_t ticket, mmpbsa_frame_terms *out,
                       mmpbsa_request_stats *stats)
{
    return mmpbsa_server_wait_summary(srv, ticket, out, stats, nullptr);
}

int mmpbsa_server_poll_summary(mmpbsa_server *srv, uint64_t ticket, mmpbsa_frame_terms *out,
                               mmpbsa_request_stats *stats, mmpbsa_request_summary *summary)
{
    if (!srv) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "unknown ticket");
    if (it->second->done < it->second->n_frames) return MMPBSA_ERR_NOT_READY;
    return finish_request(srv, it, out, stats, summary);
}

int mmpbsa_server_wait_summary(mmpbsa_server *srv, uint64_t ticket, mmpbsa_frame_terms *out,
                               mmpbsa_request_stats *stats, mmpbsa_request_summary *summary)
{
    if (!srv) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "unknown ticket");
    ServerRequest *r = it->second.get();
    srv->cv_done.wait(lk, [&] { return r->done >= r->n_frames; });
    return finish_request(srv, srv->reqs.find(ticket), out, stats, summary);
}

int mmpbsa_device_alloc(int32_t device, int64_t bytes, void **out)
{
    if (!out || bytes < 0) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "invalid argument");
    *out = nullptr;
    return guarded(nullptr, [&] {
        require_device();
        HIP_OK(hipSetDevice(device));
        HIP_OK(hipMalloc(out, (size_t)(bytes > 0 ? bytes : 1)));
    });
}

int mmpbsa_device_free(void *ptr)
{
    return guarded(nullptr, [&] {
        std::unique_lock<std::shared_mutex> g(capture_gate());
        if (ptr) HIP_OK(hipFree(ptr));
    });
}

int mmpbsa_device_upload(void *dst, const void *src, int64_t bytes)
{
    if ((!dst || !src) && bytes > 0) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null pointer");
    return guarded(nullptr, [&] {
        std::unique_lock<std::shared_mutex> g(capture_gate());
        if (bytes > 0) HIP_OK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice));
    });
}

int mmpbsa_device_download(void *dst, const void *src, int64_t bytes)
{
    if ((!dst || !src) && bytes > 0) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null pointer");
    return guarded(nullptr, [&] {
        std::unique_lock<std::shared_mutex> g(capture_gate());
        if (bytes > 0) HIP_OK(hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost));
    });
}

int mmpbsa_device_copy(void *dst, const void *src, int64_t bytes)
{
    if ((!dst || !src) && bytes > 0) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null pointer");
    if (bytes % 16) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "bytes must be a multiple of 16");
    return guarded(nullptr, [&] {
        std::unique_lock<std::shared_mutex> g(capture_gate());
        launch_copy16(src, dst, bytes / 16, nullptr);
        HIP_OK(hipDeviceSynchronize());
    });
}

int mmpbsa_tuning_set(const mmpbsa_tuning *t)
{
    Tuning v = tuning();
    if (!t) {
        v.check_cycles = 8;
        v.gb_chunks = 0;
        v.entries.clear();
        set_gb_chunks_override(0);
        set_tuning(v);
        return MMPBSA_OK;
    }
    if (t->n_entries < 0 || t->n_entries > 256 || (t->n_entries > 0 && !t->entries) ||
        t->check_cycles < 0 || t->gb_chunks < 0)
        return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "invalid tuning");
    v.check_cycles = t->check_cycles > 0 ? t->check_cycles : 8;
    v.gb_chunks = t->gb_chunks;
    v.entries.clear();
    for (int k = 0; k < t->n_entries; ++k) {
        const mmpbsa_tune_entry &e = t->entries[k];
        TuneEntry q;
        q.nodes_min = e.nodes_min;
        q.nodes_max = e.nodes_max;
        q.solves_min = e.solves_min;
        q.solves_max = e.solves_max;
        q.smoother_chunks = e.smoother_chunks;
        if (e.smoother_region < 0 || e.smoother_region > 2)
            return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "invalid smoother region");
        q.smoother_region = e.smoother_region;
        v.entries.push_back(q);
    }
    set_gb_chunks_override(v.gb_chunks);
    set_tuning(v);
    return MMPBSA_OK;
}

int mmpbsa_tuning_get(mmpbsa_tuning *t, mmpbsa_tune_entry *entries, int32_t cap)
{
    if (!t) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    const Tuning v = tuning();
    t->check_cycles = v.check_cycles;
    t->gb_chunks = v.gb_chunks;
    t->n_entries = (int32_t)v.entries.size();
    t->entries = nullptr;
    for (int k = 0; k < (int)v.entries.size() && k < cap && entries; ++k) {
        entries[k].nodes_min = v.entries[k].nodes_min;
        entries[k].nodes_max = v.entries[k].nodes_max;
        entries[k].solves_min = v.entries[k].solves_min;
        entries[k].solves_max = v.entries[k].solves_max;
        entries[k].smoother_chunks = v.entries[k].smoother_chunks;
        entries[k].smoother_region = v.entries[k].smoother_region;
    }
    return MMPBSA_OK;
}

int32_t mmpbsa_tuning_smoother_zc(int32_t nx, int32_t ny, int32_t nz, int32_t n_solves)
{
    const Tuning t = tuning();
    return tuned_smooth3_zc(t, nx, ny, nz, n_solves, tuned_smooth3_region(t, nx, ny, nz, n_solves));
}

int32_t mmpbsa_tuning_smoother_region(int32_t nx, int32_t ny, int32_t nz, int32_t n_solves)
{
    return tuned_smooth3_region(tuning(), nx, ny, nz, n_solves) + 1;
}

int mmpbsa_tuning_observe(int32_t on)
{
    tuning_observe(on != 0);
    return MMPBSA_OK;
}

int32_t mmpbsa_tuning_observed(int64_t *nodes, int32_t *solves, int64_t *count, int32_t cap)
{
    return tuning_observed(nodes, solves, count, cap);
}

int mmpbsa_profile_enable(int32_t on)
{
    kernel_profile_reset();
    std::lock_guard<std::mutex> lk(profile_mutex());
    ProfileStats &p = profile_stats();
    p.enabled = on != 0;
    p.total_ms = 0.0;
    p.launches = 0;
    p.alg_bytes = 0.0;
    return MMPBSA_OK;
}

int mmpbsa_profile_kernel(const char *name, double *total_ms, int64_t *launches, double *work)
{
    if (!name) return fail(nullptr, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(nullptr, [&] { kernel_profile_read(name, total_ms, launches, work); });
}

int mmpbsa_profile_read(double *total_ms, int64_t *launches, double *alg_bytes,
                        int64_t *grid_threads)
{
    std::lock_guard<std::mutex> lk(profile_mutex());
    const ProfileStats &p = profile_stats();
    if (total_ms) *total_ms = p.total_ms;
    if (launches) *launches = p.launches;
    if (alg_bytes) *alg_bytes = p.alg_bytes;
    if (grid_threads) *grid_threads = p.grid_threads;
    return MMPBSA_OK;
}

}  // extern "C"

// ---- GB energy without device allocation ------------------------------------
// Every temporary of mmpbsa_gb_energy (Born-radius partial sums, the radii
// when the caller does not want them, energy partials) lives in one fixed
// device array per GPU.  Host code never uses the array's name as an
// address: hipGetSymbolAddress gives each device's own copy once, and only
// that pointer is passed to kernels.
namespace {

constexpr size_t kNoallocScratchBytes = (size_t)16 << 20;

std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch_ptr;  // device -> address of its scratch copy

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

// Looks up (once per device, with that device current) the device address of
// `scratch`; throws on failure.
unsigned char *scratch_on(int device)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    int prev = 0;
    HIP_OK(hipGetDevice(&prev));
    HIP_OK(hipSetDevice(device));
    unsigned char *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(prev);
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        throw DeviceError(MMPBSA_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                                 hipGetErrorString(e));
    }
    g_scratch_ptr[device] = ptr;
    return ptr;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of scratch a GB energy of n atoms needs (radii included when the
// caller passes no output for them).
size_t gb_noalloc_bytes(int n, bool radii_in_scratch)
{
    const size_t nb = (size_t)std::max(1, ceil_div(n, kTile));
    return align256(sizeof(double) * ((size_t)gb_chunks(n) * n + 1)) +
           (radii_in_scratch ? align256(sizeof(double) * ((size_t)n + 1)) : 0) +
           align256(sizeof(double) * nb * nb);
}

}  // namespace

extern "C" {

int mmpbsa_gb_energy_noalloc(mmpbsa_ctx *ctx, const double *d_xyz, const mmpbsa_gb_params *p,
                             double *d_born_radii_out, double *d_total_out)
{
    if (!ctx || !p || !d_xyz || !d_total_out)
        return fail(ctx, MMPBSA_ERR_INVALID_ARG, "null argument");
    return guarded(ctx, [&] {
        const PairTopo v = ctx->top.view();
        const size_t need = gb_noalloc_bytes(v.n, d_born_radii_out == nullptr);
        if (need > kNoallocScratchBytes)
            throw DeviceError(MMPBSA_ERR_TOO_LARGE,
                              "GB energy of " + std::to_string(v.n) + " atoms needs " +
                                  std::to_string(need) + " bytes of scratch, above the " +
                                  std::to_string(kNoallocScratchBytes) + "-byte maximum");
        unsigned char *base = scratch_on(ctx->device);
        Worker &w = *ctx->w;
        if (v.n == 0) {
            launch_fill_f64(d_total_out, 0.0, 1, w.stream);
            w.sync();
            return;
        }
        size_t off = 0;
        double *psi = (double *)(base + off);
        off += align256(sizeof(double) * ((size_t)gb_chunks(v.n) * v.n + 1));
        double *born = d_born_radii_out;
        if (!born) {
            born = (double *)(base + off);
            off += align256(sizeof(double) * ((size_t)v.n + 1));
        }
        double *part = (double *)(base + off);
        // the same launches as mmpbsa_gb_energy (run_gb): every scratch slot
        // read is written first in this call
        launch_gb_radii(d_xyz, v, p->model, p->offset, p->rgbmax, p->cutoff, psi, born, w.stream);
        const double kappa = compute_kappa(p->salt_concentration, p->solvent_dielectric,
                                           p->temperature);
        launch_gb_energy(d_xyz, v, born, kappa, p->solute_dielectric, p->solvent_dielectric,
                         p->cutoff, part, w.stream);
        launch_sum_partials(part, gb_energy_blocks(v.n), 1, 1, d_total_out, w.stream);
        w.sync();
    });
}

}  // extern "C"
