// engine.cpp -- per-term drivers: MM, GB, SA and the PB pipeline with its
// multigrid solve.  Everything is enqueued on the worker's stream; the host
// only computes grid geometry (bounding boxes, dimensions) and reads back the
// handful of scalars the reference's control flow branches on.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "engine.h"

namespace mmpbsa {

double compute_kappa(double salt, double eps, double temperature)
{
    if (salt <= 0.0) return 0.0;
    const double kb = 0.00198688, na = 6.022e-4;
    const double factor = 8.0 * kPi * kCoulomb * na;
    return std::sqrt(factor * salt / (eps * kb * temperature));
}

// ---- MM ----------------------------------------------------------------------
void run_mm(Worker &w, const DevTopo &t, const double *d_xyz, mmpbsa_mm_terms *out)
{
    const PairTopo v = t.view();
    const int nbb = bonded_blocks(v);
    const int nmm = mm_blocks(v.n);
    DBuf<double> pb(w.pool, (size_t)nbb * 3), pm(w.pool, (size_t)nmm * 8), res(w.pool, 8);
    launch_bonded(d_xyz, v, pb.get(), w.stream);
    launch_sum_partials(pb.get(), nbb, 3, 3, res.get(), w.stream);
    if (v.n >= 2) {
        launch_mm_pairs(d_xyz, v, pm.get(), w.stream);
        launch_sum_partials_dd(pm.get(), nmm, 4, res.get() + 3, w.stream);
    } else {
        launch_fill_f64(res.get() + 3, 0.0, 4, w.stream);
    }
    HIP_OK(hipMemcpyAsync(w.pin, res.get(), sizeof(double) * 7, hipMemcpyDeviceToHost, w.stream));
    w.sync();
    out->bond = w.pin[0];
    out->angle = w.pin[1];
    out->dihedral = w.pin[2];
    out->vdw = w.pin[3];
    out->elec = w.pin[4];
    out->vdw_14 = w.pin[5];
    out->elec_14 = w.pin[6];
}

// ---- GB ----------------------------------------------------------------------
double run_gb(Worker &w, const DevTopo &t, const double *d_xyz, const mmpbsa_gb_params &p,
              double *d_born, std::vector<double> *born_out)
{
    const PairTopo v = t.view();
    if (v.n == 0) {
        if (born_out) born_out->clear();
        return 0.0;
    }
    const int chunks = gb_chunks(v.n);
    DBuf<double> psi(w.pool, (size_t)chunks * v.n + 1);
    DBuf<double> born_tmp;
    if (!d_born) {
        born_tmp = DBuf<double>(w.pool, v.n);
        d_born = born_tmp.get();
    }
    const int nbk = gb_energy_blocks(v.n);
    DBuf<double> part(w.pool, nbk), res(w.pool, 1);
    launch_gb_radii(d_xyz, v, p.model, p.offset, p.rgbmax, p.cutoff, psi.get(), d_born, w.stream);
    const double kappa = compute_kappa(p.salt_concentration, p.solvent_dielectric, p.temperature);
    launch_gb_energy(d_xyz, v, d_born, kappa, p.solute_dielectric, p.solvent_dielectric, p.cutoff,
                     part.get(), w.stream);
    launch_sum_partials(part.get(), nbk, 1, 1, res.get(), w.stream);
    HIP_OK(hipMemcpyAsync(w.pin, res.get(), sizeof(double), hipMemcpyDeviceToHost, w.stream));
    if (born_out) {
        born_out->resize(v.n);
        HIP_OK(hipMemcpyAsync(born_out->data(), d_born, sizeof(double) * v.n,
                              hipMemcpyDeviceToHost, w.stream));
    }
    w.sync();
    return w.pin[0];
}

// ---- geometry helpers (host, FP64 exactly as the reference) -----------------
static void bbox(int n, const double *xyz, double *mn, double *mx)
{
    for (int k = 0; k < 3; ++k) {
        mn[k] = DBL_MAX;
        mx[k] = -DBL_MAX;
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            if (xyz[3 * i + k] < mn[k]) mn[k] = xyz[3 * i + k];
            if (xyz[3 * i + k] > mx[k]) mx[k] = xyz[3 * i + k];
        }
}

// auto_grid (pb_grid.rs:188-215); the three spacings are equal.
static GridDesc auto_grid(const double *mn, const double *mx, double h, double buffer)
{
    GridDesc g;
    int dims[3];
    double o[3];
    for (int k = 0; k < 3; ++k) {
        const double ext = mx[k] - mn[k] + 2.0 * buffer;
        int64_t m = (int64_t)std::ceil(ext / h) + 1;
        if (m % 2 == 0) m += 1;
        dims[k] = (int)m;
        const double gext = (double)(m - 1) * h;
        const double c = 0.5 * (mn[k] + mx[k]);
        o[k] = c - 0.5 * gext;
    }
    g.nx = dims[0];
    g.ny = dims[1];
    g.nz = dims[2];
    g.ox = o[0];
    g.oy = o[1];
    g.oz = o[2];
    g.h = h;
    return g;
}

// build_coarse_grid (pb_energy.rs:92-125)
static GridDesc coarse_grid(const double *mn, const double *mx, double h, double fillratio,
                            int64_t fscale)
{
    const double cs = h * (double)fscale;
    GridDesc g;
    int dims[3];
    double o[3];
    for (int k = 0; k < 3; ++k) {
        const double ext = (mx[k] - mn[k]) * fillratio;
        int64_t m = (int64_t)std::ceil(ext / cs) + 1;
        if (m % 2 == 0) m += 1;
        dims[k] = (int)m;
        const double aext = (double)(m - 1) * cs;
        const double c = 0.5 * (mn[k] + mx[k]);
        o[k] = c - 0.5 * aext;
    }
    g.nx = dims[0];
    g.ny = dims[1];
    g.nz = dims[2];
    g.ox = o[0];
    g.oy = o[1];
    g.oz = o[2];
    g.h = cs;
    return g;
}

struct DevCells {
    DBuf<int> cell_of, start, fill, atoms;
    CellList cl;
};

static void build_cells(Worker &w, const double *d_xyz, int n, const double *mn, const double *mx,
                        double cs, DevCells &c)
{
    cs = std::max(cs, 0.5);
    const double inv = 1.0 / cs;
    int d[3];
    double o[3];
    for (int k = 0; k < 3; ++k) {
        o[k] = n > 0 ? mn[k] : 0.0;
        const double span = n > 0 ? mx[k] - mn[k] : 0.0;
        d[k] = std::max(1, std::min(1 << 10, (int)std::floor(span * inv) + 1));
    }
    const int64_t nc = (int64_t)d[0] * d[1] * d[2];
    if (nc > (1LL << 26)) throw DeviceError(MMPBSA_ERR_TOO_LARGE, "cell list too large");
    c.cell_of = DBuf<int>(w.pool, n + 1);
    c.start = DBuf<int>(w.pool, nc + 1);
    c.fill = DBuf<int>(w.pool, nc);
    c.atoms = DBuf<int>(w.pool, n + 1);
    launch_cell_build(d_xyz, n, o[0], o[1], o[2], inv, d[0], d[1], d[2], c.cell_of.get(),
                      c.start.get(), c.fill.get(), c.atoms.get(), w.stream);
    c.cl.ox = o[0];
    c.cl.oy = o[1];
    c.cl.oz = o[2];
    c.cl.inv = inv;
    c.cl.dx = d[0];
    c.cl.dy = d[1];
    c.cl.dz = d[2];
    c.cl.start = c.start.get();
    c.cl.atoms = c.atoms.get();
}

// ---- SA ----------------------------------------------------------------------
static std::vector<double> fibonacci_points(int n)
{
    std::vector<double> p(3 * (size_t)std::max(n, 1), 0.0);
    if (n == 1) {
        p[0] = 0.0;
        p[1] = 1.0;
        p[2] = 0.0;
    } else if (n > 1) {
        const double phi = kPi * (3.0 - std::sqrt(5.0));
        for (int i = 0; i < n; ++i) {
            const double y = 1.0 - ((double)i / (double)(n - 1)) * 2.0;
            const double rad = std::sqrt(1.0 - y * y);
            const double th = phi * (double)i;
            p[3 * i] = std::cos(th) * rad;
            p[3 * i + 1] = y;
            p[3 * i + 2] = std::sin(th) * rad;
        }
    }
    return p;
}

void run_sa(Worker &w, const DevTopo &t, const double *d_xyz, const double *h_xyz,
            const mmpbsa_sa_params &p, double *d_per_atom_out, std::vector<double> *per_atom,
            double *total_sasa, double *total)
{
    const int n = t.h.n;
    if (p.n_sphere_points <= 0) throw DeviceError(MMPBSA_ERR_INVALID_ARG, "n_sphere_points <= 0");
    if (n == 0) {
        if (per_atom) per_atom->clear();
        *total_sasa = 0.0;
        *total = p.offset;
        return;
    }
    const double max_ext = t.max_radius + p.probe_radius;
    double mn[3], mx[3];
    bbox(n, h_xyz, mn, mx);
    DevCells c;
    build_cells(w, d_xyz, n, mn, mx, max_ext, c);
    const std::vector<double> pts = fibonacci_points(p.n_sphere_points);
    DBuf<double> dpts(w.pool, pts.size());
    HIP_OK(hipMemcpyAsync(dpts.get(), pts.data(), sizeof(double) * pts.size(),
                          hipMemcpyHostToDevice, w.stream));
    DBuf<double> own;
    double *d_per = d_per_atom_out;
    if (!d_per) {
        own = DBuf<double>(w.pool, n);
        d_per = own.get();
    }
    DBuf<int> ovf(w.pool, 1);
    DBuf<double> res(w.pool, 1);
    launch_fill_i32(ovf.get(), 0, 1, w.stream);
    launch_sasa(d_xyz, t.radii.p, n, p.probe_radius, dpts.get(), p.n_sphere_points, c.cl, max_ext,
                d_per, ovf.get(), w.stream);
    launch_sum_partials(d_per, n, 1, 1, res.get(), w.stream);
    HIP_OK(hipMemcpyAsync(w.pin, res.get(), sizeof(double), hipMemcpyDeviceToHost, w.stream));
    HIP_OK(hipMemcpyAsync(w.pin_i, ovf.get(), sizeof(int), hipMemcpyDeviceToHost, w.stream));
    if (per_atom) {
        per_atom->resize(n);
        HIP_OK(hipMemcpyAsync(per_atom->data(), d_per, sizeof(double) * n, hipMemcpyDeviceToHost,
                              w.stream));
    }
    w.sync();
    if (w.pin_i[0]) throw DeviceError(MMPBSA_ERR_TOO_LARGE, "SASA neighbour list overflow");
    *total_sasa = w.pin[0];
    *total = p.surface_tension * w.pin[0] + p.offset;
}

// ---- PB ----------------------------------------------------------------------
static const double kFourPiEc = 4.0 * kPi * kCoulomb;

struct MapSet {
    bool uni = false;      // uniform dielectric and zero kappa
    double eps_u = 1.0;
    DBuf<double> ex, ey, ez, k2;
    DBuf<float> fx, fy, fz, k2f;
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

// assign_dielectrics + assign_kappa for one grid (pb_grid.rs:827-985)
static void build_maps(Worker &w, const DevTopo &t, const double *d_xyz, const DevCells &c,
                       const GridDesc &g, const mmpbsa_pb_params &p, double ein, double eout,
                       bool ionic, double kappa, MapSet &m)
{
    const int64_t N = g.size();
    const bool uni_eps = std::fabs(ein - eout) < DBL_EPSILON;
    if (uni_eps && !ionic) {
        m.uni = true;
        m.eps_u = ein;
        return;
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
        launch_fill_f64(m.ex.get(), ein, N, w.stream);
        launch_fill_f64(m.ey.get(), ein, N, w.stream);
        launch_fill_f64(m.ez.get(), ein, N, w.stream);
        launch_fill_f32(m.fx.get(), (float)ein, N, w.stream);
        launch_fill_f32(m.fy.get(), (float)ein, N, w.stream);
        launch_fill_f32(m.fz.get(), (float)ein, N, w.stream);
    } else {
        DBuf<uint8_t> sas(w.pool, N), inside(w.pool, N);
        const double probe = p.probe_radius;
        if (probe <= 0.0) {
            launch_inside(g, d_xyz, t.radii.p, 0.0, t.max_radius, c.cl, inside.get(), w.stream);
        } else {
            const double th2 = (probe / g.h) * (probe / g.h);
            const double T = std::floor(th2);
            if (T > 250.0)
                throw DeviceError(MMPBSA_ERR_TOO_LARGE, "probe_radius / grid_spacing above 15.8");
            launch_inside(g, d_xyz, t.radii.p, probe, t.max_radius, c.cl, sas.get(), w.stream);
            DBuf<uint8_t> ext(w.pool, N), wa(w.pool, N), wb(w.pool, N);
            DBuf<int> flag(w.pool, 1);
            launch_ses(g, sas.get(), ext.get(), wa.get(), wb.get(), (int)T, inside.get(),
                       flag.get(), w.pin_i + 8, w.stream);
        }
        DBuf<int> ovf(w.pool, 1);
        launch_fill_i32(ovf.get(), 0, 1, w.stream);
        launch_faces(g, inside.get(), d_xyz, t.radii.p, t.max_radius, c.cl, ein, eout, m.ex.get(),
                     m.ey.get(), m.ez.get(), m.fx.get(), m.fy.get(), m.fz.get(), ovf.get(),
                     w.stream);
        HIP_OK(hipMemcpyAsync(w.pin_i + 9, ovf.get(), sizeof(int), hipMemcpyDeviceToHost, w.stream));
        w.sync();
        if (w.pin_i[9]) throw DeviceError(MMPBSA_ERR_TOO_LARGE, "face interval list overflow");
    }
    if (ionic) {
        DBuf<uint8_t> ion(w.pool, N);
        launch_inside(g, d_xyz, t.radii.p, p.ion_radius, t.max_radius, c.cl, ion.get(), w.stream);
        const double kb2 = p.solvent_dielectric * kappa * kappa;
        launch_kappa_map(g, ion.get(), kb2, m.k2.get(), m.k2f.get(), w.stream);
    } else {
        launch_fill_f64(m.k2.get(), 0.0, N, w.stream);
        launch_fill_f32(m.k2f.get(), 0.0f, N, w.stream);
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
    O.rho = rho;
    return O;
}

__global__ void status_init_kernel(MgStatus *st)
{
    st->state = 0;
    st->cycles = 0;
    st->growing = 0;
    st->pad = 0;
    st->best = DBL_MAX;
    st->prev = DBL_MAX;
    st->rms = 0.0;
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

static void read_status(Worker &w, MgStatus *d_st, MgStatus &h)
{
    HIP_OK(hipMemcpyAsync(w.pin, d_st, sizeof(MgStatus), hipMemcpyDeviceToHost, w.stream));
    w.sync();
    std::memcpy(&h, w.pin, sizeof(MgStatus));
}

// solve_lpbe_initial (pb_solver.rs:373-447): Chebyshev-ramped SOR from phi.
static SolveInfo sor_solve(Worker &w, const GridDesc &g, const MapSet &m, const double *rho,
                           double *phi, double tol, int64_t max_it)
{
    SolveInfo info;
    MgOuterDev O = outer_view(g, m, rho);
    O.phi_out = phi;
    const int nbk = sor64_blocks(O);
    DBuf<double> part(w.pool, 2 * (size_t)nbk);
    DBuf<MgStatus> st(w.pool, 1);
    hipLaunchKernelGGL(status_init_kernel, dim3(1), dim3(1), 0, w.stream, st.get());
    LAUNCH_OK();
    const double nint = (g.nx > 2 && g.ny > 2 && g.nz > 2)
                            ? (double)(g.nx - 2) * (g.ny - 2) * (g.nz - 2)
                            : 0.0;
    const double rj = spectral_radius(g.nx, g.ny, g.nz);
    const double r2 = rj * rj;
    double omega = 1.0;
    int64_t k = 0;
    MgStatus hs{};
    info.final_residual = DBL_MAX;
    if (max_it <= 0) return info;
    while (k < max_it) {
        const int64_t chunk_end = std::min<int64_t>(max_it, k + 64);
        for (; k < chunk_end; ++k) {
            mg_sor64(O, 0, omega, part.get(), st.get(), w.stream);
            mg_sor64(O, 1, omega, part.get() + nbk, st.get(), w.stream);
            mg_sor_check(part.get(), 2 * nbk, nint, tol, (int)(k + 1), (int)max_it, st.get(),
                         w.stream);
            omega = (k == 0) ? 1.0 / (1.0 - r2 / 2.0) : 1.0 / (1.0 - r2 * omega / 4.0);
        }
        read_status(w, st.get(), hs);
        if (hs.state != 0) break;
    }
    info.iterations = hs.cycles;
    info.converged = hs.state == 1;
    info.final_residual = hs.rms;
    return info;
}

struct Level {
    GridDesc g;
    DBuf<float> ex, ey, ez, k2, e, f, res;
    MgLevelDev d;
    float omega;
};

static void enqueue_vcycle(Worker &w, std::vector<Level> &L, size_t l, const MgStatus *st)
{
    Level &lv = L[l];
    if (l + 1 == L.size()) {
        for (int s = 0; s < 50; ++s) {
            mg_rb_sweep(lv.d, 0, lv.omega, st, w.stream);
            mg_rb_sweep(lv.d, 1, lv.omega, st, w.stream);
        }
        return;
    }
    for (int s = 0; s < 3; ++s) {
        mg_rb_sweep(lv.d, 0, lv.omega, st, w.stream);
        mg_rb_sweep(lv.d, 1, lv.omega, st, w.stream);
    }
    mg_residual32(lv.d, st, w.stream);
    mg_restrict(lv.d, L[l + 1].d, st, w.stream);
    enqueue_vcycle(w, L, l + 1, st);
    mg_prolong(L[l + 1].d, lv.d, st, w.stream);
    for (int s = 0; s < 3; ++s) {
        mg_rb_sweep(lv.d, 0, lv.omega, st, w.stream);
        mg_rb_sweep(lv.d, 1, lv.omega, st, w.stream);
    }
}

// solve_lpbe_multigrid (pb_solver.rs:1015-1167).  phi holds the boundary
// values (interior zero) on entry and the solution on exit.
static SolveInfo mg_solve(Worker &w, const GridDesc &g, const MapSet &m, const double *rho,
                          double *phi, double tol, int64_t max_it)
{
    const int md = std::min(g.nx, std::min(g.ny, g.nz));
    if (md <= 8) return sor_solve(w, g, m, rho, phi, tol, max_it);
    const int maxv = (int)std::min<int64_t>(200, std::max<int64_t>(1, max_it / 6));
    // levels (build_multigrid_levels, pb_solver.rs:776-836; min size 5)
    std::vector<Level> L;
    {
        int d[3] = {g.nx, g.ny, g.nz};
        double h = g.h;
        for (;;) {
            Level lv;
            lv.g = g;
            lv.g.nx = d[0];
            lv.g.ny = d[1];
            lv.g.nz = d[2];
            lv.g.h = h;
            L.push_back(std::move(lv));
            if (std::min(d[0], std::min(d[1], d[2])) <= 5) break;
            for (int k = 0; k < 3; ++k) d[k] = (d[k] + 1) / 2;
            h *= 2.0;
        }
    }
    for (size_t l = 0; l < L.size(); ++l) {
        Level &lv = L[l];
        const int64_t N = lv.g.size();
        lv.e = DBuf<float>(w.pool, N);
        lv.f = DBuf<float>(w.pool, N);
        if (l + 1 < L.size()) lv.res = DBuf<float>(w.pool, N);
        lv.omega = (float)estimate_omega(lv.g.nx, lv.g.ny, lv.g.nz);
        MgLevelDev &d = lv.d;
        d.nx = lv.g.nx;
        d.ny = lv.g.ny;
        d.nz = lv.g.nz;
        d.uni = m.uni ? 1 : 0;
        d.eps_u = (float)m.eps_u;
        const double ih = 1.0 / (lv.g.h * lv.g.h);
        d.ihx = d.ihy = d.ihz = (float)ih;
        d.e = lv.e.get();
        d.f = lv.f.get();
        d.res = lv.res.get();
        if (!m.uni) {
            if (l == 0) {
                d.ex = m.fx.get();
                d.ey = m.fy.get();
                d.ez = m.fz.get();
                d.k2 = m.k2f.get();
            } else {
                lv.ex = DBuf<float>(w.pool, N);
                lv.ey = DBuf<float>(w.pool, N);
                lv.ez = DBuf<float>(w.pool, N);
                lv.k2 = DBuf<float>(w.pool, N);
                d.ex = lv.ex.get();
                d.ey = lv.ey.get();
                d.ez = lv.ez.get();
                d.k2 = lv.k2.get();
                mg_inject(L[l - 1].d, d, w.stream);
            }
        } else {
            d.ex = d.ey = d.ez = d.k2 = nullptr;
        }
    }
    const int64_t N = g.size();
    DBuf<double> buf1(w.pool, N), bc(w.pool, N);
    HIP_OK(hipMemcpyAsync(bc.get(), phi, sizeof(double) * N, hipMemcpyDeviceToDevice, w.stream));
    double *bufs[2] = {phi, buf1.get()};
    MgOuterDev O = outer_view(g, m, rho);
    O.f0 = L[0].f.get();
    const int nbk = mg_outer_blocks(O);
    DBuf<double> part(w.pool, nbk);
    O.partials = part.get();
    DBuf<MgStatus> st(w.pool, 1);
    hipLaunchKernelGGL(status_init_kernel, dim3(1), dim3(1), 0, w.stream, st.get());
    LAUNCH_OK();
    const double nint = (double)(g.nx - 2) * (g.ny - 2) * (g.nz - 2);
    // cycle 0: residual of the boundary-only potential
    O.phi_in = phi;
    O.phi_out = phi;
    O.e = nullptr;
    mg_outer64(O, st.get(), w.stream);
    mg_check(part.get(), nbk, nint, tol, 0, maxv, st.get(), w.stream);
    MgStatus hs{};
    int c = 1;
    const int chunk = 8;
    while (c <= maxv) {
        const int end = std::min(maxv, c + chunk - 1);
        for (; c <= end; ++c) {
            mg_zero(L[0].e.get(), N, st.get(), w.stream);
            enqueue_vcycle(w, L, 0, st.get());
            O.phi_in = bufs[(c - 1) & 1];
            O.phi_out = bufs[c & 1];
            O.e = L[0].e.get();
            mg_outer64(O, st.get(), w.stream);
            mg_check(part.get(), nbk, nint, tol, c, maxv, st.get(), w.stream);
        }
        read_status(w, st.get(), hs);
        if (hs.state != 0) break;
    }
    SolveInfo info;
    const int fin = hs.cycles;
    if (fin & 1)
        HIP_OK(hipMemcpyAsync(phi, bufs[1], sizeof(double) * N, hipMemcpyDeviceToDevice, w.stream));
    if (hs.state == 2) {
        // divergence: Chebyshev SOR from the partial potential, BC re-applied
        hipLaunchKernelGGL(copy_boundary_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, w.stream, g,
                           bc.get(), phi);
        LAUNCH_OK();
        return sor_solve(w, g, m, rho, phi, tol, max_it);
    }
    info.iterations = hs.state == 1 ? fin : maxv;
    info.converged = hs.state == 1;
    info.final_residual = hs.rms;
    return info;
}

static double energy_of(Worker &w, const GridDesc &g, const double *phi, int n,
                        const double *d_xyz, const double *d_q)
{
    DBuf<double> part(w.pool, std::max(1, ceil_div(n, 256))), res(w.pool, 1);
    int nbk = 0;
    launch_energy_gather(g, phi, n, d_xyz, d_q, part.get(), &nbk, w.stream);
    launch_sum_partials(part.get(), nbk, 1, 1, res.get(), w.stream);
    HIP_OK(hipMemcpyAsync(w.pin, res.get(), sizeof(double), hipMemcpyDeviceToHost, w.stream));
    w.sync();
    return 0.5 * w.pin[0];
}

// pb_solve_with_focusing (pb_energy.rs:134-322)
static void solve_focus(Worker &w, const DevTopo &t, const double *d_xyz, const DevCells &cells,
                        const GridDesc &fine, const double *fine_rho, const GridDesc *cg,
                        const double *coarse_rho, const mmpbsa_pb_params &p, double ein,
                        double eout, double kappa, bool ionic, mmpbsa_pb_solve_info *info)
{
    const int n = t.h.n;
    const double bck = ionic ? kappa : 0.0;
    const int64_t NF = fine.size();
    DBuf<double> phi(w.pool, NF);
    launch_fill_f64(phi.get(), 0.0, NF, w.stream);
    info->focused = 0;
    info->coarse_iterations = 0;
    bool have_bc = false;
    if (cg) {
        const int64_t NC = cg->size();
        MapSet cm;
        build_maps(w, t, d_xyz, cells, *cg, p, ein, eout, ionic, kappa, cm);
        DBuf<double> cphi(w.pool, NC);
        launch_fill_f64(cphi.get(), 0.0, NC, w.stream);
        launch_dh_boundary(*cg, n, d_xyz, t.q.p, bck, eout, cphi.get(), w.stream);
        const SolveInfo cs = mg_solve(w, *cg, cm, coarse_rho, cphi.get(), p.tolerance * 10.0,
                                      p.max_iterations);
        info->coarse_iterations = cs.iterations;
        bool ok = cs.converged != 0;
        if (ok) {
            DBuf<double> part(w.pool, 1024), res(w.pool, 1);
            int nbk = 0;
            launch_maxabs(cphi.get(), NC, part.get(), &nbk, w.stream);
            launch_max_partials(part.get(), nbk, res.get(), w.stream);
            HIP_OK(hipMemcpyAsync(w.pin, res.get(), sizeof(double), hipMemcpyDeviceToHost, w.stream));
            w.sync();
            ok = std::isfinite(w.pin[0]) && w.pin[0] < 1e10;
        }
        if (ok) {
            launch_interp_boundary(fine, *cg, cphi.get(), phi.get(), w.stream);
            have_bc = true;
            info->focused = 1;
        }
    }
    if (!have_bc) launch_dh_boundary(fine, n, d_xyz, t.q.p, bck, eout, phi.get(), w.stream);
    MapSet fm;
    build_maps(w, t, d_xyz, cells, fine, p, ein, eout, ionic, kappa, fm);
    const SolveInfo fs = mg_solve(w, fine, fm, fine_rho, phi.get(), p.tolerance, p.max_iterations);
    info->energy = energy_of(w, fine, phi.get(), n, d_xyz, t.q.p);
    info->iterations = fs.iterations;
    info->converged = fs.converged;
    info->final_residual = fs.final_residual;
}

// compute_pb_energy (pb_energy.rs:332-438)
void run_pb(Worker &w, const DevTopo &t, const double *d_xyz, const double *h_xyz,
            const mmpbsa_pb_params &p, mmpbsa_pb_result *out)
{
    std::memset(out, 0, sizeof(*out));
    const int n = t.h.n;
    if (!(p.grid_spacing > 0.0)) throw DeviceError(MMPBSA_ERR_INVALID_ARG, "grid_spacing <= 0");
    if (n == 0) return;
    const double kappa = compute_kappa(p.salt_concentration, p.solvent_dielectric, p.temperature);
    double mn[3], mx[3];
    bbox(n, h_xyz, mn, mx);
    GridDesc fine = auto_grid(mn, mx, p.grid_spacing, p.grid_buffer);
    if (p.fine_dims[0] > 0) {
        fine.nx = p.fine_dims[0];
        fine.ny = p.fine_dims[1] > 0 ? p.fine_dims[1] : p.fine_dims[0];
        fine.nz = p.fine_dims[2] > 0 ? p.fine_dims[2] : p.fine_dims[0];
        const int d[3] = {fine.nx, fine.ny, fine.nz};
        double o[3];
        for (int k = 0; k < 3; ++k)
            o[k] = 0.5 * (mn[k] + mx[k]) - 0.5 * ((double)(d[k] - 1) * p.grid_spacing);
        fine.ox = o[0];
        fine.oy = o[1];
        fine.oz = o[2];
    }
    const bool focus = p.fillratio > 0.0 && p.fscale > 1;
    GridDesc cg{};
    if (focus) cg = coarse_grid(mn, mx, p.grid_spacing, p.fillratio, p.fscale);
    if (fine.size() > (int64_t)1 << 31 || (focus && cg.size() > (int64_t)1 << 31))
        throw DeviceError(MMPBSA_ERR_TOO_LARGE, "PB grid above 2^31 nodes");
    out->fine_dims[0] = fine.nx;
    out->fine_dims[1] = fine.ny;
    out->fine_dims[2] = fine.nz;
    if (focus) {
        out->coarse_dims[0] = cg.nx;
        out->coarse_dims[1] = cg.ny;
        out->coarse_dims[2] = cg.nz;
    }
    DevCells cells;
    build_cells(w, d_xyz, n, mn, mx,
                t.max_radius + std::max(p.probe_radius, std::max(p.ion_radius, 0.0)), cells);
    DBuf<unsigned long long> acc(w.pool, fine.size());
    DBuf<double> rho_f(w.pool, fine.size());
    launch_spread(fine, n, d_xyz, t.q.p, acc.get(), rho_f.get(), w.stream);
    DBuf<double> rho_c;
    if (focus) {
        DBuf<unsigned long long> acc_c(w.pool, cg.size());
        rho_c = DBuf<double>(w.pool, cg.size());
        launch_spread(cg, n, d_xyz, t.q.p, acc_c.get(), rho_c.get(), w.stream);
    }
    solve_focus(w, t, d_xyz, cells, fine, rho_f.get(), focus ? &cg : nullptr, rho_c.get(), p,
                p.solute_dielectric, p.solvent_dielectric, kappa, true, &out->solv);
    solve_focus(w, t, d_xyz, cells, fine, rho_f.get(), focus ? &cg : nullptr, rho_c.get(), p,
                p.solute_dielectric, p.solute_dielectric, kappa, false, &out->ref);
    out->total = out->solv.energy - out->ref.energy;
}

}  // namespace mmpbsa
