// engine.cpp -- per-term drivers: MM, GB, SA, and the grid geometry shared
// with the PB pipeline (pb.cpp).  Everything is enqueued on the worker's stream; the host
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
        {
            // 20 FLOP per unordered pair (Coulomb + LJ + exclusion and 1-4 rules)
            KernelTimer kt(w, "mm_pair_kernel", 10.0 * (double)v.n * (v.n - 1));
            launch_mm_pairs(d_xyz, v, pm.get(), w.stream);
        }
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

// MM terms of the complex, receptor and ligand of one frame (out[0..2]) when
// the complex is exactly the two partners and no excluded, 1-4 or bonded term
// spans them: the complex's pair terms are the partners' own pair terms plus
// the receptor x ligand pairs, so its receptor-receptor and ligand-ligand
// pairs are evaluated once, not twice.  Every pair term is the one the
// complex's own pass computes; the complex's sums run over the partners' and
// the cross blocks (double-double, then rounded once), its bonded sums over
// the partners' bonded blocks.
void run_mm_parts(Worker &w, const DevTopo &ct, const DevTopo &rt, const DevTopo &lt,
                  const double *dc, const double *dr, const double *dl, const int *d_ra, int nr,
                  const int *d_la, int nl, mmpbsa_mm_terms out[3])
{
    const PairTopo vc = ct.view(), vr = rt.view(), vl = lt.view();
    const int nbr = bonded_blocks(vr), nbl = bonded_blocks(vl);
    const int nmr = vr.n >= 2 ? mm_blocks(vr.n) : 0, nml = vl.n >= 2 ? mm_blocks(vl.n) : 0;
    const int ncx = mm_cross_blocks(nr, nl);
    DBuf<double> pb(w.pool, (size_t)(nbr + nbl) * 3), pm(w.pool, (size_t)(nmr + nml + ncx) * 8 + 8),
        res(w.pool, 21);
    launch_bonded(dr, vr, pb.get(), w.stream);
    launch_bonded(dl, vl, pb.get() + (size_t)nbr * 3, w.stream);
    {
        KernelTimer kt(w, "mm_pair_kernel",
                       10.0 * ((double)vr.n * (vr.n - 1) + (double)vl.n * (vl.n - 1)));
        if (nmr) launch_mm_pairs(dr, vr, pm.get(), w.stream);
        if (nml) launch_mm_pairs(dl, vl, pm.get() + (size_t)nmr * 8, w.stream);
    }
    {
        KernelTimer kt(w, "mm_cross_kernel", 20.0 * (double)nr * nl);
        launch_mm_cross(dc, vc, d_ra, nr, d_la, nl, pm.get() + (size_t)(nmr + nml) * 8, w.stream);
    }
    // complex: every block; receptor: the first ones; ligand: the next ones
    launch_sum_partials(pb.get(), nbr + nbl, 3, 3, res.get(), w.stream);
    launch_sum_partials_dd(pm.get(), nmr + nml + ncx, 4, res.get() + 3, w.stream);
    launch_sum_partials(pb.get(), nbr, 3, 3, res.get() + 7, w.stream);
    launch_sum_partials_dd(pm.get(), nmr, 4, res.get() + 10, w.stream);
    launch_sum_partials(pb.get() + (size_t)nbr * 3, nbl, 3, 3, res.get() + 14, w.stream);
    launch_sum_partials_dd(pm.get() + (size_t)nmr * 8, nml, 4, res.get() + 17, w.stream);
    HIP_OK(hipMemcpyAsync(w.pin, res.get(), sizeof(double) * 21, hipMemcpyDeviceToHost, w.stream));
    w.sync();
    for (int k = 0; k < 3; ++k) {
        const double *v = w.pin + 7 * k;
        out[k].bond = v[0];
        out[k].angle = v[1];
        out[k].dihedral = v[2];
        out[k].vdw = v[3];
        out[k].elec = v[4];
        out[k].vdw_14 = v[5];
        out[k].elec_14 = v[6];
    }
}

// ---- GB ----------------------------------------------------------------------
// GB energies of the complex, receptor and ligand of one frame (out[0..2]) when
// the complex is exactly the two partners: each partner's psi sums (over its
// own atoms) serve both its own Born radii and, plus the sums over the other
// partner's atoms, the complex's, so the complex's intra-partner pairs are
// not evaluated a second time.  Every pair term is the one the complex's own
// pass computes; a complex atom's psi is its partner sum plus its cross sum
// (the order of the additions differs from the complex's own pass, so radii
// agree to rounding).
void run_gb_parts(Worker &w, const DevTopo &ct, const DevTopo &rt, const DevTopo &lt,
                  const double *dc, const double *dr, const double *dl, const int *d_ra, int nr,
                  const int *d_la, int nl, const mmpbsa_gb_params &p, double out[3])
{
    const PairTopo vc = ct.view(), vr = rt.view(), vl = lt.view();
    const int cr = gb_psi_chunks(nr), cl = gb_psi_chunks(nl);
    const int xr = gb_psi_cross_chunks(nr, nl), xl = gb_psi_cross_chunks(nl, nr);
    DBuf<double> psi_r(w.pool, (size_t)cr * nr + 1), psi_l(w.pool, (size_t)cl * nl + 1);
    DBuf<double> x_r(w.pool, (size_t)xr * nr + 1), x_l(w.pool, (size_t)xl * nl + 1);
    DBuf<double> born(w.pool, (size_t)vc.n + nr + nl + 1);
    double *born_c = born.get(), *born_r = born_c + vc.n, *born_l = born_r + nr;
    {
        // 32 FLOP per ordered pair, as run_gb
        KernelTimer kt(w, "gb_radii", 32.0 * ((double)nr * (nr - 1) + (double)nl * (nl - 1)));
        launch_gb_radii(dr, vr, p.model, p.offset, p.rgbmax, p.cutoff, psi_r.get(), born_r,
                        w.stream);
        launch_gb_radii(dl, vl, p.model, p.offset, p.rgbmax, p.cutoff, psi_l.get(), born_l,
                        w.stream);
    }
    {
        KernelTimer kt(w, "gb_radii_cross", 64.0 * (double)nr * nl);
        launch_gb_psi_cross(dc, vc, d_ra, nr, d_la, nl, p.offset, p.cutoff, x_r.get(), w.stream);
        launch_gb_psi_cross(dc, vc, d_la, nl, d_ra, nr, p.offset, p.cutoff, x_l.get(), w.stream);
    }
    launch_gb_finalize_parts(psi_r.get(), cr, x_r.get(), xr, d_ra, nr, vc, p.model, p.offset,
                             p.rgbmax, born_c, w.stream);
    launch_gb_finalize_parts(psi_l.get(), cl, x_l.get(), xl, d_la, nl, vc, p.model, p.offset,
                             p.rgbmax, born_c, w.stream);
    const double kappa = compute_kappa(p.salt_concentration, p.solvent_dielectric, p.temperature);
    const PairTopo *views[3] = {&vc, &vr, &vl};
    const double *xyz[3] = {dc, dr, dl};
    const double *borns[3] = {born_c, born_r, born_l};
    int nbk[3];
    for (int k = 0; k < 3; ++k) nbk[k] = views[k]->n > 0 ? gb_energy_blocks(views[k]->n) : 0;
    DBuf<double> part(w.pool, (size_t)nbk[0] + nbk[1] + nbk[2] + 1), res(w.pool, 3);
    double *pk = part.get();
    for (int k = 0; k < 3; ++k) {
        const PairTopo &v = *views[k];
        if (v.n == 0) {
            launch_fill_f64(res.get() + k, 0.0, 1, w.stream);
            continue;
        }
        {
            KernelTimer kt(w, "gb_energy_kernel", 13.0 * (double)v.n * (v.n - 1));
            launch_gb_energy(xyz[k], v, borns[k], kappa, p.solute_dielectric,
                             p.solvent_dielectric, p.cutoff, pk, w.stream);
        }
        launch_sum_partials(pk, nbk[k], 1, 1, res.get() + k, w.stream);
        pk += nbk[k];
    }
    HIP_OK(hipMemcpyAsync(w.pin, res.get(), sizeof(double) * 3, hipMemcpyDeviceToHost, w.stream));
    w.sync();
    for (int k = 0; k < 3; ++k) out[k] = w.pin[k];
}

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
    {
        // 32 FLOP per ordered pair (the HCT integral); the O(N) radius map
        // launched with it is not counted
        KernelTimer kt(w, "gb_radii", 32.0 * (double)v.n * (v.n - 1));
        launch_gb_radii(d_xyz, v, p.model, p.offset, p.rgbmax, p.cutoff, psi.get(), d_born,
                        w.stream);
    }
    const double kappa = compute_kappa(p.salt_concentration, p.solvent_dielectric, p.temperature);
    {
        // 26 FLOP per unordered pair
        KernelTimer kt(w, "gb_energy_kernel", 13.0 * (double)v.n * (v.n - 1));
        launch_gb_energy(d_xyz, v, d_born, kappa, p.solute_dielectric, p.solvent_dielectric,
                         p.cutoff, part.get(), w.stream);
    }
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

// compute_gb_energy_from_radii (gb_energy.rs:283-355): the energy kernel alone
// over given device Born radii.
double run_gb_from_radii(Worker &w, const DevTopo &t, const double *d_xyz,
                         const mmpbsa_gb_params &p, const double *d_born)
{
    const PairTopo v = t.view();
    if (v.n == 0) return 0.0;
    const int nbk = gb_energy_blocks(v.n);
    DBuf<double> part(w.pool, nbk), res(w.pool, 1);
    const double kappa = compute_kappa(p.salt_concentration, p.solvent_dielectric, p.temperature);
    launch_gb_energy(d_xyz, v, d_born, kappa, p.solute_dielectric, p.solvent_dielectric, p.cutoff,
                     part.get(), w.stream);
    launch_sum_partials(part.get(), nbk, 1, 1, res.get(), w.stream);
    HIP_OK(hipMemcpyAsync(w.pin, res.get(), sizeof(double), hipMemcpyDeviceToHost, w.stream));
    w.sync();
    return w.pin[0];
}

// ---- geometry helpers (host, FP64 exactly as the reference) -----------------
void bbox(int n, const double *xyz, double *mn, double *mx)
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

void device_bbox(Worker &w, int n, const double *d_xyz, double *mn, double *mx)
{
    DBuf<double> d(w.pool, 6);
    launch_bbox(d_xyz, n, d.get(), w.stream);
    HIP_OK(hipMemcpyAsync(w.pin, d.get(), sizeof(double) * 6, hipMemcpyDeviceToHost, w.stream));
    w.sync();
    for (int k = 0; k < 3; ++k) {
        mn[k] = w.pin[k];
        mx[k] = w.pin[3 + k];
    }
}

// auto_grid (pb_grid.rs:188-215); the three spacings are equal.
GridDesc auto_grid(const double *mn, const double *mx, double h, double buffer)
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
GridDesc coarse_grid(const double *mn, const double *mx, double h, double fillratio,
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

void build_cells(Worker &w, const double *d_xyz, int n, const double *mn, const double *mx,
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
    if (h_xyz)
        bbox(n, h_xyz, mn, mx);
    else
        device_bbox(w, n, d_xyz, mn, mx);
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
}  // namespace mmpbsa
