/*
 * mmpbsa_oracle.c -- FP64 CPU restatement of the reference MM-PBSA path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle and the timed CPU
 * baseline.  Never linked into the product library.  Build: oracle/Makefile
 * (gcc -O3 -fopenmp -ffp-contract=off, no fast-math).
 *
 * Citations are to rust-simulation-tools 0.2.4 (crates/mmpbsa/src unless a
 * path says otherwise).  Parallel sites follow the reference's rayon sites;
 * every reduction is made deterministic (partials combined in index order).
 */
#include "mmpbsa_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846
#define COULOMB 332.0522

/* Neumaier-compensated accumulation.  The MM and decomposition pair sums
 * reach 1e12 kcal/mol on near-overlapping synthetic frames; plain FP64
 * accumulation then carries ~1e-2 kcal/mol of order-dependent rounding (the
 * reference's rayon reduction order is not fixed either), so the oracle
 * accumulates those sums compensated to be a tighter reference. */
typedef struct { double s, c; } ksum;
static inline void kadd(ksum *k, double v)
{
    const double t = k->s + v;
    if (fabs(k->s) >= fabs(v))
        k->c += (k->s - t) + v;
    else
        k->c += (v - t) + k->s;
    k->s = t;
}
static inline double kval(const ksum *k) { return k->s + k->c; }

static int g_face_scan = 1;
/* > 0: the multigrid divergence test fires at this cycle (a test hook that
 * mirrors the library's MMPBSA_FORCE_DIVERGENCE_CYCLE) */
static int g_force_div = 0;
void orc_set_force_divergence(int cycle) { g_force_div = cycle > 0 ? cycle : 0; }
void orc_set_face_scan(int mode) { g_face_scan = mode; }
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* gb_energy.rs:126-139 (duplicate at pb_energy.rs:78-86)                     */
double orc_compute_kappa(double salt, double eps, double temperature)
{
    if (salt <= 0.0) return 0.0;
    const double kb = 0.00198688;
    const double na = 6.022e-4;
    const double factor = 8.0 * ORC_PI * COULOMB * na;
    return sqrt(factor * salt / (eps * kb * temperature));
}

/* gb_energy.rs:144-258: HCT descreening sum, then OBC/HCT radius map. */
void orc_gb_born_radii(const mmpbsa_topology_view *t, const double *xyz,
                       const mmpbsa_gb_params *p, double *born)
{
    const int n = t->n_atoms;
    double *rho = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    double *psi = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) rho[i] = t->radii[i] - p->offset;
    const double cut2 = p->cutoff * p->cutoff;

#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
        const double rho_i = rho[i];
        double acc = 0.0;
        for (int j = 0; j < n; j++) {
            if (i == j) continue;
            const double dx = xyz[3 * i] - xyz[3 * j];
            const double dy = xyz[3 * i + 1] - xyz[3 * j + 1];
            const double dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > cut2) continue;
            const double r = sqrt(r2);
            if (r == 0.0) continue;
            const double sj = t->screen[j] * rho[j];
            if (rho_i >= r + sj) continue;
            const double u = r + sj;
            const double rms = fabs(r - sj);
            const double l = (rho_i > rms) ? rho_i : rms;
            const double li = 1.0 / l;
            const double ui = 1.0 / u;
            const double c = 0.5 * (li - ui + 0.25 * (ui * ui - li * li) * (r - sj * sj / r) +
                                    log(l / u) / (2.0 * r));
            if (rho_i < sj - r)
                acc += c + 2.0 * (1.0 / rho_i - li);
            else
                acc += c;
        }
        psi[i] = acc;
    }

    double alpha = 1.0, beta = 0.0, gamma = 0.0; /* :220-224 */
    if (p->model == MMPBSA_GB_OBC1) { alpha = 0.8; beta = 0.0; gamma = 2.909125; }
    if (p->model == MMPBSA_GB_OBC2) { alpha = 1.0; beta = 0.8; gamma = 4.85; }
    for (int i = 0; i < n; i++) {
        const double rho_i = rho[i];
        double inv;
        if (p->model == MMPBSA_GB_HCT) {
            inv = 1.0 / rho_i - psi[i];
        } else {
            const double pr = psi[i] * rho_i;
            const double p2 = pr * pr;
            const double p3 = p2 * pr;
            const double th = tanh(alpha * pr - beta * p2 + gamma * p3);
            inv = 1.0 / rho_i - th / t->radii[i];
        }
        double R = (inv > 0.0) ? 1.0 / inv : p->rgbmax;
        if (R > p->rgbmax) R = p->rgbmax;
        born[i] = R;
    }
    free(rho);
    free(psi);
}

/* gb_energy.rs:283-355 */
double orc_gb_energy_from_radii(const mmpbsa_topology_view *t, const double *xyz,
                                const mmpbsa_gb_params *p, const double *R)
{
    const int n = t->n_atoms;
    const double kappa = orc_compute_kappa(p->salt_concentration, p->solvent_dielectric,
                                           p->temperature);
    const double cut2 = p->cutoff * p->cutoff;
    const double inv_ein = 1.0 / p->solute_dielectric;
    const double inv_eout = 1.0 / p->solvent_dielectric;
    const double KS = 0.73;
    const int no_salt = kappa <= 0.0;
    const double dns = -0.5 * (inv_ein - inv_eout);
    const double *q = t->charges_amber;

    double e = 0.0;
    for (int i = 0; i < n; i++) {
        const double qi = q[i];
        double s;
        if (no_salt)
            s = dns * qi * qi / R[i];
        else
            s = -0.5 * qi * qi * (inv_ein - exp(-KS * kappa * R[i]) * inv_eout) / R[i];
        e += s;
    }
    double *part = (double *)calloc(n > 0 ? n : 1, sizeof(double));
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
        const double qi = q[i];
        double sum = 0.0;
        for (int j = i + 1; j < n; j++) {
            const double qj = q[j];
            const double dx = xyz[3 * i] - xyz[3 * j];
            const double dy = xyz[3 * i + 1] - xyz[3 * j + 1];
            const double dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > cut2) continue;
            const double rr = R[i] * R[j];
            const double f = sqrt(r2 + rr * exp(-r2 / (4.0 * rr)));
            double pe;
            if (no_salt)
                pe = dns * qi * qj / f;
            else
                pe = -0.5 * qi * qj * (inv_ein - exp(-KS * kappa * f) * inv_eout) / f;
            sum += 2.0 * pe;
        }
        part[i] = sum;
    }
    double cross = 0.0;
    for (int i = 0; i < n; i++) cross += part[i];
    free(part);
    return e + cross;
}

/* ------------------------------------------------------------------------ */
/* Pair sets with the semantics of PairBitmap (mm_energy.rs:8-39): membership
 * of (lo, hi), lo < hi.  Stored as CSR rows over lo, sorted and unique.     */
typedef struct {
    int64_t *ptr;
    int32_t *col;
    int32_t n;
} pairset;

static int cmp_pair(const void *a, const void *b)
{
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

static void pairset_build(pairset *s, int32_t n, int64_t m, const int64_t *keys_in)
{
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
    memcpy(keys, keys_in, sizeof(int64_t) * m);
    qsort(keys, m, sizeof(int64_t), cmp_pair);
    s->n = n;
    s->ptr = (int64_t *)calloc(n + 1, sizeof(int64_t));
    s->col = (int32_t *)malloc(sizeof(int32_t) * (m > 0 ? m : 1));
    int64_t w = 0;
    for (int64_t k = 0; k < m; k++) {
        if (k > 0 && keys[k] == keys[k - 1]) continue;
        const int32_t lo = (int32_t)(keys[k] / n), hi = (int32_t)(keys[k] % n);
        s->col[w++] = hi;
        s->ptr[lo + 1]++;
    }
    for (int i = 0; i < n; i++) s->ptr[i + 1] += s->ptr[i];
    free(keys);
}

static void pairset_free(pairset *s)
{
    free(s->ptr);
    free(s->col);
}

static int pairset_has(const pairset *s, int32_t lo, int32_t hi)
{
    int64_t a = s->ptr[lo], b = s->ptr[lo + 1];
    while (a < b) {
        const int64_t m = (a + b) / 2;
        if (s->col[m] < hi) a = m + 1;
        else if (s->col[m] > hi) b = m;
        else return 1;
    }
    return 0;
}

/* build_exclusion_set, mm_energy.rs:61-83 */
static void build_exclusions(const mmpbsa_topology_view *t, pairset *s)
{
    const int32_t n = t->n_atoms;
    int64_t m = 0, off = 0;
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (t->n_excluded > 0 ? t->n_excluded : 1));
    for (int32_t i = 0; i < n; i++) {
        const int64_t cnt = t->num_excluded ? t->num_excluded[i] : 0;
        for (int64_t k = 0; k < cnt; k++) {
            if (off + k < t->n_excluded) {
                const int32_t j = t->excluded_atoms[off + k];
                if (j >= 0 && j < n && j != i) {
                    const int32_t lo = i < j ? i : j, hi = i < j ? j : i;
                    keys[m++] = (int64_t)lo * n + hi;
                }
            }
        }
        off += cnt;
    }
    pairset_build(s, n > 0 ? n : 1, m, keys);
    free(keys);
}

/* build_14_pairs, mm_energy.rs:86-95 */
static void build_14(const mmpbsa_topology_view *t, pairset *s)
{
    const int32_t n = t->n_atoms;
    int64_t m = 0;
    int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * (t->n_dihedrals > 0 ? t->n_dihedrals : 1));
    for (int32_t d = 0; d < t->n_dihedrals; d++) {
        const int32_t *q = t->dihedrals + 6 * d;
        if (q[5]) continue;
        const int32_t i = q[0], l = q[3];
        if (i == l) continue;
        const int32_t lo = i < l ? i : l, hi = i < l ? l : i;
        keys[m++] = (int64_t)lo * n + hi;
    }
    pairset_build(s, n > 0 ? n : 1, m, keys);
    free(keys);
}

/* lj_ab, mm_energy.rs:152-161 */
static void lj_ab(const mmpbsa_topology_view *t, int32_t i, int32_t j, double *a, double *b)
{
    const int64_t idx = t->nb_parm_index[(int64_t)t->n_types * t->atom_type[i] + t->atom_type[j]];
    if (idx < 1) { *a = 0.0; *b = 0.0; return; }
    *a = t->lj_acoef[idx - 1];
    *b = t->lj_bcoef[idx - 1];
}

static double angle3(const double *a, const double *b, const double *c)
{ /* mm_energy.rs:105-114 */
    const double ba[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    const double bc[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
    const double dot = ba[0] * bc[0] + ba[1] * bc[1] + ba[2] * bc[2];
    const double mba = sqrt(ba[0] * ba[0] + ba[1] * ba[1] + ba[2] * ba[2]);
    const double mbc = sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
    double ct = dot / (mba * mbc);
    if (ct < -1.0) ct = -1.0;
    if (ct > 1.0) ct = 1.0;
    return acos(ct);
}

static void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

static double dihedral4(const double *a, const double *b, const double *c, const double *d)
{ /* mm_energy.rs:116-149 */
    const double b1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double b2[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
    const double b3[3] = {d[0] - c[0], d[1] - c[1], d[2] - c[2]};
    double n1[3], n2[3], m1[3], u[3];
    cross3(b1, b2, n1);
    cross3(b2, b3, n2);
    const double len = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
    if (len < 1e-20) {
        u[0] = u[1] = u[2] = 0.0;
    } else {
        u[0] = b2[0] / len; u[1] = b2[1] / len; u[2] = b2[2] / len;
    }
    cross3(n1, u, m1);
    const double x = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
    const double y = m1[0] * n2[0] + m1[1] * n2[1] + m1[2] * n2[2];
    return atan2(-y, x);
}

/* compute_mm_energy + compute_mm_energy_with_nb, mm_energy.rs:168-267 */
void orc_mm_energy(const mmpbsa_topology_view *t, const double *xyz, mmpbsa_mm_terms *out)
{
    memset(out, 0, sizeof(*out));
    const int32_t n = t->n_atoms;
    for (int32_t b = 0; b < t->n_bonds; b++) {
        const int32_t *q = t->bonds + 3 * b;
        const double k = t->bond_k[q[2]], r0 = t->bond_r0[q[2]];
        const double dx = xyz[3 * q[0]] - xyz[3 * q[1]];
        const double dy = xyz[3 * q[0] + 1] - xyz[3 * q[1] + 1];
        const double dz = xyz[3 * q[0] + 2] - xyz[3 * q[1] + 2];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        out->bond += k * (r - r0) * (r - r0);
    }
    for (int32_t a = 0; a < t->n_angles; a++) {
        const int32_t *q = t->angles + 4 * a;
        const double k = t->angle_k[q[3]], t0 = t->angle_theta0[q[3]];
        const double th = angle3(xyz + 3 * q[0], xyz + 3 * q[1], xyz + 3 * q[2]);
        out->angle += k * (th - t0) * (th - t0);
    }
    for (int32_t d = 0; d < t->n_dihedrals; d++) {
        const int32_t *q = t->dihedrals + 6 * d;
        const double pk = t->dihedral_pk[q[4]], pn = t->dihedral_pn[q[4]];
        const double ph = t->dihedral_phase[q[4]];
        const double phi = dihedral4(xyz + 3 * q[0], xyz + 3 * q[1], xyz + 3 * q[2], xyz + 3 * q[3]);
        out->dihedral += pk * (1.0 + cos(pn * phi - ph));
    }
    if (n < 2) return;
    pairset ex, p14;
    build_exclusions(t, &ex);
    build_14(t, &p14);
    const double scee = t->scee, scnb = t->scnb;
    double *part = (double *)calloc(8 * (size_t)n, sizeof(double));
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t i = 0; i < n; i++) {
        ksum vdw = {0, 0}, elec = {0, 0}, v14 = {0, 0}, e14 = {0, 0};
        int64_t pe = ex.ptr[i], pe_end = ex.ptr[i + 1];
        int64_t p4 = p14.ptr[i], p4_end = p14.ptr[i + 1];
        const double qi = t->charges_amber[i];
        for (int32_t j = i + 1; j < n; j++) {
            const double dx = xyz[3 * i] - xyz[3 * j];
            const double dy = xyz[3 * i + 1] - xyz[3 * j + 1];
            const double dz = xyz[3 * i + 2] - xyz[3 * j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            while (p4 < p4_end && p14.col[p4] < j) p4++;
            while (pe < pe_end && ex.col[pe] < j) pe++;
            const int is14 = (p4 < p4_end && p14.col[p4] == j);
            const int isex = (pe < pe_end && ex.col[pe] == j);
            if (is14) {
                const double r = sqrt(r2);
                const double r6 = r2 * r2 * r2;
                const double r12 = r6 * r6;
                double a, b;
                lj_ab(t, i, j, &a, &b);
                kadd(&v14, (a / r12 - b / r6) / scnb);
                kadd(&e14, qi * t->charges_amber[j] / r / scee);
            } else if (!isex) {
                const double r = sqrt(r2);
                const double r6 = r2 * r2 * r2;
                const double r12 = r6 * r6;
                double a, b;
                lj_ab(t, i, j, &a, &b);
                kadd(&vdw, a / r12 - b / r6);
                kadd(&elec, qi * t->charges_amber[j] / r);
            }
        }
        part[8 * (size_t)i] = vdw.s;
        part[8 * (size_t)i + 1] = vdw.c;
        part[8 * (size_t)i + 2] = elec.s;
        part[8 * (size_t)i + 3] = elec.c;
        part[8 * (size_t)i + 4] = v14.s;
        part[8 * (size_t)i + 5] = v14.c;
        part[8 * (size_t)i + 6] = e14.s;
        part[8 * (size_t)i + 7] = e14.c;
    }
    ksum tv = {0, 0}, te = {0, 0}, t4 = {0, 0}, t5 = {0, 0};
    for (int32_t i = 0; i < n; i++) {
        kadd(&tv, part[8 * (size_t)i]);
        kadd(&tv, part[8 * (size_t)i + 1]);
        kadd(&te, part[8 * (size_t)i + 2]);
        kadd(&te, part[8 * (size_t)i + 3]);
        kadd(&t4, part[8 * (size_t)i + 4]);
        kadd(&t4, part[8 * (size_t)i + 5]);
        kadd(&t5, part[8 * (size_t)i + 6]);
        kadd(&t5, part[8 * (size_t)i + 7]);
    }
    out->vdw = kval(&tv);
    out->elec = kval(&te);
    out->vdw_14 = kval(&t4);
    out->elec_14 = kval(&t5);
    free(part);
    pairset_free(&ex);
    pairset_free(&p14);
}

/* ------------------------------------------------------------------------ */
/* Uniform cell list used for every "atoms near a point" query.  It is a
 * superset search: callers apply the reference's exact predicate.          */
typedef struct {
    double o[3];
    double inv;
    int64_t d[3];
    int64_t *start; /* [ncell+1] */
    int32_t *idx;
} cells;

static void cells_build(cells *c, int32_t n, const double *xyz, double cell)
{
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            if (xyz[3 * i + k] < mn[k]) mn[k] = xyz[3 * i + k];
            if (xyz[3 * i + k] > mx[k]) mx[k] = xyz[3 * i + k];
        }
    if (n == 0) { mn[0] = mn[1] = mn[2] = 0.0; mx[0] = mx[1] = mx[2] = 0.0; }
    c->inv = 1.0 / cell;
    for (int k = 0; k < 3; k++) {
        c->o[k] = mn[k];
        c->d[k] = (int64_t)floor((mx[k] - mn[k]) * c->inv) + 1;
    }
    const int64_t nc = c->d[0] * c->d[1] * c->d[2];
    c->start = (int64_t *)calloc(nc + 1, sizeof(int64_t));
    c->idx = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    int64_t *cid = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int32_t i = 0; i < n; i++) {
        int64_t q[3];
        for (int k = 0; k < 3; k++) {
            q[k] = (int64_t)floor((xyz[3 * i + k] - c->o[k]) * c->inv);
            if (q[k] < 0) q[k] = 0;
            if (q[k] >= c->d[k]) q[k] = c->d[k] - 1;
        }
        cid[i] = q[0] + c->d[0] * (q[1] + c->d[1] * q[2]);
        c->start[cid[i] + 1]++;
    }
    for (int64_t k = 0; k < nc; k++) c->start[k + 1] += c->start[k];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (nc > 0 ? nc : 1));
    memcpy(fill, c->start, sizeof(int64_t) * nc);
    for (int32_t i = 0; i < n; i++) c->idx[fill[cid[i]]++] = i;
    free(fill);
    free(cid);
}

static void cells_free(cells *c)
{
    free(c->start);
    free(c->idx);
}

/* Cell range [lo, hi] covering a box of half-width w around p. */
static void cells_range(const cells *c, const double *p, double w, int64_t *lo, int64_t *hi)
{
    for (int k = 0; k < 3; k++) {
        lo[k] = (int64_t)floor((p[k] - w - c->o[k]) * c->inv);
        hi[k] = (int64_t)floor((p[k] + w - c->o[k]) * c->inv);
        if (lo[k] < 0) lo[k] = 0;
        if (hi[k] >= c->d[k]) hi[k] = c->d[k] - 1;
    }
}

/* ------------------------------------------------------------------------ */
/* Shrake-Rupley SASA: core/src/sasa.rs:147-169 (points), :225-276 (count). */
void orc_sasa_per_atom(int32_t n, const double *xyz, const double *radii, double probe,
                       int32_t n_points, double *out)
{
    double *pts = (double *)malloc(sizeof(double) * 3 * (n_points > 0 ? n_points : 1));
    if (n_points == 1) {
        pts[0] = 0.0; pts[1] = 1.0; pts[2] = 0.0;
    } else if (n_points > 1) {
        const double phi = ORC_PI * (3.0 - sqrt(5.0));
        for (int32_t i = 0; i < n_points; i++) {
            const double y = 1.0 - ((double)i / (double)(n_points - 1)) * 2.0;
            const double rad = sqrt(1.0 - y * y);
            const double th = phi * (double)i;
            pts[3 * i] = cos(th) * rad;
            pts[3 * i + 1] = y;
            pts[3 * i + 2] = sin(th) * rad;
        }
    }
    double maxr = 0.0;
    for (int32_t i = 0; i < n; i++) maxr = fmax(maxr, radii[i]);
    const double max_ext = maxr + probe;
    cells c;
    cells_build(&c, n, xyz, max_ext > 0.5 ? max_ext : 0.5);
#pragma omp parallel
    {
        int32_t cap = 256, m = 0;
        int32_t *nb = (int32_t *)malloc(sizeof(int32_t) * cap);
#pragma omp for schedule(dynamic, 32)
        for (int32_t i = 0; i < n; i++) {
            const double *ci = xyz + 3 * i;
            const double ext = radii[i] + probe;
            const double ext2 = ext * ext;
            const double search = ext + max_ext;
            int64_t lo[3], hi[3];
            cells_range(&c, ci, search, lo, hi);
            m = 0;
            for (int64_t z = lo[2]; z <= hi[2]; z++)
                for (int64_t y = lo[1]; y <= hi[1]; y++)
                    for (int64_t x = lo[0]; x <= hi[0]; x++) {
                        const int64_t cc = x + c.d[0] * (y + c.d[1] * z);
                        for (int64_t k = c.start[cc]; k < c.start[cc + 1]; k++) {
                            const int32_t j = c.idx[k];
                            if (j == i) continue;
                            if (m == cap) { cap *= 2; nb = (int32_t *)realloc(nb, sizeof(int32_t) * cap); }
                            nb[m++] = j;
                        }
                    }
            int64_t acc = 0;
            for (int32_t p = 0; p < n_points; p++) {
                const double tp[3] = {ci[0] + pts[3 * p] * ext, ci[1] + pts[3 * p + 1] * ext,
                                      ci[2] + pts[3 * p + 2] * ext};
                int ok = 1;
                for (int32_t k = 0; k < m && ok; k++) {
                    const int32_t j = nb[k];
                    const double oe = radii[j] + probe;
                    const double dx = tp[0] - xyz[3 * j];
                    const double dy = tp[1] - xyz[3 * j + 1];
                    const double dz = tp[2] - xyz[3 * j + 2];
                    const double d2 = dx * dx + dy * dy + dz * dz;
                    if (!(d2 > oe * oe)) ok = 0;
                }
                acc += ok;
            }
            const double frac = (double)acc / (double)n_points;
            const double area = 4.0 * ORC_PI * ext2;
            out[i] = frac * area;
        }
        free(nb);
    }
    cells_free(&c);
    free(pts);
}

/* ------------------------------------------------------------------------ */
/* PB grid geometry. pb_grid.rs:188-215 (auto_grid), pb_energy.rs:92-125.   */
static void bbox(int32_t n, const double *xyz, double *mn, double *mx)
{
    for (int k = 0; k < 3; k++) { mn[k] = DBL_MAX; mx[k] = -DBL_MAX; }
    for (int32_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            if (xyz[3 * i + k] < mn[k]) mn[k] = xyz[3 * i + k];
            if (xyz[3 * i + k] > mx[k]) mx[k] = xyz[3 * i + k];
        }
}

void orc_auto_grid(int32_t n, const double *xyz, double spacing, double buffer, orc_grid *g)
{
    double mn[3], mx[3];
    bbox(n, xyz, mn, mx);
    for (int k = 0; k < 3; k++) {
        const double ext = mx[k] - mn[k] + 2.0 * buffer;
        int64_t m = (int64_t)ceil(ext / spacing) + 1;
        if (m % 2 == 0) m += 1;
        g->n[k] = m;
        g->h[k] = spacing;
        const double gext = (double)(m - 1) * spacing;
        const double c = 0.5 * (mn[k] + mx[k]);
        g->o[k] = c - 0.5 * gext;
    }
}

void orc_coarse_grid(int32_t n, const double *xyz, double fine_spacing, double fillratio,
                     int64_t fscale, orc_grid *g)
{
    double mn[3], mx[3];
    bbox(n, xyz, mn, mx);
    const double cs = fine_spacing * (double)fscale;
    for (int k = 0; k < 3; k++) {
        const double ext = (mx[k] - mn[k]) * fillratio;
        int64_t m = (int64_t)ceil(ext / cs) + 1;
        if (m % 2 == 0) m += 1;
        g->n[k] = m;
        g->h[k] = cs;
        const double aext = (double)(m - 1) * cs;
        const double c = 0.5 * (mn[k] + mx[k]);
        g->o[k] = c - 0.5 * aext;
    }
}

static inline int64_t gidx(const orc_grid *g, int64_t x, int64_t y, int64_t z)
{
    return x + g->n[0] * (y + g->n[1] * z);
}
static inline int64_t glen(const orc_grid *g) { return g->n[0] * g->n[1] * g->n[2]; }

/* map_charges, pb_grid.rs:220-292 (single accumulation stream, atom order) */
void orc_map_charges(const orc_grid *g, int32_t n, const double *xyz, const double *q,
                     double *out)
{
    memset(out, 0, sizeof(double) * glen(g));
    const double vol = g->h[0] * g->h[1] * g->h[2];
    for (int32_t a = 0; a < n; a++) {
        const double fx = (xyz[3 * a] - g->o[0]) / g->h[0];
        const double fy = (xyz[3 * a + 1] - g->o[1]) / g->h[1];
        const double fz = (xyz[3 * a + 2] - g->o[2]) / g->h[2];
        const double flx = floor(fx), fly = floor(fy), flz = floor(fz);
        if (flx < 0.0 || flx >= (double)(g->n[0] - 1) || fly < 0.0 ||
            fly >= (double)(g->n[1] - 1) || flz < 0.0 || flz >= (double)(g->n[2] - 1))
            continue;
        const int64_t ix = (int64_t)flx, iy = (int64_t)fly, iz = (int64_t)flz;
        const double dx = fx - (double)ix, dy = fy - (double)iy, dz = fz - (double)iz;
        const double qd = q[a] / vol;
        const double wx[2] = {1.0 - dx, dx}, wy[2] = {1.0 - dy, dy}, wz[2] = {1.0 - dz, dz};
        for (int a1 = 0; a1 < 2; a1++)
            for (int b1 = 0; b1 < 2; b1++)
                for (int c1 = 0; c1 < 2; c1++)
                    out[gidx(g, ix + a1, iy + b1, iz + c1)] += qd * wx[a1] * wy[b1] * wz[c1];
    }
}

/* AtomSpatialHash, pb_grid.rs:319-442.  Reproduced as written (cell size =
 * max radius + probe, truncating casts) so inside/outside is the same.     */
typedef struct {
    double inv;
    double o[3];
    int64_t d[3];
    int64_t *start;
    int32_t *idx;
    int empty;
} ahash;

static void ahash_build(ahash *h, int32_t n, const double *xyz, double max_eff)
{
    h->empty = (n == 0);
    h->start = NULL;
    h->idx = NULL;
    if (n == 0) return;
    const double cs = max_eff > 0.1 ? max_eff : 0.1;
    h->inv = 1.0 / cs;
    double mn[3], mx[3];
    bbox(n, xyz, mn, mx);
    for (int k = 0; k < 3; k++) {
        h->o[k] = mn[k] - max_eff;
        h->d[k] = (int64_t)ceil((mx[k] - h->o[k]) * h->inv) + 1;
    }
    const int64_t nc = h->d[0] * h->d[1] * h->d[2];
    h->start = (int64_t *)calloc(nc + 1, sizeof(int64_t));
    h->idx = (int32_t *)malloc(sizeof(int32_t) * n);
    int64_t *cid = (int64_t *)malloc(sizeof(int64_t) * n);
    for (int32_t i = 0; i < n; i++) {
        const int64_t cx = (int64_t)((xyz[3 * i] - h->o[0]) * h->inv);
        const int64_t cy = (int64_t)((xyz[3 * i + 1] - h->o[1]) * h->inv);
        const int64_t cz = (int64_t)((xyz[3 * i + 2] - h->o[2]) * h->inv);
        cid[i] = cx + h->d[0] * (cy + h->d[1] * cz);
        h->start[cid[i] + 1]++;
    }
    for (int64_t k = 0; k < nc; k++) h->start[k + 1] += h->start[k];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * nc);
    memcpy(fill, h->start, sizeof(int64_t) * nc);
    for (int32_t i = 0; i < n; i++) h->idx[fill[cid[i]]++] = i; /* ascending i per cell */
    free(fill);
    free(cid);
}

static void ahash_free(ahash *h)
{
    free(h->start);
    free(h->idx);
}

static int ahash_inside(const ahash *h, const double *pt, const double *xyz, const double *radii,
                        double probe)
{
    if (h->empty) return 0;
    const int64_t cx = (int64_t)((pt[0] - h->o[0]) * h->inv);
    const int64_t cy = (int64_t)((pt[1] - h->o[1]) * h->inv);
    const int64_t cz = (int64_t)((pt[2] - h->o[2]) * h->inv);
    for (int64_t dz = -1; dz <= 1; dz++) {
        const int64_t jz = cz + dz;
        if (jz < 0 || jz >= h->d[2]) continue;
        for (int64_t dy = -1; dy <= 1; dy++) {
            const int64_t jy = cy + dy;
            if (jy < 0 || jy >= h->d[1]) continue;
            for (int64_t dx = -1; dx <= 1; dx++) {
                const int64_t jx = cx + dx;
                if (jx < 0 || jx >= h->d[0]) continue;
                const int64_t c = jx + h->d[0] * (jy + h->d[1] * jz);
                for (int64_t k = h->start[c]; k < h->start[c + 1]; k++) {
                    const int32_t a = h->idx[k];
                    const double ddx = pt[0] - xyz[3 * a];
                    const double ddy = pt[1] - xyz[3 * a + 1];
                    const double ddz = pt[2] - xyz[3 * a + 2];
                    const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
                    const double th = radii[a] + probe;
                    if (d2 < th * th) return 1;
                }
            }
        }
    }
    return 0;
}

/* Exact 1-D squared distance transform (lower envelope of parabolas),
 * the method of pb_grid.rs:450-497. */
static void edt1(double *f, int64_t n, double *d, int64_t *v, double *zb)
{
    if (n <= 1) return;
    int64_t k = 0;
    v[0] = 0;
    zb[0] = -INFINITY;
    zb[1] = INFINITY;
    for (int64_t q = 1; q < n; q++) {
        for (;;) {
            const int64_t vk = v[k];
            const double s = (f[q] + (double)(q * q) - f[vk] - (double)(vk * vk)) /
                             (double)(2 * (q - vk));
            if (s > zb[k]) {
                k++;
                v[k] = q;
                zb[k] = s;
                zb[k + 1] = INFINITY;
                break;
            }
            if (k == 0) {
                v[0] = q;
                zb[0] = -INFINITY;
                zb[1] = INFINITY;
                break;
            }
            k--;
        }
    }
    k = 0;
    for (int64_t q = 0; q < n; q++) {
        while (zb[k + 1] < (double)q) k++;
        const double df = (double)q - (double)v[k];
        d[q] = df * df + f[v[k]];
    }
    memcpy(f, d, sizeof(double) * n);
}

static void edt3(double *dist, const int64_t *dims)
{
    const int64_t nx = dims[0], ny = dims[1], nz = dims[2];
    int64_t md = nx > ny ? nx : ny;
    if (nz > md) md = nz;
#pragma omp parallel
    {
        double *col = (double *)malloc(sizeof(double) * md);
        double *d = (double *)malloc(sizeof(double) * md);
        int64_t *v = (int64_t *)malloc(sizeof(int64_t) * md);
        double *zb = (double *)malloc(sizeof(double) * (md + 1));
#pragma omp for
        for (int64_t iz = 0; iz < nz; iz++)
            for (int64_t iy = 0; iy < ny; iy++) edt1(dist + nx * (iy + ny * iz), nx, d, v, zb);
#pragma omp for
        for (int64_t iz = 0; iz < nz; iz++)
            for (int64_t ix = 0; ix < nx; ix++) {
                for (int64_t iy = 0; iy < ny; iy++) col[iy] = dist[ix + nx * (iy + ny * iz)];
                edt1(col, ny, d, v, zb);
                for (int64_t iy = 0; iy < ny; iy++) dist[ix + nx * (iy + ny * iz)] = col[iy];
            }
#pragma omp for
        for (int64_t iy = 0; iy < ny; iy++)
            for (int64_t ix = 0; ix < nx; ix++) {
                for (int64_t iz = 0; iz < nz; iz++) col[iz] = dist[ix + nx * (iy + ny * iz)];
                edt1(col, nz, d, v, zb);
                for (int64_t iz = 0; iz < nz; iz++) dist[ix + nx * (iy + ny * iz)] = col[iz];
            }
        free(col);
        free(d);
        free(v);
        free(zb);
    }
}

/* compute_inside_ses, pb_grid.rs:576-708.  Optional outputs sas_out (inside
 * SAS) and ext_out (connected exterior) are for the GPU map parity tests. */
void orc_inside_ses(const orc_grid *g, int32_t n, const double *xyz, const double *radii,
                    double probe, uint8_t *inside, uint8_t *sas_out, uint8_t *ext_out)
{
    const int64_t nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const int64_t N = nx * ny * nz;
    double maxr = 0.0;
    for (int32_t i = 0; i < n; i++) maxr = fmax(maxr, radii[i]);
    const double max_eff = maxr + (probe > 0.0 ? probe : 0.0);
    ahash h;
    ahash_build(&h, n, xyz, max_eff);
    if (probe <= 0.0) {
#pragma omp parallel for schedule(static)
        for (int64_t idx = 0; idx < N; idx++) {
            const int64_t iz = idx / (nx * ny), iy = (idx % (nx * ny)) / nx, ix = idx % nx;
            const double pt[3] = {g->o[0] + (double)ix * g->h[0], g->o[1] + (double)iy * g->h[1],
                                  g->o[2] + (double)iz * g->h[2]};
            inside[idx] = (uint8_t)ahash_inside(&h, pt, xyz, radii, 0.0);
        }
        if (sas_out) memcpy(sas_out, inside, N);
        if (ext_out) memset(ext_out, 0, N);
        ahash_free(&h);
        return;
    }
    uint8_t *sas = (uint8_t *)malloc(N);
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < N; idx++) {
        const int64_t iz = idx / (nx * ny), iy = (idx % (nx * ny)) / nx, ix = idx % nx;
        const double pt[3] = {g->o[0] + (double)ix * g->h[0], g->o[1] + (double)iy * g->h[1],
                              g->o[2] + (double)iz * g->h[2]};
        sas[idx] = (uint8_t)ahash_inside(&h, pt, xyz, radii, probe);
    }
    ahash_free(&h);
    /* BFS from the box faces through outside-SAS nodes (:621-686) */
    uint8_t *ext = (uint8_t *)calloc(N, 1);
    int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * N);
    int64_t qh = 0, qt = 0;
    for (int64_t iz = 0; iz < nz; iz++)
        for (int64_t iy = 0; iy < ny; iy++)
            for (int64_t ix = 0; ix < nx; ix++) {
                const int bnd = ix == 0 || ix == nx - 1 || iy == 0 || iy == ny - 1 || iz == 0 ||
                                iz == nz - 1;
                if (!bnd) continue;
                const int64_t idx = gidx(g, ix, iy, iz);
                if (!sas[idx]) { ext[idx] = 1; queue[qt++] = idx; }
            }
    while (qh < qt) {
        const int64_t idx = queue[qh++];
        const int64_t iz = idx / (nx * ny), iy = (idx % (nx * ny)) / nx, ix = idx % nx;
        int64_t nb[6];
        int m = 0;
        if (ix > 0) nb[m++] = idx - 1;
        if (ix + 1 < nx) nb[m++] = idx + 1;
        if (iy > 0) nb[m++] = idx - nx;
        if (iy + 1 < ny) nb[m++] = idx + nx;
        if (iz > 0) nb[m++] = idx - nx * ny;
        if (iz + 1 < nz) nb[m++] = idx + nx * ny;
        for (int k = 0; k < m; k++)
            if (!ext[nb[k]] && !sas[nb[k]]) { ext[nb[k]] = 1; queue[qt++] = nb[k]; }
    }
    free(queue);
    /* EDT to the connected exterior (:691-707) */
    const double hh = g->h[0];
    const double th2 = (probe / hh) * (probe / hh);
    const double big = (double)N * 2.0;
    double *dist = (double *)malloc(sizeof(double) * N);
    for (int64_t i = 0; i < N; i++) dist[i] = ext[i] ? 0.0 : big;
    edt3(dist, g->n);
    for (int64_t i = 0; i < N; i++) inside[i] = (uint8_t)(dist[i] > th2);
    free(dist);
    if (sas_out) memcpy(sas_out, sas, N);
    if (ext_out) memcpy(ext_out, ext, N);
    free(sas);
    free(ext);
}

typedef struct { double a, b; } ival;

static int cmp_ival(const void *x, const void *y)
{
    const double a = ((const ival *)x)->a, b = ((const ival *)y)->a;
    return (a > b) - (a < b);
}

/* Interval merge of fraction_inside_along_axis, pb_grid.rs:780-806 */
static double merge_fraction(ival *iv, int64_t m, double len)
{
    if (m == 0) return 0.0;
    qsort(iv, m, sizeof(ival), cmp_ival);
    double tot = 0.0, lo = iv[0].a, hi = iv[0].b;
    for (int64_t k = 1; k < m; k++) {
        if (iv[k].a <= hi) {
            hi = fmax(hi, iv[k].b);
        } else {
            tot += hi - lo;
            lo = iv[k].a;
            hi = iv[k].b;
        }
    }
    tot += hi - lo;
    double f = tot / len;
    if (f < 0.0) f = 0.0;
    if (f > 1.0) f = 1.0;
    return f;
}

static inline int edge_interval(double lo, double hi, double p1, double p2, int axis,
                                const double *c, double r, double probe, ival *out)
{
    const int a1 = axis == 0 ? 1 : 0;
    const int a2 = axis == 2 ? 1 : 2;
    const double R = r + probe;
    const double d1 = p1 - c[a1], d2 = p2 - c[a2];
    const double rs = R * R - d1 * d1 - d2 * d2;
    if (rs <= 0.0) return 0;
    const double rho = sqrt(rs);
    const double ilo = fmax(c[axis] - rho, lo);
    const double ihi = fmin(c[axis] + rho, hi);
    if (ilo < ihi) { out->a = ilo; out->b = ihi; return 1; }
    return 0;
}

/* fraction_inside_along_axis, pb_grid.rs:717-807 (all-atom scan) */
double orc_fraction_inside(double lo, double hi, double p1, double p2, int axis, int32_t n,
                           const double *xyz, const double *radii, double probe)
{
    const double len = hi - lo;
    if (len <= 0.0) return 0.0;
    int64_t cap = 16, m = 0;
    ival *iv = (ival *)malloc(sizeof(ival) * cap);
    for (int32_t a = 0; a < n; a++) {
        ival t;
        if (edge_interval(lo, hi, p1, p2, axis, xyz + 3 * a, radii[a], probe, &t)) {
            if (m == cap) { cap *= 2; iv = (ival *)realloc(iv, sizeof(ival) * cap); }
            iv[m++] = t;
        }
    }
    const double f = merge_fraction(iv, m, len);
    free(iv);
    return f;
}

/* weighted_harmonic_eps, pb_grid.rs:809-817 */
static inline double harm_eps(double f, double ein, double eout)
{
    if (f <= 0.0) return eout;
    if (f >= 1.0) return ein;
    return 1.0 / (f / ein + (1.0 - f) / eout);
}

/* assign_dielectrics, pb_grid.rs:827-940 */
void orc_assign_dielectrics(const orc_grid *g, int32_t n, const double *xyz,
                            const double *radii, double probe, double ein, double eout,
                            double *ex, double *ey, double *ez)
{
    const int64_t nx = g->n[0], ny = g->n[1], nz = g->n[2];
    const int64_t N = nx * ny * nz;
    if (fabs(ein - eout) < DBL_EPSILON) {
        for (int64_t i = 0; i < N; i++) ex[i] = ey[i] = ez[i] = ein;
        return;
    }
    uint8_t *in = (uint8_t *)malloc(N);
    orc_inside_ses(g, n, xyz, radii, probe, in, NULL, NULL);
    for (int64_t i = 0; i < N; i++) ex[i] = ey[i] = ez[i] = eout;
    double maxr = 0.0;
    for (int32_t i = 0; i < n; i++) maxr = fmax(maxr, radii[i]);
    cells c;
    const double cs = maxr > 0.5 ? maxr : 0.5;
    cells_build(&c, n, xyz, cs);
#pragma omp parallel
    {
        int64_t cap = 64;
        ival *iv = (ival *)malloc(sizeof(ival) * cap);
#pragma omp for schedule(dynamic, 1)
        for (int64_t iz = 0; iz < nz; iz++)
            for (int64_t iy = 0; iy < ny; iy++)
                for (int64_t ix = 0; ix < nx; ix++) {
                    const int64_t idx = gidx(g, ix, iy, iz);
                    const int a = in[idx];
                    const double pt[3] = {g->o[0] + (double)ix * g->h[0],
                                          g->o[1] + (double)iy * g->h[1],
                                          g->o[2] + (double)iz * g->h[2]};
                    for (int axis = 0; axis < 3; axis++) {
                        const int64_t lim = axis == 0 ? nx : axis == 1 ? ny : nz;
                        const int64_t pos = axis == 0 ? ix : axis == 1 ? iy : iz;
                        if (pos + 1 >= lim) continue;
                        const int64_t step = axis == 0 ? 1 : axis == 1 ? nx : nx * ny;
                        const int b = in[idx + step];
                        double *dst = axis == 0 ? ex : axis == 1 ? ey : ez;
                        if (a == b) {
                            dst[idx] = a ? ein : eout;
                            continue;
                        }
                        const double lo = pt[axis], hi = pt[axis] + g->h[axis];
                        const double p1 = axis == 0 ? pt[1] : pt[0];
                        const double p2 = axis == 2 ? pt[1] : pt[2];
                        double f;
                        if (!g_face_scan) {
                            f = orc_fraction_inside(lo, hi, p1, p2, axis, n, xyz, radii, 0.0);
                        } else {
                            const double len = hi - lo;
                            int64_t m = 0;
                            if (len > 0.0) {
                                double mid[3] = {pt[0], pt[1], pt[2]};
                                mid[axis] = 0.5 * (lo + hi);
                                int64_t clo[3], chi[3];
                                cells_range(&c, mid, maxr + 0.5 * len + 1e-9, clo, chi);
                                for (int64_t z = clo[2]; z <= chi[2]; z++)
                                    for (int64_t y = clo[1]; y <= chi[1]; y++)
                                        for (int64_t x = clo[0]; x <= chi[0]; x++) {
                                            const int64_t cc = x + c.d[0] * (y + c.d[1] * z);
                                            for (int64_t k = c.start[cc]; k < c.start[cc + 1]; k++) {
                                                const int32_t at = c.idx[k];
                                                ival t;
                                                if (edge_interval(lo, hi, p1, p2, axis, xyz + 3 * at,
                                                                  radii[at], 0.0, &t)) {
                                                    if (m == cap) {
                                                        cap *= 2;
                                                        iv = (ival *)realloc(iv, sizeof(ival) * cap);
                                                    }
                                                    iv[m++] = t;
                                                }
                                            }
                                        }
                                f = merge_fraction(iv, m, len);
                            } else {
                                f = 0.0;
                            }
                        }
                        dst[idx] = harm_eps(f, ein, eout);
                    }
                }
        free(iv);
    }
    cells_free(&c);
    free(in);
}

/* assign_kappa, pb_grid.rs:951-985 */
void orc_assign_kappa(const orc_grid *g, int32_t n, const double *xyz, const double *radii,
                      double ion, double kappa, double eps_s, double *out)
{
    const int64_t nx = g->n[0], ny = g->n[1], N = glen(g);
    const double kbs = eps_s * kappa * kappa;
    double maxr = 0.0;
    for (int32_t i = 0; i < n; i++) maxr = fmax(maxr, radii[i]);
    ahash h;
    ahash_build(&h, n, xyz, maxr + ion);
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < N; idx++) {
        const int64_t iz = idx / (nx * ny), iy = (idx % (nx * ny)) / nx, ix = idx % nx;
        const double pt[3] = {g->o[0] + (double)ix * g->h[0], g->o[1] + (double)iy * g->h[1],
                              g->o[2] + (double)iz * g->h[2]};
        out[idx] = ahash_inside(&h, pt, xyz, radii, ion) ? 0.0 : kbs;
    }
    ahash_free(&h);
}

/* for_each_boundary_point, pb_solver.rs:31-60, as an index list. */
static int64_t boundary_list(const int64_t *d, int64_t **out)
{
    const int64_t nx = d[0], ny = d[1], nz = d[2];
    int64_t cap = 2 * nx * ny + 2 * nx * (nz > 2 ? nz - 2 : 0) +
                  2 * (ny > 2 ? ny - 2 : 0) * (nz > 2 ? nz - 2 : 0) + 8;
    int64_t *l = (int64_t *)malloc(sizeof(int64_t) * cap);
    int64_t m = 0;
    const int64_t zs[2] = {0, nz - 1};
    for (int s = 0; s < 2; s++)
        for (int64_t iy = 0; iy < ny; iy++)
            for (int64_t ix = 0; ix < nx; ix++) l[m++] = ix + nx * (iy + ny * zs[s]);
    for (int64_t iz = 1; iz < nz - 1; iz++) {
        const int64_t ys[2] = {0, ny - 1};
        for (int s = 0; s < 2; s++)
            for (int64_t ix = 0; ix < nx; ix++) l[m++] = ix + nx * (ys[s] + ny * iz);
    }
    for (int64_t iz = 1; iz < nz - 1; iz++)
        for (int64_t iy = 1; iy < ny - 1; iy++) {
            l[m++] = 0 + nx * (iy + ny * iz);
            l[m++] = (nx - 1) + nx * (iy + ny * iz);
        }
    *out = l;
    return m;
}

/* set_dh_boundary, pb_solver.rs:68-93 */
void orc_dh_boundary(const orc_grid *g, int32_t n, const double *xyz, const double *q,
                     double kappa, double eps_out, double *phi)
{
    int64_t *bl;
    const int64_t m = boundary_list(g->n, &bl);
    const int64_t nx = g->n[0], ny = g->n[1];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < m; k++) {
        const int64_t idx = bl[k];
        const int64_t iz = idx / (nx * ny), iy = (idx % (nx * ny)) / nx, ix = idx % nx;
        const double pt[3] = {g->o[0] + (double)ix * g->h[0], g->o[1] + (double)iy * g->h[1],
                              g->o[2] + (double)iz * g->h[2]};
        double s = 0.0;
        for (int32_t a = 0; a < n; a++) {
            const double dx = pt[0] - xyz[3 * a];
            const double dy = pt[1] - xyz[3 * a + 1];
            const double dz = pt[2] - xyz[3 * a + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r > 1e-10) s += COULOMB * q[a] * exp(-kappa * r) / (eps_out * r);
        }
        phi[idx] = s;
    }
    free(bl);
}

/* PbGrid::interpolate_with_data, pb_grid.rs:140-182 */
static double interp(const orc_grid *g, const double *data, const double *p)
{
    const double fx = (p[0] - g->o[0]) / g->h[0];
    const double fy = (p[1] - g->o[1]) / g->h[1];
    const double fz = (p[2] - g->o[2]) / g->h[2];
    const double flx = floor(fx), fly = floor(fy), flz = floor(fz);
    if (flx < 0.0 || flx >= (double)(g->n[0] - 1) || fly < 0.0 ||
        fly >= (double)(g->n[1] - 1) || flz < 0.0 || flz >= (double)(g->n[2] - 1))
        return 0.0;
    const int64_t ix = (int64_t)flx, iy = (int64_t)fly, iz = (int64_t)flz;
    const double dx = fx - (double)ix, dy = fy - (double)iy, dz = fz - (double)iz;
    const double c000 = data[gidx(g, ix, iy, iz)], c100 = data[gidx(g, ix + 1, iy, iz)];
    const double c010 = data[gidx(g, ix, iy + 1, iz)], c110 = data[gidx(g, ix + 1, iy + 1, iz)];
    const double c001 = data[gidx(g, ix, iy, iz + 1)], c101 = data[gidx(g, ix + 1, iy, iz + 1)];
    const double c011 = data[gidx(g, ix, iy + 1, iz + 1)];
    const double c111 = data[gidx(g, ix + 1, iy + 1, iz + 1)];
    const double c00 = c000 * (1.0 - dx) + c100 * dx;
    const double c10 = c010 * (1.0 - dx) + c110 * dx;
    const double c01 = c001 * (1.0 - dx) + c101 * dx;
    const double c11 = c011 * (1.0 - dx) + c111 * dx;
    const double c0 = c00 * (1.0 - dy) + c10 * dy;
    const double c1 = c01 * (1.0 - dy) + c11 * dy;
    return c0 * (1.0 - dz) + c1 * dz;
}

/* interpolated_boundary + set_interpolated_boundary, pb_solver.rs:100-121 */
void orc_interp_boundary(const orc_grid *fine, const orc_grid *coarse, const double *cphi,
                         double *phi)
{
    int64_t *bl;
    const int64_t m = boundary_list(fine->n, &bl);
    const int64_t nx = fine->n[0], ny = fine->n[1];
    for (int64_t k = 0; k < m; k++) {
        const int64_t idx = bl[k];
        const int64_t iz = idx / (nx * ny), iy = (idx % (nx * ny)) / nx, ix = idx % nx;
        const double pt[3] = {fine->o[0] + (double)ix * fine->h[0],
                              fine->o[1] + (double)iy * fine->h[1],
                              fine->o[2] + (double)iz * fine->h[2]};
        phi[idx] = interp(coarse, cphi, pt);
    }
    free(bl);
}

/* ------------------------------------------------------------------------ */
/* LPBE solver, pb_solver.rs:124-1186                                         */
static const double FOUR_PI_EC = 4.0 * ORC_PI * COULOMB;

static double spectral_radius(const int64_t *d)
{
    return (cos(ORC_PI / (double)d[0]) + cos(ORC_PI / (double)d[1]) + cos(ORC_PI / (double)d[2])) /
           3.0;
}

static double estimate_omega(const int64_t *d)
{
    const double r = spectral_radius(d);
    return 2.0 / (1.0 + sqrt(1.0 - r * r));
}

typedef struct {
    int64_t n[3];
    double h[3];
    const double *ex, *ey, *ez, *k2;
} op3;

/* stencil_at, pb_solver.rs:159-194 */
static inline void stencil(const op3 *A, const double *p, int64_t idx, double ihx, double ihy,
                           double ihz, double *nb, double *dg)
{
    const int64_t sy = A->n[0], sz = A->n[0] * A->n[1];
    const double exp_ = A->ex[idx], exm = A->ex[idx - 1];
    const double eyp = A->ey[idx], eym = A->ey[idx - sy];
    const double ezp = A->ez[idx], ezm = A->ez[idx - sz];
    *nb = (exp_ * p[idx + 1] + exm * p[idx - 1]) * ihx + (eyp * p[idx + sy] + eym * p[idx - sy]) * ihy +
          (ezp * p[idx + sz] + ezm * p[idx - sz]) * ihz;
    *dg = (exp_ + exm) * ihx + (eyp + eym) * ihy + (ezp + ezm) * ihz + A->k2[idx];
}

/* sor_smooth, pb_solver.rs:214-327.  Red and black updates are independent
 * within a colour, so the iterate does not depend on traversal order. */
static double sor_smooth(const op3 *A, double *p, const double *rho, double omega, int sweeps,
                         double rhs_scale)
{
    const int64_t nx = A->n[0], ny = A->n[1], nz = A->n[2];
    const double ihx = 1.0 / (A->h[0] * A->h[0]);
    const double ihy = 1.0 / (A->h[1] * A->h[1]);
    const double ihz = 1.0 / (A->h[2] * A->h[2]);
    double last = 0.0;
    double *part = (double *)calloc(nz > 0 ? nz : 1, sizeof(double));
    for (int s = 0; s < sweeps; s++) {
        double rms = 0.0;
        int64_t cnt = 0;
        for (int color = 0; color < 2; color++) {
            int64_t ccnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : ccnt)
            for (int64_t iz = 1; iz < nz - 1; iz++) {
                double loc = 0.0;
                for (int64_t iy = 1; iy < ny - 1; iy++) {
                    const int64_t x0 = 1 + ((1 + iy + iz + color) % 2);
                    for (int64_t ix = x0; ix < nx - 1; ix += 2) {
                        const int64_t idx = ix + nx * (iy + ny * iz);
                        double nb, dg;
                        stencil(A, p, idx, ihx, ihy, ihz, &nb, &dg);
                        const double rhs = rhs_scale * rho[idx];
                        const double gs = (nb + rhs) / dg;
                        const double pn = p[idx] + omega * (gs - p[idx]);
                        const double r = gs - p[idx];
                        loc += r * r;
                        ccnt++;
                        p[idx] = pn;
                    }
                }
                part[iz] = loc;
            }
            for (int64_t iz = 1; iz < nz - 1; iz++) rms += part[iz];
            cnt += ccnt;
        }
        last = cnt > 0 ? sqrt(rms / (double)cnt) : 0.0;
    }
    free(part);
    return last;
}

/* compute_residual_into, pb_solver.rs:453-509 */
static void residual_into(const op3 *A, const double *p, const double *rho, double rhs_scale,
                          double *out)
{
    const int64_t nx = A->n[0], ny = A->n[1], nz = A->n[2];
    const double ihx = 1.0 / (A->h[0] * A->h[0]);
    const double ihy = 1.0 / (A->h[1] * A->h[1]);
    const double ihz = 1.0 / (A->h[2] * A->h[2]);
    memset(out, 0, sizeof(double) * nx * ny * nz);
#pragma omp parallel for schedule(static)
    for (int64_t iz = 1; iz < nz - 1; iz++)
        for (int64_t iy = 1; iy < ny - 1; iy++)
            for (int64_t ix = 1; ix < nx - 1; ix++) {
                const int64_t idx = ix + nx * (iy + ny * iz);
                double nb, dg;
                stencil(A, p, idx, ihx, ihy, ihz, &nb, &dg);
                out[idx] = rhs_scale * rho[idx] + nb - dg * p[idx];
            }
}

/* compute_rms_residual, pb_solver.rs:946-1009 */
static double rms_residual(const op3 *A, const double *p, const double *rho, double rhs_scale)
{
    const int64_t nx = A->n[0], ny = A->n[1], nz = A->n[2];
    const double ihx = 1.0 / (A->h[0] * A->h[0]);
    const double ihy = 1.0 / (A->h[1] * A->h[1]);
    const double ihz = 1.0 / (A->h[2] * A->h[2]);
    double *part = (double *)calloc(nz > 0 ? nz : 1, sizeof(double));
    int64_t cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
    for (int64_t iz = 1; iz < nz - 1; iz++) {
        double loc = 0.0;
        for (int64_t iy = 1; iy < ny - 1; iy++)
            for (int64_t ix = 1; ix < nx - 1; ix++) {
                const int64_t idx = ix + nx * (iy + ny * iz);
                double nb, dg;
                stencil(A, p, idx, ihx, ihy, ihz, &nb, &dg);
                const double r = rhs_scale * rho[idx] + nb - dg * p[idx];
                loc += r * r;
                cnt++;
            }
        part[iz] = loc;
    }
    double s = 0.0;
    for (int64_t iz = 1; iz < nz - 1; iz++) s += part[iz];
    free(part);
    return cnt > 0 ? sqrt(s / (double)cnt) : 0.0;
}

double orc_rms_residual(const orc_grid *g, const double *phi, const double *charge,
                        const double *ex, const double *ey, const double *ez, const double *k2,
                        double rhs_scale)
{
    op3 A = {{g->n[0], g->n[1], g->n[2]}, {g->h[0], g->h[1], g->h[2]}, ex, ey, ez, k2};
    return rms_residual(&A, phi, charge, rhs_scale);
}

/* restrict_into, pb_solver.rs:578-624 */
void orc_restrict(const double *fine, const int64_t *fd, double *coarse, const int64_t *cd,
                  double scale)
{
#pragma omp parallel for schedule(static)
    for (int64_t cz = 0; cz < cd[2]; cz++)
        for (int64_t cy = 0; cy < cd[1]; cy++)
            for (int64_t cx = 0; cx < cd[0]; cx++) {
                double sum = 0.0, ws = 0.0;
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int64_t ix = 2 * cx + dx, iy = 2 * cy + dy, iz = 2 * cz + dz;
                            if (ix < 0 || ix >= fd[0] || iy < 0 || iy >= fd[1] || iz < 0 || iz >= fd[2])
                                continue;
                            const double w = (dx == 0 ? 2.0 : 1.0) * (dy == 0 ? 2.0 : 1.0) *
                                             (dz == 0 ? 2.0 : 1.0);
                            sum += w * fine[ix + fd[0] * (iy + fd[1] * iz)];
                            ws += w;
                        }
                coarse[cx + cd[0] * (cy + cd[1] * cz)] = scale * sum / ws;
            }
}

/* prolongate_add, pb_solver.rs:628-681 (adds to every fine node) */
void orc_prolongate_add(const double *coarse, const int64_t *cd, double *fine, const int64_t *fd)
{
    const int64_t cnx = cd[0], cny = cd[1];
#pragma omp parallel for schedule(static)
    for (int64_t fz = 0; fz < fd[2]; fz++)
        for (int64_t fy = 0; fy < fd[1]; fy++)
            for (int64_t fx = 0; fx < fd[0]; fx++) {
                const double xf = (double)fx / 2.0, yf = (double)fy / 2.0, zf = (double)fz / 2.0;
                int64_t x0 = (int64_t)floor(xf), y0 = (int64_t)floor(yf), z0 = (int64_t)floor(zf);
                const int64_t mx = cd[0] >= 2 ? cd[0] - 2 : 0, my = cd[1] >= 2 ? cd[1] - 2 : 0;
                const int64_t mz = cd[2] >= 2 ? cd[2] - 2 : 0;
                if (x0 > mx) x0 = mx;
                if (y0 > my) y0 = my;
                if (z0 > mz) z0 = mz;
                const double dx = xf - (double)x0, dy = yf - (double)y0, dz = zf - (double)z0;
#define CI(a, b, c) ((a) + cnx * ((b) + cny * (c)))
                const double c000 = coarse[CI(x0, y0, z0)], c100 = coarse[CI(x0 + 1, y0, z0)];
                const double c010 = coarse[CI(x0, y0 + 1, z0)], c110 = coarse[CI(x0 + 1, y0 + 1, z0)];
                const double c001 = coarse[CI(x0, y0, z0 + 1)], c101 = coarse[CI(x0 + 1, y0, z0 + 1)];
                const double c011 = coarse[CI(x0, y0 + 1, z0 + 1)];
                const double c111 = coarse[CI(x0 + 1, y0 + 1, z0 + 1)];
#undef CI
                const double c00 = c000 * (1.0 - dx) + c100 * dx;
                const double c10 = c010 * (1.0 - dx) + c110 * dx;
                const double c01 = c001 * (1.0 - dx) + c101 * dx;
                const double c11 = c011 * (1.0 - dx) + c111 * dx;
                const double c0 = c00 * (1.0 - dy) + c10 * dy;
                const double c1 = c01 * (1.0 - dy) + c11 * dy;
                fine[fx + fd[0] * (fy + fd[1] * fz)] += c0 * (1.0 - dz) + c1 * dz;
            }
}

/* coarsen_dielectrics / coarsen_kappa (injection), pb_solver.rs:684-748 */
static void inject(const double *f, const int64_t *fd, double *c, const int64_t *cd)
{
    for (int64_t cz = 0; cz < cd[2]; cz++)
        for (int64_t cy = 0; cy < cd[1]; cy++)
            for (int64_t cx = 0; cx < cd[0]; cx++) {
                int64_t x = 2 * cx, y = 2 * cy, z = 2 * cz;
                if (x > fd[0] - 1) x = fd[0] - 1;
                if (y > fd[1] - 1) y = fd[1] - 1;
                if (z > fd[2] - 1) z = fd[2] - 1;
                c[cx + cd[0] * (cy + cd[1] * cz)] = f[x + fd[0] * (y + fd[1] * z)];
            }
}

typedef struct {
    op3 A;          /* coarse operator */
    double omega;
    int64_t fd[3];  /* fine dims of this transfer */
    double *res, *rhs, *cor;
    double *own[4];
} mglevel;

/* build_multigrid_levels, pb_solver.rs:776-836 */
static int build_levels(const op3 *fine, mglevel **out)
{
    int cap = 16, m = 0;
    mglevel *L = (mglevel *)calloc(cap, sizeof(mglevel));
    op3 cur = *fine;
    for (;;) {
        int64_t md = cur.n[0];
        if (cur.n[1] < md) md = cur.n[1];
        if (cur.n[2] < md) md = cur.n[2];
        if (md <= 5) break;
        mglevel *l = &L[m++];
        int64_t cd[3] = {(cur.n[0] + 1) / 2, (cur.n[1] + 1) / 2, (cur.n[2] + 1) / 2};
        const int64_t cn = cd[0] * cd[1] * cd[2];
        const int64_t fn = cur.n[0] * cur.n[1] * cur.n[2];
        for (int k = 0; k < 4; k++) l->own[k] = (double *)malloc(sizeof(double) * cn);
        inject(cur.ex, cur.n, l->own[0], cd);
        inject(cur.ey, cur.n, l->own[1], cd);
        inject(cur.ez, cur.n, l->own[2], cd);
        inject(cur.k2, cur.n, l->own[3], cd);
        for (int k = 0; k < 3; k++) {
            l->A.n[k] = cd[k];
            l->A.h[k] = cur.h[k] * 2.0;
            l->fd[k] = cur.n[k];
        }
        l->A.ex = l->own[0]; l->A.ey = l->own[1]; l->A.ez = l->own[2]; l->A.k2 = l->own[3];
        l->omega = estimate_omega(cd);
        l->res = (double *)calloc(fn, sizeof(double));
        l->rhs = (double *)calloc(cn, sizeof(double));
        l->cor = (double *)calloc(cn, sizeof(double));
        cur = l->A;
    }
    *out = L;
    return m;
}

static void free_levels(mglevel *L, int m)
{
    for (int i = 0; i < m; i++) {
        for (int k = 0; k < 4; k++) free(L[i].own[k]);
        free(L[i].res);
        free(L[i].rhs);
        free(L[i].cor);
    }
    free(L);
}

/* v_cycle_workspace, pb_solver.rs:842-939 */
static void vcycle(const op3 *A, double *p, const double *rho, double omega, mglevel *L, int m,
                   double rhs_scale)
{
    if (m == 0) {
        sor_smooth(A, p, rho, omega, 50, rhs_scale);
        return;
    }
    sor_smooth(A, p, rho, omega, 3, rhs_scale);
    residual_into(A, p, rho, rhs_scale, L->res);
    orc_restrict(L->res, L->fd, L->rhs, L->A.n, 0.25);
    memset(L->cor, 0, sizeof(double) * L->A.n[0] * L->A.n[1] * L->A.n[2]);
    vcycle(&L->A, L->cor, L->rhs, L->omega, L + 1, m - 1, 1.0);
    orc_prolongate_add(L->cor, L->A.n, p, A->n);
    sor_smooth(A, p, rho, omega, 3, rhs_scale);
}

static void apply_bc(const orc_grid *g, int bc, int32_t n, const double *xyz, const double *q,
                     double kappa_bulk, double eps_out, double *phi, const double *saved)
{
    if (bc == 1) {
        orc_dh_boundary(g, n, xyz, q, kappa_bulk, eps_out, phi);
    } else if (bc == 2 && saved) {
        int64_t *bl;
        const int64_t m = boundary_list(g->n, &bl);
        for (int64_t k = 0; k < m; k++) phi[bl[k]] = saved[bl[k]];
        free(bl);
    }
}

/* solve_lpbe_initial, pb_solver.rs:373-447 (Chebyshev-ramped SOR) */
static void sor_solve(const orc_grid *g, const op3 *A, const double *rho, double tol,
                      int64_t max_it, double *phi, orc_solve *res)
{
    const double rj = spectral_radius(g->n);
    const double r2 = rj * rj;
    double w = 1.0, fr = DBL_MAX;
    int64_t it = 0;
    for (int64_t k = 0; k < max_it; k++) {
        fr = sor_smooth(A, phi, rho, w, 1, FOUR_PI_EC);
        it = k + 1;
        if (k == 0)
            w = 1.0 / (1.0 - r2 / 2.0);
        else
            w = 1.0 / (1.0 - r2 * w / 4.0);
        if (fr < tol) {
            res->iterations = (int32_t)it;
            res->final_residual = fr;
            res->converged = 1;
            return;
        }
    }
    res->iterations = (int32_t)it;
    res->final_residual = fr;
    res->converged = 0;
}

void orc_solve_sor(const orc_grid *g, const double *charge, const double *ex, const double *ey,
                   const double *ez, const double *k2, int bc, int32_t n, const double *xyz,
                   const double *q, double kappa_bulk, double eps_out, double tol,
                   int64_t max_it, double *phi, orc_solve *res)
{
    const int64_t N = glen(g);
    double *saved = NULL;
    if (bc == 2) {
        saved = (double *)malloc(sizeof(double) * N);
        memcpy(saved, phi, sizeof(double) * N);
    }
    memset(phi, 0, sizeof(double) * N);
    apply_bc(g, bc, n, xyz, q, kappa_bulk, eps_out, phi, saved);
    op3 A = {{g->n[0], g->n[1], g->n[2]}, {g->h[0], g->h[1], g->h[2]}, ex, ey, ez, k2};
    sor_solve(g, &A, charge, tol, max_it, phi, res);
    free(saved);
}

/* solve_lpbe_multigrid, pb_solver.rs:1015-1167.  For bc == 2 the caller puts
 * the interpolated boundary values into phi's boundary nodes on entry. */
void orc_solve_multigrid(const orc_grid *g, const double *charge, const double *ex,
                         const double *ey, const double *ez, const double *k2, int bc,
                         int32_t n, const double *xyz, const double *q, double kappa_bulk,
                         double eps_out, double tol, int64_t max_it, double *phi,
                         orc_solve *res)
{
    const int64_t N = glen(g);
    double *saved = NULL;
    if (bc == 2) {
        saved = (double *)malloc(sizeof(double) * N);
        memcpy(saved, phi, sizeof(double) * N);
    }
    memset(phi, 0, sizeof(double) * N);
    apply_bc(g, bc, n, xyz, q, kappa_bulk, eps_out, phi, saved);
    op3 A = {{g->n[0], g->n[1], g->n[2]}, {g->h[0], g->h[1], g->h[2]}, ex, ey, ez, k2};
    int64_t md = g->n[0];
    if (g->n[1] < md) md = g->n[1];
    if (g->n[2] < md) md = g->n[2];
    if (md <= 8) {
        memset(phi, 0, sizeof(double) * N);
        apply_bc(g, bc, n, xyz, q, kappa_bulk, eps_out, phi, saved);
        sor_solve(g, &A, charge, tol, max_it, phi, res);
        free(saved);
        return;
    }
    int64_t maxv = max_it / 6;
    if (maxv < 1) maxv = 1;
    if (maxv > 200) maxv = 200;
    mglevel *L;
    const int m = build_levels(&A, &L);
    const double omega = estimate_omega(g->n);
    double best = DBL_MAX, prev = DBL_MAX;
    int growing = 0;
    for (int64_t cyc = 0; cyc < maxv; cyc++) {
        vcycle(&A, phi, charge, omega, L, m, FOUR_PI_EC);
        const double r = rms_residual(&A, phi, charge, FOUR_PI_EC);
        if (r < best) {
            best = r;
            growing = 0;
        } else {
            growing++;
        }
        const int div = (g_force_div > 0 && cyc + 1 == g_force_div) || !isfinite(r) ||
                        r > best * 1e4 || (growing >= 5 && r > prev);
        if (div) {
            /* fall back to Chebyshev SOR from the partial potential, BC re-applied */
            apply_bc(g, bc, n, xyz, q, kappa_bulk, eps_out, phi, saved);
            sor_solve(g, &A, charge, tol, max_it, phi, res);
            free_levels(L, m);
            free(saved);
            return;
        }
        prev = r;
        if (r < tol) {
            res->iterations = (int32_t)(cyc + 1);
            res->final_residual = r;
            res->converged = 1;
            free_levels(L, m);
            free(saved);
            return;
        }
    }
    res->final_residual = rms_residual(&A, phi, charge, FOUR_PI_EC);
    res->iterations = (int32_t)maxv;
    res->converged = 0;
    free_levels(L, m);
    free(saved);
}

/* compute_elec_energy, pb_solver.rs:1174-1186 */
double orc_elec_energy(const orc_grid *g, const double *phi, int32_t n, const double *xyz,
                       const double *q)
{
    double e = 0.0;
    for (int32_t a = 0; a < n; a++) e += q[a] * interp(g, phi, xyz + 3 * a);
    return 0.5 * e;
}

/* pb_solve_with_focusing, pb_energy.rs:134-322 */
static void pb_solve_focus(const orc_grid *fine, const double *fine_rho, const orc_grid *cg,
                           int32_t n, const double *xyz, const double *q, const double *radii,
                           const mmpbsa_pb_params *p, double ein, double eout, double kappa,
                           int ionic, mmpbsa_pb_solve_info *info, double *phi_out)
{
    const double bck = ionic ? kappa : 0.0;
    const int64_t NF = glen(fine);
    double *phi = (double *)calloc(NF, sizeof(double));
    double *ex = (double *)malloc(sizeof(double) * NF);
    double *ey = (double *)malloc(sizeof(double) * NF);
    double *ez = (double *)malloc(sizeof(double) * NF);
    double *k2 = (double *)calloc(NF, sizeof(double));
    orc_solve rs = {0, 0, 0.0};
    int bc = 1;
    info->focused = 0;
    info->coarse_iterations = 0;
    if (cg) {
        const int64_t NC = glen(cg);
        double *crho = (double *)malloc(sizeof(double) * NC);
        double *cex = (double *)malloc(sizeof(double) * NC);
        double *cey = (double *)malloc(sizeof(double) * NC);
        double *cez = (double *)malloc(sizeof(double) * NC);
        double *ck2 = (double *)calloc(NC, sizeof(double));
        double *cphi = (double *)calloc(NC, sizeof(double));
        orc_map_charges(cg, n, xyz, q, crho);
        orc_assign_dielectrics(cg, n, xyz, radii, p->probe_radius, ein, eout, cex, cey, cez);
        if (ionic) orc_assign_kappa(cg, n, xyz, radii, p->ion_radius, kappa, p->solvent_dielectric, ck2);
        orc_solve cr = {0, 0, 0.0};
        orc_solve_multigrid(cg, crho, cex, cey, cez, ck2, 1, n, xyz, q, bck, eout,
                            p->tolerance * 10.0, p->max_iterations, cphi, &cr);
        info->coarse_iterations = cr.iterations;
        int ok = cr.converged;
        if (ok) {
            double mx = 0.0;
            for (int64_t i = 0; i < NC; i++) mx = fmax(mx, fabs(cphi[i]));
            ok = isfinite(mx) && mx < 1e10;
        }
        if (ok) {
            orc_interp_boundary(fine, cg, cphi, phi);
            bc = 2;
            info->focused = 1;
        }
        free(crho); free(cex); free(cey); free(cez); free(ck2); free(cphi);
    }
    orc_assign_dielectrics(fine, n, xyz, radii, p->probe_radius, ein, eout, ex, ey, ez);
    if (ionic) orc_assign_kappa(fine, n, xyz, radii, p->ion_radius, kappa, p->solvent_dielectric, k2);
    orc_solve_multigrid(fine, fine_rho, ex, ey, ez, k2, bc, n, xyz, q, bck, eout, p->tolerance,
                        p->max_iterations, phi, &rs);
    info->energy = orc_elec_energy(fine, phi, n, xyz, q);
    info->iterations = rs.iterations;
    info->converged = rs.converged;
    info->final_residual = rs.final_residual;
    if (phi_out) memcpy(phi_out, phi, sizeof(double) * NF);
    free(phi); free(ex); free(ey); free(ez); free(k2);
}

/* compute_pb_energy, pb_energy.rs:332-438, with the fine_dims and explicit
 * focusing-grid extensions of mmpbsa_pb_params; atom_terms (nullable) gets
 * 0.5 q_i (phi_solv - phi_ref)(r_i) per atom on the fine grid. */
static void pb_energy_impl(const mmpbsa_topology_view *t, const double *xyz,
                           const mmpbsa_pb_params *p, mmpbsa_pb_result *out, double *atom_terms)
{
    memset(out, 0, sizeof(*out));
    const int32_t n = t->n_atoms;
    const double kappa = orc_compute_kappa(p->salt_concentration, p->solvent_dielectric,
                                           p->temperature);
    orc_grid fine, cg;
    orc_auto_grid(n, xyz, p->grid_spacing, p->grid_buffer, &fine);
    if (p->fine_dims[0] > 0) {
        double mn[3], mx[3];
        bbox(n, xyz, mn, mx);
        for (int k = 0; k < 3; k++) {
            fine.n[k] = p->fine_dims[k];
            fine.o[k] = 0.5 * (mn[k] + mx[k]) - 0.5 * ((double)(fine.n[k] - 1) * p->grid_spacing);
        }
    }
    if (p->fine_centered)
        for (int k = 0; k < 3; k++)
            fine.o[k] = p->fine_center[k] - 0.5 * ((double)(fine.n[k] - 1) * p->grid_spacing);
    const int explicit_cg = p->coarse_dims[0] > 0;
    const int focus = explicit_cg || (p->fillratio > 0.0 && p->fscale > 1);
    if (explicit_cg) {
        for (int k = 0; k < 3; k++) {
            cg.n[k] = p->coarse_dims[k];
            cg.h[k] = p->coarse_spacing;
            cg.o[k] = p->coarse_center[k] - 0.5 * ((double)(cg.n[k] - 1) * p->coarse_spacing);
        }
    } else if (focus) {
        orc_coarse_grid(n, xyz, p->grid_spacing, p->fillratio, p->fscale, &cg);
    }
    for (int k = 0; k < 3; k++) {
        out->fine_dims[k] = (int32_t)fine.n[k];
        out->coarse_dims[k] = focus ? (int32_t)cg.n[k] : 0;
    }
    const int64_t NF = glen(&fine);
    double *rho = (double *)malloc(sizeof(double) * NF);
    double *ps = atom_terms ? (double *)malloc(sizeof(double) * NF) : NULL;
    double *pr = atom_terms ? (double *)malloc(sizeof(double) * NF) : NULL;
    orc_map_charges(&fine, n, xyz, t->charges, rho);
    pb_solve_focus(&fine, rho, focus ? &cg : NULL, n, xyz, t->charges, t->radii, p,
                   p->solute_dielectric, p->solvent_dielectric, kappa, 1, &out->solv, ps);
    pb_solve_focus(&fine, rho, focus ? &cg : NULL, n, xyz, t->charges, t->radii, p,
                   p->solute_dielectric, p->solute_dielectric, kappa, 0, &out->ref, pr);
    out->total = out->solv.energy - out->ref.energy;
    if (atom_terms)
        for (int32_t a = 0; a < n; a++)
            atom_terms[a] = 0.5 * t->charges[a] *
                            (interp(&fine, ps, xyz + 3 * a) - interp(&fine, pr, xyz + 3 * a));
    free(rho);
    free(ps);
    free(pr);
}

void orc_pb_energy(const mmpbsa_topology_view *t, const double *xyz, const mmpbsa_pb_params *p,
                   mmpbsa_pb_result *out)
{
    pb_energy_impl(t, xyz, p, out, NULL);
}

void orc_pb_atom_terms(const mmpbsa_topology_view *t, const double *xyz,
                       const mmpbsa_pb_params *p, mmpbsa_pb_result *out, double *atom_terms)
{
    pb_energy_impl(t, xyz, p, out, atom_terms);
}

/* decompose_binding_energy_impl pair loop, decomposition.rs:240-283 */
void orc_decomp_pairs(const mmpbsa_topology_view *t, const double *xyz, int32_t n_rec,
                      const int32_t *rec, int32_t n_lig, const int32_t *lig,
                      const int32_t *atom_res, double *res_vdw, double *res_elec)
{
    pairset ex, p14;
    build_exclusions(t, &ex);
    build_14(t, &p14);
    const double scee = t->scee, scnb = t->scnb;
    ksum *kv = (ksum *)calloc(t->n_residues > 0 ? t->n_residues : 1, sizeof(ksum));
    ksum *ke = (ksum *)calloc(t->n_residues > 0 ? t->n_residues : 1, sizeof(ksum));
    for (int32_t a = 0; a < n_rec; a++) {
        const int32_t ri = rec[a];
        for (int32_t b = 0; b < n_lig; b++) {
            const int32_t li = lig[b];
            const int32_t i = ri < li ? ri : li, j = ri < li ? li : ri;
            const double dx = xyz[3 * ri] - xyz[3 * li];
            const double dy = xyz[3 * ri + 1] - xyz[3 * li + 1];
            const double dz = xyz[3 * ri + 2] - xyz[3 * li + 2];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (r < 1e-10) continue;
            double v, e;
            const int is14 = i != j && pairset_has(&p14, i, j);
            const int isex = i != j && pairset_has(&ex, i, j);
            if (is14 || !isex) {
                const double r2 = r * r;
                const double r6 = r2 * r2 * r2;
                const double r12 = r6 * r6;
                double A, B;
                lj_ab(t, ri, li, &A, &B);
                const double qi = t->charges_amber[ri], qj = t->charges_amber[li];
                if (is14) {
                    v = (A / r12 - B / r6) / scnb;
                    e = qi * qj / r / scee;
                } else {
                    v = A / r12 - B / r6;
                    e = qi * qj / r;
                }
            } else {
                continue;
            }
            kadd(&kv[atom_res[ri]], v * 0.5);
            kadd(&kv[atom_res[li]], v * 0.5);
            kadd(&ke[atom_res[ri]], e * 0.5);
            kadd(&ke[atom_res[li]], e * 0.5);
        }
    }
    for (int32_t r = 0; r < t->n_residues; r++) {
        res_vdw[r] += kval(&kv[r]);
        res_elec[r] += kval(&ke[r]);
    }
    free(kv);
    free(ke);
    pairset_free(&ex);
    pairset_free(&p14);
}

/* ---- quasi-harmonic entropy (entropy.rs:97-275) -------------------------- */
static int desc_cmp(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    /* NaN compares equal to everything, as partial_cmp().unwrap_or(Equal) */
    if (x > y) return -1;
    if (x < y) return 1;
    return 0;
}

void orc_jacobi_eigenvalues(double *a, int32_t dim, double *eig)
{
    const double tol = 1e-10;
    for (int32_t i = 0; i < dim; ++i) eig[i] = a[(int64_t)i * dim + i];
    int changed = 1;
    for (int sweep = 0; changed && sweep < 100; ++sweep) {
        changed = 0;
        for (int32_t p = 0; p < dim; ++p)
            for (int32_t q = p + 1; q < dim; ++q) {
                const double apq = a[(int64_t)p * dim + q];
                if (fabs(apq) < tol) continue;
                /* rotation angle from theta = (a_qq - a_pp) / (2 a_pq): the
                 * smaller root t of t^2 + 2 theta t - 1 = 0 */
                const double theta = (eig[q] - eig[p]) / (2.0 * apq);
                double t;
                if (fabs(theta) > 1e15) {
                    t = 1.0 / (2.0 * theta);
                } else {
                    const double sg = theta >= 0.0 ? 1.0 : -1.0;
                    t = sg / (fabs(theta) + sqrt(1.0 + theta * theta));
                }
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double sn = t * c;
                const double h = sn / (1.0 + c);
                eig[p] -= t * apq;
                eig[q] += t * apq;
                a[(int64_t)p * dim + q] = 0.0;
                a[(int64_t)q * dim + p] = 0.0;
                for (int32_t r = 0; r < dim; ++r) {
                    if (r == p || r == q) continue;
                    const double rp = a[(int64_t)r * dim + p], rq = a[(int64_t)r * dim + q];
                    const double np_ = rp - sn * (rq + h * rp);
                    const double nq = rq + sn * (rp - h * rq);
                    a[(int64_t)r * dim + p] = np_;
                    a[(int64_t)p * dim + r] = np_;
                    a[(int64_t)r * dim + q] = nq;
                    a[(int64_t)q * dim + r] = nq;
                }
                changed = 1;
            }
    }
    qsort(eig, (size_t)dim, sizeof(double), desc_cmp);
}

int orc_quasi_harmonic_entropy(const double *xyz, int32_t nf, int32_t na, const double *masses,
                               double temperature, double *minus_tds)
{
    if (nf < 2 || na <= 0) return 0;
    const int32_t dim = 3 * na;
    double *mean = calloc((size_t)dim, sizeof(double));
    double *d = malloc(sizeof(double) * (size_t)dim);
    double *cov = calloc((size_t)dim * (size_t)dim, sizeof(double));
    double *eig = malloc(sizeof(double) * (size_t)dim);
    /* mass-weighted means, frames in order */
    for (int32_t f = 0; f < nf; ++f)
        for (int32_t i = 0; i < na; ++i) {
            const double w = sqrt(masses[i]);
            for (int c = 0; c < 3; ++c)
                mean[3 * i + c] += xyz[((int64_t)f * na + i) * 3 + c] * w;
        }
    for (int32_t k = 0; k < dim; ++k) mean[k] /= (double)nf;
    /* covariance, upper triangle accumulated frame by frame */
    for (int32_t f = 0; f < nf; ++f) {
        for (int32_t i = 0; i < na; ++i) {
            const double w = sqrt(masses[i]);
            for (int c = 0; c < 3; ++c)
                d[3 * i + c] = xyz[((int64_t)f * na + i) * 3 + c] * w - mean[3 * i + c];
        }
        for (int32_t i = 0; i < dim; ++i) {
            const double di = d[i];
            double *row = cov + (int64_t)i * dim;
            for (int32_t j = i; j < dim; ++j) row[j] += di * d[j];
        }
    }
    for (int32_t i = 0; i < dim; ++i)
        for (int32_t j = i; j < dim; ++j) {
            const double v = cov[(int64_t)i * dim + j] / (double)nf;
            cov[(int64_t)i * dim + j] = v;
            cov[(int64_t)j * dim + i] = v;
        }
    orc_jacobi_eigenvalues(cov, dim, eig);
    /* quantum harmonic oscillator entropy per mode, SI units */
    const double kbt = 1.380649e-23 * temperature;
    const double hbar = 1.054571817e-34;
    const double conv = 1e-20 * 1.66054e-27;
    double s = 0.0;
    for (int32_t k = 0; k < dim; ++k) {
        const double l = eig[k];
        if (l < 1e-10) continue;
        const double alpha = hbar / sqrt(kbt * (l * conv));
        if (alpha > 500.0) continue;
        s += alpha / (exp(alpha) - 1.0) - log(1.0 - exp(-alpha));
    }
    *minus_tds = -temperature * 0.00198688 * s;
    free(mean);
    free(d);
    free(cov);
    free(eig);
    return 1;
}
