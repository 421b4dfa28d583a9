// topo.cpp -- host topology, sub-topology extraction, device residency,
// memory pool.
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "engine.h"

namespace mmpbsa {

template <typename T>
static std::vector<T> vcopy(const T *p, int64_t n)
{
    if (!p || n <= 0) return std::vector<T>();
    return std::vector<T>(p, p + n);
}

HostTopo HostTopo::from_view(const mmpbsa_topology_view &v)
{
    if (v.n_atoms < 0 || v.n_residues < 0 || v.n_types < 0)
        throw DeviceError(MMPBSA_ERR_INVALID_ARG, "negative size in topology view");
    if (v.n_atoms > 0 && (!v.charges || !v.charges_amber || !v.radii || !v.screen || !v.atom_type))
        throw DeviceError(MMPBSA_ERR_INVALID_ARG, "topology view: null per-atom array");
    HostTopo t;
    t.n = v.n_atoms;
    t.n_res = v.n_residues;
    t.n_types = v.n_types;
    t.q = vcopy(v.charges, v.n_atoms);
    t.qa = vcopy(v.charges_amber, v.n_atoms);
    t.radii = vcopy(v.radii, v.n_atoms);
    t.screen = vcopy(v.screen, v.n_atoms);
    t.type = vcopy(v.atom_type, v.n_atoms);
    t.nbidx = vcopy(v.nb_parm_index, (int64_t)v.n_types * v.n_types);
    t.ljA = vcopy(v.lj_acoef, v.n_lj);
    t.ljB = vcopy(v.lj_bcoef, v.n_lj);
    t.res_ptr = vcopy(v.residue_pointers, v.n_residues);
    t.bonds = vcopy(v.bonds, 3LL * v.n_bonds);
    t.angles = vcopy(v.angles, 4LL * v.n_angles);
    t.dihs = vcopy(v.dihedrals, 6LL * v.n_dihedrals);
    t.bond_k = vcopy(v.bond_k, v.n_bond_types);
    t.bond_r0 = vcopy(v.bond_r0, v.n_bond_types);
    t.angle_k = vcopy(v.angle_k, v.n_angle_types);
    t.angle_t0 = vcopy(v.angle_theta0, v.n_angle_types);
    t.dih_pk = vcopy(v.dihedral_pk, v.n_dihedral_types);
    t.dih_pn = vcopy(v.dihedral_pn, v.n_dihedral_types);
    t.dih_ph = vcopy(v.dihedral_phase, v.n_dihedral_types);
    t.nex = v.num_excluded ? vcopy(v.num_excluded, v.n_atoms) : std::vector<int>(v.n_atoms, 0);
    t.exl = vcopy(v.excluded_atoms, v.n_excluded);
    t.scee = v.scee;
    t.scnb = v.scnb;
    for (int i = 0; i < t.n; ++i)
        if (t.type[i] < 0 || t.type[i] >= std::max(t.n_types, 1))
            throw DeviceError(MMPBSA_ERR_INVALID_ARG, "atom type out of range");
    for (int64_t k = 0; k < (int64_t)t.nbidx.size(); ++k)
        if (t.nbidx[k] > (int64_t)t.ljA.size())
            throw DeviceError(MMPBSA_ERR_INVALID_ARG, "nb_parm_index beyond LJ tables");
    auto check_idx = [&](const std::vector<int> &v, int stride, int n_idx, int ntype,
                         const char *what) {
        for (size_t k = 0; k < v.size(); k += stride) {
            for (int a = 0; a < n_idx; ++a)
                if (v[k + a] < 0 || v[k + a] >= t.n)
                    throw DeviceError(MMPBSA_ERR_INVALID_ARG, std::string(what) + ": atom out of range");
            if (v[k + n_idx] < 0 || v[k + n_idx] >= ntype)
                throw DeviceError(MMPBSA_ERR_INVALID_ARG, std::string(what) + ": type out of range");
        }
    };
    check_idx(t.bonds, 3, 2, (int)t.bond_k.size(), "bond");
    check_idx(t.angles, 4, 3, (int)t.angle_k.size(), "angle");
    check_idx(t.dihs, 6, 4, (int)t.dih_pk.size(), "dihedral");
    for (int r = 0; r < t.n_res; ++r)
        if (t.res_ptr[r] < 0 || t.res_ptr[r] > t.n)
            throw DeviceError(MMPBSA_ERR_INVALID_ARG, "residue pointer out of range");
    return t;
}

std::vector<int> HostTopo::atom_residue_indices() const
{
    std::vector<int> out(n, 0);
    for (int r = 0; r < n_res; ++r) {
        const int s = res_ptr[r], e = r + 1 < n_res ? res_ptr[r + 1] : n;
        for (int a = s; a < e; ++a) out[a] = r;
    }
    return out;
}

std::vector<int> HostTopo::build_selection(const std::vector<int> &res) const
{
    for (int r : res)
        if (r < 0 || r >= n_res)
            throw DeviceError(MMPBSA_ERR_BAD_RESIDUE,
                              "Residue index " + std::to_string(r) + " out of range (0-" +
                                  std::to_string(n_res - 1) + ")");
    std::vector<int> out;
    for (int r : res) {
        const int s = res_ptr[r], e = r + 1 < n_res ? res_ptr[r + 1] : n;
        for (int a = s; a < e; ++a) out.push_back(a);
    }
    return out;
}

HostTopo HostTopo::extract(const std::vector<int> &atoms) const
{
    HostTopo t;
    const int m = (int)atoms.size();
    std::vector<int> o2n(n, -1);
    for (int k = 0; k < m; ++k) o2n[atoms[k]] = k;
    t.n = m;
    t.n_types = n_types;
    t.nbidx = nbidx;
    t.ljA = ljA;
    t.ljB = ljB;
    t.bond_k = bond_k;
    t.bond_r0 = bond_r0;
    t.angle_k = angle_k;
    t.angle_t0 = angle_t0;
    t.dih_pk = dih_pk;
    t.dih_pn = dih_pn;
    t.dih_ph = dih_ph;
    t.scee = scee;
    t.scnb = scnb;
    t.q.resize(m);
    t.qa.resize(m);
    t.radii.resize(m);
    t.screen.resize(m);
    t.type.resize(m);
    for (int k = 0; k < m; ++k) {
        const int a = atoms[k];
        t.q[k] = q[a];
        t.qa[k] = qa[a];
        t.radii[k] = radii[a];
        t.screen[k] = screen[a];
        t.type[k] = type[a];
    }
    const std::vector<int> ares = atom_residue_indices();
    int last = -1;
    for (int k = 0; k < m; ++k) {
        const int r = ares[atoms[k]];
        if (r != last) {
            t.res_ptr.push_back(k);
            last = r;
        }
    }
    t.n_res = (int)t.res_ptr.size();
    for (size_t b = 0; b < bonds.size(); b += 3) {
        const int i = o2n[bonds[b]], j = o2n[bonds[b + 1]];
        if (i >= 0 && j >= 0) {
            t.bonds.push_back(i);
            t.bonds.push_back(j);
            t.bonds.push_back(bonds[b + 2]);
        }
    }
    for (size_t b = 0; b < angles.size(); b += 4) {
        const int i = o2n[angles[b]], j = o2n[angles[b + 1]], k = o2n[angles[b + 2]];
        if (i >= 0 && j >= 0 && k >= 0) {
            t.angles.insert(t.angles.end(), {i, j, k, angles[b + 3]});
        }
    }
    for (size_t b = 0; b < dihs.size(); b += 6) {
        const int i = o2n[dihs[b]], j = o2n[dihs[b + 1]], k = o2n[dihs[b + 2]],
                  l = o2n[dihs[b + 3]];
        if (i >= 0 && j >= 0 && k >= 0 && l >= 0)
            t.dihs.insert(t.dihs.end(), {i, j, k, l, dihs[b + 4], dihs[b + 5]});
    }
    t.nex.assign(m, 0);
    int64_t off = 0;
    for (int i = 0; i < n; ++i) {
        const int cnt = i < (int)nex.size() ? nex[i] : 0;
        const int ni = o2n[i];
        if (ni >= 0) {
            int c = 0;
            for (int k = 0; k < cnt; ++k) {
                if (off + k < (int64_t)exl.size()) {
                    const int j = exl[off + k];
                    if (j >= 0 && j < n && o2n[j] >= 0) {
                        t.exl.push_back(o2n[j]);
                        ++c;
                    }
                }
            }
            t.nex[ni] = c;
        }
        off += cnt;
    }
    return t;
}

// ---- device residency ------------------------------------------------------
static void build_csr(int n, std::vector<std::pair<int, int>> &pairs, std::vector<int> &ptr,
                      std::vector<int> &col)
{
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    ptr.assign(n + 1, 0);
    col.clear();
    for (auto &pr : pairs) {
        ptr[pr.first + 1]++;
        col.push_back(pr.second);
    }
    for (int i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
}

void DevTopo::build(const HostTopo &t)
{
    h = t;
    const int n = t.n;
    max_radius = 0.0;
    for (double r : t.radii) max_radius = std::max(max_radius, r);
    // exclusions, build_exclusion_set (mm_energy.rs:61-83)
    std::vector<std::pair<int, int>> ex, p14;
    int64_t off = 0;
    for (int i = 0; i < n; ++i) {
        const int cnt = i < (int)t.nex.size() ? t.nex[i] : 0;
        for (int k = 0; k < cnt; ++k) {
            if (off + k < (int64_t)t.exl.size()) {
                const int j = t.exl[off + k];
                if (j >= 0 && j < n && j != i) ex.push_back({std::min(i, j), std::max(i, j)});
            }
        }
        off += cnt;
    }
    // 1-4 pairs, build_14_pairs (mm_energy.rs:86-95)
    for (size_t d = 0; d < t.dihs.size(); d += 6) {
        if (t.dihs[d + 5]) continue;
        const int i = t.dihs[d], l = t.dihs[d + 3];
        if (i == l) continue;
        p14.push_back({std::min(i, l), std::max(i, l)});
    }
    nb = std::max(1, ceil_div(n, kTile));
    std::vector<uint8_t> flags((size_t)nb * nb, 0);
    for (auto &pr : ex) flags[(size_t)(pr.first / kTile) * nb + pr.second / kTile] = 1;
    for (auto &pr : p14) flags[(size_t)(pr.first / kTile) * nb + pr.second / kTile] = 1;
    build_csr(n, ex, h_ex_ptr, h_ex_col);
    build_csr(n, p14, h_p14_ptr, h_p14_col);
    // LJ tables expanded by type pair (lj_ab, mm_energy.rs:152-161)
    const int nt = std::max(t.n_types, 1);
    std::vector<double> A((size_t)nt * nt, 0.0), B((size_t)nt * nt, 0.0);
    for (int a = 0; a < t.n_types; ++a)
        for (int b = 0; b < t.n_types; ++b) {
            const int64_t idx = t.nbidx[(size_t)a * t.n_types + b];
            if (idx >= 1) {
                A[(size_t)a * nt + b] = t.ljA[idx - 1];
                B[(size_t)a * nt + b] = t.ljB[idx - 1];
            }
        }
    q.upload(t.q);
    qa.upload(t.qa);
    radii.upload(t.radii);
    screen.upload(t.screen);
    type.upload(t.type);
    ljA.upload(A);
    ljB.upload(B);
    ex_ptr.upload(h_ex_ptr);
    ex_col.upload(h_ex_col);
    p14_ptr.upload(h_p14_ptr);
    p14_col.upload(h_p14_col);
    tile_flag.upload(flags);
    bonds.upload(t.bonds);
    angles.upload(t.angles);
    dihs.upload(t.dihs);
    auto nz = [](std::vector<double> v) {
        if (v.empty()) v.push_back(0.0);
        return v;
    };
    bond_k.upload(nz(t.bond_k));
    bond_r0.upload(nz(t.bond_r0));
    angle_k.upload(nz(t.angle_k));
    angle_t0.upload(nz(t.angle_t0));
    dih_pk.upload(nz(t.dih_pk));
    dih_pn.upload(nz(t.dih_pn));
    dih_ph.upload(nz(t.dih_ph));
}

PairTopo DevTopo::view() const
{
    PairTopo v;
    v.n = h.n;
    v.n_types = std::max(h.n_types, 1);
    v.q_amber = qa.p;
    v.radii = radii.p;
    v.screen = screen.p;
    v.type = type.p;
    v.ljA = ljA.p;
    v.ljB = ljB.p;
    v.ex_ptr = ex_ptr.p;
    v.ex_col = ex_col.p;
    v.p14_ptr = p14_ptr.p;
    v.p14_col = p14_col.p;
    v.tile_flag = tile_flag.p;
    v.nb = nb;
    v.bonds = bonds.p;
    v.n_bonds = (int)(h.bonds.size() / 3);
    v.bond_k = bond_k.p;
    v.bond_r0 = bond_r0.p;
    v.angles = angles.p;
    v.n_angles = (int)(h.angles.size() / 4);
    v.angle_k = angle_k.p;
    v.angle_t0 = angle_t0.p;
    v.dihs = dihs.p;
    v.n_dihs = (int)(h.dihs.size() / 6);
    v.dih_pk = dih_pk.p;
    v.dih_pn = dih_pn.p;
    v.dih_ph = dih_ph.p;
    v.scee = h.scee;
    v.scnb = h.scnb;
    return v;
}

// ---- pool -------------------------------------------------------------------
std::shared_mutex &capture_gate()
{
    static std::shared_mutex m;
    return m;
}

Pool::~Pool()
{
    trim();
    std::unique_lock<std::shared_mutex> g(capture_gate());
    for (auto &kv : live_) (void)hipFree(kv.first);
}

void *Pool::get(size_t bytes)
{
    bytes = (bytes + 255) & ~(size_t)255;
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= bytes * 2) {
        void *p = it->second;
        live_[p] = it->first;
        cached_ -= it->first;
        free_.erase(it);
        return p;
    }
    // A fresh allocation waits for any graph capture in another thread to
    // end (see capture_gate); no capture region allocates, so the capturing
    // thread never gets here holding the gate shared.
    void *p = nullptr;
    hipError_t e;
    {
        std::unique_lock<std::shared_mutex> g(capture_gate());
        e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        {
            std::unique_lock<std::shared_mutex> g(capture_gate());
            e = hipMalloc(&p, bytes);
        }
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw DeviceError(MMPBSA_ERR_OOM, "device allocation of " + std::to_string(bytes) +
                                                  " bytes failed");
        }
    }
    live_[p] = bytes;
    return p;
}

void Pool::put(void *p)
{
    auto it = live_.find(p);
    if (it == live_.end()) return;
    free_.insert({it->second, p});
    cached_ += it->second;
    live_.erase(it);
}

void Pool::trim()
{
    std::unique_lock<std::shared_mutex> g(capture_gate());
    for (auto &kv : free_) (void)hipFree(kv.second);
    free_.clear();
    cached_ = 0;
}

Worker::Worker()
{
    std::unique_lock<std::shared_mutex> g(capture_gate());
    HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_OK(hipHostMalloc((void **)&pin, sizeof(double) * 64, hipHostMallocDefault));
    HIP_OK(hipHostMalloc((void **)&pin_i, sizeof(int) * 64, hipHostMallocDefault));
}

Worker::~Worker()
{
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
    std::unique_lock<std::shared_mutex> g(capture_gate());
    if (pin) (void)hipHostFree(pin);
    if (pin_i) (void)hipHostFree(pin_i);
}

void Worker::sync() { HIP_OK(hipStreamSynchronize(stream)); }

Worker &Worker::aux(int i)
{
    if (!aux_[i]) aux_[i] = std::make_unique<Worker>();
    return *aux_[i];
}

}  // namespace mmpbsa
