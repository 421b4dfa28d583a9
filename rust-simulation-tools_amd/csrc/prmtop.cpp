// prmtop.cpp -- AMBER prmtop reader (host): the topology ingest in front of
// the device path.  Follows parse_prmtop (crates/core/src/amber/prmtop.rs:
// 318-733): %FLAG sections, whitespace-separated numbers, 4-character names,
// 1-based indices made 0-based, bond / angle / dihedral entries stored as
// 3 * atom index, charges divided by 18.2223, a negative third or fourth
// dihedral index marking ignore_14, exclusion placeholder 0 -> -1.
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mmpbsa_hip.h"
#include "common.h"

using mmpbsa::DeviceError;

struct mmpbsa_prmtop {
    int32_t n_atoms = 0, n_residues = 0, n_types = 0;
    std::vector<double> charges, charges_amber, radii, screen, masses, lj_a, lj_b;
    std::vector<double> bond_k, bond_r0, angle_k, angle_t0, dih_pk, dih_pn, dih_ph;
    std::vector<int32_t> types, res_ptr, bonds, angles, dihedrals, num_excl, excl;
    std::vector<int64_t> nb_index;
    std::vector<char> atom_names, res_labels;  // records of kNameRec bytes
    double scee = 1.2, scnb = 2.0;
};

namespace {

constexpr double kAmberChargeFactor = 18.2223;  // prmtop.rs:20
constexpr int kNameRec = 5;                      // 4 characters + NUL

[[noreturn]] void io_error(const std::string &m) { throw DeviceError(MMPBSA_ERR_IO, m); }

struct Sections {
    std::unordered_map<std::string, std::vector<std::string>> s;

    const std::vector<std::string> &get(const std::string &flag) const
    {
        auto it = s.find(flag);
        if (it == s.end()) io_error("Missing section: " + flag);
        return it->second;
    }
    bool has(const std::string &flag) const { return s.count(flag) != 0; }

    std::vector<int64_t> ints(const std::string &flag) const
    {
        std::vector<int64_t> v;
        for (const std::string &line : get(flag)) {
            const char *p = line.c_str();
            for (;;) {
                while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
                if (!*p) break;
                const char *q = p;
                while (*q && *q != ' ' && *q != '\t' && *q != '\r') ++q;
                std::string tok(p, q);
                char *end = nullptr;
                errno = 0;
                const long long x = std::strtoll(tok.c_str(), &end, 10);
                if (errno || !end || *end)
                    io_error("Failed to parse integer in " + flag + ": invalid digit in '" + tok + "'");
                v.push_back(x);
                p = q;
            }
        }
        return v;
    }

    std::vector<double> floats(const std::string &flag) const
    {
        std::vector<double> v;
        for (const std::string &line : get(flag)) {
            const char *p = line.c_str();
            for (;;) {
                while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
                if (!*p) break;
                const char *q = p;
                while (*q && *q != ' ' && *q != '\t' && *q != '\r') ++q;
                std::string tok(p, q);
                char *end = nullptr;
                const double x = std::strtod(tok.c_str(), &end);
                if (!end || *end || tok.empty())
                    io_error("Failed to parse float in " + flag + ": invalid float literal '" + tok + "'");
                v.push_back(x);
                p = q;
            }
        }
        return v;
    }

    // fixed-width fields, trimmed; a trailing partial field is kept if not blank
    std::vector<std::string> strings(const std::string &flag, size_t width) const
    {
        std::vector<std::string> v;
        auto trim = [](std::string t) {
            const size_t a = t.find_first_not_of(" \t\r");
            if (a == std::string::npos) return std::string();
            const size_t b = t.find_last_not_of(" \t\r");
            return t.substr(a, b - a + 1);
        };
        for (const std::string &line : get(flag)) {
            size_t pos = 0;
            while (pos + width <= line.size()) {
                v.push_back(trim(line.substr(pos, width)));
                pos += width;
            }
            if (pos < line.size()) {
                std::string t = trim(line.substr(pos));
                if (!t.empty()) v.push_back(t);
            }
        }
        return v;
    }
};

Sections read_sections(const char *path)
{
    std::ifstream in(path);
    if (!in) io_error(std::string("Failed to open prmtop file: ") + std::strerror(errno));
    Sections sec;
    std::string line, flag;
    std::vector<std::string> data;
    bool have = false;
    while (std::getline(in, line)) {
        if (line.compare(0, 5, "%FLAG") == 0) {
            if (have) sec.s[flag] = std::move(data);
            data.clear();
            std::string f = line.substr(5);
            const size_t a = f.find_first_not_of(" \t\r");
            const size_t b = f.find_last_not_of(" \t\r");
            flag = a == std::string::npos ? std::string() : f.substr(a, b - a + 1);
            have = true;
        } else if (line.compare(0, 7, "%FORMAT") == 0 || line.compare(0, 8, "%VERSION") == 0 ||
                   line.compare(0, 8, "%COMMENT") == 0) {
            continue;
        } else if (have) {
            data.push_back(line);
        }
    }
    if (have) sec.s[flag] = std::move(data);
    return sec;
}

void put_names(const std::vector<std::string> &v, size_t n, std::vector<char> &out)
{
    out.assign(n * kNameRec, 0);
    for (size_t k = 0; k < n && k < v.size(); ++k)
        std::strncpy(out.data() + k * kNameRec, v[k].c_str(), kNameRec - 1);
}

std::vector<double> opt_floats(const Sections &s, const char *flag)
{
    return s.has(flag) ? s.floats(flag) : std::vector<double>();
}

void parse(const char *path, mmpbsa_prmtop &t)
{
    const Sections s = read_sections(path);
    const std::vector<int64_t> ptr = s.ints("POINTERS");
    if (ptr.size() < 12) io_error("POINTERS section too short");
    if (ptr[0] < 0 || ptr[1] < 0 || ptr[11] < 0)
        io_error("POINTERS contains negative values: n_atoms=" + std::to_string(ptr[0]) +
                 ", n_types=" + std::to_string(ptr[1]) + ", n_residues=" + std::to_string(ptr[11]));
    const size_t n = (size_t)ptr[0], nt = (size_t)ptr[1], nr = (size_t)ptr[11];
    t.n_atoms = (int32_t)n;
    t.n_types = (int32_t)nt;
    t.n_residues = (int32_t)nr;

    const std::vector<std::string> names = s.strings("ATOM_NAME", 4);
    if (names.size() < n)
        io_error("ATOM_NAME has " + std::to_string(names.size()) + " entries, expected " +
                 std::to_string(n));
    put_names(names, n, t.atom_names);

    const std::vector<int64_t> ti = s.ints("ATOM_TYPE_INDEX");
    for (size_t i = 0; i < ti.size() && i < n; ++i) {
        if (ti[i] < 1)
            io_error("Invalid ATOM_TYPE_INDEX at position " + std::to_string(i) + ": " +
                     std::to_string(ti[i]) + " (must be >= 1)");
        t.types.push_back((int32_t)(ti[i] - 1));
    }
    const std::vector<double> q = s.floats("CHARGE");
    for (size_t i = 0; i < q.size() && i < n; ++i) {
        t.charges_amber.push_back(q[i]);
        t.charges.push_back(q[i] / kAmberChargeFactor);
    }
    put_names(s.strings("RESIDUE_LABEL", 4), nr, t.res_labels);
    const std::vector<int64_t> rp = s.ints("RESIDUE_POINTER");
    for (size_t i = 0; i < rp.size() && i < nr; ++i) {
        if (rp[i] < 1)
            io_error("Invalid RESIDUE_POINTER at position " + std::to_string(i) + ": " +
                     std::to_string(rp[i]) + " (must be >= 1)");
        t.res_ptr.push_back((int32_t)(rp[i] - 1));
    }
    t.lj_a = s.floats("LENNARD_JONES_ACOEF");
    t.lj_b = s.floats("LENNARD_JONES_BCOEF");
    t.nb_index = s.ints("NONBONDED_PARM_INDEX");
    for (size_t i = 0; i < nt; ++i) {  // the self-interaction entries must exist
        const size_t k = nt * i + i;
        if (k >= t.nb_index.size())
            io_error("NONBONDED_PARM_INDEX out of bounds: index " + std::to_string(k) +
                     " >= length " + std::to_string(t.nb_index.size()));
    }
    for (const char *sec : {"BONDS_INC_HYDROGEN", "BONDS_WITHOUT_HYDROGEN"}) {
        if (!s.has(sec)) continue;
        const std::vector<int64_t> r = s.ints(sec);
        for (size_t c = 0; c + 3 <= r.size(); c += 3) {
            if (r[c] < 0 || r[c + 1] < 0 || r[c + 2] < 1) continue;
            t.bonds.insert(t.bonds.end(), {(int32_t)(r[c] / 3), (int32_t)(r[c + 1] / 3),
                                           (int32_t)(r[c + 2] - 1)});
        }
    }
    t.masses = opt_floats(s, "MASS");
    t.radii = opt_floats(s, "RADII");
    t.screen = opt_floats(s, "SCREEN");
    t.bond_k = opt_floats(s, "BOND_FORCE_CONSTANT");
    t.bond_r0 = opt_floats(s, "BOND_EQUIL_VALUE");
    t.angle_k = opt_floats(s, "ANGLE_FORCE_CONSTANT");
    t.angle_t0 = opt_floats(s, "ANGLE_EQUIL_VALUE");
    t.dih_pk = opt_floats(s, "DIHEDRAL_FORCE_CONSTANT");
    t.dih_pn = opt_floats(s, "DIHEDRAL_PERIODICITY");
    t.dih_ph = opt_floats(s, "DIHEDRAL_PHASE");
    for (const char *sec : {"ANGLES_INC_HYDROGEN", "ANGLES_WITHOUT_HYDROGEN"}) {
        if (!s.has(sec)) continue;
        const std::vector<int64_t> r = s.ints(sec);
        for (size_t c = 0; c + 4 <= r.size(); c += 4) {
            if (r[c] < 0 || r[c + 1] < 0 || r[c + 2] < 0 || r[c + 3] < 1) continue;
            t.angles.insert(t.angles.end(), {(int32_t)(r[c] / 3), (int32_t)(r[c + 1] / 3),
                                             (int32_t)(r[c + 2] / 3), (int32_t)(r[c + 3] - 1)});
        }
    }
    for (const char *sec : {"DIHEDRALS_INC_HYDROGEN", "DIHEDRALS_WITHOUT_HYDROGEN"}) {
        if (!s.has(sec)) continue;
        const std::vector<int64_t> r = s.ints(sec);
        for (size_t c = 0; c + 5 <= r.size(); c += 5) {
            if (r[c + 4] < 1) continue;
            auto ua = [](int64_t x) { return (int32_t)((x < 0 ? -x : x) / 3); };
            t.dihedrals.insert(t.dihedrals.end(),
                               {ua(r[c]), ua(r[c + 1]), ua(r[c + 2]), ua(r[c + 3]),
                                (int32_t)(r[c + 4] - 1), (r[c + 2] < 0 || r[c + 3] < 0) ? 1 : 0});
        }
    }
    if (s.has("NUMBER_EXCLUDED_ATOMS"))
        for (int64_t x : s.ints("NUMBER_EXCLUDED_ATOMS")) t.num_excl.push_back((int32_t)x);
    if (s.has("EXCLUDED_ATOMS_LIST"))
        for (int64_t x : s.ints("EXCLUDED_ATOMS_LIST")) t.excl.push_back(x > 0 ? (int32_t)(x - 1) : -1);
    const std::vector<double> scee = opt_floats(s, "SCEE_SCALE_FACTOR");
    const std::vector<double> scnb = opt_floats(s, "SCNB_SCALE_FACTOR");
    if (!scee.empty()) t.scee = scee[0];
    if (!scnb.empty()) t.scnb = scnb[0];

    // the device path needs every per-atom array at full length: a short one
    // is an error here (the reference would index past it later)
    auto need = [&](const char *what, size_t have) {
        if (have < n)
            io_error(std::string(what) + " has " + std::to_string(have) + " entries, expected " +
                     std::to_string(n));
    };
    need("ATOM_TYPE_INDEX", t.types.size());
    need("CHARGE", t.charges.size());
    if (t.res_ptr.size() < nr)
        io_error("RESIDUE_POINTER has " + std::to_string(t.res_ptr.size()) + " entries, expected " +
                 std::to_string(nr));
    if (t.masses.size() < n) t.masses.resize(n, 0.0);   // missing sections read as zeros
    if (t.radii.size() < n) t.radii.resize(n, 0.0);
    if (t.screen.size() < n) t.screen.resize(n, 0.0);
    if (t.num_excl.size() < n) t.num_excl.resize(n, 0);
}

std::string g_prmtop_err;

}  // namespace

extern "C" {

int mmpbsa_prmtop_read(const char *path, mmpbsa_prmtop **out)
{
    if (!path || !out) return MMPBSA_ERR_INVALID_ARG;
    *out = nullptr;
    try {
        auto t = new mmpbsa_prmtop();
        try {
            parse(path, *t);
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
        return MMPBSA_OK;
    } catch (const DeviceError &e) {
        g_prmtop_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_prmtop_err = e.what();
        return MMPBSA_ERR_IO;
    }
}

const char *mmpbsa_prmtop_last_error(void) { return g_prmtop_err.c_str(); }

int mmpbsa_prmtop_destroy(mmpbsa_prmtop *t)
{
    delete t;
    return MMPBSA_OK;
}

int mmpbsa_prmtop_view(const mmpbsa_prmtop *t, mmpbsa_topology_view *v)
{
    if (!t || !v) return MMPBSA_ERR_INVALID_ARG;
    std::memset(v, 0, sizeof(*v));
    v->n_atoms = t->n_atoms;
    v->n_residues = t->n_residues;
    v->n_types = t->n_types;
    v->charges = t->charges.data();
    v->charges_amber = t->charges_amber.data();
    v->radii = t->radii.data();
    v->screen = t->screen.data();
    v->atom_type = t->types.data();
    v->nb_parm_index = t->nb_index.data();
    v->lj_acoef = t->lj_a.data();
    v->lj_bcoef = t->lj_b.data();
    v->n_lj = (int32_t)std::min(t->lj_a.size(), t->lj_b.size());
    v->residue_pointers = t->res_ptr.data();
    v->n_bonds = (int32_t)(t->bonds.size() / 3);
    v->bonds = t->bonds.data();
    v->bond_k = t->bond_k.data();
    v->bond_r0 = t->bond_r0.data();
    v->n_bond_types = (int32_t)std::min(t->bond_k.size(), t->bond_r0.size());
    v->n_angles = (int32_t)(t->angles.size() / 4);
    v->angles = t->angles.data();
    v->angle_k = t->angle_k.data();
    v->angle_theta0 = t->angle_t0.data();
    v->n_angle_types = (int32_t)std::min(t->angle_k.size(), t->angle_t0.size());
    v->n_dihedrals = (int32_t)(t->dihedrals.size() / 6);
    v->dihedrals = t->dihedrals.data();
    v->dihedral_pk = t->dih_pk.data();
    v->dihedral_pn = t->dih_pn.data();
    v->dihedral_phase = t->dih_ph.data();
    v->n_dihedral_types =
        (int32_t)std::min(t->dih_pk.size(), std::min(t->dih_pn.size(), t->dih_ph.size()));
    v->num_excluded = t->num_excl.data();
    v->excluded_atoms = t->excl.data();
    v->n_excluded = (int64_t)t->excl.size();
    v->scee = t->scee;
    v->scnb = t->scnb;
    return MMPBSA_OK;
}

int mmpbsa_prmtop_info_get(const mmpbsa_prmtop *t, mmpbsa_prmtop_info *info)
{
    if (!t || !info) return MMPBSA_ERR_INVALID_ARG;
    info->n_atoms = t->n_atoms;
    info->n_residues = t->n_residues;
    info->n_types = t->n_types;
    info->masses = t->masses.data();
    info->atom_names = t->atom_names.data();
    info->residue_labels = t->res_labels.data();
    return MMPBSA_OK;
}

}  // extern "C"
