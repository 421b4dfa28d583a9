// trajectory.cpp -- trajectory ingest (host): DCD and AMBER mdcrd readers and
// the reimaging of periodic frames, ahead of the device path.
//   DCD:   crates/core/src/trajectory/dcd.rs:55-464 (CHARMM/NAMD Fortran
//          records; the first record's length (84) gives the byte order of the
//          record markers and header integers, coordinates and cell values
//          are read little-endian as there; coordinates float32 A -> nm)
//   mdcrd: crates/mmpbsa/src/mdcrd.rs:26-108 (title line, 8-column fields)
//   reimage_complex: crates/mmpbsa/src/binding.rs:234-308
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../../include/mmpbsa_hip.h"
#include "trajectory.h"

namespace mmpbsa {

namespace {

constexpr double kAngstromToNm = 0.1;  // dcd.rs:11

template <typename T>
T from_bytes(const unsigned char *b, bool big)
{
    unsigned char t[sizeof(T)];
    for (size_t k = 0; k < sizeof(T); ++k) t[k] = big ? b[sizeof(T) - 1 - k] : b[k];
    T v;
    std::memcpy(&v, t, sizeof(T));
    return v;
}

}  // namespace

// ---- DCD ----------------------------------------------------------------------
void DcdFile::read_exact(void *p, size_t n, const char *what)
{
    if (std::fread(p, 1, n, f) != n) throw TrajError(std::string("Failed to read ") + what + ": unexpected end of file");
}

int32_t DcdFile::read_int(const char *what)
{
    unsigned char b[4];
    read_exact(b, 4, what);
    return from_bytes<int32_t>(b, hd.big_endian != 0);
}

DcdFile::DcdFile(const char *path)
{
    f = std::fopen(path, "rb");
    if (!f) throw TrajError(std::string("Failed to open DCD file: ") + std::strerror(errno));
    try {
        read_header();
    } catch (...) {
        std::fclose(f);
        f = nullptr;
        throw;
    }
}

DcdFile::~DcdFile()
{
    if (f) std::fclose(f);
}

void DcdFile::read_header()
{
    unsigned char b[4];
    read_exact(b, 4, "block size");
    const int32_t le = from_bytes<int32_t>(b, false), be = from_bytes<int32_t>(b, true);
    if (le == 84)
        hd.big_endian = 0;
    else if (be == 84)
        hd.big_endian = 1;
    else
        throw TrajError("Invalid DCD header: first block size is " + std::to_string(le) + " (LE) or " +
                        std::to_string(be) + " (BE), expected 84");
    char magic[4];
    read_exact(magic, 4, "magic");
    if (std::memcmp(magic, "CORD", 4) != 0) throw TrajError("Invalid DCD magic");
    const int32_t nf = read_int("i32");
    if (nf < 0) throw TrajError("Invalid number of frames: " + std::to_string(nf));
    hd.n_frames = nf;
    hd.start_timestep = read_int("i32");
    hd.timestep_interval = read_int("i32");
    for (int k = 0; k < 6; ++k) (void)read_int("i32");  // n_timesteps, 4 unused, n_fixed
    unsigned char tb[4];
    read_exact(tb, 4, "f32");
    hd.timestep = from_bytes<float>(tb, false);
    hd.has_unit_cell = read_int("i32") != 0 ? 1 : 0;
    unsigned char skip[32];
    read_exact(skip, 32, "header");
    const int32_t charmm = read_int("i32");
    hd.is_charmm = charmm != 0 ? 1 : 0;
    hd.is_64bit = (charmm != 0 && (skip[16] | skip[17] | skip[18] | skip[19]) != 0) ? 1 : 0;
    (void)read_int("i32");  // end of the first record
    const int32_t b2 = read_int("i32");
    if (b2 < 4) throw TrajError("Invalid title block size: " + std::to_string(b2));
    const int32_t nt = read_int("i32");
    if (nt < 0) throw TrajError("Invalid number of titles: " + std::to_string(nt));
    if (4 + (int64_t)nt * 80 > b2)
        throw TrajError("Title block size mismatch: block says " + std::to_string(b2) + " bytes but " +
                        std::to_string(nt) + " titles need " + std::to_string(4 + nt * 80) + " bytes");
    titles.clear();
    for (int32_t k = 0; k < nt; ++k) {
        char t[81] = {0};
        read_exact(t, 80, "title");
        std::string s(t, 80);
        const size_t a = s.find_first_not_of(" \t\r\n");
        const size_t z = s.find_last_not_of(" \t\r\n");
        titles.push_back(a == std::string::npos ? std::string() : s.substr(a, z - a + 1));
    }
    const int32_t rest = b2 - (4 + nt * 80);
    if (rest > 0) std::fseek(f, rest, SEEK_CUR);
    (void)read_int("i32");  // end of the title record
    (void)read_int("i32");  // atom-count record
    const int32_t na = read_int("i32");
    if (na < 0) throw TrajError("Invalid number of atoms: " + std::to_string(na));
    if (na > 100000000)
        throw TrajError("Atom count " + std::to_string(na) + " exceeds maximum supported (100,000,000)");
    hd.n_atoms = na;
    (void)read_int("i32");
    hd.first_frame_offset = std::ftell(f);
    hd.frame_size = (int64_t)(4 * (int64_t)na + 8) * 3 + (hd.has_unit_cell ? 56 : 0);
    current = 0;
}

void DcdFile::seek(int64_t frame)
{
    if (frame < 0 || frame >= hd.n_frames)
        throw TrajError("Frame " + std::to_string(frame) + " out of range (0-" +
                        std::to_string(hd.n_frames - 1) + ")");
    if (std::fseek(f, (long)(hd.first_frame_offset + frame * hd.frame_size), SEEK_SET) != 0)
        throw TrajError("Seek failed");
    current = frame;
}

bool DcdFile::read_frame(double *xyz_nm, double *box6)
{
    if (current >= hd.n_frames) return false;
    const int n = (int)hd.n_atoms;
    if (hd.has_unit_cell) {
        (void)read_int("block size");
        unsigned char b[48];
        read_exact(b, 48, "f64");
        double c[6];
        for (int k = 0; k < 6; ++k) c[k] = from_bytes<double>(b + 8 * k, false);
        (void)read_int("block size");
        // stored a, gamma, b, beta, alpha, c; returned a, b, c (nm), alpha, beta, gamma
        if (box6) {
            box6[0] = c[0] * kAngstromToNm;
            box6[1] = c[2] * kAngstromToNm;
            box6[2] = c[5] * kAngstromToNm;
            box6[3] = c[4];
            box6[4] = c[3];
            box6[5] = c[1];
        }
    }
    buf.resize(4 * (size_t)n);
    for (int d = 0; d < 3; ++d) {
        (void)read_int("block size");
        read_exact(buf.data(), buf.size(), "f32");
        (void)read_int("block size");
        for (int i = 0; i < n; ++i)
            xyz_nm[3 * (size_t)i + d] = (double)from_bytes<float>(buf.data() + 4 * (size_t)i, false) *
                                         kAngstromToNm;
    }
    ++current;
    return true;
}

// ---- mdcrd --------------------------------------------------------------------
MdcrdFile::MdcrdFile(const char *path, int64_t n_atoms_, bool has_box_)
    : n_atoms(n_atoms_), has_box(has_box_)
{
    f = std::fopen(path, "r");
    if (!f) throw TrajError(std::string("Failed to open mdcrd: ") + std::strerror(errno));
    std::string title;
    (void)read_line(title);  // title line
}

MdcrdFile::~MdcrdFile()
{
    if (f) std::fclose(f);
}

bool MdcrdFile::read_line(std::string &line)
{
    line.clear();
    int c;
    bool any = false;
    while ((c = std::fgetc(f)) != EOF) {
        any = true;
        if (c == '\n') break;
        line.push_back((char)c);
    }
    return any;
}

bool MdcrdFile::read_frame(double *xyz_A, double *box3)
{
    const size_t nv = (size_t)n_atoms * 3 + (has_box ? 3 : 0);
    vals.clear();
    std::string line;
    while (vals.size() < nv) {
        if (!read_line(line)) {
            if (vals.empty()) return false;
            throw TrajError("Unexpected EOF at frame " + std::to_string(current) + " (got " +
                            std::to_string(vals.size()) + "/" + std::to_string(nv) + " values)");
        }
        while (!line.empty() && (line.back() == '\r')) line.pop_back();
        for (size_t pos = 0; pos < line.size() && vals.size() < nv; pos += 8) {
            std::string fld = line.substr(pos, 8);
            const size_t a = fld.find_first_not_of(" \t");
            const size_t z = fld.find_last_not_of(" \t");
            const std::string t = a == std::string::npos ? std::string() : fld.substr(a, z - a + 1);
            char *end = nullptr;
            const double v = std::strtod(t.c_str(), &end);
            if (t.empty() || !end || *end)
                throw TrajError("Failed to parse mdcrd value '" + t + "' at frame " +
                                std::to_string(current) + ": invalid float literal");
            vals.push_back(v);
        }
    }
    std::memcpy(xyz_A, vals.data(), sizeof(double) * 3 * (size_t)n_atoms);
    if (box3) {
        if (has_box)
            std::memcpy(box3, vals.data() + 3 * (size_t)n_atoms, sizeof(double) * 3);
        else
            box3[0] = box3[1] = box3[2] = 0.0;
    }
    ++current;
    return true;
}

// ---- reimage_complex (binding.rs:234-308) ---------------------------------------
void reimage_complex(double *xyz, int n, const int *bond_pairs, int n_bonds, int stride,
                     const double box[3])
{
    if (n <= 0) return;
    std::vector<std::vector<int>> adj(n);
    for (int b = 0; b < n_bonds; ++b) {
        const int i = bond_pairs[(size_t)stride * b], j = bond_pairs[(size_t)stride * b + 1];
        if (i >= 0 && i < n && j >= 0 && j < n) {
            adj[i].push_back(j);
            adj[j].push_back(i);
        }
    }
    std::vector<char> seen(n, 0);
    std::vector<std::vector<int>> comps;
    std::deque<int> q;
    for (int s = 0; s < n; ++s) {
        if (seen[s]) continue;
        std::vector<int> comp;
        q.push_back(s);
        seen[s] = 1;
        while (!q.empty()) {
            const int cur = q.front();
            q.pop_front();
            comp.push_back(cur);
            for (int nb : adj[cur]) {
                if (seen[nb]) continue;
                seen[nb] = 1;
                for (int d = 0; d < 3; ++d) {
                    const double diff = xyz[3 * (size_t)nb + d] - xyz[3 * (size_t)cur + d];
                    if (diff > box[d] * 0.5)
                        xyz[3 * (size_t)nb + d] -= box[d];
                    else if (diff < -box[d] * 0.5)
                        xyz[3 * (size_t)nb + d] += box[d];
                }
                q.push_back(nb);
            }
        }
        comps.push_back(std::move(comp));
    }
    if (comps.size() <= 1) return;
    auto center = [&](const std::vector<int> &c, double out[3]) {
        out[0] = out[1] = out[2] = 0.0;
        for (int i : c)
            for (int d = 0; d < 3; ++d) out[d] += xyz[3 * (size_t)i + d];
        for (int d = 0; d < 3; ++d) out[d] /= (double)c.size();
    };
    double ref[3];
    center(comps[0], ref);
    for (size_t k = 1; k < comps.size(); ++k) {
        double cc[3], sh[3] = {0.0, 0.0, 0.0};
        center(comps[k], cc);
        for (int d = 0; d < 3; ++d) {
            const double diff = cc[d] - ref[d];
            if (diff > box[d] * 0.5)
                sh[d] = -box[d];
            else if (diff < -box[d] * 0.5)
                sh[d] = box[d];
        }
        for (int i : comps[k])
            for (int d = 0; d < 3; ++d) xyz[3 * (size_t)i + d] += sh[d];
    }
}

}  // namespace mmpbsa

// ---- C ABI --------------------------------------------------------------------
struct mmpbsa_dcd {
    explicit mmpbsa_dcd(const char *p) : f(p) {}
    mmpbsa::DcdFile f;
};
struct mmpbsa_mdcrd {
    mmpbsa_mdcrd(const char *p, int64_t n, bool b) : f(p, n, b) {}
    mmpbsa::MdcrdFile f;
};

namespace {
thread_local std::string g_traj_err;

template <typename F>
int traj_guard(F &&fn)
{
    try {
        fn();
        return MMPBSA_OK;
    } catch (const mmpbsa::TrajError &e) {
        g_traj_err = e.what();
        return MMPBSA_ERR_IO;
    } catch (const std::bad_alloc &) {
        g_traj_err = "host allocation failed";
        return MMPBSA_ERR_OOM;
    } catch (const std::exception &e) {
        g_traj_err = e.what();
        return MMPBSA_ERR_IO;
    }
}
}  // namespace

extern "C" {

const char *mmpbsa_traj_last_error(void) { return g_traj_err.c_str(); }

int mmpbsa_dcd_open(const char *path, mmpbsa_dcd **out)
{
    if (!path || !out) return MMPBSA_ERR_INVALID_ARG;
    *out = nullptr;
    return traj_guard([&] { *out = new mmpbsa_dcd(path); });
}

int mmpbsa_dcd_close(mmpbsa_dcd *d)
{
    delete d;
    return MMPBSA_OK;
}

int mmpbsa_dcd_get_header(const mmpbsa_dcd *d, mmpbsa_dcd_header *out, int32_t *n_titles)
{
    if (!d) return MMPBSA_ERR_INVALID_ARG;
    if (out) *out = d->f.hd;
    if (n_titles) *n_titles = (int32_t)d->f.titles.size();
    return MMPBSA_OK;
}

int mmpbsa_dcd_title(const mmpbsa_dcd *d, int32_t k, char *buf, int32_t cap)
{
    if (!d || !buf || cap <= 0 || k < 0 || k >= (int32_t)d->f.titles.size())
        return MMPBSA_ERR_INVALID_ARG;
    std::strncpy(buf, d->f.titles[k].c_str(), (size_t)cap - 1);
    buf[cap - 1] = 0;
    return MMPBSA_OK;
}

int mmpbsa_dcd_seek(mmpbsa_dcd *d, int64_t frame)
{
    if (!d) return MMPBSA_ERR_INVALID_ARG;
    return traj_guard([&] { d->f.seek(frame); });
}

int64_t mmpbsa_dcd_current(const mmpbsa_dcd *d) { return d ? d->f.current : -1; }

int mmpbsa_dcd_read_frame(mmpbsa_dcd *d, double *xyz_nm, double *box6, int32_t *got)
{
    if (!d || !xyz_nm || !got) return MMPBSA_ERR_INVALID_ARG;
    *got = 0;
    return traj_guard([&] { *got = d->f.read_frame(xyz_nm, box6) ? 1 : 0; });
}

int mmpbsa_mdcrd_open(const char *path, int64_t n_atoms, int32_t has_box, mmpbsa_mdcrd **out)
{
    if (!path || !out || n_atoms < 0) return MMPBSA_ERR_INVALID_ARG;
    *out = nullptr;
    return traj_guard([&] { *out = new mmpbsa_mdcrd(path, n_atoms, has_box != 0); });
}

int mmpbsa_mdcrd_close(mmpbsa_mdcrd *m)
{
    delete m;
    return MMPBSA_OK;
}

int64_t mmpbsa_mdcrd_current(const mmpbsa_mdcrd *m) { return m ? m->f.current : -1; }

int mmpbsa_mdcrd_read_frame(mmpbsa_mdcrd *m, double *xyz_A, double *box3, int32_t *got)
{
    if (!m || !xyz_A || !got) return MMPBSA_ERR_INVALID_ARG;
    *got = 0;
    return traj_guard([&] { *got = m->f.read_frame(xyz_A, box3) ? 1 : 0; });
}

int mmpbsa_reimage(double *xyz, int32_t n_atoms, const int32_t *bonds, int32_t n_bonds,
                   const double box[3])
{
    if ((!xyz && n_atoms > 0) || (!bonds && n_bonds > 0) || !box) return MMPBSA_ERR_INVALID_ARG;
    mmpbsa::reimage_complex(xyz, n_atoms, bonds, n_bonds, 2, box);
    return MMPBSA_OK;
}

}  // extern "C"
