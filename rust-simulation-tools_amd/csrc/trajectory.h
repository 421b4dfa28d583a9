// trajectory.h -- host readers of trajectory files and periodic reimaging
// (trajectory.cpp).
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mmpbsa_hip.h"

namespace mmpbsa {

struct TrajError : std::runtime_error {
    explicit TrajError(const std::string &m) : std::runtime_error(m) {}
};

// A DCD file: header read at open, frames streamed (dcd.rs:55-135).
struct DcdFile {
    explicit DcdFile(const char *path);
    ~DcdFile();
    DcdFile(const DcdFile &) = delete;
    DcdFile &operator=(const DcdFile &) = delete;
    void seek(int64_t frame);
    // positions [n_atoms][3] in nm, box6 (may be null) a, b, c (nm), alpha,
    // beta, gamma (degrees) when the file has a unit cell; false at the end
    bool read_frame(double *xyz_nm, double *box6);

    mmpbsa_dcd_header hd{};
    std::vector<std::string> titles;
    int64_t current = 0;

private:
    void read_header();
    void read_exact(void *p, size_t n, const char *what);
    int32_t read_int(const char *what);
    std::FILE *f = nullptr;
    std::vector<unsigned char> buf;
};

// An AMBER ASCII trajectory (mdcrd.rs:26-108).
struct MdcrdFile {
    MdcrdFile(const char *path, int64_t n_atoms, bool has_box);
    ~MdcrdFile();
    MdcrdFile(const MdcrdFile &) = delete;
    MdcrdFile &operator=(const MdcrdFile &) = delete;
    // positions [n_atoms][3] in A, box3 (may be null) in A (zeros without a
    // box); false at the end
    bool read_frame(double *xyz_A, double *box3);

    int64_t n_atoms = 0;
    bool has_box = false;
    int64_t current = 0;

private:
    bool read_line(std::string &line);
    std::FILE *f = nullptr;
    std::vector<double> vals;
};

// reimage_complex (binding.rs:234-308) on AoS coordinates in place; bonds are
// (i, j) at bond_pairs[stride * b], [stride * b + 1].
void reimage_complex(double *xyz, int n, const int *bond_pairs, int n_bonds, int stride,
                     const double box[3]);

}  // namespace mmpbsa
