// Host-side model set-up of libmocgpu (native table builders): shared types.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace mocmodel {

struct Error : std::runtime_error {
    explicit Error(const std::string& m) : std::runtime_error(m) {}
};
struct EndOfFile {};

std::vector<std::string> records(const std::string& path);
std::vector<std::string> toks(const std::string& line);
double fnum(const std::string& t);
long inum(const std::string& t);

// Fortran list-directed input over a file's records
struct ListReader {
    std::vector<std::string> rec;
    size_t k = 0;
    ListReader() = default;
    explicit ListReader(const std::string& path);
    static std::vector<std::string> ltoks(const std::string& r);
    bool eof() const;
    std::vector<std::string> read(size_t n);
    std::vector<double> nums(size_t n);
};

// ph1(i, Z-1, N-1, shell-1): Verner et al. (1996) shell fits, [6][30][30][7]
#define PH1(i, z, n, s) ((((size_t)(i) * 30 + (z)) * 30 + (n)) * 7 + (s))

// The numbers of the reference's atomic data files, in the slots the
// reference reads them into.
struct AtomicData {
    std::vector<double> ph1, ph2;                 // [6][30][30][7], [7][30][30]
    std::vector<double> rrb, drb_c, drb_e;        // [6|9][31][31] by (Z, N)
    std::vector<double> rrec, rnew;               // [2|4][31][31]
    std::vector<double> diel, aldro;              // rows of 8 / 6
    int nDiel = 0, nAldro = 0;
    std::vector<double> cion_cf, cion_th;         // [5][31][31], [31][32]
    std::vector<int32_t> nauger;                  // [31][32][11]
    std::vector<double> auger;                    // [31][32][11][11]
    std::vector<int32_t> atom_idx;                // [31*32]
    std::vector<int64_t> atom_meta;               // [nAtoms][7]
    std::vector<double> atom_atp, atom_logt, atom_g, atom_e, atom_A, atom_qom;
    std::vector<int32_t> atom_rec_kind, atom_rec_valid;
    std::vector<double> atom_rec;
    std::vector<double> gam_tk[3], gam_nu[3], gam_val[3];  // H I, He I, He II
    std::vector<double> hei2phot;                 // [41][2]
    std::vector<int32_t> hl_map;                  // [9][12]
    int nHlFiles = 0;
    std::vector<double> hl_dens, hl_lines;        // [files][14], [files][14][16][9]
    std::vector<double> hbA, hbB;                 // [9][4]
    std::vector<double> rbEdge, r2aEdge;          // [2][3][26], [3][18]
    std::vector<double> heiRec;                   // [34][3][4]
    std::vector<double> heiiLyman;                // [4][2]
};

void read_atomic_data(const std::string& dir, AtomicData& d);

}  // namespace mocmodel
