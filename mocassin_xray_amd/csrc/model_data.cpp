// Readers of MOCASSIN's atomic data files (host C++).
//
// Each reader follows the read sequence of the reference routine that
// consumes the file, so that every number lands in the slot the reference
// reads it into; the files themselves are the reference's data/ directory.
// Fortran list-directed input is restated by ListReader: each read starts on
// a new record, spans as many records as it needs, skips blank ones and drops
// what is left of its last record; '!' starts a comment.

#include "model_internal.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <set>

namespace mocmodel {

// ---------------------------------------------------------------- text input
std::vector<std::string> records(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("can't open data file " + path);
    std::string raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    // universal newlines: \r\n and \r end a record as \n does
    std::vector<std::string> out;
    std::string cur;
    for (size_t i = 0; i < raw.size(); ++i) {
        const char c = raw[i];
        if (c == '\r') {
            out.push_back(cur);
            cur.clear();
            if (i + 1 < raw.size() && raw[i + 1] == '\n') ++i;
        } else if (c == '\n') {
            out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    out.push_back(cur);
    return out;
}

static bool is_sep(char c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f' || c == ',';
}

std::vector<std::string> toks(const std::string& line) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : line) {
        if (c == 'D') c = 'E';
        if (c == 'd') c = 'e';
        if (is_sep(c)) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

static bool digits(const std::string& s, size_t a, size_t b) {
    if (a >= b) return false;
    for (size_t i = a; i < b; ++i)
        if (s[i] < '0' || s[i] > '9') return false;
    return true;
}

double fnum(const std::string& t) {
    // Fortran also writes exponents without their letter: 2.066-1 is 0.2066
    size_t s = (!t.empty() && (t[0] == '+' || t[0] == '-')) ? 1 : 0;
    size_t e = t.find_first_of("+-", s);
    if (e != std::string::npos && e > s) {
        const std::string m = t.substr(s, e - s);
        const size_t dot = m.find('.');
        bool mant_ok;
        if (dot == std::string::npos) {
            mant_ok = digits(m, 0, m.size());
        } else {
            mant_ok = (dot > 0 && digits(m, 0, dot) &&
                       (dot + 1 == m.size() || digits(m, dot + 1, m.size()))) ||
                      (dot == 0 && digits(m, 1, m.size()));
        }
        if (mant_ok && digits(t, e + 1, t.size()))
            return fnum(t.substr(0, e) + "e" + t.substr(e));
    }
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (t.empty() || end != t.c_str() + t.size()) throw Error("not a number: '" + t + "'");
    return v;
}

long inum(const std::string& t) {
    char* end = nullptr;
    const long v = std::strtol(t.c_str(), &end, 10);
    if (t.empty() || end != t.c_str() + t.size()) throw Error("not an integer: '" + t + "'");
    return v;
}

// int(float(t)): truncation toward zero
static long itrunc(const std::string& t) { return (long)std::trunc(fnum(t)); }

std::vector<std::string> ListReader::ltoks(const std::string& r) {
    return toks(r.substr(0, r.find('!')));
}

bool ListReader::eof() const {
    size_t j = k;
    while (j < rec.size() && ltoks(rec[j]).empty()) ++j;
    return j >= rec.size();
}

std::vector<std::string> ListReader::read(size_t n) {
    std::vector<std::string> out;
    while (out.size() < n) {
        if (k >= rec.size()) throw EndOfFile();
        auto t = ltoks(rec[k]);
        ++k;
        out.insert(out.end(), t.begin(), t.end());
    }
    out.resize(n);
    return out;
}

std::vector<double> ListReader::nums(size_t n) {
    auto t = read(n);
    std::vector<double> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = fnum(t[i]);
    return v;
}

static std::string basename_of(const std::string& p) {
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? p : p.substr(s + 1);
}

static bool file_exists(const std::string& p) {
    std::ifstream f(p);
    return (bool)f;
}

// ---------------------------------------------------------------- the readers
static const int kNTot[30] = {1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5,
                              5, 5, 5, 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 7};

// phInit (ph_mod.f90:523-581): ph1(6,30,30,7) then ph2(7,30,30), free format
static void read_ph(const std::string& dir, AtomicData& d) {
    auto next_in = [](const std::vector<std::string>& recs) {
        std::vector<double> v;
        for (const auto& r : recs)
            for (const auto& t : toks(r)) {
                char* end = nullptr;
                const double x = std::strtod(t.c_str(), &end);
                if (end != t.c_str() + t.size()) throw Error("not a number: '" + t + "'");
                v.push_back(x);
            }
        return v;
    };
    const auto v1 = next_in(records(dir + "/ph1.dat"));
    size_t p = 0;
    d.ph1.assign(6 * 30 * 30 * 7, 0.0);
    for (int j = 1; j <= 30; ++j)
        for (int k = 1; k <= j; ++k) {
            int nt = kNTot[k - 1];
            if (j == k && k > 18) nt = 7;
            if (j == k + 1 && (j == 20 || j == 21 || j == 22 || j == 25 || j == 26)) nt = 7;
            for (int l = 1; l <= nt; ++l)
                for (int i = 0; i < 6; ++i) {
                    if (p >= v1.size()) throw Error("ph1.dat ends early");
                    d.ph1[PH1(i, j - 1, k - 1, l - 1)] = v1[p++];
                }
        }
    const auto v2 = next_in(records(dir + "/ph2.dat"));
    p = 0;
    d.ph2.assign(7 * 30 * 30, 0.0);
    for (int k = 1; k <= 30; ++k)
        for (int j = 1; j <= k; ++j)
            if (k != 15 && k != 17 && k != 19 && (k <= 20 || k == 26))
                for (int i = 0; i < 7; ++i) {
                    if (p >= v2.size()) throw Error("ph2.dat ends early");
                    d.ph2[(i * 30 + (k - 1)) * 30 + (j - 1)] = v2[p++];
                }
}

static bool starts_numeric(const std::string& t) {
    size_t i = (!t.empty() && t[0] == '-') ? 1 : 0;
    return i < t.size() && t[i] >= '0' && t[i] <= '9';
}

// initRRDR (hydro_mod.f90:334-417): Badnell RR and DR fits by (Z, N), M = 1
static void read_badnell(const std::string& dir, AtomicData& d) {
    d.rrb.assign(6 * 31 * 31, 0.0);
    const auto rr = records(dir + "/rrBadnell.dat");
    for (size_t r = 2; r < rr.size(); ++r) {
        const auto t = toks(rr[r]);
        if (t.size() < 6) continue;
        const long z = inum(t[0]), n = inum(t[1]), m = inum(t[2]);
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (size_t i = 4; i < t.size() && i < 10; ++i) v[i - 4] = fnum(t[i]);
        if (z <= 30 && m == 1 && n <= 30)
            for (int i = 0; i < 6; ++i) d.rrb[(i * 31 + z) * 31 + n] = v[i];
    }
    // coefficient block, then energy block, each after a header record
    std::vector<std::vector<std::vector<std::string>>> blocks;
    for (const auto& line : records(dir + "/drBadnell.dat")) {
        const auto t = toks(line);
        if (t.empty()) continue;
        if (!starts_numeric(t[0])) {
            blocks.emplace_back();
            continue;
        }
        if (blocks.empty()) throw Error("drBadnell.dat: data before the first header");
        blocks.back().push_back(t);
    }
    if (blocks.size() < 3) throw Error("drBadnell.dat: fewer than three blocks");
    d.drb_c.assign(9 * 31 * 31, 0.0);
    d.drb_e.assign(9 * 31 * 31, 0.0);
    auto put = [](std::vector<double>& arr, const std::vector<std::string>& t) {
        const long z = inum(t[0]), n = inum(t[1]), m = inum(t[2]);
        double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (size_t i = 4; i < t.size() && i < 13; ++i) v[i - 4] = fnum(t[i]);
        if (z <= 30 && m == 1 && n <= 30)
            for (int i = 0; i < 9; ++i) arr[(i * 31 + z) * 31 + n] = v[i];
    };
    for (int b = 0; b < 2; ++b)
        for (const auto& t : blocks[b]) put(d.drb_c, t);
    for (const auto& t : blocks[2]) put(d.drb_e, t);
}

// radRecFit (update_mod.f90:3094-3202): power laws, then Verner & Ferland fits
static void read_radrec(const std::string& dir, AtomicData& d) {
    const auto rec = records(dir + "/radrec.dat");
    size_t k = 0;
    auto nxt = [&](int n, double* out) {
        for (;;) {
            if (k >= rec.size()) throw Error("radrec.dat ends early");
            const auto t = toks(rec[k++]);
            if (t.empty()) continue;
            if ((int)t.size() < n) throw Error("radrec.dat: short record");
            for (int i = 0; i < n; ++i) out[i] = fnum(t[i]);
            return;
        }
    };
    d.rrec.assign(2 * 31 * 31, 0.0);
    d.rnew.assign(4 * 31 * 31, 0.0);
    double v[4];
    auto put = [&](std::vector<double>& a, int n, int elem, int ion) {
        nxt(n, v);
        for (int i = 0; i < n; ++i) a[(i * 31 + elem) * 31 + ion] = v[i];
    };
    for (int ion = 4; ion <= 30; ++ion)
        if (ion != 11)
            for (int elem = ion; elem <= 30; ++elem)
                if (elem != 26 || ion >= 14) put(d.rrec, 2, elem, ion);
    for (int ion = 1; ion <= 3; ++ion)
        for (int elem = ion; elem <= 30; ++elem) put(d.rnew, 4, elem, ion);
    for (int elem = 11; elem <= 30; ++elem) put(d.rnew, 4, elem, 11);
    for (int ion = 4; ion <= 10; ++ion) {
        for (int elem = std::max(6, ion); elem <= 8; ++elem) put(d.rnew, 4, elem, ion);
        put(d.rnew, 4, 10, ion);
    }
    for (int ion = 12; ion <= 26; ++ion) put(d.rnew, 4, 26, ion);
}

// dielectronic.dat / aldrovandi.dat: every record of at least ncol numbers
static std::vector<double> read_rows(const std::string& path, int ncol, int* nrows) {
    std::vector<double> out;
    int n = 0;
    for (const auto& line : records(path)) {
        const auto t = toks(line);
        if ((int)t.size() < ncol) continue;
        for (int i = 0; i < ncol; ++i) out.push_back(fnum(t[i]));
        ++n;
    }
    *nrows = n;
    return out;
}

// makeCollIonData (hydro_mod.f90:13-33): Voronov fits, then the thresholds
static void read_cion(const std::string& dir, AtomicData& d) {
    const auto rec = records(dir + "/cion.dat");
    std::vector<std::vector<std::string>> lines;
    for (const auto& r : rec) lines.push_back(toks(r));
    d.cion_cf.assign(5 * 31 * 31, 0.0);
    size_t k = 0;
    for (int elem = 1; elem <= 30; ++elem)
        for (int ion = elem; ion >= 1; --ion) {
            while (k < lines.size() && lines[k].empty()) ++k;
            if (k >= lines.size() || lines[k].size() < 5) throw Error("cion.dat: short fit table");
            for (int i = 0; i < 5; ++i) d.cion_cf[(i * 31 + elem) * 31 + ion] = fnum(lines[k][i]);
            ++k;
        }
    k += 2;  // two comment records
    d.cion_th.assign(31 * 32, 0.0);
    for (int elem = 3; elem <= 30; ++elem) {
        if (k >= lines.size() || (int)lines[k].size() < elem)
            throw Error("cion.dat: short threshold table");
        for (int i = 0; i < elem; ++i) d.cion_th[elem * 32 + 1 + i] = fnum(lines[k][i]);
        ++k;
    }
}

// makeAugerData (hydro_mod.f90:35-53): 1696 (elem, ion, shell, nelec) + 10 yields
static void read_auger(const std::string& dir, AtomicData& d) {
    std::vector<std::vector<std::string>> lines;
    for (const auto& r : records(dir + "/auger.dat")) {
        auto t = toks(r);
        if (!t.empty()) lines.push_back(t);
    }
    d.nauger.assign(31 * 32 * 11, 0);
    d.auger.assign(31 * 32 * 11 * 11, 0.0);
    size_t k = 0;
    for (int n = 0; n < 1696; ++n) {
        if (k + 1 >= lines.size() || lines[k].size() < 4 || lines[k + 1].size() < 10)
            throw Error("auger.dat ends early");
        const long elem = inum(lines[k][0]), ion = inum(lines[k][1]), shell = inum(lines[k][2]);
        const long nelec = inum(lines[k][3]);
        if (elem < 0 || elem > 30 || ion < 0 || ion > 31 || shell < 0 || shell > 10)
            throw Error("auger.dat: index out of range");
        d.nauger[(elem * 32 + ion) * 11 + shell] = (int32_t)nelec;
        for (int i = 0; i < 10; ++i)
            d.auger[((elem * 32 + ion) * 11 + shell) * 11 + 1 + i] = fnum(lines[k + 1][i]);
        k += 2;
    }
}

// one n-level file in the order equilibrium() reads it (emission_mod.f90:2505-2754)
struct AtomFile {
    int nlev = 0, ntemp = 0, iup2p = -1;
    double atp = 0.0;
    std::vector<double> logt, qom /* [nt][nlev+1][nlev+1] */, A /* [nlev+1]^2 */, g, e;
    size_t kend = 0;
};

static AtomFile read_atom(const std::vector<std::string>& rec, const std::string& name) {
    AtomFile a;
    size_t k = 0;
    auto line = [&]() -> const std::string& {
        if (k >= rec.size()) throw Error(name + ": ends early");
        return rec[k++];
    };
    auto tk = [&](size_t need) {
        auto t = toks(line());
        if (t.size() < need) throw Error(name + ": short record");
        return t;
    };
    const long numLines = inum(tk(1)[0]);
    for (long i = 0; i < numLines; ++i) line();
    {
        const auto t = tk(2);
        a.nlev = (int)inum(t[0]);
        a.ntemp = (int)inum(t[1]);
    }
    const int nl = a.nlev, nt = a.ntemp, L = nl + 1;
    if (nl < 1 || nt < 1) throw Error(name + ": bad level or temperature count");
    for (int i = 0; i < nl; ++i) line();
    for (int i = 0; i < nt; ++i) a.logt.push_back(std::log10(fnum(tk(1)[0])));
    line();  // iRats
    a.qom.assign((size_t)nt * L * L, 0.0);
    int ilow1 = 0, iup1 = 0, kt = 0;
    for (;;) {
        const auto t = tk(3);
        int il = (int)inum(t[0]), iu = (int)inum(t[1]);
        const double qx = fnum(t[2]);
        if (qx == 0.0) break;
        if (il == 0) {
            il = ilow1;
            ++kt;
        } else {
            ilow1 = il;
            kt = 1;
        }
        if (iu == 0) iu = iup1;
        else iup1 = iu;
        if (kt < 1 || kt > nt || il < 0 || il > nl || iu < 0 || iu > nl)
            throw Error(name + ": collision strength index out of range");
        a.qom[((size_t)(kt - 1) * L + il) * L + iu] = qx;
    }
    a.A.assign((size_t)L * L, 0.0);
    int jj = 0;
    for (int n = 0; n < nl * (nl - 1) / 2; ++n) {
        const auto t = tk(3);
        int i = (int)inum(t[0]), j = (int)inum(t[1]);
        const double ax = fnum(t[2]);
        if (i < 0 && j < 0) {
            a.atp = ax;
            a.iup2p = std::abs(j);
            i = std::abs(i);
            j = std::abs(j);
            if (i > nl || j > nl) throw Error(name + ": A index out of range");
            if (a.A[(size_t)j * L + i] > 0.0) ++jj;
            a.A[(size_t)j * L + i] += a.atp;
        } else {
            if (i < 0 || j < 0 || i > nl || j > nl) throw Error(name + ": A index out of range");
            a.A[(size_t)j * L + i] = ax;
        }
    }
    for (int n = 0; n < jj; ++n) {
        const auto t = tk(3);
        const int i = (int)inum(t[0]), j = (int)inum(t[1]);
        if (i < 0 || j < 0 || i > nl || j > nl) throw Error(name + ": A index out of range");
        a.A[(size_t)j * L + i] = fnum(t[2]);
    }
    a.g.assign(L, 0.0);
    a.e.assign(L, 0.0);
    for (int n = 0; n < nl; ++n) {
        const auto t = tk(3);
        const int i = (int)inum(t[0]);
        if (i < 0 || i > nl) throw Error(name + ": level index out of range");
        char* end = nullptr;
        a.g[i] = std::strtod(t[1].c_str(), &end);
        if (end != t[1].c_str() + t[1].size()) throw Error(name + ": bad statistical weight");
        a.e[i] = fnum(t[2]);
        if (i > 1 && a.e[i] == a.e[i - 1]) a.e[i] = a.e[i] + 0.1;
    }
    a.kend = k;
    return a;
}

static const std::set<std::string> kRecSpecial = {"ci.dat",   "ciii.dat", "ni.dat",
                                                   "neiv.dat", "oii.dat",  "oiii.dat",
                                                   "nii.dat",  "nev.dat",  "niv.dat"};

// the recombination data after the level block (equilibrium rec=.true.,
// emission_mod.f90:2828-2877): (kind, table[nlev+1][11], valid[nlev+1])
static int read_rec_data(const std::string& base, int nlev, ListReader& lr,
                         std::vector<double>& tab, std::vector<int32_t>& valid) {
    const int L = nlev + 1;
    tab.assign((size_t)L * 11, 0.0);
    valid.assign(L, 0);
    const bool special = kRecSpecial.count(base) > 0;
    bool any = false;
    auto row = [&](long j) -> long {  // a Python index: negative ones count from the end
        if (j < 0) j += L;
        if (j < 0 || j >= L) throw Error(base + ": recombination level out of range");
        return j;
    };
    for (int j = 2; j <= nlev; ++j) {
        if (lr.eof()) break;
        if (special) {
            const auto v = lr.nums(11);
            any = true;
            if (v[1] > 0.0) {
                for (int c = 0; c < 11; ++c) tab[(size_t)j * 11 + c] = v[c];
                valid[j] = 1;
            }
        } else {
            const size_t save = lr.k;
            const long jj0 = itrunc(lr.read(1)[0]);
            if (jj0 == -9999) break;
            lr.k = save;
            std::vector<std::string> v;
            try {
                v = lr.read(5);
            } catch (const EndOfFile&) {  // a truncated last record ends the table
                break;
            }
            const long jj = itrunc(v[0]);
            if (jj > nlev) break;
            any = true;
            const long r = row(jj);
            for (int c = 0; c < 4; ++c) tab[(size_t)r * 11 + c] = fnum(v[1 + c]);
            valid[r] = 1;
        }
    }
    if (any) valid[1] = 1;
    return any ? (special ? 1 : 2) : 0;
}

// makeElements (hydro_mod.f90:679-768): the n-level files named in fileNames.dat
static void read_atoms(const std::string& dir, AtomicData& d) {
    std::vector<std::string> names;
    for (const auto& r : records(dir + "/fileNames.dat")) {
        size_t a = 0, b = r.size();
        while (a < b && (is_sep(r[a]) && r[a] != ',')) ++a;
        while (b > a && (is_sep(r[b - 1]) && r[b - 1] != ',')) --b;
        names.push_back(r.substr(a, b - a));
    }
    d.atom_idx.assign(31 * 32, -1);
    d.atom_meta.clear();
    size_t k = 0;
    for (int elem = 3; elem <= 30; ++elem)
        for (int ion = 1; ion <= std::min(elem + 1, 31); ++ion) {
            if (k >= names.size()) throw Error("fileNames.dat ends early");
            const std::string name = names[k++];
            if (name.empty()) continue;
            const std::string base = basename_of(name);
            const std::string p = dir + "/" + base;
            if (!file_exists(p)) continue;
            const auto rec = records(p);
            AtomFile a = read_atom(rec, base);
            ListReader lr;
            lr.rec = rec;
            lr.k = a.kend;
            std::vector<double> tab;
            std::vector<int32_t> valid;
            const int kind = read_rec_data(base, a.nlev, lr, tab, valid);
            const int nl = a.nlev, nt = a.ntemp, L = nl + 1;
            d.atom_rec_kind.push_back(kind);
            d.atom_rec.insert(d.atom_rec.end(), tab.begin() + 11, tab.end());
            d.atom_rec_valid.insert(d.atom_rec_valid.end(), valid.begin() + 1, valid.end());
            d.atom_idx[elem * 32 + ion] = (int32_t)(d.atom_meta.size() / 7);
            const int64_t meta[7] = {nl, nt, (int64_t)d.atom_logt.size(), (int64_t)d.atom_g.size(),
                                     (int64_t)d.atom_A.size(), (int64_t)d.atom_qom.size(),
                                     a.iup2p};
            d.atom_meta.insert(d.atom_meta.end(), meta, meta + 7);
            d.atom_atp.push_back(a.atp);
            d.atom_logt.insert(d.atom_logt.end(), a.logt.begin(), a.logt.end());
            d.atom_g.insert(d.atom_g.end(), a.g.begin() + 1, a.g.end());
            d.atom_e.insert(d.atom_e.end(), a.e.begin() + 1, a.e.end());
            for (int i = 1; i <= nl; ++i)
                for (int j = 1; j <= nl; ++j) d.atom_A.push_back(a.A[(size_t)i * L + j]);
            // Upsilon is read into a single-precision array
            for (int t = 0; t < nt; ++t)
                for (int i = 1; i <= nl; ++i)
                    for (int j = 1; j <= nl; ++j)
                        d.atom_qom.push_back((double)(float)a.qom[((size_t)t * L + i) * L + j]);
        }
}

// the tables of the diffuse emission (initGammaCont, RecLinesEmission, ...)
static void read_emission_tables(const std::string& dir, AtomicData& d) {
    static const char* kGamma[3] = {"gammaHI.dat", "gammaHeI.dat", "gammaHeII.dat"};
    for (int g = 0; g < 3; ++g) {
        ListReader lr(dir + "/" + kGamma[g]);
        const auto h = lr.read(2);
        const long ntk = itrunc(h[0]), nlim = itrunc(h[1]);
        if (ntk < 1 || nlim < 1) throw Error(std::string(kGamma[g]) + ": bad table size");
        d.gam_tk[g] = lr.nums(ntk);
        d.gam_nu[g].assign(nlim, 0.0);
        d.gam_val[g].assign((size_t)ntk * nlim, 0.0);
        for (long i = 0; i < nlim; ++i) {
            const auto v = lr.nums(ntk + 1);
            d.gam_nu[g][i] = v[0];
            for (long t = 0; t < ntk; ++t) d.gam_val[g][(size_t)t * nlim + i] = v[1 + t];
        }
    }
    {
        ListReader lr(dir + "/HeI2phot.dat");
        d.hei2phot.clear();
        for (int i = 0; i < 41; ++i) {
            const auto v = lr.nums(2);
            d.hei2phot.insert(d.hei2phot.end(), v.begin(), v.end());
        }
    }
    // hydrogenic recombination-line tables, one file per (Z, T bin)
    std::vector<std::string> names;
    for (const auto& r : records(dir + "/hydroLinesFiles.dat")) {
        auto strip = [](std::string s, const char* set) {
            size_t a = s.find_first_not_of(set), b = s.find_last_not_of(set);
            return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
        };
        const char* ws = " \t\n\r\v\f";
        if (strip(r, ws).empty()) continue;
        names.push_back(strip(strip(strip(r, ws), "'\""), ws));
        if (names.size() == 9 * 12) break;
    }
    if (names.size() != 9 * 12) throw Error("hydroLinesFiles.dat: fewer than 108 names");
    std::set<std::string> uset(names.begin(), names.end());
    const std::vector<std::string> uniq(uset.begin(), uset.end());  // byte order, as sorted()
    d.hl_map.resize(9 * 12);
    for (int i = 0; i < 9 * 12; ++i)
        d.hl_map[i] = (int32_t)(std::lower_bound(uniq.begin(), uniq.end(), names[i]) - uniq.begin());
    const size_t nf = uniq.size();
    d.nHlFiles = (int)nf;
    d.hl_dens.assign(nf * 14, 0.0);
    d.hl_lines.assign(nf * 14 * 16 * 9, 0.0);
    for (size_t f = 0; f < nf; ++f) {
        ListReader lr(dir + "/" + basename_of(uniq[f]));
        for (int iden = 0; iden < 14; ++iden) {
            if (lr.eof()) break;
            d.hl_dens[f * 14 + iden] = fnum(lr.read(1)[0]);
            for (int iup = 15; iup >= 2; --iup) {
                const auto v = lr.nums(std::min(8, iup - 1));
                for (size_t c = 0; c < v.size(); ++c)
                    d.hl_lines[((f * 14 + iden) * 16 + iup) * 9 + 1 + c] = v[c];
            }
        }
    }
    {
        ListReader lr(dir + "/gaussfitHb.dat");
        d.hbA.assign(36, 0.0);
        d.hbB.assign(36, 0.0);
        for (int i = 0; i < 9; ++i) {
            const auto v = lr.nums(8);
            for (int c = 0; c < 4; ++c) {
                d.hbA[i * 4 + c] = v[c];
                d.hbB[i * 4 + c] = v[4 + c];
            }
        }
    }
    d.rbEdge.assign(2 * 3 * 26, 0.0);
    static const char* kRb[2] = {"r1bEdge.dat", "r2bEdge.dat"};
    for (int z = 0; z < 2; ++z) {
        ListReader lr(dir + "/" + kRb[z]);
        for (int i = 0; i < 26; ++i) {
            const auto v = lr.nums(3);
            for (int c = 0; c < 3; ++c) d.rbEdge[(z * 3 + c) * 26 + i] = v[c];
        }
    }
    {
        ListReader lr(dir + "/r2aEdge.dat");
        d.r2aEdge.assign(3 * 18, 0.0);
        for (int i = 0; i < 18; ++i) {
            const auto v = lr.nums(3);
            for (int c = 0; c < 3; ++c) d.r2aEdge[c * 18 + i] = v[c];
        }
    }
    {
        ListReader lr(dir + "/HeIrecLines.dat");
        d.heiRec.assign(34 * 3 * 4, 0.0);
        for (int iden = 0; iden < 3; ++iden)
            for (int il = 0; il < 34; ++il) {
                const auto v = lr.nums(4);
                for (int c = 0; c < 4; ++c) d.heiRec[(il * 3 + iden) * 4 + c] = v[c];
                d.heiRec[(il * 3 + iden) * 4] *= 1.e25;
            }
    }
    {
        ListReader lr(dir + "/r2a0100old.dat");
        d.heiiLyman.clear();
        for (int i = 0; i < 4; ++i) {
            const auto v = lr.nums(2);
            d.heiiLyman.insert(d.heiiLyman.end(), v.begin(), v.end());
        }
    }
}

ListReader::ListReader(const std::string& path) : rec(records(path)), k(0) {}

void read_atomic_data(const std::string& dir, AtomicData& d) {
    read_ph(dir, d);
    read_badnell(dir, d);
    read_radrec(dir, d);
    d.diel = read_rows(dir + "/dielectronic.dat", 8, &d.nDiel);
    d.aldro = read_rows(dir + "/aldrovandi.dat", 6, &d.nAldro);
    read_cion(dir, d);
    read_auger(dir, d);
    read_atoms(dir, d);
    read_emission_tables(dir, d);
}

}  // namespace mocmodel
