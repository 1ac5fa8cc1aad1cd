// potential.cpp -- reader for "<El>.<xc>.in" potential files.
//
// Semantics of rwpara's read half (src/util.F90:459-1108) for one element:
//   * 14 header lines (line 1 ignored; element, xc and gauss-smearing are checked,
//     smearing to 1e-6, src/util.F90:581-597);
//   * then "<type> <norder> <nd>" / payload record pairs; a clu record sets the
//     coefficient of every enumerated B function whose index tuple matches, records that
//     match nothing are ignored, records with nd > ndensity are skipped;
//   * defaults: crad = identity, core = 3.0, par = 1.0, c = 0 (src/util.F90:282-332).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "magace.hpp"

namespace magace {

namespace {

std::vector<std::string> split(const std::string& line) {
    std::vector<std::string> out;
    std::string cur;
    for (char ch : line) {
        if (ch == ' ' || ch == '\t' || ch == ',' || ch == '\r' || ch == '\n') {
            if (!cur.empty()) out.push_back(cur), cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

double to_d(const std::string& s) {
    std::string t = s;
    for (char& c : t)
        if (c == 'd' || c == 'D') c = 'e';
    return std::strtod(t.c_str(), nullptr);
}

template <size_t N>
uint64_t tuple_key(const std::array<int, N>& a) {
    uint64_t k = 1469598103934665603ull;
    for (int v : a) k = (k ^ (uint64_t)(v + 1)) * 1099511628211ull;
    return k;
}

}  // namespace

bool parse_potential(const std::string& path, const std::string& element, int xc, double smear,
                     Potential& pot, std::string& err) {
    std::ifstream in(path);
    if (!in) { err = "File not found: " + path; return false; }
    const char* xcname = xc == 1 ? "pbe" : xc == 2 ? "lda" : nullptr;
    if (!xcname) { err = "Unknown XC functional received by acelibini."; return false; }
    std::string line;
    std::vector<std::vector<std::string>> hdr;
    pot.header_lines.clear();
    for (int i = 0; i < 14; ++i) {
        if (!std::getline(in, line)) { err = "malformed potential header in " + path; return false; }
        pot.header_lines.push_back(line);
        hdr.push_back(split(line));
    }
    pot.element = element;
    pot.xc = xcname;
    pot.smear = smear;
    auto need = [&](int row, size_t n) { return hdr[row].size() >= n; };
    if (!need(1, 7) || !need(2, 2) || !need(3, 2) || !need(4, 2) || !need(5, 3) || !need(6, 10)) {
        err = "malformed potential header in " + path;
        return false;
    }
    if (hdr[2][1] != element) { err = "Inconsistent element information in file: " + element; return false; }
    if (hdr[3][1] != xcname) { err = "Inconsistent XC functional in file: " + element; return false; }
    if (std::fabs(smear - to_d(hdr[4][1])) > 1e-6) { err = "Inconsistent gauss-smearing in file: " + element; return false; }
    Header& h = pot.h;
    h.lambda_a = to_d(hdr[5][1]);
    h.lambda_m = to_d(hdr[5][2]);
    h.cut_a = to_d(hdr[6][4]);
    h.cut_m = to_d(hdr[6][5]);
    h.dcut_a = to_d(hdr[6][6]);
    h.dcut_m = to_d(hdr[6][7]);
    h.ecut = to_d(hdr[6][8]);
    h.decut = to_d(hdr[6][9]);
    int* rows[5][4] = {{&h.nb_a, &h.nb_m, &h.nr2_m, &h.l2m}, {&h.nr3_a, &h.nr3_m, &h.l3a, &h.l3m},
                       {&h.nr4_a, &h.nr4_m, &h.l4a, &h.l4m}, {&h.nr5_a, &h.nr5_m, &h.l5a, &h.l5m},
                       {&h.nr6_a, &h.nr6_m, &h.l6a, &h.l6m}};
    for (int r = 0; r < 5; ++r) {
        if (!need(7 + r, 6)) { err = "malformed potential header in " + path; return false; }
        for (int c = 0; c < 4; ++c) *rows[r][c] = std::atoi(hdr[7 + r][2 + c].c_str());
    }
    if (!need(12, 2) || !need(13, 5)) { err = "malformed potential header in " + path; return false; }
    h.ndensity = std::atoi(hdr[12][1].c_str());
    h.npot = std::atoi(hdr[13][1].c_str());
    h.nfemb = std::atoi(hdr[13][2].c_str());
    h.radtype_a = std::atoi(hdr[13][3].c_str());
    h.radtype_m = std::atoi(hdr[13][4].c_str());
    if (!check_header(h, err)) return false;

    pot.basis = enumerate_basis(h);
    const Basis& B = pot.basis;
    if (!B.undefined.empty()) { err = B.undefined; return false; }
    const int ndn = h.ndensity;
    pot.c1.assign(B.b1.size() * ndn, 0.0);
    pot.c2.assign(B.b2.size() * ndn, 0.0);
    pot.c3.assign(B.b3.size() * ndn, 0.0);
    pot.c4.assign(B.b4.size() * ndn, 0.0);
    pot.c5.assign(B.b5.size() * ndn, 0.0);
    pot.c6.assign(B.b6.size() * ndn, 0.0);
    pot.crad_a.assign((size_t)h.nr3_a * (h.l3a + 1) * h.nb_a, 0.0);
    for (int n = 0; n < h.nr3_a; ++n)
        for (int l = 0; l <= h.l3a; ++l) pot.crad_a[((size_t)n * (h.l3a + 1) + l) * h.nb_a + n] = 1.0;
    pot.crad_m.assign((size_t)h.nr2_m * (h.l2m + 1) * h.nb_m, 0.0);
    for (int n = 0; n < h.nr2_m; ++n)
        for (int l = 0; l <= h.l2m; ++l) pot.crad_m[((size_t)n * (h.l2m + 1) + l) * h.nb_m + n] = 1.0;
    pot.p1[0][0] = pot.p1[0][1] = pot.p1[1][0] = pot.p1[1][1] = 1.0;
    pot.core[0] = pot.core[1] = 3.0;

    // index tuple -> B function positions (a tuple can only occur once per order)
    std::multimap<uint64_t, size_t> m2, m3, m4, m5, m6;
    for (size_t k = 0; k < B.b2.size(); ++k) m2.emplace(tuple_key(B.b2[k]), k);
    for (size_t k = 0; k < B.b3.size(); ++k) m3.emplace(tuple_key(B.b3[k]), k);
    for (size_t k = 0; k < B.b4.size(); ++k) m4.emplace(tuple_key(B.b4[k]), k);
    for (size_t k = 0; k < B.b5.size(); ++k) m5.emplace(tuple_key(B.b5[k]), k);
    for (size_t k = 0; k < B.b6.size(); ++k) m6.emplace(tuple_key(B.b6[k]), k);
    auto set_match = [](auto& map, const auto& basis, const auto& tup, std::vector<double>& c, size_t o, double v) {
        auto rng = map.equal_range(tuple_key(tup));
        for (auto it = rng.first; it != rng.second; ++it)
            if (basis[it->second] == tup) c[o + it->second] = v;
    };

    while (std::getline(in, line)) {
        auto t = split(line);
        if (t.size() < 3) continue;
        const std::string type = t[0];
        const int norder = std::atoi(t[1].c_str()), nd = std::atoi(t[2].c_str());
        std::string pl;
        if (type != "clu" && type != "par" && type != "rad_atomic" && type != "rad_magnetic" && type != "core")
            continue;
        if (!std::getline(in, pl)) break;
        auto p = split(pl);
        if (type == "clu") {
            if (norder < 1 || norder > 6 || (int)p.size() < norder + 2) continue;
            bool elok = true;
            for (int e = 0; e < norder; ++e) elok = elok && p[e] == element;
            if (!elok && norder >= 2) {
                err = "Read incorrect elements for interaction of order: " + std::to_string(norder);
                return false;
            }
            if (!elok || nd < 1 || nd > h.ndensity) continue;
            const int ni = (int)p.size() - norder - 2;
            std::vector<int> iv(ni > 0 ? ni : 0);
            for (int i = 0; i < ni; ++i) iv[i] = std::atoi(p[norder + i].c_str());
            const double v = to_d(p[p.size() - 2]);
            if (norder == 1 && ni >= 1) {
                for (size_t k = 0; k < B.b1.size(); ++k)
                    if (B.b1[k][0] == iv[0]) pot.c1[(nd - 1) * B.b1.size() + k] = v;
            } else if (norder == 2 && ni >= 4) {
                std::array<int, 4> a;
                std::copy(iv.begin(), iv.begin() + 4, a.begin());
                set_match(m2, B.b2, a, pot.c2, (nd - 1) * B.b2.size(), v);
            } else if (norder == 3 && ni >= 9) {
                std::array<int, 9> a;
                std::copy(iv.begin(), iv.begin() + 9, a.begin());
                set_match(m3, B.b3, a, pot.c3, (nd - 1) * B.b3.size(), v);
            } else if (norder == 4 && ni >= 15) {
                std::array<int, 15> a;
                std::copy(iv.begin(), iv.begin() + 15, a.begin());
                set_match(m4, B.b4, a, pot.c4, (nd - 1) * B.b4.size(), v);
            } else if (norder == 5 && ni >= 21) {
                std::array<int, 21> a;
                std::copy(iv.begin(), iv.begin() + 21, a.begin());
                set_match(m5, B.b5, a, pot.c5, (nd - 1) * B.b5.size(), v);
            } else if (norder == 6 && ni >= 27) {
                std::array<int, 27> a;
                std::copy(iv.begin(), iv.begin() + 27, a.begin());
                set_match(m6, B.b6, a, pot.c6, (nd - 1) * B.b6.size(), v);
            }
        } else if (type == "par") {
            if (norder < 1 || norder > 2) { err = "norder < nvarp1 required for par."; return false; }
            if (p.size() >= 2 && p[0] == element && nd >= 1 && nd <= h.ndensity)
                pot.p1[nd - 1][norder - 1] = to_d(p[1]);
        } else if (type == "rad_atomic" || type == "rad_magnetic") {
            if (p.size() < 6) continue;
            if (p[0] != element || p[1] != element) { err = "Error when reading radial parameters."; return false; }
            const int a = std::atoi(p[2].c_str()), b = std::atoi(p[3].c_str()), l = std::atoi(p[4].c_str());
            const double v = to_d(p[5]);
            if (type == "rad_atomic") {
                if (a >= 1 && a <= h.nb_a && b >= 1 && b <= h.nr3_a && l >= 0 && l <= h.l3a)
                    pot.crad_a[((size_t)(b - 1) * (h.l3a + 1) + l) * h.nb_a + (a - 1)] = v;
            } else {
                if (a >= 1 && a <= h.nb_m && b >= 1 && b <= h.nr2_m && l >= 0 && l <= h.l2m)
                    pot.crad_m[((size_t)(b - 1) * (h.l2m + 1) + l) * h.nb_m + (a - 1)] = v;
            }
        } else if (type == "core") {
            if (norder < 1 || norder > 2) { err = "norder < 3 required for core."; return false; }
            if (p.size() < 3) continue;
            if (p[0] != element || p[1] != element) { err = "Error when reading radial parameters."; return false; }
            pot.core[norder - 1] = to_d(p[2]);
        }
    }
    return true;
}

}  // namespace magace

namespace magace {

bool write_potential(const std::string& path, const Potential& pot, std::string& err) {
    std::ofstream out(path);
    if (!out) {
        err = "cannot write " + path;
        return false;
    }
    out.precision(17);
    const Header& h = pot.h;
    const std::string& E = pot.element;
    for (const auto& l : pot.header_lines) out << l << "\n";
    const int nd = h.ndensity;
    for (int d = 0; d < nd; ++d)
        for (int k = 0; k < 2; ++k)
            out << " par " << k + 1 << " " << d + 1 << "\n " << E << " " << pot.p1[d][k] << " F\n";
    for (int k = 0; k < 2; ++k)
        out << " core " << k + 1 << " 0\n " << E << " " << E << " " << std::fabs(pot.core[k]) << " F\n";
    for (int nb = 1; nb <= h.nb_a; ++nb)
        for (int nr = 1; nr <= h.nr3_a; ++nr)
            for (int l = 0; l <= h.l3a; ++l)
                out << " rad_atomic 0 0\n " << E << " " << E << " " << nb << " " << nr << " " << l << " "
                    << pot.crad_a[((size_t)(nr - 1) * (h.l3a + 1) + l) * h.nb_a + (nb - 1)] << " T\n";
    for (int nb = 1; nb <= h.nb_m; ++nb)
        for (int nr = 1; nr <= h.nr2_m; ++nr)
            for (int l = 0; l <= h.l2m; ++l)
                out << " rad_magnetic 0 0\n " << E << " " << E << " " << nb << " " << nr << " " << l << " "
                    << pot.crad_m[((size_t)(nr - 1) * (h.l2m + 1) + l) * h.nb_m + (nb - 1)] << " T\n";
    auto clu = [&](int order, const int* idx, int n, double c, int d) {
        out << " clu " << order << " " << d + 1 << "\n";
        for (int i = 0; i < order; ++i) out << " " << E;
        for (int i = 0; i < n; ++i) out << " " << idx[i];
        out << " " << c << " T\n";
    };
    const Basis& B = pot.basis;
    for (int d = 0; d < nd; ++d) {
        for (size_t k = 0; k < B.b1.size(); ++k) clu(1, B.b1[k].data(), 1, pot.c1[d * B.b1.size() + k], d);
        for (size_t k = 0; k < B.b2.size(); ++k) clu(2, B.b2[k].data(), 4, pot.c2[d * B.b2.size() + k], d);
        for (size_t k = 0; k < B.b3.size(); ++k) clu(3, B.b3[k].data(), 9, pot.c3[d * B.b3.size() + k], d);
        for (size_t k = 0; k < B.b4.size(); ++k) clu(4, B.b4[k].data(), 15, pot.c4[d * B.b4.size() + k], d);
        for (size_t k = 0; k < B.b5.size(); ++k) clu(5, B.b5[k].data(), 21, pot.c5[d * B.b5.size() + k], d);
        for (size_t k = 0; k < B.b6.size(); ++k) clu(6, B.b6[k].data(), 27, pot.c6[d * B.b6.size() + k], d);
    }
    if (!out) {
        err = "write failed: " + path;
        return false;
    }
    return true;
}

}  // namespace magace
