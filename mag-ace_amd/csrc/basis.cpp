// basis.cpp -- basis enumeration, Clebsch-Gordan couplings and the merged term list.
//
// Enumeration follows the selection rules of setuptables/bcreate (src/util.F90:1801-3760,
// 4673-5330); couplings follow setupcg/setuphighercouplings (src/util.F90:4149-4548).
// Instead of the reference's per-B m-combination loops with stored dB/dA
// (src/b_basis.F90:91-307), every B function is expanded once, at init, into monomials
// w * prod Z[i] of the per-atom feature vector Z = [A | A2 | A0]; identical monomials are
// merged and complex-conjugate partners folded, so the device evaluates
// rho = sum_k c_k B_k and its gradient with one pass over a flat table.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <unordered_map>

#include "magace.hpp"

namespace magace {

bool check_header(const Header& h, std::string& err) {
    auto fail = [&](const char* m) { err = m; return false; };
    if (h.ndensity < 1 || h.ndensity > 2) return fail("ndensity must be 1 or 2 (src/forces.F90:1066-1085)");
    if (h.nfemb < 0 || h.nfemb > 2) return fail("embedding function not implemented (nfemb 0..2)");
    // limit checks of src/util.F90:233-272
    if (h.l3a < h.l4a || h.l3a < h.l5a || h.l3a < h.l6a) return fail(" l3max_atomic needs to be largest.");
    if (h.l2m < h.l3m || h.l2m < h.l4m || h.l2m < h.l5m || h.l2m < h.l6m)
        return fail(" l2max_magnetic needs to be largest.");
    if (h.nr3_a < h.nr4_a || h.nr3_a < h.nr5_a || h.nr3_a < h.nr6_a) return fail(" nradial_atomic needs to be largest.");
    if (h.nr2_m < h.nr3_m || h.nr2_m < h.nr4_m || h.nr2_m < h.nr5_m || h.nr2_m < h.nr6_m)
        return fail(" nradial_magnetic needs to be largest.");
    {
        // nradial = max(nradial3_atomic, nradial2_magnetic) against the orders' maxima
        const int nradial = std::max(h.nr3_a, h.nr2_m);
        if (nradial < std::max(h.nr4_a, h.nr4_m) || nradial < std::max(h.nr5_a, h.nr5_m) ||
            nradial < std::max(h.nr6_a, h.nr6_m))
            return fail(" nradial needs to be largest.");
    }
    if (h.nr3_a > h.nb_a) return fail(" nradial_atomic < nradbase_atomic required");
    if (h.nr2_m > h.nb_m) return fail(" nradial_magnetic < nradbase_magnetic required");
    if (2 * h.l6a > h.l3a || 2 * h.l6m > h.l2m) return fail("6-body couplings not worked for lmax > 6");
    // the reference's A array holds nradial3_magnetic magnetic radial functions
    // (src/a_basis.F90:346); higher orders index it with their own nradial_magnetic
    if (h.nr4_m > h.nr3_m || h.nr5_m > h.nr3_m || h.nr6_m > h.nr3_m)
        return fail("nradial_magnetic of orders 4-6 above nradial3_magnetic (the reference reads outside its A array)");
    // prepB skips the 6-body terms when the header has no 5-body terms (src/b_basis.F90:313-315)
    if ((h.nr6_a || h.nr6_m) && !h.nr5_a && !h.nr5_m)
        return fail("6-body terms without 5-body terms (the reference's prepB skips them)");
    if (h.l2m > 12 || h.l3a > 12) return fail("angular momentum above 12 is not supported");
    if (h.nb_a > 32 || h.nb_m > 32) return fail("more than 32 radial base functions");
    for (int t : {h.radtype_a, h.radtype_m})
        if (t < 3 || t > 8) return fail("radtype must be 3..8 (0-2 stop in the reference)");
    return true;
}

namespace {

// bcreate ordering key for one element (src/util.F90:4728-4730)
struct Keyer {
    int lmax, nradial;
    long operator()(int l, int nr) const { return (long)((lmax + 1) + l) * (nradial + 1) + nr; }
};

bool triangle_even(int a, int b, int c) {
    return a + b >= c && std::abs(a - b) <= c && (a + b + c) % 2 == 0;
}

// couplings L of ((l0 l1)L (l2 l3)L)0 allowed by bcreate rank 4 (src/util.F90:4810-4857)
std::vector<int> rank4_couplings(const Keyer& key, const int nr[4], const int l[4]) {
    std::vector<int> out;
    long k[4];
    for (int i = 0; i < 4; ++i) k[i] = key(l[i], nr[i]);
    int lmx = *std::max_element(l, l + 4);
    for (int L = 0; L <= 2 * lmx; ++L) {
        if (L > l[0] + l[1] || L > l[2] + l[3]) continue;
        if (L < std::abs(l[0] - l[1]) || L < std::abs(l[2] - l[3])) continue;
        if (k[0] == k[1] && (L & 1)) continue;
        if (k[2] == k[3] && (L & 1)) continue;
        out.push_back(L);
    }
    if ((l[0] + l[1] + l[2] + l[3]) & 1) out.clear();
    return out;
}

// bcreate rank 5 ((l0 l1)L12 (l2 l3)L34) l4 (src/util.F90:4960-5070): (L12, L34) pairs
std::vector<std::array<int, 2>> rank5_couplings(const Keyer& key, const int nr[5], const int l[5]) {
    std::vector<std::array<int, 2>> out;
    long k[5];
    for (int i = 0; i < 5; ++i) k[i] = key(l[i], nr[i]);
    const int lmx = *std::max_element(l, l + 5);
    for (int a = 0; a <= 2 * lmx; ++a)
        for (int b = 0; b <= 2 * lmx; ++b) {
            if (a > l[0] + l[1] || b > l[2] + l[3]) continue;
            if (a < std::abs(l[0] - l[1]) || b < std::abs(l[2] - l[3])) continue;
            if (l[4] > a + b || l[4] < std::abs(a - b)) continue;
            if (k[0] == k[1] && (a & 1)) continue;
            if (k[2] == k[3] && (b & 1)) continue;
            if (k[0] == k[2] && k[1] == k[3] && a == b && (l[4] & 1)) continue;
            out.push_back({a, b});
        }
    if ((l[0] + l[1] + l[2] + l[3] + l[4]) & 1) out.clear();
    return out;
}

// bcreate rank 6 ((l0 l1)L12 (l2 l3)L34)L56 (l4 l5)L56 (src/util.F90:5140-5230): (L12, L34, L56)
std::vector<std::array<int, 3>> rank6_couplings(const Keyer& key, const int nr[6], const int l[6]) {
    std::vector<std::array<int, 3>> out;
    long k[6];
    for (int i = 0; i < 6; ++i) k[i] = key(l[i], nr[i]);
    const int lmx = *std::max_element(l, l + 6);
    for (int a = 0; a <= 2 * lmx; ++a)
        for (int b = 0; b <= 2 * lmx; ++b)
            for (int c = 0; c <= 2 * lmx; ++c) {
                if (a > l[0] + l[1] || b > l[2] + l[3] || c > l[4] + l[5]) continue;
                if (a < std::abs(l[0] - l[1]) || b < std::abs(l[2] - l[3]) || c < std::abs(l[4] - l[5])) continue;
                if (c > a + b || c < std::abs(a - b)) continue;
                if (k[0] == k[1] && (a & 1)) continue;
                if (k[2] == k[3] && (b & 1)) continue;
                if (k[4] == k[5] && (c & 1)) continue;
                if (k[0] == k[2] && k[1] == k[3] && a == b && (c & 1)) continue;
                out.push_back({a, b, c});
            }
    if ((l[0] + l[1] + l[2] + l[3] + l[4] + l[5]) & 1) out.clear();
    return out;
}

// Odometer over digits with per-digit ranges [lo, hi], the last digit fastest (the nesting
// order of the reference's loops, outermost first).
struct Odometer {
    std::vector<int> lo, hi, v;
    bool done = false;
    Odometer(std::vector<int> l, std::vector<int> h) : lo(std::move(l)), hi(std::move(h)), v(lo) {
        for (size_t i = 0; i < lo.size(); ++i)
            if (hi[i] < lo[i]) done = true;
    }
    void next() {
        for (int i = (int)v.size() - 1; i >= 0; --i) {
            if (++v[i] <= hi[i]) return;
            v[i] = lo[i];
        }
        done = true;
    }
};

// Within each group of B functions made from one index tuple (the first entry excluded), the
// reference exchanges the internal couplings in columns [c0, c1) so that their sums ascend
// (src/util.F90:3077-3091 for order 5, :3713-3755 for order 6); replayed step by step.
template <size_t N>
void sort_group_lints(std::vector<std::array<int, N>>& v, size_t head, int nint, int c0, int c1) {
    for (int m = 1; m <= nint - 2; ++m)
        for (int k = m + 1; k <= nint - 1; ++k) {
            int s1 = 0, s2 = 0;
            for (int c = c0; c < c1; ++c) {
                s1 += v[head + m][c];
                s2 += v[head + k][c];
            }
            if (s1 > s2)
                for (int c = c0; c < c1; ++c) std::swap(v[head + m][c], v[head + k][c]);
        }
}

// 5-body functions (src/util.F90:2822-3093): index tuples ordered on combined (atomic,
// magnetic) keys, rank-4 atomic and rank-5 magnetic couplings paired one to one
void enumerate5(const Header& h, const Keyer& key, Basis& B) {
    const long kmm = (long)h.nr5_m * (h.l5m + 1);
    auto nc = [&](int na, int la, int nm, int lm) { return (kmm + 1) * ((long)h.nr5_a * la + na) + (h.nr5_m * lm + nm); };
    // digits: n1..n4, n0'..n4', l1..l4, l0'..l4'
    std::vector<int> lo(18, 1), hi(18);
    for (int i = 0; i < 4; ++i) hi[i] = h.nr5_a;
    for (int i = 4; i < 9; ++i) hi[i] = h.nr5_m;
    for (int i = 9; i < 13; ++i) lo[i] = 0, hi[i] = h.l5a;
    for (int i = 13; i < 18; ++i) lo[i] = 0, hi[i] = h.l5m;
    for (Odometer od(lo, hi); !od.done; od.next()) {
        const auto& d = od.v;
        long c[4];
        for (int i = 0; i < 4; ++i) c[i] = nc(d[i], d[9 + i], d[5 + i], d[14 + i]);
        if (c[0] < c[1] || c[1] < c[2] || c[2] < c[3]) continue;
        const int nra[4] = {d[0], d[1], d[2], d[3]}, la[4] = {d[9], d[10], d[11], d[12]};
        const int nrm[5] = {d[4], d[5], d[6], d[7], d[8]}, lm[5] = {d[13], d[14], d[15], d[16], d[17]};
        const std::vector<int> ca = rank4_couplings(key, nra, la);
        const std::vector<std::array<int, 2>> cm = rank5_couplings(key, nrm, lm);
        if (ca.empty() || cm.empty()) continue;
        if (ca.size() != cm.size() && B.undefined.empty())
            B.undefined = "5-body basis pairs " + std::to_string(ca.size()) + " atomic with " +
                          std::to_string(cm.size()) + " magnetic couplings (src/util.F90:3015-3040 fills the "
                          "rest with zero indices, which prepB reads outside its arrays)";
        const size_t head = B.b5.size(), n = std::min(ca.size(), cm.size());
        for (size_t t = 0; t < n; ++t)
            B.b5.push_back({d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11], d[12],
                            d[13], d[14], d[15], d[16], d[17], ca[t], cm[t][0], cm[t][1]});
        sort_group_lints(B.b5, head, (int)n, 19, 21);
    }
}

// 6-body functions (src/util.F90:3439-3760): the ordering keys use the 5-body maxima, as in the
// reference; rank-5 atomic and rank-6 magnetic couplings paired one to one
void enumerate6(const Header& h, const Keyer& key, Basis& B) {
    const long kmm = (long)h.nr5_m * (h.l5m + 1);
    auto nc = [&](int na, int la, int nm, int lm) { return (kmm + 1) * ((long)h.nr6_a * la + na) + (h.nr6_m * lm + nm); };
    // digits: n1..n5, n0'..n5', l1..l5, l0'..l5'
    std::vector<int> lo(22, 1), hi(22);
    for (int i = 0; i < 5; ++i) hi[i] = h.nr6_a;
    for (int i = 5; i < 11; ++i) hi[i] = h.nr6_m;
    for (int i = 11; i < 16; ++i) lo[i] = 0, hi[i] = h.l6a;
    for (int i = 16; i < 22; ++i) lo[i] = 0, hi[i] = h.l6m;
    for (Odometer od(lo, hi); !od.done; od.next()) {
        const auto& d = od.v;
        long c[5], ka[5];
        for (int i = 0; i < 5; ++i) {
            c[i] = nc(d[i], d[11 + i], d[6 + i], d[17 + i]);
            ka[i] = (long)h.nr6_a * d[11 + i] + d[i];
        }
        if (c[0] < c[1] || c[1] < c[2] || c[2] < c[3] || c[3] < c[4]) continue;
        if ((ka[0] < ka[1] || ka[1] < ka[2] || ka[2] < ka[3] || ka[3] < ka[4]) && B.undefined.empty())
            B.undefined = "Error! the atomic basis is not ordered lexicographically (6-body, src/util.F90:3525-3530)";
        const int nra[5] = {d[0], d[1], d[2], d[3], d[4]}, la[5] = {d[11], d[12], d[13], d[14], d[15]};
        const int nrm[6] = {d[5], d[6], d[7], d[8], d[9], d[10]};
        const int lm[6] = {d[16], d[17], d[18], d[19], d[20], d[21]};
        const std::vector<std::array<int, 2>> ca = rank5_couplings(key, nra, la);
        const std::vector<std::array<int, 3>> cm = rank6_couplings(key, nrm, lm);
        if (ca.empty() || cm.empty()) continue;
        if (ca.size() != cm.size() && B.undefined.empty())
            B.undefined = "6-body basis pairs " + std::to_string(ca.size()) + " atomic with " +
                          std::to_string(cm.size()) + " magnetic couplings (src/util.F90:3660-3690 fills the "
                          "rest with zero indices, which prepB reads outside its arrays)";
        const size_t head = B.b6.size(), n = std::min(ca.size(), cm.size());
        for (size_t t = 0; t < n; ++t) {
            std::array<int, 27> e;
            for (int i = 0; i < 22; ++i) e[i] = d[i];
            e[22] = ca[t][0];
            e[23] = ca[t][1];
            e[24] = cm[t][0];
            e[25] = cm[t][1];
            e[26] = cm[t][2];
            B.b6.push_back(e);
        }
        sort_group_lints(B.b6, head, (int)n, 22, 24);
        sort_group_lints(B.b6, head, (int)n, 24, 27);
    }
}

}  // namespace

Basis enumerate_basis(const Header& h) {
    Basis B;
    Keyer key{std::max(h.l3a, h.l2m), std::max(h.nr3_a, h.nr2_m)};
    for (int n = 1; n <= h.nb_m; ++n) B.b1.push_back({n});
    // order 2: rank-2 magnetic (n0', l) (n1', l), ordered
    for (int n0 = 1; n0 <= h.nr2_m; ++n0)
        for (int n1p = 1; n1p <= h.nr2_m; ++n1p)
            for (int n1 = 1; n1 <= h.nb_a; ++n1)
                for (int l = 0; l <= h.l2m; ++l)
                    if (key(l, n0) >= key(l, n1p)) B.b2.push_back({n0, n1p, n1, l});
    // order 3: pre-order on combined (atomic, magnetic) keys, then rank-2 atomic ordered and
    // rank-3 magnetic triangle/parity
    {
        const int ka_max = h.nr3_a * (h.l3a + 1), km_max = h.nr3_m * (h.l3m + 1);
        auto nc = [&](int na, int la, int nm, int lm) {
            return (long)(km_max + 1) * (h.nr3_a * la + na) + (h.nr3_m * lm + nm);
        };
        (void)ka_max;
        for (int n1 = 1; n1 <= h.nr3_a; ++n1)
        for (int n2 = 1; n2 <= h.nr3_a; ++n2)
        for (int n0p = 1; n0p <= h.nr3_m; ++n0p)
        for (int n1p = 1; n1p <= h.nr3_m; ++n1p)
        for (int n2p = 1; n2p <= h.nr3_m; ++n2p)
        for (int l1 = 0; l1 <= h.l3a; ++l1)
        for (int l0p = 0; l0p <= h.l3m; ++l0p)
        for (int l1p = 0; l1p <= h.l3m; ++l1p)
        for (int l2p = 0; l2p <= h.l3m; ++l2p) {
            if (nc(n1, l1, n1p, l1p) < nc(n2, l1, n2p, l2p)) continue;
            if (key(l1, n1) < key(l1, n2)) continue;
            if (!triangle_even(l0p, l1p, l2p)) continue;
            B.b3.push_back({n1, n2, n0p, n1p, n2p, l1, l0p, l1p, l2p});
        }
    }
    // order 4
    {
        const int km_max = h.nr4_m * (h.l4m + 1);
        auto nc = [&](int na, int la, int nm, int lm) {
            return (long)(km_max + 1) * (h.nr4_a * la + na) + (h.nr4_m * lm + nm);
        };
        for (int n1 = 1; n1 <= h.nr4_a; ++n1)
        for (int n2 = 1; n2 <= h.nr4_a; ++n2)
        for (int n3 = 1; n3 <= h.nr4_a; ++n3)
        for (int m0 = 1; m0 <= h.nr4_m; ++m0)
        for (int m1 = 1; m1 <= h.nr4_m; ++m1)
        for (int m2 = 1; m2 <= h.nr4_m; ++m2)
        for (int m3 = 1; m3 <= h.nr4_m; ++m3)
        for (int l1 = 0; l1 <= h.l4a; ++l1)
        for (int l2 = 0; l2 <= h.l4a; ++l2)
        for (int l3 = 0; l3 <= h.l4a; ++l3)
        for (int k0 = 0; k0 <= h.l4m; ++k0)
        for (int k1 = 0; k1 <= h.l4m; ++k1)
        for (int k2 = 0; k2 <= h.l4m; ++k2)
        for (int k3 = 0; k3 <= h.l4m; ++k3) {
            long c1 = nc(n1, l1, m1, k1), c2 = nc(n2, l2, m2, k2), c3 = nc(n3, l3, m3, k3);
            if (!(c1 >= c2 && c2 >= c3)) continue;
            if (!triangle_even(l1, l2, l3)) continue;
            const int nr[4] = {m0, m1, m2, m3}, ll[4] = {k0, k1, k2, k3};
            for (int L : rank4_couplings(key, nr, ll))
                B.b4.push_back({n1, n2, n3, m0, m1, m2, m3, l1, l2, l3, k0, k1, k2, k3, L});
        }
    }
    if (h.nr5_a > 0 && h.nr5_m > 0) enumerate5(h, key, B);
    if (h.nr6_a > 0 && h.nr6_m > 0) enumerate6(h, key, B);
    return B;
}

// Clebsch-Gordan <j1 m1 j2 m2 | J M> by the Racah sum (src/util.F90:4176-4249)
double clebsch_gordan(int j1, int j2, int J, int m1, int m2, int M) {
    static double fac[200];
    static bool init = false;
    if (!init) {
        fac[0] = 1.0;
        for (int i = 1; i < 200; ++i) fac[i] = fac[i - 1] * (double)i;
        init = true;
    }
    if (j1 + j2 < J || std::abs(j1 - j2) > J) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(M) > J || m1 + m2 != M) return 0.0;
    double pre = (double)(2 * J + 1) * fac[j1 + j2 - J] * fac[j1 - j2 + J] * fac[j2 - j1 + J] /
                 fac[j1 + j2 + J + 1] *
                 (fac[J + M] * fac[J - M] * fac[j1 + m1] * fac[j1 - m1] * fac[j2 + m2] * fac[j2 - m2]);
    double s = 0.0;
    for (int k = 0; k <= j1 + j2 + J + 1; ++k) {
        int a = j1 + j2 - J - k, b = j1 - m1 - k, c = j2 + m2 - k, d = J - j2 + m1 + k, e = J - j1 - m2 + k;
        if (a < 0 || b < 0 || c < 0 || d < 0 || e < 0) continue;
        double t = 1.0 / (fac[k] * fac[a] * fac[b] * fac[c] * fac[d] * fac[e]);
        s += (k & 1) ? -t : t;
    }
    return std::sqrt(pre) * s;
}

ZLayout make_layout(const Header& h) {
    ZLayout z;
    z.lsa = (h.l3a + 1) * (h.l3a + 1);
    z.lsm = (h.l2m + 1) * (h.l2m + 1);
    z.lAm = h.l3m;
    z.lsm3 = (z.lAm + 1) * (z.lAm + 1);
    z.nm3 = h.nr3_m * z.lsm3;
    z.nr2m_ = h.nr2_m;
    z.oA = 0;
    z.oA2 = h.nr3_a * z.lsa * z.nm3;
    z.oA0 = z.oA2 + h.nb_a * h.nr2_m * z.lsm;
    z.nz = z.oA0 + h.nr2_m * z.lsm;
    return z;
}

namespace {

// complex-conjugate partner of a Z entry: X_{-m,-m'} = (-1)^(m+m') conj(X_{m,m'})
// (src/a_basis.F90:303-342, spherical.F90:263-280)
struct Conj {
    std::vector<int> partner;
    std::vector<int> sign;
};

Conj conj_map(const Header& h, const ZLayout& z) {
    Conj c;
    c.partner.assign(z.nz + 1, 0);
    c.sign.assign(z.nz + 1, 1);
    for (int n = 0; n < h.nr3_a; ++n)
        for (int l = 0; l <= h.l3a; ++l)
            for (int m = -l; m <= l; ++m)
                for (int np = 0; np < h.nr3_m; ++np)
                    for (int lp = 0; lp <= z.lAm; ++lp)
                        for (int mp = -lp; mp <= lp; ++mp) {
                            int x = z.a(n, l, m, np, lp, mp);
                            c.partner[x] = z.a(n, l, -m, np, lp, -mp);
                            c.sign[x] = ((m + mp) & 1) ? -1 : 1;
                        }
    for (int n = 0; n < h.nb_a; ++n)
        for (int np = 0; np < h.nr2_m; ++np)
            for (int lp = 0; lp <= h.l2m; ++lp)
                for (int mp = -lp; mp <= lp; ++mp) {
                    int x = z.a2(n, np, lp, mp);
                    c.partner[x] = z.a2(n, np, lp, -mp);
                    c.sign[x] = (mp & 1) ? -1 : 1;
                }
    for (int np = 0; np < h.nr2_m; ++np)
        for (int lp = 0; lp <= h.l2m; ++lp)
            for (int mp = -lp; mp <= lp; ++mp) {
                int x = z.a0(np, lp, mp);
                c.partner[x] = z.a0(np, lp, -mp);
                c.sign[x] = (mp & 1) ? -1 : 1;
            }
    c.partner[z.nz] = z.nz;
    return c;
}

// canonical key of a monomial of up to six factors (indices < 2^16, padded with `one`)
using Key = unsigned __int128;
Key pack6(std::array<int, 6> v) {
    std::sort(v.begin(), v.end());
    Key k = 0;
    for (int i = 0; i < 6; ++i) k = (k << 16) | (Key)(uint16_t)v[i];
    return k;
}
std::array<int, 6> unpack6(Key k) {
    std::array<int, 6> v;
    for (int i = 5; i >= 0; --i) {
        v[i] = (int)(uint16_t)k;
        k >>= 16;
    }
    return v;
}
struct KeyHash {
    size_t operator()(Key k) const {
        const uint64_t a = (uint64_t)k, b = (uint64_t)(k >> 64);
        return std::hash<uint64_t>()(a ^ (b * 0x9e3779b97f4a7c15ull));
    }
};

}  // namespace

// LDS passes of one read instruction for one 16-lane group: lanes that read the same entry are
// served by one broadcast, distinct entries of one class x mod 16 (the same banks) take one
// pass each, so the cost is the largest number of distinct entries in one class.
//
// The 16-lane groups of a 16-byte read (ds_read_b128) on gfx950, measured with the LDS
// bank-conflict counter (a pair of lanes reading entries 0 and 16, every other lane one shared
// entry, conflicts exactly when the two lanes share a group): in each half of the wave, the
// quads (lanes 4q..4q+3, q < 8) with an even number of set bits in q form one group
// ({0-3, 12-15, 20-23, 24-27}) and those with an odd number the other.
static inline int lds_group(int lane) {
    const int q = (lane >> 2) & 7;
    return ((lane >> 5) << 1) | ((q ^ (q >> 1) ^ (q >> 2)) & 1);
}
struct BankCell {
    int n = 0;
    int val[16], mult[16];
    void add(int v) {  // at most 16 distinct entries: one per lane of the group
        for (int i = 0; i < n; ++i)
            if (val[i] == v) { ++mult[i]; return; }
        if (n == 16) std::abort();
        val[n] = v;
        mult[n++] = 1;
    }
    void remove(int v) {
        for (int i = 0; i < n; ++i)
            if (val[i] == v) {
                if (--mult[i] == 0) { val[i] = val[n - 1]; mult[i] = mult[n - 1]; --n; }
                return;
            }
    }
    bool has(int v) const {
        for (int i = 0; i < n; ++i)
            if (val[i] == v) return true;
        return false;
    }
    int load(int c) const {  // distinct entries in class c
        int k = 0;
        for (int i = 0; i < n; ++i) k += (val[i] & 15) == c;
        return k;
    }
    int passes() const {
        int cnt[16] = {0}, m = 0;
        for (int i = 0; i < n; ++i) m = std::max(m, ++cnt[val[i] & 15]);
        return m;
    }
};

// Local search after the greedy record order of one storage group (len steps x nl lanes at
// base): swaps two records of one lane between steps while that lowers the group's LDS passes
// (BankCell, for every step, 16-lane group and factor slot; padding, index `one`, is a
// broadcast and not counted).  The greedy's last steps have few records left to choose from;
// the swaps repair them (one sweep over the lanes).
static void refine_banks(TermList& T, size_t base, int len, int nl, int one) {
    const int ng = 4;
    std::vector<BankCell> cell((size_t)len * ng * 3);  // [step][group][factor]
    auto at = [&](int k, int g, int f) -> BankCell& { return cell[((size_t)k * ng + g) * 3 + f]; };
    for (int k = 0; k < len; ++k)
        for (int L = 0; L < nl; ++L)
            for (int f = 0; f < 3; ++f) {
                const int v = T.cidx[4 * (base + (size_t)k * 64 + L) + f];
                if (v != one) at(k, lds_group(L), f).add(v);
            }
    {
        for (int L = 0; L < nl; ++L) {
            const int g = lds_group(L);
            for (int k1 = 0; k1 < len; ++k1)
                for (int k2 = k1 + 1; k2 < len; ++k2) {
                    const size_t r1 = base + (size_t)k1 * 64 + L, r2 = base + (size_t)k2 * 64 + L;
                    int v1[3], v2[3];
                    bool same = true;
                    for (int f = 0; f < 3; ++f) {
                        v1[f] = T.cidx[4 * r1 + f];
                        v2[f] = T.cidx[4 * r2 + f];
                        same = same && v1[f] == v2[f];
                    }
                    if (same) continue;
                    auto move = [&](bool fwd) {  // r1 to step k2 and r2 to k1, or back
                        for (int f = 0; f < 3; ++f) {
                            BankCell &a = at(k1, g, f), &b = at(k2, g, f);
                            const int x = fwd ? v1[f] : v2[f], y = fwd ? v2[f] : v1[f];
                            // x leaves a for b, y leaves b for a; removals first, so that no
                            // cell holds more than its 16 lanes' entries
                            if (x != one) a.remove(x);
                            if (y != one) b.remove(y);
                            if (x != one) b.add(x);
                            if (y != one) a.add(y);
                        }
                    };
                    int delta = 0;
                    for (int f = 0; f < 3; ++f) delta -= at(k1, g, f).passes() + at(k2, g, f).passes();
                    move(true);
                    for (int f = 0; f < 3; ++f) delta += at(k1, g, f).passes() + at(k2, g, f).passes();
                    if (delta < 0) {
                        for (int q = 0; q < 4; ++q) std::swap(T.cidx[4 * r1 + q], T.cidx[4 * r2 + q]);
                        std::swap(T.cw[r1], T.cw[r2]);
                        std::swap(T.rec_t[r1], T.rec_t[r2]);
                    } else {
                        move(false);
                    }
                }
        }
    }
}

TermList build_terms(const Potential& pot, const ZLayout& z) {
    const Header& h = pot.h;
    const Basis& B = pot.basis;
    const int one = z.nz;  // padding factor Z[nz] = 1
    TermList T;
    // Structure first, coefficients second: every B function k of order >= 2 is expanded
    // into m-combinations u * Z[a]Z[b]Z[c]Z[d] (CG products, no coefficient), so that the
    // monomial set does not depend on c and the merged weights w_t = sum_k c_k u_{k,t} can be
    // recomputed for new coefficients (fitting).
    const Conj cj = conj_map(h, z);
    std::unordered_map<Key, double, KeyHash> one_b;  // canonical key -> u of the current B function
    std::vector<std::vector<std::pair<Key, double>>> per_b;
    auto emit6 = [&](double u, const std::array<int, 6>& v) {
        if (u == 0.0) return;
        ++T.raw;
        // fold complex-conjugate partners: Re(u P) + Re(u' s conj P) = (u + s u') Re P
        std::array<int, 6> vb;
        int sg = 1;
        for (int i = 0; i < 6; ++i) {
            vb[i] = cj.partner[v[i]];
            sg *= cj.sign[v[i]];
        }
        const Key key = pack6(v), kb = pack6(vb);
        if (kb == key) {
            // a self-conjugate monomial with s = -1 is purely imaginary: Re P == 0
            if (sg == 1) one_b[key] += u;
        } else if (kb < key) {
            one_b[kb] += sg * u;
        } else {
            one_b[key] += u;
        }
    };
    auto emit = [&](double u, int a, int b, int c, int d) { emit6(u, {a, b, c, d, one, one}); };
    auto close_b = [&]() {
        std::vector<std::pair<Key, double>> v(one_b.begin(), one_b.end());
        std::sort(v.begin(), v.end());
        per_b.push_back(std::move(v));
        one_b.clear();
    };
    // order 2: B2 = Re sum_m C(l l 0; m -m) A0[n0'][l m] A2[n1][n1'][l -m]  (b_basis.F90:95-124)
    for (size_t k = 0; k < B.b2.size(); ++k) {
        const auto& b = B.b2[k];
        const int n0 = b[0] - 1, n1p = b[1] - 1, n1 = b[2] - 1, l = b[3];
        for (int m = -l; m <= l; ++m) {
            ++T.n2;
            emit(clebsch_gordan(l, l, 0, m, -m, 0), z.a0(n0, l, m), z.a2(n1, n1p, l, -m), one, one);
        }
        close_b();
    }
    // order 3 (b_basis.F90:146-205, couplings util.F90:4292-4333)
    for (size_t k = 0; k < B.b3.size(); ++k) {
        const auto& b = B.b3[k];
        const int n1 = b[0] - 1, n2 = b[1] - 1, n0p = b[2] - 1, n1p = b[3] - 1, n2p = b[4] - 1;
        const int l1 = b[5], l0p = b[6], l1p = b[7], l2p = b[8];
        for (int m1 = -l1; m1 <= l1; ++m1) {
            const double ca = clebsch_gordan(l1, l1, 0, m1, -m1, 0);
            for (int q0 = -l0p; q0 <= l0p; ++q0)
                for (int q1 = -l1p; q1 <= l1p; ++q1) {
                    const int q2 = -(q0 + q1);
                    if (std::abs(q2) > l2p) continue;
                    ++T.n3;
                    const double cm = clebsch_gordan(l0p, l1p, l2p, q0, q1, -q2) *
                                      clebsch_gordan(l2p, l2p, 0, q2, -q2, 0);
                    emit(ca * cm, z.a0(n0p, l0p, q0), z.a(n1, l1, m1, n1p, l1p, q1),
                         z.a(n2, l1, -m1, n2p, l2p, q2), one);
                }
        }
        close_b();
    }
    // order 4 (b_basis.F90:229-301, couplings util.F90:4343-4388)
    for (size_t k = 0; k < B.b4.size(); ++k) {
        const auto& b = B.b4[k];
        const int n1 = b[0] - 1, n2 = b[1] - 1, n3 = b[2] - 1;
        const int p0 = b[3] - 1, p1 = b[4] - 1, p2 = b[5] - 1, p3 = b[6] - 1;
        const int l1 = b[7], l2 = b[8], l3 = b[9], k0 = b[10], k1 = b[11], k2 = b[12], k3 = b[13], L = b[14];
        // A holds the magnetic l' <= l3max_magnetic only (src/a_basis.F90:61,248-251: the
        // entries beyond stay zero), so a function with a factor beyond vanishes
        if (k1 > h.l3m || k2 > h.l3m || k3 > h.l3m) {
            close_b();
            continue;
        }
        for (int m1 = -l1; m1 <= l1; ++m1)
            for (int m2 = -l2; m2 <= l2; ++m2) {
                const int m3 = -(m1 + m2);
                if (std::abs(m3) > l3) continue;
                const double ca = clebsch_gordan(l1, l2, l3, m1, m2, -m3) * clebsch_gordan(l3, l3, 0, m3, -m3, 0);
                for (int M = -L; M <= L; ++M)
                    for (int q0 = -k0; q0 <= k0; ++q0) {
                        const int q1 = M - q0;
                        if (std::abs(q1) > k1) continue;
                        for (int q2 = -k2; q2 <= k2; ++q2) {
                            const int q3 = -M - q2;
                            if (std::abs(q3) > k3) continue;
                            ++T.n4;
                            const double cm = clebsch_gordan(k0, k1, L, q0, q1, M) *
                                              clebsch_gordan(k2, k3, L, q2, q3, -M) *
                                              clebsch_gordan(L, L, 0, M, -M, 0);
                            emit(ca * cm, z.a0(p0, k0, q0), z.a(n1, l1, m1, p1, k1, q1),
                                 z.a(n2, l2, m2, p2, k2, q2), z.a(n3, l3, m3, p3, k3, q3));
                        }
                    }
            }
        close_b();
    }
    // order 5 (b_basis.F90:320-414, couplings util.F90:4392-4455): atomic
    // ((l1 l2)L (l3 l4)L)0 over M = m1 + m2 = -(m3 + m4); magnetic ((l0' l1')L01 (l2' l3')L23)l4'
    // with m4' = -(M01 + M23), coupled to zero with l4'
    for (size_t k = 0; k < B.b5.size(); ++k) {
        const auto& b = B.b5[k];
        const int n[4] = {b[0] - 1, b[1] - 1, b[2] - 1, b[3] - 1}, p0 = b[4] - 1;
        const int p[4] = {b[5] - 1, b[6] - 1, b[7] - 1, b[8] - 1};
        const int l[4] = {b[9], b[10], b[11], b[12]}, k0 = b[13];
        const int kk[4] = {b[14], b[15], b[16], b[17]};
        const int L = b[18], L01 = b[19], L23 = b[20];
        if (*std::max_element(kk, kk + 4) > h.l3m) {  // a factor beyond A's l' range: zero, as above
            close_b();
            continue;
        }
        for (int M = -L; M <= L; ++M)
            for (int m1 = -l[0]; m1 <= l[0]; ++m1) {
                const int m2 = M - m1;
                if (std::abs(m2) > l[1]) continue;
                for (int m3 = -l[2]; m3 <= l[2]; ++m3) {
                    const int m4 = -M - m3;
                    if (std::abs(m4) > l[3]) continue;
                    const double ca = clebsch_gordan(l[0], l[1], L, m1, m2, M) *
                                      clebsch_gordan(l[2], l[3], L, m3, m4, -M) * clebsch_gordan(L, L, 0, M, -M, 0);
                    if (ca == 0.0) continue;
                    const int m[4] = {m1, m2, m3, m4};
                    for (int M01 = -L01; M01 <= L01; ++M01)
                        for (int M23 = -L23; M23 <= L23; ++M23) {
                            const int q4 = -(M01 + M23);
                            if (std::abs(q4) > kk[3]) continue;
                            for (int q0 = -k0; q0 <= k0; ++q0) {
                                const int q1 = M01 - q0;
                                if (std::abs(q1) > kk[0]) continue;
                                for (int q2 = -kk[1]; q2 <= kk[1]; ++q2) {
                                    const int q3 = M23 - q2;
                                    if (std::abs(q3) > kk[2]) continue;
                                    ++T.n5;
                                    const double cm = clebsch_gordan(k0, kk[0], L01, q0, q1, M01) *
                                                      clebsch_gordan(kk[1], kk[2], L23, q2, q3, M23) *
                                                      clebsch_gordan(L01, L23, kk[3], M01, M23, -q4) *
                                                      clebsch_gordan(kk[3], kk[3], 0, -q4, q4, 0);
                                    const int q[4] = {q1, q2, q3, q4};
                                    std::array<int, 6> v;
                                    v[0] = z.a0(p0, k0, q0);
                                    for (int i = 0; i < 4; ++i) v[1 + i] = z.a(n[i], l[i], m[i], p[i], kk[i], q[i]);
                                    v[5] = one;
                                    emit6(ca * cm, v);
                                }
                            }
                        }
                }
            }
        close_b();
    }
    // order 6 (b_basis.F90:420-528, couplings util.F90:4457-4548): atomic
    // ((l1 l2)L12 (l3 l4)L34)l5 over m5 = -(M12 + M34), coupled to zero with l5; magnetic
    // ((l0' l1')L01 (l2' l3')L23)L45 (l4' l5')L45 with M45 = -(M01 + M23)
    for (size_t k = 0; k < B.b6.size(); ++k) {
        const auto& b = B.b6[k];
        int n[5], p[5], l[5], kk[5];
        for (int i = 0; i < 5; ++i) {
            n[i] = b[i] - 1;
            p[i] = b[6 + i] - 1;
            l[i] = b[11 + i];
            kk[i] = b[17 + i];
        }
        const int p0 = b[5] - 1, k0 = b[16];
        const int L12 = b[22], L34 = b[23], L01 = b[24], L23 = b[25], L45 = b[26];
        if (*std::max_element(kk, kk + 5) > h.l3m) {  // a factor beyond A's l' range: zero, as above
            close_b();
            continue;
        }
        for (int M12 = -L12; M12 <= L12; ++M12)
            for (int M34 = -L34; M34 <= L34; ++M34) {
                const int m5 = -(M12 + M34);
                if (std::abs(m5) > l[4]) continue;
                for (int m1 = -l[0]; m1 <= l[0]; ++m1) {
                    const int m2 = M12 - m1;
                    if (std::abs(m2) > l[1]) continue;
                    for (int m3 = -l[2]; m3 <= l[2]; ++m3) {
                        const int m4 = M34 - m3;
                        if (std::abs(m4) > l[3]) continue;
                        const double ca = clebsch_gordan(l[0], l[1], L12, m1, m2, M12) *
                                          clebsch_gordan(l[2], l[3], L34, m3, m4, M34) *
                                          clebsch_gordan(L12, L34, l[4], M12, M34, -m5) *
                                          clebsch_gordan(l[4], l[4], 0, -m5, m5, 0);
                        if (ca == 0.0) continue;
                        const int m[5] = {m1, m2, m3, m4, m5};
                        for (int M01 = -L01; M01 <= L01; ++M01)
                            for (int M23 = -L23; M23 <= L23; ++M23) {
                                const int M45 = -(M01 + M23);
                                if (std::abs(M45) > L45) continue;
                                for (int q0 = -k0; q0 <= k0; ++q0) {
                                    const int q1 = M01 - q0;
                                    if (std::abs(q1) > kk[0]) continue;
                                    for (int q2 = -kk[1]; q2 <= kk[1]; ++q2) {
                                        const int q3 = M23 - q2;
                                        if (std::abs(q3) > kk[2]) continue;
                                        for (int q4 = -kk[3]; q4 <= kk[3]; ++q4) {
                                            const int q5 = M45 - q4;
                                            if (std::abs(q5) > kk[4]) continue;
                                            ++T.n6;
                                            const double cm = clebsch_gordan(k0, kk[0], L01, q0, q1, M01) *
                                                              clebsch_gordan(kk[1], kk[2], L23, q2, q3, M23) *
                                                              clebsch_gordan(kk[3], kk[4], L45, q4, q5, M45) *
                                                              clebsch_gordan(L01, L23, L45, M01, M23, -M45) *
                                                              clebsch_gordan(L45, L45, 0, -M45, M45, 0);
                                            const int q[5] = {q1, q2, q3, q4, q5};
                                            std::array<int, 6> v;
                                            v[0] = z.a0(p0, k0, q0);
                                            for (int i = 0; i < 5; ++i)
                                                v[1 + i] = z.a(n[i], l[i], m[i], p[i], kk[i], q[i]);
                                            emit6(ca * cm, v);
                                        }
                                    }
                                }
                            }
                    }
                }
            }
        close_b();
    }
    // the monomial set: every key with a nonzero coefficient in some B function
    std::map<Key, int> keys;
    for (const auto& v : per_b)
        for (const auto& e : v)
            if (e.second != 0.0) keys.emplace(e.first, 0);
    // device order: slot 0 = the A0 factor (every monomial has exactly one), slots 1..3 = the
    // other factors ascending, padding last; sorted lexicographically.  Five or six factors:
    // the last two (and the two before them) become products Y (TermList::ypair), numbered in
    // the order of the monomial keys
    std::map<std::pair<int, int>, int> ymap;
    auto ycol = [&](int a, int b) {
        auto it = ymap.emplace(std::make_pair(a, b), (int)ymap.size()).first;
        return one + 1 + it->second;
    };
    std::vector<std::pair<std::array<int, 4>, Key>> rows;
    rows.reserve(keys.size());
    for (const auto& kv : keys) {
        const std::array<int, 6> v = unpack6(kv.first);
        int a0 = -1;
        std::vector<int> f;
        for (int i = 0; i < 6; ++i) {
            if (a0 < 0 && v[i] >= z.oA0 && v[i] < z.nz) a0 = v[i];
            else if (v[i] != one) f.push_back(v[i]);
        }
        std::sort(f.begin(), f.end());
        std::array<int, 4> o = {a0, one, one, one};
        if (f.size() <= 3) {
            for (size_t i = 0; i < f.size(); ++i) o[1 + i] = f[i];
        } else if (f.size() == 4) {
            o[1] = f[0];
            o[2] = f[1];
            o[3] = ycol(f[2], f[3]);
        } else {
            o[1] = f[0];
            o[2] = ycol(f[1], f[2]);
            o[3] = ycol(f[3], f[4]);
        }
        std::sort(o.begin() + 1, o.end());
        rows.emplace_back(o, kv.first);
    }
    std::sort(rows.begin(), rows.end());
    T.ny = (int)ymap.size();
    const int nzx = z.nz + 1 + T.ny;  // targets: Z entries and Y entries (index nz is the padding)
    T.nzx_ = nzx;
    T.ypair.assign(2 * (size_t)T.ny, 0);
    for (const auto& kv : ymap) {
        T.ypair[2 * kv.second] = kv.first.first;
        T.ypair[2 * kv.second + 1] = kv.first.second;
    }
    {
        // dRho/dZ[x] += dRho/dY_i Z[other] for both factors of every Y_i, grouped by x
        std::vector<std::vector<std::pair<int, int>>> byx(z.nz);
        for (int i = 0; i < T.ny; ++i) {
            const int a = T.ypair[2 * i], b = T.ypair[2 * i + 1];
            byx[a].emplace_back(one + 1 + i, b);
            byx[b].emplace_back(one + 1 + i, a);
        }
        for (int x = 0; x < z.nz; ++x) {
            if (byx[x].empty()) continue;
            T.yft.push_back((int32_t)(T.yfe.size() / 2));
            T.yft.push_back(x | ((int)byx[x].size() << 12));
            for (const auto& e : byx[x]) {
                T.yfe.push_back(e.first);
                T.yfe.push_back(e.second);
            }
        }
    }
    for (size_t t = 0; t < rows.size(); ++t) keys[rows[t].second] = (int)t;
    // U (by B function and transposed by monomial), then w_t = sum_k c_k u_{k,t}
    const size_t nt = rows.size();
    const int nd = h.ndensity;
    T.nd = nd;
    T.ub_ptr.assign(1, 0);
    std::vector<std::vector<std::pair<int, double>>> byt(nt);
    // coefficients of the B functions of orders 2..6, density-major
    const size_t n2 = B.b2.size(), n3 = B.b3.size(), n4 = B.b4.size(), n5 = B.b5.size(), n6 = B.b6.size();
    const size_t nb = n2 + n3 + n4 + n5 + n6;
    std::vector<double> cvec(nb * nd);
    for (int d = 0; d < nd; ++d) {
        size_t o = d * nb;
        for (size_t k = 0; k < n2; ++k) cvec[o++] = pot.c2[d * n2 + k];
        for (size_t k = 0; k < n3; ++k) cvec[o++] = pot.c3[d * n3 + k];
        for (size_t k = 0; k < n4; ++k) cvec[o++] = pot.c4[d * n4 + k];
        for (size_t k = 0; k < n5; ++k) cvec[o++] = pot.c5[d * n5 + k];
        for (size_t k = 0; k < n6; ++k) cvec[o++] = pot.c6[d * n6 + k];
    }
    for (size_t k = 0; k < per_b.size(); ++k) {
        for (const auto& e : per_b[k]) {
            if (e.second == 0.0) continue;
            const int t = keys[e.first];
            T.ub_t.push_back(t);
            T.ub_u.push_back(e.second);
            byt[t].emplace_back((int)k, e.second);
        }
        T.ub_ptr.push_back((int32_t)T.ub_t.size());
    }
    T.ut_ptr.assign(1, 0);
    T.w.assign(nt * nd, 0.0);
    for (size_t t = 0; t < nt; ++t) {
        for (const auto& e : byt[t]) {
            T.ut_k.push_back(e.first);
            T.ut_u.push_back(e.second);
        }
        for (int d = 0; d < nd; ++d) {
            double w = 0.0;
            for (const auto& e : byt[t]) w += cvec[d * nb + e.first] * e.second;
            T.w[d * nt + t] = w;
        }
        T.ut_ptr.push_back((int32_t)T.ut_k.size());
        for (int i = 0; i < 4; ++i) T.idx.push_back((uint16_t)rows[t].first[i]);
    }
    // transposed gradient: contributions grouped by target
    // (entries: the three other factors, the monomial index)
    std::vector<std::vector<std::pair<std::array<int, 3>, int>>> by(nzx);
    for (size_t t = 0; t < nt; ++t)
        for (int s = 0; s < 4; ++s) {
            const int x = rows[t].first[s];
            if (x == one) continue;
            std::array<int, 3> o;
            int k = 0;
            for (int q = 0; q < 4; ++q)
                if (q != s) o[k++] = rows[t].first[q];
            by[x].emplace_back(o, (int)t);
        }
    // Chunks of <= kChunk contributions per row, ordered round robin across rows so that
    // neighbouring lanes flush to different targets.  A wave processes 64 consecutive
    // chunks at once; their records are stored interleaved (record k of lane L of group G
    // at G*64*len + k*64 + L), so each step of the 64 lanes reads 64 consecutive 16-byte
    // records.  Chunks are padded (factor index nz, weight 0) to the longest chunk of their
    // group, rounded up to a multiple of 4.
    struct Ch { int x, off, cnt; };
    std::vector<std::vector<Ch>> rowchunks(nzx);
    for (int x = 0; x < nzx; ++x) {
        auto& v = by[x];
        std::sort(v.begin(), v.end());
        // the sorted contributions are dealt round robin over the target's full chunks (chunk c
        // takes c, c + nf, c + 2 nf, ...), so that its chunks, which often run side by side
        // on neighbouring lanes, hold overlapping factors that the bank-aware order below can
        // align into broadcasts (modelled passes per read 1.41 -> 1.31; adjoint 1-2% faster)
        const size_t K = TermList::kChunk, nf = v.size() / K;
        if (nf > 1) {
            const std::vector<std::pair<std::array<int, 3>, int>> w(v.begin(), v.begin() + nf * K);
            for (size_t c = 0; c < nf; ++c)
                for (size_t i = 0; i < K; ++i) v[c * K + i] = w[c + nf * i];
        }
        for (size_t i = 0; i < v.size(); i += TermList::kChunk)
            rowchunks[x].push_back({x, (int)i, (int)std::min<size_t>(TermList::kChunk, v.size() - i)});
    }
    // Order of the chunks.  Chunk c runs on thread c mod L in round c / L (L = kLanes, the
    // adjoint's workgroup size), so the order is a grid of rounds (rows) by L lanes
    // (columns); the 64 chunks of a wave in a round
    // form one storage group.  Full chunks fill the first whole rounds column by column,
    // targets in decreasing number of full chunks: a lane then meets the same target in
    // consecutive rounds and keeps its sum in registers (one LDS atomic per run instead of
    // one per chunk), and the many chunks of the largest targets spread over many lanes
    // instead of crowding the atomics of a few waves.  The remaining full chunks and the
    // partial ones (one per target) follow by decreasing length, so that the chunks of a
    // wave have (nearly) equal lengths.
    std::vector<Ch> order, rest;
    {
        std::vector<int> tg;
        for (int x = 0; x < nzx; ++x)
            if (!rowchunks[x].empty()) tg.push_back(x);
        auto nfull = [&](int x) {
            int n = 0;
            for (const Ch& ch : rowchunks[x]) n += ch.cnt == TermList::kChunk;
            return n;
        };
        std::stable_sort(tg.begin(), tg.end(), [&](int a, int b) { return nfull(a) > nfull(b); });
        std::vector<Ch> full;
        for (int x : tg)
            for (const Ch& ch : rowchunks[x]) (ch.cnt == TermList::kChunk ? full : rest).push_back(ch);
        const size_t L = TermList::kLanes;
        const size_t rounds = full.size() / L;
        order.resize(rounds * L);
        for (size_t q = 0; q < rounds * L; ++q) order[(q % rounds) * L + q / rounds] = full[q];
        rest.insert(rest.begin(), full.begin() + rounds * L, full.end());
        std::stable_sort(rest.begin(), rest.end(), [](const Ch& a, const Ch& b) { return a.cnt > b.cnt; });
        order.insert(order.end(), rest.begin(), rest.end());
    }
    // Inside a chunk the records can be summed in any order.  At step k the 64 lanes read
    // Z[a], Z[b], Z[c] of their k-th records from LDS (three reads, random addresses), and
    // LDS serves a wave in lane groups whose accesses must fall in distinct banks.  The
    // order inside each chunk is therefore chosen greedily, step by step, so that the
    // distinct factor indices of each 16-lane group (lds_group) spread over the 16 classes
    // x mod 16 (BankCell): a 16-byte entry covers 4 of the 64 four-byte banks, so one pass
    // serves 16 lanes, and lanes reading the same entry share a broadcast.  The FP64 and the
    // two-atom FP32 adjoints read such entries.  With the measured groups in place of
    // consecutive lanes, the passes per read of the lmax-6 header fell from 1.83 to 1.38
    // (bank-conflict cycles per LDS instruction of the two-atom FP32 adjoint 3.2 -> 1.5;
    // adjoint 4-6% faster); choosing each record's factor order as well (which factor is
    // read by which of the three reads) takes it to 1.29.  A local search (refine_banks)
    // then repairs the last steps.  Padding (index nz) is a broadcast.
    for (size_t g0 = 0; g0 < order.size(); g0 += 64) {
        const size_t g1 = std::min(order.size(), g0 + 64);
        int len = 0;
        for (size_t c = g0; c < g1; ++c) len = std::max(len, (order[c].cnt + 3) & ~3);
        const size_t base = T.cw.size();
        T.cidx.resize((base + (size_t)len * 64) * 4, (uint16_t)one);
        T.cw.resize(base + (size_t)len * 64, 0.0);
        T.rec_t.resize(base + (size_t)len * 64, -1);
        const int nl = (int)(g1 - g0);
        std::vector<std::vector<int>> left(nl);
        for (int L = 0; L < nl; ++L)
            for (int k = 0; k < order[g0 + L].cnt; ++k) left[L].push_back(k);
        for (int k = 0; k < len; ++k) {
            BankCell sp[3][4];  // [factor slot][16-lane group] of this step
            for (int L = 0; L < nl; ++L) {
                auto& rem = left[L];
                if (rem.empty()) continue;
                const Ch& ch = order[g0 + L];
                // a record's three factors may be read in any order (the product commutes): each
                // candidate is scored in its cheapest order
                static const int perm[6][3] = {{0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {2, 0, 1}, {1, 2, 0}, {2, 1, 0}};
                int best = 0, bestp = 0, bestc = INT_MAX;
                const int g = lds_group(L);
                for (int i = 0; i < (int)rem.size(); ++i) {
                    const auto& e = by[ch.x][ch.off + rem[i]];
                    int cf[3][3];  // cost of factor j in slot f
                    for (int j = 0; j < 3; ++j)
                        for (int f = 0; f < 3; ++f) {
                            const int v = e.first[j];
                            const BankCell& bc = sp[f][g];
                            cf[j][f] = v == one || bc.has(v) ? 0 : bc.load(v & 15);
                        }
                    for (int p = 0; p < 6; ++p) {
                        const int cost = cf[perm[p][0]][0] + cf[perm[p][1]][1] + cf[perm[p][2]][2];
                        if (cost < bestc) { bestc = cost; best = i; bestp = p; }
                    }
                }
                const auto& e = by[ch.x][ch.off + rem[best]];
                const int fv[3] = {e.first[perm[bestp][0]], e.first[perm[bestp][1]], e.first[perm[bestp][2]]};
                for (int f = 0; f < 3; ++f)
                    if (fv[f] != one)
                        sp[f][g].add(fv[f]);
                const size_t r = base + (size_t)k * 64 + L;
                T.cidx[4 * r] = (uint16_t)fv[0];
                T.cidx[4 * r + 1] = (uint16_t)fv[1];
                T.cidx[4 * r + 2] = (uint16_t)fv[2];
                T.cw[r] = T.w[e.second];  // density 1; further densities appended below
                T.rec_t[r] = e.second;
                rem.erase(rem.begin() + best);
            }
        }
        refine_banks(T, base, len, nl, one);
        // every chunk of the group records the group's length: the lanes of a wave step
        // through the same number of records (those past a chunk's end have weight 0).  Bit
        // 22 marks the only chunk of its target: its sum is the target's whole gradient and
        // is stored, not added.
        for (int L = 0; L < nl; ++L) {
            const Ch& ch = order[g0 + L];
            const bool sole = rowchunks[ch.x].size() == 1;
            T.chunks.push_back((int32_t)(base + L));
            T.chunks.push_back(ch.x | (len << 12) | (sole ? 1 << 18 : 0));
        }
    }
    // partial-sum slots of the runs of targets with more than one chunk (TermList::comb)
    {
        const size_t L = TermList::kLanes, nc = T.chunks.size() / 2;
        std::vector<std::vector<std::vector<size_t>>> runs(nzx);  // target -> runs -> chunks
        for (size_t lane = 0; lane < L && lane < nc; ++lane)
            for (size_t c = lane; c < nc; c += L) {
                const int x = T.chunks[2 * c + 1] & 0xfff;
                if (rowchunks[x].size() == 1) continue;
                const bool cont = c >= L && (T.chunks[2 * (c - L) + 1] & 0xfff) == x;
                if (!cont) runs[x].emplace_back();
                runs[x].back().push_back(c);
            }
        int slot = 0;
        const int G = TermList::kGroup;
        for (int x = 0; x < nzx; ++x) {
            if (runs[x].empty()) continue;
            const int n = (int)runs[x].size(), ng = (n + G - 1) / G;
            for (int g = 0; g < ng; ++g) {
                T.comb.push_back(slot + g * G);
                T.comb.push_back(x | (std::min(G, n - g * G) << 12) | (ng == 1 ? 1 << 16 : 0));
            }
            if (ng > 1) {
                T.comb2.push_back(slot);
                T.comb2.push_back(x | (ng << 12));
            }
            for (const auto& r : runs[x]) {
                for (size_t c : r) T.chunks[2 * c + 1] = (int32_t)((uint32_t)T.chunks[2 * c + 1] | ((uint32_t)slot << 19));
                ++slot;
            }
        }
        T.nslot = slot;
    }
    if (nd > 1) {
        const size_t nr = T.rec_t.size();
        T.cw.resize(nr * nd, 0.0);
        for (int d = 1; d < nd; ++d)
            for (size_t r = 0; r < nr; ++r) T.cw[d * nr + r] = T.rec_t[r] >= 0 ? T.w[d * nt + T.rec_t[r]] : 0.0;
    }
    // representatives: A entries with m' > 0, or m' = 0 and m >= 0; A2/A0 entries with m' >= 0
    T.fold.assign(z.nz, 0u);
    for (int n = 0; n < h.nr3_a; ++n)
        for (int l = 0; l <= h.l3a; ++l)
            for (int m = -l; m <= l; ++m)
                for (int np = 0; np < h.nr3_m; ++np)
                    for (int lp = 0; lp <= z.lAm; ++lp)
                        for (int mp = -lp; mp <= lp; ++mp) {
                            const int x = z.a(n, l, m, np, lp, mp);
                            const bool rep = mp > 0 || (mp == 0 && m >= 0);
                            const bool self = m == 0 && mp == 0;
                            T.fold[x] = (uint32_t)cj.partner[x] | ((cj.sign[x] < 0 ? 1u : 0u) << 16) |
                                        ((rep ? 1u : 0u) << 17) | ((self ? 1u : 0u) << 18);
                        }
    for (int x = z.oA2; x < z.nz; ++x) {
        const int lm = (x < z.oA0 ? (x - z.oA2) : (x - z.oA0)) % z.lsm;
        int lp = 0;
        while ((lp + 1) * (lp + 1) <= lm) ++lp;
        const int mp = lm - lp * lp - lp;
        T.fold[x] = (uint32_t)cj.partner[x] | ((cj.sign[x] < 0 ? 1u : 0u) << 16) |
                    ((mp >= 0 ? 1u : 0u) << 17) | ((mp == 0 ? 1u : 0u) << 18);
    }
    return T;
}

}  // namespace magace
