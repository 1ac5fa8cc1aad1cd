// magace.hpp -- host-side model of a magnetic-ACE potential for the MI355X evaluator.
//
// Scope: the per-atom energy / force / magnetic-gradient path of libace
// (src/ace.F90:79-321 -> prepA, prepB, forceweights, forcecompute, torquecompute),
// single element, body order <= 6, ndensity 1 or 2.  The host side parses the potential file,
// enumerates the basis, builds Clebsch-Gordan couplings and FP64 spline tables, and
// compiles everything into flat tables the HIP kernels consume.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace magace {

// Potential-file header (src/util.F90:603-653).
struct Header {
    int nb_a = 0, nb_m = 0, nr2_m = 0, l2m = 0;   // "nradbase l2max"
    int nr3_a = 0, nr3_m = 0, l3a = 0, l3m = 0;   // "nradial3 l3max"
    int nr4_a = 0, nr4_m = 0, l4a = 0, l4m = 0;   // "nradial4 l4max"
    int nr5_a = 0, nr5_m = 0, l5a = 0, l5m = 0;
    int nr6_a = 0, nr6_m = 0, l6a = 0, l6m = 0;
    int ndensity = 1, npot = 0, nfemb = 0, radtype_a = 3, radtype_m = 7;
    double lambda_a = 0, lambda_m = 0, cut_a = 0, cut_m = 0, dcut_a = 0, dcut_m = 0;
    double ecut = 0, decut = 0;
};

// Enumerated basis: index tuples with the column meaning of b1index/b2index/b3index/b4index.
struct Basis {
    std::vector<std::array<int, 1>> b1;   // n0'
    std::vector<std::array<int, 4>> b2;   // n0' n1' n1 l
    std::vector<std::array<int, 9>> b3;   // n1 n2 n0' n1' n2' l1 l0' l1' l2'
    std::vector<std::array<int, 15>> b4;  // n1 n2 n3 n0'..n3' l1 l2 l3 l0'..l3' L01'
    // n1..n4 n0'..n4' l1..l4 l0'..l4' L12 L01' L23' (b5index)
    std::vector<std::array<int, 21>> b5;
    // n1..n5 n0'..n5' l1..l5 l0'..l5' L12 L34 L01' L23' L45' (b6index)
    std::vector<std::array<int, 27>> b6;
    // non-empty when the reference's enumeration pairs atomic and magnetic couplings of
    // different counts (it then fills the missing side with zero indices and reads outside
    // its arrays): the reason such a header is refused
    std::string undefined;
};

struct Potential {
    Header h;
    Basis basis;
    // expansion coefficients, one per B function and density: c[d * nvar + k]
    std::vector<double> c1, c2, c3, c4, c5, c6;
    std::vector<double> crad_a;          // [nr3_a][l3a+1][nb_a]
    std::vector<double> crad_m;          // [nr2_m][l2m+1][nb_m]
    double p1[2][2] = {{1.0, 1.0}, {1.0, 1.0}};  // per density: embedding prefactor, exponent
    double core[2] = {3.0, 3.0};         // core repulsion prefactor / exponent
    double rcut = 0, rcut_m = 0;         // global cutoffs from acelibini_
    std::string element = "Fe", xc = "pbe";
    double smear = 0.0;
    std::vector<std::string> header_lines;  // the 14 header lines as read
};

// Per-atom complex feature vector Z = [A | A2 | A0] (all m, m' stored).
// A covers the magnetic l' <= l3max_magnetic and nradial3_magnetic radial functions, as the
// reference's A does (src/a_basis.F90:61,248-251,346: entries of higher l' stay zero, so
// higher-order functions with such a factor vanish and are left out of the term list).
struct ZLayout {
    int lsa = 0, lsm = 0, lsm3 = 0;  // (l3a+1)^2, (l2m+1)^2, (lAm+1)^2
    int lAm = 0;                     // magnetic l' range of A (l3max_magnetic)
    int nm3 = 0;                     // nr3_m * lsm3 (magnetic columns of A)
    int oA = 0, oA2 = 0, oA0 = 0, nz = 0;
    int a(int n, int l, int m, int np, int lp, int mp) const {
        return oA + (n * lsa + l * l + l + m) * nm3 + np * lsm3 + lp * lp + lp + mp;
    }
    int a2(int n, int np, int lp, int mp) const { return oA2 + (n * nr2m_ + np) * lsm + lp * lp + lp + mp; }
    int a0(int np, int lp, int mp) const { return oA0 + np * lsm + lp * lp + lp + mp; }
    int nr2m_ = 0;
};

// rho = Re sum_t w_t prod_{s<4} Z[i_ts] with Z[nz] == 1 padding lower-degree terms.
// The gradient is stored transposed ("gather" form): for every target x of Z, the list of
// contributions w_t * prod_{s' != s} Z[i_ts'] of the monomials containing x, split into
// chunks of at most kChunk contributions.  rho = Re sum_{x in A0} Z[x] dRho/dZ[x], because
// every monomial has exactly one A0 factor.
struct TermList {
    static constexpr int kChunk = 32;
    static constexpr int kLanes = 512;  // threads per adjoint workgroup the chunk order is built for
    std::vector<uint16_t> idx;   // 4 per term (slot 0 = A0 factor)
    std::vector<double> w;       // merged weights, density-major: w[d * nterms + t]
    std::vector<uint16_t> cidx;  // 4 per contribution: the three other factors + padding
    std::vector<double> cw;      // weight per contribution, density-major (d * ncontrib + r)
    std::vector<int32_t> chunks; // 2 per chunk: first record, target | (count << 12) | (sole << 18)
                                 // | (slot << 19), count = the length of its group of 64 chunks
                                 // (padded with zero weights), sole = the target's only chunk,
                                 // slot = the partial-sum slot of its run (below); records of a
                                 // chunk are 64 apart (wave-interleaved layout)
    // Runs: chunk c runs on lane c % kLanes in round c / kLanes; a run is a maximal sequence of
    // a lane's chunks in consecutive rounds with one target.  The adjoint flushes one sum per
    // run: a sole chunk's sum is its target's gradient and is stored; the other runs store
    // their sums into slots, numbered so that a target's slots are consecutive.  Two passes
    // then add them up in a fixed order (the evaluation is bitwise repeatable): the first sums
    // groups of up to kGroup consecutive slots of a target (loads issued together) and writes
    // the result into the group's first slot, or straight into the gradient when the target has
    // one group; the second sums the group results of targets with several groups.
    // comb: 2 per group: first slot, target | (slots << 12) | (the target's only group << 16);
    // comb2: 2 per target with several groups: first slot, target | (groups << 12).
    static constexpr int kGroup = 8;
    std::vector<int32_t> comb, comb2;
    int nslot = 0;
    // folding of the gradient onto representatives of conjugate pairs (x, xbar):
    // partner | sign << 16 (1 = odd m+m') | representative << 17 | self-conjugate << 18
    std::vector<uint32_t> fold;
    // Products of two A entries for monomials of five or six factors (body orders 5 and 6):
    // entry i is Y_i = Z[a] Z[b] at index nz + 1 + i of the extended vector Z' = [Z | 1 | Y],
    // so that every monomial has at most four factors of Z' (a five-factor monomial
    // z0 z1 z2 z3 z4 is z0 z1 z2 Y(z3 z4), a six-factor one z0 z1 Y(z2 z3) Y(z4 z5)).
    // ypair: 2 per entry (a, b).  The gradient of a Y entry goes back to its factors after the
    // record walk: yft holds 2 per Z target with such terms (first entry, target | count << 12),
    // yfe 2 per term (Z' index of Y, the other factor): dRho/dZ[x] += dRho/dY Z[other].
    int ny = 0;
    std::vector<int32_t> ypair, yft, yfe;
    int nzx() const { return nzx_; }
    int nzx_ = 0;                // nz + 1 + ny
    // coefficient structure: B function k (orders 2..6 in file order) -> monomials,
    // rho = sum_k c_k B_k with B_k = Re sum_t u_{k,t} prod Z[idx_t]; CSR by k and by t
    std::vector<int32_t> ub_ptr, ub_t;
    std::vector<double> ub_u;
    std::vector<int32_t> ut_ptr, ut_k;
    std::vector<double> ut_u;
    std::vector<int32_t> rec_t;  // monomial of each gradient record (-1: padding)
    int n2 = 0, n3 = 0, n4 = 0, n5 = 0, n6 = 0;  // raw (before merging) m-combination counts per order
    int raw = 0;
    int nd = 1;                  // densities
    size_t nterms() const { return idx.size() / 4; }
    size_t ncontrib() const { return rec_t.size(); }
};

// Spline tables in node-major layout: row n holds [4 coeffs] for every function family.
struct Splines {
    int nlut = 10000;                    // src/radial.F90:43
    int nfa = 0, nfm = 0;                // functions per atomic / magnetic row
    int atomic_stride = 0, magnetic_stride = 0;  // doubles per row
    double rs_a = 0, rs_m = 0;           // nlut / cutoff
    // atomic row: [g_1..g_nb_a | R_{l,n} (l major, n minor) | core], each 4 coefficients
    std::vector<double> atomic;
    // magnetic row: [g^m_1..g^m_nb_m | M_{l',n'} (l' major)], each 4 coefficients
    std::vector<double> magnetic;
};

struct Model {
    Potential pot;
    ZLayout z;
    TermList terms;
    Splines spl;
};

// api.cpp: the library's last-error state (magace_last_error / magace_error_string)
void report_error(int code, const std::string& msg);
void reset_error();
// api.cpp: a counter that changes whenever state baked into a captured evaluation changes
// (a new or finalised potential, reallocated per-atom buffers, the precision mode, new
// coefficients); a graph captured at one value must not be replayed at another
unsigned long long context_generation();
// api.cpp: uploads the potential's tables and sizes the per-atom buffers for natom atoms
// (no evaluation), so that a following capture records no allocation; MAGACE_ERR_STATE
// without a potential
int context_prepare(int natom);
// noalloc.hip: the fixed device scratch of magace_compute_device_noalloc
int scratch_setup();                 // looks up the scratch address on the current device
void* scratch_device_ptr(int dev);   // nullptr until scratch_setup ran on that device
size_t scratch_bytes();
// potential.cpp
bool parse_potential(const std::string& path, const std::string& element, int xc, double smear,
                     Potential& pot, std::string& err);
// writes the potential with its current parameters in the file format it was read from
// (record order of src/util.F90:1110-1362: par, core, rad_atomic, rad_magnetic, clu 1..4)
bool write_potential(const std::string& path, const Potential& pot, std::string& err);
// basis.cpp
bool check_header(const Header& h, std::string& err);
Basis enumerate_basis(const Header& h);
double clebsch_gordan(int j1, int j2, int J, int m1, int m2, int M);
ZLayout make_layout(const Header& h);
TermList build_terms(const Potential& pot, const ZLayout& z);
// splines.cpp
void radial_basis(int radtype, double lambda, int nb, double cut, double dcut, double r,
                  double* g, double* dg);
void core_repulsion(double r, double pre, double lambda, double cut, double& cr, double& dcr);
Splines build_splines(const Potential& pot);

}  // namespace magace
