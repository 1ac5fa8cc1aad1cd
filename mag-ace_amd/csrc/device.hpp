// device.hpp -- device-side parameter block and launch interface of the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace magace {

template <typename T>
struct cplx {
    T re, im;
};

// threads of an adjoint workgroup; chunk c of the gradient records runs on thread
// c % kAdjointThreads, and the chunk order built with the records assumes this count
constexpr int kAdjointThreads = 512;
// slots per group of the adjoint's first combine pass (TermList::kGroup)
constexpr int kCombGroup = 8;

// Everything a kernel needs about the potential, passed by value as a kernel argument.
template <typename T>
struct DevParams {
    const T* lut_a = nullptr;  // atomic spline rows [nlut][stride_a]
    const T* lut_m = nullptr;  // magnetic spline rows [nlut][stride_m]
    int stride_a = 0, stride_m = 0, nlut = 0;
    double rs_a = 0, rs_m = 0;  // nlut / cutoff (index arithmetic stays FP64)
    int nb_a = 0, nb_m = 0, nr2m = 0, l2m = 0, nr3a = 0, nr3m = 0, l3a = 0, l3m = 0;
    int lsa = 0, lsm = 0, lsm3 = 0, nm3 = 0, oA2 = 0, oA0 = 0, nz = 0;
    int lha = 0, lhm = 0;  // (l+1)(l+2)/2 for l3a, l2m
    signed char mpq[32] = {0};  // m' of the magnetic column q of A (q < nm3 <= 32)
    int sa = 0;                 // k_density: neighbour split of the A tasks (0 = automatic)
    // gradient contributions, one 16-byte record each: x = a | b << 16, y = c as byte offsets
    // of the factors in the adjoint's LDS copy of Z (index * sizeof(cplx<T>)), (z, w) = weight
    // (FP64 bits as lo/hi words, or FP32 bits in z)
    const uint4* crec = nullptr;
    const uint4* crec16 = nullptr;   // FP32 only: the records with offsets of 16-byte entries (two
                                     // atoms' Z interleaved, k_adjoint_pair)
    // FP32 without the accumulation recipe: the same records with the three offsets in x, y, z
    // as they are (no unpacking) and the FP32 weight in w; 8-byte entries (crecu) and 16-byte
    // entries (crec16u, k_adjoint_pair)
    const uint4* crecu = nullptr;
    const uint4* crec16u = nullptr;
    const int2* chunks = nullptr;    // (start, target | count << 12 | sole << 18 | slot << 19), count = its
                                     // group's length; chunk c runs on thread c % kAdjointThreads
    // partial-sum combine (TermList::comb, comb2): groups of <= kCombGroup slots
    // (first slot, target | slots << 12 | only group << 16) and targets with several groups
    // (first slot, target | groups << 12)
    const int2* comb = nullptr;
    const int2* comb2 = nullptr;
    int ncomb = 0, ncomb2 = 0, nslot = 0;  // groups, such targets, partial-sum slots per atom and density
    const uint32_t* fold = nullptr;  // conjugate-pair folding of W (see TermList::fold)
    // products Y_i = Z[a] Z[b] at index nz + 1 + i (body orders 5, 6; TermList::ypair), and the
    // way of their gradient back to Z: per target (first term, target | terms << 12), per term
    // (index of Y, the other factor)
    const int2* ypair = nullptr;
    const int2* yft = nullptr;
    const int2* yfe = nullptr;
    int ny = 0, nyt = 0, nzx = 0;  // products, targets with product terms, nz + 1 + ny
    int nchunks = 0;
    int nfemb = 0;
    int nd = 1;                  // densities (1 or 2)
    const T* cw2 = nullptr;      // record weights of density 2 (same order as crec)
    const double* cw2d = nullptr;  // the same in FP64 (FP32 kernels' accumulation recipe)
    double p1[2][2] = {{1, 1}, {1, 1}};  // per density: embedding prefactor, exponent
    double ecut = 0, decut = 1;
    double c1[2][32] = {{0}};    // per density: coefficients of B1_n = g^m_n(mu_i) - 1
};

// Error bits raised by kernels (reported by the host API).
enum : int {
    kErrSmallDistance = 1,   // r < rcut/nlut ("Encountered very small distance.")
    kErrMomentRange = 2,     // |m| outside [0, rcut_magnetic*(1+1/nlut)) (reference reads out of bounds)
    kErrTooManyNeighbors = 4,
};

struct Buffers {
    void* Z = nullptr;       // [natom][nz] complex
    void* W = nullptr;       // [natom][nz] complex
    double* ehc = nullptr;   // core sum
    double* rho1 = nullptr;  // B1 part of rho, [natom][nd]
    double* dE = nullptr;    // fcut * F'(rho), [natom][nd]
    double* edf = nullptr;   // d f_in/d e_hc * F + 1
    double* rho = nullptr;   // rho_i (including the B1 part), [natom][nd]
    void* Pt = nullptr;      // adjoint partial sums, [natom][nd][nslot] x 16 bytes (partials_bytes)
    void* Wd = nullptr;      // fitting with two densities: [natom][2][nz] per-density dRho/dZ (FP64),
                             // written by the adjoint when set
    int* err = nullptr;
    size_t cap_atoms = 0;
};

// bytes of Buffers::Pt per atom
__host__ __device__ inline size_t partials_bytes(int nslot, int nd) { return (size_t)nslot * 16 * (size_t)nd; }

// atom_base: the index of the launch's first atom in the whole system (even), which fixes the
// adjoint's walk per atom, so that results do not depend on how the atoms are split into
// launches
// Density staging variants (Launch::run's dvar; -1 = the built-in rule): 0 = 32 neighbours per
// round, each pair's staging spread over 8 threads, no neighbour split of the A tasks; 1 = the
// same with the split (DevParams::sa); 2 = 64 neighbours per round, one thread per pair and
// side, with the split.
constexpr int kDensityVariants = 3;

template <typename T>
struct Launch {
    static hipError_t run(const DevParams<T>& P, int natom, int atom_base, const int* offsets,
                          const double* dlist, const double* elist, const double* dmomi,
                          const double* emomi, const double* dmom, const double* emom, double* eout,
                          double* flist, double* tlist, double* tlist0, Buffers& B, hipStream_t s,
                          int nm3_class, hipEvent_t* ev, int dvar = -1, bool acc = false);
};

// fitting (FP64): forward-mode density, per-atom adjoint scalars, monomial accumulation
struct FitBuffers {
    void* Zdot = nullptr;        // [natom][nz] complex
    double* ehc_dot = nullptr;   // [natom]
    double* alpha = nullptr;     // [natom][nd]
    double* beta = nullptr;      // [natom][nd]
    double* M = nullptr;         // [nd][nterms] complex
    void* Wd = nullptr;          // [natom][2][nz] complex: per-density dRho/dZ (two densities)
    unsigned short* idx = nullptr;  // [nterms][4] monomial factors
    int *ub_ptr = nullptr, *ub_t = nullptr;
    double* ub_u = nullptr;
    size_t cap_atoms = 0;
};
hipError_t fit_backward(const DevParams<double>& P, int cls, int natom, const int* off, const int* nlist,
                        const double* du, const double* dlist, const double* elist, const double* dmomi,
                        const double* dmom, const double* emom, const int* atom_struct, const double* gE,
                        Buffers& B, FitBuffers& F, int nterms, const int* nbo5, double* grad, hipStream_t s);
hipError_t fit_assemble(int natom, const int* off, const int* nlist, const double* flist, const double* eout,
                        const int* atom_struct, int nstruc, double* f, double* estruct, hipStream_t s);

hipError_t upload_legendre_constants();
hipError_t clear_flags(int* p, hipStream_t s);
int supported_nm3_class(int nm3, int nr2m, int nb_a, int l3a, int l2m, int nr3a);

}  // namespace magace
