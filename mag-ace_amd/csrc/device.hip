// device.hip -- HIP kernels (gfx950) for the magnetic-ACE energy / force / torque path.
//
// One evaluation (the work of libace's acelib_, src/ace.F90:280-286) is three launches:
//
//   k_density  (prepA, src/a_basis.F90:59-359)   one workgroup per atom.  Pair radial
//              functions and Y_lm of the bond and of the neighbour moment are staged in LDS
//              for a chunk of neighbours; threads then own (lm, n-block) columns of A and
//              (l'm', n-block) columns of A2 and accumulate in registers.  Writes
//              Z = [A | A2 | A0] per atom, the core sum and the B1 part of rho.
//   k_adjoint  (prepB + forceweights, src/b_basis.F90:62-307, src/forces.F90:84-349)
//              one workgroup per atom (two in FP32).  Z and dZ in LDS; each thread sums
//              chunks of gradient records w*Z[a]Z[b]Z[c] per target (the transposed
//              monomial table), partial sums meet in a fixed order (no atomics); then rho,
//              E_i, dE/drho and the scaled gradient W = dE/dZ are written.
//   k_forces   (forcecompute + torquecompute, src/forces.F90:418-623, 688-961) one lane per
//              pair.  Splines and Y_lm are recomputed in registers (nothing per pair is
//              stored between launches); F_p = dE_i/du_p and T_p = dE_i/dm_j are
//              contractions of W with the pair's derivatives; lane 0 adds T0_i = dE_i/dm_i.
//
// T = double (FP64 path) or float (mixed path: FP32 splines, Y_lm, A, B products; FP64 for
// rho, the energy and all outputs).
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.hpp"

namespace magace {

constexpr int kLmaxC = 12;
constexpr int kLhC = (kLmaxC + 1) * (kLmaxC + 2) / 2;

// Legendre recursion coefficients (src/spherical.F90:18-49): a_lm, b_lm (m <= l-2), c_l, d_l
__constant__ double c_alm_d[kLhC], c_blm_d[kLhC], c_cl_d[kLmaxC + 1], c_dl_d[kLmaxC + 1];
__constant__ float c_alm_f[kLhC], c_blm_f[kLhC], c_cl_f[kLmaxC + 1], c_dl_f[kLmaxC + 1];

template <typename T> struct LegC;
template <> struct LegC<double> {
    __device__ static double a(int k) { return c_alm_d[k]; }
    __device__ static double b(int k) { return c_blm_d[k]; }
    __device__ static double c(int l) { return c_cl_d[l]; }
    __device__ static double d(int l) { return c_dl_d[l]; }
};
template <> struct LegC<float> {
    __device__ static float a(int k) { return c_alm_f[k]; }
    __device__ static float b(int k) { return c_blm_f[k]; }
    __device__ static float c(int l) { return c_cl_f[l]; }
    __device__ static float d(int l) { return c_dl_f[l]; }
};

hipError_t upload_legendre_constants() {
    double a[kLhC] = {0}, b[kLhC] = {0}, c[kLmaxC + 1] = {0}, d[kLmaxC + 1] = {0};
    float af[kLhC], bf[kLhC], cf[kLmaxC + 1], df[kLmaxC + 1];
    for (int l = 1; l <= kLmaxC; ++l) {
        const double lsq = (double)(l * l), l1 = 4.0 * lsq - 1.0, l2 = lsq - 2.0 * l + 1.0;
        for (int m = 0; m <= l - 2; ++m) {
            const double msq = (double)(m * m);
            const int k = l * (l + 1) / 2 + m;
            a[k] = std::sqrt(l1 / (lsq - msq));
            b[k] = -std::sqrt((l2 - msq) / (4.0 * l2 - 1.0));
        }
        c[l] = -std::sqrt(1.0 + 0.5 / (double)l);
        d[l] = std::sqrt(2.0 * (double)(l - 1) + 3.0);
    }
    for (int i = 0; i < kLhC; ++i) af[i] = (float)a[i], bf[i] = (float)b[i];
    for (int i = 0; i <= kLmaxC; ++i) cf[i] = (float)c[i], df[i] = (float)d[i];
    hipError_t e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_alm_d), a, sizeof a))) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_blm_d), b, sizeof b))) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_cl_d), c, sizeof c))) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_dl_d), d, sizeof d))) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_alm_f), af, sizeof af))) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_blm_f), bf, sizeof bf))) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_cl_f), cf, sizeof cf))) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_dl_f), df, sizeof df);
}

// ---------------------------------------------------------------- small helpers
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ void cfma(cplx<T>& acc, cplx<T> a, cplx<T> b) {
    acc.re = fma(a.re, b.re, fma(-a.im, b.im, acc.re));
    acc.im = fma(a.re, b.im, fma(a.im, b.re, acc.im));
}
template <typename T>
__device__ __forceinline__ T re_mul(cplx<T> a, cplx<T> b) {  // Re(a*b)
    return fma(a.re, b.re, -a.im * b.im);
}
template <typename T>
__device__ __forceinline__ cplx<T> cconj_sign(cplx<T> a, int m) {  // (-1)^m conj(a)
    return (m & 1) ? cplx<T>{-a.re, a.im} : cplx<T>{a.re, -a.im};
}

// Kahan step: s += x, with c holding minus the part of the running sum that s has lost
// (the compensated sum is s - c)
template <typename T>
__device__ __forceinline__ void kahan_add(T& s, T& c, T x) {
    const T y = x - c;
    const T t = s + y;
    c = (t - s) - y;
    s = t;
}

// cubic Hermite row entry: value and derivative (d/dr) at fractional position w
template <typename T>
__device__ __forceinline__ void horner(const T* __restrict__ c, T w, T rs, T& v, T& dv) {
    const T c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
    v = fma(w, fma(w, fma(w, c3, c2), c1), c0);
    dv = fma(w, fma(w, (T)3 * c3, (T)2 * c2), c1) * rs;
}

// The four coefficients of one spline function as one or two 16-byte loads (rows and functions
// are 16-byte aligned: 4 coefficients per function)
__device__ __forceinline__ void load_coef(const float* __restrict__ p, float (&c)[4]) {
    const float4 v = *(const float4*)p;
    c[0] = v.x; c[1] = v.y; c[2] = v.z; c[3] = v.w;
}
__device__ __forceinline__ void load_coef(const double* __restrict__ p, double (&c)[4]) {
    const double2 a = *(const double2*)p, b = *(const double2*)(p + 2);
    c[0] = a.x; c[1] = a.y; c[2] = b.x; c[3] = b.y;
}

// Values of the spline functions koff + k, k0 <= k < k1, of one table row at fraction w
// (horner's value, in the same operation order), with the coefficient loads of eight functions
// (four in FP64) issued together so that one trip to L2 serves them all instead of one; dflt for every
// function when the row is absent.  put(k, v) stores each.  `safe` is any readable coefficient
// address, read in place of the row's where a function is past k1 or the row is absent.
template <typename T, typename F>
__device__ __forceinline__ void spline_values(const T* __restrict__ row, const T* __restrict__ safe, T w, int k0,
                                              int k1, int koff, T dflt, F&& put) {
    constexpr int B = sizeof(T) == 4 ? 8 : 4;  // FP64: four functions (32 registers) at a time
    for (int k = k0; k < k1; k += B) {
        T c[B][4];
#pragma unroll
        for (int u = 0; u < B; ++u) load_coef(row && k + u < k1 ? row + 4 * (koff + k + u) : safe, c[u]);
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (k + u < k1) put(k + u, row ? fma(w, fma(w, fma(w, c[u][3], c[u][2]), c[u][1]), c[u][0]) : dflt);
    }
}

// The same with horner's derivative as well: put(k, v, dv), dv = 0 where the row is absent
template <typename T, typename F>
__device__ __forceinline__ void spline_values_dv(const T* __restrict__ row, const T* __restrict__ safe, T w, T rs,
                                                 int k0, int k1, int koff, T dflt, F&& put) {
    constexpr int B = sizeof(T) == 4 ? 8 : 4;
    for (int k = k0; k < k1; k += B) {
        T c[B][4];
#pragma unroll
        for (int u = 0; u < B; ++u) load_coef(row && k + u < k1 ? row + 4 * (koff + k + u) : safe, c[u]);
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (k + u < k1) {
                const T v = row ? fma(w, fma(w, fma(w, c[u][3], c[u][2]), c[u][1]), c[u][0]) : dflt;
                const T dv = row ? fma(w, fma(w, (T)3 * c[u][3], (T)2 * c[u][2]), c[u][1]) * rs : (T)0;
                put(k + u, v, dv);
            }
    }
}

// Spline row and fraction for the atomic table (src/radial.F90:68-84).
// Returns nullptr when the pair is beyond the table (all functions zero).
template <typename T>
__device__ __forceinline__ const T* atomic_row(const DevParams<T>& P, double r, T& w, int* err) {
    const double x = r * P.rs_a;
    const double fl = floor(x);
    if (fl <= 0.0) { atomicOr(err, kErrSmallDistance); return nullptr; }
    if (fl >= (double)P.nlut) return nullptr;
    const int nl = (int)fl;
    w = (T)(x - fl);
    return P.lut_a + (size_t)(nl - 1) * P.stride_a;
}
// magnetic table (src/radial.F90:126-158): nl == 0 -> every function is 1, derivative 0
// (returned as row == nullptr with unit == true)
template <typename T>
__device__ __forceinline__ const T* magnetic_row(const DevParams<T>& P, double mu, T& w, bool& unit,
                                                 int* err) {
    unit = false;
    const double x = mu * P.rs_m;
    const double fl = floor(x);
    if (fl < 0.0 || fl > (double)P.nlut) { atomicOr(err, kErrMomentRange); return nullptr; }
    if (fl == 0.0) { unit = true; return nullptr; }
    w = (T)(x - fl);
    return P.lut_m + (size_t)((int)fl - 1) * P.stride_m;
}

// Normalised associated Legendre functions P[lh], dP[lh] = dP/dz for l <= L, m <= l
// (src/spherical.F90:55-155); lh = l(l+1)/2 + m.  Y_lm(x,y,z) = P[lh] (x+iy)^m.
template <typename T, int L>
__device__ __forceinline__ void legendre(T z, T* P, T* dP) {
    P[0] = (T)1;
    dP[0] = (T)0;
    if (L == 0) return;
    const T sq3 = (T)1.73205080756887719318, sq3o2 = (T)1.22474487139158894067;
    P[1] = sq3 * z;
    dP[1] = sq3;
    P[2] = -sq3o2;
    dP[2] = (T)0;
#pragma unroll
    for (int l = 2; l <= L; ++l) {
#pragma unroll
        for (int m = 0; m <= l - 2; ++m) {
            const int k = l * (l + 1) / 2 + m, k1 = (l - 1) * l / 2 + m, k2 = (l - 2) * (l - 1) / 2 + m;
            const T a = LegC<T>::a(k), b = LegC<T>::b(k);
            P[k] = a * fma(z, P[k1], b * P[k2]);
            dP[k] = a * (P[k1] + fma(z, dP[k1], b * dP[k2]));
        }
        const int kp = (l - 1) * l / 2 + (l - 1);
        const T t = LegC<T>::d(l) * P[kp];
        P[l * (l + 1) / 2 + l - 1] = t * z;
        dP[l * (l + 1) / 2 + l - 1] = t;
        P[l * (l + 1) / 2 + l] = LegC<T>::c(l) * P[kp];
        dP[l * (l + 1) / 2 + l] = (T)0;
    }
}

// Y_lm (m >= 0) and its tangential gradient (src/spherical.F90:185-260), from the Legendre
// values and the phase powers ph[m] = (x+iy)^m.
template <typename T>
__device__ __forceinline__ void ylm_grad(int m, T Plm, T dPlm, cplx<T> phm, cplx<T> phm1, T x, T y,
                                         T z, cplx<T>& Y, cplx<T> (&g)[3]) {
    Y = {Plm * phm.re, Plm * phm.im};
    cplx<T> gx, gy, gz;
    if (m == 0) {
        gx = {0, 0};
        gy = {0, 0};
        gz = {dPlm, 0};
    } else {
        const T s = (T)m * Plm;
        gx = {s * phm1.re, s * phm1.im};
        gy = {-gx.im, gx.re};
        gz = {dPlm * phm.re, dPlm * phm.im};
    }
    const cplx<T> rdy = {x * gx.re + y * gy.re + z * gz.re, x * gx.im + y * gy.im + z * gz.im};
    g[0] = {gx.re - rdy.re * x, gx.im - rdy.im * x};
    g[1] = {gy.re - rdy.re * y, gy.im - rdy.im * y};
    g[2] = {gz.re - rdy.re * z, gz.im - rdy.im * z};
}

// half set of Y_lm (m >= 0) written to out[lh]
template <typename T, int L>
__device__ __forceinline__ void ylm_half(T x, T y, T z, int lmax, cplx<T>* out) {
    T P[(L + 1) * (L + 2) / 2], dP[(L + 1) * (L + 2) / 2];
    legendre<T, L>(z, P, dP);
    cplx<T> ph = {1, 0};
    const cplx<T> e = {x, y};
#pragma unroll
    for (int m = 0; m <= L; ++m) {
#pragma unroll
        for (int l = m; l <= L; ++l) {
            if (l <= lmax) {
                const int k = l * (l + 1) / 2 + m;
                out[k] = {P[k] * ph.re, P[k] * ph.im};
            }
        }
        ph = cmul(ph, e);
    }
}

template <typename T>
__device__ __forceinline__ cplx<T> yfull(const cplx<T>* half, int l, int m) {
    const cplx<T> v = half[l * (l + 1) / 2 + (m < 0 ? -m : m)];
    return m < 0 ? cconj_sign(v, -m) : v;
}

// ============================================================================ k_density
// Shapes: NB n-block of A columns, NM3X max magnetic columns of A, NB2 n-block of A2,
// NR2X max magnetic radial index of A2.
// Shape class 0: 4 waves per SIMD in FP32 (128 registers; with the eight-function spline
// batches 127 are used, and 5 waves at 96 registers spilled and ran no faster) and 3 in FP64
// (168 registers; at 169 the occupancy halves and the kernel takes 1.85 instead of 1.39 ms).  Class 1 (27
// magnetic columns): 3 waves per SIMD in FP32 (168 registers, a 12-byte spill; on the shape
// of the example potential 0.75 -> 0.57 ms) and 2 in FP64 (256 registers with a 192-byte
// spill instead of 302 registers at 1 wave: 2.06 -> 1.33 ms).
// Stores of the m >= 0 density tasks (k_density, k_density_jvp): A[n][l,m][q] for the task's
// radial block and, for m > 0, the conjugate partners
// A[n][l,-m][qbar] = (-1)^(m+m') conj(A[n][l,m][q]) with q = n' lsm3 + l'^2 + l' + m' and
// qbar = n' lsm3 + l'^2 + l' - m'; A2[n][n'][l',m'] and A2[n][n'][l',-m'] = (-1)^m' conj(.).
template <typename T, int NB, int NM3X, int NACC, int NQ>
__device__ __forceinline__ void store_a_half(const DevParams<T>& P, const cplx<T> (&acc)[NACC][NQ], int n0, int l,
                                             int m, cplx<T>* __restrict__ Z) {
    const int lm = l * l + l + m;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int n = n0 + b;
        if (n < P.nr3a)
#pragma unroll
            for (int q = 0; q < NM3X; ++q)
                if (q < P.nm3) {
                    Z[(n * P.lsa + lm) * P.nm3 + q] = acc[b][q];
                    if (m > 0) {
                        // m' of column q (a kernel argument: q is a compile-time index); its
                        // partner is column q - 2 m'
                        const int mp = P.mpq[q];
                        const T sg = ((m + mp) & 1) ? (T)-1 : (T)1;
                        Z[(n * P.lsa + lm - 2 * m) * P.nm3 + q - 2 * mp] = {sg * acc[b][q].re, -sg * acc[b][q].im};
                    }
                }
    }
}
template <typename T, int NB2, int NR2X, int NACC, int NQ>
__device__ __forceinline__ void store_a2_half(const DevParams<T>& P, const cplx<T> (&acc)[NACC][NQ], int n0, int m,
                                              int lm, cplx<T>* __restrict__ Z) {
    const T s = (m & 1) ? (T)-1 : (T)1;
#pragma unroll
    for (int b = 0; b < NB2; ++b) {
        const int n = n0 + b;
        if (n < P.nb_a)
#pragma unroll
            for (int q = 0; q < NR2X; ++q)
                if (q < P.nr2m) {
                    cplx<T>* z = Z + P.oA2 + (n * P.nr2m + q) * P.lsm;
                    z[lm] = acc[b][q];
                    if (m > 0) z[lm - 2 * m] = {s * acc[b][q].re, -s * acc[b][q].im};
                }
    }
}

// ACC (FP32 only): A and A2 summed with Kahan compensation (the accumulation recipe of
// MAGACE_MIXED_ACC); twice the accumulator registers, so one wave per SIMD less
template <typename T, int L, int CH, int NB, int NM3X, int NB2, int NR2X, bool SPLIT, bool TASKSPLIT, bool ACC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L == 6 && NM3X <= 4 ? (sizeof(T) == 4 ? (ACC ? 3 : 4) : 3) : (sizeof(T) == 4 && !ACC ? 3 : 2), 8))) k_density(DevParams<T> P, const int* __restrict__ off,
                                                const double* __restrict__ dlist,
                                                const double* __restrict__ elist,
                                                const double* __restrict__ dmomi,
                                                const double* __restrict__ emomi,
                                                const double* __restrict__ dmom,
                                                const double* __restrict__ emom,
                                                cplx<T>* __restrict__ Zout, double* __restrict__ ehc_out,
                                                double* __restrict__ rho1_out, int* err) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int atom = blockIdx.x;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int nfa = P.nr3a * (P.l3a + 1), nfm = P.nr2m * (P.l2m + 1);
    // LDS carve-up (sizes mirror density_lds_bytes on the host).  The per-pair rows have odd
    // lengths (one entry of padding where the count is even), so that the staging threads,
    // which write the rows of consecutive pairs at once, fall in distinct banks
    const int sYs = P.lha | 1, sYms = P.lhm | 1, sHs = P.nm3 | 1, sRs = nfa | 1, sGs = P.nb_a | 1, sMs = nfm | 1;
    cplx<T>* sY = (cplx<T>*)smem;                        // [CH][sYs]
    cplx<T>* sYm = sY + CH * sYs;                         // [CH][sYms]
    cplx<T>* sH = sYm + CH * sYms;                        // [CH][sHs]
    cplx<T>* sYc = sH + CH * sHs;                         // [lhm] central moment
    T* sR = (T*)(sYc + P.lhm);                            // [CH][sRs]  R_{l,n}
    T* sG = sR + CH * sRs;                                // [CH][sGs]
    T* sM = sG + CH * sGs;                                // [CH][sMs]  M_{l',n'}
    T* sCr = sM + CH * sMs;                               // [CH]
    T* sMc = sCr + CH;                                    // [nfm + nb_m] central M and g^m
    double* sEhc = (double*)(((uintptr_t)(sMc + nfm + P.nb_m) + 15) & ~(uintptr_t)15);

    const int p0 = off[atom], nn = off[atom + 1] - p0;
    Zout += (size_t)atom * P.nz;

    // central-atom magnetic functions: A0 and B1 (src/a_basis.F90:69-106, b_basis.F90:62-83),
    // one spline function per thread and the moment's Y_lm on the last wave, so that the block
    // does not wait for one thread to evaluate them all
    for (int k = tid; k < P.nb_m + nfm; k += nthr) {
        T w;
        bool unit;
        const T* row = magnetic_row(P, dmomi[atom], w, unit, err);
        T v = unit ? (T)1 : (T)0, dv;
        if (row) horner(row + 4 * k, w, (T)P.rs_m, v, dv);
        if (k < P.nb_m)
            sMc[nfm + k] = v;
        else
            sMc[k - P.nb_m] = v;
    }
    if (tid == nthr - 64) {
        ylm_half<T, L>((T)emomi[3 * atom], (T)emomi[3 * atom + 1], (T)emomi[3 * atom + 2], P.l2m, sYc);
        *sEhc = 0.0;
    }
    __syncthreads();
    if (tid == 0) {  // B1 part of rho, in the order of the sum over k
        double rho1[2] = {0.0, 0.0};
        for (int k = 0; k < P.nb_m; ++k) {
            const double v = (double)sMc[nfm + k];
            rho1[0] += P.c1[0][k] * (v - 1.0);
            rho1[1] += P.c1[1][k] * (v - 1.0);
        }
        for (int d = 0; d < P.nd; ++d) rho1_out[(size_t)atom * P.nd + d] = rho1[d];
    }
    for (int t = tid; t < P.nr2m * P.lsm; t += nthr) {
        const int np = t / P.lsm, lm = t % P.lsm;
        int l = 0;
        while ((l + 1) * (l + 1) <= lm) ++l;
        const int m = lm - l * l - l;
        const cplx<T> y = yfull(sYc, l, m);
        const T mv = sMc[l * P.nr2m + np];
        Zout[P.oA0 + np * P.lsm + lm] = {mv * y.re, mv * y.im};
    }

    // Tasks cover m >= 0 only (P.lha and P.lhm entries per radial block): since R, M and g are
    // real and Y_{l,-m} = (-1)^m conj(Y_lm),
    //   A[n][l,-m][n',l',-m'] = (-1)^(m+m') conj(A[n][l,m][n',l',m'])  and
    //   A2[n][n'][l',-m'] = (-1)^m' conj(A2[n][n'][l',m']),
    // which are written at the store from the m > 0 sums (m = 0 tasks hold all their columns).
    const int nblkA = (P.nr3a + NB - 1) / NB, nblkA2 = (P.nb_a + NB2 - 1) / NB2;
    const int nTA = P.lha * nblkA, nTA2 = P.lhm * nblkA2;
    // TASKSPLIT: when the A and A2 column tasks leave threads idle, each A task's neighbours
    // are split over SA threads (neighbour j on part j mod SA); the parts meet in LDS at the
    // end, in the staging area, which is free by then.  SA is P.sa when that is set (a tuned
    // choice), otherwise the largest that fits (at most 4), and 1 when that is below 3 (the
    // extra barrier and scratch traffic of SA = 2 cost more than it gained with all-m tasks).
    int SA = 1;
    if constexpr (TASKSPLIT) if (nTA + nTA2 <= nthr) {
        SA = min(P.sa > 0 ? P.sa : 4, (nthr - nTA2) / max(nTA, 1));
        const size_t room = (size_t)((const char*)sEhc - (const char*)smem);
        while (SA > 1 && (size_t)(SA - 1) * nTA * NB * NM3X * sizeof(cplx<T>) > room) --SA;
        if (P.sa <= 0 && SA < 3) SA = 1;
    }
    // A2 tasks start at a wave boundary when that still fits in one round: a wave holding both
    // kinds would run both loops one after the other
    const int nAt = (TASKSPLIT ? SA : 1) * nTA;
    const int a2base = ((nAt + 63) & ~63) + nTA2 <= nthr ? (nAt + 63) & ~63 : nAt;
    for (int round = 0; round * nthr < a2base + nTA2; ++round) {
        const int task = round * nthr + tid;
        cplx<T> acc[NB > NB2 ? NB : NB2][NM3X > NR2X ? NM3X : NR2X];
#pragma unroll
        for (int a = 0; a < (NB > NB2 ? NB : NB2); ++a)
#pragma unroll
            for (int b = 0; b < (NM3X > NR2X ? NM3X : NR2X); ++b) acc[a][b] = {0, 0};
        constexpr int kCa = ACC ? (NB > NB2 ? NB : NB2) : 1, kCq = ACC ? (NM3X > NR2X ? NM3X : NR2X) : 1;
        cplx<T> cmp[kCa][kCq];
#pragma unroll
        for (int a = 0; a < kCa; ++a)
#pragma unroll
            for (int b = 0; b < kCq; ++b) cmp[a][b] = {0, 0};
        // decode task
        // only light headers split tasks; the others keep a compile-time SA = 1
        constexpr bool kSplit = TASKSPLIT;
        const int sa = kSplit ? SA : 1;
        const bool isA = task < sa * nTA, isA2 = task >= a2base && task < a2base + nTA2;
        const int ta = kSplit && isA ? task % nTA : task, part = kSplit && isA ? task / nTA : 0;
        int lh = 0, l = 0, n0 = 0;
        if (isA) { lh = ta % P.lha; n0 = (ta / P.lha) * NB; }
        if (isA2) { lh = (task - a2base) % P.lhm; n0 = ((task - a2base) / P.lhm) * NB2; }
        while ((l + 1) * (l + 2) / 2 <= lh) ++l;
        const int m = lh - l * (l + 1) / 2;  // >= 0
        const int lm = l * l + l + m;

        for (int c0 = 0; c0 < nn; c0 += CH) {
            const int nj = nn - c0 < CH ? nn - c0 : CH;
            __syncthreads();
            if constexpr (SPLIT) {
                // staging of pair c0 + j, spread over the block: thread tid works on pair
                // j = tid % CH in role tid / CH.  Roles 0 and 1, which share a wave: the bond's
                // Y_lm and the core term, and the magnetic side (Y_l'm' of the neighbour moment,
                // the M and the products H_q the A columns need), through one Y_lm call for both;
                // roles 2..nsplit + 1: a share of the radial functions g_n, R_{l,n} and, after
                // them, of the magnetic functions M_{l',n'}
                const int j = tid % CH, role = tid / CH, nsplit = nthr / CH - 2;
                const bool live = j < nj;
                const int p = p0 + c0 + (live ? j : 0);
                if (role <= 1) {
                    const bool bond = role == 0;
                    if (live) {
                        T w = 0;
                        bool unit = false;
                        const T* row = bond ? atomic_row(P, dlist[p], w, err) : magnetic_row(P, dmom[p], w, unit, err);
                        if (bond) {
                            T cr = 0, dcr;
                            if (row) horner(row + 4 * (P.nb_a + nfa), w, (T)P.rs_a, cr, dcr);
                            sCr[j] = cr;
                        }
                        const double* v3 = bond ? elist : emom;
                        cplx<T>* yo = bond ? sY + j * sYs : sYm + j * sYms;
                        // the moment's M coefficients for H (four-column class: l' <= 1, one n'),
                        // loaded before the Y_lm so that their trip overlaps it
                        constexpr bool kMh = NM3X == 4 && sizeof(T) == 4;
                        T cmh[kMh ? 2 : 1][4];
                        if constexpr (kMh)
#pragma unroll
                            for (int lp = 0; lp < 2; ++lp)
                                load_coef(!bond && row && lp <= P.l3m ? row + 4 * (P.nb_m + lp * P.nr2m) : P.lut_m, cmh[lp]);
                        ylm_half<T, L>((T)v3[3 * p], (T)v3[3 * p + 1], (T)v3[3 * p + 2], bond ? P.l3a : P.l2m, yo);
                        if (!bond) {
                            // H_q = M_{l',n'} Y_{l'm'} for the magnetic columns of A (l' <= l3m,
                            // q = n' lsm3 + l'^2 + l' + m'); the few M it needs are evaluated
                            // here rather than waiting for the spline roles
                            for (int np = 0; np * P.lsm3 < P.nm3; ++np)
                                for (int lp = 0; lp <= P.l3m; ++lp) {
                                    T mv = unit ? (T)1 : (T)0, dv;
                                    if (kMh && np == 0 && lp < 2) {
                                        if (row) horner(cmh[kMh && lp < 2 ? lp : 0], w, (T)P.rs_m, mv, dv);
                                    } else if (row) {
                                        horner(row + 4 * (P.nb_m + lp * P.nr2m + np), w, (T)P.rs_m, mv, dv);
                                    }
                                    for (int mp = -lp; mp <= lp; ++mp) {
                                        const int q = np * P.lsm3 + lp * lp + lp + mp;
                                        const cplx<T> y = yfull(yo, lp, mp);
                                        sH[j * sHs + q] = {mv * y.re, mv * y.im};
                                    }
                                }
                        }
                    } else if (bond) {
                        sCr[j] = 0;
                    }
                } else {
                    // a share of the pair's radial functions g_n, R_{l,n} and, after them, of
                    // its magnetic functions M_{l',n'}
                    if (live) {
                        const int nf = P.nb_a + nfa, per = (nf + nfm + nsplit - 1) / nsplit;
                        const int k0 = (role - 2) * per, k1 = min(nf + nfm, k0 + per);
                        if (k0 < nf) {
                            T w = 0;
                            const T* row = atomic_row(P, dlist[p], w, err);
                            spline_values<T>(row, P.lut_a, w, k0, min(k1, nf), 0, (T)0, [&](int k, T v) {
                                if (k < P.nb_a)
                                    sG[j * sGs + k] = v;
                                else
                                    sR[j * sRs + (k - P.nb_a)] = v;
                            });
                        }
                        if (k1 > nf) {
                            T w = 0;
                            bool unit;
                            const T* row = magnetic_row(P, dmom[p], w, unit, err);
                            spline_values<T>(row, P.lut_m, w, max(k0, nf), k1, P.nb_m - nf, unit ? (T)1 : (T)0,
                                             [&](int k, T v) { sM[j * sMs + (k - nf)] = v; });
                        }
                    }
                }
            } else {
                // one thread per pair and side: atomic side (threads 0..CH-1), magnetic
                // side (CH..2CH-1)
                if (tid < CH) {  // atomic side of pair c0+tid
                    const int j = tid;
                    if (j < nj) {
                        const int p = p0 + c0 + j;
                        T w = 0;
                        const T* row = atomic_row(P, dlist[p], w, err);
                        spline_values<T>(row, P.lut_a, w, 0, P.nb_a + nfa, 0, (T)0, [&](int k, T v) {
                            if (k < P.nb_a)
                                sG[j * sGs + k] = v;
                            else
                                sR[j * sRs + (k - P.nb_a)] = v;
                        });
                        T cr = 0, dcr;
                        if (row) horner(row + 4 * (P.nb_a + nfa), w, (T)P.rs_a, cr, dcr);
                        sCr[j] = cr;
                        ylm_half<T, L>((T)elist[3 * p], (T)elist[3 * p + 1], (T)elist[3 * p + 2], P.l3a,
                                       sY + j * sYs);
                    } else {
                        sCr[j] = 0;
                    }
                } else if (tid < 2 * CH) {  // magnetic side of pair c0+j
                    const int j = tid - CH;
                    if (j < nj) {
                        const int p = p0 + c0 + j;
                        T w = 0;
                        bool unit;
                        const T* row = magnetic_row(P, dmom[p], w, unit, err);
                        spline_values<T>(row, P.lut_m, w, 0, nfm, P.nb_m, unit ? (T)1 : (T)0,
                                         [&](int k, T v) { sM[j * sMs + k] = v; });
                        cplx<T>* yh = sYm + j * sYms;
                        ylm_half<T, L>((T)emom[3 * p], (T)emom[3 * p + 1], (T)emom[3 * p + 2], P.l2m, yh);
                        // H_q = M_{l',n'} Y_{l'm'} for the magnetic columns of A (l' <= l3m)
                        for (int q = 0; q < P.nm3; ++q) {
                            const int np = q / P.lsm3, lmp = q % P.lsm3;
                            int lp = 0;
                            while ((lp + 1) * (lp + 1) <= lmp) ++lp;
                            const int mp = lmp - lp * lp - lp;
                            const cplx<T> y = yfull(yh, lp, mp);
                            const T mv = sM[j * sMs + lp * P.nr2m + np];
                            sH[j * sHs + q] = {mv * y.re, mv * y.im};
                        }
                    }
                }
            }
            __syncthreads();
            if (round == 0 && tid == 0) {
                double s = 0.0;
                for (int j = 0; j < nj; ++j) s += (double)sCr[j];
                *sEhc += s;
            }
            if (isA) {
                // A[n][lm][q] += R_{l,n} Y_lm H_q   (src/a_basis.F90:248-291).  Columns past nm3
                // and radial indices past nr3a are read without a test: those reads stay inside
                // the staging area, and their sums land in slots that are never stored.
                for (int j = part; j < nj; j += sa) {
                    // the pair's Y, H and R reads issued together (R read after the products
                    // Y H had waited for LDS a second time); the wide classes keep the later
                    // reads, as they would spill
                    const cplx<T> y = sY[j * sYs + lh];
                    cplx<T> h[NM3X];
#pragma unroll
                    for (int q = 0; q < NM3X; ++q) h[q] = sH[j * sHs + q];
                    const T* rr = sR + j * sRs + l * P.nr3a + n0;
                    constexpr bool kEarly = NM3X <= 4;
                    T rv[kEarly ? NB : 1];
                    if constexpr (kEarly)
#pragma unroll
                        for (int b = 0; b < NB; ++b) rv[b] = rr[b];
                    // Y H_q once per pair and column, then two real FMAs per radial index
                    cplx<T> yh[NM3X];
#pragma unroll
                    for (int q = 0; q < NM3X; ++q) yh[q] = cmul(y, h[q]);
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const T r = kEarly ? rv[kEarly ? b : 0] : rr[b];
#pragma unroll
                        for (int q = 0; q < NM3X; ++q) {
                            if constexpr (ACC) {
                                kahan_add(acc[b][q].re, cmp[ACC ? b : 0][ACC ? q : 0].re, r * yh[q].re);
                                kahan_add(acc[b][q].im, cmp[ACC ? b : 0][ACC ? q : 0].im, r * yh[q].im);
                            } else {
                                acc[b][q].re = fma(r, yh[q].re, acc[b][q].re);
                                acc[b][q].im = fma(r, yh[q].im, acc[b][q].im);
                            }
                        }
                    }
                }
            } else if (isA2) {
                // A2[n][n'][l'm'] += g_n M_{l',n'} Y_{l'm'}   (src/a_basis.F90:193-201); unused
                // slots as above
                for (int j = 0; j < nj; ++j) {
                    // the pair's g reads issued with its Y and M reads, as in the A loop
                    const cplx<T> y = sYm[j * sYms + lh];
                    const T* mr = sM + j * sMs + l * P.nr2m;
                    const T* gr = sG + j * sGs + n0;
                    constexpr bool kEarly = NM3X <= 4;
                    T gv[kEarly ? NB2 : 1];
                    if constexpr (kEarly)
#pragma unroll
                        for (int b = 0; b < NB2; ++b) gv[b] = gr[b];
                    cplx<T> my[NR2X];
#pragma unroll
                    for (int q = 0; q < NR2X; ++q) my[q] = {mr[q] * y.re, mr[q] * y.im};
#pragma unroll
                    for (int b = 0; b < NB2; ++b) {
                        const T g = kEarly ? gv[kEarly ? b : 0] : gr[b];
#pragma unroll
                        for (int q = 0; q < NR2X; ++q) {
                            if constexpr (ACC) {
                                kahan_add(acc[b][q].re, cmp[ACC ? b : 0][ACC ? q : 0].re, g * my[q].re);
                                kahan_add(acc[b][q].im, cmp[ACC ? b : 0][ACC ? q : 0].im, g * my[q].im);
                            } else {
                                acc[b][q].re = fma(g, my[q].re, acc[b][q].re);
                                acc[b][q].im = fma(g, my[q].im, acc[b][q].im);
                            }
                        }
                    }
                }
            }
        }
        if constexpr (ACC) {  // the compensated sums
#pragma unroll
            for (int a = 0; a < kCa; ++a)
#pragma unroll
                for (int b = 0; b < kCq; ++b) {
                    acc[a][b].re -= cmp[a][b].re;
                    acc[a][b].im -= cmp[a][b].im;
                }
        }
        if (kSplit && sa > 1) {  // one round only: every thread of the block reaches these barriers
            cplx<T>* scr = (cplx<T>*)smem;  // [SA - 1][nTA][NB][NM3X]
            __syncthreads();
            if (isA && part > 0)
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int q = 0; q < NM3X; ++q) scr[(((part - 1) * nTA + ta) * NB + b) * NM3X + q] = acc[b][q];
            __syncthreads();
            if (isA && part == 0)
                for (int h = 1; h < sa; ++h)
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int q = 0; q < NM3X; ++q) {
                            const cplx<T> v = scr[(((h - 1) * nTA + ta) * NB + b) * NM3X + q];
                            acc[b][q].re += v.re;
                            acc[b][q].im += v.im;
                        }
        }
        if (isA && part == 0) {
            store_a_half<T, NB, NM3X>(P, acc, n0, l, m, Zout);
        } else if (isA2) {
            store_a2_half<T, NB2, NR2X>(P, acc, n0, m, lm, Zout);
        }
    }
    __syncthreads();
    if (tid == 0) ehc_out[atom] = *sEhc;
}

size_t density_lds_bytes(const DevParams<double>& Pd, int CH, size_t tsize) {
    const int nfa = Pd.nr3a * (Pd.l3a + 1), nfm = Pd.nr2m * (Pd.l2m + 1);
    size_t cb = 2 * tsize;
    // per-pair rows padded to odd lengths (k_density)
    size_t n = cb * ((size_t)CH * (Pd.lha | 1) + (size_t)CH * (Pd.lhm | 1) + (size_t)CH * (Pd.nm3 | 1) + Pd.lhm) +
               tsize * ((size_t)CH * (nfa | 1) + (size_t)CH * (Pd.nb_a | 1) + (size_t)CH * (nfm | 1) + CH + nfm + Pd.nb_m);
    return ((n + 15) & ~(size_t)15) + 16;
}

// ============================================================================ k_adjoint
template <typename T> __device__ __forceinline__ T rec_weight(uint4 r);
template <> __device__ __forceinline__ double rec_weight<double>(uint4 r) {
    return __hiloint2double((int)r.w, (int)r.z);
}
template <> __device__ __forceinline__ float rec_weight<float>(uint4 r) { return __uint_as_float(r.z); }

template <typename T>
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += red[i];
    return s;  // valid on thread 0
}

// embedding F(rho) and F'(rho) (src/forces.F90:1134-1233, nfemb 0..2)
__device__ inline void embed(int nfemb, double rho, double wpre, double mexp, double& F, double& dF) {
    if (nfemb == 0) { F = rho; dF = 1.0; return; }
    const double w = 10.0, eps = 1e-10;
    const double s = nfemb == 2 ? 1.0 + rho : rho;
    if (fabs(s) > eps) {
        double y1 = w * s * s;
        const double g = y1 > 30.0 ? 0.0 : exp(-y1);
        const double omg = 1.0 - g, a = fabs(s), sg = s >= 0.0 ? 1.0 : -1.0;
        y1 = pow(omg * a, mexp);
        const double y2 = mexp * g * a;
        F = nfemb == 2 ? wpre * (sg * (y1 + y2) - 1.0) : wpre * sg * (y1 + y2);
        const double dg = -2.0 * w * s * g, da = sg;
        const double dy = fabs(y1) < eps ? 0.0 : mexp * y1 / (omg * a);
        dF = wpre * sg * (dy * (-dg * a + omg * da) + mexp * (dg * a + g * da));
    } else {
        F = nfemb == 2 ? wpre * mexp * (s - 1.0) : wpre * mexp * rho;
        dF = wpre * mexp;
    }
}

// An aligned vector at LDS byte address `off`. The adjoint kernels declare no static LDS (the
// host checks this before launching them), so their dynamic LDS, and with it Z' at its start,
// begins at address 0: a record's byte offset into Z' is the address itself, and reading it
// this way saves the add of the dynamic base, which the compiler resolves only after
// instruction selection and so cannot fold away.
template <typename E, int N>
__device__ __forceinline__ E __attribute__((ext_vector_type(N))) lds_at(uint32_t off) {
    typedef E V __attribute__((ext_vector_type(N)));
    return *(const __attribute__((address_space(3))) V*)(size_t)off;
}
__device__ __forceinline__ float4 lds_f4(uint32_t off) {
    const auto v = lds_at<float, 4>(off);
    return make_float4(v.x, v.y, v.z, v.w);
}
// one complex entry of Z' at a byte offset, as a single aligned vector read
template <typename T> __device__ __forceinline__ cplx<T> lds_cplx(uint32_t off) {
    const auto v = lds_at<T, 2>(off);
    return {v.x, v.y};
}

// E_i = f_in(e_hc) sum_d F(rho_d) + e_hc (src/forces.F90:1066-1128) for nd densities whose
// B parts sum to tots[d]; writes eout, dE_i/drho_d = fc F'(rho_d) (also into scale[d]), the
// e_hc factor and rho
template <typename T>
__device__ inline void atom_energy(const DevParams<T>& P, int nd, int atom, const double* tots, double e,
                                   const double* r1, double* __restrict__ eout, double* __restrict__ dEout,
                                   double* __restrict__ edfout, double* __restrict__ rho_out, double* scale) {
    // e = e_hc of the atom, r1[d] its B1 parts of rho (loaded by the caller ahead of time)
    double Fs = 0.0, dF[2] = {0.0, 0.0};
    for (int d = 0; d < nd; ++d) {
        const double r = tots[d] + r1[d];
        rho_out[(size_t)atom * nd + d] = r;
        double F;
        embed(P.nfemb, r, P.p1[d][0], P.p1[d][1], F, dF[d]);
        Fs += F;
    }
    const double elow = P.ecut - P.decut;
    double fc = 1.0, dfc = 0.0;
    if (e >= P.ecut) { fc = 0.0; dfc = 0.0; }
    else if (e > elow) {
        const double pi = 3.14159265358979323846;
        fc = 0.5 * (1.0 + cos(pi * (e - elow) / P.decut));
        dfc = -0.5 * sin(pi * (e - elow) / P.decut) * pi / P.decut;
    }
    eout[atom] = fc * Fs + e;
    for (int d = 0; d < nd; ++d) {
        dEout[(size_t)atom * nd + d] = fc * dF[d];
        scale[d] = fc * dF[d];
    }
    edfout[atom] = dfc * Fs + 1.0;
}

__device__ __forceinline__ float4 vadd(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
template <typename T>
__device__ __forceinline__ cplx<T> vadd(cplx<T> a, cplx<T> b) {
    return {a.re + b.re, a.im + b.im};
}

// The adjoint's partial sums of one atom (or atom pair), part[slot], added up per target in a
// fixed order.  First pass: a thread sums one group of <= kCombGroup consecutive slots of a
// target, its loads issued together, into sW[x] when the group is the target's only one and
// into the group's first slot otherwise.  Second pass: a thread sums the group results of one
// target with several groups.  Two round trips to L2 instead of one per slot.
// cb, cb2: this thread's first descriptors (comb[tid], comb2[tid]), loaded ahead by the caller.
template <typename P_t, typename V>
__device__ __forceinline__ void combine_partials(const P_t& P, V* __restrict__ part, V* __restrict__ sW, int2 cb,
                                                 int2 cb2) {
    const int nthr = (int)blockDim.x, tid = (int)threadIdx.x;
    for (int k = tid; k < P.ncomb; k += nthr) {
        const int2 e = k == tid ? cb : P.comb[k];
        const int x = e.y & 0xfff, n = (e.y >> 12) & 15;
        V v[kCombGroup];
#pragma unroll
        for (int u = 0; u < kCombGroup; ++u)
            if (u < n) v[u] = part[e.x + u];
        V sum = v[0];
#pragma unroll
        for (int u = 1; u < kCombGroup; ++u)
            if (u < n) sum = vadd(sum, v[u]);
        if ((e.y >> 16) & 1)
            sW[x] = sum;
        else
            part[e.x] = sum;
    }
    __syncthreads();
    for (int k = tid; k < P.ncomb2; k += nthr) {
        const int2 e = k == tid ? cb2 : P.comb2[k];
        const int x = e.y & 0xfff, ng = e.y >> 12;
        V sum = part[e.x];
        for (int g0 = 1; g0 < ng; g0 += kCombGroup) {
            V v[kCombGroup];
#pragma unroll
            for (int u = 0; u < kCombGroup; ++u)
                if (g0 + u < ng) v[u] = part[e.x + (g0 + u) * kCombGroup];
#pragma unroll
            for (int u = 0; u < kCombGroup; ++u)
                if (g0 + u < ng) sum = vadd(sum, v[u]);
        }
        sW[x] = sum;
    }
}

// ND2: two densities (rho_d = sum_k c_{k,d} B_k, E = fc sum_d F(rho_d) + e_hc,
// src/forces.F90:976-1089): a second weight per record, a second gradient array in LDS
// ACC (FP32 only, MAGACE_MIXED_ACC): rho = sum of the A0 targets' gradient records summed in
// FP64 (double-float record weights: the FP32 weight in z plus its low part in w); in every
// round where some lane of the wave has an A0 target, the whole wave accumulates in FP64
template <typename T, bool ND2, bool ACC>
__global__ void __launch_bounds__(kAdjointThreads) k_adjoint(DevParams<T> P, int atom_base,
                                                const cplx<T>* __restrict__ Zin,
                                                cplx<T>* __restrict__ Wout, cplx<T>* __restrict__ Wd,
                                                char* __restrict__ Pt,
                                                const double* __restrict__ ehc,
                                                const double* __restrict__ rho1,
                                                double* __restrict__ eout, double* __restrict__ dEout,
                                                double* __restrict__ edfout, double* __restrict__ rho_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    // extended vector Z' = [Z | 1 | Y] (nzx entries; Y: products for body orders 5, 6)
    cplx<T>* sZ = (cplx<T>*)smem;           // [nzx], sZ[nz] = 1
    cplx<T>* sW = sZ + P.nzx;               // [nzx]
    cplx<T>* sW2 = sW + P.nzx;              // [nzx] (ND2 only)
    double* red = (double*)(sW + (ND2 ? 2 : 1) * P.nzx);
    // the energy's scale factors after the two block sums; in dynamic LDS, so that it starts at
    // address 0 and the records' offsets address sZ without an add
    double* s_scale = red + 2 * (kAdjointThreads / 64);
    const int atom = blockIdx.x, tid = threadIdx.x;
    Zin += (size_t)atom * P.nz;
    {
        // Z' from Z, with all of a thread's loads issued before the stores
        constexpr int U = 8;  // covers nzx <= 8 * kAdjointThreads = 4096 in one pass
        static_assert(U * kAdjointThreads >= 4096, "one pass over the largest feature vector");
        cplx<T> v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int x = tid + u * kAdjointThreads;
            v[u] = x < P.nz ? Zin[x] : cplx<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int x = tid + u * kAdjointThreads;
            if (x < P.nz) sZ[x] = v[u];
            if (x < P.nzx) {
                sW[x] = {0, 0};
                if (ND2) sW2[x] = {0, 0};
            }
        }
    }
    // what the tail needs from global memory, loaded now so that its latency passes during the
    // record walk: the fold words of this thread's entries, its first combine descriptors and
    // (thread 0) the atom's e_hc and B1 part of rho
    constexpr int kFold = 8;  // nz <= 8 * kAdjointThreads
    uint32_t fw[kFold];
#pragma unroll
    for (int u = 0; u < kFold; ++u) {
        const int x = tid + u * kAdjointThreads;
        fw[u] = x < P.nz ? P.fold[x] : 0u;
    }
    const int2 cb = tid < P.ncomb ? P.comb[tid] : make_int2(0, 0);
    const int2 cb2 = tid < P.ncomb2 ? P.comb2[tid] : make_int2(0, 0);
    double e_hc = 0.0, r1v[2] = {0.0, 0.0};
    if (tid == 0) {
        e_hc = ehc[atom];
        r1v[0] = rho1[(size_t)atom * P.nd];
        if (ND2) r1v[1] = rho1[(size_t)atom * P.nd + 1];
    }
    if (tid == 0) sZ[P.nz] = {1, 0};
    __syncthreads();
    if (P.ny > 0) {
        for (int i = tid; i < P.ny; i += blockDim.x) {
            const int2 pr = P.ypair[i];
            sZ[P.nz + 1 + i] = cmul(sZ[pr.x], sZ[pr.y]);
        }
        __syncthreads();
    }
    // Gather form of the gradient: a thread sums one chunk of contributions
    // w * Z[a] Z[b] Z[c] to a single target x in registers (four 16-byte records per step,
    // loads issued together; the 64 lanes of a wave read consecutive records).  Thread t
    // takes chunks t + kAdjointThreads j; the chunk -> thread mapping must match the record
    // layout built by build_terms.  While the thread's next chunk (round j + 1) has the same
    // target the sum stays in registers; at the end of such a run it is stored: into sW[x]
    // when the chunk is its target's only one, otherwise into the run's slot of the atom's
    // partial sums in global memory, which the pass after the walk adds up per target in slot
    // order.  No atomics: the evaluation is bitwise repeatable.  The walk starts at round rot,
    // a function of the atom's index in the whole system, so that the blocks of an XCD stream
    // different parts of the record table at any moment instead of all reading the same L2
    // lines; a run that crosses from round rot - 1 into round rot is met in two parts, and the
    // first part's sum is held in registers until the walk ends with the second.
    cplx<T>* part = (cplx<T>*)(Pt + (size_t)atom * partials_bytes(P.nslot, P.nd));
    cplx<T>* part2 = part + P.nslot;
    const int nthr = (int)blockDim.x;
    const int nj = (P.nchunks + nthr - 1) / nthr;
    const int rot = nj > 0 ? (int)((unsigned)(atom_base + atom) % (unsigned)nj) : 0;
    cplx<T> acc = {0, 0}, acc2 = {0, 0}, head = {0, 0}, head2 = {0, 0};
    cplx<double> acd = {0, 0}, acd2 = {0, 0};  // ACC: the run's FP64 rounds
    double rho = 0.0, rho2 = 0.0;              // ACC: summed at the flushes
    int j = rot;
    int2 ch = tid + j * nthr < P.nchunks ? P.chunks[tid + j * nthr] : make_int2(0, 0);
    bool split = false;
    if (rot > 0 && tid + rot * nthr < P.nchunks) split = ((P.chunks[tid + (rot - 1) * nthr].y ^ ch.y) & 0xfff) == 0;
    bool hold = split;
    for (int j0 = 0; j0 < nj; ++j0) {
        const int c = tid + j * nthr;
        const int jn = j + 1 < nj ? j + 1 : 0;
        const int cn = tid + jn * nthr;
        const bool has_next = j0 + 1 < nj && cn < P.nchunks;
        const int2 chn = has_next ? P.chunks[cn] : make_int2(0, 0);
        if (c < P.nchunks) {
            const int x = ch.y & 0xfff, cnt = (ch.y >> 12) & 63;
            const bool sole = (ch.y >> 18) & 1;  // the target's only chunk
            // records of this lane, stride 64; FP32 without the recipe reads the table whose
            // offsets need no unpacking
            constexpr bool kU = sizeof(T) == 4 && !ACC;
            const uint4* __restrict__ rec = (kU ? P.crecu : P.crec) + ch.x;
            const T* __restrict__ w2p = ND2 ? P.cw2 + ch.x : nullptr;
            const double* __restrict__ w2d = ND2 && ACC ? P.cw2d + ch.x : nullptr;
            if (ACC && __any(x >= P.oA0 && x < P.nz)) {
                for (int k = 0; k < cnt; k += 4) {
                    uint4 r[4];
                    double w2[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        r[u] = rec[(size_t)(k + u) * 64];
                        if (ND2) w2[u] = w2d[(size_t)(k + u) * 64];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const double w = (double)__uint_as_float(r[u].z) + (double)__uint_as_float(r[u].w);
                        const cplx<T> pr = cmul(cmul(lds_cplx<T>(r[u].x & 0xffff), lds_cplx<T>(r[u].x >> 16)),
                                                lds_cplx<T>(r[u].y));
                        acd.re = fma(w, (double)pr.re, acd.re);
                        acd.im = fma(w, (double)pr.im, acd.im);
                        if (ND2) {
                            acd2.re = fma(w2[u], (double)pr.re, acd2.re);
                            acd2.im = fma(w2[u], (double)pr.im, acd2.im);
                        }
                    }
                }
            } else
            for (int k = 0; k < cnt; k += 4) {
                uint4 r[4];
                T w2[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    r[u] = rec[(size_t)(k + u) * 64];
                    if (ND2) w2[u] = w2p[(size_t)(k + u) * 64];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const T w = kU ? (T)__uint_as_float(r[u].w) : rec_weight<T>(r[u]);
                    // records hold byte offsets into sZ, which starts at LDS address 0
                    const cplx<T> za = lds_cplx<T>(kU ? r[u].x : r[u].x & 0xffff);
                    const cplx<T> zb = lds_cplx<T>(kU ? r[u].y : r[u].x >> 16);
                    const cplx<T> zc = lds_cplx<T>(kU ? r[u].z : r[u].y);
                    const cplx<T> pr = cmul(cmul(za, zb), zc);
                    acc.re = fma(w, pr.re, acc.re);
                    acc.im = fma(w, pr.im, acc.im);
                    if (ND2) {
                        acc2.re = fma(w2[u], pr.re, acc2.re);
                        acc2.im = fma(w2[u], pr.im, acc2.im);
                    }
                }
            }
            if (!(jn == j + 1 && has_next && ((chn.y ^ ch.y) & 0xfff) == 0)) {
                if constexpr (ACC) {  // the run's FP64 and FP32 rounds; rho of an A0 target in FP64
                    const cplx<double> t = {acd.re + (double)acc.re, acd.im + (double)acc.im};
                    const cplx<double> t2 = {acd2.re + (double)acc2.re, acd2.im + (double)acc2.im};
                    if (x >= P.oA0 && x < P.nz) {
                        const cplx<T> z = sZ[x];
                        rho += (double)z.re * t.re - (double)z.im * t.im;
                        if (ND2) rho2 += (double)z.re * t2.re - (double)z.im * t2.im;
                    }
                    acc = {(T)t.re, (T)t.im};
                    acc2 = {(T)t2.re, (T)t2.im};
                    acd = {0, 0};
                    acd2 = {0, 0};
                }
                if (hold) {  // first part of the run split by the start of the walk
                    head = acc;
                    head2 = acc2;
                    hold = false;
                } else {
                    if (split && j0 + 1 == nj) {  // its second part: the whole run
                        acc.re += head.re;
                        acc.im += head.im;
                        acc2.re += head2.re;
                        acc2.im += head2.im;
                    }
                    if (sole) {
                        sW[x] = acc;
                        if (ND2) sW2[x] = acc2;
                    } else {
                        const int slot = (int)((unsigned)ch.y >> 19);
                        part[slot] = acc;
                        if (ND2) part2[slot] = acc2;
                    }
                }
                acc = {0, 0};
                acc2 = {0, 0};
            }
        }
        ch = chn;
        j = jn;
    }
    __syncthreads();  // every run stored: sole targets in sW, the others in their slots
    combine_partials(P, part, sW, cb, cb2);
    if (ND2) {
        __syncthreads();
        combine_partials(P, part2, sW2, cb, cb2);
    }
    __syncthreads();
    // products Y: dRho/dZ[x] += dRho/dY Z[other] for the Y entries that hold Z[x], one thread
    // per target in a fixed order (the targets read only Y entries, and write only their own)
    if (P.nyt > 0) {
        for (int k = tid; k < P.nyt; k += nthr) {
            const int2 e = P.yft[k];
            const int x = e.y & 0xfff, n = e.y >> 12;
            cplx<T> w = sW[x], w2 = ND2 ? sW2[x] : cplx<T>{0, 0};
            for (int u = 0; u < n; ++u) {
                const int2 t = P.yfe[e.x + u];
                cfma(w, sW[t.x], sZ[t.y]);
                if (ND2) cfma(w2, sW2[t.x], sZ[t.y]);
            }
            sW[x] = w;
            if (ND2) sW2[x] = w2;
        }
        __syncthreads();
    }
    // rho = Re sum over A0 targets of Z[x] dRho/dZ[x] (every monomial has exactly one A0
    // factor), summed in FP64 in a fixed order (with ACC: already summed at the flushes)
    if constexpr (!ACC)
        for (int x = P.oA0 + tid; x < P.nz; x += nthr) {
            rho += (double)re_mul(sZ[x], sW[x]);
            if (ND2) rho2 += (double)re_mul(sZ[x], sW2[x]);
        }
    const double tot = block_sum<T>(rho, red);
    double tot2 = 0.0;
    if (ND2) {
        __syncthreads();
        tot2 = block_sum<T>(rho2, red + kAdjointThreads / 64);
    }
    if (tid == 0) {
        const double tots[2] = {tot, tot2};
        atom_energy<T>(P, ND2 ? 2 : 1, atom, tots, e_hc, r1v, eout, dEout, edfout, rho_out, s_scale);
    }
    __syncthreads();
    // W_eff on representatives of conjugate pairs: since Z_xbar = s conj(Z_x),
    // Re(W_x dZ_x + W_xbar dZ_xbar) = Re((W_x + s conj(W_xbar)) dZ_x); zero elsewhere
    // (the reference's half-range loops with factor 2, src/forces.F90:460-614)
    // With ND2 and Wd set (fitting), each density's folded gradient dRho_d/dZ is also written,
    // unscaled: Wd[atom][d][x]
    const T sc = (T)s_scale[0], sc2 = ND2 ? (T)s_scale[1] : (T)0;
    Wout += (size_t)atom * P.nz;
    cplx<T>* wd = ND2 && Wd ? Wd + (size_t)atom * 2 * P.nz : nullptr;
#pragma unroll
    for (int u = 0; u < kFold; ++u) {
        const int x = tid + u * kAdjointThreads;
        if (x >= P.nz) break;
        const uint32_t f = fw[u];
        cplx<T> w = {0, 0}, w2 = {0, 0};
        if (f & (1u << 17)) {
            w = sW[x];
            if (ND2) w2 = sW2[x];
            if (!(f & (1u << 18))) {
                const cplx<T> b = sW[f & 0xffff];
                const T s = (f & (1u << 16)) ? (T)-1 : (T)1;
                w.re += s * b.re;
                w.im -= s * b.im;
                if (ND2) {
                    const cplx<T> b2 = sW2[f & 0xffff];
                    w2.re += s * b2.re;
                    w2.im -= s * b2.im;
                }
            }
        }
        Wout[x] = ND2 ? cplx<T>{sc * w.re + sc2 * w2.re, sc * w.im + sc2 * w2.im} : cplx<T>{sc * w.re, sc * w.im};
        if (wd) {
            wd[x] = w;
            wd[P.nz + x] = w2;
        }
    }
}

// Two atoms per block (FP32, one density): the atoms' Z and W are interleaved in LDS
// (float4 = Z0.re, Z0.im, Z1.re, Z1.im), so one 16-byte read fetches a factor for both, and
// each record (byte offsets of 16-byte entries, FP32 weight) is loaded and decoded once for
// both atoms.  Same chunk walk, carry, sole stores and rotation as k_adjoint; an odd last
// atom is paired with itself and its second copy is not written.
template <bool ACC>
__global__ void __launch_bounds__(kAdjointThreads) k_adjoint_pair(DevParams<float> P, int atom_base,
                                                     const uint4* __restrict__ crec16,
                                                     const cplx<float>* __restrict__ Zin,
                                                     cplx<float>* __restrict__ Wout, char* __restrict__ Pt,
                                                     const double* __restrict__ ehc,
                                                     const double* __restrict__ rho1, double* __restrict__ eout,
                                                     double* __restrict__ dEout, double* __restrict__ edfout,
                                                     double* __restrict__ rho_out, int natom) {
    using T = float;
    extern __shared__ __align__(16) unsigned char smem[];
    float4* sZ = (float4*)smem;                  // [nzx] (Z' of k_adjoint), sZ[nz] = (1, 0, 1, 0)
    float4* sW = sZ + P.nzx;                     // [nzx]
    double* red = (double*)(sW + P.nzx);         // [2 * waves]: one sum per atom and wave
    double* s_scale = red + 2 * (kAdjointThreads / 64);  // dynamic LDS, as in k_adjoint
    const int a0 = 2 * (int)blockIdx.x, tid = threadIdx.x;
    const bool two = a0 + 1 < natom;
    const cplx<T>* Z0 = Zin + (size_t)a0 * P.nz;
    const cplx<T>* Z1 = two ? Z0 + P.nz : Z0;
    {
        // Z' of both atoms, with all of a thread's loads issued before the stores
        constexpr int U = 8;  // covers nzx <= 8 * kAdjointThreads = 4096 in one pass
        cplx<T> a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int x = tid + u * kAdjointThreads;
            a[u] = x < P.nz ? Z0[x] : cplx<T>{0, 0};
            b[u] = x < P.nz ? Z1[x] : cplx<T>{0, 0};
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int x = tid + u * kAdjointThreads;
            if (x < P.nz) sZ[x] = make_float4(a[u].re, a[u].im, b[u].re, b[u].im);
            if (x < P.nzx) sW[x] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    // the tail's inputs from global memory, loaded now (as in k_adjoint); the two atoms'
    // energies are evaluated by threads 0 and 64 (two waves) at the same time
    constexpr int kFold = 8;
    uint32_t fw[kFold];
#pragma unroll
    for (int u = 0; u < kFold; ++u) {
        const int x = tid + u * kAdjointThreads;
        fw[u] = x < P.nz ? P.fold[x] : 0u;
    }
    const int2 cb = tid < P.ncomb ? P.comb[tid] : make_int2(0, 0);
    const int2 cb2 = tid < P.ncomb2 ? P.comb2[tid] : make_int2(0, 0);
    const int eatom = tid == 0 ? a0 : tid == 64 && two ? a0 + 1 : -1;  // the atom whose energy this thread evaluates
    double e_hc = 0.0, r1v = 0.0;
    if (eatom >= 0) {
        e_hc = ehc[eatom];
        r1v = rho1[eatom];
    }
    if (tid == 0) sZ[P.nz] = make_float4(1.f, 0.f, 1.f, 0.f);
    __syncthreads();
    if (P.ny > 0) {
        for (int i = tid; i < P.ny; i += blockDim.x) {
            const int2 pr = P.ypair[i];
            const float4 a = sZ[pr.x], b = sZ[pr.y];
            const cplx<T> y0 = cmul(cplx<T>{a.x, a.y}, cplx<T>{b.x, b.y});
            const cplx<T> y1 = cmul(cplx<T>{a.z, a.w}, cplx<T>{b.z, b.w});
            sZ[P.nz + 1 + i] = make_float4(y0.re, y0.im, y1.re, y1.im);
        }
        __syncthreads();
    }
    // the walk, runs, slots and rotation of k_adjoint; a slot holds both atoms' sums
    float4* part = (float4*)(Pt + (size_t)blockIdx.x * partials_bytes(P.nslot, 1));
    const int nthr = (int)blockDim.x;
    const int nj = (P.nchunks + nthr - 1) / nthr;
    const int rot = nj > 0 ? (int)((unsigned)((atom_base >> 1) + (int)blockIdx.x) % (unsigned)nj) : 0;
    cplx<T> acc0 = {0, 0}, acc1 = {0, 0}, head0 = {0, 0}, head1 = {0, 0};
    cplx<double> acd0 = {0, 0}, acd1 = {0, 0};  // ACC: the run's FP64 rounds (k_adjoint's ACC)
    double rho0 = 0.0, rho1s = 0.0;             // ACC: summed at the flushes
    int j = rot;
    int2 ch = tid + j * nthr < P.nchunks ? P.chunks[tid + j * nthr] : make_int2(0, 0);
    bool split = false;
    if (rot > 0 && tid + rot * nthr < P.nchunks) split = ((P.chunks[tid + (rot - 1) * nthr].y ^ ch.y) & 0xfff) == 0;
    bool hold = split;
    for (int j0 = 0; j0 < nj; ++j0) {
        const int c = tid + j * nthr;
        const int jn = j + 1 < nj ? j + 1 : 0;
        const int cn = tid + jn * nthr;
        const bool has_next = j0 + 1 < nj && cn < P.nchunks;
        const int2 chn = has_next ? P.chunks[cn] : make_int2(0, 0);
        if (c < P.nchunks) {
            const int x = ch.y & 0xfff, cnt = (ch.y >> 12) & 63;
            const bool sole = (ch.y >> 18) & 1;
            const uint4* __restrict__ rec = crec16 + ch.x;
            if (ACC && __any(x >= P.oA0 && x < P.nz)) {
                for (int k = 0; k < cnt; k += 4) {
                    uint4 r[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) r[u] = rec[(size_t)(k + u) * 64];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const double w = (double)__uint_as_float(r[u].z) + (double)__uint_as_float(r[u].w);
                        const float4 za = lds_f4(r[u].x & 0xffff);
                        const float4 zbv = lds_f4(r[u].x >> 16);
                        const float4 zc = lds_f4(r[u].y);
                        const cplx<T> p0 = cmul(cmul(cplx<T>{za.x, za.y}, cplx<T>{zbv.x, zbv.y}), cplx<T>{zc.x, zc.y});
                        const cplx<T> p1 = cmul(cmul(cplx<T>{za.z, za.w}, cplx<T>{zbv.z, zbv.w}), cplx<T>{zc.z, zc.w});
                        acd0.re = fma(w, (double)p0.re, acd0.re);
                        acd0.im = fma(w, (double)p0.im, acd0.im);
                        acd1.re = fma(w, (double)p1.re, acd1.re);
                        acd1.im = fma(w, (double)p1.im, acd1.im);
                    }
                }
            } else
            for (int k = 0; k < cnt; k += 4) {
                uint4 r[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) r[u] = rec[(size_t)(k + u) * 64];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    // without the recipe the table's offsets need no unpacking (crec16u)
                    const T w = __uint_as_float(ACC ? r[u].z : r[u].w);
                    const float4 za = lds_f4(ACC ? r[u].x & 0xffff : r[u].x);
                    const float4 zbv = lds_f4(ACC ? r[u].x >> 16 : r[u].y);
                    const float4 zc = lds_f4(ACC ? r[u].y : r[u].z);
                    const cplx<T> p0 = cmul(cmul(cplx<T>{za.x, za.y}, cplx<T>{zbv.x, zbv.y}), cplx<T>{zc.x, zc.y});
                    const cplx<T> p1 = cmul(cmul(cplx<T>{za.z, za.w}, cplx<T>{zbv.z, zbv.w}), cplx<T>{zc.z, zc.w});
                    acc0.re = fma(w, p0.re, acc0.re);
                    acc0.im = fma(w, p0.im, acc0.im);
                    acc1.re = fma(w, p1.re, acc1.re);
                    acc1.im = fma(w, p1.im, acc1.im);
                }
            }
            if (!(jn == j + 1 && has_next && ((chn.y ^ ch.y) & 0xfff) == 0)) {
                if constexpr (ACC) {
                    const cplx<double> t0 = {acd0.re + (double)acc0.re, acd0.im + (double)acc0.im};
                    const cplx<double> t1 = {acd1.re + (double)acc1.re, acd1.im + (double)acc1.im};
                    if (x >= P.oA0 && x < P.nz) {
                        const float4 z = sZ[x];
                        rho0 += (double)z.x * t0.re - (double)z.y * t0.im;
                        rho1s += (double)z.z * t1.re - (double)z.w * t1.im;
                    }
                    acc0 = {(T)t0.re, (T)t0.im};
                    acc1 = {(T)t1.re, (T)t1.im};
                    acd0 = {0, 0};
                    acd1 = {0, 0};
                }
                if (hold) {
                    head0 = acc0;
                    head1 = acc1;
                    hold = false;
                } else {
                    if (split && j0 + 1 == nj) {
                        acc0.re += head0.re;
                        acc0.im += head0.im;
                        acc1.re += head1.re;
                        acc1.im += head1.im;
                    }
                    const float4 v = make_float4(acc0.re, acc0.im, acc1.re, acc1.im);
                    if (sole)
                        sW[x] = v;
                    else
                        part[(unsigned)ch.y >> 19] = v;
                }
                acc0 = {0, 0};
                acc1 = {0, 0};
            }
        }
        ch = chn;
        j = jn;
    }
    __syncthreads();
    combine_partials(P, part, sW, cb, cb2);
    __syncthreads();
    if (P.nyt > 0) {  // products Y, as in k_adjoint
        for (int k = tid; k < P.nyt; k += nthr) {
            const int2 e = P.yft[k];
            const int x = e.y & 0xfff, n = e.y >> 12;
            float4 w = sW[x];
            for (int u = 0; u < n; ++u) {
                const int2 t = P.yfe[e.x + u];
                const float4 g = sW[t.x], z = sZ[t.y];
                cplx<T> w0 = {w.x, w.y}, w1 = {w.z, w.w};
                cfma(w0, cplx<T>{g.x, g.y}, cplx<T>{z.x, z.y});
                cfma(w1, cplx<T>{g.z, g.w}, cplx<T>{z.z, z.w});
                w = make_float4(w0.re, w0.im, w1.re, w1.im);
            }
            sW[x] = w;
        }
        __syncthreads();
    }
    if constexpr (!ACC)
        for (int x = P.oA0 + tid; x < P.nz; x += nthr) {
            const float4 z = sZ[x], w = sW[x];
            rho0 += (double)re_mul(cplx<T>{z.x, z.y}, cplx<T>{w.x, w.y});
            rho1s += (double)re_mul(cplx<T>{z.z, z.w}, cplx<T>{w.z, w.w});
        }
    const double tot0 = block_sum<T>(rho0, red);
    __syncthreads();
    double tot1 = block_sum<T>(rho1s, red + kAdjointThreads / 64);  // valid on thread 0
    if (tid == 64) {  // the second atom's sum, from the same partials in the same order
        tot1 = 0.0;
        for (int i = 0; i < kAdjointThreads / 64; ++i) tot1 += red[kAdjointThreads / 64 + i];
    }
    if (eatom >= 0) {
        double sc[1];
        const double t[1] = {eatom == a0 ? tot0 : tot1};
        atom_energy<T>(P, 1, eatom, t, e_hc, &r1v, eout, dEout, edfout, rho_out, sc);
        s_scale[eatom - a0] = sc[0];
    }
    __syncthreads();
    // W_eff on representatives of conjugate pairs, as in k_adjoint
    const T sc0 = (T)s_scale[0], sc1 = two ? (T)s_scale[1] : (T)0;
    cplx<T>* W0 = Wout + (size_t)a0 * P.nz;
    cplx<T>* W1 = W0 + P.nz;
#pragma unroll
    for (int u = 0; u < kFold; ++u) {
        const int x = tid + u * kAdjointThreads;
        if (x >= P.nz) break;
        const uint32_t f = fw[u];
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        if (f & (1u << 17)) {
            w = sW[x];
            if (!(f & (1u << 18))) {
                const float4 b = sW[f & 0xffff];
                const T sg = (f & (1u << 16)) ? (T)-1 : (T)1;
                w.x += sg * b.x;
                w.y -= sg * b.y;
                w.z += sg * b.z;
                w.w -= sg * b.w;
            }
        }
        W0[x] = cplx<T>{sc0 * w.x, sc0 * w.y};
        if (two) W1[x] = cplx<T>{sc1 * w.z, sc1 * w.w};
    }
}

// ============================================================================ k_forces
// Legendre rows streamed one l at a time (src/spherical.F90:55-155): row l of P_lm and
// dP_lm/dz for m = 0..l from rows l-1 and l-2.  m is unrolled, l is a runtime loop, so the
// rows stay in registers without indexing them dynamically.
template <typename T, int L>
struct LegRows {
    T P[L + 1], dP[L + 1], P1[L + 1], dP1[L + 1], P2[L + 1], dP2[L + 1];
    __device__ __forceinline__ void step(int l, T z) {
#pragma unroll
        for (int m = 0; m <= L; ++m) {
            P2[m] = P1[m]; dP2[m] = dP1[m];
            P1[m] = P[m]; dP1[m] = dP[m];
        }
        if (l == 0) {
#pragma unroll
            for (int m = 0; m <= L; ++m) { P[m] = (T)0; dP[m] = (T)0; }
            P[0] = (T)1;
            return;
        }
        if (l == 1) {
            P[0] = (T)1.73205080756887719318 * z;
            dP[0] = (T)1.73205080756887719318;
            if (L >= 1) { P[L >= 1 ? 1 : 0] = -(T)1.22474487139158894067; dP[L >= 1 ? 1 : 0] = (T)0; }
            return;
        }
        const int base = l * (l + 1) / 2;
#pragma unroll
        for (int m = 0; m <= L; ++m) {
            if (m <= l - 2) {
                const T a = LegC<T>::a(base + m), b = LegC<T>::b(base + m);
                const T pm = a * fma(z, P1[m], b * P2[m]);
                dP[m] = a * (P1[m] + fma(z, dP1[m], b * dP2[m]));
                P[m] = pm;
            } else if (m == l - 1) {
                const T t = LegC<T>::d(l) * P1[m];
                P[m] = t * z;
                dP[m] = t;
            } else if (m == l) {
                P[m] = LegC<T>::c(l) * P1[m > 0 ? m - 1 : 0];
                dP[m] = (T)0;
            }
        }
    }
};

// one entry P_l^m, dP_l^m/dz of the rows above (l <= L): P_m^m = prod_{k<=m} c(k), then
// P_{m+1}^m = d(m+1) z P_m^m and the three-term recurrence up to l.  The constants are loaded
// first, to the compile-time bound (factors past m are 1 and steps past l are skipped), so
// that one trip to memory serves them all instead of one per factor and step.
template <typename T, int L>
__device__ __forceinline__ void legendre_one(int l, int m, T z, T& p, T& dp) {
    static_assert(L >= 1 && L < kLmaxC, "the constant tables cover l <= kLmaxC");
    T cf[L], ca[L], cb[L];
#pragma unroll
    for (int k = 1; k <= L; ++k) cf[k - 1] = LegC<T>::c(k);  // every index is in the table
    const T dm = LegC<T>::d(m + 1);
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const int ll = min(m + 2 + i, L), k = ll * (ll + 1) / 2 + m;
        ca[i] = LegC<T>::a(k);
        cb[i] = LegC<T>::b(k);
    }
    // every step is computed and kept or dropped by a select: with early exits the compiler
    // moved each load into its branch, one trip to memory per step again
    T pmm = (T)1;
#pragma unroll
    for (int k = 1; k <= L; ++k) pmm *= k <= m ? cf[k - 1] : (T)1;
    T p2 = pmm, dp2 = (T)0;
    T p1 = dm * pmm * z, dp1 = dm * pmm;
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const T a = ca[i], b = cb[i];
        const T pn = a * fma(z, p1, b * p2);
        const T dpn = a * (p1 + fma(z, dp1, b * dp2));
        const bool on = m + 2 + i <= l;
        p2 = on ? p1 : p2;
        dp2 = on ? dp1 : dp2;
        p1 = on ? pn : p1;
        dp1 = on ? dpn : dp1;
    }
    p = l == m ? pmm : p1;
    dp = l == m ? (T)0 : dp1;
}

template <typename T>
__device__ __forceinline__ void conj_all(cplx<T>& Y, cplx<T> (&g)[3], int am) {
    Y = cconj_sign(Y, am);
    g[0] = cconj_sign(g[0], am);
    g[1] = cconj_sign(g[1], am);
    g[2] = cconj_sign(g[2], am);
}

// lane = pair.  LA/LM: compile-time bounds of l3a/l2m; NM3X: magnetic columns of A;
// NBAX: atomic radial base bound; NRAX: nr3a bound; NR2X: nr2m bound.
// W is folded onto representatives of conjugate pairs (A: m' > 0, or m' = 0 and m >= 0;
// A2, A0: m' >= 0), so only those index pairs are visited.
// Q1: the magnetic columns of A are (l', m') = (0,0), (1,-1), (1,0), (1,1) (l3m = 1, one radial
// function), so the representative columns are known at compile time: 0, 2, 3 for m >= 0 and
// 3 for m < 0, and no other W entry is read.
template <typename T, int LA, int LM, int NM3X, int NBAX, int NRAX, int NR2X, bool Q1>
__global__ void __launch_bounds__(64) k_forces(DevParams<T> P, const int* __restrict__ off,
                                               const double* __restrict__ dlist,
                                               const double* __restrict__ elist,
                                               const double* __restrict__ dmomi,
                                               const double* __restrict__ emomi,
                                               const double* __restrict__ dmom,
                                               const double* __restrict__ emom,
                                               const cplx<T>* __restrict__ Win,
                                               const double* __restrict__ dEin,
                                               const double* __restrict__ edfin,
                                               double* __restrict__ flist, double* __restrict__ tlist,
                                               double* __restrict__ tlist0, int* err) {
    const int atom = blockIdx.x, lane = threadIdx.x;
    const int p0 = off[atom], nn = off[atom + 1] - p0;
    // W of this atom staged in LDS with coalesced loads: every lane reads the same entries
    // (broadcast), which as scalar loads would thrash the scalar cache with 20-40 KB per atom
    extern __shared__ __align__(16) unsigned char smem_f[];
    cplx<T>* sW = reinterpret_cast<cplx<T>*>(smem_f);
    {
        // forty loads in flight per lane: one trip to memory serves 2560 entries (all of W for
        // lmax 6), where a load followed by its store had waited for every 64
        const cplx<T>* __restrict__ Wg = Win + (size_t)atom * P.nz;
        constexpr int U = 40;
        for (int i0 = 0; i0 < P.nz; i0 += 64 * U) {
            cplx<T> v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + 64 * u + lane;
                v[u] = i < P.nz ? Wg[i] : cplx<T>{0, 0};
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int i = i0 + 64 * u + lane;
                if (i < P.nz) sW[i] = v[u];
            }
        }
        __syncthreads();
    }
    const cplx<T>* W = sW;
    const int nfa = P.nr3a * (P.l3a + 1);
    const double edf = edfin[atom];
    // magnetic column q of A -> m' (for the representative test): column q takes part with
    // m >= 0 (use0) when m' >= 0, with m < 0 (use1, the conjugated half) when m' > 0
    bool use0[NM3X], use1[NM3X];
#pragma unroll
    for (int q = 0; q < NM3X; ++q) {
        const int lq = q % P.lsm3;
        int lp = 0;
        while ((lp + 1) * (lp + 1) <= lq) ++lp;
        const int mp = lq - lp * lp - lp;
        use0[q] = q < P.nm3 && mp >= 0;
        use1[q] = q < P.nm3 && mp > 0;
    }
    static_assert(!Q1 || NM3X == 4, "Q1 needs four magnetic columns");
    auto used = [&](int q, int sg) -> bool {
        if constexpr (Q1) return sg ? q == 3 : q != 1;
        return sg ? use1[q] : use0[q];
    };

    for (int c0 = 0; c0 < nn; c0 += 64) {
        const int j = c0 + lane;
        const bool active = j < nn;
        const int p = p0 + (active ? j : 0);
        T Fx = 0, Fy = 0, Fz = 0, Tx = 0, Ty = 0, Tz = 0;
        T wa = 0;
        // the pair's inputs, all loads issued before the first is used
        const double r_d = dlist[p], mu_d = dmom[p];
        const double ex_d = elist[3 * p], ey_d = elist[3 * p + 1], ez_d = elist[3 * p + 2];
        const double sx_d = emom[3 * p], sy_d = emom[3 * p + 1], sz_d = emom[3 * p + 2];
        const T* rowa = active ? atomic_row(P, r_d, wa, err) : nullptr;
        if (rowa) {
            const T invr = (T)(1.0 / r_d);
            const T ex = (T)ex_d, ey = (T)ey_d, ez = (T)ez_d;
            const T sx = (T)sx_d, sy = (T)sy_d, sz = (T)sz_d;
            T wm = 0;
            bool unit = false;
            const T* rowm = magnetic_row(P, mu_d, wm, unit, err);
            const T rsa = (T)P.rs_a, rsm = (T)P.rs_m;
            // the atomic loop's first spline coefficients (row l = 0); each later row is loaded
            // one l ahead in the loop.  Q1 loads row 0 here, so that it comes in the same trip
            // to memory as the M values below (the wide classes load it at the loop, where
            // their registers allow)
            T cq0[Q1 ? NRAX : 1][4];
            if constexpr (Q1)
#pragma unroll
                for (int n = 0; n < NRAX; ++n) load_coef(n < P.nr3a ? rowa + 4 * (P.nb_a + n) : P.lut_a, cq0[n]);
            // H_q = M_{l',n'}(mu_j) Y_{l'm'}(s_j), m' >= 0, for the magnetic columns of A
            cplx<T> H[NM3X], U[NM3X];
#pragma unroll
            for (int q = 0; q < NM3X; ++q) {
                H[q] = {0, 0};
                U[q] = {0, 0};
            }
            {
                // M_{l',n'}(mu_j) of the columns' (l', n'), entry n' (l3m + 1) + l' (at most nm3
                // of them), with their coefficient loads issued together
                T Mh[NM3X];
                {
                    const int nl = P.l3m + 1, nh = nl * P.nr3m;
                    T c[NM3X][4];
#pragma unroll
                    for (int i = 0; i < NM3X; ++i)
                        load_coef(rowm && i < nh ? rowm + 4 * (P.nb_m + (i % nl) * P.nr2m + i / nl) : P.lut_m, c[i]);
#pragma unroll
                    for (int i = 0; i < NM3X; ++i)
                        Mh[i] = rowm ? fma(wm, fma(wm, fma(wm, c[i][3], c[i][2]), c[i][1]), c[i][0]) : unit ? (T)1 : (T)0;
                }
                cplx<T> ph = {1, 0};
                LegRows<T, LM> Lr;
                for (int lp = 0; lp <= P.l3m; ++lp) {
                    Lr.step(lp, sz);
                    ph = {1, 0};
#pragma unroll
                    for (int am = 0; am <= LM; ++am) {
                        if (am > lp) break;
                        const cplx<T> y = {Lr.P[am] * ph.re, Lr.P[am] * ph.im};
                        for (int np = 0; np < P.nr3m; ++np) {
                            T v = 0;
                            const int ih = np * (P.l3m + 1) + lp;
#pragma unroll
                            for (int i = 0; i < NM3X; ++i)
                                if (i == ih) v = Mh[i];
                            const int q = np * P.lsm3 + lp * lp + lp + am;
#pragma unroll
                            for (int qq = 0; qq < NM3X; ++qq)
                                if (qq == q) H[qq] = {v * y.re, v * y.im};
                        }
                        ph = cmul(ph, cplx<T>{sx, sy});
                    }
                }
            }
            // ---- atomic loop: F from A and U_q = sum W_A R Y (src/forces.F90:497-620, 836-957)
            {
                cplx<T> pha[LA + 1];
                pha[0] = {1, 0};
#pragma unroll
                for (int m = 1; m <= LA; ++m) pha[m] = cmul(pha[m - 1], cplx<T>{ex, ey});
                LegRows<T, LA> Lr;
                // spline coefficients of row l are loaded one iteration ahead, so the gathers
                // from the table overlap the arithmetic of the previous l
                T cq[NRAX][4];
#pragma unroll
                for (int n = 0; n < NRAX; ++n) {
                    if constexpr (Q1) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) cq[n][k] = cq0[Q1 ? n : 0][k];
                    } else {
                        load_coef(n < P.nr3a ? rowa + 4 * (P.nb_a + n) : P.lut_a, cq[n]);
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) cq[n][k] = n < P.nr3a ? cq[n][k] : (T)0;
                }
                for (int l = 0; l <= P.l3a; ++l) {
                    Lr.step(l, ez);
                    T R[NRAX], dR[NRAX];
#pragma unroll
                    for (int n = 0; n < NRAX; ++n) {
                        R[n] = fma(wa, fma(wa, fma(wa, cq[n][3], cq[n][2]), cq[n][1]), cq[n][0]);
                        dR[n] = fma(wa, fma(wa, (T)3 * cq[n][3], (T)2 * cq[n][2]), cq[n][1]) * rsa;
                    }
                    if (l < P.l3a) {
                        const T* nrow = rowa + 4 * (P.nb_a + (l + 1) * P.nr3a);
#pragma unroll
                        for (int n = 0; n < NRAX; ++n) {
                            load_coef(n < P.nr3a ? nrow + 4 * n : P.lut_a, cq[n]);
#pragma unroll
                            for (int k = 0; k < 4; ++k) cq[n][k] = n < P.nr3a ? cq[n][k] : (T)0;
                        }
                    }
#pragma unroll
                    for (int am = 0; am <= LA; ++am) {
                        if (am > l) break;
                        cplx<T> Y0, g0[3];
                        ylm_grad<T>(am, Lr.P[am], Lr.dP[am], pha[am], pha[am > 0 ? am - 1 : 0], ex, ey, ez, Y0, g0);
#pragma unroll
                        for (int sg = 0; sg < 2; ++sg) {
                            if (sg == 1 && am == 0) break;
                            cplx<T> Y = Y0, g[3] = {g0[0], g0[1], g0[2]};
                            if (sg) conj_all(Y, g, am);
                            const int lm = l * l + l + (sg ? -am : am);
                            // V_n = sum_q W[n][lm][q] H_q.  The force needs
                            // sum_n Re[V_n (R'_n Y e + R_n/r grad Y)], the torque
                            // sum_n W[n][lm][q] R_n Y: both are summed over n first
                            // (S1 = sum_n R'_n V_n, S2 = sum_n R_n V_n, Tq = sum_n R_n W[n][lm][q])
                            // and meet Y and grad Y once per (l, m).  With the columns in
                            // registers (NM3X <= 8) the sums over n run per column instead,
                            // Tq and Dq = sum_n R'_n W[n][lm][q], and S1 = sum_q Dq H_q,
                            // S2 = sum_q Tq H_q follow after the loop: four operations per
                            // column and n in place of six, and none per n for S1, S2
                            constexpr int NTQ = NM3X <= 8 ? NM3X : 1;
                            cplx<T> S1 = {0, 0}, S2 = {0, 0}, Tq[NTQ], Dq[NTQ];
#pragma unroll
                            for (int q = 0; q < NTQ; ++q) Tq[q] = Dq[q] = {0, 0};
                            if constexpr (Q1 && sizeof(T) == 4) {
                                // FP32 rows are 32 bytes, 16-byte aligned: columns 2, 3 in one
                                // ds_read_b128 and column 0 (m >= 0 only) in another.  Each row's
                                // reads are issued kAhead rows before its sums, so that their
                                // latency passes during the sums of the rows before it (read and
                                // used at once, every row had waited for its own reads).  The
                                // loop runs to the compile-time bound, as below.
                                constexpr int kAhead = 2, kRing = kAhead + 1;
                                float4 r23[kRing], r01[kRing];
#pragma unroll
                                for (int n = 0; n < NRAX + kAhead; ++n) {
                                    if (n < NRAX) {
                                        const int nw = n >= P.nr3a ? 0 : n;
                                        const float4* Wn = (const float4*)(W + (nw * P.lsa + lm) * P.nm3);
                                        r23[n % kRing] = Wn[1];
                                        if (sg == 0) r01[n % kRing] = Wn[0];
                                    }
                                    const int k = n - kAhead;  // the row summed now
                                    if (k < 0) continue;
                                    cplx<T> wv[NM3X];
                                    wv[2] = {r23[k % kRing].x, r23[k % kRing].y};
                                    wv[3] = {r23[k % kRing].z, r23[k % kRing].w};
                                    if (sg == 0) wv[0] = {r01[k % kRing].x, r01[k % kRing].y};
#pragma unroll
                                    for (int q = 0; q < NM3X; ++q) {
                                        if (used(q, sg)) {
                                            Tq[q].re = fma(R[k], wv[q].re, Tq[q].re);
                                            Tq[q].im = fma(R[k], wv[q].im, Tq[q].im);
                                            Dq[q].re = fma(dR[k], wv[q].re, Dq[q].re);
                                            Dq[q].im = fma(dR[k], wv[q].im, Dq[q].im);
                                        }
                                    }
                                }
                            } else
#pragma unroll
                            for (int n = 0; n < NRAX; ++n) {
                                // Q1 runs to the compile-time bound (past nr3a, R = R' = 0 and
                                // row 0 is read), so that each (l, m) body is one straight block
                                // the scheduler can interleave (lmax-6 FP32 0.963 -> 0.919 ms);
                                // the general instance keeps the exit, as without it it spills
                                if constexpr (!Q1) {
                                    if (n >= P.nr3a) break;
                                }
                                const int nw = Q1 && n >= P.nr3a ? 0 : n;
                                const cplx<T>* Wn = W + (nw * P.lsa + lm) * P.nm3;
                                // the row's reads are issued together, outside the uniform
                                // branches, so their LDS latency overlaps (a read inside each
                                // branch waited for its own result)
                                // (wide classes, 27 columns, read inside the branches: the
                                // hoisted values would not fit in registers)
                                constexpr bool kHoist = NM3X <= 8;
                                cplx<T> wv[kHoist ? NM3X : 1];
                                if constexpr (kHoist) {
#pragma unroll
                                    for (int q = 0; q < NM3X; ++q)
                                        if (!Q1 || used(q, sg)) wv[q] = Wn[q < P.nm3 ? q : 0];
                                }
                                if constexpr (kHoist) {
#pragma unroll
                                    for (int q = 0; q < NM3X; ++q) {
                                        if (used(q, sg)) {
                                            const cplx<T> wq = wv[kHoist ? q : 0];
                                            const int qt = kHoist ? q : 0;
                                            Tq[qt].re = fma(R[n], wq.re, Tq[qt].re);
                                            Tq[qt].im = fma(R[n], wq.im, Tq[qt].im);
                                            Dq[qt].re = fma(dR[n], wq.re, Dq[qt].re);
                                            Dq[qt].im = fma(dR[n], wq.im, Dq[qt].im);
                                        }
                                    }
                                } else {
                                    // wide classes: per-column sums would not fit
                                    cplx<T> V = {0, 0};
                                    const cplx<T> ry = {R[n] * Y.re, R[n] * Y.im};
#pragma unroll
                                    for (int q = 0; q < NM3X; ++q) {
                                        if (used(q, sg)) {
                                            const cplx<T> wq = Wn[q];
                                            cfma(V, wq, H[q]);
                                            cfma(U[q], wq, ry);
                                        }
                                    }
                                    S1.re = fma(dR[n], V.re, S1.re);
                                    S1.im = fma(dR[n], V.im, S1.im);
                                    S2.re = fma(R[n], V.re, S2.re);
                                    S2.im = fma(R[n], V.im, S2.im);
                                }
                            }
                            if constexpr (NM3X <= 8) {
#pragma unroll
                                for (int q = 0; q < NTQ; ++q)
                                    if (used(q, sg)) {
                                        cfma(S1, Dq[q], H[q]);
                                        cfma(S2, Tq[q], H[q]);
                                        cfma(U[q], Tq[q], Y);
                                    }
                            }
                            // Re[S1 Y] e + Re[S2 grad Y] / r
                            const T a = re_mul(S1, Y);
                            Fx += fma(a, ex, invr * re_mul(S2, g[0]));
                            Fy += fma(a, ey, invr * re_mul(S2, g[1]));
                            Fz += fma(a, ez, invr * re_mul(S2, g[2]));
                        }
                    }
                }
            }
            // ---- magnetic loop over m' >= 0: A2 part of F and T, A part of T
            // (src/forces.F90:460-491, 793-957)
            {
                const T invmu = (T)(1.0 / mu_d);
                T ga[NBAX], dga[NBAX];
                T cc[4];  // the core term's coefficients, loaded with the g_n (used after this loop)
                {
                    T c[NBAX][4];  // one 16-byte load per function, issued together
#pragma unroll
                    for (int n = 0; n < NBAX; ++n) load_coef(n < P.nb_a ? rowa + 4 * n : P.lut_a, c[n]);
                    load_coef(rowa + 4 * (P.nb_a + nfa), cc);
#pragma unroll
                    for (int n = 0; n < NBAX; ++n) {
                        ga[n] = n < P.nb_a ? fma(wa, fma(wa, fma(wa, c[n][3], c[n][2]), c[n][1]), c[n][0]) : (T)0;
                        dga[n] = n < P.nb_a ? fma(wa, fma(wa, (T)3 * c[n][3], (T)2 * c[n][2]), c[n][1]) * rsa : (T)0;
                    }
                }
                T sf = 0;  // sum_n g'_n Re(W2 H2)
                cplx<T> phm[LM + 1];
                phm[0] = {1, 0};
#pragma unroll
                for (int m = 1; m <= LM; ++m) phm[m] = cmul(phm[m - 1], cplx<T>{sx, sy});
                LegRows<T, LM> Lr;
                // the coefficients of row l' are loaded one iteration ahead, as in the atomic loop
                T cm[NR2X][4];
#pragma unroll
                for (int np = 0; np < NR2X; ++np)
                    load_coef(rowm && np < P.nr2m ? rowm + 4 * (P.nb_m + np) : P.lut_m, cm[np]);
                for (int lp = 0; lp <= P.l2m; ++lp) {
                    Lr.step(lp, sz);
                    T Mv[NR2X], dMv[NR2X];
#pragma unroll
                    for (int np = 0; np < NR2X; ++np) {
                        const bool f = rowm && np < P.nr2m;
                        Mv[np] = f ? fma(wm, fma(wm, fma(wm, cm[np][3], cm[np][2]), cm[np][1]), cm[np][0]) : unit ? (T)1 : (T)0;
                        dMv[np] = f ? fma(wm, fma(wm, (T)3 * cm[np][3], (T)2 * cm[np][2]), cm[np][1]) * rsm : (T)0;
                    }
                    if (lp < P.l2m) {
#pragma unroll
                        for (int np = 0; np < NR2X; ++np)
                            load_coef(rowm && np < P.nr2m ? rowm + 4 * (P.nb_m + (lp + 1) * P.nr2m + np) : P.lut_m, cm[np]);
                    }
#pragma unroll
                    for (int am = 0; am <= LM; ++am) {
                        if (am > lp) break;
                        cplx<T> Y, g[3];
                        ylm_grad<T>(am, Lr.P[am], Lr.dP[am], phm[am], phm[am > 0 ? am - 1 : 0], sx, sy, sz, Y, g);
                        const int lmp = lp * lp + lp + am;
                        // Q1 in FP32: the W2 entries of every n' of this (l', m') are read
                        // before the first sum, so that one LDS round trip serves all n' (read
                        // per n', each n' had waited for its own reads)
                        constexpr bool kW2B = Q1 && sizeof(T) == 4;
                        cplx<T> w2b[kW2B ? NR2X : 1][kW2B ? NBAX : 1];
                        if constexpr (kW2B) {
#pragma unroll
                            for (int np = 0; np < NR2X; ++np) {
                                const cplx<T>* W2r = W + P.oA2 + (np < P.nr2m ? np : 0) * P.lsm + lmp;
#pragma unroll
                                for (int n = 0; n < NBAX; ++n)
                                    w2b[np][n] = W2r[(size_t)(n < P.nb_a ? n : 0) * P.nr2m * P.lsm];
                            }
                        }
#pragma unroll
                        for (int np = 0; np < NR2X; ++np) {
                            if (np >= P.nr2m) break;
                            const cplx<T> H2 = {Mv[np] * Y.re, Mv[np] * Y.im};
                            // U2 = sum_n g_n W2[n], D2 = sum_n g'_n W2[n]; the force's
                            // sum_n g'_n Re(W2[n] H2) is Re(D2 H2)
                            cplx<T> U2 = {0, 0}, D2 = {0, 0};
                            const cplx<T>* W2 = W + P.oA2 + np * P.lsm + lmp;
#pragma unroll
                            for (int n = 0; n < NBAX; ++n) {
                                // no early exit: entries n >= nb_a have ga = dga = 0, and a
                                // loop bounded at run time was left rolled, its register
                                // arrays indexed indirectly (FP32 forces 1.46 -> 1.30 ms)
                                const cplx<T> wv = kW2B ? w2b[kW2B ? np : 0][kW2B ? n : 0]
                                                        : W2[(size_t)(n < P.nb_a ? n : 0) * P.nr2m * P.lsm];
                                D2.re = fma(wv.re, dga[n], D2.re);
                                D2.im = fma(wv.im, dga[n], D2.im);
                                U2.re = fma(wv.re, ga[n], U2.re);
                                U2.im = fma(wv.im, ga[n], U2.im);
                            }
                            sf += re_mul(D2, H2);
                            if (lp <= P.l3m && np < P.nr3m) {
                                const int q = np * P.lsm3 + lmp;
#pragma unroll
                                for (int qq = 0; qq < NM3X; ++qq)
                                    if (qq == q) {
                                        U2.re += U[qq].re;
                                        U2.im += U[qq].im;
                                    }
                            }
                            // d(M Y)/dm = M' s Y + M/mu grad Y
                            const T a = dMv[np] * re_mul(U2, Y), b = Mv[np] * invmu;
                            Tx += fma(a, sx, b * re_mul(U2, g[0]));
                            Ty += fma(a, sy, b * re_mul(U2, g[1]));
                            Tz += fma(a, sz, b * re_mul(U2, g[2]));
                        }
                    }
                }
                // F from A2 (g'_n e) and the core term (src/forces.F90:464-494)
                const T dcr = fma(wa, fma(wa, (T)3 * cc[3], (T)2 * cc[2]), cc[1]) * rsa;
                const T s = sf + (T)edf * dcr;
                Fx = fma(s, ex, Fx);
                Fy = fma(s, ey, Fy);
                Fz = fma(s, ez, Fz);
            }
        }
        // store transposed across the wave: lane L writes element L + 64 r of the 64 pairs'
        // 192 consecutive doubles of flist (and of tlist), so that every store instruction
        // covers whole lines, also when the outputs live in host memory and the stores cross
        // PCIe.  Element i belongs to pair i / 3, component i % 3.
        const int cnt = 3 * min(64, nn - c0);
        double* __restrict__ fo = flist + 3 * (size_t)(p0 + c0);
        double* __restrict__ to = tlist + 3 * (size_t)(p0 + c0);
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int i = lane + 64 * r, src = i / 3, comp = i - 3 * src;
            const T f0 = __shfl(Fx, src), f1 = __shfl(Fy, src), f2 = __shfl(Fz, src);
            const T t0 = __shfl(Tx, src), t1 = __shfl(Ty, src), t2 = __shfl(Tz, src);
            if (i < cnt) {
                fo[i] = (double)(comp == 0 ? f0 : comp == 1 ? f1 : f2);
                to[i] = (double)(comp == 0 ? t0 : comp == 1 ? t1 : t2);
            }
        }
    }
    // ---- self term T0 = dE_i/dm_i over m' >= 0 (src/forces.F90:712-749), spread over the
    // wave: lane k takes B1 function k (nb_m <= 32) and entry k of the (l', m' >= 0) half set
    // (at most 28 entries for l2m <= LM <= 6; each lane with its own Legendre recurrence), and
    // the wave sums the three components in a fixed butterfly order.  The global loads come in
    // two batches, the atom's moment, direction and dE first, then every coefficient a lane
    // needs (its B1 function and c1, and the M rows of its (l', m')): loaded where each was
    // used, they had taken seven trips to memory at the end of every workgroup.
    static_assert((LM + 1) * (LM + 2) / 2 <= 64, "one half-set entry per lane");
    {
        const double mu_d = dmomi[atom];
        const double sxd = emomi[3 * atom], syd = emomi[3 * atom + 1], szd = emomi[3 * atom + 2];
        const double dE0 = dEin[(size_t)atom * P.nd], dE1 = P.nd > 1 ? dEin[(size_t)atom * P.nd + 1] : 0.0;
        const bool b1 = lane < P.nb_m, hs = lane < P.lhm;
        const double c10 = P.c1[0][b1 ? lane : 0], c11 = P.c1[1][b1 ? lane : 0];
        const T invmu = (T)(1.0 / mu_d);
        const T sx = (T)sxd, sy = (T)syd, sz = (T)szd;
        T wm = 0;
        bool unit = false;
        const T* rowm = magnetic_row(P, mu_d, wm, unit, err);
        const T rsm = (T)P.rs_m;
        int lp = 0;
        while ((lp + 1) * (lp + 2) / 2 <= lane) ++lp;
        const int am = lane - lp * (lp + 1) / 2;
        T cb[4], cm[NR2X][4];
        load_coef(rowm && b1 ? rowm + 4 * lane : P.lut_m, cb);
#pragma unroll
        for (int np = 0; np < NR2X; ++np)
            load_coef(rowm && hs && np < P.nr2m ? rowm + 4 * (P.nb_m + lp * P.nr2m + np) : P.lut_m, cm[np]);
        T Tx = 0, Ty = 0, Tz = 0;
        if (b1) {  // B1: dE/dg^m_n = dE/drho c1_n
            T v = unit ? (T)1 : (T)0, dv = 0;
            if (rowm) horner(cb, wm, rsm, v, dv);
            const T s1 = (T)(dE0 * c10 + dE1 * c11) * dv;
            Tx = fma(s1, sx, Tx);
            Ty = fma(s1, sy, Ty);
            Tz = fma(s1, sz, Tz);
        }
        if (hs) {
            T pl, dpl;
            legendre_one<T, LM>(lp, am, sz, pl, dpl);
            cplx<T> ph = {1, 0}, ph1 = {1, 0};
            for (int k = 0; k < am; ++k) {
                ph1 = ph;
                ph = cmul(ph, cplx<T>{sx, sy});
            }
            cplx<T> Y, g[3];
            ylm_grad<T>(am, pl, dpl, ph, ph1, sx, sy, sz, Y, g);
#pragma unroll
            for (int np = 0; np < NR2X; ++np) {
                if (np >= P.nr2m) break;
                T v = unit ? (T)1 : (T)0, dv = 0;
                if (rowm) horner(cm[np], wm, rsm, v, dv);
                const cplx<T> w0 = W[P.oA0 + np * P.lsm + lp * lp + lp + am];
                const T a = dv * re_mul(w0, Y), b = v * invmu;
                Tx += fma(a, sx, b * re_mul(w0, g[0]));
                Ty += fma(a, sy, b * re_mul(w0, g[1]));
                Tz += fma(a, sz, b * re_mul(w0, g[2]));
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            Tx += __shfl_xor(Tx, o, 64);
            Ty += __shfl_xor(Ty, o, 64);
            Tz += __shfl_xor(Tz, o, 64);
        }
        if (lane == 0) {
            tlist0[3 * (size_t)atom] = (double)Tx;
            tlist0[3 * (size_t)atom + 1] = (double)Ty;
            tlist0[3 * (size_t)atom + 2] = (double)Tz;
        }
    }
}

// ============================================================================ fitting
// Gradient of L = sum_s gE_s E_s + sum_j gF_j . F_j with respect to the expansion
// coefficients (the contraction of src/forces.F90:93-141 run backwards): with
// E_i = fc(e_hc) F(rho_i) + e_hc and rho_i = sum_k c_k B_k,i,
//   dL/dc_k = sum_i [alpha_i B_k,i + beta_i D_u B_k,i],  u = gF (a displacement field),
//   alpha_i = gE_s(i) fc F' - fc F'' D_u rho_i - F' D_u fc,   beta_i = -fc F'
// (F_j = -dE/dx_j, so sum_j gF_j . dF_j/dc = -D_u dE/dc).  B_k = Re sum_t u_{k,t} P_t with
// P_t = prod Z[idx_t]; D_u P_t follows from Zdot = D_u Z, the forward-mode derivative of the
// density along u.  FP64 only.

// Y_lm (m >= 0) and D_u Y_lm = (grad Y . du) / r for a bond of unit vector (x, y, z)
template <int L>
__device__ __forceinline__ void ylm_half_jvp(double x, double y, double z, int lmax, double dux, double duy,
                                             double duz, double invr, cplx<double>* out, cplx<double>* dout) {
    double P[(L + 1) * (L + 2) / 2], dP[(L + 1) * (L + 2) / 2];
    legendre<double, L>(z, P, dP);
    cplx<double> ph[L + 1];
    ph[0] = {1, 0};
#pragma unroll
    for (int m = 1; m <= L; ++m) ph[m] = cmul(ph[m - 1], cplx<double>{x, y});
#pragma unroll
    for (int l = 0; l <= L; ++l)
#pragma unroll
        for (int m = 0; m <= l; ++m) {
            if (l > lmax) continue;
            const int k = l * (l + 1) / 2 + m;
            cplx<double> Y, g[3];
            ylm_grad<double>(m, P[k], dP[k], ph[m], ph[m > 0 ? m - 1 : 0], x, y, z, Y, g);
            out[k] = Y;
            dout[k] = {(g[0].re * dux + g[1].re * duy + g[2].re * duz) * invr,
                       (g[0].im * dux + g[1].im * duy + g[2].im * duz) * invr};
        }
}

// Zdot = D_u Z and D_u e_hc, the pair vector of pair p = (i -> j) moving by u_i - u_j.
// Same task layout as k_density; A0 does not depend on positions.
// The wide class (27 magnetic columns) runs at 2 waves per SIMD: 256 registers with a spill
// instead of 294 at 1 wave (fit step on the shape of the example potential 41.3 -> 37.3 ms).
template <int L, int CH, int NB, int NM3X, int NB2, int NR2X>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NM3X > 8 ? 2 : 1, 8)))
k_density_jvp(DevParams<double> P, const int* __restrict__ off,
              const int* __restrict__ nlist, const double* __restrict__ du,
              const double* __restrict__ dlist,
              const double* __restrict__ elist,
              const double* __restrict__ dmom,
              const double* __restrict__ emom,
              cplx<double>* __restrict__ Zdot,
              double* __restrict__ ehc_dot, int* err) {
    using T = double;
    extern __shared__ __align__(16) unsigned char smem[];
    const int atom = blockIdx.x;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int nfa = P.nr3a * (P.l3a + 1), nfm = P.nr2m * (P.l2m + 1);
    cplx<T>* sY = (cplx<T>*)smem;        // [CH][lha]
    cplx<T>* sYd = sY + CH * P.lha;      // [CH][lha]
    cplx<T>* sYm = sYd + CH * P.lha;     // [CH][lhm]
    cplx<T>* sH = sYm + CH * P.lhm;      // [CH][nm3]
    T* sR = (T*)(sH + CH * P.nm3);       // [CH][nfa]
    T* sRd = sR + CH * nfa;              // [CH][nfa]
    T* sGd = sRd + CH * nfa;             // [CH][nb_a]
    T* sM = sGd + CH * P.nb_a;           // [CH][nfm]
    T* sCrd = sM + CH * nfm;             // [CH]
    double* sE = sCrd + CH;              // [1]

    const int p0 = off[atom], nn = off[atom + 1] - p0;
    Zdot += (size_t)atom * P.nz;
    const double uix = du[3 * (size_t)atom], uiy = du[3 * (size_t)atom + 1], uiz = du[3 * (size_t)atom + 2];
    for (int t = tid; t < P.nr2m * P.lsm; t += nthr) Zdot[P.oA0 + t] = {0, 0};
    if (tid == 0) *sE = 0.0;

    const int nblkA = (P.nr3a + NB - 1) / NB, nblkA2 = (P.nb_a + NB2 - 1) / NB2;
    // tasks cover m >= 0 only; the m < 0 entries are conjugates (see k_density)
    const int nTA = P.lha * nblkA, nTA2 = P.lhm * nblkA2;
    // A2 tasks start at a wave boundary when that fits in one round (no wave runs both loops)
    const int a2base = ((nTA + 63) & ~63) + nTA2 <= nthr ? (nTA + 63) & ~63 : nTA;
    for (int round = 0; round * nthr < a2base + nTA2; ++round) {
        const int task = round * nthr + tid;
        cplx<T> acc[NB > NB2 ? NB : NB2][NM3X > NR2X ? NM3X : NR2X];
#pragma unroll
        for (int a = 0; a < (NB > NB2 ? NB : NB2); ++a)
#pragma unroll
            for (int b = 0; b < (NM3X > NR2X ? NM3X : NR2X); ++b) acc[a][b] = {0, 0};
        const bool isA = task < nTA, isA2 = task >= a2base && task < a2base + nTA2;
        int lh = 0, l = 0, n0 = 0;
        if (isA) { lh = task % P.lha; n0 = (task / P.lha) * NB; }
        if (isA2) { lh = (task - a2base) % P.lhm; n0 = ((task - a2base) / P.lhm) * NB2; }
        while ((l + 1) * (l + 2) / 2 <= lh) ++l;
        const int m = lh - l * (l + 1) / 2;  // >= 0
        const int lm = l * l + l + m;
        for (int c0 = 0; c0 < nn; c0 += CH) {
            const int nj = nn - c0 < CH ? nn - c0 : CH;
            __syncthreads();
            // staging of pair c0 + j spread over the block (as in k_density): role 0 the bond
            // Y_lm and D_u Y_lm and the core term, roles 1..nsplit shares of the radial
            // functions and their derivatives along u, role nsplit + 1 the magnetic side
            const int j = tid % CH, role = tid / CH, nsplit = nthr / CH - 2;
            const bool live = j < nj;
            const int p = p0 + c0 + (live ? j : 0);
            if (role <= nsplit) {
                if (live) {
                    const int jn = nlist[p];
                    const double dux = uix - du[3 * (size_t)jn], duy = uiy - du[3 * (size_t)jn + 1],
                                 duz = uiz - du[3 * (size_t)jn + 2];
                    const double ex = elist[3 * p], ey = elist[3 * p + 1], ez = elist[3 * p + 2];
                    const double rdot = ex * dux + ey * duy + ez * duz;
                    T w;
                    const T* row = atomic_row(P, dlist[p], w, err);
                    if (role == 0) {
                        T cr = 0, dcr = 0;
                        if (row) horner(row + 4 * (P.nb_a + nfa), w, (T)P.rs_a, cr, dcr);
                        sCrd[j] = dcr * rdot;
                        ylm_half_jvp<L>(ex, ey, ez, P.l3a, dux, duy, duz, 1.0 / dlist[p], sY + j * P.lha,
                                        sYd + j * P.lha);
                    } else {
                        const int nf = P.nb_a + nfa, per = (nf + nsplit - 1) / nsplit;
                        const int k0 = (role - 1) * per, k1 = min(nf, k0 + per);
                        spline_values_dv<T>(row, P.lut_a, w, (T)P.rs_a, k0, k1, 0, (T)0, [&](int k, T v, T dv) {
                            if (k < P.nb_a) {
                                sGd[j * P.nb_a + k] = dv * rdot;
                            } else {
                                sR[j * nfa + (k - P.nb_a)] = v;
                                sRd[j * nfa + (k - P.nb_a)] = dv * rdot;
                            }
                        });
                    }
                } else if (role == 0) {
                    sCrd[j] = 0;
                }
            } else if (role == nsplit + 1) {  // magnetic side (does not move with positions)
                if (live) {
                    T w = 0;
                    bool unit;
                    const T* row = magnetic_row(P, dmom[p], w, unit, err);
                    spline_values<T>(row, P.lut_m, w, 0, nfm, P.nb_m, unit ? (T)1 : (T)0,
                                     [&](int k, T v) { sM[j * nfm + k] = v; });
                    cplx<T>* yh = sYm + j * P.lhm;
                    ylm_half<T, L>(emom[3 * p], emom[3 * p + 1], emom[3 * p + 2], P.l2m, yh);
                    for (int q = 0; q < P.nm3; ++q) {
                        const int np = q / P.lsm3, lmp = q % P.lsm3;
                        int lp = 0;
                        while ((lp + 1) * (lp + 1) <= lmp) ++lp;
                        const int mp = lmp - lp * lp - lp;
                        const cplx<T> y = yfull(yh, lp, mp);
                        const T mv = sM[j * nfm + lp * P.nr2m + np];
                        sH[j * P.nm3 + q] = {mv * y.re, mv * y.im};
                    }
                }
            }
            __syncthreads();
            if (round == 0 && tid == 0) {
                double sum = 0.0;
                for (int j = 0; j < nj; ++j) sum += sCrd[j];
                *sE += sum;
            }
            if (isA) {
                // unused columns and radial indices are read without a test, as in k_density
                for (int j = 0; j < nj; ++j) {
                    const cplx<T> y = sY[j * P.lha + lh], yd = sYd[j * P.lha + lh];
                    cplx<T> h[NM3X];
#pragma unroll
                    for (int q = 0; q < NM3X; ++q) h[q] = sH[j * P.nm3 + q];
                    const T* rr = sR + j * nfa + l * P.nr3a + n0;
                    const T* rdr = sRd + j * nfa + l * P.nr3a + n0;
#pragma unroll
                    for (int b = 0; b < NB; ++b) {
                        const T r = rr[b], rd = rdr[b];
                        const cplx<T> ry = {rd * y.re + r * yd.re, rd * y.im + r * yd.im};
#pragma unroll
                        for (int q = 0; q < NM3X; ++q) cfma(acc[b][q], ry, h[q]);
                    }
                }
            } else if (isA2) {
                for (int j = 0; j < nj; ++j) {
                    const cplx<T> y = sYm[j * P.lhm + lh];
                    const T* mr = sM + j * nfm + l * P.nr2m;
                    cplx<T> my[NR2X];
#pragma unroll
                    for (int q = 0; q < NR2X; ++q) my[q] = {mr[q] * y.re, mr[q] * y.im};
                    const T* gr = sGd + j * P.nb_a + n0;
#pragma unroll
                    for (int b = 0; b < NB2; ++b) {
                        const T gd = gr[b];
#pragma unroll
                        for (int q = 0; q < NR2X; ++q) {
                            acc[b][q].re = fma(gd, my[q].re, acc[b][q].re);
                            acc[b][q].im = fma(gd, my[q].im, acc[b][q].im);
                        }
                    }
                }
            }
        }
        if (isA)
            store_a_half<T, NB, NM3X>(P, acc, n0, l, m, Zdot);
        else if (isA2)
            store_a2_half<T, NB2, NR2X>(P, acc, n0, m, lm, Zdot);
    }
    __syncthreads();
    if (tid == 0) ehc_dot[atom] = *sE;
}

size_t density_jvp_lds_bytes(const DevParams<double>& Pd, int CH) {
    const int nfa = Pd.nr3a * (Pd.l3a + 1), nfm = Pd.nr2m * (Pd.l2m + 1);
    const size_t n = 16 * ((size_t)CH * 2 * Pd.lha + (size_t)CH * Pd.lhm + (size_t)CH * Pd.nm3) +
                     8 * ((size_t)CH * 2 * nfa + (size_t)CH * Pd.nb_a + (size_t)CH * nfm + CH + 1);
    return ((n + 15) & ~(size_t)15) + 16;
}

// F, F' and F'' of the embedding (src/forces.F90:1134-1233)
__device__ inline void embed2(int nfemb, double rho, double wpre, double mexp, double& F, double& dF,
                              double& d2F) {
    embed(nfemb, rho, wpre, mexp, F, dF);
    d2F = 0.0;
    if (nfemb == 0) return;
    const double w = 10.0, eps = 1e-10;
    const double s = nfemb == 2 ? 1.0 + rho : rho;
    if (!(fabs(s) > eps)) return;
    const double a = fabs(s), sg = s >= 0.0 ? 1.0 : -1.0;
    const double y0 = w * s * s;
    const double g = y0 > 30.0 ? 0.0 : exp(-y0);
    const double g1 = -2.0 * w * s * g, g2 = -2.0 * w * g + 4.0 * w * w * s * s * g;
    const double q = (1.0 - g) * a, q1 = -g1 * a + (1.0 - g) * sg, q2 = -g2 * a - 2.0 * g1 * sg;
    double y1pp = 0.0;
    if (q > 0.0) y1pp = mexp * (mexp - 1.0) * pow(q, mexp - 2.0) * q1 * q1 + mexp * pow(q, mexp - 1.0) * q2;
    const double y2pp = mexp * (g2 * a + 2.0 * g1 * sg);
    d2F = wpre * sg * (y1pp + y2pp);
}

__device__ inline void inner_cutoff(double e, double ecut, double decut, double& fc, double& dfc) {
    const double elow = ecut - decut;
    fc = 1.0;
    dfc = 0.0;
    if (e >= ecut) {
        fc = 0.0;
    } else if (e > elow) {
        const double pi = 3.14159265358979323846;
        fc = 0.5 * (1.0 + cos(pi * (e - elow) / decut));
        dfc = -0.5 * sin(pi * (e - elow) / decut) * pi / decut;
    }
}

// alpha_{i,d}, beta_{i,d} (one block per atom), the factors of B_{k,i} and D_u B_{k,i} in
// dL/dc_{k,d} for L = sum_s gE_s E_s - D_u E:
//   alpha = gE fc F'_d - fc F''_d D_u rho_d - F'_d dfc D_u e_hc,  beta = -fc F'_d.
// D_u rho_d = Re sum_x G_x Zdot_x over the representatives of conjugate pairs, G = drho_d/dZ
// folded (the conjugate symmetry of Zdot allows it): with one density G = W / (fc F'), with two
// the adjoint's per-density output Wd[atom][d].
template <int ND>
__global__ void __launch_bounds__(256) k_fit_atom(DevParams<double> P, const cplx<double>* __restrict__ W,
                                                 const cplx<double>* __restrict__ Zdot,
                                                 const double* __restrict__ rho, const double* __restrict__ ehc,
                                                 const double* __restrict__ ehc_dot,
                                                 const int* __restrict__ atom_struct,
                                                 const double* __restrict__ gE, double* __restrict__ alpha,
                                                 double* __restrict__ beta) {
    __shared__ double red[8];
    const int atom = blockIdx.x;
    double tot[ND];
    const cplx<double>* zd = Zdot + (size_t)atom * P.nz;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const cplx<double>* w = W + ((size_t)atom * ND + d) * P.nz;
        double s = 0.0;
        for (int x = threadIdx.x; x < P.nz; x += blockDim.x) s += re_mul(w[x], zd[x]);
        tot[d] = block_sum<double>(s, red);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double fc, dfc;
        inner_cutoff(ehc[atom], P.ecut, P.decut, fc, dfc);
        const double g = atom_struct ? gE[atom_struct[atom]] : 0.0;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double F, dF, d2F;
            embed2(P.nfemb, rho[(size_t)atom * ND + d], P.p1[d][0], P.p1[d][1], F, dF, d2F);
            const double sc = fc * dF;
            const double rho_dot = ND == 2 ? tot[d] : (sc != 0.0 ? tot[d] / sc : 0.0);
            alpha[(size_t)atom * ND + d] = g * fc * dF - fc * d2F * rho_dot - dF * dfc * ehc_dot[atom];
            beta[(size_t)atom * ND + d] = -fc * dF;
        }
    }
}

// M_{t,d} += sum_i alpha_{i,d} P_t,i + beta_{i,d} D_u P_t,i over a chunk of atoms (Z, Zdot of one
// atom in LDS at a time); each thread owns kFitMpt monomials so that an atom's Z and Zdot are
// staged once per kFitMpt * 256 monomials; grid = (monomial tiles, atom chunks); M is
// [ND][nterms] complex
constexpr int kFitMpt = 8;
template <int ND>
__global__ void __launch_bounds__(256) k_fit_accum(int nterms, int nz, int ny, const int2* __restrict__ ypair,
                                                  const ushort4* __restrict__ idx,
                                                  const cplx<double>* __restrict__ Z,
                                                  const cplx<double>* __restrict__ Zdot,
                                                  const double* __restrict__ alpha, const double* __restrict__ beta,
                                                  int natom, int chunk, double* __restrict__ M) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int nzx = nz + 1 + ny;              // Z' = [Z | 1 | Y] as in the adjoint
    cplx<double>* sZ = (cplx<double>*)smem;  // [nzx]
    cplx<double>* sD = sZ + nzx;              // [nzx]
    const int t0 = blockIdx.x * blockDim.x * kFitMpt + threadIdx.x;
    ushort4 ix[kFitMpt];
    double accr[ND][kFitMpt], acci[ND][kFitMpt];
#pragma unroll
    for (int j = 0; j < kFitMpt; ++j) {
        const int t = t0 + j * blockDim.x;
        ix[j] = t < nterms ? idx[t] : ushort4{0, 0, 0, 0};
#pragma unroll
        for (int d = 0; d < ND; ++d) accr[d][j] = acci[d][j] = 0.0;
    }
    const int a0 = blockIdx.y * chunk, a1 = min(natom, a0 + chunk);
    for (int a = a0; a < a1; ++a) {
        __syncthreads();
        for (int x = threadIdx.x; x < nz; x += blockDim.x) {
            sZ[x] = Z[(size_t)a * nz + x];
            sD[x] = Zdot[(size_t)a * nz + x];
        }
        if (threadIdx.x == 0) {
            sZ[nz] = {1, 0};
            sD[nz] = {0, 0};
        }
        if (ny > 0) {  // Y = Z[a] Z[b], D_u Y = D_u Z[a] Z[b] + Z[a] D_u Z[b]
            __syncthreads();
            for (int i = threadIdx.x; i < ny; i += blockDim.x) {
                const int2 pr = ypair[i];
                const cplx<double> za = sZ[pr.x], zb = sZ[pr.y];
                sZ[nz + 1 + i] = cmul(za, zb);
                cplx<double> d = cmul(sD[pr.x], zb);
                cfma(d, za, sD[pr.y]);
                sD[nz + 1 + i] = d;
            }
        }
        __syncthreads();
        double al[ND], be[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            al[d] = alpha[(size_t)a * ND + d];
            be[d] = beta[(size_t)a * ND + d];
        }
#pragma unroll
        for (int j = 0; j < kFitMpt; ++j) {
            if (t0 + j * (int)blockDim.x >= nterms) break;
            const cplx<double> z0 = sZ[ix[j].x], z1 = sZ[ix[j].y], z2 = sZ[ix[j].z], z3 = sZ[ix[j].w];
            const cplx<double> d0 = sD[ix[j].x], d1 = sD[ix[j].y], d2 = sD[ix[j].z], d3 = sD[ix[j].w];
            const cplx<double> z01 = cmul(z0, z1), z23 = cmul(z2, z3);
            const cplx<double> Pt = cmul(z01, z23);
            // product rule: d(z0 z1 z2 z3) = (d0 z1 + z0 d1) z23 + z01 (d2 z3 + z2 d3)
            cplx<double> l = cmul(d0, z1), r = cmul(d2, z3);
            cfma(l, z0, d1);
            cfma(r, z2, d3);
            cplx<double> Pd = cmul(l, z23);
            cfma(Pd, z01, r);
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                accr[d][j] += al[d] * Pt.re + be[d] * Pd.re;
                acci[d][j] += al[d] * Pt.im + be[d] * Pd.im;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < kFitMpt; ++j) {
        const int t = t0 + j * blockDim.x;
        if (t < nterms) {
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                atomicAdd(&M[2 * ((size_t)d * nterms + t)], accr[d][j]);
                atomicAdd(&M[2 * ((size_t)d * nterms + t) + 1], acci[d][j]);
            }
        }
    }
}

// gradient of the B1 coefficients: grad[d * nb_m + n] = sum_i alpha_{i,d} (g^m_n(mu_i) - 1)
__global__ void k_fit_b1(DevParams<double> P, int natom, int nd, const double* __restrict__ dmomi,
                         const double* __restrict__ alpha, double* __restrict__ grad, int* err) {
    const int atom = blockIdx.x * blockDim.x + threadIdx.x;
    if (atom >= natom) return;
    double w;
    bool unit;
    const double* row = magnetic_row(P, dmomi[atom], w, unit, err);
    for (int n = 0; n < P.nb_m; ++n) {
        double v = unit ? 1.0 : 0.0, dv;
        if (row) horner(row + 4 * n, w, P.rs_m, v, dv);
        for (int d = 0; d < nd; ++d) atomicAdd(&grad[d * P.nb_m + n], alpha[(size_t)atom * nd + d] * (v - 1.0));
    }
}

// gradient of the coefficients of orders 2..6: function k of the concatenated list (nb.n[0]
// functions of order 2, ..., nb.n[4] of order 6), density d: Re sum_t u_{k,t} M_{t,d}, stored at
// grad[nd * (functions of lower orders) + d * nb.n[o] + j] (the file's order: per body order,
// density-major); grid.y = density
struct OrderCounts {
    int n[5];
};
__global__ void k_fit_grad(int nterms, OrderCounts nb, const int* __restrict__ ptr, const int* __restrict__ tt,
                           const double* __restrict__ uu, const double* __restrict__ M, double* __restrict__ grad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x, d = blockIdx.y, nd = gridDim.y;
    int lower = 0, o = 0;
    while (o < 5 && k >= lower + nb.n[o]) lower += nb.n[o++];
    if (o == 5) return;
    const double* Md = M + 2 * (size_t)d * nterms;
    double s = 0.0;
    for (int e = ptr[k]; e < ptr[k + 1]; ++e) s += uu[e] * Md[2 * (size_t)tt[e]];
    grad[nd * lower + d * nb.n[o] + (k - lower)] = s;
}

// per-atom forces f_i = -sum_p F_p (own pairs) + sum_{p: j(p) = i} F_p, and per-structure
// energies (atomics; outputs zeroed by the caller)
__global__ void k_fit_assemble(int natom, const int* __restrict__ off, const int* __restrict__ nlist,
                               const double* __restrict__ flist, const double* __restrict__ eout,
                               const int* __restrict__ atom_struct, double* __restrict__ f,
                               double* __restrict__ estruct) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= natom) return;
    double fx = 0, fy = 0, fz = 0;
    for (int p = off[i]; p < off[i + 1]; ++p) {
        const double a = flist[3 * (size_t)p], b = flist[3 * (size_t)p + 1], c = flist[3 * (size_t)p + 2];
        fx -= a; fy -= b; fz -= c;
        const int j = nlist[p];
        atomicAdd(&f[3 * (size_t)j], a);
        atomicAdd(&f[3 * (size_t)j + 1], b);
        atomicAdd(&f[3 * (size_t)j + 2], c);
    }
    atomicAdd(&f[3 * (size_t)i], fx);
    atomicAdd(&f[3 * (size_t)i + 1], fy);
    atomicAdd(&f[3 * (size_t)i + 2], fz);
    if (estruct) atomicAdd(&estruct[atom_struct[i]], eout[i]);
}

__global__ void k_fill_double(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

template <int L, int CH, int NB, int NM3X, int NB2, int NR2X>
static hipError_t run_jvp_class(const DevParams<double>& P, int natom, const int* off, const int* nlist,
                                const double* du, const double* dlist, const double* elist, const double* dmom,
                                const double* emom, cplx<double>* Zdot, double* ehc_dot, int* err, hipStream_t s) {
    auto k = k_density_jvp<L, CH, NB, NM3X, NB2, NR2X>;
    const size_t lds = density_jvp_lds_bytes(P, CH);
    static bool set = false;
    static size_t limit = 0;
    hipError_t e;
    if (!set) {
        int dev = 0, maxsh = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) != hipSuccess)
            return e;
        hipFuncAttributes fa;
        if ((e = hipFuncGetAttributes(&fa, (const void*)k)) != hipSuccess) return e;
        limit = (size_t)maxsh - fa.sharedSizeBytes;
        if ((e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)limit)) !=
            hipSuccess)
            return e;
        set = true;
    }
    if (lds > limit) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k, dim3(natom), dim3(256), lds, s, P, off, nlist, du, dlist, elist, dmom, emom, Zdot,
                       ehc_dot, err);
    return hipGetLastError();
}

hipError_t fit_backward(const DevParams<double>& P, int cls, int natom, const int* off, const int* nlist,
                        const double* du, const double* dlist, const double* elist, const double* dmomi,
                        const double* dmom, const double* emom, const int* atom_struct, const double* gE,
                        Buffers& B, FitBuffers& F, int nterms, const int* nbo5, double* grad, hipStream_t s) {
    if (natom <= 0) return hipSuccess;
    const int nd = P.nd == 2 ? 2 : 1;
    if (nd == 2 && !B.Wd) return hipErrorInvalidValue;
    OrderCounts nbo;
    int nb234 = 0;
    for (int o = 0; o < 5; ++o) nb234 += (nbo.n[o] = nbo5[o]);
    hipError_t e = cls == 0   ? run_jvp_class<6, 32, 4, 4, 4, 2>(P, natom, off, nlist, du, dlist, elist, dmom, emom,
                                                                 (cplx<double>*)F.Zdot, F.ehc_dot, B.err, s)
                   : cls == 1 ? run_jvp_class<2, 32, 1, 27, 2, 5>(P, natom, off, nlist, du, dlist, elist, dmom, emom,
                                                                  (cplx<double>*)F.Zdot, F.ehc_dot, B.err, s)
                              : run_jvp_class<6, 32, 4, 9, 4, 5>(P, natom, off, nlist, du, dlist, elist, dmom, emom,
                                                                 (cplx<double>*)F.Zdot, F.ehc_dot, B.err, s);
    if (e != hipSuccess) return e;
    if (nd == 2)
        hipLaunchKernelGGL(k_fit_atom<2>, dim3(natom), dim3(256), 0, s, P, (const cplx<double>*)B.Wd,
                           (const cplx<double>*)F.Zdot, B.rho, B.ehc, F.ehc_dot, atom_struct, gE, F.alpha, F.beta);
    else
        hipLaunchKernelGGL(k_fit_atom<1>, dim3(natom), dim3(256), 0, s, P, (const cplx<double>*)B.W,
                           (const cplx<double>*)F.Zdot, B.rho, B.ehc, F.ehc_dot, atom_struct, gE, F.alpha, F.beta);
    const size_t nM = 2 * (size_t)nterms * nd;
    hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((nM + 255) / 256)), dim3(256), 0, s, F.M, nM);
    const size_t ng = ((size_t)P.nb_m + nb234) * nd;
    hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, s, grad, ng);
    const int chunk = 64;
    const size_t lds = 2 * (size_t)P.nzx * sizeof(cplx<double>);
    {
        // Z' and its derivative of one atom in LDS: beyond the default dynamic limit for long
        // feature vectors
        static bool set[2] = {false, false};
        const void* k = nd == 2 ? (const void*)k_fit_accum<2> : (const void*)k_fit_accum<1>;
        if (!set[nd - 1]) {
            int dev = 0, maxsh = 0;
            if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
            if ((e = hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) != hipSuccess)
                return e;
            if ((e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, maxsh)) != hipSuccess)
                return e;
            set[nd - 1] = true;
        }
    }
    dim3 grid((unsigned)((nterms + 256 * kFitMpt - 1) / (256 * kFitMpt)), (unsigned)((natom + chunk - 1) / chunk));
    if (nd == 2)
        hipLaunchKernelGGL(k_fit_accum<2>, grid, dim3(256), lds, s, nterms, P.nz, P.ny, P.ypair, (const ushort4*)F.idx,
                           (const cplx<double>*)B.Z, (const cplx<double>*)F.Zdot, F.alpha, F.beta, natom, chunk,
                           F.M);
    else
        hipLaunchKernelGGL(k_fit_accum<1>, grid, dim3(256), lds, s, nterms, P.nz, P.ny, P.ypair, (const ushort4*)F.idx,
                           (const cplx<double>*)B.Z, (const cplx<double>*)F.Zdot, F.alpha, F.beta, natom, chunk,
                           F.M);
    hipLaunchKernelGGL(k_fit_b1, dim3((natom + 255) / 256), dim3(256), 0, s, P, natom, nd, dmomi, F.alpha, grad,
                       B.err);
    if (nb234 > 0)
        hipLaunchKernelGGL(k_fit_grad, dim3((nb234 + 255) / 256, nd), dim3(256), 0, s, nterms, nbo, F.ub_ptr, F.ub_t,
                           F.ub_u, F.M, grad + (size_t)nd * P.nb_m);
    return hipGetLastError();
}

hipError_t fit_assemble(int natom, const int* off, const int* nlist, const double* flist, const double* eout,
                        const int* atom_struct, int nstruc, double* f, double* estruct, hipStream_t s) {
    const size_t nf = 3 * (size_t)natom;
    if (nf) hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, f, nf);
    if (estruct && nstruc > 0)
        hipLaunchKernelGGL(k_fill_double, dim3((unsigned)((nstruc + 255) / 256)), dim3(256), 0, s, estruct,
                           (size_t)nstruc);
    if (natom > 0)
        hipLaunchKernelGGL(k_fit_assemble, dim3((natom + 255) / 256), dim3(256), 0, s, natom, off, nlist, flist,
                           eout, atom_struct, f, estruct);
    return hipGetLastError();
}

// ============================================================================ launch
__global__ void k_clear_flags(int* p) { *p = 0; }

hipError_t clear_flags(int* p, hipStream_t s) {
    // a kernel rather than hipMemsetAsync, so that the evaluation can be captured in a graph
    hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(1), 0, s, p);
    return hipGetLastError();
}

// shape classes: 0 = small magnetic A block (nm3 <= 4, nr2m <= 2, nb_a <= 10, nr3a <= 8, l <= 6)
//                1 = wide magnetic A block  (nm3 <= 27, nr2m <= 5, nb_a <= 16, nr3a <= 4, l <= 2)
//                2 = medium magnetic A block at l <= 6 (nm3 <= 9, nr2m <= 5, nb_a <= 16, nr3a <= 8)
int supported_nm3_class(int nm3, int nr2m, int nb_a, int l3a, int l2m, int nr3a) {
    if (nb_a > 16) return -1;
    if (nm3 <= 4 && nr2m <= 2 && l3a <= 6 && l2m <= 6 && nr3a <= 8 && nb_a <= 10) return 0;
    if (nm3 <= 27 && nr2m <= 5 && l3a <= 2 && l2m <= 2 && nr3a <= 4) return 1;
    if (nm3 <= 9 && nr2m <= 5 && l3a <= 6 && l2m <= 6 && nr3a <= 8) return 2;
    return -1;
}

template <typename T, int L, int CH, bool SPLIT, bool TASKSPLIT, int NB, int NM3X, int NB2, int NR2X, int NRAX, int NBAX,
          bool ACC>
static hipError_t run_class(const DevParams<T>& P, int natom, int atom_base, const int* off, const double* dlist,
                            const double* elist, const double* dmomi, const double* emomi,
                            const double* dmom, const double* emom, double* eout, double* flist,
                            double* tlist, double* tlist0, Buffers& B, hipStream_t s, hipEvent_t* ev) {
    DevParams<double> Pd;  // only the dims are used for LDS sizing
    Pd.nr3a = P.nr3a; Pd.l3a = P.l3a; Pd.nr2m = P.nr2m; Pd.l2m = P.l2m; Pd.lha = P.lha;
    Pd.lhm = P.lhm; Pd.nm3 = P.nm3; Pd.nb_a = P.nb_a; Pd.nb_m = P.nb_m;
    const size_t lds1 = density_lds_bytes(Pd, CH, sizeof(T));
    static_assert(!ACC || sizeof(T) == 4, "the accumulation recipe belongs to the FP32 kernels");
    auto k1 = k_density<T, L, CH, NB, NM3X, NB2, NR2X, SPLIT, TASKSPLIT, ACC>;
    const bool nd2 = P.nd == 2;
    // the Q1 instantiation exists only for the four-column class.  Q1 visits every radial index
    // up to its compile-time bound, so the bounds 4 (for nr3a <= 4), 5, 6, 7 and NRAX each have
    // an instance; without Q1 all five name the general kernel
    constexpr bool kQ1 = NM3X == 4;
    const bool q1 = kQ1 && P.nm3 == 4 && P.lsm3 == 4;
    auto k3g = k_forces<T, L, L, NM3X, NBAX, NRAX, NR2X, false>;
    constexpr int kB4 = kQ1 && NRAX > 4 ? 4 : NRAX, kB5 = kQ1 && NRAX > 5 ? 5 : NRAX,
                  kB6 = kQ1 && NRAX > 6 ? 6 : NRAX, kB7 = kQ1 && NRAX > 7 ? 7 : NRAX;
    constexpr int kQB[5] = {kB4, kB5, kB6, kB7, NRAX};
    const decltype(k3g) k3qs[5] = {
        k_forces<T, L, L, NM3X, NBAX, kB4, NR2X, kQ1>, k_forces<T, L, L, NM3X, NBAX, kB5, NR2X, kQ1>,
        k_forces<T, L, L, NM3X, NBAX, kB6, NR2X, kQ1>, k_forces<T, L, L, NM3X, NBAX, kB7, NR2X, kQ1>,
        k_forces<T, L, L, NM3X, NBAX, NRAX, NR2X, kQ1>};
    int qi = 0;  // the smallest bound that covers nr3a
    while (qi < 4 && kQB[qi] < P.nr3a) ++qi;
    auto k3 = q1 ? k3qs[qi] : k3g;
    hipError_t e = hipSuccess;
    // the dynamic-LDS limit is raised once per kernel instantiation (to the workgroup
    // maximum), so that evaluations can be captured into graphs without further API calls.
    // limit[]: 0 density, 1 adjoint, 2 adjoint with two densities, 3 two-atom adjoint,
    // 4 general forces, 5 + i the Q1 forces of bound kQB[i]
    static bool attributes_set = false;
    static size_t limit[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (!attributes_set) {
        int dev = 0, maxsh = 0;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev)) != hipSuccess)
            return e;
        const void* ks[10] = {(const void*)k1, (const void*)k_adjoint<T, false, ACC>,
                              (const void*)k_adjoint<T, true, ACC>, nullptr, (const void*)k3g,
                              (const void*)k3qs[0], (const void*)k3qs[1], (const void*)k3qs[2],
                              (const void*)k3qs[3], (const void*)k3qs[4]};
        if constexpr (sizeof(T) == 4) ks[3] = (const void*)k_adjoint_pair<ACC>;
        for (int i = 0; i < 10; ++i) {
            if (!ks[i]) continue;
            hipFuncAttributes fa;
            if ((e = hipFuncGetAttributes(&fa, ks[i])) != hipSuccess) return e;
            // the adjoints address their dynamic LDS from 0 (lds_at)
            if (i >= 1 && i <= 3 && fa.sharedSizeBytes != 0) return hipErrorInvalidKernelFile;
            limit[i] = (size_t)maxsh > fa.sharedSizeBytes ? (size_t)maxsh - fa.sharedSizeBytes : 0;
            if ((e = hipFuncSetAttribute(ks[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)limit[i])) !=
                hipSuccess)
                return e;
        }
        attributes_set = true;
    }
    // FP32 with one density: two atoms per adjoint block (their Z and W interleaved in LDS)
    const bool pair = sizeof(T) == 4 && !nd2 && P.crec16 != nullptr;
    const size_t lds2p = 2 * (size_t)P.nzx * 16 + (2 * (kAdjointThreads / 64) + 2) * sizeof(double);
    if (pair && lds2p > limit[3]) return hipErrorInvalidValue;
    const size_t lds2 =
        (nd2 ? 3 : 2) * (size_t)P.nzx * sizeof(cplx<T>) + (2 * (kAdjointThreads / 64) + 2) * sizeof(double);
    const size_t lds3 = (size_t)P.nz * sizeof(cplx<T>);
    if (lds1 > limit[0] || lds2 > limit[nd2 ? 2 : 1] || lds3 > limit[q1 ? 5 + qi : 4])
        return hipErrorInvalidValue;
    if (ev) (void)hipEventRecord(ev[0], s);
    hipLaunchKernelGGL(k1, dim3(natom), dim3(256), lds1, s, P, off, dlist, elist, dmomi, emomi, dmom,
                       emom, (cplx<T>*)B.Z, B.ehc, B.rho1, B.err);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[1], s);
    if constexpr (sizeof(T) == 4) {
        if (pair)
            hipLaunchKernelGGL(k_adjoint_pair<ACC>, dim3((natom + 1) / 2), dim3(kAdjointThreads), lds2p, s, P, atom_base,
                               ACC ? P.crec16 : P.crec16u, (const cplx<T>*)B.Z, (cplx<T>*)B.W, (char*)B.Pt, B.ehc, B.rho1, eout, B.dE,
                               B.edf, B.rho, natom);
    }
    if (pair) {
    } else if (nd2)
        hipLaunchKernelGGL((k_adjoint<T, true, ACC>), dim3(natom), dim3(kAdjointThreads), lds2, s, P, atom_base,
                           (const cplx<T>*)B.Z, (cplx<T>*)B.W, (cplx<T>*)B.Wd, (char*)B.Pt, B.ehc, B.rho1, eout,
                           B.dE, B.edf, B.rho);
    else
        hipLaunchKernelGGL((k_adjoint<T, false, ACC>), dim3(natom), dim3(kAdjointThreads), lds2, s, P, atom_base,
                           (const cplx<T>*)B.Z, (cplx<T>*)B.W, (cplx<T>*)nullptr, (char*)B.Pt, B.ehc, B.rho1, eout,
                           B.dE, B.edf, B.rho);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev) (void)hipEventRecord(ev[2], s);
    hipLaunchKernelGGL(k3, dim3(natom), dim3(64), lds3, s, P, off, dlist, elist, dmomi, emomi, dmom, emom,
                       (const cplx<T>*)B.W, B.dE, B.edf, flist, tlist, tlist0, B.err);
    if (ev) (void)hipEventRecord(ev[3], s);
    return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::run(const DevParams<T>& P, int natom, int atom_base, const int* off, const double* dlist,
                          const double* elist, const double* dmomi, const double* emomi,
                          const double* dmom, const double* emom, double* eout, double* flist,
                          double* tlist, double* tlist0, Buffers& B, hipStream_t s, int cls,
                          hipEvent_t* ev, int dvar, bool acc) {
    if (natom <= 0) return hipSuccess;
    // Density staging: with many radial functions per pair (lmax-6 headers) or in FP64 the
    // per-pair staging is spread over the whole block, 32 neighbours per round; light FP32
    // headers stage 64 neighbours per round, one thread per pair and side (measured on the
    // 16000-atom workload: lmax 6 0.91 vs 1.07 ms, lmax 4 0.80 vs 0.60 ms).  Light headers
    // also split their A-column tasks over idle threads (lmax 4: FP32 0.61 -> 0.52 ms, FP64
    // 0.93 -> 0.81 ms; with lmax 6 the runtime split costs FP64 3%).
    const bool heavy = P.nb_a + P.nr3a * (P.l3a + 1) >= 40;
#define MAGACE_RUN_ACC(CH, SPLIT, TS, ACC)                                                                     \
    (cls == 0   ? run_class<T, 6, CH, SPLIT, TS, 4, 4, 4, 2, 8, 10, ACC>(P, natom, atom_base, off, dlist, elist,      \
                                                                         dmomi, emomi, dmom, emom, eout, flist,      \
                                                                         tlist, tlist0, B, s, ev)                    \
     : cls == 1 ? run_class<T, 2, CH, SPLIT, TS, 1, 27, 2, 5, 4, 16, ACC>(P, natom, atom_base, off, dlist, elist,     \
                                                                          dmomi, emomi, dmom, emom, eout, flist,     \
                                                                          tlist, tlist0, B, s, ev)                   \
                : run_class<T, 6, CH, SPLIT, TS, 4, 9, 4, 5, 8, 16, ACC>(P, natom, atom_base, off, dlist, elist,      \
                                                                         dmomi, emomi, dmom, emom, eout, flist,      \
                                                                         tlist, tlist0, B, s, ev))
    // the accumulation recipe (MAGACE_MIXED_ACC) exists for the FP32 kernels only
#define MAGACE_RUN(CH, SPLIT, TS) \
    (sizeof(T) == 4 && acc ? MAGACE_RUN_ACC(CH, SPLIT, TS, (sizeof(T) == 4)) : MAGACE_RUN_ACC(CH, SPLIT, TS, false))
    // a density variant chosen by the tuner (kDensityVariants) replaces the rule above
    if (dvar < 0) dvar = heavy ? 0 : (sizeof(T) == 8 ? 1 : 2);
    if (dvar == 0) return MAGACE_RUN(32, true, false);
    if (dvar == 1) return MAGACE_RUN(32, true, true);
    return MAGACE_RUN(64, false, true);
#undef MAGACE_RUN
#undef MAGACE_RUN_ACC
}

template struct Launch<double>;
template struct Launch<float>;

}  // namespace magace
