/*
 * magace_oracle.c -- CPU FP64 restatement of the reference libace hot path.
 *
 * TEST INFRASTRUCTURE (the oracle / checker).  It is never linked into or called by the
 * product; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
 *
 * It follows the reference data flow step by step:
 *   parser            src/util.F90:459-1108   (rwpara, read half)
 *   enumeration       src/util.F90:1613-4140  (setuptables, orders 1..6), 4673-5330 (bcreate)
 *   Clebsch-Gordan    src/util.F90:4149-4252  (setupcg, Racah formula)
 *   couplings         src/util.F90:4256-4548  (setuphighercouplings)
 *   radial basis      src/radial.F90:6-33,430-668
 *   spline tables     src/radial.F90:37-49,164-426 ; lookups :53-160
 *   Y_lm              src/spherical.F90:5-282
 *   prepA             src/a_basis.F90:5-361
 *   prepB             src/b_basis.F90:6-532   (with the stored dB/dA of the reference)
 *   energy            src/forces.F90:976-1289
 *   forceweights      src/forces.F90:5-362
 *   forcecompute      src/forces.F90:366-634
 *   torquecompute     src/forces.F90:638-972
 *   acelib_ copy-out  src/ace.F90:288-318
 * The reference keeps per-pair packs for all atoms in module arrays; here the same values
 * are kept per atom (prepA -> prepB -> weights -> forces for one atom at a time), which is
 * the same arithmetic with bounded memory and lets OpenMP run atoms in parallel.
 *
 * Parity pins (tests/): golden vectors from the reference itself (oracle/_ref),
 * sympy Clebsch-Gordan, scipy spherical harmonics, Chebyshev identities, spline-node
 * identities, enumeration counts printed by the reference, central finite differences.
 */
#define _POSIX_C_SOURCE 200809L
#include "magace_oracle.h"

#include <complex.h>
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static const double PI = 3.14159265358979323846264338327950288419;
#define NLUT 10000 /* src/radial.F90:43 */

static __thread char g_err[512];
const char *orc_error(void) { return g_err; }
#define FAIL(...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); goto fail; } while (0)

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz ? sz : 1);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

struct orc_pot {
    orc_header h;
    int lmax, nradial;         /* util.F90:224-225 */
    int lsa, lsm, lha, lhm;    /* lmaxsq/lmaxhalf atomic & magnetic, util.F90:1681-1697 */
    double fac[8 * 16 + 2];    /* util.F90:1703-1714 */
    /* enumeration */
    int nvar1, nvar2, nvar3, nvar4;
    int *b1index;                       /* [nvar1]           */
    int *b2index;                       /* [nvar2][4]        */
    int *b3index;                       /* [nvar3][9]        */
    int *b4index;                       /* [nvar4][15]       */
    int nvar3l_m, maddmax3m, *b3lindex_m /*[k][3]*/, *madd3_m, *b3lm_m /*[k][maddmax][3]*/;
    int *b3map_a, *b3map_m;
    int nvar4l_a, maddmax4a, *b4lindex_a /*[k][3]*/, *madd4_a, *b4lm_a /*[k][maddmax][3]*/;
    int nvar4l_m, maddmax4m, *b4lindex_m /*[k][5]*/, *madd4_m, *b4lm_m /*[k][maddmax][4]*/;
    int *b4map_a, *b4map_m;
    /* orders 5 and 6 (util.F90:2822-3436, 3439-4140): index tuples (b5index, b6index columns)
     * and, per function, the (atomic x magnetic) m-combinations with the product of their
     * couplings (util.F90:4392-4548); m5lm rows: m1..m4, m0'..m4'; m6lm rows: m1..m5, m0'..m5' */
    int nvar5, nvar6;
    int *b5index;   /* [nvar5][21] */
    int *b6index;   /* [nvar6][27] */
    int *m5ptr, *m5lm; double *m5cf;
    int *m6ptr, *m6lm; double *m6cf;
    /* couplings (per m-list entry, util.F90:4256-4388) */
    double *cg01;   /* [l2m+1][2*lmax+1]  indexed m+lmax */
    double *cg02a;  /* [l3a+1][2*lmax+1]  */
    double *cg02m;  /* [nvar3l_m][maddmax3m] */
    double *cg03a;  /* [nvar4l_a][maddmax4a] */
    double *cg03m;  /* [nvar4l_m][maddmax4m] */
    /* parameters */
    double *c0, *c2, *c3, *c4, *c5, *c6;  /* [nvar][ndensity] */
    double *crad_a;             /* [nr3_a][l3a+1][nb_a]  (n'' fastest) */
    double *crad_m;             /* [nr2_m][l2m+1][nb_m] */
    double p1[2][2];            /* [k][nd] */
    double core[2];
    double cutoff, cutoff_m;    /* global (acelibini) */
    /* spline tables, node-major: [n][func][4] */
    int have_luts;
    double rs_a, rs_m;
    double *lut_ga, *lut_fa, *lut_hc, *lut_gm, *lut_fm;
};

/* ----------------------------------------------------------------- index helpers */
/* shindex with m>=0 packed first (util.F90:1716-1776); 0-based */
static inline int shx(int L, int l, int m) {
    if (m >= 0) return l * (l + 1) / 2 + m;
    int half = (L + 1) * (L + 2) / 2;
    return half + l * (l - 1) / 2 + (m + l);
}
static inline int shback(int L, int k) {
    int half = (L + 1) * (L + 2) / 2;
    if (k < half) { int l = 0; while ((l + 1) * (l + 2) / 2 <= k) ++l; return l; }
    k -= half;
    int l = 1; while (l * (l + 1) / 2 <= k) ++l; return l;
}
static inline double sgnm(int m) { return (m & 1) ? -1.0 : 1.0; } /* (-1)**m */

/* ----------------------------------------------------------------- CG (util.F90:4149-4252) */
static double cg_racah(const double *fac, int j1, int j2, int J, int m1, int m2, int M) {
    if (!(j1 + j2 >= J && abs(j1 - j2) <= J)) return 0.0;
    if (abs(m1) > j1 || abs(m2) > j2 || M != m1 + m2 || abs(M) > J) return 0.0;
    double prefac = ((double)(2 * J + 1) * fac[j1 + j2 - J] * fac[j1 - j2 + J] * fac[-j1 + j2 + J] /
                     fac[j1 + j2 + J + 1]) *
                    (fac[J + M] * fac[J - M] * fac[j1 + m1] * fac[j1 - m1] * fac[j2 + m2] * fac[j2 - m2]);
    prefac = sqrt(prefac);
    double xsum = 0.0;
    for (int k = 0; k <= j1 + j2 + J + 1; ++k) {
        int a = j1 + j2 - J - k, b = j1 - m1 - k, c = j2 + m2 - k, d = J - j2 + m1 + k,
            e = J - j1 - m2 + k;
        if (a >= 0 && b >= 0 && c >= 0 && d >= 0 && e >= 0)
            xsum += ((k & 1) ? -1.0 : 1.0) / (fac[k] * fac[a] * fac[b] * fac[c] * fac[d] * fac[e]);
    }
    return prefac * xsum;
}
double orc_cg(int j1, int j2, int J, int m1, int m2, int M) {
    double fac[8 * 16 + 2];
    double x = 1.0; fac[0] = 1.0;
    for (int l = 0; l < 8 * 16 + 1; ++l) { x *= (double)(l + 1); fac[l + 1] = x; }
    return cg_racah(fac, j1, j2, J, m1, m2, M);
}

/* ----------------------------------------------------------------- bcreate (util.F90:4673-4960) */
static void bcreate(const orc_pot *p, int nrank, const int *vecnr, const int *vecl,
                    int *numlint, int *arrlint, int *isordered, int *isbasis) {
    /* single element: vecn == 1 everywhere; veck = ((nel+1-1)*(lmax+1)+l)*(nradial+1)+nr */
    int veck[8];
    for (int n = 0; n < nrank; ++n)
        veck[n] = ((1) * (p->lmax + 1) + vecl[n]) * (p->nradial + 1) + vecnr[n];
    *numlint = 1;
    *isordered = 0; *isbasis = 0;
    if (nrank == 2) {
        if (vecl[0] >= vecl[1]) *isbasis = 1;
        if (veck[0] >= veck[1]) *isordered = 1;
    } else if (nrank == 3) {
        int l1 = vecl[0], l2 = vecl[1], l3 = vecl[2];
        if (l1 + l2 >= l3 && abs(l1 - l2) <= l3 && (l1 + l2 + l3) % 2 == 0) *isbasis = 1;
        if (veck[0] >= veck[1] && veck[1] >= veck[2]) *isordered = 1;
    } else if (nrank == 4) {
        int l1 = vecl[0], l2 = vecl[1], l3 = vecl[2], l4 = vecl[3];
        int lmx = l1; if (l2 > lmx) lmx = l2; if (l3 > lmx) lmx = l3; if (l4 > lmx) lmx = l4;
        *numlint = 0;
        *isordered = !(veck[0] < veck[1] || veck[1] < veck[2] || veck[2] < veck[3]);
        for (int lint = 0; lint <= 2 * lmx; ++lint) {
            int ok = 1;
            if (lint > l1 + l2 || lint > l3 + l4) ok = 0;
            if (lint < abs(l1 - l2) || lint < abs(l3 - l4)) ok = 0;
            if (veck[0] == veck[1] && (lint % 2) != 0) ok = 0;
            if (veck[2] == veck[3] && (lint % 2) != 0) ok = 0;
            if (ok) arrlint[(*numlint)++] = lint;
        }
        if (*numlint > 0) *isbasis = 1;
        if ((l1 + l2 + l3 + l4) % 2 != 0) *isbasis = 0;
        /* the rank reduction that would use randAB is commented out (util.F90:4942) */
    }
}

/* ----------------------------------------------------------------- enumeration */
static int enumerate(orc_pot *p) {
    const orc_header *h = &p->h;
    /* order 1: util.F90:1801-1829 */
    p->nvar1 = h->nb_m;
    p->b1index = xcalloc(p->nvar1, sizeof(int));
    for (int nr = 1; nr <= h->nb_m; ++nr) p->b1index[nr - 1] = nr;

    /* order 2: util.F90:1835-1959 */
    {
        int cap = h->nr2_m * h->nr2_m * (h->nb_a > 0 ? h->nb_a : 1) * (h->l2m + 1) + 1;
        p->b2index = xcalloc((size_t)cap * 4, sizeof(int));
        int n = 0;
        for (int nr1 = 1; nr1 <= h->nr2_m; ++nr1)
            for (int nr2 = 1; nr2 <= h->nr2_m; ++nr2)
                for (int nr3 = 1; nr3 <= h->nb_a; ++nr3)
                    for (int l = 0; l <= h->l2m; ++l) {
                        int vnr[2] = {nr1, nr2}, vl[2] = {l, l}, nl, al[64], iso, isb;
                        bcreate(p, 2, vnr, vl, &nl, al, &iso, &isb);
                        if (iso && isb) {
                            int *b = p->b2index + 4 * n++;
                            b[0] = nr1; b[1] = nr2; b[2] = nr3; b[3] = l;
                        }
                    }
        p->nvar2 = n;
    }

    /* order 3: util.F90:1962-2336 */
    {
        int l3a = h->l3a, l3m = h->l3m;
        /* magnetic angular list (util.F90:2024-2050) */
        int cap = (l3m + 1) * (l3m + 1) * (l3m + 1);
        p->b3lindex_m = xcalloc((size_t)cap * 3, sizeof(int));
        int k = 0;
        for (int l2 = 0; l2 <= l3m; ++l2)
            for (int l3 = 0; l3 <= l3m; ++l3)
                for (int l4 = 0; l4 <= l3m; ++l4)
                    if (l2 + l3 >= l4 && abs(l2 - l3) <= l4 && (l2 + l3 + l4) % 2 == 0) {
                        int *b = p->b3lindex_m + 3 * k++;
                        b[0] = l2; b[1] = l3; b[2] = l4;
                    }
        p->nvar3l_m = k;
        /* basis functions (util.F90:2054-2223) */
        long capb = (long)h->nr3_a * h->nr3_a * h->nr3_m * h->nr3_m * h->nr3_m * (l3a + 1) *
                    (l3m + 1) * (l3m + 1) * (l3m + 1) + 1;
        p->b3index = xcalloc((size_t)capb * 9, sizeof(int));
        int kmax_a = h->nr3_a * (l3a + 1), kmax_m = h->nr3_m * (l3m + 1);
        int n = 0;
        for (int nr1 = 1; nr1 <= h->nr3_a; ++nr1)
        for (int nr2 = 1; nr2 <= h->nr3_a; ++nr2)
        for (int nr3 = 1; nr3 <= h->nr3_m; ++nr3)
        for (int nr4 = 1; nr4 <= h->nr3_m; ++nr4)
        for (int nr5 = 1; nr5 <= h->nr3_m; ++nr5)
        for (int l1 = 0; l1 <= l3a; ++l1)
        for (int l2 = 0; l2 <= l3m; ++l2)
        for (int l3 = 0; l3 <= l3m; ++l3)
        for (int l4 = 0; l4 <= l3m; ++l4) {
            int ka1 = h->nr3_a * l1 + nr1, ka2 = h->nr3_a * l1 + nr2;
            int km1 = h->nr3_m * l3 + nr4, km2 = h->nr3_m * l4 + nr5;
            int base = ((kmax_m + 1) * kmax_a + kmax_m) * 1;
            int nc1 = base + (kmax_m + 1) * ka1 + km1, nc2 = base + (kmax_m + 1) * ka2 + km2;
            if (!(nc1 >= nc2)) continue;
            int vnr[2] = {nr1, nr2}, vl[2] = {l1, l1}, nl, al[64], isoa, isba;
            bcreate(p, 2, vnr, vl, &nl, al, &isoa, &isba);
            int vnrm[3] = {nr3, nr4, nr5}, vlm[3] = {l2, l3, l4}, isom, isbm;
            bcreate(p, 3, vnrm, vlm, &nl, al, &isom, &isbm);
            if (isoa && isba && isbm) {
                int *b = p->b3index + 9 * n++;
                b[0] = nr1; b[1] = nr2; b[2] = nr3; b[3] = nr4; b[4] = nr5;
                b[5] = l1; b[6] = l2; b[7] = l3; b[8] = l4;
            }
        }
        p->nvar3 = n;
        p->b3map_a = xcalloc(n, sizeof(int));
        p->b3map_m = xcalloc(n, sizeof(int));
        for (int i = 0; i < n; ++i) {
            const int *b = p->b3index + 9 * i;
            p->b3map_a[i] = b[5]; /* b3lindex_atomic(k) = k-1 */
            for (int kk = 0; kk < p->nvar3l_m; ++kk) {
                const int *q = p->b3lindex_m + 3 * kk;
                if (q[0] == b[6] && q[1] == b[7] && q[2] == b[8]) p->b3map_m[i] = kk;
            }
        }
        /* magnetic m-lists (util.F90:2262-2328) */
        int mm = 0;
        p->madd3_m = xcalloc(p->nvar3l_m, sizeof(int));
        for (int kk = 0; kk < p->nvar3l_m; ++kk) {
            const int *q = p->b3lindex_m + 3 * kk;
            int c = 0;
            for (int m2 = -q[0]; m2 <= q[0]; ++m2)
                for (int m3 = -q[1]; m3 <= q[1]; ++m3)
                    if (abs(m2 + m3) <= q[2]) ++c;
            p->madd3_m[kk] = c;
            if (c > mm) mm = c;
        }
        p->maddmax3m = mm;
        p->b3lm_m = xcalloc((size_t)p->nvar3l_m * (mm ? mm : 1) * 3, sizeof(int));
        for (int kk = 0; kk < p->nvar3l_m; ++kk) {
            const int *q = p->b3lindex_m + 3 * kk;
            int c = 0;
            for (int m2 = -q[0]; m2 <= q[0]; ++m2)
                for (int m3 = -q[1]; m3 <= q[1]; ++m3) {
                    int m4 = -(m2 + m3);
                    if (abs(m4) <= q[2]) {
                        int *e = p->b3lm_m + 3 * (kk * mm + c++);
                        e[0] = m2; e[1] = m3; e[2] = m4;
                    }
                }
        }
    }

    /* order 4: util.F90:2339-2814 */
    {
        int l4a = h->l4a, l4m = h->l4m;
        int cap = (l4a + 1) * (l4a + 1) * (l4a + 1);
        p->b4lindex_a = xcalloc((size_t)cap * 3, sizeof(int));
        int k = 0;
        for (int l1 = 0; l1 <= l4a; ++l1)
            for (int l2 = 0; l2 <= l4a; ++l2)
                for (int l3 = 0; l3 <= l4a; ++l3)
                    if (l1 + l2 >= l3 && abs(l1 - l2) <= l3 && (l1 + l2 + l3) % 2 == 0) {
                        int *b = p->b4lindex_a + 3 * k++;
                        b[0] = l1; b[1] = l2; b[2] = l3;
                    }
        p->nvar4l_a = k;
        int kmax_a = h->nr4_a * (l4a + 1), kmax_m = h->nr4_m * (l4m + 1);
        int base = ((kmax_m + 1) * kmax_a + kmax_m) * 1;
        /* count first (two-pass like the reference) */
        int n = 0, capn = 1024;
        p->b4index = xcalloc((size_t)capn * 15, sizeof(int));
        for (int nr1 = 1; nr1 <= h->nr4_a; ++nr1)
        for (int nr2 = 1; nr2 <= h->nr4_a; ++nr2)
        for (int nr3 = 1; nr3 <= h->nr4_a; ++nr3)
        for (int nr4 = 1; nr4 <= h->nr4_m; ++nr4)
        for (int nr5 = 1; nr5 <= h->nr4_m; ++nr5)
        for (int nr6 = 1; nr6 <= h->nr4_m; ++nr6)
        for (int nr7 = 1; nr7 <= h->nr4_m; ++nr7)
        for (int l1 = 0; l1 <= l4a; ++l1)
        for (int l2 = 0; l2 <= l4a; ++l2)
        for (int l3 = 0; l3 <= l4a; ++l3)
        for (int l4 = 0; l4 <= l4m; ++l4)
        for (int l5 = 0; l5 <= l4m; ++l5)
        for (int l6 = 0; l6 <= l4m; ++l6)
        for (int l7 = 0; l7 <= l4m; ++l7) {
            int ka1 = h->nr4_a * l1 + nr1, ka2 = h->nr4_a * l2 + nr2, ka3 = h->nr4_a * l3 + nr3;
            int km1 = h->nr4_m * l5 + nr5, km2 = h->nr4_m * l6 + nr6, km3 = h->nr4_m * l7 + nr7;
            int nc1 = base + (kmax_m + 1) * ka1 + km1;
            int nc2 = base + (kmax_m + 1) * ka2 + km2;
            int nc3 = base + (kmax_m + 1) * ka3 + km3;
            if (!(nc1 >= nc2 && nc2 >= nc3)) continue;
            int vnr[3] = {nr1, nr2, nr3}, vl[3] = {l1, l2, l3}, nl, al[64], isoa, isba;
            bcreate(p, 3, vnr, vl, &nl, al, &isoa, &isba);
            if (!isoa) { snprintf(g_err, sizeof g_err, "Something wrong with rank = 3."); return -1; }
            int vnrm[4] = {nr4, nr5, nr6, nr7}, vlm[4] = {l4, l5, l6, l7}, nlm, alm[64], isom, isbm;
            bcreate(p, 4, vnrm, vlm, &nlm, alm, &isom, &isbm);
            if (isoa && isba && isbm) {
                for (int ni = 0; ni < nlm; ++ni) {
                    if (n >= capn) {
                        capn *= 2;
                        p->b4index = realloc(p->b4index, (size_t)capn * 15 * sizeof(int));
                    }
                    int *b = p->b4index + 15 * n++;
                    b[0] = nr1; b[1] = nr2; b[2] = nr3; b[3] = nr4; b[4] = nr5; b[5] = nr6; b[6] = nr7;
                    b[7] = l1; b[8] = l2; b[9] = l3; b[10] = l4; b[11] = l5; b[12] = l6; b[13] = l7;
                    b[14] = alm[ni];
                }
            }
        }
        p->nvar4 = n;
        p->b4map_a = xcalloc(n, sizeof(int));
        p->b4map_m = xcalloc(n, sizeof(int));
        for (int i = 0; i < n; ++i) {
            const int *b = p->b4index + 15 * i;
            for (int kk = 0; kk < p->nvar4l_a; ++kk) {
                const int *q = p->b4lindex_a + 3 * kk;
                if (q[0] == b[7] && q[1] == b[8] && q[2] == b[9]) p->b4map_a[i] = kk;
            }
        }
        /* magnetic angular groups in order of first appearance (util.F90:2615-2689) */
        p->b4lindex_m = xcalloc((size_t)(n ? n : 1) * 5, sizeof(int));
        int nm = 0;
        for (int i = 0; i < n; ++i) {
            const int *b = p->b4index + 15 * i;
            int found = -1;
            for (int kk = 0; kk < nm; ++kk) {
                const int *q = p->b4lindex_m + 5 * kk;
                if (q[0] == b[10] && q[1] == b[11] && q[2] == b[12] && q[3] == b[13] && q[4] == b[14]) {
                    found = kk; break;
                }
            }
            if (found < 0) {
                found = nm++;
                memcpy(p->b4lindex_m + 5 * found, b + 10, 5 * sizeof(int));
            }
            p->b4map_m[i] = found;
        }
        p->nvar4l_m = nm;
        /* m-lists (util.F90:2693-2810) */
        int mma = 0;
        p->madd4_a = xcalloc(p->nvar4l_a, sizeof(int));
        for (int kk = 0; kk < p->nvar4l_a; ++kk) {
            const int *q = p->b4lindex_a + 3 * kk;
            int c = 0;
            for (int m1 = -q[0]; m1 <= q[0]; ++m1)
                for (int m2 = -q[1]; m2 <= q[1]; ++m2)
                    if (abs(m1 + m2) <= q[2]) ++c;
            p->madd4_a[kk] = c; if (c > mma) mma = c;
        }
        p->maddmax4a = mma;
        p->b4lm_a = xcalloc((size_t)(p->nvar4l_a ? p->nvar4l_a : 1) * (mma ? mma : 1) * 3, sizeof(int));
        for (int kk = 0; kk < p->nvar4l_a; ++kk) {
            const int *q = p->b4lindex_a + 3 * kk;
            int c = 0;
            for (int m1 = -q[0]; m1 <= q[0]; ++m1)
                for (int m2 = -q[1]; m2 <= q[1]; ++m2) {
                    int m3 = -(m1 + m2);
                    if (abs(m3) <= q[2]) {
                        int *e = p->b4lm_a + 3 * (kk * mma + c++);
                        e[0] = m1; e[1] = m2; e[2] = m3;
                    }
                }
        }
        int mmm = 0;
        p->madd4_m = xcalloc(nm ? nm : 1, sizeof(int));
        for (int kk = 0; kk < nm; ++kk) {
            const int *q = p->b4lindex_m + 5 * kk;
            int c = 0;
            for (int M = -q[4]; M <= q[4]; ++M)
                for (int m4 = -q[0]; m4 <= q[0]; ++m4)
                    for (int m5 = -q[1]; m5 <= q[1]; ++m5)
                        for (int m6 = -q[2]; m6 <= q[2]; ++m6)
                            for (int m7 = -q[3]; m7 <= q[3]; ++m7)
                                if (M - (m4 + m5) == 0 && M + (m6 + m7) == 0) ++c;
            p->madd4_m[kk] = c; if (c > mmm) mmm = c;
        }
        p->maddmax4m = mmm;
        p->b4lm_m = xcalloc((size_t)(nm ? nm : 1) * (mmm ? mmm : 1) * 4, sizeof(int));
        for (int kk = 0; kk < nm; ++kk) {
            const int *q = p->b4lindex_m + 5 * kk;
            int c = 0;
            for (int M = -q[4]; M <= q[4]; ++M)
                for (int m4 = -q[0]; m4 <= q[0]; ++m4)
                    for (int m5 = -q[1]; m5 <= q[1]; ++m5)
                        for (int m6 = -q[2]; m6 <= q[2]; ++m6)
                            for (int m7 = -q[3]; m7 <= q[3]; ++m7)
                                if (M - (m4 + m5) == 0 && M + (m6 + m7) == 0) {
                                    int *e = p->b4lm_m + 4 * (kk * mmm + c++);
                                    e[0] = m4; e[1] = m5; e[2] = m6; e[3] = m7;
                                }
        }
    }
    return 0;
}

/* ----------------------------------------------------------------- orders 5 and 6 */
/* bcreate rank 5, coupling ((l1 l2)L12 (l3 l4)L34) l5 (util.F90:4960-5070): (L12, L34) pairs
 * into out[2*i]; returns their number, 0 when the tuple is no basis function */
static int bcreate5(const orc_pot *p, const int *vecnr, const int *vecl, int *out) {
    int veck[5], lmx = 0, n = 0;
    for (int i = 0; i < 5; ++i) {
        veck[i] = ((1) * (p->lmax + 1) + vecl[i]) * (p->nradial + 1) + vecnr[i];
        if (vecl[i] > lmx) lmx = vecl[i];
    }
    const int *l = vecl;
    for (int a = 0; a <= 2 * lmx; ++a)
        for (int b = 0; b <= 2 * lmx; ++b) {
            if (a > l[0] + l[1] || b > l[2] + l[3]) continue;
            if (a < abs(l[0] - l[1]) || b < abs(l[2] - l[3])) continue;
            if (l[4] > a + b || l[4] < abs(a - b)) continue;
            if (veck[0] == veck[1] && (a % 2) != 0) continue;
            if (veck[2] == veck[3] && (b % 2) != 0) continue;
            if (veck[0] == veck[2] && veck[1] == veck[3] && a == b && (l[4] % 2) != 0) continue;
            out[2 * n] = a; out[2 * n + 1] = b; ++n;
        }
    if ((l[0] + l[1] + l[2] + l[3] + l[4]) % 2 != 0) n = 0;
    return n;
}

/* bcreate rank 6, ((l1 l2)L12 (l3 l4)L34)L56 (l5 l6)L56 (util.F90:5140-5230) */
static int bcreate6(const orc_pot *p, const int *vecnr, const int *vecl, int *out) {
    int veck[6], lmx = 0, n = 0;
    for (int i = 0; i < 6; ++i) {
        veck[i] = ((1) * (p->lmax + 1) + vecl[i]) * (p->nradial + 1) + vecnr[i];
        if (vecl[i] > lmx) lmx = vecl[i];
    }
    const int *l = vecl;
    for (int a = 0; a <= 2 * lmx; ++a)
        for (int b = 0; b <= 2 * lmx; ++b)
            for (int c = 0; c <= 2 * lmx; ++c) {
                if (a > l[0] + l[1] || b > l[2] + l[3] || c > l[4] + l[5]) continue;
                if (a < abs(l[0] - l[1]) || b < abs(l[2] - l[3]) || c < abs(l[4] - l[5])) continue;
                if (c > a + b || c < abs(a - b)) continue;
                if (veck[0] == veck[1] && (a % 2) != 0) continue;
                if (veck[2] == veck[3] && (b % 2) != 0) continue;
                if (veck[4] == veck[5] && (c % 2) != 0) continue;
                if (veck[0] == veck[2] && veck[1] == veck[3] && a == b && (c % 2) != 0) continue;
                out[3 * n] = a; out[3 * n + 1] = b; out[3 * n + 2] = c; ++n;
            }
    if ((l[0] + l[1] + l[2] + l[3] + l[4] + l[5]) % 2 != 0) n = 0;
    return n;
}

/* next index tuple in the reference's loop order (last digit fastest) */
static int next_tuple(int *d, const int *lo, const int *hi, int n) {
    for (int i = n - 1; i >= 0; --i) {
        if (++d[i] <= hi[i]) return 1;
        d[i] = lo[i];
    }
    return 0;
}

/* the exchange of internal couplings inside one group, first entry excluded
 * (util.F90:3077-3091, 3713-3755): columns [c0, c1) of rows head+1 .. head+nint-1 */
static void sort_lints(int *idx, int width, int head, int nint, int c0, int c1) {
    for (int m = 1; m <= nint - 2; ++m)
        for (int k = m + 1; k <= nint - 1; ++k) {
            int *a = idx + (size_t)width * (head + m), *b = idx + (size_t)width * (head + k), s1 = 0, s2 = 0;
            for (int c = c0; c < c1; ++c) { s1 += a[c]; s2 += b[c]; }
            if (s1 > s2)
                for (int c = c0; c < c1; ++c) { int t = a[c]; a[c] = b[c]; b[c] = t; }
        }
}

static int grow(int **v, int *cap, int n, int width) {
    if (n < *cap) return 0;
    *cap = *cap ? 2 * *cap : 256;
    int *q = realloc(*v, (size_t)*cap * width * sizeof(int));
    if (!q) return -1;
    *v = q;
    return 0;
}

static int enumerate5(orc_pot *p) {
    const orc_header *h = &p->h;
    int lo[18], hi[18], d[18];
    for (int i = 0; i < 18; ++i) {
        lo[i] = i < 9 ? 1 : 0;
        hi[i] = i < 4 ? h->nr5_a : i < 9 ? h->nr5_m : i < 13 ? h->l5a : h->l5m;
        if (hi[i] < lo[i]) return 0;
        d[i] = lo[i];
    }
    const int kmm = h->nr5_m * (h->l5m + 1);
    int cap = 0, n = 0;
    do {
        int nc[4], ka[4];
        for (int i = 0; i < 4; ++i) {
            ka[i] = h->nr5_a * d[9 + i] + d[i];
            nc[i] = (kmm + 1) * ka[i] + (h->nr5_m * d[14 + i] + d[5 + i]);
        }
        if (nc[0] < nc[1] || nc[1] < nc[2] || nc[2] < nc[3]) continue;
        if (ka[0] < ka[1] || ka[1] < ka[2] || ka[2] < ka[3]) {
            snprintf(g_err, sizeof g_err, "Error! the atomic basis is not ordered lexicographically");
            return -1;
        }
        int vnr[4] = {d[0], d[1], d[2], d[3]}, vl[4] = {d[9], d[10], d[11], d[12]}, na, aa[64], iso, isb;
        bcreate(p, 4, vnr, vl, &na, aa, &iso, &isb);
        if (!iso) { snprintf(g_err, sizeof g_err, "Something wrong with rank = 4."); return -1; }
        int vnrm[5] = {d[4], d[5], d[6], d[7], d[8]}, vlm[5] = {d[13], d[14], d[15], d[16], d[17]}, am[2 * 1024];
        int nm = bcreate5(p, vnrm, vlm, am);
        if (!isb || nm == 0) continue;
        if (na != nm) {
            snprintf(g_err, sizeof g_err, "5-body basis pairs %d atomic with %d magnetic couplings "
                     "(the reference fills the rest with zero indices)", na, nm);
            return -1;
        }
        const int head = n;
        for (int t = 0; t < na; ++t) {
            if (grow(&p->b5index, &cap, n, 21)) return -1;
            int *b = p->b5index + 21 * n++;
            for (int i = 0; i < 18; ++i) b[i] = d[i];
            b[18] = aa[t]; b[19] = am[2 * t]; b[20] = am[2 * t + 1];
        }
        sort_lints(p->b5index, 21, head, na, 19, 21);
    } while (next_tuple(d, lo, hi, 18));
    p->nvar5 = n;
    return 0;
}

static int enumerate6(orc_pot *p) {
    const orc_header *h = &p->h;
    int lo[22], hi[22], d[22];
    for (int i = 0; i < 22; ++i) {
        lo[i] = i < 11 ? 1 : 0;
        hi[i] = i < 5 ? h->nr6_a : i < 11 ? h->nr6_m : i < 16 ? h->l6a : h->l6m;
        if (hi[i] < lo[i]) return 0;
        d[i] = lo[i];
    }
    /* the ordering keys use the 5-body maxima, as the reference's do (util.F90:3493-3494) */
    const int kmm = h->nr5_m * (h->l5m + 1);
    int cap = 0, n = 0;
    do {
        int nc[5], ka[5];
        for (int i = 0; i < 5; ++i) {
            ka[i] = h->nr6_a * d[11 + i] + d[i];
            nc[i] = (kmm + 1) * ka[i] + (h->nr6_m * d[17 + i] + d[6 + i]);
        }
        if (nc[0] < nc[1] || nc[1] < nc[2] || nc[2] < nc[3] || nc[3] < nc[4]) continue;
        if (ka[0] < ka[1] || ka[1] < ka[2] || ka[2] < ka[3] || ka[3] < ka[4]) {
            snprintf(g_err, sizeof g_err, "Error! the atomic basis is not ordered lexicographically");
            return -1;
        }
        int vnr[5] = {d[0], d[1], d[2], d[3], d[4]}, vl[5] = {d[11], d[12], d[13], d[14], d[15]}, aa[2 * 1024];
        int na = bcreate5(p, vnr, vl, aa);
        int vnrm[6] = {d[5], d[6], d[7], d[8], d[9], d[10]}, vlm[6] = {d[16], d[17], d[18], d[19], d[20], d[21]};
        int am[3 * 4096];
        int nm = bcreate6(p, vnrm, vlm, am);
        if (na == 0 || nm == 0) continue;
        if (na != nm) {
            snprintf(g_err, sizeof g_err, "6-body basis pairs %d atomic with %d magnetic couplings "
                     "(the reference fills the rest with zero indices)", na, nm);
            return -1;
        }
        const int head = n;
        for (int t = 0; t < na; ++t) {
            if (grow(&p->b6index, &cap, n, 27)) return -1;
            int *b = p->b6index + 27 * n++;
            for (int i = 0; i < 22; ++i) b[i] = d[i];
            b[22] = aa[2 * t]; b[23] = aa[2 * t + 1];
            b[24] = am[3 * t]; b[25] = am[3 * t + 1]; b[26] = am[3 * t + 2];
        }
        sort_lints(p->b6index, 27, head, na, 22, 24);
        sort_lints(p->b6index, 27, head, na, 24, 27);
    } while (next_tuple(d, lo, hi, 22));
    p->nvar6 = n;
    return 0;
}

/* per function: the atomic m-list (util.F90:3285-3312 / 3945-3975) times the magnetic one
 * (:3316-3350 / :3979-4015), each entry with the product of the couplings of
 * setuphighercouplings (:4392-4455 / :4457-4548) */
static void mlists56(orc_pot *p) {
    int cap5 = 0, n5 = 0, cap6 = 0, n6 = 0;
    p->m5ptr = xcalloc((size_t)p->nvar5 + 1, sizeof(int));
    for (int c = 0; c < p->nvar5; ++c) {
        const int *b = p->b5index + 21 * c;
        const int *l = b + 9, *lp = b + 13, L = b[18], L01 = b[19], L23 = b[20];
        for (int M = -L; M <= L; ++M)
        for (int m1 = -l[0]; m1 <= l[0]; ++m1)
        for (int m2 = -l[1]; m2 <= l[1]; ++m2)
        for (int m3 = -l[2]; m3 <= l[2]; ++m3)
        for (int m4 = -l[3]; m4 <= l[3]; ++m4) {
            if (M - (m1 + m2) != 0 || M + (m3 + m4) != 0) continue;
            double cfa = cg_racah(p->fac, l[0], l[1], L, m1, m2, M) * cg_racah(p->fac, l[2], l[3], L, m3, m4, -M) *
                         cg_racah(p->fac, L, L, 0, M, -M, 0);
            for (int M01 = -L01; M01 <= L01; ++M01)
            for (int M23 = -L23; M23 <= L23; ++M23)
            for (int q0 = -lp[0]; q0 <= lp[0]; ++q0)
            for (int q1 = -lp[1]; q1 <= lp[1]; ++q1)
            for (int q2 = -lp[2]; q2 <= lp[2]; ++q2)
            for (int q3 = -lp[3]; q3 <= lp[3]; ++q3)
            for (int q4 = -lp[4]; q4 <= lp[4]; ++q4) {
                if (M01 - (q0 + q1) != 0 || M23 - (q2 + q3) != 0 || -q4 - (M01 + M23) != 0) continue;
                double cfm = cg_racah(p->fac, lp[0], lp[1], L01, q0, q1, M01) *
                             cg_racah(p->fac, lp[2], lp[3], L23, q2, q3, M23) *
                             cg_racah(p->fac, L01, L23, lp[4], M01, M23, -q4) *
                             cg_racah(p->fac, lp[4], lp[4], 0, -q4, q4, 0);
                if (grow(&p->m5lm, &cap5, n5, 9)) return;
                int *e = p->m5lm + 9 * n5;
                e[0] = m1; e[1] = m2; e[2] = m3; e[3] = m4; e[4] = q0; e[5] = q1; e[6] = q2; e[7] = q3; e[8] = q4;
                p->m5cf = realloc(p->m5cf, (size_t)cap5 * sizeof(double));
                p->m5cf[n5++] = cfa * cfm;
            }
        }
        p->m5ptr[c + 1] = n5;
    }
    p->m6ptr = xcalloc((size_t)p->nvar6 + 1, sizeof(int));
    for (int c = 0; c < p->nvar6; ++c) {
        const int *b = p->b6index + 27 * c;
        const int *l = b + 11, *lp = b + 16, L12 = b[22], L34 = b[23], L01 = b[24], L23 = b[25], L45 = b[26];
        for (int M12 = -L12; M12 <= L12; ++M12)
        for (int M34 = -L34; M34 <= L34; ++M34)
        for (int m1 = -l[0]; m1 <= l[0]; ++m1)
        for (int m2 = -l[1]; m2 <= l[1]; ++m2)
        for (int m3 = -l[2]; m3 <= l[2]; ++m3)
        for (int m4 = -l[3]; m4 <= l[3]; ++m4)
        for (int m5 = -l[4]; m5 <= l[4]; ++m5) {
            if (M12 - (m1 + m2) != 0 || M34 - (m3 + m4) != 0 || -m5 - (M12 + M34) != 0) continue;
            double cfa = cg_racah(p->fac, l[0], l[1], L12, m1, m2, M12) * cg_racah(p->fac, l[2], l[3], L34, m3, m4, M34) *
                         cg_racah(p->fac, L12, L34, l[4], M12, M34, -m5) * cg_racah(p->fac, l[4], l[4], 0, -m5, m5, 0);
            for (int M01 = -L01; M01 <= L01; ++M01)
            for (int M23 = -L23; M23 <= L23; ++M23)
            for (int M45 = -L45; M45 <= L45; ++M45)
            for (int q0 = -lp[0]; q0 <= lp[0]; ++q0)
            for (int q1 = -lp[1]; q1 <= lp[1]; ++q1)
            for (int q2 = -lp[2]; q2 <= lp[2]; ++q2)
            for (int q3 = -lp[3]; q3 <= lp[3]; ++q3)
            for (int q4 = -lp[4]; q4 <= lp[4]; ++q4)
            for (int q5 = -lp[5]; q5 <= lp[5]; ++q5) {
                if (M01 - (q0 + q1) != 0 || M23 - (q2 + q3) != 0 || M45 - (q4 + q5) != 0 ||
                    -M45 - (M01 + M23) != 0) continue;
                double cfm = cg_racah(p->fac, lp[0], lp[1], L01, q0, q1, M01) *
                             cg_racah(p->fac, lp[2], lp[3], L23, q2, q3, M23) *
                             cg_racah(p->fac, lp[4], lp[5], L45, q4, q5, M45) *
                             cg_racah(p->fac, L01, L23, L45, M01, M23, -M45) *
                             cg_racah(p->fac, L45, L45, 0, -M45, M45, 0);
                if (grow(&p->m6lm, &cap6, n6, 11)) return;
                int *e = p->m6lm + 11 * n6;
                e[0] = m1; e[1] = m2; e[2] = m3; e[3] = m4; e[4] = m5;
                e[5] = q0; e[6] = q1; e[7] = q2; e[8] = q3; e[9] = q4; e[10] = q5;
                p->m6cf = realloc(p->m6cf, (size_t)cap6 * sizeof(double));
                p->m6cf[n6++] = cfa * cfm;
            }
        }
        p->m6ptr[c + 1] = n6;
    }
}

/* ----------------------------------------------------------------- couplings */
static void couplings(orc_pot *p) {
    const orc_header *h = &p->h;
    int L = p->lmax, W = 2 * L + 1;
    p->cg01 = xcalloc((size_t)(h->l2m + 1) * W, sizeof(double));
    for (int l = 0; l <= h->l2m; ++l)
        for (int m = -l; m <= l; ++m)
            p->cg01[l * W + m + L] = cg_racah(p->fac, l, l, 0, m, -m, 0);          /* :4277-4286 */
    p->cg02a = xcalloc((size_t)(h->l3a + 1) * W, sizeof(double));
    for (int l = 0; l <= h->l3a; ++l)
        for (int m = -l; m <= l; ++m)
            p->cg02a[l * W + m + L] = cg_racah(p->fac, l, l, 0, m, -m, 0);         /* :4292-4308 */
    p->cg02m = xcalloc((size_t)(p->nvar3l_m ? p->nvar3l_m : 1) * (p->maddmax3m ? p->maddmax3m : 1), sizeof(double));
    for (int k = 0; k < p->nvar3l_m; ++k) {
        const int *q = p->b3lindex_m + 3 * k;
        for (int n = 0; n < p->madd3_m[k]; ++n) {
            const int *e = p->b3lm_m + 3 * (k * p->maddmax3m + n);
            p->cg02m[k * p->maddmax3m + n] =
                cg_racah(p->fac, q[0], q[1], q[2], e[0], e[1], -e[2]) *
                cg_racah(p->fac, q[2], q[2], 0, e[2], -e[2], 0);                     /* :4310-4333 */
        }
    }
    p->cg03a = xcalloc((size_t)(p->nvar4l_a ? p->nvar4l_a : 1) * (p->maddmax4a ? p->maddmax4a : 1), sizeof(double));
    for (int k = 0; k < p->nvar4l_a; ++k) {
        const int *q = p->b4lindex_a + 3 * k;
        for (int n = 0; n < p->madd4_a[k]; ++n) {
            const int *e = p->b4lm_a + 3 * (k * p->maddmax4a + n);
            p->cg03a[k * p->maddmax4a + n] =
                cg_racah(p->fac, q[0], q[1], q[2], e[0], e[1], -e[2]) *
                cg_racah(p->fac, q[2], q[2], 0, e[2], -e[2], 0);                     /* :4343-4364 */
        }
    }
    p->cg03m = xcalloc((size_t)(p->nvar4l_m ? p->nvar4l_m : 1) * (p->maddmax4m ? p->maddmax4m : 1), sizeof(double));
    for (int k = 0; k < p->nvar4l_m; ++k) {
        const int *q = p->b4lindex_m + 5 * k;
        for (int n = 0; n < p->madd4_m[k]; ++n) {
            const int *e = p->b4lm_m + 4 * (k * p->maddmax4m + n);
            int M = e[0] + e[1];
            p->cg03m[k * p->maddmax4m + n] =
                cg_racah(p->fac, q[0], q[1], q[4], e[0], e[1], M) *
                cg_racah(p->fac, q[2], q[3], q[4], e[2], e[3], -M) *
                cg_racah(p->fac, q[4], q[4], 0, M, -M, 0);                            /* :4366-4388 */
        }
    }
}

orc_pot *orc_create(const orc_header *h) {
    orc_pot *p = xcalloc(1, sizeof(orc_pot));
    p->h = *h;
    if (h->ndensity < 1 || h->ndensity > 2) FAIL("ndensity must be 1 or 2 (util.F90 embedding)");
    /* limit checks (util.F90:233-272) */
    if (h->l3a < h->l4a || h->l3a < h->l5a || h->l3a < h->l6a) FAIL(" l3max_atomic needs to be largest.");
    if (h->l2m < h->l3m || h->l2m < h->l4m || h->l2m < h->l5m || h->l2m < h->l6m)
        FAIL(" l2max_magnetic needs to be largest.");
    if (h->nr3_a < h->nr4_a || h->nr3_a < h->nr5_a || h->nr3_a < h->nr6_a) FAIL(" nradial_atomic needs to be largest.");
    if (h->nr2_m < h->nr3_m || h->nr2_m < h->nr4_m || h->nr2_m < h->nr5_m || h->nr2_m < h->nr6_m)
        FAIL(" nradial_magnetic needs to be largest.");
    if (2 * h->l6a > h->l3a || 2 * h->l6m > h->l2m) FAIL("6-body couplings not worked for lmax > 6");
    /* restated here: A with the magnetic indices of order 3 only (the reference's force pass
     * differs from its energy's gradient beyond, and reads outside A for radial indices
     * beyond); 6-body terms only with 5-body terms (prepB skips them otherwise) */
    if (h->l4m > h->l3m || h->l5m > h->l3m || h->l6m > h->l3m) FAIL("magnetic l of orders 4-6 above l3max_magnetic");
    if (h->nr4_m > h->nr3_m || h->nr5_m > h->nr3_m || h->nr6_m > h->nr3_m)
        FAIL("magnetic radial functions of orders 4-6 above nradial3_magnetic");
    if ((h->nr6_a || h->nr6_m) && !h->nr5_a && !h->nr5_m) FAIL("6-body terms without 5-body terms");
    if (h->nr3_a > h->nb_a) FAIL(" nradial_atomic < nradbase_atomic required");
    if (h->nr2_m > h->nb_m) FAIL(" nradial_magnetic < nradbase_magnetic required");
    p->lmax = h->l3a > h->l2m ? h->l3a : h->l2m;
    p->nradial = h->nr3_a > h->nr2_m ? h->nr3_a : h->nr2_m;
    if (p->lmax > 16) FAIL("lmax too large");
    p->lsa = (h->l3a + 1) * (h->l3a + 1);
    p->lsm = (h->l2m + 1) * (h->l2m + 1);
    p->lha = (h->l3a + 1) * (h->l3a + 2) / 2;
    p->lhm = (h->l2m + 1) * (h->l2m + 2) / 2;
    {
        double x = 1.0; p->fac[0] = 1.0;
        for (int l = 0; l < 8 * p->lmax + 1; ++l) { x *= (double)(l + 1); p->fac[l + 1] = x; }
    }
    if (enumerate(p)) goto fail;
    if (h->nr5_a > 0 && h->nr5_m > 0 && enumerate5(p)) goto fail;
    if (h->nr6_a > 0 && h->nr6_m > 0 && enumerate6(p)) goto fail;
    couplings(p);
    mlists56(p);
    /* parameter defaults (util.F90:282-332) */
    int nd = h->ndensity;
    p->c0 = xcalloc((size_t)p->nvar1 * nd, sizeof(double));
    p->c2 = xcalloc((size_t)p->nvar2 * nd, sizeof(double));
    p->c3 = xcalloc((size_t)p->nvar3 * nd, sizeof(double));
    p->c4 = xcalloc((size_t)p->nvar4 * nd, sizeof(double));
    p->c5 = xcalloc((size_t)p->nvar5 * nd + 1, sizeof(double));
    p->c6 = xcalloc((size_t)p->nvar6 * nd + 1, sizeof(double));
    p->crad_a = xcalloc((size_t)h->nr3_a * (h->l3a + 1) * h->nb_a, sizeof(double));
    for (int m = 0; m < h->nr3_a; ++m)
        for (int l = 0; l <= h->l3a; ++l) p->crad_a[(m * (h->l3a + 1) + l) * h->nb_a + m] = 1.0;
    p->crad_m = xcalloc((size_t)h->nr2_m * (h->l2m + 1) * h->nb_m, sizeof(double));
    for (int m = 0; m < h->nr2_m; ++m)
        for (int l = 0; l <= h->l2m; ++l) p->crad_m[(m * (h->l2m + 1) + l) * h->nb_m + m] = 1.0;
    p->core[0] = p->core[1] = 3.0;
    for (int k = 0; k < 2; ++k) for (int d = 0; d < 2; ++d) p->p1[k][d] = 1.0;
    return p;
fail:
    orc_free(p);
    return NULL;
}

void orc_free(orc_pot *p) {
    if (!p) return;
    free(p->b1index); free(p->b2index); free(p->b3index); free(p->b4index);
    free(p->b3lindex_m); free(p->madd3_m); free(p->b3lm_m); free(p->b3map_a); free(p->b3map_m);
    free(p->b4lindex_a); free(p->madd4_a); free(p->b4lm_a); free(p->b4lindex_m); free(p->madd4_m);
    free(p->b4lm_m); free(p->b4map_a); free(p->b4map_m);
    free(p->cg01); free(p->cg02a); free(p->cg02m); free(p->cg03a); free(p->cg03m);
    free(p->c0); free(p->c2); free(p->c3); free(p->c4); free(p->crad_a); free(p->crad_m);
    free(p->b5index); free(p->b6index); free(p->m5ptr); free(p->m5lm); free(p->m5cf);
    free(p->m6ptr); free(p->m6lm); free(p->m6cf); free(p->c5); free(p->c6);
    free(p->lut_ga); free(p->lut_fa); free(p->lut_hc); free(p->lut_gm); free(p->lut_fm);
    free(p);
}

void orc_counts(const orc_pot *p, int nvar[6]) {
    nvar[0] = p->nvar1; nvar[1] = p->nvar2; nvar[2] = p->nvar3; nvar[3] = p->nvar4;
    nvar[4] = p->nvar5; nvar[5] = p->nvar6;
}

int orc_basis_index(const orc_pot *p, int order, int k, int *idx) {
    switch (order) {
    case 1: if (k < 0 || k >= p->nvar1) return -1; idx[0] = p->b1index[k]; return 1;
    case 2: if (k < 0 || k >= p->nvar2) return -1; memcpy(idx, p->b2index + 4 * k, 4 * sizeof(int)); return 4;
    case 3: if (k < 0 || k >= p->nvar3) return -1; memcpy(idx, p->b3index + 9 * k, 9 * sizeof(int)); return 9;
    case 4: if (k < 0 || k >= p->nvar4) return -1; memcpy(idx, p->b4index + 15 * k, 15 * sizeof(int)); return 15;
    case 5: if (k < 0 || k >= p->nvar5) return -1; memcpy(idx, p->b5index + 21 * k, 21 * sizeof(int)); return 21;
    case 6: if (k < 0 || k >= p->nvar6) return -1; memcpy(idx, p->b6index + 27 * k, 27 * sizeof(int)); return 27;
    }
    return -1;
}

/* ----------------------------------------------------------------- radial (radial.F90) */
static void calc_cheb(int n, double x, double *c1, double *dc1) {   /* :6-33 */
    double c2[64];
    double tx = 2.0 * x;
    c1[0] = 1.0; dc1[0] = 0.0; c2[0] = 1.0;
    if (n > 0) {
        c1[1] = x; c2[1] = tx;
        for (int m = 1; m <= n - 1; ++m) {
            c1[m + 1] = tx * c1[m] - c1[m - 1];
            c2[m + 1] = tx * c2[m] - c2[m - 1];
        }
        for (int m = 1; m <= n; ++m) dc1[m] = (double)m * c2[m - 1];
    }
}

void orc_radbase(int radtype, double lam, int nb, double cut, double dcut, double r,
                 double *gr, double *dgr) {                           /* :430-606 */
    double cheb[64], dcheb[64];
    if (r >= cut) {
        for (int n = 0; n < nb; ++n) { gr[n] = 0.0; dgr[n] = 0.0; }
        return;
    }
    double x, dx, x0, y0, y1, dy1, y2, env, denv;
    switch (radtype) {
    case 3:
    case 4:
        x0 = r / cut;
        y1 = exp(-lam * x0);
        dy1 = -lam * y1 / cut;
        y2 = exp(-lam);
        x = 1.0 - 2.0 * (y1 - y2) / (1.0 - y2);
        dx = -2.0 * dy1 / (1.0 - y2);
        calc_cheb(nb - 1, x, cheb, dcheb);
        gr[0] = cheb[0]; dgr[0] = dcheb[0] * dx;
        for (int n = 2; n <= nb; ++n) {
            gr[n - 1] = 0.5 - 0.5 * cheb[n - 1];
            dgr[n - 1] = -0.5 * dcheb[n - 1] * dx;
        }
        if (radtype == 3) {
            env = 0.5 * (1.0 + cos(PI * x0));
            denv = -0.5 * sin(PI * x0) * PI / cut;
            for (int n = 0; n < nb; ++n) {
                dgr[n] = gr[n] * denv + dgr[n] * env;
                gr[n] = gr[n] * env;
            }
        }
        break;
    case 5:
    case 6:
        x0 = r / cut;
        y0 = 1.0 - x0;
        dy1 = pow(y0, lam - 1.0);
        y1 = 1.0 - dy1 * y0;
        dy1 = lam / cut * dy1;
        x = 2.0 * y1 - 1.0;
        dx = 2.0 * dy1;
        if (radtype == 5) {
            calc_cheb(nb, x, cheb, dcheb);
            for (int n = 1; n <= nb; ++n) {
                gr[n - 1] = 0.5 - 0.5 * cheb[n];
                dgr[n - 1] = -0.5 * dcheb[n] * dx;
            }
        } else {
            calc_cheb(nb - 1, x, cheb, dcheb);
            gr[0] = cheb[0]; dgr[0] = dcheb[0] * dx;
            for (int n = 1; n <= nb - 1; ++n) {
                gr[n] = 0.5 - 0.5 * cheb[n];
                dgr[n] = -0.5 * dcheb[n] * dx;
            }
            env = 0.5 * (1.0 + cos(PI * x0));
            denv = -0.5 * sin(PI * x0) * PI / cut;
            for (int n = 0; n < nb; ++n) {
                dgr[n] = gr[n] * denv + dgr[n] * env;
                gr[n] = gr[n] * env;
            }
        }
        break;
    case 7:
        x0 = r / cut;
        x = 1.0 - 2.0 * x0 * x0;
        dx = -(4.0 / cut) * x0;
        calc_cheb(nb, x, cheb, dcheb);
        for (int n = 1; n <= nb; ++n) { gr[n - 1] = cheb[n]; dgr[n - 1] = dcheb[n] * dx; }
        break;
    case 8:
        x0 = r / cut;
        x = 1.0 - 2.0 * x0;
        dx = -2.0 / cut;
        calc_cheb(nb, x, cheb, dcheb);
        for (int n = 1; n <= nb; ++n) { gr[n - 1] = cheb[n] - 1.0; dgr[n - 1] = dcheb[n] * dx; }
        break;
    default:
        for (int n = 0; n < nb; ++n) { gr[n] = NAN; dgr[n] = NAN; }
        return;
    }
    dx = cut - dcut;                                                  /* :595-602 */
    if (r > dx) {
        double fcut = 0.5 * (1.0 + cos(PI * (r - dx) / dcut));
        double dfcut = -0.5 * sin(PI * (r - dx) / dcut) * PI / dcut;
        for (int n = 0; n < nb; ++n) {
            dgr[n] = gr[n] * dfcut + dgr[n] * fcut;
            gr[n] = gr[n] * fcut;
        }
    }
}

void orc_radcore(double r, double pre, double lam, double cut, double *cr, double *dcr) {
    if (r < cut) {                                                    /* :630-668 */
        pre = fabs(pre); lam = fabs(lam);
        double r2 = r * r, lr2 = lam * r2;
        if (lr2 < 50.0) {
            double y = exp(-lr2);
            *cr = pre * y / r;
            *dcr = -pre * y * (2.0 * lr2 + 1.0) / r2;
            double x0 = r / cut;
            double env = 0.5 * (1.0 + cos(PI * x0));
            double denv = -0.5 * sin(PI * x0) * PI / cut;
            *dcr = *cr * denv + *dcr * env;
            *cr = *cr * env;
        } else { *cr = 0.0; *dcr = 0.0; }
    } else { *cr = 0.0; *dcr = 0.0; }
}

static void radfunc(int nb, int nr, int lmax, const double *crad, const double *g,
                    const double *dg, double *fr, double *dfr) {     /* :610-626 */
    for (int l = 0; l <= lmax; ++l)
        for (int n = 0; n < nr; ++n) {
            const double *c = crad + (n * (lmax + 1) + l) * nb;
            double s = 0.0, ds = 0.0;
            for (int k = 0; k < nb; ++k) { s += c[k] * g[k]; ds += c[k] * dg[k]; }
            fr[n * (lmax + 1) + l] = s; dfr[n * (lmax + 1) + l] = ds;
        }
}

/* Hermite coefficients from node values walking n = nlut..1 (radial.F90:223-296) */
static void hermite_store(double *dst, double f0, double f0d1, double *f1, double *f1d1) {
    dst[0] = f0;
    dst[1] = f0d1;
    dst[2] = 3.0 * (*f1 - f0) - *f1d1 - 2.0 * f0d1;
    dst[3] = -2.0 * (*f1 - f0) + *f1d1 + f0d1;
    *f1 = dst[0]; *f1d1 = dst[1];
}

static void build_luts(orc_pot *p) {
    const orc_header *h = &p->h;
    int nb_a = h->nb_a, nfa = h->nr3_a * (h->l3a + 1);
    int nb_m = h->nb_m, nfm = h->nr2_m * (h->l2m + 1);
    p->rs_a = (double)NLUT / p->cutoff;
    p->rs_m = (double)NLUT / p->cutoff_m;
    double inv_a = 1.0 / p->rs_a, inv_m = 1.0 / p->rs_m;
    p->lut_ga = xcalloc((size_t)NLUT * (nb_a ? nb_a : 1) * 4, sizeof(double));
    p->lut_fa = xcalloc((size_t)NLUT * (nfa ? nfa : 1) * 4, sizeof(double));
    p->lut_hc = xcalloc((size_t)NLUT * 4, sizeof(double));
    p->lut_gm = xcalloc((size_t)NLUT * (nb_m ? nb_m : 1) * 4, sizeof(double));
    p->lut_fm = xcalloc((size_t)NLUT * (nfm ? nfm : 1) * 4, sizeof(double));
    double g[64], dg[64], fr[4096], dfr[4096];
    double f1g[64] = {0}, f1gd[64] = {0}, f1f[4096] = {0}, f1fd[4096] = {0}, f1c = 0, f1cd = 0;
    for (int n = NLUT; n >= 1; --n) {                                /* :223-296 */
        double x = inv_a * (double)n;
        orc_radbase(h->radtype_a, h->lambda_a, nb_a, h->cut_a, h->dcut_a, x, g, dg);
        radfunc(nb_a, h->nr3_a, h->l3a, p->crad_a, g, dg, fr, dfr);
        double cr, dcr;
        orc_radcore(x, p->core[0], p->core[1], h->cut_a, &cr, &dcr);
        for (int k = 0; k < nb_a; ++k)
            hermite_store(p->lut_ga + ((size_t)(n - 1) * nb_a + k) * 4, g[k], dg[k] * inv_a, &f1g[k], &f1gd[k]);
        /* (l outer, nr inner) in the reference; storage is per function, order irrelevant */
        for (int k = 0; k < nfa; ++k)
            hermite_store(p->lut_fa + ((size_t)(n - 1) * nfa + k) * 4, fr[k], dfr[k] * inv_a, &f1f[k], &f1fd[k]);
        hermite_store(p->lut_hc + (size_t)(n - 1) * 4, cr, dcr * inv_a, &f1c, &f1cd);
    }
    memset(f1g, 0, sizeof f1g); memset(f1gd, 0, sizeof f1gd);
    memset(f1f, 0, sizeof f1f); memset(f1fd, 0, sizeof f1fd);
    for (int n = NLUT; n >= 1; --n) {                                /* :364-417 */
        double x = inv_m * (double)n;
        orc_radbase(h->radtype_m, h->lambda_m, nb_m, h->cut_m, h->dcut_m, x, g, dg);
        radfunc(nb_m, h->nr2_m, h->l2m, p->crad_m, g, dg, fr, dfr);
        for (int k = 0; k < nb_m; ++k)
            hermite_store(p->lut_gm + ((size_t)(n - 1) * nb_m + k) * 4, g[k], dg[k] * inv_m, &f1g[k], &f1gd[k]);
        for (int k = 0; k < nfm; ++k)
            hermite_store(p->lut_fm + ((size_t)(n - 1) * nfm + k) * 4, fr[k], dfr[k] * inv_m, &f1f[k], &f1fd[k]);
    }
    p->have_luts = 1;
}

static inline void horner(const double *c, double wl, double rs, double *v, double *dv) {
    double wl2 = wl * wl, wl3 = wl2 * wl;
    *v = c[0] + c[1] * wl + c[2] * wl2 + c[3] * wl3;
    *dv = (c[1] + c[2] * (2.0 * wl) + c[3] * (3.0 * wl2)) * rs;
}

/* lookupRadspline_atomic (radial.F90:53-107); returns -1 on "very small distance" */
static int lookup_atomic(const orc_pot *p, double r, double *gr, double *dgr, double *fr,
                         double *dfr, double *cr, double *dcr) {
    const orc_header *h = &p->h;
    int nfa = h->nr3_a * (h->l3a + 1);
    double x = r * p->rs_a;
    int nl = (int)floor(x);
    if (nl <= 0) return -1;
    if (nl >= NLUT) {
        for (int k = 0; k < h->nb_a; ++k) gr[k] = dgr[k] = 0.0;
        for (int k = 0; k < nfa; ++k) fr[k] = dfr[k] = 0.0;
        *cr = *dcr = 0.0;
        return 0;
    }
    double wl = x - (double)nl;
    for (int k = 0; k < h->nb_a; ++k) horner(p->lut_ga + ((size_t)(nl - 1) * h->nb_a + k) * 4, wl, p->rs_a, &gr[k], &dgr[k]);
    for (int k = 0; k < nfa; ++k) horner(p->lut_fa + ((size_t)(nl - 1) * nfa + k) * 4, wl, p->rs_a, &fr[k], &dfr[k]);
    horner(p->lut_hc + (size_t)(nl - 1) * 4, wl, p->rs_a, cr, dcr);
    return 0;
}

/* lookupRadspline_magnetic (radial.F90:111-160).  nl==0 -> 1 (the reference's branch).
 * nl > nlut reads outside the reference table (undefined there); flagged here. */
static int lookup_magnetic(const orc_pot *p, double r, double *gr, double *dgr, double *fr,
                           double *dfr) {
    const orc_header *h = &p->h;
    int nfm = h->nr2_m * (h->l2m + 1);
    double x = r * p->rs_m;
    int nl = (int)floor(x);
    if (nl < 0) return -1;
    if (nl > NLUT) return -2;
    if (nl == 0) {
        for (int k = 0; k < h->nb_m; ++k) { gr[k] = 1.0; dgr[k] = 0.0; }
        for (int k = 0; k < nfm; ++k) { fr[k] = 1.0; dfr[k] = 0.0; }
        return 0;
    }
    double wl = x - (double)nl;
    for (int k = 0; k < h->nb_m; ++k) horner(p->lut_gm + ((size_t)(nl - 1) * h->nb_m + k) * 4, wl, p->rs_m, &gr[k], &dgr[k]);
    for (int k = 0; k < nfm; ++k) horner(p->lut_fm + ((size_t)(nl - 1) * nfm + k) * 4, wl, p->rs_m, &fr[k], &dfr[k]);
    return 0;
}

int orc_spline_eval(const orc_pot *p, int which, double r, double *val, double *dval) {
    const orc_header *h = &p->h;
    if (!p->have_luts) return -3;
    double g[64], dg[64], fr[4096], dfr[4096], cr, dcr;
    int rc;
    if (which <= 2) {
        rc = lookup_atomic(p, r, g, dg, fr, dfr, &cr, &dcr);
        if (rc) return rc;
        if (which == 0) { memcpy(val, g, h->nb_a * 8); memcpy(dval, dg, h->nb_a * 8); return h->nb_a; }
        if (which == 1) {
            int n = h->nr3_a * (h->l3a + 1);
            memcpy(val, fr, n * 8); memcpy(dval, dfr, n * 8); return n;
        }
        *val = cr; *dval = dcr; return 1;
    }
    rc = lookup_magnetic(p, r, g, dg, fr, dfr);
    if (rc) return rc;
    if (which == 3) { memcpy(val, g, h->nb_m * 8); memcpy(dval, dg, h->nb_m * 8); return h->nb_m; }
    int n = h->nr2_m * (h->l2m + 1);
    memcpy(val, fr, n * 8); memcpy(dval, dfr, n * 8); return n;
}
double orc_cutoff(const orc_pot *p, int which) { return which ? p->cutoff_m : p->cutoff; }

/* ----------------------------------------------------------------- Y_lm (spherical.F90) */
/* compute_ylm_cart with the shindex layout for lmax L (:159-282), arrays sized (L+1)^2 */
static void ylm_cart(int L, double rx, double ry, double rz, cplx *sh, cplx *dsh /*[3][(L+1)^2]*/) {
    int lsq = (L + 1) * (L + 1);
    double alm[160], blm[160], cl[20], dl[20], plm[160], dplm[160];
    for (int l = 1; l <= L; ++l) {                                   /* pre_compute :18-49 */
        double lsq_ = (double)(l * l), ld = (double)(2 * l);
        double l1 = 4.0 * lsq_ - 1.0, l2 = lsq_ - ld + 1.0;
        for (int m = 0; m <= l - 2; ++m) {
            double msq = (double)(m * m);
            int k = shx(L, l, m);
            alm[k] = sqrt(l1 / (lsq_ - msq));
            blm[k] = -sqrt((l2 - msq) / (4.0 * l2 - 1.0));
        }
        cl[l] = -sqrt(1.0 + 0.5 / (double)l);
        dl[l] = sqrt(2.0 * (double)(l - 1) + 3.0);
    }
    /* compute_barplm :55-155 */
    const double sq3 = 1.73205080756887719318, sq3o2 = 1.22474487139158894067;
    plm[shx(L, 0, 0)] = 1.0; dplm[shx(L, 0, 0)] = 0.0;
    if (L > 0) {
        plm[shx(L, 1, 0)] = sq3 * rz; dplm[shx(L, 1, 0)] = sq3;
        plm[shx(L, 1, 1)] = -sq3o2; dplm[shx(L, 1, 1)] = 0.0;
        for (int l = 2; l <= L; ++l) {
            for (int m = 0; m <= l - 2; ++m) {
                int k = shx(L, l, m), k1 = shx(L, l - 1, m), k2 = shx(L, l - 2, m);
                plm[k] = alm[k] * (rz * plm[k1] + blm[k] * plm[k2]);
                dplm[k] = alm[k] * (plm[k1] + rz * dplm[k1] + blm[k] * dplm[k2]);
            }
            int k = shx(L, l - 1, l - 1), k1 = shx(L, l, l - 1), k2 = shx(L, l, l);
            double t = dl[l] * plm[k];
            plm[k1] = t * rz; dplm[k1] = t;
            plm[k2] = cl[l] * plm[k]; dplm[k2] = 0.0;
        }
    }
    cplx *d0 = dsh, *d1 = dsh + lsq, *d2 = dsh + 2 * lsq;
    /* m = 0 */
    for (int l = 0; l <= L; ++l) {
        int k = shx(L, l, 0);
        sh[k] = plm[k];
        double dyz = dplm[k], rdy = dyz * rz;
        d0[k] = -rdy * rx; d1[k] = -rdy * ry; d2[k] = dyz - rdy * rz;
    }
    cplx phase = rx + I * ry;
    /* m = 1 */
    for (int l = 1; l <= L; ++l) {
        int k = shx(L, l, 1);
        sh[k] = phase * plm[k];
        double dyx_r = plm[k], dyy_i = plm[k];
        double dyz_r = creal(phase) * dplm[k], dyz_i = cimag(phase) * dplm[k];
        double rdy_r = rx * dyx_r + rz * dyz_r, rdy_i = ry * dyy_i + rz * dyz_i;
        d0[k] = (dyx_r - rdy_r * rx) - I * (rdy_i * rx);
        d1[k] = -rdy_r * ry + I * (dyy_i - rdy_i * ry);
        d2[k] = (dyz_r - rdy_r * rz) + I * (dyz_i - rdy_i * rz);
    }
    /* m > 1 */
    cplx phasem = phase;
    for (int m = 2; m <= L; ++m) {
        cplx mph = creal(phasem) * (double)m + I * cimag(phasem) * (double)m;
        phasem = phasem * phase;
        for (int l = m; l <= L; ++l) {
            int k = shx(L, l, m);
            sh[k] = creal(phasem) * plm[k] + I * cimag(phasem) * plm[k];
            cplx dyx = mph * plm[k];
            cplx dyy = -cimag(dyx) + I * creal(dyx);
            cplx dyz = phasem * dplm[k];
            double rdy_r = rx * creal(dyx) + ry * creal(dyy) + rz * creal(dyz);
            double rdy_i = rx * cimag(dyx) + ry * cimag(dyy) + rz * cimag(dyz);
            d0[k] = (creal(dyx) - rdy_r * rx) + I * (cimag(dyx) - rdy_i * rx);
            d1[k] = (creal(dyy) - rdy_r * ry) + I * (cimag(dyy) - rdy_i * ry);
            d2[k] = (creal(dyz) - rdy_r * rz) + I * (cimag(dyz) - rdy_i * rz);
        }
    }
    /* m < 0 fill :263-280 */
    for (int l = 1; l <= L; ++l)
        for (int m = 1; m <= l; ++m) {
            double ph = sgnm(m);
            int k = shx(L, l, -m), k1 = shx(L, l, m);
            sh[k] = conj(sh[k1]) * ph;
            d0[k] = conj(d0[k1]) * ph; d1[k] = conj(d1[k1]) * ph; d2[k] = conj(d2[k1]) * ph;
        }
}

void orc_ylm(int lmax, double x, double y, double z, double *yre, double *yim, double *dyre,
             double *dyim) {
    int lsq = (lmax + 1) * (lmax + 1);
    cplx sh[300], dsh[900];
    ylm_cart(lmax, x, y, z, sh, dsh);
    for (int l = 0; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m) {
            int o = l * l + l + m, k = shx(lmax, l, m);
            yre[o] = creal(sh[k]); yim[o] = cimag(sh[k]);
            for (int c = 0; c < 3; ++c) {
                dyre[c * lsq + o] = creal(dsh[c * lsq + k]);
                dyim[c * lsq + o] = cimag(dsh[c * lsq + k]);
            }
        }
}

/* ----------------------------------------------------------------- parser (util.F90:459-1108) */
static int tokenize(char *line, char **tok, int maxt) {
    int n = 0;
    char *s = line;
    while (*s && n < maxt) {
        while (*s && (isspace((unsigned char)*s) || *s == ',')) ++s;
        if (!*s) break;
        tok[n++] = s;
        while (*s && !isspace((unsigned char)*s) && *s != ',') ++s;
        if (*s) *s++ = '\0';
    }
    return n;
}

typedef struct { FILE *f; char buf[8192]; char *tok[128]; int n; } rdr;
static int rd_line(rdr *r) {
    for (;;) {
        if (!fgets(r->buf, sizeof r->buf, r->f)) return -1;
        r->n = tokenize(r->buf, r->tok, 128);
        return r->n;
    }
}
static double fnum(const char *s) {
    char tmp[64]; size_t n = strlen(s); if (n > 63) n = 63;
    for (size_t i = 0; i < n; ++i) tmp[i] = (s[i] == 'd' || s[i] == 'D') ? 'e' : s[i];
    tmp[n] = 0;
    return strtod(tmp, NULL);
}

static orc_pot *load_impl(const char *path, const char *element, double rcut, double rcut_m,
                          double smear, int xc) {
    rdr r; r.f = fopen(path, "r");
    orc_pot *p = NULL;
    orc_header h; memset(&h, 0, sizeof h);
    const char *xcname = xc == 1 ? "pbe" : xc == 2 ? "lda" : NULL;
    if (!xcname) { snprintf(g_err, sizeof g_err, "Unknown XC functional received by acelibini."); goto fail0; }
    if (!r.f) { snprintf(g_err, sizeof g_err, "File not found: %s", path); goto fail0; }
    if (rd_line(&r) < 0) goto bad;                               /* line 1 ignored */
    if (rd_line(&r) < 7) goto bad;                               /* created y m d h m s */
    if (rd_line(&r) < 2) goto bad;
    if (strcmp(r.tok[1], element)) { snprintf(g_err, sizeof g_err, "Inconsistent element information in file: %s", element); goto fail0; }
    if (rd_line(&r) < 2) goto bad;
    if (strcmp(r.tok[1], xcname)) { snprintf(g_err, sizeof g_err, "Inconsistent XC functional in file: %s", element); goto fail0; }
    if (rd_line(&r) < 2) goto bad;
    if (fabs(smear - fnum(r.tok[1])) > 1e-6) { snprintf(g_err, sizeof g_err, "Inconsistent gauss-smearing in file: %s", element); goto fail0; }
    if (rd_line(&r) < 3) goto bad;
    h.lambda_a = fnum(r.tok[1]); h.lambda_m = fnum(r.tok[2]);
    if (rd_line(&r) < 10) goto bad;
    h.cut_a = fnum(r.tok[4]); h.cut_m = fnum(r.tok[5]); h.dcut_a = fnum(r.tok[6]);
    h.dcut_m = fnum(r.tok[7]); h.ecut = fnum(r.tok[8]); h.decut = fnum(r.tok[9]);
    int *rows[5][4] = {{&h.nb_a, &h.nb_m, &h.nr2_m, &h.l2m}, {&h.nr3_a, &h.nr3_m, &h.l3a, &h.l3m},
                       {&h.nr4_a, &h.nr4_m, &h.l4a, &h.l4m}, {&h.nr5_a, &h.nr5_m, &h.l5a, &h.l5m},
                       {&h.nr6_a, &h.nr6_m, &h.l6a, &h.l6m}};
    for (int i = 0; i < 5; ++i) {
        if (rd_line(&r) < 6) goto bad;
        for (int j = 0; j < 4; ++j) *rows[i][j] = atoi(r.tok[2 + j]);
    }
    if (rd_line(&r) < 2) goto bad;
    h.ndensity = atoi(r.tok[1]);
    if (rd_line(&r) < 5) goto bad;
    h.npot = atoi(r.tok[1]); h.nfemb = atoi(r.tok[2]);
    h.radtype_a = atoi(r.tok[3]); h.radtype_m = atoi(r.tok[4]);

    p = orc_create(&h);
    if (!p) goto fail0;
    p->cutoff = rcut; p->cutoff_m = rcut_m;
    int nd_max = h.ndensity;
    /* parameter records */
    for (;;) {
        int n = rd_line(&r);
        if (n < 0) break;
        if (n < 3) continue;
        char ptype[32]; snprintf(ptype, sizeof ptype, "%s", r.tok[0]);
        int norder = atoi(r.tok[1]), nd = atoi(r.tok[2]);
        if (!strcmp(ptype, "clu")) {
            int nt = rd_line(&r);
            if (nt < 0) break;
            if (norder < 1 || norder > 6) continue;
            if (nt < norder + 1) continue;
            /* elecheck for orders >= 2 (util.F90:711-719); order 1 simply does not match */
            int elok = 1;
            for (int e = 0; e < norder; ++e) if (strcmp(r.tok[e], element)) elok = 0;
            if (!elok && norder >= 2) { snprintf(g_err, sizeof g_err, "Read incorrect elements for interaction of order: %d", norder); goto fail; }
            if (!elok) continue;
            if (nd > nd_max || nd < 1) continue;
            int iv[40]; int ni = nt - norder - 2;
            if (ni > 40) ni = 40;
            for (int i = 0; i < ni; ++i) iv[i] = atoi(r.tok[norder + i]);
            double xclu = fnum(r.tok[nt - 2]);
            if (norder == 1 && ni >= 1) {
                for (int k = 0; k < p->nvar1; ++k) if (p->b1index[k] == iv[0]) p->c0[k * nd_max + nd - 1] = xclu;
            } else if (norder == 2 && ni >= 4) {
                for (int k = 0; k < p->nvar2; ++k)
                    if (!memcmp(p->b2index + 4 * k, iv, 4 * sizeof(int))) p->c2[k * nd_max + nd - 1] = xclu;
            } else if (norder == 3 && ni >= 9) {
                for (int k = 0; k < p->nvar3; ++k)
                    if (!memcmp(p->b3index + 9 * k, iv, 9 * sizeof(int))) p->c3[k * nd_max + nd - 1] = xclu;
            } else if (norder == 4 && ni >= 15) {
                for (int k = 0; k < p->nvar4; ++k)
                    if (!memcmp(p->b4index + 15 * k, iv, 15 * sizeof(int))) p->c4[k * nd_max + nd - 1] = xclu;
            } else if (norder == 5 && ni >= 21) {
                for (int k = 0; k < p->nvar5; ++k)
                    if (!memcmp(p->b5index + 21 * k, iv, 21 * sizeof(int))) p->c5[k * nd_max + nd - 1] = xclu;
            } else if (norder == 6 && ni >= 27) {
                for (int k = 0; k < p->nvar6; ++k)
                    if (!memcmp(p->b6index + 27 * k, iv, 27 * sizeof(int))) p->c6[k * nd_max + nd - 1] = xclu;
            }
        } else if (!strcmp(ptype, "par")) {
            if (norder < 1 || norder > 2) { snprintf(g_err, sizeof g_err, "norder < nvarp1 required for par."); goto fail; }
            int nt = rd_line(&r);
            if (nt < 0) break;
            if (nt < 2) continue;
            if (!strcmp(r.tok[0], element) && nd >= 1 && nd <= nd_max) p->p1[norder - 1][nd - 1] = fnum(r.tok[1]);
        } else if (!strcmp(ptype, "rad_atomic") || !strcmp(ptype, "rad_magnetic")) {
            int nt = rd_line(&r);
            if (nt < 0) break;
            if (nt < 6) continue;
            if (strcmp(r.tok[0], element) || strcmp(r.tok[1], element)) { snprintf(g_err, sizeof g_err, "Error when reading radial parameters."); goto fail; }
            int a = atoi(r.tok[2]), b = atoi(r.tok[3]), l = atoi(r.tok[4]);
            double v = fnum(r.tok[5]);
            if (ptype[4] == 'a') {
                if (a >= 1 && a <= h.nb_a && b >= 1 && b <= h.nr3_a && l >= 0 && l <= h.l3a)
                    p->crad_a[((b - 1) * (h.l3a + 1) + l) * h.nb_a + (a - 1)] = v;
            } else {
                if (a >= 1 && a <= h.nb_m && b >= 1 && b <= h.nr2_m && l >= 0 && l <= h.l2m)
                    p->crad_m[((b - 1) * (h.l2m + 1) + l) * h.nb_m + (a - 1)] = v;
            }
        } else if (!strcmp(ptype, "core")) {
            if (norder < 1 || norder > 2) { snprintf(g_err, sizeof g_err, "norder < 3 required for core."); goto fail; }
            int nt = rd_line(&r);
            if (nt < 0) break;
            if (nt < 3) continue;
            if (strcmp(r.tok[0], element) || strcmp(r.tok[1], element)) { snprintf(g_err, sizeof g_err, "Error when reading radial parameters."); goto fail; }
            p->core[norder - 1] = fnum(r.tok[2]);
        }
    }
    fclose(r.f);
    /* cutoff consistency (util.F90:420-444) */
    if (h.cut_a > rcut) { snprintf(g_err, sizeof g_err, "Global cutoff too short."); orc_free(p); return NULL; }
    if (h.cut_m > rcut_m) { snprintf(g_err, sizeof g_err, "Global cutoff too short."); orc_free(p); return NULL; }
    build_luts(p);
    return p;
bad:
    snprintf(g_err, sizeof g_err, "malformed potential header in %s", path);
fail:
    orc_free(p);
fail0:
    if (r.f) fclose(r.f);
    return NULL;
}

orc_pot *orc_load(const char *path, const char *element, double rcut, double rcut_magnetic,
                  double smear, int xc) {
    return load_impl(path, element, rcut, rcut_magnetic, smear, xc);
}

/* ----------------------------------------------------------------- energy (forces.F90:976-1289) */
static void fexp(double rho, double wpre, double mexp, int nfemb, double *F, double *dF) {
    const double w = 10.0, eps = 1e-10;
    if (nfemb == 0) { *F = rho; *dF = 1.0; return; }
    double srho = nfemb == 2 ? 1.0 + rho : rho;
    if (fabs(srho) > eps) {
        double y1 = w * srho * srho;
        double g = y1 > 30.0 ? 0.0 : exp(-y1);
        double omg = 1.0 - g, a = fabs(srho);
        y1 = pow(omg * a, mexp);
        double y2 = mexp * g * a;
        double sg = srho >= 0 ? 1.0 : -1.0;
        *F = nfemb == 2 ? wpre * (sg * (y1 + y2) - 1.0) : wpre * sg * (y1 + y2);
        double dg = -2.0 * w * srho * g, da = sg;
        double dy = fabs(y1) < eps ? 0.0 : mexp * y1 / (omg * a);
        double dy1 = dy * (-dg * a + omg * da), dy2 = mexp * (dg * a + g * da);
        *dF = wpre * sg * (dy1 + dy2);
    } else {
        *F = nfemb == 2 ? wpre * mexp * (srho - 1.0) : wpre * mexp * rho;
        *dF = wpre * mexp;
    }
}

static void innercutoff(double e, double ecut, double decut, double *f, double *df) {
    double elow = ecut - decut;
    if (e >= ecut) { *f = 0.0; *df = 0.0; }
    else if (e <= elow) { *f = 1.0; *df = 0.0; }
    else {
        *f = 0.5 * (1.0 + cos(PI * (e - elow) / decut));
        *df = -0.5 * sin(PI * (e - elow) / decut) * PI / decut;
    }
}

/* ----------------------------------------------------------------- per-atom evaluation */
typedef struct {
    /* per pair packs (a_basis.F90:204-240) */
    double *fr2, *dfr2, *cr, *dcr, *fra, *dfra, *frm, *dfrm;
    cplx *sha, *dsha, *shm, *dshm;
    int cap;
    /* per atom */
    cplx *abase, *abase2, *abase2m;
    double *b;
    cplx *db2, *db2_0, *db3, *db3_0, *db4, *db4_0, *db5, *db5_0, *db6, *db6_0;
    cplx *dw, *dw2, *dw0; double *dw1;
} work;

static void work_pairs(const orc_pot *p, work *w, int n) {
    if (n <= w->cap) return;
    const orc_header *h = &p->h;
    int nfa = h->nr3_a * (h->l3a + 1), nfm = h->nr2_m * (h->l2m + 1);
    free(w->fr2); free(w->dfr2); free(w->cr); free(w->dcr); free(w->fra); free(w->dfra);
    free(w->frm); free(w->dfrm); free(w->sha); free(w->dsha); free(w->shm); free(w->dshm);
    w->cap = n;
    w->fr2 = xcalloc((size_t)n * (h->nb_a + 1), 8); w->dfr2 = xcalloc((size_t)n * (h->nb_a + 1), 8);
    w->cr = xcalloc(n, 8); w->dcr = xcalloc(n, 8);
    w->fra = xcalloc((size_t)n * nfa, 8); w->dfra = xcalloc((size_t)n * nfa, 8);
    w->frm = xcalloc((size_t)n * nfm, 8); w->dfrm = xcalloc((size_t)n * nfm, 8);
    w->sha = xcalloc((size_t)n * p->lsa, sizeof(cplx)); w->dsha = xcalloc((size_t)n * 3 * p->lsa, sizeof(cplx));
    w->shm = xcalloc((size_t)n * p->lsm, sizeof(cplx)); w->dshm = xcalloc((size_t)n * 3 * p->lsm, sizeof(cplx));
}

static void work_init(const orc_pot *p, work *w) {
    const orc_header *h = &p->h;
    memset(w, 0, sizeof *w);
    w->abase = xcalloc((size_t)h->nr3_a * p->lsa * h->nr3_m * p->lsm, sizeof(cplx));
    w->abase2 = xcalloc((size_t)(h->nb_a ? h->nb_a : 1) * h->nr2_m * p->lsm, sizeof(cplx));
    w->abase2m = xcalloc((size_t)h->nr2_m * p->lsm, sizeof(cplx));
    w->b = xcalloc((size_t)p->nvar1 + p->nvar2 + p->nvar3 + p->nvar4 + p->nvar5 + p->nvar6, 8);
    int w2 = 2 * h->l2m + 1, w3a = 2 * h->l3a + 1, w3m = 2 * h->l3m + 1, w4a = 2 * h->l4a + 1,
        w4m = 2 * h->l4m + 1;
    int w30 = 2 * (h->l3a > h->l3m ? h->l3a : h->l3m) + 1, w40 = 2 * (h->l4a > h->l4m ? h->l4a : h->l4m) + 1;
    w->db2 = xcalloc((size_t)p->nvar2 * w2, sizeof(cplx));
    w->db2_0 = xcalloc((size_t)p->nvar2 * w2, sizeof(cplx));
    w->db3 = xcalloc((size_t)p->nvar3 * 2 * w3a * w3m, sizeof(cplx));
    w->db3_0 = xcalloc((size_t)p->nvar3 * w30, sizeof(cplx));
    w->db4 = xcalloc((size_t)p->nvar4 * 3 * w4a * w4m, sizeof(cplx));
    w->db4_0 = xcalloc((size_t)p->nvar4 * w40, sizeof(cplx));
    w->db5 = xcalloc((size_t)p->nvar5 * 4 * (2 * h->l5a + 1) * (2 * h->l5m + 1) + 1, sizeof(cplx));
    w->db5_0 = xcalloc((size_t)p->nvar5 * (2 * h->l5m + 1) + 1, sizeof(cplx));
    w->db6 = xcalloc((size_t)p->nvar6 * 5 * (2 * h->l6a + 1) * (2 * h->l6m + 1) + 1, sizeof(cplx));
    w->db6_0 = xcalloc((size_t)p->nvar6 * (2 * h->l6m + 1) + 1, sizeof(cplx));
    w->dw = xcalloc((size_t)w3a * (h->l3a + 1) * h->nr3_a * (h->l3m + 1) * (h->l3m + 1) * h->nr3_m, sizeof(cplx));
    w->dw2 = xcalloc((size_t)(h->nb_a ? h->nb_a : 1) * (h->l2m + 1) * (h->l2m + 1) * h->nr2_m, sizeof(cplx));
    w->dw0 = xcalloc((size_t)(h->l2m + 1) * (h->l2m + 1) * h->nr2_m, sizeof(cplx));
    w->dw1 = xcalloc((size_t)h->nb_m, 8);
}

static void work_free(work *w) {
    free(w->fr2); free(w->dfr2); free(w->cr); free(w->dcr); free(w->fra); free(w->dfra);
    free(w->frm); free(w->dfrm); free(w->sha); free(w->dsha); free(w->shm); free(w->dshm);
    free(w->abase); free(w->abase2); free(w->abase2m); free(w->b);
    free(w->db2); free(w->db2_0); free(w->db3); free(w->db3_0); free(w->db4); free(w->db4_0);
    free(w->db5); free(w->db5_0); free(w->db6); free(w->db6_0);
    free(w->dw); free(w->dw2); free(w->dw0); free(w->dw1);
}

typedef struct {
    int natom; const int *nstart, *nstop;
    const double *dlist, *elist, *dmi, *emi, *dm, *em;
    double *eout, *flist, *tlist, *tlist0, *bout;
} job;

/* one atom: prepA + prepB + forceweights + forcecompute + torquecompute */
static int eval_atom(const orc_pot *p, work *w, const job *J, int ia) {
    const orc_header *h = &p->h;
    const int nb_a = h->nb_a, nb_m = h->nb_m, nr2m = h->nr2_m, l2m = h->l2m;
    const int nr3a = h->nr3_a, nr3m = h->nr3_m, l3a = h->l3a, l3m = h->l3m;
    const int lsa = p->lsa, lsm = p->lsm, lha = p->lha, lhm = p->lhm;
    const int nfa = nr3a * (l3a + 1), nfm = nr2m * (l2m + 1);
    const int p0 = J->nstart[ia] - 1, np = J->nstop[ia] - J->nstart[ia] + 1;
    work_pairs(p, w, np > 0 ? np : 1);
#define FA(n, l) ((n) * (l3a + 1) + (l))
#define FM(n, l) ((n) * (l2m + 1) + (l))
#define AB(nr1, k1, nr2, k2) w->abase[(((size_t)(nr1) * lsa + (k1)) * nr3m + (nr2)) * lsm + (k2)]
#define AB2(n, nr, k) w->abase2[((size_t)(n) * nr2m + (nr)) * lsm + (k)]
#define AB2M(nr, k) w->abase2m[(size_t)(nr) * lsm + (k)]

    /* ---- prepA (a_basis.F90:59-359) */
    memset(w->abase, 0, sizeof(cplx) * (size_t)nr3a * lsa * nr3m * lsm);
    memset(w->abase2, 0, sizeof(cplx) * (size_t)(nb_a ? nb_a : 1) * nr2m * lsm);
    memset(w->abase2m, 0, sizeof(cplx) * (size_t)nr2m * lsm);
    double abase2hc = 0.0;
    double gmi[64], dgmi[64], fmi[4096], dfmi[4096];
    cplx shmi[300], dshmi[900];
    double mi = J->dmi[ia];
    const double *momi = J->emi + 3 * (size_t)ia;
    if (lookup_magnetic(p, mi, gmi, dgmi, fmi, dfmi)) return -2;
    ylm_cart(l2m, momi[0], momi[1], momi[2], shmi, dshmi);
    for (int nr = 0; nr < nr2m; ++nr)
        for (int nsh = 0; nsh < lhm; ++nsh) AB2M(nr, nsh) = fmi[FM(nr, shback(l2m, nsh))] * shmi[nsh];
    for (int nr = 0; nr < nr2m; ++nr)
        for (int l = 1; l <= l2m; ++l)
            for (int m = 1; m <= l; ++m) AB2M(nr, shx(l2m, l, -m)) = sgnm(m) * conj(AB2M(nr, shx(l2m, l, m)));
    /* the central-atom packs keep only half Y and fill -m (a_basis.F90:115-126): ylm_cart
       already produced the identical -m fill */

    for (int q = 0; q < np; ++q) {
        int ij = p0 + q;
        double x = J->dlist[ij];
        const double *eb = J->elist + 3 * (size_t)ij;
        double mm = J->dm[ij];
        const double *mom = J->em + 3 * (size_t)ij;
        double *gra = w->fr2 + (size_t)q * (nb_a + 1), *dgra = w->dfr2 + (size_t)q * (nb_a + 1);
        double *fra = w->fra + (size_t)q * nfa, *dfra = w->dfra + (size_t)q * nfa;
        double *frm = w->frm + (size_t)q * nfm, *dfrm = w->dfrm + (size_t)q * nfm;
        double gm[64], dgm[64];
        if (nb_a > 0) {
            if (lookup_atomic(p, x, gra, dgra, fra, dfra, &w->cr[q], &w->dcr[q])) return -1;
        }
        if (lookup_magnetic(p, mm, gm, dgm, frm, dfrm)) return -2;
        cplx *sha = w->sha + (size_t)q * lsa, *dsha = w->dsha + (size_t)q * 3 * lsa;
        cplx *shm = w->shm + (size_t)q * lsm, *dshm = w->dshm + (size_t)q * 3 * lsm;
        ylm_cart(l3a, eb[0], eb[1], eb[2], sha, dsha);
        ylm_cart(l2m, mom[0], mom[1], mom[2], shm, dshm);
        /* abase2 (a_basis.F90:193-201) */
        for (int nr = 0; nr < nr2m; ++nr)
            for (int l = 0; l <= l2m; ++l)
                for (int m = 0; m <= l; ++m) {
                    int k = shx(l2m, l, m);
                    cplx t = frm[FM(nr, l)] * shm[k];
                    for (int n = 0; n < nb_a; ++n) AB2(n, nr, k) += gra[n] * t;
                }
        abase2hc += w->cr[q];
        /* abase (a_basis.F90:248-291) */
        for (int nr1 = 0; nr1 < nr3a; ++nr1)
            for (int nr2 = 0; nr2 < nr3m; ++nr2)
                for (int l1 = 0; l1 <= l3a; ++l1)
                    for (int l2 = 0; l2 <= l3m; ++l2)
                        for (int m2 = -l2; m2 <= l2; ++m2) {
                            int k1 = shx(l3a, l1, 0), k2 = shx(l2m, l2, m2);
                            cplx x1 = fra[FA(nr1, l1)] * frm[FM(nr2, l2)] * shm[k2];
                            AB(nr1, k1, nr2, k2) += x1 * sha[k1];
                        }
        for (int nr1 = 0; nr1 < nr3a; ++nr1)
            for (int nr2 = 0; nr2 < nr3m; ++nr2)
                for (int l1 = 1; l1 <= l3a; ++l1)
                    for (int l2 = 0; l2 <= l3m; ++l2)
                        for (int m1 = 1; m1 <= l1; ++m1)
                            for (int m2 = 0; m2 <= l2; ++m2) {
                                int k1 = shx(l3a, l1, m1), k2 = shx(l2m, l2, m2), k3 = shx(l3a, l1, -m1);
                                cplx x1 = fra[FA(nr1, l1)] * frm[FM(nr2, l2)] * shm[k2];
                                cplx y1 = sha[k1];
                                AB(nr1, k1, nr2, k2) += x1 * y1;
                                AB(nr1, k3, nr2, k2) += x1 * conj(y1) * sgnm(m1);
                            }
    }
    /* missing halves (a_basis.F90:303-342) */
    for (int nr1 = 0; nr1 < nr3a; ++nr1)
        for (int nr2 = 0; nr2 < nr3m; ++nr2)
            for (int l1 = 1; l1 <= l3a; ++l1)
                for (int l2 = 1; l2 <= l3m; ++l2)
                    for (int m1 = 1; m1 <= l1; ++m1)
                        for (int m2 = -l2; m2 <= -1; ++m2) {
                            int k1 = shx(l3a, l1, m1), k2 = shx(l3a, l1, -m1);
                            int k3 = shx(l2m, l2, m2), k4 = shx(l2m, l2, -m2);
                            double xs = sgnm(m1 + m2);
                            AB(nr1, k1, nr2, k3) = xs * conj(AB(nr1, k2, nr2, k4));
                            AB(nr1, k2, nr2, k3) = xs * conj(AB(nr1, k1, nr2, k4));
                        }
    for (int nr = 0; nr < nr2m; ++nr)
        for (int l = 1; l <= l2m; ++l)
            for (int m = 1; m <= l; ++m)
                for (int n = 0; n < nb_a; ++n)
                    AB2(n, nr, shx(l2m, l, -m)) = sgnm(m) * conj(AB2(n, nr, shx(l2m, l, m)));

    /* ---- prepB (b_basis.F90:62-307) */
    const int W2 = 2 * l2m + 1, W3A = 2 * l3a + 1, W3M = 2 * l3m + 1;
    const int W4A = 2 * h->l4a + 1, W4M = 2 * h->l4m + 1;
    const int L30 = l3a > l3m ? l3a : l3m, L40 = h->l4a > h->l4m ? h->l4a : h->l4m;
    const int W30 = 2 * L30 + 1, W40 = 2 * L40 + 1;
    const int LM = p->lmax, WL = 2 * LM + 1;
    double *b1 = w->b, *b2 = b1 + p->nvar1, *b3 = b2 + p->nvar2, *b4 = b3 + p->nvar3;
    for (int k = 0; k < p->nvar1; ++k) b1[k] = gmi[p->b1index[k] - 1] - 1.0;
    memset(w->db2, 0, sizeof(cplx) * (size_t)p->nvar2 * W2);
    memset(w->db2_0, 0, sizeof(cplx) * (size_t)p->nvar2 * W2);
    if (nr2m > 0) {
        for (int c = 0; c < p->nvar2; ++c) {
            const int *bi = p->b2index + 4 * c;
            int nr1 = bi[0] - 1, nr2 = bi[1] - 1, nr3 = bi[2] - 1, l = bi[3];
            cplx xiy = 0;
            for (int m1 = -l; m1 <= l; ++m1) {
                int m2 = -m1;
                cplx ap0 = AB2M(nr1, shx(l2m, l, m1)), ap1 = AB2(nr3, nr2, shx(l2m, l, m2));
                double cf = p->cg01[l * WL + m1 + LM];
                xiy += cf * ap0 * ap1;
                w->db2[(size_t)c * W2 + m2 + l2m] += cf * ap0;
                w->db2_0[(size_t)c * W2 + m1 + l2m] += cf * ap1;
            }
            b2[c] = creal(xiy);
        }
    }
    memset(w->db3, 0, sizeof(cplx) * (size_t)p->nvar3 * 2 * W3A * W3M);
    memset(w->db3_0, 0, sizeof(cplx) * (size_t)p->nvar3 * W30);
#define DB3(s, c, m, mp) w->db3[((((size_t)(c) * 2 + (s)) * W3A + (m) + l3a) * W3M) + (mp) + l3m]
#define DB4(s, c, m, mp) w->db4[((((size_t)(c) * 3 + (s)) * W4A + (m) + h->l4a) * W4M) + (mp) + h->l4m]
    if (nr3a > 0 || nr3m > 0) {
        for (int c = 0; c < p->nvar3; ++c) {
            const int *bi = p->b3index + 9 * c;
            int nr1 = bi[0] - 1, nr2 = bi[1] - 1, nr3 = bi[2] - 1, nr4 = bi[3] - 1, nr5 = bi[4] - 1;
            int l1 = bi[5], l2 = bi[6], l3 = bi[7], l4 = bi[8];
            int ka = p->b3map_a[c], km = p->b3map_m[c];
            cplx xiy = 0;
            for (int m1 = -l1; m1 <= l1; ++m1) {
                int m2 = -m1;
                int k1 = shx(l3a, l1, m1), k2 = shx(l3a, l1, m2);
                double cfa = p->cg02a[ka * WL + m1 + LM];
                for (int q = 0; q < p->madd3_m[km]; ++q) {
                    const int *e = p->b3lm_m + 3 * (km * p->maddmax3m + q);
                    int m3 = e[0], m4 = e[1], m5 = e[2];
                    int k3 = shx(l2m, l2, m3), k4 = shx(l2m, l3, m4), k5 = shx(l2m, l4, m5);
                    cplx ap0 = AB2M(nr3, k3), ap1 = AB(nr1, k1, nr4, k4), ap2 = AB(nr2, k2, nr5, k5);
                    double cf = cfa * p->cg02m[km * p->maddmax3m + q];
                    xiy += cf * ap0 * ap1 * ap2;
                    DB3(0, c, m1, m4) += cf * ap0 * ap2;
                    DB3(1, c, m2, m5) += cf * ap0 * ap1;
                    w->db3_0[(size_t)c * W30 + m3 + L30] += cf * ap1 * ap2;
                }
            }
            b3[c] = creal(xiy);
        }
    }
    memset(w->db4, 0, sizeof(cplx) * (size_t)p->nvar4 * 3 * W4A * W4M);
    memset(w->db4_0, 0, sizeof(cplx) * (size_t)p->nvar4 * W40);
    if (h->nr4_a > 0 || h->nr4_m > 0) {
        for (int c = 0; c < p->nvar4; ++c) {
            const int *bi = p->b4index + 15 * c;
            int nr1 = bi[0] - 1, nr2 = bi[1] - 1, nr3 = bi[2] - 1, nr4 = bi[3] - 1, nr5 = bi[4] - 1,
                nr6 = bi[5] - 1, nr7 = bi[6] - 1;
            int l1 = bi[7], l2 = bi[8], l3 = bi[9], l4 = bi[10], l5 = bi[11], l6 = bi[12], l7 = bi[13];
            int ka = p->b4map_a[c], km = p->b4map_m[c];
            cplx xiy = 0;
            for (int qa = 0; qa < p->madd4_a[ka]; ++qa) {
                const int *ea = p->b4lm_a + 3 * (ka * p->maddmax4a + qa);
                int m1 = ea[0], m2 = ea[1], m3 = ea[2];
                int k1 = shx(l3a, l1, m1), k2 = shx(l3a, l2, m2), k3 = shx(l3a, l3, m3);
                double cfa = p->cg03a[ka * p->maddmax4a + qa];
                for (int qm = 0; qm < p->madd4_m[km]; ++qm) {
                    const int *em = p->b4lm_m + 4 * (km * p->maddmax4m + qm);
                    int m4 = em[0], m5 = em[1], m6 = em[2], m7 = em[3];
                    int k4 = shx(l2m, l4, m4), k5 = shx(l2m, l5, m5), k6 = shx(l2m, l6, m6), k7 = shx(l2m, l7, m7);
                    cplx ap0 = AB2M(nr4, k4), ap1 = AB(nr1, k1, nr5, k5), ap2 = AB(nr2, k2, nr6, k6),
                         ap3 = AB(nr3, k3, nr7, k7);
                    double cf = cfa * p->cg03m[km * p->maddmax4m + qm];
                    cplx xb3 = ap0 * ap2 * ap3;
                    cplx xb4 = cf * ap1;
                    xiy += xb4 * xb3;
                    DB4(0, c, m1, m5) += cf * xb3;
                    DB4(1, c, m2, m6) += ap0 * xb4 * ap3;
                    DB4(2, c, m3, m7) += ap0 * xb4 * ap2;
                    w->db4_0[(size_t)c * W40 + m4 + L40] += xb4 * ap2 * ap3;
                }
            }
            b4[c] = creal(xiy);
        }
    }
    /* 5- and 6-body functions (b_basis.F90:320-532): B and dB/dA per factor, dB/dA0 */
    double *b5 = b4 + p->nvar4, *b6 = b5 + p->nvar5;
    const int W5A = 2 * h->l5a + 1, W5M = 2 * h->l5m + 1, W6A = 2 * h->l6a + 1, W6M = 2 * h->l6m + 1;
#define DB5(s, c, m, mp) w->db5[((((size_t)(c) * 4 + (s)) * W5A + (m) + h->l5a) * W5M) + (mp) + h->l5m]
#define DB6(s, c, m, mp) w->db6[((((size_t)(c) * 5 + (s)) * W6A + (m) + h->l6a) * W6M) + (mp) + h->l6m]
    memset(w->db5, 0, sizeof(cplx) * ((size_t)p->nvar5 * 4 * W5A * W5M + 1));
    memset(w->db5_0, 0, sizeof(cplx) * ((size_t)p->nvar5 * W5M + 1));
    for (int c = 0; c < p->nvar5; ++c) {
        const int *bi = p->b5index + 21 * c;
        const int n0 = bi[4] - 1, l0 = bi[13];
        cplx xiy = 0;
        for (int e = p->m5ptr[c]; e < p->m5ptr[c + 1]; ++e) {
            const int *mm = p->m5lm + 9 * e;
            const double cf = p->m5cf[e];
            cplx ap[5];
            ap[0] = AB2M(n0, shx(l2m, l0, mm[4]));
            for (int i = 0; i < 4; ++i)
                ap[1 + i] = AB(bi[i] - 1, shx(l3a, bi[9 + i], mm[i]), bi[5 + i] - 1, shx(l2m, bi[14 + i], mm[5 + i]));
            cplx xb1 = ap[0] * cf * ap[1] * ap[2], xb2 = ap[3] * ap[4];
            xiy += xb1 * xb2;
            DB5(0, c, mm[0], mm[5]) += ap[0] * cf * ap[2] * xb2;
            DB5(1, c, mm[1], mm[6]) += ap[0] * cf * ap[1] * xb2;
            DB5(2, c, mm[2], mm[7]) += xb1 * ap[4];
            DB5(3, c, mm[3], mm[8]) += xb1 * ap[3];
            w->db5_0[(size_t)c * W5M + mm[4] + h->l5m] += cf * ap[1] * ap[2] * xb2;
        }
        b5[c] = creal(xiy);
    }
    memset(w->db6, 0, sizeof(cplx) * ((size_t)p->nvar6 * 5 * W6A * W6M + 1));
    memset(w->db6_0, 0, sizeof(cplx) * ((size_t)p->nvar6 * W6M + 1));
    for (int c = 0; c < p->nvar6; ++c) {
        const int *bi = p->b6index + 27 * c;
        const int n0 = bi[5] - 1, l0 = bi[16];
        cplx xiy = 0;
        for (int e = p->m6ptr[c]; e < p->m6ptr[c + 1]; ++e) {
            const int *mm = p->m6lm + 11 * e;
            const double cf = p->m6cf[e];
            cplx ap[6];
            ap[0] = AB2M(n0, shx(l2m, l0, mm[5]));
            for (int i = 0; i < 5; ++i)
                ap[1 + i] = AB(bi[i] - 1, shx(l3a, bi[11 + i], mm[i]), bi[6 + i] - 1, shx(l2m, bi[17 + i], mm[6 + i]));
            cplx xb1 = ap[0] * cf * ap[1] * ap[2], xb2 = ap[3] * ap[4] * ap[5];
            xiy += xb1 * xb2;
            DB6(0, c, mm[0], mm[6]) += ap[0] * cf * ap[2] * xb2;
            DB6(1, c, mm[1], mm[7]) += ap[0] * cf * ap[1] * xb2;
            DB6(2, c, mm[2], mm[8]) += xb1 * ap[4] * ap[5];
            DB6(3, c, mm[3], mm[9]) += xb1 * ap[3] * ap[5];
            DB6(4, c, mm[4], mm[10]) += xb1 * ap[3] * ap[4];
            w->db6_0[(size_t)c * W6M + mm[5] + h->l6m] += cf * ap[1] * ap[2] * xb2;
        }
        b6[c] = creal(xiy);
    }
    if (J->bout) {
        size_t nb = (size_t)p->nvar1 + p->nvar2 + p->nvar3 + p->nvar4 + p->nvar5 + p->nvar6;
        memcpy(J->bout + nb * ia, w->b, nb * sizeof(double));
    }

    /* ---- energy + forceweights (forces.F90:84-349, 1093-1130) */
    const int nd = h->ndensity;
    double rho[2] = {0, 0}, dEdrho[2], energyi = 0.0, edf = 1.0;
    for (int d = 0; d < nd; ++d) {
        double s = 0;
        for (int k = 0; k < p->nvar1; ++k) s += p->c0[k * nd + d] * b1[k];
        rho[d] = s;
        if (nr2m > 0) { s = 0; for (int k = 0; k < p->nvar2; ++k) s += p->c2[k * nd + d] * b2[k]; rho[d] += s; }
        if (nr3a > 0 || nr3m > 0) { s = 0; for (int k = 0; k < p->nvar3; ++k) s += p->c3[k * nd + d] * b3[k]; rho[d] += s; }
        if (h->nr4_a > 0 || h->nr4_m > 0) { s = 0; for (int k = 0; k < p->nvar4; ++k) s += p->c4[k * nd + d] * b4[k]; rho[d] += s; }
        if (h->nr5_a > 0 || h->nr5_m > 0) { s = 0; for (int k = 0; k < p->nvar5; ++k) s += p->c5[k * nd + d] * b5[k]; rho[d] += s; }
        if (h->nr6_a > 0 || h->nr6_m > 0) { s = 0; for (int k = 0; k < p->nvar6; ++k) s += p->c6[k * nd + d] * b6[k]; rho[d] += s; }
    }
    for (int d = 0; d < nd; ++d) {
        double F, dF;
        fexp(rho[d], p->p1[0][d], p->p1[1][d], h->nfemb, &F, &dF);
        energyi += F; dEdrho[d] = dF;
    }
    if (nr2m > 0) {
        double fc, dfc;
        innercutoff(abase2hc, h->ecut, h->decut, &fc, &dfc);
        edf = dfc * energyi + 1.0;
        for (int d = 0; d < nd; ++d) dEdrho[d] *= fc;
        energyi = fc * energyi + abase2hc;
    }
    J->eout[ia] = energyi;

#define CEFF(carr, k) ({ double _x = 0; for (int _d = 0; _d < nd; ++_d) _x += dEdrho[_d] * (carr)[(k) * nd + _d]; _x; })
    const int DWM = 2 * l3a + 1;
#define DW(m, l, n, mp, lp, np_) w->dw[(((((size_t)(np_) * (l3m + 1) + (lp)) * (l3m + 1) + (mp)) * nr3a + (n)) * (l3a + 1) + (l)) * DWM + (m) + l3a]
#define DW2(n, m, l, nr) w->dw2[(((size_t)(nr) * (l2m + 1) + (l)) * (l2m + 1) + (m)) * (nb_a ? nb_a : 1) + (n)]
#define DW0(m, l, nr) w->dw0[((size_t)(nr) * (l2m + 1) + (l)) * (l2m + 1) + (m)]
    memset(w->dw, 0, sizeof(cplx) * (size_t)DWM * (l3a + 1) * nr3a * (l3m + 1) * (l3m + 1) * nr3m);
    memset(w->dw2, 0, sizeof(cplx) * (size_t)(nb_a ? nb_a : 1) * (l2m + 1) * (l2m + 1) * nr2m);
    memset(w->dw0, 0, sizeof(cplx) * (size_t)(l2m + 1) * (l2m + 1) * nr2m);
    memset(w->dw1, 0, sizeof(double) * nb_m);
    for (int c = 0; c < p->nvar1; ++c) w->dw1[p->b1index[c] - 1] += CEFF(p->c0, c);
    for (int c = 0; c < p->nvar2; ++c) {
        const int *bi = p->b2index + 4 * c;
        double ce = CEFF(p->c2, c);
        int l = bi[3];
        for (int m = 0; m <= l; ++m) DW2(bi[2] - 1, m, l, bi[1] - 1) += ce * w->db2[(size_t)c * W2 + m + l2m];
        for (int m = 0; m <= l; ++m) DW0(m, l, bi[0] - 1) += ce * w->db2_0[(size_t)c * W2 + m + l2m];
    }
    for (int c = 0; c < p->nvar3; ++c) {
        const int *bi = p->b3index + 9 * c;
        double ce = CEFF(p->c3, c);
        int l1 = bi[5], l2 = bi[6], l4 = bi[7], l5 = bi[8];
        for (int m1 = -l1; m1 <= l1; ++m1)
            for (int m4 = 0; m4 <= l4; ++m4) DW(m1, l1, bi[0] - 1, m4, l4, bi[3] - 1) += ce * DB3(0, c, m1, m4);
        for (int m1 = -l1; m1 <= l1; ++m1)
            for (int m5 = 0; m5 <= l5; ++m5) DW(m1, l1, bi[1] - 1, m5, l5, bi[4] - 1) += ce * DB3(1, c, m1, m5);
        for (int m2 = 0; m2 <= l2; ++m2) DW0(m2, l2, bi[2] - 1) += ce * w->db3_0[(size_t)c * W30 + m2 + L30];
    }
    for (int c = 0; c < p->nvar4; ++c) {
        const int *bi = p->b4index + 15 * c;
        double ce = CEFF(p->c4, c);
        int l1 = bi[7], l2 = bi[8], l3 = bi[9], l4 = bi[10], l5 = bi[11], l6 = bi[12], l7 = bi[13];
        for (int m1 = -l1; m1 <= l1; ++m1)
            for (int m5 = 0; m5 <= l5; ++m5) DW(m1, l1, bi[0] - 1, m5, l5, bi[4] - 1) += ce * DB4(0, c, m1, m5);
        for (int m2 = -l2; m2 <= l2; ++m2)
            for (int m6 = 0; m6 <= l6; ++m6) DW(m2, l2, bi[1] - 1, m6, l6, bi[5] - 1) += ce * DB4(1, c, m2, m6);
        for (int m3 = -l3; m3 <= l3; ++m3)
            for (int m7 = 0; m7 <= l7; ++m7) DW(m3, l3, bi[2] - 1, m7, l7, bi[6] - 1) += ce * DB4(2, c, m3, m7);
        for (int m4 = 0; m4 <= l4; ++m4) DW0(m4, l4, bi[3] - 1) += ce * w->db4_0[(size_t)c * W40 + m4 + L40];
    }
    /* 5- and 6-body weights (forces.F90:232-330) */
    for (int c = 0; c < p->nvar5; ++c) {
        const int *bi = p->b5index + 21 * c;
        double ce = CEFF(p->c5, c);
        for (int i = 0; i < 4; ++i) {
            const int l = bi[9 + i], lp = bi[14 + i];
            for (int m = -l; m <= l; ++m)
                for (int mp = 0; mp <= lp; ++mp) DW(m, l, bi[i] - 1, mp, lp, bi[5 + i] - 1) += ce * DB5(i, c, m, mp);
        }
        for (int m0 = 0; m0 <= bi[13]; ++m0) DW0(m0, bi[13], bi[4] - 1) += ce * w->db5_0[(size_t)c * W5M + m0 + h->l5m];
    }
    for (int c = 0; c < p->nvar6; ++c) {
        const int *bi = p->b6index + 27 * c;
        double ce = CEFF(p->c6, c);
        for (int i = 0; i < 5; ++i) {
            const int l = bi[11 + i], lp = bi[17 + i];
            for (int m = -l; m <= l; ++m)
                for (int mp = 0; mp <= lp; ++mp) DW(m, l, bi[i] - 1, mp, lp, bi[6 + i] - 1) += ce * DB6(i, c, m, mp);
        }
        for (int m0 = 0; m0 <= bi[16]; ++m0) DW0(m0, bi[16], bi[5] - 1) += ce * w->db6_0[(size_t)c * W6M + m0 + h->l6m];
    }

    /* ---- forcecompute (forces.F90:418-623) */
    for (int q = 0; q < np; ++q) {
        int ij = p0 + q;
        const double *eb = J->elist + 3 * (size_t)ij;
        double x = J->dlist[ij], invr = 1.0 / x;
        const double *dfr2 = w->dfr2 + (size_t)q * (nb_a + 1);
        const double *fra = w->fra + (size_t)q * nfa, *dfra = w->dfra + (size_t)q * nfa;
        const double *frm = w->frm + (size_t)q * nfm;
        const cplx *sha = w->sha + (size_t)q * lsa, *dsha = w->dsha + (size_t)q * 3 * lsa;
        const cplx *shm = w->shm + (size_t)q * lsm;
        double Fk[3] = {0, 0, 0};
        for (int n1 = 0; n1 < nb_a; ++n1)
            for (int n2 = 0; n2 < nr2m; ++n2)
                for (int l = 0; l <= l2m; ++l)
                    for (int m = 0; m <= l; ++m) {
                        cplx wei = DW2(n1, m, l, n2);
                        cplx t = frm[FM(n2, l)] * shm[shx(l2m, l, m)];
                        double f = m == 0 ? 1.0 : 2.0;
                        for (int c = 0; c < 3; ++c) {
                            cplx xv = t * (dfr2[n1] * eb[c]);
                            Fk[c] += f * (creal(wei) * creal(xv) - cimag(wei) * cimag(xv));
                        }
                    }
        for (int c = 0; c < 3; ++c) Fk[c] += edf * w->dcr[q] * eb[c];
        for (int n1 = 0; n1 < nr3a; ++n1)
            for (int n2 = 0; n2 < nr3m; ++n2)
                for (int l1 = 0; l1 <= l3a; ++l1)
                    for (int l2 = 0; l2 <= l3m; ++l2) {
                        double x2 = invr * fra[FA(n1, l1)], dr = dfra[FA(n1, l1)];
                        /* (m1,m2) list with factors as in :509-614 */
                        for (int m1 = -l1; m1 <= l1; ++m1)
                            for (int m2 = 0; m2 <= l2; ++m2) {
                                double f;
                                if (m1 == 0 && m2 == 0) f = 1.0;
                                else if (m2 == 0 && m1 < 0) continue; /* folded by symmetry */
                                else f = 2.0;
                                cplx wei = DW(m1, l1, n1, m2, l2, n2);
                                int k = shx(l3a, l1, m1), k1 = shx(l2m, l2, m2);
                                cplx pre = shm[k1] * frm[FM(n2, l2)];
                                for (int c = 0; c < 3; ++c) {
                                    cplx xv = pre * ((dr * eb[c]) * sha[k] + x2 * dsha[c * lsa + k]);
                                    Fk[c] += f * (creal(wei) * creal(xv) - cimag(wei) * cimag(xv));
                                }
                            }
                    }
        for (int c = 0; c < 3; ++c) J->flist[3 * (size_t)ij + c] = Fk[c];
    }

    /* ---- torquecompute (forces.F90:688-961) */
    {
        double Tk0[3] = {0, 0, 0};
        double invm = 1.0 / mi;
        for (int nr = 0; nr < nb_m; ++nr)
            for (int c = 0; c < 3; ++c) Tk0[c] += w->dw1[nr] * dgmi[nr] * momi[c];
        for (int nr = 0; nr < nr2m; ++nr)
            for (int l = 0; l <= l2m; ++l) {
                double dv = dfmi[FM(nr, l)], x2 = invm * fmi[FM(nr, l)];
                for (int m = 0; m <= l; ++m) {
                    cplx wei = DW0(m, l, nr);
                    int k = shx(l2m, l, m);
                    double f = m == 0 ? 1.0 : 2.0;
                    for (int c = 0; c < 3; ++c) {
                        cplx xv = shmi[k] * (dv * momi[c]) + dshmi[c * lsm + k] * x2;
                        Tk0[c] += f * (creal(wei) * creal(xv) - cimag(wei) * cimag(xv));
                    }
                }
            }
        for (int c = 0; c < 3; ++c) J->tlist0[3 * (size_t)ia + c] = Tk0[c];
    }
    for (int q = 0; q < np; ++q) {
        int ij = p0 + q;
        const double *em = J->em + 3 * (size_t)ij;
        double invm = 1.0 / J->dm[ij];
        const double *fr2 = w->fr2 + (size_t)q * (nb_a + 1);
        const double *fra = w->fra + (size_t)q * nfa;
        const double *frm = w->frm + (size_t)q * nfm, *dfrm = w->dfrm + (size_t)q * nfm;
        const cplx *sha = w->sha + (size_t)q * lsa;
        const cplx *shm = w->shm + (size_t)q * lsm, *dshm = w->dshm + (size_t)q * 3 * lsm;
        double Tk[3] = {0, 0, 0};
        for (int n1 = 0; n1 < nb_a; ++n1)
            for (int n2 = 0; n2 < nr2m; ++n2)
                for (int l = 0; l <= l2m; ++l) {
                    double x2 = invm * frm[FM(n2, l)], dv = dfrm[FM(n2, l)];
                    for (int m = 0; m <= l; ++m) {
                        cplx wei = DW2(n1, m, l, n2);
                        int k1 = shx(l2m, l, m);
                        double f = m == 0 ? 1.0 : 2.0;
                        for (int c = 0; c < 3; ++c) {
                            cplx xv = fr2[n1] * ((dv * em[c]) * shm[k1] + x2 * dshm[c * lsm + k1]);
                            Tk[c] += f * (creal(wei) * creal(xv) - cimag(wei) * cimag(xv));
                        }
                    }
                }
        for (int n1 = 0; n1 < nr3a; ++n1)
            for (int n2 = 0; n2 < nr3m; ++n2)
                for (int l1 = 0; l1 <= l3a; ++l1)
                    for (int l2 = 0; l2 <= l3m; ++l2) {
                        double x2 = invm * frm[FM(n2, l2)], dv = dfrm[FM(n2, l2)];
                        for (int m1 = -l1; m1 <= l1; ++m1)
                            for (int m2 = 0; m2 <= l2; ++m2) {
                                double f;
                                if (m1 == 0 && m2 == 0) f = 1.0;
                                else if (m2 == 0 && m1 < 0) continue;
                                else f = 2.0;
                                cplx wei = DW(m1, l1, n1, m2, l2, n2);
                                int k = shx(l3a, l1, m1), k1 = shx(l2m, l2, m2);
                                cplx pre = sha[k] * fra[FA(n1, l1)];
                                for (int c = 0; c < 3; ++c) {
                                    cplx xv = pre * ((dv * em[c]) * shm[k1] + x2 * dshm[c * lsm + k1]);
                                    Tk[c] += f * (creal(wei) * creal(xv) - cimag(wei) * cimag(xv));
                                }
                            }
                    }
        for (int c = 0; c < 3; ++c) J->tlist[3 * (size_t)ij + c] = Tk[c];
    }
    return 0;
}

static int run(orc_pot *p, job *J, int nthreads) {
    int err = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads) reduction(| : err)
#endif
    {
        work w;
        work_init(p, &w);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int ia = 0; ia < J->natom; ++ia) {
            int rc = eval_atom(p, &w, J, ia);
            if (rc == -1) err |= 1;
            if (rc == -2) err |= 2;
        }
        work_free(&w);
    }
    if (err & 1) { snprintf(g_err, sizeof g_err, "Encountered very small distance."); return -1; }
    if (err & 2) { snprintf(g_err, sizeof g_err, "moment outside the magnetic spline table"); return -2; }
    return 0;
}

int orc_acelib(orc_pot *p, int natom, int nei, const int *nstart, const int *nstop,
               const double *dlist, const double *elist, const double *dlist_momi,
               const double *elist_momi, const double *dlist_mom, const double *elist_mom,
               double *eout, double *flist, double *tlist, double *tlist0, int nthreads) {
    (void)nei;
    job J = {natom, nstart, nstop, dlist, elist, dlist_momi, elist_momi, dlist_mom, elist_mom,
             eout, flist, tlist, tlist0, NULL};
    return run(p, &J, nthreads);
}

int orc_bvectors(orc_pot *p, int natom, int nei, const int *nstart, const int *nstop,
                 const double *dlist, const double *elist, const double *dlist_momi,
                 const double *elist_momi, const double *dlist_mom, const double *elist_mom,
                 double *bout) {
    double *eout = xcalloc(natom, 8), *fl = xcalloc(3 * (size_t)nei + 1, 8);
    double *tl = xcalloc(3 * (size_t)nei + 1, 8), *t0 = xcalloc(3 * (size_t)natom, 8);
    job J = {natom, nstart, nstop, dlist, elist, dlist_momi, elist_momi, dlist_mom, elist_mom,
             eout, fl, tl, t0, bout};
    int rc = run(p, &J, 0);  /* atoms write their own rows of bout: OpenMP over atoms */
    free(eout); free(fl); free(tl); free(t0);
    return rc;
}
