/*
 * neilist_oracle.c -- CPU FP64 restatement of the reference neighbour list and of the
 * structure packing around it.  TEST INFRASTRUCTURE ONLY (the oracle): never linked into
 * the product; only tests/ and __graft_entry__.smoke() load it.
 *
 *   cluster with periodic images   neilist/neilist.F90:192-452 (buildcluster)
 *   3x3 inverse                    neilist/neilist.F90:852-902 (inv3x3)
 *   grid boxes + stencil           neilist/neilist.F90:455-544 (buildgrid)
 *   atoms -> boxes                 neilist/neilist.F90:547-636 (refreshgrid)
 *   pairs                          neilist/neilist.F90:639-796 (refreshneighbours)
 *   distance test                  neilist/neilist.F90:825-848 (fastdist)
 *
 * Pair order is the reference's: central atoms in input order; for each, the stencil
 * boxes (own box first, then the others in (k1,k2,k3) order) and, inside a box, cluster
 * atoms in ascending cluster index.  The FP64 expressions keep the reference's operation
 * order and are compiled without contraction (-ffp-contract=off), so pair sets, distances
 * and unit vectors are bit-identical to the Fortran build.
 */
#include "magace_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NL_EPS 1.0e-9 /* neilist/neilist.F90:9 */

/* normalised cross product v1 x v2 (neilist.F90:799-822); 0 if too short */
static int north(const double *v1, const double *v2, double *vn) {
    vn[0] = v1[1] * v2[2] - v1[2] * v2[1];
    vn[1] = v1[2] * v2[0] - v1[0] * v2[2];
    vn[2] = v1[0] * v2[1] - v1[1] * v2[0];
    double x = sqrt(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
    if (x < 1.0e-9) return 0;
    vn[0] /= x; vn[1] /= x; vn[2] /= x;
    return 1;
}

/* cofactor inverse with the reference's term order (neilist.F90:852-902).
 * a[i + 3*k] = a(i+1, k+1) (column-major). */
static int inv3(const double *a, double *ai) {
#define A(i, k) a[(i - 1) + 3 * (k - 1)]
    double c[9];
    c[0] = A(3, 3) * A(2, 2) - A(3, 2) * A(2, 3);
    c[3] = -(A(3, 3) * A(1, 2) - A(3, 2) * A(1, 3));
    c[6] = A(2, 3) * A(1, 2) - A(2, 2) * A(1, 3);
    c[1] = -(A(3, 3) * A(2, 1) - A(3, 1) * A(2, 3));
    c[4] = A(3, 3) * A(1, 1) - A(3, 1) * A(1, 3);
    c[7] = -(A(2, 3) * A(1, 1) - A(2, 1) * A(1, 3));
    c[2] = A(3, 2) * A(2, 1) - A(3, 1) * A(2, 2);
    c[5] = -(A(3, 2) * A(1, 1) - A(3, 1) * A(1, 2));
    c[8] = A(2, 2) * A(1, 1) - A(2, 1) * A(1, 2);
    double det = A(1, 1) * (A(3, 3) * A(2, 2) - A(3, 2) * A(2, 3))
               - A(2, 1) * (A(3, 3) * A(1, 2) - A(3, 2) * A(1, 3))
               + A(3, 1) * (A(2, 3) * A(1, 2) - A(2, 2) * A(1, 3));
#undef A
    if (!(fabs(det) > 1.0e-9)) return 0;
    for (int k = 0; k < 9; ++k) ai[k] = c[k] / det;
    return 1;
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

int orc_neilist(int natom, int periodic, double rcut, const double *cell, const double *posin,
                const double *magin, int cap_box, int cap_pair, int *counts, int *nstart,
                int *nstop, int *nlist, int *origin, int *nshift, double *dist, double *eist,
                double *dist_mom, double *eist_mom, double *dist_momi, double *eist_momi,
                double *posbox) {
    if (rcut < 1.0e-6 || rcut > 100.0) return -1;  /* neilist.F90:206-213 */
    if (natom < 0) return -1;
    const double *c1 = cell, *c2 = cell + 3, *c3 = cell + 6;
    int lc[3] = {0, 0, 0}, hc[3] = {0, 0, 0}, nglow[3], ngup[3];
    double dxl[3] = {0, 0, 0}, dxh[3] = {0, 0, 0}, dxlcc[3], dxhcc[3];
    double ci[9];
    if (periodic) {
        /* slab of thickness rcut around each pair of faces (neilist.F90:215-260) */
        const double *pairs[3][3] = {{c1, c2, c3}, {c3, c1, c2}, {c2, c3, c1}};
        const int dir[3] = {2, 1, 0};
        for (int t = 0; t < 3; ++t) {
            double v[3];
            if (!north(pairs[t][0], pairs[t][1], v)) return -2;
            double x = dot3(pairs[t][2], v);
            if (x < 1.0e-3) return -2;
            double h = 1.0 + rcut / x;
            double l = -h + 1.0;
            hc[dir[t]] = (int)ceil(h);
            lc[dir[t]] = (int)floor(l);
            dxh[dir[t]] = h;
            dxl[dir[t]] = l;
        }
        /* orthorhombic box around the cell plus rcut (neilist.F90:262-311) */
        for (int n = 0; n < 3; ++n) {
            double hi = 0.0, lo = 0.0;
            for (int i1 = 0; i1 <= 1; ++i1)
                for (int i2 = 0; i2 <= 1; ++i2)
                    for (int i3 = 0; i3 <= 1; ++i3) {
                        double x = rcut + (double)i1 * c1[n] + (double)i2 * c2[n] + (double)i3 * c3[n];
                        if (x > hi) hi = x;
                        double y = -rcut + (double)i1 * c1[n] + (double)i2 * c2[n] + (double)i3 * c3[n];
                        if (y < lo) lo = y;
                    }
            dxhcc[n] = hi;
            dxlcc[n] = lo;
        }
        if (!inv3(cell, ci)) return -2;
    } else {
        /* bounding box of the cluster (neilist.F90:313-352) */
        for (int n = 0; n < 3; ++n) { dxhcc[n] = -1.0e10; dxlcc[n] = 1.0e10; }
        for (int a = 0; a < natom; ++a)
            for (int n = 0; n < 3; ++n) {
                if (posin[3 * a + n] > dxhcc[n]) dxhcc[n] = posin[3 * a + n];
                if (posin[3 * a + n] < dxlcc[n]) dxlcc[n] = posin[3 * a + n];
            }
    }
    for (int n = 0; n < 3; ++n) {
        ngup[n] = (int)floor(dxhcc[n] / rcut + NL_EPS);
        nglow[n] = (int)floor(dxlcc[n] / rcut - NL_EPS);
    }

    /* cluster (neilist.F90:354-436) */
    int nsh = periodic ? (hc[0] - lc[0] + 1) * (hc[1] - lc[1] + 1) * (hc[2] - lc[2] + 1) : 1;
    size_t capc = (size_t)nsh * (size_t)(natom > 0 ? natom : 1);
    double *pos = malloc(sizeof(double) * 3 * capc);
    int *org = malloc(sizeof(int) * capc), *shf = malloc(sizeof(int) * 3 * capc);
    double *frac = malloc(sizeof(double) * 3 * (natom > 0 ? natom : 1));
    int nbox = 0;
    if (periodic) {
        for (int a = 0; a < natom; ++a) {
            const double *p = posin + 3 * a;
            for (int k = 0; k < 3; ++k) {
                double v = ci[k] * p[0] + ci[k + 3] * p[1] + ci[k + 6] * p[2];
                frac[3 * a + k] = v - floor(v + NL_EPS);
            }
        }
        for (int a = 0; a < natom; ++a) {
            const double *f = frac + 3 * a;
            for (int n = 0; n < 3; ++n) pos[3 * nbox + n] = c1[n] * f[0] + c2[n] * f[1] + c3[n] * f[2];
            org[nbox] = a;
            shf[3 * nbox] = shf[3 * nbox + 1] = shf[3 * nbox + 2] = 0;
            ++nbox;
        }
        for (int k1 = lc[0]; k1 <= hc[0]; ++k1)
            for (int k2 = lc[1]; k2 <= hc[1]; ++k2)
                for (int k3 = lc[2]; k3 <= hc[2]; ++k3) {
                    if (k1 == 0 && k2 == 0 && k3 == 0) continue;
                    for (int a = 0; a < natom; ++a) {
                        double r1[3] = {frac[3 * a] + (double)k1, frac[3 * a + 1] + (double)k2,
                                        frac[3 * a + 2] + (double)k3};
                        if (!(r1[0] > dxl[0] && r1[0] <= dxh[0] && r1[1] > dxl[1] && r1[1] <= dxh[1] &&
                              r1[2] > dxl[2] && r1[2] <= dxh[2]))
                            continue;
                        double r2[3];
                        for (int n = 0; n < 3; ++n) r2[n] = c1[n] * r1[0] + c2[n] * r1[1] + c3[n] * r1[2];
                        if (!(r2[0] > dxlcc[0] && r2[0] <= dxhcc[0] && r2[1] > dxlcc[1] &&
                              r2[1] <= dxhcc[1] && r2[2] > dxlcc[2] && r2[2] <= dxhcc[2]))
                            continue;
                        memcpy(pos + 3 * nbox, r2, sizeof r2);
                        org[nbox] = a;
                        shf[3 * nbox] = k1; shf[3 * nbox + 1] = k2; shf[3 * nbox + 2] = k3;
                        ++nbox;
                    }
                }
    } else {
        for (int a = 0; a < natom; ++a) {
            memcpy(pos + 3 * nbox, posin + 3 * a, 3 * sizeof(double));
            org[nbox] = a;
            shf[3 * nbox] = shf[3 * nbox + 1] = shf[3 * nbox + 2] = 0;
            ++nbox;
        }
    }

    /* grid (neilist.F90:455-636): boxes of edge rcut, atoms sorted by box, stable */
    int ng[3] = {ngup[0] - nglow[0] + 1, ngup[1] - nglow[1] + 1, ngup[2] - nglow[2] + 1};
    size_t ngb = (size_t)ng[0] * ng[1] * ng[2];
    int *gb = malloc(sizeof(int) * (nbox > 0 ? nbox : 1));
    int *bstart = calloc(ngb + 1, sizeof(int));
    int *batoms = malloc(sizeof(int) * (nbox > 0 ? nbox : 1));
    double inv = 1.0 / rcut;
    int rc = 0;
    for (int c = 0; c < nbox; ++c) {
        int g[3];
        for (int n = 0; n < 3; ++n) {
            g[n] = (int)floor(pos[3 * c + n] * inv) - nglow[n];
            if (g[n] < 0 || g[n] >= ng[n]) rc = -3;  /* the reference would index out of range */
        }
        if (rc) break;
        gb[c] = (g[0] * ng[1] + g[1]) * ng[2] + g[2];
        bstart[gb[c] + 1]++;
    }
    int nei = 0;
    if (!rc) {
        for (size_t b = 0; b < ngb; ++b) bstart[b + 1] += bstart[b];
        int *cur = malloc(sizeof(int) * (ngb > 0 ? ngb : 1));
        memcpy(cur, bstart, sizeof(int) * ngb);
        for (int c = 0; c < nbox; ++c) batoms[cur[gb[c]]++] = c;
        free(cur);

        /* pairs (neilist.F90:693-793) */
        double rcsq = rcut * rcut;
        for (int a = 0; a < natom && !rc; ++a) {
            int b1 = gb[a];
            int h[3] = {b1 / (ng[1] * ng[2]), (b1 / ng[2]) % ng[1], b1 % ng[2]};
            int sten[27], ns = 0;
            sten[ns++] = b1;
            for (int k1 = h[0] - 1; k1 <= h[0] + 1; ++k1) {
                if (k1 < 0 || k1 >= ng[0]) continue;
                for (int k2 = h[1] - 1; k2 <= h[1] + 1; ++k2) {
                    if (k2 < 0 || k2 >= ng[1]) continue;
                    for (int k3 = h[2] - 1; k3 <= h[2] + 1; ++k3) {
                        if (k3 < 0 || k3 >= ng[2]) continue;
                        if (k1 == h[0] && k2 == h[1] && k3 == h[2]) continue;
                        sten[ns++] = (k1 * ng[1] + k2) * ng[2] + k3;
                    }
                }
            }
            const double *r1 = pos + 3 * a;
            const double *m1 = magin + 3 * a;
            double mi = sqrt(m1[0] * m1[0] + m1[1] * m1[1] + m1[2] * m1[2]);
            if (nstart) nstart[a] = nei + 1;
            for (int s = 0; s < ns && !rc; ++s)
                for (int q = bstart[sten[s]]; q < bstart[sten[s] + 1]; ++q) {
                    int c = batoms[q];
                    if (c == a) continue;
                    double v[3] = {pos[3 * c] - r1[0], pos[3 * c + 1] - r1[1], pos[3 * c + 2] - r1[2]};
                    if (!(fabs(v[0]) < rcut && fabs(v[1]) < rcut && fabs(v[2]) < rcut)) continue;
                    if (!(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < rcsq)) continue;
                    if (fabs(v[0]) + fabs(v[1]) + fabs(v[2]) < NL_EPS) { rc = -4; break; }
                    if (nei < cap_pair) {
                        const double *m2 = magin + 3 * org[c];
                        double d = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                        double mm = sqrt(m2[0] * m2[0] + m2[1] * m2[1] + m2[2] * m2[2]);
                        nlist[nei] = c + 1;
                        dist[nei] = d;
                        for (int n = 0; n < 3; ++n) eist[3 * nei + n] = v[n] / d;
                        dist_mom[nei] = mm;
                        if (mm == 0.0) {
                            eist_mom[3 * nei] = 0.0; eist_mom[3 * nei + 1] = 0.0; eist_mom[3 * nei + 2] = 1.0;
                        } else {
                            for (int n = 0; n < 3; ++n) eist_mom[3 * nei + n] = m2[n] / mm;
                        }
                    }
                    ++nei;
                }
            if (nstop) nstop[a] = nei;
            dist_momi[a] = mi;
            if (mi == 0.0) {
                eist_momi[3 * a] = 0.0; eist_momi[3 * a + 1] = 0.0; eist_momi[3 * a + 2] = 1.0;
            } else {
                for (int n = 0; n < 3; ++n) eist_momi[3 * a + n] = m1[n] / mi;
            }
        }
    }
    counts[0] = nbox;
    counts[1] = nei;
    if (!rc && nbox <= cap_box) {
        for (int c = 0; c < nbox; ++c) {
            origin[c] = org[c] + 1;
            for (int n = 0; n < 3; ++n) { nshift[3 * c + n] = shf[3 * c + n]; posbox[3 * c + n] = pos[3 * c + n]; }
        }
    }
    free(pos); free(org); free(shf); free(frac); free(gb); free(bstart); free(batoms);
    if (rc) return rc;
    return (nbox > cap_box || nei > cap_pair) ? 1 : 0;
}
