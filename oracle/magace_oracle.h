/*
 * magace_oracle.h -- CPU FP64 restatement of the reference libace energy / force /
 * magnetic-gradient path.  TEST INFRASTRUCTURE ONLY: the checker for the HIP product,
 * never linked into it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.
 *
 * Every function cites the reference file:line it restates.  Single element only (the
 * shipped libace refuses more: src/ace.F90:46-49); body order <= 4 (B5/B6 out of scope).
 */
#ifndef MAGACE_ORACLE_H
#define MAGACE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Potential-file header (src/util.F90:603-653). */
typedef struct {
    int nb_a, nb_m, nr2_m, l2m;   /* "nradbase l2max" */
    int nr3_a, nr3_m, l3a, l3m;   /* "nradial3 l3max" */
    int nr4_a, nr4_m, l4a, l4m;   /* "nradial4 l4max" */
    int nr5_a, nr5_m, l5a, l5m;   /* "nradial5 l5max" */
    int nr6_a, nr6_m, l6a, l6m;   /* "nradial6 l6max" */
    int ndensity, npot, nfemb, radtype_a, radtype_m;
    double lambda_a, lambda_m, cut_a, cut_m, dcut_a, dcut_m, ecut, decut;
} orc_header;

typedef struct orc_pot orc_pot;

const char *orc_error(void);

/* Build tables (enumeration, CG, couplings) for a header; no coefficients yet. */
orc_pot *orc_create(const orc_header *h);
/* acelibini_ semantics: read <path> (header then parameters), check smearing/xc/cutoffs,
 * build spline tables.  xc: 1 pbe, 2 lda.  Returns NULL on error (see orc_error). */
orc_pot *orc_load(const char *path, const char *element, double rcut, double rcut_magnetic,
                  double smear, int xc);
void orc_free(orc_pot *p);

/* nvar[0..3] = number of B functions of orders 1..4 */
void orc_counts(const orc_pot *p, int nvar[6]);
/* index tuple (Fortran b*index column, 1-based radial, l values) of B function k (0-based) */
int orc_basis_index(const orc_pot *p, int order, int k, int *idx);

/* acelib_ semantics on 1-based nstart/nstop (src/ace.F90:79-321). Returns 0 or error. */
int orc_acelib(orc_pot *p, int natom, int nei, const int *nstart, const int *nstop,
               const double *dlist, const double *elist, const double *dlist_momi,
               const double *elist_momi, const double *dlist_mom, const double *elist_mom,
               double *eout, double *flist, double *tlist, double *tlist0, int nthreads);

/* Also return per-atom B vectors (b[natom][nvar1+nvar2+nvar3+nvar4]) */
int orc_bvectors(orc_pot *p, int natom, int nei, const int *nstart, const int *nstop,
                 const double *dlist, const double *elist, const double *dlist_momi,
                 const double *elist_momi, const double *dlist_mom, const double *elist_mom,
                 double *bout);

/* Building blocks, exposed for closed-form tests */
double orc_cg(int j1, int j2, int J, int m1, int m2, int M);
/* Y_lm and tangential gradient for m in -l..l, all l<=lmax, layout [l*l+l+m] */
void orc_ylm(int lmax, double x, double y, double z, double *yre, double *yim,
             double *dyre /*[3*(lmax+1)^2]*/, double *dyim);
void orc_radbase(int radtype, double lambda, int nb, double cut, double dcut, double r,
                 double *g, double *dg);
void orc_radcore(double r, double pre, double lambda, double cut, double *cr, double *dcr);
/* spline tables exposed for checks: which=0 atomic g, 1 atomic R, 2 core, 3 magnetic g,
 * 4 magnetic R.  Evaluate value and derivative of all functions of that family at r. */
int orc_spline_eval(const orc_pot *p, int which, double r, double *val, double *dval);
double orc_cutoff(const orc_pot *p, int which); /* 0 atomic global cutoff, 1 magnetic */

/* Reference neighbour list of one structure (neilist/neilist.F90:148-796).
 * cell[i + 3k] = component i of lattice vector k; posin/magin [3*natom].
 * Outputs (reference conventions, 1-based): nstart/nstop[natom], nlist[nei] (cluster
 * index), origin[natombox], nshift[3*natombox], dist/dist_mom[nei], eist/eist_mom[3*nei]
 * with eist = (x_j - x_i)/d, dist_momi[natom], eist_momi[3*natom], posbox[3*natombox].
 * counts = {natombox, nei}.  Returns 0, 1 if a capacity is too small (counts hold the
 * sizes), or < 0 on an input the reference stops on. */
int orc_neilist(int natom, int periodic, double rcut, const double *cell, const double *posin,
                const double *magin, int cap_box, int cap_pair, int *counts, int *nstart,
                int *nstop, int *nlist, int *origin, int *nshift, double *dist, double *eist,
                double *dist_mom, double *eist_mom, double *dist_momi, double *eist_momi,
                double *posbox);

#ifdef __cplusplus
}
#endif
#endif
