/*
 * ref_driver.c -- test-infrastructure driver around the reference libace.
 *
 * This program is part of the oracle (the checker), never of the product.
 * It links the reference Fortran objects compiled from their own sources
 * (see oracle/Makefile, outputs in oracle/_ref/) and calls the two C-ABI
 * entry points exactly the way LAMMPS pair_spin_ace does
 * (lammps/pair_spin_ace.cpp:59-62 declarations, :288 init, :545 compute):
 *
 *   acelibini_(&nele,&rcut,&rcut_magnetic,&smear,&xcnumber,elenumbers)
 *   acelib_(&natom,&nall,&nei,nlist,occ,occlist,nstart,nstop,eout,dlist,
 *           elist,flist,dlist_momi,elist_momi,dlist_mom,elist_mom,tlist,tlist0)
 *
 * Usage: ace_ref_driver <potential_dir> <input.bin> <output.bin>
 *
 * The reference reads "<El>.<xc>.in" from the current directory
 * (src/util.F90:525-531), so the driver chdir()s into <potential_dir>.
 *
 * input.bin (little endian):
 *   int32  magic = 0x4D414345 ('MACE'), natom, nall, nei, nreps, xcnumber, elenumber
 *   double rcut, rcut_magnetic, smear
 *   int32  nstart[natom], nstop[natom], nlist[nei], occ[natom], occlist[nei]  (1-based)
 *   double dlist[nei], elist[3*nei], dlist_momi[natom], elist_momi[3*natom],
 *          dlist_mom[nei], elist_mom[3*nei]
 * output.bin:
 *   double eout[natom], flist[3*nei], tlist[3*nei], tlist0[3*natom],
 *          seconds[nreps]  (wall time of each acelib_ call; init excluded)
 */
#define _POSIX_C_SOURCE 200809L
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <unistd.h>

extern void acelibini_(int *, double *, double *, double *, int *, int *);
extern void acelib_(int *, int *, int *, int *, int *, int *, int *, int *, double *,
                    double *, double *, double *, double *, double *, double *, double *,
                    double *, double *);

static void *xread(FILE *f, size_t n, size_t sz, const char *what) {
    void *p = malloc(n * sz + 16);
    if (!p) { fprintf(stderr, "ref_driver: out of memory (%s)\n", what); exit(3); }
    if (n && fread(p, sz, n, f) != n) { fprintf(stderr, "ref_driver: short read (%s)\n", what); exit(3); }
    return p;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s <potential_dir> <input.bin> <output.bin>\n", argv[0]);
        return 2;
    }
    FILE *fi = fopen(argv[2], "rb");
    if (!fi) { perror("input"); return 2; }
    int32_t hdr[7];
    if (fread(hdr, sizeof(int32_t), 7, fi) != 7 || hdr[0] != 0x4D414345) {
        fprintf(stderr, "ref_driver: bad header\n");
        return 2;
    }
    int natom = hdr[1], nall = hdr[2], nei = hdr[3], nreps = hdr[4], xc = hdr[5], ele = hdr[6];
    double cuts[3];
    if (fread(cuts, sizeof(double), 3, fi) != 3) { fprintf(stderr, "ref_driver: bad header\n"); return 2; }
    int *nstart = xread(fi, natom, 4, "nstart"), *nstop = xread(fi, natom, 4, "nstop");
    int *nlist = xread(fi, nei, 4, "nlist"), *occ = xread(fi, natom, 4, "occ");
    int *occlist = xread(fi, nei, 4, "occlist");
    double *dlist = xread(fi, nei, 8, "dlist"), *elist = xread(fi, 3 * (size_t)nei, 8, "elist");
    double *dmi = xread(fi, natom, 8, "dlist_momi"), *emi = xread(fi, 3 * (size_t)natom, 8, "elist_momi");
    double *dm = xread(fi, nei, 8, "dlist_mom"), *em = xread(fi, 3 * (size_t)nei, 8, "elist_mom");
    fclose(fi);

    double *eout = calloc(natom + 1, 8), *flist = calloc(3 * (size_t)nei + 1, 8);
    double *tlist = calloc(3 * (size_t)nei + 1, 8), *tlist0 = calloc(3 * (size_t)natom + 1, 8);
    double *secs = calloc(nreps + 1, 8);

    /* resolve the output path before changing directory */
    char outpath[4096];
    if (argv[3][0] == '/') {
        snprintf(outpath, sizeof outpath, "%s", argv[3]);
    } else {
        char cwd[2048];
        if (!getcwd(cwd, sizeof cwd)) { perror("getcwd"); return 2; }
        snprintf(outpath, sizeof outpath, "%s/%s", cwd, argv[3]);
    }
    if (chdir(argv[1]) != 0) { perror("chdir"); return 2; }

    int nele = 1;
    acelibini_(&nele, &cuts[0], &cuts[1], &cuts[2], &xc, &ele);
    for (int r = 0; r < nreps; ++r) {
        double t0 = now_s();
        acelib_(&natom, &nall, &nei, nlist, occ, occlist, nstart, nstop, eout, dlist, elist,
                flist, dmi, emi, dm, em, tlist, tlist0);
        secs[r] = now_s() - t0;
    }

    FILE *fo = fopen(outpath, "wb");
    if (!fo) { perror("output"); return 2; }
    fwrite(eout, 8, natom, fo);
    fwrite(flist, 8, 3 * (size_t)nei, fo);
    fwrite(tlist, 8, 3 * (size_t)nei, fo);
    fwrite(tlist0, 8, 3 * (size_t)natom, fo);
    fwrite(secs, 8, nreps, fo);
    fclose(fo);
    return 0;
}
