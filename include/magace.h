/*
 * magace.h -- C-ABI of libmagace.so, the MI355X (gfx950) evaluator of magnetic-ACE
 * energies, forces and magnetic gradients.
 *
 * The first two entry points replace libace's, with identical signatures and semantics,
 * so LAMMPS pair_spin_ace (or any caller of libace) links against libmagace.so unchanged.
 * All scalars are passed by reference, all indices are 1-based, arrays are plain C arrays
 * owned by the caller; nothing is retained between calls.
 */
#ifndef MAGACE_H
#define MAGACE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (the reference has none: it calls Fortran `stop`). */
#define MAGACE_OK 0
#define MAGACE_ERR_INPUT 1  /* bad file, header, sizes, geometry (e.g. very small distance) */
#define MAGACE_ERR_STATE 2  /* called before initialisation */
#define MAGACE_ERR_DEVICE 3 /* HIP runtime error */
#define MAGACE_ERR_CAPACITY 4 /* output capacity too small; the needed size is returned */

/* Precision modes */
#define MAGACE_FP64 0  /* FP64 everywhere; agrees with libace to ~1e-14 relative */
#define MAGACE_MIXED 1 /* FP32 splines/Y_lm/A/B products, FP64 rho, energy, outputs */
/* MAGACE_MIXED with the accumulation recipe: the A and A2 sums over neighbours with Kahan
 * compensation, and rho = sum_k c_k B_k accumulated in FP64 over the (FP32) monomial products
 * with FP64 weights; slower than MAGACE_MIXED, errors against FP64 reported by bench.py */
#define MAGACE_MIXED_ACC 2

/* Replaces libace acelibini_ (src/ace.F90:10-75; caller lammps/pair_spin_ace.cpp:61,288).
 * nele must be 1; elenumbers[0] = atomic number + 1; xcnumber 1 = pbe, 2 = lda; reads
 * "<El>.<xc>.in" from the current directory; smear must match the file to 1e-6.
 * Runs once per process; later calls are no-ops (src/util.F90:29-33).
 * On failure magace_last_error() != 0 (the reference stops the process instead). */
void acelibini_(int *nele, double *rcut, double *rcut_magnetic, double *smear, int *xcnumber,
                int *elenumbers);

/* Replaces libace acelib_ (src/ace.F90:79-321; caller lammps/pair_spin_ace.cpp:62,545).
 * Inputs: natom local atoms, nall >= natom, nei pairs; nstart/nstop[natom] inclusive 1-based
 * pair ranges tiling 1..nei in atom order; nlist (unused by the arithmetic), occ/occlist
 * (element index, must be 1); dlist[nei] = r, elist[3*nei] = (x_i - x_j)/r;
 * dlist_momi[natom], elist_momi[3*natom] = |m_i|, m_i/|m_i|; dlist_mom[nei],
 * elist_mom[3*nei] = |m_j|, m_j/|m_j|.
 * Outputs: eout[natom] = E_i; flist[3*nei] = dE_i/d(x_i - x_j); tlist[3*nei] = dE_i/dm_j;
 * tlist0[3*natom] = dE_i/dm_i.  Caller assembly as pair_spin_ace.cpp:587-641:
 *   f_i -= F_p, f_j += F_p, fm_i -= tlist0_i, fm_j -= tlist_p. */
void acelib_(int *natom, int *nall, int *nei, int *nlist, int *occ, int *occlist, int *nstart,
             int *nstop, double *eout, double *dlist, double *elist, double *flist,
             double *dlist_momi, double *elist_momi, double *dlist_mom, double *elist_mom,
             double *tlist, double *tlist0);

/* Error state of the last call (no counterpart in the reference, which stops). */
int magace_last_error(void);
const char *magace_error_string(void);

/* acelibini_ with an explicit file path and element symbol instead of the CWD lookup
 * (same checks as src/util.F90:525-597, 420-444).  Replaces any previous potential. */
int magace_init_file(const char *path, const char *element, double rcut, double rcut_magnetic,
                     double smear, int xcnumber);
int magace_finalize(void);

/* Select MAGACE_FP64 or MAGACE_MIXED for subsequent evaluations. */
int magace_set_precision(int mode);
int magace_get_precision(void);

/* Kernel tuning (no counterpart in the reference).  The density kernel has variants of its
 * neighbour staging and of the split of its column tasks over idle threads; the choice
 * changes timing and the order of floating-point sums (so results in the last bits), never
 * the arithmetic.  Defaults come from a table compiled into the library per potential header
 * (measured by mag-ace_amd/tune.py), else from a built-in rule.
 * v[4] = {FP64 variant, FP64 split, mixed variant, mixed split}; variant -1 = built-in rule,
 * 0..2; split 0 = automatic, 1..8.
 * magace_tune times MAGACE_TUNE_CANDIDATES (variant, split) candidates on the caller's
 * device-resident system (magace_compute_device arguments, reps evaluations each, on `stream`)
 * for the current precision, writes the average milliseconds per evaluation into ms[] (-1:
 * candidate not available for this potential), keeps the fastest and returns it in best[2]. */
#define MAGACE_TUNE_CANDIDATES 9
int magace_set_tuning(const int *v, int n);
int magace_get_tuning(int *v, int n);
int magace_tune(int natom, int nei, const int *offsets, const double *dlist, const double *elist,
                const double *dlist_momi, const double *elist_momi, const double *dlist_mom,
                const double *elist_mom, double *eout, double *flist, double *tlist, double *tlist0,
                int reps, double *ms, int *best, void *stream);

/* acelib_ on device-resident arrays (same meaning as acelib_, FP64), with 0-based CSR
 * offsets[natom+1] instead of nstart/nstop.  Enqueued on `stream` (a hipStream_t; NULL is the
 * default stream, as in HIP); no host synchronisation.  magace_check_errors() synchronises the
 * stream and reports geometry errors raised by the kernels.  From 2048 atoms the second half of
 * the atoms runs on the library's own stream, forked from `stream` and joined back to it before
 * the call's work on `stream` completes (not while `stream` is being captured into a graph;
 * MAGACE_SPLIT_STREAMS=0 turns this off).  Calls on different streams must not overlap. */
int magace_compute_device(int natom, int nei, const int *offsets, const double *dlist,
                          const double *elist, const double *dlist_momi, const double *elist_momi,
                          const double *dlist_mom, const double *elist_mom, double *eout,
                          double *flist, double *tlist, double *tlist0, void *stream);
int magace_check_errors(void *stream);

/* magace_compute_device without allocating device memory: no counterpart in the reference.
 * For callers that cannot allocate (code under graph capture, tight serving loops).  The
 * per-atom intermediates live in a fixed 128 MiB device array of the library instead of
 * buffers grown with hipMalloc, so natom is limited (about 2490 atoms with the lmax-6 header
 * in mixed precision, 1440 in FP64; 7000 and 4100 with the lmax-4 header); larger natom is
 * rejected with MAGACE_ERR_INPUT and a message giving the limit.  The potential's tables must
 * already be on the current device: magace_init_file/acelibini_ upload them when a GPU is
 * visible (otherwise run magace_compute_device once).  All calls on one GPU share the array:
 * calls that overlap on one GPU (different streams, threads) race. */
int magace_compute_device_noalloc(int natom, int nei, const int *offsets, const double *dlist,
                                  const double *elist, const double *dlist_momi,
                                  const double *elist_momi, const double *dlist_mom,
                                  const double *elist_mom, double *eout, double *flist,
                                  double *tlist, double *tlist0, void *stream);

/* Caller side of the boundary: PairSpinACE::compute (lammps/pair_spin_ace.cpp:340-669) on
 * LAMMPS-shaped arrays.  x[nall][3], sp[nall][4] (unit direction, magnitude), type[nall],
 * map[type] -> element index (must be 0), full neighbour list ilist/numneigh with the
 * neighbours of atom i at firstneigh[firstneigh_offsets[i] ...].  Accumulates into f and fm
 * ([nall][3]), energy, virial[6] (xx yy zz xy xz yz) and eatom[nall] (each may be NULL
 * except f/fm).  Returns 0 or an error code. */
int magace_pair_compute(int nlocal, int nghost, const double *x, const double *sp, const int *type,
                        const int *map, int inum, const int *ilist, const int *numneigh,
                        const int *firstneigh_offsets, const int *firstneigh, int newton_pair,
                        double *f, double *fm, double *energy, double *virial, double *eatom);

/* Per-kernel timing with HIP events recorded on the evaluation stream around each of the
 * three kernels (density, adjoint, forces).  magace_profile(1) starts recording (and clears);
 * magace_profile_read() waits for the recorded evaluations and returns the summed
 * milliseconds per kernel in ms[3] and the number of evaluations in *count. */
int magace_profile(int enable);
int magace_profile_read(double *ms, int *count);

/* Neighbour lists on the device.  Replaces the reference neighbour list
 * (neilist/neilist.F90:148-796: periodic images, grid boxes of edge rcut, 27-box stencil,
 * fastdist) run per structure and concatenated by packstructures (src/neiutil.F90:5-264),
 * with the same pair set, pair order and FP64 values.
 * Host inputs: nstruc structures of natoms[s] atoms, periodic[s] (0 = open cluster),
 * cells[9*s + 3*k + i] = component i of lattice vector k, rcut (1e-6 <= rcut <= 100).
 * Device inputs: pos[3*ntot], mag[3*ntot] (magnetic moment vectors, muB), structures one
 * after another.  Device outputs in the acelib_ conventions, ready for
 * magace_compute_device: offsets[ntot+1] (0-based global CSR), nlist[p] = global 0-based
 * atom index of the neighbour (the periodic image's origin), dlist[p] = r,
 * elist[3p] = (x_i - x_j)/r (the negated reference eist), dlist_mom/elist_mom of the
 * neighbour, dlist_momi/elist_momi of the central atom, and (optional, may be NULL)
 * shifts[3p] = integer lattice shift n of the neighbour's image, i.e. the image sits at
 * x_j + sum_k n_k a_k for the positions as passed (not wrapped), so pair vectors can be
 * recomputed from moved positions between rebuilds.  Host outputs: *npairs, natombox[s]
 * (atoms of the image cluster; may be NULL).  nlist == NULL counts only (offsets filled).
 * Synchronises `stream` once to learn the pair count; returns MAGACE_ERR_CAPACITY (with
 * *npairs set, offsets filled) when capacity < *npairs.  Errors as the reference's stops:
 * rcut out of range, extreme cell shape, identical positions. */
int magace_neighbors_device(int nstruc, const int *natoms, const int *periodic, const double *cells,
                            double rcut, const double *pos, const double *mag, long long capacity,
                            int *offsets, int *nlist, double *dlist, double *elist, double *dlist_mom,
                            double *elist_mom, double *dlist_momi, double *elist_momi, int *shifts,
                            int *npairs, int *natombox, void *stream);
/* packstructures (src/neiutil.F90:5-264) on host arrays, computed on the device, returned in
 * the reference's conventions: strucaddress[2*s] = 1-based first atom, [2*s+1] = 1-based
 * first pair; n2startpack/n2stoppack[ntot] 1-based ranges local to the structure;
 * n2listpack[p] = origin(nlist) (1-based, local); d2listpack, e2listpack (x_j - x_i)/d,
 * d2listpack_mom, e2listpack_mom, d2listpack_momi, e2listpack_momi; natombox[s] (may be
 * NULL); *npairs.  MAGACE_ERR_CAPACITY when capacity < *npairs. */
int magace_packstructures(int nstruc, const int *natoms, const int *periodic, const double *cells,
                          double rcut, const double *atompos, const double *magmom, long long capacity,
                          int *strucaddress, int *n2startpack, int *n2stoppack, int *n2listpack,
                          double *d2listpack, double *e2listpack, double *d2listpack_mom,
                          double *e2listpack_mom, double *d2listpack_momi, double *e2listpack_momi,
                          int *natombox, int *npairs);

/* Device-resident spin-lattice dynamics around the evaluator (replaces the LAMMPS-side loop
 * around pair_spin_ace: PairSpinACE::compute's assembly, lammps/pair_spin_ace.cpp:571-641, and
 * the per-spin acelib_ sweep of compute_single_pair, :677-986).  Velocity Verlet for the
 * lattice (metal units: A, ps, amu, eV) and rotations of the spins about
 * Omega_i = -(2/hbar) dE/dm_i with the field extrapolated to mid-step (second order), one
 * evaluation per step, neighbour lists (cutoff rlist =
 * r_c + skin) rebuilt every rebuild_every steps, the steps in between captured once into a
 * HIP graph when use_graph != 0.  Periodic cell[3k + i] = component i of lattice vector k;
 * x, v [3*natom]; sdir [3*natom] spin directions (normalised here), smag [natom] moment
 * magnitudes (muB, constant).  Needs the potential loaded (acelibini_ / magace_init_file);
 * uses the current precision mode. */
int magace_md_create(int natom, const double *cell, double rlist, double mass, double dt,
                     int rebuild_every, const double *x, const double *v, const double *sdir,
                     const double *smag, void **md);
int magace_md_run(void *md, int nsteps, int use_graph);
/* Copies the state back (any output may be NULL): x, v, sdir, f (forces), dEdm [3*natom],
 * eatom [natom]; epot = sum eatom and ekin (eV); virial[9] = W_ab (row-major) = sum over
 * pairs of u_a F_b with u = x_i - x_j and F = dE_i/du (eV), the tally of
 * lammps/pair_spin_ace.cpp:635-636; epot and virial are reduced on the device in FP64 in a
 * fixed order; counters = {steps, rebuilds, graph captures, pairs}. */
int magace_md_read(void *md, double *x, double *v, double *sdir, double *f, double *dEdm,
                   double *eatom, double *epot, double *ekin, double *virial, long long *counters);
int magace_md_destroy(void *md);

/* Fitting (FP64).  Coefficients are ordered as the potential file's B functions: c1 (order
 * 1, nradbase_magnetic of them), then orders 2 to 6 in enumeration order, each order
 * density-major with two densities; the reference's c0..c6 arrays (src/modules.F90) and its
 * writer (src/util.F90:1110-1362) use the same order.
 * magace_fit_set_coefficients recomputes the merged monomial weights (the contraction of
 * src/forces.F90:93-141) and takes effect for every later evaluation.
 * magace_fit_forward: an FP64 evaluation on device arrays (acelib_ meaning, 0-based CSR
 * offsets, nlist[p] = global 0-based neighbour atom) that also assembles per-atom forces
 * forces[3*natom] and per-structure energies estruct[nstruc] (atom_struct[natom] = structure
 * of each atom; may be NULL when nstruc = 0).
 * magace_fit_backward: right after magace_fit_forward on the same arrays, the gradient
 * grad[ncoef] = d/dc (sum_s gE[s] E_s + sum_j gF[3j..] . F_j) for cotangents gE (may be NULL)
 * and gF (device arrays), by forward-mode differentiation of the density along gF. */
int magace_fit_ncoef(void);
int magace_fit_get_coefficients(double *c, int n);
int magace_fit_set_coefficients(const double *c, int n);
int magace_fit_forward(int natom, int nei, const int *offsets, const int *nlist, const double *dlist,
                       const double *elist, const double *dlist_momi, const double *elist_momi,
                       const double *dlist_mom, const double *elist_mom, const int *atom_struct,
                       int nstruc, double *estruct, double *forces, void *stream);
int magace_fit_backward(int natom, int nei, const int *offsets, const int *nlist, const double *dlist,
                        const double *elist, const double *dlist_momi, const double *dlist_mom,
                        const double *elist_mom, const int *atom_struct, const double *gE,
                        const double *gF, double *grad, void *stream);
/* Writes the loaded potential with its current coefficients (same header; records par,
 * core, rad_atomic, rad_magnetic, clu as src/util.F90:1110-1362 writes them). */
int magace_write_potential(const char *path);

/* Basis introspection: counts of B functions of orders 1..6 of the loaded potential (the
 * numbers src/util.F90:1831,1929,2235,2596,3095,3760 print); writes min(n, 6) of them. */
int magace_basis_counts(int *nvar, int n);
/* Enumerate the basis of a header without a file: hdr[20] = nb_a nb_m nr2_m l2m nr3_a nr3_m
 * l3a l3m nr4_a nr4_m l4a l4m nr5_a nr5_m l5a l5m nr6_a nr6_m l6a l6m.  Writes index tuples
 * (b1index .. b6index columns) of order 1..6 when cap suffices; returns the count, or a
 * negative error code. */
int magace_enumerate(const int *hdr, int order, int *out, int cap);
/* info[0..12] = nz, merged terms, raw m-combinations, raw order 2/3/4, shape class, oA2,
 * gradient contributions, gradient chunks, raw order 5/6, products of two A entries */
int magace_model_info(long long *info, int n);
/* The merged monomial table rho = Re sum_t w_t prod_s Z[idx[4t+s]] (Z[nz] = 1) that replaces
 * the per-B m-combination loops of src/b_basis.F90:91-307; copies when cap >= count and
 * returns the count.  Host-side introspection; needs no GPU. */
int magace_model_terms(unsigned short *idx, double *w, int cap);
/* The transposed gradient table the adjoint kernel streams (dRho/dZ[x] gathered per target
 * x, the gather form of src/forces.F90:84-349): chunks[2c] = first record, chunks[2c+1] =
 * x | count << 12 | sole << 18 | slot << 19 (unsigned; count = the length of the chunk's
 * group of 64, sole = the target's only chunk, slot = the partial-sum slot of the chunk's run
 * of consecutive same-target chunks on one adjoint thread); record r holds the factors
 * idx[4r..4r+2], indices into the extended vector Z' (magace_model_products), and the weight
 * w[r] (density 1); a chunk contributes sum_j w Z'[a_j] Z'[b_j] Z'[c_j] over its count
 * records, 64 apart.  Copies when the caps suffice; returns the record count (chunk count in
 * *nchunks).  Host-side introspection. */
int magace_model_records(int *chunks, int cap_chunks, unsigned short *idx, double *w, int cap, int *nchunks);
/* The products Y of the extended vector Z' = [Z | 1 | Y] that the monomials of five and six
 * factors (body orders 5, 6) and the gradient records address: entry nz + 1 + i =
 * Z[pairs[2i]] Z[pairs[2i+1]].  Copies when cap >= the count; returns the count.  Host-side
 * introspection. */
int magace_model_products(int *pairs, int cap);

#ifdef __cplusplus
}
#endif
#endif
