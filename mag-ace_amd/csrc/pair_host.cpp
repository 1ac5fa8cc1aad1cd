// pair_host.cpp -- the caller side of the drop-in boundary: a C++ restatement of LAMMPS
// PairSpinACE::compute (lammps/pair_spin_ace.cpp:340-669) on LAMMPS-shaped arrays.
//
// It flattens a full neighbour list into the acelib_ arrays exactly as the plugin does
// (:419-543), calls acelib_ (the libmagace implementation), and scatters the results the
// same way (:571-641): f_i -= F_p, f_j += F_p, fm_i -= T0_i, fm_j -= T_p, energy via
// ev_tally_full(i, 2*eout_i) (= +eout_i), virial via ev_tally_xyz(i, j, ..., F_p, x_i - x_j).
// Deliberate differences: the plugin's pre-scatter zeroing of f/fm (:548-567, which uses a
// stale jnum and erases other pair styles' forces) is not reproduced, and errors are
// returned instead of exit(0).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/magace.h"

namespace {
constexpr int kNeighMask = 0x1FFFFFFF;  // LAMMPS NEIGHMASK (special-bond bits stripped)
}

extern "C" int magace_pair_compute(int nlocal, int nghost, const double* x, const double* sp,
                                   const int* type, const int* map, int inum, const int* ilist,
                                   const int* numneigh, const int* firstneigh_offsets,
                                   const int* firstneigh, int newton_pair, double* f, double* fm,
                                   double* energy, double* virial, double* eatom) {
    const int nall = nlocal + nghost;
    if (inum != nlocal) return MAGACE_ERR_INPUT;  // "inum: %d nlocal: %d are different."
    int nei = 0;
    for (int ii = 0; ii < inum; ++ii) nei += numneigh[ilist[ii]];
    std::vector<int> nstart(nlocal), nstop(nlocal), occ(nlocal), nlist(nei), occlist(nei);
    std::vector<double> dlist(nei), elist(3 * (size_t)nei), flist(3 * (size_t)nei + 1), tlist(3 * (size_t)nei + 1);
    std::vector<double> eout(nlocal + 1), tlist0(3 * (size_t)nlocal + 1), dlist_mom(nei), elist_mom(3 * (size_t)nei);
    std::vector<double> dlist_momi(nlocal), elist_momi(3 * (size_t)nlocal);
    int neicount = 0;
    for (int ii = 0; ii < inum; ++ii) {
        const int i = ilist[ii];
        occ[ii] = map[type[i]];
        if (occ[ii] != 0) return MAGACE_ERR_INPUT;  // "Cannot handle multiple species."
        nstart[ii] = neicount;
        dlist_momi[ii] = sp[4 * i + 3];
        for (int c = 0; c < 3; ++c) elist_momi[3 * ii + c] = sp[4 * i + c];
        const int* jlist = firstneigh + firstneigh_offsets[i];
        for (int jj = 0; jj < numneigh[i]; ++jj) {
            const int j = jlist[jj] & kNeighMask;
            if (j >= nall) return MAGACE_ERR_INPUT;  // "Atoms index too large."
            nlist[neicount] = j;
            occlist[neicount] = map[type[j]];
            const double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
            const double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const double r = std::sqrt(rsq);
            if (r < 1e-6) return MAGACE_ERR_INPUT;  // "Atoms with identical positions."
            const double inorm = 1.0 / r;
            dlist[neicount] = r;
            for (int c = 0; c < 3; ++c) elist[3 * (size_t)neicount + c] = inorm * d[c];
            dlist_mom[neicount] = sp[4 * j + 3];
            for (int c = 0; c < 3; ++c) elist_mom[3 * (size_t)neicount + c] = sp[4 * j + c];
            ++neicount;
        }
        nstop[ii] = neicount - 1;
    }
    for (int p = 0; p < neicount; ++p) { nlist[p] += 1; occlist[p] += 1; }
    for (int ii = 0; ii < nlocal; ++ii) { nstart[ii] += 1; nstop[ii] += 1; occ[ii] += 1; }
    int natom = nlocal, nall_ = nall, nei_ = nei;
    acelib_(&natom, &nall_, &nei_, nlist.data(), occ.data(), occlist.data(), nstart.data(), nstop.data(),
            eout.data(), dlist.data(), elist.data(), flist.data(), dlist_momi.data(), elist_momi.data(),
            dlist_mom.data(), elist_mom.data(), tlist.data(), tlist0.data());
    if (int rc = magace_last_error()) return rc;
    int k = 0;
    for (int ii = 0; ii < inum; ++ii) {
        const int i = ilist[ii];
        for (int c = 0; c < 3; ++c) fm[3 * i + c] -= tlist0[3 * ii + c];
        const int* jlist = firstneigh + firstneigh_offsets[i];
        for (int jj = 0; jj < numneigh[i]; ++jj, k += 3) {
            const int j = jlist[jj] & kNeighMask;
            const double F[3] = {flist[k], flist[k + 1], flist[k + 2]};
            const double d[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
            for (int c = 0; c < 3; ++c) f[3 * i + c] -= F[c];
            if (newton_pair || j < nlocal) {
                for (int c = 0; c < 3; ++c) {
                    f[3 * j + c] += F[c];
                    fm[3 * j + c] -= tlist[k + c];
                }
            }
            if (virial) {  // ev_tally_xyz with fx,fy,fz = F_p and del = x_i - x_j
                const double v[6] = {d[0] * F[0], d[1] * F[1], d[2] * F[2], d[0] * F[1], d[0] * F[2], d[1] * F[2]};
                const double w = newton_pair ? 1.0 : 0.5 * ((i < nlocal) + (j < nlocal));
                for (int c = 0; c < 6; ++c) virial[c] += w * v[c];
            }
        }
        if (energy) *energy += eout[ii];
        if (eatom) eatom[i] += eout[ii];
    }
    return 0;
}
