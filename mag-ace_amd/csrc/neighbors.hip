// neighbors.hip -- neighbour lists on the device with the pair sets, pair order and FP64
// values of the reference neighbour list (neilist/neilist.F90:148-796), for one structure
// or for a batch of structures packed one after another as packstructures packs them
// (src/neiutil.F90:5-264).
//
// Pipeline (one stream, one host synchronisation to learn the pair count):
//   k_nb_bbox     bounding box of open clusters                neilist.F90:313-336
//   k_nb_frac     fractional coordinates wrapped into the cell neilist.F90:354-368
//   k_nb_flag     image candidates inside the slab of width rc neilist.F90:382-434
//   scan          cluster index of every kept candidate (stable: identity block first,
//                 then shifts in (k1,k2,k3) order, atoms ascending inside a shift)
//   k_nb_cluster  cluster positions + grid box of edge rc       neilist.F90:385-421, 576-588
//   scan, k_nb_bfill, k_nb_bsort   atoms grouped by box, ascending cluster index inside a
//                 box (the reference's counting sort order)     neilist.F90:591-604
//   k_nb_pairs<0> one wavefront per central atom: 27-box stencil in the reference order
//                 (own box, then (k1,k2,k3) ascending), fastdist, ballot-compacted count
//   scan          CSR offsets
//   k_nb_pairs<1> same walk, writes the pairs in order          neilist.F90:693-793
//
// Contraction of multiply-add is off in this file: every FP64 expression is evaluated in
// the reference's operation order with IEEE rounding, so positions, distances and unit
// vectors are bit-identical to the Fortran build and the pair sets agree exactly at the
// cutoff.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/magace.h"
#include "magace.hpp"

namespace magace {
namespace {

constexpr double kEps = 1.0e-9;  // neilist/neilist.F90:9
constexpr int kScanThreads = 256, kScanItems = 4, kScanTile = kScanThreads * kScanItems;

enum : int { kNbErrBox = 1, kNbErrIdentical = 2 };

// Geometry of one structure, computed on the host from its cell (and, for open clusters,
// from the bounding box measured on the device).
struct NbGeom {
    double cell[9];  // cell[i + 3k] = component i of lattice vector k
    double ci[9];    // inverse, same layout
    double dxl[3], dxh[3];    // fractional window of images (lower open, upper closed)
    double lcc[3], hcc[3];    // Cartesian window
    int lc[3], nsh[3];        // image shifts lc..lc+nsh-1 per direction
    int nglow[3], ng[3];      // grid boxes nglow..nglow+ng-1 per direction
    int periodic, natom, atom_off, box_off;
    int zero_shift;           // lexicographic index of shift (0,0,0)
    int pad;
    long long cand_off;       // first image candidate of this structure
};

__device__ __forceinline__ int find_struct(const int* __restrict__ off, int n, int a) {
    // largest s with off[s] <= a (off has n+1 entries, off[0] = 0)
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= a) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int find_struct64(const long long* __restrict__ off, int n, long long i) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------- exclusive scan (int32)

__device__ __forceinline__ int block_exclusive(int v, int* sh, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    int off = 0;
    total = 0;
    for (int k = 0; k < kScanThreads / 64; ++k) {
        if (k < w) off += sh[k];
        total += sh[k];
    }
    __syncthreads();
    return off + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const int* __restrict__ in, long long n,
                                                             int* __restrict__ sums) {
    __shared__ int sh[kScanThreads / 64];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    int s = 0;
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += in[base + k];
    int total;
    (void)block_exclusive(s, sh, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// exclusive scan of the tile sums in place; sums[ntiles] and out[n] receive the total
__global__ __launch_bounds__(kScanThreads) void k_scan_sums(int* __restrict__ sums, int ntiles,
                                                            int* __restrict__ out, long long n) {
    __shared__ int sh[kScanThreads / 64];
    int carry = 0;
    for (int b = 0; b < ntiles; b += kScanThreads) {
        const int i = b + threadIdx.x;
        const int v = i < ntiles ? sums[i] : 0;
        int total;
        const int e = block_exclusive(v, sh, total);
        if (i < ntiles) sums[i] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0) {
        sums[ntiles] = carry;
        out[n] = carry;
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const int* __restrict__ in, long long n,
                                                             const int* __restrict__ sums,
                                                             int* __restrict__ out) {
    __shared__ int sh[kScanThreads / 64];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    int total;
    int e = block_exclusive(s, sh, total) + sums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = e;
        e += v[k];
    }
}

__global__ void k_fill_int(int* __restrict__ p, long long n, int v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- kernels of the build

// min / max of the positions of each open cluster (one block per structure)
__global__ __launch_bounds__(256) void k_nb_bbox(const NbGeom* __restrict__ G, const double* __restrict__ pos,
                                                 double* __restrict__ bbox) {
    __shared__ double sh[6][256];
    const NbGeom& g = G[blockIdx.x];
    double hi[3] = {-1.0e10, -1.0e10, -1.0e10}, lo[3] = {1.0e10, 1.0e10, 1.0e10};
    if (!g.periodic)
        for (int a = threadIdx.x; a < g.natom; a += blockDim.x)
            for (int n = 0; n < 3; ++n) {
                const double x = pos[3 * (size_t)(g.atom_off + a) + n];
                hi[n] = x > hi[n] ? x : hi[n];
                lo[n] = x < lo[n] ? x : lo[n];
            }
    for (int n = 0; n < 3; ++n) {
        sh[n][threadIdx.x] = hi[n];
        sh[3 + n][threadIdx.x] = lo[n];
    }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int n = 0; n < 3; ++n) {
                const double h2 = sh[n][threadIdx.x + st], l2 = sh[3 + n][threadIdx.x + st];
                if (h2 > sh[n][threadIdx.x]) sh[n][threadIdx.x] = h2;
                if (l2 < sh[3 + n][threadIdx.x]) sh[3 + n][threadIdx.x] = l2;
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) bbox[6 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// fractional coordinates wrapped into [0,1) (periodic) -- neilist.F90:356-367
// afl = the integer part removed by the wrap (image bookkeeping for pair shifts)
__global__ void k_nb_frac(const NbGeom* __restrict__ G, const int* __restrict__ atom_off, int nstruc, int ntot,
                          const double* __restrict__ pos, double* __restrict__ frac, int* __restrict__ afl) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ntot) return;
    const NbGeom& g = G[find_struct(atom_off, nstruc, a)];
    const double p0 = pos[3 * (size_t)a], p1 = pos[3 * (size_t)a + 1], p2 = pos[3 * (size_t)a + 2];
    for (int k = 0; k < 3; ++k) {
        if (g.periodic) {
            const double v = g.ci[k] * p0 + g.ci[k + 3] * p1 + g.ci[k + 6] * p2;
            const double fl = floor(v + kEps);
            frac[3 * (size_t)a + k] = v - fl;
            afl[3 * (size_t)a + k] = (int)fmax(fmin(fl, 1.0e9), -1.0e9);
        } else {
            frac[3 * (size_t)a + k] = k == 0 ? p0 : (k == 1 ? p1 : p2);
            afl[3 * (size_t)a + k] = 0;
        }
    }
}

struct Cand {
    int s, a, k0, k1, k2;
    bool identity;
};

__device__ __forceinline__ Cand decode(const NbGeom* __restrict__ G, const long long* __restrict__ cand_off,
                                       int nstruc, long long i) {
    Cand c;
    c.s = find_struct64(cand_off, nstruc, i);
    const NbGeom& g = G[c.s];
    const long long loc = i - g.cand_off;
    const int b = (int)(loc / g.natom);
    const int n = (int)(loc - (long long)b * g.natom);
    c.a = g.atom_off + n;
    c.identity = b == 0;
    c.k0 = c.k1 = c.k2 = 0;
    if (!c.identity) {
        int e = b - 1;
        if (e >= g.zero_shift) ++e;
        c.k0 = g.lc[0] + e / (g.nsh[1] * g.nsh[2]);
        c.k1 = g.lc[1] + (e / g.nsh[2]) % g.nsh[1];
        c.k2 = g.lc[2] + e % g.nsh[2];
    }
    return c;
}

// image candidate kept?  neilist.F90:394-418
__global__ void k_nb_flag(const NbGeom* __restrict__ G, const long long* __restrict__ cand_off, int nstruc,
                          long long ncand, const double* __restrict__ frac, int* __restrict__ flag) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncand) return;
    const Cand c = decode(G, cand_off, nstruc, i);
    int keep = 1;
    if (!c.identity) {
        const NbGeom& g = G[c.s];
        const double r1[3] = {frac[3 * (size_t)c.a] + (double)c.k0, frac[3 * (size_t)c.a + 1] + (double)c.k1,
                              frac[3 * (size_t)c.a + 2] + (double)c.k2};
        keep = r1[0] > g.dxl[0] && r1[0] <= g.dxh[0] && r1[1] > g.dxl[1] && r1[1] <= g.dxh[1] &&
               r1[2] > g.dxl[2] && r1[2] <= g.dxh[2];
        if (keep) {
            double r2[3];
            for (int n = 0; n < 3; ++n) r2[n] = g.cell[n] * r1[0] + g.cell[n + 3] * r1[1] + g.cell[n + 6] * r1[2];
            keep = r2[0] > g.lcc[0] && r2[0] <= g.hcc[0] && r2[1] > g.lcc[1] && r2[1] <= g.hcc[1] &&
                   r2[2] > g.lcc[2] && r2[2] <= g.hcc[2];
        }
    }
    flag[i] = keep;
}

// cluster arrays and grid box of every kept candidate -- neilist.F90:385-421, 576-588
__global__ void k_nb_cluster(const NbGeom* __restrict__ G, const long long* __restrict__ cand_off, int nstruc,
                             long long ncand, const double* __restrict__ frac, const int* __restrict__ flag,
                             const int* __restrict__ cidx, double invrc, double* __restrict__ cpos,
                             int* __restrict__ corg, int* __restrict__ cbox, int* __restrict__ cshift,
                             int* __restrict__ bcount, int* __restrict__ err) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncand || !flag[i]) return;
    const Cand c = decode(G, cand_off, nstruc, i);
    const NbGeom& g = G[c.s];
    const int ci = cidx[i];
    double r2[3];
    if (g.periodic) {
        // identity atoms: cell * wrapped fraction; images: cell * (fraction + shift)
        const double* f = frac + 3 * (size_t)c.a;
        double r1[3] = {f[0], f[1], f[2]};
        if (!c.identity) {
            r1[0] = f[0] + (double)c.k0;
            r1[1] = f[1] + (double)c.k1;
            r1[2] = f[2] + (double)c.k2;
        }
        for (int n = 0; n < 3; ++n) r2[n] = g.cell[n] * r1[0] + g.cell[n + 3] * r1[1] + g.cell[n + 6] * r1[2];
    } else {
        for (int n = 0; n < 3; ++n) r2[n] = frac[3 * (size_t)c.a + n];
    }
    int box = -1;
    double gb[3];
    bool inside = true;
    for (int n = 0; n < 3; ++n) {
        gb[n] = floor(r2[n] * invrc) - (double)g.nglow[n];
        inside = inside && gb[n] >= 0.0 && gb[n] < (double)g.ng[n];
    }
    if (inside)
        box = g.box_off + ((int)gb[0] * g.ng[1] + (int)gb[1]) * g.ng[2] + (int)gb[2];
    else
        atomicOr(err, kNbErrBox);
    for (int n = 0; n < 3; ++n) cpos[3 * (size_t)ci + n] = r2[n];
    cshift[3 * (size_t)ci] = c.k0;
    cshift[3 * (size_t)ci + 1] = c.k1;
    cshift[3 * (size_t)ci + 2] = c.k2;
    corg[ci] = c.a;
    cbox[ci] = box;
    if (box >= 0) atomicAdd(&bcount[box], 1);
}

__global__ void k_nb_bfill(const int* __restrict__ cidx_total, const int* __restrict__ cbox,
                           const int* __restrict__ bstart, int* __restrict__ bcur, int* __restrict__ batoms) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= *cidx_total) return;
    const int b = cbox[c];
    if (b < 0) return;
    const int slot = atomicAdd(&bcur[b], 1);
    batoms[bstart[b] + slot] = c;
}

// ascending cluster index inside each box (the reference fills boxes in cluster order)
__global__ void k_nb_bsort(int nbox, const int* __restrict__ bstart, int* __restrict__ batoms) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbox) return;
    const int lo = bstart[b], hi = bstart[b + 1];
    for (int i = lo + 1; i < hi; ++i) {
        const int v = batoms[i];
        int j = i - 1;
        while (j >= lo && batoms[j] > v) {
            batoms[j + 1] = batoms[j];
            --j;
        }
        batoms[j + 1] = v;
    }
}

struct PairOut {
    const int* offsets;
    int* nlist;
    double *dl, *el, *dm, *em, *dmi, *emi;
    int* shifts;  // optional: lattice shift n_p of the neighbour image (x_j + cell n_p), 3 per pair
};

// One wavefront per central atom.  Stencil lane t (< 27): t = 0 the atom's own box, t >= 1
// the other offsets of {-1,0,1}^3 in lexicographic order (neilist.F90:505-535); boxes
// outside the grid hold no atoms.  Candidates are numbered in stencil order and walked 64
// at a time; a ballot keeps the reference's order when compacting.
template <bool FILL>
__global__ __launch_bounds__(256) void k_nb_pairs(const NbGeom* __restrict__ G, const int* __restrict__ atom_off,
                                                  int nstruc, int ntot, const int* __restrict__ cidx,
                                                  const double* __restrict__ cpos, const int* __restrict__ corg,
                                                  const int* __restrict__ cbox, const int* __restrict__ bstart,
                                                  const int* __restrict__ batoms, const double* __restrict__ mag,
                                                  const int* __restrict__ cshift, const int* __restrict__ afl,
                                                  double rcut, int* __restrict__ counts, PairOut out,
                                                  int* __restrict__ err) {
    const int lane = threadIdx.x & 63;
    const int a = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (a >= ntot) return;
    const NbGeom& g = G[find_struct(atom_off, nstruc, a)];
    const int c1 = cidx[g.cand_off + (a - g.atom_off)];  // identity block: cluster index of atom a
    const int b1 = cbox[c1];
    int cnt = 0, bs = 0;
    if (b1 >= 0 && lane < 27) {
        const int bl = b1 - g.box_off;
        const int h0 = bl / (g.ng[1] * g.ng[2]), h1 = (bl / g.ng[2]) % g.ng[1], h2 = bl % g.ng[2];
        int d0 = 0, d1 = 0, d2 = 0;
        if (lane > 0) {
            int e = lane - 1;
            if (e >= 13) ++e;
            d0 = e / 9 - 1; d1 = (e / 3) % 3 - 1; d2 = e % 3 - 1;
        }
        const int k0 = h0 + d0, k1 = h1 + d1, k2 = h2 + d2;
        if (k0 >= 0 && k0 < g.ng[0] && k1 >= 0 && k1 < g.ng[1] && k2 >= 0 && k2 < g.ng[2]) {
            const int B = g.box_off + (k0 * g.ng[1] + k1) * g.ng[2] + k2;
            bs = bstart[B];
            cnt = bstart[B + 1] - bs;
        }
    }
    int incl = cnt;
    for (int o = 1; o < 32; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    const int pre = incl - cnt;
    const int total = __shfl(incl, 26, 64);
    const double r1x = cpos[3 * (size_t)c1], r1y = cpos[3 * (size_t)c1 + 1], r1z = cpos[3 * (size_t)c1 + 2];
    const double rcsq = rcut * rcut;
    int base_out = FILL ? out.offsets[a] : 0;
    int found = 0;
    for (int q0 = 0; q0 < total; q0 += 64) {
        const int q = q0 + lane;
        // largest stencil slot t with pre[t] <= q (empty slots share their successor's start)
        int t = 0;
        for (int st = 16; st > 0; st >>= 1) {
            const int cand = t + st;
            const int pv = __shfl(pre, cand < 27 ? cand : 26, 64);
            if (cand < 27 && pv <= q) t = cand;
        }
        const int pt = __shfl(pre, t, 64), bt = __shfl(bs, t, 64);
        bool pass = false;
        double v0 = 0, v1 = 0, v2 = 0;
        int c2 = -1;
        if (q < total) {
            c2 = batoms[bt + (q - pt)];
            if (c2 != c1) {
                v0 = cpos[3 * (size_t)c2] - r1x;
                v1 = cpos[3 * (size_t)c2 + 1] - r1y;
                v2 = cpos[3 * (size_t)c2 + 2] - r1z;
                // fastdist: neilist.F90:825-848
                if (fabs(v0) < rcut && fabs(v1) < rcut && fabs(v2) < rcut) pass = v0 * v0 + v1 * v1 + v2 * v2 < rcsq;
            }
        }
        const unsigned long long m = __ballot(pass);
        if (pass) {
            if (fabs(v0) + fabs(v1) + fabs(v2) < kEps) atomicOr(err, kNbErrIdentical);
            if (FILL) {
                const int p = base_out + found + __popcll(m & ((1ull << lane) - 1ull));
                const double d = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
                const int j = corg[c2];
                out.nlist[p] = j;
                out.dl[p] = d;
                // the boundary's convention (x_i - x_j)/r is the negated reference eist
                out.el[3 * (size_t)p] = -(v0 / d);
                out.el[3 * (size_t)p + 1] = -(v1 / d);
                out.el[3 * (size_t)p + 2] = -(v2 / d);
                if (out.shifts)
                    for (int n = 0; n < 3; ++n)
                        out.shifts[3 * (size_t)p + n] = cshift[3 * (size_t)c2 + n] - afl[3 * (size_t)j + n] +
                                                        afl[3 * (size_t)a + n];
                const double mx = mag[3 * (size_t)j], my = mag[3 * (size_t)j + 1], mz = mag[3 * (size_t)j + 2];
                const double mm = sqrt(mx * mx + my * my + mz * mz);
                out.dm[p] = mm;
                if (mm == 0.0) {
                    out.em[3 * (size_t)p] = 0.0; out.em[3 * (size_t)p + 1] = 0.0; out.em[3 * (size_t)p + 2] = 1.0;
                } else {
                    out.em[3 * (size_t)p] = mx / mm; out.em[3 * (size_t)p + 1] = my / mm; out.em[3 * (size_t)p + 2] = mz / mm;
                }
            }
        }
        found += __popcll(m);
    }
    if (lane == 0) {
        if (!FILL) {
            counts[a] = found;
        } else {
            // central moment -- neilist.F90:780-786
            const double mx = mag[3 * (size_t)a], my = mag[3 * (size_t)a + 1], mz = mag[3 * (size_t)a + 2];
            const double mi = sqrt(mx * mx + my * my + mz * mz);
            out.dmi[a] = mi;
            if (mi == 0.0) {
                out.emi[3 * (size_t)a] = 0.0; out.emi[3 * (size_t)a + 1] = 0.0; out.emi[3 * (size_t)a + 2] = 1.0;
            } else {
                out.emi[3 * (size_t)a] = mx / mi; out.emi[3 * (size_t)a + 1] = my / mi; out.emi[3 * (size_t)a + 2] = mz / mi;
            }
        }
    }
}

__global__ void k_nb_gather_clusters(int nstruc, const NbGeom* __restrict__ G, const int* __restrict__ cidx,
                                     long long ncand, int* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nstruc) return;
    out[s] = cidx[s < nstruc ? G[s].cand_off : ncand];
}

// ---------------------------------------------------------------- host side

template <typename T>
bool grow(T*& p, size_t& cap, size_t n) {
    if (n <= cap && p) return true;
    if (p) (void)hipFree((void*)p);
    p = nullptr;
    cap = 0;
    const size_t want = n + n / 4 + 64;
    if (hipMalloc((void**)&p, want * sizeof(T)) != hipSuccess) return false;
    cap = want;
    return true;
}

struct NbWork {
    NbGeom* geom = nullptr; size_t c_geom = 0;
    long long* cand_off = nullptr; size_t c_cand_off = 0;
    int* atom_off = nullptr; size_t c_atom_off = 0;
    double* frac = nullptr; size_t c_frac = 0;
    double* bbox = nullptr; size_t c_bbox = 0;
    int* flag = nullptr; size_t c_flag = 0;
    int* cidx = nullptr; size_t c_cidx = 0;
    double* cpos = nullptr; size_t c_cpos = 0;
    int* corg = nullptr; size_t c_corg = 0;
    int* cbox = nullptr; size_t c_cbox = 0;
    int* cshift = nullptr; size_t c_cshift = 0;
    int* afl = nullptr; size_t c_afl = 0;
    int* bcount = nullptr; size_t c_bcount = 0;
    int* bstart = nullptr; size_t c_bstart = 0;
    int* bcur = nullptr; size_t c_bcur = 0;
    int* batoms = nullptr; size_t c_batoms = 0;
    int* counts = nullptr; size_t c_counts = 0;
    int* sums = nullptr; size_t c_sums = 0;
    int* err = nullptr; size_t c_err = 0;
    int* clus = nullptr; size_t c_clus = 0;
    // staging for the host-pointer entry point
    double *pos = nullptr, *mag = nullptr; size_t c_pos = 0, c_mag = 0;
    int *off = nullptr, *nl = nullptr; size_t c_off = 0, c_nl = 0;
    double *dl = nullptr, *el = nullptr, *dm = nullptr, *em = nullptr, *dmi = nullptr, *emi = nullptr;
    size_t c_dl = 0, c_el = 0, c_dm = 0, c_em = 0, c_dmi = 0, c_emi = 0;
    std::vector<NbGeom> hgeom;
    std::vector<long long> hcand;
    std::vector<int> hatom;
};

NbWork g_nb;

int fail_hip(hipError_t e, const char* what) {
    report_error(MAGACE_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e) + " (" + what + ")");
    return MAGACE_ERR_DEVICE;
}

#define NBCHK(expr)                                     \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return fail_hip(_e, #expr); \
    } while (0)

#define NBGROW(ptr, n)                                                            \
    do {                                                                          \
        if (!grow(g_nb.ptr, g_nb.c_##ptr, (size_t)(n)))                           \
            return fail_hip(hipErrorOutOfMemory, "neighbour workspace " #ptr);     \
    } while (0)

unsigned nblocks(long long n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// out[0..n] = exclusive scan of in[0..n-1], out[n] = total
int scan(const int* in, long long n, int* out, hipStream_t s) {
    const long long ntiles = (n + kScanTile - 1) / kScanTile;
    if (ntiles > (1ll << 30)) {
        report_error(MAGACE_ERR_INPUT, "neighbour build: array too large to scan");
        return MAGACE_ERR_INPUT;
    }
    NBGROW(sums, ntiles + 1);
    if (ntiles > 0) k_scan_tiles<<<(unsigned)ntiles, kScanThreads, 0, s>>>(in, n, g_nb.sums);
    k_scan_sums<<<1, kScanThreads, 0, s>>>(g_nb.sums, (int)ntiles, out, n);
    if (ntiles > 0) k_scan_apply<<<(unsigned)ntiles, kScanThreads, 0, s>>>(in, n, g_nb.sums, out);
    NBCHK(hipGetLastError());
    return 0;
}

double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// normalised cross product (neilist.F90:799-822)
bool north(const double* v1, const double* v2, double* vn) {
    vn[0] = v1[1] * v2[2] - v1[2] * v2[1];
    vn[1] = v1[2] * v2[0] - v1[0] * v2[2];
    vn[2] = v1[0] * v2[1] - v1[1] * v2[0];
    const double x = std::sqrt(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
    if (x < 1.0e-9) return false;
    for (int n = 0; n < 3; ++n) vn[n] = vn[n] / x;
    return true;
}

// cofactor inverse, in the term order of neilist.F90:852-890 (a[i + 3k] = a(i+1,k+1))
bool inverse3(const double* a, double* ai) {
    auto A = [a](int i, int k) { return a[(i - 1) + 3 * (k - 1)]; };
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
    const double det = A(1, 1) * (A(3, 3) * A(2, 2) - A(3, 2) * A(2, 3)) -
                       A(2, 1) * (A(3, 3) * A(1, 2) - A(3, 2) * A(1, 3)) +
                       A(3, 1) * (A(2, 3) * A(1, 2) - A(2, 2) * A(1, 3));
    if (!(std::fabs(det) > 1.0e-9)) return false;
    for (int k = 0; k < 9; ++k) ai[k] = c[k] / det;
    return true;
}

// periodic geometry of one structure -- neilist.F90:215-311
bool periodic_geometry(NbGeom& g, double rcut) {
    const double *c1 = g.cell, *c2 = g.cell + 3, *c3 = g.cell + 6;
    const double* trip[3][3] = {{c1, c2, c3}, {c3, c1, c2}, {c2, c3, c1}};
    const int dir[3] = {2, 1, 0};
    for (int t = 0; t < 3; ++t) {
        double v[3];
        if (!north(trip[t][0], trip[t][1], v)) return false;
        const double x = dot3(trip[t][2], v);
        if (x < 1.0e-3) return false;  // "Extreme shape of unit cell."
        const double h = 1.0 + rcut / x;
        const double l = -h + 1.0;
        const int hc = (int)std::ceil(h), lc = (int)std::floor(l);
        g.dxh[dir[t]] = h;
        g.dxl[dir[t]] = l;
        g.lc[dir[t]] = lc;
        g.nsh[dir[t]] = hc - lc + 1;
    }
    for (int n = 0; n < 3; ++n) {
        double hi = 0.0, lo = 0.0;
        for (int i1 = 0; i1 <= 1; ++i1)
            for (int i2 = 0; i2 <= 1; ++i2)
                for (int i3 = 0; i3 <= 1; ++i3) {
                    const double x = rcut + (double)i1 * c1[n] + (double)i2 * c2[n] + (double)i3 * c3[n];
                    if (x > hi) hi = x;
                    const double y = -rcut + (double)i1 * c1[n] + (double)i2 * c2[n] + (double)i3 * c3[n];
                    if (y < lo) lo = y;
                }
        g.hcc[n] = hi;
        g.lcc[n] = lo;
    }
    return inverse3(g.cell, g.ci);
}

struct Request {
    int nstruc;
    const int* natoms;
    const int* periodic;
    const double* cells;
    double rcut;
    const double* pos;  // device
    const double* mag;  // device
    long long capacity;
    int* offsets;       // device [ntot+1]
    PairOut out;        // device (nlist == nullptr: count only)
    int* npairs;        // host
    int* natombox;      // host [nstruc] or nullptr
    hipStream_t s;
};

int build(const Request& R) {
    if (!(R.rcut >= 1.0e-6 && R.rcut <= 100.0)) {  // neilist.F90:206-213
        report_error(MAGACE_ERR_INPUT, R.rcut < 1.0e-6 ? "value for rcut to small." : "value for rcut to large.");
        return MAGACE_ERR_INPUT;
    }
    if (R.nstruc < 0 || (R.nstruc > 0 && (!R.natoms || !R.periodic || !R.cells)) || !R.offsets || !R.npairs) {
        report_error(MAGACE_ERR_INPUT, "neighbour build: missing arguments");
        return MAGACE_ERR_INPUT;
    }
    const int S = R.nstruc;
    auto& hg = g_nb.hgeom;
    auto& hc = g_nb.hcand;
    auto& ha = g_nb.hatom;
    hg.assign(S, NbGeom{});
    ha.assign(S + 1, 0);
    hc.assign(S + 1, 0);
    long long ntot = 0;
    bool any_open = false;
    for (int s = 0; s < S; ++s) {
        if (R.natoms[s] < 0) {
            report_error(MAGACE_ERR_INPUT, "neighbour build: negative atom count");
            return MAGACE_ERR_INPUT;
        }
        NbGeom& g = hg[s];
        std::memcpy(g.cell, R.cells + 9 * (size_t)s, sizeof g.cell);
        g.periodic = R.periodic[s] != 0;
        g.natom = R.natoms[s];
        g.atom_off = (int)ntot;
        ha[s] = (int)ntot;
        ntot += R.natoms[s];
        any_open = any_open || (!g.periodic && g.natom > 0);
        if (ntot >= (1ll << 31) - 1) {
            report_error(MAGACE_ERR_INPUT, "neighbour build: too many atoms");
            return MAGACE_ERR_INPUT;
        }
    }
    ha[S] = (int)ntot;
    if ((R.pos == nullptr || R.mag == nullptr) && ntot > 0) {
        report_error(MAGACE_ERR_INPUT, "neighbour build: missing positions or moments");
        return MAGACE_ERR_INPUT;
    }
    hipStream_t s = R.s;
    NBGROW(geom, S + 1);
    NBGROW(atom_off, S + 1);
    NBGROW(cand_off, S + 1);
    NBGROW(err, 4);
    std::vector<double> bbox;
    if (any_open) {
        // open clusters: the grid spans their bounding box (neilist.F90:313-352)
        NBGROW(bbox, 6 * (size_t)S);
        NBCHK(hipMemcpyAsync(g_nb.geom, hg.data(), sizeof(NbGeom) * S, hipMemcpyHostToDevice, s));
        k_nb_bbox<<<S, 256, 0, s>>>(g_nb.geom, R.pos, g_nb.bbox);
        NBCHK(hipGetLastError());
        bbox.resize(6 * (size_t)S);
        NBCHK(hipMemcpyAsync(bbox.data(), g_nb.bbox, sizeof(double) * 6 * S, hipMemcpyDeviceToHost, s));
        NBCHK(hipStreamSynchronize(s));
    }
    long long ncand = 0, nbox = 0;
    for (int si = 0; si < S; ++si) {
        NbGeom& g = hg[si];
        int nshifts = 1;
        if (g.natom == 0) {
            for (int n = 0; n < 3; ++n) { g.ng[n] = 0; g.nglow[n] = 0; g.lc[n] = 0; g.nsh[n] = 1; }
            nshifts = 0;
        } else if (g.periodic) {
            if (!periodic_geometry(g, R.rcut)) {
                report_error(MAGACE_ERR_INPUT, "Extreme shape of unit cell. (structure " + std::to_string(si) + ")");
                return MAGACE_ERR_INPUT;
            }
            nshifts = g.nsh[0] * g.nsh[1] * g.nsh[2];
            g.zero_shift = ((0 - g.lc[0]) * g.nsh[1] + (0 - g.lc[1])) * g.nsh[2] + (0 - g.lc[2]);
        } else {
            for (int n = 0; n < 3; ++n) {
                g.hcc[n] = bbox[6 * si + n];
                g.lcc[n] = bbox[6 * si + 3 + n];
                g.lc[n] = 0;
                g.nsh[n] = 1;
            }
        }
        if (g.natom > 0) {
            for (int n = 0; n < 3; ++n) {
                const double up = std::floor(g.hcc[n] / R.rcut + kEps), lo = std::floor(g.lcc[n] / R.rcut - kEps);
                if (!(up - lo + 1.0 <= 1.0e6)) {
                    report_error(MAGACE_ERR_INPUT, "neighbour build: grid too large (structure " + std::to_string(si) + ")");
                    return MAGACE_ERR_INPUT;
                }
                g.nglow[n] = (int)lo;
                g.ng[n] = (int)(up - lo) + 1;
            }
        }
        g.cand_off = ncand;
        hc[si] = ncand;
        g.box_off = (int)nbox;
        ncand += (long long)nshifts * g.natom;
        nbox += (long long)g.ng[0] * g.ng[1] * g.ng[2];
        if (ncand >= (1ll << 31) - 1 || nbox >= (1ll << 31) - 1) {
            report_error(MAGACE_ERR_INPUT, "neighbour build: image or grid count too large");
            return MAGACE_ERR_INPUT;
        }
    }
    hc[S] = ncand;
    NBGROW(frac, 3 * (size_t)ntot + 1);
    NBGROW(flag, ncand + 1);
    NBGROW(cidx, ncand + 1);
    NBGROW(cpos, 3 * (size_t)ncand + 1);
    NBGROW(corg, ncand + 1);
    NBGROW(cbox, ncand + 1);
    NBGROW(cshift, 3 * (size_t)ncand + 1);
    NBGROW(afl, 3 * (size_t)ntot + 1);
    NBGROW(bcount, nbox + 1);
    NBGROW(bstart, nbox + 1);
    NBGROW(bcur, nbox + 1);
    NBGROW(batoms, ncand + 1);
    NBGROW(counts, ntot + 1);
    NBGROW(clus, S + 1);
    NBCHK(hipMemcpyAsync(g_nb.geom, hg.data(), sizeof(NbGeom) * S, hipMemcpyHostToDevice, s));
    NBCHK(hipMemcpyAsync(g_nb.atom_off, ha.data(), sizeof(int) * (S + 1), hipMemcpyHostToDevice, s));
    NBCHK(hipMemcpyAsync(g_nb.cand_off, hc.data(), sizeof(long long) * (S + 1), hipMemcpyHostToDevice, s));
    k_fill_int<<<1, 64, 0, s>>>(g_nb.err, 1, 0);
    if (nbox > 0) {
        k_fill_int<<<nblocks(nbox, 256), 256, 0, s>>>(g_nb.bcount, nbox, 0);
        k_fill_int<<<nblocks(nbox, 256), 256, 0, s>>>(g_nb.bcur, nbox, 0);
    }
    if (ntot > 0)
        k_nb_frac<<<nblocks(ntot, 256), 256, 0, s>>>(g_nb.geom, g_nb.atom_off, S, (int)ntot, R.pos, g_nb.frac,
                                                      g_nb.afl);
    if (ncand > 0)
        k_nb_flag<<<nblocks(ncand, 256), 256, 0, s>>>(g_nb.geom, g_nb.cand_off, S, ncand, g_nb.frac, g_nb.flag);
    NBCHK(hipGetLastError());
    if (int rc = scan(g_nb.flag, ncand, g_nb.cidx, s)) return rc;
    const double invrc = 1.0 / R.rcut;
    if (ncand > 0)
        k_nb_cluster<<<nblocks(ncand, 256), 256, 0, s>>>(g_nb.geom, g_nb.cand_off, S, ncand, g_nb.frac, g_nb.flag,
                                                         g_nb.cidx, invrc, g_nb.cpos, g_nb.corg, g_nb.cbox,
                                                         g_nb.cshift, g_nb.bcount, g_nb.err);
    NBCHK(hipGetLastError());
    if (int rc = scan(g_nb.bcount, nbox, g_nb.bstart, s)) return rc;
    if (ncand > 0)
        k_nb_bfill<<<nblocks(ncand, 256), 256, 0, s>>>(g_nb.cidx + ncand, g_nb.cbox, g_nb.bstart, g_nb.bcur,
                                                       g_nb.batoms);
    if (nbox > 0) k_nb_bsort<<<nblocks(nbox, 256), 256, 0, s>>>((int)nbox, g_nb.bstart, g_nb.batoms);
    PairOut none{};
    if (ntot > 0)
        k_nb_pairs<false><<<nblocks(ntot, 4), 256, 0, s>>>(g_nb.geom, g_nb.atom_off, S, (int)ntot, g_nb.cidx,
                                                            g_nb.cpos, g_nb.corg, g_nb.cbox, g_nb.bstart,
                                                            g_nb.batoms, R.mag, g_nb.cshift, g_nb.afl, R.rcut,
                                                            g_nb.counts, none, g_nb.err);
    NBCHK(hipGetLastError());
    if (int rc = scan(g_nb.counts, ntot, R.offsets, s)) return rc;
    k_nb_gather_clusters<<<nblocks(S + 1, 64), 64, 0, s>>>(S, g_nb.geom, g_nb.cidx, ncand, g_nb.clus);
    NBCHK(hipGetLastError());
    int tail[2] = {0, 0};
    std::vector<int> clus(S + 1);
    NBCHK(hipMemcpyAsync(&tail[0], R.offsets + ntot, sizeof(int), hipMemcpyDeviceToHost, s));
    NBCHK(hipMemcpyAsync(&tail[1], g_nb.err, sizeof(int), hipMemcpyDeviceToHost, s));
    NBCHK(hipMemcpyAsync(clus.data(), g_nb.clus, sizeof(int) * (S + 1), hipMemcpyDeviceToHost, s));
    NBCHK(hipStreamSynchronize(s));
    if (tail[1] & kNbErrBox) {
        report_error(MAGACE_ERR_INPUT, "neighbour build: atom outside the grid (non-finite position?)");
        return MAGACE_ERR_INPUT;
    }
    if (tail[1] & kNbErrIdentical) {
        report_error(MAGACE_ERR_INPUT, "Atoms with identical positions.");
        return MAGACE_ERR_INPUT;
    }
    *R.npairs = tail[0];
    if (R.natombox)
        for (int si = 0; si < S; ++si) R.natombox[si] = clus[si + 1] - clus[si];
    if (!R.out.nlist) return 0;
    if ((long long)tail[0] > R.capacity) {
        report_error(MAGACE_ERR_CAPACITY, "neighbour build: " + std::to_string(tail[0]) +
                                              " pairs exceed the capacity " + std::to_string(R.capacity));
        return MAGACE_ERR_CAPACITY;
    }
    if (ntot > 0) {
        PairOut o = R.out;
        o.offsets = R.offsets;
        k_nb_pairs<true><<<nblocks(ntot, 4), 256, 0, s>>>(g_nb.geom, g_nb.atom_off, S, (int)ntot, g_nb.cidx,
                                                           g_nb.cpos, g_nb.corg, g_nb.cbox, g_nb.bstart,
                                                           g_nb.batoms, R.mag, g_nb.cshift, g_nb.afl, R.rcut,
                                                           g_nb.counts, o, g_nb.err);
        NBCHK(hipGetLastError());
    }
    return 0;
}

}  // namespace
}  // namespace magace

using namespace magace;

extern "C" {

int magace_neighbors_device(int nstruc, const int* natoms, const int* periodic, const double* cells,
                            double rcut, const double* pos, const double* mag, long long capacity,
                            int* offsets, int* nlist, double* dlist, double* elist, double* dlist_mom,
                            double* elist_mom, double* dlist_momi, double* elist_momi, int* shifts,
                            int* npairs, int* natombox, void* stream) {
    reset_error();
    Request R{nstruc, natoms, periodic, cells, rcut, pos, mag, capacity, offsets,
              PairOut{offsets, nlist, dlist, elist, dlist_mom, elist_mom, dlist_momi, elist_momi, shifts},
              npairs, natombox, (hipStream_t)stream};
    if (nlist && (!dlist || !elist || !dlist_mom || !elist_mom || !dlist_momi || !elist_momi)) {
        report_error(MAGACE_ERR_INPUT, "neighbour build: missing output arrays");
        return MAGACE_ERR_INPUT;
    }
    return build(R);
}

int magace_packstructures(int nstruc, const int* natoms, const int* periodic, const double* cells, double rcut,
                          const double* atompos, const double* magmom, long long capacity, int* strucaddress,
                          int* n2startpack, int* n2stoppack, int* n2listpack, double* d2listpack,
                          double* e2listpack, double* d2listpack_mom, double* e2listpack_mom,
                          double* d2listpack_momi, double* e2listpack_momi, int* natombox, int* npairs) {
    reset_error();
    if (nstruc < 0 || (nstruc > 0 && (!natoms || !periodic || !cells)) || !npairs) {
        report_error(MAGACE_ERR_INPUT, "packstructures: missing arguments");
        return MAGACE_ERR_INPUT;
    }
    long long ntot = 0;
    for (int s = 0; s < nstruc; ++s) ntot += natoms[s] > 0 ? natoms[s] : 0;
    if (ntot >= (1ll << 31) - 1) {
        report_error(MAGACE_ERR_INPUT, "packstructures: too many atoms");
        return MAGACE_ERR_INPUT;
    }
    hipStream_t s = nullptr;
    NBGROW(pos, 3 * ntot + 1);
    NBGROW(mag, 3 * ntot + 1);
    NBGROW(off, ntot + 1);
    if (ntot > 0) {
        NBCHK(hipMemcpy(g_nb.pos, atompos, sizeof(double) * 3 * ntot, hipMemcpyHostToDevice));
        NBCHK(hipMemcpy(g_nb.mag, magmom, sizeof(double) * 3 * ntot, hipMemcpyHostToDevice));
    }
    const long long cap = capacity > 0 ? capacity : 0;
    NBGROW(nl, cap + 1);
    NBGROW(dl, cap + 1);
    NBGROW(el, 3 * cap + 1);
    NBGROW(dm, cap + 1);
    NBGROW(em, 3 * cap + 1);
    NBGROW(dmi, ntot + 1);
    NBGROW(emi, 3 * ntot + 1);
    Request R{nstruc, natoms, periodic, cells, rcut, g_nb.pos, g_nb.mag, cap, g_nb.off,
              PairOut{g_nb.off, g_nb.nl, g_nb.dl, g_nb.el, g_nb.dm, g_nb.em, g_nb.dmi, g_nb.emi, nullptr},
              npairs, natombox, s};
    if (int rc = build(R)) return rc;
    NBCHK(hipStreamSynchronize(s));
    const long long P = *npairs;
    std::vector<int> off(ntot + 1), nl(P);
    NBCHK(hipMemcpy(off.data(), g_nb.off, sizeof(int) * (ntot + 1), hipMemcpyDeviceToHost));
    NBCHK(hipMemcpy(nl.data(), g_nb.nl, sizeof(int) * P, hipMemcpyDeviceToHost));
    if (P > 0) {
        NBCHK(hipMemcpy(d2listpack, g_nb.dl, sizeof(double) * P, hipMemcpyDeviceToHost));
        NBCHK(hipMemcpy(e2listpack, g_nb.el, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
        NBCHK(hipMemcpy(d2listpack_mom, g_nb.dm, sizeof(double) * P, hipMemcpyDeviceToHost));
        NBCHK(hipMemcpy(e2listpack_mom, g_nb.em, sizeof(double) * 3 * P, hipMemcpyDeviceToHost));
    }
    if (ntot > 0) {
        NBCHK(hipMemcpy(d2listpack_momi, g_nb.dmi, sizeof(double) * ntot, hipMemcpyDeviceToHost));
        NBCHK(hipMemcpy(e2listpack_momi, g_nb.emi, sizeof(double) * 3 * ntot, hipMemcpyDeviceToHost));
    }
    // the reference's conventions (neiutil.F90:145-196): per-structure 1-based local ranges
    // and neighbour indices, eist = (x_j - x_i)/d
    for (long long p = 0; p < 3 * P; ++p) e2listpack[p] = -e2listpack[p];
    long long a0 = 0;
    for (int si = 0; si < nstruc; ++si) {
        const int n = natoms[si];
        const int pbase = off[a0];
        strucaddress[2 * si] = (int)a0 + 1;
        strucaddress[2 * si + 1] = pbase + 1;
        for (int i = 0; i < n; ++i) {
            n2startpack[a0 + i] = off[a0 + i] - pbase + 1;
            n2stoppack[a0 + i] = off[a0 + i + 1] - pbase;
        }
        for (int p = off[a0]; p < off[a0 + n]; ++p) n2listpack[p] = nl[p] - (int)a0 + 1;
        a0 += n;
    }
    return 0;
}

}  // extern "C"
