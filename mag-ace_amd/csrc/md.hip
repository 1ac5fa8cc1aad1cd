// md.hip -- device-resident spin-lattice dynamics around the evaluator.
//
// Replaces what LAMMPS does around pair_spin_ace for a single GPU: the full-list force and
// field assembly of PairSpinACE::compute (lammps/pair_spin_ace.cpp:571-641: f_i -= F_p,
// f_j += F_p, fm_i -= tlist0_i, fm_j -= tlist_p) and the per-spin re-evaluation sweep of
// compute_single_pair (:677-986, a full acelib_ call per spin).  Here one evaluation per step
// gives every atom's force and every moment's field at once.
//
// Step n -> n+1 (one evaluation per step, states synchronous):
//   v += dt/2 f_n/m ; x += dt v ;
//   s <- R(Omega(G_mid), dt) s with G_mid = 3/2 G_n - 1/2 G_{n-1} (the field extrapolated to
//        the middle of the step, which makes the spin update second order) ;
//   pair geometry from x, s ; evaluate (k_density, k_adjoint, k_forces) ; assemble f_{n+1},
//   G_{n+1} = dE/dm ; v += dt/2 f_{n+1}/m
// with Omega_i = -(g/hbar) G_i (rad/ps; g = 2, hbar in eV ps, G in eV/muB), so
// ds_i/dt = Omega_i x s_i, which conserves |s_i| and, at fixed positions, the energy.  The
// reference's compute_single_pair instead scales the raw gradient by mub/hbar and reads
// tlist0/tlist at fixed offsets rather than at the atom's own data; the units here are
// stated once and used throughout.
//
// Neighbour lists are rebuilt every `rebuild_every` steps with cutoff r_list = r_c + skin
// (the build records each pair's lattice shift, so the pair vectors are recomputed from the
// moved positions in between; pairs beyond r_c contribute exact zeros).  The steps between
// two rebuilds are captured once into a HIP graph and replayed.  The graph bakes in the pair
// buffers and the evaluator's state at capture time (its per-atom buffers, tables, precision
// and coefficients), so it is captured again when the pair buffers grow or when that state
// changed since (context_generation).  Force and field assembly is a gather over each atom's
// own pairs and their reverse partners, so results are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/magace.h"
#include "magace.hpp"

namespace magace {
namespace {

constexpr double kHbar = 6.582119569e-4;  // eV ps
constexpr double kMvv2e = 1.0364269e-4;   // amu A^2 ps^-2 -> eV (LAMMPS metal units)
constexpr double kG = 2.0;                // gyromagnetic factor of Omega = -(g/hbar) dE/dm

struct Cell {
    double a[9];  // a[i + 3k] = component i of lattice vector k
};

__global__ void k_md_owner(int n, const int* __restrict__ off, int* __restrict__ owner) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int p = off[i]; p < off[i + 1]; ++p) owner[p] = i;
}

// partner (j -> i, -n) of every pair (i -> j, n); -1 if the list lacks it (only possible for
// pairs at the list cutoff, which lies beyond the potential's cutoff)
__global__ void k_md_reverse(int cap, const int* __restrict__ off, int n, const int* __restrict__ nlist,
                             const int* __restrict__ shifts, const int* __restrict__ owner, int* __restrict__ rev) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= cap || p >= off[n]) return;
    const int i = owner[p], j = nlist[p];
    const int s0 = -shifts[3 * (size_t)p], s1 = -shifts[3 * (size_t)p + 1], s2 = -shifts[3 * (size_t)p + 2];
    int r = -1;
    for (int q = off[j]; q < off[j + 1]; ++q)
        if (nlist[q] == i && shifts[3 * (size_t)q] == s0 && shifts[3 * (size_t)q + 1] == s1 &&
            shifts[3 * (size_t)q + 2] == s2) {
            r = q;
            break;
        }
    rev[p] = r;
}

__global__ void k_md_moments(int n, const double* __restrict__ s, const double* __restrict__ mu,
                             double* __restrict__ m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; ++k) m[3 * (size_t)i + k] = mu[i] * s[3 * (size_t)i + k];
}

// acelib_ inputs from the current positions and spins: u = x_i - (x_j + a.n_p)
__global__ void k_md_pairs(int cap, int n, const int* __restrict__ off, const int* __restrict__ owner,
                           const int* __restrict__ nlist, const int* __restrict__ shifts, Cell c,
                           const double* __restrict__ x, const double* __restrict__ s,
                           const double* __restrict__ mu, double* __restrict__ dl, double* __restrict__ el,
                           double* __restrict__ dm, double* __restrict__ em, double* __restrict__ dmi,
                           double* __restrict__ emi) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) {
        dmi[p] = mu[p];
        for (int k = 0; k < 3; ++k) emi[3 * (size_t)p + k] = s[3 * (size_t)p + k];
    }
    if (p >= cap || p >= off[n]) return;
    const int i = owner[p], j = nlist[p];
    const double n0 = shifts[3 * (size_t)p], n1 = shifts[3 * (size_t)p + 1], n2 = shifts[3 * (size_t)p + 2];
    double u[3];
    for (int k = 0; k < 3; ++k)
        u[k] = x[3 * (size_t)i + k] - (x[3 * (size_t)j + k] + (c.a[k] * n0 + c.a[k + 3] * n1 + c.a[k + 6] * n2));
    const double r = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    dl[p] = r;
    for (int k = 0; k < 3; ++k) el[3 * (size_t)p + k] = u[k] / r;
    dm[p] = mu[j];
    for (int k = 0; k < 3; ++k) em[3 * (size_t)p + k] = s[3 * (size_t)j + k];
}

// f_i = -sum_p F_p + sum_p F_rev(p);  dE/dm_i = tlist0_i + sum_p tlist_rev(p)
__global__ void k_md_assemble(int n, const int* __restrict__ off, const int* __restrict__ rev,
                              const double* __restrict__ fl, const double* __restrict__ tl,
                              const double* __restrict__ t0, double* __restrict__ f, double* __restrict__ G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double fx = 0, fy = 0, fz = 0, gx = t0[3 * (size_t)i], gy = t0[3 * (size_t)i + 1], gz = t0[3 * (size_t)i + 2];
    for (int p = off[i]; p < off[i + 1]; ++p) {
        fx -= fl[3 * (size_t)p];
        fy -= fl[3 * (size_t)p + 1];
        fz -= fl[3 * (size_t)p + 2];
        const int q = rev[p];
        if (q >= 0) {
            fx += fl[3 * (size_t)q];
            fy += fl[3 * (size_t)q + 1];
            fz += fl[3 * (size_t)q + 2];
            gx += tl[3 * (size_t)q];
            gy += tl[3 * (size_t)q + 1];
            gz += tl[3 * (size_t)q + 2];
        }
    }
    f[3 * (size_t)i] = fx; f[3 * (size_t)i + 1] = fy; f[3 * (size_t)i + 2] = fz;
    G[3 * (size_t)i] = gx; G[3 * (size_t)i + 1] = gy; G[3 * (size_t)i + 2] = gz;
}

__device__ void rotate(double* s, const double* G, double h) {
    // ds/dt = Omega x s, Omega = -(g/hbar) G: rotation about Omega by |Omega| h (Rodrigues)
    const double ox = -kG / kHbar * G[0], oy = -kG / kHbar * G[1], oz = -kG / kHbar * G[2];
    const double w = sqrt(ox * ox + oy * oy + oz * oz);
    if (!(w > 0.0)) return;
    const double kx = ox / w, ky = oy / w, kz = oz / w, th = w * h;
    const double c = cos(th), sn = sin(th);
    const double kd = kx * s[0] + ky * s[1] + kz * s[2];
    const double cx = ky * s[2] - kz * s[1], cy = kz * s[0] - kx * s[2], cz = kx * s[1] - ky * s[0];
    double r0 = s[0] * c + cx * sn + kx * kd * (1.0 - c);
    double r1 = s[1] * c + cy * sn + ky * kd * (1.0 - c);
    double r2 = s[2] * c + cz * sn + kz * kd * (1.0 - c);
    const double nrm = 1.0 / sqrt(r0 * r0 + r1 * r1 + r2 * r2);
    s[0] = r0 * nrm; s[1] = r1 * nrm; s[2] = r2 * nrm;
}

// kick + drift + spin rotation with the mid-step field; Gprev <- G
__global__ void k_md_first(int n, double dt, double inv_m, double* __restrict__ x, double* __restrict__ v,
                           double* __restrict__ s, const double* __restrict__ f, const double* __restrict__ G,
                           double* __restrict__ Gprev) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double sv[3] = {s[3 * (size_t)i], s[3 * (size_t)i + 1], s[3 * (size_t)i + 2]};
    double gm[3];
    for (int k = 0; k < 3; ++k) {
        const double g = G[3 * (size_t)i + k];
        gm[k] = 1.5 * g - 0.5 * Gprev[3 * (size_t)i + k];
        Gprev[3 * (size_t)i + k] = g;
        v[3 * (size_t)i + k] += 0.5 * dt * inv_m * f[3 * (size_t)i + k];
        x[3 * (size_t)i + k] += dt * v[3 * (size_t)i + k];
    }
    rotate(sv, gm, dt);
    for (int k = 0; k < 3; ++k) s[3 * (size_t)i + k] = sv[k];
}

__global__ void k_md_kick(int n, double dt, double inv_m, double* __restrict__ v, const double* __restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n) return;
    v[i] += 0.5 * dt * inv_m * f[i];
}

// Energy and virial of the last evaluation, FP64, in a fixed summation order (bitwise
// repeatable): E = sum_i E_i and W_ab = sum_p u_a F_b with u = x_i - x_j = r e_p, the tally
// of lammps/pair_spin_ace.cpp:635-636 (ev_tally_xyz with the pair's F_p and u_p).  Block b
// sums a fixed strided slice of atoms and pairs and reduces through LDS in a fixed tree;
// k_md_tally_final sums the block partials in order.
constexpr int kTallyBlocks = 512, kTallyThreads = 256, kTally = 10;  // E, W_xx W_xy ... W_zz

__global__ void __launch_bounds__(kTallyThreads) k_md_tally(int n, const int* __restrict__ off,
                                                            const double* __restrict__ eout,
                                                            const double* __restrict__ dl,
                                                            const double* __restrict__ el,
                                                            const double* __restrict__ fl,
                                                            double* __restrict__ part) {
    __shared__ double red[kTally][kTallyThreads];
    const int t = threadIdx.x, stride = kTallyBlocks * kTallyThreads;
    const int g = blockIdx.x * kTallyThreads + t;
    double v[kTally] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = g; i < n; i += stride) v[0] += eout[i];
    const int np = off[n];
    for (int p = g; p < np; p += stride) {
        const double r = dl[p];
        const double u[3] = {r * el[3 * (size_t)p], r * el[3 * (size_t)p + 1], r * el[3 * (size_t)p + 2]};
        const double F[3] = {fl[3 * (size_t)p], fl[3 * (size_t)p + 1], fl[3 * (size_t)p + 2]};
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) v[1 + 3 * a + b] += u[a] * F[b];
    }
    for (int k = 0; k < kTally; ++k) red[k][t] = v[k];
    __syncthreads();
    for (int h = kTallyThreads / 2; h > 0; h >>= 1) {
        if (t < h)
            for (int k = 0; k < kTally; ++k) red[k][t] += red[k][t + h];
        __syncthreads();
    }
    if (t < kTally) part[(size_t)blockIdx.x * kTally + t] = red[t][0];
}

__global__ void k_md_tally_final(const double* __restrict__ part, double* __restrict__ out) {
    const int k = threadIdx.x;
    if (k >= kTally) return;
    double s = 0.0;
    for (int b = 0; b < kTallyBlocks; ++b) s += part[(size_t)b * kTally + k];
    out[k] = s;
}

struct Md {
    int n = 0;
    Cell cell{};
    double rlist = 0, mass = 0, dt = 0;
    int every = 10;
    hipStream_t stream = nullptr;
    double *x = nullptr, *v = nullptr, *s = nullptr, *mu = nullptr, *m = nullptr, *f = nullptr, *G = nullptr;
    double* Gprev = nullptr;
    double *eout = nullptr, *t0 = nullptr, *dmi = nullptr, *emi = nullptr;
    int* off = nullptr;
    // pair buffers (capacity cap)
    long long cap = 0;
    int npairs = 0;
    int *nl = nullptr, *sh = nullptr, *owner = nullptr, *rev = nullptr;
    double *dl = nullptr, *el = nullptr, *dm = nullptr, *em = nullptr, *fl = nullptr, *tl = nullptr;
    double* tally = nullptr;  // [kTallyBlocks][kTally] block partials, then [kTally] totals
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned long long gen = 0;  // context_generation() the graph was captured at
    long long step = 0;
    int rebuilds = 0, captures = 0;
    ~Md();
};

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void drop_graph(Md& M) {
    if (M.exec) (void)hipGraphExecDestroy(M.exec);
    if (M.graph) (void)hipGraphDestroy(M.graph);
    M.exec = nullptr;
    M.graph = nullptr;
}

Md::~Md() {
    drop_graph(*this);
    dfree(x); dfree(v); dfree(s); dfree(mu); dfree(m); dfree(f); dfree(G); dfree(Gprev); dfree(eout); dfree(t0);
    dfree(dmi); dfree(emi); dfree(off); dfree(tally);
    dfree(nl); dfree(sh); dfree(owner); dfree(rev); dfree(dl); dfree(el); dfree(dm); dfree(em); dfree(fl); dfree(tl);
    if (stream) (void)hipStreamDestroy(stream);
}

int hip_fail(hipError_t e, const char* what) {
    report_error(MAGACE_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e) + " (" + what + ")");
    return MAGACE_ERR_DEVICE;
}

#define MDCHK(expr)                                     \
    do {                                                \
        hipError_t _e = (expr);                         \
        if (_e != hipSuccess) return hip_fail(_e, #expr); \
    } while (0)

unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

int alloc_pairs(Md& M, long long cap) {
    dfree(M.nl); dfree(M.sh); dfree(M.owner); dfree(M.rev); dfree(M.dl); dfree(M.el); dfree(M.dm); dfree(M.em);
    dfree(M.fl); dfree(M.tl);
    drop_graph(M);  // captured launches refer to the old buffers
    const size_t c = (size_t)cap;
    MDCHK(hipMalloc((void**)&M.nl, c * sizeof(int)));
    MDCHK(hipMalloc((void**)&M.sh, 3 * c * sizeof(int)));
    MDCHK(hipMalloc((void**)&M.owner, c * sizeof(int)));
    MDCHK(hipMalloc((void**)&M.rev, c * sizeof(int)));
    MDCHK(hipMalloc((void**)&M.dl, c * sizeof(double)));
    MDCHK(hipMalloc((void**)&M.el, 3 * c * sizeof(double)));
    MDCHK(hipMalloc((void**)&M.dm, c * sizeof(double)));
    MDCHK(hipMalloc((void**)&M.em, 3 * c * sizeof(double)));
    MDCHK(hipMalloc((void**)&M.fl, 3 * c * sizeof(double)));
    MDCHK(hipMalloc((void**)&M.tl, 3 * c * sizeof(double)));
    M.cap = cap;
    return 0;
}

int rebuild(Md& M) {
    k_md_moments<<<blocks(M.n, 256), 256, 0, M.stream>>>(M.n, M.s, M.mu, M.m);
    MDCHK(hipGetLastError());
    const int periodic = 1;
    for (int attempt = 0; attempt < 2; ++attempt) {
        int np = 0;
        const int rc = magace_neighbors_device(1, &M.n, &periodic, M.cell.a, M.rlist, M.x, M.m, M.cap, M.off, M.nl,
                                               M.dl, M.el, M.dm, M.em, M.dmi, M.emi, M.sh, &np, nullptr, M.stream);
        if (rc == 0) {
            M.npairs = np;
            break;
        }
        if (rc != MAGACE_ERR_CAPACITY || attempt) return rc;
        if (int e = alloc_pairs(M, (long long)np + np / 4 + 1024)) return e;
    }
    k_md_owner<<<blocks(M.n, 256), 256, 0, M.stream>>>(M.n, M.off, M.owner);
    k_md_reverse<<<blocks(M.cap, 256), 256, 0, M.stream>>>((int)M.cap, M.off, M.n, M.nl, M.sh, M.owner, M.rev);
    MDCHK(hipGetLastError());
    ++M.rebuilds;
    return 0;
}

// pair geometry + evaluation + assembly (the force/field evaluation of one step)
int evaluate(Md& M) {
    k_md_pairs<<<blocks(std::max<long long>(M.cap, M.n), 256), 256, 0, M.stream>>>(
        (int)M.cap, M.n, M.off, M.owner, M.nl, M.sh, M.cell, M.x, M.s, M.mu, M.dl, M.el, M.dm, M.em, M.dmi, M.emi);
    MDCHK(hipGetLastError());
    if (int rc = magace_compute_device(M.n, M.npairs, M.off, M.dl, M.el, M.dmi, M.emi, M.dm, M.em, M.eout, M.fl,
                                       M.tl, M.t0, M.stream))
        return rc;
    k_md_assemble<<<blocks(M.n, 256), 256, 0, M.stream>>>(M.n, M.off, M.rev, M.fl, M.tl, M.t0, M.f, M.G);
    MDCHK(hipGetLastError());
    return 0;
}

int one_step(Md& M) {
    const double inv_m = 1.0 / (M.mass * kMvv2e);
    k_md_first<<<blocks(M.n, 256), 256, 0, M.stream>>>(M.n, M.dt, inv_m, M.x, M.v, M.s, M.f, M.G, M.Gprev);
    MDCHK(hipGetLastError());
    if (int rc = evaluate(M)) return rc;
    k_md_kick<<<blocks(3 * (long long)M.n, 256), 256, 0, M.stream>>>(M.n, M.dt, inv_m, M.v, M.f);
    MDCHK(hipGetLastError());
    return 0;
}

int capture(Md& M) {
    drop_graph(M);
    // tables uploaded and per-atom buffers sized before the capture, so that no allocation
    // is recorded; an error here (no potential) is returned before capturing
    if (int rc = context_prepare(M.n)) return rc;
    const unsigned long long gen = context_generation();
    MDCHK(hipStreamBeginCapture(M.stream, hipStreamCaptureModeRelaxed));
    int rc = 0;
    for (int k = 0; k < M.every && !rc; ++k) rc = one_step(M);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(M.stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    M.graph = g;
    MDCHK(hipGraphInstantiate(&M.exec, M.graph, nullptr, nullptr, 0));
    M.gen = gen;
    ++M.captures;
    return 0;
}

}  // namespace
}  // namespace magace

using namespace magace;

extern "C" {

int magace_md_create(int natom, const double* cell, double rlist, double mass, double dt, int rebuild_every,
                     const double* x, const double* v, const double* sdir, const double* smag, void** handle) {
    reset_error();
    if (!handle || natom <= 0 || !cell || !x || !v || !sdir || !smag || !(mass > 0) || !(dt > 0) ||
        rebuild_every <= 0 || !(rlist > 0)) {
        report_error(MAGACE_ERR_INPUT, "magace_md_create: invalid arguments");
        return MAGACE_ERR_INPUT;
    }
    *handle = nullptr;
    auto M = std::make_unique<Md>();
    M->n = natom;
    for (int k = 0; k < 9; ++k) M->cell.a[k] = cell[k];
    M->rlist = rlist;
    M->mass = mass;
    M->dt = dt;
    M->every = rebuild_every;
    MDCHK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
    const size_t n = (size_t)natom;
    double** d3[] = {&M->x, &M->v, &M->s, &M->m, &M->f, &M->G, &M->Gprev, &M->t0, &M->emi};
    for (double** p : d3) MDCHK(hipMalloc((void**)p, 3 * n * sizeof(double)));
    double** d1[] = {&M->mu, &M->eout, &M->dmi};
    for (double** p : d1) MDCHK(hipMalloc((void**)p, n * sizeof(double)));
    MDCHK(hipMalloc((void**)&M->off, (n + 1) * sizeof(int)));
    MDCHK(hipMalloc((void**)&M->tally, (kTallyBlocks + 1) * kTally * sizeof(double)));
    std::vector<double> sd(3 * n);
    for (size_t i = 0; i < n; ++i) {
        const double a = sdir[3 * i], b = sdir[3 * i + 1], c = sdir[3 * i + 2];
        const double r = std::sqrt(a * a + b * b + c * c);
        if (!(r > 0)) {
            report_error(MAGACE_ERR_INPUT, "magace_md_create: zero spin direction");
            return MAGACE_ERR_INPUT;
        }
        sd[3 * i] = a / r; sd[3 * i + 1] = b / r; sd[3 * i + 2] = c / r;
    }
    MDCHK(hipMemcpy(M->x, x, 3 * n * sizeof(double), hipMemcpyHostToDevice));
    MDCHK(hipMemcpy(M->v, v, 3 * n * sizeof(double), hipMemcpyHostToDevice));
    MDCHK(hipMemcpy(M->s, sd.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
    MDCHK(hipMemcpy(M->mu, smag, n * sizeof(double), hipMemcpyHostToDevice));
    if (int rc = alloc_pairs(*M, 64 * (long long)natom + 1024)) return rc;
    if (int rc = rebuild(*M)) return rc;
    if (int rc = evaluate(*M)) return rc;
    MDCHK(hipMemcpyAsync(M->Gprev, M->G, 3 * n * sizeof(double), hipMemcpyDeviceToDevice, M->stream));
    if (int rc = magace_check_errors(M->stream)) return rc;
    *handle = M.release();
    return 0;
}

int magace_md_run(void* handle, int nsteps, int use_graph) {
    reset_error();
    if (!handle || nsteps < 0) {
        report_error(MAGACE_ERR_INPUT, "magace_md_run: invalid arguments");
        return MAGACE_ERR_INPUT;
    }
    Md& M = *(Md*)handle;
    int done = 0;
    while (done < nsteps) {
        if (M.step % M.every == 0) {
            if (int rc = rebuild(M)) return rc;  // may reallocate pair buffers (drops the graph)
        }
        const int left_in_block = M.every - (int)(M.step % M.every);
        const int k = std::min(left_in_block, nsteps - done);
        if (use_graph && k == M.every) {
            if (M.exec && M.gen != context_generation()) drop_graph(M);  // stale evaluator state
            if (!M.exec)
                if (int rc = capture(M)) return rc;
            MDCHK(hipGraphLaunch(M.exec, M.stream));
        } else {
            for (int i = 0; i < k; ++i)
                if (int rc = one_step(M)) return rc;
        }
        done += k;
        M.step += k;
    }
    return magace_check_errors(M.stream);
}

int magace_md_read(void* handle, double* x, double* v, double* sdir, double* f, double* dEdm, double* eatom,
                   double* epot, double* ekin, double* virial, long long* counters) {
    reset_error();
    if (!handle) {
        report_error(MAGACE_ERR_INPUT, "magace_md_read: null handle");
        return MAGACE_ERR_INPUT;
    }
    Md& M = *(Md*)handle;
    // energy and virial of the current state, reduced on the device in FP64
    k_md_tally<<<kTallyBlocks, kTallyThreads, 0, M.stream>>>(M.n, M.off, M.eout, M.dl, M.el, M.fl, M.tally);
    k_md_tally_final<<<1, 64, 0, M.stream>>>(M.tally, M.tally + (size_t)kTallyBlocks * kTally);
    MDCHK(hipGetLastError());
    double tot[kTally];
    MDCHK(hipMemcpyAsync(tot, M.tally + (size_t)kTallyBlocks * kTally, sizeof tot, hipMemcpyDeviceToHost, M.stream));
    MDCHK(hipStreamSynchronize(M.stream));
    const size_t n = (size_t)M.n;
    std::vector<double> hv(3 * n), he(n);
    MDCHK(hipMemcpy(hv.data(), M.v, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    MDCHK(hipMemcpy(he.data(), M.eout, n * sizeof(double), hipMemcpyDeviceToHost));
    if (x) MDCHK(hipMemcpy(x, M.x, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    if (v) std::copy(hv.begin(), hv.end(), v);
    if (sdir) MDCHK(hipMemcpy(sdir, M.s, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    if (f) MDCHK(hipMemcpy(f, M.f, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    if (dEdm) MDCHK(hipMemcpy(dEdm, M.G, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
    if (eatom) std::copy(he.begin(), he.end(), eatom);
    double ek = 0;
    for (size_t i = 0; i < 3 * n; ++i) ek += hv[i] * hv[i];
    if (epot) *epot = tot[0];
    if (virial)
        for (int k = 0; k < 9; ++k) virial[k] = tot[1 + k];
    if (ekin) *ekin = 0.5 * M.mass * kMvv2e * ek;
    if (counters) {
        counters[0] = M.step;
        counters[1] = M.rebuilds;
        counters[2] = M.captures;
        counters[3] = M.npairs;
    }
    return 0;
}

int magace_md_destroy(void* handle) {
    delete (Md*)handle;
    return 0;
}

}  // extern "C"
