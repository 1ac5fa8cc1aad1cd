// api.cpp -- the C-ABI of libmagace.so.
//
// acelibini_ / acelib_ keep the exact signatures and semantics of libace's entry points as
// LAMMPS pair_spin_ace calls them (lammps/pair_spin_ace.cpp:59-62, src/ace.F90:10-21,79-107),
// so the library drops in for libace.  The remaining magace_* functions are additions:
// error reporting (the reference calls Fortran `stop` instead), an explicit-path init, the
// precision switch, and a device-pointer entry point for callers whose arrays already live
// in HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <array>
#include <atomic>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/magace.h"
#include "device.hpp"
#include "host_pool.hpp"
#include "magace.hpp"

namespace magace {

static_assert(TermList::kLanes == kAdjointThreads, "the chunk order must be built for the adjoint's workgroup size");
static_assert(TermList::kGroup == kCombGroup, "combine groups must match the kernel's");

namespace {

std::string g_err;
int g_errcode = 0;

void set_error(int code, const std::string& msg) {
    g_errcode = code;
    g_err = msg;
    std::fprintf(stderr, "magace: %s\n", msg.c_str());
}
void clear_error() {
    g_errcode = 0;
    g_err.clear();
}

}  // namespace

// shared with the other translation units of the library (neighbour build)
void report_error(int code, const std::string& msg) { set_error(code, msg); }
void reset_error() { clear_error(); }

namespace {

const char* kElements[] = {
    "X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",
    "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As",
    "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn",
    "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho",
    "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po",
    "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Np", "Pu", "Am", "Cm", "Bk", "Cf", "Es", "Fm", "Md",
    "No", "Lr"};

template <typename T>
struct DevTables {
    T* lut_a = nullptr;
    T* lut_m = nullptr;
};

struct Context {
    Model model;
    int precision = MAGACE_FP64;
    int shape_class = 0;
    // density staging variant and neighbour split per precision (index 0 FP64, 1 mixed):
    // -1/0 = the built-in rule, otherwise a tuned choice (tuning table or magace_tune)
    int dvar[2] = {-1, -1};
    int dsplit[2] = {0, 0};
    DevParams<double> P64;
    DevParams<float> P32;
    DevTables<double> t64;
    DevTables<float> t32;
    uint4* crec64 = nullptr;
    uint4* crec32 = nullptr;
    uint4* crec16 = nullptr;  // FP32 weights, offsets of 16-byte entries (two-atom adjoint)
    uint4* crecu = nullptr;   // FP32, offsets unpacked (DevParams::crecu)
    uint4* crec16u = nullptr;
    double* cw2_64 = nullptr;  // density-2 record weights (ndensity 2)
    float* cw2_32 = nullptr;
    int2* chunks = nullptr;
    int2* comb = nullptr;
    int2* comb2 = nullptr;
    uint32_t* fold = nullptr;
    int2* ypair = nullptr;     // products of two A entries (body orders 5, 6), TermList::ypair
    int2* yft = nullptr;       // their gradient's way back to Z (TermList::yft, yfe)
    int2* yfe = nullptr;
    Buffers buf;
    std::vector<int> h_off;
    hipStream_t stream = nullptr;
    // acelib_ on host arrays runs in chunks of whole atoms.  Each of two slots holds one
    // chunk: its inputs, packed by the host into pinned memory and copied in one transfer to
    // the device, and its outputs, copied back in one transfer.  While the kernels run on
    // one slot, the copies and the host's packing and unpacking work on the other.
    struct Slot {
        size_t cap_in = 0, cap_out = 0;  // bytes
        char* h_in = nullptr;            // pinned: chunk inputs (layout in ChunkLayout)
        char* d_in = nullptr;            // device copy of h_in
        char* d_out = nullptr;           // device: eout | tlist0 | flist | tlist
        char* h_out = nullptr;           // pinned copy of d_out
        hipEvent_t ev_in = nullptr, ev_comp = nullptr, ev_done = nullptr;
    };
    std::array<Slot, 2> slots;
    hipStream_t s_in = nullptr, s_out = nullptr;
    // fork and join of an evaluation run as two halves (see run_device)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool uploaded = false;
    int device = -1;  // the device the tables were uploaded to
    // fitting: adjoint buffers, pair outputs of the fit forward, device coefficient structure
    FitBuffers fit;
    size_t fit_cap_pairs = 0;
    double *f_fl = nullptr, *f_tl = nullptr, *f_t0 = nullptr, *f_eout = nullptr;
    size_t fit_cap_atoms = 0;
    size_t fit_cap_wd = 0;     // atoms that fit.Wd holds
    int fit_natom = -1;
    // per-kernel timing (magace_profile): 4 events per evaluation
    bool profiling = false;
    std::vector<std::array<hipEvent_t, 4>> events;
    size_t events_used = 0;
    ~Context();
};

std::unique_ptr<Context> g_ctx;

// Tuned density variants per potential header, measured on MI355X by mag-ace_amd/tune.py
// (which writes tuning_table.inc and the record tuning.json); headers not listed use the
// built-in rule.
struct TuneEntry {
    int hdr[12];  // nb_a nb_m nr2_m l2m nr3_a nr3_m l3a l3m nr4_a nr4_m l4a l4m
    int fp64[2];  // variant, neighbour split
    int mixed[2];
};
const TuneEntry kTuning[] = {
#include "tuning_table.inc"
    {{-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1}, {-1, 0}, {-1, 0}},  // end
};

void apply_tuning_table(Context& c) {
    const Header& h = c.model.pot.h;
    const int key[12] = {h.nb_a, h.nb_m, h.nr2_m, h.l2m, h.nr3_a, h.nr3_m, h.l3a, h.l3m,
                         h.nr4_a, h.nr4_m, h.l4a, h.l4m};
    for (const TuneEntry& e : kTuning) {
        if (e.hdr[0] < 0) break;
        if (std::equal(key, key + 12, e.hdr)) {
            c.dvar[0] = e.fp64[0]; c.dsplit[0] = e.fp64[1];
            c.dvar[1] = e.mixed[0]; c.dsplit[1] = e.mixed[1];
            break;
        }
    }
    c.P64.sa = c.dsplit[0];
    c.P32.sa = c.dsplit[1];
}

// generation of the state baked into captured evaluations (context_generation)
std::atomic<unsigned long long> g_generation{1};
void bump_generation() { g_generation.fetch_add(1); }

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            set_error(MAGACE_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(_e) + \
                                             " at " #expr);                               \
            return MAGACE_ERR_DEVICE;                                                     \
        }                                                                                 \
    } while (0)

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

Context::~Context() {
    dfree(t64.lut_a); dfree(t64.lut_m); dfree(t32.lut_a); dfree(t32.lut_m);
    dfree(crec64); dfree(crec32); dfree(crec16); dfree(crecu); dfree(crec16u); dfree(chunks); dfree(comb); dfree(comb2); dfree(fold); dfree(cw2_64);
    dfree(ypair); dfree(yft); dfree(yfe);
    dfree(cw2_32);
    dfree(buf.Z); dfree(buf.W); dfree(buf.ehc); dfree(buf.rho1); dfree(buf.dE); dfree(buf.edf); dfree(buf.rho);
    dfree(buf.Pt);
    dfree(buf.err);
    dfree(fit.Zdot); dfree(fit.ehc_dot); dfree(fit.alpha); dfree(fit.beta); dfree(fit.M); dfree(fit.idx);
    dfree(fit.ub_ptr); dfree(fit.ub_t); dfree(fit.ub_u); dfree(fit.Wd);
    dfree(f_fl); dfree(f_tl); dfree(f_t0); dfree(f_eout);
    for (auto& sl : slots) {
        dfree(sl.d_in);
        dfree(sl.d_out);
        if (sl.h_in) (void)hipHostFree(sl.h_in);
        if (sl.h_out) (void)hipHostFree(sl.h_out);
        if (sl.ev_in) (void)hipEventDestroy(sl.ev_in);
        if (sl.ev_comp) (void)hipEventDestroy(sl.ev_comp);
        if (sl.ev_done) (void)hipEventDestroy(sl.ev_done);
    }
    if (s_in) (void)hipStreamDestroy(s_in);
    if (s_out) (void)hipStreamDestroy(s_out);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    for (auto& e : events)
        for (auto& x : e) (void)hipEventDestroy(x);
    if (stream) (void)hipStreamDestroy(stream);
}

template <typename T>
void set_c1(DevParams<T>& P, const Model& M);

template <typename T>
void fill_params(DevParams<T>& P, const Model& M) {
    const Header& h = M.pot.h;
    P.stride_a = M.spl.atomic_stride;
    P.stride_m = M.spl.magnetic_stride;
    P.nlut = M.spl.nlut;
    P.rs_a = M.spl.rs_a;
    P.rs_m = M.spl.rs_m;
    P.nb_a = h.nb_a; P.nb_m = h.nb_m; P.nr2m = h.nr2_m; P.l2m = h.l2m;
    P.nr3a = h.nr3_a; P.nr3m = h.nr3_m; P.l3a = h.l3a; P.l3m = M.z.lAm;  // the magnetic l range of A
    P.lsa = M.z.lsa; P.lsm = M.z.lsm; P.lsm3 = M.z.lsm3; P.nm3 = M.z.nm3;
    P.oA2 = M.z.oA2; P.oA0 = M.z.oA0; P.nz = M.z.nz;
    for (int q = 0; q < M.z.nm3 && q < 32; ++q) {
        const int lq = q % M.z.lsm3;
        int lp = 0;
        while ((lp + 1) * (lp + 1) <= lq) ++lp;
        P.mpq[q] = (signed char)(lq - lp * lp - lp);
    }
    P.lha = (h.l3a + 1) * (h.l3a + 2) / 2;
    P.lhm = (h.l2m + 1) * (h.l2m + 2) / 2;
    P.nchunks = (int)(M.terms.chunks.size() / 2);
    P.ncomb = (int)(M.terms.comb.size() / 2);
    P.ncomb2 = (int)(M.terms.comb2.size() / 2);
    P.nslot = M.terms.nslot;
    P.ny = M.terms.ny;
    P.nyt = (int)(M.terms.yft.size() / 2);
    P.nzx = M.terms.nzx();
    P.nfemb = h.nfemb;
    P.nd = h.ndensity;
    for (int d = 0; d < 2; ++d) {
        P.p1[d][0] = M.pot.p1[d][0];
        P.p1[d][1] = M.pot.p1[d][1];
    }
    P.ecut = h.ecut;
    P.decut = h.decut;
    set_c1(P, M);
}

template <typename T>
void set_c1(DevParams<T>& P, const Model& M) {
    const size_t nb1 = M.pot.basis.b1.size();
    for (int d = 0; d < 2; ++d)
        for (int k = 0; k < 32; ++k)
            P.c1[d][k] = d < M.pot.h.ndensity && k < (int)nb1 ? M.pot.c1[d * nb1 + k] : 0.0;
}

template <typename T>
int upload_table(T*& dst, const std::vector<double>& src) {
    std::vector<T> tmp(src.begin(), src.end());
    HIPCHK(hipMalloc((void**)&dst, tmp.size() * sizeof(T) + 64));
    HIPCHK(hipMemcpy(dst, tmp.data(), tmp.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

int build_context(Potential&& pot, std::unique_ptr<Context>& out) {
    auto ctx = std::make_unique<Context>();
    Model& M = ctx->model;
    M.pot = std::move(pot);
    std::string err;
    if (!check_header(M.pot.h, err)) {
        set_error(MAGACE_ERR_INPUT, err);
        return MAGACE_ERR_INPUT;
    }
    M.z = make_layout(M.pot.h);
    if (16 * (M.z.nz + 1) >= 65536) {
        set_error(MAGACE_ERR_INPUT, "feature vector too long (the adjoint addresses Z with 16-bit byte offsets)");
        return MAGACE_ERR_INPUT;
    }
    if (!M.pot.basis.undefined.empty()) {
        set_error(MAGACE_ERR_INPUT, M.pot.basis.undefined);
        return MAGACE_ERR_INPUT;
    }
    const Header& h = M.pot.h;
    ctx->shape_class = supported_nm3_class(M.z.nm3, h.nr2_m, h.nb_a, h.l3a, h.l2m, h.nr3_a);
    if (ctx->shape_class < 0) {
        set_error(MAGACE_ERR_INPUT, "potential shape outside the compiled kernel classes");
        return MAGACE_ERR_INPUT;
    }
    if (h.nb_m > 32 || M.z.nm3 > 32) {
        set_error(MAGACE_ERR_INPUT, "nradbase_magnetic > 32 or more than 32 magnetic columns");
        return MAGACE_ERR_INPUT;
    }
    M.terms = build_terms(M.pot, M.z);
    if (M.terms.ncontrib() >= (size_t(1) << 27)) {
        set_error(MAGACE_ERR_INPUT, "gradient record table too large (2^27 records)");
        return MAGACE_ERR_INPUT;
    }
    // chunk headers hold the target in 12 bits and the run's slot in 13; records address the
    // extended vector [Z | 1 | Y] with 16-bit byte offsets of 16-byte entries
    if (16 * (size_t)M.terms.nzx() >= 65536 || M.terms.nslot >= 8192) {
        set_error(MAGACE_ERR_INPUT, "gradient table too large (feature vector and products >= 4096 entries or "
                                    ">= 8192 partial sums)");
        return MAGACE_ERR_INPUT;
    }
    M.spl = build_splines(M.pot);
    fill_params(ctx->P64, M);
    fill_params(ctx->P32, M);
    apply_tuning_table(*ctx);
    out = std::move(ctx);
    return 0;
}

// 16-byte gradient records: the three factors as byte offsets into the LDS copy of Z
// (a | b << 16, c; entry size 16 bytes in FP64, 8 in FP32), then the weight as FP64 bits or
// FP32 bits.  Offsets save the kernel the index arithmetic; init checks 16 (nz + 1) < 2^16.
int upload_records(Context& c) {
    const TermList& T = c.model.terms;
    const size_t nc = T.ncontrib();
    std::vector<uint4> r64(nc), r32(nc), r16(nc), r32u(nc), r16u(nc);
    for (size_t i = 0; i < nc; ++i) {
        const uint16_t* q = &T.cidx[4 * i];
        const double w = T.cw[i];
        uint64_t wb;
        std::memcpy(&wb, &w, 8);
        const float wf = (float)w;
        uint32_t fb;
        std::memcpy(&fb, &wf, 4);
        r64[i] = make_uint4(((uint32_t)q[0] << 4) | ((uint32_t)q[1] << 20), (uint32_t)q[2] << 4,
                            (uint32_t)(wb & 0xffffffffu), (uint32_t)(wb >> 32));
        // the FP32 records carry the weight's low part too (w - (float)w), read by the
        // accumulation recipe's FP64 sums (MAGACE_MIXED_ACC)
        const float lo = (float)(w - (double)wf);
        uint32_t lb;
        std::memcpy(&lb, &lo, 4);
        r32[i] = make_uint4(((uint32_t)q[0] << 3) | ((uint32_t)q[1] << 19), (uint32_t)q[2] << 3, fb, lb);
        r16[i] = make_uint4(((uint32_t)q[0] << 4) | ((uint32_t)q[1] << 20), (uint32_t)q[2] << 4, fb, lb);
        r32u[i] = make_uint4((uint32_t)q[0] << 3, (uint32_t)q[1] << 3, (uint32_t)q[2] << 3, fb);
        r16u[i] = make_uint4((uint32_t)q[0] << 4, (uint32_t)q[1] << 4, (uint32_t)q[2] << 4, fb);
    }
    HIPCHK(hipMemcpy(c.crec64, r64.data(), nc * sizeof(uint4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.crec32, r32.data(), nc * sizeof(uint4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.crec16, r16.data(), nc * sizeof(uint4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.crecu, r32u.data(), nc * sizeof(uint4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.crec16u, r16u.data(), nc * sizeof(uint4), hipMemcpyHostToDevice));
    if (T.nd > 1) {
        if (!c.cw2_64) {
            HIPCHK(hipMalloc((void**)&c.cw2_64, nc * sizeof(double) + 64));
            HIPCHK(hipMalloc((void**)&c.cw2_32, nc * sizeof(float) + 64));
        }
        std::vector<float> w32(nc);
        for (size_t i = 0; i < nc; ++i) w32[i] = (float)T.cw[nc + i];
        HIPCHK(hipMemcpy(c.cw2_64, T.cw.data() + nc, nc * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c.cw2_32, w32.data(), nc * sizeof(float), hipMemcpyHostToDevice));
        c.P64.cw2 = c.cw2_64;
        c.P32.cw2 = c.cw2_32;
        c.P32.cw2d = c.cw2_64;
    }
    return 0;
}

// merged weights w_t = sum_k c_k u_{k,t} and record weights after a coefficient change
int refresh_weights(Context& c) {
    Potential& pot = c.model.pot;
    TermList& T = c.model.terms;
    const Basis& B = pot.basis;
    const size_t nt = T.nterms(), nr = T.ncontrib();
    for (int d = 0; d < T.nd; ++d) {
        // coefficients of the B functions of orders 2..6 (the index space of T.ut_k)
        std::vector<double> cv;
        for (const auto* c : {&pot.c2, &pot.c3, &pot.c4, &pot.c5, &pot.c6}) {
            const size_t n = c->size() / T.nd;
            cv.insert(cv.end(), c->begin() + d * n, c->begin() + (d + 1) * n);
        }
        for (size_t t = 0; t < nt; ++t) {
            double w = 0.0;
            for (int e = T.ut_ptr[t]; e < T.ut_ptr[t + 1]; ++e) w += cv[T.ut_k[e]] * T.ut_u[e];
            T.w[d * nt + t] = w;
        }
        for (size_t r = 0; r < nr; ++r) T.cw[d * nr + r] = T.rec_t[r] >= 0 ? T.w[d * nt + T.rec_t[r]] : 0.0;
    }
    set_c1(c.P64, c.model);
    set_c1(c.P32, c.model);
    bump_generation();  // c1 travels by value in the kernel arguments
    if (c.uploaded && nr) return upload_records(c);
    return 0;
}

// device tables are uploaded on first use so the host model can be built without a GPU
int upload(Context& c) {
    if (c.uploaded) return 0;
    const Model& M = c.model;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(upload_legendre_constants());
    if (scratch_setup() != 0) {
        set_error(MAGACE_ERR_DEVICE, "hipGetSymbolAddress failed for the fixed evaluation scratch");
        return MAGACE_ERR_DEVICE;
    }
    HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    if (upload_table(c.t64.lut_a, M.spl.atomic)) return MAGACE_ERR_DEVICE;
    if (upload_table(c.t64.lut_m, M.spl.magnetic)) return MAGACE_ERR_DEVICE;
    if (upload_table(c.t32.lut_a, M.spl.atomic)) return MAGACE_ERR_DEVICE;
    if (upload_table(c.t32.lut_m, M.spl.magnetic)) return MAGACE_ERR_DEVICE;
    const size_t nc = M.terms.ncontrib(), nch = M.terms.chunks.size() / 2;
    HIPCHK(hipMalloc((void**)&c.crec64, nc * sizeof(uint4) + 64));
    HIPCHK(hipMalloc((void**)&c.crec32, nc * sizeof(uint4) + 64));
    HIPCHK(hipMalloc((void**)&c.crec16, nc * sizeof(uint4) + 64));
    HIPCHK(hipMalloc((void**)&c.crecu, nc * sizeof(uint4) + 64));
    HIPCHK(hipMalloc((void**)&c.crec16u, nc * sizeof(uint4) + 64));
    HIPCHK(hipMalloc((void**)&c.chunks, nch * sizeof(int2) + 64));
    const size_t ncomb = M.terms.comb.size() / 2;
    HIPCHK(hipMalloc((void**)&c.comb, ncomb * sizeof(int2) + 64));
    if (ncomb) HIPCHK(hipMemcpy(c.comb, M.terms.comb.data(), ncomb * sizeof(int2), hipMemcpyHostToDevice));
    const size_t ncomb2 = M.terms.comb2.size() / 2;
    HIPCHK(hipMalloc((void**)&c.comb2, ncomb2 * sizeof(int2) + 64));
    if (ncomb2) HIPCHK(hipMemcpy(c.comb2, M.terms.comb2.data(), ncomb2 * sizeof(int2), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc((void**)&c.fold, M.terms.fold.size() * sizeof(uint32_t) + 64));
    HIPCHK(hipMemcpy(c.fold, M.terms.fold.data(), M.terms.fold.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    auto up_int2 = [](int2*& d, const std::vector<int32_t>& v) -> int {
        HIPCHK(hipMalloc((void**)&d, v.size() * sizeof(int32_t) + 64));
        if (!v.empty()) HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        return 0;
    };
    if (int rc = up_int2(c.ypair, M.terms.ypair)) return rc;
    if (int rc = up_int2(c.yft, M.terms.yft)) return rc;
    if (int rc = up_int2(c.yfe, M.terms.yfe)) return rc;
    if (nc) {
        if (int rc = upload_records(c)) return rc;
        HIPCHK(hipMemcpy(c.chunks, M.terms.chunks.data(), nch * sizeof(int2), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMalloc((void**)&c.buf.err, sizeof(int)));
    HIPCHK(hipMemset(c.buf.err, 0, sizeof(int)));
    c.P64.lut_a = c.t64.lut_a; c.P64.lut_m = c.t64.lut_m;
    c.P32.lut_a = c.t32.lut_a; c.P32.lut_m = c.t32.lut_m;
    c.P64.crec = c.crec64;
    c.P32.crec = c.crec32;
    c.P32.crec16 = c.crec16;
    c.P32.crecu = c.crecu;
    c.P32.crec16u = c.crec16u;
    c.P64.chunks = c.P32.chunks = c.chunks;
    c.P64.comb = c.P32.comb = c.comb;
    c.P64.comb2 = c.P32.comb2 = c.comb2;
    c.P64.fold = c.P32.fold = c.fold;
    c.P64.ypair = c.P32.ypair = c.ypair;
    c.P64.yft = c.P32.yft = c.yft;
    c.P64.yfe = c.P32.yfe = c.yfe;
    c.uploaded = true;
    c.device = dev;
    bump_generation();
    return 0;
}

int ensure_atoms(Context& c, size_t natom) {
    if (natom <= c.buf.cap_atoms) return 0;
    const size_t cap = natom + natom / 4 + 16;
    const size_t nz = (size_t)c.model.z.nz;
    dfree(c.buf.Z); dfree(c.buf.W); dfree(c.buf.ehc); dfree(c.buf.rho1); dfree(c.buf.dE); dfree(c.buf.edf);
    HIPCHK(hipMalloc(&c.buf.Z, cap * nz * 2 * sizeof(double)));
    HIPCHK(hipMalloc(&c.buf.W, cap * nz * 2 * sizeof(double)));
    HIPCHK(hipMalloc((void**)&c.buf.ehc, cap * sizeof(double)));
    const size_t nd = (size_t)c.model.pot.h.ndensity;
    HIPCHK(hipMalloc((void**)&c.buf.rho1, cap * nd * sizeof(double)));
    HIPCHK(hipMalloc((void**)&c.buf.dE, cap * nd * sizeof(double)));
    HIPCHK(hipMalloc((void**)&c.buf.edf, cap * sizeof(double)));
    dfree(c.buf.rho);
    HIPCHK(hipMalloc((void**)&c.buf.rho, cap * nd * sizeof(double)));
    dfree(c.buf.Pt);
    HIPCHK(hipMalloc(&c.buf.Pt, cap * partials_bytes(c.model.terms.nslot, (int)nd) + 256));
    c.buf.cap_atoms = cap;
    bump_generation();  // captured evaluations refer to the old buffers
    return 0;
}

// Evaluations of at least kSplitAtoms atoms run as two halves of the atom range, the second on
// the context's own stream (used otherwise only by acelib_), forked from the caller's stream
// and joined back to it; no further stream is created, since the process's streams share
// four hardware queues and a fifth one slowed acelib_'s copy pipeline.  Every
// kernel is independent per atom, so the halves only overlap: the tail of one half's launch
// and the next kernel of the other share the CUs, and an LDS-bound adjoint runs beside the
// VALU-bound forces.  The second half uses its own rows of the per-atom device buffers; pair
// arrays are indexed by absolute pair index.  Not while per-kernel times are measured, nor
// while the caller's stream is being captured into a graph, nor inside acelib_'s chunk
// pipeline, whose kernels run on that stream.  MAGACE_SPLIT_STREAMS=0 turns the split off
// (read at every call).
constexpr int kSplitAtoms = 2048;

bool split_enabled() {
    const char* e = std::getenv("MAGACE_SPLIT_STREAMS");
    return !(e && e[0] == '0' && e[1] == 0);
}

// the device buffers of atoms [a, ...)
Buffers buffers_from(const Buffers& B, size_t a, size_t zbytes, int nd, int nslot) {
    Buffers o = B;
    o.Pt = (char*)B.Pt + a * partials_bytes(nslot, nd);
    o.Z = (char*)B.Z + a * zbytes;
    o.W = (char*)B.W + a * zbytes;
    o.ehc = B.ehc + a;
    o.rho1 = B.rho1 + a * nd;
    o.dE = B.dE + a * nd;
    o.edf = B.edf + a;
    o.rho = B.rho + a * nd;
    if (B.Wd) o.Wd = (char*)B.Wd + a * 2 * zbytes;  // set only for FP64 (fitting)
    o.cap_atoms = B.cap_atoms - a;
    return o;
}

hipError_t launch(Context& c, int natom, int atom_base, const int* off, const double* dlist, const double* elist,
                  const double* dmomi, const double* emomi, const double* dmom, const double* emom,
                  double* eout, double* flist, double* tlist, double* tlist0, Buffers& B, hipStream_t s,
                  hipEvent_t* ev) {
    if (c.precision == MAGACE_FP64)
        return Launch<double>::run(c.P64, natom, atom_base, off, dlist, elist, dmomi, emomi, dmom, emom, eout, flist,
                                   tlist, tlist0, B, s, c.shape_class, ev, c.dvar[0]);
    // MAGACE_MIXED_ACC runs the FP32 kernels with the accumulation recipe and the mixed tuning
    return Launch<float>::run(c.P32, natom, atom_base, off, dlist, elist, dmomi, emomi, dmom, emom, eout, flist,
                              tlist, tlist0, B, s, c.shape_class, ev, c.dvar[1], c.precision == MAGACE_MIXED_ACC);
}

int run_device(Context& c, int natom, const int* off, const double* dlist, const double* elist,
               const double* dmomi, const double* emomi, const double* dmom, const double* emom,
               double* eout, double* flist, double* tlist, double* tlist0, hipStream_t s,
               bool clear_err = true, bool allow_split = true, Buffers* ext = nullptr, int atom_base = 0) {
    // any evaluation overwrites the per-atom buffers a fit forward left for its backward
    c.fit_natom = -1;
    if (!ext) {
        if (int rc = upload(c)) return rc;
        if (int rc = ensure_atoms(c, (size_t)natom)) return rc;
    }
    Buffers& buf = ext ? *ext : c.buf;
    if (clear_err) HIPCHK(clear_flags(c.buf.err, s));
    hipEvent_t* ev = nullptr;
    if (c.profiling && !ext) {
        if (c.events_used == c.events.size()) {
            std::array<hipEvent_t, 4> q;
            for (auto& x : q) HIPCHK(hipEventCreate(&x));
            c.events.push_back(q);
        }
        ev = c.events[c.events_used++].data();
    }
    bool split = allow_split && s != c.stream && !c.profiling && natom >= kSplitAtoms && split_enabled();
    if (split) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(s, &cs));
        split = cs == hipStreamCaptureStatusNone;
    }
    if (!split) {
        HIPCHK(launch(c, natom, atom_base, off, dlist, elist, dmomi, emomi, dmom, emom, eout, flist, tlist, tlist0,
                      buf, s, ev));
        return 0;
    }
    // an even split keeps the two-atom adjoint's atom pairs (and so every result bit) as in
    // one sequence
    const int h = (natom / 2) & ~1;
    const size_t zb = (size_t)c.model.z.nz * (c.precision == MAGACE_FP64 ? 16 : 8);
    Buffers B1 = buffers_from(buf, (size_t)h, zb, c.model.pot.h.ndensity, c.model.terms.nslot);
    HIPCHK(hipEventRecord(c.ev_fork, s));
    HIPCHK(hipStreamWaitEvent(c.stream, c.ev_fork, 0));
    HIPCHK(launch(c, h, atom_base, off, dlist, elist, dmomi, emomi, dmom, emom, eout, flist, tlist, tlist0, buf, s,
                  nullptr));
    HIPCHK(launch(c, natom - h, atom_base + h, off + h, dlist, elist, dmomi + h, emomi + 3 * (size_t)h, dmom, emom,
                  eout + h, flist, tlist, tlist0 + 3 * (size_t)h, B1, c.stream, nullptr));
    HIPCHK(hipEventRecord(c.ev_join, c.stream));
    HIPCHK(hipStreamWaitEvent(s, c.ev_join, 0));
    return 0;
}

// pairs per acelib_ chunk: MAGACE_ACELIB_CHUNK_PAIRS if set (1024 .. 2^30), else 393216; read
// at every call.  Each of the two slots holds, in pinned host memory and again on the
// device, 112 bytes per pair and 68 per atom of one chunk (44 MB at the default).  On the
// 16000-atom lmax-6 workload (917170 pairs) 393216 was the fastest of 131072 .. 524288
// (2.80 ms per call against 2.91 at 262144 and 2.99 at 524288).
int chunk_pairs() {
    if (const char* e = std::getenv("MAGACE_ACELIB_CHUNK_PAIRS")) {
        const long v = std::atol(e);
        if (v >= 1024 && v <= (1L << 30)) return (int)v;
    }
    return 393216;
}

// Byte offsets of one acelib_ chunk of na atoms and np pairs.  Inputs: dlist, elist,
// dlist_mom, elist_mom, dlist_momi, elist_momi, then the chunk's pair offsets (na + 1 ints,
// relative to its first pair); outputs: eout, tlist0, flist, tlist.
struct ChunkLayout {
    size_t na, np;
    size_t i_dl, i_el, i_dm, i_em, i_dmi, i_emi, i_off, in_bytes;
    size_t o_eout, o_t0, o_fl, o_tl, out_bytes;
    ChunkLayout(size_t na_, size_t np_) : na(na_), np(np_) {
        i_dl = 0;
        i_el = i_dl + 8 * np;
        i_dm = i_el + 24 * np;
        i_em = i_dm + 8 * np;
        i_dmi = i_em + 24 * np;
        i_emi = i_dmi + 8 * na;
        i_off = i_emi + 24 * na;
        in_bytes = i_off + 4 * (na + 1);
        o_eout = 0;
        o_t0 = o_eout + 8 * na;
        o_fl = o_t0 + 24 * na;
        o_tl = o_fl + 24 * np;
        out_bytes = o_tl + 24 * np;
    }
};

// Atom boundaries of the chunks (cut[0] = 0, ..., cut.back() = natom).  Chunks hold whole
// atoms and at most cmax pairs (unless one atom has more).  The first chunks grow from
// cmax/4 and the last ones shrink by halves down to cmax/4, so that the host's first packing
// and last unpacking, which nothing overlaps, are short; smaller chunks would cost more in
// kernel latency than they save (each chunk's three kernels take 0.3 ms at least).
std::vector<int> chunk_plan(const std::vector<int>& off, int cmax) {
    const int n = (int)off.size() - 1;
    const long total = off[n];
    const long base = std::max(1L, (long)cmax / 4);
    std::vector<int> cut{0};
    for (int i = 0, k = 0; i < n; ++k) {
        const long rest = total - off[i];
        long target = std::min<long>(cmax, base << std::min(k, 20));
        target = std::min(target, std::max(base, (rest + 1) / 2));
        const int a0 = i;
        ++i;
        while (i < n && off[i + 1] - off[a0] <= target) ++i;
        if (i < n && (i & 1)) ++i;  // chunks start at even atoms (the adjoint's atom pairs)
        cut.push_back(i);
    }
    return cut;
}

int check_kernel_errors(Context& c, hipStream_t s) {
    int err = 0;
    HIPCHK(hipMemcpyAsync(&err, c.buf.err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (err & kErrSmallDistance) {
        set_error(MAGACE_ERR_INPUT, "Encountered very small distance.");
        return MAGACE_ERR_INPUT;
    }
    if (err & kErrMomentRange) {
        set_error(MAGACE_ERR_INPUT, "magnetic moment outside [0, rcut_magnetic] (spline table range)");
        return MAGACE_ERR_INPUT;
    }
    return 0;
}

// Per-atom buffers of the allocation-free entry point, carved from the fixed device scratch:
// the adjoint's partial sums (pbytes per atom), Z and W rows of zbytes each, then ehc, rho1,
// dE, edf, rho; every region 256-byte aligned.
// Returns the number of atoms that fit.
size_t carve_scratch(char* base, size_t bytes, size_t zbytes, int nd, size_t pbytes, Buffers* B) {
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t per_atom = 2 * zbytes + pbytes + sizeof(double) * (2 + 3 * (size_t)nd);
    size_t cap = bytes / per_atom;
    while (cap > 0 && up(cap * zbytes) * 2 + up(cap * pbytes) + up(cap * 8) * 2 + up(cap * nd * 8) * 3 > bytes) --cap;
    if (B && cap > 0) {
        size_t o = 0;
        B->Pt = base + o; o += up(cap * pbytes);
        B->Z = base + o; o += up(cap * zbytes);
        B->W = base + o; o += up(cap * zbytes);
        B->ehc = (double*)(base + o); o += up(cap * 8);
        B->rho1 = (double*)(base + o); o += up(cap * nd * 8);
        B->dE = (double*)(base + o); o += up(cap * nd * 8);
        B->edf = (double*)(base + o); o += up(cap * 8);
        B->rho = (double*)(base + o); o += up(cap * nd * 8);
        B->cap_atoms = cap;
    }
    return cap;
}

}  // namespace

unsigned long long context_generation() { return g_generation.load(); }

int context_prepare(int natom) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    if (int rc = upload(*g_ctx)) return rc;
    return ensure_atoms(*g_ctx, (size_t)std::max(natom, 0));
}

}  // namespace magace

using namespace magace;

extern "C" {

int magace_last_error(void) { return g_errcode; }
const char* magace_error_string(void) { return g_err.c_str(); }

int magace_init_file(const char* path, const char* element, double rcut, double rcut_magnetic,
                     double smear, int xcnumber) {
    clear_error();
    Potential pot;
    std::string err;
    if (!parse_potential(path, element, xcnumber, smear, pot, err)) {
        set_error(MAGACE_ERR_INPUT, err);
        return MAGACE_ERR_INPUT;
    }
    // cutoff consistency (src/util.F90:420-444)
    if (pot.h.cut_a > rcut || pot.h.cut_m > rcut_magnetic) {
        set_error(MAGACE_ERR_INPUT, "Global cutoff too short.");
        return MAGACE_ERR_INPUT;
    }
    pot.rcut = rcut;
    pot.rcut_m = rcut_magnetic;
    std::unique_ptr<Context> ctx;
    int prec = g_ctx ? g_ctx->precision : MAGACE_FP64;
    if (int rc = build_context(std::move(pot), ctx)) return rc;
    ctx->precision = prec;
    g_ctx = std::move(ctx);
    bump_generation();
    // With a GPU visible the tables go to the current device now, so that the
    // allocation-free entry point finds them (and its scratch address) ready; without one
    // the host model stays usable and the upload waits for the first evaluation.
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) == hipSuccess && ndev > 0) return upload(*g_ctx);
    (void)hipGetLastError();
    return 0;
}

void acelibini_(int* nele, double* rcut, double* rcut_magnetic, double* smear, int* xcnumber,
                int* elenumbers) {
    clear_error();
    if (g_ctx) return;  // initialisation runs once per process (src/util.F90:29-33)
    if (*nele != 1) {
        set_error(MAGACE_ERR_INPUT, "Multiple elements require fixing acelibini call first.");
        return;
    }
    if (*xcnumber != 1 && *xcnumber != 2) {
        set_error(MAGACE_ERR_INPUT, "Unknown XC functional received by acelibini.");
        return;
    }
    const int z1 = elenumbers[0];
    if (z1 < 1 || z1 > 104) {
        set_error(MAGACE_ERR_INPUT, "element number out of range");
        return;
    }
    const std::string el = kElements[z1 - 1];
    const std::string file = el + "." + (*xcnumber == 1 ? "pbe" : "lda") + ".in";
    magace_init_file(file.c_str(), el.c_str(), *rcut, *rcut_magnetic, *smear, *xcnumber);
}

int magace_finalize(void) {
    g_ctx.reset();
    bump_generation();
    return 0;
}

int magace_set_precision(int mode) {
    if (mode != MAGACE_FP64 && mode != MAGACE_MIXED && mode != MAGACE_MIXED_ACC) {
        set_error(MAGACE_ERR_INPUT, "precision must be MAGACE_FP64, MAGACE_MIXED or MAGACE_MIXED_ACC");
        return MAGACE_ERR_INPUT;
    }
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    if (g_ctx->precision != mode) bump_generation();
    g_ctx->precision = mode;
    return 0;
}

int magace_get_precision(void) { return g_ctx ? g_ctx->precision : -1; }

int magace_set_tuning(const int* v, int n) {
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    if (!v || n != 4) {
        set_error(MAGACE_ERR_INPUT, "magace_set_tuning: 4 values (FP64 variant, split, mixed variant, split)");
        return MAGACE_ERR_INPUT;
    }
    for (int p = 0; p < 2; ++p)
        if (v[2 * p] < -1 || v[2 * p] >= kDensityVariants || v[2 * p + 1] < 0 || v[2 * p + 1] > 8) {
            set_error(MAGACE_ERR_INPUT, "magace_set_tuning: variant in -1..2, split in 0..8");
            return MAGACE_ERR_INPUT;
        }
    Context& c = *g_ctx;
    for (int p = 0; p < 2; ++p) {
        c.dvar[p] = v[2 * p];
        c.dsplit[p] = v[2 * p + 1];
    }
    c.P64.sa = c.dsplit[0];
    c.P32.sa = c.dsplit[1];
    bump_generation();  // captured evaluations hold the old variant
    return 0;
}

int magace_get_tuning(int* v, int n) {
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    if (!v || n != 4) {
        set_error(MAGACE_ERR_INPUT, "magace_get_tuning: 4 values");
        return MAGACE_ERR_INPUT;
    }
    for (int p = 0; p < 2; ++p) {
        v[2 * p] = g_ctx->dvar[p];
        v[2 * p + 1] = g_ctx->dsplit[p];
    }
    return 0;
}

int magace_tune(int natom, int nei, const int* offsets, const double* dlist, const double* elist,
                const double* dlist_momi, const double* elist_momi, const double* dlist_mom,
                const double* elist_mom, double* eout, double* flist, double* tlist, double* tlist0, int reps,
                double* ms, int* best, void* stream) {
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    if (natom <= 0 || nei < 0 || reps <= 0 || !ms || !best) {
        set_error(MAGACE_ERR_INPUT, "magace_tune: invalid arguments");
        return MAGACE_ERR_INPUT;
    }
    Context& c = *g_ctx;
    hipStream_t s = (hipStream_t)stream;
    const int p = c.precision == MAGACE_FP64 ? 0 : 1;
    // (variant, neighbour split) candidates, MAGACE_TUNE_CANDIDATES of them
    static const int cand[MAGACE_TUNE_CANDIDATES][2] = {{0, 0}, {1, 1}, {1, 2}, {1, 3}, {1, 4},
                                                        {2, 1}, {2, 2}, {2, 3}, {2, 4}};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int bi = -1;
    for (int k = 0; k < MAGACE_TUNE_CANDIDATES; ++k) {
        c.dvar[p] = cand[k][0];
        c.dsplit[p] = cand[k][1];
        (p ? c.P32.sa : c.P64.sa) = cand[k][1];
        ms[k] = -1.0;
        // a variant whose LDS does not fit this potential fails at launch and is skipped
        int rc = run_device(c, natom, offsets, dlist, elist, dlist_momi, elist_momi, dlist_mom, elist_mom, eout,
                            flist, tlist, tlist0, s);
        if (rc == 0) rc = check_kernel_errors(c, s);
        if (rc) {
            clear_error();
            (void)hipGetLastError();
            continue;
        }
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && rc == 0; ++r)
            rc = run_device(c, natom, offsets, dlist, elist, dlist_momi, elist_momi, dlist_mom, elist_mom, eout, flist,
                            tlist, tlist0, s);
        (void)hipEventRecord(e1, s);
        if (hipEventSynchronize(e1) != hipSuccess || rc) {
            clear_error();
            (void)hipGetLastError();
            continue;
        }
        float t = 0.f;
        (void)hipEventElapsedTime(&t, e0, e1);
        ms[k] = t / reps;
        if (bi < 0 || ms[k] < ms[bi]) bi = k;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (bi < 0) {
        set_error(MAGACE_ERR_DEVICE, "magace_tune: no variant ran");
        return MAGACE_ERR_DEVICE;
    }
    c.dvar[p] = cand[bi][0];
    c.dsplit[p] = cand[bi][1];
    (p ? c.P32.sa : c.P64.sa) = cand[bi][1];
    best[0] = cand[bi][0];
    best[1] = cand[bi][1];
    bump_generation();
    return 0;
}

int magace_compute_device(int natom, int nei, const int* offsets, const double* dlist,
                          const double* elist, const double* dlist_momi, const double* elist_momi,
                          const double* dlist_mom, const double* elist_mom, double* eout,
                          double* flist, double* tlist, double* tlist0, void* stream) {
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    if (natom < 0 || nei < 0) {
        set_error(MAGACE_ERR_INPUT, "negative sizes");
        return MAGACE_ERR_INPUT;
    }
    if (int rc = upload(*g_ctx)) return rc;
    hipStream_t s = (hipStream_t)stream;  // NULL: the default (null) stream
    if (int rc = run_device(*g_ctx, natom, offsets, dlist, elist, dlist_momi, elist_momi, dlist_mom, elist_mom,
                            eout, flist, tlist, tlist0, s))
        return rc;
    return 0;
}

int magace_compute_device_noalloc(int natom, int nei, const int* offsets, const double* dlist,
                                  const double* elist, const double* dlist_momi, const double* elist_momi,
                                  const double* dlist_mom, const double* elist_mom, double* eout,
                                  double* flist, double* tlist, double* tlist0, void* stream) {
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    Context& c = *g_ctx;
    if (natom < 0 || nei < 0) {
        set_error(MAGACE_ERR_INPUT, "negative sizes");
        return MAGACE_ERR_INPUT;
    }
    if (!c.uploaded) {
        set_error(MAGACE_ERR_STATE, "magace_compute_device_noalloc: the potential's tables are not on the device "
                                    "(initialise with a GPU visible, or run one magace_compute_device first)");
        return MAGACE_ERR_STATE;
    }
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    void* base = scratch_device_ptr(dev);
    if (dev != c.device || !base) {
        set_error(MAGACE_ERR_STATE, "magace_compute_device_noalloc: the current device is not the one the "
                                    "potential was loaded on");
        return MAGACE_ERR_STATE;
    }
    const size_t zb = (size_t)c.model.z.nz * (c.precision == MAGACE_FP64 ? 16 : 8);
    Buffers B = c.buf;  // err flag of the context (allocated at set-up)
    B.Wd = nullptr;
    const int nd = c.model.pot.h.ndensity;
    const size_t cap = carve_scratch((char*)base, scratch_bytes(), zb, nd, partials_bytes(c.model.terms.nslot, nd), &B);
    if ((size_t)natom > cap) {
        set_error(MAGACE_ERR_INPUT, "magace_compute_device_noalloc: " + std::to_string(natom) +
                                        " atoms exceed the fixed scratch (" + std::to_string(scratch_bytes() >> 20) +
                                        " MiB: at most " + std::to_string(cap) +
                                        " atoms for this potential and precision)");
        return MAGACE_ERR_INPUT;
    }
    if (natom == 0) return 0;
    return run_device(c, natom, offsets, dlist, elist, dlist_momi, elist_momi, dlist_mom, elist_mom, eout, flist,
                      tlist, tlist0, (hipStream_t)stream, true, true, &B);
}

int magace_check_errors(void* stream) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    if (int rc = upload(*g_ctx)) return rc;
    return check_kernel_errors(*g_ctx, (hipStream_t)stream);
}

void acelib_(int* natom, int* nall, int* nei, int* nlist, int* occ, int* occlist, int* nstart,
             int* nstop, double* eout, double* dlist, double* elist, double* flist, double* dlist_momi,
             double* elist_momi, double* dlist_mom, double* elist_mom, double* tlist, double* tlist0) {
    (void)nlist;  // read but unused by the arithmetic (src/forces.F90:424,754)
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "acelib_ called before acelibini_");
        return;
    }
    Context& c = *g_ctx;
    const int N = *natom, P = *nei;
    if (N < 0 || P < 0 || *nall < N) {
        set_error(MAGACE_ERR_INPUT, "Incorrect array dimension specifications.");
        return;
    }
    for (int i = 0; i < N; ++i)
        if (occ[i] != 1) {
            set_error(MAGACE_ERR_INPUT, "Cannot handle multiple species.");
            return;
        }
    {
        // occlist is as long as the pair arrays: checked by the copy threads
        HostPool& pool = HostPool::get();
        const int nt = std::min(4 * pool.size(), P / 65536 + 1);
        std::atomic<bool> other{false};
        pool.run(nt, [&](int t) {
            const long lo = (long)P * t / nt, hi = (long)P * (t + 1) / nt;
            for (long p = lo; p < hi; ++p)
                if (occlist[p] != 1) {
                    other = true;
                    return;
                }
        });
        if (other) {
            set_error(MAGACE_ERR_INPUT, "Cannot handle multiple species.");
            return;
        }
    }
    // pair ranges must tile 1..nei in atom order, as pair_spin_ace builds them
    // (lammps/pair_spin_ace.cpp:457,524)
    c.h_off.resize(N + 1);
    int expect = 1;
    for (int i = 0; i < N; ++i) {
        if (nstart[i] != expect || nstop[i] < nstart[i] - 1 || nstop[i] > P) {
            set_error(MAGACE_ERR_INPUT, "pair ranges must be contiguous and ordered (nstart/nstop)");
            return;
        }
        c.h_off[i] = nstart[i] - 1;
        expect = nstop[i] + 1;
    }
    c.h_off[N] = expect - 1;
    if (c.h_off[N] != P) {
        set_error(MAGACE_ERR_INPUT, "Something wrong with neighbor list translation: neicount");
        return;
    }
    auto fail = [&](hipError_t e, const char* what) {
        if (e == hipSuccess) return false;
        set_error(MAGACE_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e) + " (" + what + ")");
        return true;
    };
    if (upload(c)) return;
    if (!c.s_in) {
        if (fail(hipStreamCreateWithFlags(&c.s_in, hipStreamNonBlocking), "stream") ||
            fail(hipStreamCreateWithFlags(&c.s_out, hipStreamNonBlocking), "stream"))
            return;
        for (auto& sl : c.slots)
            if (fail(hipEventCreateWithFlags(&sl.ev_in, hipEventDisableTiming), "event") ||
                fail(hipEventCreateWithFlags(&sl.ev_comp, hipEventDisableTiming), "event") ||
                fail(hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming), "event"))
                return;
    }
    const std::vector<int> cut = chunk_plan(c.h_off, chunk_pairs());
    const size_t nchunk = cut.size() - 1;
    size_t need_in = 0, need_out = 0;
    for (size_t k = 0; k < nchunk; ++k) {
        const ChunkLayout L(cut[k + 1] - cut[k], c.h_off[cut[k + 1]] - c.h_off[cut[k]]);
        need_in = std::max(need_in, L.in_bytes);
        need_out = std::max(need_out, L.out_bytes);
    }
    for (auto& sl : c.slots) {
        if (need_in > sl.cap_in) {
            dfree(sl.d_in);
            if (sl.h_in) (void)hipHostFree(sl.h_in);
            sl.h_in = nullptr;
            sl.cap_in = 0;
            if (fail(hipHostMalloc((void**)&sl.h_in, need_in, hipHostMallocDefault), "pinned alloc") ||
                fail(hipMalloc((void**)&sl.d_in, need_in), "alloc"))
                return;
            sl.cap_in = need_in;
        }
        if (need_out > sl.cap_out) {
            dfree(sl.d_out);
            if (sl.h_out) (void)hipHostFree(sl.h_out);
            sl.h_out = nullptr;
            sl.cap_out = 0;
            if (fail(hipHostMalloc((void**)&sl.h_out, need_out, hipHostMallocDefault), "pinned alloc") ||
                fail(hipMalloc((void**)&sl.d_out, need_out), "alloc"))
                return;
            sl.cap_out = need_out;
        }
    }
    hipStream_t s = c.stream;
    HostPool& pool = HostPool::get();
    if (fail(clear_flags(c.buf.err, s), "clear")) return;
    // chunk k's outputs, once its kernels are done, into the caller's arrays
    auto unpack = [&](size_t k) {
        const Context::Slot& sl = c.slots[k & 1];
        if (fail(hipEventSynchronize(sl.ev_done), "sync")) return false;
        const size_t a0 = cut[k], p0 = c.h_off[a0];
        const ChunkLayout L(cut[k + 1] - a0, c.h_off[cut[k + 1]] - p0);
        const char* o = sl.h_out;
        parallel_copy(pool, {{eout + a0, o + L.o_eout, 8 * L.na},
                             {tlist0 + 3 * a0, o + L.o_t0, 24 * L.na},
                             {flist + 3 * p0, o + L.o_fl, 24 * L.np},
                             {tlist + 3 * p0, o + L.o_tl, 24 * L.np}});
        return true;
    };
    // Order of the host's work: chunk k's kernels and copy-back are enqueued, then chunk k + 1
    // is packed and its upload enqueued, and only then are chunk k - 1's outputs unpacked, so
    // that the packing of the next chunk does not wait for the previous one's kernels.  Slot
    // k & 1 is reused by chunk k + 2; the device waits guard its buffers: the upload of k + 2
    // waits for chunk k's kernels (d_in), its kernels for chunk k's copy-back (d_out); its
    // copy-back is enqueued after chunk k is unpacked (h_out), its packing after chunk k's
    // upload has finished (h_in).
    auto pack_upload = [&](size_t k) {
        Context::Slot& sl = c.slots[k & 1];
        if (k >= 2 && fail(hipEventSynchronize(sl.ev_in), "sync")) return false;  // h_in of chunk k - 2 sent
        const size_t a0 = cut[k], p0 = c.h_off[a0];
        const ChunkLayout L(cut[k + 1] - a0, c.h_off[cut[k + 1]] - p0);
        char* in = sl.h_in;
        int* off = (int*)(in + L.i_off);
        for (size_t i = 0; i <= L.na; ++i) off[i] = c.h_off[a0 + i] - (int)p0;
        parallel_copy(pool, {{in + L.i_dl, dlist + p0, 8 * L.np},
                             {in + L.i_el, elist + 3 * p0, 24 * L.np},
                             {in + L.i_dm, dlist_mom + p0, 8 * L.np},
                             {in + L.i_em, elist_mom + 3 * p0, 24 * L.np},
                             {in + L.i_dmi, dlist_momi + a0, 8 * L.na},
                             {in + L.i_emi, elist_momi + 3 * a0, 24 * L.na}});
        if (k >= 2 && fail(hipStreamWaitEvent(c.s_in, sl.ev_comp, 0), "event")) return false;  // d_in free
        return !(fail(hipMemcpyAsync(sl.d_in, in, L.in_bytes, hipMemcpyHostToDevice, c.s_in), "copy") ||
                 fail(hipEventRecord(sl.ev_in, c.s_in), "event"));
    };
    if (nchunk > 0 && !pack_upload(0)) return;
    for (size_t k = 0; k < nchunk; ++k) {
        Context::Slot& sl = c.slots[k & 1];
        const size_t a0 = cut[k], p0 = c.h_off[a0];
        const ChunkLayout L(cut[k + 1] - a0, c.h_off[cut[k + 1]] - p0);
        if (fail(hipStreamWaitEvent(s, sl.ev_in, 0), "event")) return;
        if (k >= 2 && fail(hipStreamWaitEvent(s, sl.ev_done, 0), "event")) return;  // d_out of chunk k - 2 copied
        const char* d = sl.d_in;
        char* o = sl.d_out;
        if (run_device(c, (int)L.na, (const int*)(d + L.i_off), (const double*)(d + L.i_dl),
                       (const double*)(d + L.i_el), (const double*)(d + L.i_dmi), (const double*)(d + L.i_emi),
                       (const double*)(d + L.i_dm), (const double*)(d + L.i_em), (double*)(o + L.o_eout),
                       (double*)(o + L.o_fl), (double*)(o + L.o_tl), (double*)(o + L.o_t0), s, false, false, nullptr,
                       cut[k]))
            return;
        if (fail(hipEventRecord(sl.ev_comp, s), "event") || fail(hipStreamWaitEvent(c.s_out, sl.ev_comp, 0), "event") ||
            fail(hipMemcpyAsync(sl.h_out, sl.d_out, L.out_bytes, hipMemcpyDeviceToHost, c.s_out), "copy") ||
            fail(hipEventRecord(sl.ev_done, c.s_out), "event"))
            return;
        if (k + 1 < nchunk && !pack_upload(k + 1)) return;
        if (k > 0 && !unpack(k - 1)) return;
    }
    if (nchunk > 0 && !unpack(nchunk - 1)) return;
    if (check_kernel_errors(c, s)) return;
}

int magace_profile(int enable) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    g_ctx->profiling = enable != 0;
    g_ctx->events_used = 0;
    return 0;
}

int magace_profile_read(double* ms, int* count) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    Context& c = *g_ctx;
    ms[0] = ms[1] = ms[2] = 0.0;
    for (size_t i = 0; i < c.events_used; ++i) {
        HIPCHK(hipEventSynchronize(c.events[i][3]));
        for (int k = 0; k < 3; ++k) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, c.events[i][k], c.events[i][k + 1]));
            ms[k] += t;
        }
    }
    *count = (int)c.events_used;
    c.events_used = 0;
    return 0;
}

int magace_basis_counts(int* nvar, int n) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    const Basis& B = g_ctx->model.pot.basis;
    const int v[6] = {(int)B.b1.size(), (int)B.b2.size(), (int)B.b3.size(), (int)B.b4.size(), (int)B.b5.size(),
                      (int)B.b6.size()};
    for (int i = 0; i < n && i < 6; ++i) nvar[i] = v[i];
    return 0;
}

int magace_enumerate(const int* hdr, int order, int* out, int cap) {
    clear_error();
    Header h;
    int* f[20] = {&h.nb_a, &h.nb_m, &h.nr2_m, &h.l2m, &h.nr3_a, &h.nr3_m, &h.l3a, &h.l3m,
                  &h.nr4_a, &h.nr4_m, &h.l4a, &h.l4m, &h.nr5_a, &h.nr5_m, &h.l5a, &h.l5m,
                  &h.nr6_a, &h.nr6_m, &h.l6a, &h.l6m};
    for (int i = 0; i < 20; ++i) *f[i] = hdr[i];
    std::string err;
    if (!check_header(h, err)) {
        set_error(MAGACE_ERR_INPUT, err);
        return -MAGACE_ERR_INPUT;
    }
    const Basis B = enumerate_basis(h);
    if (!B.undefined.empty()) {
        set_error(MAGACE_ERR_INPUT, B.undefined);
        return -MAGACE_ERR_INPUT;
    }
    auto copy = [&](const auto& v, int width) {
        const int n = (int)v.size();
        if (out && cap >= n * width)
            for (int k = 0; k < n; ++k)
                for (int c = 0; c < width; ++c) out[k * width + c] = v[k][c];
        return n;
    };
    switch (order) {
        case 1: return copy(B.b1, 1);
        case 2: return copy(B.b2, 4);
        case 3: return copy(B.b3, 9);
        case 4: return copy(B.b4, 15);
        case 5: return copy(B.b5, 21);
        case 6: return copy(B.b6, 27);
    }
    set_error(MAGACE_ERR_INPUT, "order must be 1..6");
    return -MAGACE_ERR_INPUT;
}

int magace_model_terms(uint16_t* idx, double* w, int cap) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return -MAGACE_ERR_STATE;
    }
    const TermList& T = g_ctx->model.terms;
    const int n = (int)T.nterms();  // weights of density 1
    if (idx && w && cap >= n) {
        std::memcpy(idx, T.idx.data(), (size_t)n * 4 * sizeof(uint16_t));
        std::memcpy(w, T.w.data(), (size_t)n * sizeof(double));
    }
    return n;
}

int magace_model_records(int* chunks, int cap_chunks, uint16_t* idx, double* w, int cap, int* nchunks) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return -MAGACE_ERR_STATE;
    }
    const TermList& T = g_ctx->model.terms;
    const int n = (int)T.ncontrib(), nc = (int)(T.chunks.size() / 2);
    if (nchunks) *nchunks = nc;
    if (chunks && idx && w && cap >= n && cap_chunks >= nc) {
        std::memcpy(chunks, T.chunks.data(), (size_t)nc * 2 * sizeof(int32_t));
        std::memcpy(idx, T.cidx.data(), (size_t)n * 4 * sizeof(uint16_t));
        std::memcpy(w, T.cw.data(), (size_t)n * sizeof(double));
    }
    return n;
}

int magace_model_products(int* pairs, int cap) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return -MAGACE_ERR_STATE;
    }
    const TermList& T = g_ctx->model.terms;
    if (pairs && cap >= T.ny && T.ny) std::memcpy(pairs, T.ypair.data(), (size_t)T.ny * 2 * sizeof(int32_t));
    return T.ny;
}

int magace_fit_ncoef(void) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return -MAGACE_ERR_STATE;
    }
    const Potential& p = g_ctx->model.pot;
    return (int)(p.c1.size() + p.c2.size() + p.c3.size() + p.c4.size() + p.c5.size() + p.c6.size());
}

int magace_fit_get_coefficients(double* c, int n) {
    clear_error();
    const int nc = magace_fit_ncoef();
    if (nc < 0) return -nc;
    if (!c || n != nc) {
        set_error(MAGACE_ERR_INPUT, "coefficient count mismatch (magace_fit_ncoef)");
        return MAGACE_ERR_INPUT;
    }
    const Potential& p = g_ctx->model.pot;
    size_t o = 0;
    for (const auto* v : {&p.c1, &p.c2, &p.c3, &p.c4, &p.c5, &p.c6})
        for (double x : *v) c[o++] = x;
    return 0;
}

int magace_fit_set_coefficients(const double* c, int n) {
    clear_error();
    const int nc = magace_fit_ncoef();
    if (nc < 0) return -nc;
    if (!c || n != nc) {
        set_error(MAGACE_ERR_INPUT, "coefficient count mismatch (magace_fit_ncoef)");
        return MAGACE_ERR_INPUT;
    }
    Potential& p = g_ctx->model.pot;
    size_t o = 0;
    for (auto* v : {&p.c1, &p.c2, &p.c3, &p.c4, &p.c5, &p.c6})
        for (double& x : *v) x = c[o++];
    g_ctx->fit_natom = -1;  // Z and W of the last forward no longer match
    return refresh_weights(*g_ctx);
}

int magace_write_potential(const char* path) {
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    std::string err;
    if (!path || !write_potential(path, g_ctx->model.pot, err)) {
        set_error(MAGACE_ERR_INPUT, path ? err : "null path");
        return MAGACE_ERR_INPUT;
    }
    return 0;
}

int magace_fit_forward(int natom, int nei, const int* offsets, const int* nlist, const double* dlist,
                       const double* elist, const double* dlist_momi, const double* elist_momi,
                       const double* dlist_mom, const double* elist_mom, const int* atom_struct, int nstruc,
                       double* estruct, double* forces, void* stream) {
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    Context& c = *g_ctx;
    if (natom < 0 || nei < 0 || nstruc < 0 || ((nstruc > 0 || estruct) && !atom_struct) || !forces) {
        set_error(MAGACE_ERR_INPUT, "magace_fit_forward: invalid sizes");
        return MAGACE_ERR_INPUT;
    }
    if (int rc = upload(c)) return rc;
    hipStream_t s = (hipStream_t)stream;  // NULL: the default (null) stream
    if ((size_t)nei > c.fit_cap_pairs) {
        dfree(c.f_fl); dfree(c.f_tl);
        const size_t cap = (size_t)nei + nei / 4 + 64;
        HIPCHK(hipMalloc((void**)&c.f_fl, 3 * cap * sizeof(double)));
        HIPCHK(hipMalloc((void**)&c.f_tl, 3 * cap * sizeof(double)));
        c.fit_cap_pairs = cap;
    }
    if ((size_t)natom > c.fit_cap_atoms) {
        dfree(c.f_t0); dfree(c.f_eout);
        const size_t cap = (size_t)natom + natom / 4 + 16;
        HIPCHK(hipMalloc((void**)&c.f_t0, 3 * cap * sizeof(double)));
        HIPCHK(hipMalloc((void**)&c.f_eout, cap * sizeof(double)));
        c.fit_cap_atoms = cap;
    }
    const int nd = c.model.pot.h.ndensity;
    if (nd == 2 && (size_t)natom > c.fit_cap_wd) {
        // with two densities the adjoint also writes each density's gradient dRho_d/dZ, which
        // the backward needs for D_u rho_d
        dfree(c.fit.Wd);
        const size_t cap = (size_t)natom + natom / 4 + 16;
        HIPCHK(hipMalloc(&c.fit.Wd, cap * 2 * c.model.z.nz * 2 * sizeof(double)));
        c.fit_cap_wd = cap;
    }
    const int saved = c.precision;
    c.precision = MAGACE_FP64;  // the adjoint runs on the FP64 Z and W of this evaluation
    c.buf.Wd = nd == 2 ? c.fit.Wd : nullptr;
    const int rc = run_device(c, natom, offsets, dlist, elist, dlist_momi, elist_momi, dlist_mom, elist_mom,
                              c.f_eout, c.f_fl, c.f_tl, c.f_t0, s);
    c.buf.Wd = nullptr;
    c.precision = saved;
    if (rc) return rc;
    HIPCHK(fit_assemble(natom, offsets, nlist, c.f_fl, c.f_eout, atom_struct, nstruc, forces, estruct, s));
    c.fit_natom = natom;
    return 0;
}

int magace_fit_backward(int natom, int nei, const int* offsets, const int* nlist, const double* dlist,
                        const double* elist, const double* dlist_momi, const double* dlist_mom,
                        const double* elist_mom, const int* atom_struct, const double* gE, const double* gF,
                        double* grad, void* stream) {
    clear_error();
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    Context& c = *g_ctx;
    if (natom != c.fit_natom) {
        set_error(MAGACE_ERR_STATE, "magace_fit_backward must follow magace_fit_forward on the same atoms "
                                    "and coefficients");
        return MAGACE_ERR_STATE;
    }
    if (nei < 0 || !gF || !grad || (gE && !atom_struct)) {
        set_error(MAGACE_ERR_INPUT, "magace_fit_backward: invalid arguments");
        return MAGACE_ERR_INPUT;
    }
    hipStream_t s = (hipStream_t)stream;  // NULL: the default (null) stream
    const TermList& T = c.model.terms;
    FitBuffers& F = c.fit;
    const size_t nd = (size_t)c.model.pot.h.ndensity;
    if (!F.idx) {
        HIPCHK(hipMalloc((void**)&F.idx, T.idx.size() * sizeof(uint16_t) + 64));
        HIPCHK(hipMemcpy(F.idx, T.idx.data(), T.idx.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc((void**)&F.ub_ptr, T.ub_ptr.size() * sizeof(int32_t) + 64));
        HIPCHK(hipMalloc((void**)&F.ub_t, T.ub_t.size() * sizeof(int32_t) + 64));
        HIPCHK(hipMalloc((void**)&F.ub_u, T.ub_u.size() * sizeof(double) + 64));
        HIPCHK(hipMemcpy(F.ub_ptr, T.ub_ptr.data(), T.ub_ptr.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(F.ub_t, T.ub_t.data(), T.ub_t.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(F.ub_u, T.ub_u.data(), T.ub_u.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMalloc((void**)&F.M, 2 * T.nterms() * nd * sizeof(double) + 64));
    }
    if ((size_t)natom > F.cap_atoms) {
        dfree(F.Zdot); dfree(F.ehc_dot); dfree(F.alpha); dfree(F.beta);
        const size_t cap = (size_t)natom + natom / 4 + 16;
        HIPCHK(hipMalloc(&F.Zdot, cap * c.model.z.nz * 2 * sizeof(double)));
        HIPCHK(hipMalloc((void**)&F.ehc_dot, cap * sizeof(double)));
        HIPCHK(hipMalloc((void**)&F.alpha, cap * nd * sizeof(double)));
        HIPCHK(hipMalloc((void**)&F.beta, cap * nd * sizeof(double)));
        F.cap_atoms = cap;
    }
    const Basis& bs = c.model.pot.basis;
    const int nbo[5] = {(int)bs.b2.size(), (int)bs.b3.size(), (int)bs.b4.size(), (int)bs.b5.size(),
                        (int)bs.b6.size()};
    Buffers B = c.buf;
    B.Wd = F.Wd;
    HIPCHK(fit_backward(c.P64, c.shape_class, natom, offsets, nlist, gF, dlist, elist, dlist_momi, dlist_mom,
                        elist_mom, gE ? atom_struct : nullptr, gE, B, F, (int)T.nterms(), nbo, grad, s));
    return 0;
}

int magace_model_info(long long* info, int n) {
    if (!g_ctx) {
        set_error(MAGACE_ERR_STATE, "library not initialised");
        return MAGACE_ERR_STATE;
    }
    const Model& M = g_ctx->model;
    long long v[13] = {(long long)M.z.nz, (long long)M.terms.nterms(), (long long)M.terms.raw, M.terms.n2,
                       M.terms.n3, M.terms.n4, g_ctx->shape_class, (long long)M.z.oA2,
                       (long long)M.terms.ncontrib(), (long long)(M.terms.chunks.size() / 2), M.terms.n5,
                       M.terms.n6, M.terms.ny};
    for (int i = 0; i < n && i < 13; ++i) info[i] = v[i];
    return 0;
}

}  // extern "C"
