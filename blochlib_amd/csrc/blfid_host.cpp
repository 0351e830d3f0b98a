// blfid_host.cpp -- planner and C ABI of the MI355X powder FID engine (include/blfid.h).
//
// The planner lowers a tensor-described SolidSys (src/QMspins/sys.h:59-62) to device tables:
//   * anisotropic terms: crystal-frame tensor C = R_PAS diag(PAS) R_PAS^T (csa.cc:276-284,
//     dip.cc:159-167) with the spin factor of their diagonal (Iz_i, T20 homo/hetero,
//     spin_ten.cc:211-236, 431-452) and flip-flop parts;
//   * isotropic diagonal (-iso Iz, csa.cc:296-300; J Iz Iz, j.cc:135-137) and strong-J
//     flip-flops (spin_ten.cc:455-465);
//   * the Hilbert-space blocks: connected components of the flip-flop graph over Zeeman
//     states.  H(t) never couples two blocks, so U, A, W and the gamma-COMPUTE sums are done
//     per block and per coupled block pair;
//   * the rotor direction at each step midpoint: tmid = (t2-t1)/2 + t1 with t1, t2 advanced
//     by tau exactly as compute::calcUFID (compute.h:257-274) and sys.cc:753-755.
// Operators rho and det are split into sparse block-pair parts; rosym is detected with the
// reference's exact comparison ro == 0.5*(det + det^H) (compute.h:75-83).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/blfid.h"
#include "blfid_dev.h"
#include "blfid_layout.h"

typedef std::complex<double> cplx;

extern "C" {
int blfid_dev_launch_exp(int prec, int M, const DevPlan *P, int Kb, int c0, const int *grp,
                         int ngrp, void *scratch, size_t slot, int nslots, void *bufU,
                         hipStream_t st);
int blfid_dev_launch_dyson(int prec, int M, const DevPlan *P, int Kb, const int *grp, int ngrp,
                           const void *bufU, void *bufA, hipStream_t st);
int blfid_dev_launch_expdyson(int M, const DevPlan *P, int Kb, int c0, const int *grp, int ngrp,
                              void *bufA, hipStream_t st);
int blfid_dev_launch_rho_prep(int prec, const DevPlan *P, const DevPlan *PP, int Kb,
                              const double2 *Aprep, const int *sblk, const int *sloc,
                              const int *nzi, const double *nzv, int nnz, void *rho_out,
                              hipStream_t st);
int blfid_dev_launch_eig(int prec, const DevPlan *P, int Kb, const int *grp, int ngrp,
                         const void *src, size_t src_stride, int mode, double2 *bufW,
                         void *bufWr, double *outv, int cls, double2 *gscr, size_t gslot,
                         int nslots, hipStream_t st);
int blfid_dev_launch_bmul(int prec, int M, const DevPlan *P, int Kb, const int *grp, int ngrp,
                          const void *bufA, const void *bufW, void *bufB, hipStream_t st);
int blfid_dev_launch_qr(int prec, int M, const DevPlan *P, int Kb, const int *pgrp, int ngrp,
                        const void *bufB, void *scratch, size_t slot, int nslots, void *Qt,
                        void *Rt, hipStream_t st);
int blfid_dev_launch_corr(int prec, const DevPlan *P, int Kb, const void *Qt, const void *Rt,
                          const double *thetas, const int *er, const int *es, double *part,
                          hipStream_t st);
int blfid_dev_launch_reduce(int Kb, int c0, int npts, int nparts, int G, const double *w,
                            double scale, const double *part, double *grp, hipStream_t st);
int blfid_dev_launch_finish(int npts, int ngroups, const double *grp, double nre, double nim,
                            double *out, hipStream_t st);
int blfid_dev_launch_zero(double *p, size_t n, hipStream_t st);
int blfid_dev_launch_prep(int K, const double *th, const double *ph, const double *ga,
                          double *RP, hipStream_t st);
int blfid_dev_launch_static_h(const DevPlan *P, int Kb, int c0, const int *grp, int ngrp,
                              double *scratch, size_t slot, int nslots, double2 *bufH,
                              hipStream_t st);
int blfid_dev_launch_state_project(long K, int N, const double2 *C, double inv_nrm, double2 *S,
                                   hipStream_t st);
int blfid_dev_launch_state_fill(long K, int N, const double2 *rho0, double2 *S, hipStream_t st);
int blfid_dev_launch_state_apply(const DevPlan *PP, int Kb, int c0, const double2 *Aprep,
                                 const int *sblk, const int *sloc, int N, int repeat, double2 *S,
                                 double2 *T, int nslots, hipStream_t st);
int blfid_dev_launch_state_trace(int K, int N, const double2 *S, const int *nzi,
                                 const double *nzv, int nnz, double *part, hipStream_t st);
int blfid_dev_launch_state_gather(int prec, const DevPlan *P, int Kb, int c0, const double2 *S,
                                  void *rho_out, hipStream_t st);
int blfid_dev_launch_static_fid(int M, const DevPlan *P, int Kb, const int *pgrp, int ngrp,
                                const double2 *bufW, const double *lam, double2 *scratch,
                                size_t slot, int nslots, double w0, int prune, double *part,
                                hipStream_t st);
}

static thread_local std::string g_err;

static blfid_status fail(blfid_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return fail(_e == hipErrorOutOfMemory ? BLFID_ERR_NOMEM : BLFID_ERR_DEVICE,       \
                        std::string(#expr) + ": " + hipGetErrorString(_e));                   \
    } while (0)

#define LAUNCHCHK(expr)                                                                       \
    do {                                                                                      \
        int _e = (expr);                                                                      \
        if (_e != 0)                                                                          \
            return fail(BLFID_ERR_DEVICE, std::string(#expr) + ": " +                         \
                                              hipGetErrorString((hipError_t)_e));             \
    } while (0)

// Process-wide launch tuning per device (blfid_tune_*): all zero (the defaults) at start.
#define TUNE_MAX_DEVICES 64
static std::mutex g_tune_mu;
static int g_tune[TUNE_MAX_DEVICES][BLFID_TUNE_LEN];

static int tune_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= TUNE_MAX_DEVICES) dev = 0;
    return dev;
}

static void tune_from_process(int *vals) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::memcpy(vals, g_tune[tune_device()], sizeof(int) * BLFID_TUNE_LEN);
}

static const double PI2 = 6.283185307179586476925286766559;
static const double DEG2RAD = 0.017453292519943295769236907684886;

namespace {

void rot3(double a, double b, double g, double R[9]) {
    double cg = cos(g), cb = cos(b), ca = cos(a), sg = sin(g), sb = sin(b), sa = sin(a);
    R[0] = ca * cb * cg - sa * sg; R[1] = sa * cb * cg + ca * sg; R[2] = -sb * cg;
    R[3] = -ca * cb * sg - sa * cg; R[4] = -sa * cb * sg + ca * cg; R[5] = sb * sg;
    R[6] = ca * sb; R[7] = sa * sb; R[8] = cb;
}

// C = R diag(p) R^T with R = rotationMatrix3D(Euler in degrees): 6 unique components
void crystal_tensor(double pxx, double pyy, double pzz, double a, double b, double g,
                    double C6[6]) {
    double R[9];
    rot3(a * DEG2RAD, b * DEG2RAD, g * DEG2RAD, R);
    double p[3] = {pxx, pyy, pzz};
    double C[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += R[3 * i + k] * p[k] * R[3 * j + k];
            C[3 * i + j] = s;
        }
    C6[0] = C[0]; C6[1] = C[4]; C6[2] = C[8];
    C6[3] = 0.5 * (C[1] + C[3]); C6[4] = 0.5 * (C[2] + C[6]); C6[5] = 0.5 * (C[5] + C[7]);
}

template <typename T> struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    hipError_t alloc(size_t cnt) {
        if (cnt <= n && p) return hipSuccess;
        release();
        n = cnt;
        if (cnt == 0) return hipSuccess;
        return hipMalloc((void **)&p, cnt * sizeof(T));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t upload(const std::vector<T> &v) {
        hipError_t e = alloc(v.size() ? v.size() : 1);
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
    }
};

struct Block {
    std::vector<int> states;
};

struct Pair {
    int I, J;
};

} // namespace

struct blfid_workspace {
    char *base = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;
    bool used = false;
};

// Per-crystallite density operators of one powder, resident in HBM (blfid_state_*): K dense
// N x N complex128 matrices, column-major, plus the powder's rotations and weights.
struct blfid_state {
    int S = 0, N = 0, K = 0;
    DevBuf<double> d_th, d_ph, d_ga, d_w, d_RP;
    DevBuf<double2> rho;   // K N^2
    DevBuf<double2> rho0;  // staging of blfid_state_set / blfid_state_load
    DevBuf<char> ws;       // propagators and product slots of blfid_state_apply
    DevBuf<int> sblk, sloc, nzi;
    DevBuf<double> nzv, part, grp, out;
    double budget = 0;
};

struct blfid_plan {
    // input
    int S = 0, N = 0;
    std::vector<int> iso;
    std::vector<double> csa, dip, jc;
    double wr = 0, rotor_deg = 0, sw = 0, t0 = 0;
    int npts = 0;
    int prec = 0;               // 0: complex128 storage, 1: complex64 storage (BLFID_C64)
    int bf16x3 = 0;             // complex64 products as three bf16 products (BLFID_C64_BF16X3)
    bool spinning = false;
    bool ptop = false;          // point-to-point pulse-program FID (blfid_plan_create_ptop)
    // rf of a ptop plan: slot per pulsed isotope, n x nrfiso x (amp/2, phase, offset)
    int nrfiso = 0;
    std::vector<int> spin_rf;
    std::vector<double> rf;
    std::vector<double> step_dt; // unequal step lengths (blfid_plan_create_pulse), else empty
    std::vector<double> step_t1; // explicit step start times, else empty (consecutive)
    // derived
    int n = 1;
    double sw_adj = 0, tau = 0;
    std::vector<double> terms, dconst, ffconst, u3;
    std::vector<int> ff_off, ff_term; // flip-flop terms of each ordered spin pair (CSR)
    std::vector<double> xd_cst, xd_fpr; // quadratic-form tables of the fused kernel's H build
    std::vector<Block> blocks;
    std::vector<int> blk_dim, blk_soff, blk_states;
    std::vector<long> blk_uoff;
    long Lsum = 0;
    int bmax = 0;
    std::map<int, std::vector<int>> blk_groups; // M -> blocks
    std::map<int, std::vector<int>> fuse_groups; // M -> blocks of <= DEV_LDS_BLOCK states
    std::vector<int> big_blocks;                 // blocks beyond DEV_LDS_BLOCK states
    // fused kernel: per block of <= DEV_LDS_BLOCK states, its flip-flop positions in the LDS plane
    // of its tile variant M, (r LDK + q) << 8 | (i S + j), LDK = 16 M + LDS_PAD (CSR over blocks)
    std::vector<int> xd_off, xd_list;
    std::map<int, std::vector<int>> eig_cls;    // eigen-solver size class -> blocks
    int exp_slots = 0, qr_slots = 0, eig_slots = 0;
    // operators
    bool have_ops = false;
    int rosym = 0;
    cplx trdd = 0;
    std::vector<Pair> pairs;
    std::vector<long> pair_eoff;
    std::vector<int> pair_doff, pair_roff, d_rowptr, d_col, r_rowptr, r_col, elem_r, elem_s;
    std::vector<double> d_val, r_val;
    long Etot = 0;
    long Estr = 0;             // Q row stride per step: Etot padded to 16 elements
    int nparts = 1;            // FID partial sums per crystallite
    bool rt_from_q = false;    // rosym spinning case: R formed from Q in k_corr
    bool rt_rev = false;       // rt_from_q and some element has a reverse element
    int corr_auto = 0;         // leading correlation chunks taking the autocorrelation kernel
    long corr_rev_elems = 0;   // elements with a reverse element (their R reads a second Q)
    std::vector<int> elem_rev;  // per element (r,s) of pair (I,J): element (s,r) of (J,I) or -1
    std::map<int, std::vector<int>> pair_groups; // M -> pairs
    // powder
    int K = 0;
    bool rp_dirty = false;      // cartPowder of the new powder still to be formed (in run)
    // device
    DevBuf<double> d_terms, d_dconst, d_ffconst, d_u3, d_RP, d_th, d_ph, d_ga, d_w;
    DevBuf<int> d_ff_off, d_ff_term;
    DevBuf<double> d_xd_cst, d_xd_fpr;
    DevBuf<int> d_blk_dim, d_blk_soff, d_blk_states, d_pair_I, d_pair_J, d_pair_doff,
        d_pair_roff, d_drow, d_dcol, d_rrow, d_rcol, d_er, d_es, d_erev;
    DevBuf<long> d_blk_uoff, d_pair_eoff;
    DevBuf<double> d_dval, d_rval;
    std::map<int, DevBuf<int>> d_bgrp, d_pgrp, d_fgrp;
    DevBuf<int> d_bigb, d_xd_off, d_xd_list;
    std::map<int, DevBuf<int>> d_eig_cls;
    DevBuf<int> d_spin_rf;
    DevBuf<double> d_rf, d_thetak;
    // pulses applied to rho before the gamma-COMPUTE FID (blfid_plan_set_prep): a pulse plan for
    // the preparation propagator of each crystallite, its state -> (block, local index) maps, and
    // the nonzero entries (a, b, re, im) of rho, from which the prepared blocks are formed
    blfid_plan *prep = nullptr;
    double prep_rotor = 0;
    DevBuf<int> d_prep_sblk, d_prep_sloc, d_rho_nzi;
    DevBuf<double> d_rho_nzv;
    int rho_nnz = 0;
    // or: each crystallite's rho taken from a state (blfid_plan_set_state), not owned
    blfid_state *state = nullptr;
    char *bufRho = nullptr;
    // working storage (carved from the plan's own arena or a shared workspace)
    DevBuf<char> own_ws;
    blfid_workspace *ws = nullptr;
    // U, A, Qt, Rt, Wr and the exp / Q-R scratch are in the storage precision (char* = bytes)
    char *bufU = nullptr, *bufA = nullptr, *bufQt = nullptr, *bufRt = nullptr, *scrQ = nullptr,
         *scrE = nullptr, *bufWr = nullptr;
    double2 *bufW = nullptr;
    double *thetas = nullptr, *part = nullptr, *grp = nullptr;
    double2 *scrEig = nullptr;
    DevBuf<double> fid_out;
    DevBuf<double> d_expflops; // propagator flop counters while profiling
    double budget = 0;
    int Kb = 0;
    hipStream_t last_stream = nullptr;
    bool ran = false;
    DevPlan dp{};
    // launch choices (blfid_tune_*): taken from the device's process table at creation
    int tn[BLFID_TUNE_LEN] = {0};
    // profiling
    bool prof = false;
    struct Ev { int phase; hipEvent_t a, b; double flops, bytes; };
    std::vector<Ev> pending;
    double ph_ms[BLFID_NSLOT] = {0}, ph_flops[BLFID_NSLOT] = {0}, ph_bytes[BLFID_NSLOT] = {0};
    int ph_n[BLFID_NSLOT] = {0};
    // HIP graph of the whole run (blfid_plan_graph_capture): valid while `gen` is unchanged and
    // the working storage and FID buffer are the ones it was recorded with
    unsigned long gen = 0, g_gen = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipStream_t cap_stream = nullptr;
    char *g_base = nullptr;
    double *g_out = nullptr;
    // the run in crystallite slices (blfid_plan_run_slice): next slice start, batch size, plan
    // generation and working storage of the slices so far
    int sl_next = -1, sl_norm = 1;
    long sl_Kb = 0;
    unsigned long sl_gen = 0;
    char *sl_base = nullptr;
    ~blfid_plan() {
        for (auto &e : pending) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        delete prep;
    }
};

// Records HIP events around one launch on `st` when profiling is on, with the launch's
// executed algorithmic flops and its algorithmic HBM bytes (what the kernel must read and
// write at the least: its inputs once, its outputs once).
struct PhaseScope {
    blfid_plan *pl; int phase; double flops, bytes; hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    PhaseScope(blfid_plan *p, int ph, double fl, double by, hipStream_t s)
        : pl(p), phase(ph), flops(fl), bytes(by), st(s) {
        if (pl->prof) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, st);
        }
    }
    ~PhaseScope() {
        if (pl->prof) {
            (void)hipEventRecord(b, st);
            pl->pending.push_back({phase, a, b, flops, bytes});
        }
    }
};

static void collect_profile(blfid_plan *pl) {
    for (auto &e : pl->pending) {
        float ms = 0.f;
        (void)hipEventSynchronize(e.b);
        (void)hipEventElapsedTime(&ms, e.a, e.b);
        pl->ph_ms[e.phase] += ms;
        pl->ph_n[e.phase] += 1;
        pl->ph_flops[e.phase] += e.flops;
        pl->ph_bytes[e.phase] += e.bytes;
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    pl->pending.clear();
}

// ------------------------------------------------------------------------------ planning
// Eigen-solver size class of a block: 0..4 keep the block in LDS sized for b <= 16, 32, 48, 64,
// 80; 5 uses global scratch.
static int eig_class(int b) {
    if (b <= 16) return 0;
    if (b <= 32) return 1;
    if (b <= 48) return 2;
    if (b <= 64) return 3;
    if (b <= DEV_LDS_BLOCK) return 4;
    return 5;
}

static blfid_status build_tables(blfid_plan *pl) {
    const int S = pl->S, N = pl->N;
    pl->terms.clear();
    pl->dconst.assign(N, 0.0);
    pl->ffconst.assign(S * S, 0.0);
    auto mz = [&](int s, int i) { return ((s >> (S - 1 - i)) & 1) ? -0.5 : 0.5; };
    const double sq3 = std::pow(3., 0.5);
    // flip-flop adjacency for the block structure
    std::vector<std::vector<char>> ffpair(S, std::vector<char>(S, 0));
    for (size_t c = 0; c < pl->csa.size() / 7; c++) {
        const double *p = &pl->csa[7 * c];
        double iso = p[0], del = p[1], eta = p[2];
        int on = (int)p[3];
        if (iso != 0) {
            double isoterm = (1. / sq3) * (-sq3 * iso);
            for (int s = 0; s < N; s++) pl->dconst[s] += isoterm * mz(s, on);
        }
        if (iso != 0 && del == 0.0) continue;
        double t[DEV_TERM_STRIDE];
        crystal_tensor(del * (eta - 1.0) / 2.0, -del * (eta + 1.0) / 2.0, 1.0 * del, p[4], p[5],
                       p[6], t);
        t[6] = 1.0;
        t[7] = 0.0;
        t[8] = on;
        t[9] = -1;
        pl->terms.insert(pl->terms.end(), t, t + DEV_TERM_STRIDE);
    }
    for (size_t d = 0; d < pl->dip.size() / 6; d++) {
        const double *p = &pl->dip[6 * d];
        double del = p[0];
        int i = (int)p[1], j = (int)p[2];
        double s6 = std::pow(6.0, 0.5);
        double t[DEV_TERM_STRIDE];
        crystal_tensor(-0.5 * del * s6, -0.5 * del * s6, 1.0 * del * s6, p[3], p[4], p[5], t);
        bool homo = pl->iso[i] == pl->iso[j];
        if (homo) {
            double f = std::pow(1. / 6., 0.5);
            t[6] = f * 2.0;
            t[7] = f * (-0.5);
            if (del != 0.0) ffpair[i][j] = ffpair[j][i] = 1;
        } else {
            t[6] = std::pow(4. / 6., 0.5);
            t[7] = 0.0;
        }
        t[8] = i;
        t[9] = j;
        pl->terms.insert(pl->terms.end(), t, t + DEV_TERM_STRIDE);
    }
    for (size_t q = 0; q < pl->jc.size() / 4; q++) {
        const double *p = &pl->jc[4 * q];
        double iso = p[0];
        int i = (int)p[1], j = (int)p[2], strong = (int)p[3];
        double f = iso * sq3 * (1. / sq3);
        for (int s = 0; s < N; s++) pl->dconst[s] += f * mz(s, i) * mz(s, j);
        if (strong) {
            pl->ffconst[i * S + j] += 0.5 * f;
            pl->ffconst[j * S + i] += 0.5 * f;
            if (f != 0.0) ffpair[i][j] = ffpair[j][i] = 1;
        }
    }
    if ((int)(pl->terms.size() / DEV_TERM_STRIDE) > DEV_MAX_TERMS)
        return fail(BLFID_ERR_UNSUPPORTED, "too many interactions (limit 128 anisotropic terms)");
    {
        std::vector<std::vector<int>> lists((size_t)S * S);
        for (size_t t = 0; t < pl->terms.size() / DEV_TERM_STRIDE; t++) {
            const double *c = &pl->terms[DEV_TERM_STRIDE * t];
            const int i = (int)c[8], j = (int)c[9];
            if (c[7] != 0.0 && j >= 0) {
                lists[(size_t)i * S + j].push_back((int)t);
                lists[(size_t)j * S + i].push_back((int)t);
            }
        }
        pl->ff_off.assign((size_t)S * S + 1, 0);
        pl->ff_term.clear();
        for (size_t q = 0; q < lists.size(); q++) {
            pl->ff_term.insert(pl->ff_term.end(), lists[q].begin(), lists[q].end());
            pl->ff_off[q + 1] = (int)pl->ff_term.size();
        }
        // The same Hamiltonian as quadratic forms in v = R_P^T u for the fused kernel: with
        // w = (vx^2, vy^2, vz^2, 2 vx vy, 2 vx vz, 2 vy vz) every term's zz_t = C_t . w, so the
        // diagonal of basis state a is xd_cst[a] . (1, w) and the flip-flop coefficient of the
        // ordered pair (i, j) is xd_fpr[i S + j] . (1, w) (entry 0 holds the constant part).
        const int nt = (int)(pl->terms.size() / DEV_TERM_STRIDE);
        pl->xd_cst.assign((size_t)N * 7, 0.0);
        for (int a = 0; a < N; a++) {
            double *o = &pl->xd_cst[(size_t)a * 7];
            o[0] = pl->dconst[a];
            for (int t = 0; t < nt; t++) {
                const double *c = &pl->terms[DEV_TERM_STRIDE * t];
                const int i = (int)c[8], j = (int)c[9];
                double f = c[6] * mz(a, i);
                if (j >= 0) f *= mz(a, j);
                for (int q = 0; q < 6; q++) o[1 + q] += f * c[q];
            }
        }
        pl->xd_fpr.assign((size_t)S * S * 7, 0.0);
        for (size_t q = 0; q < lists.size(); q++) {
            double *o = &pl->xd_fpr[q * 7];
            o[0] = pl->ffconst[q];
            for (int t : lists[q]) {
                const double *c = &pl->terms[DEV_TERM_STRIDE * t];
                for (int k = 0; k < 6; k++) o[1 + k] += c[7] * c[k];
            }
        }
    }
    // blocks: union-find over states, edges = flip-flops of coupled pairs
    std::vector<int> par(N);
    std::iota(par.begin(), par.end(), 0);
    std::function<int(int)> find = [&](int x) {
        while (par[x] != x) x = par[x] = par[par[x]];
        return x;
    };
    // rf couples the two Zeeman states of every pulsed spin (single flips)
    for (int i = 0; i < S && pl->nrfiso > 0; i++) {
        int q = pl->spin_rf[i];
        if (q < 0) continue;
        bool on = false;
        for (int k = 0; k < pl->n && !on; k++) on = pl->rf[(size_t)3 * (k * pl->nrfiso + q)] != 0.0;
        if (!on) continue;
        int bi = 1 << (S - 1 - i);
        for (int s = 0; s < N; s++) {
            int a = find(s), b = find(s ^ bi);
            if (a != b) par[a] = b;
        }
    }
    for (int i = 0; i < S; i++)
        for (int j = i + 1; j < S; j++) {
            if (!ffpair[i][j]) continue;
            int bi = 1 << (S - 1 - i), bj = 1 << (S - 1 - j);
            for (int s = 0; s < N; s++) {
                if (((s & bi) != 0) != ((s & bj) != 0)) {
                    int a = find(s), b = find(s ^ bi ^ bj);
                    if (a != b) par[a] = b;
                }
            }
        }
    std::map<int, int> root2blk;
    std::vector<Block> comps;
    for (int s = 0; s < N; s++) {
        int r = find(s);
        auto it = root2blk.find(r);
        if (it == root2blk.end()) {
            root2blk[r] = (int)comps.size();
            comps.push_back(Block());
            it = root2blk.find(r);
        }
        comps[it->second].states.push_back(s);
    }
    // Components smaller than one 16-wide tile are packed (first fit, largest first) into
    // shared blocks of at most 16 states: H stays block diagonal inside such a block, so
    // exp, products and the eigenbasis are unchanged, and tiny components stop costing a
    // workgroup each.
    pl->blocks.clear();
    std::vector<int> order(comps.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return comps[a].states.size() > comps[b].states.size();
    });
    std::vector<int> binfill;
    std::vector<int> small_bins;
    for (int ci : order) {
        const auto &st = comps[ci].states;
        if (st.size() >= 16) {
            pl->blocks.push_back(comps[ci]);
            continue;
        }
        int placed = -1;
        for (size_t q = 0; q < small_bins.size(); q++)
            if (binfill[q] + (int)st.size() <= 16) { placed = (int)q; break; }
        if (placed < 0) {
            small_bins.push_back((int)pl->blocks.size());
            binfill.push_back(0);
            pl->blocks.push_back(Block());
            placed = (int)small_bins.size() - 1;
        }
        auto &dst = pl->blocks[small_bins[placed]].states;
        dst.insert(dst.end(), st.begin(), st.end());
        binfill[placed] += (int)st.size();
    }
    pl->blk_dim.clear(); pl->blk_soff.clear(); pl->blk_uoff.clear(); pl->blk_states.clear();
    pl->blk_groups.clear();
    pl->fuse_groups.clear();
    pl->big_blocks.clear();
    long uo = 0;
    int so = 0;
    pl->bmax = 0;
    for (size_t b = 0; b < pl->blocks.size(); b++) {
        int d = (int)pl->blocks[b].states.size();
        pl->blk_dim.push_back(d);
        pl->blk_soff.push_back(so);
        pl->blk_uoff.push_back(uo);
        for (int s : pl->blocks[b].states) pl->blk_states.push_back(s);
        so += d;
        uo += (long)d * d;
        pl->bmax = std::max(pl->bmax, d);
        int M = std::min(5, (d + 15) / 16);
        pl->blk_groups[M].push_back((int)b);
        if (d <= DEV_LDS_BLOCK) pl->fuse_groups[M].push_back((int)b);
        else pl->big_blocks.push_back((int)b);
    }
    pl->Lsum = uo;
    // flip-flop positions of the fused kernel's H planes
    pl->xd_off.assign(1, 0);
    pl->xd_list.clear();
    for (size_t b = 0; b < pl->blocks.size(); b++) {
        const auto &st = pl->blocks[b].states;
        const int d = (int)st.size();
        if (d <= DEV_LDS_BLOCK) {
            const int LDK = 16 * ((d + 15) / 16) + LDS_PAD;
            for (int r = 0; r < d; r++)
                for (int q = 0; q < d; q++) {
                    const int x = st[r] ^ st[q];
                    if (__builtin_popcount((unsigned)x) != 2 || (st[r] & x) == 0 || (st[r] & x) == x)
                        continue;
                    const int p1 = 31 - __builtin_clz((unsigned)x), p0 = __builtin_ctz((unsigned)x);
                    const int i = S - 1 - p1, j = S - 1 - p0;
                    if (!ffpair[i][j]) continue;
                    pl->xd_list.push_back(((r * LDK + q) << 8) | (i * S + j));
                }
        }
        pl->xd_off.push_back((int)pl->xd_list.size());
    }
    pl->eig_cls.clear();
    for (size_t b = 0; b < pl->blocks.size(); b++)
        pl->eig_cls[eig_class(pl->blk_dim[b])].push_back((int)b);
    {
        // concurrent workgroup slots of scratch, sized so each scratch region stays ~2.5 GB
        const double bm2 = (double)pl->bmax * pl->bmax;
        auto clampi = [](double v, int lo, int hi) { return (int)std::max<double>(lo, std::min<double>(hi, v)); };
        pl->exp_slots = clampi(std::floor(2.5e9 / (9.0 * bm2 * 8.0)), 64, 16384);
        pl->qr_slots = clampi(std::floor(2.5e9 / (3.0 * bm2 * 16.0)), 64, 16384);
        pl->eig_slots = pl->eig_cls.count(5) ? clampi(std::floor(1.0e9 / (bm2 * 16.0)), 64, 2048) : 0;
    }
    if (pl->bmax > DEV_MAX_BLOCK)
        return fail(BLFID_ERR_UNSUPPORTED,
                    "largest Hilbert-space block " + std::to_string(pl->bmax) +
                        " exceeds the engine's limit of " + std::to_string(DEV_MAX_BLOCK));
    // rotor direction (row 2 of cartSpin) at each step midpoint
    pl->u3.clear();
    if (pl->spinning || pl->ptop) {
        double br = pl->rotor_deg * DEG2RAD;
        const bool uneq = !pl->step_dt.empty();
        double t1 = (uneq && !pl->step_t1.empty()) ? pl->step_t1[0] : pl->t0;
        double t2 = t1 + (uneq ? pl->step_dt[0] : pl->tau);
        for (int k = 0; k < pl->n; k++) {
            double tmid = (t2 - t1) / 2.0 + t1;
            double al = tmid * pl->wr * PI2;
            pl->u3.push_back(std::sin(al) * std::sin(br));
            pl->u3.push_back(-std::cos(al) * std::sin(br));
            pl->u3.push_back(std::cos(br));
            if (uneq && k + 1 < pl->n) {
                // the next step starts at its own time, or where this one ends
                t1 = pl->step_t1.empty() ? t2 : pl->step_t1[k + 1];
                t2 = t1 + pl->step_dt[k + 1];
            } else {
                t1 += pl->tau;
                t2 += pl->tau;
            }
        }
    } else {
        // static: Hamiltonian(0, dt) with wr = 0 -> rotor phase 0 (oneFID.h:297-299)
        double br = pl->rotor_deg * DEG2RAD;
        if (br == 0.0) {
            pl->u3 = {0.0, 0.0, 1.0};
        } else {
            pl->u3 = {0.0, -std::sin(br), std::cos(br)};
        }
    }
    return BLFID_OK;
}

static blfid_status upload_tables(blfid_plan *pl) {
    HIPCHK(pl->d_terms.upload(pl->terms));
    HIPCHK(pl->d_dconst.upload(pl->dconst));
    HIPCHK(pl->d_ffconst.upload(pl->ffconst));
    HIPCHK(pl->d_ff_off.upload(pl->ff_off));
    HIPCHK(pl->d_ff_term.upload(pl->ff_term));
    HIPCHK(pl->d_xd_cst.upload(pl->xd_cst));
    HIPCHK(pl->d_xd_fpr.upload(pl->xd_fpr));
    HIPCHK(pl->d_u3.upload(pl->u3));
    HIPCHK(pl->d_blk_dim.upload(pl->blk_dim));
    HIPCHK(pl->d_blk_soff.upload(pl->blk_soff));
    HIPCHK(pl->d_blk_states.upload(pl->blk_states));
    HIPCHK(pl->d_blk_uoff.upload(pl->blk_uoff));
    for (auto &kv : pl->blk_groups) HIPCHK(pl->d_bgrp[kv.first].upload(kv.second));
    for (auto &kv : pl->fuse_groups) HIPCHK(pl->d_fgrp[kv.first].upload(kv.second));
    HIPCHK(pl->d_bigb.upload(pl->big_blocks));
    HIPCHK(pl->d_xd_off.upload(pl->xd_off));
    HIPCHK(pl->d_xd_list.upload(pl->xd_list));
    for (auto &kv : pl->eig_cls) HIPCHK(pl->d_eig_cls[kv.first].upload(kv.second));
    HIPCHK(pl->d_spin_rf.upload(pl->spin_rf));
    HIPCHK(pl->d_rf.upload(pl->rf));
    if (!pl->step_dt.empty()) {
        std::vector<double> th(pl->step_dt.size());
        for (size_t k = 0; k < th.size(); k++) th[k] = PI2 * pl->step_dt[k];
        HIPCHK(pl->d_thetak.upload(th));
    }
    return BLFID_OK;
}

static void refresh_devplan(blfid_plan *pl) {
    DevPlan &d = pl->dp;
    d.S = pl->S;
    d.N = pl->N;
    d.n = pl->n;
    d.npts = pl->npts;
    d.nterms = (int)(pl->terms.size() / DEV_TERM_STRIDE);
    d.nblocks = (int)pl->blocks.size();
    d.npairs = (int)pl->pairs.size();
    d.Lsum = pl->Lsum;
    d.Etot = pl->Etot;
    d.rt_from_q = pl->rt_from_q ? 1 : 0;
    d.elem_rev = pl->d_erev.p;
    d.expflops = pl->prof ? pl->d_expflops.p : nullptr;
    d.bf16x3 = pl->bf16x3;
    d.Estr = pl->Estr;
    d.nparts = pl->nparts;
    d.rt_rev = pl->rt_rev ? 1 : 0;
    d.corr_auto_chunks = pl->corr_auto;
    d.theta = PI2 * pl->tau;
    d.terms = pl->d_terms.p;
    d.ff_off = pl->d_ff_off.p;
    d.ff_term = pl->d_ff_term.p;
    d.xd_cst = pl->d_xd_cst.p;
    d.xd_fpr = pl->d_xd_fpr.p;
    d.dconst = pl->d_dconst.p;
    d.ffconst = pl->d_ffconst.p;
    d.u3 = pl->d_u3.p;
    d.RP = pl->d_RP.p;
    d.blk_dim = pl->d_blk_dim.p;
    d.blk_soff = pl->d_blk_soff.p;
    d.blk_uoff = pl->d_blk_uoff.p;
    d.blk_states = pl->d_blk_states.p;
    d.pair_I = pl->d_pair_I.p;
    d.pair_J = pl->d_pair_J.p;
    d.pair_eoff = pl->d_pair_eoff.p;
    d.pair_doff = pl->d_pair_doff.p;
    d.pair_roff = pl->d_pair_roff.p;
    d.d_rowptr = pl->d_drow.p;
    d.d_col = pl->d_dcol.p;
    d.d_val = pl->d_dval.p;
    d.r_rowptr = pl->d_rrow.p;
    d.r_col = pl->d_rcol.p;
    d.r_val = pl->d_rval.p;
    d.usign = pl->ptop ? 1 : -1;
    d.akeep = pl->ptop ? 2 : pl->n;
    d.nrfiso = pl->nrfiso;
    d.spin_rf = pl->d_spin_rf.p;
    d.rf = pl->d_rf.p;
    d.thetak = pl->step_dt.empty() ? nullptr : pl->d_thetak.p;
    d.xd_off = pl->d_xd_off.p;
    d.xd_list = pl->d_xd_list.p;
    d.rho_c = pl->bufRho;
    d.elem_r = pl->d_er.p;
    d.elem_s = pl->d_es.p;
    for (int f = 0; f < 9; f++) d.tn_cap[f] = pl->tn[6 + f];
    d.tn_corr = pl->tn[5];
    d.tn_qr = pl->tn[15];
}

// ------------------------------------------------------------------------------ C ABI
extern "C" {

const char *blfid_last_error(void) { return g_err.c_str(); }
const char *blfid_version(void) { return "blfid-mi355x 0.1 (gfx950, FP64)"; }

}

// Argument checks shared by both plan kinds (the BLEXCEPTION cases of oneFID.h:78-83,
// csa.cc:413-418 and the spin-index checks of the interaction readers).
static blfid_status check_system(const blfid_system *sys, int npts, blfid_precision prec,
                                 blfid_plan **out) {
    if (!out) return fail(BLFID_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!sys) return fail(BLFID_ERR_INVALID, "sys is NULL");
    if (prec != BLFID_FP64 && prec != BLFID_C64 && prec != BLFID_C64_BF16X3)
        return fail(BLFID_ERR_UNSUPPORTED,
                    "precision not implemented (BLFID_FP64, BLFID_C64, BLFID_C64_BF16X3)");
    if (sys->nspins < 1 || sys->nspins > DEV_MAX_SPINS)
        return fail(BLFID_ERR_INVALID, "nspins must be in [1, 12]");
    if (npts <= 0)
        return fail(BLFID_ERR_INVALID,
                    "Fid size was '0' please use 'npts(int)' set set it (oneFID.h:78-83)");
    if (sys->ncsa < 0 || sys->ndip < 0 || sys->nj < 0)
        return fail(BLFID_ERR_INVALID, "negative interaction count");
    const int S = sys->nspins;
    for (int c = 0; c < sys->ncsa; c++) {
        const double *p = sys->csa + 7 * c;
        if (p[2] > 1.0 || p[2] < 0.0)
            return fail(BLFID_ERR_INVALID, " 'eta' set to a negative or value >1...");
        if ((int)p[3] < 0 || (int)p[3] >= S) return fail(BLFID_ERR_INVALID, "CSA spin out of range");
    }
    for (int d = 0; d < sys->ndip; d++) {
        const double *p = sys->dip + 6 * d;
        int i = (int)p[1], j = (int)p[2];
        if (i < 0 || j < 0 || i >= S || j >= S || i == j)
            return fail(BLFID_ERR_INVALID, "dipole spin index out of range");
    }
    for (int q = 0; q < sys->nj; q++) {
        const double *p = sys->jc + 4 * q;
        int i = (int)p[1], j = (int)p[2];
        if (i < 0 || j < 0 || i >= S || j >= S || i == j)
            return fail(BLFID_ERR_INVALID, "J spin index out of range");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(BLFID_ERR_DEVICE, "no HIP device available");
    return BLFID_OK;
}

static blfid_plan *new_plan(const blfid_system *sys, double wr_hz, double rotor_deg, int npts,
                            double t0) {
    const int S = sys->nspins;
    blfid_plan *pl = new blfid_plan();
    pl->S = S;
    pl->N = 1 << S;
    pl->iso.assign(sys->isotope, sys->isotope + S);
    pl->csa.assign(sys->csa, sys->csa + 7 * sys->ncsa);
    pl->dip.assign(sys->dip, sys->dip + 6 * sys->ndip);
    pl->jc.assign(sys->jc, sys->jc + 4 * sys->nj);
    pl->wr = wr_hz;
    pl->rotor_deg = rotor_deg;
    pl->t0 = t0;
    pl->npts = npts;
    tune_from_process(pl->tn);
    return pl;
}

static blfid_status finish_plan(blfid_plan *pl, blfid_plan **out) {
    blfid_status st = build_tables(pl);
    if (st == BLFID_OK) st = upload_tables(pl);
    if (st != BLFID_OK) {
        delete pl;
        return st;
    }
    refresh_devplan(pl);
    *out = pl;
    return BLFID_OK;
}

extern "C" {

blfid_status blfid_plan_create(const blfid_system *sys, double wr_hz, double rotor_deg,
                               double sw_hz, int npts, double t0, blfid_precision prec,
                               blfid_plan **out) {
    blfid_status cs = check_system(sys, npts, prec, out);
    if (cs != BLFID_OK) return cs;
    blfid_plan *pl = new_plan(sys, wr_hz, rotor_deg, npts, t0);
    pl->prec = prec == BLFID_FP64 ? 0 : 1;
    pl->bf16x3 = prec == BLFID_C64_BF16X3;
    pl->sw = sw_hz;
    pl->spinning = !(wr_hz == 0.0 || rotor_deg == 0.0);
    if (pl->spinning) {
        int n = (int)std::floor(sw_hz / wr_hz + 0.5);
        if (n < 1) {
            delete pl;
            return fail(BLFID_ERR_INVALID, "sw/wr rounds to zero steps per rotor period");
        }
        pl->n = n;
        pl->sw_adj = (double)(n * wr_hz);
        double tmin = t0, tmax = t0 + 1.0 / wr_hz;
        pl->tau = (tmax - tmin) / n;
        if (pl->tau <= 0) {
            delete pl;
            return fail(BLFID_ERR_INVALID, " your time for the propgator is negative");
        }
    } else {
        if (sw_hz <= 0) {
            delete pl;
            return fail(BLFID_ERR_INVALID, "sweep width must be positive");
        }
        pl->n = 1;
        pl->sw_adj = sw_hz;
        pl->tau = 1.0 / sw_hz;
    }
    return finish_plan(pl, out);
}

} // extern "C"

// Shared body of the two pulse-program plan constructors: `steps` steps of length step_dt[k]
// (or 1/(wr steps) each when step_dt is NULL), consecutive from t0.
static blfid_status create_pulse_plan(const blfid_system *sys, double wr_hz, double rotor_deg,
                                      int steps, const double *step_t1, const double *step_dt,
                                      int npts, double t0, int nrf, const blfid_rf *rf,
                                      blfid_precision prec, blfid_plan **out) {
    blfid_status cs = check_system(sys, npts, prec, out);
    if (cs != BLFID_OK) return cs;
    if (steps < 1) return fail(BLFID_ERR_INVALID, "a pulse program needs at least one step");
    if (!step_dt && (!(wr_hz > 0.0) || !std::isfinite(wr_hz)))
        return fail(BLFID_ERR_INVALID, "ptop needs a positive rotor speed (one period per point)");
    if (step_dt) {
        if (!(wr_hz >= 0.0) || !std::isfinite(wr_hz))
            return fail(BLFID_ERR_INVALID, "rotor speed must be finite and >= 0");
        for (int k = 0; k < steps; k++)
            if (!(step_dt[k] > 0.0) || !std::isfinite(step_dt[k]) ||
                (step_t1 && !std::isfinite(step_t1[k])))
                return fail(BLFID_ERR_INVALID, "step lengths must be positive and finite");
    }
    if (nrf < 0 || (nrf > 0 && !rf)) return fail(BLFID_ERR_INVALID, "bad rf channel list");
    for (int c = 0; c < nrf; c++)
        if (!rf[c].amp_hz || !rf[c].phase_rad || !rf[c].offset_hz)
            return fail(BLFID_ERR_INVALID, "rf channel with a NULL table");
    const int S = sys->nspins;
    // one slot per pulsed isotope that labels at least one spin
    std::vector<int> slot_iso;
    std::vector<int> spin_rf(S, -1);
    for (int c = 0; c < nrf; c++) {
        bool has = false;
        for (int i = 0; i < S; i++) has |= sys->isotope[i] == rf[c].isotope;
        if (!has) continue; // no spin of that symbol: the channel adds nothing
        if (std::find(slot_iso.begin(), slot_iso.end(), rf[c].isotope) == slot_iso.end())
            slot_iso.push_back(rf[c].isotope);
    }
    for (int i = 0; i < S; i++)
        for (size_t q = 0; q < slot_iso.size(); q++)
            if (sys->isotope[i] == slot_iso[q]) spin_rf[i] = (int)q;
    // the phase is applied as exp(-i phi Fz_q); that commutes with every other term unless a
    // strong J couples a pulsed spin to a spin of another isotope
    for (int q = 0; q < sys->nj; q++) {
        const double *p = sys->jc + 4 * q;
        int i = (int)p[1], j = (int)p[2];
        if ((int)p[3] != 0 && p[0] != 0.0 && sys->isotope[i] != sys->isotope[j] &&
            (spin_rf[i] >= 0 || spin_rf[j] >= 0))
            return fail(BLFID_ERR_UNSUPPORTED,
                        "strong heteronuclear J coupling on a pulsed isotope is not supported");
    }
    const int nq = (int)slot_iso.size();
    std::vector<double> tab((size_t)steps * nq * 3, 0.0);
    for (int k = 0; k < steps; k++)
        for (int q = 0; q < nq; q++) {
            // channels with amp != 0 on this isotope (pulsedata.cc:114); several channels on one
            // isotope add as amp e^{i phi}
            int cnt = 0;
            double a1 = 0, p1 = 0;
            cplx c(0.0, 0.0);
            double off = 0;
            for (int ch = 0; ch < nrf; ch++) {
                if (rf[ch].isotope != slot_iso[q]) continue;
                double a = rf[ch].amp_hz[k], ph = rf[ch].phase_rad[k], o = rf[ch].offset_hz[k];
                if (!std::isfinite(a) || !std::isfinite(ph) || !std::isfinite(o))
                    return fail(BLFID_ERR_INVALID, "non-finite rf table entry");
                if (a == 0.0) continue;
                cnt++;
                a1 = a;
                p1 = ph;
                c += a * std::exp(cplx(0.0, ph));
                off += o;
            }
            double *t = &tab[(size_t)3 * (k * nq + q)];
            if (cnt == 1) { t[0] = 0.5 * a1; t[1] = p1; }
            else if (cnt > 1) { t[0] = 0.5 * std::abs(c); t[1] = std::arg(c); }
            t[2] = off;
        }
    blfid_plan *pl = new_plan(sys, wr_hz, rotor_deg, npts, t0);
    // The period propagator of a pulse program is raised to the power m for FID point m, so its
    // error grows linearly along the FID: with complex64 storage of the step propagators alone
    // (exact FP64 values rounded, products in FP64) one crystallite of the 13C + 6 1H rf case
    // already differs from FP64 by 1.5e-5 relative L2 over 2048 points, and 2.5e-4 with the
    // FP32 Taylor series and complex64 products.  Every precision therefore forms the period
    // propagator in FP64; the rest of this path (eigenbasis FID, powder sum) is FP64 anyway.
    (void)prec;
    pl->prec = 0;
    pl->bf16x3 = 0;
    pl->ptop = true;
    pl->spinning = false;
    pl->n = steps;
    if (step_dt) {
        pl->step_dt.assign(step_dt, step_dt + steps);
        if (step_t1) pl->step_t1.assign(step_t1, step_t1 + steps);
        double T = 0;
        for (double d : pl->step_dt) T += d;
        pl->tau = T / steps;
        pl->sw = pl->sw_adj = 1.0 / T;
    } else {
        pl->tau = (1.0 / wr_hz) / steps;
        pl->sw = wr_hz;
        pl->sw_adj = wr_hz;
    }
    pl->nrfiso = nq;
    pl->spin_rf = spin_rf;
    pl->rf = tab;
    return finish_plan(pl, out);
}

extern "C" {

blfid_status blfid_plan_create_ptop(const blfid_system *sys, double wr_hz, double rotor_deg,
                                    int steps, int npts, double t0, int nrf, const blfid_rf *rf,
                                    blfid_precision prec, blfid_plan **out) {
    return create_pulse_plan(sys, wr_hz, rotor_deg, steps, nullptr, nullptr, npts, t0, nrf, rf,
                             prec, out);
}

blfid_status blfid_plan_create_pulse(const blfid_system *sys, double wr_hz, double rotor_deg,
                                     int nsteps, const double *step_t1, const double *step_dt,
                                     int npts, double t0, int nrf, const blfid_rf *rf,
                                     blfid_precision prec, blfid_plan **out) {
    if (!step_dt) {
        if (out) *out = nullptr;
        return fail(BLFID_ERR_INVALID, "step_dt is NULL");
    }
    return create_pulse_plan(sys, wr_hz, rotor_deg, nsteps, step_t1, step_dt, npts, t0, nrf, rf,
                             prec, out);
}

blfid_status blfid_plan_set_prep(blfid_plan *pl, double rotor_deg, int nsteps,
                                 const double *step_t1, const double *step_dt, int nrf,
                                 const blfid_rf *rf) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    pl->gen++; // a recorded graph no longer matches the plan
    if (nsteps < 0) return fail(BLFID_ERR_INVALID, "negative step count");
    if (nsteps > 0 && !step_dt) return fail(BLFID_ERR_INVALID, "step_dt is NULL");
    if (nsteps > 0 && (!pl->spinning || pl->ptop))
        return fail(BLFID_ERR_UNSUPPORTED,
                    "pulses before the FID are implemented for the spinning gamma-COMPUTE FID");
    if (pl->last_stream) (void)hipStreamSynchronize(pl->last_stream);
    blfid_plan *pp = nullptr;
    if (nsteps > 0) {
        const blfid_system sys{pl->S, pl->iso.data(), (int)(pl->csa.size() / 7), pl->csa.data(),
                               (int)(pl->dip.size() / 6), pl->dip.data(), (int)(pl->jc.size() / 4),
                               pl->jc.data()};
        blfid_status s = create_pulse_plan(&sys, pl->wr, rotor_deg, nsteps, step_t1, step_dt, 1,
                                           0.0, nrf, rf, BLFID_FP64, &pp);
        if (s != BLFID_OK) return s;
        std::vector<int> sblk(pl->N), sloc(pl->N);
        for (size_t b = 0; b < pp->blocks.size(); b++)
            for (size_t a = 0; a < pp->blocks[b].states.size(); a++) {
                sblk[pp->blocks[b].states[a]] = (int)b;
                sloc[pp->blocks[b].states[a]] = (int)a;
            }
        hipError_t e = pl->d_prep_sblk.upload(sblk);
        if (e == hipSuccess) e = pl->d_prep_sloc.upload(sloc);
        if (e != hipSuccess) {
            delete pp;
            return fail(BLFID_ERR_DEVICE, std::string("prep tables: ") + hipGetErrorString(e));
        }
    }
    delete pl->prep;
    pl->prep = pp;
    pl->prep_rotor = rotor_deg;
    if (pp) pl->state = nullptr; // a preparation and a state are exclusive sources of rho
    pl->have_ops = false; // the block-pair tables depend on it: operators are set again
    return BLFID_OK;
}

void blfid_plan_destroy(blfid_plan *pl) {
    if (!pl) return;
    if (pl->last_stream) (void)hipStreamSynchronize(pl->last_stream);
    else (void)hipDeviceSynchronize();
    delete pl; // DevBuf destructors free the device memory
}

double blfid_plan_sw(const blfid_plan *pl) { return pl ? pl->sw_adj : 0.0; }
int blfid_plan_steps(const blfid_plan *pl) { return pl ? pl->n : 0; }
int blfid_plan_dim(const blfid_plan *pl) { return pl ? pl->N : 0; }

blfid_status blfid_plan_set_memory_budget(blfid_plan *pl, double bytes) {
    if (!pl || bytes < 0) return fail(BLFID_ERR_INVALID, "bad plan or budget");
    pl->gen++; // a recorded graph no longer matches the plan
    pl->budget = bytes;
    return BLFID_OK;
}

blfid_status blfid_plan_set_operators(blfid_plan *pl, const double *rho_d, const double *det_d) {
    if (!pl || !rho_d || !det_d) return fail(BLFID_ERR_INVALID, "NULL argument");
    pl->gen++; // a recorded graph no longer matches the plan
    const int N = pl->N;
    const cplx *rho = (const cplx *)rho_d, *det = (const cplx *)det_d;
    auto R = [&](int i, int j) { return rho[i + (size_t)j * N]; };
    auto D = [&](int i, int j) { return det[i + (size_t)j * N]; };
    int rosym = 1;
    for (int i = 0; i < N && rosym; i++)
        for (int j = 0; j < N; j++)
            if (R(i, j) != 0.5 * (D(i, j) + std::conj(D(j, i)))) {
                rosym = 0;
                break;
            }
    // with pulses before the FID (or rho from a state) every crystallite has its own rho, which
    // couples any pair of blocks: the general (non-rosym) form, on every block pair that D couples
    const bool prepared = pl->prep != nullptr || pl->state != nullptr;
    if (prepared) rosym = 0;
    pl->rosym = rosym;
    cplx tr = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) tr += D(i, j) * std::conj(D(i, j));
    pl->trdd = tr;
    // rho_eff: det + det^H for the rosym spinning case (compute.h:381-382), rho otherwise
    // (the static and ptop FIDs propagate rho itself)
    auto RE = [&](int i, int j) {
        if (pl->spinning && rosym) return D(i, j) + std::conj(D(j, i));
        return R(i, j);
    };
    // state -> (block, local index)
    std::vector<int> sblk(N), sloc(N);
    for (size_t b = 0; b < pl->blocks.size(); b++)
        for (size_t a = 0; a < pl->blocks[b].states.size(); a++) {
            sblk[pl->blocks[b].states[a]] = (int)b;
            sloc[pl->blocks[b].states[a]] = (int)a;
        }
    const int nb = (int)pl->blocks.size();
    std::vector<char> dnz(nb * nb, 0), rnz(nb * nb, 0);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            if (D(i, j) != 0.0) dnz[sblk[i] * nb + sblk[j]] = 1;
            if (RE(i, j) != 0.0 || prepared) rnz[sblk[i] * nb + sblk[j]] = 1;
        }
    pl->pairs.clear();
    for (int I = 0; I < nb; I++)
        for (int J = 0; J < nb; J++)
            if (dnz[I * nb + J] && rnz[J * nb + I]) pl->pairs.push_back({I, J});
    if (pl->spinning && rosym) {
        // pairs without a reverse pair (J, I) first: their elements have R = conj(Q) and form
        // the leading correlation chunks that take the autocorrelation kernel
        std::vector<char> has(nb * nb, 0);
        for (auto &q : pl->pairs) has[q.I * nb + q.J] = 1;
        std::stable_partition(pl->pairs.begin(), pl->pairs.end(),
                              [&](const Pair &q) { return !has[q.J * nb + q.I]; });
    }
    pl->pair_eoff.clear(); pl->pair_doff.clear(); pl->pair_roff.clear();
    pl->d_rowptr.clear(); pl->d_col.clear(); pl->d_val.clear();
    pl->r_rowptr.clear(); pl->r_col.clear(); pl->r_val.clear();
    pl->elem_r.clear(); pl->elem_s.clear();
    pl->pair_groups.clear();
    long eo = 0;
    for (size_t p = 0; p < pl->pairs.size(); p++) {
        int I = pl->pairs[p].I, J = pl->pairs[p].J;
        const auto &sI = pl->blocks[I].states, &sJ = pl->blocks[J].states;
        int bI = (int)sI.size(), bJ = (int)sJ.size();
        pl->pair_eoff.push_back(eo);
        eo += (long)bI * bJ;
        pl->pair_doff.push_back((int)pl->d_rowptr.size());
        for (int a = 0; a < bI; a++) {
            pl->d_rowptr.push_back((int)pl->d_col.size());
            for (int b = 0; b < bJ; b++) {
                cplx v = D(sI[a], sJ[b]);
                if (v != 0.0) {
                    pl->d_col.push_back(b);
                    pl->d_val.push_back(v.real());
                    pl->d_val.push_back(v.imag());
                }
            }
        }
        pl->d_rowptr.push_back((int)pl->d_col.size());
        pl->pair_roff.push_back((int)pl->r_rowptr.size());
        for (int b = 0; b < bJ; b++) {
            pl->r_rowptr.push_back((int)pl->r_col.size());
            for (int a = 0; a < bI; a++) {
                cplx v = RE(sJ[b], sI[a]);
                if (v != 0.0) {
                    pl->r_col.push_back(a);
                    pl->r_val.push_back(v.real());
                    pl->r_val.push_back(v.imag());
                }
            }
        }
        pl->r_rowptr.push_back((int)pl->r_col.size());
        for (int s = 0; s < bJ; s++)
            for (int r = 0; r < bI; r++) {
                pl->elem_r.push_back(pl->blk_soff[I] + r);
                pl->elem_s.push_back(pl->blk_soff[J] + s);
            }
        int M = std::min(5, (std::max(bI, bJ) + 15) / 16);
        pl->pair_groups[M].push_back((int)p);
    }
    pl->Etot = eo;
    pl->Estr = (eo + 15) & ~15L;
    // rosym spinning case: R_(I,J) = B_J^H (det_JI + det_IJ^H) B_I, i.e. R_(I,J)(s,r) =
    // Q_(J,I)(s,r) + conj(Q_(I,J)(r,s)); the pair (J,I) exists exactly when det_JI != 0, so R
    // needs neither its own products nor storage: k_corr forms it from Q through elem_rev
    pl->rt_from_q = pl->spinning && rosym;
    pl->elem_rev.clear();
    if (pl->rt_from_q) {
        std::map<std::pair<int, int>, int> pidx;
        for (size_t p = 0; p < pl->pairs.size(); p++) pidx[{pl->pairs[p].I, pl->pairs[p].J}] = (int)p;
        pl->elem_rev.assign(eo, -1);
        for (size_t p = 0; p < pl->pairs.size(); p++) {
            int I = pl->pairs[p].I, J = pl->pairs[p].J;
            auto it = pidx.find({J, I});
            if (it == pidx.end()) continue;
            const int bI = pl->blk_dim[I], bJ = pl->blk_dim[J];
            const long ro = pl->pair_eoff[it->second];
            for (int s2 = 0; s2 < bJ; s2++)
                for (int r = 0; r < bI; r++)
                    pl->elem_rev[pl->pair_eoff[p] + r + (long)s2 * bI] = (int)(ro + s2 + (long)r * bJ);
        }
    }
    pl->rt_rev = false;
    long first_rev = eo;
    for (long e = 0; e < (long)pl->elem_rev.size(); e++)
        if (pl->elem_rev[e] >= 0) {
            pl->rt_rev = true;
            first_rev = std::min(first_rev, e);
        }
    // correlation chunks before the first element with a reverse element: autocorrelation
    {
        const int CH = corr_chunk(pl->n);
        const int nchunk = (int)((eo + CH - 1) / CH);
        pl->corr_auto = !pl->rt_from_q ? 0 : !pl->rt_rev ? nchunk : (int)(first_rev / CH);
        pl->corr_rev_elems = 0;
        for (int v : pl->elem_rev) pl->corr_rev_elems += v >= 0;
    }
    if (pl->spinning && !pl->ptop)
        pl->nparts = eo > 0 ? corr_segments(pl->n, eo, corr_use_fft(pl->n) ? pl->corr_auto : 0) : 1;
    else
        pl->nparts = std::max<int>(1, (int)pl->pairs.size());
    if (pl->prep) {
        // rho's nonzero entries, from which each crystallite's prepared blocks are formed
        std::vector<int> nzi;
        std::vector<double> nzv;
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++)
                if (R(i, j) != 0.0) {
                    nzi.push_back(i);
                    nzi.push_back(j);
                    nzv.push_back(R(i, j).real());
                    nzv.push_back(R(i, j).imag());
                }
        pl->rho_nnz = (int)(nzv.size() / 2);
        HIPCHK(pl->d_rho_nzi.upload(nzi));
        HIPCHK(pl->d_rho_nzv.upload(nzv));
    }
    std::vector<int> pI, pJ;
    for (auto &pp : pl->pairs) {
        pI.push_back(pp.I);
        pJ.push_back(pp.J);
    }
    HIPCHK(pl->d_pair_I.upload(pI));
    HIPCHK(pl->d_pair_J.upload(pJ));
    HIPCHK(pl->d_pair_eoff.upload(pl->pair_eoff));
    HIPCHK(pl->d_pair_doff.upload(pl->pair_doff));
    HIPCHK(pl->d_pair_roff.upload(pl->pair_roff));
    HIPCHK(pl->d_drow.upload(pl->d_rowptr));
    HIPCHK(pl->d_dcol.upload(pl->d_col));
    HIPCHK(pl->d_dval.upload(pl->d_val));
    HIPCHK(pl->d_rrow.upload(pl->r_rowptr));
    HIPCHK(pl->d_rcol.upload(pl->r_col));
    HIPCHK(pl->d_rval.upload(pl->r_val));
    HIPCHK(pl->d_er.upload(pl->elem_r));
    HIPCHK(pl->d_es.upload(pl->elem_s));
    HIPCHK(pl->d_erev.upload(pl->elem_rev));
    for (auto &kv : pl->pair_groups) HIPCHK(pl->d_pgrp[kv.first].upload(kv.second));
    pl->have_ops = true;
    refresh_devplan(pl);
    return BLFID_OK;
}

blfid_status blfid_plan_set_powder(blfid_plan *pl, const double *theta, const double *phi,
                                   const double *gamma, const double *weight, int K) {
    if (!pl || K < 0 || (K > 0 && (!theta || !phi || !gamma || !weight)))
        return fail(BLFID_ERR_INVALID, "bad powder arguments");
    pl->gen++; // a recorded graph no longer matches the plan
    pl->K = K;
    if (K == 0) return BLFID_OK;
    std::vector<double> th(theta, theta + K), ph(phi, phi + K), ga(gamma, gamma + K),
        w(weight, weight + K);
    HIPCHK(pl->d_th.upload(th));
    HIPCHK(pl->d_ph.upload(ph));
    HIPCHK(pl->d_ga.upload(ga));
    HIPCHK(pl->d_w.upload(w));
    HIPCHK(pl->d_RP.alloc((size_t)9 * K));
    // the rotations are formed by the next run, on its stream, so that setting a powder never
    // waits for work of other plans
    pl->rp_dirty = true;
    refresh_devplan(pl);
    return BLFID_OK;
}

// Scratch slots of a launch over a block group (or pair group) whose largest matrix has `e`
// entries: the plan's scratch holds `slots` slots of bm2 = bmax^2 entries, so a group of smaller
// blocks gets proportionally more slots of e entries (more workgroups in flight), up to 16384
// for the 64-thread single-wave variant (M = 1) and 8192 otherwise.
static void group_slots(size_t bm2, int slots, size_t e, int M, size_t *slot, int *n) {
    *slot = std::max<size_t>(e, 1);
    const double k = (double)slots * (double)bm2 / (double)*slot;
    *n = (int)std::min<double>(k, M == 1 ? 16384.0 : 8192.0);
}
static size_t group_e(const blfid_plan *pl, const std::vector<int> &g) {
    size_t e = 0;
    for (int b : g) e = std::max<size_t>(e, (size_t)pl->blk_dim[b] * pl->blk_dim[b]);
    return e;
}
static size_t pair_e(const blfid_plan *pl, const std::vector<int> &g) {
    size_t e = 0;
    for (int p : g)
        e = std::max<size_t>(e, (size_t)pl->blk_dim[pl->pairs[p].I] * pl->blk_dim[pl->pairs[p].J]);
    return e;
}
static double group_b3(const blfid_plan *pl, const std::vector<int> &g) {
    double s = 0;
    for (int b : g) s += (double)pl->blk_dim[b] * pl->blk_dim[b] * pl->blk_dim[b];
    return s;
}
static double group_b2(const blfid_plan *pl, const std::vector<int> &g) {
    double s = 0;
    for (int b : g) s += (double)pl->blk_dim[b] * pl->blk_dim[b];
    return s;
}
static double group_b1(const blfid_plan *pl, const std::vector<int> &g) {
    double s = 0;
    for (int b : g) s += (double)pl->blk_dim[b];
    return s;
}
// sum over the group's pairs of b_I^2 + b_J^2 (the two dense blocks a pair's product reads)
static double pair_reads(const blfid_plan *pl, const std::vector<int> &g) {
    double s = 0;
    for (int p : g) {
        double bI = pl->blk_dim[pl->pairs[p].I], bJ = pl->blk_dim[pl->pairs[p].J];
        s += bI * bI + bJ * bJ;
    }
    return s;
}
static double pair_flops(const blfid_plan *pl, const std::vector<int> &g) {
    double s = 0;
    for (int p : g) {
        double bI = pl->blk_dim[pl->pairs[p].I], bJ = pl->blk_dim[pl->pairs[p].J];
        s += 8.0 * (bI * bI * bJ + (pl->rt_from_q ? 0.0 : bJ * bJ * bI));  // Q (and R)
        if (pl->prep || pl->state) s += 8.0 * bJ * bI * bI;                // rho_c B_I
    }
    return s;
}


// Executed flops of the correlation per element and crystallite: FFT variant (n a power of two,
// 16..512) as T transforms of np = ceil(log2(n) / 3) passes, radix 8 (a 56-flop DFT and 7
// twiddle products per butterfly) but the last, of radix rl without twiddles (DFT of 4, 16 or
// 56 flops), plus the scaled product (8 per point); direct variant 8 n^2; plus twiddle staging
// (18 n) and the FID sums (14 per point).  T = 3 in general; the autocorrelation variant
// (R = conj(Q)) forms two transforms, a 4-flop |q|^2 / n product and stages Q alone (9 n).
static double corr_flops_per_elem(int n, int P, bool autoc = false) {
    double c = 8.0 * (double)n * n;
    double stage = 18.0 * n;
    if (n >= 16 && (n & (n - 1)) == 0 && n <= 512) {
        const int ln = __builtin_ctz((unsigned)n), np = (ln + 2) / 3, rl = 1 << (ln - 3 * (np - 1));
        const double dft = rl == 2 ? 4.0 : rl == 4 ? 16.0 : 56.0;
        const double T = autoc ? 2.0 : 3.0;
        c = T * (np - 1) * (n / 8.0) * 98.0 + T * ((double)n / rl) * dft + (autoc ? 4.0 : 8.0) * n;
        if (autoc) stage = 9.0 * n;
    }
    return c + stage + 14.0 * P;
}

static double pair_elems(const blfid_plan *pl, const std::vector<int> &g) {
    double s = 0;
    for (int p : g) s += (double)pl->blk_dim[pl->pairs[p].I] * pl->blk_dim[pl->pairs[p].J];
    return s;
}

static blfid_layout::Shape shape_of(const blfid_plan *pl) {
    blfid_layout::Shape s;
    s.prec = pl->prec;
    s.kind = pl->ptop ? 2 : (pl->spinning ? 0 : 1);
    s.n = pl->n;
    s.npts = pl->npts;
    s.N = pl->N;
    s.bmax = pl->bmax;
    s.rt_from_q = pl->rt_from_q ? 1 : 0;
    s.nparts = pl->nparts;
    s.exp_slots = pl->exp_slots;
    s.qr_slots = pl->qr_slots;
    s.eig_slots = pl->eig_slots;
    s.Lsum = pl->Lsum;
    s.Estr = pl->Estr;
    s.K = pl->K;
    s.rho_elems = (pl->prep || pl->state) ? pl->Etot : 0;
    return s;
}

// Layout of a preparation plan inside its FID plan's arena: only its propagators run there.
static blfid_layout::Shape prep_shape_of(const blfid_plan *pp) {
    blfid_layout::Shape s = shape_of(pp);
    s.qr_slots = 0;
    s.eig_slots = 0;
    s.nparts = 0;
    s.K = 0;
    return s;
}

// The plan's layout with its preparation's (ps, referenced by s.prep) when it has one.
static void full_shape(const blfid_plan *pl, blfid_layout::Shape &s, blfid_layout::Shape &ps) {
    s = shape_of(pl);
    if (pl->prep) {
        ps = prep_shape_of(pl->prep);
        s.prep = &ps;
    }
}

static double bytes_per_crystallite(const blfid_plan *pl) {
    blfid_layout::Shape sh, ps;
    full_shape(pl, sh, ps);
    return (double)(blfid_layout::layout_bytes(sh, 1) - blfid_layout::layout_bytes(sh, 0));
}

blfid_status blfid_plan_stats(const blfid_plan *pl, double *out) {
    if (!pl || !out) return fail(BLFID_ERR_INVALID, "NULL argument");
    double N = pl->N, n = pl->n, P = pl->npts;
    // executed FP64 flops per crystallite-step (unpadded block sizes; complex MAC = 8)
    double s3 = 0, pairs = 0, pairsQ = 0;
    for (int d : pl->blk_dim) s3 += (double)d * d * d;
    for (auto &pp : pl->pairs) {
        double bI = pl->blk_dim[pp.I], bJ = pl->blk_dim[pp.J];
        pairs += bI * bJ;
        pairsQ += bI * bJ * (bI + (pl->rt_from_q ? 0.0 : bJ));
    }
    double ex = 0;
    if (pl->ptop) {
        ex += 2.0 * 7.0 * s3;        // exp: nominal degree-8 Taylor, 7 real GEMMs
        ex += 8.0 * s3;              // Dyson
        ex += 8.0 * pairsQ / n;      // D and rho into the period eigenbasis
        ex += 8.0 * P * pairs / n;   // eigenbasis FID (complex MAC per element and point)
    } else if (pl->spinning) {
        ex += 2.0 * 7.0 * s3;   // exp: nominal degree-8 Taylor, 7 real GEMMs
        ex += 8.0 * s3;         // Dyson
        ex += 8.0 * s3;         // B = A W
        ex += 8.0 * pairsQ;     // Q and R
        const double ea = corr_use_fft(pl->n)
                              ? std::min<double>((double)pl->corr_auto * corr_chunk(pl->n), (double)pl->Etot)
                              : 0.0;
        ex += (ea * corr_flops_per_elem(pl->n, pl->npts, true) +
               ((double)pl->Etot - ea) * corr_flops_per_elem(pl->n, pl->npts, false)) / n;
    }
    out[0] = ex;
    if (pl->ptop) out[1] = 8.0 * N * N * N + 16.0 * N * N * N * P / n;
    else out[1] = 48.0 * N * N * N + 14.0 * n * N * N + 14.0 * P * N * N / n;
    out[2] = (double)pl->blocks.size();
    out[3] = (double)pl->bmax;
    out[4] = bytes_per_crystallite(pl);
    blfid_layout::Shape sh, ps;
    full_shape(pl, sh, ps);
    out[5] = (double)blfid_layout::layout_bytes(sh, 0);
    return BLFID_OK;
}

static void carve(blfid_plan *pl, char *base, const size_t *off) {
    using namespace blfid_layout;
    pl->bufU = base + off[R_U];
    pl->bufA = base + off[R_A];
    pl->bufQt = base + off[R_QT];
    pl->bufRt = base + off[R_RT];
    pl->scrE = base + off[R_EXP];
    pl->bufW = (double2 *)(base + off[R_W]);
    pl->thetas = (double *)(base + off[R_THETA]);
    pl->scrQ = base + off[R_QR];
    pl->part = (double *)(base + off[R_PART]);
    pl->scrEig = (double2 *)(base + off[R_EIG]);
    pl->bufWr = pl->prec && pl->spinning ? base + off[R_WR] : nullptr;
    pl->grp = (double *)(base + off[R_GROUP]);
    pl->bufRho = (pl->prep || pl->state) ? base + off[R_RHO] : nullptr;
}

// Enqueues the whole powder on `st` with the working storage at `base` (regions at off[],
// batches of Kb crystallites) and the normalised FID written to `out` (2 npts doubles, device).
// Every error return after the first launch leaves the launched work in the stream.
// The crystallites [c_begin, c_end) of the powder in batches of Kb; `first` starts the FID (the
// powder rotations when new, the group sums zeroed), `last` finishes it into `out`.  The group
// sums carry over between calls, so consecutive ranges give the whole run's FID bit for bit.
static blfid_status enqueue_range(blfid_plan *pl, const blfid_layout::Shape &sh, char *base,
                                  const size_t *off, long Kb, int normalize, double *out,
                                  hipStream_t st, int c_begin, int c_end, bool first, bool last) {
    const int K = pl->K;
    if (pl->state && pl->state->K != K)
        return fail(BLFID_ERR_INVALID, "the state and the plan's powder differ in size");
    pl->Kb = (int)Kb;
    carve(pl, base, off);
    refresh_devplan(pl);
    const DevPlan *P = &pl->dp;
    if (pl->rp_dirty && K > 0) {
        LAUNCHCHK(blfid_dev_launch_prep(K, pl->d_th.p, pl->d_ph.p, pl->d_ga.p, pl->d_RP.p, st));
        pl->rp_dirty = false;
    }
    const int G = blfid_layout::CRYST_GROUP;
    const int ng = (int)blfid_layout::ngroups(K);
    if (first) LAUNCHCHK(blfid_dev_launch_zero(pl->grp, (size_t)ng * 2 * 2 * pl->npts, st));
    const size_t bm2 = (size_t)pl->bmax * pl->bmax;
    const int pr = pl->prec;
    const size_t cs = pr ? 8 : 16;
    const int ff = pl->rosym == 1 ? 2 : 1;
    const double scale = (pl->spinning && !pl->ptop) ? 1.0 / (double)(pl->n / ff) : 1.0;
    for (int c0 = c_begin; c0 < c_end; c0 += (int)Kb) {
        int kb = (int)std::min<long>(Kb, c_end - c0);
        // the FID parts are all written by the correlation / eigenbasis-FID kernels, unless
        // the operators couple no block pair at all
        if (pl->pairs.empty())
            LAUNCHCHK(blfid_dev_launch_zero(pl->part, (size_t)kb * pl->nparts * 2 * pl->npts, st));
        if (pl->ptop) {
            // U_k = exp(+i theta (H_k + H_rf,k)); A_n = U_n ... U_1; its eigenbasis gives
            // rho_m = A_n^m rho A_n^-m in closed form
            for (auto &kv : pl->blk_groups) {
                PhaseScope ps(pl, 0 * 8 + kv.first, 0.0,
                              kb * (double)pl->n * group_b2(pl, kv.second) * cs, st); // flops: k_exp
                LAUNCHCHK(blfid_dev_launch_exp(pr, kv.first, P, kb, c0, pl->d_bgrp[kv.first].p,
                                               (int)kv.second.size(), pl->scrE, bm2, sh.exp_slots,
                                               pl->bufU, st));
            }
            for (auto &kv : pl->blk_groups) {
                PhaseScope ps(pl, 1 * 8 + kv.first, kb * (double)(pl->n - 1) * group_b3(pl, kv.second) * 8.0,
                              kb * (double)(pl->n + 1) * group_b2(pl, kv.second) * cs, st);
                LAUNCHCHK(blfid_dev_launch_dyson(pr, kv.first, P, kb, pl->d_bgrp[kv.first].p,
                                                 (int)kv.second.size(), pl->bufU, pl->bufA, st));
            }
            for (auto &kv : pl->eig_cls) {
                PhaseScope ps(pl, 2 * 8 + kv.first + 1, 0.0,
                              kb * (group_b2(pl, kv.second) * (cs + 16.0) + group_b1(pl, kv.second) * 8.0), st);
                LAUNCHCHK(blfid_dev_launch_eig(pr, P, kb, pl->d_eig_cls[kv.first].p,
                                               (int)kv.second.size(),
                                               pl->bufA + (size_t)((pl->n - 1) % 2) * pl->Lsum * cs,
                                               (size_t)2 * pl->Lsum, 0, pl->bufW, nullptr,
                                               pl->thetas, kv.first, pl->scrEig, bm2,
                                               sh.eig_slots, st));
            }
            if (pl->state)
                LAUNCHCHK(blfid_dev_launch_state_gather(0, P, kb, c0, pl->state->rho.p,
                                                        pl->bufRho, st));
            for (auto &kv : pl->pair_groups) {
                PhaseScope ps(pl, 7 * 8 + kv.first, kb * (pair_flops(pl, kv.second) + 8.0 * pl->npts * pair_elems(pl, kv.second)),
                              kb * (pair_reads(pl, kv.second) * 16.0 + (double)kv.second.size() * pl->npts * 16.0), st);
                LAUNCHCHK(blfid_dev_launch_static_fid(kv.first, P, kb, pl->d_pgrp[kv.first].p,
                                                      (int)kv.second.size(), pl->bufW,
                                                      pl->thetas, (double2 *)pl->scrQ, bm2,
                                                      sh.qr_slots, 1.0, 0, pl->part, st));
            }
        } else if (pl->spinning) {
            if (pl->prep) {
                // pulses before the FID: the preparation propagator of each crystallite (FP64,
                // in the preparation plan's blocks, its storage inside this arena), then the
                // prepared rho on the FID plan's block pairs
                blfid_plan *pp = pl->prep;
                size_t poff[blfid_layout::R_COUNT];
                (void)blfid_layout::layout_bytes(*sh.prep, Kb, poff);
                carve(pp, base + off[blfid_layout::R_PREP], poff);
                refresh_devplan(pp);
                pp->dp.RP = pl->d_RP.p;
                const DevPlan *PP = &pp->dp;
                const size_t pbm2 = (size_t)pp->bmax * pp->bmax;
                for (auto &kv : pp->blk_groups) {
                    PhaseScope ps(pl, 0 * 8 + kv.first, 0.0,
                                  kb * (double)pp->n * group_b2(pp, kv.second) * 16.0, st);
                    LAUNCHCHK(blfid_dev_launch_exp(0, kv.first, PP, kb, c0, pp->d_bgrp[kv.first].p,
                                                   (int)kv.second.size(), pp->scrE, pbm2,
                                                   sh.prep->exp_slots, pp->bufU, st));
                }
                for (auto &kv : pp->blk_groups) {
                    PhaseScope ps(pl, 1 * 8 + kv.first,
                                  kb * (double)(pp->n - 1) * group_b3(pp, kv.second) * 8.0,
                                  kb * (double)(pp->n + 1) * group_b2(pp, kv.second) * 16.0, st);
                    LAUNCHCHK(blfid_dev_launch_dyson(0, kv.first, PP, kb, pp->d_bgrp[kv.first].p,
                                                     (int)kv.second.size(), pp->bufU, pp->bufA, st));
                }
                LAUNCHCHK(blfid_dev_launch_rho_prep(
                    pr, P, PP, kb,
                    (const double2 *)(pp->bufA + (size_t)((pp->n - 1) % 2) * pp->Lsum * 16),
                    pl->d_prep_sblk.p, pl->d_prep_sloc.p, pl->d_rho_nzi.p, pl->d_rho_nzv.p,
                    pl->rho_nnz, pl->bufRho, st));
            } else if (pl->state) {
                // each crystallite's rho from the state, on the plan's block pairs
                LAUNCHCHK(blfid_dev_launch_state_gather(pr, P, kb, c0, pl->state->rho.p,
                                                        pl->bufRho, st));
            }
            // complex64 matrix-core plans: blocks of <= 80 states take the fused propagator +
            // Dyson chain (U_k never leaves LDS); larger blocks, the FP64 form and the BF16x3
            // form take the propagator and the Dyson chain as two kernels; a tuning may send a
            // tile class of the fused path to the two kernels
            const bool fused = pr == 1 && !pl->bf16x3;
            // (tile variant, device group list, host group list) of the two-kernel path
            std::vector<std::pair<int, std::pair<const int *, const std::vector<int> *>>> two;
            if (fused) {
                for (auto &kv : pl->fuse_groups) {
                    if (pl->tn[kv.first - 1] == 1) {
                        two.push_back({kv.first, {pl->d_fgrp[kv.first].p, &kv.second}});
                        continue;
                    }
                    PhaseScope ps(pl, 8 * 8 + kv.first, 0.0,
                                  kb * (double)pl->n * group_b2(pl, kv.second) * cs, st); // flops: device
                    LAUNCHCHK(blfid_dev_launch_expdyson(kv.first, P, kb, c0, pl->d_fgrp[kv.first].p,
                                                        (int)kv.second.size(), pl->bufA, st));
                }
                if (!pl->big_blocks.empty()) two.push_back({5, {pl->d_bigb.p, &pl->big_blocks}});
            } else {
                for (auto &kv : pl->blk_groups) two.push_back({kv.first, {pl->d_bgrp[kv.first].p, &kv.second}});
            }
            for (auto &gq : two) {
                const int M = gq.first;
                const std::vector<int> &gv = *gq.second.second;
                PhaseScope ps(pl, 0 * 8 + M, 0.0, kb * (double)pl->n * group_b2(pl, gv) * cs, st); // flops: k_exp
                size_t slot;
                int ns;
                group_slots(bm2, sh.exp_slots, group_e(pl, gv), M, &slot, &ns);
                LAUNCHCHK(blfid_dev_launch_exp(pr, M, P, kb, c0, gq.second.first, (int)gv.size(),
                                               pl->scrE, slot, ns, pl->bufU, st));
            }
            for (auto &gq : two) {
                const int M = gq.first;
                const std::vector<int> &gv = *gq.second.second;
                PhaseScope ps(pl, 1 * 8 + M, kb * (double)(pl->n - 1) * group_b3(pl, gv) * 8.0,
                              kb * 2.0 * pl->n * group_b2(pl, gv) * cs, st);
                LAUNCHCHK(blfid_dev_launch_dyson(pr, M, P, kb, gq.second.first, (int)gv.size(),
                                                 pl->bufU, pl->bufA, st));
            }
            for (auto &kv : pl->eig_cls) {
                PhaseScope ps(pl, 2 * 8 + kv.first + 1, 0.0,
                              kb * (group_b2(pl, kv.second) * (2.0 * cs) + group_b1(pl, kv.second) * 8.0), st);
                LAUNCHCHK(blfid_dev_launch_eig(pr, P, kb, pl->d_eig_cls[kv.first].p,
                                               (int)kv.second.size(),
                                               pl->bufA + (size_t)(pl->n - 1) * pl->Lsum * cs,
                                               (size_t)pl->n * pl->Lsum, 0, pl->bufW, pl->bufWr,
                                               pl->thetas, kv.first, pl->scrEig, bm2,
                                               sh.eig_slots, st));
            }
            for (auto &kv : pl->blk_groups) {
                PhaseScope ps(pl, 3 * 8 + kv.first, kb * (double)pl->n * group_b3(pl, kv.second) * 8.0,
                              kb * (2.0 * pl->n + 1.0) * group_b2(pl, kv.second) * cs, st);
                LAUNCHCHK(blfid_dev_launch_bmul(pr, kv.first, P, kb, pl->d_bgrp[kv.first].p,
                                                (int)kv.second.size(), pl->bufA,
                                                pr ? (const void *)pl->bufWr : (const void *)pl->bufW,
                                                pl->bufU, st));
            }
            for (auto &kv : pl->pair_groups) {
                PhaseScope ps(pl, 4 * 8 + kv.first, kb * (double)pl->n * pair_flops(pl, kv.second),
                              kb * (double)pl->n * cs *
                                  (pair_reads(pl, kv.second) + (pl->rt_from_q ? 1.0 : 2.0) * pair_elems(pl, kv.second)),
                              st);
                size_t slot;
                int ns;
                group_slots(bm2, sh.qr_slots, pair_e(pl, kv.second), kv.first, &slot, &ns);
                LAUNCHCHK(blfid_dev_launch_qr(pr, kv.first, P, kb, pl->d_pgrp[kv.first].p,
                                              (int)kv.second.size(), pl->bufU, pl->scrQ, slot,
                                              ns, pl->bufQt, pl->bufRt, st));
            }
            if (pl->Etot > 0) {
                // Q once per element; R too unless it is formed from Q, or the reverse element's
                // Q where there is one.  Elements of the autocorrelation chunks take that
                // variant's flop count, the others the general one.
                const double reads = pl->rt_from_q ? (double)pl->Etot + (double)pl->corr_rev_elems
                                                   : 2.0 * (double)pl->Etot;
                const double ea = corr_use_fft(pl->n)
                                      ? std::min<double>((double)pl->corr_auto * corr_chunk(pl->n), (double)pl->Etot)
                                      : 0.0;
                const double fl = ea * corr_flops_per_elem(pl->n, pl->npts, true) +
                                  ((double)pl->Etot - ea) * corr_flops_per_elem(pl->n, pl->npts, false);
                PhaseScope ps(pl, 5 * 8, kb * fl,
                              kb * ((double)pl->n * reads * cs + (double)pl->nparts * pl->npts * 16.0), st);
                LAUNCHCHK(blfid_dev_launch_corr(pr, P, kb, pl->bufQt, pl->bufRt, pl->thetas,
                                                pl->d_er.p, pl->d_es.p, pl->part, st));
            }
        } else {
            for (auto &kv : pl->blk_groups)
                LAUNCHCHK(blfid_dev_launch_static_h(P, kb, c0, pl->d_bgrp[kv.first].p,
                                                    (int)kv.second.size(), (double *)pl->scrE, bm2,
                                                    sh.exp_slots, (double2 *)pl->bufU, st));
            for (auto &kv : pl->eig_cls)
                LAUNCHCHK(blfid_dev_launch_eig(0, P, kb, pl->d_eig_cls[kv.first].p,
                                               (int)kv.second.size(), pl->bufU, (size_t)pl->Lsum, 1,
                                               pl->bufW, nullptr, pl->thetas, kv.first,
                                               pl->scrEig, bm2, sh.eig_slots, st));
            if (pl->state)
                LAUNCHCHK(blfid_dev_launch_state_gather(0, P, kb, c0, pl->state->rho.p,
                                                        pl->bufRho, st));
            for (auto &kv : pl->pair_groups) {
                PhaseScope ps(pl, 7 * 8 + kv.first, 0.0,
                              kb * (pair_reads(pl, kv.second) * 16.0 + (double)kv.second.size() * pl->npts * 16.0), st);
                LAUNCHCHK(blfid_dev_launch_static_fid(kv.first, P, kb, pl->d_pgrp[kv.first].p,
                                                      (int)kv.second.size(), pl->bufW,
                                                      pl->thetas, (double2 *)pl->scrQ, bm2,
                                                      sh.qr_slots, -PI2 / pl->sw, 1, pl->part, st));
            }
        }
        {
            PhaseScope ps(pl, 6 * 8, 4.0 * kb * 2.0 * pl->npts * pl->nparts,
                          (double)kb * pl->nparts * pl->npts * 16.0, st);
            LAUNCHCHK(blfid_dev_launch_reduce(kb, c0, pl->npts, pl->nparts, G, pl->d_w.p, scale,
                                              pl->part, pl->grp, st));
        }
    }
    if (last) {
        cplx norm = normalize ? pl->trdd : cplx(1.0, 0.0);
        LAUNCHCHK(blfid_dev_launch_finish(pl->npts, ng, pl->grp, norm.real(), norm.imag(), out, st));
    }
    return BLFID_OK;
}

static blfid_status enqueue_run(blfid_plan *pl, const blfid_layout::Shape &sh, char *base,
                                const size_t *off, long Kb, int normalize, double *out,
                                hipStream_t st) {
    return enqueue_range(pl, sh, base, off, Kb, normalize, out, st, 0, pl->K, true, true);
}

// The run of crystallites [0, c_end) on the plan's working storage (the whole powder for
// blfid_plan_run; a leading sample for the autotuner, whose FID is not a result).
static blfid_status run_upto(blfid_plan *pl, int normalize, hipStream_t st, int c_end);

blfid_status blfid_plan_run(blfid_plan *pl, int normalize, void *stream) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    return run_upto(pl, normalize, (hipStream_t)stream, pl->K);
}

static blfid_status run_upto(blfid_plan *pl, int normalize, hipStream_t st, int c_end) {
    if (!pl->have_ops) return fail(BLFID_ERR_INVALID, "operators not set");
    if (normalize && pl->trdd == 0.0) return fail(BLFID_ERR_INVALID, "Tr(det det^H) is zero");
    pl->last_stream = st;
    blfid_layout::Shape sh, psh;
    full_shape(pl, sh, psh);
    // batch size: the largest whose exact layout fits the workspace / memory budget
    double avail;
    if (pl->ws) {
        avail = (double)pl->ws->bytes;
    } else {
        avail = pl->budget;
        if (avail <= 0) {
            size_t fre = 0, tot = 0;
            HIPCHK(hipMemGetInfo(&fre, &tot));
            avail = 0.4 * (double)(fre + pl->own_ws.n);
        }
    }
    const long Kb = blfid_layout::choose_batch(sh, pl->K, avail);
    if (Kb < 1) return fail(BLFID_ERR_NOMEM, "one crystallite does not fit the memory budget");
    size_t off[blfid_layout::R_COUNT];
    const size_t need = blfid_layout::layout_bytes(sh, Kb, off);
    char *base;
    if (pl->ws) {
        if (need > pl->ws->bytes)
            return fail(BLFID_ERR_NOMEM, "working storage exceeds the workspace");
        base = pl->ws->base;
        if (pl->ws->used && pl->ws->last != st) HIPCHK(hipStreamWaitEvent(st, pl->ws->ev, 0));
    } else {
        HIPCHK(pl->own_ws.alloc(need));
        base = pl->own_ws.p;
    }
    HIPCHK(pl->fid_out.alloc((size_t)2 * pl->npts));
    blfid_status rs = enqueue_range(pl, sh, base, off, Kb, normalize, pl->fid_out.p, st, 0,
                                    c_end, true, true);
    pl->ran = rs == BLFID_OK && c_end == pl->K;
    if (pl->ws) {
        // the next user of the workspace waits for whatever was enqueued, on every exit
        hipError_t e = hipEventRecord(pl->ws->ev, st);
        pl->ws->last = st;
        pl->ws->used = true;
        if (rs == BLFID_OK && e != hipSuccess)
            return fail(BLFID_ERR_DEVICE, std::string("hipEventRecord: ") + hipGetErrorString(e));
    }
    return rs;
}

// ------------------------------------------------------------------------------ spin state
blfid_status blfid_state_create(int nspins, int K, const double *theta, const double *phi,
                                const double *gamma, const double *weight, blfid_state **out) {
    if (!out) return fail(BLFID_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (nspins < 1 || nspins > DEV_MAX_SPINS)
        return fail(BLFID_ERR_UNSUPPORTED, "spin count outside 1.." + std::to_string(DEV_MAX_SPINS));
    if (K < 0 || (K > 0 && (!theta || !phi || !gamma || !weight)))
        return fail(BLFID_ERR_INVALID, "bad powder arguments");
    std::unique_ptr<blfid_state> st(new (std::nothrow) blfid_state());
    if (!st) return fail(BLFID_ERR_NOMEM, "host allocation");
    st->S = nspins;
    st->N = 1 << nspins;
    st->K = K;
    const size_t NN = (size_t)st->N * st->N;
    HIPCHK(st->rho.alloc(std::max<size_t>((size_t)K * NN, 1)));
    HIPCHK(hipMemset(st->rho.p, 0, std::max<size_t>((size_t)K * NN, 1) * sizeof(double2)));
    if (K > 0) {
        HIPCHK(st->d_th.upload(std::vector<double>(theta, theta + K)));
        HIPCHK(st->d_ph.upload(std::vector<double>(phi, phi + K)));
        HIPCHK(st->d_ga.upload(std::vector<double>(gamma, gamma + K)));
        HIPCHK(st->d_w.upload(std::vector<double>(weight, weight + K)));
        HIPCHK(st->d_RP.alloc((size_t)9 * K));
        LAUNCHCHK(blfid_dev_launch_prep(K, st->d_th.p, st->d_ph.p, st->d_ga.p, st->d_RP.p, nullptr));
        HIPCHK(hipDeviceSynchronize());
    }
    *out = st.release();
    return BLFID_OK;
}

void blfid_state_destroy(blfid_state *st) {
    if (!st) return;
    (void)hipDeviceSynchronize();
    delete st;
}

int blfid_state_dim(const blfid_state *st) { return st ? st->N : 0; }
int blfid_state_size(const blfid_state *st) { return st ? st->K : 0; }

blfid_status blfid_state_set_memory_budget(blfid_state *st, double bytes) {
    if (!st || !(bytes >= 0)) return fail(BLFID_ERR_INVALID, "bad memory budget");
    st->budget = bytes;
    return BLFID_OK;
}

// Fills every crystallite (all == 0: the one N x N rho; all == 1: K matrices in order).
static blfid_status state_fill(blfid_state *st, const double *rho, int all, hipStream_t s) {
    if (!st || !rho) return fail(BLFID_ERR_INVALID, "NULL argument");
    const size_t NN = (size_t)st->N * st->N;
    HIPCHK(hipStreamSynchronize(s)); // the staging buffer may still feed earlier work
    if (all) {
        if (st->K > 0)
            HIPCHK(hipMemcpyAsync(st->rho.p, rho, (size_t)st->K * NN * sizeof(double2),
                                  hipMemcpyHostToDevice, s));
    } else {
        HIPCHK(st->rho0.alloc(NN));
        HIPCHK(hipMemcpy(st->rho0.p, rho, NN * sizeof(double2), hipMemcpyHostToDevice));
        LAUNCHCHK(blfid_dev_launch_state_fill(st->K, st->N, st->rho0.p, st->rho.p, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return BLFID_OK;
}

blfid_status blfid_state_set(blfid_state *st, const double *rho, void *stream) {
    return state_fill(st, rho, 0, (hipStream_t)stream);
}

blfid_status blfid_state_project(blfid_state *st, const double *op, void *stream) {
    if (!st || !op) return fail(BLFID_ERR_INVALID, "NULL argument");
    const size_t NN = (size_t)st->N * st->N;
    double nrm = 0.0;
    for (size_t e = 0; e < NN; e++) nrm += op[2 * e] * op[2 * e] + op[2 * e + 1] * op[2 * e + 1];
    if (!(nrm > 0.0)) return fail(BLFID_ERR_INVALID, "the cycler operator is zero");
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipStreamSynchronize(s)); // the staging buffer may still feed earlier work
    HIPCHK(st->rho0.alloc(NN));
    HIPCHK(hipMemcpy(st->rho0.p, op, NN * sizeof(double2), hipMemcpyHostToDevice));
    LAUNCHCHK(blfid_dev_launch_state_project(st->K, st->N, st->rho0.p, 1.0 / nrm, st->rho.p, s));
    HIPCHK(hipStreamSynchronize(s));
    return BLFID_OK;
}

blfid_status blfid_state_load(blfid_state *st, const double *rho_all, void *stream) {
    return state_fill(st, rho_all, 1, (hipStream_t)stream);
}

blfid_status blfid_state_fetch(blfid_state *st, double *rho_all, void *stream) {
    if (!st || !rho_all) return fail(BLFID_ERR_INVALID, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    if (st->K > 0)
        HIPCHK(hipMemcpyAsync(rho_all, st->rho.p,
                              (size_t)st->K * st->N * st->N * sizeof(double2),
                              hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BLFID_OK;
}

blfid_status blfid_state_apply(blfid_state *st, const blfid_system *sys, double wr_hz,
                               double rotor_deg, int nsteps, const double *step_t1,
                               const double *step_dt, int nrf, const blfid_rf *rf, int repeat,
                               void *stream) {
    if (!st || !sys) return fail(BLFID_ERR_INVALID, "NULL argument");
    if (sys->nspins != st->S)
        return fail(BLFID_ERR_INVALID, "the system and the state differ in spin count");
    if (repeat < 0) return fail(BLFID_ERR_INVALID, "negative repeat count");
    if (nsteps < 1 || !step_dt) return fail(BLFID_ERR_INVALID, "a propagator needs steps");
    hipStream_t s = (hipStream_t)stream;
    blfid_plan *ppr = nullptr;
    blfid_status cs = create_pulse_plan(sys, wr_hz, rotor_deg, nsteps, step_t1, step_dt, 1, 0.0,
                                        nrf, rf, BLFID_FP64, &ppr);
    if (cs != BLFID_OK) return cs;
    std::unique_ptr<blfid_plan> pp(ppr);
    if (repeat == 0 || st->K == 0) return BLFID_OK;
    const int N = st->N, K = st->K;
    std::vector<int> sblk(N), sloc(N);
    for (size_t b = 0; b < pp->blocks.size(); b++)
        for (size_t a = 0; a < pp->blocks[b].states.size(); a++) {
            sblk[pp->blocks[b].states[a]] = (int)b;
            sloc[pp->blocks[b].states[a]] = (int)a;
        }
    HIPCHK(hipStreamSynchronize(s)); // tables and working storage may still feed earlier work
    HIPCHK(st->sblk.upload(sblk));
    HIPCHK(st->sloc.upload(sloc));
    // working storage: the pulse plan's propagator regions for Kb crystallites (its exp scratch
    // cut to 4096 workgroup slots: a propagation is a small job), then one N x N product slot
    // per crystallite of the batch (at most 4096 slots)
    blfid_layout::Shape ps = prep_shape_of(pp.get());
    ps.exp_slots = std::min(ps.exp_slots, 4096);
    const size_t slot_bytes = (size_t)N * N * sizeof(double2);
    auto need = [&](long kb) {
        return blfid_layout::align256(blfid_layout::layout_bytes(ps, kb)) +
               (size_t)std::min<long>(kb, 4096) * slot_bytes;
    };
    double avail = st->budget;
    if (avail <= 0) {
        size_t fre = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fre, &tot));
        avail = 0.4 * (double)(fre + st->ws.n);
    }
    if ((double)need(1) > avail)
        return fail(BLFID_ERR_NOMEM, "one crystallite's propagator does not fit the memory budget");
    long lo = 1, hi = K;
    while (lo < hi) {
        const long mid = lo + (hi - lo + 1) / 2;
        if ((double)need(mid) <= avail) lo = mid;
        else hi = mid - 1;
    }
    const long Kb = lo;
    HIPCHK(st->ws.alloc(need(Kb)));
    size_t poff[blfid_layout::R_COUNT];
    const size_t pbytes = blfid_layout::align256(blfid_layout::layout_bytes(ps, Kb, poff));
    carve(pp.get(), st->ws.p, poff);
    refresh_devplan(pp.get());
    pp->dp.RP = st->d_RP.p;
    const DevPlan *PP = &pp->dp;
    double2 *T = (double2 *)(st->ws.p + pbytes);
    const int nslots = (int)std::min<long>(Kb, 4096);
    const size_t pbm2 = (size_t)pp->bmax * pp->bmax;
    for (int c0 = 0; c0 < K; c0 += (int)Kb) {
        const int kb = (int)std::min<long>(Kb, K - c0);
        for (auto &kv : pp->blk_groups)
            LAUNCHCHK(blfid_dev_launch_exp(0, kv.first, PP, kb, c0, pp->d_bgrp[kv.first].p,
                                           (int)kv.second.size(), pp->scrE, pbm2, ps.exp_slots,
                                           pp->bufU, s));
        for (auto &kv : pp->blk_groups)
            LAUNCHCHK(blfid_dev_launch_dyson(0, kv.first, PP, kb, pp->d_bgrp[kv.first].p,
                                             (int)kv.second.size(), pp->bufU, pp->bufA, s));
        LAUNCHCHK(blfid_dev_launch_state_apply(
            PP, kb, c0, (const double2 *)(pp->bufA + (size_t)((pp->n - 1) % 2) * pp->Lsum * 16),
            st->sblk.p, st->sloc.p, N, repeat, st->rho.p, T, nslots, s));
    }
    // the pulse plan's tables feed the kernels above: they finish before it is released
    HIPCHK(hipStreamSynchronize(s));
    return BLFID_OK;
}

blfid_status blfid_state_trace(blfid_state *st, const double *det_d, int normalize,
                               double *out2, void *stream) {
    if (!st || !det_d || !out2) return fail(BLFID_ERR_INVALID, "NULL argument");
    const int N = st->N, K = st->K;
    const cplx *det = (const cplx *)det_d;
    std::vector<int> nzi;
    std::vector<double> nzv;
    cplx trdd = 0;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
            const cplx v = det[i + (size_t)j * N];
            trdd += v * std::conj(v);
            if (v != 0.0) {
                nzi.push_back(i);
                nzi.push_back(j);
                nzv.push_back(v.real());
                nzv.push_back(v.imag());
            }
        }
    if (normalize && trdd == 0.0) return fail(BLFID_ERR_INVALID, "Tr(det det^H) is zero");
    out2[0] = out2[1] = 0.0;
    if (K == 0) return BLFID_OK;
    hipStream_t s = (hipStream_t)stream;
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(st->nzi.upload(nzi));
    HIPCHK(st->nzv.upload(nzv));
    const int G = blfid_layout::CRYST_GROUP;
    const int ng = (int)blfid_layout::ngroups(K);
    HIPCHK(st->part.alloc((size_t)2 * K));
    HIPCHK(st->grp.alloc((size_t)ng * 4));
    HIPCHK(st->out.alloc(2));
    const cplx norm = normalize ? trdd : cplx(1.0, 0.0);
    LAUNCHCHK(blfid_dev_launch_zero(st->grp.p, (size_t)ng * 4, s));
    LAUNCHCHK(blfid_dev_launch_state_trace(K, N, st->rho.p, st->nzi.p, st->nzv.p,
                                           (int)(nzv.size() / 2), st->part.p, s));
    LAUNCHCHK(blfid_dev_launch_reduce(K, 0, 1, 1, G, st->d_w.p, 1.0, st->part.p, st->grp.p, s));
    LAUNCHCHK(blfid_dev_launch_finish(1, ng, st->grp.p, norm.real(), norm.imag(), st->out.p, s));
    HIPCHK(hipMemcpyAsync(out2, st->out.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return BLFID_OK;
}

blfid_status blfid_plan_set_state(blfid_plan *pl, blfid_state *st) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    pl->gen++; // a recorded graph no longer matches the plan
    if (st && st->N != pl->N)
        return fail(BLFID_ERR_INVALID, "the state and the plan differ in Hilbert dimension");
    if (pl->last_stream) (void)hipStreamSynchronize(pl->last_stream);
    if (st) {
        delete pl->prep;
        pl->prep = nullptr;
    }
    pl->state = st;
    pl->have_ops = false; // the block-pair tables depend on it: operators are set again
    return BLFID_OK;
}

// ------------------------------------------------------------------------------ HIP graphs
blfid_status blfid_plan_graph_capture(blfid_plan *pl, int normalize) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    if (!pl->have_ops) return fail(BLFID_ERR_INVALID, "operators not set");
    if (normalize && pl->trdd == 0.0) return fail(BLFID_ERR_INVALID, "Tr(det det^H) is zero");
    if (pl->prof)
        return fail(BLFID_ERR_UNSUPPORTED, "profiling is on: a graph does not record phase events");
    if (pl->ws)
        return fail(BLFID_ERR_UNSUPPORTED, "graph capture needs the plan's own working storage");
    if (pl->state && pl->state->K != pl->K)
        return fail(BLFID_ERR_INVALID, "the state and the plan's powder differ in size");
    if (pl->last_stream) HIPCHK(hipStreamSynchronize(pl->last_stream));
    blfid_layout::Shape sh, psh;
    full_shape(pl, sh, psh);
    double avail = pl->budget;
    if (avail <= 0) {
        size_t fre = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fre, &tot));
        avail = 0.4 * (double)(fre + pl->own_ws.n);
    }
    const long Kb = blfid_layout::choose_batch(sh, pl->K, avail);
    if (Kb < 1) return fail(BLFID_ERR_NOMEM, "one crystallite does not fit the memory budget");
    size_t off[blfid_layout::R_COUNT];
    const size_t need = blfid_layout::layout_bytes(sh, Kb, off);
    HIPCHK(pl->own_ws.alloc(need));
    HIPCHK(pl->fid_out.alloc((size_t)2 * pl->npts));
    if (pl->gexec) (void)hipGraphExecDestroy(pl->gexec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    pl->gexec = nullptr;
    pl->graph = nullptr;
    if (!pl->cap_stream) HIPCHK(hipStreamCreateWithFlags(&pl->cap_stream, hipStreamNonBlocking));
    // the recorded run forms the powder rotations itself; an eager run still forms them too
    const bool rpd = pl->rp_dirty;
    HIPCHK(hipStreamBeginCapture(pl->cap_stream, hipStreamCaptureModeThreadLocal));
    blfid_status rs = enqueue_run(pl, sh, pl->own_ws.p, off, Kb, normalize, pl->fid_out.p,
                                  pl->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(pl->cap_stream, &g);
    pl->rp_dirty = rpd;
    if (rs != BLFID_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (rs != BLFID_OK) return rs;
        return fail(BLFID_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(BLFID_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    pl->graph = g;
    pl->gexec = x;
    pl->g_gen = pl->gen;
    pl->g_base = pl->own_ws.p;
    pl->g_out = pl->fid_out.p;
    return BLFID_OK;
}

blfid_status blfid_plan_graph_launch(blfid_plan *pl, void *stream) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    if (!pl->gexec || pl->g_gen != pl->gen || pl->own_ws.p != pl->g_base ||
        pl->fid_out.p != pl->g_out)
        return fail(BLFID_ERR_INVALID, "no graph recorded for the plan as it is now: capture again");
    hipStream_t st = (hipStream_t)stream;
    HIPCHK(hipGraphLaunch(pl->gexec, st));
    pl->last_stream = st;
    pl->ran = true;
    return BLFID_OK;
}

// ------------------------------------------------------------------------------ slices
blfid_status blfid_plan_run_slice(blfid_plan *pl, int c_begin, int c_end, int normalize,
                                  void *stream) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    if (!pl->have_ops) return fail(BLFID_ERR_INVALID, "operators not set");
    if (pl->ws)
        return fail(BLFID_ERR_UNSUPPORTED, "slices need the plan's own working storage");
    if (c_begin < 0 || c_end > pl->K || c_begin >= c_end)
        return fail(BLFID_ERR_INVALID, "bad crystallite range");
    hipStream_t st = (hipStream_t)stream;
    if (c_begin == 0) {
        if (normalize && pl->trdd == 0.0) return fail(BLFID_ERR_INVALID, "Tr(det det^H) is zero");
        blfid_layout::Shape sh, psh;
        full_shape(pl, sh, psh);
        double avail = pl->budget;
        if (avail <= 0) {
            size_t fre = 0, tot = 0;
            HIPCHK(hipMemGetInfo(&fre, &tot));
            avail = 0.4 * (double)(fre + pl->own_ws.n);
        }
        const long Kb = blfid_layout::choose_batch(sh, pl->K, avail);
        if (Kb < 1) return fail(BLFID_ERR_NOMEM, "one crystallite does not fit the memory budget");
        HIPCHK(pl->own_ws.alloc(blfid_layout::layout_bytes(sh, Kb)));
        HIPCHK(pl->fid_out.alloc((size_t)2 * pl->npts));
        pl->sl_Kb = Kb;
        pl->sl_gen = pl->gen;
        pl->sl_norm = normalize;
        pl->sl_base = pl->own_ws.p;
    } else if (c_begin != pl->sl_next || pl->sl_gen != pl->gen || pl->own_ws.p != pl->sl_base) {
        return fail(BLFID_ERR_INVALID,
                    "slices must follow each other from crystallite 0, with the plan unchanged");
    }
    blfid_layout::Shape sh, psh;
    full_shape(pl, sh, psh);
    size_t off[blfid_layout::R_COUNT];
    (void)blfid_layout::layout_bytes(sh, pl->sl_Kb, off);
    pl->last_stream = st;
    pl->sl_next = -1; // a failure leaves no slice to continue
    blfid_status rs = enqueue_range(pl, sh, pl->own_ws.p, off, pl->sl_Kb, pl->sl_norm,
                                    pl->fid_out.p, st, c_begin, c_end, c_begin == 0,
                                    c_end == pl->K);
    if (rs != BLFID_OK) return rs;
    pl->sl_next = c_end == pl->K ? -1 : c_end;
    if (c_end == pl->K) pl->ran = true;
    return BLFID_OK;
}

// ------------------------------------------------------------------------------ no-alloc run
// Working storage of blfid_plan_run_noalloc: one statically sized device array per GPU (the
// code object's global, allocated when the library's device code is loaded on that GPU and
// held until the process ends).  Host code never passes `scratch` itself to a kernel (on the
// host it only names the symbol): the device address comes from hipGetSymbolAddress, once per
// device, with that device current.
static __device__ __align__(256) char scratch[BLFID_NOALLOC_SCRATCH_BYTES];
#define NOALLOC_MAX_DEVICES 64
static std::mutex g_noalloc_mu;
static char *g_noalloc_ptr[NOALLOC_MAX_DEVICES];

static blfid_status noalloc_scratch(char **out) {
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(BLFID_ERR_DEVICE, "device index beyond the no-alloc scratch table");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(BLFID_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") +
                                              hipGetErrorString(e));
        g_noalloc_ptr[dev] = (char *)p;
    }
    *out = g_noalloc_ptr[dev];
    return BLFID_OK;
}

// Scratch-slot counts of one layout cut to an eighth of the static scratch each.
static void noalloc_cap(blfid_layout::Shape &sh, const blfid_plan *pl) {
    const double part = (double)BLFID_NOALLOC_SCRATCH_BYTES / 8.0;
    const double bm2 = (double)pl->bmax * pl->bmax;
    const double rs = pl->prec ? 4.0 : 8.0;
    const double e1 = sh.kind == 1 ? bm2 * 8.0 : 9.0 * bm2 * rs; // one exp / static-H slot
    auto cap = [&](int slots, double per) {
        return std::max(1, std::min(slots, (int)std::floor(part / per)));
    };
    sh.exp_slots = cap(sh.exp_slots, e1);
    if (sh.qr_slots > 0) sh.qr_slots = cap(sh.qr_slots, 3.0 * bm2 * 16.0);
    if (sh.eig_slots > 0) sh.eig_slots = cap(sh.eig_slots, bm2 * 16.0);
}

// The plan's layout (and its preparation's, in ps) with the scratch-slot counts cut so that
// each slot region takes at most an eighth of the static scratch (fewer workgroups in flight
// for the propagator, Q/R and large-block eigen kernels; results are unchanged, every task
// still runs).
static void noalloc_shape(const blfid_plan *pl, blfid_layout::Shape &sh, blfid_layout::Shape &ps) {
    full_shape(pl, sh, ps);
    noalloc_cap(sh, pl);
    if (pl->prep) noalloc_cap(ps, pl->prep);
}

blfid_status blfid_plan_run_noalloc(blfid_plan *pl, int normalize, double *fid_out_dev,
                                    void *stream) {
    if (!pl || !fid_out_dev) return fail(BLFID_ERR_INVALID, "plan or output is NULL");
    if (!pl->have_ops) return fail(BLFID_ERR_INVALID, "operators not set");
    if (normalize && pl->trdd == 0.0) return fail(BLFID_ERR_INVALID, "Tr(det det^H) is zero");
    char *base = nullptr;
    blfid_status s = noalloc_scratch(&base);
    if (s != BLFID_OK) return s;
    blfid_layout::Shape sh, psh;
    noalloc_shape(pl, sh, psh);
    const long Kb = blfid_layout::choose_batch(sh, pl->K, (double)BLFID_NOALLOC_SCRATCH_BYTES);
    if (Kb < 1)
        return fail(BLFID_ERR_NOMEM,
                    "one crystallite of this plan needs " +
                        std::to_string(blfid_layout::layout_bytes(sh, 1)) +
                        " bytes of working storage, beyond the no-alloc scratch of " +
                        std::to_string((size_t)BLFID_NOALLOC_SCRATCH_BYTES) + " bytes");
    size_t off[blfid_layout::R_COUNT];
    (void)blfid_layout::layout_bytes(sh, Kb, off);
    hipStream_t st = (hipStream_t)stream;
    pl->last_stream = st;
    // profiling records HIP events per launch: not done here
    const bool prof = pl->prof;
    pl->prof = false;
    s = enqueue_run(pl, sh, base, off, Kb, normalize, fid_out_dev, st);
    pl->prof = prof;
    return s;
}

blfid_status blfid_workspace_create(double bytes, blfid_workspace **out) {
    if (!out || bytes <= 0) return fail(BLFID_ERR_INVALID, "bad workspace arguments");
    *out = nullptr;
    blfid_workspace *w = new blfid_workspace();
    w->bytes = (size_t)bytes;
    hipError_t e = hipMalloc((void **)&w->base, w->bytes);
    if (e != hipSuccess) {
        delete w;
        return fail(e == hipErrorOutOfMemory ? BLFID_ERR_NOMEM : BLFID_ERR_DEVICE,
                    std::string("workspace hipMalloc: ") + hipGetErrorString(e));
    }
    e = hipEventCreateWithFlags(&w->ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipFree(w->base);
        delete w;
        return fail(BLFID_ERR_DEVICE, "workspace event");
    }
    *out = w;
    return BLFID_OK;
}

void blfid_workspace_destroy(blfid_workspace *w) {
    if (!w) return;
    (void)hipDeviceSynchronize();
    (void)hipEventDestroy(w->ev);
    (void)hipFree(w->base);
    delete w;
}

blfid_status blfid_plan_set_workspace(blfid_plan *pl, blfid_workspace *w) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    pl->gen++; // a recorded graph no longer matches the plan
    pl->ws = w;
    if (w) pl->own_ws.release();
    return BLFID_OK;
}

blfid_status blfid_plan_set_profiling(blfid_plan *pl, int on) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    pl->gen++; // a recorded graph no longer matches the plan
    collect_profile(pl);
    pl->prof = on != 0;
    for (int i = 0; i < BLFID_NSLOT; i++) {
        pl->ph_ms[i] = 0;
        pl->ph_n[i] = 0;
        pl->ph_flops[i] = 0;
        pl->ph_bytes[i] = 0;
    }
    if (pl->prof) {
        // the propagator's degree is chosen per task on the device, so its executed flops are
        // counted there (one atomic add per workgroup)
        if (!pl->d_expflops.p) HIPCHK(pl->d_expflops.alloc(16));
        HIPCHK(hipDeviceSynchronize());
        HIPCHK(hipMemset(pl->d_expflops.p, 0, 16 * sizeof(double)));
    }
    return BLFID_OK;
}

blfid_status blfid_plan_phase_times(blfid_plan *pl, double *ms, int *launches, double *flops) {
    if (!pl || !ms || !launches || !flops) return fail(BLFID_ERR_INVALID, "NULL argument");
    collect_profile(pl);
    // device-counted flops: [0..7] the propagator per tile variant (phase 0), [8..15] the fused
    // propagator + Dyson chain per tile variant (phase 8)
    double ef[16] = {0};
    if (pl->prof && pl->d_expflops.p)
        HIPCHK(hipMemcpy(ef, pl->d_expflops.p, sizeof(ef), hipMemcpyDeviceToHost));
    for (int i = 0; i < BLFID_NSLOT; i++) {
        ms[i] = pl->ph_ms[i];
        launches[i] = pl->ph_n[i];
        flops[i] = pl->ph_flops[i] + (i < 8 ? ef[i] : (i >= 64 ? ef[8 + (i - 64)] : 0.0));
    }
    return BLFID_OK;
}

blfid_status blfid_plan_phase_bytes(blfid_plan *pl, double *bytes) {
    if (!pl || !bytes) return fail(BLFID_ERR_INVALID, "NULL argument");
    collect_profile(pl);
    for (int i = 0; i < BLFID_NSLOT; i++) bytes[i] = pl->ph_bytes[i];
    return BLFID_OK;
}

// ------------------------------------------------------------------------------ launch autotuner
static blfid_status check_tuning(const int *v) {
    if (!v) return fail(BLFID_ERR_INVALID, "tuning is NULL");
    for (int M = 0; M < 5; M++)
        if (v[M] != 0 && v[M] != 1) return fail(BLFID_ERR_INVALID, "propagator method is 0 or 1");
    if (v[5] < 0 || v[5] > 2) return fail(BLFID_ERR_INVALID, "correlation variant is 0, 1 or 2");
    for (int f = 6; f < 15; f++)
        if (v[f] < 0 || v[f] > (1 << 20))
            return fail(BLFID_ERR_INVALID, "grid cap outside 0..1048576 workgroups");
    if (v[15] < 0 || v[15] > 15) return fail(BLFID_ERR_INVALID, "Q-product form bits are 0..15");
    return BLFID_OK;
}

blfid_status blfid_tune_get(int *vals) {
    if (!vals) return fail(BLFID_ERR_INVALID, "vals is NULL");
    tune_from_process(vals);
    return BLFID_OK;
}

blfid_status blfid_tune_set(const int *vals) {
    blfid_status s = check_tuning(vals);
    if (s != BLFID_OK) return s;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    std::memcpy(g_tune[tune_device()], vals, sizeof(int) * BLFID_TUNE_LEN);
    return BLFID_OK;
}

blfid_status blfid_plan_tune_get(const blfid_plan *pl, int *vals) {
    if (!pl || !vals) return fail(BLFID_ERR_INVALID, "NULL argument");
    std::memcpy(vals, pl->tn, sizeof(int) * BLFID_TUNE_LEN);
    return BLFID_OK;
}

blfid_status blfid_plan_tune_set(blfid_plan *pl, const int *vals) {
    if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
    blfid_status s = check_tuning(vals);
    if (s != BLFID_OK) return s;
    pl->gen++; // a recorded graph no longer matches the plan
    if (pl->last_stream) HIPCHK(hipStreamSynchronize(pl->last_stream));
    std::memcpy(pl->tn, vals, sizeof(int) * BLFID_TUNE_LEN);
    refresh_devplan(pl);
    return BLFID_OK;
}

// Crystallites a plan is tuned on: at least `kc`, and enough that each of its launches with one
// task per crystallite and block (or block pair) of a tile class -- the fused propagator, the
// step-walking Q products -- has 2048 workgroups, as the full runs' batches do: a smaller
// sample starves those launches and favours the forms with a task per step.
static int tune_sample(const blfid_plan *pl, int kc) {
    size_t m = 0;
    auto least = [&](const std::map<int, std::vector<int>> &g) {
        for (auto &kv : g)
            if (!kv.second.empty()) m = m ? std::min(m, kv.second.size()) : kv.second.size();
    };
    least(pl->blk_groups);
    least(pl->pair_groups);
    const int per = (int)std::max<size_t>(m, 1);
    return std::min(pl->K, std::max(kc, (2048 + per - 1) / per));
}

// Kernel time per profiling slot, summed over the plans, of one candidate tuning: each plan runs
// its tuning sample once.
static blfid_status tune_pass(blfid_plan **plans, int nplans, int kc, hipStream_t st,
                              const int *vals, double *slot_ms) {
    for (int i = 0; i < BLFID_NSLOT; i++) slot_ms[i] = 0;
    for (int p = 0; p < nplans; p++) {
        blfid_plan *pl = plans[p];
        std::memcpy(pl->tn, vals, sizeof(int) * BLFID_TUNE_LEN);
        pl->gen++;
        blfid_status s = blfid_plan_set_profiling(pl, 1);
        if (s == BLFID_OK) s = run_upto(pl, 0, st, tune_sample(pl, kc));
        if (s != BLFID_OK) return s;
        collect_profile(pl);
        for (int i = 0; i < BLFID_NSLOT; i++) slot_ms[i] += pl->ph_ms[i];
    }
    return BLFID_OK;
}

blfid_status blfid_autotune(blfid_plan **plans, int nplans, int crystallites, void *stream,
                            double *report) {
    if (!plans || nplans < 1) return fail(BLFID_ERR_INVALID, "no plans");
    if (crystallites < 1) return fail(BLFID_ERR_INVALID, "crystallites must be positive");
    for (int p = 0; p < nplans; p++) {
        const blfid_plan *pl = plans[p];
        if (!pl) return fail(BLFID_ERR_INVALID, "plan is NULL");
        if (!pl->spinning || pl->ptop)
            return fail(BLFID_ERR_UNSUPPORTED, "the autotuner takes gamma-COMPUTE plans");
        if (!pl->have_ops) return fail(BLFID_ERR_INVALID, "operators not set");
        if (pl->K < 1) return fail(BLFID_ERR_INVALID, "empty powder");
    }
    hipStream_t st = (hipStream_t)stream;
    const double keep = 0.97; // a candidate replaces the current choice when 3% faster
    int cur[BLFID_TUNE_LEN];
    std::memcpy(cur, plans[0]->tn, sizeof(cur));
    std::vector<int> orig((size_t)nplans * BLFID_TUNE_LEN);
    std::vector<char> prof0(nplans);
    for (int p = 0; p < nplans; p++) {
        std::memcpy(&orig[(size_t)p * BLFID_TUNE_LEN], plans[p]->tn, sizeof(cur));
        prof0[p] = plans[p]->prof;
    }
    std::vector<double> T(BLFID_TUNE_PASSES * BLFID_NSLOT, 0.0);
    auto slot = [&](int pass, int i) -> double & { return T[(size_t)pass * BLFID_NSLOT + i]; };
    blfid_status s = BLFID_OK;
    // passes 0-2: fused (0, 2) or two-kernel (1) propagator per tile class, default (0, 2) or
    // other (1) Q-product form per tile class, correlation variant 0, 1, 2
    for (int c = 0; c < 3 && s == BLFID_OK; c++) {
        int v[BLFID_TUNE_LEN];
        std::memcpy(v, cur, sizeof(v));
        for (int M = 0; M < 5; M++) v[M] = c == 1 ? 1 : 0;
        v[5] = c;
        v[15] = c == 1 ? 15 : 0;
        s = tune_pass(plans, nplans, crystallites, st, v, &slot(c, 0));
    }
    if (s == BLFID_OK) {
        for (int M = 1; M <= 5; M++) {
            const double fused = std::min(slot(0, 64 + M), slot(2, 64 + M));
            // the two-kernel cost of the class: its exp and Dyson slots beyond what the fused
            // pass spent there (blocks above 80 states share class 5's slots)
            const double two = slot(1, M) + slot(1, 8 + M) - slot(0, M) - slot(0, 8 + M);
            if (fused <= 0.0) continue; // no block of this class is fused in these plans
            const double now = cur[M - 1] ? two : fused, alt = cur[M - 1] ? fused : two;
            if (alt < keep * now) cur[M - 1] ^= 1;
        }
        for (int M = 1; M <= 4; M++) {
            const double dflt = std::min(slot(0, 32 + M), slot(2, 32 + M)), other = slot(1, 32 + M);
            if (dflt <= 0.0 || other <= 0.0) continue; // no pair of this class
            const bool on = (cur[15] >> (M - 1)) & 1;
            const double now = on ? other : dflt, alt = on ? dflt : other;
            if (alt < keep * now) cur[15] ^= 1 << (M - 1);
        }
        double cbest = slot(cur[5], 40);
        for (int c = 0; c < 3; c++)
            if (slot(c, 40) < keep * cbest) {
                cbest = slot(c, 40);
                cur[5] = c;
            }
    }
    // passes 3-6: one grid cap for every tuned family; each family then takes its fastest
    const int caps[4] = {1024, 2048, 4096, 8192};
    const int fams[7] = {0, 1, 2, 3, 4, 5, 8};
    for (int j = 0; j < 4 && s == BLFID_OK; j++) {
        int v[BLFID_TUNE_LEN];
        std::memcpy(v, cur, sizeof(v));
        for (int f : fams) v[6 + f] = caps[j];
        s = tune_pass(plans, nplans, crystallites, st, v, &slot(3 + j, 0));
    }
    if (s == BLFID_OK) {
        for (int f : fams) {
            double fam[4];
            for (int j = 0; j < 4; j++) {
                fam[j] = 0;
                for (int M = 0; M < 8; M++) fam[j] += slot(3 + j, f * 8 + M);
            }
            const int c0 = cur[6 + f] ? cur[6 + f] : 4096;
            int jc = 2;
            for (int j = 0; j < 4; j++)
                if (caps[j] == c0) jc = j;
            int jb = jc;
            for (int j = 0; j < 4; j++)
                if (fam[j] < keep * fam[jb]) jb = j;
            cur[6 + f] = caps[jb] == 4096 ? 0 : caps[jb];
        }
    }
    // every plan leaves with the chosen tuning (its own after a failure), its profiling switch
    // as it was and its counters cleared
    for (int p = 0; p < nplans; p++) {
        blfid_plan *pl = plans[p];
        std::memcpy(pl->tn, s == BLFID_OK ? cur : &orig[(size_t)p * BLFID_TUNE_LEN], sizeof(cur));
        pl->gen++;
        if (pl->last_stream) (void)hipStreamSynchronize(pl->last_stream);
        collect_profile(pl);
        const blfid_status s2 = blfid_plan_set_profiling(pl, prof0[p]);
        if (s == BLFID_OK) s = s2;
        refresh_devplan(pl);
        pl->ran = false;
    }
    if (s != BLFID_OK) return s;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        std::memcpy(g_tune[tune_device()], cur, sizeof(cur));
    }
    if (report)
        for (int j = 0; j < BLFID_TUNE_PASSES; j++) {
            report[j] = 0;
            for (int i = 0; i < BLFID_NSLOT; i++) report[j] += slot(j, i);
        }
    return BLFID_OK;
}

blfid_status blfid_plan_fetch(blfid_plan *pl, double *out) {
    if (!pl || !out) return fail(BLFID_ERR_INVALID, "NULL argument");
    if (!pl->ran) return fail(BLFID_ERR_INVALID, "plan has not been run");
    HIPCHK(hipStreamSynchronize(pl->last_stream));
    HIPCHK(hipMemcpy(out, pl->fid_out.p, sizeof(double) * 2 * pl->npts, hipMemcpyDeviceToHost));
    return BLFID_OK;
}

const double *blfid_plan_device_fid(const blfid_plan *pl) { return pl ? pl->fid_out.p : nullptr; }

blfid_status blfid_powder_fid(blfid_plan *pl, const double *theta, const double *phi,
                              const double *gamma, const double *weight, int K,
                              const double *rho, const double *det, int normalize,
                              double *fid_out) {
    blfid_status s = blfid_plan_set_operators(pl, rho, det);
    if (s != BLFID_OK) return s;
    s = blfid_plan_set_powder(pl, theta, phi, gamma, weight, K);
    if (s != BLFID_OK) return s;
    s = blfid_plan_run(pl, normalize, nullptr);
    if (s != BLFID_OK) return s;
    return blfid_plan_fetch(pl, fid_out);
}

} // extern "C"
