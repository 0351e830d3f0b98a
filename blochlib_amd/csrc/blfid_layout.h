// blfid_layout.h -- byte layout of a plan's device working storage (host-only, no HIP calls).
//
// The working storage of one crystallite batch is one arena carved into regions (fixed scratch
// whose size does not depend on the batch, and per-crystallite arrays scaled by the batch size
// Kb).  Every region starts on a 256-byte boundary.  The batch size is chosen as the largest
// Kb whose exact layout fits the bytes available; the arithmetic is kept here, free of device
// calls, so that it can be checked on a machine without a GPU.
#pragma once
#include <stddef.h>

#include <algorithm>

namespace blfid_layout {

enum Region {
    R_U = 0,     // U_k, then B_k = A_k W (gamma-COMPUTE); H blocks (static path)
    R_A,         // Dyson products A_k (n per crystallite, or 2 alternating slots for ptop)
    R_QT,        // Q_k elements, rows of Estr elements per step
    R_RT,        // R_k elements (only when R is not formed from Q)
    R_EXP,       // propagator scratch (9 b_max^2 reals per workgroup slot)
    R_W,         // period / static eigenvectors, FP64
    R_THETA,     // eigenphases / eigenvalues, FP64, N per crystallite
    R_QR,        // Q/R product and eigenbasis-FID scratch (3 b_max^2 complex128 per slot)
    R_PART,      // FID partial sums: nparts x npts complex128 per crystallite
    R_EIG,       // eigen-solver scratch of blocks beyond the LDS classes
    R_WR,        // W in the storage precision (complex64 plans)
    R_GROUP,     // powder-sum accumulators: per group of CRYST_GROUP crystallites, sum + comp
    R_RHO,       // prepared density operator of each crystallite: its coupled block pairs
    R_PREP,      // working storage of the preparation (pulse) propagators
    R_COUNT
};

// Crystallites per powder-sum group: each group is Kahan-summed in crystallite order by one
// thread per FID component, and the groups are then summed in order (k_reduce / k_finish).
constexpr int CRYST_GROUP = 64;

struct Shape {
    int prec = 0;       // 1: complex64 storage of U, A, B, Q, R
    int kind = 0;       // 0 gamma-COMPUTE (spinning), 1 static, 2 pulse program (ptop)
    int n = 1;          // steps per period
    int npts = 1;       // FID points
    int N = 1;          // Hilbert dimension
    int bmax = 1;       // largest block
    int rt_from_q = 0;  // R formed from Q (no Rt storage)
    int nparts = 1;     // FID partial sums per crystallite
    int exp_slots = 0, qr_slots = 0, eig_slots = 0;
    long Lsum = 0;      // sum of b^2 over blocks
    long Estr = 0;      // Q row stride per step (elements, padded)
    long K = 0;         // crystallites of the powder (sizes the group accumulators)
    long rho_elems = 0; // prepared-rho elements per crystallite (pulses before the FID), or 0
    const Shape *prep = nullptr; // layout of the preparation propagators, or none
};

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

inline size_t ngroups(long K) { return (size_t)((K + CRYST_GROUP - 1) / CRYST_GROUP); }

// Region sizes (bytes) for a batch of Kb crystallites.
inline void region_sizes(const Shape &s, long Kb, size_t sz[R_COUNT]) {
    const size_t L = (size_t)s.Lsum, bm2 = (size_t)s.bmax * s.bmax, kb = (size_t)Kb;
    const size_t c16 = 16, d8 = 8;
    const size_t cs = s.prec ? 8 : 16, rs = s.prec ? 4 : 8; // storage complex / real
    for (int i = 0; i < R_COUNT; i++) sz[i] = 0;
    if (s.kind == 2) {
        sz[R_U] = kb * s.n * L * cs;
        sz[R_A] = kb * 2 * L * cs;
        sz[R_EXP] = (size_t)s.exp_slots * 9 * bm2 * rs;
    } else if (s.kind == 0) {
        sz[R_U] = kb * s.n * L * cs;
        sz[R_A] = kb * s.n * L * cs;
        sz[R_QT] = kb * s.n * (size_t)s.Estr * cs;
        sz[R_RT] = s.rt_from_q ? 0 : kb * s.n * (size_t)s.Estr * cs;
        sz[R_EXP] = (size_t)s.exp_slots * 9 * bm2 * rs;
        if (s.prec) sz[R_WR] = kb * L * cs;
    } else {
        sz[R_U] = kb * L * c16;
        sz[R_EXP] = (size_t)s.exp_slots * bm2 * d8;
    }
    sz[R_W] = kb * L * c16;
    sz[R_THETA] = kb * s.N * d8;
    sz[R_QR] = (size_t)s.qr_slots * 3 * bm2 * c16;
    sz[R_PART] = kb * (size_t)s.nparts * 2 * s.npts * d8;
    sz[R_EIG] = (size_t)s.eig_slots * bm2 * c16;
    sz[R_GROUP] = ngroups(s.K) * 2 * (2 * (size_t)s.npts) * d8;
    sz[R_RHO] = kb * (size_t)s.rho_elems * cs;
    if (s.prep) {
        size_t psz[R_COUNT];
        region_sizes(*s.prep, Kb, psz);
        size_t o = 0;
        for (int i = 0; i < R_COUNT; i++) o += align256(psz[i]);
        sz[R_PREP] = o;
    }
}

// Total bytes of the layout for Kb crystallites; off[] receives each region's offset.
inline size_t layout_bytes(const Shape &s, long Kb, size_t off[R_COUNT]) {
    size_t sz[R_COUNT];
    region_sizes(s, Kb, sz);
    size_t o = 0;
    for (int i = 0; i < R_COUNT; i++) {
        off[i] = o;
        o += align256(sz[i]);
    }
    return o;
}

inline size_t layout_bytes(const Shape &s, long Kb) {
    size_t off[R_COUNT];
    return layout_bytes(s, Kb, off);
}

// Largest batch Kb in [1, max(K, 1)] whose layout fits `avail` bytes, or 0 when one crystallite
// does not fit.  layout_bytes is nondecreasing in Kb, so a bisection finds it exactly.
inline long choose_batch(const Shape &s, long K, double avail) {
    if (avail <= 0) return 0;
    const long hi0 = std::max<long>(K, 1);
    if ((double)layout_bytes(s, 1) > avail) return 0;
    long lo = 1, hi = hi0;
    while (lo < hi) {
        const long mid = lo + (hi - lo + 1) / 2;
        if ((double)layout_bytes(s, mid) <= avail) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

} // namespace blfid_layout
