// blfid_dev.h -- device-side plan tables shared by the host planner and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#define DEV_TERM_STRIDE 10 /* Cxx Cyy Czz Cxy Cxz Cyz  fdiag  fflip  i  j(-1 for CSA) */
#define DEV_MAX_TERMS 128
#define DEV_MAX_SPINS 12
#define DEV_MAX_BLOCK 256 /* largest Hilbert-space block (k_eig pair tables) */
#define DEV_LDS_BLOCK 80  /* blocks up to this size keep k_eig's matrix in LDS */
// Row padding (floats) of the LDS-resident product planes (chain_mfma, SymTile): rows of
// 16 M + LDS_PAD floats, a multiple of four so that every row starts on 16 bytes (8, 12 and 20
// measured no faster on the batched clusters).
#ifndef LDS_PAD
#define LDS_PAD 4
#endif

struct DevPlan {
    int S, N;      // spins, Hilbert dimension
    int n;         // steps per rotor period (1 for the static path)
    int npts;      // FID points
    int nterms;    // anisotropic interaction terms
    int nblocks;   // Hilbert-space blocks
    int npairs;    // coupled (I,J) block pairs of det / rho_eff
    long Lsum;     // sum of b^2 over blocks: per-(crystallite,step) matrix storage
    long Etot;     // sum of b_I*b_J over pairs: correlation elements per crystallite
    double theta;  // 2*pi*tau
    const double *terms;   // nterms x DEV_TERM_STRIDE
    const double *dconst;  // N: isotropic diagonal of H (-iso_csa m_i + J m_i m_j)
    const double *ffconst; // S x S: constant flip-flop coefficients (strong J / 2)
    const int *ff_off;     // S x S + 1: per ordered spin pair, its run in ff_term
    const int *ff_term;    // terms with a flip-flop part, listed under both orders of their pair
    const double *xd_cst;  // N x 7: diagonal of each basis state as (1, w) . xd_cst, w the six
                           // quadratic products of v = R_P^T u (fused kernel's H build)
    const double *xd_fpr;  // S x S x 7: flip-flop coefficient of each ordered pair, likewise
    const double *u3;      // n x 3: row 2 of cartSpin at each step midpoint
    const double *RP;      // K x 9: cartPowder of each crystallite (row-major)
    const int *blk_dim;    // nblocks
    const int *blk_soff;   // nblocks: offset of the block in the state list (sum of b)
    const long *blk_uoff;  // nblocks: offset of the block in Lsum storage
    const int *blk_states; // N: basis-state index of each block-local row
    const int *pair_I, *pair_J;
    const long *pair_eoff; // element offset of each pair (sum of bI*bJ)
    const int *pair_doff;  // start of each pair's det CSR row pointer (bI+1 entries)
    const int *pair_roff;  // start of each pair's rho_eff CSR row pointer (bJ+1 entries)
    const int *d_rowptr, *d_col;
    const double *d_val;   // complex, interleaved
    const int *r_rowptr, *r_col;
    const double *r_val;   // complex, interleaved
    // point-to-point propagation with rf
    int usign;             // step propagator exp(usign * i theta H): -1 gamma-COMPUTE, +1 ptop
    int akeep;             // Dyson products kept per crystallite (n, or 2 when only A_n is used)
    int nrfiso;            // pulsed isotopes
    const int *spin_rf;    // S: pulsed-isotope slot of each spin or -1
    const double *rf;      // n x nrfiso x 3: amp/2, phase, offset of each step
    const double *thetak;  // n: 2 pi tau_k of steps of unequal length, or NULL (all theta)
    int rt_from_q;         // 1: R(s,r) = Q_rev + conj(Q(r,s)), R is neither formed nor stored
    const int *elem_rev;   // Etot: element (s,r) of the reverse pair (J,I), or -1
    double *expflops;      // 16: executed flops per tile variant of the propagator [0..7] and
                           // the fused propagator + Dyson chain [8..15] (profiling), or NULL
    int bf16x3;            // complex64 products as three bf16 matrix-core products (BLFID_C64_BF16X3)
    long Estr;             // Q / R row stride per step: Etot padded to a multiple of 16 elements
    int nparts;            // FID partial sums per crystallite (correlation segments or block pairs)
    int rt_rev;            // rt_from_q and some element has a reverse element (elem_rev >= 0)
    const int *xd_off;     // nblocks + 1: start of each block's flip-flop positions in xd_list
    const int *xd_list;    // per block of <= DEV_LDS_BLOCK states: (r LDK + q) << 8 | (i S + j)
    const int *elem_r;     // Etot: state-list position of each pair element's column (block I)
    const int *elem_s;     // Etot: state-list position of each pair element's row (block J)
    const void *rho_c;     // prepared rho per crystallite of the batch on the block pairs (storage
                           // precision, conj(rho_JI) at element r + s b_I), or NULL
    int corr_auto_chunks;  // leading correlation chunks whose elements have no reverse element
                           // (rt_from_q plans): they take the autocorrelation kernel
    // launch choices of the autotuner (0 everywhere: the defaults)
    int tn_cap[9];         // persistent-grid cap (workgroups) per kernel family, indexed like the
                           // profiling phases: exp 0, dyson 1, eig 2, bmul 3, qr 4, corr 5,
                           // reduce 6, static 7, expdyson 8
    int tn_corr;           // k_corr variant: 0 four FID points per thread up to 1024 points
                           // (eight above) at the default occupancy, 1 the same at two waves
                           // per SIMD (no register spill), 2 eight points per thread
    int tn_qr;             // k_qr form per complex64 tile class: bit M-1 set takes the other form
                           // (per-step for classes 1-2, step-walking for 3-4)
};

// Persistent-grid cap of a kernel family: the tuned value or 4096 workgroups; single-wave
// workgroups (tile class 1) take four times as many.
#define GRID_CAP 4096
inline long tuned_cap(const DevPlan *P, int fam, int M) {
    const long c = P->tn_cap[fam] > 0 ? P->tn_cap[fam] : GRID_CAP;
    return M == 1 ? 4 * c : c;
}

// Correlation work split (k_corr), shared by the host planner and the kernel: a crystallite's
// Etot elements are cut into chunks of CH elements (CH = min(CORR_SPAN / n, 256 / ceil(n / 8)))
// and the chunks into segments of up to CORR_SEG_CHUNKS; each segment writes one FID partial sum.
#define CORR_TW 256
#define CORR_SPAN 2048
#define CORR_VTAB 512 /* complex128 entries of the period table, within the second buffer */
#define CORR_PPT 8
#define CORR_SEG_CHUNKS 64
__device__ __host__ inline int corr_seg_chunks(int nchunk) {
    return nchunk < CORR_SEG_CHUNKS ? (nchunk < 1 ? 1 : nchunk) : CORR_SEG_CHUNKS;
}
__device__ __host__ inline int corr_chunk(int n) {
    int npg = (n + CORR_PPT - 1) / CORR_PPT;
    int a = CORR_SPAN / n, b = 256 / npg;
    int ch = a < b ? a : b;
    return ch < 1 ? 1 : ch;
}
__device__ __host__ inline int corr_range_segments(int nchunk) {
    if (nchunk <= 0) return 0;
    const int spc = corr_seg_chunks(nchunk);
    return (nchunk + spc - 1) / spc;
}
// FID partial sums of a correlation: the segments of the autocorrelation chunks [0, A) and of
// the general chunks [A, nchunk), each range segmented on its own
__device__ __host__ inline int corr_segments(int n, long Etot, int auto_chunks = 0) {
    const int CH = corr_chunk(n);
    const int nchunk = (int)((Etot + CH - 1) / CH);
    return corr_range_segments(auto_chunks) + corr_range_segments(nchunk - auto_chunks);
}
__host__ __device__ inline bool corr_use_fft(int n) {
    return n >= 16 && (n & (n - 1)) == 0 && n <= 2 * CORR_TW;
}
