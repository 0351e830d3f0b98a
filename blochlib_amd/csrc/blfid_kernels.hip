// blfid_kernels.hip -- gfx950 kernels of the powder-averaged MAS FID engine.
//
// The work of one spin system is organised by the block structure of its Hamiltonian: every
// SolidSys term (CSA, secular dipolar, J; src/QMspins/sys.cc:749-770) commutes with the
// Zeeman structure that the host finds (connected components of the flip-flop graph), so all
// propagators are block diagonal and every matrix below is one block (dimension b <= N).
//
// Pipeline for a batch of crystallites (host: blfid_host.cpp):
//   k_prep      rotation of each crystallite (cartPowder, space_ten.cc:578-618)
//   k_exp       H_k block at the step midpoint (sys.cc:749-770) and U_k = exp(-i 2 pi tau H_k)
//               as cos X - i sin X by Taylor polynomials in X^2 of adaptive degree (at most 8,
//               Paterson-Stockmeyer) with scaling and squaring, in real arithmetic of the
//               storage precision (FP64, or FP32 for complex64 plans); replaces Mexp's eigen
//               route, matexp.h:239-252 (H is real symmetric)
//   k_dyson     A_k = U_k A_{k-1} along the rotor period (compute.h:257-274)
//   k_eig       U_T = A_n = W diag(lambda) W^H by joint Jacobi on (U+U^H)/2, (U-U^H)/2i
//               (replaces diag(Uf) -> corth/comqr3, compute.h:317)
//   k_bmul      B_k = A_k W
//   k_qr        Q_k = B^H det B and R_k = B^H rho_eff B on the coupled block pairs, with the
//               gamma-COMPUTE twiddles z^-j / z^i folded in (compute.h:378-427)
//   k_corr      F_p = sum_i Qt_{(i+p)%n} Rt_i (cyclic correlation) and
//               fid[t] += sum_rs F_{t%n}(r,s) z_rs^t (compute.h:402-457)
//   k_reduce    FID += w_c * fid_c / (n/ff) [/ Tr(D D^H)], Kahan-compensated, in crystallite
//               order within groups of 64, then the groups in order (k_finish)
//               (oneFID.h:396-407, sequencerun_funcs.cc:698-702)
// and, for wr == 0 or rotor == 0, the static FID (oneFID.h:292-336):
//   k_static_h  H(0, dt) block;  k_eig (Hermitian case);  k_static_fid.
//
// All matrices are column-major, complex values are double2 (re, im).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "blfid_dev.h"

typedef double2 cd;
typedef float2 cf;

// ------------------------------------------------------------------------------ helpers
// Complex helpers are templates over double2 / float2 (members x = re, y = im).
__device__ __forceinline__ cd cmk(double a, double b) { return make_double2(a, b); }
template <typename V> __device__ __forceinline__ V cadd(V a, V b) {
    V r; r.x = a.x + b.x; r.y = a.y + b.y; return r;
}
template <typename V> __device__ __forceinline__ V csub(V a, V b) {
    V r; r.x = a.x - b.x; r.y = a.y - b.y; return r;
}
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
    V r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename V> __device__ __forceinline__ V cconj(V a) { V r; r.x = a.x; r.y = -a.y; return r; }
template <typename V, typename S> __device__ __forceinline__ V cscale(V a, S s) {
    V r; r.x = a.x * s; r.y = a.y * s; return r;
}
template <typename V> __device__ __forceinline__ void cfma(V &c, V a, V b) {
    c.x = fma(a.x, b.x, c.x);
    c.x = fma(-a.y, b.y, c.x);
    c.y = fma(a.x, b.y, c.y);
    c.y = fma(a.y, b.x, c.y);
}
__device__ __forceinline__ cd cexpi(double ph) {
    double s, c;
    sincos(ph, &s, &c);
    return cmk(c, s);
}

// Storage precision R of the propagator pipeline: complex type and conversions to / from FP64.
template <typename R> struct CX;
template <> struct CX<double> {
    typedef cd T;
    static __device__ __forceinline__ cd mk(double a, double b) { return make_double2(a, b); }
    static __device__ __forceinline__ cd from(cd a) { return a; }
};
template <> struct CX<float> {
    typedef cf T;
    static __device__ __forceinline__ cf mk(float a, float b) { return make_float2(a, b); }
    static __device__ __forceinline__ cf from(cd a) { return make_float2((float)a.x, (float)a.y); }
};
__device__ __forceinline__ cd to_cd(cd a) { return a; }
__device__ __forceinline__ cd to_cd(cf a) { return make_double2((double)a.x, (double)a.y); }

template <typename T> struct Arith;
template <> struct Arith<double> {
    static __device__ __forceinline__ double zero() { return 0.0; }
    static __device__ __forceinline__ double cj(double a) { return a; }
    static __device__ __forceinline__ void mac(double &c, double a, double b) { c = fma(a, b, c); }
    static __device__ __forceinline__ double add(double a, double b) { return a + b; }
};
template <> struct Arith<float> {
    static __device__ __forceinline__ float zero() { return 0.0f; }
    static __device__ __forceinline__ float cj(float a) { return a; }
    static __device__ __forceinline__ void mac(float &c, float a, float b) { c = fmaf(a, b, c); }
    static __device__ __forceinline__ float add(float a, float b) { return a + b; }
};
template <> struct Arith<cd> {
    static __device__ __forceinline__ cd zero() { return cmk(0.0, 0.0); }
    static __device__ __forceinline__ cd cj(cd a) { return cconj(a); }
    static __device__ __forceinline__ void mac(cd &c, cd a, cd b) { cfma(c, a, b); }
    static __device__ __forceinline__ cd add(cd a, cd b) { return cadd(a, b); }
};
template <> struct Arith<cf> {
    static __device__ __forceinline__ cf zero() { return make_float2(0.0f, 0.0f); }
    static __device__ __forceinline__ cf cj(cf a) { return cconj(a); }
    static __device__ __forceinline__ void mac(cf &c, cf a, cf b) { cfma(c, a, b); }
    static __device__ __forceinline__ cf add(cf a, cf b) { return cadd(a, b); }
};

// ---------------------------------------------------------------- workgroup GEMM
// C[m x n] = op(A)[m x k] * op(B)[k x n] (+ diag * I); op = N or conjugate transpose.
// NT threads (256 or 64) form a 16 x (NT/16) grid; output tiles are 16M x 16M, thread (tx,ty)
// owns rows tx + 16 i and columns ty + (NT/16) j.  K is walked in slices of 16 staged through
// LDS with zero fill (unpadded sizes work); the next slice is prefetched into registers while
// the current one is consumed.  C must not alias A or B.  `smem` holds GemmLds<M, T>::ELEMS
// elements: two slices of 16 rows, each row padded from 16M to S elements so that the transposing
// stores (16 consecutive lanes on 16 consecutive k, one column each) fall in distinct banks.
template <int M, typename T> struct GemmLds {
    static constexpr int S = 16 * M + 16 / (int)sizeof(T);
    static constexpr int ELEMS = 2 * 16 * S;
};
template <int M, int NT, typename T>
__device__ __forceinline__ void gemm_load_slice(int m, int n, int k, int i0, int j0, int k0,
                                                const T *__restrict__ A, int lda, bool ca,
                                                const T *__restrict__ B, int ldb, bool cb,
                                                T (&ra)[16 * M * 16 / NT], T (&rb)[16 * M * 16 / NT]) {
    constexpr int TM = 16 * M, KS = 16, LPT = TM * KS / NT;
    const int tid = threadIdx.x;
#pragma unroll
    for (int l = 0; l < LPT; l++) {
        int idx = tid + NT * l;
        int r, q;
        if (!ca) { r = idx % TM; q = idx / TM; }
        else { q = idx % KS; r = idx / KS; }
        int gr = i0 + r, gq = k0 + q;
        T v = Arith<T>::zero();
        if (gr < m && gq < k)
            v = ca ? Arith<T>::cj(A[gq + (size_t)gr * lda]) : A[gr + (size_t)gq * lda];
        ra[l] = v;
    }
#pragma unroll
    for (int l = 0; l < LPT; l++) {
        int idx = tid + NT * l;
        int c, q;
        if (!cb) { q = idx % KS; c = idx / KS; }
        else { c = idx % TM; q = idx / TM; }
        int gc = j0 + c, gq = k0 + q;
        T v = Arith<T>::zero();
        if (gc < n && gq < k)
            v = cb ? Arith<T>::cj(B[gc + (size_t)gq * ldb]) : B[gq + (size_t)gc * ldb];
        rb[l] = v;
    }
}

template <int M, int NT, typename T>
__device__ __forceinline__ void wg_gemm(int m, int n, int k, const T *__restrict__ A, int lda, bool ca,
                        const T *__restrict__ B, int ldb, bool cb, T *__restrict__ C, int ldc,
                        T diag, T *smem) {
    constexpr int TM = 16 * M, KS = 16, TY = NT / 16, MJ = TM / TY, LPT = TM * KS / NT;
    constexpr int S = GemmLds<M, T>::S;
    // register prefetch of the next K slice unless the accumulators already fill the VGPRs
    // (complex64 with 80 x 80 tiles would spill with it; complex128 with 64 x 64 tiles would
    // drop to one wave per SIMD)
    constexpr bool PF = (M * MJ * sizeof(T)) < 256 && !(sizeof(T) == 8 && alignof(T) == 8 && M >= 5 &&
                                                         !__is_same(T, double));
    T *As = smem;
    T *Bs = smem + KS * S;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    for (int i0 = 0; i0 < m; i0 += TM) {
        for (int j0 = 0; j0 < n; j0 += TM) {
            T acc[M][MJ];
#pragma unroll
            for (int i = 0; i < M; i++)
#pragma unroll
                for (int j = 0; j < MJ; j++) acc[i][j] = Arith<T>::zero();
            T ra[PF ? LPT : 1], rb[PF ? LPT : 1];
            if constexpr (PF)
                gemm_load_slice<M, NT, T>(m, n, k, i0, j0, 0, A, lda, ca, B, ldb, cb, ra, rb);
            for (int k0 = 0; k0 < k; k0 += KS) {
                __syncthreads();
                if constexpr (PF) {
#pragma unroll
                    for (int l = 0; l < LPT; l++) {
                        int idx = tid + NT * l;
                        int r, q;
                        if (!ca) { r = idx % TM; q = idx / TM; }
                        else { q = idx % KS; r = idx / KS; }
                        As[q * S + r] = ra[l];
                        int c2, q2;
                        if (!cb) { q2 = idx % KS; c2 = idx / KS; }
                        else { c2 = idx % TM; q2 = idx / TM; }
                        Bs[q2 * S + c2] = rb[l];
                    }
                } else {
                    for (int idx = tid; idx < TM * KS; idx += NT) {
                        int r, q;
                        if (!ca) { r = idx % TM; q = idx / TM; }
                        else { q = idx % KS; r = idx / KS; }
                        int gr = i0 + r, gq = k0 + q;
                        T v = Arith<T>::zero();
                        if (gr < m && gq < k)
                            v = ca ? Arith<T>::cj(A[gq + (size_t)gr * lda]) : A[gr + (size_t)gq * lda];
                        As[q * S + r] = v;
                    }
                    for (int idx = tid; idx < TM * KS; idx += NT) {
                        int c, q;
                        if (!cb) { q = idx % KS; c = idx / KS; }
                        else { c = idx % TM; q = idx / TM; }
                        int gc = j0 + c, gq = k0 + q;
                        T v = Arith<T>::zero();
                        if (gc < n && gq < k)
                            v = cb ? Arith<T>::cj(B[gc + (size_t)gq * ldb]) : B[gq + (size_t)gc * ldb];
                        Bs[q * S + c] = v;
                    }
                }
                __syncthreads();
                if constexpr (PF) {
                    if (k0 + KS < k)
                        gemm_load_slice<M, NT, T>(m, n, k, i0, j0, k0 + KS, A, lda, ca, B, ldb, cb, ra, rb);
                }
#pragma unroll 4
                for (int q = 0; q < KS; q++) {
                    T a[M], b[MJ];
#pragma unroll
                    for (int i = 0; i < M; i++) a[i] = As[q * S + tx + 16 * i];
#pragma unroll
                    for (int j = 0; j < MJ; j++) b[j] = Bs[q * S + ty + TY * j];
#pragma unroll
                    for (int i = 0; i < M; i++)
#pragma unroll
                        for (int j = 0; j < MJ; j++) Arith<T>::mac(acc[i][j], a[i], b[j]);
                }
            }
#pragma unroll
            for (int i = 0; i < M; i++)
#pragma unroll
                for (int j = 0; j < MJ; j++) {
                    int r = i0 + tx + 16 * i, c = j0 + ty + TY * j;
                    if (r < m && c < n) {
                        T v = acc[i][j];
                        if (r == c) v = Arith<T>::add(v, diag);
                        C[r + (size_t)c * ldc] = v;
                    }
                }
        }
    }
    __syncthreads();
}

// Threads per workgroup for a tile variant: 64 for the 16 x 16 tiles of small blocks.
template <int M> struct NThr { static constexpr int v = (M == 1) ? 64 : 256; };

// --------------------------------------------------------- complex64 matrix-core GEMM
// C = op(A) B with op(A) = A or A^H, all column-major, complex64.  Output tiles are 16M x 16M,
// split into 16 x 16 sub-tiles dealt round robin to the NT / 64 waves.  k is staged in chunks
// of KC into LDS with k contiguous (re and im planes: At[i][q] = op(A)(i, q), Bt[j][q] = B(q, j)),
// rows padded to KC + 4 floats so that the float4 fragment reads of 8 consecutive rows fall in
// distinct banks, zero-filled past m, n, k.  Per 16-wide k slice a lane of group
// g = lane / 16 reads At[row][4g .. 4g + 3] and Bt[col][4g .. 4g + 3] (re and im) as float4
// and issues v_mfma_f32_16x16x4f32 16 times, step u supplying k = 4g + u (A and B share that
// k order, so the sum over u and g is the whole slice): Cr += Ar Br - Ai Bi,
// Ci += Ar Bi + Ai Br.  The accumulators of lane l hold rows 4 (l / 16) + j, column l % 16 of
// its sub-tile.  `smem` holds 4 * 16 M * (KC + 4) floats.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));
// bfloat16 split of four floats: hi = bf16(x) (round to nearest even), lo = bf16(x - hi), so
// x = hi + lo to 16 significand bits.
__device__ __forceinline__ unsigned short f2bf(float x) {
    unsigned int u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ void bf_split(v4f x, v4s &hi, v4s &lo) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const unsigned short h = f2bf(x[u]);
        hi[u] = (short)h;
        lo[u] = (short)f2bf(x[u] - __uint_as_float((unsigned int)h << 16));
    }
}
template <int NSUB, int M, int NW>
__device__ __forceinline__ void mfma_store(int i0, int j0, int m, int n, int w, int lane,
                                           const v4f (&cr)[NSUB], const v4f (&ci)[NSUB],
                                           cf *C, int ldc) {
    const int g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int s = 0; s < NSUB; s++) {
        const int t = w + NW * s;
        if (t < M * M) {
            const int col = j0 + 16 * (t % M) + r;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = i0 + 16 * (t / M) + 4 * g + j;
                if (row < m && col < n) C[row + (size_t)col * ldc] = make_float2(cr[s][j], ci[s][j]);
            }
        }
    }
}

template <int M, int NT, int KC, bool SPLIT>
__device__ __forceinline__ void wg_cgemm_mfma(int m, int n, int k, const cf *__restrict__ A,
                                              int lda, bool ca, const cf *__restrict__ B,
                                              int ldb, cf *__restrict__ C, int ldc,
                                              float *smem) {
    constexpr int TM = 16 * M, LDK = KC + 4, NW = NT / 64, NSUB = (M * M + NW - 1) / NW;
    float *Ar = smem, *Ai = Ar + TM * LDK, *Br = Ai + TM * LDK, *Bi = Br + TM * LDK;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, r = lane & 15;
    __syncthreads(); // operands the workgroup wrote before the call are visible to every thread
#pragma unroll 1
    for (int i0 = 0; i0 < m; i0 += TM) {
#pragma unroll 1
        for (int j0 = 0; j0 < n; j0 += TM) {
            v4f cr[NSUB], ci[NSUB];
#pragma unroll
            for (int s = 0; s < NSUB; s++) {
                cr[s] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
                ci[s] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
            }
#pragma unroll 1
            for (int q0 = 0; q0 < k; q0 += KC) {
                const int kc = min(KC, k - q0), kp = (kc + 15) & ~15;
                // staging: each thread issues its loads of both operands in groups of G (the
                // first group before the barrier that ends the previous slice's products, so
                // those loads overlap them), then stores the group to LDS; columns q >= kc are
                // zero.  The group sizes (four for tile class 3, two for 5, one otherwise) keep
                // each kernel's registers within its occupancy class (measured per class).
                constexpr int PER = TM * KC / NT, G = M == 3 ? 4 : M == 5 ? 2 : 1;
                static_assert(PER * NT == TM * KC, "staging split");
#pragma unroll 1
                for (int u0 = 0; u0 < PER; u0 += G) {
                    cf va[G], vb[G];
#pragma unroll
                    for (int t = 0; t < G; t++) {
                        const int idx = tid + (u0 + t) * NT;
                        cf v = make_float2(0.0f, 0.0f), wv = make_float2(0.0f, 0.0f);
                        if (u0 + t < PER) {
                            if (ca) {
                                const int q = idx % KC, i = idx / KC;
                                if (i0 + i < m && q < kc) {
                                    v = A[(q0 + q) + (i0 + i) * lda]; // 32-bit offsets
                                    v.y = -v.y;
                                }
                            } else {
                                // A(i, q) is contiguous in i.  Each 16 lanes load 16 consecutive
                                // rows of one column (a 128-byte segment) and a wave covers 4
                                // columns, so the 64 stores i * (KC + 4) + q of a wave spread
                                // over the banks.
                                const int rest = idx >> 6;
                                const int i = (idx & 15) + 16 * (rest % M), q = ((idx >> 4) & 3) + 4 * (rest / M);
                                if (i0 + i < m && q < kc) v = A[(i0 + i) + (q0 + q) * lda];
                            }
                            const int qb = idx % KC, j = idx / KC;
                            if (j0 + j < n && qb < kc) wv = B[(q0 + qb) + (j0 + j) * ldb];
                        }
                        va[t] = v;
                        vb[t] = wv;
                    }
                    if (u0 == 0) __syncthreads();
#pragma unroll
                    for (int t = 0; t < G; t++) {
                        if (u0 + t >= PER) continue;
                        const int idx = tid + (u0 + t) * NT;
                        int i, q;
                        if (ca) {
                            q = idx % KC;
                            i = idx / KC;
                        } else {
                            const int rest = idx >> 6;
                            i = (idx & 15) + 16 * (rest % M);
                            q = ((idx >> 4) & 3) + 4 * (rest / M);
                        }
                        Ar[i * LDK + q] = va[t].x;
                        Ai[i * LDK + q] = va[t].y;
                        const int qb = idx % KC, j = idx / KC;
                        Br[j * LDK + qb] = vb[t].x;
                        Bi[j * LDK + qb] = vb[t].y;
                    }
                }
                __syncthreads();
#pragma unroll 1
                for (int qq = 0; qq < kp; qq += 16) {
#pragma unroll
                    for (int s = 0; s < NSUB; s++) {
                        const int t = w + NW * s;
                        if (t < M * M) {
                            const int oa = (16 * (t / M) + r) * LDK + qq + 4 * g;
                            const int ob = (16 * (t % M) + r) * LDK + qq + 4 * g;
                            const v4f ar = *(const v4f *)&Ar[oa], ai = *(const v4f *)&Ai[oa];
                            const v4f br = *(const v4f *)&Br[ob], bi = *(const v4f *)&Bi[ob];
                            if constexpr (SPLIT) {
                                // BF16x3: one 16x16x16 bf16 product covers the slice; each real
                                // product is hi*hi + hi*lo + lo*hi of the split operands
                                v4s arh, arl, aih, ail, brh, brl, bih, bil;
                                bf_split(ar, arh, arl);
                                bf_split(ai, aih, ail);
                                bf_split(br, brh, brl);
                                bf_split(bi, bih, bil);
                                const v4s nsg = v4s{(short)0x8000, (short)0x8000, (short)0x8000, (short)0x8000};
                                const v4s naih = aih ^ nsg, nail = ail ^ nsg;
                                cr[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh, brh, cr[s], 0, 0, 0);
                                cr[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh, brl, cr[s], 0, 0, 0);
                                cr[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl, brh, cr[s], 0, 0, 0);
                                cr[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(naih, bih, cr[s], 0, 0, 0);
                                cr[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(naih, bil, cr[s], 0, 0, 0);
                                cr[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(nail, bih, cr[s], 0, 0, 0);
                                ci[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh, bih, ci[s], 0, 0, 0);
                                ci[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arh, bil, ci[s], 0, 0, 0);
                                ci[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(arl, bih, ci[s], 0, 0, 0);
                                ci[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih, brh, ci[s], 0, 0, 0);
                                ci[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(aih, brl, ci[s], 0, 0, 0);
                                ci[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ail, brh, ci[s], 0, 0, 0);
                            } else {
                                const v4f na = -ai;
#pragma unroll
                                for (int u = 0; u < 4; u++) {
                                    cr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[u], br[u], cr[s], 0, 0, 0);
                                    ci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[u], bi[u], ci[s], 0, 0, 0);
                                    cr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(na[u], bi[u], cr[s], 0, 0, 0);
                                    ci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[u], br[u], ci[s], 0, 0, 0);
                                }
                            }
                        }
                    }
                }
            }
            mfma_store<NSUB, M, NW>(i0, j0, m, n, w, lane, cr, ci, C, ldc);
        }
    }
    __syncthreads();
}

// Complex products: matrix cores for complex64, the VALU tile GEMM for complex128.
template <int M, int NT, typename R, bool SPLIT = false> struct CGemm {
    typedef typename CX<R>::T CT;
    static constexpr int BYTES = GemmLds<M, CT>::ELEMS * (int)sizeof(CT);
    __device__ __forceinline__ static void run(int m, int n, int k, const CT *A, int lda, bool ca, const CT *B,
                               int ldb, CT *C, int ldc, void *smem) {
        wg_gemm<M, NT, CT>(m, n, k, A, lda, ca, B, ldb, false, C, ldc, Arith<CT>::zero(),
                           (CT *)smem);
    }
};
template <int M, int NT, bool SPLIT> struct CGemm<M, NT, float, SPLIT> {
    static constexpr int KC = (M == 1) ? 16 : 32;
    static constexpr int BYTES = 4 * 16 * M * (KC + 4) * 4;
    __device__ __forceinline__ static void run(int m, int n, int k, const cf *A, int lda, bool ca, const cf *B,
                               int ldb, cf *C, int ldc, void *smem) {
        wg_cgemm_mfma<M, NT, KC, SPLIT>(m, n, k, A, lda, ca, B, ldb, C, ldc, (float *)smem);
    }
};

// Real products of the propagator (both precisions on the VALU tile GEMM: measured faster here
// than matrix cores, the propagator being bound by its staging and synchronisation).
template <int M, int NT, typename R> struct ExpGemm {
    static constexpr int SMEM = GemmLds<M, R>::ELEMS;
    __device__ __forceinline__ static void run(int b, const R *A, const R *B, R *C, R *smem) {
        wg_gemm<M, NT, R>(b, b, b, A, b, false, B, b, false, C, b, (R)0, smem);
    }
};

// ------------------------------------------------------------------------------ k_prep
// cartPowder = rotationMatrix3D(phi, theta, gamma) (space_ten.cc:578-618, set by
// SolidSys::setPowderAngles(th, ph, ga) -> Rotations::setPowderAngles(ph, th, ga)).
__global__ void k_prep(int K, const double *__restrict__ th, const double *__restrict__ ph,
                       const double *__restrict__ ga, double *__restrict__ RP) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    double a = ph[c], b = th[c], g = ga[c];
    double sa, ca, sb, cb, sg, cg;
    sincos(a, &sa, &ca);
    sincos(b, &sb, &cb);
    sincos(g, &sg, &cg);
    double *R = RP + 9 * (size_t)c;
    R[0] = ca * cb * cg - sa * sg; R[1] = sa * cb * cg + ca * sg; R[2] = -sb * cg;
    R[3] = -ca * cb * sg - sa * cg; R[4] = -sa * cb * sg + ca * cg; R[5] = sb * sg;
    R[6] = ca * sb; R[7] = sa * sb; R[8] = cb;
}

// ------------------------------------------------------------------- Hamiltonian block build
// zz of every anisotropic term for lab direction v = R_P^T u, u = row 2 of cartSpin
// (csa.cc:316-319, dip.cc:196-199): zz_t = v^T C_t v, C_t in the crystal frame.
// rfk (nullptr without rf): this step's amp/2, phase, offset per pulsed isotope.  The rf
// enters as (amp/2) on every single flip of a pulsed spin and offset * m_i on the diagonal:
// PulseData::Pulse (pulsedata.cc:110-125) with its phase taken out as the diagonal similarity
// exp(-i phi Fz) (amp (cos phi Ix + sin phi Iy) = e^{-i phi Fz} amp Ix e^{i phi Fz}), which
// k_exp applies to the propagator.
// The infinity norm of the block is returned to every thread, and the block times
// scale_of(norm) is written into H (leading dimension ld), so the propagator's scaling needs no
// separate pass.  Diagonal entries are formed once per row (s_diag); a row's off-diagonal entries are
// its flip-flop partners (bits i, j differ: s_ff[i][j]) and, with rf, its single flips of
// pulsed spins, all inside the block, so the row sums need no pass over H.
// The plan's interaction-term table (nterms x DEV_TERM_STRIDE doubles) copied into the
// workgroup's dynamic LDS once per launch: the Hamiltonian build reads it for every row and
// flip-flop pair of every task.
extern __shared__ double dyn_terms[];
template <int NT>
__device__ __forceinline__ const double *stage_terms(const DevPlan &P) {
    for (int i = threadIdx.x; i < P.nterms * DEV_TERM_STRIDE; i += NT) dyn_terms[i] = P.terms[i];
    __syncthreads();
    return dyn_terms;
}

template <typename R, int NT, typename SF>
__device__ double build_h_block(const DevPlan &P, const double *terms, const double *RPc,
                                const double *u, int blk, R *H, int ld, double *s_zz,
                                double *s_ff, double *s_v, double *s_diag, double *s_red,
                                const double *rfk, SF scale_of) {
    const int tid = threadIdx.x;
    const int S = P.S;
    if (tid == 0) {
        for (int a = 0; a < 3; a++)
            s_v[a] = RPc[0 * 3 + a] * u[0] + RPc[1 * 3 + a] * u[1] + RPc[2 * 3 + a] * u[2];
    }
    __syncthreads();
    const double vx = s_v[0], vy = s_v[1], vz = s_v[2];
    for (int t = tid; t < P.nterms; t += NT) {
        const double *c = terms + DEV_TERM_STRIDE * t;
        s_zz[t] = c[0] * vx * vx + c[1] * vy * vy + c[2] * vz * vz +
                  2.0 * (c[3] * vx * vy + c[4] * vx * vz + c[5] * vy * vz);
    }
    __syncthreads();
    // flip-flop coefficient of each ordered spin pair: constant part + its anisotropic terms (the
    // pair's list, in term order)
    for (int q = tid; q < S * S; q += NT) {
        double f = P.ffconst[q];
        for (int z = P.ff_off[q]; z < P.ff_off[q + 1]; z++) {
            const int t = P.ff_term[z];
            f += s_zz[t] * terms[DEV_TERM_STRIDE * t + 7];
        }
        s_ff[q] = f;
    }
    const int b = P.blk_dim[blk];
    const int *st = P.blk_states + P.blk_soff[blk];
    // diagonal: each row's term sum split over tpr consecutive lanes (terms dealt round
    // robin), combined by lane shuffles; every thread runs the same number of passes
    int tpr = 1;
    while (tpr < 16 && tpr * 2 * b <= NT) tpr *= 2;
    for (int r0 = 0; r0 < b; r0 += NT / tpr) {
        const int r = r0 + tid / tpr, part = tid % tpr;
        double h = 0.0;
        if (r < b) {
            const int ga = st[r];
            for (int t = part; t < P.nterms; t += tpr) {
                const double *c = terms + DEV_TERM_STRIDE * t;
                int i = (int)c[8], j = (int)c[9];
                double mi = ((ga >> (S - 1 - i)) & 1) ? -0.5 : 0.5;
                double f = c[6] * mi;
                if (j >= 0) f *= ((ga >> (S - 1 - j)) & 1) ? -0.5 : 0.5;
                h += s_zz[t] * f;
            }
            if (part == 0) {
                h += P.dconst[ga];
                if (rfk) {
                    for (int i = 0; i < S; i++) {
                        int q = P.spin_rf[i];
                        if (q >= 0) h += rfk[3 * q + 2] * (((ga >> (S - 1 - i)) & 1) ? -0.5 : 0.5);
                    }
                }
            }
        }
        for (int off = tpr / 2; off > 0; off >>= 1) h += __shfl_xor(h, off);
        if (r < b && part == 0) s_diag[r] = h;
    }
    __syncthreads();
    double rmax = 0.0;
    for (int r = tid; r < b; r += NT) {
        const int ga = st[r];
        double sm = fabs(s_diag[r]);
        for (int i = 0; i < S; i++) {
            const int bi = (ga >> (S - 1 - i)) & 1;
            for (int j = i + 1; j < S; j++)
                if (bi != ((ga >> (S - 1 - j)) & 1)) sm += fabs(s_ff[i * S + j]);
            if (rfk) {
                int q = P.spin_rf[i];
                if (q >= 0) sm += fabs(rfk[3 * q]);
            }
        }
        rmax = fmax(rmax, sm);
    }
    // infinity norm: wavefront maxima, then across the workgroup's waves
    for (int off = 32; off > 0; off >>= 1) rmax = fmax(rmax, __shfl_xor(rmax, off));
    if ((tid & 63) == 0) s_red[tid >> 6] = rmax;
    __syncthreads();
    double nrm = s_red[0];
    for (int w = 1; w < NT / 64; w++) nrm = fmax(nrm, s_red[w]);
    // H is written already multiplied by the caller's factor for this norm
    const double hs = scale_of(nrm);
    for (int e = tid; e < b * b; e += NT) {
        int ra = e % b, cb_ = e / b;
        double h = 0.0;
        if (ra == cb_) {
            h = s_diag[ra];
        } else {
            int ga = st[ra], gb = st[cb_];
            int x = ga ^ gb;
            int pc = __popc(x);
            if (pc == 2 && (ga & x) != 0 && (ga & x) != x) {
                int p1 = 31 - __clz(x);
                int p0 = __ffs(x) - 1;
                int i = S - 1 - p1, j = S - 1 - p0;
                h = s_ff[i * S + j];
            } else if (pc == 1 && rfk) {
                int q = P.spin_rf[S - __ffs(x)];
                if (q >= 0) h = rfk[3 * q];
            }
        }
        H[ra + (size_t)cb_ * ld] = (R)(h * hs);
    }
    __syncthreads();
    return nrm;
}

// ------------------------------------------------------------------------------ k_exp
// One task = (crystallite c, step k, block) of a block-size group.  U = exp(-i theta H),
// theta = 2 pi tau, as cos(X) - i sin(X) with X = theta H / 2^s, followed by s squarings
// (C, S) -> (C^2 - S^2, 2 C S).  cos X and sin(X)/X are Taylor polynomials of degree m in
// Y = X^2 (coefficients (-1)^j/(2j)! and (-1)^j/(2j+1)!), m chosen per task from
// r = ||X||_inf as the smallest one whose remainder bound r^(2m+2)/(2m+2)! is below the
// storage precision's tolerance (1e-17 for FP64, 3e-9 for FP32), and evaluated
// Paterson-Stockmeyer style with the fewest products:
//   m <= 4      powers Y..Y^m, polynomial formed elementwise         m + 1 GEMMs
//   m in 5..6   Y, Y^2, Y^3;  p = q0(Y) + Y^3 q1(Y)                      6 GEMMs
//   m in 7..8   Y..Y^4;       p = q0(Y) + Y^4 q1(Y)                      7 GEMMs
// (each count includes S = X sinc).  s is the least with r / 2^s <= 0.8 (FP64) or 2 (FP32),
// where m = 8 still meets the tolerance.  Nine b x b scratch matrices per workgroup slot.
// The output is C + i usign S, i.e. exp(usign i theta H): usign = -1 for gamma-COMPUTE
// (Mexp(H, -i 2 pi tau), compute.h:265), +1 for the pulse-program propagator (Mexp(hh,
// +i tau 2 pi), propogation.cc:273).  With rf the step's phase similarity R = diag(p_r),
// p_r = exp(-i sum_q phi_q M_q(r)) (M_q = Fz of pulsed isotope q), gives U = R U0 R^H.
template <typename R> struct ExpTol;
template <> struct ExpTol<double> { static constexpr double tol = 1e-17, rmax = 0.8; };
// complex64: a Taylor remainder below 2e-8 relative is under half an FP32 ulp of the result
template <> struct ExpTol<float> { static constexpr double tol = 2e-8, rmax = 2.0; };

// The 16 x 16 complex64 variant is held to four waves per SIMD (128 registers, no spills; three
// without the bound).  The FP64 49-64-state variant (like k_dyson's and k_qr's) is held to three
// workgroups per CU: 168 registers instead of 234 with 17 spilled, still 16% faster than two
// (k_dyson 14%, no spills; k_qr 21%, 6 spilled).
template <int M, typename R>
__global__ void __launch_bounds__(NThr<M>::v, (M == 1 && sizeof(R) == 4) ? 4 : (M == 4 && sizeof(R) == 8) ? 3 : 1) k_exp(DevPlan P, int Kb, int c0,
                                                    const int *__restrict__ grp, int ngrp,
                                                    R *__restrict__ scratch, size_t slot,
                                                    typename CX<R>::T *__restrict__ bufU) {
    constexpr int NT = NThr<M>::v;
    typedef typename CX<R>::T CT;
    __shared__ __align__(16) R smem[ExpGemm<M, NT, R>::SMEM];
    __shared__ double s_zz[DEV_MAX_TERMS];
    __shared__ double s_ff[DEV_MAX_SPINS * DEV_MAX_SPINS];
    __shared__ double s_v[4];
    __shared__ double s_red[NT / 64];
    __shared__ double s_diag[DEV_MAX_BLOCK];
    const int tid = threadIdx.x;
    R *X = scratch + slot * 9 * blockIdx.x;
    R *Y = X + slot, *Y2 = Y + slot, *Y3 = Y2 + slot, *Y4 = Y3 + slot;
    R *Hc = Y4 + slot, *Hs = Hc + slot, *Cm = Hs + slot, *Ps = Cm + slot;
    const long ntask = (long)Kb * P.n * ngrp;
    // 1/(2j)! and 1/(2j+1)! with alternating signs, j = 0..8
    const R cc[9] = {(R)1.0, (R)(-1.0 / 2), (R)(1.0 / 24), (R)(-1.0 / 720), (R)(1.0 / 40320),
                     (R)(-1.0 / 3628800), (R)(1.0 / 479001600), (R)(-1.0 / 87178291200.0),
                     (R)(1.0 / 20922789888000.0)};
    const R ss[9] = {(R)1.0, (R)(-1.0 / 6), (R)(1.0 / 120), (R)(-1.0 / 5040), (R)(1.0 / 362880),
                     (R)(-1.0 / 39916800), (R)(1.0 / 6227020800.0), (R)(-1.0 / 1307674368000.0),
                     (R)(1.0 / 355687428096000.0)};
    const double *terms = stage_terms<NT>(P);
    double nflops = 0.0; // executed GEMM flops of this workgroup (counted when profiling)
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int g = (int)(task % ngrp);
        const long ck = task / ngrp;
        const int k = (int)(ck % P.n), c = (int)(ck / P.n);
        const int blk = grp[g];
        const int b = P.blk_dim[blk];
        const int bb = b * b;
        const double theta = P.thetak ? P.thetak[k] : P.theta;
        const double *rfk = P.nrfiso > 0 ? P.rf + (size_t)3 * P.nrfiso * k : nullptr;
        // X = theta H / 2^sq, the squaring count chosen from the norm bound of theta H
        int sq = 0;
        const double nrm =
            build_h_block<R, NT>(P, terms, P.RP + 9 * (size_t)(c0 + c), P.u3 + 3 * k, blk, X, b, s_zz,
                                 s_ff, s_v, s_diag, s_red, rfk, [&](double hn) {
                                     while (hn * theta / (double)(1 << sq) > ExpTol<R>::rmax && sq < 30) sq++;
                                     return theta / (double)(1 << sq);
                                 }) * theta;
        const double rx = nrm / (double)(1 << sq);
        // smallest degree m in Y with rx^(2m+2)/(2m+2)! <= tol (uniform across the workgroup)
        int m = 1;
        {
            double term = rx * rx * rx * rx / 24.0; // m = 1: rx^4/4!
            while (m < 8 && term > ExpTol<R>::tol) {
                m++;
                term *= rx * rx / ((2.0 * m + 1.0) * (2.0 * m + 2.0));
            }
            if (m == 5) m = 6;
            if (m == 7) m = 8;
        }
        // GEMMs of the Taylor part (S = X sinc included) and three per squaring
        nflops += 2.0 * b * b * b * ((m == 6 ? 6 : m == 8 ? 7 : m + 1) + 3 * sq);
        // powers Y = X X, Y^2 = Y Y, Y^3 = Y^2 Y, Y^4 = Y^2 Y^2 as far as m needs them (one GEMM
        // call site per loop keeps the kernel small enough for its tiles to stay in registers)
        const int npow = (m == 6) ? 3 : (m < 4 ? m : 4);
#pragma unroll 1
        for (int i = 0; i < npow; i++) {
            const R *A = (i == 0) ? X : (i == 1) ? Y : Y2;
            const R *B = (i == 0) ? X : (i == 3) ? Y2 : Y;
            R *C = (i == 0) ? Y : (i == 1) ? Y2 : (i == 2) ? Y3 : Y4;
            ExpGemm<M, NT, R>::run(b, A, B, C, smem);
        }
        if (m <= 4) {
            for (int e = tid; e < bb; e += NT) {
                R d = (e % b == e / b) ? (R)1 : (R)0;
                R y = Y[e];
                R cv = cc[0] * d + cc[1] * y, sv = ss[0] * d + ss[1] * y;
                if (m >= 2) { R y2 = Y2[e]; cv += cc[2] * y2; sv += ss[2] * y2; }
                if (m >= 3) { R y3 = Y3[e]; cv += cc[3] * y3; sv += ss[3] * y3; }
                if (m >= 4) { R y4 = Y4[e]; cv += cc[4] * y4; sv += ss[4] * y4; }
                Cm[e] = cv;
                Ps[e] = sv;
            }
            __syncthreads();
        } else {
            // p = q0(Y) + Yh q1(Y): Yh = Y^3 (m = 6) or Y^4 (m = 8); the two cases are
            // spelled out so the coefficient tables are only indexed by constants
            const bool h4 = (m == 8);
            const R *Yh = h4 ? Y4 : Y3;
            for (int e = tid; e < bb; e += NT) {
                R d = (e % b == e / b) ? (R)1 : (R)0;
                R y = Y[e], y2 = Y2[e], y3 = Y3[e];
                R cv, sv;
                if (h4) {
                    R y4 = Y4[e];
                    cv = cc[4] * d + cc[5] * y + cc[6] * y2 + cc[7] * y3 + cc[8] * y4;
                    sv = ss[4] * d + ss[5] * y + ss[6] * y2 + ss[7] * y3 + ss[8] * y4;
                } else {
                    cv = cc[3] * d + cc[4] * y + cc[5] * y2 + cc[6] * y3;
                    sv = ss[3] * d + ss[4] * y + ss[5] * y2 + ss[6] * y3;
                }
                Hc[e] = cv;
                Hs[e] = sv;
            }
            __syncthreads();
#pragma unroll 1
            for (int i = 0; i < 2; i++)
                ExpGemm<M, NT, R>::run(b, Yh, i ? Hs : Hc, i ? Ps : Cm, smem);
            for (int e = tid; e < bb; e += NT) {
                R d = (e % b == e / b) ? (R)1 : (R)0;
                R y = Y[e], y2 = Y2[e];
                R cv = cc[0] * d + cc[1] * y + cc[2] * y2;
                R sv = ss[0] * d + ss[1] * y + ss[2] * y2;
                if (h4) { R y3 = Y3[e]; cv += cc[3] * y3; sv += ss[3] * y3; }
                Cm[e] += cv;
                Ps[e] += sv;
            }
            __syncthreads();
        }
        R *Sm = Y4;
        ExpGemm<M, NT, R>::run(b, X, Ps, Sm, smem);
        for (int it = 0; it < sq; it++) {
#pragma unroll 1
            for (int i = 0; i < 3; i++) {
                const R *A = (i == 1) ? Sm : Cm;
                const R *B = (i == 0) ? Cm : Sm;
                R *C = (i == 0) ? Y : (i == 1) ? Y2 : Y3;
                ExpGemm<M, NT, R>::run(b, A, B, C, smem);
            }
            for (int e = tid; e < bb; e += NT) {
                Cm[e] = Y[e] - Y2[e];
                Sm[e] = (R)2 * Y3[e];
            }
            __syncthreads();
        }
        CT *U = bufU + ((size_t)c * P.n + k) * P.Lsum + P.blk_uoff[blk];
        const double us = (double)P.usign;
        if (rfk) {
            // the GEMM staging area is free here: it holds the b phase factors (b <= 16 M)
            cd *s_ph = (cd *)smem;
            const int *st = P.blk_states + P.blk_soff[blk];
            for (int r = tid; r < b; r += NT) {
                double phi = 0.0;
                for (int i = 0; i < P.S; i++) {
                    int q = P.spin_rf[i];
                    if (q >= 0) phi += rfk[3 * q + 1] * (((st[r] >> (P.S - 1 - i)) & 1) ? -0.5 : 0.5);
                }
                s_ph[r] = cexpi(-phi);
            }
            __syncthreads();
            for (int e = tid; e < bb; e += NT)
                U[e] = CX<R>::from(cmul(cmul(s_ph[e % b], cmk((double)Cm[e], us * (double)Sm[e])),
                                        cconj(s_ph[e / b])));
        } else {
            for (int e = tid; e < bb; e += NT) U[e] = CX<R>::mk(Cm[e], (R)us * Sm[e]);
        }
        __syncthreads();
    }
    if (P.expflops && tid == 0 && nflops > 0.0) atomicAdd(&P.expflops[M], nflops);
}

// ------------------------------------------------------------------------------ k_dyson
// Chains of products of one block of at most 16 M <= 80 states on matrix cores (complex64), with
// the whole K dimension in one LDS slice and the B operand resident in LDS (Bt[j][q] = B(q, j)):
//   DYSON:  A_0 = U_0, A_k = U_k A_{k-1}, k = 1..n-1 -- the running product is written back into
//           the B operand straight from the accumulators;
//   !DYSON: C_k = X_k W, k = 0..n-1 -- W stays the B operand for every k.
// Each step then loads only its A operand (U_k or X_k), prefetched into registers one step
// ahead.  Operand layout, fragments and products as in wg_cgemm_mfma.  X_k = X0 + k xstride,
// output k at C0 + (k % ckeep) cstride.  `smem` holds 4 * 16 M * (16 M + LDS_PAD) floats.
template <int M, int NT, bool DYSON>
__device__ __forceinline__ void chain_mfma(int b, int n, const cf *__restrict__ X0, size_t xstride,
                                           const cf *__restrict__ W, cf *__restrict__ C0,
                                           size_t cstride, int ckeep, float *smem) {
    constexpr int TM = 16 * M, LDK = TM + LDS_PAD, NW = NT / 64, NSUB = (M * M + NW - 1) / NW;
    // operand elements staged per thread (the last round partial when NT does not divide TM^2)
    constexpr int PER = (TM * TM + NT - 1) / NT;
    constexpr bool EVEN = PER * NT == TM * TM;
    float *Ar = smem, *Ai = Ar + TM * LDK, *Br = Ai + TM * LDK, *Bi = Br + TM * LDK;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, r = lane & 15;
    const int kp = (b + 15) & ~15;
    // B operand (zero-filled past b): U_0, also written out as A_0, or W
    const cf *B0 = DYSON ? X0 : W;
    for (int idx = tid; idx < TM * TM; idx += NT) {
        const int q = idx % TM, j = idx / TM;
        cf v = make_float2(0.0f, 0.0f);
        if (q < b && j < b) {
            v = B0[q + j * b];
            if (DYSON) C0[q + j * b] = v;
        }
        Br[j * LDK + q] = v.x;
        Bi[j * LDK + q] = v.y;
    }
    // element u of this thread's share of the A operand: row i, column q (16 lanes on 16
    // consecutive rows of one column, as the transposing staging of wg_cgemm_mfma)
    auto at_i = [&](int u) { const int idx = tid + u * NT; return (idx & 15) + 16 * ((idx >> 6) % M); };
    auto at_q = [&](int u) { const int idx = tid + u * NT; return ((idx >> 4) & 3) + 4 * ((idx >> 6) / M); };
    cf pre[PER];
    auto load_x = [&](int k) {
        const cf *Xk = X0 + (size_t)k * xstride;
#pragma unroll
        for (int u = 0; u < PER; u++) {
            const int i = at_i(u), q = at_q(u);
            const bool in = EVEN || tid + u * NT < TM * TM;
            pre[u] = (in && i < b && q < b) ? Xk[i + q * b] : make_float2(0.0f, 0.0f);
        }
    };
    const int k0 = DYSON ? 1 : 0;
    if (k0 < n) load_x(k0);
    for (int k = k0; k < n; k++) {
#pragma unroll
        for (int u = 0; u < PER; u++) {
            if (!EVEN && tid + u * NT >= TM * TM) continue;
            const int i = at_i(u), q = at_q(u);
            Ar[i * LDK + q] = pre[u].x;
            Ai[i * LDK + q] = pre[u].y;
        }
        __syncthreads(); // A operand staged; B operand complete
        if (k + 1 < n) load_x(k + 1);
        v4f cr[NSUB], ci[NSUB];
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            cr[s] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
            ci[s] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        }
        for (int qq = 0; qq < kp; qq += 16) {
#pragma unroll
            for (int s = 0; s < NSUB; s++) {
                const int t = w + NW * s;
                if (t < M * M) {
                    const int oa = (16 * (t / M) + r) * LDK + qq + 4 * g;
                    const int ob = (16 * (t % M) + r) * LDK + qq + 4 * g;
                    const v4f ar = *(const v4f *)&Ar[oa], ai = *(const v4f *)&Ai[oa];
                    const v4f br = *(const v4f *)&Br[ob], bi = *(const v4f *)&Bi[ob];
                    const v4f na = -ai;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        cr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[u], br[u], cr[s], 0, 0, 0);
                        ci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[u], bi[u], ci[s], 0, 0, 0);
                        cr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(na[u], bi[u], cr[s], 0, 0, 0);
                        ci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[u], br[u], ci[s], 0, 0, 0);
                    }
                }
            }
        }
        __syncthreads(); // every operand read done
        cf *Ck = C0 + (size_t)(k % ckeep) * cstride;
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            const int t = w + NW * s;
            if (t < M * M) {
                const int col = 16 * (t % M) + r, row0 = 16 * (t / M) + 4 * g;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (row0 + j < b && col < b)
                        Ck[row0 + j + col * b] = make_float2(cr[s][j], ci[s][j]);
                if (DYSON) {
                    *(v4f *)&Br[col * LDK + row0] = cr[s];
                    *(v4f *)&Bi[col * LDK + row0] = ci[s];
                }
            }
        }
    }
    __syncthreads();
}

// One task = (crystallite, block) chain: A_1 = U_1, A_{k+1} = U_{k+1} A_k (compute.h:267;
// U *= X is U = X U in propogation.cc:273, 288).  Product k goes to slot k % akeep of the
// crystallite's akeep slots: all of them for gamma-COMPUTE, two alternating ones when only
// the period propagator A_n is needed.
template <int M, typename R, bool SPLIT = false>
__global__ void __launch_bounds__(NThr<M>::v, (M == 3 && sizeof(R) == 4) ? 3 : (M == 4 && sizeof(R) == 4) ? 2 : (M == 4 && sizeof(R) == 8) ? 3 : 1) k_dyson(DevPlan P, int Kb, const int *__restrict__ grp,
                                                      int ngrp,
                                                      const typename CX<R>::T *__restrict__ bufU,
                                                      typename CX<R>::T *__restrict__ bufA) {
    constexpr int NT = NThr<M>::v;
    typedef typename CX<R>::T CT;
    constexpr bool CHAIN = std::is_same<R, float>::value && !SPLIT;
    constexpr int CHB = 4 * 16 * M * (16 * M + LDS_PAD) * 4, GB = CGemm<M, NT, R, SPLIT>::BYTES;
    __shared__ __align__(16) char smem[CHAIN && CHB > GB ? CHB : GB];
    const long ntask = (long)Kb * ngrp;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int g = (int)(task % ngrp), c = (int)(task / ngrp);
        const int blk = grp[g], b = P.blk_dim[blk];
        const size_t base = (size_t)c * P.n * P.Lsum + P.blk_uoff[blk];
        const size_t abase = (size_t)c * P.akeep * P.Lsum + P.blk_uoff[blk];
        const CT *U0 = bufU + base;
        CT *A0 = bufA + abase;
        // the chain holds one 16 M x 16 M tile: blocks above 80 states (M = 5) take the tiled path
        if (CHAIN && (M < 5 || b <= 16 * M)) {
            if constexpr (CHAIN)
                chain_mfma<M, NT, true>(b, P.n, U0, P.Lsum, nullptr, A0, P.Lsum, P.akeep,
                                        (float *)smem);
        } else if constexpr (!CHAIN || M == 5) {
            for (int e = threadIdx.x; e < b * b; e += NT) A0[e] = U0[e];
            __syncthreads();
            for (int k = 1; k < P.n; k++) {
                const CT *Uk = bufU + base + (size_t)k * P.Lsum;
                const CT *Ap = bufA + abase + (size_t)((k - 1) % P.akeep) * P.Lsum;
                CT *Ak = bufA + abase + (size_t)(k % P.akeep) * P.Lsum;
                CGemm<M, NT, R, SPLIT>::run(b, b, b, Uk, b, false, Ap, b, Ak, b, smem);
            }
        }
    }
}

// ------------------------------------------------------------------------------ k_expdyson
// LDS-resident real symmetric products on FP32 matrix cores.  A plane holds a 16M x 16M symmetric
// matrix (zero past the block size b) in rows of LDK = 16M + LDS_PAD floats: plane[r LDK + q] =
// X(r, q) = X(q, r), so one layout is both the k-contiguous A operand (rows) and B operand
// (columns) of v_mfma_f32_16x16x4f32 (lane 16 g + r reads the float4 at row r, k = 4g .. 4g + 3
// of each 16-wide k slice, as wg_cgemm_mfma).  Every polynomial in X, and cos X, sin X, is
// symmetric, so the accumulators (rows 4g + j, column r of a sub-tile) are stored transposed as
// float4 runs, plane[col][row .. row + 3], with no exchange.  Sub-tiles t = w + NW s of the
// output are dealt round robin to the NW waves.
template <int M, int NT> struct SymTile {
    static constexpr int TM = 16 * M, LDK = TM + LDS_PAD, NW = NT / 64, NSUB = (M * M + NW - 1) / NW;
    static constexpr int PLANE = TM * LDK; // floats
    // symmetric products of the classes whose M x M tiles do not divide evenly over the waves
    // (3 x 3 over 4, 5 x 5 over 8) compute only the tiles I <= J and store each twice: 6 and 15
    // tiles take two per wave where 9 and 25 took three and four (for the 4 x 4 tiles over 8
    // waves the mirrored stores cost more than the 6 tiles saved: measured 1% slower)
    static constexpr bool UP = (M == 3 || M == 5);
};

// Output tile (I, J) (16 x 16 blocks of the 16M x 16M result) of sub-tile s of wave w: tile
// t = w + NW s of the M x M tiles, or of the M(M+1)/2 tiles I <= J (row by row) when UP.
template <int M, int NT, bool UP>
__device__ __forceinline__ bool sym_tile(int w, int s, int &I, int &J) {
    const int u = w + SymTile<M, NT>::NW * s;
    if constexpr (UP) {
        if (u >= M * (M + 1) / 2) return false;
        int i = 0;
#pragma unroll
        for (int q = 1; q < M; q++)
            if (u >= q * M - q * (q - 1) / 2) i = q;
        I = i;
        J = u - (i * M - i * (i - 1) / 2) + i;
        return true;
    } else {
        if (u >= M * M) return false;
        I = u / M;
        J = u % M;
        return true;
    }
}

template <int NS> __device__ __forceinline__ void sym_zero(v4f (&acc)[NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++) acc[s] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
}

// acc += sa PA PB over k < kp (the wave's sub-tiles; only tiles I <= J when the result is
// symmetric and SymTile::UP).  VIRT: the B operand is formed on the fly from two planes,
// ca I + cb PY + cc PZ (the identity masked to k < b), instead of read from PB.
template <int M, int NT, bool VIRT, int NS>
__device__ __forceinline__ void sym_mfma(const float *PA, const float *PB, const float *PY,
                                         const float *PZ, float sa, float ca, float cb, float cc,
                                         int b, int kp, v4f (&acc)[NS]) {
    typedef SymTile<M, NT> T;
    static_assert(NS == T::NSUB, "sub-tile count");
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, r = lane & 15;
    for (int qq = 0; qq < kp; qq += 16) {
#pragma unroll
        for (int s = 0; s < T::NSUB; s++) {
            int I, J;
            if (sym_tile<M, NT, T::UP>(w, s, I, J)) {
                const int col = 16 * J + r;
                const int oa = (16 * I + r) * T::LDK + qq + 4 * g;
                const int ob = col * T::LDK + qq + 4 * g;
                const v4f av = *(const v4f *)&PA[oa];
                v4f bv;
                if constexpr (VIRT) {
                    const v4f y = *(const v4f *)&PY[ob], z = *(const v4f *)&PZ[ob];
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        const int k = qq + 4 * g + u;
                        bv[u] = cb * y[u] + cc * z[u] + ((k == col && k < b) ? ca : 0.0f);
                    }
                } else {
                    bv = *(const v4f *)&PB[ob];
                }
#pragma unroll
                for (int u = 0; u < 4; u++)
                    acc[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(sa * av[u], bv[u], acc[s], 0, 0, 0);
            }
        }
    }
}

// Store v (a lane's four accumulator values of tile (I, J) of a symmetric result) into the
// plane: transposed as one float4 at block (J, I) and, for I != J, directly at block (I, J).
template <int M, int NT>
__device__ __forceinline__ void sym_put(float *PL, int I, int J, int g, int r, const v4f &v) {
    typedef SymTile<M, NT> T;
    *(v4f *)&PL[(16 * J + r) * T::LDK + 16 * I + 4 * g] = v;
    if (T::UP && I != J) {
#pragma unroll
        for (int j = 0; j < 4; j++) PL[(16 * I + 4 * g + j) * T::LDK + 16 * J + r] = v[j];
    }
}

// plane = sc * acc (symmetric result, full 16M x 16M tile including the zero padding)
template <int M, int NT, int NS>
__device__ __forceinline__ void sym_store(float *PL, const v4f (&acc)[NS], float sc) {
    typedef SymTile<M, NT> T;
    static_assert(NS == T::NSUB, "sub-tile count");
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int s = 0; s < T::NSUB; s++) {
        int I, J;
        if (sym_tile<M, NT, T::UP>(w, s, I, J)) {
            v4f v;
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = sc * acc[s][j];
            sym_put<M, NT>(PL, I, J, g, r, v);
        }
    }
}

// Hamiltonian of one step for the fused kernel, in three phases separated by barriers:
//   A: v = R_P^T u and its six quadratic products w (every thread, in registers); the diagonal
//      of each row, xd_cst[a] . (1, w) for its basis state a, and the flip-flop coefficient of
//      every ordered spin pair, xd_fpr . (1, w) (both tables formed on the host from the same
//      terms: seven FP64 multiply-adds per entry instead of a pass over every term);
//   B: each row's sum of absolute values (diagonal plus every flip-flop partner), the block's
//      infinity norm per wave;
//   C: X = theta H / 2^s rounded once from FP64 -- its diagonal and the block's fixed flip-flop
//      positions (xd_list) -- scattered into the plane, s the least with ||X||_inf <= rmax.
// Every other entry of the plane is zero and stays zero.  Returns ||H||_inf and s.  Same
// Hamiltonian as build_h_block (sys.cc:749-770; csa.cc:289-329, dip.cc:171-207), its sums
// associated differently.
template <int NT>
__device__ double xd_build_h(const DevPlan &P, const double *RPc, const double *u, int blk,
                             float *PX, int LDK, double *s_ff, double *s_diag, double *s_red,
                             double theta, double rmax_x, int &sq) {
    const int tid = threadIdx.x, S = P.S;
    const double vx = RPc[0] * u[0] + RPc[3] * u[1] + RPc[6] * u[2];
    const double vy = RPc[1] * u[0] + RPc[4] * u[1] + RPc[7] * u[2];
    const double vz = RPc[2] * u[0] + RPc[5] * u[1] + RPc[8] * u[2];
    const double w0 = vx * vx, w1 = vy * vy, w2 = vz * vz;
    const double w3 = 2.0 * vx * vy, w4 = 2.0 * vx * vz, w5 = 2.0 * vy * vz;
    auto form = [&](const double *c) {
        return c[0] + c[1] * w0 + c[2] * w1 + c[3] * w2 + c[4] * w3 + c[5] * w4 + c[6] * w5;
    };
    const int b = P.blk_dim[blk];
    const int *st = P.blk_states + P.blk_soff[blk];
    for (int x = tid; x < b + S * S; x += NT) {
        if (x < b) s_diag[x] = form(P.xd_cst + 7 * (size_t)st[x]);
        else s_ff[x - b] = form(P.xd_fpr + 7 * (x - b));
    }
    __syncthreads();
    double rmax = 0.0;
    for (int r = tid; r < b; r += NT) {
        const int ga = st[r];
        double sm = fabs(s_diag[r]);
        for (int i = 0; i < S; i++) {
            const int bi = (ga >> (S - 1 - i)) & 1;
            for (int j = i + 1; j < S; j++)
                if (bi != ((ga >> (S - 1 - j)) & 1)) sm += fabs(s_ff[i * S + j]);
        }
        rmax = fmax(rmax, sm);
    }
    for (int off = 32; off > 0; off >>= 1) rmax = fmax(rmax, __shfl_xor(rmax, off));
    if ((tid & 63) == 0) s_red[tid >> 6] = rmax;
    __syncthreads();
    double nrm = s_red[0];
    for (int w = 1; w < NT / 64; w++) nrm = fmax(nrm, s_red[w]);
    sq = 0;
    while (nrm * theta / (double)(1 << sq) > rmax_x && sq < 30) sq++;
    const double sc = theta / (double)(1 << sq);
    for (int r = tid; r < b; r += NT) PX[r * LDK + r] = (float)(sc * s_diag[r]);
    const int z0 = P.xd_off[blk], z1 = P.xd_off[blk + 1];
    for (int z = z0 + tid; z < z1; z += NT) {
        const int e = P.xd_list[z];
        PX[e >> 8] = (float)(sc * s_ff[e & 255]);
    }
    __syncthreads();
    return nrm;
}

// Fused propagator and Dyson chain of the complex64 gamma-COMPUTE path for blocks of at most
// 16 M <= 80 states: one task = (crystallite c, block) walks the n steps of the rotor period with
// every matrix in LDS (five planes: H, C, S and the running product's re / im parts):
//   X = theta H_k / 2^s, H at the step midpoint (xd_build_h; sys.cc:749-770)
//   cos X, sin X by Taylor polynomials in Y = X^2 of degree m (Paterson-Stockmeyer):
//     m = 1:  C = I - Y/2, P = I - Y/6 from the Y product's accumulators         (1 product)
//     m = 2:  C, P from Y and the Y^2 product's accumulators                    (2 products)
//     m = 4:  C = I - Y/2 + Y^2 (I/24 - Y/720 + Y^2/40320), P likewise with 1/120, -1/5040,
//             1/362880, the bracket formed on the fly as the B operand         (4 products)
//     S = X P                                                                  (+1 product)
//   s squarings (C, S) <- (C^2 - S^2, 2 C S), s the least with ||X||_inf <= 0.75 (where m = 4
//   meets the FP32 tolerance of k_exp), then
//   A_k = (C + i usign S) A_{k-1}   (compute.h:267; usign = -1: Mexp(H, -i 2 pi tau), compute.h:265)
// on v_mfma_f32_16x16x4f32 with A resident in LDS.  Only A_k goes to global memory (the layout
// k_dyson writes: product k at slot k of the crystallite's n slots).  m is chosen per step from
// the norm bound as the smallest whose remainder bound r^(2m+2)/(2m+2)! is below 2e-8.
// Replaces k_exp + k_dyson
// for these blocks (matexp.h:239-252, compute.h:257-274).  Blocks of 49-80 states run eight
// waves (two per SIMD; the LDS allows one workgroup per CU).  Tile class 3 is held to three
// workgroups per CU (167 registers instead of 195, no spills; 23% faster than two) and class 2
// to five (96 registers, 7 of them spilled to scratch; still 11% faster than four).
template <int M> struct XdThreads { static constexpr int v = M >= 4 ? 512 : NThr<M>::v; };
template <int M>
__global__ void __launch_bounds__(XdThreads<M>::v, M == 3 ? 3 : M == 2 ? 5 : 1) k_expdyson(DevPlan P, int Kb, int c0,
                                                              const int *__restrict__ grp,
                                                              int ngrp, cf *__restrict__ bufA) {
    constexpr int NT = XdThreads<M>::v;
    typedef SymTile<M, NT> T;
    constexpr int TM = T::TM, LDK = T::LDK, NSUB = T::NSUB, NW = T::NW;
    __shared__ __align__(16) float pl[5 * T::PLANE];
    float *PX = pl, *PY = PX + T::PLANE, *PZ = PY + T::PLANE, *PBr = PZ + T::PLANE,
          *PBi = PBr + T::PLANE;
    __shared__ double s_ff[DEV_MAX_SPINS * DEV_MAX_SPINS];
    __shared__ double s_red[NT / 64];
    __shared__ double s_diag[DEV_LDS_BLOCK];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, r = lane & 15;
    const float us = (float)P.usign;
    constexpr double RMAX = 0.75, TOL = 2e-8;
    double nflops = 0.0;
    const long ntask = (long)Kb * ngrp;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int gi = (int)(task % ngrp), c = (int)(task / ngrp);
        const int blk = grp[gi], b = P.blk_dim[blk];
        const int kp = (b + 15) & ~15;
        cf *A0 = bufA + (size_t)c * P.n * P.Lsum + P.blk_uoff[blk];
        // the H plane's zero entries (everything but the diagonal and the flip-flop positions,
        // padding included) are set once per task
        for (int e = tid; e < T::PLANE; e += NT) PX[e] = 0.0f;
        __syncthreads();
        for (int k = 0; k < P.n; k++) {
            const double theta = P.theta;
            int sq = 0;
            const double hn = xd_build_h<NT>(P, P.RP + 9 * (size_t)(c0 + c), P.u3 + 3 * k, blk,
                                             PX, LDK, s_ff, s_diag, s_red, theta, RMAX, sq);
            const double rx = hn * theta / (double)(1 << sq), r2 = rx * rx;
            const int m = (r2 * r2 / 24.0 <= TOL) ? 1 : (r2 * r2 * r2 / 720.0 <= TOL) ? 2 : 4;
            int ng = 1 + 3 * sq; // products of this step (Y, the squarings; more below)
            v4f acc[NSUB], acc2[NSUB];
            // Y = X X
            sym_zero(acc);
            sym_mfma<M, NT, false>(PX, PX, nullptr, nullptr, 1.0f, 0.f, 0.f, 0.f, b, kp, acc);
            if (m == 1) {
                // C = I - Y/2 into Y's plane, P = I - Y/6 into Z's (no operand of this product)
#pragma unroll
                for (int s = 0; s < NSUB; s++) {
                    int I, J;
                    if (sym_tile<M, NT, T::UP>(w, s, I, J)) {
                        const int col = 16 * J + r, row0 = 16 * I + 4 * g;
                        v4f cv, pv;
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const float d = (row0 + j == col && col < b) ? 1.0f : 0.0f;
                            cv[j] = d - 0.5f * acc[s][j];
                            pv[j] = d - (float)(1.0 / 6.0) * acc[s][j];
                        }
                        sym_put<M, NT>(PY, I, J, g, r, cv);
                        sym_put<M, NT>(PZ, I, J, g, r, pv);
                    }
                }
                __syncthreads();
            } else {
                sym_store<M, NT>(PY, acc, 1.0f);
                __syncthreads();
                if (m == 2) {
                    // Y^2; then C = I - Y/2 + Y^2/24 into Y's plane, P = I - Y/6 + Y^2/120 into Z's
                    sym_zero(acc);
                    sym_mfma<M, NT, false>(PY, PY, nullptr, nullptr, 1.0f, 0.f, 0.f, 0.f, b, kp, acc);
                    __syncthreads(); // every read of Y done
                    ng += 1;
                } else {
                    // Z = Y Y, then C and P with the bracket formed from Y and Z on the fly
                    sym_zero(acc);
                    sym_mfma<M, NT, false>(PY, PY, nullptr, nullptr, 1.0f, 0.f, 0.f, 0.f, b, kp, acc);
                    sym_store<M, NT>(PZ, acc, 1.0f);
                    __syncthreads();
                    sym_zero(acc);
                    sym_zero(acc2);
                    sym_mfma<M, NT, true>(PZ, nullptr, PY, PZ, 1.0f, (float)(1.0 / 24.0),
                                          (float)(-1.0 / 720.0), (float)(1.0 / 40320.0), b, kp, acc);
                    sym_mfma<M, NT, true>(PZ, nullptr, PY, PZ, 1.0f, (float)(1.0 / 120.0),
                                          (float)(-1.0 / 5040.0), (float)(1.0 / 362880.0), b, kp, acc2);
                    __syncthreads(); // every read of Y and Z done
                    ng += 3;
                }
                const float ac = m == 2 ? (float)(1.0 / 24.0) : 1.0f;
                const float as = m == 2 ? (float)(1.0 / 120.0) : 1.0f;
#pragma unroll
                for (int s = 0; s < NSUB; s++) {
                    int I, J;
                    if (sym_tile<M, NT, T::UP>(w, s, I, J)) {
                        const int col = 16 * J + r, row0 = 16 * I + 4 * g;
                        const int o = col * LDK + row0;
                        const v4f y = *(const v4f *)&PY[o];
                        v4f cv, pv;
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            // (I - Y/2) first: its diagonal stays below 1, where FP32 rounds
                            // twice as finely as for 1 + Z/24 (half the error on the chain)
                            const float d = (row0 + j == col && col < b) ? 1.0f : 0.0f;
                            const float p2 = m == 2 ? acc[s][j] : acc2[s][j];
                            cv[j] = (d - 0.5f * y[j]) + ac * acc[s][j];
                            pv[j] = (d - (float)(1.0 / 6.0) * y[j]) + as * p2;
                        }
                        sym_put<M, NT>(PY, I, J, g, r, cv);
                        sym_put<M, NT>(PZ, I, J, g, r, pv);
                    }
                }
                __syncthreads();
            }
            // S = X P, into P's plane
            sym_zero(acc);
            sym_mfma<M, NT, false>(PX, PZ, nullptr, nullptr, 1.0f, 0.f, 0.f, 0.f, b, kp, acc);
            __syncthreads();
            sym_store<M, NT>(PZ, acc, 1.0f);
            __syncthreads();
            ng += 1;
            // squarings: (C, S) <- (C^2 - S^2, 2 C S)
            for (int it = 0; it < sq; it++) {
                sym_zero(acc);
                sym_zero(acc2);
                sym_mfma<M, NT, false>(PY, PY, nullptr, nullptr, 1.0f, 0.f, 0.f, 0.f, b, kp, acc);
                sym_mfma<M, NT, false>(PZ, PZ, nullptr, nullptr, -1.0f, 0.f, 0.f, 0.f, b, kp, acc);
                sym_mfma<M, NT, false>(PY, PZ, nullptr, nullptr, 1.0f, 0.f, 0.f, 0.f, b, kp, acc2);
                __syncthreads();
                sym_store<M, NT>(PY, acc, 1.0f);
                sym_store<M, NT>(PZ, acc2, 2.0f);
                __syncthreads();
            }
            nflops += 2.0 * b * b * b * ng;
            // A_k = U_k A_{k-1}, U = C + i us S (C in Y's plane, S in Z's)
            cf *Ak = A0 + (size_t)k * P.Lsum;
            if (k == 0) {
                for (int e = tid; e < TM * TM; e += NT) {
                    const int jj = e / TM, qq = e - jj * TM, o = jj * LDK + qq;
                    const float cr = PY[o], ci = us * PZ[o];
                    PBr[o] = cr; // column jj of A_0 (U_0 is symmetric)
                    PBi[o] = ci;
                    if (jj < b && qq < b) Ak[qq + jj * b] = make_float2(cr, ci);
                }
                __syncthreads();
                continue;
            }
            v4f cr[NSUB], ci[NSUB];
            sym_zero(cr);
            sym_zero(ci);
            for (int qq = 0; qq < kp; qq += 16) {
#pragma unroll
                for (int s = 0; s < NSUB; s++) {
                    const int t = w + NW * s;
                    if (t < M * M) {
                        const int oa = (16 * (t / M) + r) * LDK + qq + 4 * g;
                        const int ob = (16 * (t % M) + r) * LDK + qq + 4 * g;
                        const v4f uc = *(const v4f *)&PY[oa], usn = *(const v4f *)&PZ[oa];
                        const v4f br = *(const v4f *)&PBr[ob], bi = *(const v4f *)&PBi[ob];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const float si = us * usn[u];
                            cr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(uc[u], br[u], cr[s], 0, 0, 0);
                            cr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(-si, bi[u], cr[s], 0, 0, 0);
                            ci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(uc[u], bi[u], ci[s], 0, 0, 0);
                            ci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(si, br[u], ci[s], 0, 0, 0);
                        }
                    }
                }
            }
            __syncthreads(); // every read of the running product done
#pragma unroll
            for (int s = 0; s < NSUB; s++) {
                const int t = w + NW * s;
                if (t < M * M) {
                    const int col = 16 * (t % M) + r, row0 = 16 * (t / M) + 4 * g;
                    *(v4f *)&PBr[col * LDK + row0] = cr[s];
                    *(v4f *)&PBi[col * LDK + row0] = ci[s];
                    if (col < b) {
#pragma unroll
                        for (int j = 0; j < 4; j++)
                            if (row0 + j < b) Ak[row0 + j + col * b] = make_float2(cr[s][j], ci[s][j]);
                    }
                }
            }
            nflops += 8.0 * b * b * b;
            __syncthreads();
        }
    }
    if (P.expflops && tid == 0 && nflops > 0.0) atomicAdd(&P.expflops[8 + M], nflops);
}

// ------------------------------------------------------------------------------ k_eig
// Joint Jacobi diagonalisation of a normal matrix Mx = H1 + i H2 (H1, H2 Hermitian,
// commuting): each 2x2 rotation is the Cardoso-Souloumiac angle for the pair (H1, H2)
// (SIAM J. Matrix Anal. Appl. 17 (1996) 161), pairs in round-robin order, b/2 rotations per
// round.  Mx lives in LDS; W accumulates in global memory.
// mode 0: Mx = A_n block (unitary) -> lambda_r = Mx_rr, out phase = arg(lambda_r)
// mode 1: Mx = H block (Hermitian, static path) -> out value = Re(Mx_rr)
#define EIG_MAXB 80
// The block lives in LDS (b <= MB), with LDS and threads sized for that class so that
// workgroups of small blocks share a CU; larger blocks take k_eig_blk.
// The source block may be stored in FP32 (complex64 plans); the solver always runs in FP64 and
// W stays FP64, with a storage-precision copy Wr for the B = A W products when R != double.
template <int MB, int NT, typename R>
__global__ void __launch_bounds__(NT) k_eig(DevPlan P, int Kb, const int *__restrict__ grp,
                                             int ngrp, const typename CX<R>::T *__restrict__ src,
                                             size_t src_stride, int mode, cd *__restrict__ bufW,
                                             typename CX<R>::T *__restrict__ bufWr,
                                             double *__restrict__ outv, cd *__restrict__ gscr,
                                             size_t gslot) {
    typedef typename CX<R>::T CT;
    // convergence: a sweep whose rotations all stay below tol_stop = sqrt(tol_rot) ends the
    // sweeps (Jacobi converges quadratically: the off-diagonal part such a sweep leaves is of
    // order tol_rot, which the confirming sweep of a rotations-below-tol_rot test would only
    // square); pairs below tol_off are skipped.  A complex64 period propagator carries ~1e-7
    // relative error, so its solve stops earlier.
    constexpr double tol_rot = __is_same(R, double) ? 1e-15 : 1e-10;
    constexpr double tol_stop = __is_same(R, double) ? 3.1622776601683795e-8 : 1e-5;
    constexpr double tol_off = __is_same(R, double) ? 1e-17 : 1e-12;
    static_assert(MB > 0 && MB <= EIG_MAXB, "blocks above 80 states take k_eig_blk");
    constexpr int MP = (MB + 1) / 2;
    // complex64 eigenvectors of the two largest LDS classes accumulate in LDS (FP32) when only
    // the storage-precision copy is wanted (the spinning path): the rotations then never wait on
    // global memory.  The FP64 copy W is not written in that case.  (Both classes already run
    // one 1024-thread workgroup per CU, so the extra LDS costs no occupancy; for the 16-48-state
    // classes it does, and holding W in LDS there made k_eig 3% slower on the batched clusters.)
#ifndef EIG_WLDS_MIN
#define EIG_WLDS_MIN 64
#endif
    constexpr bool WLDS = MB >= EIG_WLDS_MIN && __is_same(R, float);
    __shared__ cd Mlds[MB * (MB + 1)];
    __shared__ cf Wl[WLDS ? MB * MB : 1];
    const bool wl = WLDS && bufWr != nullptr;
    __shared__ double rc[MP];
    __shared__ cd rs[MP];
    __shared__ int rp[MP], rq[MP];
    __shared__ int s_any;
    const int tid = threadIdx.x;
    cd *Ms = Mlds;
    const long ntask = (long)Kb * ngrp;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int g = (int)(task % ngrp), c = (int)(task / ngrp);
        const int blk = grp[g], b = P.blk_dim[blk];
        const int ld = b + 1;
        const CT *S0 = src + (size_t)c * src_stride + P.blk_uoff[blk];
        cd *W = bufW + (size_t)c * P.Lsum + P.blk_uoff[blk];
        for (int e = tid; e < b * b; e += NT) {
            int r = e % b, q = e / b;
            Ms[r + q * ld] = to_cd(S0[e]);
            if (wl) Wl[e] = make_float2(r == q ? 1.0f : 0.0f, 0.0f);
            else W[e] = cmk(r == q ? 1.0 : 0.0, 0.0);
        }
        __syncthreads();
        const int bb = (b + 1) & ~1; // even count, index b is a dummy when b is odd
        const int npair = bb / 2;
        for (int sweep = 0; sweep < 40 && b > 1; sweep++) {
            if (tid == 0) s_any = 0;
            __syncthreads();
            for (int round = 0; round < bb - 1; round++) {
                for (int pi = tid; pi < npair; pi += NT) {
                    // circle method: position 0 fixed, positions 1..bb-1 rotate
                    int pa = pi, pb = bb - 1 - pi;
                    int ia = (pa == 0) ? 0 : ((pa - 1 + round) % (bb - 1)) + 1;
                    int ib = (pb == 0) ? 0 : ((pb - 1 + round) % (bb - 1)) + 1;
                    int p = min(ia, ib), q = max(ia, ib);
                    double cth = 1.0;
                    cd sth = cmk(0, 0);
                    if (q < b) {
                        cd mpp = Ms[p + p * ld], mqq = Ms[q + q * ld];
                        cd mpq = Ms[p + q * ld], mqp = Ms[q + p * ld];
                        // H1 = (M + M^H)/2, H2 = (M - M^H)/(2i)
                        cd h1pq = cscale(cadd(mpq, cconj(mqp)), 0.5);
                        cd d2 = csub(mpq, cconj(mqp));           // = 2i H2pq
                        cd h2pq = cmk(0.5 * d2.y, -0.5 * d2.x); // d2 / (2i)
                        double off = fabs(h1pq.x) + fabs(h1pq.y) + fabs(h2pq.x) + fabs(h2pq.y);
                        double scale = fabs(mpp.x) + fabs(mpp.y) + fabs(mqq.x) + fabs(mqq.y) + 1e-300;
                        if (off > tol_off * scale) {
                            double h1[3] = {mpp.x - mqq.x, 2.0 * h1pq.x, 2.0 * h1pq.y};
                            double h2[3] = {mpp.y - mqq.y, 2.0 * h2pq.x, 2.0 * h2pq.y};
                            double ka = h1[0] * h1[0] + h1[1] * h1[1] + h1[2] * h1[2];
                            double kd = h2[0] * h2[0] + h2[1] * h2[1] + h2[2] * h2[2];
                            double kb = h1[0] * h2[0] + h1[1] * h2[1] + h1[2] * h2[2];
                            double hm = 0.5 * (ka + kd);
                            double lam = hm + sqrt(0.25 * (ka - kd) * (ka - kd) + kb * kb);
                            double a1 = lam - kd, b1 = kb, a2 = kb, b2 = lam - ka;
                            double al, be;
                            if (a1 * a1 + b1 * b1 >= a2 * a2 + b2 * b2) { al = a1; be = b1; }
                            else { al = a2; be = b2; }
                            if (al == 0.0 && be == 0.0) { al = 1.0; be = 0.0; }
                            double ux = al * h1[0] + be * h2[0];
                            double uy = al * h1[1] + be * h2[1];
                            double uz = al * h1[2] + be * h2[2];
                            if (ux < 0) { ux = -ux; uy = -uy; uz = -uz; }
                            double r = sqrt(ux * ux + uy * uy + uz * uz);
                            if (r > 0.0) {
                                cth = sqrt((ux + r) / (2.0 * r));
                                double den = sqrt(2.0 * r * (ux + r));
                                sth = cmk(uy / den, -uz / den);
                                if (fabs(sth.x) + fabs(sth.y) > tol_stop) s_any = 1;
                            }
                        }
                    }
                    rc[pi] = cth;
                    rs[pi] = sth;
                    rp[pi] = p;
                    rq[pi] = q;
                }
                __syncthreads();
                // M <- R^H M R and W <- W R in one pass, R = [[c, -conj(s)], [s, c]] on each
                // pair (p, q): every 2 x 2 block (rows of pair i, columns of pair j) is rotated
                // by both of its pairs' rotations; a dummy index (odd b) rotates by identity.
                for (int e = tid; e < npair * npair; e += NT) {
                    const int pi = e % npair, pj = e / npair;
                    const int p = rp[pi], q = rq[pi], p2 = rp[pj], q2 = rq[pj];
                    const bool vq = q < b, vq2 = q2 < b;
                    const double ci = rc[pi], cj = rc[pj];
                    const cd si = rs[pi], sj = rs[pj];
                    cd a = Ms[p + p2 * ld], bb2 = cmk(0, 0), c2 = cmk(0, 0), d = cmk(0, 0);
                    if (vq2) bb2 = Ms[p + q2 * ld];
                    if (vq) c2 = Ms[q + p2 * ld];
                    if (vq && vq2) d = Ms[q + q2 * ld];
                    // rows (pair i)
                    cd a1 = cadd(cscale(a, ci), cmul(cconj(si), c2));
                    cd c1 = csub(cscale(c2, ci), cmul(si, a));
                    cd b1 = cadd(cscale(bb2, ci), cmul(cconj(si), d));
                    cd d1 = csub(cscale(d, ci), cmul(si, bb2));
                    // columns (pair j)
                    Ms[p + p2 * ld] = cadd(cscale(a1, cj), cmul(sj, b1));
                    if (vq2) Ms[p + q2 * ld] = csub(cscale(b1, cj), cmul(cconj(sj), a1));
                    if (vq) Ms[q + p2 * ld] = cadd(cscale(c1, cj), cmul(sj, d1));
                    if (vq && vq2) Ms[q + q2 * ld] = csub(cscale(d1, cj), cmul(cconj(sj), c1));
                }
                // W <- W R: consecutive threads take consecutive rows of one column pair, so
                // the global accesses of a wave are contiguous
                for (int e = tid; e < npair * b; e += NT) {
                    int row = e % b, pi = e / b;
                    int p = rp[pi], q = rq[pi];
                    if (q >= b) continue;
                    double cth = rc[pi];
                    cd sth = rs[pi];
                    if (wl) {
                        const float cf_ = (float)cth;
                        const cf sf = make_float2((float)sth.x, (float)sth.y);
                        const cf wp = Wl[row + p * b], wq = Wl[row + q * b];
                        Wl[row + p * b] = cadd(cscale(wp, cf_), cmul(sf, wq));
                        Wl[row + q * b] = csub(cscale(wq, cf_), cmul(cconj(sf), wp));
                    } else {
                        cd wp = W[row + (size_t)p * b], wq = W[row + (size_t)q * b];
                        W[row + (size_t)p * b] = cadd(cscale(wp, cth), cmul(sth, wq));
                        W[row + (size_t)q * b] = csub(cscale(wq, cth), cmul(cconj(sth), wp));
                    }
                }
                __syncthreads();
            }
            // every wave reads the flag before the next sweep's reset by thread 0
            const int any = s_any;
            __syncthreads();
            if (!any) break;
        }
        double *o = outv + (size_t)c * P.N + P.blk_soff[blk];
        for (int r = tid; r < b; r += NT) {
            cd l = Ms[r + r * ld];
            o[r] = (mode == 0) ? atan2(l.y, l.x) : l.x;
        }
        if (bufWr) {
            CT *Wr = bufWr + (size_t)c * P.Lsum + P.blk_uoff[blk];
            if constexpr (WLDS) {
                if (wl)
                    for (int e = tid; e < b * b; e += NT) Wr[e] = Wl[e];
                else
                    for (int e = tid; e < b * b; e += NT) Wr[e] = CX<R>::from(W[e]);
            } else {
                for (int e = tid; e < b * b; e += NT) Wr[e] = CX<R>::from(W[e]);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------ k_eig_blk
// Joint rotation of the index pair (p, q) of a normal-matrix iterate: the Cardoso-Souloumiac
// angle for H1 = (M + M^H)/2, H2 = (M - M^H)/2i, as k_eig forms it.  Returns false (identity)
// when the pair's off-diagonal part is below tol_off of its diagonal; big: |s| > tol_big.
__device__ __forceinline__ bool jj_rotation(cd mpp, cd mqq, cd mpq, cd mqp, double tol_off,
                                            double tol_big, double &cth, cd &sth, bool &big) {
    cth = 1.0;
    sth = cmk(0, 0);
    big = false;
    const cd h1pq = cscale(cadd(mpq, cconj(mqp)), 0.5);
    const cd d2 = csub(mpq, cconj(mqp));
    const cd h2pq = cmk(0.5 * d2.y, -0.5 * d2.x);
    const double off = fabs(h1pq.x) + fabs(h1pq.y) + fabs(h2pq.x) + fabs(h2pq.y);
    const double scale = fabs(mpp.x) + fabs(mpp.y) + fabs(mqq.x) + fabs(mqq.y) + 1e-300;
    if (!(off > tol_off * scale)) return false;
    const double h1[3] = {mpp.x - mqq.x, 2.0 * h1pq.x, 2.0 * h1pq.y};
    const double h2[3] = {mpp.y - mqq.y, 2.0 * h2pq.x, 2.0 * h2pq.y};
    const double ka = h1[0] * h1[0] + h1[1] * h1[1] + h1[2] * h1[2];
    const double kd = h2[0] * h2[0] + h2[1] * h2[1] + h2[2] * h2[2];
    const double kb = h1[0] * h2[0] + h1[1] * h2[1] + h1[2] * h2[2];
    const double hm = 0.5 * (ka + kd);
    const double lam = hm + sqrt(0.25 * (ka - kd) * (ka - kd) + kb * kb);
    const double a1 = lam - kd, b1 = kb, a2 = kb, b2 = lam - ka;
    double al, be;
    if (a1 * a1 + b1 * b1 >= a2 * a2 + b2 * b2) { al = a1; be = b1; }
    else { al = a2; be = b2; }
    if (al == 0.0 && be == 0.0) { al = 1.0; be = 0.0; }
    double ux = al * h1[0] + be * h2[0];
    double uy = al * h1[1] + be * h2[1];
    double uz = al * h1[2] + be * h2[2];
    if (ux < 0) { ux = -ux; uy = -uy; uz = -uz; }
    const double r = sqrt(ux * ux + uy * uy + uz * uz);
    if (!(r > 0.0)) return false;
    cth = sqrt((ux + r) / (2.0 * r));
    const double den = sqrt(2.0 * r * (ux + r));
    sth = cmk(uy / den, -uz / den);
    big = fabs(sth.x) + fabs(sth.y) > tol_big;
    return true;
}

// Blocks of more than 80 states (b <= 256): block Jacobi with the same rotations.  The b x b
// iterate M lives in the workgroup's global scratch slot (ld = b), W in bufW.  The indices are
// cut into column blocks of BJ_NB (the last one padded with absent indices) and the blocks
// paired by the circle method.  In each round every pair (P, Q) of the round -- up to 16 at
// once -- copies its 2 BJ_NB x 2 BJ_NB principal submatrix into LDS and runs one joint-Jacobi
// sweep on it there (15 rounds of 8 rotations), accumulating its rotation V; then
// M <- V^H M on the pair's rows (every pair), and M <- M V, W <- W V on its columns.  Sweeps
// repeat until no rotation of a sweep exceeds sqrt(tol_rot) (as in k_eig).  A round reads and
// writes M twice and W once, for b / BJ_NB - 1 rounds per sweep where the scalar rotations take
// b - 1 passes over M and W.  mode as k_eig.
#define BJ_NB 8
#define BJ_NS (2 * BJ_NB)
#define BJ_MAXPAIR 16 /* b <= 16 BJ_MAXPAIR */
template <typename R>
__global__ void __launch_bounds__(512) k_eig_blk(DevPlan P, int Kb, const int *__restrict__ grp,
                                                 int ngrp, const typename CX<R>::T *__restrict__ src,
                                                 size_t src_stride, int mode, cd *__restrict__ bufW,
                                                 typename CX<R>::T *__restrict__ bufWr,
                                                 double *__restrict__ outv, cd *__restrict__ gscr,
                                                 size_t gslot) {
    typedef typename CX<R>::T CT;
    constexpr int NT = 512, NB = BJ_NB, NS = BJ_NS, LS = NS + 1;
    constexpr int RW = NB * NB + NB * NS; // rotation work items per pair and sub-round
    constexpr double tol_stop = __is_same(R, double) ? 3.1622776601683795e-8 : 1e-5;
    constexpr double tol_off = __is_same(R, double) ? 1e-17 : 1e-12;
    __shared__ cd sS[BJ_MAXPAIR * NS * LS];
    __shared__ cd sV[BJ_MAXPAIR * NS * NS];
    __shared__ double rc[BJ_MAXPAIR * NB];
    __shared__ cd rs[BJ_MAXPAIR * NB];
    __shared__ int rp[BJ_MAXPAIR * NB], rq[BJ_MAXPAIR * NB];
    __shared__ int s_gi[BJ_MAXPAIR * NS]; // global index of each pair's local index, -1 absent
    __shared__ int s_rot[BJ_MAXPAIR];
    __shared__ int s_any;
    const int tid = threadIdx.x;
    cd *Ms = gscr + gslot * blockIdx.x;
    const long ntask = (long)Kb * ngrp;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int g = (int)(task % ngrp), c = (int)(task / ngrp);
        const int blk = grp[g], b = P.blk_dim[blk];
        const CT *S0 = src + (size_t)c * src_stride + P.blk_uoff[blk];
        cd *W = bufW + (size_t)c * P.Lsum + P.blk_uoff[blk];
        for (int e = tid; e < b * b; e += NT) {
            Ms[e] = to_cd(S0[e]);
            W[e] = cmk(e % b == e / b ? 1.0 : 0.0, 0.0);
        }
        const int nbk = 2 * ((b + NS - 1) / NS); // even number of column blocks
        const int npr = nbk / 2;                 // pairs per round, <= BJ_MAXPAIR
        __syncthreads();
        for (int sweep = 0; sweep < 40 && b > 1; sweep++) {
            if (tid == 0) s_any = 0;
            for (int round = 0; round < nbk - 1; round++) {
                for (int x = tid; x < npr * NS; x += NT) {
                    const int pi = x / NS, l = x % NS;
                    const int pa = pi, pb = nbk - 1 - pi;
                    const int ia = pa == 0 ? 0 : ((pa - 1 + round) % (nbk - 1)) + 1;
                    const int ib = pb == 0 ? 0 : ((pb - 1 + round) % (nbk - 1)) + 1;
                    const int Pb = min(ia, ib), Qb = max(ia, ib);
                    const int gi = l < NB ? Pb * NB + l : Qb * NB + (l - NB);
                    s_gi[x] = gi < b ? gi : -1;
                }
                if (tid < BJ_MAXPAIR) s_rot[tid] = 0;
                __syncthreads();
                for (int x = tid; x < npr * NS * NS; x += NT) {
                    const int pi = x / (NS * NS), e = x % (NS * NS), i = e % NS, j = e / NS;
                    const int gi = s_gi[pi * NS + i], gj = s_gi[pi * NS + j];
                    sS[pi * NS * LS + i + j * LS] =
                        (gi >= 0 && gj >= 0) ? Ms[gi + (size_t)gj * b] : cmk(0, 0);
                    sV[pi * NS * NS + e] = cmk(i == j ? 1.0 : 0.0, 0.0);
                }
                __syncthreads();
                // one joint-Jacobi sweep of every pair's submatrix (circle method over NS)
                for (int sr = 0; sr < NS - 1; sr++) {
                    for (int x = tid; x < npr * NB; x += NT) {
                        const int pi = x / NB, k = x % NB;
                        const int pa = k, pb = NS - 1 - k;
                        const int ia = pa == 0 ? 0 : ((pa - 1 + sr) % (NS - 1)) + 1;
                        const int ib = pb == 0 ? 0 : ((pb - 1 + sr) % (NS - 1)) + 1;
                        const int p = min(ia, ib), q = max(ia, ib);
                        const cd *Sp = sS + pi * NS * LS;
                        double cth = 1.0;
                        cd sth = cmk(0, 0);
                        bool big = false;
                        if (s_gi[pi * NS + p] >= 0 && s_gi[pi * NS + q] >= 0 &&
                            jj_rotation(Sp[p + p * LS], Sp[q + q * LS], Sp[p + q * LS],
                                        Sp[q + p * LS], tol_off, tol_stop, cth, sth, big)) {
                            s_rot[pi] = 1;
                            if (big) s_any = 1;
                        }
                        rc[x] = cth;
                        rs[x] = sth;
                        rp[x] = p;
                        rq[x] = q;
                    }
                    __syncthreads();
                    // S <- R^H S R on every 2 x 2 block of rotation pairs; V <- V R
                    for (int x = tid; x < npr * RW; x += NT) {
                        const int pi = x / RW, e = x % RW;
                        if (e < NB * NB) {
                            cd *Sp = sS + pi * NS * LS;
                            const int ki = pi * NB + e % NB, kj = pi * NB + e / NB;
                            const int p = rp[ki], q = rq[ki], p2 = rp[kj], q2 = rq[kj];
                            const double ci = rc[ki], cj = rc[kj];
                            const cd si = rs[ki], sj = rs[kj];
                            const cd a = Sp[p + p2 * LS], bb2 = Sp[p + q2 * LS];
                            const cd c2 = Sp[q + p2 * LS], d = Sp[q + q2 * LS];
                            const cd a1 = cadd(cscale(a, ci), cmul(cconj(si), c2));
                            const cd c1 = csub(cscale(c2, ci), cmul(si, a));
                            const cd b1 = cadd(cscale(bb2, ci), cmul(cconj(si), d));
                            const cd d1 = csub(cscale(d, ci), cmul(si, bb2));
                            Sp[p + p2 * LS] = cadd(cscale(a1, cj), cmul(sj, b1));
                            Sp[p + q2 * LS] = csub(cscale(b1, cj), cmul(cconj(sj), a1));
                            Sp[q + p2 * LS] = cadd(cscale(c1, cj), cmul(sj, d1));
                            Sp[q + q2 * LS] = csub(cscale(d1, cj), cmul(cconj(sj), c1));
                        } else {
                            cd *Vp = sV + pi * NS * NS;
                            const int f = e - NB * NB, row = f % NS, k = pi * NB + f / NS;
                            const int p = rp[k], q = rq[k];
                            const double cth = rc[k];
                            const cd sth = rs[k];
                            const cd wp = Vp[row + p * NS], wq = Vp[row + q * NS];
                            Vp[row + p * NS] = cadd(cscale(wp, cth), cmul(sth, wq));
                            Vp[row + q * NS] = csub(cscale(wq, cth), cmul(cconj(sth), wp));
                        }
                    }
                    __syncthreads();
                }
                // rows of every rotated pair: M(idx, col) <- V^H M(idx, col)
                for (int x = tid; x < npr * b; x += NT) {
                    const int pi = x / b, col = x % b;
                    if (!s_rot[pi]) continue;
                    const int *gi = s_gi + pi * NS;
                    const cd *Vp = sV + pi * NS * NS;
                    cd m[NS];
#pragma unroll
                    for (int l = 0; l < NS; l++)
                        m[l] = gi[l] >= 0 ? Ms[gi[l] + (size_t)col * b] : cmk(0, 0);
#pragma unroll
                    for (int i = 0; i < NS; i++) {
                        if (gi[i] < 0) continue;
                        cd acc = cmk(0, 0);
#pragma unroll
                        for (int l = 0; l < NS; l++) cfma(acc, cconj(Vp[l + i * NS]), m[l]);
                        Ms[gi[i] + (size_t)col * b] = acc;
                    }
                }
                __syncthreads();
                // columns of every rotated pair: M(row, idx) <- M(row, idx) V, W likewise
                for (int x = tid; x < 2 * npr * b; x += NT) {
                    const int which = x / (npr * b), y = x % (npr * b), pi = y / b, row = y % b;
                    if (!s_rot[pi]) continue;
                    const int *gi = s_gi + pi * NS;
                    const cd *Vp = sV + pi * NS * NS;
                    cd *X = which ? W : Ms;
                    cd m[NS];
#pragma unroll
                    for (int l = 0; l < NS; l++)
                        m[l] = gi[l] >= 0 ? X[row + (size_t)gi[l] * b] : cmk(0, 0);
#pragma unroll
                    for (int j = 0; j < NS; j++) {
                        if (gi[j] < 0) continue;
                        cd acc = cmk(0, 0);
#pragma unroll
                        for (int l = 0; l < NS; l++) cfma(acc, m[l], Vp[l + j * NS]);
                        X[row + (size_t)gi[j] * b] = acc;
                    }
                }
                __syncthreads();
            }
            // every thread reads the flag before the next sweep's reset by thread 0
            const int any = s_any;
            __syncthreads();
            if (!any) break;
        }
        double *o = outv + (size_t)c * P.N + P.blk_soff[blk];
        for (int r = tid; r < b; r += NT) {
            const cd l = Ms[r + (size_t)r * b];
            o[r] = (mode == 0) ? atan2(l.y, l.x) : l.x;
        }
        if (bufWr) {
            CT *Wr = bufWr + (size_t)c * P.Lsum + P.blk_uoff[blk];
            for (int e = tid; e < b * b; e += NT) Wr[e] = CX<R>::from(W[e]);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------ k_bmul
// B_k = A_k W for every (crystallite, step, block) of a group; writes into bufB.
// The complex64 chain of 65-80-state blocks runs eight waves (two per SIMD; its LDS allows one
// workgroup per CU): at four waves it needed 432 registers and ran one wave per SIMD, and eight
// waves took 18% off its time on 8-spin clusters (class 4 measured the same either way).
template <int M, typename R, bool SPLIT> struct BmThreads {
    static constexpr int v = (M == 5 && std::is_same<R, float>::value && !SPLIT) ? 512 : NThr<M>::v;
};
template <int M, typename R, bool SPLIT = false>
__global__ void __launch_bounds__((BmThreads<M, R, SPLIT>::v), (BmThreads<M, R, SPLIT>::v == 512) ? 1 : (M == 3 && sizeof(R) == 4) ? 3 : (M == 4 && sizeof(R) == 4) ? 2 : 1) k_bmul(DevPlan P, int Kb, const int *__restrict__ grp,
                                                     int ngrp,
                                                     const typename CX<R>::T *__restrict__ bufA,
                                                     const typename CX<R>::T *__restrict__ bufW,
                                                     typename CX<R>::T *__restrict__ bufB) {
    constexpr int NT = BmThreads<M, R, SPLIT>::v;
    typedef typename CX<R>::T CT;
    // complex64: one task per (crystallite, block) walks all n steps with W resident in LDS;
    // complex128 (and the BF16x3 form): one task per (crystallite, step, block)
    constexpr bool CHAIN = std::is_same<R, float>::value && !SPLIT;
    constexpr int CHB = 4 * 16 * M * (16 * M + LDS_PAD) * 4, GB = CGemm<M, NT, R, SPLIT>::BYTES;
    __shared__ __align__(16) char smem[CHAIN && CHB > GB ? CHB : GB];
    if constexpr (CHAIN) {
        // blocks above 80 states (M = 5) exceed the chain's tile: one task per step for them
        const long ntask = (long)Kb * ngrp;
        for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
            const int g = (int)(task % ngrp), c = (int)(task / ngrp);
            const int blk = grp[g], b = P.blk_dim[blk];
            const size_t base = (size_t)c * P.n * P.Lsum + P.blk_uoff[blk];
            const CT *W = bufW + (size_t)c * P.Lsum + P.blk_uoff[blk];
            if (M < 5 || b <= 16 * M) {
                chain_mfma<M, NT, false>(b, P.n, bufA + base, P.Lsum, W, bufB + base, P.Lsum,
                                         P.n, (float *)smem);
            } else if constexpr (M == 5) {
                for (int k = 0; k < P.n; k++)
                    CGemm<M, NT, R, SPLIT>::run(b, b, b, bufA + base + (size_t)k * P.Lsum, b,
                                                false, W, b, bufB + base + (size_t)k * P.Lsum, b,
                                                smem);
            }
        }
    } else {
        const long ntask = (long)Kb * P.n * ngrp;
        for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
            const int g = (int)(task % ngrp);
            const long ck = task / ngrp;
            const int k = (int)(ck % P.n), c = (int)(ck / P.n);
            const int blk = grp[g], b = P.blk_dim[blk];
            const size_t off = ((size_t)c * P.n + k) * P.Lsum + P.blk_uoff[blk];
            const CT *W = bufW + (size_t)c * P.Lsum + P.blk_uoff[blk];
            CGemm<M, NT, R, SPLIT>::run(b, b, b, bufA + off, b, false, W, b, bufB + off, b, smem);
        }
    }
}

// T[a + s rows] = sum_z val[z] B[col[z] + s ldb] for a < rows, s < ncols: the CSR rows
// rp[0..rows] (absolute offsets into col / val) of one block pair of a sparse operator times
// a dense block B (ldb x ncols, contiguous).  The rows' columns and values, and B itself when
// it fits too, are first staged in `smem` (free until the GEMM that follows, `cap` bytes): the
// gathers of B then read LDS instead of scattered global addresses.  When B does not fit it is
// gathered from global memory; when the rows do not fit either, they are read from global
// memory as well.  Ends with a barrier.
template <int NT, typename R>
__device__ __forceinline__ void sparse_rows_times(const int *__restrict__ rp,
                                                  const int *__restrict__ col,
                                                  const cd *__restrict__ val, int rows, int ncols,
                                                  const typename CX<R>::T *__restrict__ B, int ldb,
                                                  typename CX<R>::T *__restrict__ T, char *smem,
                                                  int cap) {
    typedef typename CX<R>::T CT;
    const int tid = threadIdx.x;
    const int z0 = rp[0], nnz = rp[rows] - z0;
    const int off = ((rows + 1) * 4 + 15) & ~15;
    const size_t boff = ((size_t)off + (size_t)nnz * (sizeof(CT) + 4) + 15) & ~(size_t)15;
    const size_t nb = (size_t)ldb * ncols;
    if (off + (size_t)nnz * (sizeof(CT) + 4) <= (size_t)cap) {
        int *s_rp = (int *)smem;
        CT *s_v = (CT *)(smem + off);
        int *s_c = (int *)(s_v + nnz);
        const bool bl = boff + nb * sizeof(CT) <= (size_t)cap;
        CT *s_b = (CT *)(smem + boff);
        for (int i = tid; i <= rows; i += NT) s_rp[i] = rp[i] - z0;
        for (int i = tid; i < nnz; i += NT) {
            s_c[i] = col[z0 + i];
            s_v[i] = CX<R>::from(val[z0 + i]);
        }
        if (bl)
            for (int i = tid; i < (int)nb; i += NT) s_b[i] = B[i];
        __syncthreads();
        if (bl) {
            for (int e = tid; e < rows * ncols; e += NT) {
                const int a = e % rows, s = e / rows;
                CT acc = Arith<CT>::zero();
                for (int z = s_rp[a]; z < s_rp[a + 1]; z++) cfma(acc, s_v[z], s_b[s_c[z] + s * ldb]);
                T[e] = acc;
            }
        } else {
            for (int e = tid; e < rows * ncols; e += NT) {
                const int a = e % rows, s = e / rows;
                CT acc = Arith<CT>::zero();
                for (int z = s_rp[a]; z < s_rp[a + 1]; z++)
                    cfma(acc, s_v[z], B[s_c[z] + (size_t)s * ldb]);
                T[e] = acc;
            }
        }
    } else {
        for (int e = tid; e < rows * ncols; e += NT) {
            const int a = e % rows, s = e / rows;
            CT acc = Arith<CT>::zero();
            for (int z = rp[a]; z < rp[a + 1]; z++)
                cfma(acc, CX<R>::from(val[z]), B[col[z] + (size_t)s * ldb]);
            T[e] = acc;
        }
    }
    __syncthreads();
}

// ------------------------------------------------------------------------------ k_qr
// Q products of one block pair (both blocks of at most 16 M states, complex64) for all n steps
// of a crystallite: the pair's sparse rows of D are staged in LDS once; per step B_J is copied
// to LDS for the gathers, T1 = D_IJ B_J is formed straight into the B operand planes
// (Bt[s][a] = T1(a, s)) and Q = B_I^H T1 taken on matrix cores, so T1 never goes to global
// memory.  Returns false (nothing done) when the pair's rows do not fit the LDS left for them.
// Layout after the operand planes: B_J copy (16 M x 16 M), row pointers, values, columns.
template <int M> struct QrChainLds {
    static constexpr int TM = 16 * M, PLANES = 4 * TM * (TM + 4) * 4, SB = TM * TM * 8;
    static constexpr int CSR = M == 1 ? 2304 : 4096, BYTES = PLANES + SB + CSR;
};
template <int M, int NT>
__device__ __forceinline__ bool qr_chain_mfma(int bI, int bJ, int n, const int *__restrict__ rp,
                                              const int *__restrict__ col,
                                              const cd *__restrict__ val,
                                              const cf *__restrict__ BI0,
                                              const cf *__restrict__ BJ0, size_t bstride,
                                              cf *__restrict__ Q0, size_t qstride, char *smem) {
    typedef QrChainLds<M> L;
    constexpr int TM = L::TM, LDK = TM + 4, NW = NT / 64, NSUB = (M * M + NW - 1) / NW;
    const int z0 = rp[0], nnz = rp[bI] - z0;
    const int roff = ((bI + 1) * 4 + 15) & ~15;
    if (roff + nnz * 12 > L::CSR) return false;
    float *Ar = (float *)smem, *Ai = Ar + TM * LDK, *Br = Ai + TM * LDK, *Bi = Br + TM * LDK;
    cf *sB = (cf *)(smem + L::PLANES);
    int *s_rp = (int *)(smem + L::PLANES + L::SB);
    cf *s_v = (cf *)((char *)s_rp + roff);
    int *s_c = (int *)(s_v + nnz);
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, g = lane >> 4, r = lane & 15;
    for (int i = tid; i <= bI; i += NT) s_rp[i] = rp[i] - z0;
    for (int i = tid; i < nnz; i += NT) {
        s_c[i] = col[z0 + i];
        s_v[i] = CX<float>::from(val[z0 + i]);
    }
    const int kp = (bI + 15) & ~15;
    for (int k = 0; k < n; k++) {
        const cf *BI = BI0 + (size_t)k * bstride, *BJ = BJ0 + (size_t)k * bstride;
        for (int idx = tid; idx < bJ * bJ; idx += NT) sB[idx] = BJ[idx];
        // A operand At[i][q] = conj(B_I(q, i)), zero past bI
        for (int idx = tid; idx < TM * TM; idx += NT) {
            const int q = idx % TM, i = idx / TM;
            cf v = make_float2(0.0f, 0.0f);
            if (i < bI && q < bI) {
                v = BI[q + i * bI];
                v.y = -v.y;
            }
            Ar[i * LDK + q] = v.x;
            Ai[i * LDK + q] = v.y;
        }
        __syncthreads(); // B_J copy, rows and A operand in place
        for (int idx = tid; idx < TM * TM; idx += NT) {
            const int a = idx % TM, sc = idx / TM;
            cf acc = make_float2(0.0f, 0.0f);
            if (a < bI && sc < bJ)
                for (int z = s_rp[a]; z < s_rp[a + 1]; z++) cfma(acc, s_v[z], sB[s_c[z] + sc * bJ]);
            Br[sc * LDK + a] = acc.x;
            Bi[sc * LDK + a] = acc.y;
        }
        __syncthreads(); // T1 in the B operand
        v4f cr[NSUB], ci[NSUB];
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            cr[s] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
            ci[s] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        }
        for (int qq = 0; qq < kp; qq += 16) {
#pragma unroll
            for (int s = 0; s < NSUB; s++) {
                const int t = w + NW * s;
                if (t < M * M) {
                    const int oa = (16 * (t / M) + r) * LDK + qq + 4 * g;
                    const int ob = (16 * (t % M) + r) * LDK + qq + 4 * g;
                    const v4f ar = *(const v4f *)&Ar[oa], ai = *(const v4f *)&Ai[oa];
                    const v4f br = *(const v4f *)&Br[ob], bi = *(const v4f *)&Bi[ob];
                    const v4f na = -ai;
#pragma unroll
                    for (int u = 0; u < 4; u++) {
                        cr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[u], br[u], cr[s], 0, 0, 0);
                        ci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[u], bi[u], ci[s], 0, 0, 0);
                        cr[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(na[u], bi[u], cr[s], 0, 0, 0);
                        ci[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[u], br[u], ci[s], 0, 0, 0);
                    }
                }
            }
        }
        cf *Qk = Q0 + (size_t)k * qstride;
#pragma unroll
        for (int s = 0; s < NSUB; s++) {
            const int t = w + NW * s;
            if (t < M * M) {
                const int cc = 16 * (t % M) + r, row0 = 16 * (t / M) + 4 * g;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (row0 + j < bI && cc < bJ) Qk[row0 + j + cc * bI] = make_float2(cr[s][j], ci[s][j]);
            }
        }
        __syncthreads(); // every operand read done before the next step's copies
    }
    return true;
}

// One task = (crystallite, step, block pair (I,J)).  With B = A_{k+1} W (per block):
//   Q(r,s) = sum conj(B_I(a,r)) D_IJ(a,b) B_J(b,s)        r in I, s in J
//   R(s,r) = sum conj(B_J(b,s)) rho_JI(b,a) B_I(a,r)
// (rho_eff = det + det^H when rho == (det+det^H)/2, compute.h:381-382).  Stored per element
// e = r + s*b_I at offset pair_eoff, layout [c][k][e] with rows of Estr elements:
// Qs = Q(r,s), Rs = R(s,r).
// FLIP (complex64, tile classes 1-4; a variant the launch autotuner may choose): the other of
// the two forms -- the per-step form for classes 1-2, the step-walking form for 3-4.
// The complex64 per-step products of tile classes 4-5 (blocks of 49-80 states) run eight waves,
// two workgroups per CU (class 4: 79 registers, 19% faster than four waves; class 5: 117
// registers, 3% faster); class 3 measured 13% slower at eight waves and keeps four, held to five
// workgroups per CU (90 registers, no spills; 11% faster than four).
template <int M, typename R, bool SPLIT, bool FLIP = false> struct QrThreads {
    static constexpr int v = (M >= 4 && !FLIP && std::is_same<R, float>::value && !SPLIT) ? 512 : NThr<M>::v;
};
template <int M, typename R, bool SPLIT = false, bool FLIP = false>
__global__ void __launch_bounds__((QrThreads<M, R, SPLIT, FLIP>::v), (QrThreads<M, R, SPLIT, FLIP>::v == 512) ? 2 : (M == 5 && sizeof(R) == 4) ? 3 : (M == 3 && sizeof(R) == 4 && !FLIP && !SPLIT) ? 5 : (M == 4 && sizeof(R) == 8) ? 3 : 1) k_qr(DevPlan P, int Kb, const int *__restrict__ pgrp,
                                                   int ngrp,
                                                   const typename CX<R>::T *__restrict__ bufB,
                                                   typename CX<R>::T *__restrict__ scratch,
                                                   size_t slot, typename CX<R>::T *__restrict__ Qt,
                                                   typename CX<R>::T *__restrict__ Rt) {
    constexpr int NT = QrThreads<M, R, SPLIT, FLIP>::v;
    typedef typename CX<R>::T CT;
    // complex64 pairs of small blocks with R formed from Q: one task per (crystallite, pair)
    // walks all steps (qr_chain_mfma); otherwise one task per (crystallite, step, pair)
    constexpr bool CHAIN = ((M <= 2) != FLIP) && M <= 4 && std::is_same<R, float>::value && !SPLIT;
    constexpr int GB = CGemm<M, NT, R, SPLIT>::BYTES, CB = QrChainLds<M>::BYTES;
    __shared__ __align__(16) char smem[CHAIN && CB > GB ? CB : GB];
    const int tid = threadIdx.x;
    CT *T1 = scratch + slot * 2 * blockIdx.x;
    CT *G = T1 + slot;
    if constexpr (CHAIN) {
        if (P.rt_from_q) {
            const long ntc = (long)Kb * ngrp;
            for (long task = blockIdx.x; task < ntc; task += gridDim.x) {
                const int g = (int)(task % ngrp), c = (int)(task / ngrp);
                const int pr = pgrp[g];
                const int I = P.pair_I[pr], J = P.pair_J[pr];
                const int bI = P.blk_dim[I], bJ = P.blk_dim[J];
                const size_t b0 = (size_t)c * P.n * P.Lsum;
                const size_t e0 = (size_t)c * P.n * P.Estr + P.pair_eoff[pr];
                if (qr_chain_mfma<M, NT>(bI, bJ, P.n, P.d_rowptr + P.pair_doff[pr], P.d_col,
                                         (const cd *)P.d_val, bufB + b0 + P.blk_uoff[I],
                                         bufB + b0 + P.blk_uoff[J], P.Lsum, Qt + e0, P.Estr, smem))
                    continue;
                for (int k = 0; k < P.n; k++) { // rows too many for the LDS left: per step
                    const CT *BI = bufB + b0 + (size_t)k * P.Lsum + P.blk_uoff[I];
                    const CT *BJ = bufB + b0 + (size_t)k * P.Lsum + P.blk_uoff[J];
                    sparse_rows_times<NT, R>(P.d_rowptr + P.pair_doff[pr], P.d_col,
                                             (const cd *)P.d_val, bI, bJ, BJ, bJ, T1, smem, GB);
                    CGemm<M, NT, R, SPLIT>::run(bI, bJ, bI, BI, bI, true, T1, bI,
                                                Qt + e0 + (size_t)k * P.Estr, bI, smem);
                }
            }
            return;
        }
    }
    const long ntask = (long)Kb * P.n * ngrp;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int g = (int)(task % ngrp);
        const long ck = task / ngrp;
        const int k = (int)(ck % P.n), c = (int)(ck / P.n);
        const int pr = pgrp[g];
        const int I = P.pair_I[pr], J = P.pair_J[pr];
        const int bI = P.blk_dim[I], bJ = P.blk_dim[J];
        const CT *BI = bufB + ((size_t)c * P.n + k) * P.Lsum + P.blk_uoff[I];
        const CT *BJ = bufB + ((size_t)c * P.n + k) * P.Lsum + P.blk_uoff[J];
        const size_t eo = ((size_t)c * P.n + k) * P.Estr + P.pair_eoff[pr];
        sparse_rows_times<NT, R>(P.d_rowptr + P.pair_doff[pr], P.d_col, (const cd *)P.d_val, bI,
                                 bJ, BJ, bJ, T1, smem, CGemm<M, NT, R, SPLIT>::BYTES);
        CGemm<M, NT, R, SPLIT>::run(bI, bJ, bI, BI, bI, true, T1, bI, Qt + eo, bI, smem);
        if (P.rt_from_q) continue; // R is formed from Q in k_corr
        if (P.rho_c) {
            // prepared rho of this crystallite (pulses before the FID): stored as rho_JI^H
            // (b_I x b_J, column-major), so T1 = rho_JI B_I is its conjugate transpose times B_I
            const CT *X = (const CT *)P.rho_c + (size_t)c * P.Etot + P.pair_eoff[pr];
            CGemm<M, NT, R, SPLIT>::run(bJ, bI, bI, X, bI, true, BI, bI, T1, bJ, smem);
        } else {
            sparse_rows_times<NT, R>(P.r_rowptr + P.pair_roff[pr], P.r_col, (const cd *)P.r_val,
                                     bJ, bI, BI, bI, T1, smem, CGemm<M, NT, R, SPLIT>::BYTES);
        }
        // R^T(r,s) = sum_b T1(b,r) conj(B_J(b,s)) = (T1^T conj(B_J))(r,s): computed as
        // G = B_J^H T1 (bJ x bI) then transposed into Rt
        CGemm<M, NT, R, SPLIT>::run(bJ, bI, bJ, BJ, bJ, true, T1, bJ, G, bJ, smem);
        for (int e = tid; e < bI * bJ; e += NT) {
            int r = e % bI, s = e / bI;
            Rt[eo + e] = G[s + (size_t)r * bJ];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------ k_corr
// One workgroup task = (crystallite, segment of consecutive chunks of CH elements); per chunk
// the steps below run in LDS.  Staging applies the gamma-COMPUTE
// twiddles, Qt_k = Q_k z^-k and Rt_k = R_k(s,r) z^k with z = exp(i phi_e),
// phi_e = (th_r - th_s)/n (compute.h:406-427 written as exp(-i ind2 tau w) = z^-(i+p)%n z^i),
// each thread walking a run of k with one sincos and complex multiplies.  Thread (el, pg)
// then computes F_p(el) for p in [8 pg, 8 pg + 8):
//   F_p(e) = sum_i Rt_i(e) Qt_{(i+p)%n}(e)
// For n a power of two (16 <= n <= 2 CORR_TW) the correlation is formed with FFTs instead:
//   F = IDFT( Rtilde . DFT(Qt) ),  Rtilde_k = sum_i Rt_i e^{+2 pi i ik/n},
// as in-place mixed-radix passes in LDS (radix 8, the last pass radix 2, 4 or 8): staging and
// the twiddles fuse into the first forward pass, the last forward pass, the product and the first
// inverse pass into one, and G_p = F_p z^p into the last inverse pass (corr_fwd_pass and the
// passes after it).  Thread t then forms fid[t] = sum_e F_{t%n}(e) z_e^t (compute.h:432-457)
// from G and a per-chunk table of z_e^{n j}, summed over the segment's chunks in registers and
// stored as the segment's partial FID part[c][sg][t] (accumulated chunk by chunk by the thread
// owning t when the FID is too long to hold).  No atomics: each (crystallite, segment, point) has
// one writer, and k_reduce adds the segments in order, so the FID does not depend on scheduling.
// CH = min(CORR_SPAN / n, 256 / ceil(n/8)) is computed identically on the host (blfid_dev.h).
// Q rows are Estr elements apart (Etot padded to 16 elements), so every chunk row starts on a
// 128-byte line.
// In-register DFT of RAD = 2, 4 or 8 points, X_k = sum_m x_m e^{SG 2 pi i m k / RAD}, natural
// order in and out (radix 8 as two radix-4 halves joined by e^{SG i pi k / 4}).
template <int SG, typename V> __device__ __forceinline__ V mul_si(V a) { // a * (SG i)
    V r;
    r.x = SG > 0 ? -a.y : a.y;
    r.y = SG > 0 ? a.x : -a.x;
    return r;
}
template <int RAD, int SG, typename R> struct RegDFT;
template <int SG, typename R> struct RegDFT<2, SG, R> {
    typedef typename CX<R>::T V;
    static __device__ __forceinline__ void run(V (&x)[2]) {
        const V a = x[0];
        x[0] = cadd(a, x[1]);
        x[1] = csub(a, x[1]);
    }
};
template <int SG, typename R> struct RegDFT<4, SG, R> {
    typedef typename CX<R>::T V;
    static __device__ __forceinline__ void run(V (&x)[4]) {
        const V a = cadd(x[0], x[2]), b = csub(x[0], x[2]);
        const V c = cadd(x[1], x[3]), d = mul_si<SG>(csub(x[1], x[3]));
        x[0] = cadd(a, c);
        x[2] = csub(a, c);
        x[1] = cadd(b, d);
        x[3] = csub(b, d);
    }
};
template <int SG, typename R> struct RegDFT<8, SG, R> {
    typedef typename CX<R>::T V;
    static __device__ __forceinline__ void run(V (&x)[8]) {
        V e[4] = {x[0], x[2], x[4], x[6]}, o[4] = {x[1], x[3], x[5], x[7]};
        RegDFT<4, SG, R>::run(e);
        RegDFT<4, SG, R>::run(o);
        const R h = (R)0.70710678118654752440;
        V t1, t3; // o_1 (1 + SG i) h and o_3 (-1 + SG i) h
        t1.x = h * (o[1].x - SG * o[1].y);
        t1.y = h * (o[1].y + SG * o[1].x);
        t3.x = h * (-o[3].x - SG * o[3].y);
        t3.y = h * (-o[3].y + SG * o[3].x);
        const V t2 = mul_si<SG>(o[2]);
        x[0] = cadd(e[0], o[0]);
        x[4] = csub(e[0], o[0]);
        x[1] = cadd(e[1], t1);
        x[5] = csub(e[1], t1);
        x[2] = cadd(e[2], t2);
        x[6] = csub(e[2], t2);
        x[3] = cadd(e[3], t3);
        x[7] = csub(e[3], t3);
    }
};

// Compile-time layout of the FFT variant for n = 2^LN (16 <= n <= 512): chunks of CH = 2048 / n
// elements, LDS slot of (step k, element el) = k (CH + 1) + el (rows padded by one element so that
// one element read across consecutive rows, as the FID sums do, falls in distinct banks).
// tw holds e^{-2 pi i j / n}, j < n / 2.
template <int LN> struct CorrFft {
    static constexpr int N = 1 << LN, CH = CORR_SPAN >> LN, LC = 11 - LN, ROW = CH + 1;
    static constexpr int NP = (LN + 2) / 3;         // passes per transform: radix 8 but the last
    static constexpr int LRL = LN - 3 * (NP - 1);   // log2 of the last pass's radix
    static __device__ __forceinline__ int at(int k, int el) { return k * ROW + el; }
    template <int SG, typename CT> static __device__ __forceinline__ CT twid(const CT *tw, int idx) {
        CT w = tw[idx & (N / 2 - 1)]; // e^{SG 2 pi i idx / n}, 0 <= idx < n
        if (idx & (N / 2)) { w.x = -w.x; w.y = -w.y; }
        if (SG > 0) w.y = -w.y;
        return w;
    }
};

// Mixed-radix in-place FFTs of the correlation F = IDFT(DFT+(Rt) . DFT-(Qt)): the forward
// transforms are decimation-in-frequency passes of radix 8 (the last of radix 2^LRL), output in
// digit-reversed order; the inverse runs the exact inverse passes in reverse order, so it returns
// natural order with no permutation.  A pass over sub-transforms of length L = 2^LL and radix RAD
// takes, for butterfly j < L / RAD of block blk, the RAD points at blk L + j + m L / RAD, applies
// the RAD-point DFT, multiplies output k by e^{SG 2 pi i j k / L} and stores it at
// blk L + j + k L / RAD; the inverse pass multiplies by the conjugate twiddle first and applies
// the opposite-sign DFT.  One work item = one butterfly of one element (2048 / RAD items).
template <int RAD> struct Log2R { static constexpr int v = RAD == 8 ? 3 : RAD == 4 ? 2 : 1; };
// AUTO: R~ = conj(Q~) (the autocorrelation case), only the Q transform is formed.
template <int RAD, int LN, int LL, typename R, bool AUTO>
__device__ __forceinline__ void corr_fwd_pass(typename CX<R>::T *sQ, typename CX<R>::T *sR,
                                              const typename CX<R>::T *tw) {
    typedef typename CX<R>::T CT;
    typedef CorrFft<LN> F;
    constexpr int LR = Log2R<RAD>::v, LS = LL - LR, TSH = LN - LL;
#pragma unroll
    for (int it = 0; it < (CORR_SPAN >> LR) / 256; it++) {
        const int idx = threadIdx.x + 256 * it;
        const int el = idx & (F::CH - 1), b = idx >> F::LC;
        const int j = b & ((1 << LS) - 1), base = F::at(((b >> LS) << LL) + j, el);
        CT q[RAD], r[AUTO ? 1 : RAD];
#pragma unroll
        for (int m = 0; m < RAD; m++) {
            q[m] = sQ[base + ((m << LS) * F::ROW)];
            if constexpr (!AUTO) r[m] = sR[base + ((m << LS) * F::ROW)];
        }
        RegDFT<RAD, -1, R>::run(q);
        if constexpr (!AUTO) RegDFT<RAD, +1, R>::run(r);
#pragma unroll
        for (int k = 0; k < RAD; k++) {
            if (k) {
                const CT w = F::template twid<-1>(tw, (j * k) << TSH);
                q[k] = cmul(q[k], w);
                if constexpr (!AUTO) r[k] = cmul(r[k], cconj(w));
            }
            sQ[base + ((k << LS) * F::ROW)] = q[k];
            if constexpr (!AUTO) sR[base + ((k << LS) * F::ROW)] = r[k];
        }
    }
}
template <int RAD, int LN, int LL, typename R>
__device__ __forceinline__ void corr_inv_pass(typename CX<R>::T *sQ, const typename CX<R>::T *tw) {
    typedef typename CX<R>::T CT;
    typedef CorrFft<LN> F;
    constexpr int LR = Log2R<RAD>::v, LS = LL - LR, TSH = LN - LL;
#pragma unroll
    for (int it = 0; it < (CORR_SPAN >> LR) / 256; it++) {
        const int idx = threadIdx.x + 256 * it;
        const int el = idx & (F::CH - 1), b = idx >> F::LC;
        const int j = b & ((1 << LS) - 1), base = F::at(((b >> LS) << LL) + j, el);
        CT x[RAD];
#pragma unroll
        for (int k = 0; k < RAD; k++) {
            x[k] = sQ[base + ((k << LS) * F::ROW)];
            if (k) x[k] = cmul(x[k], F::template twid<+1>(tw, (j * k) << TSH));
        }
        RegDFT<RAD, +1, R>::run(x);
#pragma unroll
        for (int m = 0; m < RAD; m++) sQ[base + ((m << LS) * F::ROW)] = x[m];
    }
}
// The last forward pass (sub-transforms of length RAD: no twiddles), the product scaled by 1/n
// and the first inverse pass, on consecutive groups of RAD points, into sQ.  AUTO: the product
// is |Q~|^2 / n (R~ = conj(Q~)).
template <int LN, typename R, bool AUTO>
__device__ __forceinline__ void corr_mid_pass(typename CX<R>::T *sQ, const typename CX<R>::T *sR) {
    typedef typename CX<R>::T CT;
    typedef CorrFft<LN> F;
    constexpr int LR = F::LRL, RAD = 1 << LR;
    const R inv_n = (R)(1.0 / (double)F::N);
#pragma unroll
    for (int it = 0; it < (CORR_SPAN >> LR) / 256; it++) {
        const int idx = threadIdx.x + 256 * it;
        const int el = idx & (F::CH - 1), base = F::at((idx >> F::LC) << LR, el);
        CT q[RAD], r[AUTO ? 1 : RAD];
#pragma unroll
        for (int m = 0; m < RAD; m++) {
            q[m] = sQ[base + m * F::ROW];
            if constexpr (!AUTO) r[m] = sR[base + m * F::ROW];
        }
        RegDFT<RAD, -1, R>::run(q);
        if constexpr (AUTO) {
#pragma unroll
            for (int m = 0; m < RAD; m++) {
                q[m].x = (q[m].x * q[m].x + q[m].y * q[m].y) * inv_n;
                q[m].y = (R)0;
            }
        } else {
            RegDFT<RAD, +1, R>::run(r);
#pragma unroll
            for (int m = 0; m < RAD; m++) q[m] = cscale(cmul(q[m], r[m]), inv_n);
        }
        RegDFT<RAD, +1, R>::run(q);
#pragma unroll
        for (int m = 0; m < RAD; m++) sQ[base + m * F::ROW] = q[m];
    }
}

// NPT FID points per thread are held over a segment (npts <= 256 NPT).  LN > 0 selects the FFT
// variant for n = 2^LN (CorrFft<LN>), LN = 0 the direct sum for any n.  The complex64 FFT variant
// with four points is held to three waves per SIMD (168 registers: at four it spilled, and its
// sixteen loads per chunk then waited on the spill reloads); the others are bound by LDS or
// registers at two.  OCC > 0 replaces that bound (a variant the launch autotuner may choose).
template <typename R, int LN, int NPT, bool AUTO, int OCC = 0>
__global__ void __launch_bounds__(256, OCC > 0 ? OCC : (LN > 0 && sizeof(R) == 4 && NPT == 4) ? (AUTO ? 4 : 3) : 2) k_corr(DevPlan P, int Kb,
                                              const typename CX<R>::T *__restrict__ Qt,
                                              const typename CX<R>::T *__restrict__ Rt,
                                              const double *__restrict__ thetas,
                                              const int *__restrict__ elem_r,
                                              const int *__restrict__ elem_s,
                                              double *__restrict__ part, int ch_lo, int ch_hi,
                                              int seg_base) {
    typedef typename CX<R>::T CT;
    constexpr bool FFT = LN > 0;
    typedef CorrFft<FFT ? LN : 4> FF;
    constexpr int SPAN = FFT ? FF::N * FF::ROW : CORR_SPAN, NE = FFT ? FF::CH : 256;
    __shared__ CT sQ[SPAN];
    __shared__ __align__(16) CT sR[SPAN];
    __shared__ double sphi[NE];
    __shared__ cd sstep[NE];
    __shared__ CT s_tw[FFT ? FF::N / 2 : 1];
    const int n = FFT ? FF::N : P.n, tid = threadIdx.x;
    const int CH = FFT ? FF::CH : corr_chunk(n);
    // LDS slot of (step k, element el).  Direct variant: rows of CH elements, the element index
    // XOR-ed with a function of the row when CH is a power of two, so that one element read
    // across consecutive rows (the FID sums) falls in distinct banks; FFT variant: padded rows.
    const int swz = (CH & (CH - 1)) == 0 ? CH - 1 : 0;
    constexpr int ssh = sizeof(CT) == 8 ? 1 : 0;
    auto at = [=](int k, int el) {
        return FFT ? FF::at(k, el) : k * CH + (el ^ ((k >> ssh) & swz));
    };
    if constexpr (FFT) {
        // e^{-2 pi i j / n}, j < n/2
        for (int j = tid; j < FF::N / 2; j += 256)
            s_tw[j] = CX<R>::from(cexpi(-2.0 * 3.14159265358979323846 * (double)j / (double)FF::N));
        __syncthreads();
    }
    const int npg = (n + CORR_PPT - 1) / CORR_PPT;
    const int nseg = 256 / CH;
    const int kpt = (n + nseg - 1) / nseg;
    // this launch's chunks [ch_lo, ch_hi) of the crystallite's elements; its segments are
    // partial sums seg_base, seg_base + 1, ...
    const int nchunk = ch_hi - ch_lo;
    // Task = (crystallite, segment of consecutive chunks); the segmentation depends on the plan
    // only (not on the batch), so a crystallite's FID is summed in the same grouping whatever
    // the batch size.  The segment's FID partial sums (t = tid + 256 m) stay in registers and
    // are stored once at the segment's end (hold), or accumulated in part[] per chunk.
    const int spc = corr_seg_chunks(nchunk);
    const int nsg = (nchunk + spc - 1) / spc;
    const long ntask = (long)Kb * nsg;
    // FFT variant: with J = ceil(npts / n) periods, z_e^t = z_e^p (z_e^n)^j for t = p + n j, so
    // the FID sums are sum_e G_p(e) V_j(e) with G = F z^p formed in place and the table
    // V_j(e) = z_e^(n j) (in the correlation's second buffer, free by then): no sine or cosine
    // per point.  Used when CH J fits CORR_VTAB entries; longer FIDs go out per chunk.
    const int J = (P.npts + n - 1) / n;
    const bool hold = P.npts <= 256 * NPT && (!FFT || CH * J <= CORR_VTAB);
    cd *sV = (cd *)sR;
    const int nth = tid < P.npts ? min(NPT, (P.npts - tid + 255) / 256) : 0;
    int hoff[NPT];
#pragma unroll
    for (int m = 0; m < NPT; m++) hoff[m] = FFT ? 0 : (tid + 256 * m) % n; // step of point tid + 256 m
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int c = (int)(task / nsg), sg = (int)(task % nsg);
        const int ch_end = ch_lo + min(nchunk, (sg + 1) * spc);
        cd hacc[NPT];
#pragma unroll
        for (int m = 0; m < NPT; m++) hacc[m] = cmk(0, 0);
        for (int ch = ch_lo + sg * spc; ch < ch_end; ch++) {
            const long e0 = (long)ch * CH;
            if constexpr (FFT) {
                // staging fused with the first forward pass (radix 8, L = n): item (el, j) loads
                // steps k = j + m n/8 of element e0 + el, applies the gamma-COMPUTE twiddles z^-k
                // (Q) and z^k (R) in FP64 and transforms both
                constexpr int LS = LN - 3;
                const int el = tid & (FF::CH - 1), j = tid >> FF::LC;
                const int base = FF::at(j, el);
                {
                    // all loads are unconditional (indexes clamped in range, masked afterwards)
                    // so that the sixteen loads of Q and R issue back to back
                    const long e = e0 + el;
                    const bool ok = e < P.Etot;
                    const long ec = ok ? e : 0;
                    const int er_ = elem_r[ec], es_ = elem_s[ec];
                    const int rv = (!AUTO && P.rt_from_q) ? P.elem_rev[ec] : -1;
                    const double *th = thetas + (size_t)c * P.N;
                    const size_t str = (size_t)(1 << LS) * P.Estr;
                    const CT *qb = Qt + ((size_t)c * FF::N + j) * P.Estr;
                    const CT *rb = P.rt_from_q ? qb : Rt + ((size_t)c * FF::N + j) * P.Estr;
                    const long rc = P.rt_from_q ? (rv >= 0 ? (long)rv : ec) : ec;
                    const double th_r = th[er_], th_s = th[es_];
                    // AUTO (R = conj(Q): rho = (D + D^H) / 2 and no element has a reverse): the
                    // R sequence is conj of the twiddled Q sequence, so only Q is loaded
                    CT qv[8], rw[AUTO ? 1 : 8];
#pragma unroll
                    for (int m = 0; m < 8; m++) qv[m] = qb[(size_t)m * str + ec];
                    if constexpr (!AUTO) {
#pragma unroll
                        for (int m = 0; m < 8; m++) rw[m] = rb[(size_t)m * str + rc];
                    }
                    const double ph = ok ? (th_r - th_s) / (double)FF::N : 0.0;
                    if (j == 0) {
                        sphi[el] = ph;
                        sstep[el] = cexpi(256.0 * ph);
                    }
                    // Q: okf masks a padded element; R = conj(Q) + reverse Q (rosym, reverse
                    // masked by rvf when the element has none) or the stored R
                    const double okf = ok ? 1.0 : 0.0;
                    const double rvf = (ok && P.rt_from_q && rv >= 0) ? 1.0 : 0.0;
                    const cd w0 = cexpi(-(double)j * ph);                // z^-k
                    const cd st = cexpi(-(double)(1 << LS) * ph);
                    CT x[8];
                    cd w = w0;
#pragma unroll
                    for (int m = 0; m < 8; m++) {
                        x[m] = CX<R>::from(cmul(cscale(to_cd(qv[m]), okf), w));
                        w = cmul(w, st);
                    }
                    RegDFT<8, -1, R>::run(x);
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        if (k) x[k] = cmul(x[k], FF::template twid<-1>(s_tw, j * k));
                        sQ[base + (k << LS) * FF::ROW] = x[k];
                    }
                    if constexpr (!AUTO) {
                        w = cconj(w0);
                        const cd sti = cconj(st);
#pragma unroll
                        for (int m = 0; m < 8; m++) {
                            cd rr;
                            if (P.rt_from_q) // R(s,r) = Q_(J,I)(s,r) + conj(Q_(I,J)(r,s))
                                rr = cadd(cscale(cconj(to_cd(qv[m])), okf), cscale(to_cd(rw[m]), rvf));
                            else
                                rr = cscale(to_cd(rw[m]), okf);
                            x[m] = CX<R>::from(cmul(rr, w));
                            w = cmul(w, sti);
                        }
                        RegDFT<8, +1, R>::run(x);
#pragma unroll
                        for (int k = 0; k < 8; k++) {
                            if (k) x[k] = cmul(x[k], FF::template twid<+1>(s_tw, j * k));
                            sR[base + (k << LS) * FF::ROW] = x[k];
                        }
                    }
                }
                __syncthreads();
                if constexpr (FF::NP == 3) {
                    corr_fwd_pass<8, LN, LN - 3, R, AUTO>(sQ, sR, s_tw);
                    __syncthreads();
                }
                corr_mid_pass<LN, R, AUTO>(sQ, sR);
                __syncthreads();
                if constexpr (FF::NP == 3) {
                    corr_inv_pass<8, LN, LN - 3, R>(sQ, s_tw);
                    __syncthreads();
                }
                // last inverse pass (radix 8, L = n) giving F_p in natural order; for the held FID
                // sums G_p = F_p z^p is stored instead, and the period table V_j(e) = z_e^(n j)
                // goes to the second buffer (free since the product)
                {
                    CT x[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        x[k] = sQ[base + (k << LS) * FF::ROW];
                        if (k) x[k] = cmul(x[k], FF::template twid<+1>(s_tw, j * k));
                    }
                    RegDFT<8, +1, R>::run(x);
                    if (hold) {
                        const double ph = sphi[el];
                        cd w = cexpi((double)j * ph);
                        const cd st = cexpi((double)(1 << LS) * ph);
#pragma unroll
                        for (int m = 0; m < 8; m++) {
                            sQ[base + (m << LS) * FF::ROW] = CX<R>::from(cmul(to_cd(x[m]), w));
                            w = cmul(w, st);
                        }
                    } else {
#pragma unroll
                        for (int m = 0; m < 8; m++) sQ[base + (m << LS) * FF::ROW] = x[m];
                    }
                }
                if (hold)
                    for (int q = tid; q < FF::CH * J; q += 256)
                        sV[q] = cexpi((double)FF::N * (double)(q % J) * sphi[q / J]);
                __syncthreads();
            } else {
                for (int el = tid; el < CH; el += 256) {
                    long e = e0 + el;
                    double ph = 0.0;
                    if (e < P.Etot) {
                        const double *th = thetas + (size_t)c * P.N;
                        ph = (th[elem_r[e]] - th[elem_s[e]]) / (double)n;
                    }
                    sphi[el] = ph;
                    sstep[el] = cexpi(256.0 * ph);
                }
                __syncthreads();
                {
                    const int el = tid % CH, seg = tid / CH;
                    const long e = e0 + el;
                    const int ka = seg * kpt, kb = min(n, ka + kpt);
                    if (seg < nseg && ka < kb) {
                        const double ph = sphi[el];
                        cd w = cexpi(-(double)ka * ph);   // z^-k
                        const cd z = cexpi(ph);
                        const cd zi = cconj(z);
                        for (int k = ka; k < kb; k++) {
                            cd q = cmk(0, 0), r = cmk(0, 0);
                            if (e < P.Etot) {
                                size_t o = ((size_t)c * n + k) * P.Estr + e;
                                q = to_cd(Qt[o]);
                                if (P.rt_from_q) {
                                    // R(s,r) = Q_(J,I)(s,r) + conj(Q_(I,J)(r,s))
                                    r = cconj(q);
                                    const int rv = P.elem_rev[e];
                                    if (rv >= 0) r = cadd(r, to_cd(Qt[((size_t)c * n + k) * P.Estr + rv]));
                                } else {
                                    r = to_cd(Rt[o]);
                                }
                            }
                            sQ[at(k, el)] = CX<R>::from(cmul(q, w));
                            sR[at(k, el)] = CX<R>::from(cmul(r, cconj(w)));
                            w = cmul(w, zi);
                        }
                    }
                }
                __syncthreads();
                CT F[CORR_PPT];
                const int el = tid % CH, pg = tid / CH;
                const int p0 = pg * CORR_PPT;
                const bool active = pg < npg;
#pragma unroll
                for (int q = 0; q < CORR_PPT; q++) F[q] = Arith<CT>::zero();
                if (active) {
                    // window slot (i + q) % 8 holds Qt_{(i + p0 + q) % n}; unrolling i by 8 makes the
                    // rotation static, so the window never moves between registers
                    CT win[CORR_PPT];
#pragma unroll
                    for (int q = 0; q < CORR_PPT; q++) win[q] = sQ[at((p0 + q) % n, el)];
                    const int n8 = n & ~(CORR_PPT - 1);
                    for (int i = 0; i < n8; i += CORR_PPT) {
#pragma unroll
                        for (int u = 0; u < CORR_PPT; u++) {
                            CT r = sR[at(i + u, el)];
#pragma unroll
                            for (int q = 0; q < CORR_PPT; q++) cfma(F[q], r, win[(u + q) % CORR_PPT]);
                            win[u] = sQ[at((i + u + p0 + CORR_PPT) % n, el)];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < CORR_PPT; u++) { // tail when 8 does not divide n
                        if (n8 + u < n) {
                            CT r = sR[at(n8 + u, el)];
#pragma unroll
                            for (int q = 0; q < CORR_PPT; q++) cfma(F[q], r, win[(u + q) % CORR_PPT]);
                            win[u] = sQ[at((n8 + u + p0 + CORR_PPT) % n, el)];
                        }
                    }
                }
                __syncthreads();
                if (active) {
#pragma unroll
                    for (int q = 0; q < CORR_PPT; q++)
                        if (p0 + q < n) sQ[at(p0 + q, el)] = F[q];
                }
                __syncthreads();
            }
            if (FFT && hold) {
                const int ne = (int)min((long)CH, P.Etot - e0);
                for (int e = 0; e < ne; e++) {
#pragma unroll
                    for (int m = 0; m < NPT; m++) {
                        if (m < nth) {
                            const int t = tid + 256 * m;
                            cfma(hacc[m], to_cd(sQ[at(t & (n - 1), e)]), sV[e * J + (t >> LN)]);
                        }
                    }
                }
            } else if (hold) {
                const int ne = (int)min((long)CH, P.Etot - e0);
                for (int e = 0; e < ne; e++) {
                    cd w = cexpi((double)tid * sphi[e]);
                    const cd st = sstep[e];
#pragma unroll
                    for (int m = 0; m < NPT; m++) {
                        if (m < nth) {
                            cfma(hacc[m], to_cd(sQ[at(FFT ? (tid + 256 * m) & (n - 1) : hoff[m], e)]), w);
                            w = cmul(w, st);
                        }
                    }
                }
            }
            for (int tb = 0; !hold && tb < P.npts; tb += 256 * NPT) {
                const int t0 = tb + tid;
                if (t0 < P.npts) {
                    const int nt = min(NPT, (P.npts - t0 + 255) / 256); // t = t0 + 256 m
                    cd acc[NPT];
#pragma unroll
                    for (int m = 0; m < NPT; m++) acc[m] = cmk(0, 0);
                    for (int e = 0; e < CH; e++) {
                        if (e0 + e >= P.Etot) break;
                        cd w = cexpi((double)t0 * sphi[e]);
                        const cd st = sstep[e];
#pragma unroll
                        for (int m = 0; m < NPT; m++) {
                            if (m < nt) {
                                cfma(acc[m], to_cd(sQ[at((t0 + 256 * m) % n, e)]), w);
                                w = cmul(w, st);
                            }
                        }
                    }
                    // the first chunk of the segment stores, the later ones add (same thread)
                    double *dst = part + 2 * (((size_t)c * P.nparts + seg_base + sg) * P.npts);
#pragma unroll
                    for (int m = 0; m < NPT; m++) {
                        if (m < nt) {
                            const int t = t0 + 256 * m;
                            if (ch == ch_lo + sg * spc) {
                                dst[2 * t] = acc[m].x;
                                dst[2 * t + 1] = acc[m].y;
                            } else {
                                dst[2 * t] += acc[m].x;
                                dst[2 * t + 1] += acc[m].y;
                            }
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (hold) {
            double *dst = part + 2 * (((size_t)c * P.nparts + seg_base + sg) * P.npts);
#pragma unroll
            for (int m = 0; m < NPT; m++) {
                if (m < nth) {
                    const int t = tid + 256 * m;
                    dst[2 * t] = hacc[m].x;
                    dst[2 * t + 1] = hacc[m].y;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------ k_reduce
// Powder sum in a fixed order (oneFID.h:396-407, sequencerun_funcs.cc:698-702): crystallites
// fall into groups of G consecutive (absolute) indices; thread (component q, group g of the
// batch) adds fid_c = sum of the crystallite's nparts partial FIDs (in part order) times
// w_c * scale to the group's Kahan sum (gsum, gcomp), crystallites in order.  A group split
// across two batches carries its (sum, compensation) pair, so the result is the same bit for bit
// whatever the batch size.  Layout of grp: [group][0: sums, 1: compensations][2 npts].
__global__ void k_reduce(int Kb, int c0, int npts, int nparts, int G, const double *__restrict__ w,
                         double scale, const double *__restrict__ part, double *__restrict__ grp) {
    const int n2 = 2 * npts;
    const int q = blockIdx.x * blockDim.x + threadIdx.x; // re/im component index
    const int g = c0 / G + blockIdx.y;                   // absolute group
    if (q >= n2) return;
    const int ca = max(c0, g * G), cb = min(c0 + Kb, (g + 1) * G);
    if (ca >= cb) return;
    double *gs = grp + (size_t)g * 2 * n2;
    double s = gs[q], cmp = gs[n2 + q];
    for (int c = ca; c < cb; c++) {
        const double *pc = part + (size_t)(c - c0) * nparts * n2 + q;
        double f = 0.0;
        for (int p = 0; p < nparts; p++) f += pc[(size_t)p * n2];
        double y = w[c] * scale * f - cmp;
        double t = s + y;
        cmp = (t - s) - y;
        s = t;
    }
    gs[q] = s;
    gs[n2 + q] = cmp;
}

// FID point t: the groups' sums added in group order (Kahan), their compensations subtracted,
// divided by norm = Tr(D D^H) (or 1).
__global__ void k_finish(int npts, int ngroups, const double *__restrict__ grp, double2 norm,
                         double *__restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    const int n2 = 2 * npts;
    double acc[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        double s = 0.0, cmp = 0.0, lost = 0.0;
        for (int g = 0; g < ngroups; g++) {
            const double *gs = grp + (size_t)g * 2 * n2;
            double y = gs[2 * t + h] - cmp;
            double u = s + y;
            cmp = (u - s) - y;
            s = u;
            lost += gs[n2 + 2 * t + h];
        }
        acc[h] = s - (cmp + lost);
    }
    cd v = cmk(acc[0], acc[1]);
    // v / norm (complex division; norm = Tr(D D^H) is real positive in practice)
    double d = norm.x * norm.x + norm.y * norm.y;
    cd r = cmk((v.x * norm.x + v.y * norm.y) / d, (v.y * norm.x - v.x * norm.y) / d);
    out[2 * t] = r.x;
    out[2 * t + 1] = r.y;
}

__global__ void k_zero(double *p, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

// ------------------------------------------------------------------------------ k_rho_prep
// Pulses before the gamma-COMPUTE FID (SequenceRun::doFID, sequencerun_funcs.cc:693-703): the
// prepared density operator rho_c = U rho U^H of crystallite c, U its preparation propagator (the
// pulse plan's A_n, FP64, block diagonal in the pulse plan's blocks: U(x, a) = 0 unless x and a
// share a block), evaluated on the FID plan's coupled block pairs only.  One thread per
// (crystallite, pair element e = r + s b_I): x = state of row s of block J, y = state of column r
// of block I, rho_c(x, y) = sum over rho's nonzeros (a, b, v) of U(x, a) v conj(U(y, b)), stored
// conjugated (the pair's block rho_JI^H, column-major b_I x b_J) in the storage precision.
template <typename R>
__global__ void __launch_bounds__(256) k_rho_prep(DevPlan P, DevPlan PP, int Kb,
                                                  const cd *__restrict__ Aprep,
                                                  const int *__restrict__ er,
                                                  const int *__restrict__ es,
                                                  const int *__restrict__ sblk,
                                                  const int *__restrict__ sloc,
                                                  const int *__restrict__ nzi,
                                                  const double *__restrict__ nzv, int nnz,
                                                  typename CX<R>::T *__restrict__ out) {
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)Kb * P.Etot) return;
    const int c = (int)(idx / P.Etot);
    const long e = idx - (long)c * P.Etot;
    const int x = P.blk_states[es[e]], y = P.blk_states[er[e]];
    const int bx = sblk[x], by = sblk[y];
    const cd *Ac = Aprep + (size_t)c * 2 * PP.Lsum; // the pulse plan keeps two product slots
    const cd *Ux = Ac + PP.blk_uoff[bx], *Uy = Ac + PP.blk_uoff[by];
    const int dx = PP.blk_dim[bx], dy = PP.blk_dim[by], lx = sloc[x], ly = sloc[y];
    cd acc = cmk(0.0, 0.0);
    for (int z = 0; z < nnz; z++) {
        const int a = nzi[2 * z], b = nzi[2 * z + 1];
        if (sblk[a] != bx || sblk[b] != by) continue;
        const cd v = cmk(nzv[2 * z], nzv[2 * z + 1]);
        const cd ua = Ux[lx + (size_t)sloc[a] * dx], ub = Uy[ly + (size_t)sloc[b] * dy];
        cfma(acc, cmul(ua, v), cconj(ub));
    }
    out[idx] = CX<R>::from(cconj(acc));
}

// ------------------------------------------------------------------------------ spin state
// Per-crystallite density operators carried through a pulse program: solid-2.0 keeps one ro per
// powder orientation and every used propagator advances it in place (subSequence::propogate,
// sequencerun.cc:164-196, applied from SequenceRun::doFID, sequencerun_funcs.cc:625-744).
// rho_c is N x N complex128, column-major: S[c N^2 + x + y N].

// Every crystallite's rho = rho0 (N x N).
__global__ void __launch_bounds__(256) k_state_fill(long total, long NN, const cd *__restrict__ rho0,
                                                    cd *__restrict__ S) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i < total) S[i] = rho0[i % NN];
}

// rho_c <- (Tr(rho_c C^H) / nrm) C for every crystallite c (nrm = Tr(C C^H), formed on the
// host): one workgroup per crystallite at a time, the trace reduced in FP64 over the N x N
// elements in a fixed order (per thread, then a tree over the workgroup).
__global__ void __launch_bounds__(256) k_state_project(long K, long NN, const cd *__restrict__ C,
                                                       double inv_nrm, cd *__restrict__ S) {
    __shared__ cd red[256];
    const int tid = threadIdx.x;
    for (long c = blockIdx.x; c < K; c += gridDim.x) {
        cd *R = S + c * NN;
        cd acc = cmk(0, 0);
        for (long e = tid; e < NN; e += 256) cfma(acc, R[e], cconj(C[e]));
        red[tid] = acc;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h) red[tid] = cadd(red[tid], red[tid + h]);
            __syncthreads();
        }
        const cd t = cscale(red[0], inv_nrm);
        for (long e = tid; e < NN; e += 256) R[e] = cmul(t, C[e]);
        __syncthreads(); // red[0] read by every thread before the next crystallite's writes
    }
}

// rho_c <- U rho_c U^H, `repeat` times, with U the crystallite's propagator from a pulse plan
// (FP64, block diagonal in that plan's blocks, read from Aprep as k_rho_prep reads it).  One
// workgroup per crystallite at a time; Tw is the workgroup's N x N slot: Tw = rho U^H, then
// rho = U Tw.  Thread e owns element (x, y) = (e mod N, e / N), so consecutive lanes read
// consecutive rows of rho and of Tw.
__global__ void __launch_bounds__(256) k_state_apply(DevPlan PP, int Kb, int c0,
                                                     const cd *__restrict__ Aprep,
                                                     const int *__restrict__ sblk,
                                                     const int *__restrict__ sloc, int N,
                                                     int repeat, cd *__restrict__ S,
                                                     cd *__restrict__ T) {
    const long NN = (long)N * N;
    cd *Tw = T + (size_t)blockIdx.x * NN;
    for (int c = blockIdx.x; c < Kb; c += gridDim.x) {
        const cd *Ac = Aprep + (size_t)c * 2 * PP.Lsum; // the pulse plan keeps two product slots
        cd *R = S + (size_t)(c0 + c) * NN;
        for (int r = 0; r < repeat; r++) {
            for (long e = threadIdx.x; e < NN; e += blockDim.x) {
                const int x = (int)(e % N), y = (int)(e / N);
                const int by = sblk[y], dy = PP.blk_dim[by], ly = sloc[y];
                const cd *Uy = Ac + PP.blk_uoff[by];
                const int *st = PP.blk_states + PP.blk_soff[by];
                cd acc = cmk(0.0, 0.0);
                for (int j = 0; j < dy; j++)
                    cfma(acc, R[x + (size_t)st[j] * N], cconj(Uy[ly + (size_t)j * dy]));
                Tw[e] = acc;
            }
            __syncthreads();
            for (long e = threadIdx.x; e < NN; e += blockDim.x) {
                const int x = (int)(e % N), y = (int)(e / N);
                const int bx = sblk[x], dx = PP.blk_dim[bx], lx = sloc[x];
                const cd *Ux = Ac + PP.blk_uoff[bx];
                const int *st = PP.blk_states + PP.blk_soff[bx];
                cd acc = cmk(0.0, 0.0);
                for (int j = 0; j < dx; j++)
                    cfma(acc, Ux[lx + (size_t)j * dx], Tw[st[j] + (size_t)y * N]);
                R[e] = acc;
            }
            __syncthreads();
        }
    }
}

// part[c] = Tr(rho_c det) = sum over det's nonzeros (a, b, v) of rho_c(b, a) v, one thread per
// crystallite (SequenceRun::doFID's trace(ro, detect), sequencerun_funcs.cc:635, 694).
__global__ void __launch_bounds__(256) k_state_trace(int K, int N, const cd *__restrict__ S,
                                                     const int *__restrict__ nzi,
                                                     const double *__restrict__ nzv, int nnz,
                                                     double *__restrict__ part) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= K) return;
    const cd *R = S + (size_t)c * N * N;
    cd acc = cmk(0.0, 0.0);
    for (int z = 0; z < nnz; z++) {
        const int a = nzi[2 * z], b = nzi[2 * z + 1];
        cfma(acc, R[b + (size_t)a * N], cmk(nzv[2 * z], nzv[2 * z + 1]));
    }
    part[2 * (size_t)c] = acc.x;
    part[2 * (size_t)c + 1] = acc.y;
}

// The FID plan's per-crystallite rho blocks taken from a state: element e of crystallite c of
// the batch holds conj(rho_{c0+c}(x, y)), x the state of row s of block J, y of column r of
// block I -- the layout k_rho_prep writes and k_qr / k_static_fid read.
template <typename R>
__global__ void __launch_bounds__(256) k_state_gather(DevPlan P, int Kb, int c0,
                                                      const cd *__restrict__ S,
                                                      typename CX<R>::T *__restrict__ out) {
    const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (idx >= (long)Kb * P.Etot) return;
    const int c = (int)(idx / P.Etot);
    const long e = idx - (long)c * P.Etot;
    const int x = P.blk_states[P.elem_s[e]], y = P.blk_states[P.elem_r[e]];
    const cd v = S[(size_t)(c0 + c) * P.N * P.N + x + (size_t)y * P.N];
    out[idx] = CX<R>::from(cconj(v));
}

// ------------------------------------------------------------------------------ static path
// H(0, dt) block of every crystallite (wr = 0: rotor phase 0), complex copy for k_eig.
__global__ void __launch_bounds__(256) k_static_h(DevPlan P, int Kb, int c0, const int *__restrict__ grp,
                                                  int ngrp, double *__restrict__ scratch,
                                                  size_t slot, cd *__restrict__ bufH) {
    __shared__ double s_zz[DEV_MAX_TERMS];
    __shared__ double s_ff[DEV_MAX_SPINS * DEV_MAX_SPINS];
    __shared__ double s_v[4];
    __shared__ double s_diag[DEV_MAX_BLOCK];
    __shared__ double s_red[4];
    double *X = scratch + slot * blockIdx.x;
    const double *terms = stage_terms<256>(P);
    const long ntask = (long)Kb * ngrp;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int g = (int)(task % ngrp), c = (int)(task / ngrp);
        const int blk = grp[g], b = P.blk_dim[blk];
        build_h_block<double, 256>(P, terms, P.RP + 9 * (size_t)(c0 + c), P.u3, blk, X, b, s_zz,
                                   s_ff, s_v, s_diag, s_red, nullptr, [](double) { return 1.0; });
        cd *H = bufH + (size_t)c * P.Lsum + P.blk_uoff[blk];
        for (int e = threadIdx.x; e < b * b; e += 256) H[e] = cmk(X[e], 0.0);
        __syncthreads();
    }
}

// Eigenbasis FID of one (crystallite, block pair): with V the eigenvectors (bufW) and lam the
// eigenvalues or eigenphases, A(i,j) = Do(i,j) sig0(j,i) and
// fid[t] += A(i,j) exp(i w0 t (lam_j - lam_i)).
//   static FID:  lam = energies of H, w0 = -2 pi dt, A kept when |A|^2 > 1e-10 (oneFID.h:
//                317-325): fid[t] += A(i,j) (conj(e_i) e_j)^t;
//   ptop FID:    lam = eigenphases of the period propagator U, w0 = 1, no pruning: this is
//                Tr(rho_t det) with rho_{t+1} = U rho_t U^H (sequencerun.cc:164-196).
// The pair's FID goes to its own partial sum part[c][pair][t] (one writer, no atomics).
template <int M>
__global__ void __launch_bounds__(NThr<M>::v) k_static_fid(DevPlan P, int Kb, const int *__restrict__ pgrp,
                                                    int ngrp, const cd *__restrict__ bufW,
                                                    const double *__restrict__ lam,
                                                    cd *__restrict__ scratch, size_t slot,
                                                    double w0, int prune,
                                                    double *__restrict__ part) {
    constexpr int NT = NThr<M>::v;
    __shared__ __align__(16) cd smem[GemmLds<M, cd>::ELEMS];
    const int tid = threadIdx.x;
    cd *T1 = scratch + slot * 3 * blockIdx.x;
    cd *Gq = T1 + slot, *Gr = Gq + slot;
    const long ntask = (long)Kb * ngrp;
    for (long task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int g = (int)(task % ngrp), c = (int)(task / ngrp);
        const int pr = pgrp[g];
        const int I = P.pair_I[pr], J = P.pair_J[pr];
        const int bI = P.blk_dim[I], bJ = P.blk_dim[J];
        const cd *VI = bufW + (size_t)c * P.Lsum + P.blk_uoff[I];
        const cd *VJ = bufW + (size_t)c * P.Lsum + P.blk_uoff[J];
        const double *lI = lam + (size_t)c * P.N + P.blk_soff[I];
        const double *lJ = lam + (size_t)c * P.N + P.blk_soff[J];
        sparse_rows_times<NT, double>(P.d_rowptr + P.pair_doff[pr], P.d_col, (const cd *)P.d_val,
                                      bI, bJ, VJ, bJ, T1, (char *)smem, (int)sizeof(smem));
        wg_gemm<M, NT, cd>(bI, bJ, bI, VI, bI, true, T1, bI, false, Gq, bI, cmk(0, 0), smem);
        if (P.rho_c) {
            // rho of this crystallite from a state: X = rho_JI^H (b_I x b_J), rho_JI V_I = X^H V_I
            const cd *X = (const cd *)P.rho_c + (size_t)c * P.Etot + P.pair_eoff[pr];
            wg_gemm<M, NT, cd>(bJ, bI, bI, X, bI, true, VI, bI, false, T1, bJ, cmk(0, 0), smem);
        } else {
            sparse_rows_times<NT, double>(P.r_rowptr + P.pair_roff[pr], P.r_col,
                                          (const cd *)P.r_val, bJ, bI, VI, bI, T1, (char *)smem,
                                          (int)sizeof(smem));
        }
        wg_gemm<M, NT, cd>(bJ, bI, bJ, VJ, bJ, true, T1, bJ, false, Gr, bJ, cmk(0, 0), smem);
        // A(i,j) = Do(i,j) sig0(j,i): i in I, j in J
        for (int e = tid; e < bI * bJ; e += NT) {
            int r = e % bI, s = e / bI;
            cd a = cmul(Gq[e], Gr[s + (size_t)r * bJ]);
            if (!prune || a.x * a.x + a.y * a.y > 1.0e-10) T1[e] = a;
            else T1[e] = cmk(0, 0);
        }
        __syncthreads();
        for (int t = tid; t < P.npts; t += NT) {
            cd acc = cmk(0, 0);
            for (int e = 0; e < bI * bJ; e++) {
                cd a = T1[e];
                if (a.x == 0.0 && a.y == 0.0) continue;
                int r = e % bI, s = e / bI;
                cfma(acc, a, cexpi(w0 * (double)t * (lJ[s] - lI[r])));
            }
            double *dst = part + 2 * (((size_t)c * P.nparts + pr) * P.npts + t);
            dst[0] = acc.x;
            dst[1] = acc.y;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------ launchers
// Complex64 plans with BF16x3 products launch the SPLIT instance of the product kernels.
#define LAUNCH_SPLIT(kern, M_, R_, grid_, nt_, st_, ...)                                         \
    do {                                                                                         \
        if (std::is_same<R_, float>::value && P->bf16x3)                                         \
            kern<M_, R_, std::is_same<R_, float>::value><<<grid_, nt_, 0, st_>>>(__VA_ARGS__);   \
        else                                                                                     \
            kern<M_, R_, false><<<grid_, nt_, 0, st_>>>(__VA_ARGS__);                            \
    } while (0)
// As LAUNCH_SPLIT, with the workgroup size of each variant from the kernel's thread traits TH.
#define LAUNCH_SPLIT_TH(kern, TH, M_, R_, grid_, st_, ...)                                       \
    do {                                                                                         \
        constexpr bool SP_ = std::is_same<R_, float>::value;                                     \
        if (SP_ && P->bf16x3)                                                                    \
            kern<M_, R_, SP_><<<grid_, TH<M_, R_, SP_>::v, 0, st_>>>(__VA_ARGS__);               \
        else                                                                                     \
            kern<M_, R_, false><<<grid_, TH<M_, R_, false>::v, 0, st_>>>(__VA_ARGS__);           \
    } while (0)
// prec: 0 = FP64 storage (double2), 1 = complex64 storage (float2) of U, A, B, Q, R and the exp
// scratch; the eigen-solver, phases, FIDs and reductions are FP64 in both.

template <typename R>
static void launch_exp(int M, const DevPlan *P, int Kb, int c0, const int *grp, int ngrp,
                       void *scratch, size_t slot, int grid, void *bufU, hipStream_t st) {
    R *sc = (R *)scratch;
    typename CX<R>::T *U = (typename CX<R>::T *)bufU;
    const size_t tb = sizeof(double) * DEV_TERM_STRIDE * (size_t)P->nterms; // staged terms
    switch (M) {
    case 1: k_exp<1, R><<<grid, 64, tb, st>>>(*P, Kb, c0, grp, ngrp, sc, slot, U); break;
    case 2: k_exp<2, R><<<grid, 256, tb, st>>>(*P, Kb, c0, grp, ngrp, sc, slot, U); break;
    case 3: k_exp<3, R><<<grid, 256, tb, st>>>(*P, Kb, c0, grp, ngrp, sc, slot, U); break;
    case 4: k_exp<4, R><<<grid, 256, tb, st>>>(*P, Kb, c0, grp, ngrp, sc, slot, U); break;
    default: k_exp<5, R><<<grid, 256, tb, st>>>(*P, Kb, c0, grp, ngrp, sc, slot, U); break;
    }
}

extern "C" int blfid_dev_launch_exp(int prec, int M, const DevPlan *P, int Kb, int c0,
                                    const int *grp, int ngrp, void *scratch, size_t slot,
                                    int nslots, void *bufU, hipStream_t st) {
    long ntask = (long)Kb * P->n * ngrp;
    const long cap = P->tn_cap[0] > 0 ? std::min<long>(tuned_cap(P, 0, M), nslots) : nslots;
    int grid = (int)(ntask < cap ? ntask : cap);
    if (grid <= 0) return 0;
    if (prec) launch_exp<float>(M, P, Kb, c0, grp, ngrp, scratch, slot, grid, bufU, st);
    else launch_exp<double>(M, P, Kb, c0, grp, ngrp, scratch, slot, grid, bufU, st);
    return (int)hipGetLastError();
}

// Fused propagator + Dyson chain (complex64 gamma-COMPUTE, blocks of at most 80 states).
extern "C" int blfid_dev_launch_expdyson(int M, const DevPlan *P, int Kb, int c0, const int *grp,
                                         int ngrp, void *bufA, hipStream_t st) {
    const long ntask = (long)Kb * ngrp;
    const long cap = tuned_cap(P, 8, M);
    const int grid = (int)(ntask < cap ? ntask : cap);
    if (grid <= 0) return 0;
    cf *A = (cf *)bufA;
    const size_t tb = 0; // the H build reads its quadratic-form tables from global memory
    switch (M) {
    case 1: k_expdyson<1><<<grid, 64, tb, st>>>(*P, Kb, c0, grp, ngrp, A); break;
    case 2: k_expdyson<2><<<grid, 256, tb, st>>>(*P, Kb, c0, grp, ngrp, A); break;
    case 3: k_expdyson<3><<<grid, 256, tb, st>>>(*P, Kb, c0, grp, ngrp, A); break;
    case 4: k_expdyson<4><<<grid, XdThreads<4>::v, tb, st>>>(*P, Kb, c0, grp, ngrp, A); break;
    default: k_expdyson<5><<<grid, XdThreads<5>::v, tb, st>>>(*P, Kb, c0, grp, ngrp, A); break;
    }
    return (int)hipGetLastError();
}

template <typename R>
static void launch_dyson(int M, const DevPlan *P, int Kb, const int *grp, int ngrp, int grid,
                         const void *bufU, void *bufA, hipStream_t st) {
    typedef typename CX<R>::T CT;
    const CT *U = (const CT *)bufU;
    CT *A = (CT *)bufA;
    switch (M) {
    case 1: LAUNCH_SPLIT(k_dyson, 1, R, grid, 64, st, *P, Kb, grp, ngrp, U, A); break;
    case 2: LAUNCH_SPLIT(k_dyson, 2, R, grid, 256, st, *P, Kb, grp, ngrp, U, A); break;
    case 3: LAUNCH_SPLIT(k_dyson, 3, R, grid, 256, st, *P, Kb, grp, ngrp, U, A); break;
    case 4: LAUNCH_SPLIT(k_dyson, 4, R, grid, 256, st, *P, Kb, grp, ngrp, U, A); break;
    default: LAUNCH_SPLIT(k_dyson, 5, R, grid, 256, st, *P, Kb, grp, ngrp, U, A); break;
    }
}

extern "C" int blfid_dev_launch_dyson(int prec, int M, const DevPlan *P, int Kb, const int *grp,
                                      int ngrp, const void *bufU, void *bufA, hipStream_t st) {
    long ntask = (long)Kb * ngrp;
    const long cap = tuned_cap(P, 1, M);
    int grid = (int)(ntask < cap ? ntask : cap);
    if (grid <= 0) return 0;
    if (prec) launch_dyson<float>(M, P, Kb, grp, ngrp, grid, bufU, bufA, st);
    else launch_dyson<double>(M, P, Kb, grp, ngrp, grid, bufU, bufA, st);
    return (int)hipGetLastError();
}

template <typename R>
static void launch_eig(const DevPlan *P, int Kb, const int *grp, int ngrp, int grid,
                       const void *src, size_t src_stride, int mode, cd *bufW, void *bufWr,
                       double *outv, int cls, cd *gscr, size_t gslot, hipStream_t st) {
    typedef typename CX<R>::T CT;
    const CT *S0 = (const CT *)src;
    CT *Wr = (CT *)bufWr;
#define EIG_ARGS *P, Kb, grp, ngrp, S0, src_stride, mode, bufW, Wr, outv, gscr, gslot
    switch (cls) {
    case 0: k_eig<16, 64, R><<<grid, 64, 0, st>>>(EIG_ARGS); break;
    case 1: k_eig<32, 128, R><<<grid, 128, 0, st>>>(EIG_ARGS); break;
    case 2: k_eig<48, 256, R><<<grid, 256, 0, st>>>(EIG_ARGS); break;
    case 3: k_eig<64, 1024, R><<<grid, 1024, 0, st>>>(EIG_ARGS); break;
    case 4: k_eig<80, 1024, R><<<grid, 1024, 0, st>>>(EIG_ARGS); break;
    default: k_eig_blk<R><<<grid, 512, 0, st>>>(EIG_ARGS); break;
    }
#undef EIG_ARGS
}

// bufWr: storage-precision copy of W (complex64 plans), or NULL.  cls: block-size class
// 0..4 (b <= 16, 32, 48, 64, 80, LDS-resident) or 5 (larger, global scratch slots).
extern "C" int blfid_dev_launch_eig(int prec, const DevPlan *P, int Kb, const int *grp, int ngrp,
                                    const void *src, size_t src_stride, int mode, cd *bufW,
                                    void *bufWr, double *outv, int cls, cd *gscr, size_t gslot,
                                    int nslots, hipStream_t st) {
    long ntask = (long)Kb * ngrp;
    long cap = cls >= 5 ? nslots : tuned_cap(P, 2, 2);
    int grid = (int)(ntask < cap ? ntask : cap);
    if (grid <= 0) return 0;
    if (prec) launch_eig<float>(P, Kb, grp, ngrp, grid, src, src_stride, mode, bufW, bufWr, outv, cls, gscr, gslot, st);
    else launch_eig<double>(P, Kb, grp, ngrp, grid, src, src_stride, mode, bufW, bufWr, outv, cls, gscr, gslot, st);
    return (int)hipGetLastError();
}

template <typename R>
static void launch_bmul(int M, const DevPlan *P, int Kb, const int *grp, int ngrp, int grid,
                        const void *bufA, const void *bufW, void *bufB, hipStream_t st) {
    typedef typename CX<R>::T CT;
    const CT *A = (const CT *)bufA, *W = (const CT *)bufW;
    CT *B = (CT *)bufB;
    switch (M) {
    case 1: LAUNCH_SPLIT(k_bmul, 1, R, grid, 64, st, *P, Kb, grp, ngrp, A, W, B); break;
    case 2: LAUNCH_SPLIT_TH(k_bmul, BmThreads, 2, R, grid, st, *P, Kb, grp, ngrp, A, W, B); break;
    case 3: LAUNCH_SPLIT_TH(k_bmul, BmThreads, 3, R, grid, st, *P, Kb, grp, ngrp, A, W, B); break;
    case 4: LAUNCH_SPLIT_TH(k_bmul, BmThreads, 4, R, grid, st, *P, Kb, grp, ngrp, A, W, B); break;
    default: LAUNCH_SPLIT_TH(k_bmul, BmThreads, 5, R, grid, st, *P, Kb, grp, ngrp, A, W, B); break;
    }
}

// bufW: W in the plan's storage precision
extern "C" int blfid_dev_launch_bmul(int prec, int M, const DevPlan *P, int Kb, const int *grp,
                                     int ngrp, const void *bufA, const void *bufW, void *bufB,
                                     hipStream_t st) {
    long ntask = (long)Kb * P->n * ngrp;
    const long cap = tuned_cap(P, 3, M);
    int grid = (int)(ntask < cap ? ntask : cap);
    if (grid <= 0) return 0;
    if (prec) launch_bmul<float>(M, P, Kb, grp, ngrp, grid, bufA, bufW, bufB, st);
    else launch_bmul<double>(M, P, Kb, grp, ngrp, grid, bufA, bufW, bufB, st);
    return (int)hipGetLastError();
}

template <typename R>
static void launch_qr(int M, const DevPlan *P, int Kb, const int *pgrp, int ngrp, int grid,
                      const void *bufB, void *scratch, size_t slot, void *Qt, void *Rt,
                      hipStream_t st) {
    typedef typename CX<R>::T CT;
    const CT *B = (const CT *)bufB;
    CT *sc = (CT *)scratch, *Q = (CT *)Qt, *Rr = (CT *)Rt;
    if constexpr (std::is_same<R, float>::value) {
        if (!P->bf16x3 && M <= 4 && ((P->tn_qr >> (M - 1)) & 1)) {
            switch (M) {
            case 1: k_qr<1, R, false, true><<<grid, 64, 0, st>>>(*P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
            case 2: k_qr<2, R, false, true><<<grid, 256, 0, st>>>(*P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
            case 3: k_qr<3, R, false, true><<<grid, 256, 0, st>>>(*P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
            default: k_qr<4, R, false, true><<<grid, 256, 0, st>>>(*P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
            }
            return;
        }
    }
    switch (M) {
    case 1: LAUNCH_SPLIT(k_qr, 1, R, grid, 64, st, *P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
    case 2: LAUNCH_SPLIT_TH(k_qr, QrThreads, 2, R, grid, st, *P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
    case 3: LAUNCH_SPLIT_TH(k_qr, QrThreads, 3, R, grid, st, *P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
    case 4: LAUNCH_SPLIT_TH(k_qr, QrThreads, 4, R, grid, st, *P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
    default: LAUNCH_SPLIT_TH(k_qr, QrThreads, 5, R, grid, st, *P, Kb, pgrp, ngrp, B, sc, slot, Q, Rr); break;
    }
}

extern "C" int blfid_dev_launch_qr(int prec, int M, const DevPlan *P, int Kb, const int *pgrp,
                                   int ngrp, const void *bufB, void *scratch, size_t slot,
                                   int nslots, void *Qt, void *Rt, hipStream_t st) {
    long ntask = (long)Kb * P->n * ngrp;
    const long cap = P->tn_cap[4] > 0 ? std::min<long>(tuned_cap(P, 4, M), nslots) : nslots;
    int grid = (int)(ntask < cap ? ntask : cap);
    if (grid <= 0) return 0;
    if (prec) launch_qr<float>(M, P, Kb, pgrp, ngrp, grid, bufB, scratch, slot, Qt, Rt, st);
    else launch_qr<double>(M, P, Kb, pgrp, ngrp, grid, bufB, scratch, slot, Qt, Rt, st);
    return (int)hipGetLastError();
}

// One launch over the chunks [lo, hi) of every crystallite of the batch, its segments written as
// partial sums seg_base...; AUTO (FFT sizes only): the autocorrelation variant, for chunks whose
// elements have R = conj(Q) (rosym, no reverse element)
template <typename R, int LN, int NPT, int OCC>
static void launch_corr_t(bool autoc, const DevPlan *P, int Kb, const void *Qt, const void *Rt,
                          const double *thetas, const int *er, const int *es, double *part,
                          int lo, int hi, int seg_base, hipStream_t st) {
    typedef typename CX<R>::T CT;
    const long ntask = (long)Kb * corr_range_segments(hi - lo);
    const long cap = tuned_cap(P, 5, 2);
    const int grid = (int)(ntask < cap ? ntask : cap);
    if (grid <= 0) return;
    if (LN > 0 && autoc)
        k_corr<R, LN, NPT, LN != 0, OCC><<<grid, 256, 0, st>>>(*P, Kb, (const CT *)Qt, (const CT *)Rt,
                                                          thetas, er, es, part, lo, hi, seg_base);
    else
        k_corr<R, LN, NPT, false, OCC><<<grid, 256, 0, st>>>(*P, Kb, (const CT *)Qt, (const CT *)Rt,
                                                        thetas, er, es, part, lo, hi, seg_base);
}
// FFT variants by log2 n (corr_use_fft(n): 16 <= n <= 512), LN = 0 the direct sum
template <typename R, int NPT, int OCC = 0>
static void launch_corr(bool autoc, const DevPlan *P, int Kb, const void *Qt, const void *Rt,
                        const double *thetas, const int *er, const int *es, double *part,
                        int lo, int hi, int seg_base, hipStream_t st) {
    const int ln = corr_use_fft(P->n) ? __builtin_ctz((unsigned)P->n) : 0;
#define CORR_ARGS autoc, P, Kb, Qt, Rt, thetas, er, es, part, lo, hi, seg_base, st
    switch (ln) {
    case 4: launch_corr_t<R, 4, NPT, OCC>(CORR_ARGS); break;
    case 5: launch_corr_t<R, 5, NPT, OCC>(CORR_ARGS); break;
    case 6: launch_corr_t<R, 6, NPT, OCC>(CORR_ARGS); break;
    case 7: launch_corr_t<R, 7, NPT, OCC>(CORR_ARGS); break;
    case 8: launch_corr_t<R, 8, NPT, OCC>(CORR_ARGS); break;
    case 9: launch_corr_t<R, 9, NPT, OCC>(CORR_ARGS); break;
    default: launch_corr_t<R, 0, NPT, OCC>(CORR_ARGS); break;
    }
#undef CORR_ARGS
}

// The correlation of a batch: the leading P->corr_auto_chunks chunks with the autocorrelation
// variant, the rest with the general one (each range segmented on its own, corr_segments).
extern "C" int blfid_dev_launch_corr(int prec, const DevPlan *P, int Kb, const void *Qt,
                                     const void *Rt, const double *thetas, const int *er,
                                     const int *es, double *part, hipStream_t st) {
    const int CH = corr_chunk(P->n);
    const int nchunk = (int)((P->Etot + CH - 1) / CH);
    const int A = corr_use_fft(P->n) ? min(max(P->corr_auto_chunks, 0), nchunk) : 0;
    const int segA = corr_range_segments(A);
    // four points per thread up to 1024 FID points, eight above (variants 0 and 1: variant 1
    // runs the complex64 four-point kernels at two waves per SIMD); variant 2: eight throughout
    for (int pass = 0; pass < 2; pass++) {
        const int lo = pass ? A : 0, hi = pass ? nchunk : A;
        if (hi <= lo) continue;
        const bool autoc = pass == 0;
        const int sb = pass ? segA : 0;
        if (P->npts <= 1024 && P->tn_corr != 2) {
            if (prec && P->tn_corr == 1)
                launch_corr<float, 4, 2>(autoc, P, Kb, Qt, Rt, thetas, er, es, part, lo, hi, sb, st);
            else if (prec) launch_corr<float, 4>(autoc, P, Kb, Qt, Rt, thetas, er, es, part, lo, hi, sb, st);
            else launch_corr<double, 4>(autoc, P, Kb, Qt, Rt, thetas, er, es, part, lo, hi, sb, st);
        } else {
            if (prec) launch_corr<float, 8>(autoc, P, Kb, Qt, Rt, thetas, er, es, part, lo, hi, sb, st);
            else launch_corr<double, 8>(autoc, P, Kb, Qt, Rt, thetas, er, es, part, lo, hi, sb, st);
        }
    }
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_rho_prep(int prec, const DevPlan *P, const DevPlan *PP, int Kb,
                                         const cd *Aprep, const int *sblk, const int *sloc,
                                         const int *nzi, const double *nzv, int nnz,
                                         void *rho_out, hipStream_t st) {
    const long n = (long)Kb * P->Etot;
    if (n <= 0) return 0;
    const int grid = (int)((n + 255) / 256);
    if (prec)
        k_rho_prep<float><<<grid, 256, 0, st>>>(*P, *PP, Kb, Aprep, P->elem_r, P->elem_s, sblk,
                                                 sloc, nzi, nzv, nnz, (cf *)rho_out);
    else
        k_rho_prep<double><<<grid, 256, 0, st>>>(*P, *PP, Kb, Aprep, P->elem_r, P->elem_s, sblk,
                                                  sloc, nzi, nzv, nnz, (cd *)rho_out);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_state_project(long K, int N, const cd *C, double inv_nrm, cd *S,
                                              hipStream_t st) {
    if (K <= 0) return 0;
    const unsigned grid = (unsigned)(K < 4096 ? K : 4096);
    k_state_project<<<grid, 256, 0, st>>>(K, (long)N * N, C, inv_nrm, S);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_state_fill(long K, int N, const cd *rho0, cd *S, hipStream_t st) {
    const long NN = (long)N * N, total = K * NN;
    if (total <= 0) return 0;
    k_state_fill<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(total, NN, rho0, S);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_state_apply(const DevPlan *PP, int Kb, int c0, const cd *Aprep,
                                            const int *sblk, const int *sloc, int N, int repeat,
                                            cd *S, cd *T, int nslots, hipStream_t st) {
    const int grid = Kb < nslots ? Kb : nslots;
    if (grid <= 0 || repeat <= 0) return 0;
    k_state_apply<<<grid, 256, 0, st>>>(*PP, Kb, c0, Aprep, sblk, sloc, N, repeat, S, T);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_state_trace(int K, int N, const cd *S, const int *nzi,
                                            const double *nzv, int nnz, double *part,
                                            hipStream_t st) {
    if (K <= 0) return 0;
    k_state_trace<<<(K + 255) / 256, 256, 0, st>>>(K, N, S, nzi, nzv, nnz, part);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_state_gather(int prec, const DevPlan *P, int Kb, int c0,
                                             const cd *S, void *rho_out, hipStream_t st) {
    const long n = (long)Kb * P->Etot;
    if (n <= 0) return 0;
    const int grid = (int)((n + 255) / 256);
    if (prec) k_state_gather<float><<<grid, 256, 0, st>>>(*P, Kb, c0, S, (cf *)rho_out);
    else k_state_gather<double><<<grid, 256, 0, st>>>(*P, Kb, c0, S, (cd *)rho_out);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_reduce(int Kb, int c0, int npts, int nparts, int G,
                                       const double *w, double scale, const double *part,
                                       double *grp, hipStream_t st) {
    if (Kb <= 0) return 0;
    const int n2 = 2 * npts;
    const int g0 = c0 / G, g1 = (c0 + Kb - 1) / G;
    dim3 grid((n2 + 255) / 256, g1 - g0 + 1);
    k_reduce<<<grid, 256, 0, st>>>(Kb, c0, npts, nparts, G, w, scale, part, grp);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_finish(int npts, int ngroups, const double *grp, double nre,
                                       double nim, double *out, hipStream_t st) {
    k_finish<<<(npts + 255) / 256, 256, 0, st>>>(npts, ngroups, grp, make_double2(nre, nim), out);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_zero(double *p, size_t n, hipStream_t st) {
    if (n == 0) return 0;
    size_t blocks = (n + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    k_zero<<<(int)blocks, 256, 0, st>>>(p, n);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_prep(int K, const double *th, const double *ph, const double *ga,
                                     double *RP, hipStream_t st) {
    if (K <= 0) return 0;
    k_prep<<<(K + 255) / 256, 256, 0, st>>>(K, th, ph, ga, RP);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_static_h(const DevPlan *P, int Kb, int c0, const int *grp,
                                         int ngrp, double *scratch, size_t slot, int nslots,
                                         cd *bufH, hipStream_t st) {
    long ntask = (long)Kb * ngrp;
    int grid = (int)(ntask < nslots ? ntask : nslots);
    if (grid <= 0) return 0;
    k_static_h<<<grid, 256, sizeof(double) * DEV_TERM_STRIDE * (size_t)P->nterms, st>>>(
        *P, Kb, c0, grp, ngrp, scratch, slot, bufH);
    return (int)hipGetLastError();
}

extern "C" int blfid_dev_launch_static_fid(int M, const DevPlan *P, int Kb, const int *pgrp,
                                           int ngrp, const cd *bufW, const double *lam,
                                           cd *scratch, size_t slot, int nslots, double w0,
                                           int prune, double *part, hipStream_t st) {
    long ntask = (long)Kb * ngrp;
    int grid = (int)(ntask < nslots ? ntask : nslots);
    if (grid <= 0) return 0;
    switch (M) {
    case 1: k_static_fid<1><<<grid, 64, 0, st>>>(*P, Kb, pgrp, ngrp, bufW, lam, scratch, slot, w0, prune, part); break;
    case 2: k_static_fid<2><<<grid, 256, 0, st>>>(*P, Kb, pgrp, ngrp, bufW, lam, scratch, slot, w0, prune, part); break;
    case 3: k_static_fid<3><<<grid, 256, 0, st>>>(*P, Kb, pgrp, ngrp, bufW, lam, scratch, slot, w0, prune, part); break;
    case 4: k_static_fid<4><<<grid, 256, 0, st>>>(*P, Kb, pgrp, ngrp, bufW, lam, scratch, slot, w0, prune, part); break;
    default: k_static_fid<5><<<grid, 256, 0, st>>>(*P, Kb, pgrp, ngrp, bufW, lam, scratch, slot, w0, prune, part); break;
    }
    return (int)hipGetLastError();
}
