/*
 * blfid_oracle.c -- FP64 CPU restatement of BlochLib's solid-2.0 powder FID path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP engine and the CPU
 * baseline timed beside it.  Nothing in the product (blochlib_amd/, include/) links, loads
 * or calls it; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
 *
 * It follows the reference algorithm step by step, in true FP64 throughout (the reference's
 * Hermitian QL runs in float internally, src/container/matrix/matdiagonalize.h:356-361):
 *
 *   spin operators      src/QMspins/singlespinop.cc:90-176, multispinop.cc:51-70
 *                       (spin 0 is the leftmost Kronecker factor, m=+1/2 first)
 *   rotations           src/QMspins/space_ten.cc:578-652 (rotationMatrix3D, 3- and 2-angle),
 *                       :717-748 (setPowderAngles(phi,theta,gamma)), :756-805 (setRotorAngles)
 *   CSA                 src/QMspins/csa.cc:223-329 (PAS, crystal-frame tensor, H=(-iso+zz)Iz),
 *                       spin_ten.cc:211-236, space_ten.cc:550-565 (A_CSA_pas)
 *   dipolar             src/QMspins/dip.cc:136-207, spin_ten.cc:431-452 (T_Dip homo/hetero)
 *   J                   src/QMspins/j.cc:129-137, spin_ten.cc:455-465 (T_J weak/strong)
 *   Hamiltonian(t1,t2)  src/QMspins/sys.cc:749-770 (midpoint rotor phase 2*pi*wr*tmid)
 *   Mexp                src/container/matrix/matexp.h:239-252 (V diag(chop(exp(mul*lambda))) V^H)
 *   Dyson + gamma-COMPUTE  src/QMspins/driver/compute.h:86-97, 213-235, 257-274, 294-461
 *   static FID          src/QMspins/driver/oneFID.h:292-336
 *   powder loop         src/QMspins/driver/oneFID.h:339-410 (w-weighted sum),
 *                       examples/solid-2.0/src/sequencerun_funcs.cc:690-704 (/Tr(D D^H))
 *
 * Linear algebra is written from the textbook methods, not from the reference's EISPACK
 * ports: Hermitian eig = Householder tridiagonalisation + implicit QL (the same method
 * family as the reference's cred/rred/complextqli, but in double); general eig of the
 * unitary period propagator = Householder-Hessenberg + single-shift complex QR (Schur).
 * For a normal matrix the Schur vectors are orthonormal eigenvectors, which is what the
 * reference's normalised comqr3 vectors approximate (matdiagonalize.h:907-914).
 */
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 2
#define ORC_ENOCONV 5

static const double PI2 = 6.283185307179586476925286766559;
static const double DEG2RAD = 0.017453292519943295769236907684886;

/* ------------------------------------------------------------------------------------ */
/* spin system                                                                          */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    int S, N;
    const int *iso; /* isotope id per spin; equal ids = homonuclear */
    int ncsa;
    const double *csa; /* per CSA: iso del eta spin alpha beta gamma (deg) */
    int ndip;
    const double *dip; /* per dipole: del on1 on2 alpha beta gamma (deg) */
    int nj;
    const double *jc; /* per J: iso on1 on2 strong */
} orc_sys;

/* rotationMatrix3D(a,b,g), space_ten.cc:578-618 (row-major 3x3) */
static void rot3(double a, double b, double g, double R[9]) {
    double cg = cos(g), cb = cos(b), ca = cos(a), sg = sin(g), sb = sin(b), sa = sin(a);
    R[0] = ca * cb * cg - sa * sg; R[1] = sa * cb * cg + ca * sg; R[2] = -sb * cg;
    R[3] = -ca * cb * sg - sa * cg; R[4] = -sa * cb * sg + ca * cg; R[5] = sb * sg;
    R[6] = ca * sb; R[7] = sa * sb; R[8] = cb;
}
/* rotationMatrix3D(phi,theta), space_ten.cc:620-652 */
static void rot2(double p, double t, double R[9]) {
    double ct = cos(t), cp = cos(p), sp = sin(p), st = sin(t);
    R[0] = cp; R[1] = sp; R[2] = 0.0;
    R[3] = -ct * sp; R[4] = cp * ct; R[5] = st;
    R[6] = sp * st; R[7] = -cp * st; R[8] = ct;
}
static void mm3(const double A[9], const double B[9], double C[9]) {
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += A[3 * i + k] * B[3 * k + j];
            T[3 * i + j] = s;
        }
    memcpy(C, T, sizeof(T));
}
static void mm3t(const double A[9], const double B[9], double C[9]) { /* A * B^T */
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += A[3 * i + k] * B[3 * j + k];
            T[3 * i + j] = s;
        }
    memcpy(C, T, sizeof(T));
}
/* R * diag(p) * R^T where R = rotationMatrix3D(euler deg), csa.cc:276-284 / dip.cc:159-167 */
static void crystal_tensor(double pxx, double pyy, double pzz, double a, double b, double g,
                           double C[9]) {
    double R[9], P[9] = {pxx, 0, 0, 0, pyy, 0, 0, 0, pzz}, T[9];
    rot3(a * DEG2RAD, b * DEG2RAD, g * DEG2RAD, R);
    mm3(R, P, T);
    mm3t(T, R, C);
}
/* (rotm * C * rotm^T)(2,2) */
static double zz_of(const double rotm[9], const double C[9]) {
    double T[9], O[9];
    mm3(rotm, C, T);
    mm3t(T, rotm, O);
    return O[8];
}

/* Zeeman-basis helpers: bit (S-1-i) of the state index is 1 when spin i has m=-1/2 */
static inline double mz(int S, int state, int spin) {
    return ((state >> (S - 1 - spin)) & 1) ? -0.5 : 0.5;
}

/* H(t1,t2,wr) for the given Rotations state.  rotb == 0 selects the cartPowder-only branch
 * of Csa::get_H / Dip::get_H (csa.cc:302-314, dip.cc:173-186).  H is col-major N x N. */
static void build_hamiltonian(const orc_sys *sy, const double cartPowder[9], double rotb_rad,
                              double rotor_phase, cplx *H) {
    const int S = sy->S, N = sy->N;
    double cartSpin[9], rotm[9];
    int beta_zero = (rotb_rad == 0.0);
    if (!beta_zero) {
        rot2(rotor_phase, rotb_rad, cartSpin);
        mm3(cartSpin, cartPowder, rotm);
    } else {
        memcpy(rotm, cartPowder, sizeof(rotm));
    }
    memset(H, 0, sizeof(cplx) * (size_t)N * N);
    const double sq3 = pow(3., 0.5);
    for (int c = 0; c < sy->ncsa; c++) {
        const double *p = sy->csa + 7 * c;
        double iso = p[0], del = p[1], eta = p[2];
        int on = (int)p[3];
        double isoterm = 0.0, zz = 0.0;
        if (iso != 0) {
            /* T00 * A_CSA_pas(iso,..,0,0) = (1/sqrt3) Iz * (-sqrt3 iso) */
            isoterm = (1. / sq3) * (-sq3 * iso);
        }
        if (!(iso != 0 && del == 0.0)) {
            if (!(beta_zero && eta == 0.0 && del == 0.0)) {
                double C[9];
                crystal_tensor(del * (eta - 1.0) / 2.0, -del * (eta + 1.0) / 2.0, 1.0 * del, p[4],
                               p[5], p[6], C);
                zz = zz_of(rotm, C);
            }
        }
        for (int s = 0; s < N; s++) H[s + (size_t)s * N] += (isoterm + zz) * mz(S, s, on);
    }
    for (int d = 0; d < sy->ndip; d++) {
        const double *p = sy->dip + 6 * d;
        double del = p[0];
        int i = (int)p[1], j = (int)p[2];
        if (beta_zero && del == 0.0) continue;
        double C[9], s6 = pow(6.0, 0.5);
        crystal_tensor(-0.5 * del * s6, -0.5 * del * s6, 1.0 * del * s6, p[3], p[4], p[5], C);
        double zz = zz_of(rotm, C);
        int homo = sy->iso[i] == sy->iso[j];
        if (homo) {
            double f = pow(1. / 6., 0.5);
            for (int s = 0; s < N; s++) {
                H[s + (size_t)s * N] += zz * f * 2.0 * mz(S, s, i) * mz(S, s, j);
                int bi = (s >> (S - 1 - i)) & 1, bj = (s >> (S - 1 - j)) & 1;
                if (bi != bj) { /* -(IxIx+IyIy) = -(1/2)(I+S- + I-S+) */
                    int t = s ^ (1 << (S - 1 - i)) ^ (1 << (S - 1 - j));
                    H[t + (size_t)s * N] += zz * f * (-0.5);
                }
            }
        } else {
            double f = pow(4. / 6., 0.5);
            for (int s = 0; s < N; s++)
                H[s + (size_t)s * N] += zz * f * mz(S, s, i) * mz(S, s, j);
        }
    }
    for (int q = 0; q < sy->nj; q++) {
        const double *p = sy->jc + 4 * q;
        double iso = p[0];
        int i = (int)p[1], j = (int)p[2], strong = (int)p[3];
        double f = iso * sq3 * (1. / sq3);
        for (int s = 0; s < N; s++) {
            H[s + (size_t)s * N] += f * mz(S, s, i) * mz(S, s, j);
            if (strong) {
                int bi = (s >> (S - 1 - i)) & 1, bj = (s >> (S - 1 - j)) & 1;
                if (bi != bj) {
                    int t = s ^ (1 << (S - 1 - i)) ^ (1 << (S - 1 - j));
                    H[t + (size_t)s * N] += f * 0.5;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* dense complex linear algebra (col-major)                                             */
/* ------------------------------------------------------------------------------------ */

#define EL(A, i, j, N) (A)[(size_t)(i) + (size_t)(j) * (N)]

/* C += A B, all N x N column-major: cache-blocked over k (panels of 128) with a 4 x 4 complex
 * register tile; each complex product a*b is a*Re(b) + swap(a)*(-Im b, +Im b) on AVX2 vectors
 * of two complex values.  Rows and columns beyond a multiple of 4 take a scalar loop. */
typedef double v4d __attribute__((vector_size(32)));
typedef long long v4l __attribute__((vector_size(32)));

static inline v4d ld4(const cplx *p) {
    v4d v;
    memcpy(&v, p, sizeof(v));
    return v;
}
static inline void st4(cplx *p, v4d v) { memcpy(p, &v, sizeof(v)); }

static void gemm_nn(int N, const cplx *A, const cplx *B, cplx *C) {
    const int KB = 128, N4 = N & ~3;
    const v4l sw = {1, 0, 3, 2};
    for (int k0 = 0; k0 < N; k0 += KB) {
        const int k1 = k0 + KB < N ? k0 + KB : N;
        for (int j = 0; j < N4; j += 4) {
            for (int i = 0; i < N4; i += 4) {
                v4d c[4][2];
                memset(c, 0, sizeof(c));
                for (int k = k0; k < k1; k++) {
                    const cplx *a = A + i + (size_t)k * N;
                    const v4d a0 = ld4(a), a1 = ld4(a + 2);
                    const v4d s0 = __builtin_shuffle(a0, sw), s1 = __builtin_shuffle(a1, sw);
                    for (int q = 0; q < 4; q++) {
                        const cplx b = B[k + (size_t)(j + q) * N];
                        const double br = creal(b), bi = cimag(b);
                        const v4d vr = {br, br, br, br}, vi = {-bi, bi, -bi, bi};
                        c[q][0] += a0 * vr + s0 * vi;
                        c[q][1] += a1 * vr + s1 * vi;
                    }
                }
                for (int q = 0; q < 4; q++) {
                    cplx *cc = C + i + (size_t)(j + q) * N;
                    st4(cc, ld4(cc) + c[q][0]);
                    st4(cc + 2, ld4(cc + 2) + c[q][1]);
                }
            }
            for (int i = N4; i < N; i++)
                for (int q = 0; q < 4; q++) {
                    cplx s = 0;
                    for (int k = k0; k < k1; k++) s += EL(A, i, k, N) * EL(B, k, j + q, N);
                    EL(C, i, j + q, N) += s;
                }
        }
        for (int j = N4; j < N; j++)
            for (int i = 0; i < N; i++) {
                cplx s = 0;
                for (int k = k0; k < k1; k++) s += EL(A, i, k, N) * EL(B, k, j, N);
                EL(C, i, j, N) += s;
            }
    }
}

/* C = op(A) * op(B); op = 'N' or 'C' (conjugate transpose).  C must not alias A or B.
 * A conjugate-transposed operand is formed explicitly (O(N^2)) and the product taken by
 * gemm_nn. */
static void zgemm(int N, char ta, char tb, const cplx *A, const cplx *B, cplx *C) {
    cplx *at = NULL, *bt = NULL;
    if (ta == 'C') {
        at = malloc(sizeof(cplx) * (size_t)N * N);
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) EL(at, i, j, N) = conj(EL(A, j, i, N));
        A = at;
    }
    if (tb == 'C') {
        bt = malloc(sizeof(cplx) * (size_t)N * N);
        for (int j = 0; j < N; j++)
            for (int i = 0; i < N; i++) EL(bt, i, j, N) = conj(EL(B, j, i, N));
        B = bt;
    }
    memset(C, 0, sizeof(cplx) * (size_t)N * N);
    gemm_nn(N, A, B, C);
    free(at);
    free(bt);
}

/* Hermitian eigen-decomposition: A = V diag(w) V^H.  Householder tridiagonalisation,
 * diagonal phase scaling to a real tridiagonal, implicit QL with Wilkinson-type shifts. */
int orc_eig_herm(int N, const cplx *Ain, double *w, cplx *V) {
    cplx *A = malloc(sizeof(cplx) * (size_t)N * N);
    cplx *v = malloc(sizeof(cplx) * N), *p = malloc(sizeof(cplx) * N);
    double *e = malloc(sizeof(double) * N);
    if (!A || !v || !p || !e) { free(A); free(v); free(p); free(e); return ORC_ENOMEM; }
    memcpy(A, Ain, sizeof(cplx) * (size_t)N * N);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) EL(V, i, j, N) = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k + 2 < N; k++) {
        int m = N - k - 1; /* length of x = A[k+1.., k] */
        double xn = 0;
        for (int i = 0; i < m; i++) xn += creal(EL(A, k + 1 + i, k, N) * conj(EL(A, k + 1 + i, k, N)));
        xn = sqrt(xn);
        if (xn == 0.0) continue;
        cplx x0 = EL(A, k + 1, k, N);
        cplx ph = (cabs(x0) > 0) ? x0 / cabs(x0) : 1.0;
        cplx alpha = -ph * xn;
        for (int i = 0; i < m; i++) v[i] = EL(A, k + 1 + i, k, N);
        v[0] -= alpha;
        double vn = 0;
        for (int i = 0; i < m; i++) vn += creal(v[i] * conj(v[i]));
        vn = sqrt(vn);
        if (vn == 0.0) continue;
        for (int i = 0; i < m; i++) v[i] /= vn;
        /* trailing block T = A[k+1..,k+1..]: T <- T - 2 v w^H - 2 w v^H, w = p - (v^H p) v */
        for (int i = 0; i < m; i++) {
            cplx s = 0;
            for (int j = 0; j < m; j++) s += EL(A, k + 1 + i, k + 1 + j, N) * v[j];
            p[i] = s;
        }
        cplx K = 0;
        for (int i = 0; i < m; i++) K += conj(v[i]) * p[i];
        for (int i = 0; i < m; i++) p[i] -= K * v[i];
        for (int j = 0; j < m; j++)
            for (int i = 0; i < m; i++)
                EL(A, k + 1 + i, k + 1 + j, N) -= 2.0 * (v[i] * conj(p[j]) + p[i] * conj(v[j]));
        for (int i = 0; i < m; i++) {
            EL(A, k + 1 + i, k, N) = (i == 0) ? alpha : 0.0;
            EL(A, k, k + 1 + i, N) = (i == 0) ? conj(alpha) : 0.0;
        }
        /* V <- V (I - 2 v v^H) on columns k+1.. */
        for (int r = 0; r < N; r++) {
            cplx s = 0;
            for (int j = 0; j < m; j++) s += EL(V, r, k + 1 + j, N) * v[j];
            for (int j = 0; j < m; j++) EL(V, r, k + 1 + j, N) -= 2.0 * s * conj(v[j]);
        }
    }
    /* real tridiagonal via diagonal phases: V <- V Phi */
    cplx phi = 1.0;
    for (int k = 0; k < N; k++) {
        w[k] = creal(EL(A, k, k, N));
        if (k > 0) {
            cplx sub = EL(A, k, k - 1, N);
            double as = cabs(sub);
            e[k - 1] = as;
            if (as > 0) phi = phi * sub / as;
            for (int r = 0; r < N; r++) EL(V, r, k, N) *= phi;
        }
    }
    if (N > 0) e[N - 1] = 0.0;
    /* implicit QL on (w, e), rotations applied to the columns of V */
    for (int l = 0; l < N; l++) {
        int iter = 0, m;
        do {
            for (m = l; m < N - 1; m++) {
                double dd = fabs(w[m]) + fabs(w[m + 1]);
                if (fabs(e[m]) <= 1e-300 || fabs(e[m]) <= 2.220446049250313e-16 * dd) break;
            }
            if (m != l) {
                if (iter++ == 60 * N + 60) { free(A); free(v); free(p); free(e); return ORC_ENOCONV; }
                double g = (w[l + 1] - w[l]) / (2.0 * e[l]);
                double r = hypot(g, 1.0);
                g = w[m] - w[l] + e[l] / (g + copysign(r, g));
                double s = 1.0, c = 1.0, pp = 0.0;
                int i;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i], b = c * e[i];
                    e[i + 1] = (r = hypot(f, g));
                    if (r == 0.0) {
                        w[i + 1] -= pp;
                        e[m] = 0.0;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = w[i + 1] - pp;
                    r = (w[i] - g) * s + 2.0 * c * b;
                    pp = s * r;
                    w[i + 1] = g + pp;
                    g = c * r - b;
                    for (int k = 0; k < N; k++) {
                        cplx fz = EL(V, k, i + 1, N);
                        EL(V, k, i + 1, N) = s * EL(V, k, i, N) + c * fz;
                        EL(V, k, i, N) = c * EL(V, k, i, N) - s * fz;
                    }
                }
                if (r == 0.0 && i >= l) continue;
                w[l] -= pp;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    free(A); free(v); free(p); free(e);
    return ORC_OK;
}

/* general complex eigen-decomposition for a normal matrix: complex Schur form A = Q T Q^H via
 * Householder-Hessenberg + single-shift QR (Wilkinson shift, exceptional shifts every 10
 * iterations).  lam = diag(T); W = Q (orthonormal eigenvectors when A is normal). */
int orc_eig_normal(int N, const cplx *Ain, cplx *lam, cplx *Q) {
    cplx *A = malloc(sizeof(cplx) * (size_t)N * N), *v = malloc(sizeof(cplx) * N);
    if (!A || !v) { free(A); free(v); return ORC_ENOMEM; }
    memcpy(A, Ain, sizeof(cplx) * (size_t)N * N);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) EL(Q, i, j, N) = (i == j) ? 1.0 : 0.0;
    for (int k = 0; k + 2 < N; k++) {
        int m = N - k - 1;
        double xn = 0;
        for (int i = 0; i < m; i++) xn += creal(EL(A, k + 1 + i, k, N) * conj(EL(A, k + 1 + i, k, N)));
        xn = sqrt(xn);
        if (xn == 0.0) continue;
        cplx x0 = EL(A, k + 1, k, N);
        cplx ph = (cabs(x0) > 0) ? x0 / cabs(x0) : 1.0;
        cplx alpha = -ph * xn;
        for (int i = 0; i < m; i++) v[i] = EL(A, k + 1 + i, k, N);
        v[0] -= alpha;
        double vn = 0;
        for (int i = 0; i < m; i++) vn += creal(v[i] * conj(v[i]));
        vn = sqrt(vn);
        if (vn == 0.0) continue;
        for (int i = 0; i < m; i++) v[i] /= vn;
        /* A <- H A : rows k+1.., columns k.. */
        for (int j = k; j < N; j++) {
            cplx s = 0;
            for (int i = 0; i < m; i++) s += conj(v[i]) * EL(A, k + 1 + i, j, N);
            for (int i = 0; i < m; i++) EL(A, k + 1 + i, j, N) -= 2.0 * v[i] * s;
        }
        /* A <- A H : all rows, columns k+1.. ; Q <- Q H */
        for (int r = 0; r < N; r++) {
            cplx s = 0, t = 0;
            for (int j = 0; j < m; j++) {
                s += EL(A, r, k + 1 + j, N) * v[j];
                t += EL(Q, r, k + 1 + j, N) * v[j];
            }
            for (int j = 0; j < m; j++) {
                EL(A, r, k + 1 + j, N) -= 2.0 * s * conj(v[j]);
                EL(Q, r, k + 1 + j, N) -= 2.0 * t * conj(v[j]);
            }
        }
        for (int i = 1; i < m; i++) EL(A, k + 1 + i, k, N) = 0.0;
    }
    const double eps = 2.220446049250313e-16;
    int hi = N - 1, iter = 0, total = 0;
    while (hi > 0) {
        int l;
        for (l = hi; l > 0; l--) {
            double s = cabs(EL(A, l - 1, l - 1, N)) + cabs(EL(A, l, l, N));
            if (s == 0.0) s = 1.0;
            if (cabs(EL(A, l, l - 1, N)) <= eps * s) {
                EL(A, l, l - 1, N) = 0.0;
                break;
            }
        }
        if (l == hi) { hi--; iter = 0; continue; }
        if (++total > 100 * N + 100) { free(A); free(v); return ORC_ENOCONV; }
        iter++;
        cplx mu;
        if (iter % 10 == 0) {
            mu = EL(A, hi, hi, N) + cabs(EL(A, hi, hi - 1, N));
        } else { /* eigenvalue of the trailing 2x2 closer to A[hi][hi] */
            cplx a = EL(A, hi - 1, hi - 1, N), b = EL(A, hi - 1, hi, N), c = EL(A, hi, hi - 1, N),
                 d = EL(A, hi, hi, N);
            cplx tr = a + d, det = a * d - b * c;
            cplx disc = csqrt(tr * tr / 4.0 - det);
            cplx l1 = tr / 2.0 + disc, l2 = tr / 2.0 - disc;
            mu = (cabs(l1 - d) < cabs(l2 - d)) ? l1 : l2;
        }
        cplx x = EL(A, l, l, N) - mu, y = EL(A, l + 1, l, N);
        for (int k = l; k < hi; k++) {
            double ax = cabs(x), ay = cabs(y), r = hypot(ax, ay);
            double c;
            cplx s;
            if (r == 0.0) { c = 1.0; s = 0.0; }
            else if (ax == 0.0) { c = 0.0; s = 1.0; }
            else { c = ax / r; s = (x / ax) * conj(y) / r; }
            /* rows k,k+1 <- G rows, G = [[c, s],[-conj(s), c]] */
            for (int j = (k > l ? k - 1 : l); j < N; j++) {
                cplx a1 = EL(A, k, j, N), a2 = EL(A, k + 1, j, N);
                EL(A, k, j, N) = c * a1 + s * a2;
                EL(A, k + 1, j, N) = -conj(s) * a1 + c * a2;
            }
            /* columns k,k+1 <- cols G^H ; same for Q */
            int rmax = (k + 2 < hi) ? k + 2 : hi;
            for (int i = 0; i <= rmax; i++) {
                cplx a1 = EL(A, i, k, N), a2 = EL(A, i, k + 1, N);
                EL(A, i, k, N) = c * a1 + conj(s) * a2;
                EL(A, i, k + 1, N) = -s * a1 + c * a2;
            }
            for (int i = 0; i < N; i++) {
                cplx a1 = EL(Q, i, k, N), a2 = EL(Q, i, k + 1, N);
                EL(Q, i, k, N) = c * a1 + conj(s) * a2;
                EL(Q, i, k + 1, N) = -s * a1 + c * a2;
            }
            if (k + 1 < hi) {
                x = EL(A, k + 1, k, N);
                y = EL(A, k + 2, k, N);
            }
        }
    }
    for (int i = 0; i < N; i++) lam[i] = EL(A, i, i, N);
    free(A); free(v);
    return ORC_OK;
}

/* chop(z, eps): zero real/imag parts below eps, complex.h:1001-1002 */
static inline cplx chopc(cplx z, double eps) {
    double re = creal(z), im = cimag(z);
    return CMPLX(fabs(re) < eps ? 0.0 : re, fabs(im) < eps ? 0.0 : im);
}

/* Mexp(H, mul) for Hermitian H: V diag(chop(exp(mul*lambda), 1e-12)) V^H, matexp.h:239-252 */
int orc_expm_herm(int N, const cplx *H, cplx mul, cplx *U) {
    double *w = malloc(sizeof(double) * N);
    cplx *V = malloc(sizeof(cplx) * (size_t)N * N), *T = malloc(sizeof(cplx) * (size_t)N * N);
    if (!w || !V || !T) { free(w); free(V); free(T); return ORC_ENOMEM; }
    int st = orc_eig_herm(N, H, w, V);
    if (st) { free(w); free(V); free(T); return st; }
    for (int j = 0; j < N; j++) {
        cplx d = chopc(cexp(mul * w[j]), 1e-12);
        for (int i = 0; i < N; i++) EL(T, i, j, N) = EL(V, i, j, N) * d;
    }
    zgemm(N, 'N', 'C', T, V, U);
    free(w); free(V); free(T);
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------ */
/* gamma-COMPUTE single-crystallite FID, compute.h:257-274 + 308-461                    */
/* ------------------------------------------------------------------------------------ */

typedef struct {
    const orc_sys *sy;
    double wr, rotb_rad, sw, t0;
    int n, npts, rosym;
    const cplx *rho, *det;
} orc_spin_cfg;

static int spinning_fid_one(const orc_spin_cfg *cf, double theta, double phi, double gamma,
                            cplx *fid) {
    const int N = cf->sy->N, n = cf->n, P = cf->npts;
    const size_t NN = (size_t)N * N;
    double cartPowder[9];
    rot3(phi, theta, gamma, cartPowder); /* setPowderAngles(th,ph,ga) -> (phi,theta,gamma) */
    double tmin = cf->t0, tmax = cf->t0 + 1.0 / cf->wr;
    double tau = (tmax - tmin) / n;
    cplx *As = malloc(sizeof(cplx) * NN * (n + 1));
    cplx *H = malloc(sizeof(cplx) * NN), *hh = malloc(sizeof(cplx) * NN);
    cplx *W = malloc(sizeof(cplx) * NN), *lam = malloc(sizeof(cplx) * N);
    cplx *Qs = malloc(sizeof(cplx) * NN * n), *Rs = malloc(sizeof(cplx) * NN * n);
    cplx *T1 = malloc(sizeof(cplx) * NN), *T2 = malloc(sizeof(cplx) * NN);
    cplx *Fave = malloc(sizeof(cplx) * NN * n), *wrs = malloc(sizeof(cplx) * NN);
    cplx *ev = malloc(sizeof(cplx) * N);
    int st = ORC_OK;
    if (!As || !H || !hh || !W || !lam || !Qs || !Rs || !T1 || !T2 || !Fave || !wrs || !ev) {
        st = ORC_ENOMEM;
        goto done;
    }
    /* calcUFID(0): gammaloop == 1 since gamma_step (10) < n is the only case driven here */
    {
        double t1 = tmin, t2 = tmin + tau;
        for (size_t i = 0; i < NN; i++) As[i] = 0;
        for (int i = 0; i < N; i++) As[i + (size_t)i * N] = 1.0;
        for (int i = 0; i < n; i++) {
            double tmid = (t2 - t1) / 2.0 + t1;
            build_hamiltonian(cf->sy, cartPowder, cf->rotb_rad, tmid * cf->wr * PI2, H);
            st = orc_expm_herm(N, H, CMPLX(0.0, -1.0) * tau * PI2, hh);
            if (st) goto done;
            zgemm(N, 'N', 'N', hh, As + NN * i, As + NN * (i + 1));
            t1 += tau;
            t2 += tau;
        }
    }
    /* calcFID */
    st = orc_eig_normal(N, As + NN * n, lam, W);
    if (st) goto done;
    {
        cplx tott = CMPLX(0.0, (double)n * tau);
        for (int i = 0; i < N; i++) ev[i] = CMPLX(log(cabs(lam[i])), carg(lam[i]));
        for (int i = 0; i < N; i++)
            for (int j = 0; j < N; j++) EL(wrs, i, j, N) = chopc((ev[i] - ev[j]) / tott, 1e-10);
    }
    for (int i = 0; i < n; i++) {
        const cplx *A = As + NN * (i + 1);
        zgemm(N, 'N', 'N', cf->det, A, T1); /* det*A */
        zgemm(N, 'C', 'N', A, T1, T2);      /* A^H det A */
        zgemm(N, 'N', 'N', T2, W, T1);
        zgemm(N, 'C', 'N', W, T1, Qs + NN * i);
        if (cf->rosym == 0) {
            zgemm(N, 'N', 'N', cf->rho, A, T1);
            zgemm(N, 'C', 'N', A, T1, T2);
            zgemm(N, 'N', 'N', T2, W, T1);
            zgemm(N, 'C', 'N', W, T1, Rs + NN * i);
        } else {
            const cplx *Q = Qs + NN * i;
            cplx *R = Rs + NN * i;
            for (int a = 0; a < N; a++)
                for (int b = 0; b < N; b++) EL(R, a, b, N) = EL(Q, a, b, N) + conj(EL(Q, b, a, N));
        }
    }
    /* Fave initialised to Fz (compute.h:403) */
    for (int p = 0; p < n; p++) {
        cplx *F = Fave + NN * p;
        memset(F, 0, sizeof(cplx) * NN);
        for (int s = 0; s < N; s++) {
            double fz = 0;
            for (int q = 0; q < cf->sy->S; q++) fz += mz(cf->sy->S, s, q);
            EL(F, s, s, N) = fz;
        }
    }
    for (int i = 0; i < n; i++) {
        for (int p = 0; p < n; p++) {
            int ind1 = (i + p) % n;
            int ind2 = (i + p >= n) ? p - n : p;
            cplx tmp1 = CMPLX(0.0, -(double)ind2 * tau);
            const cplx *Q = Qs + NN * ind1, *R = Rs + NN * i;
            cplx *F = Fave + NN * p;
            for (int r = 0; r < N; r++)
                for (int s = 0; s < N; s++)
                    EL(F, r, s, N) += EL(Q, r, s, N) * EL(R, s, r, N) * cexp(tmp1 * EL(wrs, r, s, N));
        }
    }
    {
        cplx ti = CMPLX(0.0, tau);
        for (size_t q = 0; q < NN; q++) wrs[q] = cexp(ti * wrs[q]);
        memcpy(T1, wrs, sizeof(cplx) * NN); /* tmpmm */
        for (int t = 0; t < P; t++) fid[t] = 0;
        if (P > 0)
            for (int i = 0; i < N; i++)
                for (int j = 0; j < N; j++) fid[0] += EL(Fave, i, j, N);
        for (int t = 1; t < P; t++) {
            int p = t % n;
            const cplx *F = Fave + NN * p;
            for (int r = 0; r < N; r++)
                for (int s = 0; s < N; s++) {
                    fid[t] += EL(F, r, s, N) * EL(T1, r, s, N);
                    EL(T1, r, s, N) *= EL(wrs, r, s, N);
                }
        }
        int ff = cf->rosym == 1 ? 2 : 1;
        double scale = 1. / (double)(n / ff);
        for (int t = 0; t < P; t++) fid[t] *= scale;
    }
done:
    free(As); free(H); free(hh); free(W); free(lam); free(Qs); free(Rs);
    free(T1); free(T2); free(Fave); free(wrs); free(ev);
    return st;
}

/* static FID, oneFID.h:292-336 */
typedef struct {
    const orc_sys *sy;
    double rotb_rad, sw;
    int npts;
    const cplx *rho, *det;
} orc_static_cfg;

static int static_fid_one(const orc_static_cfg *cf, double theta, double phi, double gamma,
                          cplx *fid) {
    const int N = cf->sy->N, P = cf->npts;
    const size_t NN = (size_t)N * N;
    double dt = 1. / cf->sw, cartPowder[9];
    rot3(phi, theta, gamma, cartPowder);
    cplx *H = malloc(sizeof(cplx) * NN), *V = malloc(sizeof(cplx) * NN);
    cplx *T = malloc(sizeof(cplx) * NN), *sig0 = malloc(sizeof(cplx) * NN),
         *Do = malloc(sizeof(cplx) * NN);
    cplx *Aa = malloc(sizeof(cplx) * NN), *Bb = malloc(sizeof(cplx) * NN), *e = malloc(sizeof(cplx) * N);
    double *w = malloc(sizeof(double) * N);
    int st = ORC_OK;
    if (!H || !V || !T || !sig0 || !Do || !Aa || !Bb || !e || !w) { st = ORC_ENOMEM; goto done; }
    /* Hamiltonian(0, dt) with wr = 0: rotor phase 0 */
    build_hamiltonian(cf->sy, cartPowder, cf->rotb_rad, 0.0, H);
    st = orc_eig_herm(N, H, w, V);
    if (st) goto done;
    zgemm(N, 'N', 'N', cf->rho, V, T);
    zgemm(N, 'C', 'N', V, T, sig0);
    zgemm(N, 'N', 'N', cf->det, V, T);
    zgemm(N, 'C', 'N', V, T, Do);
    cplx z = CMPLX(0.0, -dt * 2.0 * 3.14159265358979323846);
    for (int i = 0; i < N; i++) e[i] = cexp(z * w[i]);
    int pos = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            Aa[pos] = EL(Do, i, j, N) * EL(sig0, j, i, N);
            Bb[pos] = conj(e[i]) * e[j];
            double a2 = creal(Aa[pos]) * creal(Aa[pos]) + cimag(Aa[pos]) * cimag(Aa[pos]);
            if (a2 > 1.0e-10) pos++;
        }
    for (int k = 0; k < P; k++) {
        cplx s = 0;
        for (int q = 0; q < pos; q++) {
            s += Aa[q];
            Aa[q] *= Bb[q];
        }
        fid[k] = s;
    }
done:
    free(H); free(V); free(T); free(sig0); free(Do); free(Aa); free(Bb); free(e); free(w);
    return st;
}

/* ------------------------------------------------------------------------------------ */
/* powder driver (threads over crystallites; deterministic ordered final sum)          */
/* ------------------------------------------------------------------------------------ */

/* pulses applied to rho before a gamma-COMPUTE FID (SequenceRun::doFID, sequencerun_funcs.cc:
 * 693-703): rho <- U rho U^H with U the sub-sequence propagator of each crystallite */
typedef struct {
    double wr, rotb_rad;
    int nsteps, nrf;
    const double *step_t1, *step_dt, *rf;
} orc_prep;

typedef struct {
    int spinning;
    orc_spin_cfg sc;
    orc_static_cfg tc;
    const orc_prep *prep; /* NULL: rho as given */
    const cplx *rho_all;  /* non-NULL: crystallite k's rho is rho_all + k N^2 */
    const double *theta, *phi, *gamma, *wt;
    int k0, k1, P;
    cplx *acc; /* P */
    int status;
} orc_job;

static int pulse_propagator(const orc_sys *sy, const double cp[9], double wr, double rotb_rad,
                            double t0, double tau, int nsteps, const double *step_t1,
                            const double *step_dt, const double *rf, int nrf, cplx *U, cplx *H,
                            cplx *X, cplx *T);

/* rho_c = U rho U^H for one crystallite, and the reference's exact isroSYM test of rho_c */
static int prep_rho(const orc_prep *pp, const orc_sys *sy, double theta, double phi, double gamma,
                    const cplx *rho, const cplx *det, cplx *rc, int *rosym) {
    const int N = sy->N;
    const size_t NN = (size_t)N * N;
    cplx *U = malloc(sizeof(cplx) * NN), *H = malloc(sizeof(cplx) * NN);
    cplx *X = malloc(sizeof(cplx) * NN), *T = malloc(sizeof(cplx) * NN);
    int st = ORC_ENOMEM;
    if (U && H && X && T) {
        double cp[9];
        rot3(phi, theta, gamma, cp);
        st = pulse_propagator(sy, cp, pp->wr, pp->rotb_rad, 0.0, 0.0, pp->nsteps, pp->step_t1,
                              pp->step_dt, pp->rf, pp->nrf, U, H, X, T);
        if (st == ORC_OK) {
            zgemm(N, 'N', 'C', rho, U, T); /* rho U^H */
            zgemm(N, 'N', 'N', U, T, rc);  /* U rho U^H */
            int sym = 1;
            for (int i = 0; i < N && sym; i++)
                for (int j = 0; j < N; j++)
                    if (EL(rc, i, j, N) != 0.5 * (EL(det, i, j, N) + conj(EL(det, j, i, N)))) {
                        sym = 0;
                        break;
                    }
            *rosym = sym;
        }
    }
    free(U); free(H); free(X); free(T);
    return st;
}

static void *job_run(void *arg) {
    orc_job *jb = (orc_job *)arg;
    cplx *f = malloc(sizeof(cplx) * (jb->P > 0 ? jb->P : 1));
    if (!f) { jb->status = ORC_ENOMEM; return NULL; }
    for (int t = 0; t < jb->P; t++) jb->acc[t] = 0;
    cplx *rc = NULL;
    if (jb->prep) {
        const size_t N = (size_t)jb->sc.sy->N;
        rc = malloc(sizeof(cplx) * N * N);
        if (!rc) { free(f); jb->status = ORC_ENOMEM; return NULL; }
    }
    for (int k = jb->k0; k < jb->k1; k++) {
        int st;
        if (jb->rho_all) {
            /* the crystallite's own rho, with the reference's exact isroSYM test on it */
            const int N = jb->spinning ? jb->sc.sy->N : jb->tc.sy->N;
            const cplx *rk = jb->rho_all + (size_t)k * N * N;
            if (jb->spinning) {
                orc_spin_cfg cf = jb->sc;
                int sym = 1;
                for (int i = 0; i < N && sym; i++)
                    for (int j = 0; j < N; j++)
                        if (EL(rk, i, j, N) != 0.5 * (EL(cf.det, i, j, N) + conj(EL(cf.det, j, i, N)))) {
                            sym = 0;
                            break;
                        }
                cf.rho = rk;
                cf.rosym = sym;
                st = spinning_fid_one(&cf, jb->theta[k], jb->phi[k], jb->gamma[k], f);
            } else {
                orc_static_cfg tf = jb->tc;
                tf.rho = rk;
                st = static_fid_one(&tf, jb->theta[k], jb->phi[k], jb->gamma[k], f);
            }
        } else if (jb->prep) {
            orc_spin_cfg cf = jb->sc;
            st = prep_rho(jb->prep, cf.sy, jb->theta[k], jb->phi[k], jb->gamma[k], cf.rho, cf.det,
                          rc, &cf.rosym);
            cf.rho = rc;
            if (!st) st = spinning_fid_one(&cf, jb->theta[k], jb->phi[k], jb->gamma[k], f);
        } else {
            st = jb->spinning ? spinning_fid_one(&jb->sc, jb->theta[k], jb->phi[k], jb->gamma[k], f)
                              : static_fid_one(&jb->tc, jb->theta[k], jb->phi[k], jb->gamma[k], f);
        }
        if (st) { jb->status = st; break; }
        for (int t = 0; t < jb->P; t++) jb->acc[t] += jb->wt[k] * f[t];
    }
    free(rc);
    free(f);
    return NULL;
}

/* Exported entry: powder FID with BlochLib's solid-2.0 semantics.
 *   isotopes[S]; csa[7*ncsa]; dip[6*ndip]; jc[4*nj]   (see orc_sys)
 *   wr (Hz), rotor (deg), sw (Hz), npts, t0 (s)
 *   theta/phi/gamma (rad), w[K]       (powder file columns: phi theta gamma w, powder.cc:333-389)
 *   rho/det: N x N complex, column-major, interleaved re/im
 *   normalize: 1 -> divide by Tr(det det^H) (sequencerun_funcs.cc:698-702); 0 -> oneFID.h:396-402
 *   fid_out: 2*npts doubles; sw_out: adjusted sweep width (compute.h:96) */
static int powder_fid_impl(const orc_prep *prep, const double *rho_all, int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                   const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                   double sw, int npts, double t0, int K, const double *theta, const double *phi,
                   const double *gamma, const double *w, const double *rho, const double *det,
                   int normalize, int nthreads, double *fid_out, double *sw_out) {
    if (S < 1 || S > 12 || npts < 0 || K < 0) return ORC_EINVAL;
    orc_sys sy = {S, 1 << S, isotopes, ncsa, csa, ndip, dip, nj, jc};
    const int N = sy.N;
    const cplx *R = (const cplx *)rho, *D = (const cplx *)det;
    int spinning = !(wr == 0.0 || rotor_deg == 0.0);
    orc_spin_cfg sc;
    memset(&sc, 0, sizeof(sc));
    orc_static_cfg tc = {&sy, rotor_deg * DEG2RAD, sw, npts, R, D};
    double swa = sw;
    if (spinning) {
        int n = (int)floor(sw / wr + 0.5);
        if (n < 1) return ORC_EINVAL;
        swa = (double)(n * wr);
        /* isroSYM: exact element-wise ro == 0.5*(det+adjoint(det)) (compute.h:75-83) */
        int rosym = 1;
        for (int i = 0; i < N && rosym; i++)
            for (int j = 0; j < N; j++)
                if (EL(R, i, j, N) != 0.5 * (EL(D, i, j, N) + conj(EL(D, j, i, N)))) { rosym = 0; break; }
        sc.sy = &sy; sc.wr = wr; sc.rotb_rad = rotor_deg * DEG2RAD; sc.sw = swa; sc.t0 = t0;
        sc.n = n; sc.npts = npts; sc.rosym = rosym; sc.rho = R; sc.det = D;
    }
    if (sw_out) *sw_out = swa;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > K && K > 0) nthreads = K;
    if (K == 0) nthreads = 1;
    orc_job *jobs = calloc(nthreads, sizeof(orc_job));
    pthread_t *th = calloc(nthreads, sizeof(pthread_t));
    cplx *accs = calloc((size_t)nthreads * (npts > 0 ? npts : 1), sizeof(cplx));
    if (!jobs || !th || !accs) { free(jobs); free(th); free(accs); return ORC_ENOMEM; }
    for (int t = 0; t < nthreads; t++) {
        jobs[t].spinning = spinning; jobs[t].sc = sc; jobs[t].tc = tc; jobs[t].prep = prep;
        jobs[t].rho_all = (const cplx *)rho_all;
        jobs[t].theta = theta; jobs[t].phi = phi; jobs[t].gamma = gamma; jobs[t].wt = w;
        jobs[t].k0 = (int)((long)K * t / nthreads); jobs[t].k1 = (int)((long)K * (t + 1) / nthreads);
        jobs[t].P = npts; jobs[t].acc = accs + (size_t)t * (npts > 0 ? npts : 1);
    }
    for (int t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, job_run, &jobs[t]);
    job_run(&jobs[0]);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
    int st = ORC_OK;
    for (int t = 0; t < nthreads; t++) if (jobs[t].status) st = jobs[t].status;
    cplx trdd = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) trdd += EL(D, i, j, N) * conj(EL(D, i, j, N));
    for (int p = 0; p < npts; p++) {
        cplx s = 0;
        for (int t = 0; t < nthreads; t++) s += jobs[t].acc[p];
        if (normalize) s /= trdd;
        fid_out[2 * p] = creal(s);
        fid_out[2 * p + 1] = cimag(s);
    }
    free(jobs); free(th); free(accs);
    return st;
}

int orc_powder_fid(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                   const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                   double sw, int npts, double t0, int K, const double *theta, const double *phi,
                   const double *gamma, const double *w, const double *rho, const double *det,
                   int normalize, int nthreads, double *fid_out, double *sw_out) {
    return powder_fid_impl(NULL, NULL, S, isotopes, ncsa, csa, ndip, dip, nj, jc, wr, rotor_deg, sw,
                           npts, t0, K, theta, phi, gamma, w, rho, det, normalize, nthreads,
                           fid_out, sw_out);
}

/* Pulses before a gamma-COMPUTE fid() (SequenceRun::doFID, sequencerun_funcs.cc:693-703): per
 * crystallite rho <- U rho U^H with U the propagator of the nsteps pulse steps (lengths
 * step_dt, start times step_t1, rf table rf as orc_pulse_fid, rotor angle prep_rotor_deg, the
 * FID's wr), then the FID of that rho from start time t0 (oneFID::setStartTime(currentTime),
 * sequencerun_funcs.cc:611), the isroSYM test applied to the prepared rho as the reference does.
 * wr and rotor_deg must describe a spinning FID. */
int orc_prep_fid(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                 const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                 double sw, int npts, double t0, int nsteps, const double *step_t1,
                 const double *step_dt, double prep_rotor_deg, int nrf, const double *rf, int K,
                 const double *theta, const double *phi, const double *gamma, const double *w,
                 const double *rho, const double *det, int normalize, int nthreads,
                 double *fid_out, double *sw_out) {
    if (wr == 0.0 || rotor_deg == 0.0 || nsteps < 1 || !step_dt) return ORC_EINVAL;
    orc_prep pp = {wr, prep_rotor_deg * DEG2RAD, nsteps, nrf, step_t1, step_dt, rf};
    return powder_fid_impl(&pp, NULL, S, isotopes, ncsa, csa, ndip, dip, nj, jc, wr, rotor_deg, sw,
                           npts, t0, K, theta, phi, gamma, w, rho, det, normalize, nthreads,
                           fid_out, sw_out);
}

/* ------------------------------------------------------------------------------------ */
/* point-to-point FID with rf (solid-2.0 "ptop" over a rotor-period sub-sequence)       */
/* ------------------------------------------------------------------------------------ */
/* Propogation::propogator (examples/solid-2.0/src/propogation.cc:240-304): for each step
 * k of the period, hh = SolidSys::Hamiltonian(t1,t2,wr) + PulseData::Pulse (pulsedata.cc:
 * 110-125: amp (cos(ph) Ix_i + sin(ph) Iy_i) + offset Iz_i on every spin of the pulsed
 * isotope, only for channels with amp != 0), U *= Mexp(hh, +i tau 2 pi) with U *= X meaning U = X U.  The ptop doFID branch
 * (sequencerun_funcs.cc:641-652 + sequencerun.cc:164-196) records Tr(ro det) then applies
 * ro <- U ro U^H once per point.  rf[k*4*niso + 4*q + {0,1,2,3}] = isotope id, amp (Hz),
 * phase (rad), offset (Hz) for pulse channel q at step k. */
static void add_rf(const orc_sys *sy, const double *rfk, int nrf, cplx *H) {
    const int S = sy->S, N = sy->N;
    for (int q = 0; q < nrf; q++) {
        int iso = (int)rfk[4 * q];
        double amp = rfk[4 * q + 1], ph = rfk[4 * q + 2], off = rfk[4 * q + 3];
        /* a channel with zero amplitude adds nothing, its offset included (pulsedata.cc:114) */
        if (amp == 0.0) continue;
        for (int i = 0; i < S; i++) {
            if (sy->iso[i] != iso) continue;
            int bit = 1 << (S - 1 - i);
            for (int s = 0; s < N; s++) {
                int t = s ^ bit;
                int down = (s & bit) != 0; /* column state has m = -1/2 */
                /* Ix(t,s) = 1/2; Iy(t,s) = +i/2 when s has m=+1/2 (t has m=-1/2), -i/2 else
                 * (sps_Iy, singlespinop.cc:113-132) */
                cplx ix = 0.5, iy = down ? CMPLX(0.0, -0.5) : CMPLX(0.0, 0.5);
                H[t + (size_t)s * N] += amp * (cos(ph) * ix + sin(ph) * iy);
            }
            for (int s = 0; s < N; s++) H[s + (size_t)s * N] += off * mz(S, s, i);
        }
    }
}

/* step_dt: length of each step, or NULL for nsteps equal steps of one rotor period; step_t1:
 * start time of each step, or NULL for consecutive steps from t0.  A pulse segment of length T
 * split into m substeps is m entries T/m (propogation.cc:264-300: tau /= tmpp; t1 += tau;
 * t2 += tau). */
typedef struct {
    const orc_sys *sy;
    double wr, rotb_rad, t0, tau;
    int nsteps, npts, nrf;
    const double *step_t1, *step_dt, *rf;
    const cplx *rho, *det;
    const cplx *rho_all; /* non-NULL: crystallite c starts from rho_all + c N^2 */
    const double *theta, *phi, *gamma, *wt;
    int k0, k1, status;
    cplx *acc;
} orc_pulse_job;

/* U = prod_k exp(+i 2 pi tau_k (H_k + H_rf,k)) of a pulse sub-sequence for one crystallite
 * (cartPowder cp): equal steps tau from t0 when step_dt is NULL, else steps of lengths step_dt
 * starting at step_t1[k] (or consecutive from t0 when step_t1 is NULL) -- Propogation::
 * propogator, propogation.cc:240-304, U *= X meaning U = X U.  H, X, T are N x N work arrays. */
static int pulse_propagator(const orc_sys *sy, const double cp[9], double wr, double rotb_rad,
                            double t0, double tau, int nsteps, const double *step_t1,
                            const double *step_dt, const double *rf, int nrf, cplx *U, cplx *H,
                            cplx *X, cplx *T) {
    const int N = sy->N;
    const size_t NN = (size_t)N * N;
    for (size_t q = 0; q < NN; q++) U[q] = 0;
    for (int i = 0; i < N; i++) U[i + (size_t)i * N] = 1.0;
    double t1 = (step_dt && step_t1) ? step_t1[0] : t0;
    double t2 = t1 + (step_dt ? step_dt[0] : tau);
    for (int k = 0; k < nsteps; k++) {
        double dt = step_dt ? step_dt[k] : tau;
        double tmid = (t2 - t1) / 2.0 + t1;
        build_hamiltonian(sy, cp, rotb_rad, tmid * wr * PI2, H);
        add_rf(sy, rf + (size_t)k * 4 * nrf, nrf, H);
        int st = orc_expm_herm(N, H, CMPLX(0.0, 1.0) * dt * PI2, X);
        if (st) return st;
        zgemm(N, 'N', 'N', X, U, T);
        memcpy(U, T, sizeof(cplx) * NN);
        if (step_dt) {
            if (k + 1 < nsteps) {
                t1 = step_t1 ? step_t1[k + 1] : t2;
                t2 = t1 + step_dt[k + 1];
            }
        } else {
            t1 += tau;
            t2 += tau;
        }
    }
    return ORC_OK;
}

/* crystallites k0..k1-1 of a pulse-program powder: acc[m] += w_c Tr(rho_m det) */
static void *pulse_job_run(void *arg) {
    orc_pulse_job *jb = (orc_pulse_job *)arg;
    const orc_sys *sy = jb->sy;
    const int N = sy->N;
    const size_t NN = (size_t)N * N;
    const cplx *D = jb->det;
    cplx *H = malloc(sizeof(cplx) * NN), *X = malloc(sizeof(cplx) * NN), *U = malloc(sizeof(cplx) * NN);
    cplx *T = malloc(sizeof(cplx) * NN), *r = malloc(sizeof(cplx) * NN);
    for (int m = 0; m < jb->npts; m++) jb->acc[m] = 0;
    if (!H || !X || !U || !T || !r) {
        jb->status = ORC_ENOMEM;
        goto done;
    }
    for (int c = jb->k0; c < jb->k1; c++) {
        double cp[9];
        rot3(jb->phi[c], jb->theta[c], jb->gamma[c], cp);
        int st = pulse_propagator(sy, cp, jb->wr, jb->rotb_rad, jb->t0, jb->tau, jb->nsteps,
                                  jb->step_t1, jb->step_dt, jb->rf, jb->nrf, U, H, X, T);
        if (st) {
            jb->status = st;
            goto done;
        }
        memcpy(r, jb->rho_all ? jb->rho_all + (size_t)c * NN : jb->rho, sizeof(cplx) * NN);
        for (int m = 0; m < jb->npts; m++) {
            cplx tr = 0;
            for (int i = 0; i < N; i++)
                for (int j = 0; j < N; j++) tr += EL(r, i, j, N) * EL(D, j, i, N);
            jb->acc[m] += jb->wt[c] * tr;
            if (m + 1 < jb->npts) {
                zgemm(N, 'N', 'C', r, U, T);   /* r U^H */
                zgemm(N, 'N', 'N', U, T, r);   /* U r U^H */
            }
        }
    }
done:
    free(H); free(X); free(U); free(T); free(r);
    return NULL;
}

static int pulse_fid_impl(const double *rho_all, int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                  const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                  int nsteps, const double *step_t1, const double *step_dt, int npts,
                  double t0, int nrf,
                  const double *rf, int K, const double *theta, const double *phi,
                  const double *gamma, const double *w, const double *rho, const double *det,
                  int normalize, int nthreads, double *fid_out) {
    if (S < 1 || S > 12 || nsteps < 1 || npts < 0 || K < 0) return ORC_EINVAL;
    if (!step_dt && wr <= 0.0) return ORC_EINVAL;
    orc_sys sy = {S, 1 << S, isotopes, ncsa, csa, ndip, dip, nj, jc};
    const int N = sy.N;
    const cplx *D = (const cplx *)det;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > K && K > 0) nthreads = K;
    if (K == 0) nthreads = 1;
    const int P1 = npts > 0 ? npts : 1;
    orc_pulse_job *jobs = calloc(nthreads, sizeof(orc_pulse_job));
    pthread_t *th = calloc(nthreads, sizeof(pthread_t));
    cplx *accs = calloc((size_t)nthreads * P1, sizeof(cplx));
    if (!jobs || !th || !accs) { free(jobs); free(th); free(accs); return ORC_ENOMEM; }
    for (int t = 0; t < nthreads; t++) {
        orc_pulse_job *jb = &jobs[t];
        jb->sy = &sy; jb->wr = wr; jb->rotb_rad = rotor_deg * DEG2RAD; jb->t0 = t0;
        jb->tau = step_dt ? 0.0 : (1.0 / wr) / nsteps;
        jb->nsteps = nsteps; jb->npts = npts; jb->nrf = nrf;
        jb->step_t1 = step_t1; jb->step_dt = step_dt; jb->rf = rf;
        jb->rho = (const cplx *)rho; jb->det = D; jb->rho_all = (const cplx *)rho_all;
        jb->theta = theta; jb->phi = phi; jb->gamma = gamma; jb->wt = w;
        jb->k0 = (int)((long)K * t / nthreads); jb->k1 = (int)((long)K * (t + 1) / nthreads);
        jb->acc = accs + (size_t)t * P1;
    }
    for (int t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, pulse_job_run, &jobs[t]);
    pulse_job_run(&jobs[0]);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
    int st = ORC_OK;
    for (int t = 0; t < nthreads; t++) if (jobs[t].status) st = jobs[t].status;
    cplx trdd = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) trdd += EL(D, i, j, N) * conj(EL(D, i, j, N));
    for (int m = 0; m < npts; m++) {
        cplx v = 0;
        for (int t = 0; t < nthreads; t++) v += jobs[t].acc[m];
        if (normalize) v /= trdd;
        fid_out[2 * m] = creal(v);
        fid_out[2 * m + 1] = cimag(v);
    }
    free(jobs); free(th); free(accs);
    return st;
}

int orc_pulse_fid(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                  const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                  int nsteps, const double *step_t1, const double *step_dt, int npts,
                  double t0, int nrf,
                  const double *rf, int K, const double *theta, const double *phi,
                  const double *gamma, const double *w, const double *rho, const double *det,
                  int normalize, int nthreads, double *fid_out) {
    return pulse_fid_impl(NULL, S, isotopes, ncsa, csa, ndip, dip, nj, jc, wr, rotor_deg, nsteps,
                          step_t1, step_dt, npts, t0, nrf, rf, K, theta, phi, gamma, w, rho, det,
                          normalize, nthreads, fid_out);
}

/* ------------------------------------------------------------------------------------ */
/* per-crystallite rho carried through a pulse program (solid-2.0's ro per powder point) */
/* ------------------------------------------------------------------------------------ */
/* The FID of a pulse sub-sequence (as orc_pulse_fid) from crystallite c's own rho,
 * rho_all + c N^2 (column-major, interleaved). */
int orc_state_pulse_fid(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                        const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                        int nsteps, const double *step_t1, const double *step_dt, int npts,
                        double t0, int nrf, const double *rf, int K, const double *theta,
                        const double *phi, const double *gamma, const double *w,
                        const double *rho_all, const double *det, int normalize, int nthreads,
                        double *fid_out) {
    if (!rho_all) return ORC_EINVAL;
    return pulse_fid_impl(rho_all, S, isotopes, ncsa, csa, ndip, dip, nj, jc, wr, rotor_deg,
                          nsteps, step_t1, step_dt, npts, t0, nrf, rf, K, theta, phi, gamma, w,
                          rho_all, det, normalize, nthreads, fid_out);
}

/* The gamma-COMPUTE (spinning) or static FID (as orc_powder_fid) from crystallite c's own
 * rho (sequencerun_funcs.cc:699-735: hFID.FID(ro, detect, ...) with the propagated ro). */
int orc_state_fid(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                  const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                  double sw, int npts, double t0, int K, const double *theta, const double *phi,
                  const double *gamma, const double *w, const double *rho_all, const double *det,
                  int normalize, int nthreads, double *fid_out, double *sw_out) {
    if (!rho_all) return ORC_EINVAL;
    return powder_fid_impl(NULL, rho_all, S, isotopes, ncsa, csa, ndip, dip, nj, jc, wr,
                           rotor_deg, sw, npts, t0, K, theta, phi, gamma, w, rho_all, det,
                           normalize, nthreads, fid_out, sw_out);
}

/* rho_c <- U_c^repeat rho_c U_c^-repeat for every crystallite, U_c the propagator of the
 * steps at crystallite c's orientation (Propogation::propogator, propogation.cc:240-304; a
 * sub-sequence used `repeat` times applies its U that many times, sequencerun.cc:183-195). */
typedef struct {
    const orc_sys *sy;
    double wr, rotb_rad;
    int nsteps, nrf, repeat;
    const double *step_t1, *step_dt, *rf;
    const double *theta, *phi, *gamma;
    cplx *rho_all;
    int k0, k1, status;
} orc_apply_job;

static void *apply_job_run(void *arg) {
    orc_apply_job *jb = (orc_apply_job *)arg;
    const int N = jb->sy->N;
    const size_t NN = (size_t)N * N;
    cplx *H = malloc(sizeof(cplx) * NN), *X = malloc(sizeof(cplx) * NN);
    cplx *U = malloc(sizeof(cplx) * NN), *T = malloc(sizeof(cplx) * NN);
    if (!H || !X || !U || !T) {
        jb->status = ORC_ENOMEM;
        goto done;
    }
    for (int c = jb->k0; c < jb->k1; c++) {
        double cp[9];
        rot3(jb->phi[c], jb->theta[c], jb->gamma[c], cp);
        int st = pulse_propagator(jb->sy, cp, jb->wr, jb->rotb_rad, 0.0, 0.0, jb->nsteps,
                                  jb->step_t1, jb->step_dt, jb->rf, jb->nrf, U, H, X, T);
        if (st) {
            jb->status = st;
            goto done;
        }
        cplx *r = jb->rho_all + (size_t)c * NN;
        for (int m = 0; m < jb->repeat; m++) {
            zgemm(N, 'N', 'C', r, U, T); /* r U^H */
            zgemm(N, 'N', 'N', U, T, r); /* U r U^H */
        }
    }
done:
    free(H); free(X); free(U); free(T);
    return NULL;
}

int orc_state_apply(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                    const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                    int nsteps, const double *step_t1, const double *step_dt, int nrf,
                    const double *rf, int repeat, int K, const double *theta, const double *phi,
                    const double *gamma, double *rho_all, int nthreads) {
    if (S < 1 || S > 12 || nsteps < 1 || !step_dt || repeat < 0 || K < 0) return ORC_EINVAL;
    orc_sys sy = {S, 1 << S, isotopes, ncsa, csa, ndip, dip, nj, jc};
    if (nthreads < 1) nthreads = 1;
    if (nthreads > K && K > 0) nthreads = K;
    if (K == 0) return ORC_OK;
    orc_apply_job *jobs = calloc(nthreads, sizeof(orc_apply_job));
    pthread_t *th = calloc(nthreads, sizeof(pthread_t));
    if (!jobs || !th) { free(jobs); free(th); return ORC_ENOMEM; }
    for (int t = 0; t < nthreads; t++) {
        orc_apply_job *jb = &jobs[t];
        jb->sy = &sy; jb->wr = wr; jb->rotb_rad = rotor_deg * DEG2RAD;
        jb->nsteps = nsteps; jb->nrf = nrf; jb->repeat = repeat;
        jb->step_t1 = step_t1; jb->step_dt = step_dt; jb->rf = rf;
        jb->theta = theta; jb->phi = phi; jb->gamma = gamma; jb->rho_all = (cplx *)rho_all;
        jb->k0 = (int)((long)K * t / nthreads); jb->k1 = (int)((long)K * (t + 1) / nthreads);
    }
    for (int t = 1; t < nthreads; t++) pthread_create(&th[t], NULL, apply_job_run, &jobs[t]);
    apply_job_run(&jobs[0]);
    for (int t = 1; t < nthreads; t++) pthread_join(th[t], NULL);
    int st = ORC_OK;
    for (int t = 0; t < nthreads; t++) if (jobs[t].status) st = jobs[t].status;
    free(jobs); free(th);
    return st;
}

int orc_ptop_fid(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                 const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                 int nsteps, int npts, double t0, int nrf, const double *rf, int K,
                 const double *theta, const double *phi, const double *gamma, const double *w,
                 const double *rho, const double *det, int normalize, int nthreads,
                 double *fid_out) {
    return orc_pulse_fid(S, isotopes, ncsa, csa, ndip, dip, nj, jc, wr, rotor_deg, nsteps, NULL,
                         NULL, npts, t0, nrf, rf, K, theta, phi, gamma, w, rho, det, normalize,
                         nthreads,
                         fid_out);
}

/* Test hooks (exported): the pieces above on their own. */
int orc_hamiltonian(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                    const double *dip, int nj, const double *jc, double theta, double phi,
                    double gamma, double rotor_deg, double rotor_phase, double *H_out) {
    orc_sys sy = {S, 1 << S, isotopes, ncsa, csa, ndip, dip, nj, jc};
    double cp[9];
    rot3(phi, theta, gamma, cp);
    build_hamiltonian(&sy, cp, rotor_deg * DEG2RAD, rotor_phase, (cplx *)H_out);
    return ORC_OK;
}
int orc_eig_herm_x(int N, const double *A, double *w, double *V) {
    return orc_eig_herm(N, (const cplx *)A, w, (cplx *)V);
}
int orc_eig_normal_x(int N, const double *A, double *lam, double *Q) {
    return orc_eig_normal(N, (const cplx *)A, (cplx *)lam, (cplx *)Q);
}
int orc_expm_herm_x(int N, const double *H, double mul_re, double mul_im, double *U) {
    return orc_expm_herm(N, (const cplx *)H, CMPLX(mul_re, mul_im), (cplx *)U);
}
int orc_single_fid(int S, const int *isotopes, int ncsa, const double *csa, int ndip,
                   const double *dip, int nj, const double *jc, double wr, double rotor_deg,
                   double sw, int npts, double t0, double theta, double phi, double gamma,
                   const double *rho, const double *det, double *fid_out) {
    double one = 1.0, swo;
    return orc_powder_fid(S, isotopes, ncsa, csa, ndip, dip, nj, jc, wr, rotor_deg, sw, npts, t0,
                          1, &theta, &phi, &gamma, &one, rho, det, 0, 1, fid_out, &swo);
}
