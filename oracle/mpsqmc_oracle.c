/*
 * mpsqmc_oracle.c -- FP64 CPU restatement of the reference's projector-QMC
 * walker path (MPSQMC2 driver, Walker, MPSstate, MPStensor, TwoSiteObject,
 * TrotterHeisenberg, GridGenerator, HeisenbergMPO, Random).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, smoke() and the
 * bench's CPU-baseline leg may load it; the product library never does.
 *
 * It deliberately keeps the reference ALGORITHM (per coupling: 16 full-chain
 * transfer-matrix overlaps with precomputed Hermitian-conjugate term copies of
 * the trial, a full QR left-normalization after every operator application),
 * because that is what the GPU engine's results are checked against and what
 * the CPU baseline times.  Dense linear algebra (Householder QR, Jacobi
 * eigen/SVD, inverse error function, Mersenne twister) is written here from
 * the published algorithms; no LAPACK/BLAS/Boost is used.
 *
 * Storage conventions follow the reference exactly:
 *   site tensor  A[row + dimL*(col + dimR*s)]        (MPStensor.cpp:91-95 / 168-173)
 *   bond rule    VirtualD as in MPSstate.cpp:20-28
 *   coupling     J[(j*(j-1))/2 + i], i<j             (HeisenbergMPO.cpp:266-272)
 *   tables       U[i + d*(j + d*k)], VT[k + d*d*(i + d*j)]  (TrotterHeisenberg.cpp:391,425)
 *   grid         grid[i + dim*k]                     (GridGenerator.h:48-49)
 *
 * Pinning status of each function is recorded in the tests (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Mersenne twister MT19937 (Matsumoto & Nishimura 1998), u = x * 2^-32.     */
/* The reference draws from boost::random::mt19937 + uniform_real_distribution */
/* (Random.cpp:28-36,57); the build-defined contract is one stream, one 32-bit */
/* draw per double, u = x / 2^32.                                             */
/* ------------------------------------------------------------------------- */
typedef struct { uint32_t mt[624]; int idx; } orc_mt;

static void mt_seed(orc_mt *g, uint32_t s) {
    g->mt[0] = s;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static uint32_t mt_next(orc_mt *g) {
    if (g->idx >= 624) {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static double mt_uniform(orc_mt *g) { return (double)mt_next(g) * (1.0 / 4294967296.0); }

ORC_API void orc_mt_raw(uint32_t seed, int n, uint32_t *out) {
    orc_mt g; mt_seed(&g, seed);
    for (int i = 0; i < n; i++) out[i] = mt_next(&g);
}

ORC_API void orc_mt_uniforms(uint32_t seed, long skip, int n, double *out) {
    orc_mt g; mt_seed(&g, seed);
    for (long i = 0; i < skip; i++) (void)mt_next(&g);
    for (int i = 0; i < n; i++) out[i] = mt_uniform(&g);
}

/* ------------------------------------------------------------------------- */
/* Local spin operators, basis ordered smallest m first.                     */
/* OpSz.cpp:9, OpSx.cpp:11-21, OpISy.cpp:11-21, OpI.cpp:13-18, Op0.cpp:13-16  */
/* ------------------------------------------------------------------------- */
enum { OP_ZERO = 0, OP_I = 1, OP_SZ = 2, OP_SX = 3, OP_ISY = 4 };

static double ladder_coef(int d, int c) {
    return 0.25 * sqrt(d * d - 1.0 - (1.0 - d + 2 * c) * (3.0 - d + 2 * c));
}

static double op_elem(int op, int d, int i, int j) {
    switch (op) {
    case OP_I: return (i == j) ? 1.0 : 0.0;
    case OP_SZ: return (i == j) ? 0.5 * (1 - d) + i : 0.0;
    case OP_SX:
        if (i == j + 1) return ladder_coef(d, j);
        if (i + 1 == j) return ladder_coef(d, i);
        return 0.0;
    case OP_ISY:
        if (i == j + 1) return ladder_coef(d, j);
        if (i + 1 == j) return -ladder_coef(d, i);
        return 0.0;
    default: return 0.0;
    }
}

ORC_API void orc_operator(int op, int d, double *out) {
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) out[i + d * j] = op_elem(op, d, i, j);
}

/* ------------------------------------------------------------------------- */
/* MPS container.                                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    int L, d, D;       /* D = Dtrunc (max bond) */
    int *vd;           /* L+1 */
    double **t;        /* L site tensors */
} omps;

static void bond_rule(int L, int D, int d, int *vd) {  /* MPSstate.cpp:20-28 */
    vd[0] = 1;
    for (int k = 1; k < L; k++) vd[k] = (vd[k - 1] * d < D) ? vd[k - 1] * d : D;
    vd[L] = 1;
    for (int k = L - 1; k > 0; k--) {
        int a = vd[k + 1] * d; if (a > D) a = D;
        if (a < vd[k]) vd[k] = a;
    }
}

ORC_API void orc_bond_dims(int L, int D, int d, int *vd) { bond_rule(L, D, d, vd); }

static long site_size(const omps *m, int k) { return (long)m->vd[k] * m->vd[k + 1] * m->d; }

static void mps_alloc(omps *m, int L, int d, const int *vd) {
    m->L = L; m->d = d;
    m->vd = (int *)malloc(sizeof(int) * (L + 1));
    memcpy(m->vd, vd, sizeof(int) * (L + 1));
    m->D = 1;
    for (int k = 0; k <= L; k++) if (vd[k] > m->D) m->D = vd[k];
    m->t = (double **)malloc(sizeof(double *) * L);
    for (int k = 0; k < L; k++) m->t[k] = (double *)calloc((size_t)site_size(m, k), sizeof(double));
}

static void mps_free(omps *m) {
    if (!m->t) return;
    for (int k = 0; k < m->L; k++) free(m->t[k]);
    free(m->t); free(m->vd);
    m->t = NULL; m->vd = NULL;
}

static void mps_copy(omps *dst, const omps *src) {
    mps_alloc(dst, src->L, src->d, src->vd);
    dst->D = src->D;
    for (int k = 0; k < src->L; k++) memcpy(dst->t[k], src->t[k], sizeof(double) * site_size(src, k));
}

static long flat_size(int L, int d, const int *vd) {
    long n = 0;
    for (int k = 0; k < L; k++) n += (long)vd[k] * vd[k + 1] * d;
    return n;
}

static void mps_from_flat(omps *m, int L, int d, const int *vd, const double *data) {
    mps_alloc(m, L, d, vd);
    long off = 0;
    for (int k = 0; k < L; k++) {
        long s = site_size(m, k);
        memcpy(m->t[k], data + off, sizeof(double) * s);
        off += s;
    }
}

static void mps_to_flat(const omps *m, double *data) {
    long off = 0;
    for (int k = 0; k < m->L; k++) {
        long s = site_size(m, k);
        memcpy(data + off, m->t[k], sizeof(double) * s);
        off += s;
    }
}

/* ------------------------------------------------------------------------- */
/* Transfer-matrix overlap <this|other>, MPSstate.cpp:248-284.               */
/* E is dimL(other) x dimL(this), column-major; per site                     */
/*   E' = sum_s B[s]^T (E A[s]),  A = this, B = other.                       */
/* ------------------------------------------------------------------------- */
static double inner(const omps *A, const omps *B) {
    if (A->d != B->d || A->L != B->L) return 0.0;  /* MPSstate.cpp:250 */
    int L = A->L, d = A->d;
    int maxa = A->D, maxb = B->D;
    double *E = (double *)calloc((size_t)maxa * maxb, sizeof(double));
    double *E2 = (double *)calloc((size_t)maxa * maxb, sizeof(double));
    double *W = (double *)calloc((size_t)maxa * maxb, sizeof(double));
    E[0] = 1.0;
    for (int k = 0; k < L; k++) {
        int la = A->vd[k], ra = A->vd[k + 1], lb = B->vd[k], rb = B->vd[k + 1];
        memset(E2, 0, sizeof(double) * (size_t)ra * rb);
        for (int s = 0; s < d; s++) {
            const double *As = A->t[k] + (long)s * la * ra;
            const double *Bs = B->t[k] + (long)s * lb * rb;
            /* W (lb x ra) = E (lb x la) * As (la x ra), columns of E streamed */
            for (int c = 0; c < ra; c++) {
                double *w = W + (long)lb * c;
                for (int r = 0; r < lb; r++) w[r] = 0.0;
                for (int q = 0; q < la; q++) {
                    const double a = As[q + la * c];
                    const double *e = E + (long)lb * q;
                    for (int r = 0; r < lb; r++) w[r] += e[r] * a;
                }
            }
            /* E2 (rb x ra) += Bs^T (rb x lb) * W (lb x ra) */
            for (int c = 0; c < ra; c++)
                for (int r = 0; r < rb; r++) {
                    double acc = 0.0;
                    for (int q = 0; q < lb; q++) acc += Bs[q + lb * r] * W[q + lb * c];
                    E2[r + rb * c] += acc;
                }
        }
        double *sw = E; E = E2; E2 = sw;
    }
    double v = E[0];
    free(E); free(E2); free(W);
    return v;
}

/* ------------------------------------------------------------------------- */
/* Householder QR (the LAPACK dgeqrf / dorgqr convention: H_j = I - tau v v^T,*/
/* v_j = 1, beta = -sign(alpha) * ||x||).  A is m x n column-major, m >= n.   */
/* On exit A holds Q (m x n, orthonormal columns) and R (n x n upper).        */
/* ------------------------------------------------------------------------- */
static void householder_qr(int m, int n, double *A, double *R, double *tau) {
    for (int j = 0; j < n; j++) {
        double alpha = A[j + m * j];
        double xn = 0.0;
        for (int i = j + 1; i < m; i++) xn += A[i + m * j] * A[i + m * j];
        xn = sqrt(xn);
        if (xn == 0.0) { tau[j] = 0.0; continue; }
        double beta = -copysign(hypot(alpha, xn), alpha);
        tau[j] = (beta - alpha) / beta;
        double sc = 1.0 / (alpha - beta);
        for (int i = j + 1; i < m; i++) A[i + m * j] *= sc;
        A[j + m * j] = beta;
        for (int c = j + 1; c < n; c++) {
            double dot = A[j + m * c];
            for (int i = j + 1; i < m; i++) dot += A[i + m * j] * A[i + m * c];
            dot *= tau[j];
            A[j + m * c] -= dot;
            for (int i = j + 1; i < m; i++) A[i + m * c] -= dot * A[i + m * j];
        }
    }
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) R[r + n * c] = (c >= r) ? A[r + m * c] : 0.0;
    /* form Q = H_0 H_1 ... H_{n-1} [I_n; 0] backwards (dorg2r) */
    for (int j = n - 1; j >= 0; j--) {
        if (j < n - 1) {
            /* apply H_j to columns j+1..n-1 (rows j..m-1) */
            for (int c = j + 1; c < n; c++) {
                double dot = A[j + m * c];
                for (int i = j + 1; i < m; i++) dot += A[i + m * j] * A[i + m * c];
                dot *= tau[j];
                A[j + m * c] -= dot;
                for (int i = j + 1; i < m; i++) A[i + m * c] -= dot * A[i + m * j];
            }
        }
        for (int i = j + 1; i < m; i++) A[i + m * j] *= -tau[j];
        A[j + m * j] = 1.0 - tau[j];
        for (int i = 0; i < j; i++) A[i + m * j] = 0.0;
    }
}

ORC_API void orc_qr(int m, int n, double *A, double *R) {
    double *tau = (double *)malloc(sizeof(double) * n);
    householder_qr(m, n, A, R, tau);
    free(tau);
}

/* MPStensor::QR (MPStensor.cpp:100-135): rows ordered row + dimL*s. */
static void tensor_qr(int dl, int dr, int d, double *T, double *R, double *mem, double *tau) {
    int m = dl * d, n = dr;
    for (int s = 0; s < d; s++)
        for (int r = 0; r < dl; r++)
            for (int c = 0; c < dr; c++) mem[r + dl * (s + d * c)] = T[r + dl * (c + dr * s)];
    householder_qr(m, n, mem, R, tau);
    for (int s = 0; s < d; s++)
        for (int r = 0; r < dl; r++)
            for (int c = 0; c < dr; c++) T[r + dl * (c + dr * s)] = mem[r + dl * (s + d * c)];
}

/* MPStensor::LeftMultiply (MPStensor.cpp:159-175): A[s] <- R A[s]. */
static void tensor_left_multiply(int dl, int dr, int d, double *T, const double *R, double *work) {
    for (int s = 0; s < d; s++) {
        double *As = T + (long)s * dl * dr;
        for (int c = 0; c < dr; c++)
            for (int r = 0; r < dl; r++) {
                double acc = 0.0;
                for (int q = 0; q < dl; q++) acc += R[r + dl * q] * As[q + dl * c];
                work[r + dl * c] = acc;
            }
        memcpy(As, work, sizeof(double) * dl * dr);
    }
}

/* MPSstate::LeftNormalize (MPSstate.cpp:216-229): returns the final 1x1 R. */
static double left_normalize(omps *m) {
    int D = m->D, d = m->d;
    double *R = (double *)malloc(sizeof(double) * D * D);
    double *mem = (double *)malloc(sizeof(double) * D * D * d);
    double *tau = (double *)malloc(sizeof(double) * D);
    double *work = (double *)malloc(sizeof(double) * D * D);
    for (int k = 0; k < m->L; k++) {
        tensor_qr(m->vd[k], m->vd[k + 1], d, m->t[k], R, mem, tau);
        if (k != m->L - 1) tensor_left_multiply(m->vd[k + 1], m->vd[k + 2], d, m->t[k + 1], R, work);
    }
    double v = R[0];
    free(R); free(mem); free(tau); free(work);
    return v;
}

/* MPSstate::ChangePhase (MPSstate.cpp:307-315). */
static void change_phase(omps *m) {
    long n = site_size(m, 0);
    for (long i = 0; i < n; i++) m->t[0][i] = -m->t[0][i];
}

/* Walker::setOverlap (Walker.cpp:54-66). */
static double set_overlap(omps *w, const omps *trial) {
    (void)left_normalize(w);
    double v = inner(trial, w);
    if (v < 0.0) { change_phase(w); v = -v; }
    return v;
}

/* ------------------------------------------------------------------------- */
/* Symmetric Jacobi eigensolver (cyclic, Golub & Van Loan 8.5).              */
/* A n x n symmetric (destroyed); V eigenvectors as columns; w ascending.     */
/* ------------------------------------------------------------------------- */
static void jacobi_eig(int n, double *A, double *V, double *w) {
    for (int i = 0; i < n * n; i++) V[i] = 0.0;
    for (int i = 0; i < n; i++) V[i + n * i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0.0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) off += A[p + n * q] * A[p + n * q];
        if (off < 1e-300) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p + n * q];
                if (apq == 0.0) continue;
                double theta = (A[q + n * q] - A[p + n * p]) / (2.0 * apq);
                double t = copysign(1.0, theta) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k + n * p], akq = A[k + n * q];
                    A[k + n * p] = c * akp - s * akq;
                    A[k + n * q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p + n * k], aqk = A[q + n * k];
                    A[p + n * k] = c * apk - s * aqk;
                    A[q + n * k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k + n * p], vkq = V[k + n * q];
                    V[k + n * p] = c * vkp - s * vkq;
                    V[k + n * q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = A[i + n * i];
    /* sort ascending (dsyev convention) */
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++)
            if (w[j] < w[i]) {
                double tw = w[i]; w[i] = w[j]; w[j] = tw;
                for (int k = 0; k < n; k++) { double tv = V[k + n * i]; V[k + n * i] = V[k + n * j]; V[k + n * j] = tv; }
            }
}

/* ------------------------------------------------------------------------- */
/* One-sided Jacobi SVD (Hestenes), m x n column-major, any m, n.            */
/* Returns k = min(m,n) singular values descending, U (m x k), VT (k x n).    */
/* ------------------------------------------------------------------------- */
static void jacobi_svd(int m, int n, const double *Ain, double *S, double *U, double *VT) {
    int transposed = 0;
    int mm = m, nn = n;
    double *A;
    if (m < n) {  /* work on A^T so that rows >= cols */
        transposed = 1; mm = n; nn = m;
        A = (double *)malloc(sizeof(double) * (size_t)mm * nn);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < n; j++) A[j + mm * i] = Ain[i + m * j];
    } else {
        A = (double *)malloc(sizeof(double) * (size_t)mm * nn);
        memcpy(A, Ain, sizeof(double) * (size_t)mm * nn);
    }
    double *V = (double *)calloc((size_t)nn * nn, sizeof(double));
    for (int i = 0; i < nn; i++) V[i + nn * i] = 1.0;
    for (int sweep = 0; sweep < 80; sweep++) {
        int rotated = 0;
        for (int p = 0; p < nn; p++)
            for (int q = p + 1; q < nn; q++) {
                double a = 0, b = 0, g = 0;
                for (int i = 0; i < mm; i++) {
                    double x = A[i + (size_t)mm * p], y = A[i + (size_t)mm * q];
                    a += x * x; b += y * y; g += x * y;
                }
                if (fabs(g) <= 1e-15 * sqrt(a * b) || g == 0.0) continue;
                rotated = 1;
                double zeta = (b - a) / (2.0 * g);
                double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int i = 0; i < mm; i++) {
                    double x = A[i + (size_t)mm * p], y = A[i + (size_t)mm * q];
                    A[i + (size_t)mm * p] = c * x - s * y;
                    A[i + (size_t)mm * q] = s * x + c * y;
                }
                for (int i = 0; i < nn; i++) {
                    double x = V[i + (size_t)nn * p], y = V[i + (size_t)nn * q];
                    V[i + (size_t)nn * p] = c * x - s * y;
                    V[i + (size_t)nn * q] = s * x + c * y;
                }
            }
        if (!rotated) break;
    }
    double *sv = (double *)malloc(sizeof(double) * nn);
    int *ord = (int *)malloc(sizeof(int) * nn);
    for (int j = 0; j < nn; j++) {
        double s2 = 0;
        for (int i = 0; i < mm; i++) s2 += A[i + (size_t)mm * j] * A[i + (size_t)mm * j];
        sv[j] = sqrt(s2); ord[j] = j;
    }
    for (int i = 0; i < nn; i++)
        for (int j = i + 1; j < nn; j++)
            if (sv[ord[j]] > sv[ord[i]]) { int t = ord[i]; ord[i] = ord[j]; ord[j] = t; }
    /* left vectors of the worked matrix: A[:,j]/s_j; complete null directions
       by Gram-Schmidt against the already accepted ones. */
    double *Uw = (double *)calloc((size_t)mm * nn, sizeof(double));
    for (int jj = 0; jj < nn; jj++) {
        int j = ord[jj];
        double s = sv[j];
        S[jj] = s;
        double *u = Uw + (size_t)mm * jj;
        if (s > 1e-300) {
            for (int i = 0; i < mm; i++) u[i] = A[i + (size_t)mm * j] / s;
        } else {
            for (int e = 0; e < mm; e++) {  /* try unit vectors */
                for (int i = 0; i < mm; i++) u[i] = (i == e) ? 1.0 : 0.0;
                for (int pp = 0; pp < jj; pp++) {
                    double *up = Uw + (size_t)mm * pp, dt = 0;
                    for (int i = 0; i < mm; i++) dt += up[i] * u[i];
                    for (int i = 0; i < mm; i++) u[i] -= dt * up[i];
                }
                double nr = 0; for (int i = 0; i < mm; i++) nr += u[i] * u[i];
                nr = sqrt(nr);
                if (nr > 1e-8) { for (int i = 0; i < mm; i++) u[i] /= nr; break; }
            }
        }
    }
    int k = nn;  /* = min(m,n) */
    if (!transposed) {
        /* A = Uw S V^T : U = Uw (m x k), VT[r + k*c] = V[c, ord[r]] */
        for (int jj = 0; jj < k; jj++)
            for (int i = 0; i < m; i++) U[i + (size_t)m * jj] = Uw[i + (size_t)mm * jj];
        for (int r = 0; r < k; r++)
            for (int c = 0; c < n; c++) VT[r + (size_t)k * c] = V[c + (size_t)nn * ord[r]];
    } else {
        /* A^T = Uw S V^T  =>  A = V S Uw^T : U = V(:,ord) (m x k), VT = Uw^T (k x n) */
        for (int jj = 0; jj < k; jj++)
            for (int i = 0; i < m; i++) U[i + (size_t)m * jj] = V[i + (size_t)nn * ord[jj]];
        for (int r = 0; r < k; r++)
            for (int c = 0; c < n; c++) VT[r + (size_t)k * c] = Uw[c + (size_t)mm * r];
    }
    free(A); free(V); free(sv); free(ord); free(Uw);
}

ORC_API void orc_svd(int m, int n, const double *A, double *S, double *U, double *VT) {
    jacobi_svd(m, n, A, S, U, VT);
}

/* ------------------------------------------------------------------------- */
/* Trotter tables, TrotterHeisenberg.cpp:102-285 (dsyev -> P -> reshuffle ->  */
/* dgesdd).  Own Jacobi eigen/SVD: the basis inside a degenerate singular     */
/* subspace is implementation-defined (as it is for LAPACK).                  */
/* ------------------------------------------------------------------------- */
ORC_API void orc_trotter_tables(int d, double J, double dtau, double *S, double *U, double *VT, double *P_out) {
    int n = d * d;
    double *SS = (double *)malloc(sizeof(double) * n * n);
    double *A = (double *)malloc(sizeof(double) * n * n);
    double *V = (double *)malloc(sizeof(double) * n * n);
    double *lam = (double *)malloc(sizeof(double) * n);
    double *P = (double *)malloc(sizeof(double) * n * n);
    double *Q = (double *)malloc(sizeof(double) * n * n);
    for (int r1 = 0; r1 < d; r1++)
        for (int c1 = 0; c1 < d; c1++)
            for (int r2 = 0; r2 < d; r2++)
                for (int c2 = 0; c2 < d; c2++)
                    SS[r1 + d * (r2 + d * (c1 + d * c2))] =
                        op_elem(OP_SX, d, r1, c1) * op_elem(OP_SX, d, r2, c2) -
                        op_elem(OP_ISY, d, r1, c1) * op_elem(OP_ISY, d, r2, c2) +
                        op_elem(OP_SZ, d, r1, c1) * op_elem(OP_SZ, d, r2, c2);
    memcpy(A, SS, sizeof(double) * n * n);
    jacobi_eig(n, A, V, lam);
    /* P = V exp(-dtau J lam) V^T (TrotterHeisenberg.cpp:185-200) */
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) {
            double acc = 0.0;
            for (int k = 0; k < n; k++) acc += V[r + n * k] * exp(-dtau * J * lam[k]) * V[c + n * k];
            P[r + n * c] = acc;
        }
    /* reshuffle P[r1 + d r2, c1 + d c2] -> Q[r1 + d c1, r2 + d c2] (:203-219) */
    for (int r1 = 0; r1 < d; r1++)
        for (int r2 = 0; r2 < d; r2++)
            for (int c1 = 0; c1 < d; c1++)
                for (int c2 = 0; c2 < d; c2++)
                    Q[(r1 + d * c1) + n * (r2 + d * c2)] = P[(r1 + d * r2) + n * (c1 + d * c2)];
    if (P_out) memcpy(P_out, Q, sizeof(double) * n * n);
    jacobi_svd(n, n, Q, S, U, VT);
    free(SS); free(A); free(V); free(lam); free(P); free(Q);
}

/* ------------------------------------------------------------------------- */
/* Inverse error function: initial rational guess + Newton on libm erf.       */
/* (Boost.Math erf_inv is the reference's call, GridGenerator.cpp:178.)       */
/* ------------------------------------------------------------------------- */
static double erfinv_own(double y) {
    if (y <= -1.0) return -INFINITY;
    if (y >= 1.0) return INFINITY;
    if (y == 0.0) return 0.0;
    /* Giles (2010) single-precision-style initial approximation */
    double w = -log((1.0 - y) * (1.0 + y)), x;
    if (w < 5.0) {
        w -= 2.5;
        double p = 2.81022636e-08;
        p = 3.43273939e-07 + p * w; p = -3.5233877e-06 + p * w; p = -4.39150654e-06 + p * w;
        p = 0.00021858087 + p * w; p = -0.00125372503 + p * w; p = -0.00417768164 + p * w;
        p = 0.246640727 + p * w; p = 1.50140941 + p * w;
        x = p * y;
    } else {
        w = sqrt(w) - 3.0;
        double p = -0.000200214257;
        p = 0.000100950558 + p * w; p = 0.00134934322 + p * w; p = -0.00367342844 + p * w;
        p = 0.00573950773 + p * w; p = -0.0076224613 + p * w; p = 0.00943887047 + p * w;
        p = 1.00167406 + p * w; p = 2.83297682 + p * w;
        x = p * y;
    }
    for (int it = 0; it < 4; it++) {
        double err = erf(x) - y;
        x -= err / (1.1283791670955126 * exp(-x * x));
    }
    return x;
}

ORC_API double orc_erfinv(double y) { return erfinv_own(y); }

/* GridGenerator::FillMarsaglia, GridGenerator.cpp:167-225. Returns N. */
ORC_API int orc_marsaglia(int dim, int p, double *grid /* capacity dim*p^dim */) {
    double *one = (double *)malloc(sizeof(double) * p);
    for (int c = 0; c < p; c++) one[c] = erfinv_own(2 * (c + 0.5) / p - 1.0);
    long est = 1;
    for (int c = 0; c < dim; c++) est *= p;
    double *att = (double *)malloc(sizeof(double) * dim);
    int N = 0;
    for (long cnt = 0; cnt < est; cnt++) {
        long cc = cnt;
        double nrm = 0.0;
        for (int c2 = 0; c2 < dim; c2++) {
            int rem = (int)(cc % p);
            att[c2] = one[rem];
            nrm += att[c2] * att[c2];
            cc = (cc - rem) / p;
        }
        if (nrm > 1e-16) {
            nrm = 1.0 / sqrt(nrm);
            for (int c2 = 0; c2 < dim; c2++) att[c2] *= nrm;
            int keep = 1;
            for (int st = 0; st < N; st++) {
                double ds = 0.0;
                for (int co = 0; co < dim; co++) { double df = att[co] - grid[co + dim * st]; ds += df * df; }
                if (ds < 1e-16) { keep = 0; break; }
            }
            if (keep) { for (int co = 0; co < dim; co++) grid[co + dim * N] = att[co]; N++; }
        }
    }
    free(one); free(att);
    return N;
}

/* ------------------------------------------------------------------------- */
/* HeisenbergMPO W-tensor entries (HeisenbergMPO.cpp:99-253, ladder=false).   */
/* ------------------------------------------------------------------------- */
typedef struct { int L, d; const double *J; double h; int *dims; } ompo;

static double mpo_J(const ompo *M, int i, int j) {
    return (i < j) ? M->J[(j * (j - 1)) / 2 + i] : M->J[(i * (i - 1)) / 2 + j];
}

static void mpo_dims(int L, int *dims) {
    int sw = L / 2;
    dims[0] = dims[L] = 1;
    for (int s = 0; s < sw; s++) dims[s + 1] = 2 + 3 * (s + 1);
    for (int s = L - 1; s > sw; s--) dims[s] = 2 + 3 * (L - s);
}

/* returns prefactor; *op receives the operator code */
static double mpo_entry(const ompo *M, int site, int row, int col, int *op) {
    int L = M->L, sw = L / 2;
    int dl = M->dims[site], dr = M->dims[site + 1];
    double mh = -M->h;
    *op = OP_ZERO;
    if (row < 0 || row >= dl || col < 0 || col >= dr) return 0.0;
    double pf = 0.0; int o = OP_ZERO;
    const int alpha_op[3] = {OP_SX, OP_ISY, OP_SZ};
    if (site == 0) {
        if (col == 0) { pf = mh; o = OP_SZ; }
        else if (col >= 1 && col <= 3) { pf = 1.0; o = alpha_op[col - 1]; }
        else if (col == 4) { pf = 1.0; o = OP_I; }
    } else if (site == L - 1) {
        if (row == 0) { pf = 1.0; o = OP_I; }
        else if (row >= 1 && row <= 3) { pf = 1.0; o = alpha_op[row - 1]; }
        else if (row == 4) { pf = mh; o = OP_SZ; }
    } else if (site < sw) {
        if (row < dl - 1 && col == row) { pf = 1.0; o = OP_I; }
        if (row == dl - 1) {
            if (col == 0) { pf = mh; o = OP_SZ; }
            else if (col == dl - 1) { pf = 1.0; o = OP_SX; }
            else if (col == dl) { pf = 1.0; o = OP_ISY; }
            else if (col == dl + 1) { pf = 1.0; o = OP_SZ; }
            else if (col == dl + 2) { pf = 1.0; o = OP_I; }
        }
        if (col == 0 && row >= 1 && row < 1 + 3 * site) {
            int r = (row - 1) / 3, a = (row - 1) % 3;
            double v = mpo_J(M, r, site);
            pf = (a == 1) ? -v : v; o = alpha_op[a];
        }
    } else if (site == sw) {
        if (row == 0 && col == 0) { pf = 1.0; o = OP_I; }
        if (row == dl - 1 && col == 0) { pf = mh; o = OP_SZ; }
        if (row >= 1 && row < 1 + 3 * sw) {
            int r = (row - 1) / 3, a = (row - 1) % 3;
            if (col == 0) { double v = mpo_J(M, r, sw); pf = (a == 1) ? -v : v; o = alpha_op[a]; }
            else if (col >= 1 && col < 1 + 3 * (L - 1 - sw)) {
                int cc = (col - 1) / 3, b = (col - 1) % 3;
                if (b == a) { double v = mpo_J(M, r, sw + 1 + cc); pf = (a == 1) ? -v : v; o = OP_I; }
            }
        }
        if (row == dl - 1 && col >= 1 && col < 1 + 3 * (L - 1 - sw)) {
            int cc = (col - 1) / 3, b = (col - 1) % 3;
            double v = mpo_J(M, sw, sw + 1 + cc);
            pf = (b == 1) ? -v : v; o = alpha_op[b];
        }
        if (row == dl - 1 && col == 1 + 3 * (L - 1 - sw)) { pf = 1.0; o = OP_I; }
    } else { /* sw < site < L-1 */
        if (row == 0 && col == 0) { pf = 1.0; o = OP_I; }
        if (row == dl - 1 && col == 0) { pf = mh; o = OP_SZ; }
        if (col == 0 && row >= 1 && row <= 3) { pf = 1.0; o = alpha_op[row - 1]; }
        if (col >= 1 && row == 3 + col) { pf = 1.0; o = OP_I; }
        if (row == dl - 1 && col >= 1 && col < 1 + 3 * (L - 1 - site)) {
            int cc = (col - 1) / 3, b = (col - 1) % 3;
            double v = mpo_J(M, site, site + 1 + cc);
            pf = (b == 1) ? -v : v; o = alpha_op[b];
        }
    }
    if (pf == 0.0) { *op = OP_ZERO; return 0.0; }  /* HeisenbergMPO.cpp:322 */
    *op = o;
    return pf;
}

/* MPSstate::ApplyMPO (MPSstate.cpp:413-471): out = W * psi (uncompressed). */
static void apply_mpo(const ompo *M, const omps *psi, omps *out) {
    int L = psi->L, d = psi->d;
    int *vd = (int *)malloc(sizeof(int) * (L + 1));
    vd[0] = vd[L] = 1;
    for (int k = 1; k < L; k++) vd[k] = M->dims[k] * psi->vd[k];
    mps_alloc(out, L, d, vd);
    free(vd);
    for (int site = 0; site < L; site++) {
        int dls = psi->vd[site], drs = psi->vd[site + 1];
        int dlm = M->dims[site], drm = M->dims[site + 1];
        for (int ml = 0; ml < dlm; ml++)
            for (int mr = 0; mr < drm; mr++) {
                int op; double f = mpo_entry(M, site, ml, mr, &op);
                if (op == OP_ZERO) continue;
                for (int up = 0; up < d; up++) {
                    double *tg = out->t[site] + (long)up * out->vd[site] * out->vd[site + 1];
                    for (int dn = 0; dn < d; dn++) {
                        double fb = f * op_elem(op, d, up, dn);
                        if (fb == 0.0) continue;
                        const double *src = psi->t[site] + (long)dn * dls * drs;
                        for (int r = 0; r < dls; r++)
                            for (int c = 0; c < drs; c++)
                                tg[(ml * dls + r) + (dls * dlm) * (mr * drs + c)] += fb * src[r + dls * c];
                    }
                }
            }
    }
}


/* ------------------------------------------------------------------------- */
/* H|trial> through a minimal-channel MPO: channels at cut k are "nothing yet",
 * "complete", and S^a placed on a site i < k that still has a partner j >= k.
 * The resulting MPS equals the reference's ApplyMPO + lossless CompressState
 * (MPSQMC2.cpp:255-257) as a vector, with bond (2 + 3 n_open(k)) * D_T, the
 * rank of H|trial> across the cut; it gives the reference's energy cost.     */
/* ------------------------------------------------------------------------- */
static int open_sites(const double *J, int L, int k, int *list) {
    /* sites i < k with a nonzero coupling to some j >= k, in increasing order */
    int n = 0;
    for (int i = 0; i < k; i++) {
        int has = 0;
        for (int j = k; j < L && !has; j++) if (J[(j * (j - 1)) / 2 + i] != 0.0) has = 1;
        if (has) list[n++] = i;
    }
    return n;
}

static void apply_mpo_min(int L, int d, const double *J, double h, const omps *psi, omps *out) {
    int *nch = (int *)malloc(sizeof(int) * (L + 1));
    int **lst = (int **)malloc(sizeof(int *) * (L + 1));
    for (int k = 0; k <= L; k++) {
        lst[k] = (int *)malloc(sizeof(int) * (L + 1));
        int no = open_sites(J, L, k, lst[k]);
        nch[k] = (k == 0 || k == L) ? 1 : 2 + 3 * no;
        if (k > 0 && k < L) { /* keep the count in lst[k][L] */ lst[k][L] = no; }
    }
    /* channel numbering at an inner cut: 0 = nothing yet, 1 = complete, 2 + 3q + a = (lst[q], a) */
    int *vd = (int *)malloc(sizeof(int) * (L + 1));
    vd[0] = vd[L] = 1;
    for (int k = 1; k < L; k++) vd[k] = nch[k] * psi->vd[k];
    mps_alloc(out, L, d, vd);
    free(vd);
    const int ops[3] = {OP_SX, OP_ISY, OP_SZ};
    const double sg[3] = {1.0, -1.0, 1.0};
    double *O = (double *)malloc(sizeof(double) * d * d);
    for (int k = 0; k < L; k++) {
        int cl = nch[k], cr = nch[k + 1];
        int dls = psi->vd[k], drs = psi->vd[k + 1];
        int noL = (k > 0) ? lst[k][L] : 0, noR = (k + 1 < L) ? lst[k + 1][L] : 0;
        /* map a left channel / right channel to (kind, site, a) */
        for (int a = 0; a < cl; a++)
            for (int b = 0; b < cr; b++) {
                /* decode left */
                int lk, li = -1, la = -1;   /* lk: 0 nothing, 1 complete, 2 open */
                if (k == 0) lk = 0;
                else if (a == 0) lk = 0;
                else if (a == 1) lk = 1;
                else { lk = 2; li = lst[k][(a - 2) / 3]; la = (a - 2) % 3; }
                int rk, ri = -1, ra = -1;
                if (k == L - 1) rk = 1;
                else if (b == 0) rk = 0;
                else if (b == 1) rk = 1;
                else { rk = 2; ri = lst[k + 1][(b - 2) / 3]; ra = (b - 2) % 3; }
                (void)noL; (void)noR;
                double f = 0.0; int op = OP_ZERO;
                if (lk == 0 && rk == 0) { f = 1.0; op = OP_I; }
                else if (lk == 1 && rk == 1) { f = 1.0; op = OP_I; }
                else if (lk == 0 && rk == 1) { if (h != 0.0) { f = -h; op = OP_SZ; } }
                else if (lk == 0 && rk == 2) { if (ri == k) { f = 1.0; op = ops[ra]; } }
                else if (lk == 2 && rk == 2) { if (li == ri && la == ra) { f = 1.0; op = OP_I; } }
                else if (lk == 2 && rk == 1) {
                    double v = J[(k * (k - 1)) / 2 + li];
                    if (v != 0.0) { f = sg[la] * v; op = ops[la]; }
                }
                if (op == OP_ZERO) continue;
                orc_operator(op, d, O);
                for (int up = 0; up < d; up++) {
                    double *tg = out->t[k] + (long)up * out->vd[k] * out->vd[k + 1];
                    for (int dn = 0; dn < d; dn++) {
                        double fb = f * O[up + d * dn];
                        if (fb == 0.0) continue;
                        const double *src = psi->t[k] + (long)dn * dls * drs;
                        for (int r = 0; r < dls; r++)
                            for (int c = 0; c < drs; c++)
                                tg[(a * dls + r) + (dls * cl) * (b * drs + c)] += fb * src[r + dls * c];
                    }
                }
            }
    }
    free(O);
    for (int k = 0; k <= L; k++) free(lst[k]);
    free(lst); free(nch);
}

ORC_API long orc_apply_mpo_min(int L, int d, const double *J, double h, const int *vd, const double *data,
                               int *vd_out, double *out) {
    omps P, H;
    mps_from_flat(&P, L, d, vd, data);
    apply_mpo_min(L, d, J, h, &P, &H);
    memcpy(vd_out, H.vd, sizeof(int) * (L + 1));
    long n = flat_size(L, d, H.vd);
    if (out) mps_to_flat(&H, out);
    mps_free(&P); mps_free(&H);
    return n;
}

/* ------------------------------------------------------------------------- */
/* CompressState (MPSstate.cpp:379-404) with TwoSiteObject Compose/Decompose  */
/* (TwoSiteObject.cpp:71-184), Schmidt cutoff 1e-13 absolute (:16).           */
/* ------------------------------------------------------------------------- */
static void two_site_step(omps *m, int k, int Dt, int movingright) {
    int d = m->d;
    int DL = m->vd[k], DM = m->vd[k + 1], DR = m->vd[k + 2];
    int rows = DL * d, cols = DR * d, kmin = rows < cols ? rows : cols;
    double *mem = (double *)calloc((size_t)rows * cols, sizeof(double));
    /* Compose: T[sl,sr] = A[sl] B[sr]; Decompose reorders to
       mem[alpha + DL*(il + d*(beta + DR*ir))] */
    for (int il = 0; il < d; il++)
        for (int ir = 0; ir < d; ir++) {
            const double *A = m->t[k] + (long)il * DL * DM;
            const double *B = m->t[k + 1] + (long)ir * DM * DR;
            for (int a = 0; a < DL; a++)
                for (int b = 0; b < DR; b++) {
                    double acc = 0.0;
                    for (int q = 0; q < DM; q++) acc += A[a + DL * q] * B[q + DM * b];
                    mem[a + DL * (il + d * (b + DR * ir))] = acc;
                }
        }
    double *S = (double *)malloc(sizeof(double) * kmin);
    double *U = (double *)malloc(sizeof(double) * (size_t)rows * kmin);
    double *VT = (double *)malloc(sizeof(double) * (size_t)kmin * cols);
    jacobi_svd(rows, cols, mem, S, U, VT);
    int dt = Dt;
    int itr = kmin;
    for (int c = 0; c < kmin; c++) if (S[c] < 1e-13) { itr = c; break; }
    if (itr < dt) dt = itr;
    if (dt > kmin) dt = kmin;
    free(m->t[k]); free(m->t[k + 1]);
    m->vd[k + 1] = dt;
    m->t[k] = (double *)calloc((size_t)DL * dt * d, sizeof(double));
    m->t[k + 1] = (double *)calloc((size_t)dt * DR * d, sizeof(double));
    for (int ip = 0; ip < d; ip++)
        for (int a = 0; a < DL; a++)
            for (int b = 0; b < dt; b++)
                m->t[k][a + DL * (b + dt * ip)] = U[(a + DL * ip) + (size_t)rows * b] * (movingright ? 1.0 : S[b]);
    for (int ip = 0; ip < d; ip++)
        for (int a = 0; a < dt; a++)
            for (int b = 0; b < DR; b++)
                m->t[k + 1][a + dt * (b + DR * ip)] = VT[a + (size_t)kmin * (b + DR * ip)] * (movingright ? S[a] : 1.0);
    free(mem); free(S); free(U); free(VT);
}

static void compress_state(omps *m, int Dt) {
    for (int k = 0; k < m->L - 1; k++) two_site_step(m, k, Dt, 1);
    for (int k = m->L - 2; k >= 0; k--) two_site_step(m, k, Dt, 0);
    int mx = m->vd[0];
    for (int k = 1; k < m->L; k++) if (m->vd[k] > mx) mx = m->vd[k];
    m->D = mx;
}

/* ------------------------------------------------------------------------- */
/* Local operator application.                                               */
/* ------------------------------------------------------------------------- */
/* A[up] <- sum_dn O(up,dn) A[dn]  (the daxpy loops of MPSstate.cpp:505-659) */
static void apply_local(omps *m, int site, const double *O /* d x d, O[up + d*dn] */) {
    int d = m->d;
    long blk = (long)m->vd[site] * m->vd[site + 1];
    double *w = (double *)calloc((size_t)blk * d, sizeof(double));
    for (int up = 0; up < d; up++)
        for (int dn = 0; dn < d; dn++) {
            double v = O[up + d * dn];
            if (v == 0.0) continue;
            const double *src = m->t[site] + dn * blk;
            double *dst = w + up * blk;
            for (long i = 0; i < blk; i++) dst[i] += v * src[i];
        }
    memcpy(m->t[site], w, sizeof(double) * blk * d);
    free(w);
}

/* ------------------------------------------------------------------------- */
/* Standalone entry points on flat MPS (tests).                              */
/* ------------------------------------------------------------------------- */
ORC_API double orc_inner(int L, int d, const int *vda, const double *a, const int *vdb, const double *b) {
    omps A, B;
    mps_from_flat(&A, L, d, vda, a);
    mps_from_flat(&B, L, d, vdb, b);
    double v = inner(&A, &B);
    mps_free(&A); mps_free(&B);
    return v;
}

ORC_API double orc_left_normalize(int L, int d, const int *vd, double *data) {
    omps A;
    mps_from_flat(&A, L, d, vd, data);
    double v = left_normalize(&A);
    mps_to_flat(&A, data);
    mps_free(&A);
    return v;
}

/* Compress (MPSstate::CompressState(Dt)); writes new bond dims into vd_out and
   data into out (capacity flat_size of the input).  Returns the new Dtrunc. */
ORC_API int orc_compress(int L, int d, const int *vd, const double *data, int Dt, int *vd_out, double *out) {
    omps A;
    mps_from_flat(&A, L, d, vd, data);
    compress_state(&A, Dt);
    memcpy(vd_out, A.vd, sizeof(int) * (L + 1));
    mps_to_flat(&A, out);
    int D = A.D;
    mps_free(&A);
    return D;
}

/* ApplyMPO of the Heisenberg MPO; vd_out gets the bond dims, out must hold
   flat_size(vd_out) doubles (call once with out == NULL to get the size). */
ORC_API long orc_apply_mpo(int L, int d, const double *J, double h, const int *vd, const double *data,
                           int *vd_out, double *out) {
    ompo M; M.L = L; M.d = d; M.J = J; M.h = h;
    M.dims = (int *)malloc(sizeof(int) * (L + 1));
    mpo_dims(L, M.dims);
    int *vo = (int *)malloc(sizeof(int) * (L + 1));
    vo[0] = vo[L] = 1;
    for (int k = 1; k < L; k++) vo[k] = M.dims[k] * vd[k];
    long n = flat_size(L, d, vo);
    memcpy(vd_out, vo, sizeof(int) * (L + 1));
    if (out) {
        omps P, H;
        mps_from_flat(&P, L, d, vd, data);
        apply_mpo(&M, &P, &H);
        mps_to_flat(&H, out);
        mps_free(&P); mps_free(&H);
    }
    free(vo); free(M.dims);
    return n;
}

ORC_API void orc_mpo_dims(int L, int *dims) { mpo_dims(L, dims); }

ORC_API double orc_mpo_entry(int L, int d, const double *J, double h, int site, int row, int col, int *op) {
    ompo M; M.L = L; M.d = d; M.J = J; M.h = h;
    M.dims = (int *)malloc(sizeof(int) * (L + 1));
    mpo_dims(L, M.dims);
    double v = mpo_entry(&M, site, row, col, op);
    free(M.dims);
    return v;
}

/* ------------------------------------------------------------------------- */
/* The walker population (MPSQMC2 restated).                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
    int L, d, nC, n2;       /* n2 = d*d = trotterSVDsize */
    int *first, *second;    /* couplings, lexicographic (MPSQMC2.cpp:286-299) */
    double *Jc;             /* coupling value per coupling */
    double *tS, *tU, *tVT;  /* tables per coupling (copied from the distinct-J set) */
    double h, dtau;
    double *ssp;            /* single-site propagator diag (h != 0) */
    double *grid; int Ng, gdim;
    omps trial, htrial;
    omps *terms;            /* nC * n2*n2 HC term copies */
    int nw, cap, DT, DW;
    omps *w; double *wt, *ov, *eh;
    int *idx;               /* sampled grid point per (walker, coupling), last propagate */
    double *EL;             /* last local energies */
    int nthreads;
    int inner_threads;      /* threads over the 16 combos of one walker (fewer walkers than threads) */
    int emode;              /* 0: <H trial|phi> with uncompressed H|trial>; 1: per-coupling operator insertion;
                               2: <H trial|phi> with H|trial> from the minimal-channel MPO */
    omps hmin;
    double *Jfull;
    /* learned proposal (the engine's extension; no counterpart in the reference) */
    const double *pred;     /* predicted labels per (walker, coupling), 16 each, or NULL */
    double mix;
    double *lab;            /* receives combos/O per (walker, coupling), or NULL */
} orc_pop;

/* Proposal from predicted labels (the engine's learned proposal, restated):
   q(x) = (1-mix) qhat(x)/sum qhat + mix/N with qhat(x) = max(0, sum p x x),
   uniform when every qhat is 0. */
ORC_API void orc_pop_set_proposal(orc_pop *p, const double *pred, double mix) { p->pred = pred; p->mix = mix; }
ORC_API void orc_pop_set_labels(orc_pop *p, double *lab) { p->lab = lab; }

ORC_API int orc_nthreads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Two-site term for coupling c with SVD indices a (left) and b (right),
   doHC: operator elements are read transposed (MPSstate.cpp:505-573). */
static void apply_svd_term(const orc_pop *p, omps *m, int c, int a, int b, int doHC) {
    int d = p->d, n2 = p->n2;
    double *O = (double *)malloc(sizeof(double) * d * d);
    const double *U = p->tU + (long)c * n2 * n2, *VT = p->tVT + (long)c * n2 * n2, *S = p->tS + (long)c * n2;
    double sa = sqrt(S[a]), sb = sqrt(S[b]);
    for (int up = 0; up < d; up++)
        for (int dn = 0; dn < d; dn++) {
            int i = doHC ? dn : up, j = doHC ? up : dn;
            O[up + d * dn] = sa * U[i + d * (j + d * a)];
        }
    apply_local(m, p->first[c], O);
    for (int up = 0; up < d; up++)
        for (int dn = 0; dn < d; dn++) {
            int i = doHC ? dn : up, j = doHC ? up : dn;
            O[up + d * dn] = sb * VT[b + n2 * (i + d * j)];
        }
    apply_local(m, p->second[c], O);
    free(O);
}

/* Grid-point operator (MPSstate.cpp:575-633). */
static void apply_grid_term(const orc_pop *p, omps *m, int c, int g) {
    int d = p->d, n2 = p->n2, dim = p->gdim;
    const double *U = p->tU + (long)c * n2 * n2, *VT = p->tVT + (long)c * n2 * n2, *S = p->tS + (long)c * n2;
    double *OL = (double *)calloc(d * d, sizeof(double));
    double *OR = (double *)calloc(d * d, sizeof(double));
    for (int k = 0; k < dim; k++) {
        double pf = sqrt(S[k] * dim) * p->grid[k + dim * g];
        if (pf == 0.0) continue;
        for (int up = 0; up < d; up++)
            for (int dn = 0; dn < d; dn++) {
                OL[up + d * dn] += pf * U[up + d * (dn + d * k)];
                OR[up + d * dn] += pf * VT[k + n2 * (up + d * dn)];
            }
    }
    /* The reference accumulates per-k daxpys into the result; summing the
       operator first is the same linear map (differences at rounding level). */
    apply_local(m, p->first[c], OL);
    apply_local(m, p->second[c], OR);
    free(OL); free(OR);
}

static void apply_one_site_everywhere(const orc_pop *p, omps *m) {  /* MPSstate.cpp:635-659 */
    if (p->h == 0.0) return;
    int d = p->d;
    double *O = (double *)calloc(d * d, sizeof(double));
    for (int i = 0; i < d; i++) O[i + d * i] = p->ssp[i];
    for (int k = 0; k < m->L; k++) apply_local(m, k, O);
    free(O);
}

/*
 * Create the population.  J: L(L-1)/2 couplings; tables are given per
 * distinct coupling value: ndiff values dJ[], S[ndiff][n2], U[ndiff][n2*n2],
 * VT[ndiff][n2*n2] (lookup by exact equality, TrotterHeisenberg.cpp:355-357).
 * The trial is left-normalized here (MPSQMC2.cpp:239).
 */
ORC_API orc_pop *orc_pop_create(int L, int d, const int *tvd, const double *tdata, const double *J, double h,
                                double dtau, int ndiff, const double *dJ, const double *S, const double *U,
                                const double *VT, const double *grid, int Ng, int gdim, int build_terms) {
    orc_pop *p = (orc_pop *)calloc(1, sizeof(orc_pop));
    p->L = L; p->d = d; p->n2 = d * d; p->h = h; p->dtau = dtau;
    p->nthreads = 1;
    p->inner_threads = 1;
    mps_from_flat(&p->trial, L, d, tvd, tdata);
    (void)left_normalize(&p->trial);
    p->DT = p->trial.D;
    /* H|trial> (MPSQMC2.cpp:255-257) is built on first use (hamiltonian_of):
       at L = 36, D = 128 the uncompressed state holds tens of GB, and the
       operator-insertion energy (mode 1) does not need it. */
    p->Jfull = (double *)malloc(sizeof(double) * (L * (L - 1) / 2 > 0 ? L * (L - 1) / 2 : 1));
    memcpy(p->Jfull, J, sizeof(double) * (L * (L - 1) / 2));
    int nC = 0;
    for (int i = 0; i < L - 1; i++)
        for (int j = i + 1; j < L; j++) if (J[(j * (j - 1)) / 2 + i] != 0.0) nC++;
    p->nC = nC;
    p->first = (int *)malloc(sizeof(int) * (nC > 0 ? nC : 1));
    p->second = (int *)malloc(sizeof(int) * (nC > 0 ? nC : 1));
    p->Jc = (double *)malloc(sizeof(double) * (nC > 0 ? nC : 1));
    int n2 = p->n2;
    p->tS = (double *)malloc(sizeof(double) * (nC > 0 ? nC : 1) * n2);
    p->tU = (double *)malloc(sizeof(double) * (nC > 0 ? nC : 1) * n2 * n2);
    p->tVT = (double *)malloc(sizeof(double) * (nC > 0 ? nC : 1) * n2 * n2);
    nC = 0;
    for (int i = 0; i < L - 1; i++)
        for (int j = i + 1; j < L; j++) {
            double v = J[(j * (j - 1)) / 2 + i];
            if (v == 0.0) continue;
            p->first[nC] = i; p->second[nC] = j; p->Jc[nC] = v;
            int t = -1;
            for (int q = 0; q < ndiff; q++) if (dJ[q] == v) { t = q; break; }
            if (t < 0) t = 0;
            memcpy(p->tS + (long)nC * n2, S + (long)t * n2, sizeof(double) * n2);
            memcpy(p->tU + (long)nC * n2 * n2, U + (long)t * n2 * n2, sizeof(double) * n2 * n2);
            memcpy(p->tVT + (long)nC * n2 * n2, VT + (long)t * n2 * n2, sizeof(double) * n2 * n2);
            nC++;
        }
    if (h != 0.0) {  /* TrotterHeisenberg.cpp:80-88 */
        p->ssp = (double *)malloc(sizeof(double) * d);
        for (int i = 0; i < d; i++) p->ssp[i] = exp(0.5 * h * dtau * op_elem(OP_SZ, d, i, i));
    }
    p->gdim = gdim; p->Ng = Ng;
    p->grid = (double *)malloc(sizeof(double) * gdim * Ng);
    memcpy(p->grid, grid, sizeof(double) * gdim * Ng);
    if (build_terms) {  /* MPSQMC2.cpp:304-330 */
        p->terms = (omps *)calloc((size_t)p->nC * n2 * n2, sizeof(omps));
        for (int c = 0; c < p->nC; c++)
            for (int a = 0; a < n2; a++)
                for (int b = 0; b < n2; b++) {
                    omps *t = &p->terms[(long)c * n2 * n2 + a + n2 * b];
                    mps_copy(t, &p->trial);
                    apply_svd_term(p, t, c, a, b, 1);
                }
    }
    return p;
}

ORC_API void orc_pop_destroy(orc_pop *p) {
    if (!p) return;
    mps_free(&p->trial); mps_free(&p->htrial); mps_free(&p->hmin); free(p->Jfull);
    if (p->terms) {
        for (long i = 0; i < (long)p->nC * p->n2 * p->n2; i++) mps_free(&p->terms[i]);
        free(p->terms);
    }
    for (int i = 0; i < p->nw; i++) mps_free(&p->w[i]);
    free(p->w); free(p->wt); free(p->ov); free(p->eh); free(p->idx); free(p->EL);
    free(p->first); free(p->second); free(p->Jc); free(p->tS); free(p->tU); free(p->tVT);
    free(p->ssp); free(p->grid);
    free(p);
}

ORC_API void orc_pop_set_threads(orc_pop *p, int n) { p->nthreads = n < 1 ? 1 : n; }
ORC_API int orc_pop_ncouplings(const orc_pop *p) { return p->nC; }
ORC_API void orc_pop_couplings(const orc_pop *p, int *first, int *second, double *Jc) {
    for (int c = 0; c < p->nC; c++) { first[c] = p->first[c]; second[c] = p->second[c]; Jc[c] = p->Jc[c]; }
}
ORC_API int orc_pop_nwalkers(const orc_pop *p) { return p->nw; }

static void pop_reserve(orc_pop *p, int cap) {
    if (cap <= p->cap) return;
    p->w = (omps *)realloc(p->w, sizeof(omps) * cap);
    for (int i = p->cap; i < cap; i++) memset(&p->w[i], 0, sizeof(omps));
    p->wt = (double *)realloc(p->wt, sizeof(double) * cap);
    p->ov = (double *)realloc(p->ov, sizeof(double) * cap);
    p->eh = (double *)realloc(p->eh, sizeof(double) * cap);
    p->EL = (double *)realloc(p->EL, sizeof(double) * cap);
    p->cap = cap;
}

/* MPSQMC2::SetupWalkers with copyTrial = true (MPSQMC2.cpp:381-441). */
ORC_API void orc_pop_init_walkers(orc_pop *p, int N, int DW) {
    for (int i = 0; i < p->nw; i++) mps_free(&p->w[i]);
    p->nw = 0;
    pop_reserve(p, N);
    p->DW = DW;
    omps first;
    mps_copy(&first, &p->trial);
    double O = 1.0;
    if (DW < p->DT) {
        compress_state(&first, DW);
        O = set_overlap(&first, &p->trial);
    }
    for (int i = 0; i < N; i++) {
        mps_copy(&p->w[i], &first);
        p->wt[i] = 1.0; p->ov[i] = O; p->eh[i] = 0.0;
    }
    p->nw = N;
    mps_free(&first);
}

/* Load walkers from flat storage: vd[n*(L+1)], data concatenated in walker order. */
ORC_API void orc_pop_set_walkers(orc_pop *p, int n, const int *vd, const double *data, const double *w,
                                 const double *O, const double *eh) {
    for (int i = 0; i < p->nw; i++) mps_free(&p->w[i]);
    p->nw = 0;
    pop_reserve(p, n);
    long off = 0;
    for (int i = 0; i < n; i++) {
        const int *v = vd + (long)i * (p->L + 1);
        mps_from_flat(&p->w[i], p->L, p->d, v, data + off);
        off += flat_size(p->L, p->d, v);
        p->wt[i] = w[i]; p->ov[i] = O[i]; p->eh[i] = eh[i];
    }
    p->nw = n;
}

ORC_API long orc_pop_walker_size(const orc_pop *p, int i) { return flat_size(p->L, p->d, p->w[i].vd); }

ORC_API void orc_pop_get_walker(const orc_pop *p, int i, int *vd, double *data, double *w, double *O, double *eh) {
    memcpy(vd, p->w[i].vd, sizeof(int) * (p->L + 1));
    mps_to_flat(&p->w[i], data);
    *w = p->wt[i]; *O = p->ov[i]; *eh = p->eh[i];
}

ORC_API void orc_pop_get_scalars(const orc_pop *p, double *w, double *O, double *eh) {
    for (int i = 0; i < p->nw; i++) { w[i] = p->wt[i]; O[i] = p->ov[i]; eh[i] = p->eh[i]; }
}

ORC_API long orc_pop_get_trial(const orc_pop *p, int *vd, double *data) {
    memcpy(vd, p->trial.vd, sizeof(int) * (p->L + 1));
    if (data) mps_to_flat(&p->trial, data);
    return flat_size(p->L, p->d, p->trial.vd);
}

static void build_htrial(orc_pop *p);

ORC_API long orc_pop_get_htrial(orc_pop *p, int *vd, double *data) {
    build_htrial(p);
    memcpy(vd, p->htrial.vd, sizeof(int) * (p->L + 1));
    if (data) mps_to_flat(&p->htrial, data);
    return flat_size(p->L, p->d, p->htrial.vd);
}

/* One walker through one step of PropagateSeparately (MPSQMC2.cpp:532-600). */
static void propagate_walker(orc_pop *p, int i, const double *u /* nC uniforms */, int *idx,
                             double *combos, double *pdf) {
    int n2 = p->n2, N = p->Ng, dim = p->gdim;
    omps *m = &p->w[i];
    apply_one_site_everywhere(p, m);
    p->ov[i] = set_overlap(m, &p->trial);
    for (int c = 0; c < p->nC; c++) {
#pragma omp parallel for schedule(dynamic) num_threads(p->inner_threads) if (p->inner_threads > 1)
        for (int q = 0; q < n2 * n2; q++) combos[q] = inner(&p->terms[(long)c * n2 * n2 + q], m);
        double norm = 0.0;
        for (int g = 0; g < N; g++) {
            double v = 0.0;
            for (int l = 0; l < n2; l++)
                for (int r = 0; r < n2; r++) v += combos[l + n2 * r] * p->grid[l + dim * g] * p->grid[r + dim * g];
            v = v > 0.0 ? v : 0.0;
            v *= n2 / (p->ov[i] * N);
            pdf[g] = v;
            norm += v;
        }
        if (p->lab)
            for (int q = 0; q < n2 * n2; q++) p->lab[((long)i * p->nC + c) * n2 * n2 + q] = combos[q] / p->ov[i];
        double v_raw = 0.0; /* unnormalised PDF value of the learned sample */
        int lcnt = -1;
        double rn = u[c];
        if (p->pred) {
            const double *ph = p->pred + ((long)i * p->nC + c) * n2 * n2;
            double qs = 0.0;
            for (int g = 0; g < N; g++) {
                double v = 0.0;
                for (int l = 0; l < n2; l++)
                    for (int r = 0; r < n2; r++) v += ph[l + n2 * r] * p->grid[l + dim * g] * p->grid[r + dim * g];
                qs += v > 0.0 ? v : 0.0;
            }
            double a1 = qs > 0.0 ? (1.0 - p->mix) / qs : 0.0, a0 = qs > 0.0 ? p->mix / N : 1.0 / N;
            double qx;
            int cnt = 0;
            double cum = 0.0;
            for (int g = 0; g < N; g++) {
                double v = 0.0;
                for (int l = 0; l < n2; l++)
                    for (int r = 0; r < n2; r++) v += ph[l + n2 * r] * p->grid[l + dim * g] * p->grid[r + dim * g];
                cum += a1 * (v > 0.0 ? v : 0.0) + a0;
                if (g < N - 1 && cum <= rn) cnt++;
            }
            {
                double v = 0.0;
                for (int l = 0; l < n2; l++)
                    for (int r = 0; r < n2; r++) v += ph[l + n2 * r] * p->grid[l + dim * cnt] * p->grid[r + dim * cnt];
                qx = a1 * (v > 0.0 ? v : 0.0) + a0;
            }
            v_raw = pdf[cnt];
            if (v_raw > 0.0) { lcnt = cnt; p->wt[i] *= v_raw / qx; }
            else p->wt[i] = 0.0;
        }
        double inv = 1.0 / norm;
        for (int g = 0; g < N; g++) pdf[g] *= inv;
        if (!p->pred) p->wt[i] *= norm;
        int cnt = 0;
        double cum = pdf[0];
        while (cum <= rn && cnt < N - 1) { cnt++; cum += pdf[cnt]; }
        if (lcnt >= 0) cnt = lcnt;
        if (idx) idx[c] = cnt;
        apply_grid_term(p, m, c, cnt);
        p->ov[i] = set_overlap(m, &p->trial);
    }
    apply_one_site_everywhere(p, m);
    p->ov[i] = set_overlap(m, &p->trial);
}

/* MPSQMC2::PropagateSeparately: u holds nw*nC uniforms, walker-major.
   idx (optional) receives the sampled grid points, walker-major.  Returns sum w. */
ORC_API double orc_pop_propagate(orc_pop *p, const double *u, int *idx) {
    int nthr = p->nthreads;
#ifdef _OPENMP
    if (p->nw < nthr) {  /* few, large walkers: the threads share each walker's 16 overlaps */
        double *combos = (double *)malloc(sizeof(double) * p->n2 * p->n2);
        double *pdf = (double *)malloc(sizeof(double) * p->Ng);
        double s = 0.0;
        p->inner_threads = nthr;
        for (int i = 0; i < p->nw; i++) {
            propagate_walker(p, i, u + (long)i * p->nC, idx ? idx + (long)i * p->nC : NULL, combos, pdf);
            s += p->wt[i];
        }
        p->inner_threads = 1;
        free(combos); free(pdf);
        return s;
    }
    double *sums = (double *)calloc(nthr, sizeof(double));
#pragma omp parallel num_threads(nthr)
    {
        int tid = omp_get_thread_num();
        double *combos = (double *)malloc(sizeof(double) * p->n2 * p->n2);
        double *pdf = (double *)malloc(sizeof(double) * p->Ng);
#pragma omp for schedule(static)
        for (int i = 0; i < p->nw; i++) {
            propagate_walker(p, i, u + (long)i * p->nC, idx ? idx + (long)i * p->nC : NULL, combos, pdf);
            sums[tid] += p->wt[i];
        }
        free(combos); free(pdf);
    }
    double s = 0.0;
    for (int t = 0; t < nthr; t++) s += sums[t];
    free(sums);
    return s;
#else
    (void)nthr;
    double *combos = (double *)malloc(sizeof(double) * p->n2 * p->n2);
    double *pdf = (double *)malloc(sizeof(double) * p->Ng);
    double s = 0.0;
    for (int i = 0; i < p->nw; i++) {
        propagate_walker(p, i, u + (long)i * p->nC, idx ? idx + (long)i * p->nC : NULL, combos, pdf);
        s += p->wt[i];
    }
    free(combos); free(pdf);
    return s;
#endif
}

/* <trial| H |phi> as sum over couplings of J * sum_a sgn_a <trial|S^a_i S^a_j|phi>
   minus h sum_k <trial|S^z_k|phi>: the same number as <H trial|phi>
   (HeisenbergMPO.h:98-103 convention; iSy enters with a minus sign). */
static double h_insert(const orc_pop *p, const omps *phi, int nthr) {
    int d = p->d;
    const int ops[3] = {OP_SX, OP_ISY, OP_SZ};
    const double sg[3] = {1.0, -1.0, 1.0};
    double tot = 0.0;
#pragma omp parallel for schedule(dynamic) reduction(+ : tot) num_threads(nthr) if (nthr > 1)
    for (int q = 0; q < 3 * p->nC; q++) {
        const int c = q / 3, a = q % 3;
        double O[16];
        omps t; mps_copy(&t, phi);
        orc_operator(ops[a], d, O);
        apply_local(&t, p->first[c], O);
        apply_local(&t, p->second[c], O);
        tot += sg[a] * p->Jc[c] * inner(&p->trial, &t);
        mps_free(&t);
    }
    if (p->h != 0.0) {
        double O[16];
        orc_operator(OP_SZ, d, O);
        for (int k = 0; k < p->L; k++) {
            omps t; mps_copy(&t, phi);
            apply_local(&t, k, O);
            tot -= p->h * inner(&p->trial, &t);
            mps_free(&t);
        }
    }
    return tot;
}

ORC_API void orc_pop_set_energy_mode(orc_pop *p, int mode) { p->emode = mode; }

/* H|trial> with the reference's MPO (the lossless compression drops only
   Schmidt values < 1e-13 and is skipped here), and with the minimal MPO. */
static void build_htrial(orc_pop *p) {
    if (p->htrial.t) return;
    ompo M; M.L = p->L; M.d = p->d; M.J = p->Jfull; M.h = p->h;
    M.dims = (int *)malloc(sizeof(int) * (p->L + 1));
    mpo_dims(p->L, M.dims);
    apply_mpo(&M, &p->trial, &p->htrial);
    free(M.dims);
}

static void build_hmin(orc_pop *p) {
    if (!p->hmin.t) apply_mpo_min(p->L, p->d, p->Jfull, p->h, &p->trial, &p->hmin);
}

/* MPSQMC2::EnergyFunctionAndHistory (MPSQMC2.cpp:688-747), non-MPI branch.
   Returns Omega; *EP the projected energy; EL (optional) per-walker energies. */
ORC_API double orc_pop_energy(orc_pop *p, int step, int dofluct, double *EP, double *EL) {
    double s2 = 0.0, s1 = 0.0, num = 0.0, den = 0.0;
    if (p->emode == 0) build_htrial(p);
    if (p->emode == 2) build_hmin(p);
    double *e = p->EL;
    /* few walkers: operator insertion shares each walker's terms among the threads */
    const int inner_thr = (p->emode == 1 && p->nw < p->nthreads) ? p->nthreads : 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(p->nthreads) if (inner_thr == 1)
#endif
    for (int i = 0; i < p->nw; i++)
        e[i] = (p->emode == 0 ? inner(&p->htrial, &p->w[i])
                : (p->emode == 2 ? inner(&p->hmin, &p->w[i]) : h_insert(p, &p->w[i], inner_thr))) / p->ov[i];
    for (int i = 0; i < p->nw; i++) {
        num += p->wt[i] * e[i];
        den += p->wt[i];
        if (dofluct) {
            p->eh[i] += e[i];
            double a = p->eh[i] / step;
            s1 += a; s2 += a * a;
        }
        if (EL) EL[i] = e[i];
    }
    double om = 0.0;
    if (dofluct) om = (s2 - s1 * s1 / p->nw) / p->nw;
    *EP = num / den;
    return om;
}

/* MPSQMC2::SeparatePopulationControl (MPSQMC2.cpp:613-686).  u supplies the
   draws in order, consumed only by walkers with s < 0.25 or s > 1.5; *used
   receives the count.  Capacity overflow is reported as -1 (no writes past
   the arrays, unlike MPSQMC2.cpp:650-662).  Returns the new population size. */
ORC_API int orc_pop_popcontrol(orc_pop *p, double scaling, const double *u, int nu, int cap, int *used,
                               int *ncopies_out) {
    int n = p->nw, k = 0, total = 0;
    int *nc = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    double *nw = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
        double s = p->wt[i] * scaling;
        int c = 1; double w = s;
        if (w < 0.25) {
            double r = (k < nu) ? u[k] : 0.0; k++;
            c = (int)(s + r); w = 1.0;
        }
        if (w > 1.5) {
            double r = (k < nu) ? u[k] : 0.0; k++;
            c = (int)(s + r); w = s / c;
        }
        nc[i] = c; nw[i] = w; total += c;
        if (ncopies_out) ncopies_out[i] = c;
    }
    *used = k;
    if (total > cap) { free(nc); free(nw); return -1; }
    omps *neww = (omps *)calloc(total > 0 ? total : 1, sizeof(omps));
    double *wt = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
    double *ov = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
    double *eh = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
    int pos = 0;
    for (int i = 0; i < n; i++) {
        if (nc[i] == 0) { mps_free(&p->w[i]); continue; }
        neww[pos] = p->w[i];
        wt[pos] = nw[i]; ov[pos] = p->ov[i]; eh[pos] = p->eh[i];
        for (int c = 1; c < nc[i]; c++) {
            mps_copy(&neww[pos + c], &p->w[i]);
            wt[pos + c] = nw[i]; ov[pos + c] = p->ov[i]; eh[pos + c] = p->eh[i];
        }
        pos += nc[i];
    }
    free(p->w); free(p->wt); free(p->ov); free(p->eh); free(p->EL);
    p->w = neww; p->wt = wt; p->ov = ov; p->eh = eh;
    p->EL = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
    p->nw = total; p->cap = total > 0 ? total : 1;
    free(nc); free(nw);
    return total;
}

/*
 * MPSQMC2::Walk (MPSQMC2.cpp:443-513) with the build's RNG contract: one
 * mt19937 stream seeded with `seed`; per step nw*nC draws (walker-major) for
 * the propagation, then the population-control draws in walker order.
 * trace receives per step: step, N, E_P, E_T, Omega, sum_w (6 doubles);
 * E_start receives the energy at start.  Returns 0, or -1 on capacity overflow.
 */
ORC_API int orc_pop_walk(orc_pop *p, int nsteps, uint32_t seed, int Ndesired, int cap, double *E_start,
                         double *trace) {
    orc_mt g; mt_seed(&g, seed);
    double EP = 0.0;
    (void)orc_pop_energy(p, 1, 0, &EP, NULL);
    if (E_start) *E_start = EP;
    double *u = NULL; long ucap = 0;
    for (int step = 1; step <= nsteps; step++) {
        long need = (long)p->nw * p->nC;
        if (need > ucap) { ucap = need; u = (double *)realloc(u, sizeof(double) * (ucap > 0 ? ucap : 1)); }
        for (long i = 0; i < need; i++) u[i] = mt_uniform(&g);
        double sw = orc_pop_propagate(p, u, NULL);
        int Ncur = p->nw;
        double avg = sw / Ncur;
        double scaling = Ndesired / (Ncur * avg);
        double ET = log(scaling) / p->dtau;
        double om = orc_pop_energy(p, step, 1, &EP, NULL);
        double *row = trace + 6L * (step - 1);
        row[0] = step; row[1] = Ncur; row[2] = EP; row[3] = ET; row[4] = om; row[5] = sw;
        /* population control: draw lazily from the same stream */
        int n = p->nw;
        double *pu = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
        int k = 0;
        for (int i = 0; i < n; i++) {
            double s = p->wt[i] * scaling;
            if (s < 0.25 || s > 1.5) pu[k++] = mt_uniform(&g);
        }
        int used = 0;
        int r = orc_pop_popcontrol(p, scaling, pu, k, cap, &used, NULL);
        free(pu);
        if (r < 0) { free(u); return -1; }
    }
    free(u);
    return 0;
}
