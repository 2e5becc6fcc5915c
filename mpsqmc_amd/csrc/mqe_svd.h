// mqe_svd.h -- FP64 singular value decomposition of one matrix per workgroup
// (one-sided Jacobi, Hestenes), used by the device recompression of walkers
// (MPSstate::CompressState, MPSstate.cpp:379-404; two-site SVD of
// TwoSiteObject.cpp:119-184).
//
// The columns of X are rotated pairwise until they are mutually orthogonal:
// X_in V = X_out with V orthogonal, so X_in = U S V^T with U_j = X_out_j / s_j,
// s_j = |X_out_j|.  All column pairs of one round of a round-robin tournament
// are disjoint and rotate in parallel (a group of lanes per pair, dot products
// by shuffles within the group), one workgroup barrier per round; a sweep is
// one tournament, and sweeps repeat until one rotates nothing.
#ifndef MQE_SVD_H
#define MQE_SVD_H

#include "mqe_device.h"

namespace mqe {

// X: rows x cols (column-major, ld rows), V: cols x cols (ld cols), sig: cols.
// sh_flag: one int of LDS.  Whole workgroup.
__device__ inline void jacobi_svd_columns(double *X, int rows, int cols, double *V, double *sig, int *sh_flag) {
    const int tid = threadIdx.x;
    for (int e = tid; e < cols * cols; e += MQE_NT) V[e] = (e % cols == e / cols) ? 1.0 : 0.0;
    const int np = (cols + 1) / 2;   // pairs per round
    const int players = 2 * np;      // cols rounded up to even (one dummy player if odd)
    int tpp = 64;                    // lanes per pair: a power of two within one wave
    while (tpp > 1 && tpp * np > MQE_NT) tpp >>= 1;
    const int groups = MQE_NT / tpp, gl = tid & (tpp - 1), grp = tid / tpp;
    // orthogonality tolerance: the rounding floor of a dot product grows like sqrt(rows)
    const double tol = 1e-15 * fmax(1.0, sqrt((double)rows) * 0.25);
    __syncthreads();
    for (int sweep = 0; sweep < 60; ++sweep) {
        if (tid == 0) *sh_flag = 0;
        __syncthreads();
        for (int r = 0; r < players - 1; ++r) {
            for (int pr = grp; pr < np; pr += groups) {
                const int pa = pr, pb = players - 1 - pr;
                int p = pa == 0 ? 0 : ((pa - 1 + r) % (players - 1)) + 1;
                int q = pb == 0 ? 0 : ((pb - 1 + r) % (players - 1)) + 1;
                if (p >= cols || q >= cols) continue;  // paired with the dummy
                if (p > q) { const int t = p; p = q; q = t; }
                double *xp = X + (size_t)rows * p, *xq = X + (size_t)rows * q;
                double a = 0.0, b = 0.0, g = 0.0;
                for (int i = gl; i < rows; i += tpp) {
                    const double u = xp[i], v = xq[i];
                    a += u * u;
                    b += v * v;
                    g += u * v;
                }
                for (int off = 1; off < tpp; off <<= 1) {
                    a += __shfl_xor(a, off, 64);
                    b += __shfl_xor(b, off, 64);
                    g += __shfl_xor(g, off, 64);
                }
                if (g == 0.0 || fabs(g) <= tol * sqrt(a * b)) continue;
                const double zeta = (b - a) / (2.0 * g);
                const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int i = gl; i < rows; i += tpp) {
                    const double u = xp[i], v = xq[i];
                    xp[i] = c * u - s * v;
                    xq[i] = s * u + c * v;
                }
                double *vp = V + (size_t)cols * p, *vq = V + (size_t)cols * q;
                for (int i = gl; i < cols; i += tpp) {
                    const double u = vp[i], v = vq[i];
                    vp[i] = c * u - s * v;
                    vq[i] = s * u + c * v;
                }
                if (gl == 0) atomicOr(sh_flag, 1);
            }
            __syncthreads();
        }
        const int rotated = *sh_flag;
        __syncthreads();  // every thread has read the flag before it is reset
        if (!rotated) break;
    }
    for (int j = tid; j < cols; j += MQE_NT) {
        const double *x = X + (size_t)rows * j;
        double s2 = 0.0;
        for (int i = 0; i < rows; ++i) s2 += x[i] * x[i];
        sig[j] = sqrt(s2);
    }
    __syncthreads();
}

// rank[j]: position of column j when the values are sorted in decreasing order
// (ties by index); ord[rank[j]] = j.  Whole workgroup.
__device__ inline void sort_descending(const double *sig, int n, int *ord) {
    for (int j = threadIdx.x; j < n; j += MQE_NT) {
        const double s = sig[j];
        int rk = 0;
        for (int i = 0; i < n; ++i) rk += (sig[i] > s || (sig[i] == s && i < j)) ? 1 : 0;
        ord[rk] = j;
    }
    __syncthreads();
}

}  // namespace mqe

#endif
