// mqe_dmrg.cpp -- trial wave functions by two-site DMRG on the host (FP64).
//
// The reference obtains its trial state Psi_T by two-site DMRG sweeps over an
// MPSstate with a fixed bond profile (DMRG.cpp:50-100, called from
// MPSQMC2::SetupTrial, MPSQMC2.cpp:225-236), then left-normalises it.  This
// file does the same job for H = sum_c J_c S_i.S_j - h sum_i S^z_i with
// arbitrary couplings (chains, snake-mapped lattices):
//   * the Hamiltonian as a minimal-channel MPO (channels: "start", "done",
//     and S+, S-, Sz of every site whose couplings cross the bond);
//   * left/right environments, a matrix-free two-site effective Hamiltonian,
//     Lanczos with full re-orthogonalisation for its lowest eigenvector;
//   * SVD split keeping the bond profile of the initial state (no Schmidt
//     cut-off, as DMRG.cpp:81,97 decomposes with possiblyCompress = false);
//   * sweeps until the energy changes by less than tol (DMRG.cpp:59).
// Setup-time code: one trial per run.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "mqe.h"
#include "mqe_internal.h"

namespace {

using Vec = std::vector<double>;

// C(m x n) = alpha op(A) op(B) + beta C, column-major.
void gemm(bool ta, bool tb, int m, int n, int k, double alpha, const double *A, int lda, const double *B, int ldb,
          double beta, double *C, int ldc) {
    for (int j = 0; j < n; ++j) {
        double *c = C + (size_t)ldc * j;
        if (beta == 0.0)
            for (int i = 0; i < m; ++i) c[i] = 0.0;
        else if (beta != 1.0)
            for (int i = 0; i < m; ++i) c[i] *= beta;
        if (!ta) {
            for (int p = 0; p < k; ++p) {
                const double b = alpha * (tb ? B[j + (size_t)ldb * p] : B[p + (size_t)ldb * j]);
                if (b == 0.0) continue;
                const double *a = A + (size_t)lda * p;
                for (int i = 0; i < m; ++i) c[i] += a[i] * b;
            }
        } else {
            for (int i = 0; i < m; ++i) {
                const double *a = A + (size_t)lda * i;
                double s = 0.0;
                if (!tb) {
                    const double *b = B + (size_t)ldb * j;
                    for (int p = 0; p < k; ++p) s += a[p] * b[p];
                } else {
                    for (int p = 0; p < k; ++p) s += a[p] * B[j + (size_t)ldb * p];
                }
                c[i] += alpha * s;
            }
        }
    }
}

// local operators, basis ordered smallest m first (OpSz.cpp:9): op[s' + 2 s]
enum { OP_I = 0, OP_SP = 1, OP_SM = 2, OP_SZ = 3 };
const double OPS[4][4] = {
    {1.0, 0.0, 0.0, 1.0},    // I
    {0.0, 1.0, 0.0, 0.0},    // S+ : |-1/2> -> |+1/2>
    {0.0, 0.0, 1.0, 0.0},    // S- : |+1/2> -> |-1/2>
    {-0.5, 0.0, 0.0, 0.5},   // Sz
};

struct Term {  // one non-zero operator entry of an MPO site tensor
    int wl, wr;
    double coef;
    int op;
};

struct Mpo {
    std::vector<int> W;                   // channel count at bonds 0..L
    std::vector<std::vector<Term>> site;  // terms of site k
};

// Minimal-channel MPO of H = sum_c J_c S_i.S_j - h sum_i S^z_i, with
// S_i.S_j = (S+_i S-_j + S-_i S+_j)/2 + Sz_i Sz_j.  Channel 0 = nothing placed
// yet, 1 = term complete, 2 + 3p + {0,1,2} = S+, S-, Sz placed on the p-th
// site whose couplings cross the bond.
Mpo build_mpo(int L, int nC, const int *first, const int *second, const double *J, double h) {
    std::vector<std::vector<double>> Jm(L, std::vector<double>(L, 0.0));
    for (int c = 0; c < nC; ++c) {
        int i = std::min(first[c], second[c]), j = std::max(first[c], second[c]);
        Jm[i][j] += J[c];
    }
    std::vector<int> last(L, -1);  // last partner j > i of site i
    for (int i = 0; i < L; ++i)
        for (int j = i + 1; j < L; ++j)
            if (Jm[i][j] != 0.0) last[i] = j;
    // open sites at bond k: i < k with last[i] >= k
    std::vector<std::vector<int>> open(L + 1);
    for (int k = 0; k <= L; ++k)
        for (int i = 0; i < k; ++i)
            if (last[i] >= k) open[k].push_back(i);
    Mpo m;
    m.W.resize(L + 1);
    for (int k = 0; k <= L; ++k) m.W[k] = 2 + 3 * (int)open[k].size();
    m.site.resize(L);
    for (int k = 0; k < L; ++k) {
        auto &t = m.site[k];
        t.push_back({0, 0, 1.0, OP_I});
        t.push_back({1, 1, 1.0, OP_I});
        if (h != 0.0) t.push_back({0, 1, -h, OP_SZ});
        auto pos_r = [&](int i) {
            for (size_t p = 0; p < open[k + 1].size(); ++p)
                if (open[k + 1][p] == i) return (int)p;
            return -1;
        };
        const int pk = pos_r(k);
        if (pk >= 0) {
            t.push_back({0, 2 + 3 * pk + 0, 1.0, OP_SP});
            t.push_back({0, 2 + 3 * pk + 1, 1.0, OP_SM});
            t.push_back({0, 2 + 3 * pk + 2, 1.0, OP_SZ});
        }
        for (size_t p = 0; p < open[k].size(); ++p) {
            const int i = open[k][p], wl = 2 + 3 * (int)p;
            const double Jik = Jm[i][k];
            if (Jik != 0.0) {
                t.push_back({wl + 0, 1, 0.5 * Jik, OP_SM});
                t.push_back({wl + 1, 1, 0.5 * Jik, OP_SP});
                t.push_back({wl + 2, 1, Jik, OP_SZ});
            }
            const int pr = pos_r(i);
            if (pr >= 0)
                for (int q = 0; q < 3; ++q) t.push_back({wl + q, 2 + 3 * pr + q, 1.0, OP_I});
        }
    }
    return m;
}

struct State {
    int L;
    std::vector<int> vd;
    std::vector<Vec> A;  // A[k][row + dl*(col + dr*s)]
};

// environments: E[bra + D*(w + W*ket)]
struct Dmrg {
    State st;
    Mpo mpo;
    std::vector<Vec> LE, RE;  // LE[k]: bond k from the left; RE[k]: bond k from the right

    int D(int k) const { return st.vd[k]; }

    void left_update(int k) {  // LE[k+1] from LE[k] and A_k
        const int dl = D(k), dr = D(k + 1), Wl = mpo.W[k], Wr = mpo.W[k + 1];
        const Vec &E = LE[k];
        const Vec &A = st.A[k];
        // T[w,s](a', b) = LE_w(a', a) A[s](a, b)
        Vec T((size_t)Wl * 2 * dl * dr);
        for (int w = 0; w < Wl; ++w)
            for (int s = 0; s < 2; ++s)
                gemm(false, false, dl, dr, dl, 1.0, &E[(size_t)dl * w], dl * Wl, &A[(size_t)dl * dr * s], dl, 0.0,
                     &T[((size_t)w * 2 + s) * dl * dr], dl);
        // U[w', s'](a', b) = sum W[w->w'](s', s) T[w, s]
        Vec U((size_t)Wr * 2 * dl * dr, 0.0);
        for (const Term &t : mpo.site[k])
            for (int sp = 0; sp < 2; ++sp)
                for (int s = 0; s < 2; ++s) {
                    const double f = t.coef * OPS[t.op][sp + 2 * s];
                    if (f == 0.0) continue;
                    const double *src = &T[((size_t)t.wl * 2 + s) * dl * dr];
                    double *dst = &U[((size_t)t.wr * 2 + sp) * dl * dr];
                    for (size_t e = 0; e < (size_t)dl * dr; ++e) dst[e] += f * src[e];
                }
        // LE[k+1]_{w'}(b', b) = sum_s' A[s']^T U[w', s']
        Vec N((size_t)dr * Wr * dr, 0.0);
        for (int w = 0; w < Wr; ++w)
            for (int sp = 0; sp < 2; ++sp)
                gemm(true, false, dr, dr, dl, 1.0, &A[(size_t)dl * dr * sp], dl, &U[((size_t)w * 2 + sp) * dl * dr], dl,
                     1.0, &N[(size_t)dr * w], dr * Wr);
        LE[k + 1].swap(N);
    }

    void right_update(int k) {  // RE[k] from RE[k+1] and A_k
        const int dl = D(k), dr = D(k + 1), Wl = mpo.W[k], Wr = mpo.W[k + 1];
        const Vec &E = RE[k + 1];
        const Vec &A = st.A[k];
        // T[w', s](b', a) = RE_w'(b', b) A[s](a, b)^T
        Vec T((size_t)Wr * 2 * dr * dl);
        for (int w = 0; w < Wr; ++w)
            for (int s = 0; s < 2; ++s)
                gemm(false, true, dr, dl, dr, 1.0, &E[(size_t)dr * w], dr * Wr, &A[(size_t)dl * dr * s], dl, 0.0,
                     &T[((size_t)w * 2 + s) * dr * dl], dr);
        Vec U((size_t)Wl * 2 * dr * dl, 0.0);
        for (const Term &t : mpo.site[k])
            for (int sp = 0; sp < 2; ++sp)
                for (int s = 0; s < 2; ++s) {
                    const double f = t.coef * OPS[t.op][sp + 2 * s];
                    if (f == 0.0) continue;
                    const double *src = &T[((size_t)t.wr * 2 + s) * dr * dl];
                    double *dst = &U[((size_t)t.wl * 2 + sp) * dr * dl];
                    for (size_t e = 0; e < (size_t)dr * dl; ++e) dst[e] += f * src[e];
                }
        // RE[k]_w(a', a) = sum_s' A[s'](a', b') U[w, s'](b', a)
        Vec N((size_t)dl * Wl * dl, 0.0);
        for (int w = 0; w < Wl; ++w)
            for (int sp = 0; sp < 2; ++sp)
                gemm(false, false, dl, dl, dr, 1.0, &A[(size_t)dl * dr * sp], dl, &U[((size_t)w * 2 + sp) * dr * dl], dr,
                     1.0, &N[(size_t)dl * w], dl * Wl);
        RE[k].swap(N);
    }

    // out = H_eff theta on sites (k, k+1); theta[(a + dl s1) + rows (b + dr s2)]
    void apply_heff(int k, const double *th, double *out) const {
        const int dl = D(k), dr = D(k + 2), Wl = mpo.W[k], Wm = mpo.W[k + 1], Wr = mpo.W[k + 2];
        const int rows = 2 * dl;
        const size_t blk = (size_t)dl * dr;  // one (s1, s2) block (a, b)
        // theta block (s1, s2): element (a, b) at th[(a + dl s1) + rows (b + dr s2)]
        auto thp = [&](int s1, int s2) { return th + dl * s1 + (size_t)rows * dr * s2; };
        // X[w, s1, s2](a', b) = LE_w(a', a) theta[s1, s2](a, b)
        Vec X((size_t)Wl * 4 * blk);
        for (int w = 0; w < Wl; ++w)
            for (int s1 = 0; s1 < 2; ++s1)
                for (int s2 = 0; s2 < 2; ++s2)
                    gemm(false, false, dl, dr, dl, 1.0, &LE[k][(size_t)dl * w], dl * Wl, thp(s1, s2), rows, 0.0,
                         &X[((size_t)w * 4 + s1 + 2 * s2) * blk], dl);
        Vec Y((size_t)Wm * 4 * blk, 0.0);
        for (const Term &t : mpo.site[k])
            for (int sp = 0; sp < 2; ++sp)
                for (int s = 0; s < 2; ++s) {
                    const double f = t.coef * OPS[t.op][sp + 2 * s];
                    if (f == 0.0) continue;
                    for (int s2 = 0; s2 < 2; ++s2) {
                        const double *src = &X[((size_t)t.wl * 4 + s + 2 * s2) * blk];
                        double *dst = &Y[((size_t)t.wr * 4 + sp + 2 * s2) * blk];
                        for (size_t e = 0; e < blk; ++e) dst[e] += f * src[e];
                    }
                }
        Vec Z((size_t)Wr * 4 * blk, 0.0);
        for (const Term &t : mpo.site[k + 1])
            for (int sp = 0; sp < 2; ++sp)
                for (int s = 0; s < 2; ++s) {
                    const double f = t.coef * OPS[t.op][sp + 2 * s];
                    if (f == 0.0) continue;
                    for (int s1 = 0; s1 < 2; ++s1) {
                        const double *src = &Y[((size_t)t.wl * 4 + s1 + 2 * s) * blk];
                        double *dst = &Z[((size_t)t.wr * 4 + s1 + 2 * sp) * blk];
                        for (size_t e = 0; e < blk; ++e) dst[e] += f * src[e];
                    }
                }
        // out[s1, s2](a', b') = sum_w Z[w, s1, s2](a', b) RE_w(b', b)^T
        for (size_t e = 0; e < (size_t)rows * 2 * dr; ++e) out[e] = 0.0;
        for (int w = 0; w < Wr; ++w)
            for (int s1 = 0; s1 < 2; ++s1)
                for (int s2 = 0; s2 < 2; ++s2)
                    gemm(false, true, dl, dr, dr, 1.0, &Z[((size_t)w * 4 + s1 + 2 * s2) * blk], dl,
                         &RE[k + 2][(size_t)dr * w], dr * Wr, 1.0, out + dl * s1 + (size_t)rows * dr * s2, rows);
    }

    // lowest eigenpair of H_eff by Lanczos (full re-orthogonalisation); theta in/out
    double lanczos(int k, Vec &th) const {
        const size_t n = th.size();
        const int maxit = (int)std::min<size_t>(60, n);
        double E = 0.0;
        for (int restart = 0; restart < 4; ++restart) {
            std::vector<Vec> V;
            Vec a, b;
            double nrm = 0.0;
            for (double x : th) nrm += x * x;
            nrm = std::sqrt(nrm);
            if (!(nrm > 0.0)) {
                std::mt19937 g(12345u + k);
                for (double &x : th) x = (double)g() / 4294967296.0 - 0.5;
                nrm = 0.0;
                for (double x : th) nrm += x * x;
                nrm = std::sqrt(nrm);
            }
            V.push_back(th);
            for (double &x : V[0]) x /= nrm;
            Vec w(n), evec;
            bool conv = false;
            for (int it = 0; it < maxit; ++it) {
                apply_heff(k, V[it].data(), w.data());
                double al = 0.0;
                for (size_t e = 0; e < n; ++e) al += w[e] * V[it][e];
                a.push_back(al);
                for (int rep = 0; rep < 2; ++rep)
                    for (size_t q = 0; q < V.size(); ++q) {
                        double dq = 0.0;
                        for (size_t e = 0; e < n; ++e) dq += w[e] * V[q][e];
                        for (size_t e = 0; e < n; ++e) w[e] -= dq * V[q][e];
                    }
                double be = 0.0;
                for (double x : w) be += x * x;
                be = std::sqrt(be);
                // lowest eigenpair of the tridiagonal (a, b) by Jacobi rotations
                const int m = (int)a.size();
                Vec T((size_t)m * m, 0.0), Q((size_t)m * m, 0.0);
                for (int i = 0; i < m; ++i) {
                    T[i + (size_t)m * i] = a[i];
                    Q[i + (size_t)m * i] = 1.0;
                    if (i + 1 < m) T[i + (size_t)m * (i + 1)] = T[(i + 1) + (size_t)m * i] = b[i];
                }
                for (int sw = 0; sw < 100; ++sw) {
                    double off = 0.0;
                    for (int p = 0; p < m; ++p)
                        for (int q = p + 1; q < m; ++q) off += T[p + (size_t)m * q] * T[p + (size_t)m * q];
                    if (off < 1e-30) break;
                    for (int p = 0; p < m; ++p)
                        for (int q = p + 1; q < m; ++q) {
                            const double apq = T[p + (size_t)m * q];
                            if (std::fabs(apq) < 1e-300) continue;
                            const double th2 = (T[q + (size_t)m * q] - T[p + (size_t)m * p]) / (2.0 * apq);
                            const double t = (th2 >= 0 ? 1.0 : -1.0) / (std::fabs(th2) + std::sqrt(th2 * th2 + 1.0));
                            const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                            for (int r = 0; r < m; ++r) {  // columns p, q
                                const double x = T[r + (size_t)m * p], y = T[r + (size_t)m * q];
                                T[r + (size_t)m * p] = c * x - s * y;
                                T[r + (size_t)m * q] = s * x + c * y;
                            }
                            for (int r = 0; r < m; ++r) {  // rows p, q
                                const double x = T[p + (size_t)m * r], y = T[q + (size_t)m * r];
                                T[p + (size_t)m * r] = c * x - s * y;
                                T[q + (size_t)m * r] = s * x + c * y;
                            }
                            for (int r = 0; r < m; ++r) {
                                const double x = Q[r + (size_t)m * p], y = Q[r + (size_t)m * q];
                                Q[r + (size_t)m * p] = c * x - s * y;
                                Q[r + (size_t)m * q] = s * x + c * y;
                            }
                        }
                }
                int lo = 0;
                for (int i = 1; i < m; ++i)
                    if (T[i + (size_t)m * i] < T[lo + (size_t)m * lo]) lo = i;
                E = T[lo + (size_t)m * lo];
                evec.assign(Q.begin() + (size_t)m * lo, Q.begin() + (size_t)m * (lo + 1));
                const double res = std::fabs(be * evec[m - 1]);
                if (res < 1e-10 * std::max(1.0, std::fabs(E)) || be < 1e-14 || it + 1 == maxit) {
                    conv = res < 1e-10 * std::max(1.0, std::fabs(E)) || be < 1e-14;
                    break;
                }
                b.push_back(be);
                for (double &x : w) x /= be;
                V.push_back(w);
            }
            for (size_t e = 0; e < n; ++e) th[e] = 0.0;
            for (size_t q = 0; q < evec.size(); ++q)
                for (size_t e = 0; e < n; ++e) th[e] += evec[q] * V[q][e];
            if (conv) break;
        }
        return E;
    }

    // optimise the pair (k, k+1); the new bond keeps its dimension
    double optimise(int k, bool movingright) {
        const int dl = D(k), dm = D(k + 1), dr = D(k + 2), rows = 2 * dl, cols = 2 * dr;
        Vec th((size_t)rows * cols, 0.0);
        // theta[(a + dl s1) + rows (b + dr s2)] = sum_m A_k[s1](a, m) A_{k+1}[s2](m, b)
        for (int s1 = 0; s1 < 2; ++s1)
            for (int s2 = 0; s2 < 2; ++s2)
                gemm(false, false, dl, dr, dm, 1.0, &st.A[k][(size_t)dl * dm * s1], dl, &st.A[k + 1][(size_t)dm * dr * s2],
                     dm, 0.0, &th[dl * s1 + (size_t)rows * dr * s2], rows);
        const double E = lanczos(k, th);
        const int kmin = std::min(rows, cols);
        Vec s(kmin), u((size_t)rows * kmin), vt((size_t)kmin * cols);
        mqe::host_svd(rows, cols, th.data(), s.data(), u.data(), vt.data());
        const int keep = std::min(dm, kmin);
        Vec Al((size_t)dl * dm * 2, 0.0), Ar((size_t)dm * dr * 2, 0.0);
        for (int s1 = 0; s1 < 2; ++s1)
            for (int c = 0; c < keep; ++c)
                for (int a = 0; a < dl; ++a)
                    Al[a + (size_t)dl * (c + (size_t)dm * s1)] =
                        u[(a + dl * s1) + (size_t)rows * c] * (movingright ? 1.0 : s[c]);
        for (int s2 = 0; s2 < 2; ++s2)
            for (int b = 0; b < dr; ++b)
                for (int c = 0; c < keep; ++c)
                    Ar[c + (size_t)dm * (b + (size_t)dr * s2)] =
                        vt[c + (size_t)kmin * (b + dr * s2)] * (movingright ? s[c] : 1.0);
        st.A[k].swap(Al);
        st.A[k + 1].swap(Ar);
        return E;
    }
};

// right-normalise sites L-1 .. 1 (sites >= 1 right-orthonormal)
void right_normalize(State &st) {
    for (int k = st.L - 1; k >= 1; --k) {
        const int dl = st.vd[k], dr = st.vd[k + 1], cols = 2 * dr;
        // M^T (cols x dl) = Q R  =>  M = R^T Q^T
        Vec mt((size_t)cols * dl), r((size_t)dl * dl);
        for (int s = 0; s < 2; ++s)
            for (int c = 0; c < dr; ++c)
                for (int a = 0; a < dl; ++a) mt[(c + dr * s) + (size_t)cols * a] = st.A[k][a + (size_t)dl * (c + (size_t)dr * s)];
        mqe::host_qr(cols, dl, mt.data(), r.data());
        for (int s = 0; s < 2; ++s)
            for (int c = 0; c < dr; ++c)
                for (int a = 0; a < dl; ++a) st.A[k][a + (size_t)dl * (c + (size_t)dr * s)] = mt[(c + dr * s) + (size_t)cols * a];
        // A_{k-1}[s] <- A_{k-1}[s] R^T
        const int dp = st.vd[k - 1];
        Vec nt((size_t)dp * dl * 2);
        for (int s = 0; s < 2; ++s)
            gemm(false, true, dp, dl, dl, 1.0, &st.A[k - 1][(size_t)dp * dl * s], dp, r.data(), dl, 0.0,
                 &nt[(size_t)dp * dl * s], dp);
        st.A[k - 1].swap(nt);
    }
}

}  // namespace

extern "C" int mqe_dmrg(int L, int d, int n_couplings, const int *first, const int *second, const double *J,
                        double field, int D, int max_sweeps, double tol, unsigned int seed, int *vd_out,
                        double *data_out, double *energy) {
    if (L < 2 || d != 2 || D < 1 || n_couplings < 0 || (n_couplings > 0 && (!first || !second || !J)) || !vd_out ||
        !data_out || max_sweeps < 1) {
        mqe::set_error("mqe_dmrg: bad arguments (d must be 2)");
        return MQE_EINVAL;
    }
    for (int c = 0; c < n_couplings; ++c)
        if (first[c] < 0 || first[c] >= L || second[c] < 0 || second[c] >= L || first[c] == second[c]) {
            mqe::set_error("mqe_dmrg: coupling site out of range");
            return MQE_EINVAL;
        }
    Dmrg g;
    g.st.L = L;
    g.st.vd.resize(L + 1);
    mqe_bond_dims(L, D, d, g.st.vd.data());
    // random start, entries uniform in (-1, 1) from one seeded stream
    std::mt19937 rng(seed);
    g.st.A.resize(L);
    for (int k = 0; k < L; ++k) {
        g.st.A[k].resize((size_t)g.st.vd[k] * g.st.vd[k + 1] * 2);
        for (double &x : g.st.A[k]) x = 2.0 * ((double)rng() / 4294967296.0) - 1.0;
    }
    right_normalize(g.st);
    g.mpo = build_mpo(L, n_couplings, first, second, J, field);
    g.LE.assign(L + 1, Vec());
    g.RE.assign(L + 1, Vec());
    // boundaries: channel 0 ("start") on the left, channel 1 ("done") on the right
    g.LE[0].assign((size_t)g.mpo.W[0], 0.0);
    g.LE[0][0] = 1.0;
    g.RE[L].assign((size_t)g.mpo.W[L], 0.0);
    g.RE[L][1] = 1.0;
    for (int k = L - 1; k >= 2; --k) g.right_update(k);
    double Eprev = 1e300, E = 0.0;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        for (int k = 0; k <= L - 2; ++k) {  // left to right
            E = g.optimise(k, true);
            if (k < L - 2) g.left_update(k);
        }
        for (int k = L - 2; k >= 0; --k) {  // right to left
            E = g.optimise(k, false);
            if (k > 0) g.right_update(k + 1);
        }
        if (std::fabs(E - Eprev) < tol) break;
        Eprev = E;
    }
    // flat, then left-normalised as SetupTrial does (MPSQMC2.cpp:236)
    long off = 0;
    for (int k = 0; k < L; ++k) {
        std::memcpy(data_out + off, g.st.A[k].data(), sizeof(double) * g.st.A[k].size());
        off += (long)g.st.A[k].size();
    }
    std::memcpy(vd_out, g.st.vd.data(), sizeof(int) * (L + 1));
    double nrm = 0.0;
    const int rc = mqe_left_normalize(L, d, vd_out, data_out, &nrm);
    if (rc != MQE_OK) return rc;
    if (energy) *energy = E;
    return MQE_OK;
}

// MPSstate::ApplyMPO (MPSstate.cpp:413-471): H|psi> as an MPS whose bond k is
// the MPO channel w and the state's bond index r combined as w * vd[k] + r (the
// reference's index composition), with H = sum_c J_c S_i.S_j - field sum S^z_i
// in the minimal-channel MPO above (2 + 3 x the sites whose couplings cross the
// bond; the reference's HeisenbergMPO has more channels and gives the same
// vector).  The boundary channels are "nothing placed" on the left and
// "complete" on the right, so the outer bonds stay 1.
extern "C" int mqe_apply_mpo(int L, int d, int n_couplings, const int *first, const int *second, const double *J,
                             double field, const int *vd, const double *data, int *vd_out, double *out,
                             long *size) {
    if (L < 2 || d != 2 || n_couplings < 0 || !vd || !vd_out || (n_couplings > 0 && (!first || !second || !J))) {
        mqe::set_error("mqe_apply_mpo: bad arguments (d must be 2)");
        return MQE_EINVAL;
    }
    for (int c = 0; c < n_couplings; ++c)
        if (first[c] < 0 || second[c] < 0 || first[c] >= L || second[c] >= L || first[c] == second[c]) {
            mqe::set_error("mqe_apply_mpo: bad coupling");
            return MQE_EINVAL;
        }
    if (vd[0] != 1 || vd[L] != 1) { mqe::set_error("mqe_apply_mpo: boundary bonds must be 1"); return MQE_EINVAL; }
    const Mpo mpo = build_mpo(L, n_couplings, first, second, J, field);
    std::vector<int> vo(L + 1);
    for (int k = 0; k <= L; ++k) vo[k] = (k == 0 || k == L) ? 1 : mpo.W[k] * vd[k];
    long total = 0;
    for (int k = 0; k < L; ++k) total += (long)vo[k] * vo[k + 1] * d;
    std::memcpy(vd_out, vo.data(), sizeof(int) * (L + 1));
    if (size) *size = total;
    if (!out) return MQE_OK;
    std::memset(out, 0, sizeof(double) * total);
    long oin = 0, oout = 0;
    for (int k = 0; k < L; ++k) {
        const int dl = vd[k], dr = vd[k + 1], nl = vo[k], nr = vo[k + 1];
        const double *A = data + oin;
        double *B = out + oout;
        for (const Term &t : mpo.site[k]) {
            // the outer bonds keep only the boundary channels
            if (k == 0 && t.wl != 0) continue;
            if (k == L - 1 && t.wr != 1) continue;
            const int wl = k == 0 ? 0 : t.wl, wr = k == L - 1 ? 0 : t.wr;
            for (int sp = 0; sp < 2; ++sp)
                for (int s = 0; s < 2; ++s) {
                    const double f = t.coef * OPS[t.op][sp + 2 * s];
                    if (f == 0.0) continue;
                    for (int c = 0; c < dr; ++c)
                        for (int r = 0; r < dl; ++r)
                            B[(wl * dl + r) + (size_t)nl * ((wr * dr + c) + (size_t)nr * sp)] +=
                                f * A[r + (size_t)dl * (c + (size_t)dr * s)];
                }
        }
        oin += (long)dl * dr * d;
        oout += (long)nl * nr * d;
    }
    return MQE_OK;
}
