// mqe_host.cpp -- set-up code of the walker engine that runs on the host:
// bond-dimension rule, closed-form Trotter tables, the Marsaglia grid, and
// FP64 left-normalization / SVD compression of a single MPS (used once per
// run to prepare the trial and the initial walkers, as the reference does in
// MPSQMC2::SetupTrial / SetupWalkers).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "mqe.h"
#include "mqe_internal.h"

namespace {

// --- inverse error function ------------------------------------------------
// Bracketing bisection to ~1e-3, then Newton steps on std::erf; erf is smooth
// and strictly increasing, so the iteration converges to the correctly
// rounded neighbourhood of the root.
double inv_erf(double y) {
    if (y == 0.0) return 0.0;
    if (y >= 1.0) return HUGE_VAL;
    if (y <= -1.0) return -HUGE_VAL;
    double lo = -6.0, hi = 6.0;
    for (int it = 0; it < 14; ++it) {
        double mid = 0.5 * (lo + hi);
        if (std::erf(mid) < y) lo = mid; else hi = mid;
    }
    double x = 0.5 * (lo + hi);
    const double two_over_sqrt_pi = 1.12837916709551257390;
    for (int it = 0; it < 6; ++it) {
        double dx = (std::erf(x) - y) / (two_over_sqrt_pi * std::exp(-x * x));
        x -= dx;
        if (std::fabs(dx) < 1e-17 * std::fabs(x)) break;
    }
    return x;
}

// --- dense helpers (column-major) -----------------------------------------
// Householder QR of an m x n matrix (m >= n).  On exit a holds Q (m x n) and
// r the n x n upper triangle, with R_jj = beta_j of the reflector of column j.
void qr_householder(int m, int n, double *a, double *r) {
    std::vector<double> tau(n, 0.0);
    for (int j = 0; j < n; ++j) {
        double *cj = a + (size_t)m * j;
        double x2 = 0.0;
        for (int i = j + 1; i < m; ++i) x2 += cj[i] * cj[i];
        if (x2 == 0.0) { tau[j] = 0.0; continue; }
        double alpha = cj[j];
        double beta = -std::copysign(std::sqrt(alpha * alpha + x2), alpha);
        tau[j] = (beta - alpha) / beta;
        double inv = 1.0 / (alpha - beta);
        for (int i = j + 1; i < m; ++i) cj[i] *= inv;
        cj[j] = beta;
        for (int c = j + 1; c < n; ++c) {
            double *cc = a + (size_t)m * c;
            double s = cc[j];
            for (int i = j + 1; i < m; ++i) s += cj[i] * cc[i];
            s *= tau[j];
            cc[j] -= s;
            for (int i = j + 1; i < m; ++i) cc[i] -= s * cj[i];
        }
    }
    for (int c = 0; c < n; ++c)
        for (int rr = 0; rr < n; ++rr) r[rr + (size_t)n * c] = (rr <= c) ? a[rr + (size_t)m * c] : 0.0;
    // Q column c = H_0 ... H_c e_c; reflector vectors are read from a copy.
    std::vector<double> v((size_t)m * n);
    std::memcpy(v.data(), a, sizeof(double) * (size_t)m * n);
    for (int c = 0; c < n; ++c) {
        double *q = a + (size_t)m * c;
        for (int i = 0; i < m; ++i) q[i] = (i == c) ? 1.0 : 0.0;
        for (int k = c; k >= 0; --k) {
            if (tau[k] == 0.0) continue;
            const double *vk = v.data() + (size_t)m * k;
            double s = q[k];
            for (int i = k + 1; i < m; ++i) s += vk[i] * q[i];
            s *= tau[k];
            q[k] -= s;
            for (int i = k + 1; i < m; ++i) q[i] -= s * vk[i];
        }
    }
}

// One-sided Jacobi SVD of an m x n matrix with m >= n: a = U diag(s) V^T,
// s descending.  u: m x n, vt: n x n.
void svd_tall(int m, int n, const double *a_in, double *s, double *u, double *vt) {
    std::vector<double> a(a_in, a_in + (size_t)m * n), v((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[i + (size_t)n * i] = 1.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        bool changed = false;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double app = 0, aqq = 0, apq = 0;
                const double *cp = &a[(size_t)m * p], *cq = &a[(size_t)m * q];
                for (int i = 0; i < m; ++i) { app += cp[i] * cp[i]; aqq += cq[i] * cq[i]; apq += cp[i] * cq[i]; }
                if (apq == 0.0 || std::fabs(apq) <= 4e-16 * std::sqrt(app * aqq)) continue;
                changed = true;
                double tau = (aqq - app) / (2.0 * apq);
                double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
                double *wp = &a[(size_t)m * p], *wq = &a[(size_t)m * q];
                for (int i = 0; i < m; ++i) {
                    double x = wp[i], y = wq[i];
                    wp[i] = c * x - sn * y;
                    wq[i] = sn * x + c * y;
                }
                double *vp = &v[(size_t)n * p], *vq = &v[(size_t)n * q];
                for (int i = 0; i < n; ++i) {
                    double x = vp[i], y = vq[i];
                    vp[i] = c * x - sn * y;
                    vq[i] = sn * x + c * y;
                }
            }
        if (!changed) break;
    }
    std::vector<int> ord(n);
    std::vector<double> nrm(n);
    for (int j = 0; j < n; ++j) {
        double t = 0;
        for (int i = 0; i < m; ++i) t += a[i + (size_t)m * j] * a[i + (size_t)m * j];
        nrm[j] = std::sqrt(t);
        ord[j] = j;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nrm[x] > nrm[y]; });
    for (int jj = 0; jj < n; ++jj) {
        int j = ord[jj];
        s[jj] = nrm[j];
        double *uc = u + (size_t)m * jj;
        if (nrm[j] > 0.0) {
            for (int i = 0; i < m; ++i) uc[i] = a[i + (size_t)m * j] / nrm[j];
        } else {
            // complete with a unit vector orthogonal to the accepted columns
            for (int e = 0; e < m; ++e) {
                for (int i = 0; i < m; ++i) uc[i] = (i == e) ? 1.0 : 0.0;
                for (int k = 0; k < jj; ++k) {
                    const double *uk = u + (size_t)m * k;
                    double dt = 0;
                    for (int i = 0; i < m; ++i) dt += uk[i] * uc[i];
                    for (int i = 0; i < m; ++i) uc[i] -= dt * uk[i];
                }
                double t = 0;
                for (int i = 0; i < m; ++i) t += uc[i] * uc[i];
                if (t > 1e-12) {
                    t = 1.0 / std::sqrt(t);
                    for (int i = 0; i < m; ++i) uc[i] *= t;
                    break;
                }
            }
        }
        for (int c = 0; c < n; ++c) vt[jj + (size_t)n * c] = v[c + (size_t)n * j];
    }
}

// General SVD: k = min(m, n) values; u m x k, vt k x n.
void svd_any(int m, int n, const double *a, double *s, double *u, double *vt) {
    if (m >= n) { svd_tall(m, n, a, s, u, vt); return; }
    std::vector<double> at((size_t)n * m), ut((size_t)n * m), vtt((size_t)m * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) at[j + (size_t)n * i] = a[i + (size_t)m * j];
    svd_tall(n, m, at.data(), s, ut.data(), vtt.data());  // a^T = ut s vtt
    // a = vtt^T s ut^T
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < m; ++k) u[i + (size_t)m * k] = vtt[k + (size_t)m * i];
    for (int k = 0; k < m; ++k)
        for (int j = 0; j < n; ++j) vt[k + (size_t)m * j] = ut[j + (size_t)n * k];
}

struct HostMps {
    int L, d;
    std::vector<int> vd;
    std::vector<std::vector<double>> t;
    void load(int L_, int d_, const int *v, const double *data) {
        L = L_; d = d_;
        vd.assign(v, v + L + 1);
        t.resize(L);
        size_t off = 0;
        for (int k = 0; k < L; ++k) {
            size_t n = (size_t)vd[k] * vd[k + 1] * d;
            t[k].assign(data + off, data + off + n);
            off += n;
        }
    }
    void store(double *out) const {
        size_t off = 0;
        for (int k = 0; k < L; ++k) {
            std::memcpy(out + off, t[k].data(), sizeof(double) * t[k].size());
            off += t[k].size();
        }
    }
};

// Two-site SVD step of CompressState on bond k+1 (TwoSiteObject.cpp:71-184).
void two_site(HostMps &m, int k, int Dt, bool movingright) {
    const int d = m.d, DL = m.vd[k], DM = m.vd[k + 1], DR = m.vd[k + 2];
    const int rows = DL * d, cols = DR * d, kmin = std::min(rows, cols);
    std::vector<double> theta((size_t)rows * cols, 0.0);
    for (int il = 0; il < d; ++il)
        for (int ir = 0; ir < d; ++ir) {
            const double *A = m.t[k].data() + (size_t)il * DL * DM;
            const double *B = m.t[k + 1].data() + (size_t)ir * DM * DR;
            for (int b = 0; b < DR; ++b)
                for (int q = 0; q < DM; ++q) {
                    double bq = B[q + (size_t)DM * b];
                    if (bq == 0.0) continue;
                    for (int a = 0; a < DL; ++a)
                        theta[(a + (size_t)DL * il) + (size_t)rows * (b + (size_t)DR * ir)] += A[a + (size_t)DL * q] * bq;
                }
        }
    std::vector<double> s(kmin), u((size_t)rows * kmin), vt((size_t)kmin * cols);
    svd_any(rows, cols, theta.data(), s.data(), u.data(), vt.data());
    int keep = std::min(Dt, kmin);
    for (int c = 0; c < kmin; ++c)
        if (s[c] < 1e-13) { keep = std::min(keep, c); break; }
    m.vd[k + 1] = keep;
    m.t[k].assign((size_t)DL * keep * d, 0.0);
    m.t[k + 1].assign((size_t)keep * DR * d, 0.0);
    for (int ip = 0; ip < d; ++ip)
        for (int b = 0; b < keep; ++b)
            for (int a = 0; a < DL; ++a)
                m.t[k][a + (size_t)DL * (b + (size_t)keep * ip)] =
                    u[(a + (size_t)DL * ip) + (size_t)rows * b] * (movingright ? 1.0 : s[b]);
    for (int ip = 0; ip < d; ++ip)
        for (int b = 0; b < DR; ++b)
            for (int a = 0; a < keep; ++a)
                m.t[k + 1][a + (size_t)keep * (b + (size_t)DR * ip)] =
                    vt[a + (size_t)kmin * (b + (size_t)DR * ip)] * (movingright ? s[a] : 1.0);
}

}  // namespace

namespace mqe {
void host_svd(int m, int n, const double *a, double *s, double *u, double *vt) { svd_any(m, n, a, s, u, vt); }
void host_qr(int m, int n, double *a, double *r) { qr_householder(m, n, a, r); }
}  // namespace mqe

extern "C" {

void mqe_bond_dims(int L, int D, int d, int *vd) {
    vd[0] = 1;
    for (int k = 1; k < L; ++k) vd[k] = std::min(vd[k - 1] * d, D);
    vd[L] = 1;
    for (int k = L - 1; k > 0; --k) vd[k] = std::min(std::min(vd[k + 1] * d, D), vd[k]);
}

int mqe_trotter_tables(int d, double J, double dtau, double *S, double *U, double *VT) {
    if (d != 2 || !S || !U || !VT) {
        mqe::set_error("mqe_trotter_tables: only d = 2 is supported");
        return MQE_EINVAL;
    }
    // exp(-tau S.S) = a' I(x)I + c (X(x)X + Y(x)Y + Z(x)Z) with Pauli matrices and
    // tau = dtau*J:  a' = (3 e^{-tau/4} + e^{3tau/4})/4,  c = (e^{-tau/4} - e^{3tau/4})/4.
    // In the normalised operator basis P_k/sqrt(2), k in {I, X, iY, Z}
    // (iY real, Y(x)Y = -(iY)(x)(iY)), the operator-Schmidt values are
    // 2|a'| and 2|c| (threefold) and the right factors carry the signs.
    const double tau = dtau * J;
    const double em = std::exp(-0.25 * tau), ep = std::exp(0.75 * tau);
    const double a = 0.25 * (3.0 * em + ep);
    const double c = 0.25 * (em - ep);
    const double r2 = 1.0 / std::sqrt(2.0);
    // basis ordered smallest m first: Sz = diag(-1/2, +1/2)  ->  Z = diag(-1, 1)
    const double P[4][4] = {
        {1, 0, 0, 1},    // I   : (00, 10, 01, 11) = (0,0),(1,0),(0,1),(1,1)
        {0, 1, 1, 0},    // X
        {0, 1, -1, 0},   // iY  : (1,0) = +1, (0,1) = -1 (OpISy.cpp:11-21 orientation)
        {-1, 0, 0, 1},   // Z
    };
    const double sgn[4] = {a >= 0 ? 1.0 : -1.0, c >= 0 ? 1.0 : -1.0, c >= 0 ? -1.0 : 1.0, c >= 0 ? 1.0 : -1.0};
    double sv[4] = {2.0 * std::fabs(a), 2.0 * std::fabs(c), 2.0 * std::fabs(c), 2.0 * std::fabs(c)};
    int ord[4] = {0, 1, 2, 3};
    std::stable_sort(ord, ord + 4, [&](int x, int y) { return sv[x] > sv[y]; });
    for (int kk = 0; kk < 4; ++kk) {
        int k = ord[kk];
        S[kk] = sv[k];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                U[i + 2 * (j + 2 * kk)] = r2 * P[k][i + 2 * j];
                VT[kk + 4 * (i + 2 * j)] = sgn[k] * r2 * P[k][i + 2 * j];
            }
    }
    return MQE_OK;
}

int mqe_marsaglia_grid(int dim, int p, double *grid, int *n) {
    if (dim < 1 || p < 1 || !grid || !n) { mqe::set_error("mqe_marsaglia_grid: bad arguments"); return MQE_EINVAL; }
    std::vector<double> one(p);
    for (int c = 0; c < p; ++c) one[c] = inv_erf(2.0 * (c + 0.5) / p - 1.0);
    long total = 1;
    for (int c = 0; c < dim; ++c) total *= p;
    std::vector<double> x(dim);
    int count = 0;
    for (long idx = 0; idx < total; ++idx) {
        long rest = idx;
        double nrm = 0.0;
        for (int c = 0; c < dim; ++c) {  // coordinate 0 is the least significant digit
            int digit = (int)(rest % p);
            x[c] = one[digit];
            nrm += x[c] * x[c];
            rest /= p;
        }
        if (!(nrm > 1e-16)) continue;
        double inv = 1.0 / std::sqrt(nrm);
        for (int c = 0; c < dim; ++c) x[c] *= inv;
        bool dup = false;
        for (int g = 0; g < count && !dup; ++g) {
            double d2 = 0.0;
            for (int c = 0; c < dim; ++c) { double t = x[c] - grid[c + (size_t)dim * g]; d2 += t * t; }
            dup = d2 < 1e-16;
        }
        if (dup) continue;
        for (int c = 0; c < dim; ++c) grid[c + (size_t)dim * count] = x[c];
        ++count;
    }
    *n = count;
    return MQE_OK;
}

int mqe_left_normalize(int L, int d, const int *vd, double *data, double *norm) {
    if (L < 1 || d < 1 || !vd || !data) { mqe::set_error("mqe_left_normalize: bad arguments"); return MQE_EINVAL; }
    HostMps m;
    m.load(L, d, vd, data);
    double last = 1.0;
    for (int k = 0; k < L; ++k) {
        const int dl = m.vd[k], dr = m.vd[k + 1], rows = dl * d;
        if (rows < dr) { mqe::set_error("mqe_left_normalize: dimL*d < dimR"); return MQE_EINVAL; }
        std::vector<double> a((size_t)rows * dr), r((size_t)dr * dr);
        for (int s = 0; s < d; ++s)  // row index row + dimL*s (MPStensor.cpp:108)
            for (int c = 0; c < dr; ++c)
                for (int rr = 0; rr < dl; ++rr) a[(rr + (size_t)dl * s) + (size_t)rows * c] = m.t[k][rr + (size_t)dl * (c + (size_t)dr * s)];
        qr_householder(rows, dr, a.data(), r.data());
        for (int s = 0; s < d; ++s)
            for (int c = 0; c < dr; ++c)
                for (int rr = 0; rr < dl; ++rr) m.t[k][rr + (size_t)dl * (c + (size_t)dr * s)] = a[(rr + (size_t)dl * s) + (size_t)rows * c];
        if (k + 1 < L) {
            const int dn = m.vd[k + 2];
            std::vector<double> nt((size_t)dr * dn * d, 0.0);
            for (int s = 0; s < d; ++s)
                for (int c = 0; c < dn; ++c)
                    for (int q = 0; q < dr; ++q) {
                        double b = m.t[k + 1][q + (size_t)dr * (c + (size_t)dn * s)];
                        if (b == 0.0) continue;
                        for (int rr = 0; rr <= q; ++rr) nt[rr + (size_t)dr * (c + (size_t)dn * s)] += r[rr + (size_t)dr * q] * b;
                    }
            m.t[k + 1].swap(nt);
        } else {
            last = r[0];
        }
    }
    m.store(data);
    if (norm) *norm = last;
    return MQE_OK;
}

int mqe_right_normalize(int L, int d, const int *vd, double *data, double *norm) {
    if (L < 1 || d < 1 || !vd || !data) { mqe::set_error("mqe_right_normalize: bad arguments"); return MQE_EINVAL; }
    HostMps m;
    m.load(L, d, vd, data);
    double last = 1.0;
    for (int k = L - 1; k >= 0; --k) {
        // LQ of the site as a dimL x (d dimR) matrix = QR of its transpose,
        // rows (col + dimR*s)
        const int dl = m.vd[k], dr = m.vd[k + 1], cols = dr * d;
        if (cols < dl) { mqe::set_error("mqe_right_normalize: d*dimR < dimL"); return MQE_EINVAL; }
        std::vector<double> a((size_t)cols * dl), r((size_t)dl * dl);
        for (int s = 0; s < d; ++s)
            for (int c = 0; c < dr; ++c)
                for (int rr = 0; rr < dl; ++rr) a[(c + (size_t)dr * s) + (size_t)cols * rr] = m.t[k][rr + (size_t)dl * (c + (size_t)dr * s)];
        qr_householder(cols, dl, a.data(), r.data());
        for (int s = 0; s < d; ++s)
            for (int c = 0; c < dr; ++c)
                for (int rr = 0; rr < dl; ++rr) m.t[k][rr + (size_t)dl * (c + (size_t)dr * s)] = a[(c + (size_t)dr * s) + (size_t)cols * rr];
        if (k > 0) {
            // left neighbour times Lmat = R^T: B[s](a, q) <- sum_p B[s](a, p) R(q, p)
            const int dp = m.vd[k - 1];
            std::vector<double> nt((size_t)dp * dl * d, 0.0);
            for (int s = 0; s < d; ++s)
                for (int q = 0; q < dl; ++q)
                    for (int p = q; p < dl; ++p) {
                        const double rv = r[q + (size_t)dl * p];
                        if (rv == 0.0) continue;
                        for (int aa = 0; aa < dp; ++aa)
                            nt[aa + (size_t)dp * (q + (size_t)dl * s)] += m.t[k - 1][aa + (size_t)dp * (p + (size_t)dl * s)] * rv;
                    }
            m.t[k - 1].swap(nt);
        } else {
            last = r[0];
        }
    }
    m.store(data);
    if (norm) *norm = last;
    return MQE_OK;
}

int mqe_compress(int L, int d, const int *vd, const double *data, int Dt, int *vd_out, double *out) {
    if (L < 2 || d < 1 || Dt < 1 || !vd || !data || !vd_out || !out) {
        mqe::set_error("mqe_compress: bad arguments");
        return MQE_EINVAL;
    }
    HostMps m;
    m.load(L, d, vd, data);
    for (int k = 0; k < L - 1; ++k) two_site(m, k, Dt, true);
    for (int k = L - 2; k >= 0; --k) two_site(m, k, Dt, false);
    std::memcpy(vd_out, m.vd.data(), sizeof(int) * (L + 1));
    m.store(out);
    return MQE_OK;
}

struct mqe_rng {
    std::mt19937 gen;
};

int mqe_rng_create(unsigned int seed, mqe_rng **out) {
    if (!out) { mqe::set_error("mqe_rng_create: null out"); return MQE_EINVAL; }
    mqe_rng *r = new (std::nothrow) mqe_rng();
    if (!r) { mqe::set_error("mqe_rng_create: out of memory"); return MQE_EINVAL; }
    r->gen.seed(seed);
    *out = r;
    return MQE_OK;
}

void mqe_rng_destroy(mqe_rng *r) { delete r; }

int mqe_rng_uniforms(mqe_rng *r, long n, double *out) {
    if (!r || n < 0 || (n > 0 && !out)) { mqe::set_error("mqe_rng_uniforms: bad arguments"); return MQE_EINVAL; }
    for (long i = 0; i < n; ++i) out[i] = (double)r->gen() * (1.0 / 4294967296.0);
    return MQE_OK;
}

int mqe_rng_peek(mqe_rng *r, long n, double *out) {
    if (!r || n < 0 || (n > 0 && !out)) { mqe::set_error("mqe_rng_peek: bad arguments"); return MQE_EINVAL; }
    std::mt19937 g = r->gen;
    for (long i = 0; i < n; ++i) out[i] = (double)g() * (1.0 / 4294967296.0);
    return MQE_OK;
}

int mqe_rng_discard(mqe_rng *r, long k) {
    if (!r || k < 0) { mqe::set_error("mqe_rng_discard: bad arguments"); return MQE_EINVAL; }
    r->gen.discard((unsigned long long)k);
    return MQE_OK;
}

}  // extern "C"
