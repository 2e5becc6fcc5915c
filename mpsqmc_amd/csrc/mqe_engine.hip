// mqe_engine.hip -- MI355X walker engine for MPS projector QMC (gfx950).
//
// Replaces the per-walker CPU propagation of the reference's MPSQMC2 driver
// (MPSQMC2.cpp:518-747).  All walkers live on one GPU.  One 256-thread
// workgroup owns one walker for a whole step:
//
//   propagation sweep (left -> right), per coupling (i, j):
//     G   = T_i^T L_i phi_i                       (4 trial/walker physical blocks)
//     G  <- carried through sites i+1..j-1        (long-range couplings only)
//     V   = T_j R_{j+1} phi_j^T                   (4 blocks)
//     M   = <G, V>                                (16 numbers = all two-site
//                                                   mixed matrix elements)
//     combos, PDF over the grid, CDF sample, weight update (MPSQMC2.cpp:540-588)
//     apply the sampled operator on sites i and j (MPSstate.cpp:575-633), the
//     right one scaled by O / <trial|B(x) phi> so that the overlap stays O and
//     the norm stays near 1; the exact norm is taken once, in the energy
//     sweep (every quantity of the step is a ratio in which the walker's
//     scale cancels);
//     right environments R_k of sites a long-range coupling changed are
//     rebuilt when a later coupling needs them.
//   energy sweep (right -> left):
//     FP64 Householder LQ per site (right-canonical form for the next step,
//     right environments R_k), bond energies <trial|S_i.S_j|phi> from the
//     same two-site machinery, overlap and norm at site 0 (Walker.cpp:54-77).
//
// This environment algorithm yields the same overlaps, PDFs, weights and local
// energies as the reference's 16+1 full transfer-matrix overlaps per coupling
// (the quantities are gauge invariant), at O(L D^3) instead of O(L^2 D^3) per
// walker-step.  T = double (parity mode) or float (FP32 contractions; QR,
// sampling, weights, overlaps and energies stay FP64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "mqe.h"
#include "mqe_device.h"
#include "mqe_internal.h"
#include "mqe_svd.h"

#define MQE_MAXL 256
#define MQE_MAXG 4096

namespace mqe {

static thread_local std::string g_err;
void set_error(const char *msg) { g_err = msg ? msg : ""; }

enum : int { F_PREP = 1, F_PROP = 2 };

// Step control block in device memory.  The launches of a step read their
// step-dependent values from here instead of taking them as arguments, so the
// same launches can be captured once in a graph and replayed (a one-thread
// kernel advances the block at the start of every captured step).
struct Ctl {
    int step;         // device-RNG step word of the propagation; the energy's step number
    int pstep;        // device-RNG step word of population control
    long pc;          // propagations before this one (phase of the re-orthonormalisation)
    int canon_every;  // walker slot s re-orthonormalises when (s + pc) % canon_every == canon_every - 1
    int trace_cap;    // rows of the device trace
};

// Re-orthonormalisation is staggered over the walker slots: each propagation
// re-orthonormalises the slots of one phase, so every step costs the same and a
// walker that keeps its slot does it every canon_every steps.  A walker also
// carries the number of propagations since its last re-orthonormalisation
// (`since`, copied with it), and one that reaches canon_every - 1 does it
// regardless of its slot's phase: copies moved to new slots, and copies of
// copies, never go longer without it than a walker that stays put.
__host__ __device__ inline bool canon_walker(const Ctl &c, int slot, int since) {
    return c.canon_every <= 1 || since >= c.canon_every - 1 ||
           (int)((slot + c.pc) % c.canon_every) == c.canon_every - 1;
}

// Read-only description of the system (device pointers).
struct Sys {
    int L, nC, Ng, gdim, Dt, Dw;
    const int *tvd;    // L+1 trial bond dims
    const long *toff;  // L site offsets into the trial
    const int *first, *second, *nextfirst;
    const double *Jc;                // nC
    const double *tS, *tU, *tVT;     // per coupling: 4, 16, 16
    const double *grid;              // Ng * gdim
    const int *cs_ptr, *cs_idx;      // couplings grouped by their second site (CSR, L+1)
    double SS[16];                   // S.S two-site tensor, index (s'+2t')*4 + (s+2t)
    double Sz[2];
    double h;
    double ssp[2];                   // exp(h dtau m / 2)
    int tiles[8];                    // FP32 product tile shape by walker bond bucket (max bond 1-16, 17-32, ...)
    const float *trialT;             // FP32 general path: T_k[s]^T (t_{k+1} x t_k) at toff[k] + s t_k t_{k+1}
};

// bond bucket of a walker: its largest bond in steps of 16 (0: <= 16, ..., 7: > 112)
__host__ __device__ inline int bond_bucket(int maxbond) { return min(7, max(0, (maxbond - 1) / 16)); }

// The walker population (device pointers).
struct Pool {
    void *phi, *Renv, *Lenv;
    int *wvd;
    double *w, *O, *eh, *EL;
    const int *order;
    const int *n;
    int *idx;
    const double *u;       // host-supplied uniforms (n*nC, walker-major) or nullptr
    long wflat, eflat;
    const long *woff, *eoff;
    const long *moff;      // general path: right norm environments M_k at Renv + moff[k] (w_k x w_k)
    unsigned long long seed;
    const Ctl *ctl;        // step control block
    const int *list;       // population positions in launch order (nullptr: block b = position b)
    int *since;            // per slot: propagations since the last re-orthonormalisation
    int *status;
    char *scratch;
    long scratch_bytes;
    int padded;            // 1: tile32 storage (32x32 zero-padded blocks per site and physical index)
    // learned proposal (tile path)
    double *bondE;         // per slot: nC bond energies <trial|h_c|phi> and <trial|phi> (or nullptr)
    int nbe;               // entries of bondE per slot (nC + 1)
    int qlds;              // general path: doubles of dynamic LDS for the LQ workspace (0: none)
    const float *pred;     // predicted labels, row b*nC + c, 16 each (or nullptr)
    double mix;            // weight of the uniform part of the learned proposal
    float *lab;            // label rows of the sample buffer (or nullptr)
    float *plab;           // per slot and coupling: the labels of the last recorded propagation (16 each)
    int plab_on;           // the walker kernel records them in this launch
    const long *nsamp;     // rows already in the sample buffer
    long maxsamp;          // its capacity in rows
};

// Scratch layout of one workgroup (sizes in elements).
struct ScratchLayout {
    long Y, Y4, G, GcA, GcB, W, V, TMP, LM, t_total;  // T region
    long AQ, RQ, TAU, PARTS, DOTS, PDF, CUM, d_total;     // double region
};

__host__ __device__ inline ScratchLayout scratch_layout(int Dt, int Dw, int Ng) {
    ScratchLayout s;
    const long Dm = Dt > Dw ? Dt : Dw;
    long o = 0;
    s.Y = o; o += Dm * 2 * Dm;
    s.Y4 = o; o += 8L * Dt * Dw;  // the four carried channels times site k
    s.G = o; o += 4L * Dt * Dw;
    s.GcA = o; o += 4L * Dt * Dw;
    s.GcB = o; o += 4L * Dt * Dw;
    s.W = o; o += 2L * Dt * Dw;
    s.V = o; o += 4L * Dt * Dw;
    s.TMP = o; o += 2L * Dw * Dw;
    s.LM = o; o += (long)Dw * Dw;
    s.t_total = o;
    o = 0;
    s.AQ = o; o += 2L * Dw * Dw;
    s.RQ = o; o += (long)Dw * Dw;
    s.TAU = o; o += Dw;
    s.PARTS = o; o += MQE_NT;
    s.DOTS = o; o += (Dw > MQE_NT ? Dw : MQE_NT);
    s.PDF = o; o += Ng;
    s.CUM = o; o += Ng;
    s.d_total = o;
    return s;
}

template <typename T>
__host__ __device__ inline long scratch_bytes_for(int Dt, int Dw, int Ng) {
    ScratchLayout s = scratch_layout(Dt, Dw, Ng);
    long tb = s.t_total * (long)sizeof(T);
    tb = (tb + 255) & ~255L;
    return tb + s.d_total * (long)sizeof(double) + 256;
}

// ---------------------------------------------------------------------------
// per-walker context
// ---------------------------------------------------------------------------
// CH: the FP32 products are chained in registers (chain2); the host selects
// this instance only when every trial and walker bond is at most 64.
template <typename T, bool CH>
struct Ctx {
    const Sys *S;
    const T *trial;
    T *phi, *Lenv, *Renv;
    const long *woff, *eoff, *moff;
    const int *tv, *wv;  // LDS copies of bond dims
    ScratchLayout sl;
    T *ts;               // T scratch
    double *ds;          // double scratch
    double *red;         // LDS reduction scratch (>= MQE_NT + 16*MQE_WAVES)
    double *qlds;        // LDS workspace of the LQ (tau, then trailing columns), or nullptr
    long qlds_doubles;   // its size
    int tile;            // output tile shape of the FP32 matrix-core products (TILE_*)

    const float *trialT; // FP32: the trial with every T_k[s] transposed (t_{k+1} x t_k), or nullptr

    __device__ T *site(int k) const { return phi + woff[k]; }
    __device__ const T *tsite(int k) const { return trial + S->toff[k]; }
    __device__ const float *tsiteT(int k) const { return trialT + S->toff[k]; }

    __device__ T *Lk(int k) const { return Lenv + eoff[k]; }
    __device__ T *Rk(int k) const { return Renv + eoff[k]; }
    __device__ T *Mk(int k) const { return Renv + moff[k]; }  // right norm environment (w_k x w_k)
};

// Chained-product issue functions (CH instance only): they start their items at
// wave wave0, return the wave for the next independent call and leave the
// workgroup barrier to the caller.
//
// G(s', s) = T_i[s']^T L_i phi_i[s] (t_{i+1} x w_{i+1}) into out + s' ssp + s ss
// (ld ldo).  One intermediate L_i phi_i[s] feeds both s' when the walker bond
// spans two 32-column strips (four items either way).
template <typename T, bool CH>
__device__ __forceinline__ int issue_G(const Ctx<T, CH> &C, int i, const T *Lenv_i, T *out, int ldo, long ssp, long ss,
                                       int wave0) {
    if constexpr (CH) {
        const int ti = C.tv[i], tn = C.tv[i + 1], wi = C.wv[i], wn = C.wv[i + 1];
        const float *Ti = (const float *)C.tsite(i), *ph = (const float *)C.site(i);
        auto unit = [&](int u) {
            const int sp = u >> 1, s = u & 1;
            ChainOps o;
            o.A = (const float *)Lenv_i;
            o.B[0] = o.B[1] = ph + (size_t)s * wi * wn;
            o.C[0] = o.C[1] = Ti + (size_t)sp * ti * tn;
            o.D[0] = o.D[1] = (float *)out + (size_t)sp * ssp + (size_t)s * ss;
            return o;
        };
        if (wn < 32) return chain2_16<false, true, 1, 1>(wave0, 4, ti, wn, wi, tn, ti, wi, ti, ldo, false, unit);
        if (wn > 32)
            return chain2<false, true, 1, 2>(wave0, 2, ti, wn, wi, tn, ti, wi, ti, ldo, false, [&](int s) {
                ChainOps o;
                o.A = (const float *)Lenv_i;
                o.B[0] = o.B[1] = ph + (size_t)s * wi * wn;
                o.C[0] = Ti;
                o.C[1] = Ti + (size_t)ti * tn;
                o.D[0] = (float *)out + (size_t)s * ss;
                o.D[1] = (float *)out + (size_t)ssp + (size_t)s * ss;
                return o;
            });
        return chain2<false, true, 1, 1>(wave0, 4, ti, wn, wi, tn, ti, wi, ti, ldo, false, unit);
    }
    return wave0;
}

// V(t', t) = T_j[t'] R_{j+1} phi_j[t]^T (t_j x w_j) into the V scratch (see compute_V).
template <typename T, bool CH>
__device__ __forceinline__ int issue_V(const Ctx<T, CH> &C, int j, int wave0) {
    if constexpr (CH) {
        const int tj = C.tv[j], tn = C.tv[j + 1], wj = C.wv[j], wn = C.wv[j + 1];
        const float *R = (const float *)C.Rk(j + 1), *ph = (const float *)C.site(j), *TTj = C.tsiteT(j);
        float *V = (float *)(C.ts + C.sl.V);
        auto unit = [&](int u) {
            const int tp = u >> 1, t = u & 1;
            ChainOps o;
            o.A = R;
            o.B[0] = o.B[1] = ph + (size_t)t * wj * wn;
            o.C[0] = o.C[1] = TTj + (size_t)tp * tj * tn;
            o.D[0] = o.D[1] = V + (size_t)tp * tj * 2 * wj + (size_t)t * tj * wj;
            return o;
        };
        if (wj < 32) return chain2_16<true, true, 1, 1>(wave0, 4, tn, wj, wn, tj, tn, wj, tn, tj, false, unit);
        if (wj > 32)
            return chain2<true, true, 1, 2>(wave0, 2, tn, wj, wn, tj, tn, wj, tn, tj, false, [&](int t) {
                ChainOps o;
                o.A = R;
                o.B[0] = o.B[1] = ph + (size_t)t * wj * wn;
                o.C[0] = TTj;
                o.C[1] = TTj + (size_t)tj * tn;
                o.D[0] = V + (size_t)t * tj * wj;
                o.D[1] = V + (size_t)tj * 2 * wj + (size_t)t * tj * wj;
                return o;
            });
        return chain2<true, true, 1, 1>(wave0, 4, tn, wj, wn, tj, tn, wj, tn, tj, false, unit);
    }
    return wave0;
}

// Y = L_i phi_i (t_i x 2 w_{i+1});  G = T_i^T Y (2 t_{i+1} x 2 w_{i+1}), block
// (s', s) = T_i[s']^T L_i phi_i[s] at G + s' t + s w (2t), ld 2t.
template <typename T, bool CH>
__device__ __forceinline__ void compute_G(const Ctx<T, CH> &C, int i, const T *Lenv_i) {
    const int ti = C.tv[i], tn = C.tv[i + 1], wi = C.wv[i], wn = C.wv[i + 1];
    T *Y = C.ts + C.sl.Y, *G = C.ts + C.sl.G;
    if constexpr (CH) {
        issue_G(C, i, Lenv_i, G, 2 * tn, tn, (long)wn * 2 * tn, 0);
        __syncthreads();
        return;
    }
    gemm<false, false, T>(ti, 2 * wn, wi, Lenv_i, ti, C.site(i), wi, Y, ti, false, C.tile);
    __syncthreads();
    gemm<true, false, T>(2 * tn, 2 * wn, ti, C.tsite(i), ti, Y, ti, G, 2 * tn, false, C.tile);
    __syncthreads();
}

// The four blocks G(s', s) of bond i+1 as contiguous channels ch = 2 s' + s
// (t_{i+1} x w_{i+1}, ld t_{i+1}) at A.  CH: no barrier at the end (the
// caller's), the return value is the wave for the next independent call.
template <typename T, bool CH>
__device__ __forceinline__ int compute_G_channels(const Ctx<T, CH> &C, int i, T *A, int wave0 = 0) {
    const int tn = C.tv[i + 1], wn = C.wv[i + 1];
    if constexpr (CH) return issue_G(C, i, C.Lk(i), A, tn, 2L * tn * wn, (long)tn * wn, wave0);
    compute_G(C, i, C.Lk(i));
    const T *G = C.ts + C.sl.G;
    for (int e = threadIdx.x; e < 4 * tn * wn; e += MQE_NT) {
        const int ch = e / (tn * wn), r = e % (tn * wn), a = r % tn, b = r / tn;
        const int sp = ch >> 1, s = ch & 1;
        A[e] = G[(size_t)sp * tn + (size_t)s * wn * 2 * tn + a + (size_t)2 * tn * b];
    }
    __syncthreads();
    return wave0;
}

// W[t] = R_{j+1} phi_j[t]^T (t_{j+1} x w_j);  V[t'] = T_j[t'] W (t_j x 2 w_j);
// block (t', t) at V + t' (t_j 2 w_j) + t (t_j w_j), ld t_j.
template <typename T, bool CH>
__device__ __forceinline__ void compute_V(const Ctx<T, CH> &C, int j) {
    if constexpr (CH) {
        issue_V(C, j, 0);
        __syncthreads();
        return;
    }
    const int tj = C.tv[j], tn = C.tv[j + 1], wj = C.wv[j], wn = C.wv[j + 1];
    T *W = C.ts + C.sl.W, *V = C.ts + C.sl.V;
    const T *R = C.Rk(j + 1), *ph = C.site(j);
    gemm_batch<false, true, T>(2, tn, wj, wn, R, tn, 0, ph, wj, (long)wj * wn, W, tn, (long)tn * wj, false, C.tile);
    __syncthreads();
    const T *Tj = C.tsite(j);
    gemm_batch<false, false, T>(2, tj, 2 * wj, tn, Tj, tj, (long)tj * tn, W, tn, 0, V, tj, (long)tj * 2 * wj, false,
                                C.tile);
    __syncthreads();
}

// Carry four channels (t_k x w_k, ld t_k, contiguous at src) through site k
// into dst (t_{k+1} x w_{k+1}): dst_ch = sum_s T_k[s]^T src_ch phi_k[s].
template <typename T, bool CH>
// nch: the number of channels (4, or 3 for the operator channels of a bond energy).
__device__ __forceinline__ void carry_channels(const Ctx<T, CH> &C, int k, const T *src, T *dst, int nch = 4) {
    const int tk = C.tv[k], tn = C.tv[k + 1], wk = C.wv[k], wn = C.wv[k + 1];
    const T *Tk = C.tsite(k);
    if constexpr (sizeof(T) == sizeof(float)) {
        if constexpr (CH) {
            const float *ph = (const float *)C.site(k);
            auto unit = [&](int ch) {
                ChainOps o;
                o.A = (const float *)src + (size_t)ch * tk * wk;
                o.B[0] = ph;
                o.B[1] = ph + (size_t)wk * wn;
                o.C[0] = (const float *)Tk;
                o.C[1] = (const float *)Tk + (size_t)tk * tn;
                o.D[0] = o.D[1] = (float *)dst + (size_t)ch * tn * wn;
                return o;
            };
            if (wn < 32)
                chain2_16<false, true, 2, 1>(0, nch, tk, wn, wk, tn, tk, wk, tk, tn, false, unit);
            else
                chain2<false, true, 2, 1>(0, nch, tk, wn, wk, tn, tk, wk, tk, tn, false, unit);
            __syncthreads();
            return;
        }
    }
    T *Y = C.ts + C.sl.Y4;  // Y_ch = src_ch phi_k (tk x 2 wn), the four channels one after the other
    const long sy = (long)tk * 2 * wn;
    gemm_batch<false, false, T>(nch, tk, 2 * wn, wk, src, tk, (long)tk * wk, C.site(k), wk, 0, Y, tk, sy, false,
                                C.tile);
    __syncthreads();
    // dst_ch = T_k[0]^T Y_ch[0] + T_k[1]^T Y_ch[1], one accumulation chain per output tile
    gemm2<true, false, T>(tn, wn, tk, Tk, Tk + (size_t)tk * tn, tk, Y, Y + (size_t)tk * wn, tk, dst, tn, C.tile, nch, 0,
                          sy, (long)tn * wn);
    __syncthreads();
}

// Two-site mixed tensor M[(s'+2t')*4 + (s+2t)] = <G(s',s), V(t',t)> at a bond
// with trial dim tb and walker dim wb.  g[ch], ch = 2 s' + s, with leading dim gld.
template <typename T, bool CH>
// Rout (or nullptr): also writes the right environment V(0,0) + V(1,1) there.
__device__ __forceinline__ void pair_tensor(const Ctx<T, CH> &C, int tb, int wb, const T *const g[4], int gld, double *M,
                                            T *Rout = nullptr) {
    const T *V = C.ts + C.sl.V;
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) acc[q] = 0.0;
    const int nel = tb * wb;
    // two elements of a thread per pass, the loads of both before the store to
    // Rout (which the compiler cannot separate from them); the second index is
    // clamped into the block, its element skipped below (same order of sums)
    for (int e0 = threadIdx.x; e0 < nel; e0 += 2 * MQE_NT) {
        T gt[2][4], vt[2][4];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = min(e0 + u * MQE_NT, nel - 1);
            const int a = e % tb, b = e / tb;
#pragma unroll
            for (int ch = 0; ch < 4; ++ch) {
                const int tp = ch >> 1, t = ch & 1;
                gt[u][ch] = g[ch][a + (size_t)gld * b];
                vt[u][ch] = V[(size_t)tp * tb * 2 * wb + (size_t)t * tb * wb + a + (size_t)tb * b];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + u * MQE_NT;
            if (e < nel) {
                if (Rout) Rout[e] = vt[u][0] + vt[u][3];
#pragma unroll
                for (int sp = 0; sp < 2; ++sp)
#pragma unroll
                    for (int s = 0; s < 2; ++s)
#pragma unroll
                        for (int tp = 0; tp < 2; ++tp)
#pragma unroll
                            for (int t = 0; t < 2; ++t)
                                acc[(sp + 2 * tp) * 4 + (s + 2 * t)] +=
                                    (double)gt[u][2 * sp + s] * (double)vt[u][2 * tp + t];
            }
        }
    }
    block_sum_n<16>(acc, C.red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) M[q] = acc[q];
    }
    __syncthreads();
}

// Bond energy <trial|S_i . S_j|phi> from the three operator channels of site i
// carried to bond j (c[0] = S^z, c[1] = S^+, c[2] = S^- applied at site i; t_j x
// w_j, ld t_j) and the blocks V(t', t) of site j:
// S_i . S_j = S^z S^z + (S^+ S^- + S^- S^+) / 2, with S^z = diag(-1/2, 1/2),
// S^+ = |1><0|, S^- = |0><1| in the basis ordered by m (OpSz.cpp, OpSx.cpp,
// OpISy.cpp).  Rout (or nullptr): also writes R_j = V(0,0) + V(1,1) there.
template <typename T, bool CH>
__device__ __forceinline__ double pair_energy3(const Ctx<T, CH> &C, int tb, int wb, const T *c, T *Rout = nullptr) {
    const T *V = C.ts + C.sl.V;
    const size_t bs = (size_t)tb * wb;
    const T *V00 = V, *V01 = V + bs, *V10 = V + 2 * bs, *V11 = V + 3 * bs;  // V(t', t) at t' 2 bs + t bs
    double acc = 0.0;
    const int n = (int)bs;
    // two elements of a thread per pass, loads before the store to Rout (as in
    // pair_tensor; the clamped second index is skipped below)
    for (int e0 = threadIdx.x; e0 < n; e0 += 2 * MQE_NT) {
        T v00[2], v11[2], v01[2], v10[2], cz[2], cp[2], cm[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = min(e0 + u * MQE_NT, n - 1);
            v00[u] = V00[e];
            v11[u] = V11[e];
            v01[u] = V01[e];
            v10[u] = V10[e];
            cz[u] = c[e];
            cp[u] = c[bs + e];
            cm[u] = c[2 * bs + e];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = e0 + u * MQE_NT;
            if (e < n) {
                acc += (double)cz[u] * (0.5 * ((double)v11[u] - (double)v00[u])) +
                       0.5 * ((double)cp[u] * (double)v01[u] + (double)cm[u] * (double)v10[u]);
                if (Rout) Rout[e] = v00[u] + v11[u];
            }
        }
    }
    return block_sum(acc, C.red);
}

// Mixed tensor for the pair (i, j) from L_i, sites i..j and R_{j+1}.
template <typename T, bool CH>
__device__ __forceinline__ void pair_for_coupling(const Ctx<T, CH> &C, int i, int j, double *M) {
    const int tn = C.tv[i + 1], wn = C.wv[i + 1];
    const T *g[4];
    int gld;
    if (j == i + 1) {
        T *G = C.ts + C.sl.G;
        if constexpr (CH) {  // G and V in one barrier interval
            issue_V(C, j, issue_G(C, i, C.Lk(i), G, 2 * tn, tn, (long)wn * 2 * tn, 0));
            __syncthreads();
        } else {
            compute_G(C, i, C.Lk(i));
            compute_V(C, j);
        }
        for (int ch = 0; ch < 4; ++ch) {
            const int sp = ch >> 1, s = ch & 1;
            g[ch] = G + (size_t)sp * tn + (size_t)s * wn * 2 * tn;
        }
        gld = 2 * tn;
    } else {
        // the four blocks as contiguous channels (ch = 2 s' + s), carried to bond j
        T *A = C.ts + C.sl.GcA, *B = C.ts + C.sl.GcB;
        if constexpr (CH) {  // G channels and V in one barrier interval
            issue_V(C, j, compute_G_channels(C, i, A));
            __syncthreads();
        } else {
            compute_G_channels(C, i, A);
        }
        for (int k = i + 1; k < j; ++k) {
            carry_channels(C, k, A, B);
            T *t = A; A = B; B = t;
        }
        const int tb = C.tv[j], wb = C.wv[j];
        for (int ch = 0; ch < 4; ++ch) g[ch] = A + (size_t)ch * tb * wb;
        gld = tb;
        if constexpr (!CH) compute_V(C, j);
    }
    pair_tensor(C, C.tv[j], C.wv[j], g, gld, M);
}

// L_{k+1} = sum_s T_k[s]^T L_k phi_k[s]
template <typename T, bool CH>
__device__ __forceinline__ int issue_L(const Ctx<T, CH> &C, int k, int wave0) {
    if constexpr (CH) {
        const int tk = C.tv[k], tn = C.tv[k + 1], wk = C.wv[k], wn = C.wv[k + 1];
        const float *ph = (const float *)C.site(k), *Tk = (const float *)C.tsite(k);
        auto unit = [&](int) {
            ChainOps o;
            o.A = (const float *)C.Lk(k);
            o.B[0] = ph;
            o.B[1] = ph + (size_t)wk * wn;
            o.C[0] = Tk;
            o.C[1] = Tk + (size_t)tk * tn;
            o.D[0] = o.D[1] = (float *)C.Lk(k + 1);
            return o;
        };
        if (wn < 32) return chain2_16<false, true, 2, 1>(wave0, 1, tk, wn, wk, tn, tk, wk, tk, tn, true, unit);
        return chain2<false, true, 2, 1>(wave0, 1, tk, wn, wk, tn, tk, wk, tk, tn, true, unit);
    }
    return wave0;
}

template <typename T, bool CH>
__device__ __forceinline__ void advance_L(const Ctx<T, CH> &C, int k) {
    if constexpr (CH) {
        issue_L(C, k, 0);
        __syncthreads();
        return;
    }
    const int tk = C.tv[k], tn = C.tv[k + 1], wk = C.wv[k], wn = C.wv[k + 1];
    T *Y = C.ts + C.sl.Y;
    gemm<false, false, T>(tk, 2 * wn, wk, C.Lk(k), tk, C.site(k), wk, Y, tk, false, C.tile);
    __syncthreads();
    const T *Tk = C.tsite(k);
    T *Ln = C.Lk(k + 1);
    gemm2<true, false, T>(tn, wn, tk, Tk, Tk + (size_t)tk * tn, tk, Y, Y + (size_t)tk * wn, tk, Ln, tn, C.tile);
    __syncthreads();
}

// R_k = sum_t T_k[t] R_{k+1} phi_k[t]^T (the diagonal blocks of V at bond k)
template <typename T, bool CH>
__device__ __forceinline__ int issue_R(const Ctx<T, CH> &C, int k, int wave0) {
    if constexpr (CH) {
        const int tk = C.tv[k], tn = C.tv[k + 1], wk = C.wv[k], wn = C.wv[k + 1];
        const float *ph = (const float *)C.site(k), *TTk = C.tsiteT(k);
        auto unit = [&](int) {
            ChainOps o;
            o.A = (const float *)C.Rk(k + 1);
            o.B[0] = ph;
            o.B[1] = ph + (size_t)wk * wn;
            o.C[0] = TTk;
            o.C[1] = TTk + (size_t)tk * tn;
            o.D[0] = o.D[1] = (float *)C.Rk(k);
            return o;
        };
        if (wk < 32) return chain2_16<true, true, 2, 1>(wave0, 1, tn, wk, wn, tk, tn, wk, tn, tk, true, unit);
        return chain2<true, true, 2, 1>(wave0, 1, tn, wk, wn, tk, tn, wk, tn, tk, true, unit);
    }
    return wave0;
}

template <typename T, bool CH>
__device__ __forceinline__ void rebuild_R(const Ctx<T, CH> &C, int k) {
    if constexpr (CH) {
        issue_R(C, k, 0);
        __syncthreads();
        return;
    }
    compute_V(C, k);
    const int tk = C.tv[k], wk = C.wv[k];
    const T *V = C.ts + C.sl.V;
    T *Rk = C.Rk(k);
    for (int e = threadIdx.x; e < tk * wk; e += MQE_NT) Rk[e] = V[e] + V[(size_t)tk * 2 * wk + (size_t)tk * wk + e];
    __syncthreads();
}

// Right norm environment M_k = sum_s phi_k[s] M_{k+1} phi_k[s]^T (w_k x w_k):
// the norm of the walker's part right of bond k in the current gauge.
// (CH issue function: M_{k+1} is symmetric, M_k = sum_s phi_k[s] (M_{k+1} phi_k[s]^T))
template <typename T, bool CH>
__device__ __forceinline__ int issue_M(const Ctx<T, CH> &C, int k, int wave0) {
    if constexpr (CH) {
        const int wk = C.wv[k], wn = C.wv[k + 1];
        const float *ph = (const float *)C.site(k), *Mn = (const float *)C.Mk(k + 1);
        float *Mk = (float *)C.Mk(k);
        auto unit = [&](int) {
            ChainOps o;
            o.A = Mn;
            o.B[0] = o.C[0] = ph;
            o.B[1] = o.C[1] = ph + (size_t)wk * wn;
            o.D[0] = o.D[1] = Mk;
            return o;
        };
        if (wk < 32) return chain2_16<true, false, 2, 1>(wave0, 1, wn, wk, wn, wk, wn, wk, wk, wk, false, unit);
        return chain2<true, false, 2, 1>(wave0, 1, wn, wk, wn, wk, wn, wk, wk, wk, false, unit);
    }
    return wave0;
}

template <typename T, bool CH>
__device__ __forceinline__ void advance_M(const Ctx<T, CH> &C, int k) {
    const int wk = C.wv[k], wn = C.wv[k + 1];
    T *X = C.ts + C.sl.Y;
    const T *ph = C.site(k), *Mn = C.Mk(k + 1);
    T *Mk = C.Mk(k);
    if constexpr (CH) {
        issue_M(C, k, 0);
        __syncthreads();
        return;
    }
    gemm_batch<false, false, T>(2, wk, wn, wn, ph, wk, (long)wk * wn, Mn, wn, 0, X, wk, (long)wk * wn, false, C.tile);
    __syncthreads();
    gemm2<false, true, T>(wk, wk, wn, X, X + (size_t)wk * wn, wk, ph, ph + (size_t)wk * wn, wk, Mk, wk, C.tile);
    __syncthreads();
}

// Squared norm with the left norm environment N (bond k) and the right norm
// environment M_{k+1}: sum_s tr(phi_k[s]^T N phi_k[s] M_{k+1}).
template <typename T, bool CH>
__device__ __forceinline__ double norm2_at_M(const Ctx<T, CH> &C, int k, const T *N) {
    const int wk = C.wv[k], wn = C.wv[k + 1];
    T *Z = C.ts + C.sl.Y, *Z2 = C.ts + C.sl.TMP;
    const T *ph = C.site(k), *Mn = C.Mk(k + 1);
    gemm<false, false, T>(wk, 2 * wn, wk, N, wk, ph, wk, Z, wk, false, C.tile);
    __syncthreads();
    gemm_batch<false, false, T>(2, wk, wn, wn, Z, wk, (long)wk * wn, Mn, wn, 0, Z2, wk, (long)wk * wn, false, C.tile);
    __syncthreads();
    double acc = 0.0;
    for (int e = threadIdx.x; e < wk * 2 * wn; e += MQE_NT) acc += (double)ph[e] * (double)Z2[e];
    return block_sum(acc, C.red);
}

template <typename T, bool CH>
__device__ __forceinline__ void scale_site(const Ctx<T, CH> &C, int k, double f) {
    T *ph = C.site(k);
    const int n = C.wv[k] * C.wv[k + 1] * 2;
    for (int e = threadIdx.x; e < n; e += MQE_NT) ph[e] = (T)((double)ph[e] * f);
    __syncthreads();
}

// (x[e], x[e + n]) <- (c00 x[e] + c01 x[e + n], c10 x[e] + c11 x[e + n]) for
// e < n: a 2 x 2 operator on the physical index of a site stored as two blocks
// of n.  Four elements of a thread are in flight at a time, their loads issued
// before their stores (one element at a time, every load waited for the store
// before it: the compiler cannot tell that they do not overlap).
template <typename T>
__device__ __forceinline__ void mix_pair(T *x, int n, T c00, T c01, T c10, T c11) {
    for (int e0 = threadIdx.x; e0 < n; e0 += 4 * MQE_NT) {
        T a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = e0 + q * MQE_NT;
            a[q] = e < n ? x[e] : T(0);
            b[q] = e < n ? x[e + n] : T(0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = e0 + q * MQE_NT;
            if (e < n) {
                x[e] = c00 * a[q] + c01 * b[q];
                x[e + n] = c10 * a[q] + c11 * b[q];
            }
        }
    }
}

// phi_k[s] <- sum_s0 O[s + 2 s0] phi_k[s0]
template <typename T, bool CH>
__device__ __forceinline__ void apply_op(const Ctx<T, CH> &C, int k, const double *O) {
    const T o00 = (T)O[0], o10 = (T)O[1], o01 = (T)O[2], o11 = (T)O[3];
    mix_pair(C.site(k), C.wv[k] * C.wv[k + 1], o00, o01, o10, o11);
    __syncthreads();
}

// LQ of site k (w_k x 2 w_{k+1}) in FP64: phi_k <- Q (right-orthonormal),
// phi_{k-1} <- phi_{k-1} Lmat, then R_k = sum_t T_k[t] R_{k+1} Q[t]^T.
// Leaves V (all four blocks) computed for bond k.
template <typename T, bool CH>
__device__ __forceinline__ void lq_site(const Ctx<T, CH> &C, int k) {
    const int wk = C.wv[k], wn = C.wv[k + 1], wp = C.wv[k - 1];
    const int m = 2 * wn, n = wk;
    // the FP64 factorisation's sequential column steps are latency-bound: the
    // trailing columns (which receive most of the updates) live in LDS, as
    // many as fit, the rest in the global scratch
    double *R = C.ds + C.sl.RQ;
    double *tau = C.qlds ? C.qlds : C.ds + C.sl.TAU;
    int nl = 0;
    if (C.qlds) nl = min(n, (int)((C.qlds_doubles - C.S->Dw) / m));
    const QRCols col{C.ds + C.sl.AQ, C.qlds ? C.qlds + C.S->Dw : nullptr, m, n - nl};
    T *ph = C.site(k);
    for (int e = threadIdx.x; e < m * n; e += MQE_NT) {
        const int r = e % m, c = e / m;
        col(c)[r] = (double)ph[c + (size_t)wk * r];
    }
    __syncthreads();
    householder_qr(col, n, R, tau);
    for (int e = threadIdx.x; e < m * n; e += MQE_NT) {
        const int r = e % m, c = e / m;
        ph[c + (size_t)wk * r] = (T)col(c)[r];
    }
    T *LM = C.ts + C.sl.LM;
    for (int e = threadIdx.x; e < n * n; e += MQE_NT) {
        const int a = e % n, b = e / n;
        LM[e] = (T)R[b + (size_t)n * a];
    }
    __syncthreads();
    // push Lmat into the left neighbour
    T *pl = C.site(k - 1), *tmp = C.ts + C.sl.TMP;
    gemm_batch<false, false, T>(2, wp, wk, wk, pl, wp, (long)wp * wk, LM, wk, 0, tmp, wp, (long)wp * wk, false, C.tile);
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * wp * wk; e += MQE_NT) pl[e] = tmp[e];
    __syncthreads();
    // right environment of bond k
    compute_V(C, k);
    const int tk = C.tv[k];
    const T *V = C.ts + C.sl.V;
    T *Rk = C.Rk(k);
    for (int e = threadIdx.x; e < tk * wk; e += MQE_NT) Rk[e] = V[e] + V[(size_t)tk * 2 * wk + (size_t)tk * wk + e];
    __syncthreads();
}

// <trial|phi> and sum_s sz[s] <trial|S^z_0|phi> at site 0 (L_0 = 1, right part via R_1).
template <typename T, bool CH>
__device__ __forceinline__ void site0_overlap(const Ctx<T, CH> &C, double *ov, double *szv) {
    const int t1 = C.tv[1], w1 = C.wv[1];
    const T *T0 = C.tsite(0), *p0 = C.site(0), *R1 = C.Rk(1);
    double a[2] = {0.0, 0.0};
    for (int e = threadIdx.x; e < t1 * w1; e += MQE_NT) {
        const int al = e % t1, be = e / t1;
        const double r = (double)R1[al + (size_t)t1 * be];
        for (int s = 0; s < 2; ++s) a[s] += (double)T0[al + (size_t)t1 * s] * (double)p0[be + (size_t)w1 * s] * r;
    }
    block_sum_n<2>(a, C.red);
    *ov = a[0] + a[1];
    *szv = C.S->Sz[0] * a[0] + C.S->Sz[1] * a[1];
}

template <typename T>
__device__ __forceinline__ void set_one(T *p) {
    if (threadIdx.x == 0) p[0] = (T)1;
    __syncthreads();
}

// Right-canonicalise the whole walker with R environments, normalise at site 0.
// Returns the (sign-fixed) overlap of the normalised state.
template <typename T, bool CH>
__device__ __forceinline__ double right_canonicalize(const Ctx<T, CH> &C) {
    const int L = C.S->L;
    set_one(C.Rk(L));
    for (int k = L - 1; k >= 1; --k) lq_site(C, k);
    double ov, szv;
    site0_overlap(C, &ov, &szv);
    double nrm2 = 0.0;
    {
        const T *p0 = C.site(0);
        double s = 0.0;
        for (int e = threadIdx.x; e < 2 * C.wv[1]; e += MQE_NT) s += (double)p0[e] * (double)p0[e];
        nrm2 = block_sum(s, C.red);
    }
    const double nrm = sqrt(nrm2);
    double O = ov / nrm;
    double f = 1.0 / nrm;
    if (O < 0.0) { O = -O; f = -f; }
    scale_site(C, 0, f);
    return O;
}

// Energy sweep from the right end (L_k valid for all k), see file header.
// Returns e_tot = <trial|H|phi> of the state as it was on entry and writes the
// overlap of the normalised state to *O_out, <trial|phi> before normalisation to *Ou.
template <typename T, bool CH>
__device__ __forceinline__ double energy_sweep(const Ctx<T, CH> &C, double *O_out, double *Ou, bool canon) {
    const Sys &S = *C.S;
    const int L = S.L;
    __shared__ double M[16];
    set_one(C.Rk(L));
    set_one(C.Mk(L));
    double etot = 0.0;
    for (int k = L - 1; k >= 1; --k) {
        bool nn_ready = false;   // G of the coupling (k-1, k) already formed
        bool r_pending = false;  // R_k still to be written (by the next pair tensor)
        if (canon) {
            lq_site(C, k);  // V of bond k now holds T_k R_{k+1} Q_k^T blocks
        } else {
            // light sweep: no re-orthonormalisation; site k is rescaled to the
            // Frobenius norm sqrt(w_k) of a right-canonical site when it is more
            // than a factor 4 away from it, and the factor moved into site k-1
            // (the state is unchanged).  This keeps every
            // tensor, and the environments built from them, at the magnitudes of
            // the canonical form: a unit-norm target would shrink R_k by about
            // sqrt(w) per site and underflow FP32 on long chains at large D.
            // Then R_k and the right norm environment M_k.
            {
                const int wk = C.wv[k], wn = C.wv[k + 1];
                const T *ph = C.site(k);
                double s2 = 0.0;
                for (int e = threadIdx.x; e < 2 * wk * wn; e += MQE_NT) s2 += (double)ph[e] * (double)ph[e];
                const double f = sqrt(block_sum(s2, C.red) / (double)wk);
                // only a site that has drifted by more than a factor 4 is rescaled:
                // that bounds every magnitude just as well and skips the two
                // passes over the tensors at most sites of most steps
                if (f > 0.0 && isfinite(f) && (f > 4.0 || f < 0.25)) {
                    scale_site(C, k, 1.0 / f);
                    scale_site(C, k - 1, f);
                }
            }
            if constexpr (CH) {
                // V, M_k and the nearest-neighbour G of bond k (the last coupling
                // ending at k, if it is (k-1, k)) in one barrier interval
                const int wq = issue_M(C, k, issue_V(C, k, 0));
                const int ql = S.cs_ptr[k + 1] - 1;
                if (ql >= S.cs_ptr[k] && S.first[S.cs_idx[ql]] == k - 1) {
                    const int tn = C.tv[k], wn = C.wv[k];
                    issue_G(C, k - 1, C.Lk(k - 1), C.ts + C.sl.G, 2 * tn, tn, (long)wn * 2 * tn, wq);
                    nn_ready = true;
                }
                __syncthreads();
            } else {
                compute_V(C, k);
            }
            if (S.cs_ptr[k + 1] > S.cs_ptr[k]) {
                r_pending = true;  // R_k from the first pair tensor of the site
            } else {
                const int tk = C.tv[k], wk = C.wv[k];
                const T *V = C.ts + C.sl.V;
                T *Rk = C.Rk(k);
                for (int e = threadIdx.x; e < tk * wk; e += MQE_NT)
                    Rk[e] = V[e] + V[(size_t)tk * 2 * wk + (size_t)tk * wk + e];
            }
            if constexpr (!CH) advance_M(C, k);
        }
        for (int q = S.cs_ptr[k]; q < S.cs_ptr[k + 1]; ++q) {
            const int c = S.cs_idx[q], i = S.first[c];
            const int tn = C.tv[i + 1], wn = C.wv[i + 1];
            const T *g[4];
            int gld;
            if (i == k - 1) {
                if (!nn_ready) compute_G(C, i, C.Lk(i));
                const T *G = C.ts + C.sl.G;
                for (int ch = 0; ch < 4; ++ch) {
                    const int sp = ch >> 1, s = ch & 1;
                    g[ch] = G + (size_t)sp * tn + (size_t)s * wn * 2 * tn;
                }
                gld = 2 * tn;
            } else {
                // long range: only S_i . S_k is needed, so the four channels G(s', s)
                // are combined into the three operator channels S^z, S^+, S^- of
                // site i (ch = 2 s' + s: S^z = (G(1,1) - G(0,0)) / 2, S^+ = G(1,0),
                // S^- = G(0,1)) and three are carried instead of four
                T *A = C.ts + C.sl.GcA, *B = C.ts + C.sl.GcB;
                compute_G_channels(C, i, A);
                if constexpr (CH) __syncthreads();
                {
                    const int nb = tn * wn;
                    for (int e = threadIdx.x; e < nb; e += MQE_NT) {
                        const T g00 = A[e], g01 = A[nb + e], g10 = A[2 * nb + e], g11 = A[3 * nb + e];
                        A[e] = (T)0.5 * (g11 - g00);
                        A[nb + e] = g10;
                        A[2 * nb + e] = g01;
                    }
                    __syncthreads();
                }
                for (int kk = i + 1; kk < k; ++kk) {
                    carry_channels(C, kk, A, B, 3);
                    T *t = A; A = B; B = t;
                }
                const double e = pair_energy3(C, C.tv[k], C.wv[k], A, r_pending ? C.Rk(k) : nullptr);
                r_pending = false;
                etot += S.Jc[c] * e;
                continue;
            }
            pair_tensor(C, C.tv[k], C.wv[k], g, gld, M, r_pending ? C.Rk(k) : nullptr);
            r_pending = false;
            double e = 0.0;
            for (int q2 = 0; q2 < 16; ++q2) e += S.SS[q2] * M[q2];
            etot += S.Jc[c] * e;
        }
        if (S.h != 0.0) {
            // <trial|S^z_k|phi> = sum_t sz[t] <L'_k, V(t,t)>, L'_k the current-gauge left env
            compute_G(C, k - 1, C.Lk(k - 1));
            const int tk = C.tv[k], wk = C.wv[k];
            const T *G = C.ts + C.sl.G, *V = C.ts + C.sl.V;
            double a[2] = {0.0, 0.0};
            for (int e = threadIdx.x; e < tk * wk; e += MQE_NT) {
                const int al = e % tk, be = e / tk;
                const double lp = (double)G[al + (size_t)2 * tk * be] +
                                  (double)G[tk + (size_t)wk * 2 * tk + al + (size_t)2 * tk * be];
                a[0] += lp * (double)V[al + (size_t)tk * be];
                a[1] += lp * (double)V[(size_t)tk * 2 * wk + (size_t)tk * wk + al + (size_t)tk * be];
            }
            block_sum_n<2>(a, C.red);
            etot -= S.h * (S.Sz[0] * a[0] + S.Sz[1] * a[1]);
        }
    }
    double ov, szv;
    site0_overlap(C, &ov, &szv);
    if (S.h != 0.0) etot -= S.h * szv;
    double nrm2;
    if (canon) {
        const T *p0 = C.site(0);
        double s = 0.0;
        for (int e = threadIdx.x; e < 2 * C.wv[1]; e += MQE_NT) s += (double)p0[e] * (double)p0[e];
        nrm2 = block_sum(s, C.red);
    } else {
        nrm2 = norm2_at_M(C, 0, C.Lk(0));  // L_0 = 1 is also the left norm environment of site 0
    }
    const double nrm = sqrt(nrm2);
    double O = ov / nrm, f = 1.0 / nrm;
    if (O < 0.0) { O = -O; f = -f; }
    scale_site(C, 0, f);
    *O_out = O;
    *Ou = ov;
    return etot;
}

template <typename T, bool CH>
__device__ __forceinline__ void apply_field_everywhere(const Ctx<T, CH> &C) {
    const double O[4] = {C.S->ssp[0], 0.0, 0.0, C.S->ssp[1]};
    for (int k = 0; k < C.S->L; ++k) apply_op(C, k, O);
}

// combos, PDF over the grid, CDF sample and the sampled operators of one
// coupling (MPSQMC2.cpp:540-588, MPSstate.cpp:589-631) for at most 256 grid
// points: every wave computes the same result (lane q < 16 forms combo q, lane
// l owns grid points 4l .. 4l+3; wave-wide sums and scan), so no workgroup
// barrier is needed.  M: the 16 mixed matrix elements (LDS).
__device__ __forceinline__ void wave_sample(const double *M, const double *tS, const double *tU, const double *tVT,
                                            const double *grid, int Ng, double O, double u, double &pdfnorm,
                                            int &gs, double &fstar, double (&Ux)[4], double (&Vx)[4]) {
    const int lane = threadIdx.x & 63;
    double cq = 0.0;
    if (lane < 16) {
        // combos[a + 4b] = sqrt(s_a s_b) <trial| U_a (x) V_b |phi>
        const int a = lane & 3, bb = lane >> 2;
        double sacc = 0.0;
        for (int sp = 0; sp < 2; ++sp)
            for (int ss = 0; ss < 2; ++ss) {
                const double ua = tU[sp + 2 * (ss + 2 * a)];
                for (int tp = 0; tp < 2; ++tp)
                    for (int t = 0; t < 2; ++t) sacc += ua * tVT[bb + 4 * (tp + 2 * t)] * M[(sp + 2 * tp) * 4 + (ss + 2 * t)];
            }
        cq = sqrt(tS[a] * tS[bb]) * sacc;
    }
    double cb[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) cb[q] = readlane_d(cq, q);
    auto quad = [&](const double *x) {  // sum_{l,r} comb[l + 4r] x_l x_r
        double f = 0.0;
#pragma unroll
        for (int l = 0; l < 4; ++l) {
            double t = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) t += cb[l + 4 * r] * x[r];
            f += x[l] * t;
        }
        return f;
    };
    const double pref = 4.0 / (O * Ng);
    double pv[4], loc = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int g = 4 * lane + t;
        double v = 0.0;
        if (g < Ng) {
            const double f = quad(grid + 4 * g);
            v = (f > 0.0 ? f : 0.0) * pref;
        }
        pv[t] = v;
        loc += v;
    }
    pdfnorm = wave_sum(loc);
    const double inv = 1.0 / pdfnorm;
    double ln = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        pv[t] *= inv;
        ln += pv[t];
    }
    // count of the grid points g < Ng - 1 whose cumulative probability is <= u
    double run = wave_exclusive_scan(ln), cnt = 0.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        run += pv[t];
        if (4 * lane + t < Ng - 1 && run <= u) cnt += 1.0;
    }
    gs = (int)wave_sum(cnt);
    const double *xs = grid + 4 * gs;
    fstar = quad(xs);
#pragma unroll
    for (int q = 0; q < 4; ++q) Ux[q] = Vx[q] = 0.0;
    for (int k = 0; k < 4; ++k) {
        const double pf = sqrt(tS[k] * 4.0) * xs[k];  // sqrt(4 S_k) x_k (MPSstate.cpp:591,616)
        for (int up = 0; up < 2; ++up)
            for (int dn = 0; dn < 2; ++dn) {
                Ux[up + 2 * dn] += pf * tU[up + 2 * (dn + 2 * k)];
                Vx[up + 2 * dn] += pf * tVT[k + 4 * (up + 2 * dn)];
            }
    }
}

template <typename T, bool CH>
__global__ __launch_bounds__(MQE_NT, 2) void walker_kernel(Sys S, Pool P, const T *trial, int flags) {
    const int n = *P.n;
    if ((int)blockIdx.x >= n) return;
    if (*P.status & 4) return;  // the population overflowed: keep the state as it was
    const int b = P.list ? P.list[blockIdx.x] : (int)blockIdx.x;  // population position
    const int slot = P.order[b];
    const Ctl ctl = *P.ctl;
    const int L = S.L;

    __shared__ int tv[MQE_MAXL + 1], wv[MQE_MAXL + 1];
    __shared__ double red[MQE_NT + 16 * MQE_WAVES];
    __shared__ double M[16], comb[16];

    __shared__ int sh_tile;
    for (int k = threadIdx.x; k <= L; k += MQE_NT) {
        tv[k] = S.tvd[k];
        wv[k] = P.wvd[(size_t)slot * (L + 1) + k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int mb = 1;
        for (int k = 0; k <= L; ++k) mb = max(mb, wv[k]);
        sh_tile = S.tiles[bond_bucket(mb)];
    }
    __syncthreads();

    Ctx<T, CH> C;
    C.tile = sh_tile;
    C.S = &S;
    C.trial = trial;
    C.trialT = S.trialT;
    C.phi = (T *)P.phi + (size_t)slot * P.wflat;
    C.Lenv = (T *)P.Lenv + (size_t)slot * P.eflat;
    C.Renv = (T *)P.Renv + (size_t)slot * P.eflat;
    C.woff = P.woff;
    C.eoff = P.eoff;
    C.moff = P.moff;
    C.tv = tv;
    C.wv = wv;
    C.sl = scratch_layout(S.Dt, S.Dw, S.Ng);
    char *base = P.scratch + (size_t)b * P.scratch_bytes;
    C.ts = (T *)base;
    long tb = (C.sl.t_total * (long)sizeof(T) + 255) & ~255L;
    C.ds = (double *)(base + tb);
    C.red = red;
    extern __shared__ double qlds_dyn[];
    C.qlds = P.qlds > 0 ? qlds_dyn : nullptr;
    C.qlds_doubles = P.qlds;

    double w = P.w[slot];
    double O = P.O[slot];

    set_one(C.Lk(0));

    if (flags & F_PREP) {
        for (int k = 0; k < L - 1; ++k) advance_L(C, k);
        double On, Ou;
        const double e = energy_sweep(C, &On, &Ou, true);
        if (threadIdx.x == 0) {
            P.EL[slot] = e / O;  // stored overlap, MPSQMC2.cpp:447 / Walker.cpp:72-77
            P.O[slot] = On;
            P.since[slot] = 0;
        }
        return;
    }

    // ---------------- propagation (MPSQMC2.cpp:532-600) ----------------
    if (S.h != 0.0) {
        apply_field_everywhere(C);
        O = right_canonicalize(C);
    }
    // The walker is not normalised between couplings: the sampled operator of
    // the right site is scaled so that the overlap stays O (see the file
    // header), and the energy sweep takes the norm.  p: left environments L_k
    // are valid for k <= p; rv: right environments R_k for k >= rv.
    int p = 0, rv = 1;
    for (int c = 0; c < S.nC; ++c) {
        const int i = S.first[c], j = S.second[c];
        // the left environments up to L_i and the right ones down to R_{j+1}: two
        // independent chains (CH: one step of each per barrier interval)
        if constexpr (CH) {
            while (p < i || rv > j + 1) {
                int wq = 0;
                if (p < i) wq = issue_L(C, p++, wq);
                if (rv > j + 1) issue_R(C, --rv, wq);
                __syncthreads();
            }
        } else {
            while (p < i) {
                advance_L(C, p);
                ++p;
            }
            while (rv > j + 1) rebuild_R(C, --rv);
        }
        pair_for_coupling(C, i, j, M);
        const double *tS = S.tS + 4 * c, *tU = S.tU + 16 * c, *tVT = S.tVT + 16 * c;
        const double u = P.u ? P.u[(size_t)b * S.nC + c] : philox_uniform(P.seed, (uint32_t)b, (uint32_t)c, (uint32_t)ctl.step, 0u);
        double pdfnorm, fstar, Ux[4], Vx[4];
        int gs;
        if (S.Ng <= 256) {
            wave_sample(M, tS, tU, tVT, S.grid, S.Ng, O, u, pdfnorm, gs, fstar, Ux, Vx);
        } else {
            // larger grids: the workgroup forms the PDF and scans it (MPSQMC2.cpp:540-588)
            if (threadIdx.x < 16) {
                const int a = threadIdx.x & 3, bb = threadIdx.x >> 2;
                double sacc = 0.0;
                for (int sp = 0; sp < 2; ++sp)
                    for (int ss = 0; ss < 2; ++ss) {
                        const double ua = tU[sp + 2 * (ss + 2 * a)];
                        for (int tp = 0; tp < 2; ++tp)
                            for (int t = 0; t < 2; ++t)
                                sacc += ua * tVT[bb + 4 * (tp + 2 * t)] * M[(sp + 2 * tp) * 4 + (ss + 2 * t)];
                    }
                comb[threadIdx.x] = sqrt(tS[a] * tS[bb]) * sacc;
            }
            __syncthreads();
            double *pdf = C.ds + C.sl.PDF, *cum = C.ds + C.sl.CUM;
            const int Ng = S.Ng;
            const double pref = 4.0 / (O * Ng);
            double part = 0.0;
            for (int g = threadIdx.x; g < Ng; g += MQE_NT) {
                const double *x = S.grid + 4 * g;
                double f = 0.0;
                for (int l = 0; l < 4; ++l)
                    for (int r = 0; r < 4; ++r) f += comb[l + 4 * r] * x[l] * x[r];
                const double v = (f > 0.0 ? f : 0.0) * pref;
                pdf[g] = v;
                part += v;
            }
            pdfnorm = block_sum(part, red);
            const double inv = 1.0 / pdfnorm;
            for (int g = threadIdx.x; g < Ng; g += MQE_NT) pdf[g] *= inv;
            __syncthreads();
            block_inclusive_scan(Ng, pdf, cum, red);
            double cnt = 0.0;
            for (int g = threadIdx.x; g < Ng - 1; g += MQE_NT) cnt += (cum[g] <= u) ? 1.0 : 0.0;
            gs = (int)block_sum(cnt, red);
            const double *x = S.grid + 4 * gs;
            fstar = 0.0;
            for (int l = 0; l < 4; ++l)
                for (int r = 0; r < 4; ++r) fstar += comb[l + 4 * r] * x[l] * x[r];
            for (int q = 0; q < 4; ++q) Ux[q] = Vx[q] = 0.0;
            for (int k = 0; k < 4; ++k) {
                const double pf = sqrt(tS[k] * 4.0) * x[k];  // MPSstate.cpp:591,616
                for (int up = 0; up < 2; ++up)
                    for (int dn = 0; dn < 2; ++dn) {
                        Ux[up + 2 * dn] += pf * tU[up + 2 * (dn + 2 * k)];
                        Vx[up + 2 * dn] += pf * tVT[k + 4 * (up + 2 * dn)];
                    }
            }
        }
        w *= pdfnorm;
        if (threadIdx.x == 0) {
            if (P.idx) P.idx[(size_t)b * S.nC + c] = gs;
            if (!(pdfnorm > 0.0) || !isfinite(pdfnorm)) atomicOr(P.status, 1);
        }
        // the sampled operators on sites i and j (MPSstate.cpp:575-633), the right
        // one scaled by O / <trial|B(x) phi> = O / (4 f*) so that the overlap stays O;
        // for nearest neighbours whose next coupling starts right of i, the left
        // boundary moves to j: L_{i+1} = sum_{s,s0} Ux[s,s0] G(s, s0).  One pass.
        const double alpha = O / (4.0 * fstar);
        const bool nn_move = j == i + 1 && S.nextfirst[c] > i;
        {
            T *pi = C.site(i), *pj = C.site(j);
            const int bi = wv[i] * wv[i + 1], bj = wv[j] * wv[j + 1];
            const T u00 = (T)Ux[0], u10 = (T)Ux[1], u01 = (T)Ux[2], u11 = (T)Ux[3];
            const T v00 = (T)(alpha * Vx[0]), v10 = (T)(alpha * Vx[1]), v01 = (T)(alpha * Vx[2]),
                    v11 = (T)(alpha * Vx[3]);
            mix_pair(pi, bi, u00, u01, u10, u11);
            mix_pair(pj, bj, v00, v01, v10, v11);
            if (nn_move) {
                const int tn = tv[i + 1], wn = wv[i + 1];
                const T *G = C.ts + C.sl.G;
                T *Ln = C.Lk(i + 1);
                for (int e = threadIdx.x; e < tn * wn; e += MQE_NT) {
                    const int a = e % tn, bq = e / tn;
                    const size_t o = a + (size_t)2 * tn * bq;
                    const T g00 = G[o], g01 = G[o + (size_t)wn * 2 * tn], g10 = G[o + tn],
                            g11 = G[o + tn + (size_t)wn * 2 * tn];
                    Ln[a + (size_t)tn * bq] = u00 * g00 + u01 * g01 + u10 * g10 + u11 * g11;
                }
            }
            __syncthreads();
        }
        p = nn_move ? i + 1 : i;  // else site i changed: L_{i+1} is rebuilt when needed
        rv = max(rv, j + 1);  // site j changed
    }
    if (S.h != 0.0) {
        apply_field_everywhere(C);
        p = 0;
    }
    while (p < L - 1) {
        advance_L(C, p);
        ++p;
    }
    // ---------------- local energy (Walker.cpp:72-77) ----------------
    double On, Ou;
    const bool canon = canon_walker(ctl, slot, P.since[slot]);
    const double e = energy_sweep(C, &On, &Ou, canon);
    if (threadIdx.x == 0) {
        P.since[slot] = canon ? 0 : P.since[slot] + 1;
        P.w[slot] = w;
        P.O[slot] = On;
        P.EL[slot] = e / Ou;
        if (!isfinite(w)) atomicOr(P.status, 2);
    }
}

// out[b] = <other | walker_b> (transfer matrices, MPSstate.cpp:248-284).
template <typename T>
__global__ __launch_bounds__(MQE_NT) void overlap_kernel(int L, const int *ovd, const long *ooff, const T *other,
                                                       Pool P, int Dmax, double *out, int first) {
    const int b = first + blockIdx.x;  // population position; workspace blockIdx.x of P.scratch
    if (b >= *P.n) return;
    const int slot = P.order[b];
    __shared__ int ov[MQE_MAXL + 1], wv[MQE_MAXL + 1];
    for (int k = threadIdx.x; k <= L; k += MQE_NT) {
        ov[k] = ovd[k];
        wv[k] = P.wvd[(size_t)slot * (L + 1) + k];
    }
    __syncthreads();
    const T *phi = (const T *)P.phi + (size_t)slot * P.wflat;
    T *E = (T *)(P.scratch + (size_t)blockIdx.x * P.scratch_bytes);
    T *E2 = E + (size_t)Dmax * Dmax;
    T *Y = E2 + (size_t)Dmax * Dmax;
    set_one(E);
    for (int k = 0; k < L; ++k) {
        const int oa = ov[k], ob = ov[k + 1], wa = wv[k], wb = wv[k + 1];
        const T *A = other + ooff[k];
        const T *B0 = P.padded ? phi + (size_t)(2 * k) * 1024 : phi + P.woff[k];
        const T *B1 = P.padded ? phi + (size_t)(2 * k + 1) * 1024 : phi + P.woff[k] + (size_t)wa * wb;
        const int ldb = P.padded ? 32 : wa;
        // Y = E A (wa x 2 ob) with E (wa x oa)
        gemm<false, false, T>(wa, 2 * ob, oa, E, wa, A, oa, Y, wa, false);
        __syncthreads();
        // E2 = sum_s B[s]^T Y[s]  (wb x ob)
        gemm<true, false, T>(wb, ob, wa, B0, ldb, Y, wa, E2, wb, false);
        __syncthreads();
        gemm<true, false, T>(wb, ob, wa, B1, ldb, Y + (size_t)wa * ob, wa, E2, wb, true);
        __syncthreads();
        T *t = E; E = E2; E2 = t;
    }
    if (threadIdx.x == 0) out[b] = (double)E[0];
}

// ---------------------------------------------------------------------------
// Recompression of walkers (MPSstate::CompressState, MPSstate.cpp:379-404):
// a left-to-right and a right-to-left sweep of two-site SVDs
// (TwoSiteObject::Compose / Decompose, TwoSiteObject.cpp:71-184), keeping
// min(D, number of Schmidt values >= 1e-13) states per bond (:16, :147-157);
// moving right the Schmidt values go to the right tensor, moving left to the
// left one (:168, :177).  FP64 in global scratch, one workgroup per walker.
// ---------------------------------------------------------------------------
constexpr double SCHMIDT_CUT = 1e-13;

// walker tensor element (row r, column c, physical s) of site k with dims (dl, dr)
template <typename T>
__device__ __forceinline__ T *wel(const Pool &P, T *phi, int k, int dl, int dr, int s, int r, int c) {
    return P.padded ? phi + (size_t)(2 * k + s) * 1024 + r + 32 * c : phi + P.woff[k] + r + (size_t)dl * (c + (size_t)dr * s);
}

template <typename T>
__device__ void compress_bond(const Pool &P, T *phi, int *wv, int k, int D, bool movingright, double *X, double *V,
                              double *sig, int *ord, int *sh_flag) {
    const int dl = wv[k], dm = wv[k + 1], dr = wv[k + 2];
    const int m = 2 * dl, n = 2 * dr;
    const bool tr = n > m;  // Jacobi on the columns of the orientation with fewer of them
    const int rows = tr ? n : m, cols = tr ? m : n;
    // Compose: theta(alpha + dl il, beta + dr ir) = sum_mid A_k[il](alpha, mid) A_{k+1}[ir](mid, beta)
    for (int e = threadIdx.x; e < m * n; e += MQE_NT) {
        const int row = e % m, col = e / m;
        const int al = row % dl, il = row / dl, be = col % dr, ir = col / dr;
        double s = 0.0;
        for (int q = 0; q < dm; ++q)
            s += (double)*wel(P, phi, k, dl, dm, il, al, q) * (double)*wel(P, phi, k + 1, dm, dr, ir, q, be);
        if (tr) X[col + (size_t)rows * row] = s;
        else X[row + (size_t)rows * col] = s;
    }
    __syncthreads();
    jacobi_svd_columns(X, rows, cols, V, sig, sh_flag);
    sort_descending(sig, cols, ord);
    int keep = 0;
    for (int j = 0; j < cols; ++j) keep += sig[ord[j]] >= SCHMIDT_CUT ? 1 : 0;
    keep = max(1, min(keep, D));
    // left singular vector j (length m) and right one (length n) of theta
    auto u = [&](int j, int row) {
        const int c = ord[j];
        return tr ? V[row + (size_t)cols * c] : (sig[c] > 0.0 ? X[row + (size_t)rows * c] / sig[c] : 0.0);
    };
    auto vt = [&](int j, int col) {
        const int c = ord[j];
        return tr ? (sig[c] > 0.0 ? X[col + (size_t)rows * c] / sig[c] : 0.0) : V[col + (size_t)cols * c];
    };
    // Decompose: the new bond holds `keep` states (zero padding cleared on the tile layout)
    if (P.padded) {
        for (int e = threadIdx.x; e < 2 * 1024; e += MQE_NT) {
            phi[(size_t)(2 * k) * 1024 + e] = (T)0;
            phi[(size_t)(2 * k + 2) * 1024 + e] = (T)0;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < m * keep; e += MQE_NT) {
        const int row = e % m, j = e / m, al = row % dl, il = row / dl;
        const double f = movingright ? 1.0 : sig[ord[j]];
        *wel(P, phi, k, dl, keep, il, al, j) = (T)(u(j, row) * f);
    }
    for (int e = threadIdx.x; e < keep * n; e += MQE_NT) {
        const int j = e % keep, col = e / keep, be = col % dr, ir = col / dr;
        const double f = movingright ? sig[ord[j]] : 1.0;
        *wel(P, phi, k + 1, keep, dr, ir, j, be) = (T)(vt(j, col) * f);
    }
    __syncthreads();
    if (threadIdx.x == 0) wv[k + 1] = keep;
    __syncthreads();
}

// Walkers first .. first+gridDim-1 (population order) to bond targets[b];
// afterwards sites 1..L-1 are right-orthonormal and site 0 is scaled to unit
// norm (the LeftNormalize of Walker::setOverlap, Walker.cpp:54-66, up to gauge).
template <typename T>
__global__ __launch_bounds__(MQE_NT) void compress_kernel(int L, Pool P, int first, const int *targets, char *scr,
                                                          long scr_bytes, int Dmax) {
    const int b = first + blockIdx.x;
    if (b >= *P.n) return;
    const int slot = P.order[b];
    __shared__ int wv[MQE_MAXL + 1];
    __shared__ int flag;
    __shared__ double red[MQE_NT + 16 * MQE_WAVES];
    for (int k = threadIdx.x; k <= L; k += MQE_NT) wv[k] = P.wvd[(size_t)slot * (L + 1) + k];
    __syncthreads();
    T *phi = (T *)P.phi + (size_t)slot * P.wflat;
    const long n2 = 4L * Dmax * Dmax;
    double *X = (double *)(scr + (size_t)blockIdx.x * scr_bytes), *V = X + n2, *sig = V + n2;
    int *ord = (int *)(sig + 2 * Dmax);
    const int D = targets[b];
    if (D < 1) return;  // walker left as it is
    for (int k = 0; k < L - 1; ++k) compress_bond(P, phi, wv, k, D, true, X, V, sig, ord, &flag);
    for (int k = L - 2; k >= 0; --k) compress_bond(P, phi, wv, k, D, false, X, V, sig, ord, &flag);
    const int w1 = wv[1];
    double s2 = 0.0;
    for (int e = threadIdx.x; e < 2 * w1; e += MQE_NT) {
        const double v = (double)*wel(P, phi, 0, 1, w1, e / w1, 0, e % w1);
        s2 += v * v;
    }
    const double inv = 1.0 / sqrt(block_sum(s2, red));
    for (int e = threadIdx.x; e < 2 * w1; e += MQE_NT) {
        T *p = wel(P, phi, 0, 1, w1, e / w1, 0, e % w1);
        *p = (T)((double)*p * inv);
    }
    for (int k = threadIdx.x; k <= L; k += MQE_NT) P.wvd[(size_t)slot * (L + 1) + k] = wv[k];
}

// Walker::setOverlap's sign convention (Walker.cpp:59-62): a negative overlap
// with the trial flips site 0; the stored overlap is |<trial|walker>|.
template <typename T>
__global__ void overlap_sign_kernel(int L, Pool P, const double *ov, const int *targets) {
    const int b = blockIdx.x;
    if (b >= *P.n || targets[b] < 1) return;
    const int slot = P.order[b];
    const double o = ov[b];
    if (o < 0.0) {
        T *phi = (T *)P.phi + (size_t)slot * P.wflat;
        const int w1 = P.wvd[(size_t)slot * (L + 1) + 1];
        for (int e = threadIdx.x; e < 2 * w1; e += blockDim.x) {
            T *p = wel(P, phi, 0, 1, w1, e / w1, 0, e % w1);
            *p = -*p;
        }
    }
    if (threadIdx.x == 0) P.O[slot] = fabs(o);
}

}  // namespace mqe
#include "mqe_tile32.h"
namespace mqe {

// ---------------------------------------------------------------------------
// reductions, bookkeeping and population control (single workgroup, 1024 threads)
// ---------------------------------------------------------------------------
#define RED_NT 1024

__device__ double rsum(double v, double *sh) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < RED_NT / 64; ++q) t += sh[q];
    __syncthreads();
    return t;
}

// exclusive scan of ints over RED_NT threads (one value per thread); returns the total.
__device__ int rscan_excl(int v, int *out_excl, int *sh) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int s = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int o = __shfl_up(s, off, 64);
        if (lane >= off) s += o;
    }
    __syncthreads();
    if (lane == 63) sh[wid] = s;
    __syncthreads();
    int base = 0, tot = 0;
    for (int q = 0; q < RED_NT / 64; ++q) {
        if (q < wid) base += sh[q];
        tot += sh[q];
    }
    *out_excl = base + s - v;
    __syncthreads();
    return tot;
}

struct Book {
    const int *n;
    int *order, *order2;
    double *w, *O, *eh, *EL;
    double *sums;     // [0] sum w  [1] scaling  [2] E_T  [3] E_P  [4] Omega  [5] N
    double *trace;    // 6 doubles per step
    int *ncopy, *newpos, *used, *freelist, *copy_src, *copy_dst, *n_extra, *n_new, *n_used;
    double *nw;       // population control: new weights by position, committed once the copies fit
    int *status;
    const Ctl *ctl;
    int cap, n_desired;
    double dtau;
};

__global__ void set_ctl_kernel(Ctl *ctl, Ctl v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ctl = v;
}

// start of a captured step: next step number and propagation
__global__ void advance_ctl_kernel(Ctl *ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->step += 1;
        ctl->pstep += 1;
        ctl->pc += 1;
    }
}

// Population positions in launch order, most work first: the walkers that
// re-orthonormalise in this propagation (several times the work of a light
// step), then by bond bucket from the largest bonds down, each class in
// population order.  Heavy walkers start in the first round of resident
// workgroups instead of extending the last one.  key: scratch, one int per position.
__global__ __launch_bounds__(RED_NT) void launch_list_kernel(const int *n_ptr, const int *order, const int *wvd,
                                                             const int *since, int L, const Ctl *ctl, int *key,
                                                             int *list) {
    __shared__ int cnt[16], base[16], carry[16];
    __shared__ int wcnt[RED_NT / 64][16];
    const int n = *n_ptr;
    const Ctl c = *ctl;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (threadIdx.x < 16) {
        cnt[threadIdx.x] = 0;
        carry[threadIdx.x] = 0;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < n; b += RED_NT) {
        const int slot = order[b];
        const int *wv = wvd + (size_t)slot * (L + 1);
        int mb = 1;
        // eight independent loads in flight at a time (one at a time, the loop
        // waited L + 1 memory latencies per walker)
#pragma unroll 8
        for (int k = 0; k <= L; ++k) mb = max(mb, wv[k]);
        const int kk = (canon_walker(c, slot, since[slot]) ? 8 : 0) + bond_bucket(mb);
        key[b] = kk;
        atomicAdd(&cnt[kk], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int kk = 15; kk >= 0; --kk) { base[kk] = run; run += cnt[kk]; }
    }
    int present = 0;  // classes that occur (block-uniform)
    for (int kk = 0; kk < 16; ++kk) present |= (cnt[kk] > 0 ? 1 : 0) << kk;
    // stable placement in one pass: a position's rank among the positions of its
    // class before it is its rank inside its wave (ballot) plus the counts of the
    // earlier waves of the chunk and of the earlier chunks
    for (int b0 = 0; b0 < n; b0 += RED_NT) {
        const int b = b0 + threadIdx.x;
        const int kk = b < n ? key[b] : -1;
        int rank = 0;
        for (int q = 0; q < 16; ++q) {
            if (!((present >> q) & 1)) continue;
            const unsigned long long m = __ballot(kk == q);
            if (kk == q) rank = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) wcnt[wid][q] = __popcll(m);
        }
        __syncthreads();
        if (kk >= 0) {
            int off = base[kk] + carry[kk] + rank;
            for (int w = 0; w < wid; ++w) off += wcnt[w][kk];
            list[off] = b;
        }
        __syncthreads();
        if (threadIdx.x < 16 && ((present >> threadIdx.x) & 1)) {
            int t = 0;
            for (int w = 0; w < RED_NT / 64; ++w) t += wcnt[w][threadIdx.x];
            carry[threadIdx.x] += t;
        }
        __syncthreads();
    }
}

// sum of weights; if device_scaling also scaling and E_T (MPSQMC2.cpp:480-484)
__global__ __launch_bounds__(RED_NT) void sumw_kernel(Book B) {
    __shared__ double sh[RED_NT / 64];
    const int n = *B.n;
    double s = 0.0;
    for (int b = threadIdx.x; b < n; b += RED_NT) s += B.w[B.order[b]];
    s = rsum(s, sh);
    if (threadIdx.x == 0) {
        B.sums[0] = s;
        const double avg = s / n;
        const double scaling = B.n_desired / (n * avg);
        B.sums[1] = scaling;
        B.sums[2] = log(scaling) / B.dtau;
        B.sums[5] = n;
    }
}

// EnergyFunctionAndHistory (MPSQMC2.cpp:688-747).  from_ctl: step number from
// the control block and a trace row for it (a step of mqe_step / mqe_run).
__global__ __launch_bounds__(RED_NT) void energy_kernel(Book B, int step, int dofluct, int trace_row, int from_ctl) {
    __shared__ double sh[RED_NT / 64];
    const int n = *B.n;
    if (from_ctl) {
        step = B.ctl->step;
        trace_row = (step - 1) % B.ctl->trace_cap;
    }
    double num = 0.0, den = 0.0, s1 = 0.0, s2 = 0.0;
    for (int b = threadIdx.x; b < n; b += RED_NT) {
        const int sl = B.order[b];
        const double e = B.EL[sl], w = B.w[sl];
        num += w * e;
        den += w;
        if (dofluct) {
            const double h = B.eh[sl] + e;
            B.eh[sl] = h;
            const double a = h / step;
            s1 += a;
            s2 += a * a;
        }
    }
    num = rsum(num, sh);
    den = rsum(den, sh);
    s1 = rsum(s1, sh);
    s2 = rsum(s2, sh);
    if (threadIdx.x == 0) {
        B.sums[3] = num / den;
        B.sums[4] = dofluct ? (s2 - s1 * s1 / n) / n : 0.0;
        if (trace_row >= 0) {
            double *r = B.trace + 6L * trace_row;
            r[0] = step; r[1] = n; r[2] = B.sums[3]; r[3] = B.sums[2]; r[4] = B.sums[4]; r[5] = B.sums[0];
        }
    }
}

// SeparatePopulationControl (MPSQMC2.cpp:613-686): copy counts, new order,
// slot assignment for the extra copies.  scaling from host (>0) or B.sums[1].
__global__ __launch_bounds__(RED_NT) void popcontrol_kernel(Book B, double host_scaling, const double *u, int nu,
                                                            unsigned long long seed) {
    __shared__ int shi[RED_NT / 64];
    __shared__ int carry_draw, carry_pos, carry_extra, carry_free;
    const int n = *B.n;
    if (*B.status & 4) {  // an earlier overflow: leave the population as it is
        if (threadIdx.x == 0) { *B.n_extra = 0; *B.n_new = B.cap + 1; *B.n_used = 0; }  // order and n unchanged
        return;
    }
    const int step = B.ctl->pstep;
    const double scaling = host_scaling > 0.0 ? host_scaling : B.sums[1];
    if (threadIdx.x == 0) { carry_draw = 0; carry_pos = 0; carry_extra = 0; carry_free = 0; }
    for (int s = threadIdx.x; s < B.cap; s += RED_NT) B.used[s] = 0;
    __syncthreads();
    // pass 1: copy counts in population order, draws consumed in order
    for (int base = 0; base < n; base += RED_NT) {
        const int b = base + threadIdx.x;
        double sw = 0.0;
        int need = 0;
        if (b < n) {
            sw = B.w[B.order[b]] * scaling;
            need = (sw < 0.25 || sw > 1.5) ? 1 : 0;
        }
        int ex;
        const int tot = rscan_excl(need, &ex, shi);
        int c = 1;
        double nw = sw;
        if (b < n) {
            if (need) {
                const int k = carry_draw + ex;
                const double r = u ? (k < nu ? u[k] : 0.0)
                                   : philox_uniform(seed, (uint32_t)b, 0xFFFFFFFFu, (uint32_t)step, 1u);
                c = (int)(sw + r);
                nw = (sw < 0.25) ? 1.0 : sw / c;
            }
            B.ncopy[b] = c;
            B.nw[b] = nw;  // committed below once the copies are known to fit
        }
        int pos;
        const int totc = rscan_excl(b < n ? c : 0, &pos, shi);
        if (b < n) B.newpos[b] = carry_pos + pos;
        __syncthreads();
        if (threadIdx.x == 0) { carry_draw += tot; carry_pos += totc; }
        __syncthreads();
    }
    const int total = carry_pos;
    if (threadIdx.x == 0) { *B.n_used = carry_draw; *B.n_new = total; }
    if (total > B.cap) {
        if (threadIdx.x == 0) { atomicOr(B.status, 4); *B.n_extra = 0; }
        return;
    }
    // new weights (survivors; the copies read them in the copy kernel)
    for (int b = threadIdx.x; b < n; b += RED_NT) B.w[B.order[b]] = B.nw[b];
    // survivors keep their slot
    for (int b = threadIdx.x; b < n; b += RED_NT)
        if (B.ncopy[b] > 0) B.used[B.order[b]] = 1;
    __syncthreads();
    // free list (slots not used by survivors), in slot order
    for (int base = 0; base < B.cap; base += RED_NT) {
        const int s = base + threadIdx.x;
        const int f = (s < B.cap && !B.used[s]) ? 1 : 0;
        int ex;
        const int tot = rscan_excl(f, &ex, shi);
        if (f) B.freelist[carry_free + ex] = s;
        __syncthreads();
        if (threadIdx.x == 0) carry_free += tot;
        __syncthreads();
    }
    // new order and copy list
    for (int base = 0; base < n; base += RED_NT) {
        const int b = base + threadIdx.x;
        const int c = (b < n) ? B.ncopy[b] : 0;
        int ex;
        const int tot = rscan_excl(c > 1 ? c - 1 : 0, &ex, shi);
        if (b < n && c > 0) {
            const int pos = B.newpos[b];
            const int slot = B.order[b];
            B.order2[pos] = slot;
            for (int q = 1; q < c; ++q) {
                const int e = carry_extra + ex + q - 1;
                const int dst = B.freelist[e];
                B.order2[pos + q] = dst;
                B.copy_src[e] = slot;
                B.copy_dst[e] = dst;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) carry_extra += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *B.n_extra = carry_extra;
}

// Copy walker data for the extra copies; then install the new order.  The
// bulk (walker tensors and right environments of every copy) is spread over
// the whole grid in 16-byte chunks where the layout allows it, so the copies
// run at memory bandwidth instead of one workgroup per copy.
template <typename T>
__global__ __launch_bounds__(MQE_NT) void copy_walkers_kernel(Book B, Pool P, int L) {
    const int ne = *B.n_extra;
    const long gid = (long)blockIdx.x * MQE_NT + threadIdx.x, gsz = (long)gridDim.x * MQE_NT;
    for (long e = gid; e < ne; e += gsz) {  // scalars, bond dims, bond energies
        const int src = B.copy_src[e], dst = B.copy_dst[e];
        P.w[dst] = P.w[src];
        P.O[dst] = P.O[src];
        P.eh[dst] = P.eh[src];
        P.EL[dst] = P.EL[src];
        P.since[dst] = P.since[src];
        for (int k = 0; k <= L; ++k) P.wvd[(size_t)dst * (L + 1) + k] = P.wvd[(size_t)src * (L + 1) + k];
        if (P.bondE)
            for (int k = 0; k < P.nbe; ++k) P.bondE[(size_t)dst * P.nbe + k] = P.bondE[(size_t)src * P.nbe + k];
    }
    if (P.plab) {
        const long per = 16L * (P.nbe - 1);
        for (long x = gid; x < (long)ne * per; x += gsz) {
            const long e = x / per, q = x - e * per;
            P.plab[(size_t)B.copy_dst[e] * per + q] = P.plab[(size_t)B.copy_src[e] * per + q];
        }
    }
    const bool vec = (P.wflat * (long)sizeof(T)) % 16 == 0 && (P.eflat * (long)sizeof(T)) % 16 == 0;
    if (vec) {
        const long wq = P.wflat * (long)sizeof(T) / 16, eq = P.eflat * (long)sizeof(T) / 16, per = wq + eq;
        for (long x = gid; x < (long)ne * per; x += gsz) {
            const long e = x / per, q = x - e * per;
            const int src = B.copy_src[e], dst = B.copy_dst[e];
            const float4 *s4 = q < wq ? (const float4 *)P.phi + (size_t)src * wq + q
                                      : (const float4 *)P.Renv + (size_t)src * eq + (q - wq);
            float4 *d4 = q < wq ? (float4 *)P.phi + (size_t)dst * wq + q : (float4 *)P.Renv + (size_t)dst * eq + (q - wq);
            *d4 = *s4;
        }
    } else {
        const long per = P.wflat + P.eflat;
        for (long x = gid; x < (long)ne * per; x += gsz) {
            const long e = x / per, q = x - e * per;
            const int src = B.copy_src[e], dst = B.copy_dst[e];
            if (q < P.wflat)
                ((T *)P.phi)[(size_t)dst * P.wflat + q] = ((const T *)P.phi)[(size_t)src * P.wflat + q];
            else
                ((T *)P.Renv)[(size_t)dst * P.eflat + q - P.wflat] =
                    ((const T *)P.Renv)[(size_t)src * P.eflat + q - P.wflat];
        }
    }
}

// Features of the learned proposal, row b*nC + c (MQE_NFEAT floats): log O,
// log w, E_L, e_{c-1}, e_c, e_{c+1}, c/(nC-1) with e_k = bond energy / overlap
// of the last energy sweep, then the walker's 16 recorded labels of coupling c.  X: rows of this step (or nullptr); the sample
// buffer gets the same rows after its *nsamp rows if they fit.
__global__ void feature_kernel(Pool P, int nC, float *X, float *ring) {
    const int n = *P.n;
    const long base = P.nsamp ? *P.nsamp : 0;
    const bool to_ring = ring && base + (long)n * nC <= P.maxsamp;
    for (int b = blockIdx.x; b < n; b += gridDim.x) {
        const int slot = P.order[b];
        const double *be = P.bondE + (size_t)slot * (nC + 1);
        const double inv = be[nC] != 0.0 ? 1.0 / be[nC] : 0.0;  // no sweep yet: zero bond features
        const float lO = (float)log(P.O[slot]), lw = (float)log(fmax(P.w[slot], 1e-300)), EL = (float)P.EL[slot];
        for (int c = threadIdx.x; c < nC; c += blockDim.x) {
            float f[MQE_NFEAT] = {lO, lw, EL, c > 0 ? (float)(be[c - 1] * inv) : 0.f, (float)(be[c] * inv),
                                  c + 1 < nC ? (float)(be[c + 1] * inv) : 0.f, nC > 1 ? (float)c / (nC - 1) : 0.f};
            const float *pl = P.plab + ((size_t)slot * nC + c) * 16;
            for (int q = 0; q < 16; ++q) f[7 + q] = pl[q];
            const long row = (long)b * nC + c;
            for (int q = 0; q < MQE_NFEAT; ++q) {
                if (X) X[row * MQE_NFEAT + q] = f[q];
                if (to_ring) ring[(base + row) * MQE_NFEAT + q] = f[q];
            }
        }
    }
}

// The sample buffer took this step's rows (if they fitted).
__global__ void advance_samples_kernel(const int *n, int nC, long *nsamp, long maxsamp) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const long add = (long)(*n) * nC;
        if (*nsamp + add <= maxsamp) *nsamp += add;
    }
}

__global__ void install_order_kernel(Book B) {
    const int total = *B.n_new;
    if (total > B.cap) return;
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < total; b += gridDim.x * blockDim.x) B.order[b] = B.order2[b];
}

__global__ void set_n_kernel(Book B, int *n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const int t = *B.n_new;
        if (t <= B.cap) *n = t;
    }
}

}  // namespace mqe

// ===========================================================================
// host engine
// ===========================================================================
using namespace mqe;

#define HIPCHK(x)                                                                            \
    do {                                                                                     \
        hipError_t _e = (x);                                                                 \
        if (_e != hipSuccess) {                                                              \
            char _b[256];                                                                    \
            snprintf(_b, sizeof(_b), "HIP error %s at %s:%d", hipGetErrorString(_e), __FILE__, \
                     __LINE__);                                                              \
            set_error(_b);                                                                   \
            return MQE_EDEVICE;                                                              \
        }                                                                                    \
    } while (0)

struct mqe_engine {
    mqe_params p;
    int L = 0, nC = 0, Ng = 0, gdim = 0, Dt = 0, Dw = 0;
    bool fp32 = false;
    bool tile32 = false;   // FP32 MFMA path for NN chains with bonds <= 32, h = 0
    float *d_t32N = nullptr, *d_t32T = nullptr;
    float *d_trialT = nullptr;  // chained general kernel: every T_k[s] transposed
    bool chained = false;       // the general kernel with chained FP32 products (bonds <= 64)
    hipStream_t stream = nullptr;
    std::vector<int> tvd, wmax;
    std::vector<long> toff, woff, eoff, moff;
    long wflat = 0, eflat = 0, scratch_bytes = 0;
    // device
    void *d_trial = nullptr;
    int *d_tvd = nullptr, *d_first = nullptr, *d_second = nullptr, *d_nextfirst = nullptr, *d_csptr = nullptr,
        *d_csidx = nullptr;
    long *d_toff = nullptr, *d_woff = nullptr, *d_eoff = nullptr, *d_moff = nullptr;
    double *d_Jc = nullptr, *d_S = nullptr, *d_U = nullptr, *d_VT = nullptr, *d_grid = nullptr;
    void *d_phi = nullptr, *d_Renv = nullptr, *d_Lenv = nullptr;
    int *d_wvd = nullptr;
    double *d_w = nullptr, *d_O = nullptr, *d_eh = nullptr, *d_EL = nullptr;
    int *d_order = nullptr, *d_order2 = nullptr, *d_n = nullptr, *d_idx = nullptr;
    double *d_u = nullptr;
    long u_cap = 0;
    char *d_scratch = nullptr;
    int *d_status = nullptr;
    // learned proposal (tile path)
    size_t gen_lds_bytes = 0;                  // general path: dynamic LDS of the LQ workspace (0: global scratch)
    double *d_bond = nullptr;                  // per slot: nC bond energies and the overlap of the last sweep
    float *d_feat = nullptr, *d_pred = nullptr;  // rows b*nC + c of the current step
    float *d_sX = nullptr, *d_sY = nullptr;    // collected samples (features, labels)
    float *d_plab = nullptr;                   // per slot: labels of the last recorded propagation
    long *d_nsamp = nullptr;
    long max_samples = 0;
    mqe_mlp *net = nullptr;
    bool learn = false;
    double mix = 0.1;
    double *d_sums = nullptr, *d_trace = nullptr;
    int trace_cap = 0;
    int *d_ncopy = nullptr, *d_newpos = nullptr, *d_used = nullptr, *d_free = nullptr, *d_csrc = nullptr,
        *d_cdst = nullptr, *d_nextra = nullptr, *d_nnew = nullptr, *d_nused = nullptr;
    double *d_tmpout = nullptr, *d_nw = nullptr;
    Ctl *d_ctl = nullptr;                      // step control block (device)
    Ctl ctl_host{};                            // its last value set from the host
    int *d_list = nullptr, *d_key = nullptr;   // launch order of the walker kernel and its class keys
    int *d_since = nullptr;                    // per slot: propagations since the last re-orthonormalisation
    long prop_draws = 0, pop_draws = 0;        // step words of mqe_propagate / mqe_population_control without host draws
    int n_host = 0;
    bool loaded = false, prepared = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
    int last_walkers = 0;
    hipGraphExec_t graph_exec = nullptr;       // mqe_run with use_graph: graph_steps captured steps
    int graph_steps = 0;
    // per-launch timing of the walker kernel inside mqe_step / mqe_run
    int canon_every = 64;  // FP64 LQ re-orthonormalisation of every walker every this many steps (staggered)
    long prop_count = 0;
    bool timing = false;
    std::vector<hipEvent_t> tev;
    int tev_used = 0;
    Sys sys;
};

static Pool make_pool(mqe_engine *e, const double *u) {
    Pool P;
    P.phi = e->d_phi;
    P.Renv = e->d_Renv;
    P.Lenv = e->d_Lenv;
    P.wvd = e->d_wvd;
    P.w = e->d_w;
    P.O = e->d_O;
    P.eh = e->d_eh;
    P.EL = e->d_EL;
    P.order = e->d_order;
    P.n = e->d_n;
    P.idx = e->d_idx;
    P.u = u;
    P.wflat = e->wflat;
    P.eflat = e->eflat;
    P.woff = e->d_woff;
    P.eoff = e->d_eoff;
    P.moff = e->d_moff;
    P.seed = e->p.seed;
    P.ctl = e->d_ctl;
    P.list = nullptr;
    P.since = e->d_since;
    P.status = e->d_status;
    P.scratch = e->d_scratch;
    P.scratch_bytes = e->scratch_bytes;
    P.padded = e->tile32 ? 1 : 0;
    P.qlds = (int)(e->gen_lds_bytes / sizeof(double));
    P.bondE = e->d_bond;
    P.nbe = e->nC + 1;
    P.pred = e->learn ? e->d_pred : nullptr;
    P.mix = e->mix;
    P.lab = e->max_samples > 0 ? e->d_sY : nullptr;
    P.plab = e->d_plab;
    P.plab_on = (e->max_samples > 0 || e->learn) ? 1 : 0;
    P.nsamp = e->d_nsamp;
    P.maxsamp = e->max_samples;
    return P;
}

static Book make_book(mqe_engine *e) {
    Book B;
    B.n = e->d_n;
    B.order = e->d_order;
    B.order2 = e->d_order2;
    B.w = e->d_w;
    B.O = e->d_O;
    B.eh = e->d_eh;
    B.EL = e->d_EL;
    B.sums = e->d_sums;
    B.trace = e->d_trace;
    B.ncopy = e->d_ncopy;
    B.newpos = e->d_newpos;
    B.used = e->d_used;
    B.freelist = e->d_free;
    B.copy_src = e->d_csrc;
    B.copy_dst = e->d_cdst;
    B.n_extra = e->d_nextra;
    B.n_new = e->d_nnew;
    B.n_used = e->d_nused;
    B.nw = e->d_nw;
    B.status = e->d_status;
    B.ctl = e->d_ctl;
    B.cap = e->p.capacity;
    B.n_desired = e->p.n_desired;
    B.dtau = e->p.dtau;
    return B;
}

template <typename T, bool CH>
static void launch_walker(mqe_engine *e, const Pool &P, int flags) {
    hipLaunchKernelGGL((walker_kernel<T, CH>), dim3(e->p.capacity), dim3(MQE_NT), e->gen_lds_bytes, e->stream, e->sys,
                       P, (const T *)e->d_trial, flags);
}

// Copy a control block value to the device with a kernel (no memset or copy
// node, so the same call may sit in front of a graph replay).
static int push_ctl(mqe_engine *e) {
    e->ctl_host.canon_every = e->canon_every;
    e->ctl_host.trace_cap = e->trace_cap;
    hipLaunchKernelGGL(set_ctl_kernel, dim3(1), dim3(64), 0, e->stream, e->d_ctl, e->ctl_host);
    HIPCHK(hipGetLastError());
    return MQE_OK;
}

// F_PREP: canonical form and local energies of the loaded walkers.  F_PROP: one
// propagation with the step words and phase of the control block, walkers in
// the launch order of launch_list_kernel.
static int launch_walker_any(mqe_engine *e, const double *u, int flags) {
    Pool P = make_pool(e, u);
    const bool prop = (flags & F_PROP) != 0;
    if (prop) {
        ++e->prop_count;
        hipLaunchKernelGGL(launch_list_kernel, dim3(1), dim3(RED_NT), 0, e->stream, e->d_n, e->d_order, e->d_wvd,
                           e->d_since, e->L, e->d_ctl, e->d_key, e->d_list);
        HIPCHK(hipGetLastError());
        P.list = e->d_list;
    }
    if (e->tile32) {
        const bool collect = prop && e->max_samples > 0, learn = prop && e->learn;
        if (collect || (learn && e->net)) {
            hipLaunchKernelGGL(feature_kernel, dim3(std::min(e->p.capacity, 4096)), dim3(64), 0, e->stream, P, e->nC,
                               learn && e->net ? e->d_feat : nullptr, collect ? e->d_sX : nullptr);
            HIPCHK(hipGetLastError());
        }
        if (learn && e->net) {
            const int r = mlp_forward_device(e->net, e->d_feat, (long)e->p.capacity * e->nC, e->d_n, e->nC,
                                             e->d_pred, e->stream);
            if (r != MQE_OK) return r;
        }
        if (!collect) P.lab = nullptr;
        if (learn)
            hipLaunchKernelGGL(t32::walker32_kernel<true>, dim3(e->p.capacity), dim3(MQE_NT), t32::lds_bytes(),
                               e->stream, e->sys, P, (const float *)e->d_t32N, (const float *)e->d_t32T, flags);
        else
            hipLaunchKernelGGL(t32::walker32_kernel<false>, dim3(e->p.capacity), dim3(MQE_NT), t32::lds_bytes(),
                               e->stream, e->sys, P, (const float *)e->d_t32N, (const float *)e->d_t32T, flags);
        HIPCHK(hipGetLastError());
        if (collect) {
            hipLaunchKernelGGL(advance_samples_kernel, dim3(1), dim3(64), 0, e->stream, e->d_n, e->nC, e->d_nsamp,
                               e->max_samples);
        }
    } else if (e->chained) launch_walker<float, true>(e, P, flags);
    else if (e->fp32) launch_walker<float, false>(e, P, flags);
    else launch_walker<double, false>(e, P, flags);
    HIPCHK(hipGetLastError());
    return MQE_OK;
}

// the population size after the work queued so far
static int refresh_n(mqe_engine *e) {
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(&e->n_host, e->d_n, sizeof(int), hipMemcpyDeviceToHost));
    return MQE_OK;
}

// the population of the last propagation (recorded by its weight sum)
static int refresh_last_walkers(mqe_engine *e) {
    double n = 0.0;
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(&n, e->d_sums + 5, sizeof(double), hipMemcpyDeviceToHost));
    e->last_walkers = (int)n;
    return MQE_OK;
}

static int check_status(mqe_engine *e) {
    int st = 0;
    HIPCHK(hipMemcpyAsync(&st, e->d_status, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (st & 4) { set_error("population exceeds walker capacity"); return MQE_ECAPACITY; }
    if (st & 3) { set_error("zero PDF norm or non-finite weight"); return MQE_ENUMERIC; }
    return MQE_OK;
}

template <typename T>
static int upload_mps_typed(const double *src, long n, void **dst) {
    std::vector<T> tmp(n);
    for (long i = 0; i < n; ++i) tmp[i] = (T)src[i];
    HIPCHK(hipMalloc(dst, sizeof(T) * std::max(n, 1L)));
    HIPCHK(hipMemcpy(*dst, tmp.data(), sizeof(T) * n, hipMemcpyHostToDevice));
    return MQE_OK;
}

#define ECHK(x)                       \
    do {                              \
        int _r = (x);                 \
        if (_r != MQE_OK) return _r;  \
    } while (0)

template <typename T>
static void launch_copy(mqe_engine *e) {
    Book B = make_book(e);
    Pool P = make_pool(e, nullptr);
    hipLaunchKernelGGL(copy_walkers_kernel<T>, dim3(2048), dim3(MQE_NT), 0, e->stream, B, P, e->L);
}

// population control with the step word of the control block
static int enqueue_popcontrol(mqe_engine *e, double scaling, const double *du, int nu) {
    Book B = make_book(e);
    hipLaunchKernelGGL(popcontrol_kernel, dim3(1), dim3(RED_NT), 0, e->stream, B, scaling, du, nu, e->p.seed);
    HIPCHK(hipGetLastError());
    if (e->fp32) launch_copy<float>(e);
    else launch_copy<double>(e);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(install_order_kernel, dim3(16), dim3(256), 0, e->stream, B);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(set_n_kernel, dim3(1), dim3(64), 0, e->stream, B, e->d_n);
    HIPCHK(hipGetLastError());
    return MQE_OK;
}

static bool tune_cache_read(const mqe_engine *e, int *tiles);

// ---- overlaps with another MPS in a caller-provided device workspace ----------
// Layout: the other MPS (engine precision), its bond dims and site offsets,
// then one transfer-matrix workspace (E, E2: Dmax^2, Y: Dmax x 2 Dmax) per
// walker of a chunk; walkers run in chunks of as many workspaces as fit.
namespace mqe {
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct OverlapPlan {
    int Dmax;
    long tot;               // elements of the other MPS
    size_t off_vd, off_off, off_ws, per_walker;
};

static int plan_overlaps(mqe_engine *e, const mqe_mps *other, OverlapPlan *pl) {
    if (!e || !other || !other->vd || !other->data || other->L != e->L || other->d != 2) {
        set_error("mqe_overlaps: bad arguments");
        return MQE_EINVAL;
    }
    const int L = e->L;
    if (other->vd[0] != 1 || other->vd[L] != 1) { set_error("mqe_overlaps: boundary bonds must be 1"); return MQE_EINVAL; }
    pl->Dmax = e->Dw;
    pl->tot = 0;
    for (int k = 0; k < L; ++k) {
        if (other->vd[k] < 1 || other->vd[k] > 4096) { set_error("mqe_overlaps: bad bond dimension"); return MQE_EINVAL; }
        pl->tot += (long)other->vd[k] * other->vd[k + 1] * 2;
        pl->Dmax = std::max(pl->Dmax, other->vd[k]);
    }
    const size_t tsz = e->fp32 ? sizeof(float) : sizeof(double);
    pl->off_vd = align256(tsz * (size_t)pl->tot);
    pl->off_off = pl->off_vd + align256(sizeof(int) * (size_t)(L + 1));
    pl->off_ws = pl->off_off + align256(sizeof(long) * (size_t)L);
    pl->per_walker = align256(tsz * 4 * (size_t)pl->Dmax * pl->Dmax);
    return MQE_OK;
}

int overlaps_workspace(mqe_engine *e, const mqe_mps *other, long *bytes_all, long *bytes_min) {
    OverlapPlan pl;
    ECHK(plan_overlaps(e, other, &pl));
    ECHK(refresh_n(e));
    if (bytes_all) *bytes_all = (long)(pl.off_ws + pl.per_walker * (size_t)std::max(e->n_host, 1));
    if (bytes_min) *bytes_min = (long)(pl.off_ws + pl.per_walker);
    return MQE_OK;
}

int overlaps_in_workspace(mqe_engine *e, const mqe_mps *other, double *out, char *ws, long ws_bytes) {
    OverlapPlan pl;
    ECHK(plan_overlaps(e, other, &pl));
    if (!out || !ws) { set_error("mqe_overlaps: bad arguments"); return MQE_EINVAL; }
    ECHK(refresh_n(e));
    const int L = e->L, n = e->n_host;
    if ((size_t)ws_bytes < pl.off_ws + pl.per_walker) {
        set_error("mqe_overlaps: the workspace cannot hold the other MPS and one walker's transfer matrices");
        return MQE_ECAPACITY;
    }
    const int chunk = (int)std::min<size_t>((size_t)std::max(n, 1), ((size_t)ws_bytes - pl.off_ws) / pl.per_walker);
    // stage the other MPS in the engine's precision, its dims and offsets
    // (fixed-size host buffers: this path allocates nothing on host or device)
    long off[MQE_MAXL];
    long t = 0;
    for (int k = 0; k < L; ++k) {
        off[k] = t;
        t += (long)other->vd[k] * other->vd[k + 1] * 2;
    }
    if (e->fp32) {
        float tmp[4096];
        for (long i0 = 0; i0 < pl.tot; i0 += 4096) {
            const long m = std::min(4096L, pl.tot - i0);
            for (long i = 0; i < m; ++i) tmp[i] = (float)other->data[i0 + i];
            HIPCHK(hipMemcpy(ws + sizeof(float) * i0, tmp, sizeof(float) * m, hipMemcpyHostToDevice));
        }
    } else {
        HIPCHK(hipMemcpy(ws, other->data, sizeof(double) * pl.tot, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(ws + pl.off_vd, other->vd, sizeof(int) * (L + 1), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ws + pl.off_off, off, sizeof(long) * L, hipMemcpyHostToDevice));
    Pool P = make_pool(e, nullptr);
    P.scratch = ws + pl.off_ws;
    P.scratch_bytes = (long)pl.per_walker;
    const int *ovd = (const int *)(ws + pl.off_vd);
    const long *ooff = (const long *)(ws + pl.off_off);
    for (int f = 0; f < n; f += chunk) {
        if (e->fp32)
            hipLaunchKernelGGL(overlap_kernel<float>, dim3(chunk), dim3(MQE_NT), 0, e->stream, L, ovd, ooff,
                               (const float *)ws, P, pl.Dmax, e->d_tmpout, f);
        else
            hipLaunchKernelGGL(overlap_kernel<double>, dim3(chunk), dim3(MQE_NT), 0, e->stream, L, ovd, ooff,
                               (const double *)ws, P, pl.Dmax, e->d_tmpout, f);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(out, e->d_tmpout, sizeof(double) * n, hipMemcpyDeviceToHost));
    return MQE_OK;
}
}  // namespace mqe

extern "C" {

const char *mqe_last_error(void) { return g_err.c_str(); }

int mqe_create(const mqe_params *p, const mqe_mps *trial, int n_couplings, const int *first, const int *second,
               const double *J, const double *tab_S, const double *tab_U, const double *tab_VT, const double *grid,
               int n_grid, int grid_dim, mqe_engine **out) {
    if (!p || !trial || !out || !grid) { set_error("mqe_create: null argument"); return MQE_EINVAL; }
    if (p->d != 2 || trial->d != 2) { set_error("mqe_create: only d = 2 is supported"); return MQE_EINVAL; }
    if (p->L < 2 || p->L > MQE_MAXL || trial->L != p->L) { set_error("mqe_create: bad chain length"); return MQE_EINVAL; }
    if (grid_dim != 4 || n_grid < 1 || n_grid > MQE_MAXG) { set_error("mqe_create: grid must have dim d*d = 4"); return MQE_EINVAL; }
    if (p->capacity < 1 || p->n_desired < 1 || p->max_bond < 1) { set_error("mqe_create: bad sizes"); return MQE_EINVAL; }
    if (p->precision != MQE_PREC_FP64 && p->precision != MQE_PREC_FP32) { set_error("mqe_create: bad precision"); return MQE_EINVAL; }
    const int L = p->L;
    for (int c = 0; c < n_couplings; ++c) {
        if (first[c] < 0 || first[c] >= second[c] || second[c] >= L) { set_error("mqe_create: bad coupling"); return MQE_EINVAL; }
        if (c > 0 && (first[c] < first[c - 1] || (first[c] == first[c - 1] && second[c] <= second[c - 1]))) {
            set_error("mqe_create: couplings must be in lexicographic order");
            return MQE_EINVAL;
        }
    }
    for (int k = 0; k < L; ++k)
        if (trial->vd[k] < 1 || trial->vd[k] > 4096) { set_error("mqe_create: bad trial bond"); return MQE_EINVAL; }
    if (trial->vd[0] != 1 || trial->vd[L] != 1) { set_error("mqe_create: trial boundary bonds must be 1"); return MQE_EINVAL; }

    mqe_engine *e = new mqe_engine();
    e->p = *p;
    e->L = L;
    e->nC = n_couplings;
    e->Ng = n_grid;
    e->gdim = grid_dim;
    e->fp32 = p->precision == MQE_PREC_FP32;
    e->tvd.assign(trial->vd, trial->vd + L + 1);
    e->Dt = *std::max_element(e->tvd.begin(), e->tvd.end());
    e->wmax.resize(L + 1);
    mqe_bond_dims(L, p->max_bond, 2, e->wmax.data());
    e->Dw = *std::max_element(e->wmax.begin(), e->wmax.end());
    e->toff.resize(L);
    e->woff.resize(L);
    e->eoff.resize(L + 1);
    long to = 0, wo = 0, eo = 0;
    for (int k = 0; k < L; ++k) {
        e->toff[k] = to;
        to += (long)e->tvd[k] * e->tvd[k + 1] * 2;
        e->woff[k] = wo;
        wo += (long)e->wmax[k] * e->wmax[k + 1] * 2;
    }
    // every environment starts on a 4-element boundary, so that the matrix-core
    // products can read it with 16-byte loads
    for (int k = 0; k <= L; ++k) {
        e->eoff[k] = eo;
        eo += ((long)e->tvd[k] * e->wmax[k] + 3) & ~3L;
    }
    e->moff.resize(L + 1);  // general path: right norm environments after the right environments
    for (int k = 0; k <= L; ++k) {
        e->moff[k] = eo;
        eo += ((long)e->wmax[k] * e->wmax[k] + 3) & ~3L;
    }
    e->wflat = (wo + 63) & ~63L;
    e->eflat = (eo + 63) & ~63L;
    const long tsize = to;
    {
        bool nn = n_couplings == L - 1;
        for (int c = 0; nn && c < n_couplings; ++c) nn = first[c] == c && second[c] == c + 1;
        const char *no32 = getenv("MQE_NO_TILE32");
        e->tile32 = e->fp32 && nn && e->Dt <= 32 && e->Dw <= 32 && p->field == 0.0 &&
                    n_grid <= t32::MAXG_T32 && !(no32 && no32[0] == '1');
        if (e->tile32) {
            e->wflat = 2L * L * 1024;
            e->eflat = (L + 1) * 1024L;  // L_k / R_k blocks
        }
    }

    auto fail = [&](int code) { mqe_destroy(e); return code; };
#define CK(x)                                                                                       \
    do {                                                                                            \
        hipError_t _e = (x);                                                                        \
        if (_e != hipSuccess) {                                                                     \
            char _b[256];                                                                           \
            snprintf(_b, sizeof(_b), "HIP error %s at %s:%d", hipGetErrorString(_e), __FILE__, __LINE__); \
            set_error(_b);                                                                          \
            return fail(MQE_EDEVICE);                                                               \
        }                                                                                           \
    } while (0)

    CK(hipStreamCreate(&e->stream));
    CK(hipEventCreate(&e->ev0));
    CK(hipEventCreate(&e->ev1));
    {
        int r = e->fp32 ? upload_mps_typed<float>(trial->data, tsize, &e->d_trial)
                        : upload_mps_typed<double>(trial->data, tsize, &e->d_trial);
        if (r) return fail(r);
    }
    {
        const char *noch = getenv("MQE_NO_CHAIN");
        e->chained = e->fp32 && !e->tile32 && e->Dt <= 64 && e->Dw <= 64 && !(noch && noch[0] == '1');
    }
    if (e->chained) {
        // the trial with every T_k[s] transposed (t_{k+1} x t_k), same offsets:
        // the chained products read trial operands along the inner index
        std::vector<float> tt((size_t)tsize);
        for (int k = 0; k < L; ++k) {
            const int dl = e->tvd[k], dr = e->tvd[k + 1];
            for (int s = 0; s < 2; ++s)
                for (int c = 0; c < dr; ++c)
                    for (int r = 0; r < dl; ++r)
                        tt[e->toff[k] + (long)dl * dr * s + c + (long)dr * r] =
                            (float)trial->data[e->toff[k] + r + (long)dl * (c + (long)dr * s)];
        }
        CK(hipMalloc((void **)&e->d_trialT, sizeof(float) * std::max<size_t>(tt.size(), 1)));
        CK(hipMemcpy(e->d_trialT, tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice));
    }
    if (e->tile32) {
        // trial in zero-padded 32x32 blocks, plain and transposed
        std::vector<float> tn((size_t)2 * L * 1024, 0.f), tt((size_t)2 * L * 1024, 0.f);
        for (int k = 0; k < L; ++k) {
            const int dl = e->tvd[k], dr = e->tvd[k + 1];
            for (int s = 0; s < 2; ++s)
                for (int c = 0; c < dr; ++c)
                    for (int r = 0; r < dl; ++r) {
                        const float v = (float)trial->data[e->toff[k] + r + (long)dl * (c + (long)dr * s)];
                        tn[(size_t)(2 * k + s) * 1024 + r + 32 * c] = v;
                        tt[(size_t)(2 * k + s) * 1024 + c + 32 * r] = v;
                    }
        }
        CK(hipMalloc((void **)&e->d_t32N, sizeof(float) * tn.size()));
        CK(hipMalloc((void **)&e->d_t32T, sizeof(float) * tt.size()));
        CK(hipMemcpy(e->d_t32N, tn.data(), sizeof(float) * tn.size(), hipMemcpyHostToDevice));
        CK(hipMemcpy(e->d_t32T, tt.data(), sizeof(float) * tt.size(), hipMemcpyHostToDevice));
        CK(hipFuncSetAttribute((const void *)t32::walker32_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)t32::lds_bytes()));
        CK(hipFuncSetAttribute((const void *)t32::walker32_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)t32::lds_bytes()));
        CK(hipMalloc((void **)&e->d_bond, sizeof(double) * (size_t)p->capacity * (n_couplings + 1)));
        CK(hipMemset(e->d_bond, 0, sizeof(double) * (size_t)p->capacity * (n_couplings + 1)));
        CK(hipMalloc((void **)&e->d_nsamp, sizeof(long)));
        CK(hipMemset(e->d_nsamp, 0, sizeof(long)));
        CK(hipMalloc((void **)&e->d_plab, sizeof(float) * (size_t)p->capacity * std::max(n_couplings, 1) * 16));
        CK(hipMemset(e->d_plab, 0, sizeof(float) * (size_t)p->capacity * std::max(n_couplings, 1) * 16));
    }
    if (!e->tile32) {
        // the general kernel runs two workgroups per CU (launch bounds); its LQ
        // workspace (Dw tau, then up to 2 Dw x Dw matrix columns) goes to LDS, as
        // much as 72 KB holds
        const size_t need = std::min(sizeof(double) * ((size_t)2 * e->Dw * e->Dw + e->Dw), (size_t)72 * 1024);
        if (need >= sizeof(double) * (size_t)(e->Dw + 2 * e->Dw)) {
            e->gen_lds_bytes = need;
            CK(hipFuncSetAttribute((const void *)walker_kernel<float, true>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
            CK(hipFuncSetAttribute((const void *)walker_kernel<float, false>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
            CK(hipFuncSetAttribute((const void *)walker_kernel<double, false>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)need));
        }
    }
    const int nC = n_couplings;
    std::vector<int> nextfirst(std::max(nC, 1)), csptr(L + 2, 0), csidx(std::max(nC, 1));
    for (int c = 0; c < nC; ++c) nextfirst[c] = (c + 1 < nC) ? first[c + 1] : L;
    for (int c = 0; c < nC; ++c) csptr[second[c] + 1]++;
    for (int k = 0; k < L; ++k) csptr[k + 1] += csptr[k];
    {
        std::vector<int> fill(csptr.begin(), csptr.end());
        for (int c = 0; c < nC; ++c) csidx[fill[second[c]]++] = c;
    }
    auto up = [&](void **dst, const void *src, size_t bytes) -> hipError_t {
        hipError_t r = hipMalloc(dst, std::max(bytes, (size_t)8));
        if (r != hipSuccess) return r;
        if (bytes) return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
        return hipSuccess;
    };
    CK(up((void **)&e->d_tvd, e->tvd.data(), sizeof(int) * (L + 1)));
    CK(up((void **)&e->d_toff, e->toff.data(), sizeof(long) * L));
    CK(up((void **)&e->d_woff, e->woff.data(), sizeof(long) * L));
    CK(up((void **)&e->d_eoff, e->eoff.data(), sizeof(long) * (L + 1)));
    CK(up((void **)&e->d_moff, e->moff.data(), sizeof(long) * (L + 1)));
    CK(up((void **)&e->d_first, first, sizeof(int) * nC));
    CK(up((void **)&e->d_second, second, sizeof(int) * nC));
    CK(up((void **)&e->d_nextfirst, nextfirst.data(), sizeof(int) * nC));
    CK(up((void **)&e->d_csptr, csptr.data(), sizeof(int) * (L + 1)));
    CK(up((void **)&e->d_csidx, csidx.data(), sizeof(int) * nC));
    CK(up((void **)&e->d_Jc, J, sizeof(double) * nC));
    CK(up((void **)&e->d_S, tab_S, sizeof(double) * 4 * nC));
    CK(up((void **)&e->d_U, tab_U, sizeof(double) * 16 * nC));
    CK(up((void **)&e->d_VT, tab_VT, sizeof(double) * 16 * nC));
    CK(up((void **)&e->d_grid, grid, sizeof(double) * 4 * n_grid));

    const int cap = p->capacity;
    const size_t tsz = e->fp32 ? sizeof(float) : sizeof(double);
    CK(hipMalloc(&e->d_phi, tsz * e->wflat * cap));
    CK(hipMalloc(&e->d_Renv, tsz * e->eflat * cap));
    CK(hipMalloc(&e->d_Lenv, tsz * e->eflat * cap));
    CK(hipMemset(e->d_phi, 0, tsz * e->wflat * cap));
    CK(hipMemset(e->d_Renv, 0, tsz * e->eflat * cap));
    CK(hipMemset(e->d_Lenv, 0, tsz * e->eflat * cap));
    CK(hipMalloc((void **)&e->d_wvd, sizeof(int) * (L + 1) * cap));
    CK(hipMalloc((void **)&e->d_w, sizeof(double) * cap));
    CK(hipMalloc((void **)&e->d_O, sizeof(double) * cap));
    CK(hipMalloc((void **)&e->d_eh, sizeof(double) * cap));
    CK(hipMalloc((void **)&e->d_EL, sizeof(double) * cap));
    CK(hipMalloc((void **)&e->d_order, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_order2, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_n, sizeof(int)));
    CK(hipMalloc((void **)&e->d_idx, sizeof(int) * std::max(nC, 1) * (size_t)cap));
    e->scratch_bytes = e->fp32 ? scratch_bytes_for<float>(e->Dt, e->Dw, n_grid)
                               : scratch_bytes_for<double>(e->Dt, e->Dw, n_grid);
    {
        // the overlap kernel needs 3 * Dmax^2 elements of T in the same scratch
        long need = 3L * 4096 * 4096;  // upper bound used only if larger than the step scratch
        (void)need;
    }
    CK(hipMalloc((void **)&e->d_scratch, (size_t)e->scratch_bytes * cap));
    CK(hipMalloc((void **)&e->d_status, sizeof(int)));
    CK(hipMemset(e->d_status, 0, sizeof(int)));
    CK(hipMalloc((void **)&e->d_sums, sizeof(double) * 8));
    CK(hipMemset(e->d_sums, 0, sizeof(double) * 8));
    e->trace_cap = 1 << 16;
    CK(hipMalloc((void **)&e->d_trace, sizeof(double) * 6 * e->trace_cap));
    CK(hipMalloc((void **)&e->d_ncopy, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_newpos, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_used, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_free, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_csrc, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_cdst, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_nextra, sizeof(int)));
    CK(hipMalloc((void **)&e->d_nnew, sizeof(int)));
    CK(hipMalloc((void **)&e->d_nused, sizeof(int)));
    CK(hipMalloc((void **)&e->d_tmpout, sizeof(double) * cap));
    CK(hipMalloc((void **)&e->d_nw, sizeof(double) * cap));
    CK(hipMalloc((void **)&e->d_list, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_key, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_since, sizeof(int) * cap));
    CK(hipMemset(e->d_since, 0, sizeof(int) * cap));
    CK(hipMalloc((void **)&e->d_ctl, sizeof(Ctl)));

    Sys &S = e->sys;
    S.L = L;
    S.nC = nC;
    S.Ng = n_grid;
    S.gdim = grid_dim;
    S.Dt = e->Dt;
    S.Dw = e->Dw;
    S.tvd = e->d_tvd;
    S.toff = e->d_toff;
    S.first = e->d_first;
    S.second = e->d_second;
    S.nextfirst = e->d_nextfirst;
    S.Jc = e->d_Jc;
    S.tS = e->d_S;
    S.tU = e->d_U;
    S.tVT = e->d_VT;
    S.grid = e->d_grid;
    S.cs_ptr = e->d_csptr;
    S.cs_idx = e->d_csidx;
    S.trialT = e->d_trialT;
    // S.S two-site tensor: Sx Sx - iSy iSy + Sz Sz (OpSx.cpp, OpISy.cpp, OpSz.cpp, d = 2)
    const double sx[4] = {0.0, 0.5, 0.5, 0.0}, isy[4] = {0.0, 0.5, -0.5, 0.0}, sz[4] = {-0.5, 0.0, 0.0, 0.5};
    for (int sp = 0; sp < 2; ++sp)
        for (int s = 0; s < 2; ++s)
            for (int tp = 0; tp < 2; ++tp)
                for (int t = 0; t < 2; ++t)
                    S.SS[(sp + 2 * tp) * 4 + (s + 2 * t)] = sx[sp + 2 * s] * sx[tp + 2 * t] -
                                                              isy[sp + 2 * s] * isy[tp + 2 * t] +
                                                              sz[sp + 2 * s] * sz[tp + 2 * t];
    S.Sz[0] = -0.5;
    S.Sz[1] = 0.5;
    S.h = p->field;
    S.ssp[0] = std::exp(0.5 * p->field * p->dtau * (-0.5));  // TrotterHeisenberg.cpp:86
    S.ssp[1] = std::exp(0.5 * p->field * p->dtau * (0.5));
    for (int q = 0; q < 8; ++q) S.tiles[q] = q == 0 ? TILE_16 : TILE_32;  // default table
    {
        int t[8];
        if (tune_cache_read(e, t)) for (int q = 0; q < 8; ++q) S.tiles[q] = t[q];  // tuned by mqe_autotune
    }
    if (push_ctl(e) != MQE_OK) return fail(MQE_EDEVICE);
    CK(hipStreamSynchronize(e->stream));
    *out = e;
#undef CK
    return MQE_OK;
}

void mqe_destroy(mqe_engine *e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
    void *ptrs[] = {e->d_t32N, e->d_t32T, e->d_trialT, e->d_trial, e->d_tvd, e->d_first, e->d_second, e->d_nextfirst, e->d_csptr, e->d_csidx,
                    e->d_toff, e->d_woff, e->d_eoff, e->d_moff, e->d_Jc, e->d_S, e->d_U, e->d_VT, e->d_grid, e->d_phi,
                    e->d_Renv, e->d_Lenv, e->d_wvd, e->d_w, e->d_O, e->d_eh, e->d_EL, e->d_order, e->d_order2,
                    e->d_n, e->d_idx, e->d_u, e->d_scratch, e->d_status, e->d_sums, e->d_trace, e->d_ncopy,
                    e->d_newpos, e->d_used, e->d_free, e->d_csrc, e->d_cdst, e->d_nextra, e->d_nnew, e->d_nused,
                    e->d_tmpout, e->d_bond, e->d_feat, e->d_pred, e->d_sX, e->d_sY, e->d_nsamp, e->d_plab,
                    e->d_nw, e->d_list, e->d_key, e->d_since, e->d_ctl};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    for (hipEvent_t ev : e->tev) (void)hipEventDestroy(ev);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

}  // extern "C"

// Tile path: the boundary environments L_0 = 1 and R_L = 1 of every slot (the
// kernel reads them from global memory and never writes them; walker copies
// carry the right environments but not the left ones), the rest zero.
static int init_tile_env(mqe_engine *e) {
    if (!e->tile32) return MQE_OK;
    const int L = e->L;
    std::vector<float> env((size_t)e->eflat * e->p.capacity, 0.f);
    for (int sl = 0; sl < e->p.capacity; ++sl) {
        env[(size_t)sl * e->eflat] = 1.f;
        env[(size_t)sl * e->eflat + (size_t)L * 1024] = 1.f;
    }
    HIPCHK(hipMemcpy(e->d_Lenv, env.data(), sizeof(float) * env.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_Renv, env.data(), sizeof(float) * env.size(), hipMemcpyHostToDevice));
    return MQE_OK;
}

extern "C" {

int mqe_load_walkers(mqe_engine *e, int n, const int *vd, const double *data, const double *w, const double *O,
                     const double *eh) {
    if (!e || n < 1 || !vd || !data || !w || !O || !eh) { set_error("mqe_load_walkers: bad arguments"); return MQE_EINVAL; }
    if (n > e->p.capacity) { set_error("mqe_load_walkers: n exceeds capacity"); return MQE_ECAPACITY; }
    const int L = e->L;
    const size_t tsz = e->fp32 ? sizeof(float) : sizeof(double);
    std::vector<char> buf(tsz * e->wflat * n, 0);
    std::vector<int> vds((size_t)(L + 1) * n);
    long off = 0;
    for (int i = 0; i < n; ++i) {
        const int *v = vd + (size_t)i * (L + 1);
        if (v[0] != 1 || v[L] != 1) { set_error("mqe_load_walkers: boundary bonds must be 1"); return MQE_EINVAL; }
        for (int k = 0; k < L; ++k) {
            if (v[k] < 1 || v[k] > e->wmax[k] || v[k + 1] > e->wmax[k + 1] || v[k] > 2 * v[k + 1]) {
                set_error("mqe_load_walkers: walker bond dims exceed max_bond profile or dimL > d*dimR");
                return MQE_EINVAL;
            }
        }
        for (int k = 0; k <= L; ++k) vds[(size_t)i * (L + 1) + k] = v[k];
        for (int k = 0; k < L; ++k) {
            const long sz = (long)v[k] * v[k + 1] * 2;
            if (e->tile32) {
                float *dst = (float *)buf.data() + (size_t)i * e->wflat + (size_t)2 * k * 1024;
                const int dl = v[k], dr = v[k + 1];
                for (int s2 = 0; s2 < 2; ++s2)
                    for (int c = 0; c < dr; ++c)
                        for (int r = 0; r < dl; ++r)
                            dst[(size_t)s2 * 1024 + r + 32 * c] = (float)data[off + r + (long)dl * (c + (long)dr * s2)];
            } else {
                char *dst = buf.data() + tsz * ((size_t)i * e->wflat + e->woff[k]);
                for (long q = 0; q < sz; ++q) {
                    if (e->fp32) ((float *)dst)[q] = (float)data[off + q];
                    else ((double *)dst)[q] = data[off + q];
                }
            }
            off += sz;
        }
    }
    ECHK(init_tile_env(e));
    HIPCHK(hipMemcpy(e->d_phi, buf.data(), buf.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_wvd, vds.data(), sizeof(int) * vds.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_w, w, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_O, O, sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_eh, eh, sizeof(double) * n, hipMemcpyHostToDevice));
    std::vector<int> ord(e->p.capacity);
    for (int i = 0; i < e->p.capacity; ++i) ord[i] = i;
    HIPCHK(hipMemcpy(e->d_order, ord.data(), sizeof(int) * ord.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_n, &n, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(e->d_status, 0, sizeof(int)));
    e->n_host = n;
    e->loaded = true;
    e->prepared = false;
    return MQE_OK;
}

static int ensure_u(mqe_engine *e, long need) {
    if (need <= e->u_cap) return MQE_OK;
    if (e->d_u) HIPCHK(hipFree(e->d_u));
    e->d_u = nullptr;
    HIPCHK(hipMalloc((void **)&e->d_u, sizeof(double) * std::max(need, 1L)));
    e->u_cap = need;
    return MQE_OK;
}

int mqe_prepare(mqe_engine *e, double *E_start) {
    if (!e || !e->loaded) { set_error("mqe_prepare: no walkers loaded"); return MQE_ESTATE; }
    ECHK(launch_walker_any(e, nullptr, F_PREP));
    Book B = make_book(e);
    hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(RED_NT), 0, e->stream, B, 1, 0, -1, 0);
    HIPCHK(hipGetLastError());
    double sums[8];
    HIPCHK(hipMemcpyAsync(sums, e->d_sums, sizeof(sums), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (E_start) *E_start = sums[3];
    e->prepared = true;
    return check_status(e);
}

int mqe_propagate(mqe_engine *e, const double *uniforms, double *sum_w) {
    if (!e || !e->prepared) { set_error("mqe_propagate: call mqe_prepare first"); return MQE_ESTATE; }
    const double *du = nullptr;
    if (uniforms) {
        // the uniforms are sized by the current population: refresh it first
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipMemcpy(&e->n_host, e->d_n, sizeof(int), hipMemcpyDeviceToHost));
        const long need = (long)e->n_host * std::max(e->nC, 1);
        ECHK(ensure_u(e, need));
        HIPCHK(hipMemcpyAsync(e->d_u, uniforms, sizeof(double) * (size_t)e->n_host * e->nC, hipMemcpyHostToDevice,
                              e->stream));
        du = e->d_u;
    } else {
        // device draws: the k-th such call uses step word k, as step k of mqe_run does
        e->ctl_host.step = (int)++e->prop_draws;
    }
    e->ctl_host.pc = e->prop_count;
    ECHK(push_ctl(e));
    HIPCHK(hipEventRecord(e->ev0, e->stream));
    ECHK(launch_walker_any(e, du, F_PROP));
    HIPCHK(hipEventRecord(e->ev1, e->stream));
    Book B = make_book(e);
    hipLaunchKernelGGL(sumw_kernel, dim3(1), dim3(RED_NT), 0, e->stream, B);
    HIPCHK(hipGetLastError());
    double sums[8];
    HIPCHK(hipMemcpyAsync(sums, e->d_sums, sizeof(sums), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipEventElapsedTime(&e->last_ms, e->ev0, e->ev1));
    e->last_walkers = (int)sums[5];
    if (sum_w) *sum_w = sums[0];
    return check_status(e);
}

int mqe_energy(mqe_engine *e, int step, int do_fluct, double *E_P, double *omega) {
    if (!e || !e->prepared) { set_error("mqe_energy: call mqe_prepare first"); return MQE_ESTATE; }
    if (step < 1) { set_error("mqe_energy: step must be >= 1"); return MQE_EINVAL; }
    Book B = make_book(e);
    hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(RED_NT), 0, e->stream, B, step, do_fluct ? 1 : 0, -1, 0);
    HIPCHK(hipGetLastError());
    double sums[8];
    HIPCHK(hipMemcpyAsync(sums, e->d_sums, sizeof(sums), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (E_P) *E_P = sums[3];
    if (omega) *omega = sums[4];
    return MQE_OK;
}

int mqe_population_control(mqe_engine *e, double scaling, const double *uniforms, int n_uniforms, int *n_new,
                           int *n_used) {
    if (!e || !e->prepared) { set_error("mqe_population_control: call mqe_prepare first"); return MQE_ESTATE; }
    if (!(scaling > 0.0)) { set_error("mqe_population_control: scaling must be > 0"); return MQE_EINVAL; }
    const double *du = nullptr;
    if (uniforms && n_uniforms > 0) {
        ECHK(ensure_u(e, n_uniforms));
        HIPCHK(hipMemcpyAsync(e->d_u, uniforms, sizeof(double) * n_uniforms, hipMemcpyHostToDevice, e->stream));
        du = e->d_u;
    } else if (!uniforms) {
        // device draws: the k-th such call uses step word k, as step k of mqe_run does
        e->ctl_host.pstep = (int)++e->pop_draws;
        ECHK(push_ctl(e));
    }
    ECHK(enqueue_popcontrol(e, scaling, du, uniforms ? n_uniforms : 0));
    int nn = 0, nu = 0;
    HIPCHK(hipMemcpyAsync(&nn, e->d_nnew, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(&nu, e->d_nused, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (n_used) *n_used = nu;
    if (nn > e->p.capacity) {
        set_error("population exceeds walker capacity (the population is left as it was before this call)");
        return MQE_ECAPACITY;
    }
    e->n_host = nn;
    if (n_new) *n_new = nn;
    return MQE_OK;
}

// One Walk step.  Eager: the control block is set to this step; in a graph
// capture (captured = true) a kernel advances it instead, so the captured
// launches carry no step-dependent argument.
static int enqueue_step(mqe_engine *e, int step, bool captured) {
    if (captured) {
        hipLaunchKernelGGL(advance_ctl_kernel, dim3(1), dim3(64), 0, e->stream, e->d_ctl);
        HIPCHK(hipGetLastError());
    } else {
        e->ctl_host.step = e->ctl_host.pstep = step;
        e->ctl_host.pc = e->prop_count;
        e->prop_draws = e->pop_draws = step;
        ECHK(push_ctl(e));
    }
    const bool timed = e->timing && !captured;
    if (timed) {
        if (e->tev_used + 2 > (int)e->tev.size()) {
            for (int q = 0; q < 64; ++q) {
                hipEvent_t ev;
                HIPCHK(hipEventCreate(&ev));
                e->tev.push_back(ev);
            }
        }
        HIPCHK(hipEventRecord(e->tev[e->tev_used], e->stream));
    }
    ECHK(launch_walker_any(e, nullptr, F_PROP));
    if (timed) {
        HIPCHK(hipEventRecord(e->tev[e->tev_used + 1], e->stream));
        e->tev_used += 2;
    }
    Book B = make_book(e);
    hipLaunchKernelGGL(sumw_kernel, dim3(1), dim3(RED_NT), 0, e->stream, B);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(energy_kernel, dim3(1), dim3(RED_NT), 0, e->stream, B, step, 1, -1, 1);
    HIPCHK(hipGetLastError());
    ECHK(enqueue_popcontrol(e, -1.0, nullptr, 0));
    return MQE_OK;
}

int mqe_step(mqe_engine *e, int step) {
    if (!e || !e->prepared) { set_error("mqe_step: call mqe_prepare first"); return MQE_ESTATE; }
    if (step < 1) { set_error("mqe_step: step must be >= 1"); return MQE_EINVAL; }
    return enqueue_step(e, step, false);
}

// Capture n_steps steps once (kept until a run with another count) and replay
// them: the control block is set to the step before first_step, and each
// captured step advances it.
static int run_graph(mqe_engine *e, int first_step, int n_steps) {
    if (e->learn || e->max_samples > 0) {
        set_error("mqe_run: graph replay runs the reference proposal without sample collection");
        return MQE_EINVAL;
    }
    if (!e->graph_exec || e->graph_steps != n_steps) {
        if (e->graph_exec) HIPCHK(hipGraphExecDestroy(e->graph_exec));
        e->graph_exec = nullptr;
        e->graph_steps = 0;
        HIPCHK(hipStreamSynchronize(e->stream));
        const long pc0 = e->prop_count;
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int r = MQE_OK;
        for (int s = 0; s < n_steps && r == MQE_OK; ++s) r = enqueue_step(e, 0, true);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(e->stream, &g);
        e->prop_count = pc0;
        if (r != MQE_OK || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (r == MQE_OK) HIPCHK(ce);
            return r;
        }
        const hipError_t ie = hipGraphInstantiate(&e->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ie);
        e->graph_steps = n_steps;
    }
    e->ctl_host.step = e->ctl_host.pstep = first_step - 1;
    e->ctl_host.pc = e->prop_count - 1;
    ECHK(push_ctl(e));
    HIPCHK(hipGraphLaunch(e->graph_exec, e->stream));
    e->prop_count += n_steps;
    e->prop_draws = e->pop_draws = first_step + n_steps - 1;
    e->ctl_host.step = e->ctl_host.pstep = first_step + n_steps - 1;
    e->ctl_host.pc = e->prop_count - 1;
    return MQE_OK;
}

int mqe_run(mqe_engine *e, int first_step, int n_steps, int use_graph) {
    if (!e || !e->prepared) { set_error("mqe_run: call mqe_prepare first"); return MQE_ESTATE; }
    if (first_step < 1 || n_steps < 0) { set_error("mqe_run: bad step range"); return MQE_EINVAL; }
    if (n_steps == 0) return MQE_OK;
    if (use_graph) return run_graph(e, first_step, n_steps);
    for (int s = 0; s < n_steps; ++s) ECHK(enqueue_step(e, first_step + s, false));
    return MQE_OK;
}

int mqe_step_graph(mqe_engine *e, int first_step, int n_steps) { return mqe_run(e, first_step, n_steps, 1); }

int mqe_read_trace(mqe_engine *e, int first_step, int n_steps, double *out) {
    if (!e || !out || first_step < 1 || n_steps < 0 || n_steps > e->trace_cap) { set_error("mqe_read_trace: bad arguments"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int s = 0; s < n_steps; ++s) {
        const int row = (first_step - 1 + s) % e->trace_cap;
        HIPCHK(hipMemcpy(out + 6L * s, e->d_trace + 6L * row, sizeof(double) * 6, hipMemcpyDeviceToHost));
    }
    return MQE_OK;
}

int mqe_kernel_path(mqe_engine *e) { return (e && e->tile32) ? 1 : 0; }

}  // extern "C"

// ---- tile-shape autotuner -----------------------------------------------------
// Key of the tuned table: the product shapes depend on the trial and walker
// bonds, the chain length, the couplings and the precision.
static std::string tune_key(const mqe_engine *e) {
    char b[160];
    snprintf(b, sizeof(b), "L%d_nC%d_Dt%d_Dw%d_d%d_%s", e->L, e->nC, e->Dt, e->Dw, e->p.d, e->fp32 ? "fp32" : "fp64");
    return b;
}

static std::string tune_cache_path() {
    const char *p = getenv("MQE_TUNE_CACHE");
    if (p && p[0]) return p;
    const char *h = getenv("HOME");
    return std::string(h && h[0] ? h : "/tmp") + "/.cache/mqe_tiles.txt";
}

// the cached table for this engine's key, if any (lines "key t0 ... t7")
static bool tune_cache_read(const mqe_engine *e, int *tiles) {
    FILE *f = std::fopen(tune_cache_path().c_str(), "r");
    if (!f) return false;
    const std::string key = tune_key(e);
    char k[160];
    int t[8];
    bool found = false;
    while (std::fscanf(f, "%159s %d %d %d %d %d %d %d %d", k, &t[0], &t[1], &t[2], &t[3], &t[4], &t[5], &t[6],
                       &t[7]) == 9)
        if (key == k) {
            found = true;
            for (int q = 0; q < 8; ++q) tiles[q] = t[q];
        }
    std::fclose(f);
    for (int q = 0; found && q < 8; ++q)
        if (tiles[q] != TILE_16 && tiles[q] != TILE_32 && tiles[q] != TILE_WIDE) found = false;
    return found;
}

static void tune_cache_write(const mqe_engine *e, const int *tiles) {
    const std::string path = tune_cache_path();
    const size_t sl = path.rfind('/');
    if (sl != std::string::npos && sl > 0) (void)mkdir(path.substr(0, sl).c_str(), 0755);  // EEXIST is fine
    FILE *f = std::fopen(path.c_str(), "a");
    if (!f) return;  // a read-only home: the table stays in this engine only
    std::fprintf(f, "%s", tune_key(e).c_str());
    for (int q = 0; q < 8; ++q) std::fprintf(f, " %d", tiles[q]);
    std::fprintf(f, "\n");
    std::fclose(f);
}

static void set_tiles(mqe_engine *e, const int *tiles) {
    for (int q = 0; q < 8; ++q) e->sys.tiles[q] = tiles[q];
    if (e->graph_exec) {  // captured launches carry the old table
        (void)hipGraphExecDestroy(e->graph_exec);
        e->graph_exec = nullptr;
        e->graph_steps = 0;
    }
}

extern "C" {

int mqe_get_tiles(mqe_engine *e, int *tiles) {
    if (!e || !tiles) { set_error("mqe_get_tiles: bad arguments"); return MQE_EINVAL; }
    for (int q = 0; q < 8; ++q) tiles[q] = e->sys.tiles[q];
    return MQE_OK;
}

int mqe_set_tiles(mqe_engine *e, const int *tiles) {
    if (!e || !tiles) { set_error("mqe_set_tiles: bad arguments"); return MQE_EINVAL; }
    for (int q = 0; q < 8; ++q)
        if (tiles[q] != TILE_16 && tiles[q] != TILE_32 && tiles[q] != TILE_WIDE) {
            set_error("mqe_set_tiles: tile shapes are 16, 32 or 64");
            return MQE_EINVAL;
        }
    HIPCHK(hipStreamSynchronize(e->stream));
    set_tiles(e, tiles);
    return MQE_OK;
}

int mqe_autotune(mqe_engine *e, int reps, int *tiles_out, double *ms_default, double *ms_tuned) {
    if (!e || reps < 1) { set_error("mqe_autotune: bad arguments"); return MQE_EINVAL; }
    if (!e->prepared) { set_error("mqe_autotune: call mqe_prepare first"); return MQE_ESTATE; }
    int best[8];
    for (int q = 0; q < 8; ++q) best[q] = e->sys.tiles[q];
    if (!e->fp32 || e->tile32) {  // only the general path's FP32 products have tile shapes
        if (tiles_out) for (int q = 0; q < 8; ++q) tiles_out[q] = best[q];
        if (ms_default) *ms_default = 0.0;
        if (ms_tuned) *ms_tuned = 0.0;
        return MQE_OK;
    }
    ECHK(refresh_n(e));
    const int cap = e->p.capacity, L = e->L;
    // buckets present in the population
    std::vector<int> order(cap), vd((size_t)cap * (L + 1));
    HIPCHK(hipMemcpy(order.data(), e->d_order, sizeof(int) * cap, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(vd.data(), e->d_wvd, sizeof(int) * vd.size(), hipMemcpyDeviceToHost));
    bool present[8] = {false};
    for (int b = 0; b < e->n_host; ++b) {
        int mb = 1;
        for (int k = 0; k <= L; ++k) mb = std::max(mb, vd[(size_t)order[b] * (L + 1) + k]);
        present[bond_bucket(mb)] = true;
    }
    // snapshot of everything a propagation writes
    const size_t tsz = sizeof(float);
    struct Part { void *dev; size_t bytes; void *save; };
    std::vector<Part> parts = {{e->d_phi, tsz * e->wflat * cap, nullptr},  {e->d_Renv, tsz * e->eflat * cap, nullptr},
                               {e->d_Lenv, tsz * e->eflat * cap, nullptr}, {e->d_wvd, sizeof(int) * (L + 1) * cap, nullptr},
                               {e->d_w, sizeof(double) * cap, nullptr},     {e->d_O, sizeof(double) * cap, nullptr},
                               {e->d_eh, sizeof(double) * cap, nullptr},    {e->d_EL, sizeof(double) * cap, nullptr},
                               {e->d_order, sizeof(int) * cap, nullptr},    {e->d_n, sizeof(int), nullptr},
                               {e->d_since, sizeof(int) * cap, nullptr},
                               {e->d_status, sizeof(int), nullptr}};
    int rc = MQE_OK;
    for (auto &pt : parts)
        if (rc == MQE_OK) {
            if (hipMalloc(&pt.save, pt.bytes) != hipSuccess ||
                hipMemcpyAsync(pt.save, pt.dev, pt.bytes, hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
                rc = MQE_EDEVICE;
        }
    const Ctl ctl0 = e->ctl_host;
    const long pc0 = e->prop_count;
    auto restore = [&]() -> int {
        for (auto &pt : parts)
            if (pt.save && hipMemcpyAsync(pt.dev, pt.save, pt.bytes, hipMemcpyDeviceToDevice, e->stream) != hipSuccess)
                return MQE_EDEVICE;
        e->ctl_host = ctl0;
        e->prop_count = pc0;
        return MQE_OK;
    };
    // device time of `reps` propagations with the table t (state restored afterwards)
    auto trial = [&](const int *t, double *ms) -> int {
        set_tiles(e, t);
        float tot = 0.f;
        for (int r = 0; r < reps; ++r) {
            e->ctl_host.step = ctl0.step + 1 + r;
            e->ctl_host.pc = e->prop_count;
            int q = push_ctl(e);
            if (q != MQE_OK) return q;
            if (hipEventRecord(e->ev0, e->stream) != hipSuccess) return MQE_EDEVICE;
            q = launch_walker_any(e, nullptr, F_PROP);
            if (q != MQE_OK) return q;
            if (hipEventRecord(e->ev1, e->stream) != hipSuccess || hipEventSynchronize(e->ev1) != hipSuccess)
                return MQE_EDEVICE;
            float x = 0.f;
            (void)hipEventElapsedTime(&x, e->ev0, e->ev1);
            tot += x;
        }
        *ms = tot / reps;
        return restore();
    };
    double t_default = 0.0, t_best = 0.0;
    if (rc == MQE_OK) rc = trial(best, &t_default);  // also warms the code up
    if (rc == MQE_OK) rc = trial(best, &t_default);
    t_best = t_default;
    // coordinate descent over the buckets present, largest bonds first
    const int cands[3] = {TILE_16, TILE_32, TILE_WIDE};
    for (int bk = 7; rc == MQE_OK && bk >= 0; --bk) {
        if (!present[bk]) continue;
        for (int c : cands) {
            if (c == best[bk] || rc != MQE_OK) continue;
            int t[8];
            for (int q = 0; q < 8; ++q) t[q] = best[q];
            t[bk] = c;
            double ms = 0.0;
            rc = trial(t, &ms);
            if (rc == MQE_OK && ms < 0.98 * t_best) {  // keep a change only for a clear gain
                t_best = ms;
                best[bk] = c;
            }
        }
    }
    if (rc == MQE_OK) (void)restore();
    for (auto &pt : parts)
        if (pt.save) (void)hipFree(pt.save);
    HIPCHK(hipStreamSynchronize(e->stream));
    if (rc != MQE_OK) {
        set_error("mqe_autotune: device error while timing a tile table");
        return rc;
    }
    set_tiles(e, best);
    tune_cache_write(e, best);
    if (tiles_out) for (int q = 0; q < 8; ++q) tiles_out[q] = best[q];
    if (ms_default) *ms_default = t_default;
    if (ms_tuned) *ms_tuned = t_best;
    return MQE_OK;
}

int mqe_set_canonical_interval(mqe_engine *e, int steps) {
    if (!e || steps < 1) { set_error("mqe_set_canonical_interval: bad arguments"); return MQE_EINVAL; }
    e->canon_every = steps;
    e->prop_count = 0;
    return MQE_OK;
}

int mqe_sync(mqe_engine *e) {
    if (!e) { set_error("mqe_sync: null engine"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    int n = 0;
    HIPCHK(hipMemcpy(&n, e->d_n, sizeof(int), hipMemcpyDeviceToHost));
    e->n_host = n;
    return check_status(e);
}

int mqe_num_walkers(mqe_engine *e, int *n) {
    if (!e || !n) { set_error("mqe_num_walkers: bad arguments"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(n, e->d_n, sizeof(int), hipMemcpyDeviceToHost));
    e->n_host = *n;
    return MQE_OK;
}

int mqe_num_couplings(mqe_engine *e, int *n) {
    if (!e || !n) { set_error("mqe_num_couplings: bad arguments"); return MQE_EINVAL; }
    *n = e->nC;
    return MQE_OK;
}

int mqe_read_scalars(mqe_engine *e, double *w, double *O, double *eh, double *EL) {
    if (!e) { set_error("mqe_read_scalars: null engine"); return MQE_EINVAL; }
    ECHK(refresh_n(e));
    const int n = e->n_host, cap = e->p.capacity;
    std::vector<int> ord(n);
    HIPCHK(hipMemcpy(ord.data(), e->d_order, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<double> tmp(cap);
    double *dst[4] = {w, O, eh, EL};
    double *src[4] = {e->d_w, e->d_O, e->d_eh, e->d_EL};
    for (int q = 0; q < 4; ++q) {
        if (!dst[q]) continue;
        HIPCHK(hipMemcpy(tmp.data(), src[q], sizeof(double) * cap, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) dst[q][i] = tmp[ord[i]];
    }
    return MQE_OK;
}

int mqe_read_indices(mqe_engine *e, int *idx) {
    if (!e || !idx) { set_error("mqe_read_indices: bad arguments"); return MQE_EINVAL; }
    ECHK(refresh_last_walkers(e));
    HIPCHK(hipMemcpy(idx, e->d_idx, sizeof(int) * (size_t)e->last_walkers * e->nC, hipMemcpyDeviceToHost));
    return MQE_OK;
}

int mqe_read_walker(mqe_engine *e, int i, int *vd, double *data, long *size) {
    if (!e || !vd) { set_error("mqe_read_walker: bad arguments"); return MQE_EINVAL; }
    ECHK(refresh_n(e));
    if (i < 0 || i >= e->n_host) { set_error("mqe_read_walker: walker index out of range"); return MQE_EINVAL; }
    const int L = e->L;
    int slot = 0;
    HIPCHK(hipMemcpy(&slot, e->d_order + i, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(vd, e->d_wvd + (size_t)slot * (L + 1), sizeof(int) * (L + 1), hipMemcpyDeviceToHost));
    long n = 0;
    for (int k = 0; k < L; ++k) n += (long)vd[k] * vd[k + 1] * 2;
    if (size) *size = n;
    if (!data) return MQE_OK;
    const size_t tsz = e->fp32 ? sizeof(float) : sizeof(double);
    std::vector<char> buf(tsz * e->wflat);
    HIPCHK(hipMemcpy(buf.data(), (char *)e->d_phi + tsz * (size_t)slot * e->wflat, buf.size(), hipMemcpyDeviceToHost));
    long off = 0;
    for (int k = 0; k < L; ++k) {
        const long sz = (long)vd[k] * vd[k + 1] * 2;
        if (e->tile32) {
            const float *src = (const float *)buf.data() + (size_t)2 * k * 1024;
            const int dl = vd[k], dr = vd[k + 1];
            for (int s2 = 0; s2 < 2; ++s2)
                for (int c = 0; c < dr; ++c)
                    for (int r = 0; r < dl; ++r)
                        data[off + r + (long)dl * (c + (long)dr * s2)] = (double)src[(size_t)s2 * 1024 + r + 32 * c];
        } else {
            const char *src = buf.data() + tsz * e->woff[k];
            for (long q = 0; q < sz; ++q) data[off + q] = e->fp32 ? (double)((const float *)src)[q] : ((const double *)src)[q];
        }
        off += sz;
    }
    return MQE_OK;
}

int mqe_overlaps(mqe_engine *e, const mqe_mps *other, double *out) {
    long need = 0;
    ECHK(mqe::overlaps_workspace(e, other, &need, nullptr));
    char *scr = nullptr;
    HIPCHK(hipMalloc((void **)&scr, (size_t)need));
    const int r = mqe::overlaps_in_workspace(e, other, out, scr, need);
    (void)hipFree(scr);
    return r;
}

int mqe_compress_walkers(mqe_engine *e, const int *targets) {
    if (!e || !targets) { set_error("mqe_compress_walkers: bad arguments"); return MQE_EINVAL; }
    if (!e->loaded) { set_error("mqe_compress_walkers: no walkers loaded"); return MQE_ESTATE; }
    ECHK(refresh_n(e));
    const int n = e->n_host, L = e->L;
    const int Dmax = e->Dw;
    // per walker: X and V ((2 Dmax)^2 doubles each), Schmidt values and their order
    const long scr_bytes = ((8L * (2 * 4L * Dmax * Dmax + 2 * Dmax) + 4L * 2 * Dmax) + 255) & ~255L;
    const int chunk = std::max(1, std::min(n, (int)std::min(1024L, (1L << 30) / scr_bytes)));
    char *scr = nullptr;
    int *dt = nullptr;
    HIPCHK(hipMalloc((void **)&scr, (size_t)scr_bytes * chunk));
    HIPCHK(hipMalloc((void **)&dt, sizeof(int) * n));
    HIPCHK(hipMemcpy(dt, targets, sizeof(int) * n, hipMemcpyHostToDevice));
    Pool P = make_pool(e, nullptr);
    for (int f = 0; f < n; f += chunk) {
        if (e->fp32)
            hipLaunchKernelGGL(compress_kernel<float>, dim3(chunk), dim3(MQE_NT), 0, e->stream, L, P, f, dt, scr,
                               scr_bytes, Dmax);
        else
            hipLaunchKernelGGL(compress_kernel<double>, dim3(chunk), dim3(MQE_NT), 0, e->stream, L, P, f, dt, scr,
                               scr_bytes, Dmax);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    (void)hipFree(scr);
    // overlaps with the trial (transfer matrices), then the sign convention
    const int Do = std::max(e->Dw, e->Dt);
    const size_t tsz = e->fp32 ? sizeof(float) : sizeof(double);
    const size_t need = tsz * 4 * (size_t)Do * Do;
    HIPCHK(hipMalloc((void **)&scr, need * std::max(n, 1)));
    P.scratch = scr;
    P.scratch_bytes = (long)need;
    if (e->fp32) {
        hipLaunchKernelGGL(overlap_kernel<float>, dim3(n), dim3(MQE_NT), 0, e->stream, L, e->d_tvd, e->d_toff,
                           (const float *)e->d_trial, P, Do, e->d_tmpout, 0);
        hipLaunchKernelGGL(overlap_sign_kernel<float>, dim3(n), dim3(64), 0, e->stream, L, P, e->d_tmpout, dt);
    } else {
        hipLaunchKernelGGL(overlap_kernel<double>, dim3(n), dim3(MQE_NT), 0, e->stream, L, e->d_tvd, e->d_toff,
                           (const double *)e->d_trial, P, Do, e->d_tmpout, 0);
        hipLaunchKernelGGL(overlap_sign_kernel<double>, dim3(n), dim3(64), 0, e->stream, L, P, e->d_tmpout, dt);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e->stream));
    (void)hipFree(scr);
    (void)hipFree(dt);
    e->prepared = false;
    return MQE_OK;
}

int mqe_set_timing(mqe_engine *e, int on) {
    if (!e) { set_error("mqe_set_timing: null engine"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    e->timing = on != 0;
    e->tev_used = 0;
    return MQE_OK;
}

int mqe_kernel_time(mqe_engine *e, double *total_ms, int *launches) {
    if (!e) { set_error("mqe_kernel_time: null engine"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    double tot = 0.0;
    for (int q = 0; q + 1 < e->tev_used; q += 2) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e->tev[q], e->tev[q + 1]));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = e->tev_used / 2;
    return MQE_OK;
}

int mqe_last_kernel_ms(mqe_engine *e, double *ms, int *walkers) {
    if (!e) { set_error("mqe_last_kernel_ms: null engine"); return MQE_EINVAL; }
    ECHK(refresh_last_walkers(e));
    if (ms) *ms = e->last_ms;
    if (walkers) *walkers = e->last_walkers;
    return MQE_OK;
}

// ---- population snapshot ------------------------------------------------------
// File: header (magic, sizes, counters), then per walker in population order its
// slot, tensors, right and left environments (as stored, in the engine's gauge),
// bond dims, the four scalars and, on the tile path, its bond energies and
// recorded labels.  Walkers go back to their slots, so the free slots, the
// staggered re-orthonormalisation and the device random numbers continue exactly.
struct StateHeader {
    char magic[8];
    int L, d, precision, tile32;
    long wflat, eflat;
    int n, prepared;
    long prop_count, step;
    int capacity, nC;
    long prop_draws, pop_draws;
    int canon_every, pad;
};

// ---- learned proposal -------------------------------------------------------

static int need_tile(mqe_engine *e, const char *who) {
    if (!e) { set_error((std::string(who) + ": null engine").c_str()); return MQE_EINVAL; }
    if (!e->tile32) {
        set_error((std::string(who) + ": the learned proposal runs on the tile path only "
                   "(nearest-neighbour chain, bonds <= 32, FP32, h = 0)").c_str());
        return MQE_EINVAL;
    }
    return MQE_OK;
}

int mqe_collect_samples(mqe_engine *e, long max_samples) {
    ECHK(need_tile(e, "mqe_collect_samples"));
    if (max_samples < 0) { set_error("mqe_collect_samples: negative size"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    if (e->d_sX) HIPCHK(hipFree(e->d_sX));
    if (e->d_sY) HIPCHK(hipFree(e->d_sY));
    e->d_sX = e->d_sY = nullptr;
    e->max_samples = 0;
    HIPCHK(hipMemset(e->d_nsamp, 0, sizeof(long)));
    if (max_samples == 0) return MQE_OK;
    HIPCHK(hipMalloc((void **)&e->d_sX, sizeof(float) * (size_t)max_samples * MQE_NFEAT));
    HIPCHK(hipMalloc((void **)&e->d_sY, sizeof(float) * (size_t)max_samples * MQE_NLABEL));
    e->max_samples = max_samples;
    return MQE_OK;
}

int mqe_num_samples(mqe_engine *e, long *n) {
    ECHK(need_tile(e, "mqe_num_samples"));
    if (!n) { set_error("mqe_num_samples: null argument"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(n, e->d_nsamp, sizeof(long), hipMemcpyDeviceToHost));
    return MQE_OK;
}

int mqe_read_samples(mqe_engine *e, long first, long n, float *X, float *Y) {
    long have = 0;
    ECHK(mqe_num_samples(e, &have));
    if (first < 0 || n < 0 || first + n > have) { set_error("mqe_read_samples: rows out of range"); return MQE_EINVAL; }
    if (X) HIPCHK(hipMemcpy(X, e->d_sX + first * MQE_NFEAT, sizeof(float) * n * MQE_NFEAT, hipMemcpyDeviceToHost));
    if (Y) HIPCHK(hipMemcpy(Y, e->d_sY + first * MQE_NLABEL, sizeof(float) * n * MQE_NLABEL, hipMemcpyDeviceToHost));
    return MQE_OK;
}

static int check_net(mqe_mlp *m, const char *who) {
    int nin = 0, nout = 0;
    mlp_io_dims(m, &nin, &nout);
    if (nin != MQE_NFEAT || nout != MQE_NLABEL) {
        set_error((std::string(who) + ": the network must map MQE_NFEAT features to MQE_NLABEL labels").c_str());
        return MQE_EINVAL;
    }
    return MQE_OK;
}

int mqe_train_proposal(mqe_engine *e, mqe_mlp *m, int epochs, int batch, double lr, double *loss) {
    ECHK(need_tile(e, "mqe_train_proposal"));
    if (!m) { set_error("mqe_train_proposal: null network"); return MQE_EINVAL; }
    ECHK(check_net(m, "mqe_train_proposal"));
    long n = 0;
    ECHK(mqe_num_samples(e, &n));
    if (n < 1) { set_error("mqe_train_proposal: no samples collected"); return MQE_ESTATE; }
    return mlp_fit_device(m, e->d_sX, e->d_sY, n, epochs, batch, lr, 1, loss, e->stream);
}

static int ensure_pred(mqe_engine *e) {
    if (e->d_pred) return MQE_OK;
    const size_t rows = (size_t)e->p.capacity * e->nC;
    HIPCHK(hipMalloc((void **)&e->d_feat, sizeof(float) * rows * MQE_NFEAT));
    HIPCHK(hipMalloc((void **)&e->d_pred, sizeof(float) * rows * MQE_NLABEL));
    return MQE_OK;
}

int mqe_set_proposal(mqe_engine *e, mqe_mlp *m, double mix) {
    ECHK(need_tile(e, "mqe_set_proposal"));
    if (!(mix >= 0.0 && mix <= 1.0)) { set_error("mqe_set_proposal: mix must lie in [0, 1]"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    if (!m) {
        e->learn = false;
        e->net = nullptr;
        return MQE_OK;
    }
    ECHK(check_net(m, "mqe_set_proposal"));
    ECHK(ensure_pred(e));
    e->net = m;
    e->learn = true;
    e->mix = mix;
    return MQE_OK;
}

int mqe_set_proposal_table(mqe_engine *e, const float *pred, int n_walkers, double mix) {
    ECHK(need_tile(e, "mqe_set_proposal_table"));
    if (!(mix >= 0.0 && mix <= 1.0)) { set_error("mqe_set_proposal_table: mix must lie in [0, 1]"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    e->net = nullptr;
    if (!pred) {
        e->learn = false;
        return MQE_OK;
    }
    if (n_walkers < 0 || n_walkers > e->p.capacity) { set_error("mqe_set_proposal_table: bad walker count"); return MQE_EINVAL; }
    ECHK(ensure_pred(e));
    // rows beyond the table predict 0: their proposal is the uniform one
    const size_t rows = (size_t)e->p.capacity * e->nC, given = (size_t)n_walkers * e->nC;
    HIPCHK(hipMemset(e->d_pred, 0, sizeof(float) * rows * MQE_NLABEL));
    HIPCHK(hipMemcpy(e->d_pred, pred, sizeof(float) * given * MQE_NLABEL, hipMemcpyHostToDevice));
    e->learn = true;
    e->mix = mix;
    return MQE_OK;
}

int mqe_read_bond_energies(mqe_engine *e, double *out) {
    ECHK(need_tile(e, "mqe_read_bond_energies"));
    if (!out) { set_error("mqe_read_bond_energies: null argument"); return MQE_EINVAL; }
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(&e->n_host, e->d_n, sizeof(int), hipMemcpyDeviceToHost));
    const int n = e->n_host, cap = e->p.capacity, nb = e->nC + 1;
    std::vector<int> ord(n);
    HIPCHK(hipMemcpy(ord.data(), e->d_order, sizeof(int) * n, hipMemcpyDeviceToHost));
    std::vector<double> tmp((size_t)cap * nb);
    HIPCHK(hipMemcpy(tmp.data(), e->d_bond, sizeof(double) * tmp.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        const double *be = tmp.data() + (size_t)ord[i] * nb;
        const double inv = be[e->nC] != 0.0 ? 1.0 / be[e->nC] : 0.0;
        for (int c = 0; c < e->nC; ++c) out[(size_t)i * e->nC + c] = be[c] * inv;
    }
    return MQE_OK;
}

int mqe_save_state(mqe_engine *e, const char *path, long step) {
    if (!e || !path) { set_error("mqe_save_state: bad arguments"); return MQE_EINVAL; }
    if (!e->loaded) { set_error("mqe_save_state: no walkers loaded"); return MQE_ESTATE; }
    ECHK(refresh_n(e));
    const int n = e->n_host, L = e->L, nC = e->nC;
    std::vector<int> order(std::max(n, 1));
    HIPCHK(hipMemcpy(order.data(), e->d_order, sizeof(int) * n, hipMemcpyDeviceToHost));
    FILE *f = std::fopen(path, "wb");
    if (!f) { set_error("mqe_save_state: cannot open the file"); return MQE_EINVAL; }
    StateHeader h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, "MQESTAT3", 8);
    h.L = L;
    h.d = e->p.d;
    h.precision = e->p.precision;
    h.tile32 = e->tile32 ? 1 : 0;
    h.wflat = e->wflat;
    h.eflat = e->eflat;
    h.n = n;
    h.prepared = e->prepared ? 1 : 0;
    h.prop_count = e->prop_count;
    h.step = step;
    h.capacity = e->p.capacity;
    h.nC = nC;
    h.prop_draws = e->prop_draws;
    h.pop_draws = e->pop_draws;
    h.canon_every = e->canon_every;
    const size_t tsz = e->fp32 ? sizeof(float) : sizeof(double);
    std::vector<char> buf(tsz * (size_t)std::max(e->wflat, e->eflat));
    std::vector<int> vd(L + 1);
    std::vector<double> bond(nC + 1);
    std::vector<float> plab((size_t)nC * 16);
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    auto fail = [&](const char *msg) { std::fclose(f); set_error(msg); return MQE_EDEVICE; };
    for (int b = 0; ok && b < n; ++b) {
        const size_t sl = (size_t)order[b];
        const int slot = order[b];
        ok = std::fwrite(&slot, sizeof(int), 1, f) == 1;
        const struct { void *base; long len; } parts[3] = {{e->d_phi, e->wflat}, {e->d_Renv, e->eflat},
                                                           {e->d_Lenv, e->eflat}};
        for (const auto &pt : parts) {
            if (hipMemcpy(buf.data(), (char *)pt.base + tsz * sl * pt.len, tsz * pt.len, hipMemcpyDeviceToHost) !=
                hipSuccess)
                return fail("mqe_save_state: device copy failed");
            ok = ok && std::fwrite(buf.data(), tsz, pt.len, f) == (size_t)pt.len;
        }
        double sc[4];
        int since = 0;
        const bool cp = hipMemcpy(vd.data(), e->d_wvd + sl * (L + 1), sizeof(int) * (L + 1), hipMemcpyDeviceToHost) ==
                            hipSuccess &&
                        hipMemcpy(&since, e->d_since + sl, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess &&
                        hipMemcpy(&sc[0], e->d_w + sl, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
                        hipMemcpy(&sc[1], e->d_O + sl, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
                        hipMemcpy(&sc[2], e->d_eh + sl, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
                        hipMemcpy(&sc[3], e->d_EL + sl, sizeof(double), hipMemcpyDeviceToHost) == hipSuccess;
        if (!cp) return fail("mqe_save_state: device copy failed");
        ok = ok && std::fwrite(vd.data(), sizeof(int), L + 1, f) == (size_t)(L + 1);
        ok = ok && std::fwrite(sc, sizeof(double), 4, f) == 4;
        ok = ok && std::fwrite(&since, sizeof(int), 1, f) == 1;
        if (e->tile32) {
            if (hipMemcpy(bond.data(), e->d_bond + sl * (nC + 1), sizeof(double) * (nC + 1), hipMemcpyDeviceToHost) !=
                    hipSuccess ||
                hipMemcpy(plab.data(), e->d_plab + sl * nC * 16, sizeof(float) * nC * 16, hipMemcpyDeviceToHost) !=
                    hipSuccess)
                return fail("mqe_save_state: device copy failed");
            ok = ok && std::fwrite(bond.data(), sizeof(double), nC + 1, f) == (size_t)(nC + 1);
            ok = ok && std::fwrite(plab.data(), sizeof(float), (size_t)nC * 16, f) == (size_t)nC * 16;
        }
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) { set_error("mqe_save_state: write failed"); return MQE_EINVAL; }
    return MQE_OK;
}

int mqe_load_state(mqe_engine *e, const char *path, long *step) {
    if (!e || !path) { set_error("mqe_load_state: bad arguments"); return MQE_EINVAL; }
    FILE *f = std::fopen(path, "rb");
    if (!f) { set_error("mqe_load_state: cannot open the file"); return MQE_EINVAL; }
    StateHeader h;
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, "MQESTAT3", 8) != 0) {
        std::fclose(f);
        set_error("mqe_load_state: not a walker-state file");
        return MQE_EINVAL;
    }
    const int L = e->L, nC = e->nC;
    if (h.L != L || h.d != e->p.d || h.precision != e->p.precision || h.tile32 != (e->tile32 ? 1 : 0) ||
        h.wflat != e->wflat || h.eflat != e->eflat || h.capacity != e->p.capacity || h.nC != nC) {
        std::fclose(f);
        set_error("mqe_load_state: the file was written by an engine with other parameters");
        return MQE_EINVAL;
    }
    if (h.n < 1 || h.n > e->p.capacity) {
        std::fclose(f);
        set_error("mqe_load_state: population size exceeds the capacity");
        return MQE_ECAPACITY;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    {
        const int r = init_tile_env(e);  // boundary environments of every slot
        if (r != MQE_OK) { std::fclose(f); return r; }
    }
    const size_t tsz = e->fp32 ? sizeof(float) : sizeof(double);
    std::vector<char> buf(tsz * (size_t)std::max(e->wflat, e->eflat));
    std::vector<int> vd(L + 1), order(e->p.capacity, -1);
    std::vector<char> used(e->p.capacity, 0);
    std::vector<double> bond(nC + 1);
    std::vector<float> plab((size_t)nC * 16);
    bool ok = true, dev = true;
    for (int b = 0; ok && dev && b < h.n; ++b) {
        int slot = -1;
        ok = std::fread(&slot, sizeof(int), 1, f) == 1 && slot >= 0 && slot < e->p.capacity && !used[slot];
        if (!ok) break;
        used[slot] = 1;
        order[b] = slot;
        const size_t sl = (size_t)slot;
        const struct { void *base; long len; } parts[3] = {{e->d_phi, e->wflat}, {e->d_Renv, e->eflat},
                                                           {e->d_Lenv, e->eflat}};
        for (const auto &pt : parts) {
            ok = ok && std::fread(buf.data(), tsz, pt.len, f) == (size_t)pt.len;
            if (!ok) break;
            dev = dev && hipMemcpy((char *)pt.base + tsz * sl * pt.len, buf.data(), tsz * pt.len,
                                   hipMemcpyHostToDevice) == hipSuccess;
        }
        double sc[4];
        int since = 0;
        ok = ok && std::fread(vd.data(), sizeof(int), L + 1, f) == (size_t)(L + 1);
        ok = ok && std::fread(sc, sizeof(double), 4, f) == 4;
        ok = ok && std::fread(&since, sizeof(int), 1, f) == 1 && since >= 0;
        if (!ok) break;
        dev = dev && hipMemcpy(e->d_since + sl, &since, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
        for (int k = 0; k <= L; ++k)
            if (vd[k] < 1 || vd[k] > std::max(e->wmax[k], 1)) ok = false;
        if (!ok) break;
        dev = dev && hipMemcpy(e->d_wvd + sl * (L + 1), vd.data(), sizeof(int) * (L + 1), hipMemcpyHostToDevice) ==
                         hipSuccess;
        double *dst[4] = {e->d_w, e->d_O, e->d_eh, e->d_EL};
        for (int q = 0; q < 4; ++q)
            dev = dev && hipMemcpy(dst[q] + sl, &sc[q], sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
        if (e->tile32) {
            ok = ok && std::fread(bond.data(), sizeof(double), nC + 1, f) == (size_t)(nC + 1);
            ok = ok && std::fread(plab.data(), sizeof(float), (size_t)nC * 16, f) == (size_t)nC * 16;
            if (!ok) break;
            dev = dev && hipMemcpy(e->d_bond + sl * (nC + 1), bond.data(), sizeof(double) * (nC + 1),
                                   hipMemcpyHostToDevice) == hipSuccess;
            dev = dev && hipMemcpy(e->d_plab + sl * nC * 16, plab.data(), sizeof(float) * nC * 16,
                                   hipMemcpyHostToDevice) == hipSuccess;
        }
    }
    std::fclose(f);
    if (!dev) { set_error("mqe_load_state: device copy failed"); return MQE_EDEVICE; }
    if (!ok) { set_error("mqe_load_state: truncated or inconsistent file"); return MQE_EINVAL; }
    // the remaining slots follow in slot order (order beyond n is never read)
    for (int s2 = 0, b = h.n; s2 < e->p.capacity; ++s2)
        if (!used[s2]) order[b++] = s2;
    HIPCHK(hipMemcpy(e->d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_n, &h.n, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(e->d_status, 0, sizeof(int)));
    e->n_host = h.n;
    e->prop_count = h.prop_count;
    e->prop_draws = h.prop_draws;
    e->pop_draws = h.pop_draws;
    e->canon_every = h.canon_every;
    e->ctl_host.step = (int)h.prop_draws;
    e->ctl_host.pstep = (int)h.pop_draws;
    e->ctl_host.pc = h.prop_count;
    ECHK(push_ctl(e));
    e->loaded = true;
    e->prepared = h.prepared != 0;
    if (step) *step = h.step;
    return MQE_OK;
}

}  // extern "C"
