// mqe_tile32.h -- FP32 MFMA walker kernel for nearest-neighbour chains with
// all trial and walker bond dimensions <= 32 (the L=32, D=32 production case).
//
// Storage (global memory, zero-padded so every contraction is a full 32x32x32
// MFMA tile):
//   walker site k, physical s:  phi + (2k+s)*1024, element (r, c) at r + 32c
//   environment k (L_k / R_k):  env + k*1024,      element (trial a, walker b) at a + 32b
//   trial:  T32N[(2k+s)*1024] = T_k[s] (padded), T32T = T_k[s]^T (padded)
// LDS holds the working set of one walker: 32x32 blocks with leading dim 33
// (conflict-free for both row- and column-wise MFMA operand reads).
//
// One 256-thread workgroup (four wave64s) per walker.  Every 32x32x32 product
// is one wave issuing 16 v_mfma_f32_32x32x2f32; the four waves work on
// independent tiles between workgroup barriers.  QR (LQ) of the energy sweep is
// FP64 Householder in LDS with one barrier per column; sampling, weights,
// overlaps and energies are FP64.
#ifndef MQE_TILE32_H
#define MQE_TILE32_H

namespace mqe {
namespace t32 {

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int LD = 33;           // LDS leading dimension
constexpr int BLK = 32 * LD;     // floats per LDS block
constexpr int GB = 1024;         // floats per global block

// LDS map (in LDS blocks).  The walker's center tensors rotate over the block
// pairs 0-1, 2-3 and 6-7 in the propagation and alternate between 0-1 and 2-3
// in the energy sweep.  During the LQ the FP64 workspace overlays blocks 7-15.
enum : int {
    B_BUF0 = 0,  // 2 blocks
    B_BUF1 = 2,  // 2 blocks
    B_L = 4,     // left environment
    B_R = 5,     // right environment
    B_LM = 6,    // energy sweep: Lmat of the LQ or the right norm environment M_k
    B_N = 7,     // energy sweep: LQ workspace
    B_BUF2 = 6,  // propagation: third center pair (blocks 6-7)
    B_Y = 8,     // 2 blocks (G: 4 blocks)
    B_W = 10,    // 2 blocks
    B_V = 12,    // 4 blocks
    NBLK = 16
};
constexpr int MAXG_T32 = 256;                    // grid points: four per lane of one wave

constexpr size_t lds_bytes() { return sizeof(float) * (size_t)NBLK * BLK; }

__device__ __forceinline__ v16f zero16() {
    v16f z;
#pragma unroll
    for (int q = 0; q < 16; ++q) z[q] = 0.f;
    return z;
}

// acc += op(A)(32 x 32) op(B)(32 x 32); TA: A stored transposed; TB: B stored transposed.
template <bool TA, bool TB>
__device__ __forceinline__ void mma(v16f &acc, const float *A, int lda, const float *B, int ldb) {
    const int l = threadIdx.x & 63, r = l & 31, h = l >> 5;
#pragma unroll
    for (int k0 = 0; k0 < 32; k0 += 2) {
        const int k = k0 + h;
        const float a = TA ? A[k + lda * r] : A[r + lda * k];
        const float b = TB ? B[r + ldb * k] : B[k + ldb * r];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
    }
}

// A-operand fragment of a 32 x 32 block stored column-major with ld 32 (the
// padded trial blocks in global memory): a[q] = A(r, 2q + h).  Loaded ahead of
// the product so that the global-memory latency overlaps other work.
typedef float frag16[16];
__device__ __forceinline__ void load_frag(frag16 &a, const float *A) {
    const int l = threadIdx.x & 63, r = l & 31, h = l >> 5;
#pragma unroll
    for (int q = 0; q < 16; ++q) a[q] = A[r + 32 * (2 * q + h)];
}

// acc += A B with A given as a register fragment (load_frag) and B in LDS
__device__ __forceinline__ void mma_frag(v16f &acc, const frag16 &a, const float *B, int ldb) {
    const int l = threadIdx.x & 63, r = l & 31, h = l >> 5;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const float b = B[(2 * q + h) + ldb * r];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q], b, acc, 0, 0, 0);
    }
}

// MFMA result (accumulator layout: slot q of lane (j, h) holds row
// 8(q/4) + 4h + q%4 of column j) -> B-operand fragment of the same matrix
// (b[q] = row 2q + h of column j): half of the rows move between lanes l and
// l ^ 32.  v_permlane32_swap(x, y) returns (x of lanes 0-31 | y of lanes 0-31)
// and (x of lanes 32-63 | y of lanes 32-63) (checked on gfx950), which is
// exactly (b[q], b[q|2]) for x = acc[base], y = acc[base + 1].
__device__ __forceinline__ void acc_to_bfrag(const v16f &acc, frag16 &b) {
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int o = 0; o < 2; ++o) {
            const int base = 4 * g + 2 * o, q = 4 * g + o;
            const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[base]),
                                                            __float_as_uint(acc[base + 1]), false, false);
            b[q] = __uint_as_float(r[0]);
            b[q | 2] = __uint_as_float(r[1]);
        }
}

// acc += A B with both operands as register fragments
__device__ __forceinline__ void mma_rr(v16f &acc, const frag16 &a, const frag16 &b) {
#pragma unroll
    for (int q = 0; q < 16; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q], b[q], acc, 0, 0, 0);
}

// acc += op(A) B, A in LDS (TA: stored transposed), B a register fragment
template <bool TA>
__device__ __forceinline__ void mma_lr(v16f &acc, const float *A, int lda, const frag16 &b) {
    const int l = threadIdx.x & 63, r = l & 31, h = l >> 5;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int k = 2 * q + h;
        const float a = TA ? A[k + lda * r] : A[r + lda * k];
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b[q], acc, 0, 0, 0);
    }
}

// element (i, j) held by this lane in accumulator slot q
__device__ __forceinline__ int row_of(int q) { return 8 * (q >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (q & 3); }
__device__ __forceinline__ int col_of() { return threadIdx.x & 31; }

__device__ __forceinline__ void store(const v16f &acc, float *C, int ldc) {
    const int j = col_of();
#pragma unroll
    for (int q = 0; q < 16; ++q) C[row_of(q) + ldc * j] = acc[q];
}

// FP32 partial of <acc, B> over this lane's 16 elements
__device__ __forceinline__ float dot_partial(const v16f &acc, const float *B, int ldb) {
    const int j = col_of();
    float f = 0.f;
#pragma unroll
    for (int e = 0; e < 16; ++e) f += acc[e] * B[row_of(e) + ldb * j];
    return f;
}

// (x, y) -> x + y of the partner half: for lanes 0-31 x(l) + x(l + 32), for
// lanes 32-63 y(l - 32) + y(l), with one v_permlane32_swap per dword
__device__ __forceinline__ double halves_sum(double x, double y) {
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y),
                                                     false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y),
                                                     false, false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// (x, y) -> for 16-lane rows 0 and 2 x(l) + x(l + 16), for rows 1 and 3
// y(l - 16) + y(l), with one v_permlane16_swap per dword (it returns rows
// (x0, y0, x2, y2) and (x1, y1, x3, y3); checked on gfx950)
__device__ __forceinline__ double rows_sum(double x, double y) {
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x), (unsigned)__double2loint(y),
                                                     false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x), (unsigned)__double2hiint(y),
                                                     false, false);
    return __hiloint2double((int)hi[0], (int)lo[0]) + __hiloint2double((int)hi[1], (int)lo[1]);
}

// four wave-wide sums: lane l receives the sum of s_{l >> 4}
__device__ __forceinline__ double reduce4(double s0, double s1, double s2, double s3) {
    const double a0 = halves_sum(s0, s2), a1 = halves_sum(s1, s3);  // lanes < 32: s0, s1; else s2, s3
    double c0 = rows_sum(a0, a1);                                    // row r: s_r over four lanes
    c0 = group8_sum(c0);
    c0 += dpp_d<0x140>(c0);
    return c0;
}

// The four dots <acc, B_q>, B_q = base + q BLK (q = 0..3): each lane sums its
// 16 products in FP32 (the operands are FP32 MFMA results), the cross-lane
// sum is FP64.  Lane l receives dot l >> 4 (a reduce-scatter over the lane
// halves, then a DPP sum over each 16-lane row).
__device__ __forceinline__ double dot4_tile(const v16f &acc, const float *base, int ldc) {
    return reduce4(dot_partial(acc, base, ldc), dot_partial(acc, base + BLK, ldc),
                   dot_partial(acc, base + 2 * BLK, ldc), dot_partial(acc, base + 3 * BLK, ldc));
}

// global padded block -> LDS block (whole workgroup)
__device__ __forceinline__ void g2l(const float *g, float *s) {
    for (int e = threadIdx.x; e < GB; e += MQE_NT) s[(e & 31) + LD * (e >> 5)] = g[e];
}
__device__ __forceinline__ void l2g(const float *s, float *g) {
    for (int e = threadIdx.x; e < GB; e += MQE_NT) g[e] = s[(e & 31) + LD * (e >> 5)];
}
// NB global padded blocks -> registers (element tid + 256 n of block b in
// r[4 b + n]) and registers -> LDS blocks; used to fetch the next coupling's
// operands while the current one computes.
template <int NB>
__device__ __forceinline__ void pre_load(float (&r)[16], const float *const (&g)[NB]) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int n = 0; n < 4; ++n) r[4 * b + n] = g[b][threadIdx.x + MQE_NT * n];
}
template <int NB>
__device__ __forceinline__ void pre_store(const float (&r)[16], float *const (&s)[NB]) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const int e = threadIdx.x + MQE_NT * n;
            s[b][(e & 31) + LD * (e >> 5)] = r[4 * b + n];
        }
}
__device__ __forceinline__ void set_unit(float *s) {  // 1 at (0,0), 0 elsewhere
    for (int e = threadIdx.x; e < BLK; e += MQE_NT) s[e] = (e == 0) ? 1.f : 0.f;
}

struct W32 {
    const Sys *S;
    const float *TN, *TT;   // trial padded blocks
    float *phi, *Lg, *Rg;   // this walker's global storage
    float *sm;              // LDS floats
    const int *wv;          // walker bond dims (LDS)
    double *red;            // LDS reduction scratch
    double *bond;           // this walker's bond energies (nC raw values, then the overlap) or nullptr
    __device__ float *blk(int b) const { return sm + (size_t)b * BLK; }
    __device__ const float *tn(int k, int s) const { return TN + (size_t)(2 * k + s) * GB; }
    __device__ const float *tt(int k, int s) const { return TT + (size_t)(2 * k + s) * GB; }
    __device__ float *ph(int k, int s) const { return phi + (size_t)(2 * k + s) * GB; }
};

// ---------------------------------------------------------------------------
// FP64 Householder LQ of the center tensor (two LDS blocks, w_k x 2 w_{k+1}):
// phi = Lmat Q with Q right-orthonormal.  Q overwrites the center blocks and
// Lmat (lower triangular, w_k x w_k) goes to B_LM.  The transposed matrix
// A = phi^T (m = 2 w_{k+1} <= 64 rows, n = w_k <= 32 columns) is factorised
// with row index r = col + w_{k+1} s.  32 groups of 8 lanes: group g keeps
// column g in registers (lane ln holds rows ln + 8t, t = 0..7); only the
// reflector vectors and tau pass through LDS (V: 64 x 32 doubles, column j at
// V + 64 j), one workgroup barrier per column.
// ---------------------------------------------------------------------------

__device__ void lq32(const W32 &C, int wk, int wn, float *center) {
    const int m = 2 * wn, n = wk;
    double *V = (double *)C.blk(B_N), *tau = V + 64 * 32;
    const int tid = threadIdx.x, g = tid >> 3, ln = tid & 7;
    // this group's column of A = phi^T
    double a[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int r = ln + 8 * t;
        double v = 0.0;
        if (r < m && g < n) {
            const int s = r >= wn ? 1 : 0, col = r - s * wn;
            v = (double)center[s * BLK + g + LD * col];
        }
        a[t] = v;
    }
    // Householder reflector of the own column j (rows j..m-1): v_j = 1 implied,
    // v_i (i > j) written to V, beta kept in a (row j), tau to LDS.
    auto reflect = [&](int j) {
        double ss = 0.0, alpha = 0.0;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int r = ln + 8 * t;
            if (r > j) ss += a[t] * a[t];
            if (r == j) alpha = a[t];
        }
        ss = group8_sum(ss);
        alpha = group8_sum(alpha);
        double tj = 0.0, sc = 0.0, beta = alpha;
        if (ss != 0.0) {
            beta = -copysign(sqrt(alpha * alpha + ss), alpha);
            tj = (beta - alpha) / beta;
            sc = 1.0 / (alpha - beta);
        }
        double *vj = V + 64 * j;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int r = ln + 8 * t;
            if (r > j) {
                if (tj != 0.0) a[t] *= sc;
                vj[r] = (tj != 0.0) ? a[t] : 0.0;
            }
            if (r == j) a[t] = beta;
        }
        if (ln == 0) tau[j] = tj;
    };
    if (g == 0) reflect(0);
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (g > j && g < n) {
            const double tj = tau[j];
            if (tj != 0.0) {
                const double *vj = V + 64 * j;
                double vr[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int r = ln + 8 * t;
                    vr[t] = (r > j && r < m) ? vj[r] : (r == j ? 1.0 : 0.0);
                }
                double s = 0.0;
#pragma unroll
                for (int t = 0; t < 8; ++t) s += vr[t] * a[t];
                s = tj * group8_sum(s);
#pragma unroll
                for (int t = 0; t < 8; ++t) a[t] -= s * vr[t];
            }
            if (g == j + 1) reflect(j + 1);
        }
        __syncthreads();
    }
    // Lmat = R^T (lower): R(j, g) for j <= g is held in group g, row j
    float *LM = C.blk(B_LM);
    for (int e = tid; e < BLK; e += MQE_NT) LM[e] = 0.f;
    __syncthreads();
    if (g < n) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int r = ln + 8 * t;
            if (r <= g) LM[g + LD * r] = (float)a[t];  // Lmat(g, r) = R(r, g)
        }
    }
    // Q column c = H_0 H_1 ... H_c e_c, group c; written transposed into the center
    if (g < n) {
        const int c = g;
        double q[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) q[t] = (ln + 8 * t == c) ? 1.0 : 0.0;
        for (int k = c; k >= 0; --k) {
            const double tk = tau[k];
            if (tk == 0.0) continue;
            const double *vk = V + 64 * k;
            double vr[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int r = ln + 8 * t;
                vr[t] = (r > k && r < m) ? vk[r] : (r == k ? 1.0 : 0.0);
            }
            double s = 0.0;
#pragma unroll
            for (int t = 0; t < 8; ++t) s += vr[t] * q[t];
            s = tk * group8_sum(s);
#pragma unroll
            for (int t = 0; t < 8; ++t) q[t] -= s * vr[t];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int r = ln + 8 * t;
            if (r < m) {
                const int s = r >= wn ? 1 : 0, col = r - s * wn;
                center[s * BLK + c + LD * col] = (float)q[t];
            }
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// the walker kernel (flags: F_PREP or F_PROP), h = 0, couplings (i, i+1)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void scale_blocks(float *b, int nb, double f) {
    for (int e = threadIdx.x; e < nb * BLK; e += MQE_NT) b[e] = (float)((double)b[e] * f);
}

// phi[s] <- sum_s0 O[s + 2 s0] phi[s0] on a 2-block LDS tensor
__device__ __forceinline__ void apply2(float *b, const double *O) {
    const float o00 = (float)O[0], o10 = (float)O[1], o01 = (float)O[2], o11 = (float)O[3];
    // the 32 x 32 entries of each block (the padding row is never read); all
    // loads of the thread first, then the stores
    constexpr int NE = GB / MQE_NT;
    float a[NE][2];
#pragma unroll
    for (int it = 0; it < NE; ++it) {
        const int e = threadIdx.x + it * MQE_NT, o = (e & 31) + LD * (e >> 5);
        a[it][0] = b[o];
        a[it][1] = b[o + BLK];
    }
#pragma unroll
    for (int it = 0; it < NE; ++it) {
        const int e = threadIdx.x + it * MQE_NT, o = (e & 31) + LD * (e >> 5);
        b[o] = o00 * a[it][0] + o01 * a[it][1];
        b[o + BLK] = o10 * a[it][0] + o11 * a[it][1];
    }
}

// <trial|phi> with the walker's site-0 tensor in LDS block pair c0 and R_1 in
// B_R, and the squared norm sum_s <phi_0[s], phi_0[s] M_1> with the right norm
// environment M_1 (Mb; nullptr = identity, i.e. sites > 0 right-orthonormal).
__device__ void site0_32(const W32 &C, const float *c0, const float *Mb, double *ov, double *n2) {
    const float *R1 = C.blk(B_R);
    double a = 0.0, b = 0.0;
    for (int e = threadIdx.x; e < 2 * 32 * 32; e += MQE_NT) {
        const int s = e >> 10, al = (e >> 5) & 31, be = e & 31;
        const double t = (double)C.tn(0, s)[0 + 32 * al];
        const double p = (double)c0[s * BLK + 0 + LD * be];
        a += t * p * (double)R1[al + LD * be];
        if (Mb) b += (double)c0[s * BLK + 0 + LD * al] * (double)Mb[al + LD * be] * p;
        else if (al == 0) b += p * p;
    }
    double v[2] = {a, b};
    block_sum_n<2>(v, C.red);
    *ov = v[0];
    *n2 = v[1];
}

// Energy sweep (right -> left) from the center phi_{L-1} (LDS pair pc); L_k in
// global memory.  Returns <trial|H|phi>; *Ou = <trial|phi>; normalises site 0.
__device__ double energy_sweep32(const W32 &C, int pc, double *On, double *Ou, bool canon) {
    const Sys &S = *C.S;
    const int L = S.L;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    __shared__ double Mq[16];
    set_unit(C.blk(B_R));
    set_unit(C.blk(B_LM));  // M_L = 1 (light sweep)
    double etot = 0.0;
    int pj = pc == B_BUF0 ? B_BUF1 : B_BUF0;
    if (!canon) {
        // light sweep: no re-orthonormalisation; right environments R_k and
        // right norm environments M_k = sum_t phi_k[t] M_{k+1} phi_k[t]^T.
        // The last site's propagated tensor is only in LDS: store it.
        l2g(C.blk(pc), C.ph(L - 1, 0));
        l2g(C.blk(pc) + BLK, C.ph(L - 1, 1));
        // operands of site k, stored one site ahead: phi_{k-1} (into the buffer
        // that becomes free) and L_{k-1}
        float pre[16];
        {
            const float *const g[3] = {C.ph(L - 2, 0), C.ph(L - 2, 1), C.Lg + (size_t)(L - 2) * GB};
            pre_load<3>(pre, g);
            float *const d[3] = {C.blk(pj), C.blk(pj) + BLK, C.blk(B_L)};
            pre_store<3>(pre, d);
        }
        // trial operands: T_k[0], T_k[1] (waves 0,1, stage 1) and T_{k-1}[s']^T (stage 2)
        frag16 f0, f1, fg;
        if (w < 2) {
            load_frag(f0, C.tn(L - 1, 0));
            load_frag(f1, C.tn(L - 1, 1));
        }
        load_frag(fg, C.tt(L - 2, w >> 1));
        __syncthreads();
        for (int k = L - 1; k >= 1; --k) {
            float *cen = C.blk(pc), *nxt = C.blk(pj);
            if (k >= 2) {
                const float *const g[3] = {C.ph(k - 2, 0), C.ph(k - 2, 1), C.Lg + (size_t)(k - 2) * GB};
                pre_load<3>(pre, g);
            }
            // stage 1, chained in registers:
            //   waves 0,1 (t = w):     W[t] = R_{k+1} phi_k[t]^T,  V(t',t) = T_k[t'] W[t]  (to B_V + 2t' + t)
            //   waves 2,3 (t = w - 2): Z[t] = M_{k+1} phi_k[t]^T,  phi_k[t] Z[t]          (term t of M_k, to B_W + t)
            //                          Y[t] = L_{k-1} phi_{k-1}[t]                         (to B_Y + t)
            if (w < 2) {
                v16f y = zero16(), r0 = zero16(), r1 = zero16();
                mma<false, true>(y, C.blk(B_R), LD, cen + w * BLK, LD);
                frag16 yb;
                acc_to_bfrag(y, yb);
                mma_rr(r0, f0, yb);
                mma_rr(r1, f1, yb);
                store(r0, C.blk(B_V) + w * BLK, LD);
                store(r1, C.blk(B_V) + (2 + w) * BLK, LD);
            } else {
                const int t = w - 2;
                v16f z = zero16(), m = zero16(), y = zero16();
                mma<false, true>(z, C.blk(B_LM), LD, cen + t * BLK, LD);
                frag16 zb;
                acc_to_bfrag(z, zb);
                mma_lr<false>(m, cen + t * BLK, LD, zb);
                store(m, C.blk(B_W) + t * BLK, LD);
                mma<false, false>(y, C.blk(B_L), LD, nxt + t * BLK, LD);
                store(y, C.blk(B_Y) + t * BLK, LD);
            }
            __syncthreads();
            // stage 2: bond energy of (k-1, k) from G(s',s) = T_{k-1}[s']^T Y[s] (wave w: s' = w >> 1,
            // s = w & 1) and V; R_k = V(0,0) + V(1,1) and M_k = sum of its terms to LDS and global
            // memory; the next site's operands
            {
                const float *V = C.blk(B_V), *Mt = C.blk(B_W);
                v16f g = zero16();
                mma_frag(g, fg, C.blk(B_Y) + (w & 1) * BLK, LD);
                if (k >= 2) {
                    if (w < 2) {
                        load_frag(f0, C.tn(k - 1, 0));
                        load_frag(f1, C.tn(k - 1, 1));
                    }
                    load_frag(fg, C.tt(k - 2, w >> 1));
                }
                const int sp = w >> 1, ss = w & 1;
                {
                    const double d = dot4_tile(g, V, LD);
                    const int q = lane >> 4;
                    if ((lane & 15) == 0) Mq[(sp + 2 * (q >> 1)) * 4 + (ss + 2 * (q & 1))] = d;
                }
                float *R = C.blk(B_R), *Mk = C.blk(B_LM);
                float *rg = C.Rg + (size_t)k * GB;
                // all LDS loads of the thread's four elements before the first
                // store (stores to B_R / B_LM could otherwise not pass them)
                constexpr int NE = GB / MQE_NT;
                float vv[NE][2], mv[NE][2];
#pragma unroll
                for (int it = 0; it < NE; ++it) {
                    const int e = threadIdx.x + it * MQE_NT, o = (e & 31) + LD * (e >> 5);
                    vv[it][0] = V[o];
                    vv[it][1] = V[3 * BLK + o];
                    mv[it][0] = Mt[o];
                    mv[it][1] = Mt[BLK + o];
                }
#pragma unroll
                for (int it = 0; it < NE; ++it) {
                    const int e = threadIdx.x + it * MQE_NT, o = (e & 31) + LD * (e >> 5);
                    const float rv = vv[it][0] + vv[it][1];
                    R[o] = rv;
                    rg[e] = rv;
                    Mk[o] = mv[it][0] + mv[it][1];
                }
                if (k >= 2) {  // phi_{k-2} into the buffer of phi_k (not read in this stage), L_{k-2}
                    float *const d[3] = {cen, cen + BLK, C.blk(B_L)};
                    pre_store<3>(pre, d);
                }
            }
            __syncthreads();
            double e = 0.0;
            for (int q = 0; q < 16; ++q) e += S.SS[q] * Mq[q];
            etot += S.Jc[k - 1] * e;
            if (C.bond && threadIdx.x == 0) C.bond[k - 1] = S.Jc[k - 1] * e;
            int t = pc; pc = pj; pj = t;
        }
        double ov, n2;
        float *c0 = C.blk(pc);
        site0_32(C, c0, C.blk(B_LM), &ov, &n2);
        const double nrm = sqrt(n2);
        double O = ov / nrm, f = 1.0 / nrm;
        if (O < 0.0) { O = -O; f = -f; }
        scale_blocks(c0, 2, f);
        __syncthreads();
        l2g(c0, C.ph(0, 0));
        l2g(c0 + BLK, C.ph(0, 1));
        if (C.bond && threadIdx.x == 0) C.bond[S.nC] = ov;
        *On = O;
        *Ou = ov;
        return etot;
    }
    for (int k = L - 1; k >= 1; --k) {
        float *cen = C.blk(pc), *nxt = C.blk(pj);
        frag16 fv, fg;  // trial operands of stages 2 and 3
        load_frag(fv, C.tn(k, w >> 1));
        load_frag(fg, C.tt(k - 1, w >> 1));
        g2l(C.ph(k - 1, 0), nxt);
        g2l(C.ph(k - 1, 1), nxt + BLK);
        g2l(C.Lg + (size_t)(k - 1) * GB, C.blk(B_L));
        __syncthreads();
        lq32(C, C.wv[k], C.wv[k + 1], cen);  // cen <- Q_k, B_LM <- Lmat
        // stage 1: push (waves 0,1, in place), W[t] = R_{k+1} Q_k[t]^T (waves 2,3)
        {
            v16f acc = zero16();
            if (w < 2) {
                mma<false, false>(acc, nxt + w * BLK, LD, C.blk(B_LM), LD);
                store(acc, nxt + w * BLK, LD);
            } else {
                const int t = w - 2;
                mma<false, true>(acc, C.blk(B_R), LD, cen + t * BLK, LD);
                store(acc, C.blk(B_W) + t * BLK, LD);
            }
        }
        __syncthreads();
        // stage 2: V(t',t) = T_k[t'] W[t] (4 waves); Y[s] = L_{k-1} phi'_{k-1}[s] (waves 0,1)
        {
            v16f acc = zero16();
            mma_frag(acc, fv, C.blk(B_W) + (w & 1) * BLK, LD);
            store(acc, C.blk(B_V) + w * BLK, LD);
            if (w < 2) {
                v16f y = zero16();
                mma<false, false>(y, C.blk(B_L), LD, nxt + w * BLK, LD);
                store(y, C.blk(B_Y) + w * BLK, LD);
            }
        }
        // write Q_k (final for this step) to global while the products finish
        l2g(cen, C.ph(k, 0));
        l2g(cen + BLK, C.ph(k, 1));
        __syncthreads();
        // stage 3: R_k = V(0,0) + V(1,1); bond energy of (k-1, k) from G = T_{k-1}^T Y and V
        {
            const float *V = C.blk(B_V);
            float *R = C.blk(B_R);
            for (int e = threadIdx.x; e < BLK; e += MQE_NT) R[e] = V[e] + V[3 * BLK + e];
            v16f g = zero16();
            mma_frag(g, fg, C.blk(B_Y) + (w & 1) * BLK, LD);
            const int sp = w >> 1, ss = w & 1;
            {
                const double d = dot4_tile(g, V, LD);
                const int q = lane >> 4;
                if ((lane & 15) == 0) Mq[(sp + 2 * (q >> 1)) * 4 + (ss + 2 * (q & 1))] = d;
            }
        }
        __syncthreads();
        l2g(C.blk(B_R), C.Rg + (size_t)k * GB);
        double e = 0.0;
        for (int q = 0; q < 16; ++q) e += S.SS[q] * Mq[q];
        etot += S.Jc[k - 1] * e;
        if (C.bond && threadIdx.x == 0) C.bond[k - 1] = S.Jc[k - 1] * e;
        int t = pc; pc = pj; pj = t;
        __syncthreads();
    }
    double ov, n2;
    float *c0 = C.blk(pc);
    site0_32(C, c0, nullptr, &ov, &n2);
    const double nrm = sqrt(n2);
    double O = ov / nrm, f = 1.0 / nrm;
    if (O < 0.0) { O = -O; f = -f; }
    scale_blocks(c0, 2, f);
    __syncthreads();
    l2g(c0, C.ph(0, 0));
    l2g(c0 + BLK, C.ph(0, 1));
    if (C.bond && threadIdx.x == 0) C.bond[S.nC] = ov;
    *On = O;
    *Ou = ov;
    return etot;
}

// LEARN: sample from the learned proposal P.pred with the importance weight
template <bool LEARN>
__global__ __launch_bounds__(MQE_NT, 2) void walker32_kernel(Sys S, Pool P, const float *TN, const float *TT,
                                                              int flags) {
    extern __shared__ float lds32[];
    const int n = *P.n;
    if ((int)blockIdx.x >= n) return;
    if (*P.status & 4) return;  // the population overflowed: keep the state as it was
    const int b = P.list ? P.list[blockIdx.x] : (int)blockIdx.x;  // population position
    const int slot = P.order[b];
    const Ctl ctl = *P.ctl;
    const int L = S.L;
    __shared__ int wv[MQE_MAXL + 1];
    __shared__ double red[2 * MQE_WAVES];
    __shared__ double M[16];
    __shared__ double grid_s[4 * MAXG_T32];  // the Marsaglia grid
    __shared__ double tab_s[2][36];          // Trotter tables S, U, VT of the current and the next coupling
    for (int k = threadIdx.x; k <= L; k += MQE_NT) wv[k] = P.wvd[(size_t)slot * (L + 1) + k];
    W32 C;
    C.S = &S;
    C.TN = TN;
    C.TT = TT;
    C.phi = (float *)P.phi + (size_t)slot * P.wflat;
    C.Lg = (float *)P.Lenv + (size_t)slot * P.eflat;
    C.Rg = (float *)P.Renv + (size_t)slot * P.eflat;
    C.sm = lds32;
    C.wv = wv;
    C.red = red;
    C.bond = P.bondE ? P.bondE + (size_t)slot * (S.nC + 1) : nullptr;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    double wt = P.w[slot], O = P.O[slot];

    if (flags & F_PREP) {
        // left environments of the loaded state
        set_unit(C.blk(B_L));
        __syncthreads();
        float *cen = C.blk(B_BUF0);
        for (int k = 0; k < L - 1; ++k) {
            g2l(C.ph(k, 0), cen);
            g2l(C.ph(k, 1), cen + BLK);
            __syncthreads();
            if (w < 2) {
                v16f y = zero16();
                mma<false, false>(y, C.blk(B_L), LD, cen + w * BLK, LD);
                store(y, C.blk(B_Y) + w * BLK, LD);
            }
            __syncthreads();
            if (w == 0) {
                v16f a = zero16();
                mma<false, false>(a, C.tt(k, 0), 32, C.blk(B_Y), LD);
                mma<false, false>(a, C.tt(k, 1), 32, C.blk(B_Y) + BLK, LD);
                store(a, C.blk(B_L), LD);
            }
            __syncthreads();
            l2g(C.blk(B_L), C.Lg + (size_t)(k + 1) * GB);
            __syncthreads();
        }
        g2l(C.ph(L - 1, 0), cen);
        g2l(C.ph(L - 1, 1), cen + BLK);
        __syncthreads();
        double On, Ou;
        const double e = energy_sweep32(C, B_BUF0, &On, &Ou, true);
        if (threadIdx.x == 0) {
            P.EL[slot] = e / O;
            P.O[slot] = On;
            P.since[slot] = 0;
        }
        return;
    }

    // ---------------- propagation (MPSQMC2.cpp:532-600), h = 0 ----------------
    // The walker is not normalised between couplings: the right site's sampled
    // operator is scaled by O / <trial|B(x) phi> = O / (4 f*), so the overlap
    // stays O and the norm stays near 1 (every quantity of the step is a ratio
    // in which the walker's scale cancels); the energy sweep takes the exact
    // norm.  The center tensors rotate over three LDS block pairs: the current
    // left site (pc), the right site (pj) and the next coupling's right site
    // (pf), stored from registers fetched one coupling ahead.
    int pc = B_BUF0, pj = B_BUF1, pf = B_BUF2;
    g2l(C.ph(0, 0), C.blk(pc));
    g2l(C.ph(0, 1), C.blk(pc) + BLK);
    g2l(C.ph(1, 0), C.blk(pj));
    g2l(C.ph(1, 1), C.blk(pj) + BLK);
    g2l(C.Rg + (size_t)2 * GB, C.blk(B_R));
    set_unit(C.blk(B_L));
    const int Ng = S.Ng;
    for (int e = threadIdx.x; e < 4 * Ng; e += MQE_NT) grid_s[e] = S.grid[e];
    // tables of coupling c: tab[0..3] = sqrt(S), tab[4..19] = U, tab[20..35] = VT
    auto table_entry = [&](int c, int t) {
        return t < 4 ? sqrt(S.tS[4 * c + t]) : (t < 20 ? S.tU[16 * c + t - 4] : S.tVT[16 * c + t - 20]);
    };
    if (threadIdx.x < 36) tab_s[0][threadIdx.x] = table_entry(0, threadIdx.x);
    double tnext = (threadIdx.x < 36 && S.nC > 1) ? table_entry(1, threadIdx.x) : 0.0;
    __syncthreads();
    float pre[16];  // phi_{j+1} and R_{j+2}, the next coupling's walker operands
    for (int c = 0; c < S.nC; ++c) {
        const int i = c, j = c + 1;
        float *cen = C.blk(pc), *nxt = C.blk(pj);
        float plab = 0.f;  // learned proposal: predicted label lane of this coupling (lanes < 16)
        if constexpr (LEARN) plab = lane < 16 ? P.pred[((size_t)b * S.nC + c) * 16 + lane] : 0.f;
        // the coupling's uniform, drawn in stage AB so that the counter-based
        // generator (or the load) runs between the matrix-core products instead
        // of on the sampling stage's critical path (pinned there below)
        double u = P.u ? P.u[(size_t)b * S.nC + c]
                       : philox_uniform(P.seed, (uint32_t)b, (uint32_t)c, (uint32_t)ctl.step, 0u);
        // trial operands of stage AB: T_i[s']^T (waves 0,1) or T_j[t'] (waves 2,3), s', t' = 0, 1
        frag16 f0, f1;
        if (w < 2) {
            load_frag(f0, C.tt(i, 0));
            load_frag(f1, C.tt(i, 1));
        } else {
            load_frag(f0, C.tn(j, 0));
            load_frag(f1, C.tn(j, 1));
        }
        if (c + 1 < S.nC) {
            const float *const g[3] = {C.ph(j + 1, 0), C.ph(j + 1, 1), C.Rg + (size_t)(j + 2) * GB};
            pre_load<3>(pre, g);
        }
        // stage AB, chained in registers:
        //   waves 0,1 (s = w):     Y[s] = L_i phi_i[s],        G(s',s) = T_i[s']^T Y[s]
        //   waves 2,3 (t = w - 2): W[t] = R_{j+1} phi_j[t]^T,  V(t',t) = T_j[t'] W[t]
        // G(s',s) is stored at B_Y + (2s' + s), V(t',t) at B_V + (2t' + t).
        v16f r0 = zero16(), r1 = zero16();
        {
            v16f y = zero16();
            if (w < 2)
                mma<false, false>(y, C.blk(B_L), LD, cen + w * BLK, LD);
            else
                mma<false, true>(y, C.blk(B_R), LD, nxt + (w - 2) * BLK, LD);
            frag16 yb;
            acc_to_bfrag(y, yb);
            mma_rr(r0, f0, yb);
            mma_rr(r1, f1, yb);
            float *base = (w < 2) ? C.blk(B_Y) : C.blk(B_V);
            store(r0, base + (w & 1) * BLK, LD);
            store(r1, base + (2 + (w & 1)) * BLK, LD);
        }
        asm volatile("" : "+v"(u));
        __syncthreads();
        // stage C: the 16 two-site mixed matrix elements M[(s'+2t')4 + (s+2t)] = <G(s',s), V(t',t)>
        //   waves 0,1: <G(0,s), V(t',t)> for all t', t;  waves 2,3: <V(t',t), G(1,s)> for all t', s
        {
            double d;
            int idx;
            const int q = lane >> 4;
            if (w < 2) {
                const float *V = C.blk(B_V);
                d = reduce4(dot_partial(r0, V, LD), dot_partial(r0, V + BLK, LD), dot_partial(r0, V + 2 * BLK, LD),
                            dot_partial(r0, V + 3 * BLK, LD));
                idx = (2 * (q >> 1)) * 4 + (w + 2 * (q & 1));
            } else {
                const float *G1 = C.blk(B_Y) + 2 * BLK;
                d = reduce4(dot_partial(r0, G1, LD), dot_partial(r0, G1 + BLK, LD), dot_partial(r1, G1, LD),
                            dot_partial(r1, G1 + BLK, LD));
                idx = (1 + 2 * (q >> 1)) * 4 + ((q & 1) + 2 * (w - 2));
            }
            if ((lane & 15) == 0) M[idx] = d;
        }
        __syncthreads();
        // stage D: combos, PDF, CDF sample (MPSQMC2.cpp:540-588).  Every wave
        // computes the same result, so no workgroup barrier is needed: lane
        // q < 16 forms combo q, lane l owns grid points 4l .. 4l+3.
        const double *sS = tab_s[c & 1], *tU = sS + 4, *tVT = sS + 20;  // sS = sqrt(S)
        double cq = 0.0;
        if (lane < 16) {
            // combo (a, b) = sqrt(S_a S_b) sum U_a(s',s) VT_b(t',t) M[(s'+2t')4 + (s+2t)], operands loaded first
            const int a = lane & 3, bb = lane >> 2;
            double ua[4], vb[4], m[16];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                ua[q] = tU[(q & 1) + 2 * ((q >> 1) + 2 * a)];    // ua[s' + 2s] = U_a(s', s)
                vb[q] = tVT[bb + 4 * ((q & 1) + 2 * (q >> 1))];  // vb[t' + 2t] = VT_b(t', t)
            }
#pragma unroll
            for (int q = 0; q < 16; ++q) m[q] = M[q];
            double sacc = 0.0;
#pragma unroll
            for (int xsp = 0; xsp < 2; ++xsp)
#pragma unroll
                for (int xss = 0; xss < 2; ++xss) {
                    double mv = 0.0;
#pragma unroll
                    for (int tp = 0; tp < 2; ++tp)
#pragma unroll
                        for (int t = 0; t < 2; ++t) mv += vb[tp + 2 * t] * m[(xsp + 2 * tp) * 4 + (xss + 2 * t)];
                    sacc += ua[xsp + 2 * xss] * mv;
                }
            cq = sS[a] * sS[bb] * sacc;
        }
        double cb[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) cb[q] = readlane_d(cq, q);
        const double pref = 4.0 / (O * Ng);
        if (P.plab_on && w == 0 && lane < 16) {  // labels combos/O: the walker's record and the training rows
            const float lv = (float)(cq / O);
            P.plab[((size_t)slot * S.nC + c) * 16 + lane] = lv;
            if (P.lab) {
                const long base = *P.nsamp, row = base + (long)b * S.nC + c;
                if (base + (long)n * S.nC <= P.maxsamp) P.lab[row * 16 + lane] = lv;
            }
        }
        // f(x) = sum_{l,r} comb[l + 4r] x_l x_r through the symmetric part of
        // the 4 x 4 combo matrix: hs = (C_ll, C_lr + C_rl for r > l), row by row
        double hs[10];
        {
            int k = 0;
#pragma unroll
            for (int l = 0; l < 4; ++l)
#pragma unroll
                for (int r = l; r < 4; ++r) hs[k++] = r == l ? cb[5 * l] : cb[l + 4 * r] + cb[r + 4 * l];
        }
        auto quad_s = [&](const double *x) {
            double f = 0.0;
            int k = 0;
#pragma unroll
            for (int l = 0; l < 4; ++l) {
                double t = 0.0;
#pragma unroll
                for (int r = l; r < 4; ++r) t += hs[k++] * x[r];
                f += x[l] * t;
            }
            return f;
        };
        double pv[4], loc = 0.0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int gq = 4 * lane + t;
            double v = 0.0;
            if (gq < Ng) {
                const double f = quad_s(grid_s + 4 * gq);
                v = (f > 0.0 ? f : 0.0) * pref;
            }
            pv[t] = v;
            loc += v;
        }
        // one scan of the unnormalised PDF gives its sum (lane 63) and the CDF;
        // cum_g <= u is tested as cum_g * norm <= u * norm, and the count of
        // grid points that pass is a sum of four ballots
        const double incl = wave_inclusive_scan(loc);
        const double pdfnorm = readlane_d(incl, 63);
        const double ut = u * pdfnorm;
        double run = dpp_d<0x138>(incl);  // exclusive prefix (wave_shr:1)
        int gs = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            run += pv[t];
            gs += __popcll(__ballot(4 * lane + t < Ng - 1 && run <= ut));
        }
        if constexpr (LEARN) {
            // learned proposal q from the predicted labels; weight pdf(x)/q(x)
            double ph[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) ph[q] = (double)__int_as_float(__builtin_amdgcn_readlane(__float_as_int(plab), q));
            auto quad = [&](const double *cf, const double *x) {
                double f = 0.0;
#pragma unroll
                for (int l = 0; l < 4; ++l) {
                    double t = 0.0;
#pragma unroll
                    for (int r = 0; r < 4; ++r) t += cf[l + 4 * r] * x[r];
                    f += x[l] * t;
                }
                return f;
            };
            double qv[4], qloc = 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int gq = 4 * lane + t;
                const double v = gq < Ng ? fmax(quad(ph, grid_s + 4 * gq), 0.0) : 0.0;
                qv[t] = v;
                qloc += v;
            }
            const double qs = wave_sum(qloc);
            const double a1 = qs > 0.0 ? (1.0 - P.mix) / qs : 0.0, a0 = qs > 0.0 ? P.mix / Ng : 1.0 / Ng;
            double ql = 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                qv[t] = 4 * lane + t < Ng ? a1 * qv[t] + a0 : 0.0;
                ql += qv[t];
            }
            double qrun = wave_exclusive_scan(ql);
            int gl = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                qrun += qv[t];
                gl += __popcll(__ballot(4 * lane + t < Ng - 1 && qrun <= u));
            }
            const double *xl = grid_s + 4 * gl;
            const double vl = fmax(quad(cb, xl), 0.0) * pref;
            const double qgl = a1 * fmax(quad(ph, xl), 0.0) + a0;
            if (vl > 0.0) {
                gs = gl;
                wt *= vl / qgl;
            } else {
                wt = 0.0;  // continues with the reference's own sample
            }
        } else {
            wt *= pdfnorm;
        }
        if (threadIdx.x == 0) {
            if (P.idx) P.idx[(size_t)b * S.nC + c] = gs;
            if (!(pdfnorm > 0.0) || !isfinite(pdfnorm)) atomicOr(P.status, 1);
        }
        const double *xs = grid_s + 4 * gs;
        const double fstar = quad_s(xs);  // <trial|B(x*) phi> / 4
        double Ux[4] = {0.0, 0.0, 0.0, 0.0}, Vx[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const double pf = 2.0 * sS[kk] * xs[kk];  // sqrt(4 S_k) x_k (MPSstate.cpp:591,616)
#pragma unroll
            for (int up = 0; up < 2; ++up)
#pragma unroll
                for (int dn = 0; dn < 2; ++dn) {
                    Ux[up + 2 * dn] += pf * tU[up + 2 * (dn + 2 * kk)];
                    Vx[up + 2 * dn] += pf * tVT[kk + 4 * (up + 2 * dn)];
                }
        }
        {  // right site scaled so that the overlap stays O
            const double alpha = O / (4.0 * fstar);
#pragma unroll
            for (int q = 0; q < 4; ++q) Vx[q] *= alpha;
        }
        // stage F (elementwise): L_{i+1} = sum_{s',s} Ux(s',s) G(s',s) (LDS and global),
        // phi'_i = Ux phi_i to global, phi'_j = Vx phi_j in place (the next
        // coupling's left site); the next coupling's right site and R_{j+2}
        // from the registers into the free pair and B_R
        {
            const float *G = C.blk(B_Y);
            float *Ln = C.blk(B_L), *lg = C.Lg + (size_t)(i + 1) * GB;
            float *pg0 = C.ph(i, 0), *pg1 = C.ph(i, 1);
            const float u0 = (float)Ux[0], u1 = (float)Ux[1], u2 = (float)Ux[2], u3 = (float)Ux[3];
            // every LDS operand of the thread's four elements is loaded before
            // the first store (a store to B_L could otherwise not pass the loads)
            constexpr int NE = GB / MQE_NT;
            float gv[NE][4], av[NE][2];
#pragma unroll
            for (int it = 0; it < NE; ++it) {
                const int e = threadIdx.x + it * MQE_NT, o = (e & 31) + LD * (e >> 5);
#pragma unroll
                for (int q = 0; q < 4; ++q) gv[it][q] = G[q * BLK + o];
                av[it][0] = cen[o];
                av[it][1] = cen[BLK + o];
            }
#pragma unroll
            for (int it = 0; it < NE; ++it) {
                const int e = threadIdx.x + it * MQE_NT, o = (e & 31) + LD * (e >> 5);
                const float v = u0 * gv[it][0] + u2 * gv[it][1] + u1 * gv[it][2] + u3 * gv[it][3];
                Ln[o] = v;
                lg[e] = v;
                pg0[e] = u0 * av[it][0] + u2 * av[it][1];
                pg1[e] = u1 * av[it][0] + u3 * av[it][1];
            }
            apply2(nxt, Vx);
            if (c + 1 < S.nC) {
                float *const d[3] = {C.blk(pf), C.blk(pf) + BLK, C.blk(B_R)};
                pre_store<3>(pre, d);
                if (threadIdx.x < 36) {
                    tab_s[(c + 1) & 1][threadIdx.x] = tnext;
                    if (c + 2 < S.nC) tnext = table_entry(c + 2, threadIdx.x);
                }
            }
        }
        __syncthreads();
        const int t = pc; pc = pj; pj = pf; pf = t;
    }
    if (pc == B_BUF2) {  // the energy sweep takes the center in B_BUF0 / B_BUF1
        const float *src = C.blk(B_BUF2);
        float *dst = C.blk(B_BUF0);
        for (int e = threadIdx.x; e < 2 * BLK; e += MQE_NT) dst[e] = src[e];
        pc = B_BUF0;
        __syncthreads();
    }
    // ---------------- local energy (Walker.cpp:72-77) ----------------
    double On, Ou;
    const bool canon = canon_walker(ctl, slot, P.since[slot]);
    const double e = energy_sweep32(C, pc, &On, &Ou, canon);
    if (threadIdx.x == 0) {
        P.since[slot] = canon ? 0 : P.since[slot] + 1;
        P.w[slot] = wt;
        P.O[slot] = On;
        P.EL[slot] = e / Ou;
        if (!isfinite(wt)) atomicOr(P.status, 2);
    }
}

}  // namespace t32
}  // namespace mqe

#endif
