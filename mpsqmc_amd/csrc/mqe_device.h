// mqe_device.h -- workgroup-cooperative device primitives of the walker engine.
//
// One 256-thread workgroup (four wave64s) owns one walker.  All matrices are
// column-major with explicit leading dimensions and are addressed through
// generic pointers, so the same routines run on global (L2-resident) scratch
// and on LDS.
#ifndef MQE_DEVICE_H
#define MQE_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MQE_NT 256
#define MQE_WAVES (MQE_NT / 64)

namespace mqe {

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// Data-parallel-primitive (DPP) lane exchange of a double within rows of 16
// lanes; CTRL is the DPP control word (quad_perm 0x00-0xFF, row_mirror 0x140,
// row_half_mirror 0x141).  All lanes of the row must be active.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    const int lo2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, true);
    const int hi2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi2, lo2);
}

__device__ __forceinline__ double readlane_d_(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// sum over aligned groups of 8 lanes (xor 1, xor 2 inside quads, then the
// mirrored half-row), every lane of the group receives the sum
__device__ __forceinline__ double group8_sum(double v) {
    v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);  // row_half_mirror
    return v;
}

// DPP move of a double with a row mask; lanes of disabled rows receive 0
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_dm(double x) {
    const int lo = __double2loint(x), hi = __double2hiint(x);
    const int lo2 = __builtin_amdgcn_update_dpp(0, lo, CTRL, ROWS, 0xF, false);
    const int hi2 = __builtin_amdgcn_update_dpp(0, hi, CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi2, lo2);
}

// inclusive prefix sum over the 64 lanes with DPP only: row_shr 1, 2, 4, 8
// inside each 16-lane row, then row_bcast:15 (rows 1, 3) and row_bcast:31
// (rows 2, 3) carry the row totals (lane semantics checked on gfx950)
__device__ __forceinline__ double wave_inclusive_scan(double v) {
    v += dpp_d<0x111>(v);
    v += dpp_d<0x112>(v);
    v += dpp_d<0x114>(v);
    v += dpp_d<0x118>(v);
    v += dpp_dm<0x142, 0xA>(v);
    v += dpp_dm<0x143, 0xC>(v);
    return v;
}

// sum over the wave64; every lane receives the total (lane 63 of the scan)
__device__ __forceinline__ double wave_sum(double v) {
    return readlane_d_(wave_inclusive_scan(v), 63);
}

// value of v in lane l (l wave-uniform)
__device__ __forceinline__ double readlane_d(double v, int l) { return readlane_d_(v, l); }

// exclusive prefix sum over the lanes of the wave64 (lane 0 receives 0):
// the inclusive scan shifted by one lane (DPP wave_shr:1)
__device__ __forceinline__ double wave_exclusive_scan(double v) {
    return dpp_d<0x138>(wave_inclusive_scan(v));
}

// Sum over the workgroup; every thread receives the total.  sh: >= MQE_WAVES doubles.
__device__ __forceinline__ double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < MQE_WAVES; ++w) t += sh[w];
    return t;
}

// K simultaneous sums; sh: >= K*MQE_WAVES doubles.
template <int K>
__device__ __forceinline__ void block_sum_n(double (&v)[K], double *sh) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) sh[wid * K + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < MQE_WAVES; ++w) t += sh[w * K + k];
        v[k] = t;
    }
}

// Inclusive prefix sum of x[0..n) into y[0..n) (double), whole workgroup.
// sh: >= MQE_NT + MQE_WAVES doubles.
__device__ __forceinline__ void block_inclusive_scan(int n, const double *x, double *y, double *sh) {
    const int tid = threadIdx.x;
    const int per = (n + MQE_NT - 1) / MQE_NT;
    const int b0 = tid * per, b1 = min(n, b0 + per);
    double loc = 0.0;
    for (int i = b0; i < b1; ++i) loc += x[i];
    // wave inclusive scan of the per-thread totals
    const int lane = tid & 63, wid = tid >> 6;
    double s = loc;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double o = __shfl_up(s, off, 64);
        if (lane >= off) s += o;
    }
    __syncthreads();
    if (lane == 63) sh[MQE_NT + wid] = s;
    __syncthreads();
    double base = 0.0;
    for (int w = 0; w < wid; ++w) base += sh[MQE_NT + w];
    double run = base + s - loc;  // exclusive prefix of this thread's chunk
    for (int i = b0; i < b1; ++i) { run += x[i]; y[i] = run; }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// FP32 GEMM on the matrix cores: C(m x n) = sum_{s < nseg} op(A_s) op(B_s)
// (nseg = 1 or 2 terms of one shape and leading dimensions), for nb products
// at once (A_s + b sA and so on), whole workgroup.  The output is cut into
// tiles, one wave per tile at a time; the tiles of all nb products are shared
// among the four waves.  Each tile is a chain of MFMAs over the k slices of
// its terms.  A lane's operand values of a slice are one run along k, loaded
// straight into the MFMA operand registers without predication: rows and
// columns beyond the matrix are clamped to its last row / column (they only
// feed output elements that are not stored), so the wait for a slice's data
// sits right before its MFMAs, and the loads of the wave's next slice (of this
// tile or the first of its next tile) are issued before the MFMAs of the
// current one.  k not a multiple of the slice depth (16) takes a predicated
// path with zeros outside the matrices.  acc: C += instead of C =.
// ---------------------------------------------------------------------------

// Run of NV operand values at k = kb..kb+NV-1 of one row of op(X), inside the
// matrix: KC -- stored along k (p[kk]), else p[ld kk].  vec: KC with p + kb
// 16-byte aligned.
template <bool KC, int NV>
__device__ __forceinline__ void load_run(const float *p, int ld, int kb, bool vec, float (&v)[NV]) {
    if (KC && vec) {
        const float4 *q = (const float4 *)(p + kb);
#pragma unroll
        for (int j = 0; j < NV / 4; ++j) {
            const float4 x = q[j];
            v[4 * j] = x.x;
            v[4 * j + 1] = x.y;
            v[4 * j + 2] = x.z;
            v[4 * j + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int u = 0; u < NV; ++u) v[u] = KC ? p[kb + u] : p[(size_t)ld * (kb + u)];
    }
}

// The same run with zeros at k >= kmax.
template <bool KC, int NV>
__device__ __forceinline__ void load_run_pred(const float *p, int ld, int kb, int kmax, float (&v)[NV]) {
#pragma unroll
    for (int u = 0; u < NV; ++u) {
        const int kk = kb + u;
        v[u] = kk < kmax ? (KC ? p[kk] : p[(size_t)ld * kk]) : 0.f;
    }
}

// 16-byte loads along k need 16-byte aligned bases, leading dimension and batch stride
__device__ __forceinline__ bool aligned16(const float *X0, const float *X1, int ld, long s) {
    return ((((uintptr_t)X0) | ((uintptr_t)X1) | ((uintptr_t)ld * 4u) | ((uintptr_t)s * 4u)) & 15u) == 0;
}

// TS x TS output tiles (TS = 32: v_mfma_f32_32x32x2f32, TS = 16:
// v_mfma_f32_16x16x4f32), NW adjacent tiles along n per wave sharing the A
// run (NW = 2 only with TS = 32).  16-deep slices; within the slice at k0,
// MFMA u takes kk = k0 + KL (l / TS) + u in lane l (KL = 8 for TS = 32, 4 for
// TS = 16), which holds A(l % TS, kk) and B(kk, l % TS).  Accumulator slot q
// of lane l is element (8(q/4) + 4(l/32) + q%4, l % 32) of a 32 x 32 tile,
// (4(l/16) + q, l % 16) of a 16 x 16 tile.
template <int TS>
struct TileGeom {
    static constexpr int KL = TS == 32 ? 8 : 4;  // k values per lane and slice
    static constexpr int NACC = TS == 32 ? 16 : 4;
};

// Lane operand pointers of tile tb (of all nb products) and term seg: the
// lane's row of op(A) and its NW columns of op(B), clamped into the matrices.
template <bool TA, bool TB, int TS, int NW>
__device__ __forceinline__ void tile_ptrs(int tb, int seg, int m, int n, int tm, int tn, const float *A0,
                                          const float *A1, int lda, const float *B0, const float *B1, int ldb, long sA,
                                          long sB, int r, const float *&pa, const float *(&pb)[NW]) {
    const int bt = tb / (tm * tn), t = tb - bt * (tm * tn);
    const int ar = min((t % tm) * TS + r, m - 1);
    const float *A = (seg ? A1 : A0) + (size_t)bt * sA, *B = (seg ? B1 : B0) + (size_t)bt * sB;
    pa = TA ? A + (size_t)lda * ar : A + ar;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        const int bc = min((t / tm) * TS * NW + TS * q + r, n - 1);
        pb[q] = TB ? B + bc : B + (size_t)ldb * bc;
    }
}

template <int TS, int NW>
__device__ __forceinline__ void tile_store(float *C0, int ldc, long sC, int tb, int m, int n, int tm, int tn, int r,
                                           int h, bool acc,
                                           const float (&c)[NW][TileGeom<TS>::NACC]) {
    constexpr int NACC = TileGeom<TS>::NACC;
    const int bt = tb / (tm * tn), t = tb - bt * (tm * tn);
    const int r0 = (t % tm) * TS;
    float *C = C0 + (size_t)bt * sC;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        const int bc = (t / tm) * TS * NW + TS * q + r;
        if (bc < n) {
#pragma unroll
            for (int e = 0; e < NACC; ++e) {
                const int rr = TS == 32 ? r0 + 8 * (e >> 2) + 4 * h + (e & 3) : r0 + 4 * h + e;
                if (rr < m) {
                    float v = c[q][e];
                    if (acc) v += C[rr + (size_t)ldc * bc];
                    C[rr + (size_t)ldc * bc] = v;
                }
            }
        }
    }
}

template <int TS>
__device__ __forceinline__ void mfma_step(float a, float b, float (&c)[TileGeom<TS>::NACC]) {
    typedef float vacc __attribute__((ext_vector_type(TileGeom<TS>::NACC)));
    vacc x;
#pragma unroll
    for (int e = 0; e < TileGeom<TS>::NACC; ++e) x[e] = c[e];
    if constexpr (TS == 32)
        x = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, x, 0, 0, 0);
    else
        x = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, x, 0, 0, 0);
#pragma unroll
    for (int e = 0; e < TileGeom<TS>::NACC; ++e) c[e] = x[e];
}

// k a positive multiple of 16: every wave runs one pipelined chain over the
// (tile, term, slice) items of its tiles.  The operands of the item two ahead
// are loaded unconditionally (past the end the last item reloads itself) with
// a fixed instruction sequence, so the wait for an item's operands covers only
// its own loads.
// VEC: the operands stored along k are 16-byte aligned (float4 loads).
template <bool TA, bool TB, int TS, int NW, bool VEC>
__device__ __forceinline__ void gemm_tiles_full(int m, int n, int k, const float *A0, const float *A1, int lda,
                                             const float *B0, const float *B1, int ldb, float *C0, int ldc, bool acc,
                                             int nseg, int nb, long sA, long sB, long sC) {
    constexpr int KL = TileGeom<TS>::KL, NACC = TileGeom<TS>::NACC;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane % TS, h = lane / TS;
    const int tm = (m + TS - 1) / TS, tn = (n + TS * NW - 1) / (TS * NW), ntile = nb * tm * tn;
    if (w >= ntile) return;
    const int ntw = (ntile - w + MQE_WAVES - 1) / MQE_WAVES;  // tiles of this wave
    const int nsk = k >> 4, nsl = nseg * nsk, nit = ntw * nsl;
    auto load = [&](int i, float (&a)[KL], float (&b)[NW][KL]) {
        const int j = i / nsl, rem = i - j * nsl, seg = rem / nsk, sl = rem - seg * nsk;
        const float *pa, *pb[NW];
        tile_ptrs<TA, TB, TS, NW>(w + MQE_WAVES * j, seg, m, n, tm, tn, A0, A1, lda, B0, B1, ldb, sA, sB, r, pa, pb);
        const int kb = 16 * sl + KL * h;
        load_run<TA, KL>(pa, lda, kb, VEC, a);
#pragma unroll
        for (int q = 0; q < NW; ++q) load_run<!TB, KL>(pb[q], ldb, kb, VEC, b[q]);
    };
    // three operand sets in rotation (the loop is unrolled by three, so no
    // register copies wait on loads in flight): the loads run two items ahead
    // of the MFMAs
    float a0[KL], b0[NW][KL], a1[KL], b1[NW][KL], a2[KL], b2[NW][KL], c[NW][NACC];
#pragma unroll
    for (int q = 0; q < NW; ++q)
#pragma unroll
        for (int e = 0; e < NACC; ++e) c[q][e] = 0.f;
    auto step = [&](int i, const float (&a)[KL], const float (&b)[NW][KL]) {
#pragma unroll
        for (int u = 0; u < KL; ++u)
#pragma unroll
            for (int q = 0; q < NW; ++q) mfma_step<TS>(a[u], b[q][u], c[q]);
        if ((i + 1) % nsl == 0) {  // the tile is complete
            tile_store<TS, NW>(C0, ldc, sC, w + MQE_WAVES * (i / nsl), m, n, tm, tn, r, h, acc, c);
#pragma unroll
            for (int q = 0; q < NW; ++q)
#pragma unroll
                for (int e = 0; e < NACC; ++e) c[q][e] = 0.f;
        }
    };
    load(0, a0, b0);
    load(min(1, nit - 1), a1, b1);
    for (int i = 0; i < nit; i += 3) {
        load(min(i + 2, nit - 1), a2, b2);
        __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the MFMAs
        step(i, a0, b0);
        if (i + 1 >= nit) break;
        load(min(i + 3, nit - 1), a0, b0);
        __builtin_amdgcn_sched_barrier(0);
        step(i + 1, a1, b1);
        if (i + 2 >= nit) break;
        load(min(i + 4, nit - 1), a1, b1);
        __builtin_amdgcn_sched_barrier(0);
        step(i + 2, a2, b2);
    }
}

// Any k: per tile and slice, predicated loads (zeros at k >= k), no pipelining.
template <bool TA, bool TB, int TS, int NW>
__device__ __forceinline__ void gemm_tiles_pred(int m, int n, int k, const float *A0, const float *A1, int lda,
                                             const float *B0, const float *B1, int ldb, float *C0, int ldc, bool acc,
                                             int nseg, int nb, long sA, long sB, long sC) {
    constexpr int KL = TileGeom<TS>::KL, NACC = TileGeom<TS>::NACC;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane % TS, h = lane / TS;
    const int tm = (m + TS - 1) / TS, tn = (n + TS * NW - 1) / (TS * NW), ntile = nb * tm * tn;
    for (int tb = w; tb < ntile; tb += MQE_WAVES) {
        float c[NW][NACC];
#pragma unroll
        for (int q = 0; q < NW; ++q)
#pragma unroll
            for (int e = 0; e < NACC; ++e) c[q][e] = 0.f;
        for (int seg = 0; seg < nseg; ++seg) {
            const float *pa, *pb[NW];
            tile_ptrs<TA, TB, TS, NW>(tb, seg, m, n, tm, tn, A0, A1, lda, B0, B1, ldb, sA, sB, r, pa, pb);
            for (int k0 = 0; k0 < k; k0 += 16) {
                float a[KL], b[NW][KL];
                load_run_pred<TA, KL>(pa, lda, k0 + KL * h, k, a);
#pragma unroll
                for (int q = 0; q < NW; ++q) load_run_pred<!TB, KL>(pb[q], ldb, k0 + KL * h, k, b[q]);
#pragma unroll
                for (int u = 0; u < KL; ++u)
#pragma unroll
                    for (int q = 0; q < NW; ++q) mfma_step<TS>(a[u], b[q][u], c[q]);
            }
        }
        tile_store<TS, NW>(C0, ldc, sC, tb, m, n, tm, tn, r, h, acc, c);
    }
}

template <bool TA, bool TB, int TS, int NW>
__device__ __forceinline__ void gemm_tiles(int m, int n, int k, const float *A0, const float *A1, int lda,
                                           const float *B0, const float *B1, int ldb, float *C0, int ldc, bool acc,
                                           int nseg, int nb, long sA, long sB, long sC) {
    // the pipelined path needs whole slices and 16-byte aligned operands stored
    // along k (bond dimensions that are multiples of 16 give both)
    const bool vec = (!TA || aligned16(A0, nseg > 1 ? A1 : A0, lda, sA)) &&
                     (TB || aligned16(B0, nseg > 1 ? B1 : B0, ldb, sB));
    if (k >= 16 && (k & 15) == 0 && vec) {
        gemm_tiles_full<TA, TB, TS, NW, true>(m, n, k, A0, A1, lda, B0, B1, ldb, C0, ldc, acc, nseg, nb, sA, sB, sC);
    } else {
        gemm_tiles_pred<TA, TB, TS, NW>(m, n, k, A0, A1, lda, B0, B1, ldb, C0, ldc, acc, nseg, nb, sA, sB, sC);
    }
}

// Single-term products with the original entry points (also used by the
// proposal network).
template <bool TA, bool TB>
__device__ __forceinline__ void gemm_mfma(int m, int n, int k, const float *A0, int lda, const float *B0, int ldb,
                                          float *C0, int ldc, bool acc, int nb = 1, long sA = 0, long sB = 0,
                                          long sC = 0) {
    gemm_tiles<TA, TB, 32, 1>(m, n, k, A0, A0, lda, B0, B0, ldb, C0, ldc, acc, 1, nb, sA, sB, sC);
}

// Output tile shapes of the FP32 matrix-core products (chosen per walker bond
// bucket by the engine's tile table).
enum : int { TILE_16 = 16, TILE_32 = 32, TILE_WIDE = 64 };

// The matrix-core product with the output tile shape `tile`.
template <bool TA, bool TB>
__device__ __forceinline__ void gemm_f32(int m, int n, int k, const float *A0, const float *A1, int lda,
                                         const float *B0, const float *B1, int ldb, float *C, int ldc, bool acc,
                                         int nseg, int nb, long sA, long sB, long sC, int tile) {
    if (tile == TILE_16)
        gemm_tiles<TA, TB, 16, 1>(m, n, k, A0, A1, lda, B0, B1, ldb, C, ldc, acc, nseg, nb, sA, sB, sC);
    else if (tile == TILE_WIDE && n > 32)
        gemm_tiles<TA, TB, 32, 2>(m, n, k, A0, A1, lda, B0, B1, ldb, C, ldc, acc, nseg, nb, sA, sB, sC);
    else
        gemm_tiles<TA, TB, 32, 1>(m, n, k, A0, A1, lda, B0, B1, ldb, C, ldc, acc, nseg, nb, sA, sB, sC);
}

// Vector-unit product (FP64, and FP32 products too small for the matrix cores).
template <bool TA, bool TB, typename T>
__device__ __forceinline__ void gemm_valu(int m, int n, int k, const T *A0, const T *A1, int lda, const T *B0,
                                          const T *B1, int ldb, T *C, int ldc, bool acc, int nseg, int nb, long sA,
                                          long sB, long sC) {
    const int mn = m * n;
    for (int e = threadIdx.x; e < nb * mn; e += MQE_NT) {
        const int bt = e / mn, x = e - bt * mn;
        const int r = x % m, c = x / m;
        T *Cb = C + (size_t)bt * sC;
        T s = acc ? Cb[r + (size_t)ldc * c] : T(0);
        for (int sg = 0; sg < nseg; ++sg) {
            const T *Ab = (sg ? A1 : A0) + (size_t)bt * sA, *Bb = (sg ? B1 : B0) + (size_t)bt * sB;
            const T *ap = TA ? (Ab + (size_t)lda * r) : (Ab + r);
            const T *bp = TB ? (Bb + c) : (Bb + (size_t)ldb * c);
            for (int q = 0; q < k; ++q) {
                const T av = TA ? ap[q] : ap[(size_t)lda * q];
                const T bv = TB ? bp[(size_t)ldb * q] : bp[q];
                s += av * bv;
            }
        }
        Cb[r + (size_t)ldc * c] = s;
    }
}

// ---------------------------------------------------------------------------
// nb products of one shape, C_b (+)= op(A_b) op(B_b) with A_b = A + b sA (and so
// on), as one workgroup call: the output tiles of all of them are shared among
// the four waves (small products keep every wave busy) and one barrier after
// the call covers the batch.
// ---------------------------------------------------------------------------
template <bool TA, bool TB, typename T>
__device__ __forceinline__ void gemm_batch(int nb, int m, int n, int k, const T *A, int lda, long sA, const T *B,
                                           int ldb, long sB, T *C, int ldc, long sC, bool acc, int tile = TILE_32) {
    if constexpr (sizeof(T) == sizeof(float)) {
        if (m >= 8 && n >= 8 && k >= 4) {
            gemm_f32<TA, TB>(m, n, k, (const float *)A, (const float *)A, lda, (const float *)B, (const float *)B, ldb,
                             (float *)C, ldc, acc, 1, nb, sA, sB, sC, tile);
            return;
        }
    }
    gemm_valu<TA, TB, T>(m, n, k, A, A, lda, B, B, ldb, C, ldc, acc, 1, nb, sA, sB, sC);
}

// Two-term product C = op(A0) op(B0) + op(A1) op(B1) (the sums over a physical
// index): one accumulation chain per output tile.
template <bool TA, bool TB, typename T>
__device__ __forceinline__ void gemm2(int m, int n, int k, const T *A0, const T *A1, int lda, const T *B0,
                                      const T *B1, int ldb, T *C, int ldc, int tile = TILE_32, int nb = 1, long sA = 0,
                                      long sB = 0, long sC = 0) {
    if constexpr (sizeof(T) == sizeof(float)) {
        if (m >= 8 && n >= 8 && k >= 4) {
            gemm_f32<TA, TB>(m, n, k, (const float *)A0, (const float *)A1, lda, (const float *)B0, (const float *)B1,
                             ldb, (float *)C, ldc, false, 2, nb, sA, sB, sC, tile);
            return;
        }
    }
    gemm_valu<TA, TB, T>(m, n, k, A0, A1, lda, B0, B1, ldb, C, ldc, false, 2, nb, sA, sB, sC);
}

// ---------------------------------------------------------------------------
// GEMM: C(m x n) (+)= op(A) op(B).  op(A) is m x k: A(r,q) = TA ? A[q + lda*r] : A[r + lda*q].
// FP32 products of useful size run on the matrix cores with the output tile
// shape `tile`, the rest (and all of FP64) on the vector units.
// ---------------------------------------------------------------------------
template <bool TA, bool TB, typename T>
__device__ __forceinline__ void gemm(int m, int n, int k, const T *A, int lda, const T *B, int ldb, T *C, int ldc,
                                     bool acc, int tile = TILE_32) {
    gemm_batch<TA, TB, T>(1, m, n, k, A, lda, 0, B, ldb, 0, C, ldc, 0, acc, tile);
}

// ---------------------------------------------------------------------------
// Two chained FP32 matrix-core products with the intermediate in registers:
//   F_f = A op(B_f)                      (m1 x n1, inner dimension k1)
//   ND = 1:  D_0 = sum_f op(C_f) F_f     (f < NF, NF = 1, 2)
//   ND = 2:  D_d = op(C_d) F_0           (d < 2, NF = 1: two outputs of one F)
// (D_d: mD x n1; inner dimension m1 <= 64, mD <= 64) for nu independent units
// (operands of unit u from get(u)), whole workgroup, no barrier inside.  A
// wave takes (unit, 32-column strip) items: for each f it forms the strip of
// F_f in two 32 x 32 accumulator tiles (rows >= m1 zeroed) and feeds the
// accumulator slots straight into the second chain as its B operand,
// accumulating the (at most two) row tiles of each D strip in registers.  The
// second chain takes its inner index in the order the accumulators hold it:
// MFMA q of F row tile t uses rows 32t + 8(q/4) + 4h + q%4 (h = lane half),
// exactly accumulator slot q of lane (column, h), so no lane exchange is
// needed; op(C) is loaded in the same order.
// Items go to waves round-robin starting at wave0; the return value is the
// wave that takes the next item, so independent calls share the waves
// between two barriers.
// TB: B_f stored n1 x k1 (op(B) = B^T).  TC: C stored m1 x mD (op(C) = C^T,
// contiguous along the inner index), else mD x m1.  All matrices column-major.
// ---------------------------------------------------------------------------
struct ChainOps {
    const float *A;      // m1 x k1, leading dimension lda
    const float *B[2];   // B_f
    const float *C[2];   // C_f (ND = 1) or C_d (ND = 2)
    float *D[2];         // D_d, mD x n1, leading dimension ldd
};

typedef float v16f_c __attribute__((ext_vector_type(16)));

__device__ __forceinline__ v16f_c mfma32(float a, float b, v16f_c c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// Buffer resources over one walker's or the trial's arrays: the 64-bit base in
// scalar registers and 32-bit byte offsets per lane.  An offset at or beyond
// BUF_RANGE reads 0 and drops a store (the hardware range check).
constexpr int BUF_RANGE = 0x7ffffff0;
constexpr int BUF_OOB = 0x7ffffff8;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, BUF_RANGE, 0x00020000);
}
// (the b32 / b128 builtins move raw 32-bit words: bit casts, not conversions)
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, int voff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, int voff, int imm) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, imm, 0);
}
__device__ __forceinline__ void buf_ld4(__amdgpu_buffer_rsrc_t r, int voff, float *v) {
    const float4 x = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0));
    v[0] = x.x;
    v[1] = x.y;
    v[2] = x.z;
    v[3] = x.w;
}

// qsplit: the two row tiles of D are separate items (the F strips formed for
// each), for calls with too few items to keep the four waves busy.
template <bool TB, bool TC, int NF, int ND, class Get>
__device__ __forceinline__ int chain2(int wave0, int nu, int m1, int n1, int k1, int mD, int lda, int ldb, int ldc,
                                      int ldd, bool qsplit, Get get) {
    static_assert(ND == 1 || NF == 1, "two outputs take one intermediate");
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool two = m1 > 32;     // second row tile of F
    const bool twoD = mD > 32;    // second row tile of D
    const int nq = qsplit && twoD ? 2 : 1;
    const int nstrip = (n1 + 31) >> 5, nitems = nu * nstrip * nq;
    const int sa = 4 * lda, sb = TB ? 4 * ldb : 4;  // bytes per k step of A and op(B)
    const int kf = k1 & ~15;      // whole 16-deep slices
#pragma nounroll
    for (int it = (w - wave0 + 2 * MQE_WAVES) % MQE_WAVES; it < nitems; it += MQE_WAVES) {
        const int qi = it % nq, rest = it / nq;
        const int u = rest / nstrip, c0 = (rest - u * nstrip) * 32;
        const bool do0 = nq == 1 || qi == 0, do1 = twoD && (nq == 1 || qi == 1);  // D row tiles of this item
        const ChainOps op = get(u);
        // the lane coordinates as opaque values inside the loops: the per-element
        // offsets below are then formed next to their loads instead of being
        // hoisted out of the loops and held in registers throughout
        int r = lane & 31, h = lane >> 5;
        v16f_c D0, D1;  // the (at most two) row tiles of the D strip being formed
#pragma nounroll
        for (int f = 0; f < NF; ++f) {
            asm volatile("" : "+v"(r), "+v"(h));
            const int col = min(c0 + r, n1 - 1);
            const int vA0 = 4 * min(r, m1 - 1), vA1 = 4 * min(32 + r, m1 - 1);
            const int vB = TB ? 4 * col : 4 * ldb * col;
            const int row0 = min(r, mD - 1), row1 = min(32 + r, mD - 1);
            // ---- first chain: the F_f strip; lane (r, h) takes k = k0 + 8h + q of the
            // slice at k0; in the last, partial slice k >= k1 reads 0 ----
            v16f_c F0, F1;
#pragma unroll
            for (int e = 0; e < 16; ++e) F0[e] = F1[e] = 0.f;
            const auto rA = buf_rsrc(op.A), rB = buf_rsrc(op.B[f]);
            auto slice = [&](int k0, bool tail) {
                float a0[8], a1[8], bv[8];
                const int kl = k0 + 8 * h;
                // tail: k >= k1 reads 0 (offset beyond the range)
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const bool ok = !tail || kl + q < k1;
                    a0[q] = buf_ld(rA, ok ? vA0 + sa * (kl + q) : BUF_OOB);
                    bv[q] = buf_ld(rB, ok ? vB + sb * (kl + q) : BUF_OOB);
                }
                if (two) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) a1[q] = buf_ld(rA, !tail || kl + q < k1 ? vA1 + sa * (kl + q) : BUF_OOB);
                }
#pragma unroll
                for (int q = 0; q < 8; ++q) F0 = mfma32(a0[q], bv[q], F0);
                if (two) {
#pragma unroll
                    for (int q = 0; q < 8; ++q) F1 = mfma32(a1[q], bv[q], F1);
                }
            };
#pragma nounroll
            for (int k0 = 0; k0 < kf; k0 += 16) slice(k0, false);
            if (kf < k1) slice(kf, true);
            if (m1 & 31) {  // rows of F beyond m1 (clamped A rows) must not contribute
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int row = 8 * (e >> 2) + 4 * h + (e & 3);
                    if (row >= m1) F0[e] = 0.f;
                    if (32 + row >= m1) F1[e] = 0.f;
                }
            }
            // ---- second chain: D_d (+)= op(C) F_f, row tile by row tile; a D strip
            // is stored once its last term is in ----
#pragma nounroll
            for (int d = 0; d < ND; ++d) {
                asm volatile("" : "+v"(r), "+v"(h));
                if (ND == 2 || f == 0) {
#pragma unroll
                    for (int e = 0; e < 16; ++e) D0[e] = D1[e] = 0.f;
                }
                const float *Cp = op.C[ND == 1 ? f : d];
                const auto rC = buf_rsrc(Cp);
                const bool vecC = TC && (m1 & 3) == 0 && (ldc & 3) == 0 && aligned16(Cp, Cp, 0, 0);
                auto chunk = [&](int row, int t, float (&cv)[16]) {
                    if (TC && vecC) {
                        // a run at or beyond m1 meets zero rows of F: any values of C do
#pragma unroll
                        for (int g = 0; g < 4; ++g)
                            buf_ld4(rC, 4 * (ldc * row + min(32 * t + 8 * g + 4 * h, m1 - 4)), cv + 4 * g);
                    } else {
#pragma unroll
                        for (int e = 0; e < 16; ++e) {
                            const int k = min(32 * t + 8 * (e >> 2) + 4 * h + (e & 3), m1 - 1);
                            cv[e] = buf_ld(rC, TC ? 4 * (ldc * row + k) : 4 * (row + ldc * k));
                        }
                    }
                };
                auto rowtile = [&](int row, v16f_c &Dq) {
                    float c0v[16], c1v[16];
                    chunk(row, 0, c0v);
                    if (two) chunk(row, 1, c1v);
#pragma unroll
                    for (int q = 0; q < 16; ++q) Dq = mfma32(c0v[q], F0[q], Dq);
                    if (two) {
#pragma unroll
                        for (int q = 0; q < 16; ++q) Dq = mfma32(c1v[q], F1[q], Dq);
                    }
                };
                if (do0) rowtile(row0, D0);
                if (do1) rowtile(row1, D1);
                if (f == NF - 1) {
                    // rows 8(e/4) + 4h + e%4 (+32) of column c0 + r; outside D the store is dropped
                    const int vd = c0 + r < n1 ? 4 * (ldd * (c0 + r) + 4 * h) : BUF_OOB;
                    const auto rD = buf_rsrc(op.D[d]);
                    if (do0) {
#pragma unroll
                        for (int e = 0; e < 16; ++e) {
                            const int rr = 8 * (e >> 2) + 4 * h + (e & 3), imm = 4 * (8 * (e >> 2) + (e & 3));
                            buf_st(D0[e], rD, rr < mD ? vd : BUF_OOB, imm);
                        }
                    }
                    if (do1) {
#pragma unroll
                        for (int e = 0; e < 16; ++e) {
                            const int rr = 32 + 8 * (e >> 2) + 4 * h + (e & 3), imm = 4 * (32 + 8 * (e >> 2) + (e & 3));
                            buf_st(D1[e], rD, rr < mD ? vd : BUF_OOB, imm);
                        }
                    }
                }
            }
        }
    }
    return (wave0 + nitems) % MQE_WAVES;
}

// The same chained products with 16-column strips (v_mfma_f32_16x16x4f32):
// F row tiles and D row tiles of 16 rows (at most four each), lane (c, g)
// (c = lane % 16, g = lane / 16) takes k = k0 + 4g + q of a 16-deep slice, and
// accumulator slot e of lane (c, g) holds row 4g + e of its tile, so MFMA e of
// F tile t in the second chain takes rows 16t + 4g + e.  For walker bonds below
// 32 (no half-empty strips, twice the items); wider bonds keep 32-column strips
// (a bond of 48 in three 16-column strips re-reads A and C three times and
// measured 20 % slower than two 32-column strips).
// qsplit: the D row tiles 0-1 and 2-3 are separate items.
typedef float v4f_c __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f_c mfma16(float a, float b, v4f_c c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <bool TB, bool TC, int NF, int ND, class Get>
__device__ __forceinline__ int chain2_16(int wave0, int nu, int m1, int n1, int k1, int mD, int lda, int ldb, int ldc,
                                         int ldd, bool qsplit, Get get) {
    static_assert(ND == 1 || NF == 1, "two outputs take one intermediate");
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int MT = (m1 + 15) >> 4, QT = (mD + 15) >> 4;  // F and D row tiles (<= 4)
    const int nq = qsplit && QT > 2 ? 2 : 1;
    const int nstrip = (n1 + 15) >> 4, nitems = nu * nstrip * nq;
    const int sa = 4 * lda, sb = TB ? 4 * ldb : 4;  // bytes per k step of A and op(B)
    const int kf = k1 & ~15;
#pragma nounroll
    for (int it = (w - wave0 + 2 * MQE_WAVES) % MQE_WAVES; it < nitems; it += MQE_WAVES) {
        const int qi = it % nq, rest = it / nq;
        const int u = rest / nstrip, c0 = (rest - u * nstrip) * 16;
        const int q0 = nq == 1 ? 0 : 2 * qi, q1 = nq == 1 ? QT : min(QT, 2 * qi + 2);  // D row tiles of this item
        const ChainOps op = get(u);
        int c = lane & 15, g = lane >> 4;
        v4f_c D[4];
#pragma nounroll
        for (int f = 0; f < NF; ++f) {
            asm volatile("" : "+v"(c), "+v"(g));
            const int col = min(c0 + c, n1 - 1);
            const int vB = TB ? 4 * col : 4 * ldb * col;
            // ---- first chain: the F_f strip (16 columns), up to four 16-row tiles ----
            v4f_c F[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) F[t] = v4f_c{0.f, 0.f, 0.f, 0.f};
            const auto rA = buf_rsrc(op.A), rB = buf_rsrc(op.B[f]);
            auto slice = [&](int k0, bool tail) {
                const int kl = k0 + 4 * g;
                float a[4][4], bv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const bool ok = !tail || kl + q < k1;
                    bv[q] = buf_ld(rB, ok ? vB + sb * (kl + q) : BUF_OOB);
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        a[t][q] = t < MT ? buf_ld(rA, ok ? 4 * min(16 * t + c, m1 - 1) + sa * (kl + q) : BUF_OOB) : 0.f;
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int t = 0; t < 4; ++t)
                        if (t < MT) F[t] = mfma16(a[t][q], bv[q], F[t]);
            };
#pragma nounroll
            for (int k0 = 0; k0 < kf; k0 += 16) slice(k0, false);
            if (kf < k1) slice(kf, true);
            if (m1 & 15) {  // rows of F beyond m1 (clamped A rows) must not contribute
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (16 * t + 4 * g + e >= m1) F[t][e] = 0.f;
            }
            // ---- second chain: D_d (+)= op(C) F_f, 16-row tile by tile ----
#pragma nounroll
            for (int d = 0; d < ND; ++d) {
                asm volatile("" : "+v"(c), "+v"(g));
                const float *Cp = op.C[ND == 1 ? f : d];
                const auto rC = buf_rsrc(Cp);
                const bool vecC = TC && (m1 & 3) == 0 && (ldc & 3) == 0 && aligned16(Cp, Cp, 0, 0);
                if (ND == 2 || f == 0) {
#pragma unroll
                    for (int qd = 0; qd < 4; ++qd) D[qd] = v4f_c{0.f, 0.f, 0.f, 0.f};
                }
#pragma unroll
                for (int qd = 0; qd < 4; ++qd) {
                    if (qd < q0 || qd >= q1) continue;
                    const int row = min(16 * qd + c, mD - 1);
                    float cv[4][4];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (t >= MT) continue;
                        if (TC && vecC) {
                            // a run at or beyond m1 meets zero rows of F: any values of C do
                            buf_ld4(rC, 4 * (ldc * row + min(16 * t + 4 * g, m1 - 4)), cv[t]);
                        } else {
#pragma unroll
                            for (int e = 0; e < 4; ++e) {
                                const int k = min(16 * t + 4 * g + e, m1 - 1);
                                cv[t][e] = buf_ld(rC, TC ? 4 * (ldc * row + k) : 4 * (row + ldc * k));
                            }
                        }
                    }
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (t >= MT) continue;
#pragma unroll
                        for (int e = 0; e < 4; ++e) D[qd] = mfma16(cv[t][e], F[t][e], D[qd]);
                    }
                }
                if (f == NF - 1) {
                    // row 16 qd + 4g + e of column c0 + c; outside D the store is dropped
                    const int vd = c0 + c < n1 ? 4 * (ldd * (c0 + c) + 4 * g) : BUF_OOB;
                    const auto rD = buf_rsrc(op.D[d]);
#pragma unroll
                    for (int qd = 0; qd < 4; ++qd) {
                        if (qd < q0 || qd >= q1) continue;
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            buf_st(D[qd][e], rD, 16 * qd + 4 * g + e < mD ? vd : BUF_OOB, 4 * (16 * qd + e));
                    }
                }
            }
        }
    }
    return (wave0 + nitems) % MQE_WAVES;
}

// ---------------------------------------------------------------------------
// Householder QR (LAPACK geqrf/orgqr conventions), FP64, whole workgroup.
// A: m x n (m >= n), column c at Ac(c) (columns c < c_split at A + m c, the
// others at B + m (c - c_split): the caller may keep the trailing columns,
// which receive most of the updates, in LDS).  On exit the columns hold Q and
// R (n x n, ld n) the upper triangle.  tau: n.
// Column c belongs to wave c % 4 for the whole factorisation: the owner of
// column j forms reflector j from its (already updated) column, one
// workgroup barrier publishes it, and every wave applies it to its own
// columns, four columns at a time, with the 64 lanes running down the rows
// (coalesced) and wave reductions for the dot products.  Q is accumulated
// backwards the same way (one barrier per reflector).
// ---------------------------------------------------------------------------
struct QRCols {
    double *A, *B;
    int m, c_split;
    __device__ double *operator()(int c) const { return c < c_split ? A + (size_t)m * c : B + (size_t)m * (c - c_split); }
};

__device__ __forceinline__ void reflect_own_columns(const QRCols &col, int n, int j, double tj, int w, int lane) {
    const int m = col.m;
    const double *vj = col(j);
    int c = j + 1 + ((w - (j + 1)) & (MQE_WAVES - 1));  // first column > j owned by this wave
    for (; c < n; c += 4 * MQE_WAVES) {
        double *cp[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) cp[q] = col(c + MQE_WAVES * q < n ? c + MQE_WAVES * q : c);
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = j + lane; i < m; i += 64) {
            const double v = i == j ? 1.0 : vj[i];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (c + MQE_WAVES * q < n) s[q] += v * cp[q][i];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] = tj * wave_sum(s[q]);
        for (int i = j + lane; i < m; i += 64) {
            const double v = i == j ? 1.0 : vj[i];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (c + MQE_WAVES * q < n) cp[q][i] -= s[q] * v;
        }
    }
}

__device__ __forceinline__ void householder_qr(const QRCols &col, int n, double *R, double *tau) {
    const int m = col.m;
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int j = 0; j < n; ++j) {
        if (w == (j & (MQE_WAVES - 1))) {
            double *aj = col(j);
            double ss = 0.0;
            for (int i = j + 1 + lane; i < m; i += 64) {
                const double a = aj[i];
                ss += a * a;
            }
            ss = wave_sum(ss);
            const double alpha = aj[j];
            double tj = 0.0, beta = alpha, scale = 0.0;
            if (ss != 0.0) {
                beta = -copysign(sqrt(alpha * alpha + ss), alpha);
                tj = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            if (tj != 0.0) {
                for (int i = j + 1 + lane; i < m; i += 64) aj[i] *= scale;
                if (lane == 0) aj[j] = beta;
            }
            if (lane == 0) tau[j] = tj;
        }
        __syncthreads();
        const double tj = tau[j];
        if (tj != 0.0) reflect_own_columns(col, n, j, tj, w, lane);
    }
    __syncthreads();
    for (int e = tid; e < n * n; e += MQE_NT) {
        const int r = e % n, c = e / n;
        R[e] = (r <= c) ? col(c)[r] : 0.0;
    }
    __syncthreads();
    for (int j = n - 1; j >= 0; --j) {
        const double tj = tau[j];
        if (tj != 0.0) reflect_own_columns(col, n, j, tj, w, lane);
        __syncthreads();  // every wave has read reflector j
        if (w == (j & (MQE_WAVES - 1))) {
            double *aj = col(j);
            for (int i = lane; i < m; i += 64) aj[i] = (i < j) ? 0.0 : (i == j ? 1.0 - tj : -tj * aj[i]);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) counter-based RNG.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t philox_u32(uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c0;
}

__device__ __forceinline__ double philox_uniform(uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    return (double)philox_u32(key, c0, c1, c2, c3) * (1.0 / 4294967296.0);
}

}  // namespace mqe

#endif
