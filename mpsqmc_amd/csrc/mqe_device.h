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
// FP32 GEMM on the matrix cores: C(m x n) (+)= op(A) op(B), whole workgroup.
// The output is cut into 32 x 32 tiles, one wave per tile at a time; each tile
// is a chain of v_mfma_f32_32x32x2f32 over k with operands read straight from
// (L2-resident) memory, zero outside the matrices.  Within a 16-deep slice of
// k starting at k0, MFMA u of the chain takes kk = k0 + 8(l/32) + u in lane l:
// lane l holds A(l % 32, kk) and B(kk, l % 32), so each lane's eight values of
// an operand stored contiguously along k are one 32-byte run (two 16-byte
// loads when aligned), and an operand stored contiguously along its other
// index is read coalesced across the lanes.  Accumulator slot q of lane l is
// element (8(q/4) + 4(l/32) + q%4, l % 32) of the tile.
// ---------------------------------------------------------------------------
template <bool TA, bool TB>
__device__ __forceinline__ void gemm_mfma(int m, int n, int k, const float *A0, int lda, const float *B0, int ldb,
                                          float *C0, int ldc, bool acc, int nb = 1, long sA = 0, long sB = 0,
                                          long sC = 0) {
    typedef float v16 __attribute__((ext_vector_type(16)));
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 31, h = lane >> 5;
    const int tm = (m + 31) >> 5, tn = (n + 31) >> 5;
    // 16-byte loads along k need 16-byte aligned bases, leading dimensions and batch strides
    const bool avec = TA && ((((uintptr_t)A0) | ((uintptr_t)lda * 4u) | ((uintptr_t)sA * 4u)) & 15u) == 0;
    const bool bvec = !TB && ((((uintptr_t)B0) | ((uintptr_t)ldb * 4u) | ((uintptr_t)sB * 4u)) & 15u) == 0;
    for (int tb = w; tb < nb * tm * tn; tb += MQE_WAVES) {
        // tiles of all nb products of the batch are shared among the waves
        const int bt = tb / (tm * tn), t = tb - bt * (tm * tn);
        const float *A = A0 + (size_t)bt * sA, *B = B0 + (size_t)bt * sB;
        float *C = C0 + (size_t)bt * sC;
        const int r0 = (t % tm) * 32, c0 = (t / tm) * 32;
        const int ar = r0 + r, bc = c0 + r;
        const bool aok = ar < m, bok = bc < n;
        v16 c;
#pragma unroll
        for (int q = 0; q < 16; ++q) c[q] = 0.f;
        for (int k0 = 0; k0 < k; k0 += 16) {
            const int kb = k0 + 8 * h;
            const bool full = k0 + 16 <= k;
            float a[8], b[8];
            if (avec && full) {
                const float4 *p = (const float4 *)(A + (size_t)lda * (aok ? ar : 0) + kb);
                const float4 x = p[0], y = p[1];
                a[0] = aok ? x.x : 0.f; a[1] = aok ? x.y : 0.f; a[2] = aok ? x.z : 0.f; a[3] = aok ? x.w : 0.f;
                a[4] = aok ? y.x : 0.f; a[5] = aok ? y.y : 0.f; a[6] = aok ? y.z : 0.f; a[7] = aok ? y.w : 0.f;
            } else {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int kk = kb + u;
                    a[u] = (aok && kk < k) ? (TA ? A[kk + (size_t)lda * ar] : A[ar + (size_t)lda * kk]) : 0.f;
                }
            }
            if (bvec && full) {
                const float4 *p = (const float4 *)(B + (size_t)ldb * (bok ? bc : 0) + kb);
                const float4 x = p[0], y = p[1];
                b[0] = bok ? x.x : 0.f; b[1] = bok ? x.y : 0.f; b[2] = bok ? x.z : 0.f; b[3] = bok ? x.w : 0.f;
                b[4] = bok ? y.x : 0.f; b[5] = bok ? y.y : 0.f; b[6] = bok ? y.z : 0.f; b[7] = bok ? y.w : 0.f;
            } else {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int kk = kb + u;
                    b[u] = (bok && kk < k) ? (TB ? B[bc + (size_t)ldb * kk] : B[kk + (size_t)ldb * bc]) : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) c = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], b[u], c, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int rr = r0 + 8 * (q >> 2) + 4 * h + (q & 3);
            if (rr < m && bok) {
                float v = c[q];
                if (acc) v += C[rr + (size_t)ldc * bc];
                C[rr + (size_t)ldc * bc] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The same product with 16 x 16 output tiles (v_mfma_f32_16x16x4f32): small
// operands (bonds of 16 or 48) waste less of each tile and spread over all
// four waves.  Within a 16-deep slice of k starting at k0, MFMA u takes
// kk = k0 + 4(l/16) + u in lane l: lane l holds A(l % 16, kk) and B(kk, l % 16)
// (four consecutive values of an operand stored along k: one 16-byte load),
// and accumulator slot q holds element (4(l/16) + q, l % 16) of the tile.
// ---------------------------------------------------------------------------
template <bool TA, bool TB>
__device__ __forceinline__ void gemm_mfma16(int m, int n, int k, const float *A0, int lda, const float *B0, int ldb,
                                            float *C0, int ldc, bool acc, int nb = 1, long sA = 0, long sB = 0,
                                            long sC = 0) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 15, g = lane >> 4;
    const int tm = (m + 15) >> 4, tn = (n + 15) >> 4;
    const bool avec = TA && ((((uintptr_t)A0) | ((uintptr_t)lda * 4u) | ((uintptr_t)sA * 4u)) & 15u) == 0;
    const bool bvec = !TB && ((((uintptr_t)B0) | ((uintptr_t)ldb * 4u) | ((uintptr_t)sB * 4u)) & 15u) == 0;
    for (int tb = w; tb < nb * tm * tn; tb += MQE_WAVES) {
        const int bt = tb / (tm * tn), t = tb - bt * (tm * tn);
        const float *A = A0 + (size_t)bt * sA, *B = B0 + (size_t)bt * sB;
        float *C = C0 + (size_t)bt * sC;
        const int r0 = (t % tm) * 16, c0 = (t / tm) * 16;
        const int ar = r0 + r, bc = c0 + r;
        const bool aok = ar < m, bok = bc < n;
        v4 c = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < k; k0 += 16) {
            const int kb = k0 + 4 * g;
            const bool full = k0 + 16 <= k;
            float a[4], b[4];
            if (avec && full) {
                const float4 x = *(const float4 *)(A + (size_t)lda * (aok ? ar : 0) + kb);
                a[0] = aok ? x.x : 0.f; a[1] = aok ? x.y : 0.f; a[2] = aok ? x.z : 0.f; a[3] = aok ? x.w : 0.f;
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int kk = kb + u;
                    a[u] = (aok && kk < k) ? (TA ? A[kk + (size_t)lda * ar] : A[ar + (size_t)lda * kk]) : 0.f;
                }
            }
            if (bvec && full) {
                const float4 x = *(const float4 *)(B + (size_t)ldb * (bok ? bc : 0) + kb);
                b[0] = bok ? x.x : 0.f; b[1] = bok ? x.y : 0.f; b[2] = bok ? x.z : 0.f; b[3] = bok ? x.w : 0.f;
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int kk = kb + u;
                    b[u] = (bok && kk < k) ? (TB ? B[bc + (size_t)ldb * kk] : B[kk + (size_t)ldb * bc]) : 0.f;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b[u], c, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rr = r0 + 4 * g + q;
            if (rr < m && bok) {
                float v = c[q];
                if (acc) v += C[rr + (size_t)ldc * bc];
                C[rr + (size_t)ldc * bc] = v;
            }
        }
    }
}

// 32 x 64 output blocks: each wave computes two neighbouring 32 x 32 tiles that
// share the A panel (half the A reads of gemm_mfma for wide products).
template <bool TA, bool TB>
__device__ __forceinline__ void gemm_mfma_wide(int m, int n, int k, const float *A, int lda, const float *B,
                                               int ldb, float *C, int ldc, bool acc) {
    typedef float v16 __attribute__((ext_vector_type(16)));
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int r = lane & 31, h = lane >> 5;
    const int tm = (m + 31) >> 5, tn = (n + 63) >> 6;
    const bool avec = TA && ((((uintptr_t)A) | ((uintptr_t)lda * 4u)) & 15u) == 0;
    const bool bvec = !TB && ((((uintptr_t)B) | ((uintptr_t)ldb * 4u)) & 15u) == 0;
    for (int t = w; t < tm * tn; t += MQE_WAVES) {
        const int r0 = (t % tm) * 32, c0 = (t / tm) * 64;
        const bool two = c0 + 32 < n;  // wave-uniform
        const int ar = r0 + r;
        const bool aok = ar < m;
        v16 c[2];
#pragma unroll
        for (int q = 0; q < 16; ++q) { c[0][q] = 0.f; c[1][q] = 0.f; }
        for (int k0 = 0; k0 < k; k0 += 16) {
            const int kb = k0 + 8 * h;
            const bool full = k0 + 16 <= k;
            float a[8];
            if (avec && full) {
                const float4 *p = (const float4 *)(A + (size_t)lda * (aok ? ar : 0) + kb);
                const float4 x = p[0], y = p[1];
                a[0] = aok ? x.x : 0.f; a[1] = aok ? x.y : 0.f; a[2] = aok ? x.z : 0.f; a[3] = aok ? x.w : 0.f;
                a[4] = aok ? y.x : 0.f; a[5] = aok ? y.y : 0.f; a[6] = aok ? y.z : 0.f; a[7] = aok ? y.w : 0.f;
            } else {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int kk = kb + u;
                    a[u] = (aok && kk < k) ? (TA ? A[kk + (size_t)lda * ar] : A[ar + (size_t)lda * kk]) : 0.f;
                }
            }
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                if (half == 1 && !two) break;
                const int bc = c0 + 32 * half + r;
                const bool bok = bc < n;
                float b[8];
                if (bvec && full) {
                    const float4 *p = (const float4 *)(B + (size_t)ldb * (bok ? bc : 0) + kb);
                    const float4 x = p[0], y = p[1];
                    b[0] = bok ? x.x : 0.f; b[1] = bok ? x.y : 0.f; b[2] = bok ? x.z : 0.f; b[3] = bok ? x.w : 0.f;
                    b[4] = bok ? y.x : 0.f; b[5] = bok ? y.y : 0.f; b[6] = bok ? y.z : 0.f; b[7] = bok ? y.w : 0.f;
                } else {
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int kk = kb + u;
                        b[u] = (bok && kk < k) ? (TB ? B[bc + (size_t)ldb * kk] : B[kk + (size_t)ldb * bc]) : 0.f;
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) c[half] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], b[u], c[half], 0, 0, 0);
            }
        }
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            if (half == 1 && !two) break;
            const int bc = c0 + 32 * half + r;
            if (bc >= n) continue;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int rr = r0 + 8 * (q >> 2) + 4 * h + (q & 3);
                if (rr < m) {
                    float v = c[half][q];
                    if (acc) v += C[rr + (size_t)ldc * bc];
                    C[rr + (size_t)ldc * bc] = v;
                }
            }
        }
    }
}

// Output tile shapes of the FP32 matrix-core products (chosen per walker bond
// bucket by the engine's tile table).
enum : int { TILE_16 = 16, TILE_32 = 32, TILE_WIDE = 64 };

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
            if (tile == TILE_16)
                gemm_mfma16<TA, TB>(m, n, k, (const float *)A, lda, (const float *)B, ldb, (float *)C, ldc, acc, nb,
                                    sA, sB, sC);
            else
                gemm_mfma<TA, TB>(m, n, k, (const float *)A, lda, (const float *)B, ldb, (float *)C, ldc, acc, nb, sA,
                                  sB, sC);
            return;
        }
    }
    const int mn = m * n;
    for (int e = threadIdx.x; e < nb * mn; e += MQE_NT) {
        const int bt = e / mn, x = e - bt * mn;
        const int r = x % m, c = x / m;
        const T *Ab = A + (size_t)bt * sA, *Bb = B + (size_t)bt * sB;
        T *Cb = C + (size_t)bt * sC;
        T s = acc ? Cb[r + ldc * c] : T(0);
        const T *ap = TA ? (Ab + (size_t)lda * r) : (Ab + r);
        const T *bp = TB ? (Bb + c) : (Bb + (size_t)ldb * c);
        for (int q = 0; q < k; ++q) {
            const T a = TA ? ap[q] : ap[(size_t)lda * q];
            const T b = TB ? bp[(size_t)ldb * q] : bp[q];
            s += a * b;
        }
        Cb[r + ldc * c] = s;
    }
}

// ---------------------------------------------------------------------------
// GEMM: C(m x n) (+)= op(A) op(B).  op(A) is m x k: A(r,q) = TA ? A[q + lda*r] : A[r + lda*q].
// FP32 products of useful size run on the matrix cores with the output tile
// shape `tile`, the rest (and all of FP64) on the vector units.
// ---------------------------------------------------------------------------
template <bool TA, bool TB, typename T>
__device__ __forceinline__ void gemm(int m, int n, int k, const T *A, int lda, const T *B, int ldb, T *C, int ldc,
                                     bool acc, int tile = TILE_32) {
    if constexpr (sizeof(T) == sizeof(float)) {
        if (m >= 8 && n >= 8 && k >= 4) {
            if (tile == TILE_16)
                gemm_mfma16<TA, TB>(m, n, k, (const float *)A, lda, (const float *)B, ldb, (float *)C, ldc, acc);
            else if (tile == TILE_WIDE && n > 32)
                gemm_mfma_wide<TA, TB>(m, n, k, (const float *)A, lda, (const float *)B, ldb, (float *)C, ldc, acc);
            else
                gemm_mfma<TA, TB>(m, n, k, (const float *)A, lda, (const float *)B, ldb, (float *)C, ldc, acc);
            return;
        }
    }
    const int mn = m * n;
    for (int e = threadIdx.x; e < mn; e += MQE_NT) {
        const int r = e % m, c = e / m;
        T s = acc ? C[r + ldc * c] : T(0);
        const T *ap = TA ? (A + (size_t)lda * r) : (A + r);
        const T *bp = TB ? (B + c) : (B + (size_t)ldb * c);
        for (int q = 0; q < k; ++q) {
            const T a = TA ? ap[q] : ap[(size_t)lda * q];
            const T b = TB ? bp[(size_t)ldb * q] : bp[q];
            s += a * b;
        }
        C[r + ldc * c] = s;
    }
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
