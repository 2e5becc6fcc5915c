// mqe_mlp.hip -- the learned proposal's network: a small MLP
// (n_in -> H -> H -> n_out, ReLU, FP32) trained and evaluated on the GPU.
//
// The network maps per-walker, per-coupling features to the 16 normalised
// two-site mixed elements combos/O of that coupling (the quantities from which
// the reference builds the sampling PDF, MPSQMC2.cpp:540-574); the engine turns
// the prediction into a proposal over the grid and corrects the weight with the
// exact PDF value of the sampled point.  The reference has no such component.
//
// Kernels: one workgroup per tile of 64 samples.  The activations of a tile
// stay in LDS (feature-major, ld = width); every layer product is a workgroup
// MFMA GEMM (v_mfma_f32_32x32x2f32) with the weights read from global memory
// (80 KB at H = 128, L2-resident).  Training computes forward and backward of a
// tile and writes the tile's gradient to its own slot; the Adam kernel sums the
// slots in tile order (deterministic) and updates the parameters.
//
// Parameter layout (column-major): W1 [H x n_in], b1 [H], W2 [H x H], b2 [H],
// W3 [n_out x H], b3 [n_out].  Inputs are standardised with (x - x_mean) /
// x_std and outputs de-standardised with y * y_std + y_mean; the training loss
// is the mean squared error of the standardised outputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "mqe.h"
#include "mqe_device.h"
#include "mqe_internal.h"

namespace mqe {
namespace mlp {

constexpr int TS = 64;         // samples per tile
constexpr int MAX_IN = 32, MAX_H = 128, MAX_OUT = 32;

struct Net {
    int nin, H, nout;
    const float *par;     // parameters (layout above)
    const float *scale;   // x_mean[nin], x_std[nin], y_mean[nout], y_std[nout]
};

__host__ __device__ inline long off_b1(int nin, int H) { return (long)H * nin; }
__host__ __device__ inline long off_W2(int nin, int H) { return off_b1(nin, H) + H; }
__host__ __device__ inline long off_b2(int nin, int H) { return off_W2(nin, H) + (long)H * H; }
__host__ __device__ inline long off_W3(int nin, int H) { return off_b2(nin, H) + H; }
__host__ __device__ inline long off_b3(int nin, int H, int nout) { return off_W3(nin, H) + (long)nout * H; }
__host__ __device__ inline long num_params(int nin, int H, int nout) { return off_b3(nin, H, nout) + nout; }

inline size_t lds_bytes(int nin, int H, int nout, bool train) {
    // xs [nin x TS], h1, h2 [H x TS], o [nout x TS], (train) dh2 [H x TS]
    return sizeof(float) * (size_t)TS * (nin + 2 * H + nout + (train ? H : 0));
}

// bias + optional ReLU on a [rows x TS] LDS block
__device__ __forceinline__ void bias_act(float *a, int rows, const float *b, bool relu) {
    for (int e = threadIdx.x; e < rows * TS; e += MQE_NT) {
        float v = a[e] + b[e % rows];
        a[e] = relu ? fmaxf(v, 0.f) : v;
    }
}

// Rows of the tile: r < n valid.  perm (optional) maps tile rows to data rows.
// rows_dev (optional): the row count is *rows_dev * mult (device-side count).
template <bool TRAIN>
__global__ __launch_bounds__(MQE_NT) void tile_kernel(Net net, const float *X, const float *Y, long n,
                                                      const int *rows_dev, int mult, const int *perm, float *out,
                                                      float *gpart, double *loss) {
    extern __shared__ float sm[];
    const int nin = net.nin, H = net.H, nout = net.nout;
    if (rows_dev) n = (long)(*rows_dev) * mult;
    const long r0 = (long)blockIdx.x * TS;
    if (r0 >= n) return;
    float *xs = sm, *h1 = xs + nin * TS, *h2 = h1 + H * TS, *o = h2 + H * TS, *dh2 = o + nout * TS;
    const float *W1 = net.par, *b1 = W1 + off_b1(nin, H), *W2 = W1 + off_W2(nin, H), *b2 = W1 + off_b2(nin, H),
                *W3 = W1 + off_W3(nin, H), *b3 = W1 + off_b3(nin, H, nout);
    const float *xm = net.scale, *xsd = xm + nin, *ym = xsd + nin, *ysd = ym + nout;
    auto row_of = [&](int j) -> long {
        const long r = r0 + j;
        if (r >= n) return -1;
        return perm ? (long)perm[r] : r;
    };
    for (int e = threadIdx.x; e < nin * TS; e += MQE_NT) {
        const int f = e % nin, j = e / nin;
        const long r = row_of(j);
        xs[e] = r < 0 ? 0.f : (X[r * nin + f] - xm[f]) / xsd[f];
    }
    __syncthreads();
    gemm_mfma<false, false>(H, TS, nin, W1, H, xs, nin, h1, H, false);
    __syncthreads();
    bias_act(h1, H, b1, true);
    __syncthreads();
    gemm_mfma<false, false>(H, TS, H, W2, H, h1, H, h2, H, false);
    __syncthreads();
    bias_act(h2, H, b2, true);
    __syncthreads();
    gemm_mfma<false, false>(nout, TS, H, W3, nout, h2, H, o, nout, false);
    __syncthreads();
    bias_act(o, nout, b3, false);
    __syncthreads();
    if constexpr (!TRAIN) {
        for (int e = threadIdx.x; e < nout * TS; e += MQE_NT) {
            const int q = e % nout, j = e / nout;
            const long r = r0 + j;
            if (r < n) out[r * nout + q] = o[e] * ysd[q] + ym[q];
        }
        return;
    } else {
        // dL/do for L = sum (o - y)^2 / (n nout) over the whole batch of n rows
        const float g = 2.f / (float)(n * nout);
        double ls = 0.0;
        for (int e = threadIdx.x; e < nout * TS; e += MQE_NT) {
            const int q = e % nout, j = e / nout;
            const long r = row_of(j);
            float d = 0.f;
            if (r >= 0) {
                d = o[e] - (Y[r * nout + q] - ym[q]) / ysd[q];
                ls += (double)d * d;
            }
            o[e] = g * d;
        }
        __shared__ double red[MQE_WAVES];
        ls = block_sum(ls, red);
        if (threadIdx.x == 0) atomicAdd(loss, ls);
        __syncthreads();
        float *G = gpart + (size_t)blockIdx.x * num_params(nin, H, nout);
        float *gW1 = G, *gb1 = G + off_b1(nin, H), *gW2 = G + off_W2(nin, H), *gb2 = G + off_b2(nin, H),
              *gW3 = G + off_W3(nin, H), *gb3 = G + off_b3(nin, H, nout);
        // layer 3: gW3 = dO h2^T, gb3 = rows of dO, dh2 = W3^T dO masked by h2 > 0
        gemm_mfma<false, true>(nout, H, TS, o, nout, h2, H, gW3, nout, false);
        gemm_mfma<true, false>(H, TS, nout, W3, nout, o, nout, dh2, H, false);
        for (int q = threadIdx.x; q < nout; q += MQE_NT) {
            float s = 0.f;
            for (int j = 0; j < TS; ++j) s += o[q + nout * j];
            gb3[q] = s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < H * TS; e += MQE_NT)
            if (!(h2[e] > 0.f)) dh2[e] = 0.f;
        __syncthreads();
        // layer 2: gW2 = dh2 h1^T, gb2, dh1 = W2^T dh2 (into h2) masked by h1 > 0
        gemm_mfma<false, true>(H, H, TS, dh2, H, h1, H, gW2, H, false);
        gemm_mfma<true, false>(H, TS, H, W2, H, dh2, H, h2, H, false);
        for (int q = threadIdx.x; q < H; q += MQE_NT) {
            float s = 0.f;
            for (int j = 0; j < TS; ++j) s += dh2[q + H * j];
            gb2[q] = s;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < H * TS; e += MQE_NT)
            if (!(h1[e] > 0.f)) h2[e] = 0.f;
        __syncthreads();
        // layer 1: gW1 = dh1 xs^T, gb1
        gemm_mfma<false, true>(H, nin, TS, h2, H, xs, nin, gW1, H, false);
        for (int q = threadIdx.x; q < H; q += MQE_NT) {
            float s = 0.f;
            for (int j = 0; j < TS; ++j) s += h2[q + H * j];
            gb1[q] = s;
        }
    }
}

// Sum of the tile gradients (in tile order) and one Adam step.
__global__ void adam_kernel(float *par, float *m, float *v, const float *gpart, int ntiles, long np, float lr,
                            float b1, float b2, float eps, float c1, float c2, float *grad_out) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
        float g = 0.f;
        for (int t = 0; t < ntiles; ++t) g += gpart[(size_t)t * np + p];
        if (grad_out) {
            grad_out[p] = g;
            continue;
        }
        const float mm = b1 * m[p] + (1.f - b1) * g, vv = b2 * v[p] + (1.f - b2) * g * g;
        m[p] = mm;
        v[p] = vv;
        par[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    }
}

// Column mean and standard deviation of an n x dim row-major array (one
// workgroup per column; FP64 accumulation).  A zero deviation becomes 1.
__global__ __launch_bounds__(MQE_NT) void stats_kernel(const float *A, long n, int dim, float *mean, float *sd) {
    __shared__ double red[2 * MQE_WAVES];
    const int f = blockIdx.x;
    double s[2] = {0.0, 0.0};
    for (long r = threadIdx.x; r < n; r += MQE_NT) {
        const double a = A[r * dim + f];
        s[0] += a;
        s[1] += a * a;
    }
    block_sum_n<2>(s, red);
    if (threadIdx.x == 0) {
        const double mu = s[0] / (double)max(n, 1L);
        const double var = s[1] / (double)max(n, 1L) - mu * mu;
        mean[f] = (float)mu;
        sd[f] = var > 1e-30 ? (float)sqrt(var) : 1.f;
    }
}

}  // namespace mlp
}  // namespace mqe

using namespace mqe;

struct mqe_mlp {
    int nin = 0, H = 0, nout = 0;
    long np = 0;
    float *d_par = nullptr, *d_m = nullptr, *d_v = nullptr, *d_scale = nullptr, *d_gpart = nullptr;
    float *d_X = nullptr, *d_Y = nullptr, *d_out = nullptr, *d_grad = nullptr;
    int *d_perm = nullptr;
    double *d_loss = nullptr;
    long gpart_tiles = 0, data_cap = 0, perm_cap = 0;
    long adam_t = 0;
    unsigned long long seed = 0;
    hipStream_t stream = nullptr;
    mlp::Net net() const { return mlp::Net{nin, H, nout, d_par, d_scale}; }
};

#define MCHK(x)                                                                                  \
    do {                                                                                         \
        hipError_t _e = (x);                                                                     \
        if (_e != hipSuccess) {                                                                  \
            set_error(hipGetErrorString(_e));                                                    \
            return MQE_EDEVICE;                                                                  \
        }                                                                                        \
    } while (0)

namespace {

int ensure_gpart(mqe_mlp *m, long tiles) {
    if (tiles <= m->gpart_tiles) return MQE_OK;
    if (m->d_gpart) MCHK(hipFree(m->d_gpart));
    m->d_gpart = nullptr;
    MCHK(hipMalloc((void **)&m->d_gpart, sizeof(float) * (size_t)tiles * m->np));
    m->gpart_tiles = tiles;
    return MQE_OK;
}

int ensure_data(mqe_mlp *m, long n) {
    if (n <= m->data_cap) return MQE_OK;
    for (float **p : {&m->d_X, &m->d_Y, &m->d_out})
        if (*p) MCHK(hipFree(*p));
    MCHK(hipMalloc((void **)&m->d_X, sizeof(float) * (size_t)n * m->nin));
    MCHK(hipMalloc((void **)&m->d_Y, sizeof(float) * (size_t)n * m->nout));
    MCHK(hipMalloc((void **)&m->d_out, sizeof(float) * (size_t)n * m->nout));
    m->data_cap = n;
    return MQE_OK;
}

int set_lds(mqe_mlp *m) {
    MCHK(hipFuncSetAttribute((const void *)mlp::tile_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mlp::lds_bytes(m->nin, m->H, m->nout, true)));
    MCHK(hipFuncSetAttribute((const void *)mlp::tile_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mlp::lds_bytes(m->nin, m->H, m->nout, false)));
    return MQE_OK;
}

// loss and tile gradients of rows [0, n) (through perm when given) into d_gpart
int grad_pass(mqe_mlp *m, const float *dX, const float *dY, long n, const int *dperm) {
    const long tiles = (n + mlp::TS - 1) / mlp::TS;
    int r = ensure_gpart(m, tiles);
    if (r) return r;
    hipLaunchKernelGGL(mlp::tile_kernel<true>, dim3((unsigned)tiles), dim3(MQE_NT),
                       mlp::lds_bytes(m->nin, m->H, m->nout, true), m->stream, m->net(), dX, dY, n,
                       (const int *)nullptr, 0, dperm, (float *)nullptr, m->d_gpart, m->d_loss);
    MCHK(hipGetLastError());
    return MQE_OK;
}

}  // namespace

namespace mqe {

// Forward pass over device rows (count n, or *rows_dev * mult when rows_dev is
// given, at most max_rows) on the given stream.
int mlp_forward_device(mqe_mlp *m, const float *dX, long max_rows, const int *rows_dev, int mult, float *dY,
                       hipStream_t s) {
    const long tiles = (max_rows + mlp::TS - 1) / mlp::TS;
    if (tiles <= 0) return MQE_OK;
    hipLaunchKernelGGL(mlp::tile_kernel<false>, dim3((unsigned)tiles), dim3(MQE_NT),
                       mlp::lds_bytes(m->nin, m->H, m->nout, false), s, m->net(), dX, (const float *)nullptr,
                       max_rows, rows_dev, mult, (const int *)nullptr, dY, (float *)nullptr, (double *)nullptr);
    MCHK(hipGetLastError());
    return MQE_OK;
}

int mlp_io_dims(const mqe_mlp *m, int *nin, int *nout) {
    *nin = m->nin;
    *nout = m->nout;
    return MQE_OK;
}

// Adam on device-resident rows: epochs over a seeded shuffle, minibatches of
// `batch` rows; optionally standardise inputs and outputs from the data first.
int mlp_fit_device(mqe_mlp *m, const float *dX, const float *dY, long n, int epochs, int batch, double lr,
                   int standardize, double *loss_out, hipStream_t s) {
    if (n < 1 || epochs < 0 || batch < 1) { set_error("mqe_mlp_fit: bad sizes"); return MQE_EINVAL; }
    // the passes run on the caller's stream; the network's own stream is restored on every exit
    struct StreamSwap {
        mqe_mlp *m;
        hipStream_t keep;
        ~StreamSwap() { m->stream = keep; }
    } swap{m, m->stream};
    m->stream = s;
    if (standardize) {
        hipLaunchKernelGGL(mlp::stats_kernel, dim3(m->nin), dim3(MQE_NT), 0, s, dX, n, m->nin, m->d_scale,
                           m->d_scale + m->nin);
        hipLaunchKernelGGL(mlp::stats_kernel, dim3(m->nout), dim3(MQE_NT), 0, s, dY, n, m->nout,
                           m->d_scale + 2 * m->nin, m->d_scale + 2 * m->nin + m->nout);
        MCHK(hipGetLastError());
    }
    if (n > m->perm_cap) {
        if (m->d_perm) MCHK(hipFree(m->d_perm));
        MCHK(hipMalloc((void **)&m->d_perm, sizeof(int) * (size_t)n));
        m->perm_cap = n;
    }
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::mt19937_64 rng(m->seed ^ 0x9e3779b97f4a7c15ULL);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double last = 0.0;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        MCHK(hipMemcpyAsync(m->d_perm, perm.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
        MCHK(hipMemsetAsync(m->d_loss, 0, sizeof(double), s));
        for (long r0 = 0; r0 < n; r0 += batch) {
            const long nb = std::min((long)batch, n - r0);
            int r = grad_pass(m, dX, dY, nb, m->d_perm + r0);
            if (r) return r;
            ++m->adam_t;
            const float c1 = 1.f - powf(b1, (float)m->adam_t), c2 = 1.f - powf(b2, (float)m->adam_t);
            const int tiles = (int)((nb + mlp::TS - 1) / mlp::TS);
            hipLaunchKernelGGL(mlp::adam_kernel, dim3((unsigned)std::min<long>((m->np + 255) / 256, 1024)),
                               dim3(256), 0, s, m->d_par, m->d_m, m->d_v, m->d_gpart, tiles, m->np, (float)lr, b1,
                               b2, eps, c1, c2, (float *)nullptr);
            MCHK(hipGetLastError());
        }
        // the loss accumulated over the epoch's minibatches, per standardised output
        double lsum = 0.0;
        MCHK(hipMemcpyAsync(&lsum, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, s));
        MCHK(hipStreamSynchronize(s));
        last = lsum / ((double)n * m->nout);
    }
    if (loss_out) *loss_out = last;
    return MQE_OK;
}

}  // namespace mqe

extern "C" {

int mqe_mlp_create(int n_in, int hidden, int n_out, unsigned long long seed, mqe_mlp **out) {
    if (!out) { set_error("mqe_mlp_create: null argument"); return MQE_EINVAL; }
    if (n_in < 1 || n_in > mlp::MAX_IN || hidden < 1 || hidden > mlp::MAX_H || n_out < 1 || n_out > mlp::MAX_OUT) {
        set_error("mqe_mlp_create: sizes out of range (n_in <= 32, hidden <= 128, n_out <= 32)");
        return MQE_EINVAL;
    }
    mqe_mlp *m = new mqe_mlp();
    m->nin = n_in;
    m->H = hidden;
    m->nout = n_out;
    m->np = mlp::num_params(n_in, hidden, n_out);
    m->seed = seed;
    // He-normal weights (std sqrt(2 / fan_in)), zero biases, identity scaling
    std::vector<float> p(m->np, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    auto fill = [&](long off, long cnt, int fan_in) {
        const double s = std::sqrt(2.0 / fan_in);
        for (long q = 0; q < cnt; ++q) p[off + q] = (float)(s * nd(rng));
    };
    fill(0, (long)hidden * n_in, n_in);
    fill(mlp::off_W2(n_in, hidden), (long)hidden * hidden, hidden);
    fill(mlp::off_W3(n_in, hidden), (long)n_out * hidden, hidden);
    std::vector<float> sc(2 * n_in + 2 * n_out, 0.f);
    for (int f = 0; f < n_in; ++f) sc[n_in + f] = 1.f;
    for (int q = 0; q < n_out; ++q) sc[2 * n_in + n_out + q] = 1.f;
    auto fail = [&](hipError_t e) {
        set_error(hipGetErrorString(e));
        mqe_mlp_destroy(m);
        return MQE_EDEVICE;
    };
    hipError_t e;
    if ((e = hipStreamCreate(&m->stream)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->d_par, sizeof(float) * m->np)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->d_m, sizeof(float) * m->np)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->d_v, sizeof(float) * m->np)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->d_scale, sizeof(float) * sc.size())) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->d_loss, sizeof(double))) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->d_grad, sizeof(float) * m->np)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(m->d_par, p.data(), sizeof(float) * m->np, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = hipMemset(m->d_m, 0, sizeof(float) * m->np)) != hipSuccess) return fail(e);
    if ((e = hipMemset(m->d_v, 0, sizeof(float) * m->np)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(m->d_scale, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e);
    int r = set_lds(m);
    if (r) { mqe_mlp_destroy(m); return r; }
    *out = m;
    return MQE_OK;
}

void mqe_mlp_destroy(mqe_mlp *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *q : {(void *)m->d_par, (void *)m->d_m, (void *)m->d_v, (void *)m->d_scale, (void *)m->d_gpart,
                    (void *)m->d_X, (void *)m->d_Y, (void *)m->d_out, (void *)m->d_perm, (void *)m->d_loss, (void *)m->d_grad})
        if (q) (void)hipFree(q);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

long mqe_mlp_num_params(const mqe_mlp *m) { return m ? m->np : 0; }

int mqe_mlp_get_params(mqe_mlp *m, float *p) {
    if (!m || !p) { set_error("mqe_mlp_get_params: null argument"); return MQE_EINVAL; }
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(p, m->d_par, sizeof(float) * m->np, hipMemcpyDeviceToHost));
    return MQE_OK;
}

int mqe_mlp_set_params(mqe_mlp *m, const float *p) {
    if (!m || !p) { set_error("mqe_mlp_set_params: null argument"); return MQE_EINVAL; }
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(m->d_par, p, sizeof(float) * m->np, hipMemcpyHostToDevice));
    MCHK(hipMemset(m->d_m, 0, sizeof(float) * m->np));
    MCHK(hipMemset(m->d_v, 0, sizeof(float) * m->np));
    m->adam_t = 0;
    return MQE_OK;
}

int mqe_mlp_get_scaling(mqe_mlp *m, float *x_mean, float *x_std, float *y_mean, float *y_std) {
    if (!m || !x_mean || !x_std || !y_mean || !y_std) { set_error("mqe_mlp_get_scaling: null argument"); return MQE_EINVAL; }
    std::vector<float> sc(2 * m->nin + 2 * m->nout);
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(sc.data(), m->d_scale, sizeof(float) * sc.size(), hipMemcpyDeviceToHost));
    std::copy(sc.begin(), sc.begin() + m->nin, x_mean);
    std::copy(sc.begin() + m->nin, sc.begin() + 2 * m->nin, x_std);
    std::copy(sc.begin() + 2 * m->nin, sc.begin() + 2 * m->nin + m->nout, y_mean);
    std::copy(sc.begin() + 2 * m->nin + m->nout, sc.end(), y_std);
    return MQE_OK;
}

int mqe_mlp_set_scaling(mqe_mlp *m, const float *x_mean, const float *x_std, const float *y_mean,
                        const float *y_std) {
    if (!m || !x_mean || !x_std || !y_mean || !y_std) { set_error("mqe_mlp_set_scaling: null argument"); return MQE_EINVAL; }
    std::vector<float> sc;
    sc.insert(sc.end(), x_mean, x_mean + m->nin);
    sc.insert(sc.end(), x_std, x_std + m->nin);
    sc.insert(sc.end(), y_mean, y_mean + m->nout);
    sc.insert(sc.end(), y_std, y_std + m->nout);
    for (int f = 0; f < m->nin; ++f)
        if (!(x_std[f] > 0.f)) { set_error("mqe_mlp_set_scaling: x_std must be positive"); return MQE_EINVAL; }
    for (int q = 0; q < m->nout; ++q)
        if (!(y_std[q] > 0.f)) { set_error("mqe_mlp_set_scaling: y_std must be positive"); return MQE_EINVAL; }
    MCHK(hipStreamSynchronize(m->stream));
    MCHK(hipMemcpy(m->d_scale, sc.data(), sizeof(float) * sc.size(), hipMemcpyHostToDevice));
    return MQE_OK;
}

int mqe_mlp_forward(mqe_mlp *m, int n, const float *X, float *Y) {
    if (!m || !X || !Y || n < 0) { set_error("mqe_mlp_forward: bad argument"); return MQE_EINVAL; }
    if (n == 0) return MQE_OK;
    int r = ensure_data(m, n);
    if (r) return r;
    MCHK(hipMemcpyAsync(m->d_X, X, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    r = mlp_forward_device(m, m->d_X, n, nullptr, 0, m->d_out, m->stream);
    if (r) return r;
    MCHK(hipMemcpyAsync(Y, m->d_out, sizeof(float) * (size_t)n * m->nout, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return MQE_OK;
}

int mqe_mlp_loss_grad(mqe_mlp *m, int n, const float *X, const float *Y, double *loss, float *grad) {
    if (!m || !X || !Y || n < 1) { set_error("mqe_mlp_loss_grad: bad argument"); return MQE_EINVAL; }
    int r = ensure_data(m, n);
    if (r) return r;
    MCHK(hipMemcpyAsync(m->d_X, X, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemcpyAsync(m->d_Y, Y, sizeof(float) * (size_t)n * m->nout, hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemsetAsync(m->d_loss, 0, sizeof(double), m->stream));
    r = grad_pass(m, m->d_X, m->d_Y, n, nullptr);
    if (r) return r;
    const int tiles = (int)((n + mlp::TS - 1) / mlp::TS);
    // grad_out mode of the Adam kernel: the summed gradient, no update
    hipLaunchKernelGGL(mlp::adam_kernel, dim3((unsigned)std::min<long>((m->np + 255) / 256, 1024)), dim3(256), 0,
                       m->stream, m->d_par, m->d_m, m->d_v, m->d_gpart, tiles, m->np, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f,
                       m->d_grad);
    MCHK(hipGetLastError());
    double ls = 0.0;
    MCHK(hipMemcpyAsync(&ls, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    if (grad) MCHK(hipMemcpyAsync(grad, m->d_grad, sizeof(float) * m->np, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    if (loss) *loss = ls / ((double)n * m->nout);
    return MQE_OK;
}

int mqe_mlp_fit(mqe_mlp *m, int n, const float *X, const float *Y, int epochs, int batch, double lr,
                int standardize, double *loss) {
    if (!m || !X || !Y || n < 1) { set_error("mqe_mlp_fit: bad argument"); return MQE_EINVAL; }
    int r = ensure_data(m, n);
    if (r) return r;
    MCHK(hipMemcpyAsync(m->d_X, X, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    MCHK(hipMemcpyAsync(m->d_Y, Y, sizeof(float) * (size_t)n * m->nout, hipMemcpyHostToDevice, m->stream));
    return mlp_fit_device(m, m->d_X, m->d_Y, n, epochs, batch, lr, standardize, loss, m->stream);
}

}  // extern "C"
