/*
 * xo_mlp.hip — the reflectance surrogate: a fully connected ReLU network
 * trained and evaluated with bf16 MFMA matrix products (fp32 accumulation),
 * Adam on fp32 master weights. No counterpart in the reference (its
 * regression utilities, xopto/util/regression/ann.py, call TensorFlow).
 *
 * Storage: weights W_l row-major [in_l][out_l] (so a layer is H W), packed
 * with their biases as W_1, b_1, W_2, b_2, ... in fp32; the products read
 * bf16 copies. Activations between layers are bf16, rounded to nearest
 * even; the network output and the loss are fp32.
 *
 * Products: C[M][N] = op(A)[M][K] op(B)[K][N] with op = identity or
 * transpose, one 64x64 tile of C per 256-thread block (four waves, each a
 * 2x2 grid of 16x16 v_mfma_f32_16x16x32_bf16 tiles), K in steps of 32
 * staged through LDS, out-of-range elements read as zero.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/xomc.h"
#include "xo_ctx.h"
#include "xo_seedhash.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

/* fp32 -> bf16 bits, round to nearest even (NaN stays NaN). */
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

/* Two fp32 values as packed bf16 bits (low half a) with the hardware
 * conversion, which rounds to nearest even: for every non-NaN input the
 * same bits as f2bf. */
__device__ inline uint32_t bf16_pair(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    typedef float floatx2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((floatx2){a, b}, bf16x2));
}

__host__ __device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

enum Epilogue : int {
    EPI_STORE = 0,      /* C fp32                                  */
    EPI_BIAS_RELU = 1,  /* Z = C + b; H = bf16(max(Z, 0))          */
    EPI_BIAS = 2,       /* Y = C + b, fp32                         */
    EPI_MASK = 3,       /* D = bf16(C * (Hprev > 0))               */
};

struct GemmArgs {
    const uint16_t *A;
    const uint16_t *B;
    int lda, ldb;
    int M, N, K;
    float *C;           /* EPI_STORE, EPI_BIAS: fp32 out           */
    int ldc;
    const float *bias;  /* EPI_BIAS_RELU, EPI_BIAS                 */
    uint16_t *H;        /* EPI_BIAS_RELU, EPI_MASK: bf16 out       */
    int ldh;
    const uint16_t *mask; /* EPI_MASK: activations whose sign gates */
    int ldm;
};

constexpr int TM = 64, TN = 64, TK = 32, KPAD = 8;
constexpr uint32_t kMaxSplits = 32;

template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) gemm_kernel(const GemmArgs g) {
    __shared__ __attribute__((aligned(16))) uint16_t sA[TM][TK + KPAD];
    __shared__ __attribute__((aligned(16))) uint16_t sB[TN][TK + KPAD];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    floatx4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    /* split K: block z takes K range [z kc, (z + 1) kc) and, for
     * EPI_STORE, writes its partial product to C + z M ldc */
    const int kc = (g.K + (int)gridDim.z - 1) / (int)gridDim.z;
    const int kbeg = (int)blockIdx.z * kc;
    const int kend = kbeg + kc < g.K ? kbeg + kc : g.K;
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        /* stage op(A)[m0:+64][k0:+32] as sA[m][k], op(B)[k0:+32][n0:+64]
         * as sB[n][k]; each thread moves 8 elements of each */
        for (int e = tid; e < TM * TK; e += 256) {
            int r, c;
            if (TA) { c = e / TM; r = e % TM; } else { r = e / TK; c = e % TK; }
            const int m = m0 + r, k = k0 + c;
            uint16_t v = 0;
            if (m < g.M && k < kend) v = TA ? g.A[(size_t)k * g.lda + m] : g.A[(size_t)m * g.lda + k];
            sA[r][c] = v;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            int r, c; /* r: n, c: k */
            if (TB) { r = e / TK; c = e % TK; } else { c = e / TN; r = e % TN; }
            const int n = n0 + r, k = k0 + c;
            uint16_t v = 0;
            if (n < g.N && k < kend) v = TB ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n];
            sB[r][c] = v;
        }
        __syncthreads();
        /* fragments: lane l holds row (l % 16), k = 8 (l / 16) .. +7 */
        const int fr = lane & 15, fk = (lane >> 4) * 8;
        bf16x8 a[2], b[2];
        for (int i = 0; i < 2; ++i)
            a[i] = *reinterpret_cast<const bf16x8 *>(&sA[wm + i * 16 + fr][fk]);
        for (int j = 0; j < 2; ++j)
            b[j] = *reinterpret_cast<const bf16x8 *>(&sB[wn + j * 16 + fr][fk]);
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    /* result: lane l holds column (l % 16), rows 4 (l / 16) + r */
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
                const int n = n0 + wn + j * 16 + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][r];
                if (EPI == EPI_STORE) {
                    g.C[((size_t)blockIdx.z * g.M + m) * g.ldc + n] = v;
                } else if (EPI == EPI_BIAS) {
                    g.C[(size_t)m * g.ldc + n] = v + g.bias[n];
                } else if (EPI == EPI_BIAS_RELU) {
                    float z = v + g.bias[n];
                    g.H[(size_t)m * g.ldh + n] = f2bf(z > 0.f ? z : 0.f);
                } else {
                    const bool on = bf2f(g.mask[(size_t)m * g.ldm + n]) > 0.f;
                    g.H[(size_t)m * g.ldh + n] = f2bf(on ? v : 0.f);
                }
            }
}

template <bool TA, bool TB, int EPI>
hipError_t gemm(const GemmArgs &g, hipStream_t s, int splits = 1) {
    dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, splits);
    hipLaunchKernelGGL((gemm_kernel<TA, TB, EPI>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

/* C[i] = sum over z of P[z][i], z in order (deterministic). */
__global__ void sum_splits_kernel(const float *P, uint32_t splits, uint64_t n, float *C) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (uint32_t z = 0; z < splits; ++z) s += P[(size_t)z * n + i];
        C[i] = s;
    }
}

/* ------------------------------------------------------------------ */
/* Fused inference: every layer of a row block in one launch            */
/* ------------------------------------------------------------------ */
/* A 256-thread block takes 16 MT rows through the whole network: the
 * activations stay in one LDS tile (16 MT x 256 bf16), each layer is the
 * product of that tile with the layer's weights, read from L2 as bf16 MFMA
 * fragments (fragment-major copies, so a wave's fragment is 1 KB of whole
 * cache lines), bias + ReLU applied in registers and the result written
 * back over the tile. Wave w owns 64 output columns of a hidden layer
 * (16-column tiles w and w + 4 of the last layer) for all rows, so every
 * weight fragment it loads feeds MT MFMAs. The weights are the MFMA's first
 * operand and the activations its second: the 16x16 result tile is
 * transposed (columns x rows), so a lane holds four consecutive output
 * columns of one row and writes them back as one 8-byte LDS store. Layer
 * shapes are compile-time, the k loop is unrolled and the weight fragments
 * are loaded FPD k steps ahead of their use (the activation fragments one
 * step ahead), so the MFMAs of a step issue while the next steps' L2 reads
 * are in flight. Networks: input width <= 32, hidden widths 256, output
 * width <= 128, 2 to 8 layers. */
#ifndef XO_MLP_PD
#define XO_MLP_PD 2
#endif
constexpr int FHID = 256, FSTRIDE = FHID + 8, FMAXL = 8, FPD = XO_MLP_PD;
#ifndef XO_MLP_MT
#define XO_MLP_MT 4
#endif
constexpr int FMT = XO_MLP_MT, FBM = 16 * FMT;

struct FusedArgs {
    const float *X;
    float *Y;
    int M, nin, nout, nlayers;
    const uint16_t *wt[FMAXL];  /* fragment-major: 1 (first) or 8 k steps,
                                   16 (hidden) or 8 (last) column tiles */
    const float *bias[FMAXL];
};

/* One layer of a 16 FMT-row block: KS k steps of 32, NJ 16-column tiles
 * per wave; LAST writes fp32 outputs to Y, otherwise bf16(ReLU) back to
 * hb. */
template <int KS, int NJ, bool LAST>
__device__ __forceinline__ void fused_layer(const FusedArgs &f, int l, uint16_t *hb, int row0,
                                            int lane, int wave) {
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    const uint16_t *W = f.wt[l];
    const uint16_t *wp[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int tile = LAST ? wave + 4 * j : wave * 4 + j;
        wp[j] = W + ((size_t)tile * KS * 64 + lane) * 8;
    }
    floatx4 acc[FMT][NJ];
#pragma unroll
    for (int i = 0; i < FMT; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    bf16x8 b[KS][NJ], a[2][FMT];
#pragma unroll
    for (int s = 0; s < FPD && s < KS; ++s) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) b[s][j] = *reinterpret_cast<const bf16x8 *>(wp[j] + s * 512);
        __builtin_amdgcn_sched_barrier(0); /* issued in the order of use */
    }
#pragma unroll
    for (int i = 0; i < FMT; ++i)
        a[0][i] = *reinterpret_cast<const bf16x8 *>(&hb[(i * 16 + fr) * FSTRIDE + fk]);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s + FPD < KS) {
#pragma unroll
            for (int j = 0; j < NJ; ++j)
                b[s + FPD][j] = *reinterpret_cast<const bf16x8 *>(wp[j] + (s + FPD) * 512);
        }
        if (s + 1 < KS) {
#pragma unroll
            for (int i = 0; i < FMT; ++i)
                a[(s + 1) & 1][i] = *reinterpret_cast<const bf16x8 *>(
                    &hb[(i * 16 + fr) * FSTRIDE + (s + 1) * 32 + fk]);
        }
        /* keep the reads above ahead of this step's MFMAs */
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < FMT; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[s][j], a[s & 1][i],
                                                                   acc[i][j], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads(); /* every wave has read the tile */
    const float *bias = f.bias[l];
    const int nq = (lane >> 4) * 4;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int n0 = (LAST ? wave + 4 * j : wave * 4 + j) * 16 + nq;
        if (!LAST) {
            const float4 bv = *reinterpret_cast<const float4 *>(bias + n0);
#pragma unroll
            for (int i = 0; i < FMT; ++i) {
                const float z0 = acc[i][j][0] + bv.x, z1 = acc[i][j][1] + bv.y;
                const float z2 = acc[i][j][2] + bv.z, z3 = acc[i][j][3] + bv.w;
                uint2 pk;
                pk.x = bf16_pair(z0 > 0.f ? z0 : 0.f, z1 > 0.f ? z1 : 0.f);
                pk.y = bf16_pair(z2 > 0.f ? z2 : 0.f, z3 > 0.f ? z3 : 0.f);
                *reinterpret_cast<uint2 *>(&hb[(i * 16 + fr) * FSTRIDE + n0]) = pk;
            }
        } else {
#pragma unroll
            for (int i = 0; i < FMT; ++i) {
                const int m = row0 + i * 16 + fr;
                if (m >= f.M || n0 >= f.nout) continue;
                float *y = f.Y + (size_t)m * f.nout + n0;
                if ((f.nout & 3) == 0) { /* n0 + 3 < nout, 16-byte aligned */
                    *reinterpret_cast<float4 *>(y) =
                        make_float4(acc[i][j][0] + bias[n0], acc[i][j][1] + bias[n0 + 1],
                                    acc[i][j][2] + bias[n0 + 2], acc[i][j][3] + bias[n0 + 3]);
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (n0 + r < f.nout) y[r] = acc[i][j][r] + bias[n0 + r];
                }
            }
        }
    }
    __syncthreads(); /* the tile holds this layer's output */
}

#ifndef XO_MLP_WPE
#define XO_MLP_WPE 2
#endif
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(XO_MLP_WPE)))
mlp_fused_kernel(const FusedArgs f) {
    __shared__ __attribute__((aligned(16))) uint16_t hb[FBM * FSTRIDE];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row0 = blockIdx.x * FBM;
    for (int e = tid; e < FBM * 32; e += 256) {
        const int m = e >> 5, c = e & 31;
        const float v = (row0 + m < f.M && c < f.nin) ? f.X[(size_t)(row0 + m) * f.nin + c] : 0.f;
        hb[m * FSTRIDE + c] = f2bf(v);
    }
    __syncthreads();
    fused_layer<1, 4, false>(f, 0, hb, row0, lane, wave);
    for (int l = 1; l + 1 < f.nlayers; ++l) fused_layer<8, 4, false>(f, l, hb, row0, lane, wave);
    fused_layer<8, 2, true>(f, f.nlayers - 1, hb, row0, lane, wave);
}

/* The weights of one layer as bf16 MFMA fragments: for 16-column tile t,
 * 32-wide k step s and lane L, the 8 values W[k][n] with
 * n = 16 t + (L & 15), k = 32 s + 8 (L >> 4) + e (e = 0..7) are stored at
 * ((t * ksteps + s) * 64 + L) * 8 + e, zero for k >= in or n >= out. A
 * wave's fragment of one tile and step is then 1 KB of whole cache lines. */
__global__ void fragment_pad_kernel(const float *w, int in, int out, int ksteps, int ntiles,
                                    uint16_t *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntiles * ksteps * 512) return;
    const int e = i & 7, L = (i >> 3) & 63, ts = i >> 9;
    const int t = ts / ksteps, st = ts % ksteps;
    const int n = t * 16 + (L & 15), k = st * 32 + (L >> 4) * 8 + e;
    dst[i] = (k < in && n < out) ? f2bf(w[(size_t)k * out + n]) : (uint16_t)0;
}

__global__ void to_bf16_kernel(const float *x, uint16_t *y, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

/* Minibatch of rows drawn with replacement: row i of step t is
 * hash(seed, t, i) mod n. */
__global__ void gather_kernel(const float *X, const float *T, uint32_t n, uint32_t nin,
                              uint32_t nout, uint32_t B, uint64_t seed, const uint32_t *step,
                              uint16_t *Xb, float *Tb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const uint64_t h = xo_mix64(seed + 0x9E3779B97F4A7C15ull * (((uint64_t)*step << 32) + i + 1));
    const uint32_t j = (uint32_t)(h % n);
    for (uint32_t c = 0; c < nin; ++c) Xb[(size_t)i * nin + c] = f2bf(X[(size_t)j * nin + c]);
    for (uint32_t c = 0; c < nout; ++c) Tb[(size_t)i * nout + c] = T[(size_t)j * nout + c];
}

/* dY = bf16(2 (Y - T) / (B nout)); the mean squared error is added to
 * loss[0], or to loss[*step % ring] when ring > 0 (one slot per training
 * step, so a captured step needs no per-step pointer). */
__global__ void loss_kernel(const float *Y, const float *T, uint32_t count, float scale,
                            uint16_t *dY, float *loss, const uint32_t *step, uint32_t ring) {
    __shared__ float part[256];
    float s = 0.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        s += d * d;
        dY[i] = f2bf(2.f * d * scale);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(ring ? loss + (*step % ring) : loss, part[0] * scale);
}

/* db[n] = sum_m bf16 D[m][n]: 64 columns per block, four row groups with
 * eight partial sums each, combined in a fixed order. */
__global__ void __launch_bounds__(256) colsum_kernel(const uint16_t *D, uint32_t M, uint32_t N,
                                                     float *db) {
    __shared__ float part[4][64];
    const uint32_t c = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const uint32_t n = blockIdx.x * 64 + c;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (n < N) {
        uint32_t m = grp;
        for (; m + 28 < M; m += 32)
            for (int u = 0; u < 8; ++u) acc[u] += bf2f(D[(size_t)(m + 4 * u) * N + n]);
        for (; m < M; m += 4) acc[0] += bf2f(D[(size_t)m * N + n]);
    }
    part[grp][c] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    __syncthreads();
    if (grp == 0 && n < N) db[n] = (part[0][c] + part[1][c]) + (part[2][c] + part[3][c]);
}

__global__ void fill_f32_kernel(float *p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* Adam (t = *step + 1), fp32 master weights; refreshes the bf16 copies of
 * the weight matrices. */
__global__ void adam_kernel(float *w, float *m, float *v, const float *g, uint64_t n,
                            const uint32_t *step, float lr, float b1, float b2, float eps,
                            const uint8_t *is_matrix, const uint64_t *bf_index,
                            uint16_t *wbf) {
    const float t = (float)(*step + 1u);
    const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (is_matrix[i]) wbf[bf_index[i]] = f2bf(wi);
    }
}

__global__ void step_kernel(uint32_t *step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1u;
}

} // namespace

struct xo_mlp {
    xo_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    int nlayers = 0;
    std::vector<int> dims;       /* nlayers + 1 widths */
    std::vector<uint64_t> woff;  /* offset of W_l in the packed parameters */
    std::vector<uint64_t> boff;  /* offset of b_l */
    std::vector<uint64_t> wbf_off; /* offset of W_l in the bf16 copy */
    uint64_t nparams = 0, nwbf = 0;
    uint32_t max_batch = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *wbf = nullptr;
    uint8_t *is_matrix = nullptr;
    uint64_t *bf_index = nullptr;
    uint16_t *xb = nullptr;           /* [B][in] */
    float *xf = nullptr;              /* [B][in] host-provided rows */
    float *tb = nullptr;              /* [B][out] */
    std::vector<uint16_t *> h;        /* activations of hidden layers */
    uint16_t *d0 = nullptr, *d1 = nullptr; /* [B][max width] gradients */
    float *y = nullptr;               /* [B][out] */
    uint16_t *wt = nullptr;           /* padded transposed bf16 weights */
    std::vector<uint64_t> wt_off;
    bool fusable = false, wt_dirty = true;
    float *work = nullptr;            /* split-K partial products */
    float *loss = nullptr;            /* per training step */
    uint32_t loss_cap = 0;
    uint32_t *step = nullptr;
    float *data_x = nullptr, *data_t = nullptr;
    uint32_t ndata = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t graph_batch = 0;
    uint64_t graph_seed = 0;
    float graph_lr = 0.f;
    float *graph_loss = nullptr;
    std::mutex mu;
};

namespace {

#define MLP_CHECK(call)                                                               \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            xo::set_error("%s failed: %s", #call, hipGetErrorString(e_));             \
            return -1;                                                                \
        }                                                                             \
    } while (0)

int widest(const xo_mlp *N) {
    int w = 0;
    for (int d : N->dims) w = d > w ? d : w;
    return w;
}

/* Forward pass of B rows already in N->xb; output fp32 in N->y. */
int forward(xo_mlp *N, uint32_t B) {
    for (int l = 0; l < N->nlayers; ++l) {
        GemmArgs g;
        std::memset(&g, 0, sizeof(g));
        g.A = l == 0 ? N->xb : N->h[l - 1];
        g.lda = N->dims[l];
        g.B = N->wbf + N->wbf_off[l];
        g.ldb = N->dims[l + 1];
        g.M = (int)B;
        g.N = N->dims[l + 1];
        g.K = N->dims[l];
        g.bias = N->w + N->boff[l];
        if (l + 1 < N->nlayers) {
            g.H = N->h[l];
            g.ldh = N->dims[l + 1];
            MLP_CHECK((gemm<false, false, EPI_BIAS_RELU>(g, N->stream)));
        } else {
            g.C = N->y;
            g.ldc = N->dims[l + 1];
            MLP_CHECK((gemm<false, false, EPI_BIAS>(g, N->stream)));
        }
    }
    return 0;
}

/* Loss and gradient of B rows (xb, tb): loss added to loss[0] (ring 0) or
 * to the step's ring slot, gradient in N->grad. */
int backward(xo_mlp *N, uint32_t B, float *loss, uint32_t ring) {
    const int L = N->nlayers;
    const uint32_t nout = (uint32_t)N->dims[L];
    const uint32_t count = B * nout;
    hipLaunchKernelGGL(loss_kernel, dim3((count + 255) / 256 < 1024 ? (count + 255) / 256 : 1024),
                       dim3(256), 0, N->stream, N->y, N->tb, count,
                       1.0f / (float)count, N->d0, loss, N->step, ring);
    MLP_CHECK(hipGetLastError());
    uint16_t *dcur = N->d0, *dnext = N->d1;
    for (int l = L - 1; l >= 0; --l) {
        const int in = N->dims[l], out = N->dims[l + 1];
        const uint16_t *Hin = l == 0 ? N->xb : N->h[l - 1];
        /* dW_l = H_in^T dZ_l */
        GemmArgs g;
        std::memset(&g, 0, sizeof(g));
        g.A = Hin;
        g.lda = in;
        g.B = dcur;
        g.ldb = out;
        g.M = in;
        g.N = out;
        g.K = (int)B;
        /* few output tiles and a long K (the batch): split K over blocks
         * into the workspace, then add the partial products in order */
        const int tiles = ((in + TM - 1) / TM) * ((out + TN - 1) / TN);
        int splits = (int)std::min<uint32_t>(kMaxSplits, std::max<uint32_t>(1u, 256u / tiles));
        splits = std::min(splits, std::max(1, (int)B / 128));
        g.C = splits > 1 ? N->work : N->grad + N->woff[l];
        g.ldc = out;
        MLP_CHECK((gemm<true, false, EPI_STORE>(g, N->stream, splits)));
        if (splits > 1) {
            const uint64_t n = (uint64_t)in * out;
            hipLaunchKernelGGL(sum_splits_kernel, dim3(256), dim3(256), 0, N->stream, N->work,
                               (uint32_t)splits, n, N->grad + N->woff[l]);
            MLP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL(colsum_kernel, dim3((out + 63) / 64), dim3(256), 0, N->stream, dcur,
                           B, (uint32_t)out, N->grad + N->boff[l]);
        MLP_CHECK(hipGetLastError());
        if (l == 0) break;
        /* dZ_{l-1} = (dZ_l W_l^T) * (H_{l-1} > 0) */
        std::memset(&g, 0, sizeof(g));
        g.A = dcur;
        g.lda = out;
        g.B = N->wbf + N->wbf_off[l];
        g.ldb = out;
        g.M = (int)B;
        g.N = in;
        g.K = out;
        g.H = dnext;
        g.ldh = in;
        g.mask = N->h[l - 1];
        g.ldm = in;
        MLP_CHECK((gemm<false, true, EPI_MASK>(g, N->stream)));
        uint16_t *t = dcur;
        dcur = dnext;
        dnext = t;
    }
    return 0;
}

int adam(xo_mlp *N, float lr) {
    hipLaunchKernelGGL(adam_kernel, dim3(256), dim3(256), 0, N->stream, N->w, N->m, N->v,
                       N->grad, N->nparams, N->step, lr, 0.9f, 0.999f, 1e-8f, N->is_matrix,
                       N->bf_index, N->wbf);
    MLP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(step_kernel, dim3(1), dim3(64), 0, N->stream, N->step);
    MLP_CHECK(hipGetLastError());
    return 0;
}

/* One training step on a device-sampled minibatch (capturable: every
 * argument is fixed, the step counter lives on the device). */
int train_step(xo_mlp *N, uint32_t B, uint64_t seed, float lr) {
    hipLaunchKernelGGL(gather_kernel, dim3((B + 255) / 256), dim3(256), 0, N->stream,
                       N->data_x, N->data_t, N->ndata, (uint32_t)N->dims[0],
                       (uint32_t)N->dims[N->nlayers], B, seed, N->step, N->xb, N->tb);
    MLP_CHECK(hipGetLastError());
    if (forward(N, B)) return -1;
    if (backward(N, B, N->loss, N->loss_cap)) return -1;
    return adam(N, lr);
}

int refresh_bf16(xo_mlp *N) {
    /* W copies from the master weights */
    for (int l = 0; l < N->nlayers; ++l) {
        const uint64_t n = (uint64_t)N->dims[l] * N->dims[l + 1];
        hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, N->stream,
                           N->w + N->woff[l], N->wbf + N->wbf_off[l], n);
        MLP_CHECK(hipGetLastError());
    }
    return 0;
}

int refresh_transposed(xo_mlp *N) {
    for (int l = 0; l < N->nlayers; ++l) {
        const int ksteps = l == 0 ? 1 : FHID / 32, ntiles = l + 1 == N->nlayers ? 8 : FHID / 16;
        hipLaunchKernelGGL(fragment_pad_kernel, dim3(ntiles * ksteps * 2), dim3(256), 0,
                           N->stream, N->w + N->woff[l], N->dims[l], N->dims[l + 1], ksteps,
                           ntiles, N->wt + N->wt_off[l]);
        MLP_CHECK(hipGetLastError());
    }
    N->wt_dirty = false;
    return 0;
}

/* Fused forward pass of B rows in N->xf (fp32) into N->y. */
int forward_fused(xo_mlp *N, uint32_t B) {
    FusedArgs f;
    std::memset(&f, 0, sizeof(f));
    f.X = N->xf;
    f.Y = N->y;
    f.M = (int)B;
    f.nin = N->dims[0];
    f.nout = N->dims[N->nlayers];
    f.nlayers = N->nlayers;
    for (int l = 0; l < N->nlayers; ++l) {
        f.wt[l] = N->wt + N->wt_off[l];
        f.bias[l] = N->w + N->boff[l];
    }
    hipLaunchKernelGGL(mlp_fused_kernel, dim3((B + FBM - 1) / FBM), dim3(256), 0, N->stream, f);
    MLP_CHECK(hipGetLastError());
    return 0;
}

int upload_rows(xo_mlp *N, const float *x, uint32_t B) {
    MLP_CHECK(hipMemcpyAsync(N->xf, x, sizeof(float) * B * N->dims[0], hipMemcpyHostToDevice,
                             N->stream));
    hipLaunchKernelGGL(to_bf16_kernel, dim3(64), dim3(256), 0, N->stream, N->xf, N->xb,
                       (uint64_t)B * N->dims[0]);
    MLP_CHECK(hipGetLastError());
    return 0;
}

} // namespace

extern "C" {

void xo_mlp_destroy(xo_mlp *N) {
    if (!N) return;
    (void)hipSetDevice(N->ctx->device);
    if (N->stream) (void)hipStreamSynchronize(N->stream);
    if (N->exec) (void)hipGraphExecDestroy(N->exec);
    if (N->graph) (void)hipGraphDestroy(N->graph);
    void *bufs[] = {N->w, N->m, N->v, N->grad, N->wbf, N->is_matrix, N->bf_index, N->xb,
                    N->xf, N->tb, N->d0, N->d1, N->y, N->work, N->loss, N->step,
                    N->data_x, N->data_t, N->wt};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (auto p : N->h)
        if (p) (void)hipFree(p);
    if (N->stream) (void)hipStreamDestroy(N->stream);
    delete N;
}

xo_mlp *xo_mlp_create(xo_ctx *ctx, const uint32_t *dims, uint32_t nlayers, uint32_t max_batch) {
    xo::clear_error();
    if (!ctx || !dims || nlayers < 1 || nlayers > 16 || max_batch == 0 || max_batch > (1u << 20)) {
        xo::set_error("xo_mlp_create: invalid argument");
        return nullptr;
    }
    for (uint32_t l = 0; l <= nlayers; ++l)
        if (dims[l] == 0 || dims[l] > 4096) {
            xo::set_error("xo_mlp_create: layer width %u outside 1..4096", dims[l]);
            return nullptr;
        }
    xo_mlp *N = new xo_mlp();
    N->ctx = ctx;
    N->nlayers = (int)nlayers;
    N->dims.assign(dims, dims + nlayers + 1);
    N->max_batch = max_batch;
    for (uint32_t l = 0; l < nlayers; ++l) {
        N->woff.push_back(N->nparams);
        N->nparams += (uint64_t)dims[l] * dims[l + 1];
        N->boff.push_back(N->nparams);
        N->nparams += dims[l + 1];
        N->wbf_off.push_back(N->nwbf);
        N->nwbf += (uint64_t)dims[l] * dims[l + 1];
    }
    auto bad = [&](const char *what, hipError_t e) -> xo_mlp * {
        xo::set_error("xo_mlp_create: %s: %s", what, hipGetErrorString(e));
        xo_mlp_destroy(N);
        return nullptr;
    };
    hipError_t e;
    if ((e = hipSetDevice(ctx->device)) != hipSuccess) return bad("device", e);
    if ((e = hipStreamCreateWithFlags(&N->stream, hipStreamNonBlocking)) != hipSuccess)
        return bad("stream", e);
    const int W = widest(N);
    const uint64_t P = N->nparams;
    uint64_t max_matrix = 0;
    for (uint32_t l = 0; l < nlayers; ++l)
        max_matrix = std::max<uint64_t>(max_matrix, (uint64_t)dims[l] * dims[l + 1]);
    struct {
        void **p;
        size_t bytes;
    } allocs[] = {
        {(void **)&N->w, 4 * P}, {(void **)&N->m, 4 * P}, {(void **)&N->v, 4 * P},
        {(void **)&N->grad, 4 * P}, {(void **)&N->wbf, 2 * N->nwbf},
        {(void **)&N->is_matrix, P}, {(void **)&N->bf_index, 8 * P},
        {(void **)&N->xb, 2ull * max_batch * dims[0]}, {(void **)&N->xf, 4ull * max_batch * dims[0]},
        {(void **)&N->tb, 4ull * max_batch * dims[nlayers]},
        {(void **)&N->d0, 2ull * max_batch * W}, {(void **)&N->d1, 2ull * max_batch * W},
        {(void **)&N->y, 4ull * max_batch * dims[nlayers]}, {(void **)&N->step, 4},
        {(void **)&N->work, 4ull * kMaxSplits * max_matrix},
    };
    for (auto &a : allocs)
        if ((e = hipMalloc(a.p, a.bytes)) != hipSuccess) return bad("allocation", e);
    for (uint32_t l = 0; l + 1 < nlayers; ++l) {
        uint16_t *p = nullptr;
        if ((e = hipMalloc(&p, 2ull * max_batch * dims[l + 1])) != hipSuccess)
            return bad("allocation", e);
        N->h.push_back(p);
    }
    /* the fused inference kernel covers input <= 32, hidden 256, output
     * <= 128, at most 8 layers */
    N->fusable = nlayers >= 2 && nlayers <= (uint32_t)FMAXL && dims[0] <= 32 &&
                 dims[nlayers] <= 128;
    for (uint32_t l = 1; l < nlayers; ++l) N->fusable = N->fusable && dims[l] == (uint32_t)FHID;
    if (N->fusable) {
        uint64_t total = 0;
        for (uint32_t l = 0; l < nlayers; ++l) {
            N->wt_off.push_back(total);
            const uint64_t kpad = l == 0 ? 32 : FHID, npad = l + 1 == nlayers ? 128 : FHID;
            total += kpad * npad;
        }
        if ((e = hipMalloc(&N->wt, 2 * total)) != hipSuccess) return bad("allocation", e);
    }
    /* which parameters are matrix entries, and where their bf16 copy is */
    std::vector<uint8_t> ism(P, 0);
    std::vector<uint64_t> idx(P, 0);
    for (uint32_t l = 0; l < nlayers; ++l) {
        const uint64_t n = (uint64_t)dims[l] * dims[l + 1];
        for (uint64_t i = 0; i < n; ++i) {
            ism[N->woff[l] + i] = 1;
            idx[N->woff[l] + i] = N->wbf_off[l] + i;
        }
    }
    if ((e = hipMemcpy(N->is_matrix, ism.data(), P, hipMemcpyHostToDevice)) != hipSuccess)
        return bad("upload", e);
    if ((e = hipMemcpy(N->bf_index, idx.data(), 8 * P, hipMemcpyHostToDevice)) != hipSuccess)
        return bad("upload", e);
    if ((e = hipMemset(N->w, 0, 4 * P)) != hipSuccess) return bad("clear", e);
    if ((e = hipMemset(N->m, 0, 4 * P)) != hipSuccess) return bad("clear", e);
    if ((e = hipMemset(N->v, 0, 4 * P)) != hipSuccess) return bad("clear", e);
    if ((e = hipMemset(N->wbf, 0, 2 * N->nwbf)) != hipSuccess) return bad("clear", e);
    if ((e = hipMemset(N->step, 0, 4)) != hipSuccess) return bad("clear", e);
    return N;
}

uint64_t xo_mlp_param_count(xo_mlp *N) { return N ? N->nparams : 0; }

int xo_mlp_set_params(xo_mlp *N, const float *params, uint64_t n) {
    xo::clear_error();
    if (!N || !params || n != N->nparams) {
        xo::set_error("xo_mlp_set_params: expected %llu parameters",
                      (unsigned long long)(N ? N->nparams : 0));
        return -1;
    }
    std::lock_guard<std::mutex> lock(N->mu);
    MLP_CHECK(hipSetDevice(N->ctx->device));
    MLP_CHECK(hipMemcpyAsync(N->w, params, 4 * n, hipMemcpyHostToDevice, N->stream));
    MLP_CHECK(hipMemsetAsync(N->m, 0, 4 * n, N->stream));
    MLP_CHECK(hipMemsetAsync(N->v, 0, 4 * n, N->stream));
    MLP_CHECK(hipMemsetAsync(N->step, 0, 4, N->stream));
    if (refresh_bf16(N)) return -1;
    N->wt_dirty = true;
    MLP_CHECK(hipStreamSynchronize(N->stream));
    return 0;
}

int xo_mlp_get_params(xo_mlp *N, float *params, uint64_t n) {
    xo::clear_error();
    if (!N || !params || n != N->nparams) {
        xo::set_error("xo_mlp_get_params: expected %llu parameters",
                      (unsigned long long)(N ? N->nparams : 0));
        return -1;
    }
    std::lock_guard<std::mutex> lock(N->mu);
    MLP_CHECK(hipSetDevice(N->ctx->device));
    MLP_CHECK(hipMemcpyAsync(params, N->w, 4 * n, hipMemcpyDeviceToHost, N->stream));
    MLP_CHECK(hipStreamSynchronize(N->stream));
    return 0;
}

int xo_mlp_forward(xo_mlp *N, const float *x, uint64_t n, float *y, double *kernel_ms) {
    xo::clear_error();
    if (!N || (!x && n) || (!y && n)) {
        xo::set_error("xo_mlp_forward: null argument");
        return -1;
    }
    std::lock_guard<std::mutex> lock(N->mu);
    MLP_CHECK(hipSetDevice(N->ctx->device));
    const uint32_t nin = N->dims[0], nout = N->dims[N->nlayers];
    hipEvent_t e0, e1;
    MLP_CHECK(hipEventCreate(&e0));
    MLP_CHECK(hipEventCreate(&e1));
    double total = 0.0;
    int rc = 0;
    if (N->fusable && N->wt_dirty && refresh_transposed(N)) rc = -1;
    for (uint64_t r = 0; r < n && rc == 0; r += N->max_batch) {
        const uint32_t B = (uint32_t)(n - r < N->max_batch ? n - r : N->max_batch);
        if (N->fusable) {
            if (hipMemcpyAsync(N->xf, x + r * nin, sizeof(float) * B * nin,
                               hipMemcpyHostToDevice, N->stream) != hipSuccess) { rc = -1; break; }
        } else if (upload_rows(N, x + r * nin, B)) { rc = -1; break; }
        if (hipEventRecord(e0, N->stream) != hipSuccess) { rc = -1; break; }
        if (N->fusable ? forward_fused(N, B) : forward(N, B)) { rc = -1; break; }
        if (hipEventRecord(e1, N->stream) != hipSuccess) { rc = -1; break; }
        if (hipMemcpyAsync(y + r * nout, N->y, 4ull * B * nout, hipMemcpyDeviceToHost,
                           N->stream) != hipSuccess ||
            hipStreamSynchronize(N->stream) != hipSuccess) {
            xo::set_error("xo_mlp_forward: device failure");
            rc = -1;
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = total;
    return rc;
}

int xo_mlp_gradient(xo_mlp *N, const float *x, const float *t, uint32_t n, float *grad,
                    float *loss) {
    xo::clear_error();
    if (!N || !x || !t || !grad || !loss || n == 0 || n > N->max_batch) {
        xo::set_error("xo_mlp_gradient: invalid argument (1 <= n <= %u)", N ? N->max_batch : 0);
        return -1;
    }
    std::lock_guard<std::mutex> lock(N->mu);
    MLP_CHECK(hipSetDevice(N->ctx->device));
    const uint32_t nout = N->dims[N->nlayers];
    if (upload_rows(N, x, n)) return -1;
    MLP_CHECK(hipMemcpyAsync(N->tb, t, 4ull * n * nout, hipMemcpyHostToDevice, N->stream));
    float *dloss = nullptr;
    MLP_CHECK(hipMalloc(&dloss, 4));
    int rc = 0;
    if (hipMemsetAsync(dloss, 0, 4, N->stream) != hipSuccess || forward(N, n) ||
        backward(N, n, dloss, 0))
        rc = -1;
    if (rc == 0 &&
        (hipMemcpyAsync(grad, N->grad, 4 * N->nparams, hipMemcpyDeviceToHost, N->stream) !=
             hipSuccess ||
         hipMemcpyAsync(loss, dloss, 4, hipMemcpyDeviceToHost, N->stream) != hipSuccess ||
         hipStreamSynchronize(N->stream) != hipSuccess)) {
        xo::set_error("xo_mlp_gradient: device failure");
        rc = -1;
    }
    (void)hipFree(dloss);
    return rc;
}

int xo_mlp_set_data(xo_mlp *N, const float *x, const float *t, uint32_t n) {
    xo::clear_error();
    if (!N || !x || !t || n == 0) {
        xo::set_error("xo_mlp_set_data: invalid argument");
        return -1;
    }
    std::lock_guard<std::mutex> lock(N->mu);
    MLP_CHECK(hipSetDevice(N->ctx->device));
    MLP_CHECK(hipStreamSynchronize(N->stream));
    if (N->data_x) (void)hipFree(N->data_x);
    if (N->data_t) (void)hipFree(N->data_t);
    N->data_x = N->data_t = nullptr;
    const uint32_t nin = N->dims[0], nout = N->dims[N->nlayers];
    MLP_CHECK(hipMalloc(&N->data_x, 4ull * n * nin));
    MLP_CHECK(hipMalloc(&N->data_t, 4ull * n * nout));
    MLP_CHECK(hipMemcpy(N->data_x, x, 4ull * n * nin, hipMemcpyHostToDevice));
    MLP_CHECK(hipMemcpy(N->data_t, t, 4ull * n * nout, hipMemcpyHostToDevice));
    N->ndata = n;
    /* a captured step names the data buffers: capture again */
    if (N->exec) (void)hipGraphExecDestroy(N->exec);
    if (N->graph) (void)hipGraphDestroy(N->graph);
    N->exec = nullptr;
    N->graph = nullptr;
    return 0;
}

int xo_mlp_train(xo_mlp *N, uint32_t steps, uint32_t batch, float lr, uint64_t seed,
                 int use_graph, float *losses, double *kernel_ms) {
    xo::clear_error();
    if (!N || !losses || batch == 0 || batch > N->max_batch || !N->ndata) {
        xo::set_error("xo_mlp_train: invalid argument or no data (batch <= %u)",
                      N ? N->max_batch : 0);
        return -1;
    }
    if (steps == 0) return 0;
    std::lock_guard<std::mutex> lock(N->mu);
    N->wt_dirty = true;
    MLP_CHECK(hipSetDevice(N->ctx->device));
    if (steps > N->loss_cap) {
        if (N->loss) (void)hipFree(N->loss);
        N->loss = nullptr;
        N->loss_cap = 0;
        MLP_CHECK(hipMalloc(&N->loss, 4ull * steps));
        N->loss_cap = steps;
        if (N->exec) (void)hipGraphExecDestroy(N->exec);
        if (N->graph) (void)hipGraphDestroy(N->graph);
        N->exec = nullptr;
        N->graph = nullptr;
    }
    if (N->exec && (N->graph_batch != batch || N->graph_seed != seed || N->graph_lr != lr)) {
        (void)hipGraphExecDestroy(N->exec);
        (void)hipGraphDestroy(N->graph);
        N->exec = nullptr;
        N->graph = nullptr;
    }
    if (use_graph && !N->exec) {
        MLP_CHECK(hipStreamBeginCapture(N->stream, hipStreamCaptureModeThreadLocal));
        const int rc = train_step(N, batch, seed, lr);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(N->stream, &graph);
        if (rc || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (!rc) xo::set_error("xo_mlp_train: capture failed: %s", hipGetErrorString(e));
            return -1;
        }
        N->graph = graph;
        MLP_CHECK(hipGraphInstantiate(&N->exec, N->graph, nullptr, nullptr, 0));
        N->graph_batch = batch;
        N->graph_seed = seed;
        N->graph_lr = lr;
    }
    uint32_t step0 = 0;
    MLP_CHECK(hipMemcpyAsync(&step0, N->step, 4, hipMemcpyDeviceToHost, N->stream));
    /* losses start at zero: a fill kernel, outside the graph */
    hipLaunchKernelGGL(fill_f32_kernel, dim3((N->loss_cap + 255) / 256), dim3(256), 0,
                       N->stream, N->loss, N->loss_cap, 0.f);
    MLP_CHECK(hipGetLastError());
    hipEvent_t e0, e1;
    MLP_CHECK(hipEventCreate(&e0));
    MLP_CHECK(hipEventCreate(&e1));
    int rc = 0;
    if (hipEventRecord(e0, N->stream) != hipSuccess) rc = -1;
    for (uint32_t s = 0; s < steps && rc == 0; ++s) {
        if (use_graph) {
            if (hipGraphLaunch(N->exec, N->stream) != hipSuccess) rc = -1;
        } else if (train_step(N, batch, seed, lr)) {
            rc = -1;
        }
    }
    if (rc == 0 && hipEventRecord(e1, N->stream) != hipSuccess) rc = -1;
    std::vector<float> ring(N->loss_cap);
    if (rc == 0 &&
        (hipMemcpyAsync(ring.data(), N->loss, 4ull * N->loss_cap, hipMemcpyDeviceToHost,
                        N->stream) != hipSuccess ||
         hipStreamSynchronize(N->stream) != hipSuccess)) {
        xo::set_error("xo_mlp_train: device failure");
        rc = -1;
    }
    if (rc == 0)
        for (uint32_t s = 0; s < steps; ++s) losses[s] = ring[(step0 + s) % N->loss_cap];
    float ms = 0.f;
    if (rc == 0) (void)hipEventElapsedTime(&ms, e0, e1);
    if (kernel_ms) *kernel_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

/* bf16 product for tests: A, B given in fp32 (rounded to bf16 here),
 * C = op(A) op(B) in fp32. */
int xo_gemm_bf16(xo_ctx *ctx, int trans_a, int trans_b, uint32_t M, uint32_t N, uint32_t K,
                 const float *A, uint32_t lda, const float *B, uint32_t ldb, float *C,
                 uint32_t ldc) {
    xo::clear_error();
    if (!ctx || !A || !B || !C || M == 0 || N == 0 || K == 0) {
        xo::set_error("xo_gemm_bf16: invalid argument");
        return -1;
    }
    const uint64_t na = trans_a ? (uint64_t)(K - 1) * lda + M : (uint64_t)(M - 1) * lda + K;
    const uint64_t nb = trans_b ? (uint64_t)(N - 1) * ldb + K : (uint64_t)(K - 1) * ldb + N;
    if ((trans_a ? lda < M : lda < K) || (trans_b ? ldb < K : ldb < N) || ldc < N) {
        xo::set_error("xo_gemm_bf16: leading dimension too small");
        return -1;
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    MLP_CHECK(hipSetDevice(ctx->device));
    std::vector<uint16_t> ha(na), hb(nb);
    for (uint64_t i = 0; i < na; ++i) ha[i] = f2bf(A[i]);
    for (uint64_t i = 0; i < nb; ++i) hb[i] = f2bf(B[i]);
    uint16_t *da = nullptr, *db = nullptr;
    float *dc = nullptr;
    const uint64_t nc = (uint64_t)(M - 1) * ldc + N;
    MLP_CHECK(hipMalloc(&da, 2 * na));
    int rc = 0;
    if (hipMalloc(&db, 2 * nb) != hipSuccess || hipMalloc(&dc, 4 * nc) != hipSuccess) rc = -1;
    GemmArgs g;
    std::memset(&g, 0, sizeof(g));
    g.A = da;
    g.B = db;
    g.lda = (int)lda;
    g.ldb = (int)ldb;
    g.M = (int)M;
    g.N = (int)N;
    g.K = (int)K;
    g.C = dc;
    g.ldc = (int)ldc;
    hipStream_t s = ctx->stream;
    if (rc == 0 && (hipMemcpy(da, ha.data(), 2 * na, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(db, hb.data(), 2 * nb, hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemset(dc, 0, 4 * nc) != hipSuccess))
        rc = -1;
    if (rc == 0) {
        hipError_t e;
        if (!trans_a && !trans_b) e = gemm<false, false, EPI_STORE>(g, s);
        else if (trans_a && !trans_b) e = gemm<true, false, EPI_STORE>(g, s);
        else if (!trans_a && trans_b) e = gemm<false, true, EPI_STORE>(g, s);
        else e = gemm<true, true, EPI_STORE>(g, s);
        if (e != hipSuccess || hipStreamSynchronize(s) != hipSuccess) rc = -1;
    }
    if (rc == 0 && hipMemcpy(C, dc, 4 * nc, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
    if (rc) xo::set_error("xo_gemm_bf16: device failure");
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dc) (void)hipFree(dc);
    return rc;
}

} // extern "C"
