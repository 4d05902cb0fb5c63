/*
 * xo_mlp.hip — the reflectance surrogate: a fully connected ReLU network
 * trained and evaluated with bf16 MFMA matrix products (fp32 accumulation),
 * Adam on fp32 master weights. No counterpart in the reference (its
 * regression utilities, xopto/util/regression/ann.py, call TensorFlow).
 *
 * Storage: weights W_l row-major [in][out], packed with their biases as
 * W_1, b_1, W_2, b_2, ... in fp32; the products read bf16 copies of W_l
 * ([in][out]) and W_l^T ([out][in]) whose rows are padded to multiples of
 * 8 elements. Activations between layers are bf16, rounded to nearest
 * even; the network output and the loss are fp32.
 *
 * Training step (captured once as a hipGraph). Networks of input width
 * <= 32, hidden widths 256, output width <= 128 and at most six layers:
 * two fused launches, the forward and backward passes of 16-row blocks
 * with the activations resident in LDS, then every layer's weight and bias
 * gradients over the whole batch with Adam applied in place. Other shapes:
 * minibatch rows, one product per layer forward (the loss fused into the
 * last), one product per layer for the data gradients, one launch for
 * every layer's weight and bias gradients, Adam.
 * Inference: one fused kernel with the activations resident in LDS.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
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

/* The bf16 product of xo_gemm_bf16: C[M][N] = op(A) op(B), fp32, with op
 * identity or transpose, one 64x64 tile of C per 256-thread block (four
 * waves, each a 2x2 grid of 16x16 v_mfma_f32_16x16x32_bf16 tiles), K in
 * steps of 32 staged through LDS, out-of-range elements read as zero. */
struct GemmArgs {
    const uint16_t *A;
    const uint16_t *B;
    int lda, ldb;
    int M, N, K;
    float *C;
    int ldc;
};

constexpr int TM = 64, TN = 64, TK = 32, KPAD = 8;

template <bool TA, bool TB>
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
    for (int k0 = 0; k0 < g.K; k0 += TK) {
        /* stage op(A)[m0:+64][k0:+32] as sA[m][k], op(B)[k0:+32][n0:+64]
         * as sB[n][k]; each thread moves 8 elements of each */
        for (int e = tid; e < TM * TK; e += 256) {
            int r, c;
            if (TA) { c = e / TM; r = e % TM; } else { r = e / TK; c = e % TK; }
            const int m = m0 + r, k = k0 + c;
            uint16_t v = 0;
            if (m < g.M && k < g.K) v = TA ? g.A[(size_t)k * g.lda + m] : g.A[(size_t)m * g.lda + k];
            sA[r][c] = v;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            int r, c; /* r: n, c: k */
            if (TB) { r = e / TK; c = e % TK; } else { c = e / TN; r = e % TN; }
            const int n = n0 + r, k = k0 + c;
            uint16_t v = 0;
            if (n < g.N && k < g.K) v = TB ? g.B[(size_t)n * g.ldb + k] : g.B[(size_t)k * g.ldb + n];
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
                if (m < g.M && n < g.N) g.C[(size_t)m * g.ldc + n] = acc[i][j][r];
            }
}

template <bool TA, bool TB> hipError_t gemm(const GemmArgs &g, hipStream_t s) {
    dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM, 1);
    hipLaunchKernelGGL((gemm_kernel<TA, TB>), grid, dim3(256), 0, s, g);
    return hipGetLastError();
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

/* A matrix M[k][n] (k < kdim, n < ndim) as bf16 MFMA fragments: for
 * 16-column tile t, 32-wide k step s and lane L, the 8 values M[k][n] with
 * n = 16 t + (L & 15), k = 32 s + 8 (L >> 4) + e (e = 0..7) are stored at
 * ((t * ksteps + s) * 64 + L) * 8 + e, zero for k >= kdim or n >= ndim. A
 * wave's fragment of one tile and step is then 1 KB of whole cache lines.
 * M is the layer's weights W[in][out] (trans = 0: k = in, n = out) or
 * their transpose (trans = 1: k = out, n = in); ldw is the row length of W
 * as stored (out). */
__host__ __device__ inline uint64_t fragment_index(int k, int n, int ksteps) {
    return ((uint64_t)((n >> 4) * ksteps + (k >> 5)) * 64 + (n & 15) + 16 * ((k & 31) >> 3)) * 8 +
           (k & 7);
}

__global__ void fragment_pad_kernel(const float *w, int kdim, int ndim, int ldw, int trans,
                                    int ksteps, int ntiles, uint16_t *dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntiles * ksteps * 512) return;
    const int e = i & 7, L = (i >> 3) & 63, ts = i >> 9;
    const int t = ts / ksteps, st = ts % ksteps;
    const int n = t * 16 + (L & 15), k = st * 32 + (L >> 4) * 8 + e;
    const size_t src = trans ? (size_t)n * ldw + k : (size_t)k * ldw + n;
    dst[i] = (k < kdim && n < ndim) ? f2bf(w[src]) : (uint16_t)0;
}


__global__ void fill_f32_kernel(float *p, uint32_t n, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}


/* ------------------------------------------------------------------ */
/* Training step: products with both operands stored k-contiguous       */
/* ------------------------------------------------------------------ */
/* C[M][N] = A[M][K] B[K][N] with B given as its transpose BT[N][K], so both
 * operands are read 8 elements (16 bytes) per thread per load. Every product
 * of a training step has this form:
 *   forward  H_l = relu(H_{l-1} W_l + b_l)    A = H_{l-1},    BT = W_l^T
 *   data     dZ_{l-1} = (dZ_l W_l^T) . [H_{l-1} > 0]
 *                                             A = dZ_l,       BT = W_l
 *   weights  [dW_l; db_l] = [H_{l-1}^T; 1] dZ_l
 *                                             A = H_{l-1}^T,  BT = dZ_l^T
 * The weight gradient reads a row of ones below H_{l-1}^T: its product row
 * is the bias gradient, which the packed parameters hold right after W_l.
 * Forward and data products store their bf16 result row-major and
 * transposed, the two layouts the later products read. A 64x64 tile per
 * 256-thread block, K in steps of 32 through LDS, fragments and MFMA order
 * as gemm_kernel (the forward products equal it and the fused inference
 * kernel bit for bit). The weight gradients of all layers run in one launch,
 * each split over the batch into fp32 partial sums that the Adam kernel
 * adds in a fixed order. Leading dimensions are multiples of 8. */
enum NtEpi : int { NT_RELU = 0, NT_OUT = 1, NT_MASK = 2, NT_PARTIAL = 3, NT_LOSS = 4 };

struct NtProblem {
    const uint16_t *A;      /* [M][lda]                                   */
    const uint16_t *BT;     /* [N][ldb]                                   */
    const float *bias;      /* NT_RELU, NT_OUT                            */
    uint16_t *H;            /* NT_RELU, NT_MASK: [M][ldh]                 */
    uint16_t *HT;           /* NT_RELU, NT_MASK: [N][ldht]                */
    float *C;               /* NT_OUT: [M][ldc]; NT_PARTIAL: split 0 base */
    const uint16_t *mask;   /* NT_MASK: transposed activations [N][ldm]   */
    uint64_t part_stride;   /* NT_PARTIAL: floats between splits          */
    /* NT_LOSS: d = output - T[row]; H, HT receive bf16(2 d scale); the
     * squared error times scale is added to loss[0] or loss[*step % ring] */
    const float *T;
    const uint32_t *step;
    float *loss;
    uint64_t seed;
    float scale;
    uint32_t nt, ring;
    int sampled;            /* T rows of the step's minibatch, else row m  */
    int lda, ldb, M, N, K;
    int ones_row;           /* row of A read as 1.0, or -1                */
    int tiles_n, splits, block0;
    int ldh, ldht, ldc, ldm;
    int reserved;
};

/* Data row of minibatch row i at a training step. */
__device__ inline uint32_t sample_row(uint64_t seed, uint32_t step, uint32_t i, uint32_t n) {
    const uint64_t h = xo_mix64(seed + 0x9E3779B97F4A7C15ull * (((uint64_t)step << 32) + i + 1));
    return (uint32_t)(h % n);
}

/* Elements [k, k + 8) of a k-contiguous row, zero at and past kend. */
__device__ inline uint4 nt_load(const uint16_t *row, int k, int kend) {
    if (k + 8 <= kend) return *reinterpret_cast<const uint4 *>(row + k);
    uint16_t v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (k + e < kend) ? row[k + e] : (uint16_t)0;
    uint4 r;
    std::memcpy(&r, v, 16);
    return r;
}

/* The row of ones (bf16 1.0 = 0x3F80) over [k, k + 8), zero past kend. */
__device__ inline uint4 nt_ones(int k, int kend) {
    uint16_t v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (k + e < kend) ? (uint16_t)0x3F80 : (uint16_t)0;
    uint4 r;
    std::memcpy(&r, v, 16);
    return r;
}

/* K is staged through LDS NTK at a time: every thread issues NTK / 32
 * 16-byte loads per operand before the block synchronizes once. */
constexpr int NTK = 128, NTPAD = 8;
typedef uint16_t NtRow[NTK + NTPAD];

template <int EPI>
__device__ __forceinline__ void nt_tile(const NtProblem &g, int tile, NtRow *sA, NtRow *sB) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int split = tile % g.splits, t = tile / g.splits;
    const int m0 = (t / g.tiles_n) * TM, n0 = (t % g.tiles_n) * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int ksteps = (g.K + TK - 1) / TK;
    const int kc = (ksteps + g.splits - 1) / g.splits * TK;
    const int kbeg = split * kc;
    const int kend = kbeg + kc < g.K ? kbeg + kc : g.K;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    /* loader: row tid / 4 of each operand, k offsets 8 (tid % 4) + 32 s */
    const int lr = tid >> 2, lc = (tid & 3) * 8;
    const int am = m0 + lr, bn = n0 + lr;
    const bool ones = am == g.ones_row;
    const uint16_t *arow = g.A + (size_t)am * g.lda;
    const uint16_t *brow = g.BT + (size_t)bn * g.ldb;
    const bool live[2][2] = {{m0 + wm < g.M && n0 + wn < g.N, m0 + wm < g.M && n0 + wn + 16 < g.N},
                             {m0 + wm + 16 < g.M && n0 + wn < g.N,
                              m0 + wm + 16 < g.M && n0 + wn + 16 < g.N}};
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    for (int c0 = kbeg; c0 < kend; c0 += NTK) {
        uint4 va[NTK / TK], vb[NTK / TK];
#pragma unroll
        for (int s8 = 0; s8 < NTK / TK; ++s8) {
            const int k = c0 + lc + s8 * TK;
            va[s8] = uint4{0u, 0u, 0u, 0u};
            vb[s8] = uint4{0u, 0u, 0u, 0u};
            if (am < g.M) va[s8] = ones ? nt_ones(k, kend) : nt_load(arow, k, kend);
            if (bn < g.N) vb[s8] = nt_load(brow, k, kend);
        }
#pragma unroll
        for (int s8 = 0; s8 < NTK / TK; ++s8) {
            *reinterpret_cast<uint4 *>(&sA[lr][lc + s8 * TK]) = va[s8];
            *reinterpret_cast<uint4 *>(&sB[lr][lc + s8 * TK]) = vb[s8];
        }
        __syncthreads();
        const int nks = (kend - c0 + TK - 1) / TK < NTK / TK ? (kend - c0 + TK - 1) / TK : NTK / TK;
        for (int ks = 0; ks < nks; ++ks) {
            bf16x8 a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8 *>(&sA[wm + i * 16 + fr][ks * TK + fk]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const bf16x8 *>(&sB[wn + j * 16 + fr][ks * TK + fk]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (live[i][j])
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j],
                                                                           acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }

    /* lane l holds column (l % 16), rows 4 (l / 16) + r of each 16x16 */
    float sq = 0.f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int mb = m0 + wm + i * 16 + (lane >> 4) * 4;
            const int n = n0 + wn + j * 16 + (lane & 15);
            if (n >= g.N || mb >= g.M) continue;
            if (EPI == NT_PARTIAL) {
                float *c = g.C + (size_t)split * g.part_stride + (size_t)mb * g.ldc + n;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (mb + r < g.M) c[(size_t)r * g.ldc] = acc[i][j][r];
            } else if (EPI == NT_OUT) {
                const float bv = g.bias[n];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (mb + r < g.M) g.C[(size_t)(mb + r) * g.ldc + n] = acc[i][j][r] + bv;
            } else {
                uint16_t hv[4] = {0, 0, 0, 0};
                uint16_t mk[4] = {0, 0, 0, 0};
                if (EPI == NT_MASK) {
                    const uint16_t *mrow = g.mask + (size_t)n * g.ldm + mb;
                    if (mb + 4 <= g.M) {
                        const uint2 pk = *reinterpret_cast<const uint2 *>(mrow);
                        mk[0] = (uint16_t)pk.x;
                        mk[1] = (uint16_t)(pk.x >> 16);
                        mk[2] = (uint16_t)pk.y;
                        mk[3] = (uint16_t)(pk.y >> 16);
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r)
                            if (mb + r < g.M) mk[r] = mrow[r];
                    }
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (mb + r >= g.M) continue;
                    float v = acc[i][j][r];
                    if (EPI == NT_RELU) {
                        const float z = v + g.bias[n];
                        v = z > 0.f ? z : 0.f;
                    } else if (EPI == NT_MASK) {
                        v = bf2f(mk[r]) > 0.f ? v : 0.f;
                    } else {
                        const uint32_t row = g.sampled
                                                 ? sample_row(g.seed, *g.step, (uint32_t)(mb + r), g.nt)
                                                 : (uint32_t)(mb + r);
                        const float d = (v + g.bias[n]) - g.T[(size_t)row * g.N + n];
                        sq += d * d;
                        v = 2.f * d * g.scale;
                    }
                    hv[r] = f2bf(v);
                    g.H[(size_t)(mb + r) * g.ldh + n] = hv[r];
                }
                uint16_t *ht = g.HT + (size_t)n * g.ldht + mb;
                if (mb + 4 <= g.M) {
                    uint2 pk;
                    pk.x = (uint32_t)hv[0] | ((uint32_t)hv[1] << 16);
                    pk.y = (uint32_t)hv[2] | ((uint32_t)hv[3] << 16);
                    *reinterpret_cast<uint2 *>(ht) = pk;
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (mb + r < g.M) ht[r] = hv[r];
                }
            }
        }
    if (EPI == NT_LOSS) {
        /* one atomic per wave */
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
        if (lane == 0 && sq != 0.f)
            atomicAdd(g.ring ? g.loss + (*g.step % g.ring) : g.loss, sq * g.scale);
    }
}

template <int EPI>
__global__ void __launch_bounds__(256) gemm_nt_kernel(const NtProblem g) {
    __shared__ __attribute__((aligned(16))) NtRow sA[TM];
    __shared__ __attribute__((aligned(16))) NtRow sB[TN];
    nt_tile<EPI>(g, (int)blockIdx.x, sA, sB);
}

/* Weight gradients of every layer: problem p takes blocks [block0, next
 * block0). */
__global__ void __launch_bounds__(256) gemm_nt_partial_kernel(const NtProblem *table, int nprob) {
    __shared__ __attribute__((aligned(16))) NtRow sA[TM];
    __shared__ __attribute__((aligned(16))) NtRow sB[TN];
    int pi = 0;
    while (pi + 1 < nprob && (int)blockIdx.x >= table[pi + 1].block0) ++pi;
    const NtProblem g = table[pi];
    nt_tile<NT_PARTIAL>(g, (int)blockIdx.x - g.block0, sA, sB);
}

/* Input rows as bf16, row-major [B][ldx] and transposed [nin][ldxt]: the
 * minibatch of the step (sampled) or the first B rows. */
__global__ void rows_kernel(const float *X, uint32_t n, uint32_t nin, uint32_t B, int sampled,
                            uint64_t seed, const uint32_t *step, uint16_t *Xr, int ldx,
                            uint16_t *XT, int ldxt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const uint32_t j = sampled ? sample_row(seed, *step, i, n) : i;
    for (uint32_t c = 0; c < nin; ++c) {
        const uint16_t v = f2bf(X[(size_t)j * nin + c]);
        Xr[(size_t)i * ldx + c] = v;
        XT[(size_t)c * ldxt + i] = v;
    }
}

constexpr uint64_t kNotMatrix = ~0ull;

constexpr int kAdamBlocks = 256;

/* Adam (t = *step + 1) on the sum of the gradient's split partials, added
 * in split order; refreshes both bf16 copies of the weight matrices. The
 * last block to finish (counted in *done, which it returns to 0) advances
 * the step counter, after every block has read it. */
__global__ void adam_nt_kernel(float *w, float *m, float *v, const float *part, uint32_t splits,
                               uint64_t n, uint32_t *step, uint32_t *done, float lr, float b1,
                               float b2, float eps, const uint64_t *bf_index,
                               const uint64_t *bt_index, uint16_t *wbf, uint16_t *wtb) {
    const float t = (float)(*step + 1u);
    const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float gi = part[i];
        for (uint32_t s = 1; s < splits; ++s) gi += part[(size_t)s * n + i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (bf_index[i] != kNotMatrix) {
            const uint16_t h = f2bf(wi);
            wbf[bf_index[i]] = h;
            wtb[bt_index[i]] = h;
        }
    }
    /* every thread of the block read *step above; the next launch sees the
     * increment (kernel boundary), so no fence is needed */
    __syncthreads();
    if (threadIdx.x == 0 && atomicAdd(done, 1u) == gridDim.x - 1u) {
        *step += 1u;
        *done = 0u;
    }
}

/* grad[i] = sum of the split partials, in split order. */
__global__ void sum_partials_kernel(const float *part, uint32_t splits, uint64_t n, float *grad) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        float gi = part[i];
        for (uint32_t s = 1; s < splits; ++s) gi += part[(size_t)s * n + i];
        grad[i] = gi;
    }
}

/* Both bf16 copies of the weight matrices from the fp32 master weights. */
__global__ void refresh_kernel(const float *w, uint64_t n, const uint64_t *bf_index,
                               const uint64_t *bt_index, uint16_t *wbf, uint16_t *wtb) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (bf_index[i] == kNotMatrix) continue;
        const uint16_t h = f2bf(w[i]);
        wbf[bf_index[i]] = h;
        wtb[bt_index[i]] = h;
    }
}

/* ------------------------------------------------------------------ */
/* Fused training step: two launches                                    */
/* ------------------------------------------------------------------ */
/* For the networks of the fused inference kernel with at most TMAXL
 * layers. train_rows_kernel takes TBM minibatch rows per block through the
 * forward pass, the loss and the data gradients with every activation in
 * LDS, and writes what the weight gradients read: each layer's input H_l
 * and output gradient dZ_l, bf16, transposed (one row per unit, the batch
 * contiguous). The products are those of the layer-by-layer path in the
 * same k order (weights as fragment-major copies, W_l for the forward pass
 * and W_l^T for the data gradients), so activations, loss gradient and data
 * gradients are the same bits. train_weights_kernel forms each 32x32 tile
 * of every dW_l over the whole batch in one block: its WNW waves take a
 * 1/WNW share of the batch each, with the loads of the next WG k steps in
 * flight while the MFMAs of the current WG issue, and their sums are
 * added in wave order in LDS (deterministic; only the order of additions
 * differs from the split sums of the layer-by-layer path). Blocks of the
 * first row tile also form the bias gradient (a row of ones times the dZ_l
 * fragments). The tile's gradient then goes to Adam in place (or out, for
 * xo_mlp_gradient), which refreshes both fragment-major copies of the
 * updated weights. The loss: each row block writes its partial sum, the
 * first block of train_weights_kernel adds them in a fixed order (no
 * atomics: a step's loss is deterministic). */
constexpr int TMT = 1, TBM = 16 * TMT, TMAXL = 6;
/* waves per block of train_rows_kernel: wave w owns 16 / TNW of the 16
 * column tiles of a hidden-width product and one of the output layer's 8
 * (waves 8.. sit that product out). Sixteen waves: 0.0297 ms per 4096-row
 * training step against 0.0304 with eight. */
constexpr int TNW = 16, TNT = 64 * TNW;
constexpr int TXS = 40, TYS = 136; /* LDS row strides of the input and output tiles */
enum TrainEpi : int { TE_RELU = 0, TE_LOSS = 1, TE_MASK = 2 };

struct TrainArgs {
    const float *X;          /* [n][nin] rows                                  */
    const float *T;          /* [n][nout] targets                              */
    uint32_t n;
    int sampled;             /* the step's minibatch rows, else rows 0..B-1    */
    uint64_t seed;
    const uint32_t *step;    /* the step counter, read when sampled ...       */
    uint32_t *step_copy;     /* ... and copied here for the weight kernel     */
    float *loss_part;        /* per row block: its squared errors times scale  */
    float scale;             /* 1 / (B nout)                                   */
    int B, nin, nout, nlayers, ldb;
    const uint16_t *wf[TMAXL]; /* fragments of W_l                             */
    const uint16_t *wb[TMAXL]; /* fragments of W_l^T (l >= 1)                  */
    const float *bias[TMAXL];
    uint16_t *ht[TMAXL];       /* H_l^T, the input of layer l, and            */
    uint16_t *dzt[TMAXL];      /* dZ_l^T, the gradient at its output: fragments
                                  over the batch (store_transposed)            */
};

/* All of a wave's weight fragments of one product (KS k steps x NJ
 * 16-column tiles), loaded two products ahead of their use so their L2
 * latency is hidden behind the products in between (one product ahead:
 * 0.0297 ms per 4096-row training step, two: 0.0294). */
constexpr int TKS = 8, TNJ = 16 / TNW;
typedef bf16x8 TrainFrags[TKS][TNJ];

/* Column tile j of a wave: hidden-width products (16 tiles) take
 * consecutive tiles, the output layer (8 tiles) interleaves them. */
template <bool OUT, int NJ> __device__ __forceinline__ int train_tile(int wave, int j) {
    return OUT ? wave + TNW * j : wave * NJ + j;
}

template <int KS, int NJ, bool OUT>
__device__ __forceinline__ void train_wload(const uint16_t *W, int lane, int wave,
                                            TrainFrags &f) {
    if (OUT && wave >= 8) return; /* the output layer has 8 column tiles */
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
            const int tile = train_tile<OUT, NJ>(wave, j);
            f[s][j] = *reinterpret_cast<const bf16x8 *>(W + (((size_t)tile * KS + s) * 64 + lane) * 8);
        }
}

/* Targets of the loss product, per lane: rows 16 i + l % 16 of the block,
 * the four columns of each of its two column tiles (zero past nout and for
 * rows past B). Loaded before the next product's weights are requested, so
 * waiting for them does not wait for those. */
constexpr int TNJO = TNW >= 8 ? 1 : 8 / TNW; /* output-layer column tiles per wave */
typedef float4 TrainTargets[TMT][TNJO];

__device__ __forceinline__ void train_targets(const TrainArgs &a, int row0, uint32_t stp,
                                              int lane, int wave, TrainTargets &tv) {
    const int fr = lane & 15, nq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < TMT; ++i) {
        const int grow = row0 + i * 16 + fr;
        const uint32_t trow = grow < a.B ? (a.sampled ? sample_row(a.seed, stp, (uint32_t)grow, a.n)
                                                      : (uint32_t)grow)
                                         : 0u;
        const float *t = a.T + (size_t)trow * a.nout;
#pragma unroll
        for (int j = 0; j < TNJO; ++j) {
            const int n0 = train_tile<true, TNJO>(wave, j) * 16 + nq;
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = (grow < a.B && n0 + r < a.nout) ? t[n0 + r] : 0.f;
            tv[i][j] = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
}

/* One product of a TBM-row block: src (LDS, row stride ss) times the
 * fragments f, written as bf16 to dst (row stride ds): bias (LDS) + ReLU,
 * or the loss gradient 2 (y - t) scale against the targets tv (sq collects
 * the squared errors), or the product masked by mask > 0. Operand order
 * and result layout as fused_layer: lane l holds four consecutive columns
 * of row l % 16. */
template <int KS, int NJ, int EPI, bool OUT>
__device__ __forceinline__ void train_mma(const TrainArgs &a, const TrainFrags &f,
                                          const uint16_t *src, int ss, uint16_t *dst, int ds,
                                          const float *bias, const uint16_t *mask,
                                          const TrainTargets &tv, int row0, int lane, int wave,
                                          float &sq) {
    if (OUT && wave >= 8) return; /* the output layer has 8 column tiles */
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    floatx4 acc[TMT][NJ];
#pragma unroll
    for (int i = 0; i < TMT; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    bf16x8 x[2][TMT];
#pragma unroll
    for (int i = 0; i < TMT; ++i)
        x[0][i] = *reinterpret_cast<const bf16x8 *>(&src[(i * 16 + fr) * ss + fk]);
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s + 1 < KS) {
#pragma unroll
            for (int i = 0; i < TMT; ++i)
                x[(s + 1) & 1][i] = *reinterpret_cast<const bf16x8 *>(
                    &src[(i * 16 + fr) * ss + (s + 1) * 32 + fk]);
        }
#pragma unroll
        for (int i = 0; i < TMT; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(f[s][j], x[s & 1][i],
                                                                   acc[i][j], 0, 0, 0);
    }
    const int nq = (lane >> 4) * 4;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const int n0 = train_tile<OUT, NJ>(wave, j) * 16 + nq;
#pragma unroll
        for (int i = 0; i < TMT; ++i) {
            const int row = i * 16 + fr;
            float z[4] = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
            if (EPI == TE_RELU) {
                const float4 bv = *reinterpret_cast<const float4 *>(bias + n0);
                const float bb[4] = {bv.x, bv.y, bv.z, bv.w};
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float v = z[r] + bb[r];
                    z[r] = v > 0.f ? v : 0.f;
                }
            } else if (EPI == TE_LOSS) {
                const int grow = row0 + row;
                if (grow < a.B) {
                    const int jo = j < TNJO ? j : 0;
                    const float t[4] = {tv[i][jo].x, tv[i][jo].y, tv[i][jo].z, tv[i][jo].w};
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        if (n0 + r < a.nout) {
                            const float d = (z[r] + bias[n0 + r]) - t[r];
                            sq += d * d;
                            z[r] = 2.f * d * a.scale;
                        } else {
                            z[r] = 0.f;
                        }
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) z[r] = 0.f;
                }
            } else {
                const uint2 mk = *reinterpret_cast<const uint2 *>(&mask[row * FSTRIDE + n0]);
                const uint16_t mv[4] = {(uint16_t)mk.x, (uint16_t)(mk.x >> 16), (uint16_t)mk.y,
                                        (uint16_t)(mk.y >> 16)};
#pragma unroll
                for (int r = 0; r < 4; ++r) z[r] = bf2f(mv[r]) > 0.f ? z[r] : 0.f;
            }
            uint2 pk;
            pk.x = bf16_pair(z[0], z[1]);
            pk.y = bf16_pair(z[2], z[3]);
            *reinterpret_cast<uint2 *>(&dst[row * ds + n0]) = pk;
        }
    }
}

/* Columns [0, ncols) of a TBM-row LDS tile, transposed (unit c, batch
 * rows k) into dstT as MFMA fragments over the batch: unit c and batch rows
 * [8 q, 8 q + 8) are the 16 bytes at fragment_index(8 q, c, nks) (nks: k
 * steps of 32 batch rows per 16-unit tile), so the weight-gradient kernel
 * reads each 16 x 32 fragment as 1 KB of whole cache lines. Each store of
 * a wave writes one such fragment whole (lane l: slot l), which measured
 * 1.6 us per training step faster than four 16-byte slots per lane from
 * 8-byte LDS reads. */
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__device__ inline void store_transposed(const uint16_t *tile, int stride, int ncols,
                                        uint16_t *dstT, int nks, int row0) {
    /* lane l of a wave: slot l of one 16-unit tile's fragment, unit
     * 16 T + (l & 15), rows 8 (l >> 4) .. + 8 of this k step */
    constexpr int G = TBM / 8;
    const int ntiles = (ncols + 15) >> 4;
    const int nslots = ntiles * 16 * G;
    for (int it = threadIdx.x; it < nslots; it += TNT) {
        const int slot = it % (16 * G), T = it / (16 * G);
        const int c = T * 16 + (slot & 15), g8 = (slot >> 4) * 8;
        uint32_t h[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) h[e] = c < ncols ? tile[(g8 + e) * stride + c] : 0u;
        const uint4 pk = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                                    h[6] | (h[7] << 16));
        /* nontemporal: written once and read by the next kernel from
         * whichever XCD takes the tile, no reuse in this XCD's L2 (2 %
         * faster per training step) */
        const u32x4 pv = {pk.x, pk.y, pk.z, pk.w};
        __builtin_nontemporal_store(pv, reinterpret_cast<u32x4 *>(dstT + fragment_index(row0 + g8, c, nks)));
    }
}

/* The LDS tiles of a row block: the input rows, each hidden layer's
 * output, the loss gradient and every layer's bias (FHID floats each). The
 * data gradient at a hidden layer's output overwrites that layer's
 * activations, which are its mask: each lane reads the mask elements it
 * then writes, and nothing else reads the tile during the product. */
struct TrainTiles {
    uint16_t *x, *h[TMAXL - 1], *y;
    float *bias;
};

/* Product I of a network of L layers: the forward products of layers
 * 0..L-1, then the data gradients through layers L-1..1. */
template <int L, int I> struct TrainOp {
    static constexpr bool fwd = I < L;
    static constexpr int layer = fwd ? I : 2 * L - 1 - I;
    static constexpr int KS = fwd ? (layer == 0 ? 1 : 8) : (layer == L - 1 ? 4 : 8);
    static constexpr bool OUT = fwd && layer == L - 1;
    static constexpr int NJ = OUT ? TNJO : TNJ;
    static constexpr int EPI = fwd ? (layer == L - 1 ? TE_LOSS : TE_RELU) : TE_MASK;
};

template <int L, int I>
__device__ __forceinline__ void train_ops(const TrainArgs &a, TrainFrags (&f)[3],
                                          const TrainTiles &t, const TrainTargets &tv, int row0,
                                          int lane, int wave, float *wsq) {
    if constexpr (I < 2 * L - 1) {
        using O = TrainOp<L, I>;
        if constexpr (I + 2 < 2 * L - 1) {
            using Nx = TrainOp<L, I + 2>;
            train_wload<Nx::KS, Nx::NJ, Nx::OUT>(Nx::fwd ? a.wf[Nx::layer] : a.wb[Nx::layer], lane,
                                                 wave, f[(I + 2) % 3]);
        }
        const int l = O::layer;
        const uint16_t *src;
        uint16_t *dst, *outT;
        const uint16_t *mask = nullptr;
        int ss, ds = FSTRIDE, ncols = FHID;
        if constexpr (O::fwd) {
            if constexpr (O::layer == 0) {
                src = t.x;
                ss = TXS;
            } else {
                src = t.h[O::layer - 1];
                ss = FSTRIDE;
            }
            if constexpr (O::layer + 1 < L) {
                dst = t.h[l];
                outT = a.ht[l + 1];
            } else {
                dst = t.y;
                ds = TYS;
                outT = a.dzt[L - 1];
                ncols = a.nout;
            }
        } else {
            if constexpr (O::layer == L - 1) {
                src = t.y;
                ss = TYS;
            } else {
                src = t.h[O::layer];
                ss = FSTRIDE;
            }
            dst = t.h[O::layer - 1];
            mask = t.h[O::layer - 1];
            outT = a.dzt[O::layer - 1];
        }
        float sq = 0.f;
        train_mma<O::KS, O::NJ, O::EPI, O::OUT>(a, f[I % 3], src, ss, dst, ds, t.bias + l * FHID, mask,
                                        tv, row0, lane, wave, sq);
        if constexpr (O::EPI == TE_LOSS) {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off);
            if (lane == 0) wsq[wave] = sq;
        }
        __syncthreads();
        if constexpr (O::EPI == TE_LOSS) {
            /* the block's squared errors, waves added in order */
            if (threadIdx.x == 0) {
                float b = 0.f;
                for (int w = 0; w < TNW; ++w) b += wsq[w];
                a.loss_part[blockIdx.x] = b * a.scale;
            }
        }
        store_transposed(dst, ds, ncols, outT, a.ldb / 32, row0);
        train_ops<L, I + 1>(a, f, t, tv, row0, lane, wave, wsq);
    }
}

template <int L>
__global__ void __launch_bounds__(TNT) train_rows_kernel(const TrainArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t sx[TBM * TXS];
    __shared__ __attribute__((aligned(16))) uint16_t sh[L - 1][TBM * FSTRIDE];
    __shared__ __attribute__((aligned(16))) uint16_t sy[TBM * TYS];
    __shared__ __attribute__((aligned(16))) float sb[L * FHID];
    __shared__ float wsq[TNW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row0 = blockIdx.x * TBM;
    TrainFrags f[3];
    train_wload<TrainOp<L, 0>::KS, TrainOp<L, 0>::NJ, TrainOp<L, 0>::OUT>(a.wf[0], lane, wave,
                                                                          f[0]);
    train_wload<TrainOp<L, 1>::KS, TrainOp<L, 1>::NJ, TrainOp<L, 1>::OUT>(a.wf[1], lane, wave,
                                                                          f[1]);
    const uint32_t stp = a.sampled ? *a.step : 0u;
    if (a.sampled && blockIdx.x == 0 && tid == 0) *a.step_copy = stp;
    for (int e = tid; e < TBM * 32; e += TNT) {
        const int m = e >> 5, c = e & 31, grow = row0 + m;
        float v = 0.f;
        if (grow < a.B && c < a.nin) {
            const uint32_t j =
                a.sampled ? sample_row(a.seed, stp, (uint32_t)grow, a.n) : (uint32_t)grow;
            v = a.X[(size_t)j * a.nin + c];
        }
        sx[m * TXS + c] = f2bf(v);
    }
    /* the loss product's targets, requested now: their latency hides
     * behind the forward products */
    TrainTargets tv;
    train_targets(a, row0, stp, lane, wave, tv);
    for (int e = tid; e < L * FHID; e += TNT) {
        const int l = e / FHID, c = e % FHID;
        sb[e] = (l + 1 < L || c < a.nout) ? a.bias[l][c] : 0.f;
    }
    __syncthreads();
    store_transposed(sx, TXS, a.nin, a.ht[0], a.ldb / 32, row0);
    TrainTiles t;
    t.x = sx;
    for (int l = 0; l < L - 1; ++l) t.h[l] = sh[l];
    for (int l = L - 1; l < TMAXL - 1; ++l) t.h[l] = nullptr;
    t.y = sy;
    t.bias = sb;
    train_ops<L, 0>(a, f, t, tv, row0, lane, wave, wsq);
}

typedef void (*TrainRowsFn)(TrainArgs);
/* train_rows_kernel<L> for L = 2..TMAXL */
const TrainRowsFn kTrainRows[TMAXL + 1] = {nullptr,
                                           nullptr,
                                           train_rows_kernel<2>,
                                           train_rows_kernel<3>,
                                           train_rows_kernel<4>,
                                           train_rows_kernel<5>,
                                           train_rows_kernel<6>};

struct WgLayer {
    const uint16_t *A;       /* H_l^T, fragments over the batch  */
    const uint16_t *BT;      /* dZ_l^T, fragments over the batch */
    uint16_t *wf, *wb;       /* fragment copies of W_l and W_l^T (wb null for l = 0) */
    uint64_t woff, boff;
    int in, out, tiles_n, block0, ksf, ksb;
};

struct WgArgs {
    WgLayer ly[TMAXL];
    int nlayers, nks, ldb, adam;
    float *grad;             /* adam == 0: the gradient, packed like w */
    float *w, *m, *v;
    /* adam: the step is read from the row kernel's copy and its successor
     * written to the counter itself, which no block of this kernel reads:
     * no counter of finished blocks, and no atomics */
    const uint32_t *step;
    uint32_t *step_next;
    float lr, b1, b2, eps;
    const uint16_t *zrow;    /* 16 bytes of zeros, read for tiles past a layer's width */
    const float *loss_part;  /* the row blocks' partial losses ...           */
    int nparts;
    float *loss;             /* ... added in order into loss[0], or loss[step % ring] */
    uint32_t ring;
};

/* 32x32 tiles (about 170 blocks for the surrogate's layers); each wave
 * keeps two groups of WG k steps in flight */
constexpr int WTI = 2, WTJ = 2, WG = 2;
/* waves per block, each a 1/WNW share of the batch: four per SIMD, so
 * more fragment loads in flight per CU (8 waves with 5-step groups: 17.8
 * us per launch, 16 waves with 2-step groups: 15.4 us) */
constexpr int WNW = 16, WNT = 64 * WNW;
constexpr int WTM = 16 * WTI, WTN = 16 * WTJ, WEPT = WTM * WTN / WNT;

/* The k loop of a weight-gradient tile: steps [ks0, ks1) of 32 batch rows,
 * two groups of WG steps in flight. With BIAS, a row of ones times the
 * dZ_l fragments also accumulates the bias gradient (every row of accb
 * holds the column sums). */
template <bool BIAS>
__device__ __forceinline__ void wg_loop(const uint16_t *const (&ap)[WTI], const int (&ainc)[WTI],
                                        const uint16_t *const (&bp)[WTJ], const int (&binc)[WTJ],
                                        int ks0, int ks1,
                                        floatx4 (&acc)[WTI][WTJ], floatx4 (&accb)[WTJ]) {
    bf16x8 fa[2][WG][WTI], fb[2][WG][WTJ];
    bf16x8 ones;
#pragma unroll
    for (int e = 0; e < 8; ++e) ones[e] = (__bf16)1.0f;
    /* steps past ks1 read step ks1 - 1 again (in bounds) and contribute
     * nothing: their dZ fragments are replaced by zeros */
    auto load = [&](int buf, int s0) {
#pragma unroll
        for (int u = 0; u < WG; ++u) {
            const int s = s0 + u;
            const int sc = s < ks1 ? s : ks1 - 1;
#pragma unroll
            for (int i = 0; i < WTI; ++i)
                fa[buf][u][i] = *reinterpret_cast<const bf16x8 *>(ap[i] + (size_t)sc * ainc[i]);
#pragma unroll
            for (int j = 0; j < WTJ; ++j) {
                const bf16x8 v = *reinterpret_cast<const bf16x8 *>(bp[j] + (size_t)sc * binc[j]);
                fb[buf][u][j] = s < ks1 ? v : bf16x8{};
            }
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int u = 0; u < WG; ++u) {
#pragma unroll
            for (int i = 0; i < WTI; ++i)
#pragma unroll
                for (int j = 0; j < WTJ; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        fa[buf][u][i], fb[buf][u][j], acc[i][j], 0, 0, 0);
            if (BIAS) {
#pragma unroll
                for (int j = 0; j < WTJ; ++j)
                    accb[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ones, fb[buf][u][j],
                                                                     accb[j], 0, 0, 0);
            }
        }
    };
    /* the scheduling barriers keep each group's loads ahead of the MFMAs
     * of the group before it */
    load(0, ks0);
    for (int s0 = ks0; s0 < ks1; s0 += 2 * WG) {
        load(1, s0 + WG);
        __builtin_amdgcn_sched_barrier(0);
        compute(0);
        __builtin_amdgcn_sched_barrier(0);
        load(0, s0 + 2 * WG);
        __builtin_amdgcn_sched_barrier(0);
        compute(1);
        __builtin_amdgcn_sched_barrier(0);
    }
}

/* One WTM x WTN tile of dW_l (rows: inputs, columns: outputs) per block. */
__global__ void __launch_bounds__(WNT) train_weights_kernel(const WgArgs g) {
    __shared__ __attribute__((aligned(16))) float red[WNW][16 * WTI * 16 * WTJ];
    __shared__ float bred[WNW][16 * WTJ];
    /* the tile's updated weights as bf16, for the fragment stores */
    __shared__ __attribute__((aligned(16))) uint16_t sw[WTM][WTN + 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int li = 0;
    while (li + 1 < g.nlayers && (int)blockIdx.x >= g.ly[li + 1].block0) ++li;
    const WgLayer &Ly = g.ly[li];
    const int t = (int)blockIdx.x - Ly.block0;
    const int m0 = (t / Ly.tiles_n) * WTM, n0 = (t % Ly.tiles_n) * WTN;
    const bool bias_tile = m0 == 0;
    const uint32_t stp = g.adam ? *g.step : 0u;
    const int fr = lane & 15;
    /* this thread's elements of the tile (row tid / (WTN / WEPT), WEPT
     * columns) and, for the first row of tiles, one bias: their optimizer
     * state and weights are requested now, so the epilogue does not wait
     * for them */
    const int r = tid / (WTN / WEPT), cb = (tid % (WTN / WEPT)) * WEPT, in_r = m0 + r;
    const uint64_t base = Ly.woff + (uint64_t)in_r * Ly.out + n0 + cb;
    const uint64_t bidx = Ly.boff + n0 + tid;
    const bool has_bias = bias_tile && tid < WTN && n0 + tid < Ly.out;
    float mo[WEPT], vo[WEPT], wo[WEPT], mb = 0.f, vb = 0.f, wb = 0.f;
#pragma unroll
    for (int cc = 0; cc < WEPT; ++cc) {
        mo[cc] = vo[cc] = wo[cc] = 0.f;
        if (g.adam && in_r < Ly.in && n0 + cb + cc < Ly.out) {
            mo[cc] = g.m[base + cc];
            vo[cc] = g.v[base + cc];
            wo[cc] = g.w[base + cc];
        }
    }
    if (g.adam && has_bias) {
        mb = g.m[bidx];
        vb = g.v[bidx];
        wb = g.w[bidx];
    }
    if (blockIdx.x == 0 && wave == 0) {
        /* the step's loss: lane l adds parts l, l + 64, ..., then a fixed
         * butterfly over the lanes */
        float v = 0.f;
        for (int i = lane; i < g.nparts; i += 64) v += g.loss_part[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) g.loss[g.ring ? stp % g.ring : 0u] = v;
    }
    /* fragment (16 units x 32 batch rows) of unit tile T at k step s:
     * 1 KB at ((T nks + s) 64 + lane) 8 */
    const int nksb = g.ldb / 32;
    const uint16_t *ap[WTI], *bp[WTJ];
    int ainc[WTI], binc[WTJ];
#pragma unroll
    for (int i = 0; i < WTI; ++i) {
        const int T = (m0 >> 4) + i;
        const bool ok = 16 * T < Ly.in;
        ap[i] = ok ? Ly.A + ((size_t)T * nksb * 64 + lane) * 8 : g.zrow;
        ainc[i] = ok ? 512 : 0;
    }
#pragma unroll
    for (int j = 0; j < WTJ; ++j) {
        const int T = (n0 >> 4) + j;
        const bool ok = 16 * T < Ly.out;
        bp[j] = ok ? Ly.BT + ((size_t)T * nksb * 64 + lane) * 8 : g.zrow;
        binc[j] = ok ? 512 : 0;
    }
    /* the wave's k steps (uniform, so the loads below are unconditional:
     * steps past the range read zeros) */
    const int wv = __builtin_amdgcn_readfirstlane(wave);
    const int ks0 = wv * g.nks / WNW, ks1 = (wv + 1) * g.nks / WNW;
    floatx4 acc[WTI][WTJ], accb[WTJ];
#pragma unroll
    for (int i = 0; i < WTI; ++i)
#pragma unroll
        for (int j = 0; j < WTJ; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < WTJ; ++j) accb[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (bias_tile)
        wg_loop<true>(ap, ainc, bp, binc, ks0, ks1, acc, accb);
    else
        wg_loop<false>(ap, ainc, bp, binc, ks0, ks1, acc, accb);
    float bs[WTJ];
#pragma unroll
    for (int j = 0; j < WTJ; ++j) bs[j] = accb[j][0];
    /* lane l holds rows 16 i + 4 (l / 16) + q, column 16 j + l % 16 */
#pragma unroll
    for (int i = 0; i < WTI; ++i)
#pragma unroll
        for (int j = 0; j < WTJ; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                red[wave][(16 * i + (lane >> 4) * 4 + q) * WTN + 16 * j + fr] = acc[i][j][q];
    if (bias_tile && lane < 16) {
#pragma unroll
        for (int j = 0; j < WTJ; ++j) bred[wave][16 * j + lane] = bs[j];
    }
    __syncthreads();
    const float tt = (float)(stp + 1u);
    const float c1 = 1.f - powf(g.b1, tt), c2 = 1.f - powf(g.b2, tt);
    auto adam = [&](float gi, float mo, float vo, float wo, float &mi, float &vi) {
        mi = g.b1 * mo + (1.f - g.b1) * gi;
        vi = g.b2 * vo + (1.f - g.b2) * gi * gi;
        return wo - g.lr * (mi / c1) / (sqrtf(vi / c2) + g.eps);
    };
    uint16_t hw[WEPT];
#pragma unroll
    for (int cc = 0; cc < WEPT; ++cc) hw[cc] = 0u; /* padding: stays zero */
    if (in_r < Ly.in) {
        float gi[WEPT];
#pragma unroll
        for (int cc = 0; cc < WEPT; ++cc) {
            const int e = r * WTN + cb + cc;
            gi[cc] = red[0][e];
#pragma unroll
            for (int w = 1; w < WNW; ++w) gi[cc] += red[w][e];
        }
#pragma unroll
        for (int cc = 0; cc < WEPT; ++cc) {
            const int out_c = n0 + cb + cc;
            if (out_c >= Ly.out) continue;
            if (!g.adam) {
                g.grad[base + cc] = gi[cc];
                continue;
            }
            float mi, vi;
            const float wi = adam(gi[cc], mo[cc], vo[cc], wo[cc], mi, vi);
            g.m[base + cc] = mi;
            g.v[base + cc] = vi;
            g.w[base + cc] = wi;
            hw[cc] = f2bf(wi);
        }
    }
    if (g.adam) {
#pragma unroll
        for (int cc = 0; cc < WEPT; ++cc) sw[r][cb + cc] = hw[cc];
    }
    if (has_bias) {
        float gi = bred[0][tid];
#pragma unroll
        for (int w = 1; w < WNW; ++w) gi += bred[w][tid];
        if (!g.adam) {
            g.grad[bidx] = gi;
        } else {
            float mi, vi;
            const float wi = adam(gi, mb, vb, wb, mi, vi);
            g.m[bidx] = mi;
            g.v[bidx] = vi;
            g.w[bidx] = wi;
        }
    }
    if (g.adam) {
        __syncthreads();
        /* both fragment copies of the tile, one whole 16-byte slot per
         * thread: W_l (k = input, the tile's rows; n = output) and W_l^T
         * (k = output, n = input); the tile covers one k step and two
         * 16-column tiles of each */
        const int slot = tid & 63, jt = (tid >> 6) & 1;
        const int nl = 16 * jt + (slot & 15), kl = 8 * (slot >> 4);
        if (tid < 128) {
            uint32_t h[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) h[e] = sw[kl + e][nl];
            const uint4 pk = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16),
                                        h[4] | (h[5] << 16), h[6] | (h[7] << 16));
            const size_t frag = (size_t)((n0 >> 4) + jt) * Ly.ksf + (m0 >> 5);
            *reinterpret_cast<uint4 *>(Ly.wf + (frag * 64 + slot) * 8) = pk;
        } else if (tid < 256 && Ly.wb) {
            const uint4 pk = *reinterpret_cast<const uint4 *>(&sw[nl][kl]);
            const size_t frag = (size_t)((m0 >> 4) + jt) * Ly.ksb + (n0 >> 5);
            *reinterpret_cast<uint4 *>(Ly.wb + (frag * 64 + slot) * 8) = pk;
        }
        if (blockIdx.x == 0 && tid == 0) *g.step_next = stp + 1u;
    }
}

} // namespace

struct xo_mlp {
    xo_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    int nlayers = 0;
    std::vector<int> dims;       /* nlayers + 1 widths */
    std::vector<int> ld;         /* widths rounded up to multiples of 8 */
    int ldb = 0;                 /* max_batch rounded up to a multiple of 8 */
    std::vector<uint64_t> woff;  /* offset of W_l in the packed parameters */
    std::vector<uint64_t> boff;  /* offset of b_l */
    std::vector<uint64_t> wbf_off; /* offset of W_l [in][ld out] in wbf */
    std::vector<uint64_t> wtb_off; /* offset of W_l^T [out][ld in] in wtb */
    uint64_t nparams = 0, nwbf = 0, nwtb = 0;
    uint32_t max_batch = 0;
    uint32_t max_splits = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *wbf = nullptr, *wtb = nullptr;
    uint64_t *bf_index = nullptr, *bt_index = nullptr;
    float *xf = nullptr;              /* [B][in] host-provided rows */
    float *tb = nullptr;              /* [B][out] host-provided targets */
    uint16_t *xr = nullptr;           /* [B][ld in] bf16 inputs */
    uint16_t *xt = nullptr;           /* [in][ldb] */
    std::vector<uint16_t *> hr, ht;   /* hidden activations [B][ld], [w][ldb] */
    std::vector<uint16_t *> dz, dzt;  /* hidden gradients, same layouts */
    uint16_t *dy = nullptr, *dyt = nullptr; /* output gradient */
    float *y = nullptr;               /* [B][out] */
    float *part = nullptr;            /* [max_splits][nparams] partial sums */
    std::map<uint32_t, NtProblem *> dw_tables; /* weight-gradient launches per batch */
    uint16_t *wt = nullptr;           /* fragment-major bf16 W_l (inference, training) */
    std::vector<uint64_t> wt_off;
    uint16_t *wtt = nullptr;          /* fragment-major bf16 W_l^T, l >= 1 (training) */
    std::vector<uint64_t> wtt_off;
    uint16_t *zrow = nullptr;         /* zeros read for rows past a layer's width */
    float *loss_part = nullptr;       /* fused step: one partial loss per row block */
    bool fusable = false, train_fused = false, wt_dirty = true;
    float *loss = nullptr;            /* per training step */
    uint32_t loss_cap = 0;
    uint32_t *step = nullptr;         /* [0] the step counter, [1] the fused step's copy */
    uint32_t *done = nullptr;         /* Adam blocks finished this step */
    float *data_x = nullptr, *data_t = nullptr;
    uint32_t ndata = 0;
    hipGraph_t graph = nullptr;       /* graph_rem (< kGraphSteps) consecutive steps */
    hipGraphExec_t exec = nullptr;
    uint32_t graph_rem = 0;
    hipGraph_t graph_k = nullptr;     /* kGraphSteps consecutive steps */
    hipGraphExec_t exec_k = nullptr;
    uint32_t graph_batch = 0;         /* the step arguments both graphs hold */
    uint64_t graph_seed = 0;
    float graph_lr = 0.f;
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

int pad8(int x) { return (x + 7) & ~7; }

/* Splits of the batch in the weight-gradient products. */
uint32_t batch_splits(const xo_mlp *N, uint32_t B) {
    uint32_t s = B / 512u;
    if (s < 1u) s = 1u;
    return s < N->max_splits ? s : N->max_splits;
}

template <int EPI> int nt_launch(xo_mlp *N, NtProblem g) {
    g.tiles_n = (g.N + TN - 1) / TN;
    g.splits = 1;
    g.block0 = 0;
    const int blocks = ((g.M + TM - 1) / TM) * g.tiles_n;
    hipLaunchKernelGGL(gemm_nt_kernel<EPI>, dim3(blocks), dim3(256), 0, N->stream, g);
    MLP_CHECK(hipGetLastError());
    return 0;
}

/* The loss the last layer of a training forward pass is scored against:
 * target rows T (those of the step's minibatch when sampled, else the
 * first B of nt), the loss written to loss[0] (ring 0) or to the step's
 * ring slot. */
struct LossSpec {
    const float *T;
    uint32_t nt;
    int sampled;
    uint64_t seed;
    float *loss;
    uint32_t ring;
};

/* Forward pass of B rows already in xr/xt: hidden activations in hr/ht;
 * the output fp32 in y, or with a loss the output gradient in dy/dyt and
 * the loss. */
int forward(xo_mlp *N, uint32_t B, const LossSpec *loss = nullptr) {
    const int L = N->nlayers;
    for (int l = 0; l < L; ++l) {
        NtProblem g;
        std::memset(&g, 0, sizeof(g));
        g.A = l == 0 ? N->xr : N->hr[l - 1];
        g.lda = N->ld[l];
        g.BT = N->wtb + N->wtb_off[l];
        g.ldb = N->ld[l];
        g.M = (int)B;
        g.N = N->dims[l + 1];
        g.K = N->dims[l];
        g.ones_row = -1;
        g.bias = N->w + N->boff[l];
        if (l + 1 < L) {
            g.H = N->hr[l];
            g.ldh = N->ld[l + 1];
            g.HT = N->ht[l];
            g.ldht = N->ldb;
            if (nt_launch<NT_RELU>(N, g)) return -1;
        } else if (loss) {
            g.H = N->dy;
            g.ldh = N->ld[L];
            g.HT = N->dyt;
            g.ldht = N->ldb;
            g.T = loss->T;
            g.nt = loss->nt;
            g.sampled = loss->sampled;
            g.seed = loss->seed;
            g.step = N->step;
            g.loss = loss->loss;
            g.ring = loss->ring;
            g.scale = 1.0f / (float)(B * (uint32_t)N->dims[L]);
            if (nt_launch<NT_LOSS>(N, g)) return -1;
        } else {
            g.C = N->y;
            g.ldc = N->dims[L];
            if (nt_launch<NT_OUT>(N, g)) return -1;
        }
    }
    return 0;
}

/* The weight-gradient problems of a batch size, uploaded once. */
int dw_table(xo_mlp *N, uint32_t B, const NtProblem **table, int *blocks) {
    const int L = N->nlayers;
    std::vector<NtProblem> t(L);
    const uint32_t S = batch_splits(N, B);
    int total = 0;
    for (int l = 0; l < L; ++l) {
        NtProblem &g = t[l];
        std::memset(&g, 0, sizeof(g));
        g.A = l == 0 ? N->xt : N->ht[l - 1];
        g.lda = N->ldb;
        g.BT = l + 1 == L ? N->dyt : N->dzt[l];
        g.ldb = N->ldb;
        g.M = N->dims[l] + 1;
        g.N = N->dims[l + 1];
        g.K = (int)B;
        g.ones_row = N->dims[l];
        g.tiles_n = (g.N + TN - 1) / TN;
        g.splits = (int)S;
        g.block0 = total;
        g.C = N->part + N->woff[l];
        g.ldc = N->dims[l + 1];
        g.part_stride = N->nparams;
        total += ((g.M + TM - 1) / TM) * g.tiles_n * g.splits;
    }
    *blocks = total;
    auto it = N->dw_tables.find(B);
    if (it == N->dw_tables.end()) {
        NtProblem *d = nullptr;
        MLP_CHECK(hipMalloc(&d, sizeof(NtProblem) * L));
        MLP_CHECK(hipMemcpy(d, t.data(), sizeof(NtProblem) * L, hipMemcpyHostToDevice));
        it = N->dw_tables.emplace(B, d).first;
    }
    *table = it->second;
    return 0;
}

/* Gradient of the loss of the last forward pass (run with a LossSpec,
 * which left the output gradient in dy/dyt), as split partial sums in
 * part. */
int backward(xo_mlp *N, uint32_t B) {
    const int L = N->nlayers;
    const uint16_t *dcur = N->dy;
    for (int l = L - 1; l >= 1; --l) {
        /* dZ_{l-1} = (dZ_l W_l^T) * (H_{l-1} > 0) */
        NtProblem g;
        std::memset(&g, 0, sizeof(g));
        g.A = dcur;
        g.lda = N->ld[l + 1];
        g.BT = N->wbf + N->wbf_off[l];
        g.ldb = N->ld[l + 1];
        g.M = (int)B;
        g.N = N->dims[l];
        g.K = N->dims[l + 1];
        g.ones_row = -1;
        g.H = N->dz[l - 1];
        g.ldh = N->ld[l];
        g.HT = N->dzt[l - 1];
        g.ldht = N->ldb;
        g.mask = N->ht[l - 1];
        g.ldm = N->ldb;
        if (nt_launch<NT_MASK>(N, g)) return -1;
        dcur = N->dz[l - 1];
    }
    const NtProblem *table = nullptr;
    int blocks = 0;
    if (dw_table(N, B, &table, &blocks)) return -1;
    hipLaunchKernelGGL(gemm_nt_partial_kernel, dim3(blocks), dim3(256), 0, N->stream, table,
                       L);
    MLP_CHECK(hipGetLastError());
    return 0;
}

int refresh_bf16(xo_mlp *N) {
    hipLaunchKernelGGL(refresh_kernel, dim3(256), dim3(256), 0, N->stream, N->w, N->nparams,
                       N->bf_index, N->bt_index, N->wbf, N->wtb);
    MLP_CHECK(hipGetLastError());
    return 0;
}

int train_fused_step(xo_mlp *N, uint32_t B, const float *X, const LossSpec &loss, bool adam,
                     float lr);

/* One training step on a device-sampled minibatch (capturable: every
 * argument is fixed, the step counter lives on the device). */
int train_step(xo_mlp *N, uint32_t B, uint64_t seed, float lr) {
    if (N->train_fused) {
        const LossSpec loss{N->data_t, N->ndata, 1, seed, N->loss, N->loss_cap};
        return train_fused_step(N, B, N->data_x, loss, true, lr);
    }
    hipLaunchKernelGGL(rows_kernel, dim3((B + 255) / 256), dim3(256), 0, N->stream, N->data_x,
                       N->ndata, (uint32_t)N->dims[0], B, 1, seed, N->step, N->xr, N->ld[0],
                       N->xt, N->ldb);
    MLP_CHECK(hipGetLastError());
    const LossSpec loss{N->data_t, N->ndata, 1, seed, N->loss, N->loss_cap};
    if (forward(N, B, &loss)) return -1;
    if (backward(N, B)) return -1;
    hipLaunchKernelGGL(adam_nt_kernel, dim3(kAdamBlocks), dim3(256), 0, N->stream, N->w, N->m,
                       N->v, N->part, batch_splits(N, B), N->nparams, N->step, N->done, lr, 0.9f,
                       0.999f, 1e-8f, N->bf_index, N->bt_index, N->wbf, N->wtb);
    MLP_CHECK(hipGetLastError());
    return 0;
}

/* The fragment-major copies from the fp32 master weights: W_l for the
 * fused kernels' forward products and, for fused training, W_l^T for the
 * data gradients. */
int refresh_transposed(xo_mlp *N) {
    const int L = N->nlayers;
    for (int l = 0; l < L; ++l) {
        const int in = N->dims[l], out = N->dims[l + 1];
        const int ksteps = l == 0 ? 1 : FHID / 32, ntiles = l + 1 == L ? 8 : FHID / 16;
        hipLaunchKernelGGL(fragment_pad_kernel, dim3(ntiles * ksteps * 2), dim3(256), 0,
                           N->stream, N->w + N->woff[l], in, out, out, 0, ksteps, ntiles,
                           N->wt + N->wt_off[l]);
        MLP_CHECK(hipGetLastError());
        if (N->train_fused && l >= 1) {
            const int kb = l + 1 == L ? 4 : FHID / 32, nb = FHID / 16;
            hipLaunchKernelGGL(fragment_pad_kernel, dim3(nb * kb * 2), dim3(256), 0, N->stream,
                               N->w + N->woff[l], out, in, out, 1, kb, nb,
                               N->wtt + N->wtt_off[l]);
            MLP_CHECK(hipGetLastError());
        }
    }
    N->wt_dirty = false;
    return 0;
}

/* One step of the fused training kernels over B rows of X (the training
 * set's minibatch when loss.sampled, else its first B rows): Adam with
 * learning rate lr (adam), or the gradient into N->grad. Capturable. */
int train_fused_step(xo_mlp *N, uint32_t B, const float *X, const LossSpec &loss, bool adam,
                     float lr) {
    const int L = N->nlayers;
    TrainArgs a;
    std::memset(&a, 0, sizeof(a));
    a.X = X;
    a.T = loss.T;
    a.n = loss.nt;
    a.sampled = loss.sampled;
    a.seed = loss.seed;
    a.step = N->step;
    a.step_copy = N->step + 1;
    a.loss_part = N->loss_part;
    a.scale = 1.0f / (float)(B * (uint32_t)N->dims[L]);
    a.B = (int)B;
    a.nin = N->dims[0];
    a.nout = N->dims[L];
    a.nlayers = L;
    a.ldb = N->ldb;
    WgArgs g;
    std::memset(&g, 0, sizeof(g));
    int blocks = 0;
    for (int l = 0; l < L; ++l) {
        a.wf[l] = N->wt + N->wt_off[l];
        a.wb[l] = l ? N->wtt + N->wtt_off[l] : nullptr;
        a.bias[l] = N->w + N->boff[l];
        a.ht[l] = l == 0 ? N->xt : N->ht[l - 1];
        a.dzt[l] = l + 1 == L ? N->dyt : N->dzt[l];
        WgLayer &y = g.ly[l];
        y.A = a.ht[l];
        y.BT = a.dzt[l];
        y.wf = N->wt + N->wt_off[l];
        y.wb = l ? N->wtt + N->wtt_off[l] : nullptr;
        y.woff = N->woff[l];
        y.boff = N->boff[l];
        y.in = N->dims[l];
        y.out = N->dims[l + 1];
        y.tiles_n = (y.out + WTN - 1) / WTN;
        y.block0 = blocks;
        y.ksf = l == 0 ? 1 : FHID / 32;
        y.ksb = l + 1 == L ? 4 : FHID / 32;
        blocks += ((y.in + WTM - 1) / WTM) * y.tiles_n;
    }
    g.nlayers = L;
    g.nks = (int)((B + 31) / 32);
    g.ldb = N->ldb;
    g.adam = adam ? 1 : 0;
    g.grad = N->grad;
    g.w = N->w;
    g.m = N->m;
    g.v = N->v;
    g.step = N->step + 1;
    g.step_next = N->step;
    g.lr = lr;
    g.b1 = 0.9f;
    g.b2 = 0.999f;
    g.eps = 1e-8f;
    g.zrow = N->zrow;
    g.loss_part = N->loss_part;
    g.nparts = 32 * g.nks / TBM; /* the row kernel's blocks */
    g.loss = loss.loss;
    g.ring = loss.ring;
    /* row blocks through the last whole k step of 32 rows: the weight
     * kernel reads every row of it (rows past B carry zero gradients) */
    hipLaunchKernelGGL(kTrainRows[L], dim3(32 * g.nks / TBM), dim3(TNT), 0, N->stream, a);
    MLP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(train_weights_kernel, dim3(blocks), dim3(WNT), 0, N->stream, g);
    MLP_CHECK(hipGetLastError());
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

/* B host rows into xf, then as bf16 into xr and xt. */
int upload_rows(xo_mlp *N, const float *x, uint32_t B) {
    MLP_CHECK(hipMemcpyAsync(N->xf, x, sizeof(float) * B * N->dims[0], hipMemcpyHostToDevice,
                             N->stream));
    hipLaunchKernelGGL(rows_kernel, dim3((B + 255) / 256), dim3(256), 0, N->stream, N->xf, B,
                       (uint32_t)N->dims[0], B, 0, 0ull, N->step, N->xr, N->ld[0], N->xt,
                       N->ldb);
    MLP_CHECK(hipGetLastError());
    return 0;
}

/* Training steps per multi-step graph: a run of s steps is s / kGraphSteps
 * launches of one graph and at most one launch of a graph of the remaining
 * s % kGraphSteps steps, instead of one launch per step (each launch
 * boundary costs several microseconds of idle device time). */
constexpr uint32_t kGraphSteps = 64;

void drop_rem_graph(xo_mlp *N) {
    if (N->exec) (void)hipGraphExecDestroy(N->exec);
    if (N->graph) (void)hipGraphDestroy(N->graph);
    N->exec = nullptr;
    N->graph = nullptr;
    N->graph_rem = 0;
}

void drop_graphs(xo_mlp *N) {
    drop_rem_graph(N);
    if (N->exec_k) (void)hipGraphExecDestroy(N->exec_k);
    if (N->graph_k) (void)hipGraphDestroy(N->graph_k);
    N->exec_k = nullptr;
    N->graph_k = nullptr;
}

/* n consecutive training steps captured as one graph. */
int capture_steps(xo_mlp *N, uint32_t n, uint32_t batch, uint64_t seed, float lr,
                  hipGraph_t *graph, hipGraphExec_t *exec) {
    MLP_CHECK(hipStreamBeginCapture(N->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (uint32_t i = 0; i < n && rc == 0; ++i) rc = train_step(N, batch, seed, lr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(N->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (!rc) xo::set_error("xo_mlp_train: capture failed: %s", hipGetErrorString(e));
        return -1;
    }
    *graph = g;
    MLP_CHECK(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return 0;
}

} // namespace

extern "C" {

void xo_mlp_destroy(xo_mlp *N) {
    if (!N) return;
    (void)hipSetDevice(N->ctx->device);
    if (N->stream) (void)hipStreamSynchronize(N->stream);
    drop_graphs(N);
    void *bufs[] = {N->w,   N->m,    N->v,    N->grad,   N->wbf,    N->wtb,  N->bf_index,
                    N->bt_index, N->xf, N->tb, N->xr,    N->xt,     N->dy,   N->dyt,
                    N->y,   N->part, N->loss, N->step,   N->done, N->data_x, N->data_t,
                    N->wt,  N->wtt, N->zrow, N->loss_part};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    for (auto *list : {&N->hr, &N->ht, &N->dz, &N->dzt})
        for (auto p : *list)
            if (p) (void)hipFree(p);
    for (auto &kv : N->dw_tables)
        if (kv.second) (void)hipFree(kv.second);
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
    for (uint32_t l = 0; l <= nlayers; ++l) N->ld.push_back(pad8((int)dims[l]));
    N->max_batch = max_batch;
    /* transposed activations: whole k steps of 32 rows (the fused training
     * step stores them as MFMA fragments over the batch) */
    N->ldb = ((int)max_batch + 31) & ~31;
    N->max_splits = max_batch / 512u < 8u ? (max_batch / 512u > 0u ? max_batch / 512u : 1u) : 8u;
    for (uint32_t l = 0; l < nlayers; ++l) {
        N->woff.push_back(N->nparams);
        N->nparams += (uint64_t)dims[l] * dims[l + 1];
        N->boff.push_back(N->nparams);
        N->nparams += dims[l + 1];
        N->wbf_off.push_back(N->nwbf);
        N->nwbf += (uint64_t)dims[l] * N->ld[l + 1];
        N->wtb_off.push_back(N->nwtb);
        N->nwtb += (uint64_t)dims[l + 1] * N->ld[l];
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
    const uint64_t P = N->nparams, B = max_batch, LB = (uint64_t)N->ldb;
    const int L = (int)nlayers;
    /* bf16 buffers are cleared after allocation: padding columns stay 0 */
    struct {
        void **p;
        size_t bytes;
    } allocs[] = {
        {(void **)&N->w, 4 * P},
        {(void **)&N->m, 4 * P},
        {(void **)&N->v, 4 * P},
        {(void **)&N->grad, 4 * P},
        {(void **)&N->wbf, 2 * N->nwbf},
        {(void **)&N->wtb, 2 * N->nwtb},
        {(void **)&N->bf_index, 8 * P},
        {(void **)&N->bt_index, 8 * P},
        {(void **)&N->xf, 4 * B * dims[0]},
        {(void **)&N->tb, 4 * B * dims[nlayers]},
        {(void **)&N->xr, 2 * B * N->ld[0]},
        /* transposed inputs and output gradients: whole 16-unit tiles (the
         * fused step stores them as fragments) */
        {(void **)&N->xt, 2 * (uint64_t)((dims[0] + 15) & ~15u) * LB},
        {(void **)&N->dy, 2 * B * N->ld[L]},
        {(void **)&N->dyt, 2 * (uint64_t)((dims[nlayers] + 15) & ~15u) * LB},
        {(void **)&N->y, 4 * B * dims[nlayers]},
        {(void **)&N->part, 4ull * N->max_splits * P},
        /* the step counter and the fused training step's copy of it */
        {(void **)&N->step, 8},
        {(void **)&N->done, 4},
        {(void **)&N->zrow, 256},
        {(void **)&N->loss_part, 4ull * (((max_batch + 31) / 32) * 32 / TBM)},
    };
    for (auto &a : allocs) {
        if ((e = hipMalloc(a.p, a.bytes)) != hipSuccess) return bad("allocation", e);
        if ((e = hipMemset(*a.p, 0, a.bytes)) != hipSuccess) return bad("clear", e);
    }
    for (int l = 0; l + 1 < L; ++l) {
        const uint64_t rows_bytes = 2 * B * N->ld[l + 1];
        const uint64_t cols_bytes = 2 * (uint64_t)dims[l + 1] * LB;
        std::vector<uint16_t *> *lists[4] = {&N->hr, &N->ht, &N->dz, &N->dzt};
        const uint64_t sizes[4] = {rows_bytes, cols_bytes, rows_bytes, cols_bytes};
        for (int k = 0; k < 4; ++k) {
            uint16_t *p = nullptr;
            if ((e = hipMalloc(&p, sizes[k])) != hipSuccess) return bad("allocation", e);
            lists[k]->push_back(p);
            if ((e = hipMemset(p, 0, sizes[k])) != hipSuccess) return bad("clear", e);
        }
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
        /* fused training: at most TMAXL layers (the activations of a row
         * block stay in LDS) */
        N->train_fused = nlayers <= (uint32_t)TMAXL;
        if (N->train_fused) {
            uint64_t tb = 0;
            for (uint32_t l = 0; l < nlayers; ++l) {
                N->wtt_off.push_back(tb);
                if (l >= 1) tb += (uint64_t)(l + 1 == nlayers ? 128 : FHID) * FHID;
            }
            if ((e = hipMalloc(&N->wtt, 2 * tb)) != hipSuccess) return bad("allocation", e);
        }
    }
    /* where each matrix entry's bf16 copies are: W_l [in][ld out] and
     * W_l^T [out][ld in]; biases have none */
    std::vector<uint64_t> bi(P, kNotMatrix), ti(P, kNotMatrix);
    for (uint32_t l = 0; l < nlayers; ++l)
        for (uint64_t k = 0; k < dims[l]; ++k)
            for (uint64_t n = 0; n < dims[l + 1]; ++n) {
                const uint64_t i = N->woff[l] + k * dims[l + 1] + n;
                bi[i] = N->wbf_off[l] + k * N->ld[l + 1] + n;
                ti[i] = N->wtb_off[l] + n * N->ld[l] + k;
            }
    if ((e = hipMemcpy(N->bf_index, bi.data(), 8 * P, hipMemcpyHostToDevice)) != hipSuccess)
        return bad("upload", e);
    if ((e = hipMemcpy(N->bt_index, ti.data(), 8 * P, hipMemcpyHostToDevice)) != hipSuccess)
        return bad("upload", e);
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
    MLP_CHECK(hipMemsetAsync(N->step, 0, 8, N->stream));
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
    const uint32_t nin = N->dims[0], nout = N->dims[N->nlayers];
    if (N->train_fused) {
        MLP_CHECK(hipMemcpyAsync(N->xf, x, 4ull * n * nin, hipMemcpyHostToDevice, N->stream));
        if (N->wt_dirty && refresh_transposed(N)) return -1;
    } else if (upload_rows(N, x, n)) {
        return -1;
    }
    MLP_CHECK(hipMemcpyAsync(N->tb, t, 4ull * n * nout, hipMemcpyHostToDevice, N->stream));
    float *dloss = nullptr;
    MLP_CHECK(hipMalloc(&dloss, 4));
    int rc = 0;
    const LossSpec spec{N->tb, n, 0, 0ull, dloss, 0u};
    if (hipMemsetAsync(dloss, 0, 4, N->stream) != hipSuccess)
        rc = -1;
    else if (N->train_fused)
        rc = train_fused_step(N, n, N->xf, spec, false, 0.f);
    else if (forward(N, n, &spec) || backward(N, n))
        rc = -1;
    if (rc == 0 && !N->train_fused) {
        hipLaunchKernelGGL(sum_partials_kernel, dim3(256), dim3(256), 0, N->stream, N->part,
                           batch_splits(N, n), N->nparams, N->grad);
        if (hipGetLastError() != hipSuccess) rc = -1;
    }
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
    drop_graphs(N);
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
    MLP_CHECK(hipSetDevice(N->ctx->device));
    /* the fused step keeps the fragment-major copies current; the
     * layer-by-layer step leaves them to be rebuilt */
    if (N->train_fused) {
        if (N->wt_dirty && refresh_transposed(N)) return -1;
    } else {
        N->wt_dirty = true;
    }
    if (steps > N->loss_cap) {
        if (N->loss) (void)hipFree(N->loss);
        N->loss = nullptr;
        N->loss_cap = 0;
        MLP_CHECK(hipMalloc(&N->loss, 4ull * steps));
        N->loss_cap = steps;
        drop_graphs(N);
    }
    if ((N->exec || N->exec_k) &&
        (N->graph_batch != batch || N->graph_seed != seed || N->graph_lr != lr))
        drop_graphs(N);
    if (!N->train_fused) {
        /* the weight-gradient launch table is uploaded outside any capture */
        const NtProblem *table = nullptr;
        int blocks = 0;
        if (dw_table(N, batch, &table, &blocks)) return -1;
    }
    const uint32_t rem = steps % kGraphSteps;
    if (use_graph && rem && N->graph_rem != rem) {
        drop_rem_graph(N);
        if (capture_steps(N, rem, batch, seed, lr, &N->graph, &N->exec)) return -1;
        N->graph_rem = rem;
    }
    if (use_graph && steps >= kGraphSteps && !N->exec_k &&
        capture_steps(N, kGraphSteps, batch, seed, lr, &N->graph_k, &N->exec_k))
        return -1;
    N->graph_batch = batch;
    N->graph_seed = seed;
    N->graph_lr = lr;
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
    for (uint32_t s = 0; s < steps && rc == 0;) {
        if (use_graph && steps - s >= kGraphSteps) {
            if (hipGraphLaunch(N->exec_k, N->stream) != hipSuccess) rc = -1;
            s += kGraphSteps;
        } else if (use_graph) {
            if (hipGraphLaunch(N->exec, N->stream) != hipSuccess) rc = -1;
            s += rem;
        } else {
            if (train_step(N, batch, seed, lr)) rc = -1;
            ++s;
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
        if (!trans_a && !trans_b) e = gemm<false, false>(g, s);
        else if (trans_a && !trans_b) e = gemm<true, false>(g, s);
        else if (!trans_a && trans_b) e = gemm<false, true>(g, s);
        else e = gemm<true, true>(g, s);
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
