// mlp.hip — the learned surrogate: a BF16 multilayer perceptron on MFMA with
// FP32 accumulation, and its hand-written training kernels.
//
// No counterpart in the reference (it has no learned component).  The
// network maps geometry / tissue scalars plus a sampled gradient waveform to
// the normalised signal of several gradient directions; hidden layers use
// GELU (tanh form), the last layer is linear.
//
// Layout.  Every feature dimension is padded to a multiple of 128 (the GEMM
// tile) and batches to a multiple of 512 (row tiles x split-K slices of 64);
// padded weights, biases and activations are zero.  Activations are row
// major [batch][features] in bf16.  Each layer keeps fp32 master weights
// Wt[out][in] (+ bias) for Adam and two bf16 copies: Wt[out][in] (forward:
// both operands contiguous along the reduction) and W[in][out] (input
// gradient).
//
// Kernels.
//   gemm_nt<EPI, Tile>  C[m][n] = sum_k A[m][k] * B[n][k], 256 x 256 block
//                 tiles (8 waves of 64 x 128, K slabs of 64) or 128 x 128 (4
//                 waves of 64 x 64, slabs of 32), slabs double-buffered
//                 through LDS, v_mfma_f32_32x32x16_bf16; fused epilogues
//                 storing 4 adjacent columns per lane (bias + GELU, bias,
//                 GELU backward, split-K fp32 partial);
//   transpose_bf16, mse_grad, bias_grad, adam (sums the split-K partials in
//   a fixed order: deterministic), gather (minibatch rows by Philox index).
#include "internal.h"

namespace axx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBM = 128, kBN = 128;   // row and column granule of every GEMM
constexpr int kRowPad = 512;  // batches are padded to this (row tiles x split-K slices of 64)
constexpr int kSplitK = 8;    // weight-gradient GEMMs split the batch reduction
constexpr int kBiasSlices = 32;   // bias gradients: row slices summed in order by Adam

enum Epi { EPI_BIAS_GELU = 0, EPI_BIAS = 1, EPI_GELU_BWD = 2, EPI_PARTIAL = 3 };

struct EpiArgs {
    const float* __restrict__ bias;    // [N] (EPI_BIAS*)
    __bf16* __restrict__ out_bf;       // bf16 output [M][ldc]
    __bf16* __restrict__ out_z;        // pre-activation [M][ldc] (EPI_BIAS_GELU; null: not kept)
    float* __restrict__ out_f;         // fp32 output [M][ldc] (EPI_BIAS, EPI_PARTIAL: + slice * M * ldc)
    const __bf16* __restrict__ zprev;  // pre-activation for GELU backward [M][ldc]
    int ldc;
};

// GELU, tanh form: 0.5 x (1 + tanh(u)), u = sqrt(2/pi) (x + 0.044715 x^3),
// evaluated as x * s with s = 1 / (1 + exp(-2u)) (0.5 (1 + tanh u) = s exactly);
// the derivative is s + x s (1 - s) 2 u'(x) with 1 - tanh^2 u = 4 s (1 - s).
// exp and the reciprocal are the hardware's (about 1 ulp): the results are
// rounded to bf16 (8 bits) when stored.
__device__ __forceinline__ float gelu_s(float x) {
    const float c = 0.7978845608028654f;   // sqrt(2/pi)
    const float u = c * (x + 0.044715f * x * x * x);
    return __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}

__device__ __forceinline__ float gelu_f(float x) { return x * gelu_s(x); }

__device__ __forceinline__ float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    const float s = gelu_s(x);
    return s + 2.0f * x * s * (1.0f - s) * c * (1.0f + 3.0f * 0.044715f * x * x);
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint2 pack_bf16x4(float a, float b, float c, float d) {
    const bf16x4 h = {(__bf16)a, (__bf16)b, (__bf16)c, (__bf16)d};
    return __builtin_bit_cast(uint2, h);
}

__device__ __forceinline__ float4 unpack_bf16x4(uint2 v) {
    const bf16x4 h = __builtin_bit_cast(bf16x4, v);
    return make_float4((float)h[0], (float)h[1], (float)h[2], (float)h[3]);
}

// Block tile shapes of gemm_nt: WM x WN waves, each MT x NT tiles of
// 32 x 32, K slabs of BK; WPE = waves per SIMD the register budget allows.
template <int MT_, int NT_, int WM_, int WN_, int BK_, int WPE_>
struct GemmTile {
    static constexpr int MT = MT_, NT = NT_, WM = WM_, WN = WN_, BK = BK_, WPE = WPE_;
    static constexpr int BM = 32 * MT * WM, BN = 32 * NT * WN, T = 64 * WM * WN;
    static constexpr int LDK = BK + 8;       // LDS row pitch BK + 8 bf16: conflict-free b128 reads
    static constexpr int CPR = BK / 8;       // 16-B chunks per slab row
    static constexpr int RS = T / CPR;       // rows per staging pass
    static constexpr int CA = BM / RS, CB = BN / RS;   // staged chunks per thread
    static constexpr int LDS = 2 * (BM + BN) * LDK * 2;
    static_assert(CA * RS == BM && CB * RS == BN && CA <= 4 && CB <= 4, "staging");
};
// 256 x 256, 8 waves of 64 x 128, K slabs of 64 (144 KB LDS: one block per CU)
using TileBig = GemmTile<2, 4, 4, 2, 64, 2>;
// 128 x 128, 4 waves of 64 x 64, K slabs of 32 (40 KB LDS: three blocks per CU)
using TileSmall = GemmTile<2, 2, 2, 2, 32, 3>;

// C = A . B^T over k in [kbeg, kbeg + K_slice); A [M][lda], B [N][ldb] bf16.
// The K loop runs over BK-wide slabs held in two LDS stages: slab t + 1's
// global loads (made at slab t - 1) are written to the free stage after slab
// t's MFMAs, and slab t + 2's loads are issued then, so they are in flight
// during the next slab's MFMAs; one barrier per slab.  The MFMA takes the B
// fragment as its first operand, so that a lane's 16 accumulators of a tile
// are one row m (lane % 32) and columns 8q + 4 (lane / 32) + r: the epilogue
// stores 4 adjacent columns per instruction.  Grid: (M / BM) * (N / BN)
// tiles along x, ordered so that the 8 XCDs (dispatch round-robin) each take
// a contiguous run of row tiles and read each A row slab into one L2;
// split-K slices along y.  M % BM, N % BN and K_slice % BK are zero (checked
// by the host).
template <int EPI, class TL>
__global__ __launch_bounds__(TL::T) __attribute__((amdgpu_waves_per_eu(TL::WPE))) void gemm_nt(
    const __bf16* __restrict__ A, int lda, const __bf16* __restrict__ B, int ldb, int M, int N, int K_slice,
    EpiArgs E) {
    constexpr int MT = TL::MT, NT = TL::NT, BM = TL::BM, BN = TL::BN, BK = TL::BK, RS = TL::RS;
    __shared__ __align__(16) __bf16 As[2][BM][TL::LDK];
    __shared__ __align__(16) __bf16 Bs[2][BN][TL::LDK];
    const int nbn = N / BN, nblk = (M / BM) * nbn;
    int b = blockIdx.x;
    if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
    const int n0 = (b % nbn) * BN, m0 = (b / nbn) * BM;
    const int kbeg = blockIdx.y * K_slice, nk = K_slice / BK;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % TL::WM, wn = wave / TL::WM;
    // global -> register staging: thread t moves 16 B of rows t / CPR + RS i,
    // bf16 columns 8 (t % CPR) .. + 7 of the slab
    const int lr = tid / TL::CPR, lc = (tid % TL::CPR) * 8;
    const __bf16* Ag = A + (size_t)(m0 + lr) * lda + kbeg + lc;
    const __bf16* Bg = B + (size_t)(n0 + lr) * ldb + kbeg + lc;
    const size_t arow = (size_t)RS * lda, brow = (size_t)RS * ldb;
    // staged slab in named registers (an array here would be placed in LDS
    // or scratch by the compiler's private-array handling)
    uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define AXX_LA(i) if constexpr (i < TL::CA) ra##i = *reinterpret_cast<const uint4*>(Ag + (i) * arow + ko);
#define AXX_LB(i) if constexpr (i < TL::CB) rb##i = *reinterpret_cast<const uint4*>(Bg + (i) * brow + ko);
#define AXX_SA(i) if constexpr (i < TL::CA) *reinterpret_cast<uint4*>(&As[st][lr + RS * (i)][lc]) = ra##i;
#define AXX_SB(i) if constexpr (i < TL::CB) *reinterpret_cast<uint4*>(&Bs[st][lr + RS * (i)][lc]) = rb##i;
#define AXX_GEMM_LOAD(t)                                                  \
    do {                                                                  \
        const int ko = (t) * BK;                                          \
        AXX_LA(0) AXX_LA(1) AXX_LA(2) AXX_LA(3)                           \
        AXX_LB(0) AXX_LB(1) AXX_LB(2) AXX_LB(3)                           \
    } while (0)
#define AXX_GEMM_STORE(s)                                                 \
    do {                                                                  \
        const int st = (s);                                               \
        AXX_SA(0) AXX_SA(1) AXX_SA(2) AXX_SA(3)                           \
        AXX_SB(0) AXX_SB(1) AXX_SB(2) AXX_SB(3)                           \
    } while (0)
    f32x16 acc[MT][NT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
    AXX_GEMM_LOAD(0);
    AXX_GEMM_STORE(0);
    if (nk > 1) AXX_GEMM_LOAD(1);
    __syncthreads();
    for (int t = 0; t < nk; ++t) {
        const int cur = t & 1;
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            bf16x8 a[MT], bb[NT];
#pragma unroll
            for (int i = 0; i < MT; ++i)
                a[i] = *reinterpret_cast<const bf16x8*>(&As[cur][wm * 32 * MT + i * 32 + (lane & 31)][kk + (lane >> 5) * 8]);
#pragma unroll
            for (int j = 0; j < NT; ++j)
                bb[j] = *reinterpret_cast<const bf16x8*>(&Bs[cur][wn * 32 * NT + j * 32 + (lane & 31)][kk + (lane >> 5) * 8]);
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bb[j], a[i], acc[i][j], 0, 0, 0);
        }
        if (t + 1 < nk) {
            AXX_GEMM_STORE(cur ^ 1);
            if (t + 2 < nk) AXX_GEMM_LOAD(t + 2);
        }
        __syncthreads();
    }
#undef AXX_GEMM_LOAD
#undef AXX_GEMM_STORE
#undef AXX_LA
#undef AXX_LB
#undef AXX_SA
#undef AXX_SB
    // acc[i][j][4 q + r]: row m0 + 32 (MT wm + i) + lane % 32, column
    // n0 + 32 (NT wn + j) + 8 q + 4 (lane / 32) + r
    const int row0 = m0 + wm * 32 * MT + (lane & 31);
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int col = n0 + wn * 32 * NT + j * 32 + 8 * q + 4 * (lane >> 5);
            float4 bc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (EPI == EPI_BIAS_GELU || EPI == EPI_BIAS) bc = *reinterpret_cast<const float4*>(E.bias + col);
#pragma unroll
            for (int i = 0; i < MT; ++i) {
                const size_t o = (size_t)(row0 + 32 * i) * E.ldc + col;
                const f32x16& v = acc[i][j];
                if (EPI == EPI_BIAS_GELU) {
                    const float z0 = v[4 * q] + bc.x, z1 = v[4 * q + 1] + bc.y;
                    const float z2 = v[4 * q + 2] + bc.z, z3 = v[4 * q + 3] + bc.w;
                    if (E.out_z) *reinterpret_cast<uint2*>(E.out_z + o) = pack_bf16x4(z0, z1, z2, z3);
                    *reinterpret_cast<uint2*>(E.out_bf + o) = pack_bf16x4(gelu_f(z0), gelu_f(z1), gelu_f(z2), gelu_f(z3));
                } else if (EPI == EPI_BIAS) {
                    *reinterpret_cast<float4*>(E.out_f + o) =
                        make_float4(v[4 * q] + bc.x, v[4 * q + 1] + bc.y, v[4 * q + 2] + bc.z, v[4 * q + 3] + bc.w);
                } else if (EPI == EPI_GELU_BWD) {
                    const float4 z = unpack_bf16x4(*reinterpret_cast<const uint2*>(E.zprev + o));
                    *reinterpret_cast<uint2*>(E.out_bf + o) =
                        pack_bf16x4(v[4 * q] * gelu_grad(z.x), v[4 * q + 1] * gelu_grad(z.y),
                                    v[4 * q + 2] * gelu_grad(z.z), v[4 * q + 3] * gelu_grad(z.w));
                } else {
                    *reinterpret_cast<float4*>(E.out_f + (size_t)blockIdx.y * M * E.ldc + o) =
                        make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
                }
            }
        }
}

// out[c][r] = in[r][c]; R and C multiples of 64.  Grid (C / 64, R / 64), 256 threads.
__global__ __launch_bounds__(256) void transpose_bf16(const __bf16* __restrict__ in, int ldi,
                                                      __bf16* __restrict__ out, int ldo) {
    __shared__ __bf16 t[64][66];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = e >> 6, c = e & 63;
        t[r][c] = in[(size_t)(r0 + r) * ldi + c0 + c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int c = e >> 6, r = e & 63;
        out[(size_t)(c0 + c) * ldo + r0 + r] = t[r][c];
    }
}

// Mean-squared error over the first n_out of ldy outputs of n rows:
// loss_sum[0] += sum (y - t)^2 (one partial per block, added in block order
// by the host-side copy of `partial`), dy = 2 (y - t) / (n * n_out) in bf16,
// zero in the padding.
__global__ __launch_bounds__(256) void mse_grad(const float* __restrict__ y, const float* __restrict__ t,
                                                int n, int n_out, int ldy, float scale,
                                                __bf16* __restrict__ dy, float* __restrict__ partial) {
    __shared__ float sh[256];
    float s = 0.0f;
    const long long total = (long long)n * ldy;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long long)gridDim.x * 256) {
        const int r = (int)(e / ldy), c = (int)(e % ldy);
        float g = 0.0f;
        if (c < n_out) {
            const float d = y[e] - t[(size_t)r * n_out + c];
            s += d * d;
            g = 2.0f * d * scale;
        }
        dy[e] = (__bf16)g;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out[0] = scale * sum of the n partial losses (fixed tree order).
__global__ void sum_partials(const float* __restrict__ part, int n, float scale, float* __restrict__ out) {
    __shared__ float sh[256];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0] * scale;
}

// db[s][c] = sum over the rows of slice s (of kBiasSlices) of dz[r][c]:
// adjacent threads take adjacent columns.  Grid (ld / 256, kBiasSlices).
__global__ void bias_grad(const __bf16* __restrict__ dz, int n, int ld, float* __restrict__ db) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ld) return;
    const int per = (n + kBiasSlices - 1) / kBiasSlices;
    const int r0 = blockIdx.y * per, r1 = min(n, r0 + per);
    float s = 0.0f;
    for (int r = r0; r < r1; ++r) s += (float)dz[(size_t)r * ld + c];
    db[(size_t)blockIdx.y * ld + c] = s;
}

// Adam on the fp32 master weights Wt[out][in] (gradient = sum of kSplitK
// partials in slice order) and biases; writes the bf16 copies Wt and W.
struct AdamArgs {
    float* __restrict__ w;          // [outp][inp] master
    float* __restrict__ mw;
    float* __restrict__ vw;
    const float* __restrict__ gpart;   // [kSplitK][outp][inp]
    float* __restrict__ b;          // [outp]
    float* __restrict__ mb;
    float* __restrict__ vb;
    const float* __restrict__ gb;   // [kBiasSlices][outp]
    __bf16* __restrict__ wt_bf;     // [outp][inp]
    __bf16* __restrict__ w_bf;      // [inp][outp]
    int outp, inp, out_real, in_real;
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
};

__global__ __launch_bounds__(256) void adam_kernel(AdamArgs P) {
    const long long nw = (long long)P.outp * P.inp;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < nw) {
        const int o = (int)(e / P.inp), i = (int)(e % P.inp);
        float w = P.w[e];
        if (o < P.out_real && i < P.in_real) {
            float g = 0.0f;
            for (int s = 0; s < kSplitK; ++s) g += P.gpart[(size_t)s * nw + e];
            const float m = P.b1 * P.mw[e] + (1.0f - P.b1) * g;
            const float v = P.b2 * P.vw[e] + (1.0f - P.b2) * g * g;
            P.mw[e] = m;
            P.vw[e] = v;
            w -= P.lr * (m / P.c1) / (sqrtf(v / P.c2) + P.eps);
            P.w[e] = w;
        }
        P.wt_bf[e] = (__bf16)w;
        P.w_bf[(size_t)i * P.outp + o] = (__bf16)w;
    }
    if (e < P.outp && e < P.out_real) {
        float g = 0.0f;
        for (int s = 0; s < kBiasSlices; ++s) g += P.gb[(size_t)s * P.outp + e];
        const float m = P.b1 * P.mb[e] + (1.0f - P.b1) * g;
        const float v = P.b2 * P.vb[e] + (1.0f - P.b2) * g * g;
        P.mb[e] = m;
        P.vb[e] = v;
        P.b[e] -= P.lr * (m / P.c1) / (sqrtf(v / P.c2) + P.eps);
    }
}

// Minibatch rows: row j of the batch is dataset row idx_j, idx_j = Philox
// word mod n (counter (j, step, 3 << 24, 0), key = seed).
__global__ void gather_rows(const __bf16* __restrict__ X, int ldx, const float* __restrict__ T, int n_out,
                            int n_data, int batch, uint32_t step, uint32_t k0, uint32_t k1,
                            __bf16* __restrict__ xb, float* __restrict__ tb) {
    const int j = blockIdx.x;
    if (j >= batch) return;
    const u32x4 o = philox10((uint32_t)j, step, 3u << 24, 0u, k0, k1);
    const int r = (int)(o.x % (uint32_t)n_data);
    for (int c = threadIdx.x; c < ldx; c += blockDim.x) xb[(size_t)j * ldx + c] = X[(size_t)r * ldx + c];
    for (int c = threadIdx.x; c < n_out; c += blockDim.x) tb[(size_t)j * n_out + c] = T[(size_t)r * n_out + c];
}

// fp32 [n][in_real] -> bf16 [n][ldx] with zero padding.
__global__ void to_bf16_padded(const float* __restrict__ x, int n, int in_real, int ldx, __bf16* __restrict__ out) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * ldx) return;
    const int r = (int)(e / ldx), c = (int)(e % ldx);
    out[e] = (__bf16)(c < in_real ? x[(size_t)r * in_real + c] : 0.0f);
}

// Copy the first n_out columns of y [n][ldy] to out [n][n_out].
__global__ void unpad_f32(const float* __restrict__ y, int n, int ldy, int n_out, float* __restrict__ out) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * n_out) return;
    const int r = (int)(e / n_out), c = (int)(e % n_out);
    out[e] = y[(size_t)r * ldy + c];
}

}  // namespace axx

// ---------------------------------------------------------------------------
// Host side

namespace {

int pad_to(int v, int m) { return (v + m - 1) / m * m; }

struct Layer {
    int in = 0, out = 0, inp = 0, outp = 0;   // real and padded sizes
    int kin = 0;   // reduction length of the forward GEMM: inp, or for the first layer in padded to 64
                   // (the input columns beyond it are zero in both operands)
    bool gelu = true;
    float *w = nullptr, *mw = nullptr, *vw = nullptr, *b = nullptr, *mb = nullptr, *vb = nullptr;
    float *gpart = nullptr, *gb = nullptr;
    __bf16 *wt_bf = nullptr, *w_bf = nullptr;
};

}  // namespace

struct axx_mlp {
    axx_ctx* ctx = nullptr;
    std::vector<Layer> L;
    int max_batch = 0;
    int adam_t = 0;
    // activations: H[l] = input of layer l ([max_batch][L[l].inp]); Z[l] = pre-activation
    // of hidden layer l ([max_batch][L[l].outp]); Y = output fp32
    std::vector<__bf16*> H, Z;
    float* Y = nullptr;
    __bf16 *dA = nullptr, *dB = nullptr;        // gradient ping-pong [max_batch][maxdim]
    __bf16 *HT = nullptr, *dZT = nullptr;       // transposes [maxdim][max_batch]
    float* T = nullptr;                         // targets of a batch [max_batch][out]
    float* loss_part = nullptr;                 // [kLossBlocks]
    float* xin = nullptr;                       // fp32 staging [max_batch][in]
    // resident dataset
    __bf16* dX = nullptr;
    float* dT = nullptr;
    int n_data = 0;
};

namespace {

constexpr int kLossBlocks = 256;

void mlp_free(axx_mlp* m) {
    if (!m) return;
    for (Layer& l : m->L) {
        void* bufs[] = {l.w, l.mw, l.vw, l.b, l.mb, l.vb, l.gpart, l.gb, l.wt_bf, l.w_bf};
        for (void* p : bufs)
            if (p) (void)hipFree(p);
    }
    for (__bf16* p : m->H)
        if (p) (void)hipFree(p);
    for (__bf16* p : m->Z)
        if (p) (void)hipFree(p);
    void* bufs[] = {m->Y, m->dA, m->dB, m->HT, m->dZT, m->T, m->loss_part, m->xin, m->dX, m->dT};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    delete m;
}

template <typename T>
int dalloc(axx_ctx* ctx, T** p, size_t n, bool zero = true) {
    HIP_TRY(ctx, hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)));
    if (zero) HIP_TRY(ctx, hipMemset(*p, 0, std::max<size_t>(n, 1) * sizeof(T)));
    return AXX_OK;
}

template <int EPI>
void launch_epi(axx_ctx* ctx, bool big, dim3 grid, const __bf16* A, int lda, const __bf16* B, int ldb, int M,
                int N, int ks, const EpiArgs& E) {
    if (big)
        hipLaunchKernelGGL((gemm_nt<EPI, TileBig>), grid, dim3(TileBig::T), 0, ctx->stream, A, lda, B, ldb, M, N, ks, E);
    else
        hipLaunchKernelGGL((gemm_nt<EPI, TileSmall>), grid, dim3(TileSmall::T), 0, ctx->stream, A, lda, B, ldb, M, N,
                           ks, E);
}

// 256 x 256 tiles where the shape divides and the grid still holds a block
// per CU, else 128 x 128.
int launch_gemm(axx_ctx* ctx, int epi, const __bf16* A, int lda, const __bf16* B, int ldb, int M, int N, int K,
                int slices, const EpiArgs& E) {
    if (M % TileSmall::BM || N % TileSmall::BN || K % (TileSmall::BK * slices) || (E.ldc & 3))
        return fail(ctx, AXX_ERR_INVALID, "gemm shape %d x %d x %d (slices %d) not tiled", M, N, K, slices);
    const int ks = K / slices;
    const bool big = M % TileBig::BM == 0 && N % TileBig::BN == 0 && ks % TileBig::BK == 0 &&
                     (long long)(M / TileBig::BM) * (N / TileBig::BN) * slices >= 256;
    const int bm = big ? TileBig::BM : TileSmall::BM, bn = big ? TileBig::BN : TileSmall::BN;
    const dim3 grid((unsigned)((M / bm) * (N / bn)), (unsigned)slices);
    switch (epi) {
    case EPI_BIAS_GELU: launch_epi<EPI_BIAS_GELU>(ctx, big, grid, A, lda, B, ldb, M, N, ks, E); break;
    case EPI_BIAS: launch_epi<EPI_BIAS>(ctx, big, grid, A, lda, B, ldb, M, N, ks, E); break;
    case EPI_GELU_BWD: launch_epi<EPI_GELU_BWD>(ctx, big, grid, A, lda, B, ldb, M, N, ks, E); break;
    default: launch_epi<EPI_PARTIAL>(ctx, big, grid, A, lda, B, ldb, M, N, ks, E); break;
    }
    HIP_TRY(ctx, hipGetLastError());
    return AXX_OK;
}

int transpose(axx_ctx* ctx, const __bf16* in, int R, int Cc, int ldi, __bf16* out, int ldo) {
    hipLaunchKernelGGL(transpose_bf16, dim3(Cc / 64, R / 64), dim3(256), 0, ctx->stream, in, ldi, out, ldo);
    HIP_TRY(ctx, hipGetLastError());
    return AXX_OK;
}

// Forward over n rows already in H[0] (n % 128 == 0); keep_z: store the
// hidden layers' pre-activations for a backward pass.
int forward(axx_mlp* m, int n, bool keep_z) {
    axx_ctx* ctx = m->ctx;
    const int nl = (int)m->L.size();
    for (int l = 0; l < nl; ++l) {
        const Layer& y = m->L[l];
        EpiArgs E{};
        E.bias = y.b;
        E.ldc = y.outp;
        int rc;
        if (y.gelu) {
            E.out_bf = m->H[l + 1];
            E.out_z = keep_z ? m->Z[l] : nullptr;
            rc = launch_gemm(ctx, EPI_BIAS_GELU, m->H[l], y.inp, y.wt_bf, y.inp, n, y.outp, y.kin, 1, E);
        } else {
            E.out_f = m->Y;
            rc = launch_gemm(ctx, EPI_BIAS, m->H[l], y.inp, y.wt_bf, y.inp, n, y.outp, y.kin, 1, E);
        }
        if (rc) return rc;
    }
    return AXX_OK;
}

// Loss and gradients of the batch in H[0] / T (n rows): fills every layer's
// gpart (split-K partials) and gb; returns the loss partials in loss_part.
int backward(axx_mlp* m, int n) {
    axx_ctx* ctx = m->ctx;
    const int nl = (int)m->L.size();
    const Layer& last = m->L[nl - 1];
    hipLaunchKernelGGL(mse_grad, dim3(kLossBlocks), dim3(256), 0, ctx->stream, (const float*)m->Y,
                       (const float*)m->T, n, last.out, last.outp, 1.0f / ((float)n * last.out), m->dA,
                       m->loss_part);
    HIP_TRY(ctx, hipGetLastError());
    __bf16* dcur = m->dA;   // gradient w.r.t. the output of layer l (pre-activation for hidden)
    __bf16* dnext = m->dB;
    for (int l = nl - 1; l >= 0; --l) {
        const Layer& y = m->L[l];
        int rc;
        // bias gradient
        hipLaunchKernelGGL(bias_grad, dim3((y.outp + 255) / 256, kBiasSlices), dim3(256), 0, ctx->stream,
                           (const __bf16*)dcur, n, y.outp, y.gb);
        HIP_TRY(ctx, hipGetLastError());
        // weight gradient: gWt[o][i] = sum_r dZ[r][o] H[r][i] = (dZ^T)(H^T)^T
        if ((rc = transpose(ctx, dcur, n, y.outp, y.outp, m->dZT, n))) return rc;
        if ((rc = transpose(ctx, m->H[l], n, y.inp, y.inp, m->HT, n))) return rc;
        EpiArgs E{};
        E.out_f = y.gpart;
        E.ldc = y.inp;
        if ((rc = launch_gemm(ctx, EPI_PARTIAL, m->dZT, n, m->HT, n, y.outp, y.inp, n, kSplitK, E))) return rc;
        // input gradient through the previous layer's GELU
        if (l > 0) {
            EpiArgs G{};
            G.out_bf = dnext;
            G.zprev = m->Z[l - 1];
            G.ldc = y.inp;
            if ((rc = launch_gemm(ctx, EPI_GELU_BWD, dcur, y.outp, y.w_bf, y.outp, n, y.inp, y.outp, 1, G)))
                return rc;
            std::swap(dcur, dnext);
        }
    }
    return AXX_OK;
}

int adam_step(axx_mlp* m, float lr) {
    axx_ctx* ctx = m->ctx;
    m->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    for (Layer& y : m->L) {
        AdamArgs P{};
        P.w = y.w; P.mw = y.mw; P.vw = y.vw; P.gpart = y.gpart;
        P.b = y.b; P.mb = y.mb; P.vb = y.vb; P.gb = y.gb;
        P.wt_bf = y.wt_bf; P.w_bf = y.w_bf;
        P.outp = y.outp; P.inp = y.inp; P.out_real = y.out; P.in_real = y.in;
        P.lr = lr; P.b1 = b1; P.b2 = b2; P.eps = 1e-8f;
        P.c1 = 1.0f - std::pow(b1, (float)m->adam_t);
        P.c2 = 1.0f - std::pow(b2, (float)m->adam_t);
        const long long nw = (long long)y.outp * y.inp;
        hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, ctx->stream, P);
        HIP_TRY(ctx, hipGetLastError());
    }
    return AXX_OK;
}

// bf16 copies of the master weights (after set_params / init): an Adam
// kernel with lr = 0 rewrites them without changing the masters.
int refresh_bf16(axx_mlp* m) {
    axx_ctx* ctx = m->ctx;
    for (Layer& y : m->L) {
        AdamArgs P{};
        P.w = y.w; P.mw = y.mw; P.vw = y.vw; P.gpart = y.gpart;
        P.b = y.b; P.mb = y.mb; P.vb = y.vb; P.gb = y.gb;
        P.wt_bf = y.wt_bf; P.w_bf = y.w_bf;
        P.outp = y.outp; P.inp = y.inp; P.out_real = 0; P.in_real = 0;   // no update
        P.lr = 0.0f; P.b1 = 0.9f; P.b2 = 0.999f; P.eps = 1e-8f; P.c1 = 1.0f; P.c2 = 1.0f;
        const long long nw = (long long)y.outp * y.inp;
        hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, ctx->stream, P);
        HIP_TRY(ctx, hipGetLastError());
    }
    return AXX_OK;
}

int load_inputs(axx_mlp* m, const float* x_host, int n, int n_pad) {
    axx_ctx* ctx = m->ctx;
    const Layer& f = m->L[0];
    HIP_TRY(ctx, hipMemsetAsync(m->xin, 0, (size_t)n_pad * f.in * sizeof(float), ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(m->xin, x_host, (size_t)n * f.in * sizeof(float), hipMemcpyHostToDevice,
                                ctx->stream));
    const long long tot = (long long)n_pad * f.inp;
    hipLaunchKernelGGL(to_bf16_padded, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const float*)m->xin, n_pad, f.in, f.inp, m->H[0]);
    HIP_TRY(ctx, hipGetLastError());
    return AXX_OK;
}

}  // namespace

extern "C" void axx_mlp_destroy(axx_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    mlp_free(m);
}

extern "C" int axx_mlp_create(axx_ctx* ctx, int n_layers, const int* dims, int max_batch, axx_mlp** out) {
    axx_mlp* m = nullptr;   // freed by the handler if construction throws
    try {
        if (!ctx || !dims || !out || n_layers < 1 || n_layers > 64 || max_batch < 1)
            return fail(ctx, AXX_ERR_INVALID, "mlp: bad arguments");
        *out = nullptr;
        for (int l = 0; l <= n_layers; ++l)
            if (dims[l] < 1 || dims[l] > 16384) return fail(ctx, AXX_ERR_INVALID, "mlp: dims[%d] = %d", l, dims[l]);
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        m = new (std::nothrow) axx_mlp();
        if (!m) return AXX_ERR_NOMEM;
        m->ctx = ctx;
        m->max_batch = pad_to(max_batch, kRowPad);
        int maxdim = 0;
        for (int l = 0; l < n_layers; ++l) {
            Layer y;
            y.in = dims[l];
            y.out = dims[l + 1];
            // inputs of layer l are the padded outputs of layer l - 1
            y.inp = l == 0 ? pad_to(y.in, kBN) : m->L[l - 1].outp;
            y.outp = pad_to(y.out, kBN);
            y.kin = l == 0 ? std::min(y.inp, pad_to(y.in, TileBig::BK)) : y.inp;
            y.gelu = l + 1 < n_layers;
            maxdim = std::max(maxdim, std::max(y.inp, y.outp));
            m->L.push_back(y);
        }
        const size_t B = (size_t)m->max_batch;
        int rc = AXX_OK;
        for (Layer& y : m->L) {
            const size_t nw = (size_t)y.outp * y.inp;
            if ((rc = dalloc(ctx, &y.w, nw)) || (rc = dalloc(ctx, &y.mw, nw)) || (rc = dalloc(ctx, &y.vw, nw)) ||
                (rc = dalloc(ctx, &y.b, y.outp)) || (rc = dalloc(ctx, &y.mb, y.outp)) ||
                (rc = dalloc(ctx, &y.vb, y.outp)) || (rc = dalloc(ctx, &y.gpart, nw * kSplitK)) ||
                (rc = dalloc(ctx, &y.gb, (size_t)y.outp * kBiasSlices)) || (rc = dalloc(ctx, &y.wt_bf, nw)) ||
                (rc = dalloc(ctx, &y.w_bf, nw)))
                break;
        }
        m->H.assign(n_layers + 1, nullptr);
        m->Z.assign(n_layers, nullptr);
        for (int l = 0; !rc && l < n_layers; ++l) {
            rc = dalloc(ctx, &m->H[l], B * m->L[l].inp);
            if (!rc && m->L[l].gelu) {
                rc = dalloc(ctx, &m->Z[l], B * m->L[l].outp);
                if (!rc) rc = dalloc(ctx, &m->H[l + 1], B * m->L[l].outp);
            }
        }
        const Layer& last = m->L.back();
        if (!rc) rc = dalloc(ctx, &m->Y, B * last.outp);
        if (!rc) rc = dalloc(ctx, &m->dA, B * maxdim);
        if (!rc) rc = dalloc(ctx, &m->dB, B * maxdim);
        if (!rc) rc = dalloc(ctx, &m->HT, B * maxdim);
        if (!rc) rc = dalloc(ctx, &m->dZT, B * maxdim);
        if (!rc) rc = dalloc(ctx, &m->T, B * last.out);
        if (!rc) rc = dalloc(ctx, &m->loss_part, kLossBlocks);
        if (!rc) rc = dalloc(ctx, &m->xin, B * m->L[0].in);
        if (rc) {
            mlp_free(m);
            return rc;
        }
        *out = m;
        return AXX_OK;
    } catch (...) {
        if (m && !(out && *out == m)) mlp_free(m);
        if (out) *out = nullptr;
        return api_exception();
    }
}

// Parameters, packed per layer: W[out][in] row major, then b[out] (fp32).
extern "C" int axx_mlp_param_count(const axx_mlp* m, int64_t* n) {
    try {
        if (!m || !n) return AXX_ERR_INVALID;
        int64_t s = 0;
        for (const Layer& y : m->L) s += (int64_t)y.out * y.in + y.out;
        *n = s;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_mlp_set_params(axx_mlp* m, const float* p) {
    try {
        if (!m || !p) return AXX_ERR_INVALID;
        axx_ctx* ctx = m->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        for (Layer& y : m->L) {
            std::vector<float> w((size_t)y.outp * y.inp, 0.0f), b(y.outp, 0.0f);
            for (int o = 0; o < y.out; ++o)
                for (int i = 0; i < y.in; ++i) w[(size_t)o * y.inp + i] = *p++;
            for (int o = 0; o < y.out; ++o) b[o] = *p++;
            HIP_TRY(ctx, hipMemcpy(y.w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
            HIP_TRY(ctx, hipMemcpy(y.b, b.data(), b.size() * sizeof(float), hipMemcpyHostToDevice));
            HIP_TRY(ctx, hipMemset(y.mw, 0, w.size() * sizeof(float)));
            HIP_TRY(ctx, hipMemset(y.vw, 0, w.size() * sizeof(float)));
            HIP_TRY(ctx, hipMemset(y.mb, 0, b.size() * sizeof(float)));
            HIP_TRY(ctx, hipMemset(y.vb, 0, b.size() * sizeof(float)));
        }
        m->adam_t = 0;
        int rc = refresh_bf16(m);
        if (rc) return rc;
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

extern "C" int axx_mlp_get_params(axx_mlp* m, float* p) {
    try {
        if (!m || !p) return AXX_ERR_INVALID;
        axx_ctx* ctx = m->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        for (Layer& y : m->L) {
            std::vector<float> w((size_t)y.outp * y.inp), b(y.outp);
            HIP_TRY(ctx, hipMemcpy(w.data(), y.w, w.size() * sizeof(float), hipMemcpyDeviceToHost));
            HIP_TRY(ctx, hipMemcpy(b.data(), y.b, b.size() * sizeof(float), hipMemcpyDeviceToHost));
            for (int o = 0; o < y.out; ++o)
                for (int i = 0; i < y.in; ++i) *p++ = w[(size_t)o * y.inp + i];
            for (int o = 0; o < y.out; ++o) *p++ = b[o];
        }
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

// y[n][out] = network(x[n][in]), host buffers; n <= max_batch.
extern "C" int axx_mlp_forward(axx_mlp* m, const float* x, int n, float* y) {
    try {
        if (!m || !x || !y || n < 1 || n > m->max_batch) return fail(m ? m->ctx : nullptr, AXX_ERR_INVALID, "mlp forward: n");
        axx_ctx* ctx = m->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        const int n_pad = pad_to(n, kBM);
        int rc = load_inputs(m, x, n, n_pad);
        if (!rc) rc = forward(m, n_pad, false);
        if (rc) return rc;
        const Layer& last = m->L.back();
        float* tmp = m->xin;   // reuse the staging buffer for the unpadded output
        if ((size_t)n * last.out > (size_t)m->max_batch * m->L[0].in) {
            std::vector<float> full((size_t)n_pad * last.outp);
            HIP_TRY(ctx, hipMemcpyAsync(full.data(), m->Y, full.size() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            for (int r = 0; r < n; ++r)
                for (int c = 0; c < last.out; ++c) y[(size_t)r * last.out + c] = full[(size_t)r * last.outp + c];
            return AXX_OK;
        }
        const long long tot = (long long)n * last.out;
        hipLaunchKernelGGL(unpad_f32, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const float*)m->Y, n, last.outp, last.out, tmp);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipMemcpyAsync(y, tmp, (size_t)tot * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

// Forward over the n_pad rows already resident in the input buffer, timed
// with HIP events (inputs staged by a previous axx_mlp_forward).
extern "C" int axx_mlp_time_forward(axx_mlp* m, int n, float* ms) {
    try {
        if (!m || !ms || n < 1 || n > m->max_batch) return AXX_ERR_INVALID;
        axx_ctx* ctx = m->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        hipEvent_t e0, e1;
        HIP_TRY(ctx, hipEventCreate(&e0));
        HIP_TRY(ctx, hipEventCreate(&e1));
        HIP_TRY(ctx, hipEventRecord(e0, ctx->stream));
        int rc = forward(m, pad_to(n, kBM), false);
        HIP_TRY(ctx, hipEventRecord(e1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(e1));
        HIP_TRY(ctx, hipEventElapsedTime(ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return rc;
    } catch (...) {
        return api_exception();
    }
}

// Loss (mean squared error over the n x out targets) and gradients of the
// mean loss w.r.t. the packed parameters, for the batch (x, t).
extern "C" int axx_mlp_loss_grad(axx_mlp* m, const float* x, const float* t, int n, float* loss, float* grad) {
    try {
        if (!m || !x || !t || n < 1 || n > m->max_batch) return fail(m ? m->ctx : nullptr, AXX_ERR_INVALID, "mlp grad: n");
        axx_ctx* ctx = m->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        const int n_pad = pad_to(n, kRowPad);   // M tiles and split-K slices
        const Layer& last = m->L.back();
        int rc = load_inputs(m, x, n, n_pad);
        if (rc) return rc;
        HIP_TRY(ctx, hipMemsetAsync(m->T, 0, (size_t)n_pad * last.out * sizeof(float), ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(m->T, t, (size_t)n * last.out * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        if ((rc = forward(m, n_pad, true))) return rc;
        if (n_pad != n) {
            // padded rows must not contribute: their targets equal their outputs
            HIP_TRY(ctx, hipMemcpy2DAsync(m->T + (size_t)n * last.out, last.out * sizeof(float),
                                          m->Y + (size_t)n * last.outp, last.outp * sizeof(float),
                                          last.out * sizeof(float), n_pad - n, hipMemcpyDeviceToDevice,
                                          ctx->stream));
        }
        if ((rc = backward(m, n_pad))) return rc;
        std::vector<float> part(kLossBlocks);
        HIP_TRY(ctx, hipMemcpyAsync(part.data(), m->loss_part, kLossBlocks * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        double s = 0.0;
        for (float v : part) s += v;
        if (loss) *loss = (float)(s / ((double)n * last.out));
        // the scale of dY used n_pad rows; rescale to the mean over n rows
        const float rescale = (float)n_pad / (float)n;
        if (grad) {
            for (Layer& y : m->L) {
                const size_t nw = (size_t)y.outp * y.inp;
                std::vector<float> gp(nw * kSplitK), gb((size_t)y.outp * kBiasSlices);
                HIP_TRY(ctx, hipMemcpy(gp.data(), y.gpart, gp.size() * sizeof(float), hipMemcpyDeviceToHost));
                HIP_TRY(ctx, hipMemcpy(gb.data(), y.gb, gb.size() * sizeof(float), hipMemcpyDeviceToHost));
                for (int o = 0; o < y.out; ++o)
                    for (int i = 0; i < y.in; ++i) {
                        float g = 0.0f;
                        for (int sl = 0; sl < kSplitK; ++sl) g += gp[(size_t)sl * nw + (size_t)o * y.inp + i];
                        *grad++ = g * rescale;
                    }
                for (int o = 0; o < y.out; ++o) {
                    float g = 0.0f;
                    for (int sl = 0; sl < kBiasSlices; ++sl) g += gb[(size_t)sl * y.outp + o];
                    *grad++ = g * rescale;
                }
            }
        }
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

// Keep a dataset on the device for training: x[n][in], t[n][out].
extern "C" int axx_mlp_set_data(axx_mlp* m, const float* x, const float* t, int n) {
    try {
        if (!m || !x || !t || n < 1) return AXX_ERR_INVALID;
        axx_ctx* ctx = m->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (m->dX) (void)hipFree(m->dX);
        if (m->dT) (void)hipFree(m->dT);
        m->dX = nullptr;
        m->dT = nullptr;
        const Layer& f = m->L[0];
        const Layer& last = m->L.back();
        int rc = dalloc(ctx, &m->dX, (size_t)n * f.inp, false);
        if (!rc) rc = dalloc(ctx, &m->dT, (size_t)n * last.out, false);
        if (rc) return rc;
        float* tmp = nullptr;
        if ((rc = dalloc(ctx, &tmp, (size_t)n * f.in, false))) return rc;
        HIP_TRY(ctx, hipMemcpy(tmp, x, (size_t)n * f.in * sizeof(float), hipMemcpyHostToDevice));
        const long long tot = (long long)n * f.inp;
        hipLaunchKernelGGL(to_bf16_padded, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const float*)tmp, n, f.in, f.inp, m->dX);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipMemcpy(m->dT, t, (size_t)n * last.out * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(tmp);
        m->n_data = n;
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}

// Adam training on the resident dataset: `steps` minibatches of `batch`
// rows drawn with Philox (key = seed, counter = (row, global step)).
// loss_trace[s] (optional) receives each step's minibatch loss.
extern "C" int axx_mlp_train(axx_mlp* m, int steps, int batch, float lr, uint64_t seed, float* loss_trace) {
    try {
        if (!m || steps < 1 || batch < 1 || batch > m->max_batch || !m->dX)
            return fail(m ? m->ctx : nullptr, AXX_ERR_INVALID, "mlp train: arguments or no data");
        if (batch % kRowPad)
            return fail(m->ctx, AXX_ERR_INVALID, "mlp train: batch must be a multiple of %d", kRowPad);
        axx_ctx* ctx = m->ctx;
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        const Layer& f = m->L[0];
        const Layer& last = m->L.back();
        float* d_loss = nullptr;
        int rc0 = dalloc(ctx, &d_loss, (size_t)steps);
        if (rc0) return rc0;
        for (int s = 0; s < steps; ++s) {
            hipLaunchKernelGGL(gather_rows, dim3(batch), dim3(128), 0, ctx->stream, (const __bf16*)m->dX, f.inp,
                               (const float*)m->dT, last.out, m->n_data, batch, (uint32_t)(m->adam_t + 1),
                               (uint32_t)seed, (uint32_t)(seed >> 32), m->H[0], m->T);
            HIP_TRY(ctx, hipGetLastError());
            int rc = forward(m, batch, true);
            if (!rc) rc = backward(m, batch);
            if (!rc) rc = adam_step(m, lr);
            if (rc) {
                (void)hipFree(d_loss);
                return rc;
            }
            hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, ctx->stream, (const float*)m->loss_part,
                               kLossBlocks, 1.0f / ((float)batch * last.out), d_loss + s);
        }
        hipError_t e = hipStreamSynchronize(ctx->stream);
        if (e == hipSuccess && loss_trace)
            e = hipMemcpy(loss_trace, d_loss, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost);
        (void)hipFree(d_loss);
        if (e != hipSuccess) return fail(ctx, AXX_ERR_DEVICE, "mlp train: %s", hipGetErrorString(e));
        return AXX_OK;
    } catch (...) {
        return api_exception();
    }
}
