// mlp.hip -- BF16 MFMA multilayer perceptron (train + serve) for gfx950.
//
// GEMMs: 64 x 64 output tiles per 256-thread block (4 waves, each a 32 x 32
// sub-tile), K staged through LDS 32 at a time, v_mfma_f32_32x32x16_bf16
// with FP32 accumulation.  Operands are BF16 in HBM; either operand may be
// stored transposed, so forward (X W^T), input-gradient (dZ W) and
// weight-gradient (dZ^T A) products are the same kernel.  Epilogues are
// fused: bias + SiLU (forward, keeping the FP32 pre-activation for the
// backward pass), SiLU' (backward), a plain FP32 store, or an FP32 atomic add
// of a split-K partial (weight gradients: few output tiles, K = batch rows).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "mcsnow.h"

namespace mlp {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int BM = 64, BN = 64, BK = 32, LDSP = BK + 8;  // LDS row pitch (bf16)
constexpr int K0 = 32;                                    // padded input width

enum { EPI_F32 = 0, EPI_SILU = 1, EPI_DSILU = 2, EPI_SILU_NOZ = 3, EPI_ADD_F32 = 4, EPI_EXP = 5 };

__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }

struct GemmArgs {
    const bf16 *A;  // logical A [M x K]; stored [M][lda] or (A_T) [K][lda]
    const bf16 *B;  // logical B [N x K] (C = A B^T); stored [N][ldb] or (B_T) [K][ldb]
    int M, N, K, lda, ldb;
    const float *bias;  // [N] or null
    const float *oscale, *oshift, *ooff;  // EPI_EXP: C = 2^(oscale v + oshift) - ooff per column (log2 e folded in)
    float *C;           // EPI_F32: C [M][ldc]; EPI_SILU: Z [M][ldc]
    bf16 *Cb;           // EPI_SILU / EPI_SILU_NOZ: silu(z) [M][ldcb]; EPI_DSILU: dZ [M][ldcb]
    const float *Zin;   // EPI_DSILU: pre-activation [M][ldc]
    int ldc, ldcb;
    int kchunk;         // K per blockIdx.z slice (split-K; EPI_ADD_F32 adds the partial tiles)
};

template <bool A_T, bool B_T, int EPI>
__global__ void __launch_bounds__(256) k_gemm(const GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) bf16 As[BM][LDSP];
    __shared__ __attribute__((aligned(16))) bf16 Bs[BN][LDSP];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    f32x16 acc;
    for (int r = 0; r < 16; ++r) acc[r] = 0.f;
    const int kb = blockIdx.z * g.kchunk;
    const int ke = min(g.K, kb + g.kchunk);

    for (int k0 = kb; k0 < ke; k0 += BK) {
        // ---- A tile -> As[m][k]
        if (!A_T) {
            const int row = t >> 2, kc = (t & 3) * 8;
            const int m = m0 + row;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (m < g.M) v = *reinterpret_cast<const uint4 *>(g.A + (size_t)m * g.lda + k0 + kc);
            *reinterpret_cast<uint4 *>(&As[row][kc]) = v;
        } else {
            const int k = t >> 3, mc = (t & 7) * 8;
            const int m = m0 + mc;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (m + 8 <= g.M) {
                v = *reinterpret_cast<const uint4 *>(g.A + (size_t)(k0 + k) * g.lda + m);
            } else if (m < g.M) {
                bf16 tmp[8];
                for (int j = 0; j < 8; ++j) tmp[j] = (m + j < g.M) ? g.A[(size_t)(k0 + k) * g.lda + m + j] : (bf16)0.f;
                memcpy(&v, tmp, 16);
            }
            bf16 e[8];
            memcpy(e, &v, 16);
            for (int j = 0; j < 8; ++j) As[mc + j][k] = e[j];
        }
        // ---- B tile -> Bs[n][k]
        if (!B_T) {
            const int row = t >> 2, kc = (t & 3) * 8;
            const int n = n0 + row;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (n < g.N) v = *reinterpret_cast<const uint4 *>(g.B + (size_t)n * g.ldb + k0 + kc);
            *reinterpret_cast<uint4 *>(&Bs[row][kc]) = v;
        } else {
            const int k = t >> 3, nc = (t & 7) * 8;
            const int n = n0 + nc;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (n + 8 <= g.N) {
                v = *reinterpret_cast<const uint4 *>(g.B + (size_t)(k0 + k) * g.ldb + n);
            } else if (n < g.N) {
                bf16 tmp[8];
                for (int j = 0; j < 8; ++j) tmp[j] = (n + j < g.N) ? g.B[(size_t)(k0 + k) * g.ldb + n + j] : (bf16)0.f;
                memcpy(&v, tmp, 16);
            }
            bf16 e[8];
            memcpy(e, &v, 16);
            for (int j = 0; j < 8; ++j) Bs[nc + j][k] = e[j];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            // 32x32x16: lane holds row (lane % 32), k = kk + (lane / 32) * 8 .. +8
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[wm * 32 + (lane & 31)][kk + (lane >> 5) * 8]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 32 + (lane & 31)][kk + (lane >> 5) * 8]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // ---- epilogue: item r of lane -> row (r/4)*8 + (lane/32)*4 + r%4, col lane%32
    const int n = n0 + wn * 32 + (lane & 31);
    if (n >= g.N) return;
    const float bias = g.bias ? g.bias[n] : 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int m = m0 + wm * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
        if (m >= g.M) continue;
        const float v = acc[r] + bias;
        if (EPI == EPI_F32) {
            g.C[(size_t)m * g.ldc + n] = v;
        } else if (EPI == EPI_ADD_F32) {
            unsafeAtomicAdd(&g.C[(size_t)m * g.ldc + n], acc[r]);   // partial sum of one K slice
        } else if (EPI == EPI_SILU || EPI == EPI_SILU_NOZ) {
            if (EPI == EPI_SILU) g.C[(size_t)m * g.ldc + n] = v;
            g.Cb[(size_t)m * g.ldcb + n] = (bf16)(v * sigm(v));
        } else {  // EPI_DSILU: v is dA; dZ = dA * silu'(z)
            const float z = g.Zin[(size_t)m * g.ldc + n];
            const float s = sigm(z);
            g.Cb[(size_t)m * g.ldcb + n] = (bf16)(v * s * (1.f + z * (1.f - s)));
        }
    }
}


// Forward GEMM for large M (inference): C = A B^T with A [M][lda] and
// B [N][ldb] row-major BF16 (no transposes), 128 x 128 output tiles per
// 256-thread block, each wave a 64 x 64 quadrant = 2 x 2 MFMA 32x32x16
// accumulators; K staged through two LDS buffers 32 at a time, the next stage
// prefetched into registers while the current one feeds the MFMAs.
// K must be a multiple of 32 (padded layer widths are).  Epilogues: EPI_F32
// (bias, FP32 store, columns n < N), EPI_EXP (the same through the output
// transform exp(scale v + shift) - offset of log-space targets) or
// EPI_SILU_NOZ (bias + SiLU, BF16 store).
constexpr int FM = 128, FN = 128, FP = BK + 8;   // tile and LDS row pitch (bf16)

template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_fwd(const GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) bf16 As[2][FM][FP];
    __shared__ __attribute__((aligned(16))) bf16 Bs[2][FN][FP];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;        // quadrant of the 128 x 128 tile
    // XCD-aware tiles: workgroups go to the 8 XCDs round-robin by linear id,
    // each XCD with its own L2, so the column tiles of one row block are
    // given ids on the same XCD (else each of them fetches the block's rows
    // of A from HBM again)
    int bx = blockIdx.x, by = blockIdx.y;
    if ((gridDim.y & 7u) == 0u) {
        const int b = blockIdx.y * gridDim.x + blockIdx.x;
        const int s = b >> 3;
        by = (s / (int)gridDim.x) * 8 + (b & 7);
        bx = s % (int)gridDim.x;
    }
    const int m0 = by * FM, n0 = bx * FN;
    // each thread moves 2 x 16 B of A and 2 x 16 B of B per stage:
    // rows r = t / 4 and t / 4 + 64, k chunk (t % 4) * 8
    const int lr = t >> 2, lk = (t & 3) * 8;
    auto load = [&](int k0, uint4 (&ra)[2], uint4 (&rb)[2]) {
        for (int h = 0; h < 2; ++h) {
            const int m = m0 + lr + 64 * h, n = n0 + lr + 64 * h;
            ra[h] = (m < g.M) ? *reinterpret_cast<const uint4 *>(g.A + (size_t)m * g.lda + k0 + lk)
                              : make_uint4(0, 0, 0, 0);
            rb[h] = (n < g.N) ? *reinterpret_cast<const uint4 *>(g.B + (size_t)n * g.ldb + k0 + lk)
                              : make_uint4(0, 0, 0, 0);
        }
    };
    auto store = [&](int buf, const uint4 (&ra)[2], const uint4 (&rb)[2]) {
        for (int h = 0; h < 2; ++h) {
            *reinterpret_cast<uint4 *>(&As[buf][lr + 64 * h][lk]) = ra[h];
            *reinterpret_cast<uint4 *>(&Bs[buf][lr + 64 * h][lk]) = rb[h];
        }
    };
    f32x16 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
    uint4 ra[2], rb[2];
    load(0, ra, rb);
    store(0, ra, rb);
    __syncthreads();
    const int nst = g.K / BK;
    for (int st = 0; st < nst; ++st) {
        const int cur = st & 1;
        if (st + 1 < nst) load((st + 1) * BK, ra, rb);   // in flight during the MFMAs
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            const int kc = kk + (lane >> 5) * 8;
            const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(&As[cur][wm * 64 + (lane & 31)][kc]);
            const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(&As[cur][wm * 64 + 32 + (lane & 31)][kc]);
            const bf16x8 b0 = *reinterpret_cast<const bf16x8 *>(&Bs[cur][wn * 64 + (lane & 31)][kc]);
            const bf16x8 b1 = *reinterpret_cast<const bf16x8 *>(&Bs[cur][wn * 64 + 32 + (lane & 31)][kc]);
            acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc[1][1], 0, 0, 0);
        }
        if (st + 1 < nst) store(cur ^ 1, ra, rb);   // the other buffer: last read two stages ago
        __syncthreads();
    }
    // epilogue: item r of lane -> row (r/4)*8 + (lane/32)*4 + r%4, col lane%32 of a 32 x 32 tile
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int n = n0 + wn * 64 + j * 32 + (lane & 31);
        if (n >= g.N) continue;
        const float bias = g.bias ? g.bias[n] : 0.f;
        float osc = 1.f, osh = 0.f, oof = 0.f;
        if (EPI == EPI_EXP) { osc = g.oscale[n]; osh = g.oshift[n]; oof = g.ooff[n]; }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * 64 + i * 32 + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                if (m >= g.M) continue;
                const float v = acc[i][j][r] + bias;
                if (EPI == EPI_F32) g.C[(size_t)m * g.ldc + n] = v;
                else if (EPI == EPI_EXP) g.C[(size_t)m * g.ldc + n] = __builtin_amdgcn_exp2f(fmaf(osc, v, osh)) - oof;
                else g.Cb[(size_t)m * g.ldcb + n] = (bf16)(v * sigm(v));
            }
        }
    }
}

// ---------------------------------------------------------------- fused
// The whole network for 128 rows per block: input (K0 = 32 columns) -> L
// hidden layers of 256 (bias + SiLU, BF16 activations kept in LDS) -> output
// layer with the F32 or exp epilogue straight to HBM.  Weights (0.8 MB for
// the surrogate) are read from L2 as MFMA operands, no LDS staging; each of
// the 8 waves owns one column tile of 32 of a hidden layer (and every 8th
// of the output layer) and multiplies each weight operand into the block's
// four 32-row quarters, so the weight traffic per row is a quarter of that
// of 32-row blocks.  A hidden layer's tiles are computed into registers,
// then, after a barrier, written over its input (one activation buffer).
// Built for 2 waves per SIMD (up to 256 VGPRs: the four accumulators and
// the epilogue without spills).  Measured on MI355X for 1e6 queries: 2.2 ms
// (3.8 ms with the per-layer kernels; 2.3 ms with 64-row blocks, 3.4 ms
// with 32-row blocks).  Operands, k order and epilogues are those of
// k_gemm_fwd, so the results are the same.
constexpr int FUSED_ROWS = 128, FUSED_Q = FUSED_ROWS / 32, FUSED_WAVES = 8, FUSED_H = 256,
              FUSED_MAX_LAYERS = 16;
static_assert(FUSED_H / 32 == FUSED_WAVES, "one hidden column tile per wave");
struct FusedArgs {
    const bf16 *X;            // [rows][K0]
    int rows, L, n_out, np;
    const bf16 *wb;           // bf16 weights, layer l at woff[l], [prow][pcol]
    const float *p;           // fp32 parameters (biases at boff[l])
    size_t woff[FUSED_MAX_LAYERS + 1], boff[FUSED_MAX_LAYERS + 1];
    float *y;                 // [rows][n_out]
    const float *oscale, *oshift, *ooff;   // exp epilogue (nullptr: F32)
};

// acc[q] += A[32 q .. 32 q + 32)[KIN] (LDS, row pitch lda) . W[n0 .. n0 + 32)[KIN]^T
template <int KIN>
__device__ __forceinline__ void fused_tile(f32x16 (&acc)[FUSED_Q], const bf16 *A, int lda, const bf16 *W, int n0,
                                           int lane)
{
    const bf16 *wrow = W + (size_t)(n0 + (lane & 31)) * KIN + (lane >> 5) * 8;
    const bf16 *arow = A + (lane & 31) * lda + (lane >> 5) * 8;
    // weight operands in groups of up to 4 k-steps
    constexpr int NS = KIN / 16, GS = NS < 4 ? NS : 4;
#pragma unroll 1
    for (int g0 = 0; g0 < NS; g0 += GS) {
        bf16x8 b[GS];
#pragma unroll
        for (int i = 0; i < GS; ++i) b[i] = *reinterpret_cast<const bf16x8 *>(wrow + (g0 + i) * 16);
#pragma unroll
        for (int i = 0; i < GS; ++i) {
            const int k = (g0 + i) * 16;
#pragma unroll
            for (int q = 0; q < FUSED_Q; ++q) {
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(arow + 32 * q * lda + k);
                acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[i], acc[q], 0, 0, 0);
            }
        }
    }
}

__global__ void __launch_bounds__(FUSED_WAVES * 64) __attribute__((amdgpu_waves_per_eu(2)))
k_mlp_fused(const FusedArgs g)
{
    constexpr int P0 = K0 + 8, PH = FUSED_H + 8;   // LDS row pitches (bf16)
    __shared__ __attribute__((aligned(16))) bf16 X0[FUSED_ROWS][P0];
    __shared__ __attribute__((aligned(16))) bf16 Act[FUSED_ROWS][PH];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r0 = blockIdx.x * FUSED_ROWS;
    // input rows (zero beyond the last row)
    for (int i = t; i < FUSED_ROWS * K0 / 8; i += blockDim.x) {
        const int r = i / (K0 / 8), c8 = (i % (K0 / 8)) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (r0 + r < g.rows) v = *reinterpret_cast<const uint4 *>(g.X + (size_t)(r0 + r) * K0 + c8);
        *reinterpret_cast<uint4 *>(&X0[r][c8]) = v;
    }
    __syncthreads();
    const int n = wave * 32 + (lane & 31);
    for (int l = 0; l < g.L; ++l) {
        f32x16 acc[FUSED_Q];
#pragma unroll
        for (int q = 0; q < FUSED_Q; ++q)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;
        const bf16 *W = g.wb + g.woff[l];
        if (l == 0) fused_tile<K0>(acc, &X0[0][0], P0, W, wave * 32, lane);
        else fused_tile<FUSED_H>(acc, &Act[0][0], PH, W, wave * 32, lane);
        const float bn = g.p[g.boff[l] + n];
        __syncthreads();   // every wave has read this layer's input
#pragma unroll
        for (int q = 0; q < FUSED_Q; ++q) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = 32 * q + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                const float v = acc[q][r] + bn;
                Act[m][n] = (bf16)(v * sigm(v));
            }
        }
        __syncthreads();
    }
    // output layer: np / 32 column tiles over the waves
    const bf16 *W = g.wb + g.woff[g.L];
    const float *bias = g.p + g.boff[g.L];
    for (int nt = wave; nt < g.np / 32; nt += FUSED_WAVES) {
        f32x16 acc[FUSED_Q];
#pragma unroll
        for (int q = 0; q < FUSED_Q; ++q)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[q][r] = 0.f;
        fused_tile<FUSED_H>(acc, &Act[0][0], PH, W, nt * 32, lane);
        const int no = nt * 32 + (lane & 31);
        if (no >= g.n_out) continue;
        const float bo = bias[no];
        float osc = 1.f, osh = 0.f, oof = 0.f;
        if (g.oscale) { osc = g.oscale[no]; osh = g.oshift[no]; oof = g.ooff[no]; }
#pragma unroll
        for (int q = 0; q < FUSED_Q; ++q) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = r0 + 32 * q + (r >> 2) * 8 + (lane >> 5) * 4 + (r & 3);
                if (m >= g.rows) continue;
                const float v = acc[q][r] + bo;
                g.y[(size_t)m * g.n_out + no] = g.oscale ? __builtin_amdgcn_exp2f(fmaf(osc, v, osh)) - oof : v;
            }
        }
    }
}

// X rows (fp32, n_in wide) -> bf16 [rows][K0], zero padded
__global__ void k_cast_x(const float *x, int n_in, int rows, bf16 *X)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * K0) return;
    const int r = i / K0, c = i % K0;
    X[i] = (bf16)(c < n_in ? x[(size_t)r * n_in + c] : 0.f);
}

// Input features of the snowpack BRF surrogate from raw case inputs (r_eff
// [um], omega, g, theta0 [deg], tau), rows of 5 floats, FP64 arithmetic:
//   log r, log10(1 - omega), g, cos theta0, log10 tau, exp(-tau / mu0),
//   sqrt((1 - omega) / (1 - omega g)), log10(tau (1 - omega g)), log10(tau / mu0)
// each standardised as (f - mu_k) / sd_k, zero padded to K0 columns, bf16.
constexpr int SNOW_FEATURES = 9;
__global__ void k_features_snow(const float *in, int rows, const float *mu, const float *sd, bf16 *X)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float *q = in + (size_t)r * 5;
    const double reff = q[0], om = q[1], g = q[2], th = q[3], tau = q[4];
    const double mu0 = cos(th * (3.141592653589793 / 180.0));
    double f[SNOW_FEATURES];
    f[0] = log(reff);
    f[1] = log10(fmax(1.0 - om, 1e-12));
    f[2] = g;
    f[3] = mu0;
    f[4] = log10(tau);
    f[5] = exp(-tau / mu0);
    f[6] = sqrt((1.0 - om) / (1.0 - om * g));
    f[7] = log10(tau * (1.0 - om * g));
    f[8] = log10(tau / mu0);
    bf16 *row = X + (size_t)r * K0;
    for (int k = 0; k < K0; ++k)
        row[k] = (bf16)(k < SNOW_FEATURES ? (float)((f[k] - (double)mu[k]) / (double)sd[k]) : 0.f);
}

// minibatch gather from the resident dataset
__global__ void k_gather(const float *xd, const float *td, const int32_t *idx, int n_in, int n_out, int np,
                         int rows, int64_t n_data, bf16 *X, float *T)
{
    const int r = blockIdx.x;
    if (r >= rows) return;
    int64_t s = idx[r];
    if (s < 0 || s >= n_data) s = 0;
    for (int c = threadIdx.x; c < K0; c += blockDim.x)
        X[(size_t)r * K0 + c] = (bf16)(c < n_in ? xd[s * n_in + c] : 0.f);
    for (int c = threadIdx.x; c < np; c += blockDim.x)
        T[(size_t)r * np + c] = c < n_out ? td[s * n_out + c] : 0.f;
}

// dY = 2 w (y - t) / rows (bf16); part[block] = that block's share of
// sum w (y - t)^2.  Four columns per thread (np is a multiple of 64).
// k_loss_sum then adds the partials in block order (deterministic; no
// cross-block fence, which costs an L2 write-back on a multi-XCD part).
constexpr int LOSS_BLOCKS = 1024;
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_loss(const float *Y, const float *T, const float *w, int rows, int np,
                                              bf16 *dY, float *part)
{
    __shared__ float red[256];
    const float inv = 1.f / rows;
    const uint32_t np4 = (uint32_t)np >> 2, n4 = (uint32_t)rows * np4;
    float acc = 0.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += gridDim.x * blockDim.x) {
        const float4 y = reinterpret_cast<const float4 *>(Y)[i];
        const float4 t = reinterpret_cast<const float4 *>(T)[i];
        const float4 wc = reinterpret_cast<const float4 *>(w)[i % np4];
        const float d0 = y.x - t.x, d1 = y.y - t.y, d2 = y.z - t.z, d3 = y.w - t.w;
        acc += (wc.x * d0 * d0 + wc.y * d1 * d1) + (wc.z * d2 * d2 + wc.w * d3 * d3);
        bf16x4 o;
        o[0] = (bf16)(2.f * wc.x * d0 * inv);
        o[1] = (bf16)(2.f * wc.y * d1 * inv);
        o[2] = (bf16)(2.f * wc.z * d2 * inv);
        o[3] = (bf16)(2.f * wc.w * d3 * inv);
        reinterpret_cast<bf16x4 *>(dY)[i] = o;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// *loss_out = (sum of part[0..nb)) / rows, one block of 256.
__global__ void __launch_bounds__(256) k_loss_sum(const float *part, int nb, int rows, float *loss_out)
{
    __shared__ float red[256];
    float s0 = 0.f;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s0 += part[b];
    red[threadIdx.x] = s0;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss_out = red[0] / rows;
}

// db[c] += column sums of dZ (db zeroed first).  Block (64, 4): a wave per
// row phase, a lane per column; a block covers 64 columns x COLSUM_ROWS rows,
// four independent partial sums per lane, the four waves reduced through LDS
// and one atomic per column and block.
constexpr int COLSUM_ROWS = 128;
__global__ void __launch_bounds__(256) k_colsum(const bf16 *dZ, int rows, int cols, int ld, float *db)
{
    __shared__ float part[4][64];
    const int c = blockIdx.x * 64 + threadIdx.x;
    const int r0 = blockIdx.y * COLSUM_ROWS, r1 = min(rows, r0 + COLSUM_ROWS);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if (c < cols) {
        int r = r0 + threadIdx.y;
        for (; r + 12 < r1; r += 16) {
            s0 += (float)dZ[(size_t)r * ld + c];
            s1 += (float)dZ[(size_t)(r + 4) * ld + c];
            s2 += (float)dZ[(size_t)(r + 8) * ld + c];
            s3 += (float)dZ[(size_t)(r + 12) * ld + c];
        }
        for (; r < r1; r += 4) s0 += (float)dZ[(size_t)r * ld + c];
    }
    part[threadIdx.y][threadIdx.x] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (threadIdx.y == 0 && c < cols)
        unsafeAtomicAdd(&db[c], (part[0][threadIdx.x] + part[1][threadIdx.x]) + (part[2][threadIdx.x] + part[3][threadIdx.x]));
}

__global__ void k_fill_f32(float *p, size_t n, float v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Adam (decoupled weight decay) over the padded parameter buffer, refreshing
// the bf16 copies of weights; mask zeroes updates of padding.
__global__ void k_adam(float *p, const float *gr, float *m, float *v, const uint8_t *mask, size_t n,
                       float lr_t, float b1, float b2, float eps, float wd_lr, bf16 *pb, const int64_t *pb_index)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (!mask[i]) continue;
        const float g = gr[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        const float np_ = p[i] - lr_t * mi / (sqrtf(vi) + eps) - wd_lr * p[i];
        p[i] = np_;
        const int64_t j = pb_index[i];
        if (j >= 0) pb[j] = (bf16)np_;
    }
}


}  // namespace mlp

// =================================================================== host
using namespace mlp;

struct Layer {
    int rows, cols;         // logical (unpadded) shape
    int prow, pcol;         // padded device shape
    size_t w_off, b_off;    // into the padded fp32 buffer
    size_t wb_off;          // into the bf16 weight buffer
};

struct mcs_mlp {
    int device = 0;
    mcs_mlp_dims d{};
    int np = 0;  // padded outputs
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    char err[512] = {0};
    std::vector<Layer> layers;
    size_t n_pad = 0, n_flat = 0, n_wb = 0;
    float *p = nullptr, *gr = nullptr, *m = nullptr, *v = nullptr;
    uint8_t *mask = nullptr;
    int64_t *pb_index = nullptr;
    bf16 *wb = nullptr;
    // activations (max_batch rows)
    bf16 *X = nullptr, *dY = nullptr, *dZa = nullptr, *dZb = nullptr;
    std::vector<float *> Z;
    std::vector<bf16 *> A;
    float *Y = nullptr, *T = nullptr, *w = nullptr, *yout = nullptr, *xin = nullptr, *loss = nullptr;
    int32_t *idx = nullptr;
    float *lpart = nullptr;          // k_loss: per-block partial sums (LOSS_BLOCKS)
    bool fused = false;              // inference through k_mlp_fused
    float *xd = nullptr, *td = nullptr;
    int64_t n_data = 0;
    double last_ms = 0.0;
    // serving (mcs_mlp_set_serving): input feature map, its standardisation
    // (K0 floats each) and the output transform (n_out floats each)
    int32_t feature_map = MCS_FEATURES_IDENTITY;
    float *fmu = nullptr, *fsd = nullptr, *oscale = nullptr, *oshift = nullptr, *ooff = nullptr;
    bool serving = false;
};

static int merr(mcs_mlp *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define MCHK(ctx, expr)                                                                \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return merr(ctx, e_ == hipErrorOutOfMemory ? MCS_ERR_NOMEM : MCS_ERR_HIP,  \
                        "%s: %s", #expr, hipGetErrorString(e_));                       \
    } while (0)

template <class T>
static hipError_t dalloc(T **p, size_t count)
{
    return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1));
}

template <bool AT, bool BT, int EPI>
static void gemm(mcs_mlp *c, const GemmArgs &g0)
{
    GemmArgs g = g0;
    const int tiles = ((g.N + BN - 1) / BN) * ((g.M + BM - 1) / BM);
    int splits = 1;
    if (EPI == EPI_ADD_F32) {
        // long-K products with few output tiles (weight gradients): split K so
        // that ~512 blocks run, slices a multiple of BK
        splits = std::max(1, std::min(g.K / BK, (512 + tiles - 1) / tiles));
    }
    g.kchunk = ((g.K + splits - 1) / splits + BK - 1) / BK * BK;
    splits = (g.K + g.kchunk - 1) / g.kchunk;
    dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, splits);
    hipLaunchKernelGGL((k_gemm<AT, BT, EPI>), grid, dim3(256), 0, c->stream, g);
}

static GemmArgs ga(const bf16 *A, int lda, const bf16 *B, int ldb, int M, int N, int K)
{
    GemmArgs g;
    memset(&g, 0, sizeof g);
    g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.M = M; g.N = N; g.K = K;
    g.kchunk = K;
    return g;
}

template <int EPI>
static void gemm_fwd(mcs_mlp *c, const GemmArgs &g)
{
    dim3 grid((g.N + FN - 1) / FN, (g.M + FM - 1) / FM);
    hipLaunchKernelGGL((k_gemm_fwd<EPI>), grid, dim3(256), 0, c->stream, g);
}

// inference of `rows` rows already in c->X; outputs [rows][n_out] into `y`
// (through the output transform when `exp_out`)
static void infer_pass(mcs_mlp *c, int rows, float *y, bool exp_out = false)
{
    const int L = c->d.n_hidden_layers;
    if (c->fused) {
        FusedArgs f;
        memset(&f, 0, sizeof f);
        f.X = c->X; f.rows = rows; f.L = L; f.n_out = c->d.n_out; f.np = c->np;
        f.wb = c->wb; f.p = c->p; f.y = y;
        for (int l = 0; l <= L; ++l) { f.woff[l] = c->layers[l].wb_off; f.boff[l] = c->layers[l].b_off; }
        if (exp_out) { f.oscale = c->oscale; f.oshift = c->oshift; f.ooff = c->ooff; }
        hipLaunchKernelGGL(k_mlp_fused, dim3((rows + FUSED_ROWS - 1) / FUSED_ROWS), dim3(FUSED_WAVES * 64), 0, c->stream, f);
        return;
    }
    const bf16 *in = c->X;
    int ld_in = K0;
    for (int l = 0; l < L; ++l) {
        const Layer &ly = c->layers[l];
        GemmArgs g = ga(in, ld_in, c->wb + ly.wb_off, ly.pcol, rows, ly.prow, ly.pcol);
        g.bias = c->p + ly.b_off;
        g.Cb = c->A[l]; g.ldcb = ly.prow;
        gemm_fwd<EPI_SILU_NOZ>(c, g);
        in = c->A[l];
        ld_in = ly.prow;
    }
    const Layer &lo = c->layers[L];
    GemmArgs g = ga(in, ld_in, c->wb + lo.wb_off, lo.pcol, rows, c->d.n_out, lo.pcol);
    g.bias = c->p + lo.b_off;
    g.C = y; g.ldc = c->d.n_out;
    if (exp_out) {
        g.oscale = c->oscale; g.oshift = c->oshift; g.ooff = c->ooff;
        gemm_fwd<EPI_EXP>(c, g);
    } else {
        gemm_fwd<EPI_F32>(c, g);
    }
}

// training forward of `rows` rows already in c->X, keeping Z for backward
static void forward_pass(mcs_mlp *c, int rows)
{
    const int L = c->d.n_hidden_layers;
    const bf16 *in = c->X;
    int ld_in = K0;
    for (int l = 0; l < L; ++l) {
        const Layer &ly = c->layers[l];
        GemmArgs g = ga(in, ld_in, c->wb + ly.wb_off, ly.pcol, rows, ly.prow, ly.pcol);
        g.bias = c->p + ly.b_off;
        g.C = c->Z[l]; g.ldc = ly.prow;
        g.Cb = c->A[l]; g.ldcb = ly.prow;
        gemm<false, false, EPI_SILU>(c, g);
        in = c->A[l];
        ld_in = ly.prow;
    }
    const Layer &lo = c->layers[L];
    GemmArgs g = ga(in, ld_in, c->wb + lo.wb_off, lo.pcol, rows, lo.prow, lo.pcol);
    g.bias = c->p + lo.b_off;
    g.C = c->Y; g.ldc = c->np;
    gemm<false, false, EPI_F32>(c, g);
}

// loss (into *loss_out) + backward for `rows` rows (X, T in place);
// gradients into c->gr
static void backward_pass(mcs_mlp *c, int rows, float *loss_out)
{
    const int L = c->d.n_hidden_layers;
    // weight and bias gradients are sums of split-K / row-slice partials
    hipLaunchKernelGGL(k_fill_f32, dim3(256), dim3(256), 0, c->stream, c->gr, c->n_pad, 0.f);
    const int colsum_blocks_y = (rows + COLSUM_ROWS - 1) / COLSUM_ROWS;
    const int loss_blocks = (int)std::min<int64_t>(LOSS_BLOCKS, ((int64_t)rows * (c->np / 4) + 255) / 256);
    hipLaunchKernelGGL(k_loss, dim3(loss_blocks), dim3(256), 0, c->stream, c->Y, c->T, c->w, rows, c->np, c->dY,
                       c->lpart);
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, c->stream, c->lpart, loss_blocks, rows, loss_out);
    // output layer
    const Layer &lo = c->layers[L];
    const bf16 *Aprev = c->A[L - 1];
    {
        GemmArgs g = ga(c->dY, c->np, Aprev, lo.pcol, lo.prow, lo.pcol, rows);  // dW = dY^T A
        g.C = c->gr + lo.w_off; g.ldc = lo.pcol;
        gemm<true, true, EPI_ADD_F32>(c, g);
        hipLaunchKernelGGL(k_colsum, dim3((lo.prow + 63) / 64, colsum_blocks_y), dim3(64, 4), 0, c->stream, c->dY,
                           rows, lo.prow, c->np, c->gr + lo.b_off);
        GemmArgs h = ga(c->dY, c->np, c->wb + lo.wb_off, lo.pcol, rows, lo.pcol, lo.prow);  // dA = dY W
        h.Zin = c->Z[L - 1]; h.ldc = lo.pcol;
        h.Cb = c->dZa; h.ldcb = lo.pcol;
        gemm<false, true, EPI_DSILU>(c, h);
    }
    bf16 *dz = c->dZa, *dzn = c->dZb;
    for (int l = L - 1; l >= 0; --l) {
        const Layer &ly = c->layers[l];
        const bf16 *inp = (l == 0) ? c->X : c->A[l - 1];
        const int ldi = (l == 0) ? K0 : c->layers[l - 1].prow;
        GemmArgs g = ga(dz, ly.prow, inp, ldi, ly.prow, ly.pcol, rows);  // dW_l = dZ^T in
        g.C = c->gr + ly.w_off; g.ldc = ly.pcol;
        gemm<true, true, EPI_ADD_F32>(c, g);
        hipLaunchKernelGGL(k_colsum, dim3((ly.prow + 63) / 64, colsum_blocks_y), dim3(64, 4), 0, c->stream, dz,
                           rows, ly.prow, ly.prow, c->gr + ly.b_off);
        if (l > 0) {
            GemmArgs h = ga(dz, ly.prow, c->wb + ly.wb_off, ly.pcol, rows, ly.pcol, ly.prow);
            h.Zin = c->Z[l - 1]; h.ldc = ly.pcol;
            h.Cb = dzn; h.ldcb = ly.pcol;
            gemm<false, true, EPI_DSILU>(c, h);
            std::swap(dz, dzn);
        }
    }
}

static int check_dims(const mcs_mlp_dims *d)
{
    return d && d->n_in >= 1 && d->n_in <= K0 && d->hidden >= 64 && d->hidden % 64 == 0 && d->hidden <= 1024 &&
           d->n_hidden_layers >= 1 && d->n_hidden_layers <= 16 && d->n_out >= 1 && d->n_out <= 65536 &&
           d->max_batch >= 64 && d->max_batch % 64 == 0 && d->max_batch <= (1 << 20);
}

extern "C" {

int mcs_mlp_create(int device, const mcs_mlp_dims *dims, mcs_mlp **out)
{
    if (!out) return MCS_ERR_ARG;
    *out = nullptr;
    if (!check_dims(dims)) return MCS_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return MCS_ERR_HIP;
    if (device < 0 || device >= n) return MCS_ERR_ARG;
    mcs_mlp *c = new (std::nothrow) mcs_mlp();
    if (!c) return MCS_ERR_NOMEM;
    c->device = device;
    c->d = *dims;
    c->np = (dims->n_out + 63) / 64 * 64;
    // inference of 256-wide networks runs fused (k_mlp_fused); the
    // environment variable MCS_MLP_LAYERED=1 keeps the per-layer kernels
    // (the same results, for comparison)
    {
        const char *e = getenv("MCS_MLP_LAYERED");
        c->fused = dims->hidden == FUSED_H && dims->n_hidden_layers <= FUSED_MAX_LAYERS && !(e && e[0] == '1');
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->e0) != hipSuccess || hipEventCreate(&c->e1) != hipSuccess) {
        mcs_mlp_destroy(c);
        return MCS_ERR_HIP;
    }
    const int L = dims->n_hidden_layers, H = dims->hidden;
    size_t off = 0, woff = 0, flat = 0;
    for (int l = 0; l <= L; ++l) {
        Layer ly;
        ly.rows = (l == L) ? dims->n_out : H;
        ly.cols = (l == 0) ? dims->n_in : H;
        ly.prow = (l == L) ? c->np : H;
        ly.pcol = (l == 0) ? K0 : H;
        ly.w_off = off; off += (size_t)ly.prow * ly.pcol;
        ly.b_off = off; off += ly.prow;
        ly.wb_off = woff; woff += (size_t)ly.prow * ly.pcol;
        flat += (size_t)ly.rows * ly.cols + ly.rows;
        c->layers.push_back(ly);
    }
    c->n_pad = off; c->n_wb = woff; c->n_flat = flat;
    const size_t MB = dims->max_batch;
    std::vector<uint8_t> mask(off, 0);
    std::vector<int64_t> pbi(off, -1);
    for (const Layer &ly : c->layers) {
        for (int r = 0; r < ly.rows; ++r) {
            for (int k = 0; k < ly.cols; ++k) {
                mask[ly.w_off + (size_t)r * ly.pcol + k] = 1;
            }
            mask[ly.b_off + r] = 1;
        }
        for (int r = 0; r < ly.prow; ++r)
            for (int k = 0; k < ly.pcol; ++k) pbi[ly.w_off + (size_t)r * ly.pcol + k] = (int64_t)(ly.wb_off + (size_t)r * ly.pcol + k);
    }
    bool ok = dalloc(&c->p, off) == hipSuccess && dalloc(&c->gr, off) == hipSuccess &&
              dalloc(&c->m, off) == hipSuccess && dalloc(&c->v, off) == hipSuccess &&
              dalloc(&c->mask, off) == hipSuccess && dalloc(&c->pb_index, off) == hipSuccess &&
              dalloc(&c->wb, woff) == hipSuccess && dalloc(&c->X, MB * K0) == hipSuccess &&
              dalloc(&c->dY, MB * c->np) == hipSuccess && dalloc(&c->dZa, MB * H) == hipSuccess &&
              dalloc(&c->dZb, MB * H) == hipSuccess && dalloc(&c->Y, MB * c->np) == hipSuccess &&
              dalloc(&c->T, MB * c->np) == hipSuccess && dalloc(&c->w, c->np) == hipSuccess &&
              dalloc(&c->yout, MB * dims->n_out) == hipSuccess && dalloc(&c->xin, MB * dims->n_in) == hipSuccess &&
              dalloc(&c->loss, 1024) == hipSuccess && dalloc(&c->idx, MB) == hipSuccess &&
              dalloc(&c->lpart, LOSS_BLOCKS) == hipSuccess;
    for (int l = 0; ok && l < L; ++l) {
        float *z = nullptr;
        bf16 *a = nullptr;
        ok = dalloc(&z, MB * H) == hipSuccess && dalloc(&a, MB * H) == hipSuccess;
        c->Z.push_back(z);
        c->A.push_back(a);
    }
    if (!ok) {
        mcs_mlp_destroy(c);
        return MCS_ERR_NOMEM;
    }
    if (hipMemcpy(c->mask, mask.data(), off, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->pb_index, pbi.data(), off * sizeof(int64_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->p, 0, off * 4) != hipSuccess || hipMemset(c->m, 0, off * 4) != hipSuccess ||
        hipMemset(c->v, 0, off * 4) != hipSuccess || hipMemset(c->gr, 0, off * 4) != hipSuccess ||
        hipMemset(c->wb, 0, woff * 2) != hipSuccess) {
        mcs_mlp_destroy(c);
        return MCS_ERR_HIP;
    }
    *out = c;
    return MCS_OK;
}

int mcs_mlp_destroy(mcs_mlp *c)
{
    if (!c) return MCS_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    void *bufs[] = {c->p, c->gr, c->m, c->v, c->mask, c->pb_index, c->wb, c->X, c->dY, c->dZa, c->dZb,
                    c->Y, c->T, c->w, c->yout, c->xin, c->loss, c->idx, c->xd, c->td,
                    c->fmu, c->fsd, c->oscale, c->oshift, c->ooff, c->lpart};
    for (void *b : bufs) if (b) (void)hipFree(b);
    for (float *z : c->Z) if (z) (void)hipFree(z);
    for (bf16 *a : c->A) if (a) (void)hipFree(a);
    if (c->e0) (void)hipEventDestroy(c->e0);
    if (c->e1) (void)hipEventDestroy(c->e1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MCS_OK;
}

const char *mcs_mlp_last_error(const mcs_mlp *c) { return c ? c->err : "null context"; }

int mcs_mlp_param_count(const mcs_mlp *c, int64_t *out)
{
    if (!c || !out) return MCS_ERR_ARG;
    *out = (int64_t)c->n_flat;
    return MCS_OK;
}

int mcs_mlp_set_params(mcs_mlp *c, const float *params, int64_t n)
{
    if (!c || !params) return MCS_ERR_ARG;
    if (n != (int64_t)c->n_flat) return merr(c, MCS_ERR_ARG, "expected %zu parameters, got %lld", c->n_flat, (long long)n);
    for (int64_t i = 0; i < n; ++i)
        if (!std::isfinite(params[i])) return merr(c, MCS_ERR_ARG, "parameter %lld is not finite", (long long)i);
    (void)hipSetDevice(c->device);
    std::vector<float> pad(c->n_pad, 0.f);
    std::vector<bf16> wb(c->n_wb, (bf16)0.f);
    size_t k = 0;
    for (const Layer &ly : c->layers) {
        for (int r = 0; r < ly.rows; ++r)
            for (int q = 0; q < ly.cols; ++q) {
                pad[ly.w_off + (size_t)r * ly.pcol + q] = params[k];
                wb[ly.wb_off + (size_t)r * ly.pcol + q] = (bf16)params[k];
                ++k;
            }
        for (int r = 0; r < ly.rows; ++r) pad[ly.b_off + r] = params[k++];
    }
    MCHK(c, hipMemcpy(c->p, pad.data(), pad.size() * 4, hipMemcpyHostToDevice));
    MCHK(c, hipMemcpy(c->wb, wb.data(), wb.size() * 2, hipMemcpyHostToDevice));
    MCHK(c, hipMemset(c->m, 0, c->n_pad * 4));
    MCHK(c, hipMemset(c->v, 0, c->n_pad * 4));
    return MCS_OK;
}

static void unpad(const mcs_mlp *c, const std::vector<float> &pad, float *out)
{
    size_t k = 0;
    for (const Layer &ly : c->layers) {
        for (int r = 0; r < ly.rows; ++r)
            for (int q = 0; q < ly.cols; ++q) out[k++] = pad[ly.w_off + (size_t)r * ly.pcol + q];
        for (int r = 0; r < ly.rows; ++r) out[k++] = pad[ly.b_off + r];
    }
}

int mcs_mlp_get_params(mcs_mlp *c, float *params, int64_t n)
{
    if (!c || !params) return MCS_ERR_ARG;
    if (n != (int64_t)c->n_flat) return merr(c, MCS_ERR_ARG, "expected %zu parameters, got %lld", c->n_flat, (long long)n);
    (void)hipSetDevice(c->device);
    std::vector<float> pad(c->n_pad);
    MCHK(c, hipMemcpy(pad.data(), c->p, pad.size() * 4, hipMemcpyDeviceToHost));
    unpad(c, pad, params);
    return MCS_OK;
}

int mcs_mlp_forward(mcs_mlp *c, const float *x, int64_t n, float *y)
{
    if (!c || n < 0 || (n > 0 && (!x || !y))) return MCS_ERR_ARG;
    (void)hipSetDevice(c->device);
    const int MB = c->d.max_batch, ni = c->d.n_in, no = c->d.n_out;
    double total_ms = 0.0;
    for (int64_t s = 0; s < n; s += MB) {
        const int rows = (int)std::min<int64_t>(MB, n - s);
        MCHK(c, hipMemcpyAsync(c->xin, x + s * ni, sizeof(float) * rows * ni, hipMemcpyHostToDevice, c->stream));
        MCHK(c, hipEventRecord(c->e0, c->stream));
        hipLaunchKernelGGL(k_cast_x, dim3((rows * K0 + 255) / 256), dim3(256), 0, c->stream, c->xin, ni, rows, c->X);
        infer_pass(c, rows, c->yout);
        MCHK(c, hipEventRecord(c->e1, c->stream));
        MCHK(c, hipGetLastError());
        MCHK(c, hipMemcpyAsync(y + s * no, c->yout, sizeof(float) * rows * no, hipMemcpyDeviceToHost, c->stream));
        MCHK(c, hipStreamSynchronize(c->stream));
        float ms = 0.f;
        MCHK(c, hipEventElapsedTime(&ms, c->e0, c->e1));
        total_ms += ms;
    }
    c->last_ms = total_ms;
    return MCS_OK;
}

int mcs_mlp_set_serving(mcs_mlp *c, int32_t feature_map, const float *feat_mu, const float *feat_sd,
                        const float *out_scale, const float *out_shift, const float *out_offset)
{
    if (!c || !out_scale || !out_shift || !out_offset) return MCS_ERR_ARG;
    int nf = 0;
    if (feature_map == MCS_FEATURES_IDENTITY) {
        nf = c->d.n_in;
    } else if (feature_map == MCS_FEATURES_SNOW_BRF) {
        if (c->d.n_in != SNOW_FEATURES)
            return merr(c, MCS_ERR_ARG, "MCS_FEATURES_SNOW_BRF needs n_in = %d", SNOW_FEATURES);
        if (!feat_mu || !feat_sd) return merr(c, MCS_ERR_ARG, "feature standardisation required");
        nf = SNOW_FEATURES;
        for (int k = 0; k < nf; ++k)
            if (!std::isfinite(feat_mu[k]) || !(feat_sd[k] > 0) || !std::isfinite(feat_sd[k]))
                return merr(c, MCS_ERR_ARG, "feature %d: invalid standardisation", k);
    } else {
        return merr(c, MCS_ERR_ARG, "unknown feature map %d", feature_map);
    }
    const int no = c->d.n_out;
    for (int k = 0; k < no; ++k)
        if (!std::isfinite(out_scale[k]) || !std::isfinite(out_shift[k]) || !std::isfinite(out_offset[k]))
            return merr(c, MCS_ERR_ARG, "output %d: non-finite transform", k);
    (void)hipSetDevice(c->device);
    c->serving = false;
    if (!c->fmu) {
        MCHK(c, dalloc(&c->fmu, K0));
        MCHK(c, dalloc(&c->fsd, K0));
        MCHK(c, dalloc(&c->oscale, no));
        MCHK(c, dalloc(&c->oshift, no));
        MCHK(c, dalloc(&c->ooff, no));
    }
    std::vector<float> mu(K0, 0.f), sd(K0, 1.f);
    for (int k = 0; k < nf && feature_map != MCS_FEATURES_IDENTITY; ++k) { mu[k] = feat_mu[k]; sd[k] = feat_sd[k]; }
    MCHK(c, hipMemcpy(c->fmu, mu.data(), K0 * 4, hipMemcpyHostToDevice));
    MCHK(c, hipMemcpy(c->fsd, sd.data(), K0 * 4, hipMemcpyHostToDevice));
    // exp(s v + h) = 2^(s' v + h') with s' = s log2(e), h' = h log2(e): the
    // epilogue is one FMA and the hardware base-2 exponential
    std::vector<float> s2(no), h2(no);
    const double log2e = 1.0 / std::log(2.0);
    for (int k = 0; k < no; ++k) {
        s2[k] = (float)((double)out_scale[k] * log2e);
        h2[k] = (float)((double)out_shift[k] * log2e);
    }
    MCHK(c, hipMemcpy(c->oscale, s2.data(), (size_t)no * 4, hipMemcpyHostToDevice));
    MCHK(c, hipMemcpy(c->oshift, h2.data(), (size_t)no * 4, hipMemcpyHostToDevice));
    MCHK(c, hipMemcpy(c->ooff, out_offset, (size_t)no * 4, hipMemcpyHostToDevice));
    c->feature_map = feature_map;
    c->serving = true;
    return MCS_OK;
}

int mcs_mlp_serve(mcs_mlp *c, const float *d_in, int64_t n, float *d_out)
{
    if (!c || n < 0 || (n > 0 && (!d_in || !d_out))) return MCS_ERR_ARG;
    if (!c->serving) return merr(c, MCS_ERR_STATE, "mcs_mlp_set_serving must come first");
    (void)hipSetDevice(c->device);
    const int MB = c->d.max_batch, no = c->d.n_out;
    const int ni = c->feature_map == MCS_FEATURES_SNOW_BRF ? 5 : c->d.n_in;
    MCHK(c, hipEventRecord(c->e0, c->stream));
    for (int64_t s = 0; s < n; s += MB) {
        const int rows = (int)std::min<int64_t>(MB, n - s);
        if (c->feature_map == MCS_FEATURES_SNOW_BRF)
            hipLaunchKernelGGL(k_features_snow, dim3((rows + 255) / 256), dim3(256), 0, c->stream, d_in + s * ni, rows,
                               c->fmu, c->fsd, c->X);
        else
            hipLaunchKernelGGL(k_cast_x, dim3((rows * K0 + 255) / 256), dim3(256), 0, c->stream, d_in + s * ni, ni, rows,
                               c->X);
        infer_pass(c, rows, d_out + s * no, true);
    }
    MCHK(c, hipEventRecord(c->e1, c->stream));
    MCHK(c, hipGetLastError());
    MCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCHK(c, hipEventElapsedTime(&ms, c->e0, c->e1));
    c->last_ms = ms;
    return MCS_OK;
}

int mcs_mlp_last_ms(const mcs_mlp *c, double *ms)
{
    if (!c || !ms) return MCS_ERR_ARG;
    *ms = c->last_ms;
    return MCS_OK;
}

static int upload_weights(mcs_mlp *c, const float *w)
{
    std::vector<float> wp(c->np, 0.f);
    for (int k = 0; k < c->d.n_out; ++k) {
        if (!std::isfinite(w[k]) || w[k] < 0) return merr(c, MCS_ERR_ARG, "output weight %d invalid", k);
        wp[k] = w[k];
    }
    MCHK(c, hipMemcpyAsync(c->w, wp.data(), wp.size() * 4, hipMemcpyHostToDevice, c->stream));
    MCHK(c, hipStreamSynchronize(c->stream));
    return MCS_OK;
}

int mcs_mlp_grads(mcs_mlp *c, const float *x, const float *t, const float *w, int32_t n, float *grads, float *loss)
{
    if (!c || !x || !t || !w || !grads) return MCS_ERR_ARG;
    if (n <= 0 || n % 64 || n > c->d.max_batch) return merr(c, MCS_ERR_ARG, "n must be a positive multiple of 64 <= max_batch");
    (void)hipSetDevice(c->device);
    int r = upload_weights(c, w);
    if (r) return r;
    const int ni = c->d.n_in, no = c->d.n_out;
    std::vector<float> tp((size_t)n * c->np, 0.f);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < no; ++k) tp[(size_t)i * c->np + k] = t[(size_t)i * no + k];
    MCHK(c, hipMemcpy(c->xin, x, sizeof(float) * n * ni, hipMemcpyHostToDevice));
    MCHK(c, hipMemcpy(c->T, tp.data(), tp.size() * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_cast_x, dim3((n * K0 + 255) / 256), dim3(256), 0, c->stream, c->xin, ni, n, c->X);
    forward_pass(c, n);
    backward_pass(c, n, c->loss);
    MCHK(c, hipGetLastError());
    MCHK(c, hipStreamSynchronize(c->stream));
    std::vector<float> pad(c->n_pad);
    MCHK(c, hipMemcpy(pad.data(), c->gr, pad.size() * 4, hipMemcpyDeviceToHost));
    unpad(c, pad, grads);
    if (loss) MCHK(c, hipMemcpy(loss, c->loss, 4, hipMemcpyDeviceToHost));
    return MCS_OK;
}

int mcs_mlp_set_dataset(mcs_mlp *c, const float *x, const float *t, int64_t n)
{
    if (!c || !x || !t || n <= 0) return MCS_ERR_ARG;
    (void)hipSetDevice(c->device);
    if (c->xd) { (void)hipFree(c->xd); c->xd = nullptr; }
    if (c->td) { (void)hipFree(c->td); c->td = nullptr; }
    c->n_data = 0;
    MCHK(c, dalloc(&c->xd, (size_t)n * c->d.n_in));
    MCHK(c, dalloc(&c->td, (size_t)n * c->d.n_out));
    MCHK(c, hipMemcpy(c->xd, x, sizeof(float) * n * c->d.n_in, hipMemcpyHostToDevice));
    MCHK(c, hipMemcpy(c->td, t, sizeof(float) * n * c->d.n_out, hipMemcpyHostToDevice));
    c->n_data = n;
    return MCS_OK;
}

int mcs_mlp_train_steps(mcs_mlp *c, const int32_t *idx, int32_t n_steps, int32_t batch, const float *w,
                        const mcs_adam *a, float *losses)
{
    if (!c || !idx || !w || !a || n_steps <= 0) return MCS_ERR_ARG;
    if (batch <= 0 || batch % 64 || batch > c->d.max_batch) return merr(c, MCS_ERR_ARG, "batch must be a positive multiple of 64 <= max_batch");
    if (!c->xd) return merr(c, MCS_ERR_STATE, "mcs_mlp_set_dataset must come first");
    if (n_steps > 1000) return merr(c, MCS_ERR_ARG, "at most 1000 steps per call");
    if (!(a->lr >= 0) || !(a->beta1 >= 0 && a->beta1 < 1) || !(a->beta2 >= 0 && a->beta2 < 1) || !(a->eps > 0) || a->step < 0)
        return merr(c, MCS_ERR_ARG, "invalid Adam settings");
    for (int64_t i = 0; i < (int64_t)n_steps * batch; ++i)
        if (idx[i] < 0 || idx[i] >= c->n_data) return merr(c, MCS_ERR_ARG, "idx[%lld] out of range", (long long)i);
    (void)hipSetDevice(c->device);
    int r = upload_weights(c, w);
    if (r) return r;
    int32_t *didx = nullptr;
    MCHK(c, dalloc(&didx, (size_t)n_steps * batch));
    hipError_t e = hipMemcpy(didx, idx, sizeof(int32_t) * (size_t)n_steps * batch, hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(didx); return merr(c, MCS_ERR_HIP, "idx upload: %s", hipGetErrorString(e)); }
    for (int s = 0; s < n_steps; ++s) {
        hipLaunchKernelGGL(k_gather, dim3(batch), dim3(64), 0, c->stream, c->xd, c->td, didx + (size_t)s * batch,
                           c->d.n_in, c->d.n_out, c->np, batch, c->n_data, c->X, c->T);
        forward_pass(c, batch);
        backward_pass(c, batch, c->loss + 1 + s);   // step s's loss in loss[1 + s]
        const int t = a->step + s + 1;
        const float lr_t = a->lr * std::sqrt(1.f - std::pow(a->beta2, (float)t)) / (1.f - std::pow(a->beta1, (float)t));
        hipLaunchKernelGGL(k_adam, dim3(1024), dim3(256), 0, c->stream, c->p, c->gr, c->m, c->v, c->mask, c->n_pad,
                           lr_t, a->beta1, a->beta2, a->eps, a->lr * a->weight_decay, c->wb, c->pb_index);
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(didx);
    if (e != hipSuccess) return merr(c, MCS_ERR_HIP, "train: %s", hipGetErrorString(e));
    if (losses) MCHK(c, hipMemcpy(losses, c->loss + 1, sizeof(float) * n_steps, hipMemcpyDeviceToHost));
    return MCS_OK;
}

}  // extern "C"
