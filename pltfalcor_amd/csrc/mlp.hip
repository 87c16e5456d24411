// Learned BSDF surrogate (no counterpart in the reference): a 22-128-128-128-16 ReLU MLP whose
// forward pass, backward pass and Adam update run on CDNA4 matrix cores.
//
// Every product is one NT GEMM, C[M x N] = A[M x K] * Bt[N x K]^T, with both operands bf16 and
// K-contiguous so that each MFMA fragment (8 consecutive k of one row) is one 16 B load; the
// activations are kept in both sample-major [batch x features] (forward/backward-data operand) and
// feature-major [features x batch] (weight-gradient operand) layouts, written by the producing
// GEMM's epilogue. Accumulation is FP32 (v_mfma_f32_16x16x32_bf16); master weights and Adam
// moments are FP32; weight gradients are split-K partial sums reduced in a fixed order (the
// result does not depend on scheduling).
//
// Layers (parameter order of plt_surrogate_params): W1 [128 x 22], b1 [128], W2 [128 x 128], b2,
// W3 [128 x 128], b3, W4 [16 x 128], b4 [16] = 38,032 FP32 values. Inputs are padded to 32
// features with zeros (W1's padding columns stay zero).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>
#include "plt_internal.hpp"
#include "surrogate.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace plt {
namespace mlp {

constexpr int kIn = 22, kInPad = 32, kHid = 128, kOut = 16, kOutPad = 32;
constexpr int kParams = kHid * kIn + kHid + 2 * (kHid * kHid + kHid) + kOut * kHid + kOut;  // 38,032
static_assert(kParams == 38032, "parameter count");

// ---------------------------------------------------------------------------------------
// NT GEMM with fused epilogues.
// ---------------------------------------------------------------------------------------
enum Epilogue : int {
    kEpiF32 = 0,       // Cf[i, j] = acc (+ bias[j])
    kEpiBiasRelu = 1,  // h = max(acc + bias[j], 0): Cb[i, j] = bf16(h), Ct[j, i] = bf16(h) if Ct
    kEpiMask = 2,      // g = acc if H[i, j] > 0 else 0: Cb[i, j] = bf16(g), Ct[j, i] = bf16(g) if Ct
    kEpiPartial = 3,   // Cf[z][i, j] = acc over this split's K range (z = blockIdx.z)
};
struct GemmArgs {
    const __bf16* A;   // [M x K], row stride lda
    const __bf16* Bt;  // [N x K], row stride ldb
    int M, N, K, lda, ldb;
    int epi;
    const float* bias;
    float* Cf;
    __bf16* Cb;
    __bf16* Ct;
    const __bf16* H;
    int ldc, ldt, ldh;
    int k_per_split;  // K range of one split (multiple of 32)
    size_t split_stride;
    // kEpiMask: per-tile column sums of the bf16 outputs (the bias gradient's partials) into
    // bias_part[blockIdx.y * kBiasCols + bias_col + j] when set
    float* bias_part;
    int bias_col;
};
// Columns of the bias-gradient partials: b4 (16, from the loss kernel's 256-sample blocks), then
// b3, b2, b1 (128 each, from the backward GEMMs' 64-row tiles).
constexpr int kBiasCols = 16 + 3 * 128;

constexpr int kTile = 64, kKStep = 32, kLdsRow = kKStep + 8;  // +8 bf16 per row breaks bank conflicts
constexpr int kGemmThreads = 256;

__global__ void __launch_bounds__(kGemmThreads) gemm_nt_kernel(GemmArgs g) {
    __shared__ __bf16 As[kTile * kLdsRow];
    __shared__ __bf16 Bs[kTile * kLdsRow];
    __shared__ float colsum[2][kTile];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int kbeg = blockIdx.z * g.k_per_split;
    const int kend = min(g.K, kbeg + g.k_per_split);
    // staging: thread t copies 8 bf16 of row t/4, k offset 8*(t%4), of both tiles
    const int sr = t >> 2, sk = (t & 3) * 8;
    // wave w computes rows 32*(w/2)..+31 and columns 32*(w%2)..+31 of the tile: 2 x 2 MFMA blocks
    const int wr = 32 * (wave >> 1), wc = 32 * (wave & 1);
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = kbeg; k0 < kend; k0 += kKStep) {
        bf16x8 va = {}, vb = {};
        if (m0 + sr < g.M) va = *reinterpret_cast<const bf16x8*>(g.A + size_t(m0 + sr) * g.lda + k0 + sk);
        if (n0 + sr < g.N) vb = *reinterpret_cast<const bf16x8*>(g.Bt + size_t(n0 + sr) * g.ldb + k0 + sk);
        *reinterpret_cast<bf16x8*>(As + sr * kLdsRow + sk) = va;
        *reinterpret_cast<bf16x8*>(Bs + sr * kLdsRow + sk) = vb;
        __syncthreads();
        bf16x8 fa[2], fb[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) fa[a] = *reinterpret_cast<const bf16x8*>(As + (wr + 16 * a + fr) * kLdsRow + fk);
#pragma unroll
        for (int b = 0; b < 2; ++b) fb[b] = *reinterpret_cast<const bf16x8*>(Bs + (wc + 16 * b + fr) * kLdsRow + fk);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a], fb[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // D layout of a 16 x 16 block: lane holds rows 4*(lane/16) + r (r = 0..3) of column lane%16
    float cs[2] = {0.f, 0.f};  // this lane's share of its two columns' sums (kEpiMask)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = m0 + wr + 16 * a + 4 * (lane >> 4) + r;
                const int j = n0 + wc + 16 * b + (lane & 15);
                if (i >= g.M || j >= g.N) continue;
                float v = acc[a][b][r];
                switch (g.epi) {
                    case kEpiF32:
                        g.Cf[size_t(i) * g.ldc + j] = g.bias ? v + g.bias[j] : v;
                        break;
                    case kEpiBiasRelu: {
                        v = fmaxf(v + g.bias[j], 0.f);
                        const __bf16 h = static_cast<__bf16>(v);
                        g.Cb[size_t(i) * g.ldc + j] = h;
                        if (g.Ct) g.Ct[size_t(j) * g.ldt + i] = h;
                        break;
                    }
                    case kEpiMask: {
                        const float h = static_cast<float>(g.H[size_t(i) * g.ldh + j]);
                        const __bf16 d = static_cast<__bf16>(h > 0.f ? v : 0.f);
                        if (g.Cb) g.Cb[size_t(i) * g.ldc + j] = d;
                        if (g.Ct) g.Ct[size_t(j) * g.ldt + i] = d;
                        cs[b] += static_cast<float>(d);
                        break;
                    }
                    default:
                        g.Cf[blockIdx.z * g.split_stride + size_t(i) * g.ldc + j] = v;
                        break;
                }
            }
    if (g.epi == kEpiMask && g.bias_part) {  // block-uniform
        // the tile's column sums in a fixed order: lanes of one column (xor 16, 32), then the two
        // waves that share the column range
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cs[b] += __shfl_xor(cs[b], 16);
            cs[b] += __shfl_xor(cs[b], 32);
        }
        if (lane < 16) {
            colsum[wave >> 1][wc + lane] = cs[0];
            colsum[wave >> 1][wc + 16 + lane] = cs[1];
        }
        __syncthreads();
        if (t < kTile && n0 + t < g.N) g.bias_part[size_t(blockIdx.y) * kBiasCols + g.bias_col + n0 + t] = colsum[0][t] + colsum[1][t];
    }
}

int gemm(const GemmArgs& g, int splits, hipStream_t st) {
    if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.K % kKStep || g.lda % 8 || g.ldb % 8 || g.k_per_split % kKStep)
        return set_error(PLT_EINVAL, "mlp gemm: shapes must be multiples of 32 (K) and 8 (leading dimensions)");
    const dim3 grid((g.N + kTile - 1) / kTile, (g.M + kTile - 1) / kTile, splits);
    hipLaunchKernelGGL(gemm_nt_kernel, grid, dim3(kGemmThreads), 0, st, g);
    return check_launch("gemm_nt_kernel");
}

// ---------------------------------------------------------------------------------------
// Fused inference: all four layers in one persistent kernel. A workgroup of four waves loads
// the bf16 weights once into registers -- wave w holds the fragments of output features
// 32w..32w+31 of layers 1-3 and all of layer 4 -- and then walks 64-sample tiles. Each product
// is computed transposed, out[feature x sample] = W x in^T: the weights are the MFMA's A operand
// (registers), the tile's activations the B operand (LDS rows of 8 consecutive features), and the
// bias is the accumulator's initial value. The D layout then gives each lane four consecutive
// features of one sample, written to the next layer's LDS rows as one 8-byte store, and the
// output layer's four consecutive outputs of a sample are one Mueller row.
// ---------------------------------------------------------------------------------------
constexpr int kFTile = 64, kFThreads = 256;
constexpr int kXRow = kInPad + 8, kHRow = kHid + 8;  // padded LDS rows (bank spread, 16 B aligned)

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// ReLU of an MFMA result as one integer max on its bits: negative floats (sign bit set, -0
// included) read as negative integers and become +0; fmaxf would add a NaN canonicalisation.
__device__ __forceinline__ float relu_bits(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

// sinh for the decode: (e^y - e^-y) / 2 from one hardware exponential where |y| > 1/2, the Taylor
// series through y^9 below (relative error under 1e-8 there, so no cancellation).
__device__ __forceinline__ float sinh_fast(float y) {
    const float a = fabsf(y);
    if (a > 0.5f) {
        const float e = __expf(a);
        return copysignf(0.5f * (e - 1.f / e), y);
    }
    const float y2 = y * y;
    return y * (1.f + y2 * (1.f / 6.f + y2 * (1.f / 120.f + y2 * (1.f / 5040.f + y2 * (1.f / 362880.f)))));
}

// With Sout: instead of writing the raw outputs, each sample inside the training box
// (stats[32..54) / [54..76)) has its Mueller matrix decoded (M00 = exp(y0), Mk = M00 sinh(yk)
// after de-standardising with stats[0..16) / [16..32)) and applied to Sin; samples outside the box
// are left for the exact pass.
// With n_dev, the row count is min(*n_dev, n), read from device memory (a count a previous kernel
// on the stream produced), so the launch needs no host synchronisation.
__global__ void __launch_bounds__(kFThreads) mlp_infer_kernel(int n, const float* __restrict__ X, const __bf16* __restrict__ W1,
                                                              const __bf16* __restrict__ W2, const __bf16* __restrict__ W3,
                                                              const __bf16* __restrict__ W4, const float* __restrict__ params,
                                                              float* __restrict__ Y, const float* __restrict__ stats,
                                                              const float* __restrict__ Sin, float* __restrict__ Sout,
                                                              const uint32_t* __restrict__ n_dev) {
    if (n_dev) n = int(min(uint32_t(n), *n_dev));
    __shared__ __bf16 xs[kFTile * kXRow];
    __shared__ __bf16 ha[kFTile * kHRow];
    __shared__ __bf16 hb[kFTile * kHRow];
    __shared__ int outside[kFTile];  // sample outside the training box (apply mode)
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int fr = lane & 15, fg = lane >> 4, fk = 8 * fg;
    // register-resident A fragments: (feature block cb, k block kb) = W[feature][32 kb + fk .. +7]
    bf16x8 w1[2], w2[2][4], w3[2][4], w4[4];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
        const int col = 32 * w + 16 * cb + fr;
        w1[cb] = *reinterpret_cast<const bf16x8*>(W1 + size_t(col) * kInPad + fk);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            w2[cb][kb] = *reinterpret_cast<const bf16x8*>(W2 + size_t(col) * kHid + 32 * kb + fk);
            w3[cb][kb] = *reinterpret_cast<const bf16x8*>(W3 + size_t(col) * kHid + 32 * kb + fk);
        }
    }
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) w4[kb] = *reinterpret_cast<const bf16x8*>(W4 + size_t(fr) * kHid + 32 * kb + fk);
    // the biases of the lane's D rows: features 32w + 16cb + 4fg + r, outputs 4fg + r
    const float* b1 = params + kHid * kIn;
    const float* b2 = b1 + kHid + kHid * kHid;
    const float* b3 = b2 + kHid + kHid * kHid;
    const float* b4 = b3 + kHid + kOut * kHid;
    f32x4 c1[2], c2[2], c3[2], c4;
#pragma unroll
    for (int cb = 0; cb < 2; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int f = 32 * w + 16 * cb + 4 * fg + r;
            c1[cb][r] = b1[f];
            c2[cb][r] = b2[f];
            c3[cb][r] = b3[f];
        }
#pragma unroll
    for (int r = 0; r < 4; ++r) c4[r] = b4[4 * fg + r];
    // staging: thread t always loads input features 2p, 2p + 1 (p = t % 16) of rows t / 16 + 16 j
    // (j < 4) as one 8-byte load, so its box bounds are loaded once; the pair is one 4-byte LDS store
    const int sp = t & 15, sr = t >> 4;
    const bool inPair = 2 * sp < kIn;  // kIn is even: a pair is all in or all padding
    float blo[2] = {0.f, 0.f}, bhi[2] = {0.f, 0.f};
    if (Sout && inPair)
        for (int c = 0; c < 2; ++c) {
            blo[c] = stats[2 * kOut + 2 * sp + c];
            bhi[c] = stats[2 * kOut + kIn + 2 * sp + c];
        }
    const int ntiles = (n + kFTile - 1) / kFTile;
    // the thread's four input pairs of a tile, loaded back to back (the row index is clamped so every
    // load is in bounds; rows past n and padding features are zeroed when staged); the next
    // tile's are fetched while this one is computed
    const int lp = inPair ? sp : 0;
    float2 xv[kFTile / 16];
    auto fetch = [&](int tile) {
        const int r0 = tile * kFTile;
#pragma unroll
        for (int j = 0; j < kFTile / 16; ++j)
            xv[j] = *reinterpret_cast<const float2*>(X + size_t(min(r0 + sr + 16 * j, n - 1)) * kIn + 2 * lp);
    };
    if (int(blockIdx.x) < ntiles) fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int r0 = tile * kFTile;
#pragma unroll
        for (int j = 0; j < kFTile / 16; ++j) {
            const int r = sr + 16 * j;
            const bool live = r0 + r < n && inPair;
            const float v0 = live ? xv[j].x : 0.f, v1 = live ? xv[j].y : 0.f;
            bf16x2 pv;
            pv[0] = static_cast<__bf16>(v0);
            pv[1] = static_cast<__bf16>(v1);
            *reinterpret_cast<bf16x2*>(xs + r * kXRow + 2 * sp) = pv;
            if (Sout) {
                // a wave stages four rows (16 lanes each): a row is outside when any lane's pair is
                const bool out = live && !(v0 >= blo[0] && v0 <= bhi[0] && v1 >= blo[1] && v1 <= bhi[1]);
                const uint64_t m = __ballot(out);
                if ((lane & 15) == 0) outside[r] = ((m >> (lane & 48)) & 0xFFFFull) != 0 ? 1 : 0;
            }
        }
        __syncthreads();
        if (tile + int(gridDim.x) < ntiles) fetch(tile + gridDim.x);
        // hidden layers: wave w computes features 32w..32w+31 of all 64 samples
        auto hidden = [&](const __bf16* in, int in_row, int kblocks, const bf16x8 (*wf)[4], const bf16x8* wf1, const f32x4* bias,
                          __bf16* out) {
            f32x4 acc[2][4];
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int sb = 0; sb < 4; ++sb) acc[cb][sb] = bias[cb];
#pragma unroll
            for (int sb = 0; sb < 4; ++sb)
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) {
                    if (kb >= kblocks) break;
                    const bf16x8 bfrag = *reinterpret_cast<const bf16x8*>(in + (16 * sb + fr) * in_row + 32 * kb + fk);
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb) acc[cb][sb] = mfma(wf ? wf[cb][kb] : wf1[cb], bfrag, acc[cb][sb]);
                }
#pragma unroll
            for (int cb = 0; cb < 2; ++cb)
#pragma unroll
                for (int sb = 0; sb < 4; ++sb) {
                    bf16x4 h;
#pragma unroll
                    for (int r = 0; r < 4; ++r) h[r] = static_cast<__bf16>(relu_bits(acc[cb][sb][r]));
                    *reinterpret_cast<bf16x4*>(out + (16 * sb + fr) * kHRow + 32 * w + 16 * cb + 4 * fg) = h;
                }
        };
        hidden(xs, kXRow, 1, nullptr, w1, c1, ha);
        __syncthreads();
        hidden(ha, kHRow, 4, w2, nullptr, c2, hb);
        __syncthreads();
        hidden(hb, kHRow, 4, w3, nullptr, c3, ha);
        __syncthreads();
        // output layer: wave w computes samples 16w..16w+15, all 16 outputs; lane (fr, fg) holds
        // outputs 4fg..4fg+3 of sample 16w + fr: row fg of its Mueller matrix
        f32x4 acc = c4;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(w4[kb], *reinterpret_cast<const bf16x8*>(ha + (16 * w + fr) * kHRow + 32 * kb + fk), acc);
        const int srow = 16 * w + fr, row = r0 + srow;
        if (!Sout) {
            if (row < n) *reinterpret_cast<f32x4*>(Y + size_t(row) * kOut + 4 * fg) = acc;
        } else {
            // M00 = exp(y0) from the lane holding output 0 of the same sample (fg = 0)
            const float m00 = __shfl(__expf(acc[0] * stats[kOut] + stats[0]), fr);
            if (row < n && !outside[srow]) {
                const float* S = Sin + size_t(row) * 4;
                float acc4 = 0.f;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int k = 4 * fg + c;
                    const float v = k == 0 ? 1.f : sinh_fast(acc[c] * stats[kOut + k] + stats[k]);
                    acc4 += v * S[c];
                }
                Sout[size_t(row) * 4 + fg] = m00 * acc4;
            }
        }
        __syncthreads();  // the next tile overwrites xs, ha and outside
    }
}

// ---------------------------------------------------------------------------------------
// Element-wise kernels.
// ---------------------------------------------------------------------------------------
constexpr int kEw = 256;

// X f32 [n x 22] -> Xb bf16 [n x 32] (zero padded) and, if Xt, Xt [32 x ldt].
__global__ void pack_inputs_kernel(int n, const float* __restrict__ X, __bf16* __restrict__ Xb, __bf16* __restrict__ Xt, int ldt) {
    const int i = blockIdx.x * kEw + threadIdx.x;
    if (i >= n) return;
    for (int f = 0; f < kInPad; ++f) {
        const __bf16 v = static_cast<__bf16>(f < kIn ? X[size_t(i) * kIn + f] : 0.f);
        Xb[size_t(i) * kInPad + f] = v;
        if (Xt) Xt[size_t(f) * ldt + i] = v;
    }
}

// Mean-squared error over the 16 outputs: dL/dY = 2 (Y - T) / (n * 16), stored as bf16 in dY
// [n x 32] (zero padded) and dYt [16 x ldt]; per-block partial sums of the loss in FP32.
// The block's column sums of dY (the last layer's bias gradient) go to bias_part[blockIdx.x *
// kBiasCols + f] in a fixed order (lanes of a wave by xor steps, then the four waves).
__global__ void loss_grad_kernel(int n, const float* __restrict__ Y, const float* __restrict__ T, __bf16* __restrict__ dY,
                                 __bf16* __restrict__ dYt, int ldt, float* __restrict__ loss_part, float* __restrict__ bias_part) {
    __shared__ float red[kEw];
    __shared__ float colsum[kEw / 64][kOut];
    const int i = blockIdx.x * kEw + threadIdx.x;
    float l = 0.f;
    float gs[kOut];
#pragma unroll
    for (int f = 0; f < kOut; ++f) gs[f] = 0.f;
    if (i < n) {
        const float s = 2.f / (float(n) * float(kOut));
#pragma unroll
        for (int f = 0; f < kOutPad; ++f) {
            float g = 0.f;
            if (f < kOut) {
                const float d = Y[size_t(i) * kOut + f] - T[size_t(i) * kOut + f];
                l += d * d;
                g = s * d;
            }
            const __bf16 gb = static_cast<__bf16>(g);
            dY[size_t(i) * kOutPad + f] = gb;
            if (f < kOut) {
                dYt[size_t(f) * ldt + i] = gb;
                gs[f] = static_cast<float>(gb);
            }
        }
    }
    red[threadIdx.x] = l;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int f = 0; f < kOut; ++f) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) gs[f] += __shfl_xor(gs[f], o);
    }
    if (lane == 0) {
#pragma unroll
        for (int f = 0; f < kOut; ++f) colsum[wave][f] = gs[f];
    }
    __syncthreads();
    for (int s = kEw / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
    if (bias_part && threadIdx.x < kOut) {
        float c = 0.f;
        for (int w = 0; w < kEw / 64; ++w) c += colsum[w][threadIdx.x];
        bias_part[size_t(blockIdx.x) * kBiasCols + threadIdx.x] = c;
    }
}

// bf16 operand copies of one layer: W [out x in_pad] and Wt [in x out_pad] from the master
// weights [out x in] (padding zero).
__global__ void cast_weights_kernel(int out, int in, int in_pad, int out_pad, const float* __restrict__ Wm, __bf16* __restrict__ W,
                                    __bf16* __restrict__ Wt) {
    const int e = blockIdx.x * kEw + threadIdx.x;
    if (e >= out_pad * in_pad) return;
    const int o = e / in_pad, i = e % in_pad;
    const __bf16 v = static_cast<__bf16>(o < out && i < in ? Wm[size_t(o) * in + i] : 0.f);
    if (o < out) W[size_t(o) * in_pad + i] = v;
    if (i < in) Wt[size_t(i) * out_pad + o] = v;
}

}  // namespace mlp
}  // namespace plt

using namespace plt;
using namespace plt::mlp;



namespace {

// Offsets of each tensor in the public parameter order.
constexpr size_t kOffW1 = 0, kOffB1 = kOffW1 + size_t(kHid) * kIn, kOffW2 = kOffB1 + kHid, kOffB2 = kOffW2 + size_t(kHid) * kHid,
                 kOffW3 = kOffB2 + kHid, kOffB3 = kOffW3 + size_t(kHid) * kHid, kOffW4 = kOffB3 + kHid,
                 kOffB4 = kOffW4 + size_t(kOut) * kHid;
constexpr int kSplitTarget = 64;  // weight-gradient GEMM splits of the batch (K) dimension
constexpr size_t kTrainStateWords = 16;

template <class T>
int dev_alloc(plt_surrogate* s, T** p, size_t count) {
    void* q = nullptr;
    if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return set_error(PLT_ENOMEM, "plt_surrogate: device allocation failed");
    s->owned.push_back(q);
    *p = static_cast<T*>(q);
    return PLT_OK;
}

int cast_all(plt_surrogate* s, hipStream_t st) {
    struct L { size_t off; int out, in, in_pad, out_pad; __bf16 *W, *Wt; };
    const L layers[4] = {{kOffW1, kHid, kIn, kInPad, kHid, s->W1, s->W1t},
                         {kOffW2, kHid, kHid, kHid, kHid, s->W2, s->W2t},
                         {kOffW3, kHid, kHid, kHid, kHid, s->W3, s->W3t},
                         {kOffW4, kOut, kHid, kHid, kOutPad, s->W4, s->W4t}};
    for (const L& l : layers) {
        const int count = l.out_pad * l.in_pad;
        hipLaunchKernelGGL(cast_weights_kernel, dim3((count + kEw - 1) / kEw), dim3(kEw), 0, st, l.out, l.in, l.in_pad, l.out_pad,
                           s->params + l.off, l.W, l.Wt);
    }
    return check_launch("cast_weights_kernel");
}

GemmArgs nt(const __bf16* A, int lda, const __bf16* Bt, int ldb, int M, int N, int K) {
    GemmArgs g{};
    g.A = A; g.lda = lda; g.Bt = Bt; g.ldb = ldb; g.M = M; g.N = N; g.K = K;
    g.k_per_split = K;
    return g;
}

// Forward pass of n <= max_batch samples already packed in Xb; keeps H1..H3 (and their
// feature-major copies when `keep_t`) and writes the raw outputs to Y [n x 16] (FP32).
int forward(plt_surrogate* s, int n, bool keep_t, hipStream_t st) {
    const int ld = int(s->max_batch);
    int rc;
    GemmArgs g = nt(s->Xb, kInPad, s->W1, kInPad, n, kHid, kInPad);
    g.epi = kEpiBiasRelu; g.bias = s->params + kOffB1; g.Cb = s->H1; g.ldc = kHid; g.Ct = keep_t ? s->H1t : nullptr; g.ldt = ld;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    g = nt(s->H1, kHid, s->W2, kHid, n, kHid, kHid);
    g.epi = kEpiBiasRelu; g.bias = s->params + kOffB2; g.Cb = s->H2; g.ldc = kHid; g.Ct = keep_t ? s->H2t : nullptr; g.ldt = ld;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    g = nt(s->H2, kHid, s->W3, kHid, n, kHid, kHid);
    g.epi = kEpiBiasRelu; g.bias = s->params + kOffB3; g.Cb = s->H3; g.ldc = kHid; g.Ct = keep_t ? s->H3t : nullptr; g.ldt = ld;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    g = nt(s->H3, kHid, s->W4, kHid, n, kOut, kHid);
    g.epi = kEpiF32; g.bias = s->params + kOffB4; g.Cf = s->Y; g.ldc = kOut;
    return gemm(g, 1, st);
}

// The weight gradients' split-K partials: region r of s->part holds splits x (out x in) floats of
// layer 4, 3, 2, 1 in that order (layer 1 with its 32 padded inputs).
constexpr int kGradCount[4] = {kOut * kHid, kHid * kHid, kHid * kHid, kHid * kInPad};
constexpr int kGradTotal = kOut * kHid + 2 * kHid * kHid + kHid * kInPad;  // 38,912
int split_len(int n) { return (n / kKStep + kSplitTarget - 1) / kSplitTarget * kKStep; }

// dW[out x in] = sum over the batch of Gt[out, b] * Ht[in, b] (n a multiple of 32), as split-K
// partials into the layer's region of s->part; grad_reduce_kernel sums them.
int weight_grad(plt_surrogate* s, int n, const __bf16* Gt, int out, const __bf16* Ht, int in, size_t region, hipStream_t st) {
    const int ld = int(s->max_batch);
    const int per = split_len(n);
    const int splits = (n + per - 1) / per;
    GemmArgs g = nt(Gt, ld, Ht, ld, out, in, n);
    g.epi = kEpiPartial; g.Cf = s->part + region * kSplitTarget; g.ldc = in; g.k_per_split = per; g.split_stride = size_t(out) * in;
    return gemm(g, splits, st);
}

// Every gradient from its partials, in one launch. A block takes 64 elements of the
// [weights..., biases] list with four threads each: thread group g sums the element's partials
// z = g, g + 4, ... (all loads issued before the adds), and the four sums are combined in order.
// Weight elements sum their `splits` split-K partials; the kBiasCols bias elements sum their
// `rows_out` (b4) or `rows_hid` (b3, b2, b1) tile partials. Layer 1 drops its padding.
constexpr int kReduceLanes = 64, kReduceMax = 16;  // elements per block; partials per thread at most
__global__ void __launch_bounds__(256) grad_reduce_kernel(int splits, int rows_out, int rows_hid, const float* __restrict__ part,
                                                          const float* __restrict__ bias_part, float* __restrict__ grads) {
    __shared__ float sums[4][kReduceLanes];
    const int el = threadIdx.x & (kReduceLanes - 1), g = threadIdx.x >> 6;
    const int e = blockIdx.x * kReduceLanes + el;
    const bool live = e < kGradTotal + kBiasCols;
    const bool bias = e >= kGradTotal;
    int count = splits, cnt = 0, local = 0, c = 0;
    size_t base = 0;
    if (live && bias) {
        c = e - kGradTotal;
        count = c < kOut ? rows_out : rows_hid;
    } else if (live) {
        constexpr int c4 = kOut * kHid, c3 = kHid * kHid, c2 = kHid * kHid;
        const int r = e < c4 ? 0 : e < c4 + c3 ? 1 : e < c4 + c3 + c2 ? 2 : 3;
        const int first = r == 0 ? 0 : r == 1 ? c4 : r == 2 ? c4 + c3 : c4 + c3 + c2;
        cnt = r == 0 ? c4 : r == 1 ? c3 : r == 2 ? c2 : kHid * kInPad;
        local = e - first;
        base = size_t(first) * kSplitTarget;
    }
    float acc = 0.f;
    if (live) {
        // up to kReduceMax partials per thread in one batch of loads (weight partials: splits <= 64)
        for (int z0 = g; z0 < count; z0 += 4 * kReduceMax) {
            float v[kReduceMax];
#pragma unroll
            for (int q = 0; q < kReduceMax; ++q) {
                const int z = z0 + 4 * q;
                v[q] = z < count ? (bias ? bias_part[size_t(z) * kBiasCols + c] : part[base + size_t(z) * cnt + local]) : 0.f;
            }
#pragma unroll
            for (int q = 0; q < kReduceMax; ++q) acc += v[q];
        }
    }
    sums[g][el] = acc;
    __syncthreads();
    if (g != 0 || !live) return;
    const float sum = (sums[0][el] + sums[1][el]) + (sums[2][el] + sums[3][el]);
    if (bias) {
        const size_t dst = c < kOut ? kOffB4 + c : c < kOut + kHid ? kOffB3 + (c - kOut) : c < kOut + 2 * kHid ? kOffB2 + (c - kOut - kHid)
                                                                                           : kOffB1 + (c - kOut - 2 * kHid);
        grads[dst] = sum;
    } else if (cnt == kHid * kInPad) {
        if (local % kInPad < kIn) grads[kOffW1 + (local / kInPad) * kIn + local % kInPad] = sum;
    } else {
        grads[(cnt == kOut * kHid ? kOffW4 : e < kOut * kHid + kHid * kHid ? kOffW3 : kOffW2) + local] = sum;
    }
}

// Loss and gradients of n samples (X packed, T [n x 16] FP32 targets) into s->grads.
int loss_and_grads(plt_surrogate* s, int n, const float* T, float* loss_host, hipStream_t st) {
    const int ld = int(s->max_batch);
    int rc;
    if ((rc = forward(s, n, true, st)) < 0) return rc;
    const int blocks = (n + kEw - 1) / kEw;
    hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks), dim3(kEw), 0, st, n, s->Y, T, s->dY, s->dYt, ld, s->loss_part, s->bias_part);
    if ((rc = check_launch("loss_grad_kernel")) < 0) return rc;
    // backward data: G3 = (dY W4) * [H3 > 0], G2 = (G3 W3) * [H2 > 0], G1 = (G2 W2) * [H1 > 0]; each
    // epilogue also writes its tiles' column sums (the bias gradients' partials)
    GemmArgs g = nt(s->dY, kOutPad, s->W4t, kOutPad, n, kHid, kOutPad);
    g.epi = kEpiMask; g.H = s->H3; g.ldh = kHid; g.Cb = s->G3; g.ldc = kHid; g.Ct = s->G3t; g.ldt = ld;
    g.bias_part = s->bias_part; g.bias_col = kOut;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    g = nt(s->G3, kHid, s->W3t, kHid, n, kHid, kHid);
    g.epi = kEpiMask; g.H = s->H2; g.ldh = kHid; g.Cb = s->G2; g.ldc = kHid; g.Ct = s->G2t; g.ldt = ld;
    g.bias_part = s->bias_part; g.bias_col = kOut + kHid;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    g = nt(s->G2, kHid, s->W2t, kHid, n, kHid, kHid);
    g.epi = kEpiMask; g.H = s->H1; g.ldh = kHid; g.Cb = nullptr; g.Ct = s->G1t; g.ldt = ld;
    g.bias_part = s->bias_part; g.bias_col = kOut + 2 * kHid;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    // weight gradients (the batch is the K dimension), each layer's partials in its own region
    const size_t r3 = kGradCount[0], r2 = r3 + kGradCount[1], r1 = r2 + kGradCount[2];
    if ((rc = weight_grad(s, n, s->dYt, kOut, s->H3t, kHid, 0, st)) < 0) return rc;
    if ((rc = weight_grad(s, n, s->G3t, kHid, s->H2t, kHid, r3, st)) < 0) return rc;
    if ((rc = weight_grad(s, n, s->G2t, kHid, s->H1t, kHid, r2, st)) < 0) return rc;
    // W1's product is [128 x 32]; the reduction keeps the 22 real inputs
    if ((rc = weight_grad(s, n, s->G1t, kHid, s->Xt, kInPad, r1, st)) < 0) return rc;
    const int splits = (n + split_len(n) - 1) / split_len(n);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((kGradTotal + kBiasCols + kReduceLanes - 1) / kReduceLanes), dim3(256), 0, st, splits, blocks,
                       (n + kTile - 1) / kTile, s->part, s->bias_part, s->grads);
    if ((rc = check_launch("grad_reduce_kernel")) < 0) return rc;
    if (loss_host) {
        std::vector<float> lp(static_cast<size_t>(blocks));
        if (hipMemcpyAsync(lp.data(), s->loss_part, lp.size() * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return set_error(PLT_EHIP, "plt_surrogate: loss read-back failed");
        double sum = 0.0;
        for (float v : lp) sum += v;
        *loss_host = float(sum / (double(n) * kOut));
    }
    return PLT_OK;
}

// ---------------------------------------------------------------------------------------
// Training data pipeline and device-driven training (plt_surrogate_prepare / plt_surrogate_train).
// ---------------------------------------------------------------------------------------
// Loss weights per output, applied through the standardisation (an output scaled by 1/s weighs
// s^2 in the mean-squared error): the radiance row (ln M00, then M01..M03) first.
__constant__ float kOutputScale[kOut] = {4.f, 2.f, 2.f, 2.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f};

// The training targets of one exact Mueller matrix: ln M00 and asinh(Mk / M00), in double; a row
// is valid when every entry is finite and M00 > 1e-12.
__device__ bool targets_of(const float* __restrict__ Mrow, double* y) {
    bool ok = Mrow[0] > 1e-12f;
#pragma unroll
    for (int k = 0; k < kOut; ++k) ok = ok && isfinite(Mrow[k]);
    const double m00 = ok ? double(Mrow[0]) : 1.0;
    y[0] = log(m00);
#pragma unroll
    for (int k = 1; k < kOut; ++k) y[k] = asinh(double(Mrow[k]) / m00);
    return ok;
}

// Pass 1, per block of kEw rows: valid count, column sums of the targets (double), and the box of
// the valid rows' inputs. Sums run over the block's rows in order (deterministic).
__global__ void __launch_bounds__(kEw) prep_sums_kernel(int n, const float* __restrict__ X, const float* __restrict__ M,
                                                        double* __restrict__ psum, float* __restrict__ plo, float* __restrict__ phi,
                                                        uint32_t* __restrict__ pcount) {
    __shared__ double ys[kEw][kOut];
    __shared__ uint8_t valid[kEw];
    const int t = threadIdx.x, i = blockIdx.x * kEw + t;
    double y[kOut];
    bool ok = false;
    if (i < n) ok = targets_of(M + size_t(i) * kOut, y);
#pragma unroll
    for (int k = 0; k < kOut; ++k) ys[t][k] = ok ? y[k] : 0.0;
    valid[t] = ok ? 1 : 0;
    __syncthreads();
    const int rows = min(kEw, n - int(blockIdx.x) * kEw);
    if (t < kOut) {
        double acc = 0.0;
        for (int r = 0; r < rows; ++r) acc += ys[r][t];
        psum[size_t(blockIdx.x) * kOut + t] = acc;
    } else if (t >= 32 && t < 32 + kIn) {
        const int f = t - 32;
        float lo = __builtin_inff(), hi = -__builtin_inff();
        for (int r = 0; r < rows; ++r)
            if (valid[r]) {
                const float v = X[(size_t(blockIdx.x) * kEw + r) * kIn + f];
                lo = fminf(lo, v);
                hi = fmaxf(hi, v);
            }
        plo[size_t(blockIdx.x) * kIn + f] = lo;
        phi[size_t(blockIdx.x) * kIn + f] = hi;
    } else if (t == 64) {
        uint32_t c = 0;
        for (int r = 0; r < rows; ++r) c += valid[r];
        pcount[blockIdx.x] = c;
    }
}

// One block of 1024 threads: the means, the box (into stats), the total valid count, and each
// block's first slot in the compacted index list. Every sum splits the blocks into contiguous
// chunks, one per thread, and combines the chunk results in chunk order (deterministic).
constexpr int kPrepThreads = 1024;
__global__ void __launch_bounds__(kPrepThreads) prep_mean_kernel(int blocks, const double* __restrict__ psum, const float* __restrict__ plo,
                                                                 const float* __restrict__ phi, uint32_t* __restrict__ pcount,
                                                                 double* __restrict__ dstat, float* __restrict__ stats,
                                                                 uint32_t* __restrict__ n_valid) {
    __shared__ uint32_t chunk_count[kPrepThreads];
    __shared__ double col[kPrepThreads / kOut][kOut];
    __shared__ float lo_s[kPrepThreads / 32][kIn], hi_s[kPrepThreads / 32][kIn];
    const int t = threadIdx.x;
    // counts: chunk t of the blocks, then an exclusive scan over the chunks
    const int per = (blocks + kPrepThreads - 1) / kPrepThreads;
    const int b0 = min(blocks, t * per), b1 = min(blocks, b0 + per);
    uint32_t c = 0;
    for (int b = b0; b < b1; ++b) c += pcount[b];
    chunk_count[t] = c;
    // column sums: column t % 16 over chunk t / 16
    {
        const int cc = t % kOut, q = t / kOut, nq = kPrepThreads / kOut, pq = (blocks + nq - 1) / nq;
        double acc = 0.0;
        for (int b = q * pq; b < min(blocks, (q + 1) * pq); ++b) acc += psum[size_t(b) * kOut + cc];
        col[q][cc] = acc;
    }
    // box: feature t % 32 (< 22) over chunk t / 32
    {
        const int f = t % 32, q = t / 32, nq = kPrepThreads / 32, pq = (blocks + nq - 1) / nq;
        if (f < kIn) {
            float lo = __builtin_inff(), hi = -__builtin_inff();
            for (int b = q * pq; b < min(blocks, (q + 1) * pq); ++b) {
                lo = fminf(lo, plo[size_t(b) * kIn + f]);
                hi = fmaxf(hi, phi[size_t(b) * kIn + f]);
            }
            lo_s[q][f] = lo;
            hi_s[q][f] = hi;
        }
    }
    __syncthreads();
    __shared__ uint32_t total;
    if (t == 0) {
        uint32_t run = 0;
        for (int q = 0; q < kPrepThreads; ++q) {
            const uint32_t x = chunk_count[q];
            chunk_count[q] = run;
            run += x;
        }
        total = run;
        *n_valid = run;
    }
    __syncthreads();
    {
        uint32_t run = chunk_count[t];
        for (int b = b0; b < b1; ++b) {
            const uint32_t x = pcount[b];
            pcount[b] = run;
            run += x;
        }
    }
    if (t < kOut) {
        double acc = 0.0;
        for (int q = 0; q < kPrepThreads / kOut; ++q) acc += col[q][t];
        const double mean = total ? acc / double(total) : 0.0;
        dstat[t] = mean;
        stats[t] = float(mean);
    } else if (t >= 32 && t < 32 + kIn) {
        const int f = t - 32;
        float lo = __builtin_inff(), hi = -__builtin_inff();
        for (int q = 0; q < kPrepThreads / 32; ++q) {
            lo = fminf(lo, lo_s[q][f]);
            hi = fmaxf(hi, hi_s[q][f]);
        }
        stats[2 * kOut + f] = lo;
        stats[2 * kOut + kIn + f] = hi;
    }
}

// Pass 2: per block, the column sums of squared deviations from the mean (double) and the valid
// rows' indices, compacted in row order.
__global__ void __launch_bounds__(kEw) prep_dev_kernel(int n, const float* __restrict__ M, const double* __restrict__ dstat,
                                                       const uint32_t* __restrict__ boffset, double* __restrict__ psum,
                                                       uint32_t* __restrict__ idx) {
    __shared__ double ys[kEw][kOut];
    __shared__ uint32_t wcount[kEw / 64];
    const int t = threadIdx.x, i = blockIdx.x * kEw + t;
    double y[kOut];
    bool ok = false;
    if (i < n) ok = targets_of(M + size_t(i) * kOut, y);
#pragma unroll
    for (int k = 0; k < kOut; ++k) {
        const double d = ok ? y[k] - dstat[k] : 0.0;
        ys[t][k] = d * d;
    }
    const uint64_t m = __ballot(ok);
    const int lane = t & 63, wave = t >> 6;
    if (lane == 0) wcount[wave] = uint32_t(__popcll(m));
    __syncthreads();
    if (ok) {
        uint32_t before = uint32_t(__popcll(m & ((uint64_t(1) << lane) - 1u)));
        for (int w = 0; w < wave; ++w) before += wcount[w];
        idx[boffset[blockIdx.x] + before] = uint32_t(i);
    }
    const int rows = min(kEw, n - int(blockIdx.x) * kEw);
    if (t < kOut) {
        double acc = 0.0;
        for (int r = 0; r < rows; ++r) acc += ys[r][t];
        psum[size_t(blockIdx.x) * kOut + t] = acc;
    }
}

// One block of 1024 threads: the standard deviations (population), scaled by the output weights,
// into stats (chunked sums combined in chunk order).
__global__ void __launch_bounds__(kPrepThreads) prep_std_kernel(int blocks, const double* __restrict__ psum, const uint32_t* __restrict__ n_valid,
                                                                double* __restrict__ dstat, float* __restrict__ stats) {
    __shared__ double col[kPrepThreads / kOut][kOut];
    const int t = threadIdx.x;
    const int cc = t % kOut, q = t / kOut, nq = kPrepThreads / kOut, pq = (blocks + nq - 1) / nq;
    double acc = 0.0;
    for (int b = q * pq; b < min(blocks, (q + 1) * pq); ++b) acc += psum[size_t(b) * kOut + cc];
    col[q][cc] = acc;
    __syncthreads();
    if (t >= kOut) return;
    double sum = 0.0;
    for (int k = 0; k < nq; ++k) sum += col[k][t];
    const uint32_t c = *n_valid;
    double sd = c ? sqrt(sum / double(c)) : 1.0;
    sd = (sd > 1e-6 ? sd : 1.0) / double(kOutputScale[t]);
    dstat[kOut + t] = sd;
    stats[kOut + t] = float(sd);
}

// Pass 3: the standardised targets T = (y - mean) / std of every row (rows that are not valid are
// never drawn).
__global__ void __launch_bounds__(kEw) prep_targets_kernel(int n, const float* __restrict__ M, const double* __restrict__ dstat,
                                                           float* __restrict__ T) {
    const int i = blockIdx.x * kEw + threadIdx.x;
    if (i >= n) return;
    double y[kOut];
    (void)targets_of(M + size_t(i) * kOut, y);
#pragma unroll
    for (int k = 0; k < kOut; ++k) T[size_t(i) * kOut + k] = float((y[k] - dstat[k]) / dstat[kOut + k]);
}

// Device-side state of plt_surrogate_train: every launch of a training step reads what it needs
// from here, so one captured step replays as every step.
struct TrainState {
    uint32_t step, steps_per_epoch, steps_total, batch, n_valid, seed, adam_t;
    float lr0;
};
static_assert(sizeof(TrainState) <= kTrainStateWords * 4, "train state size");

__host__ __device__ inline uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}
// A bijection of [0, 2^(2 half)): a four-round balanced Feistel network keyed by `key`.
__device__ uint32_t feistel(uint32_t x, int half, uint32_t key) {
    const uint32_t mask = (1u << half) - 1u;
    uint32_t l = x >> half, r = x & mask;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        const uint32_t t = l ^ (mix32(r ^ mix32(key + k)) & mask);
        l = r;
        r = t;
    }
    return (l << half) | r;
}
// The epoch's shuffle of [0, n), n > 2^(2 half - 2): cycle-walking the Feistel bijection (the walk
// from x < n ends: the bijection's cycle through x returns to x; it takes under four steps on
// average).
__device__ uint32_t shuffle_index(uint32_t x, uint32_t n, int half, uint32_t key) {
    uint32_t y = feistel(x, half, key);
    while (y >= n) y = feistel(y, half, key);
    return y;
}

__host__ __device__ inline uint32_t epoch_key(uint32_t seed, uint32_t epoch) { return mix32(seed * 0x9E3779B9u + epoch); }
int shuffle_half(uint32_t n) {
    int bits = 2;
    while (bits < 32 && (uint64_t(1) << bits) < n) ++bits;
    return (bits + (bits & 1)) / 2;
}

// The epoch's order itself (plt_surrogate_shuffle).
__global__ void shuffle_kernel(uint32_t n, int half, uint32_t key, uint32_t* __restrict__ out) {
    const uint32_t x = blockIdx.x * kEw + threadIdx.x;
    if (x < n) out[x] = shuffle_index(x, n, half, key);
}

// One training batch: row i is the valid sample idx[shuffle(pos)] of this epoch's order, packed
// as the bf16 inputs (both layouts) and the FP32 targets.
__global__ void __launch_bounds__(kEw) gather_pack_kernel(const TrainState* __restrict__ st, const uint32_t* __restrict__ idx,
                                                          const float* __restrict__ X, const float* __restrict__ T, __bf16* __restrict__ Xb,
                                                          __bf16* __restrict__ Xt, int ldt, float* __restrict__ Tb, int half) {
    const TrainState s = *st;
    const uint32_t i = blockIdx.x * kEw + threadIdx.x;
    if (i >= s.batch) return;
    const uint32_t epoch = s.step / s.steps_per_epoch;
    const uint32_t pos = (s.step % s.steps_per_epoch) * s.batch + i;
    const uint32_t j = idx[shuffle_index(pos, s.n_valid, half, epoch_key(s.seed, epoch))];
#pragma unroll
    for (int f = 0; f < kInPad; ++f) {
        const __bf16 v = static_cast<__bf16>(f < kIn ? X[size_t(j) * kIn + f] : 0.f);
        Xb[size_t(i) * kInPad + f] = v;
        Xt[size_t(f) * ldt + i] = v;
    }
#pragma unroll
    for (int f = 0; f < kOut; ++f) Tb[size_t(i) * kOut + f] = T[size_t(j) * kOut + f];
}

// The bf16 operand copies of the four weight matrices (W [out x in_pad], Wt [in x out_pad]).
struct WeightCopies {
    __bf16 *W1, *W1t, *W2, *W2t, *W3, *W3t, *W4, *W4t;
};

// Adam on the FP32 master parameters fused with the bf16 copies of each updated weight (the
// padding of the copies stays zero). The learning rate and the bias corrections come from the
// arguments, or with `st` from the device state: the cosine-decayed rate lr0 (1 + cos(pi step /
// steps)) / 2 (at least 1e-6) and the corrections of Adam step adam_t + 1.
__global__ void adam_cast_kernel(int count, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                 float* __restrict__ v, const TrainState* __restrict__ st, float lr, float c1, float c2, float b1, float b2,
                                 float eps, WeightCopies w) {
    const int e = blockIdx.x * kEw + threadIdx.x;
    if (e >= count) return;
    if (st) {
        const TrainState s = *st;
        const double tt = double(s.adam_t) + 1.0;
        c1 = float(1.0 - pow(double(b1), tt));
        c2 = float(1.0 - pow(double(b2), tt));
        const double cs = cos(3.14159265358979323846 * double(s.step) / double(max(1u, s.steps_total)));
        lr = float(fmax(double(s.lr0) * 0.5 * (1.0 + cs), 1e-6));
    }
    const float ge = g[e];
    const float me = b1 * m[e] + (1.f - b1) * ge;
    const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    const float pe = p[e] - lr * (me / c1) / (sqrtf(ve / c2) + eps);
    p[e] = pe;
    const __bf16 h = static_cast<__bf16>(pe);
    if (size_t(e) < kOffB1) {
        const int o = e / kIn, i = e % kIn;
        w.W1[o * kInPad + i] = h;
        w.W1t[i * kHid + o] = h;
    } else if (size_t(e) >= kOffW2 && size_t(e) < kOffB2) {
        const int l = int(e - kOffW2), o = l / kHid, i = l % kHid;
        w.W2[o * kHid + i] = h;
        w.W2t[i * kHid + o] = h;
    } else if (size_t(e) >= kOffW3 && size_t(e) < kOffB3) {
        const int l = int(e - kOffW3), o = l / kHid, i = l % kHid;
        w.W3[o * kHid + i] = h;
        w.W3t[i * kHid + o] = h;
    } else if (size_t(e) >= kOffW4 && size_t(e) < kOffB4) {
        const int l = int(e - kOffW4), o = l / kHid, i = l % kHid;
        w.W4[o * kHid + i] = h;
        w.W4t[i * kOutPad + o] = h;
    }
}
WeightCopies weight_copies(const plt_surrogate* s) { return WeightCopies{s->W1, s->W1t, s->W2, s->W2t, s->W3, s->W3t, s->W4, s->W4t}; }

// End of a step: its mean loss (the loss kernel's block partials, summed in order) and the counters.
__global__ void step_end_kernel(TrainState* __restrict__ st, const float* __restrict__ loss_part, int blocks, float* __restrict__ loss_hist) {
    if (threadIdx.x != 0) return;
    double acc = 0.0;
    for (int b = 0; b < blocks; ++b) acc += double(loss_part[b]);
    loss_hist[st->step] = float(acc / (double(st->batch) * kOut));
    st->step += 1u;
    st->adam_t += 1u;
}

int pack(plt_surrogate* s, int n, const float* X, bool with_t, hipStream_t st) {
    const int ld = int(s->max_batch);
    hipLaunchKernelGGL(pack_inputs_kernel, dim3((n + kEw - 1) / kEw), dim3(kEw), 0, st, n, X, s->Xb, with_t ? s->Xt : nullptr, ld);
    return check_launch("pack_inputs_kernel");
}

}  // namespace

int plt::surrogate_infer(plt_surrogate* s, int n, const float* X, float* Y, const float* Sin, float* Sout, hipStream_t st) {
    if (n <= 0) return PLT_OK;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int tiles = (n + kFTile - 1) / kFTile;
    const int grid = std::max(1, std::min(tiles, 2 * cus));
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(kFThreads), 0, st, n, X, s->W1, s->W2, s->W3, s->W4, s->params, Y, s->stats,
                       Sin, Sout, nullptr);
    return check_launch("mlp_infer_kernel");
}

int plt::surrogate_infer_counted(plt_surrogate* s, const uint32_t* n_dev, uint32_t cap, const float* X, float* Y, hipStream_t st) {
    if (cap == 0) return PLT_OK;
    if (cap > (1u << 30)) return set_error(PLT_EINVAL, "surrogate inference: too many rows");
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int tiles = int((cap + kFTile - 1) / kFTile);
    const int grid = std::max(1, std::min(tiles, 2 * cus));
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(kFThreads), 0, st, int(cap), X, s->W1, s->W2, s->W3, s->W4, s->params, Y,
                       s->stats, nullptr, nullptr, n_dev);
    return check_launch("mlp_infer_kernel");
}

extern "C" {

int plt_surrogate_create(plt_context* ctx, uint32_t max_batch, uint32_t seed, plt_surrogate** out) {
    if (!ctx || !out || max_batch == 0 || max_batch % 32 || max_batch > (1u << 24))
        return set_error(PLT_EINVAL, "plt_surrogate_create: bad arguments (max_batch a multiple of 32, at most 2^24)");
    *out = nullptr;
    auto* s = new plt_surrogate;
    s->ctx = ctx;
    s->max_batch = max_batch;
    const size_t B = max_batch;
    int rc = PLT_OK;
    auto A = [&](auto** p, size_t count) { if (rc == PLT_OK) rc = dev_alloc(s, p, count); };
    A(&s->params, kParams); A(&s->grads, kParams); A(&s->adam_m, kParams); A(&s->adam_v, kParams);
    A(&s->W1, size_t(kHid) * kInPad); A(&s->W1t, size_t(kInPad) * kHid);
    A(&s->W2, size_t(kHid) * kHid); A(&s->W2t, size_t(kHid) * kHid);
    A(&s->W3, size_t(kHid) * kHid); A(&s->W3t, size_t(kHid) * kHid);
    A(&s->W4, size_t(kOut) * kHid); A(&s->W4t, size_t(kHid) * kOutPad);
    A(&s->Xb, B * kInPad); A(&s->Xt, B * kInPad);
    A(&s->H1, B * kHid); A(&s->H1t, B * kHid); A(&s->H2, B * kHid); A(&s->H2t, B * kHid); A(&s->H3, B * kHid); A(&s->H3t, B * kHid);
    A(&s->dY, B * kOutPad); A(&s->dYt, B * kOut);
    A(&s->G3, B * kHid); A(&s->G3t, B * kHid); A(&s->G2, B * kHid); A(&s->G2t, B * kHid); A(&s->G1t, B * kHid);
    A(&s->Y, B * kOut);
    A(&s->part, size_t(kSplitTarget) * kGradTotal);
    A(&s->loss_part, (B + kEw - 1) / kEw);
    A(&s->bias_part, (B / kTile + 1) * size_t(kBiasCols));
    A(&s->train_state, kTrainStateWords);
    A(&s->stats, 2 * kOut + 2 * kIn);
    A(&s->counter, 1);
    if (rc != PLT_OK) { plt_surrogate_destroy(s); return rc; }
    // the batch (K) padding of every feature-major operand must read as zeros
    const size_t tbytes = B * kHid * sizeof(__bf16);
    for (__bf16* p : {s->H1t, s->H2t, s->H3t, s->G3t, s->G2t, s->G1t})
        if (hipMemset(p, 0, tbytes) != hipSuccess) rc = set_error(PLT_EHIP, "plt_surrogate_create: clear failed");
    if (hipMemset(s->dYt, 0, B * kOut * sizeof(__bf16)) != hipSuccess || hipMemset(s->Xt, 0, B * kInPad * sizeof(__bf16)) != hipSuccess)
        rc = set_error(PLT_EHIP, "plt_surrogate_create: clear failed");
    // He-uniform initialisation of the weights (bound sqrt(6 / fan_in)), zero biases
    std::vector<float> p(kParams, 0.f);
    std::mt19937 gen(seed);
    std::uniform_real_distribution<float> u(-1.f, 1.f);
    auto init = [&](size_t off, int out_, int in_) {
        const float b = std::sqrt(6.f / float(in_));
        for (size_t e = 0; e < size_t(out_) * in_; ++e) p[off + e] = b * u(gen);
    };
    init(kOffW1, kHid, kIn); init(kOffW2, kHid, kHid); init(kOffW3, kHid, kHid); init(kOffW4, kOut, kHid);
    if (rc == PLT_OK && (hipMemcpy(s->params, p.data(), kParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
                         hipMemset(s->adam_m, 0, kParams * sizeof(float)) != hipSuccess ||
                         hipMemset(s->adam_v, 0, kParams * sizeof(float)) != hipSuccess))
        rc = set_error(PLT_EHIP, "plt_surrogate_create: upload failed");
    if (rc == PLT_OK) rc = cast_all(s, nullptr);
    if (rc == PLT_OK && hipDeviceSynchronize() != hipSuccess) rc = set_error(PLT_EHIP, "plt_surrogate_create: synchronize failed");
    if (rc != PLT_OK) { plt_surrogate_destroy(s); return rc; }
    *out = s;
    return PLT_OK;
}

void plt_surrogate_destroy(plt_surrogate* s) {
    if (!s) return;
    for (void* p : s->owned) (void)hipFree(p);
    delete s;
}

uint32_t plt_surrogate_param_count(void) { return uint32_t(kParams); }

int plt_surrogate_params(plt_surrogate* s, float* host, int set) {
    if (!s || !host) return set_error(PLT_EINVAL, "plt_surrogate_params: bad arguments");
    if (set) {
        if (hipMemcpy(s->params, host, kParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return set_error(PLT_EHIP, "plt_surrogate_params: upload failed");
        s->step = 0;
        if (hipMemset(s->adam_m, 0, kParams * sizeof(float)) != hipSuccess || hipMemset(s->adam_v, 0, kParams * sizeof(float)) != hipSuccess)
            return set_error(PLT_EHIP, "plt_surrogate_params: reset failed");
        const int rc = cast_all(s, nullptr);
        if (rc < 0) return rc;
        return hipDeviceSynchronize() == hipSuccess ? PLT_OK : set_error(PLT_EHIP, "plt_surrogate_params: synchronize failed");
    }
    return hipMemcpy(host, s->params, kParams * sizeof(float), hipMemcpyDeviceToHost) == hipSuccess
               ? PLT_OK
               : set_error(PLT_EHIP, "plt_surrogate_params: download failed");
}

int plt_surrogate_forward(plt_surrogate* s, uint32_t n, const float* X, float* Y, plt_stream stream) {
    if (!s || (n && (!X || !Y))) return set_error(PLT_EINVAL, "plt_surrogate_forward: bad arguments");
    if (n > (1u << 30)) return set_error(PLT_EINVAL, "plt_surrogate_forward: n too large");
    return surrogate_infer(s, int(n), X, Y, nullptr, nullptr, static_cast<hipStream_t>(stream));
}

int plt_surrogate_gradients(plt_surrogate* s, uint32_t n, const float* X, const float* T, float* grads_host, float* loss,
                            plt_stream stream) {
    if (!s || n == 0 || n % 32 || n > s->max_batch || !X || !T)
        return set_error(PLT_EINVAL, "plt_surrogate_gradients: bad arguments (n a multiple of 32, at most max_batch)");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    int rc;
    if ((rc = pack(s, int(n), X, true, st)) < 0) return rc;
    if ((rc = loss_and_grads(s, int(n), T, loss, st)) < 0) return rc;
    if (grads_host) {
        if (hipMemcpyAsync(grads_host, s->grads, kParams * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return set_error(PLT_EHIP, "plt_surrogate_gradients: download failed");
    }
    return PLT_OK;
}

int plt_surrogate_train_step(plt_surrogate* s, uint32_t n, const float* X, const float* T, float lr, float* loss, plt_stream stream) {
    if (!s || n == 0 || n % 32 || n > s->max_batch || !X || !T || !(lr > 0.f))
        return set_error(PLT_EINVAL, "plt_surrogate_train_step: bad arguments (n a multiple of 32, at most max_batch)");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    int rc;
    if ((rc = pack(s, int(n), X, true, st)) < 0) return rc;
    if ((rc = loss_and_grads(s, int(n), T, loss, st)) < 0) return rc;
    ++s->step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, float(s->step)), c2 = 1.f - std::pow(b2, float(s->step));
    hipLaunchKernelGGL(adam_cast_kernel, dim3((kParams + kEw - 1) / kEw), dim3(kEw), 0, st, kParams, s->params, s->grads, s->adam_m,
                       s->adam_v, nullptr, lr, c1, c2, b1, b2, 1e-8f, weight_copies(s));
    return check_launch("adam_cast_kernel");
}

int plt_surrogate_prepare(plt_surrogate* s, uint32_t n, const float* X, const float* M, float* T, uint32_t* idx, uint32_t* n_valid,
                          plt_stream stream) {
    if (!s || n == 0 || !X || !M || !T || !idx || !n_valid || n > (1u << 30))
        return set_error(PLT_EINVAL, "plt_surrogate_prepare: bad arguments");
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const int blocks = int((n + kEw - 1) / kEw);
    // scratch of the reductions: per-block sums, boxes and counts, the double statistics, the count
    const size_t bytes = size_t(blocks) * (kOut * 8 + 2 * kIn * 4 + 4) + 2 * kOut * 8 + 64;
    void* q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) return set_error(PLT_ENOMEM, "plt_surrogate_prepare: scratch allocation failed");
    char* b = static_cast<char*>(q);
    double* psum = reinterpret_cast<double*>(b);
    double* dstat = psum + size_t(blocks) * kOut;
    float* plo = reinterpret_cast<float*>(dstat + 2 * kOut);
    float* phi = plo + size_t(blocks) * kIn;
    uint32_t* pcount = reinterpret_cast<uint32_t*>(phi + size_t(blocks) * kIn);
    uint32_t* nv = pcount + blocks;
    int rc = PLT_OK;
    hipLaunchKernelGGL(prep_sums_kernel, dim3(blocks), dim3(kEw), 0, st, int(n), X, M, psum, plo, phi, pcount);
    hipLaunchKernelGGL(prep_mean_kernel, dim3(1), dim3(kPrepThreads), 0, st, blocks, psum, plo, phi, pcount, dstat, s->stats, nv);
    hipLaunchKernelGGL(prep_dev_kernel, dim3(blocks), dim3(kEw), 0, st, int(n), M, dstat, pcount, psum, idx);
    hipLaunchKernelGGL(prep_std_kernel, dim3(1), dim3(kPrepThreads), 0, st, blocks, psum, nv, dstat, s->stats);
    hipLaunchKernelGGL(prep_targets_kernel, dim3(blocks), dim3(kEw), 0, st, int(n), M, dstat, T);
    rc = check_launch("plt_surrogate_prepare kernels");
    if (rc == PLT_OK && (hipMemcpyAsync(n_valid, nv, sizeof(uint32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
                         hipStreamSynchronize(st) != hipSuccess))
        rc = set_error(PLT_EHIP, "plt_surrogate_prepare: read-back failed");
    (void)hipFree(q);
    if (rc == PLT_OK) s->has_stats = *n_valid > 0;
    return rc;
}

int plt_surrogate_shuffle(uint32_t n, uint32_t seed, uint32_t epoch, uint32_t* out, plt_stream stream) {
    if (n == 0 || !out) return set_error(PLT_EINVAL, "plt_surrogate_shuffle: bad arguments");
    hipLaunchKernelGGL(shuffle_kernel, dim3((n + kEw - 1) / kEw), dim3(kEw), 0, static_cast<hipStream_t>(stream), n, shuffle_half(n),
                       epoch_key(seed, epoch), out);
    return check_launch("shuffle_kernel");
}

int plt_surrogate_get_stats(const plt_surrogate* s, float* y_mean, float* y_std, float* x_lo, float* x_hi) {
    if (!s || !y_mean || !y_std || !x_lo || !x_hi) return set_error(PLT_EINVAL, "plt_surrogate_get_stats: bad arguments");
    if (!s->has_stats) return set_error(PLT_ESTATE, "plt_surrogate_get_stats: no statistics set");
    float h[2 * kOut + 2 * kIn];
    if (hipMemcpy(h, s->stats, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(PLT_EHIP, "plt_surrogate_get_stats: read-back failed");
    std::memcpy(y_mean, h, kOut * sizeof(float));
    std::memcpy(y_std, h + kOut, kOut * sizeof(float));
    std::memcpy(x_lo, h + 2 * kOut, kIn * sizeof(float));
    std::memcpy(x_hi, h + 2 * kOut + kIn, kIn * sizeof(float));
    return PLT_OK;
}

int plt_surrogate_train(plt_surrogate* s, uint32_t n_valid, const uint32_t* idx, const float* X, const float* T, uint32_t batch,
                        uint32_t epochs, float lr, uint32_t seed, int graph, float* epoch_loss, plt_stream stream) {
    if (!s || !idx || !X || !T || !epoch_loss || batch == 0 || batch % 32 || batch > s->max_batch || epochs == 0 || n_valid < batch ||
        !(lr > 0.f))
        return set_error(PLT_EINVAL, "plt_surrogate_train: bad arguments (batch a multiple of 32, at most max_batch and n_valid)");
    const uint32_t spe = n_valid / batch;
    const uint64_t total = uint64_t(epochs) * spe;
    if (total >= (uint64_t(1) << 31)) return set_error(PLT_EINVAL, "plt_surrogate_train: too many steps");
    const int half = shuffle_half(n_valid);
    const hipStream_t st = static_cast<hipStream_t>(stream);
    const TrainState ts{0u, spe, uint32_t(total), batch, n_valid, seed, s->step, lr};
    // the step's own buffers: the batch's targets and the per-step losses
    void* q = nullptr;
    if (hipMalloc(&q, size_t(batch) * kOut * 4 + size_t(total) * 4) != hipSuccess)
        return set_error(PLT_ENOMEM, "plt_surrogate_train: buffer allocation failed");
    float* Tb = static_cast<float*>(q);
    float* loss_hist = Tb + size_t(batch) * kOut;
    const int lblocks = int((batch + kEw - 1) / kEw);
    auto* state = reinterpret_cast<TrainState*>(s->train_state);
    int rc = PLT_OK;
    if (hipMemcpyAsync(state, &ts, sizeof(ts), hipMemcpyHostToDevice, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        rc = set_error(PLT_EHIP, "plt_surrogate_train: state upload failed");
    auto step = [&](hipStream_t qs) -> int {
        hipLaunchKernelGGL(gather_pack_kernel, dim3(lblocks), dim3(kEw), 0, qs, state, idx, X, T, s->Xb, s->Xt, int(s->max_batch), Tb, half);
        int r;
        if ((r = check_launch("gather_pack_kernel")) < 0) return r;
        if ((r = loss_and_grads(s, int(batch), Tb, nullptr, qs)) < 0) return r;
        hipLaunchKernelGGL(adam_cast_kernel, dim3((kParams + kEw - 1) / kEw), dim3(kEw), 0, qs, kParams, s->params, s->grads, s->adam_m,
                           s->adam_v, state, 0.f, 1.f, 1.f, 0.9f, 0.999f, 1e-8f, weight_copies(s));
        if ((r = check_launch("adam_cast_kernel")) < 0) return r;
        hipLaunchKernelGGL(step_end_kernel, dim3(1), dim3(64), 0, qs, state, s->loss_part, lblocks, loss_hist);
        return check_launch("step_end_kernel");
    };
    if (rc == PLT_OK && graph) {
        // one step captured and replayed: every step-dependent value is in the device state
        hipStream_t cs = nullptr;
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess ||
            hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = set_error(PLT_EHIP, "plt_surrogate_train: capture failed to start");
        } else {
            rc = step(cs);
            if (hipStreamEndCapture(cs, &g) != hipSuccess || !g) rc = rc < 0 ? rc : set_error(PLT_EHIP, "plt_surrogate_train: capture failed");
            if (rc == PLT_OK && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess)
                rc = set_error(PLT_EHIP, "plt_surrogate_train: graph instantiation failed");
            for (uint64_t k = 0; rc == PLT_OK && k < total; ++k)
                if (hipGraphLaunch(exec, st) != hipSuccess) rc = set_error(PLT_EHIP, "plt_surrogate_train: graph launch failed");
        }
        if (hipStreamSynchronize(st) != hipSuccess && rc == PLT_OK) rc = set_error(PLT_EHIP, "plt_surrogate_train: synchronize failed");
        if (exec) (void)hipGraphExecDestroy(exec);
        if (g) (void)hipGraphDestroy(g);
        if (cs) (void)hipStreamDestroy(cs);
    } else {
        for (uint64_t k = 0; rc == PLT_OK && k < total; ++k) rc = step(st);
    }
    std::vector<float> lh(total);
    if (rc == PLT_OK && (hipMemcpyAsync(lh.data(), loss_hist, total * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                         hipStreamSynchronize(st) != hipSuccess))
        rc = set_error(PLT_EHIP, "plt_surrogate_train: loss read-back failed");
    (void)hipStreamSynchronize(st);
    (void)hipFree(q);
    if (rc < 0) return rc;
    for (uint32_t e = 0; e < epochs; ++e) {
        double acc = 0.0;
        for (uint32_t k = 0; k < spe; ++k) acc += lh[size_t(e) * spe + k];
        epoch_loss[e] = float(acc / spe);
    }
    s->step += uint32_t(total);
    return PLT_OK;
}

// Raw GEMM (C = A * Bt^T, bf16 operands, FP32 output) for tests of the matrix-core kernel.
int plt_mlp_gemm_nt(uint32_t M, uint32_t N, uint32_t K, const void* A, const void* Bt, float* C, plt_stream stream) {
    if (!A || !Bt || !C) return set_error(PLT_EINVAL, "plt_mlp_gemm_nt: bad arguments");
    GemmArgs g = nt(static_cast<const __bf16*>(A), int(K), static_cast<const __bf16*>(Bt), int(K), int(M), int(N), int(K));
    g.epi = kEpiF32; g.Cf = C; g.ldc = int(N);
    return gemm(g, 1, static_cast<hipStream_t>(stream));
}

}  // extern "C"
