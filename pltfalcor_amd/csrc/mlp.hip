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
};

constexpr int kTile = 64, kKStep = 32, kLdsRow = kKStep + 8;  // +8 bf16 per row breaks bank conflicts
constexpr int kGemmThreads = 256;

__global__ void __launch_bounds__(kGemmThreads) gemm_nt_kernel(GemmArgs g) {
    __shared__ __bf16 As[kTile * kLdsRow];
    __shared__ __bf16 Bs[kTile * kLdsRow];
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
                        break;
                    }
                    default:
                        g.Cf[blockIdx.z * g.split_stride + size_t(i) * g.ldc + j] = v;
                        break;
                }
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
// the bf16 weights once into registers -- wave w holds the B fragments of output columns
// 32w..32w+31 of layers 1-3 (40 VGPRs... 72 with layer 1) and all of layer 4 (16 VGPRs) -- and
// then walks 64-sample tiles: each layer reads its input rows from LDS (A fragments), multiplies
// by its register-resident columns, and writes ReLU(acc + bias) back to LDS for the next layer.
// ---------------------------------------------------------------------------------------
constexpr int kFTile = 64, kFThreads = 256;
constexpr int kXRow = kInPad + 8, kHRow = kHid + 8;  // padded LDS rows (bank spread, 16 B aligned)

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

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
    const int fr = lane & 15, fk = 8 * (lane >> 4);
    // register-resident weights: B fragment (col block cb, k block kb) = W[col][32 kb + fk .. +7]
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
    const float* b1 = params + kHid * kIn;
    const float* b2 = b1 + kHid + kHid * kHid;
    const float* b3 = b2 + kHid + kHid * kHid;
    const float* b4 = b3 + kHid + kOut * kHid;
    const int ntiles = (n + kFTile - 1) / kFTile;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int r0 = tile * kFTile;
        // stage the inputs as bf16 (zero padded rows and features); flag samples outside the box
        if (t < kFTile) outside[t] = 0;
        __syncthreads();
        for (int e = t; e < kFTile * kInPad; e += kFThreads) {
            const int r = e / kInPad, f = e % kInPad;
            const float v = (r0 + r < n && f < kIn) ? X[size_t(r0 + r) * kIn + f] : 0.f;
            xs[r * kXRow + f] = static_cast<__bf16>(v);
            if (Sout && f < kIn && !(v >= stats[2 * kOut + f] && v <= stats[2 * kOut + kIn + f])) outside[r] = 1;
        }
        __syncthreads();
        // hidden layers: wave w computes columns 32w..32w+31 of all 64 rows
        auto hidden = [&](const __bf16* in, int in_row, int kblocks, const bf16x8 (*wf)[4], const bf16x8* wf1, const float* bias,
                          __bf16* out) {
            f32x4 acc[4][2];
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                for (int kb = 0; kb < 4; ++kb) {
                    if (kb >= kblocks) break;
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(in + (16 * rb + fr) * in_row + 32 * kb + fk);
#pragma unroll
                    for (int cb = 0; cb < 2; ++cb) acc[rb][cb] = mfma(a, wf ? wf[cb][kb] : wf1[cb], acc[rb][cb]);
                }
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                const int col = 32 * w + 16 * cb + fr;
                const float bv = bias[col];
#pragma unroll
                for (int rb = 0; rb < 4; ++rb)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        out[(16 * rb + 4 * (lane >> 4) + r) * kHRow + col] = static_cast<__bf16>(fmaxf(acc[rb][cb][r] + bv, 0.f));
            }
        };
        hidden(xs, kXRow, 1, nullptr, w1, b1, ha);
        __syncthreads();
        hidden(ha, kHRow, 4, w2, nullptr, b2, hb);
        __syncthreads();
        hidden(hb, kHRow, 4, w3, nullptr, b3, ha);
        __syncthreads();
        // output layer: wave w computes rows 16w..16w+15, all 16 outputs
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) acc = mfma(*reinterpret_cast<const bf16x8*>(ha + (16 * w + fr) * kHRow + 32 * kb + fk), w4[kb], acc);
        if (!Sout) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = r0 + 16 * w + 4 * (lane >> 4) + r;
                if (row < n) Y[size_t(row) * kOut + fr] = acc[r] + b4[fr];
            }
        } else {
            // gather each sample's 16 outputs in one lane through LDS (the input tile is no longer needed)
            float* ys = reinterpret_cast<float*>(xs);
#pragma unroll
            for (int r = 0; r < 4; ++r) ys[(16 * w + 4 * (lane >> 4) + r) * kOut + fr] = acc[r] + b4[fr];
            __syncthreads();
            // thread t: Stokes component r = t % 4 of sample t / 4
            const int lr = t >> 2, r = t & 3, row = r0 + lr;
            if (row < n && !outside[lr]) {
                const float* yr = ys + lr * kOut;
                const float m00 = expf(yr[0] * stats[kOut] + stats[0]);
                const float* S = Sin + size_t(row) * 4;
                float acc4 = 0.f;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int k = 4 * r + c;
                    const float v = k == 0 ? 1.f : sinhf(yr[k] * stats[kOut + k] + stats[k]);
                    acc4 += v * S[c];
                }
                Sout[size_t(row) * 4 + r] = m00 * acc4;
            }
        }
        __syncthreads();  // the next tile overwrites xs and ha
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
__global__ void loss_grad_kernel(int n, const float* __restrict__ Y, const float* __restrict__ T, __bf16* __restrict__ dY,
                                 __bf16* __restrict__ dYt, int ldt, float* __restrict__ loss_part) {
    __shared__ float red[kEw];
    const int i = blockIdx.x * kEw + threadIdx.x;
    float l = 0.f;
    if (i < n) {
        const float s = 2.f / (float(n) * float(kOut));
        for (int f = 0; f < kOutPad; ++f) {
            float g = 0.f;
            if (f < kOut) {
                const float d = Y[size_t(i) * kOut + f] - T[size_t(i) * kOut + f];
                l += d * d;
                g = s * d;
            }
            const __bf16 gb = static_cast<__bf16>(g);
            dY[size_t(i) * kOutPad + f] = gb;
            if (f < kOut) dYt[size_t(f) * ldt + i] = gb;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = kEw / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// out[e] = sum over splits of part[z * stride + e] (fixed order), for e < count.
__global__ void reduce_splits_kernel(int count, int splits, size_t stride, const float* __restrict__ part, float* __restrict__ out) {
    const int e = blockIdx.x * kEw + threadIdx.x;
    if (e >= count) return;
    float s = 0.f;
    for (int z = 0; z < splits; ++z) s += part[z * stride + e];
    out[e] = s;
}

// Bias gradient: out[j] = sum over the batch of Gt[j, 0..n) (one block per feature j).
__global__ void bias_grad_kernel(int n, const __bf16* __restrict__ Gt, int ldt, float* __restrict__ out) {
    __shared__ float red[kEw];
    const int j = blockIdx.x;
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += kEw) s += static_cast<float>(Gt[size_t(j) * ldt + i]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = kEw / 2; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = red[0];
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void adam_kernel(int count, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                            float lr, float b1, float b2, float eps, float c1, float c2) {
    const int e = blockIdx.x * kEw + threadIdx.x;
    if (e >= count) return;
    const float ge = g[e];
    const float me = b1 * m[e] + (1.f - b1) * ge;
    const float ve = b2 * v[e] + (1.f - b2) * ge * ge;
    m[e] = me;
    v[e] = ve;
    p[e] -= lr * (me / c1) / (sqrtf(ve / c2) + eps);
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

// dW[out x in] = sum over the batch of Gt[out, b] * Ht[in, b] (n a multiple of 32), as split-K
// partials reduced in a fixed order.
int weight_grad(plt_surrogate* s, int n, const __bf16* Gt, int out, const __bf16* Ht, int in, float* dst, hipStream_t st) {
    const int ld = int(s->max_batch);
    const int per = (n / kKStep + kSplitTarget - 1) / kSplitTarget * kKStep;
    const int splits = (n + per - 1) / per;
    GemmArgs g = nt(Gt, ld, Ht, ld, out, in, n);
    g.epi = kEpiPartial; g.Cf = s->part; g.ldc = in; g.k_per_split = per; g.split_stride = size_t(out) * in;
    int rc;
    if ((rc = gemm(g, splits, st)) < 0) return rc;
    const int count = out * in;
    hipLaunchKernelGGL(reduce_splits_kernel, dim3((count + kEw - 1) / kEw), dim3(kEw), 0, st, count, splits, g.split_stride, s->part, dst);
    return check_launch("reduce_splits_kernel");
}

// Loss and gradients of n samples (X packed, T [n x 16] FP32 targets) into s->grads.
int loss_and_grads(plt_surrogate* s, int n, const float* T, float* loss_host, hipStream_t st) {
    const int ld = int(s->max_batch);
    int rc;
    if ((rc = forward(s, n, true, st)) < 0) return rc;
    const int blocks = (n + kEw - 1) / kEw;
    hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks), dim3(kEw), 0, st, n, s->Y, T, s->dY, s->dYt, ld, s->loss_part);
    if ((rc = check_launch("loss_grad_kernel")) < 0) return rc;
    // backward data: G3 = (dY W4) * [H3 > 0], G2 = (G3 W3) * [H2 > 0], G1 = (G2 W2) * [H1 > 0]
    GemmArgs g = nt(s->dY, kOutPad, s->W4t, kOutPad, n, kHid, kOutPad);
    g.epi = kEpiMask; g.H = s->H3; g.ldh = kHid; g.Cb = s->G3; g.ldc = kHid; g.Ct = s->G3t; g.ldt = ld;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    g = nt(s->G3, kHid, s->W3t, kHid, n, kHid, kHid);
    g.epi = kEpiMask; g.H = s->H2; g.ldh = kHid; g.Cb = s->G2; g.ldc = kHid; g.Ct = s->G2t; g.ldt = ld;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    g = nt(s->G2, kHid, s->W2t, kHid, n, kHid, kHid);
    g.epi = kEpiMask; g.H = s->H1; g.ldh = kHid; g.Cb = nullptr; g.Ct = s->G1t; g.ldt = ld;
    if ((rc = gemm(g, 1, st)) < 0) return rc;
    // weight gradients (the batch is the K dimension)
    if ((rc = weight_grad(s, n, s->dYt, kOut, s->H3t, kHid, s->grads + kOffW4, st)) < 0) return rc;
    if ((rc = weight_grad(s, n, s->G3t, kHid, s->H2t, kHid, s->grads + kOffW3, st)) < 0) return rc;
    if ((rc = weight_grad(s, n, s->G2t, kHid, s->H1t, kHid, s->grads + kOffW2, st)) < 0) return rc;
    // W1's gradient over the 22 real inputs: the product is [128 x 32]; keep columns < 22
    if ((rc = weight_grad(s, n, s->G1t, kHid, s->Xt, kInPad, s->part + size_t(kSplitTarget + 1) * kHid * kHid, st)) < 0) return rc;
    if (hipMemcpy2DAsync(s->grads + kOffW1, kIn * sizeof(float), s->part + size_t(kSplitTarget + 1) * kHid * kHid, kInPad * sizeof(float),
                         kIn * sizeof(float), kHid, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return set_error(PLT_EHIP, "plt_surrogate: gradient copy failed");
    hipLaunchKernelGGL(bias_grad_kernel, dim3(kOut), dim3(kEw), 0, st, n, s->dYt, ld, s->grads + kOffB4);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(kHid), dim3(kEw), 0, st, n, s->G3t, ld, s->grads + kOffB3);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(kHid), dim3(kEw), 0, st, n, s->G2t, ld, s->grads + kOffB2);
    hipLaunchKernelGGL(bias_grad_kernel, dim3(kHid), dim3(kEw), 0, st, n, s->G1t, ld, s->grads + kOffB1);
    if ((rc = check_launch("bias_grad_kernel")) < 0) return rc;
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
    A(&s->part, size_t(kSplitTarget + 2) * kHid * kHid);
    A(&s->loss_part, (B + kEw - 1) / kEw);
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
    hipLaunchKernelGGL(adam_kernel, dim3((kParams + kEw - 1) / kEw), dim3(kEw), 0, st, kParams, s->params, s->grads, s->adam_m, s->adam_v,
                       lr, b1, b2, 1e-8f, c1, c2);
    if ((rc = check_launch("adam_kernel")) < 0) return rc;
    return cast_all(s, st);
}

// Raw GEMM (C = A * Bt^T, bf16 operands, FP32 output) for tests of the matrix-core kernel.
int plt_mlp_gemm_nt(uint32_t M, uint32_t N, uint32_t K, const void* A, const void* Bt, float* C, plt_stream stream) {
    if (!A || !Bt || !C) return set_error(PLT_EINVAL, "plt_mlp_gemm_nt: bad arguments");
    GemmArgs g = nt(static_cast<const __bf16*>(A), int(K), static_cast<const __bf16*>(Bt), int(K), int(M), int(N), int(K));
    g.epi = kEpiF32; g.Cf = C; g.ldc = int(N);
    return gemm(g, 1, static_cast<hipStream_t>(stream));
}

}  // extern "C"
