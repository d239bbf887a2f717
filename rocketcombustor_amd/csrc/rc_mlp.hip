// FP32 MLP for the species source term (no reference counterpart): a
// network (T, p, Y_1..Y_Nsp) -> omega_1..omega_{Nsp-1} with three tanh hidden
// layers, trained on the device by the kernels below (forward, backward,
// Adam) and evaluated per cell inside the steady step with a per-cell gate
// and the exact FP64 kinetics as fallback.
//
// Training of a width-128 network (at most 8 inputs, 16 outputs, 4 hidden
// layers): one fused kernel per 32 rows of the batch runs the forward pass,
// the loss gradient and the backward pass with the activations in LDS and
// every product an FP32 MFMA, and writes per-block partial gradients that a
// second kernel adds in block order before the Adam step (k_mlp_train_rows,
// k_mlp_reduce).  Other shapes: the dense products go through one LDS-tiled
// FP32 GEMM (64x64 tiles, 256 threads, 4x4 outputs per thread) with the
// layer's epilogue fused: bias + tanh (hidden), bias (output), or the tanh
// derivative (backward); the weight and bias gradients are split over the
// batch into fixed slices whose partials are added in slice order.  All
// reductions run in a fixed order, so training is deterministic run to run.  Inference (rc_mlp_forward and the step) of a width-128 network: one
// fused kernel per 64-row tile with the activations in LDS and FP32 MFMA
// products (k_mlp_infer); other widths take the GEMM chain.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rc_items.h"

namespace rc {

constexpr int kTile = 64, kTk = 16;

enum Epi { EPI_NONE = 0, EPI_BIAS = 1, EPI_BIAS_TANH = 2, EPI_DTANH = 3 };

// C[M x N] = op(A) op(B) (+ epilogue).  op(A) is M x K: A[m*lda + k], or
// A[k*lda + m] when TA; op(B) is K x N: B[k*ldb + n], or B[n*ldb + k] when TB.
// EPI_BIAS(_TANH): + bias[n] (then tanh); EPI_DTANH: * (1 - aux[m*ldc+n]^2).
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(256) k_gemm(const float *__restrict__ A, const float *__restrict__ B,
                                              float *__restrict__ C, int M, int N, int K, int lda, int ldb,
                                              int ldc, const float *__restrict__ bias,
                                              const float *__restrict__ aux) {
    __shared__ float As[kTk][kTile + 1];
    __shared__ float Bs[kTk][kTile + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    float acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
    for (int k0 = 0; k0 < K; k0 += kTk) {
        // 64 x 16 of op(A) and 16 x 64 of op(B), 4 elements per thread each
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = threadIdx.x + r * 256;     // 0..1023
            const int mm = e % kTile, kk = e / kTile;
            const int m = m0 + mm, k = k0 + kk;
            float v = 0.f;
            if (m < M && k < K) v = TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k];
            As[kk][mm] = v;
            const int nn = e % kTile;
            const int n = n0 + nn;
            float w = 0.f;
            if (n < N && k < K) w = TB ? B[(size_t)n * ldb + k] : B[(size_t)k * ldb + n];
            Bs[kk][nn] = w;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTk; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) { a[q] = As[kk][ty + 16 * q]; b[q] = Bs[kk][tx + 16 * q]; }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] = fmaf(a[p], b[q], acc[p][q]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int m = m0 + ty + 16 * p;
        if (m >= M) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int n = n0 + tx + 16 * q;
            if (n >= N) continue;
            float v = acc[p][q];
            if (EPI == EPI_BIAS || EPI == EPI_BIAS_TANH) v = v + bias[n];
            if (EPI == EPI_BIAS_TANH) v = tanhf(v);
            if (EPI == EPI_DTANH) {
                const float h = aux[(size_t)m * ldc + n];
                v = v * (1.f - h * h);
            }
            C[(size_t)m * ldc + n] = v;
        }
    }
}

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t st, const float *A, const float *B, float *C, int M, int N, int K, int lda, int ldb, int ldc,
          const float *bias = nullptr, const float *aux = nullptr) {
    dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
    k_gemm<TA, TB, EPI><<<grid, 256, 0, st>>>(A, B, C, M, N, K, lda, ldb, ldc, bias, aux);
}

// Weight and bias gradient of one layer, split over the batch:
// P[s][m][n] = sum over rows k of slice s of A[k*lda + m] B[k*ldb + n]
// (m < M), and row M of P[s] = sum over the slice of B[k*ldb + n] (the
// bias gradient).  Slices are `slice` rows (a multiple of kTk); blocks of
// grid row 0 form the bias row.  k_splitk_sum adds the slices in order.
__global__ void __launch_bounds__(256) k_gemm_at_splitk(const float *__restrict__ A, const float *__restrict__ B,
                                                        float *__restrict__ P, int M, int N, int K, int lda,
                                                        int ldb, int slice) {
    __shared__ float As[kTk][kTile + 1];
    __shared__ float Bs[kTk][kTile + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int kb = blockIdx.z * slice, ke = min(K, kb + slice);
    const bool bias_row = blockIdx.y == 0 && threadIdx.x < kTile;
    float acc[4][4], cs = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
    for (int k0 = kb; k0 < ke; k0 += kTk) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = threadIdx.x + r * 256;
            const int mm = e % kTile, kk = e / kTile;
            const int m = m0 + mm, k = k0 + kk, n = n0 + mm;
            As[kk][mm] = (m < M && k < ke) ? A[(size_t)k * lda + m] : 0.f;
            Bs[kk][mm] = (n < N && k < ke) ? B[(size_t)k * ldb + n] : 0.f;
        }
        __syncthreads();
        if (bias_row) {
#pragma unroll
            for (int kk = 0; kk < kTk; ++kk) cs += Bs[kk][threadIdx.x];
        }
#pragma unroll
        for (int kk = 0; kk < kTk; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) { a[q] = As[kk][ty + 16 * q]; b[q] = Bs[kk][tx + 16 * q]; }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] = fmaf(a[p], b[q], acc[p][q]);
        }
        __syncthreads();
    }
    float *Ps = P + (size_t)blockIdx.z * (M + 1) * N;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int m = m0 + ty + 16 * p;
        if (m >= M) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int n = n0 + tx + 16 * q;
            if (n < N) Ps[(size_t)m * N + n] = acc[p][q];
        }
    }
    if (bias_row && n0 + (int)threadIdx.x < N) Ps[(size_t)M * N + n0 + threadIdx.x] = cs;
}

// out[e] = sum over s < ns of P[s*len + e], in slice order
__global__ void __launch_bounds__(256) k_splitk_sum(const float *__restrict__ P, float *__restrict__ out, int len,
                                                    int ns) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= len) return;
    float s = 0.f;
    for (int k = 0; k < ns; ++k) s += P[(size_t)k * len + e];
    out[e] = s;
}

// slices of the batch for a layer gradient: about 512 blocks in all, slices
// of at least 64 rows
inline int splitk_slices(int M, int N, int K, int &slice) {
    const int tiles = ((M + kTile - 1) / kTile) * ((N + kTile - 1) / kTile);
    int ns = std::max(1, std::min(512 / tiles, (K + 63) / 64));
    slice = (((K + ns - 1) / ns) + kTk - 1) / kTk * kTk;
    ns = (K + slice - 1) / slice;
    return ns;
}
inline int splitk_max_slices(int M, int N) {
    const int tiles = ((M + kTile - 1) / kTile) * ((N + kTile - 1) / kTile);
    return std::max(1, 512 / tiles);
}

// ---------------------------------------------------------------------------
// fused forward pass of a network of width 128 (inputs <= 8, outputs <= 16):
// one block of 256 threads takes 64 rows through every layer with the
// activations in LDS (never in HBM); weights are staged in LDS 32 rows at a
// time; all products are FP32 MFMA 32x32x2.  Wave w holds rows 32 (w & 1) ..
// +31 and columns 64 (w >> 1) .. +63 (two 32x32 accumulator tiles); the
// output layer (at most 16 outputs) takes 16x16x4 tiles, and the two waves of
// a row slab split its K.  The 32x32x2 operand layout: lane l gives
// A[l % 32][l / 32] and B[l / 32][l % 32]; accumulator register r of lane l
// is D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32].
// ---------------------------------------------------------------------------
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kFW = 128;        // hidden width
constexpr int kFR = 64;         // rows per block
constexpr int kFS = kFW + 1;    // LDS row stride (odd: conflict-free column reads)
constexpr int kFK = 32;         // weight rows per LDS chunk

// tanh x = 1 - 2 / (exp(2x) + 1): exp overflows to +inf for large x (-> 1)
// and underflows to 0 for very negative x (-> -1)
__device__ __forceinline__ float fast_tanh(float x) { return 1.f - __fdividef(2.f, __expf(2.f * x) + 1.f); }

// (three waves per SIMD, as the LDS allows three blocks per CU: 115 VGPRs,
// no spill; the compiler's own choice keeps the accumulators in AGPRs and
// drops to two)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
k_mlp_infer(const float *__restrict__ X, float *__restrict__ Y, const float *__restrict__ P, int n, int nin, int nout,
            int hidden) {
    __shared__ float Hs[kFR * kFS];
    __shared__ float Ws[kFK * kFS];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int sl = w & 1, ch = w >> 1;
    const int li = lane & 31, lk = lane >> 5;
    const int row0 = blockIdx.x * kFR;
    // inputs, zero past nin (to 8 columns) and past the last row
    for (int e = t; e < kFR * 8; e += 256) {
        const int r = e >> 3, k = e & 7;
        Hs[r * kFS + k] = (k < nin && row0 + r < n) ? X[(size_t)(row0 + r) * nin + k] : 0.f;
    }
    const float *p = P;
    for (int l = 0; l <= hidden; ++l) {
        const bool last = l == hidden;
        const int fi = l == 0 ? nin : kFW, fo = last ? nout : kFW;
        const int kpad = l == 0 ? 8 : kFW;
        const float *Wl = p, *bl = p + (size_t)fi * fo;
        p = bl + fo;
        f32x16 acc0 = {}, acc1 = {};
        f32x4 out0 = {}, out1 = {};
        for (int k0 = 0; k0 < kpad; k0 += kFK) {
            const int kc = kpad - k0 < kFK ? kpad - k0 : kFK;
            __syncthreads();   // the previous chunk's (or layer's) reads of Ws and writes of Hs are done
            if (!last) {
                // 128-column rows, 16-byte loads (rows of a 128-wide W start on
                // 16-byte boundaries: every offset before one in the parameter
                // vector is a multiple of 4 floats)
                for (int e = t; e < kc * (kFW / 4); e += 256) {
                    const int k = e >> 5, n4 = (e & 31) * 4, kk = k0 + k;
                    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                    if (kk < fi) v = *reinterpret_cast<const float4 *>(Wl + (size_t)kk * kFW + n4);
                    float *w = Ws + k * kFS + n4;
                    w[0] = v.x;
                    w[1] = v.y;
                    w[2] = v.z;
                    w[3] = v.w;
                }
            } else {
                for (int e = t; e < (kc << 5); e += 256) {     // 32 columns, zero past fo
                    const int k = e >> 5, nn = e & 31, kk = k0 + k;
                    Ws[k * kFS + nn] = (kk < fi && nn < fo) ? Wl[(size_t)kk * fo + nn] : 0.f;
                }
            }
            __syncthreads();
            const float *ha = Hs + (32 * sl + li) * kFS + k0 + lk;
            if (!last) {
                const float *wb = Ws + lk * kFS + 64 * ch + li;
#pragma unroll 4
                for (int k = 0; k < kc; k += 2) {
                    const float a = ha[k];
                    const float b0 = wb[k * kFS], b1 = wb[k * kFS + 32];
                    acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b0, acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b1, acc1, 0, 0, 0);
                }
            } else {
                // one 16-column tile (MFMA 16x16x4: lane l gives A[l % 16][l / 16]
                // and B[l / 16][l % 16]) for each 16-row half of the slab; wave
                // ch takes rows [ch kc/2, (ch+1) kc/2) of the chunk's K (kc/2 is
                // 16, or 4 for a network without hidden layers)
                const int half = kc >> 1;
                const int l16 = lane & 15, lq = lane >> 4;
                const float *h0 = Hs + (32 * sl + l16) * kFS + k0 + ch * half + lq;
                const float *wq = Ws + (ch * half + lq) * kFS + l16;
#pragma unroll 4
                for (int k = 0; k < half; k += 4) {
                    const float b = wq[k * kFS];
                    out0 = __builtin_amdgcn_mfma_f32_16x16x4f32(h0[k], b, out0, 0, 0, 0);
                    out1 = __builtin_amdgcn_mfma_f32_16x16x4f32(h0[k + 16 * kFS], b, out1, 0, 0, 0);
                }
            }
        }
        __syncthreads();       // every wave has read its rows of this layer's input
        if (last) {
            // the two waves of a slab hold partial sums over the two halves of
            // K: ch 1 leaves its part in LDS (Hs is free now), ch 0 adds it
            // (fixed order) and writes the outputs.  Accumulator register r
            // of lane l is D[4 (l / 16) + r][l % 16].
            const int l16 = lane & 15, lq = lane >> 4;
            float *pp = Hs + sl * 32 * 17;
            if (ch == 1) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    pp[(4 * lq + r) * 17 + l16] = out0[r];
                    pp[(16 + 4 * lq + r) * 17 + l16] = out1[r];
                }
            }
            __syncthreads();
            if (ch == 0 && l16 < nout) {
                const float b0 = bl[l16];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int il = 4 * lq + r, i = 32 * sl + il;
                    if (row0 + i < n) Y[(size_t)(row0 + i) * nout + l16] = (out0[r] + pp[il * 17 + l16]) + b0;
                    if (row0 + i + 16 < n)
                        Y[(size_t)(row0 + i + 16) * nout + l16] = (out1[r] + pp[(il + 16) * 17 + l16]) + b0;
                }
            }
        } else {
            const int j0 = 64 * ch + li;
            const float b0 = bl[j0], b1 = bl[j0 + 32];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = 32 * sl + 8 * (r >> 2) + 4 * lk + (r & 3);
                Hs[i * kFS + j0] = fast_tanh(acc0[r] + b0);
                Hs[i * kFS + j0 + 32] = fast_tanh(acc1[r] + b1);
            }
        }
    }
}

// G = 2 (Y - T) / (M*N), and per-block partial sums of (Y - T)^2
__global__ void __launch_bounds__(256) k_mse_grad(const float *__restrict__ Y, const float *__restrict__ T,
                                                  float *__restrict__ G, float *__restrict__ partial, int count,
                                                  float scale) {
    __shared__ float sh[256];
    float s = 0.f;
    for (int e = blockIdx.x * 256 + threadIdx.x; e < count; e += gridDim.x * 256) {
        const float r = Y[e] - T[e];
        G[e] = 2.f * r * scale;
        s += r * r;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) sh[threadIdx.x] += sh[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_sum_partials(const float *__restrict__ partial, int n, float scale, float *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float s = 0.f;
        for (int k = 0; k < n; ++k) s += partial[k];
        *out = s * scale;
    }
}

// the training step counter (Adam's t, from 1) lives on the device, so a
// captured graph of several steps replays with the right counter
__global__ void k_step_count(long long *__restrict__ t) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t = *t + 1;
}

// the batch loss of the current step into the call's history (step t0 + 1
// of the call at index 0)
__global__ void k_record_loss(const float *__restrict__ loss, float *__restrict__ hist,
                              const long long *__restrict__ t, long long t0) {
    if (threadIdx.x == 0 && blockIdx.x == 0) hist[*t - t0 - 1] = *loss;
}

// Adam (Kingma & Ba) with bias correction 1 - beta^t, step t from 1
__global__ void __launch_bounds__(256) k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m1, float *__restrict__ m2, int n, float lr,
                                              float b1, float b2, float eps, const long long *__restrict__ t) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const float tf = (float)*t;
    const float c1 = 1.f - powf(b1, tf), c2 = 1.f - powf(b2, tf);
    const float gi = g[e];
    const float a = b1 * m1[e] + (1.f - b1) * gi;
    const float b = b2 * m2[e] + (1.f - b2) * gi * gi;
    m1[e] = a;
    m2[e] = b;
    p[e] = p[e] - lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// SplitMix64 (Steele, Lea, Flood 2014)
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// a minibatch of rows drawn with replacement from the resident data set:
// row r of step t from SplitMix64 at seed + t 2^32 + r
__global__ void __launch_bounds__(256) k_gather_batch(const float *__restrict__ X, const float *__restrict__ Y,
                                                      float *__restrict__ xb, float *__restrict__ yb, int batch,
                                                      long long n, int nin, int nout, uint64_t seed,
                                                      const long long *__restrict__ t) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= batch) return;
    const uint64_t s = seed + (uint64_t)*t * 0x100000000ull;
    const long long src = (long long)(splitmix64(s + (uint64_t)r) % (uint64_t)n);
    for (int k = 0; k < nin; ++k) xb[(size_t)r * nin + k] = X[(size_t)src * nin + k];
    for (int k = 0; k < nout; ++k) yb[(size_t)r * nout + k] = Y[(size_t)src * nout + k];
}

// ---------------------------------------------------------------------------
// Fused training pass of a width-128 network (inputs <= 8, outputs <= 16, at
// most kTH hidden layers): one block of 256 threads takes kTR rows of the
// batch (drawn from the resident data set as k_gather_batch draws them, or
// rows 0.. of given arrays) through the forward pass, keeping every layer's
// activations in LDS, the loss gradient and the backward pass, and writes
// its partial parameter gradient and partial squared error to its own row
// of `part`; k_mlp_reduce adds the rows in a fixed order and applies Adam.
// Every product is an FP32 MFMA 32x32x2 (operand layout as in k_mlp_infer):
// wave w owns columns 32w .. 32w+31 of a 32 x 128 product, and the m-tiles
// of column tile w of a weight gradient (a sum over the block's 32 rows).
// The activations and loss gradients (A operands, and B of the weight
// gradients) come from LDS; the weights (B operands of the products) straight
// from memory, where all blocks share them in L2: W (forward) and, for the
// backward products of the hidden layers, the transposed copy k_mlp_prep
// writes, so a lane's operand loads are consecutive in either direction.
// Rows past the batch have zero inputs and a zero loss gradient, so they
// add nothing.
// ---------------------------------------------------------------------------
constexpr int kTR = 32;          // rows per block
constexpr int kTH = 4;           // hidden layers at most
constexpr int kTW = kTR * kFS;   // floats of one 32-row activation tile in LDS
constexpr int kTO = 33;          // LDS row stride of the output layer's partial tiles

// the parameter offsets of layer l of a fused-shape network
__device__ __forceinline__ void fused_offsets(int l, int nin, int nout, int hidden, int &woff, int &boff) {
    const int first = nin * kFW + kFW, mid = kFW * kFW + kFW;
    woff = l == 0 ? 0 : first + (l - 1) * mid;
    boff = woff + (l == 0 ? nin : kFW) * (l == hidden ? nout : kFW);
}

// The step counter (when `bump`, as k_step_count) and the transposed copies
// of the 128 x 128 weights of layers 1 .. hidden-1: PT[(l-1) 128^2 + k 128 + n]
// = W_l[n][k].
__global__ void __launch_bounds__(256) k_mlp_prep(const float *__restrict__ P, float *__restrict__ PT,
                                                  long long *__restrict__ t, int bump, int nin, int nout,
                                                  int hidden) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (bump && e == 0) *t = *t + 1;
    const int l = 1 + e / (kFW * kFW), r = e % (kFW * kFW);
    if (l >= hidden) return;
    int woff, boff;
    fused_offsets(l, nin, nout, hidden, woff, boff);
    const int k = r / kFW, nn = r % kFW;
    PT[(size_t)(l - 1) * kFW * kFW + r] = P[woff + (size_t)nn * kFW + k];
}

__global__ void __launch_bounds__(256) k_mlp_train_rows(const float *__restrict__ X, const float *__restrict__ T,
                                                        long long ndata, int gather, uint64_t seed,
                                                        const long long *__restrict__ d_t, int n,
                                                        const float *__restrict__ P, const float *__restrict__ PT,
                                                        float *__restrict__ part, int nin, int nout, int hidden,
                                                        int nparam, float scale) {
    __shared__ float Hs[(kTH + 1) * kTW];   // layer inputs: Hs[0] the rows' inputs, Hs[l] hidden layer l's outputs
    __shared__ float Ds[2 * kTW];           // loss gradients of two layers (ping-pong)
    __shared__ long long src[kTR];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int li = lane & 31, lk = lane >> 5;
    const int row0 = blockIdx.x * kTR;
    float *prt = part + (size_t)blockIdx.x * (nparam + 1);
    if (t < kTR) {
        const int r = row0 + t;
        long long s = r < n ? r : -1;
        if (gather && r < n) {
            const uint64_t base = seed + (uint64_t)*d_t * 0x100000000ull;
            s = (long long)(splitmix64(base + (uint64_t)r) % (uint64_t)ndata);
        }
        src[t] = s;
    }
    __syncthreads();
    // inputs, zero past nin (to 32 columns: the input layer's gradient reads
    // a 32-wide tile) and on rows past the batch
    for (int e = t; e < kTR * 32; e += 256) {
        const int r = e >> 5, k = e & 31;
        Hs[r * kFS + k] = (k < nin && src[r] >= 0) ? X[(size_t)src[r] * nin + k] : 0.f;
    }
    // wave 0's targets (it forms the loss gradient), loaded now
    float tv[16];
    if (w == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * lk + (r & 3);
            tv[r] = (li < nout && src[i] >= 0) ? T[(size_t)src[i] * nout + li] : 0.f;
        }
    }
    // ---- forward: hidden layers (32 x 128 tiles, one 32-column tile per wave).
    // A lane's column of the next layer's B is loaded while the current layer
    // finishes (its MFMAs, epilogue and barrier), so each layer waits for at
    // most what is left of that latency.
    float bv[kFW / 2];
    auto load_col = [&](int l) {
        int wo, bo;
        fused_offsets(l, nin, nout, hidden, wo, bo);
        const float *wb = P + wo + (size_t)lk * kFW + 32 * w + li;
#pragma unroll
        for (int q = 0; q < kFW / 2; ++q) bv[q] = wb[(size_t)(2 * q) * kFW];
    };
    if (hidden > 1) load_col(1);
    for (int l = 0; l < hidden; ++l) {
        int woff, boff;
        fused_offsets(l, nin, nout, hidden, woff, boff);
        const int fi = l == 0 ? nin : kFW;
        __syncthreads();     // the layer's input rows are complete
        const float *ha = Hs + l * kTW + li * kFS + lk;
        f32x16 acc = {};
        if (l == 0) {
            const float *wb = P + woff + (size_t)lk * kFW + 32 * w + li;
#pragma unroll
            for (int k = 0; k < 8; k += 2) {
                const float b = k + lk < fi ? wb[(size_t)k * kFW] : 0.f;
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ha[k], b, acc, 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int q = 0; q < kFW / 2; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ha[2 * q], bv[q], acc, 0, 0, 0);
            if (l + 1 < hidden) load_col(l + 1);
        }
        float *out = Hs + (l + 1) * kTW;
        const int j = 32 * w + li;
        const float bj = P[boff + j];
#pragma unroll
        for (int r = 0; r < 16; ++r) out[(8 * (r >> 2) + 4 * lk + (r & 3)) * kFS + j] = fast_tanh(acc[r] + bj);
    }
    // ---- forward: output layer (32 x 32 tile, columns past nout zero); wave
    // w takes K rows 32w .. 32w+31, the four partial tiles are added in order
    int woff, boff;
    fused_offsets(hidden, nin, nout, hidden, woff, boff);
    __syncthreads();
    {
        const float *ha = Hs + hidden * kTW + li * kFS + 32 * w + lk;
        // (lanes past nout read column 0 and use zero: every load stays in W)
        const float *wb = P + woff + (size_t)(32 * w + lk) * nout + (li < nout ? li : 0);
        f32x16 acc = {};
        float bv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) bv[q] = li < nout ? wb[(size_t)(2 * q) * nout] : 0.f;
#pragma unroll
        for (int q = 0; q < 16; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ha[2 * q], bv[q], acc, 0, 0, 0);
        float *pp = Ds + (w - 1) * kTR * kTO;       // waves 1..3 leave their partial tiles in Ds
        if (w > 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) pp[(8 * (r >> 2) + 4 * lk + (r & 3)) * kTO + li] = acc[r];
        }
        __syncthreads();
        if (w == 0) {
            // outputs, loss gradient 2 (y - t) / count into Ds[1] (columns past
            // nout and rows past the batch zero), squared error
            const float bj = li < nout ? P[boff + li] : 0.f;
            float sq = 0.f;
            float g[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = 8 * (r >> 2) + 4 * lk + (r & 3);
                const float y = (((acc[r] + Ds[i * kTO + li]) + Ds[kTR * kTO + i * kTO + li]) +
                                 Ds[2 * kTR * kTO + i * kTO + li]) + bj;
                g[r] = 0.f;
                if (li < nout && src[i] >= 0) {
                    const float d = y - tv[r];
                    g[r] = 2.f * d * scale;
                    sq += d * d;
                }
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) Ds[kTW + (8 * (r >> 2) + 4 * lk + (r & 3)) * kFS + li] = g[r];
            // the wave's squared error, in a fixed order
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) sq += __shfl_xor(sq, o, 64);
            if (lane == 0) prt[nparam] = sq;
        }
    }
    // ---- backward, output layer: dW = H^T dY (128 x nout: m-tile w), db =
    // column sums of dY, dH = dY W^T (32 x 128, K = 16) times tanh'
    __syncthreads();
    float *cur = Ds + kTW, *nxt = Ds;
    {
        const float *hl = Hs + hidden * kTW;
        f32x16 acc = {};
        const float *ha = hl + lk * kFS + 32 * w + li;   // A[m][k] = H[row k][m]
        const float *gb = cur + lk * kFS + li;           // B[k][n] = dY[row k][n]
#pragma unroll 4
        for (int k = 0; k < kTR; k += 2) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ha[k * kFS], gb[k * kFS], acc, 0, 0, 0);
        if (li < nout) {
#pragma unroll
            for (int r = 0; r < 16; ++r) prt[woff + (32 * w + 8 * (r >> 2) + 4 * lk + (r & 3)) * nout + li] = acc[r];
        }
        if (t < nout) {
            float sb = 0.f;
            for (int i = 0; i < kTR; ++i) sb += cur[i * kFS + t];
            prt[boff + t] = sb;
        }
        // dH = dY W^T: B[k][n] = W[n][k] (k < nout)
        const float *ga = cur + li * kFS + lk;
        const float *wb = P + woff + (size_t)(32 * w + li) * nout;
        acc = f32x16{};
#pragma unroll
        for (int k = 0; k < 16; k += 2) {
            const float b = k + lk < nout ? wb[k + lk] : 0.f;
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ga[k], b, acc, 0, 0, 0);
        }
        const int j = 32 * w + li;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * lk + (r & 3);
            const float h = hl[i * kFS + j];
            nxt[i * kFS + j] = acc[r] * (1.f - h * h);
        }
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    // ---- backward, hidden layers l = hidden-1 .. 0 (W_l maps Hs[l] to Hs[l+1];
    // cur holds the loss gradient of layer l's pre-activations)
    for (int l = hidden - 1; l >= 0; --l) {
        fused_offsets(l, nin, nout, hidden, woff, boff);
        const int fi = l == 0 ? nin : kFW;
        __syncthreads();     // cur is complete
        const float *hl = Hs + l * kTW;
        // the lane's column of W_l^T for dH_l below, loaded during dW
        if (l > 0) {
            const float *wb = PT + (size_t)(l - 1) * kFW * kFW + (size_t)lk * kFW + 32 * w + li;
#pragma unroll
            for (int q = 0; q < kFW / 2; ++q) bv[q] = wb[(size_t)(2 * q) * kFW];
        }
        // dW = H_l^T cur: m-tiles 0..3 (0 only for the input layer) of column tile w
        const int mt = l == 0 ? 1 : 4;
        for (int q = 0; q < mt; ++q) {
            f32x16 acc = {};
            const float *ha = hl + lk * kFS + 32 * q + li;
            const float *gb = cur + lk * kFS + 32 * w + li;
#pragma unroll 4
            for (int k = 0; k < kTR; k += 2)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ha[k * kFS], gb[k * kFS], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = 32 * q + 8 * (r >> 2) + 4 * lk + (r & 3);
                if (m < fi) prt[woff + (size_t)m * kFW + 32 * w + li] = acc[r];
            }
        }
        if (t < kFW) {
            float v[kTR];
#pragma unroll
            for (int i = 0; i < kTR; ++i) v[i] = cur[i * kFS + t];
            float sb = 0.f;
#pragma unroll
            for (int i = 0; i < kTR; ++i) sb += v[i];
            prt[boff + t] = sb;
        }
        if (l == 0) break;
        // dH_l = cur W_l^T (B[k][n] = W_l[n][k], from the transposed copy),
        // times tanh'(Hs[l])
        const float *ga = cur + li * kFS + lk;
        f32x16 acc = {};
#pragma unroll
        for (int q = 0; q < kFW / 2; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ga[2 * q], bv[q], acc, 0, 0, 0);
        const int j = 32 * w + li;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * lk + (r & 3);
            const float h = hl[i * kFS + j];
            nxt[i * kFS + j] = acc[r] * (1.f - h * h);
        }
        float *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
}

// grads[p] = sum over blocks b < nb of part[b][p]; ADAM: then the Adam step
// of k_adam on parameter p.  A thread block takes 64 parameters; its four
// waves sum the blocks' rows in four consecutive quarters (eight loads in
// flight per lane), and the quarters are added in order, so the sum has one
// fixed order.  Parameter index nparam is the loss (the blocks' squared
// errors, times scale), recorded in `hist` as k_record_loss does.
constexpr int kRedP = 64;
template <bool ADAM>
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ part, int nb, int nparam,
                                                    float *__restrict__ grads, float *__restrict__ loss, float scale,
                                                    float *__restrict__ p, float *__restrict__ m1,
                                                    float *__restrict__ m2, float lr, float b1, float b2, float eps,
                                                    const long long *__restrict__ t, float *__restrict__ hist,
                                                    long long t0) {
    __shared__ float q[4][kRedP];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int e = blockIdx.x * kRedP + lane;
    const int per = (nb + 3) / 4, b0 = w * per, b1e = b0 + per < nb ? b0 + per : nb;
    float g = 0.f;
    if (e <= nparam) {
        const float *src = part + e;
        const size_t stride = (size_t)nparam + 1;
        for (int b = b0; b < b1e; b += 8) {
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = b + k < b1e ? src[(size_t)(b + k) * stride] : 0.f;
#pragma unroll
            for (int k = 0; k < 8; ++k) g += v[k];
        }
    }
    q[w][lane] = g;
    __syncthreads();
    if (w != 0 || e > nparam) return;
    g = ((q[0][lane] + q[1][lane]) + q[2][lane]) + q[3][lane];
    if (e == nparam) {
        *loss = g * scale;
        if (hist) hist[*t - t0 - 1] = g * scale;
        return;
    }
    grads[e] = g;
    if (ADAM) {
        const float tf = (float)*t;
        const float c1 = 1.f - powf(b1, tf), c2 = 1.f - powf(b2, tf);
        const float a = b1 * m1[e] + (1.f - b1) * g;
        const float bb = b2 * m2[e] + (1.f - b2) * g * g;
        m1[e] = a;
        m2[e] = bb;
        p[e] = p[e] - lr * (a / c1) / (sqrtf(bb / c2) + eps);
    }
}

// ---------------------------------------------------------------------------
// the learned source term of a case: inputs of cells [c0, c0+n) from PA,
// the network, then per cell the gate and the exact FP64 fallback
// ---------------------------------------------------------------------------
// xf: input mean (6), input std (6), output scales s_k (Nsp-1), the knee
// kappa of the output transform, raw input range lo (6), hi (6) -- inputs
// (T, log p, Y_1..Y_Nsp), raw range on (T, p, Y).  The network's output y_k
// stands for omega_k = kappa s_k sinh(y_k): linear below kappa s_k,
// logarithmic above, so a squared-error fit weighs relative errors.
constexpr int kSurIn = 2 + kNsp;
constexpr int kXfMean = 0, kXfStd = kSurIn, kXfScale = 2 * kSurIn, kXfKnee = kXfScale + kNsp - 1,
              kXfLo = kXfKnee + 1, kXfHi = kXfLo + kSurIn, kXfLen = kXfHi + kSurIn;

template <int NSP>
__global__ void __launch_bounds__(256) k_sur_inputs(const CaseDev *__restrict__ cases, int c0, int n,
                                                    const double *__restrict__ xf, float *__restrict__ x) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    const int i = c0 + r;
    const RC_G CaseDev &C = *gp(cases);
    const GState S = gview(C.S);
    double raw[kSurIn];
    raw[0] = PA_(3, i);
    raw[1] = log(PA_(0, i));
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NSP - 1; ++k) { raw[2 + k] = PA_(kNFlow + k, i); s = s + raw[2 + k]; }
    raw[1 + NSP] = 1.0 - s;
#pragma unroll
    for (int k = 0; k < kSurIn; ++k) x[(size_t)r * kSurIn + k] = (float)((raw[k] - xf[kXfMean + k]) / xf[kXfStd + k]);
}

// A calibrated trust region of the network: a grid over (T, log p, Y_1)
// whose cells are marked accepted where the network's rates on held-out
// states of that cell stayed within the calibration bound; a state outside
// the grid or in an unmarked cell is rejected.  No exact rate is formed.
struct TrustGrid {
    const unsigned char *accept = nullptr;   // dims[0]*dims[1]*dims[2], index (i0*d1 + i1)*d2 + i2
    int dims[3] = {0, 0, 0};
    double lo[3] = {0, 0, 0}, inv[3] = {0, 0, 0};   // cell of feature f: floor((f - lo) * inv)
};

__device__ __forceinline__ bool trust_accepts(const TrustGrid &g, double t, double logp, double y1) {
    const double f[3] = {t, logp, y1};
    int idx = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double u = (f[d] - g.lo[d]) * g.inv[d];
        if (!(u >= 0.0) || u >= (double)g.dims[d]) return false;
        idx = idx * g.dims[d] + (int)u;
    }
    return gp(g.accept)[idx] != 0;
}

// accept the network's source for a cell when its raw inputs lie in the
// training range and its rates keep the reaction's mass balance (the
// species rates of one reaction sum to zero): |sum w| <= tol (sum |w| +
// 1e-3 max scale); with a trust grid also when its (T, log p, Y_1) cell is
// marked; with rtol > 0 the rates must also lie within rtol of the exact
// FP64 rates, |w_k - w_k,exact| <= rtol max(|w_k,exact|, 1e-3 s_k).
// A rejected cell takes the exact FP64 rate of the source term.  Accepted
// cells are counted per block of 256 cells (acc, cumulative).
template <int NSP>
__global__ void __launch_bounds__(256) k_sur_gate(const CaseDev *__restrict__ cases, int c0, int n,
                                                  const double *__restrict__ xf, const float *__restrict__ y,
                                                  double tol, double rtol, TrustGrid trust,
                                                  long long *__restrict__ acc) {
    __shared__ int cnt[256];
    const int r = blockIdx.x * 256 + threadIdx.x;
    int ok = 0;
    if (r < n) {
        const int i = c0 + r;
        const RC_G CaseDev &C = *gp(cases);
        const GMesh M = gview(C.M);
        const GState S = gview(C.S);
        GThermo &th = *gp(C.th);
        double raw[kSurIn];
        raw[0] = PA_(3, i);
        raw[1] = PA_(0, i);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) { raw[2 + k] = PA_(kNFlow + k, i); s = s + raw[2 + k]; }
        raw[1 + NSP] = 1.0 - s;
        bool inside = true;
#pragma unroll
        for (int k = 0; k < kSurIn; ++k) inside = inside && raw[k] >= xf[kXfLo + k] && raw[k] <= xf[kXfHi + k];
        double w[NSP], wsum = 0.0, wabs = 0.0, smax = 0.0;
        const double knee = xf[kXfKnee];
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) {
            w[k] = knee * xf[kXfScale + k] * sinh((double)y[(size_t)r * (NSP - 1) + k]);
            wsum += w[k];
            wabs += fabs(w[k]);
            smax = fmax(smax, xf[kXfScale + k]);
        }
        ok = (inside && fabs(wsum) <= tol * (wabs + 1e-3 * smax)) ? 1 : 0;
        if (ok && trust.accept) ok = trust_accepts(trust, raw[0], log(raw[1]), raw[2]) ? 1 : 0;
        if (!ok || rtol > 0.0) {
            double yy[NSP], we[NSP];
#pragma unroll
            for (int k = 0; k < NSP - 1; ++k) yy[k] = raw[2 + k];
            yy[NSP - 1] = raw[1 + NSP];
            production_rate<NSP>(th, raw[0], LC_(LC_RR, i), yy, we);
            if (ok) {
#pragma unroll
                for (int k = 0; k < NSP - 1; ++k)
                    ok = ok && fabs(w[k] - we[k]) <= rtol * fmax(fabs(we[k]), 1e-3 * xf[kXfScale + k]);
            }
            if (!ok) {
#pragma unroll
                for (int k = 0; k < NSP - 1; ++k) w[k] = we[k];
            }
        }
#pragma unroll
        for (int k = 0; k < NSP - 1; ++k) S.src[k * M.sc + i] = w[k];
    }
    cnt[threadIdx.x] = ok;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) cnt[threadIdx.x] += cnt[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[c0 / 256 + blockIdx.x] += cnt[0];
}

}  // namespace rc

using namespace rc;

// device-resident network, data set and work buffers
struct rc_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    int nin = 0, nout = 0, width = 0, nlayer = 0;   // nlayer = hidden layers + 1
    std::vector<int> dims;                           // nin, width x hidden, nout
    std::vector<size_t> woff, boff;                  // offsets of W_l, b_l in the parameter vector
    size_t nparam = 0;
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    long long adam_t = 0;                            // Adam steps taken (host mirror of *d_t)
    long long *d_t = nullptr;                        // the same on the device
    float *data_x = nullptr, *data_y = nullptr;
    long long ndata = 0;
    int cap = 0;                                     // rows of the work buffers
    std::vector<float *> act;                        // act[0] = inputs, act[l] = layer l outputs (cap rows)
    std::vector<float *> dz;                         // gradients of the layer pre-activations (cap rows)
    float *ybuf = nullptr, *partial = nullptr, *loss = nullptr;
    float *fx = nullptr, *fy = nullptr;              // fused forward pass: inputs, outputs (fcap rows)
    int fcap = 0;
    float *splitk = nullptr;                         // slice partials of the layer gradients
    float *tpart = nullptr;                          // fused training: per-block partial gradients
    float *wt = nullptr;                             // fused training: transposed hidden weights
    int tcap = 0;                                    // rows tpart holds blocks for
    std::string err;
};

namespace {
thread_local std::string g_mlp_error;
int mlp_fail(rc_mlp *m, int code, const std::string &msg) {
    if (m) m->err = msg;
    g_mlp_error = msg;
    return code;
}
#define MCHK(m, x)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) return mlp_fail(m, RC_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(e_)); \
    } while (0)

constexpr int kPartials = 256;

int ensure_cap(rc_mlp *m, int rows) {
    if (rows <= m->cap) return RC_OK;
    for (float *p : m->act) (void)hipFree(p);
    for (float *p : m->dz) (void)hipFree(p);
    if (m->ybuf) (void)hipFree(m->ybuf);
    m->act.assign(m->nlayer + 1, nullptr);
    m->dz.assign(m->nlayer + 1, nullptr);
    for (int l = 0; l <= m->nlayer; ++l) {
        MCHK(m, hipMalloc(&m->act[l], sizeof(float) * (size_t)rows * m->dims[l]));
        MCHK(m, hipMalloc(&m->dz[l], sizeof(float) * (size_t)rows * m->dims[l]));
    }
    MCHK(m, hipMalloc(&m->ybuf, sizeof(float) * (size_t)rows * m->nout));
    m->cap = rows;
    return RC_OK;
}

// networks the fused forward kernel takes
bool fused_forward(const rc_mlp *m) { return m->width == kFW && m->nin <= 8 && m->nout <= 16; }

// y = f(x) for rows [0, n): the fused kernel where the shape allows, else
// the layer-by-layer GEMM chain through the activation buffers bufs
void infer(rc_mlp *m, hipStream_t st, const float *x, float *y, int n, float *const *bufs) {
    if (n <= 0) return;
    if (fused_forward(m)) {
        k_mlp_infer<<<(n + kFR - 1) / kFR, 256, 0, st>>>(x, y, m->params, n, m->nin, m->nout, m->nlayer - 1);
        return;
    }
    for (int l = 0; l < m->nlayer; ++l) {
        const float *W = m->params + m->woff[l], *b = m->params + m->boff[l];
        const int fi = m->dims[l], fo = m->dims[l + 1];
        const float *in = l == 0 ? x : bufs[l];
        float *out = l == m->nlayer - 1 ? y : bufs[l + 1];
        if (l < m->nlayer - 1)
            gemm<false, false, EPI_BIAS_TANH>(st, in, W, out, n, fo, fi, fi, fo, fo, b);
        else
            gemm<false, false, EPI_BIAS>(st, in, W, out, n, fo, fi, fi, fo, fo, b);
    }
}

// forward of rows act[0][0..n): hidden layers bias+tanh, output layer bias
// (the training pass: keeps every layer's activations for the backward pass)
void forward(rc_mlp *m, int n, const float *params) {
    for (int l = 0; l < m->nlayer; ++l) {
        const float *W = params + m->woff[l], *b = params + m->boff[l];
        const int fi = m->dims[l], fo = m->dims[l + 1];
        if (l < m->nlayer - 1)
            gemm<false, false, EPI_BIAS_TANH>(m->stream, m->act[l], W, m->act[l + 1], n, fo, fi, fi, fo, fo, b);
        else
            gemm<false, false, EPI_BIAS>(m->stream, m->act[l], W, m->act[l + 1], n, fo, fi, fi, fo, fo, b);
    }
}

// loss and parameter gradient of rows 0..n against targets ybuf
void backward(rc_mlp *m, int n) {
    const int L = m->nlayer;
    const int count = n * m->nout;
    const int nb = std::min(kPartials, (count + 255) / 256);
    k_mse_grad<<<nb, 256, 0, m->stream>>>(m->act[L], m->ybuf, m->dz[L], m->partial, count, 1.f / (float)count);
    k_sum_partials<<<1, 64, 0, m->stream>>>(m->partial, nb, 1.f / (float)count, m->loss);
    for (int l = L - 1; l >= 0; --l) {
        const int fi = m->dims[l], fo = m->dims[l + 1];
        // dW_l = act_l^T dz_{l+1}, db_l = column sums of dz_{l+1}: b_l
        // follows W_l in the parameter vector, so both are one (fi+1) x fo
        // block, formed in batch slices and summed in slice order
        int slice = 0;
        const int ns = splitk_slices(fi, fo, n, slice);
        dim3 grid((fo + kTile - 1) / kTile, (fi + kTile - 1) / kTile, ns);
        k_gemm_at_splitk<<<grid, 256, 0, m->stream>>>(m->act[l], m->dz[l + 1], m->splitk, fi, fo, n, fi, fo, slice);
        const int len = (fi + 1) * fo;
        k_splitk_sum<<<(len + 255) / 256, 256, 0, m->stream>>>(m->splitk, m->grads + m->woff[l], len, ns);
        if (l > 0)   // dz_l = (dz_{l+1} W_l^T) * (1 - act_l^2)
            gemm<false, true, EPI_DTANH>(m->stream, m->dz[l + 1], m->params + m->woff[l], m->dz[l], n, fi, fo, fo,
                                         fo, fi, nullptr, m->act[l]);
    }
}
// the fused training pass serves width-128 networks with 1 .. kTH hidden
// layers, at most 8 inputs and 16 outputs; RC_MLP_GEMM=1 takes the GEMM
// chain for every network (the same gradient to FP32 rounding)
bool fused_train(const rc_mlp *m) {
    const char *e = std::getenv("RC_MLP_GEMM");
    if (e && e[0] == '1') return false;
    const int hidden = m->nlayer - 1;
    return m->width == kFW && m->nin <= 8 && m->nout <= 16 && hidden >= 1 && hidden <= kTH;
}

int ensure_tpart(rc_mlp *m, int rows) {
    if (rows <= m->tcap) return RC_OK;
    if (m->tpart) (void)hipFree(m->tpart);
    m->tpart = nullptr;
    m->tcap = 0;
    const size_t nb = (size_t)(rows + kTR - 1) / kTR;
    MCHK(m, hipMalloc(&m->tpart, sizeof(float) * nb * (m->nparam + 1)));
    if (!m->wt) MCHK(m, hipMalloc(&m->wt, sizeof(float) * (size_t)std::max(m->nlayer - 2, 1) * kFW * kFW));
    m->tcap = rows;
    return RC_OK;
}

// loss and gradient of n rows by the fused pass (rows drawn from the data
// set at the device's step counter when `gather`, else rows 0..n-1 of X, T),
// then the block-order reduction, with the Adam step and the loss history
// entry when `adam`
// (`adam`: also the step counter, in k_mlp_prep, which k_step_count
// bumps on the GEMM chain)
void fused_pass(rc_mlp *m, const float *X, const float *T, long long ndata, bool gather, uint64_t seed, int n,
                bool adam, float lr, float b1, float b2, float eps, float *hist, long long t0) {
    const int nb = (n + kTR - 1) / kTR;
    const float scale = 1.f / (float)(n * m->nout);
    const int np = (int)m->nparam, hidden = m->nlayer - 1;
    const int nt = std::max(hidden - 1, 1) * kFW * kFW;
    k_mlp_prep<<<(nt + 255) / 256, 256, 0, m->stream>>>(m->params, m->wt, m->d_t, adam ? 1 : 0, m->nin, m->nout,
                                                       hidden);
    k_mlp_train_rows<<<nb, 256, 0, m->stream>>>(X, T, ndata, gather ? 1 : 0, seed, m->d_t, n, m->params, m->wt,
                                                m->tpart, m->nin, m->nout, hidden, np, scale);
    const int g = (np + 1 + kRedP - 1) / kRedP;
    if (adam)
        k_mlp_reduce<true><<<g, 256, 0, m->stream>>>(m->tpart, nb, np, m->grads, m->loss, scale, m->params, m->m1,
                                                     m->m2, lr, b1, b2, eps, m->d_t, hist, t0);
    else
        k_mlp_reduce<false><<<g, 256, 0, m->stream>>>(m->tpart, nb, np, m->grads, m->loss, scale, nullptr, nullptr,
                                                      nullptr, 0.f, 0.f, 0.f, 0.f, m->d_t, nullptr, 0);
}
}  // namespace

extern "C" {

const char *rc_mlp_last_error(const rc_mlp *m) { return m ? m->err.c_str() : g_mlp_error.c_str(); }

int rc_mlp_create(int device, int n_in, int width, int hidden, int n_out, rc_mlp **out) {
    if (!out || n_in < 1 || width < 1 || hidden < 1 || n_out < 1) return mlp_fail(nullptr, RC_ERR_ARGUMENT, "bad argument");
    *out = nullptr;
    rc_mlp *m = new rc_mlp();
    m->device = device;
    m->nin = n_in;
    m->nout = n_out;
    m->width = width;
    m->nlayer = hidden + 1;
    m->dims.push_back(n_in);
    for (int k = 0; k < hidden; ++k) m->dims.push_back(width);
    m->dims.push_back(n_out);
    for (int l = 0; l < m->nlayer; ++l) {
        m->woff.push_back(m->nparam);
        m->nparam += (size_t)m->dims[l] * m->dims[l + 1];
        m->boff.push_back(m->nparam);
        m->nparam += m->dims[l + 1];
    }
    auto bail = [&](int code) { rc_mlp_destroy(m); return code; };
    if (hipSetDevice(device) != hipSuccess) return bail(mlp_fail(nullptr, RC_ERR_DEVICE, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(mlp_fail(nullptr, RC_ERR_DEVICE, "stream creation failed"));
    for (float **p : {&m->params, &m->grads, &m->m1, &m->m2})
        if (hipMalloc(p, sizeof(float) * m->nparam) != hipSuccess)
            return bail(mlp_fail(nullptr, RC_ERR_MEMORY, "device allocation failed"));
    size_t nsplit = 0;
    for (int l = 0; l < m->nlayer; ++l)
        nsplit = std::max(nsplit, (size_t)splitk_max_slices(m->dims[l], m->dims[l + 1]) * (m->dims[l] + 1) *
                                      m->dims[l + 1]);
    if (hipMalloc(&m->splitk, sizeof(float) * nsplit) != hipSuccess)
        return bail(mlp_fail(nullptr, RC_ERR_MEMORY, "device allocation failed"));
    if (hipMalloc(&m->partial, sizeof(float) * kPartials) != hipSuccess ||
        hipMalloc(&m->loss, sizeof(float)) != hipSuccess || hipMalloc(&m->d_t, sizeof(long long)) != hipSuccess)
        return bail(mlp_fail(nullptr, RC_ERR_MEMORY, "device allocation failed"));
    if (hipMemsetAsync(m->d_t, 0, sizeof(long long), m->stream) != hipSuccess ||
        hipMemsetAsync(m->m1, 0, sizeof(float) * m->nparam, m->stream) != hipSuccess ||
        hipMemsetAsync(m->m2, 0, sizeof(float) * m->nparam, m->stream) != hipSuccess ||
        hipMemsetAsync(m->params, 0, sizeof(float) * m->nparam, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
        return bail(mlp_fail(nullptr, RC_ERR_DEVICE, "device initialisation failed"));
    *out = m;
    return RC_OK;
}

void rc_mlp_destroy(rc_mlp *m) {
    if (!m) return;
    if (m->stream) {
        (void)hipSetDevice(m->device);
        (void)hipStreamSynchronize(m->stream);
    }
    for (float *p : {m->params, m->grads, m->m1, m->m2, m->data_x, m->data_y, m->ybuf, m->partial, m->loss, m->splitk,
                     m->fx, m->fy, m->tpart, m->wt})
        if (p) (void)hipFree(p);
    if (m->d_t) (void)hipFree(m->d_t);
    for (float *p : m->act) if (p) (void)hipFree(p);
    for (float *p : m->dz) if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

long long rc_mlp_num_params(const rc_mlp *m) { return m ? (long long)m->nparam : -1; }

int rc_mlp_set_params(rc_mlp *m, const float *p) {
    if (!m || !p) return mlp_fail(m, RC_ERR_ARGUMENT, "NULL argument");
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipMemcpyAsync(m->params, p, sizeof(float) * m->nparam, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipMemsetAsync(m->m1, 0, sizeof(float) * m->nparam, m->stream));
    MCHK(m, hipMemsetAsync(m->m2, 0, sizeof(float) * m->nparam, m->stream));
    MCHK(m, hipMemsetAsync(m->d_t, 0, sizeof(long long), m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    m->adam_t = 0;
    return RC_OK;
}

int rc_mlp_get_params(rc_mlp *m, float *p) {
    if (!m || !p) return mlp_fail(m, RC_ERR_ARGUMENT, "NULL argument");
    MCHK(m, hipSetDevice(m->device));
    MCHK(m, hipMemcpyAsync(p, m->params, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return RC_OK;
}

int rc_mlp_forward(rc_mlp *m, const float *x, int n, float *y) {
    if (!m || !x || !y || n < 0) return mlp_fail(m, RC_ERR_ARGUMENT, "bad argument");
    if (n == 0) return RC_OK;
    MCHK(m, hipSetDevice(m->device));
    float *dx, *dy;
    if (fused_forward(m)) {
        // the fused pass keeps the activations on chip: inputs and outputs only
        if (n > m->fcap) {
            if (m->fx) (void)hipFree(m->fx);
            if (m->fy) (void)hipFree(m->fy);
            m->fx = m->fy = nullptr;
            m->fcap = 0;
            MCHK(m, hipMalloc(&m->fx, sizeof(float) * (size_t)n * m->nin));
            MCHK(m, hipMalloc(&m->fy, sizeof(float) * (size_t)n * m->nout));
            m->fcap = n;
        }
        dx = m->fx;
        dy = m->fy;
    } else {
        int rc = ensure_cap(m, n);
        if (rc) return rc;
        dx = m->act[0];
        dy = m->act[m->nlayer];
    }
    MCHK(m, hipMemcpyAsync(dx, x, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    infer(m, m->stream, dx, dy, n, m->act.data());
    MCHK(m, hipMemcpyAsync(y, dy, sizeof(float) * (size_t)n * m->nout, hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return RC_OK;
}

int rc_mlp_gradient(rc_mlp *m, const float *x, const float *t, int n, float *grad, float *loss) {
    if (!m || !x || !t || n < 1) return mlp_fail(m, RC_ERR_ARGUMENT, "bad argument");
    MCHK(m, hipSetDevice(m->device));
    int rc = ensure_cap(m, n);
    if (rc) return rc;
    MCHK(m, hipMemcpyAsync(m->act[0], x, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipMemcpyAsync(m->ybuf, t, sizeof(float) * (size_t)n * m->nout, hipMemcpyHostToDevice, m->stream));
    if (fused_train(m)) {
        if ((rc = ensure_tpart(m, n))) return rc;
        fused_pass(m, m->act[0], m->ybuf, n, false, 0, n, false, 0.f, 0.f, 0.f, 0.f, nullptr, 0);
    } else {
        forward(m, n, m->params);
        backward(m, n);
    }
    if (grad)
        MCHK(m, hipMemcpyAsync(grad, m->grads, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->stream));
    if (loss) MCHK(m, hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    return RC_OK;
}

int rc_mlp_set_data(rc_mlp *m, const float *x, const float *t, long long n) {
    if (!m || !x || !t || n < 1) return mlp_fail(m, RC_ERR_ARGUMENT, "bad argument");
    MCHK(m, hipSetDevice(m->device));
    if (m->data_x) (void)hipFree(m->data_x);
    if (m->data_y) (void)hipFree(m->data_y);
    m->data_x = m->data_y = nullptr;
    MCHK(m, hipMalloc(&m->data_x, sizeof(float) * (size_t)n * m->nin));
    MCHK(m, hipMalloc(&m->data_y, sizeof(float) * (size_t)n * m->nout));
    MCHK(m, hipMemcpyAsync(m->data_x, x, sizeof(float) * (size_t)n * m->nin, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipMemcpyAsync(m->data_y, t, sizeof(float) * (size_t)n * m->nout, hipMemcpyHostToDevice, m->stream));
    MCHK(m, hipStreamSynchronize(m->stream));
    m->ndata = n;
    return RC_OK;
}

int rc_mlp_train(rc_mlp *m, int nsteps, int batch, float lr, unsigned long long seed, int every, float *loss_out) {
    if (!m || nsteps < 0 || batch < 1) return mlp_fail(m, RC_ERR_ARGUMENT, "bad argument");
    if (!m->data_x) return mlp_fail(m, RC_ERR_ARGUMENT, "no training data (rc_mlp_set_data)");
    if (nsteps == 0) return RC_OK;
    MCHK(m, hipSetDevice(m->device));
    int rc = ensure_cap(m, batch);
    if (rc) return rc;
    const bool fused = fused_train(m);
    if (fused && (rc = ensure_tpart(m, batch))) return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float *hist = nullptr;
    MCHK(m, hipMalloc(&hist, sizeof(float) * (size_t)nsteps));
    const long long t0 = m->adam_t;
    hipStream_t st = m->stream;
    // one Adam step; every argument is fixed for the call, the step counter
    // is read on the device
    auto step = [&]() {
        if (fused) {
            fused_pass(m, m->data_x, m->data_y, m->ndata, true, seed, batch, true, lr, b1, b2, eps, hist, t0);
            return;
        }
        k_step_count<<<1, 1, 0, st>>>(m->d_t);
        k_gather_batch<<<(batch + 255) / 256, 256, 0, st>>>(m->data_x, m->data_y, m->act[0], m->ybuf, batch,
                                                            m->ndata, m->nin, m->nout, seed, m->d_t);
        forward(m, batch, m->params);
        backward(m, batch);
        k_record_loss<<<1, 1, 0, st>>>(m->loss, hist, m->d_t, t0);
        k_adam<<<(int)((m->nparam + 255) / 256), 256, 0, st>>>(m->params, m->grads, m->m1, m->m2, (int)m->nparam, lr,
                                                               b1, b2, eps, m->d_t);
    };
    // graphs of kTrainGraph steps replayed (launch-bound otherwise), the rest
    // launched directly: the same kernels with the same arguments
    constexpr int kTrainGraph = 50;
    const char *ng = std::getenv("RC_NO_GRAPH");
    const bool graphs = !(ng && ng[0] == '1') && nsteps >= kTrainGraph;
    int done = 0;
    hipError_t err = hipSuccess;
    if (graphs) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        err = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (err == hipSuccess) {
            for (int k = 0; k < kTrainGraph; ++k) step();
            err = hipStreamEndCapture(st, &g);
        }
        if (err == hipSuccess) err = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        for (; err == hipSuccess && done + kTrainGraph <= nsteps; done += kTrainGraph) err = hipGraphLaunch(ge, st);
        if (ge) (void)hipGraphExecDestroy(ge);
    }
    for (; err == hipSuccess && done < nsteps; ++done) step();
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    if (err != hipSuccess) {
        (void)hipFree(hist);
        return mlp_fail(m, RC_ERR_DEVICE, std::string("HIP error in training: ") + hipGetErrorString(err));
    }
    m->adam_t += nsteps;
    std::vector<float> h;
    if (loss_out && every > 0) {
        h.resize((size_t)nsteps);
        err = hipMemcpy(h.data(), hist, sizeof(float) * (size_t)nsteps, hipMemcpyDeviceToHost);
    }
    (void)hipFree(hist);
    if (err != hipSuccess) return mlp_fail(m, RC_ERR_DEVICE, std::string("HIP error: ") + hipGetErrorString(err));
    if (!h.empty()) {
        int nrep = 0;
        for (int s = 1; s <= nsteps; ++s)
            if (s % every == 0 || s == nsteps) loss_out[nrep++] = h[(size_t)s - 1];
    }
    return RC_OK;
}

}  // extern "C"

namespace rc {

// per-case learned-source state: the network (not owned), the transform,
// the chunk buffers of the forward pass and the acceptance counters
struct Surrogate {
    rc_mlp *m = nullptr;
    double tol = 0.0, rtol = 0.0;
    TrustGrid trust;             // device table, or none
    double *xf = nullptr;        // device, kXfLen
    float *buf[8] = {};          // inputs, (GEMM chain) activations, outputs of one chunk: dims[0..nlayer]
    int chunk = 0;
    long long *acc = nullptr;    // per block of 256 cells
    int nacc = 0;
};

int surrogate_xform_len() { return kXfLen; }

// the network's dims must be (2+Nsp) -> ... -> Nsp-1
int surrogate_create(rc_mlp *m, int ne, const double *xform, double tol, double rtol, Surrogate **out,
                     std::string &msg) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != m->device) {
        msg = "the network lives on device " + std::to_string(m->device) + ", the case on another";
        return RC_ERR_ARGUMENT;
    }
    if (m->nin != kSurIn || m->nout != kNsp - 1) {
        msg = "the network must map " + std::to_string(kSurIn) + " inputs to " + std::to_string(kNsp - 1) + " rates";
        return RC_ERR_ARGUMENT;
    }
    if (m->nlayer + 1 > 8) {
        msg = "at most 6 hidden layers";
        return RC_ERR_ARGUMENT;
    }
    Surrogate *s = new Surrogate();
    s->m = m;
    s->tol = tol;
    s->rtol = rtol;
    // the fused forward pass keeps no activations in HBM: the whole mesh in
    // one pass (up to 4M cells); the GEMM chain in chunks of 128k cells
    s->chunk = std::min(((ne + 255) / 256) * 256, fused_forward(m) ? (1 << 22) : 131072);
    s->nacc = (ne + 255) / 256;
    bool bad = hipMalloc(&s->xf, sizeof(double) * kXfLen) != hipSuccess ||
               hipMalloc(&s->acc, sizeof(long long) * s->nacc) != hipSuccess;
    // inputs and outputs; the hidden activations only for the GEMM chain
    for (int l = 0; l <= m->nlayer && !bad; ++l)
        if (l == 0 || l == m->nlayer || !fused_forward(m))
            bad = hipMalloc(&s->buf[l], sizeof(float) * (size_t)s->chunk * m->dims[l]) != hipSuccess;
    if (!bad)
        bad = hipMemcpy(s->xf, xform, sizeof(double) * kXfLen, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemset(s->acc, 0, sizeof(long long) * s->nacc) != hipSuccess;
    if (bad) {
        msg = "device allocation for the learned source failed";
        for (float *p : s->buf) if (p) (void)hipFree(p);
        if (s->xf) (void)hipFree(s->xf);
        if (s->acc) (void)hipFree(s->acc);
        delete s;
        return RC_ERR_MEMORY;
    }
    *out = s;
    return RC_OK;
}

// the calibrated trust grid of the gate (accept = NULL removes it)
int surrogate_set_trust(Surrogate *s, const int *dims, const double *lo, const double *hi,
                        const unsigned char *accept, std::string &msg) {
    if (s->trust.accept) (void)hipFree((void *)s->trust.accept);
    s->trust = TrustGrid{};
    if (!accept) return RC_OK;
    size_t cells = 1;
    for (int d = 0; d < 3; ++d) {
        if (dims[d] < 1 || dims[d] > 4096 || !(hi[d] > lo[d])) {
            msg = "trust grid: dims must be 1..4096 and hi > lo";
            return RC_ERR_ARGUMENT;
        }
        cells *= (size_t)dims[d];
    }
    unsigned char *d_acc = nullptr;
    if (hipMalloc(&d_acc, cells) != hipSuccess) {
        msg = "device allocation for the trust grid failed";
        return RC_ERR_MEMORY;
    }
    if (hipMemcpy(d_acc, accept, cells, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d_acc);
        msg = "copying the trust grid failed";
        return RC_ERR_DEVICE;
    }
    s->trust.accept = d_acc;
    for (int d = 0; d < 3; ++d) {
        s->trust.dims[d] = dims[d];
        s->trust.lo[d] = lo[d];
        s->trust.inv[d] = dims[d] / (hi[d] - lo[d]);
    }
    return RC_OK;
}

void surrogate_destroy(Surrogate *s) {
    if (!s) return;
    if (s->trust.accept) (void)hipFree((void *)s->trust.accept);
    for (float *p : s->buf) if (p) (void)hipFree(p);
    if (s->xf) (void)hipFree(s->xf);
    if (s->acc) (void)hipFree(s->acc);
    delete s;
}

// the learned source of every cell of the case (before the forward sweep)
void surrogate_enqueue(Surrogate *s, const CaseDev *dcase, int ne, hipStream_t st) {
    rc_mlp *m = s->m;
    for (int c0 = 0; c0 < ne; c0 += s->chunk) {
        const int n = std::min(s->chunk, ne - c0);
        k_sur_inputs<kNsp><<<(n + 255) / 256, 256, 0, st>>>(dcase, c0, n, s->xf, s->buf[0]);
        infer(m, st, s->buf[0], s->buf[m->nlayer], n, s->buf);
        k_sur_gate<kNsp><<<(n + 255) / 256, 256, 0, st>>>(dcase, c0, n, s->xf, s->buf[m->nlayer], s->tol, s->rtol,
                                                             s->trust, s->acc);
    }
}

// accepted cell evaluations since the surrogate was set
int surrogate_accepted(Surrogate *s, hipStream_t st, long long &acc) {
    std::vector<long long> h(s->nacc);
    if (hipMemcpyAsync(h.data(), s->acc, sizeof(long long) * s->nacc, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return RC_ERR_DEVICE;
    acc = 0;
    for (long long v : h) acc += v;
    return RC_OK;
}

}  // namespace rc
