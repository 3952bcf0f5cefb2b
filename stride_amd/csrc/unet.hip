// Kernels of the learned gradient preconditioner: a U-Net of 3x3 convolutions with BF16
// weights and activations and FP32 accumulation on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16), with the backward passes and the optimiser step needed to train
// it. The network has no counterpart in the reference; it maps the raw vp gradient of an FWI
// iteration (stride/physics/iso_acoustic/devito.py:1045-1111) to an update direction, which the
// host accepts only if the data misfit of the forward solver drops.
//
// Layouts: activations NHWC bf16 (channels innermost: the reduction axis of a convolution is
// contiguous, so a lane's 8 operand values are one 16-byte load); 3x3 weights [ky][kx][co][ci]
// bf16 with an fp32 master copy; biases fp32. Channel counts are multiples of 32, widths
// multiples of 64 and heights multiples of 4 (checked by the entry points). Every reduction is
// deterministic: partial sums per block in a fixed order, combined in block order.

#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

#define NTRY(x)                                                                      \
    do {                                                                             \
        if ((x) != hipSuccess) return sa_fail(SA_E_HIP, "HIP call failed: " #x);      \
    } while (0)

constexpr int kT = 256;

__device__ __forceinline__ bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int i = 0; i < 8; ++i) z[i] = (bf16)0.f;
    return z;
}

// ---- 3x3 convolution, implicit GEMM on the matrix cores ------------------------------------
// y[n,h,w,co] = act(sum_{ky,kx,ci} wt[ky][kx][co][ci] x[n,h+ky-1,w+kx-1,ci] + b[co]), zero
// padding. A = input pixels x channels (M = 16 pixels per tile, K = 32 channels per step),
// B = channels x output channels (N = 16 per tile). mfma_f32_16x16x32_bf16 lane layout: lane l
// holds A[l % 16][8 (l / 16) + i], B[8 (l / 16) + i][l % 16] (i < 8) and the result
// C[4 (l / 16) + i][l % 16] (i < 4).
// A workgroup of 4 waves: wave w computes row h = 4 blockIdx.y + w, pixels 64 blockIdx.x ..
// +63 (4 M tiles) and output channels 16 NT blockIdx.z' .. (NT N tiles), image n from
// blockIdx.z. amask (may be null, shaped as x): input elements whose amask is <= 0 read as zero
// (the ReLU derivative of the layer that produced x's gradient).
template <int NT>
__global__ __launch_bounds__(256) void k_conv3x3(const bf16 *__restrict__ x, const bf16 *__restrict__ amask,
                                                 const bf16 *__restrict__ wt, const float *__restrict__ bias,
                                                 bf16 *__restrict__ y, int H, int W, int Cin, int Cout, int relu,
                                                 int ncb) {
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int n = blockIdx.z / ncb;
    const int c0 = (blockIdx.z % ncb) * 16 * NT;
    const int h = blockIdx.y * 4 + wv;
    const int w0 = blockIdx.x * 64;
    const int lr = lane & 15, lq = lane >> 4;
    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const size_t img = (size_t)n * H * W;
    for (int ky = 0; ky < 3; ++ky) {
        const int hh = h + ky - 1;
        const bool hin = hh >= 0 && hh < H;
        for (int kx = 0; kx < 3; ++kx) {
            const bf16 *wk = wt + (size_t)(ky * 3 + kx) * Cout * Cin;
            for (int ci0 = 0; ci0 < Cin; ci0 += 32) {
                bf16x8 a[4], b[NT];
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const int ww = w0 + 16 * mt + lr + kx - 1;
                    a[mt] = zero8();
                    if (hin && ww >= 0 && ww < W) {
                        const size_t off = ((img + (size_t)hh * W + ww) * Cin) + ci0 + 8 * lq;
                        a[mt] = *(const bf16x8 *)(x + off);
                        if (amask) {
                            const bf16x8 mk = *(const bf16x8 *)(amask + off);
#pragma unroll
                            for (int i = 0; i < 8; ++i)
                                if (!((float)mk[i] > 0.f)) a[mt][i] = (bf16)0.f;
                        }
                    }
                }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    b[nt] = *(const bf16x8 *)(wk + (size_t)(c0 + 16 * nt + lr) * Cin + ci0 + 8 * lq);
#pragma unroll
                for (int mt = 0; mt < 4; ++mt)
#pragma unroll
                    for (int nt = 0; nt < NT; ++nt)
                        acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mt], b[nt], acc[mt][nt], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int co = c0 + 16 * nt + lr;
            const float bb = bias ? bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int ww = w0 + 16 * mt + 4 * lq + i;
                float v = acc[mt][nt][i] + bb;
                if (relu) v = fmaxf(v, 0.f);
                y[(img + (size_t)h * W + ww) * Cout + co] = (bf16)v;
            }
        }
}

// backward-data weights: wb[ky][kx][ci][co] = wt[2-ky][2-kx][co][ci]
__global__ void k_flip_weights(const bf16 *__restrict__ wt, bf16 *__restrict__ wb, int Cin, int Cout) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int total = 9 * Cin * Cout;
    if (idx >= total) return;
    const int co = idx % Cout;
    const int ci = (idx / Cout) % Cin;
    const int k = idx / (Cout * Cin);
    const int ky = k / 3, kx = k % 3;
    wb[idx] = wt[((size_t)((2 - ky) * 3 + (2 - kx)) * Cout + co) * Cin + ci];
}

// ---- weight gradient: dW[ky][kx][co][ci] = sum_p dp[p][co] x[p + (ky-1, kx-1)][ci] -----------
// dp = dy masked by the layer output (ReLU derivative) when ymask is given. A = dp^T (M = co,
// K = pixels), B = shifted x (K = pixels, N = ci). A workgroup computes one (ky, kx) and a
// TC x TC (co, ci) tile over its share of the pixels: 32 pixels per step are staged in LDS
// transposed (channel-major) so each lane's 8-pixel operand is one 16-byte LDS read; four
// waves hold 2 x 2 blocks of the tile. Partial sums go to part[split][ky][kx][co][ci].
template <int TC>
__global__ __launch_bounds__(256) void k_wgrad(const bf16 *__restrict__ dy, const bf16 *__restrict__ ymask,
                                               const bf16 *__restrict__ x, float *__restrict__ part, int N, int H,
                                               int W, int Cin, int Cout, int px_per_split, int nci) {
    constexpr int HT = TC / 2;       // per-wave tile edge
    constexpr int MT = HT / 16;      // MFMA tiles per wave edge
    __shared__ __attribute__((aligned(16))) bf16 sd[TC][32 + 8];
    __shared__ __attribute__((aligned(16))) bf16 sx[TC][32 + 8];
    const int split = blockIdx.x;
    const int kk = blockIdx.y;
    const int ky = kk / 3, kx = kk % 3;
    const int co0 = (blockIdx.z / nci) * TC, ci0 = (blockIdx.z % nci) * TC;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    const int wco = (wv >> 1) * HT, wci = (wv & 1) * HT;
    const long long npx = (long long)N * H * W;
    const long long p_begin = (long long)split * px_per_split;
    const long long p_end = p_begin + px_per_split < npx ? p_begin + px_per_split : npx;
    f32x4 acc[MT][MT];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // staging: thread t loads 8 channels of one pixel: pixel t % 32, channel chunk t / 32
    const int sp = threadIdx.x & 31, sc = threadIdx.x >> 5;  // sc < 8: up to 64 channels
    for (long long p0 = p_begin; p0 < p_end; p0 += 32) {
        if (8 * sc < TC) {
            const long long p = p0 + sp;
            bf16x8 dv = zero8(), xv = zero8();
            if (p < p_end) {
                const int w = (int)(p % W);
                const long long r = p / W;
                const int h = (int)(r % H);
                const long long n = r / H;
                const size_t doff = (size_t)p * Cout + co0 + 8 * sc;
                dv = *(const bf16x8 *)(dy + doff);
                if (ymask) {
                    const bf16x8 mk = *(const bf16x8 *)(ymask + doff);
#pragma unroll
                    for (int i = 0; i < 8; ++i)
                        if (!((float)mk[i] > 0.f)) dv[i] = (bf16)0.f;
                }
                const int hh = h + ky - 1, ww = w + kx - 1;
                if (hh >= 0 && hh < H && ww >= 0 && ww < W)
                    xv = *(const bf16x8 *)(x + ((size_t)(n * H + hh) * W + ww) * Cin + ci0 + 8 * sc);
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                sd[8 * sc + i][sp] = dv[i];
                sx[8 * sc + i][sp] = xv[i];
            }
        }
        __syncthreads();
        bf16x8 a[MT], b[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i) {
            a[i] = *(const bf16x8 *)&sd[wco + 16 * i + lr][8 * lq];
            b[i] = *(const bf16x8 *)&sx[wci + 16 * i + lr][8 * lq];
        }
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < MT; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    float *out = part + ((size_t)split * 9 + kk) * Cout * Cin;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < MT; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + wco + 16 * i + 4 * lq + r;
                const int ci = ci0 + wci + 16 * j + lr;
                out[(size_t)co * Cin + ci] = acc[i][j][r];
            }
}

// out[j] = sum_s part[s][j], s in order (double accumulation)
__global__ void k_sum_splits(const float *__restrict__ part, int nsplit, long long n, float *__restrict__ out,
                             int accumulate) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * n + j];
    out[j] = (float)s + (accumulate ? out[j] : 0.f);
}

// per-channel sums of dy (masked by ymask > 0): partial sums per block of rows, in order
__global__ void k_channel_partial(const bf16 *__restrict__ dy, const bf16 *__restrict__ ymask, long long npx,
                                  int C, int px_per_block, float *__restrict__ part) {
    const int c = threadIdx.x;  // blockDim.x = C (<= 1024)
    const long long p0 = (long long)blockIdx.x * px_per_block;
    const long long p1 = p0 + px_per_block < npx ? p0 + px_per_block : npx;
    double s = 0.0;
    for (long long p = p0; p < p1; ++p) {
        const size_t o = (size_t)p * C + c;
        const float v = (float)dy[o];
        if (!ymask || (float)ymask[o] > 0.f) s += v;
    }
    part[(size_t)blockIdx.x * C + c] = (float)s;
}

// ---- pooling, upsampling, concatenation ----------------------------------------------------
// 2 x 2 average pool: y[n,h,w,c] = mean of x[n,2h..2h+1,2w..2w+1,c]
__global__ void k_avgpool(const bf16 *__restrict__ x, bf16 *__restrict__ y, int N, int H, int W, int C) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;  // output element / 8
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    if (idx >= (long long)N * Ho * Wo * C8) return;
    const int c = (int)(idx % C8) * 8;
    long long r = idx / C8;
    const int w = (int)(r % Wo);
    r /= Wo;
    const int h = (int)(r % Ho);
    const long long n = r / Ho;
    float s[8] = {};
#pragma unroll
    for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
            const bf16x8 v = *(const bf16x8 *)(x + ((size_t)(n * H + 2 * h + dy) * W + 2 * w + dx) * C + c);
#pragma unroll
            for (int i = 0; i < 8; ++i) s[i] += (float)v[i];
        }
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = (bf16)(0.25f * s[i]);
    *(bf16x8 *)(y + ((size_t)(n * Ho + h) * Wo + w) * C + c) = o;
}

// backward of the average pool: dx[n,h,w,c] = dy[n,h/2,w/2,c] / 4 (+ dx if accumulate)
__global__ void k_avgpool_back(const bf16 *__restrict__ dy, bf16 *__restrict__ dx, int N, int H, int W, int C,
                               int accumulate) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    if (idx >= (long long)N * H * W * C8) return;
    const int c = (int)(idx % C8) * 8;
    long long r = idx / C8;
    const int w = (int)(r % W);
    r /= W;
    const int h = (int)(r % H);
    const long long n = r / H;
    const bf16x8 g = *(const bf16x8 *)(dy + ((size_t)(n * (H / 2) + h / 2) * (W / 2) + w / 2) * C + c);
    bf16x8 *o = (bf16x8 *)(dx + ((size_t)(n * H + h) * W + w) * C + c);
    bf16x8 out;
    const bf16x8 prev = accumulate ? *o : zero8();
#pragma unroll
    for (int i = 0; i < 8; ++i) out[i] = (bf16)(0.25f * (float)g[i] + (float)prev[i]);
    *o = out;
}

// nearest upsample x2 of x (N,H,W,C) written into channels [coff, coff + C) of y (N,2H,2W,Cy)
__global__ void k_upsample_into(const bf16 *__restrict__ x, bf16 *__restrict__ y, int N, int H, int W, int C,
                                int Cy, int coff) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8, Ho = 2 * H, Wo = 2 * W;
    if (idx >= (long long)N * Ho * Wo * C8) return;
    const int c = (int)(idx % C8) * 8;
    long long r = idx / C8;
    const int w = (int)(r % Wo);
    r /= Wo;
    const int h = (int)(r % Ho);
    const long long n = r / Ho;
    *(bf16x8 *)(y + ((size_t)(n * Ho + h) * Wo + w) * Cy + coff + c) =
        *(const bf16x8 *)(x + ((size_t)(n * H + h / 2) * W + w / 2) * C + c);
}

// backward of the upsample: dx[n,h,w,c] = sum of the 4 children in channels [coff, coff + C) of dy
__global__ void k_upsample_back(const bf16 *__restrict__ dy, bf16 *__restrict__ dx, int N, int H, int W, int C,
                                int Cy, int coff) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    if (idx >= (long long)N * H * W * C8) return;
    const int c = (int)(idx % C8) * 8;
    long long r = idx / C8;
    const int w = (int)(r % W);
    r /= W;
    const int h = (int)(r % H);
    const long long n = r / H;
    float s[8] = {};
#pragma unroll
    for (int dyy = 0; dyy < 2; ++dyy)
#pragma unroll
        for (int dxx = 0; dxx < 2; ++dxx) {
            const bf16x8 v =
                *(const bf16x8 *)(dy + ((size_t)(n * 2 * H + 2 * h + dyy) * (2 * W) + 2 * w + dxx) * Cy + coff + c);
#pragma unroll
            for (int i = 0; i < 8; ++i) s[i] += (float)v[i];
        }
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = (bf16)s[i];
    *(bf16x8 *)(dx + ((size_t)(n * H + h) * W + w) * C + c) = o;
}

// copy channels: dst[p, doff + c] (+)= src[p, soff + c] for c < C (concatenation and its split)
__global__ void k_copy_channels(const bf16 *__restrict__ src, int Cs, int soff, bf16 *__restrict__ dst, int Cd,
                                int doff, int C, long long npx, int accumulate) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    if (idx >= npx * C8) return;
    const int c = (int)(idx % C8) * 8;
    const long long p = idx / C8;
    const bf16x8 v = *(const bf16x8 *)(src + (size_t)p * Cs + soff + c);
    bf16x8 *o = (bf16x8 *)(dst + (size_t)p * Cd + doff + c);
    if (accumulate) {
        const bf16x8 q = *o;
        bf16x8 r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = (bf16)((float)v[i] + (float)q[i]);
        *o = r;
    } else {
        *o = v;
    }
}

// ---- network input and output --------------------------------------------------------------
// x[p][0] = bf16(s0 g0[p]), x[p][1] = bf16(s1 g1[p]) (zero without g1), channels 2 .. C-1 zero
__global__ void k_pack_input(const float *__restrict__ g0, float s0, const float *__restrict__ g1, float s1,
                             bf16 *__restrict__ x, long long npx, int C) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    if (idx >= npx * C8) return;
    const int c = (int)(idx % C8) * 8;
    const long long p = idx / C8;
    bf16x8 v = zero8();
    if (c == 0) {
        v[0] = (bf16)(s0 * g0[p]);
        if (g1) v[1] = (bf16)(s1 * g1[p]);
    }
    *(bf16x8 *)(x + (size_t)p * C + c) = v;
}

// 1 x 1 projection to one channel: out[p] = sum_c w[c] x[p][c] + b (fp32 out)
__global__ void k_head(const bf16 *__restrict__ x, const float *__restrict__ w, const float *__restrict__ b,
                       float *__restrict__ out, long long npx, int C) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx) return;
    float s = b[0];
    for (int c = 0; c < C; c += 8) {
        const bf16x8 v = *(const bf16x8 *)(x + (size_t)p * C + c);
#pragma unroll
        for (int i = 0; i < 8; ++i) s += w[c + i] * (float)v[i];
    }
    out[p] = s;
}

// backward of the projection: dx[p][c] = dout[p] w[c] (masked by x > 0: the ReLU of x's layer)
__global__ void k_head_back_data(const float *__restrict__ dout, const float *__restrict__ w,
                                 const bf16 *__restrict__ x, bf16 *__restrict__ dx, long long npx, int C) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    if (idx >= npx * C8) return;
    const int c = (int)(idx % C8) * 8;
    const long long p = idx / C8;
    const bf16x8 xv = *(const bf16x8 *)(x + (size_t)p * C + c);
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; ++i) o[i] = (bf16)((float)xv[i] > 0.f ? dout[p] * w[c + i] : 0.f);
    *(bf16x8 *)(dx + (size_t)p * C + c) = o;
}

// partial sums per block for the projection's weight and bias gradients:
// part[blk][c] = sum_p dout[p] x[p][c] (c < C), part[blk][C] = sum_p dout[p]
__global__ void k_head_back_weights(const float *__restrict__ dout, const bf16 *__restrict__ x, long long npx, int C,
                                    int px_per_block, float *__restrict__ part) {
    const int c = threadIdx.x;  // blockDim.x = C + 1
    const long long p0 = (long long)blockIdx.x * px_per_block;
    const long long p1 = p0 + px_per_block < npx ? p0 + px_per_block : npx;
    double s = 0.0;
    for (long long p = p0; p < p1; ++p) s += (double)dout[p] * (c < C ? (float)x[(size_t)p * C + c] : 1.f);
    part[(size_t)blockIdx.x * (C + 1) + c] = (float)s;
}

// ---- loss and optimiser --------------------------------------------------------------------
// d[p] = (out[p] - target[p]) * inv_n; part[blk] = sum over the block of (out - target)^2
__global__ void k_mse(const float *__restrict__ out, const float *__restrict__ target, float inv_n,
                      float *__restrict__ d, long long n, double *__restrict__ part) {
    __shared__ double sh[kT];
    const long long i = (long long)blockIdx.x * kT + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        const float r = out[i] - target[i];
        if (d) d[i] = r * inv_n;
        e = (double)r * r;
    }
    sh[threadIdx.x] = e;
    __syncthreads();
    for (int s = kT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum_doubles(const double *__restrict__ v, int n, double *__restrict__ out) {
    __shared__ double sh[kT];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kT) s += v[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int k = kT / 2; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Adam (Kingma and Ba 2015) on fp32 masters; wb (may be null) receives the bf16 copy
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, bf16 *__restrict__ wb, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = (bf16)wi;
}

__global__ void k_to_bf16(const float *__restrict__ a, bf16 *__restrict__ b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (bf16)a[i];
}

__global__ void k_from_bf16(const bf16 *__restrict__ a, float *__restrict__ b, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

int nblk(long long n, int t = kT) { return (int)((n + t - 1) / t); }

hipStream_t as_stream(void *st) { return (hipStream_t)st; }

int check_shape(int H, int W, int C) {
    if (H < 4 || H % 4 || W < 64 || W % 64) return sa_fail(SA_E_ARG, "U-Net: H must be a multiple of 4, W of 64");
    if (C < 32 || C % 32) return sa_fail(SA_E_ARG, "U-Net: channel counts must be multiples of 32");
    return SA_OK;
}

}  // namespace

// ---- C ABI (declared in include/stride_amd.h) ------------------------------------------------

extern "C" int sa_nn_conv3x3(void *stream, const void *x, const void *amask, const void *wt, const float *bias, void *y,
                             int N, int H, int W, int Cin, int Cout, int relu) {
    int rc;
    if ((rc = check_shape(H, W, Cin)) || (rc = check_shape(H, W, Cout))) return rc;
    if (!x || !wt || !y || N < 1) return sa_fail(SA_E_ARG, "U-Net conv: null buffer");
    const bool wide = Cout % 64 == 0;
    const int ncb = Cout / (wide ? 64 : 32);
    const dim3 grid(W / 64, H / 4, N * ncb);
    if (wide)
        k_conv3x3<4><<<grid, 256, 0, as_stream(stream)>>>((const bf16 *)x, (const bf16 *)amask, (const bf16 *)wt, bias,
                                                          (bf16 *)y, H, W, Cin, Cout, relu, ncb);
    else
        k_conv3x3<2><<<grid, 256, 0, as_stream(stream)>>>((const bf16 *)x, (const bf16 *)amask, (const bf16 *)wt, bias,
                                                          (bf16 *)y, H, W, Cin, Cout, relu, ncb);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_flip_weights(void *stream, const void *wt, void *wb, int Cin, int Cout) {
    if (!wt || !wb) return sa_fail(SA_E_ARG, "U-Net flip: null buffer");
    const int total = 9 * Cin * Cout;
    k_flip_weights<<<nblk(total), kT, 0, as_stream(stream)>>>((const bf16 *)wt, (bf16 *)wb, Cin, Cout);
    NTRY(hipGetLastError());
    return SA_OK;
}

// part: scratch of at least nsplit * 9 * Cin * Cout floats with nsplit = ceil(N H W / px_per_split)
extern "C" int sa_nn_wgrad3x3(void *stream, const void *dy, const void *ymask, const void *x, float *part, float *dw,
                              int N, int H, int W, int Cin, int Cout, int px_per_split, int accumulate) {
    int rc;
    if ((rc = check_shape(H, W, Cin)) || (rc = check_shape(H, W, Cout))) return rc;
    if (!dy || !x || !part || !dw || px_per_split < 32 || px_per_split % 32)
        return sa_fail(SA_E_ARG, "U-Net wgrad: bad arguments");
    const long long npx = (long long)N * H * W;
    const int nsplit = (int)((npx + px_per_split - 1) / px_per_split);
    hipStream_t st = as_stream(stream);
    const bool wide = Cin % 64 == 0 && Cout % 64 == 0;
    const int TC = wide ? 64 : 32;
    const int nci = Cin / TC;
    const dim3 grid(nsplit, 9, (Cout / TC) * nci);
    if (wide)
        k_wgrad<64><<<grid, 256, 0, st>>>((const bf16 *)dy, (const bf16 *)ymask, (const bf16 *)x, part, N, H, W, Cin,
                                          Cout, px_per_split, nci);
    else
        k_wgrad<32><<<grid, 256, 0, st>>>((const bf16 *)dy, (const bf16 *)ymask, (const bf16 *)x, part, N, H, W, Cin,
                                          Cout, px_per_split, nci);
    const long long n = 9LL * Cin * Cout;
    k_sum_splits<<<nblk(n), kT, 0, st>>>(part, nsplit, n, dw, accumulate);
    NTRY(hipGetLastError());
    return SA_OK;
}

// db[c] (+)= sum_p dy[p][c] [ymask[p][c] > 0]; part: ceil(npx / px_per_block) * C floats
extern "C" int sa_nn_bias_grad(void *stream, const void *dy, const void *ymask, float *part, float *db, long long npx,
                               int C, int px_per_block, int accumulate) {
    if (!dy || !part || !db || C < 1 || C > 1024 || px_per_block < 1) return sa_fail(SA_E_ARG, "U-Net bias grad: bad arguments");
    const int nb = (int)((npx + px_per_block - 1) / px_per_block);
    hipStream_t st = as_stream(stream);
    k_channel_partial<<<nb, C, 0, st>>>((const bf16 *)dy, (const bf16 *)ymask, npx, C, px_per_block, part);
    k_sum_splits<<<nblk(C), kT, 0, st>>>(part, nb, C, db, accumulate);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_avgpool(void *stream, const void *x, void *y, int N, int H, int W, int C) {
    if (H % 2 || W % 2 || C % 8) return sa_fail(SA_E_ARG, "U-Net pool: even H, W and C % 8 == 0");
    const long long n = (long long)N * (H / 2) * (W / 2) * (C / 8);
    k_avgpool<<<nblk(n), kT, 0, as_stream(stream)>>>((const bf16 *)x, (bf16 *)y, N, H, W, C);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_avgpool_back(void *stream, const void *dy, void *dx, int N, int H, int W, int C, int accumulate) {
    if (H % 2 || W % 2 || C % 8) return sa_fail(SA_E_ARG, "U-Net pool: even H, W and C % 8 == 0");
    const long long n = (long long)N * H * W * (C / 8);
    k_avgpool_back<<<nblk(n), kT, 0, as_stream(stream)>>>((const bf16 *)dy, (bf16 *)dx, N, H, W, C, accumulate);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_upsample(void *stream, const void *x, void *y, int N, int H, int W, int C, int Cy, int coff) {
    if (C % 8 || coff % 8 || coff + C > Cy) return sa_fail(SA_E_ARG, "U-Net upsample: bad channels");
    const long long n = (long long)N * 2 * H * 2 * W * (C / 8);
    k_upsample_into<<<nblk(n), kT, 0, as_stream(stream)>>>((const bf16 *)x, (bf16 *)y, N, H, W, C, Cy, coff);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_upsample_back(void *stream, const void *dy, void *dx, int N, int H, int W, int C, int Cy, int coff) {
    if (C % 8 || coff % 8 || coff + C > Cy) return sa_fail(SA_E_ARG, "U-Net upsample: bad channels");
    const long long n = (long long)N * H * W * (C / 8);
    k_upsample_back<<<nblk(n), kT, 0, as_stream(stream)>>>((const bf16 *)dy, (bf16 *)dx, N, H, W, C, Cy, coff);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_copy_channels(void *stream, const void *src, int Cs, int soff, void *dst, int Cd, int doff, int C,
                                   long long npx, int accumulate) {
    if (C % 8 || soff % 8 || doff % 8 || soff + C > Cs || doff + C > Cd)
        return sa_fail(SA_E_ARG, "U-Net channel copy: bad channels");
    k_copy_channels<<<nblk(npx * (C / 8)), kT, 0, as_stream(stream)>>>((const bf16 *)src, Cs, soff, (bf16 *)dst, Cd,
                                                                       doff, C, npx, accumulate);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_pack_input(void *stream, const float *g0, float s0, const float *g1, float s1, void *x,
                                long long npx, int C) {
    if (C % 8 || !g0 || !x) return sa_fail(SA_E_ARG, "U-Net input: C % 8 == 0 and g0, x given");
    k_pack_input<<<nblk(npx * (C / 8)), kT, 0, as_stream(stream)>>>(g0, s0, g1, s1, (bf16 *)x, npx, C);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_head(void *stream, const void *x, const float *w, const float *b, float *out, long long npx, int C) {
    if (C % 8) return sa_fail(SA_E_ARG, "U-Net head: C % 8 == 0");
    k_head<<<nblk(npx), kT, 0, as_stream(stream)>>>((const bf16 *)x, w, b, out, npx, C);
    NTRY(hipGetLastError());
    return SA_OK;
}

// dx = dout w (masked by x > 0); dw[c] (+)= sum_p dout x[p][c], db (+)= sum_p dout (dw has C + 1
// entries: the bias last); part: ceil(npx / px_per_block) * (C + 1) floats
extern "C" int sa_nn_head_back(void *stream, const float *dout, const float *w, const void *x, void *dx, float *part,
                               float *dw, long long npx, int C, int px_per_block, int accumulate) {
    if (C % 8 || C + 1 > 1024 || px_per_block < 1) return sa_fail(SA_E_ARG, "U-Net head: bad arguments");
    hipStream_t st = as_stream(stream);
    if (dx) k_head_back_data<<<nblk(npx * (C / 8)), kT, 0, st>>>(dout, w, (const bf16 *)x, (bf16 *)dx, npx, C);
    const int nb = (int)((npx + px_per_block - 1) / px_per_block);
    k_head_back_weights<<<nb, C + 1, 0, st>>>(dout, (const bf16 *)x, npx, C, px_per_block, part);
    k_sum_splits<<<nblk(C + 1), kT, 0, st>>>(part, nb, C + 1, dw, accumulate);
    NTRY(hipGetLastError());
    return SA_OK;
}

// loss = 0.5 sum (out - target)^2 * inv_n (returned in *loss, a host pointer, after a wait);
// d = (out - target) inv_n (may be null); part: ceil(n / 256) + 1 doubles
extern "C" int sa_nn_mse(void *stream, const float *out, const float *target, float *d, long long n, float inv_n,
                         double *part, double *loss) {
    if (!out || !target || !part || n < 1) return sa_fail(SA_E_ARG, "U-Net loss: bad arguments");
    hipStream_t st = as_stream(stream);
    const int nb = nblk(n);
    k_mse<<<nb, kT, 0, st>>>(out, target, inv_n, d, n, part);
    k_sum_doubles<<<1, kT, 0, st>>>(part, nb, part + nb);
    NTRY(hipGetLastError());
    if (loss) {
        double t = 0.0;
        NTRY(hipMemcpyAsync(&t, part + nb, sizeof(double), hipMemcpyDeviceToHost, st));
        NTRY(hipStreamSynchronize(st));
        *loss = 0.5 * t * inv_n;
    }
    return SA_OK;
}

extern "C" int sa_nn_adam(void *stream, float *w, const float *g, float *m, float *v, void *wb, long long n, float lr,
                          float b1, float b2, float eps, int step) {
    if (!w || !g || !m || !v || n < 1 || step < 1) return sa_fail(SA_E_ARG, "Adam: bad arguments");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    k_adam<<<nblk(n), kT, 0, as_stream(stream)>>>(w, g, m, v, (bf16 *)wb, n, lr, b1, b2, eps, c1, c2);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_to_bf16(void *stream, const float *a, void *b, long long n) {
    k_to_bf16<<<nblk(n), kT, 0, as_stream(stream)>>>(a, (bf16 *)b, n);
    NTRY(hipGetLastError());
    return SA_OK;
}

extern "C" int sa_nn_from_bf16(void *stream, const void *a, float *b, long long n) {
    k_from_bf16<<<nblk(n), kT, 0, as_stream(stream)>>>((const bf16 *)a, b, n);
    NTRY(hipGetLastError());
    return SA_OK;
}
