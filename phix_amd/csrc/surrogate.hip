// surrogate.hip -- learned surrogate for 2D periodic Cahn-Hilliard on gfx950.
//
// No reference counterpart (PhiX has no learned part). The network advances
// c(t) -> c(t + 100 dt):
//   x   = c - 0.5
//   a1  = relu(conv3x3(x;  1 -> 32))            FP32 VALU, BF16 NHWC out
//   a_l = relu(conv3x3(a_{l-1}; 32 -> 32)), l=2..5   BF16 MFMA implicit GEMM
//   d   = conv3x3(a5; 32 -> 1)                  VALU, FP32 out
//   c'  = c + d - mean(d)                       exact mean projection
// All convolutions use circular padding. Activations are BF16 (NHWC: the 32
// channels of a pixel are 64 contiguous bytes), accumulation is FP32.
//
// The 32->32 layers are implicit GEMMs on v_mfma_f32_16x16x32_bf16: for each
// of the 9 taps, K = 32 input channels is exactly one MFMA K step. A block
// computes a 16x16 pixel tile x 32 output channels with 4 waves (4 rows each),
// the (16+2)^2 input halo tile staged once in LDS, the weights held in VGPRs.
// Data gradients reuse the same kernel with flipped, transposed weights and a
// ReLU mask; weight gradients are a K = pixels MFMA GEMM per tap with
// per-block partials reduced in a fixed order (deterministic, no atomics).
//
// MFMA 16x16x32 bf16 lane layout (checked by tests/test_gpu_surrogate.py):
//   A (16x32): lane l holds A[l%16][8*(l/16) + j], j = 0..7
//   B (32x16): lane l holds B[8*(l/16) + j][l%16]
//   D (16x16): lane l holds D[4*(l/16) + r][l%16], r = 0..3
#include "common.hpp"

#include <hip/hip_bf16.h>

namespace phx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int CH = 32;
constexpr int NTAP = 9;
constexpr int PIXP = 40;          // LDS pixel pitch in bf16 (80 B: 16 B aligned, spreads banks)
constexpr int WG_BLOCKS = 512;    // persistent grid of the weight-gradient kernels
constexpr int MID_W = NTAP * CH * CH;   // 9216
constexpr int PART = MID_W + CH;        // per-block partial: dW + db

// parameter layout (FP32 master, PyTorch-style [out][in][kh][kw])
constexpr int P_W1 = 0, P_B1 = P_W1 + CH * 9;
constexpr int P_MID = P_B1 + CH;                 // layer l (0..3): w at P_MID + l*(MID_W+CH), b after it
constexpr int P_W6 = P_MID + 4 * (MID_W + CH), P_B6 = P_W6 + CH * 9;
constexpr int P_COUNT = P_B6 + 1;

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }

// ------------------------------------------------------------ weight pack --
// wfwd[l][tap][co][ci] = w_l[co][ci][kh][kw];  wbwd[l][tap'][ci][co] = w_l[co][ci][2-kh'][2-kw']
__global__ void k_pack(const float* __restrict__ P, __bf16* __restrict__ wfwd, __bf16* __restrict__ wbwd) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * MID_W) return;
    int l = t / MID_W, r = t % MID_W;
    int tap = r / (CH * CH), co = (r / CH) % CH, ci = r % CH;
    const float* w = P + P_MID + l * (MID_W + CH);
    wfwd[t] = f2bf(w[(co * CH + ci) * 9 + tap]);
    // backward: output channel = ci of the forward, input channel = co
    int tf = 8 - tap;
    wbwd[l * MID_W + (tap * CH + ci) * CH + co] = f2bf(w[(co * CH + ci) * 9 + tf]);
}

// --------------------------------------------------------- first layer --
__global__ __launch_bounds__(256) void k_conv_in_fwd(const float* __restrict__ c, const float* __restrict__ P,
                                                     __bf16* __restrict__ out, int B, int H, int W) {
    long long p = blockIdx.x * 256LL + threadIdx.x;
    if (p >= (long long)B * H * W) return;
    int x = (int)(p % W), y = (int)((p / W) % H);
    long long base = p - (long long)y * W - x;
    float v[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        int yy = wrap(y + t / 3 - 1, H), xx = wrap(x + t % 3 - 1, W);
        v[t] = c[base + (long long)yy * W + xx] - 0.5f;
    }
    __attribute__((aligned(16))) __bf16 o[CH];
#pragma unroll
    for (int co = 0; co < CH; ++co) {
        float s = P[P_B1 + co];
#pragma unroll
        for (int t = 0; t < 9; ++t) s += P[P_W1 + co * 9 + t] * v[t];
        o[co] = f2bf(fmaxf(s, 0.f));
    }
    uint4* dst = reinterpret_cast<uint4*>(out + p * CH);
    const uint4* src = reinterpret_cast<const uint4*>(o);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = src[q];
}

// ----------------------------------------------------------- last layer --
__global__ __launch_bounds__(256) void k_conv_out_fwd(const __bf16* __restrict__ a, const float* __restrict__ P,
                                                      float* __restrict__ d, int B, int H, int W) {
    long long p = blockIdx.x * 256LL + threadIdx.x;
    if (p >= (long long)B * H * W) return;
    int x = (int)(p % W), y = (int)((p / W) % H);
    long long base = p - (long long)y * W - x;
    float s = P[P_B6];
    for (int t = 0; t < 9; ++t) {
        int yy = wrap(y + t / 3 - 1, H), xx = wrap(x + t % 3 - 1, W);
        const uint4* src = reinterpret_cast<const uint4*>(a + (base + (long long)yy * W + xx) * CH);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint4 u = src[q];
            const __bf16* e = reinterpret_cast<const __bf16*>(&u);
#pragma unroll
            for (int j = 0; j < 8; ++j) s += P[P_W6 + (q * 8 + j) * 9 + t] * bf2f(e[j]);
        }
    }
    d[p] = s;
}

// dZ5[q][ci] = (sum_tap w6[ci][tap] * dd[q - off(tap)]) * (a5[q][ci] > 0)
__global__ __launch_bounds__(256) void k_conv_out_bwd(const float* __restrict__ dd, const __bf16* __restrict__ a5,
                                                      const float* __restrict__ P, __bf16* __restrict__ dz, int B,
                                                      int H, int W) {
    long long p = blockIdx.x * 256LL + threadIdx.x;
    if (p >= (long long)B * H * W) return;
    int x = (int)(p % W), y = (int)((p / W) % H);
    long long base = p - (long long)y * W - x;
    float g[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) {
        int yy = wrap(y - (t / 3 - 1), H), xx = wrap(x - (t % 3 - 1), W);
        g[t] = dd[base + (long long)yy * W + xx];
    }
    const uint4* am = reinterpret_cast<const uint4*>(a5 + p * CH);
    uint4* dst = reinterpret_cast<uint4*>(dz + p * CH);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint4 u = am[q];
        const __bf16* e = reinterpret_cast<const __bf16*>(&u);
        __attribute__((aligned(16))) __bf16 o[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            int ci = q * 8 + j;
            float s = 0.f;
#pragma unroll
            for (int t = 0; t < 9; ++t) s += P[P_W6 + ci * 9 + t] * g[t];
            o[j] = f2bf(bf2f(e[j]) > 0.f ? s : 0.f);
        }
        dst[q] = *reinterpret_cast<const uint4*>(o);
    }
}

// ------------------------------------------------ 32 -> 32 MFMA conv -----
// mode 0: out = relu(conv + bias)   mode 1: out = conv * (mask > 0)
__global__ __launch_bounds__(256) void k_conv_mid(const __bf16* __restrict__ in, const __bf16* __restrict__ wpk,
                                                  const float* __restrict__ bias, const __bf16* __restrict__ mask,
                                                  __bf16* __restrict__ out, int H, int W, int mode) {
    __shared__ __attribute__((aligned(16))) __bf16 S[18 * 18 * PIXP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16, n = blockIdx.z;
    const long long img = (long long)n * H * W;
    for (int t = tid; t < 18 * 18 * 4; t += 256) {
        int pi = t >> 2, q = t & 3;
        int yy = wrap(y0 - 1 + pi / 18, H), xx = wrap(x0 - 1 + pi % 18, W);
        uint4 v = *reinterpret_cast<const uint4*>(in + (img + (long long)yy * W + xx) * CH + q * 8);
        *reinterpret_cast<uint4*>(&S[pi * PIXP + q * 8]) = v;
    }
    bf16x8 bw[NTAP][2];
#pragma unroll
    for (int tap = 0; tap < NTAP; ++tap)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
            bw[tap][nb] = *reinterpret_cast<const bf16x8*>(wpk + (tap * CH + (lane & 15) + 16 * nb) * CH + 8 * (lane >> 4));
    __syncthreads();
    f32x4 acc[4][2];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tap = 0; tap < NTAP; ++tap) {
        const int dh = tap / 3, dw = tap % 3;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
            const int row = 4 * wv + mb;
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&S[((row + dh) * 18 + (lane & 15) + dw) * PIXP + 8 * (lane >> 4)]);
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
                acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[tap][nb], acc[mb][nb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
        const int y = y0 + 4 * wv + mb;
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) {
            const int co = (lane & 15) + 16 * nb;
            const float bco = mode == 0 ? bias[co] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int x = x0 + 4 * (lane >> 4) + r;
                const long long pix = img + (long long)y * W + x;
                float v = acc[mb][nb][r];
                if (mode == 0) v = fmaxf(v + bco, 0.f);
                else v = bf2f(mask[pix * CH + co]) > 0.f ? v : 0.f;
                out[pix * CH + co] = f2bf(v);
            }
        }
    }
}

// ----------------------------------------- 32 -> 32 weight gradient -------
// part[blk][tap][co][ci] = sum over the block's tiles of sum_p dz[p][co]*x[p+off(tap)][ci]
// part[blk][MID_W + co]  = sum over the block's tiles of sum_p dz[p][co]
__global__ __launch_bounds__(256) void k_conv_mid_wgrad(const __bf16* __restrict__ x, const __bf16* __restrict__ dz,
                                                        float* __restrict__ part, int B, int H, int W) {
    constexpr int XP = 18 * 18 + 4;   // row pitch of the transposed halo tile
    constexpr int ZP = 256 + 8;       // row pitch of the transposed dz tile (16 B aligned)
    __shared__ __attribute__((aligned(16))) __bf16 Xt[CH * XP];
    __shared__ __attribute__((aligned(16))) __bf16 Zt[CH * ZP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int tx = W / 16, ty = H / 16, ntiles = B * tx * ty;
    int taps[3] = {wv, wv + 4, wv == 0 ? 8 : -1};
    f32x4 acc[3][2][2];
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) acc[s][mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
    float dbs = 0.f;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int n = tile / (tx * ty), rem = tile % (tx * ty);
        const int y0 = (rem / tx) * 16, x0 = (rem % tx) * 16;
        const long long img = (long long)n * H * W;
        for (int t = tid; t < 18 * 18 * 4; t += 256) {
            int pi = t >> 2, q = t & 3;
            int yy = wrap(y0 - 1 + pi / 18, H), xx = wrap(x0 - 1 + pi % 18, W);
            uint4 v = *reinterpret_cast<const uint4*>(x + (img + (long long)yy * W + xx) * CH + q * 8);
            const __bf16* e = reinterpret_cast<const __bf16*>(&v);
#pragma unroll
            for (int j = 0; j < 8; ++j) Xt[(q * 8 + j) * XP + pi] = e[j];
        }
        for (int t = tid; t < 256 * 4; t += 256) {
            int pi = t >> 2, q = t & 3;
            const long long pix = img + (long long)(y0 + pi / 16) * W + x0 + pi % 16;
            uint4 v = *reinterpret_cast<const uint4*>(dz + pix * CH + q * 8);
            const __bf16* e = reinterpret_cast<const __bf16*>(&v);
#pragma unroll
            for (int j = 0; j < 8; ++j) Zt[(q * 8 + j) * ZP + pi] = e[j];
        }
        __syncthreads();
        if (tid < CH) {
            float s = 0.f;
            for (int p = 0; p < 256; ++p) s += bf2f(Zt[tid * ZP + p]);
            dbs += s;
        }
        for (int kk = 0; kk < 256; kk += 32) {
            const int pbase = kk + 8 * (lane >> 4);        // 8 consecutive pixels in one tile row
            const int prow = pbase >> 4, pcol = pbase & 15;
            bf16x8 a[2];
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
                a[mb] = *reinterpret_cast<const bf16x8*>(&Zt[((lane & 15) + 16 * mb) * ZP + pbase]);
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                const int tap = taps[s];
                if (tap < 0) continue;
                const int dh = tap / 3, dw = tap % 3;
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) {
                    const __bf16* src = &Xt[((lane & 15) + 16 * nb) * XP + (prow + dh) * 18 + pcol + dw];
                    bf16x8 b;
#pragma unroll
                    for (int j = 0; j < 8; ++j) b[j] = src[j];
#pragma unroll
                    for (int mb = 0; mb < 2; ++mb)
                        acc[s][mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[mb], b, acc[s][mb][nb], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    float* pp = part + (long long)blockIdx.x * PART;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int tap = taps[s];
        if (tap < 0) continue;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int co = 16 * mb + 4 * (lane >> 4) + r, ci = 16 * nb + (lane & 15);
                    pp[(tap * CH + co) * CH + ci] = acc[s][mb][nb][r];
                }
    }
    if (tid < CH) pp[MID_W + tid] = dbs;
}

// dW of layer l in the master layout [co][ci][kh][kw], db after it: fixed-order sum of partials.
__global__ void k_reduce_mid(const float* __restrict__ part, int nblk, float* __restrict__ g) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= PART) return;
    float s = 0.f;
#pragma unroll 16
    for (int b = 0; b < nblk; ++b) s += part[(long long)b * PART + t];
    if (t < MID_W) {
        int tap = t / (CH * CH), co = (t / CH) % CH, ci = t % CH;
        g[(co * CH + ci) * 9 + tap] = s;
    } else {
        g[MID_W + (t - MID_W)] = s;
    }
}

// ------------------------------------- first/last layer weight gradients --
// mode 0 (first layer): part[blk][co*9+tap] = sum_p x[p+off(tap)] * dz1[p][co]; part[blk][288+co] = sum_p dz1[p][co]
// mode 1 (last layer):  part[blk][ci*9+tap] = sum_p a5[p+off(tap)][ci] * dd[p]; part[blk][288] = sum_p dd[p]
// Persistent grid over 16x16 pixel tiles (the tile order of k_conv_mid_wgrad).
// The tile's halo and its per-pixel factor are staged in LDS. Thread t owns
// output t and sums the tile's 256 pixels in row order, so each block's
// partial is a fixed-order sum.
template <int MODE>
__global__ __launch_bounds__(320) void k_edge_wgrad(const float* __restrict__ c, const __bf16* __restrict__ act,
                                                    const float* __restrict__ dd, float* __restrict__ part, int B,
                                                    int H, int W) {
    constexpr int XP = 18 * 18 + 4;   // row pitch of the transposed a5 halo tile (mode 1)
    constexpr int ZP = 256 + 8;       // row pitch of the transposed dz1 tile (mode 0)
    __shared__ __attribute__((aligned(16))) __bf16 Ct[CH * (MODE == 0 ? ZP : XP)];
    __shared__ __attribute__((aligned(16))) float Fs[18 * 18];   // mode 0: x halo; mode 1: dd tile (256)
    const int t = threadIdx.x;   // 0..319: (channel, tap) pairs 0..287, bias 288..319
    const int tx = W / 16, ty = H / 16, ntiles = B * tx * ty;
    const int ch = t < 288 ? t / 9 : t - 288, tap = t < 288 ? t % 9 : 0;
    const int off = (tap / 3) * 18 + tap % 3;   // halo index of pixel (0, 0) shifted by the tap
    float s = 0.f;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int n = tile / (tx * ty), rem = tile % (tx * ty);
        const int y0 = (rem / tx) * 16, x0 = (rem % tx) * 16;
        const long long img = (long long)n * H * W;
        if (MODE == 0) {
            for (int i = t; i < 18 * 18; i += 320) {
                const int yy = wrap(y0 - 1 + i / 18, H), xx = wrap(x0 - 1 + i % 18, W);
                Fs[i] = c[img + (long long)yy * W + xx] - 0.5f;
            }
            for (int i = t; i < 256 * 4; i += 320) {
                const int pi = i >> 2, q = i & 3;
                const long long pix = img + (long long)(y0 + pi / 16) * W + x0 + pi % 16;
                uint4 v = *reinterpret_cast<const uint4*>(act + pix * CH + q * 8);
                const __bf16* e = reinterpret_cast<const __bf16*>(&v);
#pragma unroll
                for (int j = 0; j < 8; ++j) Ct[(q * 8 + j) * ZP + pi] = e[j];
            }
        } else {
            for (int i = t; i < 18 * 18 * 4; i += 320) {
                const int pi = i >> 2, q = i & 3;
                const int yy = wrap(y0 - 1 + pi / 18, H), xx = wrap(x0 - 1 + pi % 18, W);
                uint4 v = *reinterpret_cast<const uint4*>(act + (img + (long long)yy * W + xx) * CH + q * 8);
                const __bf16* e = reinterpret_cast<const __bf16*>(&v);
#pragma unroll
                for (int j = 0; j < 8; ++j) Ct[(q * 8 + j) * XP + pi] = e[j];
            }
            if (t < 256) Fs[t] = dd[img + (long long)(y0 + t / 16) * W + x0 + t % 16];
        }
        __syncthreads();
        if (MODE == 0) {
            const __bf16* z = &Ct[ch * ZP];
            if (t < 288) {
                for (int p8 = 0; p8 < 32; ++p8) {
                    const bf16x8 zv = *reinterpret_cast<const bf16x8*>(z + p8 * 8);
                    const float* xr = &Fs[(p8 >> 1) * 18 + (p8 & 1) * 8 + off];
#pragma unroll
                    for (int j = 0; j < 8; ++j) s += xr[j] * bf2f(zv[j]);
                }
            } else {
                for (int p8 = 0; p8 < 32; ++p8) {
                    const bf16x8 zv = *reinterpret_cast<const bf16x8*>(z + p8 * 8);
#pragma unroll
                    for (int j = 0; j < 8; ++j) s += bf2f(zv[j]);
                }
            }
        } else {
            if (t < 288) {
                const __bf16* a = &Ct[ch * XP + off];
                for (int p8 = 0; p8 < 32; ++p8) {
                    const __bf16* ar = a + (p8 >> 1) * 18 + (p8 & 1) * 8;
                    const float* dr = &Fs[p8 * 8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) s += bf2f(ar[j]) * dr[j];
                }
            } else if (t == 288) {
                for (int p = 0; p < 256; ++p) s += Fs[p];
            }
        }
        __syncthreads();
    }
    const int nout = MODE == 0 ? 320 : 289;
    if (t < nout) part[(long long)blockIdx.x * 320 + t] = s;
}

// Fixed-order sums of the per-block partials (block 0 first); the unrolled
// loop keeps many loads in flight while the adds stay in order.
__global__ void k_reduce_edge(const float* __restrict__ part, int nblk, int nout, float* __restrict__ g) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nout) return;
    float s = 0.f;
#pragma unroll 16
    for (int b = 0; b < nblk; ++b) s += part[(long long)b * 320 + t];
    g[t] = s;
}

// --------------------------------------------- projection, loss, optimiser --
// Per-segment FP64 compensated sums of v: segment s is v[s*len, (s+1)*len).
// Two levels with a fixed reduction tree (bitwise deterministic): SUM_BLOCKS
// workgroups per segment each reduce one contiguous share (Neumaier sums per
// thread, then a compensated tree), and one thread per segment combines the
// SUM_BLOCKS partial (sum, compensation) pairs in order.
constexpr int SUM_BLOCKS = 64;

__device__ __forceinline__ void neu_add(double& s, double& cmp, double x) {
    const double t = s + x;
    cmp += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

__global__ __launch_bounds__(256) void k_sum_partial(const float* __restrict__ v, long long len,
                                                     double* __restrict__ part) {
    const long long share = (len + SUM_BLOCKS - 1) / SUM_BLOCKS;
    const long long q0 = (long long)blockIdx.x * share;
    const long long q1 = min(q0 + share, len);
    const float* s0 = v + (long long)blockIdx.y * len;
    double s = 0.0, cmp = 0.0;
    for (long long q = q0 + threadIdx.x; q < q1; q += 256) neu_add(s, cmp, (double)s0[q]);
    __shared__ double S[256], Cc[256];
    S[threadIdx.x] = s;
    Cc[threadIdx.x] = cmp;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            double a = S[threadIdx.x], b = S[threadIdx.x + h], t = a + b;
            double e = fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a;
            S[threadIdx.x] = t;
            Cc[threadIdx.x] += Cc[threadIdx.x + h] + e;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* pp = part + 2 * ((long long)blockIdx.y * SUM_BLOCKS + blockIdx.x);
        pp[0] = S[0];
        pp[1] = Cc[0];
    }
}

__global__ void k_sum_final(const double* __restrict__ part, int nseg, double* __restrict__ out) {
    const int sg = blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= nseg) return;
    const double* pp = part + 2LL * sg * SUM_BLOCKS;
    double s = 0.0, cmp = 0.0;
    for (int k = 0; k < SUM_BLOCKS; ++k) {
        neu_add(s, cmp, pp[2 * k]);
        cmp += pp[2 * k + 1];
    }
    out[sg] = s + cmp;
}

static void seg_sums(const float* v, long long len, int nseg, double* part, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_partial, dim3(SUM_BLOCKS, nseg), dim3(256), 0, st, v, len, part);
    hipLaunchKernelGGL(k_sum_final, dim3((nseg + 63) / 64), dim3(64), 0, st, part, nseg, out);
}

__global__ void k_project(const float* __restrict__ c, const float* __restrict__ d, const double* __restrict__ dsum,
                          int HW, long long n, float* __restrict__ out) {
    long long p = blockIdx.x * 256LL + threadIdx.x;
    if (p >= n) return;
    const double mean = dsum[p / HW] / HW;
    out[p] = (float)((double)c[p] + ((double)d[p] - mean));
}

// g = 2*(out - tgt)/n; e2 = (out - tgt)^2 for the loss
__global__ void k_loss_grad(const float* __restrict__ out, const float* __restrict__ tgt, long long n,
                            float* __restrict__ g, float* __restrict__ e2) {
    long long p = blockIdx.x * 256LL + threadIdx.x;
    if (p >= n) return;
    float r = out[p] - tgt[p];
    g[p] = 2.f * r / (float)n;
    e2[p] = r * r;
}

// dd = g - mean_sample(g)  (adjoint of the mean projection)
__global__ void k_unproject(const float* __restrict__ g, const double* __restrict__ gsum, int HW, long long n,
                            float* __restrict__ dd) {
    long long p = blockIdx.x * 256LL + threadIdx.x;
    if (p >= n) return;
    dd[p] = (float)((double)g[p] - gsum[p / HW] / HW);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float gg = g[t];
    float mm = b1 * m[t] + (1.f - b1) * gg;
    float vv = b2 * v[t] + (1.f - b2) * gg * gg;
    m[t] = mm;
    v[t] = vv;
    p[t] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// ------------------------------------------- fused inference forward -----
// One launch evaluates layers 1-6 for a 16x16 output tile: x = c - 0.5 on the
// tile plus a 6-pixel ring (28x28, circular wrap), a1 on 26x26, a2..a5 on
// 24..18, d on the tile; each layer's activations stay in LDS (BF16, the NHWC
// pixel pitch of k_conv_mid). With circular padding every layer is
// translation invariant, so a ring pixel's activation equals the activation
// at its wrapped position.
//
// Every layer runs on the MFMA in the transposed form D^T = W^T A^T (rows =
// output channels, columns = 16 pixels), so a lane ends with 4 consecutive
// output channels of one pixel:
//   * layer 1 (1 -> 32): FP32 operands on v_mfma_f32_16x16x4_f32, the 9 taps
//     in three K steps of 4 (FP32 precision, as k_conv_in_fwd, another
//     summation order);
//   * layers 2-5 (32 -> 32): v_mfma_f32_16x16x32_bf16, one per tap and 16
//     output channels, K = the 32 input channels; the weights sit in VGPRs,
//     each layer's loaded from global memory during the layer before it;
//   * layer 6 (32 -> 1): A rows 0 and 1 hold the high and low BF16 parts of
//     the FP32 weights (~16 significant bits), d = D[0] + D[1].
// Biases start the accumulators. The region is walked as a flat list of
// pixels in blocks of 16 (the MFMA's N), two blocks per wave at a time. The
// tile's sum of d is formed in a fixed order with compensation (per lane,
// then a halving tree over the 16 lanes, then the 8 waves in order).
// Inference only: no activations are kept for a backward pass. The result
// follows the layer-by-layer forward to within FP32 summation order (the
// bias first, the MFMAs' internal order) and the split of layer 6; it is not
// the same bits.
constexpr int FT = 16, FNT = 512;
constexpr int FR0 = FT + 12;                    // 28: the x region
constexpr int FRA = FT + 10;                    // 26: the a1 region
constexpr int NX0 = FR0 * FR0;
constexpr int ACT0_PIX = FRA * FRA;             // a1 (26^2), a3 (22^2), a5 (18^2)
constexpr int ACT1_PIX = (FT + 8) * (FT + 8);   // a2 (24^2), a4 (20^2)
// layer 6's weights split into BF16 high and low parts, [tap][high, low][ci]
constexpr int FRAG6 = NTAP * 2 * CH, FRAG_N = FRAG6;

__global__ void k_pack_frag(const float* __restrict__ P, __bf16* __restrict__ frag) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < FRAG6) {
        const int tap = t / (2 * CH), part = (t / CH) % 2, ci = t % CH;
        const float w = P[P_W6 + ci * 9 + tap];
        const __bf16 hi = f2bf(w);
        frag[t] = part == 0 ? hi : f2bf(w - bf2f(hi));
    }
}

// compensated addition of (b, cb) into (a, ca)
__device__ __forceinline__ void comp_add(double& a, double& ca, double b, double cb) {
    const double t = a + b;
    const double e = fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a;
    a = t;
    ca = ca + cb + e;
}

// One step of the tile tree of k_sur_project_tiles: (a, ca) <- (a, ca) (+) (b, cb).
__device__ __forceinline__ void tree_add(double& a, double& ca, double b, double cb) {
    const double t = a + b;
    const double e = fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a;
    a = t;
    ca += cb + e;
}

// With dprev set, the input is the previous jump's unprojected state: the
// kernel first forms that jump's projection u = c + dprev - mean(dprev) at
// every point it loads, the mean from the previous launch's tile partials
// tprev with k_sur_project_tiles's operations in its order (so u is that
// kernel's bits), and writes u on its own tile to cproj.
__global__ __launch_bounds__(FNT) void k_sur_fused_fwd(const float* __restrict__ c, const float* __restrict__ dprev,
                                                       const double* __restrict__ tprev, int ntile,
                                                       float* __restrict__ cproj, const float* __restrict__ cadd,
                                                       const float* __restrict__ P,
                                                       const __bf16* __restrict__ wpk,
                                                       const __bf16* __restrict__ frag, float* __restrict__ d,
                                                       double* __restrict__ tpart, int H, int W) {
    __shared__ __attribute__((aligned(16))) __bf16 ACT0[ACT0_PIX * PIXP];
    __shared__ __attribute__((aligned(16))) __bf16 ACT1[ACT1_PIX * PIXP];
    __shared__ float X[NX0];
    __shared__ float W1[CH * 10];   // layer 1: w [co][tap], then b
    __shared__ __attribute__((aligned(16))) __bf16 W6S[FRAG6];
    __shared__ float BM[4 * CH];
    __shared__ double RW[2 * (FNT / 64)];
    __shared__ double MEAN;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, g = lane >> 4, lx = lane & 15;
    const int x0 = blockIdx.x * FT, y0 = blockIdx.y * FT;
    const long long img = (long long)blockIdx.z * H * W;
    // ---- prologue: global loads, then LDS stores
    float cin[2], dv0[2], av[2];
    long long at[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int t = min(tid + k * FNT, NX0 - 1);
        const int yy = wrap(y0 - 6 + t / FR0, H), xx = wrap(x0 - 6 + t % FR0, W);
        at[k] = img + (long long)yy * W + xx;
        cin[k] = c[at[k]];
        dv0[k] = dprev ? dprev[at[k]] : 0.f;
        av[k] = cadd ? cadd[at[k]] : 0.f;
    }
    if (dprev) {
        // the previous jump's mean of d: wave 0 holds k_sur_project_tiles's
        // 256 per-thread sums in lanes (t, t + 64, t + 128, t + 192) and runs
        // its tree (pairs h = 128, 64 in registers, then 32..1 across lanes)
        if (wv == 0) {
            const double* tp = tprev + 2LL * blockIdx.z * ntile;
            double s4[4], c4[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double sq = 0.0, cq = 0.0;
                for (int k = lane + 64 * q; k < ntile; k += 256) {
                    neu_add(sq, cq, tp[2 * k]);
                    cq += tp[2 * k + 1];
                }
                s4[q] = sq;
                c4[q] = cq;
            }
            tree_add(s4[0], c4[0], s4[2], c4[2]);
            tree_add(s4[1], c4[1], s4[3], c4[3]);
            tree_add(s4[0], c4[0], s4[1], c4[1]);
            double sv = s4[0], cc = c4[0];
#pragma unroll
            for (int h = 32; h > 0; h >>= 1) {
                const double so = __shfl_down(sv, h, 64), co = __shfl_down(cc, h, 64);
                tree_add(sv, cc, so, co);
            }
            if (lane == 0) MEAN = (sv + cc) / ((double)H * W);
        }
        __syncthreads();
        const double mean = MEAN;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            cin[k] = (float)((double)cin[k] + ((double)dv0[k] - mean));
            if (cadd) cin[k] = cin[k] + av[k];   // a correction of the projected state (FP32 add)
            const int t = tid + k * FNT, ty = t / FR0 - 6, tx = t % FR0 - 6;
            if (t < NX0 && ty >= 0 && ty < FT && tx >= 0 && tx < FT) cproj[at[k]] = cin[k];
        }
    }
    float xv[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) xv[k] = cin[k] - 0.5f;
    const uint4 w6v = tid < FRAG6 / 8 ? reinterpret_cast<const uint4*>(frag)[tid] : uint4{0, 0, 0, 0};
    const float bmv = tid < 4 * CH ? P[P_MID + (tid / CH) * (MID_W + CH) + MID_W + tid % CH] : 0.f;
    const float w1v = tid < CH * 10 ? P[P_W1 + tid] : 0.f;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int t = tid + k * FNT;
        if (t < NX0) X[t] = xv[k];
    }
    if (tid < CH * 10) W1[tid] = w1v;
    if (tid < FRAG6 / 8) reinterpret_cast<uint4*>(W6S)[tid] = w6v;
    if (tid < 4 * CH) BM[tid] = bmv;
    __syncthreads();
    // layers 2-5 keep their weights in VGPRs; a layer's are loaded during the
    // layer before it, inside that layer's first block step, so the loads are
    // issued early and waited for at their use
    bf16x8 bw[NTAP][2], bwn[NTAP][2];
    auto load_w = [&](int l, bf16x8 (&r)[NTAP][2]) {
        const __bf16* wl = wpk + l * MID_W;
#pragma unroll
        for (int tap = 0; tap < NTAP; ++tap)
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
                r[tap][mb] = *reinterpret_cast<const bf16x8*>(wl + (tap * CH + lx + 16 * mb) * CH + 8 * g);
    };
    // ---- layer 1 (1 -> 32) on the 26x26 region in FP32 on
    // v_mfma_f32_16x16x4_f32: D^T (16 channels x 16 pixels) = W^T (16 x 12)
    // X (12 x 16), K = the 9 taps in three steps of 4 (the weights of slots
    // 9-11 are 0). Lane (g, lx) holds A[lx][g] = w[lx + 16 mb][4 j + g] and
    // B[g][lx] = x of pixel lx at tap 4 j + g.
    {
        constexpr int npix = FRA * FRA, nblk = (npix + 15) / 16;
        float a1f[2][3], b1v[2][4];
        int toff[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int t = 4 * j + g;
            toff[j] = t < 9 ? (t / 3) * FR0 + t % 3 : 0;
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) a1f[mb][j] = t < 9 ? W1[(lx + 16 * mb) * 9 + t] : 0.f;
        }
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) b1v[mb][r] = W1[CH * 9 + 16 * mb + 4 * g + r];
        for (int b = wv; b < nblk; b += FNT / 64) {
            if (b == wv) load_w(0, bwn);
            const int q = 16 * b + lx, qc = min(q, npix - 1);
            const int py = qc / FRA, px = qc - py * FRA, base = py * FR0 + px;
            float xb[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) xb[j] = X[base + toff[j]];
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                f32x4 acc = {b1v[mb][0], b1v[mb][1], b1v[mb][2], b1v[mb][3]};
#pragma unroll
                for (int j = 0; j < 3; ++j) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a1f[mb][j], xb[j], acc, 0, 0, 0);
                if (q < npix) {
                    __attribute__((aligned(8))) __bf16 o[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) o[r] = f2bf(fmaxf(acc[r], 0.f));
                    *reinterpret_cast<uint2*>(&ACT0[q * PIXP + 16 * mb + 4 * g]) = *reinterpret_cast<const uint2*>(o);
                }
            }
        }
    }
    // ---- layers 2-5 (32 -> 32): input region rin, output region rin - 2
#pragma unroll
    for (int l = 0; l < 4; ++l) {
#pragma unroll
        for (int tap = 0; tap < NTAP; ++tap)
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) bw[tap][mb] = bwn[tap][mb];
        float bias[2][4];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int r = 0; r < 4; ++r) bias[mb][r] = BM[l * CH + 16 * mb + 4 * g + r];
        if (l == 0) __syncthreads();   // a1 complete
        const int rin = FRA - 2 * l, rout = rin - 2, npix = rout * rout, nblk = (npix + 15) / 16;
        const __bf16* in = (l & 1) ? ACT1 : ACT0;
        __bf16* out = (l & 1) ? ACT0 : ACT1;
        for (int nb0 = wv; nb0 < nblk; nb0 += 16) {
            if (l < 3 && nb0 == wv) load_w(l + 1, bwn);
            const bool two = nb0 + 8 < nblk;
            int pix[2], base[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                pix[k] = (nb0 + 8 * k) * 16 + lx;
                const int pc = min(pix[k], npix - 1);
                const int py = pc / rout, px = pc - py * rout;
                base[k] = (py * rin + px) * PIXP + 8 * g;
            }
            f32x4 acc[2][2];
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) acc[k][mb] = f32x4{bias[mb][0], bias[mb][1], bias[mb][2], bias[mb][3]};
            if (two) {
                bf16x8 bb[NTAP][2];
#pragma unroll
                for (int tap = 0; tap < NTAP; ++tap)
#pragma unroll
                    for (int k = 0; k < 2; ++k)
                        bb[tap][k] = *reinterpret_cast<const bf16x8*>(&in[base[k] + ((tap / 3) * rin + tap % 3) * PIXP]);
#pragma unroll
                for (int tap = 0; tap < NTAP; ++tap)
#pragma unroll
                    for (int k = 0; k < 2; ++k)
#pragma unroll
                        for (int mb = 0; mb < 2; ++mb)
                            acc[k][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[tap][mb], bb[tap][k], acc[k][mb], 0, 0, 0);
            } else {
                bf16x8 bb[NTAP];
#pragma unroll
                for (int tap = 0; tap < NTAP; ++tap)
                    bb[tap] = *reinterpret_cast<const bf16x8*>(&in[base[0] + ((tap / 3) * rin + tap % 3) * PIXP]);
#pragma unroll
                for (int tap = 0; tap < NTAP; ++tap)
#pragma unroll
                    for (int mb = 0; mb < 2; ++mb)
                        acc[0][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[tap][mb], bb[tap], acc[0][mb], 0, 0, 0);
            }
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                if (pix[k] >= npix || (k == 1 && !two)) continue;
#pragma unroll
                for (int mb = 0; mb < 2; ++mb) {
                    __attribute__((aligned(8))) __bf16 o[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) o[r] = f2bf(fmaxf(acc[k][mb][r], 0.f));
                    *reinterpret_cast<uint2*>(&out[pix[k] * PIXP + 16 * mb + 4 * g]) = *reinterpret_cast<const uint2*>(o);
                }
            }
        }
        __syncthreads();
    }
    // ---- layer 6 (32 -> 1): wave w computes tile rows w and w + 8 from a5
    // (18x18, in ACT0); lanes 0-15 hold D rows 0 and 1
    float dv[2];
    {
        bf16x8 w6[NTAP];
#pragma unroll
        for (int tap = 0; tap < NTAP; ++tap)
            w6[tap] = lx < 2 ? *reinterpret_cast<const bf16x8*>(&W6S[(tap * 2 + lx) * CH + 8 * g]) : bf16x8{};
        const float b6 = P[P_B6];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int row = wv + 8 * k;
            const int base = (row * (FT + 2) + lx) * PIXP + 8 * g;
            f32x4 acc = {g == 0 ? b6 : 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int tap = 0; tap < NTAP; ++tap) {
                const bf16x8 bx = *reinterpret_cast<const bf16x8*>(&ACT0[base + ((tap / 3) * (FT + 2) + tap % 3) * PIXP]);
                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w6[tap], bx, acc, 0, 0, 0);
            }
            dv[k] = acc[0] + acc[1];
            if (g == 0) d[img + (long long)(y0 + row) * W + x0 + lx] = dv[k];
        }
    }
    // ---- the tile's sum of d, compensated, in a fixed order
    double sv = (double)dv[0], cv = 0.0;
    comp_add(sv, cv, (double)dv[1], 0.0);
#pragma unroll
    for (int h = 8; h > 0; h >>= 1) {
        const double so = __shfl_down(sv, h, 16), co = __shfl_down(cv, h, 16);
        comp_add(sv, cv, so, co);
    }
    if (lane == 0) {
        RW[2 * wv] = sv;
        RW[2 * wv + 1] = cv;
    }
    __syncthreads();
    if (tid == 0) {
        double s = RW[0], cc = RW[1];
#pragma unroll
        for (int w = 1; w < FNT / 64; ++w) comp_add(s, cc, RW[2 * w], RW[2 * w + 1]);
        const long long tile = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        tpart[2 * tile] = s;
        tpart[2 * tile + 1] = cc;
    }
}

// c' = c + d - mean(d) per sample, the mean from the fused kernel's per-tile
// compensated sums (ntile of them per sample, summed in tile order by every
// block alike).
__global__ __launch_bounds__(256) void k_sur_project_tiles(const float* __restrict__ c, const float* __restrict__ d,
                                                           const double* __restrict__ tpart, int ntile, int HW,
                                                           long long n, float* __restrict__ out,
                                                           const float* __restrict__ cadd = nullptr) {
    __shared__ double S[256], Cc[256];
    const long long p0 = blockIdx.x * 256LL;
    const int smp = (int)(p0 / HW);   // HW is a multiple of 256: one sample per block
    const double* tp = tpart + 2LL * smp * ntile;
    double s = 0.0, cmp = 0.0;
    for (int k = threadIdx.x; k < ntile; k += 256) {
        neu_add(s, cmp, tp[2 * k]);
        cmp += tp[2 * k + 1];
    }
    S[threadIdx.x] = s;
    Cc[threadIdx.x] = cmp;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            const double a = S[threadIdx.x], b = S[threadIdx.x + h], t = a + b;
            const double e = fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a;
            S[threadIdx.x] = t;
            Cc[threadIdx.x] += Cc[threadIdx.x + h] + e;
        }
        __syncthreads();
    }
    const double mean = (S[0] + Cc[0]) / HW;
    const long long p = p0 + threadIdx.x;
    if (p < n) {
        float v = (float)((double)c[p] + ((double)d[p] - mean));
        if (cadd) v = v + cadd[p];   // a correction of the projected state (FP32 add)
        out[p] = v;
    }
}

// --------------------------------------------------------------- host ----
struct SurWS {
    // carved from the caller's workspace
    __bf16* act[5];   // a1..a5
    __bf16* dz[2];    // gradient ping-pong (BF16 NHWC)
    float* d;         // last-layer output / its gradient
    float* g;         // dL/dout
    float* e2;
    double* sums;     // per-sample sums (B)
    double* loss;     // 1
    double* spart;    // SUM_BLOCKS * 2 * B partial sums
    double* tpart;    // fused forward: 2 per 16x16 tile
    float* d2;        // rollout: the second d buffer
    double* tpart2;   // rollout: the second tile-partial buffer
    float* u[2];      // rollout: projected states between jumps
    float* part;      // WG_BLOCKS * PART
    __bf16* wfwd;     // 4 * MID_W
    __bf16* wbwd;     // 4 * MID_W
    __bf16* frag;     // FRAG_N: operand tables of the fused forward
};

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t ws_layout(int B, int H, int W, SurWS* w, char* base) {
    const size_t npix = (size_t)B * H * W;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    SurWS s;
    for (int l = 0; l < 5; ++l) s.act[l] = (__bf16*)take(npix * CH * 2);
    for (int l = 0; l < 2; ++l) s.dz[l] = (__bf16*)take(npix * CH * 2);
    s.d = (float*)take(npix * 4);
    s.g = (float*)take(npix * 4);
    s.e2 = (float*)take(npix * 4);
    s.sums = (double*)take((size_t)B * 8 + 8);
    s.loss = (double*)take(16);
    s.spart = (double*)take((size_t)SUM_BLOCKS * 2 * B * 8);
    s.tpart = (double*)take((size_t)2 * B * (H / 16) * (W / 16) * 8);
    s.d2 = (float*)take(npix * 4);
    s.tpart2 = (double*)take((size_t)2 * B * (H / 16) * (W / 16) * 8);
    for (int k = 0; k < 2; ++k) s.u[k] = (float*)take(npix * 4);
    s.part = (float*)take((size_t)WG_BLOCKS * PART * 4);
    s.wfwd = (__bf16*)take(4 * MID_W * 2);
    s.wbwd = (__bf16*)take(4 * MID_W * 2);
    s.frag = (__bf16*)take(FRAG_N * 2);
    if (w) *w = s;
    return off;
}

static int check_shape(int B, int H, int W, const char* who) {
    if (B < 1 || H < 16 || W < 16 || H % 16 || W % 16)
        return fail(PHX_ERR_INVALID, std::string(who) + ": need B >= 1 and H, W positive multiples of 16");
    if ((long long)B * H * W > (1LL << 31) / CH)
        return fail(PHX_ERR_UNSUPPORTED, std::string(who) + ": batch too large");
    return PHX_OK;
}

static int sur_forward(const float* P, SurWS& S, const float* c, float* out, int B, int H, int W, int flags,
                       hipStream_t st) {
    const long long n = (long long)B * H * W;
    const int nb = (int)((n + 255) / 256);
    if (!(flags & PHX_SUR_REUSE_WEIGHTS)) {
        hipLaunchKernelGGL(k_pack, dim3((4 * MID_W + 255) / 256), dim3(256), 0, st, P, S.wfwd, S.wbwd);
        hipLaunchKernelGGL(k_pack_frag, dim3((FRAG_N + 255) / 256), dim3(256), 0, st, P, S.frag);
    }
    if (flags & PHX_SUR_FUSED) {
        const int ntile = (H / FT) * (W / FT);
        hipLaunchKernelGGL(k_sur_fused_fwd, dim3(W / FT, H / FT, B), dim3(FNT), 0, st, c, (const float*)nullptr,
                           (const double*)nullptr, 0, (float*)nullptr, (const float*)nullptr, P, S.wfwd, S.frag, S.d,
                           S.tpart, H, W);
        hipLaunchKernelGGL(k_sur_project_tiles, dim3(nb), dim3(256), 0, st, c, S.d, S.tpart, ntile, H * W, n, out);
        PHX_LAUNCH_CHECK("phx_sur_forward");
        return PHX_OK;
    } else {
        hipLaunchKernelGGL(k_conv_in_fwd, dim3(nb), dim3(256), 0, st, c, P, S.act[0], B, H, W);
        for (int l = 0; l < 4; ++l)
            hipLaunchKernelGGL(k_conv_mid, dim3(W / 16, H / 16, B), dim3(256), 0, st, S.act[l], S.wfwd + l * MID_W,
                               P + P_MID + l * (MID_W + CH) + MID_W, (const __bf16*)nullptr, S.act[l + 1], H, W, 0);
        hipLaunchKernelGGL(k_conv_out_fwd, dim3(nb), dim3(256), 0, st, S.act[4], P, S.d, B, H, W);
    }
    seg_sums(S.d, (long long)H * W, B, S.spart, S.sums, st);
    hipLaunchKernelGGL(k_project, dim3(nb), dim3(256), 0, st, c, S.d, S.sums, H * W, n, out);
    PHX_LAUNCH_CHECK("phx_sur_forward");
    return PHX_OK;
}

// `jumps` fused forwards from c, each projection formed by the next launch
// as it loads its input (k_sur_fused_fwd with dprev), the last by
// k_sur_project_tiles: the same bits as `jumps` calls of sur_forward with
// PHX_SUR_FUSED, with one launch per jump instead of two.
static int sur_rollout(const float* P, SurWS& S, const float* c, float* out, int B, int H, int W, int jumps,
                       int flags, hipStream_t st, float* states = nullptr, int every = 0,
                       const float* corr = nullptr) {
    const long long n = (long long)B * H * W;
    const int nb = (int)((n + 255) / 256), ntile = (H / FT) * (W / FT);
    if (!(flags & PHX_SUR_REUSE_WEIGHTS)) {
        hipLaunchKernelGGL(k_pack, dim3((4 * MID_W + 255) / 256), dim3(256), 0, st, P, S.wfwd, S.wbwd);
        hipLaunchKernelGGL(k_pack_frag, dim3((FRAG_N + 255) / 256), dim3(256), 0, st, P, S.frag);
    }
    float* dbuf[2] = {S.d, S.d2};
    double* tbuf[2] = {S.tpart, S.tpart2};
    const float* base = c;   // the unprojected state the next launch reads
    for (int j = 0; j < jumps; ++j) {
        const float* dp = j ? dbuf[(j - 1) & 1] : nullptr;
        const double* tp = j ? tbuf[(j - 1) & 1] : nullptr;
        // the projected (and corrected) state after j jumps: a slot of
        // `states` at every `every` jumps, else a workspace buffer
        float* up = nullptr;
        const float* ad = nullptr;
        if (j) {
            if (states && j % every == 0) {
                up = states + (long long)(j / every - 1) * n;
                if (corr) ad = corr + (long long)(j / every - 1) * n;
            } else {
                up = S.u[j & 1];
            }
        }
        hipLaunchKernelGGL(k_sur_fused_fwd, dim3(W / FT, H / FT, B), dim3(FNT), 0, st, base, dp, tp, ntile, up, ad, P,
                           S.wfwd, S.frag, dbuf[j & 1], tbuf[j & 1], H, W);
        if (j) base = up;
    }
    float* dst = states ? states + (long long)(jumps / every - 1) * n : out;
    const float* ad = states && corr ? corr + (long long)(jumps / every - 1) * n : nullptr;
    hipLaunchKernelGGL(k_sur_project_tiles, dim3(nb), dim3(256), 0, st, base, dbuf[(jumps - 1) & 1],
                       tbuf[(jumps - 1) & 1], ntile, H * W, n, dst, ad);
    PHX_LAUNCH_CHECK("phx_sur_rollout");
    return PHX_OK;
}

}  // namespace phx

using namespace phx;

extern "C" {

int phx_sur_param_count(void) { return P_COUNT; }

size_t phx_sur_workspace_bytes(int B, int H, int W) {
    if (check_shape(B, H, W, "phx_sur_workspace_bytes")) return 0;
    return ws_layout(B, H, W, nullptr, nullptr);
}

int phx_sur_forward(const float* params, void* ws, const float* c_in, float* c_out, int B, int H, int W, int flags,
                    phx_stream s) {
    int rc = check_shape(B, H, W, "phx_sur_forward");
    if (rc) return rc;
    if (!params || !ws || !c_in || !c_out) return fail(PHX_ERR_INVALID, "phx_sur_forward: NULL argument");
    SurWS S;
    ws_layout(B, H, W, &S, (char*)ws);
    return sur_forward(params, S, c_in, c_out, B, H, W, flags, as_stream(s));
}

int phx_sur_rollout(const float* params, void* ws, const float* c_in, float* c_out, int B, int H, int W, int jumps,
                    int flags, phx_stream s) {
    int rc = check_shape(B, H, W, "phx_sur_rollout");
    if (rc) return rc;
    if (!params || !ws || !c_in || !c_out) return fail(PHX_ERR_INVALID, "phx_sur_rollout: NULL argument");
    if (jumps < 1) return fail(PHX_ERR_INVALID, "phx_sur_rollout: jumps < 1");
    SurWS S;
    ws_layout(B, H, W, &S, (char*)ws);
    return sur_rollout(params, S, c_in, c_out, B, H, W, jumps, flags, as_stream(s));
}

int phx_sur_rollout_states(const float* params, void* ws, const float* c_in, float* states, const float* corr, int B,
                           int H, int W, int jumps, int every, int flags, phx_stream s) {
    int rc = check_shape(B, H, W, "phx_sur_rollout_states");
    if (rc) return rc;
    if (!params || !ws || !c_in || !states) return fail(PHX_ERR_INVALID, "phx_sur_rollout_states: NULL argument");
    if (jumps < 1 || every < 1 || jumps % every)
        return fail(PHX_ERR_INVALID, "phx_sur_rollout_states: need jumps >= 1 and every >= 1 dividing jumps");
    // the kept states are written while c_in and corr are still read
    const size_t field = (size_t)B * H * W * sizeof(float), kept = field * (size_t)(jumps / every);
    auto overlap = [](const void* a, size_t na, const void* b, size_t nb) {
        const char *pa = (const char*)a, *pb = (const char*)b;
        return pa < pb + nb && pb < pa + na;
    };
    if (overlap(states, kept, c_in, field) || (corr && overlap(states, kept, corr, kept)))
        return fail(PHX_ERR_INVALID, "phx_sur_rollout_states: states overlaps c_in or corr");
    SurWS S;
    ws_layout(B, H, W, &S, (char*)ws);
    return sur_rollout(params, S, c_in, nullptr, B, H, W, jumps, flags, as_stream(s), states, every, corr);
}

int phx_sur_backward(const float* params, void* ws, const float* c_in, const float* c_out, const float* c_target,
                     float* grad, double* loss_dev, int B, int H, int W, phx_stream s) {
    int rc = check_shape(B, H, W, "phx_sur_backward");
    if (rc) return rc;
    if (!params || !ws || !c_in || !c_out || !c_target || !grad)
        return fail(PHX_ERR_INVALID, "phx_sur_backward: NULL argument");
    SurWS S;
    ws_layout(B, H, W, &S, (char*)ws);
    hipStream_t st = as_stream(s);
    const long long n = (long long)B * H * W;
    const int nb = (int)((n + 255) / 256);
    hipLaunchKernelGGL(k_loss_grad, dim3(nb), dim3(256), 0, st, c_out, c_target, n, S.g, S.e2);
    if (loss_dev) {
        // loss = mean of e2 = (sum over samples of per-sample sums) / n
        seg_sums(S.e2, n, 1, S.spart, S.loss, st);
    }
    seg_sums(S.g, (long long)H * W, B, S.spart, S.sums, st);
    hipLaunchKernelGGL(k_unproject, dim3(nb), dim3(256), 0, st, S.g, S.sums, H * W, n, S.d);
    // last layer: weight grads, then dz5
    hipLaunchKernelGGL(k_edge_wgrad<1>, dim3(WG_BLOCKS), dim3(320), 0, st, (const float*)nullptr, S.act[4], S.d,
                       S.part, B, H, W);
    hipLaunchKernelGGL(k_reduce_edge, dim3(2), dim3(256), 0, st, S.part, WG_BLOCKS, 289, grad + P_W6);
    hipLaunchKernelGGL(k_conv_out_bwd, dim3(nb), dim3(256), 0, st, S.d, S.act[4], params, S.dz[0], B, H, W);
    int cur = 0;
    for (int l = 3; l >= 0; --l) {
        // layer l maps act[l] -> act[l+1]; dz[cur] is the gradient at its pre-activation
        hipLaunchKernelGGL(k_conv_mid_wgrad, dim3(WG_BLOCKS), dim3(256), 0, st, S.act[l], S.dz[cur], S.part, B, H, W);
        hipLaunchKernelGGL(k_reduce_mid, dim3((PART + 255) / 256), dim3(256), 0, st, S.part, WG_BLOCKS,
                           grad + P_MID + l * (MID_W + CH));
        hipLaunchKernelGGL(k_conv_mid, dim3(W / 16, H / 16, B), dim3(256), 0, st, S.dz[cur], S.wbwd + l * MID_W,
                           (const float*)nullptr, S.act[l], S.dz[cur ^ 1], H, W, 1);
        cur ^= 1;
    }
    // first layer: weight grads from x and dz1
    hipLaunchKernelGGL(k_edge_wgrad<0>, dim3(WG_BLOCKS), dim3(320), 0, st, c_in, S.dz[cur], (const float*)nullptr,
                       S.part, B, H, W);
    hipLaunchKernelGGL(k_reduce_edge, dim3(2), dim3(256), 0, st, S.part, WG_BLOCKS, 320, grad + P_W1);
    if (loss_dev) PHX_HIP(hipMemcpyAsync(loss_dev, S.loss, sizeof(double), hipMemcpyDeviceToDevice, st));
    PHX_LAUNCH_CHECK("phx_sur_backward");
    return PHX_OK;
}

int phx_sur_adam(float* params, const float* grad, float* m, float* v, int step, double lr, double b1, double b2,
                 double eps, phx_stream s) {
    if (!params || !grad || !m || !v || step < 1) return fail(PHX_ERR_INVALID, "phx_sur_adam: bad arguments");
    const float c1 = (float)(1.0 - std::pow(b1, step)), c2 = (float)(1.0 - std::pow(b2, step));
    hipLaunchKernelGGL(k_adam, dim3((P_COUNT + 255) / 256), dim3(256), 0, as_stream(s), params, grad, m, v, P_COUNT,
                       (float)lr, (float)b1, (float)b2, (float)eps, c1, c2);
    PHX_LAUNCH_CHECK("phx_sur_adam");
    return PHX_OK;
}

// Single 32->32 layer entry points (what the network runs), exposed for
// parity tests. x, y, dz: BF16 NHWC [B][H][W][32]; w: FP32 [32][32][3][3].
int phx_conv3x3_bf16_fwd(const void* x, const float* w, const float* bias, void* y, void* wpk_ws, int B, int H, int W,
                         phx_stream s) {
    int rc = check_shape(B, H, W, "phx_conv3x3_bf16_fwd");
    if (rc) return rc;
    if (!x || !w || !bias || !y || !wpk_ws) return fail(PHX_ERR_INVALID, "phx_conv3x3_bf16_fwd: NULL argument");
    hipStream_t st = as_stream(s);
    // pack a one-layer parameter block: reuse k_pack by laying w at layer 0
    float* Pl = (float*)wpk_ws;  // [P_COUNT] floats, then packed weights
    __bf16* wf = (__bf16*)(Pl + P_COUNT);
    __bf16* wb = wf + 4 * MID_W;
    PHX_HIP(hipMemcpyAsync(Pl + P_MID, w, MID_W * sizeof(float), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_pack, dim3((4 * MID_W + 255) / 256), dim3(256), 0, st, Pl, wf, wb);
    hipLaunchKernelGGL(k_conv_mid, dim3(W / 16, H / 16, B), dim3(256), 0, st, (const __bf16*)x, wf, bias,
                       (const __bf16*)nullptr, (__bf16*)y, H, W, 0);
    PHX_LAUNCH_CHECK("phx_conv3x3_bf16_fwd");
    return PHX_OK;
}

size_t phx_conv3x3_ws_bytes(void) { return (size_t)P_COUNT * 4 + 8 * MID_W * 2 + (size_t)WG_BLOCKS * PART * 4; }

int phx_conv3x3_bf16_dgrad(const void* dy, const float* w, const void* mask, void* dx, void* wpk_ws, int B, int H,
                           int W, phx_stream s) {
    int rc = check_shape(B, H, W, "phx_conv3x3_bf16_dgrad");
    if (rc) return rc;
    if (!dy || !w || !mask || !dx || !wpk_ws) return fail(PHX_ERR_INVALID, "phx_conv3x3_bf16_dgrad: NULL argument");
    hipStream_t st = as_stream(s);
    float* Pl = (float*)wpk_ws;
    __bf16* wf = (__bf16*)(Pl + P_COUNT);
    __bf16* wb = wf + 4 * MID_W;
    PHX_HIP(hipMemcpyAsync(Pl + P_MID, w, MID_W * sizeof(float), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_pack, dim3((4 * MID_W + 255) / 256), dim3(256), 0, st, Pl, wf, wb);
    hipLaunchKernelGGL(k_conv_mid, dim3(W / 16, H / 16, B), dim3(256), 0, st, (const __bf16*)dy, wb,
                       (const float*)nullptr, (const __bf16*)mask, (__bf16*)dx, H, W, 1);
    PHX_LAUNCH_CHECK("phx_conv3x3_bf16_dgrad");
    return PHX_OK;
}

int phx_conv3x3_bf16_wgrad(const void* x, const void* dz, float* dw, float* db, void* wpk_ws, int B, int H, int W,
                           phx_stream s) {
    int rc = check_shape(B, H, W, "phx_conv3x3_bf16_wgrad");
    if (rc) return rc;
    if (!x || !dz || !dw || !db || !wpk_ws) return fail(PHX_ERR_INVALID, "phx_conv3x3_bf16_wgrad: NULL argument");
    hipStream_t st = as_stream(s);
    float* part = (float*)((char*)wpk_ws + (size_t)P_COUNT * 4 + 8 * MID_W * 2);
    float* g = (float*)wpk_ws;   // reuse the parameter area for the reduced result
    hipLaunchKernelGGL(k_conv_mid_wgrad, dim3(WG_BLOCKS), dim3(256), 0, st, (const __bf16*)x, (const __bf16*)dz, part,
                       B, H, W);
    hipLaunchKernelGGL(k_reduce_mid, dim3((PART + 255) / 256), dim3(256), 0, st, part, WG_BLOCKS, g);
    PHX_HIP(hipMemcpyAsync(dw, g, MID_W * sizeof(float), hipMemcpyDeviceToDevice, st));
    PHX_HIP(hipMemcpyAsync(db, g + MID_W, CH * sizeof(float), hipMemcpyDeviceToDevice, st));
    PHX_LAUNCH_CHECK("phx_conv3x3_bf16_wgrad");
    return PHX_OK;
}

}  // extern "C"
