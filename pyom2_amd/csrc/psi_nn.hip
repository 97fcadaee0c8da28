// psi_nn.hip -- kernels of the learned preconditioner: a 4-level
// encoder-decoder convolutional net (channels 16/32/64/128, 3x3 convolutions
// as implicit GEMM on BF16 MFMA with FP32 accumulation), its backward pass,
// Adam, the A-norm training loss and the flexible-CG helpers that use it.
//
// Activations are NHWC bfloat16 ([batch][H][W][C], C a multiple of 16).
// Convolutions: MFMA v_mfma_f32_16x16x32_bf16 (gfx950, wave64): lane l
// supplies A[l%16][8(l/16)..+7] and B[8(l/16)..+7][l%16] and receives
// D[4(l/16)+i][l%16], i = 0..3.  Weight gradients: v_mfma_f32_16x16x16_bf16
// (A[l%16][4(l/16)..+3], B[4(l/16)..+3][l%16], the same D).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/psinn.h"

namespace psinn {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ float4_t mfma16(short4_t a, short4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float4_t mfma32(uint4 a, uint4 b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                   __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Lane-layout probe: D = A (16x32) x B (32x16), bf16 in, fp32 out, row-major,
// through the convolutions' MFMA.
__global__ void k_mfma_probe(const uint16_t* A, const uint16_t* B, float* D) {
    const int l = threadIdx.x;
    uint16_t a[8], b[8];
    for (int j = 0; j < 8; ++j) {
        a[j] = A[(l & 15) * 32 + 8 * (l >> 4) + j];
        b[j] = B[(8 * (l >> 4) + j) * 16 + (l & 15)];
    }
    float4_t d = {0, 0, 0, 0};
    d = mfma32(__builtin_bit_cast(uint4, a), __builtin_bit_cast(uint4, b), d);
    for (int i = 0; i < 4; ++i) D[(4 * (l >> 4) + i) * 16 + (l & 15)] = d[i];
}

// ---------------------------------------------------------------------------
// 3x3 convolution, stride 1, zero padding, implicit GEMM on MFMA.
// Input channels are the concatenation [in1 (C1) | in2 (C2)]; in1 may be
// read through a nearest x2 upsampling (up1: in1 has H/2 x W/2).
//
// A workgroup (4 waves) computes a CV_TH x CV_TW tile of output pixels for
// all 16 NB output channels.  The input channels are taken 16 at a time: for
// each chunk the (CV_TH+2) x (CV_TW+2) x 16 input tile (with its halo) and
// the chunk's packed weights are staged in LDS once, shared by the four
// waves and reused for all 9 taps.  A K step of the MFMA covers two taps of
// the chunk (lane group g of a 16-lane quarter takes tap 2 ks + g/2,
// channels 8 (g%2) .. +7; the tenth tap is zero), so a chunk is 5 K steps.
// Wave w computes output row w of the tile: MR segments of 16 pixels x NB
// channel blocks in accumulators.  The next chunk's loads are issued into
// registers before the current chunk's MFMAs.
// Weights are pre-packed per (chunk, K step, nb) as 64 lanes x 8 bf16 (B
// fragment order, k_pack_weights).  Epilogue: + bias, optional ReLU; output
// bf16 or fp32 with cout_store channels (<= 16 NB).
// ---------------------------------------------------------------------------
#ifndef PSI_CV_TH
#define PSI_CV_TH 4
#endif
constexpr int CV_TH = PSI_CV_TH, CV_TW = 64;
constexpr int CV_NT = 64 * CV_TH;               // threads: one wave per output row
constexpr int CV_FX = CV_TW + 2, CV_FY = CV_TH + 2;
constexpr int CV_PS = 24;                       // LDS bf16 per frame pixel (16 + 8 pad)
constexpr int CV_IN16 = CV_FX * CV_FY * 2;      // 16-byte units of an input chunk

template <int NB>
struct ConvShape {
    static constexpr int MR = NB <= 4 ? 4 : (NB <= 8 ? 2 : 1);   // 16-pixel segments per wave
    static constexpr int W16 = 5 * NB * 64;                       // 16-byte units of a weight chunk
    static constexpr int PER_IN = (CV_IN16 + CV_NT - 1) / CV_NT;
    static constexpr int PER_W = (W16 + CV_NT - 1) / CV_NT;
};

// Workgroups of a small image may split the output blocks: blockIdx.z =
// image * groups + group, the group computing blocks [group NB, group NB +
// NB) of nb_total.
template <int NB>
__global__ __launch_bounds__(CV_NT) void k_conv3x3(const uint16_t* __restrict__ in1, int C1, int up1,
                                                 const uint16_t* __restrict__ in2, int C2, int H,
                                                 int W, const uint16_t* __restrict__ wpack,
                                                 int nb_total, const float* __restrict__ bias,
                                                 int relu, uint16_t* __restrict__ out16,
                                                 float* __restrict__ out32, int cout_store) {
    using Sh = ConvShape<NB>;
    constexpr int MR = Sh::MR, PER_IN = Sh::PER_IN, PER_W = Sh::PER_W, W16 = Sh::W16;
    // input frame [fy][fx][24 bf16] and the chunk's weights [ks][nb][lane][8]
    __shared__ uint4 sIn[CV_FX * CV_FY * CV_PS / 8];
    __shared__ uint4 sW[W16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x0 = blockIdx.x * CV_TW, y0 = blockIdx.y * CV_TH;
    const int groups = nb_total / NB;
    const int bimg = blockIdx.z / groups, nb0 = (blockIdx.z % groups) * NB;
    const int nchunk = (C1 + C2) >> 4;
    const int H1 = up1 ? (H >> 1) : H, W1 = up1 ? (W >> 1) : W;
    const uint16_t* img1 = in1 + (size_t)bimg * H1 * W1 * C1;
    const uint16_t* img2 = in2 ? in2 + (size_t)bimg * H * W * C2 : nullptr;

    // global -> registers: input unit u = (frame pixel, half of the 16 channels)
    auto load_in = [&](int ch, uint4 (&v)[PER_IN]) {
        const int c0 = ch * 16;
#pragma unroll
        for (int k = 0; k < PER_IN; ++k) {
            const int u = tid + k * CV_NT;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (u < CV_IN16) {
                const int px = u >> 1, half = u & 1;
                const int fy = px / CV_FX, fx = px - fy * CV_FX;
                const int yy = y0 + fy - 1, xx = x0 + fx - 1;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                    const int c = c0 + 8 * half;
                    const uint16_t* src;
                    if (c < C1) {
                        const int ys = up1 ? (yy >> 1) : yy, xs = up1 ? (xx >> 1) : xx;
                        src = img1 + ((size_t)ys * W1 + xs) * C1 + c;
                    } else {
                        src = img2 + ((size_t)yy * W + xx) * C2 + (c - C1);
                    }
                    val = *reinterpret_cast<const uint4*>(src);
                }
            }
            v[k] = val;
        }
    };
    auto load_w = [&](int ch, uint4 (&v)[PER_W]) {
        // this group's blocks of each K step: NB x 64 consecutive units
        const uint4* src = reinterpret_cast<const uint4*>(wpack) + (size_t)ch * 5 * nb_total * 64;
#pragma unroll
        for (int k = 0; k < PER_W; ++k) {
            const int u = tid + k * CV_NT;
            const int ks = u / (NB * 64), rem = u - ks * (NB * 64);
            v[k] = u < W16 ? src[(size_t)(ks * nb_total + nb0) * 64 + rem]
                           : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto store = [&](const uint4 (&vi)[PER_IN], const uint4 (&vw)[PER_W]) {
#pragma unroll
        for (int k = 0; k < PER_IN; ++k) {
            const int u = tid + k * CV_NT;
            if (u < CV_IN16) sIn[(u >> 1) * (CV_PS / 8) + (u & 1)] = vi[k];
        }
#pragma unroll
        for (int k = 0; k < PER_W; ++k) {
            const int u = tid + k * CV_NT;
            if (u < W16) sW[u] = vw[k];
        }
    };

    float4_t acc[MR][NB];
#pragma unroll
    for (int m = 0; m < MR; ++m)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[m][nb] = (float4_t){0, 0, 0, 0};

    const int g = lane >> 4, pl = lane & 15;
    // this wave's output row; its pixels [x0 + 16 m + pl] for m < MR, of the
    // CV_TW / 16 segments of the tile (waves with MR < 4 cover the tile in
    // several passes below)
    uint4 ri[PER_IN], rw[PER_W];
    load_in(0, ri);
    load_w(0, rw);
    constexpr int SEG = CV_TW / 16, PASSES = SEG / MR;
    static_assert(SEG % MR == 0, "segments per wave");
    // With MR < SEG the tile's segments are processed in PASSES rounds, each
    // over all chunks (accumulators are per round).
#pragma unroll 1
    for (int pass = 0; pass < PASSES; ++pass) {
        if (pass > 0) {
#pragma unroll
            for (int m = 0; m < MR; ++m)
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[m][nb] = (float4_t){0, 0, 0, 0};
            load_in(0, ri);
            load_w(0, rw);
        }
#pragma unroll 1
        for (int ch = 0; ch < nchunk; ++ch) {
            __syncthreads();   // the previous chunk's MFMAs have read the LDS
            store(ri, rw);
            __syncthreads();
            if (ch + 1 < nchunk) {
                load_in(ch + 1, ri);
                load_w(ch + 1, rw);
            }
#pragma unroll
            for (int ks = 0; ks < 5; ++ks) {
                const int tap = 2 * ks + (g >> 1);
                // tap 9 is padding: its A is zero, addressed at tap 8's offsets
                const int tt = tap < 9 ? tap : 8;
                const int dy = tt / 3 - 1, dx = tt - 3 * (tt / 3) - 1;
                uint4 a[MR];
#pragma unroll
                for (int m = 0; m < MR; ++m) {
                    const int fx = (pass * MR + m) * 16 + pl + dx + 1, fy = wave + dy + 1;
                    a[m] = tap < 9 ? sIn[(fy * CV_FX + fx) * (CV_PS / 8) + (g & 1)]
                                   : make_uint4(0u, 0u, 0u, 0u);
                }
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    const uint4 b = sW[(ks * NB + nb) * 64 + lane];
#pragma unroll
                    for (int m = 0; m < MR; ++m) acc[m][nb] = mfma32(a[m], b, acc[m][nb]);
                }
            }
        }
        // the output through LDS, one 16-channel block at a time (the input
        // frame's LDS is free: all MFMAs have read it): rows [pixel][16 ch]
        // in fp32, then written as 16-byte bf16 units or as consecutive
        // fp32 channels of consecutive pixels
        // (staged in bf16 for a bf16 output, in fp32 for an fp32 one)
        float* sOut = reinterpret_cast<float*>(sIn);
        uint16_t* sOut16 = reinterpret_cast<uint16_t*>(sIn);
        constexpr int TPX = CV_TH * MR * 16;   // pixels of this pass
        static_assert(TPX * 16 * 4 <= (int)sizeof(sIn), "output staging fits the frame's LDS");
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int cb = (nb0 + nb) * 16;
            const float bv = bias ? bias[cb + pl] : 0.f;
            __syncthreads();
#pragma unroll
            for (int m = 0; m < MR; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    float v = acc[m][nb][i] + bv;
                    if (relu) v = fmaxf(v, 0.f);
                    const int q = ((wave * MR + m) * 16 + 4 * g + i) * 16 + pl;
                    if (out16) sOut16[q] = f2bf(v);
                    else sOut[q] = v;
                }
            __syncthreads();
            const int nval = min(16, cout_store - cb);
            if (nval <= 0) continue;
            if (out16) {
                for (int u = tid; u < TPX * 2; u += CV_NT) {
                    const int tp = u >> 1, half = u & 1;
                    const int wy = tp / (MR * 16), px = x0 + pass * MR * 16 + (tp - wy * MR * 16);
                    const int yy = y0 + wy;
                    if (yy < H && px < W)
                        *reinterpret_cast<uint4*>(out16 + (((size_t)bimg * H + yy) * W + px) *
                                                              cout_store + cb + 8 * half) =
                            reinterpret_cast<const uint4*>(sOut16)[tp * 2 + half];
                }
            } else {
                for (int u = tid; u < TPX * nval; u += CV_NT) {
                    const int tp = u / nval, c = u - tp * nval;
                    const int wy = tp / (MR * 16), px = x0 + pass * MR * 16 + (tp - wy * MR * 16);
                    const int yy = y0 + wy;
                    if (yy < H && px < W)
                        out32[(((size_t)bimg * H + yy) * W + px) * cout_store + cb + c] =
                            sOut[tp * 16 + c];
                }
            }
        }
    }
}

// Weight gradient of k_conv3x3: dW[tap][ci][co] += sum_p in[p + tap][ci] * g[p][co]
// (MFMA with M = ci, N = co, K = pixels), dbias[co] += sum_p g[p][co].
// One wave per (tap, kb, nb) tile and pixel chunk; fp32 atomics.
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(
    const uint16_t* __restrict__ in1, int C1, int up1, const uint16_t* __restrict__ in2, int C2,
    int B, int H, int W, const uint16_t* __restrict__ g, int COUT, int chunk,
    float* __restrict__ dW, float* __restrict__ dbias) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int KB = (C1 + C2) >> 4, NB = COUT >> 4;
    const int tile = blockIdx.x * 4 + wave;
    if (tile >= 9 * KB * NB) return;
    const int tap = tile / (KB * NB), kb = (tile / NB) % KB, nb = tile % NB;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int H1 = up1 ? (H >> 1) : H, W1 = up1 ? (W >> 1) : W;
    const long npix = (long)B * H * W;
    const long p0 = (long)blockIdx.y * chunk;
    const long p1 = p0 + chunk < npix ? p0 + chunk : npix;
    const int ci = kb * 16 + (lane & 15);
    const int co = nb * 16 + (lane & 15);
    float4_t acc = {0, 0, 0, 0};
    float bsum = 0.f;
    for (long pb = p0; pb < p1; pb += 16) {
        short4_t a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const long p = pb + 4 * (lane >> 4) + j;
            uint16_t av = 0, gv = 0;
            if (p < p1) {
                const int bimg = (int)(p / ((long)H * W));
                const int rem = (int)(p % ((long)H * W));
                const int y = rem / W, x = rem % W;
                gv = g[p * COUT + co];
                const int yy = y + dy, xx = x + dx;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                    if (ci < C1) {
                        const int ys = up1 ? (yy >> 1) : yy, xs = up1 ? (xx >> 1) : xx;
                        av = in1[(((size_t)bimg * H1 + ys) * W1 + xs) * C1 + ci];
                    } else {
                        av = in2[(((size_t)bimg * H + yy) * W + xx) * C2 + (ci - C1)];
                    }
                }
            }
            a[j] = (short)av;
            b[j] = (short)gv;
            if (tap == 0 && kb == 0) bsum += bf2f(gv);
        }
        acc = mfma16(a, b, acc);
    }
    // D[m = ci][n = co]: lane holds ci rows 4(l/16)+i, co column l%16
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int cim = kb * 16 + 4 * (lane >> 4) + i;
        atomicAdd(&dW[((size_t)tap * (C1 + C2) + cim) * COUT + co], acc[i]);
    }
    if (tap == 0 && kb == 0 && dbias) {
        // lanes l, l+16, l+32, l+48 hold partial sums of the same co
        bsum += __shfl_xor(bsum, 16, 64);
        bsum += __shfl_xor(bsum, 32, 64);
        if (lane < 16) atomicAdd(&dbias[co], bsum);
    }
}

// Pack fp32 master weights w[tap][ci][co] (CIN x COUT, tap = 3 ky + kx) into
// the B fragments of k_conv3x3: per input chunk of 16 channels, 5 K steps of
// two taps (the tenth zero), per 16-channel output block nb, 64 lanes x 8
// bf16: lane l, element j holds K = (tap 2 ks + g/2, channel 16 chunk +
// 8 (g%2) + j), N = 16 nb + l%16, with g = l/16.
// transposed = 0: forward convolution (K = ci, N = co).
// transposed = 1: data-gradient convolution (K = co, N = ci, taps flipped):
// dIn[p][ci] = sum_tap sum_co dOut[p - d(tap)][co] w[tap][ci][co]
//            = sum_tap' sum_co dOut[p + d(tap')][co] w[8 - tap'][ci][co].
// K16/N16: number of 16-blocks along K and N (zero padded); the pack holds
// K16 * 5 * N16 * 512 bf16.
__global__ void k_pack_weights(const float* __restrict__ w, int CIN, int COUT, int transposed,
                               int K16, int N16, uint16_t* __restrict__ pack) {
    const long total = (long)K16 * 5 * N16 * 512;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int j = (int)(idx & 7), lane = (int)((idx >> 3) & 63);
    long t = idx >> 9;
    const int nb = (int)(t % N16);
    t /= N16;
    const int ks = (int)(t % 5);
    const int chunk = (int)(t / 5);
    const int g = lane >> 4;
    const int tap = 2 * ks + (g >> 1);
    const int k = chunk * 16 + 8 * (g & 1) + j, n = nb * 16 + (lane & 15);
    float v = 0.f;
    if (tap < 9) {
        if (!transposed) {
            if (k < CIN && n < COUT) v = w[((size_t)tap * CIN + k) * COUT + n];
        } else {
            if (k < COUT && n < CIN) v = w[((size_t)(8 - tap) * CIN + n) * COUT + k];
        }
    }
    pack[idx] = f2bf(v);
}

// 2x2 average pool, bf16 NHWC, 8 channels (16 bytes) per thread.
__global__ void k_avgpool2(const uint16_t* __restrict__ in, int B, int H, int W, int C,
                           uint16_t* __restrict__ out) {
    const int C8 = C >> 3;
    const long n = (long)B * (H / 2) * (W / 2) * C8;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int c = (int)(idx % C8) * 8;
    long t = idx / C8;
    const int x = (int)(t % (W / 2));
    t /= (W / 2);
    const int y = (int)(t % (H / 2));
    const int b = (int)(t / (H / 2));
    const uint16_t* p = in + (((size_t)b * H + 2 * y) * W + 2 * x) * C + c;
    uint16_t v[4][8], o[8];
    *reinterpret_cast<uint4*>(v[0]) = *reinterpret_cast<const uint4*>(p);
    *reinterpret_cast<uint4*>(v[1]) = *reinterpret_cast<const uint4*>(p + C);
    *reinterpret_cast<uint4*>(v[2]) = *reinterpret_cast<const uint4*>(p + (size_t)W * C);
    *reinterpret_cast<uint4*>(v[3]) = *reinterpret_cast<const uint4*>(p + (size_t)W * C + C);
#pragma unroll
    for (int k = 0; k < 8; ++k)
        o[k] = f2bf(0.25f * (((bf2f(v[0][k]) + bf2f(v[1][k])) + bf2f(v[2][k])) + bf2f(v[3][k])));
    *reinterpret_cast<uint4*>(out + (((size_t)b * (H / 2) + y) * (W / 2) + x) * C + c) =
        *reinterpret_cast<const uint4*>(o);
}

// Sum of the 4 fine children (gradient of a nearest x2 upsampling) for the
// channel slice [c0, c0+C) of a fine fp32 tensor with CF channels.
__global__ void k_sumpool2(const float* __restrict__ fine, int B, int H, int W, int CF, int c0,
                           int C, float* __restrict__ coarse) {
    const long n = (long)B * (H / 2) * (W / 2) * C;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int c = (int)(idx % C);
    long t = idx / C;
    const int x = (int)(t % (W / 2));
    t /= (W / 2);
    const int y = (int)(t % (H / 2));
    const int b = (int)(t / (H / 2));
    const float* p = fine + (((size_t)b * H + 2 * y) * W + 2 * x) * CF + c0 + c;
    coarse[idx] = p[0] + p[CF] + p[(size_t)W * CF] + p[(size_t)W * CF + CF];
}

// out_bf16[p][c] = (act[p][c] > 0) * (A[p][cA0 + c] + 0.25 * Bc[p/2][c])
// A: fp32 with CA channels (may be null); Bc: coarse fp32 with C channels
// (gradient of a 2x2 average pool, may be null); act: bf16 with C channels.
__global__ void k_grad_combine(const uint16_t* __restrict__ act, int B, int H, int W, int C,
                               const float* __restrict__ A, int CA, int cA0,
                               const float* __restrict__ Bc, uint16_t* __restrict__ out) {
    const long n = (long)B * H * W * C;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int c = (int)(idx % C);
    const long p = idx / C;
    float v = 0.f;
    if (A) v += A[p * CA + cA0 + c];
    if (Bc) {
        const int x = (int)(p % W);
        const long t = p / W;
        const int y = (int)(t % H);
        const int b = (int)(t / H);
        v += 0.25f * Bc[(((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
    }
    if (!(bf2f(act[idx]) > 0.f)) v = 0.f;
    out[idx] = f2bf(v);
}

// fp32 single channel [B][H][W] -> bf16 16-channel (channel 0 = value).
__global__ void k_expand16(const float* __restrict__ g, long n, uint16_t* __restrict__ out) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * 16) return;
    out[idx] = (idx & 15) ? (uint16_t)0 : f2bf(g[idx >> 4]);
}

// Adam with bias correction on fp32 parameters.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float gscale) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// Grid <-> network layout.  The solver grid is (ny+4) x (nx+4) with pitch
// nx+4; the network grid is H x W (multiples of 8, >= the solver grid) with
// the solver grid in its top-left corner and zeros elsewhere.
// ---------------------------------------------------------------------------

// Network input (bf16, 16 channels, 3 used): Z r * s_in (s_in from *scale,
// the reciprocal rms of Z r over active points), the normalised inverse
// depth, the active mask.  r and Z in solver layout.
template <class T>
__global__ void k_net_input(const T* __restrict__ r, const float* __restrict__ Z,
                            const float* __restrict__ hinv, int sh, int sw, int H, int W,
                            const double* __restrict__ scale, uint16_t* __restrict__ out) {
    const long n = (long)H * W;
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int y = (int)(p / W), x = (int)(p % W);
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;
    if (y < sh && x < sw) {
        const size_t o = (size_t)y * sw + x;
        const float z = Z[o];
        // scale is produced by an atomically updated device scalar: read it
        // with a device-scope load (not through this XCD's L2)
        c0 = (float)((double)z * (double)r[o] *
                     __hip_atomic_load(scale, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        c1 = hinv[o];
        c2 = z != 0.f ? 1.f : 0.f;
    }
    // the pixel's 16 channels (3 set, 13 zero) as two 16-byte stores
    uint4* q = reinterpret_cast<uint4*>(out + p * 16);
    q[0] = make_uint4((uint32_t)f2bf(c0) | ((uint32_t)f2bf(c1) << 16), (uint32_t)f2bf(c2), 0u, 0u);
    q[1] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace psinn

using namespace psinn;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static inline unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }
static int launch_status() { return hipGetLastError() == hipSuccess ? 0 : 6; }

extern "C" {

int psinn_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, void* stream) {
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, A, B, D);
    return launch_status();
}

int psinn_conv3x3(const uint16_t* in1, int C1, int up1, const uint16_t* in2, int C2, int B, int H,
                  int W, const uint16_t* wpack, int cout_pad, const float* bias, int relu,
                  uint16_t* out16, float* out32, int cout_store, void* stream) {
    if ((C1 & 15) || (C2 & 15) || (cout_pad & 15) || C1 <= 0 || B <= 0 || H <= 0 || W <= 0 ||
        (up1 && ((H & 1) || (W & 1))) || (!out16 && !out32) || cout_store > cout_pad ||
        (out16 && (cout_store & 15)))
        return 5;
    const int NB = cout_pad / 16;
    const long tiles = (long)((W + CV_TW - 1) / CV_TW) * ((H + CV_TH - 1) / CV_TH) * B;
    // a small image with many output blocks: split them over workgroups
    int sub = NB;
    if (NB > 4 && tiles < 1024) sub = (NB % 4 == 0) ? 4 : (NB % 3 == 0 ? 3 : NB);
    const dim3 grid((W + CV_TW - 1) / CV_TW, (H + CV_TH - 1) / CV_TH, B * (NB / sub)), blk(CV_NT);
    hipStream_t s = (hipStream_t)stream;
#define CONV_CASE(K)                                                                          \
    case K:                                                                                   \
        hipLaunchKernelGGL(k_conv3x3<K>, grid, blk, 0, s, in1, C1, up1, in2, C2, H, W, wpack, \
                           NB, bias, relu, out16, out32, cout_store);                         \
        break;
    switch (sub) {
        CONV_CASE(1)
        CONV_CASE(2)
        CONV_CASE(3)
        CONV_CASE(4)
        CONV_CASE(6)
        CONV_CASE(8)
        CONV_CASE(12)
        default: return 5;
    }
#undef CONV_CASE
    return launch_status();
}

int psinn_conv3x3_wgrad(const uint16_t* in1, int C1, int up1, const uint16_t* in2, int C2, int B,
                        int H, int W, const uint16_t* g, int cout_pad, float* dW, float* dbias,
                        void* stream) {
    if ((C1 & 15) || (C2 & 15) || (cout_pad & 15)) return 5;
    const int KB = (C1 + C2) / 16, NB = cout_pad / 16;
    const int tiles = 9 * KB * NB;
    const long npix = (long)B * H * W;
    const int chunk = 1024;
    const dim3 grid((tiles + 3) / 4, (unsigned)((npix + chunk - 1) / chunk)), blk(256);
    hipLaunchKernelGGL(k_conv3x3_wgrad, grid, blk, 0, (hipStream_t)stream, in1, C1, up1, in2, C2, B,
                       H, W, g, cout_pad, chunk, dW, dbias);
    return launch_status();
}

int psinn_pack_weights(const float* w, int cin, int cout, int transposed, int k16, int n16,
                       uint16_t* pack, void* stream) {
    const long total = (long)k16 * 5 * n16 * 512;
    hipLaunchKernelGGL(k_pack_weights, dim3(nb256(total)), dim3(256), 0, (hipStream_t)stream, w,
                       cin, cout, transposed, k16, n16, pack);
    return launch_status();
}

int psinn_avgpool2(const uint16_t* in, int B, int H, int W, int C, uint16_t* out, void* stream) {
    if (C & 7) return 5;
    const long n = (long)B * (H / 2) * (W / 2) * (C / 8);
    hipLaunchKernelGGL(k_avgpool2, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, in, B, H, W,
                       C, out);
    return launch_status();
}

int psinn_sumpool2(const float* fine, int B, int H, int W, int CF, int c0, int C, float* coarse,
                   void* stream) {
    const long n = (long)B * (H / 2) * (W / 2) * C;
    hipLaunchKernelGGL(k_sumpool2, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, fine, B, H, W,
                       CF, c0, C, coarse);
    return launch_status();
}

int psinn_grad_combine(const uint16_t* act, int B, int H, int W, int C, const float* A, int CA,
                       int cA0, const float* Bc, uint16_t* out, void* stream) {
    const long n = (long)B * H * W * C;
    hipLaunchKernelGGL(k_grad_combine, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, act, B, H,
                       W, C, A, CA, cA0, Bc, out);
    return launch_status();
}

int psinn_expand16(const float* g, long n, uint16_t* out, void* stream) {
    hipLaunchKernelGGL(k_expand16, dim3(nb256(n * 16)), dim3(256), 0, (hipStream_t)stream, g, n, out);
    return launch_status();
}

int psinn_adam(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2,
               float eps, int step, float gscale, void* stream) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr,
                       b1, b2, eps, c1, c2, gscale);
    return launch_status();
}

int psinn_net_input(const void* r, int r_is_f64, const float* Z, const float* hinv, int sh, int sw,
                    int H, int W, const double* scale, uint16_t* out, void* stream) {
    const long n = (long)H * W;
    if (r_is_f64)
        hipLaunchKernelGGL(k_net_input<double>, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream,
                           (const double*)r, Z, hinv, sh, sw, H, W, scale, out);
    else
        hipLaunchKernelGGL(k_net_input<float>, dim3(nb256(n)), dim3(256), 0, (hipStream_t)stream,
                           (const float*)r, Z, hinv, sh, sw, H, W, scale, out);
    return launch_status();
}

}  // extern "C"
