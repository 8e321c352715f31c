// eio_abc_train.h -- training kernels of the learned-boundary network (eio_abc.h): the
// forward of one 5x5 "valid" convolution layer with its leaky ReLU, and the two halves of its
// backward (input gradient, weight gradient), on BF16 MFMA with FP32 accumulation.
//
// Tensors are NHWC with 16 channels (layer inputs of fewer channels are zero-padded), the W
// axis the long one (a batch of frame crops: x frames are transposed by the caller), W >= 16.
// Weights per layer: [cout 16][cin 16][5][5] FP32 (rows of unused channels zero), taps
// t = 5 ty + tx; the kernels round weights and activations to BF16 at the points the
// inference kernel does.
//   forward      z[b,y,x,o] = sum_{i,ty,tx} a[b,y+ty,x+tx,i] W[o,i,ty,tx]       (Ho = H-4, Wo = W-4)
//                h = bf16(leaky(z)) (or bf16(z) for the last layer), z kept in FP32
//   input grad   da[b,y,x,i] = sum_{o,ty,tx} dz[b,y-ty,x-tx,o] W[o,i,ty,tx]    (zero outside)
//   weight grad  dW[o,i,ty,tx] = sum_{b,y,x} dz[b,y,x,o] a[b,y+ty,x+tx,i]
// Each is an implicit GEMM of v_mfma_f32_16x16x32_bf16 (M = 16 cells along W or 16 output
// channels, N = 16 channels, K = two taps x 16 channels or 32 cells).  The weight gradient
// of a launch is written as per-workgroup partial sums and reduced by a second kernel in a
// fixed order, so it is deterministic.
#pragma once
#include "eio_abc.h"

namespace eio {

constexpr int AT_NTHR = 256;   // 4 waves
constexpr int AT_CELLS = 16;   // cells along W per MFMA block

__device__ __forceinline__ float at_bf16_round(float f) {
    return __uint_as_float(unsigned(abc_bf16(f)) << 16);
}
__device__ __forceinline__ __bf16 at_to_bf16(float f) {
    return __builtin_bit_cast(__bf16, abc_bf16(f));
}

// B fragment of tap pair p for the forward (K = 16 tsel + i, N = o) or, TRANSPOSE, for the
// input gradient (K = 16 tsel + o, N = i), from the FP32 weights; lane = 16 g + n.
template <bool TRANSPOSE>
__device__ __forceinline__ abc_b8 at_wfrag(const float* __restrict__ W, int p, int lane) {
    const int n = lane & 15, g = lane >> 4;
    const int tsel = g >> 1, k0 = 8 * (g & 1);
    const int tap = 2 * p + tsel;
    abc_b8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        float w = 0.f;
        if (tap < ABC_TAPS) {
            const int o = TRANSPOSE ? k0 + j : n, i = TRANSPOSE ? n : k0 + j;
            w = W[(o * ABC_CH + i) * ABC_TAPS + tap];
        }
        v[j] = at_to_bf16(w);
    }
    return v;
}

// forward of one layer: one 16-cell block of one output row per wave and MFMA chain
__global__ __launch_bounds__(AT_NTHR) void k_abct_conv_fwd(const unsigned short* __restrict__ a, int B, int H, int Wd,
                                                            const float* __restrict__ W, int leaky,
                                                            float* __restrict__ z, unsigned short* __restrict__ h) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Ho = H - 4, Wo = Wd - 4;
    const int nbw = (Wo + AT_CELLS - 1) / AT_CELLS;
    const long nblk = long(B) * Ho * nbw;
    abc_b8 wf[ABC_PAIRS];
#pragma unroll
    for (int p = 0; p < ABC_PAIRS; ++p) wf[p] = at_wfrag<false>(W, p, lane);
    const int i = lane & 15, g = lane >> 4;
    for (long blk = long(blockIdx.x) * 4 + wave; blk < nblk; blk += long(gridDim.x) * 4) {
        const int bw = int(blk % nbw);
        const long r = blk / nbw;
        const int y = int(r % Ho), b = int(r / Ho);
        const int x0 = min(bw * AT_CELLS, Wo - AT_CELLS);
        abc_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < ABC_PAIRS; ++p) {
            const int tap = (g < 2 || 2 * p + 1 >= ABC_TAPS) ? 2 * p : 2 * p + 1;
            const int ty = tap / 5, tx = tap - 5 * ty;
            const unsigned short* src = a + ((size_t(b) * H + y + ty) * Wd + x0 + i + tx) * ABC_CH + 8 * (g & 1);
            const abc_b8 q = *reinterpret_cast<const abc_b8*>(src);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q, wf[p], acc, 0, 0, 0);
        }
        // acc[j] = z at cell x0 + 4 g + j, channel i
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t o = ((size_t(b) * Ho + y) * Wo + x0 + 4 * g + j) * ABC_CH + i;
            z[o] = acc[j];
            h[o] = abc_bf16(leaky ? abc_leaky(acc[j]) : acc[j]);
        }
    }
}

// input gradient: da at one 16-cell block of one input row; dz read in FP32 and rounded to
// BF16, zero outside the output array
__global__ __launch_bounds__(AT_NTHR) void k_abct_conv_bwd_data(const float* __restrict__ dz, int B, int H, int Wd,
                                                                 const float* __restrict__ W,
                                                                 float* __restrict__ da) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Ho = H - 4, Wo = Wd - 4;
    const int nbw = (Wd + AT_CELLS - 1) / AT_CELLS;
    const long nblk = long(B) * H * nbw;
    abc_b8 wf[ABC_PAIRS];
#pragma unroll
    for (int p = 0; p < ABC_PAIRS; ++p) wf[p] = at_wfrag<true>(W, p, lane);
    const int i = lane & 15, g = lane >> 4;
    for (long blk = long(blockIdx.x) * 4 + wave; blk < nblk; blk += long(gridDim.x) * 4) {
        const int bw = int(blk % nbw);
        const long r = blk / nbw;
        const int y = int(r % H), b = int(r / H);
        const int x0 = min(bw * AT_CELLS, Wd - AT_CELLS);
        abc_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < ABC_PAIRS; ++p) {
            const int tsel = g >> 1;
            const int tap = 2 * p + tsel;
            abc_b8 q;
            const int ty = tap / 5, tx = tap - 5 * ty;
            const int yo = y - ty, xo = x0 + i - tx;
            const bool in = tap < ABC_TAPS && yo >= 0 && yo < Ho && xo >= 0 && xo < Wo;
            const float* src = dz + ((size_t(b) * Ho + (in ? yo : 0)) * Wo + (in ? xo : 0)) * ABC_CH + 8 * (g & 1);
#pragma unroll
            for (int j = 0; j < 8; ++j) q[j] = at_to_bf16(in ? src[j] : 0.f);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q, wf[p], acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) da[((size_t(b) * H + y) * Wd + x0 + 4 * g + j) * ABC_CH + i] = acc[j];
    }
}

// weight gradient, partial sums: workgroup k sums the cells blk = k, k + G, ... (blocks of 32
// consecutive output cells of one row) into part[k][tap][o][i]; D (16 o x 16 i) += A (16 o x
// 32 cells) B (32 cells x 16 i) per tap, each wave holding the accumulators of 7 taps
__global__ __launch_bounds__(AT_NTHR) void k_abct_conv_bwd_weight(const unsigned short* __restrict__ a,
                                                                   const float* __restrict__ dz, int B, int H, int Wd,
                                                                   float* __restrict__ part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int Ho = H - 4, Wo = Wd - 4;
    const int nbw = (Wo + 31) / 32;
    const long nblk = long(B) * Ho * nbw;
    const int m = lane & 15, g = lane >> 4;   // A row = output channel o; K = cells 8 g .. 8 g + 7
    constexpr int TPW = (ABC_TAPS + 3) / 4;   // taps per wave
    abc_f4 acc[TPW];
#pragma unroll
    for (int u = 0; u < TPW; ++u) acc[u] = abc_f4{0.f, 0.f, 0.f, 0.f};
    for (long blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int bw = int(blk % nbw);
        const long r = blk / nbw;
        const int y = int(r % Ho), b = int(r / Ho);
        const int x0 = bw * 32;
        // A: dz at cells x0 + 8 g + j (zero past the row end), channel m
        abc_b8 qa;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int x = x0 + 8 * g + j;
            qa[j] = at_to_bf16(x < Wo ? dz[((size_t(b) * Ho + y) * Wo + x) * ABC_CH + m] : 0.f);
        }
#pragma unroll
        for (int u = 0; u < TPW; ++u) {
            const int tap = wave * TPW + u;
            if (tap < ABC_TAPS) {
                const int ty = tap / 5, tx = tap - 5 * ty;
                // B: a at cells x0 + 8 g + j + tx of row y + ty, channel n = m (the B column)
                abc_b8 qb;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int x = x0 + 8 * g + j;
                    const unsigned short v =
                        x < Wo ? a[((size_t(b) * H + y + ty) * Wd + x + tx) * ABC_CH + m] : (unsigned short)0;
                    qb[j] = __builtin_bit_cast(__bf16, v);
                }
                acc[u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(qa, qb, acc[u], 0, 0, 0);
            }
        }
    }
    // acc[u][j] = dW[o = 4 g + j][i = m] of tap wave*TPW + u
#pragma unroll
    for (int u = 0; u < TPW; ++u) {
        const int tap = wave * TPW + u;
        if (tap < ABC_TAPS) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                part[((size_t(blockIdx.x) * ABC_TAPS + tap) * ABC_CH + 4 * g + j) * ABC_CH + m] = acc[u][j];
        }
    }
}

// dW[o][i][tap] = sum over the partial sums of the workgroups, in workgroup order
__global__ __launch_bounds__(AT_NTHR) void k_abct_reduce_weight(const float* __restrict__ part, int nparts,
                                                                 float* __restrict__ dW) {
    const int e = blockIdx.x * AT_NTHR + threadIdx.x;   // (tap, o, i)
    if (e >= ABC_TAPS * ABC_CH * ABC_CH) return;
    float s = 0.f;
    for (int k = 0; k < nparts; ++k) s += part[size_t(k) * ABC_TAPS * ABC_CH * ABC_CH + e];
    const int tap = e / (ABC_CH * ABC_CH), o = (e / ABC_CH) % ABC_CH, i = e % ABC_CH;
    dW[(o * ABC_CH + i) * ABC_TAPS + tap] = s;
}

}  // namespace eio
