// eio_abc.h -- learned absorbing boundary: a small CNN that corrects the edge strip of the
// new time level after every fine step (BF16 MFMA, FP32 accumulation).
//
// The model keeps a narrow plain cosine sponge of `strip` cells.  After the fused step
// kernel has written the new level, each of the four edges runs its own 3-layer network on
// the cells of its strip:
//   input   16 channels per cell: the 5 fields of the new level and the 5 of the previous
//           level, each divided by a per-field scale (6 zero channels)
//   conv1   5x5, 16 -> 16, leaky ReLU (slope ABC_LEAK), no bias
//   conv2   5x5, 16 -> 16, leaky ReLU, no bias
//   conv3   5x5, 16 -> 5 (padded to 16), no bias; output times the field scale = correction
// All three convolutions are "valid" convolutions on the field arrays extended by zeros
// outside the array, so the receptive field of a strip cell is its 13 x 13 neighbourhood
// and the network is positively homogeneous (a wavefield scaled by c > 0 gets its
// correction scaled by c), which keeps it independent of the source amplitude.
//
// Ownership: the x-min and x-max strips own every update row; the z-min and z-max strips own
// the columns between them.  A tile is an output rectangle of at most 16 columns and 24 rows;
// the workgroup stages the tile's input (+6 cells) to LDS as BF16 and computes conv1 on
// (+4), conv2 on (+2) and conv3 on the tile, one 16-row block of one column per MFMA chain:
// for each of the 25 taps, A = 16 cells x 16 input channels (read from LDS), B = 16 input x
// 16 output channels (the tap's weights, held in VGPRs), with v_mfma_f32_16x16x16_bf16.
// Corrections go to a separate buffer (neighbouring tiles still read the uncorrected
// level) and k_abc_add adds them afterwards.
#pragma once
#include "eio_kernels.h"

namespace eio {

constexpr int ABC_CH = 16;
constexpr int ABC_TAPS = 25;
constexpr int ABC_LAYERS = 3;
constexpr int ABC_NTHR = 256;                       // 4 waves
constexpr int ABC_MAXOX = 16;                       // tile columns
constexpr int ABC_MAXOZ = 24;                       // tile rows
constexpr int ABC_MAXR3 = ABC_MAXOZ;                // rows computed by conv3 (>= 16)
constexpr int ABC_FRAG = 64 * 4;                    // bf16 of one 16x16 B fragment
constexpr int ABC_WEDGE = ABC_LAYERS * ABC_TAPS * ABC_FRAG;  // bf16 per edge
constexpr float ABC_LEAK = 0.1f;
// LDS: buffer 0 holds the input then conv2's output, buffer 1 conv1's output
constexpr int ABC_LDS0 = (ABC_MAXOX + 12) * (ABC_MAXR3 + 12) * ABC_CH;  // bf16
constexpr int ABC_LDS1 = (ABC_MAXOX + 8) * (ABC_MAXR3 + 8) * ABC_CH;

struct AbcTile {
    int shot;    // ordered shot index
    int edge;    // 0 x-min, 1 x-max, 2 z-min, 3 z-max
    int x0, z0;  // first output cell
    int ox, oz;  // output extent: ox <= ABC_MAXOX, oz <= ABC_MAXOZ
};

// Per-shot parameters of the network (device pointers), indexed by the ordered shot.
struct AbcShot {
    const unsigned short* w;  // [4 edges][3 layers][25 taps][64 lanes][4] bf16, MFMA B-fragment order
    float scale[5];           // per-field scale
    float inv_scale[5];
    float* corr;              // [5][nx*nz] corrections of the last step
};

typedef short abc_s4 __attribute__((ext_vector_type(4)));
typedef float abc_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short abc_bf16(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float abc_leaky(float v) { return v > 0.f ? v : ABC_LEAK * v; }

// One convolution layer on LDS: input (cI x rI cells, 16 channels), output (cI-4) x rO where
// rO = rI - 4.  LAST: write channels < 5 of the first ox x oz cells to the correction buffer.
template <bool LAST>
__device__ __forceinline__ void abc_layer(const unsigned short* __restrict__ in, int cI, int rI,
                                          unsigned short* __restrict__ out, const unsigned short* __restrict__ wl,
                                          int lane, int wave, float* corr, const AbcShot& P, unsigned n, int nz,
                                          int x0, int z0, int ox, int oz) {
    const int cO = cI - 4, rO = rI - 4;
    // B fragments of the 25 taps: lane holds W[tap][cin = 4*(lane/16) + j][cout = lane%16]
    abc_s4 b[ABC_TAPS];
#pragma unroll
    for (int t = 0; t < ABC_TAPS; ++t) b[t] = *reinterpret_cast<const abc_s4*>(wl + (t * 64 + lane) * 4);
    const int nbr = (rO + 15) >> 4;
    const int nblk = cO * nbr;
    const int i = lane & 15, g = lane >> 4;
    for (int blk = wave; blk < nblk; blk += ABC_NTHR / 64) {
        const int col = blk / nbr;
        const int rb = blk - col * nbr;
        const int row0 = min(rb * 16, rO - 16);
        abc_f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
#pragma unroll
            for (int dz = 0; dz < 5; ++dz) {
                const abc_s4 a = *reinterpret_cast<const abc_s4*>(
                    in + ((col + dx) * rI + row0 + i + dz) * ABC_CH + 4 * g);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[dx * 5 + dz], acc, 0, 0, 0);
            }
        }
        // acc[j] = output cell (col, row0 + 4g + j), channel i
        if constexpr (!LAST) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                out[(col * rO + row0 + 4 * g + j) * ABC_CH + i] = abc_bf16(abc_leaky(acc[j]));
        } else {
            if (i < 5 && col < ox) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int r = row0 + 4 * g + j;
                    if (r < oz) corr[size_t(i) * n + size_t(x0 + col) * nz + (z0 + r)] = acc[j] * P.scale[i];
                }
            }
        }
    }
}

// Corrections of fine step k (the new level is (k & 1) ? B : A, the previous one the other).
template <typename T>
__global__ __launch_bounds__(ABC_NTHR) void k_abc_cnn(const ShotDev<T>* __restrict__ shots,
                                                      const AbcShot* __restrict__ par,
                                                      const AbcTile* __restrict__ tiles, int ntile, int k) {
    __shared__ __attribute__((aligned(16))) unsigned short lds0[ABC_LDS0];
    __shared__ __attribute__((aligned(16))) unsigned short lds1[ABC_LDS1];
    const int t = blockIdx.x;
    if (t >= ntile) return;
    const AbcTile tl = tiles[t];
    const ShotDev<T>& S = shots[tl.shot];
    const AbcShot P = par[tl.shot];
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    const T* nw = (k & 1) ? S.B : S.A;
    const T* pv = (k & 1) ? S.A : S.B;
    const int r3 = max(tl.oz, 16);
    const int cin = tl.ox + 12, rin = r3 + 12;
    const int xs = tl.x0 - 6, zs = tl.z0 - 6;
    // stage the input: zero outside the array
    for (int p = threadIdx.x; p < cin * rin; p += ABC_NTHR) {
        const int c = p / rin, r = p - c * rin;
        const int x = xs + c, z = zs + r;
        unsigned short v[16];
        const bool inside = x >= 0 && x < nx && z >= 0 && z < nz;
        const size_t cell = inside ? size_t(x) * nz + z : 0;
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const float a = inside ? float(nw[size_t(f) * n + cell]) : 0.f;
            const float b = inside ? float(pv[size_t(f) * n + cell]) : 0.f;
            v[f] = abc_bf16(a * P.inv_scale[f]);
            v[5 + f] = abc_bf16(b * P.inv_scale[f]);
        }
#pragma unroll
        for (int f = 10; f < 16; ++f) v[f] = 0;
        uint4* d = reinterpret_cast<uint4*>(lds0 + size_t(p) * ABC_CH);
        d[0] = make_uint4(v[0] | (unsigned(v[1]) << 16), v[2] | (unsigned(v[3]) << 16), v[4] | (unsigned(v[5]) << 16),
                          v[6] | (unsigned(v[7]) << 16));
        d[1] = make_uint4(v[8] | (unsigned(v[9]) << 16), v[10] | (unsigned(v[11]) << 16),
                          v[12] | (unsigned(v[13]) << 16), v[14] | (unsigned(v[15]) << 16));
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned short* we = P.w + size_t(tl.edge) * ABC_WEDGE;
    abc_layer<false>(lds0, cin, rin, lds1, we, lane, wave, nullptr, P, n, nz, 0, 0, 0, 0);
    __syncthreads();
    abc_layer<false>(lds1, cin - 4, rin - 4, lds0, we + ABC_TAPS * ABC_FRAG, lane, wave, nullptr, P, n, nz, 0, 0,
                     0, 0);
    __syncthreads();
    abc_layer<true>(lds0, cin - 8, rin - 8, nullptr, we + 2 * ABC_TAPS * ABC_FRAG, lane, wave, P.corr, P, n, nz,
                    tl.x0, tl.z0, tl.ox, tl.oz);
}

// new level += corrections on every tile's output cells
template <typename T>
__global__ __launch_bounds__(ABC_NTHR) void k_abc_add(const ShotDev<T>* __restrict__ shots,
                                                      const AbcShot* __restrict__ par,
                                                      const AbcTile* __restrict__ tiles, int ntile, int k) {
    const int t = blockIdx.x;
    if (t >= ntile) return;
    const AbcTile tl = tiles[t];
    const ShotDev<T>& S = shots[tl.shot];
    const float* corr = par[tl.shot].corr;
    const int nz = S.nz;
    const size_t n = size_t(nz) * S.nx;
    T* nw = (k & 1) ? S.B : S.A;
    const int cells = tl.ox * tl.oz;
    for (int p = threadIdx.x; p < 5 * cells; p += ABC_NTHR) {
        const int f = p / cells, q = p - f * cells;
        const int c = q / tl.oz, r = q - c * tl.oz;
        const size_t idx = size_t(f) * n + size_t(tl.x0 + c) * nz + (tl.z0 + r);
        nw[idx] += T(corr[idx]);
    }
}

}  // namespace eio
