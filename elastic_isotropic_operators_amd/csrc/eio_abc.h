// eio_abc.h -- learned absorbing boundary: a small CNN that supplies the field outside the
// update region (the FAT frame the stencil reads) after every fine step (BF16 MFMA, FP32
// accumulation).
//
// Without it the stencil of the outermost update cells reads zeros in the 4-cell FAT frame:
// the edge of the grid reflects, and only the cosine sponge in front of it hides that.  With
// it, after the fused step kernel has written the new level, each of the four edges runs its
// own 3-layer network that predicts the 5 fields of the new level on its side of the frame
// (the values a wider grid would hold there), so the next step's stencil reads a continued
// wavefield instead of zeros; the model's sponge stays active.
//   input   16 channels per cell, on the update cells (zero elsewhere): the 5 fields of the
//           new level divided by a per-field scale, and either the 5 differences new -
//           previous (formed in FP32, before the BF16 rounding, which would otherwise swamp
//           them: one fine step changes a field by about a percent) or the 5 fields of the
//           previous level, divided by their own scale; 6 zero channels
//   conv1   5x5, 16 -> 16, leaky ReLU (slope ABC_LEAK), no bias
//   conv2   5x5, 16 -> 16, leaky ReLU, no bias
//   conv3   5x5, 16 -> 5 (padded to 16), no bias; output times the field scale = the frame
// All three convolutions are "valid" convolutions on the field arrays extended by zeros, so a
// frame cell sees the 13 x 13 neighbourhood around it, i.e. the first 6 update cells; the
// network is positively homogeneous (a wavefield scaled by c > 0 gets its frame scaled by c).
//
// Ownership: the x-min and x-max frames own the update rows, the z-min and z-max frames the
// update columns (the frame corners are never read by the stencil).  A tile is an output
// rectangle of the frame: 4 columns x at most 24 rows (x frames) or at most 16 columns x 4
// rows (z frames); the workgroup stages the tile's input (+6 cells) to LDS as BF16 and
// computes conv1 on (+4), conv2 on (+2) and conv3 on the tile, an MFMA block being 16 cells
// along the frame (rows of one column for the x frames, columns of one row for the z frames,
// whose LDS layout is transposed for that): the 25 taps are taken in 13 pairs (the last one
// padded with zero weights), and for each pair A = 16 cells x 32 (two taps x 16 input
// channels, read from LDS 16 bytes per lane), B = 32 x 16 output channels (the pair's weights,
// held in VGPRs), with gfx950's v_mfma_f32_16x16x32_bf16.  The network reads only update cells
// and writes only frame cells, so tiles need no ordering among themselves: one launch per step.
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
constexpr int ABC_PAIRS = (ABC_TAPS + 1) / 2;       // MFMA K steps of two taps
constexpr int ABC_FRAG = 64 * 8;                    // bf16 of one 32x16 B fragment
constexpr int ABC_WEDGE = ABC_LAYERS * ABC_PAIRS * ABC_FRAG;  // bf16 per edge
constexpr float ABC_LEAK = 0.1f;
// LDS: buffer 0 holds the input then conv2's output, buffer 1 conv1's output
constexpr int ABC_LDS0 = (ABC_MAXOX + 12) * (ABC_MAXR3 + 12) * ABC_CH;  // bf16
constexpr int ABC_LDS1 = (ABC_MAXOX + 8) * (ABC_MAXR3 + 8) * ABC_CH;

struct AbcTile {
    int shot;    // ordered shot index
    int edge;    // 0 x-min, 1 x-max, 2 z-min, 3 z-max
    int x0, z0;  // first output (frame) cell
    int ox, oz;  // output extent: ox <= ABC_MAXOX, oz <= ABC_MAXOZ
};

// Per-shot parameters of the network (device pointers), indexed by the ordered shot.
struct AbcShot {
    const unsigned short* w;  // [4 edges][3 layers][13 tap pairs][64 lanes][8] bf16, MFMA B-fragment order
    float scale[5];           // per-field scale of the new level (and of the output)
    float inv_scale[10];      // inverse scales of the 10 input channels
    int diff;                 // 1: channels 5..9 are new - previous, 0: the previous level
    int depth;                // frame rows written, counted from the update region (1..4; the rest stay 0)
};

typedef __bf16 abc_b8 __attribute__((ext_vector_type(8)));
typedef float abc_f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short abc_bf16(float f) {  // round to nearest even
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float abc_leaky(float v) { return v > 0.f ? v : ABC_LEAK * v; }

// B fragments of one layer (its 13 tap pairs) from the edge's weight buffer `rw`
__device__ __forceinline__ void abc_load_w(rsrc_t rw, int layer, int lane, abc_b8 (&w)[ABC_PAIRS]) {
#pragma unroll
    for (int q = 0; q < ABC_PAIRS; ++q)
        w[q] = __builtin_bit_cast(abc_b8, raw_buffer_load_f32x4(rw, ((layer * ABC_PAIRS + q) * 64 + lane) * 16, 0, 0));
}

// Where the last layer writes: the new level's five fields (buffer `rf`, field f at byte
// f * fbytes) and the output cells' geometry.
struct AbcOut {
    rsrc_t rf;
    unsigned fbytes;
    int nz, nx, x0, z0, ox, oz, depth;
    float scale;   // field scale of this lane's output channel (lane % 16 < 5)
};

// One convolution layer on LDS.  ZMAJ (x-min/x-max frames): cells stored [col][row], an MFMA
// block is 16 consecutive rows of one column; XMAJ (z-min/z-max frames, 4 rows tall): cells
// stored [row][col], a block is 16 consecutive columns of one row.  Input: nA x nB cells
// (A = the slow axis of the layout: columns for ZMAJ, rows for XMAJ), 16 channels; output
// (nA - 4) x (nB - 4).  Each wave runs two blocks' 13-step MFMA chains interleaved, with all
// 26 A fragments read from LDS before the chains start.  LAST: write channels < 5 of the
// first ox x oz cells, times the field scale, to the new level.
template <bool LAST, bool XMAJ, typename T>
__device__ __forceinline__ void abc_layer(const unsigned short* __restrict__ in, int nA, int nB,
                                          unsigned short* __restrict__ out, const abc_b8 (&b)[ABC_PAIRS],
                                          int lane, int wave, const AbcOut& O) {
    constexpr int NW = ABC_NTHR / 64;
    const int oA = nA - 4, oB = nB - 4;
    const int nbb = (oB + 15) >> 4;
    const int nblk = oA * nbb;
    const int i = lane & 15, g = lane >> 4;
    for (int blk = wave; blk < nblk; blk += 2 * NW) {
        const int blk2 = blk + NW;
        const bool two = blk2 < nblk;
        const int a0 = blk / nbb, b0 = min((blk - a0 * nbb) * 16, oB - 16);
        const int a1 = two ? blk2 / nbb : a0;
        const int b1 = two ? min((blk2 - a1 * nbb) * 16, oB - 16) : b0;
        abc_b8 q0[ABC_PAIRS], q1[ABC_PAIRS];
#pragma unroll
        for (int p = 0; p < ABC_PAIRS; ++p) {
            // lanes of group g < 2 read tap 2p, the others tap 2p+1 (the pad of the last pair
            // reads tap 24 again: finite values times zero weights), input channels 8 (g & 1) ..
            const int tap = (g < 2 || 2 * p + 1 >= ABC_TAPS) ? 2 * p : 2 * p + 1;
            const int dx = tap / 5, dz = tap - 5 * (tap / 5);   // cell (+dx along x, +dz along z)
            const int dA = XMAJ ? dz : dx, dB = XMAJ ? dx : dz;
            q0[p] = *reinterpret_cast<const abc_b8*>(in + ((a0 + dA) * nB + b0 + i + dB) * ABC_CH + 8 * (g & 1));
            q1[p] = *reinterpret_cast<const abc_b8*>(in + ((a1 + dA) * nB + b1 + i + dB) * ABC_CH + 8 * (g & 1));
        }
        abc_f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < ABC_PAIRS; ++p) {
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q0[p], b[p], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(q1[p], b[p], acc1, 0, 0, 0);
        }
        // acc[j] = output cell (a, b + 4g + j), channel i
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (h == 1 && !two) break;
            const int a = h ? a1 : a0, bb = h ? b1 : b0;
            const abc_f4 acc = h ? acc1 : acc0;
            if constexpr (!LAST) {
#pragma unroll
                for (int j = 0; j < 4; ++j) out[(a * oB + bb + 4 * g + j) * ABC_CH + i] = abc_bf16(abc_leaky(acc[j]));
            } else {
                if (i < 5) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int col = XMAJ ? bb + 4 * g + j : a, row = XMAJ ? a : bb + 4 * g + j;
                        // distance of the frame cell from the update region, 0 = adjacent
                        const int x = O.x0 + col, z = O.z0 + row;
                        const int dist = XMAJ ? (z < FAT ? FAT - 1 - z : z - (O.nz - FAT))
                                              : (x < FAT ? FAT - 1 - x : x - (O.nx - FAT));
                        if (col < O.ox && row < O.oz && dist < O.depth)
                            bst(O.rf, T(acc[j] * O.scale), unsigned(x * O.nz + z) * unsigned(sizeof(T)),
                                unsigned(i) * O.fbytes);
                    }
                }
            }
        }
    }
}

// Input cells staged per thread and pass: every thread loads the fields of ABC_SPT cells
// before it converts and writes any of them.
constexpr int ABC_SPT = 3;

template <bool XMAJ, typename T>
__device__ __forceinline__ void abc_tile(const ShotDev<T>& S, const AbcShot& P, const AbcTile& tl, int k,
                                         unsigned short* lds0, unsigned short* lds1, rsrc_t rw) {
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    constexpr unsigned SZ = sizeof(T);
    const unsigned fbytes = n * SZ;
    T* nw = (k & 1) ? S.B : S.A;
    const T* pv = (k & 1) ? S.A : S.B;
    const rsrc_t rnw = make_rsrc(nw, 5 * fbytes);
    const rsrc_t rpv = make_rsrc(pv, 5 * fbytes);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the first layer's B fragments, then the second's, issued ahead of the staging loads
    abc_b8 wa[ABC_PAIRS], wb[ABC_PAIRS];
    abc_load_w(rw, 0, lane, wa);
    abc_load_w(rw, 1, lane, wb);
    // layer-3 extent along the block axis is at least one block (16 cells)
    const int c3 = XMAJ ? max(tl.ox, 16) : tl.ox;
    const int r3 = XMAJ ? tl.oz : max(tl.oz, 16);
    const int cin = c3 + 12, rin = r3 + 12;
    const int ncell = cin * rin;
    const int nA = XMAJ ? rin : cin, nB = XMAJ ? cin : rin;
    const int xs = tl.x0 - 6, zs = tl.z0 - 6;
    // stage the input: update cells only (the frame and the outside read as zero); threads
    // run along z for coalesced loads; cells outside load cell 0 and are masked to zero
    for (int base = 0; base < ncell; base += ABC_SPT * ABC_NTHR) {
        T va[ABC_SPT][5], vb[ABC_SPT][5];
        bool in[ABC_SPT];
#pragma unroll
        for (int u = 0; u < ABC_SPT; ++u) {
            const int p = base + u * ABC_NTHR + int(threadIdx.x);
            const int c = p / rin, r = p - c * rin;
            const int x = xs + c, z = zs + r;
            in[u] = p < ncell && x >= FAT && x < nx - FAT && z >= FAT && z < nz - FAT;
            const unsigned voff = in[u] ? unsigned(x * nz + z) * SZ : 0u;
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                va[u][f] = bld<T>(rnw, voff, unsigned(f) * fbytes);
                vb[u][f] = bld<T>(rpv, voff, unsigned(f) * fbytes);
            }
        }
#pragma unroll
        for (int u = 0; u < ABC_SPT; ++u) {
            const int p = base + u * ABC_NTHR + int(threadIdx.x);
            if (p >= ncell) continue;
            const int c = p / rin, r = p - c * rin;
            unsigned short v[16];
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const float a = in[u] ? float(va[u][f]) : 0.f;
                const float bb = in[u] ? float(vb[u][f]) : 0.f;
                v[f] = abc_bf16(a * P.inv_scale[f]);
                v[5 + f] = abc_bf16((P.diff ? a - bb : bb) * P.inv_scale[5 + f]);
            }
#pragma unroll
            for (int f = 10; f < 16; ++f) v[f] = 0;
            const int q = XMAJ ? r * cin + c : c * rin + r;
            uint4* d = reinterpret_cast<uint4*>(lds0 + size_t(q) * ABC_CH);
            d[0] = make_uint4(v[0] | (unsigned(v[1]) << 16), v[2] | (unsigned(v[3]) << 16),
                              v[4] | (unsigned(v[5]) << 16), v[6] | (unsigned(v[7]) << 16));
            d[1] = make_uint4(v[8] | (unsigned(v[9]) << 16), v[10] | (unsigned(v[11]) << 16),
                              v[12] | (unsigned(v[13]) << 16), v[14] | (unsigned(v[15]) << 16));
        }
    }
    AbcOut O;
    O.rf = rnw;
    O.fbytes = fbytes;
    O.nz = nz;
    O.nx = nx;
    O.x0 = tl.x0;
    O.z0 = tl.z0;
    O.ox = tl.ox;
    O.oz = tl.oz;
    O.depth = P.depth;
    O.scale = P.scale[min(lane & 15, 4)];
    __syncthreads();
    abc_layer<false, XMAJ, T>(lds0, nA, nB, lds1, wa, lane, wave, O);
    abc_load_w(rw, 2, lane, wa);   // the third layer's fragments, while the second runs
    __syncthreads();
    abc_layer<false, XMAJ, T>(lds1, nA - 4, nB - 4, lds0, wb, lane, wave, O);
    __syncthreads();
    abc_layer<true, XMAJ, T>(lds0, nA - 8, nB - 8, nullptr, wa, lane, wave, O);
}

// Frame of the new level of fine step k (the new level is (k & 1) ? B : A, the previous one
// the other).  The network's parameters are read in place from global memory (`par`), never
// copied to a private struct: a lane-indexed array of such a copy would be placed in LDS.
template <typename T>
__global__ __launch_bounds__(ABC_NTHR) void k_abc_cnn(const ShotDev<T>* __restrict__ shots,
                                                      const AbcShot* __restrict__ par,
                                                      const AbcTile* __restrict__ tiles, int ntile, int k) {
    __shared__ __attribute__((aligned(16))) unsigned short lds0[ABC_LDS0];
    __shared__ __attribute__((aligned(16))) unsigned short lds1[ABC_LDS1];
    const int t = blockIdx.x;
    if (t >= ntile) return;
    const AbcTile tl = tiles[t];
    const AbcShot& P = par[tl.shot];
    // B fragments: for tap pair p, lane holds W[tap 2p + (lane/32)][cin = 8*((lane/16)&1) + j]
    // [cout = lane%16], [3 layers][13 pairs][64 lanes][8] bf16 per edge
    const rsrc_t rw = make_rsrc(P.w + size_t(tl.edge) * ABC_WEDGE, unsigned(ABC_WEDGE * sizeof(unsigned short)));
    if (tl.edge >= 2) abc_tile<true, T>(shots[tl.shot], P, tl, k, lds0, lds1, rw);
    else abc_tile<false, T>(shots[tl.shot], P, tl, k, lds0, lds1, rw);
}

}  // namespace eio
