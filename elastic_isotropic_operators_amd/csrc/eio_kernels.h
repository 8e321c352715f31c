// eio_kernels.h -- gfx950 kernels of the elastic velocity-stress propagator.
//
// One launch per fine time step advances every active shot of a batch:
//   fused forward step   = stencil (kernelsGpuElastic.cu:234-377) + source injection
//                          (:121-189) + cosine sponge (:1192-1221) + receiver recording (:7-60)
//                          [+ Born coarse decimation (:1337-1346) | Born secondary-source
//                          injection (:1363-1368)]
//   fused adjoint step   = adjoint stencil (:380-529) + data injection (:191-229) + sponge
//                          + source-trace extraction (:100-116) [+ Born adjoint-field
//                          extraction (:1369-1376)]
//   imaging kernels      = Born secondary source (:1381-1459) and imaging condition (:1538-1613)
//
// Work decomposition: a tile is 64 (z, one wave64 lane per z) x WX*R (x) update-region
// cells; a workgroup of WX waves owns one tile and each lane computes R consecutive x cells.
// The current time level is staged through LDS with a 4-cell halo so every staggered first
// derivative is read from LDS; x derivatives reuse a register window of R+8 values.  Old
// level and media are streamed once from HBM and the new level is written once (80 B per
// FP32 cell-update for the forward and adjoint steps).
//
// Memory access: each shot's state levels, media and Born slices are contiguous 5-field
// buffers addressed through SGPR buffer descriptors.  A lane keeps one 32-bit byte offset;
// the field and column offsets are scalar, so the per-load address cost is nil, and halo
// reads past the end of a buffer return 0 in hardware (reads that run past the end of a
// column land in a neighbouring column: those values only feed lanes whose results are
// discarded).
//
// Sponge fusion.  The reference damps BOTH time levels after every step
// (p0 <- d*(p0 + A p1 + s), p1 <- d*p1, then swaps).  The damped p1 is only ever used as
// the next step's older level, so the fused kernel keeps p1 undamped in memory and applies
// the second factor when it reads the older level: new = d*(d*old + A cur + s).  This is
// the same arithmetic as the reference sequence with one pass over memory.
#pragma once
#include <hip/hip_runtime.h>

#ifndef EIO_PREFETCH
#define EIO_PREFETCH 1
#endif

namespace eio {

constexpr int FAT = 4;
constexpr int TZ = 64;        // z extent of a tile = wave64 lanes
constexpr int ZH = TZ + 2 * FAT;
constexpr int NTHR = 256;     // threads of the element-wise helper kernels

// Modes of the fused forward kernel: bit flags
constexpr int F_INJ_COARSE = 1;  // source injection from coarse traces, time-interpolated on the fly
constexpr int F_INJ_FINE = 2;    // source injection from pre-fine-sampled traces
constexpr int F_REC = 4;         // receiver recording
constexpr int F_DEC = 8;         // coarse decimation of Vx, Vz into the storage ring
constexpr int F_SS = 16;         // Born secondary-source injection
constexpr int M_NL = F_INJ_COARSE | F_REC;      // NL forward
constexpr int M_BSRC = F_INJ_FINE | F_DEC;      // Born source pass
constexpr int M_BSCAT = F_SS | F_REC;           // Born scattered pass
constexpr int M_RECOMP = F_INJ_COARSE | F_DEC;  // forward recompute of a checkpointed segment
// Modes of the fused adjoint kernel
constexpr int M_NLADJ = 0;    // NL adjoint: data injection, source-trace extraction
constexpr int M_BADJ = 1;     // Born receiver pass: data injection, adjoint-field extraction

template <typename T>
struct ShotDev {
    int nz, nx, nts, sub, h, ntw, minPad;
    int surface;           // 1: free surface at row SURF (no sponge towards the top)
    T cz[4], cx[4];
    T* A;                  // state level A, 5 fields of nz*nx (older level at even steps)
    T* B;                  // state level B
    T* C;                  // levels C, D: second pair of the two-step NL forward (else NULL)
    T* D;
    const T* med;          // rhoxDtw, rhozDtw, lamb2MuDtw, lambDtw, muxzDtw (5 fields)
    const T* damp;         // [minPad] cosine sponge
    const T* filt;         // [2h] time interpolation filter, [2h] = 1/(2 dts)
    T* src_tr[5];          // source traces (NL: [n][nts], Born: [n][ntw])
    T* rec_tr[5];          // receiver traces [n][nts]
    T* ss[2];              // Born secondary sources / adjoint fields, 5 fields each, ping-pong
    T* Wvx;                // Born coarse storage: wcap slices of nz*nx, coarse index its in slot its % wcap
    T* Wvz;
    int wcap;              // nts for full storage, segment length + 3 for the checkpointed gradient
    int pad1;
    T* dm;                 // Born perturbation (forward) or per-shot imaging accumulator (adjoint)
    T* wfld;               // saved coarse-sampled wavefield [nts][5][nz*nx] of this shot, or NULL
};

// Accumulate the new values of one cell into the saved wavefield, time slices its and its+1
// with weights wa, wb (kernelsGpuElastic.cu:1299-1335).
template <typename T>
__device__ __forceinline__ void save_wfld(T* wfld, unsigned n, int its, size_t cell, const T v[5], T wa, T wb) {
    T* a = wfld + size_t(its) * 5 * n + cell;
    T* b = a + size_t(5) * n;
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        a[size_t(f) * n] += v[f] * wa;
        b[size_t(f) * n] += v[f] * wb;
    }
}

struct TileDev {
    int shot, z0, x0;
    int src_b, src_e, rec_b, rec_e;
    int pad;
};

// Sparse device entry: lcell = lz | lx << 8 (cell inside the tile), crow = row | comp << 24
struct Entry {
    int lcell;
    int crow;
};

// ------------------------------------------------------------------ buffer access
typedef __amdgpu_buffer_rsrc_t rsrc_t;

// `bytes` < 2^31: eio_model_add rejects grids whose 5-field buffers reach 2^31 bytes, so
// every descriptor size and lane offset of the kernels fits a non-negative int.
__device__ __forceinline__ rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, int(bytes), 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bld(rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float bld<float>(rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(voff), int(soff), 0));
}
template <>
__device__ __forceinline__ double bld<double>(rsrc_t r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, int(voff), int(soff), 0));
}
__device__ __forceinline__ void bst(rsrc_t r, float v, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, int(voff), int(soff), 0);
}
// nontemporal store (cache policy nt): streamed results that no later access of the launch reads
__device__ __forceinline__ void bst_nt(rsrc_t r, float v, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, int(voff), int(soff), 2);
}
__device__ __forceinline__ void bst(rsrc_t r, double v, unsigned voff, unsigned soff) {
    using u64 = decltype(__builtin_amdgcn_raw_buffer_load_b64(rsrc_t(), 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u64, v), r, int(voff), int(soff), 0);
}

// 16 bytes from a buffer (a 4-byte-aligned offset is enough): the LLVM intrinsic itself,
// since this toolchain's __builtin_amdgcn_raw_buffer_load_b128 is lowered to a 32-bit load.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ f32x4_t raw_buffer_load_f32x4(rsrc_t r, int voff, int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v4f32");

// Bijective block -> tile map that gives each of the 8 XCDs a contiguous range of tiles
// (workgroups are dispatched round-robin over XCDs), so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, j = b >> 3;
    return x * q + (x < r ? x : r) + j;
}

// kernelsGpuElastic.cu:1192-1221; with a free surface 1223-1252 (no top side)
template <typename T>
__device__ __forceinline__ T damp_at(const ShotDev<T>& S, int iz, int ix) {
    const int top = S.surface ? S.nz : iz - FAT;
    const int d = min(min(top, ix - FAT), min(S.nz - iz - 1 - FAT, S.nx - ix - 1 - FAT));
    return d < S.minPad ? S.damp[d] : T(1);
}

// ------------------------------------------------------------------ free surface
// Surface row and the top-block map M of the current level (kernelsGpuElastic.cu:536-587) on
// one column of an LDS stage whose row r is global row r: sigmazz = 0 on the surface and odd
// above it, sigmaxz odd about the half row above the surface, vx = 0 above the surface, vz = 0
// on and above it.  Its transpose clears the rows M overwrites and folds their negated values
// into the mirrored rows.  Only the top tile (z0 = FAT) of a free-surface shot touches these rows.
constexpr int SURF = 2 * FAT;

template <typename T>
__device__ __forceinline__ void fs_map_col(T* vx, T* vz, T* szz, T* sxz) {
    szz[SURF] = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) szz[SURF - 1 - k] = -szz[SURF + 1 + k];
#pragma unroll
    for (int k = 0; k < 4; ++k) sxz[SURF - k] = -sxz[SURF + 1 + k];
#pragma unroll
    for (int r = SURF - 4; r < SURF; ++r) vx[r] = T(0);
#pragma unroll
    for (int r = SURF - 4; r <= SURF; ++r) vz[r] = T(0);
}
template <typename T>
__device__ __forceinline__ void fs_map_col_transpose(T* vx, T* vz, T* szz, T* sxz) {
#pragma unroll
    for (int k = 0; k < 4; ++k) szz[SURF + 1 + k] -= szz[SURF - 1 - k];
#pragma unroll
    for (int r = SURF - 4; r <= SURF; ++r) szz[r] = T(0);
#pragma unroll
    for (int k = 0; k < 4; ++k) sxz[SURF + 1 + k] -= sxz[SURF - k];
#pragma unroll
    for (int r = SURF - 3; r <= SURF; ++r) sxz[r] = T(0);
#pragma unroll
    for (int r = SURF - 4; r < SURF; ++r) vx[r] = T(0);
#pragma unroll
    for (int r = SURF - 4; r <= SURF; ++r) vz[r] = T(0);
}

// D-: sum_k C_k (q[m+k] - q[m-1-k]) ; D+: sum_k C_k (q[m+1+k] - q[m-k])
template <typename T>
__device__ __forceinline__ T dm_q(const T* q, int m, const T* c) {
    return c[0] * (q[m] - q[m - 1]) + c[1] * (q[m + 1] - q[m - 2]) + c[2] * (q[m + 2] - q[m - 3]) +
           c[3] * (q[m + 3] - q[m - 4]);
}
template <typename T>
__device__ __forceinline__ T dp_q(const T* q, int m, const T* c) {
    return c[0] * (q[m + 1] - q[m]) + c[1] * (q[m + 2] - q[m - 1]) + c[2] * (q[m + 3] - q[m - 2]) +
           c[3] * (q[m + 4] - q[m - 3]);
}

// Time bookkeeping of fine step k (0-based launch index) for a direction.
struct StepTime {
    int its, it2;
};
__device__ __forceinline__ StepTime fwd_time(int k, int sub) {
    StepTime t;
    t.its = k / sub;
    t.it2 = k - t.its * sub + 1;  // 1..sub
    return t;
}
__device__ __forceinline__ StepTime adj_time(int k, int sub, int nts) {
    StepTime t;
    const int q = k / sub;
    t.its = nts - 2 - q;
    t.it2 = sub - 1 - (k - q * sub);  // sub-1..0
    return t;
}


// ------------------------------------------------------------------ forward step
// OCC = requested waves per SIMD (bounds the VGPR budget: 512/OCC).
template <typename T, int R, int WX, int OCC, int MODE>
__global__ __launch_bounds__(64 * WX) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_step_fwd(const ShotDev<T>* __restrict__ shots, const TileDev* __restrict__ tiles,
                const Entry* __restrict__ ents, int ntiles, int k, int sspar, int dlo, int dhi) {
    constexpr int NT = 64 * WX;
    constexpr int TX = WX * R;
    constexpr int TXH = TX + 2 * FAT;
    constexpr unsigned SZ = sizeof(T);
    __shared__ T lds[5][TXH][ZH];

    const TileDev t = tiles[xcd_remap(blockIdx.x, ntiles)];
    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    const unsigned fbytes = n * SZ, colb = unsigned(nz) * SZ;
    const StepTime st = fwd_time(k, S.sub);
    T* oldp = (k & 1) ? S.B : S.A;
    const T* curp = (k & 1) ? S.A : S.B;
    const rsrc_t rold = make_rsrc(oldp, 5 * fbytes);
    const rsrc_t rcur = make_rsrc(curp, 5 * fbytes);
    const rsrc_t rmed = make_rsrc(S.med, 5 * fbytes);

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 + lane, zz = lane + FAT;
    const int c0 = FAT + w * R, ix0 = t.x0 + w * R;
    const unsigned voff = (unsigned(ix0) * unsigned(nz) + unsigned(iz)) * SZ;

    // older level (and media, EIO_PREFETCH=2) of the own cells: issued before the halo stage
    T o[5][R], m[5][R];
#if EIO_PREFETCH >= 1
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) o[f][r] = bld<T>(rold, voff, unsigned(r) * colb + unsigned(f) * fbytes);
#endif
#if EIO_PREFETCH >= 2
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) m[f][r] = bld<T>(rmed, voff, unsigned(r) * colb + unsigned(f) * fbytes);
#endif

    // stage the current level with a 4-cell halo: every load of the stage is issued before
    // the first LDS write (register staging), so they are all in flight together
    const unsigned hbase = (unsigned(t.x0 - FAT) * unsigned(nz) + unsigned(t.z0 - FAT)) * SZ;
    {
        constexpr int NFILL = (TXH * ZH + NT - 1) / NT;
        T hf[NFILL][5];
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            const int c = e / ZH, z = e - c * ZH;
            const unsigned hv = e < TXH * ZH ? hbase + unsigned(c) * colb + unsigned(z) * SZ : 0u;
#pragma unroll
            for (int f = 0; f < 5; ++f) hf[it][f] = bld<T>(rcur, hv, unsigned(f) * fbytes);
        }
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            if (e < TXH * ZH) {
                const int c = e / ZH, z = e - c * ZH;
#pragma unroll
                for (int f = 0; f < 5; ++f) lds[f][c][z] = hf[it][f];
            }
        }
    }
    __syncthreads();

    // Free surface, top tile: the map M on every staged column (LDS row r is global row r
    // there), and the mapped rows of the own columns written back, so that this level is
    // M(cur) when it serves as the older level of the next step (the reference applies M to
    // the level in memory).  Other tiles reading these rows as halo apply M themselves; M
    // depends only on rows below the surface, which no launch writes concurrently.
    const bool fs_top = S.surface && t.z0 == FAT;
    if (fs_top) {  // uniform per tile
        for (int c = threadIdx.x; c < TXH; c += NT)
            fs_map_col<T>(&lds[0][c][0], &lds[1][c][0], &lds[3][c][0], &lds[4][c][0]);
        __syncthreads();
        const rsrc_t rcw = make_rsrc(curp, 5 * fbytes);
        for (int e = threadIdx.x; e < TX * (FAT + 1); e += NT) {
            const int cc = e / (FAT + 1), rr = SURF - FAT + (e - cc * (FAT + 1));
            if (t.x0 + cc < nx - FAT) {
                const unsigned vo = (unsigned(t.x0 + cc) * unsigned(nz) + unsigned(rr)) * SZ;
                bst(rcw, lds[0][FAT + cc][rr], vo, 0u);
                bst(rcw, lds[1][FAT + cc][rr], vo, fbytes);
                bst(rcw, lds[3][FAT + cc][rr], vo, 3 * fbytes);
                bst(rcw, lds[4][FAT + cc][rr], vo, 4 * fbytes);
            }
        }
    }

#if EIO_PREFETCH < 1
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) o[f][r] = bld<T>(rold, voff, unsigned(r) * colb + unsigned(f) * fbytes);
#endif
#if EIO_PREFETCH < 2
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) m[f][r] = bld<T>(rmed, voff, unsigned(r) * colb + unsigned(f) * fbytes);
#endif
    const T* cz = S.cz;
    const T* cx = S.cx;
    T dx_sxx[R], dx_sxz[R], dx_vx[R], dx_vz[R];
    {
        T q[R + 8];
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[2][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dx_sxx[r] = dm_q(q, r + 4, cx);
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[4][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dx_sxz[r] = dp_q(q, r + 4, cx);
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[0][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dx_vx[r] = dp_q(q, r + 4, cx);
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[1][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dx_vz[r] = dm_q(q, r + 4, cx);
    }

    const bool zok = iz < nz - FAT;
    T val[5][R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ix = ix0 + r;
        const bool ok = zok && (ix < nx - FAT);
        const int c = c0 + r;
        const T dz_sxz = dp_q(&lds[4][c][zz], 0, cz);
        const T dz_szz = dm_q(&lds[3][c][zz], 0, cz);
        const T dz_vz = dp_q(&lds[1][c][zz], 0, cz);
        const T dz_vx = dm_q(&lds[0][c][zz], 0, cz);
        const T d = ok ? damp_at(S, iz, ix) : T(1);
        T a0, a1, a2, a3, a4;
        if (!fs_top) {
            a0 = d * o[0][r] + m[0][r] * (dx_sxx[r] + dz_sxz);
            a1 = d * o[1][r] + m[1][r] * (dz_szz + dx_sxz[r]);
            a2 = d * o[2][r] + (m[2][r] * dx_vx[r] + m[3][r] * dz_vz);
            a3 = d * o[3][r] + (m[3][r] * dx_vx[r] + m[2][r] * dz_vz);
            a4 = d * o[4][r] + m[4][r] * (dz_vx + dx_vz[r]);
        } else {  // kernelsGpuElastic.cu:589-776: updated rows per component, surface modulus
            const bool on = iz == SURF, below = iz > SURF;
            a0 = d * o[0][r] + ((on || below) ? m[0][r] * (dx_sxx[r] + dz_sxz) : T(0));
            a1 = d * o[1][r] + (below ? m[1][r] * (dz_szz + dx_sxz[r]) : T(0));
            a2 = d * o[2][r] + (below ? (m[2][r] * dx_vx[r] + m[3][r] * dz_vz)
                                      : (on ? (m[2][r] - m[3][r] * m[3][r] / m[2][r]) * dx_vx[r] : T(0)));
            a3 = d * o[3][r] + (below ? (m[3][r] * dx_vx[r] + m[2][r] * dz_vz) : T(0));
            a4 = d * o[4][r] + (below ? m[4][r] * (dz_vx + dx_vz[r]) : T(0));
        }
        if (MODE & F_SS) {  // secondary source sampled at it2-1 (kernelsGpuElastic.cu:1363-1368)
            const T fl = S.filt[st.it2 - 1], fr = S.filt[S.h + st.it2 - 1];
            const rsrc_t rl = make_rsrc(S.ss[sspar], 5 * fbytes);
            const rsrc_t rr = make_rsrc(S.ss[sspar ^ 1], 5 * fbytes);
            const unsigned so = unsigned(r) * colb;
            a0 += bld<T>(rl, voff, so) * fl + bld<T>(rr, voff, so) * fr;
            a1 += bld<T>(rl, voff, so + fbytes) * fl + bld<T>(rr, voff, so + fbytes) * fr;
            a2 += bld<T>(rl, voff, so + 2 * fbytes) * fl + bld<T>(rr, voff, so + 2 * fbytes) * fr;
            a3 += bld<T>(rl, voff, so + 3 * fbytes) * fl + bld<T>(rr, voff, so + 3 * fbytes) * fr;
            a4 += bld<T>(rl, voff, so + 4 * fbytes) * fl + bld<T>(rr, voff, so + 4 * fbytes) * fr;
        }
        val[0][r] = d * a0;
        val[1][r] = d * a1;
        val[2][r] = d * a2;
        val[3][r] = d * a3;
        val[4][r] = d * a4;
    }

    const bool has_src = (MODE & (F_INJ_COARSE | F_INJ_FINE)) && (t.src_e > t.src_b);
    const bool has_rec = (MODE & F_REC) && (t.rec_e > t.rec_b);
    if (has_src || has_rec) {  // uniform per tile
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f) lds[f][c0 + r][zz] = val[f][r];
        __syncthreads();
        if (has_src) {
            for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NT) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T* tr = S.src_tr[comp];
                T s;
                if (MODE & F_INJ_COARSE) {  // kernelsGpuElastic.cu:172-189 with it2-1
                    const long b = long(row) * S.nts + st.its;
                    s = tr[b] * S.filt[st.it2 - 1] + tr[b + 1] * S.filt[S.h + st.it2 - 1];
                } else {             // pre-fine-sampled source at itw-1 = k (kernelsGpuElastic.cu:121-168)
                    s = tr[long(row) * S.ntw + k];
                }
                lds[comp][FAT + lx][FAT + lz] += damp_at(S, t.z0 + lz, t.x0 + lx) * s;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int f = 0; f < 5; ++f) val[f][r] = lds[f][c0 + r][zz];
        }
        if (has_rec) {  // kernelsGpuElastic.cu:7-60
            const T fa = S.filt[st.it2], fb = S.filt[S.h + st.it2];
            for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NT) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T v = lds[comp][FAT + lx][FAT + lz];
                T* tr = S.rec_tr[comp] + long(row) * S.nts + st.its;
                tr[0] += v * fa;
                tr[1] += v * fb;
            }
        }
    }

    // Decimation.  A slot's first contribution in the pass overwrites it (no zero fill of
    // the storage): slot its at the first fine step of interval dlo, slot its+1 at the first
    // fine step of every interval except the one ending on dhi (that slot already holds the
    // contributions of the later segment, recomputed earlier).
    rsrc_t rwx0, rwx1, rwz0, rwz1;
    T wfa = T(0), wfb = T(0);
    bool sta = false, stb = false;
    if (MODE & F_DEC) {  // coarse slices its and its+1 (kernelsGpuElastic.cu:1337-1346)
        const size_t s0 = size_t(st.its % S.wcap) * n, s1 = size_t((st.its + 1) % S.wcap) * n;
        rwx0 = make_rsrc(S.Wvx + s0, fbytes);
        rwx1 = make_rsrc(S.Wvx + s1, fbytes);
        rwz0 = make_rsrc(S.Wvz + s0, fbytes);
        rwz1 = make_rsrc(S.Wvz + s1, fbytes);
        wfa = S.filt[st.it2];
        wfb = S.filt[S.h + st.it2];
        sta = st.it2 == 1 && st.its == dlo;
        stb = st.it2 == 1 && st.its + 1 != dhi;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (zok && ix0 + r < nx - FAT) {
            const unsigned so = unsigned(r) * colb;
#pragma unroll
            for (int f = 0; f < 5; ++f) bst(rold, val[f][r], voff, so + unsigned(f) * fbytes);
            if ((MODE & F_REC) && S.wfld) {  // uniform per shot
                const T v5[5] = {val[0][r], val[1][r], val[2][r], val[3][r], val[4][r]};
                save_wfld<T>(S.wfld, n, st.its, size_t(ix0 + r) * nz + iz, v5, S.filt[st.it2],
                             S.filt[S.h + st.it2]);
            }
            if (MODE & F_DEC) {
                T xa = val[0][r] * wfa, za = val[1][r] * wfa;
                T xb = val[0][r] * wfb, zb = val[1][r] * wfb;
                if (!sta) {
                    xa = bld<T>(rwx0, voff, so) + xa;
                    za = bld<T>(rwz0, voff, so) + za;
                }
                if (!stb) {
                    xb = bld<T>(rwx1, voff, so) + xb;
                    zb = bld<T>(rwz1, voff, so) + zb;
                }
                bst(rwx0, xa, voff, so);
                bst(rwz0, za, voff, so);
                bst(rwx1, xb, voff, so);
                bst(rwz1, zb, voff, so);
            }
        }
    }
}

// ------------------------------------------------------------------ adjoint step
// LDS holds the media-scaled current level: s0 = rhoxDtw*vx, s1 = rhozDtw*vz,
// a = l2m*sxx + lam*szz, b = lam*sxx + l2m*szz, s6 = muxzDtw*sxz.  The reference applies
// the x derivative to l2m*sxx and lam*szz separately and adds; by linearity one derivative
// of their sum is the same operator.
template <typename T, int R, int WX, int OCC, int MODE>
__global__ __launch_bounds__(64 * WX) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_step_adj(const ShotDev<T>* __restrict__ shots, const TileDev* __restrict__ tiles,
                const Entry* __restrict__ ents, int ntiles, int k, int sspar) {
    constexpr int NT = 64 * WX;
    constexpr int TX = WX * R;
    constexpr int TXH = TX + 2 * FAT;
    constexpr unsigned SZ = sizeof(T);
    __shared__ T lds[5][TXH][ZH];

    const TileDev t = tiles[xcd_remap(blockIdx.x, ntiles)];
    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    const unsigned fbytes = n * SZ, colb = unsigned(nz) * SZ;
    const StepTime st = adj_time(k, S.sub, S.nts);
    T* oldp = (k & 1) ? S.B : S.A;
    const T* curp = (k & 1) ? S.A : S.B;
    const rsrc_t rold = make_rsrc(oldp, 5 * fbytes);
    const rsrc_t rcur = make_rsrc(curp, 5 * fbytes);
    const rsrc_t rmed = make_rsrc(S.med, 5 * fbytes);

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 + lane, zz = lane + FAT;
    const int c0 = FAT + w * R, ix0 = t.x0 + w * R;
    const unsigned voff = (unsigned(ix0) * unsigned(nz) + unsigned(iz)) * SZ;

    T o[5][R];
#if EIO_PREFETCH >= 1
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) o[f][r] = bld<T>(rold, voff, unsigned(r) * colb + unsigned(f) * fbytes);
#endif

    const unsigned hbase = (unsigned(t.x0 - FAT) * unsigned(nz) + unsigned(t.z0 - FAT)) * SZ;
    {
        // all stage loads in flight before the first LDS write (register staging)
        constexpr int NFILL = (TXH * ZH + NT - 1) / NT;
        T hv_[NFILL][5], hm[NFILL][5];
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            const int c = e / ZH, z = e - c * ZH;
            const unsigned hv = e < TXH * ZH ? hbase + unsigned(c) * colb + unsigned(z) * SZ : 0u;
#pragma unroll
            for (int f = 0; f < 5; ++f) hv_[it][f] = bld<T>(rcur, hv, unsigned(f) * fbytes);
#pragma unroll
            for (int f = 0; f < 5; ++f) hm[it][f] = bld<T>(rmed, hv, unsigned(f) * fbytes);
        }
        const bool fsm = S.surface && t.z0 == FAT;  // top tile of a free-surface shot
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            if (e < TXH * ZH) {
                const int c = e / ZH, z = e - c * ZH;
                if (!fsm) {
                    lds[0][c][z] = hm[it][0] * hv_[it][0];
                    lds[1][c][z] = hm[it][1] * hv_[it][1];
                    lds[2][c][z] = hm[it][2] * hv_[it][2] + hm[it][3] * hv_[it][3];
                    lds[3][c][z] = hm[it][3] * hv_[it][2] + hm[it][2] * hv_[it][3];
                    lds[4][c][z] = hm[it][4] * hv_[it][4];
                } else {
                    // the transpose of the free-surface body: each product only where the
                    // forward updates that component (LDS row z is global row z here), and the
                    // surface row's sigmaxx with its modulus
                    const bool on = z == SURF, below = z > SURF;
                    lds[0][c][z] = (on || below) ? hm[it][0] * hv_[it][0] : T(0);
                    lds[1][c][z] = below ? hm[it][1] * hv_[it][1] : T(0);
                    lds[2][c][z] = below ? hm[it][2] * hv_[it][2] + hm[it][3] * hv_[it][3]
                                         : ((on && hm[it][2] != T(0))
                                                ? (hm[it][2] - hm[it][3] * hm[it][3] / hm[it][2]) * hv_[it][2]
                                                : T(0));
                    lds[3][c][z] = below ? hm[it][3] * hv_[it][2] + hm[it][2] * hv_[it][3] : T(0);
                    lds[4][c][z] = below ? hm[it][4] * hv_[it][4] : T(0);
                }
            }
        }
    }
    __syncthreads();

#if EIO_PREFETCH < 1
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) o[f][r] = bld<T>(rold, voff, unsigned(r) * colb + unsigned(f) * fbytes);
#endif
    const T* cz = S.cz;
    const T* cx = S.cx;
    T dx_a[R], dx_s6[R], dx_s0[R], dx_s1[R];
    {
        T q[R + 8];
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[2][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dx_a[r] = dm_q(q, r + 4, cx);
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[4][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dx_s6[r] = dp_q(q, r + 4, cx);
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[0][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dx_s0[r] = dp_q(q, r + 4, cx);
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[1][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dx_s1[r] = dm_q(q, r + 4, cx);
    }

    const bool zok = iz < nz - FAT;
    T val[5][R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ix = ix0 + r;
        const bool ok = zok && (ix < nx - FAT);
        const int c = c0 + r;
        const T dz_s6 = dp_q(&lds[4][c][zz], 0, cz);
        const T dz_b = dm_q(&lds[3][c][zz], 0, cz);
        const T dz_s1 = dp_q(&lds[1][c][zz], 0, cz);
        const T dz_s0 = dm_q(&lds[0][c][zz], 0, cz);
        const T d = ok ? damp_at(S, iz, ix) : T(1);
        val[0][r] = d * (d * o[0][r] - (dx_a[r] + dz_s6));
        val[1][r] = d * (d * o[1][r] - (dz_b + dx_s6[r]));
        val[2][r] = d * (d * o[2][r] - dx_s0[r]);
        val[3][r] = d * (d * o[3][r] - dz_s1);
        val[4][r] = d * (d * o[4][r] - (dz_s0 + dx_s1[r]));
    }

    // Free surface, top tile: M^T on the new adjoint state, before the data injection (the
    // staged rows of the own columns are global rows here)
    if (S.surface && t.z0 == FAT) {  // uniform per tile
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f) lds[f][c0 + r][zz] = val[f][r];
        __syncthreads();
        for (int c = threadIdx.x; c < TX; c += NT)
            fs_map_col_transpose<T>(&lds[0][FAT + c][0], &lds[1][FAT + c][0], &lds[3][FAT + c][0],
                                    &lds[4][FAT + c][0]);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f) val[f][r] = lds[f][c0 + r][zz];
    }

    // adjoint: inject = receiver (data) entries, record = source entries
    const bool has_inj = (t.rec_e > t.rec_b);
    const bool has_ext = (MODE == M_NLADJ) && (t.src_e > t.src_b);
    if (has_inj || has_ext) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f) lds[f][c0 + r][zz] = val[f][r];
        __syncthreads();
        if (has_inj) {  // kernelsGpuElastic.cu:191-229 with it2+1
            const T fa = S.filt[st.it2 + 1], fb = S.filt[S.h + st.it2 + 1];
            for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NT) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T* tr = S.rec_tr[comp] + long(row) * S.nts + st.its;
                const T s = tr[0] * fa + tr[1] * fb;
                lds[comp][FAT + lx][FAT + lz] += damp_at(S, t.z0 + lz, t.x0 + lx) * s;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int f = 0; f < 5; ++f) val[f][r] = lds[f][c0 + r][zz];
        }
        if (has_ext) {  // kernelsGpuElastic.cu:100-116
            const T fa = S.filt[st.it2], fb = S.filt[S.h + st.it2];
            for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NT) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T v = lds[comp][FAT + lx][FAT + lz];
                T* tr = S.src_tr[comp] + long(row) * S.nts + st.its;
                tr[0] += v * fa;
                tr[1] += v * fb;
            }
        }
    }

    rsrc_t rsl, rsr;
    T efa = T(0), efb = T(0);
    // Adjoint fields of interval its go to ss[sspar] (slot its, weight F[it2]) and
    // ss[sspar^1] (slot its+1, weight F[h+it2]).  Slots are never cleared: the imaging
    // condition reads slot its+1 after the interval, and slot its starts fresh with the
    // interval's first step.  (The slot of its = nts-1 is zero from the call's set-up.)
    const bool first = st.it2 == S.sub - 1, last = st.it2 == 0;
    if (MODE == M_BADJ) {  // kernelsGpuElastic.cu:1369-1376
        rsl = make_rsrc(S.ss[sspar], 5 * fbytes);
        rsr = make_rsrc(S.ss[sspar ^ 1], 5 * fbytes);
        efa = S.filt[st.it2];
        efb = S.filt[S.h + st.it2];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (zok && ix0 + r < nx - FAT) {
            const unsigned so = unsigned(r) * colb;
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const unsigned sf = so + unsigned(f) * fbytes;
                bst(rold, val[f][r], voff, sf);
                if (MODE == M_NLADJ && S.wfld) {  // uniform per shot
                    T* wa = S.wfld + size_t(st.its) * 5 * n + size_t(f) * n + size_t(ix0 + r) * nz + iz;
                    wa[0] += val[f][r] * S.filt[st.it2];
                    wa[size_t(5) * n] += val[f][r] * S.filt[S.h + st.it2];
                }
                if (MODE == M_BADJ) {
                    // slot its: the interval's first (reverse) step overwrites it; slot its+1:
                    // its weight F[h+0] is 0 at it2 = 0 (nothing to add)
                    const T l = val[f][r] * efa;
                    bst(rsl, first ? l : bld<T>(rsl, voff, sf) + l, voff, sf);
                    if (!last) bst(rsr, bld<T>(rsr, voff, sf) + val[f][r] * efb, voff, sf);
                }
            }
        }
    }
}

// ------------------------------------------------------------------ Born imaging
// FWD (kernelsGpuElastic.cu:1381-1459): ss[which] <- secondary source at coarse time its.
// ADJ (kernelsGpuElastic.cu:1538-1613): dm += imaging condition of ss[which] at its (the
//      reference then clears the slice; here the next interval's first step overwrites it).
template <typename T, int R, int WX, bool ADJ>
__global__ __launch_bounds__(64 * WX) void k_imaging(const ShotDev<T>* __restrict__ shots,
                                                     const TileDev* __restrict__ tiles, int ntiles,
                                                     int its, int which) {
    constexpr int NT = 64 * WX;
    constexpr int TX = WX * R;
    constexpr int TXH = TX + 2 * FAT;
    constexpr unsigned SZ = sizeof(T);
    __shared__ T lds[2][TXH][ZH];

    const TileDev t = tiles[xcd_remap(blockIdx.x, ntiles)];
    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    const unsigned fbytes = n * SZ, colb = unsigned(nz) * SZ;
    // slices its-1, its, its+1 of the coarse storage ring (ends are clamped; unused there)
    const size_t sc = size_t(its % S.wcap) * n;
    const size_t sm = size_t((its > 0 ? its - 1 : its) % S.wcap) * n;
    const size_t sp = size_t((its < S.nts - 1 ? its + 1 : its) % S.wcap) * n;
    const rsrc_t rwx = make_rsrc(S.Wvx + sc, fbytes), rwz = make_rsrc(S.Wvz + sc, fbytes);
    const rsrc_t rwxm = make_rsrc(S.Wvx + sm, fbytes), rwzm = make_rsrc(S.Wvz + sm, fbytes);
    const rsrc_t rwxp = make_rsrc(S.Wvx + sp, fbytes), rwzp = make_rsrc(S.Wvz + sp, fbytes);

    // Halo fill.  Only update-region cells of a slot are ever written (the decimation
    // overwrites or accumulates them), so FAT cells are masked to the zero the wavefield
    // holds there instead of being read.
    const unsigned hbase = (unsigned(t.x0 - FAT) * unsigned(nz) + unsigned(t.z0 - FAT)) * SZ;
    for (int e = threadIdx.x; e < TXH * ZH; e += NT) {
        const int c = e / ZH, z = e - c * ZH;
        const int hz = t.z0 - FAT + z, hx = t.x0 - FAT + c;
        const bool upd = hz >= FAT && hz < nz - FAT && hx >= FAT && hx < nx - FAT;
        const unsigned hv = hbase + unsigned(c) * colb + unsigned(z) * SZ;
        const T a = bld<T>(rwx, hv, 0u), b = bld<T>(rwz, hv, 0u);
        lds[0][c][z] = upd ? a : T(0);
        lds[1][c][z] = upd ? b : T(0);
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 + lane, zz = lane + FAT;
    const int c0 = FAT + w * R, ix0 = t.x0 + w * R;
    const unsigned voff = (unsigned(ix0) * unsigned(nz) + unsigned(iz)) * SZ;
    const T* cz = S.cz;
    const T* cx = S.cx;
    const rsrc_t rss = make_rsrc(S.ss[which], 5 * fbytes);
    const rsrc_t rdm = make_rsrc(S.dm, 5 * fbytes);
    const T dinv = S.filt[2 * S.h];  // 1/(2 dts)

    T dvx_dx[R], dvz_dxm[R];
    {
        T q[R + 8];
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[0][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dvx_dx[r] = dp_q(q, r + 4, cx);
#pragma unroll
        for (int j = 0; j < R + 8; ++j) q[j] = lds[1][c0 - 4 + j][zz];
#pragma unroll
        for (int r = 0; r < R; ++r) dvz_dxm[r] = dm_q(q, r + 4, cx);
    }
    const bool zok = iz < nz - FAT;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ix = ix0 + r;
        if (!(zok && ix < nx - FAT)) continue;
        const int c = c0 + r;
        const unsigned so = unsigned(r) * colb;
        const T dvz_dz = dp_q(&lds[1][c][zz], 0, cz);
        const T dvx_dzm = dm_q(&lds[0][c][zz], 0, cz);
        const T sxz_op = dvx_dzm + dvz_dxm[r];
        T tx, tz;
        if (its == 0) {
            tx = -bld<T>(rwxp, voff, so);
            tz = -bld<T>(rwzp, voff, so);
        } else if (its == S.nts - 1) {
            tx = bld<T>(rwxm, voff, so);
            tz = bld<T>(rwzm, voff, so);
        } else {
            tx = bld<T>(rwxm, voff, so) - bld<T>(rwxp, voff, so);
            tz = bld<T>(rwzm, voff, so) - bld<T>(rwzp, voff, so);
        }
        if (!ADJ) {
            const T d0 = bld<T>(rdm, voff, so), d1 = bld<T>(rdm, voff, so + fbytes);
            const T d2 = bld<T>(rdm, voff, so + 2 * fbytes), d3 = bld<T>(rdm, voff, so + 3 * fbytes);
            const T d4 = bld<T>(rdm, voff, so + 4 * fbytes);
            const T dl = d2 * (dvx_dx[r] + dvz_dz);
            bst(rss, d0 * tx * dinv, voff, so);
            bst(rss, d1 * tz * dinv, voff, so + fbytes);
            bst(rss, dl + T(2) * d3 * dvx_dx[r], voff, so + 2 * fbytes);
            bst(rss, dl + T(2) * d3 * dvz_dz, voff, so + 3 * fbytes);
            bst(rss, d4 * sxz_op, voff, so + 4 * fbytes);
        } else {
            const T a0 = bld<T>(rss, voff, so), a1 = bld<T>(rss, voff, so + fbytes);
            const T a2 = bld<T>(rss, voff, so + 2 * fbytes), a3 = bld<T>(rss, voff, so + 3 * fbytes);
            const T a4 = bld<T>(rss, voff, so + 4 * fbytes);
            bst(rdm, bld<T>(rdm, voff, so) + a0 * tx * dinv, voff, so);
            bst(rdm, bld<T>(rdm, voff, so + fbytes) + a1 * tz * dinv, voff, so + fbytes);
            bst(rdm, bld<T>(rdm, voff, so + 2 * fbytes) + (a2 + a3) * (dvx_dx[r] + dvz_dz), voff,
                so + 2 * fbytes);
            bst(rdm, bld<T>(rdm, voff, so + 3 * fbytes) + T(2) * (dvx_dx[r] * a2 + dvz_dz * a3), voff,
                so + 3 * fbytes);
            bst(rdm, bld<T>(rdm, voff, so + 4 * fbytes) + a4 * sxz_op, voff, so + 4 * fbytes);
        }
    }
}

// out[i] += sum_s acc_s[i] (Born adjoint shot stacking, one field)
template <typename T>
__global__ __launch_bounds__(NTHR) void k_stack(T* __restrict__ out, const T* const* __restrict__ acc,
                                                int nacc, long n) {
    const long i = long(blockIdx.x) * NTHR + threadIdx.x;
    if (i >= n) return;
    T s = out[i];
    for (int a = 0; a < nacc; ++a) s += acc[a][i];
    out[i] = s;
}

// plain copy used as the bandwidth / counter calibration reference
__global__ __launch_bounds__(NTHR) void k_copy16(float4* __restrict__ dst,
                                                 const float4* __restrict__ src, long n16) {
    const long stride = long(gridDim.x) * NTHR;
    for (long i = long(blockIdx.x) * NTHR + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// achievable-bandwidth probe: one 16-byte vector per thread (no loop), nontemporal stores
__global__ __launch_bounds__(NTHR) void k_copy16_nt(float4* __restrict__ dst,
                                                    const float4* __restrict__ src, long n16) {
    const long i = long(blockIdx.x) * NTHR + threadIdx.x;
    if (i < n16) {
        const float4 v = src[i];
        __builtin_nontemporal_store(v.x, &dst[i].x);
        __builtin_nontemporal_store(v.y, &dst[i].y);
        __builtin_nontemporal_store(v.z, &dst[i].z);
        __builtin_nontemporal_store(v.w, &dst[i].w);
    }
}

// the same probe at the access width of the stencil kernels' per-cell loads and stores: one
// 4-byte word per thread (a wave moves 256 contiguous bytes per instruction)
__global__ __launch_bounds__(NTHR) void k_copy4(float* __restrict__ dst, const float* __restrict__ src, long n4) {
    const long i = long(blockIdx.x) * NTHR + threadIdx.x;
    if (i < n4) dst[i] = src[i];
}

// element copy (checkpoint save/restore; any size and alignment)
template <typename T>
__global__ __launch_bounds__(NTHR) void k_copy(T* __restrict__ dst, const T* __restrict__ src, long n) {
    const long stride = long(gridDim.x) * NTHR;
    for (long i = long(blockIdx.x) * NTHR + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// zero-fill (graph-capturable replacement for hipMemsetAsync)
template <typename T>
__global__ __launch_bounds__(NTHR) void k_fill0(T* __restrict__ p, long n) {
    const long i = long(blockIdx.x) * NTHR + threadIdx.x;
    if (i < n) p[i] = T(0);
}

}  // namespace eio
