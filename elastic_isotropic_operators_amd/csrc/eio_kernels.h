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
// Work decomposition: a tile is 64 (z, one wave64 lane per z) x 4R (x) update-region cells;
// a 256-thread workgroup (4 waves) owns one tile and each lane computes R consecutive x
// cells.  The current time level is staged through LDS with a 4-cell halo so every
// staggered first derivative is read from LDS; x derivatives reuse a register window of
// R+8 values.  Old level and media are streamed once from HBM, the new level is written
// once (80 B per FP32 cell-update for the forward and adjoint steps).
//
// Sponge fusion.  The reference damps BOTH time levels after every step
// (p0 <- d*(p0 + A p1 + s), p1 <- d*p1, then swaps).  The damped p1 is only ever used as
// the next step's older level, so the fused kernel keeps p1 undamped in memory and applies
// the second factor when it reads the older level: new = d*(d*old + A cur + s).  This is
// the same arithmetic as the reference sequence with one pass over memory.
#pragma once
#include <hip/hip_runtime.h>

namespace eio {

constexpr int FAT = 4;
constexpr int TZ = 64;        // z extent of a tile = wave64 lanes
constexpr int NTHR = 256;     // 4 waves per workgroup
constexpr int ZH = TZ + 2 * FAT;

// Modes of the fused forward kernel
constexpr int M_NL = 0;       // NL forward: coarse source injection with time interpolation, record
constexpr int M_BSRC = 1;     // Born source pass: fine source injection, coarse Vx/Vz decimation
constexpr int M_BSCAT = 2;    // Born scattered pass: secondary-source injection, record
// Modes of the fused adjoint kernel
constexpr int M_NLADJ = 0;    // NL adjoint: data injection, source-trace extraction
constexpr int M_BADJ = 1;     // Born receiver pass: data injection, adjoint-field extraction

template <typename T>
struct ShotDev {
    int nz, nx, nts, sub, h, ntw, minPad, pad0;
    T cz[4], cx[4];
    T* fa[5];              // state buffer A (older level at even steps)
    T* fb[5];              // state buffer B
    const T* med[5];       // rhoxDtw, rhozDtw, lamb2MuDtw, lambDtw, muxzDtw
    const T* damp;         // [minPad] cosine sponge
    const T* filt;         // [2h] time interpolation filter
    T* src_tr[5];          // source traces (NL: [n][nts], Born: [n][ntw])
    T* rec_tr[5];          // receiver traces [n][nts]
    T* ss[2][5];           // Born secondary sources / adjoint fields, ping-pong
    T* Wvx;                // Born coarse storage [nts][nz*nx]
    T* Wvz;
    T* dm[5];              // Born perturbation (forward) or per-shot imaging accumulator (adjoint)
};

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

// Bijective block -> tile map that gives each of the 8 XCDs a contiguous range of tiles
// (workgroups are dispatched round-robin over XCDs), so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, j = b >> 3;
    return x * q + (x < r ? x : r) + j;
}

template <typename T>
__device__ __forceinline__ T damp_at(const ShotDev<T>& S, int iz, int ix) {
    const int d = min(min(iz - FAT, ix - FAT), min(S.nz - iz - 1 - FAT, S.nx - ix - 1 - FAT));
    return d < S.minPad ? S.damp[d] : T(1);
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
template <typename T, int R, int MODE>
__global__ __launch_bounds__(NTHR) void k_step_fwd(const ShotDev<T>* __restrict__ shots,
                                                   const TileDev* __restrict__ tiles,
                                                   const Entry* __restrict__ ents, int ntiles,
                                                   int k, int sspar) {
    constexpr int TX = 4 * R;
    constexpr int TXH = TX + 2 * FAT;
    __shared__ T lds[5][TXH][ZH];

    const TileDev t = tiles[xcd_remap(blockIdx.x, ntiles)];
    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const StepTime st = fwd_time(k, S.sub);
    T* const* oldp = (k & 1) ? S.fb : S.fa;
    T* const* curp = (k & 1) ? S.fa : S.fb;

    // stage the current level with a 4-cell halo (cells beyond the array read as 0)
    for (int e = threadIdx.x; e < TXH * ZH; e += NTHR) {
        const int c = e / ZH, zz = e - c * ZH;
        const int gx = t.x0 - FAT + c, gz = t.z0 - FAT + zz;
        const bool ok = (gx < nx) && (gz < nz);
        const long gi = long(gx) * nz + gz;
#pragma unroll
        for (int f = 0; f < 5; ++f) lds[f][c][zz] = ok ? curp[f][gi] : T(0);
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 + lane, zz = lane + FAT;
    const int c0 = FAT + w * R, ix0 = t.x0 + w * R;
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

    T val[5][R];
    const bool zok = iz < nz - FAT;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ix = ix0 + r;
        const bool ok = zok && (ix < nx - FAT);
        const long i = long(ix) * nz + iz;
        const int c = c0 + r;
        const T* lsxz = &lds[4][c][zz];
        const T* lszz = &lds[3][c][zz];
        const T* lvz = &lds[1][c][zz];
        const T* lvx = &lds[0][c][zz];
        const T dz_sxz = dp_q(lsxz, 0, cz);
        const T dz_szz = dm_q(lszz, 0, cz);
        const T dz_vz = dp_q(lvz, 0, cz);
        const T dz_vx = dm_q(lvx, 0, cz);
        T o[5], m[5];
        if (ok) {
#pragma unroll
            for (int f = 0; f < 5; ++f) { o[f] = oldp[f][i]; m[f] = S.med[f][i]; }
        } else {
#pragma unroll
            for (int f = 0; f < 5; ++f) { o[f] = T(0); m[f] = T(0); }
        }
        const T d = ok ? damp_at(S, iz, ix) : T(1);
        T a0 = d * o[0] + m[0] * (dx_sxx[r] + dz_sxz);
        T a1 = d * o[1] + m[1] * (dz_szz + dx_sxz[r]);
        T a2 = d * o[2] + (m[2] * dx_vx[r] + m[3] * dz_vz);
        T a3 = d * o[3] + (m[3] * dx_vx[r] + m[2] * dz_vz);
        T a4 = d * o[4] + m[4] * (dz_vx + dx_vz[r]);
        if (MODE == M_BSCAT && ok) {  // secondary source sampled at it2-1 (kernelsGpuElastic.cu:1363-1368)
            const T fl = S.filt[st.it2 - 1], fr = S.filt[S.h + st.it2 - 1];
            T* const* sl = S.ss[sspar];
            T* const* sr = S.ss[sspar ^ 1];
            a0 += sl[0][i] * fl + sr[0][i] * fr;
            a1 += sl[1][i] * fl + sr[1][i] * fr;
            a2 += sl[2][i] * fl + sr[2][i] * fr;
            a3 += sl[3][i] * fl + sr[3][i] * fr;
            a4 += sl[4][i] * fl + sr[4][i] * fr;
        }
        val[0][r] = d * a0;
        val[1][r] = d * a1;
        val[2][r] = d * a2;
        val[3][r] = d * a3;
        val[4][r] = d * a4;
    }

    const bool has_src = (MODE != M_BSCAT) && (t.src_e > t.src_b);
    const bool has_rec = (MODE != M_BSRC) && (t.rec_e > t.rec_b);
    if (has_src || has_rec) {  // uniform per tile
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f) lds[f][c0 + r][zz] = val[f][r];
        __syncthreads();
        if (has_src) {
            for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NTHR) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T* tr = S.src_tr[comp];
                T s;
                if (MODE == M_NL) {  // kernelsGpuElastic.cu:172-189 with it2-1
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
            for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NTHR) {
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

#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ix = ix0 + r;
        if (zok && ix < nx - FAT) {
            const long i = long(ix) * nz + iz;
#pragma unroll
            for (int f = 0; f < 5; ++f) oldp[f][i] = val[f][r];
            if (MODE == M_BSRC) {  // kernelsGpuElastic.cu:1337-1346
                const long n = long(nz) * nx;
                const T fa = S.filt[st.it2], fb = S.filt[S.h + st.it2];
                T* wx = S.Wvx + long(st.its) * n + i;
                T* wz = S.Wvz + long(st.its) * n + i;
                wx[0] += val[0][r] * fa;
                wx[n] += val[0][r] * fb;
                wz[0] += val[1][r] * fa;
                wz[n] += val[1][r] * fb;
            }
        }
    }
}

// ------------------------------------------------------------------ adjoint step
// LDS holds the media-scaled current level: s0 = rhoxDtw*vx, s1 = rhozDtw*vz,
// a = l2m*sxx + lam*szz, b = lam*sxx + l2m*szz, s6 = muxzDtw*sxz.  The reference applies
// the x derivative to l2m*sxx and lam*szz separately and adds; by linearity one derivative
// of their sum is the same operator.
template <typename T, int R, int MODE>
__global__ __launch_bounds__(NTHR) void k_step_adj(const ShotDev<T>* __restrict__ shots,
                                                   const TileDev* __restrict__ tiles,
                                                   const Entry* __restrict__ ents, int ntiles,
                                                   int k, int sspar) {
    constexpr int TX = 4 * R;
    constexpr int TXH = TX + 2 * FAT;
    __shared__ T lds[5][TXH][ZH];

    const TileDev t = tiles[xcd_remap(blockIdx.x, ntiles)];
    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const StepTime st = adj_time(k, S.sub, S.nts);
    T* const* oldp = (k & 1) ? S.fb : S.fa;
    T* const* curp = (k & 1) ? S.fa : S.fb;

    for (int e = threadIdx.x; e < TXH * ZH; e += NTHR) {
        const int c = e / ZH, zz = e - c * ZH;
        const int gx = t.x0 - FAT + c, gz = t.z0 - FAT + zz;
        T s0 = T(0), s1 = T(0), sa = T(0), sb = T(0), s6 = T(0);
        if ((gx < nx) && (gz < nz)) {
            const long gi = long(gx) * nz + gz;
            const T vx = curp[0][gi], vz = curp[1][gi], sxx = curp[2][gi], szz = curp[3][gi],
                    sxz = curp[4][gi];
            const T l2m = S.med[2][gi], lam = S.med[3][gi];
            s0 = S.med[0][gi] * vx;
            s1 = S.med[1][gi] * vz;
            sa = l2m * sxx + lam * szz;
            sb = lam * sxx + l2m * szz;
            s6 = S.med[4][gi] * sxz;
        }
        lds[0][c][zz] = s0;
        lds[1][c][zz] = s1;
        lds[2][c][zz] = sa;
        lds[3][c][zz] = sb;
        lds[4][c][zz] = s6;
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 + lane, zz = lane + FAT;
    const int c0 = FAT + w * R, ix0 = t.x0 + w * R;
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

    T val[5][R];
    const bool zok = iz < nz - FAT;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ix = ix0 + r;
        const bool ok = zok && (ix < nx - FAT);
        const long i = long(ix) * nz + iz;
        const int c = c0 + r;
        const T dz_s6 = dp_q(&lds[4][c][zz], 0, cz);
        const T dz_b = dm_q(&lds[3][c][zz], 0, cz);
        const T dz_s1 = dp_q(&lds[1][c][zz], 0, cz);
        const T dz_s0 = dm_q(&lds[0][c][zz], 0, cz);
        T o[5];
        if (ok) {
#pragma unroll
            for (int f = 0; f < 5; ++f) o[f] = oldp[f][i];
        } else {
#pragma unroll
            for (int f = 0; f < 5; ++f) o[f] = T(0);
        }
        const T d = ok ? damp_at(S, iz, ix) : T(1);
        val[0][r] = d * (d * o[0] - (dx_a[r] + dz_s6));
        val[1][r] = d * (d * o[1] - (dz_b + dx_s6[r]));
        val[2][r] = d * (d * o[2] - dx_s0[r]);
        val[3][r] = d * (d * o[3] - dz_s1);
        val[4][r] = d * (d * o[4] - (dz_s0 + dx_s1[r]));
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
            for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NTHR) {
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
            for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NTHR) {
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

#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ix = ix0 + r;
        if (zok && ix < nx - FAT) {
            const long i = long(ix) * nz + iz;
#pragma unroll
            for (int f = 0; f < 5; ++f) oldp[f][i] = val[f][r];
            if (MODE == M_BADJ) {  // kernelsGpuElastic.cu:1369-1376
                const T fa = S.filt[st.it2], fb = S.filt[S.h + st.it2];
                T* const* sl = S.ss[sspar];
                T* const* sr = S.ss[sspar ^ 1];
#pragma unroll
                for (int f = 0; f < 5; ++f) {
                    sl[f][i] += val[f][r] * fa;
                    sr[f][i] += val[f][r] * fb;
                }
            }
        }
    }
}

// ------------------------------------------------------------------ Born imaging
// FWD (kernelsGpuElastic.cu:1381-1459): ss[which] <- secondary source at coarse time its.
// ADJ (kernelsGpuElastic.cu:1538-1613): dm += imaging condition of ss[which] at its, then
//      ss[which] <- 0 (the buffer becomes the next, zeroed, accumulation slice).
template <typename T, int R, bool ADJ>
__global__ __launch_bounds__(NTHR) void k_imaging(const ShotDev<T>* __restrict__ shots,
                                                  const TileDev* __restrict__ tiles, int ntiles,
                                                  int its, int which) {
    constexpr int TX = 4 * R;
    constexpr int TXH = TX + 2 * FAT;
    __shared__ T lds[2][TXH][ZH];

    const TileDev t = tiles[xcd_remap(blockIdx.x, ntiles)];
    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const long n = long(nz) * nx;
    const T* wx = S.Wvx + long(its) * n;
    const T* wz = S.Wvz + long(its) * n;

    for (int e = threadIdx.x; e < TXH * ZH; e += NTHR) {
        const int c = e / ZH, zz = e - c * ZH;
        const int gx = t.x0 - FAT + c, gz = t.z0 - FAT + zz;
        const bool ok = (gx < nx) && (gz < nz);
        const long gi = long(gx) * nz + gz;
        lds[0][c][zz] = ok ? wx[gi] : T(0);
        lds[1][c][zz] = ok ? wz[gi] : T(0);
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 + lane, zz = lane + FAT;
    const int c0 = FAT + w * R, ix0 = t.x0 + w * R;
    const T* cz = S.cz;
    const T* cx = S.cx;

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
        const long i = long(ix) * nz + iz;
        const int c = c0 + r;
        const T dvz_dz = dp_q(&lds[1][c][zz], 0, cz);
        const T dvx_dzm = dm_q(&lds[0][c][zz], 0, cz);
        const T sxz_op = dvx_dzm + dvz_dxm[r];
        T tx, tz;
        if (its == 0) {
            tx = -wx[i + n];
            tz = -wz[i + n];
        } else if (its == S.nts - 1) {
            tx = wx[i - n];
            tz = wz[i - n];
        } else {
            tx = wx[i - n] - wx[i + n];
            tz = wz[i - n] - wz[i + n];
        }
        const T dinv = S.filt[2 * S.h];  // 1/(2 dts), stored after the filter
        T* const* ss = S.ss[which];
        if (!ADJ) {
            ss[0][i] = S.dm[0][i] * tx * dinv;
            ss[1][i] = S.dm[1][i] * tz * dinv;
            const T dl = S.dm[2][i] * (dvx_dx[r] + dvz_dz);
            ss[2][i] = dl + T(2) * S.dm[3][i] * dvx_dx[r];
            ss[3][i] = dl + T(2) * S.dm[3][i] * dvz_dz;
            ss[4][i] = S.dm[4][i] * sxz_op;
        } else {
            const T a0 = ss[0][i], a1 = ss[1][i], a2 = ss[2][i], a3 = ss[3][i], a4 = ss[4][i];
            S.dm[0][i] += a0 * tx * dinv;
            S.dm[1][i] += a1 * tz * dinv;
            S.dm[2][i] += (a2 + a3) * (dvx_dx[r] + dvz_dz);
            S.dm[3][i] += T(2) * (dvx_dx[r] * a2 + dvz_dz * a3);
            S.dm[4][i] += a4 * sxz_op;
#pragma unroll
            for (int f = 0; f < 5; ++f) ss[f][i] = T(0);
        }
    }
}

// out[f][i] += sum_s acc_s[f][i] over the update region (Born adjoint shot stacking)
template <typename T>
__global__ __launch_bounds__(NTHR) void k_stack(T* __restrict__ out, const T* const* __restrict__ acc,
                                                int nacc, long n) {
    const long i = long(blockIdx.x) * NTHR + threadIdx.x;
    if (i >= n) return;
    T s = out[i];
    for (int a = 0; a < nacc; ++a) s += acc[a][i];
    out[i] = s;
}

// zero-fill (graph-capturable replacement for hipMemsetAsync)
template <typename T>
__global__ __launch_bounds__(NTHR) void k_fill0(T* __restrict__ p, long n) {
    const long i = long(blockIdx.x) * NTHR + threadIdx.x;
    if (i < n) p[i] = T(0);
}

}  // namespace eio
