// eio_step3.h -- three fine steps per launch of the FP32 NL forward (temporal blocking of
// depth 3).
//
// u1 = d(d u_old + A u_cur + s_k), u2 = d(d u_cur + A u1 + s_k+1), u3 = d(d u1 + A u2 + s_k+2)
// in one launch: u_old, u_cur and the media are read once and u2, u3 (the next launch's older
// and newer levels) written once, 100 B per THREE cell-updates (33 B) against 50 B for two
// steps per launch.  The price is redundant halo work: u2 is needed on a 4-cell halo of the
// owned cells, u1 on an 8-cell halo, so u_cur is staged with a 12-cell halo.
//
// Tile: owned rows [z0, z0+48) (z0 = 4 + 48 bz) x owned columns [x0, x0+TX), TX = WX R3.
// Lane l of every wave holds row z0-8+l for all three steps (64 rows: the owned rows and the
// 8-row halo of u1).  LDS: u_cur on (TX+24) x 72 cells from (x0-12, z0-12) (all five fields),
// and one (TX+16) x 64 area from (x0-8, z0-8) that holds u1, then u2, then u3 for the owned
// data.  Step 1 runs on the TX+16 columns [x0-8, x0+TX+8) (R3+2 per wave), step 2 on the TX+8
// columns [x0-4, x0+TX+4) (R3+1 per wave; its rows outside [z0-4, z0+52) are never read),
// step 3 on the owned cells.  The older level of step 2 is u_cur from the stage; that of step 3
// is u1, read from LDS before u2 overwrites it.
//
// Sources: each tile's list holds every source of the step-1 region (lcell relative to
// (z0-8, x0-8)); step 1 injects all of them, step 2 those of its region, step 3 the owned ones.
// Receivers are owned-only and recorded at each step.  A shot whose remaining step count is
// below 3 does 1 or 2 steps in its last launch (and stores nothing: no later launch reads it).
// Arithmetic per cell is that of k_step_fwd in mode M_NL.
#pragma once
#include "eio_step2.h"

namespace eio {

constexpr int TZ3 = 48;   // owned z rows of a three-step tile
constexpr int H3 = 8;     // lane row 0 = z0 - H3
constexpr int ZS3 = 72;   // staged rows of u_cur (z0-12 .. z0+60)

template <typename T, int R3, int WX>
__device__ __forceinline__ void step3_fwd_tile(const ShotDev<T>* __restrict__ shots, const TileDev t,
                                               const Entry* __restrict__ ents, int k, int par) {
    static_assert(sizeof(T) == 4, "the three-step forward is FP32");
    static_assert(WX == 8, "one column per wave of halo per step");
    constexpr int NT = 64 * WX;
    constexpr int TX = WX * R3;      // owned columns
    constexpr int R2 = R3 + 1;       // step-2 columns per wave
    constexpr int R1 = R3 + 2;       // step-1 columns per wave
    constexpr int TXA = TX + 16;     // columns of the u1/u2/u3 area (= step-1 columns)
    constexpr int TXC = TX + 24;     // staged u_cur columns
    constexpr unsigned SZ = sizeof(T);
    __shared__ __attribute__((aligned(16))) T lc[5][TXC][ZS3];
    __shared__ T la[5][TXA][Z1];

    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    const unsigned fbytes = n * SZ;
    const int nst = min(3, S.ntw - 1 - k);   // fine steps of this launch (1..3)
    const rsrc_t rold = make_rsrc(par ? S.C : S.A, 5 * fbytes);
    const rsrc_t rcur = make_rsrc(par ? S.D : S.B, 5 * fbytes);
    const rsrc_t rmed = make_rsrc(S.med, 5 * fbytes);
    const rsrc_t rw1 = make_rsrc(par ? S.A : S.C, 5 * fbytes);
    const rsrc_t rw2 = make_rsrc(par ? S.B : S.D, 5 * fbytes);
    const T* cz = S.cz;
    const T* cx = S.cx;

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 - H3 + lane;                     // row of this lane (all steps)
    const bool zin = iz >= FAT && iz < nz - FAT;         // update-region row
    const int izl = iz < 0 ? 0 : iz;                     // masked lanes load row 0
    // step-s columns of this wave, relative to the area origin x0-8
    const int j1 = w * R1;                               // step 1: area columns j1 ..
    const int j2 = 4 + w * R2;                           // step 2: area columns j2 ..
    const int j3 = 8 + w * R3;                           // step 3: area columns j3 ..
    const int lzb = lane < 4 ? 4 : (lane > 59 ? 59 : lane);   // step-2 rows whose z stencil is in the area
    const bool zown = lane >= H3 && lane < H3 + TZ3 && zin;

    // ---- stage u_cur on (TX+24) x 72 cells from (x0-12, z0-12), four rows per 16-byte load
    // (row quads start at z0-12+4q = 48 bz - 8 + 4q, a multiple of 4: a quad is wholly above
    // row 0 or wholly below it; the bottom edge cuts one, whose rows past the grid are masked)
    {
        constexpr int ZQ = ZS3 / 4;
        constexpr int NFILL = (TXC * ZQ + NT - 1) / NT;
        T hf[NFILL][5][4];
        unsigned hm[NFILL];
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            const int c = e / ZQ, q = e - c * ZQ;
            const int gx = t.x0 - 12 + c, gz = t.z0 - 12 + 4 * q;
            const bool in = e < TXC * ZQ && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
            const int nin = in ? (nz - gz < 4 ? nz - gz : 4) : 0;
            hm[it] = (1u << nin) - 1u;
            const unsigned hv = in ? (unsigned(gx) * unsigned(nz) + unsigned(gz)) * SZ : 0u;
#pragma unroll
            for (int f = 0; f < 5; ++f) bld4(rcur, hv, unsigned(f) * fbytes, hf[it][f]);
        }
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            if (e < TXC * ZQ) {
                const int c = e / ZQ, q = e - c * ZQ;
#pragma unroll
                for (int f = 0; f < 5; ++f) {
                    float4 v;
                    v.x = (hm[it] & 1u) ? hf[it][f][0] : 0.0f;
                    v.y = (hm[it] & 2u) ? hf[it][f][1] : 0.0f;
                    v.z = (hm[it] & 4u) ? hf[it][f][2] : 0.0f;
                    v.w = (hm[it] & 8u) ? hf[it][f][3] : 0.0f;
                    *reinterpret_cast<float4*>(&lc[f][c][4 * q]) = v;
                }
            }
        }
    }
    __syncthreads();

    // one step from the level in `src` (an LDS array [5][cols][rows] with the cell of this
    // lane's column r at src[f][c0 + r][row]) into `out`, for R columns starting at area
    // column jc; older level `old`, media loaded here
    auto step = [&](auto srcp, int srow, int c0, int jcol, const auto& old, auto& out, auto Rc) {
        constexpr int R = decltype(Rc)::value;
        T dx_sxx[R], dx_sxz[R], dx_vx[R], dx_vz[R];
        {
            T q[R + 8];
#pragma unroll
            for (int j = 0; j < R + 8; ++j) q[j] = srcp(2, c0 + j - 4, srow);
#pragma unroll
            for (int r = 0; r < R; ++r) dx_sxx[r] = dm_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R + 8; ++j) q[j] = srcp(4, c0 + j - 4, srow);
#pragma unroll
            for (int r = 0; r < R; ++r) dx_sxz[r] = dp_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R + 8; ++j) q[j] = srcp(0, c0 + j - 4, srow);
#pragma unroll
            for (int r = 0; r < R; ++r) dx_vx[r] = dp_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R + 8; ++j) q[j] = srcp(1, c0 + j - 4, srow);
#pragma unroll
            for (int r = 0; r < R; ++r) dx_vz[r] = dm_q(q, r + 4, cx);
        }
        pin_regs(dx_sxx);
        pin_regs(dx_sxz);
        pin_regs(dx_vx);
        pin_regs(dx_vz);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int ix = t.x0 - 8 + jcol + r;
            const bool ok = zin && ix >= FAT && ix < nx - FAT;
            const unsigned vo = (unsigned(ix < 0 ? 0 : (ix < nx ? ix : nx - 1)) * unsigned(nz) + unsigned(izl)) * SZ;
            T mm[5];
#pragma unroll
            for (int f = 0; f < 5; ++f) mm[f] = bld<T>(rmed, vo, unsigned(f) * fbytes);
            const T dz_sxz = dp_q(&srcp(4, c0 + r, srow), 0, cz);
            const T dz_szz = dm_q(&srcp(3, c0 + r, srow), 0, cz);
            const T dz_vz = dp_q(&srcp(1, c0 + r, srow), 0, cz);
            const T dz_vx = dm_q(&srcp(0, c0 + r, srow), 0, cz);
            const T d = ok ? damp_at(S, iz, ix) : T(1);
            const T a0 = d * old[0][r] + mm[0] * (dx_sxx[r] + dz_sxz);
            const T a1 = d * old[1][r] + mm[1] * (dz_szz + dx_sxz[r]);
            const T a2 = d * old[2][r] + (mm[2] * dx_vx[r] + mm[3] * dz_vz);
            const T a3 = d * old[3][r] + (mm[3] * dx_vx[r] + mm[2] * dz_vz);
            const T a4 = d * old[4][r] + mm[4] * (dz_vx + dx_vz[r]);
            out[0][r] = ok ? d * a0 : T(0);
            out[1][r] = ok ? d * a1 : T(0);
            out[2][r] = ok ? d * a2 : T(0);
            out[3][r] = ok ? d * a3 : T(0);
            out[4][r] = ok ? d * a4 : T(0);
#pragma unroll
            for (int f = 0; f < 5; ++f) asm volatile("" : "+v"(out[f][r]));
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    auto stage_at = [&](int f, int c, int row) -> T& { return lc[f][c][row]; };
    auto area_at = [&](int f, int c, int row) -> T& { return la[f][c][row]; };
    // sources of fine step ks whose area cell lies in [lo, hi) on both axes, added into la
    auto inject = [&](int ks, int lo_x, int hi_x, int lo_z, int hi_z) {
        const StepTime st = fwd_time(ks, S.sub);
        for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NT) {
            const Entry en = ents[e];
            const int lz = en.lcell & 255, lx = en.lcell >> 8;
            if (lz < lo_z || lz >= hi_z || lx < lo_x || lx >= hi_x) continue;
            const int comp = en.crow >> 24, row = en.crow & 0xffffff;
            const T* tr = S.src_tr[comp];
            const long b = long(row) * S.nts + st.its;
            const T sv = tr[b] * S.filt[st.it2 - 1] + tr[b + 1] * S.filt[S.h + st.it2 - 1];
            la[comp][lx][lz] += damp_at(S, t.z0 - H3 + lz, t.x0 - 8 + lx) * sv;
        }
    };
    auto record = [&](int ks) {
        const StepTime st = fwd_time(ks, S.sub);
        const T fa = S.filt[st.it2], fb = S.filt[S.h + st.it2];
        for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NT) {
            const Entry en = ents[e];
            const int lz = en.lcell & 255, lx = en.lcell >> 8;
            const int comp = en.crow >> 24, row = en.crow & 0xffffff;
            const T v = la[comp][lx][lz];
            T* tr = S.rec_tr[comp] + long(row) * S.nts + st.its;
            tr[0] += v * fa;
            tr[1] += v * fb;
        }
    };
    const bool has_src = t.src_e > t.src_b;
    const bool has_rec = t.rec_e > t.rec_b;

    // ---- step 1: u1 on area columns j1 .. j1+R1-1 (stage column = area column + 4; stage row
    // of this lane = lane + 4), older level u_old from HBM
    {
        T o[5][R1];
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int ix = t.x0 - 8 + j1 + r;
            const unsigned vo = (unsigned(ix < 0 ? 0 : (ix < nx ? ix : nx - 1)) * unsigned(nz) + unsigned(izl)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) o[f][r] = bld<T>(rold, vo, unsigned(f) * fbytes);
        }
        T u1[5][R1];
        step(stage_at, lane + 4, j1 + 4, j1, o, u1, std::integral_constant<int, R1>{});
#pragma unroll
        for (int r = 0; r < R1; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f) la[f][j1 + r][lane] = u1[f][r];
    }
    __syncthreads();
    if (has_src) {
        inject(k, 0, TXA, 0, Z1);
        __syncthreads();
    }
    if (has_rec) record(k);
    if (nst < 2) return;

    // ---- step 2: u2 on area columns j2 .. j2+R2-1, rows read clamped to lzb; older level
    // u_cur from the stage (stage column = area column + 4, row = lane + 4)
    T u1own[5][R3];   // u1 at this lane's step-3 cells: the older level of step 3
#pragma unroll
    for (int r = 0; r < R3; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) u1own[f][r] = la[f][j3 + r][lane];
    T u2[5][R2];
    {
        T oc[5][R2];
#pragma unroll
        for (int r = 0; r < R2; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f) oc[f][r] = lc[f][j2 + r + 4][lane + 4];
        step(area_at, lzb, j2, j2, oc, u2, std::integral_constant<int, R2>{});
    }
    __syncthreads();   // every wave has read u1
#pragma unroll
    for (int r = 0; r < R2; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) la[f][j2 + r][lane] = u2[f][r];
    __syncthreads();
    if (has_src) {
        inject(k + 1, 4, TX + 12, 4, 60);
        __syncthreads();
    }
    if (has_rec) record(k + 1);
    if (nst < 3) return;

    // ---- step 3: u3 on the owned columns j3 .. j3+R3-1, older level u1
    T u3[5][R3];
    step(area_at, lzb, j3, j3, u1own, u3, std::integral_constant<int, R3>{});
    // u2 at the owned cells (after its sources) for the store, then u3 staged over it when
    // the tile has sources or receivers
    T u2own[5][R3];
#pragma unroll
    for (int r = 0; r < R3; ++r)
#pragma unroll
        for (int f = 0; f < 5; ++f) u2own[f][r] = la[f][j3 + r][lane];
    if (has_src || has_rec) {
        __syncthreads();
        if (zown) {
#pragma unroll
            for (int r = 0; r < R3; ++r)
#pragma unroll
                for (int f = 0; f < 5; ++f) la[f][j3 + r][lane] = u3[f][r];
        }
        __syncthreads();
        if (has_src) {
            inject(k + 2, 8, TX + 8, H3, H3 + TZ3);
            __syncthreads();
            if (zown) {
#pragma unroll
                for (int r = 0; r < R3; ++r)
#pragma unroll
                    for (int f = 0; f < 5; ++f) u3[f][r] = la[f][j3 + r][lane];
            }
        }
        if (has_rec) record(k + 2);
    }
    // ---- store the next launch's older (u2) and newer (u3) levels on the owned cells
#pragma unroll
    for (int r = 0; r < R3; ++r) {
        const int ix = t.x0 - 8 + j3 + r;
        if (zown && ix < nx - FAT) {
            const unsigned vo = (unsigned(ix) * unsigned(nz) + unsigned(iz)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                bst(rw1, u2own[f][r], vo, unsigned(f) * fbytes);
                bst(rw2, u3[f][r], vo, unsigned(f) * fbytes);
            }
        }
    }
}

template <typename T, int R3, int WX, int OCC>
__global__ __launch_bounds__(64 * WX) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_step3_fwd(const ShotDev<T>* __restrict__ shots, const TileDev* __restrict__ tiles,
                 const Entry* __restrict__ ents, int ntiles, int k, int par) {
    step3_fwd_tile<T, R3, WX>(shots, tiles[xcd_remap(blockIdx.x, ntiles)], ents, k, par);
}

}  // namespace eio
