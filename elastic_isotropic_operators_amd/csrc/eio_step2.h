// eio_step2.h -- two fine steps per launch (temporal blocking): k_step2_fwd for the NL
// forward, the FWI forward and its segment recompute (DEC) and the Born source (FINE) and
// scattered (SS) passes; k_step2_adj for the Born receiver pass / FWI gradient and the NL
// adjoint.
//
// The single-step kernel moves 80 B per FP32 cell-update (old, current, media read; new
// written).  Two leapfrog steps u1 = d(d u_old + A u_cur + s_k), u2 = d(d u_cur + A u1 + s_k+1)
// fused in one launch read u_old, u_cur and the media once and write u1 and u2: 100 B per two
// cell-updates.  u1 is needed by step 2 on a 4-cell halo around the tile, so step 1 is
// computed redundantly on that halo (from u_cur on an 8-cell halo, staged in LDS); the
// redundant work costs L2 and VALU, not HBM.
//
// Because neighbouring tiles read u_cur's halo while this tile writes, the two results go
// to a second pair of state levels (A,B -> C,D, then C,D -> A,B): 4 levels per shot.
//
// Tile: 56 owned z rows [z0, z0+56) x TX owned columns [x0, x0+TX).  Lane l of every wave
// holds row z0-4+l (64 rows: owned rows plus the 4-row halo of u1).  Step 1 runs on the
// TX+8 columns [x0-4, x0+TX+4), R1 per wave; step 2 on the TX owned columns, R2 per wave.
// LDS: u_cur on (TX+16) x 72 cells from (x0-8, z0-8), u1 on (TX+8) x 64 from (x0-4, z0-4).
//
// Sources: each tile's list holds every source of the step-1 region (lcell relative to
// (z0-4, x0-4)), injected into its own copy of u1; only owned cells are stored.  Receivers
// are owned-only.  Arithmetic per cell is that of k_step_fwd in the matching mode.
//
// DEC (the FWI forward and the Born source pass, which keep a coarse record): both steps lie in the
// same coarse interval (pairs start at even fine steps and the fine sub-sampling is even), so
// their Vx, Vz contributions to slots its and its+1 (kernelsGpuElastic.cu:1337-1346) are
// summed in registers and each slot is read-modified-written once per launch, with the
// single-step kernel's store-first rule (dlo, dhi).  Only this tile writes the slots of its
// owned cells, so their old values are loaded with the stage, not after step 2.
#pragma once
#include "eio_kernels.h"

namespace eio {

constexpr int TZ2 = 56;       // owned z rows of a two-step tile
constexpr int Z1 = 64;        // rows of u1 (= lanes)
constexpr int ZC = 72;        // rows of u_cur staged in LDS

// Four consecutive floats from a buffer (raw_buffer_load_f32x4, eio_kernels.h).
__device__ __forceinline__ void bld4(rsrc_t r, unsigned voff, unsigned soff, float v[4]) {
    const f32x4_t q = raw_buffer_load_f32x4(r, int(voff), int(soff), 0);
    v[0] = q.x;
    v[1] = q.y;
    v[2] = q.z;
    v[3] = q.w;
}

// Forms the values of v at this point of the program: an empty asm statement that reads and
// writes them keeps the compiler from sinking the arithmetic that produces them past later
// barriers (which would keep every LDS value it reads live across the barrier).
template <typename T, int N>
__device__ __forceinline__ void pin_regs(T (&v)[N]) {
#pragma unroll
    for (int i = 0; i < N; ++i) asm volatile("" : "+v"(v[i]));
}

template <typename T>
__device__ __forceinline__ void bst_pol(rsrc_t r, T v, unsigned voff, unsigned soff, bool nt) {
    if (nt) bst_nt(r, v, voff, soff);
    else bst(r, v, voff, soff);
}

// NTS: store u1 and u2 with the nontemporal cache policy; DEC: also decimate Vx, Vz;
// FINE: the Born source pass (mode M_BSRC of k_step_fwd): sources from pre-fine-sampled traces
// (kernelsGpuElastic.cu:121-168), no receiver recording.  SS: the Born scattered pass (mode
// M_BSCAT): no point sources; every cell of both steps adds the secondary source of slots
// ss[sspar] (coarse its) and ss[sspar^1] (its+1) with the weights of its own fine step
// (kernelsGpuElastic.cu:1363-1368); both steps lie in one coarse interval, so they share the
// two slots.  NOREC: no receiver recording (the segment recompute of the FWI gradient).
// The two steps of one tile (every thread of the workgroup calls it with the same tile).
template <typename T, int R2, int WX, bool NTS, bool DEC, bool FINE, bool SS, bool NOREC, bool FS = false>
__device__ __forceinline__ void step2_fwd_tile(const ShotDev<T>* __restrict__ shots, const TileDev t,
                                               const Entry* __restrict__ ents, int k, int par, int dlo,
                                               int dhi, int sspar) {
    static_assert(8 % WX == 0, "WX must divide 8");
    static_assert(sizeof(T) == 4, "the two-step forward is FP32");
    static_assert(!FS || !SS, "the field-streamed stage has no scattered (SS) pass");
    constexpr int NT = 64 * WX;
    constexpr int TX = WX * R2;        // owned columns
    constexpr int R1 = R2 + 8 / WX;    // step-1 columns per wave
    constexpr int TX1 = TX + 8;        // step-1 columns
    constexpr int TXC = TX + 16;       // staged u_cur columns
    constexpr unsigned SZ = sizeof(T);
    // u_cur stage: all five fields, or (FS) two buffers the fields stream through
    __shared__ __attribute__((aligned(16))) T lc[FS ? 2 : 5][TXC][ZC];
    __shared__ T l1[5][TX1][Z1];

    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    const unsigned fbytes = n * SZ, colb = unsigned(nz) * SZ;
    const bool two = k + 1 < S.ntw - 1;   // the shot's last step may be a single one
    const StepTime st1 = fwd_time(k, S.sub);
    const StepTime st2 = fwd_time(k + 1, S.sub);
    const T* oldp = par ? S.C : S.A;
    const T* curp = par ? S.D : S.B;
    T* out1 = par ? S.A : S.C;
    T* out2 = par ? S.B : S.D;
    const rsrc_t rold = make_rsrc(oldp, 5 * fbytes);
    const rsrc_t rcur = make_rsrc(curp, 5 * fbytes);
    const rsrc_t rmed = make_rsrc(S.med, 5 * fbytes);
    const rsrc_t rw1 = make_rsrc(out1, 5 * fbytes);
    const rsrc_t rw2 = make_rsrc(out2, 5 * fbytes);

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 - FAT + lane;                       // row of this lane (both steps)
    const bool zin = iz >= FAT && iz < nz - FAT;            // update-region row
    const bool zown = lane >= FAT && lane < FAT + TZ2 && zin;
    const int j1 = w * R1;                                   // first step-1 column (rel x0-4)
    const int ix1 = t.x0 - FAT + j1;
    // The lane offset of a row above the grid would wrap; such lanes only produce values
    // that are masked, so their loads are pointed at row FAT instead.
    const int izl = iz < 0 ? 0 : iz;

    const T* cz = S.cz;
    const T* cx = S.cx;
    const int j2 = w * R2;                                   // first owned column of step 2 (rel x0)
    const int ix2 = t.x0 + j2;
    rsrc_t rwx0, rwx1, rwz0, rwz1;
    T f1a = T(0), f1b = T(0), f2a = T(0), f2b = T(0);
    T px0[R2], pz0[R2], px1[R2], pz1[R2];
    // decimation slots its (x0, z0) and its+1 (x1, z1) of the owned cells (DEC)
    auto load_slots = [&]() {
        const size_t s0 = size_t(st1.its % S.wcap) * n, s1 = size_t((st1.its + 1) % S.wcap) * n;
        rwx0 = make_rsrc(S.Wvx + s0, fbytes);
        rwx1 = make_rsrc(S.Wvx + s1, fbytes);
        rwz0 = make_rsrc(S.Wvz + s0, fbytes);
        rwz1 = make_rsrc(S.Wvz + s1, fbytes);
        f1a = S.filt[st1.it2];
        f1b = S.filt[S.h + st1.it2];
        f2a = S.filt[st2.it2];
        f2b = S.filt[S.h + st2.it2];
        const bool sta = st1.it2 == 1 && st1.its == dlo;
        const bool stb = st1.it2 == 1 && st1.its + 1 != dhi;
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            const unsigned vo = (unsigned(ix2 + r) * unsigned(nz) + unsigned(izl)) * SZ;
            px0[r] = sta ? T(0) : bld<T>(rwx0, vo, 0);
            pz0[r] = sta ? T(0) : bld<T>(rwz0, vo, 0);
            px1[r] = stb ? T(0) : bld<T>(rwx1, vo, 0);
            pz1[r] = stb ? T(0) : bld<T>(rwz1, vo, 0);
        }
    };
    // SS: secondary-source slots and the weights of the two steps (sampled at it2-1)
    rsrc_t rssl, rssr;
    T ssf1a = T(0), ssf1b = T(0), ssf2a = T(0), ssf2b = T(0);
    if constexpr (SS) {
        rssl = make_rsrc(S.ss[sspar], 5 * fbytes);
        rssr = make_rsrc(S.ss[sspar ^ 1], 5 * fbytes);
        ssf1a = S.filt[st1.it2 - 1];
        ssf1b = S.filt[S.h + st1.it2 - 1];
        ssf2a = S.filt[st2.it2 - 1];
        ssf2b = S.filt[S.h + st2.it2 - 1];
    }
    if constexpr (FS) {
        // ---- u_cur streamed through LDS one field at a time (two buffers of (TX+16) x 72
        // cells from (x0-8, z0-8), the same 16-byte row-quad loads as the five-field stage):
        // a field is staged while the loads of the next one are in flight, and each of its
        // derivatives at the step-1 cells goes straight into the sum the update takes
        // (S0 = dx sxx + dz sxz, S1 = dz szz + dx sxz, dx vx, dz vz, S4 = dz vx + dx vz; the
        // order sxx, szz, sxz, vx, vz keeps each sum's operand order).  LDS per tile:
        // 2 x (TX+16) x 72 + 5 x (TX+8) x 64 floats (78.8 KB at TX = 32), so two workgroups
        // share a CU.  The older level and media of the step-1 cells are loaded two phases
        // before the update that uses them (vx and vz of the older level first).
        constexpr int ZQ = ZC / 4;
        constexpr int NFILL = (TXC * ZQ + NT - 1) / NT;
        unsigned hv[NFILL];  // byte offset of each quad
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            const int c = e / ZQ, q = e - c * ZQ;
            const int gx = t.x0 - 2 * FAT + c, gz = t.z0 - 2 * FAT + 4 * q;
            const bool in = e < TXC * ZQ && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
            hv[it] = in ? (unsigned(gx) * unsigned(nz) + unsigned(gz)) * SZ : 0u;
        }
        auto load_field = [&](int f, T (&v)[NFILL][4]) {
#pragma unroll
            for (int it = 0; it < NFILL; ++it) bld4(rcur, hv[it], unsigned(f) * fbytes, v[it]);
        };
        auto stage_field = [&](int b, const T (&v)[NFILL][4]) {
#pragma unroll
            for (int it = 0; it < NFILL; ++it) {
                const int e = threadIdx.x + it * NT;
                if (e < TXC * ZQ) {
                    float4 x;
                    x.x = v[it][0];
                    x.y = v[it][1];
                    x.z = v[it][2];
                    x.w = v[it][3];
                    *reinterpret_cast<float4*>(&lc[b][0][4 * e]) = x;
                }
            }
        };
        T o[5][R1], m[5][R1];
        auto load_om = [&](int f) {
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                const int ix = ix1 + r;
                const unsigned vo = (unsigned(ix < 0 ? 0 : ix) * unsigned(nz) + unsigned(izl)) * SZ;
                o[f][r] = bld<T>(rold, vo, unsigned(f) * fbytes);
                m[f][r] = bld<T>(rmed, vo, unsigned(f) * fbytes);
            }
        };
        // x derivatives D+ / D- of buffer b at the step-1 columns (register window of R1+8)
        auto dxp = [&](int b, T (&out)[R1]) {
            T q[R1 + 8];
#pragma unroll
            for (int j = 0; j < R1 + 8; ++j) q[j] = lc[b][j1 + j][lane + FAT];
#pragma unroll
            for (int r = 0; r < R1; ++r) out[r] = dp_q(q, r + 4, cx);
        };
        auto dxm = [&](int b, T (&out)[R1]) {
            T q[R1 + 8];
#pragma unroll
            for (int j = 0; j < R1 + 8; ++j) q[j] = lc[b][j1 + j][lane + FAT];
#pragma unroll
            for (int r = 0; r < R1; ++r) out[r] = dm_q(q, r + 4, cx);
        };
        // the sums of a phase are formed before its closing barrier (pin_regs)
        T sa[NFILL][4], sb[NFILL][4];
        T s0[R1], s1[R1], dxvx[R1], dzvz[R1], s4[R1], dv[R1];
        load_field(2, sa);  // sxx
        load_field(3, sb);  // szz
        load_om(0);
        load_om(1);
        stage_field(0, sa);
        __syncthreads();
        load_field(4, sa);  // sxz
        stage_field(1, sb);
        load_field(0, sb);  // vx
        dxm(0, s0);         // sxx: D- along x
        pin_regs(s0);
        __syncthreads();
        stage_field(0, sa);
        load_field(1, sa);  // vz
#pragma unroll
        for (int r = 0; r < R1; ++r) s1[r] = dm_q(&lc[1][j1 + r + FAT][lane + FAT], 0, cz);  // szz: D- along z
        pin_regs(s1);
        __syncthreads();
        stage_field(1, sb);
        {  // sxz: D+ along x into S1, D+ along z into S0
            T tx[R1];
            dxp(0, tx);
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                s1[r] = s1[r] + tx[r];
                s0[r] = s0[r] + dp_q(&lc[0][j1 + r + FAT][lane + FAT], 0, cz);
            }
        }
        load_om(2);
#pragma unroll
        for (int r = 0; r < R1; ++r) {  // vx and vz of u1
            const int ix = ix1 + r;
            const bool ok = zin && ix >= FAT && ix < nx - FAT;
            const T d = ok ? damp_at(S, iz, ix) : T(1);
            dv[r] = d;
            const T a0 = d * o[0][r] + m[0][r] * s0[r];
            const T a1 = d * o[1][r] + m[1][r] * s1[r];
            l1[0][j1 + r][lane] = ok ? d * a0 : T(0);
            l1[1][j1 + r][lane] = ok ? d * a1 : T(0);
        }
        __syncthreads();
        stage_field(0, sa);
        load_om(3);
        load_om(4);
        dxp(1, dxvx);  // vx: D+ along x, D- along z
#pragma unroll
        for (int r = 0; r < R1; ++r) s4[r] = dm_q(&lc[1][j1 + r + FAT][lane + FAT], 0, cz);
        pin_regs(dxvx);
        pin_regs(s4);
        __syncthreads();
        {  // vz: D- along x into S4, D+ along z
            T tx[R1];
            dxm(0, tx);
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                s4[r] = s4[r] + tx[r];
                dzvz[r] = dp_q(&lc[0][j1 + r + FAT][lane + FAT], 0, cz);
            }
        }
#pragma unroll
        for (int r = 0; r < R1; ++r) {  // the stresses of u1
            const int ix = ix1 + r;
            const bool ok = zin && ix >= FAT && ix < nx - FAT;
            const T d = dv[r];
            const T a2 = d * o[2][r] + (m[2][r] * dxvx[r] + m[3][r] * dzvz[r]);
            const T a3 = d * o[3][r] + (m[3][r] * dxvx[r] + m[2][r] * dzvz[r]);
            const T a4 = d * o[4][r] + m[4][r] * s4[r];
            l1[2][j1 + r][lane] = ok ? d * a2 : T(0);
            l1[3][j1 + r][lane] = ok ? d * a3 : T(0);
            l1[4][j1 + r][lane] = ok ? d * a4 : T(0);
        }
    } else {
        // ---- u_cur on (TX+16) x 72 cells from (x0-8, z0-8) (cells off the grid are 0), four rows
        // per 16-byte load: all loads are issued first, then the older level and media of the
        // step-1 cells, and only then are the staged values written to LDS, so every load of the
        // tile is in flight at once.  Row quads start at z0-8+4q = 56 bz-4+4q, a multiple of 4:
        // a quad is wholly above row 0 or wholly below it, and only the bottom edge cuts one.
        constexpr int ZQ = ZC / 4;
        constexpr int NFILL = (TXC * ZQ + NT - 1) / NT;
        T hf[NFILL][5][4];
        unsigned hm[NFILL];  // bit k: row k of the quad is on the grid
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            const int c = e / ZQ, q = e - c * ZQ;
            const int gx = t.x0 - 2 * FAT + c, gz = t.z0 - 2 * FAT + 4 * q;
            const bool in = e < TXC * ZQ && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
            const int nin = in ? (nz - gz < 4 ? nz - gz : 4) : 0;
            hm[it] = (1u << nin) - 1u;
            const unsigned hv = in ? (unsigned(gx) * unsigned(nz) + unsigned(gz)) * SZ : 0u;
#pragma unroll
            for (int f = 0; f < 5; ++f) bld4(rcur, hv, unsigned(f) * fbytes, hf[it][f]);
        }
        T o[5][R1], m[5][R1];
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int ix = ix1 + r;
            const unsigned vo = (unsigned(ix < 0 ? 0 : ix) * unsigned(nz) + unsigned(izl)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) o[f][r] = bld<T>(rold, vo, unsigned(f) * fbytes);
#pragma unroll
            for (int f = 0; f < 5; ++f) m[f][r] = bld<T>(rmed, vo, unsigned(f) * fbytes);
        }
        // the decimation slots with the stage
        if constexpr (DEC) load_slots();
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
        __syncthreads();

        // ---- step 1 on columns j1 .. j1+R1-1 (rel x0-4) = lc columns j1+4 .. ; lane row = lc row lane+4
        {
            T dx_sxx[R1], dx_sxz[R1], dx_vx[R1], dx_vz[R1];
            {
                T q[R1 + 8];
#pragma unroll
                for (int j = 0; j < R1 + 8; ++j) q[j] = lc[2][j1 + j][lane + FAT];
#pragma unroll
                for (int r = 0; r < R1; ++r) dx_sxx[r] = dm_q(q, r + 4, cx);
#pragma unroll
                for (int j = 0; j < R1 + 8; ++j) q[j] = lc[4][j1 + j][lane + FAT];
#pragma unroll
                for (int r = 0; r < R1; ++r) dx_sxz[r] = dp_q(q, r + 4, cx);
#pragma unroll
                for (int j = 0; j < R1 + 8; ++j) q[j] = lc[0][j1 + j][lane + FAT];
#pragma unroll
                for (int r = 0; r < R1; ++r) dx_vx[r] = dp_q(q, r + 4, cx);
#pragma unroll
                for (int j = 0; j < R1 + 8; ++j) q[j] = lc[1][j1 + j][lane + FAT];
#pragma unroll
                for (int r = 0; r < R1; ++r) dx_vz[r] = dm_q(q, r + 4, cx);
            }
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                const int ix = ix1 + r;
                const bool ok = zin && ix >= FAT && ix < nx - FAT;
                const int c = j1 + r + FAT;
                const T dz_sxz = dp_q(&lc[4][c][lane + FAT], 0, cz);
                const T dz_szz = dm_q(&lc[3][c][lane + FAT], 0, cz);
                const T dz_vz = dp_q(&lc[1][c][lane + FAT], 0, cz);
                const T dz_vx = dm_q(&lc[0][c][lane + FAT], 0, cz);
                const T d = ok ? damp_at(S, iz, ix) : T(1);
                T a0 = d * o[0][r] + m[0][r] * (dx_sxx[r] + dz_sxz);
                T a1 = d * o[1][r] + m[1][r] * (dz_szz + dx_sxz[r]);
                T a2 = d * o[2][r] + (m[2][r] * dx_vx[r] + m[3][r] * dz_vz);
                T a3 = d * o[3][r] + (m[3][r] * dx_vx[r] + m[2][r] * dz_vz);
                T a4 = d * o[4][r] + m[4][r] * (dz_vx + dx_vz[r]);
                if constexpr (SS) {
                    const unsigned vo = (unsigned(ix < 0 ? 0 : ix) * unsigned(nz) + unsigned(izl)) * SZ;
                    T sl[5], sr[5];
#pragma unroll
                    for (int f = 0; f < 5; ++f) {
                        sl[f] = bld<T>(rssl, vo, unsigned(f) * fbytes);
                        sr[f] = bld<T>(rssr, vo, unsigned(f) * fbytes);
                    }
                    a0 += sl[0] * ssf1a + sr[0] * ssf1b;
                    a1 += sl[1] * ssf1a + sr[1] * ssf1b;
                    a2 += sl[2] * ssf1a + sr[2] * ssf1b;
                    a3 += sl[3] * ssf1a + sr[3] * ssf1b;
                    a4 += sl[4] * ssf1a + sr[4] * ssf1b;
                }
                l1[0][j1 + r][lane] = ok ? d * a0 : T(0);
                l1[1][j1 + r][lane] = ok ? d * a1 : T(0);
                l1[2][j1 + r][lane] = ok ? d * a2 : T(0);
                l1[3][j1 + r][lane] = ok ? d * a3 : T(0);
                l1[4][j1 + r][lane] = ok ? d * a4 : T(0);
            }
        }
    }
    __syncthreads();
    // FS: the decimation slots after step 1 (its registers are taken during the streamed
    // stage; step 1's sources and receivers and step 2 cover their latency)
    if constexpr (DEC && FS) load_slots();

    const bool has_src = !SS && t.src_e > t.src_b;
    const bool has_rec = !FINE && !NOREC && t.rec_e > t.rec_b;
    // ---- step-1 sources on the whole step-1 region (kernelsGpuElastic.cu:172-189)
    if (has_src) {
        for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NT) {
            const Entry en = ents[e];
            const int lz = en.lcell & 255, lx = en.lcell >> 8;
            const int comp = en.crow >> 24, row = en.crow & 0xffffff;
            const T* tr = S.src_tr[comp];
            T s;
            if constexpr (FINE) {
                s = tr[long(row) * S.ntw + k];
            } else {
                const long b = long(row) * S.nts + st1.its;
                s = tr[b] * S.filt[st1.it2 - 1] + tr[b + 1] * S.filt[S.h + st1.it2 - 1];
            }
            l1[comp][lx][lz] += damp_at(S, t.z0 - FAT + lz, t.x0 - FAT + lx) * s;
        }
        __syncthreads();
    }
    // ---- step-1 receivers (kernelsGpuElastic.cu:7-60)
    if (has_rec) {
        const T fa = S.filt[st1.it2], fb = S.filt[S.h + st1.it2];
        for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NT) {
            const Entry en = ents[e];
            const int lz = en.lcell & 255, lx = en.lcell >> 8;
            const int comp = en.crow >> 24, row = en.crow & 0xffffff;
            const T v = l1[comp][lx][lz];
            T* tr = S.rec_tr[comp] + long(row) * S.nts + st1.its;
            tr[0] += v * fa;
            tr[1] += v * fb;
        }
    }
    // ---- store u1 on owned cells: step-1 columns inside [x0, x0+TX)
#pragma unroll
    for (int r = 0; r < R1; ++r) {
        const int j = j1 + r, ix = ix1 + r;
        if (zown && j >= FAT && j < FAT + TX && ix < nx - FAT) {
            const unsigned vo = (unsigned(ix) * unsigned(nz) + unsigned(iz)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) bst_pol<T>(rw1, l1[f][j][lane], vo, unsigned(f) * fbytes, NTS);
        }
    }
    if (!two) return;

    // ---- step 2 on owned columns j2 .. j2+R2-1 (rel x0) = l1 columns j2+4 ..
    const int lz2 = lane < FAT ? FAT : (lane > FAT + TZ2 - 1 ? FAT + TZ2 - 1 : lane);  // in-range LDS rows
    T v2[5][R2];
    T u1x[R2], u1z[R2];  // u1's Vx, Vz on the owned cells (decimation)
    if constexpr (DEC) {
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            u1x[r] = l1[0][j2 + r + FAT][lz2];
            u1z[r] = l1[1][j2 + r + FAT][lz2];
        }
    }
    {
        T dx_sxx[R2], dx_sxz[R2], dx_vx[R2], dx_vz[R2];
        {
            T q[R2 + 8];
#pragma unroll
            for (int j = 0; j < R2 + 8; ++j) q[j] = l1[2][j2 + j][lz2];
#pragma unroll
            for (int r = 0; r < R2; ++r) dx_sxx[r] = dm_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R2 + 8; ++j) q[j] = l1[4][j2 + j][lz2];
#pragma unroll
            for (int r = 0; r < R2; ++r) dx_sxz[r] = dp_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R2 + 8; ++j) q[j] = l1[0][j2 + j][lz2];
#pragma unroll
            for (int r = 0; r < R2; ++r) dx_vx[r] = dp_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R2 + 8; ++j) q[j] = l1[1][j2 + j][lz2];
#pragma unroll
            for (int r = 0; r < R2; ++r) dx_vz[r] = dm_q(q, r + 4, cx);
        }
        if constexpr (FS) {
            pin_regs(dx_sxx);
            pin_regs(dx_sxz);
            pin_regs(dx_vx);
            pin_regs(dx_vz);
        }
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            const int ix = ix2 + r;
            const bool ok = zown && ix < nx - FAT;
            const int c = j2 + r + FAT;
            const unsigned vo = (unsigned(ix) * unsigned(nz) + unsigned(izl)) * SZ;
            T mm[5];
#pragma unroll
            for (int f = 0; f < 5; ++f) mm[f] = bld<T>(rmed, vo, unsigned(f) * fbytes);
            // older level of step 2 = u_cur at this cell (lc column j2+r+8, row lane+4)
            T oo[5];
            if constexpr (FS) {  // re-read (the stage held one field at a time; an L2 hit)
#pragma unroll
                for (int f = 0; f < 5; ++f) oo[f] = bld<T>(rcur, vo, unsigned(f) * fbytes);
            } else {
#pragma unroll
                for (int f = 0; f < 5; ++f) oo[f] = lc[f][c + FAT][lz2 + FAT];
            }
            const T dz_sxz = dp_q(&l1[4][c][lz2], 0, cz);
            const T dz_szz = dm_q(&l1[3][c][lz2], 0, cz);
            const T dz_vz = dp_q(&l1[1][c][lz2], 0, cz);
            const T dz_vx = dm_q(&l1[0][c][lz2], 0, cz);
            const T d = ok ? damp_at(S, iz, ix) : T(1);
            T a0 = d * oo[0] + mm[0] * (dx_sxx[r] + dz_sxz);
            T a1 = d * oo[1] + mm[1] * (dz_szz + dx_sxz[r]);
            T a2 = d * oo[2] + (mm[2] * dx_vx[r] + mm[3] * dz_vz);
            T a3 = d * oo[3] + (mm[3] * dx_vx[r] + mm[2] * dz_vz);
            T a4 = d * oo[4] + mm[4] * (dz_vx + dx_vz[r]);
            if constexpr (SS) {
                T sl[5], sr[5];
#pragma unroll
                for (int f = 0; f < 5; ++f) {
                    sl[f] = bld<T>(rssl, vo, unsigned(f) * fbytes);
                    sr[f] = bld<T>(rssr, vo, unsigned(f) * fbytes);
                }
                a0 += sl[0] * ssf2a + sr[0] * ssf2b;
                a1 += sl[1] * ssf2a + sr[1] * ssf2b;
                a2 += sl[2] * ssf2a + sr[2] * ssf2b;
                a3 += sl[3] * ssf2a + sr[3] * ssf2b;
                a4 += sl[4] * ssf2a + sr[4] * ssf2b;
            }
            v2[0][r] = d * a0;
            v2[1][r] = d * a1;
            v2[2][r] = d * a2;
            v2[3][r] = d * a3;
            v2[4][r] = d * a4;
            // FS (a 128-VGPR budget): one column's loads and LDS reads at a time
            if constexpr (FS) {
#pragma unroll
                for (int f = 0; f < 5; ++f) asm volatile("" : "+v"(v2[f][r]));
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    if (has_src || has_rec) {  // stage u2 in the l1 area (uniform per tile)
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R2; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f)
                if (zown) l1[f][j2 + r + FAT][lane] = v2[f][r];
        __syncthreads();
        if (has_src) {  // owned sources only
            for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NT) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                if (lz < FAT || lz >= FAT + TZ2 || lx < FAT || lx >= FAT + TX) continue;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T* tr = S.src_tr[comp];
                T s;
                if constexpr (FINE) {
                    s = tr[long(row) * S.ntw + k + 1];
                } else {
                    const long b = long(row) * S.nts + st2.its;
                    s = tr[b] * S.filt[st2.it2 - 1] + tr[b + 1] * S.filt[S.h + st2.it2 - 1];
                }
                l1[comp][lx][lz] += damp_at(S, t.z0 - FAT + lz, t.x0 - FAT + lx) * s;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R2; ++r)
#pragma unroll
                for (int f = 0; f < 5; ++f)
                    if (zown) v2[f][r] = l1[f][j2 + r + FAT][lane];
        }
        if (has_rec) {
            const T fa = S.filt[st2.it2], fb = S.filt[S.h + st2.it2];
            for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NT) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T v = l1[comp][lx][lz];
                T* tr = S.rec_tr[comp] + long(row) * S.nts + st2.its;
                tr[0] += v * fa;
                tr[1] += v * fb;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int ix = ix2 + r;
        if (zown && ix < nx - FAT) {
            const unsigned vo = (unsigned(ix) * unsigned(nz) + unsigned(iz)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) bst_pol<T>(rw2, v2[f][r], vo, unsigned(f) * fbytes, NTS);
            if constexpr (DEC) {
                const T xa = px0[r] + (u1x[r] * f1a + v2[0][r] * f2a);
                const T za = pz0[r] + (u1z[r] * f1a + v2[1][r] * f2a);
                const T xb = px1[r] + (u1x[r] * f1b + v2[0][r] * f2b);
                const T zb = pz1[r] + (u1z[r] * f1b + v2[1][r] * f2b);
                bst(rwx0, xa, vo, 0u);
                bst(rwz0, za, vo, 0u);
                bst(rwx1, xb, vo, 0u);
                bst(rwz1, zb, vo, 0u);
            }
        }
    }
}

template <typename T, int R2, int WX, int OCC, int NTS, bool DEC, bool FINE = false, bool SS = false,
          bool NOREC = false, bool FS = false>
__global__ __launch_bounds__(64 * WX) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_step2_fwd(const ShotDev<T>* __restrict__ shots, const TileDev* __restrict__ tiles,
                 const Entry* __restrict__ ents, int ntiles, int k, int par, int dlo, int dhi,
                 int sspar) {
    step2_fwd_tile<T, R2, WX, NTS != 0, DEC, FINE, SS, NOREC, FS>(shots, tiles[xcd_remap(blockIdx.x, ntiles)], ents,
                                                                   k, par, dlo, dhi, sspar);
}

// ------------------------------------------------------------------ persistent two-step forward
// The whole NL forward time loop of a batch in ONE launch (for batches whose launches are too
// short to hide the launch gap).  The grid is co-resident (cooperative launch); workgroup b owns
// the tiles xcd_remap(b + i*G) (G a multiple of 8, so a workgroup's tiles lie in the tile range
// of its own XCD) and runs them pair after pair.  Instead of a grid barrier between pairs, each
// tile waits only for its (up to 8) neighbours: before pair p it needs done[nb] >= p for every
// neighbour nb, i.e. the neighbours have written pair p-1's levels that this tile's stage reads
// (8-cell halo: only adjacent tiles) and have finished reading the levels this tile overwrites.
// A tile publishes done[t] = p+1 after its stores: every wave waits for its stores (vmcnt 0),
// the workgroup barrier, then one thread writes the L2 back (agent-scope release) and sets the
// flag.  A waiting thread polls with relaxed agent-scope loads and takes an agent-scope acquire
// (L1/L2 invalidate) before the barrier that releases the workgroup.  Deadlock-free because all
// workgroups are resident and each processes its items in (pair, tile) order; every wait is
// bounded by `spin_limit` polls, after which the workgroup sets *abort and returns, and every
// other workgroup returns at its next wait, so the grid always drains.
// the quad-lane tile of eio_step2q.h
template <typename T, bool NOREC>
__device__ __forceinline__ void step2q_fwd_tile(const ShotDev<T>* __restrict__ shots, const TileDev t,
                                                const Entry* __restrict__ ents, int k, int par);

template <typename T, int R2, int WX, int OCC, bool FS = false, bool QUAD = false>
__global__ __launch_bounds__(64 * WX) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_step2_fwd_persist(const ShotDev<T>* __restrict__ shots, const TileDev* __restrict__ tiles,
                         const Entry* __restrict__ ents, int ntiles, const int* __restrict__ nbr,
                         int* done, int* abort_flag, int npair, int spin_limit) {
    __shared__ int s_stop;
    const int G = int(gridDim.x);
    for (int p = 0; p < npair; ++p) {
        for (int v = int(blockIdx.x); v < ntiles; v += G) {
            const int ti = xcd_remap(v, ntiles);
            const TileDev t = tiles[ti];
            if (2 * p >= shots[t.shot].ntw - 1) continue;   // the tile's shot has finished (uniform)
            if (p > 0) {
                if (threadIdx.x < 8) {
                    const int nb = nbr[ti * 8 + threadIdx.x];
                    int stop = 0;
                    if (nb >= 0) {
                        int spins = 0;
                        while (__hip_atomic_load(&done[nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < p) {
                            if (++spins > spin_limit ||
                                __hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                                __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                stop = 1;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(1);
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                    // lanes 0..7 of wave 0: any lane that stopped stops the workgroup
                    const bool any = __any(stop);
                    if (threadIdx.x == 0) s_stop = any ? 1 : 0;
                }
                __syncthreads();
                if (s_stop) return;
            }
            if constexpr (QUAD) step2q_fwd_tile<T, false>(shots, t, ents, 2 * p, p & 1);
            else step2_fwd_tile<T, R2, WX, false, false, false, false, false, FS>(shots, t, ents, 2 * p, p & 1, 0, -1, 0);
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores have reached L2
            __syncthreads();
            if (threadIdx.x == 0) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
                __hip_atomic_store(&done[ti], p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// ------------------------------------------------------------------ two-step adjoint
// Two reverse fine steps per launch for the Born receiver pass and the FWI gradient (the
// arithmetic of k_step_adj in mode M_BADJ).  The adjoint stencil takes derivatives of the
// media-scaled level, so step 1 stages M u_cur on (TX+16) x 72 cells from (x0-8, z0-8)
// (kernelsGpuElastic.cu:380-529), computes u1 on the (TX+8) x 64 step-1 region, and, once
// the data of step 1 are injected, writes M u1 over that stage for step 2 on the owned
// 56 x TX cells.  u_old, u_cur and the media are read once and u1, u2 written once: 100 B per
// two cell-updates, as for the forward.
//
// Both reverse steps lie in one coarse interval (pairs start at even reverse steps and the
// fine sub-sampling of the Born operator is even), so their adjoint-field contributions
// (kernelsGpuElastic.cu:1369-1376) are summed in registers: slot its (ss[sspar]) is written
// by the interval's first pair and read-modified-written by later ones, slot its+1
// (ss[sspar^1]) read-modified-written once per pair; only this tile writes the slots of its
// owned cells, so their old values are loaded with the stage.  Levels A,B -> C,D and back
// (par), as for the forward.  LATE: the older level is loaded after the stage is written, the step-1
// media after step 1 and u_cur of the owned cells after M u1 is formed (each set of values
// is live only in the phase that uses it: fewer registers, more latency to hide).
// PS: load the slots with the stage (else after step 2: fewer registers for wide tiles).
// MODE M_NLADJ (the NL adjoint, kernelsGpuElastic.cu:100-116): no adjoint-field slots; the
// source traces are extracted from u1 and u2 at the tile's owned source cells (its source
// entries are owned-only: each source cell is read once per step).
template <typename T, int R2, int WX, int OCC, bool LATE = false, bool PS = true, int MODE = M_BADJ>
__global__ __launch_bounds__(64 * WX) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_step2_adj(const ShotDev<T>* __restrict__ shots, const TileDev* __restrict__ tiles,
                 const Entry* __restrict__ ents, int ntiles, int k, int par, int sspar) {
    static_assert(8 % WX == 0, "WX must divide 8");
    static_assert(sizeof(T) == 4, "the two-step adjoint is FP32");
    constexpr int NT = 64 * WX;
    constexpr int TX = WX * R2;        // owned columns
    constexpr int R1 = R2 + 8 / WX;    // step-1 columns per wave
    constexpr int TX1 = TX + 8;        // step-1 columns
    constexpr int TXC = TX + 16;       // staged columns
    constexpr unsigned SZ = sizeof(T);
    // M u_cur, then M u1 on the step-1 region.  Fields 0, 1, 4 (lc[0], lc[1], lc[2]) take x
    // and z derivatives and keep the whole stage; field 2 takes x derivatives only, so lcx
    // keeps the 64 rows of the step-1 region (lcx[c][r] = stage row r+4); field 3 takes z
    // derivatives only, so lcz keeps the TX+8 columns of the step-1 region (lcz[c][r] = stage
    // column c+4): 3 KB less LDS (51,968 B per 56 x 8 workgroup).
    __shared__ __attribute__((aligned(16))) T lc[3][TXC][ZC];
    __shared__ __attribute__((aligned(16))) T lcx[TXC][Z1];
    __shared__ __attribute__((aligned(16))) T lcz[TX1][ZC];
    __shared__ T l1[5][TX1][Z1];       // u1, then u2 for the owned data injection

    const TileDev t = tiles[xcd_remap(blockIdx.x, ntiles)];
    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    const unsigned fbytes = n * SZ;
    const StepTime st1 = adj_time(k, S.sub, S.nts);
    const StepTime st2 = adj_time(k + 1, S.sub, S.nts);   // same interval, it2 one lower
    const T* oldp = par ? S.C : S.A;
    const T* curp = par ? S.D : S.B;
    T* out1 = par ? S.A : S.C;
    T* out2 = par ? S.B : S.D;
    const rsrc_t rold = make_rsrc(oldp, 5 * fbytes);
    const rsrc_t rcur = make_rsrc(curp, 5 * fbytes);
    const rsrc_t rmed = make_rsrc(S.med, 5 * fbytes);
    const rsrc_t rw1 = make_rsrc(out1, 5 * fbytes);
    const rsrc_t rw2 = make_rsrc(out2, 5 * fbytes);

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = t.z0 - FAT + lane;                       // row of this lane (both steps)
    const bool zin = iz >= FAT && iz < nz - FAT;            // update-region row
    const bool zown = lane >= FAT && lane < FAT + TZ2 && zin;
    const int j1 = w * R1;                                   // first step-1 column (rel x0-4)
    const int ix1 = t.x0 - FAT + j1;
    const int izl = iz < 0 ? 0 : iz;                         // masked lanes load row 0
    const int j2 = w * R2;                                   // first owned column (rel x0)
    const int ix2 = t.x0 + j2;
    const int lz2 = lane < FAT ? FAT : (lane > FAT + TZ2 - 1 ? FAT + TZ2 - 1 : lane);

    // ---- stage M u_cur (cells off the grid are 0); the older level and media of the step-1
    // cells and u_cur of the owned cells are loaded behind the stage loads (or, LATE, after
    // the stage is written)
    T o[5][R1], m1[5][R1], oc[5][R2];
    // adjoint-field slots its (p0; not read by the interval's first pair) and its+1 (p1)
    const rsrc_t rsl = make_rsrc(S.ss[sspar], 5 * fbytes);
    const rsrc_t rsr = make_rsrc(S.ss[sspar ^ 1], 5 * fbytes);
    const bool first = st1.it2 == S.sub - 1;
    T p0[5][R2], p1[5][R2];
    auto load_slots = [&]() {
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            const unsigned vo = (unsigned(ix2 + r) * unsigned(nz) + unsigned(izl)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                p0[f][r] = first ? T(0) : bld<T>(rsl, vo, unsigned(f) * fbytes);
                p1[f][r] = bld<T>(rsr, vo, unsigned(f) * fbytes);
            }
        }
    };
    auto load_old = [&]() {
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int ix = ix1 + r;
            const unsigned vo = (unsigned(ix < 0 ? 0 : ix) * unsigned(nz) + unsigned(izl)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) o[f][r] = bld<T>(rold, vo, unsigned(f) * fbytes);
        }
    };
    auto load_med = [&]() {
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int ix = ix1 + r;
            const unsigned vo = (unsigned(ix < 0 ? 0 : ix) * unsigned(nz) + unsigned(izl)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) m1[f][r] = bld<T>(rmed, vo, unsigned(f) * fbytes);
        }
    };
    auto load_cur = [&]() {
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            const unsigned vo = (unsigned(ix2 + r) * unsigned(nz) + unsigned(izl)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) oc[f][r] = bld<T>(rcur, vo, unsigned(f) * fbytes);
        }
    };
    {
        // u_cur and media four rows per 16-byte load (row quads start at z0-8+4q, a multiple
        // of 4: wholly above row 0 or not, and only the bottom edge cuts one)
        constexpr int ZQ = ZC / 4;
        constexpr int NQ = (TXC * ZQ + NT - 1) / NT;
        T hv_[NQ][5][4], hm[NQ][5][4];
        unsigned hk[NQ];  // bit k: row k of the quad is on the grid
#pragma unroll
        for (int it = 0; it < NQ; ++it) {
            const int e = threadIdx.x + it * NT;
            const int c = e / ZQ, q = e - c * ZQ;
            const int gx = t.x0 - 2 * FAT + c, gz = t.z0 - 2 * FAT + 4 * q;
            const bool in = e < TXC * ZQ && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
            const int nin = in ? (nz - gz < 4 ? nz - gz : 4) : 0;
            hk[it] = (1u << nin) - 1u;
            const unsigned hv = in ? (unsigned(gx) * unsigned(nz) + unsigned(gz)) * SZ : 0u;
#pragma unroll
            for (int f = 0; f < 5; ++f) bld4(rcur, hv, unsigned(f) * fbytes, hv_[it][f]);
#pragma unroll
            for (int f = 0; f < 5; ++f) bld4(rmed, hv, unsigned(f) * fbytes, hm[it][f]);
        }
        if constexpr (!LATE) {
            load_old();
            load_med();
            load_cur();
            if constexpr (PS && MODE == M_BADJ) load_slots();
        }
#pragma unroll
        for (int it = 0; it < NQ; ++it) {
            const int e = threadIdx.x + it * NT;
            if (e < TXC * ZQ) {
                const int c = e / ZQ, q = e - c * ZQ;
                f32x4_t p[5];
#pragma unroll
                for (int k2 = 0; k2 < 4; ++k2) {
                    const bool in = (hk[it] >> k2) & 1u;
                    const T u0 = in ? hv_[it][0][k2] : T(0), u1_ = in ? hv_[it][1][k2] : T(0);
                    const T u2 = in ? hv_[it][2][k2] : T(0), u3 = in ? hv_[it][3][k2] : T(0);
                    const T u4 = in ? hv_[it][4][k2] : T(0);
                    p[0][k2] = hm[it][0][k2] * u0;
                    p[1][k2] = hm[it][1][k2] * u1_;
                    p[2][k2] = hm[it][2][k2] * u2 + hm[it][3][k2] * u3;
                    p[3][k2] = hm[it][3][k2] * u2 + hm[it][2][k2] * u3;
                    p[4][k2] = hm[it][4][k2] * u4;
                }
                *reinterpret_cast<f32x4_t*>(&lc[0][c][4 * q]) = p[0];
                *reinterpret_cast<f32x4_t*>(&lc[1][c][4 * q]) = p[1];
                *reinterpret_cast<f32x4_t*>(&lc[2][c][4 * q]) = p[4];
                if (q >= 1 && q <= Z1 / 4) *reinterpret_cast<f32x4_t*>(&lcx[c][4 * q - FAT]) = p[2];
                if (c >= FAT && c < FAT + TX1) *reinterpret_cast<f32x4_t*>(&lcz[c - FAT][4 * q]) = p[3];
            }
        }
    }
    if constexpr (LATE) {
        load_old();
        if constexpr (PS && MODE == M_BADJ) load_slots();
    }
    __syncthreads();

    const T* cz = S.cz;
    const T* cx = S.cx;
    // ---- step 1 on columns j1 .. j1+R1-1 (rel x0-4) = lc columns j1+4 ..; lane row = lc row lane+4
    {
        T dx_a[R1], dx_s6[R1], dx_s0[R1], dx_s1[R1];
        {
            T q[R1 + 8];
#pragma unroll
            for (int j = 0; j < R1 + 8; ++j) q[j] = lcx[j1 + j][lane];
#pragma unroll
            for (int r = 0; r < R1; ++r) dx_a[r] = dm_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R1 + 8; ++j) q[j] = lc[2][j1 + j][lane + FAT];
#pragma unroll
            for (int r = 0; r < R1; ++r) dx_s6[r] = dp_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R1 + 8; ++j) q[j] = lc[0][j1 + j][lane + FAT];
#pragma unroll
            for (int r = 0; r < R1; ++r) dx_s0[r] = dp_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R1 + 8; ++j) q[j] = lc[1][j1 + j][lane + FAT];
#pragma unroll
            for (int r = 0; r < R1; ++r) dx_s1[r] = dm_q(q, r + 4, cx);
        }
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int ix = ix1 + r;
            const bool ok = zin && ix >= FAT && ix < nx - FAT;
            const int c = j1 + r + FAT;
            const T dz_s6 = dp_q(&lc[2][c][lane + FAT], 0, cz);
            const T dz_b = dm_q(&lcz[c - FAT][lane + FAT], 0, cz);
            const T dz_s1 = dp_q(&lc[1][c][lane + FAT], 0, cz);
            const T dz_s0 = dm_q(&lc[0][c][lane + FAT], 0, cz);
            const T d = ok ? damp_at(S, iz, ix) : T(1);
            const T v0 = d * (d * o[0][r] - (dx_a[r] + dz_s6));
            const T v1 = d * (d * o[1][r] - (dz_b + dx_s6[r]));
            const T v2 = d * (d * o[2][r] - dx_s0[r]);
            const T v3 = d * (d * o[3][r] - dz_s1);
            const T v4 = d * (d * o[4][r] - (dz_s0 + dx_s1[r]));
            l1[0][j1 + r][lane] = ok ? v0 : T(0);
            l1[1][j1 + r][lane] = ok ? v1 : T(0);
            l1[2][j1 + r][lane] = ok ? v2 : T(0);
            l1[3][j1 + r][lane] = ok ? v3 : T(0);
            l1[4][j1 + r][lane] = ok ? v4 : T(0);
        }
    }
    if constexpr (LATE) load_med();
    __syncthreads();

    const bool has_inj = t.rec_e > t.rec_b;
    const bool has_ext = MODE == M_NLADJ && t.src_e > t.src_b;
    // source-trace extraction of one step from the l1 area (owned source cells)
    auto extract = [&](const StepTime& st) {
        const T fa = S.filt[st.it2], fb = S.filt[S.h + st.it2];
        for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NT) {
            const Entry en = ents[e];
            const int lz = en.lcell & 255, lx = en.lcell >> 8;
            const int comp = en.crow >> 24, row = en.crow & 0xffffff;
            const T v = l1[comp][lx][lz];
            T* tr = S.src_tr[comp] + long(row) * S.nts + st.its;
            tr[0] += v * fa;
            tr[1] += v * fb;
        }
    };
    // ---- step-1 data injection on the whole step-1 region (kernelsGpuElastic.cu:191-229, it2+1)
    if (has_inj) {
        const T fa = S.filt[st1.it2 + 1], fb = S.filt[S.h + st1.it2 + 1];
        for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NT) {
            const Entry en = ents[e];
            const int lz = en.lcell & 255, lx = en.lcell >> 8;
            const int comp = en.crow >> 24, row = en.crow & 0xffffff;
            const T* tr = S.rec_tr[comp] + long(row) * S.nts + st1.its;
            const T s = tr[0] * fa + tr[1] * fb;
            l1[comp][lx][lz] += damp_at(S, t.z0 - FAT + lz, t.x0 - FAT + lx) * s;
        }
        __syncthreads();
    }
    if (has_ext) extract(st1);  // l1 is read only until the next barrier
    // ---- M u1 over the stage (same cells: lc column j+4, row lane+4), u1 stored on owned cells
#pragma unroll
    for (int r = 0; r < R1; ++r) {
        const int j = j1 + r, ix = ix1 + r;
        T u[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) u[f] = l1[f][j][lane];
        lc[0][j + FAT][lane + FAT] = m1[0][r] * u[0];
        lc[1][j + FAT][lane + FAT] = m1[1][r] * u[1];
        lcx[j + FAT][lane] = m1[2][r] * u[2] + m1[3][r] * u[3];
        lcz[j][lane + FAT] = m1[3][r] * u[2] + m1[2][r] * u[3];
        lc[2][j + FAT][lane + FAT] = m1[4][r] * u[4];
        if (zown && j >= FAT && j < FAT + TX && ix < nx - FAT) {
            const unsigned vo = (unsigned(ix) * unsigned(nz) + unsigned(iz)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) bst(rw1, u[f], vo, unsigned(f) * fbytes);
        }
    }
    // u1 on the owned cells of this lane's step-2 columns (adjoint-field extraction)
    T u1[5][R2];
    if constexpr (MODE == M_BADJ) {
#pragma unroll
        for (int r = 0; r < R2; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f) u1[f][r] = l1[f][j2 + r + FAT][lz2];
    }
    if constexpr (LATE) load_cur();
    __syncthreads();

    // ---- step 2 on owned columns j2 .. j2+R2-1 (rel x0) = lc columns j2+8 ..; rows lz2+4
    T v2[5][R2];
    {
        T dx_a[R2], dx_s6[R2], dx_s0[R2], dx_s1[R2];
        {
            T q[R2 + 8];
#pragma unroll
            for (int j = 0; j < R2 + 8; ++j) q[j] = lcx[j2 + FAT + j][lz2];
#pragma unroll
            for (int r = 0; r < R2; ++r) dx_a[r] = dm_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R2 + 8; ++j) q[j] = lc[2][j2 + FAT + j][lz2 + FAT];
#pragma unroll
            for (int r = 0; r < R2; ++r) dx_s6[r] = dp_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R2 + 8; ++j) q[j] = lc[0][j2 + FAT + j][lz2 + FAT];
#pragma unroll
            for (int r = 0; r < R2; ++r) dx_s0[r] = dp_q(q, r + 4, cx);
#pragma unroll
            for (int j = 0; j < R2 + 8; ++j) q[j] = lc[1][j2 + FAT + j][lz2 + FAT];
#pragma unroll
            for (int r = 0; r < R2; ++r) dx_s1[r] = dm_q(q, r + 4, cx);
        }
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            const int ix = ix2 + r;
            const bool ok = zown && ix < nx - FAT;
            const int c = j2 + r + 2 * FAT;
            const T dz_s6 = dp_q(&lc[2][c][lz2 + FAT], 0, cz);
            const T dz_b = dm_q(&lcz[c - FAT][lz2 + FAT], 0, cz);
            const T dz_s1 = dp_q(&lc[1][c][lz2 + FAT], 0, cz);
            const T dz_s0 = dm_q(&lc[0][c][lz2 + FAT], 0, cz);
            const T d = ok ? damp_at(S, iz, ix) : T(1);
            v2[0][r] = d * (d * oc[0][r] - (dx_a[r] + dz_s6));
            v2[1][r] = d * (d * oc[1][r] - (dz_b + dx_s6[r]));
            v2[2][r] = d * (d * oc[2][r] - dx_s0[r]);
            v2[3][r] = d * (d * oc[3][r] - dz_s1);
            v2[4][r] = d * (d * oc[4][r] - (dz_s0 + dx_s1[r]));
        }
    }
    if (has_inj || has_ext) {  // step-2 data through the l1 area (uniform per tile), owned cells
#pragma unroll
        for (int r = 0; r < R2; ++r)
#pragma unroll
            for (int f = 0; f < 5; ++f)
                if (zown) l1[f][j2 + r + FAT][lane] = v2[f][r];
        __syncthreads();
        if (has_inj) {
            const T fa = S.filt[st2.it2 + 1], fb = S.filt[S.h + st2.it2 + 1];
            for (int e = t.rec_b + threadIdx.x; e < t.rec_e; e += NT) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                if (lz < FAT || lz >= FAT + TZ2 || lx < FAT || lx >= FAT + TX) continue;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T* tr = S.rec_tr[comp] + long(row) * S.nts + st2.its;
                const T s = tr[0] * fa + tr[1] * fb;
                l1[comp][lx][lz] += damp_at(S, t.z0 - FAT + lz, t.x0 - FAT + lx) * s;
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R2; ++r)
#pragma unroll
                for (int f = 0; f < 5; ++f)
                    if (zown) v2[f][r] = l1[f][j2 + r + FAT][lane];
        }
        if (has_ext) extract(st2);
    }
    // ---- u2 and the pair's adjoint fields: slot its gets F[a] u1 + F[a-1] u2 and slot its+1
    // F[h+a] u1 + F[h+a-1] u2 (a = st1.it2)
    const T ea1 = S.filt[st1.it2], ea2 = S.filt[st2.it2];
    const T eb1 = S.filt[S.h + st1.it2], eb2 = S.filt[S.h + st2.it2];
    if constexpr (!PS && MODE == M_BADJ) load_slots();
#pragma unroll
    for (int r = 0; r < R2; ++r) {
        const int ix = ix2 + r;
        if (zown && ix < nx - FAT) {
            const unsigned vo = (unsigned(ix) * unsigned(nz) + unsigned(iz)) * SZ;
#pragma unroll
            for (int f = 0; f < 5; ++f) {
                const unsigned sf = unsigned(f) * fbytes;
                bst(rw2, v2[f][r], vo, sf);
                if constexpr (MODE == M_BADJ) {
                    bst(rsl, p0[f][r] + (u1[f][r] * ea1 + v2[f][r] * ea2), vo, sf);
                    bst(rsr, p1[f][r] + (u1[f][r] * eb1 + v2[f][r] * eb2), vo, sf);
                }
            }
        }
    }
}

}  // namespace eio
