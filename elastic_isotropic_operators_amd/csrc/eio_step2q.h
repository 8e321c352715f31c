// eio_step2q.h -- the two-step NL forward with 16-byte lane accesses.
//
// Same arithmetic, levels, entries and tile contract as k_step2_fwd (two fine steps per
// launch, owned tile 56 rows x QTX columns, step 1 recomputed on the 4-cell halo, sources of
// the step-1 region injected into every copy of u1, owned receivers), with another mapping of
// cells to lanes: a lane owns a QUAD of 4 consecutive rows of one column instead of one row of
// several columns.  Tile row z0-4 is a multiple of 4 (z0 = 4 + 56 bz), so the 16 quads of a
// column slot cover the tile's 64 rows and every older-level, media and state access of a lane
// is one 16-byte load or store, and every z-stencil read from LDS a 16-byte row read.  Lane
// accesses of 4 bytes stream about 3.8 TB/s on the MI355X, 16-byte ones about 6.4 TB/s; the
// row-per-lane kernel's traffic runs at the 4-byte rate.
//
// Workgroup: 8 waves = 32 column slots x 16 row quads; thread (s, q) = (tid / 16, tid % 16).
// Step 1 runs on the 32 columns [x0-4, x0+QTX+4) (slot s = column x0-4+s) x 64 rows; step 2 on
// the QTX owned columns (slots 0..QTX-1) x quads 1..14 (the owned rows).  u_cur streams
// through two LDS field buffers of (QTX+16) x 72 cells as in the streamed stage of
// k_step2_fwd; u1 is kept on the step-1 region in LDS (5 x 32 x 64 floats).  LDS 63.5 KB.
//
// Cells off the grid or outside the update region are read (16-byte accesses may run past a
// column's last row into the next column, or past the buffer end, which reads 0) but only
// ever feed lanes whose results are discarded: an update cell at row iz reads rows iz-4..iz+4
// of the update region and frame, never past row nz-1.  Stores write whole quads when all
// four rows are update cells of the tile, else row by row.
#pragma once
#include "eio_step2.h"

namespace eio {

constexpr int QTX = 24;    // owned columns of a quad tile
constexpr int QNT = 512;   // threads: 32 column slots x 16 row quads

__device__ void raw_buffer_store_f32x4(f32x4_t v, rsrc_t r, int voff, int soff, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.store.v4f32");

__device__ __forceinline__ f32x4_t qld(rsrc_t r, unsigned voff, unsigned soff) {
    return raw_buffer_load_f32x4(r, int(voff), int(soff), 0);
}
// the quad's 4 rows where bit j of `rows` is set: one 16-byte store when all four are
__device__ __forceinline__ void qst(rsrc_t r, f32x4_t v, unsigned voff, unsigned soff, unsigned rows) {
    if (rows == 15u) {
        raw_buffer_store_f32x4(v, r, int(voff), int(soff), 0);
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if ((rows >> j) & 1u) bst(r, v[j], voff + 4u * unsigned(j), soff);
    }
}
__device__ __forceinline__ f32x4_t lds4(const float* p) { return *reinterpret_cast<const f32x4_t*>(p); }
__device__ __forceinline__ void lds4_st(float* p, f32x4_t v) { *reinterpret_cast<f32x4_t*>(p) = v; }

// D-/D+ along x of a quad from a window of 9 column quads (window[4] = the quad's column)
__device__ __forceinline__ f32x4_t qdm(const f32x4_t (&w)[9], const float* c) {
    return c[0] * (w[4] - w[3]) + c[1] * (w[5] - w[2]) + c[2] * (w[6] - w[1]) + c[3] * (w[7] - w[0]);
}
__device__ __forceinline__ f32x4_t qdp(const f32x4_t (&w)[9], const float* c) {
    return c[0] * (w[5] - w[4]) + c[1] * (w[6] - w[3]) + c[2] * (w[7] - w[2]) + c[3] * (w[8] - w[1]);
}
// D-/D+ along z of the quad rows 4..7 of a 12-row column window z[0..11]
__device__ __forceinline__ f32x4_t zdm(const float (&z)[12], const float* c) {
    f32x4_t r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int m = 4 + j;
        r[j] = c[0] * (z[m] - z[m - 1]) + c[1] * (z[m + 1] - z[m - 2]) + c[2] * (z[m + 2] - z[m - 3]) +
               c[3] * (z[m + 3] - z[m - 4]);
    }
    return r;
}
__device__ __forceinline__ f32x4_t zdp(const float (&z)[12], const float* c) {
    f32x4_t r;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int m = 4 + j;
        r[j] = c[0] * (z[m + 1] - z[m]) + c[1] * (z[m + 2] - z[m - 1]) + c[2] * (z[m + 3] - z[m - 2]) +
               c[3] * (z[m + 4] - z[m - 3]);
    }
    return r;
}
// the 12-row window of column `col` of an LDS array whose rows start at row0 (a multiple of 4)
__device__ __forceinline__ void zwin(const float* colp, int row0, float (&z)[12]) {
    const f32x4_t a = lds4(colp + row0), b = lds4(colp + row0 + 4), d = lds4(colp + row0 + 8);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        z[j] = a[j];
        z[4 + j] = b[j];
        z[8 + j] = d[j];
    }
}
__device__ __forceinline__ void pin4(f32x4_t& v) { asm volatile("" : "+v"(v)); }

template <typename T, bool NOREC>
__device__ __forceinline__ void step2q_fwd_tile(const ShotDev<T>* __restrict__ shots, const TileDev t,
                                                const Entry* __restrict__ ents, int k, int par) {
    static_assert(sizeof(T) == 4, "the quad kernel is FP32");
    constexpr int NT = QNT;
    constexpr int TX = QTX;
    constexpr int TX1 = TX + 8;    // step-1 columns (= 32 slots)
    constexpr int TXC = TX + 16;   // staged u_cur columns
    constexpr unsigned SZ = 4;
    static_assert(TX1 * 16 == NT, "one step-1 column per slot");
    __shared__ __attribute__((aligned(16))) float lc[2][TXC][ZC];
    __shared__ __attribute__((aligned(16))) float l1[5][TX1][Z1];

    const ShotDev<T>& S = shots[t.shot];
    const int nz = S.nz, nx = S.nx;
    const unsigned n = unsigned(nz) * unsigned(nx);
    const unsigned fbytes = n * SZ;
    const bool two = k + 1 < S.ntw - 1;
    const StepTime st1 = fwd_time(k, S.sub);
    const StepTime st2 = fwd_time(k + 1, S.sub);
    const rsrc_t rold = make_rsrc(par ? S.C : S.A, 5 * fbytes);
    const rsrc_t rcur = make_rsrc(par ? S.D : S.B, 5 * fbytes);
    const rsrc_t rmed = make_rsrc(S.med, 5 * fbytes);
    const rsrc_t rw1 = make_rsrc(par ? S.A : S.C, 5 * fbytes);
    const rsrc_t rw2 = make_rsrc(par ? S.B : S.D, 5 * fbytes);
    const T* cz = S.cz;
    const T* cx = S.cx;

    const int q = threadIdx.x & 15, s = threadIdx.x >> 4;
    const int zq = t.z0 - FAT + 4 * q;         // first row of this lane's quad (>= 0)
    const int x1 = t.x0 - FAT + s;             // step-1 column of this slot
    const int x1c = x1 < nx ? x1 : nx - 1;     // x1 >= 0 always (x0 >= FAT)
    const unsigned vq1 = (unsigned(x1c) * unsigned(nz) + unsigned(zq)) * SZ;
    unsigned ok1 = 0;                          // bit j: row zq+j of column x1 is an update cell
    {
        const bool xin = x1 >= FAT && x1 < nx - FAT;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (xin && zq + j >= FAT && zq + j < nz - FAT) ok1 |= 1u << j;
    }

    // ---- step 1: u_cur streamed field by field (sxx, szz, sxz, vx, vz) through lc[0], lc[1]
    {
        constexpr int ZQ = ZC / 4;
        constexpr int NFILL = (TXC * ZQ + NT - 1) / NT;
        unsigned hv[NFILL];
#pragma unroll
        for (int it = 0; it < NFILL; ++it) {
            const int e = threadIdx.x + it * NT;
            const int c = e / ZQ, qq = e - c * ZQ;
            const int gx = t.x0 - 2 * FAT + c, gz = t.z0 - 2 * FAT + 4 * qq;
            const bool in = e < TXC * ZQ && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
            hv[it] = in ? (unsigned(gx) * unsigned(nz) + unsigned(gz)) * SZ : 0u;
        }
        auto load_field = [&](int f, f32x4_t (&v)[NFILL]) {
#pragma unroll
            for (int it = 0; it < NFILL; ++it) v[it] = qld(rcur, hv[it], unsigned(f) * fbytes);
        };
        auto stage_field = [&](int b, const f32x4_t (&v)[NFILL]) {
#pragma unroll
            for (int it = 0; it < NFILL; ++it) {
                const int e = threadIdx.x + it * NT;
                if (e < TXC * ZQ) lds4_st(&lc[b][0][4 * e], v[it]);
            }
        };
        f32x4_t o[5], m[5];
        auto load_om = [&](int f) {
            o[f] = qld(rold, vq1, unsigned(f) * fbytes);
            m[f] = qld(rmed, vq1, unsigned(f) * fbytes);
        };
        // stage column of the slot's x is s + 4; the quad's stage rows are 4q+4 .. 4q+7
        auto xwin = [&](int b, f32x4_t (&w)[9]) {
#pragma unroll
            for (int i = 0; i < 9; ++i) w[i] = lds4(&lc[b][s + i][4 * q + 4]);
        };
        f32x4_t sa[NFILL], sb[NFILL];
        f32x4_t s0, s1, dxvx, dzvz, s4, dv;
        float z[12];
        f32x4_t w[9];
        load_field(2, sa);  // sxx
        load_field(3, sb);  // szz
        load_om(0);
        load_om(1);
        stage_field(0, sa);
        __syncthreads();
        load_field(4, sa);  // sxz
        stage_field(1, sb);
        load_field(0, sb);  // vx
        xwin(0, w);
        s0 = qdm(w, cx);    // sxx: D- along x
        pin4(s0);
        __syncthreads();
        stage_field(0, sa);
        load_field(1, sa);  // vz
        zwin(&lc[1][s + 4][0], 4 * q, z);
        s1 = zdm(z, cz);    // szz: D- along z
        pin4(s1);
        __syncthreads();
        stage_field(1, sb);
        xwin(0, w);
        s1 = s1 + qdp(w, cx);  // sxz: D+ along x into S1
        zwin(&lc[0][s + 4][0], 4 * q, z);
        s0 = s0 + zdp(z, cz);  // sxz: D+ along z into S0
#pragma unroll
        for (int j = 0; j < 4; ++j) dv[j] = ((ok1 >> j) & 1u) ? damp_at(S, zq + j, x1) : T(1);
        {  // vx and vz of u1
            const f32x4_t a0 = dv * o[0] + m[0] * s0;
            const f32x4_t a1 = dv * o[1] + m[1] * s1;
            f32x4_t u0, u1;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = (ok1 >> j) & 1u;
                u0[j] = ok ? dv[j] * a0[j] : 0.0f;
                u1[j] = ok ? dv[j] * a1[j] : 0.0f;
            }
            lds4_st(&l1[0][s][4 * q], u0);
            lds4_st(&l1[1][s][4 * q], u1);
        }
        __syncthreads();
        stage_field(0, sa);
        load_om(2);
        load_om(3);
        load_om(4);
        xwin(1, w);
        dxvx = qdp(w, cx);   // vx: D+ along x
        zwin(&lc[1][s + 4][0], 4 * q, z);
        s4 = zdm(z, cz);     // vx: D- along z
        pin4(dxvx);
        pin4(s4);
        __syncthreads();
        xwin(0, w);
        s4 = s4 + qdm(w, cx);  // vz: D- along x into S4
        zwin(&lc[0][s + 4][0], 4 * q, z);
        dzvz = zdp(z, cz);     // vz: D+ along z
        {  // the stresses of u1
            const f32x4_t a2 = dv * o[2] + (m[2] * dxvx + m[3] * dzvz);
            const f32x4_t a3 = dv * o[3] + (m[3] * dxvx + m[2] * dzvz);
            const f32x4_t a4 = dv * o[4] + m[4] * s4;
            f32x4_t u2, u3, u4;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const bool ok = (ok1 >> j) & 1u;
                u2[j] = ok ? dv[j] * a2[j] : 0.0f;
                u3[j] = ok ? dv[j] * a3[j] : 0.0f;
                u4[j] = ok ? dv[j] * a4[j] : 0.0f;
            }
            lds4_st(&l1[2][s][4 * q], u2);
            lds4_st(&l1[3][s][4 * q], u3);
            lds4_st(&l1[4][s][4 * q], u4);
        }
    }
    __syncthreads();

    const bool has_src = t.src_e > t.src_b;
    const bool has_rec = !NOREC && t.rec_e > t.rec_b;
    // ---- step-1 sources on the whole step-1 region (kernelsGpuElastic.cu:172-189)
    if (has_src) {
        for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NT) {
            const Entry en = ents[e];
            const int lz = en.lcell & 255, lx = en.lcell >> 8;
            const int comp = en.crow >> 24, row = en.crow & 0xffffff;
            const T* tr = S.src_tr[comp];
            const long b = long(row) * S.nts + st1.its;
            const T sv = tr[b] * S.filt[st1.it2 - 1] + tr[b + 1] * S.filt[S.h + st1.it2 - 1];
            l1[comp][lx][lz] += damp_at(S, t.z0 - FAT + lz, t.x0 - FAT + lx) * sv;
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
    // ---- store u1 on the owned quads: slots 4 .. 4+TX-1, quads 1 .. 14
    const bool qown = q >= 1 && q <= 14;
    if (qown && s >= FAT && s < FAT + TX && x1 < nx - FAT) {
        unsigned rows = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (zq + j < nz - FAT) rows |= 1u << j;
#pragma unroll
        for (int f = 0; f < 5; ++f) qst(rw1, lds4(&l1[f][s][4 * q]), vq1, unsigned(f) * fbytes, rows);
    }
    if (!two) return;

    // ---- step 2 on owned column s (slots 0 .. TX-1), quads 1 .. 14 (others compute a clamped
    // copy whose results are discarded)
    const int s2 = s < TX ? s : TX - 1;
    const int q2 = q < 1 ? 1 : (q > 14 ? 14 : q);
    const int x2 = t.x0 + s2;
    const int zq2 = t.z0 - FAT + 4 * q2;
    const unsigned vq2 = (unsigned(x2 < nx ? x2 : nx - 1) * unsigned(nz) + unsigned(zq2)) * SZ;
    unsigned ok2 = 0;  // bit j: row zq2+j of column x2 is an owned update cell of this lane
    if (s < TX && qown && x2 < nx - FAT) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (zq2 + j < nz - FAT) ok2 |= 1u << j;
    }
    f32x4_t v2[5];
    {
        f32x4_t mm[5], oo[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            mm[f] = qld(rmed, vq2, unsigned(f) * fbytes);
            oo[f] = qld(rcur, vq2, unsigned(f) * fbytes);  // the older level of step 2 = u_cur
        }
        // l1 column of owned column s2 is s2 + 4; the quad's l1 rows are 4 q2 .. 4 q2 + 3
        f32x4_t w[9];
        float z[12];
        auto xw = [&](int f) {
#pragma unroll
            for (int i = 0; i < 9; ++i) w[i] = lds4(&l1[f][s2 + i][4 * q2]);
        };
        xw(2);
        const f32x4_t dx_sxx = qdm(w, cx);
        xw(4);
        const f32x4_t dx_sxz = qdp(w, cx);
        xw(0);
        const f32x4_t dx_vx = qdp(w, cx);
        xw(1);
        const f32x4_t dx_vz = qdm(w, cx);
        zwin(&l1[4][s2 + 4][0], 4 * q2 - 4, z);
        const f32x4_t dz_sxz = zdp(z, cz);
        zwin(&l1[3][s2 + 4][0], 4 * q2 - 4, z);
        const f32x4_t dz_szz = zdm(z, cz);
        zwin(&l1[1][s2 + 4][0], 4 * q2 - 4, z);
        const f32x4_t dz_vz = zdp(z, cz);
        zwin(&l1[0][s2 + 4][0], 4 * q2 - 4, z);
        const f32x4_t dz_vx = zdm(z, cz);
        f32x4_t d;
#pragma unroll
        for (int j = 0; j < 4; ++j) d[j] = ((ok2 >> j) & 1u) ? damp_at(S, zq2 + j, x2) : T(1);
        const f32x4_t a0 = d * oo[0] + mm[0] * (dx_sxx + dz_sxz);
        const f32x4_t a1 = d * oo[1] + mm[1] * (dz_szz + dx_sxz);
        const f32x4_t a2 = d * oo[2] + (mm[2] * dx_vx + mm[3] * dz_vz);
        const f32x4_t a3 = d * oo[3] + (mm[3] * dx_vx + mm[2] * dz_vz);
        const f32x4_t a4 = d * oo[4] + mm[4] * (dz_vx + dx_vz);
        v2[0] = d * a0;
        v2[1] = d * a1;
        v2[2] = d * a2;
        v2[3] = d * a3;
        v2[4] = d * a4;
    }
    if (has_src || has_rec) {  // stage u2 in the l1 area (uniform per tile)
        __syncthreads();
        if (s < TX && qown) {
#pragma unroll
            for (int f = 0; f < 5; ++f) lds4_st(&l1[f][s + FAT][4 * q], v2[f]);
        }
        __syncthreads();
        if (has_src) {  // owned sources only
            for (int e = t.src_b + threadIdx.x; e < t.src_e; e += NT) {
                const Entry en = ents[e];
                const int lz = en.lcell & 255, lx = en.lcell >> 8;
                if (lz < FAT || lz >= FAT + TZ2 || lx < FAT || lx >= FAT + TX) continue;
                const int comp = en.crow >> 24, row = en.crow & 0xffffff;
                const T* tr = S.src_tr[comp];
                const long b = long(row) * S.nts + st2.its;
                const T sv = tr[b] * S.filt[st2.it2 - 1] + tr[b + 1] * S.filt[S.h + st2.it2 - 1];
                l1[comp][lx][lz] += damp_at(S, t.z0 - FAT + lz, t.x0 - FAT + lx) * sv;
            }
            __syncthreads();
            if (s < TX && qown) {
#pragma unroll
                for (int f = 0; f < 5; ++f) v2[f] = lds4(&l1[f][s + FAT][4 * q]);
            }
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
    if (ok2) {
#pragma unroll
        for (int f = 0; f < 5; ++f) qst(rw2, v2[f], vq2, unsigned(f) * fbytes, ok2);
    }
}

template <typename T, int OCC, bool NOREC = false>
__global__ __launch_bounds__(QNT) __attribute__((amdgpu_waves_per_eu(OCC)))
void k_step2q_fwd(const ShotDev<T>* __restrict__ shots, const TileDev* __restrict__ tiles,
                  const Entry* __restrict__ ents, int ntiles, int k, int par) {
    step2q_fwd_tile<T, NOREC>(shots, tiles[xcd_remap(blockIdx.x, ntiles)], ents, k, par);
}

}  // namespace eio
