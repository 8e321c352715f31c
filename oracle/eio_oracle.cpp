// eio_oracle.cpp -- CPU oracle for the 2-D isotropic elastic velocity-stress
// propagator (nonlinear and Born shot operators and their adjoints).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  The product path (HIP kernels behind include/eio.h) never calls it.
//
// It is a plain OpenMP C++ restatement of the reference arithmetic, templated
// on the real type: the double instantiation is the FP64 oracle, the float
// instantiation is the CPU timing twin.  It deliberately keeps the reference's
// UNFUSED order of operations inside a fine time step
//     step -> inject -> damp(p0 and p1) -> record/extract -> swap
// so that the fused GPU kernels are checked against the reference sequence,
// not against a copy of themselves.
//
// Reference files followed (paths relative to elastic_iso_lib/elastic_iso_float/prop):
//   FD coefficients, time filter, damping table : src/nonlinearPropElasticGpuFunctions.cu:159-202
//   forward / adjoint stencil                   : src/kernelsGpuElastic.cu:234-377 / 380-529
//   cosine damping                               : src/kernelsGpuElastic.cu:1192-1221
//   inject / record (interp in time)             : src/kernelsGpuElastic.cu:7-60, 100-116, 172-229
//   NL forward / adjoint time loops              : src/nonlinearPropElasticGpuFunctions.cu:588-617 / 966-997
//   Born source pass, scattered pass, receiver pass : src/BornElasticGpuFunctions.cu:744-819 / 1128-1191
//   Born decimation, ss injection, ss extraction : src/kernelsGpuElastic.cu:1337-1346, 1363-1376
//   Born imaging forward / adjoint               : src/kernelsGpuElastic.cu:1381-1459 / 1538-1613
//   media scaling and staggering                 : src/fdParamElastic.cpp:118-148, src/stagger.cpp:18-114
//   bilinear spatial interpolation               : src/spaceInterpGpu.cpp:78-138, 288-376
//   sinc and Gaussian spatial interpolation      : src/spaceInterpGpu.cpp:139-207 (intended stencil;
//                                                  see make_interp)
//   coarse->fine time interpolation              : src/interpTimeLinTbb.cpp:9-104
//   per-shot source scaling / packing            : src/nonlinearPropElasticGpu.cpp:43-473,
//                                                  include/seismicElasticOperator2D.cpp:27-119
//   shot stacking and Born model staggering      : src/BornElasticShotsGpu.cpp:85-378,
//                                                  src/nonlinearPropElasticShotsGpu.cpp:80-422
//   free surface (surface = 1), forward          : src/kernelsGpuElastic.cu:536-776 (top and body),
//                                                  1223-1252 (damping), src/nonlinearPropElasticGpuFunctions.cu:603-614
//   free surface, adjoint: the exact transpose of that forward step (derived below; checked by
//   the FP64 dot test), in the reference's adjoint loop order
//
// Layout conventions (same as the reference): every 2-D field is [nx][nz] with
// z fastest (index ix*nz + iz); traces are [nTrace][nts] with time fastest;
// component order vx, vz, sxx, szz, sxz.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>
#include <algorithm>
#include <omp.h>
#include <pmmintrin.h>
#include <xmmintrin.h>

namespace {

const int FAT = 4;
const double FD_C[4] = {1.196289062541883, -0.079752604188901, 0.009570312506634,
                        -6.975446437140719e-04};  // nonlinearPropElasticGpuFunctions.cu:159-167

// ---------------------------------------------------------------- parameters
struct Grid {
    int nz, nx, nts, sub;  // sub = fine sub-sampling (already doubled, fdParamElastic.cpp:28)
    int minPad;
    double dz, dx, dts, alphaCos;
    double oz, ox;         // origin of the centre grid (used by the spatial interpolation)
    int surface;           // 1: free surface at row 2*FAT (fdParamElastic.cpp:18)
};

template <typename T>
struct Consts {
    int nz, nx, nts, sub, ntw, h, minPad, surface;
    T cz[4], cx[4];
    std::vector<T> F;     // 2*(sub+1) time-interpolation filter
    std::vector<T> damp;  // minPad cosine damping coefficients
    T dtw, dts_inv;
};

template <typename T>
Consts<T> make_consts(const Grid& g) {
    Consts<T> c;
    c.nz = g.nz; c.nx = g.nx; c.nts = g.nts; c.sub = g.sub;
    c.ntw = (g.nts - 1) * g.sub + 1;
    c.h = g.sub + 1;
    c.minPad = g.minPad;
    c.surface = g.surface;
    for (int k = 0; k < 4; ++k) { c.cz[k] = T(FD_C[k] / g.dz); c.cx[k] = T(FD_C[k] / g.dx); }
    // nonlinearPropElasticGpuFunctions.cu:172-188
    c.F.assign(2 * c.h, T(0));
    const double s = 1.0 / std::sqrt(double(c.ntw) / double(c.nts));
    for (int i = 0; i < c.h; ++i) {
        const double a = 1.0 - double(i) / double(g.sub);
        c.F[i] = T(a * s);
        c.F[c.h + i] = T((1.0 - a) * s);
    }
    // nonlinearPropElasticGpuFunctions.cu:195-202
    c.damp.assign(std::max(g.minPad, 1), T(1));
    for (int j = 0; j < g.minPad; ++j) {
        const double arg = M_PI / double(g.minPad) * double(g.minPad - j);
        c.damp[j] = T(g.alphaCos + (1.0 - g.alphaCos) * std::cos(arg));
    }
    c.dtw = T(g.dts / double(g.sub));
    c.dts_inv = T(1.0 / (2.0 * g.dts));  // BornElasticGpuFunctions.cu:264-265
    return c;
}

// ---------------------------------------------------------------- state
template <typename T>
struct Fields {
    size_t n;
    std::vector<T> f[5];  // vx, vz, sxx, szz, sxz
    explicit Fields(size_t n_) : n(n_) { for (auto& v : f) v.assign(n_, T(0)); }
    void zero() { for (auto& v : f) std::fill(v.begin(), v.end(), T(0)); }
};

template <typename T>
struct Media { const T *rhox, *rhoz, *l2m, *lam, *muxz; };

// D- f[i] = sum_k C_k (f[i+k] - f[i-1-k]);  D+ f[i] = sum_k C_k (f[i+1+k] - f[i-k])
template <typename T>
inline T dminus(const T* f, long i, long s, const T* c) {
    return c[0] * (f[i] - f[i - s]) + c[1] * (f[i + s] - f[i - 2 * s]) +
           c[2] * (f[i + 2 * s] - f[i - 3 * s]) + c[3] * (f[i + 3 * s] - f[i - 4 * s]);
}
template <typename T>
inline T dplus(const T* f, long i, long s, const T* c) {
    return c[0] * (f[i + s] - f[i]) + c[1] * (f[i + 2 * s] - f[i - s]) +
           c[2] * (f[i + 3 * s] - f[i - 2 * s]) + c[3] * (f[i + 4 * s] - f[i - 3 * s]);
}

// ---------------------------------------------------------------- free surface
// Surface row and the top-block map M of the current level (kernelsGpuElastic.cu:536-587),
// column by column: sigmazz = 0 on the surface and odd above it, sigmaxz odd about the half
// row above the surface (rows s-3..s), vx = 0 above the surface, vz = 0 on and above it.
const int SURF = 2 * FAT;

template <typename T>
void fs_top_map(const Consts<T>& c, Fields<T>& p1) {
    const long nz = c.nz;
    for (int ix = FAT; ix < c.nx - FAT; ++ix) {
        const long b = long(ix) * nz;
        T* szz = p1.f[3].data() + b;
        T* sxz = p1.f[4].data() + b;
        T* vx = p1.f[0].data() + b;
        T* vz = p1.f[1].data() + b;
        szz[SURF] = 0;
        for (int k = 0; k < 4; ++k) szz[SURF - 1 - k] = -szz[SURF + 1 + k];
        for (int k = 0; k < 4; ++k) sxz[SURF - k] = -sxz[SURF + 1 + k];
        for (int r = SURF - 4; r < SURF; ++r) vx[r] = 0;
        for (int r = SURF - 4; r <= SURF; ++r) vz[r] = 0;
    }
}

// Transpose of fs_top_map (rows that M overwrites receive nothing; the mirrored source rows
// collect the negated values of their images).
template <typename T>
void fs_top_map_transpose(const Consts<T>& c, Fields<T>& p) {
    const long nz = c.nz;
    for (int ix = FAT; ix < c.nx - FAT; ++ix) {
        const long b = long(ix) * nz;
        T* szz = p.f[3].data() + b;
        T* sxz = p.f[4].data() + b;
        T* vx = p.f[0].data() + b;
        T* vz = p.f[1].data() + b;
        for (int k = 0; k < 4; ++k) szz[SURF + 1 + k] -= szz[SURF - 1 - k];
        for (int r = SURF - 4; r <= SURF; ++r) szz[r] = 0;
        for (int k = 0; k < 4; ++k) sxz[SURF + 1 + k] -= sxz[SURF - k];
        for (int r = SURF - 3; r <= SURF; ++r) sxz[r] = 0;
        for (int r = SURF - 4; r < SURF; ++r) vx[r] = 0;
        for (int r = SURF - 4; r <= SURF; ++r) vz[r] = 0;
    }
}

// Free-surface body (kernelsGpuElastic.cu:589-776): vx updated on and below the surface, vz,
// sigmazz, sigmaxz below it, sigmaxx below it and on it with the modulus
// lamb2MuDtw - lambDtw^2/lamb2MuDtw and no z derivative.
template <typename T>
void step_fwd_fs(const Consts<T>& c, const Media<T>& m, Fields<T>& p0, Fields<T>& p1) {
    fs_top_map(c, p1);
    const long nz = c.nz;
    const T *vx = p1.f[0].data(), *vz = p1.f[1].data(), *sxx = p1.f[2].data(),
            *szz = p1.f[3].data(), *sxz = p1.f[4].data();
#pragma omp parallel for schedule(static)
    for (int ix = FAT; ix < c.nx - FAT; ++ix) {
        for (int iz = SURF; iz < c.nz - FAT; ++iz) {
            const long i = long(ix) * nz + iz;
            const T dxp_vx = dplus(vx, i, nz, c.cx);
            p0.f[0][i] += m.rhox[i] * (dminus(sxx, i, nz, c.cx) + dplus(sxz, i, 1, c.cz));
            if (iz == SURF) {
                p0.f[2][i] += (m.l2m[i] - m.lam[i] * m.lam[i] / m.l2m[i]) * dxp_vx;
                continue;
            }
            const T dzp_vz = dplus(vz, i, 1, c.cz);
            p0.f[1][i] += m.rhoz[i] * (dminus(szz, i, 1, c.cz) + dplus(sxz, i, nz, c.cx));
            p0.f[2][i] += m.l2m[i] * dxp_vx + m.lam[i] * dzp_vz;
            p0.f[3][i] += m.lam[i] * dxp_vx + m.l2m[i] * dzp_vz;
            p0.f[4][i] += m.muxz[i] * (dminus(vx, i, 1, c.cz) + dminus(vz, i, nz, c.cx));
        }
    }
}

// kernelsGpuElastic.cu:298-376: p0 += A(p1), in place, update region only.
template <typename T>
void step_fwd(const Consts<T>& c, const Media<T>& m, Fields<T>& p0, const Fields<T>& p1) {
    const long nz = c.nz;
    const T *vx = p1.f[0].data(), *vz = p1.f[1].data(), *sxx = p1.f[2].data(),
            *szz = p1.f[3].data(), *sxz = p1.f[4].data();
#pragma omp parallel for schedule(static)
    for (int ix = FAT; ix < c.nx - FAT; ++ix) {
        for (int iz = FAT; iz < c.nz - FAT; ++iz) {
            const long i = long(ix) * nz + iz;
            const T dxp_vx = dplus(vx, i, nz, c.cx), dzp_vz = dplus(vz, i, 1, c.cz);
            p0.f[0][i] += m.rhox[i] * (dminus(sxx, i, nz, c.cx) + dplus(sxz, i, 1, c.cz));
            p0.f[1][i] += m.rhoz[i] * (dminus(szz, i, 1, c.cz) + dplus(sxz, i, nz, c.cx));
            p0.f[2][i] += m.l2m[i] * dxp_vx + m.lam[i] * dzp_vz;
            p0.f[3][i] += m.lam[i] * dxp_vx + m.l2m[i] * dzp_vz;
            p0.f[4][i] += m.muxz[i] * (dminus(vx, i, 1, c.cz) + dminus(vz, i, nz, c.cx));
        }
    }
}

// kernelsGpuElastic.cu:402-528: p0 -= A^T-like terms of media-scaled p1.
template <typename T>
void step_adj(const Consts<T>& c, const Media<T>& m, Fields<T>& p0, const Fields<T>& p1,
              std::vector<T> (&scr)[7]) {
    const long nz = c.nz, n = long(c.nz) * c.nx;
    // scaled copies of the current slice (the reference scales each halo cell by its own medium value)
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        scr[0][i] = m.rhox[i] * p1.f[0][i];
        scr[1][i] = m.rhoz[i] * p1.f[1][i];
        scr[2][i] = m.l2m[i] * p1.f[2][i];
        scr[3][i] = m.lam[i] * p1.f[2][i];
        scr[4][i] = m.l2m[i] * p1.f[3][i];
        scr[5][i] = m.lam[i] * p1.f[3][i];
        scr[6][i] = m.muxz[i] * p1.f[4][i];
    }
    const T *vxr = scr[0].data(), *vzr = scr[1].data(), *sxx_l2m = scr[2].data(),
            *sxx_lam = scr[3].data(), *szz_l2m = scr[4].data(), *szz_lam = scr[5].data(),
            *sxz_mu = scr[6].data();
#pragma omp parallel for schedule(static)
    for (int ix = FAT; ix < c.nx - FAT; ++ix) {
        for (int iz = FAT; iz < c.nz - FAT; ++iz) {
            const long i = long(ix) * nz + iz;
            p0.f[0][i] -= dminus(sxx_l2m, i, nz, c.cx) + dminus(szz_lam, i, nz, c.cx) +
                          dplus(sxz_mu, i, 1, c.cz);
            p0.f[1][i] -= dminus(sxx_lam, i, 1, c.cz) + dminus(szz_l2m, i, 1, c.cz) +
                          dplus(sxz_mu, i, nz, c.cx);
            p0.f[2][i] -= dplus(vxr, i, nz, c.cx);
            p0.f[3][i] -= dplus(vzr, i, 1, c.cz);
            p0.f[4][i] -= dminus(vxr, i, 1, c.cz) + dminus(vzr, i, nz, c.cx);
        }
    }
}

// Transpose of the free-surface step: p0 += B^T p1 (the body operator B restricted to the
// rows the forward updates, with the surface row's sigmaxx modulus), then p0 <- M^T p0.
// With L the damped adjoint state, L_k = D M^T (B^T L_k+1 + D L_k+2) + D R^T d_k, which in
// the reference's adjoint loop (step, inject, damp both, extract, swap) is this step
// followed by the usual injection and damping; M^T commutes with D on the top rows (their
// damping depends on x only).
template <typename T>
void step_adj_fs(const Consts<T>& c, const Media<T>& m, Fields<T>& p0, const Fields<T>& p1,
                 std::vector<T> (&scr)[7]) {
    const long nz = c.nz, n = long(c.nz) * c.nx;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        const int iz = int(i % nz);
        const bool below = iz > SURF, on = iz == SURF;
        scr[0][i] = (below || on) ? m.rhox[i] * p1.f[0][i] : T(0);
        scr[1][i] = below ? m.rhoz[i] * p1.f[1][i] : T(0);
        scr[2][i] = below ? m.l2m[i] * p1.f[2][i]
                          : (on ? (m.l2m[i] - m.lam[i] * m.lam[i] / m.l2m[i]) * p1.f[2][i] : T(0));
        scr[3][i] = below ? m.lam[i] * p1.f[2][i] : T(0);
        scr[4][i] = below ? m.l2m[i] * p1.f[3][i] : T(0);
        scr[5][i] = below ? m.lam[i] * p1.f[3][i] : T(0);
        scr[6][i] = below ? m.muxz[i] * p1.f[4][i] : T(0);
    }
    const T *vxr = scr[0].data(), *vzr = scr[1].data(), *sxx_l2m = scr[2].data(),
            *sxx_lam = scr[3].data(), *szz_l2m = scr[4].data(), *szz_lam = scr[5].data(),
            *sxz_mu = scr[6].data();
#pragma omp parallel for schedule(static)
    for (int ix = FAT; ix < c.nx - FAT; ++ix) {
        for (int iz = FAT; iz < c.nz - FAT; ++iz) {
            const long i = long(ix) * nz + iz;
            p0.f[0][i] -= dminus(sxx_l2m, i, nz, c.cx) + dminus(szz_lam, i, nz, c.cx) +
                          dplus(sxz_mu, i, 1, c.cz);
            p0.f[1][i] -= dminus(sxx_lam, i, 1, c.cz) + dminus(szz_l2m, i, 1, c.cz) +
                          dplus(sxz_mu, i, nz, c.cx);
            p0.f[2][i] -= dplus(vxr, i, nz, c.cx);
            p0.f[3][i] -= dplus(vzr, i, 1, c.cz);
            p0.f[4][i] -= dminus(vxr, i, 1, c.cz) + dminus(vzr, i, nz, c.cx);
        }
    }
    fs_top_map_transpose(c, p0);
}

// kernelsGpuElastic.cu:1192-1221 (1223-1252 with a free surface: no damping towards the top):
// both time levels, update region only.
template <typename T>
void damp_both(const Consts<T>& c, Fields<T>& p0, Fields<T>& p1) {
    const long nz = c.nz;
#pragma omp parallel for schedule(static)
    for (int ix = FAT; ix < c.nx - FAT; ++ix) {
        for (int iz = FAT; iz < c.nz - FAT; ++iz) {
            const int top = c.surface ? c.nz : iz - FAT;
            const int d = std::min(std::min(top, ix - FAT),
                                   std::min(c.nz - iz - 1 - FAT, c.nx - ix - 1 - FAT));
            if (d < c.minPad) {
                const T w = c.damp[d];
                const long i = long(ix) * nz + iz;
                for (int k = 0; k < 5; ++k) { p0.f[k][i] *= w; p1.f[k][i] *= w; }
            }
        }
    }
}

// Regular-grid device lists for the 4 staggered grids (centre, x, z, xz).
struct RegGeom {
    const int* pos[4];
    int n[4];
};
// component -> grid: vx on X, vz on Z, sxx/szz on centre, sxz on XZ
const int COMP_GRID[5] = {1, 2, 0, 0, 3};

// Coarse-sampled wavefield [nts][5][nz*nx] of the update region, accumulated like the data
// (kernelsGpuElastic.cu:1299-1335, the row/column ranges of the streamed variant)
template <typename T>
void save_wavefield(const Consts<T>& c, const Fields<T>& p, int its, T wa, T wb, T* wfld) {
    if (!wfld) return;
    const size_t n = size_t(c.nz) * c.nx;
    T* a = wfld + size_t(its) * 5 * n;
    T* b = a + 5 * n;
    for (int k = 0; k < 5; ++k)
        for (int ix = FAT; ix < c.nx - FAT; ++ix)
            for (int iz = FAT; iz < c.nz - FAT; ++iz) {
                const size_t i = size_t(ix) * c.nz + iz;
                a[k * n + i] += p.f[k][i] * wa;
                b[k * n + i] += p.f[k][i] * wb;
            }
}

// ---------------------------------------------------------------- NL (regular level)
// nonlinearPropElasticGpuFunctions.cu:561-648 ; traces are [nReg][nts], per component
// (wfld != nullptr: also the forward wavefield, nonlinearPropElasticGpuFunctions.cu:649-760)
template <typename T>
void nl_fwd_reg(const Consts<T>& c, const Media<T>& m, const RegGeom& src, const T* const mod[5],
                const RegGeom& rec, T* const dat[5], T* wfld = nullptr) {
    const size_t n = size_t(c.nz) * c.nx;
    Fields<T> a(n), b(n);
    Fields<T>*p0 = &a, *p1 = &b;
    for (int k = 0; k < 5; ++k) {
        const int gi = COMP_GRID[k];
        std::fill(dat[k], dat[k] + size_t(rec.n[gi]) * c.nts, T(0));
    }
    for (int its = 0; its < c.nts - 1; ++its) {
        for (int it2 = 1; it2 < c.sub + 1; ++it2) {
            if (c.surface) step_fwd_fs(c, m, *p0, *p1);
            else step_fwd(c, m, *p0, *p1);
            for (int k = 0; k < 5; ++k) {  // kernelsGpuElastic.cu:172-189, it2-1 lag
                const int gi = COMP_GRID[k];
                for (int is = 0; is < src.n[gi]; ++is)
                    p0->f[k][src.pos[gi][is]] += mod[k][size_t(c.nts) * is + its] * c.F[it2 - 1] +
                                                 mod[k][size_t(c.nts) * is + its + 1] * c.F[c.h + it2 - 1];
            }
            damp_both(c, *p0, *p1);
            save_wavefield(c, *p0, its, c.F[it2], c.F[c.h + it2], wfld);
            for (int k = 0; k < 5; ++k) {  // kernelsGpuElastic.cu:7-60
                const int gi = COMP_GRID[k];
                for (int ir = 0; ir < rec.n[gi]; ++ir) {
                    const T v = p0->f[k][rec.pos[gi][ir]];
                    dat[k][size_t(c.nts) * ir + its] += v * c.F[it2];
                    dat[k][size_t(c.nts) * ir + its + 1] += v * c.F[c.h + it2];
                }
            }
            std::swap(p0, p1);
        }
    }
}

// nonlinearPropElasticGpuFunctions.cu:932-1028 (wfld: the adjoint wavefield before the host
// scaling, nonlinearPropElasticGpuFunctions.cu:1029-1090)
template <typename T>
void nl_adj_reg(const Consts<T>& c, const Media<T>& m, const RegGeom& src, T* const mod[5],
                const RegGeom& rec, const T* const dat[5], T* wfld = nullptr) {
    const size_t n = size_t(c.nz) * c.nx;
    Fields<T> a(n), b(n);
    std::vector<T> scr[7];
    for (auto& v : scr) v.assign(n, T(0));
    Fields<T>*p0 = &a, *p1 = &b;
    for (int k = 0; k < 5; ++k) {
        const int gi = COMP_GRID[k];
        std::fill(mod[k], mod[k] + size_t(src.n[gi]) * c.nts, T(0));
    }
    for (int its = c.nts - 2; its > -1; --its) {
        for (int it2 = c.sub - 1; it2 > -1; --it2) {
            if (c.surface) step_adj_fs(c, m, *p0, *p1, scr);
            else step_adj(c, m, *p0, *p1, scr);
            for (int k = 0; k < 5; ++k) {  // kernelsGpuElastic.cu:191-229, it2+1
                const int gi = COMP_GRID[k];
                for (int ir = 0; ir < rec.n[gi]; ++ir)
                    p0->f[k][rec.pos[gi][ir]] += dat[k][size_t(c.nts) * ir + its] * c.F[it2 + 1] +
                                                 dat[k][size_t(c.nts) * ir + its + 1] * c.F[c.h + it2 + 1];
            }
            damp_both(c, *p0, *p1);
            save_wavefield(c, *p0, its, c.F[it2], c.F[c.h + it2], wfld);
            for (int k = 0; k < 5; ++k) {  // kernelsGpuElastic.cu:100-116
                const int gi = COMP_GRID[k];
                for (int is = 0; is < src.n[gi]; ++is) {
                    const T v = p0->f[k][src.pos[gi][is]];
                    mod[k][size_t(c.nts) * is + its] += v * c.F[it2];
                    mod[k][size_t(c.nts) * is + its + 1] += v * c.F[c.h + it2];
                }
            }
            std::swap(p0, p1);
        }
    }
}

// ---------------------------------------------------------------- Born (regular level)
// Source pass with coarse storage of Vx, Vz (BornElasticGpuFunctions.cu:748-770,
// kernelsGpuElastic.cu:121-168 with the itw-1 lag, 1337-1346 decimation).
// srcFine: [nReg][ntw] per component (already scaled and time-interpolated).
template <typename T>
void born_source_pass(const Consts<T>& c, const Media<T>& m, const RegGeom& src,
                      const T* const srcFine[5], std::vector<T>& Wvx, std::vector<T>& Wvz) {
    const size_t n = size_t(c.nz) * c.nx;
    Wvx.assign(n * c.nts, T(0));
    Wvz.assign(n * c.nts, T(0));
    Fields<T> a(n), b(n);
    Fields<T>*p0 = &a, *p1 = &b;
    for (int its = 0; its < c.nts - 1; ++its) {
        for (int it2 = 1; it2 < c.sub + 1; ++it2) {
            const int itw = its * c.sub + it2;
            step_fwd(c, m, *p0, *p1);
            for (int k = 0; k < 5; ++k) {
                const int gi = COMP_GRID[k];
                for (int is = 0; is < src.n[gi]; ++is)
                    p0->f[k][src.pos[gi][is]] += srcFine[k][size_t(c.ntw) * is + itw - 1];
            }
            damp_both(c, *p0, *p1);
            T* w0x = Wvx.data() + n * its;
            T* w0z = Wvz.data() + n * its;
#pragma omp parallel for schedule(static)
            for (int ix = FAT; ix < c.nx - FAT; ++ix)
                for (int iz = FAT; iz < c.nz - FAT; ++iz) {
                    const size_t i = size_t(ix) * c.nz + iz;
                    w0x[i] += p0->f[0][i] * c.F[it2];
                    w0x[i + n] += p0->f[0][i] * c.F[c.h + it2];
                    w0z[i] += p0->f[1][i] * c.F[it2];
                    w0z[i + n] += p0->f[1][i] * c.F[c.h + it2];
                }
            std::swap(p0, p1);
        }
    }
}

// kernelsGpuElastic.cu:1381-1459 (writes the 5 secondary sources at coarse time its)
template <typename T>
void imaging_fwd(const Consts<T>& c, const std::vector<T>& Wvx, const std::vector<T>& Wvz,
                 Fields<T>& ss, const T* const dm[5], int its) {
    const size_t n = size_t(c.nz) * c.nx;
    const long nz = c.nz;
    const T* wx = Wvx.data() + n * its;
    const T* wz = Wvz.data() + n * its;
#pragma omp parallel for schedule(static)
    for (int ix = FAT; ix < c.nx - FAT; ++ix)
        for (int iz = FAT; iz < c.nz - FAT; ++iz) {
            const long i = long(ix) * nz + iz;
            const T dvx_dx = dplus(wx, i, nz, c.cx);
            const T dvz_dz = dplus(wz, i, 1, c.cz);
            T tx, tz;
            if (its == 0) { tx = -wx[i + n]; tz = -wz[i + n]; }
            else if (its == c.nts - 1) { tx = wx[i - long(n)]; tz = wz[i - long(n)]; }
            else { tx = wx[i - long(n)] - wx[i + n]; tz = wz[i - long(n)] - wz[i + n]; }
            ss.f[0][i] = dm[0][i] * tx * c.dts_inv;
            ss.f[1][i] = dm[1][i] * tz * c.dts_inv;
            ss.f[2][i] = dm[2][i] * (dvx_dx + dvz_dz) + T(2) * dm[3][i] * dvx_dx;
            ss.f[3][i] = dm[2][i] * (dvx_dx + dvz_dz) + T(2) * dm[3][i] * dvz_dz;
            ss.f[4][i] = dm[4][i] * (dminus(wx, i, 1, c.cz) + dminus(wz, i, nz, c.cx));
        }
}

// kernelsGpuElastic.cu:1538-1613 (accumulates the imaging condition into dm)
template <typename T>
void imaging_adj(const Consts<T>& c, const std::vector<T>& Wvx, const std::vector<T>& Wvz,
                 const Fields<T>& ss, T* const dm[5], int its) {
    const size_t n = size_t(c.nz) * c.nx;
    const long nz = c.nz;
    const T* wx = Wvx.data() + n * its;
    const T* wz = Wvz.data() + n * its;
#pragma omp parallel for schedule(static)
    for (int ix = FAT; ix < c.nx - FAT; ++ix)
        for (int iz = FAT; iz < c.nz - FAT; ++iz) {
            const long i = long(ix) * nz + iz;
            const T dvx_dx = dplus(wx, i, nz, c.cx);
            const T dvz_dz = dplus(wz, i, 1, c.cz);
            T tx, tz;
            if (its == 0) { tx = -wx[i + n]; tz = -wz[i + n]; }
            else if (its == c.nts - 1) { tx = wx[i - long(n)]; tz = wz[i - long(n)]; }
            else { tx = wx[i - long(n)] - wx[i + n]; tz = wz[i - long(n)] - wz[i + n]; }
            dm[0][i] += ss.f[0][i] * tx * c.dts_inv;
            dm[1][i] += ss.f[1][i] * tz * c.dts_inv;
            dm[2][i] += (ss.f[2][i] + ss.f[3][i]) * (dvx_dx + dvz_dz);
            dm[3][i] += T(2) * (dvx_dx * ss.f[2][i] + dvz_dz * ss.f[3][i]);
            dm[4][i] += ss.f[4][i] * (dminus(wx, i, 1, c.cz) + dminus(wz, i, nz, c.cx));
        }
}

// BornElasticGpuFunctions.cu:719-830 (dm already staggered and scaled)
template <typename T>
void born_fwd_reg(const Consts<T>& c, const Media<T>& m, const RegGeom& src,
                  const T* const srcFine[5], const T* const dm[5], const RegGeom& rec,
                  T* const dat[5]) {
    const size_t n = size_t(c.nz) * c.nx;
    std::vector<T> Wvx, Wvz;
    born_source_pass(c, m, src, srcFine, Wvx, Wvz);
    Fields<T> a(n), b(n), sL(n), sR(n);
    Fields<T>*p0 = &a, *p1 = &b, *ssL = &sL, *ssR = &sR;
    for (int k = 0; k < 5; ++k) {
        const int gi = COMP_GRID[k];
        std::fill(dat[k], dat[k] + size_t(rec.n[gi]) * c.nts, T(0));
    }
    imaging_fwd(c, Wvx, Wvz, *ssL, dm, 0);
    for (int its = 0; its < c.nts - 1; ++its) {
        imaging_fwd(c, Wvx, Wvz, *ssR, dm, its + 1);
        for (int it2 = 1; it2 < c.sub + 1; ++it2) {
            step_fwd(c, m, *p0, *p1);
            const T fl = c.F[it2 - 1], fr = c.F[c.h + it2 - 1];  // kernelsGpuElastic.cu:1363-1368
#pragma omp parallel for schedule(static)
            for (int ix = FAT; ix < c.nx - FAT; ++ix)
                for (int iz = FAT; iz < c.nz - FAT; ++iz) {
                    const size_t i = size_t(ix) * c.nz + iz;
                    for (int k = 0; k < 5; ++k) p0->f[k][i] += ssL->f[k][i] * fl + ssR->f[k][i] * fr;
                }
            damp_both(c, *p0, *p1);
            for (int k = 0; k < 5; ++k) {
                const int gi = COMP_GRID[k];
                for (int ir = 0; ir < rec.n[gi]; ++ir) {
                    const T v = p0->f[k][rec.pos[gi][ir]];
                    dat[k][size_t(c.nts) * ir + its] += v * c.F[it2];
                    dat[k][size_t(c.nts) * ir + its + 1] += v * c.F[c.h + it2];
                }
            }
            std::swap(p0, p1);
        }
        std::swap(ssL, ssR);
        ssR->zero();
    }
}

// BornElasticGpuFunctions.cu:1100-1191 (accumulates into dm, staggered/scaled space)
template <typename T>
void born_adj_reg(const Consts<T>& c, const Media<T>& m, const RegGeom& src,
                  const T* const srcFine[5], T* const dm[5], const RegGeom& rec,
                  const T* const dat[5]) {
    const size_t n = size_t(c.nz) * c.nx;
    std::vector<T> Wvx, Wvz;
    born_source_pass(c, m, src, srcFine, Wvx, Wvz);
    Fields<T> a(n), b(n), sL(n), sR(n);
    std::vector<T> scr[7];
    for (auto& v : scr) v.assign(n, T(0));
    Fields<T>*p0 = &a, *p1 = &b, *ssL = &sL, *ssR = &sR;
    for (int its = c.nts - 2; its > -1; --its) {
        for (int it2 = c.sub - 1; it2 > -1; --it2) {
            step_adj(c, m, *p0, *p1, scr);
            for (int k = 0; k < 5; ++k) {
                const int gi = COMP_GRID[k];
                for (int ir = 0; ir < rec.n[gi]; ++ir)
                    p0->f[k][rec.pos[gi][ir]] += dat[k][size_t(c.nts) * ir + its] * c.F[it2 + 1] +
                                                 dat[k][size_t(c.nts) * ir + its + 1] * c.F[c.h + it2 + 1];
            }
            damp_both(c, *p0, *p1);
            const T fl = c.F[it2], fr = c.F[c.h + it2];  // kernelsGpuElastic.cu:1369-1376
#pragma omp parallel for schedule(static)
            for (int ix = FAT; ix < c.nx - FAT; ++ix)
                for (int iz = FAT; iz < c.nz - FAT; ++iz) {
                    const size_t i = size_t(ix) * c.nz + iz;
                    for (int k = 0; k < 5; ++k) {
                        ssL->f[k][i] += p0->f[k][i] * fl;
                        ssR->f[k][i] += p0->f[k][i] * fr;
                    }
                }
            std::swap(p0, p1);
        }
        imaging_adj(c, Wvx, Wvz, *ssR, dm, its + 1);
        std::swap(ssL, ssR);
        ssL->zero();
    }
    imaging_adj(c, Wvx, Wvz, *ssR, dm, 0);
}

// ---------------------------------------------------------------- host-side helpers
// stagger.cpp:38-57 (adjoint) and :18-35 (forward), X = along the slow axis
template <typename T>
void stagger_x_adj(int nz, int nx, const T* data, T* model) {
    for (int iz = 0; iz < nz; ++iz) {
        for (int ix = 1; ix < nx - 1; ++ix)
            model[size_t(ix) * nz + iz] = T(0.5) * (data[size_t(ix) * nz + iz] + data[size_t(ix - 1) * nz + iz]);
        model[iz] = T(0.5) * data[iz];
        model[size_t(nx - 1) * nz + iz] = data[size_t(nx - 1) * nz + iz] + T(0.5) * data[size_t(nx - 2) * nz + iz];
    }
}
template <typename T>
void stagger_z_adj(int nz, int nx, const T* data, T* model) {
    for (int ix = 0; ix < nx; ++ix) {
        const T* d = data + size_t(ix) * nz;
        T* mo = model + size_t(ix) * nz;
        for (int iz = 1; iz < nz - 1; ++iz) mo[iz] = T(0.5) * (d[iz] + d[iz - 1]);
        mo[0] = T(0.5) * d[0];
        mo[nz - 1] = d[nz - 1] + T(0.5) * d[nz - 2];
    }
}
template <typename T>
void stagger_x_fwd_add(int nz, int nx, const T* model, T* data) {
    for (int iz = 0; iz < nz; ++iz) {
        for (int ix = 0; ix < nx - 1; ++ix)
            data[size_t(ix) * nz + iz] += T(0.5) * (model[size_t(ix) * nz + iz] + model[size_t(ix + 1) * nz + iz]);
        data[size_t(nx - 1) * nz + iz] += model[size_t(nx - 1) * nz + iz];
    }
}
template <typename T>
void stagger_z_fwd_add(int nz, int nx, const T* model, T* data) {
    for (int ix = 0; ix < nx; ++ix) {
        const T* mo = model + size_t(ix) * nz;
        T* d = data + size_t(ix) * nz;
        for (int iz = 0; iz < nz - 1; ++iz) d[iz] += T(0.5) * (mo[iz] + mo[iz + 1]);
        d[nz - 1] += mo[nz - 1];
    }
}

// fdParamElastic.cpp:118-148 ; out[5][nz*nx] = rhoxDtw, rhozDtw, lamb2MuDtw, lambDtw, muxzDtw
template <typename T>
void media_scaled(const Grid& g, const T* elastic, T* out) {
    const size_t n = size_t(g.nz) * g.nx;
    const T* rho = elastic;
    const T* lam = elastic + n;
    const T* mu = elastic + 2 * n;
    std::vector<T> rhox(n), rhoz(n), tmp(n), muxz(n);
    stagger_x_adj(g.nz, g.nx, rho, rhox.data());
    stagger_z_adj(g.nz, g.nx, rho, rhoz.data());
    stagger_x_adj(g.nz, g.nx, mu, tmp.data());
    stagger_z_adj(g.nz, g.nx, tmp.data(), muxz.data());
    const T two_dtw = T(2.0 * (g.dts / double(g.sub)));
    for (size_t i = 0; i < n; ++i) {
        out[i] = two_dtw / rhox[i];
        out[n + i] = two_dtw / rhoz[i];
        out[2 * n + i] = two_dtw * (lam[i] + T(2) * mu[i]);
        out[3 * n + i] = two_dtw * lam[i];
        out[4 * n + i] = two_dtw * muxz[i];
    }
}

// spaceInterpGpu.cpp:78-138 (linear, with the dipole of :108-137) + :288-311 unique list
struct Interp {
    int nIrreg;
    int nf = 4;                // regular points per device
    std::vector<int> idx;      // nf per device, non-unique
    std::vector<double> w;     // nf per device
    std::vector<int> uniq;     // unique regular positions, first-seen order
    std::vector<int> map;      // nf per device -> index into uniq
};
// method 0: bilinear (spaceInterpGpu.cpp:78-138).  Methods 1 (sinc) and 2 (Gaussian,
// standard deviation 2 cells) follow calcSincWeights / calcGaussWeights
// (spaceInterpGpu.cpp:139-207): the (2 nfilt)^2 regular points whose indices run from
// int(position) - nfilt + 1 to int(position) + nfilt along z and x, stored x-major, with
// separable weights normalised to sum 1:
//   sinc:  sin(pi u)/(pi u) * sin(pi v)/(pi v)      u, v = device - point, in cells
//   gauss: exp(-u^2/8) * exp(-v^2/8)                 (the pdf constants cancel in the normalisation)
// Dipole (linear only, spaceInterpGpu.cpp:108-137): 4 more points per device, the bilinear
// stencil of the position shifted by (zsh dz, xsh dx) with negated weights (entries 4..7).
// Computed here in double from the device offsets.
Interp make_interp(int nz, double oz, double dz, double ox, double dx, int nDev, const double* zc,
                   const double* xc, int method = 0, int nfilt = 1, int dipole = 0, double zsh = 0.0,
                   double xsh = 0.0) {
    Interp ip;
    ip.nIrreg = nDev;
    const int n2 = 2 * nfilt;
    ip.nf = method == 0 ? (dipole == 1 ? 8 : 4) : n2 * n2;
    ip.idx.resize(size_t(ip.nf) * nDev);
    ip.w.resize(size_t(ip.nf) * nDev);
    for (int d = 0; d < nDev; ++d) {
        double wz = (zc[d] - oz) / dz, wx = (xc[d] - ox) / dx;
        const int zr = int(wz), xr = int(wx);
        int* id = ip.idx.data() + size_t(ip.nf) * d;
        double* ww = ip.w.data() + size_t(ip.nf) * d;
        if (method == 0) {
            wz = 1.0 - (wz - zr);
            wx = 1.0 - (wx - xr);
            const int b = xr * nz + zr;
            id[0] = b;          ww[0] = wz * wx;
            id[1] = b + 1;      ww[1] = (1.0 - wz) * wx;
            id[2] = b + nz;     ww[2] = wz * (1.0 - wx);
            id[3] = b + nz + 1; ww[3] = (1.0 - wz) * (1.0 - wx);
            if (dipole == 1) {
                double vz = (zc[d] + zsh * dz - oz) / dz, vx = (xc[d] + xsh * dx - ox) / dx;
                const int qz = int(vz), qx = int(vx);
                vz = 1.0 - (vz - qz);
                vx = 1.0 - (vx - qx);
                const int q = qx * nz + qz;
                id[4] = q;          ww[4] = -vz * vx;
                id[5] = q + 1;      ww[5] = -(1.0 - vz) * vx;
                id[6] = q + nz;     ww[6] = -vz * (1.0 - vx);
                id[7] = q + nz + 1; ww[7] = -(1.0 - vz) * (1.0 - vx);
            }
            continue;
        }
        auto kern = [method](double u) {
            if (method == 1) return u == 0.0 ? 1.0 : std::sin(M_PI * u) / (M_PI * u);
            return std::exp(-u * u / 8.0);
        };
        double sum = 0.0;
        for (int ix = 0; ix < n2; ++ix)
            for (int iz = 0; iz < n2; ++iz) {
                const int pz = zr + iz - nfilt + 1, px = xr + ix - nfilt + 1;
                const double v = kern(wz - pz) * kern(wx - px);
                id[n2 * ix + iz] = px * nz + pz;
                ww[n2 * ix + iz] = v;
                sum += v;
            }
        for (int f = 0; f < ip.nf; ++f) ww[f] /= sum;
    }
    std::map<int, int> seen;
    ip.map.resize(ip.idx.size());
    for (size_t k = 0; k < ip.idx.size(); ++k) {
        auto it = seen.find(ip.idx[k]);
        if (it == seen.end()) {
            const int u = int(ip.uniq.size());
            seen[ip.idx[k]] = u;
            ip.uniq.push_back(ip.idx[k]);
            ip.map[k] = u;
        } else {
            ip.map[k] = it->second;
        }
    }
    return ip;
}
// forward: regular [nUniq][nt] -> irregular [nIrreg][nt] (accumulates)
template <typename T>
void interp_fwd_add(const Interp& ip, int nt, const T* reg, T* irr) {
    for (int d = 0; d < ip.nIrreg; ++d)
        for (int f = 0; f < ip.nf; ++f) {
            const T w = T(ip.w[size_t(ip.nf) * d + f]);
            const T* r = reg + size_t(ip.map[size_t(ip.nf) * d + f]) * nt;
            T* o = irr + size_t(d) * nt;
            for (int it = 0; it < nt; ++it) o[it] += w * r[it];
        }
}
// adjoint: irregular -> regular (overwrites)
template <typename T>
void interp_adj(const Interp& ip, int nt, T* reg, const T* irr) {
    std::fill(reg, reg + ip.uniq.size() * size_t(nt), T(0));
    for (int d = 0; d < ip.nIrreg; ++d)
        for (int f = 0; f < ip.nf; ++f) {
            const T w = T(ip.w[size_t(ip.nf) * d + f]);
            T* r = reg + size_t(ip.map[size_t(ip.nf) * d + f]) * nt;
            const T* in = irr + size_t(d) * nt;
            for (int it = 0; it < nt; ++it) r[it] += w * in[it];
        }
}

// interpTimeLinTbb.cpp:30-58, coarse [nDev][nts] -> fine [nDev][ntw]
template <typename T>
void time_interp_fwd(int nts, double dts, int sub, int nDev, const T* coarse, T* fine) {
    const int ntw = (nts - 1) * sub + 1;
    const double dtw = dts / sub;
    const double scale = 1.0 / std::sqrt(double(ntw) / double(nts));
    for (int d = 0; d < nDev; ++d) {
        const T* m = coarse + size_t(d) * nts;
        T* o = fine + size_t(d) * ntw;
        for (int itw = 0; itw < ntw - 1; ++itw) {
            // exact rational position (the reference computes it in float; for the
            // time axes used here both give the same floor)
            const int lo = itw / sub;
            const double w = double(itw - lo * sub) / double(sub);
            (void)dtw;
            o[itw] = T((double(m[lo]) * (1.0 - w) + double(m[lo + 1]) * w) * scale);
        }
        o[ntw - 1] = T(scale * double(m[nts - 1]));
    }
}

// The four staggered grid origins (Elastic_iso_float_prop.py:224-249):
//   centre (oz, ox), X (oz, ox-dx/2), Z (oz-dz/2, ox), XZ (oz-dz/2, ox-dx/2)
struct ShotGeom {
    Interp g[4];
};
ShotGeom make_geom(const Grid& gr, int nDev, const double* zc, const double* xc) {
    ShotGeom s;
    const double oz[4] = {gr.oz, gr.oz, gr.oz - 0.5 * gr.dz, gr.oz - 0.5 * gr.dz};
    const double ox[4] = {gr.ox, gr.ox - 0.5 * gr.dx, gr.ox, gr.ox - 0.5 * gr.dx};
    for (int k = 0; k < 4; ++k) s.g[k] = make_interp(gr.nz, oz[k], gr.dz, ox[k], gr.dx, nDev, zc, xc);
    return s;
}
RegGeom reg_of(const ShotGeom& s) {
    RegGeom r;
    for (int k = 0; k < 4; ++k) { r.pos[k] = s.g[k].uniq.data(); r.n[k] = int(s.g[k].uniq.size()); }
    return r;
}

// Source preparation shared by NL and Born (nonlinearPropElasticGpu.cpp:109-144,
// seismicElasticOperator2D.cpp:80-107): irregular [5][nSrc][nts] -> regular traces,
// scaled by 2dtw (stresses) or rho{x,z}Dtw at the source point, and by 1/(dx dz).
template <typename T>
void prep_sources(const Grid& gr, const ShotGeom& sg, const T* media, const T* model,
                  int nSrc, std::vector<T> (&reg)[5]) {
    const size_t n = size_t(gr.nz) * gr.nx;
    const T two_dtw = T(2.0 * (gr.dts / double(gr.sub)));
    const T area = T(1.0 / (gr.dx * gr.dz));
    for (int k = 0; k < 5; ++k) {
        const Interp& ip = sg.g[COMP_GRID[k]];
        reg[k].assign(ip.uniq.size() * size_t(gr.nts), T(0));
        interp_adj(ip, gr.nts, reg[k].data(), model + size_t(k) * nSrc * gr.nts);
        for (size_t u = 0; u < ip.uniq.size(); ++u) {
            T s;
            if (k == 0) s = media[ip.uniq[u]];
            else if (k == 1) s = media[n + ip.uniq[u]];
            else s = two_dtw;
            for (int it = 0; it < gr.nts; ++it) reg[k][u * gr.nts + it] *= s * area;
        }
    }
}

template <typename T>
struct MediaBuf {
    std::vector<T> buf;
    Media<T> m;
    MediaBuf(const Grid& g, const T* elastic) {
        const size_t n = size_t(g.nz) * g.nx;
        buf.resize(5 * n);
        media_scaled(g, elastic, buf.data());
        m = {buf.data(), buf.data() + n, buf.data() + 2 * n, buf.data() + 3 * n, buf.data() + 4 * n};
    }
};

// ---------------------------------------------------------------- shot-level pipelines
// nonlinearPropElasticGpu.cpp:43-246 (one shot, add=false)
template <typename T>
void nl_fwd_shot(const Grid& gr, const T* elastic, int nSrc, const double* sz, const double* sx,
                 int nRec, const double* rz, const double* rx, const T* model, T* data) {
    Consts<T> c = make_consts<T>(gr);
    MediaBuf<T> mb(gr, elastic);
    ShotGeom sg = make_geom(gr, nSrc, sz, sx), rg = make_geom(gr, nRec, rz, rx);
    std::vector<T> sreg[5];
    prep_sources(gr, sg, mb.buf.data(), model, nSrc, sreg);
    std::vector<T> dreg[5];
    const T* mptr[5];
    T* dptr[5];
    for (int k = 0; k < 5; ++k) {
        dreg[k].assign(rg.g[COMP_GRID[k]].uniq.size() * size_t(gr.nts), T(0));
        mptr[k] = sreg[k].data();
        dptr[k] = dreg[k].data();
    }
    nl_fwd_reg(c, mb.m, reg_of(sg), mptr, reg_of(rg), dptr);
    std::fill(data, data + size_t(5) * nRec * gr.nts, T(0));
    for (int k = 0; k < 5; ++k)
        interp_fwd_add(rg.g[COMP_GRID[k]], gr.nts, dreg[k].data(), data + size_t(k) * nRec * gr.nts);
}

// nonlinearPropElasticGpu.cpp:248-473 (one shot, add=false)
template <typename T>
void nl_adj_shot(const Grid& gr, const T* elastic, int nSrc, const double* sz, const double* sx,
                 int nRec, const double* rz, const double* rx, T* model, const T* data) {
    Consts<T> c = make_consts<T>(gr);
    MediaBuf<T> mb(gr, elastic);
    const size_t n = size_t(gr.nz) * gr.nx;
    ShotGeom sg = make_geom(gr, nSrc, sz, sx), rg = make_geom(gr, nRec, rz, rx);
    std::vector<T> dreg[5], mreg[5];
    const T* dptr[5];
    T* mptr[5];
    for (int k = 0; k < 5; ++k) {
        const Interp& ri = rg.g[COMP_GRID[k]];
        dreg[k].assign(ri.uniq.size() * size_t(gr.nts), T(0));
        interp_adj(ri, gr.nts, dreg[k].data(), data + size_t(k) * nRec * gr.nts);
        mreg[k].assign(sg.g[COMP_GRID[k]].uniq.size() * size_t(gr.nts), T(0));
        dptr[k] = dreg[k].data();
        mptr[k] = mreg[k].data();
    }
    nl_adj_reg(c, mb.m, reg_of(sg), mptr, reg_of(rg), dptr);
    const T two_dtw = T(2.0 * (gr.dts / double(gr.sub)));
    const T area = T(1.0 / (gr.dx * gr.dz));
    std::fill(model, model + size_t(5) * nSrc * gr.nts, T(0));
    for (int k = 0; k < 5; ++k) {
        const Interp& ip = sg.g[COMP_GRID[k]];
        for (size_t u = 0; u < ip.uniq.size(); ++u) {
            T s;
            if (k == 0) s = mb.buf[ip.uniq[u]];
            else if (k == 1) s = mb.buf[n + ip.uniq[u]];
            else s = two_dtw;
            for (int it = 0; it < gr.nts; ++it) mreg[k][u * gr.nts + it] *= area * s;
        }
        interp_fwd_add(ip, gr.nts, mreg[k].data(), model + size_t(k) * nSrc * gr.nts);
    }
}

// Born sources: prep + fine-sample (seismicElasticOperator2D.cpp:27-119)
template <typename T>
void born_sources(const Grid& gr, const ShotGeom& sg, const T* media, const T* srcSig, int nSrc,
                  std::vector<T> (&fine)[5]) {
    std::vector<T> reg[5];
    prep_sources(gr, sg, media, srcSig, nSrc, reg);
    const int ntw = (gr.nts - 1) * gr.sub + 1;
    for (int k = 0; k < 5; ++k) {
        const int nu = int(sg.g[COMP_GRID[k]].uniq.size());
        fine[k].assign(size_t(nu) * ntw, T(0));
        time_interp_fwd(gr.nts, gr.dts, gr.sub, nu, reg[k].data(), fine[k].data());
    }
}

// BornElasticShotsGpu.cpp:107-150: dm [3][nx][nz] (drho, dlame, dmu) -> 5 staggered, scaled
template <typename T>
void born_model_to_slices(const Grid& gr, const T* media, const T* dm, std::vector<T>& sl) {
    const size_t n = size_t(gr.nz) * gr.nx;
    sl.assign(5 * n, T(0));
    std::vector<T> tmp(n);
    stagger_x_adj(gr.nz, gr.nx, dm, sl.data());
    stagger_z_adj(gr.nz, gr.nx, dm, sl.data() + n);
    std::copy(dm + n, dm + 2 * n, sl.data() + 2 * n);
    std::copy(dm + 2 * n, dm + 3 * n, sl.data() + 3 * n);
    stagger_x_adj(gr.nz, gr.nx, dm + 2 * n, tmp.data());
    stagger_z_adj(gr.nz, gr.nx, tmp.data(), sl.data() + 4 * n);
    const T two_dtw = T(2.0 * (gr.dts / double(gr.sub)));
    for (size_t i = 0; i < n; ++i) {
        sl[i] *= media[i];
        sl[n + i] *= media[n + i];
        sl[2 * n + i] *= two_dtw;
        sl[3 * n + i] *= two_dtw;
        sl[4 * n + i] *= two_dtw;
    }
}

// BornElasticGpu.cpp:47-124 + BornElasticShotsGpu.cpp (one shot, add=false)
template <typename T>
void born_fwd_shot(const Grid& gr, const T* elastic, int nSrc, const double* sz, const double* sx,
                   const T* srcSig, int nRec, const double* rz, const double* rx, const T* dm,
                   T* data) {
    Consts<T> c = make_consts<T>(gr);
    MediaBuf<T> mb(gr, elastic);
    const size_t n = size_t(gr.nz) * gr.nx;
    ShotGeom sg = make_geom(gr, nSrc, sz, sx), rg = make_geom(gr, nRec, rz, rx);
    std::vector<T> fine[5];
    born_sources(gr, sg, mb.buf.data(), srcSig, nSrc, fine);
    std::vector<T> sl;
    born_model_to_slices(gr, mb.buf.data(), dm, sl);
    std::vector<T> dreg[5];
    const T* fptr[5];
    const T* dmp[5];
    T* dptr[5];
    for (int k = 0; k < 5; ++k) {
        dreg[k].assign(rg.g[COMP_GRID[k]].uniq.size() * size_t(gr.nts), T(0));
        fptr[k] = fine[k].data();
        dmp[k] = sl.data() + k * n;
        dptr[k] = dreg[k].data();
    }
    born_fwd_reg(c, mb.m, reg_of(sg), fptr, dmp, reg_of(rg), dptr);
    std::fill(data, data + size_t(5) * nRec * gr.nts, T(0));
    for (int k = 0; k < 5; ++k)
        interp_fwd_add(rg.g[COMP_GRID[k]], gr.nts, dreg[k].data(), data + size_t(k) * nRec * gr.nts);
}

// One shot of the Born adjoint, accumulating into the staggered/scaled-space slices acc[5][n]
template <typename T>
void born_adj_shot_acc(const Grid& gr, const Consts<T>& c, const MediaBuf<T>& mb, int nSrc,
                       const double* sz, const double* sx, const T* srcSig, int nRec,
                       const double* rz, const double* rx, const T* data, T* acc) {
    const size_t n = size_t(gr.nz) * gr.nx;
    ShotGeom sg = make_geom(gr, nSrc, sz, sx), rg = make_geom(gr, nRec, rz, rx);
    std::vector<T> fine[5];
    born_sources(gr, sg, mb.buf.data(), srcSig, nSrc, fine);
    std::vector<T> dreg[5];
    const T* fptr[5];
    const T* dptr[5];
    T* dmp[5];
    for (int k = 0; k < 5; ++k) {
        const Interp& ri = rg.g[COMP_GRID[k]];
        dreg[k].assign(ri.uniq.size() * size_t(gr.nts), T(0));
        interp_adj(ri, gr.nts, dreg[k].data(), data + size_t(k) * nRec * gr.nts);
        fptr[k] = fine[k].data();
        dptr[k] = dreg[k].data();
        dmp[k] = acc + k * n;
    }
    born_adj_reg(c, mb.m, reg_of(sg), fptr, dmp, reg_of(rg), dptr);
}

// BornElasticShotsGpu.cpp:325-371: scale stacked slices and unstagger into dm [3][nx][nz]
template <typename T>
void born_slices_to_model(const Grid& gr, const T* media, std::vector<T>& acc, T* dm) {
    const size_t n = size_t(gr.nz) * gr.nx;
    const T two_dtw = T(2.0 * (gr.dts / double(gr.sub)));
    for (size_t i = 0; i < n; ++i) {
        acc[i] *= media[i];
        acc[n + i] *= media[n + i];
        acc[2 * n + i] *= two_dtw;
        acc[3 * n + i] *= two_dtw;
        acc[4 * n + i] *= two_dtw;
    }
    std::fill(dm, dm + 3 * n, T(0));
    stagger_x_fwd_add(gr.nz, gr.nx, acc.data(), dm);
    stagger_z_fwd_add(gr.nz, gr.nx, acc.data() + n, dm);
    std::vector<T> t1(n, T(0)), t2(n, T(0));
    stagger_z_fwd_add(gr.nz, gr.nx, acc.data() + 4 * n, t1.data());
    stagger_x_fwd_add(gr.nz, gr.nx, t1.data(), t2.data());
    for (size_t i = 0; i < n; ++i) {
        dm[n + i] += acc[2 * n + i];
        dm[2 * n + i] += acc[3 * n + i] + t2[i];
    }
}

}  // namespace

// ---------------------------------------------------------------- C ABI for ctypes
extern "C" {

struct OracleGrid {
    int nz, nx, nts, sub, minPad;
    double dz, dx, dts, alphaCos, oz, ox;
    int surface;
};

static Grid to_grid(const OracleGrid* g) {
    Grid r;
    r.nz = g->nz; r.nx = g->nx; r.nts = g->nts; r.sub = g->sub; r.minPad = g->minPad;
    r.dz = g->dz; r.dx = g->dx; r.dts = g->dts; r.alphaCos = g->alphaCos; r.oz = g->oz; r.ox = g->ox;
    r.surface = g->surface;
    return r;
}

#define EIO_ORACLE_TYPED(SUF, T)                                                                   \
    void oracle_media_##SUF(const OracleGrid* g, const T* elastic, T* out) {                       \
        media_scaled<T>(to_grid(g), elastic, out);                                                \
    }                                                                                              \
    void oracle_consts_##SUF(const OracleGrid* g, T* coef8, T* filt, T* damp) {                    \
        Consts<T> c = make_consts<T>(to_grid(g));                                                  \
        for (int k = 0; k < 4; ++k) { coef8[k] = c.cz[k]; coef8[4 + k] = c.cx[k]; }                \
        std::copy(c.F.begin(), c.F.end(), filt);                                                   \
        for (int j = 0; j < g->minPad; ++j) damp[j] = c.damp[j];                                   \
    }                                                                                              \
    /* regular-level NL: media [5][n], positions per grid, traces packed per component */          \
    void oracle_nl_fwd_reg_##SUF(const OracleGrid* g, const T* media, const int* const* spos,      \
                                 const int* sn, const T* const* mod, const int* const* rpos,       \
                                 const int* rn, T* const* dat) {                                   \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        Media<T> m = {media, media + n, media + 2 * n, media + 3 * n, media + 4 * n};              \
        RegGeom s, r;                                                                              \
        for (int k = 0; k < 4; ++k) { s.pos[k] = spos[k]; s.n[k] = sn[k]; r.pos[k] = rpos[k]; r.n[k] = rn[k]; } \
        nl_fwd_reg(c, m, s, mod, r, dat);                                                          \
    }                                                                                              \
    void oracle_nl_fwd_reg_wfld_##SUF(const OracleGrid* g, const T* media, const int* const* spos, \
                                      const int* sn, const T* const* mod, const int* const* rpos,  \
                                      const int* rn, T* const* dat, T* wfld) {                     \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        Media<T> m = {media, media + n, media + 2 * n, media + 3 * n, media + 4 * n};              \
        RegGeom s, r;                                                                              \
        for (int k = 0; k < 4; ++k) { s.pos[k] = spos[k]; s.n[k] = sn[k]; r.pos[k] = rpos[k]; r.n[k] = rn[k]; } \
        std::fill(wfld, wfld + size_t(gr.nts) * 5 * n, T(0));                                      \
        nl_fwd_reg(c, m, s, mod, r, dat, wfld);                                                    \
    }                                                                                              \
    void oracle_nl_adj_reg_wfld_##SUF(const OracleGrid* g, const T* media, const int* const* spos, \
                                      const int* sn, T* const* mod, const int* const* rpos,        \
                                      const int* rn, const T* const* dat, T* wfld) {               \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        Media<T> m = {media, media + n, media + 2 * n, media + 3 * n, media + 4 * n};              \
        RegGeom s, r;                                                                              \
        for (int k = 0; k < 4; ++k) { s.pos[k] = spos[k]; s.n[k] = sn[k]; r.pos[k] = rpos[k]; r.n[k] = rn[k]; } \
        std::fill(wfld, wfld + size_t(gr.nts) * 5 * n, T(0));                                      \
        nl_adj_reg(c, m, s, mod, r, dat, wfld);                                                    \
    }                                                                                              \
    void oracle_nl_adj_reg_##SUF(const OracleGrid* g, const T* media, const int* const* spos,      \
                                 const int* sn, T* const* mod, const int* const* rpos,             \
                                 const int* rn, const T* const* dat) {                             \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        Media<T> m = {media, media + n, media + 2 * n, media + 3 * n, media + 4 * n};              \
        RegGeom s, r;                                                                              \
        for (int k = 0; k < 4; ++k) { s.pos[k] = spos[k]; s.n[k] = sn[k]; r.pos[k] = rpos[k]; r.n[k] = rn[k]; } \
        nl_adj_reg(c, m, s, mod, r, dat);                                                          \
    }                                                                                              \
    /* regular-level Born: srcFine[5] ([nReg][ntw]), dm[5][n] staggered+scaled */                   \
    void oracle_born_fwd_reg_##SUF(const OracleGrid* g, const T* media, const int* const* spos,    \
                                   const int* sn, const T* const* srcFine, const T* dm5,           \
                                   const int* const* rpos, const int* rn, T* const* dat) {         \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        Media<T> m = {media, media + n, media + 2 * n, media + 3 * n, media + 4 * n};              \
        RegGeom s, r;                                                                              \
        for (int k = 0; k < 4; ++k) { s.pos[k] = spos[k]; s.n[k] = sn[k]; r.pos[k] = rpos[k]; r.n[k] = rn[k]; } \
        const T* dmp[5] = {dm5, dm5 + n, dm5 + 2 * n, dm5 + 3 * n, dm5 + 4 * n};                   \
        born_fwd_reg(c, m, s, srcFine, dmp, r, dat);                                               \
    }                                                                                              \
    void oracle_born_adj_reg_##SUF(const OracleGrid* g, const T* media, const int* const* spos,    \
                                   const int* sn, const T* const* srcFine, T* dm5,                 \
                                   const int* const* rpos, const int* rn, const T* const* dat) {   \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        Media<T> m = {media, media + n, media + 2 * n, media + 3 * n, media + 4 * n};              \
        RegGeom s, r;                                                                              \
        for (int k = 0; k < 4; ++k) { s.pos[k] = spos[k]; s.n[k] = sn[k]; r.pos[k] = rpos[k]; r.n[k] = rn[k]; } \
        T* dmp[5] = {dm5, dm5 + n, dm5 + 2 * n, dm5 + 3 * n, dm5 + 4 * n};                         \
        born_adj_reg(c, m, s, srcFine, dmp, r, dat);                                               \
    }                                                                                              \
    /* shot-level (irregular geometry, elastic [3][nx][nz] = rho, lambda, mu) */                   \
    void oracle_nl_fwd_shot_##SUF(const OracleGrid* g, const T* elastic, int nSrc,                 \
                                  const double* sz, const double* sx, int nRec, const double* rz,  \
                                  const double* rx, const T* model, T* data) {                     \
        nl_fwd_shot<T>(to_grid(g), elastic, nSrc, sz, sx, nRec, rz, rx, model, data);              \
    }                                                                                              \
    void oracle_nl_adj_shot_##SUF(const OracleGrid* g, const T* elastic, int nSrc,                 \
                                  const double* sz, const double* sx, int nRec, const double* rz,  \
                                  const double* rx, T* model, const T* data) {                     \
        nl_adj_shot<T>(to_grid(g), elastic, nSrc, sz, sx, nRec, rz, rx, model, data);              \
    }                                                                                              \
    void oracle_born_fwd_shot_##SUF(const OracleGrid* g, const T* elastic, int nSrc,               \
                                    const double* sz, const double* sx, const T* srcSig, int nRec, \
                                    const double* rz, const double* rx, const T* dm, T* data) {    \
        born_fwd_shot<T>(to_grid(g), elastic, nSrc, sz, sx, srcSig, nRec, rz, rx, dm, data);       \
    }                                                                                              \
    /* Born adjoint over nShot shots (per-shot source x-coords sx[shot*nSrc + i]) -> dm [3][n] */   \
    void oracle_born_adj_shots_##SUF(const OracleGrid* g, const T* elastic, int nShot, int nSrc,   \
                                     const double* sz, const double* sx, const T* srcSig,          \
                                     int nRec, const double* rz, const double* rx, T* dm,          \
                                     const T* data) {                                              \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        MediaBuf<T> mb(gr, elastic);                                                               \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        std::vector<T> acc(5 * n, T(0));                                                           \
        for (int s = 0; s < nShot; ++s)                                                            \
            born_adj_shot_acc<T>(gr, c, mb, nSrc, sz + size_t(s) * nSrc, sx + size_t(s) * nSrc,    \
                                 srcSig, nRec, rz, rx, data + size_t(s) * 5 * nRec * gr.nts,       \
                                 acc.data());                                                      \
        born_slices_to_model<T>(gr, mb.buf.data(), acc, dm);                                       \
    }                                                                                              \
    void oracle_stagger_##SUF(int which, int nz, int nx, const T* in, T* out) {                    \
        const size_t n = size_t(nz) * nx;                                                          \
        if (which == 0) stagger_x_adj<T>(nz, nx, in, out);                                         \
        else if (which == 1) stagger_z_adj<T>(nz, nx, in, out);                                    \
        else { std::fill(out, out + n, T(0));                                                      \
               if (which == 2) stagger_x_fwd_add<T>(nz, nx, in, out);                              \
               else stagger_z_fwd_add<T>(nz, nx, in, out); }                                       \
    }                                                                                              \
    void oracle_time_interp_##SUF(int nts, double dts, int sub, int nDev, const T* coarse,         \
                                  T* fine) {                                                       \
        time_interp_fwd<T>(nts, dts, sub, nDev, coarse, fine);                                     \
    }                                                                                              \
    /* test hook: `nsteps` fine steps of the NL loop body without sources or receivers (step,      \
       damping of both levels, swap; nonlinearPropElasticGpuFunctions.cu:588-617) from the state   \
       pair state[0] = older, state[1] = current ([2][5][n]), in place */                          \
    void oracle_free_steps_##SUF(const OracleGrid* g, const T* media, T* state, int nsteps) {     \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        Media<T> m = {media, media + n, media + 2 * n, media + 3 * n, media + 4 * n};              \
        Fields<T> a(n), b(n);                                                                      \
        for (int k = 0; k < 5; ++k) {                                                              \
            std::copy(state + k * n, state + (k + 1) * n, a.f[k].begin());                         \
            std::copy(state + (5 + k) * n, state + (6 + k) * n, b.f[k].begin());                   \
        }                                                                                          \
        Fields<T>*p0 = &a, *p1 = &b;                                                               \
        for (int it = 0; it < nsteps; ++it) {                                                      \
            if (c.surface) step_fwd_fs(c, m, *p0, *p1);                                            \
            else step_fwd(c, m, *p0, *p1);                                                         \
            damp_both(c, *p0, *p1);                                                                \
            std::swap(p0, p1);                                                                     \
        }                                                                                          \
        for (int k = 0; k < 5; ++k) {                                                              \
            std::copy(p0->f[k].begin(), p0->f[k].end(), state + k * n);                            \
            std::copy(p1->f[k].begin(), p1->f[k].end(), state + (5 + k) * n);                      \
        }                                                                                          \
    }                                                                                              \
    /* test hook: the factor damp_both applies to each cell (fields of ones after it) */           \
    void oracle_damp_field_##SUF(const OracleGrid* g, T* out) {                                    \
        Grid gr = to_grid(g);                                                                      \
        Consts<T> c = make_consts<T>(gr);                                                          \
        const size_t n = size_t(gr.nz) * gr.nx;                                                    \
        Fields<T> a(n), b(n);                                                                      \
        for (int k = 0; k < 5; ++k) {                                                              \
            std::fill(a.f[k].begin(), a.f[k].end(), T(1));                                         \
            std::fill(b.f[k].begin(), b.f[k].end(), T(1));                                         \
        }                                                                                          \
        damp_both(c, a, b);                                                                        \
        std::copy(a.f[0].begin(), a.f[0].end(), out);                                              \
    }

EIO_ORACLE_TYPED(f64, double)
EIO_ORACLE_TYPED(f32, float)

// Flush-to-zero / denormals-are-zero on every OpenMP thread (the FP32 timing twin otherwise
// spends most of its time on denormals in the quiet part of the wavefield).
void oracle_set_ftz(int on) {
#pragma omp parallel
    {
        _MM_SET_FLUSH_ZERO_MODE(on ? _MM_FLUSH_ZERO_ON : _MM_FLUSH_ZERO_OFF);
        _MM_SET_DENORMALS_ZERO_MODE(on ? _MM_DENORMALS_ZERO_ON : _MM_DENORMALS_ZERO_OFF);
    }
    _MM_SET_FLUSH_ZERO_MODE(on ? _MM_FLUSH_ZERO_ON : _MM_FLUSH_ZERO_OFF);
    _MM_SET_DENORMALS_ZERO_MODE(on ? _MM_DENORMALS_ZERO_ON : _MM_DENORMALS_ZERO_OFF);
}

int oracle_num_threads(void) { return omp_get_max_threads(); }

// Spatial interpolation of one grid: returns the number of unique regular points;
// fills idx/w (4 per device) and uniq (capacity 4*nDev).
int oracle_interp(int nz, double oz, double dz, double ox, double dx, int nDev, const double* zc,
                  const double* xc, int* idx, double* w, int* uniq) {
    Interp ip = make_interp(nz, oz, dz, ox, dx, nDev, zc, xc);
    std::copy(ip.idx.begin(), ip.idx.end(), idx);
    std::copy(ip.w.begin(), ip.w.end(), w);
    std::copy(ip.uniq.begin(), ip.uniq.end(), uniq);
    return int(ip.uniq.size());
}
// Same for method 0 (bilinear), 1 (sinc), 2 (Gaussian) with nfilt: nf = 4 or (2 nfilt)^2
// entries per device in idx/w (capacity nf*nDev for uniq).
int oracle_interp_m(int nz, double oz, double dz, double ox, double dx, int nDev, const double* zc,
                    const double* xc, int method, int nfilt, int* idx, double* w, int* uniq) {
    Interp ip = make_interp(nz, oz, dz, ox, dx, nDev, zc, xc, method, nfilt);
    std::copy(ip.idx.begin(), ip.idx.end(), idx);
    std::copy(ip.w.begin(), ip.w.end(), w);
    std::copy(ip.uniq.begin(), ip.uniq.end(), uniq);
    return int(ip.uniq.size());
}
// Bilinear with a dipole (8 entries per device; shifts in cells).
int oracle_interp_dipole(int nz, double oz, double dz, double ox, double dx, int nDev, const double* zc,
                         const double* xc, double zsh, double xsh, int* idx, double* w, int* uniq) {
    Interp ip = make_interp(nz, oz, dz, ox, dx, nDev, zc, xc, 0, 1, 1, zsh, xsh);
    std::copy(ip.idx.begin(), ip.idx.end(), idx);
    std::copy(ip.w.begin(), ip.w.end(), w);
    std::copy(ip.uniq.begin(), ip.uniq.end(), uniq);
    return int(ip.uniq.size());
}

}  // extern "C"
