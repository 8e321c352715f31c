// host_ops.h -- C++ host layer of the drop-in operators (templated on the real type).
//
// Mirrors the reference's host classes on the path of the nonlinear and Born shot operators
// (paths relative to elastic_iso_lib/elastic_iso_float/prop):
//   FdParam           src/fdParamElastic.cpp:12-157, checks :256-324
//   stagger helpers   src/stagger.cpp:18-114
//   SpaceInterpT      src/spaceInterpGpu.cpp:11-138 (linear + dipole), :139-207 (sinc, gauss), :288-376,
//                     checks :378-403
//   time_interp_fwd   src/interpTimeLinTbb.cpp:30-58
//   NonlinearShots    src/nonlinearPropElasticShotsGpu.cpp:80-422 + src/nonlinearPropElasticGpu.cpp:43-473
//   BornShots         src/BornElasticShotsGpu.cpp:85-378 + src/BornElasticGpu.cpp:47-194 +
//                     include/seismicElasticOperator2D.cpp:27-119
// Propagation goes through the C ABI of include/eio.h (all shots of a call in one batched
// device time loop on one GPU).  Errors are thrown as std::runtime_error.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "eio.h"

namespace eioh {

constexpr int FAT = 4;
constexpr int SUB_MAX = 100;     // varDeclare.h:10
constexpr int GRID_OF_COMP[5] = {1, 2, 0, 0, 3};

struct Axis {
    int n = 1;
    double o = 0.0, d = 1.0;
};

// Parameter source (parfile / command-line key=value), with the reference defaults semantics:
// a missing key without default is an error.
struct Params {
    std::map<std::string, std::string> kv;
    bool has(const std::string& k) const { return kv.count(k) > 0; }
    int getInt(const std::string& k) const {
        if (!has(k)) throw std::runtime_error("missing parameter " + k);
        return std::stoi(kv.at(k));
    }
    int getInt(const std::string& k, int def) const { return has(k) ? std::stoi(kv.at(k)) : def; }
    double getFloat(const std::string& k) const {
        if (!has(k)) throw std::runtime_error("missing parameter " + k);
        return std::stod(kv.at(k));
    }
    double getFloat(const std::string& k, double def) const { return has(k) ? std::stod(kv.at(k)) : def; }
    std::string getString(const std::string& k, const std::string& def) const {
        return has(k) ? kv.at(k) : def;
    }
    std::vector<int> getInts(const std::string& k, const std::vector<int>& def) const {
        if (!has(k)) return def;
        std::vector<int> out;
        std::string s = kv.at(k);
        size_t p = 0;
        while (p <= s.size()) {
            size_t q = s.find(',', p);
            if (q == std::string::npos) q = s.size();
            if (q > p) out.push_back(std::stoi(s.substr(p, q - p)));
            p = q + 1;
        }
        return out;
    }
};

inline void check_eio(int rc) {
    if (rc != EIO_OK) throw std::runtime_error(std::string("eio: ") + eio_last_error());
}

// Page-locked allocator (eio_host_alloc) for the per-call trace buffers the library copies to
// or from the device: the copies run as direct DMA, not through a pageable staging pass.
template <typename T>
struct PinnedAlloc {
    using value_type = T;
    PinnedAlloc() = default;
    template <typename U>
    PinnedAlloc(const PinnedAlloc<U>&) {}
    T* allocate(size_t n) {
        void* p = nullptr;
        check_eio(eio_host_alloc((unsigned long long)n * sizeof(T), &p));
        if (!p && n) throw std::bad_alloc();
        return static_cast<T*>(p);
    }
    void deallocate(T* p, size_t) { (void)eio_host_free(p); }
    template <typename U>
    bool operator==(const PinnedAlloc<U>&) const { return true; }
    template <typename U>
    bool operator!=(const PinnedAlloc<U>&) const { return false; }
};
template <typename T>
using PinnedVec = std::vector<T, PinnedAlloc<T>>;

// ---------------------------------------------------------------- staggering (stagger.cpp)
template <typename T>
void stagger_x_adj(int nz, int nx, const T* d, T* m) {  // stagger.cpp:38-57
    for (int iz = 0; iz < nz; ++iz) {
        for (int ix = 1; ix < nx - 1; ++ix)
            m[size_t(ix) * nz + iz] = T(0.5) * (d[size_t(ix) * nz + iz] + d[size_t(ix - 1) * nz + iz]);
        m[iz] = T(0.5) * d[iz];
        m[size_t(nx - 1) * nz + iz] = d[size_t(nx - 1) * nz + iz] + T(0.5) * d[size_t(nx - 2) * nz + iz];
    }
}
template <typename T>
void stagger_z_adj(int nz, int nx, const T* d, T* m) {  // stagger.cpp:95-114
    for (int ix = 0; ix < nx; ++ix) {
        const T* a = d + size_t(ix) * nz;
        T* b = m + size_t(ix) * nz;
        for (int iz = 1; iz < nz - 1; ++iz) b[iz] = T(0.5) * (a[iz] + a[iz - 1]);
        b[0] = T(0.5) * a[0];
        b[nz - 1] = a[nz - 1] + T(0.5) * a[nz - 2];
    }
}
template <typename T>
void stagger_x_fwd_add(int nz, int nx, const T* m, T* d) {  // stagger.cpp:18-35
    for (int iz = 0; iz < nz; ++iz) {
        for (int ix = 0; ix < nx - 1; ++ix)
            d[size_t(ix) * nz + iz] += T(0.5) * (m[size_t(ix) * nz + iz] + m[size_t(ix + 1) * nz + iz]);
        d[size_t(nx - 1) * nz + iz] += m[size_t(nx - 1) * nz + iz];
    }
}
template <typename T>
void stagger_z_fwd_add(int nz, int nx, const T* m, T* d) {  // stagger.cpp:75-93
    for (int ix = 0; ix < nx; ++ix) {
        const T* a = m + size_t(ix) * nz;
        T* b = d + size_t(ix) * nz;
        for (int iz = 0; iz < nz - 1; ++iz) b[iz] += T(0.5) * (a[iz] + a[iz + 1]);
        b[nz - 1] += a[nz - 1];
    }
}

// ---------------------------------------------------------------- fdParamElastic
template <typename T>
struct FdParam {
    int nts, sub, ntw, nz, nx, zPadPlus, zPadMinus, xPadPlus, xPadMinus, zPad, xPad, minPad;
    int surfaceCondition, mod_par, blockSize, fat, saveWavefield;
    double dts, ots, dtw, otw, dz, dx, oz, ox, fMax, alphaCos, errorTolerance;
    float minVpVs, maxVpVs, Courant, dispersionRatio;
    std::vector<T> media;  // 5*nz*nx: rhoxDtw, rhozDtw, lamb2MuDtw, lambDtw, muxzDtw

    // elastic: [3][nx][nz] (rho, lambda, mu); axes of the elastic-parameter hypercube
    FdParam(const T* elastic, const std::vector<Axis>& ax, const Params& par) {
        surfaceCondition = par.getInt("surfaceCondition", 0);
        nts = par.getInt("nts");
        dts = par.getFloat("dts", 0.0);
        ots = par.getFloat("ots", 0.0);
        sub = par.getInt("sub");
        sub = sub * 2;  // fdParamElastic.cpp:28
        if (sub > SUB_MAX)
            throw std::runtime_error("**** ERROR: 2*_sub is greater than the allowed SUB_MAX value. " +
                                     std::to_string(sub) + " > " + std::to_string(SUB_MAX) + " ****");
        ntw = (nts - 1) * sub + 1;
        // fdParamElastic.cpp:34 in the library's real type (float: float dts / float sub)
        dtw = double(T(dts) / T(sub));
        otw = ots;
        nz = par.getInt("nz");
        zPadPlus = par.getInt("zPadPlus");
        zPadMinus = par.getInt("zPadMinus");
        if (surfaceCondition == 0) zPad = std::min(zPadMinus, zPadPlus);
        else if (surfaceCondition == 1) zPad = zPadPlus;
        else throw std::runtime_error("ERROR NO IMPROPER FREE SURFACE PARAM PROVIDED");
        mod_par = par.getInt("mod_par", 1);
        dz = par.getFloat("dz", -1.0);
        oz = ax.size() > 0 ? ax[0].o : 0.0;
        nx = par.getInt("nx");
        xPadPlus = par.getInt("xPadPlus");
        xPadMinus = par.getInt("xPadMinus");
        xPad = std::min(xPadMinus, xPadPlus);
        dx = par.getFloat("dx", -1.0);
        ox = ax.size() > 1 ? ax[1].o : 0.0;
        if (mod_par == 2) { dz *= 1000.; ox *= 1000.; dx *= 1000.; oz *= 1000.; }
        fMax = par.getFloat("fMax", 1000.0);
        blockSize = par.getInt("blockSize");
        fat = par.getInt("fat", 4);
        minPad = std::min(zPad, xPad);
        saveWavefield = par.getInt("saveWavefield", 0);
        alphaCos = par.getFloat("alphaCos", 0.99);
        errorTolerance = par.getFloat("errorTolerance", 0.000001);
        if (surfaceCondition != 0 && surfaceCondition != 1)
            throw std::runtime_error("ERROR NO IMPROPER FREE SURFACE PARAM PROVIDED");
        if (fat != FAT) throw std::runtime_error("fat must be 4");
        // consistency with the elastic parameter hypercube (fdParamElastic.cpp:296-324)
        const double scale = (mod_par == 2) ? 1000.0 : 1.0;
        if (ax.size() < 3 || ax[0].n != nz) throw std::runtime_error("**** ERROR: nz not consistent with parfile ****");
        if (std::abs(dz - ax[0].d * scale) > errorTolerance) throw std::runtime_error("**** ERROR: dz not consistent with parfile ****");
        if (ax[1].n != nx) throw std::runtime_error("**** ERROR nx not consistent with parfile ****");
        if (std::abs(dx - ax[1].d * scale) > errorTolerance) throw std::runtime_error("**** ERROR: dx not consistent with parfile ****");
        set_model(elastic);
    }

    // vmin/vmax scan (fdParamElastic.cpp:81-98, same loop bounds), checks, media scaling
    void set_model(const T* elastic) {
        const size_t n = size_t(nz) * nx;
        const T* rho = elastic;
        const T* lam = elastic + n;
        const T* mu = elastic + 2 * n;
        minVpVs = 10000.f;
        maxVpVs = -1.f;
        for (int ix = fat; ix < nx - 2 * fat; ix++)
            for (int iz = fat; iz < nz - 2 * fat; iz++) {
                const size_t i = size_t(ix) * nz + iz;
                const float vp = std::sqrt(float((lam[i] + 2 * mu[i]) / rho[i]));
                const float vs = std::sqrt(float(mu[i] / rho[i]));
                if (vp < minVpVs) minVpVs = vp;
                if (vp > maxVpVs) maxVpVs = vp;
                if (vs < minVpVs && vs != 0) minVpVs = vs;
                if (vs > maxVpVs) maxVpVs = vs;
            }
        // checkFdStability (fdParamElastic.cpp:256-267)
        Courant = float(maxVpVs * dtw * 2.0 / std::min(dz, dx));
        if (Courant > 0.45f)
            throw std::runtime_error("**** ERROR: Courant is too big: " + std::to_string(Courant) + " ****");
        // checkFdDispersion (:269-281)
        dispersionRatio = float(minVpVs / (fMax * std::max(dz, dx)));
        if (dispersionRatio < 3.0f)
            throw std::runtime_error("**** ERROR: Dispersion is too small: " + std::to_string(dispersionRatio) + " > 3 ****");
        // checkModelSize (:283-294)
        if ((nz - 2 * fat) % blockSize != 0) throw std::runtime_error("**** ERROR: nz-2*_fat not a multiple of block size ****");
        if ((nx - 2 * fat) % blockSize != 0) throw std::runtime_error("**** ERROR: nx-2*_fat not a multiple of block size ****");
        if (blockSize != 16) throw std::runtime_error("**** ERROR: Block size for time stepper is not consistent with parfile ****");
        if (minPad >= 500) throw std::runtime_error("**** ERROR: Padding value is too high ****");
        // staggered and scaled media (fdParamElastic.cpp:118-148)
        media.assign(5 * n, T(0));
        std::vector<T> tmp(n), muxz(n);
        stagger_x_adj(nz, nx, rho, media.data());
        stagger_z_adj(nz, nx, rho, media.data() + n);
        stagger_x_adj(nz, nx, mu, tmp.data());
        stagger_z_adj(nz, nx, tmp.data(), muxz.data());
        const T two = T(2.0 * dtw);
        for (size_t i = 0; i < n; ++i) {
            media[i] = two / media[i];
            media[n + i] = two / media[n + i];
            media[2 * n + i] = two * (lam[i] + T(2) * mu[i]);
            media[3 * n + i] = two * lam[i];
            media[4 * n + i] = two * muxz[i];
        }
    }

    size_t ncell() const { return size_t(nz) * nx; }
    void getInfo() const {  // fdParamElastic.cpp:159-254 (abridged)
        std::cerr << "nts = " << nts << ", dts = " << dts << ", sub (fine) = " << sub << ", ntw = " << ntw
                  << ", dtw = " << dtw << "\nnz = " << nz << ", nx = " << nx << ", dz = " << dz << ", dx = " << dx
                  << "\nCourant = " << Courant << ", dispersion = " << dispersionRatio << ", vmin = " << minVpVs
                  << ", vmax = " << maxVpVs << ", minPad = " << minPad << std::endl;
    }
};

// ---------------------------------------------------------------- spaceInterpGpu
// R: the library's real type, in which coordinates and weights are computed (the FP32 library
// in float, its FP64 copy in double: elastic_iso_double/prop/src/spaceInterpGpu.cpp:83-84).
template <typename R>
struct SpaceInterpT {
    std::vector<R> zCoord, xCoord;
    std::vector<Axis> hyper;
    int nt = 0, nFilt = 1, nFiltTotal = 4, dipole = 0, nDeviceIrreg = 0, nz = 0;
    R zDipoleShift = 0, xDipoleShift = 0;
    std::string method;
    std::vector<int> gridPointIndex;       // nFiltTotal per device
    std::vector<R> weight;
    std::vector<int> uniq;                 // unique regular points, first-seen order
    std::vector<int> map;                  // per (device, filter) -> unique index

    SpaceInterpT(std::vector<R> z, std::vector<R> x, std::vector<Axis> h, int nt_, std::string m,
                 int nf, int dip, R zs, R xs)
        : zCoord(std::move(z)), xCoord(std::move(x)), hyper(std::move(h)), nt(nt_), method(std::move(m)) {
        if (zCoord.size() != xCoord.size()) throw std::runtime_error("zCoord and xCoord sizes differ");
        if (hyper.size() < 2) throw std::runtime_error("spaceInterpGpu needs a 2-D (or 3-D) hypercube");
        if (method != "linear" && method != "sinc" && method != "gauss")
            throw std::runtime_error("**** ERROR: Space interp method not defined ****");
        if (method != "linear" && nf < 1) throw std::runtime_error("**** ERROR: nFilt must be >= 1 ****");
        nFilt = method == "linear" ? 1 : nf;
        dipole = dip;
        zDipoleShift = R(zs * hyper[0].d);
        xDipoleShift = R(xs * hyper[1].d);
        nDeviceIrreg = int(zCoord.size());
        nz = hyper[0].n;
        check_out_of_bounds();
        if (method == "linear") {
            nFiltTotal = dipole == 1 ? 8 : 4;
            calc_linear_weights();
        } else {
            if (dipole == 1)
                throw std::runtime_error(method == "sinc"
                    ? "**** ERROR: Spatial sinc interpolation and dipole acquisition not implemented ****"
                    : "**** ERROR: Gaussian spatial interpolation and dipole acquisition not implemented ****");
            calc_filter_weights(method == "sinc");
        }
        convert_irreg_to_reg();
    }

    void check_out_of_bounds() const {  // spaceInterpGpu.cpp:378-403
        const R zMin = R(hyper[0].o), xMin = R(hyper[1].o);
        const R zMax = zMin + (hyper[0].n - 1) * R(hyper[0].d);
        const R xMax = xMin + (hyper[1].n - 1) * R(hyper[1].d);
        const R zb = R(hyper[0].d) * nFilt, xb = R(hyper[1].d) * nFilt;
        for (int i = 0; i < nDeviceIrreg; ++i) {
            if (zCoord[i] >= zMax - zb || zCoord[i] <= zMin + zb)
                throw std::runtime_error("**** ERROR: One of the device is out of bounds in the z direction ****");
            if (xCoord[i] >= xMax - xb || xCoord[i] <= xMin + xb)
                throw std::runtime_error("**** ERROR: One of the device is out of bounds in the x direction ****");
        }
    }

    void calc_linear_weights() {  // spaceInterpGpu.cpp:78-138
        gridPointIndex.assign(size_t(nFiltTotal) * nDeviceIrreg, 0);
        weight.assign(size_t(nFiltTotal) * nDeviceIrreg, R(0));
        const R oz = R(hyper[0].o), dz = R(hyper[0].d), ox = R(hyper[1].o), dx = R(hyper[1].d);
        const R one = R(1);
        for (int d = 0; d < nDeviceIrreg; ++d) {
            const int i1 = d * nFiltTotal;
            for (int pole = 0; pole < (dipole == 1 ? 2 : 1); ++pole) {
                R wz = (zCoord[d] + (pole ? zDipoleShift : R(0)) - oz) / dz;
                R wx = (xCoord[d] + (pole ? xDipoleShift : R(0)) - ox) / dx;
                const int zReg = int(wz), xReg = int(wx);
                wz = one - (wz - zReg);
                wx = one - (wx - xReg);
                const R sgn = pole ? -one : one;
                const int b = i1 + 4 * pole;
                gridPointIndex[b] = xReg * nz + zReg;
                weight[b] = sgn * wz * wx;
                gridPointIndex[b + 1] = gridPointIndex[b] + 1;
                weight[b + 1] = sgn * (one - wz) * wx;
                gridPointIndex[b + 2] = gridPointIndex[b] + nz;
                weight[b + 2] = sgn * wz * (one - wx);
                gridPointIndex[b + 3] = gridPointIndex[b] + nz + 1;
                weight[b + 3] = sgn * (one - wz) * (one - wx);
            }
        }
    }

    // sin(x)/x with a Taylor guard near 0 (the series 1 - x^2/6 + x^4/120 where sin(x)/x
    // would lose digits), evaluated in double.
    static double sinc_pi(double x) {
        const double ax = std::fabs(x);
        if (ax >= 1.2207e-4) return std::sin(x) / x;
        const double x2 = x * x;
        return 1.0 - x2 / 6.0 + x2 * x2 / 120.0;
    }
    static double normal_pdf(double x, double mean, double sd) {
        const double e = (x - mean) * (x - mean) / (2.0 * sd * sd);
        return std::exp(-e) / (sd * 2.5066282746310002);  // sqrt(2 pi)
    }

    // spaceInterpGpu.cpp:139-207: per device the (2 nFilt)^2 regular points whose indices run
    // from int(position) - nFilt + 1 to int(position) + nFilt, x-major, with separable sinc
    // weights (sinc(pi dz_cells) sinc(pi dx_cells)) or the product of two normal densities of
    // standard deviation 2 cells, normalised to sum 1.  Coordinates are in R and the
    // kernels double, then rounded to R, as there.  The arrays hold the whole stencil:
    // the reference's constructor resets its per-device count to 4 after choosing nFilt
    // (spaceInterpGpu.cpp:38-41), which is consistent only for nFilt = 1.
    void calc_filter_weights(bool sinc) {
        const int n2 = 2 * nFilt;
        nFiltTotal = n2 * n2;
        gridPointIndex.assign(size_t(nFiltTotal) * nDeviceIrreg, 0);
        weight.assign(size_t(nFiltTotal) * nDeviceIrreg, R(0));
        const R oz = R(hyper[0].o), dz = R(hyper[0].d), ox = R(hyper[1].o), dx = R(hyper[1].d);
        for (int d = 0; d < nDeviceIrreg; ++d) {
            const size_t off = size_t(d) * nFiltTotal;
            const R zf = zCoord[d], xf = xCoord[d];
            const int zr = int((zf - oz) / dz), xr = int((xf - ox) / dx);
            R sum = 0;
            for (int ix = 0; ix < n2; ++ix)
                for (int iz = 0; iz < n2; ++iz) {
                    const int px = xr + ix - nFilt + 1, pz = zr + iz - nFilt + 1;
                    const R xcur = R(px) * dx + ox, zcur = R(pz) * dz + oz;
                    double v;
                    if (sinc) {
                        const R xin = (xf - xcur) / dx, zin = (zf - zcur) / dz;
                        v = sinc_pi(M_PI * zin) * sinc_pi(M_PI * xin);
                    } else {
                        v = normal_pdf(xcur, xf, double(R(2) * dx)) * normal_pdf(zcur, zf, double(R(2) * dz));
                    }
                    const size_t k = off + size_t(n2) * ix + iz;
                    weight[k] = R(v);
                    gridPointIndex[k] = nz * px + pz;
                    sum += weight[k];
                }
            for (int i = 0; i < nFiltTotal; ++i) weight[off + i] = weight[off + i] / sum;
        }
    }

    void convert_irreg_to_reg() {  // spaceInterpGpu.cpp:288-311
        std::map<int, int> seen;
        uniq.clear();
        map.assign(gridPointIndex.size(), 0);
        for (size_t k = 0; k < gridPointIndex.size(); ++k) {
            auto it = seen.find(gridPointIndex[k]);
            if (it == seen.end()) {
                seen[gridPointIndex[k]] = int(uniq.size());
                map[k] = int(uniq.size());
                uniq.push_back(gridPointIndex[k]);
            } else {
                map[k] = it->second;
            }
        }
    }

    int nReg() const { return int(uniq.size()); }

    // regular [nReg][nt] -> irregular [nIrreg][nt] (accumulates), spaceInterpGpu.cpp:313-340
    template <typename T>
    void forward_add(const T* reg, T* irr) const {
        for (int d = 0; d < nDeviceIrreg; ++d)
            for (int f = 0; f < nFiltTotal; ++f) {
                const size_t k = size_t(d) * nFiltTotal + f;
                const T w = T(weight[k]);
                const T* r = reg + size_t(map[k]) * nt;
                T* o = irr + size_t(d) * nt;
                for (int it = 0; it < nt; ++it) o[it] += w * r[it];
            }
    }
    // irregular -> regular (overwrites), spaceInterpGpu.cpp:342-376
    template <typename T>
    void adjoint(T* reg, const T* irr) const {
        std::fill(reg, reg + uniq.size() * size_t(nt), T(0));
        for (int d = 0; d < nDeviceIrreg; ++d)
            for (int f = 0; f < nFiltTotal; ++f) {
                const size_t k = size_t(d) * nFiltTotal + f;
                const T w = T(weight[k]);
                T* r = reg + size_t(map[k]) * nt;
                const T* in = irr + size_t(d) * nt;
                for (int it = 0; it < nt; ++it) r[it] += w * in[it];
            }
    }
    void getInfo() const {
        std::cerr << "---------- Space Interp Info ----------\n_nDeviceIrreg= " << nDeviceIrreg
                  << "\n_nDeviceRreg= " << uniq.size() << "\n";
    }
};

// ---------------------------------------------------------------- interpTimeLinTbb::forward
// Times, weights and scale in T (the FP32 library in float, its FP64 copy in double:
// elastic_iso_double/prop/src/interpTimeLinTbb.cpp:49-55).
template <typename T>
void time_interp_fwd(int nts, double dts, double ots, int sub, int nDev, const T* coarse, T* fine) {
    const int ntw = (nts - 1) * sub + 1;
    const T dtw = T(dts) / T(sub);
    T scale = T(ntw) / T(nts);
    scale = T(1.0 / std::sqrt(double(scale)));
    for (int d = 0; d < nDev; ++d) {
        const T* m = coarse + size_t(d) * nts;
        T* o = fine + size_t(d) * ntw;
        for (int itw = 0; itw < ntw - 1; ++itw) {
            const T tw = T(ots) + itw * dtw;
            T w = (tw - T(ots)) / T(dts);
            const int lo = int(w);
            w = w - lo;
            o[itw] = (m[lo] * T(1.0 - w) + m[lo + 1] * T(w)) * scale;
        }
        o[ntw - 1] = scale * m[nts - 1];
    }
}

// ---------------------------------------------------------------- device context
struct Device {
    eio_ctx* ctx = nullptr;
    int id = 0;
    Device(int device, int prec) {
        id = device;
        check_eio(eio_ctx_create(device, prec, &ctx));
        if (prec == 4) {
            // FP32: two fine steps per launch for the forward and the adjoint passes, tile
            // width chosen per call by the batch size; passes these kernels do not cover
            // (wavefield saving, free surface, odd fine sub-sampling) fall back to one step
            // per launch by themselves
            check_eio(eio_ctx_set_time_block(ctx, 2, EIO_VARIANT_AUTO));
            check_eio(eio_ctx_set_adjoint_time_block(ctx, 2, EIO_VARIANT_AUTO));
        }
    }
    ~Device() {
        if (ctx) eio_ctx_destroy(ctx);
    }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
};

inline int pick_device(const Params& par) {  // createGpuIdList (nonlinearPropElasticShotsGpu.cpp:45-77)
    int nGpu = par.getInt("nGpu", -1);
    std::vector<int> list = par.getInts("iGpu", {-1});
    if (nGpu <= 0 && list[0] < 0)
        throw std::runtime_error("**** ERROR: Please provide a list of GPUs to be used ****");
    if (list[0] >= 0) {
        std::vector<int> s = list;
        std::sort(s.begin(), s.end());
        if (std::unique(s.begin(), s.end()) != s.end())
            throw std::runtime_error("**** ERROR: Please make sure there are no duplicates in the GPU Id list ****");
        return list[0];
    }
    return 0;  // all shots run on one device; see the batched C ABI
}

template <typename T>
using InterpVec = std::vector<std::shared_ptr<SpaceInterpT<T>>>;

template <typename T>
void scale_sources(const FdParam<T>& fd, const SpaceInterpT<T>& ip, int k, T* reg) {
    // nonlinearPropElasticGpu.cpp:116-144: stresses *2dtw, vx *rhoxDtw, vz *rhozDtw, all /(dx dz)
    const size_t n = fd.ncell();
    const T two = T(2.0 * fd.dtw);
    const T area = T(1.0 / (fd.dx * fd.dz));
    for (int u = 0; u < ip.nReg(); ++u) {
        T s;
        if (k == 0) s = fd.media[size_t(ip.uniq[u])];
        else if (k == 1) s = fd.media[n + size_t(ip.uniq[u])];
        else s = two;
        T* r = reg + size_t(u) * ip.nt;
        for (int it = 0; it < ip.nt; ++it) r[it] = (r[it] * s) * area;
    }
}

template <typename T>
inline void fill_positions(eio_shot& s, const InterpVec<T>& src4, const InterpVec<T>& rec4) {
    // order of the four interpolators: centre, X, Z, XZ
    for (int g = 0; g < 4; ++g) {
        s.src_pos[g] = src4[g]->uniq.data();
        s.n_src[g] = src4[g]->nReg();
        s.rec_pos[g] = rec4[g]->uniq.data();
        s.n_rec[g] = rec4[g]->nReg();
    }
}

// ---------------------------------------------------------------- nonlinear shot operator
template <typename T>
class NonlinearShots {
   public:
    NonlinearShots(const T* elastic, std::vector<Axis> ax, Params par, std::vector<InterpVec<T>> sources,
                   std::vector<InterpVec<T>> receivers)
        : ax_(std::move(ax)), par_(std::move(par)), src_(std::move(sources)), rec_(std::move(receivers)) {
        nExp_ = par_.getInt("nExp");
        fd_ = std::make_unique<FdParam<T>>(elastic, ax_, par_);
        if (int(src_.size()) < nExp_) throw std::runtime_error("fewer source geometries than nExp");
        if (rec_.size() != 1 && int(rec_.size()) < nExp_)
            throw std::runtime_error("receiver geometries must be 1 or nExp");
        const int device = pick_device(par_);
        if (par_.getInt("info", 0) == 1) fd_->getInfo();
        dev_ = std::make_unique<Device>(device, int(sizeof(T)));
        register_model();
    }
    ~NonlinearShots() {
        if (model_ >= 0 && dev_) eio_model_remove(dev_->ctx, model_);
    }

    const FdParam<T>& fd() const { return *fd_; }

    // setBackground (nonlinearPropElasticShotsGpu.h:61): new fdParamElastic, same geometry
    void setBackground(const T* elastic) {
        fd_ = std::make_unique<FdParam<T>>(elastic, ax_, par_);
        const size_t n = fd_->ncell();
        const void* media[5] = {fd_->media.data(), fd_->media.data() + n, fd_->media.data() + 2 * n,
                                fd_->media.data() + 3 * n, fd_->media.data() + 4 * n};
        check_eio(eio_model_set_media(dev_->ctx, model_, media, 0));
    }

    // model [nSrcSig][5][nSrc][nts] with nSrcSig = 1 or nExp; data [nExp][5][nRec][nts].
    // save: also keep the wavefield of shot wavefieldShotNumber when saveWavefield != 0
    // (nonlinearPropElasticShotsGpu.cpp:185-300).
    void forward(bool add, const T* model, int nModelShots, int nSrc, T* data, int nRec, bool save = false) {
        const int nts = fd_->nts;
        if (nModelShots != 1 && nModelShots != nExp_) throw std::runtime_error("model shot axis must be 1 or nExp");
        std::vector<eio_shot> shots = shot_table();
        for (int e = 0; e < nExp_; ++e) check_counts(src_[e], nSrc, rec_of(e), nRec);
        std::vector<PinnedVec<T>>& srcReg = srcReg_;
        std::vector<PinnedVec<T>>& recReg = recReg_;
        // per (shot, component): disjoint buffers, so the interpolation runs on all host cores
#pragma omp parallel for collapse(2) schedule(dynamic)
        for (int e = 0; e < nExp_; ++e)
            for (int k = 0; k < 5; ++k) {
                const T* m = model + size_t(nModelShots == 1 ? 0 : e) * 5 * nSrc * nts;
                const SpaceInterpT<T>& ip = *src_[e][GRID_OF_COMP[k]];
                PinnedVec<T>& r = srcReg[size_t(e) * 5 + k];
                r.resize(size_t(ip.nReg()) * nts);
                ip.adjoint(r.data(), m + size_t(k) * nSrc * nts);
                scale_sources(*fd_, ip, k, r.data());
                // output: the library writes every regular receiver trace of the call
                PinnedVec<T>& d = recReg[size_t(e) * 5 + k];
                d.resize(size_t(rec_of(e)[GRID_OF_COMP[k]]->nReg()) * nts);
                shots[e].src_tr[k] = r.data();
                shots[e].rec_tr[k] = d.data();
            }
        if (save_shot(save) >= 0) {
            wavefield_.assign(size_t(nts) * 5 * fd_->ncell(), T(0));
            check_eio(eio_nl_forward_wavefield(dev_->ctx, nExp_, shots.data(), 0, save_shot(save), wavefield_.data()));
        } else {
            check_eio(eio_nl_forward(dev_->ctx, nExp_, shots.data(), 0));
        }
        if (!add) std::fill(data, data + size_t(nExp_) * 5 * nRec * nts, T(0));
#pragma omp parallel for collapse(2) schedule(dynamic)
        for (int e = 0; e < nExp_; ++e)
            for (int k = 0; k < 5; ++k)
                rec_of(e)[GRID_OF_COMP[k]]->forward_add(recReg[size_t(e) * 5 + k].data(),
                                                        data + (size_t(e) * 5 + k) * nRec * nts);
    }

    void adjoint(bool add, T* model, int nModelShots, int nSrc, const T* data, int nRec, bool save = false) {
        const int nts = fd_->nts;
        if (nModelShots != 1 && nModelShots != nExp_) throw std::runtime_error("model shot axis must be 1 or nExp");
        std::vector<eio_shot> shots = shot_table();
        for (int e = 0; e < nExp_; ++e) check_counts(src_[e], nSrc, rec_of(e), nRec);
        std::vector<PinnedVec<T>>& srcReg = srcReg_;
        std::vector<PinnedVec<T>>& recReg = recReg_;
#pragma omp parallel for collapse(2) schedule(dynamic)
        for (int e = 0; e < nExp_; ++e)
            for (int k = 0; k < 5; ++k) {
                const SpaceInterpT<T>& rp = *rec_of(e)[GRID_OF_COMP[k]];
                PinnedVec<T>& d = recReg[size_t(e) * 5 + k];
                d.resize(size_t(rp.nReg()) * nts);
                rp.adjoint(d.data(), data + (size_t(e) * 5 + k) * nRec * nts);
                // output: the library writes every regular source trace of the call
                PinnedVec<T>& r = srcReg[size_t(e) * 5 + k];
                r.resize(size_t(src_[e][GRID_OF_COMP[k]]->nReg()) * nts);
                shots[e].src_tr[k] = r.data();
                shots[e].rec_tr[k] = d.data();
            }
        if (save_shot(save) >= 0) {
            const size_t n = fd_->ncell();
            wavefield_.assign(size_t(nts) * 5 * n, T(0));
            check_eio(eio_nl_adjoint_wavefield(dev_->ctx, nExp_, shots.data(), 0, save_shot(save), wavefield_.data()));
            // nonlinearPropElasticGpu.cpp:432-448: rhoxDtw, rhozDtw, 2*dtw, all / (dx*dz)
            const T area = T(1.0 / (fd_->dx * fd_->dz));
            const T two = T(2.0 * fd_->dtw);
            for (int it = 0; it < nts; ++it) {
                T* w = wavefield_.data() + size_t(it) * 5 * n;
                for (size_t i = 0; i < n; ++i) {
                    w[i] *= fd_->media[i] * area;
                    w[n + i] *= fd_->media[n + i] * area;
                    w[2 * n + i] *= two * area;
                    w[3 * n + i] *= two * area;
                    w[4 * n + i] *= two * area;
                }
            }
        } else {
            check_eio(eio_nl_adjoint(dev_->ctx, nExp_, shots.data(), 0));
        }
        if (!add) std::fill(model, model + size_t(nModelShots) * 5 * nSrc * nts, T(0));
        // one signal for all shots: every shot adds into the same slot, so shots stay in order
        // and only the components run in parallel
        const int nGroup = nModelShots == 1 ? 1 : nExp_;
        for (int e0 = 0; e0 < nExp_; e0 += nGroup) {
            const int e1 = std::min(nExp_, e0 + nGroup);
#pragma omp parallel for collapse(2) schedule(dynamic)
            for (int e = e0; e < e1; ++e)
                for (int k = 0; k < 5; ++k) {
                    T* m = model + size_t(nModelShots == 1 ? 0 : e) * 5 * nSrc * nts;
                    const SpaceInterpT<T>& ip = *src_[e][GRID_OF_COMP[k]];
                    PinnedVec<T>& r = srcReg[size_t(e) * 5 + k];
                    scale_sources(*fd_, ip, k, r.data());  // same diagonal scaling (nonlinearPropElasticGpu.cpp:406-430)
                    ip.forward_add(r.data(), m + size_t(k) * nSrc * nts);
                }
        }
    }

    int nExp() const { return nExp_; }
    // the last saved wavefield [nts][5][nx][nz] (empty before any saving call)
    const std::vector<T>& wavefield() const { return wavefield_; }
    int nSrcIrreg() const { return src_[0][0]->nDeviceIrreg; }
    int nRecIrreg() const { return rec_[0][0]->nDeviceIrreg; }
    double last_ms() const {
        double t = 0;
        eio_ctx_last_timing(dev_->ctx, &t, nullptr, nullptr);
        return t;
    }

   private:
    const InterpVec<T>& rec_of(int e) const { return rec_.size() == 1 ? rec_[0] : rec_[e]; }
    // the call's shot descriptors (positions, model id; traces are set by the caller) and
    // one regular-trace buffer per shot and component, kept between calls
    std::vector<eio_shot> shot_table() {
        std::vector<eio_shot> shots(nExp_);
        for (int e = 0; e < nExp_; ++e) {
            std::memset(&shots[e], 0, sizeof(eio_shot));
            shots[e].model = model_;
            fill_positions(shots[e], src_[e], rec_of(e));
        }
        srcReg_.resize(size_t(nExp_) * 5);
        recReg_.resize(size_t(nExp_) * 5);
        return shots;
    }
    // shot whose wavefield a saving call keeps, or -1
    int save_shot(bool save) const {
        if (!save || par_.getInt("saveWavefield", 0) == 0) return -1;
        const int s = par_.getInt("wavefieldShotNumber", 0);
        if (s < 0 || s >= nExp_) throw std::runtime_error("wavefieldShotNumber out of range");
        return s;
    }
    void register_model() {
        eio_grid g;
        g.nz = fd_->nz; g.nx = fd_->nx; g.nts = fd_->nts; g.sub_fd = fd_->sub; g.minPad = fd_->minPad;
        g.dz = fd_->dz; g.dx = fd_->dx; g.dts = fd_->dts; g.alphaCos = fd_->alphaCos;
        g.surface = fd_->surfaceCondition;
        const size_t n = fd_->ncell();
        const void* media[5] = {fd_->media.data(), fd_->media.data() + n, fd_->media.data() + 2 * n,
                                fd_->media.data() + 3 * n, fd_->media.data() + 4 * n};
        check_eio(eio_model_add(dev_->ctx, &g, media, 0, &model_));
    }
    void check_counts(const InterpVec<T>& S, int nSrc, const InterpVec<T>& Rv, int nRec) const {
        if (S.size() != 4 || Rv.size() != 4) throw std::runtime_error("geometry must have 4 grids");
        for (int g = 0; g < 4; ++g) {
            if (S[g]->nDeviceIrreg != nSrc) throw std::runtime_error("source count not consistent with the model vector");
            if (Rv[g]->nDeviceIrreg != nRec) throw std::runtime_error("receiver count not consistent with the data vector");
            if (S[g]->nt != fd_->nts || Rv[g]->nt != fd_->nts) throw std::runtime_error("nts not consistent with parfile");
        }
    }

    std::vector<Axis> ax_;
    Params par_;
    std::vector<InterpVec<T>> src_, rec_;
    int nExp_ = 1;
    std::unique_ptr<Device> dev_;
    std::unique_ptr<FdParam<T>> fd_;
    int model_ = -1;
    std::vector<T> wavefield_;
    std::vector<PinnedVec<T>> srcReg_, recReg_;  // regular traces per (shot, component), page-locked
};

// ---------------------------------------------------------------- Born shot operator
template <typename T>
class BornShots {
   public:
    // sourcesSignals: list of [5][nSrc][nts] (1 = same for all shots, or nExp)
    BornShots(const T* elastic, std::vector<Axis> ax, Params par, std::vector<std::vector<T>> sourcesSignals,
              std::vector<InterpVec<T>> sources, std::vector<InterpVec<T>> receivers)
        : ax_(std::move(ax)), par_(std::move(par)), sig_(std::move(sourcesSignals)), src_(std::move(sources)),
          rec_(std::move(receivers)) {
        nExp_ = par_.getInt("nExp");
        fd_ = std::make_unique<FdParam<T>>(elastic, ax_, par_);
        if (fd_->surfaceCondition != 0)
            throw std::runtime_error("surfaceCondition=1 (free surface) is available for the nonlinear operator only");
        if (int(src_.size()) < nExp_) throw std::runtime_error("fewer source geometries than nExp");
        if (rec_.size() != 1 && int(rec_.size()) < nExp_) throw std::runtime_error("receiver geometries must be 1 or nExp");
        if (sig_.empty() || (sig_.size() != 1 && int(sig_.size()) < nExp_))
            throw std::runtime_error("source signals must be 1 or nExp");
        const int device = pick_device(par_);
        if (par_.getInt("info", 0) == 1) fd_->getInfo();
        dev_ = std::make_unique<Device>(device, int(sizeof(T)));
        register_model();
    }
    ~BornShots() {
        if (model_ >= 0 && dev_) eio_model_remove(dev_->ctx, model_);
    }
    const FdParam<T>& fd() const { return *fd_; }

    void setBackground(const T* elastic) {
        fd_ = std::make_unique<FdParam<T>>(elastic, ax_, par_);
        const size_t n = fd_->ncell();
        const void* media[5] = {fd_->media.data(), fd_->media.data() + n, fd_->media.data() + 2 * n,
                                fd_->media.data() + 3 * n, fd_->media.data() + 4 * n};
        check_eio(eio_model_set_media(dev_->ctx, model_, media, 0));
    }

    // model [3][nx][nz] (drho, dlambda, dmu); data [nExp][5][nRec][nts]
    void forward(bool add, const T* model, T* data, int nRec) {
        const int nts = fd_->nts;
        const size_t n = fd_->ncell();
        // BornElasticShotsGpu.cpp:107-150
        std::vector<T> sl(5 * n, T(0)), tmp(n);
        stagger_x_adj(fd_->nz, fd_->nx, model, sl.data());
        stagger_z_adj(fd_->nz, fd_->nx, model, sl.data() + n);
        std::copy(model + n, model + 2 * n, sl.data() + 2 * n);
        std::copy(model + 2 * n, model + 3 * n, sl.data() + 3 * n);
        stagger_x_adj(fd_->nz, fd_->nx, model + 2 * n, tmp.data());
        stagger_z_adj(fd_->nz, fd_->nx, tmp.data(), sl.data() + 4 * n);
        scale_slices(sl);
        std::vector<eio_shot> shots = build_shots(nRec, nullptr);
        const void* dm[5] = {sl.data(), sl.data() + n, sl.data() + 2 * n, sl.data() + 3 * n, sl.data() + 4 * n};
        check_eio(eio_born_forward(dev_->ctx, nExp_, shots.data(), dm, 0));
        if (!add) std::fill(data, data + size_t(nExp_) * 5 * nRec * nts, T(0));
#pragma omp parallel for collapse(2) schedule(dynamic)
        for (int e = 0; e < nExp_; ++e)
            for (int k = 0; k < 5; ++k)
                rec_of(e)[GRID_OF_COMP[k]]->forward_add(recReg_[size_t(e) * 5 + k].data(),
                                                        data + (size_t(e) * 5 + k) * nRec * nts);
    }

    void adjoint(bool add, T* model, const T* data, int nRec) {
        const size_t n = fd_->ncell();
        std::vector<eio_shot> shots = build_shots(nRec, data);
        std::vector<T> acc(5 * n, T(0));
        void* dm[5] = {acc.data(), acc.data() + n, acc.data() + 2 * n, acc.data() + 3 * n, acc.data() + 4 * n};
        check_eio(eio_born_adjoint(dev_->ctx, nExp_, shots.data(), dm, 0));
        // BornElasticShotsGpu.cpp:325-371
        scale_slices(acc);
        if (!add) std::fill(model, model + 3 * n, T(0));
        stagger_x_fwd_add(fd_->nz, fd_->nx, acc.data(), model);
        stagger_z_fwd_add(fd_->nz, fd_->nx, acc.data() + n, model);
        std::vector<T> t1(n, T(0)), t2(n, T(0));
        stagger_z_fwd_add(fd_->nz, fd_->nx, acc.data() + 4 * n, t1.data());
        stagger_x_fwd_add(fd_->nz, fd_->nx, t1.data(), t2.data());
        for (size_t i = 0; i < n; ++i) {
            model[n + i] += acc[2 * n + i];
            model[2 * n + i] += acc[3 * n + i] + t2[i];
        }
    }

    int nExp() const { return nExp_; }
    // device time of the last forward or adjoint call (its passes and imaging), ms
    double last_ms() const {
        double t = 0;
        eio_ctx_last_timing(dev_->ctx, &t, nullptr, nullptr);
        return t;
    }

   private:
    void scale_slices(std::vector<T>& sl) const {
        const size_t n = fd_->ncell();
        const T two = T(2.0 * fd_->dtw);
        for (size_t i = 0; i < n; ++i) {
            sl[i] *= fd_->media[i];
            sl[n + i] *= fd_->media[n + i];
            sl[2 * n + i] *= two;
            sl[3 * n + i] *= two;
            sl[4 * n + i] *= two;
        }
    }

    // Born sources: interpolate, scale and fine-sample (seismicElasticOperator2D.cpp:27-119)
    const InterpVec<T>& rec_of(int e) const { return rec_.size() == 1 ? rec_[0] : rec_[e]; }

    // The call's shot descriptors with fine-sampled source traces and regular receiver
    // traces (interpolated from `data`, or zero), in buffers kept between calls.
    std::vector<eio_shot> build_shots(int nRec, const T* data) {
        const int nts = fd_->nts, ntw = fd_->ntw;
        std::vector<eio_shot> shots(nExp_);
        for (int e = 0; e < nExp_; ++e) {
            const std::vector<T>& sig = sig_.size() == 1 ? sig_[0] : sig_[e];
            const int nSrc = src_[e][0]->nDeviceIrreg;
            if (sig.size() != size_t(5) * nSrc * nts) throw std::runtime_error("source signal size not consistent");
            for (int g = 0; g < 4; ++g) {
                if (rec_of(e)[g]->nDeviceIrreg != nRec)
                    throw std::runtime_error("receiver count not consistent with the data vector");
                if (src_[e][g]->nt != nts || rec_of(e)[g]->nt != nts)
                    throw std::runtime_error("nts not consistent with parfile");
            }
            std::memset(&shots[e], 0, sizeof(eio_shot));
            shots[e].model = model_;
            fill_positions(shots[e], src_[e], rec_of(e));
        }
        fine_.resize(size_t(nExp_) * 5);
        recReg_.resize(size_t(nExp_) * 5);
        // per (shot, component): disjoint buffers, so the interpolation runs on all host cores
#pragma omp parallel for collapse(2) schedule(dynamic)
        for (int e = 0; e < nExp_; ++e)
            for (int k = 0; k < 5; ++k) {
                const std::vector<T>& sig = sig_.size() == 1 ? sig_[0] : sig_[e];
                const int nSrc = src_[e][0]->nDeviceIrreg;
                const SpaceInterpT<T>& ip = *src_[e][GRID_OF_COMP[k]];
                std::vector<T> reg(size_t(ip.nReg()) * nts);
                ip.adjoint(reg.data(), sig.data() + size_t(k) * nSrc * nts);
                scale_sources(*fd_, ip, k, reg.data());
                PinnedVec<T>& f = fine_[size_t(e) * 5 + k];
                f.resize(size_t(ip.nReg()) * ntw);
                time_interp_fwd(nts, fd_->dts, fd_->ots, fd_->sub, ip.nReg(), reg.data(), f.data());
                shots[e].src_tr[k] = f.data();
                const SpaceInterpT<T>& rp = *rec_of(e)[GRID_OF_COMP[k]];
                PinnedVec<T>& d = recReg_[size_t(e) * 5 + k];
                d.resize(size_t(rp.nReg()) * nts);
                // forward: an output the library writes in full
                if (data) rp.adjoint(d.data(), data + (size_t(e) * 5 + k) * nRec * nts);
                shots[e].rec_tr[k] = d.data();
            }
        return shots;
    }

    void register_model() {
        eio_grid g;
        g.nz = fd_->nz; g.nx = fd_->nx; g.nts = fd_->nts; g.sub_fd = fd_->sub; g.minPad = fd_->minPad;
        g.dz = fd_->dz; g.dx = fd_->dx; g.dts = fd_->dts; g.alphaCos = fd_->alphaCos;
        g.surface = fd_->surfaceCondition;
        const size_t n = fd_->ncell();
        const void* media[5] = {fd_->media.data(), fd_->media.data() + n, fd_->media.data() + 2 * n,
                                fd_->media.data() + 3 * n, fd_->media.data() + 4 * n};
        check_eio(eio_model_add(dev_->ctx, &g, media, 0, &model_));
    }

    std::vector<Axis> ax_;
    Params par_;
    std::vector<std::vector<T>> sig_;
    std::vector<InterpVec<T>> src_, rec_;
    int nExp_ = 1;
    std::unique_ptr<Device> dev_;
    std::unique_ptr<FdParam<T>> fd_;
    int model_ = -1;
    // source (fine) and receiver traces per (shot, component), page-locked
    std::vector<PinnedVec<T>> fine_, recReg_;
};

}  // namespace eioh
