// eio_capi.hip -- C ABI (include/eio.h) of the gfx950 elastic propagator.
//
// Host orchestration of the fused kernels in eio_kernels.h:
//   * a context owns a stream, timing events, registered models and a grow-only device
//     workspace (288 GB of HBM: buffers are kept across calls instead of malloc/free per shot);
//   * a call propagates all of its shots in ONE batched time loop: every launch advances
//     every active shot by one fine step (shots are ordered by decreasing trace length so
//     the active shots are always a prefix of the tile table);
//   * source/receiver devices are turned into per-tile sparse entry lists so injection and
//     recording happen inside the stencil launch.
// Time loops follow nonlinearPropElasticGpuFunctions.cu:588-617 / 966-997 and
// BornElasticGpuFunctions.cu:744-819 / 1128-1191.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/eio.h"
#include "eio_kernels.h"
#include "eio_step2.h"
#include "eio_step2q.h"
#include "eio_step3.h"
#include "eio_abc.h"
#include "eio_abc_train.h"

// Workspace of eio_nl_forward_noalloc: device memory the library never allocates at call
// time.  Each GPU holds its own copy of this array for the life of the process; in host code
// its name is not a device address, so host code reaches it only through the pointer that
// hipGetSymbolAddress returns on that device (g_scratch_ptr, taken once per device when a
// context is created there).
constexpr size_t EIO_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[EIO_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

struct EioError {
    int code;
    std::string msg;
};

#define EIO_CHECK_HIP(expr)                                                                   \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw EioError{EIO_ERR_RUNTIME, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
    } while (0)

void fail(const std::string& m) { throw EioError{EIO_ERR_ARG, m}; }

// Device address of the scratch array on each device, taken with that device current the
// first time a context is created on it.
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_ptr;

unsigned char* scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (int(g_scratch_ptr.size()) <= device) g_scratch_ptr.resize(size_t(device) + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        EIO_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw EioError{EIO_ERR_RUNTIME, "hipGetSymbolAddress returned no address for the scratch array"};
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    return g_scratch_ptr[device];
}

constexpr double FD_C[4] = {1.196289062541883, -0.079752604188901, 0.009570312506634,
                            -6.975446437140719e-04};  // nonlinearPropElasticGpuFunctions.cu:159-167
constexpr int GRID_OF_COMP[5] = {1, 2, 0, 0, 3};
constexpr int PAD_MAX = 500;   // varDeclare.h:9
constexpr int SUB_MAX = 100;   // varDeclare.h:10

// Step-kernel variants: R x-cells per lane, WX waves per workgroup (tile = 64 x WX*R),
// OCC requested waves per SIMD.  FP32 has a table the autotuner chooses from; FP64 (dot
// tests and oracle comparisons only) has one variant.
struct Variant {
    int R, WX, OCC;
};
constexpr Variant F32_VARIANTS[] = {{4, 4, 2}, {2, 8, 4}, {2, 4, 4}, {4, 8, 2}, {8, 4, 1}, {2, 8, 2}};
constexpr int N_F32_VARIANTS = int(sizeof(F32_VARIANTS) / sizeof(F32_VARIANTS[0]));
constexpr Variant F64_VARIANT = {2, 4, 1};

// Two-step (temporally blocked) NL forward variants: owned tile = 56 x WX*R2, WX waves, OCC
// waves per SIMD.  LAYOUT 0: lane l owns row z0-4+l of R2 columns, u_cur staged whole in LDS;
// 1: the same lanes, u_cur streamed through two LDS field buffers (78.8 KB instead of 120 KB
// at 56 x 32: two workgroups per CU); 2: a lane owns a 4-row quad of one column (k_step2q_fwd:
// 16-byte loads and stores; R2 = 3 stands for its 24 owned columns).  Layouts 1 and 2 serve
// the NL forward and the recompute; the other passes of such a variant run a layout-0 kernel
// of the same tile width.  NTS: nontemporal stores.  FP32 only.
struct Variant2 {
    int R2, WX, OCC, NTS, LAYOUT;
};
constexpr Variant2 F32_VARIANTS2[] = {{2, 8, 2, 0, 0}, {2, 8, 4, 0, 0}, {4, 4, 2, 0, 0}, {1, 8, 4, 0, 0},
                                      {4, 8, 1, 0, 0}, {2, 8, 4, 1, 0}, {2, 8, 2, 1, 0}, {4, 8, 1, 1, 0},
                                      {4, 8, 4, 0, 1}, {8, 4, 2, 0, 1}, {2, 8, 6, 0, 1}, {3, 8, 2, 0, 2},
                                      {1, 8, 8, 0, 1}};
constexpr int N_F32_VARIANTS2 = int(sizeof(F32_VARIANTS2) / sizeof(F32_VARIANTS2[0]));
// Three-step NL forward variants (k_step3_fwd): R3 owned x-cells per lane, WX waves, OCC waves
// per SIMD; owned tile = 48 x WX*R3.
struct Variant3 {
    int R3, WX, OCC;
};
constexpr Variant3 F32_VARIANTS3[] = {{3, 8, 1}, {4, 8, 1}, {1, 8, 4}};
constexpr int N_F32_VARIANTS3 = int(sizeof(F32_VARIANTS3) / sizeof(F32_VARIANTS3[0]));
// Two-step adjoint (Born receiver pass, FWI gradient) variants: R2, WX, OCC as above.
struct Variant2A {
    int R2, WX, OCC, LATE, PS;  // LATE: per-cell loads after the stage is written; PS: slots with the stage
};
constexpr Variant2A F32_VARIANTS2A[] = {{2, 8, 2, 0, 1}, {4, 8, 1, 0, 1}, {2, 8, 4, 0, 1}, {4, 4, 2, 0, 1},
                                        {2, 8, 4, 1, 1}, {2, 8, 2, 1, 1}, {4, 8, 1, 1, 1}, {4, 8, 1, 1, 0},
                                        {4, 8, 1, 0, 0}, {2, 8, 2, 1, 0}, {1, 8, 4, 1, 0}, {1, 8, 4, 0, 1}};
constexpr int N_F32_VARIANTS2A = int(sizeof(F32_VARIANTS2A) / sizeof(F32_VARIANTS2A[0]));

template <typename T>
int tile_x(int variant) {
    if (sizeof(T) == 8) return F64_VARIANT.R * F64_VARIANT.WX;
    return F32_VARIANTS[variant].R * F32_VARIANTS[variant].WX;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct ModelBase {
    eio_grid g;
    int ntw, h;
    long n;
    void* med = nullptr;    // 5*n
    void* damp = nullptr;   // minPad
    void* filt = nullptr;   // 2h + 1 (last entry: 1/(2 dts))
    double cz[4], cx[4];
    int abc_on = 0;         // learned absorbing boundary attached: 1 difference inputs, 2 level inputs
    int abc_depth = 4;      // frame rows the network writes
    void* abc_w = nullptr;  // [4][3][13][64][8] bf16 B fragments (tap pairs)
    float abc_scale[10] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
};

// Forward-wavefield store of the checkpointed FWI gradient: the batch state at the start of
// every segment of L coarse intervals but the first (which starts from rest), plus the coarse
// Vx/Vz ring (L+3 slices per shot) that already holds the last segment.
struct FwiStore {
    int prec = 4, nshot = 0, model = 0, nts = 0, sub = 0, L = 0, nseg = 0, wcap = 0;
    unsigned long long sig = 0;  // hash of the shot geometry the store was made for
    size_t state_elems = 0;      // elements of one batch state (10 fields per shot)
    void* ck = nullptr;          // (nseg-1) batch states
    void* ring = nullptr;        // per ordered shot: Vx then Vz, wcap slices each
    bool ring_fresh = false;     // the ring holds the last segment (true right after the forward)
    bool ws = false;             // ring and checkpoints borrowed from the context workspace
    size_t bytes = 0;
};

// One kernel launch of a time loop, recorded instead of launched (graph mode): the kernel,
// its shape and its arguments packed as the kernel takes them.
struct LaunchRec {
    const void* fn = nullptr;
    dim3 grid, block;
    std::vector<unsigned char> blob;
    std::vector<unsigned> offs;
};

// An executable graph of one launch sequence shape (kernels and shapes in order); its kernel
// nodes' arguments are updated in place when a later call has the same shape.
struct GraphEntry {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipGraphNode_t> nodes;
    std::vector<LaunchRec> log;
    unsigned long long last_use = 0;
};

struct Ctx {
    int device = 0;
    int prec = 4;
    unsigned char* scratch_dev = nullptr;  // this device's scratch array (scratch_address)
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4];
    std::map<int, ModelBase> models;
    int next_id = 1;
    std::map<std::string, DevBuf> ws;
    // while a no-allocation call runs: its workspace is carved from the device's scratch array
    unsigned char* bump = nullptr;
    size_t bump_cap = 0, bump_off = 0;
    double last_total_ms = 0.0, last_step_ms = 0.0;
    long long last_nstep = 0;
    int use_graph = 0;
    bool recording = false;                 // graph mode: launches go to `log`
    std::vector<LaunchRec> log;
    std::map<unsigned long long, GraphEntry> graphs;
    unsigned long long graph_clock = 0;
    long long graph_builds = 0, graph_launches = 0, graph_node_updates = 0;
    int variant = 1;  // best measured FP32 variant on the 64-shot layered batch
    int time_block = 1;  // fine steps per launch of the FP32 NL forward (1, 2 or 3)
    int variant3 = 0;    // three-step kernel variant, or EIO_VARIANT_AUTO
    int variant2 = 0;    // two-step kernel variant, or EIO_VARIANT_AUTO (chosen per call)
    int adj_time_block = 1;  // reverse fine steps per launch of the FP32 Born adjoint (1 or 2)
    int variant2a = 0;       // two-step adjoint kernel variant, or EIO_VARIANT_AUTO
    int step_limit = 0;  // > 0: NL loops stop after this many fine steps (tuning runs)
    long long group_bytes = EIO_GROUP_AUTO;  // shot-group working-set budget of the NL forward
    int last_groups = 0;                     // shot groups of the last NL forward call
    int group_streams = 2;                   // streams the shot groups of a call run on (1 or 2)
    hipStream_t side = nullptr;              // the library's second stream (default priority)
    hipEvent_t fork = nullptr, join = nullptr;
    int persistent = 0;  // two-step NL forward in one persistent launch: 0 off, 1 on, 2 when it fits
    int spin_limit = 1 << 20;  // polls per neighbour wait of the persistent launch before it aborts
    std::map<int, FwiStore> fwi;
    int next_fwi = 1;
    bool fwi_ws_busy = false;  // a store holds the workspace's "fwi_ring" / "fwi_ck" buffers
};

void* ws_get(Ctx* c, const std::string& name, size_t bytes) {
    if (c->bump) {  // no-allocation call: 256-byte aligned pieces of the scratch array
        const size_t off = (c->bump_off + 255) / 256 * 256;
        if (off > c->bump_cap || bytes > c->bump_cap - off)
            fail("eio_nl_forward_noalloc: the call needs more than the " + std::to_string(c->bump_cap >> 20) +
                 " MiB scratch workspace (buffer \"" + name + "\" of " + std::to_string(bytes) +
                 " bytes at offset " + std::to_string(off) + "); use eio_nl_forward");
        c->bump_off = off + bytes;
        return c->bump + off;
    }
    DevBuf& b = c->ws[name];
    if (b.bytes < bytes) {
        if (b.p) {
            EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
            EIO_CHECK_HIP(hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
        }
        EIO_CHECK_HIP(hipMalloc(&b.p, bytes));
        b.bytes = bytes;
    }
    return b.p;
}

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int d) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != d) EIO_CHECK_HIP(hipSetDevice(d));
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// ------------------------------------------------------------------ kernel launches
// Every launch of a time loop goes through klaunch: launched at once on the context stream,
// or, while the context records (graph mode), appended to its log with the arguments
// converted to the kernel's parameter types and packed as the kernel-node parameters.
template <typename Tup, size_t... I>
void tuple_ptrs(Tup& t, void** p, std::index_sequence<I...>) {
    ((p[I] = static_cast<void*>(&std::get<I>(t))), ...);
}

template <typename... P, typename... A>
void klaunch(Ctx* c, void (*fn)(P...), dim3 grid, dim3 block, A&&... a) {
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    static_assert((std::is_trivially_copyable<P>::value && ...), "kernel arguments must be trivially copyable");
    std::tuple<P...> t(std::forward<A>(a)...);
    if (!c->recording) {
        void* ptrs[sizeof...(P)];
        tuple_ptrs(t, ptrs, std::index_sequence_for<P...>{});
        EIO_CHECK_HIP(hipLaunchKernel(reinterpret_cast<const void*>(fn), grid, block, ptrs, 0, c->stream));
        return;
    }
    LaunchRec r;
    r.fn = reinterpret_cast<const void*>(fn);
    r.grid = grid;
    r.block = block;
    void* ptrs[sizeof...(P)];
    tuple_ptrs(t, ptrs, std::index_sequence_for<P...>{});
    constexpr size_t sizes[] = {sizeof(P)...};
    constexpr size_t aligns[] = {alignof(P)...};
    size_t off = 0;
    for (size_t i = 0; i < sizeof...(P); ++i) {
        off = (off + aligns[i] - 1) / aligns[i] * aligns[i];
        r.offs.push_back(unsigned(off));
        off += sizes[i];
    }
    r.blob.resize(off);
    for (size_t i = 0; i < sizeof...(P); ++i) std::memcpy(r.blob.data() + r.offs[i], ptrs[i], sizes[i]);
    c->log.push_back(std::move(r));
}

template <typename T>
void fill0(Ctx* c, T* p, long n) {
    if (n <= 0) return;
    const long nb = (n + eio::NTHR - 1) / eio::NTHR;
    klaunch(c, eio::k_fill0<T>, dim3(unsigned(nb)), dim3(eio::NTHR), p, n);
}

// ------------------------------------------------------------------ model registration
template <typename T>
void upload_model(Ctx* c, ModelBase& m, const void* const media[5], int flags) {
    const size_t nb = size_t(m.n) * sizeof(T);
    for (int k = 0; k < 5; ++k) {
        if (!media[k]) fail("media pointer is NULL");
        const int kind = (flags & EIO_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        EIO_CHECK_HIP(hipMemcpyAsync(static_cast<char*>(m.med) + k * nb, media[k], nb,
                                     hipMemcpyKind(kind), c->stream));
    }
    EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
}

// Grid checks of eio_model_add (host only; no device call).
void check_grid(const eio_grid* g, int prec) {
    if (!g) fail("grid is NULL");
    if (prec != 4 && prec != 8) fail("precision must be 4 or 8");
    if (g->nz < 2 * eio::FAT + 1 || g->nx < 2 * eio::FAT + 1) fail("nz and nx must exceed 2*FAT");
    if (g->nts < 2) fail("nts must be >= 2");
    if (g->sub_fd < 1) fail("sub_fd must be >= 1");
    if (g->sub_fd > SUB_MAX) fail("Subsampling parameter too high");  // fdParamElastic.cpp:29-32
    if (g->minPad < 0 || g->minPad >= PAD_MAX) fail("Padding value is too high");  // NLF:191-194
    if (!(g->dz > 0) || !(g->dx > 0) || !(g->dts > 0)) fail("dz, dx, dts must be > 0");
    // the kernels address a 5-field buffer (state level, media) through one buffer descriptor
    // with 32-bit byte offsets: its size must stay below 2^31 bytes
    if (5LL * g->nz * g->nx * prec >= (1LL << 31))
        fail("grid too large: 5*nz*nx*" + std::to_string(prec) +
             " bytes must stay below 2^31 (32-bit buffer offsets)");
    if (g->surface != 0 && g->surface != 1) fail("surface must be 0 or 1");
    // the free-surface adjoint needs rows up to 2*FAT+4 damped independently of z
    if (g->surface == 1 && g->nz - 1 - eio::FAT - (eio::SURF + eio::FAT) < g->minPad)
        fail("free surface: nz too small for the bottom sponge");
}

template <typename T>
int model_add(Ctx* c, const eio_grid* g, const void* const media[5], int flags) {
    check_grid(g, int(sizeof(T)));
    ModelBase m;
    m.g = *g;
    m.n = long(g->nz) * g->nx;
    m.ntw = (g->nts - 1) * g->sub_fd + 1;
    m.h = g->sub_fd + 1;
    for (int k = 0; k < 4; ++k) { m.cz[k] = FD_C[k] / g->dz; m.cx[k] = FD_C[k] / g->dx; }
    // time interpolation filter (nonlinearPropElasticGpuFunctions.cu:172-188) + 1/(2 dts)
    std::vector<T> filt(2 * m.h + 1);
    const double s = 1.0 / std::sqrt(double(m.ntw) / double(g->nts));
    for (int i = 0; i < m.h; ++i) {
        const double a = 1.0 - double(i) / double(g->sub_fd);
        filt[i] = T(a * s);
        filt[m.h + i] = T((1.0 - a) * s);
    }
    filt[2 * m.h] = T(1.0 / (2.0 * g->dts));  // BornElasticGpuFunctions.cu:264-265
    // cosine sponge (nonlinearPropElasticGpuFunctions.cu:195-202)
    std::vector<T> damp(std::max(g->minPad, 1), T(1));
    for (int j = 0; j < g->minPad; ++j) {
        const double arg = M_PI / double(g->minPad) * double(g->minPad - j);
        damp[j] = T(g->alphaCos + (1.0 - g->alphaCos) * std::cos(arg));
    }
    EIO_CHECK_HIP(hipMalloc(&m.med, size_t(5) * m.n * sizeof(T)));
    EIO_CHECK_HIP(hipMalloc(&m.damp, damp.size() * sizeof(T)));
    EIO_CHECK_HIP(hipMalloc(&m.filt, filt.size() * sizeof(T)));
    EIO_CHECK_HIP(hipMemcpy(m.damp, damp.data(), damp.size() * sizeof(T), hipMemcpyHostToDevice));
    EIO_CHECK_HIP(hipMemcpy(m.filt, filt.data(), filt.size() * sizeof(T), hipMemcpyHostToDevice));
    upload_model<T>(c, m, media, flags);
    const int id = c->next_id++;
    c->models[id] = m;
    return id;
}

void model_free(ModelBase& m) {
    if (m.med) (void)hipFree(m.med);
    if (m.damp) (void)hipFree(m.damp);
    if (m.filt) (void)hipFree(m.filt);
    if (m.abc_w) (void)hipFree(m.abc_w);
    m.med = m.damp = m.filt = m.abc_w = nullptr;
}

ModelBase& get_model(Ctx* c, int id) {
    auto it = c->models.find(id);
    if (it == c->models.end()) fail("unknown model id " + std::to_string(id));
    return it->second;
}

// ------------------------------------------------------------------ batch planning
enum class Op { NL_FWD, NL_ADJ, BORN_FWD, BORN_ADJ, FWI_FWD, FWI_GRAD };

struct Plan {
    std::vector<int> order;               // shot indices sorted by decreasing ntw
    std::vector<eio::TileDev> tiles;
    std::vector<eio::Entry> ents;
    std::vector<int> tile_end;            // tile count of the first i+1 ordered shots
    std::vector<int> nsteps;              // per ordered shot: ntw-1
};

template <typename T>
void build_entries(const ModelBase& m, const int* const pos[4], const int nreg[4], int TX,
                   std::vector<std::vector<eio::Entry>>& per_tile, int nbz, const char* what) {
    const int nz = m.g.nz, nx = m.g.nx;
    for (int k = 0; k < 5; ++k) {
        const int gi = GRID_OF_COMP[k];
        const int cnt = nreg[gi];
        if (cnt < 0) fail(std::string(what) + ": negative device count");
        if (cnt > 0 && !pos[gi]) fail(std::string(what) + ": NULL position list");
        if (cnt >= (1 << 24)) fail(std::string(what) + ": too many devices");
        for (int u = 0; u < cnt; ++u) {
            const int p = pos[gi][u];
            if (p < 0 || long(p) >= m.n) fail(std::string(what) + ": device position out of the grid");
            const int iz = p % nz, ix = p / nz;
            if (iz < eio::FAT || iz >= nz - eio::FAT || ix < eio::FAT || ix >= nx - eio::FAT)
                fail(std::string(what) +
                     ": device inside the FAT halo (only update-region positions are supported)");
            const int bz = (iz - eio::FAT) / eio::TZ, lz = (iz - eio::FAT) % eio::TZ;
            const int bx = (ix - eio::FAT) / TX, lx = (ix - eio::FAT) % TX;
            eio::Entry e;
            e.lcell = lz | (lx << 8);
            e.crow = u | (k << 24);
            per_tile[size_t(bx) * nbz + bz].push_back(e);
        }
    }
}

// Entries of a two-step tile: sources of the step-1 region [x0-4, x0+TX+4) x [z0-4, z0+60)
// (every tile whose region holds the source gets it), receivers of the owned cells only;
// lcell relative to (z0-4, x0-4).
void build_entries2(const ModelBase& m, const int* const pos[4], const int nreg[4], int TX, bool halo,
                    std::vector<std::vector<eio::Entry>>& per_tile, int nbz, int nbx, const char* what) {
    const int nz = m.g.nz, nx = m.g.nx;
    for (int k = 0; k < 5; ++k) {
        const int gi = GRID_OF_COMP[k];
        const int cnt = nreg[gi];
        if (cnt < 0) fail(std::string(what) + ": negative device count");
        if (cnt > 0 && !pos[gi]) fail(std::string(what) + ": NULL position list");
        if (cnt >= (1 << 24)) fail(std::string(what) + ": too many devices");
        for (int u = 0; u < cnt; ++u) {
            const int p = pos[gi][u];
            if (p < 0 || long(p) >= m.n) fail(std::string(what) + ": device position out of the grid");
            const int iz = p % nz, ix = p / nz;
            if (iz < eio::FAT || iz >= nz - eio::FAT || ix < eio::FAT || ix >= nx - eio::FAT)
                fail(std::string(what) +
                     ": device inside the FAT halo (only update-region positions are supported)");
            const int h = halo ? eio::FAT : 0;
            const int bz_lo = std::max(0, (iz - eio::FAT - eio::TZ2 - h + 1 + eio::TZ2 - 1) / eio::TZ2);
            const int bz_hi = std::min(nbz - 1, (iz - eio::FAT + h) / eio::TZ2);
            const int bx_lo = std::max(0, (ix - eio::FAT - TX - h + 1 + TX - 1) / TX);
            const int bx_hi = std::min(nbx - 1, (ix - eio::FAT + h) / TX);
            for (int bx = bx_lo; bx <= bx_hi; ++bx)
                for (int bz = bz_lo; bz <= bz_hi; ++bz) {
                    const int z0 = eio::FAT + bz * eio::TZ2, x0 = eio::FAT + bx * TX;
                    const int lz = iz - (z0 - eio::FAT), lx = ix - (x0 - eio::FAT);
                    if (lz < eio::FAT - h || lz >= eio::FAT + eio::TZ2 + h || lx < eio::FAT - h ||
                        lx >= eio::FAT + TX + h)
                        continue;
                    eio::Entry e;
                    e.lcell = lz | (lx << 8);
                    e.crow = u | (k << 24);
                    per_tile[size_t(bx) * nbz + bz].push_back(e);
                }
        }
    }
}

// Entries of a three-step tile: sources of the step-1 region [x0-8, x0+TX+8) x [z0-8, z0+56)
// (every tile whose region holds the source gets it), receivers of the owned cells only;
// lcell relative to (z0-8, x0-8).
void build_entries3(const ModelBase& m, const int* const pos[4], const int nreg[4], int TX, bool halo,
                    std::vector<std::vector<eio::Entry>>& per_tile, int nbz, int nbx, const char* what) {
    const int nz = m.g.nz, nx = m.g.nx;
    constexpr int O = eio::H3, TZ = eio::TZ3;
    for (int k = 0; k < 5; ++k) {
        const int gi = GRID_OF_COMP[k];
        const int cnt = nreg[gi];
        if (cnt < 0) fail(std::string(what) + ": negative device count");
        if (cnt > 0 && !pos[gi]) fail(std::string(what) + ": NULL position list");
        if (cnt >= (1 << 24)) fail(std::string(what) + ": too many devices");
        for (int u = 0; u < cnt; ++u) {
            const int p = pos[gi][u];
            if (p < 0 || long(p) >= m.n) fail(std::string(what) + ": device position out of the grid");
            const int iz = p % nz, ix = p / nz;
            if (iz < eio::FAT || iz >= nz - eio::FAT || ix < eio::FAT || ix >= nx - eio::FAT)
                fail(std::string(what) +
                     ": device inside the FAT halo (only update-region positions are supported)");
            const int h = halo ? O : 0;
            // candidate tiles (a superset of those whose region [FAT + bz TZ - h,
            // FAT + (bz+1) TZ + h) holds iz, likewise along x); the exact test follows
            const int bz_lo = std::max(0, (iz - eio::FAT - h) / TZ - 1);
            const int bz_hi = std::min(nbz - 1, (iz - eio::FAT + h) / TZ);
            const int bx_lo = std::max(0, (ix - eio::FAT - h) / TX - 1);
            const int bx_hi = std::min(nbx - 1, (ix - eio::FAT + h) / TX);
            for (int bx = bx_lo; bx <= bx_hi; ++bx)
                for (int bz = bz_lo; bz <= bz_hi; ++bz) {
                    const int z0 = eio::FAT + bz * TZ, x0 = eio::FAT + bx * TX;
                    const int lz = iz - (z0 - O), lx = ix - (x0 - O);
                    if (lz < O - h || lz >= O + TZ + h || lx < O - h || lx >= O + TX + h) continue;
                    eio::Entry e;
                    e.lcell = lz | (lx << 8);
                    e.crow = u | (k << 24);
                    per_tile[size_t(bx) * nbz + bz].push_back(e);
                }
        }
    }
}

// Tile plans: one-step kernels (64 x TX tiles), the two-step forward (56 x TX tiles, source
// entries on the step-1 region, receivers owned-only) and the two-step adjoint (56 x TX
// tiles, receiver entries on the step-1 region: they are injected in both steps; source
// entries owned-only: the NL adjoint reads them in both steps).
enum class PlanKind { ONE, FWD2, ADJ2, FWD3 };

template <typename T>
Plan make_plan(Ctx* c, int nshot, const eio_shot* shots, Op op, PlanKind kind = PlanKind::ONE, int v2 = 0) {
    const bool two_step = kind == PlanKind::FWD2 || kind == PlanKind::ADJ2;
    const int TX = kind == PlanKind::FWD2   ? F32_VARIANTS2[v2].R2 * F32_VARIANTS2[v2].WX
                   : kind == PlanKind::ADJ2 ? F32_VARIANTS2A[v2].R2 * F32_VARIANTS2A[v2].WX
                   : kind == PlanKind::FWD3 ? F32_VARIANTS3[v2].R3 * F32_VARIANTS3[v2].WX
                                            : tile_x<T>(c->variant);
    const int TZ = kind == PlanKind::FWD3 ? eio::TZ3 : two_step ? eio::TZ2 : eio::TZ;
    Plan p;
    p.order.resize(nshot);
    std::iota(p.order.begin(), p.order.end(), 0);
    std::stable_sort(p.order.begin(), p.order.end(), [&](int a, int b) {
        return get_model(c, shots[a].model).ntw > get_model(c, shots[b].model).ntw;
    });
    for (int oi = 0; oi < nshot; ++oi) {
        const eio_shot& s = shots[p.order[oi]];
        const ModelBase& m = get_model(c, s.model);
        const int nbz = (m.g.nz - 2 * eio::FAT + TZ - 1) / TZ;
        const int nbx = (m.g.nx - 2 * eio::FAT + TX - 1) / TX;
        std::vector<std::vector<eio::Entry>> src_t(size_t(nbz) * nbx), rec_t(size_t(nbz) * nbx);
        if (kind == PlanKind::FWD3) {
            build_entries3(m, s.src_pos, s.n_src, TX, true, src_t, nbz, nbx, "source");
            build_entries3(m, s.rec_pos, s.n_rec, TX, false, rec_t, nbz, nbx, "receiver");
        } else if (kind == PlanKind::FWD2) {
            build_entries2(m, s.src_pos, s.n_src, TX, true, src_t, nbz, nbx, "source");
            build_entries2(m, s.rec_pos, s.n_rec, TX, false, rec_t, nbz, nbx, "receiver");
        } else if (kind == PlanKind::ADJ2) {
            build_entries2(m, s.src_pos, s.n_src, TX, false, src_t, nbz, nbx, "source");
            build_entries2(m, s.rec_pos, s.n_rec, TX, true, rec_t, nbz, nbx, "receiver");
        } else {
            build_entries<T>(m, s.src_pos, s.n_src, TX, src_t, nbz, "source");
            build_entries<T>(m, s.rec_pos, s.n_rec, TX, rec_t, nbz, "receiver");
        }
        for (int bx = 0; bx < nbx; ++bx)
            for (int bz = 0; bz < nbz; ++bz) {
                const size_t ti = size_t(bx) * nbz + bz;
                eio::TileDev t;
                t.shot = oi;
                t.z0 = eio::FAT + bz * TZ;
                t.x0 = eio::FAT + bx * TX;
                t.src_b = int(p.ents.size());
                p.ents.insert(p.ents.end(), src_t[ti].begin(), src_t[ti].end());
                t.src_e = int(p.ents.size());
                t.rec_b = int(p.ents.size());
                p.ents.insert(p.ents.end(), rec_t[ti].begin(), rec_t[ti].end());
                t.rec_e = int(p.ents.size());
                t.pad = 0;
                p.tiles.push_back(t);
            }
        p.tile_end.push_back(int(p.tiles.size()));
        p.nsteps.push_back(m.ntw - 1);
    }
    if (p.tiles.size() >= (1u << 31)) fail("too many tiles");
    (void)op;
    return p;
}

// ------------------------------------------------------------------ execution
template <typename T>
struct Run {
    Ctx* c;
    int v2 = 0;   // two-step forward variant of this call's FWD2 plans
    int v2a = 0;  // two-step adjoint variant of this call's ADJ2 plan
    int v3 = 0;   // three-step forward variant of this call's FWD3 plan
    Plan plan;
    std::vector<eio::ShotDev<T>> sd;
    eio::ShotDev<T>* d_shots = nullptr;
    eio::TileDev* d_tiles = nullptr;
    eio::Entry* d_ents = nullptr;
    // second plan of a Born adjoint / FWI gradient run with the two-step adjoint kernel
    // (the one-step plan above serves the forward passes and the imaging condition)
    Plan plan2;
    eio::TileDev* d_tiles2 = nullptr;
    eio::Entry* d_ents2 = nullptr;
    // two-step forward plan of the Born source pass (Born runs whose other passes take one step)
    Plan plan3;
    eio::TileDev* d_tiles3 = nullptr;
    eio::Entry* d_ents3 = nullptr;
    // host copies for the D2H of outputs
    struct Out {
        void* host;
        T* dev;
        size_t bytes;
    };
    std::vector<Out> outs;
};

// The two-step forward variant of a call.  EIO_VARIANT_AUTO: a batch that fills the chip with
// fewer than AUTO_SMALL_TILES tiles of 56 x 32 (four per CU) takes 56 x 8 tiles, which spread
// its launches over every CU: the streamed stage at 8 waves per SIMD (variant 12; the
// coarse-record pass at 6, the scattered pass, which has no streamed form, variant 3's row-lane
// 56 x 8 kernel); larger batches take 56 x 32 (variant 4), whose wider tiles re-read less
// halo, unless they run in shot groups.  Measured: 8 Born shots of 280 x 472 (600 wide tiles)
// run their source and scattered passes 11 % faster on 56 x 8 than on 56 x 32, and 8 NL
// forward shots of that grid 16 % faster on variant 12 than on variant 3 (one shot: the same);
// the 64-shot layered batch in one group runs 19 % slower on 56 x 8.
constexpr int AUTO_SMALL_TILES = 1024;
// shot-group working-set budget of EIO_GROUP_AUTO, of the MI355X's 256 MB Infinity Cache
constexpr long long GROUP_AUTO_BYTES = 192ll << 20;
// tiles of 56 x 32 over the update regions of a batch
long wide_tiles(Ctx* c, int nshot, const eio_shot* shots) {
    long tiles = 0;
    for (int i = 0; i < nshot; ++i) {
        const ModelBase& m = get_model(c, shots[i].model);
        tiles += long((m.g.nz - 2 * eio::FAT + eio::TZ2 - 1) / eio::TZ2) * ((m.g.nx - 2 * eio::FAT + 31) / 32);
    }
    return tiles;
}
// `grouped`: a plain NL forward whose shots run in more than one shot group; their data stays
// in the Infinity Cache between launches, the kernel is bound by the latency of its load
// phases, and 56 x 8 tiles with the streamed stage at 8 waves per SIMD (variant 12, four
// workgroups per CU) are the fastest form (64-shot batch: 0.502 ms per fine step against 0.518
// for 56 x 16 at 6 waves and 0.633 for 56 x 32).
int variant2_for(Ctx* c, int nshot, const eio_shot* shots, bool grouped = false) {
    if (c->variant2 != EIO_VARIANT_AUTO) return c->variant2;
    if (grouped) return 12;
    return wide_tiles(c, nshot, shots) < AUTO_SMALL_TILES ? 12 : 4;
}

// The two-step adjoint variant of a call: EIO_VARIANT_AUTO takes 56 x 8 tiles (variant 10)
// under the same threshold, else 56 x 32 (variant 7).
int variant2a_for(Ctx* c, int nshot, const eio_shot* shots) {
    if (c->variant2a != EIO_VARIANT_AUTO) return c->variant2a;
    return wide_tiles(c, nshot, shots) < AUTO_SMALL_TILES ? 10 : 7;
}

// The three-step forward variant of a call: EIO_VARIANT_AUTO takes 48 x 8 tiles (variant 2,
// two workgroups per CU) for a batch of fewer than AUTO_SMALL_TILES tiles of 56 x 32, else
// 48 x 32 (variant 1).
int variant3_for(Ctx* c, int nshot, const eio_shot* shots) {
    if (c->variant3 != EIO_VARIANT_AUTO) return c->variant3;
    return wide_tiles(c, nshot, shots) < AUTO_SMALL_TILES ? 2 : 1;
}

// Whether a plain FP32 NL forward of these shots splits into more than one shot group: their
// working set (`per_cell` bytes per cell) exceeds the budget of one group.
bool nl_grouped(Ctx* c, int nshot, const eio_shot* shots, long long per_cell, long long budget) {
    if (budget <= 0) return false;
    long long total = 0;
    for (int i = 0; i < nshot; ++i) {
        const ModelBase& m = get_model(c, shots[i].model);
        total += (long long)m.g.nz * m.g.nx * per_cell;
    }
    return total > budget;
}

// Builds the plan, the per-shot descriptors and the device tables of one call, uploads
// input traces and zeroes the state and the outputs.  FWI calls bring their own coarse
// storage ring (`ring`, `wcap` slices per component and shot).  `two_step`: the NL forward
// runs the two-step kernel; `adj2`: the adjoint passes run the two-step adjoint kernel;
// `src2`: the Born source pass runs the two-step forward kernel; `grouped`: a plain NL forward
// that runs in more than one shot group (its automatic two-step variant differs).
template <typename T>
Run<T> prepare(Ctx* c, int nshot, const eio_shot* shots, Op op, int flags,
               const void* const dm_fwd[5], T* ring = nullptr, int wcap = 0, bool two_step = false,
               bool adj2 = false, bool src2 = false, bool three_step = false, bool grouped = false) {
    if (nshot <= 0) fail("nshot must be > 0");
    if (!shots) fail("shots is NULL");
    // the learned boundary is part of the NL forward only: the adjoint, Born and FWI passes
    // would silently run another operator than the forward they are paired with
    if (op != Op::NL_FWD)
        for (int i = 0; i < nshot; ++i)
            if (get_model(c, shots[i].model).abc_on)
                fail("a learned boundary is attached to model " + std::to_string(shots[i].model) +
                     ": it applies to the nonlinear forward only; detach it (eio_model_set_abc with on = 0) "
                     "for adjoint, Born and FWI calls");
    Run<T> run;
    run.c = c;
    run.v2 = variant2_for(c, nshot, shots, grouped);
    run.v2a = variant2a_for(c, nshot, shots);
    run.v3 = variant3_for(c, nshot, shots);
    run.plan = three_step ? make_plan<T>(c, nshot, shots, op, PlanKind::FWD3, run.v3)
                          : make_plan<T>(c, nshot, shots, op, two_step ? PlanKind::FWD2 : PlanKind::ONE, run.v2);
    if (adj2) run.plan2 = make_plan<T>(c, nshot, shots, op, PlanKind::ADJ2, run.v2a);
    if (src2) run.plan3 = make_plan<T>(c, nshot, shots, op, PlanKind::FWD2, run.v2);
    const int nlev = (two_step || three_step || adj2 || src2) ? 4 : 2;  // state levels per shot
    const bool born = (op == Op::BORN_FWD || op == Op::BORN_ADJ);
    const bool fwi = (op == Op::FWI_FWD || op == Op::FWI_GRAD);
    const bool has_ss = born || op == Op::FWI_GRAD;
    const bool has_acc = op == Op::BORN_ADJ || op == Op::FWI_GRAD;
    if (born || fwi) {
        for (int i = 1; i < nshot; ++i)
            if (shots[i].model != shots[0].model) fail("Born and FWI shots of one call must share a model");
        if (get_model(c, shots[0].model).g.surface)
            fail("the free surface is available for the nonlinear operator only");
    }
    if (fwi && !ring) fail("FWI call without a coarse storage ring");
    const bool dev_ptrs = (flags & EIO_DEVICE_PTRS) != 0;

    // ---- sizes
    size_t state_elems = 0, tr_elems = 0, ss_elems = 0, w_elems = 0, acc_elems = 0;
    for (int oi = 0; oi < nshot; ++oi) {
        const eio_shot& s = shots[run.plan.order[oi]];
        const ModelBase& m = get_model(c, s.model);
        state_elems += 5 * nlev * size_t(m.n);
        if (has_ss) ss_elems += 10 * size_t(m.n);
        if (born) w_elems += 2 * size_t(m.n) * m.g.nts;
        if (has_acc) acc_elems += 5 * size_t(m.n);
        if (!dev_ptrs) {
            const int src_ld = born ? m.ntw : m.g.nts;
            for (int k = 0; k < 5; ++k) {
                tr_elems += size_t(s.n_src[GRID_OF_COMP[k]]) * src_ld;
                tr_elems += size_t(s.n_rec[GRID_OF_COMP[k]]) * m.g.nts;
            }
        }
    }
    T* state = static_cast<T*>(ws_get(c, "state", state_elems * sizeof(T)));
    T* trbuf = tr_elems ? static_cast<T*>(ws_get(c, "traces", tr_elems * sizeof(T))) : nullptr;
    T* ssbuf = ss_elems ? static_cast<T*>(ws_get(c, "ss", ss_elems * sizeof(T))) : nullptr;
    T* wbuf = w_elems ? static_cast<T*>(ws_get(c, "W", w_elems * sizeof(T))) : nullptr;
    T* accbuf = acc_elems ? static_cast<T*>(ws_get(c, "acc", acc_elems * sizeof(T))) : nullptr;

    // ---- shot descriptors
    run.sd.resize(nshot);
    size_t so = 0, to = 0, sso = 0, wo = 0, ao = 0;
    const hipMemcpyKind h2d = hipMemcpyHostToDevice;
    for (int oi = 0; oi < nshot; ++oi) {
        const eio_shot& s = shots[run.plan.order[oi]];
        const ModelBase& m = get_model(c, s.model);
        eio::ShotDev<T>& d = run.sd[oi];
        std::memset(&d, 0, sizeof(d));
        d.nz = m.g.nz; d.nx = m.g.nx; d.nts = m.g.nts; d.sub = m.g.sub_fd; d.h = m.h;
        d.ntw = m.ntw; d.minPad = m.g.minPad; d.surface = m.g.surface;
        for (int k = 0; k < 4; ++k) { d.cz[k] = T(m.cz[k]); d.cx[k] = T(m.cx[k]); }
        d.A = state + so;
        d.B = state + so + 5 * size_t(m.n);
        if (two_step || three_step || adj2 || src2) {
            d.C = state + so + 10 * size_t(m.n);
            d.D = state + so + 15 * size_t(m.n);
        }
        d.med = static_cast<const T*>(m.med);
        so += 5 * nlev * size_t(m.n);
        d.damp = static_cast<const T*>(m.damp);
        d.filt = static_cast<const T*>(m.filt);
        const int src_ld = born ? m.ntw : m.g.nts;
        const bool src_in = (op != Op::NL_ADJ);
        const bool rec_in = (op == Op::NL_ADJ || op == Op::BORN_ADJ || op == Op::FWI_GRAD);
        for (int k = 0; k < 5; ++k) {
            const size_t ns = size_t(s.n_src[GRID_OF_COMP[k]]) * src_ld;
            const size_t nr = size_t(s.n_rec[GRID_OF_COMP[k]]) * m.g.nts;
            if (ns && !s.src_tr[k]) fail("source trace pointer is NULL");
            if (nr && !s.rec_tr[k]) fail("receiver trace pointer is NULL");
            if (dev_ptrs) {
                d.src_tr[k] = static_cast<T*>(s.src_tr[k]);
                d.rec_tr[k] = static_cast<T*>(s.rec_tr[k]);
                if (!src_in) fill0<T>(c, d.src_tr[k], long(ns));
                if (!rec_in) fill0<T>(c, d.rec_tr[k], long(nr));
            } else {
                d.src_tr[k] = trbuf + to;
                to += ns;
                d.rec_tr[k] = trbuf + to;
                to += nr;
                if (src_in) {
                    if (ns) EIO_CHECK_HIP(hipMemcpyAsync(d.src_tr[k], s.src_tr[k], ns * sizeof(T), h2d, c->stream));
                } else {
                    fill0<T>(c, d.src_tr[k], long(ns));
                    run.outs.push_back({s.src_tr[k], d.src_tr[k], ns * sizeof(T)});
                }
                if (rec_in) {
                    if (nr) EIO_CHECK_HIP(hipMemcpyAsync(d.rec_tr[k], s.rec_tr[k], nr * sizeof(T), h2d, c->stream));
                } else {
                    fill0<T>(c, d.rec_tr[k], long(nr));
                    run.outs.push_back({s.rec_tr[k], d.rec_tr[k], nr * sizeof(T)});
                }
            }
        }
        if (has_ss) {
            d.ss[0] = ssbuf + sso;
            d.ss[1] = ssbuf + sso + 5 * size_t(m.n);
            sso += 10 * size_t(m.n);
        }
        if (born) {
            d.Wvx = wbuf + wo;
            d.Wvz = wbuf + wo + size_t(m.n) * m.g.nts;
            d.wcap = m.g.nts;
            wo += 2 * size_t(m.n) * m.g.nts;
        } else if (fwi) {
            d.Wvx = ring + wo;
            d.Wvz = ring + wo + size_t(m.n) * wcap;
            d.wcap = wcap;
            wo += 2 * size_t(m.n) * wcap;
        }
        if (has_acc) {
            d.dm = accbuf + ao;
            ao += 5 * size_t(m.n);
        } else if (op == Op::BORN_FWD) {
            d.dm = const_cast<T*>(static_cast<const T*>(dm_fwd[0]));
        }
    }
    // ---- tables
    run.d_shots = static_cast<eio::ShotDev<T>*>(ws_get(c, "shots", sizeof(eio::ShotDev<T>) * nshot));
    run.d_tiles = static_cast<eio::TileDev*>(ws_get(c, "tiles", sizeof(eio::TileDev) * run.plan.tiles.size()));
    run.d_ents = static_cast<eio::Entry*>(ws_get(c, "ents", sizeof(eio::Entry) * std::max<size_t>(1, run.plan.ents.size())));
    EIO_CHECK_HIP(hipMemcpyAsync(run.d_shots, run.sd.data(), sizeof(eio::ShotDev<T>) * nshot, h2d, c->stream));
    EIO_CHECK_HIP(hipMemcpyAsync(run.d_tiles, run.plan.tiles.data(), sizeof(eio::TileDev) * run.plan.tiles.size(), h2d, c->stream));
    if (!run.plan.ents.empty())
        EIO_CHECK_HIP(hipMemcpyAsync(run.d_ents, run.plan.ents.data(), sizeof(eio::Entry) * run.plan.ents.size(), h2d, c->stream));
    if (src2) {
        const Plan& p3 = run.plan3;
        run.d_tiles3 = static_cast<eio::TileDev*>(ws_get(c, "tiles3", sizeof(eio::TileDev) * p3.tiles.size()));
        run.d_ents3 = static_cast<eio::Entry*>(ws_get(c, "ents3", sizeof(eio::Entry) * std::max<size_t>(1, p3.ents.size())));
        EIO_CHECK_HIP(hipMemcpyAsync(run.d_tiles3, p3.tiles.data(), sizeof(eio::TileDev) * p3.tiles.size(), h2d, c->stream));
        if (!p3.ents.empty())
            EIO_CHECK_HIP(hipMemcpyAsync(run.d_ents3, p3.ents.data(), sizeof(eio::Entry) * p3.ents.size(), h2d, c->stream));
    }
    if (adj2) {
        const Plan& p2 = run.plan2;
        run.d_tiles2 = static_cast<eio::TileDev*>(ws_get(c, "tiles2", sizeof(eio::TileDev) * p2.tiles.size()));
        run.d_ents2 = static_cast<eio::Entry*>(ws_get(c, "ents2", sizeof(eio::Entry) * std::max<size_t>(1, p2.ents.size())));
        EIO_CHECK_HIP(hipMemcpyAsync(run.d_tiles2, p2.tiles.data(), sizeof(eio::TileDev) * p2.tiles.size(), h2d, c->stream));
        if (!p2.ents.empty())
            EIO_CHECK_HIP(hipMemcpyAsync(run.d_ents2, p2.ents.data(), sizeof(eio::Entry) * p2.ents.size(), h2d, c->stream));
    }
    // ---- zero state (and Born buffers; the coarse storage needs no clearing: the
    // decimation overwrites each slot with its first contribution)
    fill0<T>(c, state, long(state_elems));
    if (ssbuf) fill0<T>(c, ssbuf, long(ss_elems));
    if (accbuf) fill0<T>(c, accbuf, long(acc_elems));
    return run;
}

template <typename T>
void finish(Run<T>& run) {
    Ctx* c = run.c;
    for (auto& o : run.outs)
        if (o.bytes) EIO_CHECK_HIP(hipMemcpyAsync(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost, c->stream));
    EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
}

// number of tiles active at step k (shots sorted by decreasing step count)
inline int active_tiles(const Plan& p, int k) {
    int n = 0;
    for (size_t i = 0; i < p.nsteps.size(); ++i)
        if (k < p.nsteps[i]) n = p.tile_end[i];
    return n;
}

// Shots [s0, s1) of a plan (in plan order) that run their steps together.
struct ShotGroup {
    int s0, s1;
};

// Consecutive groups of the plan's shots whose working set (`per_cell` bytes per cell of the
// grid) fits `budget` bytes (EIO_GROUP_AUTO: GROUP_AUTO_BYTES; 0: one group); a shot larger
// than the budget is a group of its own.
template <typename T>
std::vector<ShotGroup> shot_groups(const Run<T>& run, long long budget, long long per_cell) {
    const int n = int(run.plan.nsteps.size());
    if (budget == EIO_GROUP_AUTO) budget = GROUP_AUTO_BYTES;
    if (budget <= 0 || n == 0) return {{0, n}};
    std::vector<ShotGroup> g;
    int s0 = 0;
    long long used = 0;
    for (int i = 0; i < n; ++i) {
        const long long b = (long long)run.sd[i].nz * run.sd[i].nx * per_cell;
        if (i > s0 && used + b > budget) {
            g.push_back({s0, i});
            s0 = i;
            used = 0;
        }
        used += b;
    }
    g.push_back({s0, n});
    return g;
}

// tiles of group g active at step k, counted from the group's first tile (the group's shots
// are sorted by decreasing step count, so the active ones come first)
inline int group_tiles(const Plan& p, const ShotGroup& g, int k) {
    const int t0 = g.s0 ? p.tile_end[g.s0 - 1] : 0;
    int n = 0;
    for (int i = g.s0; i < g.s1; ++i)
        if (k < p.nsteps[i]) n = p.tile_end[i] - t0;
    return n;
}

#define EIO_F32_VARIANT_SWITCH(V, CALL)                                                  \
    switch (V) {                                                                          \
        case 0: CALL(4, 4, 2); break;                                                     \
        case 1: CALL(2, 8, 4); break;                                                     \
        case 2: CALL(2, 4, 4); break;                                                     \
        case 3: CALL(4, 8, 2); break;                                                     \
        case 4: CALL(8, 4, 1); break;                                                     \
        default: CALL(2, 8, 2); break;                                                    \
    }

// dlo/dhi: decimation segment bounds (coarse index of the first interval, coarse index of
// the slot shared with the next segment, or -1); `dsh` replaces the run's shot table.
template <typename T, int MODE>
void launch_fwd(Run<T>& run, int ntile, int k, int sspar, int dlo = 0, int dhi = -1,
                const eio::ShotDev<T>* dsh = nullptr, const eio::TileDev* tiles = nullptr) {
    const eio::ShotDev<T>* tab = dsh ? dsh : run.d_shots;
    const eio::TileDev* tt = tiles ? tiles : run.d_tiles;
#define EIO_L(R_, WX_, OCC_)                                                                  \
    klaunch(run.c, eio::k_step_fwd<T, R_, WX_, OCC_, MODE>, dim3(ntile), dim3(64 * WX_), tab, tt, \
            run.d_ents, ntile, k, sspar, dlo, dhi)
    if constexpr (sizeof(T) == 8) {
        EIO_L(2, 4, 1);
    } else {
        EIO_F32_VARIANT_SWITCH(run.c->variant, EIO_L)
    }
#undef EIO_L
}
template <typename T, int MODE>
void launch_adj(Run<T>& run, int ntile, int k, int sspar) {
#define EIO_L(R_, WX_, OCC_)                                                                  \
    klaunch(run.c, eio::k_step_adj<T, R_, WX_, OCC_, MODE>, dim3(ntile), dim3(64 * WX_), run.d_shots, \
            run.d_tiles, run.d_ents, ntile, k, sspar)
    if constexpr (sizeof(T) == 8) {
        EIO_L(2, 4, 1);
    } else {
        EIO_F32_VARIANT_SWITCH(run.c->variant, EIO_L)
    }
#undef EIO_L
}
// `tiles`, `ents`: another plan's tables (the Born passes of a one-step run); `sspar`: the
// secondary-source slot of coarse step its (SS)
template <bool DEC = false, bool FINE = false, bool SS = false, bool NOREC = false>
void launch_fwd2(Run<float>& run, int ntile, int k, int par, int dlo = 0, int dhi = -1,
                 const eio::TileDev* tiles = nullptr, const eio::Entry* ents = nullptr, int sspar = 0,
                 const eio::ShotDev<float>* dsh = nullptr) {
    const eio::TileDev* tt = tiles ? tiles : run.d_tiles;
    const eio::Entry* te = ents ? ents : run.d_ents;
    const eio::ShotDev<float>* sh = dsh ? dsh : run.d_shots;
#define EIO_L2(R_, WX_, OCC_, NT_)                                                                    \
    klaunch(run.c, eio::k_step2_fwd<float, R_, WX_, OCC_, NT_, DEC, FINE, SS, NOREC>, dim3(ntile),      \
            dim3(64 * WX_), sh, tt, te, ntile, k, par, dlo, dhi, sspar)
// streamed-stage variants (layout 1): the coarse-record (DEC) pass runs the streamed kernel at
// occupancy DOCC_ (its slot registers do not fit the plain pass's budget without spills; 0:
// no streamed form), the scattered (SS) pass, which has none, the row-lane kernel of the same
// tile width (R_ * WX_ = FR_ * FWX_: the plan's tiles and entries were built for that width)
#define EIO_L2FS(R_, WX_, OCC_, DOCC_, FR_, FWX_, FOCC_)                                              \
    static_assert(R_ * WX_ == FR_ * FWX_, "fallback kernel of another tile width");                 \
    if constexpr (!DEC && !SS)                                                                        \
        klaunch(run.c, eio::k_step2_fwd<float, R_, WX_, OCC_, 0, false, FINE, false, NOREC, true>,      \
                dim3(ntile), dim3(64 * WX_), sh, tt, te, ntile, k, par, dlo, dhi, sspar);             \
    else if constexpr (DEC && !SS && DOCC_ > 0)                                                      \
        klaunch(run.c, eio::k_step2_fwd<float, R_, WX_, DOCC_, 0, true, FINE, false, NOREC, true>,      \
                dim3(ntile), dim3(64 * WX_), sh, tt, te, ntile, k, par, dlo, dhi, sspar);             \
    else                                                                                              \
        EIO_L2(FR_, FWX_, FOCC_, 0)
#define EIO_L2Q(OCC_)                                                                                 \
    if constexpr (!DEC && !SS && !FINE)                                                               \
        klaunch(run.c, eio::k_step2q_fwd<float, OCC_, NOREC>, dim3(ntile), dim3(eio::QNT), sh, tt, te, ntile, \
                k, par);                                                                              \
    else                                                                                              \
        EIO_L2(3, 8, 2, 0)
    switch (run.v2) {
        case 0: EIO_L2(2, 8, 2, 0); break;
        case 1: EIO_L2(2, 8, 4, 0); break;
        case 2: EIO_L2(4, 4, 2, 0); break;
        case 3: EIO_L2(1, 8, 4, 0); break;
        case 4: EIO_L2(4, 8, 1, 0); break;
        case 5: EIO_L2(2, 8, 4, 1); break;
        case 6: EIO_L2(2, 8, 2, 1); break;
        case 7: EIO_L2(4, 8, 1, 1); break;
        case 8: EIO_L2FS(4, 8, 4, 4, 4, 8, 1); break;
        case 9: EIO_L2FS(8, 4, 2, 0, 4, 8, 1); break;
        case 10: EIO_L2FS(2, 8, 6, 6, 2, 8, 4); break;
        case 12: EIO_L2FS(1, 8, 8, 6, 1, 8, 4); break;
        default: EIO_L2Q(2); break;
    }
#undef EIO_L2Q
#undef EIO_L2FS
#undef EIO_L2
}

// Fine steps k, k+1, k+2 of the FP32 NL forward (levels A,B -> C,D when par = 0); `tiles`
// replaces the run's tile table (a shot group's first tile).
void launch_fwd3(Run<float>& run, int ntile, int k, int par, const eio::TileDev* tiles = nullptr) {
    if (ntile <= 0) return;
    const eio::TileDev* tt = tiles ? tiles : run.d_tiles;
    switch (run.v3) {
        case 0:
            klaunch(run.c, eio::k_step3_fwd<float, 3, 8, 1>, dim3(ntile), dim3(512), run.d_shots, tt,
                    run.d_ents, ntile, k, par);
            break;
        case 2:
            klaunch(run.c, eio::k_step3_fwd<float, 1, 8, 4>, dim3(ntile), dim3(512), run.d_shots, tt,
                    run.d_ents, ntile, k, par);
            break;
        default:
            klaunch(run.c, eio::k_step3_fwd<float, 4, 8, 1>, dim3(ntile), dim3(512), run.d_shots, tt,
                    run.d_ents, ntile, k, par);
            break;
    }
}

// Reverse fine steps k, k+1 (k even) of the Born receiver pass (adjoint-field slots as in
// launch_adj, sspar: slot of coarse step its) or of the NL adjoint.
template <int MODE = eio::M_BADJ>
void launch_adj2(Run<float>& run, int k, int sspar) {
    const int ntile = int(run.plan2.tiles.size());
    const int par = (k >> 1) & 1;
#define EIO_L2A(R_, WX_, OCC_, LATE_, PS_)                                                               \
    klaunch(run.c, eio::k_step2_adj<float, R_, WX_, OCC_, LATE_, PS_, MODE>, dim3(ntile), dim3(64 * WX_),   \
            run.d_shots, run.d_tiles2, run.d_ents2, ntile, k, par, sspar)
    switch (run.v2a) {
        case 0: EIO_L2A(2, 8, 2, false, true); break;
        case 1: EIO_L2A(4, 8, 1, false, true); break;
        case 2: EIO_L2A(2, 8, 4, false, true); break;
        case 3: EIO_L2A(4, 4, 2, false, true); break;
        case 4: EIO_L2A(2, 8, 4, true, true); break;
        case 5: EIO_L2A(2, 8, 2, true, true); break;
        case 6: EIO_L2A(4, 8, 1, true, true); break;
        case 7: EIO_L2A(4, 8, 1, true, false); break;
        case 8: EIO_L2A(4, 8, 1, false, false); break;
        case 9: EIO_L2A(2, 8, 2, true, false); break;
        case 10: EIO_L2A(1, 8, 4, true, false); break;
        default: EIO_L2A(1, 8, 4, false, true); break;
    }
#undef EIO_L2A
}

// The Born adjoint pass over coarse interval its (reverse fine steps (nts-2-its)*sub ..),
// adjoint fields of the interval into slots ss[p] (its) and ss[p^1] (its+1).
template <typename T>
void adjoint_interval(Run<T>& run, int ntile, bool two, int its, int nts, int sub, int p) {
    const int k0 = (nts - 2 - its) * sub;
    if constexpr (sizeof(T) == 4) {
        if (two) {
            for (int j = 0; j < sub; j += 2) launch_adj2(run, k0 + j, p);
            return;
        }
    }
    for (int j = 0; j < sub; ++j) launch_adj<T, eio::M_BADJ>(run, ntile, k0 + j, p);
}

template <typename T, bool ADJ>
void launch_img(Run<T>& run, int ntile, int its, int which) {
#define EIO_L(R_, WX_, OCC_)                                                                  \
    klaunch(run.c, eio::k_imaging<T, R_, WX_, ADJ>, dim3(ntile), dim3(64 * WX_), run.d_shots, run.d_tiles, \
            ntile, its, which)
    if constexpr (sizeof(T) == 8) {
        EIO_L(2, 4, 1);
    } else {
        EIO_F32_VARIANT_SWITCH(run.c->variant, EIO_L)
    }
#undef EIO_L
}

// ------------------------------------------------------------------ graph mode
// A time loop is recorded as a list of launches, then run from an executable hipGraph (a
// chain of kernel nodes) cached per launch-sequence shape: the first call of a shape builds
// and instantiates it; later calls with the same kernels, grids and argument sizes replay it
// after updating only the kernel nodes whose arguments changed.  Device tables (shot, tile
// and entry tables) are uploaded by each call before the graph launch, on the same stream, so
// a replay reads the call's own tables.  No memset nodes: buffers are cleared by kernels.
constexpr size_t GRAPH_CACHE_MAX = 4;

unsigned long long log_shape(const std::vector<LaunchRec>& log) {
    unsigned long long h = 1469598103934665603ULL;
    auto mix = [&](unsigned long long v) {
        h ^= v;
        h *= 1099511628211ULL;
    };
    mix(log.size());
    for (const LaunchRec& r : log) {
        mix(reinterpret_cast<unsigned long long>(r.fn));
        mix(r.grid.x), mix(r.grid.y), mix(r.grid.z), mix(r.block.x), mix(r.block.y), mix(r.block.z);
        mix(r.blob.size());
    }
    return h;
}

bool same_shape(const std::vector<LaunchRec>& a, const std::vector<LaunchRec>& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i) {
        const LaunchRec &x = a[i], &y = b[i];
        if (x.fn != y.fn || x.blob.size() != y.blob.size() || x.grid.x != y.grid.x || x.grid.y != y.grid.y ||
            x.grid.z != y.grid.z || x.block.x != y.block.x || x.block.y != y.block.y || x.block.z != y.block.z)
            return false;
    }
    return true;
}

void graph_free(GraphEntry& e) {
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (e.graph) (void)hipGraphDestroy(e.graph);
    e.exec = nullptr;
    e.graph = nullptr;
    e.nodes.clear();
}

hipKernelNodeParams node_params(LaunchRec& r, std::vector<void*>& ptrs) {
    ptrs.resize(r.offs.size());
    for (size_t i = 0; i < r.offs.size(); ++i) ptrs[i] = r.blob.data() + r.offs[i];
    hipKernelNodeParams p{};
    p.func = const_cast<void*>(r.fn);
    p.gridDim = r.grid;
    p.blockDim = r.block;
    p.sharedMemBytes = 0;
    p.kernelParams = ptrs.data();
    p.extra = nullptr;
    return p;
}

void graph_launch(Ctx* c) {
    std::vector<LaunchRec> log;
    log.swap(c->log);
    if (log.empty()) return;
    const unsigned long long key = log_shape(log);
    auto it = c->graphs.find(key);
    if (it != c->graphs.end() && !same_shape(it->second.log, log)) {  // hash collision
        graph_free(it->second);
        c->graphs.erase(it);
        it = c->graphs.end();
    }
    std::vector<void*> ptrs;
    if (it == c->graphs.end()) {
        if (c->graphs.size() >= GRAPH_CACHE_MAX) {  // evict the least recently used graph
            auto lru = c->graphs.begin();
            for (auto j = c->graphs.begin(); j != c->graphs.end(); ++j)
                if (j->second.last_use < lru->second.last_use) lru = j;
            graph_free(lru->second);
            c->graphs.erase(lru);
        }
        GraphEntry e;
        e.log.swap(log);
        try {
            EIO_CHECK_HIP(hipGraphCreate(&e.graph, 0));
            e.nodes.resize(e.log.size());
            for (size_t i = 0; i < e.log.size(); ++i) {
                const hipKernelNodeParams p = node_params(e.log[i], ptrs);
                EIO_CHECK_HIP(hipGraphAddKernelNode(&e.nodes[i], e.graph, i ? &e.nodes[i - 1] : nullptr, i ? 1 : 0,
                                                    &p));
            }
            EIO_CHECK_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
        } catch (...) {
            graph_free(e);
            throw;
        }
        ++c->graph_builds;
        it = c->graphs.emplace(key, std::move(e)).first;
    } else {
        GraphEntry& e = it->second;
        for (size_t i = 0; i < log.size(); ++i) {
            if (log[i].blob == e.log[i].blob) continue;
            e.log[i].blob.swap(log[i].blob);
            const hipKernelNodeParams p = node_params(e.log[i], ptrs);
            EIO_CHECK_HIP(hipGraphExecKernelNodeSetParams(e.exec, e.nodes[i], &p));
            ++c->graph_node_updates;
        }
    }
    it->second.last_use = ++c->graph_clock;
    EIO_CHECK_HIP(hipGraphLaunch(it->second.exec, c->stream));
    ++c->graph_launches;
}

// Runs `body` (a sequence of launches) directly, or, in graph mode, records it and runs it
// from the context's cached executable graph.  `graph_ok` = false for bodies that issue
// anything but kernel launches (they always run directly).
template <typename F>
void run_loop(Ctx* c, F&& body, bool graph_ok = true) {
    if (!c->use_graph || !graph_ok) {
        body();
        EIO_CHECK_HIP(hipGetLastError());
        return;
    }
    c->log.clear();
    c->recording = true;
    try {
        body();
    } catch (...) {
        c->recording = false;
        c->log.clear();
        throw;
    }
    c->recording = false;
    graph_launch(c);
}

template <typename T>
void timing_begin(Ctx* c) { EIO_CHECK_HIP(hipEventRecord(c->ev[0], c->stream)); }
template <typename T>
void timing_end(Ctx* c, long long nstep) {
    EIO_CHECK_HIP(hipEventRecord(c->ev[1], c->stream));
    EIO_CHECK_HIP(hipEventSynchronize(c->ev[1]));
    float tot = 0.f, stp = 0.f;
    EIO_CHECK_HIP(hipEventElapsedTime(&tot, c->ev[0], c->ev[1]));
    EIO_CHECK_HIP(hipEventElapsedTime(&stp, c->ev[2], c->ev[3]));
    c->last_total_ms = tot;
    c->last_nstep = nstep;
    c->last_step_ms = nstep > 0 ? double(stp) / double(nstep) : 0.0;
}

// Saved wavefield of one shot of a call: [nts][5][nz*nx] on the device, zeroed, attached to
// the shot's descriptor (re-uploaded); copied to `out` (host, or device with EIO_DEVICE_PTRS)
// when the call finishes.
template <typename T>
void attach_wavefield(Run<T>& run, int shot, void* out, int flags) {
    Ctx* c = run.c;
    if (!out) fail("wavefield pointer is NULL");
    const int ns = int(run.sd.size());
    if (shot < 0 || shot >= ns) fail("wavefield shot index out of range");
    int oi = 0;
    while (run.plan.order[oi] != shot) ++oi;
    eio::ShotDev<T>& d = run.sd[oi];
    const size_t elems = size_t(d.nts) * 5 * size_t(d.nz) * d.nx;
    T* w = (flags & EIO_DEVICE_PTRS) ? static_cast<T*>(out) : static_cast<T*>(ws_get(c, "wfld", elems * sizeof(T)));
    fill0<T>(c, w, long(elems));
    d.wfld = w;
    if (!(flags & EIO_DEVICE_PTRS)) run.outs.push_back({out, w, elems * sizeof(T)});
    EIO_CHECK_HIP(hipMemcpyAsync(run.d_shots + oi, &run.sd[oi], sizeof(eio::ShotDev<T>), hipMemcpyHostToDevice,
                                 c->stream));
}

// Learned absorbing boundary of the shots of a run whose model has one attached: frame tile
// list (grouped by ordered shot, so the tiles of the active shots are a prefix) and per-shot
// network parameters.
struct AbcRun {
    int ntile = 0;
    std::vector<int> shot_end;  // per ordered shot: end of its tiles
    std::vector<int> nsteps;
    eio::AbcTile* d_tiles = nullptr;
    eio::AbcShot* d_par = nullptr;
    int active(int k) const {
        int e = 0;
        for (size_t i = 0; i < nsteps.size(); ++i)
            if (k < nsteps[i]) e = shot_end[i];
        return e;
    }
};

bool any_abc(Ctx* c, int nshot, const eio_shot* shots) {
    for (int i = 0; i < nshot && shots; ++i)
        if (get_model(c, shots[i].model).abc_on) return true;
    return false;
}

template <typename T>
AbcRun abc_prepare(Run<T>& run, const eio_shot* shots) {
    AbcRun a;
    if constexpr (sizeof(T) != 4) {
        return a;
    } else {
        Ctx* c = run.c;
        const int ns = int(run.sd.size());
        std::vector<eio::AbcTile> tiles;
        std::vector<eio::AbcShot> par(ns);
        a.shot_end.resize(ns);
        a.nsteps = run.plan.nsteps;
        const int F = eio::FAT;
        for (int oi = 0; oi < ns; ++oi) {
            const ModelBase& m = get_model(c, shots[run.plan.order[oi]].model);
            const int nz = m.g.nz, nx = m.g.nx;
            if (m.abc_on) {
                eio::AbcShot& P = par[oi];
                P.w = static_cast<const unsigned short*>(m.abc_w);
                for (int f = 0; f < 5; ++f) P.scale[f] = m.abc_scale[f];
                for (int f = 0; f < 10; ++f) P.inv_scale[f] = 1.0f / m.abc_scale[f];
                P.diff = m.abc_on == 1;
                P.depth = m.abc_depth;
                for (int e = 0; e < 2; ++e)  // x-min, x-max frames: the update rows
                    for (int z0 = F; z0 < nz - F; z0 += eio::ABC_MAXOZ)
                        tiles.push_back({oi, e, e == 0 ? 0 : nx - F, z0, F, std::min(eio::ABC_MAXOZ, nz - F - z0)});
                for (int e = 2; e < 4; ++e)  // z-min, z-max frames: the update columns
                    for (int x0 = F; x0 < nx - F; x0 += eio::ABC_MAXOX)
                        tiles.push_back({oi, e, x0, e == 2 ? 0 : nz - F, std::min(eio::ABC_MAXOX, nx - F - x0), F});
            }
            a.shot_end[oi] = int(tiles.size());
        }
        a.ntile = int(tiles.size());
        if (a.ntile == 0) return a;
        a.d_tiles = static_cast<eio::AbcTile*>(ws_get(c, "abc_tiles", sizeof(eio::AbcTile) * tiles.size()));
        a.d_par = static_cast<eio::AbcShot*>(ws_get(c, "abc_par", sizeof(eio::AbcShot) * ns));
        EIO_CHECK_HIP(hipMemcpyAsync(a.d_tiles, tiles.data(), sizeof(eio::AbcTile) * tiles.size(), hipMemcpyHostToDevice,
                                     c->stream));
        EIO_CHECK_HIP(hipMemcpyAsync(a.d_par, par.data(), sizeof(eio::AbcShot) * ns, hipMemcpyHostToDevice, c->stream));
        return a;
    }
}

// frames of the new level of fine step k of the active shots
template <typename T>
void launch_abc(Run<T>& run, const AbcRun& a, int k) {
    if constexpr (sizeof(T) == 4) {
        const int nt = a.active(k);
        if (nt <= 0) return;
        klaunch(run.c, eio::k_abc_cnn<float>, dim3(nt), dim3(eio::ABC_NTHR), run.d_shots, a.d_par, a.d_tiles, nt, k);
    }
}

// ------------------------------------------------------------------ persistent launch
// Neighbour table of a FWD2 plan: per tile the (up to 8) adjacent tiles of the same shot (the
// halo of a two-step tile is 8 cells, no wider than a tile), -1 where there is none.
std::vector<int> plan_neighbours(Ctx* c, const Plan& p, const eio_shot* shots, int TX) {
    std::vector<int> nb(p.tiles.size() * 8, -1);
    int base = 0;
    for (size_t oi = 0; oi < p.order.size(); ++oi) {
        const ModelBase& m = get_model(c, shots[p.order[oi]].model);
        const int nbz = (m.g.nz - 2 * eio::FAT + eio::TZ2 - 1) / eio::TZ2;
        const int nbx = (m.g.nx - 2 * eio::FAT + TX - 1) / TX;
        for (int bx = 0; bx < nbx; ++bx)
            for (int bz = 0; bz < nbz; ++bz) {
                int* o = nb.data() + size_t(base + bx * nbz + bz) * 8;
                int q = 0;
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dz = -1; dz <= 1; ++dz) {
                        if (!dx && !dz) continue;
                        const int x = bx + dx, z = bz + dz;
                        if (x >= 0 && x < nbx && z >= 0 && z < nbz) o[q] = base + x * nbz + z;
                        ++q;
                    }
            }
        base = p.tile_end[oi];
    }
    return nb;
}

// Co-resident workgroups of a kernel on the context's device.
int coresident_blocks(Ctx* c, const void* fn, int threads) {
    int per_cu = 0, ncu = 0;
    EIO_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, 0));
    EIO_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    return per_cu * ncu;
}

// The persistent kernel of two-step variant v (same tile width; the variants with nontemporal
// stores map to their cached-store twins: results are re-read by the next pair).
template <typename F>
void with_persist_kernel(int v, F&& f) {
    switch (v) {
        case 0: case 1: case 5: case 6: f(eio::k_step2_fwd_persist<float, 2, 8, 2>, 8); break;  // 2,8,4 spills
        case 2: f(eio::k_step2_fwd_persist<float, 4, 4, 2>, 4); break;
        case 3: f(eio::k_step2_fwd_persist<float, 1, 8, 4>, 8); break;
        case 12: f(eio::k_step2_fwd_persist<float, 1, 8, 4, true>, 8); break;  // 8 waves/SIMD spills
        case 8: f(eio::k_step2_fwd_persist<float, 4, 8, 4, true>, 8); break;
        case 9: f(eio::k_step2_fwd_persist<float, 8, 4, 2, true>, 4); break;
        case 10: f(eio::k_step2_fwd_persist<float, 2, 8, 6, true>, 8); break;
        case 11: f(eio::k_step2_fwd_persist<float, 3, 8, 2, false, true>, 8); break;
        case 4: case 7: f(eio::k_step2_fwd_persist<float, 4, 8, 1>, 8); break;
        default: fail("no persistent kernel for two-step variant " + std::to_string(v));
    }
}

// Whether the two-step NL forward of this run takes the persistent launch.
bool use_persistent(Ctx* c, const Run<float>& run) {
    if (c->persistent == 0) return false;
    if (c->persistent == 1) return true;
    bool fits = false;
    with_persist_kernel(run.v2, [&](auto kern, int wx) {
        fits = long(run.plan.tiles.size()) <= coresident_blocks(c, reinterpret_cast<const void*>(kern), 64 * wx);
    });
    return fits;
}

// The whole time loop of a two-step NL forward run in one cooperative launch; returns after
// checking the abort flag.
void launch_persistent(Run<float>& run, const eio_shot* shots, int kmax) {
    Ctx* c = run.c;
    const Plan& p = run.plan;
    const int ntile = int(p.tiles.size());
    const int TX = F32_VARIANTS2[run.v2].R2 * F32_VARIANTS2[run.v2].WX;
    const std::vector<int> nb = plan_neighbours(c, p, shots, TX);
    int* d_nb = static_cast<int*>(ws_get(c, "persist_nbr", nb.size() * sizeof(int)));
    int* d_done = static_cast<int*>(ws_get(c, "persist_done", (size_t(ntile) + 64) * sizeof(int)));
    int* d_abort = d_done + ntile;
    EIO_CHECK_HIP(hipMemcpyAsync(d_nb, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    fill0<int>(c, d_done, long(ntile) + 1);
    const int npair = (kmax + 1) / 2;
    with_persist_kernel(run.v2, [&](auto kern, int wx) {
        const int cap = coresident_blocks(c, reinterpret_cast<const void*>(kern), 64 * wx);
        if (cap < 1) fail("persistent launch: the kernel does not fit on a CU");
        int G = std::min(ntile, cap);
        if (G < ntile && G >= 8) G -= G % 8;  // several tiles per workgroup: keep them on one XCD
        const eio::ShotDev<float>* sh = run.d_shots;
        const eio::TileDev* tt = run.d_tiles;
        const eio::Entry* te = run.d_ents;
        int nt = ntile, np = npair, sl = c->spin_limit;
        void* args[] = {&sh, &tt, &te, &nt, &d_nb, &d_done, &d_abort, &np, &sl};
        EIO_CHECK_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(G), dim3(64 * wx), args, 0,
                                                 c->stream));
    });
    int aborted = 0;
    EIO_CHECK_HIP(hipMemcpyAsync(&aborted, d_abort, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (aborted)
        throw EioError{EIO_ERR_RUNTIME, "persistent launch: a tile waited longer than the spin limit for a "
                                        "neighbour; the launch was abandoned (outputs are incomplete)"};
}

// Fine-step snapshots of one shot: after step steps[i] (after the boundary correction) the
// older then the newer level, [nsnap][2][5][nx*nz].
template <typename T>
struct Snap {
    int oi = -1, nsnap = 0, next = 0;
    const int* steps = nullptr;
    T* out = nullptr;
};

template <typename T>
void nl_forward(Ctx* c, int nshot, const eio_shot* shots, int flags, int wshot = -1, void* wfld = nullptr,
                int snap_shot = -1, int nsnap = 0, const int* snap_steps = nullptr, void* snap_out = nullptr) {
    // the saved wavefield, snapshots, the free surface and the learned boundary use the
    // one-step kernel
    bool any_fs = false;
    for (int i = 0; i < nshot && shots; ++i) any_fs = any_fs || get_model(c, shots[i].model).g.surface;
    const bool abc = any_abc(c, nshot, shots);
    const bool plain = sizeof(T) == 4 && !wfld && !any_fs && !abc && nsnap == 0;
    const bool three = plain && c->time_block == 3;
    const bool two = plain && c->time_block == 2;
    // plain runs: shot groups, each through all its steps before the next; without graph
    // recording the groups run on two streams, each group with half the budget (the cache holds
    // one group of each stream), so one stream's launch tails and gaps are filled by the other's
    const long long per_cell = (two || three ? 25 : 15) * (long long)sizeof(T);
    const long long budget = c->group_bytes == EIO_GROUP_AUTO ? GROUP_AUTO_BYTES : c->group_bytes;
    const int nstream = plain && budget > 0 && c->group_streams > 1 && !c->use_graph ? 2 : 1;
    const long long gbudget = budget > 0 ? std::max(1LL, budget / nstream) : 0;
    const bool grouped = plain && nl_grouped(c, nshot, shots, per_cell, gbudget);
    Run<T> run = prepare<T>(c, nshot, shots, Op::NL_FWD, flags, nullptr, nullptr, 0, two, false, false, three,
                            grouped);
    if (wfld) attach_wavefield(run, wshot, wfld, flags);
    AbcRun ar;
    if (abc) ar = abc_prepare(run, shots);
    int kmax = run.plan.nsteps.empty() ? 0 : run.plan.nsteps[0];
    if (c->step_limit > 0) kmax = std::min(kmax, c->step_limit);
    Snap<T> sn;
    size_t snap_n = 0;
    if (nsnap > 0) {
        if (!snap_steps || !snap_out) fail("snapshot steps or output pointer is NULL");
        if (!(flags & EIO_DEVICE_PTRS)) fail("snapshots need EIO_DEVICE_PTRS (device output)");
        if (snap_shot < 0 || snap_shot >= nshot) fail("snapshot shot index out of range");
        int oi = 0;
        while (run.plan.order[oi] != snap_shot) ++oi;
        sn.oi = oi;
        sn.nsnap = nsnap;
        sn.steps = snap_steps;
        sn.out = static_cast<T*>(snap_out);
        snap_n = size_t(run.sd[oi].nz) * run.sd[oi].nx;
        for (int i = 0; i < nsnap; ++i) {
            if (snap_steps[i] < 0 || snap_steps[i] >= run.plan.nsteps[oi] || snap_steps[i] >= kmax)
                fail("snapshot step out of range");
            if (i > 0 && snap_steps[i] <= snap_steps[i - 1]) fail("snapshot steps must increase");
        }
    }
    long long nlaunch = 0;
    if constexpr (sizeof(T) == 4) {
        if (two && use_persistent(c, run)) {
            timing_begin<T>(c);
            EIO_CHECK_HIP(hipEventRecord(c->ev[2], c->stream));
            launch_persistent(run, shots, kmax);
            EIO_CHECK_HIP(hipEventRecord(c->ev[3], c->stream));
            timing_end<T>(c, 1);
            finish(run);
            return;
        }
    }
    const std::vector<ShotGroup> groups =
        plain ? shot_groups(run, gbudget, per_cell) : std::vector<ShotGroup>{{0, nshot}};
    c->last_groups = int(groups.size());
    // stream of each group: the one with less work so far (cell-updates), in plan order
    std::vector<int> gstream(groups.size(), 0);
    if (nstream > 1 && groups.size() > 1) {
        long long work[2] = {0, 0};
        for (size_t gi = 0; gi < groups.size(); ++gi) {
            long long w = 0;
            for (int i = groups[gi].s0; i < groups[gi].s1; ++i)
                w += (long long)run.sd[i].nz * run.sd[i].nx * run.plan.nsteps[i];
            gstream[gi] = work[1] < work[0] ? 1 : 0;
            work[gstream[gi]] += w;
        }
    }
    const bool side = nstream > 1 && groups.size() > 1;
    timing_begin<T>(c);
    EIO_CHECK_HIP(hipEventRecord(c->ev[2], c->stream));
    run_loop(c, [&]() {
        if constexpr (sizeof(T) == 4) {
            if (plain) {
                // the side stream starts after what the call queued so far; the call's stream
                // waits for it at the end (also on errors, so no launch outlives the call)
                struct SideStream {
                    Ctx* c;
                    hipStream_t main;
                    bool on;
                    ~SideStream() {
                        c->stream = main;
                        if (on && hipEventRecord(c->join, c->side) == hipSuccess)
                            (void)hipStreamWaitEvent(main, c->join, 0);
                    }
                } ss{c, c->stream, side};
                if (side) {
                    EIO_CHECK_HIP(hipEventRecord(c->fork, ss.main));
                    EIO_CHECK_HIP(hipStreamWaitEvent(c->side, c->fork, 0));
                }
                for (size_t gi = 0; gi < groups.size(); ++gi) {
                    const ShotGroup& g = groups[gi];
                    c->stream = gstream[gi] ? c->side : ss.main;
                    const eio::TileDev* tg = run.d_tiles + (g.s0 ? run.plan.tile_end[g.s0 - 1] : 0);
                    int kg = run.plan.nsteps[g.s0];
                    if (c->step_limit > 0) kg = std::min(kg, c->step_limit);
                    if (two) {
                        // pairs (k, k+1); a shot whose last pair has one step does only step k
                        for (int k = 0; k < kg; k += 2, ++nlaunch)
                            launch_fwd2(run, group_tiles(run.plan, g, k), k, (k >> 1) & 1, 0, -1, tg);
                    } else if (three) {
                        // triples (k, k+1, k+2); a shot's last launch may take one or two steps
                        for (int k = 0; k < kg; k += 3, ++nlaunch)
                            launch_fwd3(run, group_tiles(run.plan, g, k), k, (k / 3) & 1, tg);
                    } else {
                        for (int k = 0; k < kg; ++k, ++nlaunch)
                            launch_fwd<T, eio::M_NL>(run, group_tiles(run.plan, g, k), k, 0, 0, -1, nullptr, tg);
                    }
                }
                return;
            }
        }
        for (int k = 0; k < kmax; ++k, ++nlaunch) {
            launch_fwd<T, eio::M_NL>(run, active_tiles(run.plan, k), k, 0);
            if (abc) launch_abc(run, ar, k);
            if (sn.next < sn.nsnap && sn.steps[sn.next] == k) {
                const eio::ShotDev<T>& d = run.sd[sn.oi];
                const T* newer = (k & 1) ? d.B : d.A;
                const T* older = (k & 1) ? d.A : d.B;
                T* o = sn.out + size_t(sn.next) * 10 * snap_n;
                EIO_CHECK_HIP(hipMemcpyAsync(o, older, 5 * snap_n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
                EIO_CHECK_HIP(hipMemcpyAsync(o + 5 * snap_n, newer, 5 * snap_n * sizeof(T), hipMemcpyDeviceToDevice,
                                             c->stream));
                ++sn.next;
            }
        }
    }, nsnap == 0);  // snapshot copies run directly
    EIO_CHECK_HIP(hipEventRecord(c->ev[3], c->stream));
    timing_end<T>(c, nlaunch);
    finish(run);
}

// eio_nl_forward whose workspace (state levels, shot/tile/entry tables, the persistent
// launch's neighbour table and flags) is carved from this device's scratch array instead of
// the context's hipMalloc'd workspace; the loop runs without a graph (instantiating one
// allocates).  The state is cleared by a fill kernel each call; calls that overlap on one
// device share the array and race.
template <typename T>
void nl_forward_noalloc(Ctx* c, int nshot, const eio_shot* shots, int flags) {
    if (!(flags & EIO_DEVICE_PTRS)) fail("eio_nl_forward_noalloc needs EIO_DEVICE_PTRS (device traces)");
    if (!c->scratch_dev) fail("eio_nl_forward_noalloc: no scratch array on this device");
    struct Restore {
        Ctx* c;
        int graph;
        ~Restore() {
            c->bump = nullptr;
            c->bump_cap = c->bump_off = 0;
            c->use_graph = graph;
        }
    } restore{c, c->use_graph};
    c->use_graph = 0;
    c->bump = c->scratch_dev;
    c->bump_cap = EIO_SCRATCH_BYTES;
    c->bump_off = 0;
    nl_forward<T>(c, nshot, shots, flags);
}

// One fine step (no sources or receivers) of `nstate` states of one model, in place:
// states [nstate][2][5][nx*nz] hold (older, newer); afterwards (newest, newer).  The
// model's learned boundary, when attached, corrects the new level as in a propagation.
template <typename T>
void step_states(Ctx* c, int model, int nstate, void* states, int flags) {
    if (nstate <= 0) fail("nstate must be > 0");
    if (!states) fail("states pointer is NULL");
    if (!(flags & EIO_DEVICE_PTRS)) fail("eio_nl_step_states needs EIO_DEVICE_PTRS (device states)");
    const ModelBase& m = get_model(c, model);
    std::vector<eio_shot> sh(nstate);
    for (auto& s : sh) {
        std::memset(&s, 0, sizeof(s));
        s.model = model;
    }
    const int limit = c->step_limit;
    c->step_limit = 0;
    Run<T> run = prepare<T>(c, nstate, sh.data(), Op::NL_FWD, flags | EIO_DEVICE_PTRS, nullptr);
    c->step_limit = limit;
    T* st = static_cast<T*>(states);
    for (int oi = 0; oi < nstate; ++oi) {
        T* base = st + size_t(run.plan.order[oi]) * 10 * size_t(m.n);
        run.sd[oi].A = base;
        run.sd[oi].B = base + 5 * size_t(m.n);
    }
    EIO_CHECK_HIP(hipMemcpyAsync(run.d_shots, run.sd.data(), sizeof(eio::ShotDev<T>) * nstate, hipMemcpyHostToDevice,
                                 c->stream));
    AbcRun ar;
    if (m.abc_on) ar = abc_prepare(run, sh.data());
    launch_fwd<T, eio::M_NL>(run, int(run.plan.tiles.size()), 0, 0);
    if (m.abc_on) launch_abc(run, ar, 0);
    EIO_CHECK_HIP(hipGetLastError());
    finish(run);
}

// Attach (on = 1: inputs new level and new - previous; on = 2: new and previous level; plus
// 10 x depth to write only the `depth` frame rows next to the update region, default all 4)
// or detach (on = 0) a learned boundary to a model.  weights: FP32, per edge
// (x-min, x-max, z-min, z-max) the three layers in torch Conv2d layout [out][in][5][5] with
// (out, in) = (16, 10), (16, 16), (5, 16), the kernel's first axis along x and second along z;
// scale: 5 per-field scales of the new level (and output), then 5 of the differences.
void set_abc(Ctx* c, int model, int on, const float* w, const float* scale) {
    ModelBase& m = get_model(c, model);
    if (c->prec != 4) fail("the learned boundary runs in FP32 contexts only");
    const int depth = on / 10 ? on / 10 : eio::FAT;
    on %= 10;
    if (on < 0 || on > 2 || depth < 1 || depth > eio::FAT) fail("on must be 0, 1 or 2, plus 10 x depth (1..4)");
    if (on) {
        if (m.g.surface) fail("the learned boundary is not available with the free surface");
        if (!w || !scale) fail("weights or scale is NULL");
        for (int f = 0; f < 10; ++f)
            if (!(scale[f] > 0.f) || !std::isfinite(scale[f])) fail("scales must be positive and finite");
    }
    if (m.abc_w) {
        EIO_CHECK_HIP(hipFree(m.abc_w));
        m.abc_w = nullptr;
    }
    m.abc_on = 0;
    if (!on) return;
    const int cin[3] = {10, 16, 16}, cout[3] = {16, 16, 5};
    std::vector<unsigned short> frag(4 * size_t(eio::ABC_WEDGE), 0);
    auto bf = [](float f) {
        unsigned u;
        std::memcpy(&u, &f, 4);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return (unsigned short)(u >> 16);
    };
    size_t wo = 0;
    for (int e = 0; e < 4; ++e)
        for (int l = 0; l < 3; ++l) {
            for (int o = 0; o < cout[l]; ++o)
                for (int i = 0; i < cin[l]; ++i)
                    for (int dx = 0; dx < 5; ++dx)
                        for (int dz = 0; dz < 5; ++dz) {
                            const float v = w[wo + ((size_t(o) * cin[l] + i) * 5 + dx) * 5 + dz];
                            const int tap = dx * 5 + dz;
                            // B fragment of tap pair tap/2 (32 x 16): K row 16 (tap % 2) + i sits in
                            // lane 16 g + o with g = 2 (tap % 2) + i / 8, element i % 8
                            const int lane = 16 * (2 * (tap % 2) + i / 8) + o;
                            frag[size_t(e) * eio::ABC_WEDGE + size_t(l) * eio::ABC_PAIRS * eio::ABC_FRAG +
                                 size_t(tap / 2) * eio::ABC_FRAG + lane * 8 + (i % 8)] = bf(v);
                        }
            wo += size_t(cout[l]) * cin[l] * 25;
        }
    DeviceGuard dg(c->device);
    EIO_CHECK_HIP(hipMalloc(&m.abc_w, frag.size() * sizeof(unsigned short)));
    EIO_CHECK_HIP(hipMemcpy(m.abc_w, frag.data(), frag.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    for (int f = 0; f < 10; ++f) m.abc_scale[f] = scale[f];
    m.abc_on = on;
    m.abc_depth = depth;
}

template <typename T>
void nl_adjoint(Ctx* c, int nshot, const eio_shot* shots, int flags, int wshot = -1, void* wfld = nullptr) {
    if (nshot <= 0 || !shots) fail("no shots");
    // FP32 without free surface or saved wavefield: two reverse fine steps per launch
    bool any_fs = false;
    for (int i = 0; i < nshot; ++i) any_fs = any_fs || get_model(c, shots[i].model).g.surface;
    int kmax0 = get_model(c, shots[0].model).ntw - 1;
    if (c->step_limit > 0) kmax0 = std::min(kmax0, c->step_limit);
    const bool two = sizeof(T) == 4 && c->adj_time_block == 2 && !wfld && !any_fs && kmax0 % 2 == 0;
    Run<T> run = prepare<T>(c, nshot, shots, Op::NL_ADJ, flags, nullptr, nullptr, 0, false, two);
    if (wfld) attach_wavefield(run, wshot, wfld, flags);
    // The adjoint runs backwards from each shot's own last sample: shots of different
    // lengths are aligned at their end, so a shot joins the batch when its remaining step
    // count is reached.  Keep it simple and exact: all shots of a call share ntw here.
    for (size_t i = 1; i < run.plan.nsteps.size(); ++i)
        if (run.plan.nsteps[i] != run.plan.nsteps[0]) fail("adjoint shots of one call must share nts and sub");
    int kmax = run.plan.nsteps.empty() ? 0 : run.plan.nsteps[0];
    if (c->step_limit > 0) kmax = std::min(kmax, c->step_limit);
    const int ntile = int(run.plan.tiles.size());
    timing_begin<T>(c);
    EIO_CHECK_HIP(hipEventRecord(c->ev[2], c->stream));
    run_loop(c, [&]() {
        if constexpr (sizeof(T) == 4) {
            if (two) {
                for (int k = 0; k < kmax; k += 2) launch_adj2<eio::M_NLADJ>(run, k, 0);
                return;
            }
        }
        for (int k = 0; k < kmax; ++k) launch_adj<T, eio::M_NLADJ>(run, ntile, k, 0);
    });
    EIO_CHECK_HIP(hipEventRecord(c->ev[3], c->stream));
    timing_end<T>(c, kmax);
    finish(run);
}

// The Born scattered pass over coarse interval its (fine steps its*sub ..): two fine steps
// per launch on the run's third plan (`two`), secondary sources from slots ss[p], ss[p^1].
template <typename T>
void scattered_interval(Run<T>& run, int ntile, bool two, int its, int sub, int p) {
    if constexpr (sizeof(T) == 4) {
        if (two) {
            const int nt3 = int(run.plan3.tiles.size());
            for (int j = 0; j < sub; j += 2) {
                const int k = its * sub + j;
                launch_fwd2<false, false, true>(run, nt3, k, (k >> 1) & 1, 0, -1, run.d_tiles3, run.d_ents3, p);
            }
            return;
        }
    }
    for (int it2 = 1; it2 <= sub; ++it2) launch_fwd<T, eio::M_BSCAT>(run, ntile, its * sub + it2 - 1, p);
}

// The Born source pass (fine sources, Vx/Vz decimated into the whole record): two fine
// steps per launch on the run's third plan when it has one (FP32, even step count).
template <typename T>
void born_source_pass(Run<T>& run, int ntile, int kmax) {
    if constexpr (sizeof(T) == 4) {
        if (!run.plan3.tiles.empty()) {
            const int nt3 = int(run.plan3.tiles.size());
            for (int k = 0; k < kmax; k += 2)
                launch_fwd2<true, true>(run, nt3, k, (k >> 1) & 1, 0, -1, run.d_tiles3, run.d_ents3);
            return;
        }
    }
    for (int k = 0; k < kmax; ++k) launch_fwd<T, eio::M_BSRC>(run, ntile, k, 0, 0, -1);
}

template <typename T>
void born_forward(Ctx* c, int nshot, const eio_shot* shots, const void* const dm[5], int flags) {
    if (!dm) fail("dm is NULL");
    const ModelBase& m = get_model(c, shots[0].model);
    // the kernels address dm as one contiguous 5-field buffer
    const size_t nb = size_t(m.n) * sizeof(T);
    T* b = static_cast<T*>(ws_get(c, "dm", 5 * nb));
    const hipMemcpyKind kind = (flags & EIO_DEVICE_PTRS) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    for (int f = 0; f < 5; ++f) {
        if (!dm[f]) fail("dm pointer is NULL");
        EIO_CHECK_HIP(hipMemcpyAsync(b + size_t(f) * m.n, dm[f], nb, kind, c->stream));
    }
    const void* dmp[5] = {b, b + m.n, b + 2 * m.n, b + 3 * m.n, b + 4 * m.n};
    const int kmax = m.ntw - 1, sub = m.g.sub_fd, nts = m.g.nts;
    // pairs must stay inside a coarse interval (decimation, secondary-source slots): even sub_fd
    const bool src2 = sizeof(T) == 4 && c->time_block >= 2 && sub % 2 == 0;
    Run<T> run = prepare<T>(c, nshot, shots, Op::BORN_FWD, flags, dmp, nullptr, 0, false, false, src2);
    const int ntile = int(run.plan.tiles.size());
    long long nstep = 0;
    timing_begin<T>(c);
    EIO_CHECK_HIP(hipEventRecord(c->ev[2], c->stream));
    run_loop(c, [&]() {
        born_source_pass(run, ntile, kmax);
        // scattered pass restarts from a zero state (BornElasticGpuFunctions.cu:773-774)
        for (int oi = 0; oi < nshot; ++oi) {
            fill0<T>(c, run.sd[oi].A, 5 * m.n);
            fill0<T>(c, run.sd[oi].B, 5 * m.n);
            if (src2) {
                fill0<T>(c, run.sd[oi].C, 5 * m.n);
                fill0<T>(c, run.sd[oi].D, 5 * m.n);
            }
        }
        int p = 0;
        launch_img<T, false>(run, ntile, 0, p);
        for (int its = 0; its < nts - 1; ++its) {
            launch_img<T, false>(run, ntile, its + 1, p ^ 1);
            scattered_interval(run, ntile, src2, its, sub, p);
            p ^= 1;
        }
    });
    nstep = 2LL * kmax;
    EIO_CHECK_HIP(hipEventRecord(c->ev[3], c->stream));
    timing_end<T>(c, nstep);
    finish(run);
}

template <typename T>
void stack_into(Run<T>& run, const ModelBase& m, int nshot, void* const dm[5], int flags);

template <typename T>
void born_adjoint(Ctx* c, int nshot, const eio_shot* shots, void* const dm[5], int flags) {
    if (!dm) fail("dm is NULL");
    const ModelBase& m = get_model(c, shots[0].model);
    const int kmax = m.ntw - 1, sub = m.g.sub_fd, nts = m.g.nts;
    const bool two = sizeof(T) == 4 && c->adj_time_block == 2 && sub % 2 == 0;
    // pairs must stay inside a coarse interval (decimation, secondary-source slots): even sub_fd
    const bool src2 = sizeof(T) == 4 && c->time_block >= 2 && sub % 2 == 0;
    Run<T> run = prepare<T>(c, nshot, shots, Op::BORN_ADJ, flags, nullptr, nullptr, 0, false, two, src2);
    const int ntile = int(run.plan.tiles.size());
    timing_begin<T>(c);
    EIO_CHECK_HIP(hipEventRecord(c->ev[2], c->stream));
    run_loop(c, [&]() {
        born_source_pass(run, ntile, kmax);
        for (int oi = 0; oi < nshot; ++oi) {
            fill0<T>(c, run.sd[oi].A, 5 * m.n);
            fill0<T>(c, run.sd[oi].B, 5 * m.n);
            if (two) {
                fill0<T>(c, run.sd[oi].C, 5 * m.n);
                fill0<T>(c, run.sd[oi].D, 5 * m.n);
            }
        }
        int p = 0;
        for (int its = nts - 2; its >= 0; --its) {
            adjoint_interval(run, ntile, two, its, nts, sub, p);
            launch_img<T, true>(run, ntile, its + 1, p ^ 1);
            p ^= 1;
        }
        launch_img<T, true>(run, ntile, 0, p ^ 1);
    });
    EIO_CHECK_HIP(hipEventRecord(c->ev[3], c->stream));
    timing_end<T>(c, 2LL * kmax);
    stack_into(run, m, nshot, dm, flags);
    finish(run);
}

// dm[f] += sum over shots of the per-shot imaging accumulators (BornElasticShotsGpu.cpp:313-323)
template <typename T>
void stack_into(Run<T>& run, const ModelBase& m, int nshot, void* const dm[5], int flags) {
    Ctx* c = run.c;
    const size_t nb = size_t(m.n) * sizeof(T);
    T* out = static_cast<T*>(ws_get(c, "dmout", 5 * nb));
    std::vector<const T*> accp(size_t(nshot) * 5);
    for (int oi = 0; oi < nshot; ++oi)
        for (int f = 0; f < 5; ++f) accp[size_t(f) * nshot + oi] = run.sd[oi].dm + size_t(f) * m.n;
    const T** d_accp = static_cast<const T**>(ws_get(c, "accp", accp.size() * sizeof(T*)));
    EIO_CHECK_HIP(hipMemcpyAsync(d_accp, accp.data(), accp.size() * sizeof(T*), hipMemcpyHostToDevice, c->stream));
    for (int f = 0; f < 5; ++f) {
        T* o = (flags & EIO_DEVICE_PTRS) ? static_cast<T*>(dm[f]) : out + size_t(f) * m.n;
        if (!(flags & EIO_DEVICE_PTRS))
            EIO_CHECK_HIP(hipMemcpyAsync(o, dm[f], nb, hipMemcpyHostToDevice, c->stream));
        const long nbk = (m.n + eio::NTHR - 1) / eio::NTHR;
        hipLaunchKernelGGL(eio::k_stack<T>, dim3(unsigned(nbk)), dim3(eio::NTHR), 0, c->stream, o,
                           d_accp + size_t(f) * nshot, nshot, m.n);
        EIO_CHECK_HIP(hipGetLastError());
        if (!(flags & EIO_DEVICE_PTRS))
            EIO_CHECK_HIP(hipMemcpyAsync(dm[f], o, nb, hipMemcpyDeviceToHost, c->stream));
    }
}

// ------------------------------------------------------------------ checkpointed FWI gradient
// The reference's FWI gradient is the Born adjoint of the data residual at the current model
// (efwiMain.py:229-476 -> BornElasticShotsGpu.cpp:228-378), whose source pass stores Vx and
// Vz at every coarse sample (nz*nx*nts per component and shot).  Here the NL forward that
// produces the predicted data also saves the batch state every L coarse intervals; the
// gradient call sweeps backwards and, just before the imaging condition first needs a
// segment, recomputes that segment from its checkpoint into a ring of L+3 coarse slices.
// The source injection of the NL forward (coarse traces, time-interpolated in the step) is
// the same arithmetic as the Born source pass on pre-fine-sampled traces.

unsigned long long geometry_hash(int nshot, const eio_shot* shots) {
    unsigned long long h = 1469598103934665603ULL;
    auto mix = [&](long long v) {
        h ^= (unsigned long long)v;
        h *= 1099511628211ULL;
    };
    mix(nshot);
    for (int i = 0; i < nshot; ++i) {
        mix(shots[i].model);
        for (int g = 0; g < 4; ++g) {
            mix(shots[i].n_src[g]);
            mix(shots[i].n_rec[g]);
            for (int u = 0; u < shots[i].n_src[g] && shots[i].src_pos[g]; ++u) mix(shots[i].src_pos[g][u]);
            for (int u = 0; u < shots[i].n_rec[g] && shots[i].rec_pos[g]; ++u) mix(shots[i].rec_pos[g][u]);
        }
    }
    return h;
}

// A store's buffers: the first store alive borrows the context's grow-only workspace (no
// allocation when consecutive stores fit: allocating and freeing tens of GB per gradient costs
// seconds); stores made while it is held allocate their own.
void fwi_free(Ctx* c, FwiStore& f) {
    if (f.ws) {
        c->fwi_ws_busy = false;
    } else {
        if (f.ck) (void)hipFree(f.ck);
        if (f.ring) (void)hipFree(f.ring);
    }
    f.ck = f.ring = nullptr;
    f.ws = false;
}

template <typename T>
void copy_dev(Ctx* c, T* dst, const T* src, size_t n) {
    if (!n) return;
    const unsigned nb = unsigned(std::min<size_t>((n + eio::NTHR - 1) / eio::NTHR, 256 * 64));
    klaunch(c, eio::k_copy<T>, dim3(nb), dim3(eio::NTHR), dst, src, long(n));
}

inline int seg_begin(const FwiStore& f, int j) { return j * f.L; }
inline int seg_end(const FwiStore& f, int j) { return std::min((j + 1) * f.L, f.nts - 1); }

template <typename T>
int fwi_forward(Ctx* c, int nshot, const eio_shot* shots, int flags, int segment) {
    if (nshot <= 0 || !shots) fail("no shots");
    const ModelBase& m = get_model(c, shots[0].model);
    const int nts = m.g.nts, sub = m.g.sub_fd;
    FwiStore f;
    f.prec = int(sizeof(T));
    f.nshot = nshot;
    f.model = shots[0].model;
    f.nts = nts;
    f.sub = sub;
    f.L = segment > 0 ? segment : int(std::ceil(std::sqrt(double(nts - 1))));
    f.L = std::max(1, std::min(f.L, nts - 1));
    f.nseg = (nts - 1 + f.L - 1) / f.L;
    f.wcap = std::min(f.L + 3, nts);
    f.sig = geometry_hash(nshot, shots);
    f.state_elems = size_t(nshot) * 10 * size_t(m.n);
    const size_t ring_elems = size_t(nshot) * 2 * size_t(m.n) * f.wcap;
    const size_t ck_elems = size_t(f.nseg - 1) * f.state_elems;
    f.bytes = (ring_elems + ck_elems) * sizeof(T);
    const std::string oom = "out of device memory for the store (" + std::to_string(f.nseg - 1) +
                            " checkpoints, " + std::to_string(f.wcap) + " coarse slices); pass another segment";
    if (!c->fwi_ws_busy) {
        f.ws = true;
        c->fwi_ws_busy = true;
        try {
            f.ring = ws_get(c, "fwi_ring", ring_elems * sizeof(T));
            if (ck_elems) f.ck = ws_get(c, "fwi_ck", ck_elems * sizeof(T));
        } catch (const EioError&) {
            (void)hipGetLastError();
            fwi_free(c, f);
            throw EioError{EIO_ERR_RUNTIME, oom};
        }
    } else {
        if (hipMalloc(&f.ring, ring_elems * sizeof(T)) != hipSuccess ||
            (ck_elems && hipMalloc(&f.ck, ck_elems * sizeof(T)) != hipSuccess)) {
            (void)hipGetLastError();
            fwi_free(c, f);
            throw EioError{EIO_ERR_RUNTIME, oom};
        }
    }
    try {
        // FP32: the two-step kernel (with the decimation in the last segment); pairs of fine
        // steps stay inside one coarse interval because the fine sub-sampling is even, and
        // segments start at even fine steps
        const bool two = sizeof(T) == 4 && c->time_block >= 2 && sub % 2 == 0;
        Run<T> run = prepare<T>(c, nshot, shots, Op::FWI_FWD, flags, nullptr, static_cast<T*>(f.ring), f.wcap, two);
        const int ntile = int(run.plan.tiles.size());
        T* state = run.sd[0].A;  // the batch state is one contiguous buffer in plan order
        timing_begin<T>(c);
        EIO_CHECK_HIP(hipEventRecord(c->ev[2], c->stream));
        run_loop(c, [&]() {
            if constexpr (sizeof(T) == 4) {
                if (two) {
                    for (int j = 0; j < f.nseg; ++j) {
                        const int k0 = seg_begin(f, j) * sub, k1 = seg_end(f, j) * sub;
                        // checkpoint (older, current level) of each shot: levels A,B before
                        // an even pair, C,D before an odd one
                        if (j > 0)
                            for (int oi = 0; oi < nshot; ++oi) {
                                const size_t n10 = 10 * size_t(run.sd[oi].nz) * run.sd[oi].nx;
                                copy_dev<T>(c, static_cast<T*>(f.ck) + size_t(j - 1) * f.state_elems + oi * n10,
                                            ((k0 >> 1) & 1) ? run.sd[oi].C : run.sd[oi].A, n10);
                            }
                        if (j + 1 < f.nseg) {
                            for (int k = k0; k < k1; k += 2) launch_fwd2<false>(run, ntile, k, (k >> 1) & 1);
                        } else {  // the last segment also fills the ring
                            for (int k = k0; k < k1; k += 2)
                                launch_fwd2<true>(run, ntile, k, (k >> 1) & 1, seg_begin(f, j), -1);
                        }
                    }
                    return;
                }
            }
            for (int j = 0; j < f.nseg; ++j) {
                if (j > 0) copy_dev<T>(c, static_cast<T*>(f.ck) + size_t(j - 1) * f.state_elems, state, f.state_elems);
                const int k0 = seg_begin(f, j) * sub, k1 = seg_end(f, j) * sub;
                if (j + 1 < f.nseg) {
                    for (int k = k0; k < k1; ++k) launch_fwd<T, eio::M_NL>(run, ntile, k, 0);
                } else {  // the last segment also fills the ring
                    for (int k = k0; k < k1; ++k)
                        launch_fwd<T, eio::M_NL | eio::F_DEC>(run, ntile, k, 0, seg_begin(f, j), -1);
                }
            }
        });
        EIO_CHECK_HIP(hipEventRecord(c->ev[3], c->stream));
        timing_end<T>(c, (long long)(nts - 1) * sub);
        finish(run);
    } catch (...) {
        (void)hipStreamSynchronize(c->stream);
        fwi_free(c, f);
        throw;
    }
    f.ring_fresh = true;
    const int id = c->next_fwi++;
    c->fwi[id] = f;
    return id;
}

template <typename T>
void fwi_gradient(Ctx* c, int handle, int nshot, const eio_shot* shots, void* const dm[5], int flags) {
    if (!dm) fail("dm is NULL");
    auto it = c->fwi.find(handle);
    if (it == c->fwi.end()) fail("unknown FWI store handle " + std::to_string(handle));
    FwiStore& f = it->second;
    if (nshot != f.nshot || !shots || geometry_hash(nshot, shots) != f.sig)
        fail("shots differ from those of the FWI forward call of this store");
    const ModelBase& m = get_model(c, f.model);
    if (m.g.nts != f.nts || m.g.sub_fd != f.sub) fail("model time axis changed since the FWI forward call");
    const bool two = sizeof(T) == 4 && c->adj_time_block == 2 && f.sub % 2 == 0;
    // FP32: segments are recomputed two fine steps per launch (four levels per shot), the
    // kernel of the forward, so that a recomputed ring equals the forward's bit for bit
    const bool rtwo = sizeof(T) == 4 && c->time_block >= 2 && f.sub % 2 == 0;
    Run<T> run = prepare<T>(c, nshot, shots, Op::FWI_GRAD, flags, nullptr, static_cast<T*>(f.ring), f.wcap,
                            false, two, rtwo);
    const int ntile = int(run.plan.tiles.size());
    const int nts = f.nts, sub = f.sub;
    // forward-recompute state and its shot table
    const size_t fstate_elems = (rtwo ? 2 : 1) * f.state_elems;
    T* fstate = static_cast<T*>(ws_get(c, "fstate", fstate_elems * sizeof(T)));
    std::vector<eio::ShotDev<T>> sdf = run.sd;
    size_t off = 0;
    for (int oi = 0; oi < nshot; ++oi) {
        const size_t n = size_t(sdf[oi].nz) * sdf[oi].nx;
        sdf[oi].A = fstate + off;
        sdf[oi].B = fstate + off + 5 * n;
        sdf[oi].C = rtwo ? fstate + off + 10 * n : nullptr;
        sdf[oi].D = rtwo ? fstate + off + 15 * n : nullptr;
        off += (rtwo ? 20 : 10) * n;
    }
    auto* d_sdf = static_cast<eio::ShotDev<T>*>(ws_get(c, "shots_f", sizeof(eio::ShotDev<T>) * nshot));
    EIO_CHECK_HIP(hipMemcpyAsync(d_sdf, sdf.data(), sizeof(eio::ShotDev<T>) * nshot, hipMemcpyHostToDevice,
                                 c->stream));
    long long nrecomp = 0;
    for (int j = 0; j + 1 < f.nseg; ++j) nrecomp += (long long)(seg_end(f, j) - seg_begin(f, j)) * sub;
    timing_begin<T>(c);
    EIO_CHECK_HIP(hipEventRecord(c->ev[2], c->stream));
    auto recompute = [&](int j) {
        const int dhi = (j + 1 < f.nseg) ? seg_end(f, j) : -1;
        const int k0 = seg_begin(f, j) * sub, k1 = seg_end(f, j) * sub;
        if constexpr (sizeof(T) == 4) {
            if (rtwo) {
                // the checkpoint (older, current) goes to the level pair the segment's first
                // pair reads; the other pair is overwritten on the update region, and every
                // level's cells outside it stay zero from the clear before the sweep
                if (j == 0) {
                    fill0<T>(c, fstate, long(fstate_elems));
                } else {
                    for (int oi = 0; oi < nshot; ++oi) {
                        const size_t n10 = 10 * size_t(sdf[oi].nz) * sdf[oi].nx;
                        copy_dev<T>(c, ((k0 >> 1) & 1) ? sdf[oi].C : sdf[oi].A,
                                    static_cast<const T*>(f.ck) + size_t(j - 1) * f.state_elems + oi * n10, n10);
                    }
                }
                const int nt3 = int(run.plan3.tiles.size());
                for (int k = k0; k < k1; k += 2)
                    launch_fwd2<true, false, false, true>(run, nt3, k, (k >> 1) & 1, seg_begin(f, j), dhi,
                                                          run.d_tiles3, run.d_ents3, 0, d_sdf);
                return;
            }
        }
        if (j == 0) fill0<T>(c, fstate, long(f.state_elems));
        else copy_dev<T>(c, fstate, static_cast<const T*>(f.ck) + size_t(j - 1) * f.state_elems, f.state_elems);
        for (int k = k0; k < k1; ++k) launch_fwd<T, eio::M_RECOMP>(run, ntile, k, 0, seg_begin(f, j), dhi, d_sdf);
    };
    const bool fresh = f.ring_fresh;
    f.ring_fresh = false;
    if (!fresh) nrecomp += (long long)(seg_end(f, f.nseg - 1) - seg_begin(f, f.nseg - 1)) * sub;
    run_loop(c, [&]() {
        if (rtwo) fill0<T>(c, fstate, long(fstate_elems));
        // a store used before no longer holds the last segment in its ring
        if (!fresh) recompute(f.nseg - 1);
        int p = 0;
        int jn = f.nseg - 1;  // next segment start to cross going backwards
        for (int its = nts - 2; its >= 0; --its) {
            if (jn >= 1 && its == seg_begin(f, jn)) {
                // segment jn-1 is recomputed before the imaging at its+1 first needs slot its
                recompute(jn - 1);
                --jn;
            }
            adjoint_interval(run, ntile, two, its, nts, sub, p);
            launch_img<T, true>(run, ntile, its + 1, p ^ 1);
            p ^= 1;
        }
        launch_img<T, true>(run, ntile, 0, p ^ 1);
    });
    EIO_CHECK_HIP(hipEventRecord(c->ev[3], c->stream));
    timing_end<T>(c, (long long)(nts - 1) * sub + nrecomp);
    stack_into(run, m, nshot, dm, flags);
    finish(run);
}

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return EIO_OK;
    } catch (const EioError& e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        g_err = e.what();
        return EIO_ERR_RUNTIME;
    }
}

Ctx* as_ctx(eio_ctx* p) {
    if (!p) fail("context is NULL");
    return reinterpret_cast<Ctx*>(p);
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* eio_last_error(void) { return g_err.c_str(); }

int eio_ctx_create(int device, int precision, eio_ctx** out) {
    return guarded([&]() {
        if (!out) fail("out is NULL");
        if (precision != 4 && precision != 8) fail("precision must be 4 or 8");
        int ndev = 0;
        EIO_CHECK_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) fail("invalid device id");
        DeviceGuard dg(device);
        auto* c = new Ctx();
        c->device = device;
        c->prec = precision;
        try {
            c->scratch_dev = scratch_address(device);
        } catch (...) {
            delete c;
            throw;
        }
        EIO_CHECK_HIP(hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking));
        c->stream = c->own;
        for (auto& e : c->ev) EIO_CHECK_HIP(hipEventCreate(&e));
        EIO_CHECK_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        EIO_CHECK_HIP(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
        EIO_CHECK_HIP(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
        *out = reinterpret_cast<eio_ctx*>(c);
    });
}

int eio_ctx_destroy(eio_ctx* ctx) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto& kv : c->models) model_free(kv.second);
        for (auto& kv : c->fwi) fwi_free(c, kv.second);
        for (auto& kv : c->graphs) graph_free(kv.second);
        for (auto& kv : c->ws)
            if (kv.second.p) (void)hipFree(kv.second.p);
        for (auto& e : c->ev) (void)hipEventDestroy(e);
        (void)hipStreamSynchronize(c->side);
        (void)hipEventDestroy(c->fork);
        (void)hipEventDestroy(c->join);
        (void)hipStreamDestroy(c->side);
        (void)hipStreamDestroy(c->own);
        delete c;
    });
}

int eio_ctx_set_stream(eio_ctx* ctx, void* s) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        c->stream = s ? static_cast<hipStream_t>(s) : c->own;
    });
}

int eio_ctx_set_graph(eio_ctx* ctx, int use_graph) {
    return guarded([&]() {
        as_ctx(ctx)->use_graph = use_graph ? 1 : 0;
    });
}

int eio_ctx_graph_info(eio_ctx* ctx, long long* builds, long long* launches, long long* node_updates) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        if (builds) *builds = c->graph_builds;
        if (launches) *launches = c->graph_launches;
        if (node_updates) *node_updates = c->graph_node_updates;
    });
}

int eio_ctx_set_variant(eio_ctx* ctx, int variant) {
    return guarded([&]() {
        if (variant < 0 || variant >= N_F32_VARIANTS) fail("variant out of range");
        as_ctx(ctx)->variant = variant;
    });
}

int eio_ctx_variant_info(int variant, int* R, int* WX, int* occ) {
    return guarded([&]() {
        if (variant < 0 || variant >= N_F32_VARIANTS) fail("variant out of range");
        if (R) *R = F32_VARIANTS[variant].R;
        if (WX) *WX = F32_VARIANTS[variant].WX;
        if (occ) *occ = F32_VARIANTS[variant].OCC;
    });
}

int eio_ctx_num_variants(void) { return N_F32_VARIANTS; }

int eio_ctx_set_time_block(eio_ctx* ctx, int steps, int variant) {
    return guarded([&]() {
        if (steps < 1 || steps > 3) fail("time block must be 1, 2 or 3 fine steps");
        const int nv = steps == 3 ? N_F32_VARIANTS3 : N_F32_VARIANTS2;
        if (variant != EIO_VARIANT_AUTO && (variant < 0 || variant >= nv)) fail("time-block variant out of range");
        Ctx* c = as_ctx(ctx);
        c->time_block = steps;
        if (steps == 3) {
            c->variant3 = variant;
            c->variant2 = EIO_VARIANT_AUTO;   // the other forward passes take two steps per launch
        } else {
            c->variant2 = variant;
        }
    });
}

int eio_ctx_set_shot_group_bytes(eio_ctx* ctx, long long bytes) {
    return guarded([&]() {
        if (bytes < 0 && bytes != EIO_GROUP_AUTO) fail("shot-group budget must be >= 0 or EIO_GROUP_AUTO");
        as_ctx(ctx)->group_bytes = bytes;
    });
}

int eio_ctx_set_shot_group_streams(eio_ctx* ctx, int streams) {
    return guarded([&]() {
        if (streams != 1 && streams != 2) fail("shot-group streams must be 1 or 2");
        as_ctx(ctx)->group_streams = streams;
    });
}

int eio_ctx_last_shot_groups(eio_ctx* ctx, int* groups) {
    return guarded([&]() {
        if (!groups) fail("groups is NULL");
        *groups = as_ctx(ctx)->last_groups;
    });
}

int eio_ctx_time_block_info(int variant, int* R2, int* WX, int* occ) {
    return guarded([&]() {
        if (variant < 0 || variant >= N_F32_VARIANTS2) fail("two-step variant out of range");
        if (R2) *R2 = F32_VARIANTS2[variant].R2;
        if (WX) *WX = F32_VARIANTS2[variant].WX;
        if (occ) *occ = F32_VARIANTS2[variant].OCC;
    });
}

int eio_ctx_num_time_block_variants(void) { return N_F32_VARIANTS2; }

int eio_ctx_num_time_block3_variants(void) { return N_F32_VARIANTS3; }

int eio_ctx_time_block3_info(int variant, int* R3, int* WX, int* occ) {
    return guarded([&]() {
        if (variant < 0 || variant >= N_F32_VARIANTS3) fail("three-step variant out of range");
        if (R3) *R3 = F32_VARIANTS3[variant].R3;
        if (WX) *WX = F32_VARIANTS3[variant].WX;
        if (occ) *occ = F32_VARIANTS3[variant].OCC;
    });
}

int eio_ctx_time_block_layout(int variant) {
    return variant >= 0 && variant < N_F32_VARIANTS2 ? F32_VARIANTS2[variant].LAYOUT : -1;
}

int eio_ctx_set_adjoint_time_block(eio_ctx* ctx, int steps, int variant) {
    return guarded([&]() {
        if (steps != 1 && steps != 2) fail("adjoint time block must be 1 or 2 fine steps");
        if (variant != EIO_VARIANT_AUTO && (variant < 0 || variant >= N_F32_VARIANTS2A))
            fail("two-step adjoint variant out of range");
        Ctx* c = as_ctx(ctx);
        c->adj_time_block = steps;
        c->variant2a = variant;
    });
}

int eio_ctx_adjoint_time_block_info(int variant, int* R2, int* WX, int* occ) {
    return guarded([&]() {
        if (variant < 0 || variant >= N_F32_VARIANTS2A) fail("two-step adjoint variant out of range");
        if (R2) *R2 = F32_VARIANTS2A[variant].R2;
        if (WX) *WX = F32_VARIANTS2A[variant].WX;
        if (occ) *occ = F32_VARIANTS2A[variant].OCC;
    });
}

int eio_ctx_num_adjoint_time_block_variants(void) { return N_F32_VARIANTS2A; }

int eio_ctx_set_persistent(eio_ctx* ctx, int mode, int spin_limit) {
    return guarded([&]() {
        if (mode < 0 || mode > 2) fail("persistent mode must be 0, 1 or 2");
        if (spin_limit < 0) fail("spin limit must be >= 0");
        Ctx* c = as_ctx(ctx);
        c->persistent = mode;
        if (spin_limit > 0) c->spin_limit = spin_limit;
    });
}

int eio_ctx_set_step_limit(eio_ctx* ctx, int max_steps) {
    return guarded([&]() {
        if (max_steps < 0) fail("step limit must be >= 0");
        as_ctx(ctx)->step_limit = max_steps;
    });
}

int eio_ctx_last_timing(eio_ctx* ctx, double* total_ms, long long* n_step, double* step_ms) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        if (total_ms) *total_ms = c->last_total_ms;
        if (n_step) *n_step = c->last_nstep;
        if (step_ms) *step_ms = c->last_step_ms;
    });
}

int eio_grid_check(const eio_grid* grid, int precision) {
    return guarded([&]() { check_grid(grid, precision); });
}

int eio_model_add(eio_ctx* ctx, const eio_grid* grid, const void* const media[5], int flags,
                  int* model_id) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        if (!model_id) fail("model_id is NULL");
        if (!media) fail("media is NULL");
        DeviceGuard dg(c->device);
        *model_id = (c->prec == 4) ? model_add<float>(c, grid, media, flags)
                                   : model_add<double>(c, grid, media, flags);
    });
}

int eio_model_set_media(eio_ctx* ctx, int model_id, const void* const media[5], int flags) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        if (!media) fail("media is NULL");
        DeviceGuard dg(c->device);
        ModelBase& m = get_model(c, model_id);
        if (c->prec == 4) upload_model<float>(c, m, media, flags);
        else upload_model<double>(c, m, media, flags);
    });
}

int eio_model_set_damping(eio_ctx* ctx, int model_id, const void* table) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        if (!table) fail("damping table is NULL");
        DeviceGuard dg(c->device);
        ModelBase& m = get_model(c, model_id);
        const int n = m.g.minPad;
        if (n < 1) fail("the model has no sponge (minPad = 0)");
        std::vector<unsigned char> h(size_t(n) * c->prec);
        std::memcpy(h.data(), table, h.size());
        for (int j = 0; j < n; ++j) {
            const double v = c->prec == 4 ? double(reinterpret_cast<const float*>(h.data())[j])
                                           : reinterpret_cast<const double*>(h.data())[j];
            if (!(v > 0.0 && v <= 1.0)) fail("damping factors must lie in (0, 1]");
        }
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
        EIO_CHECK_HIP(hipMemcpy(m.damp, h.data(), h.size(), hipMemcpyHostToDevice));
    });
}

int eio_model_remove(eio_ctx* ctx, int model_id) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        ModelBase& m = get_model(c, model_id);
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
        model_free(m);
        c->models.erase(model_id);
    });
}

int eio_nl_forward(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (c->prec == 4) nl_forward<float>(c, nshot, shots, flags);
        else nl_forward<double>(c, nshot, shots, flags);
    });
}

int eio_nl_forward_noalloc(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (c->prec == 4) nl_forward_noalloc<float>(c, nshot, shots, flags);
        else nl_forward_noalloc<double>(c, nshot, shots, flags);
    });
}

int eio_nl_forward_wavefield(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags, int shot,
                             void* wavefield) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (!wavefield) fail("wavefield is NULL");
        if (c->prec == 4) nl_forward<float>(c, nshot, shots, flags, shot, wavefield);
        else nl_forward<double>(c, nshot, shots, flags, shot, wavefield);
    });
}

int eio_nl_forward_snapshots(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags, int shot, int nsnap,
                             const int* steps, void* out) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (nsnap <= 0) fail("nsnap must be > 0");
        if (c->prec == 4) nl_forward<float>(c, nshot, shots, flags, -1, nullptr, shot, nsnap, steps, out);
        else nl_forward<double>(c, nshot, shots, flags, -1, nullptr, shot, nsnap, steps, out);
    });
}

int eio_nl_step_states(eio_ctx* ctx, int model, int nstate, void* states, int flags) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (c->prec == 4) step_states<float>(c, model, nstate, states, flags);
        else step_states<double>(c, model, nstate, states, flags);
    });
}

int eio_model_set_abc(eio_ctx* ctx, int model, int on, const float* weights, const float* scale) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        set_abc(c, model, on, weights, scale);
    });
}

int eio_nl_adjoint_wavefield(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags, int shot,
                             void* wavefield) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (!wavefield) fail("wavefield is NULL");
        if (c->prec == 4) nl_adjoint<float>(c, nshot, shots, flags, shot, wavefield);
        else nl_adjoint<double>(c, nshot, shots, flags, shot, wavefield);
    });
}

int eio_nl_adjoint(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (c->prec == 4) nl_adjoint<float>(c, nshot, shots, flags);
        else nl_adjoint<double>(c, nshot, shots, flags);
    });
}

int eio_born_forward(eio_ctx* ctx, int nshot, const eio_shot* shots, const void* const dm[5],
                     int flags) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (nshot <= 0 || !shots) fail("no shots");
        if (c->prec == 4) born_forward<float>(c, nshot, shots, dm, flags);
        else born_forward<double>(c, nshot, shots, dm, flags);
    });
}

int eio_born_adjoint(eio_ctx* ctx, int nshot, const eio_shot* shots, void* const dm[5], int flags) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (nshot <= 0 || !shots) fail("no shots");
        if (c->prec == 4) born_adjoint<float>(c, nshot, shots, dm, flags);
        else born_adjoint<double>(c, nshot, shots, dm, flags);
    });
}

int eio_fwi_forward(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags, int segment, int* handle) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        if (!handle) fail("handle is NULL");
        if (segment < 0) fail("segment must be >= 0");
        DeviceGuard dg(c->device);
        *handle = (c->prec == 4) ? fwi_forward<float>(c, nshot, shots, flags, segment)
                                 : fwi_forward<double>(c, nshot, shots, flags, segment);
    });
}

int eio_fwi_gradient(eio_ctx* ctx, int handle, int nshot, const eio_shot* shots, void* const dm[5], int flags) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        if (c->prec == 4) fwi_gradient<float>(c, handle, nshot, shots, dm, flags);
        else fwi_gradient<double>(c, handle, nshot, shots, dm, flags);
    });
}

int eio_fwi_info(eio_ctx* ctx, int handle, int* segment, int* nckpt, unsigned long long* bytes) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        auto it = c->fwi.find(handle);
        if (it == c->fwi.end()) fail("unknown FWI store handle " + std::to_string(handle));
        if (segment) *segment = it->second.L;
        if (nckpt) *nckpt = it->second.nseg - 1;
        if (bytes) *bytes = it->second.bytes;
    });
}

int eio_fwi_release(eio_ctx* ctx, int handle) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        auto it = c->fwi.find(handle);
        if (it == c->fwi.end()) fail("unknown FWI store handle " + std::to_string(handle));
        DeviceGuard dg(c->device);
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
        fwi_free(c, it->second);
        c->fwi.erase(it);
    });
}

int eio_ctx_release_workspace(eio_ctx* ctx) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
        for (auto it = c->ws.begin(); it != c->ws.end();) {
            const bool held = c->fwi_ws_busy && (it->first == "fwi_ring" || it->first == "fwi_ck");
            if (!held && it->second.p) {
                EIO_CHECK_HIP(hipFree(it->second.p));
                it = c->ws.erase(it);
            } else {
                ++it;
            }
        }
    });
}

int eio_device_alloc(eio_ctx* ctx, unsigned long long bytes, void** ptr) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        if (!ptr) fail("ptr is NULL");
        DeviceGuard dg(c->device);
        EIO_CHECK_HIP(hipMalloc(ptr, bytes));
    });
}

int eio_device_free(eio_ctx* ctx, void* ptr) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
        EIO_CHECK_HIP(hipFree(ptr));
    });
}

int eio_host_alloc(unsigned long long bytes, void** ptr) {
    return guarded([&]() {
        if (!ptr) fail("ptr is NULL");
        *ptr = nullptr;
        if (bytes == 0) return;
        EIO_CHECK_HIP(hipHostMalloc(ptr, bytes, hipHostMallocPortable));
    });
}

int eio_host_free(void* ptr) {
    return guarded([&]() {
        if (ptr) EIO_CHECK_HIP(hipHostFree(ptr));
    });
}

int eio_memcpy(eio_ctx* ctx, void* dst, const void* src, unsigned long long bytes, int kind) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        hipMemcpyKind k;
        if (kind == 1) k = hipMemcpyHostToDevice;
        else if (kind == 2) k = hipMemcpyDeviceToHost;
        else if (kind == 3) k = hipMemcpyDeviceToDevice;
        else fail("invalid copy kind");
        EIO_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, k, c->stream));
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
    });
}

int eio_copy_probe(eio_ctx* ctx, void* dst, const void* src, unsigned long long bytes, int reps, int kind,
                   double* ms) {
    return guarded([&]() {
        if (kind < 0 || kind > 2) fail("copy probe kind must be 0, 1 or 2");
        Ctx* c = as_ctx(ctx);
        if (!dst || !src || bytes % 16 || reps < 1) fail("invalid copy probe arguments");
        DeviceGuard dg(c->device);
        const long n16 = long(bytes / 16);
        const long nb = (n16 + eio::NTHR - 1) / eio::NTHR;
        const long nb4 = (4 * n16 + eio::NTHR - 1) / eio::NTHR;
        if (nb4 >= (1L << 31)) fail("copy probe too large");
        EIO_CHECK_HIP(hipEventRecord(c->ev[0], c->stream));
        for (int r = 0; r < reps; ++r) {
            if (kind == 2)
                hipLaunchKernelGGL(eio::k_copy4, dim3(unsigned(nb4)), dim3(eio::NTHR), 0, c->stream,
                                   static_cast<float*>(dst), static_cast<const float*>(src), 4 * n16);
            else if (kind == 1)
                hipLaunchKernelGGL(eio::k_copy16_nt, dim3(unsigned(nb)), dim3(eio::NTHR), 0, c->stream,
                                   static_cast<float4*>(dst), static_cast<const float4*>(src), n16);
            else
                hipLaunchKernelGGL(eio::k_copy16, dim3(unsigned(std::min<long>(nb, 256L * 64))), dim3(eio::NTHR), 0,
                                   c->stream, static_cast<float4*>(dst), static_cast<const float4*>(src), n16);
        }
        EIO_CHECK_HIP(hipGetLastError());
        EIO_CHECK_HIP(hipEventRecord(c->ev[1], c->stream));
        EIO_CHECK_HIP(hipEventSynchronize(c->ev[1]));
        float t = 0.f;
        EIO_CHECK_HIP(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
        if (ms) *ms = double(t) / reps;
    });
}

// ---- training kernels of the learned-boundary network (device pointers, synchronous)
namespace {
void abct_check(Ctx* c, int B, int H, int W) {
    if (c->prec != 4) fail("the learned-boundary training kernels run in FP32 contexts");
    if (B < 1 || H < 5 || W < 20) fail("layer input must have B >= 1, H >= 5 and W >= 20 (16 output cells)");
    if (size_t(B) * H * W * eio::ABC_CH >= (size_t(1) << 31)) fail("layer too large");
}
int abct_grid(long nblk) { return int(std::min<long>((nblk + 3) / 4, 65536L)); }
}  // namespace

int eio_abc_conv_forward(eio_ctx* ctx, const void* a, int B, int H, int W, const float* weights, int leaky,
                         float* z, void* h) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        abct_check(c, B, H, W);
        if (!a || !weights || !z || !h) fail("NULL pointer");
        DeviceGuard dg(c->device);
        const long nblk = long(B) * (H - 4) * ((W - 4 + eio::AT_CELLS - 1) / eio::AT_CELLS);
        hipLaunchKernelGGL(eio::k_abct_conv_fwd, dim3(abct_grid(nblk)), dim3(eio::AT_NTHR), 0, c->stream,
                           static_cast<const unsigned short*>(a), B, H, W, weights, leaky, z,
                           static_cast<unsigned short*>(h));
        EIO_CHECK_HIP(hipGetLastError());
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
    });
}

int eio_abc_conv_backward(eio_ctx* ctx, const void* a, const float* dz, int B, int H, int W, const float* weights,
                          float* da, float* dweights) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        abct_check(c, B, H, W);
        if (!a || !dz || !weights || !dweights) fail("NULL pointer");
        DeviceGuard dg(c->device);
        if (da) {
            const long nblk = long(B) * H * ((W + eio::AT_CELLS - 1) / eio::AT_CELLS);
            hipLaunchKernelGGL(eio::k_abct_conv_bwd_data, dim3(abct_grid(nblk)), dim3(eio::AT_NTHR), 0, c->stream,
                               dz, B, H, W, weights, da);
            EIO_CHECK_HIP(hipGetLastError());
        }
        const long nblk = long(B) * (H - 4) * ((W - 4 + 31) / 32);
        const int nparts = int(std::min<long>(nblk, 1024L));
        const size_t per = size_t(eio::ABC_TAPS) * eio::ABC_CH * eio::ABC_CH;
        float* part = static_cast<float*>(ws_get(c, "abct_part", size_t(nparts) * per * sizeof(float)));
        hipLaunchKernelGGL(eio::k_abct_conv_bwd_weight, dim3(nparts), dim3(eio::AT_NTHR), 0, c->stream,
                           static_cast<const unsigned short*>(a), dz, B, H, W, part);
        EIO_CHECK_HIP(hipGetLastError());
        hipLaunchKernelGGL(eio::k_abct_reduce_weight, dim3(unsigned((per + eio::AT_NTHR - 1) / eio::AT_NTHR)),
                           dim3(eio::AT_NTHR), 0, c->stream, part, nparts, dweights);
        EIO_CHECK_HIP(hipGetLastError());
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
    });
}

int eio_synchronize(eio_ctx* ctx) {
    return guarded([&]() {
        Ctx* c = as_ctx(ctx);
        DeviceGuard dg(c->device);
        EIO_CHECK_HIP(hipStreamSynchronize(c->stream));
    });
}

}  // extern "C"
