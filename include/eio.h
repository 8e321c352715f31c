/* eio.h -- C ABI of the MI355X (gfx950) elastic velocity-stress propagator.
 *
 * Plain C: opaque context, POD descriptors of plain pointers and sizes, int status
 * codes (0 = success) and a thread-local error string.  No torch or C++ types.
 *
 * Every entry point replaces a piece of the reference's C-style GPU seam
 * (paths relative to elastic_iso_lib/elastic_iso_float/prop/include/):
 *   nonlinearPropElasticGpuFunctions.h:5-136  initNonlinearElasticGpu, allocateNonlinearElasticGpu,
 *                                             propShotsElasticFwdGpu, propShotsElasticAdjGpu,
 *                                             deallocateNonlinearElasticGpu
 *   BornElasticGpuFunctions.h:5-66            initBornGpu, allocateBornElasticGpu,
 *                                             BornShotsFwdGpu, BornShotsAdjGpu, deallocateBornElasticGpu
 * Unlike the reference (global per-GPU state, one shot per call, exit(-1) on error),
 * the state lives in a context, a call propagates any number of shots in one batched
 * time loop on one device, and errors are returned.
 *
 * Conventions (identical to the reference):
 *   fields are [nx][nz] with z fastest, index ix*nz + iz, nz/nx include padding + FAT=4;
 *   regular device positions are such linear indices; traces are [nTrace][nt], time fastest;
 *   component k = 0..4 is vx, vz, sxx, szz, sxz; grid g = 0..3 is centre, X, Z, XZ and
 *   component k lives on grid EIO_GRID_OF_COMP[k] = {1, 2, 0, 0, 3}.
 *   Real type is float (precision 4) or double (precision 8), fixed per context.
 */
#ifndef EIO_H
#define EIO_H

#ifdef __cplusplus
extern "C" {
#endif

#define EIO_OK 0
#define EIO_ERR_ARG 1       /* invalid argument (message in eio_last_error) */
#define EIO_ERR_RUNTIME 2   /* HIP runtime error */
#define EIO_ERR_STATE 3     /* call sequence error */

/* flags */
#define EIO_DEVICE_PTRS 1   /* trace / field pointers are device pointers (no H2D/D2H copies) */

typedef struct eio_ctx eio_ctx;

/* FD grid of one model.  sub_fd is the fine sub-sampling already doubled
 * (fdParamElastic.cpp:28-37): ntw = (nts-1)*sub_fd + 1, dtw = dts/sub_fd. */
typedef struct {
    int nz, nx;          /* including padding and FAT */
    int nts, sub_fd;     /* coarse samples, fine steps per coarse step */
    int minPad;          /* sponge width, min over the four sides */
    double dz, dx, dts;  /* spatial sampling [m], coarse time sampling [s] */
    double alphaCos;     /* sponge coefficient, 0.99 by default */
    int surface;         /* 0: sponge on all four sides; 1: free surface at row 2*FAT and no
                          * sponge towards the top (surfaceCondition, fdParamElastic.cpp:18;
                          * NL forward and adjoint only) */
} eio_grid;

/* One shot: regular source/receiver positions per grid and trace pointers per component.
 * NL forward : src_tr = source traces [n_src[g]][nts] (already scaled, coarse), rec_tr = output
 * NL adjoint : rec_tr = input data, src_tr = output model traces
 * Born       : src_tr = fine-sampled source traces [n_src[g]][ntw], rec_tr = data in/out */
typedef struct {
    int model;                 /* id from eio_model_add */
    const int* src_pos[4];
    int n_src[4];
    const int* rec_pos[4];
    int n_rec[4];
    void* src_tr[5];
    void* rec_tr[5];
} eio_shot;

/* Error text of the last failing call on this thread. */
const char* eio_last_error(void);

/* Context on HIP device `device` with real size `precision` (4 or 8).
 * Replaces the per-GPU global pointer tables built by initNonlinearElasticGpu /
 * initBornGpu (nonlinearPropElasticGpuFunctions.cu:85-232, BornElasticGpuFunctions.cu:88-267). */
int eio_ctx_create(int device, int precision, eio_ctx** out);
int eio_ctx_destroy(eio_ctx* ctx);
/* Free the context's device workspace (state, traces, FWI stores' buffers once released): the
 * context keeps and reuses it across calls; the next call allocates again.  No counterpart in
 * the reference, which allocates and frees per call (nonlinearPropElasticGpuFunctions.cu:385-403). */
int eio_ctx_release_workspace(eio_ctx* ctx);
/* Run on an external stream (e.g. torch's current stream); NULL = the context's own stream. */
int eio_ctx_set_stream(eio_ctx* ctx, void* hip_stream);
/* Timing of the last propagation call, from HIP events on the context stream:
 * total device time [ms], number of fused stencil launches, and the mean time of one
 * stencil launch [ms] measured over a bracketed run of them. */
int eio_ctx_last_timing(eio_ctx* ctx, double* total_ms, long long* n_step_launches,
                        double* step_ms);
/* 0 = launch each fine step (default); 1 = run each call's time loop from an executable
 * hipGraph (a chain of kernel nodes) kept by the context per launch-sequence shape: built and
 * instantiated by the first call of a shape, replayed by later ones after updating only the
 * kernel nodes whose arguments changed (the 4 most recently used shapes are kept, until the
 * context is destroyed; every argument of every node is compared, so a kept graph stays valid
 * when workspace buffers move).  No
 * counterpart in the reference, which launches every kernel of every fine step
 * (nonlinearPropElasticGpuFunctions.cu:588-617). */
int eio_ctx_set_graph(eio_ctx* ctx, int use_graph);
/* Graph-mode counters of the context: graphs built (instantiated), graph launches, kernel
 * nodes updated in place before a replay. */
int eio_ctx_graph_info(eio_ctx* ctx, long long* builds, long long* launches, long long* node_updates);

/* FP32 step-kernel variant (tile of 64 x WX*R cells, WX waves per workgroup, R x-cells per
 * lane, requested occupancy) used by the next calls.  FP64 contexts ignore it.  No counterpart
 * in the reference (fixed 16x16 blocks, varDeclare.h:5). */
int eio_ctx_num_variants(void);
int eio_ctx_variant_info(int variant, int* R, int* WX, int* occ);
int eio_ctx_set_variant(eio_ctx* ctx, int variant);
/* Fine steps per launch of the FP32 forward passes: 1 (one fused step per launch) or 2 (two
 * steps per launch, temporally blocked: 50 instead of 80 bytes of HBM traffic per
 * cell-update, four state levels per shot), and the variant of the two-step kernel
 * (owned tile 56 x WX*R2), or EIO_VARIANT_AUTO: per call, 56 x 8 tiles with the streamed
 * stage at 8 waves per SIMD (variant 12) for a batch of fewer than 1024 tiles of 56 x 32 and
 * for an NL forward that runs in several shot groups (eio_ctx_set_shot_group_bytes), else
 * 56 x 32 (variant 4).
 * It applies to the NL forward (without
 * saved wavefield, free surface or learned boundary), the whole-record FWI forward and the
 * Born source and scattered passes.  No counterpart in the reference. */
#define EIO_VARIANT_AUTO (-1)
int eio_ctx_num_time_block_variants(void);
int eio_ctx_time_block_info(int variant, int* R2, int* WX, int* occ);
int eio_ctx_set_time_block(eio_ctx* ctx, int steps, int variant);
/* steps = 3: the FP32 NL forward (without saved wavefield, free surface, learned boundary or
 * snapshots) takes THREE fine steps per launch (k_step3_fwd: 100 B per three cell-updates,
 * owned tile 48 x WX*R3, u_cur staged with a 12-cell halo), `variant` indexing the three-step
 * table (EIO_VARIANT_AUTO: 48 x 8 tiles below 1024 tiles of 56 x 32, else 48 x 32); the
 * other forward passes then take two steps per launch with the automatic two-step variant.
 * No counterpart in the reference. */
int eio_ctx_num_time_block3_variants(void);
int eio_ctx_time_block3_info(int variant, int* R3, int* WX, int* occ);
/* Lane layout of two-step variant `variant`: 0 = lane l owns grid row z0-4+l of several
 * columns and the current level is staged whole in LDS; 1 = the same lanes, the current level
 * streamed through two LDS field buffers (two workgroups per CU at 56 x 32); 2 = a lane owns a
 * 4-row quad of one column (16-byte loads and stores), streamed stage.  Layouts 1 and 2 serve
 * the NL forward and the FWI segment recompute; the other passes of such a variant run a
 * layout-0 kernel of the same tile width.  -1 for an unknown variant.  No counterpart in the
 * reference. */
int eio_ctx_time_block_layout(int variant);
/* Reverse fine steps per launch of the FP32 adjoint passes (eio_born_adjoint,
 * eio_fwi_gradient and eio_nl_adjoint without free surface or saved wavefield): 1 (one fused
 * step per launch, default) or 2 (both steps of a pair in one launch, four state levels per
 * shot; the Born adjoint fields of the pair are extracted with at most one read-modify-write
 * of each coarse slot), and the variant of the two-step adjoint kernel (owned tile
 * 56 x WX*R2), or EIO_VARIANT_AUTO: per call, 56 x 8 tiles (variant 10) for a batch of fewer
 * than 1024 tiles of 56 x 32, else 56 x 32 (variant 7).  FP64 contexts ignore it.  No counterpart in the reference (its receiver pass,
 * BornElasticGpuFunctions.cu:262-370, and its NL adjoint take one step per launch). */
int eio_ctx_num_adjoint_time_block_variants(void);
int eio_ctx_adjoint_time_block_info(int variant, int* R2, int* WX, int* occ);
int eio_ctx_set_adjoint_time_block(eio_ctx* ctx, int steps, int variant);
/* The FP32 two-step NL forward (the time block 2 path) in ONE persistent launch per call
 * instead of one launch per pair of fine steps: mode 0 = off (default), 1 = on, 2 = when every
 * tile of the batch gets its own co-resident workgroup.  The grid is launched cooperatively;
 * tiles wait for their neighbours' previous pair through per-tile flags (no grid barrier).
 * spin_limit > 0 replaces the number of polls a wait may take before the launch is abandoned
 * with EIO_ERR_RUNTIME (default 2^20).  No counterpart in the reference, which launches every
 * kernel of every fine step (nonlinearPropElasticGpuFunctions.cu:588-617). */
int eio_ctx_set_persistent(eio_ctx* ctx, int mode, int spin_limit);
/* Tuning runs: NL forward and adjoint loops stop after `max_steps` fine steps (0: no limit).
 * Outputs of a limited call are partial.  No counterpart in the reference. */
int eio_ctx_set_step_limit(eio_ctx* ctx, int max_steps);
/* Shot groups of the FP32 NL forward: the ordered shots of a call are split into consecutive
 * groups whose per-step working set (the state levels a step touches and the media: 15 values
 * per cell with one fine step per launch, 25 with two or three) fits `bytes`, and each group
 * runs all its steps before the next one starts, so its working set stays in the MI355X's
 * 256 MB Infinity Cache between steps.  EIO_GROUP_AUTO (the default): 192 MiB (shared by the
 * groups running at once on the call's streams, eio_ctx_set_shot_group_streams); 0: one group
 * (every shot advances each launch).  Results are those of the ungrouped call; the persistent
 * launch and runs with snapshots, a saved wavefield, a free surface or a learned boundary take
 * one group.  No counterpart in the reference (one shot per call, NLF:561-647). */
#define EIO_GROUP_AUTO (-1LL)
int eio_ctx_set_shot_group_bytes(eio_ctx* ctx, long long bytes);
/* Streams the shot groups of a call run on: 2 (the default) runs the groups on the context's
 * stream and a second stream of the library (default priority), each group with half the
 * budget, so that one stream's launch tails and gaps are filled by the other's work (the
 * cache holds one group of each); the call's stream waits for both before the call returns.
 * 1: one stream.  Graph recording (eio_ctx_set_graph) always uses one.  No counterpart in the
 * reference. */
int eio_ctx_set_shot_group_streams(eio_ctx* ctx, int streams);
/* Number of shot groups of the context's last NL forward call.  No counterpart in the reference. */
int eio_ctx_last_shot_groups(eio_ctx* ctx, int* groups);

/* The grid checks of eio_model_add, on the host only (no device, no context): sizes, sampling,
 * sponge width (the reference's messages, fdParamElastic.cpp:29-32,
 * nonlinearPropElasticGpuFunctions.cu:191-194), and 5*nz*nx*precision < 2^31 bytes (the
 * kernels address each 5-field buffer with 32-bit offsets).  EIO_ERR_ARG when a check fails. */
int eio_grid_check(const eio_grid* grid, int precision);

/* Register a model: grid constants (FD coefficients, time filter, damping table) and the
 * five scaled media rhoxDtw, rhozDtw, lamb2MuDtw, lambDtw, muxzDtw, each nz*nx.
 * Replaces initNonlinearElasticGpu + allocateNonlinearElasticGpu
 * (nonlinearPropElasticGpuFunctions.h:8-16; nonlinearPropElasticGpuFunctions.cu:85-265). */
int eio_model_add(eio_ctx* ctx, const eio_grid* grid, const void* const media[5], int flags,
                  int* model_id);
/* Replace the media of a model (setBackground, nonlinearPropElasticShotsGpu.h:61). */
int eio_model_set_media(eio_ctx* ctx, int model_id, const void* const media[5], int flags);
int eio_model_remove(eio_ctx* ctx, int model_id);
/* Replace the model's sponge table (minPad factors in the context's precision, host pointer;
 * entry j multiplies both time levels of the cells j cells in from the grid edge's update
 * ring, as the cosine table of kernelsGpuElastic.cu:1192-1221 does).  Factors must lie in
 * (0, 1], so any table only removes energy.  No counterpart in the reference, whose table is
 * the cosine taper (used to fit a narrow sponge to wide-sponge runs). */
int eio_model_set_damping(eio_ctx* ctx, int model_id, const void* table);

/* Nonlinear forward of nshot shots in one batched time loop; rec_tr is overwritten.
 * Replaces propShotsElasticFwdGpu (nonlinearPropElasticGpuFunctions.h:20-26, .cu:561-648). */
int eio_nl_forward(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags);
/* eio_nl_forward that allocates nothing: no counterpart in the reference; it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The call's workspace (state levels, shot/tile/entry
 * tables) comes from a 64 MiB device array of the library (one per GPU, held until the
 * process ends) instead of hipMalloc; a batch needing more is rejected with EIO_ERR_ARG
 * before any launch.  EIO_DEVICE_PTRS is required (traces on the device); no graph is used.
 * Models must be registered beforehand (eio_model_add allocates).  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other threads or contexts) race. */
int eio_nl_forward_noalloc(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags);
/* Nonlinear adjoint; src_tr is overwritten.
 * Replaces propShotsElasticAdjGpu (nonlinearPropElasticGpuFunctions.h:28-34, .cu:932-1028). */
int eio_nl_adjoint(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags);
/* The same calls, also saving the coarse-sampled wavefield of shot `shot` (index into shots)
 * into `wavefield` [nts][5][nx][nz] (host pointer, or device pointer with EIO_DEVICE_PTRS):
 * every fine step's update-region values accumulated into slices its and its+1 with the time
 * filter, as the receivers are; the adjoint wavefield is unscaled (the reference scales it on
 * the host by rhoxDtw, rhozDtw, 2*dtw and 1/(dx*dz)).  Replace
 * propShotsElastic{Fwd,Adj}GpuWavefield[Streams] (nonlinearPropElasticGpuFunctions.h:38-136,
 * .cu:649-927, 1029-1263; rows and columns of the streamed variant). */
int eio_nl_forward_wavefield(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags, int shot,
                             void* wavefield);
int eio_nl_adjoint_wavefield(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags, int shot,
                             void* wavefield);

/* Learned absorbing boundary (a method change beyond the reference; no counterpart there,
 * whose only boundary is the cosine sponge of kernelsGpuElastic.cu:1192-1221).  Attach
 * (on = 1 or 2) or detach (on = 0) a small CNN that, after every fine step of an FP32 NL forward
 * of this model, writes the new level's 4-cell FAT frame (the cells outside the update
 * region that the stencil reads, zero without it) from the first update cells of the new and
 * the previous level; the model's own sponge (minPad) stays active.  weights: FP32, per edge
 * (x-min, x-max, z-min, z-max) the three 5x5 layers in torch Conv2d layout [out][in][5][5]
 * with (out, in) = (16, 10), (16, 16), (5, 16), first kernel axis along x; inputs are the 5
 * fields of the new level divided by scale[0..4] and, divided by scale[5..9], the 5
 * differences new - previous level (on = 1) or the 5 fields of the previous level (on = 2);
 * outputs are times scale[0..4].  on + 10 * depth (depth 1..4) writes only the `depth` frame
 * rows next to the update region and leaves the outer ones zero.  Rounded to BF16 on upload.  Not
 * available with the free surface or in FP64. */
int eio_model_set_abc(eio_ctx* ctx, int model_id, int on, const float* weights, const float* scale);

/* eio_nl_forward that also copies the raw state of shot `shot` after each fine step
 * steps[i] (0-based, increasing): out [nsnap][2][5][nx][nz], the older then the newer level
 * (device pointer; EIO_DEVICE_PTRS required).  The older level is stored as the fused step
 * kernel keeps it: without its second sponge factor (applied when it is next read).
 * No counterpart in the reference: training data of the learned boundary. */
int eio_nl_forward_snapshots(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags, int shot, int nsnap,
                             const int* steps, void* out);

/* Training kernels of the learned-boundary network (no counterpart in the reference): one
 * 5x5 "valid" convolution layer of 16 channels, NHWC, BF16 activations and weights, FP32
 * accumulation on MFMA, the long axis W (>= 20).  Device pointers; the calls return when the
 * kernels have finished.  weights: FP32 [16 out][16 in][5][5] (unused channels zero).
 * forward:  a [B][H][W][16] BF16 -> z [B][H-4][W-4][16] FP32 and h = BF16(leaky(z)) (leaky = 1,
 *           slope 0.1) or BF16(z) (leaky = 0), the same rounding points as the inference kernel.
 * backward: dz [B][H-4][W-4][16] FP32 (the gradient with respect to z) -> da [B][H][W][16] FP32
 *           (NULL: skipped) and dweights [16][16][5][5] FP32, both computed with dz rounded to
 *           BF16; the weight gradient is summed in a fixed order (deterministic). */
int eio_abc_conv_forward(eio_ctx* ctx, const void* a, int B, int H, int W, const float* weights, int leaky,
                         float* z, void* h);
int eio_abc_conv_backward(eio_ctx* ctx, const void* a, const float* dz, int B, int H, int W, const float* weights,
                          float* da, float* dweights);

/* One fine step without sources or receivers of nstate states of model_id, in place:
 * states [nstate][2][5][nx][nz] (device, EIO_DEVICE_PTRS) hold (older, newer) as the
 * snapshots do and become (newest, newer); an attached learned boundary is applied.
 * No counterpart in the reference: training data of the learned boundary. */
int eio_nl_step_states(eio_ctx* ctx, int model_id, int nstate, void* states, int flags);

/* Born forward: dm = 5 staggered and scaled perturbations (drhox*rhoxDtw, drhoz*rhozDtw,
 * dlame*2dtw, dmu*2dtw, dmuxz*2dtw), nz*nx each, shared by all shots (all shots must use
 * the same model).  rec_tr is overwritten.
 * Replaces BornShotsFwdGpu (BornElasticGpuFunctions.h:12-23, .cu:719-1098). */
int eio_born_forward(eio_ctx* ctx, int nshot, const eio_shot* shots, const void* const dm[5],
                     int flags);
/* Born adjoint: accumulates (+=) the imaging condition of all shots into dm[5].
 * Replaces BornShotsAdjGpu (BornElasticGpuFunctions.h:25-36, .cu:1100-1454). */
int eio_born_adjoint(eio_ctx* ctx, int nshot, const eio_shot* shots, void* const dm[5], int flags);

/* Checkpointed FWI gradient, in two calls around the residual.  The reference computes the
 * gradient as the Born adjoint of the data residual at the current model (efwiMain.py:229-476
 * calling BornElasticShotsGpu::adjoint, BornElasticShotsGpu.cpp:228-378, whose source pass keeps
 * Vx, Vz at every coarse sample, BornElasticGpuFunctions.cu:1128-1191); these entry points replace
 * the NL forward + Born adjoint pair of one gradient.
 *
 * eio_fwi_forward: NL forward of all shots (same model; src_tr coarse [nReg][nts] as for
 * eio_nl_forward; rec_tr is overwritten with the predicted data), saving the batch state every
 * `segment` coarse intervals (0: ceil(sqrt(nts-1))) into a store owned by the context.
 * *handle names the store. */
int eio_fwi_forward(eio_ctx* ctx, int nshot, const eio_shot* shots, int flags, int segment,
                    int* handle);
/* eio_fwi_gradient: Born adjoint at the model of the store, with rec_tr = the adjoint source
 * (data residual, regular traces) and src_tr = the same sources as the forward call; the
 * forward wavefield is recomputed segment by segment from the checkpoints.  Accumulates (+=)
 * the stacked imaging condition into dm[5] (same slices as eio_born_adjoint).  The shots must
 * have the geometry and model of the forward call, and the model's media must not change in
 * between.  The store stays valid for further gradient calls until released. */
int eio_fwi_gradient(eio_ctx* ctx, int handle, int nshot, const eio_shot* shots, void* const dm[5],
                     int flags);
/* Store facts: segment length (coarse intervals), number of checkpoints, device bytes held. */
int eio_fwi_info(eio_ctx* ctx, int handle, int* segment, int* nckpt, unsigned long long* bytes);
int eio_fwi_release(eio_ctx* ctx, int handle);

/* Diagnostics: copy `bytes` (multiple of 16) from src to dst (device pointers) `reps` times on
 * the context stream; *ms = mean kernel time.  kind 1: one 16-byte vector per thread with
 * nontemporal stores (the achievable HBM copy bandwidth); kind 0: a grid-stride 16-byte copy
 * (the reference kernel of the HBM byte-counter calibration); kind 2: one 4-byte word per
 * thread (the access width of the stencil kernels' per-cell loads and stores).  No
 * counterpart in the reference. */
int eio_copy_probe(eio_ctx* ctx, void* dst, const void* src, unsigned long long bytes, int reps, int kind,
                   double* ms);

/* Device memory helpers for callers without their own allocator (bytes). */
int eio_device_alloc(eio_ctx* ctx, unsigned long long bytes, void** ptr);
int eio_device_free(eio_ctx* ctx, void* ptr);
int eio_memcpy(eio_ctx* ctx, void* dst, const void* src, unsigned long long bytes, int kind);
/* kind: 1 = host->device, 2 = device->host, 3 = device->device */
int eio_synchronize(eio_ctx* ctx);
/* Page-locked host memory (bytes; *ptr = NULL for 0 bytes) for buffers that calls with host
 * pointers copy to or from the device every time, such as regular-grid traces: the copies
 * are direct DMA instead of a pass through the runtime's pageable staging buffer.  Usable
 * with every context; free with eio_host_free (which waits for the device).  No
 * counterpart in the reference, whose host trace buffers are pageable. */
int eio_host_alloc(unsigned long long bytes, void** ptr);
int eio_host_free(void* ptr);

#ifdef __cplusplus
}
#endif
#endif /* EIO_H */
