/*
 * feynsimul.h -- C ABI of the MI355X PIMC engine (libfeynsimul.so).
 *
 * Plain pointers and sizes only.  Device buffers are allocated by the caller
 * (the engine borrows them); every call that touches the GPU is ordered on
 * the hipStream_t passed as `stream` (NULL = the null stream).  Every entry
 * point returns FS_OK (0) or an error code; fs_last_error() gives the
 * message of the last failure on the calling thread.
 *
 * The reference drives one OpenCL program per configuration from Python
 * (FeynSimul/kernel.py); the entry points below are what its driver would
 * bind in place of pyopencl:
 *   fs_create / fs_destroy        <- PIMCKernel.__init__ program build,
 *                                    kernel.py:346-382 (cl.Program(...).build)
 *   fs_build_log                  <- PIMCKernel.getBuildLog, kernel.py:778-785
 *   fs_run                        <- PIMCKernel.run, kernel.py:459-506
 *                                    (the `metropolis` launch, kernel.c:514-837)
 *   fs_seed_xorshift              <- seed table fill, kernel.py:421-426
 *   fs_ranlux_init                <- initRanlux / groupInitRanlux,
 *                                    kernel.py:508-554, kernel.c:152-166
 *   fs_xorshift_draw, fs_ranlux_draw, fs_rng_bench, fs_eval_system
 *                                 <- stream and expression dumps of
 *                                    examples/ranlux_test_kernel.c:50-212
 *                                    (test support)
 */
#ifndef FEYNSIMUL_H
#define FEYNSIMUL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_ABI_VERSION 6

enum {
    FS_OK = 0,
    FS_ERR_INVALID = 1,   /* bad argument or configuration */
    FS_ERR_COMPILE = 2,   /* hiprtc rejected the system source; see fs_build_log */
    FS_ERR_HIP = 3,       /* a HIP runtime call failed */
    FS_ERR_NO_DEVICE = 4  /* no GPU visible */
};

enum { FS_SAMPLER_EXACT = 0, FS_SAMPLER_WAVE = 1 };

/* Where the wave sampler keeps a walker's path during a launch: AUTO picks
 * registers when 64 <= N <= 1024 and the path plus its potential cache take
 * at most 64 VGPRs per lane, else memory (then pot_cache is required). */
enum { FS_STORAGE_AUTO = 0, FS_STORAGE_MEMORY = 1, FS_STORAGE_REGISTERS = 2 };

/* Memory storage keeps the potential of every slice in pot_cache (read for
 * the old slice of a move, written on acceptance) instead of evaluating it
 * again; that pays when the potential is dearer than 8 bytes of HBM traffic
 * per slice-update.  AUTO: cache when DOF > 1. */
enum { FS_POTCACHE_AUTO = 0, FS_POTCACHE_ON = 1, FS_POTCACHE_OFF = 2 };

/* One configuration = one compiled module (the reference rebuilds its
 * program for every change of these, kernel.py:213-382). */
typedef struct fs_config {
    int32_t abi_version;        /* FS_ABI_VERSION */
    int32_t sampler;            /* FS_SAMPLER_* */
    int32_t N;                  /* slices, power of two */
    int32_t S;                  /* bisection level, 2^S <= N */
    int32_t DOF;                /* coordinates per slice */
    int32_t walkers_per_group;  /* nbrOfWalkersPerWorkGroup */
    int64_t n_walkers;          /* nbrOfWalkers */
    int32_t parallelize_path;   /* enableParallelizePath (exact sampler) */
    int32_t local_size;         /* work-group size that fixes group rows when !parallelize_path */
    int32_t double_precision;   /* enableDouble */
    int32_t ranlux;             /* enableRanlux */
    int32_t ranlux_lux;         /* RANLUXCL luxury value: 0..4, or p >= 24 */
    int32_t fast_math;          /* relaxed math as with -cl-fast-relaxed-math */
    uint64_t ranlux_int_range;  /* ranluxIntMax + 1 (scale of integer draws) */
    int32_t n_ops;              /* operators */
    int32_t n_corr;             /* correlators */
    int32_t cor_lags;           /* correlator lags (wave sampler; exact uses N/2) */
    int32_t bins;               /* enableBins */
    int32_t bin_res;            /* binResolutionPerDOF */
    int32_t waves_per_block;    /* wave sampler launch shape (0 = automatic) */
    double beta;
    double x_min, x_max;
    double op_norm;             /* exact sampler: 1/(operatorRuns N walkers/threads) */
    int32_t path_storage;       /* wave sampler: FS_STORAGE_* */
    int32_t n_hist_obs;         /* wave sampler: scalar observables binned into one histogram */
    int32_t hist_bins;          /* bins of that histogram */
    double hist_lo, hist_hi;    /* its range; values outside are not counted */
    int32_t no_bisection;       /* !enableBisection */
    int32_t path_shift;         /* enablePathShift (exact sampler, one item per walker) */
    int32_t single_node;        /* enableSingleNodeMove (exact sampler, one item per walker) */
    int32_t max_waves_per_simd; /* wave sampler: occupancy cap (0 = as registers and LDS allow), see
                                  * fs_set_waves_per_simd */
    double alpha;               /* single-node step width */
    double ps_alpha;            /* path-shift offset width */
    int32_t potential_cache;    /* wave sampler, memory storage: FS_POTCACHE_* */
} fs_config;

/* Device buffers of one launch; unused entries may be NULL.  Layouts:
 *   paths     FLOAT [walkers][DOF][N]
 *   exact:  accepts u32[threads]; seeds u32[threads+1][4]; seeds_snap same;
 *           rl, grl 112-byte RANLUX states [threads]; old_paths FLOAT
 *           [threads][(2^S-1) DOF]; op_means FLOAT[threads][n_ops];
 *           cor_means FLOAT[walkers][n_corr][N/2]
 *   wave:   accepts u32[walkers]; seeds u32[walkers*64 + walkers][4] (lane rows,
 *           then group rows); rl [walkers*64]; grl [walkers];
 *           op_means f64[walkers][n_ops]; op_totals f64[walkers][n_ops][2];
 *           series f32[measurements][walkers][n_ops];
 *           cor_means f64[walkers][n_corr][cor_lags][2];
 *           pot_cache FLOAT[walkers][N] (memory storage)
 *   bins u32[bin_res^DOF] */
typedef struct fs_buffers {
    void *paths;
    void *accepts;
    void *seeds;
    void *seeds_snap;
    void *rl;
    void *grl;
    void *old_paths;
    void *op_means;
    void *op_totals;
    void *series;
    void *cor_means;
    void *bins;
    void *pot_cache;
    void *obs_hist;             /* u32[hist_bins] (wave sampler, n_hist_obs > 0) */
} fs_buffers;

typedef struct fs_engine fs_engine;

int fs_abi_version(void);
const char *fs_last_error(void);
int fs_device_count(int *count);

/* Compile the per-configuration module.  `system_src` is the system section:
 * the user's helper code and the functions potential / fs_operator /
 * fs_correlator with their *_v array wrappers.  Needs a GPU to load the
 * module unless `compile_only` is set (then only the build log is kept). */
int fs_create(const fs_config *cfg, const char *system_src, int compile_only, fs_engine **out);
int fs_destroy(fs_engine *e);
/* The path storage the module was built with (FS_STORAGE_REGISTERS or _MEMORY). */
int fs_path_storage(const fs_engine *e, int *storage);
/* The path storage a module for this configuration would use, without building it (the batched
 * scheduler groups types by it).  No reference counterpart: the reference always stages paths the
 * way its flags say (kernel.c:582-603). */
int fs_storage_of(const fs_config *cfg, int *storage);
/* Copies the build log (NUL-terminated, truncated to len) and returns its full length in *needed. */
int fs_build_log(const fs_engine *e, char *buf, size_t len, size_t *needed);
/* Copies the generated module source (for verbose dumps). */
int fs_source(const fs_engine *e, char *buf, size_t len, size_t *needed);

/* Occupancy cap of the wave sampler's launches: at most `cap` wavefronts per
 * SIMD (0 = no cap), enforced by padding each block's LDS request; caps that
 * would need fewer than two blocks per CU give two.  No reference counterpart
 * (the reference fixes its work-group shape per build, kernel.py:192-211); the
 * autotuner uses it. */
int fs_set_waves_per_simd(fs_engine *e, int32_t cap);

/* Launch shape and resource use of the module's sampler kernel against the
 * device limits: what the reference's getMaxKernelWGSize and exceedsLimits
 * read from OpenCL (kernel.py:732-775: WORK_GROUP_SIZE, LOCAL_MEM_SIZE,
 * MAX_WORK_GROUP_SIZE, MAX_WORK_ITEM_SIZES). */
typedef struct fs_launch_info_t {
    int32_t threads_per_block;          /* work-items of one launch's work-group */
    int32_t max_threads_per_block;      /* the kernel's limit (its WORK_GROUP_SIZE) */
    int32_t lds_bytes;                  /* static + dynamic LDS of one work-group */
    int32_t registers;                  /* VGPRs per work-item */
    int32_t lds_limit;                  /* device LDS per work-group */
    int32_t device_max_threads_per_block;
    int32_t device_max_block_dim[3];
} fs_launch_info_t;
int fs_launch_info(const fs_engine *e, fs_launch_info_t *info);

/* One launch of operator_runs x metro_steps sweeps (the reference `run`). */
int fs_run(fs_engine *e, const fs_buffers *b, int32_t operator_runs, int32_t metro_steps,
           int32_t series_offset, void *stream);
/* fs_run with flags.  FS_RUN_POT_CACHE_VALID: the caller vouches that b->pot_cache still holds the potential
 * at every slice of b->paths, as the previous launch of this engine on these buffers left it (the paths have
 * not been written since); the wave sampler with memory storage then skips recomputing it at the start of
 * the launch (a pass over every path; the results are the same bit for bit).  No counterpart in the
 * reference, which evaluates the potential afresh for every move (kernel.c:376-377). */
#define FS_RUN_POT_CACHE_VALID 1
int fs_run_flags(fs_engine *e, const fs_buffers *b, int32_t operator_runs, int32_t metro_steps,
                 int32_t series_offset, int32_t flags, void *stream);

/* Evaluate the compiled potential and operators at n points x[n][DOF]
 * (FLOAT); writes f64 pot[n] and ops[n][n_ops]. */
int fs_eval_system(fs_engine *e, const void *x, void *pot, void *ops, int64_t n, void *stream);

/* rows[i] = xorshift128 state from splitmix64 state seed + first + i. */
int fs_seed_xorshift(void *rows, int64_t n, uint64_t seed, int64_t first, void *stream);
/* tab[i] = ranluxcl_init(seeds[i]); seeds is a device u64 array. */
int fs_ranlux_init(void *tab, const void *seeds, int64_t n, void *stream);
/* ndraw draws per row without advancing rows: u32, f32 and f64 (any may be NULL). */
int fs_xorshift_draw(const void *rows, void *out_u32, void *out_f32, void *out_f64, int64_t n, int32_t ndraw,
                     void *stream);
/* nquad four-value RANLUX draws per state (advancing the states); lux 0..4. */
int fs_ranlux_draw(void *tab, void *out_f32, void *out_f64, int64_t n, int32_t nquad, int32_t lux, void *stream);
/* Stream throughput, as the reference's RNG benchmark kernels
 * (examples/ranlux_test_kernel.c:90-200, timings in
 * examples/benchmark_xorshift_vs_ranlux.py:12-26): each of n streams draws
 * `count` uniforms through the samplers' wrappers (xorshift128 randFloat,
 * kernel.c:233-237; RANLUX ranluxWrapper with its per-launch float4 buffer,
 * kernel.c:180-207), FP32 or FP64, and writes the XOR of their bit patterns
 * to check[i] (u32) instead of the values; the states advance.  kind is
 * FS_RNG_XORSHIFT (states: uint32 rows [n][4]) or FS_RNG_RANLUX (states:
 * 112-byte RANLUX states [n], lux 0..4). */
enum { FS_RNG_XORSHIFT = 0, FS_RNG_RANLUX = 1 };
int fs_rng_bench(int32_t kind, int32_t lux, int32_t double_precision, void *states, void *check, int64_t n,
                 int64_t count, void *stream);

/* HBM ceilings, no reference counterpart: the measured bandwidths the
 * samplers' rooflines are read against.  A grid-stride stream of n_vec16
 * 16-byte vectors in `blocks` workgroups of 256 work-items that copies src to
 * dst (kind 0), only reads src (kind 1) or only writes dst (kind 2). */
int fs_hbm_probe(int32_t kind, const void *src, void *dst, int64_t n_vec16, int32_t blocks, void *stream);

/* ---- batched systems: one persistent launch over several configurations ----
 * No reference counterpart as a call: the reference compiles one program per
 * (system, N, S) and its run loop relaunches them one after another
 * (kernel.py:346-382, pimc_utils.py:230-260).  Every configuration must use
 * the wave sampler and agree on precision, RNG kind and luxury; buffers are
 * given per type in type order.  `order` lists type ids in dispatch order
 * (NULL: as given); `grid_blocks` 0 = enough blocks to fill the device. */
typedef struct fs_batch fs_batch;
typedef struct fs_graph fs_graph;
/* kernel_of_type[t] in [0, 4) puts type t into one of up to four persistent
 * sweep kernels (numbered without gaps), so types of different register
 * demand run at their own occupancy; the kernels run one after another on
 * the launch stream.  NULL: one kernel for every type.  waves_of_kernel[j]
 * in [0, 8] asks the compiler to fit kernel j's registers to that many
 * wavefronts per SIMD (0 or NULL: its own choice). */
int fs_batch_create(const fs_config *cfgs, const char *const *system_srcs, int32_t ntypes,
                    int32_t waves_per_block, const int32_t *kernel_of_type, const int32_t *waves_of_kernel,
                    int compile_only, fs_batch **out);
int fs_batch_destroy(fs_batch *b);
int fs_batch_path_storage(const fs_batch *b, int32_t type, int *storage);
int fs_batch_build_log(const fs_batch *b, char *buf, size_t len, size_t *needed);
int fs_batch_run(fs_batch *b, const fs_buffers *bufs, const int32_t *order, int32_t operator_runs,
                 int32_t metro_steps, int32_t series_offset, int32_t grid_blocks, void *stream);
/* fs_batch_run that allocates nothing and uses no device memory of the
 * batch's own: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The dispatch tables and the work
 * counter live in one fixed 16 KiB static device array per GPU (at most 64
 * types), which each call overwrites (tables copied on `stream`, counter
 * cleared by a kernel) and every call on that GPU shares: calls that overlap
 * on one GPU race.  The batch's device must be current. */
int fs_batch_run_noalloc(fs_batch *b, const fs_buffers *bufs, const int32_t *order, int32_t operator_runs,
                         int32_t metro_steps, int32_t series_offset, int32_t grid_blocks, void *stream);
/* Records n_launches consecutive fs_batch_run launches (measurement series
 * offsets advancing by operator_runs) into a HIP graph; replay with
 * fs_graph_launch.  Buffers and order are fixed at capture. */
int fs_batch_capture(fs_batch *b, const fs_buffers *bufs, const int32_t *order, int32_t operator_runs,
                     int32_t metro_steps, int32_t series_offset, int32_t n_launches, int32_t grid_blocks,
                     void *stream, fs_graph **out);
int fs_graph_launch(fs_graph *g, void *stream);
int fs_graph_destroy(fs_graph *g);

/* ---- learned bisection proposal (BF16 MFMA MLP, MH-corrected) ----
 * No reference counterpart: it replaces the Levy midpoint draw of
 * kernel.c:338-360 with a conditional Gaussian from an MLP
 * (2 DOF + 1 -> 128 -> 128 -> 128 -> 2 DOF) and extends the accept test of
 * kernel.c:376-377 with the proposal density ratio, so any weights keep
 * detailed balance.  FP32 paths only, 1 <= S <= 6 (or the whole-path move of
 * fs_lp_config.coarse), DOF <= 15.  Random numbers
 * come from Philox4x32-10 keyed by `seed`.  The caller allocates the buffers
 * with the sizes of fs_lp_info. */
typedef struct fs_lp fs_lp;
typedef struct fs_lp_config {
    int32_t abi_version;      /* FS_ABI_VERSION */
    int32_t atoms3d;          /* inputs centred on the atoms' mean (DOF = 3 x atoms) */
    int32_t batch;            /* training rows per Adam step, a multiple of 2048 */
    int32_t segs_per_wave;    /* segments per wavefront in a sweep (0 = automatic) */
    float in_scale;           /* input scale of the coordinates */
    float lr, beta1, beta2, adam_eps;
    uint64_t seed;            /* Philox key */
    int32_t coarse;           /* 0: segment proposal (levels 1 .. S of 2^S-slice segments);
                                 1: whole-path move, N = 64 x 2^k: the network proposes the nodes every
                                 64 slices (levels 1 .. k of the whole cyclic path), a Levy bridge the
                                 slices between them, one Metropolis-Hastings decision per path */
    int32_t width;            /* active hidden units, a multiple of 32 up to 128 (0 = 128): inference and
                                 training compute on these only; the parameters keep the 128-wide layout,
                                 and units beyond the active width must be zero (training keeps them zero,
                                 their gradients being exactly zero) */
    float bridge_omega2;      /* whole-path move: omega^2 of a harmonic reference, omega^2 |x|^2 / 2 per slice,
                                 whose exact bridge fills the intervals between the nodes (the accept test then
                                 scores V - omega^2 |x|^2 / 2); 0 = the free-particle (Levy) bridge (ABI 6) */
} fs_lp_config;
typedef struct fs_lp_info {
    int64_t n_params;         /* FP32 master parameters (and each Adam moment) */
    int64_t image_bytes;      /* packed inference image */
    int64_t train_bf16;       /* BF16 training matrices */
    int64_t train_bias;       /* FP32 training biases */
    int64_t workspace_bytes;  /* training workspace */
    int64_t loss_parts;       /* FP32 partial losses of a training step (sum / batch = mean NLL) */
    int32_t segs_per_wave;
    int32_t waves_per_block;  /* of a sweep launch (the whole-path move: its network kernel, or the fused
                                 move for paths of at most 1024 floats) */
    int32_t lds_bytes;        /* dynamic LDS of a sweep block */
    int32_t grid_blocks;      /* sweep blocks (persistent, except the fused move's one block per 4 walkers) */
} fs_lp_info;
typedef struct fs_lp_buffers {
    void *params, *adam_m, *adam_v;  /* FP32 [n_params] */
    void *image;                     /* [image_bytes] */
    void *train_w;                   /* BF16 [train_bf16] */
    void *train_b;                   /* FP32 [train_bias] */
    void *workspace;                 /* [workspace_bytes] */
    void *loss_parts;                /* FP32 [loss_parts] */
    void *trace;                     /* optional FP32 [walkers][N / 2^S][4] of the last sweep:
                                        log q(old) - log q(new), dK, dV, log alpha (test support) */
} fs_lp_buffers;
/* Compile the proposal module for one configuration (wave-sampler system
 * source, FP32). */
int fs_lp_create(const fs_config *cfg, const char *system_src, const fs_lp_config *lp, int compile_only,
                 fs_lp **out);
int fs_lp_destroy(fs_lp *p);
int fs_lp_get_info(const fs_lp *p, fs_lp_info *info);
int fs_lp_build_log(const fs_lp *p, char *buf, size_t len, size_t *needed);
/* Rebuild the BF16 inference image and training matrices from params. */
int fs_lp_pack(fs_lp *p, const fs_lp_buffers *b, void *stream);
/* n_sweeps learned sweeps (sweep counters sweep0 ...) over paths [walkers][DOF][N];
 * accepted segments (whole-path move: accepted paths) are added to accepts[walkers].
 * The whole-path move of a path of at most 1024 floats makes all n_sweeps moves in one launch. */
int fs_lp_sweep(fs_lp *p, void *paths, void *accepts, const fs_lp_buffers *b, uint64_t sweep0, int32_t n_sweeps,
                void *stream);
/* fs_lp_sweep of the whole-path move (n_sweeps >= 1 moves), whose last move also records one measurement of the
 * kernel's operators on the resulting paths into kb (op_means, op_totals, series at series_offset, the buffers
 * of fs_run), as fs_run(operator_runs = 1, metro_steps = 0) records it, without that launch's read of every
 * path; for kernels with operators and no correlators, bins or observable histograms.  No counterpart in the
 * reference (its measurement follows its own sweeps, kernel.c:478-506). */
int fs_lp_sweep_measure(fs_lp *p, void *paths, void *accepts, const fs_lp_buffers *b, const fs_buffers *kb,
                        uint64_t sweep0, int32_t n_sweeps, int64_t series_offset, void *stream);
/* One Adam step (step counts from 1) on `batch` rows drawn from paths, then a repack. */
int fs_lp_train(fs_lp *p, const void *paths, const fs_lp_buffers *b, uint64_t step, void *stream);
/* The network alone: X [rows][32] FP32 features -> O [rows][32] FP32 (logical outputs 0 .. 2 DOF - 1). */
int fs_lp_eval(fs_lp *p, const fs_lp_buffers *b, const void *X, void *O, int64_t rows, void *stream);
/* Weight gradients of the last fs_lp_train in logical parameter order, FP32 [n_params]
 * (a device buffer); test support. */
int fs_lp_grad(fs_lp *p, const fs_lp_buffers *b, void *grad, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FEYNSIMUL_H */
