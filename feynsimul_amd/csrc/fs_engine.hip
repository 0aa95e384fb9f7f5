// fs_engine.hip -- host side of libfeynsimul.so: the C ABI of include/feynsimul.h.
//
// fs_create turns a configuration plus the system section into one hiprtc
// module (the sampler template fs_sampler.h is baked into this library), the
// role cl.Program(...).build plays in the reference (FeynSimul/kernel.py:346-382).
// Launch geometry follows the reference for the exact sampler
// (kernel.py:192-211) and is the wave sampler's own otherwise.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/feynsimul.h"
#include "device/fs_abi_args.h"
#include "fs_rng_kernels.hip"

namespace {

const char kRngHeader[] = {
#embed "device/fs_rng.h"
    , 0};
const char kArgsHeader[] = {
#embed "device/fs_abi_args.h"
    , 0};
const char kSamplerHeader[] = {
#embed "device/fs_sampler.h"
    , 0};
const char kWaveHeader[] = {
#embed "device/fs_wave.h"
    , 0};
const char kStreamHeader[] = {
#embed "device/fs_stream.h"
    , 0};
const char kLaneHeader[] = {
#embed "device/fs_lane.h"
    , 0};

thread_local std::string g_error;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define FS_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(FS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

std::string lit(double v)
{
    char b[64];
    snprintf(b, sizeof(b), "%1.17e", v);
    return b;
}

int copy_out(const std::string &s, char *buf, size_t len, size_t *needed)
{
    if (needed) *needed = s.size() + 1;
    if (buf && len) {
        size_t n = s.size() < len - 1 ? s.size() : len - 1;
        memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return FS_OK;
}

}  // namespace

struct fs_engine {
    fs_config cfg{};
    std::string source;
    std::string log;
    std::vector<char> code;
    hipModule_t module = nullptr;
    hipFunction_t fn_exact = nullptr;
    hipFunction_t fn_wave = nullptr;
    hipFunction_t fn_eval = nullptr;
    int wpb = 4;
    size_t wave_smem = 0;
    int device = -1;
    int lds_per_cu = 0;        // bytes of LDS per CU (device attribute)
    int waves_per_simd = 0;    // occupancy cap of wave launches (0 = none)
};

namespace {

int validate(const fs_config *c)
{
    if (!c) return fail(FS_ERR_INVALID, "config is NULL");
    if (c->abi_version != FS_ABI_VERSION)
        return fail(FS_ERR_INVALID, "ABI version %d, library has %d", c->abi_version, FS_ABI_VERSION);
    if (c->N <= 0 || (c->N & (c->N - 1)))
        return fail(FS_ERR_INVALID, "Path node number, N, must be a power of 2.");
    if (c->S < 1 || (1 << c->S) > c->N) return fail(FS_ERR_INVALID, "2^S must be less than or equal to N.");
    if (c->DOF < 1 || c->DOF > 64) return fail(FS_ERR_INVALID, "DOF must be in [1, 64]");
    if (c->n_walkers < 1) return fail(FS_ERR_INVALID, "nbrOfWalkers must be positive");
    if (c->walkers_per_group < 1 || c->n_walkers % c->walkers_per_group)
        return fail(FS_ERR_INVALID, "The number of walkers must be a multiple of the number of walkers per work group.");
    if (c->n_ops < 0 || c->n_corr < 0) return fail(FS_ERR_INVALID, "negative operator count");
    if (c->ranlux) {
        const int l = c->ranlux_lux;
        if (l < 0 || (l > 4 && l < 24)) return fail(FS_ERR_INVALID, "RANLUX luxury value %d gives a negative skip", l);
        if (l >= 24 && ((l - 24) % 4 != 0 || (l - 24 > 0 && l - 24 < 24)))
            return fail(FS_ERR_INVALID, "RANLUX skip %d must be 0 or a multiple of 4 that is >= 24", l - 24);
    }
    if (c->sampler == FS_SAMPLER_EXACT) {
        if (c->parallelize_path) {
            const long long T = c->N >> c->S;
            if (T * c->walkers_per_group > 1024)
                return fail(FS_ERR_INVALID, "work-group of %lld items exceeds 1024", T * c->walkers_per_group);
        } else if (c->local_size < 1) {
            return fail(FS_ERR_INVALID, "local_size must be positive");
        }
    } else if (c->sampler == FS_SAMPLER_WAVE) {
        if ((1 << c->S) > 64) return fail(FS_ERR_INVALID, "wave sampler needs 2^S <= 64");
        if (c->n_corr > 0 && (c->cor_lags < 1 || c->cor_lags > c->N / 2))
            return fail(FS_ERR_INVALID, "cor_lags must be in [1, N/2]");
        if (c->n_hist_obs < 0 || (c->n_hist_obs > 0 && (c->hist_bins < 1 || c->hist_bins > 8192 ||
                                                      !(c->hist_hi > c->hist_lo))))
            return fail(FS_ERR_INVALID, "observable histogram needs 1 <= hist_bins <= 8192 and hist_hi > hist_lo");
        if (c->potential_cache < FS_POTCACHE_AUTO || c->potential_cache > FS_POTCACHE_OFF)
            return fail(FS_ERR_INVALID, "unknown potential_cache %d", c->potential_cache);
        if (c->max_waves_per_simd < 0 || c->max_waves_per_simd > 16)
            return fail(FS_ERR_INVALID, "max_waves_per_simd must be in [0, 16]");
        if (c->path_storage < FS_STORAGE_AUTO || c->path_storage > FS_STORAGE_REGISTERS)
            return fail(FS_ERR_INVALID, "unknown path_storage %d", c->path_storage);
        if (c->path_storage == FS_STORAGE_REGISTERS && (c->N < 64 || c->N / 64 > 16))
            return fail(FS_ERR_INVALID, "register storage needs 64 <= N <= 1024");
    } else {
        return fail(FS_ERR_INVALID, "unknown sampler %d", c->sampler);
    }
    if (c->bins && (c->bin_res < 1 || !(c->x_max > c->x_min)))
        return fail(FS_ERR_INVALID, "histogram needs bin_res >= 1 and x_max > x_min");
    if (c->path_shift || c->single_node) {
        // kernel.py:193-198: these moves exist only with one work-item per walker
        if (c->sampler != FS_SAMPLER_EXACT || c->parallelize_path)
            return fail(FS_ERR_INVALID, "path-shift and single-node moves need the exact sampler without "
                                        "parallelize_path");
    }
    if (c->no_bisection && c->sampler == FS_SAMPLER_WAVE)
        return fail(FS_ERR_INVALID, "the wave sampler's move is the bisection");
    return FS_OK;
}

// Register storage of the wave sampler: lane j keeps N/64 slices of DOF
// coordinates plus their potential for the whole launch.  Chosen automatically
// only while that costs at most 64 VGPRs, which keeps several wavefronts per
// SIMD; larger paths are streamed through the cache hierarchy instead.
bool resident_storage(const fs_config *c)
{
    if (c->sampler != FS_SAMPLER_WAVE || c->N < 64 || c->N / 64 > 16) return false;
    if (c->path_storage == FS_STORAGE_REGISTERS) return true;
    if (c->path_storage == FS_STORAGE_MEMORY) return false;
    const int words = c->double_precision ? 2 : 1;
    return (c->N / 64) * (c->DOF + 1) * words <= 64;
}

// Memory storage with the per-slice potential cache (see FS_POTCACHE_*).
bool potential_cached(const fs_config *c)
{
    if (resident_storage(c)) return false;
    return c->potential_cache == FS_POTCACHE_ON || (c->potential_cache == FS_POTCACHE_AUTO && c->DOF > 1);
}

int nskip_of(const fs_config *c)
{
    const int l = c->ranlux_lux;
    return c->ranlux ? (l == 0 ? 0 : l == 1 ? 24 : l == 2 ? 76 : l == 3 ? 200 : l == 4 ? 380 : l - 24) : 0;
}

// Definitions shared by every module: precision, random-stream kind, the
// OpenCL C conveniences the reference's system strings may use
// (physical_systems/__init__.py:6-9) and the fast sin(2 pi x) helpers.
std::string common_prelude(const fs_config *c)
{
    std::string s;
    auto def = [&](const char *k, const std::string &v) { s += "#define " + std::string(k) + " " + v + "\n"; };
    def("FLOAT_TYPE", c->double_precision ? "double" : "float");
    def("FS_DOUBLE", c->double_precision ? "1" : "0");
    def("FS_RANLUX", c->ranlux ? "1" : "0");
    def("FS_NSKIP", std::to_string(nskip_of(c)));
    char b[64];
    snprintf(b, sizeof(b), "%.9Ef", (double)c->ranlux_int_range);  // as kernel.py:334 formats it
    def("FS_ISCALE", b);
    s += R"(
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_PI_2
#define M_PI_2 1.57079632679489661923
#endif
#define native_exp exp
#define native_log log
#define native_sqrt sqrt
#define native_sin sin
#define native_cos cos
#define native_powr pow
#define native_recip(a) (1.0 / (a))
#define native_divide(a, b) ((a) / (b))
#define half_exp exp
#define half_log log
#define half_sqrt sqrt
#define mad(a, b, c) ((a) * (b) + (c))
#define pown(a, n) pow((a), (FLOAT_TYPE)(n))
)";
    // sin(2 pi x) and cos(2 pi x): the hardware sine/cosine take their argument
    // in revolutions; used when relaxed math is allowed in FP32
    if (c->fast_math && !c->double_precision)
        s += "__device__ __forceinline__ float fs_sin2pi(float x) { return __builtin_amdgcn_sinf(x); }\n"
             "__device__ __forceinline__ float fs_cos2pi(float x) { return __builtin_amdgcn_cosf(x); }\n";
    else
        s += "__device__ __forceinline__ FLOAT_TYPE fs_sin2pi(FLOAT_TYPE x) { return sin(2.0 * M_PI * x); }\n"
             "__device__ __forceinline__ FLOAT_TYPE fs_cos2pi(FLOAT_TYPE x) { return cos(2.0 * M_PI * x); }\n";
    return s;
}

// Trait class of one system for fs_wave_walker (see fs_wave.h).
std::string wave_trait(const fs_config *c)
{
    const bool res = resident_storage(c);
    std::string s = "struct sys {\n    static constexpr int N = " + std::to_string(c->N) +
                    ", S = " + std::to_string(c->S) + ", DOF = " + std::to_string(c->DOF) +
                    ", NOPS = " + std::to_string(c->n_ops) + ", NCORR = " + std::to_string(c->n_corr) +
                    ", CORLAGS = " + std::to_string(c->cor_lags > 0 ? c->cor_lags : c->N / 2) +
                    ", NHOBS = " + std::to_string(c->n_hist_obs) +
                    ", HBINS = " + std::to_string(c->hist_bins > 0 ? c->hist_bins : 1) +
                    ", RESIDENT = " + std::to_string(res ? 1 : 0) +
                    ", POTCACHE = " + std::to_string(potential_cached(c) || res ? 1 : 0) +
                    ", BINS = " + std::to_string(c->bins ? 1 : 0) +
                    ", BINRES = " + std::to_string(c->bin_res > 0 ? c->bin_res : 1) + ";\n";
    s += "    static constexpr FLOAT_TYPE EPS = " + lit(c->beta / (double)c->N) + ", HLO = " + lit(c->hist_lo) +
         ", HINV = " + lit(c->n_hist_obs > 0 ? (double)c->hist_bins / (c->hist_hi - c->hist_lo) : 1.0) +
         ", XMIN = " + lit(c->x_min) + ", XMAX = " + lit(c->x_max) +
         ", INVBIN = " + lit(c->bins ? (double)c->bin_res / (c->x_max - c->x_min) : 1.0) + ";\n";
    s += "    static FS_DEV FLOAT_TYPE V(const FLOAT_TYPE *x) { return fs_potential_v(x); }\n"
         "    static FS_DEV void ops(const FLOAT_TYPE *x, FLOAT_TYPE *acc) { fs_operator_v(x, acc); }\n"
         "    static FS_DEV void corr(const FLOAT_TYPE *x, FLOAT_TYPE *prod) { fs_correlator_v(x, prod); }\n"
         "    static FS_DEV void hobs(const FLOAT_TYPE *x, FLOAT_TYPE *out) { fs_hist_obs_v(x, out); }\n};\n";
    return s;
}

size_t wave_lds_bytes(const fs_config *c)
{
    // the same figure as fs_wave_lds_bytes<Sys>() in device/fs_wave.h
    const int fsize = c->double_precision ? 8 : 4;
    const int lags = c->cor_lags > 0 ? c->cor_lags : c->N / 2;
    const int cor = c->n_corr > 0 ? fs_cor_stage_len(c->N, lags) * fsize : 0;
    const int hist = c->n_hist_obs > 0 ? c->hist_bins * 4 : 0;
    return (size_t)fs_wave_meas_total(cor, hist, c->n_ops) + (c->ranlux ? sizeof(fs_ranlux) : 0);
}

std::vector<std::string> compile_options(const fs_config *c, const std::string &arch)
{
    std::vector<std::string> opts = {"--offload-arch=" + arch, "-O3", "-std=c++17", "-Xclang", "-target-feature",
                                     "-Xclang", "-packed-fp32-ops", "-fno-math-errno", "-ffp-contract=fast"};
    if (!c->double_precision) {
        opts.push_back("-Xclang");
        opts.push_back("-cl-single-precision-constant");
    }
    // FS_MIN_WAVES=n: register budget of the wave sampler for n wavefronts per SIMD
    if (const char *mw = getenv("FS_MIN_WAVES"); mw && atoi(mw) > 0)
        opts.push_back("-DFS_MIN_WAVES=" + std::to_string(atoi(mw)));
    // FS_RESOURCE_USAGE=1: registers, spills, LDS and occupancy of every kernel in the build log
    if (const char *ru = getenv("FS_RESOURCE_USAGE"); ru && ru[0] == '1')
        opts.push_back("-Rpass-analysis=kernel-resource-usage");
    if (c->fast_math) {
        opts.push_back("-freciprocal-math");
        opts.push_back("-fno-signed-zeros");
        opts.push_back("-fno-trapping-math");
        opts.push_back("-fapprox-func");
        // FP32 subnormals flush to zero, so exp and log map to one v_exp_f32 /
        // v_log_f32 each instead of a range-scaled sequence of six instructions
        // (below 2^-126 no result of a move, a potential or an estimator is
        // distinguishable from zero)
        opts.push_back("-fgpu-flush-denormals-to-zero");
    }
    return opts;
}

std::string build_source(const fs_config *c, const char *system_src, int wpb)
{
    std::string s;
    auto def = [&](const char *k, const std::string &v) { s += "#define " + std::string(k) + " " + v + "\n"; };
    auto defi = [&](const char *k, long long v) { def(k, std::to_string(v)); };
    s += "// generated by fs_create\n";
    s += common_prelude(c);
    defi("FS_N", c->N);
    defi("FS_S", c->S);
    defi("FS_DOF", c->DOF);
    defi("FS_NOPS", c->n_ops);
    defi("FS_NCORR", c->n_corr);
    defi("FS_CORLAGS", c->cor_lags > 0 ? c->cor_lags : c->N / 2);
    defi("FS_BINS", c->bins ? 1 : 0);
    defi("FS_BINRES", c->bin_res > 0 ? c->bin_res : 1);
    def("FS_XMIN", lit(c->x_min));
    def("FS_XMAX", lit(c->x_max));
    def("FS_INVBIN", lit(c->bins ? (double)c->bin_res / (c->x_max - c->x_min) : 1.0));
    def("FS_OPNORM", lit(c->op_norm));
    defi("FS_PARALLEL", c->parallelize_path ? 1 : 0);
    defi("FS_W", c->walkers_per_group);
    defi("FS_LOCAL", c->local_size > 0 ? c->local_size : 1);
    defi("FS_WPB", wpb);
    defi("FS_SAMPLER", c->sampler);
    defi("FS_NHOBS", c->sampler == FS_SAMPLER_WAVE ? c->n_hist_obs : 0);
    defi("FS_HBINS", c->hist_bins > 0 ? c->hist_bins : 1);
    def("FS_HLO", lit(c->hist_lo));
    def("FS_HINV", lit(c->n_hist_obs > 0 ? (double)c->hist_bins / (c->hist_hi - c->hist_lo) : 1.0));
    defi("FS_RESIDENT", resident_storage(c) ? 1 : 0);
    defi("FS_POTCACHE", potential_cached(c) ? 1 : 0);
    defi("FS_BISECTION", c->no_bisection ? 0 : 1);
    defi("FS_PATH_SHIFT", c->path_shift ? 1 : 0);
    defi("FS_SINGLE_NODE", c->single_node ? 1 : 0);
    def("FS_ALPHA", lit(c->alpha));
    def("FS_PSALPHA", lit(c->ps_alpha));
    def("FS_EPS_INV2", lit(((double)c->N / c->beta) * ((double)c->N / c->beta)));  // kernel.py:299
    def("eps", lit(c->beta / (double)c->N));
    s += "#pragma clang force_cuda_host_device begin\n";
    s += system_src ? system_src : "";
    s += "\n#pragma clang force_cuda_host_device end\n#include \"fs_sampler.h\"\n";
    return s;
}

std::string device_arch()
{
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess) {
        std::string a = p.gcnArchName;
        const size_t colon = a.find(':');
        if (colon != std::string::npos) a = a.substr(0, colon);
        if (!a.empty()) return a;
    }
    return "gfx950";
}

}  // namespace

extern "C" {

int fs_abi_version(void) { return FS_ABI_VERSION; }

const char *fs_last_error(void) { return g_error.c_str(); }

int fs_device_count(int *count)
{
    if (!count) return fail(FS_ERR_INVALID, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return FS_OK;
}

int fs_create(const fs_config *cfg, const char *system_src, int compile_only, fs_engine **out)
{
    if (!out) return fail(FS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
    if (!compile_only && ndev == 0) return fail(FS_ERR_NO_DEVICE, "no GPU visible");

    fs_engine *e = new fs_engine();
    e->cfg = *cfg;
    int wpb = cfg->waves_per_block > 0 ? cfg->waves_per_block : 4;
    if (cfg->sampler == FS_SAMPLER_WAVE) {
        // per wavefront: correlator staging, then a private observable histogram
        const size_t per = wave_lds_bytes(cfg);
        while (wpb > 1 && per * wpb > 65536) wpb--;
        if (per > 65536) {
            delete e;
            return fail(FS_ERR_INVALID, "measurement staging needs %zu bytes of LDS per wavefront", per);
        }
        e->wave_smem = per * wpb;
    }
    e->wpb = wpb;
    e->source = build_source(cfg, system_src, wpb);

    hiprtcProgram prog;
    const char *hdrs[] = {kRngHeader, kArgsHeader, kSamplerHeader, kWaveHeader, kStreamHeader, kLaneHeader};
    const char *names[] = {"fs_rng.h", "fs_abi_args.h", "fs_sampler.h", "fs_wave.h", "fs_stream.h", "fs_lane.h"};
    if (hiprtcCreateProgram(&prog, e->source.c_str(), "fs_module.hip", 6, hdrs, names) != HIPRTC_SUCCESS) {
        delete e;
        return fail(FS_ERR_COMPILE, "hiprtcCreateProgram failed");
    }
    std::vector<std::string> opts = compile_options(cfg, ndev ? device_arch() : std::string("gfx950"));
    std::vector<const char *> copts;
    for (auto &o : opts) copts.push_back(o.c_str());
    const hiprtcResult cr = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    if (ls > 1) {
        std::vector<char> log(ls + 1, 0);
        hiprtcGetProgramLog(prog, log.data());
        e->log = log.data();
    }
    if (cr != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        g_error = "hiprtc compilation failed:\n" + e->log;
        *out = e;  // keep the engine so the caller can read the build log
        return FS_ERR_COMPILE;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    e->code.resize(cs);
    hiprtcGetCode(prog, e->code.data());
    hiprtcDestroyProgram(&prog);

    if (!compile_only) {
        (void)hipGetDevice(&e->device);
        hipError_t le = hipModuleLoadData(&e->module, e->code.data());
        if (le != hipSuccess) {
            delete e;
            return fail(FS_ERR_HIP, "hipModuleLoadData failed: %s", hipGetErrorString(le));
        }
        const char *kname = cfg->sampler == FS_SAMPLER_WAVE ? "fs_wave_sweep" : "fs_exact_sweep";
        hipFunction_t fn = nullptr;
        hipError_t fe = hipModuleGetFunction(&fn, e->module, kname);
        if (fe != hipSuccess) {
            (void)hipModuleUnload(e->module);
            delete e;
            return fail(FS_ERR_HIP, "module has no %s: %s", kname, hipGetErrorString(fe));
        }
        (cfg->sampler == FS_SAMPLER_WAVE ? e->fn_wave : e->fn_exact) = fn;
        if (hipDeviceGetAttribute(&e->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, e->device) !=
            hipSuccess)
            e->lds_per_cu = 0;
        e->waves_per_simd = cfg->max_waves_per_simd;
        (void)hipModuleGetFunction(&e->fn_eval, e->module, "fs_eval_system");
        (void)hipGetLastError();
    }
    *out = e;
    return FS_OK;
}

int fs_destroy(fs_engine *e)
{
    if (!e) return FS_OK;
    if (e->module) (void)hipModuleUnload(e->module);
    delete e;
    return FS_OK;
}

int fs_storage_of(const fs_config *cfg, int *storage)
{
    if (!cfg || !storage) return fail(FS_ERR_INVALID, "config or storage is NULL");
    if (int rc = validate(cfg)) return rc;
    *storage = cfg->sampler == FS_SAMPLER_WAVE && resident_storage(cfg) ? FS_STORAGE_REGISTERS : FS_STORAGE_MEMORY;
    return FS_OK;
}

int fs_path_storage(const fs_engine *e, int *storage)
{
    if (!e || !storage) return fail(FS_ERR_INVALID, "engine or storage is NULL");
    *storage = e->cfg.sampler == FS_SAMPLER_WAVE
                   ? (resident_storage(&e->cfg) ? FS_STORAGE_REGISTERS : FS_STORAGE_MEMORY)
                   : FS_STORAGE_MEMORY;
    return FS_OK;
}

int fs_build_log(const fs_engine *e, char *buf, size_t len, size_t *needed)
{
    if (!e) return fail(FS_ERR_INVALID, "engine is NULL");
    return copy_out(e->log, buf, len, needed);
}

int fs_source(const fs_engine *e, char *buf, size_t len, size_t *needed)
{
    if (!e) return fail(FS_ERR_INVALID, "engine is NULL");
    return copy_out(e->source, buf, len, needed);
}

int fs_run(fs_engine *e, const fs_buffers *b, int32_t operator_runs, int32_t metro_steps, int32_t series_offset,
           void *stream)
{
    return fs_run_flags(e, b, operator_runs, metro_steps, series_offset, 0, stream);
}

int fs_run_flags(fs_engine *e, const fs_buffers *b, int32_t operator_runs, int32_t metro_steps,
                 int32_t series_offset, int32_t flags, void *stream)
{
    if (flags & ~FS_RUN_POT_CACHE_VALID) return fail(FS_ERR_INVALID, "unknown run flags 0x%x", flags);
    if (!e || !b) return fail(FS_ERR_INVALID, "engine or buffers is NULL");
    if (!e->module) return fail(FS_ERR_INVALID, "engine was created compile-only");
    if (operator_runs < 0 || metro_steps < 0) return fail(FS_ERR_INVALID, "negative run length");
    const fs_config &c = e->cfg;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (!b->paths || !b->accepts) return fail(FS_ERR_INVALID, "paths and accepts are required");
    if (c.ranlux ? (!b->rl || !b->grl) : !b->seeds) return fail(FS_ERR_INVALID, "random state buffers missing");
    if (c.n_ops > 0 && !b->op_means) return fail(FS_ERR_INVALID, "op_means required");
    if (c.n_corr > 0 && !b->cor_means) return fail(FS_ERR_INVALID, "cor_means required");
    if (c.bins && !b->bins) return fail(FS_ERR_INVALID, "bins required");
    if (c.sampler == FS_SAMPLER_WAVE && c.n_hist_obs > 0 && !b->obs_hist)
        return fail(FS_ERR_INVALID, "obs_hist required");
    if (c.sampler == FS_SAMPLER_WAVE && !resident_storage(&c) && !b->pot_cache)
        return fail(FS_ERR_INVALID, "pot_cache required by the wave sampler with memory storage");

    fs_kargs a{};
    a.paths = b->paths;
    a.accepts = static_cast<unsigned *>(b->accepts);
    a.seeds = static_cast<unsigned *>(b->seeds);
    a.seedsSnap = static_cast<unsigned *>(b->seeds_snap);
    a.rl = static_cast<fs_ranlux *>(b->rl);
    a.grl = static_cast<fs_ranlux *>(b->grl);
    a.oldPaths = b->old_paths;
    a.opMeans = b->op_means;
    a.opTotals = static_cast<double *>(b->op_totals);
    a.series = static_cast<float *>(b->series);
    a.corMeans = b->cor_means;
    a.bins = static_cast<unsigned *>(b->bins);
    a.potCache = b->pot_cache;
    a.obsHist = static_cast<unsigned *>(b->obs_hist);
    a.nWalkers = c.n_walkers;
    a.operatorRuns = operator_runs;
    a.metroSteps = metro_steps;
    a.seriesOffset = series_offset;
    a.flags = (flags & FS_RUN_POT_CACHE_VALID) ? FS_KF_POT_VALID : 0;
    void *args[] = {&a};

    if (c.sampler == FS_SAMPLER_EXACT) {
        if (!b->old_paths) return fail(FS_ERR_INVALID, "old_paths required by the exact sampler");
        if (!c.ranlux) {
            if (!b->seeds_snap) return fail(FS_ERR_INVALID, "seeds_snap required by the exact sampler");
            const long long threads = c.parallelize_path ? c.n_walkers * (long long)(c.N >> c.S) : c.n_walkers;
            FS_HIP(hipMemcpyAsync(b->seeds_snap, b->seeds, (size_t)(threads + 1) * 16, hipMemcpyDeviceToDevice, st));
        }
        unsigned grid, block;
        if (c.parallelize_path) {
            block = (unsigned)((c.N >> c.S) * c.walkers_per_group);
            grid = (unsigned)(c.n_walkers / c.walkers_per_group);
        } else {
            block = 64;
            grid = (unsigned)((c.n_walkers + 63) / 64);
        }
        FS_HIP(hipModuleLaunchKernel(e->fn_exact, grid, 1, 1, block, 1, 1, 0, st, args, nullptr));
    } else {
        if (!e->fn_wave) return fail(FS_ERR_INVALID, "module has no wave sampler");
        const unsigned grid = (unsigned)((c.n_walkers + e->wpb - 1) / e->wpb);
        size_t smem = e->wave_smem;
        if (e->waves_per_simd > 0 && e->lds_per_cu > 0) {
            // at most B blocks of wpb wavefronts per CU (4 SIMDs): request more than 1/(B+1) of its LDS
            const int B = std::max(2, e->waves_per_simd * 4 / e->wpb);
            const size_t pad = ((size_t)e->lds_per_cu / (size_t)(B + 1) + 16) & ~(size_t)15;
            if (pad <= 65536) smem = std::max(smem, pad);
        }
        FS_HIP(hipModuleLaunchKernel(e->fn_wave, grid, 1, 1, 64 * e->wpb, 1, 1, (unsigned)smem, st, args, nullptr));
    }
    return FS_OK;
}

int fs_launch_info(const fs_engine *e, fs_launch_info_t *info)
{
    if (!e || !info) return fail(FS_ERR_INVALID, "engine and info are required");
    if (!e->module) return fail(FS_ERR_INVALID, "engine was created compile-only");
    const fs_config &c = e->cfg;
    hipFunction_t fn = c.sampler == FS_SAMPLER_EXACT ? e->fn_exact : e->fn_wave;
    fs_launch_info_t r{};
    int v = 0;
    FS_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, fn));
    r.max_threads_per_block = v;
    FS_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, fn));
    r.lds_bytes = v;
    FS_HIP(hipFuncGetAttribute(&v, HIP_FUNC_ATTRIBUTE_NUM_REGS, fn));
    r.registers = v;
    if (c.sampler == FS_SAMPLER_EXACT) {
        r.threads_per_block = c.parallelize_path ? (c.N >> c.S) * c.walkers_per_group : 64;
    } else {
        r.threads_per_block = 64 * e->wpb;
        r.lds_bytes += (int32_t)e->wave_smem;
    }
    int dev = e->device >= 0 ? e->device : 0;
    FS_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    r.lds_limit = v;
    FS_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxThreadsPerBlock, dev));
    r.device_max_threads_per_block = v;
    FS_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxBlockDimX, dev));
    r.device_max_block_dim[0] = v;
    FS_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxBlockDimY, dev));
    r.device_max_block_dim[1] = v;
    FS_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxBlockDimZ, dev));
    r.device_max_block_dim[2] = v;
    *info = r;
    return FS_OK;
}

int fs_set_waves_per_simd(fs_engine *e, int32_t cap)
{
    if (!e) return fail(FS_ERR_INVALID, "engine is NULL");
    if (cap < 0 || cap > 16) return fail(FS_ERR_INVALID, "waves per SIMD must be in [0, 16]");
    e->waves_per_simd = cap;
    return FS_OK;
}

int fs_eval_system(fs_engine *e, const void *x, void *pot, void *ops, int64_t n, void *stream)
{
    if (!e || !x || !pot) return fail(FS_ERR_INVALID, "engine, x and pot are required");
    if (!e->module) return fail(FS_ERR_INVALID, "engine was created compile-only");
    if (e->cfg.n_ops > 0 && !ops) return fail(FS_ERR_INVALID, "ops required");
    if (n <= 0) return FS_OK;
    long long nn = n;
    void *args[] = {&x, &pot, &ops, &nn};
    const unsigned grid = (unsigned)((n + 255) / 256);
    FS_HIP(hipModuleLaunchKernel(e->fn_eval, grid, 1, 1, 256, 1, 1, 0, reinterpret_cast<hipStream_t>(stream), args,
                                 nullptr));
    return FS_OK;
}

// ---- batched systems ----
}  // extern "C"

// Device copy of one dispatch, in one allocation: the work counters (one per
// sweep kernel, in the first 256 bytes), the per-type kernel arguments, and
// per sweep kernel the prefix sums of its dispatch slots and the type of each
// slot.  A batch keeps one for plain launches; every captured graph owns its
// own, so later changes of buffers or order never reach a graph.
struct fs_dispatch_layout {
    int nt, nk;
    constexpr fs_dispatch_layout(int types, int kernels) : nt(types), nk(kernels) {}
    static constexpr size_t up(size_t v) { return (v + 255) & ~(size_t)255; }
    constexpr size_t types() const { return 256; }
    constexpr size_t per_kernel() const { return up(sizeof(long long) * (size_t)(nt + 1) + sizeof(int) * (size_t)nt); }
    constexpr size_t start(int j) const { return up(256 + sizeof(fs_kargs) * (size_t)nt) + per_kernel() * (size_t)j; }
    constexpr size_t type_of(int j) const { return start(j) + sizeof(long long) * (size_t)(nt + 1); }
    constexpr size_t bytes() const { return start(nk); }
};

// Dispatch tables of fs_batch_run_noalloc: one fixed array per GPU, held
// until the process ends, shared by every call on that GPU.  Host code uses
// the device address that hipGetSymbolAddress returns for the current
// device (scratch_address), never the array's name.
constexpr int kBatchMaxTypes = 64;
constexpr int kBatchMaxKernels = 4;
constexpr size_t kScratchBytes = 16384;
static_assert(fs_dispatch_layout(kBatchMaxTypes, kBatchMaxKernels).bytes() <= kScratchBytes,
              "scratch too small for 64 types in 4 sweep kernels");
static __device__ __align__(256) char scratch[kScratchBytes];

struct fs_batch {
    std::vector<fs_config> cfgs;
    std::string source;
    std::string log;
    std::vector<char> code;
    hipModule_t module = nullptr;
    // one persistent sweep kernel per register class of types (kernel_of_type)
    int nk = 1;
    std::vector<int> kernel_of_type;
    std::vector<int> types_in_kernel;
    std::vector<hipFunction_t> fn_sweep;
    std::vector<int> max_blocks;
    hipFunction_t fn_reset = nullptr;
    int wpb = 4;
    size_t smem = 0;
    int device = -1;
    // dispatch tables of plain launches and the host image last uploaded
    char *d_tables = nullptr;
    std::vector<char> h_tables;
    // register classes after the first run on streams of their own (default priority), forked from and joined
    // back to the launch stream with events, so one class's blocks fill the tail of another's
    std::vector<hipStream_t> side;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
};

struct fs_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    char *d_tables = nullptr;  // this graph's own dispatch tables
};

namespace {

std::string batch_source(const std::vector<fs_config> &cfgs, const std::vector<int> &kernel_of_type, int nk,
                         const std::vector<int> &waves_of_kernel, const char *const *srcs, int wpb, size_t lds)
{
    std::string s = "// generated by fs_batch_create\n";
    s += common_prelude(&cfgs[0]);
    s += "#define FS_WPB " + std::to_string(wpb) + "\n#define FS_LDS_PER_WAVE " + std::to_string(lds) + "\n";
    s += "#include \"fs_wave.h\"\n";
    for (size_t k = 0; k < cfgs.size(); k++) {
        s += "namespace fs_ty" + std::to_string(k) + " {\n#define eps " + lit(cfgs[k].beta / (double)cfgs[k].N) +
             "\n#pragma clang force_cuda_host_device begin\n";
        s += srcs[k] ? srcs[k] : "";
        s += "\n#pragma clang force_cuda_host_device end\n" + wave_trait(&cfgs[k]) + "#undef eps\n}\n";
    }
    s += R"(
extern "C" __global__ void fs_batch_reset(unsigned long long *counter, int n)
{
    if ((int)threadIdx.x < n) counter[threadIdx.x] = 0ull;
}

// Persistent sweep of one register class of types: each wavefront takes
// walker numbers from the class's counter and runs that walker's whole launch
// with its type's specialisation.
#define FS_BATCH_LOOP_BEGIN                                                   \
    const int lane = threadIdx.x & 63;                                        \
    extern __shared__ char fs_smem[];                                         \
    char *lds = fs_smem + (threadIdx.x >> 6) * FS_LDS_PER_WAVE;               \
    const long long total = b.start[b.ntypes];                                \
    for (;;) {                                                                \
        unsigned long long item = 0;                                          \
        if (lane == 0) item = atomicAdd(b.counter, 1ull);                     \
        item = __shfl(item, 0, 64);                                           \
        if ((long long)item >= total) break;                                  \
        int d = 0;                                                            \
        while (d + 1 < b.ntypes && (long long)item >= b.start[d + 1]) d++;    \
        const int t = b.typeOf[d];                                            \
        fs_kargs a = b.types[t];                                              \
        a.operatorRuns = b.operatorRuns;                                      \
        a.metroSteps = b.metroSteps;                                          \
        a.seriesOffset = b.seriesOffset;                                      \
        const long w = (long)((long long)item - b.start[d]);                  \
        switch (t) {
#define FS_BATCH_LOOP_END \
        default: break;   \
        }                 \
    }
)";
    for (int j = 0; j < nk; j++) {
        if (waves_of_kernel[j] > 0)
            s += "__attribute__((amdgpu_waves_per_eu(" + std::to_string(waves_of_kernel[j]) + ")))\n";
        s += "extern \"C\" __global__ __launch_bounds__(64 * FS_WPB) void fs_batch_sweep_" + std::to_string(j) +
             "(fs_batch_args b)\n{\n    FS_BATCH_LOOP_BEGIN\n";
        for (size_t k = 0; k < cfgs.size(); k++)
            if (kernel_of_type[k] == j)
                s += "        case " + std::to_string(k) + ": fs_wave_walker<fs_ty" + std::to_string(k) +
                     "::sys>(a, w, lane, lds); break;\n";
        s += "    FS_BATCH_LOOP_END\n}\n";
    }
    return s;
}

// Host image of the dispatch tables (fs_dispatch_layout, counter bytes zero).
int batch_tables(const fs_batch *b, const fs_buffers *bufs, const int32_t *order, int32_t operator_runs,
                 int32_t metro_steps, int32_t series_offset, std::vector<char> &image)
{
    const int nt = (int)b->cfgs.size();
    const fs_dispatch_layout L(nt, b->nk);
    image.assign(L.bytes(), 0);
    fs_kargs *types = reinterpret_cast<fs_kargs *>(image.data() + L.types());
    for (int t = 0; t < nt; t++) {
        const fs_config &c = b->cfgs[t];
        const fs_buffers &x = bufs[t];
        if (!x.paths || !x.accepts) return fail(FS_ERR_INVALID, "type %d: paths and accepts are required", t);
        if (c.ranlux ? (!x.rl || !x.grl) : !x.seeds) return fail(FS_ERR_INVALID, "type %d: random states missing", t);
        if (c.n_ops > 0 && !x.op_means) return fail(FS_ERR_INVALID, "type %d: op_means required", t);
        if (c.n_corr > 0 && !x.cor_means) return fail(FS_ERR_INVALID, "type %d: cor_means required", t);
        if (c.bins && !x.bins) return fail(FS_ERR_INVALID, "type %d: bins required", t);
        if (c.n_hist_obs > 0 && !x.obs_hist) return fail(FS_ERR_INVALID, "type %d: obs_hist required", t);
        if (!resident_storage(&c) && !x.pot_cache) return fail(FS_ERR_INVALID, "type %d: pot_cache required", t);
        fs_kargs a{};
        a.paths = x.paths;
        a.accepts = static_cast<unsigned *>(x.accepts);
        a.seeds = static_cast<unsigned *>(x.seeds);
        a.rl = static_cast<fs_ranlux *>(x.rl);
        a.grl = static_cast<fs_ranlux *>(x.grl);
        a.opMeans = x.op_means;
        a.opTotals = static_cast<double *>(x.op_totals);
        a.series = static_cast<float *>(x.series);
        a.corMeans = x.cor_means;
        a.bins = static_cast<unsigned *>(x.bins);
        a.potCache = x.pot_cache;
        a.obsHist = static_cast<unsigned *>(x.obs_hist);
        a.nWalkers = c.n_walkers;
        a.operatorRuns = operator_runs;
        a.metroSteps = metro_steps;
        a.seriesOffset = series_offset;
        memcpy(&types[t], &a, sizeof(a));
    }
    std::vector<char> seen(nt, 0);
    std::vector<int> dispatch(nt);
    for (int d = 0; d < nt; d++) {
        const int t = order ? order[d] : d;
        if (t < 0 || t >= nt || seen[t]) return fail(FS_ERR_INVALID, "order must be a permutation of the types");
        seen[t] = 1;
        dispatch[d] = t;
    }
    // each sweep kernel dispatches its own types in the given order
    for (int j = 0; j < b->nk; j++) {
        long long *start = reinterpret_cast<long long *>(image.data() + L.start(j));
        int *type_of = reinterpret_cast<int *>(image.data() + L.type_of(j));
        int n = 0;
        start[0] = 0;
        for (int d = 0; d < nt; d++) {
            const int t = dispatch[d];
            if (b->kernel_of_type[t] != j) continue;
            type_of[n] = t;
            start[n + 1] = start[n] + b->cfgs[t].n_walkers;
            n++;
        }
    }
    return FS_OK;
}

// Copies the tables (not the counter) of `image` to device memory `dst`,
// ordered on `st`.  The source is pageable host memory, which the runtime
// stages before the call returns.
int upload_tables(char *dst, const std::vector<char> &image, hipStream_t st)
{
    const size_t off = 256;
    FS_HIP(hipMemcpyAsync(dst + off, image.data() + off, image.size() - off, hipMemcpyHostToDevice, st));
    return FS_OK;
}

int batch_launch(fs_batch *b, char *tables, int32_t operator_runs, int32_t metro_steps, int32_t series_offset,
                 int32_t grid_blocks, hipStream_t st)
{
    const int nt = (int)b->cfgs.size();
    const fs_dispatch_layout L(nt, b->nk);
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(tables);
    int nk = b->nk;
    void *rargs[] = {&counter, &nk};
    FS_HIP(hipModuleLaunchKernel(b->fn_reset, 1, 1, 1, 64, 1, 1, 0, st, rargs, nullptr));
    const bool concurrent = !b->side.empty();
    if (concurrent) FS_HIP(hipEventRecord(b->ev_fork, st));
    for (int j = 0; j < b->nk; j++) {
        hipStream_t sj = concurrent && j > 0 ? b->side[j - 1] : st;
        if (sj != st) FS_HIP(hipStreamWaitEvent(sj, b->ev_fork, 0));
        fs_batch_args ba{};
        ba.types = reinterpret_cast<const fs_kargs *>(tables + L.types());
        ba.start = reinterpret_cast<const long long *>(tables + L.start(j));
        ba.typeOf = reinterpret_cast<const int *>(tables + L.type_of(j));
        ba.counter = counter + j;
        ba.ntypes = b->types_in_kernel[j];
        ba.operatorRuns = operator_runs;
        ba.metroSteps = metro_steps;
        ba.seriesOffset = series_offset;
        void *args[] = {&ba};
        const unsigned grid = (unsigned)(grid_blocks > 0 ? grid_blocks : b->max_blocks[j]);
        FS_HIP(hipModuleLaunchKernel(b->fn_sweep[j], grid, 1, 1, 64 * b->wpb, 1, 1, (unsigned)b->smem, sj, args,
                                     nullptr));
        if (sj != st) FS_HIP(hipEventRecord(b->ev_join[j - 1], sj));
    }
    if (concurrent)
        for (int j = 1; j < b->nk; j++) FS_HIP(hipStreamWaitEvent(st, b->ev_join[j - 1], 0));
    return FS_OK;
}

void batch_free(fs_batch *b)
{
    if (!b) return;
    for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
    for (hipEvent_t ev : b->ev_join) (void)hipEventDestroy(ev);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->d_tables) (void)hipFree(b->d_tables);
    if (b->module) (void)hipModuleUnload(b->module);
    delete b;
}

}  // namespace

extern "C" {

int fs_batch_create(const fs_config *cfgs, const char *const *system_srcs, int32_t ntypes, int32_t waves_per_block,
                    const int32_t *kernel_of_type, const int32_t *waves_of_kernel, int compile_only, fs_batch **out)
{
    if (!out || !cfgs || !system_srcs) return fail(FS_ERR_INVALID, "cfgs, system_srcs and out are required");
    *out = nullptr;
    if (ntypes < 1 || ntypes > kBatchMaxTypes) return fail(FS_ERR_INVALID, "ntypes must be in [1, %d]", kBatchMaxTypes);
    int nk = 1;
    std::vector<int> kot(ntypes, 0), per(kBatchMaxKernels, 0);
    if (kernel_of_type) {
        for (int t = 0; t < ntypes; t++) {
            if (kernel_of_type[t] < 0 || kernel_of_type[t] >= kBatchMaxKernels)
                return fail(FS_ERR_INVALID, "kernel_of_type[%d] must be in [0, %d)", t, kBatchMaxKernels);
            kot[t] = kernel_of_type[t];
            nk = std::max(nk, kot[t] + 1);
        }
    }
    for (int t = 0; t < ntypes; t++) per[kot[t]]++;
    for (int j = 0; j < nk; j++)
        if (!per[j]) return fail(FS_ERR_INVALID, "sweep kernel %d has no type (kernel_of_type must number its kernels "
                                                 "0, 1, ... without gaps)", j);
    std::vector<int> wok(nk, 0);
    if (waves_of_kernel)
        for (int j = 0; j < nk; j++) {
            if (waves_of_kernel[j] < 0 || waves_of_kernel[j] > 8)
                return fail(FS_ERR_INVALID, "waves_of_kernel[%d] must be in [0, 8]", j);
            wok[j] = waves_of_kernel[j];
        }
    for (int t = 0; t < ntypes; t++) {
        if (int rc = validate(&cfgs[t])) return rc;
        const fs_config &c = cfgs[t], &c0 = cfgs[0];
        if (c.sampler != FS_SAMPLER_WAVE) return fail(FS_ERR_INVALID, "type %d: batches need the wave sampler", t);
        if (c.double_precision != c0.double_precision || c.ranlux != c0.ranlux ||
            (c.ranlux && c.ranlux_lux != c0.ranlux_lux) || c.fast_math != c0.fast_math ||
            c.ranlux_int_range != c0.ranlux_int_range)
            return fail(FS_ERR_INVALID, "type %d: precision, RNG and math settings must match type 0", t);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
    if (!compile_only && ndev == 0) return fail(FS_ERR_NO_DEVICE, "no GPU visible");
    fs_batch *b = new fs_batch();
    b->cfgs.assign(cfgs, cfgs + ntypes);
    b->nk = nk;
    b->kernel_of_type = kot;
    b->types_in_kernel.assign(per.begin(), per.begin() + nk);
    size_t lds = 0;
    for (auto &c : b->cfgs) lds = std::max(lds, wave_lds_bytes(&c));
    lds = (lds + 15) & ~(size_t)15;
    int wpb = waves_per_block > 0 ? waves_per_block : 4;
    while (wpb > 1 && lds * wpb > 65536) wpb--;
    if (lds > 65536) {
        batch_free(b);
        return fail(FS_ERR_INVALID, "measurement staging needs %zu bytes of LDS per wavefront", lds);
    }
    b->wpb = wpb;
    b->smem = lds * wpb;
    b->source = batch_source(b->cfgs, b->kernel_of_type, nk, wok, system_srcs, wpb, lds);
    hiprtcProgram prog;
    const char *hdrs[] = {kRngHeader, kArgsHeader, kWaveHeader, kStreamHeader, kLaneHeader};
    const char *names[] = {"fs_rng.h", "fs_abi_args.h", "fs_wave.h", "fs_stream.h", "fs_lane.h"};
    if (hiprtcCreateProgram(&prog, b->source.c_str(), "fs_batch.hip", 5, hdrs, names) != HIPRTC_SUCCESS) {
        batch_free(b);
        return fail(FS_ERR_COMPILE, "hiprtcCreateProgram failed");
    }
    std::vector<std::string> opts = compile_options(&cfgs[0], ndev ? device_arch() : std::string("gfx950"));
    std::vector<const char *> copts;
    for (auto &o : opts) copts.push_back(o.c_str());
    const hiprtcResult cr = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    if (ls > 1) {
        std::vector<char> log(ls + 1, 0);
        hiprtcGetProgramLog(prog, log.data());
        b->log = log.data();
    }
    if (cr != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        g_error = "hiprtc compilation failed:\n" + b->log;
        *out = b;  // the caller can read the build log, then destroy
        return FS_ERR_COMPILE;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    b->code.resize(cs);
    hiprtcGetCode(prog, b->code.data());
    hiprtcDestroyProgram(&prog);
    if (!compile_only) {
        hipError_t e = hipModuleLoadData(&b->module, b->code.data());
        if (e == hipSuccess) e = hipModuleGetFunction(&b->fn_reset, b->module, "fs_batch_reset");
        if (e == hipSuccess) e = hipMalloc(&b->d_tables, fs_dispatch_layout(ntypes, nk).bytes());
        int dev = 0, cus = 0;
        if (e == hipSuccess) e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        b->fn_sweep.assign(nk, nullptr);
        b->max_blocks.assign(nk, 1);
        for (int j = 0; j < nk && e == hipSuccess; j++) {
            const std::string name = "fs_batch_sweep_" + std::to_string(j);
            e = hipModuleGetFunction(&b->fn_sweep[j], b->module, name.c_str());
            int per_cu = 0;
            if (e == hipSuccess)
                e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, b->fn_sweep[j], 64 * wpb, b->smem);
            b->max_blocks[j] = std::max(1, per_cu) * std::max(1, cus);
        }
        // FS_BATCH_SERIAL=1: the register classes run one after the other on the launch stream
        const char *serial = getenv("FS_BATCH_SERIAL");
        if (e == hipSuccess && nk > 1 && !(serial && serial[0] == '1')) {
            e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
            for (int j = 1; j < nk && e == hipSuccess; j++) {
                hipStream_t s = nullptr;
                hipEvent_t ev = nullptr;
                e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
                if (e == hipSuccess) b->side.push_back(s);
                if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
                if (e == hipSuccess) b->ev_join.push_back(ev);
            }
        }
        if (e != hipSuccess) {
            batch_free(b);
            return fail(FS_ERR_HIP, "batch module setup failed: %s", hipGetErrorString(e));
        }
        b->device = dev;
    }
    *out = b;
    return FS_OK;
}

int fs_batch_destroy(fs_batch *b)
{
    batch_free(b);
    return FS_OK;
}

int fs_batch_path_storage(const fs_batch *b, int32_t type, int *storage)
{
    if (!b || !storage || type < 0 || type >= (int)b->cfgs.size()) return fail(FS_ERR_INVALID, "bad type");
    *storage = resident_storage(&b->cfgs[type]) ? FS_STORAGE_REGISTERS : FS_STORAGE_MEMORY;
    return FS_OK;
}

int fs_batch_build_log(const fs_batch *b, char *buf, size_t len, size_t *needed)
{
    if (!b) return fail(FS_ERR_INVALID, "batch is NULL");
    return copy_out(b->log, buf, len, needed);
}

int fs_batch_run(fs_batch *b, const fs_buffers *bufs, const int32_t *order, int32_t operator_runs,
                 int32_t metro_steps, int32_t series_offset, int32_t grid_blocks, void *stream)
{
    if (!b || !bufs) return fail(FS_ERR_INVALID, "batch and buffers are required");
    if (!b->module) return fail(FS_ERR_INVALID, "batch was created compile-only");
    if (operator_runs < 0 || metro_steps < 0) return fail(FS_ERR_INVALID, "negative run length");
    std::vector<char> image;
    if (int rc = batch_tables(b, bufs, order, operator_runs, metro_steps, series_offset, image)) return rc;
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // the tables are tiny; an unchanged set needs no copy
    if (image != b->h_tables) {
        if (int rc = upload_tables(b->d_tables, image, st)) return rc;
        b->h_tables.swap(image);
    }
    return batch_launch(b, b->d_tables, operator_runs, metro_steps, series_offset, grid_blocks, st);
}

}  // extern "C"

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) on the first call.
int scratch_address(int device, char **out)
{
    int cur = -1;
    FS_HIP(hipGetDevice(&cur));
    if (cur != device)
        return fail(FS_ERR_INVALID, "the batch was created on device %d; make it current (device %d is)", device, cur);
    if (device < 0 || device >= kMaxDevices) return fail(FS_ERR_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(FS_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<char *>(g_scratch_ptr[device]);
    return FS_OK;
}

}  // namespace

extern "C" {

int fs_batch_run_noalloc(fs_batch *b, const fs_buffers *bufs, const int32_t *order, int32_t operator_runs,
                         int32_t metro_steps, int32_t series_offset, int32_t grid_blocks, void *stream)
{
    if (!b || !bufs) return fail(FS_ERR_INVALID, "batch and buffers are required");
    if (!b->module) return fail(FS_ERR_INVALID, "batch was created compile-only");
    if (operator_runs < 0 || metro_steps < 0) return fail(FS_ERR_INVALID, "negative run length");
    const int nt = (int)b->cfgs.size();
    if (nt > kBatchMaxTypes || b->nk > kBatchMaxKernels || fs_dispatch_layout(nt, b->nk).bytes() > kScratchBytes)
        return fail(FS_ERR_INVALID, "%d types in %d sweep kernels need %zu bytes of dispatch tables; the static "
                    "scratch holds %zu (at most %d types, %d kernels)", nt, b->nk,
                    fs_dispatch_layout(nt, b->nk).bytes(), kScratchBytes, kBatchMaxTypes, kBatchMaxKernels);
    std::vector<char> image;
    if (int rc = batch_tables(b, bufs, order, operator_runs, metro_steps, series_offset, image)) return rc;
    char *tables = nullptr;
    if (int rc = scratch_address(b->device, &tables)) return rc;
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    // every call uploads its tables: the last call on this GPU may have left another batch's
    if (int rc = upload_tables(tables, image, st)) return rc;
    // the counter is cleared by the reset kernel of batch_launch
    return batch_launch(b, tables, operator_runs, metro_steps, series_offset, grid_blocks, st);
}

int fs_batch_capture(fs_batch *b, const fs_buffers *bufs, const int32_t *order, int32_t operator_runs,
                     int32_t metro_steps, int32_t series_offset, int32_t n_launches, int32_t grid_blocks,
                     void *stream, fs_graph **out)
{
    if (!b || !bufs || !out) return fail(FS_ERR_INVALID, "batch, buffers and out are required");
    *out = nullptr;
    if (!b->module) return fail(FS_ERR_INVALID, "batch was created compile-only");
    if (!stream) return fail(FS_ERR_INVALID, "capture needs a non-default stream");
    if (n_launches < 1 || operator_runs < 0 || metro_steps < 0) return fail(FS_ERR_INVALID, "bad run length");
    std::vector<char> image;
    if (int rc = batch_tables(b, bufs, order, operator_runs, metro_steps, series_offset, image)) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    fs_graph *g = new fs_graph();
    auto drop = [&](int rc) {
        (void)fs_graph_destroy(g);
        return rc;
    };
    if (hipMalloc(&g->d_tables, image.size()) != hipSuccess) {
        g->d_tables = nullptr;
        return drop(fail(FS_ERR_HIP, "graph dispatch tables: out of device memory"));
    }
    if (int rc = upload_tables(g->d_tables, image, st)) return drop(rc);
    if (hipStreamSynchronize(st) != hipSuccess) return drop(fail(FS_ERR_HIP, "stream synchronisation failed"));
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return drop(fail(FS_ERR_HIP, "hipStreamBeginCapture failed"));
    int rc = FS_OK;
    for (int k = 0; k < n_launches && rc == FS_OK; k++)
        rc = batch_launch(b, g->d_tables, operator_runs, metro_steps, series_offset + k * operator_runs,
                          grid_blocks, st);
    const hipError_t ec = hipStreamEndCapture(st, &g->graph);
    if (rc != FS_OK) return drop(rc);
    if (ec != hipSuccess) return drop(fail(FS_ERR_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(ec)));
    const hipError_t ei = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        g->exec = nullptr;
        return drop(fail(FS_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei)));
    }
    *out = g;
    return FS_OK;
}

int fs_graph_launch(fs_graph *g, void *stream)
{
    if (!g || !g->exec) return fail(FS_ERR_INVALID, "graph is NULL");
    FS_HIP(hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream)));
    return FS_OK;
}

int fs_graph_destroy(fs_graph *g)
{
    if (!g) return FS_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_tables) (void)hipFree(g->d_tables);
    delete g;
    return FS_OK;
}

// ---- ahead-of-time random-stream kernels (fs_rng_kernels.hip) ----
static int need_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(FS_ERR_NO_DEVICE, "no GPU visible");
    return FS_OK;
}

int fs_seed_xorshift(void *rows, int64_t n, uint64_t seed, int64_t first, void *stream)
{
    if (!rows || n < 0) return fail(FS_ERR_INVALID, "bad rows");
    if (int rc = need_device()) return rc;
    if (n == 0) return FS_OK;
    hipLaunchKernelGGL(fs_k_seed_xorshift, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<uint4 *>(rows), (long long)n,
                       (unsigned long long)seed, (long long)first);
    FS_HIP(hipGetLastError());
    return FS_OK;
}

int fs_ranlux_init(void *tab, const void *seeds, int64_t n, void *stream)
{
    if (!tab || !seeds || n < 0) return fail(FS_ERR_INVALID, "bad ranlux table");
    if (int rc = need_device()) return rc;
    if (n == 0) return FS_OK;
    hipLaunchKernelGGL(fs_k_ranlux_init, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<fs_ranlux *>(tab),
                       static_cast<const unsigned long long *>(seeds), (long long)n);
    FS_HIP(hipGetLastError());
    return FS_OK;
}

int fs_xorshift_draw(const void *rows, void *out_u32, void *out_f32, void *out_f64, int64_t n, int32_t ndraw,
                     void *stream)
{
    if (!rows || n < 0 || ndraw < 0) return fail(FS_ERR_INVALID, "bad draw request");
    if (int rc = need_device()) return rc;
    if (n == 0 || ndraw == 0) return FS_OK;
    hipLaunchKernelGGL(fs_k_xs_draw, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<const uint4 *>(rows),
                       static_cast<unsigned *>(out_u32), static_cast<float *>(out_f32),
                       static_cast<double *>(out_f64), (long long)n, (int)ndraw);
    FS_HIP(hipGetLastError());
    return FS_OK;
}

int fs_ranlux_draw(void *tab, void *out_f32, void *out_f64, int64_t n, int32_t nquad, int32_t lux, void *stream)
{
    if (!tab || n < 0 || nquad < 0 || lux < 0 || lux > 4 || (!out_f32 && !out_f64))
        return fail(FS_ERR_INVALID, "bad ranlux draw request");
    if (int rc = need_device()) return rc;
    if (n == 0 || nquad == 0) return FS_OK;
    hipLaunchKernelGGL(fs_k_rl_draw, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                       reinterpret_cast<hipStream_t>(stream), static_cast<fs_ranlux *>(tab),
                       static_cast<float *>(out_f32), static_cast<double *>(out_f64), (long long)n, (int)nquad,
                       (int)lux);
    FS_HIP(hipGetLastError());
    return FS_OK;
}

int fs_hbm_probe(int32_t kind, const void *src, void *dst, int64_t n_vec16, int32_t blocks, void *stream)
{
    if (kind < 0 || kind > 2 || n_vec16 < 0 || blocks < 1 || (kind != 2 && !src) || !dst)
        return fail(FS_ERR_INVALID, "bad bandwidth probe request");
    if (int rc = need_device()) return rc;
    if (n_vec16 == 0) return FS_OK;
    hipLaunchKernelGGL(fs_k_hbm_probe, dim3((unsigned)blocks), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       static_cast<const float4 *>(src), static_cast<float4 *>(dst), (long long)n_vec16, (int)kind);
    FS_HIP(hipGetLastError());
    return FS_OK;
}

// Stream throughput: `count` uniforms per stream through the sampler's
// wrappers, XOR of their bits per stream in check[n]; states advance.
int fs_rng_bench(int32_t kind, int32_t lux, int32_t double_precision, void *states, void *check, int64_t n,
                 int64_t count, void *stream)
{
    if (!states || !check || n < 0 || count < 0 || (kind != FS_RNG_XORSHIFT && kind != FS_RNG_RANLUX) ||
        (kind == FS_RNG_RANLUX && (lux < 0 || lux > 4)))
        return fail(FS_ERR_INVALID, "bad stream benchmark request");
    if (int rc = need_device()) return rc;
    if (n == 0) return FS_OK;
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (kind == FS_RNG_XORSHIFT)
        hipLaunchKernelGGL(fs_k_xs_bench, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           static_cast<uint4 *>(states), static_cast<unsigned *>(check), (long long)n,
                           (long long)count, (int)double_precision);
    else
        hipLaunchKernelGGL(fs_k_rl_bench, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                           static_cast<fs_ranlux *>(states), static_cast<unsigned *>(check), (long long)n,
                           (long long)count, (int)lux, (int)double_precision);
    FS_HIP(hipGetLastError());
    return FS_OK;
}

}  // extern "C"

#include "fs_lp_host.inc"
