"""ctypes binding of libfeynsimul.so (declarations in include/feynsimul.h).

The library is the only compute path: importing this module fails loudly if
it has not been built (``python -c "import __graft_entry__ as g; g.build()"``
or ``make -C feynsimul_amd/csrc``).
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfeynsimul.so")

FS_ABI_VERSION = 6
FS_RUN_POT_CACHE_VALID = 1  # fs_run_flags: the potential cache is current
FS_OK, FS_ERR_INVALID, FS_ERR_COMPILE, FS_ERR_HIP, FS_ERR_NO_DEVICE = 0, 1, 2, 3, 4
FS_SAMPLER_EXACT, FS_SAMPLER_WAVE = 0, 1
FS_RNG_XORSHIFT, FS_RNG_RANLUX = 0, 1
FS_STORAGE_AUTO, FS_STORAGE_MEMORY, FS_STORAGE_REGISTERS = 0, 1, 2
FS_POTCACHE_AUTO, FS_POTCACHE_ON, FS_POTCACHE_OFF = 0, 1, 2

EXPORTS = ("fs_abi_version", "fs_last_error", "fs_device_count", "fs_create", "fs_destroy", "fs_path_storage",
           "fs_batch_create", "fs_batch_destroy", "fs_batch_path_storage", "fs_batch_build_log", "fs_batch_run",
           "fs_storage_of", "fs_launch_info", "fs_hbm_probe",
           "fs_batch_run_noalloc",
           "fs_batch_capture", "fs_graph_launch", "fs_graph_destroy", "fs_build_log",
           "fs_source", "fs_run", "fs_run_flags", "fs_set_waves_per_simd", "fs_eval_system", "fs_seed_xorshift", "fs_ranlux_init", "fs_xorshift_draw",
           "fs_ranlux_draw", "fs_rng_bench", "fs_lp_create", "fs_lp_destroy", "fs_lp_get_info", "fs_lp_build_log", "fs_lp_pack",
           "fs_lp_sweep", "fs_lp_sweep_measure", "fs_lp_train", "fs_lp_eval", "fs_lp_grad")


class FsConfig(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_int32), ("sampler", ctypes.c_int32), ("N", ctypes.c_int32),
        ("S", ctypes.c_int32), ("DOF", ctypes.c_int32), ("walkers_per_group", ctypes.c_int32),
        ("n_walkers", ctypes.c_int64), ("parallelize_path", ctypes.c_int32), ("local_size", ctypes.c_int32),
        ("double_precision", ctypes.c_int32), ("ranlux", ctypes.c_int32), ("ranlux_lux", ctypes.c_int32),
        ("fast_math", ctypes.c_int32), ("ranlux_int_range", ctypes.c_uint64), ("n_ops", ctypes.c_int32),
        ("n_corr", ctypes.c_int32), ("cor_lags", ctypes.c_int32), ("bins", ctypes.c_int32),
        ("bin_res", ctypes.c_int32), ("waves_per_block", ctypes.c_int32), ("beta", ctypes.c_double),
        ("x_min", ctypes.c_double), ("x_max", ctypes.c_double), ("op_norm", ctypes.c_double),
        ("path_storage", ctypes.c_int32), ("n_hist_obs", ctypes.c_int32), ("hist_bins", ctypes.c_int32),
        ("hist_lo", ctypes.c_double), ("hist_hi", ctypes.c_double), ("no_bisection", ctypes.c_int32),
        ("path_shift", ctypes.c_int32), ("single_node", ctypes.c_int32), ("max_waves_per_simd", ctypes.c_int32),
        ("alpha", ctypes.c_double), ("ps_alpha", ctypes.c_double), ("potential_cache", ctypes.c_int32)]


class FsBuffers(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "paths", "accepts", "seeds", "seeds_snap", "rl", "grl", "old_paths", "op_means", "op_totals", "series",
        "cor_means", "bins", "pot_cache", "obs_hist")]


class FsLpConfig(ctypes.Structure):
    _fields_ = [("abi_version", ctypes.c_int32), ("atoms3d", ctypes.c_int32), ("batch", ctypes.c_int32),
                ("segs_per_wave", ctypes.c_int32), ("in_scale", ctypes.c_float), ("lr", ctypes.c_float),
                ("beta1", ctypes.c_float), ("beta2", ctypes.c_float), ("adam_eps", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("coarse", ctypes.c_int32),
                ("width", ctypes.c_int32), ("bridge_omega2", ctypes.c_float)]


class FsLpInfo(ctypes.Structure):
    _fields_ = [("n_params", ctypes.c_int64), ("image_bytes", ctypes.c_int64), ("train_bf16", ctypes.c_int64),
                ("train_bias", ctypes.c_int64), ("workspace_bytes", ctypes.c_int64), ("loss_parts", ctypes.c_int64),
                ("segs_per_wave", ctypes.c_int32), ("waves_per_block", ctypes.c_int32),
                ("lds_bytes", ctypes.c_int32), ("grid_blocks", ctypes.c_int32)]


class FsLpBuffers(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "params", "adam_m", "adam_v", "image", "train_w", "train_b", "workspace", "loss_parts", "trace")]


class FsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libfeynsimul.so is not built (%s); run __graft_entry__.build() or "
                          "make -C feynsimul_amd/csrc" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
    L.fs_abi_version.restype = ctypes.c_int
    L.fs_last_error.restype = ctypes.c_char_p
    L.fs_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.fs_create.argtypes = [ctypes.POINTER(FsConfig), ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(vp)]
    L.fs_destroy.argtypes = [vp]
    L.fs_path_storage.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.fs_batch_create.argtypes = [ctypes.POINTER(FsConfig), ctypes.POINTER(ctypes.c_char_p), i32, i32,
                                  ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_int,
                                  ctypes.POINTER(vp)]
    L.fs_batch_destroy.argtypes = [vp]
    L.fs_storage_of.argtypes = [ctypes.POINTER(FsConfig), ctypes.POINTER(ctypes.c_int)]
    L.fs_launch_info.argtypes = [vp, ctypes.POINTER(FsLaunchInfo)]
    L.fs_hbm_probe.argtypes = [i32, vp, vp, ctypes.c_int64, i32, vp]
    L.fs_batch_path_storage.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int)]
    L.fs_batch_build_log.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.fs_batch_run.argtypes = [vp, ctypes.POINTER(FsBuffers), ctypes.POINTER(ctypes.c_int32), i32, i32, i32, i32, vp]
    L.fs_batch_run_noalloc.argtypes = [vp, ctypes.POINTER(FsBuffers), ctypes.POINTER(ctypes.c_int32), i32, i32, i32,
                                       i32, vp]
    L.fs_batch_capture.argtypes = [vp, ctypes.POINTER(FsBuffers), ctypes.POINTER(ctypes.c_int32), i32, i32, i32, i32,
                                   i32, vp, ctypes.POINTER(vp)]
    L.fs_graph_launch.argtypes = [vp, vp]
    L.fs_graph_destroy.argtypes = [vp]
    L.fs_build_log.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.fs_source.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.fs_run.argtypes = [vp, ctypes.POINTER(FsBuffers), i32, i32, i32, vp]
    L.fs_run_flags.argtypes = [vp, ctypes.POINTER(FsBuffers), i32, i32, i32, i32, vp]
    L.fs_set_waves_per_simd.argtypes = [vp, i32]
    L.fs_eval_system.argtypes = [vp, vp, vp, vp, i64, vp]
    L.fs_seed_xorshift.argtypes = [vp, i64, u64, i64, vp]
    L.fs_ranlux_init.argtypes = [vp, vp, i64, vp]
    L.fs_xorshift_draw.argtypes = [vp, vp, vp, vp, i64, i32, vp]
    L.fs_ranlux_draw.argtypes = [vp, vp, vp, i64, i32, i32, vp]
    L.fs_rng_bench.argtypes = [i32, i32, i32, vp, vp, i64, i64, vp]
    L.fs_lp_create.argtypes = [ctypes.POINTER(FsConfig), ctypes.c_char_p, ctypes.POINTER(FsLpConfig), ctypes.c_int,
                               ctypes.POINTER(vp)]
    L.fs_lp_destroy.argtypes = [vp]
    L.fs_lp_get_info.argtypes = [vp, ctypes.POINTER(FsLpInfo)]
    L.fs_lp_build_log.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.fs_lp_pack.argtypes = [vp, ctypes.POINTER(FsLpBuffers), vp]
    L.fs_lp_sweep.argtypes = [vp, vp, vp, ctypes.POINTER(FsLpBuffers), u64, i32, vp]
    L.fs_lp_sweep_measure.argtypes = [vp, vp, vp, ctypes.POINTER(FsLpBuffers), ctypes.POINTER(FsBuffers), u64, i32,
                                      i64, vp]
    L.fs_lp_train.argtypes = [vp, vp, ctypes.POINTER(FsLpBuffers), u64, vp]
    L.fs_lp_eval.argtypes = [vp, ctypes.POINTER(FsLpBuffers), vp, vp, i64, vp]
    L.fs_lp_grad.argtypes = [vp, ctypes.POINTER(FsLpBuffers), vp, vp]
    for name in EXPORTS:
        if name not in ("fs_abi_version", "fs_last_error"):
            getattr(L, name).restype = ctypes.c_int
    if L.fs_abi_version() != FS_ABI_VERSION:
        raise ImportError("libfeynsimul.so ABI %d, expected %d" % (L.fs_abi_version(), FS_ABI_VERSION))
    _lib = L
    return L


def check(rc):
    if rc != FS_OK:
        raise FsError(rc, lib().fs_last_error().decode(errors="replace"))


def device_count():
    n = ctypes.c_int(0)
    check(lib().fs_device_count(ctypes.byref(n)))
    return n.value


def _text(fn, handle):
    need = ctypes.c_size_t(0)
    check(fn(handle, None, 0, ctypes.byref(need)))
    buf = ctypes.create_string_buffer(need.value)
    check(fn(handle, buf, need.value, ctypes.byref(need)))
    return buf.value.decode(errors="replace")


class Engine:
    """One compiled sampler module (owner of the fs_engine handle)."""

    def __init__(self, cfg: FsConfig, system_src: str, compile_only=False):
        self._h = ctypes.c_void_p()
        L = lib()
        rc = L.fs_create(ctypes.byref(cfg), system_src.encode(), int(compile_only), ctypes.byref(self._h))
        if rc != FS_OK:
            msg = L.fs_last_error().decode(errors="replace")
            self.build_log = _text(L.fs_build_log, self._h) if self._h.value else ""
            if self._h.value:
                L.fs_destroy(self._h)
                self._h = ctypes.c_void_p()
            raise FsError(rc, msg)
        self.cfg = cfg
        self.build_log = _text(L.fs_build_log, self._h)

    @property
    def path_storage(self):
        v = ctypes.c_int(0)
        check(lib().fs_path_storage(self._h, ctypes.byref(v)))
        return v.value

    @property
    def source(self):
        return _text(lib().fs_source, self._h)

    def launch_info(self):
        info = FsLaunchInfo()
        check(lib().fs_launch_info(self._h, ctypes.byref(info)))
        return info

    def set_waves_per_simd(self, cap):
        check(lib().fs_set_waves_per_simd(self._h, int(cap)))

    def run(self, bufs: FsBuffers, operator_runs, metro_steps, series_offset, stream, flags=0):
        check(lib().fs_run_flags(self._h, ctypes.byref(bufs), operator_runs, metro_steps, series_offset, flags,
                                 stream))

    def eval_system(self, x_ptr, pot_ptr, ops_ptr, n, stream):
        check(lib().fs_eval_system(self._h, x_ptr, pot_ptr, ops_ptr, n, stream))

    def close(self):
        if self._h.value:
            lib().fs_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FsLaunchInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("threads_per_block", "max_threads_per_block", "lds_bytes", "registers",
                                               "lds_limit", "device_max_threads_per_block")] + [
        ("device_max_block_dim", ctypes.c_int32 * 3)]


def storage_of(cfg):
    """The path storage a wave-sampler module for this configuration uses (FS_STORAGE_REGISTERS or _MEMORY)."""
    v = ctypes.c_int(0)
    check(lib().fs_storage_of(ctypes.byref(cfg), ctypes.byref(v)))
    return v.value


class Batch:
    """One compiled batched module (owner of the fs_batch handle)."""

    def __init__(self, cfgs, system_srcs, waves_per_block=0, compile_only=False, kernel_of_type=None,
                 waves_of_kernel=None):
        n = len(cfgs)
        self._h = ctypes.c_void_p()
        arr = (FsConfig * n)(*cfgs)
        srcs = (ctypes.c_char_p * n)(*[x.encode() for x in system_srcs])
        kot = None if kernel_of_type is None else (ctypes.c_int32 * n)(*[int(v) for v in kernel_of_type])
        wok = None if waves_of_kernel is None else (ctypes.c_int32 * len(waves_of_kernel))(
            *[int(v) for v in waves_of_kernel])
        L = lib()
        rc = L.fs_batch_create(arr, srcs, n, int(waves_per_block), kot, wok, int(compile_only), ctypes.byref(self._h))
        if rc != FS_OK:
            msg = L.fs_last_error().decode(errors="replace")
            if self._h.value:
                self.build_log = _text(L.fs_batch_build_log, self._h)
                L.fs_batch_destroy(self._h)
                self._h = ctypes.c_void_p()
            raise FsError(rc, msg)
        self.ntypes = n
        self.build_log = _text(L.fs_batch_build_log, self._h)

    def path_storage(self, t):
        v = ctypes.c_int(0)
        check(lib().fs_batch_path_storage(self._h, t, ctypes.byref(v)))
        return v.value

    @staticmethod
    def _order(order, n):
        return None if order is None else (ctypes.c_int32 * n)(*order)

    def run(self, bufs, order, operator_runs, metro_steps, series_offset, grid_blocks, stream):
        arr = (FsBuffers * self.ntypes)(*bufs)
        check(lib().fs_batch_run(self._h, arr, self._order(order, self.ntypes), operator_runs, metro_steps,
                                 series_offset, grid_blocks, stream))

    def run_noalloc(self, bufs, order, operator_runs, metro_steps, series_offset, grid_blocks, stream):
        """fs_batch_run_noalloc: dispatch tables in the library's static per-GPU array."""
        arr = (FsBuffers * self.ntypes)(*bufs)
        check(lib().fs_batch_run_noalloc(self._h, arr, self._order(order, self.ntypes), operator_runs, metro_steps,
                                         series_offset, grid_blocks, stream))

    def capture(self, bufs, order, operator_runs, metro_steps, series_offset, n_launches, grid_blocks, stream):
        arr = (FsBuffers * self.ntypes)(*bufs)
        g = ctypes.c_void_p()
        check(lib().fs_batch_capture(self._h, arr, self._order(order, self.ntypes), operator_runs, metro_steps,
                                     series_offset, n_launches, grid_blocks, stream, ctypes.byref(g)))
        return Graph(g)

    def close(self):
        if self._h.value:
            lib().fs_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """An instantiated HIP graph of batched launches."""

    def __init__(self, handle):
        self._h = handle

    def launch(self, stream):
        check(lib().fs_graph_launch(self._h, stream))

    def close(self):
        if self._h.value:
            lib().fs_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Proposal:
    """One compiled learned-proposal module (owner of the fs_lp handle)."""

    def __init__(self, cfg: FsConfig, system_src: str, lp: FsLpConfig, compile_only=False):
        self._h = ctypes.c_void_p()
        L = lib()
        rc = L.fs_lp_create(ctypes.byref(cfg), system_src.encode(), ctypes.byref(lp), int(compile_only),
                            ctypes.byref(self._h))
        if rc != FS_OK:
            msg = L.fs_last_error().decode(errors="replace")
            self.build_log = _text(L.fs_lp_build_log, self._h) if self._h.value else ""
            if self._h.value:
                L.fs_lp_destroy(self._h)
                self._h = ctypes.c_void_p()
            raise FsError(rc, msg)
        self.cfg, self.lp = cfg, lp
        self.build_log = _text(L.fs_lp_build_log, self._h)
        self.info = FsLpInfo()
        check(L.fs_lp_get_info(self._h, ctypes.byref(self.info)))

    def pack(self, bufs, stream):
        check(lib().fs_lp_pack(self._h, ctypes.byref(bufs), stream))

    def sweep(self, paths, accepts, bufs, sweep0, n_sweeps, stream):
        check(lib().fs_lp_sweep(self._h, paths, accepts, ctypes.byref(bufs), sweep0, n_sweeps, stream))

    def sweep_measure(self, paths, accepts, bufs, kbufs, sweep0, n_sweeps, series_offset, stream):
        check(lib().fs_lp_sweep_measure(self._h, paths, accepts, ctypes.byref(bufs), ctypes.byref(kbufs), sweep0,
                                        n_sweeps, series_offset, stream))

    def train(self, paths, bufs, step, stream):
        check(lib().fs_lp_train(self._h, paths, ctypes.byref(bufs), step, stream))

    def eval(self, bufs, X, O, rows, stream):
        check(lib().fs_lp_eval(self._h, ctypes.byref(bufs), X, O, rows, stream))

    def grad(self, bufs, out, stream):
        check(lib().fs_lp_grad(self._h, ctypes.byref(bufs), out, stream))

    def close(self):
        if self._h.value:
            lib().fs_lp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
