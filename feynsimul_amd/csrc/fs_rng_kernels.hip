// fs_rng_kernels.hip -- system-independent random-stream kernels, compiled
// ahead of time for gfx950 into libfeynsimul.so: stream seeding, RANLUX
// initialisation (the reference's ranlux_init_kernel / ranlux_group_init_kernel,
// FeynSimul/GPUsrc/kernel.c:152-166, with a distinct seed per stream) and
// stream dumps used by the bit-exactness tests.
#include <hip/hip_runtime.h>

#include "device/fs_rng.h"

// Row i of a xorshift128 table: two splitmix64 outputs from state seed+first+i.
extern "C" __global__ void fs_k_seed_xorshift(uint4 *__restrict__ rows, long long n, fs_u64 seed, long long first)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fs_u64 st = seed + (fs_u64)(first + i);
    const fs_u64 a = fs_splitmix64_next(st);
    const fs_u64 b = fs_splitmix64_next(st);
    uint4 v = make_uint4((fs_u32)a, (fs_u32)(a >> 32), (fs_u32)b, (fs_u32)(b >> 32));
    if ((v.x | v.y | v.z | v.w) == 0u) v.x = 1u;  // the all-zero state is absorbing
    rows[i] = v;
}

// tab[i] = ranluxcl_init(seeds[i]).
extern "C" __global__ void fs_k_ranlux_init(fs_ranlux *__restrict__ tab, const fs_u64 *__restrict__ seeds, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fs_ranlux r;
    fs_rl_init(seeds[i], r);
    fs_rl_store(tab, i, r);
}

// ndraw consecutive draws per row: randInt (uint), randFloat in FP32 and FP64.
extern "C" __global__ void fs_k_xs_draw(const uint4 *__restrict__ rows, fs_u32 *__restrict__ outInt,
                                        float *__restrict__ outF32, double *__restrict__ outF64, long long n,
                                        int ndraw)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint4 v = rows[i];
    fs_xs128 a{v.x, v.y, v.z, v.w}, b = a, c = a;
    for (int k = 0; k < ndraw; k++) {
        const long long o = i * ndraw + k;
        if (outInt) outInt[o] = fs_xs_int(a);
        if (outF32) outF32[o] = fs_xs_float32(b);
        if (outF64) outF64[o] = fs_xs_float64(c);
    }
}

template <int NSKIP>
__device__ void fs_rl_draw_one(fs_ranlux *tab, float *out32, double *out64, long long i, int nquad)
{
    fs_ranlux r;
    fs_rl_load(tab, i, r);
    for (int q = 0; q < nquad; q++) {
        const long long o = (i * nquad + q) * 4;
        if (out64) {
            double v[4];
            fs_rl_next4_double<NSKIP>(r, v);
            for (int k = 0; k < 4; k++) out64[o + k] = v[k];
        } else {
            float v[4];
            fs_rl_next4<NSKIP>(r, v);
            for (int k = 0; k < 4; k++) out32[o + k] = v[k];
        }
    }
    fs_rl_store(tab, i, r);
}

// nquad ranluxcl32 (or ranluxcl64) calls per state; the states advance.
extern "C" __global__ void fs_k_rl_draw(fs_ranlux *__restrict__ tab, float *__restrict__ out32,
                                        double *__restrict__ out64, long long n, int nquad, int lux)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    switch (lux) {
    case 0: fs_rl_draw_one<fs_ranlux_nskip(0)>(tab, out32, out64, i, nquad); break;
    case 1: fs_rl_draw_one<fs_ranlux_nskip(1)>(tab, out32, out64, i, nquad); break;
    case 2: fs_rl_draw_one<fs_ranlux_nskip(2)>(tab, out32, out64, i, nquad); break;
    case 3: fs_rl_draw_one<fs_ranlux_nskip(3)>(tab, out32, out64, i, nquad); break;
    default: fs_rl_draw_one<fs_ranlux_nskip(4)>(tab, out32, out64, i, nquad); break;
    }
}

// ---- stream throughput (the reference's xorshift_test_kernel and
// ranlux_test_kernel, examples/ranlux_test_kernel.c:90-200) ----
// Every stream draws `count` uniforms through the sampler's wrapper (RANLUX:
// buffered float4, counter reset at launch start, kernel.c:180-207) and keeps
// the XOR of their bit patterns instead of storing them; the stream state
// advances.  One stream per lane.
template <typename F> __device__ __forceinline__ fs_u32 fs_bits_of(F v)
{
    if constexpr (sizeof(F) == 8) {
        const fs_u64 b = (fs_u64)__double_as_longlong(v);
        return (fs_u32)b ^ (fs_u32)(b >> 32);
    } else {
        return __float_as_uint(v);
    }
}

template <typename F>
__device__ void fs_xs_bench_one(uint4 *rows, fs_u32 *check, long long i, long long count)
{
    const uint4 v = rows[i];
    fs_xs128 s{v.x, v.y, v.z, v.w};
    fs_u32 acc = 0u;
    for (long long k = 0; k < count; k++) acc ^= fs_bits_of(fs_xs_float<F>(s));
    rows[i] = make_uint4(s.x, s.y, s.z, s.w);
    check[i] = acc;
}

template <typename F, int NSKIP>
__device__ void fs_rl_bench_one(fs_ranlux *tab, fs_u32 *check, long long i, long long count)
{
    fs_rl_wrap<F, NSKIP> g;
    fs_rl_load(tab, i, g.st);
    g.cnt = 0u;
    fs_u32 acc = 0u;
    for (long long k = 0; k < count; k++) acc ^= fs_bits_of(g.next_float());
    fs_rl_store(tab, i, g.st);
    check[i] = acc;
}

template <typename F>
__device__ void fs_rl_bench_lux(fs_ranlux *tab, fs_u32 *check, long long i, long long count, int lux)
{
    switch (lux) {
    case 0: fs_rl_bench_one<F, fs_ranlux_nskip(0)>(tab, check, i, count); break;
    case 1: fs_rl_bench_one<F, fs_ranlux_nskip(1)>(tab, check, i, count); break;
    case 2: fs_rl_bench_one<F, fs_ranlux_nskip(2)>(tab, check, i, count); break;
    case 3: fs_rl_bench_one<F, fs_ranlux_nskip(3)>(tab, check, i, count); break;
    default: fs_rl_bench_one<F, fs_ranlux_nskip(4)>(tab, check, i, count); break;
    }
}

extern "C" __global__ void fs_k_xs_bench(uint4 *__restrict__ rows, fs_u32 *__restrict__ check, long long n,
                                         long long count, int dbl)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dbl) fs_xs_bench_one<double>(rows, check, i, count);
    else fs_xs_bench_one<float>(rows, check, i, count);
}

extern "C" __global__ void fs_k_rl_bench(fs_ranlux *__restrict__ tab, fs_u32 *__restrict__ check, long long n,
                                         long long count, int lux, int dbl)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dbl) fs_rl_bench_lux<double>(tab, check, i, count, lux);
    else fs_rl_bench_lux<float>(tab, check, i, count, lux);
}

// HBM ceilings the samplers' rooflines are read against: a grid-stride stream
// of 16-byte vectors that copies (kind 0), only reads (kind 1, the sum of the
// values is written only if it equals an impossible sentinel) or only writes
// (kind 2) n vectors.
extern "C" __global__ void fs_k_hbm_probe(const float4 *__restrict__ src, float4 *__restrict__ dst, long long n,
                                          int kind)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (kind == 0) {
        for (; i < n; i += stride) dst[i] = src[i];
    } else if (kind == 1) {
        float s = 0.0f;
        for (; i < n; i += stride) {
            const float4 v = src[i];
            s += (v.x + v.y) + (v.z + v.w);
        }
        if (s == -1.2345e-38f) dst[0] = make_float4(s, s, s, s);
    } else {
        for (; i < n; i += stride) dst[i] = make_float4(1.0f, 2.0f, 3.0f, 4.0f);
    }
}
