// fs_stream.h -- a sampler's view of one random stream: xorshift128 rows or
// RANLUX states with the per-launch buffered wrapper (kernel.c:171-239).
// Needs FLOAT_TYPE, FS_RANLUX, FS_NSKIP and FS_ISCALE.
#pragma once

#include "fs_rng.h"

typedef FLOAT_TYPE fs_f;

struct fs_stream {
#if FS_RANLUX
    fs_rl_wrap<fs_f, FS_NSKIP> g;
    FS_DEV void load(const fs_ranlux *tab, long idx)
    {
        fs_rl_load(tab, idx, g.st);
        g.cnt = 0u;
    }
    FS_DEV void store(fs_ranlux *tab, long idx) const { fs_rl_store(tab, idx, g.st); }
    FS_DEV fs_f uniform() { return g.next_float(); }
    // K uniforms in a row, with at most two refill sites
    template <int K> FS_DEV void uniforms(fs_f u[K]) { g.template draw<K>(u); }
    FS_DEV fs_u32 integer() { return g.next_int((fs_f)FS_ISCALE); }
    // Uniform index in [0, n), n a power of two <= 2^24: the top log2(n) bits
    // of one buffered uniform.  ranluxIntWrapper's (uint)(u * 2^31) & (n - 1)
    // is 0 whenever u lies on the 2^-24 grid (kernel.c:206, 715), which would
    // freeze the segment offsets; u * n is exact and below n.
    FS_DEV fs_u32 below(fs_u32 n) { return (fs_u32)(g.next_float() * (fs_f)n); }
#else
    fs_xs128 g;
    FS_DEV void load(const fs_u32 *rows, long idx)
    {
        const uint4 v = reinterpret_cast<const uint4 *>(rows)[idx];
        g.x = v.x; g.y = v.y; g.z = v.z; g.w = v.w;
    }
    FS_DEV void store(fs_u32 *rows, long idx) const
    {
        reinterpret_cast<uint4 *>(rows)[idx] = make_uint4(g.x, g.y, g.z, g.w);
    }
    FS_DEV fs_f uniform() { return fs_xs_float<fs_f>(g); }
    template <int K> FS_DEV void uniforms(fs_f u[K])
    {
#pragma unroll
        for (int k = 0; k < K; k++) u[k] = uniform();
    }
    FS_DEV fs_u32 integer() { return fs_xs_int(g); }
    // kernel.c:715: the low bits of one xorshift word
    FS_DEV fs_u32 below(fs_u32 n) { return fs_xs_int(g) & (n - 1u); }
#endif
};

#if FS_RANLUX
// The RANLUX stream that every lane of a wavefront holds in the same state
// (the wave sampler's group stream, loaded by every lane from one table
// entry).  The state sits in LDS between refills, in register format, so its
// 26 words take registers only while a refill runs instead of through the
// whole sweep; the cycle position is read into a scalar register, so the
// refill's position switch is a uniform branch.  Lane 0 writes the state
// back; the fences keep the next refill from reusing a register copy.
struct fs_wave_stream {
    typedef fs_f vec4 __attribute__((ext_vector_type(4)));
    fs_ranlux *sh;
    vec4 buf;
    fs_u32 cnt;

    FS_DEV void load(const fs_ranlux *tab, long idx, fs_ranlux *lds)
    {
        sh = lds;
        fs_ranlux st;
        fs_rl_load(tab, idx, st);
        if (__lane_id() == 0) *sh = st;
        cnt = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // lane 0 writes the table entry
    FS_DEV void store(fs_ranlux *tab, long idx) const
    {
        if (__lane_id() == 0) {
            const fs_ranlux st = *sh;
            fs_rl_store(tab, idx, st);
        }
    }
    FS_DEV fs_f uniform()
    {
        if ((cnt & 3u) == 0u) {
            fs_ranlux st = *sh;
            st.stepnr = __builtin_amdgcn_readfirstlane(st.stepnr);
            st.in24 = __builtin_amdgcn_readfirstlane(st.in24);
            if constexpr (sizeof(fs_f) == 8) {
                double v[4];
                fs_rl_next4_double<FS_NSKIP>(st, v);
                buf = vec4{(fs_f)v[0], (fs_f)v[1], (fs_f)v[2], (fs_f)v[3]};
            } else {
                float v[4];
                fs_rl_next4<FS_NSKIP>(st, v);
                buf = vec4{v[0], v[1], v[2], v[3]};
            }
            // the four values are the same in every lane: scalar registers
            if constexpr (sizeof(fs_f) == 4) {
                buf.x = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(buf.x)));
                buf.y = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(buf.y)));
                buf.z = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(buf.z)));
                buf.w = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(buf.w)));
            }
            if (__lane_id() == 0) *sh = st;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const fs_u32 k = (cnt++) & 3u;
        const vec4 b = buf;
        fs_f v = b.x;
        v = (k == 1u) ? b.y : v;
        v = (k == 2u) ? b.z : v;
        v = (k == 3u) ? b.w : v;
        return v;
    }
    FS_DEV fs_u32 below(fs_u32 n) { return (fs_u32)(uniform() * (fs_f)n); }
};
#endif
