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
    FS_DEV fs_u32 integer() { return fs_xs_int(g); }
    // kernel.c:715: the low bits of one xorshift word
    FS_DEV fs_u32 below(fs_u32 n) { return fs_xs_int(g) & (n - 1u); }
#endif
};

