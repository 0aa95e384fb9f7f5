// fs_rng.h -- device random number generators of the PIMC engine (gfx950).
//
// Shared by the ahead-of-time utility kernels (fs_rng_kernels.hip) and by the
// sampler template that hiprtc compiles per system (fs_sampler.h).
//
//  * xorshift128 (Marsaglia 2003) with the reference's conversions,
//    FeynSimul/GPUsrc/kernel.c:214-239.
//  * RANLUX as implemented by RANLUXCL v1.3.1 (24-bit subtract-with-borrow,
//    lags 10 and 24, luxury skipping in the "planar" and "planar shift"
//    schemes, ranluxcl_init seeding), examples/pyopencl-ranluxcl.cl:265-725,
//    893-920, and the per-launch float4 buffering of kernel.c:180-207.
//  * splitmix64 (Steele, Lea and Flood 2014) to derive per-stream seeds.
//
// The RANLUX lag table is a 24-entry register array; every index below is a
// compile-time constant so the state never leaves VGPRs.
#pragma once

typedef unsigned int fs_u32;
typedef unsigned long long fs_u64;

#ifndef FS_DEV
#define FS_DEV __device__ __forceinline__
#endif

// ---------------------------------------------------------------------------
// splitmix64: stream row r is seeded from the generator state seed + r.
// ---------------------------------------------------------------------------
FS_DEV fs_u64 fs_splitmix64_next(fs_u64 &state)
{
    state += 0x9E3779B97F4A7C15ull;
    fs_u64 z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ---------------------------------------------------------------------------
// xorshift128
// ---------------------------------------------------------------------------
struct fs_xs128 {
    fs_u32 x, y, z, w;
};

// w' = w ^ (w >> 19) ^ t ^ (t >> 8), t = x ^ (x << 11): the exclusive or of
// w, w >> 19 and t is one v_bitop3_b32 (truth table 0x96 = a ^ b ^ c), so a
// step costs six instructions instead of seven.
FS_DEV void fs_xs_step(fs_xs128 &s)
{
    const fs_u32 t = s.x ^ (s.x << 11);
    s.x = s.y;
    s.y = s.z;
    s.z = s.w;
    s.w = __builtin_amdgcn_bitop3_b32(s.w, s.w >> 19, t, 0x96) ^ (t >> 8);
}

FS_DEV fs_u32 fs_xs_int(fs_xs128 &s)
{
    fs_xs_step(s);
    return s.w;
}

// FP32: w rounded to float times the literal 2.328306437080797e-10 read as a
// float, which is exactly 2^-32, so the value lies in [0, 1].
FS_DEV float fs_xs_float32(fs_xs128 &s)
{
    fs_xs_step(s);
    return (float)s.w * 0x1p-32f;
}

// FP64: w times the double literal 2.328306437080797e-10 (about 1/(2^32-1)).
FS_DEV double fs_xs_float64(fs_xs128 &s)
{
    fs_xs_step(s);
    return (double)s.w * 2.328306437080797e-10;
}

template <typename F> FS_DEV F fs_xs_float(fs_xs128 &s);
template <> FS_DEV float fs_xs_float<float>(fs_xs128 &s) { return fs_xs_float32(s); }
template <> FS_DEV double fs_xs_float<double>(fs_xs128 &s) { return fs_xs_float64(s); }

// ---------------------------------------------------------------------------
// RANLUX
// ---------------------------------------------------------------------------
// In-memory layout (112 bytes): s01..s24, carry, dummy, in24, stepnr -- the
// field order of ranluxcl_state_t, so a state table is interchangeable.  In
// registers `carry` is kept in units of 2^-24 (0 or 1); fs_rl_load and
// fs_rl_store convert.
struct fs_ranlux {
    float s[24];
    float carry;
    float dummy;
    int in24;
    int stepnr;
};

#define FS_TWOM24 0x1p-24f
#define FS_TWOM12 0x1p-12f

// Luxury level -> values skipped after each 24 produced.
__host__ __device__ constexpr int fs_ranlux_nskip(int lux)
{
    return lux == 0 ? 0 : lux == 1 ? 24 : lux == 2 ? 76 : lux == 3 ? 200 : lux == 4 ? 380 : lux - 24;
}

// One SWB step at cycle position T: x = s[lag 10] - s[lag 24] - carry, taken
// mod 1 with a borrow; the result replaces the lag-24 entry.  Values below
// 2^-12 get 24 more low bits from the next lag-10 entry (never 0 returned).
// Every value is a multiple of 2^-24 in (-1, 1), so each operation below is
// exact: fract(x) is x + 1 for the negative differences (the borrow), the new
// carry in units of 2^-24 is fract(x) - x, and the low-bit fill is one FMA
// whose result is 0 or at least 2^-48, so the max replaces exactly a 0 by
// 2^-48.  Four instructions per step and no branches.
template <int T> FS_DEV float fs_rl_step(fs_ranlux &r)
{
    constexpr int tgt = 23 - T;
    constexpr int lag10 = (9 - T + 24) % 24;
    constexpr int lag9 = (8 - T + 24) % 24;
    const float x = fmaf(r.carry, -FS_TWOM24, r.s[lag10] - r.s[tgt]);
    const float uni = __builtin_amdgcn_fractf(x);
    r.carry = uni - x;
    r.s[tgt] = uni;
    const float out = fmaf(uni < FS_TWOM12 ? FS_TWOM24 : 0.0f, r.s[lag9], uni);
    return fmaxf(out, FS_TWOM24 * FS_TWOM24);
}

// Steps [FIRST, LAST) of one cycle, outputs discarded.
template <int FIRST, int LAST> FS_DEV void fs_rl_run(fs_ranlux &r)
{
    if constexpr (FIRST < LAST) {
        (void)fs_rl_step<FIRST>(r);
        fs_rl_run<FIRST + 1, LAST>(r);
    }
}

template <int T0> FS_DEV void fs_rl_quad(fs_ranlux &r, float &a, float &b, float &c, float &d)
{
    a = fs_rl_step<T0 + 0>(r);
    b = fs_rl_step<T0 + 1>(r);
    c = fs_rl_step<T0 + 2>(r);
    d = fs_rl_step<T0 + 3>(r);
}

// Luxury skip of NSKIP values starting at cycle position T0 (a multiple of 4):
// finish the cycle, run the whole cycles, then the first `rest` steps of the
// next one; every count is a compile-time constant for a given T0.
template <int NSKIP, int T0> FS_DEV void fs_rl_skip_from(fs_ranlux &r)
{
    constexpr int initial = T0 ? 24 - T0 : 0;
    constexpr int cycles = (NSKIP - initial) / 24;
    constexpr int rest = NSKIP - initial - cycles * 24;
    static_assert(NSKIP >= initial && rest % 4 == 0, "skips restart at a multiple of four");
    if constexpr (T0 != 0) fs_rl_run<T0, 24>(r);
#pragma unroll 1
    for (int c = 0; c < cycles; c++) fs_rl_run<0, 24>(r);
    if constexpr (rest > 0) fs_rl_run<0, rest>(r);
    r.stepnr = rest;
}

// Luxury skip of NSKIP values starting at the current cycle position (one
// branch on the position, which every lane of a sampler shares).
template <int NSKIP> FS_DEV void fs_rl_skip(fs_ranlux &r)
{
    switch (r.stepnr) {
    case 4: fs_rl_skip_from<NSKIP, 4>(r); break;
    case 8: fs_rl_skip_from<NSKIP, 8>(r); break;
    case 12: fs_rl_skip_from<NSKIP, 12>(r); break;
    case 16: fs_rl_skip_from<NSKIP, 16>(r); break;
    case 20: fs_rl_skip_from<NSKIP, 20>(r); break;
    default: fs_rl_skip_from<NSKIP, 0>(r); break;
    }
}

// ranluxcl32: four values in (0, 1).
template <int NSKIP> FS_DEV void fs_rl_next4(fs_ranlux &r, float out[4])
{
    const int t0 = r.stepnr;
    switch (t0) {
    case 0: fs_rl_quad<0>(r, out[0], out[1], out[2], out[3]); break;
    case 4: fs_rl_quad<4>(r, out[0], out[1], out[2], out[3]); break;
    case 8: fs_rl_quad<8>(r, out[0], out[1], out[2], out[3]); break;
    case 12: fs_rl_quad<12>(r, out[0], out[1], out[2], out[3]); break;
    case 16: fs_rl_quad<16>(r, out[0], out[1], out[2], out[3]); break;
    default: fs_rl_quad<20>(r, out[0], out[1], out[2], out[3]); break;
    }
    r.stepnr = (t0 == 20) ? 0 : t0 + 4;
    if constexpr (NSKIP % 24 == 0) {
        if constexpr (NSKIP > 0) {
            if (t0 == 20) fs_rl_skip<NSKIP>(r);
        }
    } else {
        r.in24 += 4;
        if (r.in24 == 24) {
            r.in24 = 0;
            fs_rl_skip<NSKIP>(r);
        }
    }
}

// ranluxcl64: two float4 draws make four 48-bit doubles.
template <int NSKIP> FS_DEV void fs_rl_next4_double(fs_ranlux &r, double out[4])
{
#pragma unroll
    for (int h = 0; h < 2; h++) {
        float v[4];
        fs_rl_next4<NSKIP>(r, v);
        if (v[0] == FS_TWOM24 * FS_TWOM24) v[0] = 0.0f;
        if (v[2] == FS_TWOM24 * FS_TWOM24) v[2] = 0.0f;
        out[2 * h + 0] = (double)v[0] + (double)v[1] / 16777216;
        out[2 * h + 1] = (double)v[2] + (double)v[3] / 16777216;
    }
}

// ranluxcl_init: xorshift64 (13, 7, 17) expands the seed into 24 fractions of
// 24 bits each, then 32 cycles of warm-up.
FS_DEV void fs_rl_init(fs_u64 x, fs_ranlux &r)
{
    const float inv = 0x1p-24f;
#pragma unroll
    for (int blk = 0; blk < 3; blk++) {
        fs_u64 a, b, c;
        x ^= (x << 13); x ^= (x >> 7); x ^= (x << 17); a = x;
        x ^= (x << 13); x ^= (x >> 7); x ^= (x << 17); b = x;
        x ^= (x << 13); x ^= (x >> 7); x ^= (x << 17); c = x;
        r.s[8 * blk + 0] = (float)(a >> 40) * inv;
        r.s[8 * blk + 1] = (float)((a & 0x000000FFFFFFFFFFull) >> 16) * inv;
        r.s[8 * blk + 2] = (float)(((a & 0x000000000000FFFFull) << 8) + (b >> 56)) * inv;
        r.s[8 * blk + 3] = (float)((b & 0x00FFFFFFFFFFFFFFull) >> 32) * inv;
        r.s[8 * blk + 4] = (float)((b & 0x00000000FFFFFFFFull) >> 8) * inv;
        r.s[8 * blk + 5] = (float)(((b & 0x00000000000000FFull) << 16) + (c >> 48)) * inv;
        r.s[8 * blk + 6] = (float)((c & 0x0000FFFFFFFFFFFFull) >> 24) * inv;
        r.s[8 * blk + 7] = (float)(c & 0x0000000000FFFFFFull) * inv;
    }
    r.dummy = 0.0f;
    r.in24 = 0;
    r.stepnr = 0;
    r.carry = (r.s[23] == 0.0f) ? 1.0f : 0.0f;
    for (int i = 0; i < 32; i++) fs_rl_run<0, 24>(r);
}

FS_DEV void fs_rl_load(const fs_ranlux *__restrict__ tab, long idx, fs_ranlux &r)
{
    const float4 *p = reinterpret_cast<const float4 *>(tab + idx);
#pragma unroll
    for (int q = 0; q < 6; q++) {
        float4 v = p[q];
        r.s[4 * q + 0] = v.x;
        r.s[4 * q + 1] = v.y;
        r.s[4 * q + 2] = v.z;
        r.s[4 * q + 3] = v.w;
    }
    float4 t = p[6];
    r.carry = t.x * 16777216.0f;
    r.dummy = t.y;
    r.in24 = __float_as_int(t.z);
    r.stepnr = __float_as_int(t.w);
}

FS_DEV void fs_rl_store(fs_ranlux *__restrict__ tab, long idx, const fs_ranlux &r)
{
    float4 *p = reinterpret_cast<float4 *>(tab + idx);
#pragma unroll
    for (int q = 0; q < 6; q++) p[q] = make_float4(r.s[4 * q], r.s[4 * q + 1], r.s[4 * q + 2], r.s[4 * q + 3]);
    p[6] = make_float4(r.carry * FS_TWOM24, r.dummy, __int_as_float(r.in24), __int_as_float(r.stepnr));
}

// Per-launch buffered draw (kernel.c:180-207): refill four values whenever the
// launch-local counter is a multiple of four.
template <typename F, int NSKIP> struct fs_rl_wrap {
    typedef F vec4 __attribute__((ext_vector_type(4)));
    fs_ranlux st;
    vec4 buf;
    fs_u32 cnt;

    FS_DEV F next_float()
    {
        if ((cnt & 3u) == 0u) {
            if constexpr (sizeof(F) == 8) {
                double v[4];
                fs_rl_next4_double<NSKIP>(st, v);
                buf = vec4{v[0], v[1], v[2], v[3]};
            } else {
                float v[4];
                fs_rl_next4<NSKIP>(st, v);
                buf = vec4{v[0], v[1], v[2], v[3]};
            }
        }
        const fs_u32 k = (cnt++) & 3u;
        // selects on the buffer's value: an index into the buffer in memory
        // (or a select between its addresses, which the optimiser may form
        // from selects between loads) would keep the whole stream state in
        // scratch memory
        const vec4 b = buf;
        F v = b.x;
        v = (k == 1u) ? b.y : v;
        v = (k == 2u) ? b.z : v;
        v = (k == 3u) ? b.w : v;
        return v;
    }

    // K draws in a row, the same values next_float gives K times.  The
    // buffer's phase p = cnt mod 4 is the same in every lane, m = (4 - p) mod
    // 4 values are left in it, and the rest come from ceil((K - m) / 4)
    // refills: RMIN of them always, one more for the phases that need it.  So
    // a block has at most two refill sites instead of one per draw.
    template <int K> FS_DEV void draw(F u[K])
    {
        constexpr int RMIN = K > 3 ? (K - 3 + 3) / 4 : 0;
        constexpr int RMAX = (K + 3) / 4;
        const fs_u32 p = cnt & 3u;
        const int m = (int)((4u - p) & 3u);
        const vec4 old = buf;
        vec4 fresh[RMAX];
#pragma unroll
        for (int r = 0; r < RMIN; r++) fresh[r] = refill();
        if constexpr (RMAX > RMIN) {
            if (K - m > 4 * RMIN) fresh[RMAX - 1] = refill();
            else fresh[RMAX - 1] = old;
        }
        // value j of the concatenation (old[p..3], fresh...)
#pragma unroll
        for (int k = 0; k < K; k++) {
            F v = fresh[k / 4][k % 4];  // p = 0
            if (k < 3) {
                const F o1 = old[(k + 1) & 3], o2 = old[(k + 2) & 3], o3 = old[(k + 3) & 3];
                v = p == 1u ? (k < 3 ? o1 : v) : v;
                v = p == 2u ? (k < 2 ? o2 : v) : v;
                v = p == 3u ? (k < 1 ? o3 : v) : v;
            }
            if (k >= 1) v = p == 3u ? fresh[(k - 1) / 4][(k - 1) % 4] : v;
            if (k >= 2) v = p == 2u ? fresh[(k - 2) / 4][(k - 2) % 4] : v;
            if (k >= 3) v = p == 1u ? fresh[(k - 3) / 4][(k - 3) % 4] : v;
            u[k] = v;
        }
        // the group the next draw continues from
        const int R = (K - m + 3) / 4;
        if (R > 0) buf = R == RMAX ? fresh[RMAX - 1] : fresh[RMIN > 0 ? RMIN - 1 : 0];
        cnt += (fs_u32)K;
    }

    FS_DEV vec4 refill()
    {
        if constexpr (sizeof(F) == 8) {
            double v[4];
            fs_rl_next4_double<NSKIP>(st, v);
            return vec4{v[0], v[1], v[2], v[3]};
        } else {
            float v[4];
            fs_rl_next4<NSKIP>(st, v);
            return vec4{v[0], v[1], v[2], v[3]};
        }
    }

    // ranluxIntWrapper: (uint)(u * float(ranluxIntMax + 1)).
    FS_DEV fs_u32 next_int(F scale) { return (fs_u32)(next_float() * scale); }
};
