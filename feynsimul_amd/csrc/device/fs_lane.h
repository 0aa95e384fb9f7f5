// fs_lane.h -- wavefront lane exchanges and scalar-base addressing (gfx950),
// shared by the wave sampler and the learned proposal.
#pragma once

#ifndef FS_DEV
#define FS_DEV __device__ __forceinline__
#endif

template <typename T> FS_DEV T fs_lane_read(T v, int src) { return __shfl(v, src, 64); }

// The value of lane (l + 1) mod 64 on every lane l: the DPP wave rotation
// (wave_rol:1), one VALU lane exchange per 32-bit word.
template <typename T> FS_DEV T fs_lane_next(T v)
{
    static_assert(sizeof(T) % 4 == 0, "32-bit words");
    union words {
        T v;
        int w[sizeof(T) / 4];
    } a;
    a.v = v;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); k++) a.w[k] = __builtin_amdgcn_mov_dpp(a.w[k], 0x134, 0xF, 0xF, true);
    return a.v;
}

// Inclusive prefix sum over the 64 lanes with six DPP additions and no LDS:
// a log-step scan inside each 16-lane row (row_shr 1, 2, 4, 8; lanes shifted
// in from outside the row read 0), then row 0's total (lane 15) into row 1 and
// row 2's (lane 47) into row 3 (row_bcast:15), and the first half's total
// (lane 31) into rows 2 and 3 (row_bcast:31).
#define FS_DPP_F(x, ctrl, rows) \
    __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), (ctrl), (rows), 0xF, false))
FS_DEV float fs_lane_scan(float v)
{
    v += FS_DPP_F(v, 0x111, 0xF);  // row_shr:1
    v += FS_DPP_F(v, 0x112, 0xF);  // row_shr:2
    v += FS_DPP_F(v, 0x114, 0xF);  // row_shr:4
    v += FS_DPP_F(v, 0x118, 0xF);  // row_shr:8
    v += FS_DPP_F(v, 0x142, 0xA);  // row_bcast:15 into rows 1 and 3
    v += FS_DPP_F(v, 0x143, 0xC);  // row_bcast:31 into rows 2 and 3
    return v;
}

// Inclusive decayed prefix sum w_i = sum_{m <= i} q^(i - m) v_m: the additions of fs_lane_scan with
// each shifted term weighted by q to the power of its distance, uniform inside the rows (q1 = q,
// q2 = q^2, q4, q8) and per lane for the row broadcasts (wb1 = q^(i - 15) in row 1, q^(i - 47) in
// row 3; wb2 = q^(i - 31) in rows 2 and 3).  With every weight 1 it equals fs_lane_scan bit for bit.
FS_DEV float fs_lane_scan_decay(float v, float q1, float q2, float q4, float q8, float wb1, float wb2)
{
    v = fmaf(q1, FS_DPP_F(v, 0x111, 0xF), v);   // row_shr:1
    v = fmaf(q2, FS_DPP_F(v, 0x112, 0xF), v);   // row_shr:2
    v = fmaf(q4, FS_DPP_F(v, 0x114, 0xF), v);   // row_shr:4
    v = fmaf(q8, FS_DPP_F(v, 0x118, 0xF), v);   // row_shr:8
    v = fmaf(wb1, FS_DPP_F(v, 0x142, 0xA), v);  // row_bcast:15 into rows 1 and 3
    v = fmaf(wb2, FS_DPP_F(v, 0x143, 0xC), v);  // row_bcast:31 into rows 2 and 3
    return v;
}
#undef FS_DPP_F

// The value of lane l (wave-uniform l) in scalar registers, one v_readlane_b32
// per 32-bit word.
template <typename T> FS_DEV T fs_readlane(T v, int l)
{
    static_assert(sizeof(T) % 4 == 0, "32-bit words");
    union words {
        T v;
        int w[sizeof(T) / 4];
    } a;
    a.v = v;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); k++) a.w[k] = __builtin_amdgcn_readlane(a.w[k], l);
    return a.v;
}

// Element i of an array whose base is wave-uniform, addressed with a 32-bit
// unsigned byte offset: the load or store then takes the base in SGPRs and the
// offset in one VGPR, with no 64-bit address arithmetic per lane.
template <typename T> FS_DEV T &fs_at(T *base, int i)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (unsigned)i * (unsigned)sizeof(T));
}
template <typename T>
FS_DEV __attribute__((address_space(1))) T &fs_at(__attribute__((address_space(1))) T *base, int i)
{
    return *(__attribute__((address_space(1))) T *)((__attribute__((address_space(1))) char *)base +
                                                      (unsigned)i * (unsigned)sizeof(T));
}

// A wave-uniform value moved to scalar registers (the compiler cannot prove
// that threadIdx.x / 64 or a value broadcast by a shuffle is uniform).
FS_DEV long fs_uniform(long v)
{
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)((unsigned long long)v >> 32));
    return (long)(((unsigned long long)hi << 32) | lo);
}

// One stage of an xor butterfly, v(l) + v(l ^ OFF), for a butterfly whose
// stages with smaller offsets are done, so aligned groups of OFF lanes hold
// equal values.  Every stage is a VALU lane exchange (no LDS round trip):
// OFF 1 and 2 are quad permutations; 4 and 8 the half-row and row mirrors
// (lane i <- 7 - i, 15 - i: a lane of the partner group, which holds the
// partner's value); 16 and 32 the gfx950 row and half-wave swaps, after which
// one register holds the even and the other the odd member of each pair.
// Addition is commutative, so every lane ends with the bits of the shuffle
// butterfly v += shfl_xor(v, OFF).
template <int OFF, typename T> FS_DEV T fs_bfly_stage(T v)
{
    static_assert(sizeof(T) % 4 == 0, "32-bit words");
    constexpr int NW = sizeof(T) / 4;
    union words {
        T v;
        unsigned w[NW];
    };
    words a, b;
    a.v = v;
#pragma unroll
    for (int k = 0; k < NW; k++) {
        if constexpr (OFF == 1) {
            b.w[k] = __builtin_amdgcn_update_dpp(0u, a.w[k], 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
        } else if constexpr (OFF == 2) {
            b.w[k] = __builtin_amdgcn_update_dpp(0u, a.w[k], 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
        } else if constexpr (OFF == 4) {
            b.w[k] = __builtin_amdgcn_update_dpp(0u, a.w[k], 0x141, 0xF, 0xF, false);  // row_half_mirror
        } else if constexpr (OFF == 8) {
            b.w[k] = __builtin_amdgcn_update_dpp(0u, a.w[k], 0x140, 0xF, 0xF, false);  // row_mirror
        } else if constexpr (OFF == 16) {
            const auto sw = __builtin_amdgcn_permlane16_swap(a.w[k], a.w[k], false, false);
            a.w[k] = sw[0];
            b.w[k] = sw[1];
        } else {
            static_assert(OFF == 32, "offsets 1 .. 32");
            const auto sw = __builtin_amdgcn_permlane32_swap(a.w[k], a.w[k], false, false);
            a.w[k] = sw[0];
            b.w[k] = sw[1];
        }
    }
    if constexpr (OFF >= 16) return a.v + b.v;
    else return v + b.v;
}

// Sum over aligned groups of G lanes (xor butterfly, offsets 1, 2, ..., G/2);
// every lane of a group ends with the same bits.
template <int G, typename T> FS_DEV T fs_group_sum(T v)
{
    static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "groups of 1 .. 64 lanes");
    if constexpr (G > 1) v = fs_bfly_stage<1>(v);
    if constexpr (G > 2) v = fs_bfly_stage<2>(v);
    if constexpr (G > 4) v = fs_bfly_stage<4>(v);
    if constexpr (G > 8) v = fs_bfly_stage<8>(v);
    if constexpr (G > 16) v = fs_bfly_stage<16>(v);
    if constexpr (G > 32) v = fs_bfly_stage<32>(v);
    return v;
}
