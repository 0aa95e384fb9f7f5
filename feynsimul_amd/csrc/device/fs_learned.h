// fs_learned.h -- the learned bisection proposal (gfx950, BF16 MFMA).
//
// No counterpart in the reference: it replaces the Levy midpoint draw of
// FeynSimul/GPUsrc/kernel.c:338-360 by a conditional Gaussian whose mean and
// log-width come from a small MLP, and extends the accept test of
// kernel.c:376-377 with the exact proposal density ratio, so detailed balance
// holds for any weights:
//
//   level n, node m with ends a = x(m - h), b = x(m + h), h = 2^(S-n),
//   sigma_n^2 = 2^S eps / (2 * 2^n)                      (kernel.c:605-619)
//   f = [a', b', n / S]     (a', b': centred on the atoms' mean for clusters,
//                            times inScale; zero padded to 32 inputs)
//   o = W4 relu(W3 relu(W2 relu(W1 f + b1) + b2) + b3) + b4
//   mu_d = (a_d + b_d)/2 + sigma_n o_d,  s_d = clamp(o_{D+d}, +-SMAX),
//   x_d = mu_d + sigma_n e^{s_d} z_d,    log q = sum_d (-z_d^2/2 - s_d)
//   log alpha = -(K_new - K_old) - eps (V_new - V_old) + log q_old - log q_new
//   accept iff !(exp(log alpha) < u)
//
// Inference (fs_lp_sweep): one launch is one sweep.  A wavefront owns SPW
// consecutive segments of one walker and walks the S levels; the rows of a
// level (its proposals, then the reverse rows that score the old path) go
// through the MLP in tiles of 32 FS_LP_CB rows.  The network is computed transposed,
// C^T[features x rows] = W[features x in] . H[in x rows], with
// v_mfma_f32_32x32x16_bf16: weights are the A operand (read from LDS, where a
// block keeps the active units' part of the image, 82 KB at the full width),
// activations the B operand.  The C
// layout of one layer is the B layout of the next once the next layer's input
// features are permuted to match (the pack kernel does this), so hidden
// activations never leave registers.  Proposals are staged in LDS; the old
// path is read from HBM (cached).  Normals and uniforms come from
// Philox4x32-10 (Salmon et al., SC'11) keyed by the seed and counted by
// (slice or segment, walker, sweep, purpose): no per-chain state, no order
// dependence between rows.
//
// Training (fs_lp_gather, fs_lp_gemm_*, fs_lp_loss, fs_lp_colsum, fs_lp_adam):
// rows (walker, segment start, node) are drawn with Philox; the MLP forward
// and backward are plain BF16 GEMMs on MFMA with FP32 accumulation and
// ReLU/mask epilogues; weight gradients are split over the batch and summed
// in a fixed order; Adam (Kingma & Ba) updates FP32 master weights, and
// fs_lp_pack rebuilds the BF16 image.
//
// Needs: a trait `sys` (N, S, DOF, EPS, V), FS_LP_SPW (segments per
// wavefront), FS_LP_WAVES (wavefronts per block), FS_LP_ATOMS3D, and
// FS_LP_CB (row tile = 32 FS_LP_CB rows: 2 reuses each weight operand for two
// MFMAs, 1 halves the accumulators and activations so two wavefronts fit a
// SIMD).
#pragma once

#include "fs_lane.h"
#include "fs_lp_abi.h"
#include "fs_rng.h"

typedef __attribute__((ext_vector_type(8))) __bf16 fs_bf8;
typedef __attribute__((ext_vector_type(16))) float fs_f16v;

#define FS_LP_KIN 32
// Active hidden width of inference (a multiple of 32, at most FS_LP_HID): the image keeps the
// FS_LP_HID-wide layout, and a network whose units beyond the active width are zero (weights in and
// out, and biases; training leaves them zero: their ReLU masks stop every gradient) is evaluated on
// the active units only.
#ifndef FS_LP_ACT
#define FS_LP_ACT FS_LP_HID
#endif
#define FS_LP_OB (FS_LP_ACT / 32)  // 32-unit output blocks of a hidden layer
#define FS_LP_KS (FS_LP_ACT / 16)  // 16-unit k-steps of the layer after it
// A block's LDS copy of the image holds only the active units: W1p [OB][2 ks], W2p and W3p
// [OB][KS], W4p [KS] (64 lanes x 8 BF16 = 1024 bytes per block entry), then the biases b1, b2, b3
// (FS_LP_ACT floats each) and b4p (32 floats).  At the full width this is the global layout.
#define FS_LP_L_W1 0
#define FS_LP_L_W2 (FS_LP_L_W1 + FS_LP_OB * 2 * 1024)
#define FS_LP_L_W3 (FS_LP_L_W2 + FS_LP_OB * FS_LP_KS * 1024)
#define FS_LP_L_W4 (FS_LP_L_W3 + FS_LP_OB * FS_LP_KS * 1024)
#define FS_LP_L_B1 (FS_LP_L_W4 + FS_LP_KS * 1024)
#define FS_LP_L_B2 (FS_LP_L_B1 + 4 * FS_LP_ACT)
#define FS_LP_L_B3 (FS_LP_L_B2 + 4 * FS_LP_ACT)
#define FS_LP_L_B4 (FS_LP_L_B3 + 4 * FS_LP_ACT)
#define FS_LP_L_BYTES FS_LP_LDS_IMG_BYTES(FS_LP_ACT)
static_assert(FS_LP_L_B4 + 128 == FS_LP_L_BYTES, "LDS image layout");
static_assert(FS_LP_ACT % 32 == 0 && FS_LP_ACT <= FS_LP_HID, "active width");
#define FS_LP_SMAX 3.0f
#define FS_LP_TAG_NORMAL 0x100u
#define FS_LP_TAG_ACCEPT 0x200u
#define FS_LP_TAG_OFFSET 0x300u
#define FS_LP_TAG_TRAIN 0x400u

// ---- packed inference image (offsets in fs_lp_abi.h) ----
// W1p [4 ob][2 ks][64 lanes][8], W2p/W3p [4][8][64][8], W4p [8 ks][64][8] (BF16),
// then FP32 biases b1, b2, b3 (logical order) and b4p (output slot order).

// ---- Philox4x32-10 ----
FS_DEV uint4 fs_philox(uint4 c, fs_u32 k0, fs_u32 k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32 x 32 -> 64-bit product per multiplier (v_mad_u64_u32) gives both halves, and
        // each three-input exclusive or is one v_bitop3_b32 (truth table 0x96)
        const fs_u64 p0 = (fs_u64)0xD2511F53u * c.x;
        const fs_u64 p1 = (fs_u64)0xCD9E8D57u * c.z;
        c = make_uint4(__builtin_amdgcn_bitop3_b32((fs_u32)(p1 >> 32), c.y, k0, 0x96), (fs_u32)p1,
                       __builtin_amdgcn_bitop3_b32((fs_u32)(p0 >> 32), c.w, k1, 0x96), (fs_u32)p0);
    }
    return c;
}

// top 24 bits as a uniform in [0, 1), or (0, 1) with the half offset
FS_DEV float fs_lp_u24(fs_u32 x) { return (float)(x >> 8) * 0x1p-24f; }
FS_DEV float fs_lp_u24h(fs_u32 x) { return ((float)(x >> 8) + 0.5f) * 0x1p-24f; }

// Normal for coordinate d = 4 blk + 2 pair + t of a slice: Box-Muller on the
// Philox words (2 pair, 2 pair + 1); t = 0 takes the cosine, t = 1 the sine.
FS_DEV float fs_lp_normal(uint4 x, int pair, int t)
{
    const fs_u32 a = pair ? x.z : x.x, b = pair ? x.w : x.y;
    const float rho = sqrtf(-2.0f * logf(fs_lp_u24h(a)));
    const float v = fs_lp_u24(b);
    return rho * (t ? fs_sin2pi(v) : fs_cos2pi(v));
}

FS_DEV fs_f16v fs_lp_mfma(fs_bf8 a, fs_bf8 b, fs_f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }

// Copy the active part of the global image (FS_LP_IMG_BYTES, full-width layout) into the block's
// LDS (FS_LP_L_BYTES, the FS_LP_L_* layout), 16 bytes per step; the caller synchronises.
FS_DEV void fs_lp_load_image(const char *img_g, char *lds)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(img_g);
    uint4 *dst = reinterpret_cast<uint4 *>(lds);
    constexpr int W1 = FS_LP_OB * 2 * 64, WH = FS_LP_OB * FS_LP_KS * 64, W4 = FS_LP_KS * 64, BC = FS_LP_ACT / 4;
    for (int i = threadIdx.x; i < FS_LP_L_BYTES / 16; i += blockDim.x) {
        int c = i, s;
        if (FS_LP_ACT == FS_LP_HID || c < W1) {
            s = c;  // W1p [ob][ks]: the active output blocks come first
        } else if ((c -= W1) < 2 * WH) {
            const int l = c >= WH, r = c - l * WH, q = r >> 6;
            s = (l ? FS_LP_OFF_W3 : FS_LP_OFF_W2) / 16 + ((q / FS_LP_KS) * 8 + q % FS_LP_KS) * 64 + (r & 63);
        } else if ((c -= 2 * WH) < W4) {
            s = FS_LP_OFF_W4 / 16 + c;
        } else if ((c -= W4) < 3 * BC) {
            s = FS_LP_OFF_B1 / 16 + (c / BC) * (4 * FS_LP_HID / 16) + c % BC;
        } else {
            s = FS_LP_OFF_B4 / 16 + (c - 3 * BC);
        }
        dst[i] = src[s];
    }
}

// ---- the network on a tile of 32 FS_LP_CB rows held by one wavefront ----
// xin[ks][cb]: layer-1 B operand (inputs ks*16 + 8g + e of row cb*32 + lane%32).
// out[cb][j]: output slot (g, j) of row cb*32 + lane%32; slot j of lane group g
// holds mu-offset (j even) or s (j odd) of coordinate d = 2 (j/2) + g.
FS_DEV void fs_lp_bias_init(const float *b, int ob, int g, fs_f16v (&c)[FS_LP_CB])
{
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float4 v = *reinterpret_cast<const float4 *>(b + ob * 32 + 8 * q + 4 * g);
#pragma unroll
        for (int cb = 0; cb < FS_LP_CB; cb++) {
            c[cb][4 * q + 0] = v.x;
            c[cb][4 * q + 1] = v.y;
            c[cb][4 * q + 2] = v.z;
            c[cb][4 * q + 3] = v.w;
        }
    }
}

// ReLU after the rounding, on packed BF16 pairs: pairs round to BF16 in one
// v_cvt_pk_bf16_f32, and a BF16 with its sign bit set is a negative 16-bit
// integer, so one v_pk_max_i16 against 0 zeroes exactly the negative values
// (-0 included) of two outputs.  The bits equal those of max-then-round, in
// 64 instructions per 32 x 128 block instead of 96 (an FP32 integer max per
// value before the conversion).
FS_DEV void fs_lp_relu_pack(const fs_f16v (&acc)[4][FS_LP_CB], fs_bf8 (&h)[8][FS_LP_CB])
{
#pragma unroll
    for (int ob = 0; ob < FS_LP_OB; ob++)
#pragma unroll
        for (int cb = 0; cb < FS_LP_CB; cb++)
#pragma unroll
            for (int hh = 0; hh < 2; hh++) {
                typedef float fs_f2 __attribute__((ext_vector_type(2)));
                typedef __bf16 fs_b2 __attribute__((ext_vector_type(2)));
                typedef short fs_s2 __attribute__((ext_vector_type(2)));
                typedef unsigned fs_w4 __attribute__((ext_vector_type(4)));
                fs_w4 w;
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    const fs_f2 x = {acc[ob][cb][8 * hh + 2 * p], acc[ob][cb][8 * hh + 2 * p + 1]};
                    const fs_s2 v = __builtin_elementwise_max(__builtin_bit_cast(fs_s2, __builtin_convertvector(x, fs_b2)), (fs_s2)0);
                    w[p] = __builtin_bit_cast(unsigned, v);
                }
                h[2 * ob + hh][cb] = __builtin_bit_cast(fs_bf8, w);
            }
}

FS_DEV void fs_lp_hidden(const char *img, int woff, int boff, const fs_bf8 (&hin)[8][FS_LP_CB],
                         fs_f16v (&acc)[4][FS_LP_CB],
                         int lane)
{
    const int g = lane >> 5;
    const fs_bf8 *W = reinterpret_cast<const fs_bf8 *>(img + woff);
    const float *b = reinterpret_cast<const float *>(img + boff);
#pragma unroll
    for (int ob = 0; ob < FS_LP_OB; ob++) fs_lp_bias_init(b, ob, g, acc[ob]);
#pragma unroll
    for (int ks = 0; ks < FS_LP_KS; ks++)
#pragma unroll
        for (int ob = 0; ob < FS_LP_OB; ob++) {
            const fs_bf8 A = W[(ob * FS_LP_KS + ks) * 64 + lane];
#pragma unroll
            for (int cb = 0; cb < FS_LP_CB; cb++) acc[ob][cb] = fs_lp_mfma(A, hin[ks][cb], acc[ob][cb]);
        }
}

FS_DEV void fs_lp_mlp(const char *img, const fs_bf8 (&xin)[2][FS_LP_CB], fs_f16v (&out)[FS_LP_CB], int lane)
{
    const int g = lane >> 5;
    fs_f16v acc[4][FS_LP_CB];
    {
        const fs_bf8 *W = reinterpret_cast<const fs_bf8 *>(img + FS_LP_L_W1);
        const float *b = reinterpret_cast<const float *>(img + FS_LP_L_B1);
#pragma unroll
        for (int ob = 0; ob < FS_LP_OB; ob++) fs_lp_bias_init(b, ob, g, acc[ob]);
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
#pragma unroll
            for (int ob = 0; ob < FS_LP_OB; ob++) {
                const fs_bf8 A = W[(ob * 2 + ks) * 64 + lane];
#pragma unroll
                for (int cb = 0; cb < FS_LP_CB; cb++) acc[ob][cb] = fs_lp_mfma(A, xin[ks][cb], acc[ob][cb]);
            }
    }
    // With 64-row tiles the layer boundaries are scheduling barriers: without
    // them the weight loads of later layers are hoisted and the live set
    // outgrows the registers.  32-row tiles leave room, and letting the
    // scheduler overlap layers there is 4 % faster (He3, 174 VGPRs).
    fs_bf8 h[8][FS_LP_CB];
    fs_lp_relu_pack(acc, h);
    if constexpr (FS_LP_CB == 2) __builtin_amdgcn_sched_barrier(0);
    fs_lp_hidden(img, FS_LP_L_W2, FS_LP_L_B2, h, acc, lane);
    fs_lp_relu_pack(acc, h);
    if constexpr (FS_LP_CB == 2) __builtin_amdgcn_sched_barrier(0);
    fs_lp_hidden(img, FS_LP_L_W3, FS_LP_L_B3, h, acc, lane);
    fs_lp_relu_pack(acc, h);
    if constexpr (FS_LP_CB == 2) __builtin_amdgcn_sched_barrier(0);
    {
        const fs_bf8 *W = reinterpret_cast<const fs_bf8 *>(img + FS_LP_L_W4);
        const float *b = reinterpret_cast<const float *>(img + FS_LP_L_B4);
        fs_lp_bias_init(b, 0, g, out);
#pragma unroll
        for (int ks = 0; ks < FS_LP_KS; ks++) {
            const fs_bf8 A = W[ks * 64 + lane];
#pragma unroll
            for (int cb = 0; cb < FS_LP_CB; cb++) out[cb] = fs_lp_mfma(A, h[ks][cb], out[cb]);
        }
    }
}

// Feature vector of one row from its ends (registers, compile-time D).
template <int D>
FS_DEV void fs_lp_features(const float (&a)[D], const float (&b)[D], float level, float in_scale, float (&f)[32])
{
    float c[3] = {0.0f, 0.0f, 0.0f};
#if FS_LP_ATOMS3D
#pragma unroll
    for (int d = 0; d < D; d++) c[d % 3] += a[d] + b[d];
#pragma unroll
    for (int k = 0; k < 3; k++) c[k] *= 1.0f / (float)(2 * (D / 3));
#endif
#pragma unroll
    for (int i = 0; i < 32; i++) f[i] = 0.0f;
#pragma unroll
    for (int d = 0; d < D; d++) {
        f[d] = (a[d] - c[d % 3]) * in_scale;
        f[D + d] = (b[d] - c[d % 3]) * in_scale;
    }
    f[2 * D] = level;
}

// layer-1 B operand of a lane from its row's features.  Both halves are
// converted with constant indices and the lane group selects whole vectors, so
// the feature array never needs a runtime index (which would put it in scratch).
FS_DEV void fs_lp_xin(const float (&f)[32], int g, fs_bf8 &k0, fs_bf8 &k1)
{
    fs_bf8 q[4];
#pragma unroll
    for (int h = 0; h < 4; h++)
#pragma unroll
        for (int e = 0; e < 8; e++) q[h][e] = (__bf16)f[8 * h + e];
    k0 = g ? q[1] : q[0];
    k1 = g ? q[3] : q[2];
}



template <class Sys> struct fs_lp_geom {
    static constexpr int N = Sys::N, S = Sys::S, D = Sys::DOF, SEG = 1 << S, NSEG = N / SEG;
    static constexpr int SPW = FS_LP_SPW;
    static constexpr int UPW = NSEG / SPW;                                // units per walker
    static constexpr int STAGE = SPW * (SEG - 1) * D;                     // floats of staged proposals
    static constexpr int WAVE_BYTES = ((STAGE + 2 * SPW) * 4 + 15) / 16 * 16; // + log q accumulators
    static_assert(S >= 1 && S <= 6, "learned proposal: 1 <= S <= 6");
    static_assert(2 * D + 1 <= 32 && D <= 15, "learned proposal: DOF <= 15");
    static_assert(NSEG % SPW == 0, "segments per wavefront must divide N / 2^S");
};

// One wavefront, one unit: segments k0 .. k0 + SPW - 1 of walker w, one sweep.
template <class Sys>
FS_DEV void fs_lp_unit(const fs_lp_args &a, const char *img, float *stage, long w, int k0, int lane)
{
    typedef fs_lp_geom<Sys> G;
    constexpr int N = G::N, S = G::S, D = G::D, SEG = G::SEG, SPW = G::SPW;
    const int g = lane >> 5, c32 = lane & 31;
    float *path = a.paths + w * (long)(N * D);
    // log q(old) - log q(new) per segment, in two halves: [k] collects the proposal rows,
    // [SPW + k] the reverse rows.  Within one instruction the leader lanes of a segment's
    // proposal rows and of its reverse rows then update different words (a shared word
    // would lose one of the two read-modify-writes).
    float *lq = stage + G::STAGE;
    const fs_u32 wl = (fs_u32)w;
    const int r = (int)(fs_philox(make_uint4(0u, wl, a.sweep, FS_LP_TAG_OFFSET), a.key0, a.key1).x & (SEG - 1));
    if (lane < 2 * SPW) lq[lane] = 0.0f;
    auto slice_of = [&](int k, int pos) { return (r + (k0 + k) * SEG + pos) & (N - 1); };
    auto value = [&](int k, int pos, bool old, int d) -> float {
        if (old || pos == 0 || pos == SEG) return path[d * N + slice_of(k, pos)];
        return stage[(k * (SEG - 1) + pos - 1) * D + d];
    };
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();

#pragma unroll 1
    for (int n = 1; n <= S; n++) {
        const int h = SEG >> n, nn = 1 << (n - 1);
        const int NR = SPW * nn, total = n == 1 ? NR : 2 * NR;
        const float sig = sqrtf((float)SEG * (float)Sys::EPS / (2.0f * (float)(1 << n)));
        const float level = (float)n / (float)S;
#pragma unroll 1
        for (int t0 = 0; t0 < total; t0 += 32 * FS_LP_CB) {
            fs_bf8 xin[2][FS_LP_CB];
            // midpoint of the row's ends for this lane's coordinates d = 2q + g, kept
            // across the network instead of reading the ends again
            float mid[FS_LP_CB][8];
            // what the row needs that does not depend on the network, computed before it
            // so that it overlaps the MFMA chain: the proposal's normals and the old
            // coordinates that the density of the old path needs
            float zr[FS_LP_CB][8], xo[FS_LP_CB][8];
#pragma unroll
            for (int cb = 0; cb < FS_LP_CB; cb++) {
                const int R = t0 + cb * 32 + c32;
                float f[32];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    mid[cb][q] = 0.0f;
                    zr[cb][q] = 0.0f;
                    xo[cb][q] = 0.0f;
                }
                if (R < total) {
                    const bool old = R >= NR;
                    const int Rl = old ? R - NR : R, k = Rl >> (n - 1), m = (2 * (Rl & (nn - 1)) + 1) * h;
                    const int s = slice_of(k, m);
                    // the ends' loads are issued first; the proposal's normals (integer and
                    // transcendental work that needs no memory) are computed while they are in flight
                    float ea[D], eb[D];
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        ea[d] = value(k, m - h, old, d);
                        eb[d] = value(k, m + h, old, d);
                    }
                    // Philox block b holds the normals of coordinates 4b .. 4b + 3.  Both lane halves
                    // hold the same row, so they compute two blocks at once, lanes 0-31 block b and
                    // lanes 32-63 block b + 1, and v_permlane32_swap hands each half the other's words.
                    constexpr int NB = (D + 3) / 4;
                    uint4 pb[NB];
                    if (!old) {
#pragma unroll
                        for (int b0 = 0; b0 < NB; b0 += 2) {
                            const bool pair = b0 + 1 < NB;
                            const uint4 x = fs_philox(
                                make_uint4((fs_u32)s, wl, a.sweep, FS_LP_TAG_NORMAL | (fs_u32)(pair ? b0 + g : b0)),
                                a.key0, a.key1);
                            if (pair) {
                                const auto sx = __builtin_amdgcn_permlane32_swap(x.x, x.x, false, false);
                                const auto sy = __builtin_amdgcn_permlane32_swap(x.y, x.y, false, false);
                                const auto sz = __builtin_amdgcn_permlane32_swap(x.z, x.z, false, false);
                                const auto sw = __builtin_amdgcn_permlane32_swap(x.w, x.w, false, false);
                                pb[b0] = make_uint4(sx[0], sy[0], sz[0], sw[0]);
                                pb[b0 + 1] = make_uint4(sx[1], sy[1], sz[1], sw[1]);
                            } else {
                                pb[b0] = x;
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int d = 2 * q + g;
                        if (2 * q < D) {
                            if (d < D) {
                                if (!old) zr[cb][q] = fs_lp_normal(pb[q / 2], q % 2, g);
                                if (old || n == 1) xo[cb][q] = path[d * N + s];
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const float m0 = 2 * q < D ? 0.5f * (ea[2 * q < D ? 2 * q : 0] + eb[2 * q < D ? 2 * q : 0]) : 0.0f;
                        const float m1 =
                            2 * q + 1 < D ? 0.5f * (ea[2 * q + 1 < D ? 2 * q + 1 : 0] + eb[2 * q + 1 < D ? 2 * q + 1 : 0]) : 0.0f;
                        mid[cb][q] = g ? m1 : m0;
                    }
                    fs_lp_features<D>(ea, eb, level, a.inScale, f);
                } else {
#pragma unroll
                    for (int i = 0; i < 32; i++) f[i] = 0.0f;
                }
                fs_lp_xin(f, g, xin[0][cb], xin[1][cb]);
            }
            fs_f16v out[FS_LP_CB];
            fs_lp_mlp(img, xin, out, lane);
#pragma unroll
            for (int cb = 0; cb < FS_LP_CB; cb++) {
                const int R = t0 + cb * 32 + c32;
                const bool valid = R < total, old = R >= NR;
                const int Rl = old ? R - NR : R, k = Rl >> (n - 1), m = (2 * (Rl & (nn - 1)) + 1) * h;
                float contrib = 0.0f;
                if (valid) {
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int d = 2 * q + g;
                        if (d < D) {
                            const float mu = mid[cb][q] + sig * out[cb][2 * q];
                            const float sv = fminf(fmaxf(out[cb][2 * q + 1], -FS_LP_SMAX), FS_LP_SMAX);
                            const float sd = sig * expf(sv);
                            if (!old) {
                                const float z = zr[cb][q];
                                stage[(k * (SEG - 1) + m - 1) * D + d] = mu + sd * z;
                                contrib -= -0.5f * z * z - sv;
                            }
                            if (old || n == 1) {
                                const float zo = (xo[cb][q] - mu) / sd;
                                contrib += -0.5f * zo * zo - sv;
                            }
                        }
                    }
                }
                contrib += __shfl_xor(contrib, 32, 64);
                for (int o = 1; o < nn; o <<= 1) contrib += __shfl_xor(contrib, o, 64);
                if (valid && g == 0 && (c32 & (nn - 1)) == 0) lq[(old ? SPW : 0) + k] += contrib;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }

    // ---- accept or reject each segment (lanes = positions 0 .. SEG - 1) ----
    constexpr int SPP = 64 / SEG;
    const float inv2e = 0.5f / (float)Sys::EPS;
    unsigned nacc = 0;
#pragma unroll 1
    for (int k1 = 0; k1 < SPW; k1 += SPP) {
        const int k = k1 + lane / SEG, p = lane & (SEG - 1);
        const bool valid = k < SPW;
        float dk = 0.0f, dv = 0.0f;
        if (valid) {
            float xn[D], xo[D];
#pragma unroll
            for (int d = 0; d < D; d++) {
                const float n0 = value(k, p, false, d), n1 = value(k, p + 1, false, d);
                const float o0 = value(k, p, true, d), o1 = value(k, p + 1, true, d);
                dk += (n1 - n0) * (n1 - n0) - (o1 - o0) * (o1 - o0);
                xn[d] = n0;
                xo[d] = o0;
            }
            dk *= inv2e;
            if (p > 0) dv = (float)Sys::V(xn) - (float)Sys::V(xo);
        }
#pragma unroll
        for (int o = 1; o < SEG; o <<= 1) {
            dk += __shfl_xor(dk, o, 64);
            dv += __shfl_xor(dv, o, 64);
        }
        if (valid) {
            const float lqk = lq[k] + lq[SPW + k];
            const float logA = -dk - (float)Sys::EPS * dv + lqk;
            if (a.trace && p == 0) {
                float *tr = a.trace + ((long)w * G::NSEG + k0 + k) * 4;
                tr[0] = lqk;
                tr[1] = dk;
                tr[2] = dv;
                tr[3] = logA;
            }
            const float u = fs_lp_u24(fs_philox(make_uint4((fs_u32)(k0 + k), wl, a.sweep, FS_LP_TAG_ACCEPT), a.key0, a.key1).x);
            if (!(expf(logA) < u)) {
                if (p > 0) {
#pragma unroll
                    for (int d = 0; d < D; d++) path[d * N + slice_of(k, p)] = value(k, p, false, d);
                } else {
                    nacc++;
                }
            }
        }
    }
    // every lane 0 of a segment counted its own acceptances
    for (int o = 1; o < 64; o <<= 1) nacc += __shfl_xor(nacc, o, 64);
    if (lane == 0 && nacc) atomicAdd(a.accepts + w, nacc);
}

// ---- whole-path coarse move (FS_LP_COARSE) ----
// One wavefront moves FS_LP_SPW walkers, one MH decision per path.  With L = FS_LPC_L slices per
// fine interval and M = N / L, the coarse nodes are the slices c_j = r + j L (j = 0 .. M - 1,
// node M = node 0) for a random anchor r.  Levels 1 .. LC (M = 2^LC) of the bisection of the whole
// cyclic path (2^S = N) draw the new coarse nodes from the network, exactly as a segment's levels
// are drawn; each interval between neighbouring new nodes is then filled with a free-particle
// (Levy) bridge, built in closed form from a lane prefix sum of Gaussian increments of variance
// eps: x(jL + i) = c'_j + (c'_{j+1} - c'_j) i / L + W_i - (i / L) W_L.  The bridge samples the
// kinetic action of the fine links exactly given the nodes, so the Metropolis-Hastings ratio keeps
// only the coarse part of the kinetic action:
//   log alpha = -eps (V_new - V_old) - (Keff_new - Keff_old) + log q(old nodes) - log q(new nodes),
//   Keff = sum_j |c_{j+1} - c_j|^2 / (2 eps L).
// With a harmonic reference omega^2 (hw = omega^2 / 2 > 0) the intervals take the exact Gaussian bridge
// of the kinetic links plus eps omega^2 |x|^2 / 2 on their interior (an AR(1) chain conditioned on its
// end, built from a decayed lane scan): V is then scored less omega^2 |x|^2 / 2 on every slice, and Keff
// becomes kc sum_j |c_{j+1} - rho^L c_j|^2 with kc = 1 / (2 V_L), V_L the variance of a node given the one
// before under the reference (fs_lpc_dkin).  A reference close to the potential's curvature leaves little
// of the fine slices' action to the accept test.
// Random numbers are counted by (node or lane and block, walker, sweep, purpose) as in the segment
// move.  A lane keeps its new slices in registers until the decision when they fit (M DOF <= 64);
// otherwise the bridge's normals are drawn twice (to score the path, and to write it once accepted).
// One move is two launches: fs_lp_sweep runs the network on the coarse levels (FS_LP_SPW walkers per
// wavefront, sharing row tiles, at the network's register budget) and leaves each walker's new nodes
// and log q ratio in the workspace; fs_lpc_fine, one wavefront per walker and few registers, builds the
// bridges, decides and writes.
#define FS_LP_TAG_FINE 0x500u
#ifndef FS_LPC_L
#define FS_LPC_L 64
#endif
template <class Sys, int WPW_ = FS_LP_SPW> struct fs_lpc_geom {
    static constexpr int N = Sys::N, S = Sys::S, D = Sys::DOF, L = FS_LPC_L, M = N / L, WPW = WPW_;
    static constexpr int NODES = WPW * M * D;                                // floats per node set
    static constexpr int SCR = M * D + 2;                                    // workspace floats per walker
    static constexpr int WAVE_BYTES = ((2 * NODES + 2 * WPW) * 4 + 15) / 16 * 16;
    static_assert(L == 64, "coarse move: one lane per fine slice");
    static_assert((1 << S) == N && M >= 2, "coarse move: N = 64 x 2^LC, LC >= 1");
    static_assert(2 * D + 1 <= 32 && D <= 15, "learned proposal: DOF <= 15");
};

// Normal k (k = interval * D + d) of lane `lane` of one walker's bridge, from Philox block k / 4.
FS_DEV float fs_lpc_fine_normal(uint4 &px, int k, int lane, fs_u32 wl, const fs_lp_args &a)
{
    if ((k & 3) == 0)
        px = fs_philox(make_uint4((fs_u32)((k >> 2) * 64 + lane), wl, a.sweep, FS_LP_TAG_FINE), a.key0, a.key1);
    return fs_lp_normal(px, (k >> 1) & 1, k & 1);
}

// The bridge between neighbouring nodes, per lane (slice k = lane + 1 of the interval).  hw = 0: the Levy
// bridge x_k = a + (b - a) k / 64 + W_k - (k / 64) W_64, W the prefix sums of increments sqrt(eps) z.
// hw = omega^2 / 2 > 0: the harmonic reference bridge, the AR(1) chain W_k = rho W_{k-1} + sqrt(rho eps) z_k
// (rho + 1/rho = 2 + eps^2 omega^2) conditioned on its end, x_k = a rho^k + W_k + g_k (b - a rho^64 - W_64),
// g_k = rho^(64 - k) (1 - rho^(2k)) / (1 - rho^128): the exact Gaussian bridge of the links' kinetic action
// plus eps omega^2 |x|^2 / 2 on the interior.
struct fs_lpc_brg {
    float s, fi, wb1, wb2, pk, g;  // increment scale; per lane: k / 64, broadcast weights, rho^k, g_k
};
template <bool HARM> FS_DEV fs_lpc_brg fs_lpc_bridge_coefs(const fs_lp_args &a, float eps, int lane)
{
    fs_lpc_brg c;
    c.fi = (float)(lane + 1) * (1.0f / (float)FS_LPC_L);
    c.s = HARM ? a.bs : sqrtf(eps);
    c.wb1 = c.wb2 = c.pk = c.g = 1.0f;
    if constexpr (HARM) {
        const float l = a.l2q;
        c.wb1 = exp2f((float)((lane & 15) + 1) * l);
        c.wb2 = exp2f((float)(lane > 31 ? lane - 31 : 0) * l);
        c.pk = a.brg[lane];
        c.g = a.brg[FS_LPC_L + lane];
    }
    return c;
}
// Fine slice of this lane from its normal z and the interval's nodes ca, cb (lane 63: cb).
template <bool HARM>
FS_DEV float fs_lpc_bridge_x(const fs_lpc_brg &c, const fs_lp_args &a, float z, float ca, float cb, int lane)
{
    if constexpr (!HARM) {
        const float v = fs_lane_scan(c.s * z);
        const float tot = fs_readlane(v, 63);
        return lane == 63 ? cb : ca + (cb - ca) * c.fi + v - c.fi * tot;
    }
    const float v = fs_lane_scan_decay(c.s * z, a.q1, a.q2, a.q4, a.q8, c.wb1, c.wb2);
    const float tot = fs_readlane(v, 63);
    return lane == 63 ? cb : ca * c.pk + v + c.g * (cb - ca * a.rhoL - tot);
}
// The potential the fine pass scores: V, less hw |x|^2 under the harmonic reference.
template <class Sys, bool HARM> FS_DEV float fs_lpc_veff(const float *x, float hw)
{
    float v = (float)Sys::V(x);
    if constexpr (HARM) {
        float q = 0.0f;
#pragma unroll
        for (int d = 0; d < Sys::DOF; d++) q += x[d] * x[d];
        v -= hw * q;
    }
    return v;
}

// The new fine slices of interval j on every lane (lane 63: node j + 1).
template <int D, bool HARM>
FS_DEV void fs_lpc_bridge(const float *cn, int j, int M, int lane, uint4 &px, fs_u32 wl, const fs_lp_args &a,
                          const fs_lpc_brg &c, float (&x)[D])
{
#pragma unroll
    for (int d = 0; d < D; d++)
        x[d] = fs_lpc_bridge_x<HARM>(c, a, fs_lpc_fine_normal(px, j * D + d, lane, wl, a), cn[j * D + d],
                                     cn[((j + 1) & (M - 1)) * D + d], lane);
}

// The move's anchor r of walker w: node j of the coarse level is slice r + 64 j.
template <int N> FS_DEV int fs_lpc_anchor(const fs_lp_args &a, long w)
{
    return (int)(fs_philox(make_uint4(0u, (fs_u32)w, a.sweep, FS_LP_TAG_OFFSET), a.key0, a.key1).x & (N - 1));
}

// Levels 1 .. LC of the coarse bisection of WPW walkers w0 .. w0 + WPW - 1: network proposals and reverse
// rows, as in the segment move.  The caller has put each walker's old nodes in cold [WPW][M][D], its anchor
// node in cnew[k][0] and zeros in lq [2 WPW] (wavefront LDS); the new nodes are left in cnew and the parts of
// the log q ratio in lq (proposal rows, then reverse rows).
template <class Sys, int WPW>
FS_DEV void fs_lpc_levels(const fs_lp_args &a, const char *img, float *cnew, const float *cold, float *lq, long w0,
                          int lane)
{
    typedef fs_lpc_geom<Sys, WPW> G;
    constexpr int N = G::N, S = G::S, D = G::D, M = G::M, LC = S - 6;
    const int g = lane >> 5, c32 = lane & 31;
#pragma unroll 1
    for (int n = 1; n <= LC; n++) {
        const int h = M >> n, nn = 1 << (n - 1);
        const int NR = WPW * nn, total = n == 1 ? NR : 2 * NR;
        const float sig = sqrtf((float)N * (float)Sys::EPS / (2.0f * (float)(1 << n)));
        const float level = (float)n / (float)S;
#pragma unroll 1
        for (int t0 = 0; t0 < total; t0 += 32 * FS_LP_CB) {
            fs_bf8 xin[2][FS_LP_CB];
            float mid[FS_LP_CB][8], zr[FS_LP_CB][8], xo[FS_LP_CB][8];
#pragma unroll
            for (int cb = 0; cb < FS_LP_CB; cb++) {
                const int R = t0 + cb * 32 + c32;
                float f[32];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    mid[cb][q] = 0.0f;
                    zr[cb][q] = 0.0f;
                    xo[cb][q] = 0.0f;
                }
                if (R < total) {
                    const bool old = R >= NR;
                    const int Rl = old ? R - NR : R, k = Rl >> (n - 1), j = (2 * (Rl & (nn - 1)) + 1) * h;
                    const float *src = (old ? cold : cnew) + k * M * D;
                    float ea[D], eb[D];
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        ea[d] = src[(j - h) * D + d];
                        eb[d] = src[((j + h) & (M - 1)) * D + d];
                    }
                    constexpr int NB = (D + 3) / 4;
                    uint4 pb[NB];
                    if (!old) {
#pragma unroll
                        for (int b0 = 0; b0 < NB; b0 += 2) {
                            const bool pair = b0 + 1 < NB;
                            const uint4 x = fs_philox(make_uint4((fs_u32)j, (fs_u32)(w0 + k), a.sweep,
                                                                 FS_LP_TAG_NORMAL | (fs_u32)(pair ? b0 + g : b0)),
                                                      a.key0, a.key1);
                            if (pair) {
                                const auto sx = __builtin_amdgcn_permlane32_swap(x.x, x.x, false, false);
                                const auto sy = __builtin_amdgcn_permlane32_swap(x.y, x.y, false, false);
                                const auto sz = __builtin_amdgcn_permlane32_swap(x.z, x.z, false, false);
                                const auto sw = __builtin_amdgcn_permlane32_swap(x.w, x.w, false, false);
                                pb[b0] = make_uint4(sx[0], sy[0], sz[0], sw[0]);
                                pb[b0 + 1] = make_uint4(sx[1], sy[1], sz[1], sw[1]);
                            } else {
                                pb[b0] = x;
                            }
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int d = 2 * q + g;
                        if (2 * q < D && d < D) {
                            if (!old) zr[cb][q] = fs_lp_normal(pb[q / 2], q % 2, g);
                            if (old || n == 1) xo[cb][q] = cold[k * M * D + j * D + d];
                        }
                    }
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const float m0 = 2 * q < D ? 0.5f * (ea[2 * q < D ? 2 * q : 0] + eb[2 * q < D ? 2 * q : 0]) : 0.0f;
                        const float m1 =
                            2 * q + 1 < D ? 0.5f * (ea[2 * q + 1 < D ? 2 * q + 1 : 0] + eb[2 * q + 1 < D ? 2 * q + 1 : 0]) : 0.0f;
                        mid[cb][q] = g ? m1 : m0;
                    }
                    fs_lp_features<D>(ea, eb, level, a.inScale, f);
                } else {
#pragma unroll
                    for (int i = 0; i < 32; i++) f[i] = 0.0f;
                }
                fs_lp_xin(f, g, xin[0][cb], xin[1][cb]);
            }
            fs_f16v out[FS_LP_CB];
            fs_lp_mlp(img, xin, out, lane);
#pragma unroll
            for (int cb = 0; cb < FS_LP_CB; cb++) {
                const int R = t0 + cb * 32 + c32;
                const bool valid = R < total, old = R >= NR;
                const int Rl = old ? R - NR : R, k = Rl >> (n - 1), j = (2 * (Rl & (nn - 1)) + 1) * h;
                float contrib = 0.0f;
                if (valid) {
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int d = 2 * q + g;
                        if (d < D) {
                            const float mu = mid[cb][q] + sig * out[cb][2 * q];
                            const float sv = fminf(fmaxf(out[cb][2 * q + 1], -FS_LP_SMAX), FS_LP_SMAX);
                            const float sd = sig * expf(sv);
                            if (!old) {
                                const float z = zr[cb][q];
                                cnew[k * M * D + j * D + d] = mu + sd * z;
                                contrib -= -0.5f * z * z - sv;
                            }
                            if (old || n == 1) {
                                const float zo = (xo[cb][q] - mu) / sd;
                                contrib += -0.5f * zo * zo - sv;
                            }
                        }
                    }
                }
                // a walker's rows of this level within one 32-row group: nn, or 32 when its rows span groups
                // (levels 7 and deeper, more than 64 coarse nodes); each group adds its own part
                const int span = nn < 32 ? nn : 32;
                contrib += __shfl_xor(contrib, 32, 64);
                for (int o = 1; o < span; o <<= 1) contrib += __shfl_xor(contrib, o, 64);
                if (valid && g == 0 && (c32 & (span - 1)) == 0) lq[(old ? WPW : 0) + k] += contrib;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// Change of the coarse action of walker k of a node set, Keff_new - Keff_old, on every lane (lane i takes
// the links i, i + 64, ...): Keff = sum_j |c_{j+1} - c_j|^2 / (2 eps L) for the Levy bridge, and
// kc sum_j |c_{j+1} - rho^L c_j|^2 under the harmonic reference: minus log of the product of the
// intervals' bridge normalisations, less the reference potential eps hw |c_j|^2 of the nodes, which the
// fine pass scores with every other slice's V - hw |x|^2.
template <class Sys, int WPW>
FS_DEV float fs_lpc_dkin(const fs_lp_args &a, const float *cnew, const float *cold, int k, int lane)
{
    typedef fs_lpc_geom<Sys, WPW> G;
    constexpr int D = G::D, M = G::M;
    const bool harm = a.hw > 0.0f;
    const float pL = harm ? a.rhoL : 1.0f;
    float dk = 0.0f;
    for (int i = lane; i < M * D; i += 64) {
        const int j = i / D, d = i % D, i1 = ((j + 1) & (M - 1)) * D + d;
        const float en = cnew[k * M * D + i1] - pL * cnew[k * M * D + i];
        const float eo = cold[k * M * D + i1] - pL * cold[k * M * D + i];
        dk += en * en - eo * eo;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) dk += __shfl_xor(dk, o, 64);
    return dk * (harm ? a.kc : 0.5f / ((float)Sys::EPS * (float)G::L));
}

template <class Sys>
FS_DEV void fs_lpc_unit(const fs_lp_args &a, const char *img, float *stage, long w0, int lane)
{
    typedef fs_lpc_geom<Sys> G;
    constexpr int N = G::N, D = G::D, L = G::L, M = G::M, WPW = G::WPW;
    float *cnew = stage, *cold = stage + G::NODES, *lq = stage + 2 * G::NODES;
    // old coarse nodes; the anchor (node 0) is also the first new node.  Each node is a 128-byte line of its
    // path: when a walker's nodes fit a lane each (M DOF <= 64), the WPW walkers' loads are all issued
    // before the first is waited for
    if constexpr (M * D <= 64) {
        const int j = lane / D, d = lane % D;
        float v[WPW];
#pragma unroll
        for (int k = 0; k < WPW; k++) {
            const int rk = fs_lpc_anchor<N>(a, w0 + k);
            v[k] = lane < M * D ? a.paths[(w0 + k) * (long)(N * D) + d * N + ((rk + j * L) & (N - 1))] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < WPW; k++) {
            if (lane < M * D) cold[k * M * D + lane] = v[k];
            if (lane < D) cnew[k * M * D + lane] = v[k];
        }
    } else {
#pragma unroll 1
        for (int k = 0; k < WPW; k++) {
            const float *path = a.paths + (w0 + k) * (long)(N * D);
            const int rk = fs_lpc_anchor<N>(a, w0 + k);
            for (int i = lane; i < M * D; i += 64) {
                const int j = i / D, d = i % D;
                const float v = path[d * N + ((rk + j * L) & (N - 1))];
                cold[k * M * D + i] = v;
                if (j == 0) cnew[k * M * D + i] = v;
            }
        }
    }
    if (lane < 2 * WPW) lq[lane] = 0.0f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    fs_lpc_levels<Sys, WPW>(a, img, cnew, cold, lq, w0, lane);
    // the new nodes, the log q ratio and the change of the coarse kinetic action of every walker,
    // for the fine pass
#pragma unroll 1
    for (int k = 0; k < WPW; k++) {
        float *dst = a.nodes + (w0 + k) * (long)G::SCR;
        for (int i = lane; i < M * D; i += 64) dst[i] = cnew[k * M * D + i];
        const float dk = fs_lpc_dkin<Sys, WPW>(a, cnew, cold, k, lane);
        if (lane == 0) {
            dst[M * D] = lq[k] + lq[WPW + k];
            dst[M * D + 1] = dk;
        }
    }
}

// One measurement of walker w's operators from this lane's slice sums part[NOPS], recorded as a
// measure-only launch of the sampler records it (fs_wave.h): the FP32 lane sums summed over the wavefront,
// the slice mean into opMeans (and series), the slice sum added to the FP64 Kahan pair of opTotals.
template <class Sys> FS_DEV void fs_lpc_record(const fs_lp_args &a, long w, const float *part, int lane)
{
#pragma unroll
    for (int k = 0; k < Sys::NOPS; k++) {
        float m = part[k];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) m += __shfl_xor(m, o, 64);
        if (lane == 0) {
            a.opMeans[w * Sys::NOPS + k] = (double)m / (double)Sys::N;
            if (a.opTotals != nullptr) {
                double *tk = a.opTotals + (w * Sys::NOPS + k) * 2;
                const double y = (double)m - tk[1];
                const double t = tk[0] + y;
                tk[1] = (t - tk[0]) - y;
                tk[0] = t;
            }
            if (a.series != nullptr)
                a.series[((long)a.seriesOffset * a.nWalkers + w) * Sys::NOPS + k] = m / (float)Sys::N;
        }
    }
}
// The kernel's operators can be measured by the move when it has operators and no correlator, grid bins or
// observable histogram.
template <class Sys> constexpr bool fs_lpc_measurable()
{
    return Sys::NOPS > 0 && Sys::NCORR == 0 && !Sys::BINS && Sys::NHOBS == 0;
}

// The fine pass of one walker (one wavefront) given its new coarse nodes cn [M][D] (LDS), its log q ratio lqk
// and coarse kinetic change dk: bridges between the nodes, log alpha, the decision, and the write of an
// accepted path into path [D][N] (global memory, or the walker's LDS copy).  Returns the decision, the same on
// every lane.  HARM: the harmonic reference bridge (hw > 0), else the Levy bridge.  MEAS: the operators of
// the resulting path are recorded (fs_lpc_record), from sums over the old and the new slices.
template <class Sys, bool HARM, bool MEAS>
FS_DEV bool fs_lpc_fine_path_t(const fs_lp_args &a, const float *cn, float lqk, float dk, float *path, long w,
                               int lane)
{
    typedef fs_lpc_geom<Sys> G;
    constexpr int N = G::N, D = G::D, L = G::L, M = G::M, NO = Sys::NOPS > 0 ? Sys::NOPS : 1;
    float po[NO], pn[NO];
#pragma unroll
    for (int k = 0; k < NO; k++) po[k] = pn[k] = 0.0f;
    const fs_u32 wl = (fs_u32)w;
    const int rk = fs_lpc_anchor<N>(a, w);
    const float eps = (float)Sys::EPS;
    const float u = fs_lp_u24(fs_philox(make_uint4(0u, wl, a.sweep, FS_LP_TAG_ACCEPT), a.key0, a.key1).x);
    const fs_lpc_brg bc = fs_lpc_bridge_coefs<HARM>(a, eps, lane);
    const float hw = a.hw;
    float dv = 0.0f;
    uint4 px = make_uint4(0u, 0u, 0u, 0u);
    // the old slices of the next interval are loaded while this one is computed
    float xnext[D];
#pragma unroll
    for (int d = 0; d < D; d++) xnext[d] = path[d * N + ((rk + lane + 1) & (N - 1))];
    auto decide = [&](float dvs) {
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) dvs += __shfl_xor(dvs, o, 64);
        const float logA = -dk - eps * dvs + lqk;
        if (a.trace && lane == 0) {
            float *tr = a.trace + w * 4;
            tr[0] = lqk;
            tr[1] = dk;
            tr[2] = dvs;
            tr[3] = logA;
        }
        return !(expf(logA) < u);
    };
    if constexpr (M * D <= 64) {
        // a lane's new slices fit in registers: kept there until the decision.  The old slices only
        // enter through -V; the bridges are drawn one Philox block (four normals) at a time.
#pragma unroll 4
        for (int j = 0; j < M; j++) {
            float xold[D];
#pragma unroll
            for (int d = 0; d < D; d++) xold[d] = path[d * N + ((rk + j * L + lane + 1) & (N - 1))];
            dv -= fs_lpc_veff<Sys, HARM>(xold, hw);
            if constexpr (MEAS) Sys::ops(xold, po);
        }
        float xs[M][D];
#pragma unroll
        for (int b = 0; b < (M * D + 3) / 4; b++) {
            const uint4 x =
                fs_philox(make_uint4((fs_u32)(b * 64 + lane), wl, a.sweep, FS_LP_TAG_FINE), a.key0, a.key1);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k = 4 * b + t, j = k / D, d = k % D;
                if (k < M * D) {
                    xs[j][d] = fs_lpc_bridge_x<HARM>(bc, a, fs_lp_normal(x, t >> 1, t & 1), cn[j * D + d],
                                                     cn[((j + 1) & (M - 1)) * D + d], lane);
                    // a slice is scored as soon as its last coordinate is drawn (slices in order, as before)
                    if (d == D - 1) dv += fs_lpc_veff<Sys, HARM>(xs[j], hw);
                }
            }
        }
        const bool acc = decide(dv);
        if (acc) {
            // an accepted path's operators are summed while it is written
#pragma unroll
            for (int j = 0; j < M; j++) {
                const int s = (rk + j * L + lane + 1) & (N - 1);
#pragma unroll
                for (int d = 0; d < D; d++) path[d * N + s] = xs[j][d];
                if constexpr (MEAS) {
                    // the slice through an empty asm, so that the operators are computed here rather than
                    // from products kept live since the potential was scored
                    float xt[D];
#pragma unroll
                    for (int d = 0; d < D; d++) {
                        xt[d] = xs[j][d];
                        asm volatile("" : "+v"(xt[d]));
                    }
                    Sys::ops(xt, pn);
                }
            }
        }
        if constexpr (MEAS) {
            float pm[NO];
#pragma unroll
            for (int k = 0; k < NO; k++) pm[k] = acc ? pn[k] : po[k];
            fs_lpc_record<Sys>(a, w, pm, lane);
        }
        return acc;
    } else {
        // otherwise the bridges are drawn again (the same normals) to write an accepted path
#pragma unroll 2
        for (int j = 0; j < M; j++) {
            float xn[D], xold[D];
#pragma unroll
            for (int d = 0; d < D; d++) {
                xold[d] = xnext[d];
                if (j + 1 < M) xnext[d] = path[d * N + ((rk + (j + 1) * L + lane + 1) & (N - 1))];
            }
            fs_lpc_bridge<D, HARM>(cn, j, M, lane, px, wl, a, bc, xn);
            dv += fs_lpc_veff<Sys, HARM>(xn, hw) - fs_lpc_veff<Sys, HARM>(xold, hw);
            if constexpr (MEAS) {
                Sys::ops(xold, po);
                Sys::ops(xn, pn);
            }
        }
        const bool acc = decide(dv);
        if constexpr (MEAS) {
            float pm[NO];
#pragma unroll
            for (int k = 0; k < NO; k++) pm[k] = acc ? pn[k] : po[k];
            fs_lpc_record<Sys>(a, w, pm, lane);
        }
        if (acc) {
            px = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll 2
            for (int j = 0; j < M; j++) {
                float xn[D];
                fs_lpc_bridge<D, HARM>(cn, j, M, lane, px, wl, a, bc, xn);
                const int s = (rk + j * L + lane + 1) & (N - 1);
#pragma unroll
                for (int d = 0; d < D; d++) path[d * N + s] = xn[d];
            }
            return true;
        }
        return false;
    }
}

template <class Sys>
FS_DEV bool fs_lpc_fine_path(const fs_lp_args &a, const float *cn, float lqk, float dk, float *path, long w, int lane)
{
    if constexpr (fs_lpc_measurable<Sys>()) {
        if (a.measure)
            return a.hw > 0.0f ? fs_lpc_fine_path_t<Sys, true, true>(a, cn, lqk, dk, path, w, lane)
                               : fs_lpc_fine_path_t<Sys, false, true>(a, cn, lqk, dk, path, w, lane);
    }
    return a.hw > 0.0f ? fs_lpc_fine_path_t<Sys, true, false>(a, cn, lqk, dk, path, w, lane)
                       : fs_lpc_fine_path_t<Sys, false, false>(a, cn, lqk, dk, path, w, lane);
}

// The fine pass of one walker from the network kernel's workspace (new nodes, log q ratio, coarse kinetic
// change); cn: LDS for the walker's new nodes.
template <class Sys>
FS_DEV void fs_lpc_fine_walker(const fs_lp_args &a, float *cn, long w, int lane)
{
    typedef fs_lpc_geom<Sys> G;
    constexpr int N = G::N, D = G::D, M = G::M;
    const float *src = a.nodes + w * (long)G::SCR;
    for (int i = lane; i < M * D; i += 64) cn[i] = src[i];
    const float lqk = src[M * D], dk = src[M * D + 1];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (fs_lpc_fine_path<Sys>(a, cn, lqk, dk, a.paths + w * (long)(N * D), w, lane) && lane == 0) a.accepts[w] += 1u;
}

// ---- the whole-path move fused for small paths (N DOF <= FS_LPC_FUSED_MAX) ----
// One wavefront per walker holds the walker's path in LDS for the whole launch and makes nMoves moves in
// a row: the coarse nodes from the LDS path, the network's levels (rows of one walker per tile), the fine
// pass on the LDS path, then the next move; the path returns to HBM at the end.  The random numbers are
// counted exactly as in the two-kernel move, and every row's network output and every sum is formed in the
// same order, so the result equals that of nMoves two-kernel moves.  For small paths the two-kernel move is
// bound by launch gaps and by one wavefront per SIMD (four walkers per wavefront).
#define FS_LPC_FUSED_MAX 1024
#define FS_LPC_FUSED_WAVES 4
template <class Sys> struct fs_lpc_fused_geom {
    typedef fs_lpc_geom<Sys, 1> G;
    static constexpr int PATH = Sys::N * Sys::DOF;
    static constexpr int WAVE_BYTES = ((PATH + 2 * G::NODES + 2) * 4 + 15) / 16 * 16;
};

template <class Sys>
FS_DEV void fs_lpc_fused_walker(const fs_lp_args &a, const char *img, float *wb, long w, int nMoves, int lane)
{
    typedef fs_lpc_geom<Sys, 1> G;
    constexpr int N = G::N, D = G::D, L = G::L, M = G::M, PATH = N * D;
    float *path = wb, *cnew = wb + PATH, *cold = cnew + G::NODES, *lq = cold + G::NODES;
    float *gp = a.paths + w * (long)PATH;
    for (int i = lane; i < PATH; i += 64) path[i] = gp[i];
    fs_u32 nacc = 0;
    fs_lp_args am = a;
    am.measure = 0;
#pragma unroll 1
    for (int m = 0; m < nMoves; m++) {
        am.sweep = a.sweep + (unsigned)m;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int rk = fs_lpc_anchor<N>(am, w);
        for (int i = lane; i < M * D; i += 64) {
            const int j = i / D, d = i % D;
            const float v = path[d * N + ((rk + j * L) & (N - 1))];
            cold[i] = v;
            if (j == 0) cnew[i] = v;
        }
        if (lane < 2) lq[lane] = 0.0f;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        fs_lpc_levels<Sys, 1>(am, img, cnew, cold, lq, w, lane);
        const float dk = fs_lpc_dkin<Sys, 1>(am, cnew, cold, 0, lane);
        const float lqk = lq[0] + lq[1];
        if (fs_lpc_fine_path<Sys>(am, cnew, lqk, dk, path, w, lane)) nacc++;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if constexpr (fs_lpc_measurable<Sys>()) {
        if (a.measure) {
            // the operators of the final path, slice by slice from LDS
            float part[Sys::NOPS];
#pragma unroll
            for (int k = 0; k < Sys::NOPS; k++) part[k] = 0.0f;
            for (int t = lane; t < N; t += 64) {
                float x[D];
#pragma unroll
                for (int d = 0; d < D; d++) x[d] = path[d * N + t];
                Sys::ops(x, part);
            }
            fs_lpc_record<Sys>(a, w, part, lane);
        }
    }
    for (int i = lane; i < PATH; i += 64) gp[i] = path[i];
    if (lane == 0) a.accepts[w] += nacc;
}

// ---- pack: FP32 master parameters (logical) -> inference image and training matrices ----
// Logical parameter order: W1 [128][2D+1], b1 [128], W2 [128][128], b2, W3 [128][128], b3,
// W4 [2D][128], b4 [2D].  Training matrices (BF16): W1b [128][32], W2b, W3b [128][128],
// W4b [32][128]; then FP32 biases b1..b3 [128] and b4 [32] (zero padded).
template <int D> struct fs_lp_params {
    static constexpr int NIN = 2 * D + 1, NOUT = 2 * D, H = FS_LP_HID;
    static constexpr int W1 = 0, B1 = W1 + H * NIN, W2 = B1 + H, B2 = W2 + H * H, W3 = B2 + H, B3 = W3 + H * H,
                         W4 = B3 + H, B4 = W4 + NOUT * H, COUNT = B4 + NOUT;
    // training matrices, element offsets in the BF16 part
    static constexpr int T_W1 = 0, T_W2 = T_W1 + H * 32, T_W3 = T_W2 + H * H, T_W4 = T_W3 + H * H,
                         T_BF16 = T_W4 + 32 * H;
    static constexpr int T_BIAS_FLOATS = 3 * H + 32;
};

// logical output row of output slot f (-1: unused)
template <int D> FS_DEV int fs_lp_out_row(int f)
{
    const int g = (f >> 2) & 1, j = 4 * (f >> 3) + (f & 3), d = 2 * (j >> 1) + g;
    return d < D ? ((j & 1) ? D + d : d) : -1;
}

// input feature read by element e of lane group g at k-step ks of a hidden layer
FS_DEV int fs_lp_hidden_in(int ks, int g, int e) { return (ks >> 1) * 32 + 16 * (ks & 1) + 8 * (e >> 2) + 4 * g + (e & 3); }

template <int D> FS_DEV void fs_lp_pack_body(const float *P, char *img, __bf16 *tw, float *tb, int i)
{
    typedef fs_lp_params<D> Q;
    // inference image: 40960 BF16 weights, one per thread
    if (i < 40960) {
        __bf16 *wimg = reinterpret_cast<__bf16 *>(img);
        float v = 0.0f;
        if (i < 4096) {  // W1p [ob][ks][lane][e]
            const int e = i & 7, lane = (i >> 3) & 63, ks = (i >> 9) & 1, ob = i >> 10;
            const int o = ob * 32 + (lane & 31), in = ks * 16 + (lane >> 5) * 8 + e;
            v = in < Q::NIN ? P[Q::W1 + o * Q::NIN + in] : 0.0f;
        } else if (i < 36864) {  // W2p, W3p
            const int L = i < 20480 ? 0 : 1, j = i - (L ? 20480 : 4096);
            const int e = j & 7, lane = (j >> 3) & 63, ks = (j >> 9) & 7, ob = j >> 12;
            const int o = ob * 32 + (lane & 31), in = fs_lp_hidden_in(ks, lane >> 5, e);
            v = P[(L ? Q::W3 : Q::W2) + o * Q::H + in];
        } else {  // W4p [ks][lane][e]
            const int j = i - 36864, e = j & 7, lane = (j >> 3) & 63, ks = j >> 9;
            const int row = fs_lp_out_row<D>(lane & 31), in = fs_lp_hidden_in(ks, lane >> 5, e);
            v = row >= 0 ? P[Q::W4 + row * Q::H + in] : 0.0f;
        }
        wimg[i] = (__bf16)v;
    }
    if (i < 416) {
        float *bimg = reinterpret_cast<float *>(img + FS_LP_OFF_B1);
        float v;
        if (i < 384) {
            v = P[(i < 128 ? Q::B1 : i < 256 ? Q::B2 : Q::B3) + (i & 127)];
        } else {
            const int row = fs_lp_out_row<D>(i - 384);
            v = row >= 0 ? P[Q::B4 + row] : 0.0f;
        }
        bimg[i] = v;
        // training biases (logical, zero padded)
        tb[i] = i < 384 ? v : ((i - 384) < Q::NOUT ? P[Q::B4 + (i - 384)] : 0.0f);
    }
    // training matrices
    if (i < Q::T_BF16) {
        float v = 0.0f;
        if (i < Q::T_W2) {
            const int o = i / 32, in = i % 32;
            v = in < Q::NIN ? P[Q::W1 + o * Q::NIN + in] : 0.0f;
        } else if (i < Q::T_W3) {
            v = P[Q::W2 + (i - Q::T_W2)];
        } else if (i < Q::T_W4) {
            v = P[Q::W3 + (i - Q::T_W3)];
        } else {
            const int j = i - Q::T_W4, o = j / Q::H, in = j % Q::H;
            v = o < Q::NOUT ? P[Q::W4 + o * Q::H + in] : 0.0f;
        }
        tw[i] = (__bf16)v;
    }
}

// ---- training: rows, GEMM, loss, column sums, Adam ----


template <class Sys> FS_DEV void fs_lp_gather_row(const fs_lp_train_args &a, int b)
{
    constexpr int N = Sys::N, S = Sys::S, D = Sys::DOF, SEG = 1 << S;
    const uint4 x = fs_philox(make_uint4((fs_u32)b, a.stepLo, a.stepHi, FS_LP_TAG_TRAIN), a.key0, a.key1);
    const long w = (long)(((unsigned long long)x.x * (unsigned long long)a.nWalkers) >> 32);
    const int s0 = (int)(x.y & (N - 1));
#if FS_LP_COARSE
    // the whole-path move learns only the coarse nodes, every FS_LPC_L slices
    const int p = FS_LPC_L * (1 + (int)(((unsigned long long)x.z * (unsigned long long)(SEG / FS_LPC_L - 1)) >> 32));
#else
    const int p = 1 + (int)(((unsigned long long)x.z * (unsigned long long)(SEG - 1)) >> 32);
#endif
    const int h = p & -p;
    const int n = S - (31 - __builtin_clz((unsigned)h));
    const float *path = a.paths + w * (long)(N * D);
    float ea[D], eb[D], f[32];
    float *T = a.T + (long)b * (3 * D + 1);
#pragma unroll
    for (int d = 0; d < D; d++) {
        ea[d] = path[d * N + ((s0 + p - h) & (N - 1))];
        eb[d] = path[d * N + ((s0 + p + h) & (N - 1))];
        T[d] = ea[d];
        T[D + d] = eb[d];
        T[2 * D + d] = path[d * N + ((s0 + p) & (N - 1))];
    }
    T[3 * D] = sqrtf((float)SEG * (float)Sys::EPS / (2.0f * (float)(1 << n)));
    fs_lp_features<D>(ea, eb, (float)n / (float)S, a.inScale, f);
    __bf16 *X0 = reinterpret_cast<__bf16 *>(a.X0);
#pragma unroll
    for (int i = 0; i < 32; i++) X0[(long)b * 32 + i] = (__bf16)f[i];
}

// C[M x N] = A[M x K] B[K x N] on one 32 x 32 tile per wavefront; K range split
// over blockIdx.z (partials at C + z M N).  AK / BK: K is the contiguous index.
// Epilogues: 0 FP32 (+ bias[n]), 1 BF16 relu(acc + bias[n]), 2 BF16 acc * (H[m][n] > 0).


template <bool AK, bool BK> FS_DEV void fs_lp_gemm_body(const fs_lp_gemm_args &p)
{
    const int lane = threadIdx.x & 63, g = lane >> 5, c32 = lane & 31;
    const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    const int kc = p.K / gridDim.z, kb = blockIdx.z * kc;
    fs_f16v acc;
#pragma unroll
    for (int j = 0; j < 16; j++) acc[j] = 0.0f;
    const __bf16 *PA = reinterpret_cast<const __bf16 *>(p.A), *PB = reinterpret_cast<const __bf16 *>(p.B);
    const __bf16 *PH = reinterpret_cast<const __bf16 *>(p.H);
    const long am = (long)(m0 + c32) * p.lda_m, bn = (long)(n0 + c32) * p.ldb_n;
    for (int k = kb; k < kb + kc; k += 16) {
        const int kk = k + 8 * g;
        fs_bf8 A, B;
        if (AK) {
            A = *reinterpret_cast<const fs_bf8 *>(PA + am + kk);
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) A[e] = PA[am + (long)(kk + e) * p.lda_k];
        }
        if (BK) {
            B = *reinterpret_cast<const fs_bf8 *>(PB + bn + kk);
        } else {
#pragma unroll
            for (int e = 0; e < 8; e++) B[e] = PB[bn + (long)(kk + e) * p.ldb_k];
        }
        acc = fs_lp_mfma(A, B, acc);
    }
    const int n = n0 + c32;
    const float bias = p.bias ? p.bias[n] : 0.0f;
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int m = m0 + 8 * (j >> 2) + 4 * g + (j & 3);
        const long ci = (long)m * p.ldc + n;
        if (p.mode == 0) {
            reinterpret_cast<float *>(p.C)[(long)blockIdx.z * p.zstride + ci] = acc[j] + bias;
        } else if (p.mode == 1) {
            reinterpret_cast<__bf16 *>(p.C)[ci] = (__bf16)fmaxf(acc[j] + bias, 0.0f);
        } else {
            const bool on = (float)PH[ci] > 0.0f;
            reinterpret_cast<__bf16 *>(p.C)[ci] = (__bf16)(on ? acc[j] : 0.0f);
        }
    }
}

// Loss and output gradient of row b: O [B][32] FP32 logical outputs -> dO [B][32] BF16.
template <int D> FS_DEV float fs_lp_loss_row(const float *O, const float *T, __bf16 *dO, int b, float invB)
{
    const float *o = O + (long)b * 32, *t = T + (long)b * (3 * D + 1);
    const float sig = t[3 * D];
    float loss = 0.0f;
    __bf16 *g = dO + (long)b * 32;
#pragma unroll
    for (int d = 0; d < D; d++) {
        const float mid = 0.5f * (t[d] + t[D + d]);
        const float mu = mid + sig * o[d];
        const float sraw = o[D + d];
        const float sv = fminf(fmaxf(sraw, -FS_LP_SMAX), FS_LP_SMAX);
        const float z = (t[2 * D + d] - mu) / (sig * expf(sv));
        loss += 0.5f * z * z + sv;
        g[d] = (__bf16)(-z * expf(-sv) * invB);
        g[D + d] = (__bf16)((sraw > -FS_LP_SMAX && sraw < FS_LP_SMAX) ? (1.0f - z * z) * invB : 0.0f);
    }
#pragma unroll
    for (int i = 2 * D; i < 32; i++) g[i] = (__bf16)0.0f;
    return loss;
}

// Adam on one logical parameter.  grad = sum of the weight-gradient partials
// (or the bias column sum); step counts from 1.


template <int D> FS_DEV void fs_lp_adam_body(const fs_lp_adam_args &a, int i)
{
    typedef fs_lp_params<D> Q;
    if (i >= Q::COUNT) return;
    const float *src = nullptr;
    long stride = 0, idx = 0;
    float grad = 0.0f;
    // a weight into or out of a hidden unit beyond the active width has exactly zero gradient (the
    // unit's ReLU mask is zero); training does not compute those partials, so they are not read
    constexpr int A = FS_LP_ACT;
    if (i < Q::B1) { if ((i / Q::NIN) < A) { src = a.dW1; stride = 128 * 32; idx = (i / Q::NIN) * 32 + i % Q::NIN; } }
    else if (i < Q::W2) { idx = i - Q::B1; if (idx < A) { src = a.db; stride = FS_LP_NBIAS; } }
    else if (i < Q::B2) { idx = i - Q::W2; if (idx / 128 < A && idx % 128 < A) { src = a.dW2; stride = 128 * 128; } }
    else if (i < Q::W3) { idx = i - Q::B2; if (idx < A) { src = a.db; stride = FS_LP_NBIAS; idx += 128; } }
    else if (i < Q::B3) { idx = i - Q::W3; if (idx / 128 < A && idx % 128 < A) { src = a.dW3; stride = 128 * 128; } }
    else if (i < Q::W4) { idx = i - Q::B3; if (idx < A) { src = a.db; stride = FS_LP_NBIAS; idx += 256; } }
    else if (i < Q::B4) { idx = i - Q::W4; if (idx % 128 < A) { src = a.dW4; stride = 32 * 128; } }
    else { src = a.db; stride = FS_LP_NBIAS; idx = 384 + i - Q::B4; }
    // the partials are added in chunk order; the unrolled loop keeps sixteen loads in flight
    if (src) {
#pragma unroll 16
        for (int z = 0; z < a.splits; z++) grad += src[z * stride + idx];
    }
    if (a.gradOut) a.gradOut[i] = grad;
    if (!a.apply) return;
    float m = a.m[i] * a.b1 + (1.0f - a.b1) * grad;
    float v = a.v[i] * a.b2 + (1.0f - a.b2) * grad * grad;
    a.m[i] = m;
    a.v[i] = v;
    a.P[i] -= a.lr * (m / a.c1) / (sqrtf(v / a.c2) + a.aeps);
}

// ---- kernels of one module (instantiated for the trait `sys`) ----
#ifndef FS_LP_NO_KERNELS
#if FS_LP_COARSE
// One launch = one whole-path coarse move of every walker (FS_LP_SPW walkers per wavefront).  The
// module allows up to FS_LP_WAVES wavefronts per block; the launch takes blockDim.x / 64 of them,
// as many as fill a CU best beside the (width-dependent) LDS image.
extern "C" __global__ __launch_bounds__(64 * FS_LP_WAVES) void fs_lp_sweep(fs_lp_args a)
{
    typedef fs_lpc_geom<sys> G;
    extern __shared__ __attribute__((aligned(16))) char fs_lp_smem[];
    fs_lp_load_image(a.img, fs_lp_smem);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, waves = blockDim.x >> 6;
    float *stage = reinterpret_cast<float *>(fs_lp_smem + FS_LP_L_BYTES + wave * G::WAVE_BYTES);
    for (long long u = (long long)blockIdx.x * waves + wave; u < a.nUnits; u += (long long)gridDim.x * waves)
        fs_lpc_unit<sys>(a, fs_lp_smem, stage, (long)(a.w0 + u * G::WPW), lane);
}
// The fine pass of the whole-path move: one wavefront per walker, FS_LPC_FINE_WAVES per block.
#define FS_LPC_FINE_WAVES 4
extern "C" __global__ __launch_bounds__(64 * FS_LPC_FINE_WAVES) void fs_lpc_fine(fs_lp_args a)
{
    typedef fs_lpc_geom<sys> G;
    __shared__ float nodes[FS_LPC_FINE_WAVES][G::M * G::D];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long w = (long)blockIdx.x * FS_LPC_FINE_WAVES + wave;
    if (w < a.nMove) fs_lpc_fine_walker<sys>(a, nodes[wave], (long)(a.w0 + w), lane);
}
// nMoves whole-path moves per launch for small paths: one wavefront per walker, FS_LPC_FUSED_WAVES per
// block sharing the LDS image of the network's active units.
extern "C" __global__ __launch_bounds__(64 * FS_LPC_FUSED_WAVES) void fs_lpc_fused(fs_lp_args a, int nMoves)
{
    if constexpr (sys::N * sys::DOF <= FS_LPC_FUSED_MAX) {
        extern __shared__ __attribute__((aligned(16))) char fs_lp_smem[];
        fs_lp_load_image(a.img, fs_lp_smem);
        __syncthreads();
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const long w = (long)blockIdx.x * FS_LPC_FUSED_WAVES + wave;
        if (w >= a.nMove) return;  // whole wavefronts; no block barrier follows
        float *wb = reinterpret_cast<float *>(fs_lp_smem + FS_LP_L_BYTES + wave * fs_lpc_fused_geom<sys>::WAVE_BYTES);
        fs_lpc_fused_walker<sys>(a, fs_lp_smem, wb, (long)(a.w0 + w), nMoves, lane);
    }
}
#else
extern "C" __global__ __launch_bounds__(64 * FS_LP_WAVES) void fs_lp_sweep(fs_lp_args a)
{
    typedef fs_lp_geom<sys> G;
    extern __shared__ __attribute__((aligned(16))) char fs_lp_smem[];
    fs_lp_load_image(a.img, fs_lp_smem);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float *stage = reinterpret_cast<float *>(fs_lp_smem + FS_LP_L_BYTES + wave * G::WAVE_BYTES);
    for (long long u = (long long)blockIdx.x * FS_LP_WAVES + wave; u < a.nUnits;
         u += (long long)gridDim.x * FS_LP_WAVES)
        fs_lp_unit<sys>(a, fs_lp_smem, stage, (long)(u / G::UPW), (int)(u % G::UPW) * G::SPW, lane);
}
#endif

// The network alone on given feature rows (FP32 [rows][32]) -> logical outputs [rows][32].
extern "C" __global__ __launch_bounds__(64 * FS_LP_WAVES) void fs_lp_eval(const char *img_g, const float *X,
                                                                          float *O, int rows)
{
    extern __shared__ __attribute__((aligned(16))) char fs_lp_smem[];
    fs_lp_load_image(img_g, fs_lp_smem);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 5, c32 = lane & 31;
#pragma unroll 1
    for (int t0 = (blockIdx.x * FS_LP_WAVES + wave) * 32 * FS_LP_CB; t0 < rows;
         t0 += gridDim.x * FS_LP_WAVES * 32 * FS_LP_CB) {
        fs_bf8 xin[2][FS_LP_CB];
#pragma unroll
        for (int cb = 0; cb < FS_LP_CB; cb++) {
            const int R = t0 + cb * 32 + c32;
            float f[32];
#pragma unroll
            for (int i = 0; i < 32; i++) f[i] = R < rows ? X[(long)R * 32 + i] : 0.0f;
            fs_lp_xin(f, g, xin[0][cb], xin[1][cb]);
        }
        fs_f16v out[FS_LP_CB];
        fs_lp_mlp(fs_lp_smem, xin, out, lane);
#pragma unroll
        for (int cb = 0; cb < FS_LP_CB; cb++) {
            const int R = t0 + cb * 32 + c32;
            if (R < rows) {
#pragma unroll
                for (int j = 0; j < 16; j++) {
                    const int row = fs_lp_out_row<sys::DOF>(8 * (j >> 2) + 4 * g + (j & 3));
                    if (row >= 0) O[(long)R * 32 + row] = out[cb][j];
                }
            }
        }
    }
}

extern "C" __global__ __launch_bounds__(256) void fs_lp_pack(const float *P, char *img, __bf16 *tw, float *tb)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    fs_lp_pack_body<sys::DOF>(P, img, tw, tb, i);
}

extern "C" __global__ __launch_bounds__(256) void fs_lp_gather(fs_lp_train_args a)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b < a.batch) fs_lp_gather_row<sys>(a, b);
}

extern "C" __global__ __launch_bounds__(64) void fs_lp_gemm_kk(fs_lp_gemm_args p) { fs_lp_gemm_body<true, true>(p); }
extern "C" __global__ __launch_bounds__(64) void fs_lp_gemm_ks(fs_lp_gemm_args p) { fs_lp_gemm_body<true, false>(p); }
extern "C" __global__ __launch_bounds__(64) void fs_lp_gemm_sk(fs_lp_gemm_args p) { fs_lp_gemm_body<false, true>(p); }
extern "C" __global__ __launch_bounds__(64) void fs_lp_gemm_ss(fs_lp_gemm_args p) { fs_lp_gemm_body<false, false>(p); }

// loss partial per block (256 rows) -> lossPart[blockIdx.x]
extern "C" __global__ __launch_bounds__(256) void fs_lp_loss(const float *O, const float *T, __bf16 *dO,
                                                            float *lossPart, int batch)
{
    __shared__ float red[256];
    const int b = blockIdx.x * 256 + threadIdx.x;
    const float l = b < batch ? fs_lp_loss_row<sys::DOF>(O, T, dO, b, 1.0f / (float)batch) : 0.0f;
    red[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) lossPart[blockIdx.x] = red[0];
}

// Bias gradients: column sums of dO (-> b4), dA3 (b3), dA2 (b2), dA1 (b1), one partial per chunk of
// FS_LP_SPLIT_ROWS rows (blockIdx.y), written as db [chunk][b1 128][b2 128][b3 128][b4 32] and summed
// over the chunks in a fixed order by Adam.  A block sums 32 columns (blockIdx.x) of its chunk: lane
// t & 31 takes one column and t >> 5 every eighth row, so a wavefront reads two rows' 64 contiguous
// bytes per load.  Training computes the hidden layers on the FS_LP_ACT active units only; the other
// columns of dA1..dA3 are not written, and their bias gradients are exactly zero.
extern "C" __global__ __launch_bounds__(256) void fs_lp_colsum(const __bf16 *dA1, const __bf16 *dA2,
                                                              const __bf16 *dA3, const __bf16 *dO, float *db)
{
    __shared__ float red[8][33];
    const int c = blockIdx.x * 32 + (threadIdx.x & 31), r0 = threadIdx.x >> 5;  // c: 0 .. 415
    const __bf16 *src;
    int ld, col;
    if (c < 128) { src = dA1; ld = 128; col = c; }
    else if (c < 256) { src = dA2; ld = 128; col = c - 128; }
    else if (c < 384) { src = dA3; ld = 128; col = c - 256; }
    else { src = dO; ld = 32; col = c - 384; }
    float s = 0.0f;
    if (c >= 384 || col < FS_LP_ACT) {
        const long b0 = (long)blockIdx.y * FS_LP_SPLIT_ROWS;
#pragma unroll 4
        for (int r = r0; r < FS_LP_SPLIT_ROWS; r += 8) s += (float)src[(b0 + r) * ld + col];
    }
    red[r0][threadIdx.x & 31] = s;
    __syncthreads();
    if (threadIdx.x < 32) {
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; k++) t += red[k][threadIdx.x];
        db[(long)blockIdx.y * FS_LP_NBIAS + c] = t;
    }
}

extern "C" __global__ __launch_bounds__(256) void fs_lp_adam(fs_lp_adam_args a)
{
    fs_lp_adam_body<sys::DOF>(a, blockIdx.x * 256 + threadIdx.x);
}
#endif
