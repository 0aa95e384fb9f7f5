/*
 * pimc_oracle.c -- CPU restatement of FeynSimul's Metropolis PIMC kernel.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (feynsimul_amd/)
 * may link, load or call this file; it is the checker that the HIP engine is
 * compared against, and the CPU baseline timed beside it.
 *
 * This file is compiled once per physical system: a generated header holding
 * the system's potential / operator / correlator strings (the same text the
 * reference pastes into its kernel, FeynSimul/kernel.py:253-348) is included
 * first via `-include`.  That header must define
 *     FLOAT_TYPE, FS_DOF, FS_NOPS, FS_NCORR,
 *     static inline FLOAT_TYPE fs_potential_v(const FLOAT_TYPE *x);
 *     static inline void fs_operator_v(const FLOAT_TYPE *x, FLOAT_TYPE *acc);
 *     static inline void fs_correlator_v(const FLOAT_TYPE *x, FLOAT_TYPE *prod);
 * The FP32 build uses gcc -fsingle-precision-constant, mirroring the
 * reference's -cl-single-precision-constant (kernel.py:376-377), and
 * <tgmath.h> so that sqrt/exp/log/sin on float arguments are the float
 * routines, as in OpenCL C.
 *
 * Functions and the reference lines they follow:
 *   fs_xs_step / fs_xs_float / fs_xs_int  kernel.c:214-239
 *   fs_rl_os / fs_rl_next4 / fs_rl_init   examples/pyopencl-ranluxcl.cl:363-557, 597-725
 *   rl_wrap_float / rl_wrap_int           kernel.c:180-207
 *   bisect_move                           kernel.c:258-399 (sigmaN table 605-619)
 *   kin_est / shift_diff / path_shift_move / single_node_move
 *                                         kernel.c:245-249, 407-473, 640-705
 *   fs_oracle_run_reference               kernel.c:514-837, launch geometry kernel.py:192-211
 *   fs_oracle_run_wave                    the engine's slice-per-lane ordering of the same
 *                                         Levy-bisection move (no reference counterpart;
 *                                         its law is the reference's, see the wave notes)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef FLOAT_TYPE
#error "FLOAT_TYPE must come from the generated system header"
#endif

#define FS_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* xorshift128 (Marsaglia 2003), kernel.c:214-239                              */
/* ------------------------------------------------------------------------- */
typedef struct { uint32_t x, y, z, w; } fs_xs128;

static inline void fs_xs_step(fs_xs128 *s)
{
    uint32_t t = s->x ^ (s->x << 11);
    s->x = s->y;
    s->y = s->z;
    s->z = s->w;
    s->w = s->w ^ (s->w >> 19) ^ (t ^ (t >> 8));
}

/* w * 2.328306437080797e-10: under single-precision constants the literal is
 * exactly 2^-32 and w is rounded to float first, so the result lies in [0,1]. */
static inline FLOAT_TYPE fs_xs_float(fs_xs128 *s)
{
    fs_xs_step(s);
    return s->w * 2.328306437080797e-10;
}

static inline uint32_t fs_xs_int(fs_xs128 *s)
{
    fs_xs_step(s);
    return s->w;
}

/* ------------------------------------------------------------------------- */
/* RANLUXCL (Luescher's RANLUX, 24-bit subtract-with-borrow, lags 10 and 24)  */
/* pyopencl-ranluxcl.cl:265-557, 597-725 (v1.3.1 as vendored in examples/).   */
/* ------------------------------------------------------------------------- */
typedef struct {
    float s[24];        /* s01..s24: s[k-1] holds x_{n-k} at cycle position 0 */
    float carry;
    float dummy;
    int32_t in24;
    int32_t stepnr;
} fs_ranlux;            /* 112 bytes, same field order as ranluxcl_state_t */

#define FS_TWOM24 ((float)(1.0 / 16777216.0))
#define FS_TWOM12 ((float)(1.0 / 4096.0))

static inline float fs_rl_os(float sj24m1, float sj24, float *si24, float *carry)
{
    float uni = sj24 - (*si24) - (*carry);
    if (uni < 0.0f) {
        uni += 1.0f;
        *carry = FS_TWOM24;
    } else {
        *carry = 0.0f;
    }
    *si24 = uni;
    float out = uni;
    if (uni < FS_TWOM12) {
        out += FS_TWOM24 * sj24m1;
        if (out == 0.0f) out = FS_TWOM24 * FS_TWOM24;
    }
    return out;
}

/* One step at cycle position t (0..23).  Position t overwrites s(24-t) with
 * s(10-t) - s(24-t) - carry (1-based, wrapping), and reads s(9-t) for the
 * small-value fill; this is the unrolled schedule of ranluxcl32. */
static inline float fs_rl_step(fs_ranlux *r, int t)
{
    int tgt = 23 - t;
    int lag10 = (9 - t + 24) % 24;
    int lag9 = (8 - t + 24) % 24;
    return fs_rl_os(r->s[lag9], r->s[lag10], &r->s[tgt], &r->carry);
}

static inline void fs_rl_skip(fs_ranlux *r, int nskip)
{
    int initialskips = r->stepnr ? (24 - r->stepnr) : 0;
    int bulkskips = ((nskip - initialskips) / 24) * 24;
    int remainingskips = nskip - initialskips - bulkskips;
    for (int t = r->stepnr; t < 24 && initialskips; t++) fs_rl_step(r, t);
    for (int b = 0; b < bulkskips / 24; b++)
        for (int t = 0; t < 24; t++) fs_rl_step(r, t);
    for (int t = 0; t < remainingskips; t++) fs_rl_step(r, t);
    r->stepnr = remainingskips;
}

/* ranluxcl32: four outputs, then the luxury skip after every 24 outputs.
 * nskip % 24 == 0 is the "planar" scheme (skip when the cycle wraps); any
 * other multiple of 4 is "planar shift" (skip when in24 reaches 24). */
static void fs_rl_next4(fs_ranlux *r, int nskip, float out[4])
{
    int t0 = r->stepnr;
    for (int k = 0; k < 4; k++) out[k] = fs_rl_step(r, t0 + k);
    r->stepnr = (t0 == 20) ? 0 : t0 + 4;
    if (nskip % 24 == 0) {
        if (t0 == 20 && nskip) fs_rl_skip(r, nskip);
    } else {
        r->in24 += 4;
        if (r->in24 == 24) {
            r->in24 = 0;
            fs_rl_skip(r, nskip);
        }
    }
}

/* ranluxcl64: two float4 draws combined into 48-bit doubles
 * (pyopencl-ranluxcl.cl:893-920). */
static void fs_rl_next4_double(fs_ranlux *r, int nskip, double out[4])
{
    float v[4];
    for (int h = 0; h < 2; h++) {
        fs_rl_next4(r, nskip, v);
        if (v[0] == FS_TWOM24 * FS_TWOM24) v[0] = 0.0f;
        if (v[2] == FS_TWOM24 * FS_TWOM24) v[2] = 0.0f;
        out[2 * h + 0] = (double)v[0] + (double)v[1] / 16777216;
        out[2 * h + 1] = (double)v[2] + (double)v[3] / 16777216;
    }
}

/* ranluxcl_init: xorshift64 (13, 7, 17) expansion of a 64-bit seed into 24
 * 24-bit fractions, then 32 x 24 warm-up steps (pyopencl-ranluxcl.cl:597-725). */
static void fs_rl_expand(uint64_t x, fs_ranlux *r)
{
    const float inv = (float)(1.0 / 16777216.0);
    for (int blk = 0; blk < 3; blk++) {
        uint64_t x1, x2, x3;
        x ^= (x << 13); x ^= (x >> 7); x ^= (x << 17); x1 = x;
        x ^= (x << 13); x ^= (x >> 7); x ^= (x << 17); x2 = x;
        x ^= (x << 13); x ^= (x >> 7); x ^= (x << 17); x3 = x;
        float *s = r->s + 8 * blk;
        s[0] = (float)(x1 >> 40) * inv;
        s[1] = (float)((x1 & 0x000000FFFFFFFFFFull) >> 16) * inv;
        s[2] = (float)(((x1 & 0x000000000000FFFFull) << 8) + (x2 >> 56)) * inv;
        s[3] = (float)((x2 & 0x00FFFFFFFFFFFFFFull) >> 32) * inv;
        s[4] = (float)((x2 & 0x00000000FFFFFFFFull) >> 8) * inv;
        s[5] = (float)(((x2 & 0x00000000000000FFull) << 16) + (x3 >> 48)) * inv;
        s[6] = (float)((x3 & 0x0000FFFFFFFFFFFFull) >> 24) * inv;
        s[7] = (float)(x3 & 0x0000000000FFFFFFull) * inv;
    }
    r->dummy = 0.0f;
    r->in24 = 0;
    r->stepnr = 0;
    r->carry = 0.0f;
    if (r->s[23] == 0.0f) r->carry = FS_TWOM24;
}

static void fs_rl_init(uint64_t x, fs_ranlux *r)
{
    fs_rl_expand(x, r);
    for (int i = 0; i < 32; i++)
        for (int t = 0; t < 24; t++) fs_rl_step(r, t);
}

/* Per-launch buffered wrapper (kernel.c:180-207): a float4 (double4 in FP64)
 * is refilled when the counter is a multiple of 4; the counter starts at 0
 * at every launch, so values left in the buffer are dropped. */
typedef struct {
    fs_ranlux *st;
    uint32_t cnt;
    FLOAT_TYPE buf[4];
    int nskip;
} fs_rlwrap;

static inline FLOAT_TYPE rl_wrap_float(fs_rlwrap *w)
{
    if ((w->cnt & 3) == 0) {
#ifdef FS_DOUBLE
        double v[4];
        fs_rl_next4_double(w->st, w->nskip, v);
#else
        float v[4];
        fs_rl_next4(w->st, w->nskip, v);
#endif
        for (int k = 0; k < 4; k++) w->buf[k] = (FLOAT_TYPE)v[k];
    }
    return w->buf[(w->cnt++) & 3];
}

static inline uint32_t rl_wrap_int(fs_rlwrap *w, FLOAT_TYPE scale)
{
    FLOAT_TYPE u = rl_wrap_float(w);
    return (uint32_t)(u * scale);
}

/* ------------------------------------------------------------------------- */
/* Exported RNG entry points (golden vectors and device cross-checks)         */
/* ------------------------------------------------------------------------- */
FS_EXPORT void fs_oracle_xs_stream(const uint32_t *rows, int nrows, int ndraw,
                                   uint32_t *out_int, FLOAT_TYPE *out_float)
{
    for (int r = 0; r < nrows; r++) {
        fs_xs128 s = {rows[4 * r], rows[4 * r + 1], rows[4 * r + 2], rows[4 * r + 3]};
        fs_xs128 s2 = s;
        for (int k = 0; k < ndraw; k++) {
            if (out_int) out_int[(size_t)r * ndraw + k] = fs_xs_int(&s);
            if (out_float) out_float[(size_t)r * ndraw + k] = fs_xs_float(&s2);
        }
    }
}

FS_EXPORT void fs_oracle_ranlux_init(const uint64_t *seeds, int n, fs_ranlux *states)
{
    for (int i = 0; i < n; i++) fs_rl_init(seeds[i], &states[i]);
}

/* the seed expansion of ranluxcl_init without its warm-up cycles */
FS_EXPORT void fs_oracle_ranlux_expand(const uint64_t *seeds, int n, fs_ranlux *states)
{
    for (int i = 0; i < n; i++) fs_rl_expand(seeds[i], &states[i]);
}

FS_EXPORT void fs_oracle_ranlux_stream(fs_ranlux *states, int n, int nskip, int nquad,
                                       float *out)
{
    for (int i = 0; i < n; i++)
        for (int q = 0; q < nquad; q++) fs_rl_next4(&states[i], nskip, out + ((size_t)i * nquad + q) * 4);
}

FS_EXPORT void fs_oracle_ranlux_stream_double(fs_ranlux *states, int n, int nskip, int nquad,
                                              double *out)
{
    for (int i = 0; i < n; i++)
        for (int q = 0; q < nquad; q++)
            fs_rl_next4_double(&states[i], nskip, out + ((size_t)i * nquad + q) * 4);
}

FS_EXPORT int fs_oracle_float_bytes(void) { return (int)sizeof(FLOAT_TYPE); }
FS_EXPORT int fs_oracle_dof(void) { return FS_DOF; }

FS_EXPORT double fs_oracle_potential(const FLOAT_TYPE *x) { return (double)fs_potential_v(x); }

FS_EXPORT void fs_oracle_operators(const FLOAT_TYPE *x, FLOAT_TYPE *acc) { fs_operator_v(x, acc); }

/* ------------------------------------------------------------------------- */
/* Run configuration shared by both orderings                                */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t N, S, DOF;
    int32_t nWalkers, walkersPerGroup;
    int32_t parallelizePath;
    int32_t ranlux;          /* 0: xorshift128, 1: RANLUXCL */
    int32_t nskip;           /* RANLUXCL skip count per 24 outputs */
    int32_t operatorRuns, metroSteps;
    int32_t enableOperator, nOps;
    int32_t enableCorrelator, nCorr, corLags;
    int32_t enableBins, binRes;
    int32_t localSize;       /* work-group size of the one-walker-per-item mode */
    int32_t nThreads;        /* work-items in the launch */
    int32_t nHistObs;        /* wave ordering: observables binned into one histogram */
    int32_t histBins;
    double epsilon;          /* beta / N */
    double opNorm;
    double ranluxIntScale;   /* float(ranluxIntMax + 1) */
    double xMin, xMax, invBinSize;
    double histLo, histInv;
    int32_t bisection, pathShift, singleNode, pad_;
    double alpha, psAlpha;   /* single-node and path-shift step widths */
    double epsInv2;          /* (N / beta)^2, kinEnergyEst's scale (kernel.py:299) */
} fs_oracle_cfg;

typedef struct {
    fs_xs128 xs;
    fs_rlwrap rl;
    int ranlux;
    FLOAT_TYPE iscale;
} fs_stream;

static inline FLOAT_TYPE st_float(fs_stream *s)
{
    return s->ranlux ? rl_wrap_float(&s->rl) : fs_xs_float(&s->xs);
}

static inline uint32_t st_int(fs_stream *s)
{
    return s->ranlux ? rl_wrap_int(&s->rl, s->iscale) : fs_xs_int(&s->xs);
}

/* Offset draw in [0, n), n a power of two.  xorshift: the low bits of one word
 * (kernel.c:715).  RANLUX: the top log2(n) bits of one uniform; the reference's
 * st_int(..) & (n - 1) is 0 for every u on the 2^-24 grid (kernel.c:206), so
 * its offsets never move. */
static inline uint32_t st_below(fs_stream *s, uint32_t n)
{
    return s->ranlux ? (uint32_t)(rl_wrap_float(&s->rl) * (FLOAT_TYPE)n) : (fs_xs_int(&s->xs) & (n - 1u));
}

/* Gather the DOF coordinates of one slice (walker path is DOF-major). */
static inline void gather(const FLOAT_TYPE *path, int N, int s, FLOAT_TYPE *x)
{
    for (int d = 0; d < FS_DOF; d++) x[d] = path[(size_t)d * N + s];
}

/* doBisectMove, kernel.c:258-399.  `path` is one walker's DOF-major path. */
static void bisect_move(FLOAT_TYPE *path, int N, int S, uint32_t startPoint, fs_stream *st,
                        uint32_t *accepts, const FLOAT_TYPE *sigmaN, FLOAT_TYPE epsv,
                        FLOAT_TYPE *oldPath)
{
    const int P = 1 << S;
    FLOAT_TYPE x[FS_DOF];
    FLOAT_TYPE actionOld = 0.0;
    for (int i = 1; i < P; i++) {
        int iCurr = i + (int)startPoint;
        if (iCurr >= N) iCurr -= N;
        for (int d = 0; d < FS_DOF; d++) oldPath[(i - 1) * FS_DOF + d] = path[(size_t)d * N + iCurr];
        gather(path, N, iCurr, x);
        actionOld += epsv * fs_potential_v(x);
    }
    for (int n = 1; n <= S; n++) {
        for (int i = 0; i < (1 << (n - 1)); i++) {
            uint32_t pm = (uint32_t)((1 + 2 * i) * (1 << (S - n))) + startPoint;
            uint32_t pl = (uint32_t)(i * (1 << (S - n + 1))) + startPoint;
            uint32_t pr = (uint32_t)((i + 1) * (1 << (S - n + 1))) + startPoint;
            if (pm >= (uint32_t)N) pm -= N;
            if (pl >= (uint32_t)N) pl -= N;
            if (pr >= (uint32_t)N) pr -= N;
            for (int d = 0; d < FS_DOF; d++) {
                size_t nd = (size_t)d * N;
                FLOAT_TYPE u = st_float(st);
                FLOAT_TYPE v = st_float(st);
                FLOAT_TYPE rz;
                if (v == 0.0)
                    rz = 0.0;
                else
                    rz = sigmaN[n] * sqrt(-2.0 * log(v)) * sin(2.0 * M_PI * u);
                path[pm + nd] = 0.5 * (path[pl + nd] + path[pr + nd]) + rz;
            }
        }
    }
    FLOAT_TYPE actionNew = 0.0;
    for (int i = 1; i < P; i++) {
        int iCurr = i + (int)startPoint;
        if (iCurr >= N) iCurr -= N;
        gather(path, N, iCurr, x);
        actionNew += epsv * fs_potential_v(x);
    }
    if (exp(actionOld - actionNew) < st_float(st)) {
        for (int i = 1; i < P; i++) {
            int iCurr = i + (int)startPoint;
            if (iCurr >= N) iCurr -= N;
            for (int d = 0; d < FS_DOF; d++) path[(size_t)d * N + iCurr] = oldPath[(i - 1) * FS_DOF + d];
        }
    } else {
        (*accepts)++;
    }
}

static void sigma_table(int S, FLOAT_TYPE epsv, FLOAT_TYPE *sigmaN)
{
    int Ns = 1 << S;
    int twoPow = 1;
    sigmaN[0] = ((FLOAT_TYPE)Ns) * epsv;
    for (int i = 1; i <= S; i++) {
        twoPow *= 2;
        sigmaN[i] = sqrt(((FLOAT_TYPE)Ns * epsv / (2.0 * (FLOAT_TYPE)twoPow)));
    }
}

/* histogram(), kernel.c:478-506: DOF-dimensional grid over raw coordinates. */
static inline void histogram_slice(const FLOAT_TYPE *path, int N, int s, const fs_oracle_cfg *c,
                                   uint32_t *bins)
{
    int binIndex = 0, inc = 1;
    FLOAT_TYPE xMin = (FLOAT_TYPE)c->xMin, xMax = (FLOAT_TYPE)c->xMax;
    FLOAT_TYPE ib = (FLOAT_TYPE)c->invBinSize;
    for (int p = 0; p < FS_DOF; p++) {
        FLOAT_TYPE v = path[(size_t)p * N + s];
        if (v < xMin || v >= xMax) return;
        binIndex += (int)((uint32_t)((v - xMin) * ib)) * inc;
        inc *= c->binRes;
    }
#pragma omp atomic
    bins[binIndex] += 1u;
}

/* ------------------------------------------------------------------------- */
/* fs_oracle_run_reference: one launch of the reference `metropolis` kernel.  */
/*                                                                           */
/* Work-item geometry (kernel.py:192-211):                                   */
/*   ParallelizePath: local (T, W), global ((walkers/W) T, W), T = N / 2^S;   */
/*     item (gid0, gid1): walker = (gid0 / T) W + gid1, lid0 = gid0 % T,      */
/*     threadId = gid0 + gid1 (walkers/W) T, group row = gid0/T + gid1 (walkers/W).
 *   otherwise: one item per walker; threadId = walker; group row =          */
/*     walker / localSize.                                                   */
/* Group xorshift rows are read from a snapshot taken before the launch: in  */
/* the reference they alias other items' own rows, which are rewritten at    */
/* the end of the launch, so their value there depends on scheduling.        */
/* ------------------------------------------------------------------------- */
typedef struct {
    int walker, lid0, threadId, groupRow;
} fs_item;

static fs_item item_of(const fs_oracle_cfg *c, int walker, int lid0)
{
    fs_item it;
    it.walker = walker;
    it.lid0 = lid0;
    if (c->parallelizePath) {
        int T = c->N >> c->S, W = c->walkersPerGroup, nG = c->nWalkers / W;
        int g = walker / W, l = walker % W;
        int gid0 = g * T + lid0;
        it.threadId = gid0 + l * nG * T;
        it.groupRow = g + l * nG;
    } else {
        it.threadId = walker;
        it.groupRow = walker / c->localSize;
    }
    return it;
}

/* kinEnergyEst (kernel.c:245-249): half the squared step in units of 1/eps^2. */
static inline FLOAT_TYPE kin_est(FLOAT_TYPE a, FLOAT_TYPE b, FLOAT_TYPE epsInv2)
{
    FLOAT_TYPE delta = a - b;
    return 0.5 * (delta * delta) * epsInv2;
}

/* shiftPathEnergyDiff (kernel.c:407-473): adds `offset` to coordinate `degree`
 * of slices left..right (wrapping past N - 1 when left > right) and returns the
 * change of kinetic estimator plus potential, in the reference's order. */
static FLOAT_TYPE shift_diff(FLOAT_TYPE *path, int N, int degree, FLOAT_TYPE offset, int left, int right,
                             FLOAT_TYPE epsInv2)
{
    if (left == right) return 0.0;
    FLOAT_TYPE *xd = path + (size_t)degree * N;
    FLOAT_TYPE kinDiff = 0.0, potDiff = 0.0, x[FS_DOF];
    int ll, rr;
    if (left < right) {
        ll = left == 0 ? N - 1 : left - 1;
        rr = right == N - 1 ? 0 : right + 1;
    } else {
        ll = left - 1;
        rr = right + 1;
    }
    kinDiff -= kin_est(xd[left], xd[ll], epsInv2);
    kinDiff -= kin_est(xd[right], xd[rr], epsInv2);
    int count = left < right ? right - left + 1 : N - left + right + 1;
    for (int t = 0; t < count; t++) {
        int i = left + t;
        if (i >= N) i -= N;
        gather(path, N, i, x);
        potDiff -= fs_potential_v(x);
        xd[i] = xd[i] + offset;
        gather(path, N, i, x);
        potDiff += fs_potential_v(x);
    }
    kinDiff += kin_est(xd[left], xd[ll], epsInv2);
    kinDiff += kin_est(xd[right], xd[rr], epsInv2);
    return kinDiff + potDiff;
}

/* Path-shift move (kernel.c:640-663): two integer draws from the group stream
 * pick the interval, a uniform the offset in [-PSAlpha, PSAlpha], a second
 * uniform the decision; a rejected shift is undone by shifting back. */
static void path_shift_move(FLOAT_TYPE *path, int N, int degree, fs_stream *own, fs_stream *grp, uint32_t *acc,
                            FLOAT_TYPE epsv, FLOAT_TYPE epsInv2, FLOAT_TYPE psAlpha)
{
    int left = (int)st_below(grp, (uint32_t)N);
    int right = (int)st_below(grp, (uint32_t)N);
    FLOAT_TYPE offset = psAlpha * (2.0 * st_float(own) - 1.0);
    FLOAT_TYPE dE = shift_diff(path, N, degree, offset, left, right, epsInv2);
    if (exp(-epsv * dE) < st_float(own))
        shift_diff(path, N, degree, -offset, left, right, epsInv2);
    else
        (*acc)++;
}

/* Single-node move (kernel.c:665-703): node `step` mod (N DOF) of the walker's
 * DOF-major path, uniform step in [-alpha, alpha]. */
static void single_node_move(FLOAT_TYPE *path, int N, uint32_t step, fs_stream *own, uint32_t *acc,
                             FLOAT_TYPE epsv, FLOAT_TYPE epsInv2, FLOAT_TYPE alpha)
{
    int modPoint = (int)(step % (uint32_t)N);
    int modPathPoint = (int)(step % (uint32_t)(N * FS_DOF));
    int rightIndex = modPathPoint + 1, leftIndex = modPathPoint - 1;
    if (modPoint == 0)
        leftIndex += N;
    else if (modPoint == N - 1)
        rightIndex -= N;
    FLOAT_TYPE oldX = path[modPathPoint];
    FLOAT_TYPE modX = oldX + (2.0 * st_float(own) - 1.0) * alpha;
    FLOAT_TYPE diffE = kin_est(modX, path[rightIndex], epsInv2) + kin_est(modX, path[leftIndex], epsInv2) -
                       kin_est(path[modPathPoint], path[leftIndex], epsInv2) -
                       kin_est(path[modPathPoint], path[rightIndex], epsInv2);
    FLOAT_TYPE x[FS_DOF];
    gather(path, N, modPoint, x);
    diffE -= fs_potential_v(x);
    path[modPathPoint] = modX;
    gather(path, N, modPoint, x);
    diffE += fs_potential_v(x);
    path[modPathPoint] = oldX;
    if (exp(-epsv * diffE) > st_float(own)) {
        path[modPathPoint] = modX;
        (*acc)++;
    }
}

FS_EXPORT int fs_oracle_run_reference(const fs_oracle_cfg *c, FLOAT_TYPE *paths, uint32_t *accepts,
                                      uint32_t *seeds, fs_ranlux *rlTab, fs_ranlux *groupRlTab,
                                      FLOAT_TYPE *opMeans, FLOAT_TYPE *corMeans, uint32_t *bins,
                                      int nthreads)
{
    const int N = c->N, S = c->S, P = 1 << S;
    const int T = c->parallelizePath ? (N >> S) : 1;
    const FLOAT_TYPE epsv = (FLOAT_TYPE)c->epsilon;
    const FLOAT_TYPE epsInv2 = (FLOAT_TYPE)c->epsInv2;
    if (c->DOF != FS_DOF) return -1;
    if ((c->pathShift || c->singleNode) && c->parallelizePath) return -3;
    uint32_t *snap = NULL;
    if (!c->ranlux) {
        size_t rows = (size_t)c->nThreads + 1;
        snap = (uint32_t *)malloc(rows * 4 * sizeof(uint32_t));
        if (!snap) return -2;
        memcpy(snap, seeds, rows * 4 * sizeof(uint32_t));
    }
    FLOAT_TYPE sigmaN[32];
    sigma_table(S, epsv, sigmaN);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int w = 0; w < c->nWalkers; w++) {
        FLOAT_TYPE *path = paths + (size_t)w * N * FS_DOF;
        fs_stream *own = (fs_stream *)calloc((size_t)T, sizeof(fs_stream));
        fs_stream *grp = (fs_stream *)calloc((size_t)T, sizeof(fs_stream));
        fs_item *items = (fs_item *)calloc((size_t)T, sizeof(fs_item));
        FLOAT_TYPE *opAcc = (FLOAT_TYPE *)calloc((size_t)T * (FS_NOPS + 1), sizeof(FLOAT_TYPE));
        FLOAT_TYPE *oldPath = (FLOAT_TYPE *)calloc((size_t)P * FS_DOF, sizeof(FLOAT_TYPE));
        uint32_t *acc = (uint32_t *)calloc((size_t)T, sizeof(uint32_t));
        if (!own || !grp || !items || !opAcc || !oldPath || !acc) {
#pragma omp atomic write
            err = -2;
            free(own); free(grp); free(items); free(opAcc); free(oldPath); free(acc);
            continue;
        }
        for (int k = 0; k < T; k++) {
            items[k] = item_of(c, w, k);
            own[k].ranlux = grp[k].ranlux = c->ranlux;
            own[k].iscale = grp[k].iscale = (FLOAT_TYPE)c->ranluxIntScale;
            if (c->ranlux) {
                own[k].rl.st = &rlTab[items[k].threadId];
                own[k].rl.cnt = 0;
                own[k].rl.nskip = c->nskip;
                grp[k].rl.st = &groupRlTab[items[k].threadId];
                grp[k].rl.cnt = 0;
                grp[k].rl.nskip = c->nskip;
            } else {
                const uint32_t *r = seeds + 4 * (size_t)items[k].threadId;
                own[k].xs = (fs_xs128){r[0], r[1], r[2], r[3]};
                const uint32_t *g = snap + 4 * (size_t)items[k].groupRow;
                grp[k].xs = (fs_xs128){g[0], g[1], g[2], g[3]};
            }
        }
        for (int i = 0; i < c->operatorRuns; i++) {
            for (int j = 0; j < c->metroSteps; j++) {
                for (int k = 0; k < T; k++) {
                    /* per step: path shift, single node, bisection (kernel.c:640-721) */
                    if (c->pathShift)
                        path_shift_move(path, N, (int)(((uint32_t)c->metroSteps * i + j) % FS_DOF), &own[k],
                                        &grp[k], &acc[k], epsv, epsInv2, (FLOAT_TYPE)c->psAlpha);
                    if (c->singleNode)
                        single_node_move(path, N, (uint32_t)(i * c->metroSteps + j), &own[k], &acc[k], epsv,
                                         epsInv2, (FLOAT_TYPE)c->alpha);
                    if (!c->bisection) continue;
                    uint32_t start;
                    if (c->parallelizePath)
                        start = (st_below(&grp[k], (uint32_t)P) + (uint32_t)(P * k)) % (uint32_t)N;
                    else
                        start = st_below(&grp[k], (uint32_t)N);
                    bisect_move(path, N, S, start, &own[k], &acc[k], sigmaN, epsv, oldPath);
                }
            }
            if (c->enableOperator || c->enableBins || c->enableCorrelator) {
                int span = c->parallelizePath ? P : N;
                FLOAT_TYPE x[FS_DOF], x2[FS_DOF];
                for (int k = 0; k < T; k++) {
                    for (int ts = k * span; ts < (k + 1) * span; ts++) {
                        if (c->enableOperator) {
                            gather(path, N, ts, x);
                            fs_operator_v(x, opAcc + (size_t)k * (FS_NOPS + 1));
                        }
                        if (c->enableBins) histogram_slice(path, N, ts, c, bins);
                    }
                }
                if (c->enableCorrelator) {
                    /* Reference order: local slice index, then lag step, then
                     * item; each (slice, lag step) adds one term per bin. */
                    const int half = N / 2;
                    FLOAT_TYPE *cw = corMeans + (size_t)w * c->nCorr * half;
                    for (int tl = 0; tl < span; tl++) {
                        for (int corTI = 0; corTI < half; corTI++) {
                            for (int k = 0; k < T; k++) {
                                int ts = k * span + tl;
                                int corT = corTI + (c->parallelizePath ? k : 0);
                                if (corT >= half) corT -= half;
                                int ts2 = ts + corT < N ? ts + corT : ts + corT - N;
                                FLOAT_TYPE prod[FS_NCORR + 1];
                                for (int q = 0; q <= FS_NCORR; q++) prod[q] = 1.0;
                                gather(path, N, ts, x);
                                fs_correlator_v(x, prod);
                                gather(path, N, ts2, x2);
                                fs_correlator_v(x2, prod);
                                for (int q = 0; q < c->nCorr; q++) cw[(size_t)q * half + corT] += prod[q];
                            }
                        }
                    }
                }
            }
        }
        for (int k = 0; k < T; k++) {
            int tid = items[k].threadId;
            if (c->enableOperator)
                for (int op = 0; op < c->nOps; op++)
                    opMeans[op + (size_t)tid * c->nOps] =
                        opAcc[(size_t)k * (FS_NOPS + 1) + op] * (FLOAT_TYPE)c->opNorm;
            accepts[tid] = acc[k];
            if (!c->ranlux) {
                uint32_t *r = seeds + 4 * (size_t)tid;
                r[0] = own[k].xs.x; r[1] = own[k].xs.y; r[2] = own[k].xs.z; r[3] = own[k].xs.w;
            }
        }
        free(own); free(grp); free(items); free(opAcc); free(oldPath); free(acc);
    }
    free(snap);
    return err;
}

/* ------------------------------------------------------------------------- */
/* fs_oracle_run_wave: the engine's wave-parallel ordering of a sweep.        */
/*                                                                           */
/* Same move and acceptance law as ParallelizePath bisection (every segment  */
/* of a walker is moved once per step; the tiling offset r in [0, 2^S) is    */
/* shared), with a different assignment of random numbers:                   */
/*  - walker w owns a group stream (row w of the group table) that supplies  */
/*    r = int & (2^S - 1) once per step, and 64 lane streams (rows 64w + q);  */
/*  - slices are handled in blocks of B = min(64, N) consecutive slices      */
/*    starting at r + gB; position p of a block is slice r + gB + p, in      */
/*    segment p / 2^S at segment position i = p % 2^S;                       */
/*  - every lane stream draws 2 ceil(DOF/2) uniforms per block; position p   */
/*    takes the draws of stream (r + p) mod B, the residue of its slice, so  */
/*    a slice keeps its stream whatever the offset; a position with i != 0   */
/*    uses them as pairs (u, v) and forms both Box-Muller normals            */
/*    rho sin(2 pi u), rho cos(2 pi u) for coordinates 2k and 2k+1; the      */
/*    position with i == 0 (the fixed left end) keeps its first draw as the  */
/*    segment's acceptance uniform;                                          */
/*  - the action change of a segment is epsv * sum_i (V_new - V_old), summed   */
/*    by an xor butterfly over the segment's positions (the end adds 0).     */
/* Per-walker outputs: accepted move count, FP64 operator means, FP64         */
/* correlator bins.                                                          */
/* ------------------------------------------------------------------------- */
static inline int lvl_of(int i, int S)
{
    /* level n at which segment position i (1 <= i < 2^S) is generated */
    int tz = __builtin_ctz((unsigned)i);
    return S - tz;
}

FS_EXPORT int fs_oracle_run_wave(const fs_oracle_cfg *c, FLOAT_TYPE *paths, uint32_t *accepts,
                                 uint32_t *laneSeeds, uint32_t *groupSeeds, fs_ranlux *laneRl,
                                 fs_ranlux *groupRl, double *opMeans, double *corBins, uint32_t *obsHist,
                                 int nthreads)
{
    const int N = c->N, S = c->S, SEG = 1 << S;
    const int B = N < 64 ? N : 64;
    const int NB = N / B;
    const FLOAT_TYPE epsv = (FLOAT_TYPE)c->epsilon;
    if (c->DOF != FS_DOF || SEG > B) return -1;
    FLOAT_TYPE sigmaN[32];
    sigma_table(S, epsv, sigmaN);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int w = 0; w < c->nWalkers; w++) {
        FLOAT_TYPE *path = paths + (size_t)w * N * FS_DOF;
        fs_stream lane[64], grp;
        memset(lane, 0, sizeof(lane));
        memset(&grp, 0, sizeof(grp));
        grp.ranlux = c->ranlux;
        grp.iscale = (FLOAT_TYPE)c->ranluxIntScale;
        if (c->ranlux) {
            grp.rl.st = &groupRl[w];
            grp.rl.nskip = c->nskip;
        } else {
            uint32_t *g = groupSeeds + 4 * (size_t)w;
            grp.xs = (fs_xs128){g[0], g[1], g[2], g[3]};
        }
        for (int p = 0; p < 64; p++) {
            lane[p].ranlux = c->ranlux;
            if (c->ranlux) {
                lane[p].rl.st = &laneRl[(size_t)w * 64 + p];
                lane[p].rl.nskip = c->nskip;
            } else {
                uint32_t *r = laneSeeds + 4 * ((size_t)w * 64 + p);
                lane[p].xs = (fs_xs128){r[0], r[1], r[2], r[3]};
            }
        }
        double opS[FS_NOPS + 1], opC[FS_NOPS + 1];
        for (int q = 0; q <= FS_NOPS; q++) opS[q] = opC[q] = 0.0;
        uint32_t nacc = 0;
        FLOAT_TYPE xo[64][FS_DOF], xn[64][FS_DOF], xr[FS_DOF], z[64][FS_DOF + 1];
        FLOAT_TYPE dv[64];
        for (int i = 0; i < c->operatorRuns; i++) {
            for (int j = 0; j < c->metroSteps; j++) {
                uint32_t r = st_below(&grp, (uint32_t)SEG);
                for (int g = 0; g < NB; g++) {
                    int base = (int)r + g * B;
                    for (int p = 0; p < B; p++) gather(path, N, (base + p) % N, xo[p]);
                    gather(path, N, (base + B) % N, xr);
                    /* every one of the 64 streams draws KDRAW uniforms per block, so
                     * buffered streams stay in step; position p uses the stream of
                     * its slice residue (r + p) mod B; interior positions use the
                     * draws as Box-Muller pairs, a segment's left end keeps the
                     * first as the acceptance uniform */
                    const int KDRAW = 2 * ((FS_DOF + 1) / 2);
                    FLOAT_TYPE ust[64][FS_DOF + 1];
                    FLOAT_TYPE ufirst[64];
                    for (int q = 0; q < 64; q++)
                        for (int k = 0; k < KDRAW; k++) ust[q][k] = st_float(&lane[q]);
                    for (int p = 0; p < B; p++) {
                        const FLOAT_TYPE *u = ust[((int)r + p) % B];
                        ufirst[p] = u[0];
                        int ip = p & (SEG - 1);
                        memcpy(xn[p], xo[p], sizeof(xo[p]));
                        dv[p] = 0.0;
                        if (ip == 0) continue;
                        FLOAT_TYPE sg = sigmaN[lvl_of(ip, S)];
                        for (int d = 0; d < FS_DOF; d += 2) {
                            FLOAT_TYPE uu = u[d];
                            FLOAT_TYPE v = u[d + 1];
                            FLOAT_TYPE z0 = 0.0, z1 = 0.0;
                            if (v != 0.0) {
                                FLOAT_TYPE rho = sqrt(-2.0 * log(v));
                                FLOAT_TYPE ph = 2.0 * M_PI * uu;
                                z0 = rho * sin(ph);
                                z1 = rho * cos(ph);
                            }
                            z[p][d] = sg * z0;
                            z[p][d + 1] = sg * z1;
                        }
                    }
                    for (int n = 1; n <= S; n++) {
                        int h = SEG >> n;
                        for (int p = 0; p < B; p++) {
                            int ip = p & (SEG - 1);
                            if ((ip & (2 * h - 1)) != h) continue;
                            const FLOAT_TYPE *L = xn[p - h];
                            const FLOAT_TYPE *R = (p + h == B) ? xr : xn[p + h];
                            for (int d = 0; d < FS_DOF; d++) xn[p][d] = 0.5 * (L[d] + R[d]) + z[p][d];
                        }
                    }
                    for (int p = 0; p < B; p++) {
                        int ip = p & (SEG - 1);
                        if (ip) dv[p] = fs_potential_v(xn[p]) - fs_potential_v(xo[p]);
                    }
                    for (int q = 0; q < B / SEG; q++) {
                        /* xor butterfly over the segment's lanes */
                        FLOAT_TYPE red[64];
                        for (int k = 0; k < SEG; k++) red[k] = dv[q * SEG + k];
                        for (int off = 1; off < SEG; off <<= 1) {
                            FLOAT_TYPE nxt[64];
                            for (int k = 0; k < SEG; k++) nxt[k] = red[k] + red[k ^ off];
                            memcpy(red, nxt, sizeof(FLOAT_TYPE) * SEG);
                        }
                        FLOAT_TYPE dS = epsv * red[0];
                        FLOAT_TYPE ua = ufirst[q * SEG];
                        if (!(exp(-dS) < ua)) {
                            nacc++;
                            for (int k = 1; k < SEG; k++) {
                                int p = q * SEG + k;
                                int s = (base + p) % N;
                                for (int d = 0; d < FS_DOF; d++) path[(size_t)d * N + s] = xn[p][d];
                            }
                        }
                    }
                }
            }
            if (c->enableOperator) {
                /* each lane sums its slices p, p+B, ... in FP32, then a
                 * butterfly over the B lanes; FP64 Kahan across measurements */
                FLOAT_TYPE part[64][FS_NOPS + 1];
                for (int p = 0; p < B; p++) {
                    for (int q = 0; q <= FS_NOPS; q++) part[p][q] = 0.0;
                    FLOAT_TYPE x[FS_DOF];
                    for (int t = p; t < N; t += B) {
                        gather(path, N, t, x);
                        fs_operator_v(x, part[p]);
                    }
                }
                for (int q = 0; q < c->nOps; q++) {
                    FLOAT_TYPE red[64];
                    for (int p = 0; p < B; p++) red[p] = part[p][q];
                    for (int off = 1; off < B; off <<= 1) {
                        FLOAT_TYPE nxt[64];
                        for (int p = 0; p < B; p++) nxt[p] = red[p] + red[p ^ off];
                        memcpy(red, nxt, sizeof(FLOAT_TYPE) * B);
                    }
                    double yk = (double)red[0] - opC[q];
                    double tk = opS[q] + yk;
                    opC[q] = (tk - opS[q]) - yk;
                    opS[q] = tk;
                }
            }
            if (c->nHistObs > 0) {
                /* histogram of scalar observables at every slice */
                FLOAT_TYPE x[FS_DOF], ob[FS_NHOBS];
                for (int t = 0; t < N; t++) {
                    gather(path, N, t, x);
                    fs_hist_obs_v(x, ob);
                    for (int o = 0; o < c->nHistObs; o++) {
                        FLOAT_TYPE v = (ob[o] - (FLOAT_TYPE)c->histLo) * (FLOAT_TYPE)c->histInv;
                        if (v >= 0.0 && v < (FLOAT_TYPE)c->histBins) {
#pragma omp atomic
                            obsHist[(int)v] += 1u;
                        }
                    }
                }
            }
            if (c->enableCorrelator) {
                int L = c->corLags;
                double *cb = corBins + (size_t)w * c->nCorr * L;
                FLOAT_TYPE x[FS_DOF];
                FLOAT_TYPE *cv = (FLOAT_TYPE *)malloc(sizeof(FLOAT_TYPE) * (size_t)N * (FS_NCORR + 1));
                for (int t = 0; t < N; t++) {
                    FLOAT_TYPE prod[FS_NCORR + 1];
                    for (int q = 0; q <= FS_NCORR; q++) prod[q] = 1.0;
                    gather(path, N, t, x);
                    fs_correlator_v(x, prod);
                    for (int q = 0; q < c->nCorr; q++) cv[(size_t)q * N + t] = prod[q];
                }
                for (int q = 0; q < c->nCorr; q++)
                    for (int tau = 0; tau < L; tau++) {
                        double sum = 0.0;
                        for (int t = 0; t < N; t++)
                            sum += (double)cv[(size_t)q * N + t] * (double)cv[(size_t)q * N + (t + tau) % N];
                        cb[(size_t)q * L + tau] += sum;
                    }
                free(cv);
            }
        }
        accepts[w] = nacc;
        if (c->enableOperator)
            for (int q = 0; q < c->nOps; q++)
                opMeans[(size_t)w * c->nOps + q] = opS[q] / ((double)c->operatorRuns * N);
        if (!c->ranlux) {
            uint32_t *g = groupSeeds + 4 * (size_t)w;
            g[0] = grp.xs.x; g[1] = grp.xs.y; g[2] = grp.xs.z; g[3] = grp.xs.w;
            for (int p = 0; p < 64; p++) {
                uint32_t *rr = laneSeeds + 4 * ((size_t)w * 64 + p);
                rr[0] = lane[p].xs.x; rr[1] = lane[p].xs.y; rr[2] = lane[p].xs.z; rr[3] = lane[p].xs.w;
            }
        }
    }
    return err;
}
