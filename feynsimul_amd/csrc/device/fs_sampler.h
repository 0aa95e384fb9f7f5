// fs_sampler.h -- Metropolis bisection samplers of the PIMC engine (gfx950).
//
// Compiled by hiprtc once per (system, N, S, DOF, operators, flags), after a
// generated prelude that defines the configuration macros below and the
// system section (the user's potential / operator / correlator strings
// wrapped as fs_potential_v, fs_operator_v, fs_correlator_v).  This mirrors
// the reference, which pastes the same strings into its OpenCL template and
// rebuilds per configuration (FeynSimul/kernel.py:253-382).
//
// Configuration macros: FS_N, FS_S, FS_DOF, FS_NOPS, FS_NCORR, FS_CORLAGS,
// FLOAT_TYPE, FS_DOUBLE, FS_RANLUX, FS_NSKIP, FS_ISCALE, FS_BINS, FS_BINRES,
// FS_XMIN, FS_XMAX, FS_INVBIN, FS_OPNORM, FS_PARALLEL, FS_W, FS_LOCAL,
// FS_WPB (waves per block of the wave sampler), FS_SAMPLER (0 exact, 1 wave),
// FS_RESIDENT (wave path storage), FS_NHOBS, FS_HBINS, FS_HLO, FS_HINV
// (observable histogram of the wave sampler).
//
// Two samplers share the move:
//
//  fs_exact_sweep -- one work-item per reference work-item, with the
//    reference's work-item numbering, stream rows, draw order and
//    accumulation order (kernel.c:514-837; geometry kernel.py:192-211).  It
//    is the parity anchor against the CPU restatement of the reference.
//
//  fs_wave_sweep -- the throughput sampler.  One wavefront owns one walker
//    for the whole launch; a block of 64 consecutive slices is spread one
//    slice per lane, so the 2^S-1 midpoints of a segment, the potential
//    evaluations and the action sum all run in parallel across lanes and
//    the segment never leaves registers.  Lane j always holds the slices
//    congruent to j mod 64 and draws from its own stream; with tiling
//    offset r it plays block position (j - r) mod 64.  The offset of every
//    step comes from a per-walker group stream (the ParallelizePath scheme
//    of kernel.c:713-716).  The proposal and the acceptance rule are the
//    reference's; only the assignment of random numbers differs.
#pragma once

#include "fs_abi_args.h"
#include "fs_rng.h"

typedef FLOAT_TYPE fs_f;

#define FS_SEG (1 << FS_S)
#define FS_B (FS_N < 64 ? FS_N : 64)
#define FS_NB (FS_N / FS_B)
#define FS_KPAIR ((FS_DOF + 1) / 2)
#define FS_KDRAW (2 * FS_KPAIR)
#define FS_NOPS1 (FS_NOPS > 0 ? FS_NOPS : 1)
#define FS_NCORR1 (FS_NCORR > 0 ? FS_NCORR : 1)

#if FS_SEG > FS_N
#error "2^S must not exceed N"
#endif

// ---------------------------------------------------------------------------
// Random streams as the samplers see them
// ---------------------------------------------------------------------------
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
#endif
};

// sigmaN[n] = sqrt(2^S eps / (2 2^n)), the Levy midpoint width (kernel.c:605-619).
FS_DEV fs_f fs_sigma(int n)
{
    return sqrt(((fs_f)FS_SEG * eps / ((fs_f)2.0 * (fs_f)(1 << n))));
}

FS_DEV void fs_gather(const fs_f *__restrict__ path, int s, fs_f x[FS_DOF])
{
#pragma unroll
    for (int d = 0; d < FS_DOF; d++) x[d] = path[d * FS_N + s];
}

// DOF-dimensional grid histogram over raw coordinates (kernel.c:478-506).
FS_DEV void fs_histogram(const fs_f x[FS_DOF], fs_u32 *__restrict__ bins)
{
#if FS_BINS
    int idx = 0, inc = 1;
#pragma unroll
    for (int p = 0; p < FS_DOF; p++) {
        if (x[p] < (fs_f)FS_XMIN || x[p] >= (fs_f)FS_XMAX) return;
        idx += (int)((fs_u32)((x[p] - (fs_f)FS_XMIN) * (fs_f)FS_INVBIN)) * inc;
        inc *= FS_BINRES;
    }
    atomicAdd(bins + idx, 1u);
#endif
}

// ===========================================================================
// fs_exact_sweep
// ===========================================================================
#if FS_SAMPLER == 0
FS_DEV void fs_exact_move(fs_f *__restrict__ path, fs_f *__restrict__ oldPath, fs_u32 startPoint,
                          fs_stream &st, fs_u32 &accepts)
{
    fs_f x[FS_DOF];
    fs_f actionOld = (fs_f)0.0;
    for (int i = 1; i < FS_SEG; i++) {
        int iCurr = i + (int)startPoint;
        if (iCurr >= FS_N) iCurr -= FS_N;
        fs_gather(path, iCurr, x);
#pragma unroll
        for (int d = 0; d < FS_DOF; d++) oldPath[(i - 1) * FS_DOF + d] = x[d];
        actionOld += eps * fs_potential_v(x);
    }
    for (int n = 1; n <= FS_S; n++) {
        const fs_f sig = fs_sigma(n);
        for (int i = 0; i < (1 << (n - 1)); i++) {
            fs_u32 pm = (fs_u32)((1 + 2 * i) << (FS_S - n)) + startPoint;
            fs_u32 pl = (fs_u32)(i << (FS_S - n + 1)) + startPoint;
            fs_u32 pr = (fs_u32)((i + 1) << (FS_S - n + 1)) + startPoint;
            if (pm >= (fs_u32)FS_N) pm -= FS_N;
            if (pl >= (fs_u32)FS_N) pl -= FS_N;
            if (pr >= (fs_u32)FS_N) pr -= FS_N;
#pragma unroll
            for (int d = 0; d < FS_DOF; d++) {
                const fs_f u = st.uniform();
                const fs_f v = st.uniform();
                fs_f rz = (fs_f)0.0;
                if (v != (fs_f)0.0) rz = sig * sqrt(-2.0 * log(v)) * sin(2.0 * M_PI * u);
                path[pm + d * FS_N] = 0.5 * (path[pl + d * FS_N] + path[pr + d * FS_N]) + rz;
            }
        }
    }
    fs_f actionNew = (fs_f)0.0;
    for (int i = 1; i < FS_SEG; i++) {
        int iCurr = i + (int)startPoint;
        if (iCurr >= FS_N) iCurr -= FS_N;
        fs_gather(path, iCurr, x);
        actionNew += eps * fs_potential_v(x);
    }
    if (exp(actionOld - actionNew) < st.uniform()) {
        for (int i = 1; i < FS_SEG; i++) {
            int iCurr = i + (int)startPoint;
            if (iCurr >= FS_N) iCurr -= FS_N;
#pragma unroll
            for (int d = 0; d < FS_DOF; d++) path[d * FS_N + iCurr] = oldPath[(i - 1) * FS_DOF + d];
        }
    } else {
        accepts++;
    }
}

extern "C" __global__ void fs_exact_sweep(fs_kargs a)
{
#if FS_PARALLEL
    // local (T, W), global ((walkers/W) T, W); lid0 varies fastest.
    constexpr int T = FS_N >> FS_S;
    const int lid0 = threadIdx.x % T;
    const int lid1 = threadIdx.x / T;
    const long nG = a.nWalkers / FS_W;
    const long g = blockIdx.x;
    const long walker = g * FS_W + lid1;
    const long threadId = (g * T + lid0) + (long)lid1 * nG * T;
    const long groupRow = g + (long)lid1 * nG;
#else
    const long walker = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (walker >= a.nWalkers) return;
    const int lid0 = 0;
    const long threadId = walker;
    const long groupRow = walker / FS_LOCAL;
#endif
    fs_f *path = reinterpret_cast<fs_f *>(a.paths) + walker * (long)(FS_N * FS_DOF);
    fs_f *oldPath = reinterpret_cast<fs_f *>(a.oldPaths) + threadId * (long)((FS_SEG - 1) * FS_DOF);

    fs_stream own, grp;
#if FS_RANLUX
    own.load(a.rl, threadId);
    grp.load(a.grl, threadId);
#else
    own.load(a.seeds, threadId);
    grp.load(a.seedsSnap, groupRow);
#endif
    fs_u32 accepts = 0;
    fs_f opAccum[FS_NOPS1];
#pragma unroll
    for (int k = 0; k < FS_NOPS1; k++) opAccum[k] = (fs_f)0.0;

    for (int i = 0; i < a.operatorRuns; i++) {
#if FS_PARALLEL
        __syncthreads();
#endif
        for (int j = 0; j < a.metroSteps; j++) {
#if FS_PARALLEL
            const fs_u32 start = ((grp.integer() & (fs_u32)(FS_SEG - 1)) + (fs_u32)(FS_SEG * lid0)) % (fs_u32)FS_N;
#else
            const fs_u32 start = (grp.integer() & (fs_u32)(FS_N - 1)) % (fs_u32)FS_N;
#endif
            fs_exact_move(path, oldPath, start, own, accepts);
#if FS_PARALLEL
            __syncthreads();
#endif
        }
#if FS_NOPS > 0 || FS_BINS || FS_NCORR > 0
#if FS_PARALLEL
        const int t0 = lid0 * FS_SEG, t1 = (lid0 + 1) * FS_SEG;
#else
        const int t0 = 0, t1 = FS_N;
#endif
        for (int ts = t0; ts < t1; ts++) {
            fs_f x[FS_DOF];
            fs_gather(path, ts, x);
#if FS_NOPS > 0
            fs_operator_v(x, opAccum);
#endif
#if FS_BINS
            fs_histogram(x, a.bins);
#endif
#if FS_NCORR > 0
            fs_f *cor = reinterpret_cast<fs_f *>(a.corMeans) + walker * (long)(FS_NCORR * (FS_N / 2));
            for (int corTI = 0; corTI < FS_N / 2; corTI++) {
                int corT = corTI + lid0;
                if (corT >= FS_N / 2) corT -= FS_N / 2;
                const int ts2 = (ts + corT < FS_N) ? ts + corT : ts + corT - FS_N;
                fs_f prod[FS_NCORR1], x2[FS_DOF];
#pragma unroll
                for (int c = 0; c < FS_NCORR1; c++) prod[c] = (fs_f)1.0;
                fs_correlator_v(x, prod);
                fs_gather(path, ts2, x2);
                fs_correlator_v(x2, prod);
#pragma unroll
                for (int c = 0; c < FS_NCORR; c++) cor[c * (FS_N / 2) + corT] += prod[c];
#if FS_PARALLEL
                __syncthreads();
#endif
            }
#endif
        }
#endif
    }
#if FS_NOPS > 0
    fs_f *opm = reinterpret_cast<fs_f *>(a.opMeans);
#pragma unroll
    for (int k = 0; k < FS_NOPS; k++) opm[k + threadId * FS_NOPS] = opAccum[k] * (fs_f)FS_OPNORM;
#endif
    a.accepts[threadId] = accepts;
#if FS_RANLUX
    own.store(a.rl, threadId);
    grp.store(a.grl, threadId);
#else
    own.store(a.seeds, threadId);
#endif
}
#endif  // FS_SAMPLER == 0

// ===========================================================================
// fs_wave_sweep
// ===========================================================================
#if FS_SAMPLER == 1 && FS_SEG <= 64

template <typename T> FS_DEV T fs_lane_read(T v, int src) { return __shfl(v, src, 64); }

FS_DEV float fs_readlane(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
FS_DEV double fs_readlane(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over aligned groups of G lanes (xor butterfly); every lane of a group
// ends with the same bits.
template <int G, typename T> FS_DEV T fs_group_sum(T v)
{
#pragma unroll
    for (int off = 1; off < G; off <<= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Path storage of the wave sampler.  When the walker's whole path fits in
// registers (N >= 64, at most 16 blocks of 64 slices), lane j keeps slices
// j, j + 64, ... for the whole launch together with the potential at each of
// them, so a launch touches HBM once on entry and once on exit; the block
// index is uniform, so the slot of a block is a register-indexed move.
// Otherwise every block is loaded from and stored to the path in memory.
// FS_RESIDENT (1: registers, 0: memory) is decided by the host.
// In memory storage the potential at every slice is cached in potCache
// ([walkers][N], refreshed at the start of each launch), so a move evaluates
// the potential only at its proposed slices, as in register storage.
#if FS_RESIDENT
typedef fs_f fs_slots __attribute__((ext_vector_type(FS_NB)));
#endif

extern "C" __global__ __launch_bounds__(64 * FS_WPB) void fs_wave_sweep(fs_kargs a)
{
    const int lane = threadIdx.x & 63;
    const long w = (long)blockIdx.x * FS_WPB + (threadIdx.x >> 6);
    if (w >= a.nWalkers) return;  // whole wavefronts only

    fs_f *path = reinterpret_cast<fs_f *>(a.paths) + w * (long)(FS_N * FS_DOF);
    const bool active = lane < FS_B;
    fs_f sigtab[FS_S + 1];
#pragma unroll
    for (int n = 1; n <= FS_S; n++) sigtab[n] = fs_sigma(n);
    sigtab[0] = sigtab[1];

    // lane j draws from stream 64 w + j, i.e. the stream follows the slice residue
    fs_stream rng, grp;
#if FS_RANLUX
    rng.load(a.rl, w * 64 + lane);
    grp.load(a.grl, w);
#else
    rng.load(a.seeds, w * 64 + lane);
    grp.load(a.seeds, a.nWalkers * 64 + w);
#endif

    fs_u32 nacc = 0;
    double opS[FS_NOPS1], opC[FS_NOPS1];
#pragma unroll
    for (int k = 0; k < FS_NOPS1; k++) opS[k] = opC[k] = 0.0;

    // dynamic LDS: per wavefront, correlator staging then a private histogram
    extern __shared__ char fs_smem[];
#if FS_NCORR > 0
    fs_f *cbuf = reinterpret_cast<fs_f *>(fs_smem) + (threadIdx.x >> 6) * (FS_N * FS_NCORR);
#endif
#if FS_NHOBS > 0
    unsigned *hbuf = reinterpret_cast<unsigned *>(fs_smem + FS_WPB * (FS_N * FS_NCORR * (int)sizeof(fs_f))) +
                     (threadIdx.x >> 6) * FS_HBINS;
    for (int b = lane; b < FS_HBINS; b += 64) hbuf[b] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
#endif

#if FS_RESIDENT
    fs_slots xs[FS_DOF];
    fs_slots vc;
#pragma unroll 1
    for (int k = 0; k < FS_NB; k++) {
        fs_f x[FS_DOF];
        fs_gather(path, k * 64 + lane, x);
#pragma unroll
        for (int d = 0; d < FS_DOF; d++) xs[d][k] = x[d];
        vc[k] = fs_potential_v(x);
    }
#else
    // lane j only ever touches the cache entries of slices congruent to j mod B
    fs_f *pot = reinterpret_cast<fs_f *>(a.potCache) + w * (long)FS_N;
    if (active) {
#pragma unroll 1
        for (int k = 0; k < FS_NB; k++) {
            fs_f x[FS_DOF];
            fs_gather(path, k * FS_B + lane, x);
            pot[k * FS_B + lane] = fs_potential_v(x);
        }
    }
#endif

    for (int i = 0; i < a.operatorRuns; i++) {
        for (int j = 0; j < a.metroSteps; j++) {
            const int r = (int)(grp.integer() & (fs_u32)(FS_SEG - 1));
            // this lane plays block position p; its slice is r + gB + p
            const int p = active ? ((lane - r) & (FS_B - 1)) : lane;
            const int ip = p & (FS_SEG - 1);
            const bool wrap = lane < r;                 // slice lies in the next block of slots
            const int toPos = (lane + r) & (FS_B - 1);  // lane that plays position `lane`
            const int lvl = ip ? FS_S - __builtin_ctz((unsigned)ip) : 0;
            fs_f sig = sigtab[0];
#pragma unroll
            for (int n = 1; n <= FS_S; n++) sig = (lvl == n) ? sigtab[n] : sig;
            // The Levy bisection of a segment, written in closed form: the
            // midpoint recursion x_m = (x_{m-h} + x_{m+h}) / 2 + sigma_n z_m
            // resolves to the linear interpolation of the two ends plus, for
            // each level n, the level-n node covering position i times the tent
            // 1 - |i - m_n| / h_n (zero at nodes of coarser levels and at the
            // ends).  Same bridge, no per-level masking.
            const int segL = p - ip;
            const int laneL = (segL + r) & (FS_B - 1);
            const int laneR = (segL + FS_SEG + r) & (FS_B - 1);  // lane r when the end is the block end
            const fs_f frac = (fs_f)ip * ((fs_f)1.0 / (fs_f)FS_SEG);
            int srcN[FS_S + 1];
            fs_f wN[FS_S + 1];
#pragma unroll
            for (int n = 1; n <= FS_S; n++) {
                const int h = FS_SEG >> n;
                const int m = (ip & ~(2 * h - 1)) + h;
                srcN[n] = (segL + m + r) & (FS_B - 1);
                wN[n] = (fs_f)1.0 - (fs_f)abs(ip - m) * ((fs_f)1.0 / (fs_f)h);
            }

#if !FS_RESIDENT
            // block g + 1 is loaded while block g is moved: its position 0 is the
            // right end of block g, which no move of this step changes
            fs_f xnext[FS_DOF];
            int snext = (lane + (wrap ? FS_B : 0)) & (FS_N - 1);
            fs_gather(path, snext, xnext);
#endif
#if FS_RESIDENT
#pragma unroll 1
#endif
            for (int g = 0; g < FS_NB; g++) {
                fs_f xo[FS_DOF], xn[FS_DOF], xr[FS_DOF], z[FS_KDRAW];
                fs_f vOld;
#if FS_RESIDENT
                const int g1 = (g + 1) & (FS_NB - 1);
#pragma unroll
                for (int d = 0; d < FS_DOF; d++) {
                    const fs_f a0 = xs[d][g], a1 = xs[d][g1];
                    xo[d] = wrap ? a1 : a0;
                    xr[d] = fs_readlane(a1, r);
                }
                {
                    const fs_f v0 = vc[g], v1 = vc[g1];
                    vOld = wrap ? v1 : v0;
                }
#else
                const int s = snext;
#pragma unroll
                for (int d = 0; d < FS_DOF; d++) xo[d] = xnext[d];
                snext = ((g + 1) * FS_B + lane + (wrap ? FS_B : 0)) & (FS_N - 1);
                fs_gather(path, snext, xnext);
#pragma unroll
                for (int d = 0; d < FS_DOF; d++) xr[d] = fs_readlane(xnext[d], r);
                vOld = pot[s];
#endif
                // every lane draws the same count so buffered streams stay in step
                fs_f u[FS_KDRAW];
#pragma unroll
                for (int k = 0; k < FS_KDRAW; k++) u[k] = rng.uniform();
#pragma unroll
                for (int k = 0; k < FS_KPAIR; k++) {
                    // v == 0 gives a zero displacement, as kernel.c:346-349 does
                    const fs_f v = u[2 * k + 1];
                    const fs_f rho = v != (fs_f)0.0 ? sig * sqrt(-2.0 * log(v)) : (fs_f)0.0;
                    z[2 * k] = rho * fs_sin2pi(u[2 * k]);
                    z[2 * k + 1] = rho * fs_cos2pi(u[2 * k]);
                }
#pragma unroll
                for (int d = 0; d < FS_DOF; d++) {
                    const fs_f xL = fs_lane_read(xo[d], laneL);
                    const fs_f xR = fs_lane_read(p == 0 ? xr[d] : xo[d], laneR);
                    fs_f x = xL + (xR - xL) * frac;
#pragma unroll
                    for (int n = 1; n <= FS_S; n++) x += wN[n] * fs_lane_read(z[d], srcN[n]);
                    xn[d] = x;
                }
                const fs_f vNew = fs_potential_v(xn);
                const fs_f dv = (ip != 0 && active) ? vNew - vOld : (fs_f)0.0;
                // segment sums and decisions in position space (lane L = position L)
                const fs_f dvp = fs_group_sum<FS_SEG>(fs_lane_read(dv, toPos));
                const fs_f u0p = fs_lane_read(u[0], toPos);
                const fs_f ua = fs_lane_read(u0p, lane & ~(FS_SEG - 1));
                const bool accp = !(exp(-eps * dvp) < ua);
                if (accp && (lane & (FS_SEG - 1)) == 0 && active) nacc++;
                const bool acc = fs_lane_read((int)accp, p) != 0;
                const bool put = acc && ip != 0 && active;
#if FS_RESIDENT
                if constexpr (FS_NB == 1) {
#pragma unroll
                    for (int d = 0; d < FS_DOF; d++) xs[d][0] = put ? xn[d] : xs[d][0];
                    vc[0] = put ? vNew : vc[0];
                } else {
#pragma unroll
                    for (int d = 0; d < FS_DOF; d++) {
                        const fs_f a0 = xs[d][g], a1 = xs[d][g1];
                        xs[d][g] = (put && !wrap) ? xn[d] : a0;
                        xs[d][g1] = (put && wrap) ? xn[d] : a1;
                    }
                    const fs_f v0 = vc[g], v1 = vc[g1];
                    vc[g] = (put && !wrap) ? vNew : v0;
                    vc[g1] = (put && wrap) ? vNew : v1;
                }
#else
                if (put) {
#pragma unroll
                    for (int d = 0; d < FS_DOF; d++) path[d * FS_N + s] = xn[d];
                    pot[s] = vNew;
                }
#endif
            }
        }

        // ---- measurement over all N slices (lane j: slices j, j + B, ...) ----
#if FS_NOPS > 0 || FS_BINS || FS_NCORR > 0 || FS_NHOBS > 0
        fs_f part[FS_NOPS1];
#pragma unroll
        for (int k = 0; k < FS_NOPS1; k++) part[k] = (fs_f)0.0;
        if (active) {
#pragma unroll 1
            for (int k = 0; k < FS_NB; k++) {
                const int t = k * FS_B + lane;
                fs_f x[FS_DOF];
#if FS_RESIDENT
#pragma unroll
                for (int d = 0; d < FS_DOF; d++) x[d] = xs[d][k];
#else
                fs_gather(path, t, x);
#endif
#if FS_NOPS > 0
                fs_operator_v(x, part);
#endif
#if FS_BINS
                fs_histogram(x, a.bins);
#endif
#if FS_NHOBS > 0
                {
                    fs_f ob[FS_NHOBS];
                    fs_hist_obs_v(x, ob);
#pragma unroll
                    for (int o = 0; o < FS_NHOBS; o++) {
                        const fs_f v = (ob[o] - (fs_f)FS_HLO) * (fs_f)FS_HINV;
                        if (v >= (fs_f)0.0 && v < (fs_f)FS_HBINS) atomicAdd(hbuf + (int)v, 1u);
                    }
                }
#endif
#if FS_NCORR > 0
                fs_f prod[FS_NCORR1];
#pragma unroll
                for (int c = 0; c < FS_NCORR1; c++) prod[c] = (fs_f)1.0;
                fs_correlator_v(x, prod);
#pragma unroll
                for (int c = 0; c < FS_NCORR; c++) cbuf[c * FS_N + t] = prod[c];
#endif
            }
        }
#if FS_NOPS > 0
#pragma unroll
        for (int k = 0; k < FS_NOPS; k++) {
            const fs_f m = fs_group_sum<FS_B>(part[k]);
            const double y = (double)m - opC[k];
            const double t = opS[k] + y;
            opC[k] = (t - opS[k]) - y;
            opS[k] = t;
            if (a.series != nullptr && lane == 0)
                a.series[((long)(a.seriesOffset + i) * a.nWalkers + w) * FS_NOPS + k] = (float)(m / (fs_f)FS_N);
        }
#endif
#if FS_NCORR > 0
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        {
            double *cb = reinterpret_cast<double *>(a.corMeans) + w * (long)(FS_NCORR * FS_CORLAGS * 2);
#pragma unroll
            for (int c = 0; c < FS_NCORR; c++) {
                const fs_f *cv = cbuf + c * FS_N;
                for (int tau = lane; tau < FS_CORLAGS; tau += 64) {
                    fs_f acc = (fs_f)0.0;
                    for (int t = 0; t < FS_N; t++) acc += cv[t] * cv[(t + tau) & (FS_N - 1)];
                    double *bin = cb + (c * FS_CORLAGS + tau) * 2;
                    const double sum = bin[0], comp = bin[1];
                    const double y = (double)acc - comp;
                    const double t2 = sum + y;
                    bin[1] = (t2 - sum) - y;
                    bin[0] = t2;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#endif
#endif
    }

    // ---- launch outputs ----
#if FS_NHOBS > 0
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int b = lane; b < FS_HBINS; b += 64) {
        const unsigned c = hbuf[b];
        if (c) atomicAdd(a.obsHist + b, c);
    }
#endif
#if FS_RESIDENT
#pragma unroll 1
    for (int k = 0; k < FS_NB; k++) {
#pragma unroll
        for (int d = 0; d < FS_DOF; d++) path[d * FS_N + k * 64 + lane] = xs[d][k];
    }
#endif
    const fs_u32 tot = fs_group_sum<64>(nacc);
    if (lane == 0) a.accepts[w] = tot;
#if FS_NOPS > 0
    if (lane == 0) {
        double *opm = reinterpret_cast<double *>(a.opMeans);
#pragma unroll
        for (int k = 0; k < FS_NOPS; k++) {
            opm[w * FS_NOPS + k] = opS[k] / ((double)a.operatorRuns * (double)FS_N);
            if (a.opTotals != nullptr) {
                double *tk = a.opTotals + (w * FS_NOPS + k) * 2;
                const double y = (opS[k] - opC[k]) - tk[1];
                const double t = tk[0] + y;
                tk[1] = (t - tk[0]) - y;
                tk[0] = t;
            }
        }
    }
#endif
#if FS_RANLUX
    rng.store(a.rl, w * 64 + lane);
    if (lane == 0) grp.store(a.grl, w);
#else
    rng.store(a.seeds, w * 64 + lane);
    if (lane == 0) grp.store(a.seeds, a.nWalkers * 64 + w);
#endif
}
#endif  // FS_SAMPLER == 1

// ===========================================================================
// Helpers for parity tests of the compiled system section
// ===========================================================================
extern "C" __global__ void fs_eval_system(const fs_f *__restrict__ x, double *__restrict__ pot,
                                          double *__restrict__ ops, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fs_f v[FS_DOF];
#pragma unroll
    for (int d = 0; d < FS_DOF; d++) v[d] = x[i * FS_DOF + d];
    pot[i] = (double)fs_potential_v(v);
    fs_f acc[FS_NOPS1];
#pragma unroll
    for (int k = 0; k < FS_NOPS1; k++) acc[k] = (fs_f)0.0;
#if FS_NOPS > 0
    fs_operator_v(v, acc);
#endif
#pragma unroll
    for (int k = 0; k < FS_NOPS; k++) ops[i * FS_NOPS + k] = (double)acc[k];
}
