// fs_wave.h -- the wave-parallel Metropolis bisection sampler (gfx950).
//
// One wavefront owns one walker for a whole launch.  A block of 64
// consecutive slices is spread one slice per lane, so the 2^S - 1 midpoints
// of a segment, the potential evaluations and the action sum all run in
// parallel across lanes and a segment never leaves registers.  Lane j always
// holds the slices congruent to j mod 64 and draws from its own stream; with
// tiling offset r it plays block position (j - r) mod 64.  The offset of
// every step comes from a per-walker group stream (the ParallelizePath scheme
// of FeynSimul/GPUsrc/kernel.c:713-716).  The proposal and the acceptance
// rule are the reference's (kernel.c:258-399); only the assignment of random
// numbers differs.
//
// Everything system-specific comes from a trait class Sys:
//   constants N, S, DOF, NOPS, NCORR, CORLAGS, NHOBS, HBINS, RESIDENT, POTCACHE, BINS,
//   BINRES and EPS, HLO, HINV, XMIN, XMAX, INVBIN;
//   V(x), ops(x, acc), corr(x, prod), hobs(x, out) over one slice x[DOF].
// A module for one system instantiates fs_wave_walker once; the batched
// scheduler instantiates it once per system type.
//
// Path storage: with RESIDENT, lane j keeps slices j, j + 64, ... and their
// potential in registers for the whole launch (HBM touched on entry and
// exit); otherwise blocks stream through the cache hierarchy and the next
// block is prefetched.  With POTCACHE the potential at every slice is cached
// (registers, or potCache [walkers][N] refreshed at the start of each launch),
// so a move evaluates the potential only at its proposed slices; without it
// (memory storage, cheap potentials) the old slice's potential is evaluated
// again, which costs less than reading and writing the cache.
#pragma once

#include "fs_abi_args.h"
#include "fs_lane.h"
#include "fs_rng.h"
#include "fs_stream.h"

// Bytes of LDS one wavefront uses for a system: correlator staging, a
// private histogram of the scalar observables and the FP64 operator sums
// (fs_wave_meas_total); then the group stream's RANLUX state.
template <class Sys> constexpr int fs_wave_cor_bytes()
{
    return Sys::NCORR > 0 ? fs_cor_stage_len(Sys::N, Sys::CORLAGS) * (int)sizeof(FLOAT_TYPE) : 0;
}
template <class Sys> constexpr int fs_wave_hist_bytes() { return Sys::NHOBS > 0 ? Sys::HBINS * 4 : 0; }
template <class Sys> constexpr int fs_wave_meas_bytes()
{
    return fs_wave_meas_total(fs_wave_cor_bytes<Sys>(), fs_wave_hist_bytes<Sys>(), Sys::NOPS);
}
template <class Sys> constexpr int fs_wave_lds_bytes()
{
    return fs_wave_meas_bytes<Sys>() + (FS_RANLUX ? (int)sizeof(fs_ranlux) : 0);
}

FS_DEV void fs_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lag sums acc[k] += sum_t c_t c_{t+tau0+k}, k < CT, for t = 0..nt-1 in
// ascending order (each lag keeps the sequential FP32 order).  bc(t) points
// at c_t .. c_{t+CT-1} (the same words for every lane), win(i) at
// c_{i+tau0} .. c_{i+tau0+CT-1}; both are CT-aligned runs in LDS, read as one
// vector each.  The window of 2 CT values alternates between two register
// halves (the loop is unrolled by two), so a step of CT slices costs two
// vector LDS reads and CT^2 FMAs.  nt is a multiple of 2 CT.
template <int CT, typename F> struct fs_run_of {
    typedef F type __attribute__((ext_vector_type(CT)));
};
template <typename F> struct fs_run_of<1, F> {
    typedef F type;
};
template <int CT, typename F> FS_DEV void fs_load_run(const F *p, F (&v)[CT])
{
    typedef typename fs_run_of<CT, F>::type R;
    const R r = *reinterpret_cast<const R *>(p);
    if constexpr (CT == 1) {
        v[0] = r;
    } else {
#pragma unroll
        for (int k = 0; k < CT; k++) v[k] = r[k];
    }
}
template <int CT, typename F>
FS_DEV void fs_cor_step(F (&acc)[CT], const F (&av)[CT], const F (&lo)[CT], const F (&hi)[CT])
{
#pragma unroll
    for (int j = 0; j < CT; j++) {
#pragma unroll
        for (int k = 0; k < CT; k++) acc[k] += av[j] * (j + k < CT ? lo[j + k] : hi[j + k - CT]);
    }
}
template <int CT, typename F, typename BC, typename WIN>
FS_DEV void fs_cor_accumulate(F (&acc)[CT], int nt, BC bc, WIN win)
{
    F w0[CT], w1[CT], av[CT];
    fs_load_run<CT>(win(0), w0);
#pragma unroll 1
    for (int t = 0; t < nt; t += 2 * CT) {
        fs_load_run<CT>(bc(t), av);
        fs_load_run<CT>(win(t + CT), w1);
        fs_cor_step<CT>(acc, av, w0, w1);
        fs_load_run<CT>(bc(t + CT), av);
        fs_load_run<CT>(win(t + 2 * CT), w0);
        fs_cor_step<CT>(acc, av, w1, w0);
    }
}

template <class Sys>
FS_DEV void fs_wave_walker(const fs_kargs &a, const long wIn, const int lane, char *lds)
{
    const long w = fs_uniform(wIn);
    typedef FLOAT_TYPE F;
    constexpr int N = Sys::N, S = Sys::S, DOF = Sys::DOF;
    constexpr int SEG = 1 << S;
    constexpr int B = N < 64 ? N : 64;
    constexpr int NB = N / B;
    constexpr int KPAIR = (DOF + 1) / 2;
    constexpr int KDRAW = 2 * KPAIR;
    constexpr int NOPS1 = Sys::NOPS > 0 ? Sys::NOPS : 1;
    constexpr int NCORR1 = Sys::NCORR > 0 ? Sys::NCORR : 1;
    static_assert(SEG <= B, "2^S must not exceed min(N, 64)");
    static_assert(!Sys::RESIDENT || (N >= 64 && NB <= 16), "register storage needs 64 <= N <= 1024");

    F *path = reinterpret_cast<F *>(a.paths) + w * (long)(N * DOF);
    // one scalar base per coordinate row (opaque to the compiler, so a row access
    // stays base + 32-bit lane offset instead of 64-bit per-lane arithmetic)
    __attribute__((address_space(1))) F *row[DOF];
#pragma unroll
    for (int d = 0; d < DOF; d++)
        row[d] = (__attribute__((address_space(1))) F *)fs_uniform((long)(path + d * N));
    const bool active = lane < B;
    F sigtab[S + 1];
#pragma unroll
    for (int n = 1; n <= S; n++) sigtab[n] = sqrt(((F)SEG * (F)Sys::EPS / ((F)2.0 * (F)(1 << n))));

    // lane j draws from stream 64 w + j, i.e. the stream follows the slice residue
#if FS_RANLUX
    fs_stream rng;
    fs_wave_stream grp;
    rng.load(a.rl, w * 64 + lane);
    grp.load(a.grl, w, reinterpret_cast<fs_ranlux *>(lds + fs_wave_meas_bytes<Sys>()));
#else
    fs_stream rng, grp;
    rng.load(a.seeds, w * 64 + lane);
    grp.load(a.seeds, a.nWalkers * 64 + w);
#endif

    fs_u32 nacc = 0;
    // the running FP64 operator sums and their Kahan compensations live in LDS
    // (lane 0 updates them once per measurement), not in registers that would
    // stay allocated through the whole sweep loop
    double *opAcc = reinterpret_cast<double *>(lds + fs_wave_ops_off(fs_wave_cor_bytes<Sys>(), fs_wave_hist_bytes<Sys>()));
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 2 * Sys::NOPS; k++) opAcc[k] = 0.0;
    }

    F *cbuf = reinterpret_cast<F *>(lds);
    unsigned *hbuf = reinterpret_cast<unsigned *>(
        lds + (Sys::NCORR > 0 ? fs_cor_stage_len(N, Sys::CORLAGS) * (int)sizeof(F) : 0));
    if constexpr (Sys::NHOBS > 0) {
        for (int b = lane; b < Sys::HBINS; b += 64) hbuf[b] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }

    typedef F slots_t __attribute__((ext_vector_type(NB)));
    slots_t xs[Sys::RESIDENT ? DOF : 1];
    slots_t vc;
    F *pot = reinterpret_cast<F *>(a.potCache) + w * (long)N;
    if constexpr (Sys::RESIDENT) {
#pragma unroll 1
        for (int k = 0; k < NB; k++) {
            F x[DOF];
#pragma unroll
            for (int d = 0; d < DOF; d++) x[d] = fs_at(path, d * N + k * 64 + lane);
#pragma unroll
            for (int d = 0; d < DOF; d++) xs[d][k] = x[d];
            vc[k] = Sys::V(x);
        }
    } else if (Sys::POTCACHE && active && !(a.flags & FS_KF_POT_VALID)) {
        // (skipped when the caller vouches that the cache still holds the potentials the previous launch left)
        // lane j only ever touches the cache entries of slices congruent to j mod B
#pragma unroll 1
        for (int k = 0; k < NB; k++) {
            F x[DOF];
#pragma unroll
            for (int d = 0; d < DOF; d++) x[d] = fs_at(path, d * N + k * B + lane);
            fs_at(pot, k * B + lane) = Sys::V(x);
        }
    }

    for (int i = 0; i < a.operatorRuns; i++) {
        for (int j = 0; j < a.metroSteps; j++) {
            // the group stream is the walker's, so r is wave-uniform: keep it in a scalar register
            const int r = __builtin_amdgcn_readfirstlane((int)grp.below((fs_u32)SEG));
            const int srcN1 = ((SEG >> 1) + r) & (B - 1);  // the level-1 node when SEG == 64
            // this lane plays block position p; its slice is r + gB + p
            const int p = active ? ((lane - r) & (B - 1)) : lane;
            const int ip = p & (SEG - 1);
            const bool wrap = lane < r;              // slice lies in the next block of slots
            const int toPos = (lane + r) & (B - 1);  // lane that plays position `lane`
            // The Levy bisection of a segment, written in closed form: the
            // midpoint recursion x_m = (x_{m-h} + x_{m+h}) / 2 + sigma_n z_m
            // resolves to the linear interpolation of the two ends plus, for
            // each level n, the level-n node covering position i times the tent
            // 1 - |i - m_n| / h_n (zero at nodes of coarser levels and at the
            // ends).  Same bridge, no per-level masking.  The node's sigma_n is
            // folded into the weight, so the normals z are drawn with unit variance.
            const int segL = p - ip;
            const int laneL = (segL + r) & (B - 1);
            const int laneR = (segL + SEG + r) & (B - 1);  // lane r when the end is the block end
            const F frac = (F)ip * ((F)1.0 / (F)SEG);
            int srcN[S + 1];
            F wN[S + 1];
#pragma unroll
            for (int n = 1; n <= S; n++) {
                const int h = SEG >> n;
                const int m = (ip & ~(2 * h - 1)) + h;
                srcN[n] = (segL + m + r) & (B - 1);
                wN[n] = ((F)1.0 - (F)abs(ip - m) * ((F)1.0 / (F)h)) * sigtab[n];
            }
            // With 64-slice blocks the finest node (h = 1) of position p is p + 1 for even p
            // and p itself for odd p, i.e. lane j + 1 (mod 64) or lane j: a wave rotation, no
            // LDS permute.  Its tent weight is split between the two sources.
            constexpr bool FINE_DPP = B == 64;
            const bool pEven = (p & 1) == 0;
            const F wFineNext = pEven ? wN[S] : (F)0.0, wFineOwn = pEven ? (F)0.0 : wN[S];

            // memory storage: block g + 1 is loaded while block g is moved; its
            // position 0 is the right end of block g, which no move of this step changes
            F xnext[DOF];
            int snext = (lane + (wrap ? B : 0)) & (N - 1);
            if constexpr (!Sys::RESIDENT) {
#pragma unroll
                for (int d = 0; d < DOF; d++) xnext[d] = fs_at(row[d], snext);
            }
            // one segment per block: block g's left end (its position 0, on lane r) is block
            // g - 1's right end, which no move of this step changes, so it is carried from
            // block to block in scalar registers instead of read again
            F xEnd[DOF];
            if constexpr (SEG == 64) {
#pragma unroll
                for (int d = 0; d < DOF; d++) {
                    if constexpr (Sys::RESIDENT) xEnd[d] = fs_readlane(xs[d][0], r);
                    else xEnd[d] = fs_readlane(xnext[d], r);
                }
            }
            // two blocks per iteration in memory storage: the loaded block and the
            // stream state then change registers in place, with no copies at the loop edge.
            // Register storage unrolls every block: the blocks of a step are independent (a segment's
            // ends are positions no move of the step changes), so one block's acceptance chain overlaps
            // the next block's draws, and the path's register slots are indexed statically instead of
            // through lane-uniform dynamic register indexing (harmonic oscillator, 32-sweep launches: four
            // blocks 0.120 -> 0.087 ms, eight blocks 0.781 -> 0.485 ms, sixteen 0.298 -> 0.278 ms)
            constexpr int BLOCK_UNROLL = Sys::RESIDENT ? NB : (NB >= 2 ? 2 : 1);
#pragma unroll BLOCK_UNROLL
            for (int g = 0; g < NB; g++) {
                // xb: this lane's slot in block g + 1; on lane r (position 0) that is the
                // right end of the block's last segment, read from there by the bridge
                F xo[DOF], xn[DOF], xb[DOF], z[KDRAW];
                F vOld;
                const int g1 = (g + 1) & (NB - 1);
                int s = 0;
                if constexpr (Sys::RESIDENT) {
#pragma unroll
                    for (int d = 0; d < DOF; d++) {
                        const F a0 = xs[d][g], a1 = xs[d][g1];
                        xo[d] = wrap ? a1 : a0;
                        xb[d] = a1;
                    }
                    const F v0 = vc[g], v1 = vc[g1];
                    vOld = wrap ? v1 : v0;
                } else {
                    s = snext;
#pragma unroll
                    for (int d = 0; d < DOF; d++) xo[d] = xnext[d];
                    snext = ((g + 1) * B + lane + (wrap ? B : 0)) & (N - 1);
#pragma unroll
                    for (int d = 0; d < DOF; d++) xnext[d] = fs_at(row[d], snext);
#pragma unroll
                    for (int d = 0; d < DOF; d++) xb[d] = xnext[d];
                    if constexpr (Sys::POTCACHE) vOld = fs_at(pot, s);
                    else vOld = Sys::V(xo);
                }
                // every lane draws the same count so buffered streams stay in step
                F u[KDRAW];
                rng.template uniforms<KDRAW>(u);
#pragma unroll
                for (int k = 0; k < KPAIR; k++) {
                    // v == 0 gives a zero displacement, as kernel.c:346-349 does; -2 ln v is
                    // written with the base-2 logarithm the hardware computes
                    const F v = u[2 * k + 1];
                    const F rho = v != (F)0.0 ? sqrt((F)(-2.0 * 0.69314718055994531) * log2(v)) : (F)0.0;
                    z[2 * k] = rho * fs_sin2pi(u[2 * k]);
                    z[2 * k + 1] = rho * fs_cos2pi(u[2 * k]);
                }
#pragma unroll
                for (int d = 0; d < DOF; d++) {
                    F x;
                    if constexpr (SEG == 64) {
                        // one segment per block: both ends (lane r of this block and of the
                        // next) and the level-1 node are the same for every lane, so they
                        // come from scalar lane reads instead of LDS permutes
                        const F xL = xEnd[d];
                        const F xR = fs_readlane(xb[d], r);
                        xEnd[d] = xR;
                        // (1 - frac) xL + frac xR as two FMAs that read the scalar ends directly
                        // (written as fma calls: frac xR of one block is frac xL of the next,
                        // and a shared product would cost a multiply and a subtraction)
                        x = fma(-frac, xL, xL);
                        x = fma(frac, xR, x);
                    } else {
                        const F xL = fs_lane_read(xo[d], laneL);
                        const F xR = fs_lane_read(p == 0 ? xb[d] : xo[d], laneR);
                        x = xL + (xR - xL) * frac;
                    }
#pragma unroll
                    for (int n = 1; n <= S; n++) {
                        if (FINE_DPP && n == S) {
                            x += wFineNext * fs_lane_next(z[d]);
                            x += wFineOwn * z[d];
                        } else if (SEG == 64 && n == 1) {
                            x += wN[1] * fs_readlane(z[d], srcN1);
                        } else {
                            x += wN[n] * fs_lane_read(z[d], srcN[n]);
                        }
                    }
                    xn[d] = x;
                }
                const F vNew = Sys::V(xn);
                const F dv = (ip != 0 && active) ? vNew - vOld : (F)0.0;
                // segment sums and decisions in position space (lane L = position L)
                const F dvp = fs_group_sum<SEG>(fs_lane_read(dv, toPos));
                // the acceptance uniform: first draw of the lane that plays the segment's left end
                F ua;
                if constexpr (SEG == 64) ua = fs_readlane(u[0], r);
                else ua = fs_lane_read(u[0], ((lane & ~(SEG - 1)) + r) & (B - 1));
                // exp(-eps dS) as exp2 with log2(e) folded into the constant
                const bool accp = !(exp2((F)(-1.4426950408889634 * (double)Sys::EPS) * dvp) < ua);
                if (accp && (lane & (SEG - 1)) == 0 && active) nacc++;
                // the decision is equal on every position of a segment; lane j takes it at its position p
                const bool acc = ((__ballot(accp) >> p) & 1ull) != 0;
                const bool put = acc && ip != 0 && active;
                if constexpr (Sys::RESIDENT) {
                    if constexpr (NB == 1) {
#pragma unroll
                        for (int d = 0; d < DOF; d++) xs[d][0] = put ? xn[d] : xs[d][0];
                        vc[0] = put ? vNew : vc[0];
                    } else {
#pragma unroll
                        for (int d = 0; d < DOF; d++) {
                            const F a0 = xs[d][g], a1 = xs[d][g1];
                            xs[d][g] = (put && !wrap) ? xn[d] : a0;
                            xs[d][g1] = (put && wrap) ? xn[d] : a1;
                        }
                        const F v0 = vc[g], v1 = vc[g1];
                        vc[g] = (put && !wrap) ? vNew : v0;
                        vc[g1] = (put && wrap) ? vNew : v1;
                    }
                } else if (put) {
                    // a fresh copy of the slice index inside the branch: the stores then form their
                    // 32-bit offset here and take the scalar-base form, instead of reusing the 64-bit
                    // address the loads above left in registers
                    int sl = s;
                    __asm__ volatile("" : "+v"(sl));
#pragma unroll
                    for (int d = 0; d < DOF; d++) fs_at(row[d], sl) = xn[d];
                    if constexpr (Sys::POTCACHE) fs_at(pot, sl) = vNew;
                }
            }
        }

        // ---- measurement over all N slices (lane j: slices j, j + B, ...) ----
        if constexpr (Sys::NOPS > 0 || Sys::BINS || Sys::NCORR > 0 || Sys::NHOBS > 0) {
            F part[NOPS1];
#pragma unroll
            for (int k = 0; k < NOPS1; k++) part[k] = (F)0.0;
            if (active) {
#pragma unroll 1
                for (int k = 0; k < NB; k++) {
                    const int t = k * B + lane;
                    F x[DOF];
                    if constexpr (Sys::RESIDENT) {
#pragma unroll
                        for (int d = 0; d < DOF; d++) x[d] = xs[d][k];
                    } else {
#pragma unroll
                        for (int d = 0; d < DOF; d++) x[d] = fs_at(path, d * N + t);
                    }
                    if constexpr (Sys::NOPS > 0) Sys::ops(x, part);
                    if constexpr (Sys::BINS) {
                        // DOF-dimensional grid histogram over raw coordinates (kernel.c:478-506)
                        int idx = 0, inc = 1;
                        bool ok = true;
#pragma unroll
                        for (int q = 0; q < DOF; q++) {
                            ok = ok && x[q] >= (F)Sys::XMIN && x[q] < (F)Sys::XMAX;
                            idx += (int)((fs_u32)((x[q] - (F)Sys::XMIN) * (F)Sys::INVBIN)) * inc;
                            inc *= Sys::BINRES;
                        }
                        if (ok) atomicAdd(a.bins + idx, 1u);
                    }
                    if constexpr (Sys::NHOBS > 0) {
                        F ob[Sys::NHOBS > 0 ? Sys::NHOBS : 1];
                        Sys::hobs(x, ob);
#pragma unroll
                        for (int o = 0; o < Sys::NHOBS; o++) {
                            const F v = (ob[o] - (F)Sys::HLO) * (F)Sys::HINV;
                            if (v >= (F)0.0 && v < (F)Sys::HBINS) atomicAdd(hbuf + (int)v, 1u);
                        }
                    }
                }
            }
            if constexpr (Sys::NOPS > 0) {
#pragma unroll
                for (int k = 0; k < Sys::NOPS; k++) {
                    const F m = fs_group_sum<B>(part[k]);
                    if (lane == 0) {
                        const double sk = opAcc[2 * k], ck = opAcc[2 * k + 1];
                        const double y = (double)m - ck;
                        const double t = sk + y;
                        opAcc[2 * k + 1] = (t - sk) - y;
                        opAcc[2 * k] = t;
                    }
                    if (a.series != nullptr && lane == 0)
                        a.series[((long)(a.seriesOffset + i) * a.nWalkers + w) * Sys::NOPS + k] = (float)(m / (F)N);
                }
            }
            if constexpr (Sys::NCORR > 0) {
                // two-time correlators C_c(tau) = sum_t c(x_t) c(x_{t+tau}), tau < CORLAGS, one
                // FP32 sum per measurement added to FP64 Kahan bins (sum, compensation)
                constexpr int L = Sys::CORLAGS;
                constexpr int CT = fs_cor_lanelags(L);
                constexpr int LSPAN = 64 * CT;
                constexpr int NLC = (L + LSPAN - 1) / LSPAN;
                constexpr bool FULL = N <= FS_COR_FULL_N;
                static_assert(FULL || !Sys::RESIDENT, "register storage implies a fully staged correlator");
                double *cb = reinterpret_cast<double *>(a.corMeans) + w * (long)(Sys::NCORR * L * 2);
                // correlator c at slice s
                auto cval = [&](int c, int s) -> F {
                    F x[DOF], prod[NCORR1];
#pragma unroll
                    for (int d = 0; d < DOF; d++) x[d] = fs_at(path, d * N + s);
#pragma unroll
                    for (int q = 0; q < NCORR1; q++) prod[q] = (F)1.0;
                    Sys::corr(x, prod);
                    F v = prod[0];
#pragma unroll
                    for (int q = 1; q < NCORR1; q++) v = (c == q) ? prod[q] : v;
                    return v;
                };
                auto bin = [&](int c, int tau0, const F (&acc)[CT]) {
#pragma unroll
                    for (int k = 0; k < CT; k++) {
                        const int tau = tau0 + k;
                        if (tau < L) {
                            double *bn = cb + (c * L + tau) * 2;
                            const double sum = bn[0], comp = bn[1];
                            const double y = (double)acc[k] - comp;
                            const double t2 = sum + y;
                            bn[1] = (t2 - sum) - y;
                            bn[0] = t2;
                        }
                    }
                };
#pragma unroll 1
                for (int c = 0; c < Sys::NCORR; c++) {
                    if constexpr (FULL) {
                        // the whole row c_0..c_{N-1}; lane j stages slices j, j + B, ...
                        if (active) {
                            if constexpr (Sys::RESIDENT) {
#pragma unroll
                                for (int k = 0; k < NB; k++) {
                                    F x[DOF], prod[NCORR1];
#pragma unroll
                                    for (int d = 0; d < DOF; d++) x[d] = xs[d][k];
#pragma unroll
                                    for (int q = 0; q < NCORR1; q++) prod[q] = (F)1.0;
                                    Sys::corr(x, prod);
                                    F v = prod[0];
#pragma unroll
                                    for (int q = 1; q < NCORR1; q++) v = (c == q) ? prod[q] : v;
                                    cbuf[k * B + lane] = v;
                                }
                            } else {
#pragma unroll 1
                                for (int k = 0; k < NB; k++) cbuf[k * B + lane] = cval(c, k * B + lane);
                            }
                        }
                        fs_wave_sync();
#pragma unroll 1
                        for (int lc = 0; lc < NLC; lc++) {
                            const int tau0 = lc * LSPAN + lane * CT;
                            F acc[CT];
#pragma unroll
                            for (int k = 0; k < CT; k++) acc[k] = (F)0.0;
                            fs_cor_accumulate<CT>(acc, N, [&](int t) { return cbuf + t; },
                                                  [&](int i) { return cbuf + ((i + tau0) & (N - 1)); });
                            bin(c, tau0, acc);
                        }
                        fs_wave_sync();
                    } else {
                        // time chunks [t0, t0 + CH): bc = c_{t0..}, wn = c_{t0 + lag chunk start ..}
                        constexpr int CH = FS_COR_CH;
                        F *bcv = cbuf;
                        F *wnv = cbuf + CH;
#pragma unroll 1
                        for (int lc = 0; lc < NLC; lc++) {
                            const int lag0 = lc * LSPAN;
                            const int tau0 = lag0 + lane * CT;
                            F acc[CT];
#pragma unroll
                            for (int k = 0; k < CT; k++) acc[k] = (F)0.0;
#pragma unroll 1
                            for (int t0 = 0; t0 < N; t0 += CH) {
#pragma unroll 1
                                for (int i = lane; i < CH; i += 64) bcv[i] = cval(c, t0 + i);
#pragma unroll 1
                                for (int i = lane; i < CH + LSPAN; i += 64)
                                    wnv[i] = cval(c, (t0 + lag0 + i) & (N - 1));
                                fs_wave_sync();
                                fs_cor_accumulate<CT>(acc, CH, [&](int t) { return bcv + t; },
                                                      [&](int i) { return wnv + i + lane * CT; });
                                fs_wave_sync();
                            }
                            bin(c, tau0, acc);
                        }
                    }
                }
            }
        }
    }

    // ---- launch outputs ----
    if constexpr (Sys::NHOBS > 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int b = lane; b < Sys::HBINS; b += 64) {
            const unsigned c = hbuf[b];
            if (c) atomicAdd(a.obsHist + b, c);
        }
    }
    if constexpr (Sys::RESIDENT) {
#pragma unroll 1
        for (int k = 0; k < NB; k++) {
#pragma unroll
            for (int d = 0; d < DOF; d++) fs_at(path, d * N + k * 64 + lane) = xs[d][k];
        }
    }
    const fs_u32 tot = fs_group_sum<64>(nacc);
    if (lane == 0) a.accepts[w] = tot;
    if constexpr (Sys::NOPS > 0) {
        if (lane == 0) {
            double *opm = reinterpret_cast<double *>(a.opMeans);
#pragma unroll
            for (int k = 0; k < Sys::NOPS; k++) {
                const double sk = opAcc[2 * k], ck = opAcc[2 * k + 1];
                opm[w * Sys::NOPS + k] = sk / ((double)a.operatorRuns * (double)N);
                if (a.opTotals != nullptr) {
                    double *tk = a.opTotals + (w * Sys::NOPS + k) * 2;
                    const double y = (sk - ck) - tk[1];
                    const double t = tk[0] + y;
                    tk[1] = (t - tk[0]) - y;
                    tk[0] = t;
                }
            }
        }
    }
#if FS_RANLUX
    rng.store(a.rl, w * 64 + lane);
    grp.store(a.grl, w);
#else
    rng.store(a.seeds, w * 64 + lane);
    if (lane == 0) grp.store(a.seeds, a.nWalkers * 64 + w);
#endif
}
