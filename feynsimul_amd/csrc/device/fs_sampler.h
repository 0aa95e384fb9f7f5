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
// FS_RESIDENT (wave path storage), FS_POTCACHE (wave potential cache), FS_NHOBS, FS_HBINS, FS_HLO, FS_HINV
// (observable histogram of the wave sampler).
//
// Two samplers share the move:
//
//  fs_exact_sweep -- one work-item per reference work-item, with the
//    reference's work-item numbering, stream rows, draw order and
//    accumulation order (kernel.c:514-837; geometry kernel.py:192-211).  It
//    is the parity anchor against the CPU restatement of the reference.
//
//  fs_wave_sweep -- the throughput sampler (fs_wave.h): one wavefront per
//    walker, one slice per lane, the segment in registers.
#pragma once

#include "fs_abi_args.h"
#include "fs_rng.h"

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

#include "fs_stream.h"

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

#if FS_PATH_SHIFT || FS_SINGLE_NODE
// kinEnergyEst (kernel.c:245-249): half the squared step in units of 1/eps^2.
FS_DEV fs_f fs_kin_est(fs_f a, fs_f b)
{
    const fs_f delta = a - b;
    return 0.5 * (delta * delta) * (fs_f)FS_EPS_INV2;
}
#endif

#if FS_PATH_SHIFT
// shiftPathEnergyDiff (kernel.c:407-473): adds `offset` to coordinate `degree`
// of slices left..right (wrapping past N - 1 when left > right); returns the
// change of kinetic estimator plus potential, accumulated in the reference's order.
FS_DEV fs_f fs_shift_diff(fs_f *__restrict__ path, int degree, fs_f offset, int left, int right)
{
    if (left == right) return (fs_f)0.0;
    fs_f *xd = path + degree * FS_N;
    fs_f kinDiff = (fs_f)0.0, potDiff = (fs_f)0.0;
    int ll, rr;
    if (left < right) {
        ll = left == 0 ? FS_N - 1 : left - 1;
        rr = right == FS_N - 1 ? 0 : right + 1;
    } else {
        ll = left - 1;
        rr = right + 1;
    }
    kinDiff -= fs_kin_est(xd[left], xd[ll]);
    kinDiff -= fs_kin_est(xd[right], xd[rr]);
    const int count = left < right ? right - left + 1 : FS_N - left + right + 1;
    for (int t = 0; t < count; t++) {
        int i = left + t;
        if (i >= FS_N) i -= FS_N;
        fs_f x[FS_DOF];
        fs_gather(path, i, x);
        potDiff -= fs_potential_v(x);
        xd[i] = xd[i] + offset;
        fs_gather(path, i, x);
        potDiff += fs_potential_v(x);
    }
    kinDiff += fs_kin_est(xd[left], xd[ll]);
    kinDiff += fs_kin_est(xd[right], xd[rr]);
    return kinDiff + potDiff;
}
#endif

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
            // per step: path shift, single node, bisection (kernel.c:640-721)
#if FS_PATH_SHIFT
            {
                const int degree = (int)(((fs_u32)a.metroSteps * (fs_u32)i + (fs_u32)j) % (fs_u32)FS_DOF);
                const int left = (int)grp.below((fs_u32)FS_N);
                const int right = (int)grp.below((fs_u32)FS_N);
                const fs_f offset = (fs_f)FS_PSALPHA * ((fs_f)2.0 * own.uniform() - (fs_f)1.0);
                if (exp(-eps * fs_shift_diff(path, degree, offset, left, right)) < own.uniform())
                    fs_shift_diff(path, degree, -offset, left, right);
                else
                    accepts++;
            }
#endif
#if FS_SINGLE_NODE
            {
                const fs_u32 step = (fs_u32)i * (fs_u32)a.metroSteps + (fs_u32)j;
                const int modPoint = (int)(step % (fs_u32)FS_N);
                const int modPathPoint = (int)(step % (fs_u32)(FS_N * FS_DOF));
                int rightIndex = modPathPoint + 1, leftIndex = modPathPoint - 1;
                if (modPoint == 0)
                    leftIndex += FS_N;
                else if (modPoint == FS_N - 1)
                    rightIndex -= FS_N;
                const fs_f oldX = path[modPathPoint];
                const fs_f modX = oldX + ((fs_f)2.0 * own.uniform() - (fs_f)1.0) * (fs_f)FS_ALPHA;
                fs_f diffE = fs_kin_est(modX, path[rightIndex]) + fs_kin_est(modX, path[leftIndex]) -
                             fs_kin_est(oldX, path[leftIndex]) - fs_kin_est(oldX, path[rightIndex]);
                fs_f x[FS_DOF];
                fs_gather(path, modPoint, x);
                diffE -= fs_potential_v(x);
                path[modPathPoint] = modX;
                fs_gather(path, modPoint, x);
                diffE += fs_potential_v(x);
                path[modPathPoint] = oldX;
                if (exp(-eps * diffE) > own.uniform()) {
                    path[modPathPoint] = modX;
                    accepts++;
                }
            }
#endif
#if FS_BISECTION
#if FS_PARALLEL
            const fs_u32 start = (grp.below((fs_u32)FS_SEG) + (fs_u32)(FS_SEG * lid0)) % (fs_u32)FS_N;
#else
            const fs_u32 start = grp.below((fs_u32)FS_N);
#endif
            fs_exact_move(path, oldPath, start, own, accepts);
#endif
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
// fs_wave_sweep: the wave sampler of fs_wave.h for the module's one system
// ===========================================================================
#if FS_SAMPLER == 1
#include "fs_wave.h"

struct fs_sys_main {
    static constexpr int N = FS_N, S = FS_S, DOF = FS_DOF, NOPS = FS_NOPS, NCORR = FS_NCORR,
                         CORLAGS = FS_CORLAGS, NHOBS = FS_NHOBS, HBINS = FS_HBINS, RESIDENT = FS_RESIDENT,
                         POTCACHE = FS_POTCACHE || FS_RESIDENT, BINS = FS_BINS, BINRES = FS_BINRES;
    static constexpr FLOAT_TYPE EPS = eps, HLO = FS_HLO, HINV = FS_HINV, XMIN = FS_XMIN, XMAX = FS_XMAX,
                                INVBIN = FS_INVBIN;
    static FS_DEV FLOAT_TYPE V(const FLOAT_TYPE *x) { return fs_potential_v(x); }
    static FS_DEV void ops(const FLOAT_TYPE *x, FLOAT_TYPE *acc) { fs_operator_v(x, acc); }
    static FS_DEV void corr(const FLOAT_TYPE *x, FLOAT_TYPE *prod) { fs_correlator_v(x, prod); }
    static FS_DEV void hobs(const FLOAT_TYPE *x, FLOAT_TYPE *out) { fs_hist_obs_v(x, out); }
};

#if defined(FS_MIN_WAVES) && FS_MIN_WAVES > 0
#define FS_WAVE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(FS_MIN_WAVES)))
#else
#define FS_WAVE_OCCUPANCY
#endif
extern "C" __global__ __launch_bounds__(64 * FS_WPB) FS_WAVE_OCCUPANCY void fs_wave_sweep(fs_kargs a)
{
    const int lane = threadIdx.x & 63;
    const long w = (long)blockIdx.x * FS_WPB + (threadIdx.x >> 6);
    if (w >= a.nWalkers) return;  // whole wavefronts only
    extern __shared__ char fs_smem[];
    fs_wave_walker<fs_sys_main>(a, w, lane, fs_smem + (threadIdx.x >> 6) * fs_wave_lds_bytes<fs_sys_main>());
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
