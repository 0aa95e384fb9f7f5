// fs_abi_args.h -- the kernel-argument block shared by the host engine
// (compiled by hipcc) and the per-system sampler module (compiled by hiprtc).
// Plain C layout: pointers and 32/64-bit scalars only.
#pragma once

struct fs_ranlux;

struct fs_kargs {
    void *paths;          // FLOAT_TYPE [walkers][DOF][N]  (DOF-major per walker)
    unsigned *accepts;    // exact: [threads]; wave: [walkers]
    unsigned *seeds;      // xorshift rows (uint32 x4). exact: [threads+1]; wave: lanes [walkers*64] then groups [walkers]
    unsigned *seedsSnap;  // exact: copy of `seeds` taken before the launch (group rows are read from it)
    fs_ranlux *rl;        // RANLUX states. exact: [threads]; wave: [walkers*64]
    fs_ranlux *grl;       // RANLUX group states. exact: [threads]; wave: [walkers]
    void *oldPaths;       // exact: FLOAT_TYPE [threads][(2^S-1)*DOF]
    void *opMeans;        // exact: FLOAT_TYPE [threads*nOps]; wave: double [walkers][nOps]
    double *opTotals;     // wave, optional: running FP64 (sum, compensation) [walkers][nOps][2]
    float *series;        // wave, optional: per-measurement slice means [measurement][walkers][nOps]
    void *corMeans;       // exact: FLOAT_TYPE [walkers][nCorr][N/2]; wave: double [walkers][nCorr][lags][2]
    unsigned *bins;       // histogram counts
    void *potCache;       // wave, memory storage: FLOAT_TYPE [walkers][N] potential at every slice
    unsigned *obsHist;    // wave: histogram of the scalar observables [hist_bins]
    long long nWalkers;
    int operatorRuns;
    int metroSteps;
    int seriesOffset;     // index of the first measurement of this launch in `series`
    int flags;            // FS_KF_POT_VALID: potCache already holds V at every slice of the current paths
};
#define FS_KF_POT_VALID 1

// Correlator staging of the wave sampler (shared by the host, which sizes the
// dynamic LDS, and the device code).  A lane owns fs_cor_lanelags(lags)
// consecutive lags of a lag chunk of 64 times as many.  Paths of up to
// FS_COR_FULL_N slices stage the whole correlator row in LDS; longer ones
// stage time chunks of FS_COR_CH slices plus the lag window that follows.
#define FS_COR_FULL_N 1024
#define FS_COR_CH 512
constexpr int fs_cor_lanelags(int lags)
{
    int t = 1;
    while (t < 4 && 64 * t < lags) t <<= 1;
    return t;
}
constexpr int fs_cor_stage_len(int N, int lags)
{
    return N <= FS_COR_FULL_N ? N : 2 * FS_COR_CH + 64 * fs_cor_lanelags(lags);
}

// LDS of one wave-sampler wavefront's measurements: correlator staging
// (cor_bytes), then the private histogram of the scalar observables
// (hist_bytes), then the FP64 running operator sums and their Kahan
// compensations (nops x 16 bytes, 8-byte aligned), rounded up to 16 bytes.
constexpr int fs_wave_ops_off(int cor_bytes, int hist_bytes) { return (cor_bytes + hist_bytes + 7) & ~7; }
constexpr int fs_wave_meas_total(int cor_bytes, int hist_bytes, int nops)
{
    return (fs_wave_ops_off(cor_bytes, hist_bytes) + nops * 16 + 15) & ~15;
}

// Persistent batched launch over several system types (one wave per walker).
// Walkers are numbered in dispatch order: type typeOf[d] owns numbers
// [start[d], start[d + 1]); wavefronts take numbers from *counter until
// start[ntypes] is reached.  The run length is the same for every type.
struct fs_batch_args {
    const fs_kargs *types;     // [ntypes], indexed by type id
    const long long *start;    // [ntypes + 1], prefix sums of walkers in dispatch order
    const int *typeOf;         // [ntypes], type id of each dispatch slot
    unsigned long long *counter;
    int ntypes;
    int operatorRuns;
    int metroSteps;
    int seriesOffset;
};
