// fs_lp_abi.h -- argument blocks of the learned-proposal kernels, shared by
// the host engine (hipcc) and the proposal module (hiprtc).  Plain C layout;
// BF16 arrays are passed as unsigned short.
#pragma once

#define FS_LP_HID 128
#define FS_LP_OFF_W1 0
#define FS_LP_OFF_W2 8192
#define FS_LP_OFF_W3 40960
#define FS_LP_OFF_W4 73728
#define FS_LP_OFF_B1 81920
#define FS_LP_OFF_B2 (FS_LP_OFF_B1 + 512)
#define FS_LP_OFF_B3 (FS_LP_OFF_B2 + 512)
#define FS_LP_OFF_B4 (FS_LP_OFF_B3 + 512)
#define FS_LP_IMG_BYTES (FS_LP_OFF_B4 + 128)
// Bytes of a block's LDS copy of the image at active width A (a multiple of 32): the weight blocks
// and biases of the active units only; FS_LP_IMG_BYTES at A = FS_LP_HID.
#define FS_LP_LDS_IMG_BYTES(A) (1024 * (((A) / 32) * 2 + 2 * ((A) / 32) * ((A) / 16) + (A) / 16) + 12 * (A) + 128)
#define FS_LP_BATCH_QUANTUM 2048  // training batches are multiples of this many rows
#define FS_LP_SPLIT_ROWS 256      // training rows per weight- and bias-gradient partial
#define FS_LP_NBIAS 416           // bias gradients b1, b2, b3 (128 each), b4 (32)

struct fs_lp_args {
    float *paths;           // [walkers][DOF][N]
    const char *img;        // packed inference image (FS_LP_IMG_BYTES)
    unsigned *accepts;      // [walkers], accepted segments are added
    long long nWalkers;
    long long nUnits;       // walkers * (N / 2^S) / SPW
    unsigned key0, key1;    // Philox key (the 64-bit seed)
    unsigned sweep;         // sweep counter (low 32 bits)
    float inScale;
    float *trace;           // optional [walkers][N / 2^S][4]: log q ratio, dK, dV, log alpha (test support)
    float *nodes;           // whole-path move: [walkers][M DOF + 2] new coarse nodes, log q ratio, dKeff
    long long w0;           // whole-path move: first walker of this launch (walkers w0 .. w0 + nMove - 1)
    long long nMove;        // whole-path move: walkers this launch moves
    // whole-path move, harmonic reference bridge (hw = omega^2 / 2 > 0; 0: the Levy bridge): the AR(1)
    // factor rho, its powers q1 = rho, q2, q4, q8 and rhoL = rho^64, log2 rho, bs = sqrt(rho eps) and
    // kc = 1 / (2 V_64), V_64 the variance of a node given the one before under the reference
    float hw, bs, q1, q2, q4, q8, rhoL, kc, l2q;
    const float *brg;       // [2][64]: rho^k and the end weights g_k of the bridge's slices k = 1 .. 64
    // whole-path move, measure != 0: after its move the fine pass (the fused move: after its last move) records one
    // measurement of the kernel's operators on the resulting path, as a measure-only fs_run does: opMeans
    // [walkers][NOPS] = slice mean, opTotals [walkers][NOPS][2] += slice sum (Kahan), series [seriesOffset]
    // [walkers][NOPS] = slice mean (opTotals and series optional)
    double *opMeans, *opTotals;
    float *series;
    long long seriesOffset;
    int measure;
};

struct fs_lp_train_args {
    const float *paths;
    unsigned short *X0;     // BF16 [B][32] features
    float *T;               // [B][3 DOF + 1]: ends a, b, target x, sigma_n
    long long nWalkers;
    int batch;
    unsigned key0, key1;
    unsigned stepLo, stepHi;
    float inScale;
};

// C[M x N] = A[M x K] B[K x N]; element (m, k) of A at A[m lda_m + k lda_k],
// (k, n) of B at B[k ldb_k + n ldb_n], (m, n) of C at m ldc + n.
struct fs_lp_gemm_args {
    const unsigned short *A;
    const unsigned short *B;
    void *C;
    const float *bias;        // [N] or NULL
    const unsigned short *H;  // mode 2: ReLU mask source, same layout as C
    long long lda_m, lda_k, ldb_k, ldb_n, ldc;
    long long zstride;        // mode 0: elements between the partials of consecutive blockIdx.z
    int M, N, K, mode;        // 0 FP32 (+bias) partial per blockIdx.z, 1 BF16 relu(acc + bias), 2 BF16 masked
};

struct fs_lp_adam_args {
    float *P, *m, *v;
    const float *dW1, *dW2, *dW3, *dW4;  // partials [splits][rows][cols] (padded shapes)
    const float *db;                     // bias-gradient partials [splits][FS_LP_NBIAS]
    float *gradOut;                      // optional: summed gradient, logical order
    int splits;
    float lr, b1, b2, aeps, c1, c2;      // c1 = 1 - b1^t, c2 = 1 - b2^t
    int apply;                           // 0: only write gradOut
};
