/*
 * engine_internal.h -- device-buffer layouts and kernel launch entry points
 * of the AMICA engine (internal to libcudamica_amd).
 */
#ifndef AMICA_ENGINE_INTERNAL_H_
#define AMICA_ENGINE_INTERNAL_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amica {

/* Sets the message amica_last_error() returns; returns code. */
int report_error(int code, const char* msg);

/* Per-(h, j, i) FP32 constants of the E-step, table [M][kNumCst][m][NP]. */
enum {
    kCstSb = 0,     /* sqrt(beta) */
    kCstNsbmu = 1,  /* -sqrt(beta) mu            y = sb b + nsbmu */
    kCstRm2 = 2,    /* rho - 2 */
    kCstCq2 = 3,    /* (log alpha + .5 log beta - log2 - lgamma(1 + 1/rho)) log2(e) */
    kCstRsb = 4,    /* rho sqrt(beta)             g = sum_j rsb u sign(y)|y|^(rho-1) */
    kNumCst = 5
};

/* Raw E-step statistics per (h, j, i), in the order they are accumulated. */
enum {
    kRawU = 0,      /* sum u */
    kRawUs = 1,     /* sum u sy          sy = sign(y)|y|^(rho-1) */
    kRawUs2 = 2,    /* sum u sy^2 */
    kRawUm2 = 3,    /* sum u |y|^(rho-2) */
    kRawUr = 4,     /* sum u |y|^rho */
    kRawUr2 = 5,    /* sum u |y|^(2 rho) */
    kRawUrl = 6,    /* sum u |y|^rho log2|y| */
    kStatsPerJ = 7
};
/* Raw statistics per (h, i) after the m*kStatsPerJ block. */
enum { kRawS2 = 0, kRawCnew = 1, kRawGsum = 2, kStatsPerI = 3 };

constexpr int kMaxEstepWaves = 8;   /* M * ceil(n/64) <= 8 for the fused E-step */
/* Zero rows after the T sample rows of x, b and b64: the fused E-step runs
 * whole 8-sample groups and prefetches one tile ahead, so it reads up to 9
 * rows past the end of its range unchecked. */
constexpr int kEstepRowPad = 16;

/* FP64 per-(h, j, i) constants of the FP64 unmixing path, [M][2][m][NP]:
 * sqrt(beta) and -sqrt(beta) mu, so that y is formed in FP64. */
enum { kCst64Sb = 0, kCst64Nsbmu = 1, kNumCst64 = 2 };

/* A work item of a launch over many recordings: recording `rec`, the range
 * [t0, t1) of its samples (or of its 16- or 32-sample tiles), written into
 * its partial slot `slot`. */
struct BatchItem {
    int rec;
    int slot;
    long t0;
    long t1;
};

/* Per-recording operands of the batched unmixing and statistic GEMMs. */
struct UnmixArgs {
    const void* xa;         /* bf16 hi part (or FP8 q), [T][NP] */
    const uint16_t* xlo;    /* bf16 lo part, or null (FP8) */
    const uint16_t* Whi;
    const uint16_t* Wlo;
    const float* wc;
    float* b;
    long T;
    int ldb;
};
struct GxArgs {
    const uint16_t* g8;
    const uint16_t* x8;
    double* part;
};

/* E-step launches (estep.hip): fused, or the split passes */
enum { kEstepFused = 0, kEstepLsum = 1, kEstepStats = 2 };

struct EstepArgs {
    const float* b;        /* [T][ldb] (FP32 unmixing) */
    const double* b64;     /* [T][ldb] (FP64 unmixing), or null */
    const double* cst64;   /* [M][kNumCst64][m][NP] when b64 is set */
    const float* x;        /* [T][NP] whitened data */
    void* g;               /* fp32 [T][ldb], or bf16 [T32/8][ldb][8] (8 samples packed) */
    const float* cst;      /* [M][kNumCst][m][NP] */
    const double* lconst;  /* [M] log gm + ldet + ldetS */
    double* part;          /* [nblocks][pstride] */
    float* ltall;          /* optional [M][T], L_ht - max_h const (natural log) */
    long T;
    long chunk;            /* samples per workgroup, multiple of 8 */
    size_t pstride;
    int n, NP, M, CW, ldb;
    int g_bf16;
    /* split launch (M CW > kMaxEstepWaves): workgroups of mg models x cwg
     * channel blocks, blockIdx.y = model group * ncg + channel group; lq
     * [T][M][ncg] holds each group's sum of q (log2 units), vq [T][M] the
     * model posteriors */
    int split, mg, cwg, ncg, mj;
    float* lq;
    float* vq;
};

/* Workgroup shape of the split E-step for M models of CW channel blocks:
 * models per group, channel blocks per group, channel groups, groups. */
void estep_split_shape(int M, int CW, int* mg, int* cwg, int* ncg, int* ngroups);
int launch_estep(const EstepArgs& a, int m, int nblocks, hipStream_t st);
/* Single-model E-steps of many recordings (BF16 path, same NP) in one
 * launch, one workgroup per work item (device arrays). */
int launch_estep_m1_batch(const EstepArgs* args, const BatchItem* items, int nitems, int m, int NP,
                          hipStream_t st);
/* Allows the E-step variants for m components `lds_bytes` of dynamic LDS. */
int estep_configure(int m, size_t lds_bytes);
/* Dynamic LDS of one E-step workgroup of `threads` threads. */
size_t estep_lds_bytes(int m, int threads, bool m1b64 = false);
/* Resident E-step workgroups per CU (registers and LDS), or -1; m1 for the
 * single-model kernel (M = 1, no per-sample L recorded), split for the
 * split launch's statistics pass. */
int estep_blocks_per_cu(int m, bool b64, int threads, bool m1, bool split = false);

/* GEMMs ------------------------------------------------------------------ */
/* b[t][c] = sum_k x[t][k] W[c][k] - wc[c]    (c < ldb, k < NP) */
int launch_unmix_f64(const double* x, const double* Wg, const double* wc, double* b,
                     long T, int NP, int ldb, int cus, hipStream_t st);
/* b = x_hi W_hi + x_hi W_lo + x_lo W_hi (three MFMAs, ~2^-16 relative) with
 * W given as bf16 hi and lo, x as the split bf16 copy xbf ([T][NP] hi, then
 * [T][NP] lo) or, for NP > 256, float32 x split on the fly. */
/* Plain FP32 unmixing b = x W^T - W c on FP32 MFMA (Wg: [ldb][NP] FP32). */
int launch_unmix_f32(const float* x, const float* Wg, const float* wc, float* b, long T, int NP,
                     int ldb, int cus, hipStream_t st);
int launch_unmix_bf16x3(const float* x, const uint16_t* xbf, const uint16_t* Whi,
                        const uint16_t* Wlo, const float* wc, float* b, long T, int NP, int ldb,
                        int cus, hipStream_t st);
/* Split-BF16 unmixing of many recordings (same NP and ldb, NP <= 256) in one
 * launch, one workgroup per work item of 16-sample tiles. */
int launch_unmix_batch(const UnmixArgs* args, const BatchItem* items, int nitems, int NP, int ldb,
                       hipStream_t st);
/* Packed BF16 statistic GEMM of many recordings (same NP and ldb) in one
 * launch, one workgroup column per work item of 32-sample blocks. */
int launch_gx_bf16_batch(const GxArgs* args, const BatchItem* items, int nitems, int NP, int ldb,
                         hipStream_t st);
/* FP8 path: b = q W_hi + q W_lo with q the FP8 e4m3 data [T][NP] and W~ the
 * split of W diag(scale) (NP <= 256). */
int launch_unmix_fp8(const uint8_t* xq, const uint16_t* Whi, const uint16_t* Wlo, const float* wc,
                     float* b, long T, int NP, int ldb, int cus, hipStream_t st);
/* part[chunk][c][k] = sum_{t in chunk} g[t][c] x[t][k]   (FP64 partials) */
int launch_gx_f32(const float* g, const float* x, double* part, long T, int NP, int ldb,
                  int nchunks, hipStream_t st);
/* BF16 path: g and x packed 8 samples per column ([T32/8][cols][8], T32 =
 * T rounded up to 32, zero padded); part is [nchunks][ldb][NP]. */
int launch_gx_bf16(const uint16_t* g8, const uint16_t* x8, double* part, long T, int NP, int ldb,
                   int nchunks, hipStream_t st);
/* Number of sample chunks (partials) launch_gx_bf16 / launch_gx_f32 should
 * be given for a budget of `cus` CUs (nondecreasing in cus). */
int gx_bf16_chunks(long T, int NP, int ldb, int cus);
int gx_f32_chunks(long T, int NP, int ldb, int cus);
/* x8[(t/8) NP + i][t%8] = xhi[t][i] (the packed B operand of the BF16
 * statistic GEMM); padding stays as allocated (zero). */
int launch_pack8(const uint16_t* xhi, long T, int NP, uint16_t* x8, hipStream_t st);

/* Reductions and the M-step ------------------------------------------------ */
struct DevState {
    int iter;              /* completed iterations since set_state */
    int numdec;
    int no_newt;
    int pad;
    double lrate;
    double lratemax_local;
    double ll_prev;
    double ldetS;
};

struct MstepArgs {
    int n, NP, M, m, ldb, prec_bf16;
    long T;
    int log_cap;
    /* options */
    int do_newton, update_rho, rho_start_iter, newt_start_iter;
    int update_gm, update_alpha, update_A, update_mu, update_beta, do_reparm;
    double lratemax, lnatrate, lratefact, rholrate, rhomin, rhomax;
    int maxdec;
    /* parameters (FP64) */
    double *A, *c, *alpha, *mu, *beta, *rho, *gm, *W, *ldet, *lconst;
    /* GEMM copies of W (rows h*NP + i, columns k): FP64 path Wg64 + wc64,
     * BF16 path Whi + Wlo (bf16) + wc, plain FP32 path Wg32 + wc */
    double* Wg64;          /* [ldb][NP] */
    float* Wg32;           /* [ldb][NP], or null (not the plain FP32 path) */
    double* wc64;          /* [ldb] */
    const float* xscale;   /* [NP] per-channel FP8 scales folded into Whi/Wlo, or null */
    uint16_t* Whi;         /* [ldb][NP] */
    uint16_t* Wlo;         /* [ldb][NP] */
    float* wc;             /* [ldb] */
    float* cst;            /* E-step constants */
    double* cst64;         /* FP64 E-step constants (FP64 path) */
    /* statistics */
    double* stats;         /* canonical, M*stride + 1 */
    long stride;
    /* work */
    double *sigma2, *cnew, *phis, *kappa, *lambda, *B, *work, *lu;
    int* bflag;            /* [M] */
    int* newton;           /* [M] */
    DevState* ds;
    /* logs */
    double* log_ll; double* log_lrate; int* log_bflag; int* log_newton; int* log_lldec;
    double* log_vsum;
};

/* Sum E-step partials and the statistic-GEMM partials, then write the
 * canonical statistics (including phi through the unmixing identity). */
/* Partials are summed in up to kSumGroups groups first: tmp holds
 * kSumGroups * max(pstride, ldb NP) doubles. */
constexpr int kSumGroups = 64;
int launch_reduce(const double* part, int nparts, size_t pstride, const double* gxpart,
                  int nchunks, double* raw, double* gxtot, double* tmp, const MstepArgs& ma,
                  hipStream_t st);
int launch_mstep(const MstepArgs& ma, hipStream_t st);
/* W = A^-1, ldet, GEMM-precision W, W c, E-step constants. */
int launch_prepare(const MstepArgs& ma, hipStream_t st);
/* Allows the LDS inverse its dynamic shared memory (once per engine). */
int mstep_configure(int n);
/* doubles of M-step scratch the inverse uses after the first 8 + M n n */
size_t inverse_work_doubles(int n, int M);
/* Ainv_h = A_h^-1 and ldet_inv[h] = -log|det A_h| for M column-major n x n
 * matrices by the blocked Gauss-Jordan of the M-step; work holds
 * 8 + M n n + inverse_work_doubles(n, M) doubles. */
int launch_gj_inverse(const double* A, double* Ainv, double* ldet_inv, int n, int M, double* work,
                      hipStream_t st);

/* Preprocessing -------------------------------------------------------------- */
int launch_colmean(const float* x, long T, int n, double* part, int nchunks, double* mean,
                   hipStream_t st);
int launch_cov(const float* x, const double* mean, long T, int n, double* part, int nchunks,
               double* C, hipStream_t st);
int launch_eigh(double* C, double* V, double* d, int n, double* work, hipStream_t st);
int launch_sphere_matrix(const double* V, const double* d, const double* C, int n,
                         const float* x, const double* mean, long T, int do_sphere,
                         double* S, double* ldetS_out, double* part, int nchunks,
                         hipStream_t st);
/* x_white = S (x - mean): float32 copy xw, optional FP64 copy xw64 and bf16
 * copy xbf, all [T][NP] with zero padding. */
int launch_whiten(const float* x, const double* mean, const double* S, long T, int n, int NP,
                  float* xw, double* xw64, uint16_t* xbf, hipStream_t st);
int launch_white_from_f64(const double* xin, long T, int n, int NP, float* xw, double* xw64,
                          uint16_t* xbf, hipStream_t st);
int launch_bf16_from_f32(const float* x, size_t count, uint16_t* out, hipStream_t st);
/* FP8 storage: per-channel scales (part: nchunks x NP floats of scratch),
 * xq = e4m3(x / scale), x <- scale dequant(xq), x8 = bf16(dequant(xq))
 * packed 8 samples per channel. */
int launch_quantize_fp8(float* xw, long T, int NP, float* scale, uint8_t* xq, uint16_t* x8,
                        float* part, int nchunks, hipStream_t st);
/* Sinv = S^-1 (LU), A_out = colnorm(Sinv A_h), c_out = Sinv c_h + mean. */
int launch_output_transform(const double* S, const double* mean, const double* A,
                            const double* c, int n, int M, double* lu, double* Sinv,
                            double* A_out, double* c_out, hipStream_t st);

}  // namespace amica

#endif
