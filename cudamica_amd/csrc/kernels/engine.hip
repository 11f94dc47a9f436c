/*
 * engine.hip -- host side of the MI355X AMICA engine (C ABI in
 * include/amica_engine.h).
 *
 * One engine owns one recording on one GPU.  Device memory (sizes for
 * n channels padded to NP = 64*ceil(n/64), ldb = M*NP):
 *   x        float32 [T][NP]   whitened data (+ bf16 copy in BF16 mode)
 *   b        float32 [T][ldb]  unmixed data of all models (b = W x - W c)
 *   g        fp32/bf16 [T][ldb] score g = sum_j sqrt(beta) u f
 *   part     FP64 [nblocks][pstride] E-step partial statistics
 *   gxpart   FP64 [nchunks][ldb][NP] partials of G = sum_t g x^T
 *   params   FP64 A, c, alpha, mu, beta, rho, gm, W, ldet (+ GEMM copy of W)
 * One EM iteration is eight launches on one stream, with every decision of
 * the reference's host loop (LL decrease, Newton test, lrate schedule) made
 * on the device, so an iteration can be replayed as one hipGraph.
 */
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/amica_batch.h"
#include "../../../include/amica_engine.h"
#include "common.h"
#include "engine_internal.h"

using namespace amica;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace

int amica::report_error(int code, const char* msg) { return fail(code, msg); }

namespace {

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t _e = (call);                                                       \
        if (_e != hipSuccess)                                                         \
            return fail(AMICA_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define LCHK(call)                                                                    \
    do {                                                                              \
        if ((call) != 0) {                                                            \
            hipError_t _e = hipGetLastError();                                        \
            return fail(AMICA_E_DEVICE, std::string("launch failed: ") + #call + ": " + \
                                            hipGetErrorString(_e));                   \
        }                                                                             \
    } while (0)

long canonical_stride(int n, int m) { return 1 + 8L * n * m + 2L * n + (long)n * n; }

/* A roctx range over the host code that enqueues one EM stage (rocprofv3
 * --marker-trace shows it and attributes the stage's kernels to it). */
struct StageRange {
    explicit StageRange(const char* name) { roctxRangePushA(name); }
    ~StageRange() { roctxRangePop(); }
    StageRange(const StageRange&) = delete;
    StageRange& operator=(const StageRange&) = delete;
};

}  // namespace

struct amica_engine {
    amica_options o;
    int NP = 0, ldb = 0, CW = 0;
    int bf16 = 0;
    int plain = 0;                 /* plain FP32 path: FP32 MFMA GEMMs, FP32 b and g */
    int cus = 0;
    hipStream_t st = nullptr;
    /* data */
    float* xraw = nullptr;
    float* xw = nullptr;           /* float32 whitened data (both paths) */
    double* xw64 = nullptr;        /* FP64 whitened data (FP32 path) */
    uint16_t* xbf = nullptr;       /* bf16 whitened data, hi then lo part (BF16 path) */
    uint16_t* x8 = nullptr;        /* bf16 hi part packed 8 samples per column (BF16 path) */
    int fp8 = 0;                   /* FP8 storage of the whitened data */
    uint8_t* xq = nullptr;         /* FP8 e4m3 data, [T][NP] */
    float* xscale = nullptr;       /* per-channel FP8 scales, [NP] */
    float* b = nullptr;            /* FP32 unmixed data (BF16 path) */
    double* b64 = nullptr;         /* FP64 unmixed data (FP32 path) */
    void* g = nullptr;
    /* partials */
    int nblocks = 0;
    int nblocks_cap = 0;           /* E-step partial slots allocated */
    /* split E-step (all models' channel blocks exceed one workgroup) */
    int split = 0, sp_mg = 0, sp_cwg = 0, sp_ncg = 0, sp_groups = 1;
    float* lq = nullptr;           /* [T32][M][ncg] per-group channel sums of q */
    float* vq = nullptr;           /* [T32][M] model posteriors */
    int nchunks_cap = 0;           /* statistic-GEMM partial slots allocated */
    int cu_budget = 0;             /* CUs the grids are sized for (default: all) */
    int estep_wg_per_cu = 0;
    long chunk = 0;
    size_t pstride = 0;
    double* part = nullptr;
    int nchunks = 0;
    double* gxpart = nullptr;
    double* raw = nullptr;
    double* sumtmp = nullptr;      /* first-pass group sums of the partials */
    double* gxtot = nullptr;
    /* parameters and work */
    double *A = nullptr, *c = nullptr, *alpha = nullptr, *mu = nullptr, *beta = nullptr,
           *rho = nullptr, *gm = nullptr, *W = nullptr, *ldet = nullptr, *lconst = nullptr;
    double* Wg64 = nullptr;
    float* Wg32 = nullptr;
    double* wc64 = nullptr;
    uint16_t* Whi = nullptr;
    uint16_t* Wlo = nullptr;
    float* wc = nullptr;
    float* cst = nullptr;
    double* cst64 = nullptr;
    double* stats = nullptr;
    long stride = 0;
    double *sigma2 = nullptr, *cnew = nullptr, *phis = nullptr, *kappa = nullptr,
           *lambda = nullptr, *Bm = nullptr, *work = nullptr, *lu = nullptr;
    int *bflag = nullptr, *newton = nullptr;
    DevState* ds = nullptr;
    double *log_ll = nullptr, *log_lrate = nullptr, *log_vsum = nullptr;
    int *log_bflag = nullptr, *log_newton = nullptr, *log_lldec = nullptr;
    float* ltall = nullptr;
    int ltall_on = 0;
    /* preprocessing */
    double *S = nullptr, *mean = nullptr, *ldetS_d = nullptr;
    double ldetS = 0.0;
    int have_data = 0, have_mean = 0, have_white = 0, have_state = 0;
    /* graph */
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    /* timing: one set of 5 events per timed iteration, read back lazily */
    int timing = 0;
    std::vector<std::array<hipEvent_t, 6>> evs;
    size_t ev_used = 0;
    hipEvent_t* ev = nullptr;

    /* Grid sizes for a budget of `budget` CUs: the E-step gets budget x
     * (resident workgroups per CU) workgroups of whole 32-sample chunks, the
     * statistic GEMM its chunk count; both stay within the partial slots
     * allocated for the whole device. */
    void size_grids(int budget) {
        cu_budget = std::max(1, std::min(budget, cus));
        const long T = o.T;
        long target = (long)cu_budget * estep_wg_per_cu;
        if (split) {
            /* grid.y runs the workgroup groups of one sample chunk; the
             * partial slots (one per chunk) stay within 2 GiB */
            target = std::max(1L, target / sp_groups);
            const long by_mem = std::max(1L, (2L << 30) / (long)(pstride_for() * sizeof(double)));
            target = std::min(target, by_mem);
        }
        nblocks = (int)std::max(1L, std::min(target, (T + 31) / 32));
        chunk = (T + nblocks - 1) / nblocks;
        chunk = ((chunk + 31) / 32) * 32;     /* whole packed 8-sample groups */
        nblocks = (int)((T + chunk - 1) / chunk);
        nchunks = bf16 ? gx_bf16_chunks(T, NP, ldb, cu_budget) : gx_f32_chunks(T, NP, ldb, cu_budget);
        if (gexec) {   /* the captured grids are stale */
            (void)hipGraphExecDestroy(gexec);
            (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
        }
    }

    size_t pstride_for() const {
        const int NS = kStatsPerJ * o.m + kStatsPerI;
        return (((size_t)o.M * NS * NP + o.M + 1) + 31) & ~(size_t)31;
    }

    MstepArgs margs() const {
        MstepArgs a;
        std::memset(&a, 0, sizeof(a));
        a.n = o.n; a.NP = NP; a.M = o.M; a.m = o.m; a.ldb = ldb; a.prec_bf16 = bf16;
        a.T = o.T; a.log_cap = o.max_iter;
        a.do_newton = o.do_newton; a.update_rho = o.update_rho;
        a.rho_start_iter = o.rho_start_iter; a.newt_start_iter = o.newt_start_iter;
        a.update_gm = o.update_gm; a.update_alpha = o.update_alpha; a.update_A = o.update_A;
        a.update_mu = o.update_mu; a.update_beta = o.update_beta; a.do_reparm = o.do_reparm;
        a.lratemax = o.lratemax; a.lnatrate = o.lnatrate; a.lratefact = o.lratefact;
        a.rholrate = o.rholrate; a.rhomin = o.rhomin; a.rhomax = o.rhomax; a.maxdec = o.maxdec;
        a.A = A; a.c = c; a.alpha = alpha; a.mu = mu; a.beta = beta; a.rho = rho; a.gm = gm;
        a.W = W; a.ldet = ldet; a.lconst = lconst;
        a.Wg64 = Wg64; a.Wg32 = Wg32; a.wc64 = wc64; a.Whi = Whi; a.Wlo = Wlo; a.wc = wc;
        a.xscale = fp8 ? xscale : nullptr;
        a.cst = cst; a.cst64 = cst64;
        a.stats = stats; a.stride = stride;
        a.sigma2 = sigma2; a.cnew = cnew; a.phis = phis; a.kappa = kappa; a.lambda = lambda;
        a.B = Bm; a.work = work; a.lu = lu;
        a.bflag = bflag; a.newton = newton; a.ds = ds;
        a.log_ll = log_ll; a.log_lrate = log_lrate; a.log_bflag = log_bflag;
        a.log_newton = log_newton; a.log_lldec = log_lldec; a.log_vsum = log_vsum;
        return a;
    }

    EstepArgs eargs(float* lt) const {
        EstepArgs a;
        std::memset(&a, 0, sizeof(a));
        a.b = b; a.b64 = b64; a.cst64 = cst64;
        a.x = xw; a.g = g; a.cst = cst; a.lconst = lconst; a.part = part; a.ltall = lt;
        a.T = o.T; a.chunk = chunk; a.pstride = pstride;
        a.n = o.n; a.NP = NP; a.M = o.M; a.CW = CW; a.ldb = ldb; a.g_bf16 = bf16;
        a.split = split; a.mg = sp_mg; a.cwg = sp_cwg; a.ncg = sp_ncg; a.mj = o.m;
        a.lq = lq; a.vq = vq;
        return a;
    }

    /* W = A^-1, log|det W| and the GEMM copies of W and the E-step constants */
    int enqueue_prepare() {
        LCHK(launch_prepare(margs(), st));
        return 0;
    }

    /* the per-recording E-step partials (nparts slots) and statistic-GEMM
     * partials (nch slots) into the canonical statistics */
    int enqueue_reduce(int nparts, int nch) {
        LCHK(launch_reduce(part, nparts, pstride, gxpart, nch, raw, gxtot, sumtmp, margs(), st));
        return 0;
    }

    /* prepare, unmixing GEMM, E-step, statistic GEMM, reductions */
    int enqueue_estep(bool with_timing) {
        if (with_timing) HIPCHK(hipEventRecord(ev[0], st));
        int rc;
        {
            StageRange r("amica: inverse + GEMM copies of W");
            rc = enqueue_prepare();
        }
        if (rc) return rc;
        if (with_timing) HIPCHK(hipEventRecord(ev[1], st));
        {
        StageRange r("amica: unmixing GEMM");
        if (bf16)
        {
            if (fp8) LCHK(launch_unmix_fp8(xq, Whi, Wlo, wc, b, o.T, NP, ldb, cu_budget, st));
            else LCHK(launch_unmix_bf16x3(xw, xbf, Whi, Wlo, wc, b, o.T, NP, ldb, cu_budget, st));
        }
        else if (plain)
            LCHK(launch_unmix_f32(xw, Wg32, wc, b, o.T, NP, ldb, cu_budget, st));
        else
            LCHK(launch_unmix_f64(xw64, Wg64, wc64, b64, o.T, NP, ldb, cu_budget, st));
        }
        if (with_timing) HIPCHK(hipEventRecord(ev[2], st));
        {
            StageRange r("amica: E-step");
            LCHK(launch_estep(eargs(ltall_on ? ltall : nullptr), o.m, nblocks, st));
        }
        if (with_timing) HIPCHK(hipEventRecord(ev[3], st));
        {
            StageRange r("amica: statistic GEMM");
            if (bf16)
                LCHK(launch_gx_bf16((const uint16_t*)g, x8, gxpart, o.T, NP, ldb, nchunks, st));
            else
                LCHK(launch_gx_f32((const float*)g, xw, gxpart, o.T, NP, ldb, nchunks, st));
        }
        if (with_timing) HIPCHK(hipEventRecord(ev[4], st));
        StageRange r("amica: reductions");
        return enqueue_reduce(nblocks, nchunks);
    }

    /* The batch's launches over many recordings cover the BF16 single-model
     * path: unmix_lds_kernel shapes (NP <= 256), estep_m1_kernel, the packed
     * statistic GEMM. */
    bool groupable() const {
        return bf16 && !fp8 && o.M == 1 && !split && !ltall_on && NP <= 256;
    }

    int enqueue_mstep() {
        StageRange r("amica: M-step");
        LCHK(launch_mstep(margs(), st));
        return 0;
    }

    int enqueue_iteration(bool with_timing) {
        StageRange r("amica: EM iteration");
        int rc = enqueue_estep(with_timing);
        if (rc) return rc;
        rc = enqueue_mstep();
        if (rc) return rc;
        if (with_timing) HIPCHK(hipEventRecord(ev[5], st));
        return 0;
    }

    int next_event_set() {
        if (ev_used == evs.size()) {
            std::array<hipEvent_t, 6> s;
            for (auto& v : s) HIPCHK(hipEventCreate(&v));
            evs.push_back(s);
        }
        ev = evs[ev_used++].data();
        return 0;
    }
};

extern "C" {

const char* amica_last_error(void) { return g_last_error.c_str(); }

void amica_options_default(amica_options* o, int n, long T, int M, int m) {
    std::memset(o, 0, sizeof(*o));
    o->n = n; o->T = T; o->M = M; o->m = m;
    o->precision = AMICA_PREC_FP32;
    o->max_iter = 200;
    o->device = 0;
    o->do_newton = 1; o->update_rho = 1; o->rho_start_iter = 1; o->newt_start_iter = 50;
    o->update_gm = o->update_alpha = o->update_A = o->update_mu = o->update_beta = 1;
    o->do_reparm = 1;
    o->lrate0 = 0.1; o->lratemax = 1.0; o->lnatrate = 0.1; o->lratefact = 0.5;
    o->rholrate = 0.1; o->rhomin = 1.0; o->rhomax = 2.0;
    o->maxdec = 3;
}

int amica_engine_create(const amica_options* opts, amica_engine** out) {
    if (!opts || !out) return fail(AMICA_E_INVALID, "null argument");
    const amica_options& o = *opts;
    if (o.n < 1 || o.n > 1024) return fail(AMICA_E_INVALID, "n must be in [1, 1024]");
    if (o.T < 2) return fail(AMICA_E_INVALID, "T must be >= 2");
    if (o.M < 1 || o.M > AMICA_MAX_M) return fail(AMICA_E_INVALID, "M must be in [1, 16]");
    if (o.m < 1 || o.m > AMICA_MAX_MJ) return fail(AMICA_E_INVALID, "m must be in [1, 8]");
    if (o.precision != AMICA_PREC_FP32 && o.precision != AMICA_PREC_FP32_PLAIN &&
        o.precision != AMICA_PREC_BF16 && o.precision != AMICA_PREC_FP8)
        return fail(AMICA_E_INVALID, "precision must be 32 (fp32), 33 (plain fp32), 16 (bf16) or 8 (fp8)");
    const int NP = ((o.n + 63) / 64) * 64;
    const int CW = NP / 64;
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (o.device < 0 || o.device >= ndev) return fail(AMICA_E_INVALID, "no such HIP device");
    HIPCHK(hipSetDevice(o.device));

    amica_engine* e = new amica_engine();
    e->o = o;
    if (e->o.max_iter < 1) e->o.max_iter = 1;
    e->NP = NP;
    e->CW = CW;
    e->ldb = o.M * NP;
    /* the fused E-step holds every model's channel blocks in one workgroup
     * of at most kMaxEstepWaves waves; larger shapes run the split launch */
    e->split = o.M * CW > kMaxEstepWaves;
    if (e->split) estep_split_shape(o.M, CW, &e->sp_mg, &e->sp_cwg, &e->sp_ncg, &e->sp_groups);
    e->bf16 = (o.precision == AMICA_PREC_BF16 || o.precision == AMICA_PREC_FP8);
    e->fp8 = (o.precision == AMICA_PREC_FP8);
    e->plain = (o.precision == AMICA_PREC_FP32_PLAIN);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, o.device));
    e->cus = prop.multiProcessorCount;
    HIPCHK(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking));

    const long T = o.T;
    const int n = o.n, M = o.M, m = o.m;
    auto alloc = [&](void** p, size_t bytes) -> int {
        if (bytes == 0) bytes = 8;
        hipError_t r = hipMalloc(p, bytes);
        if (r != hipSuccess) return fail(AMICA_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        r = hipMemset(*p, 0, bytes);
        if (r != hipSuccess) return fail(AMICA_E_DEVICE, "hipMemset failed");
        return 0;
    };
#define ALLOC(ptr, bytes) do { int _r = alloc((void**)&(ptr), (bytes)); if (_r) { amica_engine_destroy(e); return _r; } } while (0)
    const size_t Tp = (size_t)T + kEstepRowPad;   /* rows the E-step may read */
    ALLOC(e->xw, Tp * NP * sizeof(float));
    if (e->fp8) {
        ALLOC(e->xq, (size_t)T * NP);
        ALLOC(e->xscale, (size_t)NP * sizeof(float));
        ALLOC(e->b, Tp * e->ldb * sizeof(float));
    } else if (e->bf16) {
        ALLOC(e->xbf, 2 * (size_t)T * NP * sizeof(uint16_t));   /* hi, then lo */
        ALLOC(e->b, Tp * e->ldb * sizeof(float));
    } else if (e->plain) {
        ALLOC(e->b, Tp * e->ldb * sizeof(float));
    } else {
        ALLOC(e->xw64, (size_t)T * NP * sizeof(double));
        ALLOC(e->b64, Tp * e->ldb * sizeof(double));
    }
    /* BF16: g and x packed [T32/8][cols][8] for the statistic GEMM; the
     * padding samples up to T32 stay zero */
    const long T32 = (T + 31) / 32 * 32;
    if (e->bf16) {
        ALLOC(e->x8, (size_t)T32 * NP * sizeof(uint16_t));
        ALLOC(e->g, (size_t)T32 * e->ldb * sizeof(uint16_t));
    } else {
        ALLOC(e->g, (size_t)T * e->ldb * sizeof(float));
    }
    if (e->split) {
        ALLOC(e->lq, (size_t)T32 * M * e->sp_ncg * sizeof(float));
        ALLOC(e->vq, (size_t)T32 * M * sizeof(float));
    }
    /* E-step grid: exactly as many workgroups as the CUs hold at once, each
     * streaming one contiguous range of samples. */
    const int waves = e->split ? e->sp_mg * e->sp_cwg : M * CW;
    {
        const int rc = estep_configure(m, estep_lds_bytes(m, 64 * waves, M == 1 && !e->split && !e->bf16 &&
                                                                           !e->plain));
        if (rc) {
            amica_engine_destroy(e);
            return fail(AMICA_E_DEVICE, "E-step LDS configuration failed");
        }
    }
    if (mstep_configure(n)) {
        amica_engine_destroy(e);
        return fail(AMICA_E_DEVICE, "M-step LDS configuration failed");
    }
    const int wg_per_cu = estep_blocks_per_cu(m, e->b64 != nullptr, 64 * waves, M == 1 && !e->split,
                                              e->split != 0);
    if (wg_per_cu < 1) {
        amica_engine_destroy(e);
        return fail(AMICA_E_DEVICE, "E-step occupancy query failed");
    }
    e->estep_wg_per_cu = wg_per_cu;
    /* partial slots for the largest grids (the whole device) */
    e->pstride = e->pstride_for();
    e->size_grids(e->cus);
    e->nblocks_cap = e->nblocks;
    e->nchunks_cap = e->nchunks;
    ALLOC(e->part, (size_t)e->nblocks_cap * e->pstride * sizeof(double));
    ALLOC(e->gxpart, (size_t)e->nchunks_cap * e->ldb * NP * sizeof(double));
    ALLOC(e->raw, e->pstride * sizeof(double));
    ALLOC(e->sumtmp, (size_t)kSumGroups * std::max(e->pstride, (size_t)e->ldb * NP) * sizeof(double));
    ALLOC(e->gxtot, (size_t)e->ldb * NP * sizeof(double));
    const size_t nn = (size_t)n * n, nm = (size_t)n * m;
    ALLOC(e->A, M * nn * sizeof(double));
    ALLOC(e->c, M * (size_t)n * sizeof(double));
    ALLOC(e->alpha, M * nm * sizeof(double));
    ALLOC(e->mu, M * nm * sizeof(double));
    ALLOC(e->beta, M * nm * sizeof(double));
    ALLOC(e->rho, M * nm * sizeof(double));
    ALLOC(e->gm, M * sizeof(double));
    ALLOC(e->W, M * nn * sizeof(double));
    ALLOC(e->ldet, M * sizeof(double));
    ALLOC(e->lconst, M * sizeof(double));
    if (e->bf16) {
        ALLOC(e->Whi, (size_t)e->ldb * NP * sizeof(uint16_t));
        ALLOC(e->Wlo, (size_t)e->ldb * NP * sizeof(uint16_t));
    } else if (e->plain) {
        ALLOC(e->Wg32, (size_t)e->ldb * NP * sizeof(float));
    } else {
        ALLOC(e->Wg64, (size_t)e->ldb * NP * sizeof(double));
        ALLOC(e->wc64, (size_t)e->ldb * sizeof(double));
        ALLOC(e->cst64, (size_t)M * kNumCst64 * m * NP * sizeof(double));
    }
    ALLOC(e->wc, (size_t)e->ldb * sizeof(float));
    ALLOC(e->cst, (size_t)M * kNumCst * m * NP * sizeof(float));
    e->stride = canonical_stride(n, m);
    ALLOC(e->stats, ((size_t)M * e->stride + 1) * sizeof(double));
    ALLOC(e->sigma2, M * (size_t)n * sizeof(double));
    ALLOC(e->cnew, M * (size_t)n * sizeof(double));
    ALLOC(e->phis, M * nn * sizeof(double));
    ALLOC(e->kappa, M * (size_t)n * sizeof(double));
    ALLOC(e->lambda, M * (size_t)n * sizeof(double));
    ALLOC(e->Bm, M * nn * sizeof(double));
    ALLOC(e->work, (8 + M * nn + inverse_work_doubles(n, M)) * sizeof(double));
    ALLOC(e->lu, std::max((size_t)M, (size_t)2) * nn * sizeof(double));
    ALLOC(e->bflag, M * sizeof(int));
    ALLOC(e->newton, M * sizeof(int));
    ALLOC(e->ds, sizeof(DevState));
    const size_t cap = e->o.max_iter;
    ALLOC(e->log_ll, cap * sizeof(double));
    ALLOC(e->log_lrate, cap * sizeof(double));
    ALLOC(e->log_vsum, cap * M * sizeof(double));
    ALLOC(e->log_bflag, cap * M * sizeof(int));
    ALLOC(e->log_newton, cap * M * sizeof(int));
    ALLOC(e->log_lldec, cap * sizeof(int));
    ALLOC(e->S, nn * sizeof(double));
    ALLOC(e->mean, n * sizeof(double));
    ALLOC(e->ldetS_d, sizeof(double));
#undef ALLOC
    *out = e;
    return 0;
}

int amica_engine_destroy(amica_engine* e) {
    if (!e) return 0;
    (void)hipSetDevice(e->o.device);
    if (e->st) (void)hipStreamSynchronize(e->st);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    void* ptrs[] = {e->xraw, e->xw, e->xw64, e->xbf, e->x8, e->xq, e->xscale, e->b, e->b64, e->g, e->part, e->gxpart,
                    e->raw, e->sumtmp, e->gxtot, e->A, e->c, e->alpha, e->mu, e->beta, e->rho, e->gm, e->W,
                    e->ldet, e->lconst, e->Wg64, e->Wg32, e->wc64, e->Whi, e->Wlo, e->wc, e->cst,
                    e->cst64, e->stats, e->sigma2, e->cnew, e->phis,
                    e->kappa, e->lambda, e->Bm, e->work, e->lu, e->bflag, e->newton, e->ds,
                    e->log_ll, e->log_lrate, e->log_vsum, e->log_bflag, e->log_newton,
                    e->log_lldec, e->ltall, e->S, e->mean, e->ldetS_d, e->lq, e->vq};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& s : e->evs)
        for (auto& v : s) (void)hipEventDestroy(v);
    if (e->st) (void)hipStreamDestroy(e->st);
    delete e;
    return 0;
}

int amica_engine_cu_count(amica_engine* e) { return e ? e->cus : 0; }

int amica_engine_set_cu_budget(amica_engine* e, int cus) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    e->size_grids(cus > 0 ? cus : e->cus);
    return 0;
}

int amica_engine_load_raw(amica_engine* e, const float* x, int x_on_device) {
    if (!e || !x) return fail(AMICA_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->o.device));
    const size_t bytes = (size_t)e->o.T * e->o.n * sizeof(float);
    if (!e->xraw) HIPCHK(hipMalloc((void**)&e->xraw, bytes));
    HIPCHK(hipMemcpyAsync(e->xraw, x, bytes, x_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    e->have_data = 1;
    return 0;
}

int amica_engine_center(amica_engine* e) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    if (!e->have_data) return fail(AMICA_E_STATE, "no raw data loaded");
    HIPCHK(hipSetDevice(e->o.device));
    const int n = e->o.n;
    const long T = e->o.T;
    const int nchunks = (int)std::max(1L, std::min((long)e->cus * 2, T / 64 + 1));
    double* part = nullptr;
    HIPCHK(hipMalloc((void**)&part, (size_t)nchunks * n * sizeof(double)));
    const int rc = launch_colmean(e->xraw, T, n, part, nchunks, e->mean, e->st);
    const hipError_t se = hipStreamSynchronize(e->st);
    (void)hipFree(part);
    if (rc) return fail(AMICA_E_DEVICE, "centering launch failed");
    if (se != hipSuccess) return fail(AMICA_E_DEVICE, std::string("centering: ") + hipGetErrorString(se));
    e->have_mean = 1;
    return 0;
}

/* ldetS enters the log-likelihood constant through the device state; data
 * whitened after set_state must reach it there too */
static int publish_ldetS(amica_engine* e) {
    if (!e->have_state) return 0;
    HIPCHK(hipMemcpy((char*)e->ds + offsetof(DevState, ldetS), &e->ldetS, sizeof(double),
                     hipMemcpyHostToDevice));
    return 0;
}

int amica_engine_whiten(amica_engine* e, int do_sphere) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    if (!e->have_mean) return fail(AMICA_E_STATE, "whiten before center");
    HIPCHK(hipSetDevice(e->o.device));
    const int n = e->o.n;
    const long T = e->o.T;
    const int nchunks = (int)std::max(1L, std::min((long)e->cus * 2, T / 64 + 1));
    double *part = nullptr, *C = nullptr, *V = nullptr, *d = nullptr;
    /* scratch of the covariance and sphering partials, and of the FP8 channel maxima */
    const size_t partn = std::max(std::max((size_t)nchunks * n * n, (size_t)nchunks * 2 * n),
                                  (size_t)nchunks * e->NP);
    HIPCHK(hipMalloc((void**)&part, partn * sizeof(double)));
    HIPCHK(hipMalloc((void**)&C, (size_t)n * n * sizeof(double)));
    HIPCHK(hipMalloc((void**)&V, (size_t)n * n * sizeof(double)));
    HIPCHK(hipMalloc((void**)&d, (size_t)n * sizeof(double)));
    int rc = 0;
    do {
        if (do_sphere == 1) {
            if (launch_cov(e->xraw, e->mean, T, n, part, nchunks, C, e->st)) { rc = -1; break; }
            if (launch_eigh(C, V, d, n, nullptr, e->st)) { rc = -1; break; }
        }
        if (launch_sphere_matrix(V, d, C, n, e->xraw, e->mean, T, do_sphere, e->S, e->ldetS_d,
                                 part, nchunks, e->st)) { rc = -1; break; }
        if (launch_whiten(e->xraw, e->mean, e->S, T, n, e->NP, e->xw, e->xw64, e->xbf, e->st)) { rc = -1; break; }
        if (e->x8 && !e->fp8 && launch_pack8(e->xbf, T, e->NP, e->x8, e->st)) { rc = -1; break; }
        if (e->fp8 && launch_quantize_fp8(e->xw, T, e->NP, e->xscale, e->xq, e->x8, (float*)part,
                                          nchunks, e->st)) { rc = -1; break; }
    } while (0);
    const hipError_t se = hipStreamSynchronize(e->st);
    (void)hipFree(part); (void)hipFree(C); (void)hipFree(V); (void)hipFree(d);
    if (rc) return fail(AMICA_E_DEVICE, std::string("sphering launch failed: ") + hipGetErrorString(hipGetLastError()));
    if (se != hipSuccess) return fail(AMICA_E_DEVICE, std::string("sphering: ") + hipGetErrorString(se));
    HIPCHK(hipMemcpy(&e->ldetS, e->ldetS_d, sizeof(double), hipMemcpyDeviceToHost));
    e->have_white = 1;
    return publish_ldetS(e);
}

int amica_engine_preprocess(amica_engine* e, int do_sphere) {
    int rc = amica_engine_center(e);
    if (rc) return rc;
    return amica_engine_whiten(e, do_sphere);
}

int amica_engine_get_sphere(amica_engine* e, double* S, double* mean, double* ldetS) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    HIPCHK(hipSetDevice(e->o.device));
    const int n = e->o.n;
    if (S) HIPCHK(hipMemcpy(S, e->S, (size_t)n * n * sizeof(double), hipMemcpyDeviceToHost));
    if (mean) HIPCHK(hipMemcpy(mean, e->mean, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    if (ldetS) *ldetS = e->ldetS;
    return 0;
}

int amica_engine_load_white(amica_engine* e, const double* xw, double ldetS) {
    if (!e || !xw) return fail(AMICA_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->o.device));
    const size_t cnt = (size_t)e->o.T * e->o.n;
    double* tmp = nullptr;
    HIPCHK(hipMalloc((void**)&tmp, cnt * sizeof(double)));
    HIPCHK(hipMemcpy(tmp, xw, cnt * sizeof(double), hipMemcpyHostToDevice));
    int rc = launch_white_from_f64(tmp, e->o.T, e->o.n, e->NP, e->xw, e->xw64, e->xbf, e->st);
    float* qpart = nullptr;
    const int qchunks = (int)std::max(1L, std::min((long)e->cus * 2, e->o.T / 64 + 1));
    if (!rc && e->fp8) {
        if (hipMalloc((void**)&qpart, (size_t)qchunks * e->NP * sizeof(float)) != hipSuccess) rc = -1;
        else rc = launch_quantize_fp8(e->xw, e->o.T, e->NP, e->xscale, e->xq, e->x8, qpart, qchunks, e->st);
    } else if (!rc && e->x8) {
        rc = launch_pack8(e->xbf, e->o.T, e->NP, e->x8, e->st);
    }
    hipError_t se = hipStreamSynchronize(e->st);
    (void)hipFree(tmp);
    if (qpart) (void)hipFree(qpart);
    if (rc || se != hipSuccess) return fail(AMICA_E_DEVICE, "load_white failed");
    e->ldetS = ldetS;
    /* S and mean are unknown for externally whitened data: identity, zero */
    std::vector<double> I((size_t)e->o.n * e->o.n, 0.0);
    for (int i = 0; i < e->o.n; ++i) I[(size_t)i * e->o.n + i] = 1.0;
    HIPCHK(hipMemcpy(e->S, I.data(), I.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(e->mean, 0, e->o.n * sizeof(double)));
    e->have_white = 1;
    return publish_ldetS(e);
}

int amica_engine_get_white(amica_engine* e, float* x) {
    if (!e || !x) return fail(AMICA_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipMemcpy2D(x, (size_t)e->o.n * sizeof(float), e->xw, (size_t)e->NP * sizeof(float),
                       (size_t)e->o.n * sizeof(float), e->o.T, hipMemcpyDeviceToHost));
    return 0;
}

int amica_engine_set_state(amica_engine* e, const double* A, const double* c, const double* alpha,
                           const double* mu, const double* beta, const double* rho,
                           const double* gm) {
    if (!e || !A || !c || !alpha || !mu || !beta || !rho || !gm)
        return fail(AMICA_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->o.device));
    const int n = e->o.n, M = e->o.M, m = e->o.m;
    const size_t nn = (size_t)n * n, nm = (size_t)n * m;
    HIPCHK(hipStreamSynchronize(e->st));
    HIPCHK(hipMemcpy(e->A, A, M * nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->c, c, M * (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->alpha, alpha, M * nm * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->mu, mu, M * nm * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->beta, beta, M * nm * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->rho, rho, M * nm * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->gm, gm, M * sizeof(double), hipMemcpyHostToDevice));
    DevState ds;
    std::memset(&ds, 0, sizeof(ds));
    ds.lrate = e->o.lrate0;
    ds.lratemax_local = e->o.lratemax;
    ds.ll_prev = -HUGE_VAL;
    ds.ldetS = e->ldetS;
    HIPCHK(hipMemcpy(e->ds, &ds, sizeof(ds), hipMemcpyHostToDevice));
    e->have_state = 1;
    return 0;
}

int amica_engine_get_state(amica_engine* e, double* A, double* c, double* alpha, double* mu,
                           double* beta, double* rho, double* gm, double* W) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    const int n = e->o.n, M = e->o.M, m = e->o.m;
    const size_t nn = (size_t)n * n, nm = (size_t)n * m;
    if (A) HIPCHK(hipMemcpy(A, e->A, M * nn * sizeof(double), hipMemcpyDeviceToHost));
    if (c) HIPCHK(hipMemcpy(c, e->c, M * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    if (alpha) HIPCHK(hipMemcpy(alpha, e->alpha, M * nm * sizeof(double), hipMemcpyDeviceToHost));
    if (mu) HIPCHK(hipMemcpy(mu, e->mu, M * nm * sizeof(double), hipMemcpyDeviceToHost));
    if (beta) HIPCHK(hipMemcpy(beta, e->beta, M * nm * sizeof(double), hipMemcpyDeviceToHost));
    if (rho) HIPCHK(hipMemcpy(rho, e->rho, M * nm * sizeof(double), hipMemcpyDeviceToHost));
    if (gm) HIPCHK(hipMemcpy(gm, e->gm, M * sizeof(double), hipMemcpyDeviceToHost));
    if (W) HIPCHK(hipMemcpy(W, e->W, M * nn * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

static int check_ready(amica_engine* e) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    if (!e->have_white) return fail(AMICA_E_STATE, "no whitened data (load_raw + preprocess, or load_white)");
    if (!e->have_state) return fail(AMICA_E_STATE, "no model state (set_state)");
    return 0;
}

int amica_engine_iterate(amica_engine* e, int k) {
    int rc = check_ready(e);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->o.device));
    for (int it = 0; it < k; ++it) {
        if (e->timing) {
            rc = e->next_event_set();
            if (rc) return rc;
        }
        rc = e->enqueue_iteration(e->timing != 0);
        if (rc) return rc;
    }
    return 0;
}

int amica_engine_iterate_graph(amica_engine* e, int k) {
    int rc = check_ready(e);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->o.device));
    if (!e->gexec) {
        HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
        rc = e->enqueue_iteration(false);
        hipGraph_t gr = nullptr;
        hipError_t ce = hipStreamEndCapture(e->st, &gr);
        if (rc) return rc;
        if (ce != hipSuccess) return fail(AMICA_E_DEVICE, std::string("capture: ") + hipGetErrorString(ce));
        e->graph = gr;
        HIPCHK(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
    }
    for (int it = 0; it < k; ++it) HIPCHK(hipGraphLaunch(e->gexec, e->st));
    return 0;
}

int amica_engine_sync(amica_engine* e) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
}

int amica_engine_iterations(amica_engine* e) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    DevState ds;
    if (hipSetDevice(e->o.device) != hipSuccess) return AMICA_E_DEVICE;
    if (hipStreamSynchronize(e->st) != hipSuccess) return AMICA_E_DEVICE;
    if (hipMemcpy(&ds, e->ds, sizeof(ds), hipMemcpyDeviceToHost) != hipSuccess) return AMICA_E_DEVICE;
    return ds.iter;
}

int amica_engine_get_log(amica_engine* e, int n_iter, double* ll, double* lrate, int* bflag,
                         int* newton, int* lldec, double* vsum) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    if (n_iter < 0 || n_iter > e->o.max_iter) return fail(AMICA_E_INVALID, "n_iter exceeds the log capacity");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    const size_t M = e->o.M;
    if (ll) HIPCHK(hipMemcpy(ll, e->log_ll, n_iter * sizeof(double), hipMemcpyDeviceToHost));
    if (lrate) HIPCHK(hipMemcpy(lrate, e->log_lrate, n_iter * sizeof(double), hipMemcpyDeviceToHost));
    if (bflag) HIPCHK(hipMemcpy(bflag, e->log_bflag, n_iter * M * sizeof(int), hipMemcpyDeviceToHost));
    if (newton) HIPCHK(hipMemcpy(newton, e->log_newton, n_iter * M * sizeof(int), hipMemcpyDeviceToHost));
    if (lldec) HIPCHK(hipMemcpy(lldec, e->log_lldec, n_iter * sizeof(int), hipMemcpyDeviceToHost));
    if (vsum) HIPCHK(hipMemcpy(vsum, e->log_vsum, n_iter * M * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int amica_engine_get_log_at(amica_engine* e, int it, double* ll, double* lrate, int* bflag,
                            int* newton, int* lldec, double* vsum) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    if (it < 1 || it > e->o.max_iter) return fail(AMICA_E_INVALID, "iteration outside the log capacity");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    const size_t k = (size_t)it - 1, M = e->o.M;
    if (ll) HIPCHK(hipMemcpy(ll, e->log_ll + k, sizeof(double), hipMemcpyDeviceToHost));
    if (lrate) HIPCHK(hipMemcpy(lrate, e->log_lrate + k, sizeof(double), hipMemcpyDeviceToHost));
    if (bflag) HIPCHK(hipMemcpy(bflag, e->log_bflag + k * M, M * sizeof(int), hipMemcpyDeviceToHost));
    if (newton) HIPCHK(hipMemcpy(newton, e->log_newton + k * M, M * sizeof(int), hipMemcpyDeviceToHost));
    if (lldec) HIPCHK(hipMemcpy(lldec, e->log_lldec + k, sizeof(int), hipMemcpyDeviceToHost));
    if (vsum) HIPCHK(hipMemcpy(vsum, e->log_vsum + k * M, M * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int amica_engine_estep_stats(amica_engine* e, double* stats) {
    int rc = check_ready(e);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->o.device));
    rc = e->enqueue_estep(false);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->st));
    if (stats)
        HIPCHK(hipMemcpy(stats, e->stats, ((size_t)e->o.M * e->stride + 1) * sizeof(double),
                         hipMemcpyDeviceToHost));
    return 0;
}

int amica_engine_record_ltall(amica_engine* e, int on) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    HIPCHK(hipSetDevice(e->o.device));
    if (on && !e->ltall) {
        HIPCHK(hipMalloc((void**)&e->ltall, (size_t)e->o.M * e->o.T * sizeof(float)));
        if (e->gexec) {   /* the captured graph does not record ltall */
            (void)hipGraphExecDestroy(e->gexec);
            (void)hipGraphDestroy(e->graph);
            e->gexec = nullptr;
            e->graph = nullptr;
        }
    }
    if (e->ltall_on != (on != 0) && e->gexec) {
        (void)hipGraphExecDestroy(e->gexec);
        (void)hipGraphDestroy(e->graph);
        e->gexec = nullptr;
        e->graph = nullptr;
    }
    e->ltall_on = on != 0;
    return 0;
}

int amica_engine_get_ltall(amica_engine* e, double* lt) {
    if (!e || !lt) return fail(AMICA_E_INVALID, "null argument");
    if (!e->ltall) return fail(AMICA_E_STATE, "ltall recording was not enabled");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    const size_t cnt = (size_t)e->o.M * e->o.T;
    std::vector<float> tmp(cnt);
    HIPCHK(hipMemcpy(tmp.data(), e->ltall, cnt * sizeof(float), hipMemcpyDeviceToHost));
    double c0 = 0.0;
    HIPCHK(hipMemcpy(&c0, e->work, sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < cnt; ++k) lt[k] = (double)tmp[k] + c0;
    return 0;
}

int amica_engine_output(amica_engine* e, double* A_out, double* c_out) {
    if (!e || !A_out || !c_out) return fail(AMICA_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->o.device));
    const int n = e->o.n, M = e->o.M;
    const size_t nn = (size_t)n * n;
    double *lu = nullptr, *Sinv = nullptr, *Ao = nullptr, *co = nullptr;
    HIPCHK(hipMalloc((void**)&lu, 2 * nn * sizeof(double)));
    HIPCHK(hipMalloc((void**)&Sinv, nn * sizeof(double)));
    HIPCHK(hipMalloc((void**)&Ao, M * nn * sizeof(double)));
    HIPCHK(hipMalloc((void**)&co, (size_t)M * n * sizeof(double)));
    int rc = launch_output_transform(e->S, e->mean, e->A, e->c, n, M, lu, Sinv, Ao, co, e->st);
    hipError_t se = hipStreamSynchronize(e->st);
    if (!rc && se == hipSuccess) {
        se = hipMemcpy(A_out, Ao, M * nn * sizeof(double), hipMemcpyDeviceToHost);
        if (se == hipSuccess) se = hipMemcpy(c_out, co, (size_t)M * n * sizeof(double), hipMemcpyDeviceToHost);
    }
    (void)hipFree(lu); (void)hipFree(Sinv); (void)hipFree(Ao); (void)hipFree(co);
    if (rc || se != hipSuccess) return fail(AMICA_E_DEVICE, "output transform failed");
    return 0;
}

int amica_engine_debug_buffer(amica_engine* e, const char* name, void* out, size_t bytes) {
    if (!e || !name || !out) return fail(AMICA_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    const long T = e->o.T;
    const size_t NP = e->NP, ldb = e->ldb;
    struct Buf { const char* n; const void* p; size_t size; };
    const Buf bufs[] = {
        {"x", e->xw, T * NP * sizeof(float)},
        {"x64", e->xw64, e->xw64 ? T * NP * sizeof(double) : 0},
        {"xbf", e->xbf, e->xbf ? 2 * T * NP * sizeof(uint16_t) : 0},
        {"b", e->b, e->b ? T * ldb * sizeof(float) : 0},
        {"b64", e->b64, e->b64 ? T * ldb * sizeof(double) : 0},
        {"g", e->g, e->bf16 ? (T + 31) / 32 * 32 * ldb * sizeof(uint16_t) : T * ldb * sizeof(float)},
        {"gx", e->gxtot, ldb * NP * sizeof(double)},
        {"raw", e->raw, e->pstride * sizeof(double)},
        {"Wg64", e->Wg64, e->Wg64 ? ldb * NP * sizeof(double) : 0},
        {"Wg32", e->Wg32, e->Wg32 ? ldb * NP * sizeof(float) : 0},
        {"wc64", e->wc64, e->wc64 ? ldb * sizeof(double) : 0},
        {"wc", e->wc, ldb * sizeof(float)},
        {"ldet", e->ldet, (size_t)e->o.M * sizeof(double)},
    };
    for (const Buf& b : bufs) {
        if (std::strcmp(b.n, name) != 0) continue;
        if (!b.p || b.size == 0) return fail(AMICA_E_STATE, std::string("buffer not allocated: ") + name);
        if (bytes != b.size) return fail(AMICA_E_INVALID, std::string("size mismatch for ") + name +
                                                             ": " + std::to_string(b.size));
        HIPCHK(hipMemcpy(out, b.p, bytes, hipMemcpyDeviceToHost));
        return 0;
    }
    return fail(AMICA_E_INVALID, std::string("unknown buffer ") + name);
}

int amica_engine_set_timing(amica_engine* e, int on) {
    if (!e) return fail(AMICA_E_INVALID, "null engine");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    e->timing = on;
    e->ev_used = 0;
    return 0;
}

int amica_engine_get_timing(amica_engine* e, double* ms, int* count) {
    if (!e || !ms) return fail(AMICA_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->o.device));
    HIPCHK(hipStreamSynchronize(e->st));
    double sum[5] = {0, 0, 0, 0, 0};
    for (size_t s = 0; s < e->ev_used; ++s) {
        const auto& v = e->evs[s];
        for (int k = 0; k < 5; ++k) {
            float t = 0.f;
            HIPCHK(hipEventElapsedTime(&t, v[k], v[k + 1]));
            sum[k] += t;
        }
    }
    for (int k = 0; k < 5; ++k) ms[k] = e->ev_used ? sum[k] / (double)e->ev_used : 0.0;
    if (count) *count = (int)e->ev_used;
    return 0;
}

}  // extern "C"

/* ------------------------------------------------------------------------ */
/* batches of recordings (include/amica_batch.h)                              */
/* ------------------------------------------------------------------------ */

struct amica_batch {
    std::vector<amica_engine*> eng;
    int device = 0;
    int cus = 0;
    hipStream_t cap = nullptr;          /* the stream the batch graph is launched on */
    hipEvent_t fork = nullptr;
    hipEvent_t done = nullptr;          /* end of the last launch; engine streams wait on it */
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int mode = 0;                       /* launch mode of the current policy */
    long item_samples = 0;              /* mode 2: samples per work item (0: each recording's own partition) */
    hipEvent_t fork2 = nullptr;
    std::vector<hipEvent_t> join2;

    /* mode 2: the recordings whose heavy kernels share launches (same NP and
     * m), with their per-recording operands and work lists on the device */
    struct Group {
        int NP = 0, ldb = 0, m = 0;
        std::vector<int> recs;
        void* dmem = nullptr;
        UnmixArgs* ua = nullptr;
        EstepArgs* ea = nullptr;
        GxArgs* ga = nullptr;
        BatchItem *iu = nullptr, *ie = nullptr, *ig = nullptr;
        int nu = 0, ne = 0, ng = 0;
    };
    std::vector<Group> groups;
    std::vector<int> gr_nparts, gr_nch;   /* per recording: E-step and GEMM partial slots used */

    void drop_groups() {
        for (Group& g : groups)
            if (g.dmem) (void)hipFree(g.dmem);
        groups.clear();
    }

    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        drop_groups();
    }

    /* Work lists of the grouped launches: every recording's samples in
     * chunks of item_samples (at least T / its partial slots, whole 32-sample
     * blocks), the unmixing in the same ranges as 16-sample tiles, the
     * statistic GEMM as 32-sample blocks. */
    int build_groups() {
        drop_groups();
        const size_t R = eng.size();
        gr_nparts.assign(R, 0);
        gr_nch.assign(R, 0);
        for (size_t r = 0; r < R; ++r) {
            const amica_engine* e = eng[r];
            if (!e->groupable()) continue;
            Group* g = nullptr;
            for (Group& q : groups)
                if (q.NP == e->NP && q.m == e->o.m) g = &q;
            if (!g) {
                groups.emplace_back();
                g = &groups.back();
                g->NP = e->NP;
                g->ldb = e->ldb;
                g->m = e->o.m;
            }
            g->recs.push_back((int)r);
        }
        for (Group& g : groups) {
            std::vector<UnmixArgs> ua;
            std::vector<EstepArgs> ea;
            std::vector<GxArgs> ga;
            std::vector<BatchItem> iu, ie, ig;
            for (size_t k = 0; k < g.recs.size(); ++k) {
                const int r = g.recs[k];
                amica_engine* e = eng[r];
                const long T = e->o.T;
                /* item_samples 0: the recording's own grid partition, so that
                 * every partial sum covers the samples it covers in the
                 * recording's own launches (identical results) */
                long chunk = item_samples > 0
                                 ? std::max(item_samples, (T + e->nblocks_cap - 1) / e->nblocks_cap)
                                 : e->chunk;
                chunk = (chunk + 31) / 32 * 32;
                const int nparts = (int)((T + chunk - 1) / chunk);
                for (int i = 0; i < nparts; ++i) {
                    const long t0 = (long)i * chunk, t1 = std::min(T, t0 + chunk);
                    ie.push_back(BatchItem{(int)k, i, t0, t1});
                    iu.push_back(BatchItem{(int)k, 0, t0 / 16, (t1 + 15) / 16});
                }
                const long nblk = (T + 31) / 32;
                const long per = item_samples > 0
                                     ? std::max(chunk / 32, (nblk + e->nchunks_cap - 1) / e->nchunks_cap)
                                     : (nblk + e->nchunks - 1) / e->nchunks;
                const int nch = (int)((nblk + per - 1) / per);
                for (int i = 0; i < nch; ++i)
                    ig.push_back(BatchItem{(int)k, i, (long)i * per, std::min(nblk, (long)(i + 1) * per)});
                gr_nparts[r] = nparts;
                gr_nch[r] = nch;
                EstepArgs a = e->eargs(nullptr);
                a.chunk = chunk;
                ea.push_back(a);
                ua.push_back(UnmixArgs{e->xbf, e->xbf + (size_t)T * e->NP, e->Whi, e->Wlo, e->wc, e->b, T, e->ldb});
                ga.push_back(GxArgs{(const uint16_t*)e->g, e->x8, e->gxpart});
            }
            auto bytes = [](size_t n, size_t sz) { return (n * sz + 255) / 256 * 256; };
            const size_t b_ua = bytes(ua.size(), sizeof(UnmixArgs)), b_ea = bytes(ea.size(), sizeof(EstepArgs)),
                         b_ga = bytes(ga.size(), sizeof(GxArgs)), b_iu = bytes(iu.size(), sizeof(BatchItem)),
                         b_ie = bytes(ie.size(), sizeof(BatchItem)), b_ig = bytes(ig.size(), sizeof(BatchItem));
            HIPCHK(hipMalloc(&g.dmem, b_ua + b_ea + b_ga + b_iu + b_ie + b_ig));
            char* p = (char*)g.dmem;
            g.ua = (UnmixArgs*)p; p += b_ua;
            g.ea = (EstepArgs*)p; p += b_ea;
            g.ga = (GxArgs*)p; p += b_ga;
            g.iu = (BatchItem*)p; p += b_iu;
            g.ie = (BatchItem*)p; p += b_ie;
            g.ig = (BatchItem*)p;
            HIPCHK(hipMemcpy(g.ua, ua.data(), ua.size() * sizeof(UnmixArgs), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(g.ea, ea.data(), ea.size() * sizeof(EstepArgs), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(g.ga, ga.data(), ga.size() * sizeof(GxArgs), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(g.iu, iu.data(), iu.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(g.ie, ie.data(), ie.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(g.ig, ig.data(), ig.size() * sizeof(BatchItem), hipMemcpyHostToDevice));
            g.nu = (int)iu.size();
            g.ne = (int)ie.size();
            g.ng = (int)ig.size();
        }
        return 0;
    }

    /* mode 2: per recording W = A^-1 and the constants (branches); on the
     * capture's origin stream, per group one launch for each of the
     * unmixing GEMM, the E-step and the statistic GEMM over all its
     * recordings; then per recording the reductions and the M-step
     * (branches).  Recordings outside the groups run their whole iteration
     * on their own branch.  The groups' launches stay on the origin
     * stream: streams created for them and first used inside the capture
     * ended the process with a segmentation fault in the first grouped
     * iteration. */
    int enqueue_grouped() {
        int rc = hipEventRecord(fork, cap) == hipSuccess ? 0 : AMICA_E_DEVICE;
        for (size_t r = 0; r < eng.size() && rc == 0; ++r) {
            amica_engine* e = eng[r];
            if (hipStreamWaitEvent(e->st, fork, 0) != hipSuccess) { rc = AMICA_E_DEVICE; break; }
            rc = e->groupable() ? e->enqueue_prepare() : e->enqueue_iteration(false);
            if (rc) break;
            if (hipEventRecord(join[r], e->st) != hipSuccess) { rc = AMICA_E_DEVICE; break; }
            if (e->groupable() && hipStreamWaitEvent(cap, join[r], 0) != hipSuccess) rc = AMICA_E_DEVICE;
        }
        for (const Group& g : groups) {
            if (rc) break;
            if (launch_unmix_batch(g.ua, g.iu, g.nu, g.NP, g.ldb, cap) ||
                launch_estep_m1_batch(g.ea, g.ie, g.ne, g.m, g.NP, cap) ||
                launch_gx_bf16_batch(g.ga, g.ig, g.ng, g.NP, g.ldb, cap))
                rc = fail(AMICA_E_DEVICE, std::string("grouped launch failed: ") + hipGetErrorString(hipGetLastError()));
        }
        if (!rc && hipEventRecord(fork2, cap) != hipSuccess) rc = AMICA_E_DEVICE;
        for (size_t r = 0; r < eng.size() && rc == 0; ++r) {
            amica_engine* e = eng[r];
            if (!e->groupable()) {
                if (hipStreamWaitEvent(cap, join[r], 0) != hipSuccess) rc = AMICA_E_DEVICE;
                continue;
            }
            if (hipStreamWaitEvent(e->st, fork2, 0) != hipSuccess) { rc = AMICA_E_DEVICE; break; }
            rc = e->enqueue_reduce(gr_nparts[r], gr_nch[r]);
            if (!rc) rc = e->enqueue_mstep();
            if (rc) break;
            if (hipEventRecord(join2[r], e->st) != hipSuccess ||
                hipStreamWaitEvent(cap, join2[r], 0) != hipSuccess)
                rc = AMICA_E_DEVICE;
        }
        return rc;
    }

    /* one iteration of every recording: fork from `cap`, each recording on
     * its own stream (mode 0), or the grouped launches (mode 2), join back
     * into `cap` */
    int capture() {
        if (mode == 2) {
            const int brc = build_groups();
            if (brc) return brc;
        }
        HIPCHK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        int rc = 0;
        if (mode == 2) rc = enqueue_grouped();
        else rc = hipEventRecord(fork, cap) == hipSuccess ? 0 : AMICA_E_DEVICE;
        for (size_t r = 0; mode != 2 && r < eng.size() && rc == 0; ++r) {
            if (hipStreamWaitEvent(eng[r]->st, fork, 0) != hipSuccess) { rc = AMICA_E_DEVICE; break; }
            rc = eng[r]->enqueue_iteration(false);
            if (rc) break;
            if (hipEventRecord(join[r], eng[r]->st) != hipSuccess ||
                hipStreamWaitEvent(cap, join[r], 0) != hipSuccess)
                rc = AMICA_E_DEVICE;
        }
        hipGraph_t gr = nullptr;
        const hipError_t ce = hipStreamEndCapture(cap, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc == AMICA_E_DEVICE ? fail(AMICA_E_DEVICE, "batch capture failed") : rc;
        }
        if (ce != hipSuccess) return fail(AMICA_E_DEVICE, std::string("batch capture: ") + hipGetErrorString(ce));
        graph = gr;
        HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        return 0;
    }
};

extern "C" {

int amica_batch_destroy(amica_batch* b) {
    if (!b) return fail(AMICA_E_INVALID, "null batch");
    (void)hipSetDevice(b->device);
    if (b->cap) (void)hipStreamSynchronize(b->cap);
    b->drop_graph();
    for (hipEvent_t ev : b->join)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : b->join2)
        if (ev) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->fork2) (void)hipEventDestroy(b->fork2);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    for (amica_engine* e : b->eng)
        if (e) amica_engine_destroy(e);
    delete b;
    return 0;
}

int amica_batch_create(const amica_options* opts, int nrec, amica_batch** out) {
    if (!opts || !out || nrec < 1) return fail(AMICA_E_INVALID, "need options for nrec >= 1 recordings");
    *out = nullptr;
    for (int r = 1; r < nrec; ++r)
        if (opts[r].device != opts[0].device) return fail(AMICA_E_INVALID, "all recordings of a batch share one device");
    amica_batch* b = new amica_batch;
    b->device = opts[0].device;
    for (int r = 0; r < nrec; ++r) {
        amica_engine* e = nullptr;
        const int rc = amica_engine_create(&opts[r], &e);
        if (rc) {
            const std::string msg = "recording " + std::to_string(r) + ": " + g_last_error;
            amica_batch_destroy(b);
            return fail(rc, msg);
        }
        b->eng.push_back(e);
    }
    b->cus = b->eng[0]->cus;
    b->join.assign(nrec, nullptr);
    b->join2.assign(nrec, nullptr);
    if (hipSetDevice(b->device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->fork2, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
        amica_batch_destroy(b);
        return fail(AMICA_E_DEVICE, "batch stream or event creation failed");
    }
    for (int r = 0; r < nrec; ++r)
        if (hipEventCreateWithFlags(&b->join[r], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->join2[r], hipEventDisableTiming) != hipSuccess) {
            amica_batch_destroy(b);
            return fail(AMICA_E_DEVICE, "batch event creation failed");
        }
    *out = b;
    return 0;
}

int amica_batch_size(amica_batch* b) { return b ? (int)b->eng.size() : 0; }

int amica_batch_engine(amica_batch* b, int r, amica_engine** eng) {
    if (!b || !eng || r < 0 || r >= (int)b->eng.size()) return fail(AMICA_E_INVALID, "no such recording");
    *eng = b->eng[r];
    return 0;
}

int amica_batch_set_budgets(amica_batch* b, const int* cu_budget) {
    if (!b) return fail(AMICA_E_INVALID, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->cap));
    for (size_t r = 0; r < b->eng.size(); ++r) {
        HIPCHK(hipStreamSynchronize(b->eng[r]->st));
        b->eng[r]->size_grids(cu_budget ? cu_budget[r] : b->cus);
    }
    b->drop_graph();
    return 0;
}

int amica_batch_set_policy(amica_batch* b, int p) {
    if (!b) return fail(AMICA_E_INVALID, "null batch");
    if (p < 0 || p >= AMICA_BATCH_POLICIES) return fail(AMICA_E_INVALID, "no such policy");
    if (p >= 8) {
        /* grouped launches over all recordings; work items of each recording's
         * own partition, or of 512, 1024 or 2048 samples */
        static const long kItem[4] = {0, 512, 1024, 2048};
        b->drop_graph();
        b->mode = 2;
        b->item_samples = kItem[p - 8];
        return amica_batch_set_budgets(b, nullptr);
    }
    b->mode = p / 4;
    p %= 4;
    const int R = (int)b->eng.size(), cus = b->cus;
    double total = 0.0;
    std::vector<double> work(R);
    for (int r = 0; r < R; ++r) {
        const amica_options& o = b->eng[r]->o;
        work[r] = (double)o.M * o.m * o.n * (double)o.T;
        total += work[r];
    }
    std::vector<int> bud(R);
    for (int r = 0; r < R; ++r) {
        const int share = std::max(std::max(1, cus / 16), (int)std::lround(cus * work[r] / total));
        switch (p) {
            case 0: bud[r] = cus; break;
            case 1: bud[r] = std::min(cus, share); break;
            case 2: bud[r] = std::min(cus, 2 * share); break;
            default: bud[r] = std::max(1, cus / 2); break;
        }
    }
    return amica_batch_set_budgets(b, bud.data());
}

int amica_batch_iterate(amica_batch* b, int k) {
    if (!b) return fail(AMICA_E_INVALID, "null batch");
    for (amica_engine* e : b->eng) {
        const int rc = check_ready(e);
        if (rc) return rc;
    }
    HIPCHK(hipSetDevice(b->device));
    if (b->mode == 1) {
        /* each recording replays its own graph on its own stream */
        HIPCHK(hipEventRecord(b->fork, b->cap));
        for (size_t r = 0; r < b->eng.size(); ++r) {
            amica_engine* e = b->eng[r];
            HIPCHK(hipStreamWaitEvent(e->st, b->fork, 0));
            const int rc = amica_engine_iterate_graph(e, k);
            if (rc) return rc;
            HIPCHK(hipEventRecord(b->join[r], e->st));
            HIPCHK(hipStreamWaitEvent(b->cap, b->join[r], 0));
        }
        return 0;
    }
    if (!b->gexec) {
        const int rc = b->capture();
        if (rc) return rc;
    }
    for (int it = 0; it < k; ++it) HIPCHK(hipGraphLaunch(b->gexec, b->cap));
    /* work later enqueued on a recording's own stream (reads of its logs and
     * state) orders after the batch */
    HIPCHK(hipEventRecord(b->done, b->cap));
    for (amica_engine* e : b->eng) HIPCHK(hipStreamWaitEvent(e->st, b->done, 0));
    return 0;
}

int amica_batch_sync(amica_batch* b) {
    if (!b) return fail(AMICA_E_INVALID, "null batch");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->cap));
    return 0;
}

int amica_batch_tune(amica_batch* b, int iters, int* policy, double* ms) {
    if (!b || iters < 1) return fail(AMICA_E_INVALID, "null batch or iters < 1");
    for (amica_engine* e : b->eng) {
        const int rc = check_ready(e);
        if (rc) return rc;
        if (e->o.max_iter < iters + 1) return fail(AMICA_E_INVALID, "max_iter of every recording must cover the tuning iterations");
    }
    HIPCHK(hipSetDevice(b->device));
    /* the state every recording had on entry */
    struct Snap { std::vector<double> A, c, alpha, mu, beta, rho, gm, W; };
    std::vector<Snap> snap(b->eng.size());
    for (size_t r = 0; r < b->eng.size(); ++r) {
        const amica_options& o = b->eng[r]->o;
        const size_t nn = (size_t)o.n * o.n * o.M, nm = (size_t)o.n * o.m * o.M;
        Snap& s = snap[r];
        s.A.resize(nn); s.W.resize(nn); s.c.resize((size_t)o.n * o.M);
        s.alpha.resize(nm); s.mu.resize(nm); s.beta.resize(nm); s.rho.resize(nm); s.gm.resize(o.M);
        const int rc = amica_engine_get_state(b->eng[r], s.A.data(), s.c.data(), s.alpha.data(),
                                              s.mu.data(), s.beta.data(), s.rho.data(), s.gm.data(),
                                              s.W.data());
        if (rc) return rc;
    }
    auto restore = [&]() -> int {
        for (size_t r = 0; r < b->eng.size(); ++r) {
            const Snap& s = snap[r];
            const int rc = amica_engine_set_state(b->eng[r], s.A.data(), s.c.data(), s.alpha.data(),
                                                  s.mu.data(), s.beta.data(), s.rho.data(), s.gm.data());
            if (rc) return rc;
        }
        return 0;
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = HUGE_VAL;
    int bestp = 0, rc = 0;
    for (int p = 0; p < AMICA_BATCH_POLICIES && rc == 0; ++p) {
        rc = amica_batch_set_policy(b, p);
        if (!rc) rc = amica_batch_iterate(b, 1);            /* capture + untimed replay */
        if (!rc && hipEventRecord(e0, b->cap) != hipSuccess) rc = AMICA_E_DEVICE;
        if (!rc) rc = amica_batch_iterate(b, iters);
        if (!rc && hipEventRecord(e1, b->cap) != hipSuccess) rc = AMICA_E_DEVICE;
        if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = AMICA_E_DEVICE;
        float el = 0.f;
        if (!rc && hipEventElapsedTime(&el, e0, e1) != hipSuccess) rc = AMICA_E_DEVICE;
        if (!rc) {
            const double t = el / iters;
            if (ms) ms[p] = t;
            if (t < best) { best = t; bestp = p; }
            rc = restore();
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc == AMICA_E_DEVICE ? fail(AMICA_E_DEVICE, "batch tuning failed") : rc;
    rc = amica_batch_set_policy(b, bestp);
    if (rc) return rc;
    if (policy) *policy = bestp;
    return restore();
}

}  // extern "C"
