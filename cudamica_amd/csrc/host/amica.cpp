/*
 * amica.cpp -- runamica(), the AMICA EM driver (fraimondo/cudamica
 * src/amica.cu:624-2843), on the MI355X engine.
 *
 * Sequence kept from the reference: load the .fdt data, center, whiten,
 * initialise the models (amica.cu:135-398), run maxiter EM iterations
 * printing the same per-iteration lines, then write W, A, c, ltall, gm,
 * alpha, beta, mu, rho, LL and the sphere matrix as raw float64 files.
 * Each EM iteration runs entirely on the GPU (include/amica_engine.h); the
 * host only reads that iteration's log to print it.
 *
 * Output semantics kept from the reference: W is pinv(A) from the start of
 * the last iteration (amica.cu:2552-2562); A and c go through
 * c = S^+ c + mean, A = colnorm(S^+ A) (amica.cu:2578-2668); alpha, mu,
 * beta, rho are the model's values after the last M-step.  The LL file
 * holds LL of iterations 1..maxiter (the reference writes entries
 * 0..maxiter-1 of its history, whose entry 0 is never set, amica.cu:2806).
 * The initial state comes from `ainitfile` / `cinitfile` when given, else
 * from the engine's deterministic PRNG (`seed`) with the reference's
 * recipe, since its cuRAND draws cannot be reproduced.
 */
#include "../../../include/amica.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <vector>

#include "../../../include/amica_synth.h"
#include "../../../include/config.h"
#include "../../../include/device.h"
#include "../../../include/helpers.h"
#include "../../../include/preprocess.h"

extern "C" {
model_block_t* blocks = NULL;
natural nblocks = 0;
}

namespace {

struct ModelState {
    std::vector<double> A, c, alpha, mu, beta, rho, gm, W;
};

error initialize_models(ModelState* s) {
    const natural n = currentConfig.nchannels, M = currentConfig.nmmodels, m = currentConfig.nsdm;
    s->A.resize(n * n * M);
    s->c.resize(n * M);
    s->alpha.resize(n * m * M);
    s->mu.resize(n * m * M);
    s->beta.resize(n * m * M);
    s->rho.resize(n * m * M);
    s->gm.resize(M);
    s->W.resize(n * n * M);
    if (amica_synth_init((int)n, (int)M, (int)m, currentConfig.seed, (int)(currentConfig.fix_init != 0),
                         currentConfig.rho0, s->A.data(), s->c.data(), s->alpha.data(), s->mu.data(),
                         s->beta.data(), s->rho.data(), s->gm.data()) != 0)
        return ERRORINVALIDPARAM;
    if (currentConfig.ainitfile) {
        error e = loadMatrix(s->A.data(), n * M, n, currentConfig.ainitfile, true);
        if (e != SUCCESS) {
            fprintf(stderr, "ERROR: cannot read ainitfile %s\n", currentConfig.ainitfile);
            return e;
        }
    }
    if (currentConfig.cinitfile) {
        error e = loadMatrix(s->c.data(), M, n, currentConfig.cinitfile, true);
        if (e != SUCCESS) {
            fprintf(stderr, "ERROR: cannot read cinitfile %s\n", currentConfig.cinitfile);
            return e;
        }
    }
    /* sample-block table of the reference (amica.cu:406-421) */
    const natural bs = currentConfig.block_size ? currentConfig.block_size : currentConfig.nsamples;
    nblocks = (currentConfig.nsamples + bs - 1) / bs;
    blocks = (model_block_t*)calloc(nblocks, sizeof(model_block_t));
    for (natural b = 0; b < nblocks; ++b) {
        blocks[b].device = 0;
        blocks[b].models = M;
        blocks[b].start = b * bs;
        blocks[b].end = std::min((b + 1) * bs, currentConfig.nsamples);
        blocks[b].block_size = blocks[b].end - blocks[b].start;
    }
    return SUCCESS;
}

error write_result(const char* path, const double* data, size_t xsize, size_t ysize) {
    if (!path) return SUCCESS;
    amica_pitched p = {(void*)data, xsize * sizeof(double), xsize, ysize};
    error e = writeMatrix(p, (char*)path);
    if (e != SUCCESS) fprintf(stderr, "ERROR: cannot write %s\n", path);
    return e;
}

}  // namespace

extern "C" error runamica(void) {
    fprintf(stdout, "====================================\n");
    fprintf(stdout, "Running AMICA\n");
    fprintf(stdout, "====================================\n");
    const natural n = currentConfig.nchannels, T = currentConfig.nsamples;
    const natural M = currentConfig.nmmodels, m = currentConfig.nsdm;
    const natural maxiter = currentConfig.maxiter;

    C_CHECK_RETURN(initializeDevs());
    amica_engine* eng = gpus[0].engine;
    {
        std::vector<float> host(n * T);
        C_CHECK_RETURN(loadMatrixF32(host.data(), T, n, currentConfig.datafile));
        if (amica_engine_load_raw(eng, host.data(), 0) != 0) {
            fprintf(stderr, "ERROR: %s\n", amica_last_error());
            exit(1);
        }
    }
    C_CHECK_RETURN(center());
    C_CHECK_RETURN(whiten());

    ModelState st;
    C_CHECK_RETURN(initialize_models(&st));
    if (amica_engine_set_state(eng, st.A.data(), st.c.data(), st.alpha.data(), st.mu.data(),
                               st.beta.data(), st.rho.data(), st.gm.data()) != 0) {
        fprintf(stderr, "ERROR: %s\n", amica_last_error());
        exit(1);
    }

    std::vector<double> ll(maxiter, 0.0);
    /* the E-step's algorithmic bytes per iteration, M n T (b read + g
     * written): b is FP64 on the FP32 path, g bf16 on the BF16 and FP8 paths */
    const natural prec = currentConfig.precision;
    const double estep_bytes = (double)M * n * T * ((prec == 32 ? 8 : 4) + (prec == 16 || prec == 8 ? 2 : 4));
    std::vector<double> vs(M);
    std::vector<int> bfs(M), nws(M);
    FILE* jsonlog = nullptr;
    if (currentConfig.jsonlogfile && !(jsonlog = fopen(currentConfig.jsonlogfile, "w"))) {
        fprintf(stderr, "ERROR: cannot open %s\n", currentConfig.jsonlogfile);
        return ERRORNOFILE;
    }
    fprintf(stdout, "Starting %lu AMICA iterations\n", maxiter);
    for (natural iter = 1; iter <= maxiter; ++iter) {
        const clock_t c0 = clock();
        const auto t0 = std::chrono::steady_clock::now();
        /* one captured hipGraph per iteration; the last iteration records the
         * per-sample L_h when ltall is written (not part of the graph), and
         * with a JSON log every iteration runs eagerly between HIP events,
         * which time its stages */
        const bool last_ltall = iter == maxiter && currentConfig.ltallfile;
        if (last_ltall) amica_engine_record_ltall(eng, 1);
        if (jsonlog) amica_engine_set_timing(eng, 1);     /* this iteration's events only */
        const int irc = (last_ltall || jsonlog) ? amica_engine_iterate(eng, 1)
                                                : amica_engine_iterate_graph(eng, 1);
        if (irc != 0 || amica_engine_sync(eng) != 0) {
            fprintf(stderr, "ERROR: iteration %lu: %s\n", iter, amica_last_error());
            exit(1);
        }
        /* this iteration's log entry only */
        double lli = 0.0, lri = 0.0;
        int deci = 0;
        if (amica_engine_get_log_at(eng, (int)iter, &lli, &lri, bfs.data(), nws.data(), &deci, vs.data()) != 0) {
            fprintf(stderr, "ERROR: %s\n", amica_last_error());
            exit(1);
        }
        ll[iter - 1] = lli;
        for (natural h = 0; h < M; ++h)
            fprintf(stdout, "mmodel %lu -> vsum %.16f\n", h, vs[h]);
        fprintf(stdout, "Iteration %lu: lrate %.16f LL %.16f \n", iter, lri, lli);
        if (deci) fprintf(stdout, "Likelihood decreasing!\n");
        for (natural h = 0; h < M; ++h)
            if (bfs[h] != 0 && iter >= currentConfig.newt_start_iter)
                fprintf(stdout, "Hessian non positive definite! Using natural gradient!\n");
        const auto t1 = std::chrono::steady_clock::now();
        const double msd = std::chrono::duration<double, std::milli>(t1 - t0).count();
        fprintf(stdout, "Iteration %lu times = %lu ticks, %lu ms \n", iter,
                (unsigned long)(clock() - c0), (unsigned long)msd);
        fflush(stdout);
        if (jsonlog) {
            double sms[5] = {0, 0, 0, 0, 0};
            int cnt = 0;
            if (amica_engine_get_timing(eng, sms, &cnt) != 0) {
                fprintf(stderr, "ERROR: %s\n", amica_last_error());
                exit(1);
            }
            fprintf(jsonlog, "{\"iteration\": %lu, \"ll\": %.17g, \"lrate\": %.17g, \"ll_decrease\": %d, "
                             "\"ms\": %.4f, \"stages_ms\": {\"inverse\": %.4f, \"unmix\": %.4f, "
                             "\"estep\": %.4f, \"stat_gemm\": %.4f, \"reduce_mstep\": %.4f}, "
                             "\"estep_bytes\": %.0f, \"models\": [",
                    iter, lli, lri, deci, msd, sms[0], sms[1], sms[2], sms[3], sms[4], estep_bytes);
            for (natural h = 0; h < M; ++h)
                fprintf(jsonlog, "%s{\"vsum\": %.17g, \"bflag\": %d, \"newton\": %d}", h ? ", " : "", vs[h],
                        bfs[h], nws[h]);
            fprintf(jsonlog, "]}\n");
            fflush(jsonlog);
        }
    }
    if (jsonlog) fclose(jsonlog);

    /* results (amica.cu:2542-2816) */
    if (amica_engine_get_state(eng, st.A.data(), st.c.data(), st.alpha.data(), st.mu.data(),
                               st.beta.data(), st.rho.data(), st.gm.data(), st.W.data()) != 0) {
        fprintf(stderr, "ERROR: %s\n", amica_last_error());
        exit(1);
    }
    std::vector<double> Aout(n * n * M), cout(n * M);
    if (amica_engine_output(eng, Aout.data(), cout.data()) != 0) {
        fprintf(stderr, "ERROR: %s\n", amica_last_error());
        exit(1);
    }
    std::vector<double> ltall;
    if (currentConfig.ltallfile) {
        ltall.resize(M * T);
        if (amica_engine_get_ltall(eng, ltall.data()) != 0) {
            fprintf(stderr, "ERROR: %s\n", amica_last_error());
            exit(1);
        }
    }
    std::vector<double> S(n * n), mean(n);
    double ldetS = 0.0;
    amica_engine_get_sphere(eng, S.data(), mean.data(), &ldetS);
    fprintf(stdout, "Finished!\n");
    printf("Writing results\n");
    C_CHECK_RETURN(write_result(currentConfig.weightsfile, st.W.data(), n, n * M));
    C_CHECK_RETURN(write_result(currentConfig.afile, Aout.data(), n, n * M));
    C_CHECK_RETURN(write_result(currentConfig.cfile, cout.data(), n, M));
    C_CHECK_RETURN(write_result(currentConfig.ltallfile, ltall.data(), T, M));
    C_CHECK_RETURN(write_result(currentConfig.gmfile, st.gm.data(), M, 1));
    C_CHECK_RETURN(write_result(currentConfig.alphafile, st.alpha.data(), n, m * M));
    C_CHECK_RETURN(write_result(currentConfig.betafile, st.beta.data(), n, m * M));
    C_CHECK_RETURN(write_result(currentConfig.mufile, st.mu.data(), n, m * M));
    C_CHECK_RETURN(write_result(currentConfig.rhofile, st.rho.data(), n, m * M));
    C_CHECK_RETURN(write_result(currentConfig.llfile, ll.data(), maxiter, 1));
    C_CHECK_RETURN(write_result(currentConfig.spherefile, S.data(), n, n));

    C_CHECK_RETURN(finalizeDevs());
    free(blocks);
    blocks = NULL;
    nblocks = 0;
    printf("Done!\n");
    return SUCCESS;
}
