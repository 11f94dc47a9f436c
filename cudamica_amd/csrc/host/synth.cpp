/*
 * synth.cpp -- deterministic synthetic recordings and initial model states
 * (declared in include/amica_synth.h).
 *
 * PRNG: xoshiro256** seeded through splitmix64 (Blackman & Vigna).
 * Normals: Marsaglia polar method.  Gamma(a): Marsaglia & Tsang (2000),
 * with Gamma(a) = Gamma(a+1) U^(1/a) for a < 1.  Generalized-Gaussian
 * source of shape rho: s = +-G^(1/rho) / sqrt(Gamma(3/rho)/Gamma(1/rho)),
 * G ~ Gamma(1/rho, 1), which has unit variance.  No std:: distributions
 * are used, so the output is the same across C++ libraries; samples are
 * drawn in blocks of 4096 with one stream per block, so it is also the
 * same for any thread count.
 */
#include "../../../include/amica_synth.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

inline uint64_t splitmix64(uint64_t* x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Rng {
    uint64_t s[4];
    bool has_spare = false;
    double spare = 0.0;
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& v : s) v = splitmix64(&x);
    }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    /* uniform on the open interval (0, 1) */
    double u01() { return ((double)(next() >> 11) + 0.5) * 0x1.0p-53; }
    double normal() {
        if (has_spare) { has_spare = false; return spare; }
        double u, v, q;
        do {
            u = 2.0 * u01() - 1.0;
            v = 2.0 * u01() - 1.0;
            q = u * u + v * v;
        } while (q >= 1.0 || q == 0.0);
        const double f = std::sqrt(-2.0 * std::log(q) / q);
        spare = v * f;
        has_spare = true;
        return u * f;
    }
    double gamma(double a) {
        if (a < 1.0) return gamma(a + 1.0) * std::pow(u01(), 1.0 / a);
        const double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
        for (;;) {
            double x, v;
            do {
                x = normal();
                v = 1.0 + c * x;
            } while (v <= 0.0);
            v = v * v * v;
            const double u = u01();
            if (u < 1.0 - 0.0331 * x * x * x * x) return d * v;
            if (std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v;
        }
    }
};

/* Eigenvalues of a symmetric matrix by cyclic Jacobi (column-major). */
void sym_eigenvalues(int n, std::vector<double> a, std::vector<double>* ev) {
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int q = 0; q < n; ++q)
            for (int p = 0; p < n; ++p) {
                const double x = a[p + (size_t)q * n];
                if (p == q) dia += x * x; else off += x * x;
            }
        if (off <= 1e-28 * dia) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[p + (size_t)q * n];
                if (apq == 0.0) continue;
                const double th = (a[q + (size_t)q * n] - a[p + (size_t)p * n]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double x = a[k + (size_t)p * n], y = a[k + (size_t)q * n];
                    a[k + (size_t)p * n] = c * x - s * y;
                    a[k + (size_t)q * n] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    const double x = a[p + (size_t)k * n], y = a[q + (size_t)k * n];
                    a[p + (size_t)k * n] = c * x - s * y;
                    a[q + (size_t)k * n] = s * x + c * y;
                }
            }
    }
    ev->resize(n);
    for (int i = 0; i < n; ++i) (*ev)[i] = a[i + (size_t)i * n];
}

double cond_of(int n, const double* A) {
    std::vector<double> ata((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i <= j; ++i) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += A[k + (size_t)i * n] * A[k + (size_t)j * n];
            ata[i + (size_t)j * n] = s;
            ata[j + (size_t)i * n] = s;
        }
    std::vector<double> ev;
    sym_eigenvalues(n, ata, &ev);
    double lo = ev[0], hi = ev[0];
    for (double v : ev) { lo = std::min(lo, v); hi = std::max(hi, v); }
    if (lo <= 0.0) return HUGE_VAL;
    return std::sqrt(hi / lo);
}

uint64_t derive(uint64_t seed, uint64_t a, uint64_t b) {
    uint64_t x = seed ^ (a * 0xD1B54A32D192ED03ull) ^ (b * 0x9E3779B97F4A7C15ull);
    return splitmix64(&x);
}

}  // namespace

extern "C" {

double amica_synth_cond(int n, const double* A) { return cond_of(n, A); }

int amica_synth_data(const amica_synth_spec* sp, float* x, double* A_true, double* rho_true,
                     int nthreads) {
    if (!sp || !x || sp->n < 1 || sp->T < 1 || sp->nmix < 1) return -3;
    const int n = sp->n, nmix = sp->nmix;
    const long T = sp->T;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    Rng main(derive(sp->seed, 1, 0));
    /* source shapes */
    std::vector<double> rho(n);
    const int nsub = (int)std::lround(sp->sub_frac * n);
    for (int i = 0; i < n; ++i)
        rho[i] = (i < nsub) ? sp->sub_lo + (sp->sub_hi - sp->sub_lo) * main.u01()
                            : sp->rho_lo + (sp->rho_hi - sp->rho_lo) * main.u01();
    if (rho_true) std::copy(rho.begin(), rho.end(), rho_true);
    /* mixing matrices: attempt r of matrix q has its own stream; the first
     * attempt that meets the condition bound is kept */
    std::vector<double> A((size_t)nmix * n * n);
    for (int q = 0; q < nmix; ++q) {
        const int batch = std::max(1, nthreads);
        bool found = false;
        for (long r0 = 0; !found; r0 += batch) {
            std::vector<std::vector<double>> cand(batch, std::vector<double>((size_t)n * n));
            std::vector<double> cnd(batch, HUGE_VAL);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
            for (int b = 0; b < batch; ++b) {
                Rng r(derive(sp->seed, 2 + q, (uint64_t)(r0 + b)));
                for (size_t k = 0; k < (size_t)n * n; ++k) cand[b][k] = r.normal();
                cnd[b] = (sp->cond_max > 0) ? cond_of(n, cand[b].data()) : 0.0;
            }
            for (int b = 0; b < batch; ++b) {
                if (sp->cond_max <= 0 || cnd[b] <= sp->cond_max) {
                    std::copy(cand[b].begin(), cand[b].end(), A.begin() + (size_t)q * n * n);
                    found = true;
                    break;
                }
            }
            if (r0 > 2000000) return -4;
        }
    }
    if (A_true) std::copy(A.begin(), A.end(), A_true);
    /* block -> matrix assignment */
    const long blk = (nmix > 1 && sp->block > 0) ? sp->block : T;
    const long nb = (T + blk - 1) / blk;
    std::vector<int> which(nb, 0);
    for (long b = 0; b < nb; ++b)
        which[b] = (nmix == 1) ? 0 : (sp->alternate ? (int)(b % nmix) : std::min(nmix - 1, (int)(main.u01() * nmix)));
    std::vector<double> sd(n);
    for (int i = 0; i < n; ++i)
        sd[i] = std::sqrt(std::exp(std::lgamma(3.0 / rho[i]) - std::lgamma(1.0 / rho[i])));
    const long SB = 4096;
    const long nsb = (T + SB - 1) / SB;
#pragma omp parallel num_threads(nthreads)
    {
        std::vector<double> s(n), xv(n);
#pragma omp for schedule(dynamic, 4)
        for (long sb = 0; sb < nsb; ++sb) {
            Rng r(derive(sp->seed, 1000003, (uint64_t)sb));
            const long te = std::min(T, (sb + 1) * SB);
            for (long t = sb * SB; t < te; ++t) {
                for (int i = 0; i < n; ++i) {
                    const double g = r.gamma(1.0 / rho[i]);
                    const double sign = (r.next() >> 63) ? -1.0 : 1.0;
                    s[i] = sign * std::pow(g, 1.0 / rho[i]) / sd[i];
                }
                const double* Aq = &A[(size_t)which[t / blk] * n * n];
                for (int i = 0; i < n; ++i) xv[i] = 0.0;
                for (int k = 0; k < n; ++k) {
                    const double sk = s[k];
                    const double* col = Aq + (size_t)k * n;
                    for (int i = 0; i < n; ++i) xv[i] += col[i] * sk;
                }
                if (sp->noise_std > 0)
                    for (int i = 0; i < n; ++i) xv[i] += sp->noise_std * r.normal();
                float* xt = x + (size_t)t * n;
                for (int i = 0; i < n; ++i) xt[i] = (float)xv[i];
            }
        }
    }
    return 0;
}

int amica_synth_init(int n, int M, int m, unsigned long long seed, int fix_init, double rho0,
                     double* A, double* c, double* alpha, double* mu, double* beta, double* rho,
                     double* gm) {
    if (n < 1 || M < 1 || m < 1) return -3;
    Rng r(derive(seed, 7, 0));
    const size_t nn = (size_t)n * n, nm = (size_t)n * m;
    for (int h = 0; h < M; ++h) {
        double* Ah = A + h * nn;
        for (size_t k = 0; k < nn; ++k) Ah[k] = 0.05 * (r.u01() - 0.5);
        for (int i = 0; i < n; ++i) Ah[i + (size_t)i * n] += 1.0;
        for (int k = 0; k < n; ++k) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += Ah[i + (size_t)k * n] * Ah[i + (size_t)k * n];
            const double nu = std::sqrt(s);
            for (int i = 0; i < n; ++i) Ah[i + (size_t)k * n] /= nu;
        }
        for (int i = 0; i < n; ++i) c[(size_t)h * n + i] = r.normal();
        gm[h] = 1.0 / (double)M;
        for (size_t q = 0; q < nm; ++q) alpha[h * nm + q] = 1.0 / (double)m;
        if (fix_init) {
            for (int j = 0; j < m; ++j)
                for (int i = 0; i < n; ++i) {
                    mu[h * nm + (size_t)j * n + i] = (double)j - (m - 1.0) / 2.0;
                    beta[h * nm + (size_t)j * n + i] = 1.0;
                }
        } else {
            for (size_t q = 0; q < nm; ++q) mu[h * nm + q] = (m > 1) ? 0.1 * r.normal() : 0.0;
            for (size_t q = 0; q < nm; ++q) beta[h * nm + q] = 1.0 + 0.1 * r.normal();
        }
        for (size_t q = 0; q < nm; ++q) rho[h * nm + q] = rho0;
    }
    return 0;
}

}  // extern "C"
