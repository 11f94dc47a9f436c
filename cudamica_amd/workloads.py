"""Synthetic AMICA workloads: data recipe, EM options and seed of each.

Each workload is named by its shape.  Data: generalized-Gaussian sources
mixed by N(0,1) matrices with a bounded condition number (see
include/amica_synth.h).  The initial state uses the seed + 0x1000 stream.
"""
from dataclasses import dataclass, field

import numpy as np

from . import engine as _e

_U64 = (1 << 64) - 1


@dataclass
class Workload:
    name: str
    n: int
    T: int
    M: int
    m: int
    iters: int
    seed: int
    rho: tuple = (1.0, 2.0)
    sub_frac: float = 0.0
    cond_max: float = 100.0
    noise_std: float = 0.0
    nmix: int = 1
    block: int = 0
    alternate: int = 0
    newt_start_iter: int = 50
    do_sphere: int = 1
    precision: int = _e.PREC_FP32
    extra: dict = field(default_factory=dict)

    def spec(self, T=None):
        return _e.synth_spec(self.n, self.T if T is None else T, self.seed, nmix=self.nmix,
                             block=self.block, alternate=self.alternate, rho=self.rho,
                             sub_frac=self.sub_frac, cond_max=self.cond_max, noise_std=self.noise_std)

    def init_state(self):
        return _e.synth_init(self.n, self.M, self.m, (self.seed + 0x1000) & _U64, fix_init=1,
                             rho0=1.5)

    def options(self, T=None, precision=None, max_iter=None, **kw):
        return _e.options(self.n, self.T if T is None else T, self.M, self.m,
                          precision=self.precision if precision is None else precision,
                          max_iter=self.iters if max_iter is None else max_iter,
                          newt_start_iter=self.newt_start_iter, **kw)


WORKLOADS = {
    # plain AMICA, FP64 host reference case: 16 ch x 20k samples, 50 iterations
    "n16_T20k": Workload("n16_T20k", n=16, T=20_000, M=1, m=3, iters=50, seed=0x5EED0000,
                         rho=(1.0, 1.5), cond_max=50.0, noise_std=0.01),
    # plain AMICA FP32, EEG-like 64 ch x 1M samples, Newton from iteration 25
    "n64_T1M": Workload("n64_T1M", n=64, T=1_000_000, M=1, m=3, iters=200, seed=0x5EED0001,
                        sub_frac=0.1, newt_start_iter=25),
    # mixed precision, non-stationary: 128 ch x 2M samples, 4 models
    "n128_T2M_M4": Workload("n128_T2M_M4", n=128, T=2_000_000, M=4, m=3, iters=200,
                            seed=0x5EED0002, sub_frac=0.1, nmix=4, block=50_000, alternate=0,
                            precision=_e.PREC_BF16),
    # 256 ch x 4M samples, 2 models alternating in 100k blocks, 4 components,
    # whitened data stored as FP8 e4m3 with per-channel scales
    "n256_T4M_M2": Workload("n256_T4M_M2", n=256, T=4_000_000, M=2, m=4, iters=200,
                            seed=0x5EED0004, nmix=2, block=100_000, alternate=1, cond_max=0.0,
                            precision=_e.PREC_FP8),
}


def batch_workload(count=32, seed=0x5EED0003, precision=_e.PREC_FP32, t_range=(100_000, 1_500_000),
                   n_choices=(32, 64, 128), iters=200):
    """Many independent recordings per EM iteration: n drawn uniformly from
    n_choices, T uniformly from t_range, M = 1, m = 3, rho ~ U[1, 2], mixing
    condition number <= 100.  Recording r uses the data seed
    seed ^ ((r + 1) * 0x9E3779B97F4A7C15) (64-bit)."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(count):
        n = int(rng.choice(n_choices))
        T = int(rng.integers(t_range[0], t_range[1] + 1))
        s = (seed ^ ((r + 1) * 0x9E3779B97F4A7C15)) & _U64
        out.append(Workload(f"rec{r}_n{n}_T{T}", n=n, T=T, M=1, m=3, iters=iters, seed=s,
                            rho=(1.0, 2.0), cond_max=100.0, precision=precision))
    return out
