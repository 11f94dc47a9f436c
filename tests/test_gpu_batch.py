"""Batches of recordings (include/amica_batch.h) on the GPU: one hipGraph per
EM iteration over recordings of uneven n, T and M.

* With every recording's grids sized for the whole GPU, the batch computes
  exactly what each recording's own engine computes (same partitions, fixed
  reduction order): identical LL, lrate and parameters, bit for bit, both
  as one graph and as one graph per recording on its own stream.
* The measured tuner times every CU-budget policy, applies the fastest and
  leaves each recording in the state it had; iterating afterwards gives the
  first LL of the solo engine to 1e-6 (other grids sum in another order).
"""
import numpy as np
import pytest

import cudamica_amd as ca

pytestmark = pytest.mark.gpu

RECS = [(32, 40000, 1), (64, 25000, 1), (128, 31000, 1), (48, 17003, 1), (64, 30000, 2)]


def _say(msg):
    print(msg, flush=True)


def _setup(precision, iters):
    opts, data, states = [], [], []
    for r, (n, T, M) in enumerate(RECS):
        x = ca.synth_data(ca.synth_spec(n, T, seed=0x5EED0300 + r, nmix=M, block=5000,
                                        rho=(1.0, 2.0), cond_max=100.0))
        opts.append(ca.options(n, T, M, 3, precision=precision, max_iter=iters))
        data.append(x)
        states.append(ca.synth_init(n, M, 3, 0x5EED1300 + r))
    return opts, data, states


def _load(eng, x, st):
    eng.load_raw(x)
    eng.preprocess(1)
    eng.set_state(st)


@pytest.mark.parametrize("precision,policy", [(ca.PREC_FP32, 0), (ca.PREC_BF16, 0), (ca.PREC_FP32, 4)])
def test_batch_equals_solo_engines(built, precision, policy):
    iters = 6
    _say(f"[batch] {len(RECS)} recordings, precision {precision}, {iters} iterations, "
         f"{ca.BATCH_POLICIES[policy]}")
    opts, data, states = _setup(precision, iters)
    solo = []
    for o, x, st in zip(opts, data, states):
        with ca.Engine(o) as e:
            _load(e, x, st)
            e.iterate(iters)
            e.sync()
            solo.append((e.log(iters), e.get_state()))
    with ca.Batch(opts) as b:
        for e, x, st in zip(b.engines, data, states):
            _load(e, x, st)
        b.set_policy(policy)
        b.iterate(iters)
        b.sync()
        for r, e in enumerate(b.engines):
            lg, st = e.log(iters), e.get_state()
            assert np.array_equal(lg["ll"], solo[r][0]["ll"]), r
            assert np.array_equal(lg["lrate"], solo[r][0]["lrate"]), r
            for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm"):
                assert np.array_equal(st[k], solo[r][1][k]), (r, k)
            _say(f"  recording {r} (n={RECS[r][0]}, T={RECS[r][1]}, M={RECS[r][2]}): identical, "
                 f"final LL {lg['ll'][-1]:.10f}")


def test_batch_tuner_restores_state(built):
    opts, data, states = _setup(ca.PREC_FP32, 8)
    first = []
    for o, x, st in zip(opts, data, states):
        with ca.Engine(o) as e:
            _load(e, x, st)
            e.iterate(1)
            e.sync()
            first.append(e.log(1)["ll"][0])
    with ca.Batch(opts) as b:
        for e, x, st in zip(b.engines, data, states):
            _load(e, x, st)
        p, ms = b.tune(iters=2)
        _say(f"[batch tune] ms per batched iteration by policy: "
             + ", ".join(f"{name} {t:.3f}" for name, t in zip(ca.BATCH_POLICIES, ms))
             + f"; chose {ca.BATCH_POLICIES[p]}")
        assert 0 <= p < len(ca.BATCH_POLICIES) and np.all(np.isfinite(ms)) and np.all(ms > 0)
        assert ms[p] == ms.min()
        b.iterate(3)
        b.sync()
        for r, e in enumerate(b.engines):
            assert e.iterations() == 3
            lg = e.log(3)
            assert abs(lg["ll"][0] - first[r]) <= 1e-6 * abs(first[r]), r
            assert np.all(np.isfinite(lg["ll"]))


@pytest.mark.parametrize("policy", [8, 10])
def test_grouped_launches_match_solo_engines(built, policy):
    """Grouped launches (one unmixing GEMM, one E-step and one statistic GEMM
    per channel width over a work list of every single-model BF16
    recording).  With each recording's own partition (policy 8) every
    partial sum covers the samples it covers in the recording's own
    launches: all 6 iterations bit for bit.  With 1024-sample items (policy
    10) the partial sums cover other sample ranges and the FP32 level of the
    sums rounds differently: the first iteration's LL (the E-step from the
    same state) within 1e-7, the second within 1e-5.  The 2-model recording
    runs its own branch in this mode: bit for bit."""
    iters = 6
    _say(f"[batch grouped] {len(RECS)} recordings, BF16, {iters} iterations, {ca.BATCH_POLICIES[policy]}")
    opts, data, states = _setup(ca.PREC_BF16, iters)
    solo = []
    for o, x, st in zip(opts, data, states):
        with ca.Engine(o) as e:
            _load(e, x, st)
            e.iterate(iters)
            e.sync()
            solo.append((e.log(iters), e.get_state()))
    with ca.Batch(opts) as b:
        for e, x, st in zip(b.engines, data, states):
            _load(e, x, st)
        b.set_policy(policy)
        b.iterate(iters)
        b.sync()
        for r, e in enumerate(b.engines):
            lg, st = e.log(iters), e.get_state()
            rel = np.abs(lg["ll"] - solo[r][0]["ll"]) / np.abs(solo[r][0]["ll"])
            _say(f"  recording {r} (n={RECS[r][0]}, T={RECS[r][1]}, M={RECS[r][2]}): LL rel. "
                 f"difference to the solo engine {np.array2string(rel, precision=2)}")
            if RECS[r][2] > 1 or policy == 8:
                assert np.array_equal(lg["ll"], solo[r][0]["ll"]), r
                for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm"):
                    assert np.array_equal(st[k], solo[r][1][k]), (r, k)
            else:
                assert rel[0] <= 1e-7, r
                assert rel[1] <= 1e-5, r
