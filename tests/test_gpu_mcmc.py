"""Learned-proposal MCMC on the device likelihood (config1: 20 x 120 nt)."""
import sys

import numpy as np
import pytest

import oracle
from messi_amd import lib, mcmc, model, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[gpu-test]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c1():
    a = synth.config1()
    return a, synth.dataset_of(a)


def test_posterior_matches_oracle(c1):
    a, ds = c1
    _say("posterior batch of 4, fp64, config1")
    post = mcmc.Posterior(ds, a.params, 4, prec="fp64")
    try:
        rng = np.random.default_rng(0)
        th = post.theta_of(a.params) + rng.normal(0, 0.2, (4, mcmc.DIM))
        ll = post.log_likelihood(th)
        lp = post(th)
        for i in range(4):
            p15, bl = post.params_of(th[i])
            mp = model.getparams(p15, bl)
            u = oracle.site_loglik(ds, model.site_components(mp))
            M = oracle.pair_loglik(ds, model.pair_components(mp))
            z = oracle.inside(u, M)[0, 0, -1]
            assert abs(ll[i] - z) <= 1e-9 * abs(z)
            jac = np.sum(th[i, 0:8]) + np.sum(np.log(post.unpack(th[i])[2])) + th[i, 11]
            want = z + model.computeprior(mp, samplebranchlengths=True) + jac
            assert lp[i] == pytest.approx(want, rel=1e-12, abs=1e-8)
    finally:
        post.close()


def test_learned_proposal_beats_random_walk(c1):
    a, ds = c1
    rng = np.random.default_rng(1)
    post = mcmc.Posterior(ds, a.params, 16, prec="fp64")
    net = lib.Mlp(mcmc.DIM, 128, seed=2)
    try:
        th0 = post.theta_of(a.params) + rng.normal(0, 0.05, (16, mcmc.DIM))
        res = mcmc.run_experiment(post, th0, net, rng, burn=1000, steps=500, long_steps=1500,
                                  train_steps=1500, log=_say)
        _say("ESS/s gain (min over parameters)", res["ess_per_s_gain_min"],
             "max |z| of means", res["max_abs_z_learned"])
        assert res["ess_per_s_gain_min"] >= 2.0
        assert res["max_abs_z_learned"] < 3.5
    finally:
        net.close()
        post.close()


def test_config2_posterior_and_learned_kernel_exactness():
    """50 x 1500, 12 parameters: the device log posterior against the oracle's
    full evaluation (FP64 within 1e-9, FP32 within 1e-6 relative); a short learned
    chain on it records, at every step, the log posterior of the state it holds
    (a fresh evaluation of that state gives the same value), and the proposal
    density it uses is the network's, which matches the numpy restatement."""
    from oracle import mlp as omlp
    a = synth.config2()
    ds = synth.dataset_of(a)
    _say("config2 posterior, 2 vectors, fp64 vs oracle")
    post = mcmc.Posterior(ds, a.params, 2, prec="fp64")
    post32 = mcmc.Posterior(ds, a.params, 2, prec="fp32")
    net = lib.Mlp(mcmc.DIM, 128, seed=3)
    try:
        rng = np.random.default_rng(4)
        th = post.theta_of(a.params) + rng.normal(0, 0.05, (2, mcmc.DIM))
        ll = post.log_likelihood(th)
        for i in range(2):
            p15, bl = post.params_of(th[i])
            mp = model.getparams(p15, bl)
            u = oracle.site_loglik(ds, model.site_components(mp))
            M = oracle.pair_loglik(ds, model.pair_components(mp))
            z = oracle.inside(u, M)[0, 0, -1]
            _say(f"config2 vector {i}: device {ll[i]:.9f} oracle {z:.9f}")
            assert abs(ll[i] - z) <= 1e-9 * abs(z)
        ll32 = post32.log_likelihood(th)
        assert np.max(np.abs(ll32 - ll) / np.abs(ll)) < 1e-6
        _say("config2 learned chain, 24 steps, fp32")
        pool = th[None] + rng.normal(0, 0.02, (64, 2, mcmc.DIM))
        wh = mcmc.Whitening(pool)
        lp0 = post32(th)
        tr = mcmc.learned_chain(post32, th, lp0, 24, net, wh, rng, 0.01 * np.eye(mcmc.DIM), 1.0, rw_prob=0.3)
        assert tr.samples.shape == (24, 2, mcmc.DIM)
        for it in (0, 11, 23):
            assert np.array_equal(post32(tr.samples[it]), tr.logp[it])
        z = wh.z(tr.samples[-1]).astype(np.float32)
        want, _ = omlp.forward(net.get_params(), z, mcmc.DIM, 128, emulate=True)
        got = net.forward(z)
        assert np.max(np.abs(got - want)) <= 1e-3 * np.max(np.abs(want)) + 1e-6
    finally:
        net.close()
        post.close()
        post32.close()
