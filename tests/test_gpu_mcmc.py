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
