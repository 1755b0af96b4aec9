"""Sweep surrogate on the GPU, at reduced size: the trained ensemble predicts held-out variants'
readout marginals better than the method without it (the training-set mean, and the same
ensemble untrained), it answers far faster than plain simulation of matching accuracy, and the
spread gate hands every rejected point to the simulator, whose marginals are returned unchanged."""
import numpy as np
import pytest

from upmabossg_amd import Context, Model, synth
from upmabossg_amd.surrogate import ParamSpace, SweepSurrogate, simulate_marginals


@pytest.mark.gpu
def test_surrogate_gain_and_fallback():
    net = synth.config_surrogate(1)
    base = Model(net.bnd(), net.cfg())
    space = ParamSpace(net)
    with Context(0) as ctx:
        mu, md, ist = space.sample(4096, 11)
        Y, _, _ = simulate_marginals(ctx, base, mu, md, ist, 512, seed=5)
        sur = SweepSurrogate(ctx, base, members=2, seed=3)
        sur.fit(mu, md, ist, Y, epochs=1, stages=1)
        tmu, tmd, tist = space.sample(32, 12)
        early, _, _, _ = sur.ens.predict(ParamSpace.features(tmu, tmd, tist))
        sur.fit(mu, md, ist, Y, epochs=40)
        R, _, _ = simulate_marginals(ctx, base, tmu, tmd, tist, 16384, seed=7, variant_base=10 ** 6)

        pred, spread, _, wall = sur.ens.predict(ParamSpace.features(tmu, tmd, tist))
        rms = lambda a: float(np.sqrt(((a - R) ** 2).mean()))
        rms_sur, rms_mean, rms_early = rms(pred), rms(np.broadcast_to(Y.mean(axis=0), R.shape)), rms(early)
        assert rms_sur < 0.7 * rms_mean, (rms_sur, rms_mean)
        assert rms_sur < 0.7 * rms_early, (rms_sur, rms_early)

        # plain simulation at the trajectory count that matches the surrogate's RMS error
        for T in (16, 64, 256, 1024, 4096):
            P, _, wall_p = simulate_marginals(ctx, base, tmu, tmd, tist, T, seed=9, chunk=32, variant_base=2 * 10 ** 6)
            if rms(P) <= rms_sur:
                break
        assert len(tmu) / wall > 10 * len(tmu) / wall_p

        # gate: with a threshold between the spreads, both paths are taken
        th = float(np.median(spread))
        out, acc, info = sur.screen(tmu, tmd, tist, 256, threshold=th, seed=21, variant_base=500)
        assert acc.any() and (~acc).any()
        assert np.array_equal(out[acc], pred[acc])
        rej = np.flatnonzero(~acc)
        D, _, _ = simulate_marginals(ctx, base, tmu[rej[:1]], tmd[rej[:1]], tist[rej[:1]], 256, seed=21,
                                     variant_base=500 + int(rej[0]))
        assert np.array_equal(out[rej[0]], D[0])
        assert info["fallback_s"] > 0
