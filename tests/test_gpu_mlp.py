"""Pair-screening surrogate on the GPU.

Parity: the GPU MLP forward and backward (BF16 operands, FP32 accumulation;
the network output is a residual on the base estimate in feature 15)
against the FP64 oracle evaluated with the same BF16 rounding points
(oracle.mlp_loss_grad(emulate_bf16=True)); what remains is FP32-vs-FP64
accumulation order and the rare bf16 rounding flip, so predictions are held to
1e-3 relative and gradients to 2e-2 in norm per parameter block.  Features
are recomputed in numpy from the FP64 B.  Screening: the exact kernel's pair
energies on the kept pairs are bitwise those of the unscreened run.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BLOCKS = {"W1": (0, 1024), "b1": (1024, 1088), "W2": (1088, 5184), "b2": (5184, 5248), "w3": (5248, 5312),
          "b3": (5312, 5313)}


def say(*a):
    print(*a, flush=True)


def case(seed, n):
    rng = np.random.default_rng(seed)
    P = (rng.standard_normal(5313) * 0.2).astype(np.float32)
    F = (rng.standard_normal((n, 16)) * 2 + 1).astype(np.float32)
    T = (rng.standard_normal(n) + 3).astype(np.float32)
    norm = np.concatenate([np.full(16, 1.0), np.full(16, 2.0), [3.0, 1.5]]).astype(np.float32)
    return P, F, T, norm


def load(ctx, P, norm):
    ctx.mlp_init(1)
    ctx.mlp_set_norm(norm[:16], norm[16:32], norm[32], norm[33])
    ctx.mlp_set_params(P)


@pytest.mark.parametrize("n", [1, 63, 64, 200])
def test_forward_matches_oracle(ctx, oracle, n):
    say("mlp forward", n)
    P, F, T, norm = case(10 + n, n)
    load(ctx, P, norm)
    pred = ctx.mlp_predict(F)
    _, _, ref = oracle.mlp_loss_grad(F, None, P, norm, emulate_bf16=True, want_grad=False)
    assert np.max(np.abs(pred - ref)) <= 1e-3 * np.max(np.abs(ref)) + 1e-4


@pytest.mark.parametrize("n", [64, 130, 1000])
def test_gradient_matches_oracle(ctx, oracle, n):
    say("mlp gradient", n)
    P, F, T, norm = case(20 + n, n)
    load(ctx, P, norm)
    loss, g = ctx.mlp_grad(F, T)
    lref, gref, _ = oracle.mlp_loss_grad(F, T, P, norm, emulate_bf16=True)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    for name, (a, b) in BLOCKS.items():
        err = np.linalg.norm(g[a:b] - gref[a:b])
        assert err <= 2e-2 * np.linalg.norm(gref[a:b]) + 1e-7, (name, err, np.linalg.norm(gref[a:b]))


def test_training_fits_a_smooth_target_and_is_deterministic(ctx):
    say("mlp training")
    rng = np.random.default_rng(5)
    F = rng.standard_normal((8192, 16)).astype(np.float32)
    w = rng.standard_normal(16)
    T = (np.tanh(F @ w / 4) * 2 + 0.3 * F[:, 0] ** 2).astype(np.float32)
    norm_mean, norm_std = F.mean(0), F.std(0)
    runs = []
    for rep in range(2):
        ctx.mlp_init(7)
        ctx.mlp_set_norm(norm_mean, norm_std, float(T.mean()), float(T.std()))
        losses = [ctx.mlp_train_step(F[k:k + 1024], T[k:k + 1024], 3e-3) for ep in range(12)
                  for k in range(0, 8192, 1024)]
        runs.append((losses, ctx.mlp_params()))
    l0, l1 = runs[0][0][0], runs[0][0][-1]
    say("loss", l0, "->", l1)
    assert l1 < 0.1 * l0
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def test_features_match_numpy(ctx, rimp2):
    naux, nvirt, nocc = 96, 20, 5
    say("features")
    P, V, eo, ev = rimp2.gen_fragment(41, naux, nvirt, nocc)
    ctx.energy(P, V, eo, ev, "fp64")
    B = ctx.get_B(naux, nvirt, nocc)
    F = ctx.pair_features(eo, ev)
    r = np.sqrt((B ** 2).sum(axis=1))  # [naux, nocc]
    slab = (np.arange(naux) * 8) // naux
    k = 0
    for j in range(nocc):
        for i in range(j + 1):
            fs = np.array([np.sum(r[slab == s, i] * r[slab == s, j]) for s in range(8)])
            tot = fs.sum()
            s2 = np.sum((r[:, i] * r[:, j]) ** 2)
            ni, nj = np.sqrt((B[:, :, i] ** 2).sum()), np.sqrt((B[:, :, j] ** 2).sum())
            gmean = 2 * ev.mean() - eo[i] - eo[j]
            exp = np.concatenate([[np.log(ni * nj / tot), np.log(ni / nj)], np.log(fs / tot),
                                  [np.log(tot * tot / s2), eo[i] + eo[j], 2 * ev.min() - eo[i] - eo[j], gmean,
                                   float(i == j), np.log(s2 / gmean)]])
            np.testing.assert_allclose(F[k], exp, rtol=2e-5, atol=2e-5)
            k += 1


def test_screened_energy_comes_from_the_exact_kernel(ctx, rimp2):
    """Locality-decay fragments (labelled variant of the generator): train on
    12 fragments, screen a 13th; kept pairs carry the exact kernel's energies."""
    from rimp2_gpu_amd import surrogate
    nocc, nvirt, naux, lam = 32, 64, 256, 1.5
    say("screening")
    F, T = surrogate.training_set(ctx, 12, nocc=nocc, nvirt=nvirt, naux=naux, seed0=500, lam=lam, mode="fp64")
    hist = surrogate.train(ctx, F, T, epochs=6, batch=512, lr=3e-3)
    say("train loss", hist)
    assert hist[-1] < hist[0]
    P, V, eo, ev = rimp2.gen_fragment_local(999, naux, nvirt, nocc, lam)
    E, eps = ctx.energy(P, V, eo, ev, "fp64")
    thresh = 1e-6
    Es, eps_s, nskip = ctx.screened_energy(eo, ev, "fp64", thresh)
    kept = eps_s != 0
    np.testing.assert_array_equal(eps_s[kept], eps[kept])
    skipped = ~kept
    say("skipped", nskip, "of", nocc * (nocc + 1) // 2, "dE", Es - E, "max skipped |eps|",
        np.abs(eps[skipped]).max() if skipped.any() else 0.0)
    assert nskip > 0
    assert np.abs(eps[skipped]).max() < 100 * thresh
    assert abs(Es - E) < 2 * nskip * 100 * thresh


def test_network_gains_over_the_base_estimate_across_shapes(ctx, rimp2):
    """Trained on 32/128/512 locality fragments, applied to a 48/192/768 one
    (another shape): at the same threshold the network's screen skips about
    the pairs an oracle screen would and loses less energy than the network-
    free screen on the base estimate (feature 15) alone."""
    from rimp2_gpu_amd import surrogate
    lam, thresh = 1.5, 1e-6
    say("gain over base estimate")
    F, T = surrogate.training_set(ctx, 48, nocc=32, nvirt=128, naux=512, seed0=700, lam=lam, mode="fp64")
    surrogate.train(ctx, F, T, epochs=6, batch=1024, lr=3e-3)
    rep = surrogate.screen_report(ctx, 48, 192, 768, 4242, lam, thresh, mode="fp64")
    say({k: rep[k] for k in ("mlp_skipped", "bound_skipped", "ideal_skipped", "dE_screened", "dE_bound_screen",
                             "dE_ideal_screen", "pred_log_rmse", "base_log_rmse")})
    assert rep["ideal_skipped"] > 0
    assert rep["pred_log_rmse"] < rep["base_log_rmse"]
    assert abs(rep["dE_screened"]) < abs(rep["dE_bound_screen"])
    assert rep["mlp_skipped"] >= 0.9 * rep["ideal_skipped"]


def test_training_set_cache_resumes_without_recomputing(ctx, rimp2, tmp_path):
    """With cache_dir, a repeated training-set run reads every fragment back
    (the pair kernel is not called again) and returns the same arrays; a
    partial cache (an interrupted run) is completed."""
    from rimp2_gpu_amd import surrogate
    kw = dict(nocc=8, nvirt=32, naux=128, seed0=900, lam=1.5, mode="fp64")
    say("training-set cache")
    F0, T0 = surrogate.training_set(ctx, 4, **kw)
    d = str(tmp_path / "cache")
    F1, T1 = surrogate.training_set(ctx, 2, cache_dir=d, **kw)  # the interrupted run
    assert len(os.listdir(d)) == 2
    np.testing.assert_array_equal(F1, F0[:len(F1)])
    np.testing.assert_array_equal(T1, T0[:len(T1)])

    class CountEnergy:
        def __init__(self, c):
            self.c, self.calls = c, 0

        def __getattr__(self, name):
            return getattr(self.c, name)

        def energy_device(self, *a):
            self.calls += 1
            return self.c.energy_device(*a)

    proxy = CountEnergy(ctx)
    F2, T2 = surrogate.training_set(proxy, 4, cache_dir=d, **kw)
    assert proxy.calls == 2 and len(os.listdir(d)) == 4
    np.testing.assert_array_equal(F2, F0)
    np.testing.assert_array_equal(T2, T0)
    proxy = CountEnergy(ctx)
    F3, T3 = surrogate.training_set(proxy, 4, cache_dir=d, **kw)
    assert proxy.calls == 0
    np.testing.assert_array_equal(F3, F0)
    np.testing.assert_array_equal(T3, T0)
