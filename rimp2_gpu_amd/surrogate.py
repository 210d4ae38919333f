"""Pair-energy surrogate: training-set synthesis, training and screening.

The network (BF16 MLP 16 -> 64 -> 64 -> 1, GELU) and its forward, backward and
Adam kernels live in the HIP library (mlp.hip); this module drives them:

  * training_set: fragments are generated on the GPU, their exact pair
    energies computed by the pair kernel, and the 16 per-pair features taken
    from the same transformed B (rimp2_pair_features); optionally cached per
    fragment, so an interrupted run resumes where it stopped;
  * train: standardises features / target (log(-eps_ij)) and runs Adam steps
    on minibatches;
  * screen_report: runs the exact energy, the MLP-screened energy and a
    network-free screen (the feature-15 base estimate at the same threshold)
    and reports skipped pairs, energy errors and times.

No part of the energy ever comes from the network: screened energies are the
exact kernel's energies over the kept pairs.
"""
import os
import time

import numpy as np

from . import Context, pair_flops  # noqa: F401

NFEAT = 16


def _device_fragment(ctx, seed, naux, nvirt, nocc, lam, bufs):
    dP, dV, deo, dev = bufs
    ctx.gen_device(seed, naux, nvirt, nocc, dP, dV, deo, dev, lam)
    eo = ctx.from_device(deo, (nocc,))
    ev = ctx.from_device(dev, (nvirt,))
    return eo, ev


def _cache_path(cache_dir, seed, nocc, nvirt, naux, lam, mode):
    return os.path.join(cache_dir, f"frag_{seed}_{nocc}_{nvirt}_{naux}_{lam!r}_{mode}.npz")


def training_set(ctx, nfrag, nocc=64, nvirt=512, naux=2048, seed0=100000, lam=0.0, mode="i8s5", log=None,
                 min_log_eps=-46.0, cache_dir=None):
    """Features [npairs, 16] and targets log(-eps_ij) over nfrag
    device-generated fragments (seeds seed0 .. seed0+nfrag-1).  Pairs with
    -eps_ij below exp(min_log_eps) (1e-20 by default, far below any screening
    threshold; some underflow to zero) are left out: they cannot change a
    screening decision and would dominate the residual's statistics.

    cache_dir: checkpoint of a long training-set run.  Each fragment's
    features and targets are written there (one .npz per fragment, named by
    seed, shape, lam and mode; written to a temporary name and renamed, so a
    crash leaves no partial file) and read back instead of recomputed when
    the run is repeated."""
    bufs = [ctx.alloc(naux * nvirt * nocc * 8), ctx.alloc(naux * naux * 8), ctx.alloc(nocc * 8), ctx.alloc(nvirt * 8)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    # unique pairs i <= j, ordered j-major as the features are
    jj = np.concatenate([np.full(j + 1, j) for j in range(nocc)])
    ii = np.concatenate([np.arange(j + 1) for j in range(nocc)])
    if cache_dir:
        os.makedirs(cache_dir, exist_ok=True)
    F, T = [], []
    reused = 0
    try:
        for f in range(nfrag):
            cp = _cache_path(cache_dir, seed0 + f, nocc, nvirt, naux, lam, mode) if cache_dir else None
            if cp and os.path.exists(cp):
                with np.load(cp) as z:
                    F.append(z["F"])
                    T.append(z["T"])
                reused += 1
                continue
            eo, ev = _device_fragment(ctx, seed0 + f, naux, nvirt, nocc, lam, bufs)
            ctx.energy_device(naux, nvirt, nocc, bufs[0], bufs[1], bufs[2], bufs[3], mode, dpe, det)
            ctx.sync()
            eps = ctx.from_device(dpe, (nocc, nocc))
            F.append(ctx.pair_features(eo, ev))
            T.append(np.log(np.maximum(-eps[ii, jj], 1e-300)).astype(np.float32))
            if cp:
                tmp = cp + ".tmp.npz"
                np.savez(tmp, F=F[-1], T=T[-1])
                os.replace(tmp, cp)
            if log and (f % 32 == 0 or f == nfrag - 1):
                log(f"[surrogate] training fragment {f + 1}/{nfrag}")
    finally:
        for p in bufs + [dpe, det]:
            ctx.free(p)
    if log and reused:
        log(f"[surrogate] {reused} of {nfrag} training fragments read from {cache_dir}")
    F, T = np.concatenate(F), np.concatenate(T)
    keep = T > min_log_eps
    return F[keep], T[keep]


def standardise(F, T):
    """Feature statistics, and those of the residual target T - F[:, 15] (the
    network corrects the base estimate carried in feature 15)."""
    mean = F.mean(axis=0).astype(np.float32)
    std = F.std(axis=0).astype(np.float32)
    std[std < 1e-6] = 1.0
    res = T - F[:, 15]
    return mean, std, np.float32(res.mean()), np.float32(max(res.std(), 1e-6))


def train(ctx, F, T, epochs=4, batch=4096, lr=2e-3, seed=0, log=None):
    """Adam on minibatches; returns the per-epoch mean training loss."""
    mean, std, tm, ts = standardise(F, T)
    ctx.mlp_init(seed)
    ctx.mlp_set_norm(mean, std, tm, ts)
    rng = np.random.default_rng(seed)
    n = F.shape[0]
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        losses = []
        step_lr = lr * (0.5 ** ep)
        for b0 in range(0, n, batch):
            idx = perm[b0:b0 + batch]
            losses.append(ctx.mlp_train_step(F[idx], T[idx], step_lr))
        hist.append(float(np.mean(losses)))
        if log:
            log(f"[surrogate] epoch {ep + 1}/{epochs} loss {hist[-1]:.5f}")
    return hist


def screen_report(ctx, nocc, nvirt, naux, seed, lam, thresh, mode="i8s5"):
    """Exact vs MLP-screened vs bound-screened energies of one fragment."""
    bufs = [ctx.alloc(naux * nvirt * nocc * 8), ctx.alloc(naux * naux * 8), ctx.alloc(nocc * 8), ctx.alloc(nvirt * 8)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    try:
        eo, ev = _device_fragment(ctx, seed, naux, nvirt, nocc, lam, bufs)
        ctx.energy_device(naux, nvirt, nocc, bufs[0], bufs[1], bufs[2], bufs[3], mode, dpe, det)
        ctx.sync()
        E = float(ctx.from_device(det, (1,))[0])
        eps = ctx.from_device(dpe, (nocc, nocc))
        # time the exact pair stage alone (same transform in the context)
        t0 = time.perf_counter()
        E2, _ = ctx.pair_energies(eo, ev, mode)
        t_exact = time.perf_counter() - t0
        F = ctx.pair_features(eo, ev)
        pred = ctx.mlp_predict(F)
        t0 = time.perf_counter()
        Es, eps_s, nskip = ctx.screened_energy(eo, ev, mode, thresh)
        t_screen = time.perf_counter() - t0
    finally:
        for p in bufs + [dpe, det]:
            ctx.free(p)
    jj = np.concatenate([np.full(j + 1, j) for j in range(nocc)])
    ii = np.concatenate([np.arange(j + 1) for j in range(nocc)])
    true = -eps[ii, jj]
    mult = np.where(ii == jj, 1.0, 2.0)
    # network-free screen: the base estimate exp(feature 15) at the same threshold
    bound_keep = np.exp(F[:, 15].astype(np.float64)) >= thresh
    # E_screened - E_exact = sum over skipped pairs of mult * |eps_ij| (eps_ij < 0)
    dE_bound = float(np.sum(mult * true * (~bound_keep)))
    mlp_keep = np.exp(pred.astype(np.float64)) >= thresh
    # a screen that knows the exact pair energies (the best any screen at this threshold can do)
    ideal_keep = true >= thresh
    dE_ideal = float(np.sum(mult * true * (~ideal_keep)))
    # prediction error over the pairs that can matter (some far pairs underflow to zero)
    rel = true > 1e-20
    ltrue = np.log(true[rel])
    return dict(
        E_exact=E, E_exact_again=E2, E_screened=Es, dE_screened=Es - E, nskipped=int(nskip),
        npairs=int(len(true)), t_exact_s=t_exact, t_screened_s=t_screen,
        mlp_skipped=int((~mlp_keep).sum()), bound_skipped=int((~bound_keep).sum()),
        ideal_skipped=int((~ideal_keep).sum()), dE_bound_screen=dE_bound, dE_ideal_screen=dE_ideal,
        pred_log_rmse=float(np.sqrt(np.mean((pred[rel] - ltrue) ** 2))),
        base_log_rmse=float(np.sqrt(np.mean((F[rel, 15].astype(np.float64) - ltrue) ** 2))),
    )
