"""The benchmark configurations at full size, through the C ABI, against the
CPU oracle.

  C2  64/512/2048        every precision mode, every pair energy and E
                         against the oracle's full FP64 energy
  C3  256 DEC fragments  (nocc ~ U{4..32}, nvirt ~ U{32..256}, naux = 4 nvirt)
                         every fragment's E in fp64, fp32 and i8s5 against the
                         oracle; in fp64 and fp32 under every tile candidate
  C4  128/1024/3072      surrogate trained on the exact pair energies of 512
                         seeded 64/512/2048 fragments, then the screened
                         energy of a 128/1024/3072 fragment: every kept pair
                         carries the exact kernel's energy, and a seeded sample
                         of the exact kernel's pair energies matches the oracle
  C5  256/2048/6144      B has 3.2e9 elements (> 2^31): fp64 pair energies of a
                         seeded subset (the first and the last units of the
                         launch included) against the oracle's pair loops on an
                         independent LAPACK metric (rel 1e-11); i8s5 within
                         1 uEh of fp64; fp32 and bf16x3 (the modes this
                         configuration compares) at their tolerances below

Tolerances (E_ref: the oracle's FP64 energy):
  fp64   |dE| <= 1e-11 |E_ref|, pair energies rel 1e-10
  i8s5   |dE| < 1 uEh; every pair energy within 1 uEh
  i8s4   |dE| < 1 uEh (C2)
  fp32   |dE| <= 2e-7 |E_ref|, pair energies rel 2e-6
  bf16x3 |dE| <= 2e-5 |E_ref|, pair energies rel 1e-4
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_E_REL = {"fp64": 1e-11, "fp32": 2e-7, "bf16x3": 2e-5}
TOL_E_ABS = {"i8s5": 1e-6, "i8s4": 1e-6}
TOL_PAIR_REL = {"fp64": 1e-10, "fp32": 2e-6, "bf16x3": 1e-4}
TOL_PAIR_ABS = {"i8s5": 1e-6, "i8s4": 1e-6}


def say(*a):
    print(*a, flush=True)


def check_E(mode, E, Eref):
    if mode in TOL_E_ABS:
        assert abs(E - Eref) < TOL_E_ABS[mode], (mode, E, Eref, E - Eref)
    else:
        assert abs(E - Eref) <= TOL_E_REL[mode] * abs(Eref), (mode, E, Eref, (E - Eref) / Eref)


def check_pairs(mode, eps, ref):
    if mode in TOL_PAIR_ABS:
        assert np.max(np.abs(eps - ref)) < TOL_PAIR_ABS[mode], (mode, np.max(np.abs(eps - ref)))
    else:
        np.testing.assert_allclose(eps, ref, rtol=TOL_PAIR_REL[mode], atol=0)


# ---------------------------------------------------------------------------
# C2: 64/512/2048, every mode against the full oracle
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(rimp2, oracle):
    naux, nvirt, nocc = 2048, 512, 64
    say("c2: generating and running the oracle", naux, nvirt, nocc)
    P, V, eo, ev = rimp2.gen_fragment(2, naux, nvirt, nocc)
    B = oracle.metric_transform(V, P)
    Eref, epsref = oracle.pair_energies(B, eo, ev)
    say("c2 oracle E", Eref)
    return P, V, eo, ev, Eref, epsref


@pytest.mark.parametrize("mode", ["fp64", "i8s5", "i8s4", "fp32", "bf16x3"])
def test_c2_every_mode_against_the_oracle(ctx, c2, mode):
    P, V, eo, ev, Eref, epsref = c2
    E, eps = ctx.energy(P, V, eo, ev, mode)
    say("c2", mode, "E", E, "dE uEh", (E - Eref) * 1e6, "max pair rel",
        float(np.max(np.abs(eps - epsref) / np.abs(epsref))))
    check_E(mode, E, Eref)
    check_pairs(mode, eps, epsref)
    np.testing.assert_array_equal(eps, eps.T)
    assert np.all(eps < 0)


# ---------------------------------------------------------------------------
# C3: the 256-fragment DEC batch, every fragment against the oracle
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c3(rimp2, oracle):
    o, v, q = rimp2.gen_batch_shapes(3, 256)
    say("c3: generating 256 fragments and running the oracle")
    frags = [rimp2.gen_fragment(3 + f, int(q[f]), int(v[f]), int(o[f])) for f in range(256)]
    ref = np.array([oracle.pair_energies(oracle.metric_transform(V, P), eo, ev)[0] for P, V, eo, ev in frags])
    return frags, ref


@pytest.mark.parametrize("mode", ["fp64", "fp32", "i8s5"])
def test_c3_every_fragment_against_the_oracle(ctx, c3, mode):
    frags, ref = c3
    ef, pp = ctx.batch(frags, mode)
    if mode in TOL_E_ABS:
        err = np.abs(ef - ref)
        say("c3", mode, "max |dE| uEh", err.max() * 1e6, "tile", ctx.batch_tile())
        assert err.max() < TOL_E_ABS[mode], int(np.argmax(err))
    else:
        rel = np.abs(ef - ref) / np.abs(ref)
        say("c3", mode, "max rel dE", rel.max(), "tile", ctx.batch_tile())
        assert rel.max() <= TOL_E_REL[mode], int(np.argmax(rel))
    # every fragment's pair matrix is symmetric and sums to its energy
    off = 0
    for k, (P, V, eo, ev) in enumerate(frags):
        n = eo.size
        eps = pp[off:off + n * n].reshape(n, n, order="F")
        off += n * n
        np.testing.assert_array_equal(eps, eps.T)
        assert abs(eps.sum() - ef[k]) <= 1e-12 * abs(ef[k])


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_c3_every_tile_candidate_against_the_oracle(rimp2, c3, mode):
    """The 256 fragments under each fixed tile candidate (every fragment 32-
    or 64-wide, and the per-fragment mix, whose launch carries both tile
    paths): every fragment's energy against the oracle, and under the mix each
    fragment equal bit for bit to the one-tile layout of its own tile."""
    frags, ref = c3
    res = {}
    with rimp2.Context(0) as c:
        for cand in (32, 64, -1):
            c.batch_set_tile(cand)
            c.batch_prepare(frags, mode)
            assert c.batch_tile() == cand
            c.batch_run()
            res[cand], _ = c.batch_result(want_pairs=False)
            rel = np.abs(res[cand] - ref) / np.abs(ref)
            say("c3", mode, "candidate", cand, "max rel dE", rel.max())
            assert rel.max() <= TOL_E_REL[mode], (cand, int(np.argmax(rel)))
    w32 = lambda v: -(-v // 32) * 32
    w64 = lambda v: -(-v // 64) * 64
    n32 = 0
    for k, (P, V, eo, ev) in enumerate(frags):
        nv = P.shape[1]
        t = 32 if w64(nv) ** 2 >= 1.5 * w32(nv) ** 2 else 64
        n32 += t == 32
        assert res[-1][k] == res[t][k], (k, nv, t)
    assert 0 < n32 < len(frags)


# ---------------------------------------------------------------------------
# C4: surrogate trained on 512 C2-shape fragments, screening a C4 fragment
# ---------------------------------------------------------------------------
def test_c4_surrogate_trained_on_512_fragments_screens_with_exact_energies(ctx, rimp2, oracle):
    import scipy.linalg as sl
    from rimp2_gpu_amd import surrogate
    thresh = 1e-6
    say("c4: training set from 512 seeded 64/512/2048 fragments (i8s5 exact pair energies)")
    F, T = surrogate.training_set(ctx, 512, seed0=100000, mode="i8s5", log=say)
    assert len(T) == 512 * 2080
    hist = surrogate.train(ctx, F, T, epochs=4, batch=4096, lr=2e-3, log=say)
    assert hist[-1] < hist[0]
    naux, nvirt, nocc = 3072, 1024, 128
    say("c4: screening", naux, nvirt, nocc)
    b = [ctx.alloc(naux * nvirt * nocc * 8), ctx.alloc(naux * naux * 8), ctx.alloc(nocc * 8), ctx.alloc(nvirt * 8)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    try:
        ctx.gen_device(4, naux, nvirt, nocc, *b)
        ctx.energy_device(naux, nvirt, nocc, *b, "i8s5", dpe, det)
        ctx.sync()
        E = float(ctx.from_device(det, (1,))[0])
        eps = ctx.from_device(dpe, (nocc, nocc))
        eo, ev = ctx.from_device(b[2], (nocc,)), ctx.from_device(b[3], (nvirt,))
        Fc = ctx.pair_features(eo, ev)
        pred = ctx.mlp_predict(Fc)
        Es, eps_s, nskip = ctx.screened_energy(eo, ev, "i8s5", thresh)
        # inputs of a seeded sample of pairs for the oracle (the last pair included)
        rng = np.random.default_rng(4)
        pi = np.concatenate([[0, 127, 3], rng.integers(0, nocc, 4)])
        pj = np.concatenate([[0, 127, 120], rng.integers(0, nocc, 4)])
        orbs = np.unique(np.concatenate([pi, pj]))
        V = ctx.from_device(b[1], (naux, naux))
        X = np.empty((naux, nvirt, len(orbs)), order="F")
        for k, i in enumerate(orbs):
            X[:, :, k] = ctx.from_device(b[0] + int(i) * naux * nvirt * 8, (naux, nvirt))
    finally:
        for p in b + [dpe, det]:
            ctx.free(p)
    jj = np.concatenate([np.full(j + 1, j) for j in range(nocc)])
    ii = np.concatenate([np.arange(j + 1) for j in range(nocc)])
    ltrue = np.log(-eps[ii, jj])
    rmse = float(np.sqrt(np.mean((pred - ltrue) ** 2)))
    base = float(np.sqrt(np.mean((Fc[:, 15].astype(np.float64) - ltrue) ** 2)))
    say("c4 E", E, "screened", Es, "skipped", nskip, "of", len(ii), "pred log rmse", rmse, "base", base,
        "min |eps|", float(np.min(-eps)))
    # the exact kernel at this shape: sampled i8s5 pair energies against the
    # oracle's pair loops on an independent LAPACK metric, within 1 uEh
    L = sl.cholesky(V, lower=True)
    B = np.asfortranarray(sl.solve_triangular(L, X.reshape(naux, -1, order="F"), lower=True).reshape(X.shape,
                                                                                                     order="F"))
    loc = {int(i): k for k, i in enumerate(orbs)}
    ref = oracle.pair_energies_subset(B, eo[orbs], ev, np.array([loc[int(i)] for i in pi]),
                                      np.array([loc[int(j)] for j in pj]))
    say("c4 sampled pairs |d| vs oracle", np.abs(eps[pi, pj] - ref))
    check_pairs("i8s5", eps[pi, pj], ref)
    # the network has learned the pair energies of an unseen, larger shape
    assert rmse < 0.5 and rmse < base
    # kept pairs carry the exact kernel's energies; a skipped pair contributes 0
    kept = eps_s != 0
    np.testing.assert_array_equal(eps_s[kept], eps[kept])
    assert nskip == int((~kept[ii, jj]).sum())
    mult = np.where(ii == jj, 1.0, 2.0)
    assert abs(Es - np.sum(mult * eps_s[ii, jj])) <= 1e-12 * abs(Es)
    # on this generator no pair is below the threshold: nothing may be skipped
    if np.min(-eps) > 100 * thresh:
        assert nskip == 0 and Es == E


def test_c4_shape_locality_variant_network_screen_beats_the_network_free_screen(ctx, rimp2):
    """The same pipeline on the locality-decay variant of the generator (orbitals
    on a chain, lambda = 1.5), where small pairs exist: trained on 128 seeded
    64/512/2048 fragments, applied to a 128/1024/3072 fragment.  At 1e-6 Eh the
    network skips about the pairs an oracle screen (one that knows every exact
    pair energy) skips, predicts log(-eps_ij) better than the base estimate
    (feature 15) and loses less energy than the network-free screen on that
    estimate; the skipped pairs are what the screened energy leaves out, and
    the screened pair stage is faster than the exact one."""
    from rimp2_gpu_amd import surrogate
    lam, thresh = 1.5, 1e-6
    say("c4 locality: training set from 128 seeded 64/512/2048 fragments, lambda", lam)
    F, T = surrogate.training_set(ctx, 128, seed0=100000, lam=lam, mode="i8s5", log=say)
    hist = surrogate.train(ctx, F, T, epochs=4, batch=4096, lr=2e-3, log=say)
    assert hist[-1] < hist[0]
    rep = surrogate.screen_report(ctx, 128, 1024, 3072, 4, lam, thresh, mode="i8s5")
    say("c4 locality", {k: rep[k] for k in ("npairs", "nskipped", "mlp_skipped", "bound_skipped", "ideal_skipped",
                                            "dE_screened", "dE_bound_screen", "dE_ideal_screen", "pred_log_rmse",
                                            "base_log_rmse", "t_exact_s", "t_screened_s")})
    assert rep["E_exact_again"] == rep["E_exact"]
    assert rep["ideal_skipped"] > rep["npairs"] // 2
    assert rep["nskipped"] == rep["mlp_skipped"]
    assert rep["mlp_skipped"] >= 0.9 * rep["ideal_skipped"]
    assert rep["pred_log_rmse"] < rep["base_log_rmse"]
    # E_screened - E_exact = the (positive) energy of the skipped pairs
    assert rep["dE_screened"] >= 0.0
    assert rep["dE_screened"] < rep["dE_bound_screen"]
    assert rep["t_screened_s"] < 0.5 * rep["t_exact_s"]


# ---------------------------------------------------------------------------
# C5: 256/2048/6144, B beyond 2^31 elements
# ---------------------------------------------------------------------------
def test_c5_pair_energies_and_exact_slice_energy(ctx, oracle, rimp2):
    import scipy.linalg as sl
    naux, nvirt, nocc = 6144, 2048, 256
    say("c5", naux, nvirt, nocc, "B elements", naux * nvirt * nocc)
    per = naux * nvirt * 8
    b = [ctx.alloc(per * nocc), ctx.alloc(naux * naux * 8), ctx.alloc(nocc * 8), ctx.alloc(nvirt * 8)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    # pairs: the first and the last units of the launch, the last diagonal
    # pair, and a seeded sample
    rng = np.random.default_rng(5)
    pi = np.concatenate([[0, 0, 254, 255, 1], rng.integers(0, nocc, 3)])
    pj = np.concatenate([[0, 1, 255, 255, 200], rng.integers(0, nocc, 3)])
    orbs = np.unique(np.concatenate([pi, pj]))
    try:
        ctx.gen_device(5, naux, nvirt, nocc, *b)
        say("c5 fp64")
        ctx.energy_device(naux, nvirt, nocc, *b, "fp64", dpe, det)
        ctx.sync()
        E64 = float(ctx.from_device(det, (1,))[0])
        eps64 = ctx.from_device(dpe, (nocc, nocc))
        say("c5 i8s5")
        ctx.energy_device(naux, nvirt, nocc, *b, "i8s5", dpe, det)
        ctx.sync()
        E8 = float(ctx.from_device(det, (1,))[0])
        eps8 = ctx.from_device(dpe, (nocc, nocc))
        # the modes this configuration compares with fp64: FP32 and split-BF16
        # MFMA with FP32 accumulation (packed operands of 12.9 and 25.8 GB)
        other = {}
        for mode in ("fp32", "bf16x3"):
            say("c5", mode)
            ctx.energy_device(naux, nvirt, nocc, *b, mode, dpe, det)
            ctx.sync()
            other[mode] = (float(ctx.from_device(det, (1,))[0]), ctx.from_device(dpe, (nocc, nocc)))
        V = ctx.from_device(b[1], (naux, naux))
        eo, ev = ctx.from_device(b[2], (nocc,)), ctx.from_device(b[3], (nvirt,))
        X = np.empty((naux, nvirt, len(orbs)), order="F")
        for k, i in enumerate(orbs):
            X[:, :, k] = ctx.from_device(b[0] + int(i) * per, (naux, nvirt))
    finally:
        for p in b + [dpe, det]:
            ctx.free(p)
    say("c5 E64", E64, "E_i8s5", E8, "dE uEh", (E8 - E64) * 1e6)
    assert abs(E8 - E64) < 1e-6
    assert np.max(np.abs(eps8 - eps64)) < 1e-6
    np.testing.assert_array_equal(eps64, eps64.T)
    assert np.all(eps64 < 0)
    # oracle: LAPACK metric (dpotrf + dtrsm) on the needed orbitals, then the
    # oracle's pair loops
    L = sl.cholesky(V, lower=True)
    B = sl.solve_triangular(L, X.reshape(naux, -1, order="F"), lower=True).reshape(X.shape, order="F")
    B = np.asfortranarray(B)
    loc = {int(i): k for k, i in enumerate(orbs)}
    li = np.array([loc[int(i)] for i in pi])
    lj = np.array([loc[int(j)] for j in pj])
    ref = oracle.pair_energies_subset(B, eo[orbs], ev, li, lj)
    got = eps64[pi, pj]
    say("c5 pair rel err", np.abs(got - ref) / np.abs(ref))
    np.testing.assert_allclose(got, ref, rtol=1e-11)
    # the total is the sum of the pair matrix
    assert abs(eps64.sum() - E64) <= 1e-12 * abs(E64)
    # FP32 and split BF16: total against the fp64 total (checked above against
    # the oracle to 1e-11) and the sampled pairs against the oracle, at their
    # stated relative tolerances
    for mode, (Em, epsm) in other.items():
        say("c5", mode, "E", Em, "dE", Em - E64, "rel", (Em - E64) / E64, "pair rel",
            np.abs(epsm[pi, pj] - ref) / np.abs(ref))
        check_E(mode, Em, E64)
        check_pairs(mode, epsm[pi, pj], ref)
        np.testing.assert_array_equal(epsm, epsm.T)
