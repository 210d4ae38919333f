"""Batched DEC fragments through one persistent, graph-captured launch.

Parity: every fragment's energy and pair matrix against the CPU oracle (mode
tolerances as in test_gpu_parity.py).  The captured graph is replayed several
times with eager work and synchronisations between replays; every replay
must reproduce the first result bit for bit.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_REL = {"fp64": 1e-11, "fp32": 2e-7, "bf16x3": 2e-5}
TOL_ABS = {"i8s5": 1e-6, "i8s4": 1e-6}


def say(*a):
    print(*a, flush=True)


def make_batch(rimp2, seed, nfrag, scale=1):
    o, v, q = rimp2.gen_batch_shapes(seed, nfrag)
    frags = []
    for f in range(nfrag):
        nocc, nvirt, naux = int(o[f]) // scale or 1, int(v[f]) // scale or 1, int(q[f]) // scale or 1
        frags.append(rimp2.gen_fragment(seed + f, naux, nvirt, nocc))
    return frags


def oracle_energies(oracle, frags):
    out = []
    for P, V, eo, ev in frags:
        B = oracle.metric_transform(V, P)
        out.append(oracle.pair_energies(B, eo, ev))
    return out


@pytest.mark.parametrize("mode", ["fp64", "fp32", "i8s5", "i8s4", "bf16x3"])
def test_batch_matches_oracle(ctx, oracle, rimp2, mode):
    frags = make_batch(rimp2, 11, 9, scale=4)
    say("batch", mode, [f[0].shape for f in frags])
    ref = oracle_energies(oracle, frags)
    ef, pp = ctx.batch(frags, mode)
    off = 0
    for k, (P, V, eo, ev) in enumerate(frags):
        Eref, epsref = ref[k]
        n = eo.size
        eps = pp[off:off + n * n].reshape(n, n, order="F")
        off += n * n
        if mode in TOL_ABS:
            assert abs(ef[k] - Eref) < TOL_ABS[mode], (k, ef[k], Eref)
        else:
            assert abs(ef[k] - Eref) <= TOL_REL[mode] * abs(Eref), (k, ef[k], Eref)
        if mode == "fp64":
            np.testing.assert_allclose(eps, epsref, rtol=1e-10, atol=1e-15)
        np.testing.assert_array_equal(eps, eps.T)
    assert ctx.batch_tile() in (-1, 32, 64)


def test_batch_graph_replays_are_identical(ctx, rimp2):
    frags = make_batch(rimp2, 21, 12, scale=2)
    say("graph replay", len(frags))
    ctx.batch_prepare(frags, "fp32")
    ctx.batch_run()
    e0, p0 = ctx.batch_result()
    P, V, eo, ev = frags[0]
    for rep in range(3):
        # eager work on the same context between replays
        E_single, _ = ctx.energy(P, V, eo, ev, "fp64")
        assert E_single < 0
        ctx.sync()
        ctx.batch_run()
        e1, p1 = ctx.batch_result()
        np.testing.assert_array_equal(e1, e0)
        np.testing.assert_array_equal(p1, p0)


def test_batch_equals_single_fragment_path(ctx, rimp2):
    frags = make_batch(rimp2, 31, 6, scale=2)
    say("batch vs single", len(frags))
    ef, _ = ctx.batch(frags, "i8s5")
    for k, (P, V, eo, ev) in enumerate(frags):
        E, _ = ctx.energy(P, V, eo, ev, "i8s5")
        assert abs(E - ef[k]) < 1e-9 * abs(E) + 1e-9


def test_batch_errors(ctx, rimp2):
    say("batch errors")
    P, V, eo, ev = rimp2.gen_fragment(3, 40, 10, 3)
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.batch([(P, V, eo, ev - 10.0)], "fp64")
    assert ei.value.code == rimp2.ERR_NUMERIC
    # one NaN orbital energy among finite ones (max / min would skip it)
    for k, which in [(1, "eo"), (4, "ev")]:
        eon, evn = eo.copy(), ev.copy()
        (eon if which == "eo" else evn)[k] = np.nan
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.batch([(P, V, eo, ev), (P, V, eon, evn)], "fp64")
        assert ei.value.code == rimp2.ERR_NUMERIC
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.batch([], "fp64")
    assert ei.value.code == rimp2.ERR_ARG
    # still usable
    ef, _ = ctx.batch([(P, V, eo, ev)], "fp64")
    assert ef[0] < 0


@pytest.mark.parametrize("mode", ["fp64", "i8s5"])
def test_batch_mixes_fitted_and_metric_fragments(ctx, oracle, rimp2, mode):
    """A descriptor with V = NULL carries fitted coefficients (B = Pia)."""
    frags = make_batch(rimp2, 23, 6, scale=4)
    rng = np.random.default_rng(5)
    mixed, ref = [], []
    for k, (P, V, eo, ev) in enumerate(frags):
        if k % 2:
            C = np.asfortranarray(rng.standard_normal(P.shape) * 0.05)
            mixed.append((C, None, eo, ev))
            ref.append(oracle.pair_energies(C, eo, ev)[0])
        else:
            mixed.append((P, V, eo, ev))
            ref.append(oracle.pair_energies(oracle.metric_transform(V, P), eo, ev)[0])
    say("batch mixed", mode, [f[0].shape for f in mixed])
    ef, _ = ctx.batch(mixed, mode, want_pairs=False)
    for k in range(len(mixed)):
        if mode in TOL_ABS:
            assert abs(ef[k] - ref[k]) < TOL_ABS[mode], (k, ef[k], ref[k])
        else:
            assert abs(ef[k] - ref[k]) <= TOL_REL[mode] * abs(ref[k]), (k, ef[k], ref[k])


def _to_device_descs(ctx, frags):
    descs, ptrs = [], []
    for P, V, eo, ev in frags:
        dP, deo, dev = ctx.to_device(P), ctx.to_device(eo), ctx.to_device(ev)
        dV = None if V is None else ctx.to_device(V)
        ptrs += [p for p in (dP, dV, deo, dev) if p is not None]
        naux, nvirt, nocc = P.shape
        descs.append((naux, nvirt, nocc, dP, dV, deo, dev))
    return descs, ptrs


@pytest.mark.parametrize("mode", ["fp64", "i8s5"])
def test_batch_prepare_device_matches_host(ctx, rimp2, mode):
    """Descriptors holding device pointers (inputs already in HBM) give the
    fragment energies and pair matrices of host descriptors bit for bit (the
    same kernels run on the same data); a fitted fragment (V = NULL) is among
    them."""
    frags = make_batch(rimp2, 41, 7, scale=4)
    P, V, eo, ev = frags[3]
    frags[3] = (np.asfortranarray(np.random.default_rng(2).standard_normal(P.shape) * 0.05), None, eo, ev)
    say("batch prepare device", mode, [f[0].shape for f in frags])
    ctx.batch_prepare(frags, mode)
    ctx.batch_run()
    ef_h, pp_h = ctx.batch_result()
    descs, ptrs = _to_device_descs(ctx, frags)
    try:
        ctx.batch_prepare_device(descs, mode)
        ctx.batch_run()
        ef_d, pp_d = ctx.batch_result()
    finally:
        for p in ptrs:
            ctx.free(p)
    np.testing.assert_array_equal(ef_d, ef_h)
    np.testing.assert_array_equal(pp_d, pp_h)


@pytest.mark.parametrize("mode", ["fp64", "i8s5"])
def test_batch_prepare_device_errors(ctx, rimp2, mode):
    """The device-side input checks name the failing fragment: a metric that
    is not positive definite, a non-finite coefficient, a denominator >= 0;
    the context stays usable.  In the exact-slice modes the (P|ia) check is
    part of the int8 transform's slicing."""
    good = rimp2.gen_fragment(3, 40, 10, 3)
    P, V, eo, ev = good
    Vbad = V.copy(order="F")
    Vbad[5, 5] = -1.0
    Pbad = P.copy(order="F")
    Pbad[1, 2, 0] = np.nan
    eon = eo.copy()
    eon[1] = np.nan
    cases = [((P, Vbad, eo, ev), rimp2.ERR_NOT_SPD), ((Pbad, V, eo, ev), rimp2.ERR_NUMERIC),
             ((P, V, eo, ev - 10.0), rimp2.ERR_NUMERIC), ((P, V, eon, ev), rimp2.ERR_NUMERIC)]
    for bad, code in cases:
        say("batch prepare device error", mode, code)
        descs, ptrs = _to_device_descs(ctx, [good, bad])
        try:
            with pytest.raises(rimp2.Rimp2Error) as ei:
                ctx.batch_prepare_device(descs, mode)
        finally:
            for p in ptrs:
                ctx.free(p)
        assert ei.value.code == code
        assert "fragment 1" in str(ei.value)
    ef, _ = ctx.batch([good], mode)
    assert ef[0] < 0


def test_batch_autotuner_times_every_candidate(oracle, rimp2):
    """A fresh context times the three fp32 candidates (every fragment 32- or
    64-wide, and the per-fragment mix) on the batch and keeps the fastest; a
    second prepare of the same size class uses the cached choice (no
    measurements).  The chosen layout gives the oracle's fragment energies
    within the fp32 tolerance (the 32-wide tile runs 16x16x4 MFMAs)."""
    frags = make_batch(rimp2, 57, 5, scale=4)
    with rimp2.Context(0) as c:
        c.batch_prepare(frags, "fp32")
        tuning = c.batch_tuning()
        say("autotune fp32", tuning, "chosen", c.batch_tile())
        assert sorted(t for t, _ in tuning) == [-1, 32, 64]
        assert all(ms > 0 for _, ms in tuning)
        assert c.batch_tile() == min(tuning, key=lambda x: x[1])[0]
        c.batch_run()
        ef, _ = c.batch_result(want_pairs=False)
        c.batch_prepare(frags, "fp32")
        assert c.batch_tuning() == []
    for k, (Eref, _) in enumerate(oracle_energies(oracle, frags)):
        assert abs(ef[k] - Eref) <= TOL_REL["fp32"] * abs(Eref), (k, ef[k], Eref)


# The mixed candidate's rule: the 32-wide tile where the 64-wide one executes
# at least 1.5x its MFMA work.  nvirt 20 and 70 (32 / 96 wide against 64 / 128:
# 4x and 1.78x) take it; 150 (160 against 192: 1.44x) and 40, 100, 200 (the
# same width either way) keep the 64-wide tile.
MIXED_SHAPES = [(5, 20, 60), (3, 40, 100), (4, 70, 130), (2, 100, 260), (3, 150, 170), (2, 200, 140)]


def mixed_tile(nvirt):
    w32, w64 = -(-nvirt // 32) * 32, -(-nvirt // 64) * 64
    return 32 if w64 * w64 >= 1.5 * w32 * w32 else 64


@pytest.mark.parametrize("mode", ["fp64", "fp32"])
def test_batch_mixed_tiles_per_fragment(oracle, rimp2, mode):
    """Each tile candidate fixed by rimp2_batch_set_tile, on fragments of both
    tile classes: every fragment's energy and pair matrix match the oracle in
    each layout, and under the mixed candidate each fragment reproduces, bit
    for bit, the one-tile layout of its own class (same kernel path, same
    unit order, same reduction)."""
    frags = [rimp2.gen_fragment(300 + k, q, v, o) for k, (o, v, q) in enumerate(MIXED_SHAPES)]
    assert [mixed_tile(v) for _, v, _ in MIXED_SHAPES] == [32, 64, 32, 64, 64, 64]
    ref = oracle_energies(oracle, frags)
    res = {}
    with rimp2.Context(0) as c:
        for cand in (32, 64, -1):
            c.batch_set_tile(cand)
            c.batch_prepare(frags, mode)
            assert c.batch_tile() == cand and c.batch_tuning() == []
            c.batch_run()
            res[cand] = c.batch_result()
            say("batch", mode, "candidate", cand, res[cand][0])
        c.batch_set_tile(0)
    for cand, (ef, pp) in res.items():
        off = 0
        for k, (P, V, eo, ev) in enumerate(frags):
            Eref, epsref = ref[k]
            n = eo.size
            eps = pp[off:off + n * n].reshape(n, n, order="F")
            off += n * n
            assert abs(ef[k] - Eref) <= TOL_REL[mode] * abs(Eref), (cand, k, ef[k], Eref)
            if mode == "fp64":
                np.testing.assert_allclose(eps, epsref, rtol=1e-10, atol=1e-15)
    off = 0
    for k, (o, v, q) in enumerate(MIXED_SHAPES):
        same = res[mixed_tile(v)]
        assert res[-1][0][k] == same[0][k], (k, v)
        np.testing.assert_array_equal(res[-1][1][off:off + o * o], same[1][off:off + o * o])
        off += o * o


def test_batch_set_tile_rejects_unavailable_candidates(rimp2):
    frags = [rimp2.gen_fragment(400, 64, 20, 3)]
    with rimp2.Context(0) as c:
        with pytest.raises(rimp2.Rimp2Error) as ei:
            c.batch_set_tile(7)
        assert ei.value.code == rimp2.ERR_ARG
        c.batch_set_tile(-1)
        with pytest.raises(rimp2.Rimp2Error) as ei:
            c.batch_prepare(frags, "i8s5")  # the exact-slice modes have the 64-wide tile only
        assert ei.value.code == rimp2.ERR_ARG and "not available" in str(ei.value)
        c.batch_set_tile(0)
        c.batch_prepare(frags, "i8s5")
        assert c.batch_tile() == 64


class CountingCtx:
    """Passes batch calls through to a Context and counts the fragments."""

    def __init__(self, c):
        self.c, self.nfrag = c, 0

    def batch(self, frags, mode, want_pairs=True):
        self.nfrag += len(frags)
        return self.c.batch(frags, mode, want_pairs)


@pytest.mark.parametrize("mode", ["fp64", "i8s5"])
def test_batch_results_file_resumes_to_the_straight_result(rimp2, tmp_path, mode):
    """A run over 12 fragments whose first 7 an earlier (interrupted) run
    recorded: only the other 5 are computed, and every fragment's energy and
    pair matrix equals a straight batch of all 12 bit for bit (one tile
    fixed, so a fragment's result does not depend on the batch around it)."""
    from rimp2_gpu_amd import results
    frags = make_batch(rimp2, 31, 12, scale=4)
    keys = [f"seed{31 + k}" for k in range(12)]
    say("results file", mode, [f[0].shape for f in frags])
    p = str(tmp_path / "c3.jsonl")
    with rimp2.Context(0) as c:
        c.batch_set_tile(64)
        ef, pp = c.batch(frags, mode)
        results.run_batch(c, frags[:7], keys[:7], p, mode=mode, chunk=4)
        cc = CountingCtx(c)
        E, P = results.run_batch(cc, frags, keys, p, mode=mode, chunk=4)
    assert cc.nfrag == 5
    np.testing.assert_array_equal(E, ef)
    off = 0
    for k, (Pk, V, eo, ev) in enumerate(frags):
        n = eo.size
        np.testing.assert_array_equal(P[k], pp[off:off + n * n].reshape(n, n, order="F"))
        off += n * n
