"""GPU parity of the DEC local-basis 4-index arrays (Major Steps 5-8,
RIMP2.F90:67-170) and of the general FP64 GEMM (cublasDgemm_v2 semantics,
gpu_interfaces.F90:24-32), through the C ABI.

Tolerances: the reference's own check, |test - got| <= 1e-10 absolute for
every element (RIMP2.F90:262 and its three siblings), and in addition
<= 1e-12 of the largest oracle entry (summation order is the only
difference: FP64 MFMA against the oracle's scalar loops).  GEMM: 1e-13 of the
largest |A||B| product sum.
"""
import numpy as np
import pytest

from test_oracle import FIXTURES, first_order_inputs, ref_fixture

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def orth(rng, n, k):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return np.asfortranarray(q[:, :k])


def random_inputs(seed, nvirt, nocc, noccEOS, nvirtEOS, nba):
    """Physically shaped: Calpha ~ 0.1 N(0,1), occupied below virtual
    energies, orthonormal columns in every U (orbital rotations)."""
    rng = np.random.default_rng(seed)
    C = np.asfortranarray(0.1 * rng.standard_normal((nba, nvirt, nocc)))
    eo = rng.uniform(-2.0, -0.3, nocc)
    ev = rng.uniform(0.1, 3.0, nvirt)
    return C, eo, ev, orth(rng, nocc, noccEOS), orth(rng, nvirt, nvirt), orth(rng, nocc, nocc), \
        orth(rng, nvirt, nvirtEOS)


def check_arrays(got, exp):
    names = ("toccEOS", "tvirtEOS", "goccEOS", "gvirtEOS")
    for name, g, e in zip(names, got, exp):
        assert g.shape == e.shape, (name, g.shape, e.shape)
        err = float(np.max(np.abs(g - e)))
        scale = float(np.max(np.abs(e)))
        say(f"  {name}: max|d| = {err:.3e} (max|ref| {scale:.3e})")
        assert err <= 1e-10, (name, err)
        assert err <= 1e-12 * scale, (name, err, scale)


# the reference's own fixtures (generate.F90:41-46, 58-63), fed through the
# GPU: every element equals the closed form within the reference's 1e-10
@pytest.mark.parametrize("dims,expect", FIXTURES)
def test_eos_reference_fixture(ctx, dims, expect):
    nvirt, nocc, noccEOS, nvirtEOS, nba = dims
    say(f"eos reference fixture {dims}")
    ctx.eos_set_chunk(0)
    outs = ctx.eos_arrays(*ref_fixture(nvirt, nocc, noccEOS, nvirtEOS, nba))
    for arr, val in zip(outs, expect):
        assert np.all(np.isfinite(arr))
        assert float(np.max(np.abs(arr - val))) <= 1e-10


# shapes: the reference's two fixture sizes, a one-orbital case, ragged
# tiles (65 = 64 + 1 virtuals), full EOS spaces (noccEOS = nocc,
# nvirtEOS = nvirt), more than 8 virtual blocks; chunk = occupied orbitals
# per amplitude slab (0: automatic, the whole-P form; > 0: per-chunk P)
SHAPES = [
    ((74, 10, 4, 15, 273), 0),
    ((74, 10, 4, 15, 273), 1),
    ((74, 10, 4, 15, 273), 3),
    ((104, 20, 4, 15, 373), 0),
    ((1, 1, 1, 1, 1), 0),
    ((65, 3, 2, 64, 17), 2),
    ((33, 7, 7, 33, 130), 0),
    # 520 virtuals = 9 blocks of 64: the pair kernel's super-tile unit order
    ((520, 6, 3, 10, 64), 0),
    ((520, 6, 3, 10, 64), 4),
]


@pytest.mark.parametrize("dims,chunk", SHAPES)
def test_eos_random_vs_oracle(ctx, oracle, dims, chunk):
    say(f"eos random {dims} chunk={chunk}")
    args = random_inputs(17 + sum(dims), *dims)
    ctx.eos_set_chunk(chunk)
    try:
        got = ctx.eos_arrays(*args)
    finally:
        ctx.eos_set_chunk(0)
    check_arrays(got, oracle.steps5to8(*args))


# Exact-slice amplitude slab (rimp2_eos_set_mode i8s5): V to ~1e-12 relative;
# on arrays of O(1) values the reference's 1e-10 absolute check still holds.
@pytest.mark.parametrize("dims", [(74, 10, 4, 15, 273), (104, 20, 4, 15, 373), (520, 6, 3, 10, 64)])
def test_eos_i8s5_slab_vs_oracle(ctx, oracle, dims):
    say(f"eos i8s5 {dims}")
    args = random_inputs(23 + sum(dims), *dims)
    ctx.eos_set_mode("i8s5")
    try:
        got = ctx.eos_arrays(*args)
    finally:
        ctx.eos_set_mode("fp64")
    for name, g, e in zip(("toccEOS", "tvirtEOS", "goccEOS", "gvirtEOS"), got, oracle.steps5to8(*args)):
        err, scale = float(np.max(np.abs(g - e))), float(np.max(np.abs(e)))
        say(f"  {name}: max|d| = {err:.3e} (max|ref| {scale:.3e})")
        assert err <= 1e-10, (name, err)
        assert err <= 1e-9 * scale, (name, err, scale)


def test_eos_mode_rejects_others(ctx, rimp2):
    for m in ("fp32", "bf16x3", "i8s4"):
        with pytest.raises(rimp2.Rimp2Error) as e:
            ctx.eos_set_mode(m)
        assert e.value.code == rimp2.ERR_ARG


def test_eos_device_entry_matches_host_entry(ctx):
    dims = (74, 10, 4, 15, 273)
    nvirt, nocc, noccEOS, nvirtEOS, nba = dims
    args = random_inputs(5, *dims)
    say("eos device entry")
    host = ctx.eos_arrays(*args)
    d_in = [ctx.to_device(x) for x in args]
    shapes = [(nvirt, noccEOS, nvirt, noccEOS), (nvirtEOS, nocc, nvirtEOS, nocc)] * 2
    d_out = [ctx.alloc(int(np.prod(s)) * 8) for s in shapes]
    try:
        ctx.eos_arrays_device(nvirt, nocc, noccEOS, nvirtEOS, nba, d_in, d_out)
        ctx.sync()
        dev = [ctx.from_device(p, s) for p, s in zip(d_out, shapes)]
    finally:
        for p in d_in + d_out:
            ctx.free(p)
    for h, d in zip(host, dev):
        np.testing.assert_array_equal(h, d)


def test_eos_rejects_bad_dimensions(ctx, rimp2):
    args = random_inputs(1, 8, 3, 2, 4, 5)
    C, eo, ev, UoE, Uv, Uo, UvE = args
    with pytest.raises(rimp2.Rimp2Error) as e:
        ctx.eos_arrays(C, eo, ev, np.asfortranarray(np.ones((3, 4))), Uv, Uo, UvE)  # noccEOS 4 > nocc 3
    assert e.value.code == rimp2.ERR_ARG
    with pytest.raises(rimp2.Rimp2Error) as e:
        ctx.eos_set_chunk(-1)
    assert e.value.code == rimp2.ERR_ARG


# ---------------------------------------------------------------------------
# first-order djik / blad (RIMP2.F90:172-251) against the oracle, which keeps
# the reference's K = ldb = nocc CalphaVV product (RIMP2.F90:231-235)
# ---------------------------------------------------------------------------
# (6, 6, ...): nocc = nvirt, where the K = ldb = nocc product reads UvirtT up to its last element
FO_SHAPES = [(74, 10, 4, 15, 273), (104, 20, 4, 15, 373), (1, 1, 1, 1, 1), (65, 33, 2, 64, 17), (6, 6, 2, 3, 7)]


@pytest.mark.parametrize("dims", FO_SHAPES)
def test_first_order_vs_oracle(ctx, oracle, dims):
    nvirt, nocc, noccEOS, nvirtEOS, nba = dims
    say(f"first order {dims}")
    rng = np.random.default_rng(41 + sum(dims))
    args = [0.1 * a if a.ndim == 3 else a for a in first_order_inputs(rng, nba, nvirt, nocc, noccEOS, nvirtEOS)]
    got = ctx.first_order(*args)
    exp = oracle.first_order(*args)
    for name, g, e in zip(("djik", "blad"), got, exp):
        err, scale = float(np.max(np.abs(g - e))), float(np.max(np.abs(e)))
        say(f"  {name}: max|d| = {err:.3e} (max|ref| {scale:.3e})")
        assert err <= 1e-12 * scale + 1e-14


def test_first_order_rejects_out_of_bounds_shapes(ctx, rimp2):
    rng = np.random.default_rng(2)
    C, Co, Cvv, UoE, Uv, Uo, UvE = first_order_inputs(rng, 5, 6, 3, 2, 4)
    with pytest.raises(rimp2.Rimp2Error) as e:  # nocctot != nocc
        ctx.first_order(C, np.asfortranarray(np.ones((5, 3, 4))), Cvv, UoE, Uv, Uo, UvE)
    assert e.value.code == rimp2.ERR_ARG
    C, Co, Cvv, UoE, Uv, Uo, UvE = first_order_inputs(rng, 5, 3, 4, 2, 2)  # nocc 4 > nvirt 3
    with pytest.raises(rimp2.Rimp2Error) as e:
        ctx.first_order(C, Co, Cvv, UoE, Uv, Uo, UvE)
    assert e.value.code == rimp2.ERR_ARG


# ---------------------------------------------------------------------------
# rimp2_dgemm: all four op combinations, ragged sizes, odd and even leading
# dimensions (scalar and 16-byte load paths), alpha / beta, k = 0, and beta = 0
# not reading C (NaN there stays unseen)
# ---------------------------------------------------------------------------
GEMM_CASES = [(1, 1, 1, 0), (70, 33, 129, 3), (128, 64, 16, 0), (65, 130, 47, 1), (5, 200, 0, 2), (200, 5, 300, 4)]


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
@pytest.mark.parametrize("m,n,k,pad", GEMM_CASES)
def test_dgemm_matches_numpy(ctx, ta, tb, m, n, k, pad):
    rng = np.random.default_rng(m * 7 + n * 3 + k + ta + 2 * tb)
    ra, ca = (k, m) if ta else (m, k)
    rb, cb = (n, k) if tb else (k, n)
    lda, ldb, ldc = max(1, ra) + pad, max(1, rb) + pad, max(1, m) + pad
    A = np.asfortranarray(rng.standard_normal((lda, max(ca, 1))))
    B = np.asfortranarray(rng.standard_normal((ldb, max(cb, 1))))
    C0 = np.asfortranarray(rng.standard_normal((ldc, n)))
    opA = A[:ra, :ca].T if ta else A[:ra, :ca]
    opB = B[:rb, :cb].T if tb else B[:rb, :cb]
    say(f"dgemm ta={ta} tb={tb} m={m} n={n} k={k} pad={pad}")
    dA, dB = ctx.to_device(A), ctx.to_device(B)
    try:
        for alpha, beta in ((1.0, 0.0), (1.5, -0.5)):
            C = C0.copy(order="F")
            if beta == 0.0:
                C[:m, :] = np.nan
            dC = ctx.to_device(C)
            try:
                ctx.dgemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc)
                ctx.sync()
                got = ctx.from_device(dC, C.shape)
            finally:
                ctx.free(dC)
            exp = C0.copy(order="F")
            exp[:m, :] = alpha * (opA @ opB) + (beta * C0[:m, :] if beta != 0.0 else 0.0)
            scale = max(1.0, float(np.max(np.abs(opA) @ np.abs(opB))) if k else 1.0)
            assert np.all(np.isfinite(got[:m, :]))
            assert float(np.max(np.abs(got[:m, :] - exp[:m, :]))) <= 1e-13 * scale * max(1.0, abs(alpha))
            # rows below m (the padding of ldc) are untouched
            np.testing.assert_array_equal(got[m:, :], C0[m:, :])
    finally:
        ctx.free(dA)
        ctx.free(dB)


def test_dgemm_alpha_zero_reads_neither_operand(ctx):
    """BLAS semantics: alpha = 0 gives beta C without touching A or B (NaN there)."""
    m, n, k = 33, 17, 9
    A = np.full((m, k), np.nan, order="F")
    B = np.full((k, n), np.nan, order="F")
    C0 = np.asfortranarray(np.random.default_rng(3).standard_normal((m, n)))
    dA, dB, dC = ctx.to_device(A), ctx.to_device(B), ctx.to_device(C0)
    try:
        ctx.dgemm(0, 0, m, n, k, 0.0, dA, m, dB, k, 0.5, dC, m)
        ctx.sync()
        got = ctx.from_device(dC, (m, n))
    finally:
        for p in (dA, dB, dC):
            ctx.free(p)
    np.testing.assert_array_equal(got, 0.5 * C0)


def test_dgemm_rejects_bad_arguments(ctx, rimp2):
    p = ctx.alloc(64 * 8)
    try:
        for args in ((5, 0, 4, 4, 4, 4, 4, 4), (0, 0, 4, 4, 4, 3, 4, 4), (0, 0, 4, 4, 4, 4, 3, 4),
                     (0, 0, 4, 4, 4, 4, 4, 3), (0, 0, -1, 4, 4, 4, 4, 4), (1, 0, 4, 4, 5, 4, 4, 4)):
            ta, tb, m, n, k, lda, ldb, ldc = args
            with pytest.raises(rimp2.Rimp2Error) as e:
                ctx.dgemm(ta, tb, m, n, k, 1.0, p, lda, p, ldb, 0.0, p, ldc)
            assert e.value.code == rimp2.ERR_ARG, args
    finally:
        ctx.free(p)


@pytest.mark.parametrize("mode", ["fp64", "i8s5"])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_eos_rejects_non_finite_inputs(ctx, rimp2, mode, bad):
    """A NaN or Inf in any input fails with RIMP2_ERR_NUMERIC before an output
    is written (the exact-slice packer would otherwise turn it into finite
    slices), in the blocking and the device forms; the context stays usable."""
    dims = (40, 6, 3, 5, 50)
    nvirt, nocc, noccEOS, nvirtEOS, nba = dims
    say("eos non-finite", mode, bad)
    args = list(random_inputs(81, *dims))
    ctx.eos_set_mode(mode)
    try:
        for which, idx in ((0, (7, 3, 1)), (1, (2,)), (4, (5, 9))):
            a = [x.copy(order="F") for x in args]
            a[which][idx] = bad
            with pytest.raises(rimp2.Rimp2Error) as ei:
                ctx.eos_arrays(*a)
            assert ei.value.code == rimp2.ERR_NUMERIC, which
        # device form: reported by the next sync
        a = [x.copy(order="F") for x in args]
        a[0][0, 0, 0] = bad
        d_in = [ctx.to_device(x) for x in a]
        shapes = [(nvirt, noccEOS, nvirt, noccEOS), (nvirtEOS, nocc, nvirtEOS, nocc)] * 2
        d_out = [ctx.alloc(int(np.prod(s)) * 8) for s in shapes]
        ctx.eos_arrays_device(*dims, d_in, d_out)
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.sync()
        assert ei.value.code == rimp2.ERR_NUMERIC
        for p in d_in + d_out:
            ctx.free(p)
        # first order: NaN in CalphaVV
        fo = list(first_order_inputs(np.random.default_rng(82), nba, nvirt, nocc, noccEOS, nvirtEOS))
        fo[2] = fo[2].copy(order="F")
        fo[2][1, 2, 3] = bad
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.first_order(*fo)
        assert ei.value.code == rimp2.ERR_NUMERIC
        outs = ctx.eos_arrays(*args)
        assert all(np.all(np.isfinite(o)) for o in outs)
    finally:
        ctx.eos_set_mode("fp64")
