"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the
committed golden fixtures.

Tolerances (stated per mode; E_ref is the oracle's FP64 energy):
  fp64  : |dE| <= 1e-11 |E_ref|, pair energies rel 1e-10 (summation order only)
  i8s5  : |dE| <  1 microhartree (1e-6 Eh) -- the exact-slice target
  i8s4  : |dE| <  1 microhartree at these sizes
  fp32  : |dE| <= 2e-7 |E_ref|   (FP32 inputs and accumulation)
  bf16x3: |dE| <= 2e-5 |E_ref|   (2-way BF16 split, FP32 accumulation)
Metric: Cholesky / inverse / B rel 1e-10 of the largest entry.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")

TOL_REL = {"fp64": 1e-11, "fp32": 2e-7, "bf16x3": 2e-5}
TOL_ABS = {"i8s5": 1e-6, "i8s4": 1e-6}


def say(*a):
    print(*a, flush=True)


def check_E(mode, E, Eref):
    if mode in TOL_ABS:
        assert abs(E - Eref) < TOL_ABS[mode], (mode, E, Eref, E - Eref)
    else:
        assert abs(E - Eref) <= TOL_REL[mode] * abs(Eref), (mode, E, Eref, (E - Eref) / Eref)


# 64-wide Cholesky steps with ragged last blocks of 2, 8, 1 and 60 rows
# (130, 200, 257, 700), and several levels of the inverse's doubling with a
# ragged last group
@pytest.mark.parametrize("naux", [1, 5, 64, 100, 128, 130, 200, 257, 512, 700])
def test_metric_factor_and_inverse(ctx, oracle, rimp2, naux):
    say("metric naux", naux)
    _, V, _, _ = rimp2.gen_fragment(100 + naux, naux, 2, 1)
    ctx.metric(V)
    L = ctx.get_cholesky(naux)
    Lref = oracle.cholesky(V)
    assert np.max(np.abs(L - Lref)) <= 1e-10 * np.abs(Lref).max()
    Li = ctx.get_metric_inverse(naux)
    np.testing.assert_allclose(Li @ Lref, np.eye(naux), rtol=0, atol=1e-10)
    assert np.all(np.triu(Li, 1) == 0) and np.all(np.triu(L, 1) == 0)


# B = L^{-1} X relative to max |B|: FP64 MFMA (fp64 and the emulated modes),
# six int8 slices per operand with an FP64 epilogue (the exact-slice modes;
# about 1e-13 at these sizes)
TRANSFORM_TOL = {"fp64": 1e-10, "i8s5": 1e-11}


@pytest.mark.parametrize("mode", ["fp64", "i8s5"])
@pytest.mark.parametrize("shape", [(64, 16, 3), (100, 17, 2), (130, 70, 3), (300, 130, 2)])
def test_transform_matches_oracle(ctx, oracle, rimp2, shape, mode):
    naux, nvirt, nocc = shape
    say("transform", shape, mode)
    P, V, eo, ev = rimp2.gen_fragment(7, naux, nvirt, nocc)
    ctx.metric(V)
    ctx.transform(P, mode)
    B = ctx.get_B(naux, nvirt, nocc)
    Bref = oracle.metric_transform(V, P)
    err = np.max(np.abs(B - Bref)) / np.abs(Bref).max()
    say("relative error", err)
    assert err <= TRANSFORM_TOL[mode]


GOLDEN = sorted(f for f in os.listdir(GOLD) if f.startswith("frag_"))


@pytest.mark.parametrize("fname", GOLDEN)
def test_fp64_against_golden(ctx, rimp2, fname):
    g = np.load(os.path.join(GOLD, fname))
    say("golden", fname)
    P, V, eo, ev = rimp2.gen_fragment(int(g["seed"]), int(g["naux"]), int(g["nvirt"]), int(g["nocc"]))
    E, eps = ctx.energy(P, V, eo, ev, "fp64")
    check_E("fp64", E, float(g["E"]))
    np.testing.assert_allclose(eps, g["eps_pair"], rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("mode", ["fp64", "i8s5", "i8s4", "fp32", "bf16x3"])
@pytest.mark.parametrize("fname", ["frag_4_512_128_16.npz", "frag_3_128_70_4.npz"])
def test_modes_against_golden(ctx, rimp2, mode, fname):
    g = np.load(os.path.join(GOLD, fname))
    say("mode", mode, fname)
    P, V, eo, ev = rimp2.gen_fragment(int(g["seed"]), int(g["naux"]), int(g["nvirt"]), int(g["nocc"]))
    E, eps = ctx.energy(P, V, eo, ev, mode)
    check_E(mode, E, float(g["E"]))
    np.testing.assert_array_equal(eps, eps.T)
    assert np.all(eps < 0)


# The last two run more than one super-tile column of 8 x 8 virtual blocks
# (700: 11 blocks of 64, the last column 3 wide; 1090: 18 blocks, last 2 wide).
EDGE = [(16, 1, 1), (8, 3, 1), (20, 33, 2), (37, 65, 3), (64, 130, 5), (5, 2, 4), (48, 100, 1), (96, 700, 3),
        (40, 1090, 2)]


@pytest.mark.parametrize("mode", ["fp64", "i8s5", "fp32", "bf16x3"])
@pytest.mark.parametrize("shape", EDGE)
def test_ragged_shapes(ctx, oracle, rimp2, mode, shape):
    naux, nvirt, nocc = shape
    say("edge", mode, shape)
    P, V, eo, ev = rimp2.gen_fragment(1000 + naux + nvirt + nocc, naux, nvirt, nocc)
    B = oracle.metric_transform(V, P)
    Eref, epsref = oracle.pair_energies(B, eo, ev)
    E, eps = ctx.energy(P, V, eo, ev, mode)
    check_E(mode, E, Eref)
    if mode == "fp64":
        np.testing.assert_allclose(eps, epsref, rtol=1e-10, atol=1e-15)


def test_device_resident_path_matches_host(ctx, rimp2):
    naux, nvirt, nocc = 256, 96, 6
    say("device path", naux, nvirt, nocc)
    P, V, eo, ev = rimp2.gen_fragment(77, naux, nvirt, nocc)
    for mode in ["fp64", "i8s5"]:
        E_host, eps_host = ctx.energy(P, V, eo, ev, mode)
        d = [ctx.to_device(x) for x in (P, V, eo, ev)]
        dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
        ctx.energy_device(naux, nvirt, nocc, *d, mode, dpe, det)
        ctx.sync()
        E_dev = float(ctx.from_device(det, (1,))[0])
        eps_dev = ctx.from_device(dpe, (nocc, nocc))
        for p in d + [dpe, det]:
            ctx.free(p)
        assert E_dev == E_host
        np.testing.assert_array_equal(eps_dev, eps_host)


def test_staged_host_copies_are_exact(ctx, rimp2):
    """Host arrays of 1 MiB and more reach the device through two pinned chunk
    buffers: byte-exact at chunk-ragged sizes, twice in a row (buffer reuse),
    and an energy from host arrays above the chunk size equals the one from
    device arrays."""
    say("staged host copies")
    rng = np.random.default_rng(11)
    for nbytes in (40 * 2**20 + 13, 9 * 2**20 + 1, 2**24, 3 * 2**19 + 7, 2**20, 4096):
        for _ in range(2):
            x = rng.integers(0, 256, nbytes, dtype=np.uint8)
            d = ctx.to_device(x)
            back = ctx.from_device(d, (nbytes,), np.uint8)
            ctx.free(d)
            np.testing.assert_array_equal(back, x)
    naux, nvirt, nocc = 512, 128, 40  # (P|ia) 21 MB: two chunks, the second partial
    P, V, eo, ev = rimp2.gen_fragment(21, naux, nvirt, nocc)
    E_host, eps_host = ctx.energy(P, V, eo, ev, "i8s5")
    d = [ctx.to_device(x) for x in (P, V, eo, ev)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    ctx.energy_device(naux, nvirt, nocc, *d, "i8s5", dpe, det)
    ctx.sync()
    E_dev = float(ctx.from_device(det, (1,))[0])
    eps_dev = ctx.from_device(dpe, (nocc, nocc))
    for p in d + [dpe, det]:
        ctx.free(p)
    assert E_dev == E_host
    np.testing.assert_array_equal(eps_dev, eps_host)
    E_fit, _ = ctx.energy_fitted(ctx.get_B(naux, nvirt, nocc), eo, ev, "i8s5")
    assert E_fit == E_host


def test_host_energy_pipeline_by_orbital_chunks(ctx, rimp2):
    """rimp2_energy from host arrays copies (P|ia) in chunks of occupied
    orbitals under the Cholesky and transforms each chunk as it lands: the
    energy equals the device-resident path's bit for bit, a NaN/Inf in any
    chunk or in V is rejected before results are copied out, and the context
    keeps working afterwards."""
    naux, nvirt, nocc = 1024, 512, 20  # 4 MiB per orbital: chunks of 8, 8 and 4 orbitals
    say("host energy pipeline", naux, nvirt, nocc)
    P, V, eo, ev = rimp2.gen_fragment(31, naux, nvirt, nocc)
    d = [ctx.to_device(x) for x in (P, V, eo, ev)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    for mode in ["i8s5", "fp64"]:
        E_host, eps_host = ctx.energy(P, V, eo, ev, mode)
        ctx.energy_device(naux, nvirt, nocc, *d, mode, dpe, det)
        ctx.sync()
        assert float(ctx.from_device(det, (1,))[0]) == E_host
        np.testing.assert_array_equal(ctx.from_device(dpe, (nocc, nocc)), eps_host)
        E_split, _ = ctx.energy_fitted(ctx.get_B(naux, nvirt, nocc), eo, ev, mode)
        assert E_split == E_host
    for p in d + [dpe, det]:
        ctx.free(p)
    for occ, val in [(0, np.nan), (8, np.inf), (19, -np.inf)]:
        Pn = P.copy(order="F")
        Pn[naux - 1, nvirt - 1, occ] = val
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.energy(Pn, V, eo, ev, "i8s5")
        assert ei.value.code == rimp2.ERR_NUMERIC
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.get_B(naux, nvirt, nocc)  # a rejected call leaves no coefficients
        assert ei.value.code == rimp2.ERR_STATE
    Vn = V.copy(order="F")
    Vn[3, 5] = np.nan
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.energy(P, Vn, eo, ev, "i8s5")
    assert ei.value.code == rimp2.ERR_NUMERIC
    Vb = V.copy(order="F")
    Vb[7, 7] = -10.0
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.energy(P, Vb, eo, ev, "i8s5")
    assert ei.value.code == rimp2.ERR_NOT_SPD
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.transform(P, "i8s5")  # no metric factored after the failures
    assert ei.value.code == rimp2.ERR_STATE
    E_again, _ = ctx.energy(P, V, eo, ev, "i8s5")
    ctx.metric(V)
    ctx.transform(P, "i8s5")
    E_steps, _ = ctx.pair_energies(eo, ev, "i8s5")
    assert E_again == E_steps
    E_ref, _ = ctx.energy(P, V, eo, ev, "fp64")
    assert abs(E_again - E_ref) < 1e-6


@pytest.mark.parametrize("mode", ["fp64", "fp32", "bf16x3", "i8s4", "i8s5"])
def test_host_energy_chunks_at_ragged_sizes(ctx, rimp2, mode):
    """naux and nvirt off every tile multiple, 5.6 MB per occupied orbital:
    (P|ia) arrives in chunks of 5, 5 and 2 orbitals, each packed into padded
    slabs; the host path equals the device path bit for bit in every mode."""
    naux, nvirt, nocc = 1001, 700, 12
    say("ragged host chunks", mode)
    P, V, eo, ev = rimp2.gen_fragment(41, naux, nvirt, nocc)
    E_host, eps_host = ctx.energy(P, V, eo, ev, mode)
    d = [ctx.to_device(x) for x in (P, V, eo, ev)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    ctx.energy_device(naux, nvirt, nocc, *d, mode, dpe, det)
    ctx.sync()
    E_dev = float(ctx.from_device(det, (1,))[0])
    eps_dev = ctx.from_device(dpe, (nocc, nocc))
    for p in d + [dpe, det]:
        ctx.free(p)
    assert E_dev == E_host
    np.testing.assert_array_equal(eps_dev, eps_host)


def test_deterministic(ctx, rimp2):
    P, V, eo, ev = rimp2.gen_fragment(5, 192, 80, 7)
    say("determinism")
    for mode in ["fp64", "i8s5", "fp32"]:
        E1, e1 = ctx.energy(P, V, eo, ev, mode)
        E2, e2 = ctx.energy(P, V, eo, ev, mode)
        assert E1 == E2
        np.testing.assert_array_equal(e1, e2)


def test_error_paths(ctx, rimp2):
    say("error paths")
    P, V, eo, ev = rimp2.gen_fragment(9, 32, 8, 2)
    bad = V.copy(order="F")
    bad[5, 5] = -10.0
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.metric(bad)
    assert ei.value.code == rimp2.ERR_NOT_SPD
    ctx.metric(V)
    ctx.transform(P, "fp64")
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.pair_energies(eo, ev, "i8s5")  # transform holds fp64
    assert ei.value.code == rimp2.ERR_STATE
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.pair_energies(eo, ev - 5.0, "fp64")  # denominators >= 0
    assert ei.value.code == rimp2.ERR_NUMERIC
    Pn = P.copy(order="F")
    Pn[3, 2, 1] = np.nan
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.transform(Pn, "fp64")
    assert ei.value.code == rimp2.ERR_NUMERIC
    with pytest.raises(ValueError):
        ctx.transform(P, "fp16")
    # context is still usable after errors
    ctx.metric(V)
    E, _ = ctx.energy(P, V, eo, ev, "fp64")
    assert E < 0


def test_device_path_flags_bad_input(ctx, rimp2):
    say("device error flags")
    naux, nvirt, nocc = 32, 8, 2
    P, V, eo, ev = rimp2.gen_fragment(9, naux, nvirt, nocc)
    d = [ctx.to_device(x) for x in (P, V, eo, ev - 10.0)]
    det = ctx.alloc(8)
    ctx.energy_device(naux, nvirt, nocc, *d, "fp64", None, det)
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.sync()
    assert ei.value.code == rimp2.ERR_NUMERIC
    for p in d + [det]:
        ctx.free(p)
    ctx.sync()  # status was cleared


def test_int8_transform_at_the_largest_exact_slice_naux(ctx, rimp2):
    """naux = 16384, the exact-slice modes' limit (the int8 transform's int32
    diagonals hold at most 6 x 127^2 x naux < 2^31): its K loops run up to 512
    stages.  B against the FP64 transform of the same device-generated inputs,
    and E within 1 uEh of fp64; one more auxiliary function is refused."""
    naux, nvirt, nocc = 16384, 32, 2
    say("int8 transform at naux", naux)
    b = [ctx.alloc(naux * nvirt * nocc * 8), ctx.alloc(naux * naux * 8), ctx.alloc(nocc * 8), ctx.alloc(nvirt * 8)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    E, B = {}, {}
    try:
        ctx.gen_device(41, naux, nvirt, nocc, *b)
        for mode in ("fp64", "i8s5"):
            ctx.energy_device(naux, nvirt, nocc, *b, mode, dpe, det)
            ctx.sync()
            E[mode] = float(ctx.from_device(det, (1,))[0])
            B[mode] = ctx.get_B(naux, nvirt, nocc)
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.energy_device(naux + 1, nvirt, nocc, *b, "i8s5", dpe, det)
        assert ei.value.code == rimp2.ERR_ARG
    finally:
        for p in b + [dpe, det]:
            ctx.free(p)
    err = np.max(np.abs(B["i8s5"] - B["fp64"])) / np.abs(B["fp64"]).max()
    say("B relative error", err, "dE uEh", (E["i8s5"] - E["fp64"]) * 1e6)
    assert err <= TRANSFORM_TOL["i8s5"]
    assert abs(E["i8s5"] - E["fp64"]) < 1e-6


@pytest.mark.parametrize("shape", [(33, 5, 3), (256, 129, 33)])
def test_non_finite_pia_is_found_at_any_position(ctx, rimp2, shape):
    """The device check of (P|ia) reads 16-byte pairs on a capped grid: a NaN
    or Inf is found as the first, a middle and the last element (for odd
    sizes the unpaired tail), and in a pointer shifted by one element (the
    unaligned path); the clean array passes."""
    naux, nvirt, nocc = shape
    n = naux * nvirt * nocc
    say("non-finite positions", shape, n)
    P, V, eo, ev = rimp2.gen_fragment(13, naux, nvirt, nocc)
    dV, deo, dev = ctx.to_device(V), ctx.to_device(eo), ctx.to_device(ev)
    det = ctx.alloc(8)
    try:
        for pos, bad in [(None, 0.0), (0, np.nan), (n // 2 + 1, np.inf), (n - 1, np.nan), (n - 2, -np.inf)]:
            for shift in (0, 1):
                Q = np.asfortranarray(P).ravel(order="F").copy()
                if pos is not None:
                    Q[pos] = bad
                dQ = ctx.to_device(np.concatenate([np.zeros(shift), Q]))
                try:
                    ctx.energy_device(naux, nvirt, nocc, dQ + 8 * shift, dV, deo, dev, "fp64", None, det)
                    if pos is None:
                        ctx.sync()
                    else:
                        with pytest.raises(rimp2.Rimp2Error) as ei:
                            ctx.sync()
                        assert ei.value.code == rimp2.ERR_NUMERIC, (pos, shift)
                finally:
                    ctx.free(dQ)
    finally:
        for p in (dV, deo, dev, det):
            ctx.free(p)
    ctx.sync()


@pytest.mark.parametrize("mode", ["fp64", "i8s5", "fp32", "bf16x3"])
def test_fitted_coefficients_match_oracle(ctx, oracle, rimp2, mode):
    """Calpha read as the reference reads it (already fitted, B = Calpha)."""
    rng = np.random.default_rng(61)
    naux, nvirt, nocc = 200, 45, 7
    say("fitted", mode, naux, nvirt, nocc)
    C = np.asfortranarray(rng.standard_normal((naux, nvirt, nocc)) * 0.05)
    eo = rng.uniform(-2.0, -0.3, nocc)
    ev = rng.uniform(0.1, 3.0, nvirt)
    Eref, epsref = oracle.pair_energies(C, eo, ev)
    E, eps = ctx.energy_fitted(C, eo, ev, mode)
    check_E(mode, E, Eref)
    if mode == "fp64":
        np.testing.assert_allclose(eps, epsref, rtol=1e-10, atol=1e-15)
        np.testing.assert_array_equal(ctx.get_B(naux, nvirt, nocc), C)


def test_fitted_device_path_matches_host(ctx, rimp2):
    rng = np.random.default_rng(62)
    naux, nvirt, nocc = 160, 70, 5
    C = np.asfortranarray(rng.standard_normal((naux, nvirt, nocc)) * 0.05)
    eo = rng.uniform(-2.0, -0.3, nocc)
    ev = rng.uniform(0.1, 3.0, nvirt)
    say("fitted device")
    for mode in ["fp64", "i8s5"]:
        E_host, eps_host = ctx.energy_fitted(C, eo, ev, mode)
        d = [ctx.to_device(x) for x in (C, eo, ev)]
        dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
        ctx.energy_fitted_device(naux, nvirt, nocc, *d, mode, dpe, det)
        ctx.sync()
        E_dev = float(ctx.from_device(det, (1,))[0])
        eps_dev = ctx.from_device(dpe, (nocc, nocc))
        for p in d + [dpe, det]:
            ctx.free(p)
        assert E_dev == E_host
        np.testing.assert_array_equal(eps_dev, eps_host)


def test_fitted_equals_metric_path_on_transformed_B(ctx, rimp2):
    """Loading the B that the metric path produced gives the same energies."""
    naux, nvirt, nocc = 192, 64, 6
    P, V, eo, ev = rimp2.gen_fragment(63, naux, nvirt, nocc)
    say("fitted vs metric path")
    for mode in ["fp64", "i8s5"]:
        E1, e1 = ctx.energy(P, V, eo, ev, mode)
        B = ctx.get_B(naux, nvirt, nocc)
        E2, e2 = ctx.energy_fitted(B, eo, ev, mode)
        assert E1 == E2
        np.testing.assert_array_equal(e1, e2)


def test_fitted_state_errors(ctx, rimp2):
    say("fitted state errors")
    P, V, eo, ev = rimp2.gen_fragment(64, 48, 10, 3)
    ctx.metric(V)
    ctx.load_fitted(P[:40], "fp64")  # another naux: the factored metric no longer applies
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.transform(P, "fp64")
    assert ei.value.code == rimp2.ERR_STATE
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.pair_energies(eo, ev, "i8s5")  # loaded for fp64
    assert ei.value.code == rimp2.ERR_STATE
    Pn = P.copy(order="F")
    Pn[0, 0, 0] = np.inf
    with pytest.raises(rimp2.Rimp2Error) as ei:
        ctx.load_fitted(Pn, "fp64")
    assert ei.value.code == rimp2.ERR_NUMERIC
    E, _ = ctx.energy_fitted(P, eo, ev, "fp64")
    assert E < 0


def test_get_dev_mem_both_names(ctx, rimp2):
    import ctypes
    say("get_dev_mem")
    t, f = rimp2.get_dev_mem()
    assert t > 200 * 2**30 and 0 < f <= t  # 288 GB of HBM3E
    t2, f2 = ctypes.c_size_t(), ctypes.c_size_t()
    rimp2.lib().get_dev_mem(ctypes.byref(t2), ctypes.byref(f2))
    assert t2.value == t and 0 < f2.value <= t


@pytest.mark.parametrize("lam", [0.0, 1.5])
def test_device_generator_matches_host(ctx, rimp2, lam):
    """The device generator evaluates the host generator's counter-based
    formulas; values agree to rounding (libm, FMA contraction of the affine
    maps, GEMM summation order of V)."""
    naux, nvirt, nocc, seed = 96, 40, 6, 17
    say("device generator", lam)
    P, V, eo, ev = rimp2.gen_fragment_local(seed, naux, nvirt, nocc, lam)
    b = [ctx.alloc(naux * nvirt * nocc * 8), ctx.alloc(naux * naux * 8), ctx.alloc(nocc * 8), ctx.alloc(nvirt * 8)]
    ctx.gen_device(seed, naux, nvirt, nocc, *b, lam)
    Pd = ctx.from_device(b[0], (naux, nvirt, nocc))
    Vd = ctx.from_device(b[1], (naux, naux))
    eod, evd = ctx.from_device(b[2], (nocc,)), ctx.from_device(b[3], (nvirt,))
    for p in b:
        ctx.free(p)
    np.testing.assert_allclose(eod, eo, rtol=1e-15, atol=0)
    np.testing.assert_allclose(evd, ev, rtol=1e-15, atol=0)
    np.testing.assert_allclose(Pd, P, rtol=1e-13, atol=1e-16)
    np.testing.assert_allclose(Vd, V, rtol=1e-13, atol=1e-14)


def test_launches_beyond_32bit_work_items(ctx, oracle, rimp2):
    """256/1536/768: the pair launch covers 9.5M work units x 512 threads
    (> 2^32 work-items, the width of a dispatch's grid size) and the packing /
    generator kernels cover 302M elements (> their capped grids), so every
    grid-stride and unit loop runs more than once.  Pair energies of a seeded
    subset against the oracle; i8s5 against fp64."""
    naux, nvirt, nocc = 768, 1536, 256
    say("beyond 32-bit work-items", naux, nvirt, nocc)
    b = [ctx.alloc(naux * nvirt * nocc * 8), ctx.alloc(naux * naux * 8), ctx.alloc(nocc * 8), ctx.alloc(nvirt * 8)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    try:
        ctx.gen_device(31, naux, nvirt, nocc, *b)
        ctx.energy_device(naux, nvirt, nocc, *b, "i8s5", dpe, det)
        ctx.sync()
        E8 = float(ctx.from_device(det, (1,))[0])
        ctx.energy_device(naux, nvirt, nocc, *b, "fp64", dpe, det)
        ctx.sync()
        E64 = float(ctx.from_device(det, (1,))[0])
        eps = ctx.from_device(dpe, (nocc, nocc))
        eo, ev = ctx.from_device(b[2], (nocc,)), ctx.from_device(b[3], (nvirt,))
        B = ctx.get_B(naux, nvirt, nocc)
    finally:
        for p in b + [dpe, det]:
            ctx.free(p)
    rng = np.random.default_rng(3)
    pi, pj = rng.integers(0, nocc, 8), rng.integers(0, nocc, 8)
    pi[0], pj[0] = nocc - 1, nocc - 1  # the last units of the launch
    ref = oracle.pair_energies_subset(B, eo, ev, pi, pj)
    del B
    np.testing.assert_allclose(eps[pi, pj], ref, rtol=1e-11)
    say("E64", E64, "E_i8s5", E8, "dE uEh", (E8 - E64) * 1e6)
    assert abs(E8 - E64) < 1e-6


def test_mismatched_shapes_raise_before_the_library(ctx, rimp2):
    """V, eps_o and eps_v must match (P|ia) [naux, nvirt, nocc]: the library
    reads V as naux x naux and the orbital energies as nocc / nvirt values."""
    say("shape checks")
    P, V, eo, ev = rimp2.gen_fragment(12, 48, 10, 3)
    for args in ((P, V[:40, :40], eo, ev), (P, V, eo[:2], ev), (P, V, eo, ev[:9]), (P[:, :, 0], V, eo, ev),
                 (P, V.ravel(), eo, ev)):
        with pytest.raises(ValueError):
            ctx.energy(*args)
    with pytest.raises(ValueError):
        ctx.energy_fitted(P, eo, ev[:5])
    with pytest.raises(ValueError):
        ctx.metric(V[:, :40])
    with pytest.raises(ValueError):
        ctx.batch([(P, V, eo[:1], ev)], "fp64")
    E, _ = ctx.energy(P, V, eo, ev, "fp64")
    assert E < 0


def test_async_error_is_not_dropped_by_a_later_blocking_call(ctx, rimp2):
    """energy_device with a non-SPD metric and no sync: the next blocking call
    that uses the context's state (pair_energies) or starts a new energy reports
    the pending error instead of clearing it; the state it left is invalid."""
    say("pending async error")
    naux, nvirt, nocc = 48, 12, 3
    P, V, eo, ev = rimp2.gen_fragment(13, naux, nvirt, nocc)
    bad = V.copy(order="F")
    bad[7, 7] = -50.0
    d = [ctx.to_device(x) for x in (P, bad, eo, ev)]
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    try:
        ctx.energy_device(naux, nvirt, nocc, *d, "fp64", dpe, det)
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.pair_energies(eo, ev, "fp64")
        assert ei.value.code == rimp2.ERR_NOT_SPD
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.pair_energies(eo, ev, "fp64")  # the state is invalid now, not silently reused
        assert ei.value.code == rimp2.ERR_STATE
        ctx.energy_device(naux, nvirt, nocc, *d, "fp64", dpe, det)
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.energy(P, V, eo, ev, "fp64")
        assert ei.value.code == rimp2.ERR_NOT_SPD
    finally:
        for p in d + [dpe, det]:
            ctx.free(p)
    E, _ = ctx.energy(P, V, eo, ev, "fp64")
    assert E < 0


def test_energy_larger_than_free_memory_is_refused_before_allocation(ctx, rimp2):
    """A request whose device buffers exceed hipMemGetInfo's free memory fails
    with RIMP2_ERR_NOMEM before any allocation or read of the (P|ia) and V
    arrays (which are deliberately one element long here): naux = nvirt =
    16384, nocc = 4096 needs ~17 TB."""
    import ctypes
    say("nomem precheck")
    naux, nvirt, nocc = 16384, 16384, 4096
    t0, f0 = rimp2.get_dev_mem()
    one = np.zeros(1)
    eo = np.linspace(-2.0, -0.3, nocc)
    ev = np.linspace(0.1, 3.0, nvirt)
    e = np.zeros(1)
    dp = ctypes.POINTER(ctypes.c_double)
    L = rimp2.lib()
    for mode in (0, 4):
        st = L.rimp2_energy(ctx.h, naux, nvirt, nocc, one.ctypes.data_as(dp), one.ctypes.data_as(dp),
                            eo.ctypes.data_as(dp), ev.ctypes.data_as(dp), mode, None, e.ctypes.data_as(dp))
        assert st == rimp2.ERR_NOMEM, (st, rimp2.last_error(ctx.h))
        assert "GiB" in rimp2.last_error(ctx.h)
    dummy = ctx.alloc(8)
    try:
        st = L.rimp2_energy_device(ctx.h, naux, nvirt, nocc, dummy, dummy, dummy, dummy, 4, None, dummy)
        assert st == rimp2.ERR_NOMEM, (st, rimp2.last_error(ctx.h))
    finally:
        ctx.free(dummy)
    t1, f1 = rimp2.get_dev_mem()
    assert f1 >= f0 - (64 << 20)  # nothing was allocated
    P, V, eo2, ev2 = rimp2.gen_fragment(14, 40, 8, 2)
    E, _ = ctx.energy(P, V, eo2, ev2, "fp64")
    assert E < 0


@pytest.mark.parametrize("limit_mib", [1, 4])
def test_allocation_failure_is_reported_and_the_context_recovers(rimp2, limit_mib):
    """RIMP2_FAULT_ALLOC_BYTES makes every library buffer of at least that
    size fail as hipMalloc would. On a fresh context at naux 512, nvirt 128,
    nocc 16 a 1 MiB limit fails the first buffer (the 2 MiB metric) and a
    4 MiB limit fails later, after the metric buffers exist (the 8 MiB (P|ia)
    and B). Both surface as RIMP2_ERR_NOMEM, and with the variable cleared the
    same context then forms the golden energy."""
    g = np.load(os.path.join(GOLD, "frag_4_512_128_16.npz"))
    say("injected allocation failure, limit MiB", limit_mib)
    P, V, eo, ev = rimp2.gen_fragment(int(g["seed"]), int(g["naux"]), int(g["nvirt"]), int(g["nocc"]))
    c = rimp2.Context(0)
    try:
        os.environ["RIMP2_FAULT_ALLOC_BYTES"] = str(limit_mib << 20)
        try:
            for mode in ("fp64", "i8s5"):
                with pytest.raises(rimp2.Rimp2Error) as ei:
                    c.energy(P, V, eo, ev, mode)
                assert ei.value.code == rimp2.ERR_NOMEM
                assert "injected" in str(ei.value)
        finally:
            del os.environ["RIMP2_FAULT_ALLOC_BYTES"]
        for mode in ("fp64", "i8s5"):
            E, eps = c.energy(P, V, eo, ev, mode)
            check_E(mode, E, float(g["E"]))
        np.testing.assert_allclose(c.energy(P, V, eo, ev, "fp64")[1], g["eps_pair"], rtol=1e-10, atol=1e-14)
    finally:
        c.close()
