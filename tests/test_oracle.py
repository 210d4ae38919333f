"""Pins the CPU oracle (CPU-only tests).

Each check is chosen so that a plausible mistake in the restatement (wrong
index order, transposed U, wrong sign of Delta, dropped exchange term) fails
at least one of them:
  * closed-form values of the reference's own fixtures (generate.F90:77-128
    pushed through Major Steps 5-8, RIMP2.F90:67-170), abs tol 1e-10 as the
    reference's check (RIMP2.F90:262);
  * an independent numpy einsum of Steps 5-8 on random, non-constant inputs;
  * pair-energy brute force, closed forms and invariants;
  * LAPACK (scipy) for the metric step.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sl


# ---------------------------------------------------------------------------
# reference fixtures (generate.F90:77-128), evaluated in the same order
# ---------------------------------------------------------------------------
def ref_fixture(nvirt, nocc, noccEOS, nvirtEOS, nba):
    A = np.arange(1, nba + 1, dtype=np.float64)[:, None, None]
    B = np.arange(1, nvirt + 1, dtype=np.float64)[None, :, None]
    C = np.arange(1, nocc + 1, dtype=np.float64)[None, None, :]
    Calpha = np.asfortranarray(A * 0.00124235636 / A + B * 0.01010101 / B + C * 0.014321 / C)
    c = np.arange(1, nocc + 1, dtype=np.float64)
    EVocc = c * 1.934321 / c
    c = np.arange(1, nvirt + 1, dtype=np.float64)
    EVvirt = c * 1.91321 / c

    def U(n1, n2, x, y):
        a = np.arange(1, n1 + 1, dtype=np.float64)[:, None]
        b = np.arange(1, n2 + 1, dtype=np.float64)[None, :]
        return np.asfortranarray(a * x / a + b * y / b)

    UoccEOST = U(nocc, noccEOS, 0.030321, 0.101121)
    UvirtT = U(nvirt, nvirt, 0.021321, 0.01321)
    UoccT = U(nocc, nocc, 0.032321, 0.01521)
    UvirtEOST = U(nvirt, nvirtEOS, 0.023321, 0.01421)
    return Calpha, EVocc, EVvirt, UoccEOST, UvirtT, UoccT, UvirtEOST


def closed_form(nvirt, nocc, nba):
    c = 0.00124235636 + 0.01010101 + 0.014321
    d = 2 * 1.934321 - 2 * 1.91321
    uoE, uv, uo, uvE = 0.030321 + 0.101121, 0.021321 + 0.01321, 0.032321 + 0.01521, 0.023321 + 0.01421
    tocc = nba * c * c / d * (nocc * uoE) ** 2 * (nvirt * uv) ** 2
    tvirt = nba * c * c / d * (nvirt * uvE) ** 2 * (nocc * uo) ** 2
    gocc = nba * (nvirt * uv * nocc * uoE * c) ** 2
    gvirt = nba * (nvirt * uvE * nocc * uo * c) ** 2
    return tocc, tvirt, gocc, gvirt


# file1: generate.F90:41-46, file2: generate.F90:58-63; values tabulated from
# the fixture formulas.
FIXTURES = [
    ((74, 10, 4, 15, 273), (48.0435406249666, 7.4213438296809, 2.02849437226733, 0.313343979176785)),
    ((104, 20, 4, 15, 373), (518.614220438611, 80.1109659023915, 21.8969296153589, 3.38244520233076)),
]


@pytest.mark.parametrize("dims,expect", FIXTURES)
def test_reference_fixture_closed_form(oracle, dims, expect):
    nvirt, nocc, noccEOS, nvirtEOS, nba = dims
    cf = closed_form(nvirt, nocc, nba)
    np.testing.assert_allclose(cf, expect, rtol=0, atol=1e-10)
    outs = oracle.steps5to8(*ref_fixture(nvirt, nocc, noccEOS, nvirtEOS, nba))
    shapes = [(nvirt, noccEOS, nvirt, noccEOS), (nvirtEOS, nocc, nvirtEOS, nocc),
              (nvirt, noccEOS, nvirt, noccEOS), (nvirtEOS, nocc, nvirtEOS, nocc)]
    for arr, shp, val in zip(outs, shapes, expect):
        assert arr.shape == shp
        assert np.all(np.isfinite(arr))
        # the reference's own check: ABS(test - got) > 1e-10 -> ERROR3
        assert np.max(np.abs(arr - val)) <= 1e-10


def numpy_steps5to8(C, eo, ev, UoE, Uv, Uo, UvE):
    """Independent restatement in einsum form (canonical basis -> local)."""
    g = np.einsum("qai,qbj->aibj", C, C)  # g(a,i,b,j)
    D = eo[None, :, None, None] + eo[None, None, None, :] - ev[:, None, None, None] - ev[None, None, :, None]
    t = g / D
    # toccEOS(Al,Il,Bl,Jl) = sum U_v(A,Al) U_oE(I,Il) U_v(B,Bl) U_oE(J,Jl) t(A,I,B,J)
    toccEOS = np.einsum("aibj,aA,iI,bB,jJ->AIBJ", t, Uv, UoE, Uv, UoE)
    # tvirtEOS(Al,Il,Bl,Jl) = sum U_vE(A,Al) U_o(I,Il) U_vE(B,Bl) U_o(J,Jl) t(A,I,B,J)
    tvirtEOS = np.einsum("aibj,aA,iI,bB,jJ->AIBJ", t, UvE, Uo, UvE, Uo)
    goccEOS = np.einsum("aibj,aA,iI,bB,jJ->AIBJ", g, Uv, UoE, Uv, UoE)
    gvirtEOS = np.einsum("aibj,aA,iI,bB,jJ->AIBJ", g, UvE, Uo, UvE, Uo)
    return toccEOS, tvirtEOS, goccEOS, gvirtEOS


def test_steps5to8_random_inputs(oracle):
    rng = np.random.default_rng(11)
    nba, nvirt, nocc, noccEOS, nvirtEOS = 7, 6, 4, 3, 5
    C = np.asfortranarray(rng.standard_normal((nba, nvirt, nocc)))
    eo = rng.uniform(-2, -0.5, nocc)
    ev = rng.uniform(0.2, 2, nvirt)
    UoE = np.asfortranarray(rng.standard_normal((nocc, noccEOS)))
    Uv = np.asfortranarray(rng.standard_normal((nvirt, nvirt)))
    Uo = np.asfortranarray(rng.standard_normal((nocc, nocc)))
    UvE = np.asfortranarray(rng.standard_normal((nvirt, nvirtEOS)))
    got = oracle.steps5to8(C, eo, ev, UoE, Uv, Uo, UvE)
    exp = numpy_steps5to8(C, eo, ev, UoE, Uv, Uo, UvE)
    for g, e in zip(got, exp):
        np.testing.assert_allclose(g, e, rtol=1e-12, atol=1e-12)


def test_identity_U_reduces_to_canonical(oracle):
    rng = np.random.default_rng(3)
    nba, nvirt, nocc = 9, 5, 3
    C = np.asfortranarray(rng.standard_normal((nba, nvirt, nocc)))
    eo, ev = rng.uniform(-2, -0.5, nocc), rng.uniform(0.2, 2, nvirt)
    t = oracle.toccA(C, eo, ev, np.eye(nocc))  # tocc(I,J,A,B)
    g = oracle.gocc(C)                          # g(A,I,B,J)
    D = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    np.testing.assert_allclose(t, np.einsum("aibj->ijab", g) / D, rtol=1e-13)
    # gocc == gvirt when U = I and EOS = all orbitals
    np.testing.assert_array_equal(oracle.gvirt(C), g)
    # V^{ij}_{ab} = V^{ji}_{ba}
    np.testing.assert_allclose(g, np.einsum("aibj->bjai", g), rtol=1e-14)
    # tvirtA with U = I carries the same amplitudes in (I,J,A,B) order
    tv = oracle.tvirtA(C, eo, ev, np.eye(nvirt))
    np.testing.assert_allclose(tv, t, rtol=1e-13)


def test_gen4dimfo_einsum(oracle):
    rng = np.random.default_rng(5)
    C3 = rng.standard_normal((6, 4, 3))
    C4 = rng.standard_normal((6, 5, 2))
    np.testing.assert_allclose(oracle.gen4dimfo(C3, C4), np.einsum("qab,qcd->abcd", C3, C4), rtol=1e-13)


# ---------------------------------------------------------------------------
# first-order intermediates djik / blad (RIMP2.F90:172-251)
# ---------------------------------------------------------------------------
def numpy_first_order(C, Co, Cvv, UoE, Uv, Uo, UvE):
    """Independent restatement.  RIMP2.F90:231-235 calls dgemm with K = ldb =
    nocc on CalphaVV(NBA,nvirt,nvirt) and UvirtT(nvirt,nvirt): only the first
    nocc virtual columns of CalphaVV are summed, against UvirtT's storage read
    as an nocc x nvirt matrix."""
    nba, nvirt, nocc = C.shape
    C3 = np.einsum("qbj,jJ,bB->qBJ", C, UoE, Uv)
    C4 = np.einsum("qik,kK,iI->qIK", Co, Uo, UoE)
    djik = np.einsum("qab,qcd->abcd", C3, C4)
    B3 = np.einsum("qbj,jJ,bB->qBJ", C, Uo, UvE)
    Uq = Uv.ravel(order="F")[: nocc * nvirt].reshape((nocc, nvirt), order="F")
    B4 = np.einsum("qak,kb,aA->qAb", Cvv[:, :, :nocc], Uq, UvE)
    blad = np.einsum("qab,qcd->abcd", B3, B4)
    return djik, blad


def first_order_inputs(rng, nba, nvirt, nocc, noccEOS, nvirtEOS):
    C = np.asfortranarray(rng.standard_normal((nba, nvirt, nocc)))
    Co = np.asfortranarray(rng.standard_normal((nba, nocc, nocc)))
    Cvv = np.asfortranarray(rng.standard_normal((nba, nvirt, nvirt)))
    U = [np.asfortranarray(rng.standard_normal(s)) for s in
         ((nocc, noccEOS), (nvirt, nvirt), (nocc, nocc), (nvirt, nvirtEOS))]
    return C, Co, Cvv, *U


def test_first_order_random_inputs(oracle):
    args = first_order_inputs(np.random.default_rng(23), 7, 9, 4, 3, 5)
    got = oracle.first_order(*args)
    exp = numpy_first_order(*args)
    for g, e in zip(got, exp):
        assert g.shape == e.shape
        np.testing.assert_allclose(g, e, rtol=1e-12, atol=1e-12)


def test_first_order_closed_form(oracle):
    nvirt, nocc, noccEOS, nvirtEOS, nba = 11, 5, 2, 3, 13
    Cal, _, _, UoE, Uv, Uo, UvE = ref_fixture(nvirt, nocc, noccEOS, nvirtEOS, nba)[:7]
    cO, cV = 0.0173, 0.0291
    Co = np.full((nba, nocc, nocc), cO, order="F")
    Cvv = np.full((nba, nvirt, nvirt), cV, order="F")
    djik, blad = oracle.first_order(Cal, Co, Cvv, UoE, Uv, Uo, UvE)
    c = 0.00124235636 + 0.01010101 + 0.014321
    uoE, uv, uo, uvE = 0.030321 + 0.101121, 0.021321 + 0.01321, 0.032321 + 0.01521, 0.023321 + 0.01421
    d_val = nba * (c * nocc * uoE * nvirt * uv) * (cO * nocc * uo * nocc * uoE)
    # K = nocc in the CalphaVV product: nocc terms, not nvirt
    b_val = nba * (c * nocc * uo * nvirt * uvE) * (nvirt * uvE * cV * nocc * uv)
    assert np.max(np.abs(djik - d_val)) <= 1e-13 * d_val
    assert np.max(np.abs(blad - b_val)) <= 1e-13 * b_val


def test_first_order_quirk_is_kept(oracle):
    """A dgemm with K = nvirt (the evident intent) gives a different blad."""
    args = first_order_inputs(np.random.default_rng(29), 6, 8, 3, 2, 4)
    C, Co, Cvv, UoE, Uv, Uo, UvE = args
    _, blad = oracle.first_order(*args)
    B3 = np.einsum("qbj,jJ,bB->qBJ", C, Uo, UvE)
    B4 = np.einsum("qak,kb,aA->qAb", Cvv, Uv, UvE)
    intended = np.einsum("qab,qcd->abcd", B3, B4)
    assert np.max(np.abs(blad - intended)) > 1e-3 * np.max(np.abs(intended))


# ---------------------------------------------------------------------------
# pair energies
# ---------------------------------------------------------------------------
def brute_energy(B, eo, ev):
    nocc = B.shape[2]
    eps = np.zeros((nocc, nocc))
    for i in range(nocc):
        for j in range(nocc):
            V = B[:, :, i].T @ B[:, :, j]
            D = eo[i] + eo[j] - ev[:, None] - ev[None, :]
            terms = (V / D) * (2 * V - V.T)
            eps[i, j] = math.fsum(terms.ravel())
    return math.fsum(eps.ravel()), eps


def rand_frag(rng, naux, nvirt, nocc):
    B = np.asfortranarray(rng.standard_normal((naux, nvirt, nocc)) * 0.1)
    return B, rng.uniform(-2, -0.3, nocc), rng.uniform(0.1, 3.0, nvirt)


@pytest.mark.parametrize("shape", [(7, 5, 3), (10, 6, 4), (33, 9, 5), (1, 1, 1), (16, 1, 3), (5, 7, 1)])
def test_pair_energy_brute_force(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    B, eo, ev = rand_frag(rng, *shape)
    E, eps = oracle.pair_energies(B, eo, ev)
    Eb, epsb = brute_energy(B, eo, ev)
    np.testing.assert_allclose(eps, epsb, rtol=1e-13, atol=1e-16)
    assert abs(E - Eb) <= 1e-13 * abs(Eb) + 1e-16
    np.testing.assert_array_equal(eps, eps.T)


def test_pair_energy_single_orbital_closed_form(oracle):
    B = np.asfortranarray(np.array([0.3, -0.2, 0.5]).reshape(3, 1, 1))
    eo, ev = np.array([-1.1]), np.array([0.7])
    V = 0.09 + 0.04 + 0.25
    E, _ = oracle.pair_energies(B, eo, ev)
    assert abs(E - V * V / (2 * -1.1 - 2 * 0.7)) < 1e-16


def test_pair_energy_invariants(oracle):
    rng = np.random.default_rng(7)
    B, eo, ev = rand_frag(rng, 20, 8, 5)
    E, eps = oracle.pair_energies(B, eo, ev)
    assert np.all(eps <= 0.0) and E < 0
    s = 1.7
    E2, _ = oracle.pair_energies(B * s, eo, ev)
    assert abs(E2 - s ** 4 * E) <= 1e-12 * abs(E2)
    po, pv = rng.permutation(5), rng.permutation(8)
    E3, eps3 = oracle.pair_energies(np.asfortranarray(B[:, pv][:, :, po]), eo[po], ev[pv])
    assert abs(E3 - E) <= 1e-12 * abs(E)
    np.testing.assert_allclose(eps3, eps[np.ix_(po, po)], rtol=1e-12)
    # E = sum_i eps_ii + 2 sum_{i<j} eps_ij
    assert abs(E - (np.trace(eps) + 2 * np.sum(np.triu(eps, 1)))) <= 1e-12 * abs(E)


def test_pair_energy_subset_matches_full(oracle):
    rng = np.random.default_rng(9)
    B, eo, ev = rand_frag(rng, 24, 10, 6)
    _, eps = oracle.pair_energies(B, eo, ev)
    pi, pj = np.array([0, 2, 5, 1]), np.array([0, 3, 5, 4])
    np.testing.assert_array_equal(oracle.pair_energies_subset(B, eo, ev, pi, pj), eps[pi, pj])


def test_exchange_term_matters(oracle):
    """A restatement that dropped -V_ba (or used V_ab twice) fails here."""
    rng = np.random.default_rng(13)
    B, eo, ev = rand_frag(rng, 12, 6, 3)
    E, _ = oracle.pair_energies(B, eo, ev)
    nocc = 3
    E_no_x = 0.0
    for i in range(nocc):
        for j in range(nocc):
            V = B[:, :, i].T @ B[:, :, j]
            D = eo[i] + eo[j] - ev[:, None] - ev[None, :]
            E_no_x += np.sum(2 * V * V / D)
    assert abs(E - E_no_x) > 1e-3 * abs(E)


# ---------------------------------------------------------------------------
# metric
# ---------------------------------------------------------------------------
def spd(rng, n):
    A = rng.standard_normal((n, n))
    return np.asfortranarray(A @ A.T / n + np.eye(n))


@pytest.mark.parametrize("n", [1, 5, 64, 130])
def test_cholesky_vs_lapack(oracle, n):
    rng = np.random.default_rng(n)
    V = spd(rng, n)
    L = oracle.cholesky(V)
    np.testing.assert_allclose(L, sl.cholesky(V, lower=True), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(L @ L.T, V, rtol=0, atol=1e-13 * n * np.abs(V).max())
    assert np.all(np.triu(L, 1) == 0)


def test_cholesky_rejects_indefinite(oracle):
    V = np.asfortranarray(np.array([[1.0, 2.0], [2.0, 1.0]]))
    with pytest.raises(np.linalg.LinAlgError):
        oracle.cholesky(V)


def test_metric_transform_identities(oracle):
    rng = np.random.default_rng(17)
    naux, nvirt, nocc = 40, 6, 3
    V = spd(rng, naux)
    X = np.asfortranarray(rng.standard_normal((naux, nvirt, nocc)))
    B = oracle.metric_transform(V, X)
    X2, B2 = X.reshape(naux, -1, order="F"), B.reshape(naux, -1, order="F")
    L = sl.cholesky(V, lower=True)
    np.testing.assert_allclose(B2, sl.solve_triangular(L, X2, lower=True), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(B2.T @ B2, X2.T @ np.linalg.solve(V, X2), rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(oracle.metric_transform(np.asfortranarray(np.eye(naux)), X), X, rtol=0, atol=0)


# ---------------------------------------------------------------------------
# screening MLP oracle
# ---------------------------------------------------------------------------
def mlp_case(seed, n=37):
    rng = np.random.default_rng(seed)
    import oracle as o
    npar = o.lib().oracle_mlp_num_params()
    assert npar == 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1
    P = (rng.standard_normal(npar) * 0.2).astype(np.float32)
    F = rng.standard_normal((n, 16)).astype(np.float32) * 2 + 1
    T = rng.standard_normal(n).astype(np.float32) + 3
    norm = np.concatenate([np.full(16, 1.0), np.full(16, 2.0), [3.0, 1.5]]).astype(np.float32)
    return P, F, T, norm


def test_mlp_oracle_finite_differences(oracle):
    P, F, T, norm = mlp_case(1)
    loss, g, _ = oracle.mlp_loss_grad(F, T, P, norm)
    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(1024, 6), 1024 + rng.choice(64, 3), 1088 + rng.choice(4096, 8),
                          5184 + rng.choice(64, 3), 5248 + rng.choice(64, 3), [5312]])
    for k in idx:
        h = 1e-3 * max(1.0, abs(float(P[k])))
        Pp, Pm = P.astype(np.float64).copy(), P.astype(np.float64).copy()
        # parameters are float32 at the interface: step by an exactly representable amount
        Pp[k] = np.float32(P[k] + h)
        Pm[k] = np.float32(P[k] - h)
        lp, _, _ = oracle.mlp_loss_grad(F, T, Pp.astype(np.float32), norm, want_grad=False)
        lm, _, _ = oracle.mlp_loss_grad(F, T, Pm.astype(np.float32), norm, want_grad=False)
        fd = (lp - lm) / (float(Pp[k]) - float(Pm[k]))
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(g[k]), (k, fd, g[k])
    assert loss > 0


def test_gelu_and_bf16_closed_forms(oracle):
    import math as m
    assert oracle.gelu(0.0) == 0.0
    assert abs(oracle.gelu(1.0) - 0.5 * (1 + m.erf(1 / m.sqrt(2)))) < 1e-16
    for x in [-3.0, -0.7, 0.2, 2.5]:
        assert abs(oracle.gelu(x) - oracle.gelu(-x) - x) < 1e-15  # GELU(x) - GELU(-x) = x
        h = 1e-6
        assert abs((oracle.gelu(x + h) - oracle.gelu(x - h)) / (2 * h) - oracle.gelu_d(x)) < 1e-8
    assert oracle.gelu_d(0.0) == 0.5
    assert oracle.bf16(1.0 + 2.0 ** -8) == 1.0            # tie -> even
    assert oracle.bf16(1.0 + 3 * 2.0 ** -9) == 1.0 + 2.0 ** -7
    assert oracle.bf16(1.0 + 2.0 ** -7 + 2.0 ** -8) == 1.0 + 2.0 ** -6  # tie -> even (up)
    assert oracle.bf16(-2.5) == -2.5


def _bf16_np(a):
    """Round-to-nearest-even to bf16 on float32 bit patterns (NaN-free inputs)."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _mlp_forward_np(P, F, T, norm, bf16=False):
    """Independent numpy restatement of the network, from the documented
    parameter order W1[16][64] | b1[64] | W2[64][64] (input k -> output m) |
    b2[64] | w3[64] | b3 and norm = mean[16] | std[16] | tmean | tstd:
      x  = clip((F - mean)/std, -4, 4)
      y  = w3 . gelu(W2^T gelu(W1^T x + b1) + b2) + b3
      pred = F[:,15] + y tstd + tmean,  loss = mean (y - (T - F[:,15] - tmean)/tstd)^2
    With bf16=True, x, the weights and both hidden activations are rounded to
    bf16 (the MFMA operands); biases and sums stay FP64."""
    from scipy.special import erf
    r = _bf16_np if bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    P = P.astype(np.float64)
    W1 = r(P[:1024].reshape(16, 64))
    b1 = P[1024:1088]
    W2 = r(P[1088:5184].reshape(64, 64))
    b2 = P[5184:5248]
    w3 = r(P[5248:5312])
    b3 = P[5312]
    nm = norm.astype(np.float64)
    x = r(np.clip((F.astype(np.float64) - nm[:16]) / nm[16:32], -4.0, 4.0))
    gelu = lambda z: 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))
    h1 = r(gelu(x @ W1 + b1))
    h2 = r(gelu(h1 @ W2 + b2))
    y = h2 @ w3 + b3
    base = F[:, 15].astype(np.float64)
    pred = base + y * nm[33] + nm[32]
    loss = np.mean((y - (T.astype(np.float64) - base - nm[32]) / nm[33]) ** 2)
    return loss, pred


@pytest.mark.parametrize("bf16", [False, True])
def test_mlp_oracle_forward_matches_numpy_restatement(oracle, bf16):
    P, F, T, norm = mlp_case(5, n=200)
    F[:7, 3] = 40.0  # standardised input beyond +4: saturates
    F[7:9, 5] = -40.0
    loss, _, pred = oracle.mlp_loss_grad(F, T, P, norm, emulate_bf16=bf16, want_grad=False)
    lref, pref = _mlp_forward_np(P, F, T, norm, bf16=bf16)
    np.testing.assert_allclose(pred, pref, rtol=1e-12, atol=1e-12)
    assert abs(loss - lref) <= 1e-12 * lref
    # a transposed W2 (the layout mistake both implementations could share)
    # moves the prediction far beyond that tolerance; an input scale that
    # saturates a feature for every sample is clamped alike in both
    Pt = P.copy()
    Pt[1088:5184] = P[1088:5184].reshape(64, 64).T.ravel()
    _, pt = _mlp_forward_np(Pt, F, T, norm, bf16=bf16)
    assert np.max(np.abs(pt - pred)) > 1e-3
    nm2 = norm.copy()
    nm2[16 + 3] = 1e-3  # input 3 now saturates for every sample
    _, _, psat = oracle.mlp_loss_grad(F, T, P, nm2, emulate_bf16=bf16, want_grad=False)
    _, psat_ref = _mlp_forward_np(P, F, T, nm2, bf16=bf16)
    np.testing.assert_allclose(psat, psat_ref, rtol=1e-12, atol=1e-12)


def test_mlp_bf16_emulation_is_close_to_fp64(oracle):
    P, F, T, norm = mlp_case(3, n=64)
    l64, g64, p64 = oracle.mlp_loss_grad(F, T, P, norm)
    lbf, gbf, pbf = oracle.mlp_loss_grad(F, T, P, norm, emulate_bf16=True)
    assert np.linalg.norm(gbf - g64) <= 5e-2 * np.linalg.norm(g64)
    assert np.max(np.abs(pbf - p64)) < 0.1 * np.max(np.abs(p64))


def test_generator_locality_variant(rimp2):
    P0, V0, eo0, ev0 = rimp2.gen_fragment(8, 64, 16, 6)
    P1, V1, eo1, ev1 = rimp2.gen_fragment_local(8, 64, 16, 6, 0.0)
    np.testing.assert_array_equal(P0, P1)
    np.testing.assert_array_equal(V0, V1)
    P2, V2, _, _ = rimp2.gen_fragment_local(8, 64, 16, 6, 1.0)
    # occupied 0 lives at the start of the chain: far auxiliary functions are damped
    assert np.abs(P2[-8:, :, 0]).max() < 1e-2 * np.abs(P2[:8, :, 0]).max()
    assert np.linalg.eigvalsh(V2).min() > 0.99
