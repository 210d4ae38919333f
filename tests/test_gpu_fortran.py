"""The Fortran host driver (rimp2_driver.x) over the C ABI.

The driver reads a fragment file in the reference's record order
(RIMP2.F90:45-63: header nvirt,nocc,noccEOS,nvirtEOS,NBA,nocctot as default
integers; Calpha(NBA,nvirt,nocc); EVocc; EVvirt; then the U matrices and,
as generate.F90:333-336 writes them, the expected toccEOS, tvirtEOS,
gvirtEOS, goccEOS) or generates a seeded fragment.  It forms the four
local-basis arrays on the GPU and checks them element by element at the
reference's 1e-10 (RIMP2.F90:253-323), then runs the FP64 energy and one other
precision mode when the denominators allow it.  Files are written here as
sequential unformatted records with 4-byte length markers.

Parity: the FP64 energy against the oracle (metric transform + pair loops) to
1e-12 relative; the other mode against the FP64 energy within its accuracy.
The two failure cases (missing file, bad usage) stop with the reference's
failure word ERROR3 before any GPU call, so they run without a GPU.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "rimp2_gpu_amd", "fortran", "rimp2_driver.x")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="rimp2_driver.x not built (amdflang missing)")


def say(*a):
    print(*a, flush=True)


def write_records(path, records):
    with open(path, "wb") as f:
        for rec in records:
            b = rec.tobytes(order="F")
            f.write(struct.pack("<i", len(b)))
            f.write(b)
            f.write(struct.pack("<i", len(b)))


def run(*args, timeout=120):
    r = subprocess.run([EXE] + [str(a) for a in args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    return r.returncode, r.stdout


def parse(out):
    vals = {}
    for line in out.splitlines():
        w = line.split()
        if not w:
            continue
        if w[0] == "E_MP2" and len(w) == 3:
            vals["mode"], vals["E"] = int(w[1]), float(w[2])
        elif w[0] in ("E_MP2_fp64", "dE_uEh", "eps_11"):
            vals[w[0]] = float(w[1])
    return vals


@needs_exe
def test_missing_file_stops_with_error3(tmp_path):
    rc, out = run("file", tmp_path / "absent.data", 0)
    assert rc != 0 and "ERROR3" in out


@needs_exe
def test_bad_usage_stops_with_error3():
    rc, out = run("nonsense")
    assert rc != 0 and "ERROR3" in out


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("mode,tol_uEh", [(4, 1e-3), (0, 0.0)])
def test_generated_fragment(oracle, rimp2, mode, tol_uEh):
    seed, naux, nvirt, nocc = 7, 256, 48, 12
    say("fortran gen", seed, naux, nvirt, nocc, "mode", mode)
    rc, out = run("gen", seed, naux, nvirt, nocc, mode)
    assert rc == 0, out
    v = parse(out)
    P, V, eo, ev = rimp2.gen_fragment(seed, naux, nvirt, nocc)
    B = oracle.metric_transform(V, P)
    E_ref, eps_ref = oracle.pair_energies(B, eo, ev)
    assert v["mode"] == mode
    assert abs(v["E_MP2_fp64"] - E_ref) <= 1e-12 * abs(E_ref)
    assert abs(v["E"] - v["E_MP2_fp64"]) * 1e6 <= tol_uEh + 1e-12 * abs(E_ref) * 1e6
    assert abs(v["eps_11"] - eps_ref[0, 0]) <= 1e-9 * abs(eps_ref[0, 0])


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 2, 4])
def test_reference_record_file(tmp_path, oracle, mode):
    rng = np.random.default_rng(31 + mode)
    nvirt, nocc, noccEOS, nvirtEOS, naux = 40, 9, 3, 5, 160
    nocctot = nocc + 2
    Calpha = np.asfortranarray(rng.standard_normal((naux, nvirt, nocc)) * 0.05)
    eo = np.sort(rng.uniform(-2.0, -0.3, nocc))
    ev = np.sort(rng.uniform(0.1, 3.0, nvirt))
    recs = [np.array([nvirt, nocc, noccEOS, nvirtEOS, naux, nocctot], dtype=np.int32), Calpha, eo, ev,
            rng.standard_normal((nocc, noccEOS)), rng.standard_normal((nvirt, nvirt)),
            rng.standard_normal((nocc, nocc)), rng.standard_normal((nvirt, nvirtEOS))]
    path = tmp_path / "Calpha1.data"
    write_records(path, recs)
    say("fortran file", path, "mode", mode)
    rc, out = run("file", path, mode)
    assert rc == 0, out
    v = parse(out)
    assert "nvirt,nocc,noccEOS,nvirtEOS,NBA,nocctot" in out
    E_ref, eps_ref = oracle.pair_energies(Calpha, eo, ev)
    assert abs(v["E_MP2_fp64"] - E_ref) <= 1e-12 * abs(E_ref)
    rel = {0: 0.0, 2: 2e-5, 4: 1e-9}[mode]
    assert abs(v["E"] - E_ref) <= rel * abs(E_ref) + 1e-12 * abs(E_ref)


def reference_fixture_records(oracle, dims):
    """A fragment file exactly as generate.F90 writes it (first_order off,
    generate.F90:75-128, 333-336): header, the fixture inputs, and the four
    arrays its CPU Steps 5-8 produce (here: the oracle's)."""
    from test_oracle import ref_fixture
    nvirt, nocc, noccEOS, nvirtEOS, nba = dims
    args = ref_fixture(*dims)
    to, tv, go, gv = oracle.steps5to8(*args)
    hdr = np.array([nvirt, nocc, noccEOS, nvirtEOS, nba, nocc], dtype=np.int32)
    return [hdr, *args, to, tv, gv, go], (to, tv, go, gv)


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(74, 10, 4, 15, 273), (104, 20, 4, 15, 373)])
def test_reference_fixture_file_steps5to8(tmp_path, oracle, dims):
    recs, _ = reference_fixture_records(oracle, dims)
    path = tmp_path / "Calpha2.data"
    write_records(path, recs)
    say("fortran steps 5-8 on the reference fixture", dims)
    rc, out = run("file", path, 4)
    assert rc == 0, out
    assert "EOS arrays match the file within 1e-10" in out
    assert "E_MP2 skipped" in out  # the fixture's denominators are positive


@needs_exe
@pytest.mark.gpu
def test_reference_fixture_file_mismatch_stops_with_error3(tmp_path, oracle):
    dims = (74, 10, 4, 15, 273)
    recs, _ = reference_fixture_records(oracle, dims)
    tv = recs[9].copy(order="F")
    tv[3, 2, 1, 4] += 5e-10  # beyond the reference's 1e-10
    recs[9] = tv
    path = tmp_path / "Calpha1.data"
    write_records(path, recs)
    say("fortran steps 5-8 mismatch")
    rc, out = run("file", path, 0)
    assert rc != 0 and "ERROR3" in out
    assert "tvirtEOS" in out
    assert "4 3 2 5" in " ".join(out.split())  # 1-based index of the changed element


@needs_exe
@pytest.mark.gpu
def test_reference_fixture_file_nan_stops_with_error3(tmp_path, oracle):
    """A NaN in the expected arrays fails the element check (the reference's
    `ABS(test - got) > 1e-10` would let it pass)."""
    recs, _ = reference_fixture_records(oracle, (74, 10, 4, 15, 273))
    go = recs[11].copy(order="F")
    go[0, 1, 2, 3] = np.nan
    recs[11] = go
    path = tmp_path / "Calpha_nan.data"
    write_records(path, recs)
    rc, out = run("file", path, 0)
    assert rc != 0 and "ERROR3" in out and "goccEOS" in out


@needs_exe
@pytest.mark.gpu
def test_record_file_eos_sums(tmp_path, oracle):
    """Random inputs with U records but no expected arrays: the driver forms
    the four arrays and prints their sums, checked against the oracle."""
    rng = np.random.default_rng(5)
    nvirt, nocc, noccEOS, nvirtEOS, naux = 40, 9, 3, 5, 160
    C = np.asfortranarray(rng.standard_normal((naux, nvirt, nocc)) * 0.05)
    eo = np.sort(rng.uniform(-2.0, -0.3, nocc))
    ev = np.sort(rng.uniform(0.1, 3.0, nvirt))
    U = [np.asfortranarray(rng.standard_normal(s)) for s in
         ((nocc, noccEOS), (nvirt, nvirt), (nocc, nocc), (nvirt, nvirtEOS))]
    write_records(tmp_path / "f.data", [np.array([nvirt, nocc, noccEOS, nvirtEOS, naux, nocc], dtype=np.int32),
                                        C, eo, ev, *U])
    rc, out = run("file", tmp_path / "f.data", 0)
    assert rc == 0, out
    line = [ln for ln in out.splitlines() if ln.startswith("EOS_sums")][0]
    got = [float(x) for x in line.split()[1:]]
    exp = oracle.steps5to8(C, eo, ev, *U)
    for g, e in zip(got, exp):
        assert abs(g - float(np.sum(e))) <= 1e-12 * float(np.sum(np.abs(e))), (g, float(np.sum(e)))


@needs_exe
@pytest.mark.gpu
def test_first_order_record_file(tmp_path, oracle):
    """The first-order record layout (RIMP2.F90:189, 230; generate.F90:262,
    308, 333-339): the driver forms toccEOS..goccEOS and djik/blad on the GPU
    and checks every element against the file at 1e-10."""
    from test_gpu_eos import random_inputs
    nvirt, nocc, noccEOS, nvirtEOS, nba = 40, 9, 3, 5, 160
    C, eo, ev, UoE, Uv, Uo, UvE = random_inputs(77, nvirt, nocc, noccEOS, nvirtEOS, nba)
    rng = np.random.default_rng(78)
    Co = np.asfortranarray(0.1 * rng.standard_normal((nba, nocc, nocc)))
    Cvv = np.asfortranarray(0.1 * rng.standard_normal((nba, nvirt, nvirt)))
    to, tv, go, gv = oracle.steps5to8(C, eo, ev, UoE, Uv, Uo, UvE)
    dj, bl = oracle.first_order(C, Co, Cvv, UoE, Uv, Uo, UvE)
    hdr = np.array([nvirt, nocc, noccEOS, nvirtEOS, nba, nocc], dtype=np.int32)
    path = tmp_path / "Calpha_fo.data"
    write_records(path, [hdr, C, eo, ev, UoE, Uv, Uo, UvE, Co, Cvv, to, tv, gv, go, dj, bl])
    say("fortran first-order file")
    rc, out = run("file", path, 4, "fo")
    assert rc == 0, out
    assert "EOS arrays match the file within 1e-10" in out
    assert "first-order arrays match the file within 1e-10" in out
    # a planted difference in blad stops the run as the reference's check does
    bl2 = bl.copy(order="F")
    bl2[1, 2, 3, 4] += 3e-10
    write_records(path, [hdr, C, eo, ev, UoE, Uv, Uo, UvE, Co, Cvv, to, tv, gv, go, dj, bl2])
    rc, out = run("file", path, 4, "fo")
    assert rc != 0 and "ERROR3" in out and "blad" in out


def parse_lines(out, key):
    return [line.split()[1:] for line in out.splitlines() if line.split()[:1] == [key]]


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 4])
def test_batch_mode_against_the_oracle(oracle, rimp2, mode):
    """rimp2_batch from host arrays and rimp2_batch_prepare_device / _run /
    _result on device copies, driven from Fortran: every fragment's energy
    against the oracle, and the two paths bit for bit."""
    seed, nfrag = 21, 12
    say("fortran batch", seed, nfrag, "mode", mode)
    rc, out = run("batch", seed, nfrag, mode, timeout=300)
    assert rc == 0, out
    rows = parse_lines(out, "E_FRAG")
    assert len(rows) == nfrag
    o, v, q = rimp2.gen_batch_shapes(seed, nfrag)
    tol = 1e-11 if mode == 0 else None
    for k, (idx, e_host, e_dev) in enumerate(rows):
        assert int(idx) == k
        assert float(e_host) == float(e_dev)
        P, V, eo, ev = rimp2.gen_fragment(seed + k, int(q[k]), int(v[k]), int(o[k]))
        Eref, _ = oracle.pair_energies(oracle.metric_transform(V, P), eo, ev)
        if tol:
            assert abs(float(e_host) - Eref) <= tol * abs(Eref), (k, e_host, Eref)
        else:
            assert abs(float(e_host) - Eref) < 1e-6, (k, e_host, Eref)
    tile = parse_lines(out, "BATCH_TILE")[0]
    assert int(tile[0]) in (-1, 32, 64)  # a tile candidate (-1: per fragment)


@needs_exe
@pytest.mark.gpu
def test_screen_mode_keeps_exact_pair_energies(oracle, rimp2):
    """Surrogate training and screening driven from Fortran (locality-decay
    generator, lam = 1.5): the exact energy against the oracle, the screened
    energy equal to the oracle's energy less the skipped pairs, and no kept pair
    differing from the exact kernel's (the driver stops with ERROR3 otherwise)."""
    seed, naux, nvirt, nocc, lam, ntrain, thresh = 999, 256, 64, 32, 1.5, 12, 1e-6
    say("fortran screen")
    rc, out = run("screen", seed, naux, nvirt, nocc, 0, lam, ntrain, thresh, timeout=300)
    assert rc == 0, out
    vals = {k: float(parse_lines(out, k)[0][0]) for k in ("E_EXACT", "E_SCREENED", "NSKIPPED")}
    skips = [(int(a), int(b)) for a, b, _ in parse_lines(out, "SKIP")]
    assert len(skips) == int(vals["NSKIPPED"]) > 0
    P, V, eo, ev = rimp2.gen_fragment_local(seed, naux, nvirt, nocc, lam)
    Eref, epsref = oracle.pair_energies(oracle.metric_transform(V, P), eo, ev)
    assert abs(vals["E_EXACT"] - Eref) <= 1e-11 * abs(Eref)
    dropped = sum((1.0 if i == j else 2.0) * epsref[i, j] for i, j in skips)
    assert abs(vals["E_SCREENED"] - (Eref - dropped)) <= 1e-11 * abs(Eref)
    assert max(abs(epsref[i, j]) for i, j in skips) < 100 * thresh
    losses = [float(r[1]) for r in parse_lines(out, "EPOCH_LOSS")]
    assert len(losses) == 6 and losses[-1] < losses[0]


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 4])
def test_device_mode_against_the_oracle(oracle, rimp2, mode):
    """Device buffers, the device generator, rimp2_energy_device + rimp2_sync,
    the inspection calls and rimp2_energy_fitted_device, all from Fortran."""
    seed, naux, nvirt, nocc = 8, 192, 40, 6
    say("fortran device", mode)
    rc, out = run("device", seed, naux, nvirt, nocc, mode)
    assert rc == 0, out
    vals = {k: float(parse_lines(out, k)[0][0]) for k in ("E_DEVICE", "E_FITTED_DEVICE", "PAIR_SUM",
                                                          "MAX_LINV_L_MINUS_I")}
    P, V, eo, ev = rimp2.gen_fragment(seed, naux, nvirt, nocc)
    Eref, _ = oracle.pair_energies(oracle.metric_transform(V, P), eo, ev)
    # the device generator agrees with the host one to rounding
    tol = 1e-10 * abs(Eref) + (1e-6 if mode == 4 else 0.0)
    assert abs(vals["E_DEVICE"] - Eref) <= tol
    assert vals["E_FITTED_DEVICE"] == vals["E_DEVICE"]
    assert abs(vals["PAIR_SUM"] - vals["E_DEVICE"]) <= 1e-12 * abs(Eref)
    assert vals["MAX_LINV_L_MINUS_I"] < 1e-10
    n, ms = parse_lines(out, "PAIR_KERNEL_LAUNCHES_MS")[0]
    assert int(n) == 1 and float(ms) > 0
