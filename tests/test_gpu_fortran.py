"""The Fortran host driver (rimp2_driver.x) over the C ABI.

The driver reads a fragment file in the reference's record order
(RIMP2.F90:45-63: header nvirt,nocc,noccEOS,nvirtEOS,NBA,nocctot as default
integers; Calpha(NBA,nvirt,nocc); EVocc; EVvirt; then the U matrices, which
the energy path does not read) or generates a seeded fragment, runs the FP64
energy and one other precision mode, and prints both.  Files are written here
as sequential unformatted records with 4-byte length markers.

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
