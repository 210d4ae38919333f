"""Per-fragment results file (checkpoint / resume of batched runs), on the CPU:
exact float round trip, torn-write handling, and an interrupted run resumed to
the results of an uninterrupted one, with a stand-in for the GPU context."""
import json
import os

import numpy as np
import pytest

from rimp2_gpu_amd import results


def rec(key, E, nocc=2, mode="fp64", pairs=True):
    r = {"key": key, "mode": mode, "nocc": nocc, "nvirt": 8, "naux": 16, "E": E}
    if pairs:
        r["pairs"] = [E / (k + 1) for k in range(nocc * nocc)]
    return r


def test_round_trip_is_bit_exact(tmp_path):
    p = str(tmp_path / "r.jsonl")
    vals = [-1234.5678901234567, -5e-324, -2.2250738585072014e-308, -1.0000000000000002, -0.1, -7.0]
    results.append(p, [rec(f"f{k}", v) for k, v in enumerate(vals)])
    got = results.load(p)
    for k, v in enumerate(vals):
        r = got[f"f{k}"]
        assert r["E"] == v and np.float64(r["E"]).tobytes() == np.float64(v).tobytes()
        assert r["pairs"] == [v / (m + 1) for m in range(4)]


def test_torn_last_line_is_ignored_then_cut_off(tmp_path):
    p = str(tmp_path / "r.jsonl")
    results.append(p, [rec("a", -1.5), rec("b", -2.5)])
    full = open(p, "rb").read()
    line_b = full.split(b"\n")[1]
    # a crash in the middle of writing the second record
    with open(p, "wb") as fh:
        fh.write(full[: len(full) - len(line_b) // 2 - 1])
    assert set(results.load(p)) == {"a"}
    # the next append starts where the last whole line ended
    results.append(p, [rec("c", -3.5)])
    lines = open(p, "rb").read().split(b"\n")
    assert lines[-1] == b"" and len(lines) == 3
    assert set(results.load(p)) == {"a", "c"}
    # a torn record that happens to be valid JSON but lacks its newline is not taken
    with open(p, "ab") as fh:
        fh.write(json.dumps(rec("d", -4.5)).encode())
    assert set(results.load(p)) == {"a", "c"}


def test_bad_lines_raise(tmp_path):
    p = str(tmp_path / "r.jsonl")
    results.append(p, [rec("a", -1.5)])
    with open(p, "ab") as fh:
        fh.write(b"{not json\n")
    results.append(p, [rec("b", -2.5)])
    with pytest.raises(results.ResultsFileError, match="line 2"):
        results.load(p)
    q = str(tmp_path / "q.jsonl")
    bad = rec("a", -1.5)
    del bad["naux"]
    with open(q, "w") as fh:
        fh.write(json.dumps(bad) + "\n")
    with pytest.raises(results.ResultsFileError, match="naux"):
        results.load(q)
    short = rec("a", -1.5, nocc=3)
    short["pairs"] = short["pairs"][:5]
    with open(q, "w") as fh:
        fh.write(json.dumps(short) + "\n")
    with pytest.raises(results.ResultsFileError, match="pair energies"):
        results.load(q)


class Frags:
    """Lazily generated fragments; records which ones were generated."""

    def __init__(self, n, seed=3):
        rng = np.random.default_rng(seed)
        self.shapes = [(int(rng.integers(1, 5)), int(rng.integers(2, 9)), int(rng.integers(4, 17))) for _ in range(n)]
        self.made = []

    def __len__(self):
        return len(self.shapes)

    def __getitem__(self, k):
        self.made.append(k)
        nocc, nvirt, naux = self.shapes[k]
        rng = np.random.default_rng(1000 + k)
        P = rng.standard_normal((naux, nvirt, nocc))
        return P, np.eye(naux), -1.0 - rng.random(nocc), 1.0 + rng.random(nvirt)


class FakeCtx:
    """Stands in for Context.batch: per-fragment values that depend only on
    the fragment (as the library's do), and a crash after `fail_after` calls."""

    def __init__(self, fail_after=None):
        self.calls, self.nfrag, self.fail_after = 0, 0, fail_after

    def batch(self, frags, mode, want_pairs=True):
        if self.fail_after is not None and self.calls >= self.fail_after:
            raise RuntimeError("simulated crash")
        self.calls += 1
        self.nfrag += len(frags)
        ef, pp = [], []
        for P, V, eo, ev in frags:
            n = eo.size
            M = np.add.outer(eo, eo) * float(np.sum(P ** 2))
            ef.append(float(M.sum()))
            pp.append(M.reshape(-1, order="F"))
        return np.array(ef), (np.concatenate(pp) if want_pairs else None)


def test_interrupted_run_resumes_to_the_uninterrupted_result(tmp_path):
    n = 23
    keys = [f"seed{k}" for k in range(n)]
    E0, P0 = results.run_batch(FakeCtx(), Frags(n), keys, str(tmp_path / "straight.jsonl"), mode="fp32", chunk=5)
    p = str(tmp_path / "r.jsonl")
    with pytest.raises(RuntimeError, match="simulated crash"):
        results.run_batch(FakeCtx(fail_after=2), Frags(n), keys, p, mode="fp32", chunk=5)
    assert len(results.load(p)) == 10  # two whole chunks recorded
    fr, ctx = Frags(n), FakeCtx()
    E1, P1 = results.run_batch(ctx, fr, keys, p, mode="fp32", chunk=5, shapes=Frags(n).shapes)
    assert ctx.nfrag == n - 10 and sorted(fr.made) == list(range(10, n))  # finished ones not regenerated
    np.testing.assert_array_equal(E1, E0)
    for a, b in zip(P1, P0):
        np.testing.assert_array_equal(a, b)
    # a third run computes nothing
    ctx = FakeCtx()
    E2, _ = results.run_batch(ctx, Frags(n), keys, p, mode="fp32", chunk=5)
    assert ctx.calls == 0
    np.testing.assert_array_equal(E2, E0)


def test_mode_shape_and_key_checks(tmp_path):
    p = str(tmp_path / "r.jsonl")
    keys = [f"k{k}" for k in range(4)]
    results.run_batch(FakeCtx(), Frags(4), keys, p, mode="fp64", chunk=3)
    with pytest.raises(results.ResultsFileError, match="mode fp64"):
        results.run_batch(FakeCtx(), Frags(4), keys, p, mode="i8s5")
    other = [(s[0] + 1, s[1], s[2]) for s in Frags(4).shapes]
    with pytest.raises(results.ResultsFileError, match="nocc, nvirt, naux"):
        results.run_batch(FakeCtx(), Frags(4), keys, p, mode="fp64", shapes=other)
    with pytest.raises(ValueError, match="unique key"):
        results.run_batch(FakeCtx(), Frags(4), ["a", "a", "b", "c"], p)
    # energies without pair energies are completed when pairs are asked for
    q = str(tmp_path / "q.jsonl")
    results.run_batch(FakeCtx(), Frags(4), keys, q, mode="fp64", want_pairs=False)
    ctx = FakeCtx()
    E, P = results.run_batch(ctx, Frags(4), keys, q, mode="fp64")
    assert ctx.nfrag == 4 and all(x.shape == (s[0], s[0]) for x, s in zip(P, Frags(4).shapes))
    assert os.path.getsize(q) > 0


def test_torn_line_longer_than_the_scan_block(tmp_path):
    """A torn record longer than append's backward-scan block is still cut
    back to the last whole line."""
    p = str(tmp_path / "r.jsonl")
    results.append(p, [rec("a", -1.5)])
    big = rec("b", -2.5, nocc=200)  # ~40000 pair energies: several scan blocks
    line = json.dumps(big).encode()
    assert len(line) > 3 * (1 << 16)
    with open(p, "ab") as fh:
        fh.write(line[: len(line) - 10])  # no newline: torn
    results.append(p, [rec("c", -3.5)])
    assert set(results.load(p)) == {"a", "c"}
    assert open(p, "rb").read().count(b"\n") == 2
