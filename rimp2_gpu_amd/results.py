"""Per-fragment results file: checkpoint and resume for long batched runs.

A run over many DEC fragments (one batched launch per chunk of fragments)
appends one JSON line per finished fragment to a results file:

    {"key": "...", "mode": "i8s5", "nocc": 12, "nvirt": 96, "naux": 384,
     "E": -1234.5678, "pairs": [... the fragment's nocc*nocc pair energies ...]}

Floats are written with Python's shortest round-trip repr, so a value read
back is bit-identical to the one computed.  Each chunk's lines go out in one
write followed by fsync, so after a crash the file holds every chunk that
finished and at most a torn last line, which `load` ignores and the next
`append` cuts off.  `run_batch` skips the fragments whose key is already in
the file in the same mode, runs the rest chunk by chunk, and returns every
fragment's results in the caller's order: an interrupted run restarted with
the same arguments ends with the results of an uninterrupted one.

The reference keeps no state between runs beyond its input files
(RIMP2.F90:36 loops over fragments in one process); this is the resume path
of the batched driver.
"""
import json
import os

import numpy as np

__all__ = ["load", "append", "run_batch"]


class ResultsFileError(ValueError):
    pass


def load(path):
    """Records of a results file by key ({} if the file does not exist).  The
    text after the last newline is a torn write and is ignored (`append`
    truncates it before writing); a bad line anywhere else raises
    ResultsFileError.  A key recorded more than once takes its last record."""
    if not os.path.exists(path):
        return {}
    with open(path, "rb") as fh:
        data = fh.read()
    body = data.split(b"\n")[:-1]
    out = {}
    for n, raw in enumerate(body):
        if not raw.strip():
            continue
        try:
            rec = json.loads(raw)
        except json.JSONDecodeError as e:
            raise ResultsFileError(f"{path}: line {n + 1} is not a JSON record ({e})") from None
        _check_record(path, n + 1, rec)
        out[rec["key"]] = rec
    return out


def _check_record(path, line, rec):
    need = {"key": str, "mode": str, "nocc": int, "nvirt": int, "naux": int, "E": float}
    for f, t in need.items():
        if f not in rec or not isinstance(rec[f], t):
            raise ResultsFileError(f"{path}: line {line}: field {f!r} missing or not {t.__name__}")
    if "pairs" in rec and len(rec["pairs"]) != rec["nocc"] ** 2:
        raise ResultsFileError(f"{path}: line {line}: {len(rec['pairs'])} pair energies for nocc {rec['nocc']}")


def append(path, records):
    """Appends records (dicts as above) as one write, then fsync.  A torn
    last line left by an earlier crash is cut off first."""
    if not records:
        return
    text = "".join(json.dumps(r, separators=(",", ":"), allow_nan=False) + "\n" for r in records)
    fd = os.open(path, os.O_RDWR | os.O_CREAT, 0o644)
    try:
        size = os.fstat(fd).st_size
        keep = _after_last_newline(fd, size)
        if keep != size:
            os.ftruncate(fd, keep)
        os.pwrite(fd, text.encode(), keep)
        os.fsync(fd)
    finally:
        os.close(fd)


def _after_last_newline(fd, size, block=1 << 16):
    """Offset just past the last newline of the first `size` bytes (0 if none),
    reading backwards from the end one block at a time."""
    end = size
    while end > 0:
        start = max(0, end - block)
        k = os.pread(fd, end - start, start).rfind(b"\n")
        if k >= 0:
            return start + k + 1
        end = start
    return 0


def _shape(frag):
    Pia = frag[0]
    naux, nvirt, nocc = (int(x) for x in np.shape(Pia))
    return nocc, nvirt, naux


def run_batch(ctx, frags, keys, path, mode="fp32", chunk=64, want_pairs=True, shapes=None, log=None):
    """Batched energies of `frags` (a sequence of (Pia, V, eps_o, eps_v), as
    Context.batch takes; it may produce each fragment lazily in __getitem__)
    with per-fragment checkpointing in `path`.

    keys: one unique string per fragment naming it across runs (for example
    its seed).  A fragment whose key is already recorded in the same mode is
    not recomputed (nor generated); one recorded in another mode raises
    ResultsFileError, and so does one recorded with another shape than
    `shapes[k]` = (nocc, nvirt, naux) when shapes are given.  The rest run
    `chunk` fragments per batched launch, each chunk's records appended as it
    finishes.

    Returns (E [nfrag], pairs: list of [nocc, nocc] arrays or None); the
    pair energies keep the library's column-major order."""
    n = len(frags)
    keys = [str(k) for k in keys]
    if len(keys) != n or len(set(keys)) != n:
        raise ValueError("run_batch needs one unique key per fragment")
    if chunk < 1:
        raise ValueError("chunk must be at least 1")
    done = load(path)
    todo = []
    for k in range(n):
        rec = done.get(keys[k])
        if rec is None:
            todo.append(k)
            continue
        if rec["mode"] != mode:
            raise ResultsFileError(f"{path}: fragment {keys[k]!r} was computed in mode {rec['mode']}, not {mode}")
        if shapes is not None and tuple(shapes[k]) != (rec["nocc"], rec["nvirt"], rec["naux"]):
            raise ResultsFileError(f"{path}: fragment {keys[k]!r} was computed at (nocc, nvirt, naux) = "
                                   f"{(rec['nocc'], rec['nvirt'], rec['naux'])}, not {tuple(shapes[k])}")
        if want_pairs and "pairs" not in rec:
            todo.append(k)  # recomputed with its pair energies
            continue
    if log:
        log(f"[results] {n - len(todo)} of {n} fragments already in {path}; {len(todo)} to run")
    for c0 in range(0, len(todo), chunk):
        idx = todo[c0:c0 + chunk]
        sub = [frags[k] for k in idx]
        ef, pp = ctx.batch(sub, mode, want_pairs)
        recs, off = [], 0
        for m, k in enumerate(idx):
            nocc, nvirt, naux = _shape(sub[m])
            rec = {"key": keys[k], "mode": mode, "nocc": nocc, "nvirt": nvirt, "naux": naux, "E": float(ef[m])}
            if want_pairs:
                rec["pairs"] = [float(x) for x in pp[off:off + nocc * nocc]]
            off += nocc * nocc
            recs.append(rec)
        append(path, recs)
        for r in recs:
            done[r["key"]] = r
        if log:
            log(f"[results] chunk of {len(idx)} fragments recorded ({c0 + len(idx)}/{len(todo)})")
    E = np.empty(n)
    pairs = [] if want_pairs else None
    for k in range(n):
        rec = done[keys[k]]
        E[k] = rec["E"]
        if want_pairs:
            nocc = rec["nocc"]
            pairs.append(np.asarray(rec["pairs"], dtype=np.float64).reshape(nocc, nocc, order="F"))
    return E, pairs
