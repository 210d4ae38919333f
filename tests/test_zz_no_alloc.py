"""The allocation-free entry point (rimp2_energy_fitted_device_noalloc) against
rimp2_energy_fitted_device, through the C ABI.

Both run the same packing, pair kernel and reductions in the same order, so
energies and pair matrices must be identical bit for bit, on two calls in a
row with the second smaller than the first (the static scratch keeps the
first call's contents).  The entry point allocates nothing (free device memory
does not drop once the stream has run a kernel), refuses a fragment beyond its scratch with RIMP2_ERR_NOMEM,
and reports a non-finite input through rimp2_sync.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(512, 128, 16), (192, 70, 5)]  # (naux, nvirt, nocc): first larger, then smaller


def say(*a):
    print(*a, flush=True)


def fitted_inputs(seed, naux, nvirt, nocc):
    rng = np.random.default_rng(seed)
    C = np.asfortranarray(rng.standard_normal((naux, nvirt, nocc)) * 0.05)
    return C, rng.uniform(-2.0, -0.3, nocc), rng.uniform(0.1, 3.0, nvirt)


def run(ctx, fn, dims, d, nocc):
    dpe, det = ctx.alloc(nocc * nocc * 8), ctx.alloc(8)
    try:
        fn(*dims, *d, dpe, det)
        ctx.sync()
        return float(ctx.from_device(det, (1,))[0]), ctx.from_device(dpe, (nocc, nocc))
    finally:
        ctx.free(dpe)
        ctx.free(det)


@pytest.mark.parametrize("mode", ["fp64", "i8s5", "fp32", "bf16x3", "i8s4"])
def test_noalloc_matches_the_allocating_entry_point(rimp2, mode):
    ctx = rimp2.Context(0)
    try:
        inputs = [fitted_inputs(41 + k, *s) for k, s in enumerate(SHAPES)]
        dev = [[ctx.to_device(x) for x in inp] for inp in inputs]
        outs = [ctx.alloc(s[2] * s[2] * 8) for s in SHAPES] + [ctx.alloc(8) for _ in SHAPES]
        ctx.sync()
        # the HIP runtime sets up a stream's kernel-argument pool on its first
        # launch (4 MiB); one call beforehand keeps that out of the reading
        naux, nvirt, nocc = SHAPES[1]
        ctx.energy_fitted_device_noalloc(naux, nvirt, nocc, *dev[1], mode, outs[1], outs[len(SHAPES) + 1])
        ctx.sync()
        t0, f0 = rimp2.get_dev_mem()
        got = []
        for k, (naux, nvirt, nocc) in enumerate(SHAPES):
            ctx.energy_fitted_device_noalloc(naux, nvirt, nocc, *dev[k], mode, outs[k], outs[len(SHAPES) + k])
            ctx.sync()
            got.append((float(ctx.from_device(outs[len(SHAPES) + k], (1,))[0]),
                        ctx.from_device(outs[k], (nocc, nocc))))
        t1, f1 = rimp2.get_dev_mem()
        say("noalloc", mode, "free before", f0, "after", f1)
        assert f1 >= f0  # nothing was allocated
        for k, (naux, nvirt, nocc) in enumerate(SHAPES):
            E_ref, eps_ref = run(ctx, lambda *a: ctx.energy_fitted_device(*a[:6], mode, *a[6:]),
                                 (naux, nvirt, nocc), dev[k], nocc)
            E, eps = got[k]
            say("  shape", (naux, nvirt, nocc), "E", E, "ref", E_ref)
            assert E == E_ref
            np.testing.assert_array_equal(eps, eps_ref)
            assert E < 0
        for p in [q for d in dev for q in d] + outs:
            ctx.free(p)
    finally:
        ctx.close()


def test_noalloc_refuses_beyond_its_scratch_and_reports_bad_input(rimp2):
    ctx = rimp2.Context(0)
    try:
        dummy = ctx.alloc(64)
        # 8192 x 2048 x 8 in fp64 packs 1 GiB: refused before anything is queued
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.energy_fitted_device_noalloc(8192, 2048, 8, dummy, dummy, dummy, "fp64", None, dummy)
        assert ei.value.code == rimp2.ERR_NOMEM
        assert "RIMP2_NOALLOC_SCRATCH_BYTES" in str(ei.value)
        ctx.free(dummy)
        naux, nvirt, nocc = 96, 40, 4
        C, eo, ev = fitted_inputs(7, naux, nvirt, nocc)
        C[3, 5, 1] = np.nan
        d = [ctx.to_device(x) for x in (C, eo, ev)]
        det = ctx.alloc(8)
        ctx.energy_fitted_device_noalloc(naux, nvirt, nocc, *d, "i8s5", None, det)
        with pytest.raises(rimp2.Rimp2Error) as ei:
            ctx.sync()
        assert ei.value.code == rimp2.ERR_NUMERIC
        ctx.sync()  # read once, then cleared
        C[3, 5, 1] = 0.01
        ctx.free(d[0])
        d[0] = ctx.to_device(C)
        ctx.energy_fitted_device_noalloc(naux, nvirt, nocc, *d, "i8s5", None, det)
        ctx.sync()
        assert float(ctx.from_device(det, (1,))[0]) < 0
        for p in d + [det]:
            ctx.free(p)
    finally:
        ctx.close()
