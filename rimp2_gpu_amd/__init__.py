"""rimp2_gpu_amd -- MI355X-native RI-MP2 energy engine (Python host binding).

A thin ctypes binding over the C ABI of ``librimp2_hip.so`` (declared in
``include/rimp2_hip.h``).  Arrays are numpy float64, column-major (Fortran
order) exactly as the reference's ``Calpha(NBA,nvirt,nocc)``
(RIMP2.F90:48).  There is no CPU fallback: every compute call goes through the
HIP library and raises ``Rimp2Error`` on failure.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librimp2_hip.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_NOT_SPD, ERR_STATE, ERR_NOMEM, ERR_NUMERIC = 1, 2, 3, 4, 5, 6

MODES = {"fp64": 0, "fp32": 1, "bf16x3": 2, "i8s4": 3, "i8s5": 4}
MODE_NAMES = {v: k for k, v in MODES.items()}
KERNELS = {"pair": 0, "transform": 1, "pack": 2, "cholesky": 3, "trtri": 4, "reduce": 5,
           "batch": 6, "mlp": 7, "features": 8, "eos": 9,
           "eos_ampl": 10}
NFEAT = 16
NOALLOC_SCRATCH_BYTES = 128 << 20

_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p


class Rimp2Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[rimp2 status {code}] {msg}")
        self.code = code


class FragDesc(ctypes.Structure):
    _fields_ = [("nocc", _i64), ("nvirt", _i64), ("naux", _i64),
                ("Pia", _dp), ("V", _dp), ("eps_o", _dp), ("eps_v", _dp)]


_lib = None

# name -> (restype, argtypes); every symbol declared in include/rimp2_hip.h
SIGNATURES = {
    "rimp2_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "rimp2_ctx_destroy": (ctypes.c_int, [_vp]),
    "rimp2_get_dev_mem": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "rimp2_last_error": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t]),
    "rimp2_metric": (ctypes.c_int, [_vp, _i64, _dp]),
    "rimp2_transform": (ctypes.c_int, [_vp, _i64, _i64, _i64, _dp, ctypes.c_int]),
    "rimp2_pair_energies": (ctypes.c_int, [_vp, _i64, _i64, _dp, _dp, ctypes.c_int, _dp, _dp]),
    "rimp2_energy": (ctypes.c_int, [_vp, _i64, _i64, _i64, _dp, _dp, _dp, _dp, ctypes.c_int, _dp, _dp]),
    "rimp2_energy_device": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "get_dev_mem": (None, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "rimp2_load_fitted": (ctypes.c_int, [_vp, _i64, _i64, _i64, _dp, ctypes.c_int]),
    "rimp2_energy_fitted": (ctypes.c_int, [_vp, _i64, _i64, _i64, _dp, _dp, _dp, ctypes.c_int, _dp, _dp]),
    "rimp2_energy_fitted_device": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _vp, _vp, ctypes.c_int, _vp, _vp]),
    "rimp2_energy_fitted_device_noalloc": (ctypes.c_int, [_vp, _i64, _i64, _i64, _vp, _vp, _vp, ctypes.c_int, _vp,
                                                          _vp]),
    "rimp2_get_cholesky": (ctypes.c_int, [_vp, _dp]),
    "rimp2_get_metric_inverse": (ctypes.c_int, [_vp, _dp]),
    "rimp2_get_B": (ctypes.c_int, [_vp, _dp]),
    "rimp2_batch": (ctypes.c_int, [_vp, _i64, ctypes.POINTER(FragDesc), ctypes.c_int, _dp, _dp]),
    "rimp2_batch_prepare": (ctypes.c_int, [_vp, _i64, ctypes.POINTER(FragDesc), ctypes.c_int]),
    "rimp2_batch_prepare_device": (ctypes.c_int, [_vp, _i64, ctypes.POINTER(FragDesc), ctypes.c_int]),
    "rimp2_batch_run": (ctypes.c_int, [_vp, _vp]),
    "rimp2_batch_result": (ctypes.c_int, [_vp, _dp, _dp]),
    "rimp2_batch_tile": (ctypes.c_int, [_vp, ctypes.POINTER(_i64)]),
    "rimp2_batch_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_i64), _dp]),
    "rimp2_batch_set_tile": (ctypes.c_int, [_vp, _i64]),
    "rimp2_pair_features": (ctypes.c_int, [_vp, _i64, _i64, _dp, _dp, _fp]),
    "rimp2_mlp_init": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "rimp2_mlp_set_norm": (ctypes.c_int, [_vp, _fp, _fp, ctypes.c_float, ctypes.c_float]),
    "rimp2_mlp_train_step": (ctypes.c_int, [_vp, _i64, _fp, _fp, ctypes.c_float, _fp]),
    "rimp2_mlp_predict": (ctypes.c_int, [_vp, _i64, _fp, _fp]),
    "rimp2_mlp_get_params": (ctypes.c_int, [_vp, _fp, _i64]),
    "rimp2_mlp_set_params": (ctypes.c_int, [_vp, _fp, _i64]),
    "rimp2_mlp_num_params": (_i64, []),
    "rimp2_mlp_grad": (ctypes.c_int, [_vp, _i64, _fp, _fp, _fp, _fp]),
    "rimp2_screened_energy": (ctypes.c_int, [_vp, _i64, _i64, _dp, _dp, ctypes.c_int, ctypes.c_double, _dp, _dp,
                                             ctypes.POINTER(_i64)]),
    "rimp2_device_alloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "rimp2_device_free": (ctypes.c_int, [_vp, _vp]),
    "rimp2_copy_h2d": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "rimp2_copy_d2h": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "rimp2_sync": (ctypes.c_int, [_vp]),
    "rimp2_timing_enable": (ctypes.c_int, [_vp, ctypes.c_int]),
    "rimp2_timing_get": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.POINTER(_i64)]),
    "rimp2_timing_reset": (ctypes.c_int, [_vp]),
    "rimp2_gen_fragment": (ctypes.c_int, [ctypes.c_uint64, _i64, _i64, _i64, _dp, _dp, _dp, _dp]),
    "rimp2_gen_fragment_local": (ctypes.c_int, [ctypes.c_uint64, _i64, _i64, _i64, ctypes.c_double, _dp, _dp, _dp,
                                                _dp]),
    "rimp2_gen_fragment_device": (ctypes.c_int, [_vp, ctypes.c_uint64, _i64, _i64, _i64, ctypes.c_double, _vp, _vp,
                                                 _vp, _vp]),
    "rimp2_gen_batch_shapes": (ctypes.c_int, [ctypes.c_uint64, _i64, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                              ctypes.POINTER(_i64)]),
    "rimp2_dgemm": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _i64, _i64, _i64, _dp, _vp, _i64, _vp, _i64, _dp,
                                   _vp, _i64]),
    "rimp2_eos_arrays": (ctypes.c_int, [_vp, _i64, _i64, _i64, _i64, _i64] + [_dp] * 11),
    "rimp2_eos_arrays_device": (ctypes.c_int, [_vp, _i64, _i64, _i64, _i64, _i64] + [_vp] * 11),
    "rimp2_eos_set_chunk": (ctypes.c_int, [_vp, _i64]),
    "rimp2_eos_set_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "rimp2_first_order": (ctypes.c_int, [_vp, _i64, _i64, _i64, _i64, _i64, _i64] + [_dp] * 9),
    "rimp2_first_order_device": (ctypes.c_int, [_vp, _i64, _i64, _i64, _i64, _i64, _i64] + [_vp] * 9),
}


def lib():
    """Loads librimp2_hip.so (raises if it was not built: no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run __graft_entry__.build() or python -m rimp2_gpu_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a, ct=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ct))


def _f64(a, shape=None):
    a = np.asarray(a, dtype=np.float64)
    if shape is not None and a.size != int(np.prod(shape)):
        raise ValueError(f"expected {shape} elements, got {a.shape}")
    return np.asfortranarray(a) if a.ndim > 1 else np.ascontiguousarray(a)


def _fragment(Pia, V, eps_o, eps_v):
    """Checks one fragment's arrays against the shape of (P|ia) [naux, nvirt,
    nocc] before any pointer reaches the library (which reads V as
    naux x naux, eps_o as nocc and eps_v as nvirt values); V may be None (an
    already-fitted tensor)."""
    Pia = _f64(Pia)
    if Pia.ndim != 3:
        raise ValueError(f"(P|ia) must be [naux, nvirt, nocc], got shape {Pia.shape}")
    naux, nvirt, nocc = Pia.shape
    if V is not None:
        V = np.asarray(V)
        if V.shape != (naux, naux):
            raise ValueError(f"V must be ({naux}, {naux}) for naux={naux}, got {V.shape}")
        V = _f64(V)
    eo, ev = np.asarray(eps_o), np.asarray(eps_v)
    if eo.shape != (nocc,) or ev.shape != (nvirt,):
        raise ValueError(f"eps_o must hold nocc={nocc} and eps_v nvirt={nvirt} values, got {eo.shape} and {ev.shape}")
    return Pia, V, _f64(eo), _f64(ev)


def _mode(m):
    if isinstance(m, str):
        if m not in MODES:
            raise ValueError(f"unknown mode {m!r}; choose from {sorted(MODES)}")
        return MODES[m]
    return int(m)


def last_error(ctx=None):
    buf = ctypes.create_string_buffer(2048)
    lib().rimp2_last_error(ctx, buf, 2048)
    return buf.value.decode(errors="replace")


def check(st, ctx=None):
    if st != OK:
        raise Rimp2Error(st, last_error(ctx))


def gen_fragment(seed, naux, nvirt, nocc, want_V=True):
    """Seeded synthetic fragment: Pia [naux,nvirt,nocc] (F order), V, eps_o, eps_v."""
    Pia = np.empty((naux, nvirt, nocc), dtype=np.float64, order="F")
    V = np.empty((naux, naux), dtype=np.float64, order="F") if want_V else None
    eo = np.empty(nocc)
    ev = np.empty(nvirt)
    check(lib().rimp2_gen_fragment(seed, naux, nvirt, nocc, _ptr(Pia), _ptr(V), _ptr(eo), _ptr(ev)))
    return Pia, V, eo, ev


def gen_fragment_local(seed, naux, nvirt, nocc, lam):
    """Locality-decay variant of gen_fragment (lam <= 0 reproduces it exactly)."""
    Pia = np.empty((naux, nvirt, nocc), dtype=np.float64, order="F")
    V = np.empty((naux, naux), dtype=np.float64, order="F")
    eo = np.empty(nocc)
    ev = np.empty(nvirt)
    check(lib().rimp2_gen_fragment_local(seed, naux, nvirt, nocc, lam, _ptr(Pia), _ptr(V), _ptr(eo), _ptr(ev)))
    return Pia, V, eo, ev


def gen_batch_shapes(seed, nfrag):
    o = np.empty(nfrag, np.int64)
    v = np.empty(nfrag, np.int64)
    q = np.empty(nfrag, np.int64)
    p = ctypes.POINTER(_i64)
    check(lib().rimp2_gen_batch_shapes(seed, nfrag, o.ctypes.data_as(p), v.ctypes.data_as(p), q.ctypes.data_as(p)))
    return o, v, q


def get_dev_mem():
    t, f = ctypes.c_size_t(), ctypes.c_size_t()
    check(lib().rimp2_get_dev_mem(ctypes.byref(t), ctypes.byref(f)))
    return t.value, f.value


class Context:
    """One HIP context (stream + device buffers), the replacement of the
    reference's cuBLAS handle (cublasCreate_v2, RIMP2.GPU.F90:81)."""

    def __init__(self, device=0):
        h = _vp()
        check(lib().rimp2_ctx_create(device, ctypes.byref(h)))
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib().rimp2_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        check(st, self.h)

    # ---- energy path ----
    def metric(self, V):
        V = _f64(V)
        if V.ndim != 2 or V.shape[0] != V.shape[1]:
            raise ValueError(f"V must be square, got shape {V.shape}")
        n = V.shape[0]
        self._chk(lib().rimp2_metric(self.h, n, _ptr(V)))

    def transform(self, Pia, mode="fp64"):
        Pia = _f64(Pia)
        if Pia.ndim != 3:
            raise ValueError(f"(P|ia) must be [naux, nvirt, nocc], got shape {Pia.shape}")
        naux, nvirt, nocc = Pia.shape
        self._chk(lib().rimp2_transform(self.h, naux, nvirt, nocc, _ptr(Pia), _mode(mode)))

    def load_fitted(self, Calpha, mode="fp64"):
        """Already-fitted Calpha [naux, nvirt, nocc] becomes B (no metric)."""
        Calpha = _f64(Calpha)
        if Calpha.ndim != 3:
            raise ValueError(f"Calpha must be [naux, nvirt, nocc], got shape {Calpha.shape}")
        naux, nvirt, nocc = Calpha.shape
        self._chk(lib().rimp2_load_fitted(self.h, naux, nvirt, nocc, _ptr(Calpha), _mode(mode)))

    def energy_fitted(self, Calpha, eps_o, eps_v, mode="fp64", want_pairs=True):
        Calpha, _, eo, ev = _fragment(Calpha, None, eps_o, eps_v)
        naux, nvirt, nocc = Calpha.shape
        ep = np.zeros((nocc, nocc), order="F") if want_pairs else None
        et = np.zeros(1)
        self._chk(lib().rimp2_energy_fitted(self.h, naux, nvirt, nocc, _ptr(Calpha), _ptr(eo), _ptr(ev), _mode(mode),
                                            _ptr(ep), _ptr(et)))
        return float(et[0]), ep

    def energy_fitted_device(self, naux, nvirt, nocc, dB, deo, dev, mode, dpairs, detot):
        self._chk(lib().rimp2_energy_fitted_device(self.h, naux, nvirt, nocc, dB, deo, dev, _mode(mode), dpairs, detot))

    def energy_fitted_device_noalloc(self, naux, nvirt, nocc, dB, deo, dev, mode, dpairs, detot):
        """rimp2_energy_fitted_device with no device allocation: the scratch
        is the device's static array (NOALLOC_SCRATCH_BYTES)."""
        self._chk(lib().rimp2_energy_fitted_device_noalloc(self.h, naux, nvirt, nocc, dB, deo, dev, _mode(mode),
                                                           dpairs, detot))

    def pair_energies(self, eps_o, eps_v, mode="fp64", want_pairs=True):
        eo, ev = _f64(eps_o), _f64(eps_v)
        nocc, nvirt = eo.size, ev.size
        ep = np.zeros((nocc, nocc), order="F") if want_pairs else None
        et = np.zeros(1)
        self._chk(lib().rimp2_pair_energies(self.h, nocc, nvirt, _ptr(eo), _ptr(ev), _mode(mode), _ptr(ep), _ptr(et)))
        return float(et[0]), ep

    def energy(self, Pia, V, eps_o, eps_v, mode="fp64", want_pairs=True):
        if V is None:
            raise ValueError("V is required (use energy_fitted for an already-fitted tensor)")
        Pia, V, eo, ev = _fragment(Pia, V, eps_o, eps_v)
        naux, nvirt, nocc = Pia.shape
        ep = np.zeros((nocc, nocc), order="F") if want_pairs else None
        et = np.zeros(1)
        self._chk(lib().rimp2_energy(self.h, naux, nvirt, nocc, _ptr(Pia), _ptr(V), _ptr(eo), _ptr(ev), _mode(mode),
                                     _ptr(ep), _ptr(et)))
        return float(et[0]), ep

    def get_cholesky(self, n):
        L = np.empty((n, n), order="F")
        self._chk(lib().rimp2_get_cholesky(self.h, _ptr(L)))
        return L

    def get_metric_inverse(self, n):
        L = np.empty((n, n), order="F")
        self._chk(lib().rimp2_get_metric_inverse(self.h, _ptr(L)))
        return L

    def get_B(self, naux, nvirt, nocc):
        B = np.empty((naux, nvirt, nocc), order="F")
        self._chk(lib().rimp2_get_B(self.h, _ptr(B)))
        return B

    # ---- device memory ----
    def alloc(self, nbytes):
        p = _vp()
        self._chk(lib().rimp2_device_alloc(self.h, nbytes, ctypes.byref(p)))
        return p.value

    def free(self, p):
        self._chk(lib().rimp2_device_free(self.h, p))

    def to_device(self, a):
        a = np.ascontiguousarray(a) if a.ndim <= 1 else np.asfortranarray(a)
        p = self.alloc(a.nbytes)
        self._chk(lib().rimp2_copy_h2d(self.h, p, a.ctypes.data, a.nbytes))
        return p

    def from_device(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype, order="F")
        self._chk(lib().rimp2_copy_d2h(self.h, out.ctypes.data, p, out.nbytes))
        return out

    def sync(self):
        self._chk(lib().rimp2_sync(self.h))

    def gen_device(self, seed, naux, nvirt, nocc, dPia, dV, deo, dev, lam=0.0):
        self._chk(lib().rimp2_gen_fragment_device(self.h, seed, naux, nvirt, nocc, lam, dPia, dV, deo, dev))

    def energy_device(self, naux, nvirt, nocc, dPia, dV, deo, dev, mode, dpairs, detot):
        self._chk(lib().rimp2_energy_device(self.h, naux, nvirt, nocc, dPia, dV, deo, dev, _mode(mode), dpairs, detot))

    # ---- timers ----
    def timing(self, on=True):
        self._chk(lib().rimp2_timing_enable(self.h, 1 if on else 0))

    def timer(self, name):
        ms = ctypes.c_double()
        n = _i64()
        self._chk(lib().rimp2_timing_get(self.h, KERNELS[name], ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def timing_reset(self):
        self._chk(lib().rimp2_timing_reset(self.h))

    # ---- batch ----
    def _descs(self, frags):
        arr = (FragDesc * len(frags))()
        keep = []
        for k, (Pia, V, eo, ev) in enumerate(frags):
            # V None: Pia is already fitted (B = Pia), no metric
            Pia, V, eo, ev = _fragment(Pia, V, eo, ev)
            keep += [Pia, V, eo, ev]
            naux, nvirt, nocc = Pia.shape
            arr[k] = FragDesc(nocc, nvirt, naux, _ptr(Pia), _ptr(V), _ptr(eo), _ptr(ev))
        return arr, keep

    def batch(self, frags, mode="fp32", want_pairs=True):
        arr, keep = self._descs(frags)
        ef = np.zeros(len(frags))
        npp = sum(int(f[2].size) ** 2 for f in frags)
        pp = np.zeros(npp) if want_pairs else None
        self._chk(lib().rimp2_batch(self.h, len(frags), arr, _mode(mode), _ptr(ef), _ptr(pp)))
        return ef, pp

    def batch_prepare(self, frags, mode="fp32"):
        arr, keep = self._descs(frags)
        self._chk(lib().rimp2_batch_prepare(self.h, len(frags), arr, _mode(mode)))
        self._batch_n = len(frags)
        self._batch_npp = sum(int(f[2].size) ** 2 for f in frags)

    def batch_prepare_device(self, frags, mode="fp32"):
        """frags: (naux, nvirt, nocc, d_Pia, d_V or None, d_eps_o, d_eps_v)
        with device pointers (ints, from alloc)."""
        arr = (FragDesc * len(frags))()
        for k, (naux, nvirt, nocc, dP, dV, deo, dev) in enumerate(frags):
            arr[k] = FragDesc(nocc, nvirt, naux, ctypes.cast(dP, _dp), None if dV is None else ctypes.cast(dV, _dp),
                              ctypes.cast(deo, _dp), ctypes.cast(dev, _dp))
        self._chk(lib().rimp2_batch_prepare_device(self.h, len(frags), arr, _mode(mode)))
        self._batch_n = len(frags)
        self._batch_npp = sum(int(f[2]) ** 2 for f in frags)

    def batch_run(self, d_e_frag=None):
        self._chk(lib().rimp2_batch_run(self.h, d_e_frag))

    def batch_result(self, want_pairs=True):
        ef = np.zeros(self._batch_n)
        pp = np.zeros(self._batch_npp) if want_pairs else None
        self._chk(lib().rimp2_batch_result(self.h, _ptr(ef), _ptr(pp)))
        return ef, pp

    def batch_tuning(self):
        """The autotuner's last measurements: [(candidate, ms), ...] ([] when
        cached or fixed); candidates as batch_tile."""
        n = ctypes.c_int()
        tiles = (_i64 * 4)()
        ms = (ctypes.c_double * 4)()
        self._chk(lib().rimp2_batch_tuning(self.h, ctypes.byref(n), tiles, ms))
        return [(int(tiles[k]), float(ms[k])) for k in range(n.value)]

    def batch_set_tile(self, tile):
        """Fixes the batch tile candidate (32, 64, -1 = mixed per fragment);
        0 returns to the autotuner."""
        self._chk(lib().rimp2_batch_set_tile(self.h, int(tile)))

    def batch_tile(self):
        """Tile candidate of the prepared batch: 32, 64 or -1 (mixed)."""
        t = _i64()
        self._chk(lib().rimp2_batch_tile(self.h, ctypes.byref(t)))
        return t.value

    # ---- DEC local-basis 4-index arrays (RIMP2.F90:67-170) ----
    def eos_arrays(self, Calpha, EVocc, EVvirt, UoccEOST, UvirtT, UoccT, UvirtEOST):
        """Major Steps 5-8 on the GPU: returns (toccEOS, tvirtEOS, goccEOS,
        gvirtEOS) as Fortran-ordered arrays of the reference's shapes."""
        C = _f64(Calpha)
        if C.ndim != 3:
            raise ValueError("Calpha must be (NBA, nvirt, nocc)")
        nba, nvirt, nocc = C.shape
        UoE, Uv, Uo, UvE = _f64(UoccEOST), _f64(UvirtT), _f64(UoccT), _f64(UvirtEOST)
        noccEOS, nvirtEOS = UoE.shape[1], UvE.shape[1]
        if (UoE.shape[0] != nocc or Uv.shape != (nvirt, nvirt) or Uo.shape != (nocc, nocc)
                or UvE.shape[0] != nvirt):
            raise ValueError("U matrix shapes do not match (nocc, nvirt)")
        eo, ev = _f64(EVocc, (nocc,)), _f64(EVvirt, (nvirt,))
        to = np.zeros((nvirt, noccEOS, nvirt, noccEOS), order="F")
        tv = np.zeros((nvirtEOS, nocc, nvirtEOS, nocc), order="F")
        go = np.zeros((nvirt, noccEOS, nvirt, noccEOS), order="F")
        gv = np.zeros((nvirtEOS, nocc, nvirtEOS, nocc), order="F")
        self._chk(lib().rimp2_eos_arrays(self.h, nvirt, nocc, noccEOS, nvirtEOS, nba, _ptr(C), _ptr(eo), _ptr(ev),
                                         _ptr(UoE), _ptr(Uv), _ptr(Uo), _ptr(UvE), _ptr(to), _ptr(tv), _ptr(go),
                                         _ptr(gv)))
        return to, tv, go, gv

    def eos_arrays_device(self, nvirt, nocc, noccEOS, nvirtEOS, nba, d_in, d_out):
        """d_in: 7 device pointers (Calpha, EVocc, EVvirt, UoccEOST, UvirtT,
        UoccT, UvirtEOST); d_out: 4 (toccEOS, tvirtEOS, goccEOS, gvirtEOS)."""
        self._chk(lib().rimp2_eos_arrays_device(self.h, nvirt, nocc, noccEOS, nvirtEOS, nba, *d_in, *d_out))

    def first_order(self, Calpha, CalphaOcc, CalphaVV, UoccEOST, UvirtT, UoccT, UvirtEOST):
        """First-order intermediates (RIMP2.F90:172-251) on the GPU: returns
        (djik, blad) as Fortran-ordered arrays of the reference's shapes."""
        C, Co, Cvv = _f64(Calpha), _f64(CalphaOcc), _f64(CalphaVV)
        nba, nvirt, nocc = C.shape
        nocctot = Co.shape[2]
        UoE, Uv, Uo, UvE = _f64(UoccEOST), _f64(UvirtT), _f64(UoccT), _f64(UvirtEOST)
        noccEOS, nvirtEOS = UoE.shape[1], UvE.shape[1]
        if Co.shape[:2] != (nba, nocc) or Cvv.shape != (nba, nvirt, nvirt):
            raise ValueError("CalphaOcc must be (NBA, nocc, nocctot) and CalphaVV (NBA, nvirt, nvirt)")
        if (UoE.shape[0] != nocc or Uv.shape != (nvirt, nvirt) or Uo.shape != (nocc, nocc)
                or UvE.shape[0] != nvirt):
            raise ValueError("U matrix shapes do not match (nocc, nvirt)")
        djik = np.zeros((nvirt, noccEOS, noccEOS, nocctot), order="F")
        blad = np.zeros((nvirtEOS, nocc, nvirtEOS, nvirt), order="F")
        self._chk(lib().rimp2_first_order(self.h, nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot, _ptr(C), _ptr(Co),
                                          _ptr(Cvv), _ptr(UoE), _ptr(Uv), _ptr(Uo), _ptr(UvE), _ptr(djik),
                                          _ptr(blad)))
        return djik, blad

    def first_order_device(self, nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot, *ptrs):
        """ptrs: Calpha, CalphaOcc, CalphaVV, UoccEOST, UvirtT, UoccT,
        UvirtEOST, djik, blad as device pointers (asynchronous)."""
        self._chk(lib().rimp2_first_order_device(self.h, nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot, *ptrs))

    def eos_set_mode(self, mode):
        """Amplitude slab precision of eos_arrays: "fp64" (default) or "i8s5"."""
        self._chk(lib().rimp2_eos_set_mode(self.h, _mode(mode)))

    def eos_set_chunk(self, nocc_per_chunk):
        self._chk(lib().rimp2_eos_set_chunk(self.h, int(nocc_per_chunk)))

    def dgemm(self, transa, transb, m, n, k, alpha, d_A, lda, d_B, ldb, beta, d_C, ldc):
        """cublasDgemm_v2 semantics on device pointers (asynchronous)."""
        a, b = ctypes.c_double(alpha), ctypes.c_double(beta)
        self._chk(lib().rimp2_dgemm(self.h, int(transa), int(transb), m, n, k, ctypes.byref(a), d_A, lda, d_B, ldb,
                                    ctypes.byref(b), d_C, ldc))

    # ---- surrogate ----
    def pair_features(self, eps_o, eps_v):
        eo, ev = _f64(eps_o), _f64(eps_v)
        n = eo.size * (eo.size + 1) // 2
        F = np.zeros((n, NFEAT), dtype=np.float32)
        self._chk(lib().rimp2_pair_features(self.h, eo.size, ev.size, _ptr(eo), _ptr(ev), _ptr(F, ctypes.c_float)))
        return F

    def mlp_init(self, seed=0):
        self._chk(lib().rimp2_mlp_init(self.h, seed))

    def mlp_set_norm(self, mean, std, tmean, tstd):
        m = np.ascontiguousarray(mean, dtype=np.float32)
        sd = np.ascontiguousarray(std, dtype=np.float32)
        self._chk(lib().rimp2_mlp_set_norm(self.h, _ptr(m, ctypes.c_float), _ptr(sd, ctypes.c_float), float(tmean),
                                           float(tstd)))

    def mlp_train_step(self, feats, target, lr):
        F = np.ascontiguousarray(feats, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = ctypes.c_float()
        self._chk(lib().rimp2_mlp_train_step(self.h, F.shape[0], _ptr(F, ctypes.c_float), _ptr(t, ctypes.c_float),
                                             lr, ctypes.byref(loss)))
        return loss.value

    def mlp_predict(self, feats):
        F = np.ascontiguousarray(feats, dtype=np.float32)
        out = np.zeros(F.shape[0], dtype=np.float32)
        self._chk(lib().rimp2_mlp_predict(self.h, F.shape[0], _ptr(F, ctypes.c_float), _ptr(out, ctypes.c_float)))
        return out

    def mlp_grad(self, feats, target):
        F = np.ascontiguousarray(feats, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(lib().rimp2_mlp_num_params(), dtype=np.float32)
        loss = ctypes.c_float()
        self._chk(lib().rimp2_mlp_grad(self.h, F.shape[0], _ptr(F, ctypes.c_float), _ptr(t, ctypes.c_float),
                                       ctypes.byref(loss), _ptr(g, ctypes.c_float)))
        return loss.value, g

    def mlp_params(self):
        n = lib().rimp2_mlp_num_params()
        p = np.zeros(n, dtype=np.float32)
        self._chk(lib().rimp2_mlp_get_params(self.h, _ptr(p, ctypes.c_float), n))
        return p

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._chk(lib().rimp2_mlp_set_params(self.h, _ptr(p, ctypes.c_float), p.size))

    def screened_energy(self, eps_o, eps_v, mode, thresh, want_pairs=True):
        eo, ev = _f64(eps_o), _f64(eps_v)
        ep = np.zeros((eo.size, eo.size), order="F") if want_pairs else None
        et = np.zeros(1)
        ns = _i64()
        self._chk(lib().rimp2_screened_energy(self.h, eo.size, ev.size, _ptr(eo), _ptr(ev), _mode(mode), thresh,
                                              _ptr(ep), _ptr(et), ctypes.byref(ns)))
        return float(et[0]), ep, ns.value


def pair_flops(nocc, nvirt, naux):
    """Algorithmic FLOPs of the pair contraction over unique pairs i <= j."""
    return 2.0 * naux * nvirt * nvirt * nocc * (nocc + 1) / 2.0


def eos_chunk(nvirt, nocc, noccEOS, nvirtEOS):
    """Occupied orbitals per amplitude slab chosen by rimp2_eos_arrays when no
    chunk is set: the slab and its tvirtEOS partial transform (plus the
    toccEOS partial when that is not kept whole, above 32 GiB) within 8 GiB."""
    p_all = nvirt * nvirt * noccEOS * nocc * 8 <= (32 << 30)
    per = (nvirt * nvirt * nocc + (0 if p_all else nvirt * nvirt * noccEOS) + nvirtEOS * nvirt * nocc) * 8
    return max(1, min(nocc, int((8 << 30) // per)))


def eos_flops(nvirt, nocc, noccEOS, nvirtEOS, nba):
    """Algorithmic FLOPs of rimp2_eos_arrays in its order of work.  Returns
    (total, amplitude build): the amplitude build forms each unordered
    occupied pair once (nocc (nocc + 1) / 2 pairs, 2 NBA nvirt^2 each), then
    each array's transforms (the index that shrinks most first) and Steps 7-8
    as C3^T C3."""
    nv, no, nE, vE = float(nvirt), float(nocc), float(noccEOS), float(nvirtEOS)
    npair = no * (no + 1) / 2
    ampl = 2 * nba * nv * nv * npair
    tocc = 2 * nv * nv * npair * nE + 2 * nv * nv * nE * no * nE + 4 * nv ** 3 * nE * nE
    tvirt = 2 * vE * nv * nv * npair + 2 * vE * vE * nv * npair + 4 * vE * vE * no ** 3
    gocc = 2 * nba * nv * no * nE + 2 * nba * nv * nv * nE + 2 * (nv * nE) ** 2 * nba
    gvirt = 2 * nba * nv * no * no + 2 * nba * nv * vE * no + 2 * (vE * no) ** 2 * nba
    return ampl + tocc + tvirt + gocc + gvirt, ampl


def first_order_flops(nvirt, nocc, noccEOS, nvirtEOS, nba):
    """FLOPs of rimp2_first_order (nocctot = nocc): the 3-index transforms and
    the two gen4DimFO contractions; the CalphaVV chain (K = nocc) is summed
    with its virtual contraction first, on the nocc columns it uses."""
    nv, no, nE, vE = float(nvirt), float(nocc), float(noccEOS), float(nvirtEOS)
    djik = 2 * nba * nv * no * nE + 2 * nba * nv * nv * nE + 2 * nba * no * no * no + 2 * nba * no * nE * no \
        + 2 * (nv * nE) * (nE * no) * nba
    blad = 2 * nba * nv * no * no + 2 * nba * nv * vE * no + 2 * nba * nv * vE * no + 2 * nba * vE * no * nv \
        + 2 * (vE * no) * (vE * nv) * nba
    return djik + blad


def metric_flops(nocc, nvirt, naux):
    """Triangular metric transform L^{-1} X plus Cholesky (n^3/3) and inverse (n^3/3)."""
    return float(naux) * naux * nvirt * nocc + 2.0 * naux ** 3 / 3.0
