"""CPU FP64 oracle for the RI-MP2 hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only to compare the GPU results against (or as the timed
CPU baseline); the product (rimp2_gpu_amd) never imports it.

Functions follow the reference (januseriksen/rimp2_gpu) line by line as cited
in rimp2_oracle.c; the MLP oracle follows mlp_oracle.c.

Parity pins (tests/test_oracle.py):
  toccA/tvirtA/toccB/tvirtB/transalpha/gocc/gvirt/steps5to8
      -> closed-form values of the reference fixtures (generate.F90:77-128)
         at the reference's 1e-10 tolerance, and an independent numpy einsum
         on random non-constant inputs (index order / transposes).
  pair_energies -> numpy brute force + math.fsum, nocc=nvirt=1 closed form,
         s^4 scaling, orbital-permutation invariance, eps_ij <= 0.
  cholesky / trsm -> scipy.linalg (LAPACK dpotrf / dtrtrs), L L^T = V,
         B^T B = X^T V^-1 X, V = I gives B = X.
  gen4dimfo -> numpy einsum.
  first_order -> an independent numpy einsum that spells the reference's
         K = ldb = nocc product (RIMP2.F90:231-235) out as a reinterpretation
         of UvirtT's storage, closed forms on constant inputs, and a check
         that the quirk changes the result (so a "corrected" call fails).  The
         reference's generator writes no first-order records
         (generate.F90:130), so no reference fixture reaches this block.
  mlp_* -> finite differences and a closed-form GELU (erf form).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "librimp2_oracle.so")
LIB_PATH_V4 = os.path.join(_HERE, "build", "librimp2_oracle_v4.so")
_lib = None
_lib_path = None


def cpu_info():
    """Model name and the x86-64 ISA level of the running CPU."""
    model, flags = "unknown", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    v4 = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    return model, ("x86-64-v4" if v4 else "x86-64-v3")

_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    """Loads the oracle built for the running CPU's ISA level (AVX-512 where
    the CPU has it, AVX2 + FMA otherwise)."""
    global _lib, _lib_path
    if _lib is None:
        if not (os.path.exists(LIB_PATH) and os.path.exists(LIB_PATH_V4)):
            build()
        _lib_path = LIB_PATH_V4 if cpu_info()[1] == "x86-64-v4" else LIB_PATH
        L = ctypes.CDLL(_lib_path)
        ci = ctypes.c_int
        sig = {
            "oracle_toccA": [ci, ci, ci, ci, _dp, _dp, _dp, _dp, _dp],
            "oracle_tvirtA": [ci, ci, ci, ci, _dp, _dp, _dp, _dp, _dp],
            "oracle_toccB": [ci, ci, _dp, _dp, _dp],
            "oracle_tvirtB": [ci, ci, _dp, _dp, _dp],
            "oracle_transalpha1": [ci, ci, ci, _dp, _dp, _dp],
            "oracle_transalpha2": [ci, ci, ci, ci, _dp, _dp, _dp],
            "oracle_gocc": [ci, ci, ci, _dp, _dp],
            "oracle_gvirt": [ci, ci, ci, _dp, _dp],
            "oracle_gen4dimfo": [ci, _dp, ci, ci, _dp, ci, ci, _dp],
            "oracle_first_order": [ci, ci, ci, ci, ci, ci] + [_dp] * 9,
            "oracle_steps5to8": [ci, ci, ci, ci, ci, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp],
            "oracle_pair_energies": [_i64, _i64, _i64, _dp, _dp, _dp, _dp, _dp],
            "oracle_pair_energies_subset": [_i64, _i64, _dp, _dp, _dp, _i64, ctypes.POINTER(_i64),
                                            ctypes.POINTER(_i64), _dp],
            "oracle_pair_V": [_i64, _i64, _dp, _i64, _i64, _dp],
            "oracle_cholesky": [_i64, _dp, _dp],
            "oracle_trsm_lower": [_i64, _i64, _dp, _dp, _dp],
            "oracle_metric_transform": [_i64, _i64, _dp, _dp, _dp],
            "oracle_num_threads": [],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int if name in ("oracle_cholesky", "oracle_metric_transform",
                                                 "oracle_num_threads") else None
        for name, (res, args) in MLP_SIG.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
        _lib = L
    return _lib


MLP_SIG = {
    "oracle_mlp_loss_grad": (ctypes.c_double, [_i64, _fp, _fp, _fp, _fp, ctypes.c_int, _dp, _dp]),
    "oracle_gelu": (ctypes.c_double, [ctypes.c_double]),
    "oracle_gelu_d": (ctypes.c_double, [ctypes.c_double]),
    "oracle_bf16": (ctypes.c_float, [ctypes.c_float]),
    "oracle_mlp_num_params": (_i64, []),
}


def mlp_loss_grad(feats, target, params, norm, emulate_bf16=False, want_grad=True):
    """FP64 loss, gradient and raw predictions of the screening MLP."""
    F = np.ascontiguousarray(feats, dtype=np.float32)
    t = None if target is None else np.ascontiguousarray(target, dtype=np.float32)
    P = np.ascontiguousarray(params, dtype=np.float32)
    N = np.ascontiguousarray(norm, dtype=np.float32)
    n = F.shape[0]
    g = np.zeros(lib().oracle_mlp_num_params()) if want_grad else None
    pred = np.zeros(n)
    fp = ctypes.POINTER(ctypes.c_float)
    loss = lib().oracle_mlp_loss_grad(n, F.ctypes.data_as(fp), None if t is None else t.ctypes.data_as(fp),
                                      P.ctypes.data_as(fp), N.ctypes.data_as(fp), 1 if emulate_bf16 else 0,
                                      None if g is None else g.ctypes.data_as(_dp), pred.ctypes.data_as(_dp))
    return loss, g, pred


def gelu(x):
    return lib().oracle_gelu(x)


def gelu_d(x):
    return lib().oracle_gelu_d(x)


def bf16(x):
    return lib().oracle_bf16(x)


def _p(a):
    return a.ctypes.data_as(_dp)


def _F(a):
    return np.asfortranarray(a, dtype=np.float64)


def toccA(Calpha, EVocc, EVvirt, UoccEOST):
    nba, nvirt, nocc = Calpha.shape
    noccEOS = UoccEOST.shape[1]
    t = np.zeros((nocc, noccEOS, nvirt, nvirt), order="F")
    Calpha, EVocc, EVvirt, UoccEOST = _F(Calpha), _F(EVocc), _F(EVvirt), _F(UoccEOST)
    lib().oracle_toccA(nvirt, nocc, noccEOS, nba, _p(Calpha), _p(EVocc), _p(EVvirt), _p(t), _p(UoccEOST))
    return t


def tvirtA(Calpha, EVocc, EVvirt, UvirtEOST):
    nba, nvirt, nocc = Calpha.shape
    nvirtEOS = UvirtEOST.shape[1]
    t = np.zeros((nocc, nocc, nvirtEOS, nvirt), order="F")
    Calpha, EVocc, EVvirt, UvirtEOST = _F(Calpha), _F(EVocc), _F(EVvirt), _F(UvirtEOST)
    lib().oracle_tvirtA(nvirt, nocc, nvirtEOS, nba, _p(Calpha), _p(EVocc), _p(EVvirt), _p(t), _p(UvirtEOST))
    return t


def gocc(C3):
    nba, nv, no = C3.shape
    g = np.zeros((nv, no, nv, no), order="F")
    C3 = _F(C3)
    lib().oracle_gocc(nv, no, nba, _p(C3), _p(g))
    return g


def gvirt(C3):
    nba, nv, no = C3.shape
    g = np.zeros((nv, no, nv, no), order="F")
    C3 = _F(C3)
    lib().oracle_gvirt(nv, no, nba, _p(C3), _p(g))
    return g


def gen4dimfo(C3, C4):
    nba, n1, n2 = C3.shape
    _, n3, n4 = C4.shape
    out = np.zeros((n1, n2, n3, n4), order="F")
    C3, C4 = _F(C3), _F(C4)
    lib().oracle_gen4dimfo(nba, _p(C3), n1, n2, _p(C4), n3, n4, _p(out))
    return out


def steps5to8(Calpha, EVocc, EVvirt, UoccEOST, UvirtT, UoccT, UvirtEOST):
    nba, nvirt, nocc = Calpha.shape
    noccEOS, nvirtEOS = UoccEOST.shape[1], UvirtEOST.shape[1]
    args = [_F(x) for x in (Calpha, EVocc, EVvirt, UoccEOST, UvirtT, UoccT, UvirtEOST)]
    toccEOS = np.zeros((nvirt, noccEOS, nvirt, noccEOS), order="F")
    tvirtEOS = np.zeros((nvirtEOS, nocc, nvirtEOS, nocc), order="F")
    goccEOS = np.zeros((nvirt, noccEOS, nvirt, noccEOS), order="F")
    gvirtEOS = np.zeros((nvirtEOS, nocc, nvirtEOS, nocc), order="F")
    lib().oracle_steps5to8(nvirt, nocc, noccEOS, nvirtEOS, nba, *[_p(a) for a in args], _p(toccEOS), _p(tvirtEOS),
                           _p(goccEOS), _p(gvirtEOS))
    return toccEOS, tvirtEOS, goccEOS, gvirtEOS


def first_order(Calpha, CalphaOcc, CalphaVV, UoccEOST, UvirtT, UoccT, UvirtEOST):
    """RIMP2.F90:172-251: returns (djik, blad) with the reference's own call
    arguments (nocctot = nocc required)."""
    nba, nvirt, nocc = Calpha.shape
    nocctot = CalphaOcc.shape[2]
    if nocctot != nocc:
        raise ValueError("the reference's first-order block is in bounds only for nocctot == nocc")
    noccEOS, nvirtEOS = UoccEOST.shape[1], UvirtEOST.shape[1]
    args = [_F(x) for x in (Calpha, CalphaOcc, CalphaVV, UoccEOST, UvirtT, UoccT, UvirtEOST)]
    djik = np.zeros((nvirt, noccEOS, noccEOS, nocctot), order="F")
    blad = np.zeros((nvirtEOS, nocc, nvirtEOS, nvirt), order="F")
    lib().oracle_first_order(nvirt, nocc, noccEOS, nvirtEOS, nba, nocctot, *[_p(a) for a in args], _p(djik),
                             _p(blad))
    return djik, blad


def pair_energies(B, eps_o, eps_v):
    """Canonical RI-MP2: returns (E, eps_pair[nocc,nocc])."""
    naux, nvirt, nocc = B.shape
    B, eo, ev = _F(B), _F(eps_o), _F(eps_v)
    ep = np.zeros((nocc, nocc), order="F")
    et = np.zeros(1)
    lib().oracle_pair_energies(nocc, nvirt, naux, _p(B), _p(eo), _p(ev), _p(ep), _p(et))
    return float(et[0]), ep


def pair_energies_subset(B, eps_o, eps_v, pi, pj):
    naux, nvirt, nocc = B.shape
    B, eo, ev = _F(B), _F(eps_o), _F(eps_v)
    pi = np.ascontiguousarray(pi, dtype=np.int64)
    pj = np.ascontiguousarray(pj, dtype=np.int64)
    out = np.zeros(pi.size)
    P = ctypes.POINTER(_i64)
    lib().oracle_pair_energies_subset(nvirt, naux, _p(B), _p(eo), _p(ev), pi.size, pi.ctypes.data_as(P),
                                      pj.ctypes.data_as(P), _p(out))
    return out


def pair_V(B, i, j):
    naux, nvirt, nocc = B.shape
    B = _F(B)
    V = np.zeros((nvirt, nvirt), order="F")
    lib().oracle_pair_V(nvirt, naux, _p(B), i, j, _p(V))
    return V


def cholesky(V):
    n = V.shape[0]
    V = _F(V)
    L = np.zeros((n, n), order="F")
    st = lib().oracle_cholesky(n, _p(V), _p(L))
    if st:
        raise np.linalg.LinAlgError(f"not positive definite at pivot {st}")
    return L


def trsm_lower(L, X):
    n = L.shape[0]
    X2 = _F(X).reshape(n, -1, order="F")
    B = np.zeros_like(X2, order="F")
    lib().oracle_trsm_lower(n, X2.shape[1], _p(_F(L)), _p(X2), _p(B))
    return B.reshape(X.shape, order="F")


def metric_transform(V, X):
    """B = L^{-1} X with V = L L^T (X: [naux, ...])."""
    n = V.shape[0]
    X2 = _F(X).reshape(n, -1, order="F")
    B = np.zeros_like(X2, order="F")
    st = lib().oracle_metric_transform(n, X2.shape[1], _p(_F(V)), _p(X2), _p(B))
    if st:
        raise np.linalg.LinAlgError(f"not positive definite at pivot {st}")
    return B.reshape(X.shape, order="F")


def num_threads():
    return lib().oracle_num_threads()


def build_flavour():
    """The -march level of the loaded oracle build."""
    lib()
    return "x86-64-v4" if _lib_path == LIB_PATH_V4 else "x86-64-v3"
