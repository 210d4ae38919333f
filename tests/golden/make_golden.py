"""Writes the committed golden fixtures under tests/golden/.

* generator_pin.npz : sha256 of the seeded generator output (seed 42,
  naux 48, nvirt 10, nocc 3) and a few raw values -- pins the input synthesis
  bit for bit across machines.
* frag_<seed>_<naux>_<nvirt>_<nocc>.npz : small seeded fragments with the
  oracle's FP64 reference results (B = L^-1 X, E, eps_pair), so GPU tests can
  check against committed numbers without recomputing.

Run:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
import rimp2_gpu_amd as r  # noqa: E402

FRAGS = [(1, 64, 16, 3), (2, 96, 40, 5), (3, 128, 70, 4), (4, 512, 128, 16)]


def main():
    P, V, eo, ev = r.gen_fragment(42, 48, 10, 3)
    h = hashlib.sha256(np.concatenate([P.ravel(order="F"), V.ravel(order="F"), eo, ev]).tobytes()).hexdigest()
    np.savez(os.path.join(HERE, "generator_pin.npz"), sha256=h, P_head=P[:4, :3, 0])
    for seed, naux, nvirt, nocc in FRAGS:
        P, V, eo, ev = r.gen_fragment(seed, naux, nvirt, nocc)
        B = oracle.metric_transform(V, P)
        E, eps = oracle.pair_energies(B, eo, ev)
        name = f"frag_{seed}_{naux}_{nvirt}_{nocc}.npz"
        # inputs are regenerated from the seed (pinned above); store a checksum
        hin = hashlib.sha256(np.concatenate([P.ravel(order="F"), V.ravel(order="F"), eo, ev]).tobytes()).hexdigest()
        np.savez_compressed(os.path.join(HERE, name), seed=seed, naux=naux, nvirt=nvirt, nocc=nocc,
                            inputs_sha256=hin, E=E, eps_pair=eps, B_head=B[:8, :4, 0])
        print(name, E)


if __name__ == "__main__":
    main()
