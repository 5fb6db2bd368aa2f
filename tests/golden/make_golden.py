"""Writes the golden fixtures of this directory with the CPU oracle.

Each fixture holds, for a small seeded workload, a checksum of the inputs
(to detect a change of the generator) and the interior cv and bv after a few
steps of the oracle.  The workloads are the repository's own seeded
generators (combustioncodegpu_amd/cases.py) at reduced size; the reference
ships no fixtures for this path.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from combustioncodegpu_amd import abi, cases  # noqa: E402

FIXTURES = {
    "c0_32x16_rk3_5steps": (lambda: cases.c0(32, 16), 5),
    "c0_32x16_euler_5steps": (lambda: cases.c0(32, 16, rk_stages=1), 5),
    "c1_10x8x6_ns9_vanalbada_2steps": (lambda: cases.c1(10, 8, 6), 2),
    "c1_10x8x6_ns2_minmod_curved_2steps": (lambda: cases.c1(10, 8, 6, n_spec=2, limiter=0, perturb=0.2), 2),
    "lattice_2x1x1_ns9_2steps": (lambda: cases.lattice([[6, 7], [8], [6]], n_spec=9), 2),
}


def input_digest(case):
    h = hashlib.sha256()
    for b, bv, yk in zip(case.blocks, case.bv0, case.yk0):
        for a in (b.x, b.y, b.z, bv) + ((yk,) if yk is not None else ()):
            h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def run_oracle(case, steps):
    from oracle.pyoracle import Oracle
    o = Oracle(case)
    o.upload_initial()
    o.init_conservative()
    o.step(steps)
    out = {}
    for b in range(len(case.blocks)):
        out[f"cv{b}"] = o.download(b, abi.CV)[..., 2:-2, 2:-2, 2:-2]
        out[f"bv{b}"] = o.download(b, abi.BV)[..., 2:-2, 2:-2, 2:-2]
    o.close()
    return out


def main():
    for name, (make, steps) in FIXTURES.items():
        case = make()
        out = run_oracle(case, steps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), digest=np.array(input_digest(case)),
                            steps=np.array(steps), **out)
        print(name, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
