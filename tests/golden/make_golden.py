"""Writes tests/golden/elements.npz: input/output vectors of the reference's own
element kernels (ComputeMembrane.cpp, ComputeInertial.cpp, ComputeBending.cpp),
compiled from their sources by oracle/Makefile into oracle/_ref/libref_elements.so.

Run in the build container (where /root/reference exists):
    make -C oracle ref && python tests/golden/make_golden.py
Also writes jitter_seed1.npy from std::mt19937 + std::uniform_real_distribution
(libstdc++), the collision jitter stream of boxTriCollision.cpp:649-659, via the
small C++ program make_jitter.cpp next to this script.
"""
import os
import subprocess
import tempfile
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402


def cases(rng, n):
    """Random elements of varied shape, scale and deformation plus special cases."""
    out = []
    for k in range(n):
        s = 10 ** rng.uniform(-3, 0)
        Xa = np.zeros(2)
        Xb = np.array([1.0, 0.0]) + rng.normal(0, 0.2, 2)
        Xc = np.array([0.3, 1.0]) + rng.normal(0, 0.2, 2)
        Xs = [Xa * s, Xb * s, Xc * s]
        if k % 8 == 0:      # rest state
            xs = [np.r_[X, 0.0] for X in Xs]
        elif k % 8 == 1:    # mirrored in-plane (det Fbar < 0 branch of poldec)
            xs = [np.r_[X[0], -X[1], 0.0] for X in Xs]
        else:
            amp = [0.05, 0.3, 1.0][k % 3]
            xs = [np.r_[X, 0.0] + rng.normal(0, amp * s, 3) for X in Xs]
        out.append((xs, Xs))
    return out


def main():
    R = O.ref_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libref_elements.so missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20240531)
    g = np.array([0.0, 0.0, -9.8])
    rec = {k: [] for k in ("tri_x", "tri_X", "tri_P", "tri_Q", "mem_W", "mem_f", "mem_K",
                           "ine_W", "ine_f", "ine_M", "hin_x", "hin_X", "hin_W", "hin_f", "hin_K")}
    for xs, Xs in cases(rng, 96):
        P, Q = O.face_PQ(*xs, *Xs)
        W, f, K = O.membrane(*xs, *Xs, 500.0, 0.3, P, Q, impl=R)
        Wi, fi, Mi = O.inertial(*xs, *Xs, g, 0.1, impl=R)
        rec["tri_x"].append(np.array(xs)); rec["tri_X"].append(np.array(Xs))
        rec["tri_P"].append(P); rec["tri_Q"].append(Q)
        rec["mem_W"].append(W); rec["mem_f"].append(f); rec["mem_K"].append(K)
        rec["ine_W"].append(Wi); rec["ine_f"].append(fi); rec["ine_M"].append(Mi)
    for k in range(96):
        s = 10 ** rng.uniform(-3, 0)
        Xs = [np.zeros(2), np.array([1.0, 0.0]), np.array([0.5, 0.8]), np.array([0.5, -0.8])]
        Xs = [(X + (rng.normal(0, 0.1, 2) if i else 0)) * s for i, X in enumerate(Xs)]
        if k % 6 == 0:      # flat hinge
            xs = [np.r_[X, 0.0] for X in Xs]
        elif k % 6 == 1:    # sharply folded
            xs = [np.r_[X, 0.0] for X in Xs]
            xs[2] = np.r_[Xs[2][0], 0.1 * Xs[2][1], 0.9 * s]
        else:
            amp = [0.05, 0.3, 1.0, 0.1][k % 4]
            xs = [np.r_[X, 0.0] + rng.normal(0, amp * s, 3) for X in Xs]
        W, f, K = O.bending(*xs, *Xs, 1e-5, impl=R)
        rec["hin_x"].append(np.array(xs)); rec["hin_X"].append(np.array(Xs))
        rec["hin_W"].append(W); rec["hin_f"].append(f); rec["hin_K"].append(K)
    np.savez_compressed(os.path.join(HERE, "elements.npz"), **{k: np.array(v) for k, v in rec.items()})
    # jitter stream
    with tempfile.TemporaryDirectory() as tmp:
        exe = os.path.join(tmp, "make_jitter")
        subprocess.check_call(["g++", "-O2", "-o", exe, os.path.join(HERE, "make_jitter.cpp")])
        raw = subprocess.check_output([exe, "3000"])
    np.save(os.path.join(HERE, "jitter_seed1.npy"), np.frombuffer(raw, dtype=np.float64).copy())
    print("wrote elements.npz and jitter_seed1.npy")


if __name__ == "__main__":
    main()
