"""Writes the golden fixtures of the reference's single-core acoustic solver.

Inputs: the outputs of `oracle/ref_build.sh run`, i.e. the reference
(singlecore/src/main_cpp/cpu_acoustic.cpp, built from its own sources with
-O3 -ffp-contract=off) run for its hard-wired case: 600x600 periodic,
dx = 0.008, dt = 1e-4, Nt = 60000, Ricker f = 5, t0 = 0.25 on P at (200,200),
receivers at (400,400) on Vx, Vy, P; printed with "% 16.15e".

Output: tests/golden/acoustic_singlecore_<prec>.npz with
  E_head   E(it) for it < 3000 (full resolution)
  E_every  E(it) for it = 0, 50, 100, ... (whole run)
  trace_it step indices 15000, 15025, ... 59975 (the wave reaches the receiver
           after step ~17000)
  Vx, Vy, P  receiver traces at trace_it

and, from `oracle/ref_build.sh run-cpst` (the same program behind
oracle/ref_cli.cpp, which applies the reference's own option parser, run with
-Nt 20000 -cpst 3 and -cpst 6, i.e. the Fast2Sum and 2Sum updates of
singlecore/src/cpp/namespaces.hpp:633-668), for prec f64 and f32:
tests/golden/acoustic_singlecore_<prec>_cpst<c>.npz with
  E_head   E(it) for it < 3000
  E_every  E(it) for it = 0, 50, ..., 19950
  trace_it step indices 15000, 15025, ... 19975
  Vx, Vy, P  receiver traces at trace_it

Usage: python tests/golden/make_golden.py [f64 f32 f16 | cpst]
"""
import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main(precs=("f64", "f32", "f16")):
    for p in precs:
        hits = glob.glob(os.path.join(ROOT, "oracle", "_ref", "run_" + p, "result", "**", "E.txt"), recursive=True)
        if not hits:
            print("no reference output for", p)
            continue
        d = os.path.dirname(hits[0])
        E = np.loadtxt(os.path.join(d, "E.txt"))
        tr = {k: np.loadtxt(os.path.join(d, k + ".txt")) for k in ("Vx", "Vy", "P")}
        if len(E) != 60000:
            print("incomplete run for", p)
            continue
        it = np.arange(15000, 60000, 25)
        out = os.path.join(HERE, "acoustic_singlecore_%s.npz" % p)
        np.savez_compressed(out, E_head=E[:3000], E_every=E[::50], trace_it=it,
                            Vx=tr["Vx"][it], Vy=tr["Vy"][it], P=tr["P"][it])
        print("wrote", out, os.path.getsize(out), "bytes")


def main_cpst(precs=("f64", "f32"), cpsts=("3", "6"), nt=20000):
    for p in precs:
        for c in cpsts:
            d = os.path.join(ROOT, "oracle", "_ref")
            pre = os.path.join(d, "golden_%s_cpst%s_" % (p, c))
            if not os.path.exists(pre + "E.txt"):
                print("no reference output for", p, "cpst", c)
                continue
            E = np.loadtxt(pre + "E.txt")
            tr = {k: np.loadtxt(pre + k + ".txt") for k in ("Vx", "Vy", "P")}
            if len(E) != nt:
                print("incomplete run for", p, "cpst", c)
                continue
            it = np.arange(15000, nt, 25)
            out = os.path.join(HERE, "acoustic_singlecore_%s_cpst%s.npz" % (p, c))
            np.savez_compressed(out, E_head=E[:3000], E_every=E[::50], trace_it=it,
                                Vx=tr["Vx"][it], Vy=tr["Vy"][it], P=tr["P"][it])
            print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    if sys.argv[1:] == ["cpst"]:
        main_cpst()
    else:
        main(sys.argv[1:] or ("f64", "f32", "f16"))
