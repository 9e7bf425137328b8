"""Writes tests/golden/ref_scalars.npz: the reference's own r_sqrt and
r_greater (Working/AStrO_r_overloadFunctions.f90:6-43), compiled from source
by `make -C oracle ref-scalars`, on a fixed set of inputs.

Run in a container that has the reference sources; the GPU box and the tests
only read the committed fixture.
"""

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "astro_ref_scalars")


def inputs():
    rng = np.random.default_rng(7)
    # squared lengths as the shell frame and rotation code normalise them
    x = np.concatenate([np.logspace(-12, 8, 241), rng.uniform(0.0, 4.0, 200), [0.0, 1.0, 4.0, 2.0, 0.5]])
    a = rng.normal(size=300) * 10.0 ** rng.integers(-6, 6, 300)
    b = rng.normal(size=300) * 10.0 ** rng.integers(-6, 6, 300)
    eq = rng.normal(size=40)
    ulp_up = eq + np.spacing(np.abs(eq))   # one ulp above
    big = np.array([0.0, 0.0, 1e100, 1.0, 5e-4, 5e-4, 1e-5])
    small = np.array([0.0, -1e-300, 0.0, 1e100, 5e-4, 4.9e-4, 1e-5 * (1 + 1e-12)])
    A = np.concatenate([a, eq, ulp_up, big])
    B = np.concatenate([b, eq, eq, small])
    return x, np.stack([A, B], axis=1)


def main():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "ref-scalars"], cwd=os.path.join(ROOT, "oracle"), check=True)
    x, ab = inputs()
    with tempfile.TemporaryDirectory() as wd:
        fin = os.path.join(wd, "in.f64")
        with open(fin, "wb") as f:
            np.array([len(x)], "<f8").tofile(f)
            x.astype("<f8").tofile(f)
            np.array([len(ab)], "<f8").tofile(f)
            ab.astype("<f8").reshape(-1).tofile(f)  # (a, b) pairs, column-major (2, m) in Fortran
        subprocess.run([EXE, fin, os.path.join(wd, "o_")], check=True)
        s = np.fromfile(os.path.join(wd, "o_sqrt.f64"), "<f8")
        g = np.fromfile(os.path.join(wd, "o_greater.i32"), "<i4")
    np.savez(os.path.join(HERE, "ref_scalars.npz"), x=x, r_sqrt=s, a=ab[:, 0], b=ab[:, 1], r_greater=g)


if __name__ == "__main__":
    main()
