"""Write the oracle's tracking results on the benchmark workloads as fixtures.

The GPU parity tests compare the HIP engine with these files at sizes the
oracle needs minutes for; tests/test_libm_gate.py re-runs the oracle on the
config-0 grid and on every 16th config-1 particle and checks that it still
reproduces them bit for bit, so a fixture cannot silently go stale.

  oracle_config0_grid_1000.npz   config 0: the 5 x 205 DA grid on the
                                 998-element ring, Track, 1,000 turns
  oracle_config1_256_1000.npz    config 1: the first 256 Gaussian particles,
                                 Track_spin with 6-D beam-beam, 1,000 turns

Arrays: x0 (input), x (coordinates after tracking or at the loss point),
stable, lost_turn, lost_post.  Usage: python tests/golden/make_oracle_fixtures.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import pyoracle  # noqa: E402
from simtrack_amd import workloads as W  # noqa: E402

NTURN = 1000


def config0_grid():
    gp, ln, x = W.config0()
    return gp, ln.records(), x, False


def config1_sample():
    gp, ln, x = W.config1(n_particles=256)
    return gp, ln.records(), x, True


FIXTURES = {
    "oracle_config0_grid_1000.npz": config0_grid,
    "oracle_config1_256_1000.npz": config1_sample,
}


def load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def main():
    for name, make in FIXTURES.items():
        gp, rec, x, spin = make()
        xf, s, lt, lp = pyoracle.Oracle(gp, rec).track(x.copy(), NTURN, spin=spin)
        np.savez_compressed(os.path.join(HERE, name), x0=x, x=xf, stable=s, lost_turn=lt, lost_post=lp)
        print(name, "survivors", int(s.sum()), "of", len(s), flush=True)


if __name__ == "__main__":
    main()
