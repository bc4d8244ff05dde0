"""Writes ensemble_128_64.json: the 64 members of the batched-ensemble
workload (128^3, two fields).  g^2 is log-uniform in [1e4, 1e6] and phi0
uniform in [0.8, 1.2], drawn from numpy's PCG64 with seed 20241203.  Member m
uses Park-Miller seed m + 1: seed 0 is the reference's debug mode in which
every draw returns 0.33 (initialize.cpp:15), so the members' seeds 0..63 are
shifted by one to give 64 independent streams."""
import json
import os

import numpy as np

PRNG_SEED = 20241203


def main():
    rng = np.random.default_rng(PRNG_SEED)
    log_g2 = rng.uniform(4.0, 6.0, 64)
    phi0 = rng.uniform(0.8, 1.2, 64)
    members = [{"member": m, "seed": m + 1, "g2": float(10 ** log_g2[m]), "phi0": float(phi0[m])}
               for m in range(64)]
    out = {"prng": "numpy PCG64", "prng_seed": PRNG_SEED, "N": 128, "members": members}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ensemble_128_64.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print(path)


if __name__ == "__main__":
    main()
