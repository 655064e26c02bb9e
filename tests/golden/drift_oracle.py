"""NVE energy drift of the reference algorithm (the oracle: IMD's
neighbour-list pair/EAM path restated in FP64 C, serial, IMD's operation
order) on two inputs, written to drift_oracle.json next to this script.

* lj4000: the 4,000-atom fcc LJ crystal (10^3 cells, a = 1.5496 sigma, LJ
  table r^2 0.25-6.25 with 2,000 points, kT 0.7 Maxwell from seed 1000),
  dt 0.002, skin 0.3, 10,000 steps;
* eam32k: a 32,000-atom subset of the 256,000-atom EAM workload (20^3 fcc
  cells, a = 3.615 A, displacements N(0, 0.02 A), 600 K, seed 1001,
  Johnson-form tables 5,000 points r_c 5.0 A), dt 1 fs, skin 0.4 A, 10,000
  steps.

E_tot per atom = (Epot + Ekin) / N every 10 steps; drift = least-squares
slope x 1000 steps / |<E_tot>| (the protocol of the GPU drift runs, which
use the same inputs: bench.py --drift N --drift-config lj4000|eam32k).

Run: python tests/golden/drift_oracle.py  (several minutes on one core).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from imd_amd import workloads as W  # noqa: E402


def drift_of(th, n, every=10):
    e = (th[::every, 0] + th[::every, 1]) / n
    t = np.arange(0, th.shape[0], every, dtype=np.float64)
    slope, icept = np.polyfit(t, e, 1)
    resid = e - (slope * t + icept)
    return {"drift_per_1000_steps": float(slope * 1000 / abs(e.mean())),
            "etot_per_atom_mean": float(e.mean()),
            "etot_per_atom_span": float(e.max() - e.min()),
            "etot_resid_rms": float(np.sqrt(np.mean(resid ** 2)))}


def main(steps=10000):
    out = {}
    s = W.lj_crystal()
    orc = O.OracleSystem(s, O.OracleTable.from_table(W.lj_table()), None,
                         None, nbl_margin=0.3, timestep=0.002, ntypes=1)
    t0 = time.time()
    th = orc.run(steps)
    out["lj4000"] = dict(drift_of(th, s.n), natoms=s.n, steps=steps,
                         seconds=time.time() - t0)
    eam = W.johnson_eam()
    s = W.eam_crystal(n_cells=20, temp_k=600.0, seed=1001, displace=0.02)
    orc = O.OracleSystem(s, O.OracleTable.from_table(eam.pair),
                         O.OracleTable.from_table(eam.rho),
                         O.OracleTable.from_table(eam.embed, radial=False),
                         nbl_margin=0.4, timestep=W.FS, ntypes=1)
    t0 = time.time()
    th = orc.run(steps)
    out["eam32k"] = dict(drift_of(th, s.n), natoms=s.n, steps=steps,
                         seconds=time.time() - t0)
    with open(os.path.join(HERE, "drift_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
