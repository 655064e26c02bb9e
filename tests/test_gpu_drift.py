"""NVE energy drift over 10,000 steps on the inputs whose drift under the
reference algorithm (the oracle, FP64, IMD's operation order) is recorded in
tests/golden/drift_oracle.json (made by tests/golden/drift_oracle.py): the
4,000-atom LJ crystal and a 32,000-atom subset of the 256,000-atom EAM
workload.

Drift = least-squares slope of E_tot per atom (sampled every 10 steps) x 1000
steps / |<E_tot>|.  Bars:
* |drift| <= 1e-6 per 1000 steps in both precisions;
* no worse than the reference's: |drift| <= |drift_ref| + 3 SE, SE the
  slope's standard error from the fit residuals with one independent sample
  per 10 (the energy fluctuations are correlated over ~100 steps);
* the FP64 path follows the reference's trajectory: its drift within half a
  standard error of the reference's, its mean energy within 1e-8.
"""
import json
import os

import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W

from helpers import eam_tables, lj

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REF = json.load(open(os.path.join(HERE, "golden", "drift_oracle.json")))
STEPS = 10000


def _drift(ctx, n):
    e = []
    done = 0
    while done < STEPS:
        k = min(1000, STEPS - done)
        ctx.run(k)
        h = ctx.thermo_history(done, k)[:, 0, :]
        e.extend((h[::10, 1] + h[::10, 2]) / n)
        done += k
    y = np.array(e)
    t = np.arange(0, STEPS, 10, dtype=np.float64)
    slope, icept = np.polyfit(t, y, 1)
    resid = y - (slope * t + icept)
    neff = len(y) / 10.0
    se = np.sqrt(np.mean(resid ** 2) / neff) / np.std(t)
    scale = 1000.0 / abs(y.mean())
    return slope * scale, se * scale, y.mean()


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("name", ["lj4000", "eam32k"])
def test_nve_drift_against_reference(name, precision):
    ref = REF[name]
    if name == "lj4000":
        s = W.lj_crystal()
        ctx = G.pair_context(lj(), [s], precision=precision, ensemble=G.NVE,
                             timestep=0.002, nbl_margin=0.3)
    else:
        s = W.eam_crystal(n_cells=20, temp_k=600.0, seed=1001, displace=0.02)
        ctx = G.eam_context(eam_tables(), [s], precision=precision,
                            ensemble=G.NVE, timestep=W.FS, nbl_margin=0.4)
    with ctx:
        d, se, emean = _drift(ctx, s.n)
    dref = ref["drift_per_1000_steps"]
    assert abs(d) <= 1e-6, d
    assert abs(d) <= abs(dref) + 3 * se, (d, dref, se)
    if precision == G.FP64:
        assert abs(d - dref) <= 0.5 * se, (d, dref, se)
        assert abs(emean - ref["etot_per_atom_mean"]) <= \
            1e-8 * abs(ref["etot_per_atom_mean"])
