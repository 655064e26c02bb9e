"""The captured step (IMDG_GRAPH=1): replays of the one-step graph, with
eager work and synchronisations between them, must give bit-identical
states and thermo to the eager launches of the same steps (same kernels in
the same order), across list rebuilds and a list overflow."""
import os

import numpy as np
import pytest

from imd_amd import imdg as G
from imd_amd import workloads as W
from helpers import eam_tables

pytestmark = pytest.mark.gpu


def _ctx(graph, systems, **kw):
    old = os.environ.get("IMDG_GRAPH")
    os.environ["IMDG_GRAPH"] = "1" if graph else "0"
    try:
        return G.eam_context(eam_tables(), systems, precision=G.MIXED,
                             ensemble=G.NVE, **kw)
    finally:
        if old is None:
            del os.environ["IMDG_GRAPH"]
        else:
            os.environ["IMDG_GRAPH"] = old


@pytest.mark.parametrize("nbl_size", [1.1, 0.3])
def test_graph_replays_match_eager(nbl_size):
    # a hot crystal rebuilds its list every few steps; nbl_size 0.3 also
    # forces list overflows, which re-capture the step
    s = W.eam_crystal(n_cells=6, temp_k=1500.0, seed=5, displace=0.05)
    with _ctx(True, [s], nbl_size=nbl_size) as g, \
            _ctx(False, [s], nbl_size=nbl_size) as e:
        done = 0
        for chunk in (1, 7, 13, 5, 20):
            g.run(chunk)
            e.run(chunk)
            # eager work between replays: a readback of every state
            ga, ea = g.atoms(), e.atoms()
            for key in ("x", "p", "f", "epot"):
                assert np.array_equal(ga[key], ea[key]), (done, chunk, key)
            hg = g.thermo_history(done, chunk)
            he = e.thermo_history(done, chunk)
            assert np.array_equal(hg, he), (done, chunk)
            done += chunk
        assert g.info()["nbl_count"] == e.info()["nbl_count"] > 2
