"""The GPU's Nose-Hoover step against the closed-form free-particle
recurrence (src/imd_integrate.c:907-908, 1020-1027, 1103, 1140-1141): no
partner inside the cutoff, eta0 != 0 and a finite tau, fused on the device
for 50 steps in both precisions (forces vanish, so the FP64 integration
decides everything), and the same run through the unfused entry points."""
import numpy as np
import pytest

from imd_amd import imdg as G

from test_oracle_integrator import (NH_FREE, nh_free_case, nose_hoover_free,
                                    short_table)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [G.FP64, G.MIXED])
@pytest.mark.parametrize("fused", [True, False])
def test_gpu_nose_hoover_free_particles(precision, fused):
    s, T = nh_free_case()
    s.temperature = T
    dt, eta0 = NH_FREE["dt"], NH_FREE["eta0"]
    isq = 1.0 / NH_FREE["tau"] ** 2
    n = 50
    x, p, ek, etas = nose_hoover_free(s, dt, T, eta0, isq, n)
    ctx = G.Context(precision=precision, ensemble=G.NVT, ntypes=1,
                    timestep=dt, nbl_margin=0.4)
    with ctx as gpu:
        gpu.set_tables(G.TableArrays.from_format2(short_table()), None, None)
        gpu.set_systems([s], eta=eta0, isq_tau_eta=isq)
        if fused:
            gpu.run(n)
            hist = gpu.thermo_history(0, n)[:, 0, :]
            ek_g, eta_g = hist[:, 2], hist[:, 4]
        else:
            ek_g, eta_g = [], []
            for st in range(n):
                gpu.calc_forces(st)
                gpu.move_atoms()
                gpu.check_nblist()
                th = gpu.thermo()
                ek_g.append(th["ekin"])
                eta_g.append(th["eta"])
            ek_g, eta_g = np.array(ek_g), np.array(eta_g)
        a = gpu.atoms()
    assert np.allclose(eta_g, etas, rtol=1e-12, atol=1e-14)
    assert np.allclose(ek_g, ek, rtol=1e-12, atol=0)
    assert np.allclose(a["p"], p, rtol=1e-12, atol=1e-16)
    L = 30.0
    d = a["x"] - x
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-11
    assert np.all(a["f"] == 0.0)
