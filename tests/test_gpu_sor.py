"""SOR induced dipoles on the GPU (scf_algorithm = SOR, and the SOR fallback
after a failed PCG) against the oracle's SOR, which is pinned against the PCG
solution (tests/test_oracle_sor.py)."""
import dataclasses

import numpy as np
import pytest

from forcefieldx_amd import ffi
from forcefieldx_amd.energy import ForceFieldEnergy
from forcefieldx_amd.forcefield import NonbondedParams

pytestmark = pytest.mark.gpu

REL = 1e-8


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def run(oracle, xyz, box, top, p, graph=False, steps=1):
    eo, go, uo, upo = oracle.evaluate(oracle.System(xyz, box, top), p)
    ffe = ForceFieldEnergy(top, box, p, precision=ffi.NB_PREC_FP64)
    try:
        if graph:
            ffe._handle.set_graph(True)
        for _ in range(steps):
            e, g = ffe.energyAndGradient(xyz)
        last = dict(ffe._last)
        u, up = ffe.particleMeshEwald.getInducedDipoles()
        replays = ffe._handle.graph_replays()
    finally:
        ffe.destroy()
    return eo, go, uo, last, g.reshape(-1, 3), u, replays


@pytest.mark.parametrize("kw", [dict(scf_algorithm=1), dict(scf_algorithm=1, polar_sor=0.55),
                                dict(max_scf_iter=2)])
def test_sor_fp64_matches_oracle(oracle, water648, kw):
    xyz, box, top = water648
    p = dataclasses.replace(NonbondedParams(vdw_cutoff=9.0, grid=(24, 24, 24), polar_eps=1e-7), **kw)
    print(f"SOR {kw}", flush=True)
    eo, go, uo, e, g, u, _ = run(oracle, xyz, box, top, p)
    assert eo["status"] == 0
    assert e["pcg_iters"] == eo["pcg_iters"]
    for k in ("pol", "perm", "total"):
        assert abs(e[k] - eo[k]) <= REL * abs(eo[k]), k
    assert rel_err(g, go) <= REL
    assert rel_err(u, uo) <= REL


def test_sor_in_graph_mode_is_not_recorded(oracle, water648):
    xyz, box, top = water648
    p = NonbondedParams(vdw_cutoff=9.0, grid=(24, 24, 24), polar_eps=1e-7, scf_algorithm=1)
    eo, go, uo, e, g, u, replays = run(oracle, xyz, box, top, p, graph=True, steps=3)
    assert replays == 0
    assert rel_err(g, go) <= REL
