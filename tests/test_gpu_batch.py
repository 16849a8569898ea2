"""nb_evaluate_batch over independent small triclinic cells (image sums where
a cell is thinner than twice the cutoff): the batched, graph-replayed results
equal one-at-a-time evaluations, and the FP64 form matches the oracle."""
import numpy as np
import pytest

from forcefieldx_amd import ffi
from forcefieldx_amd.forcefield import NonbondedParams
from forcefieldx_amd.systems import crystal_batch

pytestmark = pytest.mark.gpu

PARAMS = NonbondedParams(ewald_cutoff=6.0, vdw_cutoff=6.0, polar_eps=1e-8)


def handles(systems, precision, graph):
    out = []
    for xyz, box, top, grid in systems:
        h = ffi.Handle(ffi.params_from(PARAMS, grid, precision), top)
        h.set_cell(box)
        h.set_positions(xyz)
        if graph:
            h.set_graph(True)
        out.append(h)
    return out


@pytest.mark.parametrize("precision", [ffi.NB_PREC_FP64, ffi.NB_PREC_MIXED])
def test_batch_equals_single_evaluations(oracle, precision):
    systems = crystal_batch(12, seed=3)
    print(f"batch of {len(systems)} cells, {sum(s[2].n for s in systems)} atoms, precision={precision}", flush=True)
    hb, hs = handles(systems, precision, True), handles(systems, precision, False)
    try:
        for sweep in range(3):
            eb, rc = ffi.evaluate_batch(hb, ffi.NB_GRAD)
            assert rc == ffi.NB_OK
        assert all(h.graph_replays() >= 2 for h in hb)
        for b, (h1, h2) in enumerate(zip(hb, hs)):
            e2, _ = h2.evaluate(ffi.NB_GRAD)
            assert abs(eb[b]["total"] - e2["total"]) <= 1e-12 * max(abs(e2["perm"]), abs(e2["vdw"])), b
            g1, g2 = h1.gradient(), h2.gradient()
            assert np.abs(g1 - g2).max() <= 1e-12 * np.abs(g2).max(), b
        if precision == ffi.NB_PREC_FP64:
            for b in (0, 5, 11):
                xyz, box, top, grid = systems[b]
                import dataclasses
                eo, go, uo, upo = oracle.evaluate(oracle.System(xyz, box, top), dataclasses.replace(PARAMS, grid=grid))
                scale = max(abs(eo[k]) for k in ("perm", "pol", "vdw"))
                assert abs(eb[b]["total"] - eo["total"]) <= 1e-8 * scale, b
                g = hb[b].gradient().reshape(-1, 3)
                assert np.linalg.norm(g - go) <= 1e-8 * np.linalg.norm(go), b
    finally:
        for h in hb + hs:
            h.close()
