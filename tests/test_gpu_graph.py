"""Graph mode (nb_set_graph): recorded evaluations replayed with the PCG under
device-side control give the results of the regular path, across a
trajectory of small perturbations, list rebuilds and flag changes."""
import dataclasses

import numpy as np
import pytest

from forcefieldx_amd import ffi
from forcefieldx_amd.forcefield import NonbondedParams
from forcefieldx_amd.systems import chain_in_water, water_box

pytestmark = pytest.mark.gpu

TOL = 1e-12   # same kernels; only the order of the FP64 energy and gradient atomics differs
# The mixed form's grids are fixed-point sums, independent of the order of the
# spread's atomics, so its replays hold the same bar
TOL_MIXED = 1e-12


def make(top, box, p, precision, graph):
    h = ffi.Handle(ffi.params_from(p, p.resolved_grid(box), precision), top)
    h.set_cell(box)
    if graph:
        h.set_graph(True)
    return h


def run(h, x, flags):
    h.set_positions(x)
    e, rc = h.evaluate(flags)
    out = dict(e=e, g=h.gradient().reshape(-1).copy())
    if flags & ffi.NB_VIRIAL:
        out["W"] = h.virial()
    if flags & ffi.NB_DUDL:
        out["l"] = h.dudx_dl().reshape(-1).copy()
    return out


def close(a, b, tol=TOL):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision", [ffi.NB_PREC_FP64, ffi.NB_PREC_MIXED])
def test_graph_replay_matches_regular_path(water648, precision):
    xyz, box, top = water648
    p = NonbondedParams(vdw_cutoff=9.0, ewald_cutoff=7.0, grid=(24, 24, 24), polar_eps=1e-6)
    ref, gr = make(top, box, p, precision, False), make(top, box, p, precision, True)
    tol = TOL if precision == ffi.NB_PREC_FP64 else TOL_MIXED
    rng = np.random.default_rng(11)
    x = xyz.reshape(-1).copy()
    try:
        for step in range(24):
            print(f"graph precision={precision} step={step}", flush=True)
            if step == 16:
                x = x + rng.normal(0.0, 0.7, x.shape)   # forces a list rebuild
            else:
                x = x + rng.normal(0.0, 0.01, x.shape)
            flags = ffi.NB_GRAD | (ffi.NB_VIRIAL if step >= 20 else 0)
            a, b = run(ref, x, flags), run(gr, x, flags)
            for k in ("total", "perm", "pol", "vdw"):
                assert abs(a["e"][k] - b["e"][k]) <= tol * abs(a["e"][k]), (step, k)
            assert a["e"]["pcg_iters"] == b["e"]["pcg_iters"]
            assert close(b["g"], a["g"], tol)
            if "W" in a:
                assert close(b["W"], a["W"], tol)
        assert gr.graph_replays() >= 18
    finally:
        ref.close()
        gr.close()


def test_graph_replay_with_refinement_corrections(water648):
    """Mixed form at 1e-8 D, where the FP32-operator solve needs FP64
    refinement corrections: the graph records the corrections the regular
    path took, replays run them under device-side control, and a replay whose
    last residual fails is repeated on the regular path; every evaluation
    equals the regular path's."""
    xyz, box, top = water648
    p = NonbondedParams(vdw_cutoff=9.0, ewald_cutoff=7.0, grid=(24, 24, 24), polar_eps=1e-8)
    ref, gr = make(top, box, p, ffi.NB_PREC_MIXED, False), make(top, box, p, ffi.NB_PREC_MIXED, True)
    rng = np.random.default_rng(5)
    x = xyz.reshape(-1).copy()
    try:
        for step in range(16):
            print(f"graph refinement step={step}", flush=True)
            x = x + rng.normal(0.0, 0.01, x.shape)
            a, b = run(ref, x, ffi.NB_GRAD), run(gr, x, ffi.NB_GRAD)
            assert a["e"]["pcg_eps"] < 1e-8 and b["e"]["pcg_eps"] < 1e-8
            for k in ("total", "perm", "pol", "vdw"):
                assert abs(a["e"][k] - b["e"][k]) <= TOL_MIXED * abs(a["e"][k]), (step, k)
            assert a["e"]["pcg_iters"] == b["e"]["pcg_iters"]
            assert close(b["g"], a["g"], TOL_MIXED)
        print(f"replays {gr.graph_replays()} of 16", flush=True)
        assert gr.graph_replays() >= 8
    finally:
        ref.close()
        gr.close()


def test_graph_replay_with_ost_children(oracle):
    """OST handles record one graph per phase; dU/dL and dU/dX/dL included."""
    xyz, box, top = chain_in_water(24, 180, 18.0, seed=5, n_soft=6)
    p = NonbondedParams(vdw_cutoff=7.0, ewald_cutoff=7.0, grid=(24, 24, 24), polar_eps=1e-8, lambda_value=0.7,
                        elec_lambda=True)
    ref, gr = make(top, box, p, ffi.NB_PREC_FP64, False), make(top, box, p, ffi.NB_PREC_FP64, True)
    rng = np.random.default_rng(5)
    x = xyz.reshape(-1).copy()
    flags = ffi.NB_GRAD | ffi.NB_DUDL
    try:
        for step in range(6):
            print(f"graph OST step={step}", flush=True)
            x = x + rng.normal(0.0, 0.01, x.shape)
            a, b = run(ref, x, flags), run(gr, x, flags)
            for k in ("total", "dudl", "d2udl2"):
                assert abs(a["e"][k] - b["e"][k]) <= TOL * max(abs(a["e"][k]), 1.0), (step, k)
            assert close(b["g"], a["g"])
            assert close(b["l"], a["l"])
        assert gr.graph_replays() >= 4
        # the replayed result is the oracle's
        eo, go, lgo = oracle.evaluate_lambda(oracle.System(x.reshape(-1, 3), box, top), p)
        assert abs(b["e"]["total"] - eo["total"]) <= 1e-8 * abs(eo["total"])
        assert np.linalg.norm(b["g"] - go.reshape(-1)) <= 1e-8 * np.linalg.norm(go)
    finally:
        ref.close()
        gr.close()
