"""NB_FINITE_PCG on the GPU: the gradient of the dipoles the PCG returned
(iteration count fixed) against the oracle's, whose adjoint terms are pinned
by central differences (tests/test_oracle_finite_pcg.py), and against central
differences of the GPU energy itself.  FP64 at 1e-8 of the oracle; the mixed
form within 1e-5 of FP64."""
import dataclasses

import numpy as np
import pytest

from forcefieldx_amd import ffi
from forcefieldx_amd.forcefield import NonbondedParams
from forcefieldx_amd.systems import chain_in_water, small_crystal, water_box

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def gpu(xyz, box, top, p, flags, precision=ffi.NB_PREC_FP64):
    h = ffi.Handle(ffi.params_from(p, p.resolved_grid(box) if box is not None else (16, 16, 16), precision), top)
    try:
        h.set_cell(box)
        h.set_positions(xyz.reshape(-1))
        e, rc = h.evaluate(flags)
        return e, h.gradient().reshape(-1, 3)
    finally:
        h.close()


def cases():
    xyz, box, top = water_box(64, 12.42, seed=2)
    yield "water", xyz, box, top, NonbondedParams(vdw_cutoff=6.0, ewald_cutoff=6.0, grid=(16, 16, 16), polar_eps=0.05)
    xyz, box, top = chain_in_water(24, 60, 13.0, seed=5, n_soft=0)
    pol = top.polarity.copy()
    pol[::4] = 0.0
    yield "chain, unpolarizable atoms", xyz, box, dataclasses.replace(top, polarity=pol), \
        NonbondedParams(vdw_cutoff=6.0, ewald_cutoff=6.0, grid=(16, 16, 16), polar_eps=0.02)
    cell = [9.5, 10.2, 11.0, 80.0, 95.0, 103.0]
    xyz, box, top = small_crystal(50, cell, seed=2)
    yield "thin triclinic", xyz, box, top, NonbondedParams(ewald_cutoff=6.0, vdw_cutoff=6.0, grid=(16, 16, 16),
                                                           polar_eps=0.02)
    xyz, box, top = chain_in_water(16, 20, 13.0, seed=3, n_soft=0)
    yield "aperiodic", xyz, None, top, NonbondedParams(vdw_cutoff=50.0, ewald_cutoff=50.0, polar_eps=0.02)


CASES = list(cases())


@pytest.mark.parametrize("case", range(len(CASES)), ids=[c[0] for c in CASES])
def test_finite_gradient_matches_oracle(oracle, case):
    name, xyz, box, top, p = CASES[case]
    print(f"finite PCG: {name}, n={top.n}", flush=True)
    s = oracle.System(xyz.reshape(-1, 3), box, top)
    eo, go, _, _ = oracle.evaluate(s, p, finite_pcg=True)
    gs_o = oracle.evaluate(s, p)[1]
    e, g = gpu(xyz, box, top, p, ffi.NB_GRAD | ffi.NB_FINITE_PCG)
    assert e["pcg_iters"] == eo["pcg_iters"]
    assert abs(e["total"] - eo["total"]) <= 1e-8 * abs(eo["total"])
    assert rel(g, go) <= 1e-8, rel(g, go)
    assert rel(gs_o, go) >= 1e-5   # the case is one where the truncation matters


def test_finite_gradient_central_differences():
    """GPU energy differences at h = 1e-5 A (iteration count unchanged)."""
    name, xyz, box, top, p = CASES[0]
    p = dataclasses.replace(p, do_vdw=False)
    xyz = xyz.reshape(-1, 3)
    e0, g = gpu(xyz, box, top, p, ffi.NB_GRAD | ffi.NB_FINITE_PCG)
    h = 1e-5
    for i in (0, 17, 101):
        for d in range(3):
            xp = xyz.copy(); xp[i, d] += h
            xm = xyz.copy(); xm[i, d] -= h
            ep, _ = gpu(xp, box, top, p, 0)
            em, _ = gpu(xm, box, top, p, 0)
            assert ep["pcg_iters"] == e0["pcg_iters"] == em["pcg_iters"]
            fd = (ep["total"] - em["total"]) / (2 * h)
            assert abs(g[i, d] - fd) <= 1e-6 * np.abs(g).max(), (i, d, g[i, d], fd)


def test_finite_mixed_against_fp64():
    name, xyz, box, top, p = CASES[0]
    e64, g64 = gpu(xyz, box, top, p, ffi.NB_GRAD | ffi.NB_FINITE_PCG)
    emx, gmx = gpu(xyz, box, top, p, ffi.NB_GRAD | ffi.NB_FINITE_PCG, ffi.NB_PREC_MIXED)
    assert rel(gmx, g64) <= 1e-5


def test_converged_finite_equals_stationary():
    xyz, box, top = water_box(64, 12.42, seed=2)
    p = NonbondedParams(vdw_cutoff=6.0, ewald_cutoff=6.0, grid=(16, 16, 16), polar_eps=1e-10)
    _, gs = gpu(xyz, box, top, p, ffi.NB_GRAD)
    _, gf = gpu(xyz, box, top, p, ffi.NB_GRAD | ffi.NB_FINITE_PCG)
    assert rel(gf, gs) <= 1e-8


def test_finite_with_virial_is_rejected():
    name, xyz, box, top, p = CASES[0]
    h = ffi.Handle(ffi.params_from(p, p.resolved_grid(box), ffi.NB_PREC_FP64), top)
    try:
        h.set_cell(box)
        h.set_positions(xyz.reshape(-1))
        e, rc = h.evaluate(ffi.NB_GRAD | ffi.NB_FINITE_PCG | ffi.NB_VIRIAL, check=False)
        assert rc == ffi.NB_ERR_INVALID
    finally:
        h.close()
