"""Cells thinner than twice the cutoff on the GPU (image sums with the lattice
translation in each list entry) against the oracle, whose image sums are
pinned by 2x2x2 replicated supercells (tests/test_oracle_images.py)."""
import dataclasses

import numpy as np
import pytest

from forcefieldx_amd import ffi
from forcefieldx_amd.energy import ForceFieldEnergy
from forcefieldx_amd.forcefield import NonbondedParams
from forcefieldx_amd.systems import replicate, small_crystal

pytestmark = pytest.mark.gpu

REL = 1e-8
CELLS = [([9.5, 10.2, 11.0, 80.0, 95.0, 103.0], 2), ([8.4, 12.5, 9.1, 101.0, 72.0, 88.0], 3),
         ([14.0, 8.2, 16.5, 70.0, 110.0, 96.0], 4)]


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def params(pol=2, grid=(16, 16, 16)):
    return NonbondedParams(ewald_cutoff=6.0, vdw_cutoff=6.0, grid=grid, polar_eps=1e-8, polarization=pol)


@pytest.mark.parametrize("pol", [0, 1, 2])
@pytest.mark.parametrize("cell,seed", CELLS)
def test_thin_cell_fp64_matches_oracle(oracle, cell, seed, pol):
    xyz, box, top = small_crystal(int(0.05 * np.prod(cell[:3])), cell, seed=seed)
    p = params(pol)
    print(f"thin cell {cell} n={top.n} polarization={pol}", flush=True)
    eo, go, uo, upo = oracle.evaluate(oracle.System(xyz, box, top), p)
    ffe = ForceFieldEnergy(top, box, p, precision=ffi.NB_PREC_FP64, direct_fallback=False)
    try:
        e, g, W = ffe.energyGradientAndVirial(xyz)
        last = dict(ffe._last)
        u, up = ffe.particleMeshEwald.getInducedDipoles()
    finally:
        ffe.destroy()
    keys = ("perm_real", "perm_self", "perm_recip", "pol", "vdw", "total")
    scale = max(abs(eo[k]) for k in keys)
    for k in keys:
        assert abs(last[k] - eo[k]) <= REL * scale, (k, last[k], eo[k])
    assert last["n_vdw"] == eo["n_vdw"]
    assert rel_err(g.reshape(-1, 3), go) <= REL
    if pol:
        assert rel_err(u, uo) <= REL
        assert last["pcg_iters"] == eo["pcg_iters"]
    assert np.abs(W - eo["vir"]).max() <= REL * np.abs(eo["vir"]).max()


def test_thin_cell_equals_replicated_supercell_on_gpu():
    """GPU image sums on the cell against the GPU minimum-image path on the
    2x2x2 supercell (grid doubled): 8x the energy, the same gradients."""
    cell, seed = CELLS[0]
    xyz, box, top = small_crystal(48, cell, seed=seed)
    X, B, T = replicate(xyz, box, top, (2, 2, 2))
    out = []
    for x, b, t, grid in ((xyz, box, top, (16, 16, 16)), (X, B, T, (32, 32, 32))):
        ffe = ForceFieldEnergy(t, b, params(2, grid), precision=ffi.NB_PREC_FP64, direct_fallback=False)
        try:
            e, g = ffe.energyAndGradient(x)
            out.append((dict(ffe._last), g.reshape(-1, 3)))
        finally:
            ffe.destroy()
    (e1, g1), (e8, g8) = out
    assert abs(8 * e1["total"] - e8["total"]) <= 1e-9 * max(abs(e8["perm"]), abs(e8["vdw"]))
    assert rel_err(g8[:top.n], g1) <= 1e-9


def test_thin_cell_mixed_against_fp64():
    cell, seed = CELLS[1]
    xyz, box, top = small_crystal(int(0.05 * np.prod(cell[:3])), cell, seed=seed)
    out = {}
    for prec in (ffi.NB_PREC_FP64, ffi.NB_PREC_MIXED):
        ffe = ForceFieldEnergy(top, box, params(), precision=prec, direct_fallback=False)
        try:
            out[prec] = ffe.energyAndGradient(xyz)[1]
        finally:
            ffe.destroy()
    assert rel_err(out[ffi.NB_PREC_MIXED], out[ffi.NB_PREC_FP64]) <= 1e-5
