"""The five benchmark configurations at their full sizes.

* C2, water-20736 (6912 waters, 59.1 A, 64^3, Ewald 7 A, vdW 9 A): the FP64
  form equals the oracle at 1e-8 (energies, gradient, induced dipoles,
  interaction and PCG iteration counts).
* C3, chain-23557 (2,500-atom chain in 7,019 waters, 62.2 A, 64^3, Ewald
  7 A, 20 soft atoms): mixed against FP64 for the forces and dU/dL at
  L in {0, 0.25, 0.5, 0.75, 1}; FP64 against the oracle at L = 0.5.
* C4, crystals-512 (512 triclinic cells of 30-300 atoms): through the batch
  kernel, with fewer persistent workgroups than cells, so that workgroups
  take a second cell from the queue; FP64 equals the oracle on every cell,
  mixed is held to 1e-5 of FP64 over all cells.
* C5, water-288000 (96,000 waters, 144.3 A, 144^3): graph replays under
  Gaussian steps of sigma 0.01 A applied to the coordinates on the device;
  mixed against FP64 after the steps, and a replay equals the regular path
  on the same coordinates.
"""
import dataclasses

import numpy as np
import pytest

from forcefieldx_amd import ffi
from forcefieldx_amd.energy import ForceFieldEnergy
from forcefieldx_amd.forcefield import NonbondedParams
from forcefieldx_amd.systems import chain_in_water, crystal_batch, water_box

pytestmark = pytest.mark.gpu

REL = 1e-8


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_c2_water20736_fp64_matches_oracle(oracle):
    xyz, box, top = water_box(6912, 59.1, seed=2)
    p = NonbondedParams(vdw_cutoff=9.0, grid=(64, 64, 64), polar_eps=1e-5)
    print("C2 water-20736 FP64 against the oracle", flush=True)
    eo, go, uo, upo = oracle.evaluate(oracle.System(xyz, box, top), p)
    ffe = ForceFieldEnergy(top, box, p, precision=ffi.NB_PREC_FP64)
    try:
        e, g = ffe.energyAndGradient(xyz)
        u, up = ffe.particleMeshEwald.getInducedDipoles()
        last = dict(ffe._last)
    finally:
        ffe.destroy()
    print(f"C2 FP64: total {e:.10f} oracle {eo['total']:.10f}; PCG {last['pcg_iters']} / {eo['pcg_iters']}",
          flush=True)
    assert last["pcg_iters"] == eo["pcg_iters"]
    assert last["n_vdw"] == eo["n_vdw"] and last["n_perm"] == eo["n_perm"]
    scale = abs(eo["total"])
    for k in ("perm_real", "perm_self", "perm_recip", "pol_real", "pol_self", "pol_recip", "vdw", "total"):
        assert abs(last[k] - eo[k]) <= REL * max(abs(eo[k]), 1e-3 * scale), (k, last[k], eo[k])
    assert rel_err(g.reshape(-1, 3), go) <= REL
    assert rel_err(u, uo) <= REL and rel_err(up, upo) <= REL


def test_c2_refinement_brings_mixed_dipoles_to_fp64():
    """FP64 iterative refinement of the mixed solve (nb_set_refinement): at a
    tolerance where the FP32 field operator limits the mixed dipoles, the
    FP64 residual test restarts the solve, and the refined dipoles lie closer
    to the converged FP64 ones than the FP32-operator solve's.  At the C2
    tolerance (1e-5 D) the refined mixed form stays within 1e-5 of FP64."""
    xyz, box, top = water_box(6912, 59.1, seed=2)
    ref = NonbondedParams(vdw_cutoff=9.0, grid=(64, 64, 64), polar_eps=1e-10)
    fp64 = {}
    for eps in (1e-10, 1e-5):
        print(f"C2 refinement: FP64 at {eps} D", flush=True)
        p = dataclasses.replace(ref, polar_eps=eps)
        h = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
        try:
            h.set_cell(box)
            h.set_positions(xyz)
            h.evaluate(ffi.NB_GRAD)
            fp64[eps] = (h.induced()[0], h.gradient())
        finally:
            h.close()
    u64, g64 = fp64[1e-10]
    res = {}
    for eps in (1e-7, 1e-5):
        p = dataclasses.replace(ref, polar_eps=eps)
        for on in (False, True):
            h = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_MIXED), top)
            try:
                h.set_cell(box)
                h.set_refinement(on)
                h.set_positions(xyz)
                e, _ = h.evaluate(ffi.NB_GRAD)
                res[eps, on] = (e, h.induced()[0], h.gradient())
            finally:
                h.close()
            e, u, g = res[eps, on]
            print(f"polar_eps {eps} refinement {on}: PCG {e['pcg_iters']}, eps {e['pcg_eps']:.3e}, dipole err "
                  f"{rel_err(u, u64):.3e}, force err {rel_err(g, g64):.3e}", flush=True)
    (e_off, u_off, _), (e_on, u_on, _) = res[1e-7, False], res[1e-7, True]
    assert e_on["pcg_eps"] < 1e-7
    assert rel_err(u_on, u64) < rel_err(u_off, u64)
    for on in (False, True):
        assert rel_err(res[1e-5, on][2], fp64[1e-5][1]) <= 1e-5


C3_LAMBDAS = (0.0, 0.25, 0.5, 0.75, 1.0)


def c3_case():
    xyz, box, top = chain_in_water(2500, 7019, 62.2, seed=7, n_soft=20)
    p = NonbondedParams(vdw_cutoff=9.0, ewald_cutoff=7.0, grid=(64, 64, 64), polar_eps=1e-5, elec_lambda=True)
    return xyz.reshape(-1, 3), box, top, p


def c3_eval(xyz, box, top, p, precision, lams):
    ffe = ForceFieldEnergy(top, box, p, precision=precision, direct_fallback=False)
    out = {}
    try:
        for lam in lams:
            ffe.setLambda(lam)
            _, g = ffe.energyAndGradient(xyz.reshape(-1))
            out[lam] = (dict(ffe._last), g.reshape(-1, 3), ffe.getdEdXdL().reshape(-1, 3))
    finally:
        ffe.destroy()
    return out


def test_c3_chain23557_mixed_against_fp64():
    xyz, box, top, p = c3_case()
    print("C3 chain-23557: FP64 and mixed at five lambdas", flush=True)
    r64 = c3_eval(xyz, box, top, p, ffi.NB_PREC_FP64, C3_LAMBDAS)
    rmx = c3_eval(xyz, box, top, p, ffi.NB_PREC_MIXED, C3_LAMBDAS)
    for lam in C3_LAMBDAS:
        (e64, g64, l64), (emx, gmx, lmx) = r64[lam], rmx[lam]
        print(f"C3 lambda={lam}: dU/dL {e64['dudl']:.6f} / {emx['dudl']:.6f}, force err {rel_err(gmx, g64):.2e}, "
              f"dU/dX/dL err {rel_err(lmx, l64):.2e}", flush=True)
        assert rel_err(gmx, g64) <= 1e-5, lam
        # dU/dL against its own size, with a floor of 10 kcal/mol where it passes through zero
        assert abs(emx["dudl"] - e64["dudl"]) <= 1e-5 * max(abs(e64["dudl"]), 10.0), lam
        # dU/dX/dL is a lambda derivative of the gradient, not a force: its
        # terms partly cancel at L = 1 (1.05e-5 measured there), so it is held
        # to 1e-4 of its own size
        if np.linalg.norm(l64) > 0.0:
            assert rel_err(lmx, l64) <= 1e-4, lam


def test_c3_chain23557_fp64_matches_oracle(oracle):
    xyz, box, top, p = c3_case()
    lam = 0.5
    print("C3 chain-23557: FP64 against the oracle at lambda 0.5", flush=True)
    eo, go, lgo = oracle.evaluate_lambda(oracle.System(xyz, box, top), dataclasses.replace(p, lambda_value=lam))
    e, g, lg = c3_eval(xyz, box, top, p, ffi.NB_PREC_FP64, (lam,))[lam]
    scale = abs(eo["total"])
    for k in ("perm_real", "perm_self", "perm_recip", "pol_real", "pol_self", "pol_recip", "vdw", "total", "dudl",
              "d2udl2"):
        assert abs(e[k] - eo[k]) <= REL * max(abs(eo[k]), 1e-3 * scale), (k, e[k], eo[k])
    assert rel_err(g, go) <= REL
    assert rel_err(lg, lgo) <= REL


def c4_handles(systems, params, precision):
    out = []
    for xyz, box, top, grid in systems:
        h = ffi.Handle(ffi.params_from(params, grid, precision), top)
        h.set_cell(box)
        h.set_positions(xyz)
        out.append(h)
    return out


def test_c4_crystals512_batch_kernel(oracle):
    systems = crystal_batch(512, seed=2)
    natoms = sum(s[2].n for s in systems)
    counts = [s[2].n for s in systems]
    assert min(counts) >= 30 and max(counts) <= 300
    p = NonbondedParams(ewald_cutoff=6.0, vdw_cutoff=6.0, polar_eps=1e-5)
    print(f"C4: {len(systems)} cells, {natoms} atoms", flush=True)
    h64 = c4_handles(systems, p, ffi.NB_PREC_FP64)
    try:
        e64, rc = ffi.evaluate_batch(h64, ffi.NB_GRAD)
        assert rc == ffi.NB_OK
        cfg = ffi.batch_config()
        print(f"C4 FP64 batch kernel: {cfg}", flush=True)
        # fewer persistent workgroups than cells: workgroups take further cells from the queue
        assert cfg["systems"] == len(systems) and cfg["workgroups"] < len(systems)
        g64 = [h.gradient() for h in h64]
        u64 = [h.induced()[0] for h in h64]
    finally:
        for h in h64:
            h.close()
    worst = 0.0
    for b, (xyz, box, top, grid) in enumerate(systems):
        eo, go, uo, _ = oracle.evaluate(oracle.System(xyz, box, top), dataclasses.replace(p, grid=grid))
        scale = max(abs(eo[k]) for k in ("perm", "pol", "vdw"))
        for k in ("perm_real", "perm_self", "perm_recip", "pol", "vdw", "total"):
            assert abs(e64[b][k] - eo[k]) <= REL * scale, (b, k, e64[b][k], eo[k])
        assert e64[b]["pcg_iters"] == eo["pcg_iters"], b
        assert e64[b]["n_vdw"] == eo["n_vdw"], b
        err = rel_err(g64[b], go)
        worst = max(worst, err)
        assert err <= REL, (b, err)
        assert rel_err(u64[b], uo) <= REL, b
    print(f"C4 FP64 = oracle on all {len(systems)} cells, worst gradient error {worst:.2e}", flush=True)
    hmx = c4_handles(systems, p, ffi.NB_PREC_MIXED)
    try:
        emx, rc = ffi.evaluate_batch(hmx, ffi.NB_GRAD)
        assert rc == ffi.NB_OK
        assert ffi.batch_config()["workgroups"] < len(systems)
        gmx = [h.gradient() for h in hmx]
    finally:
        for h in hmx:
            h.close()
    err = rel_err(np.concatenate(gmx), np.concatenate(g64))
    print(f"C4 mixed against FP64: force rel RMS error {err:.2e}", flush=True)
    assert err <= 1e-5
    # energies against the largest term summed into them: the per-cell totals
    # of these small cells cancel to well below their reciprocal and self terms
    worst_e = 0.0
    for c, (a, b) in enumerate(zip(emx, e64)):
        scale = max(abs(b[k]) for k in ("perm_real", "perm_self", "perm_recip", "pol_real", "pol_self",
                                        "pol_recip", "vdw"))
        worst_e = max(worst_e, abs(a["total"] - b["total"]) / scale)
        assert abs(a["total"] - b["total"]) <= 1e-5 * scale, (c, a["total"], b["total"], scale)
    print(f"C4 mixed against FP64: worst energy error {worst_e:.2e} of the largest term", flush=True)


def test_c5_water288000_graph_replays_under_device_steps():
    xyz, box, top = water_box(96000, 144.3, seed=2)
    p = NonbondedParams(vdw_cutoff=9.0, grid=(144, 144, 144), polar_eps=1e-5)
    steps = 20
    # a random walk of sigma 0.01 A steps, every frame resident on the device
    rng = np.random.default_rng(4321)
    x0 = np.asarray(xyz, np.float64).reshape(-1)
    frames = ffi.DeviceArray(x0.size, frames=steps)
    x = x0.copy()
    for k in range(steps):
        x += rng.normal(0.0, 0.01, x.shape)
        frames.upload(x, k)
    res = {}
    for prec in (ffi.NB_PREC_FP64, ffi.NB_PREC_MIXED):
        print(f"C5 water-288000 precision={prec}: {steps} graph replays", flush=True)
        h = ffi.Handle(ffi.params_from(p, p.grid, prec), top)
        try:
            h.set_cell(box)
            h.set_graph(True)
            for k in range(steps):
                h.set_positions_device(frames.ptr(k))
                e, _ = h.evaluate(ffi.NB_GRAD)
            replays = h.graph_replays()
            g_replay = h.gradient()
            u_replay = h.induced()[0]
            # the same coordinates on the regular path
            h.set_graph(False)
            h.set_positions_device(frames.ptr(steps - 1))
            e_reg, _ = h.evaluate(ffi.NB_GRAD)
            g_reg = h.gradient()
            res[prec] = (e, g_replay, u_replay, e_reg, g_reg, replays)
        finally:
            h.close()
        print(f"C5 precision={prec}: {replays} replays, total {e['total']:.6f}, PCG {e['pcg_iters']}", flush=True)
        assert replays >= steps - 2
        tol = 1e-12
        assert np.abs(g_replay - g_reg).max() <= tol * np.abs(g_reg).max()
        assert abs(e["total"] - e_reg["total"]) <= tol * abs(e_reg["perm_self"])
    e64, g64, u64 = res[ffi.NB_PREC_FP64][:3]
    emx, gmx, umx = res[ffi.NB_PREC_MIXED][:3]
    print(f"C5 mixed against FP64: force {rel_err(gmx, g64):.2e}, dipoles {rel_err(umx, u64):.2e}", flush=True)
    assert rel_err(gmx, g64) <= 1e-5
    assert rel_err(umx, u64) <= 1e-5
    assert abs(emx["total"] - e64["total"]) <= 1e-6 * abs(e64["perm_self"])
    frames.close()
