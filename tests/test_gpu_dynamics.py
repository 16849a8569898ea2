"""The ASPC SCF predictor and velocity-Verlet dynamics on the device.

* Predictor (SCFPredictorParameters.java:79-121, 269-302): on a water box
  moving along a smooth path, every solve after the sixth starts from the
  direct dipoles plus Kolafa's extrapolation of the last six (u - u_direct);
  the oracle, given the same starting point (built here from its own earlier
  solutions), returns the same iteration counts, energies, gradients and
  dipoles at 1e-8, and the predicted solves take fewer iterations.
* Velocity Verlet (VelocityVerlet.java:66-113) on the reference's ubiquitin
  structure with its valence terms: three device steps equal three numpy steps
  driven by the oracle's gradients at 1e-9; energy is conserved over a short
  NVE run; graph replays give the regular path's trajectory."""
import dataclasses

import numpy as np
import pytest

from forcefieldx_amd import ffi, tinker
from forcefieldx_amd.forcefield import NonbondedParams
from forcefieldx_amd.systems import water_box
from test_tinker import GOLDEN, load_fixture

pytestmark = pytest.mark.gpu

ASPC = np.array([22.0 / 7.0, -55.0 / 14.0, 55.0 / 21.0, -22.0 / 21.0, 5.0 / 21.0, -1.0 / 42.0])


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_aspc_predictor_matches_oracle(oracle):
    xyz, box, top = water_box(216, 18.643, seed=1)
    p = NonbondedParams(vdw_cutoff=9.0, ewald_cutoff=7.0, grid=(24, 24, 24), polar_eps=1e-6)
    rng = np.random.default_rng(3)
    vel = rng.normal(0.0, 0.004, xyz.shape)          # A per step, a straight path
    h = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
    cold = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
    hist = []
    try:
        h.set_cell(box)
        cold.set_cell(box)
        h.set_scf_predictor(ffi.NB_PRED_ASPC)
        for t in range(10):
            x = xyz + t * vel
            guess = np.tensordot(ASPC, np.array(hist), axes=1) if len(hist) == 6 else None
            eo, go, uo, upo, direct = oracle.evaluate(oracle.System(x, box, top), p, scf_guess=guess,
                                                      want_direct=True)
            hist = ([np.stack([uo - direct[0], upo - direct[1]])] + hist)[:6]
            h.set_positions(x.reshape(-1))
            e, _ = h.evaluate(ffi.NB_GRAD)
            u, up = h.induced()
            cold.set_positions(x.reshape(-1))
            ec, _ = cold.evaluate(ffi.NB_GRAD)
            print(f"step {t}: iterations predicted {e['pcg_iters']} oracle {eo['pcg_iters']} cold {ec['pcg_iters']}",
                  flush=True)
            assert e["pcg_iters"] == eo["pcg_iters"], t
            assert abs(e["total"] - eo["total"]) <= 1e-8 * abs(eo["total"]), t
            assert rel(h.gradient(), go) <= 1e-8, t
            assert rel(u, uo) <= 1e-8 and rel(up, upo) <= 1e-8, t
            if t >= 6:
                assert e["pcg_iters"] < ec["pcg_iters"], t
    finally:
        h.close()
        cold.close()


def ubiquitin(polar_eps=None):
    xyz, box, top, p, terms = load_fixture("ubiquitin", valence=True)
    if polar_eps is not None:
        p = dataclasses.replace(p, polar_eps=polar_eps)
    d = GOLDEN + "/ubiquitin/"
    ff = tinker.read_prm(d + "amoeba_bio_2009_ubiquitin.prm.gz")
    mass = tinker.masses(tinker.read_xyz(d + "ubiquitin.xyz.gz").types, ff)
    return xyz, box, top, p, terms, mass


def md_handle(top, box, p, terms, graph=False, predictor=ffi.NB_PRED_NONE):
    h = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
    h.set_cell(box)
    h.set_valence(terms)
    if predictor != ffi.NB_PRED_NONE:
        h.set_scf_predictor(predictor)
    if graph:
        h.set_graph(True)
    return h


def velocities(mass, seed=5, temperature=298.15):
    rng = np.random.default_rng(seed)
    sd = np.sqrt(0.0019872042586408316 * temperature * 418.4 / mass)
    return rng.normal(size=(len(mass), 3)) * sd[:, None]


def test_velocity_verlet_matches_oracle(oracle):
    xyz, box, top, p, terms, mass = ubiquitin()
    v0 = velocities(mass)
    dt = 0.001
    h = md_handle(top, box, p, terms)
    try:
        h.set_positions(xyz.reshape(-1))
        h.md_init(mass, v0.reshape(-1))
        rep = h.md_step(3, dt)
        xg, vg = h.md_state()
    finally:
        h.close()

    def grad(x):
        e, g, _, _ = oracle.evaluate(oracle.System(x, box, top), p)
        ev, gv = oracle.valence(terms, x)
        return e["total"] + ev["total"], g + gv

    x, v = xyz.copy(), v0.copy()
    m = mass[:, None]
    a = -418.4 * grad(x)[1] / m
    for _ in range(3):
        v = v + a * (0.5 * dt)
        x = x + v * dt
        pe, g = grad(x)
        a = -418.4 * g / m
        v = v + a * (0.5 * dt)
    ke = 0.5 * float((m * v * v).sum()) / 418.4
    print(f"after 3 steps: KE {rep['kinetic']:.8f} / {ke:.8f}, PE {rep['potential']:.8f} / {pe:.8f}", flush=True)
    assert np.abs(xg - x).max() <= 1e-9
    assert rel(vg, v) <= 1e-9
    assert abs(rep["kinetic"] - ke) <= 1e-9 * ke
    assert abs(rep["potential"] - pe) <= 1e-8 * abs(pe)


def test_nve_energy_conservation_and_graph_replay():
    """40 steps of 0.5 fs with the predictor and graph replays: the total
    energy stays within 0.1 % of the kinetic energy of its start, and the
    trajectory equals the regular path's."""
    xyz, box, top, p, terms, mass = ubiquitin(polar_eps=1e-7)
    v0 = velocities(mass, seed=9)
    runs = []
    for graph in (True, False):
        h = md_handle(top, box, p, terms, graph=graph, predictor=ffi.NB_PRED_ASPC)
        try:
            h.set_positions(xyz.reshape(-1))
            h.md_init(mass, v0.reshape(-1))
            r0 = h.md_step(0, 0.0005)
            reps = [h.md_step(10, 0.0005) for _ in range(4)]
            x, v = h.md_state()
            runs.append((r0, reps, x, v, h.graph_replays()))
        finally:
            h.close()
    (r0, reps, x, v, replays), (s0, sreps, sx, sv, _) = runs
    drift = max(abs(r["total"] - r0["total"]) for r in reps)
    iters = [r["pcg_iters"] / r["steps"] for r in reps]
    print(f"E0 {r0['total']:.4f} KE0 {r0['kinetic']:.2f} max drift {drift:.4f}; PCG iterations per step {iters}; "
          f"graph replays {replays}", flush=True)
    assert drift <= 1e-3 * r0["kinetic"]
    assert replays >= 30
    assert np.abs(x - sx).max() <= 1e-9 and rel(v, sv) <= 1e-9


def test_berendsen_scales_velocities_after_the_full_step():
    """One step under the Berendsen thermostat (Berendsen.java:99-101) equals
    one NVE step with the velocities scaled by sqrt(1 + dt/tau (T0/T - 1)),
    T = sum m v^2 / (kB (3n - 3)) from the NVE step's kinetic energy
    (Thermostat.java:167, 262-274); a time step above tau is refused."""
    xyz, box, top, p, terms, mass = ubiquitin()
    v0 = velocities(mass)
    dt, tau, t0 = 0.001, 0.02, 350.0
    out = []
    for kind in (ffi.NB_THERMO_NONE, ffi.NB_THERMO_BERENDSEN):
        h = md_handle(top, box, p, terms)
        try:
            h.set_positions(xyz.reshape(-1))
            h.md_set_thermostat(kind, t0, tau)
            h.md_init(mass, v0.reshape(-1))
            rep = h.md_step(1, dt)
            out.append((rep, *h.md_state()))
            if kind == ffi.NB_THERMO_BERENDSEN:
                with pytest.raises(ffi.NbError):
                    h.md_step(1, 2 * tau)
        finally:
            h.close()
    (r1, x1, v1), (r2, x2, v2) = out
    kb, dof = 0.831446261815324, 3 * top.n - 3
    temp = lambda ke: 2.0 * ke * 418.4 / (kb * dof)
    s = np.sqrt(1.0 + dt / tau * (t0 / temp(r1["kinetic"]) - 1.0))
    assert s > 1.0
    # equal but for the summation order of the grid's atomic adds
    assert np.abs(x2 - x1).max() <= 1e-12 * np.abs(x1).max()
    assert np.abs(v2 - s * v1).max() <= 1e-12 * np.abs(v1).max()
    assert abs(r2["kinetic"] - s * s * r1["kinetic"]) <= 1e-12 * r1["kinetic"]


def test_predictor_history_skips_rejected_replays():
    """A graph replay whose list needs rebuilding is repeated on the regular
    path; the replay's solve must not enter the predictor's history, or the
    history of the following solves is one step off.  Along a path with a
    jump that forces a rebuild, a graph-mode handle and a regular handle with
    the ASPC predictor return the same iteration counts and dipoles (a loose
    polar-eps makes any difference in the starting points visible)."""
    xyz, box, top = water_box(216, 18.643, seed=1)
    p = NonbondedParams(vdw_cutoff=9.0, ewald_cutoff=7.0, grid=(24, 24, 24), polar_eps=1e-4)
    rng = np.random.default_rng(8)
    vel = rng.normal(0.0, 0.004, xyz.shape)
    hs = []
    try:
        for graph in (True, False):
            h = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
            h.set_cell(box)
            h.set_scf_predictor(ffi.NB_PRED_ASPC)
            h.set_graph(graph)
            hs.append(h)
        rebuilt = []
        for t in range(16):
            x = (xyz + t * vel).reshape(-1, 3)
            if t >= 10:
                x = x.copy()
                x[5] += 1.3          # more than half the 2 A buffer: the list is rebuilt at t = 10
            out = []
            for h in hs:
                h.set_positions(x.reshape(-1))
                e, _ = h.evaluate(ffi.NB_GRAD)
                out.append((e, h.induced()[0]))
            (eg, ug), (er, ur) = out
            rebuilt.append(eg["list_rebuilt"])
            print(f"step {t}: iterations graph {eg['pcg_iters']} regular {er['pcg_iters']}, rebuilt "
                  f"{eg['list_rebuilt']}", flush=True)
            assert eg["pcg_iters"] == er["pcg_iters"], t
            assert rel(ug, ur) <= 1e-9, t
            assert abs(eg["total"] - er["total"]) <= 1e-9 * abs(er["total"]), t
        assert rebuilt[10] == 1
        assert hs[0].graph_replays() >= 8
    finally:
        for h in hs:
            h.close()


def test_berendsen_leaves_massless_atoms_alone():
    """Atoms with mass <= 0 do not move (drift and kick skip them), so the
    thermostat leaves their velocities as they are and counts 3m - 3 degrees
    of freedom over the m atoms with mass > 0."""
    xyz, box, top, p, terms, mass = ubiquitin()
    mass = mass.copy()
    frozen = np.arange(0, 40, 4)
    mass[frozen] = 0.0
    v0 = velocities(np.where(mass > 0.0, mass, 1.0))
    dt, tau, t0 = 0.001, 0.02, 350.0
    out = []
    for kind in (ffi.NB_THERMO_NONE, ffi.NB_THERMO_BERENDSEN):
        h = md_handle(top, box, p, terms)
        try:
            h.set_positions(xyz.reshape(-1))
            h.md_set_thermostat(kind, t0, tau)
            h.md_init(mass, v0.reshape(-1))
            rep = h.md_step(1, dt)
            out.append((rep, *h.md_state()))
        finally:
            h.close()
    (r1, x1, v1), (r2, x2, v2) = out
    active = mass > 0.0
    kb, dof = 0.831446261815324, 3 * int(active.sum()) - 3
    temp = 2.0 * r1["kinetic"] * 418.4 / (kb * dof)
    s = np.sqrt(1.0 + dt / tau * (t0 / temp - 1.0))
    assert np.array_equal(v2[frozen], v0[frozen]) and np.array_equal(x2[frozen], xyz.reshape(-1, 3)[frozen])
    assert np.abs(v2[active] - s * v1[active]).max() <= 1e-12 * np.abs(v1).max()


def test_batch_of_predictor_handles_keeps_the_history():
    """Handles with an SCF predictor are evaluated by nb_evaluate_batch on the
    regular path (the batch kernel neither applies nor extends the history):
    switching between batch and single evaluations gives the same sequence
    as single evaluations alone."""
    xyz, box, top = water_box(216, 18.643, seed=1)
    p = NonbondedParams(vdw_cutoff=9.0, ewald_cutoff=7.0, grid=(24, 24, 24), polar_eps=1e-4)
    rng = np.random.default_rng(5)
    vel = rng.normal(0.0, 0.004, xyz.shape)
    mk = lambda: ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
    a, b = mk(), mk()
    try:
        for h in (a, b):
            h.set_cell(box)
            h.set_scf_predictor(ffi.NB_PRED_ASPC)
        for t in range(12):
            x = (xyz + t * vel).reshape(-1)
            a.set_positions(x)
            b.set_positions(x)
            if t % 3 == 1:
                ea, rc = ffi.evaluate_batch([a], ffi.NB_GRAD)
                assert rc == ffi.NB_OK
                ea = ea[0]
            else:
                ea, _ = a.evaluate(ffi.NB_GRAD)
            eb, _ = b.evaluate(ffi.NB_GRAD)
            assert ea["pcg_iters"] == eb["pcg_iters"], t
            assert rel(a.induced()[0], b.induced()[0]) <= 1e-9, t
    finally:
        a.close()
        b.close()


def ls_coefficients(hist, start, count):
    """The least-squares predictor's fit as the reference defines it
    (SCFPredictorParameters.java:144-213, LeastSquaresPredictor.value and
    updateJacobianAndTarget, :341-418): the newest entry of the ring (u and
    uCR parts, all 6n values) against the five before it, over the entries
    filled so far; the other coefficients stay 0."""
    c = np.zeros(6)
    if count < 2:
        return c
    m = min(count - 1, 5)
    e = [hist[(start + k) % 6].reshape(-1) for k in range(6)]
    A = np.stack([e[1 + j] for j in range(m)], axis=1)
    c[:m] = np.linalg.lstsq(A, e[0], rcond=None)[0]
    return c


def test_ls_predictor_matches_oracle(oracle):
    """NB_PRED_LS along a random walk (steps of sigma 0.02 A): each prediction's
    coefficients equal numpy's least-squares fit of the history the device
    held before the solve (1e-8), and the oracle started from the same
    extrapolation of its own earlier solutions returns the same iteration
    counts, energies, gradients and dipoles at 1e-8."""
    xyz, box, top = water_box(216, 18.643, seed=1)
    p = NonbondedParams(vdw_cutoff=9.0, ewald_cutoff=7.0, grid=(24, 24, 24), polar_eps=1e-6)
    rng = np.random.default_rng(8)
    h = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
    ohist, prev, predicted = [], None, 0
    x = xyz.reshape(-1, 3).copy()
    try:
        h.set_cell(box)
        h.set_scf_predictor(ffi.NB_PRED_LS)
        for t in range(10):
            x = x + rng.normal(0.0, 0.02, x.shape)
            h.set_positions(x.reshape(-1))
            e, _ = h.evaluate(ffi.NB_GRAD)
            u, up = h.induced()
            coef, hist, start, count = h.scf_predictor()
            guess = None
            if prev is not None and prev[2] >= 2:
                expect = ls_coefficients(*prev)
                print(f"step {t}: {prev[2]} entries, LS coefficients {np.round(coef[:5], 6)}", flush=True)
                assert np.abs(coef - expect).max() <= 1e-8 * max(1.0, np.abs(expect).max()), t
                guess = sum(coef[k] * ohist[k] for k in range(min(5, len(ohist))))
                predicted += 1
            eo, go, uo, upo, direct = oracle.evaluate(oracle.System(x, box, top), p, scf_guess=guess,
                                                      want_direct=True)
            ohist = ([np.stack([uo - direct[0], upo - direct[1]])] + ohist)[:6]
            print(f"step {t}: iterations {e['pcg_iters']} oracle {eo['pcg_iters']}", flush=True)
            assert e["pcg_iters"] == eo["pcg_iters"], t
            assert abs(e["total"] - eo["total"]) <= 1e-8 * abs(eo["total"]), t
            assert rel(h.gradient(), go) <= 1e-8, t
            assert rel(u, uo) <= 1e-8 and rel(up, upo) <= 1e-8, t
            assert count == min(t + 1, 6)
            prev = (hist, start, count)
        assert predicted == 8
    finally:
        h.close()


def test_ls_predictor_saves_iterations_on_a_smooth_path():
    """Along a straight path the LS extrapolation of the last solutions
    starts the solve closer to its answer than the direct dipoles do."""
    xyz, box, top = water_box(216, 18.643, seed=1)
    p = NonbondedParams(vdw_cutoff=9.0, ewald_cutoff=7.0, grid=(24, 24, 24), polar_eps=1e-6)
    vel = np.random.default_rng(3).normal(0.0, 0.004, xyz.shape)
    its = {}
    for kind in (ffi.NB_PRED_NONE, ffi.NB_PRED_LS):
        h = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
        try:
            h.set_cell(box)
            if kind != ffi.NB_PRED_NONE:
                h.set_scf_predictor(kind)
            its[kind] = []
            for t in range(10):
                h.set_positions((xyz + t * vel).reshape(-1))
                e, _ = h.evaluate(ffi.NB_GRAD)
                its[kind].append(e["pcg_iters"])
        finally:
            h.close()
    print(f"PCG iterations: none {its[ffi.NB_PRED_NONE]}, LS {its[ffi.NB_PRED_LS]}", flush=True)
    assert sum(its[ffi.NB_PRED_LS][4:]) < sum(its[ffi.NB_PRED_NONE][4:])


def test_md_step_graph_with_list_rebuilds():
    """Graph mode records a whole velocity-Verlet step (drift, evaluation,
    kick) as one graph.  With a 0.3 A list buffer ubiquitin rebuilds its
    list every few steps: those replays are not kept (their kick returns at
    once) and the step is finished on the regular path from the drifted
    positions.  The trajectory equals the regular path's, with the LS
    predictor's history kept through the rejected replays."""
    xyz, box, top, p, terms, mass = ubiquitin(polar_eps=1e-6)
    p = dataclasses.replace(p, list_buffer=0.3)
    v0 = velocities(mass, seed=12)
    runs = []
    for graph in (True, False):
        h = md_handle(top, box, p, terms, graph=graph, predictor=ffi.NB_PRED_LS)
        try:
            h.set_positions(xyz.reshape(-1))
            h.md_init(mass, v0.reshape(-1))
            reps = [h.md_step(1, 0.001) for _ in range(30)]
            x, v = h.md_state()
            runs.append((reps, x, v, h.graph_replays()))
        finally:
            h.close()
    (reps, x, v, replays), (sreps, sx, sv, _) = runs
    rebuilds = sum(int(r["last"]["list_rebuilt"]) for r in sreps)
    print(f"graph replays {replays} of 30, list rebuilds {rebuilds}, PCG iterations "
          f"{[r['pcg_iters'] for r in reps]}", flush=True)
    assert rebuilds >= 2
    assert 10 <= replays <= 30 - 1 - rebuilds
    for r, s in zip(reps, sreps):
        assert r["pcg_iters"] == s["pcg_iters"]
        assert abs(r["total"] - s["total"]) <= 1e-9 * abs(s["total"])
    assert np.abs(x - sx).max() <= 1e-9 and rel(v, sv) <= 1e-9
