"""AMOEBA valence terms (forcefieldx_amd.bonded) and their oracle
(oracle/valence_oracle.c).

Pinned by the reference's EnergyTest values for DHFR (EnergyTest.java:148-180)
and ubiquitin (:363-394): the number of terms of every kind and each kind's
energy, printed by the reference to 1e-8 kcal/mol; by central differences of
the energy (gradient and strain derivative) on distorted geometries, where
every term is away from its minimum; and by scipy's periodic and natural
cubic splines for the torsion-torsion grid derivatives."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from forcefieldx_amd import bonded
from test_tinker import SYSTEMS, expected, load_fixture

KINDS = ("bond", "angle", "strbnd", "urey", "opbend", "torsion", "pitors", "tortor")


def fixture(name):
    return load_fixture(name, valence=True)


@pytest.mark.parametrize("name", SYSTEMS)
def test_term_counts_and_energies_match_energy_test(oracle, name):
    xyz, box, top, p, terms = fixture(name)
    ref = expected(name)["valence"]
    counts = terms.counts()
    e, g = oracle.valence(terms, xyz)
    for k in KINDS:
        assert counts[k] == ref[k][1], (k, counts[k], ref[k][1])
        assert abs(e[k] - ref[k][0]) <= 1e-7, (k, e[k], ref[k][0])
    # translation invariance of every term
    assert np.abs(g.sum(axis=0)).max() <= 1e-9 * np.abs(g).max()


def distorted(name, seed=1, amp=0.08):
    xyz, box, top, p, terms = fixture(name)
    rng = np.random.default_rng(seed)
    return xyz + rng.normal(0.0, amp, xyz.shape), terms


def atoms_of_every_kind(terms, rng):
    picks = set()
    for arr in (terms.bond, terms.angle[terms.angle[:, 3] >= 0], terms.angle[terms.angle[:, 3] < 0], terms.strbnd,
                terms.urey, terms.opbend, terms.torsion, terms.pitors, terms.tortor[:, :5]):
        row = arr[rng.integers(len(arr))]
        picks.update(int(a) for a in row if a >= 0)
    tt = terms.tortor[terms.tortor[:, 5] >= 0]
    picks.update(int(a) for a in tt[rng.integers(len(tt))][:5])
    return sorted(picks)


def test_gradient_central_differences(oracle):
    xyz, terms = distorted("ubiquitin")
    e0, g = oracle.valence(terms, xyz)
    rng = np.random.default_rng(5)
    atoms = atoms_of_every_kind(terms, rng)
    assert len(atoms) >= 20
    h = 1e-6
    for i in atoms:
        for d in range(3):
            xp = xyz.copy(); xp[i, d] += h
            xm = xyz.copy(); xm[i, d] -= h
            fd = (oracle.valence(terms, xp)[0]["total"] - oracle.valence(terms, xm)[0]["total"]) / (2 * h)
            assert abs(g[i, d] - fd) <= 1e-5 * max(1.0, abs(fd)), (i, d, g[i, d], fd)


def test_strain_derivative_central_differences(oracle):
    """W_ab = dE/d eps_ab with every coordinate scaled by (1 + eps)."""
    xyz, terms = distorted("ubiquitin", seed=2)
    e0, g, vir = oracle.valence(terms, xyz, virial=True)
    h = 1e-7
    for a in range(3):
        for b in range(3):
            eps = np.zeros((3, 3))
            eps[a, b] = h
            ep = oracle.valence(terms, xyz @ (np.eye(3) + eps).T)[0]["total"]
            em = oracle.valence(terms, xyz @ (np.eye(3) - eps).T)[0]["total"]
            fd = (ep - em) / (2 * h)
            assert abs(vir[a, b] - fd) <= 1e-5 * np.abs(vir).max(), (a, b, vir[a, b], fd)


def test_in_plane_angle_and_out_of_plane_atom():
    """In-plane angles exist exactly at trigonal centres and take the fourth
    atom of the centre as their projection atom."""
    xyz, box, top, p, terms = fixture("dhfr")
    ip = terms.angle[terms.angle[:, 3] >= 0]
    assert len(ip) > 1000
    for a, b, c, d in ip[:200]:
        assert len({a, b, c, d}) == 4
    ob = {(int(a), int(b), int(c)): int(d) for a, b, c, d in terms.opbend}
    for a, b, c, d in ip:
        assert ob.get((int(a), int(b), int(c))) == d


def test_periodic_and_natural_spline_slopes():
    rng = np.random.default_rng(3)
    x = np.linspace(-180.0, 180.0, 25)
    y = rng.normal(size=25)
    y[-1] = y[0]
    ref = CubicSpline(x, y, bc_type="periodic").derivative()(x)
    assert np.abs(bonded.periodic_spline_slopes(x, y) - ref).max() <= 1e-12 * np.abs(ref).max()
    xn = np.sort(rng.uniform(0, 10, 9))
    yn = rng.normal(size=9)
    refn = CubicSpline(xn, yn, bc_type="natural").derivative()(xn)
    assert np.abs(bonded.natural_spline_slopes(xn, yn) - refn).max() <= 1e-12 * np.abs(refn).max()


def test_tortor_grid_reproduces_knot_values(oracle):
    """At a grid point the bicubic patch returns the tabulated energy."""
    xyz, box, top, p, terms = fixture("dhfr")
    gr = terms.grids[0]
    assert gr.nx == 25 and gr.ny == 25 and gr.tx[0] == -180.0 and gr.tx[-1] == 180.0
    # slopes of a periodic table agree at both ends of each row and column
    dx = gr.dx.reshape(gr.ny, gr.nx)
    assert np.abs(dx[:, 0] - dx[:, -1]).max() <= 1e-9
