"""Space-group crystals expanded to P1 (forcefieldx_amd.crystal,
tinker.load_crystal): the five crystal benchmarks of the reference's
EnergyTest (EnergyTest.java:435-614: acetanilide in Pbca, ethylparaben,
methylparaben, paracetamol and phenacetin in P21/c), read from
tests/golden/crystals.  Per asymmetric unit (P1 energy / number of
operators) the oracle reproduces each test's vdW energy and every valence
term but the angles to 1e-8 kcal/mol, the precision the test prints, which
pins the operators; the permanent and polarization energies and the angle
energies lie within the test's 1e-2 kcal/mol.  The test's angle energies are
those of in-plane angles at every trigonal centre: taking those reproduces
them to 1e-8 as well, while the parameter patches hold only `angle` records,
which AngleType assigns the normal form (Angle.java:176-190)."""
import json
import os

import numpy as np
import pytest

from forcefieldx_amd import tinker
from forcefieldx_amd.crystal import SPACE_GROUPS, cell_matrix, expand_to_p1

CRYSTALS = os.path.join(os.path.dirname(__file__), "golden", "crystals")
NAMES = ("acetanilide", "ethylparaben", "methylparaben", "paracetamol", "phenacetin")
KINDS = ("bond", "angle", "strbnd", "urey", "opbend", "torsion", "pitors", "tortor")


def load_crystal(name):
    return tinker.load_crystal(*(os.path.join(CRYSTALS, name + ext) for ext in (".xyz", ".key", ".prm")))


def expected(name):
    with open(os.path.join(CRYSTALS, "expected.json")) as f:
        return json.load(f)[name]


def in_plane_at_trigonal_centres(terms, n):
    adj = [set() for _ in range(n)]
    for i, j in terms.bond:
        adj[i].add(int(j))
        adj[j].add(int(i))
    for r in terms.angle:
        if len(adj[r[1]]) == 3:
            r[3] = min(adj[r[1]] - {int(r[0]), int(r[2])})
    return terms


@pytest.mark.parametrize("name", NAMES)
def test_per_asymmetric_unit_energies_match_energy_test(oracle, name):
    xyz, box, top, p, terms, ns = load_crystal(name)
    ref = expected(name)
    assert top.n == ns * ref["atoms"]
    e = oracle.evaluate(oracle.System(xyz, box, top), p)[0]
    assert abs(e["vdw"] / ns - ref["vdw"][0]) <= 1e-8
    assert abs(e["perm"] / ns - ref["perm"][0]) <= ref["tolerance"]
    if "pol" in ref:
        assert abs(e["pol"] / ns - ref["pol"][0]) <= ref["tolerance"]
    v, g = oracle.valence(terms, xyz)
    counts = terms.counts()
    for k in KINDS:
        assert counts[k] == ns * ref["valence"][k][1], k
        tol = ref["tolerance"] if k == "angle" else 1e-8
        assert abs(v[k] / ns - ref["valence"][k][0]) <= tol, (k, v[k] / ns, ref["valence"][k][0])
    v = oracle.valence(in_plane_at_trigonal_centres(terms, top.n), xyz)[0]
    assert abs(v["angle"] / ns - ref["valence"]["angle"][0]) <= 1e-8


@pytest.mark.parametrize("group", sorted(SPACE_GROUPS))
def test_operators_form_a_group_modulo_lattice(group):
    """Composing any two operators gives a third, up to a lattice vector."""
    ops = [(np.array(r, float), np.array(t)) for r, t in SPACE_GROUPS[group]]
    for r1, t1 in ops:
        for r2, t2 in ops:
            r, t = r1 * r2, r1 * t2 + t1
            assert any(np.array_equal(r, r3) and np.allclose((t - t3) - np.round(t - t3), 0.0) for r3, t3 in ops)


def test_expansion_keeps_molecules_whole_and_preserves_internal_geometry():
    xf = tinker.read_xyz(os.path.join(CRYSTALS, "acetanilide.xyz"))
    box = [19.640, 9.483, 7.979, 90.0, 90.0, 90.0]
    x, t, b, names, ns = expand_to_p1(xf.xyz, xf.types, xf.bonds, xf.names, box, "Pbca")
    n = len(xf.names)
    assert ns == 8 and x.shape == (8 * n, 3) and len(b) == 8 * len(xf.bonds)
    d0 = np.linalg.norm(xf.xyz[:, None] - xf.xyz[None], axis=-1)
    for k in range(ns):
        c = x[k * n:(k + 1) * n]
        assert np.abs(np.linalg.norm(c[:, None] - c[None], axis=-1) - d0).max() <= 1e-9
    # the inversion copy is the asymmetric unit through the origin
    assert np.abs(x[4 * n:5 * n] + xf.xyz).max() <= 1e-9
    with pytest.raises(ValueError):
        expand_to_p1(xf.xyz, xf.types, xf.bonds, xf.names, box, "R-3")


def test_cell_matrix_reproduces_lengths_and_angles():
    box = [10.0, 12.0, 14.0, 80.0, 95.0, 110.0]
    m = cell_matrix(box)
    ln = np.linalg.norm(m, axis=1)
    assert np.allclose(ln, box[:3])
    ang = [np.degrees(np.arccos(m[i] @ m[j] / (ln[i] * ln[j]))) for i, j in ((1, 2), (0, 2), (0, 1))]
    assert np.allclose(ang, box[3:])
    assert m[0, 1] == m[0, 2] == m[1, 2] == 0.0


def test_folded_gradient_central_differences(oracle):
    """The per-unit gradient of the whole nonbonded energy with respect to
    asymmetric-unit coordinates (all images moving with it), against central
    differences of the P1 energy / nsymm, on a monoclinic P21/c cell."""
    import dataclasses
    from forcefieldx_amd.crystal import fold_gradient
    name = "paracetamol"
    files = [os.path.join(CRYSTALS, name + ext) for ext in (".xyz", ".key", ".prm")]
    xf = tinker.read_xyz(files[0])
    xyz, box, top, p, terms, ns = tinker.load_crystal(*files)
    p = dataclasses.replace(p, polar_eps=1e-11)
    assert box[4] != 90.0

    def energy(asu):
        x = expand_to_p1(asu, xf.types, xf.bonds, xf.names, box, "P21/c")[0]
        return oracle.evaluate(oracle.System(x, box, top), p)[0]["total"] / ns

    e, g = oracle.evaluate(oracle.System(xyz, box, top), p)[:2]
    gu = fold_gradient(g, box, "P21/c")
    h = 1e-5
    for i in (0, 7, 13, 19):
        for d in range(3):
            xp = xf.xyz.copy(); xp[i, d] += h
            xm = xf.xyz.copy(); xm[i, d] -= h
            fd = (energy(xp) - energy(xm)) / (2 * h)
            assert abs(gu[i, d] - fd) <= 1e-5 * max(1.0, abs(fd)), (i, d, gu[i, d], fd)
