"""Writing synthetic systems as Tinker files (forcefieldx_amd.tinker.
write_system, systems.write_named_config) and reading them back.

The reference writes a coordinate file with its key file and parameter
patch (XYZFilter.writeFile, XYZFilter.java:639; KeyFilter.java) and reads
them back through the same parser as any other input.  Here each system is
written, read back by tinker.load (the reader the DHFR and ubiquitin
fixtures go through) and held to:

* the same Topology arrays: frames and axis atoms by the reference's
  multipole matching rules, local moments after the Bohr round trip,
  polarizabilities and groups, vdW classes, pair table and reductions,
  masks, molecules and soft atoms;
* the same settings (every NonbondedParams field that has a keyword);
* the same oracle energies, gradient and induced dipoles at the written
  coordinates (1e-12), and for the soft chain the same dU/dL and dU/dX/dL,
  so that a dropped unit conversion, a frame written with the wrong signs,
  a missing group type or reduction shows as an energy difference; a
  planted mistake in the written patch shows that the comparison sees one.
"""
import dataclasses

import numpy as np
import pytest

from forcefieldx_amd import tinker
from forcefieldx_amd.forcefield import (FRAME_BISECTOR, FRAME_ZONLY, FRAME_ZTHENX, NonbondedParams,
                                        masks_from_bonds)
from forcefieldx_amd.systems import (chain_in_water, crystal_batch, elements_of, named_config, water_box,
                                     write_named_config)

INT_FIELDS = ("frame", "axis", "vdw_class", "red_index", "mask_off", "mask_atom", "mask_rel", "soft", "molecule")
FLOAT_FIELDS = ("mpole", "polarity", "thole", "pdamp", "vdw_table", "red_value")
ENERGY_KEYS = ("perm_real", "perm_self", "perm_recip", "pol_real", "pol_self", "pol_recip", "vdw", "total")
REL = 1e-12


def rel_err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def groups(g):
    first = {}
    return np.array([first.setdefault(v, a) for a, v in enumerate(np.asarray(g).tolist())])


def round_trip(tmp_path, name, xyz, box, top, p, elements=None):
    out = tinker.write_system(str(tmp_path / name), xyz, box, top, p, title=name, elements=elements)
    x2, b2, t2, p2 = tinker.load(out["xyz"], out["key"], alchemical_atoms=out["alchemical_atoms"],
                                 lambda_value=out["lambda_value"])
    assert b2 == [float(v) for v in box]
    assert np.abs(x2.reshape(-1) - np.asarray(xyz).reshape(-1)).max() <= 5.1e-9
    for k in INT_FIELDS:
        assert np.array_equal(getattr(t2, k), getattr(top, k)), k
    for k in FLOAT_FIELDS:
        a, b = np.asarray(getattr(t2, k)), np.asarray(getattr(top, k))
        assert np.allclose(a, b, rtol=4e-16, atol=1e-17), (k, np.abs(a - b).max())
    assert np.array_equal(groups(t2.group), groups(top.group))
    assert p2 == p
    return x2.reshape(-1), t2, p2, out


def energies_agree(oracle, xyz, box, top_a, top_b, p):
    ea, ga, ua, _ = oracle.evaluate(oracle.System(xyz, box, top_a), p)
    eb, gb, ub, _ = oracle.evaluate(oracle.System(xyz, box, top_b), p)
    scale = max(abs(ea[k]) for k in ("perm", "pol", "vdw"))
    for k in ENERGY_KEYS:
        assert abs(ea[k] - eb[k]) <= REL * scale, (k, ea[k], eb[k])
    assert ea["pcg_iters"] == eb["pcg_iters"] and ea["n_vdw"] == eb["n_vdw"]
    assert rel_err(gb, ga) <= REL and rel_err(ub, ua) <= REL
    return ea


def test_water648_round_trip(oracle, tmp_path):
    (name, xyz, box, top, p), = named_config("water-648")
    x2, t2, p2, out = round_trip(tmp_path, name, xyz, box, top, p, elements_of(top))
    xf = tinker.read_xyz(out["xyz"])
    # two types, O and H, each with one multipole record: the oxygen's a
    # bisector on its two hydrogens, the hydrogen's a Z-then-X frame on its
    # oxygen and the other hydrogen
    assert sorted(set(xf.types.tolist())) == [1, 2] and xf.names[:3] == ["O", "H", "H"]
    ff = tinker.read_prm(out["prm"])
    assert ff.multipole[(1, 2, 2)][0] == FRAME_BISECTOR and ff.multipole[(2, 1, 2)][0] == FRAME_ZTHENX
    assert ff.multipole[(1, 2, 2)][1][0] == -0.51966 and ff.multipole[(1, 2, 2)][1][3] == 0.14279
    assert ff.polarize == {1: (0.837, 0.39, (2,)), 2: (0.496, 0.39, (1,))}
    assert ff.vdw == {1: (3.405, 0.11, 0.0), 2: (2.655, 0.0135, 0.91)} and not ff.vdwpr
    assert tinker.read_keys(out["key"])["pme-grid"] == ["24", "24", "24"]
    e = energies_agree(oracle, x2, box, top, t2, p2)
    assert e["pcg_iters"] > 0


def test_planted_patch_mistake_changes_the_energy(oracle, tmp_path):
    """The energy comparison sees a mistake in a written patch: the water
    hydrogen's polarize record without its group type leaves every atom a
    group of its own, which changes the polarization energy."""
    xyz, box, top = water_box(27, 9.32, seed=3)
    p = NonbondedParams(vdw_cutoff=4.6, ewald_cutoff=4.6, grid=(16, 16, 16))
    x2, t2, p2, out = round_trip(tmp_path, "w81", xyz, box, top, p, elements_of(top))
    text = open(out["prm"]).read()
    bad = text.replace("polarize      2        0.496         0.39      1", "polarize      2        0.496         0.39")
    assert bad != text
    with open(out["prm"], "w") as f:
        f.write(bad)
    t3 = tinker.load(out["xyz"], out["key"])[2]
    assert not np.array_equal(groups(t3.group), groups(top.group))
    ea, _, _, _ = oracle.evaluate(oracle.System(x2, box, top), p2)
    eb, _, _, _ = oracle.evaluate(oracle.System(x2, box, t3), p2)
    assert abs(ea["pol"] - eb["pol"]) > 1e-3 * abs(ea["pol"])


def test_soft_chain_round_trip_with_lambda(oracle, tmp_path):
    """A chain with one atom type per atom (random moments), 6 soft atoms and
    every keyword set away from its default."""
    xyz, box, top = chain_in_water(40, 150, 18.0, seed=3, n_soft=6)
    p = NonbondedParams(ewald_cutoff=6.5, ewald_alpha=0.6, vdw_cutoff=8.0, vdw_taper=0.85, list_buffer=1.5,
                        grid=(25, 24, 28), polar_eps=3e-6, precond_cutoff=4.0, precond_scale=1.5,
                        m_scale=(0.0, 0.0, 0.5, 0.9), p_scale=(0.0, 0.1, 0.9, 1.0), intra14=0.4, d11=0.1,
                        v_scale=(0.0, 0.1, 0.9), halgren_delta=0.071, halgren_gamma=0.121, lambda_value=0.45,
                        vdw_lambda_alpha=0.3, vdw_lambda_exponent=2.5, vdw_lambda_end=0.9,
                        intermolecular_softcore=True, perm_lambda_alpha=1.5, perm_lambda_exponent=2.0,
                        perm_lambda_start=0.35, perm_lambda_end=0.95, pol_lambda_exponent=2.5,
                        pol_lambda_start=0.7, pol_lambda_end=0.98, elec_lambda=True, ligand_vapor_elec=False,
                        scf_algorithm=1, polar_sor=0.65, sor_fallback=False)
    x2, t2, p2, out = round_trip(tmp_path, "chain", xyz, box, top, p)
    assert out["alchemical_atoms"] == "18-23" and out["lambda_value"] == 0.45
    assert len(set(tinker.read_xyz(out["xyz"]).types.tolist())) == 40 + 2
    # the same settings with no keyword stay at the library's defaults
    assert (p2.max_scf_iter, p2.precision) == (p.max_scf_iter, p.precision)
    ea, ga, la = oracle.evaluate_lambda(oracle.System(x2, box, top), p2)
    eb, gb, lb = oracle.evaluate_lambda(oracle.System(x2, box, t2), p2)
    scale = max(abs(ea[k]) for k in ("perm", "pol", "vdw"))
    for k in ENERGY_KEYS + ("dudl", "d2udl2"):
        assert abs(ea[k] - eb[k]) <= REL * scale, (k, ea[k], eb[k])
    assert rel_err(gb, ga) <= REL and rel_err(lb, la) <= REL
    assert abs(ea["dudl"]) > 0.0


def test_crystal_cells_round_trip(oracle, tmp_path):
    """Triclinic cells of star molecules: one type per atom, frames on 1-2
    and 1-3 atoms, one polarization group per molecule."""
    for k, (xyz, box, top, grid) in enumerate(crystal_batch(3, seed=5)):
        p = NonbondedParams(ewald_cutoff=6.0, vdw_cutoff=6.0, polar_eps=1e-5, grid=tuple(grid))
        x2, t2, p2, _ = round_trip(tmp_path, f"cell{k}", xyz, box, top, p, elements_of(top))
        assert len(np.unique(t2.molecule)) == len(np.unique(top.group))
        energies_agree(oracle, x2, box, top, t2, p2)


def test_named_configs_write_and_read(tmp_path):
    """chain-23557 at its full size and the first cells of crystals-512 go
    through write_named_config and read back to the same topology."""
    (name, xyz, box, top, p), = named_config("chain-23557")
    res = write_named_config("chain-23557", str(tmp_path / "c3"))
    assert len(res) == 1 and res[0]["alchemical_atoms"] == "1241-1260" and res[0]["lambda_value"] == 0.5
    x2, b2, t2, p2 = tinker.load(res[0]["xyz"], res[0]["key"], alchemical_atoms=res[0]["alchemical_atoms"],
                                 lambda_value=0.5)
    assert t2.n == 23557 and p2 == p and int(t2.soft.sum()) == 20
    for k in INT_FIELDS:
        assert np.array_equal(getattr(t2, k), getattr(top, k)), k
    for k in FLOAT_FIELDS:
        assert np.allclose(getattr(t2, k), getattr(top, k), rtol=4e-16, atol=1e-17), k
    cells = named_config("crystals-512")
    assert len(cells) == 512 and cells[0][0] == "cell000"
    for label, xyz, box, top, p in cells[:4]:
        round_trip(tmp_path, label, xyz, box, top, p, elements_of(top))


def test_atom_ranges():
    assert tinker.parse_atom_ranges("1-3,7,9-N", 10).tolist() == [0, 1, 2, 6, 8, 9]
    assert tinker.parse_atom_ranges("ALL", 3).tolist() == [0, 1, 2]
    assert tinker.parse_atom_ranges("none", 3).size == 0
    assert tinker.format_atom_ranges([0, 1, 2, 6, 8, 9]) == "1-3,7,9-10"
    assert tinker.format_atom_ranges([]) == "NONE"
    with pytest.raises(ValueError):
        tinker.parse_atom_ranges("4-12", 10)


def test_unwritable_topologies_are_refused():
    """A frame on a 1-4 atom and a polarization group that bonds do not
    connect have no multipole / polarize records that give them back."""
    xyz, box, top = chain_in_water(8, 0, 20.0, seed=1)
    bad = dataclasses.replace(top, axis=top.axis.copy())
    bad.axis[0, 1] = 3               # x axis atom three bonds away
    with pytest.raises(ValueError):
        tinker.system_types(bad, elements_of(bad))
    bad = dataclasses.replace(top, group=top.group.copy())
    bad.group[5] = bad.group[0]      # atoms 1 and 6 in one group, not bonded through it
    with pytest.raises(ValueError):
        tinker.system_types(bad, elements_of(bad))
    # a Z-only frame on a bonded atom is expressible
    ok = dataclasses.replace(top, frame=top.frame.copy(), axis=top.axis.copy())
    ok.frame[3], ok.axis[3] = FRAME_ZONLY, (4, -1, -1)
    types, ff, bonds, _ = tinker.system_types(ok, elements_of(ok))
    assert len(bonds) == 7 and np.array_equal(masks_from_bonds(8, bonds)[2], top.mask_rel)
