"""Two periodic AMOEBA-BIO-2009 systems of the reference's EnergyTest on the
GPU, read from the committed Tinker fixtures: "DHFR Benchmark"
(EnergyTest.java:148-180, 23,558 atoms, PME 80^3) and "Ubiquitin Benchmark"
(:363-394, 9,737 atoms, PME 70 x 54 x 54, whose radix-7 and radix-3 line
transforms this exercises).  The FP64 form reproduces each test's nonbonded
energies within its 1e-2 kcal/mol tolerance and its interaction counts
exactly, by SOR and by PCG; the mixed form's forces are within 1e-5 of FP64."""
import dataclasses

import numpy as np
import pytest

from forcefieldx_amd import ffi
from test_tinker import expected, load_fixture

pytestmark = pytest.mark.gpu


def run(prec, name="dhfr", **kw):
    xyz, box, top, p = load_fixture(name)
    p = dataclasses.replace(p, **kw)
    h = ffi.Handle(ffi.params_from(p, p.grid, prec), top)
    try:
        h.set_cell(box)
        h.set_positions(xyz.reshape(-1))
        e, rc = h.evaluate(ffi.NB_GRAD)
        return e, h.gradient().reshape(-1, 3)
    finally:
        h.close()


@pytest.mark.parametrize("name,scf", [("dhfr", 1), ("dhfr", 0), ("ubiquitin", 0), ("ubiquitin", 1)],
                         ids=["dhfr-sor", "dhfr-pcg", "ubiquitin-pcg", "ubiquitin-sor"])
def test_energy_test_fp64(name, scf):
    print(f"{name} FP64 scf={scf}", flush=True)
    e, g = run(ffi.NB_PREC_FP64, name, scf_algorithm=scf)
    ref = expected(name)
    for k in ("vdw", "perm", "pol"):
        assert abs(e[k] - ref[k]) <= ref["tolerance"], (k, e[k], ref[k])
    assert e["n_vdw"] == ref["n_vdw"] and e["n_perm"] == ref["n_perm"]
    assert abs(e["vdw"] - ref["vdw"]) <= 1e-6 and abs(e["perm"] - ref["perm"]) <= 1e-6


def test_ubiquitin_matches_oracle(oracle):
    """Gradient and induced dipoles at 1e-8 of the oracle on the 70 x 54 x 54
    grid."""
    xyz, box, top, p = load_fixture("ubiquitin")
    eo, go, uo, _ = oracle.evaluate(oracle.System(xyz, box, top), p)
    h = ffi.Handle(ffi.params_from(p, p.grid, ffi.NB_PREC_FP64), top)
    try:
        h.set_cell(box)
        h.set_positions(xyz.reshape(-1))
        e, rc = h.evaluate(ffi.NB_GRAD)
        g = h.gradient().reshape(-1, 3)
        u, _ = h.induced()
    finally:
        h.close()
    assert e["pcg_iters"] == eo["pcg_iters"]
    for k in ("perm_real", "perm_recip", "pol", "vdw", "total"):
        assert abs(e[k] - eo[k]) <= 1e-8 * abs(eo["perm"]), (k, e[k], eo[k])
    assert np.linalg.norm(g - go) <= 1e-8 * np.linalg.norm(go)
    assert np.linalg.norm(u - uo) <= 1e-8 * np.linalg.norm(uo)


@pytest.mark.parametrize("name", ["dhfr", "ubiquitin"])
def test_mixed_against_fp64(name):
    e64, g64 = run(ffi.NB_PREC_FP64, name, scf_algorithm=0, polar_eps=1e-5)
    emx, gmx = run(ffi.NB_PREC_MIXED, name, scf_algorithm=0, polar_eps=1e-5)
    err = np.linalg.norm(gmx - g64) / np.linalg.norm(g64)
    print(f"{name} mixed force rel RMS err {err:.2e}", flush=True)
    assert err <= 1e-5
    ref = expected(name)
    assert abs(emx["vdw"] - ref["vdw"]) <= ref["tolerance"]
    assert abs(emx["perm"] - ref["perm"]) <= 0.1
