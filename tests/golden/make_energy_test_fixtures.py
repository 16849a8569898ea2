"""Writes fixtures of two periodic AMOEBA-BIO-2009 systems of the reference's
EnergyTest (modules/potential/src/test/java/ffx/potential/commands/
EnergyTest.java): "DHFR Benchmark" (:148-180) and "Ubiquitin Benchmark"
(:363-394).  Each holds the coordinates and keywords the test reads
(modules/potential/src/test/resources/<name>.xyz, <name>.key), the records of
the AMOEBA_BIO_2009 parameter file (modules/potential/src/main/java/ffx/
potential/parameters/ff/AMOEBA_BIO_2009) for the atom types and classes the
system uses, in the same Tinker formats, and the test's expected valence and
nonbonded energies and term / interaction counts (tolerance 1.0e-2 kcal/mol,
ParentEnergyTest.java:101).

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_energy_test_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from forcefieldx_amd import tinker  # noqa: E402

REF = "/root/reference/modules/potential/src"

SYSTEMS = {
    "dhfr": {
        "source": "EnergyTest.java:148-180 (DHFR Benchmark), tolerance 1.0e-2 (ParentEnergyTest.java:101)",
        "atoms": 23558,
        "valence": {"bond": [6623.70878543, 16569], "angle": [4748.78155255, 11584],
                    "strbnd": [-19.78973386, 4031], "urey": [-132.74781407, 7023],
                    "opbend": [110.84355342, 1566], "torsion": [433.69593807, 6701],
                    "pitors": [58.75073806, 292], "tortor": [-41.99981847, 147]},
        "vdw": 31718.84803063, "n_vdw": 3480619,
        "perm": -78752.17402965919, "n_perm": 1463589,
        "pol": -31790.15161303, "n_pol": 1463589,
    },
    "ubiquitin": {
        "source": "EnergyTest.java:363-394 (Ubiquitin Benchmark), tolerance 1.0e-2 (ParentEnergyTest.java:101); "
                  "the polarization energy the test lists for grids with factors 6 and 7",
        "atoms": 9737,
        "valence": {"bond": [2856.59378034, 6908], "angle": [2018.16571190, 5094],
                    "strbnd": [-6.98794687, 1958], "urey": [-55.63461732, 2835],
                    "opbend": [65.00809012, 651], "torsion": [224.37608631, 3297],
                    "pitors": [20.43063520, 106], "tortor": [-31.48011891, 71]},
        "vdw": 12862.34983956, "n_vdw": 1482946,
        "perm": -32737.101745561464, "n_perm": 623627,
        "pol": -12934.94219259, "n_pol": 623627,
    },
}


def main():
    ff = tinker.read_prm(REF + "/main/java/ffx/potential/parameters/ff/AMOEBA_BIO_2009")
    for name, exp in SYSTEMS.items():
        out = os.path.join(HERE, name)
        os.makedirs(out, exist_ok=True)
        xf = tinker.read_xyz(REF + f"/test/resources/{name}.xyz")
        keys = tinker.read_keys(REF + f"/test/resources/{name}.key")
        tinker.write_xyz(os.path.join(out, f"{name}.xyz.gz"), xf)
        with open(os.path.join(out, f"{name}.key"), "w") as f:
            for k, v in keys.items():
                f.write(f"{k:<20s}{' '.join(v)}\n")
        tinker.write_prm(os.path.join(out, f"amoeba_bio_2009_{name}.prm.gz"), ff,
                         types=set(int(t) for t in xf.types))
        with open(os.path.join(out, "expected.json"), "w") as f:
            json.dump(dict(exp, tolerance=1.0e-2), f, indent=1)


if __name__ == "__main__":
    main()
