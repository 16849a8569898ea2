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

It also writes tests/golden/crystals: the asymmetric units, keywords and
parameter patches of the five crystal benchmarks (EnergyTest.java:435-614,
modules/potential/src/test/resources/<name>.xyz, .properties, .patch) with
the tests' per-asymmetric-unit energies and counts; and tests/golden/snare:
the asymmetric unit of the SNARE protein crystal in P212121 (EnergyTest.java:
74-109, "SNARE P212121", 1n7s.P212121.xyz / .key) with the records of the
AMOEBA_PROTEIN_2004 parameter file its atom types use.

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


# P1 crystals of the reference's EnergyTest (EnergyTest.java:435-615): the
# asymmetric unit, its keywords (space group and cell) and its parameter file
CRYSTALS = {
    "acetanilide": (":435-470 (Acetanilide Benchmark)", "acetanilide.patch",
                    {"bond": [0.67814832, 19], "angle": [1.06817134, 30], "strbnd": [-0.06290501, 26],
                     "urey": [0.0, 0], "opbend": [0.13033970, 24], "torsion": [-8.01262794, 38],
                     "pitors": [0.0, 0], "tortor": [0.0, 0]},
                    (-3.62572228, 7295), (-26.0720131242538, 2231), (-1.9856761645305, 2231)),
    "ethylparaben": (":471-506 (Ethylparaben Benchmark)", "ethylparaben.patch",
                     {"bond": [1.37991170, 44], "angle": [2.43600964, 70], "strbnd": [0.20006782, 58],
                      "urey": [0.0, 0], "opbend": [0.04434686, 42], "torsion": [-7.14298308, 88],
                      "pitors": [0.0, 0], "tortor": [0.0, 0]},
                     (-4.816366774263829, 32644), (-45.55012417320137, 5012), None),
    "methylparaben": (":507-542 (Methylparaben Benchmark)", "methylparaben.patch",
                      {"bond": [0.42501694, 19], "angle": [1.23611203, 29], "strbnd": [0.00351228, 24],
                       "urey": [0.0, 0], "opbend": [0.01058196, 21], "torsion": [-3.72750990, 35],
                       "pitors": [0.0, 0], "tortor": [0.0, 0]},
                      (0.20573979, 7639), (-26.558577471708134, 2314), (-2.35904592, 2314)),
    "paracetamol": (":543-578 (Paracetamol Benchmark)", "paracetamol.patch",
                    {"bond": [0.55316710, 20], "angle": [1.51848706, 31], "strbnd": [0.05096337, 27],
                     "urey": [0.0, 0], "opbend": [0.46651316, 24], "torsion": [-7.40125527, 40],
                     "pitors": [0.0, 0], "tortor": [0.0, 0]},
                    (0.76621272, 7832), (-32.02011297249507, 2357), (-4.96649992387851, 2357)),
    "phenacetin": (":579-614 (Phenacetin Benchmark)", "phenacetin.patch",
                   {"bond": [0.56492623, 26], "angle": [2.68273715, 43], "strbnd": [0.00409248, 35],
                    "urey": [0.0, 0], "opbend": [0.33385711, 24], "torsion": [-6.49787814, 52],
                    "pitors": [0.0, 0], "tortor": [0.0, 0]},
                   (-5.792072251098926, 20384), (-25.42173449818894, 3138), None),
}


def write_crystals():
    out = os.path.join(HERE, "crystals")
    os.makedirs(out, exist_ok=True)
    expected = {}
    for name, (lines, patch, valence, vdw, perm, pol) in CRYSTALS.items():
        xf = tinker.read_xyz(REF + f"/test/resources/{name}.xyz")
        keys = tinker.read_keys(REF + f"/test/resources/{name}.properties")
        ff = tinker.read_prm(REF + f"/test/resources/{patch}")
        tinker.write_xyz(os.path.join(out, f"{name}.xyz"), xf, digits=8)
        with open(os.path.join(out, f"{name}.key"), "w") as f:
            for k, v in keys.items():
                if k != "parameters":
                    f.write(f"{k:<20s}{' '.join(v)}\n")
        tinker.write_prm(os.path.join(out, f"{name}.prm"), ff)
        e = {"source": "EnergyTest.java" + lines + ", tolerance 1.0e-2 (ParentEnergyTest.java:101)",
             "atoms": len(xf.names), "valence": valence, "vdw": list(vdw), "perm": list(perm), "tolerance": 1.0e-2}
        if pol is not None:
            e["pol"] = list(pol)
        expected[name] = e
    with open(os.path.join(out, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1)


SNARE = {
    "source": "EnergyTest.java:74-109 (SNARE P212121), per asymmetric unit, tolerance 1.0e-2 "
              "(ParentEnergyTest.java:101); the P1 row (EnergyTest.java:110-145) holds four times these",
    "atoms": 5357, "spacegroup": "P212121",
    "valence": {"bond": [351.32142483, 5040], "angle": [744.19251364, 8255], "strbnd": [6.42015744, 6924],
                "urey": [2.50163831, 322], "opbend": [135.24919366, 2487], "torsion": [417.89244418, 11449],
                "pitors": [39.85643934, 370], "tortor": [-560.99576469, 268]},
    "vdw": 4003.27183547, "n_vdw": 741643,
    "perm": -12305.43400643136, "n_perm": 332114,
    "pol": -2811.45683671, "n_pol": 332114,
}


def write_snare():
    out = os.path.join(HERE, "snare")
    os.makedirs(out, exist_ok=True)
    ff = tinker.read_prm(REF + "/main/java/ffx/potential/parameters/ff/AMOEBA_PROTEIN_2004")
    xf = tinker.read_xyz(REF + "/test/resources/1n7s.P212121.xyz")
    keys = tinker.read_keys(REF + "/test/resources/1n7s.P212121.key")
    tinker.write_xyz(os.path.join(out, "1n7s.xyz.gz"), xf)
    with open(os.path.join(out, "1n7s.key"), "w") as f:
        for k, v in keys.items():
            f.write(f"{k:<20s}{' '.join(v)}\n")
    tinker.write_prm(os.path.join(out, "amoeba_protein_2004_1n7s.prm.gz"), ff, types=set(int(t) for t in xf.types))
    with open(os.path.join(out, "expected.json"), "w") as f:
        json.dump(dict(SNARE, tolerance=1.0e-2), f, indent=1)


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
    write_crystals()
    write_snare()
