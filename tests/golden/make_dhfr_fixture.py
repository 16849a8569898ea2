"""Writes the DHFR fixture of the reference's EnergyTest ("DHFR Benchmark",
modules/potential/src/test/java/ffx/potential/commands/EnergyTest.java:148-180):
the coordinates and keywords the test reads (modules/potential/src/test/
resources/dhfr.xyz, dhfr.key), the records of the AMOEBA_BIO_2009 parameter
file (modules/potential/src/main/java/ffx/potential/parameters/ff/
AMOEBA_BIO_2009) for the atom types DHFR uses, in the same Tinker formats,
and the test's expected nonbonded energies and interaction counts.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_dhfr_fixture.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from forcefieldx_amd import tinker  # noqa: E402

REF = "/root/reference/modules/potential/src"


def main():
    out = os.path.join(HERE, "dhfr")
    os.makedirs(out, exist_ok=True)
    xf = tinker.read_xyz(REF + "/test/resources/dhfr.xyz")
    keys = tinker.read_keys(REF + "/test/resources/dhfr.key")
    ff = tinker.read_prm(REF + "/main/java/ffx/potential/parameters/ff/AMOEBA_BIO_2009")
    tinker.write_xyz(os.path.join(out, "dhfr.xyz.gz"), xf)
    with open(os.path.join(out, "dhfr.key"), "w") as f:
        for k, v in keys.items():
            f.write(f"{k:<20s}{' '.join(v)}\n")
    tinker.write_prm(os.path.join(out, "amoeba_bio_2009_dhfr.prm.gz"), ff, types=set(int(t) for t in xf.types))
    expected = {
        "source": "EnergyTest.java:148-180 (DHFR Benchmark), tolerance 1.0e-2 (ParentEnergyTest.java:101)",
        "atoms": 23558,
        "vdw": 31718.84803063, "n_vdw": 3480619,
        "perm": -78752.17402965919, "n_perm": 1463589,
        "pol": -31790.15161303, "n_pol": 1463589,
        "tolerance": 1.0e-2,
    }
    with open(os.path.join(out, "expected.json"), "w") as f:
        json.dump(expected, f, indent=1)


if __name__ == "__main__":
    main()
