"""Host-side mirror of the reference's molecular-dynamics driver for the
velocity-Verlet integrator (MolecularDynamics.dynamic, MolecularDynamics.java:
1285-1299; VelocityVerlet.java:66-113) over the device integrator of the HIP
library (nb_md_*): positions, velocities and accelerations stay in device
memory, each step's evaluation is the ForceFieldEnergy handle's (a graph
replay in graph mode), and the SCF predictor carries the induced dipoles from
step to step (ParticleMeshEwald's scfPredictor, ASPC)."""
from __future__ import annotations

import numpy as np

from . import ffi

KCAL_TO_GRAM_ANG2_PER_PS2 = 418.40     # Constants.java:205
KB = 0.0019872042586408316             # kcal/mol/K (Constants.java, R / 1000 in kcal)


class MolecularDynamics:
    """Velocity Verlet on a ForceFieldEnergy: adiabatic (NVE) by default, or
    with the Berendsen thermostat."""

    def __init__(self, ffe, masses, scf_predictor=ffi.NB_PRED_ASPC, graph=True, berendsen=None):
        """berendsen: (temperature K, tau ps) for the Berendsen thermostat
        (Berendsen.java:99-101), None for NVE."""
        self.ffe = ffe
        self.mass = np.ascontiguousarray(masses, np.float64)
        self._h = ffe._handle
        if scf_predictor is not None:
            self._h.set_scf_predictor(scf_predictor)
        if graph:
            self._h.set_graph(True)
        if berendsen is not None:
            self._h.md_set_thermostat(ffi.NB_THERMO_BERENDSEN, *berendsen)
        self._report = None

    def initVelocities(self, temperature, seed=0):
        """Maxwell-Boltzmann velocities (A/ps) at the temperature (K), with the
        centre-of-mass motion removed (Thermostat.maxwell, Thermostat.java)."""
        rng = np.random.default_rng(seed)
        sd = np.sqrt(KB * temperature * KCAL_TO_GRAM_ANG2_PER_PS2 / self.mass)
        v = rng.normal(size=(len(self.mass), 3)) * sd[:, None]
        v -= (self.mass[:, None] * v).sum(axis=0) / self.mass.sum()
        return v

    def init(self, x, v):
        """Start from coordinates x (n, 3) and velocities v (n, 3)."""
        self.ffe.setCoordinates(np.asarray(x, float).reshape(-1))
        self._h.md_init(self.mass, np.asarray(v, float).reshape(-1))

    def dynamic(self, nSteps, timeStep):
        """nSteps steps of timeStep femtoseconds; returns the last report."""
        self._report = self._h.md_step(nSteps, timeStep * 1.0e-3)
        self.ffe._last = self._report["last"]
        return self._report

    def getCoordinates(self):
        return self._h.md_state()[0]

    def getVelocities(self):
        return self._h.md_state()[1]

    def getKineticEnergy(self):
        return self._report["kinetic"]

    def getPotentialEnergy(self):
        return self._report["potential"]

    def getTotalEnergy(self):
        return self._report["total"]
