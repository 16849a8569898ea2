"""Host-side mirror of the reference's energy API for the nonbonded path.

``ForceFieldEnergy`` implements ``energy(x)`` / ``energyAndGradient(x, g)``
(Potential interface, OptimizationInterface.java:69-129;
ForceFieldEnergy.java:1806-1893) over the HIP library, and exposes the
per-term objects ``ParticleMeshEwald`` and ``VanDerWaals`` with the
reference's getters (ParticleMeshEwald.java:1014-1127, VanDerWaals.java:
573-583).  Errors from the library are raised as ``EnergyException`` the way
the reference throws them (RealSpaceEnergyRegion.java:384-391,
PCGSolver.java:553-569, ReduceRegion.java:208-229); on an SCF failure the
caller-visible policy of ParticleMeshEwald.selfConsistentField
(:2300-2325) is applied: retry with direct polarization when
``direct_fallback`` is set.
"""
from __future__ import annotations

import numpy as np

from . import ffi
from .forcefield import NonbondedParams


class EnergyException(RuntimeError):
    """ffx.potential.utils.EnergyException"""


class VanDerWaals:
    def __init__(self, ffe):
        self._ffe = ffe

    def getEnergy(self):
        return self._ffe._last["vdw"]

    def getInteractions(self):
        return int(self._ffe._last["n_vdw"])

    def energy(self, gradient=False, print_=False):
        self._ffe._evaluate(gradient)
        return self.getEnergy()


class ParticleMeshEwald:
    def __init__(self, ffe):
        self._ffe = ffe

    def energy(self, gradient=False, print_=False):
        self._ffe._evaluate(gradient)
        return self.getTotalElectrostaticEnergy()

    def _e(self, k):
        return self._ffe._last[k]

    def getPermanentEnergy(self):
        return self._e("perm")

    def getPermRealEnergy(self):
        return self._e("perm_real")

    def getPermSelfEnergy(self):
        return self._e("perm_self")

    def getPermRecipEnergy(self):
        return self._e("perm_recip")

    def getPolarizationEnergy(self):
        return self._e("pol")

    def getIndRealEnergy(self):
        return self._e("pol_real")

    def getIndSelfEnergy(self):
        return self._e("pol_self")

    def getIndRecipEnergy(self):
        return self._e("pol_recip")

    def getTotalElectrostaticEnergy(self):
        return self._e("perm") + self._e("pol")

    def getInteractions(self):
        return int(self._e("n_perm"))

    def getScfIterations(self):
        return int(self._e("pcg_iters"))

    def getInducedDipoles(self):
        """(u, uCR) in caller order, (n, 3) each."""
        return self._ffe._handle.induced()


class EnergyTerm:
    """One valence kind's share of the last evaluation (terms/EnergyTerm.java:
    getEnergy :182, getNumberOfTerms :114)."""

    def __init__(self, ffe, kind, count):
        self._ffe, self._kind, self._count = ffe, kind, count

    def getEnergy(self):
        return self._ffe._last["valence_" + self._kind]

    def getNumberOfTerms(self):
        return self._count


class ForceFieldEnergy:
    """ForceFieldEnergy on the MI355X library: the nonbonded terms and, when
    given, the AMOEBA valence terms."""

    def __init__(self, topology, box, params: NonbondedParams | None = None, precision=None, device=0,
                 direct_fallback=None, finite_pcg_gradient=False, valence=None):
        """finite_pcg_gradient: gradients of the energy of the dipoles the PCG
        returned (NB_FINITE_PCG) instead of the stationary-dipole gradient.
        valence: forcefieldx_amd.bonded.ValenceTerms evaluated with every
        energy (the bonded-term region of ForceFieldEnergy.java:1670-1677)."""
        self.params = params or NonbondedParams()
        self.finite_pcg_gradient = bool(finite_pcg_gradient)
        self.topology = topology
        self.box = np.asarray(box if box is not None else [0, 0, 0, 90, 90, 90], float)
        grid = self.params.resolved_grid(self.box) if self.box[0] > 0 else (16, 16, 16)
        self.grid = grid
        self._nbp = ffi.params_from(self.params, grid, precision, device)
        self._handle = ffi.Handle(self._nbp, topology, device)
        self._handle.set_cell(self.box)
        self._x = None
        self._last = None
        self._lambda = None     # set by setLambda: evaluations then carry dU/dL
        self._fallback_grad = self._fallback_dudx = None
        # ParticleMeshEwald.java:563-565: direct dipoles after a failed SCF only
        # when the SOR fallback (done inside the library) is off
        self.direct_fallback = (not self.params.sor_fallback) if direct_fallback is None else direct_fallback
        self.vanDerWaals = VanDerWaals(self)
        self.particleMeshEwald = ParticleMeshEwald(self)
        self.valence = valence
        counts = valence.counts() if valence is not None else {k: 0 for k in ffi.VALENCE_KINDS}
        self._terms = {k: EnergyTerm(self, k, counts[k]) for k in ffi.VALENCE_KINDS}
        if valence is not None:
            self._handle.set_valence(valence)

    # -- valence terms (ForceFieldEnergy.java:2955-3045) --------------------
    def getBondPotentialEnergy(self):
        return self._terms["bond"]

    def getAnglePotentialEnergy(self):
        return self._terms["angle"]

    def getStretchBendPotentialEnergy(self):
        return self._terms["strbnd"]

    def getUreyBradleyPotentialEnergy(self):
        return self._terms["urey"]

    def getOutOfPlaneBendPotentialEnergy(self):
        return self._terms["opbend"]

    def getTorsionPotentialEnergy(self):
        return self._terms["torsion"]

    def getPiOrbitalTorsionPotentialEnergy(self):
        return self._terms["pitors"]

    def getTorsionTorsionPotentialEnergy(self):
        return self._terms["tortor"]

    def getVanDerWaalsEnergy(self):
        return self._last["vdw"]

    def getPermanentMultipoleEnergy(self):
        return self._last["perm"]

    def getPolarizationEnergy(self):
        return self._last["pol"]

    def getTotalEnergy(self):
        return self._last["total"]

    # -- Potential interface ------------------------------------------------
    def getNumberOfVariables(self):
        return 3 * self.topology.n

    def getCoordinates(self, x=None):
        if x is None:
            return None if self._x is None else self._x.copy()
        x[:] = self._x
        return x

    def setCoordinates(self, x):
        self._x = np.array(x, dtype=np.float64).reshape(-1)
        self._handle.set_positions(self._x)

    def energy(self, x=None, verbose=False):
        if x is not None:
            self.setCoordinates(x)
        self._evaluate(False)
        return self._last["total"]

    def energyAndGradient(self, x, g=None, verbose=False):
        self.setCoordinates(x)
        self._evaluate(True)
        grad = (self._fallback_grad if self._fallback_grad is not None else self._handle.gradient()).reshape(-1)
        if g is not None:
            g[:] = grad
        return self._last["total"], grad

    def getVanDerWaalsEnergy(self):
        return self._last["vdw"]

    def getPermanentMultipoleEnergy(self):
        return self._last["perm"]

    def getPolarizationEnergy(self):
        return self._last["pol"]

    def getTotalEnergy(self):
        return self._last["total"]

    def destroy(self):
        self._handle.close()

    # -- LambdaInterface (ForceFieldEnergy.java setLambda / getdEdL /
    #    getd2EdL2 / getdEdXdL; the vdW softcore of VanDerWaals.java) --------
    def setLambda(self, lam):
        self._handle.set_lambda(lam)
        self._lambda = float(lam)

    def getLambda(self):
        return 1.0 if self._lambda is None else self._lambda

    def getdEdL(self):
        return self._last["dudl"]

    def getd2EdL2(self):
        return self._last["d2udl2"]

    def getdEdXdL(self, g=None):
        """Adds dU/dX/dL of the last gradient evaluation into g (allocated
        when None), as LambdaInterface.getdEdXdL accumulates."""
        d = (self._fallback_dudx if self._fallback_dudx is not None else self._handle.dudx_dl()).reshape(-1)
        if g is None:
            return d.copy()
        g += d
        return g

    # -- internals -----------------------------------------------------------
    def energyGradientAndVirial(self, x):
        """Energy, gradient and the strain derivative W_ab = dE/d eps_ab
        (kcal/mol).  The reference has no analytic virial (it differentiates
        the lattice numerically, XtalEnergy.java:495-511)."""
        self.setCoordinates(x)
        self._evaluate(True, virial=True)
        if self._fallback_grad is not None:
            raise EnergyException("no virial after the direct-polarization fallback")
        return self._last["total"], self._handle.gradient().reshape(-1), self._handle.virial()

    def _evaluate(self, gradient, virial=False):
        self._fallback_grad = self._fallback_dudx = None
        flags = (ffi.NB_GRAD if gradient else 0) | (ffi.NB_VIRIAL if virial else 0) | \
            (ffi.NB_DUDL if self._lambda is not None else 0) | \
            (ffi.NB_FINITE_PCG if gradient and self.finite_pcg_gradient else 0)
        e, rc = self._handle.evaluate(flags, check=False)
        if rc in (ffi.NB_ERR_SCF_DIVERGED, ffi.NB_ERR_SCF_MAXITER) and self.direct_fallback:
            # direct polarization for this evaluation only; the mutual handle is kept
            p = ffi.params_from(self.params, self.grid, self._nbp.precision, self._nbp.device)
            p.polarization = 1
            tmp = ffi.Handle(p, self.topology, self._nbp.device)
            try:
                tmp.set_cell(self.box)
                tmp.set_positions(self._x)
                if self.valence is not None:
                    tmp.set_valence(self.valence)
                if self._lambda is not None:
                    tmp.set_lambda(self._lambda)
                e, rc = tmp.evaluate(flags & ~ffi.NB_VIRIAL, check=False)
                if rc == ffi.NB_OK:
                    self._fallback_grad = tmp.gradient()
                    if self._lambda is not None and gradient:
                        self._fallback_dudx = tmp.dudx_dl()
            finally:
                tmp.close()
        if rc != ffi.NB_OK:
            raise EnergyException(self._handle.L.nb_last_error(self._handle.h).decode())
        self._last = e
        return e
