"""AMOEBA force-field parameters and combining rules for the nonbonded path.

Mirrors the parts of the reference's parameter layer that the nonbonded
energy reads:

* ``MultipoleType`` local-frame moments, stored in e, e*A, e*A^2 after the
  Bohr conversion (MultipoleType.java:1116-1119) and frame definitions
  (MultipoleType.java:813-838),
* ``PolarizeType`` polarizability, Thole and pdamp = alpha^(1/6)
  (PolarizeType.java:145-147),
* ``VanDerWaalsForm`` radius/epsilon combining rules (VanDerWaalsForm.java:
  476-534), R-min diameter convention and the reduction factor of hydrogens,
* the AMOEBA masking scale factors (ScaleParameters.java:306-390).

Parameter values for water are those of AMOEBA-WATER-2003
(modules/potential/src/main/java/ffx/potential/parameters/ff/AMOEBA_WATER_2003).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

BOHR = 0.529177210903  # Constants.java:209
ELECTRIC = 332.0637133  # Constants.java:192

FRAME_NONE, FRAME_ZONLY, FRAME_ZTHENX, FRAME_BISECTOR, FRAME_ZTHENBISECTOR, FRAME_THREEFOLD = range(6)


def bohr_to_angstrom(q, d, quad):
    """Local moments read from a parameter file in Bohr units -> e*A, e*A^2."""
    d = np.asarray(d, float) * BOHR
    quad = np.asarray(quad, float) * BOHR * BOHR
    return np.concatenate([[q], d, quad])


@dataclass
class NonbondedParams:
    """Scalar settings of the nonbonded path (reference defaults in comments)."""
    ewald_cutoff: float = 7.0          # EwaldParameters.java:55
    ewald_alpha: float = 0.545         # EwaldParameters.java:71
    vdw_cutoff: float = 12.0           # VanDerWaalsForm.java:85
    vdw_taper: float = 0.9             # VanDerWaalsForm.java:80
    list_buffer: float = 2.0           # ForceFieldEnergy.java:244
    pme_order: int = 5                 # ReciprocalSpace.java:141
    grid: tuple = (0, 0, 0)            # 0 -> density rule (ReciprocalSpace.java:936-969)
    pme_mesh_density: float = 1.2
    polar_eps: float = 1.0e-6          # ParticleMeshEwald.java:171 (Debye)
    electric: float = ELECTRIC
    precond_cutoff: float = 4.5        # PCGSolver.java:300
    precond_scale: float = 2.0         # PCGSolver.java:316
    m_scale: tuple = (0.0, 0.0, 0.4, 0.8)
    p_scale: tuple = (0.0, 0.0, 1.0, 1.0)
    intra14: float = 0.5
    d11: float = 0.0
    v_scale: tuple = (0.0, 0.0, 1.0)
    halgren_delta: float = 0.07        # VDWType.java:91-95
    halgren_gamma: float = 0.12
    polarization: int = 2              # 0 none, 1 direct, 2 mutual
    do_vdw: bool = True
    do_pme: bool = True
    max_scf_iter: int = 1000
    precision: int = 0                 # 0 FP64, 1 mixed
    lambda_value: float = 1.0
    vdw_lambda_alpha: float = 0.25     # VanDerWaals.java:599-624
    vdw_lambda_exponent: float = 3.0
    vdw_lambda_end: float = 1.0        # VanDerWaals.java:605-609
    intermolecular_softcore: bool = False
    intramolecular_softcore: bool = False
    perm_lambda_alpha: float = 2.0     # AlchemicalParameters.java:171-249
    perm_lambda_exponent: float = 3.0
    perm_lambda_start: float = 0.4
    perm_lambda_end: float = 1.0
    pol_lambda_exponent: float = 3.0
    pol_lambda_start: float = 0.75
    pol_lambda_end: float = 1.0
    elec_lambda: bool = False          # lambda also couples the soft atoms' electrostatics (OST)
    ligand_vapor_elec: bool = True     # AlchemicalParameters.java:249
    no_ligand_condensed_scf: bool = True  # AlchemicalParameters.java:251
    scf_algorithm: int = 0             # 0 PCG, 1 SOR (ParticleMeshEwald.java:2288-2299)
    polar_sor: float = 0.7             # SORRegion.java:65-72
    sor_fallback: bool = True          # ParticleMeshEwald.java:563

    def resolved_grid(self, box):
        """Grid dimensions: explicit, or the reference density rule."""
        if all(g > 0 for g in self.grid):
            return tuple(int(g) for g in self.grid)
        out = []
        for length in box[:3]:
            n = int(np.floor(length * self.pme_mesh_density)) + 1
            if n % 2:
                n += 1
            while not _preferred_dimension(n):
                n += 2
            out.append(max(n, 16))
        return tuple(out)


def _preferred_dimension(n):
    """Complex.preferredDimension: factors over {7,6,5,4,3,2} only."""
    if n < 2:
        return False
    for f in (7, 6, 5, 4, 3, 2):
        while n % f == 0:
            n //= f
    return n == 1


def combined_radius(ri, rj, rule="CUBIC-MEAN"):
    if rule == "ARITHMETIC":
        return ri + rj
    if rule == "GEOMETRIC":
        return 2.0 * np.sqrt(ri) * np.sqrt(rj)
    return 2.0 * (ri ** 3 + rj ** 3) / (ri ** 2 + rj ** 2)


def combined_eps(ei, ej, ri, rj, rule="HHG"):
    sei, sej = np.sqrt(ei), np.sqrt(ej)
    if rule == "GEOMETRIC":
        return sei * sej
    if rule == "W-H":
        rr = ri * rj
        return 2.0 * sei * sej * rr ** 3 / (ri ** 6 + rj ** 6)
    if ei * ej == 0.0:
        return 0.0
    return 4.0 * (ei * ej) / ((sei + sej) ** 2)


def vdw_pair_table(radius, eps, radius_rule="CUBIC-MEAN", eps_rule="HHG", radius_size="DIAMETER",
                   radius_type="R-MIN"):
    """(1/Rmin, eps) table over classes, as VanDerWaalsForm builds radEps."""
    scale = 0.5 if radius_size == "DIAMETER" else 1.0
    if radius_type == "SIGMA":
        scale *= 1.122462048309372981
    r = np.asarray(radius, float) * scale
    e = np.asarray(eps, float)
    nc = len(r)
    tab = np.zeros((nc, nc, 2))
    for i in range(nc):
        for j in range(nc):
            rm = combined_radius(r[i], r[j], radius_rule)
            tab[i, j, 0] = 1.0 / rm if rm > 0 else 0.0
            tab[i, j, 1] = combined_eps(e[i], e[j], r[i], r[j], eps_rule)
    return tab


# AMOEBA-WATER-2003 (AMOEBA_WATER_2003:22-72)
WATER_O_MPOLE = bohr_to_angstrom(-0.51966, [0.0, 0.0, 0.14279],
                                 [0.37928, -0.41809, 0.03881, 0.0, 0.0, 0.0])
WATER_H_MPOLE = bohr_to_angstrom(0.25983, [-0.03859, 0.0, -0.05818],
                                 [-0.03673, -0.10739, 0.14412, 0.0, -0.00203, 0.0])
WATER_POLARITY = {"O": 0.837, "H": 0.496}
WATER_THOLE = 0.39
WATER_VDW = {"O": (3.405, 0.110, 0.0), "H": (2.655, 0.0135, 0.91)}
WATER_OH = 0.9572
WATER_HOH = 108.50


@dataclass
class Topology:
    """Per-atom arrays of the nonbonded path (structure-of-arrays)."""
    n: int
    frame: np.ndarray
    axis: np.ndarray          # (n,3) int, -1 unused
    mpole: np.ndarray         # (n,10) local frame
    polarity: np.ndarray
    thole: np.ndarray
    pdamp: np.ndarray
    group: np.ndarray
    molecule: np.ndarray
    vdw_class: np.ndarray
    vdw_table: np.ndarray     # (nc,nc,2)
    red_index: np.ndarray
    red_value: np.ndarray
    mask_off: np.ndarray      # CSR of bonded relations
    mask_atom: np.ndarray
    mask_rel: np.ndarray
    soft: np.ndarray = field(default=None)

    def __post_init__(self):
        if self.soft is None:
            self.soft = np.zeros(self.n, np.int32)


def masks_from_bonds(n, bonds, max_rel=5):
    """1-2 .. 1-5 relations from a bond list by breadth-first search."""
    adj = [[] for _ in range(n)]
    for a, b in bonds:
        adj[a].append(b)
        adj[b].append(a)
    off = [0]
    atoms, rels = [], []
    for i in range(n):
        seen = {i: 1}
        frontier = [i]
        for depth in range(2, max_rel + 1):
            nxt = []
            for a in frontier:
                for b in adj[a]:
                    if b not in seen:
                        seen[b] = depth
                        nxt.append(b)
            frontier = nxt
        items = sorted((k, r) for k, r in seen.items() if k != i)
        atoms.extend(k for k, _ in items)
        rels.extend(r for _, r in items)
        off.append(len(atoms))
    return (np.asarray(off, np.int32), np.asarray(atoms, np.int32), np.asarray(rels, np.int32))
