"""Seeded synthetic AMOEBA systems for the nonbonded path.

* ``water_box``: rigid AMOEBA-water-2003 molecules (O-H 0.9572 A, H-O-H
  108.5 deg) with oxygens placed by random sequential insertion (no O-O pair
  closer than ``min_oo``) and uniformly random orientations.
* ``chain_in_water``: a self-avoiding random-walk chain with random moments
  solvated by waters.

Every generator takes a seed for numpy's PCG64 and is deterministic.
"""
from __future__ import annotations

import numpy as np

from .forcefield import (FRAME_BISECTOR, FRAME_NONE, FRAME_ZTHENX, WATER_H_MPOLE, WATER_HOH, WATER_O_MPOLE,
                         WATER_OH, WATER_POLARITY, WATER_THOLE, WATER_VDW, Topology, masks_from_bonds,
                         vdw_pair_table)


def _random_rotations(rng, n):
    """Uniform random rotation matrices from uniform unit quaternions."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _sysgen():
    import ctypes as C
    import os
    from .build import BUILD
    path = os.path.join(BUILD, "libsysgen.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    L.rsa_place.argtypes = [dp, C.c_int, dp, C.c_int, C.c_int, dp, C.c_double]
    L.rsa_place.restype = C.c_int
    return L


def sequential_insertion(rng, n, box, min_dist, max_tries=2000):
    """Random sequential insertion in an orthorhombic box.

    Candidates are drawn in batches of 4096 from ``rng``; each is accepted if
    no placed point lies closer than ``min_dist`` (minimum image)."""
    L = _sysgen()
    if L is not None:
        import ctypes as C
        b = np.ascontiguousarray(np.asarray(box, float)[:3])
        pts = np.zeros((n, 3))
        count = 0
        dp = C.POINTER(C.c_double)
        while count < n:
            cand = np.ascontiguousarray(rng.random((4096, 3)) * b)
            count = L.rsa_place(cand.ctypes.data_as(dp), 4096, pts.ctypes.data_as(dp), count, n,
                                b.ctypes.data_as(dp), float(min_dist))
            max_tries -= 1
            if max_tries <= 0:
                raise RuntimeError("sequential insertion did not converge")
        return pts
    return _sequential_insertion_py(rng, n, box, min_dist, max_tries)


def _sequential_insertion_py(rng, n, box, min_dist, max_tries=2000):
    """Pure-Python form of sequential_insertion (same acceptance sequence)."""
    box = np.asarray(box, float)[:3]
    ncell = np.maximum((box / min_dist).astype(int), 1)
    csize = box / ncell
    grid = {}
    pts = np.empty((n, 3))
    md2 = min_dist * min_dist
    count = 0
    batch = 4096
    while count < n:
        cand = rng.random((batch, 3)) * box
        for c in cand:
            ci = (c // csize).astype(int) % ncell
            ok = True
            for dx in (-1, 0, 1):
                if not ok:
                    break
                for dy in (-1, 0, 1):
                    if not ok:
                        break
                    for dz in (-1, 0, 1):
                        key = ((ci[0] + dx) % ncell[0], (ci[1] + dy) % ncell[1], (ci[2] + dz) % ncell[2])
                        lst = grid.get(key)
                        if not lst:
                            continue
                        d = pts[lst] - c
                        d -= box * np.round(d / box)
                        if np.any(np.einsum("ij,ij->i", d, d) < md2):
                            ok = False
                            break
            if ok:
                pts[count] = c
                grid.setdefault((ci[0], ci[1], ci[2]), []).append(count)
                count += 1
                if count == n:
                    break
        max_tries -= 1
        if max_tries <= 0:
            raise RuntimeError("sequential insertion did not converge")
    return pts


def water_geometry():
    """Body-frame water: O at the origin, bisector along +z, molecule in xz."""
    half = np.deg2rad(WATER_HOH) / 2.0
    h1 = WATER_OH * np.array([np.sin(half), 0.0, np.cos(half)])
    h2 = WATER_OH * np.array([-np.sin(half), 0.0, np.cos(half)])
    return np.zeros(3), h1, h2


def water_topology(n_waters, moment_scale=(1.0, 1.0), atom_offset=0, molecule_offset=0):
    """Topology arrays for ``n_waters`` AMOEBA waters in O,H,H order."""
    n = 3 * n_waters
    frame = np.zeros(n, np.int32)
    axis = -np.ones((n, 3), np.int32)
    mp = np.zeros((n, 10))
    pol = np.zeros(n)
    vcls = np.zeros(n, np.int32)
    red_i = np.arange(n, dtype=np.int32) + atom_offset
    red_v = np.zeros(n)
    o_m = WATER_O_MPOLE.copy()
    h_m = WATER_H_MPOLE.copy()
    for m in (o_m, h_m):
        m[1:4] *= moment_scale[0]
        m[4:10] *= moment_scale[1]
    bonds = []
    for w in range(n_waters):
        o, h1, h2 = 3 * w, 3 * w + 1, 3 * w + 2
        g = atom_offset
        frame[o] = FRAME_BISECTOR
        axis[o] = (h1 + g, h2 + g, -1)
        frame[h1] = FRAME_ZTHENX
        axis[h1] = (o + g, h2 + g, -1)
        frame[h2] = FRAME_ZTHENX
        axis[h2] = (o + g, h1 + g, -1)
        mp[o] = o_m
        mp[h1] = h_m
        mp[h2] = h_m
        pol[o] = WATER_POLARITY["O"]
        pol[h1] = pol[h2] = WATER_POLARITY["H"]
        vcls[o] = 0
        vcls[h1] = vcls[h2] = 1
        red_i[h1] = o + g
        red_i[h2] = o + g
        red_v[h1] = red_v[h2] = WATER_VDW["H"][2]
        bonds += [(o, h1), (o, h2)]
    off, atoms, rels = masks_from_bonds(n, bonds)
    atoms = atoms + atom_offset
    tab = vdw_pair_table([WATER_VDW["O"][0], WATER_VDW["H"][0]], [WATER_VDW["O"][1], WATER_VDW["H"][1]])
    mol = np.repeat(np.arange(n_waters, dtype=np.int32), 3) + molecule_offset
    return Topology(n=n, frame=frame, axis=axis, mpole=mp, polarity=pol, thole=np.full(n, WATER_THOLE),
                    pdamp=pol ** (1.0 / 6.0), group=mol.copy(), molecule=mol, vdw_class=vcls, vdw_table=tab,
                    red_index=red_i, red_value=red_v, mask_off=off, mask_atom=atoms, mask_rel=rels)


def water_box(n_waters, box_length, seed=0, min_oo=2.6, moment_scale=(1.0, 1.0)):
    """Seeded AMOEBA water box.  Returns (xyz (3n,), box[6], Topology)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    box = np.array([box_length, box_length, box_length, 90.0, 90.0, 90.0])
    opos = sequential_insertion(rng, n_waters, box, min_oo)
    R = _random_rotations(rng, n_waters)
    o, h1, h2 = water_geometry()
    xyz = np.empty((n_waters, 3, 3))
    xyz[:, 0] = opos
    xyz[:, 1] = opos + R @ h1
    xyz[:, 2] = opos + R @ h2
    top = water_topology(n_waters, moment_scale)
    return xyz.reshape(-1), box, top


WATER_CONFIGS = {
    # name: (n_waters, box length, grid, ewald cutoff, vdw cutoff, polar eps)
    "water-648": (216, 18.643, (24, 24, 24), 7.0, 9.0, 1.0e-6),
    "water-20736": (6912, 59.1, (64, 64, 64), 7.0, 9.0, 1.0e-5),
    "water-288000": (96000, 144.3, (144, 144, 144), 7.0, 9.0, 1.0e-5),
}


CHAIN_VDW = (3.8, 0.1)     # chain atom vdW radius (R-min, A) and well depth (kcal/mol)


def _random_chain(rng, n, box, bond=1.5, min_dist=2.6, tries=200, max_backtracks=10000):
    """Self-avoiding random walk of ``n`` points with fixed bond length: each
    new point keeps ``min_dist`` (minimum image) from every earlier point
    except its bonded predecessor.  A walk that cannot be continued steps back
    a few points and grows again."""
    box = np.asarray(box, float)[:3]
    pts = np.empty((n, 3))
    pts[0] = rng.random(3) * box
    md2 = min_dist * min_dist
    i, backtracks = 1, 0
    while i < n:
        placed = False
        for _ in range(tries):
            v = rng.normal(size=3)
            c = pts[i - 1] + bond * v / np.linalg.norm(v)
            if i >= 2:
                d = pts[:i - 1] - c
                d -= box * np.round(d / box)
                if np.einsum("ij,ij->i", d, d).min() < md2:
                    continue
            pts[i] = c
            placed = True
            break
        if placed:
            i += 1
            continue
        backtracks += 1
        if backtracks > max_backtracks:
            raise RuntimeError("random chain could not be continued")
        i = max(1, i - int(rng.integers(2, 8)))
    return pts


def _random_traceless(rng, n, norm):
    """Random symmetric traceless 3x3 tensors with Frobenius norm ``norm``, as
    (xx, yy, zz, xy, xz, yz)."""
    a = rng.normal(size=(n, 3, 3))
    q = 0.5 * (a + a.transpose(0, 2, 1))
    q -= np.eye(3) * (np.trace(q, axis1=1, axis2=2) / 3.0)[:, None, None]
    q *= norm / np.linalg.norm(q, axis=(1, 2))[:, None, None]
    return np.stack([q[:, 0, 0], q[:, 1, 1], q[:, 2, 2], q[:, 0, 1], q[:, 0, 2], q[:, 1, 2]], axis=1)


def chain_in_water(n_chain, n_waters, box_length, seed=0, n_soft=0, soft_start=None, min_dist=2.6):
    """Random-walk chain with random moments, solvated by AMOEBA waters.

    Chain: bond 1.5 A, self-avoiding at ``min_dist``; charges N(0, 0.3) e made
    neutral; local dipoles of 0.1 e*A and traceless quadrupoles of 0.3 e*A^2
    (Frobenius) in Z-then-X frames on the chain neighbours; polarizabilities
    U(0.3, 1.5) A^3, Thole 0.39; 1-2..1-5 masks from the chain bonds; each
    atom its own polarization group.  The waters' oxygens are inserted at
    ``min_dist`` from every chain atom and from each other.  ``n_soft``
    contiguous chain atoms (from ``soft_start``, default the middle) are
    marked soft.  Returns (xyz (3n,), box[6], Topology); the chain is
    molecule 0 and comes first."""
    rng = np.random.Generator(np.random.PCG64(seed))
    box = np.array([box_length, box_length, box_length, 90.0, 90.0, 90.0])
    chain = _random_chain(rng, n_chain, box, min_dist=min_dist)
    q = rng.normal(0.0, 0.3, n_chain)
    q -= q.mean()
    dv = rng.normal(size=(n_chain, 3))
    dv *= 0.1 / np.linalg.norm(dv, axis=1, keepdims=True)
    quad = _random_traceless(rng, n_chain, 0.3)
    pol = rng.uniform(0.3, 1.5, n_chain)
    # waters around the chain
    L = _sysgen()
    n_tot = n_chain + n_waters
    if L is None:
        raise RuntimeError("libsysgen.so is not built")
    import ctypes as C
    dp = C.POINTER(C.c_double)
    b = np.ascontiguousarray(box[:3])
    pts = np.zeros((n_tot, 3))
    pts[:n_chain] = chain
    count, tries = n_chain, 4000
    while count < n_tot:
        cand = np.ascontiguousarray(rng.random((4096, 3)) * b)
        count = L.rsa_place(cand.ctypes.data_as(dp), 4096, pts.ctypes.data_as(dp), count, n_tot,
                            b.ctypes.data_as(dp), float(min_dist))
        tries -= 1
        if tries <= 0:
            raise RuntimeError("water insertion did not converge")
    opos = pts[n_chain:]
    R = _random_rotations(rng, n_waters)
    o, h1, h2 = water_geometry()
    wxyz = np.empty((n_waters, 3, 3))
    wxyz[:, 0] = opos
    wxyz[:, 1] = opos + R @ h1
    wxyz[:, 2] = opos + R @ h2
    wt = water_topology(n_waters, atom_offset=n_chain, molecule_offset=1)
    # chain topology
    frame = np.full(n_chain, FRAME_ZTHENX, np.int32)
    axis = -np.ones((n_chain, 3), np.int32)
    for i in range(n_chain):
        # z toward the next atom (the previous one at the end), x toward the
        # previous one (the one after z at the start, the one before z at the end)
        if i + 1 < n_chain:
            z, x = i + 1, (i - 1 if i > 0 else i + 2)
        else:
            z, x = i - 1, i - 2
        axis[i, 0], axis[i, 1] = z, x
    mp = np.concatenate([q[:, None], dv, quad], axis=1)
    bonds = [(i, i + 1) for i in range(n_chain - 1)]
    off, atoms, rels = masks_from_bonds(n_chain, bonds)
    # vdW classes: 0 water O, 1 water H, 2 chain
    tab = vdw_pair_table([WATER_VDW["O"][0], WATER_VDW["H"][0], CHAIN_VDW[0]],
                         [WATER_VDW["O"][1], WATER_VDW["H"][1], CHAIN_VDW[1]])
    n = n_tot = n_chain + 3 * n_waters
    cat = np.concatenate
    mask_off = cat([off, off[-1] + wt.mask_off[1:]]).astype(np.int32)
    top = Topology(n=n, frame=cat([frame, wt.frame]).astype(np.int32), axis=cat([axis, wt.axis]).astype(np.int32),
                   mpole=cat([mp, wt.mpole]), polarity=cat([pol, wt.polarity]),
                   thole=cat([np.full(n_chain, WATER_THOLE), wt.thole]),
                   pdamp=cat([pol ** (1.0 / 6.0), wt.pdamp]),
                   group=cat([np.arange(n_chain, dtype=np.int32) + 10 * n, wt.group]).astype(np.int32),
                   molecule=cat([np.zeros(n_chain, np.int32), wt.molecule]).astype(np.int32),
                   vdw_class=cat([np.full(n_chain, 2, np.int32), wt.vdw_class]).astype(np.int32), vdw_table=tab,
                   red_index=cat([np.arange(n_chain, dtype=np.int32), wt.red_index]).astype(np.int32),
                   red_value=cat([np.zeros(n_chain), wt.red_value]),
                   mask_off=mask_off, mask_atom=cat([atoms, wt.mask_atom]).astype(np.int32),
                   mask_rel=cat([rels, wt.mask_rel]).astype(np.int32))
    if n_soft:
        s0 = (n_chain - n_soft) // 2 if soft_start is None else soft_start
        top.soft[s0:s0 + n_soft] = 1
    xyz = np.concatenate([chain, wxyz.reshape(-1, 3)]).reshape(-1)
    return xyz, box, top


def lattice_vectors(box):
    """Rows a, b, c of a cell given as (a, b, c, alpha, beta, gamma) in degrees,
    with a along x and b in the xy plane (the convention of Crystal.java)."""
    a, b, c, al, be, ga = [float(v) for v in box]
    if al == 90.0 and be == 90.0 and ga == 90.0:
        return np.diag([a, b, c])
    ca, cb, cg = np.cos(np.deg2rad([al, be, ga]))
    sg, sb = np.sin(np.deg2rad(ga)), np.sin(np.deg2rad(be))
    bt = (ca - cb * cg) / sg
    gt = np.sqrt(sb * sb - bt * bt)
    return np.array([[a, 0.0, 0.0], [b * cg, b * sg, 0.0], [c * cb, c * bt, c * gt]])


def replicate(xyz, box, top, reps):
    """Supercell of reps = (l, m, n) copies of the cell (ReplicatesCrystal):
    replica r = (i, j, k) holds atoms r * n_atoms + 0..n_atoms-1 translated by
    i a + j b + k c.  Returns (xyz (3 N,), box[6], Topology)."""
    l, m, nn = reps
    L = lattice_vectors(box)
    x = np.asarray(xyz, float).reshape(-1, 3)
    n = top.n
    shifts = [i * L[0] + j * L[1] + k * L[2] for k in range(nn) for j in range(m) for i in range(l)]
    R = len(shifts)
    X = np.concatenate([x + s for s in shifts])
    off = lambda a, r: np.where(a >= 0, a + r * n, a)
    moff = np.concatenate([top.mask_off[:-1] + r * top.mask_off[-1] for r in range(R)] + [[R * top.mask_off[-1]]])
    gmax = int(max(top.group.max(), top.molecule.max())) + 1
    t = Topology(n=R * n, frame=np.tile(top.frame, R), axis=np.concatenate([off(top.axis, r) for r in range(R)]),
                 mpole=np.tile(top.mpole, (R, 1)), polarity=np.tile(top.polarity, R), thole=np.tile(top.thole, R),
                 pdamp=np.tile(top.pdamp, R), group=np.concatenate([top.group + r * gmax for r in range(R)]),
                 molecule=np.concatenate([top.molecule + r * gmax for r in range(R)]),
                 vdw_class=np.tile(top.vdw_class, R), vdw_table=top.vdw_table,
                 red_index=np.concatenate([top.red_index + r * n for r in range(R)]),
                 red_value=np.tile(top.red_value, R), mask_off=moff.astype(np.int32),
                 mask_atom=np.concatenate([top.mask_atom + r * n for r in range(R)]).astype(np.int32),
                 mask_rel=np.tile(top.mask_rel, R).astype(np.int32),
                 soft=np.tile(top.soft, R))
    sbox = np.array([box[0] * l, box[1] * m, box[2] * nn, box[3], box[4], box[5]], float)
    return X.reshape(-1), sbox, t


def small_crystal(n_target, box, seed=0, min_dist=2.6, sizes=(3, 4, 5), tries=4000):
    """P1 cell of small rigid molecules with random multipoles (the batched
    small-crystal workload): each molecule is a 3-5 atom chain (bond 1.5 A,
    bond angles >= 109.5 deg) with charges N(0, 0.3) e made neutral per
    molecule, local dipoles of 0.1 e*A and traceless quadrupoles of 0.3 e*A^2
    in Z-then-X frames, polarizabilities U(0.3, 1.5) A^3 and Thole 0.39; one
    polarization group per molecule.  Molecules are placed at random with every
    atom at least ``min_dist`` from every image of the atoms already placed,
    until ``n_target`` atoms or ``tries`` failed attempts.  The cell may be
    thinner than twice the cutoff (image sums).  Returns (xyz, box, Topology)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    box = np.asarray(box, float)
    L = lattice_vectors(box)
    # neighbouring translations suffice for a spacing below every cell width
    img = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], float) @ L
    placed = np.zeros((0, 3))
    mols = []
    fails = 0
    while placed.shape[0] < n_target and fails < tries:
        na = int(rng.choice(sizes))
        if placed.shape[0] + na > n_target + 2:
            na = max(3, n_target - placed.shape[0])
        # rigid chain geometry
        geo = [np.zeros(3)]
        d = rng.normal(size=3)
        geo.append(1.5 * d / np.linalg.norm(d))
        while len(geo) < na:
            v = rng.normal(size=3)
            c = geo[-1] + 1.5 * v / np.linalg.norm(v)
            back = geo[-2] - geo[-1]
            if np.dot(back, c - geo[-1]) / (1.5 * 1.5) > np.cos(np.deg2rad(109.5)):
                continue
            geo.append(c)
        geo = np.array(geo) - np.mean(geo, axis=0)
        R = _random_rotations(rng, 1)[0]
        pos = geo @ R.T + rng.random(3) @ L
        if placed.shape[0]:
            ok = True
            for p in pos:
                dd = placed[:, None, :] - p[None, None, :] + img[None, :, :]
                if np.einsum("ijk,ijk->ij", dd, dd).min() < min_dist ** 2:
                    ok = False
                    break
            if not ok:
                fails += 1
                continue
        mols.append(pos)
        placed = np.concatenate([placed, pos])
    n = placed.shape[0]
    frame = np.full(n, FRAME_ZTHENX, np.int32)
    axis = -np.ones((n, 3), np.int32)
    mp = np.zeros((n, 10))
    mol = np.zeros(n, np.int32)
    bonds = []
    a0 = 0
    for m, pos in enumerate(mols):
        na = len(pos)
        q = rng.normal(0.0, 0.3, na)
        q -= q.mean()
        dv = rng.normal(size=(na, 3))
        dv *= 0.1 / np.linalg.norm(dv, axis=1, keepdims=True)
        mp[a0:a0 + na] = np.concatenate([q[:, None], dv, _random_traceless(rng, na, 0.3)], axis=1)
        for j in range(na):
            i = a0 + j
            if j + 1 < na:
                axis[i, 0], axis[i, 1] = i + 1, (i - 1 if j > 0 else i + 2)
            else:
                axis[i, 0], axis[i, 1] = i - 1, i - 2
            if j + 1 < na:
                bonds.append((i, i + 1))
        mol[a0:a0 + na] = m
        a0 += na
    pol = rng.uniform(0.3, 1.5, n)
    off, atoms, rels = masks_from_bonds(n, bonds)
    tab = vdw_pair_table([CHAIN_VDW[0]], [CHAIN_VDW[1]])
    top = Topology(n=n, frame=frame, axis=axis, mpole=mp, polarity=pol, thole=np.full(n, WATER_THOLE),
                   pdamp=pol ** (1.0 / 6.0), group=mol.copy(), molecule=mol, vdw_class=np.zeros(n, np.int32),
                   vdw_table=tab, red_index=np.arange(n, dtype=np.int32), red_value=np.zeros(n), mask_off=off,
                   mask_atom=atoms, mask_rel=rels)
    return placed.reshape(-1), box, top


def _smooth_at_least(x, lo=16, hi=32):
    """Smallest 2,3,5-smooth integer >= x, clamped to [lo, hi]."""
    n = max(lo, int(np.ceil(x)))
    while True:
        m = n
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        if m == 1 or n >= hi:
            return min(n, hi) if m == 1 or n <= hi else hi
        n += 1


MOLECULE_BOND = 1.1        # centre-substituent bond of the packed small molecules (A)


def _star_geometry(m):
    """Body-frame geometry of a compact m-atom molecule (3 <= m <= 5): a centre
    at the origin and m - 1 substituents at MOLECULE_BOND along tetrahedral
    directions (m = 3: bent at 109.47 deg, m = 4: pyramidal, m = 5:
    tetrahedral)."""
    t = np.array([[1.0, 1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]]) / np.sqrt(3.0)
    return np.concatenate([np.zeros((1, 3)), MOLECULE_BOND * t[:m - 1]])


def _molecule_sizes(rng, n):
    """Sizes in {3, 4, 5} that add up to n (n >= 6 or n in {3, 4, 5})."""
    sizes = []
    rem = n
    while rem > 5:
        choices = [k for k in (3, 4, 5) if rem - k not in (1, 2)]
        k = int(rng.choice(choices))
        sizes.append(k)
        rem -= k
    sizes.append(rem)
    return sizes


def molecular_crystal(n_atoms, box, seed=0, min_dist=1.9, orient_tries=60):
    """P1 cell of exactly ``n_atoms`` atoms in compact rigid molecules of 3-5
    atoms (a centre with substituents at 1.1 A) packed on a lattice: molecule
    centres sit on a g0 x g1 x g2 grid of fractional sites (spacing as even as
    the cell allows, about 3.4 A at 0.1 atoms/A^3) with a small jitter, each
    in a random orientation drawn until every atom is at least ``min_dist``
    from every image of the atoms of other molecules already placed; a site
    that admits no orientation is left empty and the molecule takes a spare
    site.  Charges N(0, 0.3) e made neutral per molecule, local dipoles of
    0.1 e*A and traceless quadrupoles of 0.3 e*A^2 in Z-then-X frames,
    polarizabilities U(0.3, 1.5) A^3, Thole 0.39, one polarization group per
    molecule; centres take AMOEBA water oxygen vdW parameters and
    substituents water hydrogen parameters with their 0.91 reduction toward
    the centre.  Raises RuntimeError when the atoms do not fit.  Returns
    (xyz, box, Topology)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    box = np.asarray(box, float)
    L = lattice_vectors(box)
    vol = abs(np.linalg.det(L))
    sizes = _molecule_sizes(rng, n_atoms)
    M = len(sizes)
    # site grid: about M * 1.15 sites, spaced in proportion to the cell lengths
    lens = box[:3]
    s = (vol / (1.15 * M)) ** (1.0 / 3.0)
    g = np.maximum(1, np.round(lens / s)).astype(int)
    while np.prod(g) < M + max(2, M // 10):
        g[np.argmax(lens / g)] += 1
    sites = np.array([(i, j, k) for i in range(g[0]) for j in range(g[1]) for k in range(g[2])], float)
    sites = (sites + 0.5) / g
    rng.shuffle(sites)
    img = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], float) @ L
    placed = np.zeros((0, 3))
    mols = []
    site = 0
    for m in sizes:
        geo = _star_geometry(m)
        ok = False
        while not ok and site < len(sites):
            centre = (sites[site] + rng.uniform(-0.08, 0.08, 3) / g) @ L
            site += 1
            # images of placed atoms that a molecule at this centre can reach
            near = np.zeros((0, 3))
            if placed.shape[0]:
                cand = (placed[:, None, :] + img[None, :, :]).reshape(-1, 3)
                reach = MOLECULE_BOND + min_dist
                near = cand[np.einsum("ij,ij->i", cand - centre, cand - centre) < reach * reach]
            for _ in range(orient_tries):
                pos = geo @ _random_rotations(rng, 1)[0].T + centre
                if near.shape[0]:
                    dd = near[:, None, :] - pos[None, :, :]
                    if np.einsum("ijk,ijk->ij", dd, dd).min() < min_dist ** 2:
                        continue
                ok = True
                break
        if not ok:
            raise RuntimeError(f"molecular_crystal: {n_atoms} atoms do not fit the cell {box.tolist()}")
        mols.append(pos)
        placed = np.concatenate([placed, pos])
    n = placed.shape[0]
    frame = np.full(n, FRAME_ZTHENX, np.int32)
    axis = -np.ones((n, 3), np.int32)
    mp = np.zeros((n, 10))
    mol = np.zeros(n, np.int32)
    vcls = np.zeros(n, np.int32)
    red_i = np.arange(n, dtype=np.int32)
    red_v = np.zeros(n)
    bonds = []
    a0 = 0
    for k, pos in enumerate(mols):
        na = len(pos)
        q = rng.normal(0.0, 0.3, na)
        q -= q.mean()
        dv = rng.normal(size=(na, 3))
        dv *= 0.1 / np.linalg.norm(dv, axis=1, keepdims=True)
        mp[a0:a0 + na] = np.concatenate([q[:, None], dv, _random_traceless(rng, na, 0.3)], axis=1)
        c = a0
        axis[c, 0], axis[c, 1] = c + 1, c + 2
        for j in range(1, na):
            i = a0 + j
            axis[i, 0], axis[i, 1] = c, (a0 + 1 if j != 1 else a0 + 2)
            bonds.append((c, i))
            vcls[i] = 1
            red_i[i] = c
            red_v[i] = WATER_VDW["H"][2]
        mol[a0:a0 + na] = k
        a0 += na
    pol = rng.uniform(0.3, 1.5, n)
    off, atoms, rels = masks_from_bonds(n, bonds)
    tab = vdw_pair_table([WATER_VDW["O"][0], WATER_VDW["H"][0]], [WATER_VDW["O"][1], WATER_VDW["H"][1]])
    top = Topology(n=n, frame=frame, axis=axis, mpole=mp, polarity=pol, thole=np.full(n, WATER_THOLE),
                   pdamp=pol ** (1.0 / 6.0), group=mol.copy(), molecule=mol, vdw_class=vcls, vdw_table=tab,
                   red_index=red_i, red_value=red_v, mask_off=off, mask_atom=atoms, mask_rel=rels)
    return placed.reshape(-1), box, top


CRYSTAL_DENSITY_MAX = 0.1  # atoms per A^3 a batch cell may hold


def crystal_batch(count, seed=0):
    """The batched small-crystal workload: ``count`` independent P1 triclinic
    cells.  Each draws its atom count U{30..300}, then cells with lengths
    U(8, 20) A and angles U(70, 110) deg until one has a positive volume that
    holds the count at no more than 0.1 atoms/A^3 (a cell that cannot is
    redrawn), and fills it by molecular_crystal.  Per-cell PME grids are the
    next 2,3,5-smooth size >= 1.2 L, clamped to 16..32.  512 cells of seed 0
    hold 83,375 atoms.  Returns a list of (xyz, box, Topology, grid)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for c in range(count):
        n_atoms = int(rng.integers(30, 301))
        while True:
            while True:
                box = np.concatenate([rng.uniform(8.0, 20.0, 3), rng.uniform(70.0, 110.0, 3)])
                ca, cb, cg = np.cos(np.deg2rad(box[3:]))
                v2 = 1 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg
                if v2 > 0.05 and np.prod(box[:3]) * np.sqrt(v2) * CRYSTAL_DENSITY_MAX >= n_atoms:
                    break
            try:
                xyz, box, top = molecular_crystal(n_atoms, box, seed=int(rng.integers(1 << 30)))
                break
            except RuntimeError:
                continue
        grid = tuple(_smooth_at_least(1.2 * L) for L in box[:3])
        out.append((xyz, box, top, grid))
    return out


# --------------------------------------------------------------------------
# the named workloads
# --------------------------------------------------------------------------
def elements_of(top):
    """(symbol, atomic number, mass) per atom of the synthetic systems by vdW
    class: water oxygens and molecule centres (class 0) O, water hydrogens
    and substituents (class 1) H, chain atoms (class 2) C."""
    table = {0: ("O", 8, 15.999), 1: ("H", 1, 1.008), 2: ("C", 6, 12.011)}
    return [table[int(c)] for c in top.vdw_class]


NAMED_CONFIGS = ("water-648", "water-20736", "chain-23557", "crystals-512", "water-288000")


def named_config(name, seed=None):
    """The systems of a named workload as a list of (label, xyz (3n,), box,
    Topology, NonbondedParams with the grid and state set):

    * water-648: 216 waters, 18.643 A, PME 24^3, vdW 9 A, PCG 1e-6 D (seed 1);
    * water-20736: 6912 waters, 59.1 A, 64^3, vdW 9 A, 1e-5 D (seed 2);
    * chain-23557: a 2500-atom chain in 7019 waters, 62.2 A, 64^3, 20 soft
      atoms, lambda 0.5 with the soft atoms' electrostatics coupled (seed 7);
    * crystals-512: 512 triclinic cells of 30-300 atoms, Ewald and vdW 6 A,
      per-cell grids, 1e-5 D (seed 2), one entry per cell;
    * water-288000: 96000 waters, 144.3 A, 144^3, vdW 9 A, 1e-5 D (seed 2)."""
    from .forcefield import NonbondedParams
    if name == "water-648":
        xyz, box, top = water_box(216, 18.643, seed=1 if seed is None else seed)
        return [(name, xyz, box, top, NonbondedParams(vdw_cutoff=9.0, grid=(24, 24, 24), polar_eps=1e-6))]
    if name == "water-20736":
        xyz, box, top = water_box(6912, 59.1, seed=2 if seed is None else seed)
        return [(name, xyz, box, top, NonbondedParams(vdw_cutoff=9.0, grid=(64, 64, 64), polar_eps=1e-5))]
    if name == "chain-23557":
        xyz, box, top = chain_in_water(2500, 7019, 62.2, seed=7 if seed is None else seed, n_soft=20)
        return [(name, xyz, box, top, NonbondedParams(vdw_cutoff=9.0, grid=(64, 64, 64), polar_eps=1e-5,
                                                      elec_lambda=True, lambda_value=0.5))]
    if name == "crystals-512":
        cells = crystal_batch(512, seed=2 if seed is None else seed)
        return [(f"cell{k:03d}", xyz, box, top, NonbondedParams(ewald_cutoff=6.0, vdw_cutoff=6.0, polar_eps=1e-5,
                                                                grid=tuple(grid)))
                for k, (xyz, box, top, grid) in enumerate(cells)]
    if name == "water-288000":
        xyz, box, top = water_box(96000, 144.3, seed=2 if seed is None else seed)
        return [(name, xyz, box, top, NonbondedParams(vdw_cutoff=9.0, grid=(144, 144, 144), polar_eps=1e-5))]
    raise ValueError(f"unknown workload {name}; one of {', '.join(NAMED_CONFIGS)}")


def write_named_config(name, folder, seed=None, digits=8):
    """Write a named workload as Tinker files under ``folder``: one
    xyz/key/prm set per system (``folder/<label>.*``).  Returns the list of
    forcefieldx_amd.tinker.write_system results; their alchemical_atoms and
    lambda_value are the options that tinker.load takes for the soft atoms
    and the state."""
    import os

    from .tinker import write_system
    os.makedirs(folder, exist_ok=True)
    return [write_system(os.path.join(folder, label), xyz, box, top, p, title=f"{name} {label}".strip(),
                         elements=elements_of(top), digits=digits)
            for label, xyz, box, top, p in named_config(name, seed)]


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="Write a named workload as Tinker xyz/key/prm files.")
    ap.add_argument("name", choices=NAMED_CONFIGS)
    ap.add_argument("folder")
    ap.add_argument("--seed", type=int, default=None)
    a = ap.parse_args()
    for r in write_named_config(a.name, a.folder, a.seed):
        opts = "".join(f" --ac {r['alchemical_atoms']}" for _ in [0] if r["alchemical_atoms"])
        print(f"{r['xyz']} {r['key']} {r['prm']} (lambda {r['lambda_value']}{opts})")
