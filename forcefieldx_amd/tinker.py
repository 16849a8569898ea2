"""Tinker-format input for the nonbonded path, read as Force Field X reads it.

* Coordinate files (.xyz): atom count and title, an optional line of six cell
  parameters, then one line per atom: index, name, x, y, z, atom type, bonded
  atoms (parsers/XYZFilter.java).
* Keyword files (.key / .properties): `keyword value ...` lines, `#` comments
  (parsers/KeyFilter.java; keyword semantics ForceField.java:238-763).
* AMOEBA parameter files: the `atom`, `vdw`, `vdwpr`, `multipole` (five-line
  records), `polarize` records and the scalar keywords of the vdW form, the
  mask scales and the polarization model (parsers/ForceFieldFilter.java;
  MultipoleType.parse, MultipoleType.java:787-898); the valence records
  `bond`, `angle`, `anglep`, `strbnd`, `ureybrad`, `opbend`, `torsion`,
  `pitors` and `tortors` (with its grid lines) and their unit and polynomial
  keywords, kept as records for forcefieldx_amd.bonded.

`build` assembles a Topology and NonbondedParams from them:

* multipole types by the reference's matching order (MultipoleType.
  multipoleTypeFactory, MultipoleType.java:547-725): the atom's type alone,
  then 1, 2 and 3 reference atoms among its 1-2 neighbours, then a 1-2 and
  one or two 1-3 atoms, neighbours sorted by index; the matched atoms are the
  frame's z, x, y atoms; Bohr -> Angstrom units (MultipoleType.java:265);
* polarization groups grown over bonds between atoms whose types the
  `polarize` records list (PolarizeType.assignPolarizationGroups,
  PolarizeType.java:174-215, 360-390), pdamp = alpha^(1/6);
* reduced vdW sites on atoms with one bond and a reduction factor
  (VanDerWaals.java:944-956), pair table by the file's combining rules with
  `vdwpr` overrides;
* the cell from the keywords (a-axis .. gamma), PME grid by the density rule,
  cutoffs, Ewald coefficient, convergence and SCF algorithm, the vdW taper,
  preconditioner, SOR and alchemical keywords; molecules are the bonded
  components.

`write_system` goes the other way for systems held as arrays (the synthetic
workloads): coordinates, a key file and a parameter patch whose records give
back the same Topology under the matching rules above.
"""
from __future__ import annotations

import gzip
import io
from dataclasses import dataclass, field, replace

import numpy as np

from .forcefield import (BOHR, FRAME_BISECTOR, FRAME_NONE, FRAME_THREEFOLD, FRAME_ZONLY, FRAME_ZTHENBISECTOR,
                         FRAME_ZTHENX, NonbondedParams, Topology, combined_eps, combined_radius, masks_from_bonds)


def _open(src):
    if hasattr(src, "read"):
        return src
    if str(src).endswith(".gz"):
        return io.TextIOWrapper(gzip.open(src, "rb"), encoding="utf-8")
    return open(src, "r", encoding="utf-8", errors="replace")


# --------------------------------------------------------------------------
# coordinates
# --------------------------------------------------------------------------
@dataclass
class XYZFile:
    title: str
    names: list
    xyz: np.ndarray        # (n, 3)
    types: np.ndarray      # (n,) atom types
    bonds: list            # (i, j) pairs, 0-based, i < j
    box: list | None = None


def read_xyz(src) -> XYZFile:
    with _open(src) as f:
        lines = [ln.rstrip("\n") for ln in f]
    head = lines[0].split()
    n = int(head[0])
    title = lines[0].split(None, 1)[1] if len(head) > 1 else ""
    k = 1
    box = None
    toks = lines[1].split() if len(lines) > 1 else []
    if len(toks) == 6:
        try:
            vals = [float(t) for t in toks]
            if not toks[0].lstrip("-").isdigit():
                box = vals
                k = 2
        except ValueError:
            pass
    names, xyz, types, bonds = [], np.zeros((n, 3)), np.zeros(n, np.int64), set()
    for a in range(n):
        t = lines[k + a].split()
        idx = int(t[0]) - 1
        if idx != a:
            raise ValueError(f"atom {a + 1}: index {t[0]} out of order")
        names.append(t[1])
        xyz[a] = [float(t[2]), float(t[3]), float(t[4])]
        types[a] = int(t[5])
        for b in t[6:]:
            j = int(b) - 1
            if j != a:
                bonds.add((min(a, j), max(a, j)))
    return XYZFile(title, names, xyz, types, sorted(bonds), box)


def _create(path):
    """Text output; a .gz path is compressed with a fixed header time, so the
    same content gives the same bytes."""
    if str(path).endswith(".gz"):
        return io.TextIOWrapper(gzip.GzipFile(path, "wb", mtime=0), encoding="utf-8")
    return open(path, "w")


def write_xyz(path, xf: XYZFile, digits=6):
    adj = [[] for _ in range(len(xf.names))]
    for i, j in xf.bonds:
        adj[i].append(j)
        adj[j].append(i)
    with _create(path) as f:
        f.write(f"{len(xf.names):6d}  {xf.title}\n")
        if xf.box is not None:
            f.write(" ".join(f"{v:.6f}" for v in xf.box) + "\n")
        for a, nm in enumerate(xf.names):
            x, y, z = xf.xyz[a]
            nb = "".join(f"{j + 1:6d}" for j in sorted(adj[a]))
            f.write(f"{a + 1:6d}  {nm:<3s} {x:{digits + 5}.{digits}f} {y:{digits + 5}.{digits}f} {z:{digits + 5}.{digits}f}"
                    f" {int(xf.types[a]):5d}{nb}\n")


# --------------------------------------------------------------------------
# keywords and parameters
# --------------------------------------------------------------------------
def read_keys(src) -> dict:
    """keyword (lower case) -> list of value strings (the last occurrence wins)."""
    keys = {}
    with _open(src) as f:
        for ln in f:
            ln = ln.split("#")[0].strip()
            if not ln:
                continue
            t = ln.split()
            keys[t[0].lower()] = t[1:]
    return keys


@dataclass
class ForceFieldParams:
    keys: dict = field(default_factory=dict)          # scalar keywords
    atom: dict = field(default_factory=dict)          # type -> (class, name)
    element: dict = field(default_factory=dict)       # type -> (description, atomic number, mass, valence)
    valence: dict = field(default_factory=dict)       # record keyword -> [(classes, values)]
    tortors: list = field(default_factory=list)       # [(5 classes, nx, ny, [(t1, t2, e)] in file order)]
    vdw: dict = field(default_factory=dict)           # class -> (radius, eps, reduction)
    vdwpr: dict = field(default_factory=dict)         # (c1, c2) -> (radius, eps)
    multipole: dict = field(default_factory=dict)     # (types...) -> (frame, [q, d(3), Q(6)] in Bohr units)
    polarize: dict = field(default_factory=dict)      # type -> (alpha, thole, (group types))

    def key(self, name, default=None):
        v = self.keys.get(name)
        return default if v is None else v[0]


def _frame_of(signed):
    """Frame from the signs of the reference types (MultipoleType.java:818-842)."""
    n = len(signed)
    if n == 1:
        return FRAME_NONE
    if n == 2:
        return FRAME_ZONLY
    if n == 3:
        return FRAME_BISECTOR if (signed[1] < 0 or signed[2] < 0) else FRAME_ZTHENX
    if n == 4 and signed[2] < 0 and signed[3] < 0:
        return FRAME_THREEFOLD if signed[1] < 0 else FRAME_ZTHENBISECTOR
    return FRAME_ZTHENX


def read_prm(src) -> ForceFieldParams:
    ff = ForceFieldParams()
    with _open(src) as f:
        lines = [ln.split("#")[0].rstrip() for ln in f]
    i = 0
    while i < len(lines):
        raw = lines[i]
        i += 1
        t = raw.replace("\\", " ").split()
        if not t:
            continue
        kw = t[0].lower()
        if kw == "atom":
            ff.atom[int(t[1])] = (int(t[2]), t[3])
            q = raw.split('"')
            if len(q) >= 3:
                rest = q[2].split()
                if len(rest) >= 3:
                    ff.element[int(t[1])] = (q[1], int(rest[0]), float(rest[1]), int(rest[2]))
        elif kw in _VALENCE:
            nc = _VALENCE[kw]
            ff.valence.setdefault(kw, []).append((tuple(int(x) for x in t[1:1 + nc]),
                                                  [float(x) for x in t[1 + nc:]]))
        elif kw == "tortors":
            cls, nx, ny = tuple(int(x) for x in t[1:6]), int(t[6]), int(t[7])
            vals = []
            while len(vals) < 3 * nx * ny:
                vals += [float(x) for x in lines[i].replace("\\", " ").split()]
                i += 1
            ff.tortors.append((cls, nx, ny, [tuple(vals[3 * k:3 * k + 3]) for k in range(nx * ny)]))
        elif kw == "vdw":
            red = float(t[4]) if len(t) > 4 else 0.0
            ff.vdw[int(t[1])] = (float(t[2]), float(t[3]), red)
        elif kw in ("vdwpr", "vdwpair"):
            a, b = int(t[1]), int(t[2])
            ff.vdwpr[(min(a, b), max(a, b))] = (float(t[3]), float(t[4]))
        elif kw == "multipole":
            signed = [int(x) for x in t[1:-1] if int(x) != 0]
            charge = float(t[-1])
            vals = []
            while len(vals) < 9:
                vals += [float(x) for x in lines[i].replace("\\", " ").split()]
                i += 1
            d = vals[0:3]
            qxx, qyx, qyy, qzx, qzy, qzz = vals[3:9]
            ff.multipole[tuple(abs(x) for x in signed)] = (_frame_of(signed),
                                                           [charge] + d + [qxx, qyy, qzz, qyx, qzx, qzy])
        elif kw == "polarize":
            ff.polarize[int(t[1])] = (float(t[2]), float(t[3]), tuple(int(x) for x in t[4:]))
        elif len(t) >= 2 and kw in _SCALARS:
            ff.keys[kw] = t[1:]
    return ff


_SCALARS = {"vdwtype", "radiusrule", "radiustype", "radiussize", "epsilonrule", "polarization", "dielectric",
            "vdw-12-scale", "vdw-13-scale", "vdw-14-scale", "vdw-15-scale", "mpole-12-scale", "mpole-13-scale",
            "mpole-14-scale", "mpole-15-scale", "polar-12-scale", "polar-13-scale", "polar-14-scale",
            "polar-15-scale", "polar-12-intra", "polar-13-intra", "polar-14-intra", "polar-15-intra",
            "direct-11-scale", "direct-12-scale", "direct-13-scale", "direct-14-scale", "mutual-11-scale",
            "mutual-12-scale", "mutual-13-scale", "mutual-14-scale", "forcefield",
            "bondunit", "bond-cubic", "bond-quartic", "angleunit", "angle-cubic", "angle-quartic", "angle-pentic",
            "angle-sextic", "strbndunit", "ureyunit", "urey-cubic", "urey-quartic", "opbendtype", "opbendunit",
            "opbend-cubic", "opbend-quartic", "opbend-pentic", "opbend-sextic", "torsionunit", "pitorsunit",
            "tortorunit", "torsion-scale", "halgren-delta", "halgren-gamma", "electric"}

# valence records and their number of atom classes (parameters/*Type.parse)
_VALENCE = {"bond": 2, "angle": 3, "anglep": 3, "strbnd": 3, "ureybrad": 3, "opbend": 4, "torsion": 4,
            "pitors": 2}


def write_prm(path, ff: ForceFieldParams, types=None, exact=False):
    """Parameter file in the same record formats (only the records of the
    given atom types and their classes when types is set).  With exact set,
    the vdW, multipole and polarize values are written as the shortest
    decimal that reads back to the same double, instead of the files' fixed
    five and nine decimals."""
    tset = None if types is None else set(int(t) for t in types)
    cset = None if tset is None else {ff.atom[t][0] for t in tset}
    if exact:
        mp = rad = pol = lambda v: f"{float(v)!r:>12s}"
    else:
        mp, rad, pol = (lambda v: f"{v:12.5f}"), (lambda v: f"{v:.9f}"), (lambda v: f"{v:.5f}")
    with _create(path) as f:
        for k, v in ff.keys.items():
            f.write(f"{k:<23s} {' '.join(v)}\n")
        for t, (c, nm) in sorted(ff.atom.items()):
            if tset is None or t in tset:
                el = ff.element.get(t)
                tail = f'  "{el[0]}" {el[1]:4d} {el[2]:9.4f} {el[3]:3d}' if el else ""
                f.write(f"atom {t:6d} {c:6d}  {nm}{tail}\n")
        for kw in _VALENCE:
            for cls, vals in ff.valence.get(kw, []):
                if cset is None or all(c in cset or c == 0 for c in cls):
                    f.write(f"{kw:<8s}" + "".join(f" {c:6d}" for c in cls) + "".join(f" {v:.9g}" for v in vals)
                            + "\n")
        for cls, nx, ny, grid in ff.tortors:
            if cset is None or all(c in cset for c in cls):
                f.write("tortors " + "".join(f" {c:6d}" for c in cls) + f" {nx:3d} {ny:3d}\n")
                for t1, t2, e in grid:
                    f.write(f"  {t1:.9g} {t2:.9g} {e:.9g}\n")
        for c, (r, e, red) in sorted(ff.vdw.items()):
            if cset is None or c in cset:
                red_s = (f" {float(red)!r}" if exact else f" {red:.3f}") if red > 0 else ""
                f.write(f"vdw {c:6d} {rad(r)} {rad(e)}{red_s}\n")
        for (a, b), (r, e) in sorted(ff.vdwpr.items()):
            if cset is None or (a in cset and b in cset):
                f.write(f"vdwpr {a:6d} {b:6d} {rad(r)} {rad(e)}\n")
        for key, (fr, v) in sorted(ff.multipole.items()):
            if tset is not None and key[0] not in tset:
                continue
            signed = list(key)
            if fr == FRAME_BISECTOR:
                signed[1], signed[2] = -signed[1], -signed[2]
            elif fr == FRAME_ZTHENBISECTOR:
                signed[2], signed[3] = -signed[2], -signed[3]
            elif fr == FRAME_THREEFOLD:
                signed[1], signed[2], signed[3] = -signed[1], -signed[2], -signed[3]
            q, d, Q = v[0], v[1:4], v[4:10]   # Qxx Qyy Qzz Qxy Qxz Qyz
            f.write("multipole " + " ".join(f"{x:6d}" for x in signed) + f" {mp(q)}\n")
            f.write(f"    {mp(d[0])} {mp(d[1])} {mp(d[2])}\n")
            f.write(f"    {mp(Q[0])}\n    {mp(Q[3])} {mp(Q[1])}\n    {mp(Q[4])} {mp(Q[5])} {mp(Q[2])}\n")
        for t, (a, th, grp) in sorted(ff.polarize.items()):
            if tset is None or t in tset:
                f.write(f"polarize {t:6d} {pol(a)} {pol(th)}" + "".join(f" {g:6d}" for g in grp) + "\n")


# --------------------------------------------------------------------------
# topology
# --------------------------------------------------------------------------
def _neighbours(n, bonds):
    n12 = [[] for _ in range(n)]
    for i, j in bonds:
        n12[i].append(j)
        n12[j].append(i)
    n12 = [sorted(a) for a in n12]
    n13 = []
    for i in range(n):
        s = set(n12[i])
        lst = [b for a in n12[i] for b in n12[a] if b != i and b not in s]
        n13.append(sorted(lst))
    return n12, n13


def assign_multipoles(types, n12, n13, ff: ForceFieldParams):
    """Frame, axis atoms and local multipole (Bohr units) per atom, in the
    matching order of MultipoleType.multipoleTypeFactory."""
    n = len(types)
    frame = np.zeros(n, np.int32)
    axis = -np.ones((n, 3), np.int32)
    mp = np.zeros((n, 10))
    mt = ff.multipole
    for i in range(n):
        t = int(types[i])
        hit = None
        if (t,) in mt:
            hit = ((t,), ())
        if hit is None:
            for a in n12[i]:
                key = (t, int(types[a]))
                if key in mt:
                    hit = (key, (a,))
                    break
        if hit is None:
            for a in n12[i]:
                for b in n12[i]:
                    if a == b:
                        continue
                    key = (t, int(types[a]), int(types[b]))
                    if key in mt:
                        hit = (key, (a, b))
                        break
                if hit:
                    break
        if hit is None:
            for a in n12[i]:
                for b in n12[i]:
                    if a == b:
                        continue
                    for c in n12[i]:
                        if c == a or c == b:
                            continue
                        key = (t, int(types[a]), int(types[b]), int(types[c]))
                        if key in mt:
                            hit = (key, (a, b, c))
                            break
                    if hit:
                        break
                if hit:
                    break
        if hit is None:
            for a in n12[i]:
                for b in n13[i]:
                    key = (t, int(types[a]), int(types[b]))
                    if key in mt:
                        hit = (key, (a, b))
                        break
                    for c in n13[i]:
                        if c != b:
                            key = (t, int(types[a]), int(types[b]), int(types[c]))
                            if key in mt:
                                hit = (key, (a, b, c))
                                break
                    if hit:
                        break
                if hit:
                    break
        if hit is None:
            raise ValueError(f"no multipole type for atom {i + 1} (type {t})")
        key, atoms = hit
        fr, vals = mt[key]
        frame[i] = fr
        for k, a in enumerate(atoms):
            axis[i, k] = a
        mp[i] = vals
    return frame, axis, mp


def polarization_groups(types, n12, ff: ForceFieldParams):
    """Group id per atom from PolarizeType.assignPolarizationGroups (the
    group of each atom grown over bonds to atoms of the listed types)."""
    n = len(types)
    gid = -np.ones(n, np.int64)
    for i in range(n):
        pol = ff.polarize.get(int(types[i]))
        allowed = [int(types[i])] + ([g for g in pol[2]] if pol else [])
        group = [i]
        stack = [i]
        if pol and pol[2]:
            while stack:
                seed = stack.pop()
                for j in n12[seed]:
                    if int(types[j]) in allowed and j not in group:
                        group.append(j)
                        pj = ff.polarize.get(int(types[j]))
                        for g in (pj[2] if pj else ()):
                            if g not in allowed:
                                allowed.append(g)
                        stack.append(j)
        gid[i] = min(group)
    return gid


def _f(keys, name, default):
    v = keys.get(name)
    return float(v[0]) if v else default


def _flag(keys, fk, name, default):
    """A boolean keyword of the key file, else of the parameter file: a bare
    keyword is true, a value reads as true/false (ForceField.getBoolean)."""
    v = keys.get(name, fk.get(name))
    if v is None:
        return default
    return True if not v else v[0].lower() in ("true", "1", "yes")


def parse_atom_ranges(spec, n):
    """0-based atom indices of a range list in the form of the reference's
    --ac option (AlchemicalOptions.java:298): ALL, NONE, or comma-separated
    1-based atoms and ranges such as "1-20,31,40-N"."""
    s = str(spec).strip().upper()
    if s in ("", "NONE"):
        return np.zeros(0, np.int64)
    if s == "ALL":
        return np.arange(n)
    out = []
    for part in s.split(","):
        a, _, b = part.strip().partition("-")
        lo = n if a == "N" else int(a)
        hi = lo if not b else (n if b == "N" else int(b))
        if not 1 <= lo <= hi <= n:
            raise ValueError(f"atom range {part!r} outside 1..{n}")
        out.extend(range(lo - 1, hi))
    return np.unique(np.asarray(out, np.int64))


def format_atom_ranges(idx):
    """The inverse of parse_atom_ranges for a set of 0-based indices."""
    idx = sorted(int(i) for i in idx)
    if not idx:
        return "NONE"
    parts, start, prev = [], idx[0], idx[0]
    for i in idx[1:] + [None]:
        if i is not None and i == prev + 1:
            prev = i
            continue
        parts.append(f"{start + 1}" if start == prev else f"{start + 1}-{prev + 1}")
        if i is not None:
            start = prev = i
    return ",".join(parts)


def molecules_from_bonds(n, bonds):
    """Molecule number per atom: bonded components numbered in the order of
    their first atom (the molecules of a MolecularAssembly)."""
    parent = list(range(n))

    def root(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i, j in bonds:
        ri, rj = root(i), root(j)
        if ri != rj:
            parent[max(ri, rj)] = min(ri, rj)
    mol = np.empty(n, np.int32)
    num = {}
    for a in range(n):
        mol[a] = num.setdefault(root(a), len(num))
    return mol


def build(xf: XYZFile, ff: ForceFieldParams, keys: dict, alchemical_atoms=None, lambda_value=None):
    """(xyz (n,3), box or None, Topology, NonbondedParams).

    Settings come from the key file, then the parameter file's scalar
    keywords, then the reference's defaults; `alchemical_atoms` (a range
    list, as the reference's --ac option) marks the soft atoms and
    `lambda_value` sets the state (-l)."""
    n = len(xf.names)
    types = xf.types
    n12, n13 = _neighbours(n, xf.bonds)
    frame, axis, mpb = assign_multipoles(types, n12, n13, ff)
    mpole = np.zeros((n, 10))
    mpole[:, 0] = mpb[:, 0]
    mpole[:, 1:4] = mpb[:, 1:4] * BOHR
    mpole[:, 4:10] = mpb[:, 4:10] * BOHR * BOHR
    polarity = np.array([ff.polarize.get(int(t), (0.0, 0.0, ()))[0] for t in types])
    thole = np.array([ff.polarize.get(int(t), (0.0, 0.0, ()))[1] for t in types])
    pdamp = np.where(polarity > 0, polarity ** (1.0 / 6.0), 0.0)
    group = polarization_groups(types, n12, ff)
    # vdW classes (compacted) and reduced sites
    cls = np.array([ff.atom[int(t)][0] for t in types])
    used = sorted(set(int(c) for c in cls))
    cidx = {c: k for k, c in enumerate(used)}
    vclass = np.array([cidx[int(c)] for c in cls], np.int32)
    red_index = np.arange(n, dtype=np.int32)
    red_value = np.zeros(n)
    for i in range(n):
        r = ff.vdw[int(cls[i])][2]
        if r > 0.0 and len(n12[i]) == 1:
            red_index[i] = n12[i][0]
            red_value[i] = r
    rule_r = ff.key("radiusrule", "CUBIC-MEAN").upper()
    rule_e = ff.key("epsilonrule", "HHG").upper()
    size = ff.key("radiussize", "DIAMETER").upper()
    rtype = ff.key("radiustype", "R-MIN").upper()
    scale = 0.5 if size == "DIAMETER" else 1.0
    if rtype == "SIGMA":
        scale *= 1.122462048309372981
    nc = len(used)
    tab = np.zeros((nc, nc, 2))
    for a, ca in enumerate(used):
        ra, ea = ff.vdw[ca][0] * scale, ff.vdw[ca][1]
        for b, cb in enumerate(used):
            rb, eb = ff.vdw[cb][0] * scale, ff.vdw[cb][1]
            pr = ff.vdwpr.get((min(ca, cb), max(ca, cb)))
            if pr is not None:
                # a pair record gives Rmin itself (VanDerWaalsForm.java:442-460)
                rm, e = pr[0], pr[1]
            else:
                rm = combined_radius(ra, rb, rule_r)
                e = combined_eps(ea, eb, ra, rb, rule_e)
            tab[a, b] = (1.0 / rm if rm > 0 else 0.0, e)
    off, atoms, rels = masks_from_bonds(n, xf.bonds)
    top = Topology(n=n, frame=frame, axis=axis, mpole=mpole, polarity=polarity, thole=thole, pdamp=pdamp,
                   group=group.astype(np.int32), molecule=molecules_from_bonds(n, xf.bonds), vdw_class=vclass,
                   vdw_table=tab, red_index=red_index, red_value=red_value, mask_off=off, mask_atom=atoms,
                   mask_rel=rels)
    if alchemical_atoms is not None:
        top.soft[parse_atom_ranges(alchemical_atoms, n)] = 1
    # cell and settings
    box = xf.box
    if "a-axis" in keys:
        a = _f(keys, "a-axis", 0.0)
        box = [a, _f(keys, "b-axis", a), _f(keys, "c-axis", a), _f(keys, "alpha", 90.0), _f(keys, "beta", 90.0),
               _f(keys, "gamma", 90.0)]
    fk = ff.keys
    kv = {**fk, **keys}     # the key file overrides the parameter file

    def num(name, default):
        return _f(kv, name, default)

    pol = (kv.get("polarization") or ["MUTUAL"])[0].upper()
    if not _flag(keys, fk, "polarizeterm", True):
        pol = "NONE"
    p = NonbondedParams(
        ewald_cutoff=num("ewald-cutoff", 7.0), ewald_alpha=num("ewald-alpha", 0.545),
        vdw_cutoff=num("vdw-cutoff", 12.0), vdw_taper=num("vdw-taper", 0.9), list_buffer=num("list-buffer", 2.0),
        pme_order=int(num("pme-order", 5)), pme_mesh_density=num("pme-mesh-density", 1.2),
        polar_eps=num("polar-eps", 1e-6), electric=num("electric", NonbondedParams.electric),
        m_scale=tuple(num(f"mpole-1{k}-scale", d) for k, d in zip("2345", (0.0, 0.0, 0.4, 0.8))),
        p_scale=tuple(num(f"polar-1{k}-scale", d) for k, d in zip("2345", (0.0, 0.0, 1.0, 1.0))),
        intra14=num("polar-14-intra", 0.5), d11=num("direct-11-scale", 0.0),
        v_scale=tuple(num(f"vdw-1{k}-scale", d) for k, d in zip("234", (0.0, 0.0, 1.0))),
        halgren_delta=num("halgren-delta", 0.07), halgren_gamma=num("halgren-gamma", 0.12),
        polarization={"NONE": 0, "DIRECT": 1, "MUTUAL": 2}.get(pol, 2),
        do_vdw=_flag(keys, fk, "vdwterm", True), do_pme=_flag(keys, fk, "mpoleterm", True),
        scf_algorithm=1 if (kv.get("scf-algorithm") or ["CG"])[0].upper() == "SOR" else 0,
        polar_sor=num("polar-sor", 0.7), sor_fallback=_flag(keys, fk, "sor-scf-fallback", True),
        precond_cutoff=num("cg-preconditioner-cutoff", 4.5), precond_scale=num("cg-preconditioner-scale", 2.0),
        # alchemical settings (VanDerWaals.java:326-357, AlchemicalParameters.java:171-251)
        vdw_lambda_alpha=num("vdw-lambda-alpha", 0.25), vdw_lambda_exponent=num("vdw-lambda-exponent", 3.0),
        vdw_lambda_end=num("vdw-lambda-end", 1.0),
        intermolecular_softcore=_flag(keys, fk, "intermolecular-softcore", False),
        intramolecular_softcore=_flag(keys, fk, "intramolecular-softcore", False),
        perm_lambda_alpha=num("permanent-lambda-alpha", 2.0),
        perm_lambda_exponent=num("permanent-lambda-exponent", 3.0),
        perm_lambda_start=num("permanent-lambda-start", 0.4), perm_lambda_end=num("permanent-lambda-end", 1.0),
        pol_lambda_exponent=num("polarization-lambda-exponent", 3.0),
        pol_lambda_start=num("polarization-lambda-start", 0.75),
        pol_lambda_end=num("polarization-lambda-end", 1.0),
        elec_lambda=_flag(keys, fk, "elec-lambdaterm", False),
        ligand_vapor_elec=_flag(keys, fk, "ligand-vapor-elec", True),
        no_ligand_condensed_scf=_flag(keys, fk, "no-ligand-condensed-scf", True))
    if lambda_value is not None:
        p = replace(p, lambda_value=float(lambda_value))
    if "pme-grid" in keys:
        g = [int(v) for v in keys["pme-grid"]]
        p = replace(p, grid=tuple(g if len(g) == 3 else g * 3))
    if box is not None and box[0] > 0:
        p = replace(p, grid=p.resolved_grid(box))
    return xf.xyz.copy(), box, top, p


def masses(types, ff: ForceFieldParams):
    """Atomic masses (g/mol) of the atoms' types from the `atom` records."""
    return np.array([ff.element[int(t)][2] for t in types])


def parameters_path(key_path, keys):
    """The parameter file a key file names with `parameters`, relative to the
    key file's folder."""
    import os
    v = keys.get("parameters")
    if not v:
        raise ValueError(f"{key_path}: no parameters keyword and no parameter file given")
    name = " ".join(v)
    return name if os.path.isabs(name) else os.path.join(os.path.dirname(os.path.abspath(str(key_path))), name)


def load(xyz_path, key_path, prm_path=None, valence=False, alchemical_atoms=None, lambda_value=None):
    """Read the three files and build (xyz, box, Topology, NonbondedParams),
    with the valence terms (forcefieldx_amd.bonded.ValenceTerms) appended
    when asked.  Without prm_path the key file's `parameters` file is read."""
    keys = read_keys(key_path)
    if prm_path is None:
        prm_path = parameters_path(key_path, keys)
    xf, ff = read_xyz(xyz_path), read_prm(prm_path)
    out = build(xf, ff, keys, alchemical_atoms=alchemical_atoms, lambda_value=lambda_value)
    if not valence:
        return out
    from .bonded import assign
    return out + (assign(xf.types, xf.bonds, ff, keys),)


def load_crystal(xyz_path, key_path, prm_path):
    """A crystal of the key file's space group expanded to its P1 cell
    (forcefieldx_amd.crystal): (xyz, box, Topology, NonbondedParams,
    ValenceTerms, nsymm).  Per-asymmetric-unit energies are the P1 energies
    divided by nsymm; the cell comes from the keywords."""
    from .bonded import assign
    from .crystal import expand_to_p1
    xf, ff, keys = read_xyz(xyz_path), read_prm(prm_path), read_keys(key_path)
    box = [_f(keys, "a-axis", 0.0), _f(keys, "b-axis", 0.0), _f(keys, "c-axis", 0.0), _f(keys, "alpha", 90.0),
           _f(keys, "beta", 90.0), _f(keys, "gamma", 90.0)]
    box[1] = box[1] or box[0]
    box[2] = box[2] or box[0]
    sg = (keys.get("spacegroup") or ["P1"])[0]
    x, t, b, names, nsymm = expand_to_p1(xf.xyz, xf.types, xf.bonds, xf.names, box, sg)
    p1 = XYZFile(xf.title, names, x, t, b, box)
    keys = {k: v for k, v in keys.items() if k != "spacegroup"}
    xyz, bx, top, params = build(p1, ff, keys)
    return xyz, bx, top, params, assign(p1.types, p1.bonds, ff, keys), nsymm


# --------------------------------------------------------------------------
# writing a system
# --------------------------------------------------------------------------
HYDROGEN = ("H", 1, 1.008)
CARBON = ("C", 6, 12.011)
OXYGEN = ("O", 8, 15.999)


def bonds_of(top: Topology):
    """(i, j) pairs, i < j, of the Topology's 1-2 relations."""
    out = []
    for i in range(top.n):
        for k in range(int(top.mask_off[i]), int(top.mask_off[i + 1])):
            j = int(top.mask_atom[k])
            if int(top.mask_rel[k]) == 2 and j > i:
                out.append((i, j))
    return out


def _canonical_groups(group):
    """Each atom's group named by the group's first atom."""
    first = {}
    for a, g in enumerate(np.asarray(group).tolist()):
        first.setdefault(g, a)
    return np.array([first[g] for g in np.asarray(group).tolist()])


def _tidy(v):
    """v, or its 12-significant-digit form when that is within 2e-16 of it
    (values that went through a unit conversion read back as written)."""
    v = float(v)
    w = float(f"{v:.12g}")
    return w if abs(w - v) <= 2e-16 * abs(v) else v


def _records(top, types, n12, elements):
    """ForceFieldParams holding one atom/multipole/polarize record set for
    the given per-atom types (multipoles back in Bohr units)."""
    ff = ForceFieldParams()
    for i in range(top.n):
        t = int(types[i])
        el = elements[i]
        ff.atom.setdefault(t, (int(top.vdw_class[i]) + 1, el[0]))
        ff.element.setdefault(t, (el[0], int(el[1]), float(el[2]), len(n12[i])))
        key = (t,) + tuple(int(types[a]) for a in top.axis[i] if a >= 0)
        v = np.asarray(top.mpole[i], float)
        ff.multipole[key] = (int(top.frame[i]), [_tidy(x) for x in [v[0]] + (v[1:4] / BOHR).tolist()
                                                  + (v[4:10] / (BOHR * BOHR)).tolist()])
        grp = {int(types[j]) for j in n12[i] if top.group[j] == top.group[i]}
        old = ff.polarize.get(t)
        if old is not None:
            grp |= set(old[2])
        ff.polarize[t] = (float(top.polarity[i]), float(top.thole[i]), tuple(sorted(grp)))
    return ff


def system_types(top: Topology, elements):
    """Atom types and the ForceFieldParams records (atom, multipole, polarize)
    that give back the Topology's frames, axis atoms, local moments,
    polarizabilities and polarization groups under the reference's matching
    rules (assign_multipoles, polarization_groups).

    Atoms that agree in element, vdW class, reduction, polarizability, Thole
    factor, frame and moments, and whose axis atoms agree in the same, share a
    type; an atom the shared records would match differently (another axis
    atom, a group grown too far) gets a type of its own, until every atom
    matches.  A frame whose axis atoms are not 1-2 or 1-3 neighbours, or a
    polarization group that is not connected by bonds, cannot be written and
    raises ValueError."""
    n = top.n
    bonds = bonds_of(top)
    n12, n13 = _neighbours(n, bonds)
    own = [(elements[i], int(top.vdw_class[i]), float(top.red_value[i]), float(top.polarity[i]),
            float(top.thole[i]), int(top.frame[i]), tuple(np.asarray(top.mpole[i], float).tolist()))
           for i in range(n)]
    sig = [(own[i],) + tuple(own[a] for a in top.axis[i] if a >= 0) for i in range(n)]
    ids = {}
    types = np.array([ids.setdefault(s, len(ids) + 1) for s in sig], np.int64)
    want_groups = _canonical_groups(top.group)
    for attempt in range(4):
        ff = _records(top, types, n12, elements)
        try:
            frame, axis, mpb = assign_multipoles(types, n12, n13, ff)
        except ValueError as exc:
            raise ValueError(f"a frame cannot be written as multipole records: {exc}") from exc
        mp = np.concatenate([mpb[:, :1], mpb[:, 1:4] * BOHR, mpb[:, 4:10] * (BOHR * BOHR)], axis=1)
        bad = (frame != top.frame) | np.any(axis != top.axis, axis=1)
        bad |= ~np.all(np.isclose(mp, top.mpole, rtol=1e-13, atol=1e-15), axis=1)
        bad |= _canonical_groups(polarization_groups(types, n12, ff)) != want_groups
        if not bad.any():
            return types, ff, bonds, n12
        if attempt == 3:
            break
        nxt = int(types.max()) + 1
        for i in np.nonzero(bad)[0]:
            types[i] = nxt
            nxt += 1
    raise ValueError(f"{int(bad.sum())} atoms keep frames or polarization groups the records cannot express "
                     f"(first: atom {int(np.nonzero(bad)[0][0]) + 1})")


def _vdw_records(top: Topology, n12, ff: ForceFieldParams):
    """vdw records per class from the pair table's diagonal (R-min diameter,
    CUBIC-MEAN / HHG rules), vdwpr records where the table differs from the
    combined value, and the class's reduction factor."""
    used = sorted(set(int(c) for c in top.vdw_class))
    red = {}
    for i in range(top.n):
        c = int(top.vdw_class[i])
        rv = float(top.red_value[i])
        if rv > 0.0:
            if len(n12[i]) != 1 or int(top.red_index[i]) != n12[i][0]:
                raise ValueError(f"atom {i + 1}: a reduced vdW site must have one bond and reduce toward it")
        if len(n12[i]) == 1 or rv > 0.0:
            if red.setdefault(c, rv) != rv:
                raise ValueError(f"vdW class {c}: singly bonded atoms with different reductions")
    tab = np.asarray(top.vdw_table, float)
    for c in used:
        inv, eps = tab[c, c]
        if inv <= 0.0:
            raise ValueError(f"vdW class {c}: zero radius")
        ff.vdw[c + 1] = (_tidy(1.0 / inv), _tidy(eps), red.get(c, 0.0))
    for a in used:
        for b in used:
            if b < a:
                continue
            ra, ea = ff.vdw[a + 1][0] * 0.5, ff.vdw[a + 1][1]
            rb, eb = ff.vdw[b + 1][0] * 0.5, ff.vdw[b + 1][1]
            rm = combined_radius(ra, rb, "CUBIC-MEAN")
            e = combined_eps(ea, eb, ra, rb, "HHG")
            inv, ee = tab[a, b]
            if not (np.isclose(1.0 / rm, inv, rtol=1e-14, atol=0.0) and np.isclose(e, ee, rtol=1e-14, atol=0.0)):
                ff.vdwpr[(a + 1, b + 1)] = (1.0 / inv, float(ee))
    ff.keys.update({"vdwtype": ["BUFFERED-14-7"], "radiusrule": ["CUBIC-MEAN"], "radiustype": ["R-MIN"],
                    "radiussize": ["DIAMETER"], "epsilonrule": ["HHG"]})


def _key_lines(box, p: NonbondedParams, prm_name, title):
    def r(v):
        return repr(_tidy(v))

    out = [f"parameters {prm_name}"]
    if title:
        out.append(f"# {title}")
    if box is not None:
        for k, v in zip(("a-axis", "b-axis", "c-axis", "alpha", "beta", "gamma"), box):
            out.append(f"{k} {r(v)}")
    vals = [("ewald-cutoff", p.ewald_cutoff), ("ewald-alpha", p.ewald_alpha), ("vdw-cutoff", p.vdw_cutoff),
            ("vdw-taper", p.vdw_taper), ("list-buffer", p.list_buffer), ("pme-order", p.pme_order),
            ("polar-eps", p.polar_eps), ("polar-sor", p.polar_sor), ("electric", p.electric),
            ("cg-preconditioner-cutoff", p.precond_cutoff), ("cg-preconditioner-scale", p.precond_scale)]
    for k, v in vals:
        out.append(f"{k} {v if isinstance(v, int) else r(v)}")
    if all(int(g) > 0 for g in p.grid):
        out.append("pme-grid " + " ".join(str(int(g)) for g in p.grid))
    else:
        out.append(f"pme-mesh-density {r(p.pme_mesh_density)}")
    out.append(f"polarization {('NONE', 'DIRECT', 'MUTUAL')[int(p.polarization)]}")
    out.append(f"scf-algorithm {('CG', 'SOR')[int(p.scf_algorithm)]}")
    tf = lambda b: "true" if b else "false"
    for k, v in (("sor-scf-fallback", p.sor_fallback), ("vdwterm", p.do_vdw), ("mpoleterm", p.do_pme)):
        out.append(f"{k} {tf(v)}")
    for k, v in zip(("2", "3", "4", "5"), p.m_scale):
        out.append(f"mpole-1{k}-scale {r(v)}")
    for k, v in zip(("2", "3", "4", "5"), p.p_scale):
        out.append(f"polar-1{k}-scale {r(v)}")
    for k, v in zip(("2", "3", "4"), p.v_scale):
        out.append(f"vdw-1{k}-scale {r(v)}")
    out += [f"polar-14-intra {r(p.intra14)}", f"direct-11-scale {r(p.d11)}",
            f"halgren-delta {r(p.halgren_delta)}", f"halgren-gamma {r(p.halgren_gamma)}"]
    alch = [("vdw-lambda-alpha", p.vdw_lambda_alpha), ("vdw-lambda-exponent", p.vdw_lambda_exponent),
            ("vdw-lambda-end", p.vdw_lambda_end), ("permanent-lambda-alpha", p.perm_lambda_alpha),
            ("permanent-lambda-exponent", p.perm_lambda_exponent), ("permanent-lambda-start", p.perm_lambda_start),
            ("permanent-lambda-end", p.perm_lambda_end), ("polarization-lambda-exponent", p.pol_lambda_exponent),
            ("polarization-lambda-start", p.pol_lambda_start), ("polarization-lambda-end", p.pol_lambda_end)]
    for k, v in alch:
        out.append(f"{k} {r(v)}")
    for k, v in (("intermolecular-softcore", p.intermolecular_softcore),
                 ("intramolecular-softcore", p.intramolecular_softcore), ("elec-lambdaterm", p.elec_lambda),
                 ("ligand-vapor-elec", p.ligand_vapor_elec), ("no-ligand-condensed-scf", p.no_ligand_condensed_scf)):
        out.append(f"{k} {tf(v)}")
    return out


def write_system(prefix, xyz, box, top: Topology, params: NonbondedParams, title="", elements=None, digits=8):
    """Write a system held as arrays (a Topology, coordinates, cell and
    settings) as Tinker files that `load` reads back to the same energy
    function, the way the reference writes a coordinate file with its key
    file and parameter patch (XYZFilter.writeFile, XYZFilter.java:639;
    KeyFilter.java): `prefix.xyz` (coordinates with `digits` decimals, atom
    types, bonds from the 1-2 relations), `prefix.key` (cell, cutoffs, grid,
    convergence, mask scales, alchemical settings; its `parameters` line names
    the patch) and `prefix.prm` (atom, vdw, vdwpr, multipole and polarize
    records, values written exactly).

    `elements` gives (symbol, atomic number, mass) per atom (default: H for
    reduced vdW sites, C otherwise).  The soft atoms and the state are not
    keywords in the reference (its --ac and -l options): they are returned as
    `alchemical_atoms` and `lambda_value`, the arguments `load` takes for
    them.  Returns {"xyz", "key", "prm": paths, "alchemical_atoms",
    "lambda_value"}."""
    import os
    n = top.n
    if elements is None:
        elements = [HYDROGEN if top.red_value[i] > 0 else CARBON for i in range(n)]
    elements = [tuple(e) for e in elements]
    types, ff, bonds, n12 = system_types(top, elements)
    _vdw_records(top, n12, ff)
    x = np.asarray(xyz, float).reshape(n, 3)
    paths = {"xyz": f"{prefix}.xyz", "key": f"{prefix}.key", "prm": f"{prefix}.prm"}
    write_xyz(paths["xyz"], XYZFile(title or "system", [e[0] for e in elements], x, types, bonds,
                                    None if box is None else [float(v) for v in box]), digits=digits)
    write_prm(paths["prm"], ff, exact=True)
    with _create(paths["key"]) as f:
        f.write("\n".join(_key_lines(box, params, os.path.basename(paths["prm"]), title)) + "\n")
    soft = np.nonzero(np.asarray(top.soft))[0]
    paths["alchemical_atoms"] = format_atom_ranges(soft) if soft.size else None
    paths["lambda_value"] = float(params.lambda_value)
    return paths
