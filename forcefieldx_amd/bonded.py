"""AMOEBA valence terms of a Tinker system, assigned as Force Field X assigns
them (the terms ForceFieldEnergy adds beside the nonbonded path).

Enumeration follows MSGroup.finalize (bonded/MSGroup.java:190-390) over the
whole bond graph:

* bonds: every bond of the coordinate file, type by the sorted class pair
  (BondType.sortKey, BondType.java:249-262);
* angles: every pair of bonds at an atom; type by the classes with the ends
  sorted (AngleType.sortKey, AngleType.java:360-370), from the `angle`
  records, else the `anglep` (in-plane) records (ForceField.getAngleType,
  ForceField.java:470-487); the ideal angle is chosen by the number of
  hydrogens on the centre outside the angle (Angle.setAngleType,
  Angle.java:390-410);
* stretch-bends and Urey-Bradleys: one per angle whose key has a record
  (StretchBend.java:95-103, 193-208; UreyBradley.java:95-104);
* out-of-plane bends: one per angle at an atom with three bonds, the
  fourth atom out of plane, key (fourth, centre, end, end), then the ends
  swapped, then (fourth, centre, 0, 0) (OutOfPlaneBend.java:134-150,
  ForceField.java:978-1000); an in-plane angle takes that fourth atom as its
  projection atom;
* torsions: around every bond, every pair of outer bonds, with the single
  and double wild-card lookups of ForceField.getTorsionType
  (ForceField.java:1280-1309; key order TorsionType.sortKey, :303-340);
* pi-orbital torsions: every bond between two atoms with three bonds each
  (PiOrbitalTorsion.java:89-141);
* torsion-torsions: every five-atom chain around an angle whose classes
  match a `tortors` record forwards or reversed (MSGroup.java:353-378,
  TorsionTorsion.java:142-162), with the chirality atom of the centre
  (TorsionTorsion.chktor, :489-555) and the grid's spline derivatives
  (TorsionTorsionType.java:107-284).

Units: angle and out-of-plane force constants x (pi/180)^2, stretch-bend
x pi/180, torsion amplitudes x `torsionunit` (AngleType.java:98,
OutOfPlaneBendType.java:163, StretchBendType.java:82, ForceField.java:1238).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MAX_TORSION_TERMS = 6


@dataclass
class TortorGrid:
    nx: int
    ny: int
    tx: np.ndarray      # (nx,) degrees
    ty: np.ndarray      # (ny,)
    e: np.ndarray       # (nx * ny,) energy at (tx[i], ty[j]) in e[j * nx + i]
    dx: np.ndarray      # dE/dt1 at the grid points
    dy: np.ndarray      # dE/dt2
    dxy: np.ndarray     # d2E/dt1 dt2


@dataclass
class ValenceTerms:
    n: int
    bond: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    bond_k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    bond_r0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    bond_poly: tuple = (0.0, 0.0)                       # cubic, quartic
    angle: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))   # a, centre, c, in-plane or -1
    angle_k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    angle_t0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    angle_poly: tuple = (0.0, 0.0, 0.0, 0.0)            # cubic .. sextic
    strbnd: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    strbnd_k: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    strbnd_r0: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    strbnd_t0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    urey: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int32))
    urey_k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    urey_r0: np.ndarray = field(default_factory=lambda: np.zeros(0))
    urey_poly: tuple = (0.0, 0.0)
    opbend: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))  # a, centre, c, out-of-plane
    opbend_k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    opbend_poly: tuple = (0.0, 0.0, 0.0, 0.0)
    torsion: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    torsion_amp: np.ndarray = field(default_factory=lambda: np.zeros((0, MAX_TORSION_TERMS)))
    torsion_cos: np.ndarray = field(default_factory=lambda: np.zeros((0, MAX_TORSION_TERMS)))
    torsion_sin: np.ndarray = field(default_factory=lambda: np.zeros((0, MAX_TORSION_TERMS)))
    pitors: np.ndarray = field(default_factory=lambda: np.zeros((0, 6), np.int32))
    pitors_k: np.ndarray = field(default_factory=lambda: np.zeros(0))
    tortor: np.ndarray = field(default_factory=lambda: np.zeros((0, 6), np.int32))  # 5 atoms, chirality atom or -1
    tortor_grid: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    grids: list = field(default_factory=list)

    def flat(self):
        """The terms as contiguous arrays in the layouts of the C interfaces:
        int32 atom indices, float64 values, the grids concatenated (tx then ty
        of each grid in `grid_t`, values in `grid_e` / `grid_dx` / `grid_dy` /
        `grid_dxy`, with per-grid offsets)."""
        i32 = lambda a: np.ascontiguousarray(a, np.int32).reshape(-1)
        f64 = lambda a: np.ascontiguousarray(a, np.float64).reshape(-1)
        g = self.grids
        nx = np.array([x.nx for x in g], np.int32)
        ny = np.array([x.ny for x in g], np.int32)
        off_t = np.concatenate([[0], np.cumsum(nx + ny)[:-1]]).astype(np.int32) if g else np.zeros(0, np.int32)
        off_e = np.concatenate([[0], np.cumsum(nx * ny)[:-1]]).astype(np.int32) if g else np.zeros(0, np.int32)
        cat = lambda parts: f64(np.concatenate(parts)) if parts else np.zeros(0)
        return dict(
            bond=i32(self.bond), bond_k=f64(self.bond_k), bond_r0=f64(self.bond_r0),
            angle=i32(self.angle), angle_k=f64(self.angle_k), angle_t0=f64(self.angle_t0),
            strbnd=i32(self.strbnd), strbnd_k=f64(self.strbnd_k), strbnd_r0=f64(self.strbnd_r0),
            strbnd_t0=f64(self.strbnd_t0),
            urey=i32(self.urey), urey_k=f64(self.urey_k), urey_r0=f64(self.urey_r0),
            opbend=i32(self.opbend), opbend_k=f64(self.opbend_k),
            torsion=i32(self.torsion), torsion_amp=f64(self.torsion_amp), torsion_cos=f64(self.torsion_cos),
            torsion_sin=f64(self.torsion_sin),
            pitors=i32(self.pitors), pitors_k=f64(self.pitors_k),
            tortor=i32(self.tortor), tortor_grid=i32(self.tortor_grid),
            grid_nx=nx, grid_ny=ny, grid_off_t=off_t, grid_off_e=off_e,
            grid_t=cat([np.concatenate([x.tx, x.ty]) for x in g]), grid_e=cat([x.e for x in g]),
            grid_dx=cat([x.dx for x in g]), grid_dy=cat([x.dy for x in g]), grid_dxy=cat([x.dxy for x in g]))

    def counts(self):
        return {"bond": len(self.bond), "angle": len(self.angle), "strbnd": len(self.strbnd),
                "urey": len(self.urey), "opbend": len(self.opbend), "torsion": len(self.torsion),
                "pitors": len(self.pitors), "tortor": len(self.tortor)}


# ---------------------------------------------------------------------------
# splines of the torsion-torsion grids
# ---------------------------------------------------------------------------
def periodic_spline_slopes(x, y):
    """First derivatives at the knots of the periodic cubic spline through
    (x[k], y[k]), k = 0..n, with y[0] and y[n] replaced by their mean (the
    curve TorsionTorsionType.cspline builds, :822-871).  Solves for the knot
    second derivatives M (h[i-1] M[i-1] + 2 (h[i-1] + h[i]) M[i] + h[i] M[i+1]
    = 6 (s[i] - s[i-1]), indices mod n) by a dense solve."""
    x = np.asarray(x, float)
    y = np.array(y, float)
    n = len(x) - 1
    m = 0.5 * (y[0] + y[n])
    y[0] = y[n] = m
    h = np.diff(x)
    s = np.diff(y) / h
    A = np.zeros((n, n))
    r = np.zeros(n)
    for i in range(n):
        hp, hn = h[i - 1], h[i]           # h[-1] = h[n-1]: the interval before knot 0
        A[i, (i - 1) % n] += hp
        A[i, i] += 2.0 * (hp + hn)
        A[i, (i + 1) % n] += hn
        r[i] = 6.0 * (s[i] - s[i - 1])
    M = np.linalg.solve(A, r)
    M = np.append(M, M[0])
    b = s - h * (2.0 * M[:-1] + M[1:]) / 6.0
    return np.append(b, b[0])


def natural_spline_slopes(x, y):
    """First derivatives at the knots of the natural cubic spline (zero second
    derivatives at both ends; TorsionTorsionType.nspline with y21 = y2n = 0)."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    n = len(x) - 1
    h = np.diff(x)
    s = np.diff(y) / h
    A = np.zeros((n + 1, n + 1))
    r = np.zeros(n + 1)
    A[0, 0] = A[n, n] = 1.0
    for i in range(1, n):
        A[i, i - 1], A[i, i], A[i, i + 1] = h[i - 1], 2.0 * (h[i - 1] + h[i]), h[i]
        r[i] = 6.0 * (s[i] - s[i - 1])
    M = np.linalg.solve(A, r)
    b = s - h * (2.0 * M[:-1] + M[1:]) / 6.0
    return np.append(b, s[-1] + h[-1] * (M[-2] + 2.0 * M[-1]) / 6.0)


def tortor_grid(nx, ny, points):
    """Grid and spline derivatives of one `tortors` record (points in file
    order, the first angle running fastest; TorsionTorsionType.java:107-284)."""
    t1 = np.array([p[0] for p in points])
    t2 = np.array([p[1] for p in points])
    e = np.array([p[2] for p in points])
    tx = np.unique(t1)[:nx]
    ty = np.unique(t2)[:ny]
    cyclic = not (abs(tx[0] - tx[-1]) - 360.0 > 1e-4)
    E = e.reshape(ny, nx)
    if cyclic and (np.abs(E[:, 0] - E[:, -1]).max() > 1e-4 or np.abs(E[0, :] - E[-1, :]).max() > 1e-4):
        cyclic = False
    fit = periodic_spline_slopes if cyclic else natural_spline_slopes
    dx = np.array([fit(tx, E[j, :]) for j in range(ny)])            # (ny, nx)
    dy = np.array([fit(ty, E[:, i]) for i in range(nx)]).T          # (ny, nx)
    dxy = np.array([fit(ty, dx[:, i]) for i in range(nx)]).T
    return TortorGrid(nx, ny, tx, ty, e.copy(), dx.reshape(-1), dy.reshape(-1), dxy.reshape(-1))


# ---------------------------------------------------------------------------
# assignment
# ---------------------------------------------------------------------------
def _fkey(ff, keys, name, default):
    v = keys.get(name) or ff.keys.get(name)
    return float(v[0]) if v else default


def _sorted_ends(c):
    return (c[0], c[1], c[2]) if c[0] <= c[2] else (c[2], c[1], c[0])


def _torsion_key(c):
    """Torsion lookup key (TorsionType.sortKey)."""
    c0, c1, c2, c3 = c
    if c1 < c2:
        return (c0, c1, c2, c3)
    if c2 < c1:
        return (c3, c2, c1, c0)
    return (c0, c1, c2, c3) if c0 <= c3 else (c3, c2, c1, c0)


def _table(records, keyfn):
    """Record lookup by the reference's key; a later record of the same key
    replaces an earlier one (ForceField.addForceFieldType, ForceField.java:253-276)."""
    out = {}
    for cls, vals in records:
        out[keyfn(cls)] = (cls, vals)
    return out


def assign(types, bonds, ff, keys=None):
    """ValenceTerms of atoms with the given types and 0-based bonds (i < j)."""
    keys = keys or {}
    n = len(types)
    types = [int(t) for t in types]
    cls = [ff.atom[t][0] for t in types]
    z = [ff.element[t][1] if t in ff.element else 0 for t in types]
    adj = [[] for _ in range(n)]
    for i, j in bonds:
        adj[i].append(j)
        adj[j].append(i)
    adj = [sorted(a) for a in adj]
    V = ff.valence
    bond_t = _table(V.get("bond", []), lambda c: tuple(sorted(c)))
    angle_t = _table(V.get("angle", []), _sorted_ends)
    anglep_t = _table(V.get("anglep", []), _sorted_ends)
    strbnd_t = _table(V.get("strbnd", []), _sorted_ends)
    urey_t = _table(V.get("ureybrad", []), _sorted_ends)
    opbend_t = _table(V.get("opbend", []), tuple)
    torsion_t = _table(V.get("torsion", []), _torsion_key)
    pitors_t = _table(V.get("pitors", []), lambda c: tuple(sorted(c)))
    tortor_t = {}
    for k, (c5, nx, ny, pts) in enumerate(ff.tortors):
        tortor_t[tuple(c5)] = k

    deg = np.pi / 180.0
    T = ValenceTerms(n=n)
    T.bond_poly = (_fkey(ff, keys, "bond-cubic", 0.0), _fkey(ff, keys, "bond-quartic", 0.0))
    T.angle_poly = tuple(_fkey(ff, keys, f"angle-{p}", 0.0) for p in ("cubic", "quartic", "pentic", "sextic"))
    T.opbend_poly = tuple(_fkey(ff, keys, f"opbend-{p}", 0.0) for p in ("cubic", "quartic", "pentic", "sextic"))
    T.urey_poly = (_fkey(ff, keys, "urey-cubic", 0.0), _fkey(ff, keys, "urey-quartic", 0.0))
    bond_unit = _fkey(ff, keys, "bondunit", 1.0)
    angle_unit = _fkey(ff, keys, "angleunit", deg * deg)
    strbnd_unit = _fkey(ff, keys, "strbndunit", deg)
    urey_unit = _fkey(ff, keys, "ureyunit", 1.0)
    opbend_unit = _fkey(ff, keys, "opbendunit", deg * deg)
    torsion_unit = _fkey(ff, keys, "torsionunit", 1.0)
    pitors_unit = _fkey(ff, keys, "pitorsunit", 1.0)
    tortor_unit = _fkey(ff, keys, "tortorunit", 1.0)
    torsion_scale = _fkey(ff, keys, "torsion-scale", None) if ("torsion-scale" in keys or
                                                              "torsion-scale" in ff.keys) else None

    def missing(kind, atoms):
        raise ValueError(f"no {kind} parameters for atoms {[a + 1 for a in atoms]} "
                         f"(classes {[cls[a] for a in atoms]})")

    # bonds
    r0 = {}
    bl, bk, br = [], [], []
    for i, j in bonds:
        rec = bond_t.get(tuple(sorted((cls[i], cls[j]))))
        if rec is None:
            missing("bond", (i, j))
        bl.append((i, j))
        bk.append(bond_unit * rec[1][0])
        br.append(rec[1][1])
        r0[(i, j)] = r0[(j, i)] = rec[1][1]
    T.bond, T.bond_k, T.bond_r0 = np.array(bl, np.int32).reshape(-1, 2), np.array(bk), np.array(br)

    # angles (centre by centre, bond pairs in neighbour order)
    angles = []          # (a, b, c, key, in_plane, theta0)
    for b in range(n):
        nb = adj[b]
        for x in range(len(nb)):
            for y in range(x + 1, len(nb)):
                a, c = nb[x], nb[y]
                key = _sorted_ends((cls[a], cls[b], cls[c]))
                rec, ip = angle_t.get(key), False
                if rec is None:
                    rec, ip = anglep_t.get(key), True
                if rec is None:
                    missing("angle", (a, b, c))
                vals = rec[1]
                nh = sum(1 for d in nb if d != a and d != c and z[d] == 1)
                while len(vals) - 1 <= nh:
                    nh -= 1
                angles.append((a, b, c, key, ip, vals[1 + nh], vals[0]))
    fourth = {}
    ol, ok = [], []
    for a, b, c, key, ip, t0, k in angles:
        if len(adj[b]) != 3:
            continue
        d = [x for x in adj[b] if x != a and x != c][0]
        rec = (opbend_t.get((cls[d], cls[b], cls[a], cls[c])) or opbend_t.get((cls[d], cls[b], cls[c], cls[a]))
               or opbend_t.get((cls[d], cls[b], 0, 0)))
        if rec is None:
            continue
        fourth[(a, b, c)] = d
        ol.append((a, b, c, d))
        ok.append(opbend_unit * rec[1][0])
    al, ak, at = [], [], []
    for a, b, c, key, ip, t0, k in angles:
        d = -1
        if ip:
            if (a, b, c) not in fourth:
                raise ValueError(f"in-plane angle {a + 1}-{b + 1}-{c + 1} has no out-of-plane atom")
            d = fourth[(a, b, c)]
        al.append((a, b, c, d))
        ak.append(angle_unit * k)
        at.append(t0)
    T.angle, T.angle_k, T.angle_t0 = np.array(al, np.int32).reshape(-1, 4), np.array(ak), np.array(at)
    T.opbend, T.opbend_k = np.array(ol, np.int32).reshape(-1, 4), np.array(ok)

    # stretch-bends and Urey-Bradleys
    sl, sk, sr, st, ul, uk, ur = [], [], [], [], [], [], []
    for a, b, c, key, ip, t0, k in angles:
        rec = strbnd_t.get(key)
        if rec is not None:
            c0 = min(rec[0][0], rec[0][2])
            f = rec[1][:2] if rec[0][0] <= rec[0][2] else rec[1][1::-1]
            f0, f1 = (f[0], f[1]) if cls[a] == c0 else (f[1], f[0])
            sl.append((a, b, c))
            sk.append((strbnd_unit * f0, strbnd_unit * f1))
            sr.append((r0[(a, b)], r0[(b, c)]))
            st.append(t0)
        rec = urey_t.get(key)
        if rec is not None:
            ul.append((a, c))
            uk.append(urey_unit * rec[1][0])
            ur.append(rec[1][1])
    T.strbnd, T.strbnd_k = np.array(sl, np.int32).reshape(-1, 3), np.array(sk).reshape(-1, 2)
    T.strbnd_r0, T.strbnd_t0 = np.array(sr).reshape(-1, 2), np.array(st)
    T.urey, T.urey_k, T.urey_r0 = np.array(ul, np.int32).reshape(-1, 2), np.array(uk), np.array(ur)

    # torsions around every bond
    tl, tamp, tcos, tsin = [], [], [], []
    for b, c in bonds:
        for a in adj[b]:
            if a == c:
                continue
            for d in adj[c]:
                if d == b:
                    continue
                c4 = (cls[a], cls[b], cls[c], cls[d])
                rec = torsion_t.get(_torsion_key(c4))
                if rec is None:
                    first, second = ((c4[0], c4[1], c4[2], 0), (0, c4[1], c4[2], c4[3])) if c4[0] > c4[3] else \
                        ((0, c4[1], c4[2], c4[3]), (c4[0], c4[1], c4[2], 0))
                    rec = torsion_t.get(_torsion_key(first)) or torsion_t.get(_torsion_key(second))
                if rec is None:
                    rec = torsion_t.get(_torsion_key((0, c4[1], c4[2], 0)))
                if rec is None:
                    missing("torsion", (a, b, c, d))
                vals = rec[1]
                amp, ph = np.zeros(MAX_TORSION_TERMS), np.zeros(MAX_TORSION_TERMS)
                for q in range(len(vals) // 3):
                    per = int(round(vals[3 * q + 2]))
                    if not 1 <= per <= MAX_TORSION_TERMS:
                        raise ValueError(f"torsion periodicity {per} out of range")
                    amp[per - 1], ph[per - 1] = vals[3 * q], vals[3 * q + 1]
                scale = 1.0
                if torsion_scale is not None and not any(z[x] == 1 or z[x] in (9, 17, 35, 53) for x in (a, d)):
                    scale = torsion_scale
                tl.append((a, b, c, d))
                tamp.append(scale * torsion_unit * amp)
                tcos.append(np.cos(np.radians(ph)))
                tsin.append(np.sin(np.radians(ph)))
    T.torsion = np.array(tl, np.int32).reshape(-1, 4)
    T.torsion_amp = np.array(tamp).reshape(-1, MAX_TORSION_TERMS)
    T.torsion_cos = np.array(tcos).reshape(-1, MAX_TORSION_TERMS)
    T.torsion_sin = np.array(tsin).reshape(-1, MAX_TORSION_TERMS)

    # pi-orbital torsions
    pl, pk = [], []
    for b, c in bonds:
        if len(adj[b]) != 3 or len(adj[c]) != 3:
            continue
        rec = pitors_t.get(tuple(sorted((cls[b], cls[c]))))
        if rec is None:
            continue
        ob = [x for x in adj[b] if x != c]
        oc = [x for x in adj[c] if x != b]
        pl.append((ob[0], ob[1], b, c, oc[0], oc[1]))
        pk.append(pitors_unit * rec[1][0])
    T.pitors, T.pitors_k = np.array(pl, np.int32).reshape(-1, 6), np.array(pk)

    # torsion-torsions around every angle
    ql, qg, used = [], [], {}
    for a1, a2, a3, key, ip, t0, k in angles:
        for a0 in adj[a1]:
            if a0 == a2 or a0 == a3:
                continue
            for a4 in adj[a3]:
                if a4 in (a0, a1, a2):
                    continue
                c5 = (cls[a0], cls[a1], cls[a2], cls[a3], cls[a4])
                g = tortor_t.get(c5)
                chain = (a0, a1, a2, a3, a4)
                if g is None:
                    g = tortor_t.get(c5[::-1])
                    chain = chain[::-1]
                if g is None:
                    continue
                if g not in used:
                    used[g] = len(used)
                centre = chain[2]
                chiral = -1
                if len(adj[centre]) == 4:
                    o = [x for x in adj[centre] if x != chain[1] and x != chain[3]]
                    p, q = o[0], o[1]
                    if types[p] > types[q]:
                        chiral = p
                    if types[q] > types[p]:
                        chiral = q
                    if z[p] > z[q]:
                        chiral = p
                    if z[q] > z[p]:
                        chiral = q
                ql.append(chain + (chiral,))
                qg.append(used[g])
    T.tortor = np.array(ql, np.int32).reshape(-1, 6)
    T.tortor_grid = np.array(qg, np.int32)
    order = sorted(used, key=used.get)
    T.grids = []
    for g in order:
        c5, nx, ny, pts = ff.tortors[g]
        gr = tortor_grid(nx, ny, pts)
        gr.e, gr.dx, gr.dy, gr.dxy = (tortor_unit * gr.e, tortor_unit * gr.dx, tortor_unit * gr.dy,
                                      tortor_unit * gr.dxy)
        T.grids.append(gr)
    return T
