"""ctypes binding of the FP64 CPU oracle (oracle/amoeba_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker.  The product path never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")


def build(force=False):
    """Compile the oracle with gcc (-O2, OpenMP)."""
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    srcs = [os.path.join(HERE, f) for f in ("amoeba_oracle.c", "valence_oracle.c", "amoeba_oracle.h",
                                            "valence_oracle.h")]
    if not force and os.path.exists(LIB_PATH) and all(os.path.getmtime(LIB_PATH) >= os.path.getmtime(f)
                                                      for f in srcs):
        return LIB_PATH
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=gnu11", "-o", LIB_PATH] + \
        [f for f in srcs if f.endswith(".c")] + ["-lm"]
    subprocess.check_call(cmd)
    return LIB_PATH


class OraSystem(C.Structure):
    _fields_ = [("n", C.c_int), ("xyz", C.POINTER(C.c_double)), ("frame", C.POINTER(C.c_int)),
                ("axis", C.POINTER(C.c_int)), ("mpole", C.POINTER(C.c_double)),
                ("polarity", C.POINTER(C.c_double)), ("thole", C.POINTER(C.c_double)),
                ("pdamp", C.POINTER(C.c_double)), ("group", C.POINTER(C.c_int)),
                ("molecule", C.POINTER(C.c_int)), ("vdw_class", C.POINTER(C.c_int)), ("n_class", C.c_int),
                ("vdw_table", C.POINTER(C.c_double)), ("red_index", C.POINTER(C.c_int)),
                ("red_value", C.POINTER(C.c_double)), ("mask_off", C.POINTER(C.c_int)),
                ("mask_atom", C.POINTER(C.c_int)), ("mask_rel", C.POINTER(C.c_int)),
                ("box", C.POINTER(C.c_double)), ("soft", C.POINTER(C.c_int))]


class OraParams(C.Structure):
    _fields_ = [("ewald_cutoff", C.c_double), ("ewald_alpha", C.c_double), ("vdw_cutoff", C.c_double),
                ("vdw_taper", C.c_double), ("list_buffer", C.c_double), ("pme_order", C.c_int),
                ("grid", C.c_int * 3), ("polar_eps", C.c_double), ("electric", C.c_double),
                ("precond_cutoff", C.c_double), ("precond_scale", C.c_double), ("m_scale", C.c_double * 4),
                ("p_scale", C.c_double * 4), ("intra14", C.c_double), ("d11", C.c_double),
                ("v_scale", C.c_double * 3), ("halgren_delta", C.c_double), ("halgren_gamma", C.c_double),
                ("polarization", C.c_int), ("do_vdw", C.c_int), ("do_pme", C.c_int), ("nthreads", C.c_int),
                ("max_scf_iter", C.c_int), ("lambda_", C.c_double), ("vdw_lambda_alpha", C.c_double),
                ("vdw_lambda_exponent", C.c_double), ("vdw_lambda_end", C.c_double),
                ("intermolecular_softcore", C.c_int), ("intramolecular_softcore", C.c_int),
                ("elec_lambda", C.c_int), ("perm_lambda_alpha", C.c_double), ("perm_lambda_exponent", C.c_double),
                ("perm_lambda_start", C.c_double), ("perm_lambda_end", C.c_double),
                ("pol_lambda_exponent", C.c_double), ("pol_lambda_start", C.c_double),
                ("pol_lambda_end", C.c_double), ("ligand_vapor_elec", C.c_int), ("no_ligand_condensed_scf", C.c_int),
                ("scf_algorithm", C.c_int), ("polar_sor", C.c_double), ("sor_fallback", C.c_int),
                ("finite_pcg", C.c_int), ("scf_guess", C.POINTER(C.c_double)),
                ("direct_out", C.POINTER(C.c_double))]


class OraAlchemy(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("lpow_perm", "dlpow_perm", "d2lpow_perm", "lalpha", "dlalpha", "d2lalpha",
                                          "lpow_pol", "dlpow_pol", "d2lpow_pol")]


class OraEnergies(C.Structure):
    _fields_ = [("perm_real", C.c_double), ("perm_self", C.c_double), ("perm_recip", C.c_double),
                ("perm_charge_corr", C.c_double), ("pol_real", C.c_double), ("pol_self", C.c_double),
                ("pol_recip", C.c_double), ("vdw", C.c_double), ("perm", C.c_double), ("pol", C.c_double),
                ("total", C.c_double), ("pcg_iters", C.c_int), ("pcg_eps", C.c_double), ("n_vdw", C.c_int),
                ("n_perm", C.c_int), ("status", C.c_int), ("vir", C.c_double * 9),
                ("dudl", C.c_double), ("d2udl2", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["vir"] = np.array(list(self.vir)).reshape(3, 3)
        return d


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.ora_evaluate.argtypes = [C.POINTER(OraSystem), C.POINTER(OraParams), C.POINTER(OraEnergies), dp, dp, dp]
        L.ora_evaluate.restype = C.c_int
        L.ora_evaluate_lambda.argtypes = [C.POINTER(OraSystem), C.POINTER(OraParams), C.POINTER(OraEnergies), dp, dp,
                                          dp, dp]
        L.ora_evaluate_lambda.restype = C.c_int
        L.ora_rotate_multipoles.argtypes = [C.POINTER(OraSystem), dp]
        L.ora_bspline.argtypes = [C.c_double, C.c_int, C.c_int, dp]
        L.ora_perm_pair.argtypes = [dp, dp, dp, C.c_double, C.c_double, dp, dp, dp]
        L.ora_perm_pair.restype = C.c_double
        L.ora_pol_pair.argtypes = [dp, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.ora_pol_pair.restype = C.c_double
        L.ora_convolution.argtypes = [dp, dp, C.c_int, C.c_int, C.c_int]
        L.ora_fft3d.argtypes = [dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ora_ewald_recip_exact.argtypes = [C.POINTER(OraSystem), C.POINTER(OraParams), C.c_int]
        L.ora_ewald_recip_exact.restype = C.c_double
        L.ora_influence.argtypes = [dp, C.POINTER(C.c_int), C.c_int, C.c_double, dp]
        L.ora_reduced_xyz.argtypes = [C.POINTER(OraSystem), dp]
        L.ora_alchemical_factors.argtypes = [C.POINTER(OraParams), C.c_double, C.POINTER(OraAlchemy)]
        L.ora_valence_evaluate.argtypes = [C.POINTER(OraValence), dp, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class System:
    """Holds numpy arrays alive and exposes the C struct."""

    def __init__(self, xyz, box, top):
        self.xyz = np.array(xyz, dtype=np.float64).reshape(-1)
        self.box = np.ascontiguousarray(box if box is not None else [0, 0, 0, 90, 90, 90], np.float64)
        self.top = top
        t = top
        self._keep = dict(
            frame=np.ascontiguousarray(t.frame, np.int32), axis=np.ascontiguousarray(t.axis, np.int32).reshape(-1),
            mpole=np.ascontiguousarray(t.mpole, np.float64).reshape(-1),
            polarity=np.ascontiguousarray(t.polarity, np.float64), thole=np.ascontiguousarray(t.thole, np.float64),
            pdamp=np.ascontiguousarray(t.pdamp, np.float64), group=np.ascontiguousarray(t.group, np.int32),
            molecule=np.ascontiguousarray(t.molecule, np.int32),
            vdw_class=np.ascontiguousarray(t.vdw_class, np.int32),
            vdw_table=np.ascontiguousarray(t.vdw_table, np.float64).reshape(-1),
            red_index=np.ascontiguousarray(t.red_index, np.int32),
            red_value=np.ascontiguousarray(t.red_value, np.float64),
            mask_off=np.ascontiguousarray(t.mask_off, np.int32), mask_atom=np.ascontiguousarray(t.mask_atom, np.int32),
            mask_rel=np.ascontiguousarray(t.mask_rel, np.int32),
            soft=np.ascontiguousarray(t.soft if getattr(t, "soft", None) is not None else np.zeros(t.n), np.int32))
        k = self._keep
        self.c = OraSystem(t.n, _dp(self.xyz), _ip(k["frame"]), _ip(k["axis"]), _dp(k["mpole"]), _dp(k["polarity"]),
                           _dp(k["thole"]), _dp(k["pdamp"]), _ip(k["group"]), _ip(k["molecule"]),
                           _ip(k["vdw_class"]), int(t.vdw_table.shape[0]), _dp(k["vdw_table"]),
                           _ip(k["red_index"]), _dp(k["red_value"]), _ip(k["mask_off"]), _ip(k["mask_atom"]),
                           _ip(k["mask_rel"]), _dp(self.box), _ip(k["soft"]))

    def set_xyz(self, xyz):
        self.xyz[:] = np.asarray(xyz, np.float64).reshape(-1)


def make_params(p, box=None, nthreads=0, finite_pcg=False):
    """OraParams from a forcefieldx_amd.forcefield.NonbondedParams; finite_pcg
    asks for the gradient of the truncated PCG's dipoles (adjoint terms)."""
    grid = p.resolved_grid(box) if box is not None and box[0] > 0 else (16, 16, 16)
    return OraParams(p.ewald_cutoff, p.ewald_alpha, p.vdw_cutoff, p.vdw_taper, p.list_buffer, p.pme_order,
                     (C.c_int * 3)(*grid), p.polar_eps, p.electric, p.precond_cutoff, p.precond_scale,
                     (C.c_double * 4)(*p.m_scale), (C.c_double * 4)(*p.p_scale), p.intra14, p.d11,
                     (C.c_double * 3)(*p.v_scale), p.halgren_delta, p.halgren_gamma, p.polarization,
                     int(p.do_vdw), int(p.do_pme), nthreads, p.max_scf_iter, p.lambda_value, p.vdw_lambda_alpha,
                     p.vdw_lambda_exponent, p.vdw_lambda_end, int(p.intermolecular_softcore),
                     int(p.intramolecular_softcore), int(p.elec_lambda), p.perm_lambda_alpha,
                     p.perm_lambda_exponent, p.perm_lambda_start, p.perm_lambda_end, p.pol_lambda_exponent,
                     p.pol_lambda_start, p.pol_lambda_end, int(p.ligand_vapor_elec),
                     int(p.no_ligand_condensed_scf), int(p.scf_algorithm), p.polar_sor, int(p.sor_fallback),
                     int(finite_pcg))


def alchemical_factors(params, lam):
    """lPowPerm, lAlpha, lPowPol and their lambda derivatives (dict)."""
    f = OraAlchemy()
    lib().ora_alchemical_factors(C.byref(make_params(params)), float(lam), C.byref(f))
    return {k: getattr(f, k) for k, _ in OraAlchemy._fields_}


def evaluate(system, params, nthreads=0, finite_pcg=False, scf_guess=None, want_direct=False):
    """Run the oracle; returns (energies dict, gradient (n,3), u (n,3), uCR (n,3)).
    With finite_pcg the gradient is that of the energy of the dipoles the PCG
    returned (iteration count fixed) and uCR is replaced by the adjoint dipoles.
    scf_guess (2, n, 3): added to the direct dipoles as the solve's starting
    point; want_direct appends the direct dipoles (2, n, 3) to the result."""
    n = system.top.n
    prm = make_params(params, system.box, nthreads, finite_pcg) if not isinstance(params, OraParams) else params
    keep = []
    if scf_guess is not None:
        g6 = np.ascontiguousarray(scf_guess, np.float64).reshape(-1)
        keep.append(g6)
        prm.scf_guess = _dp(g6)
    d6 = np.zeros(6 * n) if want_direct else None
    if want_direct:
        prm.direct_out = _dp(d6)
    out = OraEnergies()
    g = np.zeros(3 * n)
    u = np.zeros(3 * n)
    up = np.zeros(3 * n)
    lib().ora_evaluate(C.byref(system.c), C.byref(prm), C.byref(out), _dp(g), _dp(u), _dp(up))
    res = (out.as_dict(), g.reshape(n, 3), u.reshape(n, 3), up.reshape(n, 3))
    return res + (d6.reshape(2, n, 3),) if want_direct else res


def evaluate_lambda(system, params, nthreads=0):
    """As evaluate, plus dU/dX/dlambda (n,3); energies carry dudl, d2udl2."""
    n = system.top.n
    prm = make_params(params, system.box, nthreads) if not isinstance(params, OraParams) else params
    out = OraEnergies()
    g, u, up, lg = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
    lib().ora_evaluate_lambda(C.byref(system.c), C.byref(prm), C.byref(out), _dp(g), _dp(u), _dp(up), _dp(lg))
    return out.as_dict(), g.reshape(n, 3), lg.reshape(n, 3)


def rotate(system):
    g = np.zeros(10 * system.top.n)
    lib().ora_rotate_multipoles(C.byref(system.c), _dp(g))
    return g.reshape(-1, 10)


def bspline(x, order=5, nd=3):
    th = np.zeros(order * (nd + 1))
    lib().ora_bspline(float(x), order, nd, _dp(th))
    return th.reshape(order, nd + 1)


def perm_pair(r, mi, mk, aewald, mscale=1.0):
    r = np.ascontiguousarray(r, float)
    mi = np.ascontiguousarray(mi, float)
    mk = np.ascontiguousarray(mk, float)
    gi, ti, tk = np.zeros(3), np.zeros(3), np.zeros(3)
    e = lib().ora_perm_pair(_dp(r), _dp(mi), _dp(mk), aewald, mscale, _dp(gi), _dp(ti), _dp(tk))
    return e, gi, ti, tk


def pol_pair(r, mi, mk, ui, uip, uk, ukp, aewald, pdi=0.0, pdk=0.0, tholei=0.0, tholek=0.0, dscale=1.0,
             pscale=1.0):
    arrs = [np.ascontiguousarray(a, float) for a in (r, mi, mk, ui, uip, uk, ukp)]
    gi, ti, tk = np.zeros(3), np.zeros(3), np.zeros(3)
    e = lib().ora_pol_pair(*[_dp(a) for a in arrs], aewald, pdi, pdk, tholei, tholek, dscale, pscale, _dp(gi),
                           _dp(ti), _dp(tk))
    return e, gi, ti, tk


def convolution(data, recip, nx, ny, nz):
    d = np.ascontiguousarray(data, np.float64).copy()
    r = np.ascontiguousarray(recip, np.float64)
    lib().ora_convolution(_dp(d), _dp(r), nx, ny, nz)
    return d


def fft3d(data, nx, ny, nz, inverse=False):
    d = np.ascontiguousarray(data, np.float64).copy()
    lib().ora_fft3d(_dp(d), nx, ny, nz, int(inverse))
    return d


def ewald_recip_exact(system, params, kmax):
    prm = make_params(params, system.box)
    return lib().ora_ewald_recip_exact(C.byref(system.c), C.byref(prm), kmax)


def influence(box, grid, order, aewald):
    g = (C.c_int * 3)(*grid)
    out = np.zeros(grid[0] * grid[1] * grid[2])
    lib().ora_influence(_dp(np.ascontiguousarray(box, float)), g, order, aewald, _dp(out))
    return out


# ---------------------------------------------------------------------------
# valence terms (valence_oracle.c)
# ---------------------------------------------------------------------------
_PI, _PD = C.POINTER(C.c_int), C.POINTER(C.c_double)


class OraValence(C.Structure):
    _fields_ = [("n_bond", C.c_int), ("bond", _PI), ("bond_k", _PD), ("bond_r0", _PD),
                ("bond_cubic", C.c_double), ("bond_quartic", C.c_double),
                ("n_angle", C.c_int), ("angle", _PI), ("angle_k", _PD), ("angle_t0", _PD),
                ("angle_poly", C.c_double * 4),
                ("n_strbnd", C.c_int), ("strbnd", _PI), ("strbnd_k", _PD), ("strbnd_r0", _PD), ("strbnd_t0", _PD),
                ("n_urey", C.c_int), ("urey", _PI), ("urey_k", _PD), ("urey_r0", _PD),
                ("urey_cubic", C.c_double), ("urey_quartic", C.c_double),
                ("n_opbend", C.c_int), ("opbend", _PI), ("opbend_k", _PD), ("opbend_poly", C.c_double * 4),
                ("n_torsion", C.c_int), ("torsion", _PI), ("torsion_amp", _PD), ("torsion_cos", _PD),
                ("torsion_sin", _PD),
                ("n_pitors", C.c_int), ("pitors", _PI), ("pitors_k", _PD),
                ("n_tortor", C.c_int), ("tortor", _PI), ("tortor_grid", _PI),
                ("n_grid", C.c_int), ("grid_nx", _PI), ("grid_ny", _PI), ("grid_off_t", _PI), ("grid_off_e", _PI),
                ("grid_t", _PD), ("grid_e", _PD), ("grid_dx", _PD), ("grid_dy", _PD), ("grid_dxy", _PD)]


VALENCE_KINDS = ("bond", "angle", "strbnd", "urey", "opbend", "torsion", "pitors", "tortor")


def valence(terms, xyz, virial=False):
    """Energies of the eight valence kinds (dict), gradient (n, 3) and, when
    asked, the strain derivative sum_i g_i x_i^T (3, 3)."""
    f = terms.flat()
    c = terms.counts()
    s = OraValence()
    keep = []
    for name, arr in f.items():
        keep.append(arr)
        setattr(s, name, _ip(arr) if arr.dtype == np.int32 else _dp(arr))
    for k in VALENCE_KINDS:
        setattr(s, "n_" + k, c[k])
    s.n_grid = len(terms.grids)
    s.bond_cubic, s.bond_quartic = terms.bond_poly
    s.urey_cubic, s.urey_quartic = terms.urey_poly
    s.angle_poly = (C.c_double * 4)(*terms.angle_poly)
    s.opbend_poly = (C.c_double * 4)(*terms.opbend_poly)
    x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
    e = np.zeros(8)
    g = np.zeros(terms.n * 3)
    vir = np.zeros(9) if virial else None
    lib().ora_valence_evaluate(C.byref(s), _dp(x), _dp(e), _dp(g), _dp(vir) if virial else None)
    out = dict(zip(VALENCE_KINDS, e))
    out["total"] = float(e.sum())
    return (out, g.reshape(-1, 3), vir.reshape(3, 3)) if virial else (out, g.reshape(-1, 3))
