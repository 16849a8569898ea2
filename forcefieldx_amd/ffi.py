"""ctypes binding of libamoeba_nb.so (include/amoeba_nb.h).

This is the Python counterpart of the FFM downcall stubs a Java host would
declare.  Structures mirror the C header field by field.
The library is the HIP build in forcefieldx_amd/_build; there is no CPU
fallback: loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libamoeba_nb.so")

NB_OK, NB_ERR_NAN_ENERGY, NB_ERR_SCF_DIVERGED, NB_ERR_SCF_MAXITER, NB_ERR_BAD_TORQUE, NB_ERR_HIP, \
    NB_ERR_INVALID, NB_ERR_NO_DEVICE = range(8)
NB_GRAD, NB_VIRIAL, NB_DUDL, NB_FINITE_PCG, NB_BATCH_SEPARATE = 1, 2, 4, 8, 16
NB_PREC_FP64, NB_PREC_MIXED = 0, 1

EXPORTED = ["nb_create", "nb_set_cell", "nb_set_positions", "nb_set_positions_device", "nb_set_lambda",
            "nb_evaluate", "nb_get_gradient", "nb_gradient_device", "nb_get_induced", "nb_get_virial",
            "nb_get_dudx_dl", "nb_last_error", "nb_destroy", "nb_abi_version", "nb_device_count",
            "nb_set_profiling", "nb_set_profile_filter", "nb_profile_count", "nb_profile_entry", "nb_struct_sizes",
            "nb_set_graph", "nb_graph_replays", "nb_evaluate_batch", "nb_batch_config", "nb_set_valence",
            "nb_valence_struct_size", "nb_set_scf_predictor", "nb_md_init", "nb_md_step", "nb_md_get_state",
            "nb_md_set_thermostat", "nb_set_refinement", "nb_evaluate_batch_static", "nb_get_scf_predictor"]
NB_BATCH_SCRATCH_MIB = 64
NB_PRED_NONE, NB_PRED_ASPC, NB_PRED_LS = 0, 1, 2
NB_THERMO_NONE, NB_THERMO_BERENDSEN = 0, 1
VALENCE_KINDS = ("bond", "angle", "strbnd", "urey", "opbend", "torsion", "pitors", "tortor")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class NbParams(C.Structure):
    _fields_ = [("ewald_cutoff", C.c_double), ("ewald_alpha", C.c_double), ("vdw_cutoff", C.c_double),
                ("vdw_taper", C.c_double), ("list_buffer", C.c_double), ("pme_order", C.c_int32),
                ("grid", C.c_int32 * 3), ("polar_eps", C.c_double), ("electric", C.c_double),
                ("precond_cutoff", C.c_double), ("precond_scale", C.c_double), ("m_scale", C.c_double * 4),
                ("p_scale", C.c_double * 4), ("intra14", C.c_double), ("d11", C.c_double),
                ("v_scale", C.c_double * 3), ("halgren_delta", C.c_double), ("halgren_gamma", C.c_double),
                ("polarization", C.c_int32), ("do_vdw", C.c_int32), ("do_pme", C.c_int32),
                ("precision", C.c_int32), ("max_scf_iter", C.c_int32), ("device", C.c_int32),
                ("lambda_", C.c_double), ("vdw_lambda_alpha", C.c_double), ("vdw_lambda_exponent", C.c_double),
                ("vdw_lambda_end", C.c_double), ("intermolecular_softcore", C.c_int32),
                ("intramolecular_softcore", C.c_int32), ("elec_lambda", C.c_int32),
                ("perm_lambda_alpha", C.c_double), ("perm_lambda_exponent", C.c_double),
                ("perm_lambda_start", C.c_double), ("perm_lambda_end", C.c_double),
                ("pol_lambda_exponent", C.c_double), ("pol_lambda_start", C.c_double),
                ("pol_lambda_end", C.c_double), ("ligand_vapor_elec", C.c_int32),
                ("no_ligand_condensed_scf", C.c_int32), ("scf_algorithm", C.c_int32), ("polar_sor", C.c_double),
                ("sor_fallback", C.c_int32)]


class NbTopology(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("frame", _ip), ("axis", _ip), ("multipole", _dp),
                ("polarizability", _dp), ("thole", _dp), ("pdamp", _dp), ("group", _ip), ("vdw_class", _ip),
                ("n_vdw_class", C.c_int32), ("vdw_table", _dp), ("red_index", _ip), ("red_value", _dp),
                ("mask_offset", _ip), ("mask_atom", _ip), ("mask_rel", _ip), ("soft", _ip), ("molecule", _ip)]


class NbEnergies(C.Structure):
    _fields_ = [("perm_real", C.c_double), ("perm_self", C.c_double), ("perm_recip", C.c_double),
                ("perm_charge_corr", C.c_double), ("pol_real", C.c_double), ("pol_self", C.c_double),
                ("pol_recip", C.c_double), ("vdw", C.c_double), ("perm", C.c_double), ("pol", C.c_double),
                ("total", C.c_double), ("dudl", C.c_double), ("d2udl2", C.c_double), ("pcg_iters", C.c_int32),
                ("pcg_eps", C.c_double), ("n_vdw", C.c_int64), ("n_perm", C.c_int64),
                ("list_rebuilt", C.c_int32), ("valence", C.c_double * 8)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "valence"}
        d.update({"valence_" + k: self.valence[i] for i, k in enumerate(VALENCE_KINDS)})
        d["valence"] = sum(self.valence)
        return d


class NbMdReport(C.Structure):
    _fields_ = [("kinetic", C.c_double), ("potential", C.c_double), ("total", C.c_double), ("steps", C.c_int32),
                ("pcg_iters", C.c_int32), ("last", NbEnergies)]


class NbValence(C.Structure):
    _fields_ = [("n_bond", C.c_int32), ("bond", _ip), ("bond_k", _dp), ("bond_r0", _dp),
                ("bond_cubic", C.c_double), ("bond_quartic", C.c_double),
                ("n_angle", C.c_int32), ("angle", _ip), ("angle_k", _dp), ("angle_t0", _dp),
                ("angle_poly", C.c_double * 4),
                ("n_strbnd", C.c_int32), ("strbnd", _ip), ("strbnd_k", _dp), ("strbnd_r0", _dp),
                ("strbnd_t0", _dp),
                ("n_urey", C.c_int32), ("urey", _ip), ("urey_k", _dp), ("urey_r0", _dp),
                ("urey_cubic", C.c_double), ("urey_quartic", C.c_double),
                ("n_opbend", C.c_int32), ("opbend", _ip), ("opbend_k", _dp), ("opbend_poly", C.c_double * 4),
                ("n_torsion", C.c_int32), ("torsion", _ip), ("torsion_amp", _dp), ("torsion_cos", _dp),
                ("torsion_sin", _dp),
                ("n_pitors", C.c_int32), ("pitors", _ip), ("pitors_k", _dp),
                ("n_tortor", C.c_int32), ("tortor", _ip), ("tortor_grid", _ip),
                ("n_grid", C.c_int32), ("grid_nx", _ip), ("grid_ny", _ip), ("grid_off_t", _ip),
                ("grid_off_e", _ip), ("grid_t", _dp), ("grid_e", _dp), ("grid_dx", _dp), ("grid_dy", _dp),
                ("grid_dxy", _dp)]


_lib = None


def load(path=LIB_PATH):
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with forcefieldx_amd.build.build() "
                           "(the nonbonded path has no CPU fallback)")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.nb_create.argtypes = [C.POINTER(NbParams), C.POINTER(NbTopology), C.POINTER(vp)]
    L.nb_set_cell.argtypes = [vp, _dp]
    L.nb_set_positions.argtypes = [vp, _dp]
    L.nb_set_positions_device.argtypes = [vp, vp]
    L.nb_set_lambda.argtypes = [vp, C.c_double]
    L.nb_set_graph.argtypes = [vp, C.c_int]
    L.nb_set_refinement.argtypes = [vp, C.c_int]
    L.nb_set_refinement.restype = C.c_int
    L.nb_evaluate_batch.argtypes = [C.POINTER(vp), C.c_int, C.c_uint32, C.POINTER(NbEnergies)]
    L.nb_evaluate_batch_static.argtypes = [C.POINTER(vp), C.c_int, C.c_uint32, C.POINTER(NbEnergies)]
    L.nb_evaluate_batch_static.restype = C.c_int
    L.nb_batch_config.argtypes = [C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    L.nb_graph_replays.argtypes = [vp]
    L.nb_struct_sizes.argtypes = [_ip]
    L.nb_struct_sizes.restype = C.c_int
    L.nb_evaluate.argtypes = [vp, C.c_uint32, C.POINTER(NbEnergies)]
    L.nb_get_gradient.argtypes = [vp, _dp]
    L.nb_gradient_device.argtypes = [vp]
    L.nb_gradient_device.restype = vp
    L.nb_get_induced.argtypes = [vp, _dp, _dp]
    L.nb_get_virial.argtypes = [vp, _dp]
    L.nb_get_dudx_dl.argtypes = [vp, _dp]
    L.nb_last_error.argtypes = [vp]
    L.nb_last_error.restype = C.c_char_p
    L.nb_destroy.argtypes = [vp]
    L.nb_destroy.restype = None
    for f in ("nb_create", "nb_set_cell", "nb_set_positions", "nb_set_positions_device", "nb_set_lambda",
              "nb_evaluate", "nb_get_gradient", "nb_get_induced", "nb_get_virial", "nb_get_dudx_dl",
              "nb_abi_version", "nb_device_count"):
        getattr(L, f).restype = C.c_int
    L.nb_set_profiling.argtypes = [vp, C.c_int]
    L.nb_set_profiling.restype = C.c_int
    L.nb_set_profile_filter.argtypes = [vp, C.c_char_p]
    L.nb_set_profile_filter.restype = C.c_int
    L.nb_profile_count.argtypes = [vp]
    L.nb_profile_count.restype = C.c_int
    L.nb_profile_entry.argtypes = [vp, C.c_int, C.c_char_p, C.c_int, _dp, C.POINTER(C.c_int64)]
    L.nb_profile_entry.restype = C.c_int
    L.nb_set_scf_predictor.argtypes = [vp, C.c_int]
    L.nb_set_scf_predictor.restype = C.c_int
    L.nb_get_scf_predictor.argtypes = [vp, _dp, _dp, C.POINTER(C.c_int32)]
    L.nb_get_scf_predictor.restype = C.c_int
    L.nb_md_init.argtypes = [vp, _dp, _dp]
    L.nb_md_init.restype = C.c_int
    L.nb_md_step.argtypes = [vp, C.c_int, C.c_double, C.POINTER(NbMdReport)]
    L.nb_md_step.restype = C.c_int
    L.nb_md_get_state.argtypes = [vp, _dp, _dp]
    L.nb_md_get_state.restype = C.c_int
    L.nb_md_set_thermostat.argtypes = [vp, C.c_int, C.c_double, C.c_double]
    L.nb_md_set_thermostat.restype = C.c_int
    L.nb_set_valence.argtypes = [vp, C.POINTER(NbValence)]
    L.nb_set_valence.restype = C.c_int
    L.nb_valence_struct_size.restype = C.c_int
    L.nb_test_convolution.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.nb_test_convolution.restype = C.c_int
    _lib = L
    return L


class NbError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"amoeba_nb error {code}: {msg}")
        self.code = code


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Handle:
    """Owns one nb handle and keeps the topology arrays alive for its lifetime."""

    def __init__(self, params: NbParams, topology, device=0):
        L = load()
        t = topology
        self._arrays = dict(
            frame=np.ascontiguousarray(t.frame, np.int32), axis=np.ascontiguousarray(t.axis, np.int32).reshape(-1),
            mpole=np.ascontiguousarray(t.mpole, np.float64).reshape(-1),
            polarity=np.ascontiguousarray(t.polarity, np.float64), thole=np.ascontiguousarray(t.thole, np.float64),
            pdamp=np.ascontiguousarray(t.pdamp, np.float64), group=np.ascontiguousarray(t.group, np.int32),
            vclass=np.ascontiguousarray(t.vdw_class, np.int32),
            vtab=np.ascontiguousarray(t.vdw_table, np.float64).reshape(-1),
            redi=np.ascontiguousarray(t.red_index, np.int32), redv=np.ascontiguousarray(t.red_value, np.float64),
            moff=np.ascontiguousarray(t.mask_off, np.int32), matom=np.ascontiguousarray(t.mask_atom, np.int32),
            mrel=np.ascontiguousarray(t.mask_rel, np.int32),
            mol=np.ascontiguousarray(t.molecule, np.int32))
        soft = getattr(t, "soft", None)
        has_soft = soft is not None and np.any(np.asarray(soft) != 0)
        if has_soft:
            self._arrays["soft"] = np.ascontiguousarray(soft, np.int32)
        a = self._arrays
        top = NbTopology(t.n, _p(a["frame"], C.c_int32), _p(a["axis"], C.c_int32), _p(a["mpole"], C.c_double),
                         _p(a["polarity"], C.c_double), _p(a["thole"], C.c_double), _p(a["pdamp"], C.c_double),
                         _p(a["group"], C.c_int32), _p(a["vclass"], C.c_int32), int(t.vdw_table.shape[0]),
                         _p(a["vtab"], C.c_double), _p(a["redi"], C.c_int32), _p(a["redv"], C.c_double),
                         _p(a["moff"], C.c_int32), _p(a["matom"], C.c_int32), _p(a["mrel"], C.c_int32),
                         _p(a["soft"], C.c_int32) if has_soft else None, _p(a["mol"], C.c_int32))
        params.device = device
        self.params = params
        self.n = t.n
        h = C.c_void_p()
        rc = L.nb_create(C.byref(params), C.byref(top), C.byref(h))
        if rc != NB_OK:
            raise NbError(rc, "nb_create failed")
        self.h = h
        self.L = L

    def _check(self, rc):
        if rc != NB_OK:
            raise NbError(rc, self.L.nb_last_error(self.h).decode())

    def set_cell(self, box):
        b = np.ascontiguousarray(box if box is not None else [0, 0, 0, 90, 90, 90], np.float64)
        self._check(self.L.nb_set_cell(self.h, _p(b, C.c_double)))

    def set_positions(self, xyz):
        x = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        self._check(self.L.nb_set_positions(self.h, _p(x, C.c_double)))

    def set_positions_device(self, ptr):
        self._check(self.L.nb_set_positions_device(self.h, C.c_void_p(ptr)))

    def evaluate(self, flags=NB_GRAD, check=True):
        e = NbEnergies()
        rc = self.L.nb_evaluate(self.h, flags, C.byref(e))
        if check:
            self._check(rc)
        return e.as_dict(), rc

    def gradient(self):
        g = np.zeros(3 * self.n)
        self._check(self.L.nb_get_gradient(self.h, _p(g, C.c_double)))
        return g.reshape(self.n, 3)

    def set_valence(self, terms):
        """Valence terms (forcefieldx_amd.bonded.ValenceTerms) evaluated with
        every later evaluation; None removes them."""
        if terms is None:
            self._check(self.L.nb_set_valence(self.h, None))
            return
        f = terms.flat()
        c = terms.counts()
        v = NbValence()
        for name, arr in f.items():
            setattr(v, name, _p(arr, C.c_int32) if arr.dtype == np.int32 else _p(arr, C.c_double))
        for k in VALENCE_KINDS:
            setattr(v, "n_" + k, c[k])
        v.n_grid = len(terms.grids)
        v.bond_cubic, v.bond_quartic = terms.bond_poly
        v.urey_cubic, v.urey_quartic = terms.urey_poly
        v.angle_poly = (C.c_double * 4)(*terms.angle_poly)
        v.opbend_poly = (C.c_double * 4)(*terms.opbend_poly)
        self._check(self.L.nb_set_valence(self.h, C.byref(v)))   # the library copies the arrays

    def set_scf_predictor(self, kind=NB_PRED_ASPC):
        self._check(self.L.nb_set_scf_predictor(self.h, int(kind)))

    def scf_predictor(self):
        """(coefficients [6], history [6, 2, n, 3] by ring slot in caller order,
        newest slot, entries held) of the SCF predictor."""
        c = np.zeros(6)
        hist = np.zeros(36 * self.n)
        st = np.zeros(2, np.int32)
        self._check(self.L.nb_get_scf_predictor(self.h, _p(c, C.c_double), _p(hist, C.c_double),
                                                _p(st, C.c_int32)))
        return c, hist.reshape(6, 2, self.n, 3), int(st[0]), int(st[1])

    def md_init(self, mass, v0=None):
        m = np.ascontiguousarray(mass, np.float64).reshape(-1)
        v = None if v0 is None else np.ascontiguousarray(v0, np.float64).reshape(-1)
        self._check(self.L.nb_md_init(self.h, _p(m, C.c_double), None if v is None else _p(v, C.c_double)))

    def md_step(self, nsteps, dt_ps):
        """nsteps velocity-Verlet steps of dt_ps; returns the report as a dict."""
        r = NbMdReport()
        self._check(self.L.nb_md_step(self.h, int(nsteps), float(dt_ps), C.byref(r)))
        return dict(kinetic=r.kinetic, potential=r.potential, total=r.total, steps=r.steps,
                    pcg_iters=r.pcg_iters, last=r.last.as_dict())

    def md_set_thermostat(self, kind, temperature=298.15, tau_ps=0.1):
        """Berendsen.java:99-101 after each full step (NB_THERMO_BERENDSEN),
        or none (NB_THERMO_NONE)."""
        self._check(self.L.nb_md_set_thermostat(self.h, int(kind), float(temperature), float(tau_ps)))

    def md_state(self):
        x = np.zeros(3 * self.n)
        v = np.zeros(3 * self.n)
        self._check(self.L.nb_md_get_state(self.h, _p(x, C.c_double), _p(v, C.c_double)))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def set_lambda(self, lam):
        self._check(self.L.nb_set_lambda(self.h, float(lam)))

    def set_graph(self, on=True):
        """Record evaluations into a HIP graph and replay them."""
        self._check(self.L.nb_set_graph(self.h, int(on)))

    def set_refinement(self, on=True):
        """FP64 iterative refinement of the mixed form's solve (default on)."""
        self._check(self.L.nb_set_refinement(self.h, int(on)))

    def graph_replays(self):
        return self.L.nb_graph_replays(self.h)

    def dudx_dl(self):
        """dU/dX/dlambda (n, 3) of the last evaluation made with NB_DUDL."""
        g = np.zeros(3 * self.n)
        self._check(self.L.nb_get_dudx_dl(self.h, _p(g, C.c_double)))
        return g.reshape(self.n, 3)

    def virial(self):
        """Strain derivative W_ab = dE/d eps_ab (kcal/mol, 3x3) of the last
        evaluation made with NB_VIRIAL."""
        v = np.zeros(9)
        self._check(self.L.nb_get_virial(self.h, _p(v, C.c_double)))
        return v.reshape(3, 3)

    def gradient_device_ptr(self):
        return self.L.nb_gradient_device(self.h)

    def induced(self):
        u = np.zeros(3 * self.n)
        up = np.zeros(3 * self.n)
        self._check(self.L.nb_get_induced(self.h, _p(u, C.c_double), _p(up, C.c_double)))
        return u.reshape(self.n, 3), up.reshape(self.n, 3)

    def set_profiling(self, on=True, kernel_filter=""):
        self._check(self.L.nb_set_profiling(self.h, int(on)))
        self._check(self.L.nb_set_profile_filter(self.h, kernel_filter.encode()))

    def profile(self):
        """{kernel name: (total ms, launches)} accumulated since set_profiling."""
        out = {}
        buf = C.create_string_buffer(256)
        ms = C.c_double()
        cnt = C.c_int64()
        for i in range(self.L.nb_profile_count(self.h)):
            self._check(self.L.nb_profile_entry(self.h, i, buf, 256, C.byref(ms), C.byref(cnt)))
            out[buf.value.decode()] = (ms.value, cnt.value)
        return out

    def close(self):
        if getattr(self, "h", None):
            self.L.nb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def evaluate_batch(handles, flags=NB_GRAD):
    """nb_evaluate_batch over Handle objects; returns (list of energy dicts, rc)."""
    L = load()
    arr = (C.c_void_p * len(handles))(*[h.h for h in handles])
    outs = (NbEnergies * len(handles))()
    rc = L.nb_evaluate_batch(arr, len(handles), flags, outs)
    return [o.as_dict() for o in outs], rc


def evaluate_batch_static(handles, flags=NB_GRAD):
    """nb_evaluate_batch_static (the batch kernel with its buffers in the
    library's static device scratch, no allocation) over Handle objects;
    returns (list of energy dicts, rc)."""
    L = load()
    arr = (C.c_void_p * len(handles))(*[h.h for h in handles])
    outs = (NbEnergies * len(handles))()
    rc = L.nb_evaluate_batch_static(arr, len(handles), flags, outs)
    return [o.as_dict() for o in outs], rc


def batch_config(device=0):
    """Shape of the last batch-kernel launch: threads per workgroup, lanes per
    atom, persistent workgroups, systems, and its duration in ms."""
    L = load()
    cfg = (C.c_int32 * 4)()
    ms = C.c_double(0.0)
    rc = L.nb_batch_config(device, cfg, C.byref(ms))
    if rc != NB_OK:
        raise NbError(rc, "nb_batch_config failed")
    return dict(threads=cfg[0], lanes=cfg[1], workgroups=cfg[2], systems=cfg[3], kernel_ms=ms.value)


def params_from(p, grid, precision=None, device=0):
    """NbParams from forcefieldx_amd.forcefield.NonbondedParams."""
    prec = p.precision if precision is None else precision
    return NbParams(p.ewald_cutoff, p.ewald_alpha, p.vdw_cutoff, p.vdw_taper, p.list_buffer, p.pme_order,
                    (C.c_int32 * 3)(*grid), p.polar_eps, p.electric, p.precond_cutoff, p.precond_scale,
                    (C.c_double * 4)(*p.m_scale), (C.c_double * 4)(*p.p_scale), p.intra14, p.d11,
                    (C.c_double * 3)(*p.v_scale), p.halgren_delta, p.halgren_gamma, p.polarization, int(p.do_vdw),
                    int(p.do_pme), prec, p.max_scf_iter, device, p.lambda_value, p.vdw_lambda_alpha,
                    p.vdw_lambda_exponent, p.vdw_lambda_end, int(p.intermolecular_softcore),
                    int(p.intramolecular_softcore), int(p.elec_lambda), p.perm_lambda_alpha,
                    p.perm_lambda_exponent, p.perm_lambda_start, p.perm_lambda_end, p.pol_lambda_exponent,
                    p.pol_lambda_start, p.pol_lambda_end, int(p.ligand_vapor_elec),
                    int(p.no_ligand_condensed_scf), int(p.scf_algorithm), p.polar_sor, int(p.sor_fallback))


def device_count():
    try:
        return load().nb_device_count()
    except Exception:
        return 0


_hip = None


def _hip_runtime():
    """The HIP runtime the library itself links (loaded once per process, so
    buffers made here live in the same device context as the library's)."""
    global _hip
    if _hip is None:
        load()
        H = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipSetDevice.argtypes = [C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipSetDevice", "hipDeviceSynchronize"):
            getattr(H, f).restype = C.c_int
        _hip = H
    return _hip


def device_synchronize(device=0):
    """Wait for all work on the device (the caller's timing boundary)."""
    H = _hip_runtime()
    H.hipSetDevice(device)
    rc = H.hipDeviceSynchronize()
    if rc != 0:
        raise NbError(NB_ERR_HIP, f"hipDeviceSynchronize returned {rc}")


class DeviceArray:
    """A caller-owned FP64 buffer in device memory, for the *_device entry
    points (coordinates already resident in HBM).  Frames of equal size can be
    addressed by index: ptr(k) is the address of frame k."""

    H2D, D2H = 1, 2

    def __init__(self, count, frames=1, device=0):
        H = _hip_runtime()
        self.count, self.frames, self.device = int(count), int(frames), device
        self._p = C.c_void_p()
        self._chk(H.hipSetDevice(device), "hipSetDevice")
        self._chk(H.hipMalloc(C.byref(self._p), 8 * self.count * self.frames), "hipMalloc")

    @staticmethod
    def _chk(rc, what):
        if rc != 0:
            raise NbError(NB_ERR_HIP, f"{what} returned {rc}")

    def ptr(self, frame=0):
        if not 0 <= frame < self.frames:
            raise IndexError(frame)
        return self._p.value + 8 * self.count * frame

    def upload(self, a, frame=0):
        a = np.ascontiguousarray(a, np.float64).reshape(-1)
        if a.size != self.count:
            raise ValueError(f"expected {self.count} values, got {a.size}")
        self._chk(_hip.hipMemcpy(C.c_void_p(self.ptr(frame)), a.ctypes.data_as(C.c_void_p), 8 * a.size, self.H2D),
                  "hipMemcpy")

    def download(self, frame=0):
        out = np.empty(self.count, np.float64)
        self._chk(_hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr(frame)), 8 * out.size, self.D2H),
                  "hipMemcpy")
        return out

    def close(self):
        if self._p.value:
            _hip.hipFree(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def test_convolution(data, recip, nx, ny, nz, precision=NB_PREC_FP64):
    """GPU convolution of the library's FFT kernels (test entry); precision
    NB_PREC_MIXED runs the FP32 transform."""
    d = np.ascontiguousarray(data, np.float64).copy()
    r = np.ascontiguousarray(recip, np.float64)
    rc = load().nb_test_convolution(_p(d, C.c_double), _p(r, C.c_double), nx, ny, nz, precision)
    if rc != NB_OK:
        raise NbError(rc, "nb_test_convolution failed")
    return d
