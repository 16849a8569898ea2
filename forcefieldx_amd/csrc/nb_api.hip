// nb_api.hip -- the C ABI of include/amoeba_nb.h and the evaluation driver.
//
// nb_evaluate follows ForceFieldEnergy.energy -> VanDerWaals.energy ->
// ParticleMeshEwald.energy (ForceFieldEnergy.java:1638-1800;
// ParticleMeshEwald.java:837-954, 1871-2183):
//   list check/rebuild, sorted coordinates, multipole rotation,
//   vdW, permanent PME (spread, convolution, gather), permanent real-space
//   field, direct dipoles, PCG, reciprocal and real-space energies,
//   torques -> gradients in caller order.
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <algorithm>
#include <vector>

#include "nb_kern.h"

namespace {


void cell_from_box(const double *box, CellD &c) {
  memset(&c, 0, sizeof(c));
  c.kmask = NB_INDEX_MASK;
  if (!box || box[0] <= 0.0) {
    c.aperiodic = 1;
    return;
  }
  double a = box[0], b = box[1], cc = box[2], al = box[3], be = box[4], ga = box[5];
  double cos_a, sin_b, cos_b, sin_g, cos_g, beta_term, gamma_term;
  if (al == 90.0 && be == 90.0 && ga == 90.0) {
    cos_a = 0.0; cos_b = 0.0; sin_b = 1.0; sin_g = 1.0; cos_g = 0.0; beta_term = 0.0; gamma_term = 1.0;
  } else {
    double d2r = M_PI / 180.0;
    cos_a = cos(al * d2r);
    sin_b = sin(be * d2r); cos_b = cos(be * d2r);
    sin_g = sin(ga * d2r); cos_g = cos(ga * d2r);
    beta_term = (cos_a - cos_b * cos_g) / sin_g;
    gamma_term = sqrt(sin_b * sin_b - beta_term * beta_term);
  }
  c.vol = sin_g * gamma_term * a * b * cc;
  double Ai[3][3] = {{a, 0.0, 0.0}, {b * cos_g, b * sin_g, 0.0}, {cc * cos_b, cc * beta_term, cc * gamma_term}};
  double det = Ai[0][0] * (Ai[1][1] * Ai[2][2] - Ai[1][2] * Ai[2][1]) - Ai[0][1] * (Ai[1][0] * Ai[2][2] - Ai[1][2] * Ai[2][0]) +
               Ai[0][2] * (Ai[1][0] * Ai[2][1] - Ai[1][1] * Ai[2][0]);
  double id = 1.0 / det;
  double A[3][3];
  A[0][0] = (Ai[1][1] * Ai[2][2] - Ai[1][2] * Ai[2][1]) * id;
  A[0][1] = (Ai[0][2] * Ai[2][1] - Ai[0][1] * Ai[2][2]) * id;
  A[0][2] = (Ai[0][1] * Ai[1][2] - Ai[0][2] * Ai[1][1]) * id;
  A[1][0] = (Ai[1][2] * Ai[2][0] - Ai[1][0] * Ai[2][2]) * id;
  A[1][1] = (Ai[0][0] * Ai[2][2] - Ai[0][2] * Ai[2][0]) * id;
  A[1][2] = (Ai[0][2] * Ai[1][0] - Ai[0][0] * Ai[1][2]) * id;
  A[2][0] = (Ai[1][0] * Ai[2][1] - Ai[1][1] * Ai[2][0]) * id;
  A[2][1] = (Ai[0][1] * Ai[2][0] - Ai[0][0] * Ai[2][1]) * id;
  A[2][2] = (Ai[0][0] * Ai[1][1] - Ai[0][1] * Ai[1][0]) * id;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) c.A[i][j] = A[i][j];
  c.A00 = A[0][0]; c.A10 = A[1][0]; c.A11 = A[1][1]; c.A20 = A[2][0]; c.A21 = A[2][1]; c.A22 = A[2][2];
  c.Ai00 = Ai[0][0]; c.Ai10 = Ai[1][0]; c.Ai11 = Ai[1][1]; c.Ai20 = Ai[2][0]; c.Ai21 = Ai[2][1]; c.Ai22 = Ai[2][2];
}

template <typename T>
void upload(T **dst, const T *src, size_t count) {
  HIPCHECK(hipMalloc((void **)dst, sizeof(T) * (count > 0 ? count : 1)));
  if (count) HIPCHECK(hipMemcpy(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice));
}

template <typename T>
void alloc(T **dst, size_t count) {
  HIPCHECK(hipMalloc((void **)dst, sizeof(T) * (count > 0 ? count : 1)));
}

void free_all(NbHandle *h) {
  nb_valence_free(h);
  nb_md_free(h);
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
  if (h->child_env) { free_all(h->child_env); delete h->child_env; h->child_env = nullptr; }
  if (h->child_vap) { free_all(h->child_vap); delete h->child_vap; h->child_vap = nullptr; }
  void *lam[] = {h->d_lgrad_s, h->d_ltorque_s, h->d_lgrad_r, h->d_ltorque_r, h->d_vap_map, h->d_tape, h->d_tape_sc,
                 h->d_fin, h->d_gfin, h->d_vfin};
  for (void *p : lam)
    if (p) hipFree(p);
  void *ptrs[] = {h->d_frame, h->d_axis, h->d_group0, h->d_vclass0, h->d_redidx, h->d_mpole_local, h->d_polar0,
                  h->d_thole0, h->d_ipd0, h->d_redval, h->d_mask_off, h->d_mask_atom, h->d_mask_rel, h->d_vdw_table,
                  h->d_xyz, h->d_xyz_list, h->d_perm, h->d_iperm, h->d_cellid, h->d_cellcount, h->d_cellstart,
                  h->d_nbr, h->d_ew, h->d_pc, h->d_sp, h->d_nsp, h->d_nnb, h->d_new, h->d_npc, h->d_flags, h->d_pos4, h->d_wrap, h->d_red4, h->d_mp,
                  h->d_gmp, h->d_polar, h->d_group, h->d_vclass, h->d_grid, h->d_recip, h->d_theta, h->d_igrid,
                  h->d_recipf, h->d_twf, h->d_bcount, h->d_bstart, h->d_bid, h->d_brank, h->d_batom, h->d_bbuf,
                  h->d_fmp, h->d_phip, h->d_phiu, h->d_phiup, h->d_tw, h->d_field, h->d_fieldp, h->d_u, h->d_up,
                  h->d_udir, h->d_updir, h->d_r, h->d_rp, h->d_z, h->d_zp, h->d_p, h->d_pp, h->d_vec, h->d_vecp,
                  h->d_scal, h->d_part, h->d_dip, h->d_dip2, h->d_dip64, h->d_theta64, h->d_vpart, h->d_vir, h->d_soft0, h->d_mol0, h->d_soft, h->d_mol, h->d_gred_l, h->d_dudx, h->d_grad_r, h->d_torque_r, h->d_grad_s, h->d_torque_s, h->d_gred_s, h->d_grad, h->d_energy, h->d_count};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (h->h_pinned) hipHostFree(h->h_pinned);
  if (h->d_pin) hipFree(h->d_pin);
  for (auto &r : h->prof_pool) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  if (h->ev_fork) hipEventDestroy(h->ev_fork);
  if (h->ev_join) hipEventDestroy(h->ev_join);
  if (h->ev_eps) hipEventDestroy(h->ev_eps);
  if (h->stream2) hipStreamDestroy(h->stream2);
  if (h->stream3) hipStreamDestroy(h->stream3);
  if (h->ev_fork3) hipEventDestroy(h->ev_fork3);
  if (h->ev_join3) hipEventDestroy(h->ev_join3);
  if (h->stream) hipStreamDestroy(h->stream);
}

void drop_graph(NbHandle *h) {
  if (h->gexec) hipGraphExecDestroy(h->gexec);
  if (h->graph) hipGraphDestroy(h->graph);
  h->gexec = nullptr;
  h->graph = nullptr;
}

int fail(NbHandle *h, const std::exception &e, int code) {
  if (h) h->err = e.what();
  return code;
}

// dst[3 map[j] + d] += src[3 j + d] (map null: identity)
__global__ void k_add_mapped(int m, const int32_t *__restrict__ map, const double *__restrict__ src,
                             double *__restrict__ dst) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * m) return;
  int j = t / 3, d = t % 3;
  int i = map ? map[j] : j;
  dst[3 * i + d] += src[t];
}

// dst[3 j + d] = src[3 map[j] + d]
__global__ void k_gather_mapped(int m, const int32_t *__restrict__ map, const double *__restrict__ src,
                                double *__restrict__ dst) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * m) return;
  int j = t / 3, d = t % 3;
  dst[t] = src[3 * map[j] + d];
}

void alloc_lambda_arrays(NbHandle *h) {
  size_t n3 = 3 * (size_t)h->n;
  double **v[] = {&h->d_lgrad_s, &h->d_ltorque_s, &h->d_lgrad_r, &h->d_ltorque_r};
  for (double **p : v) {
    alloc(p, n3);
    HIPCHECK(hipMemset(*p, 0, sizeof(double) * n3));
  }
  h->ost = true;
}

// OST lambda factors (AlchemicalParameters.update, AlchemicalParameters.java:316-385)
struct Alchemy {
  double lpow_perm = 1.0, dlpow_perm = 0.0, d2lpow_perm = 0.0;
  double lalpha = 0.0, dlalpha = 0.0, d2lalpha = 0.0;
  double lpow_pol = 1.0, dlpow_pol = 0.0, d2lpow_pol = 0.0;
};

Alchemy alchemy_of(const nb_params &p, double lam) {
  Alchemy f;
  if (lam < p.perm_lambda_start) {
    f.lpow_perm = 0.0;
  } else if (lam <= p.perm_lambda_end) {
    double sc = 1.0 / (p.perm_lambda_end - p.perm_lambda_start);
    double L = sc * (lam - p.perm_lambda_start), a = p.perm_lambda_alpha, b = p.perm_lambda_exponent;
    f.lalpha = a * (1.0 - L) * (1.0 - L);
    f.dlalpha = a * (1.0 - L) * sc;
    f.d2lalpha = -a * sc * sc;
    f.lpow_perm = pow(L, b);
    f.dlpow_perm = b * pow(L, b - 1.0) * sc;
    f.d2lpow_perm = b >= 2.0 ? b * (b - 1.0) * pow(L, b - 2.0) * sc * sc : 0.0;
  }
  if (lam < p.pol_lambda_start) {
    f.lpow_pol = 0.0;
  } else if (lam <= p.pol_lambda_end) {
    double sc = 1.0 / (p.pol_lambda_end - p.pol_lambda_start);
    double L = sc * (lam - p.pol_lambda_start), b = p.pol_lambda_exponent;
    if (b > 0.0) {
      f.lpow_pol = pow(L, b);
      if (b >= 1.0) {
        f.dlpow_pol = b * pow(L, b - 1.0) * sc;
        if (b >= 2.0) f.d2lpow_pol = b * (b - 1.0) * pow(L, b - 2.0) * sc * sc;
      }
    }
  }
  return f;
}

// Phase scales of the three OST evaluations at the handle's lambda
// (condensedEnergy / condensedNoLigandSCF / ligandElec,
// ParticleMeshEwald.java:1609-1790).
void set_ost_phases(NbHandle *h) {
  const nb_params &p = h->prm;
  double lam = h->lambda;
  Alchemy f = alchemy_of(p, lam);
  ElecPhase p1;
  p1.soft_pairs = 1;
  p1.l2 = f.lpow_perm; p1.lalpha = f.lalpha; p1.dlalpha = f.dlalpha; p1.d2lalpha = f.d2lalpha;
  p1.perm = f.lpow_perm; p1.dperm = f.dlpow_perm; p1.d2perm = f.d2lpow_perm;
  if (lam < p.pol_lambda_start) { p1.polar = 0; p1.pol = 0.0; }
  else { p1.pol = f.lpow_pol; p1.dpol = f.dlpow_pol; p1.d2pol = f.d2lpow_pol; }
  h->ph = p1;
  if (h->child_env) {
    ElecPhase p2;
    p2.perm_real = 0;
    p2.perm = 1.0 - f.lpow_perm; p2.dperm = -f.dlpow_perm; p2.d2perm = -f.d2lpow_perm;
    if (lam <= p.pol_lambda_end && p.no_ligand_condensed_scf) {
      p2.pol = 1.0 - f.lpow_pol; p2.dpol = -f.dlpow_pol; p2.d2pol = -f.d2lpow_pol;
    } else {
      p2.polar = 0; p2.pol = 0.0;
    }
    h->child_env->ph = p2;
  }
  if (h->child_vap) {
    ElecPhase p3;
    p3.soft_pairs = 1; p3.vapor = 1; p3.sor = 1;
    p3.inter = p.intermolecular_softcore; p3.intra = p.intramolecular_softcore;
    p3.l2 = 1.0 - f.lpow_perm;
    p3.perm = 1.0 - f.lpow_perm; p3.dperm = -f.dlpow_perm; p3.d2perm = -f.d2lpow_perm;
    if (lam <= p.pol_lambda_end) {
      p3.pol = 1.0 - f.lpow_pol; p3.dpol = -f.dlpow_pol; p3.d2pol = -f.d2lpow_pol;
    } else {
      p3.polar = 0; p3.pol = 0.0;
    }
    h->child_vap->ph = p3;
  }
}

}  // namespace

// for the other translation units (nb_set_valence)
void nb_drop_graph(NbHandle *h) { drop_graph(h); }

extern "C" int nb_create(const nb_params *params, const nb_topology *top, void **handle);
extern "C" int nb_set_cell(void *handle, const double box[6]);

namespace {

// The no-ligand phase is this system with the soft atoms' multipoles and
// polarizabilities removed (an atom with use[i] = false contributes nothing
// else); the vapor phase is the soft atoms plus their frame atoms (the
// latter without multipole or polarizability) in an aperiodic cell with no
// Ewald screening and no cutoff (ParticleMeshEwald.java:1704-1790; its SOR
// solve is replaced by PCG to the same tolerance).
void create_ost_children(NbHandle *h, const nb_params *params, const nb_topology *top) {
  int n = top->n_atoms;
  nb_params cp = *params;
  cp.elec_lambda = 0;
  cp.do_vdw = 0;
  cp.lambda = 1.0;
  // no-ligand phase
  std::vector<double> mp(top->multipole, top->multipole + 10 * (size_t)n), pol(top->polarizability,
                                                                              top->polarizability + n);
  int nsoft = 0;
  for (int i = 0; i < n; i++)
    if (top->soft[i]) {
      nsoft++;
      std::fill(mp.begin() + 10 * (size_t)i, mp.begin() + 10 * (size_t)i + 10, 0.0);
      pol[i] = 0.0;
    }
  if (nsoft < n) {
    nb_topology t2 = *top;
    t2.multipole = mp.data();
    t2.polarizability = pol.data();
    t2.soft = nullptr;
    void *c = nullptr;
    int rc = nb_create(&cp, &t2, &c);
    if (rc != NB_OK) throw std::runtime_error("could not create the no-ligand phase");
    h->child_env = (NbHandle *)c;
    alloc_lambda_arrays(h->child_env);
  }
  // vapor phase
  if (params->ligand_vapor_elec) {
    std::vector<int32_t> map, inv(n, -1);
    for (int i = 0; i < n; i++)
      if (top->soft[i]) { inv[i] = (int)map.size(); map.push_back(i); }
    for (int i = 0; i < n; i++)
      if (top->soft[i])
        for (int a = 0; a < 3; a++) {
          int k = top->axis[3 * i + a];
          if (k >= 0 && inv[k] < 0) { inv[k] = (int)map.size(); map.push_back(k); }
        }
    int m = (int)map.size();
    std::vector<int32_t> fr(m, NB_FRAME_NONE), ax(3 * (size_t)m, -1), grp(m), vc(m, 0), ri(m), sf(m), mol(m);
    std::vector<double> vmp(10 * (size_t)m, 0.0), vpol(m, 0.0), th(m), pd(m), rv(m, 0.0);
    std::vector<int32_t> moff(m + 1, 0), mat, mrel;
    for (int j = 0; j < m; j++) {
      int i = map[j];
      th[j] = top->thole[i];
      pd[j] = top->pdamp[i];
      grp[j] = top->group[i];
      ri[j] = j;
      sf[j] = top->soft[i] != 0;
      mol[j] = top->molecule ? top->molecule[i] : i;
      if (top->soft[i]) {
        std::copy(top->multipole + 10 * (size_t)i, top->multipole + 10 * (size_t)i + 10, vmp.begin() + 10 * (size_t)j);
        vpol[j] = top->polarizability[i];
        fr[j] = top->frame[i];
        for (int a = 0; a < 3; a++) {
          int k = top->axis[3 * i + a];
          ax[3 * j + a] = k >= 0 ? inv[k] : -1;
        }
      }
      for (int q = top->mask_offset[i]; q < top->mask_offset[i + 1]; q++)
        if (inv[top->mask_atom[q]] >= 0) {
          mat.push_back(inv[top->mask_atom[q]]);
          mrel.push_back(top->mask_rel[q]);
        }
      moff[j + 1] = (int32_t)mat.size();
    }
    nb_topology t3 = *top;
    t3.n_atoms = m;
    t3.frame = fr.data(); t3.axis = ax.data(); t3.multipole = vmp.data(); t3.polarizability = vpol.data();
    t3.thole = th.data(); t3.pdamp = pd.data(); t3.group = grp.data(); t3.vdw_class = vc.data();
    t3.red_index = ri.data(); t3.red_value = rv.data(); t3.mask_offset = moff.data();
    t3.mask_atom = mat.data(); t3.mask_rel = mrel.data(); t3.soft = sf.data(); t3.molecule = mol.data();
    nb_params vp = cp;
    vp.ewald_cutoff = 1.0e4;   // setEwaldParameters(Double.MAX_VALUE, 0.0)
    vp.vdw_cutoff = 0.0;
    void *c = nullptr;
    int rc = nb_create(&vp, &t3, &c);
    if (rc != NB_OK) throw std::runtime_error("could not create the vapor phase");
    h->child_vap = (NbHandle *)c;
    alloc_lambda_arrays(h->child_vap);
    if (nb_set_cell(c, nullptr) != NB_OK) throw std::runtime_error("could not set the vapor cell");
    upload(&h->d_vap_map, map.data(), m);
    h->n_vap = m;
  }
}

}  // namespace

// the next pair of timing events of the profiled evaluation
int nb_prof_slot(NbHandle *h, const char *name) {
  if (h->prof_n == (int)h->prof_pool.size()) {
    NbHandle::ProfRec r;
    HIPCHECK(hipEventCreate(&r.a));
    HIPCHECK(hipEventCreate(&r.b));
    h->prof_pool.push_back(r);
  }
  int slot = h->prof_n++;
  h->prof_pool[slot].name = name;
  return slot;
}

// accumulate elapsed times by kernel name; requires a completed stream
void nb_prof_harvest(NbHandle *h) {
  for (int i = 0; i < h->prof_n; i++) {
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, h->prof_pool[i].a, h->prof_pool[i].b));
    const std::string &nm = h->prof_pool[i].name;
    size_t k = 0;
    while (k < h->prof_names.size() && h->prof_names[k] != nm) k++;
    if (k == h->prof_names.size()) {
      h->prof_names.push_back(nm);
      h->prof_ms.push_back(0.0);
      h->prof_count.push_back(0);
    }
    h->prof_ms[k] += ms;
    h->prof_count[k] += 1;
  }
  h->prof_n = 0;
}

extern "C" {

int nb_set_profiling(void *handle, int on) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  h->profiling = on != 0;
  h->prof_names.clear();
  h->prof_ms.clear();
  h->prof_count.clear();
  h->prof_n = 0;
  return NB_OK;
}

int nb_set_profile_filter(void *handle, const char *substring) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  h->prof_filter = substring ? substring : "";
  return NB_OK;
}

int nb_profile_count(void *handle) {
  NbHandle *h = (NbHandle *)handle;
  return h ? (int)h->prof_names.size() : 0;
}

int nb_profile_entry(void *handle, int idx, char *name, int cap, double *total_ms, int64_t *launches) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || idx < 0 || idx >= (int)h->prof_names.size() || !name || cap <= 0) return NB_ERR_INVALID;
  snprintf(name, cap, "%s", h->prof_names[idx].c_str());
  if (total_ms) *total_ms = h->prof_ms[idx];
  if (launches) *launches = h->prof_count[idx];
  return NB_OK;
}

int nb_abi_version(void) { return NB_ABI_VERSION; }

int nb_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

int nb_create(const nb_params *params, const nb_topology *top, void **handle) {
  if (!params || !top || !handle || top->n_atoms <= 0) return NB_ERR_INVALID;
  int ndev = nb_device_count();
  if (ndev <= 0) return NB_ERR_NO_DEVICE;
  NbHandle *h = new NbHandle();
  try {
    h->prm = *params;
    h->n = top->n_atoms;
    h->device = params->device;
    HIPCHECK(hipSetDevice(h->device));
    HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&h->stream3, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&h->ev_fork3, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&h->ev_join3, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&h->ev_eps, hipEventDisableTiming));
    h->cur = h->stream;
    {
      hipDeviceProp_t prop;
      HIPCHECK(hipGetDeviceProperties(&prop, h->device));
      h->lds_max = std::min<size_t>(prop.sharedMemPerBlock, 65536);
    }
    int n = h->n;
    if (n >= (int)NB_INDEX_MASK) throw std::runtime_error("too many atoms");
    for (int d = 0; d < 3; d++) {
      int m = params->grid[d];
      if (params->do_pme && m < 8) throw std::runtime_error("PME grid must be at least 8 per dimension");
      while (m > 1 && m % 2 == 0) m /= 2;
      while (m > 1 && m % 3 == 0) m /= 3;
      while (m > 1 && m % 5 == 0) m /= 5;
      while (m > 1 && m % 7 == 0) m /= 7;
      // the factors Complex.preferredDimension admits (2, 3, 5, 6, 7)
      if (params->do_pme && m != 1) throw std::runtime_error("PME grid dimensions must factor over 2, 3, 5 and 7");
    }
    if (params->pme_order != 5) throw std::runtime_error("only B-spline order 5 is implemented");
    // every index a kernel dereferences is validated here
    for (int i = 0; i < n; i++) {
      int f = top->frame[i];
      if (f < NB_FRAME_NONE || f > NB_FRAME_THREEFOLD) throw std::runtime_error("invalid frame type");
      int need = f == NB_FRAME_NONE ? 0 : (f == NB_FRAME_ZONLY ? 1 : (f == NB_FRAME_ZTHENBISECTOR || f == NB_FRAME_THREEFOLD ? 3 : 2));
      for (int a = 0; a < need; a++) {
        int k = top->axis[3 * i + a];
        if (k < 0 || k >= n || k == i) throw std::runtime_error("invalid frame axis atom");
      }
      int r = top->red_index[i];
      if (r < 0 || r >= n) throw std::runtime_error("invalid reduction index");
      if (top->vdw_class[i] < 0 || top->vdw_class[i] >= top->n_vdw_class) throw std::runtime_error("invalid vdW class");
      if (top->mask_offset[i + 1] < top->mask_offset[i]) throw std::runtime_error("invalid mask offsets");
    }
    if (top->mask_offset[0] != 0) throw std::runtime_error("invalid mask offsets");
    for (int t = 0; t < top->mask_offset[n]; t++)
      if (top->mask_atom[t] < 0 || top->mask_atom[t] >= n) throw std::runtime_error("invalid mask atom");
    std::vector<double> ipd(n);
    for (int i = 0; i < n; i++) ipd[i] = top->pdamp[i] > 0.0 ? 1.0 / top->pdamp[i] : INFINITY;
    upload(&h->d_frame, top->frame, n);
    upload(&h->d_axis, top->axis, 3 * (size_t)n);
    upload(&h->d_group0, top->group, n);
    upload(&h->d_vclass0, top->vdw_class, n);
    upload(&h->d_redidx, top->red_index, n);
    upload(&h->d_mpole_local, top->multipole, 10 * (size_t)n);
    upload(&h->d_polar0, top->polarizability, n);
    upload(&h->d_thole0, top->thole, n);
    upload(&h->d_ipd0, ipd.data(), n);
    upload(&h->d_redval, top->red_value, n);
    int nnz = top->mask_offset[n];
    upload(&h->d_mask_off, top->mask_offset, n + 1);
    upload(&h->d_mask_atom, top->mask_atom, nnz);
    upload(&h->d_mask_rel, top->mask_rel, nnz);
    h->n_class = top->n_vdw_class;
    h->lambda = params->lambda;
    if (!(h->lambda >= 0.0 && h->lambda <= 1.0)) throw std::runtime_error("lambda must lie in [0, 1]");
    if (top->soft) {
      for (int i = 0; i < n; i++) h->has_soft = h->has_soft || top->soft[i] != 0;
    }
    if (h->has_soft) {
      std::vector<int32_t> mol(n);
      for (int i = 0; i < n; i++) mol[i] = top->molecule ? top->molecule[i] : i;
      upload(&h->d_soft0, top->soft, n);
      upload(&h->d_mol0, mol.data(), n);
      alloc(&h->d_soft, n);
      alloc(&h->d_mol, n);
    }
    upload(&h->d_vdw_table, top->vdw_table, 2 * (size_t)h->n_class * h->n_class);
    size_t n3 = 3 * (size_t)n;
    alloc(&h->d_xyz, n3);
    alloc(&h->d_xyz_list, n3);
    alloc(&h->d_perm, n);
    alloc(&h->d_iperm, n);
    alloc(&h->d_cellid, n);
    alloc(&h->d_nnb, n);
    alloc(&h->d_new, n);
    alloc(&h->d_npc, n);
    alloc(&h->d_nsp, n);
    {
      int mx = 0;
      for (int i = 0; i < n; i++) mx = std::max(mx, top->mask_offset[i + 1] - top->mask_offset[i]);
      h->maxsp = mx + 16;
      alloc(&h->d_sp, (size_t)h->maxsp * n);
    }
    alloc(&h->d_flags, 8);
    HIPCHECK(hipMemset(h->d_flags, 0, 8 * sizeof(int32_t)));
    alloc(&h->d_pos4, n);
    alloc(&h->d_wrap, 3 * (size_t)n);
    HIPCHECK(hipMemset(h->d_wrap, 0, sizeof(int32_t) * 3 * n));
    alloc(&h->d_red4, n);
    alloc(&h->d_mp, n);
    HIPCHECK(hipMemset(h->d_mp, 0, sizeof(AtomMP) * n));
    alloc(&h->d_gmp, 10 * (size_t)n);
    alloc(&h->d_polar, n);
    alloc(&h->d_group, n);
    alloc(&h->d_vclass, n);
    alloc(&h->d_theta, 60 * (size_t)n);
    alloc(&h->d_bid, n);
    alloc(&h->d_brank, n);
    alloc(&h->d_batom, n);
    alloc(&h->d_igrid, n3);
    alloc(&h->d_fmp, 10 * (size_t)n);
    alloc(&h->d_phip, 30 * (size_t)n);
    alloc(&h->d_phiu, 30 * (size_t)n);
    alloc(&h->d_phiup, 30 * (size_t)n);
    HIPCHECK(hipMemset(h->d_phip, 0, sizeof(double) * 30 * n));
    HIPCHECK(hipMemset(h->d_phiu, 0, sizeof(double) * 30 * n));
    HIPCHECK(hipMemset(h->d_phiup, 0, sizeof(double) * 30 * n));
    double **vecs[] = {&h->d_field, &h->d_fieldp, &h->d_u, &h->d_up, &h->d_udir, &h->d_updir, &h->d_r, &h->d_rp,
                       &h->d_z, &h->d_zp, &h->d_p, &h->d_pp, &h->d_vec, &h->d_vecp, &h->d_grad_s, &h->d_torque_s,
                       &h->d_gred_s, &h->d_grad, &h->d_grad_r, &h->d_torque_r, &h->d_gred_l, &h->d_dudx};
    for (double **v : vecs) {
      alloc(v, n3);
      HIPCHECK(hipMemset(*v, 0, sizeof(double) * n3));
    }
    alloc(&h->d_scal, S_NSLOT);
    alloc(&h->d_vir, 16);
    alloc(&h->d_dip, 12 * (size_t)n);
    alloc(&h->d_dip2, 12 * (size_t)n);
    if (params->precision == NB_PREC_MIXED) {
      // FP64 iterative refinement of the solve (on by default in the mixed form)
      alloc(&h->d_dip64, 12 * (size_t)n);
      alloc(&h->d_theta64, 60 * (size_t)n);
      h->refine = true;
    }
    // producers of PCG partial sums: vector kernels (3n threads), the gather
    // (8 lanes per atom) and the field kernels (NB_LPA_PCG lanes per atom; 16
    // lanes bound every row kernel)
    h->part_stride = (int)std::max({grid_for(3 * (long)n, 256).x, xcd_grid((long)n * 8, 256).x,
                                    xcd_grid((long)n * 16, 256).x});
    alloc(&h->d_part, 16 * (size_t)h->part_stride);
    alloc(&h->d_energy, 16);
    alloc(&h->d_count, 4);
    HIPCHECK(hipHostMalloc((void **)&h->h_pinned, sizeof(double) * P_NSLOT, hipHostMallocDefault));
    alloc(&h->d_pin, P_NSLOT);
    memset(&h->last, 0, sizeof(h->last));
    if (params->elec_lambda && h->has_soft && params->do_pme) {
      if (!(params->perm_lambda_end > params->perm_lambda_start && params->pol_lambda_end > params->pol_lambda_start))
        throw std::runtime_error("invalid alchemical lambda windows");
      alloc_lambda_arrays(h);
      h->elec_lambda = true;
      create_ost_children(h, params, top);
    }
    *handle = h;
    return NB_OK;
  } catch (std::exception &e) {
    int code = NB_ERR_HIP;
    if (std::string(e.what()).find("hip") == std::string::npos) code = NB_ERR_INVALID;
    free_all(h);
    delete h;
    return code;
  }
}

int nb_set_cell(void *handle, const double box[6]) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  try {
    HIPCHECK(hipSetDevice(h->device));
    double zero[6] = {0, 0, 0, 90, 90, 90};
    const double *b = box ? box : zero;
    for (int d = 0; d < 6; d++) h->box[d] = b[d];
    drop_graph(h);
    cell_from_box(b, h->cell);
    if (!h->cell.aperiodic) {
      // image sums when a perpendicular width is below twice the interaction
      // cutoff (the reference then builds a ReplicatesCrystal,
      // ForceFieldEnergy.java:893-901)
      double cmax = std::max(h->prm.ewald_cutoff, h->prm.vdw_cutoff);
      double lcut = cmax + h->prm.list_buffer;
      double w[3];
      for (int d = 0; d < 3; d++) {
        const double(*A)[3] = h->cell.A;
        w[d] = 1.0 / sqrt(A[0][d] * A[0][d] + A[1][d] * A[1][d] + A[2][d] * A[2][d]);
        if (0.5 * w[d] < cmax) h->cell.multi = 1;
      }
      for (int d = 0; d < 3; d++)
        if (h->cell.multi && (int)floor(lcut / w[d]) + 1 > NB_IMG_R)
          throw std::runtime_error("cell too thin for image sums at this cutoff");
      if (h->cell.multi) {
        if (h->n >= (1 << NB_IMG_SHIFT)) throw std::runtime_error("too many atoms for image sums");
        h->cell.kmask = (1u << NB_IMG_SHIFT) - 1u;
      }
    }
    h->cell_set = true;
    h->list_valid = false;
    if (h->prm.do_pme && !h->cell.aperiodic && h->prm.ewald_alpha > 0.0) nb_pme_setup(h);
    if (h->child_env) {
      int rc = nb_set_cell(h->child_env, box);
      if (rc != NB_OK) throw std::runtime_error(h->child_env->err);
    }
    return NB_OK;
  } catch (std::exception &e) {
    return fail(h, e, NB_ERR_HIP);
  }
}

int nb_set_positions(void *handle, const double *xyz) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || !xyz) return NB_ERR_INVALID;
  try {
    HIPCHECK(hipSetDevice(h->device));
    HIPCHECK(hipMemcpyAsync(h->d_xyz, xyz, sizeof(double) * 3 * h->n, hipMemcpyHostToDevice, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    h->positions_set = true;
    return NB_OK;
  } catch (std::exception &e) {
    return fail(h, e, NB_ERR_HIP);
  }
}

int nb_set_positions_device(void *handle, const void *d_xyz) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || !d_xyz) return NB_ERR_INVALID;
  try {
    HIPCHECK(hipSetDevice(h->device));
    HIPCHECK(hipMemcpyAsync(h->d_xyz, d_xyz, sizeof(double) * 3 * h->n, hipMemcpyDeviceToDevice, h->stream));
    h->positions_set = true;
    return NB_OK;
  } catch (std::exception &e) {
    return fail(h, e, NB_ERR_HIP);
  }
}

int nb_set_lambda(void *handle, double lambda) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  if (!(lambda >= 0.0 && lambda <= 1.0)) {
    h->err = "lambda must lie in [0, 1]";
    return NB_ERR_INVALID;
  }
  if (lambda != h->lambda) drop_graph(h);
  h->lambda = lambda;
  return NB_OK;
}

int nb_set_graph(void *handle, int on) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  h->graph_on = on != 0;
  if (!h->graph_on) drop_graph(h);
  if (h->child_env) nb_set_graph(h->child_env, on);
  if (h->child_vap) nb_set_graph(h->child_vap, on);
  return NB_OK;
}

int nb_set_refinement(void *handle, int on) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  if (on && !h->d_dip64) {
    h->err = "FP64 refinement applies to the mixed form only";
    return NB_ERR_INVALID;
  }
  if ((on != 0) != h->refine) drop_graph(h);
  h->refine = on != 0;
  if (h->child_env) nb_set_refinement(h->child_env, on);
  if (h->child_vap) nb_set_refinement(h->child_vap, on);
  return NB_OK;
}

int nb_graph_replays(void *handle) {
  NbHandle *h = (NbHandle *)handle;
  return h ? (int)h->graph_replays : -1;
}

int nb_struct_sizes(int32_t sizes[3]) {
  if (!sizes) return NB_ERR_INVALID;
  sizes[0] = (int32_t)sizeof(nb_params);
  sizes[1] = (int32_t)sizeof(nb_topology);
  sizes[2] = (int32_t)sizeof(nb_energies);
  return NB_OK;
}

// Enqueue one evaluation on the handle's streams, ending with the energies,
// counts, flags (and virial, PCG scalars) copied into pinned host memory.
// capture: the list check only raises flags[3] (a list built for other
// coordinates is not rebuilt), the PCG runs graph_iters iterations under
// device-side control, and nothing synchronises with the host, so the
// sequence can be recorded into a HIP graph.
static int issue_evaluation(NbHandle *h, bool capture, nb_energies &e, bool &rebuilt) {
  int status = NB_OK;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  h->vpart_used = 0;
  h->used_sor = false;
  h->finite_active = false;
  rebuilt = false;
  if (capture) {
    // the cleared flags include the list check's, which k_max_disp raises
    nb_zero_outputs(h);
    nb_check_rebuild_async(h);
    nb_gather_sorted(h);
  } else {
    rebuilt = nb_need_rebuild(h);
    if (rebuilt) nb_build_lists(h);
    else nb_gather_sorted(h);
    nb_zero_outputs(h);
  }
  if (h->prm.do_vdw || h->prm.do_pme) nb_rotate(h);
  // vdW and the valence terms run on the third stream, overlapping everything
  // up to the gradient assembly
  const bool third = h->prm.do_vdw || h->has_valence;
  if (third) {
    nb_fork3(h);
    NbThird t3(h);
    if (h->prm.do_vdw) nb_vdw(h);
    nb_valence_terms(h);
  }
  bool polar = h->prm.polarization != NB_POL_NONE && h->ph.polar;
  if (h->prm.do_pme) {
    // Ewald / preconditioner / masked rows and the real-space permanent
    // field (side) || permanent PME (main)
    nb_fork(h);
    {
      NbSide side(h);
      nb_filter_lists(h);
      nb_perm_field(h);
    }
    if (use_recip) nb_pme_permanent(h);
    nb_join(h);
    if (polar) {
      nb_direct(h);
      if (h->prm.polarization == NB_POL_MUTUAL) {
        // the SCF predictor's extrapolation is the solve's starting point
        // (not for the finite-PCG gradient, which starts from the direct dipoles)
        if (!h->want_finite) nb_pred_apply(h);
        if (capture) {
          status = nb_pcg_device(h, h->graph_iters);
          nb_pred_save(h, true);
        } else if (h->prm.scf_algorithm == NB_SCF_SOR || h->ph.sor) {
          status = nb_sor(h, h->prm.polar_sor, &e);
          h->used_sor = true;
          if (status == NB_OK) nb_pred_save(h, false);
        } else {
          h->tape_on = h->want_finite;
          h->tape_K = 0;
          status = nb_pcg(h, &e);
          h->tape_on = false;
          if ((status == NB_ERR_SCF_DIVERGED || status == NB_ERR_SCF_MAXITER) && h->prm.sor_fallback) {
            // ParticleMeshEwald.java:2300-2312: SOR from the direct dipoles with
            // the conservative factor 0.65
            nb_reset_dipoles(h);
            status = nb_sor(h, 0.65, &e);
            h->used_sor = true;
          } else if (status == NB_OK && h->want_finite) {
            // gradient of the dipoles this solve returned: uCR -> adjoint
            nb_finite_adjoint(h);
            h->finite_active = true;
          }
          if (status == NB_OK && !h->finite_active) nb_pred_save(h, false);
        }
      } else if (use_recip) {
        // direct polarization: only the reciprocal potentials of u_direct are needed
        nb_pme_induced(h, h->d_u, h->d_up, true);
      }
    }
    if (h->want_vir && use_recip) nb_grid_virial(h, polar);
    // permanent real-space energy (side) || reciprocal energy, then the
    // polarization real-space energy added into the reciprocal arrays (main)
    nb_fork(h);
    {
      NbSide side(h);
      nb_real_energy(h, polar, use_recip ? NB_RE_PERM : NB_RE_ALL);
    }
    if (use_recip) {
      nb_recip_energy(h, polar);
      if (polar) nb_real_energy(h, polar, NB_RE_POL);
    }
  }
  nb_join(h);
  if (third) nb_join3(h);
  if (h->finite_active) nb_finite_add(h);
  nb_finalize_gradient(h);
  if (h->has_valence) nb_valence_gather(h);
  if (h->want_vir) nb_vir_reduce(h);
  // energies, counts, flags (and the virial, valence energies, PCG scalars)
  // packed into h_pinned's layout on the device, then one transfer
  nb_pack_outputs(h, capture);
  HIPCHECK(hipMemcpyAsync(h->h_pinned, h->d_pin, sizeof(double) * P_NSLOT, hipMemcpyDeviceToHost, h->stream));
  return status;
}

// Record the evaluation sequence into a graph (not launched here).
static void capture_graph(NbHandle *h, uint32_t flags, int iters) {
  drop_graph(h);
  h->graph_iters = iters;
  h->capturing = true;
  nb_energies e;
  memset(&e, 0, sizeof(e));
  bool rebuilt = false;
  hipGraph_t g = nullptr;
  const bool md = (flags & NB_FLAG_MD_STEP) != 0;
  HIPCHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  try {
    // a recorded MD step: the drift before the evaluation, the guarded kick,
    // kinetic energy and thermostat after it
    if (md) nb_md_drift(h);
    issue_evaluation(h, true, e, rebuilt);
    if (md) nb_md_kick(h, true);
  } catch (...) {
    hipStreamEndCapture(h->stream, &g);
    if (g) hipGraphDestroy(g);
    h->capturing = false;
    throw;
  }
  h->capturing = false;
  HIPCHECK(hipStreamEndCapture(h->stream, &g));
  h->graph = g;
  HIPCHECK(hipGraphInstantiate(&h->gexec, h->graph, nullptr, nullptr, 0));
  h->graph_flags = flags;
  if (md) {
    h->md_graph_dt = h->md_dt;
    h->md_graph_thermo = h->md_thermo;
    h->md_graph_t0 = h->md_t0;
    h->md_graph_tau = h->md_tau;
  }
}

static int nb_evaluate_one(void *handle, uint32_t flags, nb_energies *out) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  if (!h->cell_set || !h->positions_set) {
    h->err = "nb_set_cell and nb_set_positions must be called before nb_evaluate";
    return NB_ERR_INVALID;
  }
  nb_energies e;
  memset(&e, 0, sizeof(e));
  int status = NB_OK;
  h->induced_caller = false;
  try {
    HIPCHECK(hipSetDevice(h->device));
    int n = h->n;
    double en[16];
    int64_t cnt[4];
    int32_t fl[8];
    bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
    int rebuilt = 0;
    h->want_vir = (flags & NB_VIRIAL) != 0;
    h->vir_valid = false;
    h->want_dudl = (flags & NB_DUDL) != 0;
    h->dudl_valid = false;
    h->want_finite = (flags & NB_FINITE_PCG) != 0 && h->prm.polarization == NB_POL_MUTUAL;
    if (h->want_finite && !h->d_gfin) HIPCHECK(hipMalloc(&h->d_gfin, sizeof(double) * 3 * n));
    if (h->want_vir) {
      // an aperiodic system has pair and frame terms only
      // 9 partial sums per block of every contributing launch
      long pb = xcd_grid((long)n * 16, 256).x, sb = xcd_grid((long)n * 4, 256).x, ab = (n + 127) / 128;
      long need = 5 * pb + 2 * sb + 2 * ab + (h->ntot + 255) / 256 + ((long)n * 8 + 255) / 256 + 64 + NB_VIR_STAGE;
      if (need > h->vpart_cap) {
        drop_graph(h);
        if (h->d_vpart) hipFree(h->d_vpart);
        h->d_vpart = nullptr;
        HIPCHECK(hipMalloc(&h->d_vpart, sizeof(double) * 9 * need));
        h->vpart_cap = need;
      }
    }
    bool mutual = h->prm.do_pme && h->prm.polarization == NB_POL_MUTUAL && h->ph.polar;
    bool replayed = false, shrink_graph = false, drifted = false;
    const bool md = (flags & NB_FLAG_MD_STEP) != 0;
    if (md && h->gexec && (h->graph_flags & NB_FLAG_MD_STEP) &&
        (h->md_graph_dt != h->md_dt || h->md_graph_thermo != h->md_thermo || h->md_graph_t0 != h->md_t0 ||
         h->md_graph_tau != h->md_tau))
      drop_graph(h);
    if (h->graph_on && h->gexec && h->graph_flags == flags && !h->profiling && !h->want_finite) {
      // replay; a list that needs rebuilding, an overflowed filtered row or a
      // solve that did not stop within the recorded iterations sends the
      // evaluation down the regular path below
      if (!h->graph_launched) HIPCHECK(hipGraphLaunch(h->gexec, h->stream));
      h->graph_launched = false;
      // a recorded MD step has drifted the positions whether or not its
      // evaluation is kept
      drifted = md;
      HIPCHECK(hipStreamSynchronize(h->stream));
      memcpy(fl, h->h_pinned + 12, sizeof(fl));
      const double *sc = h->h_pinned + 40;
      bool ok = fl[3] == 0 && !(h->prm.do_pme && nb_grow_filtered_rows(h, fl)) &&
                (!mutual || (sc[S_DONE] != 0.0 && sc[S_STAT] == 0.0));
      if (ok) {
        replayed = true;
        h->graph_replays++;
        if (mutual) {
          e.pcg_iters = (int)sc[S_ITER];
          e.pcg_eps = sc[S_EPSV];
          status = (int)sc[S_STAT];
          // the recorded budget's surplus iterations still cost their
          // (early-returning) launches; when solves have become cheaper (an
          // SCF predictor warming up) the graph is recorded again at the next
          // evaluation (against the first solve's budget: the recorded
          // refinement corrections only run when a residual asks for them)
          if (e.pcg_iters + 2 < h->graph_iters) shrink_graph = true;
          // recorded refinement corrections that replays keep skipping
          // (S_RDONE after the first residual) still cost their launches:
          // after eight such replays in a row the graph is recorded again
          if (h->graph_refine_n > 0) {
            if ((int)sc[S_RCYC] < h->graph_refine_n) {
              if (++h->refine_idle >= 8) shrink_graph = true;
            } else {
              h->refine_idle = 0;
            }
          }
        }
      } else {
        drop_graph(h);
      }
    }
    if (replayed) {
      memcpy(en, h->h_pinned, 8 * sizeof(double));
      en[E_DUDL] = h->h_pinned[32];
      en[E_D2UDL2] = h->h_pinned[33];
      memcpy(cnt, h->h_pinned + 8, sizeof(cnt));
    } else {
      // the regular path may grow or rebuild the rows a recorded graph points to
      drop_graph(h);
      if (md && !drifted) nb_md_drift(h);
      // A filtered (Ewald / preconditioner / masked) row that overflowed its
      // capacity is only detected when the flags come back with the energies;
      // the rows are then grown and the evaluation repeated.
      for (int attempt = 0;; attempt++) {
        bool rb = false;
        status = issue_evaluation(h, false, e, rb);
        rebuilt |= rb ? 1 : 0;
        HIPCHECK(hipStreamSynchronize(h->stream));
        if (h->profiling) nb_prof_harvest(h);
        memcpy(en, h->h_pinned, 8 * sizeof(double));
        en[E_DUDL] = h->h_pinned[32];
        en[E_D2UDL2] = h->h_pinned[33];
        memcpy(cnt, h->h_pinned + 8, sizeof(cnt));
        memcpy(fl, h->h_pinned + 12, sizeof(fl));
        if (!h->prm.do_pme || !nb_grow_filtered_rows(h, fl)) break;
        if (attempt == 3) throw std::runtime_error("Ewald list capacity could not be established");
        memset(&e, 0, sizeof(e));
      }
    }
    e.list_rebuilt = rebuilt;
    if (h->want_vir) {
      for (int q = 0; q < 9; q++) h->vir[q] = h->h_pinned[16 + q];
    }
    e.perm_real = en[E_PERM_REAL];
    e.pol_real = en[E_POL_REAL];
    e.vdw = en[E_VDW];
    e.dudl = en[E_DUDL];
    e.d2udl2 = en[E_D2UDL2];
    h->dudl_valid = h->want_dudl;
    if (use_recip && h->prm.do_pme) {
      e.perm_self = en[E_PERM_SELF];
      e.perm_recip = en[E_PERM_RECIP];
      e.pol_self = en[E_POL_SELF];
      e.pol_recip = en[E_POL_RECIP];
      double q = en[E_QTOT];
      if (q != 0.0) {
        double cc = -0.5 * h->prm.electric * M_PI * q * q / (h->cell.vol * h->prm.ewald_alpha * h->prm.ewald_alpha);
        e.perm_charge_corr = h->ph.perm * cc;
        if (h->ost) {
          e.dudl += h->ph.dperm * cc;
          e.d2udl2 += h->ph.d2perm * cc;
        }
      }
    }
    if (h->want_vir) {
      // the net-charge correction scales as 1 / V
      for (int a = 0; a < 3; a++) h->vir[4 * a] -= e.perm_charge_corr;
      h->vir_valid = true;
    }
    e.n_vdw = cnt[0] / 2;
    e.n_perm = cnt[1] / 2 + ((use_recip && h->prm.do_pme) ? n : 0);
    e.perm = e.perm_real + e.perm_self + e.perm_recip + e.perm_charge_corr;
    e.pol = e.pol_real + e.pol_self + e.pol_recip;
    e.total = e.perm + e.pol + e.vdw;
    if (h->has_valence)
      for (int k = 0; k < 8; k++) {
        e.valence[k] = h->h_pinned[P_VALENCE + k];
        e.total += e.valence[k];
      }
    if (status == NB_OK && fl[4]) status = NB_ERR_BAD_TORQUE;
    if (status == NB_OK && !std::isfinite(e.total)) status = NB_ERR_NAN_ENERGY;
    // the step's kick on the regular path (the recorded kick of a replay that
    // was not kept returned at once)
    if (md && !replayed && status == NB_OK) {
      nb_md_kick(h, false);
      HIPCHECK(hipStreamSynchronize(h->stream));
    }
    // record a graph for the next evaluations: the PCG budget is this
    // evaluation's iteration count plus one, and so is the budget of each
    // FP64 refinement correction it took.  A replay whose FP64 residual then
    // fails with no correction left is repeated on the regular path (which
    // takes the corrections and records them); no spare correction is
    // recorded, since its early-returning launches cost about as much as
    // two PCG iterations on every replay
    if (h->graph_on && !replayed && status == NB_OK && !h->profiling && !h->used_sor && !h->want_finite) {
      int first = e.pcg_iters;
      const bool refining = mutual && h->refine && h->prm.precision == NB_PREC_MIXED;
      h->graph_refine_n = 0;
      if (refining) {
        for (int c = 0; c < h->refine_cycles; c++) {
          h->graph_refine_iters[c] = h->refine_iters[c] + 1;
          first -= h->refine_iters[c];
        }
        h->graph_refine_n = h->refine_cycles;
      }
      h->refine_idle = 0;
      capture_graph(h, flags, mutual ? first + 1 : 0);
    }
    if (shrink_graph) drop_graph(h);
    h->last = e;
    if (out) *out = e;
    if (status != NB_OK) {
      const char *msg[] = {"", "NaN energy", "SCF convergence failure (residual increased)",
                           "maximum SCF iterations reached", "undefined torque"};
      h->err = status < 5 ? msg[status] : "error";
    }
    return status;
  } catch (std::exception &ex) {
    h->capturing = false;
    drop_graph(h);
    return fail(h, ex, NB_ERR_HIP);
  }
}

// With electrostatic alchemistry the energy is the sum of the condensed phase
// (this handle) and the no-ligand and vapor phases (its two child handles),
// ParticleMeshEwald.java:894-935; gradients, dU/dX/dL and the virial add up.
int nb_evaluate(void *handle, uint32_t flags, nb_energies *out) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  if (flags & NB_FLAG_MD_STEP) {
    h->err = "unknown evaluation flag";
    return NB_ERR_INVALID;
  }
  if (!h->elec_lambda) return nb_evaluate_one(h, flags, out);
  if (!h->cell_set || !h->positions_set) {
    h->err = "nb_set_cell and nb_set_positions must be called before nb_evaluate";
    return NB_ERR_INVALID;
  }
  try {
    HIPCHECK(hipSetDevice(h->device));
    set_ost_phases(h);
    nb_energies e;
    int status = nb_evaluate_one(h, flags, &e);
    if (status != NB_OK) return status;
    uint32_t cf = flags & (NB_GRAD | NB_VIRIAL | NB_DUDL | NB_FINITE_PCG);
    NbHandle *kids[2] = {h->child_env, h->child_vap};
    for (int c = 0; c < 2; c++) {
      NbHandle *k = kids[c];
      if (!k) continue;
      int m = k->n;
      if (c == 0) {
        HIPCHECK(hipMemcpyAsync(k->d_xyz, h->d_xyz, sizeof(double) * 3 * m, hipMemcpyDeviceToDevice, k->stream));
      } else {
        hipLaunchKernelGGL(k_gather_mapped, dim3((3 * m + 255) / 256), dim3(256), 0, k->stream, m, h->d_vap_map,
                           h->d_xyz, k->d_xyz);
        HIPCHECK(hipGetLastError());
      }
      k->positions_set = true;
      nb_energies ek;
      int st = nb_evaluate_one(k, cf, &ek);
      if (st != NB_OK) {
        h->err = std::string(c ? "vapor phase: " : "no-ligand phase: ") + k->err;
        return st;
      }
      e.perm_real += ek.perm_real;
      e.perm_self += ek.perm_self;
      e.perm_recip += ek.perm_recip;
      e.perm_charge_corr += ek.perm_charge_corr;
      e.pol_real += ek.pol_real;
      e.pol_self += ek.pol_self;
      e.pol_recip += ek.pol_recip;
      e.dudl += ek.dudl;
      e.d2udl2 += ek.d2udl2;
      const int32_t *map = c ? h->d_vap_map : nullptr;
      dim3 G((3 * m + 255) / 256);
      hipLaunchKernelGGL(k_add_mapped, G, dim3(256), 0, h->stream, m, map, (const double *)k->d_grad, h->d_grad);
      if (flags & NB_DUDL)
        hipLaunchKernelGGL(k_add_mapped, G, dim3(256), 0, h->stream, m, map, (const double *)k->d_dudx, h->d_dudx);
      HIPCHECK(hipGetLastError());
      if (flags & NB_VIRIAL)
        for (int q = 0; q < 9; q++) h->vir[q] += k->vir[q];
    }
    HIPCHECK(hipStreamSynchronize(h->stream));
    e.perm = e.perm_real + e.perm_self + e.perm_recip + e.perm_charge_corr;
    e.pol = e.pol_real + e.pol_self + e.pol_recip;
    e.total = e.perm + e.pol + e.vdw;
    for (int k = 0; k < 8; k++) e.total += e.valence[k];
    if (!std::isfinite(e.total)) {
      h->err = "NaN energy";
      status = NB_ERR_NAN_ENERGY;
    }
    h->last = e;
    if (out) *out = e;
    return status;
  } catch (std::exception &ex) {
    return fail(h, ex, NB_ERR_HIP);
  }
}

// Evaluate several independent systems (one handle each) together.  When
// every handle qualifies (small periodic systems with the same parameters,
// no alchemistry, virial or SOR), one persistent launch of the batch kernel
// evaluates them all; a handle whose solve failed there is evaluated again on
// the regular path, which applies the SOR fallback.  Otherwise, or with
// NB_BATCH_SEPARATE, the recorded graphs of the handles are launched
// together, each on its own stream, then each is completed (falling back to
// the regular path where its replay is not valid); handles without a valid
// graph, and alchemical ones, are evaluated one after the other.
int nb_evaluate_batch(void *const *handles, int count, uint32_t flags, nb_energies *out) {
  if (!handles || count < 0) return NB_ERR_INVALID;
  for (int b = 0; b < count; b++)
    if (!handles[b]) return NB_ERR_INVALID;
  bool separate_all = (flags & NB_BATCH_SEPARATE) != 0;
  flags &= ~(uint32_t)NB_BATCH_SEPARATE;
  int status = NB_OK;
  if (!separate_all && count > 0) {
    std::vector<char> again;
    int st = NB_OK;
    bool done;
    try {
      done = nb_batch_evaluate((NbHandle **)handles, count, flags, out, again, st);
    } catch (std::exception &ex) {
      return fail((NbHandle *)handles[0], ex, NB_ERR_HIP);
    }
    if (done) {
      status = st;
      for (int b = 0; b < count; b++) {
        if (!again[b]) continue;
        int s2 = nb_evaluate((void *)handles[b], flags, out ? out + b : nullptr);
        if (s2 != NB_OK && status == NB_OK) status = s2;
      }
      return status;
    }
  }
  std::vector<char> launched(count, 0);
  for (int b = 0; b < count; b++) {
    NbHandle *h = (NbHandle *)handles[b];
    if (h->elec_lambda || !h->graph_on || !h->gexec || h->graph_flags != flags || h->profiling || !h->cell_set ||
        !h->positions_set)
      continue;
    if (hipSetDevice(h->device) != hipSuccess || hipGraphLaunch(h->gexec, h->stream) != hipSuccess) {
      h->err = "graph launch failed";
      return NB_ERR_HIP;
    }
    h->graph_launched = true;
    launched[b] = 1;
  }
  for (int b = 0; b < count; b++) {
    int st = nb_evaluate((void *)handles[b], flags, out ? out + b : nullptr);
    if (st != NB_OK && status == NB_OK) status = st;
  }
  return status;
}

int nb_evaluate_batch_static(void *const *handles, int count, uint32_t flags, nb_energies *out) {
  if (!handles || count <= 0) return NB_ERR_INVALID;
  for (int b = 0; b < count; b++)
    if (!handles[b]) return NB_ERR_INVALID;
  NbHandle *h0 = (NbHandle *)handles[0];
  if (flags & ~(uint32_t)NB_GRAD) {
    h0->err = "nb_evaluate_batch_static: flags may hold at most NB_GRAD";
    return NB_ERR_INVALID;
  }
  try {
    return nb_batch_evaluate_static((NbHandle **)handles, count, flags, out);
  } catch (std::exception &ex) {
    return fail(h0, ex, NB_ERR_HIP);
  }
}

int nb_batch_config(int device, int32_t cfg[4], double *kernel_ms) {
  if (!cfg) return NB_ERR_INVALID;
  try {
    return nb_batch_last_config(device, cfg, kernel_ms);
  } catch (std::exception &) {
    return NB_ERR_HIP;
  }
}

int nb_get_gradient(void *handle, double *g) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || !g) return NB_ERR_INVALID;
  try {
    HIPCHECK(hipMemcpyAsync(g, h->d_grad, sizeof(double) * 3 * h->n, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    return NB_OK;
  } catch (std::exception &e) {
    return fail(h, e, NB_ERR_HIP);
  }
}

const void *nb_gradient_device(void *handle) {
  NbHandle *h = (NbHandle *)handle;
  return h ? h->d_grad : nullptr;
}

int nb_get_induced(void *handle, double *u, double *ucr) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  try {
    int n = h->n;
    std::vector<double> us(3 * n), ups(3 * n);
    std::vector<int32_t> perm(n);
    HIPCHECK(hipMemcpyAsync(us.data(), h->d_u, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipMemcpyAsync(ups.data(), h->d_up, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipMemcpyAsync(perm.data(), h->d_perm, sizeof(int32_t) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    if (h->induced_caller)
      for (int s = 0; s < n; s++) perm[s] = s;   // the batch kernel works in caller order
    for (int s = 0; s < n; s++) {
      int i = perm[s];
      for (int d = 0; d < 3; d++) {
        if (u) u[3 * i + d] = us[3 * s + d];
        if (ucr) ucr[3 * i + d] = ups[3 * s + d];
      }
    }
    return NB_OK;
  } catch (std::exception &e) {
    return fail(h, e, NB_ERR_HIP);
  }
}

int nb_get_virial(void *handle, double v[9]) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || !v) return NB_ERR_INVALID;
  if (!h->vir_valid) {
    h->err = "the last evaluation did not compute the virial (pass NB_VIRIAL)";
    return NB_ERR_INVALID;
  }
  for (int q = 0; q < 9; q++) v[q] = h->vir[q];
  return NB_OK;
}

int nb_get_dudx_dl(void *handle, double *g) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || !g) return NB_ERR_INVALID;
  if (!h->dudl_valid) {
    h->err = "the last evaluation did not compute lambda derivatives (pass NB_DUDL)";
    return NB_ERR_INVALID;
  }
  try {
    HIPCHECK(hipSetDevice(h->device));
    HIPCHECK(hipMemcpyAsync(g, h->d_dudx, sizeof(double) * 3 * h->n, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    return NB_OK;
  } catch (std::exception &e) {
    return fail(h, e, NB_ERR_HIP);
  }
}

const char *nb_last_error(void *handle) {
  NbHandle *h = (NbHandle *)handle;
  return h ? h->err.c_str() : "invalid handle";
}

void nb_destroy(void *handle) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return;
  hipSetDevice(h->device);
  free_all(h);
  delete h;
}

}  // extern "C"

// one velocity-Verlet step of nb_md_step (h->md_dt set): drift, evaluation
// and kick, recorded and replayed as one graph in graph mode
int nb_evaluate_md_step(NbHandle *h, nb_energies *out) {
  return nb_evaluate_one(h, NB_GRAD | NB_FLAG_MD_STEP, out);
}
