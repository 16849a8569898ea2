// nb_state.h -- device state of one AMOEBA nonbonded handle.
//
// Layout in HBM (all per-atom arrays in *sorted* order unless noted):
//   perm[s] / iperm[i]       sorted position <-> caller atom index
//   pos4[s]                  double4 x, y, z, 0 (Cartesian, caller coordinates)
//   mp[s]                    AtomMP: q, d, Q/3 (global frame), 1/pdamp, thole
//   polar[s], group[s]       polarizability, polarization group
//   nbr[s * maxnb + t]       full Verlet list, one row per atom (lanes of an
//                            atom read consecutive entries), k_sorted | code << 27
//   ew / pc [s * max + t]    Ewald (r <= off) and preconditioner (r < 4.5) rows
//   grid                     interleaved complex PME grid (x fastest)
// Gradients are returned in caller order.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <stdint.h>
#include <string>
#include <map>
#include <vector>

#include "../../include/amoeba_nb.h"
#include "nb_device.h"

#define NB_CODE_SHIFT 27
#define NB_INDEX_MASK ((1u << NB_CODE_SHIFT) - 1u)

struct MaskScales { double m[32], p[32], d[32], v[32]; };

// Scales of one electrostatics evaluation.  A plain evaluation has every
// scale 1 and no lambda terms.  Under OST alchemistry the energy is the sum
// of three evaluations (ParticleMeshEwald.java:894-935): permanent real-space
// pairs with a soft atom are softcore-shifted (r -> sqrt(r^2 + lalpha) in the
// radial functions) and scaled by l2 (RealSpaceEnergyRegion.java:875-908);
// reciprocal and self terms are scaled by perm, all polarization terms by pol
// (ReciprocalEnergyRegion.java:254, 390-391, 555-556); d* / d2* are their
// first and second lambda derivatives, sign included.
struct ElecPhase {
  int perm_real = 1;     // permanent real-space energy on
  int polar = 1;         // polarization on
  int sor = 0;           // solve the dipoles by SOR (the vapor phase, ParticleMeshEwald.java:1744-1746)
  int soft_pairs = 0;    // pairs with a soft atom use l2 and the softcore shift
  int vapor = 0;         // leave out both-soft pairs per the inter / intramolecular flags
  int inter = 0, intra = 0;
  double l2 = 1.0, lalpha = 0.0, dlalpha = 0.0, d2lalpha = 0.0;
  double perm = 1.0, dperm = 0.0, d2perm = 0.0;
  double pol = 1.0, dpol = 0.0, d2pol = 0.0;
};

struct NbHandle {
  nb_params prm;
  int n = 0;
  int device = 0;
  hipStream_t stream = nullptr;   // main stream (PME chain, PCG vectors, bookkeeping)
  hipStream_t stream2 = nullptr;  // side stream (real-space pair kernels that overlap the PME chain)
  hipStream_t stream3 = nullptr;  // third stream (vdW and valence terms, joined at the gradient assembly)
  hipStream_t cur = nullptr;      // stream that NB_LAUNCH uses
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_eps = nullptr, ev_fork3 = nullptr, ev_join3 = nullptr;
  std::string err;

  // topology (caller order)
  int32_t *d_frame = nullptr, *d_axis = nullptr, *d_group0 = nullptr, *d_vclass0 = nullptr, *d_redidx = nullptr;
  double *d_mpole_local = nullptr, *d_polar0 = nullptr, *d_thole0 = nullptr, *d_ipd0 = nullptr, *d_redval = nullptr;
  int32_t *d_mask_off = nullptr, *d_mask_atom = nullptr, *d_mask_rel = nullptr;
  int n_class = 0;
  double *d_vdw_table = nullptr;
  int32_t *d_soft0 = nullptr, *d_mol0 = nullptr;   // caller order (soft: null when no atom is soft)
  int32_t *d_soft = nullptr, *d_mol = nullptr;     // sorted order
  bool has_soft = false;
  double lambda = 1.0;
  double *d_gred_l = nullptr;     // [3n] dU/dX/dlambda on reduced vdW sites (sorted)
  double *d_dudx = nullptr;       // [3n] dU/dX/dlambda, caller order
  bool want_dudl = false, dudl_valid = false;
  MaskScales scales;

  // electrostatic alchemistry (OST): this handle's phase scales, the lambda
  // gradient and torque of its electrostatics (sorted), and the two handles
  // that evaluate the no-ligand and vapor phases for it
  ElecPhase ph;
  bool ost = false;               // this handle's electrostatics carry lambda terms
  bool elec_lambda = false;       // this handle composes the three OST phases
  double *d_lgrad_s = nullptr, *d_ltorque_s = nullptr, *d_lgrad_r = nullptr, *d_ltorque_r = nullptr;
  NbHandle *child_env = nullptr, *child_vap = nullptr;
  int32_t *d_vap_map = nullptr;   // vapor-handle atom -> atom of this handle (caller order)
  int n_vap = 0;

  // cell
  CellD cell;
  double box[6];
  bool cell_set = false;

  // coordinates
  double *d_xyz = nullptr;        // caller order, interleaved
  double *d_xyz_list = nullptr;   // coordinates at the last list build
  bool positions_set = false;

  // sorting / lists
  int32_t *d_perm = nullptr, *d_iperm = nullptr, *d_cellid = nullptr, *d_cellcount = nullptr, *d_cellstart = nullptr;
  int ncell[3] = {1, 1, 1};
  int ncell_tot = 0, ncell_cap = 0;
  uint32_t *d_nbr = nullptr;  int maxnb = 0;
  uint32_t *d_ew = nullptr;   int maxew = 0;
  uint32_t *d_pc = nullptr;   int maxpc = 0;
  uint32_t *d_sp = nullptr;   int maxsp = 0;   // masked (bonded / same-group) pairs within the Ewald cutoff
  int32_t *d_nnb = nullptr, *d_new = nullptr, *d_npc = nullptr, *d_nsp = nullptr;
  int32_t *d_flags = nullptr;     // [0] overflow nbr, [1] overflow ew, [2] overflow pc, [3] rebuild needed
  bool list_valid = false;

  // sorted per-atom data
  double4 *d_pos4 = nullptr;
  int32_t *d_wrap = nullptr;      // [3n] lattice translation subtracted from each sorted atom (image sums)
  double4 *d_red4 = nullptr;      // reduced vdW sites
  AtomMP *d_mp = nullptr;
  double *d_gmp = nullptr;        // [10n] global multipoles (full Q), sorted
  double *d_polar = nullptr;
  int32_t *d_group = nullptr, *d_vclass = nullptr;

  // PME
  int K[3] = {0, 0, 0};
  long ntot = 0;
  double *d_grid = nullptr;       // interleaved complex [2 ntot] (FP64 form) or float2 [ntot] (mixed form)
  double *d_recip = nullptr;      // influence function [ntot]
  float *d_recipf = nullptr;      // influence function, FP32 copy (mixed form)
  float *d_twf = nullptr;         // twiddle tables, FP32 copy
  // grid bricks: bricks along axis d cover [b K_d / nbk_d, (b+1) K_d / nbk_d);
  // each atom is binned by its first B-spline point; the spread workgroup of a
  // brick sums the stencils of its bin into the brick and its 4-point spill
  // (edge bext), and the x transform assembles every point from its brick
  // and the spills of the previous bricks
  int nbk[3] = {0, 0, 0};
  int nbricks = 0, bext = 0;
  double *d_bbuf = nullptr;       // [nbricks][2][bext^3] brick buffer (grid real type)
  int32_t *d_bcount = nullptr;    // [nbricks] atoms per brick
  int32_t *d_bstart = nullptr;    // [nbricks + 1] exclusive scan
  int32_t *d_bid = nullptr;       // [n] brick of each sorted atom
  int32_t *d_brank = nullptr;     // [n] rank of the atom within its brick
  int32_t *d_batom = nullptr;     // [n] sorted atoms grouped by brick
  size_t lds_max = 65536;         // dynamic LDS available to one workgroup
  double *d_theta = nullptr;      // [n][3][5][4]
  int32_t *d_igrid = nullptr;     // [n][3]
  double *d_fmp = nullptr;        // [10n] fractional multipoles
  double *d_phip = nullptr;       // [30n] permanent phi: 20 fractional + 10 Cartesian
  double *d_phiu = nullptr;       // [30n] induced u phi
  double *d_phiup = nullptr;      // [30n] induced uCR phi
  double *d_tw = nullptr;         // twiddle tables
  double afrac[3][3];             // a[f][c] = K_f A[c][f]
  bool recip_valid = false;

  // fields / dipoles / PCG vectors (sorted, [3n] each, xyz interleaved)
  double *d_field = nullptr, *d_fieldp = nullptr;
  double *d_u = nullptr, *d_up = nullptr, *d_udir = nullptr, *d_updir = nullptr;
  double *d_r = nullptr, *d_rp = nullptr, *d_z = nullptr, *d_zp = nullptr, *d_p = nullptr, *d_pp = nullptr;
  double *d_vec = nullptr, *d_vecp = nullptr;
  double *d_scal = nullptr;       // PCG scalars
  double *d_dip = nullptr;        // [n] packed dipole records of the Ewald field kernel (96 B each)
  double *d_dip2 = nullptr;       // [n] packed records of polar * r, read by the preconditioner field
  // FP64 iterative refinement of the mixed form's solve: FP64 records of the
  // dipoles and FP64 B-spline weights for the FP64 residual
  bool refine = false;
  double *d_dip64 = nullptr;      // [n] DipRec<double>
  double *d_theta64 = nullptr;    // [n][3][5][4]
  int refine_cycles = 0;          // corrections taken by the last solve
  int refine_iters[4] = {0, 0, 0, 0};  // PCG iterations of each of those corrections
  double *d_part = nullptr;       // [S_NSLOT][part_stride] per-block partial dot products
  // strain derivative (NB_VIRIAL): each contributing kernel writes 9 partial
  // sums per block into its own region of d_vpart; k_vir_reduce sums them
  double *d_vpart = nullptr;
  long vpart_cap = 0, vpart_used = 0;
  double *d_vir = nullptr;        // [9]
  double *d_fgrid = nullptr;      // forward transform of the permanent grid (virial)
  bool want_vir = false, vir_valid = false;
  double vir[9];
  int part_stride = 0;            // partial sums per slot: the largest block count of their producers

  // outputs
  double *d_grad_s = nullptr;     // [3n] sorted direct gradient
  double *d_torque_s = nullptr;   // [3n]
  double *d_grad_r = nullptr;     // [3n] reciprocal-space gradient (written concurrently with d_grad_s)
  double *d_torque_r = nullptr;   // [3n] reciprocal-space torque
  double *d_gred_s = nullptr;     // [3n] vdW gradient on reduced sites
  double *d_grad = nullptr;       // [3n] caller order
  double *d_energy = nullptr;     // per-block energy partials
  int64_t *d_count = nullptr;
  int energy_blocks = 0;
  double *h_pinned = nullptr;
  double *d_pin = nullptr;        // [P_NSLOT] the evaluation's outputs in h_pinned's layout (one transfer)

  nb_energies last;
  int fmaxa = 0, fmaxb = 0;       // row capacities of the batch kernel (preconditioner part, the rest)
  // gradient of the truncated PCG (NB_FINITE_PCG)
  bool want_finite = false;       // this evaluation asks for it
  bool tape_on = false;           // the PCG records its iterations
  bool finite_active = false;     // the energy kernels use the adjoint terms
  double *d_tape = nullptr;       // [tape_cap][4][3n] r, z, p, q of each iteration (sorted)
  double *d_tape_sc = nullptr;    // [tape_cap][2] r.z, p.q
  int tape_cap = 0, tape_K = 0;
  double *d_fin = nullptr;        // adjoint work vectors and tensor pairs
  size_t fin_cap = 0;
  double *d_gfin = nullptr;       // [3n] tensor-derivative gradient terms (sorted)
  double *d_vfin = nullptr;       // [vfin_cap][9] strain-derivative partial sums of those terms
  long vfin_cap = 0;
  bool induced_caller = false;    // d_u / d_up hold caller order (written by the batch kernel)

  // valence terms (nb_set_valence): per-term gradient contributions go to
  // slots, each atom sums its slots in a fixed order (no atomics)
  bool has_valence = false;
  long val_terms = 0, val_slots = 0;
  void *val_dev = nullptr;        // device-side term table (VDev, nb_valence.hip)
  int32_t *d_val_i = nullptr;     // atom indices and grid integers
  double *d_val_d = nullptr;      // parameters and grids
  double *d_val_slot = nullptr;   // [3 val_slots]
  int32_t *d_val_off = nullptr;   // [n + 1] slots of each atom
  int32_t *d_val_list = nullptr;  // [val_slots]
  double *d_val_eterm = nullptr;  // [val_terms]
  double *d_val_energy = nullptr; // [8]

  // SCF predictor (nb_set_scf_predictor) and dynamics (nb_md_*)
  int predictor = 0;              // NB_PRED_*
  double *d_pred = nullptr;       // [6][6n] u - u_direct history (u then uCR), caller order
  int32_t *d_pred_ctl = nullptr;  // [2] ring start, fill count (advanced on the device)
  double *d_pred_coef = nullptr;  // [6] coefficients of the prediction (ASPC constants or the LS fit)
  double *d_pred_part = nullptr;  // [21][NB_PRED_BLOCKS] Gram-matrix partial sums of the LS fit
  double *d_md_v = nullptr, *d_md_a = nullptr, *d_md_mass = nullptr, *d_md_part = nullptr, *d_md_ke = nullptr;
  bool md_ready = false, md_accel = false;
  int md_thermo = 0;                 // NB_THERMO_*
  int md_active = 0;                 // atoms with mass > 0 (degrees of freedom 3 md_active - 3)
  double md_t0 = 0.0, md_tau = 0.0;  // K, ps
  // a recorded MD step (graph flags with NB_FLAG_MD_STEP): drift, evaluation,
  // kick and thermostat with these settings
  double md_dt = 0.0;                // ps, of the step being taken
  double md_graph_dt = 0.0;
  int md_graph_thermo = 0;
  double md_graph_t0 = 0.0, md_graph_tau = 0.0;

  // graph mode (nb_set_graph): the evaluation sequence is recorded once and
  // replayed; the PCG then runs graph_iters iterations under device-side
  // control (kernels that only serve a stopped solve return at once)
  bool graph_on = false, capturing = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  uint32_t graph_flags = 0;
  int graph_iters = 0;
  // recorded FP64 refinement: corrections and the PCG budget of each
  int graph_refine_n = 0;
  int refine_idle = 0;            // replays in a row that skipped a recorded correction
  int graph_refine_iters[4] = {0, 0, 0, 0};
  long graph_replays = 0;
  bool graph_launched = false;    // the graph of this evaluation was launched by nb_evaluate_batch
  bool used_sor = false;          // the last evaluation's dipoles came from SOR (not recorded)
  const double *ctl = nullptr;    // PCG control slots while the captured solve is being recorded

  // profiling
  bool profiling = false;
  std::string prof_filter;  // when set, only launches whose kernel name contains it are timed
  struct ProfRec { std::string name; hipEvent_t a, b; };
  std::vector<ProfRec> prof_pool;
  int prof_n = 0;
  std::vector<std::string> prof_names;
  std::vector<double> prof_ms;
  std::vector<long long> prof_count;
};

// launch helpers implemented in the kernel files
void nb_build_lists(NbHandle *h);
bool nb_need_rebuild(NbHandle *h);
void nb_gather_sorted(NbHandle *h);
void nb_rotate(NbHandle *h);
void nb_pme_setup(NbHandle *h);
void nb_pme_permanent(NbHandle *h);
void nb_pme_induced(NbHandle *h, const double *u, const double *up, bool full);
void nb_pme_induced64(NbHandle *h, const double *u, const double *up);
void nb_dip_field64(NbHandle *h, double *f, double *fp);
void nb_filter_lists(NbHandle *h);
bool nb_grow_filtered_rows(NbHandle *h, const int32_t *fl);
void nb_perm_field(NbHandle *h);
void nb_direct(NbHandle *h);
void nb_induced_field(NbHandle *h, const double *u, const double *up, double *f, double *fp, bool full);
void nb_precond_field(NbHandle *h, double *f, double *fp);
int nb_pcg(NbHandle *h, nb_energies *out);
int nb_pcg_device(NbHandle *h, int iters);
int nb_sor(NbHandle *h, double w, nb_energies *out);
void nb_reset_dipoles(NbHandle *h);
void nb_check_rebuild_async(NbHandle *h);
void nb_recip_energy(NbHandle *h, bool polar);
// passes of nb_real_energy
enum { NB_RE_ALL = 0, NB_RE_PERM = 1, NB_RE_POL = 2 };
void nb_real_energy(NbHandle *h, bool polar, int which);
void nb_vdw(NbHandle *h);
void nb_finalize_gradient(NbHandle *h);
void nb_zero_outputs(NbHandle *h);
void nb_pack_outputs(NbHandle *h, bool scal);
double *nb_vir_region(NbHandle *h, long nblocks);
// rows of d_vpart kept at its end for the first pass of nb_vir_reduce
constexpr long NB_VIR_STAGE = 512;
void nb_grid_virial(NbHandle *h, bool polar);
// grid part of the reciprocal strain derivative of scale * (a . T_rec b)
// terms (the mutual term's form), 9 partial sums per block into vpart
long nb_grid_virial_blocks(const NbHandle *h);
void nb_grid_virial_pair(NbHandle *h, const double *a, const double *b, double scale, double *vpart);
void nb_vir_reduce(NbHandle *h);
void nb_tape_push(NbHandle *h, int nrz, int npq);
void nb_pred_apply(NbHandle *h);
void nb_pred_save(NbHandle *h, bool device_ctl);
void nb_md_free(NbHandle *h);
// internal evaluation flag: the evaluation is one velocity-Verlet step (drift
// before, kick and thermostat after, all in the recorded graph)
constexpr uint32_t NB_FLAG_MD_STEP = 1u << 30;
constexpr int NB_PRED_BLOCKS = 256;
void nb_md_drift(NbHandle *h);
void nb_md_kick(NbHandle *h, bool guarded);
int nb_evaluate_md_step(NbHandle *h, nb_energies *out);
void nb_valence_terms(NbHandle *h);
void nb_valence_gather(NbHandle *h);
void nb_valence_free(NbHandle *h);
void nb_finite_adjoint(NbHandle *h);
void nb_finite_add(NbHandle *h);
bool nb_batch_evaluate(NbHandle **hv, int count, uint32_t flags, nb_energies *out, std::vector<char> &separate,
                       int &status);
int nb_batch_last_config(int device, int32_t cfg[4], double *kernel_ms);
int nb_batch_evaluate_static(NbHandle **hv, int count, uint32_t flags, nb_energies *out);

#define HIPCHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Optional per-launch timing (nb_set_profiling): the launch carries a start
// and a stop event that the runtime records with the kernel's own dispatch
// (hipExtLaunchKernelGGL), so the measured time is the kernel's, without the
// launch gaps that separately recorded events would add; harvested after the
// evaluation's final synchronisation.
int nb_prof_slot(NbHandle *h, const char *name);
static inline bool nb_prof_wanted(const NbHandle *h, const char *name) {
  return h->profiling && (h->prof_filter.empty() || h->prof_filter.find(name) != std::string::npos ||
                          std::string(name).find(h->prof_filter) != std::string::npos);
}
void nb_prof_harvest(NbHandle *h);

#define NB_LAUNCH(KERNEL, G, B, SH, ...)                                                              \
  do {                                                                                                \
    if (nb_prof_wanted(h, #KERNEL)) {                                                                 \
      const int slot_ = nb_prof_slot(h, #KERNEL);                                                     \
      hipExtLaunchKernelGGL(KERNEL, G, B, SH, h->cur, h->prof_pool[slot_].a, h->prof_pool[slot_].b, 0, \
                            __VA_ARGS__);                                                             \
    } else {                                                                                          \
      hipLaunchKernelGGL(KERNEL, G, B, SH, h->cur, __VA_ARGS__);                                      \
    }                                                                                                 \
  } while (0)

// Fork / join between the main and the side stream.  nb_fork makes the side
// stream wait for everything issued so far on the main stream; nb_join makes
// the main stream wait for everything issued so far on the side stream.
// Inside an NbSide scope, NB_LAUNCH targets the side stream.
static inline void nb_fork(NbHandle *h) {
  HIPCHECK(hipEventRecord(h->ev_fork, h->stream));
  HIPCHECK(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
}
static inline void nb_join(NbHandle *h) {
  HIPCHECK(hipEventRecord(h->ev_join, h->stream2));
  HIPCHECK(hipStreamWaitEvent(h->stream, h->ev_join, 0));
}
struct NbSide {
  NbHandle *h;
  explicit NbSide(NbHandle *hh) : h(hh) { h->cur = h->stream2; }
  ~NbSide() { h->cur = h->stream; }
};
// The same for the third stream.
static inline void nb_fork3(NbHandle *h) {
  HIPCHECK(hipEventRecord(h->ev_fork3, h->stream));
  HIPCHECK(hipStreamWaitEvent(h->stream3, h->ev_fork3, 0));
}
static inline void nb_join3(NbHandle *h) {
  HIPCHECK(hipEventRecord(h->ev_join3, h->stream3));
  HIPCHECK(hipStreamWaitEvent(h->stream, h->ev_join3, 0));
}
struct NbThird {
  NbHandle *h;
  explicit NbThird(NbHandle *hh) : h(hh) { h->cur = h->stream3; }
  ~NbThird() { h->cur = h->stream; }
};
