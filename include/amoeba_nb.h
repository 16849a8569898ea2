/*
 * amoeba_nb.h -- C ABI of the MI355X AMOEBA nonbonded energy-and-gradient
 * library (libamoeba_nb.so, HIP kernels for gfx950).
 *
 * This is the drop-in boundary below Force Field X's ForceFieldEnergy /
 * ParticleMeshEwald / VanDerWaals objects.  Plain pointers and sizes only;
 * all host arrays are caller-owned and copied.  One device context per
 * handle; calls on one handle are not thread-safe, separate handles may be
 * used concurrently.
 *
 * Units follow the reference: Angstrom, e, kcal/mol, Debye for polar-eps.
 * Coordinates and gradients are interleaved x,y,z per atom, exactly the
 * double[] x / double[] g of Potential.energyAndGradient
 * (modules/numerics/src/main/java/ffx/numerics/OptimizationInterface.java:91).
 */
#ifndef AMOEBA_NB_H
#define AMOEBA_NB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NB_ABI_VERSION 9

/* Return codes.  Map onto the reference's failure behaviour:
 * EnergyException on NaN energy (RealSpaceEnergyRegion.java:384-391),
 * SCF divergence / iteration cap (PCGSolver.java:553-569), undefined torque
 * (ReduceRegion.java:208-229). */
enum {
  NB_OK = 0,
  NB_ERR_NAN_ENERGY = 1,
  NB_ERR_SCF_DIVERGED = 2,
  NB_ERR_SCF_MAXITER = 3,
  NB_ERR_BAD_TORQUE = 4,
  NB_ERR_HIP = 5,
  NB_ERR_INVALID = 6,
  NB_ERR_NO_DEVICE = 7
};

/* Local frame definitions (MultipoleType.MultipoleFrameDefinition order). */
enum { NB_FRAME_NONE = 0, NB_FRAME_ZONLY = 1, NB_FRAME_ZTHENX = 2, NB_FRAME_BISECTOR = 3,
       NB_FRAME_ZTHENBISECTOR = 4, NB_FRAME_THREEFOLD = 5 };

/* Polarization (Polarization.java) */
enum { NB_POL_NONE = 0, NB_POL_DIRECT = 1, NB_POL_MUTUAL = 2 };

/* Arithmetic form. */
enum { NB_PREC_FP64 = 0, NB_PREC_MIXED = 1 };

/* nb_evaluate flags.  NB_FINITE_PCG (no counterpart in the reference, whose
 * gradient assumes stationary dipoles, RealSpaceEnergyRegion.java:1700-2177):
 * the gradient of the energy of the dipoles the PCG actually returned, with
 * its iteration count held fixed -- the adjoint of the recorded iterations
 * replaces uCR and the mutual terms.  Mutual polarization solved by PCG only
 * (no effect otherwise; a phase solved by SOR keeps the stationary form).
 * With NB_VIRIAL the strain derivative of the same energy; under OST phases
 * the polarization energy of each phase is its pol factor times the energy of
 * lambda-independent dipoles, so NB_DUDL and dU/dX/dL follow.  Evaluated on
 * the regular path (no graph replay).  After it, nb_get_induced returns the
 * adjoint dipoles as ucr. */
enum { NB_GRAD = 1u, NB_VIRIAL = 2u, NB_DUDL = 4u, NB_FINITE_PCG = 8u, NB_BATCH_SEPARATE = 16u };

/* SCF algorithms (SCFAlgorithm.java) */
enum { NB_SCF_PCG = 0, NB_SCF_SOR = 1 };

typedef struct {
  double ewald_cutoff;      /* real-space Ewald cutoff, A (EwaldParameters.java:55) */
  double ewald_alpha;       /* Ewald coefficient, 1/A (EwaldParameters.java:71) */
  double vdw_cutoff;        /* A (VanDerWaalsForm.java:85) */
  double vdw_taper;         /* taper start as a fraction of the cutoff (0.9) */
  double list_buffer;       /* neighbor-list buffer, A (ForceFieldEnergy.java:244) */
  int32_t pme_order;        /* B-spline order (5) */
  int32_t grid[3];          /* PME grid, each a product of 2, 3, 5 and 7 (Complex.java preferredDimension) */
  double polar_eps;         /* PCG convergence, RMS Debye (ParticleMeshEwald.java:171) */
  double electric;          /* 332.0637133 (Constants.java:192) */
  double precond_cutoff;    /* 4.5 A (PCGSolver.java:300) */
  double precond_scale;     /* 2.0 (PCGSolver.java:316) */
  double m_scale[4];        /* mpole-12..15-scale */
  double p_scale[4];        /* polar-12..15-scale */
  double intra14;           /* polar-14-intra */
  double d11;               /* direct-11-scale */
  double v_scale[3];        /* vdw-12..14-scale */
  double halgren_delta;     /* 0.07 */
  double halgren_gamma;     /* 0.12 */
  int32_t polarization;     /* NB_POL_* */
  int32_t do_vdw;
  int32_t do_pme;
  int32_t precision;        /* NB_PREC_* */
  int32_t max_scf_iter;     /* 1000 */
  int32_t device;           /* HIP device ordinal */
  /* alchemical vdW softcore (VanDerWaals.setLambda, VanDerWaals.java:599-653) */
  double lambda;            /* initial state, 1 = fully coupled (nb_set_lambda changes it) */
  double vdw_lambda_alpha;  /* softcore shift alpha (0.25) */
  double vdw_lambda_exponent; /* eps scaling exponent beta (3) */
  double vdw_lambda_end;    /* lambda at which the vdW coupling is complete (1) */
  int32_t intermolecular_softcore; /* both-soft pairs of different molecules are soft (0) */
  int32_t intramolecular_softcore; /* both-soft pairs of one molecule are soft (0) */
  /* electrostatic alchemistry in OST mode (AlchemicalParameters.java:171-251):
   * with elec_lambda set and soft atoms present, nb_evaluate sums the
   * condensed, no-ligand and vapor phases of ParticleMeshEwald.java:894-935 */
  int32_t elec_lambda;      /* 0: lambda applies to vdW only */
  double perm_lambda_alpha; /* permanent softcore shift (2) */
  double perm_lambda_exponent; /* (3) */
  double perm_lambda_start, perm_lambda_end; /* permanent window (0.4, 1) */
  double pol_lambda_exponent; /* (3) */
  double pol_lambda_start, pol_lambda_end; /* polarization window (0.75, 1) */
  int32_t ligand_vapor_elec; /* vapor phase of the soft atoms (1) */
  int32_t no_ligand_condensed_scf; /* polarization of the no-ligand phase (1) */
  /* SCF algorithm (ParticleMeshEwald.java:2247-2325) */
  int32_t scf_algorithm;    /* NB_SCF_PCG or NB_SCF_SOR */
  double polar_sor;         /* SOR factor, polar-sor (0.7; SORRegion.java:65-72) */
  int32_t sor_fallback;     /* SOR with factor 0.65 after a failed PCG (1; ParticleMeshEwald.java:563) */
} nb_params;

typedef struct {
  int32_t n_atoms;
  const int32_t *frame;        /* [n] NB_FRAME_* */
  const int32_t *axis;         /* [3n] z, x, y frame atoms, -1 when unused */
  const double *multipole;     /* [10n] local frame q, d(3), Q(xx,yy,zz,xy,xz,yz) in e, e*A, e*A^2 */
  const double *polarizability;/* [n] A^3 */
  const double *thole;         /* [n] */
  const double *pdamp;         /* [n] polarizability^(1/6); <= 0: undamped */
  const int32_t *group;        /* [n] polarization group id (ip11 equivalence class) */
  const int32_t *vdw_class;    /* [n] row of vdw_table */
  int32_t n_vdw_class;
  const double *vdw_table;     /* [n_vdw_class^2 * 2] combined (1/Rmin, eps) per class pair */
  const int32_t *red_index;    /* [n] heavy atom a reduced hydrogen rides on (self otherwise) */
  const double *red_value;     /* [n] reduction factor (0 = none) */
  const int32_t *mask_offset;  /* [n+1] CSR offsets of bonded relations */
  const int32_t *mask_atom;    /* [nnz] partner */
  const int32_t *mask_rel;     /* [nnz] 2 = 1-2, 3 = 1-3, 4 = 1-4, 5 = 1-5 */
  const int32_t *soft;         /* [n] alchemical atoms (Atom.applyLambda), NULL: none */
  const int32_t *molecule;     /* [n] molecule ids, needed with soft (NULL: all distinct) */
} nb_topology;

typedef struct {
  double perm_real, perm_self, perm_recip, perm_charge_corr;
  double pol_real, pol_self, pol_recip;
  double vdw;
  double perm, pol, total;
  double dudl, d2udl2;
  int32_t pcg_iters;
  double pcg_eps;
  int64_t n_vdw, n_perm;
  int32_t list_rebuilt;
  /* valence terms (nb_set_valence), NB_VAL_* order; total includes them */
  double valence[8];
} nb_energies;

/* Valence term kinds, the order of nb_energies.valence (the reference's
 * EnergyTest columns: bond, angle, stretch-bend, Urey-Bradley, out-of-plane
 * bend, torsion, pi-orbital torsion, torsion-torsion). */
enum { NB_VAL_BOND = 0, NB_VAL_ANGLE, NB_VAL_STRBND, NB_VAL_UREY, NB_VAL_OPBEND, NB_VAL_TORSION, NB_VAL_PITORS,
       NB_VAL_TORTOR };

/* SCF predictors (SCFPredictor.java) */
enum { NB_PRED_NONE = 0, NB_PRED_ASPC = 1, NB_PRED_LS = 2 };

/* Result of nb_md_step: the energies after the last step. */
typedef struct {
  double kinetic, potential, total;   /* kcal/mol */
  int32_t steps;                      /* steps taken */
  int32_t pcg_iters;                  /* PCG iterations summed over the steps */
  nb_energies last;                   /* the last step's evaluation */
} nb_md_report;

/* AMOEBA valence terms, assigned by the caller (atom indices 0-based; force
 * constants already multiplied by their unit keywords; ideal lengths in A,
 * ideal angles in degrees).  Forms follow Bond.java:242-283,
 * Angle.java:240-345, StretchBend.java:117-160, UreyBradley.java:123-153,
 * OutOfPlaneBend.java:172-250, Torsion.java:247-330,
 * PiOrbitalTorsion.java:145-240 and TorsionTorsion.java:285-430. */
typedef struct {
  int32_t n_bond;    const int32_t *bond;     /* [2 n] */
  const double *bond_k, *bond_r0;
  double bond_cubic, bond_quartic;
  int32_t n_angle;   const int32_t *angle;    /* [4 n] end, centre, end, projection atom of an in-plane angle or -1 */
  const double *angle_k, *angle_t0;
  double angle_poly[4];                       /* cubic, quartic, pentic, sextic */
  int32_t n_strbnd;  const int32_t *strbnd;   /* [3 n] */
  const double *strbnd_k, *strbnd_r0;         /* [2 n] each: first and second bond */
  const double *strbnd_t0;
  int32_t n_urey;    const int32_t *urey;     /* [2 n] the angle's end atoms */
  const double *urey_k, *urey_r0;
  double urey_cubic, urey_quartic;
  int32_t n_opbend;  const int32_t *opbend;   /* [4 n] end, trigonal centre, end, out-of-plane atom */
  const double *opbend_k;
  double opbend_poly[4];
  int32_t n_torsion; const int32_t *torsion;  /* [4 n] */
  const double *torsion_amp, *torsion_cos, *torsion_sin;  /* [6 n]: periodicity 1..6, cos / sin of the phase */
  int32_t n_pitors;  const int32_t *pitors;   /* [6 n] two neighbours of c, c, d, two neighbours of d */
  const double *pitors_k;
  int32_t n_tortor;  const int32_t *tortor;   /* [6 n] five chain atoms, chirality atom of the centre or -1 */
  const int32_t *tortor_grid;                 /* [n] grid of each term */
  int32_t n_grid;    const int32_t *grid_nx, *grid_ny;     /* [n_grid] */
  const int32_t *grid_off_t;                  /* [n_grid] offset into grid_t of tx (nx values) then ty (ny) */
  const int32_t *grid_off_e;                  /* [n_grid] offset into the value arrays (nx * ny each) */
  const double *grid_t;                       /* degrees, ascending */
  const double *grid_e, *grid_dx, *grid_dy, *grid_dxy;   /* e[j nx + i] at (tx[i], ty[j]); spline derivatives */
} nb_valence;

/* Build a handle: uploads topology, allocates device buffers.
 * Replaces the construction of ParticleMeshEwald/VanDerWaals
 * (ParticleMeshEwald.java:400-700, VanDerWaals.java:308-404). */
int nb_create(const nb_params *params, const nb_topology *top, void **handle);

/* Unit cell a, b, c (A), alpha, beta, gamma (deg); a <= 0: aperiodic.
 * Replaces ParticleMeshEwald.setCrystal (ParticleMeshEwald.java:1244-1283). */
int nb_set_cell(void *handle, const double box[6]);

/* Host coordinates [3n]; replaces the coordinate load of
 * InitializationRegion (InitializationRegion.java:316-321). */
int nb_set_positions(void *handle, const double *xyz);

/* Device-resident coordinates [3n] (copied device-to-device). */
int nb_set_positions_device(void *handle, const void *d_xyz);

/* Alchemical state in [0, 1]; replaces LambdaInterface.setLambda
 * (LambdaInterface.java:73).  Softcore applies to vdW pairs with exactly one
 * soft atom (and to both-soft pairs per the inter/intramolecular flags). */
int nb_set_lambda(void *handle, double lambda);

/* One energy (+ gradient) evaluation: ForceFieldEnergy.energy(gradient, print)
 * restricted to VanDerWaals.energy + ParticleMeshEwald.energy
 * (ForceFieldEnergy.java:1748-1755).  out may be NULL. */
int nb_evaluate(void *handle, uint32_t flags, nb_energies *out);

/* Gradient of the last evaluation [3n] (ForceFieldEnergy.fillGradient,
 * ForceFieldEnergy.java:2848-2884). */
int nb_get_gradient(void *handle, double *g);

/* Device pointer of the gradient [3n] doubles of the last evaluation. */
const void *nb_gradient_device(void *handle);

/* Induced dipoles u (direct-masked) and uCR (energy-masked) [3n] each;
 * ParticleMeshEwald.inducedDipole / inducedDipoleCR (ParticleMeshEwald.java:276). */
int nb_get_induced(void *handle, double *u, double *ucr);

/* Virial (3x3, row-major, kcal/mol); not provided by the reference. */
int nb_get_virial(void *handle, double v[9]);

/* dU/dX/dlambda [3n] of the last evaluation made with NB_DUDL (vdW softcore;
 * VanDerWaals.getdEdXdL and the lambdaGrad of VanDerWaals.java:1869-1881). */
int nb_get_dudx_dl(void *handle, double *g);

/* Valence terms evaluated with every later nb_evaluate, beside the
 * nonbonded terms: their energies in nb_energies.valence (and total), their
 * gradient and strain derivative added to the handle's.  Replaces the
 * bonded-term setup and region of ForceFieldEnergy (ForceFieldEnergy.java:
 * 935-1100, 1670-1677).  Arrays are copied; NULL or all counts zero removes
 * the terms.  Coordinates are used as given (no periodic images), as the
 * reference's bonded terms do.  NB_ERR_INVALID for an index out of range or
 * a periodicity table beyond 6 terms. */
int nb_set_valence(void *handle, const nb_valence *terms);

/* sizeof(nb_valence) as compiled. */
int nb_valence_struct_size(void);

/* Initial guess of the mutual solve from earlier solutions: NB_PRED_ASPC is
 * the always-stable predictor-corrector of SCFPredictorParameters.
 * aspcPredictor (SCFPredictorParameters.java:79-121; wired at
 * ParticleMeshEwald.java:2272-2281, saved at :1952-1955): after six converged
 * solves, the next starts from the direct dipoles plus sum_k c_k (u -
 * u_direct) of the last six, c = 22/7, -55/14, 55/21, -22/21, 5/21, -1/42, for
 * u and uCR alike.  NB_PRED_LS is the least-squares predictor
 * (SCFPredictorParameters.leastSquaresPredictor, :144-213, with the model of
 * LeastSquaresPredictor.value, :341-375): from two saved solves on, the
 * coefficients c_0..c_4 that best fit the newest (u - u_direct, uCR -
 * uCR_direct) by the five before it (in least squares over all 6n values,
 * the normal equations solved on the device; slots not yet filled and
 * columns that depend linearly on earlier ones within 1e-12 get c = 0) are
 * applied to the five newest, u = u_direct + sum_k c_k (u - u_direct)_(t-k).
 * The reference minimises the same sum of squares with a Levenberg-Marquardt
 * optimiser stopped at a relative change of 1e-4, so its coefficients are
 * this minimiser to that tolerance.  The history is kept per handle in
 * caller order and reset by this call; graph replays carry it.  Not applied
 * with NB_FINITE_PCG. */
int nb_set_scf_predictor(void *handle, int kind);

/* The predictor's state: coef [6] the coefficients of the last prediction
 * (the ASPC constants, or the last LS fit), history [6][2][n][3] the saved
 * (u - u_direct, uCR - uCR_direct) in caller order by ring slot, state [2]
 * the slot of the newest entry and the number of entries held (the
 * reference's predictorStartIndex and predictorCount,
 * SCFPredictorParameters.java:269-302).  Any pointer may be NULL. */
int nb_get_scf_predictor(void *handle, double *coef, double *history, int32_t *state);

/* Velocity-Verlet dynamics on the device (VelocityVerlet.java:66-113, driven
 * as MolecularDynamics.java:1285-1299): nb_md_init takes the masses [n]
 * (g/mol) and velocities [3n] (A/ps, NULL = zero); nb_md_step advances nsteps
 * steps of dt ps from the handle's current positions, one evaluation per step,
 * positions and velocities staying in device memory; nb_md_get_state copies
 * them out (either may be NULL).  The first step after nb_md_init evaluates
 * the starting accelerations.  In graph mode (nb_set_graph) a whole step --
 * drift, evaluation, kick, kinetic energy and thermostat -- is recorded as
 * one graph and replayed; the kick kernels run only when the evaluation is
 * valid (no list rebuild pending, no filtered row overflowed, the solve
 * converged within the recorded iterations), and a step whose replay is not
 * valid is finished on the regular path from its drifted positions.  Handles
 * with alchemical electrostatics (the OST environment and vapour handles)
 * step with a graph per evaluation. */
int nb_md_init(void *handle, const double *mass, const double *v0);
int nb_md_step(void *handle, int nsteps, double dt, nb_md_report *out);
int nb_md_get_state(void *handle, double *x, double *v);

/* Thermostat of nb_md_step (replaces the Berendsen class,
 * modules/algorithms/.../dynamics/thermostats/Berendsen.java:99-101, applied
 * after each full step as MolecularDynamics.java:1577-1583).  kind
 * NB_THERMO_BERENDSEN: after each step's second half kick the velocities are
 * scaled by sqrt(1 + dt/tau (T0/T - 1)), T = sum m v^2 / (kB (3m - 3)) over
 * the m atoms with mass > 0 (atoms with mass <= 0 do not move),
 * temperature T0 in K, tau in ps (dt <= tau, else nb_md_step returns
 * NB_ERR_INVALID); the reported kinetic energy is the scaled one.
 * NB_THERMO_NONE (the default) gives NVE. */
enum { NB_THERMO_NONE = 0, NB_THERMO_BERENDSEN = 1 };
int nb_md_set_thermostat(void *handle, int kind, double temperature, double tau);

/* Human-readable message for the last error on this handle. */
const char *nb_last_error(void *handle);

/* Release all device memory of the handle. */
void nb_destroy(void *handle);

/* Graph mode (no counterpart in the reference): with on != 0, an evaluation
 * records its launch sequence into a HIP graph, and later evaluations with
 * the same flags replay it.  The PCG of a replay runs under device-side
 * control with a budget of the recorded evaluation's iterations plus one;
 * a replay whose list needs rebuilding, whose filtered rows overflow or whose
 * solve did not stop within the budget is repeated on the regular path and
 * recorded again; so is one whose solve stopped two or more iterations short of
 * the budget (the surplus iterations' kernels still launch, and return at
 * once).
 * nb_set_cell, nb_set_lambda and list rebuilds drop the graph.
 * Results equal those of the regular path. */
int nb_set_graph(void *handle, int on);

/* Evaluate count independent systems, one handle each, with the same flags
 * (no counterpart in the reference, which evaluates one system at a time).
 * When every handle qualifies -- periodic, at most 4096 atoms, grids of at
 * most 64 per axis, the same parameters, flags of at most NB_GRAD, no
 * alchemistry, no SOR, no valence terms -- one persistent launch evaluates all of them: a
 * workgroup takes one system at a time (largest first) through the whole
 * evaluation, with explicit image sums in every cell.  A handle whose solve
 * fails there is evaluated again on the regular path (SOR fallback).  The
 * kernel's workgroup size and lanes per atom are chosen by timing every
 * candidate on the first batch of a given size (nb_batch_config reports the
 * choice).  Otherwise, or with NB_BATCH_SEPARATE in flags, handles in graph
 * mode with a recorded graph are launched together, each on its own stream,
 * then completed; the others are evaluated in turn.  out (may be NULL)
 * receives count records.  Returns the first non-OK status. */
int nb_evaluate_batch(void *const *handles, int count, uint32_t flags, nb_energies *out);

/* Allocation-free batch evaluation: the batch kernel of nb_evaluate_batch
 * with its workspace, system descriptors, outputs and queue counter in a
 * static device buffer of NB_BATCH_SCRATCH_MIB MiB that each GPU holds for the
 * life of the process, instead of device and pinned memory the library
 * allocates.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory during evaluation (code under HIP graph capture,
 * tight serving loops); the handles themselves are created beforehand.
 * Every handle must qualify for the batch kernel as in nb_evaluate_batch and
 * flags may hold at most NB_GRAD; the kernel's shape is fixed (256 threads,
 * 4 lanes per atom; no tuner).  NB_ERR_INVALID when a handle does not
 * qualify or the batch needs more than the static buffer (nb_last_error of
 * handles[0] says which); a system the batch kernel hands back (failed solve,
 * list cutoff beyond the checked image range) is not re-evaluated: its out
 * record is left untouched, its nb_last_error says so, and the call returns
 * NB_ERR_SCF_DIVERGED.  Calls on one GPU share its buffer and are serialised
 * within the process; calls that overlap from different host threads wait
 * for each other, and nothing else may use the buffer. */
#define NB_BATCH_SCRATCH_MIB 64
int nb_evaluate_batch_static(void *const *handles, int count, uint32_t flags, nb_energies *out);

/* The shape of the last batch-kernel launch on a device: cfg[0] threads per
 * workgroup, cfg[1] lanes per atom in the pair loops, cfg[2] persistent
 * workgroups, cfg[3] systems; kernel_ms (may be NULL) its duration from HIP
 * events around the launch (no counterpart in the reference). */
int nb_batch_config(int device, int32_t cfg[4], double *kernel_ms);

/* FP64 iterative refinement of the mixed form's induced-dipole solve (on by
 * default in NB_PREC_MIXED; no counterpart in the reference, whose solve is
 * FP64 throughout).  After the PCG with the FP32 field operator converges, the
 * residual (PCGSolver.java:609-663) is formed again with the FP64 operator
 * (FP64 real-space field and FP64 PME) and its eps (PCGSolver.java:543-547)
 * must be below polar_eps; otherwise the PCG restarts from the current dipoles
 * with the FP64 residual, up to three times.  The FP64 reciprocal potentials
 * of that last residual are the ones the energy and gradient use;
 * nb_energies.pcg_eps reports the FP64 eps and pcg_iters counts the restarts'
 * iterations too.  In graph mode the recorded evaluation carries the restarts
 * its solve took (at least one), each with its iteration count + 1 as budget;
 * a replay whose FP64 residual passes earlier skips the rest.
 * on = 0 turns it off (the FP32-operator solve stands); NB_ERR_INVALID for an
 * FP64 handle. */
int nb_set_refinement(void *handle, int on);

/* Number of evaluations of this handle served by a graph replay. */
int nb_graph_replays(void *handle);

/* ABI version (NB_ABI_VERSION). */
int nb_abi_version(void);

/* sizeof(nb_params), sizeof(nb_topology), sizeof(nb_energies) as compiled,
 * for bindings to check their struct mirrors (no counterpart in the
 * reference, whose interface is Java objects). */
int nb_struct_sizes(int32_t sizes[3]);

/* Number of HIP devices visible (0 without a GPU). */
int nb_device_count(void);

/* Kernel timing with HIP events on the handle's stream (no counterpart in the
 * reference, whose timers are PMETimings.java:47-120 and
 * ReciprocalSpace.printTimings :645-718).  When on, every launch of an
 * evaluation carries a start and a stop event that the runtime records with
 * the kernel's dispatch, so each time is the kernel's own; totals accumulate
 * per kernel name until profiling is switched again. */
int nb_set_profiling(void *handle, int on);
/* Restrict the timing to kernels whose name contains `substring` ("" = all). */
int nb_set_profile_filter(void *handle, const char *substring);
int nb_profile_count(void *handle);
int nb_profile_entry(void *handle, int idx, char *name, int cap, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
