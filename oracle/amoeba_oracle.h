/*
 * amoeba_oracle.h -- FP64 CPU restatement of Force Field X's AMOEBA nonbonded
 * energy-and-gradient path (PME permanent multipoles, PCG induced dipoles,
 * buffered-14-7 vdW).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * library (include/amoeba_nb.h).  It is called from tests/, from the
 * smoke() check and from the CPU-baseline leg of bench.py, never from the
 * product path.
 *
 * Units follow the reference: Angstrom, electron charges, kcal/mol.
 * Multipoles are given in the local frame, quadrupoles as full traceless
 * Cartesian tensors (no 1/3 factor), ordered
 *   q, dx, dy, dz, Qxx, Qyy, Qzz, Qxy, Qxz, Qyz
 * as in MultipoleType.java:109-145 (reference stores e*A and e*A^2 after the
 * Bohr conversion at MultipoleType.java:1116-1119).
 */
#ifndef AMOEBA_ORACLE_H
#define AMOEBA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Local frame definitions, MultipoleType.java:1504-1506 (enum order). */
enum { ORA_FRAME_NONE = 0, ORA_FRAME_ZONLY = 1, ORA_FRAME_ZTHENX = 2,
       ORA_FRAME_BISECTOR = 3, ORA_FRAME_ZTHENBISECTOR = 4, ORA_FRAME_THREEFOLD = 5 };

/* Polarization model, Polarization.java. */
enum { ORA_POL_NONE = 0, ORA_POL_DIRECT = 1, ORA_POL_MUTUAL = 2 };

typedef struct {
  int n;                    /* atoms */
  const double *xyz;        /* [3n] interleaved */
  const int *frame;         /* [n] */
  const int *axis;          /* [3n] z, x, y axis atoms (-1 if unused) */
  const double *mpole;      /* [10n] local-frame multipoles */
  const double *polarity;   /* [n] A^3 */
  const double *thole;      /* [n] */
  const double *pdamp;      /* [n] polarity^(1/6); <= 0 means undamped */
  const int *group;         /* [n] polarization group id */
  const int *molecule;      /* [n] molecule id (unused by the periodic path) */
  const int *vdw_class;     /* [n] index into vdw_table */
  int n_class;
  const double *vdw_table;  /* [n_class*n_class*2] (1/Rmin, eps) combined pair table */
  const int *red_index;     /* [n] heavy atom of a reduced hydrogen (self otherwise) */
  const double *red_value;  /* [n] reduction factor (0 = no reduction) */
  const int *mask_off;      /* [n+1] CSR offsets of bonded relations */
  const int *mask_atom;     /* [nnz] partner atom */
  const int *mask_rel;      /* [nnz] 2 = 1-2, 3 = 1-3, 4 = 1-4, 5 = 1-5 */
  const double *box;        /* [6] a b c alpha beta gamma (deg); a <= 0: aperiodic */
  const int *soft;          /* [n] alchemical (softcore) atoms, NULL: none */
} ora_system;

typedef struct {
  double ewald_cutoff;      /* A */
  double ewald_alpha;       /* 1/A ; 0 disables reciprocal space */
  double vdw_cutoff;        /* A */
  double vdw_taper;         /* fraction of cutoff where the taper starts (0.9) */
  double list_buffer;       /* A */
  int pme_order;
  int grid[3];
  double polar_eps;         /* Debye */
  double electric;          /* 332.0637133 */
  double precond_cutoff;    /* 4.5 */
  double precond_scale;     /* 2.0 */
  double m_scale[4];        /* mpole 12 13 14 15 */
  double p_scale[4];        /* polar 12 13 14 15 */
  double intra14;           /* polar-14-intra */
  double d11;               /* direct-11-scale */
  double v_scale[3];        /* vdw 12 13 14 */
  double halgren_delta, halgren_gamma;
  int polarization;         /* ORA_POL_* */
  int do_vdw;
  int do_pme;
  int nthreads;             /* 0 = OpenMP default */
  int max_scf_iter;         /* 1000 */
  /* vdW softcore (VanDerWaals.setLambda, VanDerWaals.java:599-653) */
  double lambda;            /* 1 = fully coupled */
  double vdw_lambda_alpha;  /* 0.25 */
  double vdw_lambda_exponent; /* 3 */
  double vdw_lambda_end;    /* 1 */
  int intermolecular_softcore, intramolecular_softcore; /* both-soft pairs (0, 0) */
  /* electrostatic alchemistry in OST mode (AlchemicalParameters.java:171-249,
   * 316-385; ParticleMeshEwald.java:894-935, 1609-1790) */
  int elec_lambda;          /* apply lambda to the electrostatics of the soft atoms (0) */
  double perm_lambda_alpha; /* 2 */
  double perm_lambda_exponent; /* 3 */
  double perm_lambda_start, perm_lambda_end; /* 0.4, 1 */
  double pol_lambda_exponent; /* 3 */
  double pol_lambda_start, pol_lambda_end; /* 0.75, 1 */
  int ligand_vapor_elec;    /* 1 */
  int no_ligand_condensed_scf; /* 1 */
  /* SCF algorithm (ParticleMeshEwald.java:2247-2325, SORRegion.java) */
  int scf_algorithm;        /* 0 PCG, 1 SOR */
  double polar_sor;         /* SOR factor (0.7) */
  int sor_fallback;         /* SOR with factor 0.65 after a failed PCG (1) */
  /* gradient of the energy of the dipoles the PCG actually returned (its
   * iteration count held fixed): the adjoint of the truncated iterations
   * replaces the stationary-dipole terms (0) */
  int finite_pcg;
  /* initial guess of a mutual solve (ParticleMeshEwald.java:2272-2281): when
   * set, [6n] (u then uCR) added to the direct dipoles before the solve, as an
   * SCF predictor adds its extrapolation (NULL: start from the direct dipoles) */
  const double *scf_guess;
  /* when set, [6n] receives the direct dipoles (u then uCR) */
  double *direct_out;
} ora_params;

typedef struct {
  double perm_real, perm_self, perm_recip, perm_charge_corr;
  double pol_real, pol_self, pol_recip;
  double vdw;
  double perm, pol, total;
  int pcg_iters;
  double pcg_eps;
  int n_vdw, n_perm;
  int status;               /* 0 ok, 1 nan energy, 2 scf diverged, 3 scf maxiter, 4 bad torque */
  double vir[9];            /* strain derivative W_ab = dE/d(eps_ab), row-major, kcal/mol */
  double dudl, d2udl2;      /* lambda derivatives (vdW softcore and, with elec_lambda, electrostatics) */
} ora_energies;

/* Full evaluation.  grad [3n] (kcal/mol/A), ind, indcr [3n] (e*A) may be NULL. */
/* As ora_evaluate, plus dU/dX/dlambda [3n] (may be NULL). */
int ora_evaluate_lambda(const ora_system *sys, const ora_params *prm, ora_energies *out, double *grad,
                        double *ind, double *indcr, double *dudx_dl);
int ora_evaluate(const ora_system *sys, const ora_params *prm, ora_energies *out,
                 double *grad, double *ind, double *indcr);

/* OST lambda factors of AlchemicalParameters.update (lPowPerm, lAlpha,
 * lPowPol and their first and second lambda derivatives). */
typedef struct {
  double lpow_perm, dlpow_perm, d2lpow_perm;
  double lalpha, dlalpha, d2lalpha;
  double lpow_pol, dlpow_pol, d2lpow_pol;
} ora_alchemy;
void ora_alchemical_factors(const ora_params *prm, double lambda, ora_alchemy *f);

/* Global-frame multipoles [10n] (RotateMultipolesLoop). */
void ora_rotate_multipoles(const ora_system *sys, double *gmp);

/* Order-p uniform B-spline weights and derivatives 0..nd for fractional
 * offset x (UniformBSpline.bSplineDerivatives); theta[k*(nd+1)+j]. */
void ora_bspline(double x, int order, int nd, double *theta);

/* Real-space pair terms with unit prefactor (no electric factor),
 * r = x_k - x_i.  Multipoles are global-frame, full quadrupoles.
 * Returns the energy; forces are dE/dx_i... see the implementation notes. */
double ora_perm_pair(const double r[3], const double *mi, const double *mk,
                     double aewald, double mscale, double *gi, double *ti, double *tk);
double ora_pol_pair(const double r[3], const double *mi, const double *mk,
                    const double ui[3], const double uip[3], const double uk[3], const double ukp[3],
                    double aewald, double pdi, double pdk, double tholei, double tholek,
                    double dscale, double pscale, double *gi, double *ti, double *tk);

/* 3D complex convolution with an influence array (Complex3DParallel.convolution):
 * forward FFT, multiply by recip[K], inverse (un-normalised) FFT, in place on
 * interleaved complex data [2*nx*ny*nz], x fastest. */
void ora_convolution(double *data, const double *recip, int nx, int ny, int nz);

/* Forward or inverse 3D FFT on interleaved complex data (inverse un-normalised). */
void ora_fft3d(double *data, int nx, int ny, int nz, int inverse);

/* Exact reciprocal-space permanent energy by direct k-space summation (no
 * grid): 1/(2 pi V) sum_{m != 0} exp(-pi^2 m^2/a^2)/m^2 |S(m)|^2 times
 * electric, over |m_i| <= kmax.  Used to pin PME. */
double ora_ewald_recip_exact(const ora_system *sys, const ora_params *prm, int kmax);

/* Influence function of the reference (generalizedInfluenceFunction). */
void ora_influence(const double *box, const int *grid, int order, double aewald, double *recip);

/* Reduced (vdW) coordinates of the reference. */
void ora_reduced_xyz(const ora_system *sys, double *red);

#ifdef __cplusplus
}
#endif
#endif
