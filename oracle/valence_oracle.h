/*
 * valence_oracle.h -- FP64 CPU restatement of Force Field X's AMOEBA valence
 * terms (bond, angle and in-plane angle, stretch-bend, Urey-Bradley,
 * out-of-plane bend, torsion, pi-orbital torsion, torsion-torsion).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the valence kernels of the HIP
 * library (include/amoeba_nb.h, nb_set_valence).  Called from tests/, never
 * from the product path.
 *
 * Terms arrive assigned (atom indices, force constants already multiplied by
 * their unit keywords, ideal values in Angstrom and degrees); energies in
 * kcal/mol.  Coordinates are used as given (no periodic images), as the
 * reference's bonded terms do.
 */
#ifndef VALENCE_ORACLE_H
#define VALENCE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_V_BOND = 0, ORA_V_ANGLE, ORA_V_STRBND, ORA_V_UREY, ORA_V_OPBEND, ORA_V_TORSION, ORA_V_PITORS,
       ORA_V_TORTOR, ORA_V_N };

typedef struct {
  int n_bond;    const int *bond;    const double *bond_k, *bond_r0;      /* [2 n] atoms */
  double bond_cubic, bond_quartic;
  int n_angle;   const int *angle;   const double *angle_k, *angle_t0;    /* [4 n]: a, centre, c, in-plane or -1 */
  double angle_poly[4];                                                   /* cubic, quartic, pentic, sextic */
  int n_strbnd;  const int *strbnd;  const double *strbnd_k, *strbnd_r0, *strbnd_t0;  /* [3 n]; k, r0 [2 n] */
  int n_urey;    const int *urey;    const double *urey_k, *urey_r0;      /* [2 n] */
  double urey_cubic, urey_quartic;
  int n_opbend;  const int *opbend;  const double *opbend_k;             /* [4 n]: a, centre, c, out-of-plane */
  double opbend_poly[4];
  int n_torsion; const int *torsion; const double *torsion_amp, *torsion_cos, *torsion_sin;  /* [4 n]; [6 n] */
  int n_pitors;  const int *pitors;  const double *pitors_k;             /* [6 n] */
  int n_tortor;  const int *tortor;  const int *tortor_grid;             /* [6 n]: 5 atoms, chirality atom or -1 */
  int n_grid;    const int *grid_nx, *grid_ny;
  const int *grid_off_t;             /* [n_grid] offset into grid_t: tx (nx values) then ty (ny values) */
  const int *grid_off_e;             /* [n_grid] offset into the four value arrays (nx * ny each) */
  const double *grid_t, *grid_e, *grid_dx, *grid_dy, *grid_dxy;
} ora_valence;

/* Energies of the eight term kinds (ORA_V_*) into energy[8]; the gradient
 * is added to grad[3n] and the strain derivative sum_i g_i x_i^T to vir[9]
 * (either may be NULL). */
void ora_valence_evaluate(const ora_valence *v, const double *xyz, double *energy, double *grad, double *vir);

#ifdef __cplusplus
}
#endif
#endif
