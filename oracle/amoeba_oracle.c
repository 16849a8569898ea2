/*
 * amoeba_oracle.c -- FP64 CPU restatement of Force Field X's AMOEBA
 * nonbonded energy and gradient.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never linked
 * into it.  Each section names the reference file:line it restates.  The
 * arithmetic is written from the published forms (Smith's multipole
 * gl-function expansion as used by Tinker/FFX, Essmann smooth PME with
 * Sagui's multipole extension, Halgren buffered 14-7) in this file's own
 * organisation: the radial-function-linear pair terms are evaluated once per
 * radial set (screened and bare) and combined, instead of being expanded
 * twice by hand.
 *
 * Threading: OpenMP over atoms with per-thread accumulators, which mirrors
 * the reference's per-thread arrays (AtomicDoubleArray3D / MultiDoubleArray).
 */
#include "amoeba_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SQRT_PI 1.77245385090551602729
#define ELEC_ANG_TO_DEBYE 4.80321

/* ------------------------------------------------------------------------ */
/* small vector helpers                                                      */
/* ------------------------------------------------------------------------ */
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double *a, const double *b, double *c) {
  double x = a[1] * b[2] - a[2] * b[1];
  double y = a[2] * b[0] - a[0] * b[2];
  double z = a[0] * b[1] - a[1] * b[0];
  c[0] = x; c[1] = y; c[2] = z;
}
static inline void axpy3(double a, const double *x, double *y) { y[0] += a * x[0]; y[1] += a * x[1]; y[2] += a * x[2]; }
static inline double norm3(const double *a) { return sqrt(dot3(a, a)); }
static inline void scale3(double s, double *a) { a[0] *= s; a[1] *= s; a[2] *= s; }
static inline void unit3(double *a) { scale3(1.0 / norm3(a), a); }
static inline void sub3(const double *a, const double *b, double *c) { c[0] = a[0] - b[0]; c[1] = a[1] - b[1]; c[2] = a[2] - b[2]; }
/* symmetric 3x3 stored xx yy zz xy xz yz, times vector */
static inline void symv(const double *q, const double *v, double *o) {
  double x = q[0] * v[0] + q[3] * v[1] + q[4] * v[2];
  double y = q[3] * v[0] + q[1] * v[1] + q[5] * v[2];
  double z = q[4] * v[0] + q[5] * v[1] + q[2] * v[2];
  o[0] = x; o[1] = y; o[2] = z;
}
static inline void sym_full(const double *q, double m[3][3]) {
  m[0][0] = q[0]; m[1][1] = q[1]; m[2][2] = q[2];
  m[0][1] = m[1][0] = q[3]; m[0][2] = m[2][0] = q[4]; m[1][2] = m[2][1] = q[5];
}

static int nthreads_of(const ora_params *p) {
#ifdef _OPENMP
  return p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#else
  (void)p;
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* unit cell: Crystal.java:1425-1545 (matrices), 1039-1083 (image)           */
/* ------------------------------------------------------------------------ */
typedef struct {
  int aperiodic;
  int multi;       /* cell thinner than twice the cutoff: explicit image sums */
  double Ai[3][3]; /* rows: lattice vectors a, b, c in Cartesian */
  double A[3][3];  /* inverse of Ai: fractional = cart . A */
  double vol;
} cell_t;

static void inv3(double m[3][3], double o[3][3]) {
  double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  double id = 1.0 / det;
  o[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * id;
  o[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
  o[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
  o[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * id;
  o[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
  o[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
  o[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * id;
  o[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
  o[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
}

static void cell_init(cell_t *c, const double *box) {
  memset(c, 0, sizeof(*c));
  if (!box || box[0] <= 0.0) {
    c->aperiodic = 1;
    return;
  }
  double a = box[0], b = box[1], cc = box[2];
  double al = box[3], be = box[4], ga = box[5];
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
  c->vol = sin_g * gamma_term * a * b * cc;
  c->Ai[0][0] = a; c->Ai[0][1] = 0.0; c->Ai[0][2] = 0.0;
  c->Ai[1][0] = b * cos_g; c->Ai[1][1] = b * sin_g; c->Ai[1][2] = 0.0;
  c->Ai[2][0] = cc * cos_b; c->Ai[2][1] = cc * beta_term; c->Ai[2][2] = cc * gamma_term;
  inv3(c->Ai, c->A);
}

/* minimum image by independent rounding of fractional components (half away
 * from zero), Crystal.image */
static inline double cell_image(const cell_t *c, double *d) {
  if (c->aperiodic) return dot3(d, d);
  double xf = d[0] * c->A[0][0] + d[1] * c->A[1][0] + d[2] * c->A[2][0];
  double yf = d[1] * c->A[1][1] + d[2] * c->A[2][1];
  double zf = d[2] * c->A[2][2];
  xf -= copysign(floor(fabs(xf) + 0.5), xf);
  yf -= copysign(floor(fabs(yf) + 0.5), yf);
  zf -= copysign(floor(fabs(zf) + 0.5), zf);
  d[0] = xf * c->Ai[0][0] + yf * c->Ai[1][0] + zf * c->Ai[2][0];
  d[1] = yf * c->Ai[1][1] + zf * c->Ai[2][1];
  d[2] = zf * c->Ai[2][2];
  return dot3(d, d);
}

/* Perpendicular width of the cell along lattice direction d. */
static double cell_width(const cell_t *c, int d) {
  return 1.0 / sqrt(c->A[0][d] * c->A[0][d] + c->A[1][d] * c->A[1][d] + c->A[2][d] * c->A[2][d]);
}

/* Image sums for cells thinner than twice the cutoff (the reference builds a
 * ReplicatesCrystal, ReplicatesCrystal.java:172-202, applied when the
 * interfacial radius is below the cutoff, ForceFieldEnergy.java:893-901).
 * Here a list entry names the partner atom k and the lattice translation T of
 * the image: e = k + n * code(T).  Bonded masks apply only to the pair as
 * given (T = 0), as the reference masks only the asymmetric unit's own pairs. */
#define IMG_R 8
#define IMG_W (2 * IMG_R + 1)
#define IMG_ORIGIN (IMG_R + IMG_W * (IMG_R + IMG_W * IMG_R))

static inline int img_encode(const int *T) { return (T[0] + IMG_R) + IMG_W * ((T[1] + IMG_R) + IMG_W * (T[2] + IMG_R)); }

/* displacement (x_k + T) - x_i of list entry e of atom i; returns r^2 and
 * sets *k and *orig (the entry is the pair as given: masks apply) */
static inline double entry_vec(const cell_t *c, int n, const double *xyz, int i, int e, int *k, double *r,
                               int *orig) {
  if (!c->multi) {
    *k = e;
    *orig = 1;
    sub3(xyz + 3 * e, xyz + 3 * i, r);
    return cell_image(c, r);
  }
  int kk = e % n, code = e / n;
  int T[3] = {code % IMG_W - IMG_R, (code / IMG_W) % IMG_W - IMG_R, code / (IMG_W * IMG_W) - IMG_R};
  *k = kk;
  *orig = code == IMG_ORIGIN;
  sub3(xyz + 3 * kk, xyz + 3 * i, r);
  for (int a = 0; a < 3; a++) r[a] += T[0] * c->Ai[0][a] + T[1] * c->Ai[1][a] + T[2] * c->Ai[2][a];
  return dot3(r, r);
}

/* ------------------------------------------------------------------------ */
/* local frames: MultipoleType.getRotationMatrix (MultipoleType.java:385-540) */
/* and rotateMultipole (:1031-1058)                                           */
/* ------------------------------------------------------------------------ */
static void frame_axes(int frame, const double *o, const double *pz, const double *px, const double *py,
                       double R[3][3]) {
  double z[3], x[3], y[3];
  memset(R, 0, 9 * sizeof(double));
  R[0][0] = R[1][1] = R[2][2] = 1.0;
  if (frame == ORA_FRAME_NONE) return;
  switch (frame) {
    case ORA_FRAME_BISECTOR: {
      sub3(pz, o, z); unit3(z);
      sub3(px, o, x); unit3(x);
      double zb[3] = {x[0] + z[0], x[1] + z[1], x[2] + z[2]};
      unit3(zb);
      memcpy(z, zb, sizeof z);
      axpy3(-dot3(x, z), z, x);
      unit3(x);
      break;
    }
    case ORA_FRAME_ZTHENBISECTOR: {
      sub3(pz, o, z); unit3(z);
      double a[3], b[3];
      sub3(px, o, a); unit3(a);
      sub3(py, o, b); unit3(b);
      x[0] = a[0] + b[0]; x[1] = a[1] + b[1]; x[2] = a[2] + b[2];
      unit3(x);
      axpy3(-dot3(x, z), z, x);
      unit3(x);
      break;
    }
    case ORA_FRAME_ZONLY: {
      sub3(pz, o, z); unit3(z);
      x[0] = 1.0; x[1] = 0.0; x[2] = 0.0;
      double dt = z[0];
      if (fabs(dt) > 0.866) { x[0] = 0.0; x[1] = 1.0; dt = z[1]; }
      axpy3(-dt, z, x);
      unit3(x);
      break;
    }
    case ORA_FRAME_THREEFOLD: {
      double a[3], b[3], c[3];
      sub3(pz, o, a); unit3(a);
      sub3(px, o, b); unit3(b);
      sub3(py, o, c); unit3(c);
      z[0] = a[0] + b[0] + c[0]; z[1] = a[1] + b[1] + c[1]; z[2] = a[2] + b[2] + c[2];
      unit3(z);
      memcpy(x, a, sizeof x);
      axpy3(-dot3(a, z), z, x);
      unit3(x);
      break;
    }
    default: { /* Z-then-X */
      sub3(pz, o, z); unit3(z);
      sub3(px, o, x);
      axpy3(-dot3(x, z), z, x);
      unit3(x);
      break;
    }
  }
  cross3(z, x, y);
  for (int i = 0; i < 3; i++) { R[i][0] = x[i]; R[i][1] = y[i]; R[i][2] = z[i]; }
}

static void rotate_one(const double R[3][3], const double *in, double *out) {
  out[0] = in[0];
  for (int i = 0; i < 3; i++) out[1 + i] = R[i][0] * in[1] + R[i][1] * in[2] + R[i][2] * in[3];
  double Q[3][3], T[3][3], O[3][3];
  sym_full(in + 4, Q);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[i][j] = R[i][0] * Q[0][j] + R[i][1] * Q[1][j] + R[i][2] * Q[2][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) O[i][j] = T[i][0] * R[j][0] + T[i][1] * R[j][1] + T[i][2] * R[j][2];
  out[4] = O[0][0]; out[5] = O[1][1]; out[6] = O[2][2];
  out[7] = O[0][1]; out[8] = O[0][2]; out[9] = O[1][2];
}

void ora_rotate_multipoles(const ora_system *s, double *gmp) {
  for (int i = 0; i < s->n; i++) {
    const double *o = s->xyz + 3 * i;
    const int *ax = s->axis + 3 * i;
    const double *pz = ax[0] >= 0 ? s->xyz + 3 * ax[0] : o;
    const double *px = ax[1] >= 0 ? s->xyz + 3 * ax[1] : o;
    const double *py = ax[2] >= 0 ? s->xyz + 3 * ax[2] : o;
    double R[3][3];
    frame_axes(s->frame[i], o, pz, px, py, R);
    rotate_one(R, s->mpole + 10 * i, gmp + 10 * i);
  }
}

/* ------------------------------------------------------------------------ */
/* torque -> frame-atom gradients: Torque.torque (Torque.java:111-316)        */
/* ------------------------------------------------------------------------ */
static int torque_to_grad(int frame, const double *o, const double *pz, const double *px, const double *py,
                          const double *trq, double g[4][3]) {
  memset(g, 0, 12 * sizeof(double));
  if (frame == ORA_FRAME_NONE) return 0;
  double u[3], v[3], w[3];
  sub3(pz, o, u);
  double ru = norm3(u);
  if (frame != ORA_FRAME_ZONLY) {
    sub3(px, o, v);
  } else {
    v[0] = 1.0; v[1] = 0.0; v[2] = 0.0;
    if (fabs(u[0] / ru) > 0.866) { v[0] = 0.0; v[1] = 1.0; }
  }
  double rv = norm3(v);
  if (frame == ORA_FRAME_ZTHENBISECTOR || frame == ORA_FRAME_THREEFOLD) {
    sub3(py, o, w);
  } else {
    cross3(u, v, w);
  }
  double rw = norm3(w);
  scale3(1.0 / ru, u); scale3(1.0 / rv, v); scale3(1.0 / rw, w);
  double uv[3], uw[3], vw[3];
  cross3(v, u, uv); unit3(uv);   /* perpendicular to x and z */
  cross3(w, u, uw); unit3(uw);
  cross3(w, v, vw); unit3(vw);
  double cuv = dot3(u, v);
  double suv = sqrt(1.0 - cuv * cuv);
  double dphidu = -dot3(trq, u), dphidv = -dot3(trq, v), dphidw = -dot3(trq, w);
  switch (frame) {
    case ORA_FRAME_ZONLY:
      for (int j = 0; j < 3; j++) {
        double dz = uv[j] * dphidv / (ru * suv) + uw[j] * dphidw / ru;
        g[0][j] = dz; g[3][j] = -dz;
      }
      break;
    case ORA_FRAME_ZTHENX:
      for (int j = 0; j < 3; j++) {
        double dz = uv[j] * dphidv / (ru * suv) + uw[j] * dphidw / ru;
        double dx = -uv[j] * dphidu / (rv * suv);
        g[0][j] = dz; g[1][j] = dx; g[3][j] = -(dz + dx);
      }
      break;
    case ORA_FRAME_BISECTOR:
      for (int j = 0; j < 3; j++) {
        double dz = uv[j] * dphidv / (ru * suv) + 0.5 * uw[j] * dphidw / ru;
        double dx = -uv[j] * dphidu / (rv * suv) + 0.5 * vw[j] * dphidw / rv;
        g[0][j] = dz; g[1][j] = dx; g[3][j] = -(dz + dx);
      }
      break;
    case ORA_FRAME_ZTHENBISECTOR: {
      double r[3] = {v[0] + w[0], v[1] + w[1], v[2] + w[2]};
      double s[3];
      cross3(u, r, s);
      unit3(r); unit3(s);
      double ur[3], us[3], vs[3], ws[3];
      cross3(r, u, ur); unit3(ur);
      cross3(s, u, us); unit3(us);
      cross3(s, v, vs); unit3(vs);
      cross3(s, w, ws); unit3(ws);
      (void)vs; (void)ws;
      double cur = dot3(u, r), sur = sqrt(1.0 - cur * cur);
      double cvs = dot3(v, s), svs = sqrt(1.0 - cvs * cvs);
      double cws = dot3(w, s), sws = sqrt(1.0 - cws * cws);
      double t1[3] = {v[0] - cvs * s[0], v[1] - cvs * s[1], v[2] - cvs * s[2]};
      double t2[3] = {w[0] - cws * s[0], w[1] - cws * s[1], w[2] - cws * s[2]};
      unit3(t1); unit3(t2);
      double ut1c = dot3(u, t1), ut1s = sqrt(1.0 - ut1c * ut1c);
      double ut2c = dot3(u, t2), ut2s = sqrt(1.0 - ut2c * ut2c);
      double dphidr = -dot3(trq, r), dphids = -dot3(trq, s);
      for (int j = 0; j < 3; j++) {
        double dz = ur[j] * dphidr / (ru * sur) + us[j] * dphids / ru;
        double dx = (svs * s[j] - cvs * t1[j]) * dphidu / (rv * (ut1s + ut2s));
        double dy = (sws * s[j] - cws * t2[j]) * dphidu / (rw * (ut1s + ut2s));
        g[0][j] = dz; g[1][j] = dx; g[2][j] = dy; g[3][j] = -(dz + dx + dy);
      }
      break;
    }
    case ORA_FRAME_THREEFOLD: {
      double sum[3] = {u[0] + v[0] + w[0], u[1] + v[1] + w[1], u[2] + v[2] + w[2]};
      double rsum = norm3(sum);
      scale3(1.0 / rsum, sum);
      double cws = dot3(w, sum), cus = dot3(u, sum), cvs = dot3(v, sum);
      const double *axes[3] = {u, v, w};
      double rr[3] = {ru, rv, rw};
      double csum[3] = {cus, cvs, cws};
      /* for each axis, rotate about the bisector of the other two */
      const int other[3][2] = {{1, 2}, {2, 0}, {0, 1}};
      /* reference pairings: Z uses (x+y), X uses (z+y), Y uses (z+x) */
      for (int a = 0; a < 3; a++) {
        const double *e = axes[a];
        const double *p = axes[other[a][0]], *q = axes[other[a][1]];
        double r[3] = {p[0] + q[0], p[1] + q[1], p[2] + q[2]};
        unit3(r);
        double c = dot3(r, e), sn = sqrt(1.0 - c * c);
        double dphidr = -dot3(trq, r);
        double del[3], eps[3];
        cross3(r, e, del); unit3(del);
        double dphiddel = -dot3(trq, del);
        cross3(del, e, eps);
        for (int j = 0; j < 3; j++) {
          double d = del[j] * dphidr / (rr[a] * sn) + eps[j] * dphiddel * csum[a] / (rr[a] * rsum);
          g[a][j] = d;
          g[3][j] -= d;
        }
      }
      break;
    }
    default:
      return -1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* B-splines: UniformBSpline.bSplineDerivatives.  Restated via the cardinal  */
/* recursion M_p(u) = [u M_{p-1}(u) + (p-u) M_{p-1}(u-1)]/(p-1) on the        */
/* shifted weights w_k = M_p(x + p - 1 - k), with derivatives as finite      */
/* differences of lower-order weights.                                      */
/* ------------------------------------------------------------------------ */
void ora_bspline(double x, int order, int nd, double *theta) {
  double w[8][9];
  memset(w, 0, sizeof w);
  /* w[p][k+1] holds order-p weight k (index 0 and p+1 are zero guards) */
  w[1][1] = 1.0;
  for (int p = 2; p <= order; p++) {
    for (int k = 0; k < p; k++) {
      double a = (x + p - 1 - k) * w[p - 1][k];      /* weight k-1 of order p-1 */
      double b = (k + 1 - x) * w[p - 1][k + 1];      /* weight k of order p-1 */
      w[p][k + 1] = (a + b) / (p - 1);
    }
  }
  for (int k = 0; k < order; k++) {
    for (int j = 0; j <= nd; j++) {
      /* d^j theta_k = sum_m (-1)^m C(j,m) w^(order-j)_{k-j+m} */
      int p = order - j;
      double acc = 0.0, binom = 1.0;
      for (int m = 0; m <= j; m++) {
        int idx = k - j + m + 1; /* array slot of weight k-j+m */
        double wv = (idx >= 1 && idx <= p) ? w[p][idx] : 0.0;
        acc += ((m & 1) ? -binom : binom) * wv;
        binom = binom * (j - m) / (m + 1);
      }
      theta[k * (nd + 1) + j] = acc;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* FFT: recursive mixed-radix complex transform for each axis               */
/* ------------------------------------------------------------------------ */
/* decimation-in-time mixed radix: n = p * m, p the smallest prime factor.
 * tw holds exp(sign 2 pi i j / N) for the top-level length N; at a level of
 * length n the decimation stride equals N / n, so the twiddle of angle
 * index q (mod n) is tw[q * stride]. */
static void dft_recursive(const double *in, int stride, double *out, int n, const double *tw, double *scratch) {
  if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
  int p = 0;
  for (int f = 2; f <= n; f++) if (n % f == 0) { p = f; break; }
  int m = n / p;
  for (int r = 0; r < p; r++) dft_recursive(in + 2 * r * stride, stride * p, scratch + 2 * r * m, m, tw, out);
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      int kk = k + q * m;
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; r++) {
        long j = ((long)r * kk % n) * stride;
        double c = tw[2 * j], s = tw[2 * j + 1];
        double xr = scratch[2 * (r * m + k)], xi = scratch[2 * (r * m + k) + 1];
        sr += xr * c - xi * s;
        si += xr * s + xi * c;
      }
      out[2 * kk] = sr; out[2 * kk + 1] = si;
    }
  }
}

static double *twiddles(int n, int sign) {
  double *t = malloc(2 * n * sizeof(double));
  for (int j = 0; j < n; j++) {
    double a = sign * 2.0 * M_PI * j / n;
    t[2 * j] = cos(a); t[2 * j + 1] = sin(a);
  }
  return t;
}

static void fft1d_line(double *line, int n, const double *tw, double *work, double *scratch) {
  dft_recursive(line, 1, work, n, tw, scratch);
  memcpy(line, work, 2 * n * sizeof(double));
}

void ora_fft3d(double *data, int nx, int ny, int nz, int inverse) {
  int sign = inverse ? 1 : -1;
  int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
  double *twx = twiddles(nx, sign), *twy = twiddles(ny, sign), *twz = twiddles(nz, sign);
  #pragma omp parallel
  {
    double *line = malloc(2 * nmax * sizeof(double));
    double *work = malloc(2 * nmax * sizeof(double));
    double *scr = malloc(2 * nmax * sizeof(double));
    /* x lines */
    #pragma omp for
    for (int jk = 0; jk < ny * nz; jk++) {
      double *p = data + 2L * jk * nx;
      fft1d_line(p, nx, twx, work, scr);
    }
    /* y lines */
    #pragma omp for
    for (int ik = 0; ik < nx * nz; ik++) {
      int i = ik % nx, k = ik / nx;
      for (int j = 0; j < ny; j++) {
        long idx = 2L * (i + nx * (j + (long)ny * k));
        line[2 * j] = data[idx]; line[2 * j + 1] = data[idx + 1];
      }
      fft1d_line(line, ny, twy, work, scr);
      for (int j = 0; j < ny; j++) {
        long idx = 2L * (i + nx * (j + (long)ny * k));
        data[idx] = line[2 * j]; data[idx + 1] = line[2 * j + 1];
      }
    }
    /* z lines */
    #pragma omp for
    for (int ij = 0; ij < nx * ny; ij++) {
      for (int k = 0; k < nz; k++) {
        long idx = 2L * (ij + (long)nx * ny * k);
        line[2 * k] = data[idx]; line[2 * k + 1] = data[idx + 1];
      }
      fft1d_line(line, nz, twz, work, scr);
      for (int k = 0; k < nz; k++) {
        long idx = 2L * (ij + (long)nx * ny * k);
        data[idx] = line[2 * k]; data[idx + 1] = line[2 * k + 1];
      }
    }
    free(line); free(work); free(scr);
  }
  free(twx); free(twy); free(twz);
}

void ora_convolution(double *data, const double *recip, int nx, int ny, int nz) {
  ora_fft3d(data, nx, ny, nz, 0);
  long tot = (long)nx * ny * nz;
  for (long i = 0; i < tot; i++) { data[2 * i] *= recip[i]; data[2 * i + 1] *= recip[i]; }
  ora_fft3d(data, nx, ny, nz, 1);
}

/* ------------------------------------------------------------------------ */
/* PME influence function: ReciprocalSpace.generalizedInfluenceFunction     */
/* (ReciprocalSpace.java:1045-1121) with the B-spline modulus and optimal    */
/* zeta correction of discreteFTMod (:435-492).                              */
/* ------------------------------------------------------------------------ */
static void bspline_modulus(double *bsmod, int n, int order) {
  double th[8 * 4];
  ora_bspline(0.0, order, 0, th);
  double *arr = calloc(n > order + 1 ? n : order + 1, sizeof(double));
  for (int k = 0; k < order; k++) arr[1 + k] = th[k];
  for (int i = 0; i < n; i++) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < n; j++) {
      double ang = 2.0 * M_PI * (double)((long)i * j % n) / n;
      s1 += arr[j] * cos(ang);
      s2 += arr[j] * sin(ang);
    }
    bsmod[i] = s1 * s1 + s2 * s2;
  }
  free(arr);
  const double eps = 1.0e-7;
  if (bsmod[0] < eps) bsmod[0] = 0.5 * bsmod[1];
  for (int i = 1; i < n - 1; i++)
    if (bsmod[i] < eps) bsmod[i] = 0.5 * (bsmod[i - 1] + bsmod[i + 1]);
  if (bsmod[n - 1] < eps) bsmod[n - 1] = 0.5 * bsmod[n - 2];
  for (int i = 0; i < n; i++) {
    int k = i > n / 2 ? i - n : i;
    double zeta = 1.0;
    if (k != 0) {
      double s1 = 1.0, s2 = 1.0, f = M_PI * k / n;
      for (int j = 1; j <= 50; j++) {
        double a = f / (f + M_PI * j);
        s1 += pow(a, order); s2 += pow(a, 2 * order);
      }
      for (int j = 1; j <= 50; j++) {
        double a = f / (f - M_PI * j);
        s1 += pow(a, order); s2 += pow(a, 2 * order);
      }
      zeta = s2 / s1;
    }
    bsmod[i] *= zeta * zeta;
  }
}

void ora_influence(const double *box, const int *grid, int order, double aewald, double *recip) {
  cell_t c;
  cell_init(&c, box);
  int nx = grid[0], ny = grid[1], nz = grid[2];
  double *mx = malloc(nx * sizeof(double)), *my = malloc(ny * sizeof(double)), *mz = malloc(nz * sizeof(double));
  bspline_modulus(mx, nx, order);
  bspline_modulus(my, ny, order);
  bspline_modulus(mz, nz, order);
  double piterm = (M_PI / aewald) * (M_PI / aewald);
  double volterm = M_PI * c.vol;
  for (int kz = 0; kz < nz; kz++) {
    int l = kz >= (nz + 1) / 2 ? kz - nz : kz;
    for (int ky = 0; ky < ny; ky++) {
      int k = ky >= (ny + 1) / 2 ? ky - ny : ky;
      for (int kx = 0; kx < nx; kx++) {
        int h = kx >= (nx + 1) / 2 ? kx - nx : kx;
        double s[3];
        for (int d = 0; d < 3; d++) s[d] = c.A[d][0] * h + c.A[d][1] * k + c.A[d][2] * l;
        double s2 = dot3(s, s);
        double term = -piterm * s2;
        double v = 0.0;
        if (term > -50.0 && s2 > 0.0) v = exp(term) / (s2 * volterm * mx[kx] * my[ky] * mz[kz]);
        recip[kx + (long)nx * (ky + (long)ny * kz)] = v;
      }
    }
  }
  recip[0] = 0.0;
  free(mx); free(my); free(mz);
}

/* ------------------------------------------------------------------------ */
/* PME grid helpers: BSplineRegion (ReciprocalSpace.java:1344-1397),         */
/* FractionalMultipoleRegion/toFractionalMultipole (:2778-2834, 1123-1161),  */
/* transformFieldMatrix (:1163-1205), spread (:1445-1505, 1555-1643),        */
/* gather (:2264-2455, 2467-2776).                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
  int K[3], order, nthreads;
  long ntot;
  double a[3][3];    /* a[f][c] = K_f * A[c][f]: Cartesian -> grid-fractional */
  double *theta;     /* [n][3][order][4] */
  int *igrid;        /* [n][3] */
  double *grid;      /* interleaved complex [2 ntot] */
  double *recip;     /* [ntot] */
} pme_t;

static void pme_setup(pme_t *p, const cell_t *c, const ora_params *prm, int n, int nthreads) {
  p->order = prm->pme_order;
  p->nthreads = nthreads;
  for (int d = 0; d < 3; d++) p->K[d] = prm->grid[d];
  p->ntot = (long)p->K[0] * p->K[1] * p->K[2];
  for (int f = 0; f < 3; f++)
    for (int cc = 0; cc < 3; cc++) p->a[f][cc] = p->K[f] * c->A[cc][f];
  p->theta = malloc((size_t)n * 3 * p->order * 4 * sizeof(double));
  p->igrid = malloc((size_t)n * 3 * sizeof(int));
  p->grid = malloc(2 * p->ntot * sizeof(double));
  p->recip = malloc(p->ntot * sizeof(double));
}

static void pme_free(pme_t *p) {
  free(p->theta); free(p->igrid); free(p->grid); free(p->recip);
}

static void pme_bsplines(pme_t *p, const cell_t *c, int n, const double *xyz) {
  int o = p->order;
  for (int i = 0; i < n; i++) {
    const double *r = xyz + 3 * i;
    for (int d = 0; d < 3; d++) {
      double w = r[0] * c->A[0][d] + r[1] * c->A[1][d] + r[2] * c->A[2][d];
      /* Math.round semantics: floor(w + 0.5) */
      double u = w - floor(w + 0.5) + 0.5;
      double fr = p->K[d] * u;
      int ifr = (int)fr;
      double bx = fr - ifr;
      p->igrid[3 * i + d] = ifr - o;
      ora_bspline(bx, o, 3, p->theta + ((size_t)i * 3 + d) * o * 4);
    }
  }
}

static inline int wrap(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }

/* fractional multipole (quadrupole x 1/3, off-diagonal doubled) */
static void frac_multipole(const pme_t *p, const double *g, double *f) {
  f[0] = g[0];
  for (int j = 0; j < 3; j++) f[1 + j] = p->a[j][0] * g[1] + p->a[j][1] * g[2] + p->a[j][2] * g[3];
  double Q[3][3], T[3][3], F[3][3];
  sym_full(g + 4, Q);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[i][j] = p->a[i][0] * Q[0][j] + p->a[i][1] * Q[1][j] + p->a[i][2] * Q[2][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) F[i][j] = T[i][0] * p->a[j][0] + T[i][1] * p->a[j][1] + T[i][2] * p->a[j][2];
  const double t = 1.0 / 3.0;
  f[4] = F[0][0] * t; f[5] = F[1][1] * t; f[6] = F[2][2] * t;
  f[7] = 2.0 * F[0][1] * t; f[8] = 2.0 * F[0][2] * t; f[9] = 2.0 * F[1][2] * t;
}

static void frac_dipole(const pme_t *p, const double *u, double *f) {
  for (int j = 0; j < 3; j++) f[j] = p->a[j][0] * u[0] + p->a[j][1] * u[1] + p->a[j][2] * u[2];
}

/* spread: ncomp=10 permanent (real part), or 3+3 induced (real/imag) */
static void pme_spread(pme_t *p, int n, const double *fm, const double *fu, const double *fp) {
  int o = p->order, nx = p->K[0], ny = p->K[1], nz = p->K[2];
  memset(p->grid, 0, 2 * p->ntot * sizeof(double));
  for (int a = 0; a < n; a++) {
    const double *tx = p->theta + ((size_t)a * 3 + 0) * o * 4;
    const double *ty = p->theta + ((size_t)a * 3 + 1) * o * 4;
    const double *tz = p->theta + ((size_t)a * 3 + 2) * o * 4;
    const int *ig = p->igrid + 3 * a;
    for (int kz = 0; kz < o; kz++) {
      int k = wrap(ig[2] + 1 + kz, nz);
      const double *v = tz + 4 * kz;
      for (int ky = 0; ky < o; ky++) {
        int j = wrap(ig[1] + 1 + ky, ny);
        const double *u = ty + 4 * ky;
        double t0, t1, t2, s0 = 0, s1 = 0;
        if (fm) {
          const double *m = fm + 10 * a;
          t0 = (m[0] * v[0] + m[3] * v[1] + m[6] * v[2]) * u[0] + (m[2] * v[0] + m[9] * v[1]) * u[1] + m[5] * v[0] * u[2];
          t1 = (m[1] * v[0] + m[8] * v[1]) * u[0] + m[7] * v[0] * u[1];
          t2 = m[4] * v[0] * u[0];
        } else {
          const double *d = fu + 3 * a, *q = fp + 3 * a;
          t0 = d[2] * v[1] * u[0] + d[1] * v[0] * u[1];
          t1 = d[0] * v[0] * u[0];
          t2 = 0.0;
          s0 = q[2] * v[1] * u[0] + q[1] * v[0] * u[1];
          s1 = q[0] * v[0] * u[0];
        }
        for (int kx = 0; kx < o; kx++) {
          int i = wrap(ig[0] + 1 + kx, nx);
          const double *w = tx + 4 * kx;
          long idx = 2L * (i + (long)nx * (j + (long)ny * k));
          p->grid[idx] += w[0] * t0 + w[1] * t1 + w[2] * t2;
          if (!fm) p->grid[idx + 1] += w[0] * s0 + w[1] * s1;
        }
      }
    }
  }
}

/* fractional potential derivatives t000..t111 (20) from the grid part `im` */
static void pme_gather(const pme_t *p, int a, int im, double *phi) {
  int o = p->order, nx = p->K[0], ny = p->K[1], nz = p->K[2];
  const double *tx = p->theta + ((size_t)a * 3 + 0) * o * 4;
  const double *ty = p->theta + ((size_t)a * 3 + 1) * o * 4;
  const double *tz = p->theta + ((size_t)a * 3 + 2) * o * 4;
  const int *ig = p->igrid + 3 * a;
  memset(phi, 0, 20 * sizeof(double));
  for (int kz = 0; kz < o; kz++) {
    int k = wrap(ig[2] + 1 + kz, nz);
    const double *v = tz + 4 * kz;
    double tu[10] = {0};
    for (int ky = 0; ky < o; ky++) {
      int j = wrap(ig[1] + 1 + ky, ny);
      const double *u = ty + 4 * ky;
      double t[4] = {0, 0, 0, 0};
      for (int kx = 0; kx < o; kx++) {
        int i = wrap(ig[0] + 1 + kx, nx);
        double q = p->grid[2L * (i + (long)nx * (j + (long)ny * k)) + im];
        const double *w = tx + 4 * kx;
        t[0] += q * w[0]; t[1] += q * w[1]; t[2] += q * w[2]; t[3] += q * w[3];
      }
      /* tu[ab] = sum t[a] u[b]: 00 10 01 20 11 02 30 21 12 03 */
      tu[0] += t[0] * u[0]; tu[1] += t[1] * u[0]; tu[2] += t[0] * u[1];
      tu[3] += t[2] * u[0]; tu[4] += t[1] * u[1]; tu[5] += t[0] * u[2];
      tu[6] += t[3] * u[0]; tu[7] += t[2] * u[1]; tu[8] += t[1] * u[2]; tu[9] += t[0] * u[3];
    }
    /* order: 000 100 010 001 200 020 002 110 101 011 300 030 003 210 201 120 021 102 012 111 */
    phi[0] += tu[0] * v[0]; phi[1] += tu[1] * v[0]; phi[2] += tu[2] * v[0]; phi[3] += tu[0] * v[1];
    phi[4] += tu[3] * v[0]; phi[5] += tu[5] * v[0]; phi[6] += tu[0] * v[2];
    phi[7] += tu[4] * v[0]; phi[8] += tu[1] * v[1]; phi[9] += tu[2] * v[1];
    phi[10] += tu[6] * v[0]; phi[11] += tu[9] * v[0]; phi[12] += tu[0] * v[3];
    phi[13] += tu[7] * v[0]; phi[14] += tu[3] * v[1]; phi[15] += tu[8] * v[0];
    phi[16] += tu[5] * v[1]; phi[17] += tu[1] * v[2]; phi[18] += tu[2] * v[2]; phi[19] += tu[4] * v[1];
  }
}

/* fractional -> Cartesian potential, first 10 components */
static void phi_to_cart(const pme_t *p, const double *f, double *c) {
  /* b[c][f] = K_f A[c][f] = a[f][c] */
  c[0] = f[0];
  for (int j = 0; j < 3; j++) c[1 + j] = p->a[0][j] * f[1] + p->a[1][j] * f[2] + p->a[2][j] * f[3];
  double F[3][3] = {{f[4], f[7], f[8]}, {f[7], f[5], f[9]}, {f[8], f[9], f[6]}};
  double T[3][3], C[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) T[i][j] = p->a[0][i] * F[0][j] + p->a[1][i] * F[1][j] + p->a[2][i] * F[2][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[i][j] = T[i][0] * p->a[0][j] + T[i][1] * p->a[1][j] + T[i][2] * p->a[2][j];
  c[4] = C[0][0]; c[5] = C[1][1]; c[6] = C[2][2]; c[7] = C[0][1]; c[8] = C[0][2]; c[9] = C[1][2];
}

/* ------------------------------------------------------------------------ */
/* neighbor list: half list within cutoff + buffer (NeighborList.java)      */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  int *off;   /* [n+1] */
  int *nbr;   /* pairs */
} nlist_t;

static void nlist_build_images(nlist_t *L, const cell_t *c, int n, const double *xyz, double cut) {
  /* half list over atom-image pairs: every k > i with each translation T in
   * range, and k == i with T in the positive half space */
  double cut2 = cut * cut;
  int m[3];
  for (int d = 0; d < 3; d++) m[d] = (int)ceil(cut / cell_width(c, d)) + 1;
  L->n = n;
  L->off = malloc((n + 1) * sizeof(int));
  int *cnt = calloc(n, sizeof(int));
  int **lists = malloc(n * sizeof(int *));
  #pragma omp parallel for schedule(dynamic, 4)
  for (int i = 0; i < n; i++) {
    int cap = 256, q = 0;
    int *lst = malloc(cap * sizeof(int));
    for (int k = i; k < n; k++) {
      double d0[3];
      sub3(xyz + 3 * k, xyz + 3 * i, d0);
      int T0[3];
      for (int a = 0; a < 3; a++) {
        double f = d0[0] * c->A[0][a] + d0[1] * c->A[1][a] + d0[2] * c->A[2][a];
        T0[a] = -(int)floor(f + 0.5);
      }
      for (int dz = -m[2]; dz <= m[2]; dz++)
        for (int dy = -m[1]; dy <= m[1]; dy++)
          for (int dx = -m[0]; dx <= m[0]; dx++) {
            int T[3] = {T0[0] + dx, T0[1] + dy, T0[2] + dz};
            if (k == i) {
              /* self images: one of each pair T / -T */
              if (T[0] == 0 && T[1] == 0 && T[2] == 0) continue;
              if (T[2] < 0 || (T[2] == 0 && (T[1] < 0 || (T[1] == 0 && T[0] < 0)))) continue;
            }
            double r[3];
            for (int a = 0; a < 3; a++) r[a] = d0[a] + T[0] * c->Ai[0][a] + T[1] * c->Ai[1][a] + T[2] * c->Ai[2][a];
            if (dot3(r, r) > cut2) continue;
            if (abs(T[0]) > IMG_R || abs(T[1]) > IMG_R || abs(T[2]) > IMG_R) {
              fprintf(stderr, "oracle: image translation out of range\n");
              abort();
            }
            if (q == cap) { cap *= 2; lst = realloc(lst, cap * sizeof(int)); }
            lst[q++] = k + n * img_encode(T);
          }
    }
    cnt[i] = q;
    lists[i] = lst;
  }
  L->off[0] = 0;
  for (int i = 0; i < n; i++) L->off[i + 1] = L->off[i] + cnt[i];
  L->nbr = malloc((size_t)(L->off[n] > 0 ? L->off[n] : 1) * sizeof(int));
  for (int i = 0; i < n; i++) {
    memcpy(L->nbr + L->off[i], lists[i], cnt[i] * sizeof(int));
    free(lists[i]);
  }
  free(lists); free(cnt);
}

static void nlist_build(nlist_t *L, const cell_t *c, int n, const double *xyz, double cut) {
  if (c->multi) {
    nlist_build_images(L, c, n, xyz, cut);
    return;
  }
  double cut2 = cut * cut;
  int cap = 1024;
  L->n = n;
  L->off = malloc((n + 1) * sizeof(int));
  int *cnt = calloc(n, sizeof(int));
  int **lists = malloc(n * sizeof(int *));
  int ncell[3] = {1, 1, 1};
  int use_cells = 0;
  if (!c->aperiodic) {
    use_cells = 1;
    for (int d = 0; d < 3; d++) {
      double wid = 1.0 / sqrt(c->A[0][d] * c->A[0][d] + c->A[1][d] * c->A[1][d] + c->A[2][d] * c->A[2][d]);
      ncell[d] = (int)floor(wid / cut);
      if (ncell[d] < 3) use_cells = 0;
    }
  }
  int *cellof = NULL, *cstart = NULL, *corder = NULL;
  int nc = ncell[0] * ncell[1] * ncell[2];
  if (nc < 1) nc = 1;
  if (use_cells) {
    cellof = malloc(n * sizeof(int));
    cstart = calloc(nc + 1, sizeof(int));
    corder = malloc(n * sizeof(int));
    for (int i = 0; i < n; i++) {
      int ci[3];
      for (int d = 0; d < 3; d++) {
        const double *r = xyz + 3 * i;
        double f = r[0] * c->A[0][d] + r[1] * c->A[1][d] + r[2] * c->A[2][d];
        f -= floor(f);
        int q = (int)(f * ncell[d]);
        if (q >= ncell[d]) q = ncell[d] - 1;
        ci[d] = q;
      }
      cellof[i] = ci[0] + ncell[0] * (ci[1] + ncell[1] * ci[2]);
      cstart[cellof[i] + 1]++;
    }
    for (int q = 0; q < nc; q++) cstart[q + 1] += cstart[q];
    int *fill = calloc((size_t)nc, sizeof(int));
    for (int i = 0; i < n; i++) corder[cstart[cellof[i]] + fill[cellof[i]]++] = i;
    free(fill);
  }
  #pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; i++) {
    int mycap = cap, m = 0;
    int *lst = malloc(mycap * sizeof(int));
    const double *ri = xyz + 3 * i;
    if (!use_cells) {
      for (int k = i + 1; k < n; k++) {
        double d[3];
        sub3(xyz + 3 * k, ri, d);
        if (cell_image(c, d) <= cut2) {
          if (m == mycap) { mycap *= 2; lst = realloc(lst, mycap * sizeof(int)); }
          lst[m++] = k;
        }
      }
    } else {
      int q = cellof[i];
      int cx = q % ncell[0], cy = (q / ncell[0]) % ncell[1], cz = q / (ncell[0] * ncell[1]);
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            int qq = wrap(cx + dx, ncell[0]) + ncell[0] * (wrap(cy + dy, ncell[1]) + ncell[1] * wrap(cz + dz, ncell[2]));
            for (int t = cstart[qq]; t < cstart[qq + 1]; t++) {
              int k = corder[t];
              if (k <= i) continue;
              double d[3];
              sub3(xyz + 3 * k, ri, d);
              if (cell_image(c, d) <= cut2) {
                if (m == mycap) { mycap *= 2; lst = realloc(lst, mycap * sizeof(int)); }
                lst[m++] = k;
              }
            }
          }
    }
    cnt[i] = m;
    lists[i] = lst;
  }
  L->off[0] = 0;
  for (int i = 0; i < n; i++) L->off[i + 1] = L->off[i] + cnt[i];
  L->nbr = malloc((size_t)(L->off[n] > 0 ? L->off[n] : 1) * sizeof(int));
  for (int i = 0; i < n; i++) {
    memcpy(L->nbr + L->off[i], lists[i], cnt[i] * sizeof(int));
    free(lists[i]);
  }
  free(lists); free(cnt);
  free(cellof); free(cstart); free(corder);
}

static void nlist_free(nlist_t *L) { free(L->off); free(L->nbr); }

/* ------------------------------------------------------------------------ */
/* Ewald radial functions and Thole damping                                 */
/* ------------------------------------------------------------------------ */
/* bn[0..nmax]: bn0 = erfc(ar)/r, bn_{k} = ((2k-1) bn_{k-1} + an_{k-1} e^{-a^2r^2})/r^2
 * with an_k = (2a^2)^{k+1}/(sqrt(pi) a)  (EwaldParameters.java:100-125). */
static void ewald_bn(double r, double aewald, int nmax, double *bn) {
  double rr1 = 1.0 / r, rr2 = rr1 * rr1;
  if (aewald > 0.0) {
    double ra = aewald * r;
    double ex = exp(-ra * ra);
    double alsq2 = 2.0 * aewald * aewald;
    double an = alsq2 / (SQRT_PI * aewald);
    bn[0] = erfc(ra) * rr1;
    for (int k = 1; k <= nmax; k++) {
      bn[k] = ((2 * k - 1) * bn[k - 1] + an * ex) * rr2;
      an *= alsq2;
    }
  } else {
    bn[0] = rr1;
    for (int k = 1; k <= nmax; k++) bn[k] = (2 * k - 1) * bn[k - 1] * rr2;
  }
}

/* bare Coulomb rr[k] = (2k-1)!! / r^(2k+1) */
static void coulomb_rr(double r, int nmax, double *rr) { ewald_bn(r, 0.0, nmax, rr); }

typedef struct { double s3, s5, s7; double d3[3], d5[3], d7[3]; } thole_t;

/* Thole screening of damped dipoles: u = r/(pd_i pd_k), a = min(thole),
 * damp = -a u^3; s3 = 1 - e^damp, s5 = 1 - e^damp (1 - damp),
 * s7 = 1 - e^damp (1 - damp + 0.6 damp^2) (PermanentFieldRegion.java:726-739),
 * and the gradient factors of RealSpaceEnergyRegion.java:960-980. */
static void thole_terms(double r, const double *rv, double ipdi, double ipdk, double ti, double tk, thole_t *t) {
  t->s3 = t->s5 = t->s7 = 1.0;
  memset(t->d3, 0, sizeof t->d3); memset(t->d5, 0, sizeof t->d5); memset(t->d7, 0, sizeof t->d7);
  double pg = ti < tk ? ti : tk;
  double rd = r * ipdi * ipdk;
  double damp = -pg * rd * rd * rd;
  if (damp > -50.0) {
    double e = exp(damp);
    t->s3 = 1.0 - e;
    t->s5 = 1.0 - e * (1.0 - damp);
    t->s7 = 1.0 - e * (1.0 - damp + 0.6 * damp * damp);
    if (rv) {
      double c3 = -3.0 * damp * e / (r * r);
      for (int d = 0; d < 3; d++) {
        t->d3[d] = c3 * rv[d];
        t->d5[d] = -damp * t->d3[d];
        t->d7[d] = (-0.2 - 0.6 * damp) * t->d5[d];
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* multipole fields (PermanentFieldRegion.java:704-808)                      */
/* ------------------------------------------------------------------------ */
/* field at displacement R from a multipole (c, d, Q/3 in q3) for radial set
 * (B1, B2, B3): E = R (B1 c + B2 d.R + B3 R.Q.R) - B1 d - 2 B2 Q R */
static void mp_field(const double *R, double c, const double *d, const double *q3, double B1, double B2, double B3,
                     double *E) {
  double qr[3];
  symv(q3, R, qr);
  double s = B1 * c + B2 * dot3(d, R) + B3 * dot3(R, qr);
  for (int k = 0; k < 3; k++) E[k] = R[k] * s - B1 * d[k] - 2.0 * B2 * qr[k];
}

/* field of a point dipole u at displacement R: -B1 u + B2 (u.R) R */
static inline void dip_field(const double *R, const double *u, double B1, double B2, double *E) {
  double s = B2 * dot3(u, R);
  for (int k = 0; k < 3; k++) E[k] = -B1 * u[k] + s * R[k];
}

/* ------------------------------------------------------------------------ */
/* permanent multipole pair energy: the gl-function expansion               */
/* (RealSpaceEnergyRegion.permanentPair, :1190-1460).  The terms are linear  */
/* in the radial set B0..B5; the driver evaluates them once with the Ewald   */
/* set and once with the bare Coulomb set.                                   */
/* ------------------------------------------------------------------------ */
typedef struct { double c, d[3], q[6]; } mp3_t; /* q = full quadrupole / 3 */

static void mp3_load(const double *g, mp3_t *m) {
  m->c = g[0];
  m->d[0] = g[1]; m->d[1] = g[2]; m->d[2] = g[3];
  for (int k = 0; k < 6; k++) m->q[k] = g[4 + k] / 3.0;
}

typedef struct {
  double sc3, sc4, sc5, sc6;
  double gl0, gl16, gl278, gl35, gl4;
  double qir[3], qkr[3], qiqkr[3], qkqir[3], qidk[3], qkdi[3];
  double dixdk[3], dixr[3], dkxr[3], qixqk[3], rxqir[3], rxqkr[3], rxqikr[3], rxqkir[3], qkrxqir[3];
  double dixqkr[3], dkxqir[3], rxqidk[3], rxqkdi[3];
} permgeo_t;

static void perm_geometry(const double *r, const mp3_t *I, const mp3_t *K, permgeo_t *g) {
  symv(I->q, r, g->qir);
  symv(K->q, r, g->qkr);
  symv(I->q, g->qkr, g->qiqkr);
  symv(K->q, g->qir, g->qkqir);
  symv(I->q, K->d, g->qidk);
  symv(K->q, I->d, g->qkdi);
  cross3(I->d, K->d, g->dixdk);
  cross3(I->d, r, g->dixr);
  cross3(K->d, r, g->dkxr);
  double Qi[3][3], Qk[3][3];
  sym_full(I->q, Qi); sym_full(K->q, Qk);
  for (int a = 0; a < 3; a++) {
    int b = (a + 1) % 3, c = (a + 2) % 3;
    g->qixqk[a] = 0.0;
    for (int e = 0; e < 3; e++) g->qixqk[a] += Qi[b][e] * Qk[c][e] - Qi[c][e] * Qk[b][e];
  }
  cross3(r, g->qir, g->rxqir);
  cross3(r, g->qkr, g->rxqkr);
  cross3(r, g->qiqkr, g->rxqikr);
  cross3(r, g->qkqir, g->rxqkir);
  cross3(g->qkr, g->qir, g->qkrxqir);
  cross3(I->d, g->qkr, g->dixqkr);
  cross3(K->d, g->qir, g->dkxqir);
  cross3(r, g->qidk, g->rxqidk);
  cross3(r, g->qkdi, g->rxqkdi);
  double sc2 = dot3(I->d, K->d);
  g->sc3 = dot3(I->d, r);
  g->sc4 = dot3(K->d, r);
  g->sc5 = dot3(g->qir, r);
  g->sc6 = dot3(g->qkr, r);
  double sc7 = dot3(g->qir, K->d);
  double sc8 = dot3(g->qkr, I->d);
  double sc9 = dot3(g->qir, g->qkr);
  double sc10 = I->q[0] * K->q[0] + I->q[1] * K->q[1] + I->q[2] * K->q[2] +
                2.0 * (I->q[3] * K->q[3] + I->q[4] * K->q[4] + I->q[5] * K->q[5]);
  g->gl0 = I->c * K->c;
  double gl1 = K->c * g->sc3 - I->c * g->sc4;
  double gl2 = I->c * g->sc6 + K->c * g->sc5 - g->sc3 * g->sc4;
  double gl3 = g->sc3 * g->sc6 - g->sc4 * g->sc5;
  g->gl4 = g->sc5 * g->sc6;
  double gl5 = -4.0 * sc9;
  double gl6 = sc2;
  double gl7 = 2.0 * (sc7 - sc8);
  double gl8 = 2.0 * sc10;
  g->gl16 = gl1 + gl6;
  g->gl278 = gl2 + gl7 + gl8;
  g->gl35 = gl3 + gl5;
}

/* energy, dE/dx_i (ftm), torque on i (ti) and on k (tk) for radial set B */
static double perm_terms(const double *r, const mp3_t *I, const mp3_t *K, const permgeo_t *g, const double *B,
                         double *ftm, double *ti, double *tk) {
  double e = g->gl0 * B[0] + g->gl16 * B[1] + g->gl278 * B[2] + g->gl35 * B[3] + g->gl4 * B[4];
  if (!ftm) return e;
  double gf1 = g->gl0 * B[1] + g->gl16 * B[2] + g->gl278 * B[3] + g->gl35 * B[4] + g->gl4 * B[5];
  double gf2 = -K->c * B[1] + g->sc4 * B[2] - g->sc6 * B[3];
  double gf3 = I->c * B[1] + g->sc3 * B[2] + g->sc5 * B[3];
  double gf4 = 2.0 * B[2];
  double gf5 = 2.0 * (-K->c * B[2] + g->sc4 * B[3] - g->sc6 * B[4]);
  double gf6 = 2.0 * (-I->c * B[2] - g->sc3 * B[3] - g->sc5 * B[4]);
  double gf7 = 4.0 * B[3];
  for (int a = 0; a < 3; a++) {
    ftm[a] = gf1 * r[a] + gf2 * I->d[a] + gf3 * K->d[a] + gf4 * (g->qkdi[a] - g->qidk[a]) + gf5 * g->qir[a] +
             gf6 * g->qkr[a] + gf7 * (g->qiqkr[a] + g->qkqir[a]);
    ti[a] = -B[1] * g->dixdk[a] + gf2 * g->dixr[a] +
            gf4 * (g->dixqkr[a] + g->dkxqir[a] + g->rxqidk[a] - 2.0 * g->qixqk[a]) - gf5 * g->rxqir[a] -
            gf7 * (g->rxqikr[a] + g->qkrxqir[a]);
    tk[a] = B[1] * g->dixdk[a] + gf3 * g->dkxr[a] -
            gf4 * (g->dixqkr[a] + g->dkxqir[a] + g->rxqkdi[a] - 2.0 * g->qixqk[a]) - gf6 * g->rxqkr[a] -
            gf7 * (g->rxqkir[a] - g->qkrxqir[a]);
  }
  return e;
}

double ora_perm_pair(const double r[3], const double *mi, const double *mk, double aewald, double mscale,
                     double *gi, double *ti, double *tk) {
  mp3_t I, K;
  mp3_load(mi, &I);
  mp3_load(mk, &K);
  permgeo_t g;
  perm_geometry(r, &I, &K, &g);
  double rl = norm3(r);
  double bn[7], rr[7];
  ewald_bn(rl, aewald, 6, bn);
  coulomb_rr(rl, 6, rr);
  double f1[3], t1[3], u1[3], f2[3], t2[3], u2[3];
  double e = perm_terms(r, &I, &K, &g, bn, f1, t1, u1);
  double s1 = 1.0 - mscale;
  if (s1 != 0.0) {
    e -= s1 * perm_terms(r, &I, &K, &g, rr, f2, t2, u2);
    for (int a = 0; a < 3; a++) { f1[a] -= s1 * f2[a]; t1[a] -= s1 * t2[a]; u1[a] -= s1 * u2[a]; }
  }
  if (gi) for (int a = 0; a < 3; a++) { gi[a] = f1[a]; ti[a] = t1[a]; tk[a] = u1[a]; }
  return e;
}

/* ------------------------------------------------------------------------ */
/* polarization pair: RealSpaceEnergyRegion.polarizationPair (:1700-2177)   */
/* ------------------------------------------------------------------------ */
typedef struct {
  const double *bn;    /* bn0..bn4 */
  const double *rr;    /* rr1..rr9 (index k -> rr_{2k+1}) */
  thole_t th;
  double scalep, scaled;
  int direct;          /* 1 for DIRECT polarization */
} polctx_t;

static double pol_terms(const double *r, const mp3_t *I, const mp3_t *K, const double *ui, const double *pi,
                        const double *uk, const double *pk, const polctx_t *cx, int grad, double *ftm, double *ti,
                        double *tk) {
  const double *bn = cx->bn, *rr = cx->rr;
  double dsc3 = 1.0 - cx->th.s3 * cx->scaled, dsc5 = 1.0 - cx->th.s5 * cx->scaled, dsc7 = 1.0 - cx->th.s7 * cx->scaled;
  double psc3 = 1.0 - cx->th.s3 * cx->scalep, psc5 = 1.0 - cx->th.s5 * cx->scalep, psc7 = 1.0 - cx->th.s7 * cx->scalep;
  double usc3 = 1.0 - cx->th.s3, usc5 = 1.0 - cx->th.s5;
  double qir[3], qkr[3];
  symv(I->q, r, qir);
  symv(K->q, r, qkr);
  double sc3 = dot3(I->d, r), sc4 = dot3(K->d, r), sc5 = dot3(qir, r), sc6 = dot3(qkr, r);
  double sci1 = dot3(ui, K->d) + dot3(I->d, uk);
  double sci3 = dot3(ui, r), sci4 = dot3(uk, r);
  double sci7 = dot3(qir, uk), sci8 = dot3(qkr, ui);
  double scip1 = dot3(pi, K->d) + dot3(I->d, pk);
  double scip2 = dot3(ui, pk) + dot3(pi, uk);
  double scip3 = dot3(pi, r), scip4 = dot3(pk, r);
  double scip7 = dot3(qir, pk), scip8 = dot3(qkr, pi);
  double gli1 = K->c * sci3 - I->c * sci4, gli2 = -sc3 * sci4 - sci3 * sc4, gli3 = sci3 * sc6 - sci4 * sc5;
  double gli6 = sci1, gli7 = 2.0 * (sci7 - sci8);
  double glip1 = K->c * scip3 - I->c * scip4, glip2 = -sc3 * scip4 - scip3 * sc4, glip3 = scip3 * sc6 - scip4 * sc5;
  double glip6 = scip1, glip7 = 2.0 * (scip7 - scip8);
  double ereal = (gli1 + gli6) * bn[1] + (gli2 + gli7) * bn[2] + gli3 * bn[3];
  double efix = (gli1 + gli6) * rr[1] * psc3 + (gli2 + gli7) * rr[2] * psc5 + gli3 * rr[3] * psc7;
  double e = 0.5 * (ereal - efix);
  if (!grad) return e;

  double qiuk[3], qkui[3], qiukp[3], qkuip[3];
  symv(I->q, uk, qiuk); symv(K->q, ui, qkui); symv(I->q, pk, qiukp); symv(K->q, pi, qkuip);
  double dixuk[3], dkxui[3], dixukp[3], dkxuip[3], dixr[3], dkxr[3], rxqir[3], rxqkr[3];
  cross3(I->d, uk, dixuk); cross3(K->d, ui, dkxui); cross3(I->d, pk, dixukp); cross3(K->d, pi, dkxuip);
  cross3(I->d, r, dixr); cross3(K->d, r, dkxr); cross3(r, qir, rxqir); cross3(r, qkr, rxqkr);
  double uixqkr[3], ukxqir[3], uixqkrp[3], ukxqirp[3], rxqiuk[3], rxqkui[3], rxqiukp[3], rxqkuip[3];
  cross3(ui, qkr, uixqkr); cross3(uk, qir, ukxqir); cross3(pi, qkr, uixqkrp); cross3(pk, qir, ukxqirp);
  cross3(r, qiuk, rxqiuk); cross3(r, qkui, rxqkui); cross3(r, qiukp, rxqiukp); cross3(r, qkuip, rxqkuip);

  /* screened (Ewald) part */
  double gfi1 = 0.5 * bn[2] * (gli1 + glip1 + gli6 + glip6 + scip2) +
                0.5 * bn[3] * (gli2 + glip2 + gli7 + glip7 - sci3 * scip4 - scip3 * sci4) + 0.5 * bn[4] * (gli3 + glip3);
  double gfi2 = -K->c * bn[1] + sc4 * bn[2] - sc6 * bn[3];
  double gfi3 = I->c * bn[1] + sc3 * bn[2] + sc5 * bn[3];
  double gfi4 = 2.0 * bn[2];
  double gfi5 = bn[3] * (sci4 + scip4);
  double gfi6 = -bn[3] * (sci3 + scip3);
  double gti2 = 0.5 * bn[2] * (sci4 + scip4), gti3 = 0.5 * bn[2] * (sci3 + scip3);
  double f[3], t2[3], t3[3];
  for (int a = 0; a < 3; a++) {
    f[a] = gfi1 * r[a] +
           0.5 * (gfi2 * (ui[a] + pi[a]) + bn[2] * (sci4 * pi[a] + scip4 * ui[a]) + gfi3 * (uk[a] + pk[a]) +
                  bn[2] * (sci3 * pk[a] + scip3 * uk[a]) + (sci4 + scip4) * bn[2] * I->d[a] +
                  (sci3 + scip3) * bn[2] * K->d[a] + gfi4 * (qkui[a] + qkuip[a] - qiuk[a] - qiukp[a])) +
           gfi5 * qir[a] + gfi6 * qkr[a];
    t2[a] = -0.5 * bn[1] * (dixuk[a] + dixukp[a]) + gti2 * dixr[a] - gfi5 * rxqir[a] +
            0.5 * gfi4 * (ukxqir[a] + rxqiuk[a] + ukxqirp[a] + rxqiukp[a]);
    t3[a] = -0.5 * bn[1] * (dkxui[a] + dkxuip[a]) + gti3 * dkxr[a] - gfi6 * rxqkr[a] -
            0.5 * gfi4 * (uixqkr[a] + rxqkui[a] + uixqkrp[a] + rxqkuip[a]);
  }
  /* bare Coulomb part weighted by the screening complements */
  if (psc3 != 0.0 || dsc3 != 0.0 || usc3 != 0.0) {
    double gfr1 = 0.5 * rr[2] * ((gli1 + gli6) * psc3 + (glip1 + glip6) * dsc3 + scip2 * usc3) +
                  0.5 * rr[3] * ((gli7 + gli2) * psc5 + (glip7 + glip2) * dsc5 - (sci3 * scip4 + scip3 * sci4) * usc5) +
                  0.5 * rr[4] * (gli3 * psc7 + glip3 * dsc7);
    double gfr4 = 2.0 * rr[2];
    double gfr5 = rr[3] * (sci4 * psc7 + scip4 * dsc7);
    double gfr6 = -rr[3] * (sci3 * psc7 + scip3 * dsc7);
    double gtr2 = 0.5 * rr[2] * (sci4 * psc5 + scip4 * dsc5);
    double gtr3 = 0.5 * rr[2] * (sci3 * psc5 + scip3 * dsc5);
    for (int a = 0; a < 3; a++) {
      double fr = gfr1 * r[a] +
                  0.5 * (-rr[1] * K->c * (ui[a] * psc3 + pi[a] * dsc3) + rr[2] * sc4 * (ui[a] * psc5 + pi[a] * dsc5) -
                         rr[3] * sc6 * (ui[a] * psc7 + pi[a] * dsc7)) +
                  0.5 * (rr[1] * I->c * (uk[a] * psc3 + pk[a] * dsc3) + rr[2] * sc3 * (uk[a] * psc5 + pk[a] * dsc5) +
                         rr[3] * sc5 * (uk[a] * psc7 + pk[a] * dsc7)) +
                  0.5 * rr[2] * usc5 * (sci4 * pi[a] + scip4 * ui[a] + sci3 * pk[a] + scip3 * uk[a]) +
                  0.5 * (sci4 * psc5 + scip4 * dsc5) * rr[2] * I->d[a] +
                  0.5 * (sci3 * psc5 + scip3 * dsc5) * rr[2] * K->d[a] +
                  0.5 * gfr4 * ((qkui[a] - qiuk[a]) * psc5 + (qkuip[a] - qiukp[a]) * dsc5) + gfr5 * qir[a] +
                  gfr6 * qkr[a];
      double tr2 = -0.5 * rr[1] * (dixuk[a] * psc3 + dixukp[a] * dsc3) + gtr2 * dixr[a] - gfr5 * rxqir[a] +
                   0.5 * gfr4 * ((ukxqir[a] + rxqiuk[a]) * psc5 + (ukxqirp[a] + rxqiukp[a]) * dsc5);
      double tr3 = -0.5 * rr[1] * (dkxui[a] * psc3 + dkxuip[a] * dsc3) + gtr3 * dkxr[a] - gfr6 * rxqkr[a] -
                   0.5 * gfr4 * ((uixqkr[a] + rxqkui[a]) * psc5 + (uixqkrp[a] + rxqkuip[a]) * dsc5);
      f[a] -= fr; t2[a] -= tr2; t3[a] -= tr3;
    }
  }
  /* derivatives of the Thole screening functions */
  double a3 = 0.5 * rr[1] * ((gli1 + gli6) * cx->scalep + (glip1 + glip6) * cx->scaled);
  double a5 = 0.5 * rr[2] * ((gli2 + gli7) * cx->scalep + (glip2 + glip7) * cx->scaled);
  double a7 = 0.5 * rr[3] * (gli3 * cx->scalep + glip3 * cx->scaled);
  double b3 = 0.5 * rr[1] * scip2, b5 = -0.5 * rr[2] * (sci3 * scip4 + scip3 * sci4);
  double findmp[3];
  for (int a = 0; a < 3; a++) {
    double fridmp = a3 * cx->th.d3[a] + a5 * cx->th.d5[a] + a7 * cx->th.d7[a];
    findmp[a] = b3 * cx->th.d3[a] + b5 * cx->th.d5[a];
    f[a] -= fridmp + findmp[a];
  }
  if (cx->direct) {
    double gfd = 0.5 * (bn[2] * scip2 - bn[3] * (scip3 * sci4 + sci3 * scip4));
    double gfdr = 0.5 * (rr[2] * scip2 * usc3 - rr[3] * (scip3 * sci4 + sci3 * scip4) * usc5);
    for (int a = 0; a < 3; a++) {
      double v = sci4 * pi[a] + scip4 * ui[a] + sci3 * pk[a] + scip3 * uk[a];
      f[a] -= gfd * r[a] + 0.5 * bn[2] * v;
      f[a] += gfdr * r[a] + 0.5 * usc5 * rr[2] * v + findmp[a];
    }
  }
  for (int a = 0; a < 3; a++) { ftm[a] = f[a]; ti[a] = t2[a]; tk[a] = t3[a]; }
  return e;
}

double ora_pol_pair(const double r[3], const double *mi, const double *mk, const double ui[3], const double uip[3],
                    const double uk[3], const double ukp[3], double aewald, double pdi, double pdk, double tholei,
                    double tholek, double dscale, double pscale, double *gi, double *ti, double *tk) {
  mp3_t I, K;
  mp3_load(mi, &I);
  mp3_load(mk, &K);
  double rl = norm3(r);
  double bn[7], rr[7];
  ewald_bn(rl, aewald, 6, bn);
  coulomb_rr(rl, 6, rr);
  polctx_t cx;
  cx.bn = bn; cx.rr = rr;
  double ipdi = pdi > 0 ? 1.0 / pdi : INFINITY, ipdk = pdk > 0 ? 1.0 / pdk : INFINITY;
  thole_terms(rl, r, ipdi, ipdk, tholei, tholek, &cx.th);
  cx.scalep = pscale; cx.scaled = dscale; cx.direct = 0;
  double f[3], a[3], b[3];
  double e = pol_terms(r, &I, &K, ui, uip, uk, ukp, &cx, 1, f, a, b);
  if (gi) for (int c = 0; c < 3; c++) { gi[c] = f[c]; ti[c] = a[c]; tk[c] = b[c]; }
  return e;
}

/* ------------------------------------------------------------------------ */
/* masks: RealSpaceEnergyRegion.applyMask (:308-352) and                    */
/* PermanentFieldRegion.applyMask (:232-259)                                 */
/* ------------------------------------------------------------------------ */
typedef struct { double m, p, d; } scales_t;

static void apply_masks(const ora_system *s, const ora_params *prm, int i, double *m, double *p, double *d) {
  for (int t = s->mask_off[i]; t < s->mask_off[i + 1]; t++) {
    int k = s->mask_atom[t], rel = s->mask_rel[t];
    if (rel < 2 || rel > 5) continue;
    m[k] = prm->m_scale[rel - 2];
    if (rel <= 4) {
      double ps = prm->p_scale[rel - 2];
      if (rel == 4 && s->group[i] == s->group[k]) ps *= prm->intra14;
      p[k] = ps;
    }
  }
  (void)d;
}

static void remove_masks(const ora_system *s, int i, double *m, double *p) {
  for (int t = s->mask_off[i]; t < s->mask_off[i + 1]; t++) {
    int k = s->mask_atom[t];
    m[k] = 1.0; p[k] = 1.0;
  }
}

/* ------------------------------------------------------------------------ */
/* evaluation context                                                       */
/* ------------------------------------------------------------------------ */
/* Scale factors of one electrostatics evaluation.  A plain evaluation has
 * every scale 1 and no lambda terms.  Under OST alchemistry the energy is
 * the sum of three such evaluations (ParticleMeshEwald.java:894-935):
 * permanent real-space pairs with a soft atom are softcore-shifted
 * (r -> sqrt(r^2 + lalpha)) and scaled by l2 (RealSpaceEnergyRegion.java:
 * 875-908); reciprocal, self and charge-correction terms are scaled by perm,
 * all polarization terms by pol (ReciprocalEnergyRegion.java:254, 390-391,
 * 555-556; RealSpaceEnergyRegion.java:1818, 2176); d* and d2* are the
 * first and second lambda derivatives of those scales, sign included. */
typedef struct {
  int perm_real;      /* permanent real-space energy on */
  int polar;          /* polarization on */
  int sor;            /* solve the dipoles by SOR (the vapor phase, ParticleMeshEwald.java:1744-1746) */
  int soft_pairs;     /* pairs with a soft atom use l2 and the softcore shift */
  int vapor;          /* skip both-soft pairs per the inter/intramolecular flags */
  double l2, lalpha, dlalpha, d2lalpha;
  double perm, dperm, d2perm;
  double pol, dpol, d2pol;
} phase_t;

static const phase_t PHASE_PLAIN = {1, 1, 0, 0, 0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0};

typedef struct {
  const ora_system *s;
  const ora_params *prm;
  const phase_t *ph;
  int n, nth;
  cell_t cell;
  double aewald, off2;
  double *gmp;          /* [10n] global multipoles */
  double *ipd;          /* [n] 1/pdamp */
  nlist_t nl;           /* half list at list cutoff */
  int *ew_off, *ew;     /* ewald list (CSR, half) */
  int *pc_off, *pc;     /* preconditioner list */
  int use_recip;
  pme_t pme;
  double *fmp;          /* [10n] fractional multipoles */
  double *phi_p_f, *phi_p_c;     /* permanent: [20n] fractional, [10n] Cartesian */
  double *phi_u_f, *phi_u_c, *phi_pp_f, *phi_pp_c; /* induced u / uCR */
  double *field, *fieldp;        /* [3n] */
  double *ind, *indp, *dir, *dirp;
  double *grad, *torque;
  double *lgrad, *ltorque;       /* dU/dX/dlambda and its torque */
  int n_perm;
  struct pcg_tape *tape;         /* the u-system iterations of the PCG (finite_pcg) */
  int finite_grad;               /* gradient terms of direct polarization with uCR -> adjoint */
} ctx_t;

/* The PCG iterations of the u system, for the adjoint of the truncated solve:
 * at iteration k the residual r_k, the preconditioned residual z_k, the
 * direction p_k, q_k = A p_k, and the scalars a_k = gam_k / del_k
 * (gam_k = r_k.z_k, del_k = p_k.q_k) and beta_k. */
typedef struct pcg_tape {
  int K, cap;
  double *r, *z, *p, *q;   /* [cap][3n] */
  double *a, *b, *gam, *del;
} pcg_tape_t;

static void tape_push(pcg_tape_t *t, int n3, const double *r, const double *z, const double *p, const double *q,
                      double a, double gam, double del) {
  if (t->K == t->cap) {
    t->cap = t->cap ? 2 * t->cap : 8;
    double **vs[4] = {&t->r, &t->z, &t->p, &t->q};
    for (int v = 0; v < 4; v++) *vs[v] = realloc(*vs[v], (size_t)t->cap * n3 * sizeof(double));
    double **ss[4] = {&t->a, &t->b, &t->gam, &t->del};
    for (int v = 0; v < 4; v++) *ss[v] = realloc(*ss[v], (size_t)t->cap * sizeof(double));
  }
  int k = t->K++;
  memcpy(t->r + (size_t)k * n3, r, n3 * sizeof(double));
  memcpy(t->z + (size_t)k * n3, z, n3 * sizeof(double));
  memcpy(t->p + (size_t)k * n3, p, n3 * sizeof(double));
  memcpy(t->q + (size_t)k * n3, q, n3 * sizeof(double));
  t->a[k] = a;
  t->b[k] = 0.0;
  t->gam[k] = gam;
  t->del[k] = del;
}

static void tape_free(pcg_tape_t *t) {
  free(t->r); free(t->z); free(t->p); free(t->q); free(t->a); free(t->b); free(t->gam); free(t->del);
}

/* induced dipole real-space field over a (half) list with Thole screening,
 * no masks: InducedDipoleFieldRegion.java:308-450 and the preconditioner of
 * PCGSolver.java:1270-1390 (aewald = 0). */
static void induced_real_field(ctx_t *c, const int *off, const int *lst, double aewald, const double *u,
                               const double *up, double *f, double *fp) {
  int n = c->n;
  int nth = c->nth;
  double *acc = calloc((size_t)nth * 6 * n, sizeof(double));
  #pragma omp parallel num_threads(nth)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *af = acc + (size_t)tid * 6 * n;
    #pragma omp for schedule(dynamic, 32)
    for (int i = 0; i < n; i++) {
      const double *xi = c->s->xyz + 3 * i;
      double fi[3] = {0, 0, 0}, pi_[3] = {0, 0, 0};
      for (int t = off[i]; t < off[i + 1]; t++) {
        int k, orig;
        double r[3];
        double r2 = entry_vec(&c->cell, n, c->s->xyz, i, lst[t], &k, r, &orig);
        double rl = sqrt(r2);
        double bn[3];
        ewald_bn(rl, aewald, 2, bn);
        thole_t th;
        thole_terms(rl, NULL, c->ipd[i], c->ipd[k], c->s->thole[i], c->s->thole[k], &th);
        double rr3 = 1.0 / (rl * r2), rr5 = 3.0 * rr3 / r2;
        double B1 = bn[1] - (1.0 - th.s3) * rr3, B2 = bn[2] - (1.0 - th.s5) * rr5;
        double E[3];
        dip_field(r, u + 3 * k, B1, B2, E);
        axpy3(1.0, E, fi);
        dip_field(r, up + 3 * k, B1, B2, E);
        axpy3(1.0, E, pi_);
        dip_field(r, u + 3 * i, B1, B2, E);
        axpy3(1.0, E, af + 3 * k);
        dip_field(r, up + 3 * i, B1, B2, E);
        axpy3(1.0, E, af + 3 * n + 3 * k);
      }
      axpy3(1.0, fi, af + 3 * i);
      axpy3(1.0, pi_, af + 3 * n + 3 * i);
    }
  }
  for (int t = 0; t < nth; t++)
    for (int i = 0; i < 3 * n; i++) {
      f[i] += acc[(size_t)t * 6 * n + i];
      fp[i] += acc[(size_t)t * 6 * n + 3 * n + i];
    }
  free(acc);
}

/* reciprocal + self field of induced dipoles; fills induced phi arrays.
 * ParticleMeshEwald.computeInduceDipoleField (:763-804),
 * InducedDipoleFieldReduceRegion.java:159-215 */
static void induced_field(ctx_t *c, const double *u, const double *up, double *f, double *fp) {
  int n = c->n;
  memset(f, 0, 3 * n * sizeof(double));
  memset(fp, 0, 3 * n * sizeof(double));
  induced_real_field(c, c->ew_off, c->ew, c->aewald, u, up, f, fp);
  if (c->use_recip) {
    pme_t *p = &c->pme;
    double *fu = malloc(3 * n * sizeof(double)), *fpp = malloc(3 * n * sizeof(double));
    for (int i = 0; i < n; i++) { frac_dipole(p, u + 3 * i, fu + 3 * i); frac_dipole(p, up + 3 * i, fpp + 3 * i); }
    pme_spread(p, n, NULL, fu, fpp);
    ora_convolution(p->grid, p->recip, p->K[0], p->K[1], p->K[2]);
    #pragma omp parallel for num_threads(c->nth)
    for (int i = 0; i < n; i++) {
      pme_gather(p, i, 0, c->phi_u_f + 20 * i);
      pme_gather(p, i, 1, c->phi_pp_f + 20 * i);
      phi_to_cart(p, c->phi_u_f + 20 * i, c->phi_u_c + 10 * i);
      phi_to_cart(p, c->phi_pp_f + 20 * i, c->phi_pp_c + 10 * i);
    }
    double a3 = 4.0 / 3.0 * pow(c->aewald, 3.0) / SQRT_PI;
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) {
        f[3 * i + d] += a3 * u[3 * i + d] - c->phi_u_c[10 * i + 1 + d];
        fp[3 * i + d] += a3 * up[3 * i + d] - c->phi_pp_c[10 * i + 1 + d];
      }
    free(fu); free(fpp);
  }
}

/* permanent field (PermanentFieldRegion real space + DirectRegion) and the
 * Ewald / preconditioner lists */
static void permanent_field(ctx_t *c) {
  int n = c->n, nth = c->nth;
  const ora_system *s = c->s;
  const ora_params *prm = c->prm;
  double *acc = calloc((size_t)nth * 6 * n, sizeof(double));
  int *ewcnt = calloc(n, sizeof(int)), *pccnt = calloc(n, sizeof(int));
  int **ewl = malloc(n * sizeof(int *)), **pcl = malloc(n * sizeof(int *));
  int npairs_tot = 0;
  double pc2 = prm->precond_cutoff;
  #pragma omp parallel num_threads(nth) reduction(+ : npairs_tot)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *af = acc + (size_t)tid * 6 * n;
    double *pm = malloc(n * sizeof(double)), *mm = malloc(n * sizeof(double));
    for (int k = 0; k < n; k++) { pm[k] = 1.0; mm[k] = 1.0; }
    #pragma omp for schedule(dynamic, 32)
    for (int i = 0; i < n; i++) {
      apply_masks(s, prm, i, mm, pm, NULL);
      int gi = s->group[i];
      const double *xi = s->xyz + 3 * i;
      mp3_t I;
      mp3_load(c->gmp + 10 * i, &I);
      int cap = c->nl.off[i + 1] - c->nl.off[i];
      int *ew = malloc((cap > 0 ? cap : 1) * sizeof(int)), *pcq = malloc((cap > 0 ? cap : 1) * sizeof(int));
      int ne = 0, np = 0;
      double fi[3] = {0, 0, 0}, fip[3] = {0, 0, 0};
      for (int t = c->nl.off[i]; t < c->nl.off[i + 1]; t++) {
        int e = c->nl.nbr[t], k, orig;
        double r[3];
        double r2 = entry_vec(&c->cell, n, s->xyz, i, e, &k, r, &orig);
        if (r2 > c->off2) continue;
        npairs_tot++;
        ew[ne++] = e;
        double rl = sqrt(r2);
        if (rl < pc2) pcq[np++] = e;
        mp3_t K;
        mp3_load(c->gmp + 10 * k, &K);
        double bn[4];
        ewald_bn(rl, c->aewald, 3, bn);
        thole_t th;
        thole_terms(rl, NULL, c->ipd[i], c->ipd[k], s->thole[i], s->thole[k], &th);
        double dsc = (orig && s->group[k] == gi) ? prm->d11 : 1.0;
        double psc = orig ? pm[k] : 1.0;
        double rr3 = 1.0 / (rl * r2), rr5 = 3.0 * rr3 / r2, rr7 = 5.0 * rr5 / r2;
        double D1 = bn[1] - (1.0 - th.s3 * dsc) * rr3, D2 = bn[2] - (1.0 - th.s5 * dsc) * rr5,
               D3 = bn[3] - (1.0 - th.s7 * dsc) * rr7;
        double P1 = bn[1] - (1.0 - th.s3 * psc) * rr3, P2 = bn[2] - (1.0 - th.s5 * psc) * rr5,
               P3 = bn[3] - (1.0 - th.s7 * psc) * rr7;
        double E[3], mr[3] = {-r[0], -r[1], -r[2]};
        /* field at k from i */
        mp_field(r, I.c, I.d, I.q, D1, D2, D3, E);
        axpy3(1.0, E, af + 3 * k);
        mp_field(r, I.c, I.d, I.q, P1, P2, P3, E);
        axpy3(1.0, E, af + 3 * n + 3 * k);
        /* field at i from k */
        mp_field(mr, K.c, K.d, K.q, D1, D2, D3, E);
        axpy3(1.0, E, fi);
        mp_field(mr, K.c, K.d, K.q, P1, P2, P3, E);
        axpy3(1.0, E, fip);
      }
      axpy3(1.0, fi, af + 3 * i);
      axpy3(1.0, fip, af + 3 * n + 3 * i);
      ewcnt[i] = ne; pccnt[i] = np; ewl[i] = ew; pcl[i] = pcq;
      remove_masks(s, i, mm, pm);
    }
    free(pm); free(mm);
  }
  c->n_perm = npairs_tot;
  memset(c->field, 0, 3 * n * sizeof(double));
  memset(c->fieldp, 0, 3 * n * sizeof(double));
  for (int t = 0; t < nth; t++)
    for (int i = 0; i < 3 * n; i++) {
      c->field[i] += acc[(size_t)t * 6 * n + i];
      c->fieldp[i] += acc[(size_t)t * 6 * n + 3 * n + i];
    }
  free(acc);
  c->ew_off = malloc((n + 1) * sizeof(int));
  c->pc_off = malloc((n + 1) * sizeof(int));
  c->ew_off[0] = c->pc_off[0] = 0;
  for (int i = 0; i < n; i++) { c->ew_off[i + 1] = c->ew_off[i] + ewcnt[i]; c->pc_off[i + 1] = c->pc_off[i] + pccnt[i]; }
  c->ew = malloc((c->ew_off[n] + 1) * sizeof(int));
  c->pc = malloc((c->pc_off[n] + 1) * sizeof(int));
  for (int i = 0; i < n; i++) {
    memcpy(c->ew + c->ew_off[i], ewl[i], ewcnt[i] * sizeof(int));
    memcpy(c->pc + c->pc_off[i], pcl[i], pccnt[i] * sizeof(int));
    free(ewl[i]); free(pcl[i]);
  }
  free(ewl); free(pcl); free(ewcnt); free(pccnt);
}

/* ------------------------------------------------------------------------ */
/* PCG: PCGSolver.scfByPCG (PCGSolver.java:487-604) and its regions          */
/* ------------------------------------------------------------------------ */
static int pcg_solve(ctx_t *c, ora_energies *out) {
  int n = c->n;
  const double *pol = c->s->polarity;
  double psc = c->prm->precond_scale;
  size_t sz = 3 * (size_t)n * sizeof(double);
  double *r = malloc(sz), *rp = malloc(sz), *z = malloc(sz), *zp = malloc(sz), *p = malloc(sz), *pp = malloc(sz);
  double *vec = malloc(sz), *vecp = malloc(sz), *f = malloc(sz), *fp = malloc(sz), *au = malloc(sz), *aup = malloc(sz);
  /* field of the initial guess */
  induced_field(c, c->ind, c->indp, f, fp);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      int q = 3 * i + d;
      if (pol[i] > 0.0) {
        r[q] = (c->dir[q] - c->ind[q]) / pol[i] + f[q];
        rp[q] = (c->dirp[q] - c->indp[q]) / pol[i] + fp[q];
      } else {
        r[q] = rp[q] = 0.0;
      }
    }
  /* preconditioner field of alpha*r */
  #define PRECOND()                                                                                  \
    do {                                                                                             \
      for (int i = 0; i < n; i++)                                                                    \
        for (int d = 0; d < 3; d++) { au[3 * i + d] = pol[i] * r[3 * i + d]; aup[3 * i + d] = pol[i] * rp[3 * i + d]; } \
      memset(f, 0, sz); memset(fp, 0, sz);                                                           \
      induced_real_field(c, c->pc_off, c->pc, 0.0, au, aup, f, fp);                                  \
    } while (0)
  PRECOND();
  for (int q = 0; q < 3 * n; q++) {
    int i = q / 3;
    z[q] = pol[i] * (f[q] + psc * r[q]);
    zp[q] = pol[i] * (fp[q] + psc * rp[q]);
    p[q] = z[q]; pp[q] = zp[q];
  }
  int iter = 0, maxit = c->prm->max_scf_iter > 0 ? c->prm->max_scf_iter : 1000;
  double eps = 100.0, prev;
  int status = 0;
  for (;;) {
    /* field of the search direction p */
    induced_field(c, p, pp, f, fp);
    double pap = 0, papp = 0, rz = 0, rzp = 0;
    for (int q = 0; q < 3 * n; q++) {
      int i = q / 3;
      /* vec = A p = p / alpha - E(p) */
      if (pol[i] > 0.0) {
        vec[q] = p[q] / pol[i] - f[q];
        vecp[q] = pp[q] / pol[i] - fp[q];
      } else {
        c->ind[q] = c->indp[q] = vec[q] = vecp[q] = 0.0;
      }
      pap += p[q] * vec[q]; papp += pp[q] * vecp[q];
      rz += r[q] * z[q]; rzp += rp[q] * zp[q];
    }
    double al = pap != 0.0 ? rz / pap : rz, alp = papp != 0.0 ? rzp / papp : rzp;
    if (c->tape) tape_push(c->tape, 3 * n, r, z, p, vec, al, rz, pap);
    double e1 = 0, e2 = 0;
    for (int q = 0; q < 3 * n; q++) {
      r[q] -= al * vec[q]; rp[q] -= alp * vecp[q];
      c->ind[q] += al * p[q]; c->indp[q] += alp * pp[q];
      e1 += r[q] * r[q]; e2 += rp[q] * rp[q];
    }
    prev = eps;
    eps = e1 > e2 ? e1 : e2;
    iter++;
    eps = ELEC_ANG_TO_DEBYE * sqrt(eps / n);
    if (eps > prev) { status = 2; break; }
    if (iter >= maxit) { status = 3; break; }
    if (eps < c->prm->polar_eps) break;
    PRECOND();
    double rz2 = 0, rzp2 = 0;
    for (int q = 0; q < 3 * n; q++) {
      int i = q / 3;
      z[q] = pol[i] * (f[q] + psc * r[q]);
      zp[q] = pol[i] * (fp[q] + psc * rp[q]);
      rz2 += r[q] * z[q]; rzp2 += rp[q] * zp[q];
    }
    double b = rz2 / (rz != 0.0 ? rz : 1.0), bp = rzp2 / (rzp != 0.0 ? rzp : 1.0);
    if (c->tape) c->tape->b[c->tape->K - 1] = b;
    for (int q = 0; q < 3 * n; q++) {
      p[q] = z[q] + b * p[q];
      pp[q] = zp[q] + bp * pp[q];
    }
  }
  #undef PRECOND
  out->pcg_iters = iter;
  out->pcg_eps = eps;
  /* final induced field and reciprocal potentials of the converged dipoles */
  induced_field(c, c->ind, c->indp, f, fp);
  free(r); free(rp); free(z); free(zp); free(p); free(pp); free(vec); free(vecp); free(f); free(fp); free(au);
  free(aup);
  return status;
}

/* ------------------------------------------------------------------------ */
/* gradient of the truncated PCG (finite_pcg)                               */
/* ------------------------------------------------------------------------ */
/* The polarization energy is linear in the dipoles the solve returns,
 * E = -(el/2) u_K . F_p.  With u_K the K-th PCG iterate (K fixed), the
 * reverse sweep through the recorded iterations, seeded with g = F_p, gives
 * the adjoint of the direct field, Fhat = alpha (g + T rbar_0), and pairs of
 * vectors (a, b) whose tensor derivatives a . dT b (Ewald field operator T,
 * A = 1/alpha - T) and a . dM b (preconditioner M = alpha (T_pre alpha + s))
 * complete the gradient:
 *   dE = DirectGrad(u_K, Fhat) - (el/2) [sum_k -Qbar_k . dT p_k
 *        + rbar_0 . dT u_0 + sum_k Zbar_k . dM r_k],
 * where DirectGrad is the gradient of direct polarization with uCR replaced
 * by Fhat (u . dF_p + Fhat . dF_d, no mutual term). */
typedef struct { double *a, *b; double w; int pre; } bipair_t;

static void apply_T(ctx_t *c, const double *v, double *out) {
  size_t sz = 3 * (size_t)c->n * sizeof(double);
  double *fp = malloc(sz);
  induced_field(c, v, v, out, fp);
  free(fp);
}

/* M v = alpha (T_pre (alpha v) + s v) */
static void apply_M(ctx_t *c, const double *v, double *out) {
  int n = c->n;
  const double *pol = c->s->polarity;
  size_t sz = 3 * (size_t)n * sizeof(double);
  double *av = malloc(sz), *f = calloc(3 * n, sizeof(double)), *fp = calloc(3 * n, sizeof(double));
  for (int q = 0; q < 3 * n; q++) av[q] = pol[q / 3] * v[q];
  induced_real_field(c, c->pc_off, c->pc, 0.0, av, av, f, fp);
  for (int q = 0; q < 3 * n; q++) out[q] = pol[q / 3] * (f[q] + c->prm->precond_scale * v[q]);
  free(av); free(f); free(fp);
}

/* A v = v / alpha - T v on the polarizable atoms, 0 elsewhere */
static void apply_A(ctx_t *c, const double *v, double *out) {
  const double *pol = c->s->polarity;
  apply_T(c, v, out);
  for (int q = 0; q < 3 * c->n; q++) out[q] = pol[q / 3] > 0.0 ? v[q] / pol[q / 3] - out[q] : 0.0;
}

static double vdot(int m, const double *a, const double *b) {
  double s = 0.0;
  for (int q = 0; q < m; q++) s += a[q] * b[q];
  return s;
}

static double *vdup(int m, const double *a) {
  double *o = malloc(m * sizeof(double));
  memcpy(o, a, m * sizeof(double));
  return o;
}

/* reverse sweep; returns the pairs (count in *np) and writes Fhat */
static bipair_t *pcg_adjoint(ctx_t *c, const double *g, double *fhat, int *np) {
  const pcg_tape_t *t = c->tape;
  int n3 = 3 * c->n, K = t->K;
  const double *pol = c->s->polarity;
  size_t vs = (size_t)n3;
  double *Pb = calloc((size_t)K * vs, sizeof(double)), *Zb = calloc((size_t)K * vs, sizeof(double));
  double *Rb = calloc((size_t)(K + 1) * vs, sizeof(double)), *Gb = calloc(K + 1, sizeof(double));
  double *tmp = malloc(vs * sizeof(double)), *Qb = malloc(vs * sizeof(double));
  bipair_t *pairs = calloc(2 * K + 2, sizeof(bipair_t));
  int m = 0;
#define TV(arr, k) (t->arr + (size_t)(k) * vs)
  for (int k = K - 1; k >= 0; k--) {
    double *Rk1 = Rb + (size_t)(k + 1) * vs;
    if (k < K - 1) {
      double *Pk1 = Pb + (size_t)(k + 1) * vs, *Zk1 = Zb + (size_t)(k + 1) * vs;
      /* p_{k+1} = z_{k+1} + beta_k p_k */
      for (int q = 0; q < n3; q++) Zk1[q] += Pk1[q];
      double bb = vdot(n3, Pk1, TV(p, k));
      for (int q = 0; q < n3; q++) Pb[(size_t)k * vs + q] += t->b[k] * Pk1[q];
      /* beta_k = gam_{k+1} / gam_k */
      Gb[k + 1] += bb / t->gam[k];
      Gb[k] -= bb * t->gam[k + 1] / (t->gam[k] * t->gam[k]);
      /* gam_{k+1} = r_{k+1} . z_{k+1} */
      for (int q = 0; q < n3; q++) {
        Rk1[q] += Gb[k + 1] * TV(z, k + 1)[q];
        Zk1[q] += Gb[k + 1] * TV(r, k + 1)[q];
      }
      /* z_{k+1} = M r_{k+1} */
      apply_M(c, Zk1, tmp);
      for (int q = 0; q < n3; q++) Rk1[q] += tmp[q];
      pairs[m++] = (bipair_t){vdup(n3, Zk1), vdup(n3, TV(r, k + 1)), 1.0, 1};
    }
    /* r_{k+1} = r_k - a_k q_k;  u_{k+1} = u_k + a_k p_k */
    double *Rk = Rb + (size_t)k * vs, *Pk = Pb + (size_t)k * vs;
    for (int q = 0; q < n3; q++) Rk[q] += Rk1[q];
    double ab = -vdot(n3, Rk1, TV(q, k)) + vdot(n3, g, TV(p, k));
    for (int q = 0; q < n3; q++) {
      Qb[q] = -t->a[k] * Rk1[q];
      Pk[q] += t->a[k] * g[q];
    }
    /* a_k = gam_k / del_k,  del_k = p_k . q_k */
    Gb[k] += ab / t->del[k];
    double db = -ab * t->gam[k] / (t->del[k] * t->del[k]);
    for (int q = 0; q < n3; q++) {
      Pk[q] += db * TV(q, k)[q];
      Qb[q] += db * TV(p, k)[q];
    }
    /* q_k = A p_k */
    apply_A(c, Qb, tmp);
    for (int q = 0; q < n3; q++) Pk[q] += tmp[q];
    pairs[m++] = (bipair_t){vdup(n3, Qb), vdup(n3, TV(p, k)), -1.0, 0};
  }
  /* gam_0 = r_0 . z_0;  p_0 = z_0;  z_0 = M r_0 */
  double *R0 = Rb, *Z0 = Zb;
  for (int q = 0; q < n3; q++) {
    R0[q] += Gb[0] * TV(z, 0)[q];
    Z0[q] += Gb[0] * TV(r, 0)[q] + Pb[q];
  }
  apply_M(c, Z0, tmp);
  for (int q = 0; q < n3; q++) R0[q] += tmp[q];
  pairs[m++] = (bipair_t){vdup(n3, Z0), vdup(n3, TV(r, 0)), 1.0, 1};
  /* r_0 = T u_0;  u_0 = alpha F_d */
  apply_T(c, R0, tmp);
  for (int q = 0; q < n3; q++) fhat[q] = pol[q / 3] * (g[q] + tmp[q]);
  pairs[m++] = (bipair_t){vdup(n3, R0), vdup(n3, c->dir), 1.0, 0};
#undef TV
  free(Pb); free(Zb); free(Rb); free(Gb); free(tmp); free(Qb);
  *np = m;
  return pairs;
}

/* sum over the pairs of w grad(a . T b) (pre: w grad(alpha a . T_pre alpha b)),
 * scaled by coef, into grad: real space over the Ewald (preconditioner) list,
 * reciprocal space through the PME potentials of a and b */
static void bilinear_grad(ctx_t *c, const bipair_t *pairs, int np, double coef, double *grad) {
  int n = c->n;
  const double *pol = c->s->polarity;
  const ora_system *s = c->s;
  double **A = malloc(np * sizeof(double *)), **B = malloc(np * sizeof(double *));
  for (int j = 0; j < np; j++) {
    if (pairs[j].pre) {
      A[j] = malloc(3 * n * sizeof(double));
      B[j] = malloc(3 * n * sizeof(double));
      for (int q = 0; q < 3 * n; q++) { A[j][q] = pol[q / 3] * pairs[j].a[q]; B[j][q] = pol[q / 3] * pairs[j].b[q]; }
    } else {
      A[j] = pairs[j].a;
      B[j] = pairs[j].b;
    }
  }
  for (int pre = 0; pre < 2; pre++) {
    const int *off = pre ? c->pc_off : c->ew_off, *lst = pre ? c->pc : c->ew;
    double aw = pre ? 0.0 : c->aewald;
    for (int i = 0; i < n; i++)
      for (int e = off[i]; e < off[i + 1]; e++) {
        int k, orig;
        double R[3];
        double r2 = entry_vec(&c->cell, n, s->xyz, i, lst[e], &k, R, &orig);
        double rl = sqrt(r2);
        double bn[4], rr[5];
        ewald_bn(rl, aw, 3, bn);
        coulomb_rr(rl, 4, rr);
        thole_t th;
        thole_terms(rl, R, c->ipd[i], c->ipd[k], s->thole[i], s->thole[k], &th);
        double B1, B2, C1, C2;
        if (pre) {
          B1 = th.s3 * rr[1]; B2 = th.s5 * rr[2]; C1 = th.s3 * rr[2]; C2 = th.s5 * rr[3];
        } else {
          B1 = bn[1] - (1.0 - th.s3) * rr[1]; B2 = bn[2] - (1.0 - th.s5) * rr[2];
          C1 = bn[2] - (1.0 - th.s3) * rr[2]; C2 = bn[3] - (1.0 - th.s5) * rr[3];
        }
        double gW[3] = {0, 0, 0};
        for (int j = 0; j < np; j++) {
          if (pairs[j].pre != pre) continue;
          const double *ai = A[j] + 3 * i, *ak = A[j] + 3 * k, *bi = B[j] + 3 * i, *bk = B[j] + 3 * k;
          double S1 = dot3(ai, bk) + dot3(ak, bi);
          double air = dot3(ai, R), akr = dot3(ak, R), bir = dot3(bi, R), bkr = dot3(bk, R);
          double S2 = air * bkr + akr * bir;
          for (int d = 0; d < 3; d++) {
            double dB1 = -C1 * R[d] + rr[1] * th.d3[d], dB2 = -C2 * R[d] + rr[2] * th.d5[d];
            gW[d] += pairs[j].w * (-dB1 * S1 + dB2 * S2 + B2 * (bkr * ai[d] + air * bk[d] + bir * ak[d] + akr * bi[d]));
          }
        }
        for (int d = 0; d < 3; d++) {
          grad[3 * i + d] -= coef * gW[d];
          grad[3 * k + d] += coef * gW[d];
        }
      }
  }
  if (c->use_recip) {
    pme_t *p = &c->pme;
    double *fa = malloc(3 * n * sizeof(double)), *fb = malloc(3 * n * sizeof(double));
    double *pa = malloc(20 * n * sizeof(double)), *pb = malloc(20 * n * sizeof(double));
    for (int j = 0; j < np; j++) {
      if (pairs[j].pre) continue;
      for (int i = 0; i < n; i++) { frac_dipole(p, A[j] + 3 * i, fa + 3 * i); frac_dipole(p, B[j] + 3 * i, fb + 3 * i); }
      pme_spread(p, n, NULL, fa, fb);
      ora_convolution(p->grid, p->recip, p->K[0], p->K[1], p->K[2]);
      for (int i = 0; i < n; i++) {
        pme_gather(p, i, 0, pa + 20 * i);
        pme_gather(p, i, 1, pb + 20 * i);
      }
      /* the mutual term of the reciprocal gradient with (u, uCR) -> (a, b):
       * it is the reciprocal part of -(el/2) a . dT b */
      for (int i = 0; i < n; i++) {
        const double *u = fa + 3 * i, *v = fb + 3 * i, *fi = pa + 20 * i, *fip = pb + 20 * i;
        double gg[3];
        gg[0] = u[0] * fip[4] + v[0] * fi[4] + u[1] * fip[7] + v[1] * fi[7] + u[2] * fip[8] + v[2] * fi[8];
        gg[1] = u[0] * fip[7] + v[0] * fi[7] + u[1] * fip[5] + v[1] * fi[5] + u[2] * fip[9] + v[2] * fi[9];
        gg[2] = u[0] * fip[8] + v[0] * fi[8] + u[1] * fip[9] + v[1] * fi[9] + u[2] * fip[6] + v[2] * fi[6];
        for (int a = 0; a < 3; a++) {
          double df = 0.5 * c->prm->electric * (p->a[0][a] * gg[0] + p->a[1][a] * gg[1] + p->a[2][a] * gg[2]);
          grad[3 * i + a] += pairs[j].w * (coef / (-0.5 * c->prm->electric)) * df;
        }
      }
    }
    free(fa); free(fb); free(pa); free(pb);
  }
  for (int j = 0; j < np; j++)
    if (pairs[j].pre) { free(A[j]); free(B[j]); }
  free(A); free(B);
}

/* SOR (ParticleMeshEwald.scfBySOR, ParticleMeshEwald.java:2331-2423, with
 * the update of SORRegion.java:236-258): from the direct dipoles,
 * u <- u + w (u_direct + alpha E(u) - u) for u and uCR; the RMS step
 * 4.80321 sqrt(max(sum du^2, sum duCR^2) / n) must fall below polar_eps; a
 * rise fails the solve, and 1000 cycles are the limit. */
static int sor_solve(ctx_t *c, ora_energies *out, double w) {
  int n = c->n;
  const double *pol = c->s->polarity;
  size_t sz = 3 * (size_t)n * sizeof(double);
  double *f = malloc(sz), *fp = malloc(sz);
  memcpy(c->ind, c->dir, sz);
  memcpy(c->indp, c->dirp, sz);
  double eps = 100.0, prev;
  int iter = 0, status = 0;
  for (;;) {
    induced_field(c, c->ind, c->indp, f, fp);
    double e1 = 0.0, e2 = 0.0;
    for (int q = 0; q < 3 * n; q++) {
      double a = pol[q / 3];
      double du = w * (c->dir[q] + a * f[q] - c->ind[q]);
      double dv = w * (c->dirp[q] + a * fp[q] - c->indp[q]);
      c->ind[q] += du;
      c->indp[q] += dv;
      e1 += du * du;
      e2 += dv * dv;
    }
    iter++;
    prev = eps;
    eps = ELEC_ANG_TO_DEBYE * sqrt((e1 > e2 ? e1 : e2) / n);
    if (eps > prev) { status = 2; break; }
    if (iter >= 1000) { status = 3; break; }
    if (eps < c->prm->polar_eps) break;
  }
  out->pcg_iters = iter;
  out->pcg_eps = eps;
  induced_field(c, c->ind, c->indp, f, fp);
  free(f); free(fp);
  return status;
}

/* ------------------------------------------------------------------------ */
/* reciprocal strain derivative, grid part: for every wave vector m the      */
/* energy weight w(m) = el/2 recip(m) [|Fp|^2 + Re Fp*(Fu + Fv) (+ Re Fu* Fv  */
/* for mutual polarization)] times (2 (1 + pi^2 m^2/a^2) m_a m_b / m^2 -     */
/* delta_ab), the derivative of exp(-pi^2 m^2/a^2) / (V m^2) under strain.   */
/* Fp, Fu, Fv: transforms of the permanent, u and uCR grids; u and uCR share */
/* one complex grid and are separated with F(-m) = conj F(m) of a real grid. */
/* ------------------------------------------------------------------------ */
static void recip_grid_virial(ctx_t *c, int polar, int mutual, double *vir) {
  pme_t *p = &c->pme;
  int n = c->n, nx = p->K[0], ny = p->K[1], nz = p->K[2];
  long nt = p->ntot;
  double el = c->prm->electric, aw = c->aewald;
  double *fpm = malloc(2 * nt * sizeof(double)), *fuv = NULL;
  pme_spread(p, n, c->fmp, NULL, NULL);
  memcpy(fpm, p->grid, 2 * nt * sizeof(double));
  ora_fft3d(fpm, nx, ny, nz, 0);
  if (polar) {
    double *fu = malloc(3 * n * sizeof(double)), *fv = malloc(3 * n * sizeof(double));
    for (int i = 0; i < n; i++) {
      frac_dipole(p, c->ind + 3 * i, fu + 3 * i);
      frac_dipole(p, c->indp + 3 * i, fv + 3 * i);
    }
    pme_spread(p, n, NULL, fu, fv);
    fuv = malloc(2 * nt * sizeof(double));
    memcpy(fuv, p->grid, 2 * nt * sizeof(double));
    ora_fft3d(fuv, nx, ny, nz, 0);
    free(fu); free(fv);
  }
  double pt = (M_PI / aw) * (M_PI / aw);
  double w9[9] = {0};
  for (int kz = 0; kz < nz; kz++) {
    int l = kz >= (nz + 1) / 2 ? kz - nz : kz;
    for (int ky = 0; ky < ny; ky++) {
      int k = ky >= (ny + 1) / 2 ? ky - ny : ky;
      for (int kx = 0; kx < nx; kx++) {
        int h = kx >= (nx + 1) / 2 ? kx - nx : kx;
        long idx = kx + (long)nx * (ky + (long)ny * kz);
        double r = p->recip[idx];
        if (r == 0.0) continue;
        double m[3];
        for (int d = 0; d < 3; d++) m[d] = c->cell.A[d][0] * h + c->cell.A[d][1] * k + c->cell.A[d][2] * l;
        double m2 = dot3(m, m);
        double pr = fpm[2 * idx], pi = fpm[2 * idx + 1];
        double ws = c->ph->perm * (pr * pr + pi * pi), wq = 0.0;
        if (polar) {
          long nidx = ((nx - kx) % nx) + (long)nx * (((ny - ky) % ny) + (long)ny * ((nz - kz) % nz));
          double ar = fuv[2 * idx], ai = fuv[2 * idx + 1], br = fuv[2 * nidx], bi = -fuv[2 * nidx + 1];
          double ur = 0.5 * (ar + br), ui = 0.5 * (ai + bi);   /* Fu = (F(m) + conj F(-m)) / 2 */
          double vr = 0.5 * (ai - bi), vi = -0.5 * (ar - br);  /* Fv = (F(m) - conj F(-m)) / 2i */
          wq = pr * (ur + vr) + pi * (ui + vi);
          if (mutual) wq += ur * vr + ui * vi;
          ws += c->ph->pol * wq;
        }
        double w = 0.5 * el * r * ws, f = 2.0 * (1.0 + pt * m2) / m2;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) w9[3 * a + b] += w * (f * m[a] * m[b] - (a == b ? 1.0 : 0.0));
      }
    }
  }
  for (int q = 0; q < 9; q++) vir[q] += w9[q];
  free(fpm);
  free(fuv);
}

/* ------------------------------------------------------------------------ */
/* reciprocal energy and gradient: ReciprocalEnergyRegion.java:381-746      */
/* ------------------------------------------------------------------------ */
static void recip_energy(ctx_t *c, ora_energies *out, int polar) {
  int n = c->n;
  double el = c->prm->electric, aw = c->aewald;
  pme_t *p = &c->pme;
  double aewald1 = -el * aw / SQRT_PI, aewald2 = 2.0 * aw * aw;
  double aewald3 = -2.0 / 3.0 * el * aw * aw * aw / SQRT_PI, aewald4 = -2.0 * aewald3;
  const phase_t *phs = c->ph;
  double eself = 0, erecip = 0, qtot = 0;
  for (int i = 0; i < n; i++) {
    const double *m = c->gmp + 10 * i;
    qtot += m[0];
    double cii = m[0] * m[0];
    double dii = m[1] * m[1] + m[2] * m[2] + m[3] * m[3];
    double qii = m[4] * m[4] + m[5] * m[5] + m[6] * m[6] + 2.0 * (m[7] * m[7] + m[8] * m[8] + m[9] * m[9]);
    eself += aewald1 * (cii + aewald2 * (dii / 3.0 + 2.0 * aewald2 * qii / 45.0));
  }
  /* gradient of a fractional-multipole potential: (g . A) with g scaled by K */
  for (int i = 0; i < n; i++) {
    const double *m = c->gmp + 10 * i, *ph = c->phi_p_c + 10 * i, *fm = c->fmp + 10 * i, *fp = c->phi_p_f + 20 * i;
    erecip += m[0] * ph[0] + m[1] * ph[1] + m[2] * ph[2] + m[3] * ph[3] +
              (m[4] * ph[4] + m[5] * ph[5] + m[6] * ph[6] + 2.0 * (m[7] * ph[7] + m[8] * ph[8] + m[9] * ph[9])) / 3.0;
    double g[3];
    g[0] = fm[0] * fp[1] + fm[1] * fp[4] + fm[2] * fp[7] + fm[3] * fp[8] + fm[4] * fp[10] + fm[5] * fp[15] +
           fm[6] * fp[17] + fm[7] * fp[13] + fm[8] * fp[14] + fm[9] * fp[19];
    g[1] = fm[0] * fp[2] + fm[1] * fp[7] + fm[2] * fp[5] + fm[3] * fp[9] + fm[4] * fp[13] + fm[5] * fp[11] +
           fm[6] * fp[18] + fm[7] * fp[15] + fm[8] * fp[19] + fm[9] * fp[16];
    g[2] = fm[0] * fp[3] + fm[1] * fp[8] + fm[2] * fp[9] + fm[3] * fp[6] + fm[4] * fp[14] + fm[5] * fp[16] +
           fm[6] * fp[12] + fm[7] * fp[19] + fm[8] * fp[17] + fm[9] * fp[18];
    double df[3];
    for (int a = 0; a < 3; a++) df[a] = p->a[0][a] * g[0] + p->a[1][a] * g[1] + p->a[2][a] * g[2];
    double tq[3];
    tq[0] = -m[2] * ph[3] + m[3] * ph[2];
    tq[1] = -m[3] * ph[1] + m[1] * ph[3];
    tq[2] = -m[1] * ph[2] + m[2] * ph[1];
    const double tt = 2.0 / 3.0;
    tq[0] -= tt * (m[7] * ph[8] + m[5] * ph[9] + m[9] * ph[6] - m[8] * ph[7] - m[9] * ph[5] - m[6] * ph[9]);
    tq[1] -= tt * (m[8] * ph[4] + m[9] * ph[7] + m[6] * ph[8] - m[4] * ph[8] - m[7] * ph[9] - m[8] * ph[6]);
    tq[2] -= tt * (m[4] * ph[7] + m[7] * ph[5] + m[8] * ph[9] - m[7] * ph[4] - m[5] * ph[7] - m[9] * ph[8]);
    for (int a = 0; a < 3; a++) {
      c->grad[3 * i + a] += phs->perm * el * df[a];
      c->torque[3 * i + a] += phs->perm * el * tq[a];
      c->lgrad[3 * i + a] += phs->dperm * el * df[a];
      c->ltorque[3 * i + a] += phs->dperm * el * tq[a];
    }
  }
  double ecc = 0.0;
  if (qtot != 0.0) ecc = -0.5 * el * M_PI * qtot * qtot / (c->cell.vol * aw * aw);
  double eperm = eself + 0.5 * el * erecip + ecc;
  out->perm_self += phs->perm * eself;
  out->perm_recip += phs->perm * 0.5 * el * erecip;
  out->perm_charge_corr += phs->perm * ecc;
  out->dudl += phs->dperm * eperm;
  out->d2udl2 += phs->d2perm * eperm;
  /* strain derivative: the charge correction scales as 1/V; with the global
   * multipoles held fixed, the wave vectors' strain dependence gives
   * -d_b dphi/dx_a - (2/3) Q_bc d2phi/dx_a dx_c per atom (Cartesian
   * multipoles, potentials of the permanent grid) */
  for (int a = 0; a < 3; a++) out->vir[4 * a] -= phs->perm * ecc;
  for (int i = 0; i < n; i++) {
    const double *m = c->gmp + 10 * i, *ph = c->phi_p_c + 10 * i;
    double Q[3][3], H[3][3];
    sym_full(m + 4, Q);
    sym_full(ph + 4, H);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double t = m[1 + b] * ph[1 + a];
        for (int cc = 0; cc < 3; cc++) t += 2.0 / 3.0 * Q[b][cc] * H[a][cc];
        out->vir[3 * a + b] -= phs->perm * el * t;
      }
  }
  if (!polar) return;
  /* polarization: permanent multipoles in the potentials of u + uCR, u + uCR
   * in the permanent potential, and (mutual) u in the potential of uCR and
   * uCR in that of u; each with weight 1/2 */
  for (int i = 0; i < n; i++) {
    const double *m = c->gmp + 10 * i, *pp = c->phi_p_c + 10 * i;
    const double *u = c->ind + 3 * i, *v = c->indp + 3 * i;
    const double *cu = c->phi_u_c + 10 * i, *cv = c->phi_pp_c + 10 * i;
    double Q[3][3], Hw[3][3], hw[6];
    sym_full(m + 4, Q);
    for (int t = 0; t < 6; t++) hw[t] = cu[4 + t] + cv[4 + t];
    sym_full(hw, Hw);
    int mutual = c->prm->polarization == ORA_POL_MUTUAL && !c->finite_grad;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double t = (u[b] + v[b]) * pp[1 + a] + m[1 + b] * (cu[1 + a] + cv[1 + a]);
        for (int cc = 0; cc < 3; cc++) t += 2.0 / 3.0 * Q[b][cc] * Hw[a][cc];
        if (mutual) t += v[b] * cu[1 + a] + u[b] * cv[1 + a];
        out->vir[3 * a + b] -= phs->pol * 0.5 * el * t;
      }
  }
  /* induced dipoles */
  double eselfi = 0, erecipi = 0;
  for (int i = 0; i < n; i++) {
    const double *m = c->gmp + 10 * i, *u = c->ind + 3 * i, *up = c->indp + 3 * i;
    eselfi += aewald3 * (u[0] * m[1] + u[1] * m[2] + u[2] * m[3]);
    double us[3] = {0.5 * (u[0] + up[0]), 0.5 * (u[1] + up[1]), 0.5 * (u[2] + up[2])};
    double t[3];
    cross3(m + 1, us, t);
    axpy3(phs->pol * aewald4, t, c->torque + 3 * i);
    axpy3(phs->dpol * aewald4, t, c->ltorque + 3 * i);
  }
  for (int i = 0; i < n; i++) {
    const double *m = c->gmp + 10 * i, *fm = c->fmp + 10 * i, *fp = c->phi_p_f + 20 * i;
    double fu[3], fpp[3];
    frac_dipole(p, c->ind + 3 * i, fu);
    frac_dipole(p, c->indp + 3 * i, fpp);
    erecipi += fu[0] * fp[1] + fu[1] * fp[2] + fu[2] * fp[3];
    const double *fi = c->phi_u_f + 20 * i, *fip = c->phi_pp_f + 20 * i;
    const double *ci = c->phi_u_c + 10 * i, *cip = c->phi_pp_c + 10 * i;
    double sp[10], sf[20];
    for (int t = 0; t < 10; t++) sp[t] = 0.5 * (ci[t] + cip[t]);
    for (int t = 0; t < 20; t++) sf[t] = fi[t] + fip[t];
    double s[3] = {fu[0] + fpp[0], fu[1] + fpp[1], fu[2] + fpp[2]};
    double g[3];
    g[0] = s[0] * fp[4] + s[1] * fp[7] + s[2] * fp[8];
    g[1] = s[0] * fp[7] + s[1] * fp[5] + s[2] * fp[9];
    g[2] = s[0] * fp[8] + s[1] * fp[9] + s[2] * fp[6];
    if (c->prm->polarization == ORA_POL_MUTUAL && !c->finite_grad) {
      g[0] += fu[0] * fip[4] + fpp[0] * fi[4] + fu[1] * fip[7] + fpp[1] * fi[7] + fu[2] * fip[8] + fpp[2] * fi[8];
      g[1] += fu[0] * fip[7] + fpp[0] * fi[7] + fu[1] * fip[5] + fpp[1] * fi[5] + fu[2] * fip[9] + fpp[2] * fi[9];
      g[2] += fu[0] * fip[8] + fpp[0] * fi[8] + fu[1] * fip[9] + fpp[1] * fi[9] + fu[2] * fip[6] + fpp[2] * fi[6];
    }
    g[0] += fm[0] * sf[1] + fm[1] * sf[4] + fm[2] * sf[7] + fm[3] * sf[8] + fm[4] * sf[10] + fm[5] * sf[15] +
            fm[6] * sf[17] + fm[7] * sf[13] + fm[8] * sf[14] + fm[9] * sf[19];
    g[1] += fm[0] * sf[2] + fm[1] * sf[7] + fm[2] * sf[5] + fm[3] * sf[9] + fm[4] * sf[13] + fm[5] * sf[11] +
            fm[6] * sf[18] + fm[7] * sf[15] + fm[8] * sf[19] + fm[9] * sf[16];
    g[2] += fm[0] * sf[3] + fm[1] * sf[8] + fm[2] * sf[9] + fm[3] * sf[6] + fm[4] * sf[14] + fm[5] * sf[16] +
            fm[6] * sf[12] + fm[7] * sf[19] + fm[8] * sf[17] + fm[9] * sf[18];
    double df[3];
    for (int a = 0; a < 3; a++) df[a] = 0.5 * el * (p->a[0][a] * g[0] + p->a[1][a] * g[1] + p->a[2][a] * g[2]);
    double tq[3];
    tq[0] = -m[2] * sp[3] + m[3] * sp[2];
    tq[1] = -m[3] * sp[1] + m[1] * sp[3];
    tq[2] = -m[1] * sp[2] + m[2] * sp[1];
    const double tt = 2.0 / 3.0;
    tq[0] -= tt * (m[7] * sp[8] + m[5] * sp[9] + m[9] * sp[6] - m[8] * sp[7] - m[9] * sp[5] - m[6] * sp[9]);
    tq[1] -= tt * (m[8] * sp[4] + m[9] * sp[7] + m[6] * sp[8] - m[4] * sp[8] - m[7] * sp[9] - m[8] * sp[6]);
    tq[2] -= tt * (m[4] * sp[7] + m[7] * sp[5] + m[8] * sp[9] - m[7] * sp[4] - m[5] * sp[7] - m[9] * sp[8]);
    for (int a = 0; a < 3; a++) {
      c->grad[3 * i + a] += phs->pol * df[a];
      c->torque[3 * i + a] += phs->pol * el * tq[a];
      c->lgrad[3 * i + a] += phs->dpol * df[a];
      c->ltorque[3 * i + a] += phs->dpol * el * tq[a];
    }
  }
  out->pol_self += phs->pol * eselfi;
  out->pol_recip += phs->pol * 0.5 * el * erecipi;
  out->dudl += phs->dpol * (eselfi + 0.5 * el * erecipi);
  out->d2udl2 += phs->d2pol * (eselfi + 0.5 * el * erecipi);
}

/* ------------------------------------------------------------------------ */
/* real-space energy and gradient: RealSpaceEnergyRegion.realSpaceChunk      */
/* (:803-1183) over the Ewald list with AMOEBA masks                         */
/* ------------------------------------------------------------------------ */
static void real_energy(ctx_t *c, ora_energies *out, int polar) {
  int n = c->n, nth = c->nth;
  const ora_system *s = c->s;
  const ora_params *prm = c->prm;
  const phase_t *ph = c->ph;
  double el = prm->electric;
  double *acc = calloc((size_t)nth * 12 * n, sizeof(double));
  double ep = 0.0, ei = 0.0, dl1 = 0.0, dl2 = 0.0;
  int cnt = 0;
  double w[9] = {0};
  #pragma omp parallel num_threads(nth) reduction(+ : ep, ei, dl1, dl2, cnt, w[:9])
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *ag = acc + (size_t)tid * 12 * n, *at = ag + 3 * n, *lg = ag + 6 * n, *lt = ag + 9 * n;
    double *mm = malloc(n * sizeof(double)), *pm = malloc(n * sizeof(double));
    for (int k = 0; k < n; k++) { mm[k] = 1.0; pm[k] = 1.0; }
    #pragma omp for schedule(dynamic, 32)
    for (int i = 0; i < n; i++) {
      apply_masks(s, prm, i, mm, pm, NULL);
      const double *xi = s->xyz + 3 * i;
      mp3_t I;
      mp3_load(c->gmp + 10 * i, &I);
      for (int t = c->ew_off[i]; t < c->ew_off[i + 1]; t++) {
        int k, orig;
        double r[3];
        double r2 = entry_vec(&c->cell, n, s->xyz, i, c->ew[t], &k, r, &orig);
        int soft = ph->soft_pairs && s->soft && (s->soft[i] || s->soft[k]);
        if (ph->vapor && s->soft && s->soft[i] && s->soft[k]) {
          /* vapor phase: both-soft pairs that the condensed phase softens
           * are left out (RealSpaceEnergyRegion.java:864-870) */
          int same = s->molecule ? s->molecule[i] == s->molecule[k] : 0;
          if ((prm->intermolecular_softcore && !same) || (prm->intramolecular_softcore && same)) continue;
        }
        double rl = sqrt(r2);
        mp3_t K;
        mp3_load(c->gmp + 10 * k, &K);
        double bn[7], rr[7];
        ewald_bn(rl, c->aewald, 6, bn);
        coulomb_rr(rl, 6, rr);
        if (ph->perm_real) {
          /* permanent pair; a soft pair sees r_eff = sqrt(r^2 + lalpha) in
           * the radial functions and the unshifted r in the angular ones */
          double sb[7], sr[7];
          const double *B = bn, *R = rr;
          double l2 = 1.0;
          if (soft) {
            double re = sqrt(r2 + ph->lalpha);
            ewald_bn(re, c->aewald, 6, sb);
            coulomb_rr(re, 6, sr);
            B = sb; R = sr;
            l2 = ph->l2;
          }
          permgeo_t g;
          perm_geometry(r, &I, &K, &g);
          double f1[3], t1[3], u1[3], f2[3], t2[3], u2[3];
          double e = perm_terms(r, &I, &K, &g, B, f1, t1, u1);
          double s1 = orig ? 1.0 - mm[k] : 0.0;
          if (s1 != 0.0) {
            e -= s1 * perm_terms(r, &I, &K, &g, R, f2, t2, u2);
            for (int a = 0; a < 3; a++) { f1[a] -= s1 * f2[a]; t1[a] -= s1 * t2[a]; u1[a] -= s1 * u2[a]; }
          }
          if (e != 0.0) cnt++;
          ep += l2 * e;
          /* pair strain derivative: dE/dr = -(i-side gradient), r = x_k - x_i */
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) w[3 * a + b] -= el * l2 * f1[a] * r[b];
          for (int a = 0; a < 3; a++) {
            ag[3 * i + a] += el * l2 * f1[a];
            ag[3 * k + a] -= el * l2 * f1[a];
            at[3 * i + a] += el * l2 * t1[a];
            at[3 * k + a] += el * l2 * u1[a];
          }
          if (soft) {
            /* d(radial)/dlambda: bn_j -> bn_{j+1} times dlalpha, since
             * d bn_j / d(r^2) = -bn_{j+1}/2 and d lalpha/dlambda = -2 dlalpha
             * (RealSpaceEnergyRegion.java:1477-1690) */
            double fr[3], tr[3], ur[3], fq[3], tq2[3], uq[3];
            double dR = perm_terms(r, &I, &K, &g, B + 1, fr, tr, ur);
            double d2R = perm_terms(r, &I, &K, &g, B + 2, NULL, NULL, NULL);
            if (s1 != 0.0) {
              dR -= s1 * perm_terms(r, &I, &K, &g, R + 1, fq, tq2, uq);
              d2R -= s1 * perm_terms(r, &I, &K, &g, R + 2, NULL, NULL, NULL);
              for (int a = 0; a < 3; a++) { fr[a] -= s1 * fq[a]; tr[a] -= s1 * tq2[a]; ur[a] -= s1 * uq[a]; }
            }
            dl1 += ph->dperm * e + l2 * ph->dlalpha * dR;
            dl2 += ph->d2perm * e + 2.0 * ph->dperm * ph->dlalpha * dR + l2 * ph->d2lalpha * dR +
                   l2 * ph->dlalpha * ph->dlalpha * d2R;
            double ca = ph->dperm, cb = l2 * ph->dlalpha;
            for (int a = 0; a < 3; a++) {
              double fl = el * (ca * f1[a] + cb * fr[a]);
              lg[3 * i + a] += fl;
              lg[3 * k + a] -= fl;
              lt[3 * i + a] += el * (ca * t1[a] + cb * tr[a]);
              lt[3 * k + a] += el * (ca * u1[a] + cb * ur[a]);
            }
          }
        }
        if (polar) {
          polctx_t cx;
          cx.bn = bn; cx.rr = rr;
          thole_terms(rl, r, c->ipd[i], c->ipd[k], s->thole[i], s->thole[k], &cx.th);
          cx.scalep = orig ? pm[k] : 1.0;
          cx.scaled = (orig && s->group[k] == s->group[i]) ? prm->d11 : 1.0;
          cx.direct = (prm->polarization == ORA_POL_DIRECT) || c->finite_grad;
          double f[3], a2[3], a3[3];
          double e = pol_terms(r, &I, &K, c->ind + 3 * i, c->indp + 3 * i, c->ind + 3 * k, c->indp + 3 * k, &cx, 1,
                               f, a2, a3);
          ei += ph->pol * e;
          dl1 += ph->dpol * e;
          dl2 += ph->d2pol * e;
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) w[3 * a + b] -= el * ph->pol * f[a] * r[b];
          for (int a = 0; a < 3; a++) {
            ag[3 * i + a] += el * ph->pol * f[a];
            ag[3 * k + a] -= el * ph->pol * f[a];
            at[3 * i + a] += el * ph->pol * a2[a];
            at[3 * k + a] += el * ph->pol * a3[a];
            lg[3 * i + a] += el * ph->dpol * f[a];
            lg[3 * k + a] -= el * ph->dpol * f[a];
            lt[3 * i + a] += el * ph->dpol * a2[a];
            lt[3 * k + a] += el * ph->dpol * a3[a];
          }
        }
      }
      remove_masks(s, i, mm, pm);
    }
    free(mm); free(pm);
  }
  for (int t = 0; t < nth; t++)
    for (int q = 0; q < 3 * n; q++) {
      const double *a = acc + (size_t)t * 12 * n;
      c->grad[q] += a[q];
      c->torque[q] += a[3 * n + q];
      c->lgrad[q] += a[6 * n + q];
      c->ltorque[q] += a[9 * n + q];
    }
  free(acc);
  out->perm_real += el * ep;
  out->pol_real += el * ei;
  out->dudl += el * dl1;
  out->d2udl2 += el * dl2;
  out->n_perm += cnt;
  for (int q = 0; q < 9; q++) out->vir[q] += w[q];
}

/* ------------------------------------------------------------------------ */
/* buffered 14-7 vdW: VanDerWaals.VanDerWaalsLoop (VanDerWaals.java:1638-1900)*/
/* with VanDerWaalsForm combining rules and MultiplicativeSwitch taper       */
/* ------------------------------------------------------------------------ */
void ora_reduced_xyz(const ora_system *s, double *red) {
  for (int i = 0; i < s->n; i++) {
    int h = s->red_index[i];
    double a = s->red_value[i];
    for (int d = 0; d < 3; d++) {
      double x = s->xyz[3 * i + d];
      if (h >= 0 && h != i && a > 0.0) {
        double rx = s->xyz[3 * h + d];
        red[3 * i + d] = a * (x - rx) + rx;
      } else {
        red[3 * i + d] = x;
      }
    }
  }
}

/* Softcore factors of VanDerWaals.setLambda (:599-653): lambda is rescaled by
 * the end point, sc1 = alpha (1 - L)^2 shifts both buffered denominators and
 * sc2 = L^beta scales eps; derivatives are with respect to the rescaled L. */
typedef struct { double sc1, dsc1, d2sc1, sc2, dsc2, d2sc2; } softcore_t;

static softcore_t softcore_factors(const ora_params *prm) {
  softcore_t f;
  double L = prm->lambda, end = prm->vdw_lambda_end > 0.0 ? prm->vdw_lambda_end : 1.0;
  L = L > end ? 1.0 : L / end;
  double a = prm->vdw_lambda_alpha, b = prm->vdw_lambda_exponent;
  f.sc1 = a * (1.0 - L) * (1.0 - L);
  f.dsc1 = -2.0 * a * (1.0 - L);
  f.d2sc1 = 2.0 * a;
  f.sc2 = pow(L, b);
  f.dsc2 = b * pow(L, b - 1.0);
  f.d2sc2 = b >= 2.0 ? b * (b - 1.0) * pow(L, b - 2.0) : 0.0;
  return f;
}

/* softcore pair: exactly one atom soft, or both soft under the inter- /
 * intramolecular flags (VanDerWaals.java:1689-1737) */
static int soft_pair(const ora_system *s, const ora_params *prm, int i, int k) {
  if (!s->soft) return 0;
  int si = s->soft[i] != 0, sk = s->soft[k] != 0;
  if (si != sk) return 1;
  if (si && sk) {
    int same = s->molecule[i] == s->molecule[k];
    return (prm->intermolecular_softcore && !same) || (prm->intramolecular_softcore && same);
  }
  return 0;
}

static double vdw_energy(ctx_t *c, double *grad, int *count, double *vir, double *dudl, double *d2udl2,
                         double *lgrad) {
  const ora_system *s = c->s;
  const ora_params *prm = c->prm;
  int n = c->n, nth = c->nth;
  double *red = malloc(3 * n * sizeof(double));
  ora_reduced_xyz(s, red);
  double off = prm->vdw_cutoff, cut = prm->vdw_taper * off;
  double off2 = off * off, cut2 = cut * cut;
  /* quintic switch f(cut) = 1, f(off) = 0 */
  double a = cut, b = off, a2 = a * a, b2 = b * b, den = pow(b - a, 5.0);
  double c0 = b * b2 * (b2 - 5.0 * a * b + 10.0 * a2) / den, c1 = -30.0 * a2 * b2 / den;
  double c2 = 30.0 * b * a * (b + a) / den, c3 = -10.0 * (a2 + 4.0 * a * b + b2) / den;
  double c4 = 15.0 * (a + b) / den, c5 = -6.0 / den;
  double dl = prm->halgren_delta, gm = prm->halgren_gamma;
  double t1n = pow(1.0 + dl, 7.0), g1 = 1.0 + gm;
  softcore_t sf = softcore_factors(prm);
  double e = 0.0, el1 = 0.0, el2 = 0.0;
  int cnt = 0;
  double w[9] = {0};
  double *acc = calloc((size_t)nth * 6 * n, sizeof(double));
  #pragma omp parallel num_threads(nth) reduction(+ : e, cnt, w[:9], el1, el2)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double *ag = acc + (size_t)tid * 6 * n, *al = ag + 3 * n;
    double *mask = malloc(n * sizeof(double));
    for (int k = 0; k < n; k++) mask[k] = 1.0;
    #pragma omp for schedule(dynamic, 32)
    for (int i = 0; i < n; i++) {
      for (int t = s->mask_off[i]; t < s->mask_off[i + 1]; t++) {
        int rel = s->mask_rel[t];
        if (rel >= 2 && rel <= 4) mask[s->mask_atom[t]] = prm->v_scale[rel - 2];
      }
      int ci = s->vdw_class[i];
      int ri = s->red_index[i];
      double rv_i = s->red_value[i];
      for (int tt = c->nl.off[i]; tt < c->nl.off[i + 1]; tt++) {
        int k, orig;
        double d[3];
        double r2 = entry_vec(&c->cell, n, red, i, c->nl.nbr[tt], &k, d, &orig);
        scale3(-1.0, d);   /* d = red_i - (red_k + T) */
        int ck = s->vdw_class[k];
        const double *pe = s->vdw_table + 2 * (ci * s->n_class + ck);
        double irv = pe[0];
        double mk = orig ? mask[k] : 1.0;
        if (!(r2 <= off2 && mk > 0.0 && irv > 0.0)) continue;
        double r = sqrt(r2);
        int soft = soft_pair(s, prm, i, k);
        double sc1 = soft ? sf.sc1 : 0.0, sc2 = soft ? sf.sc2 : 1.0;
        double ev = mk * pe[1];
        double rho = r * irv;
        double rho6 = pow(rho, 6.0), rho7 = rho6 * rho;
        double rd = rho + dl, rd6 = pow(rd, 6.0), rd7 = rd6 * rd;
        double t1d = 1.0 / (sc1 + rd7), t2d = 1.0 / (sc1 + rho7 + gm);
        double t1 = t1n * t1d, t2a = g1 * t2d, t2 = t2a - 2.0;
        double eik = sc2 * ev * t1 * t2;
        double tp = 1.0, dtp = 0.0;
        if (r2 > cut2) {
          double r3 = r2 * r, r4 = r2 * r2, r5 = r2 * r3;
          tp = c5 * r5 + c4 * r4 + c3 * r3 + c2 * r2 + c1 * r + c0;
          dtp = 5.0 * c5 * r4 + 4.0 * c4 * r3 + 3.0 * c3 * r2 + 2.0 * c2 * r + c1;
        }
        e += eik * tp;
        cnt++;
        double dt1 = t1 * 7.0 * rd6 * irv * t1d;   /* -d t1 / dr */
        double dt2 = t2a * 7.0 * rho6 * irv * t2d; /* -d t2 / dr */
        double ir = 1.0 / r;
        double dldr = 0.0;
        if (soft) {
          /* dU/dL, d2U/dL2 and d(dU/dL)/dr of VanDerWaals.java:1842-1881 */
          double lt1 = -t1 * t1d * sf.dsc1, lt2 = -t2a * t2d * sf.dsc1;
          double dedl = ev * (sf.dsc2 * t1 * t2 + sf.sc2 * lt1 * t2 + sf.sc2 * t1 * lt2);
          el1 += dedl * tp;
          double t1d2 = -sf.dsc1 * t1d * t1d, t2d2 = -sf.dsc1 * t2d * t2d;
          double d2t1 = -lt1 * t1d * sf.dsc1 - t1 * t1d * sf.d2sc1 - t1 * t1d2 * sf.dsc1;
          double d2t2 = -lt2 * t2d * sf.dsc1 - t2a * t2d * sf.d2sc1 - t2a * t2d2 * sf.dsc1;
          double df1 = sf.d2sc2 * t1 * t2 + sf.dsc2 * lt1 * t2 + sf.dsc2 * t1 * lt2;
          double df2 = sf.dsc2 * lt1 * t2 + sf.sc2 * d2t1 * t2 + sf.sc2 * lt1 * lt2;
          double df3 = sf.dsc2 * t1 * lt2 + sf.sc2 * lt1 * lt2 + sf.sc2 * t1 * d2t2;
          el2 += ev * (df1 + df2 + df3) * tp;
          double a1 = -sf.dsc2 * t2 * dt1, a2 = -sf.dsc2 * t1 * dt2;
          double a3 = 2.0 * sf.sc2 * t2 * dt1 * sf.dsc1 * t1d, a4 = 2.0 * sf.sc2 * t1 * dt2 * sf.dsc1 * t2d;
          double a5 = -sf.sc2 * lt2 * dt1, a6 = -sf.sc2 * lt1 * dt2;
          double dedldr = ev * (a1 + a2 + a3 + a4 + a5 + a6);
          dldr = dedl * dtp + dedldr * tp;
        }
        if (!grad) continue;
        double dedr = -sc2 * ev * (dt1 * t2 + t1 * dt2);
        double ds = eik * dtp + dedr * tp;
        int rk = s->red_index[k];
        double rv_k = s->red_value[k];
        /* pair strain derivative (dE/dr)_a r_b with r = red_k - red_i = -d */
        for (int a2 = 0; a2 < 3; a2++)
          for (int b2 = 0; b2 < 3; b2++) w[3 * a2 + b2] += ds * ir * d[a2] * d[b2];
        for (int q = 0; q < 3; q++) {
          double de = ds * d[q] * ir, dl2 = dldr * d[q] * ir;
          if (ri >= 0 && ri != i && rv_i > 0.0) {
            ag[3 * i + q] += de * rv_i;
            ag[3 * ri + q] += de * (1.0 - rv_i);
            al[3 * i + q] += dl2 * rv_i;
            al[3 * ri + q] += dl2 * (1.0 - rv_i);
          } else {
            ag[3 * i + q] += de;
            al[3 * i + q] += dl2;
          }
          if (rk >= 0 && rk != k && rv_k > 0.0) {
            ag[3 * k + q] -= de * rv_k;
            ag[3 * rk + q] -= de * (1.0 - rv_k);
            al[3 * k + q] -= dl2 * rv_k;
            al[3 * rk + q] -= dl2 * (1.0 - rv_k);
          } else {
            ag[3 * k + q] -= de;
            al[3 * k + q] -= dl2;
          }
        }
      }
      for (int t = s->mask_off[i]; t < s->mask_off[i + 1]; t++) mask[s->mask_atom[t]] = 1.0;
    }
    free(mask);
  }
  if (grad)
    for (int t = 0; t < nth; t++)
      for (int q = 0; q < 3 * n; q++) {
        grad[q] += acc[(size_t)t * 6 * n + q];
        if (lgrad) lgrad[q] += acc[(size_t)t * 6 * n + 3 * n + q];
      }
  free(acc);
  free(red);
  *count = cnt;
  if (dudl) *dudl += el1;
  if (d2udl2) *d2udl2 += el2;
  if (vir)
    for (int q = 0; q < 9; q++) vir[q] += w[q];
  return e;
}

/* ------------------------------------------------------------------------ */
/* exact Ewald reciprocal sum for pinning PME                               */
/* ------------------------------------------------------------------------ */
double ora_ewald_recip_exact(const ora_system *s, const ora_params *prm, int kmax) {
  cell_t c;
  cell_init(&c, s->box);
  int n = s->n;
  double *g = malloc(10 * n * sizeof(double));
  ora_rotate_multipoles(s, g);
  double aw = prm->ewald_alpha, e = 0.0;
  for (int h = -kmax; h <= kmax; h++)
    for (int k = -kmax; k <= kmax; k++)
      for (int l = -kmax; l <= kmax; l++) {
        if (!h && !k && !l) continue;
        double m[3];
        for (int d = 0; d < 3; d++) m[d] = c.A[d][0] * h + c.A[d][1] * k + c.A[d][2] * l;
        double m2 = dot3(m, m);
        double w = exp(-M_PI * M_PI * m2 / (aw * aw)) / m2;
        if (w < 1e-30) continue;
        double sr = 0.0, si = 0.0;
        for (int a = 0; a < n; a++) {
          const double *x = s->xyz + 3 * a, *q = g + 10 * a;
          /* structure factor of a multipole: q + 2 pi i m.d - (4 pi^2/3) m.Q.m (Q full) */
          double qm[3];
          symv(q + 4, m, qm);
          double re = q[0] - (4.0 * M_PI * M_PI / 3.0) * dot3(m, qm);
          double im = 2.0 * M_PI * dot3(m, q + 1);
          double ph = 2.0 * M_PI * dot3(m, x);
          double cs = cos(ph), sn = sin(ph);
          sr += re * cs - im * sn;
          si += re * sn + im * cs;
        }
        e += w * (sr * sr + si * si);
      }
  free(g);
  return prm->electric * e / (2.0 * M_PI * c.vol);
}

/* ------------------------------------------------------------------------ */
/* driver: ForceFieldEnergy.energy -> VanDerWaals.energy, PME.energy         */
/* ------------------------------------------------------------------------ */
int ora_evaluate(const ora_system *s, const ora_params *prm, ora_energies *out, double *grad, double *ind,
                 double *indcr) {
  return ora_evaluate_lambda(s, prm, out, grad, ind, indcr, NULL);
}

static void ctx_open(ctx_t *c, const ora_system *s, const ora_params *prm, const phase_t *ph) {
  memset(c, 0, sizeof *c);
  c->s = s; c->prm = prm; c->ph = ph; c->n = s->n;
  c->nth = nthreads_of(prm);
  int n = s->n;
  cell_init(&c->cell, s->box);
  if (!c->cell.aperiodic) {
    double cmax = prm->ewald_cutoff > prm->vdw_cutoff ? prm->ewald_cutoff : prm->vdw_cutoff;
    for (int d = 0; d < 3; d++)
      if (0.5 * cell_width(&c->cell, d) < cmax) c->cell.multi = 1;
  }
  c->aewald = c->cell.aperiodic ? 0.0 : prm->ewald_alpha;
  c->off2 = prm->ewald_cutoff * prm->ewald_cutoff;
  c->grad = calloc(3 * n, sizeof(double));
  c->torque = calloc(3 * n, sizeof(double));
  c->lgrad = calloc(3 * n, sizeof(double));
  c->ltorque = calloc(3 * n, sizeof(double));
  double lcut = (prm->vdw_cutoff > prm->ewald_cutoff ? prm->vdw_cutoff : prm->ewald_cutoff) + prm->list_buffer;
  nlist_build(&c->nl, &c->cell, n, s->xyz, lcut);
}

static void ctx_close(ctx_t *c) {
  free(c->grad); free(c->torque); free(c->lgrad); free(c->ltorque);
  nlist_free(&c->nl);
}

/* torques to frame-atom gradients (ReduceRegion.TorqueLoop), for the energy
 * gradient (with the frame-rotation virial) and for dU/dX/dlambda */
static int reduce_torques(ctx_t *c, ora_energies *out) {
  const ora_system *s = c->s;
  int status = 0;
  for (int pass = 0; pass < 2; pass++) {
    double *tqa = pass ? c->ltorque : c->torque, *ga = pass ? c->lgrad : c->grad;
    for (int i = 0; i < c->n; i++) {
      const double *o = s->xyz + 3 * i;
      const int *ax = s->axis + 3 * i;
      const double *pz = ax[0] >= 0 ? s->xyz + 3 * ax[0] : o;
      const double *px = ax[1] >= 0 ? s->xyz + 3 * ax[1] : o;
      const double *py = ax[2] >= 0 ? s->xyz + 3 * ax[2] : o;
      double g4[4][3];
      const double *tq = tqa + 3 * i;
      if (!isfinite(tq[0]) || !isfinite(tq[1]) || !isfinite(tq[2])) status = status ? status : 4;
      if (torque_to_grad(s->frame[i], o, pz, px, py, tq, g4) == 0 && s->frame[i] != ORA_FRAME_NONE) {
        int idx[4] = {ax[0], ax[1], ax[2], i};
        for (int j = 0; j < 4; j++)
          if (idx[j] >= 0)
            for (int d = 0; d < 3; d++) ga[3 * idx[j] + d] += g4[j][d];
        /* frame rotation under strain: the frame-atom gradients against the
         * frame vectors (they sum to zero with the centre's) */
        if (!pass)
          for (int j = 0; j < 3; j++)
            if (idx[j] >= 0)
              for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) out->vir[3 * a + b] += g4[j][a] * (s->xyz[3 * idx[j] + b] - o[b]);
      }
    }
  }
  return status;
}

/* One electrostatics evaluation (ParticleMeshEwald.computeEnergy) of the
 * system in c under the phase scales c->ph; energies and the virial are
 * added into out, gradients into c->grad / c->lgrad. */
static int elec_phase(ctx_t *c, ora_energies *out, double *ind, double *indcr) {
  const ora_system *s = c->s;
  const ora_params *prm = c->prm;
  int n = c->n;
  size_t s3 = 3 * (size_t)n * sizeof(double);
  c->gmp = malloc(10 * s3 / 3);
  ora_rotate_multipoles(s, c->gmp);
  c->ipd = malloc(n * sizeof(double));
  for (int i = 0; i < n; i++) c->ipd[i] = s->pdamp[i] > 0.0 ? 1.0 / s->pdamp[i] : INFINITY;
  c->field = malloc(s3); c->fieldp = malloc(s3);
  c->ind = calloc(3 * n, sizeof(double)); c->indp = calloc(3 * n, sizeof(double));
  c->dir = calloc(3 * n, sizeof(double)); c->dirp = calloc(3 * n, sizeof(double));
  c->use_recip = (!c->cell.aperiodic && c->aewald > 0.0);
  c->phi_p_f = calloc(20 * n, sizeof(double)); c->phi_p_c = calloc(10 * n, sizeof(double));
  c->phi_u_f = calloc(20 * n, sizeof(double)); c->phi_u_c = calloc(10 * n, sizeof(double));
  c->phi_pp_f = calloc(20 * n, sizeof(double)); c->phi_pp_c = calloc(10 * n, sizeof(double));
  c->fmp = calloc(10 * n, sizeof(double));
  if (c->use_recip) {
    pme_setup(&c->pme, &c->cell, prm, n, c->nth);
    ora_influence(s->box, prm->grid, prm->pme_order, c->aewald, c->pme.recip);
    pme_bsplines(&c->pme, &c->cell, n, s->xyz);
    for (int i = 0; i < n; i++) frac_multipole(&c->pme, c->gmp + 10 * i, c->fmp + 10 * i);
    pme_spread(&c->pme, n, c->fmp, NULL, NULL);
    ora_convolution(c->pme.grid, c->pme.recip, c->pme.K[0], c->pme.K[1], c->pme.K[2]);
    for (int i = 0; i < n; i++) {
      pme_gather(&c->pme, i, 0, c->phi_p_f + 20 * i);
      phi_to_cart(&c->pme, c->phi_p_f + 20 * i, c->phi_p_c + 10 * i);
    }
  }
  permanent_field(c);
  int polar = prm->polarization != ORA_POL_NONE && c->ph->polar;
  int status = 0;
  if (polar) {
    double a3 = 4.0 / 3.0 * pow(c->aewald, 3.0) / SQRT_PI;
    for (int i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++) {
        if (c->use_recip) {
          double fx = a3 * c->gmp[10 * i + 1 + d] - c->phi_p_c[10 * i + 1 + d];
          c->field[3 * i + d] += fx;
          c->fieldp[3 * i + d] += fx;
        }
        c->ind[3 * i + d] = c->dir[3 * i + d] = s->polarity[i] * c->field[3 * i + d];
        c->indp[3 * i + d] = c->dirp[3 * i + d] = s->polarity[i] * c->fieldp[3 * i + d];
      }
    }
    if (prm->direct_out) {
      memcpy(prm->direct_out, c->dir, s3);
      memcpy(prm->direct_out + 3 * n, c->dirp, s3);
    }
    if (prm->polarization == ORA_POL_MUTUAL && prm->scf_guess) {
      for (int q = 0; q < 3 * n; q++) {
        c->ind[q] += prm->scf_guess[q];
        c->indp[q] += prm->scf_guess[3 * n + q];
      }
    }
    if (prm->polarization == ORA_POL_MUTUAL) {
      if (prm->scf_algorithm == 1 || c->ph->sor) {
        status = sor_solve(c, out, prm->polar_sor);
      } else {
        pcg_tape_t tape;
        memset(&tape, 0, sizeof tape);
        if (prm->finite_pcg) c->tape = &tape;
        status = pcg_solve(c, out);
        /* ParticleMeshEwald.java:2300-2312: a failed PCG is repeated by SOR
         * from the direct dipoles with the conservative factor 0.65 */
        if (status && prm->sor_fallback) status = sor_solve(c, out, 0.65);
        else if (!status && prm->finite_pcg && tape.K > 0) {
          /* gradient of the truncated solve: uCR is replaced by the adjoint
           * of the direct field, the mutual terms by the tensor pairs */
          double *g = malloc(s3), *fhat = malloc(s3), *f = malloc(s3), *fp = malloc(s3);
          for (int q = 0; q < 3 * n; q++) g[q] = s->polarity[q / 3] > 0.0 ? c->fieldp[q] : 0.0;
          int np = 0;
          bipair_t *pairs = pcg_adjoint(c, g, fhat, &np);
          memcpy(c->indp, fhat, s3);
          induced_field(c, c->ind, c->indp, f, fp);
          bilinear_grad(c, pairs, np, -0.5 * prm->electric * c->ph->pol, c->grad);
          c->finite_grad = 1;
          for (int j = 0; j < np; j++) { free(pairs[j].a); free(pairs[j].b); }
          free(pairs); free(g); free(fhat); free(f); free(fp);
        }
        c->tape = NULL;
        tape_free(&tape);
      }
    } else {
      double *f = malloc(s3), *fp = malloc(s3);
      induced_field(c, c->ind, c->indp, f, fp);
      free(f); free(fp);
    }
  }
  if (c->use_recip) {
    recip_energy(c, out, polar);
    recip_grid_virial(c, polar, prm->polarization == ORA_POL_MUTUAL, out->vir);
  }
  real_energy(c, out, polar);
  out->n_perm += c->use_recip ? n : 0;
  int st = reduce_torques(c, out);
  if (!status) status = st;
  if (ind) memcpy(ind, c->ind, s3);
  if (indcr) memcpy(indcr, c->indp, s3);
  free(c->gmp); free(c->ipd); free(c->field); free(c->fieldp); free(c->ind); free(c->indp);
  free(c->dir); free(c->dirp); free(c->phi_p_f); free(c->phi_p_c); free(c->phi_u_f); free(c->phi_u_c);
  free(c->phi_pp_f); free(c->phi_pp_c); free(c->fmp);
  free(c->ew_off); free(c->ew); free(c->pc_off); free(c->pc);
  if (c->use_recip) pme_free(&c->pme);
  return status;
}

/* AlchemicalParameters.update (AlchemicalParameters.java:316-385), OST mode */
void ora_alchemical_factors(const ora_params *prm, double lambda, ora_alchemy *f) {
  memset(f, 0, sizeof *f);
  f->lpow_perm = 1.0;
  if (lambda < prm->perm_lambda_start) {
    f->lpow_perm = 0.0;
  } else if (lambda <= prm->perm_lambda_end) {
    double sc = 1.0 / (prm->perm_lambda_end - prm->perm_lambda_start);
    double L = sc * (lambda - prm->perm_lambda_start), a = prm->perm_lambda_alpha, b = prm->perm_lambda_exponent;
    f->lalpha = a * (1.0 - L) * (1.0 - L);
    f->dlalpha = a * (1.0 - L) * sc;
    f->d2lalpha = -a * sc * sc;
    f->lpow_perm = pow(L, b);
    f->dlpow_perm = b * pow(L, b - 1.0) * sc;
    f->d2lpow_perm = b >= 2.0 ? b * (b - 1.0) * pow(L, b - 2.0) * sc * sc : 0.0;
  }
  f->lpow_pol = 1.0;
  if (lambda < prm->pol_lambda_start) {
    f->lpow_pol = 0.0;
  } else if (lambda <= prm->pol_lambda_end) {
    double sc = 1.0 / (prm->pol_lambda_end - prm->pol_lambda_start);
    double L = sc * (lambda - prm->pol_lambda_start), b = prm->pol_lambda_exponent;
    if (b > 0.0) {
      f->lpow_pol = pow(L, b);
      if (b >= 1.0) {
        f->dlpow_pol = b * pow(L, b - 1.0) * sc;
        if (b >= 2.0) f->d2lpow_pol = b * (b - 1.0) * pow(L, b - 2.0) * sc * sc;
      }
    }
  }
}

/* The three OST electrostatics phases (ParticleMeshEwald.java:894-935):
 * (1) condensed phase, all atoms (condensedEnergy, :1609-1643);
 * (2) condensed phase without the soft atoms, reciprocal and polarization
 *     only, scaled by 1 - lambda terms (condensedNoLigandSCF, :1650-1697); an
 *     atom with use[i] = false is the same as one with no multipole and no
 *     polarizability, which is how it is built here;
 * (3) the soft atoms in vacuum, no Ewald screening or cutoff (ligandElec,
 *     :1704-1790), built as the soft atoms plus their frame atoms (the
 *     latter with no multipole or polarizability) in an aperiodic cell,
 *     whose dipoles are solved by SOR as in the reference. */
static int ost_electrostatics(ctx_t *c1, ora_energies *out, double *ind, double *indcr) {
  const ora_system *s = c1->s;
  const ora_params *prm = c1->prm;
  int n = s->n;
  double lam = prm->lambda;
  ora_alchemy f;
  ora_alchemical_factors(prm, lam, &f);
  int status = 0;
  /* phase 1 */
  phase_t p1 = PHASE_PLAIN;
  p1.soft_pairs = 1;
  p1.l2 = f.lpow_perm; p1.lalpha = f.lalpha; p1.dlalpha = f.dlalpha; p1.d2lalpha = f.d2lalpha;
  p1.perm = f.lpow_perm; p1.dperm = f.dlpow_perm; p1.d2perm = f.d2lpow_perm;
  if (lam < prm->pol_lambda_start) { p1.polar = 0; p1.pol = 0.0; }
  else { p1.pol = f.lpow_pol; p1.dpol = f.dlpow_pol; p1.d2pol = f.d2lpow_pol; }
  c1->ph = &p1;
  status = elec_phase(c1, out, ind, indcr);

  /* phase 2: no soft atoms */
  int nsoft = 0;
  for (int i = 0; i < n; i++) nsoft += s->soft[i] != 0;
  if (nsoft < n) {
    double *mp = malloc(10 * (size_t)n * sizeof(double)), *pol = malloc((size_t)n * sizeof(double));
    memcpy(mp, s->mpole, 10 * (size_t)n * sizeof(double));
    memcpy(pol, s->polarity, (size_t)n * sizeof(double));
    for (int i = 0; i < n; i++)
      if (s->soft[i]) { memset(mp + 10 * i, 0, 10 * sizeof(double)); pol[i] = 0.0; }
    ora_system s2 = *s;
    s2.mpole = mp; s2.polarity = pol;
    phase_t p2 = PHASE_PLAIN;
    p2.perm_real = 0;
    p2.perm = 1.0 - f.lpow_perm; p2.dperm = -f.dlpow_perm; p2.d2perm = -f.d2lpow_perm;
    if (lam <= prm->pol_lambda_end && prm->no_ligand_condensed_scf) {
      p2.pol = 1.0 - f.lpow_pol; p2.dpol = -f.dlpow_pol; p2.d2pol = -f.d2lpow_pol;
    } else {
      p2.polar = 0; p2.pol = 0.0;
    }
    ctx_t c2;
    ctx_open(&c2, &s2, prm, &p2);
    ora_energies e2;
    memset(&e2, 0, sizeof e2);
    int st = elec_phase(&c2, &e2, NULL, NULL);
    if (!status) status = st;
    for (int q = 0; q < 3 * n; q++) { c1->grad[q] += c2.grad[q]; c1->lgrad[q] += c2.lgrad[q]; }
    out->perm_real += e2.perm_real; out->perm_self += e2.perm_self; out->perm_recip += e2.perm_recip;
    out->perm_charge_corr += e2.perm_charge_corr;
    out->pol_real += e2.pol_real; out->pol_self += e2.pol_self; out->pol_recip += e2.pol_recip;
    out->dudl += e2.dudl; out->d2udl2 += e2.d2udl2;
    for (int q = 0; q < 9; q++) out->vir[q] += e2.vir[q];
    ctx_close(&c2);
    free(mp); free(pol);
  }

  /* phase 3: soft atoms in vacuum */
  if (prm->ligand_vapor_elec && nsoft > 0) {
    int *map = malloc((size_t)n * sizeof(int)), *inv = malloc((size_t)n * sizeof(int));
    for (int i = 0; i < n; i++) inv[i] = -1;
    int m = 0;
    for (int i = 0; i < n; i++)
      if (s->soft[i]) { inv[i] = m; map[m++] = i; }
    for (int i = 0; i < n; i++)
      if (s->soft[i])
        for (int j = 0; j < 3; j++) {
          int a = s->axis[3 * i + j];
          if (a >= 0 && inv[a] < 0) { inv[a] = m; map[m++] = a; }
        }
    double *xyz = malloc(3 * (size_t)m * sizeof(double)), *mp = calloc(10 * (size_t)m, sizeof(double));
    double *pol = calloc((size_t)m, sizeof(double)), *th = malloc((size_t)m * sizeof(double));
    double *pd = malloc((size_t)m * sizeof(double)), *rv = calloc((size_t)m, sizeof(double));
    int *fr = calloc((size_t)m, sizeof(int)), *ax = malloc(3 * (size_t)m * sizeof(int));
    int *grp = malloc((size_t)m * sizeof(int)), *mol = malloc((size_t)m * sizeof(int));
    int *sf = malloc((size_t)m * sizeof(int)), *ri = malloc((size_t)m * sizeof(int));
    int *vc = calloc((size_t)m, sizeof(int));
    int *moff = calloc((size_t)m + 1, sizeof(int));
    int nnz = s->mask_off[n];
    int *mat = malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int)), *mrel = malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    int t = 0;
    for (int j = 0; j < m; j++) {
      int i = map[j];
      memcpy(xyz + 3 * j, s->xyz + 3 * i, 3 * sizeof(double));
      th[j] = s->thole[i]; pd[j] = s->pdamp[i]; grp[j] = s->group[i];
      mol[j] = s->molecule ? s->molecule[i] : i;
      sf[j] = s->soft[i] != 0; ri[j] = j;
      if (s->soft[i]) {
        memcpy(mp + 10 * j, s->mpole + 10 * i, 10 * sizeof(double));
        pol[j] = s->polarity[i];
        fr[j] = s->frame[i];
        for (int d = 0; d < 3; d++) ax[3 * j + d] = s->axis[3 * i + d] >= 0 ? inv[s->axis[3 * i + d]] : -1;
      } else {
        for (int d = 0; d < 3; d++) ax[3 * j + d] = -1;
      }
      for (int q = s->mask_off[i]; q < s->mask_off[i + 1]; q++)
        if (inv[s->mask_atom[q]] >= 0) { mat[t] = inv[s->mask_atom[q]]; mrel[t] = s->mask_rel[q]; t++; }
      moff[j + 1] = t;
    }
    double vbox[6] = {0, 0, 0, 90, 90, 90};
    ora_system s3 = *s;
    s3.n = m; s3.xyz = xyz; s3.frame = fr; s3.axis = ax; s3.mpole = mp; s3.polarity = pol; s3.thole = th;
    s3.pdamp = pd; s3.group = grp; s3.molecule = mol; s3.vdw_class = vc; s3.red_index = ri; s3.red_value = rv;
    s3.mask_off = moff; s3.mask_atom = mat; s3.mask_rel = mrel; s3.box = vbox; s3.soft = sf;
    ora_params q3 = *prm;
    q3.ewald_cutoff = 1.0e4;   /* setEwaldParameters(Double.MAX_VALUE, 0.0) */
    q3.vdw_cutoff = 0.0;
    q3.list_buffer = 0.0;
    q3.do_vdw = 0;
    phase_t p3 = PHASE_PLAIN;
    p3.soft_pairs = 1; p3.vapor = 1; p3.sor = 1;
    p3.l2 = 1.0 - f.lpow_perm;
    p3.perm = 1.0 - f.lpow_perm; p3.dperm = -f.dlpow_perm; p3.d2perm = -f.d2lpow_perm;
    if (lam <= prm->pol_lambda_end) {
      p3.pol = 1.0 - f.lpow_pol; p3.dpol = -f.dlpow_pol; p3.d2pol = -f.d2lpow_pol;
    } else {
      p3.polar = 0; p3.pol = 0.0;
    }
    ctx_t c3;
    ctx_open(&c3, &s3, &q3, &p3);
    ora_energies e3;
    memset(&e3, 0, sizeof e3);
    int st = elec_phase(&c3, &e3, NULL, NULL);
    if (!status) status = st;
    for (int j = 0; j < m; j++)
      for (int d = 0; d < 3; d++) {
        c1->grad[3 * map[j] + d] += c3.grad[3 * j + d];
        c1->lgrad[3 * map[j] + d] += c3.lgrad[3 * j + d];
      }
    out->perm_real += e3.perm_real;
    out->pol_real += e3.pol_real;
    out->dudl += e3.dudl; out->d2udl2 += e3.d2udl2;
    for (int q = 0; q < 9; q++) out->vir[q] += e3.vir[q];
    ctx_close(&c3);
    free(map); free(inv); free(xyz); free(mp); free(pol); free(th); free(pd); free(rv); free(fr); free(ax);
    free(grp); free(mol); free(sf); free(ri); free(vc); free(moff); free(mat); free(mrel);
  }
  return status;
}

int ora_evaluate_lambda(const ora_system *s, const ora_params *prm, ora_energies *out, double *grad, double *ind,
                        double *indcr, double *dudx_dl) {
  memset(out, 0, sizeof(*out));
  int n = s->n;
  size_t s3 = 3 * (size_t)n * sizeof(double);
  ctx_t c;
  ctx_open(&c, s, prm, &PHASE_PLAIN);
  if (prm->do_vdw) out->vdw = vdw_energy(&c, c.grad, &out->n_vdw, out->vir, &out->dudl, &out->d2udl2, c.lgrad);
  int status = 0;
  if (prm->do_pme) {
    int nsoft = 0;
    if (s->soft)
      for (int i = 0; i < n; i++) nsoft += s->soft[i] != 0;
    if (prm->elec_lambda && nsoft > 0) status = ost_electrostatics(&c, out, ind, indcr);
    else status = elec_phase(&c, out, ind, indcr);
  }
  out->status = status;
  out->perm = out->perm_real + out->perm_self + out->perm_recip + out->perm_charge_corr;
  out->pol = out->pol_real + out->pol_self + out->pol_recip;
  out->total = out->perm + out->pol + out->vdw;
  if (!isfinite(out->total) && out->status == 0) out->status = 1;
  if (grad) memcpy(grad, c.grad, s3);
  if (dudx_dl) memcpy(dudx_dl, c.lgrad, s3);
  ctx_close(&c);
  return out->status;
}
