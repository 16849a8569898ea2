/*
 * valence_oracle.c -- FP64 CPU restatement of the AMOEBA valence terms of
 * Force Field X.  TEST INFRASTRUCTURE ONLY (see valence_oracle.h).
 *
 * Each term follows the reference's energy() method:
 *   bond              Bond.java:242-283
 *   angle / in-plane  Angle.java:240-345
 *   stretch-bend      StretchBend.java:117-160
 *   Urey-Bradley      UreyBradley.java:123-153
 *   out-of-plane bend OutOfPlaneBend.java:172-250 (Allinger form)
 *   torsion           Torsion.java:247-330
 *   pi-orbital        PiOrbitalTorsion.java:145-240
 *   torsion-torsion   TorsionTorsion.java:285-430, chktor :489-555,
 *                     bicubic patch of bcuint/bcuint1 :565-620
 */
#include "valence_oracle.h"

#include <math.h>
#include <string.h>

#define DEG (180.0 / M_PI)

typedef struct { double x, y, z; } v3;

static v3 at(const double *xyz, int i) { v3 r = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}; return r; }
static v3 sub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 add(v3 a, v3 b) { v3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static v3 scl(v3 a, double s) { v3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 crs(v3 a, v3 b) { v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; return r; }
static double len(v3 a) { return sqrt(dot(a, a)); }
static double clamp1(double c) { return c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c); }

typedef struct { double *grad, *vir; const double *xyz; } acc_t;

static void put(acc_t *a, int i, v3 g) {
  if (a->grad) {
    a->grad[3 * i] += g.x;
    a->grad[3 * i + 1] += g.y;
    a->grad[3 * i + 2] += g.z;
  }
  if (a->vir) {
    const double gg[3] = {g.x, g.y, g.z}, *x = a->xyz + 3 * i;
    for (int p = 0; p < 3; p++)
      for (int q = 0; q < 3; q++) a->vir[3 * p + q] += gg[p] * x[q];
  }
}

static double poly4(const double *c, double dv, double *dpoly) {
  double dv2 = dv * dv, dv3 = dv2 * dv, dv4 = dv2 * dv2;
  *dpoly = dv * (2.0 + 3.0 * c[0] * dv + 4.0 * c[1] * dv2 + 5.0 * c[2] * dv3 + 6.0 * c[3] * dv4);
  return dv2 * (1.0 + c[0] * dv + c[1] * dv2 + c[2] * dv3 + c[3] * dv4);
}

/* gradient of the angle at b between legs u = a - b, w = c - b, scaled by
 * dE/dtheta (radians): returns ga and gc; gb = -(ga + gc) */
static void angle_grad(v3 u, v3 w, double dedt, v3 *ga, v3 *gc) {
  v3 p = crs(w, u);
  double rp = len(p);
  if (rp < 1e-6) rp = 1e-6;
  *ga = scl(crs(u, p), -dedt / (dot(u, u) * rp));
  *gc = scl(crs(w, p), dedt / (dot(w, w) * rp));
}

/* Bond.java:242-283 */
static double e_bond(const ora_valence *v, int t, acc_t *A) {
  int ia = v->bond[2 * t], ib = v->bond[2 * t + 1];
  v3 d = sub(at(A->xyz, ia), at(A->xyz, ib));
  double r = len(d), dv = r - v->bond_r0[t], k = v->bond_k[t];
  double e = k * dv * dv * (1.0 + v->bond_cubic * dv + v->bond_quartic * dv * dv);
  double dedr = 2.0 * k * dv * (1.0 + 1.5 * v->bond_cubic * dv + 2.0 * v->bond_quartic * dv * dv);
  v3 g = scl(d, r > 0.0 ? dedr / r : 0.0);
  put(A, ia, g);
  put(A, ib, scl(g, -1.0));
  return e;
}

/* Angle.java:240-345: normal and projected in-plane angle */
static double e_angle(const ora_valence *v, int t, acc_t *A) {
  const int *q = v->angle + 4 * t;
  v3 xa = at(A->xyz, q[0]), xb = at(A->xyz, q[1]), xc = at(A->xyz, q[2]);
  double k = v->angle_k[t], t0 = v->angle_t0[t], dp;
  if (q[3] < 0) {
    v3 u = sub(xa, xb), w = sub(xc, xb);
    double uu = dot(u, u), ww = dot(w, w);
    if (uu == 0.0 || ww == 0.0) return 0.0;
    double th = acos(clamp1(dot(u, w) / sqrt(uu * ww))) * DEG;
    double e = k * poly4(v->angle_poly, th - t0, &dp);
    v3 ga, gc;
    angle_grad(u, w, k * dp * DEG, &ga, &gc);
    put(A, q[0], ga);
    put(A, q[1], scl(add(ga, gc), -1.0));
    put(A, q[2], gc);
    return e;
  }
  /* in-plane: the centre projected onto the plane of a, c and the fourth atom */
  v3 xd = at(A->xyz, q[3]);
  v3 ad = sub(xa, xd), bd = sub(xb, xd), cd = sub(xc, xd);
  v3 n = crs(ad, cd);
  double nn = dot(n, n);
  double delta = -dot(n, bd) / nn;
  v3 ip = add(scl(n, delta), bd);              /* projected centre, relative to d */
  v3 jp = sub(ad, ip), kp = sub(cd, ip);
  double jj = dot(jp, jp), kk = dot(kp, kp);
  if (jj == 0.0 || kk == 0.0) return 0.0;
  double th = acos(clamp1(dot(jp, kp) / sqrt(jj * kk))) * DEG;
  double e = k * poly4(v->angle_poly, th - t0, &dp);
  v3 g0, g2;
  angle_grad(jp, kp, k * dp * DEG, &g0, &g2);
  v3 gp = scl(add(g0, g2), -1.0);              /* gradient on the projected centre */
  /* the projection p = b + delta n moves with the plane normal n = ad x cd;
   * gp lies in the plane (gp . n = 0), so only delta dn contributes:
   * a gets delta (cd x gp), c gets delta (gp x ad), b gets gp itself */
  v3 ga = add(g0, scl(crs(cd, gp), delta));
  v3 gc = add(g2, scl(crs(gp, ad), delta));
  v3 gb = gp;
  put(A, q[0], ga);
  put(A, q[1], gb);
  put(A, q[2], gc);
  put(A, q[3], scl(add(ga, add(gb, gc)), -1.0));
  return e;
}

/* StretchBend.java:117-160 */
static double e_strbnd(const ora_valence *v, int t, acc_t *A) {
  const int *q = v->strbnd + 3 * t;
  v3 u = sub(at(A->xyz, q[0]), at(A->xyz, q[1])), w = sub(at(A->xyz, q[2]), at(A->xyz, q[1]));
  double uu = dot(u, u), ww = dot(w, w);
  if (uu == 0.0 || ww == 0.0) return 0.0;
  double ru = sqrt(uu), rw = sqrt(ww);
  double th = acos(clamp1(dot(u, w) / (ru * rw))) * DEG;
  double k0 = v->strbnd_k[2 * t], k1 = v->strbnd_k[2 * t + 1];
  double dr = k0 * (ru - v->strbnd_r0[2 * t]) + k1 * (rw - v->strbnd_r0[2 * t + 1]);
  double dt = th - v->strbnd_t0[t];
  v3 ga, gc;
  angle_grad(u, w, dr * DEG, &ga, &gc);
  ga = add(ga, scl(u, k0 * dt / ru));
  gc = add(gc, scl(w, k1 * dt / rw));
  put(A, q[0], ga);
  put(A, q[1], scl(add(ga, gc), -1.0));
  put(A, q[2], gc);
  return dr * dt;
}

/* UreyBradley.java:123-153 */
static double e_urey(const ora_valence *v, int t, acc_t *A) {
  int ia = v->urey[2 * t], ic = v->urey[2 * t + 1];
  v3 d = sub(at(A->xyz, ia), at(A->xyz, ic));
  double r = len(d), dv = r - v->urey_r0[t], k = v->urey_k[t];
  double e = k * dv * dv * (1.0 + v->urey_cubic * dv + v->urey_quartic * dv * dv);
  double de = 2.0 * k * dv * (1.0 + 1.5 * v->urey_cubic * dv + 2.0 * v->urey_quartic * dv * dv);
  v3 g = scl(d, r > 0.0 ? de / r : 0.0);
  put(A, ia, g);
  put(A, ic, scl(g, -1.0));
  return e;
}

/* OutOfPlaneBend.java:172-250: angle of the bond centre -> d with the plane
 * through a, c and d */
static double e_opbend(const ora_valence *v, int t, acc_t *A) {
  const int *q = v->opbend + 4 * t;
  v3 xa = at(A->xyz, q[0]), xb = at(A->xyz, q[1]), xc = at(A->xyz, q[2]), xd = at(A->xyz, q[3]);
  v3 ab = sub(xa, xb), cb = sub(xc, xb), db = sub(xd, xb), ad = sub(xa, xd), cd = sub(xc, xd);
  double db2 = dot(db, db), ad2 = dot(ad, ad), cd2 = dot(cd, cd);
  double ee = dot(ab, crs(cb, db));
  double ac = dot(ad, cd);
  double cc = ad2 * cd2 - ac * ac;
  if (db2 == 0.0 || cc == 0.0) return 0.0;
  double bkk2 = db2 - ee * ee / cc;
  double th = acos(clamp1(sqrt(bkk2 / db2))) * DEG;
  double dp, k = v->opbend_k[t];
  double e = k * poly4(v->opbend_poly, th, &dp);
  double dedcos = -k * dp * DEG * (ee < 0.0 ? -1.0 : 1.0) / sqrt(cc * bkk2);
  double term = ee / cc;
  v3 dcda = scl(sub(scl(ad, cd2), scl(cd, ac)), term);
  v3 dcdc = scl(sub(scl(cd, ad2), scl(ad, ac)), term);
  v3 dcdd = scl(add(dcda, dcdc), -1.0);
  v3 deda = crs(db, cb), dedc = crs(ab, db), dedd = add(crs(cb, ab), scl(db, ee / db2));
  v3 ga = scl(add(dcda, deda), dedcos), gc = scl(add(dcdc, dedc), dedcos), gd = scl(add(dcdd, dedd), dedcos);
  put(A, q[0], ga);
  put(A, q[1], scl(add(ga, add(gc, gd)), -1.0));
  put(A, q[2], gc);
  put(A, q[3], gd);
  return e;
}

/* dihedral a-b-c-d: cosine, sine and the gradient for dE/dphi (radians) */
typedef struct { v3 ba, cb, dc, t, u; double rt2, ru2, rcb, cosine, sine; } dihed_t;

static void dihedral(v3 xa, v3 xb, v3 xc, v3 xd, double tol, dihed_t *h) {
  h->ba = sub(xb, xa);
  h->cb = sub(xc, xb);
  h->dc = sub(xd, xc);
  h->t = crs(h->ba, h->cb);
  h->u = crs(h->cb, h->dc);
  h->rt2 = fmax(dot(h->t, h->t), tol);
  h->ru2 = fmax(dot(h->u, h->u), tol);
  double rr = sqrt(h->rt2 * h->ru2);
  h->rcb = fmax(len(h->cb), tol);
  h->cosine = dot(h->t, h->u) / rr;
  h->sine = dot(h->cb, crs(h->t, h->u)) / (h->rcb * rr);
}

static void dihedral_grad(const dihed_t *h, v3 xa, v3 xb, v3 xc, v3 xd, double dedphi, v3 g[4]) {
  v3 ca = sub(xc, xa), db = sub(xd, xb);
  v3 dt = scl(crs(h->t, h->cb), dedphi / (h->rt2 * h->rcb));
  v3 du = scl(crs(h->u, h->cb), -dedphi / (h->ru2 * h->rcb));
  g[0] = crs(dt, h->cb);
  g[1] = add(crs(ca, dt), crs(du, h->dc));
  g[2] = add(crs(dt, h->ba), crs(db, du));
  g[3] = crs(du, h->cb);
}

/* Torsion.java:247-330 */
static double e_torsion(const ora_valence *v, int t, acc_t *A) {
  const int *q = v->torsion + 4 * t;
  v3 x[4];
  for (int k = 0; k < 4; k++) x[k] = at(A->xyz, q[k]);
  dihed_t h;
  dihedral(x[0], x[1], x[2], x[3], 1e-4, &h);
  const double *amp = v->torsion_amp + 6 * t, *tc = v->torsion_cos + 6 * t, *ts = v->torsion_sin + 6 * t;
  double e = 0.0, dedphi = 0.0, cn = h.cosine, sn = h.sine;
  for (int k = 0; k < 6; k++) {
    if (k > 0) {
      double c2 = h.cosine * cn - h.sine * sn, s2 = h.sine * cn + h.cosine * sn;
      cn = c2;
      sn = s2;
    }
    e += amp[k] * (1.0 + cn * tc[k] + sn * ts[k]);
    dedphi += amp[k] * (k + 1) * (cn * ts[k] - sn * tc[k]);
  }
  v3 g[4];
  dihedral_grad(&h, x[0], x[1], x[2], x[3], dedphi, g);
  for (int k = 0; k < 4; k++) put(A, q[k], g[k]);
  return e;
}

/* PiOrbitalTorsion.java:145-240: torsion p-c-d-q between the pi-orbital
 * directions p = c + (a - d) x (b - d) and q = d + (e - c) x (f - c) */
static double e_pitors(const ora_valence *v, int t, acc_t *A) {
  const int *q = v->pitors + 6 * t;
  v3 xa = at(A->xyz, q[0]), xb = at(A->xyz, q[1]), xc = at(A->xyz, q[2]);
  v3 xd = at(A->xyz, q[3]), xe = at(A->xyz, q[4]), xf = at(A->xyz, q[5]);
  v3 ad = sub(xa, xd), bd = sub(xb, xd), ec = sub(xe, xc), fc = sub(xf, xc);
  v3 xp = add(crs(ad, bd), xc), xq = add(crs(ec, fc), xd);
  v3 pc = sub(xc, xp), dc = sub(xd, xc), dq = sub(xq, xd);
  v3 tt = crs(pc, dc), uu = crs(dc, dq);
  double rt2 = dot(tt, tt), ru2 = dot(uu, uu);
  if (rt2 * ru2 == 0.0) return 0.0;
  double rr = sqrt(rt2 * ru2), rdc = len(dc);
  double sine = dot(dc, crs(tt, uu)) / (rdc * rr);
  double cosine = clamp1(dot(tt, uu) / rr);
  double k = v->pitors_k[t];
  double e = k * (1.0 - (cosine * cosine - sine * sine));
  double dedphi = k * 4.0 * cosine * sine;
  v3 dp = sub(xd, xp), qc = sub(xq, xc);
  v3 vdt = scl(crs(tt, dc), dedphi / (rt2 * rdc));
  v3 vdu = scl(crs(uu, dc), -dedphi / (ru2 * rdc));
  v3 gp = crs(vdt, dc);
  v3 gcc = add(crs(dp, vdt), crs(vdu, dq));
  v3 gdd = add(crs(vdt, pc), crs(qc, vdu));
  v3 gq = crs(vdu, dc);
  v3 ga = crs(bd, gp), gb = crs(gp, ad), ge = crs(fc, gq), gf = crs(gq, ec);
  put(A, q[0], ga);
  put(A, q[1], gb);
  put(A, q[2], sub(sub(add(gcc, gp), ge), gf));
  put(A, q[3], sub(sub(add(gdd, gq), ga), gb));
  put(A, q[4], ge);
  put(A, q[5], gf);
  return e;
}

/* bicubic Hermite patch on [xl, xu] x [yl, yu] from corner values,
 * first and cross derivatives (corners counter-clockwise from (xl, yl));
 * value and the two first derivatives at (t1, t2) */
static double bicubic(double xl, double xu, double yl, double yu, double t1, double t2, const double *e,
                      const double *dx, const double *dy, const double *dxy, double *d1, double *d2) {
  double hx = xu - xl, hy = yu - yl;
  double u = (t1 - xl) / hx, w = (t2 - yl) / hy;
  /* cubic Hermite basis on [0, 1] and its derivative: value at 0, slope at 0,
   * value at 1, slope at 1 */
  double bu[4] = {(2 * u - 3) * u * u + 1, ((u - 2) * u + 1) * u, (3 - 2 * u) * u * u, (u - 1) * u * u};
  double du[4] = {6 * u * (u - 1), (3 * u - 4) * u + 1, 6 * u * (1 - u), (3 * u - 2) * u};
  double bw[4] = {(2 * w - 3) * w * w + 1, ((w - 2) * w + 1) * w, (3 - 2 * w) * w * w, (w - 1) * w * w};
  double dw[4] = {6 * w * (w - 1), (3 * w - 4) * w + 1, 6 * w * (1 - w), (3 * w - 2) * w};
  /* corner k sits at (cx[k], cy[k]) of the unit square */
  const int cx[4] = {0, 1, 1, 0}, cy[4] = {0, 0, 1, 1};
  double f = 0.0, fu = 0.0, fw = 0.0;
  for (int k = 0; k < 4; k++) {
    int iu = 2 * cx[k], iw = 2 * cy[k];
    double vals[4] = {e[k], dx[k] * hx, dy[k] * hy, dxy[k] * hx * hy};
    /* (value, u-slope, w-slope, cross) x basis */
    f += vals[0] * bu[iu] * bw[iw] + vals[1] * bu[iu + 1] * bw[iw] + vals[2] * bu[iu] * bw[iw + 1]
         + vals[3] * bu[iu + 1] * bw[iw + 1];
    fu += vals[0] * du[iu] * bw[iw] + vals[1] * du[iu + 1] * bw[iw] + vals[2] * du[iu] * bw[iw + 1]
          + vals[3] * du[iu + 1] * bw[iw + 1];
    fw += vals[0] * bu[iu] * dw[iw] + vals[1] * bu[iu + 1] * dw[iw] + vals[2] * bu[iu] * dw[iw + 1]
          + vals[3] * bu[iu + 1] * dw[iw + 1];
  }
  *d1 = fu / hx;
  *d2 = fw / hy;
  return f;
}

/* TorsionTorsion.java:285-430 */
static double e_tortor(const ora_valence *v, int t, acc_t *A) {
  const int *q = v->tortor + 6 * t;
  v3 x[5];
  for (int k = 0; k < 5; k++) x[k] = at(A->xyz, q[k]);
  v3 ba = sub(x[1], x[0]), cb = sub(x[2], x[1]), dc = sub(x[3], x[2]), ed = sub(x[4], x[3]);
  v3 tt = crs(ba, cb), uu = crs(cb, dc), vv = crs(dc, ed);
  double rt2 = dot(tt, tt), ru2 = dot(uu, uu), rv2 = dot(vv, vv);
  if (rt2 * ru2 == 0.0 || ru2 * rv2 == 0.0) return 0.0;
  double a1 = acos(clamp1(dot(tt, uu) / sqrt(rt2 * ru2))) * DEG;
  if (dot(ba, uu) < 0.0) a1 = -a1;
  double a2 = acos(clamp1(dot(uu, vv) / sqrt(ru2 * rv2))) * DEG;
  if (dot(cb, vv) < 0.0) a2 = -a2;
  /* chirality of the centre (chktor): sign of the volume spanned from it */
  double sign = 1.0;
  if (q[5] >= 0) {
    v3 c = x[2], c0 = sub(at(A->xyz, q[5]), c), c1 = sub(x[1], c), c2 = sub(x[3], c);
    if (dot(c0, crs(c1, c2)) < 0.0) sign = -1.0;
  }
  double t1 = sign * a1, t2 = sign * a2;
  int gi = v->tortor_grid[t], nx = v->grid_nx[gi], ny = v->grid_ny[gi];
  const double *tx = v->grid_t + v->grid_off_t[gi], *ty = tx + nx;
  int lo = 0, hi = nx - 1;
  while (hi - lo > 1) {
    int m = (hi + lo) / 2;
    if (tx[m] > t1) hi = m; else lo = m;
  }
  int xl = lo;
  lo = 0;
  hi = ny - 1;
  while (hi - lo > 1) {
    int m = (hi + lo) / 2;
    if (ty[m] > t2) hi = m; else lo = m;
  }
  int yl = lo;
  const int o = v->grid_off_e[gi];
  int p1 = yl * nx + xl, p2 = (yl + 1) * nx + xl;
  int idx[4] = {p1, p1 + 1, p2 + 1, p2};
  double ce[4], cdx[4], cdy[4], cdxy[4];
  for (int k = 0; k < 4; k++) {
    ce[k] = v->grid_e[o + idx[k]];
    cdx[k] = v->grid_dx[o + idx[k]];
    cdy[k] = v->grid_dy[o + idx[k]];
    cdxy[k] = v->grid_dxy[o + idx[k]];
  }
  double d1, d2;
  double e = bicubic(tx[xl], tx[xl + 1], ty[yl], ty[yl + 1], t1, t2, ce, cdx, cdy, cdxy, &d1, &d2);
  double de1 = sign * d1 * DEG, de2 = sign * d2 * DEG;
  double rcb = len(cb), rdc = len(dc);
  /* first torsion a-b-c-d */
  v3 ca = sub(x[2], x[0]), db = sub(x[3], x[1]), ec = sub(x[4], x[2]);
  v3 dt = scl(crs(tt, cb), de1 / (rt2 * rcb)), du = scl(crs(uu, cb), -de1 / (ru2 * rcb));
  put(A, q[0], crs(dt, cb));
  put(A, q[1], add(crs(ca, dt), crs(du, dc)));
  put(A, q[2], add(crs(db, du), crs(dt, ba)));
  put(A, q[3], crs(du, cb));
  /* second torsion b-c-d-e */
  dt = scl(crs(uu, dc), de2 / (ru2 * rdc));
  du = scl(crs(vv, dc), -de2 / (rv2 * rdc));
  put(A, q[1], crs(dt, dc));
  put(A, q[2], add(crs(db, dt), crs(du, ed)));
  put(A, q[3], add(crs(dt, cb), crs(ec, du)));
  put(A, q[4], crs(du, dc));
  return e;
}

void ora_valence_evaluate(const ora_valence *v, const double *xyz, double *energy, double *grad, double *vir) {
  acc_t A = {grad, vir, xyz};
  memset(energy, 0, ORA_V_N * sizeof(double));
  for (int t = 0; t < v->n_bond; t++) energy[ORA_V_BOND] += e_bond(v, t, &A);
  for (int t = 0; t < v->n_angle; t++) energy[ORA_V_ANGLE] += e_angle(v, t, &A);
  for (int t = 0; t < v->n_strbnd; t++) energy[ORA_V_STRBND] += e_strbnd(v, t, &A);
  for (int t = 0; t < v->n_urey; t++) energy[ORA_V_UREY] += e_urey(v, t, &A);
  for (int t = 0; t < v->n_opbend; t++) energy[ORA_V_OPBEND] += e_opbend(v, t, &A);
  for (int t = 0; t < v->n_torsion; t++) energy[ORA_V_TORSION] += e_torsion(v, t, &A);
  for (int t = 0; t < v->n_pitors; t++) energy[ORA_V_PITORS] += e_pitors(v, t, &A);
  for (int t = 0; t < v->n_tortor; t++) energy[ORA_V_TORTOR] += e_tortor(v, t, &A);
}
