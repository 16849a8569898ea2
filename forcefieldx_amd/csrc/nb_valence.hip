// nb_valence.hip -- AMOEBA valence terms beside the nonbonded path: bond,
// angle (normal and projected in-plane), stretch-bend, Urey-Bradley,
// out-of-plane bend, torsion, pi-orbital torsion and torsion-torsion, in the
// forms of Bond.java:242-283, Angle.java:240-345, StretchBend.java:117-160,
// UreyBradley.java:123-153, OutOfPlaneBend.java:172-250,
// Torsion.java:247-330, PiOrbitalTorsion.java:145-240 and
// TorsionTorsion.java:285-430 (replacing the reference's bonded-term region,
// ForceFieldEnergy.java:1670-1677, and terms/EnergyTermRegion.java).
//
//  * k_valence: one thread per term over all kinds (kinds in contiguous
//    ranges, so a wave diverges only at a range boundary); writes the term's
//    energy and one gradient contribution per atom into its own slots.
//  * k_valence_gather: eight lanes per atom sum its slots (a CSR list in slot
//    order, built at nb_set_valence) into the gradient, and its strain
//    derivative g x^T into per-block partials; no atomics, so the result is
//    the same on every run and every graph replay.
//  * k_valence_energy: one workgroup per kind sums the term energies.
// All FP64: the terms are ~1e-3 of the evaluation's work.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "nb_kern.h"

void nb_drop_graph(NbHandle *h);

namespace {

constexpr int NKIND = 8;
constexpr int NATOMS[NKIND] = {2, 4, 3, 2, 4, 4, 6, 5};   // gradient slots per term of each kind (host)
constexpr int NIDX[NKIND] = {2, 4, 3, 2, 4, 4, 6, 6};     // integers per term (tortor: + chirality atom)
__device__ __forceinline__ int natoms_of(int k) {
  return k == 0 || k == 3 ? 2 : (k == 2 ? 3 : (k == 6 ? 6 : (k == 7 ? 5 : 4)));
}

struct VDev {
  int start[NKIND + 1];          // first global term index of each kind
  long sbase[NKIND];             // first slot of each kind
  const int32_t *idx[NKIND];
  const double *bond_k, *bond_r0;
  double bond_c3, bond_c4;
  const double *ang_k, *ang_t0;
  double ang_p[4];
  const double *sb_k, *sb_r0, *sb_t0;
  const double *ub_k, *ub_r0;
  double ub_c3, ub_c4;
  const double *op_k;
  double op_p[4];
  const double *tor_amp, *tor_cos, *tor_sin;
  const double *pi_k;
  const int32_t *tt_grid, *g_nx, *g_ny, *g_off_t, *g_off_e;
  const double *g_t, *g_e, *g_dx, *g_dy, *g_dxy;
};

struct D3 { double x, y, z; };
__device__ __forceinline__ D3 ld(const double *xyz, int i) { return D3{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 crs(D3 a, D3 b) {
  return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double clamp1(double c) { return fmin(1.0, fmax(-1.0, c)); }

constexpr double RAD2DEG = 57.29577951308232087680;

// k dv^2 (1 + c3 dv + c4 dv^2 + c5 dv^3 + c6 dv^4) and its dv-derivative / k
__device__ __forceinline__ double sextic(const double *c, double dv, double &d) {
  double dv2 = dv * dv;
  d = dv * (2.0 + dv * (3.0 * c[0] + dv * (4.0 * c[1] + dv * (5.0 * c[2] + dv * 6.0 * c[3]))));
  return dv2 * (1.0 + dv * (c[0] + dv * (c[1] + dv * (c[2] + dv * c[3]))));
}

// d(theta)/d(end a) and d(theta)/d(end c) of the angle between legs u = a - b
// and w = c - b, times dedt (per radian)
__device__ __forceinline__ void angle_legs(D3 u, D3 w, double dedt, D3 &ga, D3 &gc) {
  D3 p = crs(w, u);
  double rp = fmax(sqrt(dot(p, p)), 1e-6);
  ga = crs(u, p) * (-dedt / (dot(u, u) * rp));
  gc = crs(w, p) * (dedt / (dot(w, w) * rp));
}

// gradient of a dihedral a-b-c-d for dE/dphi = dedphi (per radian), given
// t = (b - a) x (c - b), u = (c - b) x (d - c) and their squared norms
__device__ __forceinline__ void dihedral_grad(D3 ba, D3 cb, D3 dc, D3 ca, D3 db, D3 t, D3 u, double rt2, double ru2,
                                              double rcb, double dedphi, D3 *g) {
  D3 dt = crs(t, cb) * (dedphi / (rt2 * rcb));
  D3 du = crs(u, cb) * (-dedphi / (ru2 * rcb));
  g[0] = crs(dt, cb);
  g[1] = crs(ca, dt) + crs(du, dc);
  g[2] = crs(dt, ba) + crs(db, du);
  g[3] = crs(du, cb);
}

__device__ double term_bond(const VDev &v, int t, const double *xyz, D3 *g) {
  const int32_t *q = v.idx[0] + 2 * t;
  D3 d = ld(xyz, q[0]) - ld(xyz, q[1]);
  double r = sqrt(dot(d, d)), dv = r - v.bond_r0[t], k = v.bond_k[t];
  double e = k * dv * dv * (1.0 + dv * (v.bond_c3 + dv * v.bond_c4));
  double dedr = 2.0 * k * dv * (1.0 + dv * (1.5 * v.bond_c3 + 2.0 * dv * v.bond_c4));
  g[0] = d * (r > 0.0 ? dedr / r : 0.0);
  g[1] = g[0] * -1.0;
  return e;
}

__device__ double term_angle(const VDev &v, int t, const double *xyz, D3 *g) {
  const int32_t *q = v.idx[1] + 4 * t;
  D3 xa = ld(xyz, q[0]), xb = ld(xyz, q[1]), xc = ld(xyz, q[2]);
  double k = v.ang_k[t], t0 = v.ang_t0[t], dp;
  g[0] = g[1] = g[2] = g[3] = D3{0.0, 0.0, 0.0};
  if (q[3] < 0) {
    D3 u = xa - xb, w = xc - xb;
    double uu = dot(u, u), ww = dot(w, w);
    if (uu == 0.0 || ww == 0.0) return 0.0;
    double th = acos(clamp1(dot(u, w) / sqrt(uu * ww))) * RAD2DEG;
    double e = k * sextic(v.ang_p, th - t0, dp);
    angle_legs(u, w, k * dp * RAD2DEG, g[0], g[2]);
    g[1] = (g[0] + g[2]) * -1.0;
    return e;
  }
  // in-plane: the angle at the centre's projection p = b + delta n onto the
  // plane through a, c and the fourth atom d (n = (a - d) x (c - d)); the
  // projection's gradient gp lies in that plane, so a and c also receive
  // delta (cd x gp) and delta (gp x ad) through the motion of n
  D3 xd = ld(xyz, q[3]);
  D3 ad = xa - xd, bd = xb - xd, cd = xc - xd;
  D3 n = crs(ad, cd);
  double delta = -dot(n, bd) / dot(n, n);
  D3 ip = n * delta + bd;
  D3 jp = ad - ip, kp = cd - ip;
  double jj = dot(jp, jp), kk = dot(kp, kp);
  if (jj == 0.0 || kk == 0.0) return 0.0;
  double th = acos(clamp1(dot(jp, kp) / sqrt(jj * kk))) * RAD2DEG;
  double e = k * sextic(v.ang_p, th - t0, dp);
  D3 g0, g2;
  angle_legs(jp, kp, k * dp * RAD2DEG, g0, g2);
  D3 gp = (g0 + g2) * -1.0;
  g[0] = g0 + crs(cd, gp) * delta;
  g[2] = g2 + crs(gp, ad) * delta;
  g[1] = gp;
  g[3] = (g[0] + g[1] + g[2]) * -1.0;
  return e;
}

__device__ double term_strbnd(const VDev &v, int t, const double *xyz, D3 *g) {
  const int32_t *q = v.idx[2] + 3 * t;
  D3 xb = ld(xyz, q[1]);
  D3 u = ld(xyz, q[0]) - xb, w = ld(xyz, q[2]) - xb;
  double uu = dot(u, u), ww = dot(w, w);
  g[0] = g[1] = g[2] = D3{0.0, 0.0, 0.0};
  if (uu == 0.0 || ww == 0.0) return 0.0;
  double ru = sqrt(uu), rw = sqrt(ww);
  double th = acos(clamp1(dot(u, w) / (ru * rw))) * RAD2DEG;
  double k0 = v.sb_k[2 * t], k1 = v.sb_k[2 * t + 1];
  double dr = k0 * (ru - v.sb_r0[2 * t]) + k1 * (rw - v.sb_r0[2 * t + 1]);
  double dt = th - v.sb_t0[t];
  angle_legs(u, w, dr * RAD2DEG, g[0], g[2]);
  g[0] = g[0] + u * (k0 * dt / ru);
  g[2] = g[2] + w * (k1 * dt / rw);
  g[1] = (g[0] + g[2]) * -1.0;
  return dr * dt;
}

__device__ double term_urey(const VDev &v, int t, const double *xyz, D3 *g) {
  const int32_t *q = v.idx[3] + 2 * t;
  D3 d = ld(xyz, q[0]) - ld(xyz, q[1]);
  double r = sqrt(dot(d, d)), dv = r - v.ub_r0[t], k = v.ub_k[t];
  double e = k * dv * dv * (1.0 + dv * (v.ub_c3 + dv * v.ub_c4));
  double de = 2.0 * k * dv * (1.0 + dv * (1.5 * v.ub_c3 + 2.0 * dv * v.ub_c4));
  g[0] = d * (r > 0.0 ? de / r : 0.0);
  g[1] = g[0] * -1.0;
  return e;
}

// Allinger out-of-plane angle: sin(theta) = h / |d - b|, h the distance of
// the centre b from the plane through a, c and d: with ee the triple product
// (a - b) . ((c - b) x (d - b)) and cc = |(a - d) x (c - d)|^2, h = ee / sqrt(cc)
__device__ double term_opbend(const VDev &v, int t, const double *xyz, D3 *g) {
  const int32_t *q = v.idx[4] + 4 * t;
  D3 xa = ld(xyz, q[0]), xb = ld(xyz, q[1]), xc = ld(xyz, q[2]), xd = ld(xyz, q[3]);
  D3 ab = xa - xb, cb = xc - xb, db = xd - xb, ad = xa - xd, cd = xc - xd;
  double db2 = dot(db, db), ad2 = dot(ad, ad), cd2 = dot(cd, cd);
  double ee = dot(ab, crs(cb, db));
  double ac = dot(ad, cd);
  double cc = ad2 * cd2 - ac * ac;
  g[0] = g[1] = g[2] = g[3] = D3{0.0, 0.0, 0.0};
  if (db2 == 0.0 || cc == 0.0) return 0.0;
  double bkk2 = db2 - ee * ee / cc;
  double th = acos(clamp1(sqrt(bkk2 / db2))) * RAD2DEG;
  double dp, k = v.op_k[t];
  double e = k * sextic(v.op_p, th, dp);
  // dtheta = sign(ee) / sqrt(cc bkk2) [dee - ee/(2 cc) dcc - ee/(2 db2) d(db2)]
  double f = k * dp * RAD2DEG * (ee < 0.0 ? -1.0 : 1.0) / sqrt(cc * bkk2);
  double s = ee / cc;
  D3 ga = (crs(cb, db) - (ad * cd2 - cd * ac) * s) * f;
  D3 gc = (crs(db, ab) - (cd * ad2 - ad * ac) * s) * f;
  D3 gd = (crs(ab, cb) + (ad * cd2 - cd * ac + cd * ad2 - ad * ac) * s - db * (ee / db2)) * f;
  g[0] = ga;
  g[2] = gc;
  g[3] = gd;
  g[1] = (ga + gc + gd) * -1.0;
  return e;
}

__device__ double term_torsion(const VDev &v, int t, const double *xyz, D3 *g) {
  const int32_t *q = v.idx[5] + 4 * t;
  D3 xa = ld(xyz, q[0]), xb = ld(xyz, q[1]), xc = ld(xyz, q[2]), xd = ld(xyz, q[3]);
  D3 ba = xb - xa, cb = xc - xb, dc = xd - xc;
  D3 tt = crs(ba, cb), uu = crs(cb, dc);
  double rt2 = fmax(dot(tt, tt), 1e-4), ru2 = fmax(dot(uu, uu), 1e-4);
  double rr = sqrt(rt2 * ru2), rcb = fmax(sqrt(dot(cb, cb)), 1e-4);
  double c1 = dot(tt, uu) / rr, s1 = dot(cb, crs(tt, uu)) / (rcb * rr);
  const double *amp = v.tor_amp + 6 * t, *tc = v.tor_cos + 6 * t, *ts = v.tor_sin + 6 * t;
  double e = 0.0, dedphi = 0.0, cn = c1, sn = s1;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (k > 0) {
      double c2 = c1 * cn - s1 * sn, s2 = s1 * cn + c1 * sn;
      cn = c2;
      sn = s2;
    }
    e += amp[k] * (1.0 + cn * tc[k] + sn * ts[k]);
    dedphi += amp[k] * (k + 1) * (cn * ts[k] - sn * tc[k]);
  }
  dihedral_grad(ba, cb, dc, xc - xa, xd - xb, tt, uu, rt2, ru2, rcb, dedphi, g);
  return e;
}

// torsion p-c-d-q of the pi-orbital directions p = c + (a - d) x (b - d),
// q = d + (e - c) x (f - c); E = k (1 - cos 2 phi)
__device__ double term_pitors(const VDev &v, int t, const double *xyz, D3 *g) {
  const int32_t *q = v.idx[6] + 6 * t;
  D3 xa = ld(xyz, q[0]), xb = ld(xyz, q[1]), xc = ld(xyz, q[2]);
  D3 xd = ld(xyz, q[3]), xe = ld(xyz, q[4]), xf = ld(xyz, q[5]);
  D3 ad = xa - xd, bd = xb - xd, ec = xe - xc, fc = xf - xc;
  D3 xp = crs(ad, bd) + xc, xq = crs(ec, fc) + xd;
  D3 pc = xc - xp, dc = xd - xc, dq = xq - xd;
  D3 tt = crs(pc, dc), uu = crs(dc, dq);
  double rt2 = dot(tt, tt), ru2 = dot(uu, uu);
  for (int k = 0; k < 6; k++) g[k] = D3{0.0, 0.0, 0.0};
  if (rt2 * ru2 == 0.0) return 0.0;
  double rr = sqrt(rt2 * ru2), rdc = sqrt(dot(dc, dc));
  double sine = dot(dc, crs(tt, uu)) / (rdc * rr);
  double cosine = clamp1(dot(tt, uu) / rr);
  double k = v.pi_k[t];
  double e = k * (1.0 - (cosine * cosine - sine * sine));
  double dedphi = 4.0 * k * cosine * sine;
  D3 gq4[4];
  dihedral_grad(pc, dc, dq, xd - xp, xq - xc, tt, uu, rt2, ru2, rdc, dedphi, gq4);
  // gq4: gradient on p, c, d, q of the torsion p-c-d-q
  D3 gp = gq4[0], gq = gq4[3];
  D3 ga = crs(bd, gp), gb = crs(gp, ad), ge = crs(fc, gq), gf = crs(gq, ec);
  g[0] = ga;
  g[1] = gb;
  g[2] = gq4[1] + gp - ge - gf;
  g[3] = gq4[2] + gq - ga - gb;
  g[4] = ge;
  g[5] = gf;
  return e;
}

// bicubic Hermite patch from the corner values, slopes and cross slopes
// (corners counter-clockwise from (xl, yl)); value and both derivatives
__device__ double bicubic(double hx, double hy, double u, double w, const double *e, const double *dx,
                          const double *dy, const double *dxy, double &d1, double &d2) {
  double bu[4] = {(2 * u - 3) * u * u + 1, ((u - 2) * u + 1) * u, (3 - 2 * u) * u * u, (u - 1) * u * u};
  double du[4] = {6 * u * (u - 1), (3 * u - 4) * u + 1, 6 * u * (1 - u), (3 * u - 2) * u};
  double bw[4] = {(2 * w - 3) * w * w + 1, ((w - 2) * w + 1) * w, (3 - 2 * w) * w * w, (w - 1) * w * w};
  double dw[4] = {6 * w * (w - 1), (3 * w - 4) * w + 1, 6 * w * (1 - w), (3 * w - 2) * w};
  const int cx[4] = {0, 2, 2, 0}, cy[4] = {0, 0, 2, 2};
  double f = 0.0, fu = 0.0, fw = 0.0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int iu = cx[k], iw = cy[k];
    double a0 = e[k], a1 = dx[k] * hx, a2 = dy[k] * hy, a3 = dxy[k] * hx * hy;
    f += (a0 * bu[iu] + a1 * bu[iu + 1]) * bw[iw] + (a2 * bu[iu] + a3 * bu[iu + 1]) * bw[iw + 1];
    fu += (a0 * du[iu] + a1 * du[iu + 1]) * bw[iw] + (a2 * du[iu] + a3 * du[iu + 1]) * bw[iw + 1];
    fw += (a0 * bu[iu] + a1 * bu[iu + 1]) * dw[iw] + (a2 * bu[iu] + a3 * bu[iu + 1]) * dw[iw + 1];
  }
  d1 = fu / hx;
  d2 = fw / hy;
  return f;
}

__device__ int bracket(const double *x, int n, double v) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int m = (hi + lo) >> 1;
    if (x[m] > v) hi = m; else lo = m;
  }
  return lo;
}

__device__ double term_tortor(const VDev &v, int t, const double *xyz, D3 *g) {
  const int32_t *q = v.idx[7] + 6 * t;
  D3 x0 = ld(xyz, q[0]), x1 = ld(xyz, q[1]), x2 = ld(xyz, q[2]), x3 = ld(xyz, q[3]), x4 = ld(xyz, q[4]);
  D3 ba = x1 - x0, cb = x2 - x1, dc = x3 - x2, ed = x4 - x3;
  D3 tt = crs(ba, cb), uu = crs(cb, dc), vv = crs(dc, ed);
  double rt2 = dot(tt, tt), ru2 = dot(uu, uu), rv2 = dot(vv, vv);
  for (int k = 0; k < 5; k++) g[k] = D3{0.0, 0.0, 0.0};
  if (rt2 * ru2 == 0.0 || ru2 * rv2 == 0.0) return 0.0;
  double a1 = acos(clamp1(dot(tt, uu) / sqrt(rt2 * ru2))) * RAD2DEG;
  if (dot(ba, uu) < 0.0) a1 = -a1;
  double a2 = acos(clamp1(dot(uu, vv) / sqrt(ru2 * rv2))) * RAD2DEG;
  if (dot(cb, vv) < 0.0) a2 = -a2;
  // chirality at the centre: sign of the volume spanned by (chiral - c),
  // (b - c), (d - c)
  double sign = 1.0;
  if (q[5] >= 0) {
    D3 c0 = ld(xyz, q[5]) - x2, c1 = x1 - x2, c2 = x3 - x2;
    if (dot(c0, crs(c1, c2)) < 0.0) sign = -1.0;
  }
  double t1 = sign * a1, t2 = sign * a2;
  int gi = v.tt_grid[t], nx = v.g_nx[gi], ny = v.g_ny[gi];
  const double *tx = v.g_t + v.g_off_t[gi], *ty = tx + nx;
  int xl = bracket(tx, nx, t1), yl = bracket(ty, ny, t2);
  const int o = v.g_off_e[gi], p1 = o + yl * nx + xl, p2 = p1 + nx;
  const int id[4] = {p1, p1 + 1, p2 + 1, p2};
  double ce[4], cdx[4], cdy[4], cdxy[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    ce[k] = v.g_e[id[k]];
    cdx[k] = v.g_dx[id[k]];
    cdy[k] = v.g_dy[id[k]];
    cdxy[k] = v.g_dxy[id[k]];
  }
  double hx = tx[xl + 1] - tx[xl], hy = ty[yl + 1] - ty[yl];
  double d1, d2;
  double e = bicubic(hx, hy, (t1 - tx[xl]) / hx, (t2 - ty[yl]) / hy, ce, cdx, cdy, cdxy, d1, d2);
  double rcb = sqrt(dot(cb, cb)), rdc = sqrt(dot(dc, dc));
  D3 ga[4], gb[4];
  dihedral_grad(ba, cb, dc, x2 - x0, x3 - x1, tt, uu, rt2, ru2, rcb, sign * d1 * RAD2DEG, ga);
  dihedral_grad(cb, dc, ed, x3 - x1, x4 - x2, uu, vv, ru2, rv2, rdc, sign * d2 * RAD2DEG, gb);
  g[0] = ga[0];
  g[1] = ga[1] + gb[0];
  g[2] = ga[2] + gb[1];
  g[3] = ga[3] + gb[2];
  g[4] = gb[3];
  return e;
}

__global__ void __launch_bounds__(256) k_valence(VDev v, const double *__restrict__ xyz,
                                                 double *__restrict__ slot, double *__restrict__ eterm) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= v.start[NKIND]) return;
  int k = 0;
  while (gid >= v.start[k + 1]) k++;
  const int t = gid - v.start[k];
  D3 g[6];
  double e;
  switch (k) {
    case 0: e = term_bond(v, t, xyz, g); break;
    case 1: e = term_angle(v, t, xyz, g); break;
    case 2: e = term_strbnd(v, t, xyz, g); break;
    case 3: e = term_urey(v, t, xyz, g); break;
    case 4: e = term_opbend(v, t, xyz, g); break;
    case 5: e = term_torsion(v, t, xyz, g); break;
    case 6: e = term_pitors(v, t, xyz, g); break;
    default: e = term_tortor(v, t, xyz, g); break;
  }
  eterm[gid] = e;
  const int na = natoms_of(k);
  double *s = slot + 3 * (v.sbase[k] + (long)t * na);
  for (int j = 0; j < na; j++) {
    s[3 * j] = g[j].x;
    s[3 * j + 1] = g[j].y;
    s[3 * j + 2] = g[j].z;
  }
}

// GL lanes per atom: lane j sums slots j, j + GL, ... of the atom's list,
// then a fixed butterfly over the GL lanes (the same order on every run)
constexpr int GL = 8;

__global__ void __launch_bounds__(256) k_valence_gather(int n, const int32_t *__restrict__ off,
                                                        const int32_t *__restrict__ list,
                                                        const double *__restrict__ slot, const double *__restrict__ xyz,
                                                        double *__restrict__ grad, double *__restrict__ vpart) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = g / GL, lane = g % GL;
  double gx = 0.0, gy = 0.0, gz = 0.0;
  if (i < n) {
    for (int k = off[i] + lane; k < off[i + 1]; k += GL) {
      const double *s = slot + 3 * (long)list[k];
      gx += s[0];
      gy += s[1];
      gz += s[2];
    }
  }
#pragma unroll
  for (int o = GL / 2; o > 0; o >>= 1) {
    gx += __shfl_xor(gx, o, 64);
    gy += __shfl_xor(gy, o, 64);
    gz += __shfl_xor(gz, o, 64);
  }
  if (i < n && lane == 0) {
    grad[3 * i] += gx;
    grad[3 * i + 1] += gy;
    grad[3 * i + 2] += gz;
  }
  if (vpart) {
    double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n && lane == 0) {
      const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
      w[0] = gx * x; w[1] = gx * y; w[2] = gx * z;
      w[3] = gy * x; w[4] = gy * y; w[5] = gy * z;
      w[6] = gz * x; w[7] = gz * y; w[8] = gz * z;
    }
    block_store9(w, vpart);
  }
}

// one workgroup per kind: the kind's term energies in a fixed order
__global__ void __launch_bounds__(256) k_valence_energy(VDev v, const double *__restrict__ eterm,
                                                        double *__restrict__ out) {
  __shared__ double part[4];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int t = v.start[k] + threadIdx.x; t < v.start[k + 1]; t += blockDim.x) s += eterm[t];
  s = wave_sum_d(s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[k] = part[0] + part[1] + part[2] + part[3];
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void nb_valence_terms(NbHandle *h) {
  if (!h->has_valence) return;
  const VDev &v = *(const VDev *)h->val_dev;
  NB_LAUNCH(k_valence, dim3((unsigned)((h->val_terms + 255) / 256)), dim3(256), 0, v, (const double *)h->d_xyz,
            h->d_val_slot, h->d_val_eterm);
}

void nb_valence_gather(NbHandle *h) {
  if (!h->has_valence) return;
  const VDev &v = *(const VDev *)h->val_dev;
  const int nb = (int)(((long)h->n * GL + 255) / 256);
  double *vp = h->want_vir ? nb_vir_region(h, nb) : nullptr;
  NB_LAUNCH(k_valence_gather, dim3(nb), dim3(256), 0, h->n, (const int32_t *)h->d_val_off,
            (const int32_t *)h->d_val_list, (const double *)h->d_val_slot, (const double *)h->d_xyz, h->d_grad, vp);
  NB_LAUNCH(k_valence_energy, dim3(NKIND), dim3(256), 0, v, (const double *)h->d_val_eterm, h->d_val_energy);
}

void nb_valence_free(NbHandle *h) {
  void *p[] = {h->d_val_i, h->d_val_d, h->d_val_slot, h->d_val_off, h->d_val_list, h->d_val_eterm, h->d_val_energy};
  for (void *q : p)
    if (q) hipFree(q);
  h->d_val_i = nullptr;
  h->d_val_d = nullptr;
  h->d_val_slot = nullptr;
  h->d_val_off = nullptr;
  h->d_val_list = nullptr;
  h->d_val_eterm = nullptr;
  h->d_val_energy = nullptr;
  delete (VDev *)h->val_dev;
  h->val_dev = nullptr;
  h->has_valence = false;
  h->val_terms = h->val_slots = 0;
}

int nb_valence_struct_size(void) { return (int)sizeof(nb_valence); }

int nb_set_valence(void *handle, const nb_valence *v) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  try {
    HIPCHECK(hipSetDevice(h->device));
    HIPCHECK(hipStreamSynchronize(h->stream));
    nb_drop_graph(h);
    nb_valence_free(h);
    if (!v) return NB_OK;
    const int n = h->n;
    const int cnt[NKIND] = {v->n_bond, v->n_angle, v->n_strbnd, v->n_urey, v->n_opbend, v->n_torsion, v->n_pitors,
                            v->n_tortor};
    const int32_t *idx[NKIND] = {v->bond, v->angle, v->strbnd, v->urey, v->opbend, v->torsion, v->pitors, v->tortor};
    long total = 0;
    for (int k = 0; k < NKIND; k++) {
      if (cnt[k] < 0 || (cnt[k] > 0 && !idx[k])) throw std::invalid_argument("valence: bad term count or array");
      total += cnt[k];
    }
    if (total == 0) return NB_OK;
    if (total > (1L << 30)) throw std::invalid_argument("valence: too many terms");
    // every index a kernel dereferences is checked here
    for (int k = 0; k < NKIND; k++)
      for (long t = 0; t < cnt[k]; t++)
        for (int j = 0; j < NIDX[k]; j++) {
          int a = idx[k][NIDX[k] * t + j];
          bool optional = (k == 1 && j == 3) || (k == 7 && j == 5);
          if (a < -1 || a >= n || (a == -1 && !optional))
            throw std::invalid_argument("valence: atom index out of range");
        }
    long len_t = 0, len_e = 0;
    if (v->n_tortor > 0) {
      if (v->n_grid <= 0 || !v->tortor_grid || !v->grid_nx || !v->grid_ny || !v->grid_off_t || !v->grid_off_e ||
          !v->grid_t || !v->grid_e || !v->grid_dx || !v->grid_dy || !v->grid_dxy)
        throw std::invalid_argument("valence: torsion-torsion terms without grids");
      for (int g = 0; g < v->n_grid; g++) {
        if (v->grid_nx[g] < 2 || v->grid_ny[g] < 2 || v->grid_off_t[g] < 0 || v->grid_off_e[g] < 0)
          throw std::invalid_argument("valence: bad torsion-torsion grid");
        len_t = std::max(len_t, (long)v->grid_off_t[g] + v->grid_nx[g] + v->grid_ny[g]);
        len_e = std::max(len_e, (long)v->grid_off_e[g] + (long)v->grid_nx[g] * v->grid_ny[g]);
      }
      for (int t = 0; t < v->n_tortor; t++)
        if (v->tortor_grid[t] < 0 || v->tortor_grid[t] >= v->n_grid)
          throw std::invalid_argument("valence: torsion-torsion grid index out of range");
    }
    const int ngrid = v->n_tortor > 0 ? v->n_grid : 0;
    // integers: indices of every kind, tortor grid ids, grid nx, ny, offsets
    std::vector<int32_t> I;
    long ioff[NKIND], iextra;
    for (int k = 0; k < NKIND; k++) {
      ioff[k] = (long)I.size();
      I.insert(I.end(), idx[k], idx[k] + (long)NIDX[k] * cnt[k]);
    }
    iextra = (long)I.size();
    if (ngrid) {
      I.insert(I.end(), v->tortor_grid, v->tortor_grid + v->n_tortor);
      I.insert(I.end(), v->grid_nx, v->grid_nx + ngrid);
      I.insert(I.end(), v->grid_ny, v->grid_ny + ngrid);
      I.insert(I.end(), v->grid_off_t, v->grid_off_t + ngrid);
      I.insert(I.end(), v->grid_off_e, v->grid_off_e + ngrid);
    }
    // doubles, in a fixed order
    std::vector<double> D;
    auto put = [&](const double *p, long m, const char *what) -> long {
      long o = (long)D.size();
      if (m > 0 && !p) throw std::invalid_argument(std::string("valence: missing ") + what);
      if (m > 0) D.insert(D.end(), p, p + m);
      return o;
    };
    long o_bk = put(v->bond_k, cnt[0], "bond_k"), o_br = put(v->bond_r0, cnt[0], "bond_r0");
    long o_ak = put(v->angle_k, cnt[1], "angle_k"), o_at = put(v->angle_t0, cnt[1], "angle_t0");
    long o_sk = put(v->strbnd_k, 2L * cnt[2], "strbnd_k"), o_sr = put(v->strbnd_r0, 2L * cnt[2], "strbnd_r0");
    long o_st = put(v->strbnd_t0, cnt[2], "strbnd_t0");
    long o_uk = put(v->urey_k, cnt[3], "urey_k"), o_ur = put(v->urey_r0, cnt[3], "urey_r0");
    long o_ok = put(v->opbend_k, cnt[4], "opbend_k");
    long o_ta = put(v->torsion_amp, 6L * cnt[5], "torsion_amp"), o_tc = put(v->torsion_cos, 6L * cnt[5], "torsion_cos");
    long o_ts = put(v->torsion_sin, 6L * cnt[5], "torsion_sin");
    long o_pk = put(v->pitors_k, cnt[6], "pitors_k");
    long o_gt = put(v->grid_t, ngrid ? len_t : 0, "grid_t"), o_ge = put(v->grid_e, ngrid ? len_e : 0, "grid_e");
    long o_gx = put(v->grid_dx, ngrid ? len_e : 0, "grid_dx"), o_gy = put(v->grid_dy, ngrid ? len_e : 0, "grid_dy");
    long o_gxy = put(v->grid_dxy, ngrid ? len_e : 0, "grid_dxy");
    // slots and the per-atom lists
    long sbase[NKIND], nslot = 0;
    for (int k = 0; k < NKIND; k++) {
      sbase[k] = nslot;
      nslot += (long)NATOMS[k] * cnt[k];
    }
    std::vector<int32_t> off(n + 1, 0);
    for (int k = 0; k < NKIND; k++)
      for (long t = 0; t < cnt[k]; t++)
        for (int j = 0; j < NATOMS[k]; j++) {
          int a = idx[k][NIDX[k] * t + j];
          if (a >= 0) off[a + 1]++;
        }
    for (int i = 0; i < n; i++) off[i + 1] += off[i];
    std::vector<int32_t> fill(off.begin(), off.end() - 1), list(off[n]);
    for (int k = 0; k < NKIND; k++)
      for (long t = 0; t < cnt[k]; t++)
        for (int j = 0; j < NATOMS[k]; j++) {
          int a = idx[k][NIDX[k] * t + j];
          if (a >= 0) list[fill[a]++] = (int32_t)(sbase[k] + t * NATOMS[k] + j);
        }
    HIPCHECK(hipMalloc(&h->d_val_i, sizeof(int32_t) * I.size()));
    HIPCHECK(hipMalloc(&h->d_val_d, sizeof(double) * std::max<size_t>(D.size(), 1)));
    HIPCHECK(hipMalloc(&h->d_val_slot, sizeof(double) * 3 * std::max<long>(nslot, 1)));
    HIPCHECK(hipMalloc(&h->d_val_off, sizeof(int32_t) * (n + 1)));
    HIPCHECK(hipMalloc(&h->d_val_list, sizeof(int32_t) * std::max<size_t>(list.size(), 1)));
    HIPCHECK(hipMalloc(&h->d_val_eterm, sizeof(double) * total));
    HIPCHECK(hipMalloc(&h->d_val_energy, sizeof(double) * NKIND));
    HIPCHECK(hipMemcpy(h->d_val_i, I.data(), sizeof(int32_t) * I.size(), hipMemcpyHostToDevice));
    if (!D.empty()) HIPCHECK(hipMemcpy(h->d_val_d, D.data(), sizeof(double) * D.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h->d_val_off, off.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
    if (!list.empty())
      HIPCHECK(hipMemcpy(h->d_val_list, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice));
    VDev *d = new VDev();
    h->val_dev = d;
    d->start[0] = 0;
    for (int k = 0; k < NKIND; k++) {
      d->start[k + 1] = d->start[k] + cnt[k];
      d->sbase[k] = sbase[k];
      d->idx[k] = h->d_val_i + ioff[k];
    }
    const double *B = h->d_val_d;
    d->bond_k = B + o_bk;
    d->bond_r0 = B + o_br;
    d->bond_c3 = v->bond_cubic;
    d->bond_c4 = v->bond_quartic;
    d->ang_k = B + o_ak;
    d->ang_t0 = B + o_at;
    d->sb_k = B + o_sk;
    d->sb_r0 = B + o_sr;
    d->sb_t0 = B + o_st;
    d->ub_k = B + o_uk;
    d->ub_r0 = B + o_ur;
    d->ub_c3 = v->urey_cubic;
    d->ub_c4 = v->urey_quartic;
    d->op_k = B + o_ok;
    for (int q = 0; q < 4; q++) {
      d->ang_p[q] = v->angle_poly[q];
      d->op_p[q] = v->opbend_poly[q];
    }
    d->tor_amp = B + o_ta;
    d->tor_cos = B + o_tc;
    d->tor_sin = B + o_ts;
    d->pi_k = B + o_pk;
    const int32_t *X = h->d_val_i + iextra;
    d->tt_grid = X;
    d->g_nx = X + (ngrid ? v->n_tortor : 0);
    d->g_ny = d->g_nx + ngrid;
    d->g_off_t = d->g_ny + ngrid;
    d->g_off_e = d->g_off_t + ngrid;
    d->g_t = B + o_gt;
    d->g_e = B + o_ge;
    d->g_dx = B + o_gx;
    d->g_dy = B + o_gy;
    d->g_dxy = B + o_gxy;
    h->val_terms = total;
    h->val_slots = nslot;
    h->has_valence = true;
    return NB_OK;
  } catch (std::invalid_argument &e) {
    nb_valence_free(h);
    h->err = e.what();
    return NB_ERR_INVALID;
  } catch (std::exception &e) {
    nb_valence_free(h);
    h->err = e.what();
    return NB_ERR_HIP;
  }
}
