// nb_atom.h -- per-atom AMOEBA terms shared by the per-system kernels and the
// batched small-cell kernel: frame rotation, fractional multipoles and
// potentials, the reciprocal energy / gradient / torque of one atom, and the
// torque-to-frame-gradient rules.
#pragma once
#include "nb_kern.h"

struct Frac { double a[3][3]; };  // a[f][c] = K_f A[c][f]

__device__ __forceinline__ V3<double> ld3(const double *x, int i) { return V3<double>{x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
__device__ __forceinline__ V3<double> unit(V3<double> a) { return (1.0 / sqrt(dot(a, a))) * a; }

// Global-frame multipole of atom i (MultipoleType.getRotationMatrix,
// MultipoleType.java:385-540, and rotateMultipole, :1031-1058): g = q, d(3),
// full Q (xx, yy, zz, xy, xz, yz).
__device__ __forceinline__ void rotate_multipole(int i, const double *__restrict__ xyz,
                                                 const int32_t *__restrict__ frame, const int32_t *__restrict__ axis,
                                                 const double *__restrict__ mloc, double g[10]) {
  V3<double> o = ld3(xyz, i);
  int f = frame[i];
  double R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  if (f != NB_FRAME_NONE) {
    int az = axis[3 * i], ax = axis[3 * i + 1], ay = axis[3 * i + 2];
    V3<double> z, x;
    V3<double> pz = ld3(xyz, az);
    if (f == NB_FRAME_BISECTOR) {
      z = unit(pz - o);
      x = unit(ld3(xyz, ax) - o);
      z = unit(x + z);
      x = unit(x - dot(x, z) * z);
    } else if (f == NB_FRAME_ZTHENBISECTOR) {
      z = unit(pz - o);
      V3<double> a = unit(ld3(xyz, ax) - o), b = unit(ld3(xyz, ay) - o);
      x = unit(a + b);
      x = unit(x - dot(x, z) * z);
    } else if (f == NB_FRAME_ZONLY) {
      z = unit(pz - o);
      x = V3<double>{1.0, 0.0, 0.0};
      double dt = z.x;
      if (fabs(dt) > 0.866) { x = V3<double>{0.0, 1.0, 0.0}; dt = z.y; }
      x = unit(x - dt * z);
    } else if (f == NB_FRAME_THREEFOLD) {
      V3<double> a = unit(pz - o), b = unit(ld3(xyz, ax) - o), c = unit(ld3(xyz, ay) - o);
      z = unit(a + b + c);
      x = unit(a - dot(a, z) * z);
    } else {
      z = unit(pz - o);
      x = ld3(xyz, ax) - o;
      x = unit(x - dot(x, z) * z);
    }
    V3<double> y = cross(z, x);
    R[0][0] = x.x; R[1][0] = x.y; R[2][0] = x.z;
    R[0][1] = y.x; R[1][1] = y.y; R[2][1] = y.z;
    R[0][2] = z.x; R[1][2] = z.y; R[2][2] = z.z;
  }
  const double *m = mloc + 10 * i;
  double d[3], Q[3][3] = {{m[4], m[7], m[8]}, {m[7], m[5], m[9]}, {m[8], m[9], m[6]}}, T[3][3], O[3][3];
  for (int a = 0; a < 3; a++) d[a] = R[a][0] * m[1] + R[a][1] * m[2] + R[a][2] * m[3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) T[a][b] = R[a][0] * Q[0][b] + R[a][1] * Q[1][b] + R[a][2] * Q[2][b];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) O[a][b] = T[a][0] * R[b][0] + T[a][1] * R[b][1] + T[a][2] * R[b][2];
  g[0] = m[0]; g[1] = d[0]; g[2] = d[1]; g[3] = d[2];
  g[4] = O[0][0]; g[5] = O[1][1]; g[6] = O[2][2]; g[7] = O[0][1]; g[8] = O[0][2]; g[9] = O[1][2];
}

// fractional multipole of a global one (ReciprocalSpace.toFractionalMultipole,
// ReciprocalSpace.java:2778-2834; quadrupole with the 1/3 pre-factor, :1234-1238)
__device__ __forceinline__ void frac_multipole(const Frac &fr, const double *g, double *f) {
  f[0] = g[0];
  for (int j = 0; j < 3; j++) f[1 + j] = fr.a[j][0] * g[1] + fr.a[j][1] * g[2] + fr.a[j][2] * g[3];
  double Q[3][3] = {{g[4], g[7], g[8]}, {g[7], g[5], g[9]}, {g[8], g[9], g[6]}}, T[3][3], F[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) T[a][b] = fr.a[a][0] * Q[0][b] + fr.a[a][1] * Q[1][b] + fr.a[a][2] * Q[2][b];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) F[a][b] = T[a][0] * fr.a[b][0] + T[a][1] * fr.a[b][1] + T[a][2] * fr.a[b][2];
  const double t = 1.0 / 3.0;
  f[4] = F[0][0] * t; f[5] = F[1][1] * t; f[6] = F[2][2] * t;
  f[7] = 2.0 * F[0][1] * t; f[8] = 2.0 * F[0][2] * t; f[9] = 2.0 * F[1][2] * t;
}

template <typename T>
__device__ __forceinline__ void phi_cart(const Frac &fr, const T *f, double *c) {
  c[0] = f[0];
  for (int j = 0; j < 3; j++) c[1 + j] = fr.a[0][j] * f[1] + fr.a[1][j] * f[2] + fr.a[2][j] * f[3];
  double F[3][3] = {{f[4], f[7], f[8]}, {f[7], f[5], f[9]}, {f[8], f[9], f[6]}}, T3[3][3], C[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) T3[a][b] = fr.a[0][a] * F[0][b] + fr.a[1][a] * F[1][b] + fr.a[2][a] * F[2][b];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) C[a][b] = T3[a][0] * fr.a[0][b] + T3[a][1] * fr.a[1][b] + T3[a][2] * fr.a[2][b];
  c[4] = C[0][0]; c[5] = C[1][1]; c[6] = C[2][2]; c[7] = C[0][1]; c[8] = C[0][2]; c[9] = C[1][2];
}


// Reciprocal-space terms of atom s (ReciprocalEnergyRegion.java:381-746):
// permanent self and reciprocal energy, gradient df and torque tq; with polar,
// the induced self / reciprocal energy, gradient dq and torque tqq (u and uCR
// in the permanent potential, the multipoles in the potentials of u + uCR,
// and for mutual polarization u / uCR in each other's potential).  w9 (may be
// null) receives the atom's strain-derivative terms scaled by sperm / spol.
struct RecipAtom { double eself, erec, qt, eselfi, ereci, df[3], tq[3], dq[3], tqq[3]; };

__device__ __forceinline__ void recip_atom(int s, const double *__restrict__ gmp, const double *__restrict__ fmp,
                                           const double *__restrict__ phip, const double *__restrict__ phiu,
                                           const double *__restrict__ phiup, const double *__restrict__ u,
                                           const double *__restrict__ up, const Frac &fr, double el, double aw,
                                           int polar, int mutual, double sperm, double spol, double *w9,
                                           RecipAtom &o) {
  double eself = 0, erec = 0, qt = 0, eselfi = 0, ereci = 0;
  {
    const double *m = gmp + 10 * s, *fm = fmp + 10 * s;
    const double *fp = phip + 30 * (size_t)s, *ph = fp + 20;
    double aewald1 = -el * aw / NB_SQRT_PI, aewald2 = 2.0 * aw * aw;
    double aewald3 = -2.0 / 3.0 * el * aw * aw * aw / NB_SQRT_PI, aewald4 = -2.0 * aewald3;
    qt = m[0];
    double cii = m[0] * m[0], dii = m[1] * m[1] + m[2] * m[2] + m[3] * m[3];
    double qii = m[4] * m[4] + m[5] * m[5] + m[6] * m[6] + 2.0 * (m[7] * m[7] + m[8] * m[8] + m[9] * m[9]);
    eself = aewald1 * (cii + aewald2 * (dii / 3.0 + 2.0 * aewald2 * qii / 45.0));
    erec = m[0] * ph[0] + m[1] * ph[1] + m[2] * ph[2] + m[3] * ph[3] +
           (m[4] * ph[4] + m[5] * ph[5] + m[6] * ph[6] + 2.0 * (m[7] * ph[7] + m[8] * ph[8] + m[9] * ph[9])) / 3.0;
    double g[3];
    g[0] = fm[0] * fp[1] + fm[1] * fp[4] + fm[2] * fp[7] + fm[3] * fp[8] + fm[4] * fp[10] + fm[5] * fp[15] +
           fm[6] * fp[17] + fm[7] * fp[13] + fm[8] * fp[14] + fm[9] * fp[19];
    g[1] = fm[0] * fp[2] + fm[1] * fp[7] + fm[2] * fp[5] + fm[3] * fp[9] + fm[4] * fp[13] + fm[5] * fp[11] +
           fm[6] * fp[18] + fm[7] * fp[15] + fm[8] * fp[19] + fm[9] * fp[16];
    g[2] = fm[0] * fp[3] + fm[1] * fp[8] + fm[2] * fp[9] + fm[3] * fp[6] + fm[4] * fp[14] + fm[5] * fp[16] +
           fm[6] * fp[12] + fm[7] * fp[19] + fm[8] * fp[17] + fm[9] * fp[18];
    const double tt = 2.0 / 3.0;
    double df[3], tq[3];
    for (int a = 0; a < 3; a++) df[a] = el * (fr.a[0][a] * g[0] + fr.a[1][a] * g[1] + fr.a[2][a] * g[2]);
    tq[0] = -m[2] * ph[3] + m[3] * ph[2] - tt * (m[7] * ph[8] + m[5] * ph[9] + m[9] * ph[6] - m[8] * ph[7] - m[9] * ph[5] - m[6] * ph[9]);
    tq[1] = -m[3] * ph[1] + m[1] * ph[3] - tt * (m[8] * ph[4] + m[9] * ph[7] + m[6] * ph[8] - m[4] * ph[8] - m[7] * ph[9] - m[8] * ph[6]);
    tq[2] = -m[1] * ph[2] + m[2] * ph[1] - tt * (m[4] * ph[7] + m[7] * ph[5] + m[8] * ph[9] - m[7] * ph[4] - m[5] * ph[7] - m[9] * ph[8]);
    for (int a = 0; a < 3; a++) tq[a] *= el;
    double dq[3] = {0, 0, 0}, tqq[3] = {0, 0, 0};   // polarization part
    if (polar) {
      const double *ui = u + 3 * s, *upi = up + 3 * s;
      eselfi = aewald3 * (ui[0] * m[1] + ui[1] * m[2] + ui[2] * m[3]);
      double us[3] = {0.5 * (ui[0] + upi[0]), 0.5 * (ui[1] + upi[1]), 0.5 * (ui[2] + upi[2])};
      tqq[0] = aewald4 * (m[2] * us[2] - m[3] * us[1]);
      tqq[1] = aewald4 * (m[3] * us[0] - m[1] * us[2]);
      tqq[2] = aewald4 * (m[1] * us[1] - m[2] * us[0]);
      double fu[3], fpp[3];
      for (int j = 0; j < 3; j++) {
        fu[j] = fr.a[j][0] * ui[0] + fr.a[j][1] * ui[1] + fr.a[j][2] * ui[2];
        fpp[j] = fr.a[j][0] * upi[0] + fr.a[j][1] * upi[1] + fr.a[j][2] * upi[2];
      }
      ereci = fu[0] * fp[1] + fu[1] * fp[2] + fu[2] * fp[3];
      const double *fi = phiu + 30 * (size_t)s, *fip = phiup + 30 * (size_t)s;
      const double *ci = fi + 20, *cip = fip + 20;
      double sp[10], sf[20];
      for (int t = 0; t < 10; t++) sp[t] = 0.5 * (ci[t] + cip[t]);
      for (int t = 0; t < 20; t++) sf[t] = fi[t] + fip[t];
      double sv[3] = {fu[0] + fpp[0], fu[1] + fpp[1], fu[2] + fpp[2]};
      double h[3];
      h[0] = sv[0] * fp[4] + sv[1] * fp[7] + sv[2] * fp[8];
      h[1] = sv[0] * fp[7] + sv[1] * fp[5] + sv[2] * fp[9];
      h[2] = sv[0] * fp[8] + sv[1] * fp[9] + sv[2] * fp[6];
      if (mutual) {
        h[0] += fu[0] * fip[4] + fpp[0] * fi[4] + fu[1] * fip[7] + fpp[1] * fi[7] + fu[2] * fip[8] + fpp[2] * fi[8];
        h[1] += fu[0] * fip[7] + fpp[0] * fi[7] + fu[1] * fip[5] + fpp[1] * fi[5] + fu[2] * fip[9] + fpp[2] * fi[9];
        h[2] += fu[0] * fip[8] + fpp[0] * fi[8] + fu[1] * fip[9] + fpp[1] * fi[9] + fu[2] * fip[6] + fpp[2] * fi[6];
      }
      h[0] += fm[0] * sf[1] + fm[1] * sf[4] + fm[2] * sf[7] + fm[3] * sf[8] + fm[4] * sf[10] + fm[5] * sf[15] +
              fm[6] * sf[17] + fm[7] * sf[13] + fm[8] * sf[14] + fm[9] * sf[19];
      h[1] += fm[0] * sf[2] + fm[1] * sf[7] + fm[2] * sf[5] + fm[3] * sf[9] + fm[4] * sf[13] + fm[5] * sf[11] +
              fm[6] * sf[18] + fm[7] * sf[15] + fm[8] * sf[19] + fm[9] * sf[16];
      h[2] += fm[0] * sf[3] + fm[1] * sf[8] + fm[2] * sf[9] + fm[3] * sf[6] + fm[4] * sf[14] + fm[5] * sf[16] +
              fm[6] * sf[12] + fm[7] * sf[19] + fm[8] * sf[17] + fm[9] * sf[18];
      for (int a = 0; a < 3; a++) dq[a] = 0.5 * el * (fr.a[0][a] * h[0] + fr.a[1][a] * h[1] + fr.a[2][a] * h[2]);
      tqq[0] += el * (-m[2] * sp[3] + m[3] * sp[2] - tt * (m[7] * sp[8] + m[5] * sp[9] + m[9] * sp[6] - m[8] * sp[7] - m[9] * sp[5] - m[6] * sp[9]));
      tqq[1] += el * (-m[3] * sp[1] + m[1] * sp[3] - tt * (m[8] * sp[4] + m[9] * sp[7] + m[6] * sp[8] - m[4] * sp[8] - m[7] * sp[9] - m[8] * sp[6]));
      tqq[2] += el * (-m[1] * sp[2] + m[2] * sp[1] - tt * (m[4] * sp[7] + m[7] * sp[5] + m[8] * sp[9] - m[7] * sp[4] - m[5] * sp[7] - m[9] * sp[8]));
    }
    for (int a = 0; a < 3; a++) {
      o.df[a] = df[a];
      o.tq[a] = tq[a];
      o.dq[a] = dq[a];
      o.tqq[a] = tqq[a];
    }
    if (w9) {
      // Strain derivative with the global multipoles held fixed: the wave
      // vectors' strain dependence gives -d_b dphi_a - (2/3) Q_bc dphi_ac per
      // atom; polarization adds the same for u + uCR in the permanent
      // potential, the multipoles in the potentials of u + uCR and (mutual)
      // u / uCR in each other's potential, each with weight 1/2.
      double Q[3][3] = {{m[4], m[7], m[8]}, {m[7], m[5], m[9]}, {m[8], m[9], m[6]}};
      double H[3][3] = {{ph[4], ph[7], ph[8]}, {ph[7], ph[5], ph[9]}, {ph[8], ph[9], ph[6]}};
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          double t = m[1 + b] * ph[1 + a];
          for (int c = 0; c < 3; c++) t += 2.0 / 3.0 * Q[b][c] * H[a][c];
          w9[3 * a + b] -= sperm * el * t;
        }
      if (polar) {
        const double *ui = u + 3 * s, *vi = up + 3 * s;
        const double *cu = phiu + 30 * (size_t)s + 20, *cv = phiup + 30 * (size_t)s + 20;
        double hw[6];
        for (int t = 0; t < 6; t++) hw[t] = cu[4 + t] + cv[4 + t];
        double Hw[3][3] = {{hw[0], hw[3], hw[4]}, {hw[3], hw[1], hw[5]}, {hw[4], hw[5], hw[2]}};
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            double t = (ui[b] + vi[b]) * ph[1 + a] + m[1 + b] * (cu[1 + a] + cv[1 + a]);
            for (int c = 0; c < 3; c++) t += 2.0 / 3.0 * Q[b][c] * Hw[a][c];
            if (mutual) t += vi[b] * cu[1 + a] + ui[b] * cv[1 + a];
            w9[3 * a + b] -= spol * 0.5 * el * t;
          }
      }
    }
    }
  o.eself = eself;
  o.erec = erec;
  o.qt = qt;
  o.eselfi = eselfi;
  o.ereci = ereci;
}

__device__ __forceinline__ V3<double> unit_d(V3<double> a) { return (1.0 / sqrt(dot(a, a))) * a; }

// Torque on atom i to gradients on its frame atoms z, x, y (Torque.torque,
// Torque.java:111-316); the centre takes -(gz + gx + gy).  vz, vx, vy are the
// frame vectors (for the strain term).
__device__ void torque_frame(int fr, int i, int iz, int ix, int iy, const double *__restrict__ xyz, V3<double> trq,
                             V3<double> &gz, V3<double> &gx, V3<double> &gy, V3<double> &vz, V3<double> &vx,
                             V3<double> &vy) {
  V3<double> o = ld3d(xyz, i);
  V3<double> u = ld3d(xyz, iz) - o;
  vz = u;
  vx = V3<double>{0, 0, 0};
  vy = V3<double>{0, 0, 0};
  double ru = sqrt(dot(u, u));
  V3<double> v, w;
  if (fr != NB_FRAME_ZONLY) {
    v = ld3d(xyz, ix) - o;
    vx = v;
  } else {
    v = V3<double>{1.0, 0.0, 0.0};
    if (fabs(u.x / ru) > 0.866) v = V3<double>{0.0, 1.0, 0.0};
  }
  double rv = sqrt(dot(v, v));
  if (fr == NB_FRAME_ZTHENBISECTOR || fr == NB_FRAME_THREEFOLD) {
    w = ld3d(xyz, iy) - o;
    vy = w;
  } else {
    w = cross(u, v);
  }
  double rw = sqrt(dot(w, w));
  u = (1.0 / ru) * u; v = (1.0 / rv) * v; w = (1.0 / rw) * w;
  V3<double> uv = unit_d(cross(v, u)), uw = unit_d(cross(w, u)), vw = unit_d(cross(w, v));
  double cuv = dot(u, v), suv = sqrt(1.0 - cuv * cuv);
  double dpu = -dot(trq, u), dpv = -dot(trq, v), dpw = -dot(trq, w);
  gz = gx = gy = V3<double>{0, 0, 0};
  if (fr == NB_FRAME_ZONLY) {
    gz = (dpv / (ru * suv)) * uv + (dpw / ru) * uw;
  } else if (fr == NB_FRAME_ZTHENX) {
    gz = (dpv / (ru * suv)) * uv + (dpw / ru) * uw;
    gx = (-dpu / (rv * suv)) * uv;
  } else if (fr == NB_FRAME_BISECTOR) {
    gz = (dpv / (ru * suv)) * uv + (0.5 * dpw / ru) * uw;
    gx = (-dpu / (rv * suv)) * uv + (0.5 * dpw / rv) * vw;
  } else if (fr == NB_FRAME_ZTHENBISECTOR) {
    V3<double> r = v + w;
    V3<double> sv = unit_d(cross(u, r));
    r = unit_d(r);
    V3<double> ur = unit_d(cross(r, u)), us = unit_d(cross(sv, u));
    double cur = dot(u, r), sur = sqrt(1.0 - cur * cur);
    double cvs = dot(v, sv), svs = sqrt(1.0 - cvs * cvs);
    double cws = dot(w, sv), sws = sqrt(1.0 - cws * cws);
    V3<double> t1 = unit_d(v - cvs * sv), t2 = unit_d(w - cws * sv);
    double ut1c = dot(u, t1), ut1s = sqrt(1.0 - ut1c * ut1c);
    double ut2c = dot(u, t2), ut2s = sqrt(1.0 - ut2c * ut2c);
    double dpr = -dot(trq, r), dps = -dot(trq, sv);
    gz = (dpr / (ru * sur)) * ur + (dps / ru) * us;
    gx = (dpu / (rv * (ut1s + ut2s))) * (svs * sv - cvs * t1);
    gy = (dpu / (rw * (ut1s + ut2s))) * (sws * sv - cws * t2);
  } else if (fr == NB_FRAME_THREEFOLD) {
    V3<double> sum = u + v + w;
    double rs = sqrt(dot(sum, sum));
    sum = (1.0 / rs) * sum;
    V3<double> ax[3] = {u, v, w};
    double rr3[3] = {ru, rv, rw};
    V3<double> out[3];
    for (int a = 0; a < 3; a++) {
      V3<double> e = ax[a];
      V3<double> r = unit_d(ax[(a + 1) % 3] + ax[(a + 2) % 3]);
      double c = dot(r, e), sn = sqrt(1.0 - c * c);
      double dpr = -dot(trq, r);
      V3<double> del = unit_d(cross(r, e));
      double dpd = -dot(trq, del);
      V3<double> eps = cross(del, e);
      out[a] = (dpr / (rr3[a] * sn)) * del + (dpd * dot(e, sum) / (rr3[a] * rs)) * eps;
    }
    gz = out[0]; gx = out[1]; gy = out[2];
  }
}

