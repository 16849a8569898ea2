// nb_pair.h -- real-space pair terms shared by the per-system pair kernels
// and the batched small-cell kernel: permanent field, permanent and
// polarization energy / force / torque of one (i, k) pair, i-side only (full
// neighbour rows visit every pair from both atoms).
#pragma once
#include "nb_kern.h"

// one permanent-field pair in arithmetic U, accumulated in FP64
template <typename U>
__device__ __forceinline__ void perm_field_pair(const double4 &pi, const AtomMP &ai, const double4 &pk,
                                                const AtomMP &ak, int code, const CellD &cell, double aewald,
                                                const ScaleTab &sc, V3<double> &fd, V3<double> &fp) {
  double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
  double r2 = image(cell, dx, dy, dz);
  MP<U> K = load_mp<U>(ak);
  U r, rinv;
  r_rinv(r2, r, rinv);
  V3<U> R{(U)-dx, (U)-dy, (U)-dz};  // from k to i
  U bn[4];
  ewald_bn<U, 3>(r, rinv, (U)aewald, bn);
  Thole<U> th = thole<U, false>(r, R, (U)ai.ipd, (U)ak.ipd, (U)ai.thole, (U)ak.thole);
  U rr3 = rinv * rinv * rinv, rr5 = (U)3 * rr3 * rinv * rinv, rr7 = (U)5 * rr5 * rinv * rinv;
  U ds = (U)sc.d[code], ps = (U)sc.p[code];
  acc3(fd, mp_field(R, K, bn[1] - ((U)1 - th.s3 * ds) * rr3, bn[2] - ((U)1 - th.s5 * ds) * rr5,
                    bn[3] - ((U)1 - th.s7 * ds) * rr7));
  acc3(fp, mp_field(R, K, bn[1] - ((U)1 - th.s3 * ps) * rr3, bn[2] - ((U)1 - th.s5 * ps) * rr5,
                    bn[3] - ((U)1 - th.s7 * ps) * rr7));
}

// Strain derivative of one full-row visit: with r = x_k - x_i and f the i-side
// gradient, dE/dr = -f and the pair term (dE/dr)_a r_b is shared by the two
// visits of the pair.
template <typename U>
__device__ __forceinline__ void pair_strain(double *w9, const V3<U> &f, const V3<U> &r) {
  double fa[3] = {(double)f.x, (double)f.y, (double)f.z}, rb[3] = {(double)r.x, (double)r.y, (double)r.z};
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) w9[3 * a + b] -= 0.5 * fa[a] * rb[b];
}

// Real-space energy, force and torque, split into a permanent kernel and a
// polarization kernel (the combined pair function needs more than 128 VGPRs
// and ran one wavefront per SIMD).  Both add into the same sorted gradient and
// torque arrays, one after the other on the handle's stream.
template <typename U>
__device__ __forceinline__ void perm_energy_pair(const double4 &pi, const AtomMP &ai, const double4 &pk,
                                                 const AtomMP &ak, int code, const CellD &cell, double aewald,
                                                 const ScaleTab &sc, double &ep, long long &cnt, V3<double> &F,
                                                 V3<double> &Tq, double *w9) {
  double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
  double r2 = image(cell, dx, dy, dz);
  MP<U> I = load_mp<U>(ai), K = load_mp<U>(ak);
  V3<U> r{(U)dx, (U)dy, (U)dz};
  U rl, rinv;
  r_rinv(r2, rl, rinv);
  U bn[6], rr[6];
  ewald_bn<U, 5>(rl, rinv, (U)aewald, bn);
  PermGeo<U> g;
  perm_geometry(r, I, K, g);
  V3<U> f{0, 0, 0}, tq{0, 0, 0};
  U epair = perm_terms<U, true>(r, I, K, g, bn, (U)1, f, tq);
  U s1 = (U)1 - (U)sc.m[code];
  if (s1 != (U)0) {
    coulomb_rr<U, 5>(rinv, rr);
    epair -= s1 * perm_terms<U, true>(r, I, K, g, rr, -s1, f, tq);
  }
  if (epair != (U)0) cnt++;
  ep += (double)epair;
  acc3(F, f);
  acc3(Tq, tq);
  if (w9) pair_strain(w9, f, r);
}

template <typename U>
__device__ __forceinline__ void pol_energy_pair(const double4 &pi, const AtomMP &ai, const double *u, const double *up,
                                                int s, const double4 &pk, const AtomMP &ak, int k, int code,
                                                const CellD &cell, double aewald, const ScaleTab &sc, int direct,
                                                double &ei, V3<double> &F, V3<double> &Tq, double *w9) {
  double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
  double r2 = image(cell, dx, dy, dz);
  MP<U> I = load_mp<U>(ai), K = load_mp<U>(ak);
  V3<U> r{(U)dx, (U)dy, (U)dz};
  U rl, rinv;
  r_rinv(r2, rl, rinv);
  U bn[5], rr[5];
  ewald_bn<U, 4>(rl, rinv, (U)aewald, bn);
  coulomb_rr<U, 4>(rinv, rr);
  V3<U> f{0, 0, 0}, tq{0, 0, 0};
  Thole<U> th = thole<U, true>(rl, r, (U)ai.ipd, (U)ak.ipd, (U)ai.thole, (U)ak.thole);
  ei += (double)pol_terms<U, true>(r, I, K, ld3t<U>(u, s), ld3t<U>(up, s), ld3t<U>(u, k), ld3t<U>(up, k), bn, rr, th,
                                   (U)sc.p[code], (U)sc.d[code], direct != 0, f, tq);
  acc3(F, f);
  acc3(Tq, tq);
  if (w9) pair_strain(w9, f, r);
}

// Buffered 14-7 pair (VanDerWaalsForm, Halgren's gamma = gg and delta = d1 with
// g1 = 1 + gg, t1n = (1 + d1)^7) with the softcore shift sc1 and scale sc2,
// and the MultiplicativeSwitch taper (f(cut) = 1, f(off) = 0, zero first and
// second derivatives at both ends) in the normalised variable x = (r - cut)
// iw: f = 1 - 10 x^3 + 15 x^4 - 6 x^5, the polynomial of
// MultiplicativeSwitch.java:74-96 without its monomial cancellation (which
// FP32 cannot absorb).  ds is the radial factor of the i-side gradient
// (gradient = ds (red_i - red_k)); the intermediates serve the lambda terms.
template <typename T> struct Vdw147 { T rr, eik, tp, dtp, t1, t2, t2a, t1d, t2d, dt1, dt2, ds; };

template <typename T>
__device__ __forceinline__ Vdw147<T> vdw147(double r2, T irv, T ev, T sc1, T sc2, T d1, T g1, T gg, T t1n, T wcut,
                                            T iw, bool tapered) {
  Vdw147<T> v;
  T rinv;
  r_rinv(r2, v.rr, rinv);
  T rho = v.rr * irv;
  T rho2 = rho * rho, rho6 = rho2 * rho2 * rho2, rho7 = rho6 * rho;
  T rd = rho + d1, rd2 = rd * rd, rd6 = rd2 * rd2 * rd2, rd7 = rd6 * rd;
  v.t1d = nb_rcp(sc1 + rd7);
  v.t2d = nb_rcp(sc1 + rho7 + gg);
  v.t1 = t1n * v.t1d;
  v.t2a = g1 * v.t2d;
  v.t2 = v.t2a - (T)2;
  v.eik = sc2 * ev * v.t1 * v.t2;
  v.tp = 1;
  v.dtp = 0;
  if (tapered) {
    T x = (v.rr - wcut) * iw, x2 = x * x;
    v.tp = (T)1 + x2 * x * ((T)-10 + x * ((T)15 - (T)6 * x));
    v.dtp = x2 * ((T)-30 + x * ((T)60 - (T)30 * x)) * iw;
  }
  v.dt1 = v.t1 * (T)7 * rd6 * irv * v.t1d;
  v.dt2 = v.t2a * (T)7 * rho6 * irv * v.t2d;
  T dedr = -sc2 * ev * (v.dt1 * v.t2 + v.t1 * v.dt2);
  v.ds = (v.eik * v.dtp + dedr * v.tp) * rinv;
  return v;
}

// Field at i of the two dipole vectors of record k (InducedDipoleFieldRegion.
// java:365-450): Ewald-screened minus the Thole-damped bare part, or (PRE) the
// Thole-damped bare Coulomb field of the preconditioner (PCGSolver.java:
// 1270-1390).  e is the row entry (lattice translation of an image sum).
template <typename T, bool PRE>
__device__ __forceinline__ void dip_field_pair(const DipRec<T> &ri, const DipRec<T> &rk, uint32_t e,
                                               const CellD &cell, double aewald, V3<double> &a, V3<double> &b) {
  double dx = rk.x - ri.x, dy = rk.y - ri.y, dz = rk.z - ri.z;
  entry_shift(cell, e, dx, dy, dz);
  double r2 = image(cell, dx, dy, dz);
  T r, rinv;
  r_rinv(r2, r, rinv);
  V3<T> R{(T)dx, (T)dy, (T)dz};
  Thole<T> th = thole<T, false>(r, R, ri.ipd, rk.ipd, ri.thole, rk.thole);
  T rr3 = rinv * rinv * rinv, rr5 = (T)3 * rr3 * rinv * rinv;
  T B1, B2;
  if (PRE) {
    B1 = th.s3 * rr3;
    B2 = th.s5 * rr5;
  } else {
    T bn[3];
    ewald_bn<T, 2>(r, rinv, (T)aewald, bn);
    B1 = bn[1] - ((T)1 - th.s3) * rr3;
    B2 = bn[2] - ((T)1 - th.s5) * rr5;
  }
  acc3(a, dip_field(R, V3<T>{rk.u[0], rk.u[1], rk.u[2]}, B1, B2));
  acc3(b, dip_field(R, V3<T>{rk.v[0], rk.v[1], rk.v[2]}, B1, B2));
}
