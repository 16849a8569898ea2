// nb_pairs.hip -- real-space pair kernels (gfx950).
//
// LPA consecutive lanes of a wavefront share one (spatially sorted) atom and
// walk its neighbor row in stride LPA, so a wavefront reads LPA consecutive
// list entries per atom and the chip sees n * LPA threads.  Per-lane partial
// sums are FP64 (also in the mixed form: the pair arithmetic is FP32, the
// accumulation is not) and are combined with cross-lane shuffles; each atom's
// field / force / torque is written once, by lane 0, without atomics.
//
//  * k_filter: PermanentFieldRegion's side effect of building the Ewald
//    (r <= off) and preconditioner (r < 4.5 A) neighbor lists
//    (PermanentFieldRegion.java:679-714), as an ordered ballot compaction.
//  * k_perm_field: permanent real-space field and fieldCR
//    (PermanentFieldRegion.java:704-808).
//  * k_dip_field: induced-dipole real-space field (InducedDipoleFieldRegion.
//    java:365-450) and the preconditioner field (PCGSolver.java:1270-1390).
//  * k_real_energy: permanent and polarization energy, force and torque
//    (RealSpaceEnergyRegion.java:803-2177).
//  * k_vdw: buffered 14-7 (VanDerWaals.java:1638-1900).
#include <cmath>
#include <stdexcept>

#include "nb_kern.h"
#include "nb_pair.h"

namespace {

template <int LPA>
__device__ __forceinline__ unsigned long long group_mask(unsigned long long ballot) {
  int base = (threadIdx.x & 63) & ~(LPA - 1);
  unsigned long long m = LPA == 64 ? ballot : (ballot >> base) & ((1ull << LPA) - 1ull);
  return m;
}

template <int LPA>
__global__ void __launch_bounds__(256) k_filter(int n, const double4 *__restrict__ pos, const uint32_t *__restrict__ nbr,
                                                const int32_t *__restrict__ nnb, int maxnb, CellD cell, double off2,
                                                double pcut, ScaleTab sc, uint32_t *__restrict__ ew, int maxew,
                                                int32_t *__restrict__ new_, uint32_t *__restrict__ pc, int maxpc,
                                                int32_t *__restrict__ npc, uint32_t *__restrict__ sp, int maxsp,
                                                int32_t *__restrict__ nsp, int32_t *__restrict__ flags) {
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt / LPA, lane = gt % LPA;
  bool valid = s < n;
  int cnt = valid ? nnb[s] : 0;
  double4 pi = valid ? pos[s] : make_double4(0, 0, 0, 0);
  int ne = 0, np = 0, ns = 0;
  const uint32_t *row = nbr + (size_t)(valid ? s : 0) * maxnb;
  uint32_t *erow = ew + (size_t)(valid ? s : 0) * maxew;
  uint32_t *prow = pc + (size_t)(valid ? s : 0) * maxpc;
  uint32_t *srow = sp + (size_t)(valid ? s : 0) * maxsp;
  const double pc2 = pcut * pcut, plo = pc2 * (1.0 - 1e-12), phi = pc2 * (1.0 + 1e-12);
  // the entry's class: kept by the Ewald cutoff, the preconditioner cutoff
  // (r < pcut, by sqrt only within rounding of the boundary), masked
  auto classify = [&](uint32_t e, double4 pk, bool &ke, bool &kp, bool &ks) {
    entry_shift(cell, e, pk.x, pk.y, pk.z);
    double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
    double r2 = image(cell, dx, dy, dz);
    ke = r2 <= off2;
    kp = ke && (r2 < plo || (r2 <= phi && sqrt(r2) < pcut));
    ks = ke && sc.masked[e >> NB_CODE_SHIFT];
  };
  for (int base = 0; base < cnt; base += 2 * LPA) {
    // two chunks of the row per trip: both row loads and both position
    // gathers are issued before either is used
    const int t0 = base + lane, t1 = base + LPA + lane;
    const uint32_t e0 = t0 < cnt ? row[t0] : 0u, e1 = t1 < cnt ? row[t1] : 0u;
    const double4 p0 = t0 < cnt ? pos[e0 & cell.kmask] : pi, p1 = t1 < cnt ? pos[e1 & cell.kmask] : pi;
    for (int half = 0; half < 2; half++) {
      if (half == 1 && base + LPA >= cnt) break;   // uniform across the atom's lanes
      const int t = half ? t1 : t0;
      const uint32_t e = half ? e1 : e0;
      bool ke = false, kp = false, ks = false;
      if (t < cnt) classify(e, half ? p1 : p0, ke, kp, ks);
      unsigned long long me = group_mask<LPA>(__ballot(ke));
      unsigned long long mp = group_mask<LPA>(__ballot(kp));
      unsigned long long ms = group_mask<LPA>(__ballot(ks));
      unsigned long long below = (1ull << lane) - 1ull;
      if (ke) {
        int q = ne + __popcll(me & below);
        if (q < maxew) erow[q] = e;
      }
      if (kp) {
        int q = np + __popcll(mp & below);
        if (q < maxpc) prow[q] = e;
      }
      if (ks) {
        int q = ns + __popcll(ms & below);
        if (q < maxsp) srow[q] = e;
      }
      ne += __popcll(me);
      np += __popcll(mp);
      ns += __popcll(ms);
    }
  }
  if (valid && lane == 0) {
    // counts are clamped to the row capacity (rows are never read past it);
    // the flags carry the true maxima so the caller can grow the rows and redo
    new_[s] = min(ne, maxew);
    npc[s] = min(np, maxpc);
    nsp[s] = min(ns, maxsp);
    if (ne > maxew) atomicMax(&flags[1], ne);
    if (np > maxpc) atomicMax(&flags[2], np);
    if (ns > maxsp) atomicMax(&flags[5], ns);
  }
}


// LIST 0: the whole Ewald row (FP64 form); LIST 1: the row without the masked
// pairs, whose screened-minus-bare radial differences cancel catastrophically
// at bonded distances (mixed form, T = float); LIST 2: only the masked row in
// FP64, added to the fields LIST 1 wrote (mixed form)
template <typename T, int LPA, int LIST>
__global__ void __launch_bounds__(256) k_perm_field(int n, const double4 *__restrict__ pos, const AtomMP *__restrict__ mp,
                                                    const uint32_t *__restrict__ lst, const int32_t *__restrict__ lcnt,
                                                    int maxl, CellD cell, double aewald, ScaleTab sc,
                                                    double *__restrict__ field, double *__restrict__ fieldp) {
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt / LPA, lane = gt % LPA;
  bool valid = s < n;
  V3<double> fd{0, 0, 0}, fp{0, 0, 0};
  if (valid) {
    double4 pi = pos[s];
    AtomMP ai = mp[s];
    int m = lcnt[s];
    const uint32_t *row = lst + (size_t)s * maxl;
    for (int t = lane; t < m; t += LPA) {
      uint32_t e = row[t];
      int k = e & cell.kmask, code = e >> NB_CODE_SHIFT;
      if (LIST == 1 && sc.masked[code]) continue;
      double4 pk = pos[k];
      entry_shift(cell, e, pk.x, pk.y, pk.z);
      perm_field_pair<T>(pi, ai, pk, mp[k], code, cell, aewald, sc, fd, fp);
    }
  }
  group_sum3<LPA>(fd);
  group_sum3<LPA>(fp);
  if (valid && lane == 0) {
    if (LIST == 2) {
      add3(field, s, fd);
      add3(fieldp, s, fp);
    } else {
      st3(field, s, fd);
      st3(fieldp, s, fp);
    }
  }
}

// real-space field of the packed dipoles over one list: Ewald-screened over
// the Ewald rows, or (PRE) bare Coulomb + Thole over the preconditioner rows
template <typename T, int LPA, bool PRE>
__global__ void __launch_bounds__(256) k_dip_field(int n, const DipRec<T> *__restrict__ rec,
                                                   const uint32_t *__restrict__ lst, const int32_t *__restrict__ cnt,
                                                   int maxl, CellD cell, double aewald, double *__restrict__ f,
                                                   double *__restrict__ fp, const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt / LPA, lane = gt % LPA;
  bool valid = s < n;
  V3<double> a{0, 0, 0}, b{0, 0, 0};
  if (valid) {
    const DipRec<T> ri = rec[s];
    int m = cnt[s];
    const uint32_t *row = lst + (size_t)s * maxl;
    if constexpr (sizeof(T) == 4) {
      // two entries per trip: both row loads, then both record gathers, are in
      // flight together, which halves the dependent memory round trips of the
      // latency-bound loop (71 VGPRs, 7 waves per SIMD; the FP64 form would
      // drop to 3 waves and keeps one entry per trip)
      for (int t = lane; t < m; t += 2 * LPA) {
        const bool two = t + LPA < m;
        const uint32_t e0 = row[t], e1 = two ? row[t + LPA] : e0;
        const DipRec<T> r0 = rec[e0 & cell.kmask], r1 = rec[e1 & cell.kmask];
        dip_field_pair<T, PRE>(ri, r0, e0, cell, aewald, a, b);
        if (two) dip_field_pair<T, PRE>(ri, r1, e1, cell, aewald, a, b);
      }
    } else {
      for (int t = lane; t < m; t += LPA) {
        uint32_t e = row[t];
        dip_field_pair<T, PRE>(ri, rec[e & cell.kmask], e, cell, aewald, a, b);
      }
    }
  }
  group_sum3<LPA>(a);
  group_sum3<LPA>(b);
  if (valid && lane == 0) {
    st3(f, s, a);
    st3(fp, s, b);
  }
}


// The preconditioner field of polar * r (records recA) fused with the
// preconditioned residual z = polar (E_pre + scale r) (PCGSolver.java:
// 1183-1571, the z of UpdateConjugate): lane 0 of each atom writes z, its
// record in recB (the dipoles whose Ewald field comes next) and the partial
// r.z into S_RZN / S_RZPN.
template <typename T, int LPA>
__global__ void __launch_bounds__(256) k_precond_z(int n, const DipRec<T> *__restrict__ rec,
                                                   const uint32_t *__restrict__ lst, const int32_t *__restrict__ cnt,
                                                   int maxl, CellD cell, const double *__restrict__ polar, double psc,
                                                   const double *__restrict__ r, const double *__restrict__ rp,
                                                   double *__restrict__ z, double *__restrict__ zp, void *recb,
                                                   int mixed, double *__restrict__ part, int stride, int neps,
                                                   double *__restrict__ scal, int control, double polar_eps,
                                                   int maxit, const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt / LPA, lane = gt % LPA;
  bool valid = s < n;
  V3<double> a{0, 0, 0}, b{0, 0, 0};
  if (valid) {
    const DipRec<T> ri = rec[s];
    int m = cnt[s];
    const uint32_t *row = lst + (size_t)s * maxl;
    if constexpr (sizeof(T) == 4) {
      for (int t = lane; t < m; t += 2 * LPA) {
        const bool two = t + LPA < m;
        const uint32_t e0 = row[t], e1 = two ? row[t + LPA] : e0;
        const DipRec<T> r0 = rec[e0 & cell.kmask], r1 = rec[e1 & cell.kmask];
        dip_field_pair<T, true>(ri, r0, e0, cell, 0.0, a, b);
        if (two) dip_field_pair<T, true>(ri, r1, e1, cell, 0.0, a, b);
      }
    } else {
      for (int t = lane; t < m; t += LPA) {
        uint32_t e = row[t];
        dip_field_pair<T, true>(ri, rec[e & cell.kmask], e, cell, 0.0, a, b);
      }
    }
  }
  group_sum3<LPA>(a);
  group_sum3<LPA>(b);
  double v[2] = {0.0, 0.0};
  if (valid && lane == 0) {
    const double al = polar[s];
    const double fa[3] = {a.x, a.y, a.z}, fb[3] = {b.x, b.y, b.z};
    for (int d = 0; d < 3; d++) {
      const size_t q = 3 * (size_t)s + d;
      const double zz = al * (fa[d] + psc * r[q]), zzp = al * (fb[d] + psc * rp[q]);
      z[q] = zz;
      zp[q] = zzp;
      dip_put(recb, mixed, s, d, zz, zzp);
      v[0] += r[q] * zz;
      v[1] += rp[q] * zzp;
    }
  }
  const int sl[2] = {S_RZN, S_RZPN};
  block_partials<2>(v, part, sl, stride);
  // block 0 reduces the residual norms of the update before it (neps partial
  // sums) and, for a recorded solve, applies the convergence test; the solve
  // stops after this kernel, whose z is then not used
  if (blockIdx.x == 0 && neps > 0) {
    const int in[2] = {S_EPS, S_EPSP};
    double t[2];
    block_totals<2>(part, in, stride, neps, t);
    if (threadIdx.x == 0) {
      scal[S_EPS] = t[0];
      scal[S_EPSP] = t[1];
      if (control) pcg_control(scal, t, n, polar_eps, maxit);
    }
  }
}

// Permanent pair under an OST phase.  A soft pair sees r_eff = sqrt(r^2 +
// lalpha) in the radial functions and the unshifted r in the angular ones, and
// is scaled by l2.  Its lambda derivatives use d bn_j / d(r^2) = -bn_{j+1} / 2
// with d lalpha / dlambda = -2 dlalpha: the radial set shifted by one index is
// dE/dlambda / dlalpha, by two the second derivative (RealSpaceEnergyRegion.java
// :1477-1690); the force of the once-shifted set is its coordinate derivative.
template <typename U>
__device__ __forceinline__ void perm_energy_pair_ost(const double4 &pi, const AtomMP &ai, const double4 &pk,
                                                     const AtomMP &ak, int code, bool soft, const CellD &cell,
                                                     double aewald, const ScaleTab &sc, const ElecPhase &ph, double &ep,
                                                     long long &cnt, V3<double> &F, V3<double> &Tq, double *w9,
                                                     double &dl1, double &dl2, V3<double> &FL, V3<double> &TL) {
  double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
  double r2 = image(cell, dx, dy, dz);
  MP<U> I = load_mp<U>(ai), K = load_mp<U>(ak);
  V3<U> r{(U)dx, (U)dy, (U)dz};
  U rl, rinv;
  r_rinv(soft ? r2 + ph.lalpha : r2, rl, rinv);
  U bn[7], rr[7];
  ewald_bn<U, 6>(rl, rinv, (U)aewald, bn);
  coulomb_rr<U, 6>(rinv, rr);
  PermGeo<U> g;
  perm_geometry(r, I, K, g);
  V3<U> f{0, 0, 0}, tq{0, 0, 0};
  U epair = perm_terms<U, true>(r, I, K, g, bn, (U)1, f, tq);
  U s1 = (U)1 - (U)sc.m[code];
  if (s1 != (U)0) epair -= s1 * perm_terms<U, true>(r, I, K, g, rr, -s1, f, tq);
  U l2 = soft ? (U)ph.l2 : (U)1;
  if (epair != (U)0) cnt++;
  ep += (double)(l2 * epair);
  acc3(F, l2 * f);
  acc3(Tq, l2 * tq);
  if (w9) pair_strain(w9, l2 * f, r);
  if (soft) {
    V3<U> fr{0, 0, 0}, tr{0, 0, 0}, dummy{0, 0, 0};
    U dR = perm_terms<U, true>(r, I, K, g, bn + 1, (U)1, fr, tr);
    U d2R = perm_terms<U, false>(r, I, K, g, bn + 2, (U)1, dummy, dummy);
    if (s1 != (U)0) {
      dR -= s1 * perm_terms<U, true>(r, I, K, g, rr + 1, -s1, fr, tr);
      d2R -= s1 * perm_terms<U, false>(r, I, K, g, rr + 2, (U)1, dummy, dummy);
    }
    double e = (double)epair, d1 = (double)dR, d2 = (double)d2R, l2d = ph.l2;
    dl1 += ph.dperm * e + l2d * ph.dlalpha * d1;
    dl2 += ph.d2perm * e + 2.0 * ph.dperm * ph.dlalpha * d1 + l2d * ph.d2lalpha * d1 +
           l2d * ph.dlalpha * ph.dlalpha * d2;
    U ca = (U)ph.dperm, cb = (U)(l2d * ph.dlalpha);
    acc3(FL, ca * f + cb * fr);
    acc3(TL, ca * tq + cb * tr);
  }
}

// Polarization pair under an OST phase: unshifted, scaled by pol; its lambda
// terms are dpol / d2pol times the pair.
template <typename U>
__device__ __forceinline__ void pol_energy_pair_ost(const double4 &pi, const AtomMP &ai, const double *u,
                                                    const double *up, int s, const double4 &pk, const AtomMP &ak,
                                                    int k, int code, const CellD &cell, double aewald,
                                                    const ScaleTab &sc, int direct, const ElecPhase &ph, double &ei,
                                                    V3<double> &F, V3<double> &Tq, double *w9, double &dl1,
                                                    double &dl2, V3<double> &FL, V3<double> &TL) {
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
  double e = (double)pol_terms<U, true>(r, I, K, ld3t<U>(u, s), ld3t<U>(up, s), ld3t<U>(u, k), ld3t<U>(up, k), bn, rr,
                                        th, (U)sc.p[code], (U)sc.d[code], direct != 0, f, tq);
  U w = (U)ph.pol, dw = (U)ph.dpol;
  ei += ph.pol * e;
  dl1 += ph.dpol * e;
  dl2 += ph.d2pol * e;
  acc3(F, w * f);
  acc3(Tq, w * tq);
  acc3(FL, dw * f);
  acc3(TL, dw * tq);
  if (w9) pair_strain(w9, w * f, r);
}

// POL = false: permanent energy (slot E_PERM_REAL, pair count); POL = true:
// polarization energy (slot E_POL_REAL).  LIST 0: the whole Ewald row (FP64
// form); LIST 1: the Ewald row without the masked pairs (mixed form, T =
// float); LIST 2: only the masked row, in FP64 (mixed form).
// VIR: also the pair strain derivative, 9 partial sums per block into vpart.
// OST: the phase scales of ph, with dU/dL, d2U/dL2 (slots E_DUDL, E_D2UDL2)
// and the lambda gradient and torque added into lgrad / ltorque; soft (sorted,
// may be null) marks the alchemical atoms and mol their molecules.
template <typename T, int LPA, bool POL, int LIST, bool VIR, bool OST>
__global__ void __launch_bounds__(256) k_real_energy(int n, const double4 *__restrict__ pos, const AtomMP *__restrict__ mp,
                                                     const uint32_t *__restrict__ lst, const int32_t *__restrict__ lcnt,
                                                     int maxl, CellD cell, double aewald, ScaleTab sc, int direct,
                                                     const double *__restrict__ u, const double *__restrict__ up,
                                                     double el, double *__restrict__ grad, double *__restrict__ torque,
                                                     double *__restrict__ energy, int64_t *__restrict__ count,
                                                     double *__restrict__ vpart, ElecPhase ph,
                                                     const int32_t *__restrict__ soft, const int32_t *__restrict__ mol,
                                                     double *__restrict__ lgrad, double *__restrict__ ltorque) {
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt / LPA, lane = gt % LPA;
  bool valid = s < n;
  double e = 0.0, dl1 = 0.0, dl2 = 0.0;
  long long cnt = 0;
  V3<double> F{0, 0, 0}, Tq{0, 0, 0}, FL{0, 0, 0}, TL{0, 0, 0};
  double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double *wp = VIR ? w9 : nullptr;
  if (valid) {
    double4 pi = pos[s];
    AtomMP ai = mp[s];
    int m = lcnt[s];
    int si = (OST && soft) ? soft[s] : 0, moli = (OST && soft) ? mol[s] : 0;
    const uint32_t *row = lst + (size_t)s * maxl;
    for (int t = lane; t < m; t += LPA) {
      uint32_t en = row[t];
      int k = en & cell.kmask, code = en >> NB_CODE_SHIFT;
      if (LIST == 1 && sc.masked[code]) continue;
      double4 pk = pos[k];
      entry_shift(cell, en, pk.x, pk.y, pk.z);
      if (OST) {
        int sk = soft ? soft[k] : 0;
        if (ph.vapor && si && sk) {
          // vapor phase: both-soft pairs that the condensed phase softens are
          // left out (RealSpaceEnergyRegion.java:864-870)
          bool same = moli == mol[k];
          if ((ph.inter && !same) || (ph.intra && same)) continue;
        }
        bool sp = ph.soft_pairs && (si || sk);
        if (POL)
          pol_energy_pair_ost<T>(pi, ai, u, up, s, pk, mp[k], k, code, cell, aewald, sc, direct, ph, e, F, Tq, wp,
                                 dl1, dl2, FL, TL);
        else
          perm_energy_pair_ost<T>(pi, ai, pk, mp[k], code, sp, cell, aewald, sc, ph, e, cnt, F, Tq, wp, dl1, dl2,
                                  FL, TL);
      } else {
        if (POL) pol_energy_pair<T>(pi, ai, u, up, s, pk, mp[k], k, code, cell, aewald, sc, direct, e, F, Tq, wp);
        else perm_energy_pair<T>(pi, ai, pk, mp[k], code, cell, aewald, sc, e, cnt, F, Tq, wp);
      }
    }
  }
  if (OST) {
    group_sum3<LPA>(FL);
    group_sum3<LPA>(TL);
    if (valid && lane == 0) {
      add3(lgrad, s, el * FL);
      add3(ltorque, s, el * TL);
    }
    block_add(0.5 * el * dl1, energy + E_DUDL);
    block_add(0.5 * el * dl2, energy + E_D2UDL2);
  }
  if (VIR) {
#pragma unroll
    for (int q = 0; q < 9; q++) w9[q] *= el;
    block_store9(w9, vpart);
  }
  group_sum3<LPA>(F);
  group_sum3<LPA>(Tq);
  if (valid && lane == 0) {
    add3(grad, s, el * F);
    add3(torque, s, el * Tq);
  }
  block_add(0.5 * el * e, energy + (POL ? E_POL_REAL : E_PERM_REAL));
  if (!POL) block_add_ll(cnt, count + 1);
}

// LAM: softcore pairs (exactly one atom soft, or both soft under the inter /
// intramolecular flags; VanDerWaals.java:1689-1737) with dU/dL, d2U/dL2 and the
// lambda gradient dU/dX/dL (VanDerWaals.java:1842-1881) on the reduced sites
template <typename T, int LPA, bool VIR, bool LAM>
__global__ void __launch_bounds__(256) k_vdw(int n, const double4 *__restrict__ red, const int32_t *__restrict__ vclass,
                                             const uint32_t *__restrict__ nbr, const int32_t *__restrict__ nnb, int maxnb,
                                             CellD cell, const double *__restrict__ vtab, int nclass, ScaleTab sc,
                                             double off, double cut, double dl, double gm, double *__restrict__ gred,
                                             double *__restrict__ energy, int64_t *__restrict__ count,
                                             double *__restrict__ vpart, const int32_t *__restrict__ soft,
                                             const int32_t *__restrict__ mol, Softcore sf,
                                             double *__restrict__ gred_l) {
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt / LPA, lane = gt % LPA;
  bool valid = s < n;
  double esum = 0.0, el1 = 0.0, el2 = 0.0;
  long long cnt = 0;
  V3<double> G{0, 0, 0}, GL{0, 0, 0};
  double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (valid) {
    double off2 = off * off, cut2 = cut * cut;
    // MultiplicativeSwitch (f(cut) = 1, f(off) = 0, zero first and second
    // derivatives at both ends) in the normalised variable x = (r - cut) /
    // (off - cut): f = 1 - 10 x^3 + 15 x^4 - 6 x^5.  Identical polynomial to
    // the monomial form of MultiplicativeSwitch.java:74-96, but without its
    // cancellation, which FP32 cannot absorb.
    T wcut = (T)cut, iw = (T)(1.0 / (off - cut));
    T d1 = (T)dl, g1 = (T)1 + (T)gm, gg = (T)gm;
    T t1n = (T)pow(1.0 + dl, 7.0);
    double4 pi = red[s];
    int ci = vclass[s];
    int m = nnb[s];
    int softi = LAM ? soft[s] : 0, moli = LAM ? mol[s] : 0;
    const uint32_t *row = nbr + (size_t)s * maxnb;
    // one pair; pk and ck are the partner's reduced site and vdW class
    auto pair = [&](uint32_t en, double4 pk, int ck) {
      int k = en & cell.kmask, code = en >> NB_CODE_SHIFT;
      double vs = sc.v[code];
      if (!(vs > 0.0)) return;
      entry_shift(cell, en, pk.x, pk.y, pk.z);
      double dx = pi.x - pk.x, dy = pi.y - pk.y, dz = pi.z - pk.z;
      double r2 = image(cell, dx, dy, dz);
      if (r2 > off2) return;
      const double *pe = vtab + 2 * (ci * nclass + ck);
      T irv = (T)pe[0];
      if (!(irv > (T)0)) return;
      cnt++;
      bool spair = false;
      if (LAM) {
        int sk = soft[k];
        spair = softi != sk || (softi && sk && ((sf.inter && moli != mol[k]) || (sf.intra && moli == mol[k])));
      }
      T sc1 = spair ? (T)sf.sc1 : (T)0, sc2 = spair ? (T)sf.sc2 : (T)1;
      T ev = (T)(vs * pe[1]);
      Vdw147<T> v = vdw147<T>(r2, irv, ev, sc1, sc2, d1, g1, gg, t1n, wcut, iw, r2 > cut2);
      T rr = v.rr, eik = v.eik, tp = v.tp, dtp = v.dtp, t1 = v.t1, t2 = v.t2, t2a = v.t2a;
      T t1d = v.t1d, t2d = v.t2d, dt1 = v.dt1, dt2 = v.dt2, ds = v.ds;
      esum += (double)(eik * tp);
      acc3(G, V3<T>{ds * (T)dx, ds * (T)dy, ds * (T)dz});
      if (LAM && spair) {
        T dsc1 = (T)sf.dsc1, dsc2 = (T)sf.dsc2, d2sc1 = (T)sf.d2sc1, d2sc2 = (T)sf.d2sc2;
        T lt1 = -t1 * t1d * dsc1, lt2 = -t2a * t2d * dsc1;
        T dedl = ev * (dsc2 * t1 * t2 + sc2 * lt1 * t2 + sc2 * t1 * lt2);
        el1 += (double)(dedl * tp);
        T t1d2 = -dsc1 * t1d * t1d, t2d2 = -dsc1 * t2d * t2d;
        T d2t1 = -lt1 * t1d * dsc1 - t1 * t1d * d2sc1 - t1 * t1d2 * dsc1;
        T d2t2 = -lt2 * t2d * dsc1 - t2a * t2d * d2sc1 - t2a * t2d2 * dsc1;
        T df1 = d2sc2 * t1 * t2 + dsc2 * lt1 * t2 + dsc2 * t1 * lt2;
        T df2 = dsc2 * lt1 * t2 + sc2 * d2t1 * t2 + sc2 * lt1 * lt2;
        T df3 = dsc2 * t1 * lt2 + sc2 * lt1 * lt2 + sc2 * t1 * d2t2;
        el2 += (double)(ev * (df1 + df2 + df3) * tp);
        T dedldr = ev * (-dsc2 * t2 * dt1 - dsc2 * t1 * dt2 + (T)2 * sc2 * t2 * dt1 * dsc1 * t1d +
                         (T)2 * sc2 * t1 * dt2 * dsc1 * t2d - sc2 * lt2 * dt1 - sc2 * lt1 * dt2);
        T dl2 = (dedl * dtp + dedldr * tp) / rr;
        acc3(GL, V3<T>{dl2 * (T)dx, dl2 * (T)dy, dl2 * (T)dz});
      }
      if (VIR) {
        // d = red_i - red_k and the i-side gradient ds d: half of ds d_a d_b per visit
        double dd[3] = {dx, dy, dz}, h = 0.5 * (double)ds;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int b = 0; b < 3; b++) w9[3 * a + b] += h * dd[a] * dd[b];
      }
    };
    if constexpr (sizeof(T) == 4 && !LAM) {
      // two entries per trip: the row loads, then the site and class gathers
      // of both partners, are in flight together (the loop is bound by these
      // dependent loads, not by arithmetic)
      for (int t = lane; t < m; t += 2 * LPA) {
        const bool two = t + LPA < m;
        const uint32_t e0 = row[t], e1 = two ? row[t + LPA] : e0;
        const int k0 = e0 & cell.kmask, k1 = e1 & cell.kmask;
        const double4 p0 = red[k0], p1 = red[k1];
        const int c0 = vclass[k0], c1 = vclass[k1];
        pair(e0, p0, c0);
        if (two) pair(e1, p1, c1);
      }
    } else {
      for (int t = lane; t < m; t += LPA) {
        uint32_t en = row[t];
        int k = en & cell.kmask;
        pair(en, red[k], vclass[k]);
      }
    }
  }
  group_sum3<LPA>(G);
  if (LAM) group_sum3<LPA>(GL);
  if (valid && lane == 0) {
    st3(gred, s, G);
    if (LAM) st3(gred_l, s, GL);
  }
  block_add(0.5 * esum, energy + E_VDW);
  block_add_ll(cnt, count);
  if (LAM) {
    block_add(0.5 * el1, energy + E_DUDL);
    block_add(0.5 * el2, energy + E_D2UDL2);
  }
  if (VIR) block_store9(w9, vpart);
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// Ewald / preconditioner / masked rows of this evaluation from the Verlet list
// (PermanentFieldRegion.java:679-714).  No host synchronisation: overflow is
// reported in flags[1], [2], [5] (cleared by k_zero_outputs at the start of
// the evaluation) and handled by nb_grow_filtered_rows after the evaluation.
void nb_filter_lists(NbHandle *h) {
  int n = h->n;
  ScaleTab sc = scale_tab(h);
  dim3 B(256), G = xcd_grid((long)n * NB_LPA, 256);
  double off2 = h->prm.ewald_cutoff * h->prm.ewald_cutoff;
  NB_LAUNCH(k_filter<NB_LPA>, G, B, 0, n, h->d_pos4, h->d_nbr, h->d_nnb, h->maxnb, h->cell, off2,
            h->prm.precond_cutoff, sc, h->d_ew, h->maxew, h->d_new, h->d_pc, h->maxpc, h->d_npc, h->d_sp, h->maxsp,
            h->d_nsp, h->d_flags);
}

// fl: the 8 flags read back after an evaluation.  Returns true (after growing
// the rows) when a filtered row overflowed, in which case the evaluation must
// be repeated.
bool nb_grow_filtered_rows(NbHandle *h, const int32_t *fl) {
  int n = h->n;
  bool grow = false;
  if (fl[1] > h->maxew) {
    hipFree(h->d_ew);
    h->d_ew = nullptr;
    h->maxew = fl[1] + fl[1] / 4 + 16;
    HIPCHECK(hipMalloc(&h->d_ew, sizeof(uint32_t) * (size_t)h->maxew * n));
    grow = true;
  }
  if (fl[2] > h->maxpc) {
    hipFree(h->d_pc);
    h->d_pc = nullptr;
    h->maxpc = fl[2] + fl[2] / 4 + 16;
    HIPCHECK(hipMalloc(&h->d_pc, sizeof(uint32_t) * (size_t)h->maxpc * n));
    grow = true;
  }
  if (fl[5] > h->maxsp) {
    hipFree(h->d_sp);
    h->d_sp = nullptr;
    h->maxsp = fl[5] + fl[5] / 4 + 16;
    HIPCHECK(hipMalloc(&h->d_sp, sizeof(uint32_t) * (size_t)h->maxsp * n));
    grow = true;
  }
  return grow;
}

// real-space permanent field and fieldCR over the filtered Ewald rows
void nb_perm_field(NbHandle *h) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  ScaleTab sc = scale_tab(h);
  dim3 B(256), G = xcd_grid((long)n * NB_LPA, 256);
  if (h->prm.precision == NB_PREC_MIXED) {
    NB_LAUNCH((k_perm_field<float, NB_LPA, 1>), G, B, 0, n, h->d_pos4, h->d_mp, h->d_ew, h->d_new, h->maxew, h->cell, aw,
              sc, h->d_field, h->d_fieldp);
    NB_LAUNCH((k_perm_field<double, 4, 2>), xcd_grid((long)n * 4, 256), B, 0, n, h->d_pos4, h->d_mp, h->d_sp, h->d_nsp,
              h->maxsp, h->cell, aw, sc, h->d_field, h->d_fieldp);
  } else {
    NB_LAUNCH((k_perm_field<double, NB_LPA, 0>), G, B, 0, n, h->d_pos4, h->d_mp, h->d_ew, h->d_new, h->maxew, h->cell, aw,
              sc, h->d_field, h->d_fieldp);
  }
}

// field of the dipoles currently held in the packed records (d_dip): the
// Ewald-screened field over the Ewald rows, or (precond) the bare Coulomb +
// Thole field over the preconditioner rows
void nb_dip_field_real(NbHandle *h, double *f, double *fp, bool precond) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = (use_recip && !precond) ? h->prm.ewald_alpha : 0.0;
  dim3 B(256), G = xcd_grid((long)n * NB_LPA_PCG, 256);
  const uint32_t *lst = precond ? h->d_pc : h->d_ew;
  const int32_t *cnt = precond ? h->d_npc : h->d_new;
  int maxl = precond ? h->maxpc : h->maxew;
  bool mixed = h->prm.precision == NB_PREC_MIXED;
  const DipRec<float> *rf = (const DipRec<float> *)h->d_dip;
  const DipRec<double> *rd = (const DipRec<double> *)h->d_dip;
  const double *ctl = h->ctl;
  if (mixed && precond)
    NB_LAUNCH((k_dip_field<float, NB_LPA_PCG, true>), G, B, 0, n, rf, lst, cnt, maxl, h->cell, aw, f, fp, ctl);
  else if (mixed)
    NB_LAUNCH((k_dip_field<float, NB_LPA_PCG, false>), G, B, 0, n, rf, lst, cnt, maxl, h->cell, aw, f, fp, ctl);
  else if (precond)
    NB_LAUNCH((k_dip_field<double, NB_LPA_PCG, true>), G, B, 0, n, rd, lst, cnt, maxl, h->cell, aw, f, fp, ctl);
  else
    NB_LAUNCH((k_dip_field<double, NB_LPA_PCG, false>), G, B, 0, n, rd, lst, cnt, maxl, h->cell, aw, f, fp, ctl);
}

// z of the PCG from the preconditioner field of the records d_dip2 (polar *
// r); z goes to d_z / d_zp and the records d_dip; neps > 0: also the
// residual norms of the update before (its partial sums), with the
// convergence test when control is set.  Returns the number of partial sums
// of r.z.
int nb_precond_z(NbHandle *h, int neps, bool control) {
  int n = h->n;
  dim3 B(256), G = xcd_grid((long)n * NB_LPA_PCG, 256);
  int mixed = h->prm.precision == NB_PREC_MIXED ? 1 : 0;
  double psc = h->prm.precond_scale;
  int maxit = h->prm.max_scf_iter > 0 ? h->prm.max_scf_iter : 1000;
  if (mixed)
    NB_LAUNCH((k_precond_z<float, NB_LPA_PCG>), G, B, 0, n, (const DipRec<float> *)h->d_dip2, h->d_pc, h->d_npc,
              h->maxpc, h->cell, h->d_polar, psc, h->d_r, h->d_rp, h->d_z, h->d_zp, (void *)h->d_dip, mixed,
              h->d_part, h->part_stride, neps, h->d_scal, (int)control, h->prm.polar_eps, maxit, h->ctl);
  else
    NB_LAUNCH((k_precond_z<double, NB_LPA_PCG>), G, B, 0, n, (const DipRec<double> *)h->d_dip2, h->d_pc, h->d_npc,
              h->maxpc, h->cell, h->d_polar, psc, h->d_r, h->d_rp, h->d_z, h->d_zp, (void *)h->d_dip, mixed,
              h->d_part, h->part_stride, neps, h->d_scal, (int)control, h->prm.polar_eps, maxit, h->ctl);
  return (int)G.x;
}

// FP64 Ewald real-space field of the FP64 records d_dip64 (the FP64 residual
// of the mixed form's refinement)
void nb_dip_field64(NbHandle *h, double *f, double *fp) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  dim3 B(256), G = xcd_grid((long)n * NB_LPA_PCG, 256);
  NB_LAUNCH((k_dip_field<double, NB_LPA_PCG, false>), G, B, 0, n, (const DipRec<double> *)h->d_dip64, h->d_ew, h->d_new,
            h->maxew, h->cell, aw, f, fp, h->ctl);
}

#define NB_LPA_SP 4

// which: NB_RE_ALL both passes into the real-space arrays (grad_s, torque_s);
// NB_RE_PERM the permanent pass (and, in the mixed form, the FP64 masked
// polarization pairs); NB_RE_POL the rest of the polarization pass, added
// into the reciprocal arrays (grad_r, torque_r, which k_recip_energy stores
// first on the same stream), so it can run beside the permanent pass
template <bool VIR, bool OST>
static void real_energy_t(NbHandle *h, bool polar, int which) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  ScaleTab sc = scale_tab(h);
  // with the finite-PCG gradient, uCR holds the adjoint and the mutual terms are replaced
  int direct = h->prm.polarization == NB_POL_DIRECT || h->finite_active;
  dim3 B(256), G = xcd_grid((long)n * NB_LPA, 256), Gs = xcd_grid((long)n * NB_LPA_SP, 256);
  double el = h->prm.electric;
  bool perm = (!OST || h->ph.perm_real) && which != NB_RE_POL;
  // the mixed form's FP64 pass over the masked polarization pairs runs with
  // the permanent passes (side stream, into the real-space arrays), which
  // balances the two streams: the FP32 polarization pass follows the
  // reciprocal energy on the main stream
  const bool masked_pol = polar && which != NB_RE_POL;
  polar = polar && which != NB_RE_PERM;
  const bool to_r = which == NB_RE_POL;
  double *g = to_r ? h->d_grad_r : h->d_grad_s, *tq = to_r ? h->d_torque_r : h->d_torque_s;
  double *lg = to_r ? h->d_lgrad_r : h->d_lgrad_s, *ltq = to_r ? h->d_ltorque_r : h->d_ltorque_s;
  auto vr = [&](dim3 gr) { return VIR ? nb_vir_region(h, gr.x) : (double *)nullptr; };
#define NB_RE(T, L, POL, LIST, GRID, ROWS, CNT, MAXR)                                                               \
  NB_LAUNCH((k_real_energy<T, L, POL, LIST, VIR, OST>), GRID, B, 0, n, h->d_pos4, h->d_mp, ROWS, CNT, MAXR, h->cell,  \
            aw, sc, direct, h->d_u, h->d_up, el, g, tq, h->d_energy, h->d_count, vr(GRID), h->ph, h->d_soft, h->d_mol, \
            lg, ltq)
  if (h->prm.precision == NB_PREC_MIXED) {
    if (perm) {
      NB_RE(float, NB_LPA, false, 1, G, h->d_ew, h->d_new, h->maxew);
      NB_RE(double, NB_LPA_SP, false, 2, Gs, h->d_sp, h->d_nsp, h->maxsp);
    }
    if (polar) NB_RE(float, NB_LPA, true, 1, G, h->d_ew, h->d_new, h->maxew);
    if (masked_pol) NB_RE(double, NB_LPA_SP, true, 2, Gs, h->d_sp, h->d_nsp, h->maxsp);
  } else {
    if (perm) NB_RE(double, NB_LPA, false, 0, G, h->d_ew, h->d_new, h->maxew);
    if (polar) NB_RE(double, NB_LPA, true, 0, G, h->d_ew, h->d_new, h->maxew);
  }
#undef NB_RE
}

void nb_real_energy(NbHandle *h, bool polar, int which) {
  if (h->ost) {
    if (h->want_vir) real_energy_t<true, true>(h, polar, which);
    else real_energy_t<false, true>(h, polar, which);
  } else {
    if (h->want_vir) real_energy_t<true, false>(h, polar, which);
    else real_energy_t<false, false>(h, polar, which);
  }
}

void nb_vdw(NbHandle *h) {
  int n = h->n;
  ScaleTab sc = scale_tab(h);
  double off = h->prm.vdw_cutoff, cut = h->prm.vdw_taper * off;
  dim3 B(256), G = xcd_grid((long)n * NB_LPA, 256);
  bool mixed = h->prm.precision == NB_PREC_MIXED, vir = h->want_vir;
  // softcore only matters when some atom is soft and lambda < 1 or its derivatives are wanted
  bool lam = h->has_soft && (h->lambda < 1.0 || h->want_dudl);
  double *vp = vir ? nb_vir_region(h, G.x) : nullptr;
  Softcore sf = softcore_of(h);
#define NB_VDW(T, V, L)                                                                                           \
  NB_LAUNCH((k_vdw<T, NB_LPA, V, L>), G, B, 0, n, h->d_red4, h->d_vclass, h->d_nbr, h->d_nnb, h->maxnb, h->cell,    \
            h->d_vdw_table, h->n_class, sc, off, cut, h->prm.halgren_delta, h->prm.halgren_gamma, h->d_gred_s,     \
            h->d_energy, h->d_count, vp, h->d_soft, h->d_mol, sf, h->d_gred_l)
  if (mixed) {
    if (lam) { if (vir) NB_VDW(float, true, true); else NB_VDW(float, false, true); }
    else { if (vir) NB_VDW(float, true, false); else NB_VDW(float, false, false); }
  } else {
    if (lam) { if (vir) NB_VDW(double, true, true); else NB_VDW(double, false, true); }
    else { if (vir) NB_VDW(double, true, false); else NB_VDW(double, false, false); }
  }
#undef NB_VDW
}
