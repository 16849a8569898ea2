// nb_real.hip -- real-space pair kernels, direct field, PCG, energies.
//
// One thread per sorted atom walks its own full-list row, so every output
// (field, force, torque) belongs to the thread and is written without
// atomics.  Pair arithmetic is templated on T (double: FP64 form; float:
// mixed form) and displacements are formed in FP64 first.
//
//  * k_perm_field: PermanentFieldRegion real space (PermanentFieldRegion.
//    java:587-808) -- field (group mask d11) and fieldCR (energy mask p) --
//    and, as a side effect, the Ewald (r <= off) and preconditioner
//    (r < 4.5 A) rows, as the reference builds realSpaceLists (:679-714).
//  * k_direct: DirectRegion (DirectRegion.java:170-245).
//  * k_ind_field / k_precond_field / k_ind_reduce: InducedDipoleFieldRegion
//    (:308-450), PCGSolver.PreconditionerRegion (PCGSolver.java:1270-1390),
//    InducedDipoleFieldReduceRegion (:159-215).
//  * k_pcg_*: PCGSolver.scfByPCG and its regions (PCGSolver.java:487-1178).
//  * k_recip_energy: ReciprocalEnergyRegion (ReciprocalEnergyRegion.java:381-746).
//  * k_real_energy: RealSpaceEnergyRegion (RealSpaceEnergyRegion.java:803-2177).
//  * k_vdw: VanDerWaals.VanDerWaalsLoop (VanDerWaals.java:1638-1900).
//  * k_finalize: torque -> frame-atom gradients (Torque.java:111-316),
//    reduced-hydrogen chain rule, and the return to caller order.
#include <cmath>
#include <stdexcept>

#include "nb_state.h"

namespace {

struct ScaleTab { double m[16], p[16], d[16], v[16]; };

// energy slots
enum { E_PERM_REAL = 0, E_PERM_SELF, E_PERM_RECIP, E_POL_REAL, E_POL_SELF, E_POL_RECIP, E_VDW, E_QTOT, E_NSLOT };
// PCG scalar slots
enum { S_PAP = 0, S_PAPP, S_RZ, S_RZP, S_EPS, S_EPSP, S_RZN, S_RZPN, S_NSLOT };

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block sum then one atomic per block; blockDim.x <= 1024
__device__ __forceinline__ void block_add(double v, double *dst) {
  __shared__ double part[16];
  v = wave_sum(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) part[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++) s += part[i];
    atomicAdd(dst, s);
  }
}

__device__ __forceinline__ void block_add_ll(long long v, int64_t *dst) {
  __shared__ long long part[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) part[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long s = 0;
    int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++) s += part[i];
    atomicAdd((unsigned long long *)dst, (unsigned long long)s);
  }
}

__device__ __forceinline__ V3<double> ld3d(const double *x, int i) { return V3<double>{x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
template <typename T> __device__ __forceinline__ V3<T> ld3t(const double *x, int i) {
  return V3<T>{(T)x[3 * i], (T)x[3 * i + 1], (T)x[3 * i + 2]};
}
__device__ __forceinline__ void st3(double *x, int i, V3<double> v) { x[3 * i] = v.x; x[3 * i + 1] = v.y; x[3 * i + 2] = v.z; }
template <typename T> __device__ __forceinline__ void add3(double *x, int i, V3<T> v) {
  x[3 * i] += (double)v.x; x[3 * i + 1] += (double)v.y; x[3 * i + 2] += (double)v.z;
}

__global__ void k_gather_pos(int n, const int32_t *__restrict__ perm, const double *__restrict__ xyz,
                             double4 *__restrict__ pos) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  pos[s] = make_double4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.0);
}

// ---------------------------------------------------------------------------
// permanent field and Ewald / preconditioner rows
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(128) k_perm_field(int n, const double4 *__restrict__ pos, const AtomMP *__restrict__ mp,
                                                    const uint32_t *__restrict__ nbr, const int32_t *__restrict__ nnb,
                                                    CellD cell, double off2, double pcut, double aewald, ScaleTab sc,
                                                    double *__restrict__ field, double *__restrict__ fieldp,
                                                    uint32_t *__restrict__ ew, int maxew, int32_t *__restrict__ new_,
                                                    uint32_t *__restrict__ pc, int maxpc, int32_t *__restrict__ npc,
                                                    int32_t *__restrict__ flags, int64_t *__restrict__ count) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  long long cnt = 0;
  if (s < n) {
    double4 pi = pos[s];
    AtomMP ai = mp[s];
    T ipdi = (T)ai.ipd, ti = (T)ai.thole;
    V3<T> fd{0, 0, 0}, fp{0, 0, 0};
    int ne = 0, np = 0;
    int m = nnb[s];
    for (int t = 0; t < m; t++) {
      uint32_t e = nbr[(size_t)t * n + s];
      int k = e & NB_INDEX_MASK, code = e >> NB_CODE_SHIFT;
      double4 pk = pos[k];
      double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
      double r2 = image(cell, dx, dy, dz);
      if (r2 > off2) continue;
      if (ne < maxew) ew[(size_t)ne * n + s] = e;
      ne++;
      double rd = sqrt(r2);
      if (rd < pcut) {
        if (np < maxpc) pc[(size_t)np * n + s] = e;
        np++;
      }
      cnt++;
      AtomMP ak = mp[k];
      MP<T> K = load_mp<T>(ak);
      T r = (T)rd, rinv = (T)1 / r;
      V3<T> R{(T)-dx, (T)-dy, (T)-dz};  // from k to i
      T bn[4];
      ewald_bn<T, 3>(r, rinv, (T)aewald, bn);
      Thole<T> th = thole<T, false>(r, R, ipdi, (T)ak.ipd, ti, (T)ak.thole);
      T rr3 = rinv * rinv * rinv, rr5 = (T)3 * rr3 * rinv * rinv, rr7 = (T)5 * rr5 * rinv * rinv;
      T ds = (T)sc.d[code], ps = (T)sc.p[code];
      fd += mp_field(R, K, bn[1] - ((T)1 - th.s3 * ds) * rr3, bn[2] - ((T)1 - th.s5 * ds) * rr5,
                     bn[3] - ((T)1 - th.s7 * ds) * rr7);
      fp += mp_field(R, K, bn[1] - ((T)1 - th.s3 * ps) * rr3, bn[2] - ((T)1 - th.s5 * ps) * rr5,
                     bn[3] - ((T)1 - th.s7 * ps) * rr7);
    }
    st3(field, s, V3<double>{(double)fd.x, (double)fd.y, (double)fd.z});
    st3(fieldp, s, V3<double>{(double)fp.x, (double)fp.y, (double)fp.z});
    new_[s] = ne;
    npc[s] = np;
    if (ne > maxew) atomicMax(&flags[1], ne);
    if (np > maxpc) atomicMax(&flags[2], np);
  }
  block_add_ll(cnt, count);
}

// self + reciprocal permanent field and the direct dipoles
__global__ void k_direct(int n, const double *__restrict__ gmp, const double *__restrict__ phip, double a3, int recip,
                         const double *__restrict__ polar, double *__restrict__ field, double *__restrict__ fieldp,
                         double *__restrict__ u, double *__restrict__ up, double *__restrict__ ud,
                         double *__restrict__ upd) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double a = polar[s];
  for (int d = 0; d < 3; d++) {
    double f = field[3 * s + d], fp = fieldp[3 * s + d];
    if (recip) {
      double x = a3 * gmp[10 * s + 1 + d] - phip[30 * (size_t)s + 21 + d];
      f += x;
      fp += x;
      field[3 * s + d] = f;
      fieldp[3 * s + d] = fp;
    }
    u[3 * s + d] = ud[3 * s + d] = a * f;
    up[3 * s + d] = upd[3 * s + d] = a * fp;
  }
}

// real-space field of induced dipole pairs (u, uCR) over a row list.
// PRE: dipoles are polar * r (preconditioner) and aewald = 0.
template <typename T, bool PRE>
__global__ void __launch_bounds__(128) k_dip_field(int n, const double4 *__restrict__ pos, const AtomMP *__restrict__ mp,
                                                   const uint32_t *__restrict__ lst, const int32_t *__restrict__ cnt,
                                                   CellD cell, double aewald, const double *__restrict__ u,
                                                   const double *__restrict__ up, const double *__restrict__ polar,
                                                   double *__restrict__ f, double *__restrict__ fp) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double4 pi = pos[s];
  T ipdi = (T)mp[s].ipd, ti = (T)mp[s].thole;
  V3<T> a{0, 0, 0}, b{0, 0, 0};
  int m = cnt[s];
  for (int t = 0; t < m; t++) {
    int k = lst[(size_t)t * n + s] & NB_INDEX_MASK;
    double4 pk = pos[k];
    double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
    double r2 = image(cell, dx, dy, dz);
    T r = (T)sqrt(r2), rinv = (T)1 / r;
    V3<T> R{(T)dx, (T)dy, (T)dz};
    T bn[3];
    ewald_bn<T, 2>(r, rinv, (T)aewald, bn);
    Thole<T> th = thole<T, false>(r, R, ipdi, (T)mp[k].ipd, ti, (T)mp[k].thole);
    T rr3 = rinv * rinv * rinv, rr5 = (T)3 * rr3 * rinv * rinv;
    T B1 = bn[1] - ((T)1 - th.s3) * rr3, B2 = bn[2] - ((T)1 - th.s5) * rr5;
    V3<T> uk = ld3t<T>(u, k), pk3 = ld3t<T>(up, k);
    if (PRE) {
      T pol = (T)polar[k];
      uk = pol * uk;
      pk3 = pol * pk3;
    }
    a += dip_field(R, uk, B1, B2);
    b += dip_field(R, pk3, B1, B2);
  }
  st3(f, s, V3<double>{(double)a.x, (double)a.y, (double)a.z});
  st3(fp, s, V3<double>{(double)b.x, (double)b.y, (double)b.z});
}

__global__ void k_ind_reduce(int n, double a3, const double *__restrict__ u, const double *__restrict__ up,
                             const double *__restrict__ phiu, const double *__restrict__ phiup, double *__restrict__ f,
                             double *__restrict__ fp) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  for (int d = 0; d < 3; d++) {
    f[3 * s + d] += a3 * u[3 * s + d] - phiu[30 * (size_t)s + 21 + d];
    fp[3 * s + d] += a3 * up[3 * s + d] - phiup[30 * (size_t)s + 21 + d];
  }
}

// ---------------------------------------------------------------------------
// PCG vector kernels (PCGSolver.java:609-1178)
// ---------------------------------------------------------------------------
__global__ void k_fill(int m, double *x, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) x[i] = v;
}

__global__ void k_pcg_init_r(int n, const double *__restrict__ polar, const double *__restrict__ ud,
                             const double *__restrict__ upd, const double *__restrict__ u, const double *__restrict__ up,
                             const double *__restrict__ f, const double *__restrict__ fp, double *__restrict__ r,
                             double *__restrict__ rp) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  double a = polar[q / 3];
  if (a > 0.0) {
    r[q] = (ud[q] - u[q]) / a + f[q];
    rp[q] = (upd[q] - up[q]) / a + fp[q];
  } else {
    r[q] = rp[q] = 0.0;
  }
}

// z = polar (E_pre + scale r); optionally p = z; accumulates r.z into slots
__global__ void k_pcg_z(int n, const double *__restrict__ polar, double psc, const double *__restrict__ f,
                        const double *__restrict__ fp, const double *__restrict__ r, const double *__restrict__ rp,
                        double *__restrict__ z, double *__restrict__ zp, double *__restrict__ p, double *__restrict__ pp,
                        int setp, double *__restrict__ scal, int slot) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  double rz = 0.0, rzp = 0.0;
  if (q < 3 * n) {
    double a = polar[q / 3];
    double zz = a * (f[q] + psc * r[q]), zzp = a * (fp[q] + psc * rp[q]);
    z[q] = zz;
    zp[q] = zzp;
    if (setp) { p[q] = zz; pp[q] = zzp; }
    rz = r[q] * zz;
    rzp = rp[q] * zzp;
  }
  block_add(rz, scal + slot);
  block_add(rzp, scal + slot + 1);
}

// vec = A p = p / polar - E(p); accumulates p.Ap and r.z
__global__ void k_pcg_ap(int n, const double *__restrict__ polar, const double *__restrict__ p,
                         const double *__restrict__ pp, const double *__restrict__ f, const double *__restrict__ fp,
                         const double *__restrict__ r, const double *__restrict__ rp, const double *__restrict__ z,
                         const double *__restrict__ zp, double *__restrict__ vec, double *__restrict__ vecp,
                         double *__restrict__ scal) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  double a1 = 0, a2 = 0, a3 = 0, a4 = 0;
  if (q < 3 * n) {
    double a = polar[q / 3];
    double v = 0.0, vp = 0.0;
    if (a > 0.0) {
      v = p[q] / a - f[q];
      vp = pp[q] / a - fp[q];
    }
    vec[q] = v;
    vecp[q] = vp;
    a1 = p[q] * v;
    a2 = pp[q] * vp;
    a3 = r[q] * z[q];
    a4 = rp[q] * zp[q];
  }
  block_add(a1, scal + S_PAP);
  block_add(a2, scal + S_PAPP);
  block_add(a3, scal + S_RZ);
  block_add(a4, scal + S_RZP);
}

__global__ void k_pcg_update(int n, const double *__restrict__ scal_in, const double *__restrict__ vec,
                             const double *__restrict__ vecp, const double *__restrict__ p, const double *__restrict__ pp,
                             double *__restrict__ r, double *__restrict__ rp, double *__restrict__ u,
                             double *__restrict__ up, double *__restrict__ scal) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  double pap = scal_in[S_PAP], papp = scal_in[S_PAPP];
  double al = scal_in[S_RZ], alp = scal_in[S_RZP];
  if (pap != 0.0) al /= pap;
  if (papp != 0.0) alp /= papp;
  double e1 = 0.0, e2 = 0.0;
  if (q < 3 * n) {
    double rr = r[q] - al * vec[q], rrp = rp[q] - alp * vecp[q];
    r[q] = rr;
    rp[q] = rrp;
    u[q] += al * p[q];
    up[q] += alp * pp[q];
    e1 = rr * rr;
    e2 = rrp * rrp;
  }
  block_add(e1, scal + S_EPS);
  block_add(e2, scal + S_EPSP);
}

__global__ void k_pcg_p(int n, const double *__restrict__ scal, const double *__restrict__ z,
                        const double *__restrict__ zp, double *__restrict__ p, double *__restrict__ pp) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  double prev = scal[S_RZ], prevp = scal[S_RZP];
  if (prev == 0.0) prev = 1.0;
  if (prevp == 0.0) prevp = 1.0;
  double b = scal[S_RZN] / prev, bp = scal[S_RZPN] / prevp;
  p[q] = z[q] + b * p[q];
  pp[q] = zp[q] + bp * pp[q];
}

// ---------------------------------------------------------------------------
// reciprocal energy, gradient and torque per atom
// ---------------------------------------------------------------------------
struct Frac2 { double a[3][3]; };

__global__ void k_recip_energy(int n, const double *__restrict__ gmp, const double *__restrict__ fmp,
                               const double *__restrict__ phip, const double *__restrict__ phiu,
                               const double *__restrict__ phiup, const double *__restrict__ u,
                               const double *__restrict__ up, Frac2 fr, double el, double aw, int polar, int mutual,
                               double *__restrict__ grad, double *__restrict__ torque, double *__restrict__ energy) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double eself = 0, erec = 0, qt = 0, eselfi = 0, ereci = 0;
  if (s < n) {
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
    if (polar) {
      const double *ui = u + 3 * s, *upi = up + 3 * s;
      eselfi = aewald3 * (ui[0] * m[1] + ui[1] * m[2] + ui[2] * m[3]);
      double us[3] = {0.5 * (ui[0] + upi[0]), 0.5 * (ui[1] + upi[1]), 0.5 * (ui[2] + upi[2])};
      tq[0] += aewald4 * (m[2] * us[2] - m[3] * us[1]);
      tq[1] += aewald4 * (m[3] * us[0] - m[1] * us[2]);
      tq[2] += aewald4 * (m[1] * us[1] - m[2] * us[0]);
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
      for (int a = 0; a < 3; a++) df[a] += 0.5 * el * (fr.a[0][a] * h[0] + fr.a[1][a] * h[1] + fr.a[2][a] * h[2]);
      tq[0] += el * (-m[2] * sp[3] + m[3] * sp[2] - tt * (m[7] * sp[8] + m[5] * sp[9] + m[9] * sp[6] - m[8] * sp[7] - m[9] * sp[5] - m[6] * sp[9]));
      tq[1] += el * (-m[3] * sp[1] + m[1] * sp[3] - tt * (m[8] * sp[4] + m[9] * sp[7] + m[6] * sp[8] - m[4] * sp[8] - m[7] * sp[9] - m[8] * sp[6]));
      tq[2] += el * (-m[1] * sp[2] + m[2] * sp[1] - tt * (m[4] * sp[7] + m[7] * sp[5] + m[8] * sp[9] - m[7] * sp[4] - m[5] * sp[7] - m[9] * sp[8]));
    }
    for (int a = 0; a < 3; a++) { grad[3 * s + a] += df[a]; torque[3 * s + a] += tq[a]; }
  }
  block_add(eself, energy + E_PERM_SELF);
  block_add(0.5 * el * erec, energy + E_PERM_RECIP);
  block_add(qt, energy + E_QTOT);
  block_add(eselfi, energy + E_POL_SELF);
  block_add(0.5 * el * ereci, energy + E_POL_RECIP);
}

// ---------------------------------------------------------------------------
// real-space permanent + polarization energy, i-side force and torque
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(128) k_real_energy(int n, const double4 *__restrict__ pos, const AtomMP *__restrict__ mp,
                                                     const uint32_t *__restrict__ ew, const int32_t *__restrict__ new_,
                                                     CellD cell, double aewald, ScaleTab sc, int polar, int direct,
                                                     const double *__restrict__ u, const double *__restrict__ up,
                                                     double el, double *__restrict__ grad, double *__restrict__ torque,
                                                     double *__restrict__ energy, int64_t *__restrict__ count) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double ep = 0.0, ei = 0.0;
  long long cnt = 0;
  if (s < n) {
    double4 pi = pos[s];
    AtomMP ai = mp[s];
    MP<T> I = load_mp<T>(ai);
    V3<T> ui = ld3t<T>(u, s), pi3 = ld3t<T>(up, s);
    T ipdi = (T)ai.ipd, ti = (T)ai.thole;
    V3<T> F{0, 0, 0}, Tq{0, 0, 0};
    int m = new_[s];
    for (int t = 0; t < m; t++) {
      uint32_t e = ew[(size_t)t * n + s];
      int k = e & NB_INDEX_MASK, code = e >> NB_CODE_SHIFT;
      double4 pk = pos[k];
      double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
      double r2 = image(cell, dx, dy, dz);
      AtomMP ak = mp[k];
      MP<T> K = load_mp<T>(ak);
      V3<T> r{(T)dx, (T)dy, (T)dz};
      T rl = (T)sqrt(r2), rinv = (T)1 / rl;
      T bn[7], rr[7];
      ewald_bn<T, 6>(rl, rinv, (T)aewald, bn);
      coulomb_rr<T, 6>(rinv, rr);
      PermGeo<T> g;
      perm_geometry(r, I, K, g);
      T epair = perm_terms<T, true>(r, I, K, g, bn, (T)1, F, Tq);
      T s1 = (T)1 - (T)sc.m[code];
      if (s1 != (T)0) epair -= s1 * perm_terms<T, true>(r, I, K, g, rr, -s1, F, Tq);
      if (epair != (T)0) cnt++;
      ep += (double)epair;
      if (polar) {
        Thole<T> th = thole<T, true>(rl, r, (T)ai.ipd, (T)ak.ipd, ti, (T)ak.thole);
        ei += (double)pol_terms<T, true>(r, I, K, ui, pi3, ld3t<T>(u, k), ld3t<T>(up, k), bn, rr, th, (T)sc.p[code],
                                         (T)sc.d[code], direct != 0, F, Tq);
      }
    }
    (void)ipdi;
    add3(grad, s, (T)el * F);
    add3(torque, s, (T)el * Tq);
  }
  block_add(0.5 * el * ep, energy + E_PERM_REAL);
  block_add(0.5 * el * ei, energy + E_POL_REAL);
  block_add_ll(cnt, count + 1);
}

// ---------------------------------------------------------------------------
// buffered 14-7 vdW on reduced sites; gradient on the reduced site
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(128) k_vdw(int n, const double4 *__restrict__ red, const int32_t *__restrict__ vclass,
                                             const uint32_t *__restrict__ nbr, const int32_t *__restrict__ nnb, CellD cell,
                                             const double *__restrict__ vtab, int nclass, ScaleTab sc, double off,
                                             double cut, double dl, double gm, double *__restrict__ gred,
                                             double *__restrict__ energy, int64_t *__restrict__ count) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double esum = 0.0;
  long long cnt = 0;
  if (s < n) {
    double off2 = off * off, cut2 = cut * cut;
    double a = cut, b = off, a2 = a * a, b2 = b * b;
    double den = (b - a) * (b - a) * (b - a) * (b - a) * (b - a);
    T c0 = (T)(b * b2 * (b2 - 5.0 * a * b + 10.0 * a2) / den), c1 = (T)(-30.0 * a2 * b2 / den);
    T c2 = (T)(30.0 * b * a * (b + a) / den), c3 = (T)(-10.0 * (a2 + 4.0 * a * b + b2) / den);
    T c4 = (T)(15.0 * (a + b) / den), c5 = (T)(-6.0 / den);
    T d1 = (T)dl, g1 = (T)1 + (T)gm, gg = (T)gm;
    T t1n = (T)pow(1.0 + dl, 7.0);
    double4 pi = red[s];
    int ci = vclass[s];
    V3<T> G{0, 0, 0};
    T e = 0;
    int m = nnb[s];
    for (int t = 0; t < m; t++) {
      uint32_t en = nbr[(size_t)t * n + s];
      int k = en & NB_INDEX_MASK, code = en >> NB_CODE_SHIFT;
      double vs = sc.v[code];
      if (!(vs > 0.0)) continue;
      double4 pk = red[k];
      double dx = pi.x - pk.x, dy = pi.y - pk.y, dz = pi.z - pk.z;
      double r2 = image(cell, dx, dy, dz);
      if (r2 > off2) continue;
      const double *pe = vtab + 2 * (ci * nclass + vclass[k]);
      T irv = (T)pe[0];
      if (!(irv > (T)0)) continue;
      cnt++;
      T rr = (T)sqrt(r2);
      T ev = (T)(vs * pe[1]);
      T rho = rr * irv;
      T rho2 = rho * rho, rho6 = rho2 * rho2 * rho2, rho7 = rho6 * rho;
      T rd = rho + d1, rd2 = rd * rd, rd6 = rd2 * rd2 * rd2, rd7 = rd6 * rd;
      T t1d = (T)1 / rd7, t2d = (T)1 / (rho7 + gg);
      T t1 = t1n * t1d, t2a = g1 * t2d, t2 = t2a - (T)2;
      T eik = ev * t1 * t2;
      T tp = 1, dtp = 0;
      if (r2 > cut2) {
        T r2t = (T)r2, r3 = r2t * rr, r4 = r2t * r2t, r5 = r2t * r3;
        tp = c5 * r5 + c4 * r4 + c3 * r3 + c2 * r2t + c1 * rr + c0;
        dtp = (T)5 * c5 * r4 + (T)4 * c4 * r3 + (T)3 * c3 * r2t + (T)2 * c2 * rr + c1;
      }
      e += eik * tp;
      T dt1 = t1 * (T)7 * rd6 * irv * t1d;
      T dt2 = t2a * (T)7 * rho6 * irv * t2d;
      T dedr = -ev * (dt1 * t2 + t1 * dt2);
      T ds = (eik * dtp + dedr * tp) / rr;
      G += ds * V3<T>{(T)dx, (T)dy, (T)dz};
    }
    esum = 0.5 * (double)e;
    gred[3 * s] = (double)G.x;
    gred[3 * s + 1] = (double)G.y;
    gred[3 * s + 2] = (double)G.z;
  }
  block_add(esum, energy + E_VDW);
  block_add_ll(cnt, count);
}

// ---------------------------------------------------------------------------
// gradient assembly in caller order
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3<double> unit_d(V3<double> a) { return (1.0 / sqrt(dot(a, a))) * a; }

__device__ __forceinline__ void aadd(double *g, int i, V3<double> v) {
  atomicAdd(&g[3 * i], v.x);
  atomicAdd(&g[3 * i + 1], v.y);
  atomicAdd(&g[3 * i + 2], v.z);
}

__global__ void k_finalize(int n, const int32_t *__restrict__ perm, const double *__restrict__ xyz,
                           const int32_t *__restrict__ frame, const int32_t *__restrict__ axis,
                           const int32_t *__restrict__ redidx, const double *__restrict__ redval,
                           const double *__restrict__ grad_s, const double *__restrict__ torque_s,
                           const double *__restrict__ gred_s, int do_vdw, int do_pme, double *__restrict__ grad,
                           int32_t *__restrict__ flags) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  V3<double> g = do_pme ? ld3d(grad_s, s) : V3<double>{0, 0, 0};
  if (do_vdw) {
    V3<double> gv = ld3d(gred_s, s);
    int h = redidx[i];
    double f = redval[i];
    if (h >= 0 && h != i && f > 0.0) {
      g += f * gv;
      aadd(grad, h, (1.0 - f) * gv);
    } else {
      g += gv;
    }
  }
  aadd(grad, i, g);
  if (!do_pme) return;
  int fr = frame[i];
  if (fr == NB_FRAME_NONE) return;
  V3<double> trq = ld3d(torque_s, s);
  if (!isfinite(trq.x) || !isfinite(trq.y) || !isfinite(trq.z)) atomicOr(&flags[4], 1);
  int iz = axis[3 * i], ix = axis[3 * i + 1], iy = axis[3 * i + 2];
  V3<double> o = ld3d(xyz, i);
  V3<double> u = ld3d(xyz, iz) - o;
  double ru = sqrt(dot(u, u));
  V3<double> v, w;
  if (fr != NB_FRAME_ZONLY) {
    v = ld3d(xyz, ix) - o;
  } else {
    v = V3<double>{1.0, 0.0, 0.0};
    if (fabs(u.x / ru) > 0.866) v = V3<double>{0.0, 1.0, 0.0};
  }
  double rv = sqrt(dot(v, v));
  if (fr == NB_FRAME_ZTHENBISECTOR || fr == NB_FRAME_THREEFOLD) w = ld3d(xyz, iy) - o;
  else w = cross(u, v);
  double rw = sqrt(dot(w, w));
  u = (1.0 / ru) * u; v = (1.0 / rv) * v; w = (1.0 / rw) * w;
  V3<double> uv = unit_d(cross(v, u)), uw = unit_d(cross(w, u)), vw = unit_d(cross(w, v));
  double cuv = dot(u, v), suv = sqrt(1.0 - cuv * cuv);
  double dpu = -dot(trq, u), dpv = -dot(trq, v), dpw = -dot(trq, w);
  V3<double> gz{0, 0, 0}, gx{0, 0, 0}, gy{0, 0, 0};
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
  aadd(grad, iz, gz);
  if (ix >= 0 && fr != NB_FRAME_ZONLY) aadd(grad, ix, gx);
  if (iy >= 0 && (fr == NB_FRAME_ZTHENBISECTOR || fr == NB_FRAME_THREEFOLD)) aadd(grad, iy, gy);
  aadd(grad, i, -(gz + gx + gy));
}

ScaleTab scale_tab(const NbHandle *h) {
  ScaleTab t;
  for (int c = 0; c < 16; c++) {
    int rel = c & 7, sg = (c >> 3) & 1;
    t.m[c] = (rel >= 2 && rel <= 5) ? h->prm.m_scale[rel - 2] : 1.0;
    double p = 1.0;
    if (rel >= 2 && rel <= 4) {
      p = h->prm.p_scale[rel - 2];
      if (rel == 4 && sg) p *= h->prm.intra14;
    }
    t.p[c] = p;
    t.d[c] = sg ? h->prm.d11 : 1.0;
    t.v[c] = (rel >= 2 && rel <= 4) ? h->prm.v_scale[rel - 2] : 1.0;
  }
  return t;
}

}  // namespace

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline dim3 grid_for(int n, int b) { return dim3((n + b - 1) / b); }

void nb_gather_sorted(NbHandle *h) {
  hipLaunchKernelGGL(k_gather_pos, grid_for(h->n, 256), dim3(256), 0, h->stream, h->n, h->d_perm, h->d_xyz, h->d_pos4);
}

void nb_zero_outputs(NbHandle *h) {
  int n = h->n;
  hipLaunchKernelGGL(k_fill, grid_for(3 * n, 256), dim3(256), 0, h->stream, 3 * n, h->d_grad_s, 0.0);
  hipLaunchKernelGGL(k_fill, grid_for(3 * n, 256), dim3(256), 0, h->stream, 3 * n, h->d_torque_s, 0.0);
  hipLaunchKernelGGL(k_fill, grid_for(3 * n, 256), dim3(256), 0, h->stream, 3 * n, h->d_gred_s, 0.0);
  hipLaunchKernelGGL(k_fill, grid_for(3 * n, 256), dim3(256), 0, h->stream, 3 * n, h->d_grad, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, h->stream, 16, h->d_energy, 0.0);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, h->stream, 4, (double *)h->d_count, 0.0);
}

void nb_perm_field(NbHandle *h) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  ScaleTab sc = scale_tab(h);
  for (int attempt = 0; attempt < 4; attempt++) {
    int32_t zeros[3] = {0, 0, 0};
    HIPCHECK(hipMemcpyAsync(h->d_flags + 1, zeros, sizeof(zeros) - sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, h->stream, 1, (double *)(h->d_count + 2), 0.0);
    double off2 = h->prm.ewald_cutoff * h->prm.ewald_cutoff;
    if (h->prm.precision == NB_PREC_MIXED)
      hipLaunchKernelGGL(k_perm_field<float>, grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_pos4, h->d_mp, h->d_nbr,
                         h->d_nnb, h->cell, off2, h->prm.precond_cutoff, aw, sc, h->d_field, h->d_fieldp, h->d_ew,
                         h->maxew, h->d_new, h->d_pc, h->maxpc, h->d_npc, h->d_flags, h->d_count + 2);
    else
      hipLaunchKernelGGL(k_perm_field<double>, grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_pos4, h->d_mp,
                         h->d_nbr, h->d_nnb, h->cell, off2, h->prm.precond_cutoff, aw, sc, h->d_field, h->d_fieldp,
                         h->d_ew, h->maxew, h->d_new, h->d_pc, h->maxpc, h->d_npc, h->d_flags, h->d_count + 2);
    int32_t f[2];
    HIPCHECK(hipMemcpyAsync(f, h->d_flags + 1, sizeof(f), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    if (f[0] <= h->maxew && f[1] <= h->maxpc) return;
    if (f[0] > h->maxew) {
      hipFree(h->d_ew);
      h->maxew = f[0] + f[0] / 4 + 16;
      HIPCHECK(hipMalloc(&h->d_ew, sizeof(uint32_t) * (size_t)h->maxew * n));
    }
    if (f[1] > h->maxpc) {
      hipFree(h->d_pc);
      h->maxpc = f[1] + f[1] / 4 + 16;
      HIPCHECK(hipMalloc(&h->d_pc, sizeof(uint32_t) * (size_t)h->maxpc * n));
    }
  }
  throw std::runtime_error("Ewald list capacity could not be established");
}

void nb_direct(NbHandle *h) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  double a3 = 4.0 / 3.0 * aw * aw * aw / NB_SQRT_PI;
  hipLaunchKernelGGL(k_direct, grid_for(n, 256), dim3(256), 0, h->stream, n, h->d_gmp, h->d_phip, a3, (int)use_recip,
                     h->d_polar, h->d_field, h->d_fieldp, h->d_u, h->d_up, h->d_udir, h->d_updir);
}

void nb_induced_field(NbHandle *h, const double *u, const double *up, double *f, double *fp) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  if (h->prm.precision == NB_PREC_MIXED)
    hipLaunchKernelGGL((k_dip_field<float, false>), grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_pos4, h->d_mp,
                       h->d_ew, h->d_new, h->cell, aw, u, up, h->d_polar, f, fp);
  else
    hipLaunchKernelGGL((k_dip_field<double, false>), grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_pos4, h->d_mp,
                       h->d_ew, h->d_new, h->cell, aw, u, up, h->d_polar, f, fp);
  if (use_recip) {
    nb_pme_induced(h, u, up);
    double a3 = 4.0 / 3.0 * aw * aw * aw / NB_SQRT_PI;
    hipLaunchKernelGGL(k_ind_reduce, grid_for(n, 256), dim3(256), 0, h->stream, n, a3, u, up, h->d_phiu, h->d_phiup, f,
                       fp);
  }
}

void nb_precond_field(NbHandle *h, const double *r, const double *rp, double *f, double *fp) {
  int n = h->n;
  if (h->prm.precision == NB_PREC_MIXED)
    hipLaunchKernelGGL((k_dip_field<float, true>), grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_pos4, h->d_mp,
                       h->d_pc, h->d_npc, h->cell, 0.0, r, rp, h->d_polar, f, fp);
  else
    hipLaunchKernelGGL((k_dip_field<double, true>), grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_pos4, h->d_mp,
                       h->d_pc, h->d_npc, h->cell, 0.0, r, rp, h->d_polar, f, fp);
}

int nb_pcg(NbHandle *h, nb_energies *out) {
  int n = h->n;
  dim3 B(256), G = grid_for(3 * n, 256);
  double psc = h->prm.precond_scale;
  double *f = h->d_field, *fp = h->d_fieldp;  // permanent fields are no longer needed
  nb_induced_field(h, h->d_u, h->d_up, f, fp);
  hipLaunchKernelGGL(k_pcg_init_r, G, B, 0, h->stream, n, h->d_polar, h->d_udir, h->d_updir, h->d_u, h->d_up, f, fp,
                     h->d_r, h->d_rp);
  nb_precond_field(h, h->d_r, h->d_rp, f, fp);
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, h->stream, S_NSLOT, h->d_scal, 0.0);
  hipLaunchKernelGGL(k_pcg_z, G, B, 0, h->stream, n, h->d_polar, psc, f, fp, h->d_r, h->d_rp, h->d_z, h->d_zp, h->d_p,
                     h->d_pp, 1, h->d_scal, (int)S_RZN);
  int maxit = h->prm.max_scf_iter > 0 ? h->prm.max_scf_iter : 1000;
  double eps = 100.0, prev;
  int iter = 0, status = NB_OK;
  double sc[S_NSLOT];
  for (;;) {
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, h->stream, S_NSLOT, h->d_scal, 0.0);
    nb_induced_field(h, h->d_p, h->d_pp, f, fp);
    hipLaunchKernelGGL(k_pcg_ap, G, B, 0, h->stream, n, h->d_polar, h->d_p, h->d_pp, f, fp, h->d_r, h->d_rp, h->d_z,
                       h->d_zp, h->d_vec, h->d_vecp, h->d_scal);
    hipLaunchKernelGGL(k_pcg_update, G, B, 0, h->stream, n, h->d_scal, h->d_vec, h->d_vecp, h->d_p, h->d_pp, h->d_r,
                       h->d_rp, h->d_u, h->d_up, h->d_scal);
    HIPCHECK(hipMemcpyAsync(h->h_pinned, h->d_scal, sizeof(double) * S_NSLOT, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    for (int q = 0; q < S_NSLOT; q++) sc[q] = h->h_pinned[q];
    prev = eps;
    double e = sc[S_EPS] > sc[S_EPSP] ? sc[S_EPS] : sc[S_EPSP];
    iter++;
    eps = 4.80321 * sqrt(e / n);
    if (!(eps <= prev)) { status = NB_ERR_SCF_DIVERGED; break; }
    if (iter >= maxit) { status = NB_ERR_SCF_MAXITER; break; }
    if (eps < h->prm.polar_eps) break;
    nb_precond_field(h, h->d_r, h->d_rp, f, fp);
    hipLaunchKernelGGL(k_pcg_z, G, B, 0, h->stream, n, h->d_polar, psc, f, fp, h->d_r, h->d_rp, h->d_z, h->d_zp, h->d_p,
                       h->d_pp, 0, h->d_scal, (int)S_RZN);
    hipLaunchKernelGGL(k_pcg_p, G, B, 0, h->stream, n, h->d_scal, h->d_z, h->d_zp, h->d_p, h->d_pp);
  }
  out->pcg_iters = iter;
  out->pcg_eps = eps;
  // final induced field: fills the induced reciprocal potentials for the energy
  nb_induced_field(h, h->d_u, h->d_up, f, fp);
  return status;
}

void nb_recip_energy(NbHandle *h, bool polar) {
  int n = h->n;
  Frac2 fr;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) fr.a[a][b] = h->afrac[a][b];
  hipLaunchKernelGGL(k_recip_energy, grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_gmp, h->d_fmp, h->d_phip,
                     h->d_phiu, h->d_phiup, h->d_u, h->d_up, fr, h->prm.electric, h->prm.ewald_alpha, (int)polar,
                     (int)(h->prm.polarization == NB_POL_MUTUAL), h->d_grad_s, h->d_torque_s, h->d_energy);
}

void nb_real_energy(NbHandle *h, bool polar) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  ScaleTab sc = scale_tab(h);
  int direct = h->prm.polarization == NB_POL_DIRECT;
  if (h->prm.precision == NB_PREC_MIXED)
    hipLaunchKernelGGL(k_real_energy<float>, grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_pos4, h->d_mp, h->d_ew,
                       h->d_new, h->cell, aw, sc, (int)polar, direct, h->d_u, h->d_up, h->prm.electric, h->d_grad_s,
                       h->d_torque_s, h->d_energy, h->d_count);
  else
    hipLaunchKernelGGL(k_real_energy<double>, grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_pos4, h->d_mp, h->d_ew,
                       h->d_new, h->cell, aw, sc, (int)polar, direct, h->d_u, h->d_up, h->prm.electric, h->d_grad_s,
                       h->d_torque_s, h->d_energy, h->d_count);
}

void nb_vdw(NbHandle *h) {
  int n = h->n;
  ScaleTab sc = scale_tab(h);
  double off = h->prm.vdw_cutoff, cut = h->prm.vdw_taper * off;
  if (h->prm.precision == NB_PREC_MIXED)
    hipLaunchKernelGGL(k_vdw<float>, grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_red4, h->d_vclass, h->d_nbr,
                       h->d_nnb, h->cell, h->d_vdw_table, h->n_class, sc, off, cut, h->prm.halgren_delta,
                       h->prm.halgren_gamma, h->d_gred_s, h->d_energy, h->d_count);
  else
    hipLaunchKernelGGL(k_vdw<double>, grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_red4, h->d_vclass, h->d_nbr,
                       h->d_nnb, h->cell, h->d_vdw_table, h->n_class, sc, off, cut, h->prm.halgren_delta,
                       h->prm.halgren_gamma, h->d_gred_s, h->d_energy, h->d_count);
}

void nb_finalize_gradient(NbHandle *h) {
  int n = h->n;
  hipLaunchKernelGGL(k_finalize, grid_for(n, 128), dim3(128), 0, h->stream, n, h->d_perm, h->d_xyz, h->d_frame,
                     h->d_axis, h->d_redidx, h->d_redval, h->d_grad_s, h->d_torque_s, h->d_gred_s, (int)h->prm.do_vdw,
                     (int)h->prm.do_pme, h->d_grad, h->d_flags);
}
