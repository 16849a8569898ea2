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

#include "nb_kern.h"
#include "nb_atom.h"

namespace {


__global__ void k_gather_pos(int n, const int32_t *__restrict__ perm, const double *__restrict__ xyz,
                             const AtomMP *__restrict__ mp, double4 *__restrict__ pos, void *rec, void *rec2,
                             int mixed, CellD cell, const int32_t *__restrict__ wrap) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int i = perm[s];
  double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  if (cell.multi) {
    // the lattice translation fixed at the last list build (image sums)
    int wx = wrap[3 * s], wy = wrap[3 * s + 1], wz = wrap[3 * s + 2];
    x -= wx * cell.Ai00 + wy * cell.Ai10 + wz * cell.Ai20;
    y -= wy * cell.Ai11 + wz * cell.Ai21;
    z -= wz * cell.Ai22;
  }
  pos[s] = make_double4(x, y, z, 0.0);
  dip_put_static(rec, mixed, s, x, y, z, mp[s].ipd, mp[s].thole);
  dip_put_static(rec2, mixed, s, x, y, z, mp[s].ipd, mp[s].thole);
}

// self + reciprocal permanent field and the direct dipoles
__global__ void k_direct(int n, const double *__restrict__ gmp, const double *__restrict__ phip, double a3, int recip,
                         const double *__restrict__ polar, double *__restrict__ field, double *__restrict__ fieldp,
                         double *__restrict__ u, double *__restrict__ up, double *__restrict__ ud,
                         double *__restrict__ upd, void *rec, int mixed) {
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
    dip_put(rec, mixed, s, d, a * f, a * fp);
  }
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
// PCG vector kernels (PCGSolver.java:609-1178).  Dot products are reduced
// without atomics: every block writes its partial sum to part[slot][block],
// and each consumer block re-sums the partials in a fixed order (so the
// result is deterministic and no buffer has to be cleared).
// ---------------------------------------------------------------------------
__global__ void k_fill(int m, double *x, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) x[i] = v;
}

// r0 = (u_direct - u0) / polar + E(u0); the packed records receive polar * r0
__global__ void k_pcg_init_r(int n, const double *__restrict__ polar, const double *__restrict__ ud,
                             const double *__restrict__ upd, const double *__restrict__ u, const double *__restrict__ up,
                             const double *__restrict__ f, const double *__restrict__ fp, double *__restrict__ r,
                             double *__restrict__ rp, void *rec, int mixed) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  double a = polar[q / 3];
  double rr = 0.0, rrp = 0.0;
  if (a > 0.0) {
    rr = (ud[q] - u[q]) / a + f[q];
    rrp = (upd[q] - up[q]) / a + fp[q];
  }
  r[q] = rr;
  rp[q] = rrp;
  dip_put(rec, mixed, q / 3, q % 3, a * rr, a * rrp);
}

// Conjugate direction without reciprocal space (aperiodic systems): E(z) is
// the real-space field f alone; partial p.q into S_PAP / S_PAPP.
__global__ void __launch_bounds__(256) k_pcg_dir(int n, const double *__restrict__ polar, const double *__restrict__ z,
                                                 const double *__restrict__ zp, const double *__restrict__ f,
                                                 const double *__restrict__ fp, double *__restrict__ p,
                                                 double *__restrict__ pp, double *__restrict__ vec,
                                                 double *__restrict__ vecp, double *__restrict__ part, int stride,
                                                 int nrz, const double *__restrict__ scal, int first,
                                                 const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;
  double beta[2];
  pcg_beta(part, stride, nrz, scal, first, beta);
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (q < 3 * n) {
    double a = polar[q / 3];
    double pn = p[q], qn = vec[q], pnp = pp[q], qnp = vecp[q];
    pcg_direction(a, z[q], f[q], beta[0], first, pn, qn);
    pcg_direction(a, zp[q], fp[q], beta[1], first, pnp, qnp);
    p[q] = pn;
    vec[q] = qn;
    pp[q] = pnp;
    vecp[q] = qnp;
    v[0] = pn * qn;
    v[1] = pnp * qnp;
  }
  __syncthreads();
  const int sl[2] = {S_PAP, S_PAPP};
  block_partials<2>(v, part, sl, stride);
}

// totals of two pairs of partial rows (counts na, nb; a negative count: the
// collapsed totals in scal) read in one pass with one barrier; the summation
// order is block_totals'
static __device__ __forceinline__ void totals_2x2(const double *__restrict__ part, int stride,
                                                  const double *__restrict__ scal, int na, int a0, int a1, int ta0,
                                                  int ta1, int nb, int b0, int b1, int tb0, int tb1, double (&ta)[2],
                                                  double (&tb)[2]) {
  __shared__ double red[4][4];
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const int m = max(na, nb);
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    if (i < na) {
      v[0] += part[(size_t)a0 * stride + i];
      v[1] += part[(size_t)a1 * stride + i];
    }
    if (i < nb) {
      v[2] += part[(size_t)b0 * stride + i];
      v[3] += part[(size_t)b1 * stride + i];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) v[k] = wave_sum_d(v[k]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 4; k++) red[k][w] = v[k];
  __syncthreads();
  double t[4];
#pragma unroll
  for (int k = 0; k < 4; k++) t[k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
  ta[0] = na < 0 ? scal[ta0] : t[0];
  ta[1] = na < 0 ? scal[ta1] : t[1];
  tb[0] = nb < 0 ? scal[tb0] : t[2];
  tb[1] = nb < 0 ? scal[tb1] : t[3];
}

// r -= alpha q, u += alpha p with alpha = r.z / p.q per system; the records
// of the preconditioner field receive polar * r; partial r.r into S_EPS /
// S_EPSP (reduced by the next k_precond_z); block 0 keeps the r.z totals as
// the denominators of the next conjugation.  The thread's vector loads are
// issued before the totals are read, so the two memory latencies overlap
// (one of the PCG iteration's latency-bound launches).
__global__ void __launch_bounds__(256) k_pcg_update(int n, const double *__restrict__ polar,
                                                    const double *__restrict__ vec, const double *__restrict__ vecp,
                                                    const double *__restrict__ p, const double *__restrict__ pp,
                                                    double *__restrict__ r, double *__restrict__ rp,
                                                    double *__restrict__ u, double *__restrict__ up,
                                                    double *__restrict__ part, int stride, int npq, int nrz,
                                                    void *rec, int mixed, double *__restrict__ scal,
                                                    const double *__restrict__ ctl) {
  // a recorded solve that has stopped keeps its dipoles and residuals
  if (ctl && ctl[0] != 0.0) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = q < 3 * n;
  double r0 = 0.0, rp0 = 0.0, v0 = 0.0, vp0 = 0.0, u0 = 0.0, up0 = 0.0, p0 = 0.0, pp0 = 0.0, a = 0.0;
  if (ok) {
    r0 = r[q];
    rp0 = rp[q];
    v0 = vec[q];
    vp0 = vecp[q];
    u0 = u[q];
    up0 = up[q];
    p0 = p[q];
    pp0 = pp[q];
    a = polar[q / 3];
  }
  double pq[2], rz[2];
  totals_2x2(part, stride, scal, npq, S_PAP, S_PAPP, S_PQT, S_PQTP, nrz, S_RZN, S_RZPN, S_RZT, S_RZTP, pq, rz);
  double al = rz[0], alp = rz[1];
  if (pq[0] != 0.0) al /= pq[0];
  if (pq[1] != 0.0) alp /= pq[1];
  double v[2] = {0.0, 0.0};
  if (ok) {
    const double rr = r0 - al * v0, rrp = rp0 - alp * vp0;
    r[q] = rr;
    rp[q] = rrp;
    u[q] = u0 + al * p0;
    up[q] = up0 + alp * pp0;
    dip_put(rec, mixed, q / 3, q % 3, a * rr, a * rrp);
    v[0] = rr * rr;
    v[1] = rrp * rrp;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    scal[S_RZOLD] = rz[0];
    scal[S_RZOLDP] = rz[1];
  }
  __syncthreads();
  const int out[2] = {S_EPS, S_EPSP};
  block_partials<2>(v, part, out, stride);
}

// FP64 records of the dipoles u, up for the FP64 residual of the refinement
__global__ void k_rec64(int n, const double4 *__restrict__ pos, const AtomMP *__restrict__ mp,
                        const double *__restrict__ u, const double *__restrict__ up, DipRec<double> *__restrict__ rec,
                        const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded refinement already finished
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  DipRec<double> r;
  double4 p = pos[s];
  r.x = p.x; r.y = p.y; r.z = p.z;
  r.ipd = mp[s].ipd;
  r.thole = mp[s].thole;
  for (int d = 0; d < 3; d++) {
    r.u[d] = u[3 * s + d];
    r.v[d] = up[3 * s + d];
  }
  r.pad = 0.0;
  rec[s] = r;
}

// FP64 residual of the solve (the r of PCGSolver.java:609-663 with every
// field term in FP64): r = (u_direct - u) / polar + f + a3 u - phi, f the
// FP64 real-space field, phi the FP64 reciprocal field (phiu null: none); the
// records of the preconditioner field receive polar * r (a restarted solve
// starts from them); partial r.r into S_EPS / S_EPSP.
__global__ void __launch_bounds__(256) k_refine_residual(int n, const double *__restrict__ polar, double a3,
                                                         const double *__restrict__ ud, const double *__restrict__ upd,
                                                         const double *__restrict__ u, const double *__restrict__ up,
                                                         const double *__restrict__ f, const double *__restrict__ fp,
                                                         const double *__restrict__ phiu,
                                                         const double *__restrict__ phiup, double *__restrict__ r,
                                                         double *__restrict__ rp, void *rec, int mixed,
                                                         double *__restrict__ part, int stride,
                                                         const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded refinement already finished
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (q < 3 * n) {
    int s = q / 3, d = q - 3 * s;
    double a = polar[s];
    double ft = f[q] + a3 * u[q], fpt = fp[q] + a3 * up[q];
    if (phiu) {
      ft -= phiu[30 * (size_t)s + 21 + d];
      fpt -= phiup[30 * (size_t)s + 21 + d];
    }
    double rr = 0.0, rrp = 0.0;
    if (a > 0.0) {
      rr = (ud[q] - u[q]) / a + ft;
      rrp = (upd[q] - up[q]) / a + fpt;
    }
    r[q] = rr;
    rp[q] = rrp;
    dip_put(rec, mixed, s, d, a * rr, a * rrp);
    v[0] = rr * rr;
    v[1] = rrp * rrp;
  }
  const int out[2] = {S_EPS, S_EPSP};
  block_partials<2>(v, part, out, stride);
}

// eps of the FP64 residual (the reference's measure, PCGSolver.java:543-547)
// for a recorded solve.  restart = 1 (before a recorded correction): a solve
// whose FP64 residual is not below polar_eps is reopened for the correction's
// iterations (S_DONE cleared, the eps rise test restarted); one that passes,
// or has failed, ends the refinement (S_RDONE), and the later corrections
// and residuals return at once.  restart = 0 (the last check, unless the
// refinement ended before): a solve whose FP64 residual is not below
// polar_eps is marked unfinished (S_DONE cleared), so the replay is repeated
// on the regular path, which takes further corrections.
__global__ void __launch_bounds__(256) k_refine_check(const double *__restrict__ part, int stride, int nb,
                                                      double *__restrict__ scal, int n, double polar_eps,
                                                      int restart) {
  const int in[2] = {S_EPS, S_EPSP};
  double t[2];
  block_totals<2>(part, in, stride, nb, t);
  if (threadIdx.x == 0 && scal[S_RDONE] == 0.0) {
    double e = t[0] > t[1] ? t[0] : t[1];
    double eps = 4.80321 * sqrt(e / n);
    const bool pass = eps < polar_eps, failed = scal[S_STAT] != 0.0;
    scal[S_EPS] = t[0];
    scal[S_EPSP] = t[1];
    if (!failed) scal[S_EPSV] = eps;
    if (restart) {
      if (pass || failed) {
        scal[S_RDONE] = 1.0;
      } else {
        scal[S_DONE] = 0.0;
        scal[S_PREV] = 100.0;
        scal[S_RCYC] += 1.0;
      }
    } else if (!failed && !pass) {
      scal[S_DONE] = 0.0;
    }
  }
}

// residual norms of an SOR cycle: scal[S_EPS], scal[S_EPSP]
__global__ void __launch_bounds__(256) k_pcg_eps(const double *__restrict__ part, int stride, int nb,
                                                 double *__restrict__ scal) {
  const int in[2] = {S_EPS, S_EPSP};
  double t[2];
  block_totals<2>(part, in, stride, nb, t);
  if (threadIdx.x == 0) {
    scal[S_EPS] = t[0];
    scal[S_EPSP] = t[1];
  }
}

__global__ void k_pcg_ctl_init(double *__restrict__ scal) {
  if (threadIdx.x == 0) {
    scal[S_DONE] = 0.0;
    scal[S_ITER] = 0.0;
    scal[S_PREV] = 100.0;
    scal[S_STAT] = 0.0;
    scal[S_EPSV] = 0.0;
    scal[S_RDONE] = 0.0;
    scal[S_RCYC] = 0.0;
  }
}

// SOR step (SORRegion.java:236-258): u <- u + w (u_direct + polar E(u) - u)
// for u and uCR; the packed records receive the new dipoles; partial sums of
// the squared steps go to the residual slots
__global__ void __launch_bounds__(256) k_sor_update(int n, double w, const double *__restrict__ polar,
                                                    const double *__restrict__ ud, const double *__restrict__ upd,
                                                    const double *__restrict__ f, const double *__restrict__ fp,
                                                    double *__restrict__ u, double *__restrict__ up,
                                                    double *__restrict__ part, int stride, void *rec, int mixed) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (q < 3 * n) {
    double a = polar[q / 3];
    double du = w * (ud[q] + a * f[q] - u[q]), dv = w * (upd[q] + a * fp[q] - up[q]);
    double nu = u[q] + du, nv = up[q] + dv;
    u[q] = nu;
    up[q] = nv;
    dip_put(rec, mixed, q / 3, q % 3, nu, nv);
    v[0] = du * du;
    v[1] = dv * dv;
  }
  const int out[2] = {S_EPS, S_EPSP};
  block_partials<2>(v, part, out, stride);
}

// u = u_direct (and the packed records), the start of an SOR solve
__global__ void k_reset_dipoles(int n, const double *__restrict__ ud, const double *__restrict__ upd,
                                double *__restrict__ u, double *__restrict__ up, void *rec, int mixed) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  u[q] = ud[q];
  up[q] = upd[q];
  dip_put(rec, mixed, q / 3, q % 3, ud[q], upd[q]);
}

// ---------------------------------------------------------------------------
// reciprocal energy, gradient and torque per atom
// ---------------------------------------------------------------------------
__global__ void k_recip_energy(int n, const double *__restrict__ gmp, const double *__restrict__ fmp,
                               const double *__restrict__ phip, const double *__restrict__ phiu,
                               const double *__restrict__ phiup, const double *__restrict__ u,
                               const double *__restrict__ up, Frac fr, double el, double aw, int polar, int mutual,
                               double *__restrict__ grad, double *__restrict__ torque, double *__restrict__ energy,
                               double *__restrict__ vpart, ElecPhase phs, double *__restrict__ lgrad,
                               double *__restrict__ ltorque) {
  // permanent terms are scaled by phs.perm, polarization terms by phs.pol; with
  // lgrad set, their lambda derivatives go to lgrad / ltorque and E_DUDL
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double eself = 0, erec = 0, qt = 0, eselfi = 0, ereci = 0;
  double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (s < n) {
    RecipAtom o;
    recip_atom(s, gmp, fmp, phip, phiu, phiup, u, up, fr, el, aw, polar, mutual, phs.perm, phs.pol,
               vpart ? w9 : nullptr, o);
    eself = o.eself; erec = o.erec; qt = o.qt; eselfi = o.eselfi; ereci = o.ereci;
    for (int a = 0; a < 3; a++) {
      grad[3 * s + a] = phs.perm * o.df[a] + phs.pol * o.dq[a];
      torque[3 * s + a] = phs.perm * o.tq[a] + phs.pol * o.tqq[a];
    }
    if (lgrad)
      for (int a = 0; a < 3; a++) {
        lgrad[3 * s + a] = phs.dperm * o.df[a] + phs.dpol * o.dq[a];
        ltorque[3 * s + a] = phs.dperm * o.tq[a] + phs.dpol * o.tqq[a];
      }
  }
  if (vpart) block_store9(w9, vpart);
  block_add(phs.perm * eself, energy + E_PERM_SELF);
  block_add(phs.perm * 0.5 * el * erec, energy + E_PERM_RECIP);
  block_add(qt, energy + E_QTOT);
  block_add(phs.pol * eselfi, energy + E_POL_SELF);
  block_add(phs.pol * 0.5 * el * ereci, energy + E_POL_RECIP);
  if (lgrad) {
    double ep = eself + 0.5 * el * erec, eq = eselfi + 0.5 * el * ereci;
    block_add(phs.dperm * ep + phs.dpol * eq, energy + E_DUDL);
    block_add(phs.d2perm * ep + phs.d2pol * eq, energy + E_D2UDL2);
  }
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
                           const double *__restrict__ grad_r, const double *__restrict__ torque_r, int recip,
                           const double *__restrict__ gred_s, int do_vdw, int do_pme, double *__restrict__ grad,
                           int32_t *__restrict__ flags, double *__restrict__ vpart, const double *__restrict__ gred_l,
                           double *__restrict__ dudx, const double *__restrict__ lgrad_s,
                           const double *__restrict__ ltorque_s, const double *__restrict__ lgrad_r,
                           const double *__restrict__ ltorque_r) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (s < n) {
    int i = perm[s];
    V3<double> g = do_pme ? ld3d(grad_s, s) : V3<double>{0, 0, 0};
    if (do_pme && recip) g += ld3d(grad_r, s);
    if (do_vdw) {
      V3<double> gv = ld3d(gred_s, s);
      int h = redidx[i];
      double f = redval[i];
      bool reduced = h >= 0 && h != i && f > 0.0;
      if (reduced) {
        g += f * gv;
        aadd(grad, h, (1.0 - f) * gv);
      } else {
        g += gv;
      }
      if (dudx) {
        // the lambda gradient follows the same reduced-site chain rule
        V3<double> gl = ld3d(gred_l, s);
        if (reduced) {
          aadd(dudx, i, f * gl);
          aadd(dudx, h, (1.0 - f) * gl);
        } else {
          aadd(dudx, i, gl);
        }
      }
    }
    aadd(grad, i, g);
    int fr = frame[i];
    if (do_pme && fr != NB_FRAME_NONE) {
      V3<double> trq = ld3d(torque_s, s);
      if (recip) trq += ld3d(torque_r, s);
      if (!isfinite(trq.x) || !isfinite(trq.y) || !isfinite(trq.z)) atomicOr(&flags[4], 1);
      int iz = axis[3 * i], ix = axis[3 * i + 1], iy = axis[3 * i + 2];
      V3<double> gz, gx, gy, vz, vx, vy;
      torque_frame(fr, i, iz, ix, iy, xyz, trq, gz, gx, gy, vz, vx, vy);
      aadd(grad, iz, gz);
      if (ix >= 0 && fr != NB_FRAME_ZONLY) aadd(grad, ix, gx);
      if (iy >= 0 && (fr == NB_FRAME_ZTHENBISECTOR || fr == NB_FRAME_THREEFOLD)) aadd(grad, iy, gy);
      aadd(grad, i, -(gz + gx + gy));
      if (vpart) {
        // frame rotation under strain: frame-atom gradients against the frame vectors
        double G3[3][3] = {{gz.x, gz.y, gz.z}, {gx.x, gx.y, gx.z}, {gy.x, gy.y, gy.z}};
        double V3a[3][3] = {{vz.x, vz.y, vz.z}, {vx.x, vx.y, vx.z}, {vy.x, vy.y, vy.z}};
        for (int j = 0; j < 3; j++)
          for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) w9[3 * a + b] += G3[j][a] * V3a[j][b];
      }
    }
    if (dudx && lgrad_s) {
      // electrostatic lambda gradient and torque (OST), same frame rules
      V3<double> gl = ld3d(lgrad_s, s);
      if (recip) gl += ld3d(lgrad_r, s);
      aadd(dudx, i, gl);
      if (fr != NB_FRAME_NONE) {
        V3<double> lt = ld3d(ltorque_s, s);
        if (recip) lt += ld3d(ltorque_r, s);
        int iz = axis[3 * i], ix = axis[3 * i + 1], iy = axis[3 * i + 2];
        V3<double> gz, gx, gy, vz, vx, vy;
        torque_frame(fr, i, iz, ix, iy, xyz, lt, gz, gx, gy, vz, vx, vy);
        aadd(dudx, iz, gz);
        if (ix >= 0 && fr != NB_FRAME_ZONLY) aadd(dudx, ix, gx);
        if (iy >= 0 && (fr == NB_FRAME_ZTHENBISECTOR || fr == NB_FRAME_THREEFOLD)) aadd(dudx, iy, gy);
        aadd(dudx, i, -(gz + gx + gy));
      }
    }
  }
  if (vpart) block_store9(w9, vpart);
}

// Sum of rows of 9 partial sums: block g sums rows [g chunk, (g + 1) chunk)
// of `part` into out[9 g ..].  252 threads (28 x 9) read the rows as one
// coalesced stream, thread t always meeting component t % 9; the 28 partial
// sums of each component are added in thread order.  A first pass over many
// blocks and a second over its block sums give a fixed order at any size.
constexpr int VIR_THREADS = 252;
constexpr long VIR_MAX_BLOCKS = NB_VIR_STAGE;
__global__ void __launch_bounds__(VIR_THREADS) k_vir_reduce(long nrows, long chunk, const double *__restrict__ part,
                                                            double *__restrict__ out) {
  __shared__ double red[VIR_THREADS];
  const long r0 = (long)blockIdx.x * chunk, r1 = r0 + chunk < nrows ? r0 + chunk : nrows;
  const long nv = r1 > r0 ? 9 * (r1 - r0) : 0;
  const double *p = part + 9 * r0;
  double t = 0.0;
  for (long v = threadIdx.x; v < nv; v += VIR_THREADS) t += p[v];
  red[threadIdx.x] = t;
  __syncthreads();
  if (threadIdx.x < 9) {
    double s = 0.0;
    for (int k = threadIdx.x; k < VIR_THREADS; k += 9) s += red[k];
    out[9 * (size_t)blockIdx.x + threadIdx.x] = s;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

static int mixed_of(const NbHandle *h) { return h->prm.precision == NB_PREC_MIXED ? 1 : 0; }

void nb_gather_sorted(NbHandle *h) {
  NB_LAUNCH(k_gather_pos, grid_for(h->n, 256), dim3(256), 0, h->n, h->d_perm, h->d_xyz, h->d_mp, h->d_pos4,
            (void *)h->d_dip, (void *)h->d_dip2, mixed_of(h), h->cell, (const int32_t *)h->d_wrap);
}

// one launch clears the accumulated outputs of an evaluation, with (when
// given) the status flags 1-5 (list check, row overflows, undefined torque)
// and the per-brick atom counts of the PME binning
__global__ void k_zero_outputs(int m, double *__restrict__ a, double *__restrict__ b, double *__restrict__ c,
                               double *__restrict__ d, double *__restrict__ energy, int64_t *__restrict__ count,
                               double *__restrict__ e, double *__restrict__ f, int32_t *__restrict__ flags,
                               int32_t *__restrict__ bcount, int nbricks) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) { a[i] = 0.0; b[i] = 0.0; c[i] = 0.0; d[i] = 0.0; e[i] = 0.0; f[i] = 0.0; }
  if (i < 16) energy[i] = 0.0;
  if (i < 4) count[i] = 0;
  if (flags && i >= 1 && i <= 5) flags[i] = 0;
  if (bcount)
    for (int k = i; k < nbricks; k += gridDim.x * blockDim.x) bcount[k] = 0;
}

// The evaluation's host-visible outputs in the pinned block's layout (P_*
// offsets): energies, counts and flags always; the virial, the PCG scalars
// and the valence energies when given (their slots are zero otherwise).
// Words are moved as 64-bit patterns, so counts and flags keep their bits.
__global__ void k_pack_outputs(const double *__restrict__ energy, const int64_t *__restrict__ count,
                               const int32_t *__restrict__ flags, const double *__restrict__ vir,
                               const double *__restrict__ scal, const double *__restrict__ val,
                               double *__restrict__ pin) {
  const int t = threadIdx.x;
  if (t >= P_NSLOT) return;
  const unsigned long long *src = nullptr;
  int k = 0;
  if (t < P_COUNT) { src = (const unsigned long long *)energy; k = t; }
  else if (t < P_FLAGS) { src = (const unsigned long long *)count; k = t - P_COUNT; }
  else if (t < P_VIR) { src = (const unsigned long long *)flags; k = t - P_FLAGS; }
  else if (t < P_VIR + 9) { src = (const unsigned long long *)vir; k = t - P_VIR; }
  else if (t >= P_DUDL && t < P_DUDL + 2) { src = (const unsigned long long *)energy; k = 8 + t - P_DUDL; }
  else if (t >= P_SCAL && t < P_SCAL + S_NSLOT) { src = (const unsigned long long *)scal; k = t - P_SCAL; }
  else if (t >= P_VALENCE && t < P_VALENCE + 8) { src = (const unsigned long long *)val; k = t - P_VALENCE; }
  ((unsigned long long *)pin)[t] = src ? src[k] : 0ull;
}

void nb_pack_outputs(NbHandle *h, bool scal) {
  NB_LAUNCH(k_pack_outputs, dim3(1), dim3(128), 0, h->d_energy, h->d_count, h->d_flags,
            h->want_vir ? h->d_vir : (const double *)nullptr, scal ? h->d_scal : (const double *)nullptr,
            h->has_valence ? h->d_val_energy : (const double *)nullptr, h->d_pin);
}

void nb_zero_outputs(NbHandle *h) {
  int n3 = 3 * h->n;
  const bool bins = h->prm.do_pme && !h->cell.aperiodic && h->prm.ewald_alpha > 0.0 && h->d_bcount;
  NB_LAUNCH(k_zero_outputs, grid_for(n3 > 16 ? n3 : 16, 256), dim3(256), 0, n3, h->d_grad_s, h->d_torque_s,
            h->d_gred_s, h->d_grad, h->d_energy, h->d_count, h->d_gred_l, h->d_dudx, h->d_flags,
            bins ? h->d_bcount : (int32_t *)nullptr, bins ? h->nbricks : 0);
  if (h->ost)
    NB_LAUNCH(k_zero_outputs, grid_for(n3 > 16 ? n3 : 16, 256), dim3(256), 0, n3, h->d_lgrad_s, h->d_ltorque_s,
              h->d_lgrad_r, h->d_ltorque_r, h->d_energy, h->d_count, h->d_lgrad_s, h->d_ltorque_s,
              (int32_t *)nullptr, (int32_t *)nullptr, 0);
}

void nb_direct(NbHandle *h) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  double a3 = 4.0 / 3.0 * aw * aw * aw / NB_SQRT_PI;
  NB_LAUNCH(k_direct, grid_for(n, 256), dim3(256), 0, n, h->d_gmp, h->d_phip, a3, (int)use_recip,
                     h->d_polar, h->d_field, h->d_fieldp, h->d_u, h->d_up, h->d_udir, h->d_updir, (void *)h->d_dip,
                     mixed_of(h));
}

void nb_dip_field_real(NbHandle *h, double *f, double *fp, bool precond);

// total induced field: real space + reciprocal + self (ParticleMeshEwald.
// computeInduceDipoleField, :763-804).  The real-space part reads the packed
// records, which must hold (u, up); full = false gathers only the field,
// full = true also fills the 20 potential derivatives.
void nb_induced_field(NbHandle *h, const double *u, const double *up, double *f, double *fp, bool full) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  // real space on the side stream, reciprocal space on the main stream
  nb_fork(h);
  {
    NbSide side(h);
    nb_dip_field_real(h, f, fp, false);
  }
  if (use_recip) nb_pme_induced(h, u, up, full);
  nb_join(h);
  if (use_recip) {
    double a3 = 4.0 / 3.0 * aw * aw * aw / NB_SQRT_PI;
    NB_LAUNCH(k_ind_reduce, grid_for(n, 256), dim3(256), 0, n, a3, u, up, h->d_phiu, h->d_phiup, f, fp);
  }
}

// preconditioner field of polar * r (the packed records must hold it)
void nb_precond_field(NbHandle *h, double *f, double *fp) { nb_dip_field_real(h, f, fp, true); }

int nb_precond_z(NbHandle *h, int neps, bool control);
int nb_pme_pcg_direction(NbHandle *h, double a3, int nrz, int first);

// the totals of two rows of partial sums into scal (one block, the fixed
// order of block_totals, so the consumers see the same values either way)
__global__ void __launch_bounds__(256) k_part_totals(const double *__restrict__ part, int stride, int nb, int in0,
                                                     int in1, double *__restrict__ scal, int out0, int out1,
                                                     const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  const int in[2] = {in0, in1};
  double t[2];
  block_totals<2>(part, in, stride, nb, t);
  if (threadIdx.x == 0) {
    scal[out0] = t[0];
    scal[out1] = t[1];
  }
}

// nparts partial sums of rows in0 / in1 for the next consumers: passed on as
// they are (the count), or collapsed into scal[out0 / out1] on h->cur when
// there are more than NB_TOTALS_MIN (-1)
int nb_collapse(NbHandle *h, int nparts, int in0, int in1, int out0, int out1) {
  if (nparts <= NB_TOTALS_MIN) return nparts;
  NB_LAUNCH(k_part_totals, dim3(1), dim3(256), 0, (const double *)h->d_part, h->part_stride, nparts, in0, in1,
            h->d_scal, out0, out1, h->ctl);
  return -1;
}

// PCG iterations (PCGSolver.scfByPCG, PCGSolver.java:487-604) from r, z (and
// the records of z) already in place; nrz: partial sums of the current r.z.
//   E(z): real space (side stream) || spread + transforms (main stream),
//   joined before the gather, which forms p = z + beta p and q = A p
//   (A z + beta q)                                         k_gather_ap
//   alpha = r.z / p.q; r -= alpha q; u += alpha p         k_pcg_update
//   z = polar (E_pre(polar r) + scale r), residual norms  k_precond_z
// control: `iters` iterations are queued and the convergence test runs in
// k_precond_z (a recorded solve); otherwise the host reads the norms back
// each iteration.  iter / eps receive the count and the last eps.
static int pcg_iterations(NbHandle *h, bool control, int iters, int nrz, int &iter, double &eps) {
  int n = h->n;
  dim3 B(256), G = grid_for(3 * n, 256);
  const int nb = (int)G.x;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  double a3 = 4.0 / 3.0 * aw * aw * aw / NB_SQRT_PI;
  double *f = h->d_field, *fp = h->d_fieldp;
  int maxit = h->prm.max_scf_iter > 0 ? h->prm.max_scf_iter : 1000;
  int mx = mixed_of(h);
  double prev;
  eps = 100.0;
  iter = 0;
  for (int it = 0; control ? it < iters : true; it++) {
    nb_fork(h);
    {
      NbSide side(h);
      // r.z totals for the conjugation and the update, beside the spread
      nrz = nb_collapse(h, nrz, S_RZN, S_RZPN, S_RZT, S_RZTP);
      nb_dip_field_real(h, f, fp, false);
    }
    int npq;
    if (use_recip) {
      npq = nb_pme_pcg_direction(h, a3, nrz, it == 0);
    } else {
      nb_join(h);
      NB_LAUNCH(k_pcg_dir, G, B, 0, n, h->d_polar, h->d_z, h->d_zp, f, fp, h->d_p, h->d_pp, h->d_vec, h->d_vecp,
                h->d_part, h->part_stride, nrz, h->d_scal, (int)(it == 0), h->ctl);
      npq = nb;
    }
    npq = nb_collapse(h, npq, S_PAP, S_PAPP, S_PQT, S_PQTP);
    if (h->tape_on) nb_tape_push(h, nrz, npq);
    NB_LAUNCH(k_pcg_update, G, B, 0, n, h->d_polar, h->d_vec, h->d_vecp, h->d_p, h->d_pp, h->d_r, h->d_rp, h->d_u,
              h->d_up, h->d_part, h->part_stride, npq, nrz, (void *)h->d_dip2, mx, h->d_scal, h->ctl);
    // the next z and the norms of this update; when the test ends the loop,
    // this z, its records and partial sums are not read after the solve
    nrz = nb_precond_z(h, nb, control);
    if (control) continue;
    HIPCHECK(hipMemcpyAsync(h->h_pinned, h->d_scal, sizeof(double) * S_NSLOT, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipEventRecord(h->ev_eps, h->stream));
    HIPCHECK(hipEventSynchronize(h->ev_eps));
    prev = eps;
    double e = h->h_pinned[S_EPS] > h->h_pinned[S_EPSP] ? h->h_pinned[S_EPS] : h->h_pinned[S_EPSP];
    iter++;
    eps = 4.80321 * sqrt(e / n);
    if (!(eps <= prev)) return NB_ERR_SCF_DIVERGED;
    if (iter >= maxit) return NB_ERR_SCF_MAXITER;
    if (eps < h->prm.polar_eps) return NB_OK;
  }
  return NB_OK;
}

// FP64 residual of the current dipoles and their FP64 reciprocal potentials
// (all derivatives, which are then the final potentials): FP64 records and
// real-space field (side stream) || FP64 spread, transforms and gather (main
// stream), then k_refine_residual; partial r.r are left in S_EPS / S_EPSP.
static void refine_residual(NbHandle *h) {
  int n = h->n;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double aw = use_recip ? h->prm.ewald_alpha : 0.0;
  double a3 = 4.0 / 3.0 * aw * aw * aw / NB_SQRT_PI;
  NB_LAUNCH(k_rec64, grid_for(n, 256), dim3(256), 0, n, h->d_pos4, h->d_mp, h->d_u, h->d_up,
            (DipRec<double> *)h->d_dip64, h->ctl);
  nb_fork(h);
  {
    NbSide side(h);
    nb_dip_field64(h, h->d_field, h->d_fieldp);
  }
  if (use_recip) nb_pme_induced64(h, h->d_u, h->d_up);
  nb_join(h);
  NB_LAUNCH(k_refine_residual, grid_for(3 * n, 256), dim3(256), 0, n, h->d_polar, a3, h->d_udir, h->d_updir, h->d_u,
            h->d_up, h->d_field, h->d_fieldp, use_recip ? h->d_phiu : (const double *)nullptr, h->d_phiup, h->d_r,
            h->d_rp, (void *)h->d_dip2, mixed_of(h), h->d_part, h->part_stride, h->ctl);
}

constexpr int NB_REFINE_MAX = 3;   // corrections of one solve

// One induced-dipole solve from the dipoles u0 held in d_u / d_up and the
// records d_dip: r0 = (u_direct - u0) / polar + E(u0) (k_pcg_init_r), z0, then
// the iterations.  In the mixed form with refinement (h->refine) the FP32
// operator's solve is followed by FP64 iterative refinement: the residual is
// formed with the FP64 operator and its eps (PCGSolver.java:543-547) must be
// below polar_eps; otherwise the solve is restarted from the current dipoles
// with r = the FP64 residual (FP32 corrections), up to NB_REFINE_MAX times.
// The FP64 residual's reciprocal potentials are the final ones.
static int pcg_solve(NbHandle *h, nb_energies *out, bool control, int iters) {
  int n = h->n;
  dim3 B(256), G = grid_for(3 * n, 256);
  const int nb = (int)G.x;
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  int mx = mixed_of(h);
  nb_induced_field(h, h->d_u, h->d_up, h->d_field, h->d_fieldp, false);
  NB_LAUNCH(k_pcg_init_r, G, B, 0, n, h->d_polar, h->d_udir, h->d_updir, h->d_u, h->d_up, h->d_field, h->d_fieldp,
            h->d_r, h->d_rp, (void *)h->d_dip2, mx);
  if (control) {
    NB_LAUNCH(k_pcg_ctl_init, dim3(1), dim3(64), 0, h->d_scal);
    h->ctl = h->d_scal + S_DONE;
  }
  int nrz = nb_precond_z(h, 0, control);
  int iter = 0;
  double eps = 100.0;
  int status = pcg_iterations(h, control, iters, nrz, iter, eps);
  h->ctl = nullptr;
  h->refine_cycles = 0;
  const bool refine = h->refine && mx && !h->tape_on;
  if (refine && control) {
    // the corrections the recorded evaluation took, each with a budget of its
    // iteration count + 1; once a residual passes, the remaining corrections
    // and residuals return at once (S_RDONE); a last residual that fails
    // leaves the solve unfinished, and the evaluation is repeated on the
    // regular path
    for (int c = 0; c < h->graph_refine_n; c++) {
      h->ctl = h->d_scal + S_RDONE;
      refine_residual(h);
      NB_LAUNCH(k_refine_check, dim3(1), B, 0, h->d_part, h->part_stride, nb, h->d_scal, n, h->prm.polar_eps, 1);
      h->ctl = h->d_scal + S_DONE;
      int nrz2 = nb_precond_z(h, 0, true);
      int it2 = 0;
      double eps2 = 100.0;
      pcg_iterations(h, true, h->graph_refine_iters[c], nrz2, it2, eps2);
    }
    h->ctl = h->d_scal + S_RDONE;
    refine_residual(h);
    h->ctl = nullptr;
    NB_LAUNCH(k_refine_check, dim3(1), B, 0, h->d_part, h->part_stride, nb, h->d_scal, n, h->prm.polar_eps, 0);
    return NB_OK;
  }
  if (refine && status == NB_OK) {
    for (int cycle = 0;; cycle++) {
      refine_residual(h);
      NB_LAUNCH(k_pcg_eps, dim3(1), B, 0, h->d_part, h->part_stride, nb, h->d_scal);
      HIPCHECK(hipMemcpyAsync(h->h_pinned, h->d_scal, sizeof(double) * S_NSLOT, hipMemcpyDeviceToHost, h->stream));
      HIPCHECK(hipStreamSynchronize(h->stream));
      double e = h->h_pinned[S_EPS] > h->h_pinned[S_EPSP] ? h->h_pinned[S_EPS] : h->h_pinned[S_EPSP];
      eps = 4.80321 * sqrt(e / n);
      if (eps < h->prm.polar_eps || cycle == NB_REFINE_MAX) break;
      h->refine_cycles = cycle + 1;
      // restart from the current dipoles with the FP64 residual (its records
      // of polar * r were written by k_refine_residual)
      nrz = nb_precond_z(h, 0, false);
      int it2 = 0;
      double eps2 = 100.0;
      status = pcg_iterations(h, false, 0, nrz, it2, eps2);
      h->refine_iters[cycle] = it2;
      iter += it2;
      if (status != NB_OK) break;
    }
  }
  if (out) {
    out->pcg_iters = iter;
    out->pcg_eps = eps;
  }
  // reciprocal potentials of the final dipoles (20 derivatives each) for the
  // energy and gradient, unless the FP64 residual of the refinement gave them
  if (use_recip && !(refine && status == NB_OK)) nb_pme_induced(h, h->d_u, h->d_up, true);
  return status;
}

int nb_pcg(NbHandle *h, nb_energies *out) { return pcg_solve(h, out, false, 0); }

// The same solve with its control on the device, for recording into a graph:
// a fixed budget of iterations is queued; once k_pcg_update has stopped the
// solve, the update and the field, spread, gather and z kernels of the
// remaining iterations return at once (the transforms still run on stale data
// that nothing reads).  The result equals that of nb_pcg.
int nb_pcg_device(NbHandle *h, int iters) { return pcg_solve(h, nullptr, true, iters); }

void nb_reset_dipoles(NbHandle *h) {
  int n = h->n;
  NB_LAUNCH(k_reset_dipoles, grid_for(3 * n, 256), dim3(256), 0, n, h->d_udir, h->d_updir, h->d_u, h->d_up,
            (void *)h->d_dip, mixed_of(h));
}

// SOR solve (ParticleMeshEwald.scfBySOR, ParticleMeshEwald.java:2331-2423) from
// the current dipoles: eps = 4.80321 sqrt(max(sum du^2, sum duCR^2) / n); a rise
// fails, 1000 cycles are the limit, eps < polar_eps converges.
int nb_sor(NbHandle *h, double w, nb_energies *out) {
  int n = h->n;
  dim3 B(256), G = grid_for(3 * n, 256);
  const int nb = (int)G.x, stride = h->part_stride;
  if (nb > stride) throw std::runtime_error("PCG partial-sum buffer too small");
  bool use_recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
  double *f = h->d_field, *fp = h->d_fieldp;
  void *rec = (void *)h->d_dip;
  int mx = mixed_of(h);
  double eps = 100.0, prev;
  int iter = 0, status = NB_OK;
  for (;;) {
    nb_induced_field(h, h->d_u, h->d_up, f, fp, false);
    NB_LAUNCH(k_sor_update, G, B, 0, n, w, h->d_polar, h->d_udir, h->d_updir, f, fp, h->d_u, h->d_up, h->d_part,
              stride, rec, mx);
    NB_LAUNCH(k_pcg_eps, dim3(1), B, 0, h->d_part, stride, nb, h->d_scal);
    HIPCHECK(hipMemcpyAsync(h->h_pinned, h->d_scal, sizeof(double) * S_NSLOT, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    prev = eps;
    double e = h->h_pinned[S_EPS] > h->h_pinned[S_EPSP] ? h->h_pinned[S_EPS] : h->h_pinned[S_EPSP];
    iter++;
    eps = 4.80321 * sqrt(e / n);
    if (eps > prev) { status = NB_ERR_SCF_DIVERGED; break; }
    if (iter >= 1000) { status = NB_ERR_SCF_MAXITER; break; }
    if (eps < h->prm.polar_eps) break;
  }
  out->pcg_iters = iter;
  out->pcg_eps = eps;
  if (use_recip) nb_pme_induced(h, h->d_u, h->d_up, true);
  return status;
}

void nb_recip_energy(NbHandle *h, bool polar) {
  int n = h->n;
  Frac fr;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) fr.a[a][b] = h->afrac[a][b];
  NB_LAUNCH(k_recip_energy, grid_for(n, 128), dim3(128), 0, n, h->d_gmp, h->d_fmp, h->d_phip,
                     h->d_phiu, h->d_phiup, h->d_u, h->d_up, fr, h->prm.electric, h->prm.ewald_alpha, (int)polar,
                     (int)(h->prm.polarization == NB_POL_MUTUAL && !h->finite_active), h->d_grad_r, h->d_torque_r,
                     h->d_energy,
                     h->want_vir ? nb_vir_region(h, grid_for(n, 128).x) : (double *)nullptr, h->ph,
                     h->ost ? h->d_lgrad_r : (double *)nullptr, h->d_ltorque_r);
}

void nb_finalize_gradient(NbHandle *h) {
  int n = h->n;
  int recip = (!h->cell.aperiodic && h->prm.ewald_alpha > 0.0) ? 1 : 0;
  NB_LAUNCH(k_finalize, grid_for(n, 128), dim3(128), 0, n, h->d_perm, h->d_xyz, h->d_frame,
                     h->d_axis, h->d_redidx, h->d_redval, h->d_grad_s, h->d_torque_s, h->d_grad_r, h->d_torque_r,
                     recip, h->d_gred_s, (int)h->prm.do_vdw, (int)h->prm.do_pme, h->d_grad, h->d_flags,
                     h->want_vir ? nb_vir_region(h, grid_for(n, 128).x) : (double *)nullptr, h->d_gred_l,
                     h->want_dudl ? h->d_dudx : (double *)nullptr,
                     (h->ost && h->prm.do_pme) ? h->d_lgrad_s : (const double *)nullptr, h->d_ltorque_s, h->d_lgrad_r,
                     h->d_ltorque_r);
}

// next free region of nblocks x 9 partial sums in d_vpart (host bookkeeping;
// the regions are laid out in launch order and reset every evaluation)
double *nb_vir_region(NbHandle *h, long nblocks) {
  if (h->vpart_used + nblocks > h->vpart_cap - NB_VIR_STAGE)
    throw std::runtime_error("virial partial-sum buffer too small");
  double *p = h->d_vpart + 9 * h->vpart_used;
  h->vpart_used += nblocks;
  return p;
}

// the evaluation's virial from every region's partial sums: many blocks,
// then one (the stage buffer is the last VIR_MAX_BLOCKS rows of d_vpart)
void nb_vir_reduce(NbHandle *h) {
  const long used = h->vpart_used;
  long chunk = (used + VIR_MAX_BLOCKS - 1) / VIR_MAX_BLOCKS;
  if (chunk < 64) chunk = 64;
  const long g1 = (used + chunk - 1) / chunk;
  if (used + VIR_MAX_BLOCKS > h->vpart_cap) throw std::runtime_error("virial partial-sum buffer too small");
  double *stage = h->d_vpart + 9 * (h->vpart_cap - VIR_MAX_BLOCKS);
  NB_LAUNCH(k_vir_reduce, dim3(g1 > 0 ? g1 : 1), dim3(VIR_THREADS), 0, used, chunk, (const double *)h->d_vpart,
            stage);
  NB_LAUNCH(k_vir_reduce, dim3(1), dim3(VIR_THREADS), 0, g1 > 0 ? g1 : 1, g1 > 0 ? g1 : 1, (const double *)stage,
            h->d_vir);
}
