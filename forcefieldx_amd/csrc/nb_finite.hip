// nb_finite.hip -- gradient of the truncated PCG solve (NB_FINITE_PCG).
//
// The reference's gradient (RealSpaceEnergyRegion.java:1700-2177,
// ReciprocalEnergyRegion.java:545-733) assumes stationary dipoles: it pairs
// u with uCR = A^-1 F_p.  A solve stopped at a finite tolerance returns u_K,
// the K-th PCG iterate (PCGSolver.java:487-604), whose energy
// E = -(el/2) u_K . F_p has the exact gradient
//   dE = DirectGrad(u_K, Fhat) - (el/2) [sum_k -Qbar_k . dT p_k
//        + rbar_0 . dT u_0 + sum_k Zbar_k . dM r_k]
// with the iteration count held fixed.  DirectGrad is the gradient of direct
// polarization (u . dF_p + Fhat . dF_d, no mutual term) with uCR replaced by
// Fhat; Fhat, Qbar, rbar, Zbar come from the reverse sweep through the
// recorded iterations seeded with g = F_p; T is the Ewald dipole-field
// operator (A = 1/alpha - T) and M = alpha (T_pre alpha + s) the
// preconditioner.
//
//  * nb_tape_push: the iteration's r, z, p, q = A p and r.z, p.q (u system).
//  * nb_finite_adjoint: the reverse sweep (operators applied with the solve's
//    own field kernels), Fhat into d_up, and the tensor-derivative terms into
//    d_gfin: k_bilinear over the Ewald / preconditioner rows and, per T pair,
//    the reciprocal term from the PME potentials of the pair.
//  * With the virial, the same terms' strain derivatives: the pair part
//    sum (dW/dR)_a R_b of every row entry (k_bilinear<VIR>), and per T pair
//    (a, b) the reciprocal mutual-term strain derivative with (u, uCR) ->
//    (a, b): the atom part -(el/2)(b_b dphi_a(a) + a_b dphi_a(b))
//    (k_recip_bilinear_vir) and the grid part from the transforms of a and b
//    (nb_grid_virial_pair).  DirectGrad's virial comes from the energy
//    kernels in their direct form, with uCR replaced by Fhat.
#include <cmath>
#include <stdexcept>
#include <vector>

#include "nb_kern.h"
#include "nb_atom.h"

void nb_dip_field_real(NbHandle *h, double *f, double *fp, bool precond);

namespace {

constexpr int FIN_LPA = 16;

// totals of the r.z and p.q partial sums of an iteration (slots S_RZN, S_PAP),
// or the collapsed totals in scal when a count is negative
__global__ void k_tape_scal(const double *__restrict__ part, int stride, int nrz, int npq,
                            const double *__restrict__ scal, double *__restrict__ out) {
  __shared__ double red[2][4];
  double v[2] = {0.0, 0.0};
  for (int i = threadIdx.x; i < nrz; i += blockDim.x) v[0] += part[(size_t)S_RZN * stride + i];
  for (int i = threadIdx.x; i < npq; i += blockDim.x) v[1] += part[(size_t)S_PAP * stride + i];
  for (int k = 0; k < 2; k++) v[k] = wave_sum_d(v[k]);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { red[0][w] = v[0]; red[1][w] = v[1]; }
  __syncthreads();
  if (threadIdx.x < 2) {
    const bool collapsed = threadIdx.x == 0 ? nrz < 0 : npq < 0;
    out[threadIdx.x] = collapsed ? scal[threadIdx.x == 0 ? S_RZT : S_PQT]
                                 : red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] +
                                       red[threadIdx.x][3];
  }
}

// out = c0 x0 + c1 x1 + c2 x2 (null terms skipped; out may be x0)
__global__ void k_lin3(int m, double *out, double c0, const double *x0, double c1, const double *x1, double c2,
                       const double *x2) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  double v = 0.0;
  if (x0) v += c0 * x0[q];
  if (x1) v += c1 * x1[q];
  if (x2) v += c2 * x2[q];
  out[q] = v;
}

// per-block partial sums of up to three dot products -> part[3][nblk]
__global__ void __launch_bounds__(256) k_dots3(int m, const double *x0, const double *y0, const double *x1,
                                               const double *y1, const double *x2, const double *y2,
                                               double *__restrict__ part) {
  __shared__ double red[3][4];
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (q < m) {
    if (x0) v[0] = x0[q] * y0[q];
    if (x1) v[1] = x1[q] * y1[q];
    if (x2) v[2] = x2[q] * y2[q];
  }
  for (int k = 0; k < 3; k++) v[k] = wave_sum_d(v[k]);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < 3; k++) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < 3)
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
        red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

// packed records <- (s0 v0, s0 v1), s0 = polar when scaled
__global__ void k_rec_put(int n, const double *__restrict__ v0, const double *__restrict__ v1,
                          const double *__restrict__ polar, int scaled, void *rec, int mixed) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  double a = scaled ? polar[q / 3] : 1.0;
  dip_put(rec, mixed, q / 3, q % 3, a * v0[q], a * (v1 ? v1[q] : 0.0));
}

// T v = real-space field f + self a3 v - reciprocal phi
__global__ void k_T_combine(int n, double a3, const double *__restrict__ v, const double *__restrict__ f,
                            const double *__restrict__ phi, int recip, double *__restrict__ out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  int s = q / 3, d = q - 3 * s;
  double x = f[q];
  if (recip) x += a3 * v[q] - phi[30 * (size_t)s + 21 + d];
  out[q] = x;
}

// M v = alpha (T_pre (alpha v) + psc v), or A v = v / alpha - T v (alpha > 0)
__global__ void k_MA_combine(int n, const double *__restrict__ polar, double psc, const double *__restrict__ v,
                             const double *__restrict__ f, int mode_a, double *__restrict__ out) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  double a = polar[q / 3];
  if (mode_a) out[q] = a > 0.0 ? v[q] / a - f[q] : 0.0;
  else out[q] = a * (f[q] + psc * v[q]);
}

// seed g = F_p = uCR_direct / alpha on the polarizable atoms
__global__ void k_seed(int n, const double *__restrict__ polar, const double *__restrict__ updir,
                       double *__restrict__ g) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 3 * n) return;
  double a = polar[q / 3];
  g[q] = a > 0.0 ? updir[q] / a : 0.0;
}

// Sum over the pairs j of w_j grad(a_j . T b_j) on the i side of every row
// entry, times coef, added to gout.  T is the Ewald dipole-field tensor minus
// its Thole-damped bare part (B1 = bn1 - (1 - s3) rr3, B2 = bn2 - (1 - s5)
// rr5), or (PRE) the Thole-damped bare tensor of the preconditioner with the
// vectors scaled by alpha.  For a pair, W = a_i.T b_k + a_k.T b_i =
// -B1 S1 + B2 S2 with S1 = a_i.b_k + a_k.b_i, S2 = (a_i.R)(b_k.R) +
// (a_k.R)(b_i.R), R = x_k - x_i, and dW/dx_i = -grad_R W.
// VIR: also the strain derivative of the same terms, half of coef G_a R_b per
// visit of a row entry (G the pair's sum of w_j grad_R W_j; the i-side
// gradient of the pair is -coef G and R = x_k - x_i; the pair convention of
// the energy kernels, (dE/dR)_a R_b), 9 partial sums per block
template <int LPA, bool PRE, bool VIR>
__global__ void __launch_bounds__(256) k_bilinear(int n, const double4 *__restrict__ pos, const AtomMP *__restrict__ mp,
                                                  const uint32_t *__restrict__ lst, const int32_t *__restrict__ cnt,
                                                  int maxl, CellD cell, double aewald, const double *__restrict__ A,
                                                  const double *__restrict__ B, const double *__restrict__ w, int J,
                                                  const double *__restrict__ polar, double coef,
                                                  double *__restrict__ gout, double *__restrict__ vpart) {
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt / LPA, lane = gt % LPA;
  bool valid = s < n;
  V3<double> G{0, 0, 0};
  double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const size_t n3 = 3 * (size_t)n;
  if (valid) {
    double4 pi = pos[s];
    AtomMP ai = mp[s];
    double asc = PRE ? polar[s] : 1.0;
    int m = cnt[s];
    const uint32_t *row = lst + (size_t)s * maxl;
    for (int t = lane; t < m; t += LPA) {
      uint32_t e = row[t];
      int k = e & cell.kmask;
      double4 pk = pos[k];
      entry_shift(cell, e, pk.x, pk.y, pk.z);
      double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
      double r2 = image(cell, dx, dy, dz);
      double r = sqrt(r2), rinv = 1.0 / r;
      V3<double> R{dx, dy, dz};
      AtomMP ak = mp[k];
      double bn[4], rr[4];
      ewald_bn<double, 3>(r, rinv, PRE ? 0.0 : aewald, bn);
      coulomb_rr<double, 3>(rinv, rr);
      Thole<double> th = thole<double, true>(r, R, ai.ipd, ak.ipd, ai.thole, ak.thole);
      double B2, C1, C2;
      if (PRE) {
        B2 = th.s5 * rr[2]; C1 = th.s3 * rr[2]; C2 = th.s5 * rr[3];
      } else {
        B2 = bn[2] - (1.0 - th.s5) * rr[2];
        C1 = bn[2] - (1.0 - th.s3) * rr[2];
        C2 = bn[3] - (1.0 - th.s5) * rr[3];
      }
      V3<double> dB1 = rr[1] * th.d3 - C1 * R, dB2 = rr[2] * th.d5 - C2 * R;
      double ksc = PRE ? polar[k] : 1.0;
      V3<double> Gp{0, 0, 0};
      for (int j = 0; j < J; j++) {
        const double *Aj = A + j * n3, *Bj = B + j * n3;
        V3<double> a_i = asc * ld3d(Aj, s), b_i = asc * ld3d(Bj, s), a_k = ksc * ld3d(Aj, k), b_k = ksc * ld3d(Bj, k);
        double S1 = dot(a_i, b_k) + dot(a_k, b_i);
        double air = dot(a_i, R), akr = dot(a_k, R), bir = dot(b_i, R), bkr = dot(b_k, R);
        double S2 = air * bkr + akr * bir;
        V3<double> gW = (-S1) * dB1 + S2 * dB2 + B2 * (bkr * a_i + air * b_k + bir * a_k + akr * b_i);
        Gp += w[j] * gW;
      }
      G += Gp;
      if (VIR) {
        const double r3[3] = {R.x, R.y, R.z}, g3[3] = {Gp.x, Gp.y, Gp.z}, hc = 0.5 * coef;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int b = 0; b < 3; b++) w9[3 * a + b] += hc * g3[a] * r3[b];
      }
    }
  }
  if (VIR) block_store9(w9, vpart);
  group_sum3<LPA>(G);
  if (valid && lane == 0) add3(gout, s, (-coef) * G);
}

// Atom part of the reciprocal strain derivative of a T pair (a, b): the
// mutual term of the reciprocal virial with (u, uCR) -> (a, b),
// -scale (b_b dphi_a(a) + a_b dphi_a(b)) per atom, from the Cartesian
// potential derivatives of the two (phia / phib [30 s + 21 + a]); 9 partial
// sums per block
__global__ void __launch_bounds__(128) k_recip_bilinear_vir(int n, const double *__restrict__ phia,
                                                            const double *__restrict__ phib,
                                                            const double *__restrict__ a, const double *__restrict__ b,
                                                            double scale, double *__restrict__ vpart) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (s < n) {
    const double *ca = phia + 30 * (size_t)s + 21, *cb = phib + 30 * (size_t)s + 21;
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) w9[3 * x + y] = -scale * (b[3 * s + y] * ca[x] + a[3 * s + y] * cb[x]);
  }
  block_store9(w9, vpart);
}

// scale times the sum of nblk blocks of 9 partial sums, in block order, into out[9]
__global__ void k_vir_sum(long nblk, const double *__restrict__ part, double scale, double *__restrict__ out) {
  __shared__ double red[9][64];
  const int q = threadIdx.x / 64, l = threadIdx.x % 64;
  double t = 0.0;
  if (q < 9)
    for (long b = l; b < nblk; b += 64) t += part[9 * b + q];
  if (q < 9) red[q][l] = t;
  __syncthreads();
  if (threadIdx.x < 9) {
    double v = 0.0;
    for (int i = 0; i < 64; i++) v += red[threadIdx.x][i];
    out[threadIdx.x] = scale * v;
  }
}

// Reciprocal part of coef_j a . dT b from the 20 fractional potential
// derivatives of a (phia) and b (phib): the mutual term of the reciprocal
// gradient (ReciprocalEnergyRegion.java:690-733) with (u, uCR) -> (a, b),
// which is the reciprocal part of -(el/2) a . dT b
__global__ void k_recip_bilinear(int n, Frac fr, const double *__restrict__ phia, const double *__restrict__ phib,
                                 const double *__restrict__ a, const double *__restrict__ b, double scale,
                                 double *__restrict__ gout) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double u[3], v[3];
  for (int j = 0; j < 3; j++) {
    u[j] = fr.a[j][0] * a[3 * s] + fr.a[j][1] * a[3 * s + 1] + fr.a[j][2] * a[3 * s + 2];
    v[j] = fr.a[j][0] * b[3 * s] + fr.a[j][1] * b[3 * s + 1] + fr.a[j][2] * b[3 * s + 2];
  }
  const double *fi = phia + 30 * (size_t)s, *fip = phib + 30 * (size_t)s;
  double g[3];
  g[0] = u[0] * fip[4] + v[0] * fi[4] + u[1] * fip[7] + v[1] * fi[7] + u[2] * fip[8] + v[2] * fi[8];
  g[1] = u[0] * fip[7] + v[0] * fi[7] + u[1] * fip[5] + v[1] * fi[5] + u[2] * fip[9] + v[2] * fi[9];
  g[2] = u[0] * fip[8] + v[0] * fi[8] + u[1] * fip[9] + v[1] * fi[9] + u[2] * fip[6] + v[2] * fi[6];
  for (int c = 0; c < 3; c++) gout[3 * s + c] += scale * (fr.a[0][c] * g[0] + fr.a[1][c] * g[1] + fr.a[2][c] * g[2]);
}

// y += c x, and (y2 non-null) y2 += c2 x
__global__ void k_add_vec(int m, const double *__restrict__ x, double c, double *__restrict__ y, double c2,
                          double *__restrict__ y2) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  y[q] += c * x[q];
  if (y2) y2[q] += c2 * x[q];
}

}  // namespace

static int mixed_rec(const NbHandle *h) { return h->prm.precision == NB_PREC_MIXED ? 1 : 0; }

static void ensure(double **p, size_t &cap, size_t need) {
  if (need <= cap) return;
  if (*p) hipFree(*p);
  *p = nullptr;
  HIPCHECK(hipMalloc(p, need * sizeof(double)));
  cap = need;
}

// record iteration k of the u system (called after the conjugation, before
// the update); nrz / npq: partial sums of r.z and p.q
void nb_tape_push(NbHandle *h, int nrz, int npq) {
  size_t n3 = 3 * (size_t)h->n;
  int k = h->tape_K;
  if (k + 1 > h->tape_cap) {
    int cap = h->tape_cap ? 2 * h->tape_cap : 16;
    double *nv = nullptr, *ns = nullptr;
    HIPCHECK(hipMalloc(&nv, sizeof(double) * 4 * n3 * cap));
    HIPCHECK(hipMalloc(&ns, sizeof(double) * 2 * cap));
    if (h->d_tape) {
      HIPCHECK(hipMemcpyAsync(nv, h->d_tape, sizeof(double) * 4 * n3 * h->tape_cap, hipMemcpyDeviceToDevice,
                              h->stream));
      HIPCHECK(hipMemcpyAsync(ns, h->d_tape_sc, sizeof(double) * 2 * h->tape_cap, hipMemcpyDeviceToDevice,
                              h->stream));
      HIPCHECK(hipStreamSynchronize(h->stream));
      hipFree(h->d_tape);
      hipFree(h->d_tape_sc);
    }
    h->d_tape = nv;
    h->d_tape_sc = ns;
    h->tape_cap = cap;
  }
  double *slot = h->d_tape + 4 * n3 * k;
  const double *src[4] = {h->d_r, h->d_z, h->d_p, h->d_vec};
  for (int v = 0; v < 4; v++)
    HIPCHECK(hipMemcpyAsync(slot + v * n3, src[v], sizeof(double) * n3, hipMemcpyDeviceToDevice, h->stream));
  NB_LAUNCH(k_tape_scal, dim3(1), dim3(256), 0, h->d_part, h->part_stride, nrz, npq, (const double *)h->d_scal,
            h->d_tape_sc + 2 * k);
  h->tape_K = k + 1;
}

namespace {

struct Fin {
  NbHandle *h;
  int n, n3;
  dim3 G, B;
  bool recip;
  double a3;
  double *f, *fp;

  explicit Fin(NbHandle *hh) : h(hh) {
    n = h->n;
    n3 = 3 * n;
    B = dim3(256);
    G = grid_for(n3, 256);
    recip = !h->cell.aperiodic && h->prm.ewald_alpha > 0.0;
    double aw = recip ? h->prm.ewald_alpha : 0.0;
    a3 = 4.0 / 3.0 * aw * aw * aw / NB_SQRT_PI;
    f = h->d_field;
    fp = h->d_fieldp;
  }
  // out = T v
  void T(const double *v, double *out) {
    NB_LAUNCH(k_rec_put, G, B, 0, n, v, (const double *)nullptr, (const double *)nullptr, 0, (void *)h->d_dip,
              mixed_rec(h));
    nb_dip_field_real(h, f, fp, false);
    if (recip) nb_pme_induced(h, v, v, false);
    NB_LAUNCH(k_T_combine, G, B, 0, n, a3, v, f, (const double *)h->d_phiu, (int)recip, out);
  }
  // out = M v
  void M(const double *v, double *out) {
    NB_LAUNCH(k_rec_put, G, B, 0, n, v, (const double *)nullptr, (const double *)h->d_polar, 1, (void *)h->d_dip,
              mixed_rec(h));
    nb_dip_field_real(h, f, fp, true);
    NB_LAUNCH(k_MA_combine, G, B, 0, n, (const double *)h->d_polar, h->prm.precond_scale, v, (const double *)f, 0,
              out);
  }
  // out = A v
  void A(const double *v, double *tmp, double *out) {
    T(v, tmp);
    NB_LAUNCH(k_MA_combine, G, B, 0, n, (const double *)h->d_polar, 0.0, v, (const double *)tmp, 1, out);
  }
  void lin(double *out, double c0, const double *x0, double c1 = 0.0, const double *x1 = nullptr,
           double c2 = 0.0, const double *x2 = nullptr) {
    NB_LAUNCH(k_lin3, G, B, 0, n3, out, c0, x0, c1, x1, c2, x2);
  }
  void copy(double *out, const double *x) {
    HIPCHECK(hipMemcpyAsync(out, x, sizeof(double) * n3, hipMemcpyDeviceToDevice, h->stream));
  }
  // up to three dot products, summed on the host in a fixed order
  void dots(double *res, const double *x0, const double *y0, const double *x1 = nullptr,
            const double *y1 = nullptr, const double *x2 = nullptr, const double *y2 = nullptr) {
    int nb = (int)G.x;
    NB_LAUNCH(k_dots3, G, B, 0, n3, x0, y0, x1, y1, x2, y2, h->d_part);
    std::vector<double> part(3 * (size_t)nb);
    HIPCHECK(hipMemcpyAsync(part.data(), h->d_part, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < 3; k++) {
      double s = 0.0;
      for (int b = 0; b < nb; b++) s += part[(size_t)k * nb + b];
      res[k] = s;
    }
  }
};

}  // namespace

// Reverse sweep through the recorded iterations (the oracle's pcg_adjoint in
// device form) and the tensor-derivative gradient terms, into d_gfin; Fhat is
// left in d_up with the reciprocal potentials of (u, Fhat) in d_phiu / d_phiup.
void nb_finite_adjoint(NbHandle *h) {
  Fin F(h);
  const int K = h->tape_K, n3 = F.n3;
  if (K <= 0) throw std::runtime_error("finite-PCG gradient without recorded iterations");
  if ((int)F.G.x * 3 > 16 * h->part_stride) throw std::runtime_error("partial-sum buffer too small");
  std::vector<double> sc(2 * (size_t)K);
  HIPCHECK(hipMemcpyAsync(sc.data(), h->d_tape_sc, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  std::vector<double> gam(K), del(K), a(K), beta(K, 0.0);
  for (int k = 0; k < K; k++) {
    gam[k] = sc[2 * k];
    del[k] = sc[2 * k + 1];
    a[k] = del[k] != 0.0 ? gam[k] / del[k] : gam[k];
  }
  for (int k = 0; k + 1 < K; k++) beta[k] = gam[k + 1] / (gam[k] != 0.0 ? gam[k] : 1.0);
  // work vectors: g, Pn, Pc, R, tmp, tmp2, and the pairs: T (K + 1) and M (K), a and b each
  const int JT = K + 1, JM = K;
  size_t need = (size_t)n3 * (6 + 2 * JT + 2 * JM) + JT + JM;
  ensure(&h->d_fin, h->fin_cap, need);
  double *g = h->d_fin, *Pn = g + n3, *Pc = Pn + n3, *R = Pc + n3, *tmp = R + n3, *tmp2 = tmp + n3;
  double *AT = tmp2 + n3, *BT = AT + (size_t)JT * n3, *AM = BT + (size_t)JT * n3, *BM = AM + (size_t)JM * n3;
  double *wT = BM + (size_t)JM * n3, *wM = wT + JT;
  auto tv = [&](int v, int k) { return h->d_tape + ((size_t)4 * k + v) * n3; };  // 0 r, 1 z, 2 p, 3 q
  NB_LAUNCH(k_seed, F.G, F.B, 0, F.n, (const double *)h->d_polar, (const double *)h->d_updir, g);
  HIPCHECK(hipMemsetAsync(Pn, 0, sizeof(double) * n3, h->stream));
  HIPCHECK(hipMemsetAsync(R, 0, sizeof(double) * n3, h->stream));
  std::vector<double> Gb(K + 1, 0.0), w_t(JT), w_m(JM, 1.0);
  for (int k = K - 1; k >= 0; k--) {
    HIPCHECK(hipMemsetAsync(Pc, 0, sizeof(double) * n3, h->stream));
    if (k < K - 1) {
      // p_{k+1} = z_{k+1} + beta_k p_k;  beta_k = gam_{k+1} / gam_k;
      // gam_{k+1} = r_{k+1} . z_{k+1};  z_{k+1} = M r_{k+1}
      double d3[3];
      F.dots(d3, Pn, tv(2, k));
      double bb = d3[0];
      F.lin(Pc, beta[k], Pn);
      Gb[k + 1] += bb / gam[k];
      Gb[k] -= bb * gam[k + 1] / (gam[k] * gam[k]);
      double *Z = AM + (size_t)(k + 1) * n3;
      F.lin(Z, 1.0, Pn, Gb[k + 1], tv(0, k + 1));
      F.lin(R, 1.0, R, Gb[k + 1], tv(1, k + 1));
      F.M(Z, tmp);
      F.lin(R, 1.0, R, 1.0, tmp);
      F.copy(BM + (size_t)(k + 1) * n3, tv(0, k + 1));
    }
    // r_{k+1} = r_k - a_k q_k;  u_{k+1} = u_k + a_k p_k;  a_k = gam_k / del_k;
    // del_k = p_k . q_k;  q_k = A p_k
    double d3[3];
    F.dots(d3, R, tv(3, k), g, tv(2, k));
    double ab = -d3[0] + d3[1];
    Gb[k] += ab / del[k];
    double db = -ab * gam[k] / (del[k] * del[k]);
    double *Q = AT + (size_t)k * n3;
    F.lin(Q, -a[k], R, db, tv(2, k));
    F.lin(Pc, 1.0, Pc, a[k], g, db, tv(3, k));
    F.A(Q, tmp2, tmp);
    F.lin(Pc, 1.0, Pc, 1.0, tmp);
    F.copy(BT + (size_t)k * n3, tv(2, k));
    w_t[k] = -1.0;
    std::swap(Pn, Pc);
  }
  // gam_0 = r_0 . z_0;  p_0 = z_0;  z_0 = M r_0;  r_0 = T u_0;  u_0 = alpha F_d
  double *Z0 = AM;
  F.lin(Z0, 1.0, Pn, Gb[0], tv(0, 0));
  F.lin(R, 1.0, R, Gb[0], tv(1, 0));
  F.M(Z0, tmp);
  F.lin(R, 1.0, R, 1.0, tmp);
  F.copy(BM, tv(0, 0));
  double *RT = AT + (size_t)K * n3;
  F.copy(RT, R);
  F.copy(BT + (size_t)K * n3, h->d_udir);
  w_t[K] = 1.0;
  F.T(R, tmp);
  F.lin(tmp, 1.0, g, 1.0, tmp);
  HIPCHECK(hipMemcpyAsync(wT, w_t.data(), sizeof(double) * JT, hipMemcpyHostToDevice, h->stream));
  HIPCHECK(hipMemcpyAsync(wM, w_m.data(), sizeof(double) * JM, hipMemcpyHostToDevice, h->stream));
  // Fhat = alpha (g + T rbar_0) into d_up
  NB_LAUNCH(k_MA_combine, F.G, F.B, 0, F.n, (const double *)h->d_polar, 0.0, (const double *)tmp, (const double *)tmp,
            0, h->d_up);
  // tensor-derivative terms, coefficient -(el / 2)
  HIPCHECK(hipMemsetAsync(h->d_gfin, 0, sizeof(double) * n3, h->stream));
  double coef = -0.5 * h->prm.electric;
  dim3 Gp = xcd_grid((long)F.n * FIN_LPA, 256);
  double aw = F.recip ? h->prm.ewald_alpha : 0.0;
  // with the virial: 9 partial sums per block of every launch below in
  // d_vfin, summed at the end into one block of the evaluation's virial
  const bool vir = h->want_vir;
  const long nbil = Gp.x, natom = grid_for(F.n, 128).x, ngrid = F.recip ? nb_grid_virial_blocks(h) : 0;
  const long vneed = vir ? 2 * nbil + (F.recip ? (long)JT * (natom + ngrid) : 0) : 0;
  if (vir && vneed > h->vfin_cap) {
    HIPCHECK(hipStreamSynchronize(h->stream));
    if (h->d_vfin) (void)hipFree(h->d_vfin);
    h->d_vfin = nullptr;
    HIPCHECK(hipMalloc(&h->d_vfin, sizeof(double) * 9 * vneed));
    h->vfin_cap = vneed;
  }
  long vused = 0;
  auto vregion = [&](long nblk) {
    double *p = vir ? h->d_vfin + 9 * vused : nullptr;
    vused += vir ? nblk : 0;
    return p;
  };
  if (vir) {
    NB_LAUNCH((k_bilinear<FIN_LPA, false, true>), Gp, F.B, 0, F.n, h->d_pos4, h->d_mp, h->d_ew, h->d_new, h->maxew,
              h->cell, aw, (const double *)AT, (const double *)BT, (const double *)wT, JT,
              (const double *)h->d_polar, coef, h->d_gfin, vregion(nbil));
    NB_LAUNCH((k_bilinear<FIN_LPA, true, true>), Gp, F.B, 0, F.n, h->d_pos4, h->d_mp, h->d_pc, h->d_npc, h->maxpc,
              h->cell, 0.0, (const double *)AM, (const double *)BM, (const double *)wM, JM,
              (const double *)h->d_polar, coef, h->d_gfin, vregion(nbil));
  } else {
    NB_LAUNCH((k_bilinear<FIN_LPA, false, false>), Gp, F.B, 0, F.n, h->d_pos4, h->d_mp, h->d_ew, h->d_new, h->maxew,
              h->cell, aw, (const double *)AT, (const double *)BT, (const double *)wT, JT,
              (const double *)h->d_polar, coef, h->d_gfin, (double *)nullptr);
    NB_LAUNCH((k_bilinear<FIN_LPA, true, false>), Gp, F.B, 0, F.n, h->d_pos4, h->d_mp, h->d_pc, h->d_npc, h->maxpc,
              h->cell, 0.0, (const double *)AM, (const double *)BM, (const double *)wM, JM,
              (const double *)h->d_polar, coef, h->d_gfin, (double *)nullptr);
  }
  if (F.recip) {
    Frac fr;
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) fr.a[x][y] = h->afrac[x][y];
    for (int j = 0; j < JT; j++) {
      const double *Aj = AT + (size_t)j * n3, *Bj = BT + (size_t)j * n3;
      const double scale = w_t[j] * 0.5 * h->prm.electric;
      nb_pme_induced(h, Aj, Bj, true);
      NB_LAUNCH(k_recip_bilinear, grid_for(F.n, 128), dim3(128), 0, F.n, fr, (const double *)h->d_phiu,
                (const double *)h->d_phiup, Aj, Bj, scale, h->d_gfin);
      if (vir) {
        NB_LAUNCH(k_recip_bilinear_vir, grid_for(F.n, 128), dim3(128), 0, F.n, (const double *)h->d_phiu,
                  (const double *)h->d_phiup, Aj, Bj, scale, vregion(natom));
        nb_grid_virial_pair(h, Aj, Bj, scale, vregion(ngrid));
      }
    }
    // reciprocal potentials of (u, Fhat) for the energy and DirectGrad
    nb_pme_induced(h, h->d_u, h->d_up, true);
  }
  // the polarization energy of an OST phase is scaled by its pol factor
  if (vir)
    NB_LAUNCH(k_vir_sum, dim3(1), dim3(576), 0, vused, (const double *)h->d_vfin, h->ph.pol, nb_vir_region(h, 1));
}

// the tensor terms into the gradient, scaled by the phase's pol factor (1
// outside OST), and under OST their lambda derivative (dpol) into the lambda
// gradient: the dipoles do not depend on lambda, only the energy's scale does
void nb_finite_add(NbHandle *h) {
  NB_LAUNCH(k_add_vec, grid_for(3 * h->n, 256), dim3(256), 0, 3 * h->n, (const double *)h->d_gfin, h->ph.pol,
            h->d_grad_s, h->ph.dpol, h->ost ? h->d_lgrad_s : (double *)nullptr);
}
