// nb_md.hip -- a molecular-dynamics caller of the evaluation and its SCF
// predictor.
//
//  * ASPC predictor (SCFPredictorParameters.java:79-121, saved by
//    saveMutualInducedDipoles :269-302, wired at ParticleMeshEwald.java:
//    2272-2281 and 1952-1955): after each converged mutual solve the
//    difference between the induced and the direct dipoles (u and uCR) is
//    stored in a ring of six, in caller order (the sorted order changes at
//    list rebuilds); once six are held, the next solve starts from the
//    direct dipoles plus sum_k c_k (u - u_direct)_(t-1-k) with Kolafa's
//    coefficients 22/7, -55/14, 55/21, -22/21, 5/21, -1/42.  The ring's start
//    and fill count live in device memory and are advanced by a kernel, so a
//    recorded graph replays the predictor unchanged.
//  * Least-squares predictor (SCFPredictorParameters.java:144-213 and its
//    LeastSquaresPredictor model, :315-418): the newest entry is fitted by
//    the five before it; the 21 products of the six entries are reduced in
//    a fixed order (k_pred_gram, k_pred_ls) and the 5 x 5 normal equations
//    are solved by a Cholesky factorisation that drops columns depending on
//    earlier ones, all on the device; the coefficients then act on the five
//    newest entries through the same apply kernel as ASPC.
//  * Velocity Verlet (VelocityVerlet.java:66-113, MolecularDynamics.java:
//    1285-1299): v += a dt/2, x += v dt; evaluate; a = -418.4 g / m,
//    v += a dt/2; the kinetic energy sum m v^2 / (2 418.4) is reduced in a
//    fixed order.  Positions and velocities stay on the device between steps;
//    each step's evaluation is the handle's (a graph replay in graph mode).
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "nb_kern.h"

namespace {

// kcal/mol -> g/mol A^2/ps^2 (Constants.java:205)
constexpr double KCAL_TO_GRAM_ANG2_PER_PS2 = 4.1840e2;

__device__ __forceinline__ void pred_coefs(double *c) {
  c[0] = 22.0 / 7.0;
  c[1] = -55.0 / 14.0;
  c[2] = 55.0 / 21.0;
  c[3] = -22.0 / 21.0;
  c[4] = 5.0 / 21.0;
  c[5] = -1.0 / 42.0;
}

// u += sum_k c_k hist[(start + k) mod 6] once min_count entries are held;
// the packed records follow u
__global__ void k_pred_apply(int n, const int32_t *__restrict__ perm, const double *__restrict__ hist,
                             const int32_t *__restrict__ ctl, const double *__restrict__ coef, int min_count,
                             double *__restrict__ u, double *__restrict__ up, void *rec, int mixed) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || ctl[1] < min_count) return;
  double c[6];
  for (int k = 0; k < 6; k++) c[k] = coef[k];
  const int i = perm[s], start = ctl[0];
  const size_t stride = 6 * (size_t)n;
  for (int d = 0; d < 3; d++) {
    double du = 0.0, dup = 0.0;
    for (int k = 0; k < 6; k++) {
      const double *h = hist + (size_t)((start + k) % 6) * stride;
      du += c[k] * h[3 * i + d];
      dup += c[k] * h[3 * (size_t)n + 3 * i + d];
    }
    const double a = u[3 * s + d] + du, b = up[3 * s + d] + dup;
    u[3 * s + d] = a;
    up[3 * s + d] = b;
    dip_put(rec, mixed, s, d, a, b);
  }
}

// A solve's dipoles enter the history only when the evaluation that produced
// them is kept: its filtered rows did not overflow (flags 1, 2, 5: the
// evaluation is repeated with larger rows), no list rebuild was pending
// (flag 3: a replay is repeated on the regular path), and, for a solve under
// device control (scal non-null), the solve stopped converged.
__device__ __forceinline__ bool pred_keep(const int32_t *__restrict__ flags, const double *__restrict__ scal) {
  if (flags[1] | flags[2] | flags[3] | flags[5]) return false;
  return !scal || (scal[S_DONE] != 0.0 && scal[S_STAT] == 0.0);
}

// the converged u - u_direct into the slot before the current start
__global__ void k_pred_save(int n, const int32_t *__restrict__ perm, const double *__restrict__ u,
                            const double *__restrict__ up, const double *__restrict__ ud,
                            const double *__restrict__ upd, const int32_t *__restrict__ ctl,
                            double *__restrict__ hist, const int32_t *__restrict__ flags,
                            const double *__restrict__ scal) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !pred_keep(flags, scal)) return;
  const int slot = (ctl[0] + 5) % 6, i = perm[s];
  double *h = hist + (size_t)slot * 6 * (size_t)n;
  for (int d = 0; d < 3; d++) {
    h[3 * i + d] = u[3 * s + d] - ud[3 * s + d];
    h[3 * (size_t)n + 3 * i + d] = up[3 * s + d] - upd[3 * s + d];
  }
}

__global__ void k_pred_aspc_coefs(double *coef) {
  if (threadIdx.x == 0) pred_coefs(coef);
}

// index of the product of the entries of ages a <= b among the 21
__host__ __device__ constexpr int gram_index(int a, int b) { return a * 6 - a * (a - 1) / 2 + (b - a); }

// the 21 products of the six history entries ordered by age (0 the newest),
// as per-block partial sums part[21][gridDim.x] (grid-stride over the 6n values)
__global__ void __launch_bounds__(256) k_pred_gram(size_t n6, const double *__restrict__ hist,
                                                   const int32_t *__restrict__ ctl, double *__restrict__ part) {
  if (ctl[1] < 2) return;
  const int start = ctl[0];
  const double *e[6];
  for (int m = 0; m < 6; m++) e[m] = hist + (size_t)((start + m) % 6) * n6;
  double acc[21];
  for (int t = 0; t < 21; t++) acc[t] = 0.0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n6; q += (size_t)gridDim.x * blockDim.x) {
    double v[6];
    for (int m = 0; m < 6; m++) v[m] = e[m][q];
    int t = 0;
    for (int a = 0; a < 6; a++)
      for (int b = a; b < 6; b++) acc[t++] += v[a] * v[b];
  }
  __shared__ double red[21][4];
  for (int t = 0; t < 21; t++) {
    const double w = wave_sum_d(acc[t]);
    if ((threadIdx.x & 63) == 0) red[t][threadIdx.x >> 6] = w;
  }
  __syncthreads();
  if (threadIdx.x < 21)
    part[(size_t)threadIdx.x * gridDim.x + blockIdx.x] =
        red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

// Least-squares coefficients: minimise |e_0 - sum_j c_j e_(1+j)|^2 over the
// m = min(count - 1, 5) filled older entries, by the normal equations
// A c = r, A_jl = e_(1+j).e_(1+l), r_j = e_0.e_(1+j).  Cholesky, with a column
// whose part orthogonal to the kept earlier ones has squared norm below
// 1e-12 of its own dropped (c = 0); unfilled columns keep c = 0.
__global__ void k_pred_ls(const double *__restrict__ part, int nblk, const int32_t *__restrict__ ctl,
                          double *__restrict__ coef) {
  __shared__ double G[21];
  const int count = ctl[1];
  if (count < 2) return;
  if (threadIdx.x < 21) {
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(size_t)threadIdx.x * nblk + b];
    G[threadIdx.x] = s;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int m = count - 1 < 5 ? count - 1 : 5;
  double A[5][5], r[5], L[5][5], c[6] = {0, 0, 0, 0, 0, 0}, y[5];
  bool keep[5];
  for (int j = 0; j < m; j++) {
    r[j] = G[gram_index(0, 1 + j)];
    for (int l = 0; l < m; l++) A[j][l] = G[j <= l ? gram_index(1 + j, 1 + l) : gram_index(1 + l, 1 + j)];
  }
  for (int j = 0; j < m; j++) {
    double d = A[j][j];
    for (int k = 0; k < j; k++)
      if (keep[k]) d -= L[j][k] * L[j][k];
    keep[j] = A[j][j] > 0.0 && d > 1e-12 * A[j][j];
    L[j][j] = keep[j] ? sqrt(d) : 0.0;
    for (int i = j + 1; i < m; i++) {
      double v = 0.0;
      if (keep[j]) {
        v = A[i][j];
        for (int k = 0; k < j; k++)
          if (keep[k]) v -= L[i][k] * L[j][k];
        v /= L[j][j];
      }
      L[i][j] = v;
    }
  }
  for (int j = 0; j < m; j++) {
    if (!keep[j]) continue;
    double v = r[j];
    for (int k = 0; k < j; k++)
      if (keep[k]) v -= L[j][k] * y[k];
    y[j] = v / L[j][j];
  }
  for (int j = m - 1; j >= 0; j--) {
    if (!keep[j]) continue;
    double v = y[j];
    for (int k = j + 1; k < m; k++)
      if (keep[k]) v -= L[k][j] * c[k];
    c[j] = v / L[j][j];
  }
  for (int k = 0; k < 6; k++) coef[k] = c[k];
}

__global__ void k_pred_advance(int32_t *ctl, const int32_t *__restrict__ flags, const double *__restrict__ scal) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && pred_keep(flags, scal)) {
    ctl[0] = (ctl[0] + 5) % 6;
    ctl[1] = ctl[1] < 6 ? ctl[1] + 1 : 6;
  }
}

__global__ void k_md_drift(int n3, double dt, const double *__restrict__ mass, double *__restrict__ x,
                           double *__restrict__ v, const double *__restrict__ a) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n3 || !(mass[q / 3] > 0.0)) return;
  const double vq = v[q] + a[q] * (0.5 * dt);
  v[q] = vq;
  x[q] += vq * dt;
}

// A kick recorded into a step's graph runs only when the evaluation before it
// is kept (flags non-null: the pred_keep test of the evaluation's flags and,
// for a mutual solve, its control scalars); the host finishes the step
// otherwise.  first: a from the gradient only (no half kick); partial kinetic
// sums per block
__global__ void __launch_bounds__(256) k_md_kick(int n3, double dt, int first, const double *__restrict__ mass,
                                                 const double *__restrict__ g, double *__restrict__ v,
                                                 double *__restrict__ a, double *__restrict__ part,
                                                 const int32_t *__restrict__ flags,
                                                 const double *__restrict__ scal) {
  if (flags && !pred_keep(flags, scal)) return;
  __shared__ double red[4];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  double ke = 0.0;
  if (q < n3) {
    const double m = mass[q / 3];
    if (m > 0.0) {
      const double aq = -KCAL_TO_GRAM_ANG2_PER_PS2 * g[q] / m;
      a[q] = aq;
      double vq = v[q];
      if (!first) {
        vq += aq * (0.5 * dt);
        v[q] = vq;
      }
      ke = m * vq * vq;
    }
  }
  ke = wave_sum_d(ke);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ke;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void k_md_ke(int nb, const double *__restrict__ part, double *__restrict__ ke,
                        const int32_t *__restrict__ flags, const double *__restrict__ scal) {
  if (flags && !pred_keep(flags, scal)) return;
  __shared__ double red[4];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
  s = wave_sum_d(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) ke[0] = 0.5 * (red[0] + red[1] + red[2] + red[3]) / KCAL_TO_GRAM_ANG2_PER_PS2;
}

// Berendsen weak coupling (Berendsen.java:99-101) after the full step: T from
// the kinetic energy over 3m - 3 degrees of freedom, m the atoms with mass > 0
// (the active variables of Thermostat.java:167, 262-274),
// v *= sqrt(1 + dt/tau (T0/T - 1)); a system at rest is left as it is.
constexpr double KB_GRAM_ANG2_PER_PS2_K = 0.831446261815324;  // Boltzmann constant, g A^2 / (ps^2 mol K)

__device__ __forceinline__ double berendsen_scale(double ke, double dof, double t0, double dt_tau) {
  if (!(ke > 0.0)) return 1.0;
  const double temp = 2.0 * ke * KCAL_TO_GRAM_ANG2_PER_PS2 / (KB_GRAM_ANG2_PER_PS2_K * dof);
  return sqrt(1.0 + dt_tau * (t0 / temp - 1.0));
}

// only atoms with mass > 0 move (Berendsen.fullStep skips the others, as the
// drift and kick do)
__global__ void __launch_bounds__(256) k_md_berendsen(int n3, double dof, double t0, double dt_tau,
                                                      const double *__restrict__ ke, const double *__restrict__ mass,
                                                      double *__restrict__ v, const int32_t *__restrict__ flags,
                                                      const double *__restrict__ scal) {
  if (flags && !pred_keep(flags, scal)) return;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n3 && mass[q / 3] > 0.0) v[q] *= berendsen_scale(ke[0], dof, t0, dt_tau);
}

// the kinetic energy after scaling (every block above has read the old value)
__global__ void k_md_ke_scale(double dof, double t0, double dt_tau, double *__restrict__ ke,
                              const int32_t *__restrict__ flags, const double *__restrict__ scal) {
  if (flags && !pred_keep(flags, scal)) return;
  if (threadIdx.x == 0) {
    const double s = berendsen_scale(ke[0], dof, t0, dt_tau);
    ke[0] *= s * s;
  }
}

}  // namespace

int nb_md_set_thermostat(void *handle, int kind, double temperature, double tau) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  if (kind == NB_THERMO_NONE) {
    h->md_thermo = NB_THERMO_NONE;
    return NB_OK;
  }
  if (kind != NB_THERMO_BERENDSEN || !(temperature > 0.0) || !(tau > 0.0) || h->n < 2) {
    h->err = "thermostat: kind NB_THERMO_BERENDSEN with temperature > 0 K and tau > 0 ps, at least two atoms";
    return NB_ERR_INVALID;
  }
  h->md_thermo = kind;
  h->md_t0 = temperature;
  h->md_tau = tau;
  return NB_OK;
}

// ---------------------------------------------------------------------------
// predictor, called from the evaluation sequence
// ---------------------------------------------------------------------------
static int mixed_flag(const NbHandle *h) { return h->prm.precision == NB_PREC_MIXED ? 1 : 0; }

void nb_pred_apply(NbHandle *h) {
  if (h->predictor == NB_PRED_NONE || !h->d_pred) return;
  if (h->predictor == NB_PRED_LS) {
    NB_LAUNCH(k_pred_gram, dim3(NB_PRED_BLOCKS), dim3(256), 0, 6 * (size_t)h->n, (const double *)h->d_pred,
              (const int32_t *)h->d_pred_ctl, h->d_pred_part);
    NB_LAUNCH(k_pred_ls, dim3(1), dim3(64), 0, (const double *)h->d_pred_part, NB_PRED_BLOCKS,
              (const int32_t *)h->d_pred_ctl, h->d_pred_coef);
  }
  NB_LAUNCH(k_pred_apply, grid_for(h->n, 256), dim3(256), 0, h->n, (const int32_t *)h->d_perm,
            (const double *)h->d_pred, (const int32_t *)h->d_pred_ctl, (const double *)h->d_pred_coef,
            h->predictor == NB_PRED_LS ? 2 : 6, h->d_u, h->d_up, (void *)h->d_dip, mixed_flag(h));
}

void nb_pred_save(NbHandle *h, bool device_ctl) {
  if (h->predictor == NB_PRED_NONE || !h->d_pred) return;
  const double *scal = device_ctl ? h->d_scal : nullptr;
  NB_LAUNCH(k_pred_save, grid_for(h->n, 256), dim3(256), 0, h->n, (const int32_t *)h->d_perm,
            (const double *)h->d_u, (const double *)h->d_up, (const double *)h->d_udir, (const double *)h->d_updir,
            (const int32_t *)h->d_pred_ctl, h->d_pred, (const int32_t *)h->d_flags, scal);
  NB_LAUNCH(k_pred_advance, dim3(1), dim3(64), 0, h->d_pred_ctl, (const int32_t *)h->d_flags, scal);
}

void nb_md_free(NbHandle *h) {
  void *p[] = {h->d_pred, h->d_pred_ctl, h->d_pred_coef, h->d_pred_part, h->d_md_v, h->d_md_a, h->d_md_mass,
               h->d_md_part, h->d_md_ke};
  for (void *q : p)
    if (q) (void)hipFree(q);
  h->d_pred = h->d_pred_coef = h->d_pred_part = h->d_md_v = h->d_md_a = h->d_md_mass = h->d_md_part = h->d_md_ke =
      nullptr;
  h->d_pred_ctl = nullptr;
  h->md_ready = h->md_accel = false;
}

void nb_drop_graph(NbHandle *h);

int nb_set_scf_predictor(void *handle, int kind) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  if (kind != NB_PRED_NONE && kind != NB_PRED_ASPC && kind != NB_PRED_LS) {
    h->err = "unknown SCF predictor";
    return NB_ERR_INVALID;
  }
  try {
    HIPCHECK(hipSetDevice(h->device));
    HIPCHECK(hipStreamSynchronize(h->stream));
    nb_drop_graph(h);
    h->predictor = kind;
    if (kind != NB_PRED_NONE) {
      if (!h->d_pred) HIPCHECK(hipMalloc(&h->d_pred, sizeof(double) * 36 * (size_t)h->n));
      if (!h->d_pred_ctl) HIPCHECK(hipMalloc(&h->d_pred_ctl, sizeof(int32_t) * 2));
      if (!h->d_pred_coef) HIPCHECK(hipMalloc(&h->d_pred_coef, sizeof(double) * 6));
      if (!h->d_pred_part) HIPCHECK(hipMalloc(&h->d_pred_part, sizeof(double) * 21 * NB_PRED_BLOCKS));
      // an empty history: the next solves start from the direct dipoles
      int32_t zero[2] = {0, 0};
      HIPCHECK(hipMemcpy(h->d_pred_ctl, zero, sizeof(zero), hipMemcpyHostToDevice));
      HIPCHECK(hipMemset(h->d_pred, 0, sizeof(double) * 36 * (size_t)h->n));
      HIPCHECK(hipMemset(h->d_pred_coef, 0, sizeof(double) * 6));
      if (kind == NB_PRED_ASPC) hipLaunchKernelGGL(k_pred_aspc_coefs, dim3(1), dim3(64), 0, h->stream, h->d_pred_coef);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(h->stream));
    }
    return NB_OK;
  } catch (std::exception &e) {
    h->err = e.what();
    return NB_ERR_HIP;
  }
}

int nb_get_scf_predictor(void *handle, double *coef, double *history, int32_t *state) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  if (h->predictor == NB_PRED_NONE || !h->d_pred) {
    h->err = "no SCF predictor is set";
    return NB_ERR_INVALID;
  }
  try {
    HIPCHECK(hipSetDevice(h->device));
    HIPCHECK(hipStreamSynchronize(h->stream));
    if (coef) HIPCHECK(hipMemcpy(coef, h->d_pred_coef, sizeof(double) * 6, hipMemcpyDeviceToHost));
    if (history) HIPCHECK(hipMemcpy(history, h->d_pred, sizeof(double) * 36 * (size_t)h->n, hipMemcpyDeviceToHost));
    if (state) HIPCHECK(hipMemcpy(state, h->d_pred_ctl, sizeof(int32_t) * 2, hipMemcpyDeviceToHost));
    return NB_OK;
  } catch (std::exception &e) {
    h->err = e.what();
    return NB_ERR_HIP;
  }
}

// ---------------------------------------------------------------------------
// dynamics
// ---------------------------------------------------------------------------
// the drift of a step (x += v dt after the first half kick), on the main stream
void nb_md_drift(NbHandle *h) {
  const int n3 = 3 * h->n, nb = (n3 + 255) / 256;
  NB_LAUNCH(k_md_drift, dim3(nb), dim3(256), 0, n3, h->md_dt, (const double *)h->d_md_mass, h->d_xyz, h->d_md_v,
            (const double *)h->d_md_a);
}

// the second half kick, kinetic energy and thermostat of a step; guarded: the
// kernels check the evaluation's flags (and the solve's control scalars)
// themselves, as they must in a recorded step
void nb_md_kick(NbHandle *h, bool guarded) {
  const int n3 = 3 * h->n, nb = (n3 + 255) / 256;
  const bool mutual = h->prm.do_pme && h->prm.polarization == NB_POL_MUTUAL && h->ph.polar;
  const int32_t *fl = guarded ? (const int32_t *)h->d_flags : nullptr;
  const double *sc = guarded && mutual ? (const double *)h->d_scal : nullptr;
  const double dt = h->md_dt;
  NB_LAUNCH(k_md_kick, dim3(nb), dim3(256), 0, n3, dt, 0, (const double *)h->d_md_mass, (const double *)h->d_grad,
            h->d_md_v, h->d_md_a, h->d_md_part, fl, sc);
  NB_LAUNCH(k_md_ke, dim3(1), dim3(256), 0, nb, (const double *)h->d_md_part, h->d_md_ke, fl, sc);
  if (h->md_thermo == NB_THERMO_BERENDSEN) {
    const double dof = 3.0 * h->md_active - 3.0, dt_tau = dt / h->md_tau;
    NB_LAUNCH(k_md_berendsen, dim3(nb), dim3(256), 0, n3, dof, h->md_t0, dt_tau, (const double *)h->d_md_ke,
              (const double *)h->d_md_mass, h->d_md_v, fl, sc);
    NB_LAUNCH(k_md_ke_scale, dim3(1), dim3(64), 0, dof, h->md_t0, dt_tau, h->d_md_ke, fl, sc);
  }
  HIPCHECK(hipGetLastError());
}
int nb_md_init(void *handle, const double *mass, const double *v0) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || !mass) return NB_ERR_INVALID;
  try {
    HIPCHECK(hipSetDevice(h->device));
    const int n = h->n, nb = (3 * n + 255) / 256;
    if (!h->d_md_v) {
      HIPCHECK(hipMalloc(&h->d_md_v, sizeof(double) * 3 * n));
      HIPCHECK(hipMalloc(&h->d_md_a, sizeof(double) * 3 * n));
      HIPCHECK(hipMalloc(&h->d_md_mass, sizeof(double) * n));
      HIPCHECK(hipMalloc(&h->d_md_part, sizeof(double) * nb));
      HIPCHECK(hipMalloc(&h->d_md_ke, sizeof(double)));
    }
    HIPCHECK(hipMemcpy(h->d_md_mass, mass, sizeof(double) * n, hipMemcpyHostToDevice));
    h->md_active = 0;
    for (int i = 0; i < n; i++) h->md_active += mass[i] > 0.0 ? 1 : 0;
    if (v0) {
      HIPCHECK(hipMemcpy(h->d_md_v, v0, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    } else {
      std::vector<double> zero(3 * (size_t)n, 0.0);
      HIPCHECK(hipMemcpy(h->d_md_v, zero.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    }
    h->md_ready = true;
    h->md_accel = false;
    return NB_OK;
  } catch (std::exception &e) {
    h->err = e.what();
    return NB_ERR_HIP;
  }
}

int nb_md_step(void *handle, int nsteps, double dt, nb_md_report *out) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || nsteps < 0 || !(dt > 0.0)) return NB_ERR_INVALID;
  if (h->md_thermo == NB_THERMO_BERENDSEN && dt > h->md_tau) {
    h->err = "thermostat: the time step must not exceed tau";
    return NB_ERR_INVALID;
  }
  if (!h->md_ready || !h->cell_set || !h->positions_set) {
    h->err = "nb_md_init, nb_set_cell and nb_set_positions must be called before nb_md_step";
    return NB_ERR_INVALID;
  }
  if (h->md_thermo == NB_THERMO_BERENDSEN && h->md_active < 2) {
    h->err = "thermostat: at least two atoms with mass > 0";
    return NB_ERR_INVALID;
  }
  try {
    HIPCHECK(hipSetDevice(h->device));
    const int n3 = 3 * h->n, nb = (n3 + 255) / 256;
    nb_energies e;
    std::memset(&e, 0, sizeof(e));
    long iters = 0;
    h->md_dt = dt;
    if (!h->md_accel) {
      // MolecularDynamics.init: accelerations of the starting coordinates
      int rc = nb_evaluate(h, NB_GRAD, &e);
      if (rc != NB_OK) return rc;
      hipLaunchKernelGGL(k_md_kick, dim3(nb), dim3(256), 0, h->stream, n3, dt, 1, (const double *)h->d_md_mass,
                         (const double *)h->d_grad, h->d_md_v, h->d_md_a, h->d_md_part, (const int32_t *)nullptr,
                         (const double *)nullptr);
      hipLaunchKernelGGL(k_md_ke, dim3(1), dim3(256), 0, h->stream, nb, (const double *)h->d_md_part, h->d_md_ke,
                         (const int32_t *)nullptr, (const double *)nullptr);
      HIPCHECK(hipGetLastError());
      h->md_accel = true;
    }
    for (int step = 0; step < nsteps; step++) {
      int rc;
      if (h->graph_on && !h->elec_lambda) {
        // the whole step in one recorded graph (drift, evaluation, kick)
        rc = nb_evaluate_md_step(h, &e);
      } else {
        nb_md_drift(h);
        HIPCHECK(hipGetLastError());
        rc = nb_evaluate(h, NB_GRAD, &e);
        if (rc == NB_OK) nb_md_kick(h, false);
      }
      if (rc != NB_OK) return rc;
      iters += e.pcg_iters;
    }
    double ke = 0.0;
    HIPCHECK(hipMemcpyAsync(&ke, h->d_md_ke, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    if (out) {
      out->kinetic = ke;
      out->potential = e.total;
      out->total = ke + e.total;
      out->steps = nsteps;
      out->pcg_iters = (int32_t)iters;
      out->last = e;
    }
    return NB_OK;
  } catch (std::exception &ex) {
    h->err = ex.what();
    return NB_ERR_HIP;
  }
}

int nb_md_get_state(void *handle, double *x, double *v) {
  NbHandle *h = (NbHandle *)handle;
  if (!h || !h->md_ready) return NB_ERR_INVALID;
  try {
    HIPCHECK(hipSetDevice(h->device));
    if (x) HIPCHECK(hipMemcpyAsync(x, h->d_xyz, sizeof(double) * 3 * h->n, hipMemcpyDeviceToHost, h->stream));
    if (v) HIPCHECK(hipMemcpyAsync(v, h->d_md_v, sizeof(double) * 3 * h->n, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    return NB_OK;
  } catch (std::exception &ex) {
    h->err = ex.what();
    return NB_ERR_HIP;
  }
}
