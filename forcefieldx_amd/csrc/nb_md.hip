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

// u += sum_k c_k hist[(start + k) mod 6]; the packed records follow u
__global__ void k_pred_apply(int n, const int32_t *__restrict__ perm, const double *__restrict__ hist,
                             const int32_t *__restrict__ ctl, double *__restrict__ u, double *__restrict__ up,
                             void *rec, int mixed) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || ctl[1] < 6) return;
  double c[6];
  pred_coefs(c);
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

// first: a from the gradient only (no half kick); partial kinetic sums per block
__global__ void __launch_bounds__(256) k_md_kick(int n3, double dt, int first, const double *__restrict__ mass,
                                                 const double *__restrict__ g, double *__restrict__ v,
                                                 double *__restrict__ a, double *__restrict__ part) {
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

__global__ void k_md_ke(int nb, const double *__restrict__ part, double *__restrict__ ke) {
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
                                                      double *__restrict__ v) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n3 && mass[q / 3] > 0.0) v[q] *= berendsen_scale(ke[0], dof, t0, dt_tau);
}

// the kinetic energy after scaling (every block above has read the old value)
__global__ void k_md_ke_scale(double dof, double t0, double dt_tau, double *__restrict__ ke) {
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
  NB_LAUNCH(k_pred_apply, grid_for(h->n, 256), dim3(256), 0, h->n, (const int32_t *)h->d_perm,
            (const double *)h->d_pred, (const int32_t *)h->d_pred_ctl, h->d_u, h->d_up, (void *)h->d_dip,
            mixed_flag(h));
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
  void *p[] = {h->d_pred, h->d_pred_ctl, h->d_md_v, h->d_md_a, h->d_md_mass, h->d_md_part, h->d_md_ke};
  for (void *q : p)
    if (q) (void)hipFree(q);
  h->d_pred = h->d_md_v = h->d_md_a = h->d_md_mass = h->d_md_part = h->d_md_ke = nullptr;
  h->d_pred_ctl = nullptr;
  h->md_ready = h->md_accel = false;
}

void nb_drop_graph(NbHandle *h);

int nb_set_scf_predictor(void *handle, int kind) {
  NbHandle *h = (NbHandle *)handle;
  if (!h) return NB_ERR_INVALID;
  if (kind != NB_PRED_NONE && kind != NB_PRED_ASPC) {
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
      // an empty history: the next solves start from the direct dipoles
      int32_t zero[2] = {0, 0};
      HIPCHECK(hipMemcpy(h->d_pred_ctl, zero, sizeof(zero), hipMemcpyHostToDevice));
    }
    return NB_OK;
  } catch (std::exception &e) {
    h->err = e.what();
    return NB_ERR_HIP;
  }
}

// ---------------------------------------------------------------------------
// dynamics
// ---------------------------------------------------------------------------
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
    auto kick = [&](int first) {
      hipLaunchKernelGGL(k_md_kick, dim3(nb), dim3(256), 0, h->stream, n3, dt, first,
                         (const double *)h->d_md_mass, (const double *)h->d_grad, h->d_md_v, h->d_md_a,
                         h->d_md_part);
      hipLaunchKernelGGL(k_md_ke, dim3(1), dim3(256), 0, h->stream, nb, (const double *)h->d_md_part, h->d_md_ke);
      HIPCHECK(hipGetLastError());
    };
    if (!h->md_accel) {
      // MolecularDynamics.init: accelerations of the starting coordinates
      int rc = nb_evaluate(h, NB_GRAD, &e);
      if (rc != NB_OK) return rc;
      kick(1);
      h->md_accel = true;
    }
    for (int step = 0; step < nsteps; step++) {
      hipLaunchKernelGGL(k_md_drift, dim3(nb), dim3(256), 0, h->stream, n3, dt, (const double *)h->d_md_mass,
                         h->d_xyz, h->d_md_v, (const double *)h->d_md_a);
      HIPCHECK(hipGetLastError());
      int rc = nb_evaluate(h, NB_GRAD, &e);
      if (rc != NB_OK) return rc;
      iters += e.pcg_iters;
      kick(0);
      if (h->md_thermo == NB_THERMO_BERENDSEN) {
        const double dof = 3.0 * h->md_active - 3.0, dt_tau = dt / h->md_tau;
        hipLaunchKernelGGL(k_md_berendsen, dim3(nb), dim3(256), 0, h->stream, n3, dof, h->md_t0, dt_tau,
                           (const double *)h->d_md_ke, (const double *)h->d_md_mass, h->d_md_v);
        hipLaunchKernelGGL(k_md_ke_scale, dim3(1), dim3(64), 0, h->stream, dof, h->md_t0, dt_tau, h->d_md_ke);
        HIPCHECK(hipGetLastError());
      }
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
