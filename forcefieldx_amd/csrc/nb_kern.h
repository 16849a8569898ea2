// nb_kern.h -- helpers shared by the kernel translation units.
#pragma once
#include "nb_state.h"

// mask-code -> scale tables (code = bonded relation 0/2..5 + 8 * same group)
struct ScaleTab { double m[16], p[16], d[16], v[16]; int masked[16]; };

// energy slots of NbHandle::d_energy
enum { E_PERM_REAL = 0, E_PERM_SELF, E_PERM_RECIP, E_POL_REAL, E_POL_SELF, E_POL_RECIP, E_VDW, E_QTOT, E_DUDL,
       E_D2UDL2, E_NSLOT };
// PCG scalar slots of NbHandle::d_scal
// PCG scalars; S_DONE .. S_EPSV are the device-side control of a recorded solve;
// S_RZOLD / S_RZOLDP hold the r.z totals of the last update (the denominators
// of the next conjugation)
// S_RDONE stops the FP64 refinement residuals recorded after a solve once
// one of them has passed the test; S_RCYC counts the recorded corrections a
// replay reopened the solve for.  S_RZT / S_PQT hold the r.z and p.q totals
// when their partial sums were collapsed by k_part_totals (a negative count
// passed to the consumers says so).
enum { S_PAP = 0, S_PAPP, S_RZ, S_RZP, S_EPS, S_EPSP, S_RZN, S_RZPN, S_DONE, S_ITER, S_PREV, S_STAT, S_EPSV,
       S_RZOLD, S_RZOLDP, S_RDONE, S_RCYC, S_RZT, S_RZTP, S_PQT, S_PQTP, S_NSLOT };
// Partial sums of more blocks than this are collapsed into totals by one
// block before the PCG kernels that need them: otherwise every consumer block
// would read every producer block's partial sum (at 288,000 atoms the 9,000
// blocks of the gather would each read the 18,000 r.z partials).
constexpr int NB_TOTALS_MIN = 2048;
// pinned host slots of an evaluation's read-back (doubles)
enum { P_ENERGY = 0, P_COUNT = 8, P_FLAGS = 12, P_VIR = 16, P_DUDL = 32, P_SCAL = 40, P_VALENCE = 64, P_NSLOT = 80 };
static_assert(P_SCAL + S_NSLOT <= P_VALENCE, "PCG scalars overlap the valence energies in the read-back");

static inline ScaleTab scale_tab(const NbHandle *h) {
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
    t.masked[c] = (t.m[c] != 1.0 || t.p[c] != 1.0 || t.d[c] != 1.0) ? 1 : 0;
  }
  return t;
}

static inline dim3 grid_for(long n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

// lanes per atom of the real-space pair kernels: 16 for the kernels that run
// once per evaluation over the 9-11 A rows (lists, permanent field, energies,
// vdW); 8 for the PCG's dipole-field kernels, which run every iteration over
// the shorter Ewald and preconditioner rows (at 20,736 waters the whole
// evaluation is 4 % faster mixed and 2 % FP64 than at 16; 32 lanes are slower)
#define NB_LPA 16
#define NB_LPA_PCG 8

// Workgroups are dispatched to the 8 XCDs of the device round-robin, and each
// XCD has its own L2.  xcd_block() gives XCD x the contiguous logical blocks
// [x per, (x + 1) per), so spatially sorted atoms that share neighbours are
// processed behind one L2.  Launch with xcd_grid (a multiple of 8 blocks);
// logical blocks past the end fall out through the kernels' bounds checks.
constexpr int NB_XCD = 8;
static inline dim3 xcd_grid(long n, int b) {
  long nb = (n + b - 1) / b, per = (nb + NB_XCD - 1) / NB_XCD;
  return dim3((unsigned)(per * NB_XCD));
}
static __device__ __forceinline__ int xcd_block() {
  int per = gridDim.x / NB_XCD;
  return (blockIdx.x % NB_XCD) * per + blockIdx.x / NB_XCD;
}

static __device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block sum then one atomic per block; every thread of the block must call
static __device__ __forceinline__ void block_add(double v, double *dst) {
  __shared__ double part[16];
  v = wave_sum_d(v);
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

static __device__ __forceinline__ void block_add_ll(long long v, int64_t *dst) {
  __shared__ long long partl[16];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) partl[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long s = 0;
    int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; i++) s += partl[i];
    atomicAdd((unsigned long long *)dst, (unsigned long long)s);
  }
}

// block sums of 9 values (a 3x3 strain derivative) -> part[9 * blockIdx.x + q];
// every thread of the block must call
static __device__ __forceinline__ void block_store9(double (&w)[9], double *part) {
  __shared__ double red9[9][16];
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int q = 0; q < 9; q++) {
    double v = wave_sum_d(w[q]);
    if (lane == 0) red9[q][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 9) {
    double t = 0.0;
    for (int i = 0; i < nw; i++) t += red9[threadIdx.x][i];
    part[9 * (size_t)blockIdx.x + threadIdx.x] = t;
  }
}

// block sums of NV values -> part[slot_k * stride + blockIdx.x]
template <int NV>
static __device__ __forceinline__ void block_partials(double (&v)[NV], double *part, const int (&slot)[NV], int stride) {
  __shared__ double red[NV][4];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum_d(v[k]);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) red[k][w] = v[k];
  __syncthreads();
  if (threadIdx.x < NV) {
    int k = threadIdx.x;
    part[(size_t)slot[k] * stride + blockIdx.x] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
  }
}

// totals of NV partial rows (each of nb entries); every thread receives them
template <int NV>
static __device__ __forceinline__ void block_totals(const double *part, const int (&slot)[NV], int stride, int nb,
                                             double (&tot)[NV]) {
  __shared__ double red[NV][4];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) {
    v[k] = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v[k] += part[(size_t)slot[k] * stride + i];
    v[k] = wave_sum_d(v[k]);
  }
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) red[k][w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) tot[k] = red[k][0] + red[k][1] + red[k][2] + red[k][3];
}

// The PCG's conjugate direction and its operator image, formed from the field
// of the preconditioned residual z instead of the field of p (the operator is
// linear): p = z + beta p and q = A p = A z + beta q, with A z = z / polar -
// E(z); beta = 0 on the first iteration (PCGSolver.java:991-1178 forms E(p)
// directly; the two agree to rounding).
static __device__ __forceinline__ void pcg_direction(double a, double zq, double ez, double beta, int first,
                                                     double &p, double &q) {
  const double az = a > 0.0 ? zq / a - ez : 0.0;
  if (first) {
    p = zq;
    q = az;
  } else {
    p = zq + beta * p;
    q = az + beta * q;
  }
}

// beta of the two systems: this iteration's r.z totals over the last
// update's (a zero denominator counts as 1, as the reference's guard)
static __device__ __forceinline__ void pcg_beta(const double *__restrict__ part, int stride, int nrz,
                                                const double *__restrict__ scal, int first, double (&beta)[2]) {
  beta[0] = beta[1] = 0.0;
  if (first) return;
  const int in[2] = {S_RZN, S_RZPN};
  double t[2];
  if (nrz < 0) {
    t[0] = scal[S_RZT];
    t[1] = scal[S_RZTP];
  } else {
    block_totals<2>(part, in, stride, nrz, t);
  }
  double d0 = scal[S_RZOLD], d1 = scal[S_RZOLDP];
  beta[0] = t[0] / (d0 == 0.0 ? 1.0 : d0);
  beta[1] = t[1] / (d1 == 0.0 ? 1.0 : d1);
}

// The convergence test of PCGSolver.java:543-547 on the residual totals t:
// eps = 4.80321 sqrt(max(r.r, rCR.rCR) / n); a rise of eps ends the solve as
// diverged, reaching maxit as out of iterations, eps < polar_eps as converged;
// S_DONE then stops the following iterations of a recorded solve.
static __device__ __forceinline__ void pcg_control(double *__restrict__ scal, const double (&t)[2], int n, double polar_eps,
                                            int maxit) {
  if (scal[S_DONE] != 0.0) return;
  double e = t[0] > t[1] ? t[0] : t[1];
  double eps = 4.80321 * sqrt(e / n), prev = scal[S_PREV];
  double iter = scal[S_ITER] + 1.0;
  scal[S_ITER] = iter;
  scal[S_EPSV] = eps;
  scal[S_PREV] = eps;
  if (!(eps <= prev)) {
    scal[S_STAT] = (double)NB_ERR_SCF_DIVERGED;
    scal[S_DONE] = 1.0;
  } else if (iter >= (double)maxit) {
    scal[S_STAT] = (double)NB_ERR_SCF_MAXITER;
    scal[S_DONE] = 1.0;
  } else if (eps < polar_eps) {
    scal[S_DONE] = 1.0;
  }
}

// sum over the LPA consecutive lanes that share one atom
template <int LPA> static __device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = LPA / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPA);
  return v;
}
template <int LPA> static __device__ __forceinline__ int group_sum_i(int v) {
#pragma unroll
  for (int o = LPA / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPA);
  return v;
}
template <int LPA> static __device__ __forceinline__ void group_sum3(V3<double> &a) {
  a.x = group_sum<LPA>(a.x);
  a.y = group_sum<LPA>(a.y);
  a.z = group_sum<LPA>(a.z);
}

// bonded relation of caller atoms i,k (0 if none): linear scan of i's CSR row
static __device__ __forceinline__ int relation(const int32_t *__restrict__ off, const int32_t *__restrict__ atom,
                                        const int32_t *__restrict__ rel, int i, int k) {
  int b = off[i], e = off[i + 1];
  for (int t = b; t < e; t++)
    if (atom[t] == k) return rel[t];
  return 0;
}

static __device__ __forceinline__ V3<double> ld3d(const double *x, int i) {
  return V3<double>{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
}
template <typename T> static __device__ __forceinline__ V3<T> ld3t(const double *x, int i) {
  return V3<T>{(T)x[3 * i], (T)x[3 * i + 1], (T)x[3 * i + 2]};
}
static __device__ __forceinline__ void st3(double *x, int i, V3<double> v) {
  x[3 * i] = v.x; x[3 * i + 1] = v.y; x[3 * i + 2] = v.z;
}
template <typename T> static __device__ __forceinline__ void acc3(V3<double> &a, const V3<T> &b) {
  a.x += (double)b.x; a.y += (double)b.y; a.z += (double)b.z;
}
static __device__ __forceinline__ void add3(double *x, int i, V3<double> v) {
  x[3 * i] += v.x; x[3 * i + 1] += v.y; x[3 * i + 2] += v.z;
}

// Packed per-atom record read by the dipole-field kernels: FP64 position,
// Thole parameters and the two dipole vectors in the arithmetic type (FP32 in
// the mixed form), so a neighbour costs one 64 B (FP32) or 96 B (FP64) record.
// The static part is written with the sorted coordinates; the dipoles are
// written by whichever kernel produces the vectors whose field comes next
// (u_direct, polar * r, or the conjugate direction p).
template <typename T> struct DipRec;
template <> struct alignas(16) DipRec<float> { double x, y, z; float ipd, thole, u[3], v[3]; };
template <> struct alignas(16) DipRec<double> { double x, y, z, ipd, thole, u[3], v[3], pad; };

__device__ __forceinline__ void dip_put(void *rec, int mixed, int s, int d, double a, double b) {
  if (mixed) {
    DipRec<float> *r = (DipRec<float> *)rec + s;
    r->u[d] = (float)a;
    r->v[d] = (float)b;
  } else {
    DipRec<double> *r = (DipRec<double> *)rec + s;
    r->u[d] = a;
    r->v[d] = b;
  }
}

__device__ __forceinline__ void dip_put_static(void *rec, int mixed, int s, double x, double y, double z, double ipd,
                                               double thole) {
  if (mixed) {
    DipRec<float> *r = (DipRec<float> *)rec + s;
    r->x = x; r->y = y; r->z = z;
    r->ipd = (float)ipd;
    r->thole = (float)thole;
  } else {
    DipRec<double> *r = (DipRec<double> *)rec + s;
    r->x = x; r->y = y; r->z = z;
    r->ipd = ipd;
    r->thole = thole;
  }
}

// vdW softcore factors (VanDerWaals.setLambda, VanDerWaals.java:599-653):
// lambda is rescaled by the end point, sc1 = alpha (1 - L)^2 shifts both
// buffered denominators and sc2 = L^beta scales eps; derivatives with respect
// to the rescaled L
struct Softcore { double sc1, dsc1, d2sc1, sc2, dsc2, d2sc2; int inter, intra; };

static inline Softcore softcore_of(const NbHandle *h) {
  Softcore f;
  double L = h->lambda, end = h->prm.vdw_lambda_end > 0.0 ? h->prm.vdw_lambda_end : 1.0;
  L = L > end ? 1.0 : L / end;
  double a = h->prm.vdw_lambda_alpha, b = h->prm.vdw_lambda_exponent;
  f.sc1 = a * (1.0 - L) * (1.0 - L);
  f.dsc1 = -2.0 * a * (1.0 - L);
  f.d2sc1 = 2.0 * a;
  f.sc2 = pow(L, b);
  f.dsc2 = b * pow(L, b - 1.0);
  f.d2sc2 = b >= 2.0 ? b * (b - 1.0) * pow(L, b - 2.0) : 0.0;
  f.inter = h->prm.intermolecular_softcore;
  f.intra = h->prm.intramolecular_softcore;
  return f;
}
