// nb_pme.hip -- reciprocal-space PME for multipoles and induced dipoles.
//
//  * k_rotate: local -> global multipoles (MultipoleType.getRotationMatrix,
//    MultipoleType.java:385-540; rotateMultipole :1031-1058), written into the
//    sorted AtomMP records together with the reduced vdW sites.
//  * k_bspline_frac: order-5 B-splines with 3 derivatives and fractional
//    multipoles (ReciprocalSpace.java:1344-1397, 2778-2834).
//  * k_spread_*: 125-point scatter with FP64 atomics into the interleaved
//    complex grid (real part: u / permanent, imaginary part: uCR), which is
//    the SpatialPermanentLoop / SpatialInducedLoop density
//    (ReciprocalSpace.java:1445-1643) without octant colouring.
//  * k_fft_*: in-LDS Stockham FFT along x, y and a fused
//    forward-z / influence multiply / inverse-z kernel (Complex3DParallel.
//    convolution, Complex3DParallel.java:445-453, 667-798); the inverse is
//    un-normalised as in the reference.
//  * k_gather_*: 20 fractional potential derivatives and the Cartesian
//    transform (ReciprocalSpace.java:2264-2455, 2467-2776, 1163-1205).
#include <cmath>
#include <stdexcept>
#include <vector>

#include "nb_kern.h"

namespace {

struct Frac { double a[3][3]; };  // a[f][c] = K_f A[c][f]

// ---------------------------------------------------------------------------
// local frames
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3<double> ld3(const double *x, int i) { return V3<double>{x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }
__device__ __forceinline__ V3<double> unit(V3<double> a) { return (1.0 / sqrt(dot(a, a))) * a; }

__global__ void k_rotate(int n, const double *__restrict__ xyz, const int32_t *__restrict__ frame,
                         const int32_t *__restrict__ axis, const double *__restrict__ mloc,
                         const int32_t *__restrict__ iperm, const int32_t *__restrict__ redidx,
                         const double *__restrict__ redval, AtomMP *__restrict__ mp, double *__restrict__ gmp,
                         double4 *__restrict__ red4) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
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
  int s = iperm[i];
  double *g = gmp + 10 * s;
  g[0] = m[0]; g[1] = d[0]; g[2] = d[1]; g[3] = d[2];
  g[4] = O[0][0]; g[5] = O[1][1]; g[6] = O[2][2]; g[7] = O[0][1]; g[8] = O[0][2]; g[9] = O[1][2];
  const double t = 1.0 / 3.0;
  AtomMP &a = mp[s];
  a.c = m[0]; a.dx = d[0]; a.dy = d[1]; a.dz = d[2];
  a.qxx = O[0][0] * t; a.qyy = O[1][1] * t; a.qzz = O[2][2] * t;
  a.qxy = O[0][1] * t; a.qxz = O[0][2] * t; a.qyz = O[1][2] * t;
  // reduced vdW site (VanDerWaals.java:1510-1540)
  int h = redidx[i];
  double rv = redval[i];
  double4 rs = make_double4(o.x, o.y, o.z, 0.0);
  if (h >= 0 && h != i && rv > 0.0) {
    V3<double> ph = ld3(xyz, h);
    rs.x = rv * (o.x - ph.x) + ph.x;
    rs.y = rv * (o.y - ph.y) + ph.y;
    rs.z = rv * (o.z - ph.z) + ph.z;
  }
  red4[s] = rs;
}

// ---------------------------------------------------------------------------
// B-splines and fractional multipoles
// ---------------------------------------------------------------------------
__global__ void k_bspline_frac(int n, const double4 *__restrict__ pos, const double *__restrict__ gmp, CellD cell,
                               int K0, int K1, int K2, Frac fr, double *__restrict__ theta,
                               int32_t *__restrict__ igrid, double *__restrict__ fmp) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double4 p = pos[s];
  int K[3] = {K0, K1, K2};
  for (int d = 0; d < 3; d++) {
    double w = p.x * cell.A[0][d] + p.y * cell.A[1][d] + p.z * cell.A[2][d];
    double u = w - floor(w + 0.5) + 0.5;
    double frx = K[d] * u;
    int ifr = (int)frx;
    double bx = frx - ifr;
    igrid[3 * s + d] = ifr - 5;
    double th[5][4];
    bspline4<5>(bx, th);
    double *o = theta + ((size_t)s * 3 + d) * 20;
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 4; j++) o[4 * k + j] = th[k][j];
  }
  const double *g = gmp + 10 * s;
  double *f = fmp + 10 * s;
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

__device__ __forceinline__ int wrapg(int i, int n) { return i >= n ? i - n : (i < 0 ? i + n : i); }

// One thread per (atom, grid point [, real/imaginary part]): consecutive
// lanes hit consecutive x points (and, for the induced grid, the two
// interleaved parts of a point), so every atomic instruction of a wavefront
// touches a handful of cache lines instead of one line per lane.
__global__ void __launch_bounds__(256) k_spread_perm(int n, const double *__restrict__ theta,
                                                     const int32_t *__restrict__ igrid, const double *__restrict__ fmp,
                                                     int K0, int K1, int K2, double *__restrict__ grid) {
  long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int s = (int)(gid / 125), r = (int)(gid - (long)s * 125);
  if (s >= n) return;
  int kz = r / 25, ky = (r / 5) % 5, kx = r % 5;
  const double *tx = theta + (size_t)s * 60, *u = tx + 20 + 4 * ky, *v = tx + 40 + 4 * kz, *w = tx + 4 * kx;
  const double *m = fmp + 10 * (size_t)s;
  int k = wrapg(igrid[3 * s + 2] + 1 + kz, K2), j = wrapg(igrid[3 * s + 1] + 1 + ky, K1);
  int i = wrapg(igrid[3 * s] + 1 + kx, K0);
  double t0 = (m[0] * v[0] + m[3] * v[1] + m[6] * v[2]) * u[0] + (m[2] * v[0] + m[9] * v[1]) * u[1] + m[5] * v[0] * u[2];
  double t1 = (m[1] * v[0] + m[8] * v[1]) * u[0] + m[7] * v[0] * u[1];
  double t2 = m[4] * v[0] * u[0];
  atomicAdd(&grid[2 * ((size_t)K0 * (j + (size_t)K1 * k) + i)], w[0] * t0 + w[1] * t1 + w[2] * t2);
}

// induced dipoles: u in the real part, uCR in the imaginary part
__global__ void __launch_bounds__(256) k_spread_ind(int n, const double *__restrict__ theta,
                                                    const int32_t *__restrict__ igrid, const double *__restrict__ u3,
                                                    const double *__restrict__ up3, Frac fr, int K0, int K1, int K2,
                                                    double *__restrict__ grid) {
  long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int s = (int)(gid / 250), r = (int)(gid - (long)s * 250);
  if (s >= n) return;
  int part = r & 1;
  r >>= 1;
  int kz = r / 25, ky = (r / 5) % 5, kx = r % 5;
  const double *tx = theta + (size_t)s * 60, *u = tx + 20 + 4 * ky, *v = tx + 40 + 4 * kz, *w = tx + 4 * kx;
  const double *src = part ? up3 : u3;
  double d[3];
  for (int c = 0; c < 3; c++)
    d[c] = fr.a[c][0] * src[3 * s] + fr.a[c][1] * src[3 * s + 1] + fr.a[c][2] * src[3 * s + 2];
  int k = wrapg(igrid[3 * s + 2] + 1 + kz, K2), j = wrapg(igrid[3 * s + 1] + 1 + ky, K1);
  int i = wrapg(igrid[3 * s] + 1 + kx, K0);
  double t0 = d[2] * v[1] * u[0] + d[1] * v[0] * u[1], t1 = d[0] * v[0] * u[0];
  atomicAdd(&grid[2 * ((size_t)K0 * (j + (size_t)K1 * k) + i) + part], w[0] * t0 + w[1] * t1);
}

__device__ __forceinline__ void phi_cart(const Frac &fr, const double *f, double *c) {
  c[0] = f[0];
  for (int j = 0; j < 3; j++) c[1 + j] = fr.a[0][j] * f[1] + fr.a[1][j] * f[2] + fr.a[2][j] * f[3];
  double F[3][3] = {{f[4], f[7], f[8]}, {f[7], f[5], f[9]}, {f[8], f[9], f[6]}}, T[3][3], C[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) T[a][b] = fr.a[0][a] * F[0][b] + fr.a[1][a] * F[1][b] + fr.a[2][a] * F[2][b];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) C[a][b] = T[a][0] * fr.a[0][b] + T[a][1] * fr.a[1][b] + T[a][2] * fr.a[2][b];
  c[4] = C[0][0]; c[5] = C[1][1]; c[6] = C[2][2]; c[7] = C[0][1]; c[8] = C[0][2]; c[9] = C[1][2];
}

// Gather: 8 lanes per atom, lane kz < 5 sums one z plane (25 rows of 5
// points) into fractional derivatives, then the lanes are combined with
// shuffles.  FULL: all 20 derivatives t000..t111 plus the Cartesian
// transform (out[30 s] = 20 fractional + 10 Cartesian).  Otherwise only the
// first derivatives are formed and the Cartesian field components are
// written to out[30 s + 21..23] (what the PCG iterations consume).
template <int NP, bool FULL>
__global__ void __launch_bounds__(256) k_gather(int n, const double *__restrict__ theta, const int32_t *__restrict__ igrid,
                                                const double *__restrict__ grid, Frac fr, int K0, int K1, int K2,
                                                double *__restrict__ out0, double *__restrict__ out1) {
  constexpr int NC = FULL ? 20 : 4;
  int gt = blockIdx.x * blockDim.x + threadIdx.x;
  int s = gt >> 3, lane = gt & 7;
  bool valid = s < n;
  double phi[NP][NC];
#pragma unroll
  for (int p = 0; p < NP; p++)
#pragma unroll
    for (int c = 0; c < NC; c++) phi[p][c] = 0.0;
  if (valid && lane < 5) {
    const double *tx = theta + (size_t)s * 60, *ty = tx + 20, *v = tx + 40 + 4 * lane;
    int k = wrapg(igrid[3 * s + 2] + 1 + lane, K2);
    int j0 = igrid[3 * s + 1] + 1, i0 = igrid[3 * s] + 1;
    int ix[5];
#pragma unroll
    for (int kx = 0; kx < 5; kx++) ix[kx] = wrapg(i0 + kx, K0);
    double tu[NP][10];
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
      for (int c = 0; c < 10; c++) tu[p][c] = 0.0;
    for (int ky = 0; ky < 5; ky++) {
      int j = wrapg(j0 + ky, K1);
      size_t row = (size_t)K0 * (j + (size_t)K1 * k);
      double t[NP][4];
#pragma unroll
      for (int p = 0; p < NP; p++) t[p][0] = t[p][1] = t[p][2] = t[p][3] = 0.0;
#pragma unroll
      for (int kx = 0; kx < 5; kx++) {
        const double *w = tx + 4 * kx;
#pragma unroll
        for (int p = 0; p < NP; p++) {
          double q = grid[2 * (row + ix[kx]) + p];
          t[p][0] += q * w[0]; t[p][1] += q * w[1];
          if (FULL) { t[p][2] += q * w[2]; t[p][3] += q * w[3]; }
        }
      }
      const double *uu = ty + 4 * ky;
#pragma unroll
      for (int p = 0; p < NP; p++) {
        tu[p][0] += t[p][0] * uu[0]; tu[p][1] += t[p][1] * uu[0]; tu[p][2] += t[p][0] * uu[1];
        if (FULL) {
          tu[p][3] += t[p][2] * uu[0]; tu[p][4] += t[p][1] * uu[1]; tu[p][5] += t[p][0] * uu[2];
          tu[p][6] += t[p][3] * uu[0]; tu[p][7] += t[p][2] * uu[1]; tu[p][8] += t[p][1] * uu[2];
          tu[p][9] += t[p][0] * uu[3];
        }
      }
    }
#pragma unroll
    for (int p = 0; p < NP; p++) {
      double *ph = phi[p];
      const double *a = tu[p];
      if (FULL) {
        ph[0] = a[0] * v[0]; ph[1] = a[1] * v[0]; ph[2] = a[2] * v[0]; ph[3] = a[0] * v[1];
        ph[4] = a[3] * v[0]; ph[5] = a[5] * v[0]; ph[6] = a[0] * v[2];
        ph[7] = a[4] * v[0]; ph[8] = a[1] * v[1]; ph[9] = a[2] * v[1];
        ph[10] = a[6] * v[0]; ph[11] = a[9] * v[0]; ph[12] = a[0] * v[3];
        ph[13] = a[7] * v[0]; ph[14] = a[3] * v[1]; ph[15] = a[8] * v[0];
        ph[16] = a[5] * v[1]; ph[17] = a[1] * v[2]; ph[18] = a[2] * v[2]; ph[19] = a[4] * v[1];
      } else {
        ph[1] = a[1] * v[0]; ph[2] = a[2] * v[0]; ph[3] = a[0] * v[1];
      }
    }
  }
#pragma unroll
  for (int p = 0; p < NP; p++)
#pragma unroll
    for (int c = 0; c < NC; c++) phi[p][c] = group_sum<8>(phi[p][c]);
  if (valid && lane == 0) {
    double *outs[2] = {out0, out1};
#pragma unroll
    for (int p = 0; p < NP; p++) {
      double *o = outs[p] + 30 * (size_t)s;
      if (FULL) {
        for (int c = 0; c < 20; c++) o[c] = phi[p][c];
        phi_cart(fr, phi[p], o + 20);
      } else {
        for (int c = 0; c < 3; c++)
          o[21 + c] = fr.a[0][c] * phi[p][1] + fr.a[1][c] * phi[p][2] + fr.a[2][c] * phi[p][3];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// FFT: in-LDS Stockham autosort, radix 2/3/4/5.  One workgroup transforms
// L lines of length N held in LDS (ping-pong).  Lines along y and z are
// grouped by consecutive x so global loads and stores are coalesced.
// ---------------------------------------------------------------------------
struct Radices { int r[12]; int nr; };

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// y[m] = sum_k x[k] exp(sign 2 pi i m k / R) for R in {2,3,4,5}
template <int R>
__device__ __forceinline__ void small_dft(double2 *v, double sign) {
  double2 o[R];
#pragma unroll
  for (int m = 0; m < R; m++) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < R; k++) {
      int q = (m * k) % R;
      double c, s;
      if (q == 0) { c = 1.0; s = 0.0; }
      else { sincospi(2.0 * q / R, &s, &c); s *= sign; }
      acc.x += v[k].x * c - v[k].y * s;
      acc.y += v[k].x * s + v[k].y * c;
    }
    o[m] = acc;
  }
#pragma unroll
  for (int m = 0; m < R; m++) v[m] = o[m];
}

// one Stockham pass over all L lines in LDS
template <int R>
__device__ __forceinline__ void stockham_pass(const double2 *__restrict__ x, double2 *__restrict__ y, int N, int L,
                                              int Ns, const double2 *__restrict__ tw, double sign) {
  int nb = N / R;
  int stride = N / (Ns * R);
  for (int b = threadIdx.x; b < L * nb; b += blockDim.x) {
    int line = b / nb, j = b - line * nb;
    const double2 *xl = x + line * N;
    double2 *yl = y + line * N;
    int jm = j % Ns;
    double2 v[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
      double2 a = xl[j + k * nb];
      if (k > 0 && jm > 0) {
        double2 w = tw[(jm * k * stride) % N];
        w.y *= sign;
        a = cmul(a, w);
      }
      v[k] = a;
    }
    small_dft<R>(v, sign);
    int base = (j / Ns) * Ns * R + jm;
#pragma unroll
    for (int k = 0; k < R; k++) yl[base + k * Ns] = v[k];
  }
}

// in-LDS transform of L lines; returns the buffer holding the result
__device__ double2 *lds_fft(double2 *a, double2 *b, int N, int L, const Radices &rad, const double2 *tw, double sign) {
  int Ns = 1;
  for (int p = 0; p < rad.nr; p++) {
    __syncthreads();
    int R = rad.r[p];
    if (R == 4) stockham_pass<4>(a, b, N, L, Ns, tw, sign);
    else if (R == 2) stockham_pass<2>(a, b, N, L, Ns, tw, sign);
    else if (R == 3) stockham_pass<3>(a, b, N, L, Ns, tw, sign);
    else stockham_pass<5>(a, b, N, L, Ns, tw, sign);
    Ns *= R;
    double2 *t = a; a = b; b = t;
  }
  __syncthreads();
  return a;
}

// axis transform.  Line `l` of the workgroup: AXIS 0 -> (y,z) rows, contiguous
// elements; AXIS 1 -> (x,z) with x fastest across lines; AXIS 2 -> (x,y).
// FUSE_CONV (AXIS 2 only): forward, multiply by influence, inverse.
template <int AXIS, bool FUSE_CONV>
__global__ void __launch_bounds__(256) k_fft_axis(double2 *__restrict__ grid, int K0, int K1, int K2, int L, Radices rad,
                                                  const double2 *__restrict__ twg, double sign,
                                                  const double *__restrict__ recip) {
  extern __shared__ double2 lds[];
  int N = AXIS == 0 ? K0 : (AXIS == 1 ? K1 : K2);
  // tw[j] = exp(+2 pi i j / N); passes conjugate it for sign < 0
  double2 *a = lds, *b = lds + (size_t)L * N, *tw = lds + 2 * (size_t)L * N;
  for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
  // line bookkeeping
  int nlines = AXIS == 0 ? K1 * K2 : (AXIS == 1 ? K0 * K2 : K0 * K1);
  int first = blockIdx.x * L;
  int cnt = min(L, nlines - first);
  // element (line l, position e) -> grid index
  auto gidx = [&](int l, int e) -> size_t {
    int line = first + l;
    if (AXIS == 0) return (size_t)line * K0 + e;
    if (AXIS == 1) {
      int x = line % K0, z = line / K0;
      return (size_t)x + (size_t)K0 * (e + (size_t)K1 * z);
    }
    return (size_t)line + (size_t)K0 * K1 * e;
  };
  for (int t = threadIdx.x; t < L * N; t += blockDim.x) {
    int l, e;
    if (AXIS == 0) { l = t / N; e = t - l * N; }
    else { e = t / L; l = t - e * L; }
    a[l * N + e] = (l < cnt) ? grid[gidx(l, e)] : make_double2(0.0, 0.0);
  }
  double2 *res = lds_fft(a, b, N, L, rad, tw, sign);
  if (FUSE_CONV) {
    for (int t = threadIdx.x; t < L * N; t += blockDim.x) {
      int e = t / L, l = t - e * L;
      if (l < cnt) {
        double g = recip[(size_t)(first + l) + (size_t)K0 * K1 * e];
        res[l * N + e].x *= g;
        res[l * N + e].y *= g;
      }
    }
    double2 *other = (res == a) ? b : a;
    res = lds_fft(res, other, N, L, rad, tw, -sign);
  }
  for (int t = threadIdx.x; t < L * N; t += blockDim.x) {
    int l, e;
    if (AXIS == 0) { l = t / N; e = t - l * N; }
    else { e = t / L; l = t - e * L; }
    if (l < cnt) grid[gidx(l, e)] = res[l * N + e];
  }
}

Radices factor(int N) {
  Radices r;
  r.nr = 0;
  int m = N;
  while (m % 4 == 0) { r.r[r.nr++] = 4; m /= 4; }
  while (m % 2 == 0) { r.r[r.nr++] = 2; m /= 2; }
  while (m % 3 == 0) { r.r[r.nr++] = 3; m /= 3; }
  while (m % 5 == 0) { r.r[r.nr++] = 5; m /= 5; }
  if (m != 1) throw std::runtime_error("PME grid dimension must factor over 2, 3, 5");
  return r;
}

int lines_per_block(int N) {
  int L = 512 / N;
  if (L > 32) L = 32;
  if (L < 1) L = 1;
  return L;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static Frac frac_of(const NbHandle *h) {
  Frac f;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) f.a[a][b] = h->afrac[a][b];
  return f;
}

// influence function of the reference (generalizedInfluenceFunction,
// ReciprocalSpace.java:1045-1121; discreteFTMod :435-492), computed on the host
static void bspline_modulus(std::vector<double> &bsmod, int n, int order) {
  double th[5][4];
  // weights at x = 0 from the same recursion as the kernels
  double w[8][8] = {{0}};
  w[1][1] = 1.0;
  for (int p = 2; p <= order; p++)
    for (int k = 0; k < p; k++) w[p][k + 1] = ((0.0 + p - 1 - k) * w[p - 1][k] + (k + 1 - 0.0) * w[p - 1][k + 1]) / (p - 1);
  (void)th;
  std::vector<double> arr(n > order + 1 ? n : order + 1, 0.0);
  for (int k = 0; k < order; k++) arr[1 + k] = w[order][k + 1];
  bsmod.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < n; j++) {
      double ang = 2.0 * M_PI * (double)((long)i * j % n) / n;
      s1 += arr[j] * cos(ang);
      s2 += arr[j] * sin(ang);
    }
    bsmod[i] = s1 * s1 + s2 * s2;
  }
  const double eps = 1.0e-7;
  if (bsmod[0] < eps) bsmod[0] = 0.5 * bsmod[1];
  for (int i = 1; i < n - 1; i++)
    if (bsmod[i] < eps) bsmod[i] = 0.5 * (bsmod[i - 1] + bsmod[i + 1]);
  if (bsmod[n - 1] < eps) bsmod[n - 1] = 0.5 * bsmod[n - 2];
  for (int i = 0; i < n; i++) {
    int k = i > n / 2 ? i - n : i;
    double zeta = 1.0;
    if (k != 0) {
      double s1 = 1.0, s2 = 1.0, f = M_PI * k / n;
      for (int j = 1; j <= 50; j++) {
        double a = f / (f + M_PI * j);
        s1 += pow(a, order); s2 += pow(a, 2 * order);
      }
      for (int j = 1; j <= 50; j++) {
        double a = f / (f - M_PI * j);
        s1 += pow(a, order); s2 += pow(a, 2 * order);
      }
      zeta = s2 / s1;
    }
    bsmod[i] *= zeta * zeta;
  }
}

void nb_pme_setup(NbHandle *h) {
  int K0 = h->prm.grid[0], K1 = h->prm.grid[1], K2 = h->prm.grid[2];
  long ntot = (long)K0 * K1 * K2;
  if (ntot != h->ntot || !h->d_grid) {
    if (h->d_grid) { hipFree(h->d_grid); hipFree(h->d_recip); hipFree(h->d_tw); }
    HIPCHECK(hipMalloc(&h->d_grid, sizeof(double) * 2 * ntot));
    HIPCHECK(hipMalloc(&h->d_recip, sizeof(double) * ntot));
    HIPCHECK(hipMalloc(&h->d_tw, sizeof(double) * 2 * (K0 + K1 + K2)));
  }
  h->K[0] = K0; h->K[1] = K1; h->K[2] = K2;
  h->ntot = ntot;
  int K[3] = {K0, K1, K2};
  for (int f = 0; f < 3; f++)
    for (int c = 0; c < 3; c++) h->afrac[f][c] = K[f] * h->cell.A[c][f];
  std::vector<double> mx, my, mz;
  bspline_modulus(mx, K0, 5);
  bspline_modulus(my, K1, 5);
  bspline_modulus(mz, K2, 5);
  std::vector<double> rec(ntot);
  double aw = h->prm.ewald_alpha;
  double piterm = (M_PI / aw) * (M_PI / aw), volterm = M_PI * h->cell.vol;
  for (int kz = 0; kz < K2; kz++) {
    int l = kz >= (K2 + 1) / 2 ? kz - K2 : kz;
    for (int ky = 0; ky < K1; ky++) {
      int k = ky >= (K1 + 1) / 2 ? ky - K1 : ky;
      for (int kx = 0; kx < K0; kx++) {
        int hh = kx >= (K0 + 1) / 2 ? kx - K0 : kx;
        double s[3];
        for (int d = 0; d < 3; d++) s[d] = h->cell.A[d][0] * hh + h->cell.A[d][1] * k + h->cell.A[d][2] * l;
        double s2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
        double term = -piterm * s2, v = 0.0;
        if (term > -50.0 && s2 > 0.0) v = exp(term) / (s2 * volterm * mx[kx] * my[ky] * mz[kz]);
        rec[kx + (size_t)K0 * (ky + (size_t)K1 * kz)] = v;
      }
    }
  }
  rec[0] = 0.0;
  HIPCHECK(hipMemcpy(h->d_recip, rec.data(), sizeof(double) * ntot, hipMemcpyHostToDevice));
  std::vector<double> tw(2 * (K0 + K1 + K2));
  int off = 0;
  for (int d = 0; d < 3; d++) {
    for (int j = 0; j < K[d]; j++) {
      double a = 2.0 * M_PI * j / K[d];  // exp(+2 pi i j / K)
      tw[2 * (off + j)] = cos(a);
      tw[2 * (off + j) + 1] = sin(a);
    }
    off += K[d];
  }
  HIPCHECK(hipMemcpy(h->d_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
  h->recip_valid = true;
}

// forward x, forward y, fused z (forward * influence * inverse), inverse y, inverse x
static void convolution(NbHandle *h) {
  int K0 = h->K[0], K1 = h->K[1], K2 = h->K[2];
  double2 *g = (double2 *)h->d_grid;
  const double2 *twx = (const double2 *)h->d_tw, *twy = twx + K0, *twz = twy + K1;
  Radices rx = factor(K0), ry = factor(K1), rz = factor(K2);
  int Lx = lines_per_block(K0), Ly = lines_per_block(K1), Lz = lines_per_block(K2);
  size_t shx = sizeof(double2) * (2 * (size_t)Lx * K0 + K0);
  size_t shy = sizeof(double2) * (2 * (size_t)Ly * K1 + K1);
  size_t shz = sizeof(double2) * (2 * (size_t)Lz * K2 + K2);
  dim3 B(128);
  dim3 Gx((K1 * K2 + Lx - 1) / Lx), Gy((K0 * K2 + Ly - 1) / Ly), Gz((K0 * K1 + Lz - 1) / Lz);
  NB_LAUNCH((k_fft_axis<0, false>), Gx, B, shx, g, K0, K1, K2, Lx, rx, twx, 1.0, nullptr);
  NB_LAUNCH((k_fft_axis<1, false>), Gy, B, shy, g, K0, K1, K2, Ly, ry, twy, 1.0, nullptr);
  NB_LAUNCH((k_fft_axis<2, true>), Gz, B, shz, g, K0, K1, K2, Lz, rz, twz, 1.0, h->d_recip);
  NB_LAUNCH((k_fft_axis<1, false>), Gy, B, shy, g, K0, K1, K2, Ly, ry, twy, -1.0, nullptr);
  NB_LAUNCH((k_fft_axis<0, false>), Gx, B, shx, g, K0, K1, K2, Lx, rx, twx, -1.0, nullptr);
}

void nb_pme_permanent(NbHandle *h) {
  int n = h->n;
  dim3 B(128), G((n + 127) / 128);
  Frac fr = frac_of(h);
  NB_LAUNCH(k_bspline_frac, G, B, 0, n, h->d_pos4, h->d_gmp, h->cell, h->K[0], h->K[1], h->K[2], fr, h->d_theta,
            h->d_igrid, h->d_fmp);
  HIPCHECK(hipMemsetAsync(h->d_grid, 0, sizeof(double) * 2 * h->ntot, h->stream));
  NB_LAUNCH(k_spread_perm, grid_for((long)n * 125, 256), dim3(256), 0, n, h->d_theta, h->d_igrid, h->d_fmp, h->K[0],
            h->K[1], h->K[2], h->d_grid);
  convolution(h);
  NB_LAUNCH((k_gather<1, true>), grid_for((long)n * 8, 256), dim3(256), 0, n, h->d_theta, h->d_igrid, h->d_grid, fr,
            h->K[0], h->K[1], h->K[2], h->d_phip, (double *)nullptr);
}

void nb_pme_induced(NbHandle *h, const double *u, const double *up, bool full) {
  int n = h->n;
  Frac fr = frac_of(h);
  HIPCHECK(hipMemsetAsync(h->d_grid, 0, sizeof(double) * 2 * h->ntot, h->stream));
  NB_LAUNCH(k_spread_ind, grid_for((long)n * 250, 256), dim3(256), 0, n, h->d_theta, h->d_igrid, u, up, fr, h->K[0],
            h->K[1], h->K[2], h->d_grid);
  convolution(h);
  if (full)
    NB_LAUNCH((k_gather<2, true>), grid_for((long)n * 8, 256), dim3(256), 0, n, h->d_theta, h->d_igrid, h->d_grid, fr,
              h->K[0], h->K[1], h->K[2], h->d_phiu, h->d_phiup);
  else
    NB_LAUNCH((k_gather<2, false>), grid_for((long)n * 8, 256), dim3(256), 0, n, h->d_theta, h->d_igrid, h->d_grid, fr,
              h->K[0], h->K[1], h->K[2], h->d_phiu, h->d_phiup);
}

void nb_rotate(NbHandle *h) {
  int n = h->n;
  dim3 B(128), G((n + 127) / 128);
  NB_LAUNCH(k_rotate, G, B, 0, n, h->d_xyz, h->d_frame, h->d_axis, h->d_mpole_local, h->d_iperm,
                     h->d_redidx, h->d_redval, h->d_mp, h->d_gmp, h->d_red4);
}

// standalone convolution entry used by the FFT parity test
extern "C" int nb_test_convolution(double *data, const double *recip, int nx, int ny, int nz) {
  try {
    NbHandle h;
    HIPCHECK(hipStreamCreate(&h.stream));
    long ntot = (long)nx * ny * nz;
    h.K[0] = nx; h.K[1] = ny; h.K[2] = nz; h.ntot = ntot;
    HIPCHECK(hipMalloc(&h.d_grid, sizeof(double) * 2 * ntot));
    HIPCHECK(hipMalloc(&h.d_recip, sizeof(double) * ntot));
    HIPCHECK(hipMalloc(&h.d_tw, sizeof(double) * 2 * (nx + ny + nz)));
    std::vector<double> tw(2 * (nx + ny + nz));
    int K[3] = {nx, ny, nz}, off = 0;
    for (int d = 0; d < 3; d++) {
      for (int j = 0; j < K[d]; j++) {
        tw[2 * (off + j)] = cos(2.0 * M_PI * j / K[d]);
        tw[2 * (off + j) + 1] = sin(2.0 * M_PI * j / K[d]);
      }
      off += K[d];
    }
    HIPCHECK(hipMemcpy(h.d_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h.d_grid, data, sizeof(double) * 2 * ntot, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h.d_recip, recip, sizeof(double) * ntot, hipMemcpyHostToDevice));
    convolution(&h);
    HIPCHECK(hipStreamSynchronize(h.stream));
    HIPCHECK(hipMemcpy(data, h.d_grid, sizeof(double) * 2 * ntot, hipMemcpyDeviceToHost));
    hipFree(h.d_grid); hipFree(h.d_recip); hipFree(h.d_tw);
    hipStreamDestroy(h.stream);
    h.d_grid = nullptr;
    return 0;
  } catch (std::exception &e) {
    return NB_ERR_HIP;
  }
}
