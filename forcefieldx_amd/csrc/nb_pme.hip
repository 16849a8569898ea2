// nb_pme.hip -- reciprocal-space PME for multipoles and induced dipoles.
//
//  * k_rotate: local -> global multipoles (MultipoleType.getRotationMatrix,
//    MultipoleType.java:385-540; rotateMultipole :1031-1058), written into the
//    sorted AtomMP records together with the reduced vdW sites.
//  * k_bspline_frac: order-5 B-splines with 3 derivatives and fractional
//    multipoles (ReciprocalSpace.java:1344-1397, 2778-2834), plus the binning
//    of atoms into grid bricks.
//  * k_spread_brick: the SpatialPermanentLoop / SpatialInducedLoop density
//    (ReciprocalSpace.java:1445-1643).  Instead of the reference's 8-colour
//    octant decomposition, one workgroup owns one brick of the grid: it
//    accumulates the 125-point stencils of its atoms in LDS (exact sums of
//    quantized values, independent of the order of the LDS atomics) and
//    writes the brick, including its 4-point spill into the next bricks, to a
//    brick buffer; the forward x transform sums, for every grid point, its
//    own brick and the spill of up to 7 neighbours into the interleaved
//    complex grid (real part: permanent / u, imaginary part: uCR).  No global
//    atomics.
//  * k_ct_* / k_fft_lines: in-place in-LDS Stockham FFT (radix 2/3/4/5/7) along
//    x (fused with the brick combine on the forward pass) and y, and a fused
//    forward-z / influence multiply / inverse-z kernel
//    (Complex3DParallel.convolution, Complex3DParallel.java:445-453, 667-798);
//    the inverse is un-normalised as in the reference.
//  * k_gather: 20 fractional potential derivatives and the Cartesian
//    transform (ReciprocalSpace.java:2264-2455, 2467-2776, 1163-1205).
//
// Every grid kernel is templated on the grid real T: double in the FP64 form,
// float in the mixed form (spread, FFT, convolution and gather arithmetic in
// FP32; the potentials handed back are FP64).
#include <cmath>
#include <stdexcept>
#include <vector>

#include "nb_kern.h"
#include "nb_atom.h"
#include "nb_fft.h"

namespace {

struct Frac2v { double a[3][3]; };  // cell.A (fractional = Cartesian . A)


// ---------------------------------------------------------------------------
// local frames
// ---------------------------------------------------------------------------

__global__ void k_rotate(int n, const double *__restrict__ xyz, const int32_t *__restrict__ frame,
                         const int32_t *__restrict__ axis, const double *__restrict__ mloc,
                         const int32_t *__restrict__ iperm, const int32_t *__restrict__ redidx,
                         const double *__restrict__ redval, AtomMP *__restrict__ mp, double *__restrict__ gmp,
                         double4 *__restrict__ red4, CellD cell, const int32_t *__restrict__ wrap) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3<double> o = ld3(xyz, i);
  double gl[10];
  rotate_multipole(i, xyz, frame, axis, mloc, gl);
  int s = iperm[i];
  double *g = gmp + 10 * s;
  for (int q = 0; q < 10; q++) g[q] = gl[q];
  const double t = 1.0 / 3.0;
  AtomMP &a = mp[s];
  a.c = gl[0]; a.dx = gl[1]; a.dy = gl[2]; a.dz = gl[3];
  a.qxx = gl[4] * t; a.qyy = gl[5] * t; a.qzz = gl[6] * t;
  a.qxy = gl[7] * t; a.qxz = gl[8] * t; a.qyz = gl[9] * t;
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
  if (cell.multi) {
    int wx = wrap[3 * s], wy = wrap[3 * s + 1], wz = wrap[3 * s + 2];
    rs.x -= wx * cell.Ai00 + wy * cell.Ai10 + wz * cell.Ai20;
    rs.y -= wy * cell.Ai11 + wz * cell.Ai21;
    rs.z -= wz * cell.Ai22;
  }
  red4[s] = rs;
}

// ---------------------------------------------------------------------------
// bricks: axis d is cut into nb bricks, brick b = [b K / nb, (b+1) K / nb)
// ---------------------------------------------------------------------------
struct Bricks { int K[3], nb[3], ext; };
// longest brick edge (nb = max(2, ceil(K / 8)) keeps every edge at or below
// it); a spread workgroup's LDS brick holds the edge plus the 4-point spill
#define NB_BRICK_MAX 8
#define NB_BRICK_EXT (NB_BRICK_MAX + 4)

__host__ __device__ __forceinline__ int brick_of(int p, int K, int nb) { return ((p + 1) * nb - 1) / K; }
__host__ __device__ __forceinline__ int brick_start(int b, int K, int nb) { return (b * K) / nb; }


// B-splines, grid start, fractional multipoles and brick binning.  The
// weights are stored in the grid's real type TH: the spread and the gather
// compute in that type; theta64 (when given) also receives them in FP64, for
// the FP64 refinement of the mixed solve.
template <typename TH, bool BIN = true>
__global__ void k_bspline_frac(int n, const double4 *__restrict__ pos, const double *__restrict__ gmp, CellD cell,
                               Bricks bk, Frac fr, TH *__restrict__ theta, int32_t *__restrict__ igrid,
                               double *__restrict__ fmp, int32_t *__restrict__ bcount, int32_t *__restrict__ bid,
                               int32_t *__restrict__ brank, double *__restrict__ theta64) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double4 p = pos[s];
  int b3[3];
  for (int d = 0; d < 3; d++) {
    double w = p.x * cell.A[0][d] + p.y * cell.A[1][d] + p.z * cell.A[2][d];
    double u = w - floor(w + 0.5) + 0.5;
    double frx = bk.K[d] * u;
    int ifr = (int)frx;
    double bx = frx - ifr;
    if (BIN) igrid[3 * s + d] = ifr - 5;
    b3[d] = brick_of(wrapg(ifr - 4, bk.K[d]), bk.K[d], bk.nb[d]);
    double th[5][4];
    bspline4<5>(bx, th);
    TH *o = theta + ((size_t)s * 3 + d) * 20;
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 4; j++) o[4 * k + j] = (TH)th[k][j];
    if (theta64) {
      double *o64 = theta64 + ((size_t)s * 3 + d) * 20;
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 4; j++) o64[4 * k + j] = th[k][j];
    }
  }
  if (BIN) {
    int b = b3[0] + bk.nb[0] * (b3[1] + bk.nb[1] * b3[2]);
    bid[s] = b;
    brank[s] = atomicAdd(&bcount[b], 1);
    frac_multipole(fr, gmp + 10 * s, fmp + 10 * s);
  }
}

// exclusive scan of the brick counts (one workgroup of 1024)
__global__ void __launch_bounds__(1024) k_brick_scan(int nb, const int32_t *__restrict__ bcount, int32_t *__restrict__ bstart) {
  __shared__ int32_t part[1024];
  int t = threadIdx.x;
  int per = (nb + 1023) / 1024;
  int lo = t * per, hi = min(nb, lo + per);
  int sum = 0;
  for (int b = lo; b < hi; b++) sum += bcount[b];
  part[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int v = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;
  for (int b = lo; b < hi; b++) {
    bstart[b] = run;
    run += bcount[b];
  }
  if (t == 1023) bstart[nb] = part[1023];
}

__global__ void k_brick_fill(int n, const int32_t *__restrict__ bid, const int32_t *__restrict__ brank,
                             const int32_t *__restrict__ bstart, int32_t *__restrict__ batom) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int q = bstart[bid[s]] + brank[s];
  if (q < n) batom[q] = s;
}

// Bound of the stencil values of atom s: the order-5 B-spline values, first
// and second derivatives are at most 0.599, 0.546 and 1.25, so every product
// of three factors in a stencil value is below 1 and the values are bounded by
// sum|m| (permanent fractional multipole) or sum|d| (fractional dipole).
__device__ __forceinline__ double spread_bound(const double *__restrict__ fmp, const double *__restrict__ v3,
                                               const Frac &fr, int s, bool ind) {
  if (!ind) {
    const double *m = fmp + 10 * (size_t)s;
    double b = 0.0;
#pragma unroll
    for (int k = 0; k < 10; k++) b += fabs(m[k]);
    return b;
  }
  const double *d = v3 + 3 * (size_t)s;
  double b = 0.0;
#pragma unroll
  for (int a = 0; a < 3; a++) b += fabs(fr.a[a][0] * d[0] + fr.a[a][1] * d[1] + fr.a[a][2] * d[2]);
  return b;
}

// One workgroup per brick: it sums the full 125-point stencils of the atoms
// binned in its brick into an LDS brick of edge ext (its points plus the
// 4-point spill into the next bricks) and stores it, in the grid's real type,
// to the brick buffer; the forward x transform assembles each grid point from
// its own brick and the spills of up to 7 previous bricks (brick_point), in a
// fixed order.  Thread (atom, ky, kz) forms the y-z weight and adds 5
// consecutive x points.
//
// The LDS sums are exact, so they do not depend on the order in which the
// atomics land and every evaluation of the same input gives the same grid,
// in the FP32 grid of the mixed form as in the FP64 form (a rounded
// floating-point sum changes with the order, and one FP32 ulp of a grid point
// moves the solved dipoles by up to ~1e-7).  Every stencil value is rounded
// to a multiple of a quantum before its FP64 LDS atomic add; while all
// partial sums stay below 2^53 quanta, FP64 adds of such multiples are exact.
// The quantum is chosen per brick and part: a point receives at most one
// value per atom, so with na atoms in the bin and bmax the largest bound of
// their values (spread_bound) every partial sum stays below na bmax < 2^e,
// and the quantum 2^(e-52) keeps it below 2^52 (+ na / 2 of rounding) quanta,
// with 52 - log2(na) significant bits against the largest value.
template <typename T, bool IND>
__global__ void __launch_bounds__(256) k_spread_brick(int n, Bricks bk, const int32_t *__restrict__ bstart,
                                                      const int32_t *__restrict__ batom, const T *__restrict__ theta,
                                                      const int32_t *__restrict__ igrid,
                                                      const double *__restrict__ fmp, const double *__restrict__ u3,
                                                      const double *__restrict__ up3, Frac fr, T *__restrict__ bbuf,
                                                      const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  constexpr int NP = IND ? 2 : 1, EM = NB_BRICK_EXT * NB_BRICK_EXT * NB_BRICK_EXT;
  __shared__ double lds[NP][EM];
  __shared__ double bmax_w[NP][4];
  const int e = bk.ext, e3 = e * e * e;
  const int b = blockIdx.x;
  const int a0 = bstart[b], na = min(bstart[b + 1], n) - a0;
  // the brick's first grid point per axis (runtime divisions once per block)
  int st[3];
  {
    const int bx = b % bk.nb[0], r = b / bk.nb[0];
    const int bb[3] = {bx, r % bk.nb[1], r / bk.nb[1]};
#pragma unroll
    for (int d = 0; d < 3; d++) st[d] = brick_start(bb[d], bk.K[d], bk.nb[d]);
  }
  for (int l = threadIdx.x; l < e3; l += blockDim.x)
#pragma unroll
    for (int p = 0; p < NP; p++) lds[p][l] = 0.0;
  double bm[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) bm[p] = 0.0;
  for (int a = threadIdx.x; a < na; a += blockDim.x) {
    const int s = batom[a0 + a];
#pragma unroll
    for (int p = 0; p < NP; p++) bm[p] = fmax(bm[p], spread_bound(fmp, p ? up3 : u3, fr, s, IND));
  }
#pragma unroll
  for (int p = 0; p < NP; p++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bm[p] = fmax(bm[p], __shfl_xor(bm[p], o, 64));
    if ((threadIdx.x & 63) == 0) bmax_w[p][threadIdx.x >> 6] = bm[p];
  }
  __syncthreads();
  double qs[NP], qi[NP];   // 2^(52 - e) and the quantum 2^(e - 52)
#pragma unroll
  for (int p = 0; p < NP; p++) {
    const double m = fmax(fmax(bmax_w[p][0], bmax_w[p][1]), fmax(bmax_w[p][2], bmax_w[p][3])) * (double)na;
    int ex = 0;
    if (m > 0.0) frexp(m, &ex);
    qs[p] = ldexp(1.0, 52 - ex);
    qi[p] = ldexp(1.0, ex - 52);
  }
  for (int q = threadIdx.x; q < na * 25; q += blockDim.x) {
    const int a = q / 25, yz = q - a * 25, ky = yz % 5, kz = yz / 5;
    const int s = batom[a0 + a];
    int o[3];   // local offset of the atom's first spline point inside the brick
#pragma unroll
    for (int d = 0; d < 3; d++) o[d] = wrapg(igrid[3 * s + d] + 1, bk.K[d]) - st[d];
    const T *tx = theta + (size_t)s * 60, *uu = tx + 20 + 4 * ky, *vv = tx + 40 + 4 * kz;
    const int base = ((o[2] + kz) * e + (o[1] + ky)) * e + o[0];
    // every global load is issued before the first LDS atomic: the compiler
    // does not move loads across atomics, so interleaving them serialises
    // five load latencies per thread
    T wx[5][3];
#pragma unroll
    for (int kx = 0; kx < 5; kx++) {
      wx[kx][0] = tx[4 * kx];
      wx[kx][1] = tx[4 * kx + 1];
      wx[kx][2] = IND ? (T)0 : tx[4 * kx + 2];
    }
    T val[NP][5];
    if (!IND) {
      const double *m = fmp + 10 * (size_t)s;
      T u0 = uu[0], u1 = uu[1], u2 = uu[2], v0 = vv[0], v1 = vv[1], v2 = vv[2];
      T t0 = ((T)m[0] * v0 + (T)m[3] * v1 + (T)m[6] * v2) * u0 + ((T)m[2] * v0 + (T)m[9] * v1) * u1 + (T)m[5] * v0 * u2;
      T t1 = ((T)m[1] * v0 + (T)m[8] * v1) * u0 + (T)m[7] * v0 * u1;
      T t2 = (T)m[4] * v0 * u0;
#pragma unroll
      for (int kx = 0; kx < 5; kx++) val[0][kx] = wx[kx][0] * t0 + wx[kx][1] * t1 + wx[kx][2] * t2;
    } else {
      T u0 = uu[0], u1 = uu[1], v0 = vv[0], v1 = vv[1];
#pragma unroll
      for (int p = 0; p < NP; p++) {
        const double *src3 = (p ? up3 : u3) + 3 * (size_t)s;
        T d0 = (T)(fr.a[0][0] * src3[0] + fr.a[0][1] * src3[1] + fr.a[0][2] * src3[2]);
        T d1 = (T)(fr.a[1][0] * src3[0] + fr.a[1][1] * src3[1] + fr.a[1][2] * src3[2]);
        T d2 = (T)(fr.a[2][0] * src3[0] + fr.a[2][1] * src3[1] + fr.a[2][2] * src3[2]);
        T t0 = d2 * v1 * u0 + d1 * v0 * u1, t1 = d0 * v0 * u0;
#pragma unroll
        for (int kx = 0; kx < 5; kx++) val[p][kx] = wx[kx][0] * t0 + wx[kx][1] * t1;
      }
    }
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
      for (int kx = 0; kx < 5; kx++)
        atomicAdd(&lds[p][base + kx], rint((double)val[p][kx] * qs[p]));
  }
  __syncthreads();
  T *out = bbuf + (size_t)b * NP * e3;
  for (int l = threadIdx.x; l < e3; l += blockDim.x)
#pragma unroll
    for (int p = 0; p < NP; p++) out[(size_t)p * e3 + l] = (T)(lds[p][l] * qi[p]);
}

// grid point (x, y, z) = own brick at local offset o plus, along each axis
// where o < 4, the spill of the previous brick at local offset len_prev + o
template <typename T, int NP>
__device__ __forceinline__ CX<T> brick_point(const Bricks &bk, const T *__restrict__ bbuf, int cx, int cy, int cz) {
  int c[3] = {cx, cy, cz};
  int bo[3], lo[3], bp[3], lp[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    int K = bk.K[d], nb = bk.nb[d];
    bo[d] = brick_of(c[d], K, nb);
    int st = brick_start(bo[d], K, nb);
    lo[d] = c[d] - st;
    bp[d] = bo[d] == 0 ? nb - 1 : bo[d] - 1;
    int lenp = (bo[d] == 0 ? K : st) - brick_start(bp[d], K, nb);
    lp[d] = lenp + lo[d];
  }
  const int e = bk.ext, e3 = e * e * e;
  T acc[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) acc[p] = (T)0;
#pragma unroll
  for (int m = 0; m < 8; m++) {
    int dx = m & 1, dy = (m >> 1) & 1, dz = m >> 2;
    if ((dx && lo[0] >= 4) || (dy && lo[1] >= 4) || (dz && lo[2] >= 4)) continue;
    int bx = dx ? bp[0] : bo[0], by = dy ? bp[1] : bo[1], bz = dz ? bp[2] : bo[2];
    int x = dx ? lp[0] : lo[0], y = dy ? lp[1] : lo[1], z = dz ? lp[2] : lo[2];
    size_t bi = (size_t)bx + (size_t)bk.nb[0] * (by + (size_t)bk.nb[1] * bz);
    const T *src = bbuf + bi * NP * e3 + (z * e + y) * e + x;
#pragma unroll
    for (int p = 0; p < NP; p++) acc[p] += src[(size_t)p * e3];
  }
  CX<T> v;
  v.x = acc[0];
  v.y = NP == 2 ? acc[NP - 1] : (T)0;
  return v;
}

// the grid assembled from the brick buffer (for lengths without a fused x
// transform)
template <typename T, int NP>
__global__ void __launch_bounds__(256) k_brick_combine(Bricks bk, const T *__restrict__ bbuf, CX<T> *__restrict__ grid,
                                                       const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long ntot = (long)bk.K[0] * bk.K[1] * bk.K[2];
  if (g >= ntot) return;
  int cx = (int)(g % bk.K[0]);
  long r = g / bk.K[0];
  grid[g] = brick_point<T, NP>(bk, bbuf, cx, (int)(r % bk.K[1]), (int)(r / bk.K[1]));
}

// Gather: 8 lanes per atom, lane kz < 5 sums one z plane (25 rows of 5
// points) into fractional derivatives, then the lanes are combined with
// shuffles.  FULL: all 20 derivatives t000..t111 plus the Cartesian
// transform (out[30 s] = 20 fractional + 10 Cartesian).  Otherwise only the
// first derivatives are formed and the Cartesian field components are
// written to out[30 s + 21..23] (what the PCG iterations consume).
template <typename T, int NP, bool FULL>
__global__ void __launch_bounds__(256) k_gather(int n, const T *__restrict__ theta, const int32_t *__restrict__ igrid,
                                                const CX<T> *__restrict__ grid, Frac fr, int K0, int K1, int K2,
                                                double *__restrict__ out0, double *__restrict__ out1,
                                                const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt >> 3, lane = gt & 7;
  gather_atom<T, NP, FULL, false, T>(s, lane, s < n, theta, igrid, grid, fr, K0, K1, K2, out0, out1);
}

// The gather of a PCG iteration fused with the conjugation: after the field
// of z is gathered (the real-space part f is complete when this runs), lane 0
// of each atom forms E(z) = f + a3 z - phi, A z = z / polar - E(z), and the
// new direction p = z + beta p with its image q = A z + beta q
// (pcg_direction); partial p.q into S_PAP / S_PAPP.
template <typename T>
__global__ void __launch_bounds__(256) k_gather_ap(int n, const T *__restrict__ theta,
                                                   const int32_t *__restrict__ igrid, const CX<T> *__restrict__ grid,
                                                   Frac fr, int K0, int K1, int K2, double *__restrict__ phiu,
                                                   double *__restrict__ phiup, double a3,
                                                   const double *__restrict__ polar, const double *__restrict__ z,
                                                   const double *__restrict__ zp, const double *__restrict__ f,
                                                   const double *__restrict__ fp, double *__restrict__ p,
                                                   double *__restrict__ pp, double *__restrict__ vec,
                                                   double *__restrict__ vecp, double *__restrict__ part, int stride,
                                                   int nrz, const double *__restrict__ scal, int first,
                                                   const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  int gt = xcd_block() * blockDim.x + threadIdx.x;
  int s = gt >> 3, lane = gt & 7;
  // lane d < 3 of the atom's 8 updates component d; its vectors are loaded
  // first, so their latency overlaps the totals and the grid gather
  const bool mine = s < n && lane < 3;
  const size_t q = 3 * (size_t)s + (lane < 3 ? lane : 0);
  double a = 0.0, zq = 0.0, zpq = 0.0, fq = 0.0, fpq = 0.0, pn = 0.0, qn = 0.0, pnp = 0.0, qnp = 0.0;
  if (mine) {
    a = polar[s];
    zq = z[q];
    zpq = zp[q];
    fq = f[q];
    fpq = fp[q];
    pn = p[q];
    qn = vec[q];
    pnp = pp[q];
    qnp = vecp[q];
  }
  double beta[2];
  pcg_beta(part, stride, nrz, scal, first, beta);
  double fld[6];
  gather_atom<T, 2, false, false, T>(s, lane, s < n, theta, igrid, grid, fr, K0, K1, K2, phiu, phiup, fld);
  double v[2] = {0.0, 0.0};
  if (mine) {
    const int d = lane;
    pcg_direction(a, zq, fq + a3 * zq - fld[d], beta[0], first, pn, qn);
    pcg_direction(a, zpq, fpq + a3 * zpq - fld[3 + d], beta[1], first, pnp, qnp);
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

// lines along one axis, L lines per workgroup (fft_chunk of chunk blockIdx.x)
template <typename T, int EPT, int AXIS, bool CONV>
__global__ void __launch_bounds__(256) k_fft_lines(CX<T> *__restrict__ grid, int K0, int K1, int K2, int L, Radices rad,
                                                   const CX<T> *__restrict__ twg, T s, const T *__restrict__ recip,
                                                   const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int N = AXIS == 0 ? K0 : (AXIS == 1 ? K1 : K2);
  CX<T> *buf = (CX<T> *)smem_raw, *tw = buf + (size_t)L * N;
  for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
  fft_chunk<T, EPT, AXIS, CONV>(grid, K0, K1, K2, L, blockIdx.x, rad, buf, tw, s, recip);
}

Radices factor(int N) {
  Radices r;
  r.nr = 0;
  int m = N;
  while (m % 4 == 0) { r.r[r.nr++] = 4; m /= 4; }
  while (m % 2 == 0) { r.r[r.nr++] = 2; m /= 2; }
  while (m % 3 == 0) { r.r[r.nr++] = 3; m /= 3; }
  while (m % 5 == 0) { r.r[r.nr++] = 5; m /= 5; }
  while (m % 7 == 0) { r.r[r.nr++] = 7; m /= 7; }
  if (m != 1) throw std::runtime_error("PME grid dimension must factor over 2, 3, 5, 7");
  return r;
}

constexpr int LINE_EPT = 8;    // elements per thread of the line kernels (256 threads)

int lines_per_block(int N) {
  int L = 256 * LINE_EPT / N;
  if (L > 16) L = 16;
  if (L < 1) throw std::runtime_error("PME grid dimension too large for the line FFT");
  return L;
}


// ---------------------------------------------------------------------------
// Line FFTs with the length N, the lines per workgroup and the block size
// known at compile time (every index division is by a constant and the radix
// schedule is unrolled).  Used for the grid sizes listed in NB_CT_SIZES; other
// sizes take the runtime-N kernels above.
// ---------------------------------------------------------------------------
constexpr int CT_BS = 256;

template <int M> struct RadixOf {
  static constexpr int v = (M % 4 == 0) ? 4 : (M % 2 == 0) ? 2 : (M % 3 == 0) ? 3 : (M % 5 == 0) ? 5 : 7;
};
template <int N> struct CtLines {
  static constexpr int raw = 1024 / N;
  static constexpr int x = raw < 1 ? 1 : (raw > 32 ? 32 : raw);
  static constexpr int yz = raw < 1 ? 1 : (raw > 16 ? 16 : raw);
};

template <int R, int N, int Ns, int NL, int SL, int SE, bool LF, typename T, int BS = CT_BS>
__device__ __forceinline__ void ct_pass(CX<T> *buf, const CX<T> *tw, T s) {
  constexpr int nb = N / R, total = NL * nb, stride = N / (Ns * R);
  constexpr int MAXB = (total + BS - 1) / BS;
  CX<T> v[MAXB][R];
  int where[MAXB];
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    const int b = q * BS + threadIdx.x;
    where[q] = -1;
    if (b < total) {
      const int line = LF ? b % NL : b / nb;
      const int j = LF ? b / NL : b % nb;
      const int jm = j % Ns;
      const CX<T> *x = buf + line * SL;
#pragma unroll
      for (int k = 0; k < R; k++) {
        CX<T> a = x[(j + k * nb) * SE];
        if (Ns > 1 && k > 0) {  // tw[0] = 1 exactly, so jm = 0 needs no test
          CX<T> w = tw[jm * k * stride];
          w.y *= s;
          a = cmul(a, w);
        }
        v[q][k] = a;
      }
      small_dft<R, T>(v[q], s);
      where[q] = line * SL + ((j / Ns) * Ns * R + jm) * SE;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    if (where[q] >= 0) {
#pragma unroll
      for (int k = 0; k < R; k++) buf[where[q] + k * Ns * SE] = v[q][k];
    }
  }
  __syncthreads();
}

template <int N, int Ns, int NL, int SL, int SE, bool LF, typename T, int BS = CT_BS>
__device__ __forceinline__ void ct_fft(CX<T> *buf, const CX<T> *tw, T s) {
  if constexpr (Ns < N) {
    constexpr int R = RadixOf<N / Ns>::v;
    ct_pass<R, N, Ns, NL, SL, SE, LF, T, BS>(buf, tw, s);
    ct_fft<N, Ns * R, NL, SL, SE, LF, T, BS>(buf, tw, s);
  }
}

// Whole XY planes of the FP32 grid in LDS (rows padded by one element), one
// workgroup of PL_BS threads per z: the forward kernel (INV false) assembles
// the plane from the brick buffer (MODE 1 / 2) or reads it (MODE 0) and
// transforms x then y; the inverse kernel transforms y then x.  With the
// fused z kernel between them a convolution takes three launches instead of
// five (x, y, z, y, x), for planes of up to 64 KB of LDS (FP32 complex).
constexpr int PL_BS = 1024;
template <typename T, int N0, int N1, int MODE, bool INV>
__global__ void __launch_bounds__(PL_BS) k_plane(CX<T> *__restrict__ grid, const CX<T> *__restrict__ twx,
                                                 const CX<T> *__restrict__ twy, T s, Bricks bk,
                                                 const T *__restrict__ bbuf, const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  constexpr int SL = N0 + 1, TOT = N0 * N1, PER = (TOT + PL_BS - 1) / PL_BS;
  __shared__ CX<T> buf[N1 * SL];
  __shared__ CX<T> tx[N0], ty[N1];
  const int z = blockIdx.x;
  CX<T> *plane = grid + (size_t)TOT * z;
  for (int t = threadIdx.x; t < N0; t += PL_BS) tx[t] = twx[t];
  for (int t = threadIdx.x; t < N1; t += PL_BS) ty[t] = twy[t];
  // all loads of the thread are issued before the first LDS store
  CX<T> v[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int t = q * PL_BS + threadIdx.x;
    if (t < TOT) {
      if constexpr (MODE == 0) v[q] = plane[t];
      else v[q] = brick_point<T, MODE>(bk, bbuf, t % N0, t / N0, z);
    }
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int t = q * PL_BS + threadIdx.x;
    if (t < TOT) buf[(t / N0) * SL + t % N0] = v[q];
  }
  __syncthreads();
  if (!INV) {
    ct_fft<N0, 1, N1, SL, 1, false, T, PL_BS>(buf, tx, s);
    ct_fft<N1, 1, N0, 1, SL, true, T, PL_BS>(buf, ty, s);
  } else {
    ct_fft<N1, 1, N0, 1, SL, true, T, PL_BS>(buf, ty, s);
    ct_fft<N0, 1, N1, SL, 1, false, T, PL_BS>(buf, tx, s);
  }
  for (int t = threadIdx.x; t < TOT; t += PL_BS) plane[t] = buf[(t / N0) * SL + t % N0];
}

// x lines (contiguous), rows padded by one element in LDS.  MODE 0 reads the
// grid; MODE 1 / 2 assembles each point from the brick buffer (1 or 2 parts),
// which fuses the brick combine into the forward transform.
template <typename T, int N, int MODE>
__global__ void __launch_bounds__(CT_BS) k_ct_x(CX<T> *__restrict__ grid, int K1, int K2,
                                                const CX<T> *__restrict__ twg, T s, Bricks bk,
                                                const T *__restrict__ bbuf, const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  constexpr int NL = CtLines<N>::x, SL = N + 1;
  __shared__ CX<T> buf[NL * SL];
  __shared__ CX<T> tw[N];
  constexpr int TOT = NL * N, PER = (TOT + CT_BS - 1) / CT_BS;
  const int first = blockIdx.x * NL, nlines = K1 * K2;
  for (int t = threadIdx.x; t < N; t += CT_BS) tw[t] = twg[t];
  // all loads of the thread are issued before the first LDS store
  T re[PER], im[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int t = q * CT_BS + threadIdx.x, l = t / N, e = t - l * N, line = first + l;
    re[q] = (T)0;
    im[q] = (T)0;
    if (t < TOT && line < nlines) {
      CX<T> v;
      if constexpr (MODE == 0) v = grid[(size_t)line * N + e];
      else v = brick_point<T, MODE>(bk, bbuf, e, line % K1, line / K1);
      re[q] = v.x;
      im[q] = v.y;
    }
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int t = q * CT_BS + threadIdx.x, l = t / N, e = t - l * N;
    if (t < TOT) buf[l * SL + e] = CX<T>{re[q], im[q]};
  }
  __syncthreads();
  ct_fft<N, 1, NL, SL, 1, false, T>(buf, tw, s);
  for (int t = threadIdx.x; t < NL * N; t += CT_BS) {
    int l = t / N, e = t - l * N, line = first + l;
    if (line < nlines) grid[(size_t)line * N + e] = buf[l * SL + e];
  }
}

// y lines: NL consecutive x at one z
template <typename T, int N>
__global__ void __launch_bounds__(CT_BS) k_ct_y(CX<T> *__restrict__ grid, int K0, const CX<T> *__restrict__ twg, T s,
                                                const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  constexpr int NL = CtLines<N>::yz;
  __shared__ CX<T> buf[NL * N];
  __shared__ CX<T> tw[N];
  const int nxb = (K0 + NL - 1) / NL, xb = (blockIdx.x % nxb) * NL, z = blockIdx.x / nxb;
  const int cnt = min(NL, K0 - xb);
  constexpr int TOT = NL * N, PER = (TOT + CT_BS - 1) / CT_BS;
  for (int t = threadIdx.x; t < N; t += CT_BS) tw[t] = twg[t];
  T re[PER], im[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int t = q * CT_BS + threadIdx.x, e = t / NL, l = t - e * NL;
    re[q] = (T)0;
    im[q] = (T)0;
    if (t < TOT && l < cnt) {
      CX<T> v = grid[(size_t)(xb + l) + (size_t)K0 * (e + (size_t)N * z)];
      re[q] = v.x;
      im[q] = v.y;
    }
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int t = q * CT_BS + threadIdx.x;
    if (t < TOT) buf[t] = CX<T>{re[q], im[q]};
  }
  __syncthreads();
  ct_fft<N, 1, NL, 1, NL, true, T>(buf, tw, s);
  for (int t = threadIdx.x; t < NL * N; t += CT_BS) {
    int e = t / NL, l = t - e * NL;
    if (l < cnt) grid[(size_t)(xb + l) + (size_t)K0 * (e + (size_t)N * z)] = buf[t];
  }
}

// z lines: forward, influence multiply, un-normalised inverse (CONV), or the
// forward transform alone
template <typename T, int N, bool CONV>
__global__ void __launch_bounds__(CT_BS) k_ct_z(CX<T> *__restrict__ grid, int K01, const CX<T> *__restrict__ twg,
                                                const T *__restrict__ recip, const double *__restrict__ ctl) {
  if (ctl && ctl[0] != 0.0) return;  // recorded solve already stopped
  constexpr int NL = CtLines<N>::yz;
  __shared__ CX<T> buf[NL * N];
  __shared__ CX<T> tw[N];
  const int first = blockIdx.x * NL, cnt = min(NL, K01 - first);
  constexpr int TOT = NL * N, PER = (TOT + CT_BS - 1) / CT_BS;
  for (int t = threadIdx.x; t < N; t += CT_BS) tw[t] = twg[t];
  T re[PER], im[PER], gr[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int t = q * CT_BS + threadIdx.x, e = t / NL, l = t - e * NL;
    re[q] = (T)0;
    im[q] = (T)0;
    gr[q] = (T)0;
    if (t < TOT && l < cnt) {
      size_t gi = (size_t)(first + l) + (size_t)K01 * e;
      CX<T> v = grid[gi];
      re[q] = v.x;
      im[q] = v.y;
      if (CONV) gr[q] = recip[gi];
    }
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    int t = q * CT_BS + threadIdx.x;
    if (t < TOT) buf[t] = CX<T>{re[q], im[q]};
  }
  __syncthreads();
  ct_fft<N, 1, NL, 1, NL, true, T>(buf, tw, (T)1);
  if (CONV) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      int t = q * CT_BS + threadIdx.x;
      if (t < TOT) {
        buf[t].x *= gr[q];
        buf[t].y *= gr[q];
      }
    }
    __syncthreads();
    ct_fft<N, 1, NL, 1, NL, true, T>(buf, tw, (T)-1);
  }
  for (int t = threadIdx.x; t < NL * N; t += CT_BS) {
    int e = t / NL, l = t - e * NL;
    if (l < cnt) grid[(size_t)(first + l) + (size_t)K01 * e] = buf[t];
  }
}

#define NB_CT_SIZES(X) X(16) X(20) X(24) X(30) X(32) X(40) X(48) X(60) X(64) X(72) X(80) X(96) X(120) X(128) X(144)

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

static Bricks bricks_of(const NbHandle *h) {
  Bricks b;
  for (int d = 0; d < 3; d++) { b.K[d] = h->K[d]; b.nb[d] = h->nbk[d]; }
  b.ext = h->bext;
  return b;
}

// influence function of the reference (generalizedInfluenceFunction,
// ReciprocalSpace.java:1045-1121; discreteFTMod :435-492), computed on the host
static void bspline_modulus(std::vector<double> &bsmod, int n, int order) {
  // weights at x = 0 from the same recursion as the kernels
  double w[8][8] = {{0}};
  w[1][1] = 1.0;
  for (int p = 2; p <= order; p++)
    for (int k = 0; k < p; k++) w[p][k + 1] = ((0.0 + p - 1 - k) * w[p - 1][k] + (k + 1 - 0.0) * w[p - 1][k + 1]) / (p - 1);
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

static void twiddles(const int K[3], std::vector<double> &tw) {
  tw.assign(2 * (K[0] + K[1] + K[2]), 0.0);
  int off = 0;
  for (int d = 0; d < 3; d++) {
    for (int j = 0; j < K[d]; j++) {
      double a = 2.0 * M_PI * j / K[d];  // exp(+2 pi i j / K)
      tw[2 * (off + j)] = cos(a);
      tw[2 * (off + j) + 1] = sin(a);
    }
    off += K[d];
  }
}

static void free_grid(NbHandle *h) {
  void *ptrs[] = {h->d_grid, h->d_recip, h->d_recipf, h->d_tw, h->d_twf, h->d_bcount, h->d_bstart,
                  h->d_bbuf, h->d_fgrid};
  for (void *p : ptrs)
    if (p) hipFree(p);
  h->d_fgrid = nullptr;
  h->d_grid = nullptr; h->d_recip = nullptr; h->d_recipf = nullptr; h->d_tw = nullptr; h->d_twf = nullptr;
  h->d_bcount = nullptr; h->d_bstart = nullptr; h->d_bbuf = nullptr;
}

void nb_pme_setup(NbHandle *h) {
  int K0 = h->prm.grid[0], K1 = h->prm.grid[1], K2 = h->prm.grid[2];
  int K[3] = {K0, K1, K2};
  for (int d = 0; d < 3; d++)
    if (K[d] < 8) throw std::runtime_error("PME grid dimensions must be at least 8");
  for (int d = 0; d < 3; d++) factor(K[d]);
  long ntot = (long)K0 * K1 * K2;
  if (ntot != h->ntot || !h->d_grid) {
    free_grid(h);
    // bricks: at least 2 per axis, edges of at most NB_BRICK_MAX and at
    // least 4 points (the 4-point spill of a brick lands in the next brick
    // only); the LDS brick of the spread holds the longest edge + 4
    int nbk[3], ext = 0;
    for (int d = 0; d < 3; d++) {
      nbk[d] = std::max(2, (K[d] + NB_BRICK_MAX - 1) / NB_BRICK_MAX);
      for (int b = 0; b < nbk[d]; b++) {
        int len = brick_start(b + 1, K[d], nbk[d]) - brick_start(b, K[d], nbk[d]);
        if (len > NB_BRICK_MAX || len < 4) throw std::runtime_error("PME brick edge out of range");
        ext = std::max(ext, len + 4);
      }
      h->nbk[d] = nbk[d];
    }
    h->bext = ext;
    h->nbricks = nbk[0] * nbk[1] * nbk[2];
    HIPCHECK(hipMalloc(&h->d_grid, sizeof(double) * 2 * ntot));
    HIPCHECK(hipMalloc(&h->d_recip, sizeof(double) * ntot));
    HIPCHECK(hipMalloc(&h->d_recipf, sizeof(float) * ntot));
    HIPCHECK(hipMalloc(&h->d_tw, sizeof(double) * 2 * (K0 + K1 + K2)));
    HIPCHECK(hipMalloc(&h->d_twf, sizeof(float) * 2 * (K0 + K1 + K2)));
    HIPCHECK(hipMalloc(&h->d_bcount, sizeof(int32_t) * h->nbricks));
    HIPCHECK(hipMalloc(&h->d_bstart, sizeof(int32_t) * (h->nbricks + 1)));
    HIPCHECK(hipMalloc(&h->d_bbuf, sizeof(double) * 2 * (size_t)h->nbricks * ext * ext * ext));
  }
  h->K[0] = K0; h->K[1] = K1; h->K[2] = K2;
  h->ntot = ntot;
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
  std::vector<float> recf(rec.begin(), rec.end());
  HIPCHECK(hipMemcpy(h->d_recip, rec.data(), sizeof(double) * ntot, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(h->d_recipf, recf.data(), sizeof(float) * ntot, hipMemcpyHostToDevice));
  std::vector<double> tw;
  twiddles(K, tw);
  std::vector<float> twf(tw.begin(), tw.end());
  HIPCHECK(hipMemcpy(h->d_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(h->d_twf, twf.data(), sizeof(float) * twf.size(), hipMemcpyHostToDevice));
  h->recip_valid = true;
}

// Compile-time line kernels when the axis length is one of NB_CT_SIZES,
// runtime-N line kernels otherwise.  Each returns false when it has no
// specialisation for the length.
template <typename T>
static bool ct_x(NbHandle *h, int mode, CX<T> *g, const CX<T> *tw, T s) {
  int K1 = h->K[1], K2 = h->K[2];
  const Bricks bk = bricks_of(h);
  const T *bb = (const T *)h->d_bbuf;
  switch (h->K[0]) {
#define NB_X(N)                                                                                       \
  case N: {                                                                                           \
    dim3 G((K1 * K2 + CtLines<N>::x - 1) / CtLines<N>::x);                                            \
    if (mode == 0) NB_LAUNCH((k_ct_x<T, N, 0>), G, dim3(CT_BS), 0, g, K1, K2, tw, s, bk, bb, h->ctl); \
    else if (mode == 1) NB_LAUNCH((k_ct_x<T, N, 1>), G, dim3(CT_BS), 0, g, K1, K2, tw, s, bk, bb, h->ctl); \
    else NB_LAUNCH((k_ct_x<T, N, 2>), G, dim3(CT_BS), 0, g, K1, K2, tw, s, bk, bb, h->ctl); \
    return true;                                                                                      \
  }
    NB_CT_SIZES(NB_X)
#undef NB_X
    default: return false;
  }
}

template <typename T>
static bool ct_y(NbHandle *h, CX<T> *g, const CX<T> *tw, T s) {
  int K0 = h->K[0], K2 = h->K[2];
  switch (h->K[1]) {
#define NB_Y(N)                                                                                      \
  case N: {                                                                                          \
    dim3 G(((K0 + CtLines<N>::yz - 1) / CtLines<N>::yz) * K2);                                       \
    NB_LAUNCH((k_ct_y<T, N>), G, dim3(CT_BS), 0, g, K0, tw, s, h->ctl);                                \
    return true;                                                                                     \
  }
    NB_CT_SIZES(NB_Y)
#undef NB_Y
    default: return false;
  }
}

template <typename T>
static bool ct_z(NbHandle *h, CX<T> *g, const CX<T> *tw, const T *recip) {
  int K01 = h->K[0] * h->K[1];
  switch (h->K[2]) {
#define NB_Z(N)                                                                                      \
  case N: {                                                                                          \
    dim3 G((K01 + CtLines<N>::yz - 1) / CtLines<N>::yz);                                             \
    if (recip) NB_LAUNCH((k_ct_z<T, N, true>), G, dim3(CT_BS), 0, g, K01, tw, recip, h->ctl);                \
    else NB_LAUNCH((k_ct_z<T, N, false>), G, dim3(CT_BS), 0, g, K01, tw, recip, h->ctl);                    \
    return true;                                                                                     \
  }
    NB_CT_SIZES(NB_Z)
#undef NB_Z
    default: return false;
  }
}

// XY-plane kernels for square FP32 planes of these edges; returns false
// (nothing launched) for other grids and for FP64 grids
#define NB_PLANE_SIZES(X) X(24) X(32) X(40) X(48) X(60) X(64) X(72) X(80)
template <typename T>
static bool plane_t(NbHandle *h, int np, bool inv, CX<T> *g, const CX<T> *twx, const CX<T> *twy, T s) {
  if (sizeof(T) != sizeof(float) || h->K[0] != h->K[1]) return false;
  const Bricks bk = bricks_of(h);
  const T *bb = (const T *)h->d_bbuf;
  dim3 G(h->K[2]);
  switch (h->K[0]) {
#define NB_P(N)                                                                                               \
  case N:                                                                                                     \
    if (inv) NB_LAUNCH((k_plane<T, N, N, 0, true>), G, dim3(PL_BS), 0, g, twx, twy, s, bk, bb, h->ctl);        \
    else if (np == 0) NB_LAUNCH((k_plane<T, N, N, 0, false>), G, dim3(PL_BS), 0, g, twx, twy, s, bk, bb, h->ctl); \
    else if (np == 1) NB_LAUNCH((k_plane<T, N, N, 1, false>), G, dim3(PL_BS), 0, g, twx, twy, s, bk, bb, h->ctl); \
    else NB_LAUNCH((k_plane<T, N, N, 2, false>), G, dim3(PL_BS), 0, g, twx, twy, s, bk, bb, h->ctl);           \
    return true;
    NB_PLANE_SIZES(NB_P)
#undef NB_P
    default: return false;
  }
}

// forward x (assembling the grid from the bricks when np > 0), forward y,
// fused z (forward * influence * inverse), inverse y, inverse x; with XY-plane
// kernels: forward plane, fused z, inverse plane
template <typename T>
static void transform_t(NbHandle *h, int np, CX<T> *g, const CX<T> *tw0, const T *recip) {
  int K0 = h->K[0], K1 = h->K[1], K2 = h->K[2];
  const CX<T> *twx = tw0, *twy = twx + K0, *twz = twy + K1;
  const Bricks bk = bricks_of(h);
  const T *bb = (const T *)h->d_bbuf;
  int Lx = lines_per_block(K0), Ly = lines_per_block(K1), Lz = lines_per_block(K2);
  size_t shx = sizeof(CX<T>) * ((size_t)Lx * K0 + K0), shy = sizeof(CX<T>) * ((size_t)Ly * K1 + K1);
  size_t shz = sizeof(CX<T>) * ((size_t)Lz * K2 + K2);
  dim3 B(256);
  dim3 Gx((K1 * K2 + Lx - 1) / Lx), Gy(((K0 + Ly - 1) / Ly) * K2), Gz((K0 * K1 + Lz - 1) / Lz);
  const T *nul = nullptr;
  if (plane_t<T>(h, np, false, g, twx, twy, (T)1)) {
    if (!ct_z<T>(h, g, twz, recip))
      NB_LAUNCH((k_fft_lines<T, LINE_EPT, 2, true>), Gz, B, shz, g, K0, K1, K2, Lz, factor(K2), twz, (T)1, recip, h->ctl);
    plane_t<T>(h, 0, true, g, twx, twy, (T)-1);
    return;
  }
  for (int pass = 0; pass < 2; pass++) {
    T s = pass == 0 ? (T)1 : (T)-1;
    if (pass == 0 && !ct_x<T>(h, np, g, twx, s)) {
      if (np == 1) NB_LAUNCH((k_brick_combine<T, 1>), grid_for(h->ntot, 256), dim3(256), 0, bk, bb, g, h->ctl);
      if (np == 2) NB_LAUNCH((k_brick_combine<T, 2>), grid_for(h->ntot, 256), dim3(256), 0, bk, bb, g, h->ctl);
      NB_LAUNCH((k_fft_lines<T, LINE_EPT, 0, false>), Gx, B, shx, g, K0, K1, K2, Lx, factor(K0), twx, s, nul, h->ctl);
    }
    if (!ct_y<T>(h, g, twy, s))
      NB_LAUNCH((k_fft_lines<T, LINE_EPT, 1, false>), Gy, B, shy, g, K0, K1, K2, Ly, factor(K1), twy, s, nul, h->ctl);
    if (pass == 0) {
      if (!ct_z<T>(h, g, twz, recip))
        NB_LAUNCH((k_fft_lines<T, LINE_EPT, 2, true>), Gz, B, shz, g, K0, K1, K2, Lz, factor(K2), twz, (T)1, recip, h->ctl);
    } else if (!ct_x<T>(h, 0, g, twx, s)) {
      NB_LAUNCH((k_fft_lines<T, LINE_EPT, 0, false>), Gx, B, shx, g, K0, K1, K2, Lx, factor(K0), twx, s, nul, h->ctl);
    }
  }
}

// forward transform only (x assembled from the bricks when np > 0), used for
// the strain derivative
template <typename T>
static void forward_t(NbHandle *h, int np, CX<T> *g, const CX<T> *tw0) {
  int K0 = h->K[0], K1 = h->K[1], K2 = h->K[2];
  const CX<T> *twx = tw0, *twy = twx + K0, *twz = twy + K1;
  const Bricks bk = bricks_of(h);
  const T *bb = (const T *)h->d_bbuf;
  int Lx = lines_per_block(K0), Ly = lines_per_block(K1), Lz = lines_per_block(K2);
  size_t shx = sizeof(CX<T>) * ((size_t)Lx * K0 + K0), shy = sizeof(CX<T>) * ((size_t)Ly * K1 + K1);
  size_t shz = sizeof(CX<T>) * ((size_t)Lz * K2 + K2);
  dim3 B(256), Gx((K1 * K2 + Lx - 1) / Lx), Gy(((K0 + Ly - 1) / Ly) * K2), Gz((K0 * K1 + Lz - 1) / Lz);
  const T *nul = nullptr;
  if (plane_t<T>(h, np, false, g, twx, twy, (T)1)) {
    if (!ct_z<T>(h, g, twz, nul))
      NB_LAUNCH((k_fft_lines<T, LINE_EPT, 2, false>), Gz, B, shz, g, K0, K1, K2, Lz, factor(K2), twz, (T)1, nul, h->ctl);
    return;
  }
  if (!ct_x<T>(h, np, g, twx, (T)1)) {
    if (np == 1) NB_LAUNCH((k_brick_combine<T, 1>), grid_for(h->ntot, 256), dim3(256), 0, bk, bb, g, h->ctl);
    if (np == 2) NB_LAUNCH((k_brick_combine<T, 2>), grid_for(h->ntot, 256), dim3(256), 0, bk, bb, g, h->ctl);
    NB_LAUNCH((k_fft_lines<T, LINE_EPT, 0, false>), Gx, B, shx, g, K0, K1, K2, Lx, factor(K0), twx, (T)1, nul, h->ctl);
  }
  if (!ct_y<T>(h, g, twy, (T)1))
    NB_LAUNCH((k_fft_lines<T, LINE_EPT, 1, false>), Gy, B, shy, g, K0, K1, K2, Ly, factor(K1), twy, (T)1, nul, h->ctl);
  if (!ct_z<T>(h, g, twz, nul))
    NB_LAUNCH((k_fft_lines<T, LINE_EPT, 2, false>), Gz, B, shz, g, K0, K1, K2, Lz, factor(K2), twz, (T)1, nul, h->ctl);
}

// Spread into the brick buffer (IND: u and uCR of the dipoles u3 / up3,
// otherwise the permanent fractional multipoles); theta of type T unless
// given.  The next forward transform (np = 1, or 2 with IND) assembles the grid.
template <typename T, bool IND>
static void spread_into(NbHandle *h, const double *u, const double *up, const T *theta = nullptr) {
  NB_LAUNCH((k_spread_brick<T, IND>), dim3(h->nbricks), dim3(256), 0, h->n, bricks_of(h), h->d_bstart, h->d_batom,
            theta ? theta : (const T *)h->d_theta, h->d_igrid, h->d_fmp, u, up, frac_of(h), (T *)h->d_bbuf, h->ctl);
}

// Reciprocal strain derivative, grid part: at every wave vector m the energy
// weight w(m) = el/2 recip(m) [|Fp|^2 + Re Fp*(Fu + Fv) (+ Re Fu* Fv, mutual)]
// times 2 (1 + pi^2 m^2 / a^2) m_a m_b / m^2 - delta_ab.  fp: transform of the
// permanent grid; fi: transform of the induced grid (u real, uCR imaginary;
// the two are separated with F(-m) = conj F(m) of a real grid) or null.
template <typename T>
__global__ void __launch_bounds__(256) k_grid_virial(int K0, int K1, int K2, const CX<T> *__restrict__ fp,
                                                     const CX<T> *__restrict__ fi, int mutual,
                                                     const double *__restrict__ recip, Frac2v A, double pt, double el,
                                                     double sperm, double spol, double *__restrict__ vpart) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long ntot = (long)K0 * K1 * K2;
  double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (g < ntot && recip[g] != 0.0) {
    int kx = (int)(g % K0);
    long r1 = g / K0;
    int ky = (int)(r1 % K1), kz = (int)(r1 / K1);
    int hh = kx >= (K0 + 1) / 2 ? kx - K0 : kx, k = ky >= (K1 + 1) / 2 ? ky - K1 : ky;
    int l = kz >= (K2 + 1) / 2 ? kz - K2 : kz;
    double m[3];
    for (int d = 0; d < 3; d++) m[d] = A.a[d][0] * hh + A.a[d][1] * k + A.a[d][2] * l;
    double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    double pr = (double)fp[g].x, pim = (double)fp[g].y;
    double ws = sperm * (pr * pr + pim * pim);
    if (fi) {
      long ng = (long)((K0 - kx) % K0) + (long)K0 * (((K1 - ky) % K1) + (long)K1 * ((K2 - kz) % K2));
      double ar = (double)fi[g].x, ai = (double)fi[g].y, br = (double)fi[ng].x, bi = -(double)fi[ng].y;
      double ur = 0.5 * (ar + br), ui = 0.5 * (ai + bi);
      double vr = 0.5 * (ai - bi), vi = -0.5 * (ar - br);
      double wq = pr * (ur + vr) + pim * (ui + vi);
      if (mutual) wq += ur * vr + ui * vi;
      ws += spol * wq;
    }
    double w = 0.5 * el * recip[g] * ws, f = 2.0 * (1.0 + pt * m2) / m2;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) w9[3 * a + b] = w * (f * m[a] * m[b] - (a == b ? 1.0 : 0.0));
  }
  block_store9(w9, vpart);
}

// The mutual part alone for a pair of dipole sets packed as one grid (a real,
// b imaginary): w = scale recip(m) Re Fa* Fb (2 (1 + pi^2 m^2 / a^2) m_a m_b /
// m^2 - delta_ab)
template <typename T>
__global__ void __launch_bounds__(256) k_grid_virial_pair(int K0, int K1, int K2, const CX<T> *__restrict__ fi,
                                                          const double *__restrict__ recip, Frac2v A, double pt,
                                                          double scale, double *__restrict__ vpart) {
  long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long ntot = (long)K0 * K1 * K2;
  double w9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (g < ntot && recip[g] != 0.0) {
    int kx = (int)(g % K0);
    long r1 = g / K0;
    int ky = (int)(r1 % K1), kz = (int)(r1 / K1);
    int hh = kx >= (K0 + 1) / 2 ? kx - K0 : kx, k = ky >= (K1 + 1) / 2 ? ky - K1 : ky;
    int l = kz >= (K2 + 1) / 2 ? kz - K2 : kz;
    double m[3];
    for (int d = 0; d < 3; d++) m[d] = A.a[d][0] * hh + A.a[d][1] * k + A.a[d][2] * l;
    double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    long ng = (long)((K0 - kx) % K0) + (long)K0 * (((K1 - ky) % K1) + (long)K1 * ((K2 - kz) % K2));
    double ar = (double)fi[g].x, ai = (double)fi[g].y, br = (double)fi[ng].x, bi = -(double)fi[ng].y;
    double ur = 0.5 * (ar + br), ui = 0.5 * (ai + bi);
    double vr = 0.5 * (ai - bi), vi = -0.5 * (ar - br);
    double w = scale * recip[g] * (ur * vr + ui * vi), f = 2.0 * (1.0 + pt * m2) / m2;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) w9[3 * a + b] = w * (f * m[a] * m[b] - (a == b ? 1.0 : 0.0));
  }
  block_store9(w9, vpart);
}

template <typename T>
static void grid_virial_pair_t(NbHandle *h, const double *a, const double *b, double scale, double *vpart,
                               const CX<T> *tw) {
  CX<T> *g = (CX<T> *)h->d_grid;
  spread_into<T, true>(h, a, b);
  forward_t<T>(h, 2, g, tw);
  Frac2v A;
  for (int x = 0; x < 3; x++)
    for (int y = 0; y < 3; y++) A.a[x][y] = h->cell.A[x][y];
  double aw = h->prm.ewald_alpha, pt = (M_PI / aw) * (M_PI / aw);
  NB_LAUNCH((k_grid_virial_pair<T>), grid_for(h->ntot, 256), dim3(256), 0, h->K[0], h->K[1], h->K[2],
            (const CX<T> *)g, h->d_recip, A, pt, scale, vpart);
}

template <typename T>
static void grid_virial_t(NbHandle *h, bool polar, const CX<T> *tw) {
  CX<T> *g = (CX<T> *)h->d_grid, *fpg = (CX<T> *)h->d_fgrid;
  spread_into<T, false>(h, nullptr, nullptr);
  forward_t<T>(h, 1, fpg, tw);
  if (polar) {
    spread_into<T, true>(h, h->d_u, h->d_up);
    forward_t<T>(h, 2, g, tw);
  }
  Frac2v A;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) A.a[a][b] = h->cell.A[a][b];
  double aw = h->prm.ewald_alpha, pt = (M_PI / aw) * (M_PI / aw);
  dim3 G = grid_for(h->ntot, 256);
  NB_LAUNCH((k_grid_virial<T>), G, dim3(256), 0, h->K[0], h->K[1], h->K[2], (const CX<T> *)fpg,
            polar ? (const CX<T> *)g : (const CX<T> *)nullptr,
            (int)(h->prm.polarization == NB_POL_MUTUAL && !h->finite_active), h->d_recip,
            A, pt, h->prm.electric, h->ph.perm, h->ph.pol, nb_vir_region(h, G.x));
}

static bool mixed_grid(const NbHandle *h) { return h->prm.precision == NB_PREC_MIXED; }

// np = 0: the grid already holds the data; np = 1, 2: it is assembled from the bricks
static void convolution(NbHandle *h, int np) {
  if (mixed_grid(h))
    transform_t<float>(h, np, (CX<float> *)h->d_grid, (const CX<float> *)h->d_twf, h->d_recipf);
  else
    transform_t<double>(h, np, (CX<double> *)h->d_grid, (const CX<double> *)h->d_tw, h->d_recip);
}

template <typename T, bool IND>
static void spread_t(NbHandle *h, const double *u, const double *up) {
  spread_into<T, IND>(h, u, up);
}

template <typename T, int NP, bool FULL>
static void gather_t(NbHandle *h, double *o0, double *o1) {
  Frac fr = frac_of(h);
  NB_LAUNCH((k_gather<T, NP, FULL>), xcd_grid((long)h->n * 8, 256), dim3(256), 0, h->n, (const T *)h->d_theta, h->d_igrid,
            (const CX<T> *)h->d_grid, fr, h->K[0], h->K[1], h->K[2], o0, o1, h->ctl);
}

void nb_pme_permanent(NbHandle *h) {
  int n = h->n;
  dim3 B(128), G((n + 127) / 128);
  Frac fr = frac_of(h);
  // the brick counts were cleared by k_zero_outputs at the start of the evaluation
  // the FP64 weights of the mixed form's refinement come from the same launch
  const bool w64 = mixed_grid(h) && h->refine && h->prm.polarization == NB_POL_MUTUAL;
  if (mixed_grid(h))
    NB_LAUNCH(k_bspline_frac<float>, G, B, 0, n, h->d_pos4, h->d_gmp, h->cell, bricks_of(h), fr, (float *)h->d_theta,
              h->d_igrid, h->d_fmp, h->d_bcount, h->d_bid, h->d_brank, w64 ? h->d_theta64 : (double *)nullptr);
  else
    NB_LAUNCH(k_bspline_frac<double>, G, B, 0, n, h->d_pos4, h->d_gmp, h->cell, bricks_of(h), fr, h->d_theta,
              h->d_igrid, h->d_fmp, h->d_bcount, h->d_bid, h->d_brank, (double *)nullptr);
  NB_LAUNCH(k_brick_scan, dim3(1), dim3(1024), 0, h->nbricks, h->d_bcount, h->d_bstart);
  NB_LAUNCH(k_brick_fill, grid_for(n, 256), dim3(256), 0, n, h->d_bid, h->d_brank, h->d_bstart, h->d_batom);
  if (mixed_grid(h)) {
    spread_t<float, false>(h, nullptr, nullptr);
    convolution(h, 1);
    gather_t<float, 1, true>(h, h->d_phip, nullptr);
  } else {
    spread_t<double, false>(h, nullptr, nullptr);
    convolution(h, 1);
    gather_t<double, 1, true>(h, h->d_phip, nullptr);
  }
}

// FP64 reciprocal potentials (all 20 derivatives and the Cartesian field) of
// the dipoles u, up in the mixed form: FP64 spread with the FP64 weights,
// FP64 transforms and influence function, FP64 gather
void nb_pme_induced64(NbHandle *h, const double *u, const double *up) {
  CX<double> *g = (CX<double> *)h->d_grid;
  spread_into<double, true>(h, u, up, h->d_theta64);
  transform_t<double>(h, 2, g, (const CX<double> *)h->d_tw, h->d_recip);
  NB_LAUNCH((k_gather<double, 2, true>), xcd_grid((long)h->n * 8, 256), dim3(256), 0, h->n,
            (const double *)h->d_theta64, h->d_igrid, (const CX<double> *)g, frac_of(h), h->K[0], h->K[1], h->K[2],
            h->d_phiu, h->d_phiup, h->ctl);
}

void nb_pme_induced(NbHandle *h, const double *u, const double *up, bool full) {
  if (mixed_grid(h)) {
    spread_t<float, true>(h, u, up);
    convolution(h, 2);
    if (full) gather_t<float, 2, true>(h, h->d_phiu, h->d_phiup);
    else gather_t<float, 2, false>(h, h->d_phiu, h->d_phiup);
  } else {
    spread_t<double, true>(h, u, up);
    convolution(h, 2);
    if (full) gather_t<double, 2, true>(h, h->d_phiu, h->d_phiup);
    else gather_t<double, 2, false>(h, h->d_phiu, h->d_phiup);
  }
}

// PME of the PCG iteration's z (spread, convolution), then the join with the
// side stream's real-space field of z, then the gather fused with the
// conjugation (k_gather_ap); returns the number of partial sums of p.q.
int nb_pme_pcg_direction(NbHandle *h, double a3, int nrz, int first) {
  Frac fr = frac_of(h);
  dim3 G = xcd_grid((long)h->n * 8, 256);
  auto gap = [&](auto tag) {
    using T = decltype(tag);
    spread_t<T, true>(h, h->d_z, h->d_zp);
    convolution(h, 2);
    nb_join(h);
    NB_LAUNCH((k_gather_ap<T>), G, dim3(256), 0, h->n, (const T *)h->d_theta, h->d_igrid, (const CX<T> *)h->d_grid, fr,
              h->K[0], h->K[1], h->K[2], h->d_phiu, h->d_phiup, a3, h->d_polar, h->d_z, h->d_zp, h->d_field,
              h->d_fieldp, h->d_p, h->d_pp, h->d_vec, h->d_vecp, h->d_part, h->part_stride, nrz, h->d_scal, first,
              h->ctl);
  };
  if (mixed_grid(h)) gap(0.0f);
  else gap(0.0);
  return (int)G.x;
}

void nb_rotate(NbHandle *h) {
  int n = h->n;
  dim3 B(128), G((n + 127) / 128);
  NB_LAUNCH(k_rotate, G, B, 0, n, h->d_xyz, h->d_frame, h->d_axis, h->d_mpole_local, h->d_iperm,
                     h->d_redidx, h->d_redval, h->d_mp, h->d_gmp, h->d_red4, h->cell, (const int32_t *)h->d_wrap);
}

// standalone convolution entry used by the FFT parity test (precision:
// NB_PREC_FP64 or NB_PREC_MIXED selects the FP64 or FP32 transform)
extern "C" int nb_test_convolution(double *data, const double *recip, int nx, int ny, int nz, int precision) {
  NbHandle h;
  try {
    HIPCHECK(hipStreamCreate(&h.stream));
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHECK(hipGetDeviceProperties(&prop, dev));
    h.lds_max = std::min<size_t>(prop.sharedMemPerBlock, 65536);
    h.prm.precision = precision;
    long ntot = (long)nx * ny * nz;
    h.K[0] = nx; h.K[1] = ny; h.K[2] = nz; h.ntot = ntot;
    HIPCHECK(hipMalloc(&h.d_grid, sizeof(double) * 2 * ntot));
    HIPCHECK(hipMalloc(&h.d_recip, sizeof(double) * ntot));
    HIPCHECK(hipMalloc(&h.d_recipf, sizeof(float) * ntot));
    HIPCHECK(hipMalloc(&h.d_tw, sizeof(double) * 2 * (nx + ny + nz)));
    HIPCHECK(hipMalloc(&h.d_twf, sizeof(float) * 2 * (nx + ny + nz)));
    int K[3] = {nx, ny, nz};
    std::vector<double> tw;
    twiddles(K, tw);
    std::vector<float> twf(tw.begin(), tw.end());
    HIPCHECK(hipMemcpy(h.d_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(h.d_twf, twf.data(), sizeof(float) * twf.size(), hipMemcpyHostToDevice));
    if (precision == NB_PREC_MIXED) {
      std::vector<float> gf(data, data + 2 * ntot), rf(recip, recip + ntot);
      HIPCHECK(hipMemcpy(h.d_grid, gf.data(), sizeof(float) * 2 * ntot, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(h.d_recipf, rf.data(), sizeof(float) * ntot, hipMemcpyHostToDevice));
      convolution(&h, 0);
      HIPCHECK(hipStreamSynchronize(h.stream));
      HIPCHECK(hipMemcpy(gf.data(), h.d_grid, sizeof(float) * 2 * ntot, hipMemcpyDeviceToHost));
      for (long i = 0; i < 2 * ntot; i++) data[i] = gf[i];
    } else {
      HIPCHECK(hipMemcpy(h.d_grid, data, sizeof(double) * 2 * ntot, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(h.d_recip, recip, sizeof(double) * ntot, hipMemcpyHostToDevice));
      convolution(&h, 0);
      HIPCHECK(hipStreamSynchronize(h.stream));
      HIPCHECK(hipMemcpy(data, h.d_grid, sizeof(double) * 2 * ntot, hipMemcpyDeviceToHost));
    }
    free_grid(&h);
    hipStreamDestroy(h.stream);
    return 0;
  } catch (std::exception &e) {
    free_grid(&h);
    if (h.stream) hipStreamDestroy(h.stream);
    return NB_ERR_HIP;
  }
}

// reciprocal grid part of the strain derivative (main stream, after the
// induced potentials have been gathered)
void nb_grid_virial(NbHandle *h, bool polar) {
  if (!h->d_fgrid) HIPCHECK(hipMalloc(&h->d_fgrid, sizeof(double) * 2 * h->ntot));
  if (mixed_grid(h)) grid_virial_t<float>(h, polar, (const CX<float> *)h->d_twf);
  else grid_virial_t<double>(h, polar, (const CX<double> *)h->d_tw);
}

long nb_grid_virial_blocks(const NbHandle *h) { return grid_for(h->ntot, 256).x; }

void nb_grid_virial_pair(NbHandle *h, const double *a, const double *b, double scale, double *vpart) {
  if (mixed_grid(h)) grid_virial_pair_t<float>(h, a, b, scale, vpart, (const CX<float> *)h->d_twf);
  else grid_virial_pair_t<double>(h, a, b, scale, vpart, (const CX<double> *)h->d_tw);
}

Radices nb_factor(int N) { return factor(N); }
