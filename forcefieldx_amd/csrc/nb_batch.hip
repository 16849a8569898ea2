// nb_batch.hip -- many small periodic systems evaluated in one persistent
// launch (gfx950).
//
// A workgroup takes one system at a time from a queue (largest first) and
// carries it through the whole evaluation with workgroup barriers between the
// phases, so a batch of hundreds of cells costs one launch instead of a few
// hundred small launches per cell:
//
//   atoms   local -> global multipoles (MultipoleType.java:385-540, 1031-1058),
//           reduced vdW sites (VanDerWaals.java:1510-1540), positions wrapped
//           into the cell, order-5 B-splines and fractional multipoles
//           (ReciprocalSpace.java:1344-1397, 2778-2834);
//   rows    every partner at every lattice translation within the cutoff
//           (explicit image sums, as ReplicatesCrystal provides for thin
//           cells; masks only for the pair as given), split into the
//           preconditioner part (r < 4.5 A) and the rest
//           (PermanentFieldRegion.java:679-714);
//   field   permanent real-space field (PermanentFieldRegion.java:704-808)
//           and the permanent PME potential (spread with atomics into the
//           system's grid, in-LDS line FFTs with the influence multiply fused,
//           gather; ReciprocalSpace.java:1445-1505, Complex3DParallel.java:
//           445-453, ReciprocalSpace.java:2264-2455);
//   solve   direct dipoles (DirectRegion.java:170-245) and the PCG
//           (PCGSolver.java:487-1178) with workgroup reductions, the
//           convergence test of the reference (:543-547) taken in the kernel;
//   energy  real-space permanent and polarization energy / force / torque
//           (RealSpaceEnergyRegion.java:803-2177), buffered 14-7 vdW
//           (VanDerWaals.java:1638-1900), reciprocal energy / gradient /
//           torque (ReciprocalEnergyRegion.java:381-746), torques to
//           frame-atom gradients and the reduced-site chain rule
//           (Torque.java:111-316, VanDerWaals.java:1826-1833).
//
// The per-pair and per-atom arithmetic is shared with the per-system kernels
// (nb_pair.h, nb_atom.h, nb_fft.h); T is double (FP64 form) or float (mixed
// form: FP32 pair, spread and FFT arithmetic, FP64 accumulation, masked pairs
// of the permanent terms in FP64).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <stdexcept>
#include <vector>

#include "nb_kern.h"
#include "nb_pair.h"
#include "nb_fft.h"

namespace {

struct BatchSys {
  int n, K0, K1, K2;
  int maxa, maxb;           // row capacities: preconditioner part, the rest
  int Lx, Ly, Lz;           // FFT lines per chunk along x, y, z
  double lcut;              // list cutoff the image range was checked for
  Radices rad[3];
  CellD cell;               // image sums (multi = 1)
  Frac fr;
  const double *xyz, *mloc, *polar, *thole, *ipd, *redval, *vtab;
  const int32_t *frame, *axis, *group, *vclass, *redidx, *moff, *matom, *mrel;
  int nclass;
  const void *recip, *tw;   // influence function and twiddles in the grid type
  double *grad, *u, *up;    // outputs (caller order)
  char *ws;                 // workspace
};

struct BatchPrm {
  ScaleTab sc;
  double aewald, ewald_cut, vdw_off, vdw_cut, pcut, psc, polar_eps, electric, hdelta, hgamma;
  int polarization, maxit, do_vdw, do_pme, recip;
};

// per-system workspace
template <typename T> struct Ws {
  double4 *pos, *red;
  AtomMP *mp;
  double *gmp, *theta, *fmp, *phip, *phiu, *phiup;
  double *f, *fp, *ud, *upd, *r, *rp, *z, *zp, *p, *pp, *vec, *vecp, *gacc, *tacc, *gred;
  int32_t *igrid, *wrap, *na, *nb;
  uint32_t *ra, *rb;
  DipRec<T> *rec;
  CX<T> *grid;
};

template <typename T>
__host__ __device__ size_t ws_map(char *base, int n, int maxa, int maxb, long ntot, Ws<T> *w) {
  size_t off = 0;
  auto take = [&](size_t bytes) -> char * {
    char *q = base + off;
    off += (bytes + 255) & ~(size_t)255;
    return q;
  };
  size_t d3 = sizeof(double) * 3 * (size_t)n;
  Ws<T> v;
  v.pos = (double4 *)take(sizeof(double4) * n);
  v.red = (double4 *)take(sizeof(double4) * n);
  v.mp = (AtomMP *)take(sizeof(AtomMP) * n);
  v.gmp = (double *)take(sizeof(double) * 10 * n);
  v.theta = (double *)take(sizeof(double) * 60 * n);
  v.fmp = (double *)take(sizeof(double) * 10 * n);
  v.phip = (double *)take(sizeof(double) * 30 * n);
  v.phiu = (double *)take(sizeof(double) * 30 * n);
  v.phiup = (double *)take(sizeof(double) * 30 * n);
  v.f = (double *)take(d3);
  v.fp = (double *)take(d3);
  v.ud = (double *)take(d3);
  v.upd = (double *)take(d3);
  v.r = (double *)take(d3);
  v.rp = (double *)take(d3);
  v.z = (double *)take(d3);
  v.zp = (double *)take(d3);
  v.p = (double *)take(d3);
  v.pp = (double *)take(d3);
  v.vec = (double *)take(d3);
  v.vecp = (double *)take(d3);
  v.gacc = (double *)take(d3);
  v.tacc = (double *)take(d3);
  v.gred = (double *)take(d3);
  v.igrid = (int32_t *)take(sizeof(int32_t) * 3 * n);
  v.wrap = (int32_t *)take(sizeof(int32_t) * 3 * n);
  v.na = (int32_t *)take(sizeof(int32_t) * n);
  v.nb = (int32_t *)take(sizeof(int32_t) * n);
  v.ra = (uint32_t *)take(sizeof(uint32_t) * (size_t)n * maxa);
  v.rb = (uint32_t *)take(sizeof(uint32_t) * (size_t)n * maxb);
  v.rec = (DipRec<T> *)take(sizeof(DipRec<T>) * n);
  v.grid = (CX<T> *)take(sizeof(CX<T>) * (size_t)ntot);
  if (w) *w = v;
  return off;
}

// per-system outputs: energy slots E_PERM_REAL .. E_QTOT, then these
enum { O_ITER = 8, O_EPS, O_STAT, O_NVDW, O_NPERM, O_FLAGS, O_ROWA, O_ROWB, O_N };
// O_FLAGS bits
enum { F_ROWS = 1, F_TORQUE = 2, F_CUTOFF = 4 };

// LDS line-buffer elements per workgroup: two 28 KB buffers (ping-pong), so
// that the workgroup's LDS stays below 64 KB.  (A 135 KB layout, within the
// 160 KB the device reports per workgroup, gave wrong results: the variables
// placed after the buffers read back altered.)
template <typename T> struct FftCap { static constexpr int value = 28672 / sizeof(CX<T>); };

// One out-of-place Stockham pass of radix R over nlines lines of length N
// (line l at l * sl, element e at e * se), src -> dst.  Every butterfly is
// read, combined and written by one thread with nothing held across the
// pass, so a chunk of any size costs no registers beyond one butterfly.
template <int R, bool LINE_FAST, typename T>
__device__ __forceinline__ void pp_pass(const CX<T> *src, CX<T> *dst, int N, int nlines, int sl, int se, int Ns,
                                        const CX<T> *tw, T s) {
  const int nb = N / R, total = nlines * nb, stride = N / (Ns * R);
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    int line, j;
    if (LINE_FAST) { line = b % nlines; j = b / nlines; }
    else { j = b % nb; line = b / nb; }
    int jm = j % Ns;
    const CX<T> *x = src + line * sl;
    CX<T> v[R];
#pragma unroll
    for (int k = 0; k < R; k++) {
      CX<T> a = x[(j + k * nb) * se];
      if (k > 0 && jm > 0) {
        CX<T> w = tw[jm * k * stride];
        w.y *= s;
        a = cmul(a, w);
      }
      v[k] = a;
    }
    small_dft<R, T>(v, s);
    CX<T> *y = dst + line * sl + ((j / Ns) * Ns * R + jm) * se;
#pragma unroll
    for (int k = 0; k < R; k++) y[k * Ns * se] = v[k];
  }
  __syncthreads();
}

// all passes of one transform; returns the buffer that holds the result
template <bool LINE_FAST, typename T>
__device__ CX<T> *pp_fft(CX<T> *a, CX<T> *b, int N, int nlines, int sl, int se, const Radices &rad,
                         const CX<T> *tw, T s) {
  int Ns = 1;
  for (int p = 0; p < rad.nr; p++) {
    int R = rad.r[p];
    if (R == 4) pp_pass<4, LINE_FAST, T>(a, b, N, nlines, sl, se, Ns, tw, s);
    else if (R == 2) pp_pass<2, LINE_FAST, T>(a, b, N, nlines, sl, se, Ns, tw, s);
    else if (R == 3) pp_pass<3, LINE_FAST, T>(a, b, N, nlines, sl, se, Ns, tw, s);
    else if (R == 5) pp_pass<5, LINE_FAST, T>(a, b, N, nlines, sl, se, Ns, tw, s);
    else pp_pass<7, LINE_FAST, T>(a, b, N, nlines, sl, se, Ns, tw, s);
    Ns *= R;
    CX<T> *t = a;
    a = b;
    b = t;
  }
  return a;
}

// One chunk of L lines of the system's grid through LDS (AXIS 0: rows; AXIS
// 3: lines along y over consecutive (x, z) indices; AXIS 2: lines along z over
// consecutive (x, y) indices, with the influence multiply and the inverse
// transform fused when CONV).
template <typename T, int NT, int AXIS, bool CONV>
__device__ void pp_chunk(CX<T> *__restrict__ grid, int K0, int K1, int K2, int L, int chunk, const Radices &rad,
                         CX<T> *b0, CX<T> *b1, const CX<T> *tw, T s, const T *__restrict__ recip) {
  const int N = AXIS == 0 ? K0 : (AXIS == 3 ? K1 : K2);
  const int nlt = AXIS == 0 ? K1 * K2 : (AXIS == 3 ? K0 * K2 : K0 * K1);
  const int first = chunk * L, cnt = min(L, nlt - first);
  auto gidx = [&](int l, int e) -> size_t {
    int q = first + l;
    if (AXIS == 0) return (size_t)q * K0 + e;
    if (AXIS == 3) return (size_t)(q % K0) + (size_t)K0 * (e + (size_t)K1 * (q / K0));
    return (size_t)q + (size_t)K0 * K1 * e;
  };
  // every load of the chunk is issued before the first is consumed (one
  // memory latency per chunk, not one per element)
  {
    constexpr int MAXE = (FftCap<T>::value + NT - 1) / NT;
    CX<T> v[MAXE];
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      int t = q * NT + threadIdx.x;
      if (t < cnt * N) {
        int l, e;
        if (AXIS == 0) { l = t / N; e = t - l * N; }
        else { e = t / cnt; l = t - e * cnt; }
        v[q] = grid[gidx(l, e)];
      }
    }
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      int t = q * NT + threadIdx.x;
      if (t < cnt * N) b0[t] = v[q];
    }
  }
  __syncthreads();
  CX<T> *r;
  if (AXIS == 0) r = pp_fft<false, T>(b0, b1, N, cnt, N, 1, rad, tw, s);
  else r = pp_fft<true, T>(b0, b1, N, cnt, 1, cnt, rad, tw, s);
  if (CONV) {
    for (int t = threadIdx.x; t < cnt * N; t += blockDim.x) {
      int e = t / cnt, l = t - e * cnt;
      T g = recip[(size_t)(first + l) + (size_t)K0 * K1 * e];
      r[t].x *= g;
      r[t].y *= g;
    }
    __syncthreads();
    CX<T> *o = r == b0 ? b1 : b0;
    r = pp_fft<true, T>(r, o, N, cnt, 1, cnt, rad, tw, -s);
  }
  for (int t = threadIdx.x; t < cnt * N; t += blockDim.x) {
    int l, e;
    if (AXIS == 0) { l = t / N; e = t - l * N; }
    else { e = t / cnt; l = t - e * cnt; }
    grid[gidx(l, e)] = r[AXIS == 0 ? l * N + e : e * cnt + l];
  }
  __syncthreads();
}
constexpr int TW_CAP = 3 * 64;  // LDS twiddle entries (three axes of at most 64)

// sums of NV values over the workgroup, returned to every thread (fixed order)
template <int NT, int NV>
__device__ __forceinline__ void block_sums(double (&v)[NV], double *red) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum_d(v[k]);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) red[k * NW + w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
    for (int q = 0; q < NW; q++) s += red[k * NW + q];
    v[k] = s;
  }
}

// maxima of NV values over the workgroup, returned to every thread
template <int NT, int NV>
__device__ __forceinline__ void block_max(double (&v)[NV], double *red) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int k = 0; k < NV; k++)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] = fmax(v[k], __shfl_xor(v[k], o, 64));
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) red[k * NW + w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double m = red[k * NW];
    for (int q = 1; q < NW; q++) m = fmax(m, red[k * NW + q]);
    v[k] = m;
  }
}

// Phases read what other threads of the workgroup wrote to global memory
// with plain stores: the workgroup barrier orders and publishes them.
__device__ __forceinline__ void phase_sync() { __syncthreads(); }

// The grid is spread with atomics, which are performed in L2 and leave any
// copy of the line in the CU's vector L1 as it was: the device-scope fence
// waits for the cleared grid to reach L2 before the atomics, and for the
// atomics, with L1 invalidated, before the transform reads the grid.
__device__ __forceinline__ void atomic_sync() {
  __threadfence();
  __syncthreads();
}

// exclusive prefix over the lower lanes of an LPA group
template <int LPA>
__device__ __forceinline__ int group_excl_scan(int v) {
  int lane = threadIdx.x % LPA, x = v;
#pragma unroll
  for (int o = 1; o < LPA; o <<= 1) {
    int y = __shfl_up(x, o, LPA);
    if (lane >= o) x += y;
  }
  return x - v;
}

// row entry t of atom i: the preconditioner part first, then the rest
template <typename T>
__device__ __forceinline__ uint32_t row_entry(const BatchSys &S, const Ws<T> &w, int i, int na, int t) {
  return t < na ? w.ra[(size_t)i * S.maxa + t] : w.rb[(size_t)i * S.maxb + (t - na)];
}

// ---------------------------------------------------------------------------
// atoms: frames, multipoles, reduced sites, wrapped positions, B-splines.
// Returns the largest reduced-site displacement among this thread's atoms.
// ---------------------------------------------------------------------------
template <typename T, int NT>
__device__ double phase_atoms(const BatchSys &S, const BatchPrm &P, const Ws<T> &w) {
  const CellD &c = S.cell;
  double dmax = 0.0;
  for (int i = threadIdx.x; i < S.n; i += NT) {
    V3<double> o = ld3(S.xyz, i);
    double f0 = o.x * c.A00 + o.y * c.A10 + o.z * c.A20, f1 = o.y * c.A11 + o.z * c.A21, f2 = o.z * c.A22;
    int wx = (int)floor(f0), wy = (int)floor(f1), wz = (int)floor(f2);
    w.wrap[3 * i] = wx;
    w.wrap[3 * i + 1] = wy;
    w.wrap[3 * i + 2] = wz;
    double sx = wx * c.Ai00 + wy * c.Ai10 + wz * c.Ai20, sy = wy * c.Ai11 + wz * c.Ai21, sz = wz * c.Ai22;
    double4 p = make_double4(o.x - sx, o.y - sy, o.z - sz, 0.0);
    w.pos[i] = p;
    double g[10];
    rotate_multipole(i, S.xyz, S.frame, S.axis, S.mloc, g);
    for (int q = 0; q < 10; q++) w.gmp[10 * i + q] = g[q];
    const double t = 1.0 / 3.0;
    AtomMP a;
    a.c = g[0]; a.dx = g[1]; a.dy = g[2]; a.dz = g[3];
    a.qxx = g[4] * t; a.qyy = g[5] * t; a.qzz = g[6] * t;
    a.qxy = g[7] * t; a.qxz = g[8] * t; a.qyz = g[9] * t;
    a.ipd = S.ipd[i];
    a.thole = S.thole[i];
    w.mp[i] = a;
    int h = S.redidx[i];
    double rv = S.redval[i];
    double4 rs = make_double4(o.x, o.y, o.z, 0.0);
    if (h >= 0 && h != i && rv > 0.0) {
      V3<double> ph = ld3(S.xyz, h);
      rs.x = rv * (o.x - ph.x) + ph.x;
      rs.y = rv * (o.y - ph.y) + ph.y;
      rs.z = rv * (o.z - ph.z) + ph.z;
    }
    double ex = rs.x - o.x, ey = rs.y - o.y, ez = rs.z - o.z;
    dmax = fmax(dmax, sqrt(ex * ex + ey * ey + ez * ez));
    w.red[i] = make_double4(rs.x - sx, rs.y - sy, rs.z - sz, 0.0);
    dip_put_static((void *)w.rec, sizeof(T) == 4, i, p.x, p.y, p.z, a.ipd, a.thole);
    if (P.recip) {
      int Kd[3] = {S.K0, S.K1, S.K2};
      for (int d = 0; d < 3; d++) {
        double wf = p.x * c.A[0][d] + p.y * c.A[1][d] + p.z * c.A[2][d];
        double u = wf - floor(wf + 0.5) + 0.5;
        double frx = Kd[d] * u;
        int ifr = (int)frx;
        double bx = frx - ifr;
        w.igrid[3 * i + d] = ifr - 5;
        double th[5][4];
        bspline4<5>(bx, th);
        double *ot = w.theta + ((size_t)i * 3 + d) * 20;
        for (int k = 0; k < 5; k++)
          for (int j = 0; j < 4; j++) ot[4 * k + j] = th[k][j];
      }
      frac_multipole(S.fr, g, w.fmp + 10 * i);
    }
    for (int d = 0; d < 3; d++) {
      w.gacc[3 * i + d] = 0.0;
      w.tacc[3 * i + d] = 0.0;
      w.gred[3 * i + d] = 0.0;
      S.grad[3 * i + d] = 0.0;
    }
  }
  return dmax;
}

// ---------------------------------------------------------------------------
// rows.  Lane l of an atom's group scans the partner block [l n / LPA,
// (l + 1) n / LPA) and, per partner, the lattice translations whose
// fractional offset can reach the cutoff.  A counting pass, a group scan for
// each lane's offset and a writing pass list the partners of every row in
// ascending order, translations in (z, y, x) order.  Entry: partner | image
// code << 18 | mask code << 27, the format of the per-system image rows.
// ---------------------------------------------------------------------------
template <typename T, bool WRITE>
__device__ __forceinline__ void scan_partners(const BatchSys &S, const Ws<T> &w, int i, int t0, int t1,
                                              double lcut2, double ecut2, double pcut, const double cr[3], int &ca,
                                              int &cb, uint32_t *rowa, uint32_t *rowb, int capa, int capb) {
  const CellD &c = S.cell;
  double4 pi = w.pos[i];
  double fi0 = pi.x * c.A00 + pi.y * c.A10 + pi.z * c.A20, fi1 = pi.y * c.A11 + pi.z * c.A21, fi2 = pi.z * c.A22;
  int gi = S.group[i];
  int wi0 = w.wrap[3 * i], wi1 = w.wrap[3 * i + 1], wi2 = w.wrap[3 * i + 2];
  for (int t = t0; t < t1; t++) {
    double4 pk = w.pos[t];
    double f0 = pk.x * c.A00 + pk.y * c.A10 + pk.z * c.A20 - fi0, f1 = pk.y * c.A11 + pk.z * c.A21 - fi1;
    double f2 = pk.z * c.A22 - fi2;
    int lo0 = max((int)ceil(-cr[0] - f0), -NB_IMG_R), hi0 = min((int)floor(cr[0] - f0), NB_IMG_R);
    int lo1 = max((int)ceil(-cr[1] - f1), -NB_IMG_R), hi1 = min((int)floor(cr[1] - f1), NB_IMG_R);
    int lo2 = max((int)ceil(-cr[2] - f2), -NB_IMG_R), hi2 = min((int)floor(cr[2] - f2), NB_IMG_R);
    int o0 = 0, o1 = 0, o2 = 0, rel = -1;
    if (WRITE) {
      o0 = w.wrap[3 * t] - wi0;
      o1 = w.wrap[3 * t + 1] - wi1;
      o2 = w.wrap[3 * t + 2] - wi2;
    }
    for (int tz = lo2; tz <= hi2; tz++)
      for (int ty = lo1; ty <= hi1; ty++)
        for (int tx = lo0; tx <= hi0; tx++) {
          if (t == i && tx == 0 && ty == 0 && tz == 0) continue;
          double x = pk.x + tx * c.Ai00 + ty * c.Ai10 + tz * c.Ai20, y = pk.y + ty * c.Ai11 + tz * c.Ai21;
          double z = pk.z + tz * c.Ai22;
          double dx = x - pi.x, dy = y - pi.y, dz = z - pi.z;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 > lcut2) continue;
          bool pre = r2 <= ecut2 && sqrt(r2) < pcut;
          if (WRITE) {
            int code = 0;
            if (tx == o0 && ty == o1 && tz == o2) {
              if (rel < 0) rel = relation(S.moff, S.matom, S.mrel, i, t);
              code = rel + (S.group[t] == gi ? 8 : 0);
            }
            uint32_t img = (uint32_t)((tx + NB_IMG_R) + NB_IMG_W * ((ty + NB_IMG_R) + NB_IMG_W * (tz + NB_IMG_R)));
            uint32_t e = (uint32_t)t | (img << NB_IMG_SHIFT) | ((uint32_t)code << NB_CODE_SHIFT);
            if (pre) {
              if (ca < capa) rowa[ca] = e;
            } else {
              if (cb < capb) rowb[cb] = e;
            }
          }
          if (pre) ca++;
          else cb++;
        }
  }
}

// returns (per thread) the largest true row lengths seen, which exceed the
// capacities when a row overflowed (its entries past the capacity are dropped)
template <typename T, int NT, int LPA>
__device__ void phase_rows(const BatchSys &S, const BatchPrm &P, const Ws<T> &w, double lcut, int &mxa, int &mxb) {
  const CellD &c = S.cell;
  double cr[3];
  for (int d = 0; d < 3; d++)
    cr[d] = lcut * sqrt(c.A[0][d] * c.A[0][d] + c.A[1][d] * c.A[1][d] + c.A[2][d] * c.A[2][d]);
  double lcut2 = lcut * lcut, ecut2 = P.ewald_cut * P.ewald_cut;
  const int groups = NT / LPA, g = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  mxa = 0;
  mxb = 0;
  for (int base = 0; base < S.n; base += groups) {
    int i = base + g;
    bool valid = i < S.n;
    int t0 = (int)(((long)lane * S.n) / LPA), t1 = (int)(((long)(lane + 1) * S.n) / LPA);
    int ca = 0, cb = 0;
    if (valid) scan_partners<T, false>(S, w, i, t0, t1, lcut2, ecut2, P.pcut, cr, ca, cb, nullptr, nullptr, 0, 0);
    int oa = group_excl_scan<LPA>(ca), ob = group_excl_scan<LPA>(cb);
    int ta = __shfl(oa + ca, LPA - 1, LPA), tb = __shfl(ob + cb, LPA - 1, LPA);
    if (valid) {
      int wa = 0, wb = 0;
      uint32_t *rowa = w.ra + (size_t)i * S.maxa, *rowb = w.rb + (size_t)i * S.maxb;
      // each lane writes from its offset, bounded by what is left of the row
      if (oa < S.maxa && ob < S.maxb)
        scan_partners<T, true>(S, w, i, t0, t1, lcut2, ecut2, P.pcut, cr, wa, wb, rowa + oa, rowb + ob,
                               S.maxa - oa, S.maxb - ob);
      if (lane == 0) {
        w.na[i] = min(ta, S.maxa);
        w.nb[i] = min(tb, S.maxb);
        mxa = max(mxa, ta);
        mxb = max(mxb, tb);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// real-space pair loops: LPA lanes per atom walk the row (preconditioner part,
// then the rest)
// ---------------------------------------------------------------------------
template <typename T, int NT, int LPA>
__device__ void phase_perm_field(const BatchSys &S, const BatchPrm &P, const Ws<T> &w) {
  const int groups = NT / LPA, g = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  const double ecut2 = P.ewald_cut * P.ewald_cut, aw = P.recip ? P.aewald : 0.0;
  for (int base = 0; base < S.n; base += groups) {
    int i = base + g;
    bool valid = i < S.n;
    V3<double> fd{0, 0, 0}, fp{0, 0, 0};
    if (valid) {
      double4 pi = w.pos[i];
      AtomMP ai = w.mp[i];
      int na = w.na[i], m = na + w.nb[i];
      for (int t = lane; t < m; t += LPA) {
        uint32_t e = row_entry(S, w, i, na, t);
        int k = e & S.cell.kmask, code = e >> NB_CODE_SHIFT;
        double4 pk = w.pos[k];
        entry_shift(S.cell, e, pk.x, pk.y, pk.z);
        double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
        if (dx * dx + dy * dy + dz * dz > ecut2) continue;
        if (sizeof(T) == 4 && P.sc.masked[code])
          perm_field_pair<double>(pi, ai, pk, w.mp[k], code, S.cell, aw, P.sc, fd, fp);
        else
          perm_field_pair<T>(pi, ai, pk, w.mp[k], code, S.cell, aw, P.sc, fd, fp);
      }
    }
    group_sum3<LPA>(fd);
    group_sum3<LPA>(fp);
    if (valid && lane == 0) {
      st3(w.f, i, fd);
      st3(w.fp, i, fp);
    }
  }
}

// field of the vectors held in the packed records: Ewald-screened over the
// Ewald part of the rows, or (PRE) bare Coulomb + Thole over the
// preconditioner part
template <typename T, int NT, int LPA, bool PRE>
__device__ void phase_dip_field(const BatchSys &S, const BatchPrm &P, const Ws<T> &w) {
  const int groups = NT / LPA, g = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  const double ecut2 = P.ewald_cut * P.ewald_cut, aw = (P.recip && !PRE) ? P.aewald : 0.0;
  for (int base = 0; base < S.n; base += groups) {
    int i = base + g;
    bool valid = i < S.n;
    V3<double> a{0, 0, 0}, b{0, 0, 0};
    if (valid) {
      const DipRec<T> ri = w.rec[i];
      int na = w.na[i], m = PRE ? na : na + w.nb[i];
      for (int t = lane; t < m; t += LPA) {
        uint32_t e = row_entry(S, w, i, na, t);
        const DipRec<T> rk = w.rec[e & S.cell.kmask];
        if (!PRE && t >= na) {
          double dx = rk.x - ri.x, dy = rk.y - ri.y, dz = rk.z - ri.z;
          entry_shift(S.cell, e, dx, dy, dz);
          if (dx * dx + dy * dy + dz * dz > ecut2) continue;
        }
        dip_field_pair<T, PRE>(ri, rk, e, S.cell, aw, a, b);
      }
    }
    group_sum3<LPA>(a);
    group_sum3<LPA>(b);
    if (valid && lane == 0) {
      st3(w.f, i, a);
      st3(w.fp, i, b);
    }
  }
}

// ---------------------------------------------------------------------------
// reciprocal space
// ---------------------------------------------------------------------------
template <typename T, int NT>
__device__ void grid_zero(const BatchSys &S, const Ws<T> &w) {
  long ntot = (long)S.K0 * S.K1 * S.K2;
  for (long q = threadIdx.x; q < ntot; q += NT) w.grid[q] = CX<T>{(T)0, (T)0};
}

// 125-point stencils added to the grid with atomics: the permanent fractional
// multipoles into the real part (IND false), or the dipoles s0 into the real
// and s1 into the imaginary part (ReciprocalSpace.java:1445-1643)
template <typename T, int NT, bool IND>
__device__ void grid_spread(const BatchSys &S, const Ws<T> &w, const double *__restrict__ s0,
                            const double *__restrict__ s1) {
  const int K0 = S.K0, K1 = S.K1, K2 = S.K2;
  const Frac &fr = S.fr;
  for (int q = threadIdx.x; q < S.n * 25; q += NT) {
    int a = q / 25, yz = q - a * 25;
    int ky = yz % 5, kz = yz / 5;
    const double *tx = w.theta + (size_t)a * 60, *uu = tx + 20 + 4 * ky, *vv = tx + 40 + 4 * kz;
    int gz = wrapg(w.igrid[3 * a + 2] + 1 + kz, K2), gy = wrapg(w.igrid[3 * a + 1] + 1 + ky, K1);
    int gx0 = w.igrid[3 * a] + 1;
    CX<T> *row = w.grid + (size_t)K0 * (gy + (size_t)K1 * gz);
    if (!IND) {
      const double *m = w.fmp + 10 * (size_t)a;
      T u0 = (T)uu[0], u1 = (T)uu[1], u2 = (T)uu[2], v0 = (T)vv[0], v1 = (T)vv[1], v2 = (T)vv[2];
      T t0 = ((T)m[0] * v0 + (T)m[3] * v1 + (T)m[6] * v2) * u0 + ((T)m[2] * v0 + (T)m[9] * v1) * u1 + (T)m[5] * v0 * u2;
      T t1 = ((T)m[1] * v0 + (T)m[8] * v1) * u0 + (T)m[7] * v0 * u1;
      T t2 = (T)m[4] * v0 * u0;
#pragma unroll
      for (int kx = 0; kx < 5; kx++) {
        T val = (T)tx[4 * kx] * t0 + (T)tx[4 * kx + 1] * t1 + (T)tx[4 * kx + 2] * t2;
        unsafeAtomicAdd(&row[wrapg(gx0 + kx, K0)].x, val);
      }
    } else {
      T u0 = (T)uu[0], u1 = (T)uu[1], v0 = (T)vv[0], v1 = (T)vv[1];
      T t0[2], t1[2];
#pragma unroll
      for (int p = 0; p < 2; p++) {
        const double *src = (p ? s1 : s0) + 3 * (size_t)a;
        T d0 = (T)(fr.a[0][0] * src[0] + fr.a[0][1] * src[1] + fr.a[0][2] * src[2]);
        T d1 = (T)(fr.a[1][0] * src[0] + fr.a[1][1] * src[1] + fr.a[1][2] * src[2]);
        T d2 = (T)(fr.a[2][0] * src[0] + fr.a[2][1] * src[1] + fr.a[2][2] * src[2]);
        t0[p] = d2 * v1 * u0 + d1 * v0 * u1;
        t1[p] = d0 * v0 * u0;
      }
#pragma unroll
      for (int kx = 0; kx < 5; kx++) {
        T wx0 = (T)tx[4 * kx], wx1 = (T)tx[4 * kx + 1];
        CX<T> *pt = &row[wrapg(gx0 + kx, K0)];
        unsafeAtomicAdd(&pt->x, wx0 * t0[0] + wx1 * t1[0]);
        unsafeAtomicAdd(&pt->y, wx0 * t0[1] + wx1 * t1[1]);
      }
    }
  }
}

// forward x, forward y, fused forward z / influence / inverse z, inverse y,
// inverse x (un-normalised inverse, Complex.java:436-470)
template <typename T, int NT>
__device__ void grid_convolve(const BatchSys &S, const Ws<T> &w, CX<T> *b0, CX<T> *b1, const CX<T> *tw) {
  const int K0 = S.K0, K1 = S.K1, K2 = S.K2;
  const CX<T> *twx = tw, *twy = tw + K0, *twz = tw + K0 + K1;
  const T *recip = (const T *)S.recip;
  const T *nul = nullptr;
  const int nx = (K1 * K2 + S.Lx - 1) / S.Lx, ny = (K0 * K2 + S.Ly - 1) / S.Ly, nz = (K0 * K1 + S.Lz - 1) / S.Lz;
  for (int c = 0; c < nx; c++) pp_chunk<T, NT, 0, false>(w.grid, K0, K1, K2, S.Lx, c, S.rad[0], b0, b1, twx, (T)1, nul);
  for (int c = 0; c < ny; c++) pp_chunk<T, NT, 3, false>(w.grid, K0, K1, K2, S.Ly, c, S.rad[1], b0, b1, twy, (T)1, nul);
  for (int c = 0; c < nz; c++) pp_chunk<T, NT, 2, true>(w.grid, K0, K1, K2, S.Lz, c, S.rad[2], b0, b1, twz, (T)1, recip);
  for (int c = 0; c < ny; c++) pp_chunk<T, NT, 3, false>(w.grid, K0, K1, K2, S.Ly, c, S.rad[1], b0, b1, twy, (T)-1, nul);
  for (int c = 0; c < nx; c++) pp_chunk<T, NT, 0, false>(w.grid, K0, K1, K2, S.Lx, c, S.rad[0], b0, b1, twx, (T)-1, nul);
}

template <typename T, int NT, int NP, bool FULL>
__device__ void grid_gather(const BatchSys &S, const Ws<T> &w, double *o0, double *o1) {
  const int groups = NT / 8, g = threadIdx.x / 8, lane = threadIdx.x % 8;
  for (int base = 0; base < S.n; base += groups) {
    int i = base + g;
    // the FP32 form prefetches the stencil plane (in FP64 the extra registers spill)
    gather_atom<T, NP, FULL, sizeof(T) == 4>(i, lane, i < S.n, w.theta, w.igrid, w.grid, S.fr, S.K0, S.K1, S.K2, o0,
                                             o1);
  }
}

// reciprocal potential of the dipoles s0 (u) and s1 (uCR): FULL gives the 20
// derivatives each, otherwise the field only
template <typename T, int NT, bool FULL>
__device__ void recip_induced(const BatchSys &S, const Ws<T> &w, const double *s0, const double *s1, CX<T> *buf,
                              const CX<T> *tw) {
  grid_zero<T, NT>(S, w);
  atomic_sync();
  grid_spread<T, NT, true>(S, w, s0, s1);
  atomic_sync();
  grid_convolve<T, NT>(S, w, buf, buf + FftCap<T>::value, tw);
  grid_gather<T, NT, 2, FULL>(S, w, w.phiu, w.phiup);
  phase_sync();
}

// total induced field of (s0, s1) into f / fp: real space over the Ewald rows
// (the packed records must hold s0 / s1), self term and reciprocal potential
// (InducedDipoleFieldRegion + InducedDipoleFieldReduceRegion)
template <typename T, int NT, int LPA>
__device__ void induced_field(const BatchSys &S, const BatchPrm &P, const Ws<T> &w, const double *s0,
                              const double *s1, CX<T> *buf, const CX<T> *tw) {
  phase_dip_field<T, NT, LPA, false>(S, P, w);
  if (P.recip) {
    recip_induced<T, NT, false>(S, w, s0, s1, buf, tw);
    const double a3 = 4.0 / 3.0 * P.aewald * P.aewald * P.aewald / NB_SQRT_PI;
    for (int q = threadIdx.x; q < 3 * S.n; q += NT) {
      int s = q / 3, d = q - 3 * s;
      w.f[q] += a3 * s0[q] - w.phiu[30 * (size_t)s + 21 + d];
      w.fp[q] += a3 * s1[q] - w.phiup[30 * (size_t)s + 21 + d];
    }
  }
  phase_sync();
}

// ---------------------------------------------------------------------------
// PCG (PCGSolver.java:487-604, the iteration of nb_pcg in device form).
// Returns the status; iterations and final eps through it / eps.
// ---------------------------------------------------------------------------
template <typename T, int NT, int LPA>
__device__ int phase_pcg(const BatchSys &S, const BatchPrm &P, const Ws<T> &w, CX<T> *buf, const CX<T> *tw,
                         double *red, int &iters, double &eps_out) {
  const int n = S.n, n3 = 3 * n;
  const int mixed = sizeof(T) == 4;
  void *rec = (void *)w.rec;
  double *u = S.u, *up = S.up;
  // E(u0) with u0 = u_direct (the records hold it)
  induced_field<T, NT, LPA>(S, P, w, u, up, buf, tw);
  for (int q = threadIdx.x; q < n3; q += NT) {
    double a = S.polar[q / 3];
    double rr = 0.0, rrp = 0.0;
    if (a > 0.0) {
      rr = (w.ud[q] - u[q]) / a + w.f[q];
      rrp = (w.upd[q] - up[q]) / a + w.fp[q];
    }
    w.r[q] = rr;
    w.rp[q] = rrp;
    dip_put(rec, mixed, q / 3, q % 3, a * rr, a * rrp);
  }
  phase_sync();
  phase_dip_field<T, NT, LPA, true>(S, P, w);
  phase_sync();
  double rz[2] = {0.0, 0.0};
  for (int q = threadIdx.x; q < n3; q += NT) {
    double a = S.polar[q / 3];
    double zz = a * (w.f[q] + P.psc * w.r[q]), zzp = a * (w.fp[q] + P.psc * w.rp[q]);
    w.z[q] = zz;
    w.zp[q] = zzp;
    w.p[q] = zz;
    w.pp[q] = zzp;
    dip_put(rec, mixed, q / 3, q % 3, zz, zzp);
    rz[0] += w.r[q] * zz;
    rz[1] += w.rp[q] * zzp;
  }
  block_sums<NT, 2>(rz, red);
  phase_sync();
  double eps = 100.0;
  int iter = 0, status = NB_OK;
  const double a3 = P.recip ? 4.0 / 3.0 * P.aewald * P.aewald * P.aewald / NB_SQRT_PI : 0.0;
  (void)a3;
  for (;;) {
    induced_field<T, NT, LPA>(S, P, w, w.p, w.pp, buf, tw);
    double v4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < n3; q += NT) {
      double a = S.polar[q / 3];
      double x = 0.0, xp = 0.0;
      if (a > 0.0) {
        x = w.p[q] / a - w.f[q];
        xp = w.pp[q] / a - w.fp[q];
      }
      w.vec[q] = x;
      w.vecp[q] = xp;
      v4[0] += w.p[q] * x;
      v4[1] += w.pp[q] * xp;
      v4[2] += w.r[q] * w.z[q];
      v4[3] += w.rp[q] * w.zp[q];
    }
    block_sums<NT, 4>(v4, red);
    double al = v4[2], alp = v4[3];
    if (v4[0] != 0.0) al /= v4[0];
    if (v4[1] != 0.0) alp /= v4[1];
    double rr2[2] = {0.0, 0.0};
    for (int q = threadIdx.x; q < n3; q += NT) {
      double rr = w.r[q] - al * w.vec[q], rrp = w.rp[q] - alp * w.vecp[q];
      w.r[q] = rr;
      w.rp[q] = rrp;
      u[q] += al * w.p[q];
      up[q] += alp * w.pp[q];
      double a = S.polar[q / 3];
      dip_put(rec, mixed, q / 3, q % 3, a * rr, a * rrp);
      rr2[0] += rr * rr;
      rr2[1] += rrp * rrp;
    }
    block_sums<NT, 2>(rr2, red);
    double prev = eps;
    double e = rr2[0] > rr2[1] ? rr2[0] : rr2[1];
    iter++;
    eps = 4.80321 * sqrt(e / n);
    if (!(eps <= prev)) { status = NB_ERR_SCF_DIVERGED; break; }
    if (iter >= P.maxit) { status = NB_ERR_SCF_MAXITER; break; }
    if (eps < P.polar_eps) break;
    phase_sync();
    phase_dip_field<T, NT, LPA, true>(S, P, w);
    phase_sync();
    double rzn[2] = {0.0, 0.0};
    for (int q = threadIdx.x; q < n3; q += NT) {
      double a = S.polar[q / 3];
      double zz = a * (w.f[q] + P.psc * w.r[q]), zzp = a * (w.fp[q] + P.psc * w.rp[q]);
      w.z[q] = zz;
      w.zp[q] = zzp;
      rzn[0] += w.r[q] * zz;
      rzn[1] += w.rp[q] * zzp;
    }
    block_sums<NT, 2>(rzn, red);
    double prev0 = v4[2] != 0.0 ? v4[2] : 1.0, prev1 = v4[3] != 0.0 ? v4[3] : 1.0;
    double b = rzn[0] / prev0, bp = rzn[1] / prev1;
    for (int q = threadIdx.x; q < n3; q += NT) {
      double x = w.z[q] + b * w.p[q], xp = w.zp[q] + bp * w.pp[q];
      w.p[q] = x;
      w.pp[q] = xp;
      dip_put(rec, mixed, q / 3, q % 3, x, xp);
    }
    phase_sync();
  }
  phase_sync();
  iters = iter;
  eps_out = eps;
  return status;
}

// ---------------------------------------------------------------------------
// energies, forces and torques
// ---------------------------------------------------------------------------
template <typename T, int NT, int LPA, bool POL>
__device__ void phase_real_energy(const BatchSys &S, const BatchPrm &P, const Ws<T> &w, double &esum,
                                  long long &cnt) {
  const int groups = NT / LPA, g = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  const double ecut2 = P.ewald_cut * P.ewald_cut, aw = P.recip ? P.aewald : 0.0, el = P.electric;
  const int direct = P.polarization == NB_POL_DIRECT;
  for (int base = 0; base < S.n; base += groups) {
    int i = base + g;
    bool valid = i < S.n;
    V3<double> F{0, 0, 0}, Tq{0, 0, 0};
    double e = 0.0;
    if (valid) {
      double4 pi = w.pos[i];
      AtomMP ai = w.mp[i];
      int na = w.na[i], m = na + w.nb[i];
      for (int t = lane; t < m; t += LPA) {
        uint32_t en = row_entry(S, w, i, na, t);
        int k = en & S.cell.kmask, code = en >> NB_CODE_SHIFT;
        double4 pk = w.pos[k];
        entry_shift(S.cell, en, pk.x, pk.y, pk.z);
        double dx = pk.x - pi.x, dy = pk.y - pi.y, dz = pk.z - pi.z;
        if (dx * dx + dy * dy + dz * dz > ecut2) continue;
        bool dbl = sizeof(T) == 4 && P.sc.masked[code];
        if (POL) {
          if (dbl)
            pol_energy_pair<double>(pi, ai, S.u, S.up, i, pk, w.mp[k], k, code, S.cell, aw, P.sc, direct, e, F, Tq,
                                    nullptr);
          else
            pol_energy_pair<T>(pi, ai, S.u, S.up, i, pk, w.mp[k], k, code, S.cell, aw, P.sc, direct, e, F, Tq,
                               nullptr);
        } else {
          if (dbl) perm_energy_pair<double>(pi, ai, pk, w.mp[k], code, S.cell, aw, P.sc, e, cnt, F, Tq, nullptr);
          else perm_energy_pair<T>(pi, ai, pk, w.mp[k], code, S.cell, aw, P.sc, e, cnt, F, Tq, nullptr);
        }
      }
    }
    group_sum3<LPA>(F);
    group_sum3<LPA>(Tq);
    if (valid && lane == 0) {
      add3(w.gacc, i, el * F);
      add3(w.tacc, i, el * Tq);
    }
    esum += 0.5 * el * e;
  }
}

template <typename T, int NT, int LPA>
__device__ void phase_vdw(const BatchSys &S, const BatchPrm &P, const Ws<T> &w, double &esum, long long &cnt) {
  const int groups = NT / LPA, g = threadIdx.x / LPA, lane = threadIdx.x % LPA;
  double off = P.vdw_off, cut = P.vdw_cut, off2 = off * off, cut2 = cut * cut;
  T wcut = (T)cut, iw = (T)(1.0 / (off - cut));
  T d1 = (T)P.hdelta, g1 = (T)1 + (T)P.hgamma, gg = (T)P.hgamma;
  T t1n = (T)pow(1.0 + P.hdelta, 7.0);
  for (int base = 0; base < S.n; base += groups) {
    int i = base + g;
    bool valid = i < S.n;
    V3<double> G{0, 0, 0};
    double e = 0.0;
    if (valid) {
      double4 pi = w.red[i];
      int ci = S.vclass[i];
      int na = w.na[i], m = na + w.nb[i];
      for (int t = lane; t < m; t += LPA) {
        uint32_t en = row_entry(S, w, i, na, t);
        int k = en & S.cell.kmask, code = en >> NB_CODE_SHIFT;
        double vs = P.sc.v[code];
        if (!(vs > 0.0)) continue;
        double4 pk = w.red[k];
        entry_shift(S.cell, en, pk.x, pk.y, pk.z);
        double dx = pi.x - pk.x, dy = pi.y - pk.y, dz = pi.z - pk.z;
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > off2) continue;
        const double *pe = S.vtab + 2 * (ci * S.nclass + S.vclass[k]);
        T irv = (T)pe[0];
        if (!(irv > (T)0)) continue;
        cnt++;
        T ev = (T)(vs * pe[1]);
        Vdw147<T> v = vdw147<T>(r2, irv, ev, (T)0, (T)1, d1, g1, gg, t1n, wcut, iw, r2 > cut2);
        e += (double)(v.eik * v.tp);
        acc3(G, V3<T>{v.ds * (T)dx, v.ds * (T)dy, v.ds * (T)dz});
      }
    }
    group_sum3<LPA>(G);
    if (valid && lane == 0) st3(w.gred, i, G);
    esum += 0.5 * e;
  }
}

// torques to frame-atom gradients, reduced-site chain rule, caller-order
// gradient (atomics: frame and parent atoms receive from several atoms)
template <typename T, int NT>
__device__ int phase_finalize(const BatchSys &S, const BatchPrm &P, const Ws<T> &w) {
  int bad = 0;
  for (int i = threadIdx.x; i < S.n; i += NT) {
    V3<double> g = ld3d(w.gacc, i);
    if (P.do_vdw) {
      V3<double> gv = ld3d(w.gred, i);
      int h = S.redidx[i];
      double f = S.redval[i];
      if (h >= 0 && h != i && f > 0.0) {
        g += f * gv;
        V3<double> gh = (1.0 - f) * gv;
        unsafeAtomicAdd(&S.grad[3 * h], gh.x);
        unsafeAtomicAdd(&S.grad[3 * h + 1], gh.y);
        unsafeAtomicAdd(&S.grad[3 * h + 2], gh.z);
      } else {
        g += gv;
      }
    }
    V3<double> gi = g;
    int fr = S.frame[i];
    if (P.do_pme && fr != NB_FRAME_NONE) {
      V3<double> trq = ld3d(w.tacc, i);
      if (!isfinite(trq.x) || !isfinite(trq.y) || !isfinite(trq.z)) bad = 1;
      int iz = S.axis[3 * i], ix = S.axis[3 * i + 1], iy = S.axis[3 * i + 2];
      V3<double> gz, gx, gy, vz, vx, vy;
      torque_frame(fr, i, iz, ix, iy, S.xyz, trq, gz, gx, gy, vz, vx, vy);
      unsafeAtomicAdd(&S.grad[3 * iz], gz.x);
      unsafeAtomicAdd(&S.grad[3 * iz + 1], gz.y);
      unsafeAtomicAdd(&S.grad[3 * iz + 2], gz.z);
      if (ix >= 0 && fr != NB_FRAME_ZONLY) {
        unsafeAtomicAdd(&S.grad[3 * ix], gx.x);
        unsafeAtomicAdd(&S.grad[3 * ix + 1], gx.y);
        unsafeAtomicAdd(&S.grad[3 * ix + 2], gx.z);
      }
      if (iy >= 0 && (fr == NB_FRAME_ZTHENBISECTOR || fr == NB_FRAME_THREEFOLD)) {
        unsafeAtomicAdd(&S.grad[3 * iy], gy.x);
        unsafeAtomicAdd(&S.grad[3 * iy + 1], gy.y);
        unsafeAtomicAdd(&S.grad[3 * iy + 2], gy.z);
      }
      gi = gi - (gz + gx + gy);
    }
    unsafeAtomicAdd(&S.grad[3 * i], gi.x);
    unsafeAtomicAdd(&S.grad[3 * i + 1], gi.y);
    unsafeAtomicAdd(&S.grad[3 * i + 2], gi.z);
  }
  return bad;
}

// ---------------------------------------------------------------------------
// one system
// ---------------------------------------------------------------------------
template <typename T, int NT, int LPA>
__device__ void run_system(const BatchSys &S, const BatchPrm &P, const Ws<T> &w, double *__restrict__ out,
                           CX<T> *buf, CX<T> *tw, double *red) {
  const int mixed = sizeof(T) == 4;
  const bool polar = P.do_pme && P.polarization != NB_POL_NONE;
  const bool mutual = polar && P.polarization == NB_POL_MUTUAL;
  int flags = 0;
  if (P.recip) {
    const CX<T> *twg = (const CX<T> *)S.tw;
    for (int t = threadIdx.x; t < S.K0 + S.K1 + S.K2; t += NT) tw[t] = twg[t];
  }
  // the list cutoff covers the Ewald cutoff and the vdW cutoff between
  // reduced sites (each within its largest displacement of the atom)
  double dmax[1] = {phase_atoms<T, NT>(S, P, w)};
  block_max<NT, 1>(dmax, red);
  double lcut = fmax(P.ewald_cut, P.do_vdw ? P.vdw_off + 2.0 * dmax[0] : 0.0);
  if (lcut > S.lcut) flags |= F_CUTOFF;
  phase_sync();
  double rows[2];
  {
    int mxa, mxb;
    phase_rows<T, NT, LPA>(S, P, w, fmin(lcut, S.lcut), mxa, mxb);
    rows[0] = mxa;
    rows[1] = mxb;
    block_max<NT, 2>(rows, red);
    if (rows[0] > S.maxa || rows[1] > S.maxb) flags |= F_ROWS;
  }
  phase_sync();
  double e_pr = 0.0, e_pol = 0.0, e_vdw = 0.0, e_self = 0.0, e_rec = 0.0, e_q = 0.0, e_selfi = 0.0, e_reci = 0.0;
  long long c_perm = 0, c_vdw = 0;
  int iters = 0, status = NB_OK;
  double eps = 0.0;
  if (P.do_pme && !(flags & (F_ROWS | F_CUTOFF))) {
    phase_perm_field<T, NT, LPA>(S, P, w);
    if (P.recip) {
      grid_zero<T, NT>(S, w);
      atomic_sync();
      grid_spread<T, NT, false>(S, w, nullptr, nullptr);
      atomic_sync();
      grid_convolve<T, NT>(S, w, buf, buf + FftCap<T>::value, tw);
      grid_gather<T, NT, 1, true>(S, w, w.phip, nullptr);
    }
    phase_sync();
    if (polar) {
      const double a3 = P.recip ? 4.0 / 3.0 * P.aewald * P.aewald * P.aewald / NB_SQRT_PI : 0.0;
      for (int s = threadIdx.x; s < S.n; s += NT) {
        double a = S.polar[s];
        for (int d = 0; d < 3; d++) {
          double f = w.f[3 * s + d], fp = w.fp[3 * s + d];
          if (P.recip) {
            double x = a3 * w.gmp[10 * s + 1 + d] - w.phip[30 * (size_t)s + 21 + d];
            f += x;
            fp += x;
          }
          S.u[3 * s + d] = w.ud[3 * s + d] = a * f;
          S.up[3 * s + d] = w.upd[3 * s + d] = a * fp;
          dip_put((void *)w.rec, mixed, s, d, a * f, a * fp);
        }
      }
      phase_sync();
      if (mutual) status = phase_pcg<T, NT, LPA>(S, P, w, buf, tw, red, iters, eps);
      if (P.recip) recip_induced<T, NT, true>(S, w, S.u, S.up, buf, tw);
    }
    phase_real_energy<T, NT, LPA, false>(S, P, w, e_pr, c_perm);
    phase_sync();
    if (polar) {
      phase_real_energy<T, NT, LPA, true>(S, P, w, e_pol, c_perm);
      phase_sync();
    }
    if (P.recip) {
      for (int s = threadIdx.x; s < S.n; s += NT) {
        RecipAtom o;
        recip_atom(s, w.gmp, w.fmp, w.phip, w.phiu, w.phiup, S.u, S.up, S.fr, P.electric, P.aewald, (int)polar,
                   (int)mutual, 1.0, 1.0, nullptr, o);
        e_self += o.eself;
        e_rec += o.erec;
        e_q += o.qt;
        e_selfi += o.eselfi;
        e_reci += o.ereci;
        for (int a = 0; a < 3; a++) {
          w.gacc[3 * s + a] += o.df[a] + o.dq[a];
          w.tacc[3 * s + a] += o.tq[a] + o.tqq[a];
        }
      }
    }
  }
  if (P.do_vdw && !(flags & (F_ROWS | F_CUTOFF))) phase_vdw<T, NT, LPA>(S, P, w, e_vdw, c_vdw);
  phase_sync();
  int bad = 0;
  if (!(flags & (F_ROWS | F_CUTOFF))) bad = phase_finalize<T, NT>(S, P, w);
  double v[11] = {e_pr, e_self, 0.5 * P.electric * e_rec, e_pol, e_selfi, 0.5 * P.electric * e_reci, e_vdw, e_q,
                  (double)c_vdw, (double)c_perm, (double)bad};
  block_sums<NT, 11>(v, red);
  if (v[10] != 0.0) flags |= F_TORQUE;
  if (threadIdx.x == 0) {
    out[E_PERM_REAL] = v[0];
    out[E_PERM_SELF] = v[1];
    out[E_PERM_RECIP] = v[2];
    out[E_POL_REAL] = v[3];
    out[E_POL_SELF] = v[4];
    out[E_POL_RECIP] = v[5];
    out[E_VDW] = v[6];
    out[E_QTOT] = v[7];
    out[O_ITER] = (double)iters;
    out[O_EPS] = eps;
    out[O_STAT] = (double)status;
    out[O_NVDW] = v[8];
    out[O_NPERM] = v[9];
    out[O_FLAGS] = (double)flags;
    out[O_ROWA] = rows[0];
    out[O_ROWB] = rows[1];
  }
}

// Persistent workgroups take systems from the queue (host order: largest
// first) until it is empty; every wave of a workgroup leaves the loop on the
// same index.
template <typename T, int NT, int LPA>
__global__ void __launch_bounds__(NT) k_batch(const BatchSys *__restrict__ sys, int nsys, BatchPrm P,
                                               double *__restrict__ out, int *__restrict__ next) {
  __shared__ __align__(16) CX<T> buf[2 * FftCap<T>::value];
  __shared__ __align__(16) CX<T> tw[TW_CAP];
  __shared__ double red[16 * (NT / 64)];
  // the system's descriptor, its workspace pointers and the parameters live
  // in LDS: every thread reads them with uniform addresses, and they do not
  // occupy registers across the phases
  __shared__ BatchSys Sl;
  __shared__ Ws<T> wl;
  __shared__ BatchPrm Pl;
  __shared__ int take;
  if (threadIdx.x == 0) Pl = P;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) {
      take = atomicAdd(next, 1);
      if (take < nsys) {
        Sl = sys[take];
        ws_map<T>(Sl.ws, Sl.n, Sl.maxa, Sl.maxb, (long)Sl.K0 * Sl.K1 * Sl.K2, &wl);
      }
    }
    __syncthreads();
    int b = take;
    if (b >= nsys) break;
    run_system<T, NT, LPA>(Sl, Pl, wl, out + (size_t)O_N * b, buf, tw, red);
  }
}

__global__ void k_queue_reset(int *next) {
  if (threadIdx.x == 0) next[0] = 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
Radices nb_factor(int N);

namespace {

// kernel shapes the tuner chooses from: threads per workgroup, lanes per atom
struct Shape { int nt, lpa; };
constexpr Shape kShapes[] = {{256, 8}, {256, 16}, {512, 8}, {512, 16}, {256, 4}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

template <typename T>
static const void *kernel_of(int v) {
  switch (v) {
    case 0: return (const void *)k_batch<T, 256, 8>;
    case 1: return (const void *)k_batch<T, 256, 16>;
    case 2: return (const void *)k_batch<T, 512, 8>;
    case 3: return (const void *)k_batch<T, 512, 16>;
    default: return (const void *)k_batch<T, 256, 4>;
  }
}

template <typename T>
static void launch_shape(int v, int blocks, hipStream_t st, const BatchSys *sys, int nsys, const BatchPrm &P,
                         double *out, int *next) {
  hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(64), 0, st, next);
  dim3 G(blocks);
  switch (v) {
    case 0: hipLaunchKernelGGL((k_batch<T, 256, 8>), G, dim3(256), 0, st, sys, nsys, P, out, next); break;
    case 1: hipLaunchKernelGGL((k_batch<T, 256, 16>), G, dim3(256), 0, st, sys, nsys, P, out, next); break;
    case 2: hipLaunchKernelGGL((k_batch<T, 512, 8>), G, dim3(512), 0, st, sys, nsys, P, out, next); break;
    case 3: hipLaunchKernelGGL((k_batch<T, 512, 16>), G, dim3(512), 0, st, sys, nsys, P, out, next); break;
    default: hipLaunchKernelGGL((k_batch<T, 256, 4>), G, dim3(256), 0, st, sys, nsys, P, out, next); break;
  }
  HIPCHECK(hipGetLastError());
}

// Per-device resources of the batch kernel, shared by all batches on the
// device and guarded by a mutex: workspace, descriptors, outputs, queue
// counter, stream, and the tuner's choice per batch signature.
struct Pool {
  std::mutex mu;
  int device = 0, cus = 0;
  char *ws = nullptr;
  size_t ws_cap = 0;
  BatchSys *d_sys = nullptr, *h_sys = nullptr;
  int sys_cap = 0;
  double *d_out = nullptr, *h_out = nullptr;
  int *d_next = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr, ready = nullptr;
  std::map<std::string, int> tuned;
  int occ[2][kNumShapes] = {};
  int last[4] = {0, 0, 0, 0};
  double last_ms = 0.0;   // duration of the last launch (HIP events on the pool stream)
};

std::mutex g_pools_mu;
std::map<int, Pool *> g_pools;

// The device buffers one batch launch uses (the pool's, or the static scratch
// of the allocation-free entry point) and its host mirrors.
struct BatchBuf {
  char *ws = nullptr;
  BatchSys *d_sys = nullptr, *h_sys = nullptr;
  double *d_out = nullptr, *h_out = nullptr;
  int *d_next = nullptr;
  hipStream_t stream = nullptr;
  int cus = 0;
  int (*occ)[kNumShapes] = nullptr;   // occupancy cache [form][shape]
  int shape = -1;                     // >= 0: this shape, no tuner
  Pool *tune = nullptr;               // the tuner's choices and timing events
  int used_shape = 0, used_blocks = 0;
};

// workspace bytes of the systems in queue order, and each one's offset
template <typename T>
static size_t batch_ws_bytes(std::vector<NbHandle *> &hs, const std::vector<int> &order, std::vector<size_t> &off) {
  size_t total = 0;
  off.assign(hs.size(), 0);
  for (size_t q = 0; q < hs.size(); q++) {
    NbHandle *h = hs[order[q]];
    long ntot = h->prm.do_pme ? (long)h->K[0] * h->K[1] * h->K[2] : 1;
    off[q] = total;
    total += ws_map<T>(nullptr, h->n, h->fmaxa, h->fmaxb, ntot, nullptr);
  }
  return total;
}

static Pool *pool_of(int device) {
  std::lock_guard<std::mutex> lk(g_pools_mu);
  auto it = g_pools.find(device);
  if (it != g_pools.end()) return it->second;
  Pool *p = new Pool();
  p->device = device;
  hipDeviceProp_t prop;
  HIPCHECK(hipGetDeviceProperties(&prop, device));
  p->cus = prop.multiProcessorCount;
  HIPCHECK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  HIPCHECK(hipEventCreate(&p->e0));
  HIPCHECK(hipEventCreate(&p->e1));
  HIPCHECK(hipEventCreateWithFlags(&p->ready, hipEventDisableTiming));
  HIPCHECK(hipMalloc(&p->d_next, sizeof(int) * 64));
  g_pools[device] = p;
  return p;
}

static bool same_params(const nb_params &a, const nb_params &b) {
  bool ok = a.precision == b.precision && a.do_vdw == b.do_vdw && a.do_pme == b.do_pme &&
            a.polarization == b.polarization && a.ewald_cutoff == b.ewald_cutoff && a.vdw_cutoff == b.vdw_cutoff &&
            a.ewald_alpha == b.ewald_alpha && a.vdw_taper == b.vdw_taper && a.precond_cutoff == b.precond_cutoff &&
            a.precond_scale == b.precond_scale && a.polar_eps == b.polar_eps && a.electric == b.electric &&
            a.halgren_delta == b.halgren_delta && a.halgren_gamma == b.halgren_gamma &&
            a.max_scf_iter == b.max_scf_iter && a.list_buffer == b.list_buffer && a.d11 == b.d11 &&
            a.intra14 == b.intra14;
  for (int q = 0; q < 4; q++) ok = ok && a.m_scale[q] == b.m_scale[q];
  for (int q = 0; q < 3; q++) ok = ok && a.p_scale[q] == b.p_scale[q] && a.v_scale[q] == b.v_scale[q];
  return ok;
}

// a handle the batch kernel can evaluate: a small periodic system without
// alchemistry, virial or SOR, and a cell whose image range covers the list
// cutoff
static bool batchable(const NbHandle *h, uint32_t flags) {
  if (!h->cell_set || !h->positions_set || h->cell.aperiodic || h->elec_lambda || h->ost || h->profiling ||
      h->has_valence)
    return false;
  // the batch kernel neither applies nor extends an SCF predictor's history
  if (h->predictor != NB_PRED_NONE) return false;
  if (flags & ~(uint32_t)NB_GRAD) return false;
  if (h->has_soft && h->lambda < 1.0) return false;
  if (h->prm.scf_algorithm != NB_SCF_PCG) return false;
  if (h->n > 4096) return false;
  if (h->prm.do_pme) {
    if (!(h->prm.ewald_alpha > 0.0) || !h->recip_valid) return false;
    for (int d = 0; d < 3; d++)
      if (h->K[d] > 64) return false;
  }
  double lcut = std::max(h->prm.ewald_cutoff, h->prm.vdw_cutoff) + h->prm.list_buffer;
  for (int d = 0; d < 3; d++) {
    const double(*A)[3] = h->cell.A;
    double w = 1.0 / sqrt(A[0][d] * A[0][d] + A[1][d] * A[1][d] + A[2][d] * A[2][d]);
    if ((int)floor(lcut / w) + 1 > NB_IMG_R) return false;
  }
  return true;
}

static int row_estimate(const NbHandle *h, double r) {
  double est = h->n / h->cell.vol * 4.0 / 3.0 * M_PI * r * r * r * 1.5 + 32.0;
  return (int)std::min(est, (double)h->n * 343.0);
}

// grow the pool's buffers to the batch and point b at them
template <typename T>
static void pool_buffers(Pool *pl, std::vector<NbHandle *> &hs, const std::vector<int> &order,
                         std::vector<size_t> &off, BatchBuf &b) {
  int nsys = (int)hs.size();
  size_t total = batch_ws_bytes<T>(hs, order, off);
  if (total > pl->ws_cap) {
    if (pl->ws) HIPCHECK(hipFree(pl->ws));
    pl->ws = nullptr;
    pl->ws_cap = total + total / 4;
    HIPCHECK(hipMalloc(&pl->ws, pl->ws_cap));
  }
  if (nsys > pl->sys_cap) {
    if (pl->d_sys) HIPCHECK(hipFree(pl->d_sys));
    if (pl->h_sys) HIPCHECK(hipHostFree(pl->h_sys));
    if (pl->d_out) HIPCHECK(hipFree(pl->d_out));
    if (pl->h_out) HIPCHECK(hipHostFree(pl->h_out));
    pl->sys_cap = nsys + nsys / 4 + 8;
    HIPCHECK(hipMalloc(&pl->d_sys, sizeof(BatchSys) * pl->sys_cap));
    HIPCHECK(hipHostMalloc((void **)&pl->h_sys, sizeof(BatchSys) * pl->sys_cap, hipHostMallocDefault));
    HIPCHECK(hipMalloc(&pl->d_out, sizeof(double) * O_N * pl->sys_cap));
    HIPCHECK(hipHostMalloc((void **)&pl->h_out, sizeof(double) * O_N * pl->sys_cap, hipHostMallocDefault));
  }
  b.ws = pl->ws;
  b.d_sys = pl->d_sys;
  b.h_sys = pl->h_sys;
  b.d_out = pl->d_out;
  b.h_out = pl->h_out;
  b.d_next = pl->d_next;
  b.stream = pl->stream;
  b.cus = pl->cus;
  b.occ = pl->occ;
  b.tune = pl;
}

// One launch of the batch kernel over the systems in queue order, with the
// buffers of b (workspace offsets off).  Returns 1 when some rows overflowed
// (their capacities are raised; the batch must be repeated).
template <typename T>
static int run_batch(BatchBuf &bf, std::vector<NbHandle *> &hs, BatchPrm &P, std::vector<int> &order,
                     const std::vector<size_t> &off, std::vector<char> &retry) {
  int nsys = (int)hs.size();
  const bool mixed = sizeof(T) == 4;
  for (int q = 0; q < nsys; q++) {
    NbHandle *h = hs[order[q]];
    BatchSys &S = bf.h_sys[q];
    memset(&S, 0, sizeof(S));
    S.n = h->n;
    S.K0 = h->prm.do_pme ? h->K[0] : 1;
    S.K1 = h->prm.do_pme ? h->K[1] : 1;
    S.K2 = h->prm.do_pme ? h->K[2] : 1;
    S.maxa = h->fmaxa;
    S.maxb = h->fmaxb;
    if (h->prm.do_pme) {
      const int cap = FftCap<T>::value;
      S.Lx = std::max(1, std::min(cap / S.K0, S.K1 * S.K2));
      S.Ly = std::max(1, std::min(cap / S.K1, S.K0 * S.K2));
      S.Lz = std::max(1, std::min(cap / S.K2, S.K0 * S.K1));
      for (int d = 0; d < 3; d++) S.rad[d] = nb_factor(h->K[d]);
    }
    S.lcut = std::max(h->prm.ewald_cutoff, h->prm.vdw_cutoff) + h->prm.list_buffer;
    S.cell = h->cell;
    S.cell.multi = 1;
    S.cell.kmask = (1u << NB_IMG_SHIFT) - 1u;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) S.fr.a[a][b] = h->afrac[a][b];
    S.xyz = h->d_xyz;
    S.mloc = h->d_mpole_local;
    S.polar = h->d_polar0;
    S.thole = h->d_thole0;
    S.ipd = h->d_ipd0;
    S.redval = h->d_redval;
    S.vtab = h->d_vdw_table;
    S.frame = h->d_frame;
    S.axis = h->d_axis;
    S.group = h->d_group0;
    S.vclass = h->d_vclass0;
    S.redidx = h->d_redidx;
    S.moff = h->d_mask_off;
    S.matom = h->d_mask_atom;
    S.mrel = h->d_mask_rel;
    S.nclass = h->n_class;
    S.recip = mixed ? (const void *)h->d_recipf : (const void *)h->d_recip;
    S.tw = mixed ? (const void *)h->d_twf : (const void *)h->d_tw;
    S.grad = h->d_grad;
    S.u = h->d_u;
    S.up = h->d_up;
    S.ws = bf.ws + off[q];
  }
  HIPCHECK(hipMemcpyAsync(bf.d_sys, bf.h_sys, sizeof(BatchSys) * nsys, hipMemcpyHostToDevice, bf.stream));
  // the handles' own streams may still hold position uploads
  for (NbHandle *h : hs) {
    hipError_t q = hipStreamQuery(h->stream);
    if (q == hipErrorNotReady) {
      HIPCHECK(hipEventRecord(h->ev_join, h->stream));
      HIPCHECK(hipStreamWaitEvent(bf.stream, h->ev_join, 0));
    } else if (q != hipSuccess) {
      HIPCHECK(q);
    }
  }
  const int ti = mixed ? 1 : 0;
  auto blocks_for = [&](int v) {
    if (!bf.occ[ti][v]) {
      int per = 0;
      HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel_of<T>(v), kShapes[v].nt, 0));
      bf.occ[ti][v] = std::max(per, 1);
    }
    return std::max(1, std::min(nsys, bf.occ[ti][v] * bf.cus));
  };
  int best = bf.shape;
  if (best < 0) {
    // tuner: the first batch of a signature times every shape once (after a
    // warm-up launch) and keeps the fastest
    Pool *pl = bf.tune;
    long natoms = 0;
    for (NbHandle *h : hs) natoms += h->n;
    std::string sig = std::to_string(ti) + ":" + std::to_string(nsys) + ":" + std::to_string(natoms);
    auto it = pl->tuned.find(sig);
    if (it == pl->tuned.end()) {
      float bt = 0.0f;
      best = 0;
      for (int v = 0; v < kNumShapes; v++) {
        launch_shape<T>(v, blocks_for(v), bf.stream, bf.d_sys, nsys, P, bf.d_out, bf.d_next);
        HIPCHECK(hipEventRecord(pl->e0, bf.stream));
        launch_shape<T>(v, blocks_for(v), bf.stream, bf.d_sys, nsys, P, bf.d_out, bf.d_next);
        HIPCHECK(hipEventRecord(pl->e1, bf.stream));
        HIPCHECK(hipEventSynchronize(pl->e1));
        float ms = 0.0f;
        HIPCHECK(hipEventElapsedTime(&ms, pl->e0, pl->e1));
        if (v == 0 || ms < bt) {
          bt = ms;
          best = v;
        }
      }
      pl->tuned[sig] = best;
    } else {
      best = it->second;
    }
  }
  int blocks = blocks_for(best);
  bf.used_shape = best;
  bf.used_blocks = blocks;
  if (bf.tune) {
    Pool *pl = bf.tune;
    HIPCHECK(hipEventRecord(pl->e0, bf.stream));
    launch_shape<T>(best, blocks, bf.stream, bf.d_sys, nsys, P, bf.d_out, bf.d_next);
    HIPCHECK(hipEventRecord(pl->e1, bf.stream));
    pl->last[0] = kShapes[best].nt;
    pl->last[1] = kShapes[best].lpa;
    pl->last[2] = blocks;
    pl->last[3] = nsys;
  } else {
    launch_shape<T>(best, blocks, bf.stream, bf.d_sys, nsys, P, bf.d_out, bf.d_next);
  }
  HIPCHECK(hipMemcpyAsync(bf.h_out, bf.d_out, sizeof(double) * O_N * nsys, hipMemcpyDeviceToHost, bf.stream));
  HIPCHECK(hipStreamSynchronize(bf.stream));
  if (bf.tune) {
    float ms = 0.0f;
    HIPCHECK(hipEventElapsedTime(&ms, bf.tune->e0, bf.tune->e1));
    bf.tune->last_ms = ms;
  }
  // rows that overflowed get larger capacities and the batch is repeated
  int grow = 0;
  for (int q = 0; q < nsys; q++) {
    NbHandle *h = hs[order[q]];
    const double *o = bf.h_out + (size_t)O_N * q;
    int fl = (int)o[O_FLAGS];
    if (fl & F_ROWS) {
      h->fmaxa = std::max(h->fmaxa, (int)o[O_ROWA] + (int)o[O_ROWA] / 4 + 8);
      h->fmaxb = std::max(h->fmaxb, (int)o[O_ROWB] + (int)o[O_ROWB] / 4 + 8);
      grow = 1;
    }
    if (fl & F_CUTOFF) retry[order[q]] = 1;
  }
  return grow;
}

}  // namespace

namespace {

// parameters of a batch, row capacities of new handles, and the queue order
// (largest work first: pairs ~ n^2 / V, grid ~ K^3)
static void batch_setup(std::vector<NbHandle *> &hs, BatchPrm &P, std::vector<int> &order) {
  const nb_params &pr = hs[0]->prm;
  memset(&P, 0, sizeof(P));
  P.sc = scale_tab(hs[0]);
  P.aewald = pr.ewald_alpha;
  P.ewald_cut = pr.ewald_cutoff;
  P.vdw_off = pr.vdw_cutoff;
  P.vdw_cut = pr.vdw_taper * pr.vdw_cutoff;
  P.pcut = pr.precond_cutoff;
  P.psc = pr.precond_scale;
  P.polar_eps = pr.polar_eps;
  P.electric = pr.electric;
  P.hdelta = pr.halgren_delta;
  P.hgamma = pr.halgren_gamma;
  P.polarization = pr.polarization;
  P.maxit = pr.max_scf_iter > 0 ? pr.max_scf_iter : 1000;
  P.do_vdw = pr.do_vdw;
  P.do_pme = pr.do_pme;
  P.recip = pr.do_pme && pr.ewald_alpha > 0.0;
  for (NbHandle *h : hs) {
    double lcut = std::max(pr.ewald_cutoff, pr.vdw_cutoff) + pr.list_buffer;
    if (h->fmaxa <= 0) h->fmaxa = row_estimate(h, std::min(pr.precond_cutoff, pr.ewald_cutoff));
    if (h->fmaxb <= 0) h->fmaxb = row_estimate(h, lcut);
  }
  const int count = (int)hs.size();
  order.resize(count);
  std::vector<double> cost(count);
  for (int b = 0; b < count; b++) {
    const NbHandle *h = hs[b];
    double ntot = pr.do_pme ? (double)h->K[0] * h->K[1] * h->K[2] : 0.0;
    cost[b] = (double)h->n * h->n / h->cell.vol * 40.0 + ntot * 2.0;
    order[b] = b;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
}

// energies of the systems from the launch's outputs (queue order); systems
// the batch kernel hands back (a failed solve, a cutoff beyond the checked
// image range) are marked in `separate`
static void batch_results(std::vector<NbHandle *> &hs, const BatchPrm &P, const std::vector<int> &order,
                          const double *h_out, nb_energies *out, std::vector<char> &separate, int &status) {
  const nb_params &pr = hs[0]->prm;
  const int count = (int)hs.size();
  separate.assign(count, 0);
  status = NB_OK;
  for (int q = 0; q < count; q++) {
    int b = order[q];
    NbHandle *h = hs[b];
    const double *o = h_out + (size_t)O_N * q;
    int st = (int)o[O_STAT], fl = (int)o[O_FLAGS];
    if (st != NB_OK || (fl & F_CUTOFF)) {
      separate[b] = 1;
      continue;
    }
    nb_energies e;
    memset(&e, 0, sizeof(e));
    e.perm_real = o[E_PERM_REAL];
    e.pol_real = o[E_POL_REAL];
    e.vdw = o[E_VDW];
    if (P.recip) {
      e.perm_self = o[E_PERM_SELF];
      e.perm_recip = o[E_PERM_RECIP];
      e.pol_self = o[E_POL_SELF];
      e.pol_recip = o[E_POL_RECIP];
      double qt = o[E_QTOT];
      if (qt != 0.0) e.perm_charge_corr = -0.5 * pr.electric * M_PI * qt * qt / (h->cell.vol * pr.ewald_alpha * pr.ewald_alpha);
    }
    e.n_vdw = (int64_t)(o[O_NVDW] + 0.5) / 2;
    e.n_perm = (int64_t)(o[O_NPERM] + 0.5) / 2 + (P.recip ? h->n : 0);
    e.pcg_iters = (int)o[O_ITER];
    e.pcg_eps = o[O_EPS];
    e.list_rebuilt = 1;
    e.perm = e.perm_real + e.perm_self + e.perm_recip + e.perm_charge_corr;
    e.pol = e.pol_real + e.pol_self + e.pol_recip;
    e.total = e.perm + e.pol + e.vdw;
    int s = NB_OK;
    if (fl & F_TORQUE) s = NB_ERR_BAD_TORQUE;
    else if (!std::isfinite(e.total)) s = NB_ERR_NAN_ENERGY;
    if (s != NB_OK) {
      h->err = s == NB_ERR_BAD_TORQUE ? "undefined torque" : "NaN energy";
      if (status == NB_OK) status = s;
    }
    h->want_vir = false;
    h->vir_valid = false;
    h->want_dudl = false;
    h->dudl_valid = false;
    h->induced_caller = true;
    h->last = e;
    if (out) out[b] = e;
  }
}

static bool batch_qualifies(std::vector<NbHandle *> &hs, uint32_t flags) {
  for (NbHandle *h : hs)
    if (!h || !batchable(h, flags) || h->device != hs[0]->device || !same_params(h->prm, hs[0]->prm)) return false;
  return true;
}

// ---------------------------------------------------------------------------
// Static scratch of the allocation-free entry point (nb_evaluate_batch_static):
// one launch's workspace, system descriptors, outputs and queue counter.  Each
// GPU the library runs on holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress gives on that
// device (scratch_base), never through the array's name.
// ---------------------------------------------------------------------------
constexpr size_t NB_SCRATCH_BYTES = (size_t)NB_BATCH_SCRATCH_MIB << 20;
static __device__ __align__(256) char scratch[NB_SCRATCH_BYTES];

std::mutex g_scratch_mu;          // guards g_scratch
std::mutex g_scratch_run_mu;      // one allocation-free batch at a time in the process
std::map<int, char *> g_scratch;   // device -> address of its scratch

// the device's scratch address (looked up once per device, with it current)
static char *scratch_base(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(device);
  if (it != g_scratch.end()) return it->second;
  char *ptr = nullptr;
  HIPCHECK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw std::runtime_error("scratch symbol has no device address");
  g_scratch[device] = ptr;
  return ptr;
}

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// Evaluate the handles with the batch kernel.  Returns false (nothing done)
// when some handle does not qualify; otherwise fills out / the handles, and
// marks in `separate` the handles whose evaluation must be repeated on the
// regular path (a failed solve, which the regular path hands to SOR, or a
// list cutoff beyond the checked image range).
bool nb_batch_evaluate(NbHandle **hv, int count, uint32_t flags, nb_energies *out, std::vector<char> &separate,
                       int &status) {
  if (count <= 0) return false;
  std::vector<NbHandle *> hs(hv, hv + count);
  if (!batch_qualifies(hs, flags)) return false;
  HIPCHECK(hipSetDevice(hs[0]->device));
  Pool *pl = pool_of(hs[0]->device);
  std::lock_guard<std::mutex> lk(pl->mu);
  BatchPrm P;
  std::vector<int> order;
  batch_setup(hs, P, order);
  const bool mixed = hs[0]->prm.precision == NB_PREC_MIXED;
  std::vector<char> retry(count, 0);
  std::vector<size_t> off;
  BatchBuf bf;
  for (int attempt = 0;; attempt++) {
    int grow;
    if (mixed) {
      pool_buffers<float>(pl, hs, order, off, bf);
      grow = run_batch<float>(bf, hs, P, order, off, retry);
    } else {
      pool_buffers<double>(pl, hs, order, off, bf);
      grow = run_batch<double>(bf, hs, P, order, off, retry);
    }
    if (!grow) break;
    if (attempt == 2) throw std::runtime_error("batch row capacity could not be established");
  }
  batch_results(hs, P, order, pl->h_out, out, separate, status);
  return true;
}

// The batch kernel with its buffers in the static scratch; no device or
// pinned allocation.  See nb_evaluate_batch_static in amoeba_nb.h.
int nb_batch_evaluate_static(NbHandle **hv, int count, uint32_t flags, nb_energies *out) {
  if (count <= 0) return NB_ERR_INVALID;
  std::vector<NbHandle *> hs(hv, hv + count);
  if (!batch_qualifies(hs, flags)) {
    hs[0]->err = "nb_evaluate_batch_static: every handle must qualify for the batch kernel (small periodic "
                 "systems with the same parameters on one device, NB_GRAD only, no alchemistry, valence terms, "
                 "SCF predictor or profiling)";
    return NB_ERR_INVALID;
  }
  const int device = hs[0]->device;
  HIPCHECK(hipSetDevice(device));
  char *base = scratch_base(device);
  std::lock_guard<std::mutex> run_lock(g_scratch_run_mu);
  BatchPrm P;
  std::vector<int> order;
  batch_setup(hs, P, order);
  const bool mixed = hs[0]->prm.precision == NB_PREC_MIXED;
  std::vector<BatchSys> h_sys(count);
  std::vector<double> h_out((size_t)O_N * count);
  std::vector<char> retry(count, 0);
  std::vector<size_t> off;
  BatchBuf bf;
  int cus = 0;
  HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  static int occ[2][kNumShapes] = {};
  for (int attempt = 0;; attempt++) {
    // layout: queue counter | descriptors | outputs | workspace
    size_t ws = mixed ? batch_ws_bytes<float>(hs, order, off) : batch_ws_bytes<double>(hs, order, off);
    size_t o_sys = 256, o_out = o_sys + up256(sizeof(BatchSys) * count), o_ws = o_out + up256(sizeof(double) * O_N * count);
    if (o_ws + ws > NB_SCRATCH_BYTES) {
      hs[0]->err = "nb_evaluate_batch_static: the batch needs more device memory than the static scratch holds "
                   "(NB_BATCH_SCRATCH_MIB); use nb_evaluate_batch";
      return NB_ERR_INVALID;
    }
    bf.d_next = (int *)base;
    bf.d_sys = (BatchSys *)(base + o_sys);
    bf.d_out = (double *)(base + o_out);
    bf.ws = base + o_ws;
    bf.h_sys = h_sys.data();
    bf.h_out = h_out.data();
    bf.stream = hs[0]->stream;
    bf.cus = cus;
    bf.occ = occ;
    bf.shape = kNumShapes - 1;   // 256 threads, 4 lanes per atom
    bf.tune = nullptr;
    int grow = mixed ? run_batch<float>(bf, hs, P, order, off, retry) : run_batch<double>(bf, hs, P, order, off, retry);
    if (!grow) break;
    if (attempt == 2) throw std::runtime_error("batch row capacity could not be established");
  }
  std::vector<char> separate;
  int status = NB_OK;
  batch_results(hs, P, order, h_out.data(), out, separate, status);
  for (int b = 0; b < count; b++)
    if (separate[b]) {
      hs[b]->err = "nb_evaluate_batch_static: the batch kernel handed this system back (failed solve or a "
                   "cutoff beyond the image range); evaluate it with nb_evaluate";
      if (status == NB_OK) status = NB_ERR_SCF_DIVERGED;
    }
  return status;
}

int nb_batch_last_config(int device, int32_t cfg[4], double *kernel_ms) {
  Pool *pl = pool_of(device);
  std::lock_guard<std::mutex> lk(pl->mu);
  for (int q = 0; q < 4; q++) cfg[q] = pl->last[q];
  if (kernel_ms) *kernel_ms = pl->last_ms;
  return NB_OK;
}
