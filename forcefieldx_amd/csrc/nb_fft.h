// nb_fft.h -- grid routines shared by the per-system PME kernels and the
// batched small-cell kernel: in-LDS Stockham line FFTs (radix 2/3/4/5/7) over a
// chunk of grid lines, and the potential gather of one atom.
#pragma once
#include "nb_atom.h"

template <typename T> struct alignas(2 * sizeof(T)) CX { T x, y; };

__device__ __forceinline__ int wrapg(int i, int n) { return i >= n ? i - n : (i < 0 ? i + n : i); }

// ---------------------------------------------------------------------------
// FFT: in-place Stockham autosort in LDS, radix 2/3/4/5/7.  A pass reads every
// butterfly input of the workgroup into registers, synchronises, and writes
// the outputs back to the same buffer, so one LDS buffer suffices.  Lines
// live at buf[l * sl + e * se]; LINE_FAST puts consecutive threads on
// consecutive lines (used when lines are the contiguous direction).
// ---------------------------------------------------------------------------
struct Radices { int r[12]; int nr; };

template <typename T>
__device__ __forceinline__ CX<T> cmul(CX<T> a, CX<T> b) {
  return CX<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ CX<T> cadd(CX<T> a, CX<T> b) { return CX<T>{a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ CX<T> csub(CX<T> a, CX<T> b) { return CX<T>{a.x - b.x, a.y - b.y}; }
// i * s * a for s = +-1
template <typename T> __device__ __forceinline__ CX<T> cis(CX<T> a, T s) { return CX<T>{-s * a.y, s * a.x}; }

// y[m] = sum_k x[k] exp(s 2 pi i m k / R), s = +-1
template <int R, typename T>
__device__ __forceinline__ void small_dft(CX<T> *v, T s) {
  if (R == 2) {
    CX<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 3) {
    const T c = (T)-0.5, sn = (T)0.86602540378443864676;  // cos, sin of 2 pi / 3
    CX<T> t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    CX<T> m{v[0].x + c * t.x, v[0].y + c * t.y};
    CX<T> r = cis(CX<T>{sn * d.x, sn * d.y}, s);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  } else if (R == 4) {
    CX<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]), c = cadd(v[1], v[3]), d = cis(csub(v[1], v[3]), s);
    v[0] = cadd(a, c);
    v[1] = cadd(b, d);
    v[2] = csub(a, c);
    v[3] = csub(b, d);
  } else if (R == 7) {
    // X_m = v0 + sum_k cos(2 pi k m / 7) t_k + i s sum_k sin(2 pi k m / 7) d_k,
    // t_k = v_k + v_(7-k), d_k = v_k - v_(7-k); X_(7-m) takes the minus sign
    const T c1 = (T)0.62348980185873353, c2 = (T)-0.22252093395631440, c3 = (T)-0.90096886790241913;
    const T s1 = (T)0.78183148246802981, s2 = (T)0.97492791218182361, s3 = (T)0.43388373911755812;
    CX<T> t1 = cadd(v[1], v[6]), d1 = csub(v[1], v[6]);
    CX<T> t2 = cadd(v[2], v[5]), d2 = csub(v[2], v[5]);
    CX<T> t3 = cadd(v[3], v[4]), d3 = csub(v[3], v[4]);
    CX<T> m1{v[0].x + c1 * t1.x + c2 * t2.x + c3 * t3.x, v[0].y + c1 * t1.y + c2 * t2.y + c3 * t3.y};
    CX<T> m2{v[0].x + c2 * t1.x + c3 * t2.x + c1 * t3.x, v[0].y + c2 * t1.y + c3 * t2.y + c1 * t3.y};
    CX<T> m3{v[0].x + c3 * t1.x + c1 * t2.x + c2 * t3.x, v[0].y + c3 * t1.y + c1 * t2.y + c2 * t3.y};
    CX<T> r1 = cis(CX<T>{s1 * d1.x + s2 * d2.x + s3 * d3.x, s1 * d1.y + s2 * d2.y + s3 * d3.y}, s);
    CX<T> r2 = cis(CX<T>{s2 * d1.x - s3 * d2.x - s1 * d3.x, s2 * d1.y - s3 * d2.y - s1 * d3.y}, s);
    CX<T> r3 = cis(CX<T>{s3 * d1.x - s1 * d2.x + s2 * d3.x, s3 * d1.y - s1 * d2.y + s2 * d3.y}, s);
    v[0] = cadd(v[0], cadd(t1, cadd(t2, t3)));
    v[1] = cadd(m1, r1);
    v[6] = csub(m1, r1);
    v[2] = cadd(m2, r2);
    v[5] = csub(m2, r2);
    v[3] = cadd(m3, r3);
    v[4] = csub(m3, r3);
  } else {
    // cos and sin of 2 pi / 5 and 4 pi / 5
    const T c1 = (T)0.30901699437494742410, s1 = (T)0.95105651629515357212;
    const T c2 = (T)-0.80901699437494742410, s2 = (T)0.58778525229247312917;
    CX<T> t1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]), t2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
    CX<T> m1{v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
    CX<T> m2{v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
    CX<T> r1 = cis(CX<T>{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y}, s);
    CX<T> r2 = cis(CX<T>{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y}, s);
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(m1, r1);
    v[4] = csub(m1, r1);
    v[2] = cadd(m2, r2);
    v[3] = csub(m2, r2);
  }
}

// one in-place pass; EPT bounds the elements per thread (N * nlines <= EPT * blockDim)
template <int R, int EPT, bool LINE_FAST, typename T>
__device__ __forceinline__ void lds_pass(CX<T> *buf, int N, int nlines, int sl, int se, int Ns, const CX<T> *tw, T s) {
  constexpr int MAXB = (EPT + R - 1) / R;
  const int nb = N / R, total = nlines * nb, stride = N / (Ns * R);
  CX<T> v[MAXB][R];
  int where[MAXB];
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    int b = q * blockDim.x + threadIdx.x;
    where[q] = -1;
    if (b < total) {
      int line, j;
      if (LINE_FAST) { line = b % nlines; j = b / nlines; }
      else { j = b % nb; line = b / nb; }
      int jm = j % Ns;
      const CX<T> *x = buf + line * sl;
#pragma unroll
      for (int k = 0; k < R; k++) {
        CX<T> a = x[(j + k * nb) * se];
        if (k > 0 && jm > 0) {
          CX<T> w = tw[jm * k * stride];  // jm k stride < N
          w.y *= s;
          a = cmul(a, w);
        }
        v[q][k] = a;
      }
      small_dft<R, T>(v[q], s);
      where[q] = line * sl + ((j / Ns) * Ns * R + jm) * se;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; q++) {
    if (where[q] >= 0) {
#pragma unroll
      for (int k = 0; k < R; k++) buf[where[q] + k * Ns * se] = v[q][k];
    }
  }
  __syncthreads();
}

template <int EPT, bool LINE_FAST, typename T>
__device__ void lds_fft(CX<T> *buf, int N, int nlines, int sl, int se, const Radices &rad, const CX<T> *tw, T s) {
  int Ns = 1;
  for (int p = 0; p < rad.nr; p++) {
    int R = rad.r[p];
    if (R == 4) lds_pass<4, EPT, LINE_FAST, T>(buf, N, nlines, sl, se, Ns, tw, s);
    else if (R == 2) lds_pass<2, EPT, LINE_FAST, T>(buf, N, nlines, sl, se, Ns, tw, s);
    else if (R == 3) lds_pass<3, EPT, LINE_FAST, T>(buf, N, nlines, sl, se, Ns, tw, s);
    else if (R == 5) lds_pass<5, EPT, LINE_FAST, T>(buf, N, nlines, sl, se, Ns, tw, s);
    else lds_pass<7, EPT, LINE_FAST, T>(buf, N, nlines, sl, se, Ns, tw, s);
    Ns *= R;
  }
}

// One chunk of L lines along one axis, processed by the whole workgroup:
// AXIS 0 -> L consecutive rows; AXIS 1 -> an x block of L at fixed z; AXIS 2 ->
// L consecutive (x, y) line indices, with the influence multiply and inverse
// transform fused (CONV); AXIS 3 -> lines along y over L consecutive (x, z)
// line indices x + K0 z (several z-planes per chunk).  buf holds L * N
// values, tw the N twiddles.
template <typename T, int EPT, int AXIS, bool CONV>
__device__ __forceinline__ void fft_chunk(CX<T> *__restrict__ grid, int K0, int K1, int K2, int L, int chunk,
                                          const Radices &rad, CX<T> *buf, const CX<T> *tw, T s,
                                          const T *__restrict__ recip) {
  const int N = (AXIS == 0) ? K0 : ((AXIS == 1 || AXIS == 3) ? K1 : K2);
  int cnt, first = 0, xb = 0, z = 0;
  if (AXIS == 0) {
    first = chunk * L;
    cnt = min(L, K1 * K2 - first);
  } else if (AXIS == 3) {
    first = chunk * L;
    cnt = min(L, K0 * K2 - first);
  } else if (AXIS == 1) {
    int nxb = (K0 + L - 1) / L;
    xb = (chunk % nxb) * L;
    z = chunk / nxb;
    cnt = min(L, K0 - xb);
  } else {
    first = chunk * L;
    cnt = min(L, K0 * K1 - first);
  }
  auto gidx = [&](int l, int e) -> size_t {
    if (AXIS == 0) return (size_t)(first + l) * K0 + e;
    if (AXIS == 3) {
      int q = first + l, x = q % K0, zz = q / K0;
      return (size_t)x + (size_t)K0 * (e + (size_t)K1 * zz);
    }
    if (AXIS == 1) return (size_t)(xb + l) + (size_t)K0 * (e + (size_t)K1 * z);
    return (size_t)(first + l) + (size_t)K0 * K1 * e;
  };
  // LDS layout: AXIS 0 lines contiguous (buf[l N + e]); otherwise lines fastest (buf[e L + l])
  for (int t = threadIdx.x; t < L * N; t += blockDim.x) {
    int l, e;
    if (AXIS == 0) { l = t / N; e = t - l * N; }
    else { e = t / L; l = t - e * L; }
    CX<T> v{(T)0, (T)0};
    if (l < cnt) v = grid[gidx(l, e)];
    buf[t] = v;
  }
  __syncthreads();
  if (AXIS == 0) lds_fft<EPT, false, T>(buf, N, L, N, 1, rad, tw, s);
  else lds_fft<EPT, true, T>(buf, N, L, 1, L, rad, tw, s);
  if (CONV) {
    for (int t = threadIdx.x; t < L * N; t += blockDim.x) {
      int e = t / L, l = t - e * L;
      if (l < cnt) {
        T g = recip[(size_t)(first + l) + (size_t)K0 * K1 * e];
        buf[t].x *= g;
        buf[t].y *= g;
      }
    }
    __syncthreads();
    lds_fft<EPT, true, T>(buf, N, L, 1, L, rad, tw, -s);
  }
  for (int t = threadIdx.x; t < L * N; t += blockDim.x) {
    int l, e;
    if (AXIS == 0) { l = t / N; e = t - l * N; }
    else { e = t / L; l = t - e * L; }
    if (l < cnt) grid[gidx(l, e)] = buf[t];
  }
}

template <typename T>
__device__ __forceinline__ T group8(T v) {
#pragma unroll
  for (int o = 4; o > 0; o >>= 1) v += __shfl_xor(v, o, 8);
  return v;
}

// Potential of atom s from the convolved grid (ReciprocalSpace.java:2264-2455,
// 2467-2776): 8 consecutive lanes share the atom, lane z < 5 sums the (x, y)
// stencil plane z; FULL gives the 20 fractional derivatives and the Cartesian
// transform (out[30 s ..]), otherwise only the Cartesian field gradient
// (out[30 s + 21 ..]).  NP = 2 reads u from the real and uCR from the
// imaginary part.  Every lane of the 8 must call (cross-lane sums).
template <typename T, int NP, bool FULL, bool PREFETCH = false, typename TH = double>
__device__ __forceinline__ void gather_atom(int s, int lane, bool valid, const TH *__restrict__ theta,
                                            const int32_t *__restrict__ igrid, const CX<T> *__restrict__ grid,
                                            const Frac &fr, int K0, int K1, int K2, double *__restrict__ out0,
                                            double *__restrict__ out1, double *fld = nullptr) {
  constexpr int NC = FULL ? 20 : 4;
  T phi[NP][NC];
#pragma unroll
  for (int p = 0; p < NP; p++)
#pragma unroll
    for (int c = 0; c < NC; c++) phi[p][c] = (T)0;
  if (valid && lane < 5) {
    const TH *tx = theta + (size_t)s * 60, *ty = tx + 20, *vd = tx + 40 + 4 * lane;
    T v[4] = {(T)vd[0], (T)vd[1], (T)vd[2], (T)vd[3]};
    T wx[5][4];
#pragma unroll
    for (int kx = 0; kx < 5; kx++)
#pragma unroll
      for (int q = 0; q < 4; q++) wx[kx][q] = (T)tx[4 * kx + q];
    int k = wrapg(igrid[3 * s + 2] + 1 + lane, K2);
    int j0 = igrid[3 * s + 1] + 1, i0 = igrid[3 * s] + 1;
    int ix[5];
#pragma unroll
    for (int kx = 0; kx < 5; kx++) ix[kx] = wrapg(i0 + kx, K0);
    T tu[NP][10];
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
      for (int c = 0; c < 10; c++) tu[p][c] = (T)0;
    // PREFETCH: the 25 grid values of the lane's plane are loaded before the
    // first is used (one memory latency instead of five)
    CX<T> pre[PREFETCH ? 5 : 1][PREFETCH ? 5 : 1];
    if (PREFETCH) {
#pragma unroll
      for (int ky = 0; ky < 5; ky++) {
        size_t row = (size_t)K0 * (wrapg(j0 + ky, K1) + (size_t)K1 * k);
#pragma unroll
        for (int kx = 0; kx < 5; kx++) pre[PREFETCH ? ky : 0][PREFETCH ? kx : 0] = grid[row + ix[kx]];
      }
    }
#pragma unroll
    for (int ky = 0; ky < 5; ky++) {
      int j = wrapg(j0 + ky, K1);
      size_t row = (size_t)K0 * (j + (size_t)K1 * k);
      T t[NP][4];
#pragma unroll
      for (int p = 0; p < NP; p++) t[p][0] = t[p][1] = t[p][2] = t[p][3] = (T)0;
#pragma unroll
      for (int kx = 0; kx < 5; kx++) {
        CX<T> q2 = PREFETCH ? pre[PREFETCH ? ky : 0][PREFETCH ? kx : 0] : grid[row + ix[kx]];
        T qv[2] = {q2.x, q2.y};
#pragma unroll
        for (int p = 0; p < NP; p++) {
          T q = qv[p];
          t[p][0] += q * wx[kx][0]; t[p][1] += q * wx[kx][1];
          if (FULL) { t[p][2] += q * wx[kx][2]; t[p][3] += q * wx[kx][3]; }
        }
      }
      const TH *ud = ty + 4 * ky;
      T uu[4] = {(T)ud[0], (T)ud[1], (T)ud[2], (T)ud[3]};
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
      T *ph = phi[p];
      const T *a = tu[p];
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
    for (int c = 0; c < NC; c++) phi[p][c] = group8(phi[p][c]);
  // fld (when given): the Cartesian field of each part in every lane of the 8
  if (!FULL && fld) {
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
      for (int c = 0; c < 3; c++)
        fld[3 * p + c] =
            fr.a[0][c] * (double)phi[p][1] + fr.a[1][c] * (double)phi[p][2] + fr.a[2][c] * (double)phi[p][3];
  }
  if (valid && lane == 0) {
    double *outs[2] = {out0, out1};
#pragma unroll
    for (int p = 0; p < NP; p++) {
      double *o = outs[p] + 30 * (size_t)s;
      if (FULL) {
        for (int c = 0; c < 20; c++) o[c] = (double)phi[p][c];
        phi_cart<T>(fr, phi[p], o + 20);
      } else {
        for (int c = 0; c < 3; c++)
          o[21 + c] = fr.a[0][c] * (double)phi[p][1] + fr.a[1][c] * (double)phi[p][2] + fr.a[2][c] * (double)phi[p][3];
      }
    }
  }
}


