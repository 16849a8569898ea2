// nb_device.h -- device math shared by the AMOEBA nonbonded kernels (gfx950).
//
// Pair terms are templated on the arithmetic type T (double for the FP64
// form, float for the mixed form).  Displacements are always formed in FP64
// and rounded to T afterwards, so the mixed form never loses the relative
// precision of coordinates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NB_SQRT_PI 1.77245385090551602729

// ---------------------------------------------------------------------------
// unit cell (Crystal.java:1425-1545, image at :1039-1083)
// ---------------------------------------------------------------------------
struct CellD {
  double A00, A10, A11, A20, A21, A22;        // fractional = cart . A (A01=A02=A12=0)
  double Ai00, Ai10, Ai11, Ai20, Ai21, Ai22;  // cart = frac . Ai
  double A[3][3];
  double vol;
  int aperiodic;
  // Cells thinner than twice the cutoff use explicit image sums: a list entry
  // then carries the lattice translation of its partner in bits 18..26 and
  // the partner index in bits 0..17 (kmask); otherwise the minimum image.
  int multi;
  uint32_t kmask;
};

#define NB_IMG_SHIFT 18
#define NB_IMG_R 3
#define NB_IMG_W 7
#define NB_IMG_ORIGIN (NB_IMG_R + NB_IMG_W * (NB_IMG_R + NB_IMG_W * NB_IMG_R))

// lattice translation of a list entry's partner (image sums only)
__device__ __forceinline__ void entry_shift(const CellD &c, uint32_t en, double &x, double &y, double &z) {
  if (!c.multi) return;
  int code = (int)((en >> NB_IMG_SHIFT) & 511u);
  int tx = code % NB_IMG_W - NB_IMG_R, ty = (code / NB_IMG_W) % NB_IMG_W - NB_IMG_R;
  int tz = code / (NB_IMG_W * NB_IMG_W) - NB_IMG_R;
  x += tx * c.Ai00 + ty * c.Ai10 + tz * c.Ai20;
  y += ty * c.Ai11 + tz * c.Ai21;
  z += tz * c.Ai22;
}

__device__ __forceinline__ double round_half_away(double x) { return copysign(floor(fabs(x) + 0.5), x); }

// minimum image with the reference's independent fractional rounding (image
// sums: the displacement already holds its translation)
__device__ __forceinline__ double image(const CellD &c, double &dx, double &dy, double &dz) {
  if (c.aperiodic || c.multi) return dx * dx + dy * dy + dz * dz;
  double xf = dx * c.A00 + dy * c.A10 + dz * c.A20;
  double yf = dy * c.A11 + dz * c.A21;
  double zf = dz * c.A22;
  xf -= round_half_away(xf);
  yf -= round_half_away(yf);
  zf -= round_half_away(zf);
  dx = xf * c.Ai00 + yf * c.Ai10 + zf * c.Ai20;
  dy = yf * c.Ai11 + zf * c.Ai21;
  dz = zf * c.Ai22;
  return dx * dx + dy * dy + dz * dz;
}

template <typename T> struct V3 { T x, y, z; };
template <typename T> __device__ __forceinline__ V3<T> mk3(T a, T b, T c) { return V3<T>{a, b, c}; }
template <typename T> __device__ __forceinline__ T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> __device__ __forceinline__ V3<T> cross(const V3<T> &a, const V3<T> &b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> __device__ __forceinline__ V3<T> operator+(const V3<T> &a, const V3<T> &b) { return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator-(const V3<T> &a, const V3<T> &b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator*(T s, const V3<T> &a) { return V3<T>{s * a.x, s * a.y, s * a.z}; }
template <typename T> __device__ __forceinline__ V3<T> operator-(const V3<T> &a) { return V3<T>{-a.x, -a.y, -a.z}; }
template <typename T> __device__ __forceinline__ void operator+=(V3<T> &a, const V3<T> &b) { a.x += b.x; a.y += b.y; a.z += b.z; }
template <typename T> __device__ __forceinline__ void operator-=(V3<T> &a, const V3<T> &b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; }

// symmetric 3x3 (xx yy zz xy xz yz)
template <typename T> struct S3 { T xx, yy, zz, xy, xz, yz; };
template <typename T> __device__ __forceinline__ V3<T> mul(const S3<T> &q, const V3<T> &v) {
  return V3<T>{q.xx * v.x + q.xy * v.y + q.xz * v.z, q.xy * v.x + q.yy * v.y + q.yz * v.z,
               q.xz * v.x + q.yz * v.y + q.zz * v.z};
}

// multipole with the quadrupole pre-divided by 3 (as the real-space loops use it)
template <typename T> struct MP { T c; V3<T> d; S3<T> q; };

// Per-atom record in sorted order: 10 multipole components (Q/3), 1/pdamp, thole
struct AtomMP { double c, dx, dy, dz, qxx, qyy, qzz, qxy, qxz, qyz, ipd, thole; };

template <typename T> __device__ __forceinline__ MP<T> load_mp(const AtomMP &a) {
  MP<T> m;
  m.c = (T)a.c;
  m.d = V3<T>{(T)a.dx, (T)a.dy, (T)a.dz};
  m.q = S3<T>{(T)a.qxx, (T)a.qyy, (T)a.qzz, (T)a.qxy, (T)a.qxz, (T)a.qyz};
  return m;
}

// ---------------------------------------------------------------------------
// radial functions: bn (Ewald, EwaldParameters.java:100-125) and bare rr
// ---------------------------------------------------------------------------
__device__ __forceinline__ double nb_erfc(double x) { return erfc(x); }
__device__ __forceinline__ float nb_erfc(float x) { return erfcf(x); }
__device__ __forceinline__ double nb_exp(double x) { return exp(x); }
__device__ __forceinline__ float nb_exp(float x) { return __expf(x); }
__device__ __forceinline__ double nb_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float nb_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double nb_rsqrt(double x) { return 1.0 / sqrt(x); }
__device__ __forceinline__ float nb_rsqrt(float x) { return rsqrtf(x); }

// r and 1/r from r^2, and reciprocals inside the pair loops: FP64 by the
// correctly rounded sqrt and division; FP32 by the hardware reciprocal square
// root and reciprocal (v_rsq_f32 / v_rcp_f32, about 1 ulp), which replace the
// ~30-instruction correctly rounded sqrt + division sequence of the FP32 form
// (the mixed form's pair error budget is 1e-5, its pair terms err ~1e-7)
__device__ __forceinline__ void r_rinv(double r2, double &r, double &rinv) {
  r = sqrt(r2);
  rinv = 1.0 / r;
}
__device__ __forceinline__ void r_rinv(double r2, float &r, float &rinv) {
  float x = (float)r2;
  rinv = __builtin_amdgcn_rsqf(x);
  r = x * rinv;
}
__device__ __forceinline__ double nb_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float nb_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// erfc(x) given e^{-x^2}: FP64 calls erfc; the FP32 form uses the rational
// approximation of Abramowitz & Stegun, Handbook of Mathematical Functions,
// eq. 7.1.26 (after Hastings), erfc(x) = t (a1 + t (a2 + t (a3 + t (a4 + t a5))))
// e^{-x^2} with t = 1 / (1 + p x), absolute error <= 1.5e-7 for x >= 0, which
// reuses the exponential the Ewald recursion needs anyway.
__device__ __forceinline__ double nb_erfc_ex(double x, double) { return erfc(x); }
__device__ __forceinline__ float nb_erfc_ex(float x, float ex) {
  float t = nb_rcp(1.0f + 0.3275911f * x);
  float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f + t * 1.061405429f))));
  return poly * ex;
}

// B[0..NMAX]: B0 = erfc(a r)/r, Bk = ((2k-1) B(k-1) + an(k-1) e^{-a^2 r^2}) / r^2
template <typename T, int NMAX>
__device__ __forceinline__ void ewald_bn(T r, T rinv, T aewald, T *B) {
  T rr2 = rinv * rinv;
  if (aewald > (T)0) {
    T ra = aewald * r;
    T ex = nb_exp(-ra * ra);
    T alsq2 = (T)2 * aewald * aewald;
    T an = alsq2 / ((T)NB_SQRT_PI * aewald);
    B[0] = nb_erfc_ex(ra, ex) * rinv;
#pragma unroll
    for (int k = 1; k <= NMAX; k++) {
      B[k] = ((T)(2 * k - 1) * B[k - 1] + an * ex) * rr2;
      an *= alsq2;
    }
  } else {
    B[0] = rinv;
#pragma unroll
    for (int k = 1; k <= NMAX; k++) B[k] = (T)(2 * k - 1) * B[k - 1] * rr2;
  }
}

template <typename T, int NMAX>
__device__ __forceinline__ void coulomb_rr(T rinv, T *R) {
  T rr2 = rinv * rinv;
  R[0] = rinv;
#pragma unroll
  for (int k = 1; k <= NMAX; k++) R[k] = (T)(2 * k - 1) * R[k - 1] * rr2;
}

// Thole screening (PermanentFieldRegion.java:726-739) and its gradient
// factors (RealSpaceEnergyRegion.java:960-980)
template <typename T> struct Thole { T s3, s5, s7; V3<T> d3, d5, d7; };

template <typename T, bool GRAD>
__device__ __forceinline__ Thole<T> thole(T r, const V3<T> &rv, T ipdi, T ipdk, T ti, T tk) {
  Thole<T> t;
  t.s3 = t.s5 = t.s7 = (T)1;
  if (GRAD) { t.d3 = t.d5 = t.d7 = V3<T>{0, 0, 0}; }
  T pg = ti < tk ? ti : tk;
  T rd = r * ipdi * ipdk;
  T damp = -pg * rd * rd * rd;
  if (damp > (T)-50) {
    T e = nb_exp(damp);
    t.s3 = (T)1 - e;
    t.s5 = (T)1 - e * ((T)1 - damp);
    t.s7 = (T)1 - e * ((T)1 - damp + (T)0.6 * damp * damp);
    if (GRAD) {
      T c3 = (T)-3 * damp * e / (r * r);
      t.d3 = c3 * rv;
      t.d5 = (-damp) * t.d3;
      t.d7 = ((T)-0.2 - (T)0.6 * damp) * t.d5;
    }
  }
  return t;
}

// field at displacement R from a multipole for radial set (B1,B2,B3):
// E = R (B1 c + B2 d.R + B3 R.Q.R) - B1 d - 2 B2 Q R   (PermanentFieldRegion.java:760-808)
template <typename T>
__device__ __forceinline__ V3<T> mp_field(const V3<T> &R, const MP<T> &m, T B1, T B2, T B3) {
  V3<T> qr = mul(m.q, R);
  T s = B1 * m.c + B2 * dot(m.d, R) + B3 * dot(R, qr);
  return V3<T>{R.x * s - B1 * m.d.x - (T)2 * B2 * qr.x, R.y * s - B1 * m.d.y - (T)2 * B2 * qr.y,
               R.z * s - B1 * m.d.z - (T)2 * B2 * qr.z};
}

// field of a point dipole: -B1 u + B2 (u.R) R   (InducedDipoleFieldRegion.java:400-450)
template <typename T>
__device__ __forceinline__ V3<T> dip_field(const V3<T> &R, const V3<T> &u, T B1, T B2) {
  T s = B2 * dot(u, R);
  return V3<T>{-B1 * u.x + s * R.x, -B1 * u.y + s * R.y, -B1 * u.z + s * R.z};
}

// ---------------------------------------------------------------------------
// permanent multipole pair (gl-function expansion of RealSpaceEnergyRegion.
// permanentPair, :1190-1460).  Linear in the radial set; the i-side force
// (dE/dx_i) and the torque on i are produced.  With a full neighbor list the
// torque on k comes from k's own row.
// ---------------------------------------------------------------------------
template <typename T> struct PermGeo {
  T sc3, sc4, sc5, sc6;
  T gl0, gl16, gl278, gl35, gl4;
  V3<T> qir, qkr, qiqkr, qkqir, qidk, qkdi;
  V3<T> dixdk, dixr, qixqk, rxqir, rxqikr, qkrxqir, dixqkr, dkxqir, rxqidk;
};

template <typename T>
__device__ __forceinline__ void perm_geometry(const V3<T> &r, const MP<T> &I, const MP<T> &K, PermGeo<T> &g) {
  g.qir = mul(I.q, r);
  g.qkr = mul(K.q, r);
  g.qiqkr = mul(I.q, g.qkr);
  g.qkqir = mul(K.q, g.qir);
  g.qidk = mul(I.q, K.d);
  g.qkdi = mul(K.q, I.d);
  g.dixdk = cross(I.d, K.d);
  g.dixr = cross(I.d, r);
  // qixqk_a = eps_abc Qi_be Qk_ce
  g.qixqk.x = I.q.xy * K.q.xz + I.q.yy * K.q.yz + I.q.yz * K.q.zz - I.q.xz * K.q.xy - I.q.yz * K.q.yy - I.q.zz * K.q.yz;
  g.qixqk.y = I.q.xz * K.q.xx + I.q.yz * K.q.xy + I.q.zz * K.q.xz - I.q.xx * K.q.xz - I.q.xy * K.q.yz - I.q.xz * K.q.zz;
  g.qixqk.z = I.q.xx * K.q.xy + I.q.xy * K.q.yy + I.q.xz * K.q.yz - I.q.xy * K.q.xx - I.q.yy * K.q.xy - I.q.yz * K.q.xz;
  g.rxqir = cross(r, g.qir);
  g.rxqikr = cross(r, g.qiqkr);
  g.qkrxqir = cross(g.qkr, g.qir);
  g.dixqkr = cross(I.d, g.qkr);
  g.dkxqir = cross(K.d, g.qir);
  g.rxqidk = cross(r, g.qidk);
  T sc2 = dot(I.d, K.d);
  g.sc3 = dot(I.d, r);
  g.sc4 = dot(K.d, r);
  g.sc5 = dot(g.qir, r);
  g.sc6 = dot(g.qkr, r);
  T sc7 = dot(g.qir, K.d);
  T sc8 = dot(g.qkr, I.d);
  T sc9 = dot(g.qir, g.qkr);
  T sc10 = I.q.xx * K.q.xx + I.q.yy * K.q.yy + I.q.zz * K.q.zz + (T)2 * (I.q.xy * K.q.xy + I.q.xz * K.q.xz + I.q.yz * K.q.yz);
  g.gl0 = I.c * K.c;
  T gl1 = K.c * g.sc3 - I.c * g.sc4;
  T gl2 = I.c * g.sc6 + K.c * g.sc5 - g.sc3 * g.sc4;
  T gl3 = g.sc3 * g.sc6 - g.sc4 * g.sc5;
  g.gl4 = g.sc5 * g.sc6;
  g.gl16 = gl1 + sc2;
  g.gl278 = gl2 + (T)2 * (sc7 - sc8) + (T)2 * sc10;
  g.gl35 = gl3 - (T)4 * sc9;
}

// energy; adds w * (dE/dx_i) to F and w * torque_i to Tq
template <typename T, bool GRAD>
__device__ __forceinline__ T perm_terms(const V3<T> &r, const MP<T> &I, const MP<T> &K, const PermGeo<T> &g,
                                        const T *B, T w, V3<T> &F, V3<T> &Tq) {
  T e = g.gl0 * B[0] + g.gl16 * B[1] + g.gl278 * B[2] + g.gl35 * B[3] + g.gl4 * B[4];
  if (GRAD) {
    T gf1 = g.gl0 * B[1] + g.gl16 * B[2] + g.gl278 * B[3] + g.gl35 * B[4] + g.gl4 * B[5];
    T gf2 = -K.c * B[1] + g.sc4 * B[2] - g.sc6 * B[3];
    T gf3 = I.c * B[1] + g.sc3 * B[2] + g.sc5 * B[3];
    T gf4 = (T)2 * B[2];
    T gf5 = (T)2 * (-K.c * B[2] + g.sc4 * B[3] - g.sc6 * B[4]);
    T gf6 = (T)2 * (-I.c * B[2] - g.sc3 * B[3] - g.sc5 * B[4]);
    T gf7 = (T)4 * B[3];
    V3<T> f = gf1 * r + gf2 * I.d + gf3 * K.d + gf4 * (g.qkdi - g.qidk) + gf5 * g.qir + gf6 * g.qkr +
              gf7 * (g.qiqkr + g.qkqir);
    V3<T> t = (-B[1]) * g.dixdk + gf2 * g.dixr + gf4 * (g.dixqkr + g.dkxqir + g.rxqidk - (T)2 * g.qixqk) -
              gf5 * g.rxqir - gf7 * (g.rxqikr + g.qkrxqir);
    F += w * f;
    Tq += w * t;
  }
  return e;
}

// ---------------------------------------------------------------------------
// polarization pair (RealSpaceEnergyRegion.polarizationPair, :1700-2177),
// i-side force and torque.  B: bn0..bn4, R: rr1..rr9 (index k -> rr_{2k+1}).
// ---------------------------------------------------------------------------
template <typename T, bool GRAD>
__device__ __forceinline__ T pol_terms(const V3<T> &r, const MP<T> &I, const MP<T> &K, const V3<T> &ui,
                                       const V3<T> &pi, const V3<T> &uk, const V3<T> &pk, const T *bn, const T *rr,
                                       const Thole<T> &th, T scalep, T scaled, bool direct, V3<T> &F,
                                       V3<T> &Tq) {
  T dsc3 = (T)1 - th.s3 * scaled, dsc5 = (T)1 - th.s5 * scaled, dsc7 = (T)1 - th.s7 * scaled;
  T psc3 = (T)1 - th.s3 * scalep, psc5 = (T)1 - th.s5 * scalep, psc7 = (T)1 - th.s7 * scalep;
  T usc3 = (T)1 - th.s3, usc5 = (T)1 - th.s5;
  V3<T> qir = mul(I.q, r), qkr = mul(K.q, r);
  T sc3 = dot(I.d, r), sc4 = dot(K.d, r), sc5 = dot(qir, r), sc6 = dot(qkr, r);
  T sci1 = dot(ui, K.d) + dot(I.d, uk);
  T sci3 = dot(ui, r), sci4 = dot(uk, r);
  T sci7 = dot(qir, uk), sci8 = dot(qkr, ui);
  T scip1 = dot(pi, K.d) + dot(I.d, pk);
  T scip2 = dot(ui, pk) + dot(pi, uk);
  T scip3 = dot(pi, r), scip4 = dot(pk, r);
  T scip7 = dot(qir, pk), scip8 = dot(qkr, pi);
  T gli1 = K.c * sci3 - I.c * sci4, gli2 = -sc3 * sci4 - sci3 * sc4, gli3 = sci3 * sc6 - sci4 * sc5;
  T gli6 = sci1, gli7 = (T)2 * (sci7 - sci8);
  T glip1 = K.c * scip3 - I.c * scip4, glip2 = -sc3 * scip4 - scip3 * sc4, glip3 = scip3 * sc6 - scip4 * sc5;
  T glip6 = scip1, glip7 = (T)2 * (scip7 - scip8);
  T ereal = (gli1 + gli6) * bn[1] + (gli2 + gli7) * bn[2] + gli3 * bn[3];
  T efix = (gli1 + gli6) * rr[1] * psc3 + (gli2 + gli7) * rr[2] * psc5 + gli3 * rr[3] * psc7;
  T e = (T)0.5 * (ereal - efix);
  if (!GRAD) return e;

  V3<T> qiuk = mul(I.q, uk), qkui = mul(K.q, ui), qiukp = mul(I.q, pk), qkuip = mul(K.q, pi);
  V3<T> dixuk = cross(I.d, uk), dixukp = cross(I.d, pk);
  V3<T> dixr = cross(I.d, r), rxqir = cross(r, qir);
  V3<T> ukxqir = cross(uk, qir), ukxqirp = cross(pk, qir), rxqiuk = cross(r, qiuk), rxqiukp = cross(r, qiukp);

  T gfi1 = (T)0.5 * bn[2] * (gli1 + glip1 + gli6 + glip6 + scip2) +
           (T)0.5 * bn[3] * (gli2 + glip2 + gli7 + glip7 - sci3 * scip4 - scip3 * sci4) + (T)0.5 * bn[4] * (gli3 + glip3);
  T gfi2 = -K.c * bn[1] + sc4 * bn[2] - sc6 * bn[3];
  T gfi3 = I.c * bn[1] + sc3 * bn[2] + sc5 * bn[3];
  T gfi4 = (T)2 * bn[2];
  T gfi5 = bn[3] * (sci4 + scip4);
  T gfi6 = -bn[3] * (sci3 + scip3);
  T gti2 = (T)0.5 * bn[2] * (sci4 + scip4);
  V3<T> f = gfi1 * r +
            (T)0.5 * (gfi2 * (ui + pi) + bn[2] * (sci4 * pi + scip4 * ui) + gfi3 * (uk + pk) +
                      bn[2] * (sci3 * pk + scip3 * uk) + ((sci4 + scip4) * bn[2]) * I.d +
                      ((sci3 + scip3) * bn[2]) * K.d + gfi4 * (qkui + qkuip - qiuk - qiukp)) +
            gfi5 * qir + gfi6 * qkr;
  V3<T> t2 = ((T)-0.5 * bn[1]) * (dixuk + dixukp) + gti2 * dixr - gfi5 * rxqir +
             ((T)0.5 * gfi4) * (ukxqir + rxqiuk + ukxqirp + rxqiukp);
  if (psc3 != (T)0 || dsc3 != (T)0 || usc3 != (T)0) {
    T gfr1 = (T)0.5 * rr[2] * ((gli1 + gli6) * psc3 + (glip1 + glip6) * dsc3 + scip2 * usc3) +
             (T)0.5 * rr[3] * ((gli7 + gli2) * psc5 + (glip7 + glip2) * dsc5 - (sci3 * scip4 + scip3 * sci4) * usc5) +
             (T)0.5 * rr[4] * (gli3 * psc7 + glip3 * dsc7);
    T gfr4 = (T)2 * rr[2];
    T gfr5 = rr[3] * (sci4 * psc7 + scip4 * dsc7);
    T gfr6 = -rr[3] * (sci3 * psc7 + scip3 * dsc7);
    T gtr2 = (T)0.5 * rr[2] * (sci4 * psc5 + scip4 * dsc5);
    V3<T> fr = gfr1 * r +
               (T)0.5 * ((-rr[1] * K.c) * (psc3 * ui + dsc3 * pi) + (rr[2] * sc4) * (psc5 * ui + dsc5 * pi) -
                         (rr[3] * sc6) * (psc7 * ui + dsc7 * pi)) +
               (T)0.5 * ((rr[1] * I.c) * (psc3 * uk + dsc3 * pk) + (rr[2] * sc3) * (psc5 * uk + dsc5 * pk) +
                         (rr[3] * sc5) * (psc7 * uk + dsc7 * pk)) +
               ((T)0.5 * rr[2] * usc5) * (sci4 * pi + scip4 * ui + sci3 * pk + scip3 * uk) +
               ((T)0.5 * (sci4 * psc5 + scip4 * dsc5) * rr[2]) * I.d +
               ((T)0.5 * (sci3 * psc5 + scip3 * dsc5) * rr[2]) * K.d +
               ((T)0.5 * gfr4) * (psc5 * (qkui - qiuk) + dsc5 * (qkuip - qiukp)) + gfr5 * qir + gfr6 * qkr;
    V3<T> tr2 = ((T)-0.5 * rr[1]) * (psc3 * dixuk + dsc3 * dixukp) + gtr2 * dixr - gfr5 * rxqir +
                ((T)0.5 * gfr4) * (psc5 * (ukxqir + rxqiuk) + dsc5 * (ukxqirp + rxqiukp));
    f -= fr;
    t2 -= tr2;
  }
  T a3 = (T)0.5 * rr[1] * ((gli1 + gli6) * scalep + (glip1 + glip6) * scaled);
  T a5 = (T)0.5 * rr[2] * ((gli2 + gli7) * scalep + (glip2 + glip7) * scaled);
  T a7 = (T)0.5 * rr[3] * (gli3 * scalep + glip3 * scaled);
  T b3 = (T)0.5 * rr[1] * scip2, b5 = (T)-0.5 * rr[2] * (sci3 * scip4 + scip3 * sci4);
  V3<T> findmp = b3 * th.d3 + b5 * th.d5;
  f -= a3 * th.d3 + a5 * th.d5 + a7 * th.d7 + findmp;
  if (direct) {
    T gfd = (T)0.5 * (bn[2] * scip2 - bn[3] * (scip3 * sci4 + sci3 * scip4));
    T gfdr = (T)0.5 * (rr[2] * scip2 * usc3 - rr[3] * (scip3 * sci4 + sci3 * scip4) * usc5);
    V3<T> v = sci4 * pi + scip4 * ui + sci3 * pk + scip3 * uk;
    f -= gfd * r + ((T)0.5 * bn[2]) * v;
    f += gfdr * r + ((T)0.5 * usc5 * rr[2]) * v + findmp;
  }
  F += f;
  Tq += t2;
  return e;
}

// ---------------------------------------------------------------------------
// B-spline weights and derivatives 0..3 (UniformBSpline.bSplineDerivatives),
// cardinal recursion on shifted weights with finite-difference derivatives.
// ---------------------------------------------------------------------------
template <int ORDER>
__device__ __forceinline__ void bspline4(double x, double th[ORDER][4]) {
  double w[ORDER + 2][ORDER + 2];
#pragma unroll
  for (int p = 0; p < ORDER + 2; p++)
#pragma unroll
    for (int k = 0; k < ORDER + 2; k++) w[p][k] = 0.0;
  w[1][1] = 1.0;
#pragma unroll
  for (int p = 2; p <= ORDER; p++) {
    double inv = 1.0 / (p - 1);
#pragma unroll
    for (int k = 0; k < p; k++) w[p][k + 1] = ((x + p - 1 - k) * w[p - 1][k] + (k + 1 - x) * w[p - 1][k + 1]) * inv;
  }
#pragma unroll
  for (int k = 0; k < ORDER; k++) {
    th[k][0] = w[ORDER][k + 1];
    th[k][1] = w[ORDER - 1][k] - w[ORDER - 1][k + 1];
    double a = (k - 1 >= 0) ? w[ORDER - 2][k - 1] : 0.0;
    th[k][2] = a - 2.0 * w[ORDER - 2][k] + w[ORDER - 2][k + 1];
    double b = (k - 2 >= 0) ? w[ORDER - 3][k - 2] : 0.0;
    double c = (k - 1 >= 0) ? w[ORDER - 3][k - 1] : 0.0;
    th[k][3] = b - 3.0 * c + 3.0 * w[ORDER - 3][k] - w[ORDER - 3][k + 1];
  }
}
