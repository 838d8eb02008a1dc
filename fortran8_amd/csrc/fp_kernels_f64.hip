// Plain FP64 path of the FENE-P / IBM SMAC step (the parity path).
// Compiled with -ffp-contract=off: every expression is evaluated in the
// order written, which is the order of the reference statement it restates.
#include "fp_device.h"
#include <algorithm>
#include <map>
#include <cmath>

namespace fp {

static inline int nblk(long n, int t) { return (int)((n + t - 1) / t); }

// ============================================================ FENE-P ======
// CpxCnx/CpyCny/CpzCnz selection (fenep.f90:169-382) for all three axes of
// one cell; the chosen candidate per axis is packed in 2 bits.
template <bool WL>
__global__ __launch_bounds__(256) void k_select(Grid g, const double* __restrict__ C, uint8_t* __restrict__ sel) {
  const long L = xcd_cell();
  if (L >= g.n) return;
  long m;
  int i, j, k;
  band_cell(g, L, m, i, j, k);
  uint8_t s = 0;
#pragma unroll 1
  for (int d = 0; d < 3; d++) {
    double cm[6], c[6], cp[6];
    int di = d == 0, dj = d == 1, dk = d == 2;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const double* Cq = C + q * g.n;
      cm[q] = ldbw<WL>(Cq, g, i - di, j - dj, k - dk, BC_C);
      c[q] = Cq[m];
      cp[q] = ldbw<WL>(Cq, g, i + di, j + dj, k + dk, BC_C);
    }
    s |= (uint8_t)(select_d(cm, c, cp) << (2 * d));
  }
  sel[m] = s;
}

// face value of cell a (wrapped) along axis d with its own selection:
// left (Cp) when left=true, right (Cn) otherwise, component q.
template <bool WL>
__device__ __forceinline__ void face_vals(const Grid& g, const double* __restrict__ C, const uint8_t* __restrict__ sel,
                                          int d, int i, int j, int k, bool left, double* out) {
  i = wrapi(i, g.nx); j = wrapi(j, g.ny); k = wrapi(k, g.nz);
  long m = cidx(g, i, j, k);
  int l = (sel[m] >> (2 * d)) & 3;
  int di = d == 0, dj = d == 1, dk = d == 2;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const double* Cq = C + q * g.n;
    double c = Cq[m];
    double cm = l == 3 ? 0.0 : ldbw<WL>(Cq, g, i - di, j - dj, k - dk, BC_C);
    double cp = l == 3 ? 0.0 : ldbw<WL>(Cq, g, i + di, j + dj, k + dk, BC_C);
    out[q] = left ? cand_p(l, cm, c, cp) : cand_n(l, cm, c, cp);
  }
}

// Cstar (fenep.f90:464-507) + Lyapunov Newton (fenep.f90:518-585) + the
// Peterlin factor of the new C (f, fenep.f90:510-515, used by polymer_stress).
template <bool WL>
__global__ __launch_bounds__(128) void k_cstar_newton(Grid g, const double* __restrict__ C, const uint8_t* __restrict__ sel,
                                                      const double* __restrict__ U, const double* __restrict__ V,
                                                      const double* __restrict__ W, double* __restrict__ Cn,
                                                      double* __restrict__ fC, int* __restrict__ nitr,
                                                      double dc, double eps, int itrmax) {
  const long L = xcd_cell();
  if (L >= g.n) return;
  long m;
  int i, j, k;
  band_cell(g, L, m, i, j, k);
  const double dt = g.dt;
  double cx[6];
#pragma unroll 1
  for (int d = 0; d < 3; d++) {
    double h = d == 0 ? g.dx : (d == 1 ? g.dy : g.dz);
    double v1, v2;
    if (d == 0) { v1 = ldbw<WL>(U, g, i - 1, j, k, BC_U); v2 = U[m]; }
    else if (d == 1) { v1 = ldbw<WL>(V, g, i, j - 1, k, BC_V); v2 = ldbw<WL>(V, g, i, j, k, BC_V); }
    else { v1 = ldbw<WL>(W, g, i, j, k - 1, BC_W); v2 = W[m]; }
    double s1 = ((-v2 + fabs(v2)) * dt) / (2.0 * h);
    double s2 = 1.0 / 6.0 + ((-v2 - fabs(v2)) * dt) / (2.0 * h);
    double s3 = 1.0 / 6.0 + ((v1 - fabs(v1)) * dt) / (2.0 * h);
    double s4 = ((v1 + fabs(v1)) * dt) / (2.0 * h);
    int di = d == 0, dj = d == 1, dk = d == 2;
    double pl[6], pr[6], nl[6], nr[6];
    face_vals<WL>(g, C, sel, d, i, j, k, true, pl);                    // Cp(a-1): own left face
    face_vals<WL>(g, C, sel, d, i + di, j + dj, k + dk, true, pr);     // Cp(a): left face of a+1
    face_vals<WL>(g, C, sel, d, i - di, j - dj, k - dk, false, nl);    // Cn(a-1): right face of a-1
    face_vals<WL>(g, C, sel, d, i, j, k, false, nr);                   // Cn(a): own right face
#pragma unroll
    for (int q = 0; q < 6; q++) {
      double acc = d == 0 ? s3 * pl[q] : cx[q] + s3 * pl[q];
      acc = acc + s1 * pr[q];
      acc = acc + s4 * nl[q];
      acc = acc + s2 * nr[q];
      cx[q] = acc;
    }
  }
  // Lyapunov
  double b[6];
#pragma unroll
  for (int q = 0; q < 6; q++) b[q] = cx[q];
  b[0] = b[0] + dt / g.wi;
  b[3] = b[3] + dt / g.wi;
  b[5] = b[5] + dt / g.wi;
  const double dx = g.dx, dy = g.dy, dz = g.dz;
#define UL(a, bb, c) ldbw<WL>(U, g, a, bb, c, BC_U)
#define VL(a, bb, c) ldbw<WL>(V, g, a, bb, c, BC_V)
#define WL(a, bb, c) ldbw<WL>(W, g, a, bb, c, BC_W)
  double a11 = -((dt * (-UL(i - 1, j, k) + UL(i, j, k))) / dx);
  double a12 = -((dt * (((-VL(i - 1, j - 1, k) + VL(i + 1, j - 1, k)) - VL(i - 1, j, k)) + VL(i + 1, j, k))) / (4.0 * dx));
  double a13 = -((dt * (((-WL(i - 1, j, k - 1) + WL(i + 1, j, k - 1)) - WL(i - 1, j, k)) + WL(i + 1, j, k))) / (4.0 * dx));
  double a21 = -((dt * (((-UL(i - 1, j - 1, k) - UL(i, j - 1, k)) + UL(i - 1, j + 1, k)) + UL(i, j + 1, k))) / (4.0 * dy));
  double a22 = -((dt * (-VL(i, j - 1, k) + VL(i, j, k))) / dy);
  double a23 = -((dt * (((-WL(i, j - 1, k - 1) + WL(i, j + 1, k - 1)) - WL(i, j - 1, k)) + WL(i, j + 1, k))) / (4.0 * dy));
  double a31 = -((dt * (((-UL(i - 1, j, k - 1) - UL(i, j, k - 1)) + UL(i - 1, j, k + 1)) + UL(i, j, k + 1))) / (4.0 * dz));
  double a32 = -((dt * (((-VL(i, j - 1, k - 1) - VL(i, j, k - 1)) + VL(i, j - 1, k + 1)) + VL(i, j, k + 1))) / (4.0 * dz));
  double a33 = -((dt * (-WL(i, j, k - 1) + WL(i, j, k))) / dz);
#undef UL
#undef VL
#undef WL
  double A[9] = {a11, a12, a13, a21, a22, a23, a31, a32, a33};  // transpose(A0), column-major
  double x[6];
#pragma unroll
  for (int q = 0; q < 6; q++) x[q] = C[q * g.n + m];
  const int it = newton_d(A, b, x, dt, g.wi, g.lp, dc, eps, itrmax);
#pragma unroll
  for (int q = 0; q < 6; q++) Cn[q * g.n + m] = x[q];
  fC[m] = peterlin_d(g.lp, x[0], x[3], x[5]);
  nitr[m] = it > itrmax ? -itrmax : it;   // negative: stopped at the cap
}

// ========================================================= predictor ======
// C0 = f(C) C - E at (i,j,k) with ghosts (polymer_stress, fenep.f90:655-659)
template <bool WL>
__device__ __forceinline__ double c0_at(const Grid& g, const double* __restrict__ C, const double* __restrict__ fC,
                                        int q, int i, int j, int k) {
  const double E = (q == 0 || q == 3 || q == 5) ? 1.0 : 0.0;
  if (g.walls && (j < 0 || j >= g.ny)) {
    double c0 = ldbw<WL>(C, g, i, j, k, BC_C), c3 = ldbw<WL>(C + 3 * g.n, g, i, j, k, BC_C), c5 = ldbw<WL>(C + 5 * g.n, g, i, j, k, BC_C);
    double f = peterlin_d(g.lp, c0, c3, c5);
    return f * ldbw<WL>(C + q * g.n, g, i, j, k, BC_C) - E;
  }
  long m = cidx(g, wrapi(i, g.nx), wrapi(j, g.ny), wrapi(k, g.nz));
  return fC[m] * C[q * g.n + m] - E;
}

// convection (fenep.f90:694-734), viscous (736-770), polymer_stress
// (645-691) and the predictor right-hand side of navier (803-826),
// fft_navier (1466-1511) or ibm_preliminary (1994-2013), with the history
// updates A0=A, T0=T, B0=B done in place.
template <bool WL>
__global__ __launch_bounds__(256) void k_predictor(Grid g, int method, int first, int polymer,
                                                   const double* __restrict__ U, const double* __restrict__ V,
                                                   const double* __restrict__ W, const double* __restrict__ P,
                                                   const double* __restrict__ C, const double* __restrict__ fC,
                                                   double* __restrict__ A0, double* __restrict__ T0, double* __restrict__ B0,
                                                   double* __restrict__ Bout, const double* __restrict__ F,
                                                   double bf0, double bf1, double bf2, double* __restrict__ Q) {
  const long L = xcd_cell();
  if (L >= g.n) return;
  long m;
  int i, j, k;
  band_cell(g, L, m, i, j, k);
  const double dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
#define Ug(a, b, c) ldbw<WL>(U, g, a, b, c, BC_U)
#define Vg(a, b, c) ldbw<WL>(V, g, a, b, c, BC_V)
#define Wg(a, b, c) ldbw<WL>(W, g, a, b, c, BC_W)
  double a[3], bv[3], t[3];
  {
    double h1 = (Ug(i - 1, j, k) + Ug(i, j, k)) / 2.0, h2 = (Ug(i, j, k) + Ug(i + 1, j, k)) / 2.0;
    double ax = ((-(h1 * h1) + h2 * h2) / dx
                 + (-((((Vg(i, j - 1, k) + Vg(i + 1, j - 1, k)) / 2.0) * (Ug(i, j - 1, k) + Ug(i, j, k))) / 2.0)
                    + (((Vg(i, j, k) + Vg(i + 1, j, k)) / 2.0) * (Ug(i, j, k) + Ug(i, j + 1, k))) / 2.0) / dy)
                + (-((((Wg(i, j, k - 1) + Wg(i + 1, j, k - 1)) / 2.0) * (Ug(i, j, k - 1) + Ug(i, j, k))) / 2.0)
                   + (((Wg(i, j, k) + Wg(i + 1, j, k)) / 2.0) * (Ug(i, j, k) + Ug(i, j, k + 1))) / 2.0) / dz;
    double g1 = (Vg(i, j - 1, k) + Vg(i, j, k)) / 2.0, g2 = (Vg(i, j, k) + Vg(i, j + 1, k)) / 2.0;
    double ay = ((-((((Ug(i - 1, j, k) + Ug(i - 1, j + 1, k)) / 2.0) * (Vg(i - 1, j, k) + Vg(i, j, k))) / 2.0)
                  + (((Ug(i, j, k) + Ug(i, j + 1, k)) / 2.0) * (Vg(i, j, k) + Vg(i + 1, j, k))) / 2.0) / dx
                 + (-(g1 * g1) + g2 * g2) / dy)
                + (-((((Wg(i, j, k - 1) + Wg(i, j + 1, k - 1)) / 2.0) * (Vg(i, j, k - 1) + Vg(i, j, k))) / 2.0)
                   + (((Wg(i, j, k) + Wg(i, j + 1, k)) / 2.0) * (Vg(i, j, k) + Vg(i, j, k + 1))) / 2.0) / dz;
    double e1 = (Wg(i, j, k - 1) + Wg(i, j, k)) / 2.0, e2 = (Wg(i, j, k) + Wg(i, j, k + 1)) / 2.0;
    double az = ((-((((Ug(i - 1, j, k) + Ug(i - 1, j, k + 1)) / 2.0) * (Wg(i - 1, j, k) + Wg(i, j, k))) / 2.0)
                  + (((Ug(i, j, k) + Ug(i, j, k + 1)) / 2.0) * (Wg(i, j, k) + Wg(i + 1, j, k))) / 2.0) / dx
                 + (-((((Vg(i, j - 1, k) + Vg(i, j - 1, k + 1)) / 2.0) * (Wg(i, j - 1, k) + Wg(i, j, k))) / 2.0)
                    + (((Vg(i, j, k) + Vg(i, j, k + 1)) / 2.0) * (Wg(i, j, k) + Wg(i, j + 1, k))) / 2.0) / dy)
                + (-(e1 * e1) + e2 * e2) / dz;
    a[0] = -ax; a[1] = -ay; a[2] = -az;
  }
  const double cb = g.beta / g.re;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double* X = c == 0 ? U : (c == 1 ? V : W);
    int kind = c == 0 ? BC_U : (c == 1 ? BC_V : BC_W);
    double x0 = ldbw<WL>(X, g, i, j, k, kind);
    double bb = (((ldbw<WL>(X, g, i - 1, j, k, kind) - 2.0 * x0) + ldbw<WL>(X, g, i + 1, j, k, kind)) / (dx * dx)
                 + ((ldbw<WL>(X, g, i, j - 1, k, kind) - 2.0 * x0) + ldbw<WL>(X, g, i, j + 1, k, kind)) / (dy * dy))
                + ((ldbw<WL>(X, g, i, j, k - 1, kind) - 2.0 * x0) + ldbw<WL>(X, g, i, j, k + 1, kind)) / (dz * dz);
    bv[c] = cb * bb;
  }
  if (polymer) {
#define C0g(q, aa, bb, cc) c0_at<WL>(g, C, fC, q, aa, bb, cc)
    double tx = ((-C0g(0, i, j, k) + C0g(0, i + 1, j, k)) / dx
                 + (((-C0g(1, i, j - 1, k) - C0g(1, i + 1, j - 1, k)) + C0g(1, i, j + 1, k)) + C0g(1, i + 1, j + 1, k)) / (4.0 * dy))
                + (((-C0g(2, i, j, k - 1) - C0g(2, i + 1, j, k - 1)) + C0g(2, i, j, k + 1)) + C0g(2, i + 1, j, k + 1)) / (4.0 * dz);
    double ty = ((((-C0g(1, i - 1, j, k) + C0g(1, i + 1, j, k)) - C0g(1, i - 1, j + 1, k)) + C0g(1, i + 1, j + 1, k)) / (4.0 * dx)
                 + (-C0g(3, i, j, k) + C0g(3, i, j + 1, k)) / dy)
                + (((-C0g(4, i, j, k - 1) - C0g(4, i, j + 1, k - 1)) + C0g(4, i, j, k + 1)) + C0g(4, i, j + 1, k + 1)) / (4.0 * dz);
    double tz = ((((-C0g(2, i - 1, j, k) + C0g(2, i + 1, j, k)) - C0g(2, i - 1, j, k + 1)) + C0g(2, i + 1, j, k + 1)) / (4.0 * dx)
                 + (((-C0g(4, i, j - 1, k) + C0g(4, i, j + 1, k)) - C0g(4, i, j - 1, k + 1)) + C0g(4, i, j + 1, k + 1)) / (4.0 * dy))
                + (-C0g(5, i, j, k) + C0g(5, i, j, k + 1)) / dz;
#undef C0g
    const double coef = (1.0 - g.beta) / (g.re * g.wi);
    t[0] = coef * tx; t[1] = coef * ty; t[2] = coef * tz;
  } else {
    t[0] = t[1] = t[2] = 0.0;
  }
  const double bfs[3] = {bf0, bf1, bf2};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double* X = c == 0 ? U : (c == 1 ? V : W);
    int kind = c == 0 ? BC_U : (c == 1 ? BC_V : BC_W);
    long o = c * g.n + m;
    double a0 = first ? a[c] : A0[o];
    double t0 = first ? t[c] : T0[o];
    double b0 = first ? bv[c] : B0[o];
    double u = ldbw<WL>(X, g, i, j, k, kind);
    double p1 = c == 0 ? ldbw<WL>(P, g, i + 1, j, k, BC_P) : (c == 1 ? ldbw<WL>(P, g, i, j + 1, k, BC_P) : ldbw<WL>(P, g, i, j, k + 1, BC_P));
    double h = c == 0 ? dx : (c == 1 ? dy : dz);
    double dp = (dt * (-P[m] + p1)) / h;
    double fv = (F ? F[o] : 0.0) + bfs[c];
    double r;
    if (method == FP_METHOD_EXPLICIT) {
      r = (((u - dp) + (dt * (3.0 * (a[c] + bv[c]) - (a0 + b0))) / 2.0) + (dt * (t[c] + t0)) / 2.0) + dt * fv;
    } else if (method == FP_METHOD_FFT) {
      r = ((((u - dp) + (dt * (3.0 * a[c] - a0)) / 2.0) + (dt * bv[c]) / 2.0) + (dt * (t[c] + t0)) / 2.0) + dt * fv;
    } else {
      r = ((((u - dp) + (dt * (3.0 * a[c] - a0)) / 2.0) + dt * bv[c]) + (dt * (t[c] + t0)) / 2.0) + dt * fv;
    }
    Q[o] = r;
    A0[o] = a[c];
    T0[o] = t[c];
    B0[o] = bv[c];
    if (Bout) Bout[o] = bv[c];
  }
#undef Ug
#undef Vg
#undef Wg
}

// ============================================================ FFT ==========
// Stockham radix-2 autosort (forward e^{-2 pi i jk/n}, unnormalised; the
// inverse uses the conjugate twiddles) over `nl` lines resident in LDS.
// x and y are LDS buffers of nl*n complex; line l occupies [l*n, (l+1)*n).
// Returns the buffer holding the result.
__device__ double2* stockham_lds(double2* x, double2* y, int n, int logn, int nl, const double2* __restrict__ tw, bool inv) {
  int len = n, s = 1, lgs = 0;
  const int half = n >> 1;
  while (len > 1) {
    int m = len >> 1;
    for (int b = threadIdx.x; b < nl * half; b += blockDim.x) {
      int line = b >> (logn - 1);
      int t = b & (half - 1);
      int p = t >> lgs, q = t & (s - 1);
      double2 w = tw[p * s];
      double wr = w.x, wi = inv ? -w.y : w.y;
      const double2* xl = x + line * n;
      double2* yl = y + line * n;
      double2 a = xl[q + s * p], bb = xl[q + s * (p + m)];
      double2 y0, y1;
      y0.x = a.x + bb.x;
      y0.y = a.y + bb.y;
      double dr = a.x - bb.x, di = a.y - bb.y;
      y1.x = dr * wr - di * wi;
      y1.y = dr * wi + di * wr;
      yl[q + s * (2 * p)] = y0;
      yl[q + s * (2 * p + 1)] = y1;
    }
    __syncthreads();
    double2* tt = x; x = y; y = tt;
    len = m; s <<= 1; lgs++;
  }
  return x;
}

__device__ __forceinline__ int ilog2(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

// real x-lines -> half spectrum rows (fftr2c along x)
__global__ __launch_bounds__(256) void k_fft_x_fwd(Grid g, const double* __restrict__ q, double2* __restrict__ spec,
                                                   const double2* __restrict__ tw, int lines_per_block) {
  extern __shared__ double2 lds[];
  const int n = g.nx, nl = lines_per_block;
  double2* x = lds;
  double2* y = lds + nl * n;
  long nlines = (long)g.ny * g.nz;
  long l0 = (long)blockIdx.x * nl;
  for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
    int l = e / n, i = e % n;
    long L = l0 + l;
    double v = L < nlines ? q[L * n + i] : 0.0;
    x[e] = make_double2(v, 0.0);
  }
  __syncthreads();
  double2* r = stockham_lds(x, y, n, ilog2(n), nl, tw, false);
  const int nxh = g.nxh;
  for (int e = threadIdx.x; e < nl * nxh; e += blockDim.x) {
    int l = e / nxh, i = e % nxh;
    long L = l0 + l;
    if (L >= nlines) continue;
    long j = L % g.ny, k = L / g.ny;
    spec[i + (long)nxh * ((j + g.jo) + g.sy * k)] = r[l * n + i];
  }
}

// half spectrum rows -> real x-lines (fftc2r along x)
__global__ __launch_bounds__(256) void k_fft_x_inv(Grid g, const double2* __restrict__ spec, double* __restrict__ out,
                                                   const double2* __restrict__ tw, int lines_per_block) {
  extern __shared__ double2 lds[];
  const int n = g.nx, nl = lines_per_block, h = n / 2, nxh = g.nxh;
  double2* x = lds;
  double2* y = lds + nl * n;
  long nlines = (long)g.ny * g.nz;
  long l0 = (long)blockIdx.x * nl;
  for (int e = threadIdx.x; e < nl * nxh; e += blockDim.x) {
    int l = e / nxh, i = e % nxh;
    long L = l0 + l;
    double2 v = make_double2(0.0, 0.0);
    if (L < nlines) {
      long j = L % g.ny, k = L / g.ny;
      v = spec[i + (long)nxh * ((j + g.jo) + g.sy * k)];
    }
    if (i == 0 || i == h) v.y = 0.0;
    x[l * n + i] = v;
    if (i > 0 && i < h) x[l * n + (n - i)] = make_double2(v.x, -v.y);
  }
  __syncthreads();
  double2* r = stockham_lds(x, y, n, ilog2(n), nl, tw, true);
  for (int e = threadIdx.x; e < nl * n; e += blockDim.x) {
    int l = e / n, i = e % n;
    long L = l0 + l;
    if (L < nlines) out[L * n + i] = r[e].x;
  }
}

// complex lines along y (axis 1) or z (axis 2) of the spectrum.  mode 0:
// forward, 1: inverse, 2: forward, divide by the periodic LHS of `kind`, then
// inverse (fft_solve's divide fused between the z transforms).
__global__ __launch_bounds__(256) void k_fft_yz(Grid g, double2* __restrict__ spec, const double2* __restrict__ tw,
                                                int axis, int mode, int kind, int row0, int nrows, int ti,
                                                const double* __restrict__ kx2, const double* __restrict__ ky2,
                                                const double* __restrict__ kz2) {
  extern __shared__ double2 lds[];
  const int n = axis == 1 ? g.ny : g.nz;
  const int nxh = g.nxh;
  const long stride = axis == 1 ? (long)nxh : (long)nxh * g.sy;
  const int ntile = (nxh + ti - 1) / ti;
  int tile = blockIdx.x % ntile;
  long other = blockIdx.x / ntile;  // k for axis 1, row for axis 2
  int i0 = tile * ti;
  double2* base;
  if (axis == 1) base = spec + (long)nxh * (g.jo + g.sy * other);
  else base = spec + (long)nxh * (row0 + other);
  double2* x = lds;
  double2* y = lds + ti * n;
  // LDS layout: line l (i = i0+l) contiguous
  for (int e = threadIdx.x; e < ti * n; e += blockDim.x) {
    int l = e % ti, r = e / ti;
    int i = i0 + l;
    x[l * n + r] = i < nxh ? base[i + stride * r] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  int logn = ilog2(n);
  double2* res = stockham_lds(x, y, n, logn, ti, tw, mode == 1);
  if (mode == 2) {
    double2* other_buf = res == x ? y : x;
    const double ntot = (double)((long)g.nx * g.ny * g.nz);
    const double coef = ((g.dt * g.beta) / g.re) / 2.0;
    for (int e = threadIdx.x; e < ti * n; e += blockDim.x) {
      int l = e / n, kk = e % n;
      int i = i0 + l;
      if (i >= nxh) continue;
      long jj = other;  // axis 2: row index == j for periodic grids
      double lhs;
      if (kind == 3) lhs = -((kx2[i] + ky2[jj]) + kz2[kk]);
      else lhs = 1.0 + coef * ((kx2[i] + ky2[jj]) + kz2[kk]);
      double2 s = res[e];
      double re = s.x / ntot, im = s.y / ntot;
      if (fabs(lhs) < 1.0e-16) { s.x = 0.0; s.y = 0.0; }
      else { s.x = re / lhs; s.y = im / lhs; }
      res[e] = s;
    }
    __syncthreads();
    res = stockham_lds(res, other_buf, n, logn, ti, tw, true);
  }
  for (int e = threadIdx.x; e < ti * n; e += blockDim.x) {
    int l = e % ti, r = e / ti;
    int i = i0 + l;
    if (i < nxh) base[i + stride * r] = res[l * n + r];
  }
}

// tridiagonal line solve along y with the wall rows of fenep_2d.f90:1222-1331
// (tdma, fenep_2d.f90:1145-1177), one thread per (kx, kz) line.
// Tridiagonal line solve along y of each (kx, kz) mode (tdma with the wall
// rows of fft_*_kabe, fenep_2d.f90:1145-1177, 1222-1331).  One block of 256
// threads solves `tl` consecutive kx lines of one kz plane: the rows are
// staged through LDS with coalesced loads, one thread per line runs the
// recurrence from LDS, and the rows are stored back.  The arithmetic is the
// Thomas recurrence of the reference, in its order.
__global__ __launch_bounds__(256) void k_thomas(Grid g, double2* __restrict__ spec, int kind, int tl,
                                                const double* __restrict__ kx2, const double* __restrict__ kz2) {
  extern __shared__ double2 tq[];                 // [size][tl] right-hand sides / Q
  const int nxh = g.nxh, ny = g.ny, size = ny + 2;
  double* tp = (double*)(tq + (long)size * tl);   // [size][tl] P
  const int ngrp = (nxh + tl - 1) / tl;
  const int i0 = (int)(blockIdx.x % ngrp) * tl;
  const long k = blockIdx.x / ngrp;
  const int nl = min(tl, nxh - i0);
  double2* base = spec + (long)nxh * g.sy * k + i0;  // row r of line l at base[nxh*r + l]
  for (int e = threadIdx.x; e < size * tl; e += blockDim.x) {
    const int r = e / tl, l = e - r * tl;
    if (l < nl) tq[e] = base[(long)nxh * r + l];
  }
  __syncthreads();
  if (threadIdx.x < nl) {
    const int l = threadIdx.x, i = i0 + l;
    const double nxz = (double)((long)g.nx * g.nz);
    const double coef = ((g.dt * g.beta) / g.re) / 2.0;
    const double idy2 = g.dy * g.dy;
    const double kk = kx2[i] + kz2[k];
    auto coeffs = [&](int r, double& a, double& b, double& c, bool& rhs) {
      rhs = false;
      if (kind == 3) {
        if (r == 0) { a = 1.0; b = -1.0; c = 0.0; }
        else if (r == size - 1) { a = 1.0; b = 0.0; c = -1.0; }
        else { a = -(kk + 2.0 / idy2); b = 1.0 / idy2; c = 1.0 / idy2; rhs = true; }
      } else if (kind == 1) {
        if (r == 0) { a = 1.0; b = 0.0; c = 0.0; }
        else if (r >= ny) { a = 1.0; b = 0.0; c = 0.0; }
        else { a = 1.0 + coef * (kk + 2.0 / idy2); b = -(coef / idy2); c = -(coef / idy2); rhs = true; }
      } else {
        if (r == 0) { a = 1.0; b = 1.0; c = 0.0; }
        else if (r == size - 1) { a = 1.0; b = 0.0; c = 1.0; }
        else { a = 1.0 + coef * (kk + 2.0 / idy2); b = -(coef / idy2); c = -(coef / idy2); rhs = true; }
      }
    };
    double a, b, c; bool rhs;
    coeffs(0, a, b, c, rhs);
    double Pp = -b / a;
    double Qr = 0.0 / a, Qi = 0.0 / a;
    tp[l] = Pp;
    tq[l] = make_double2(Qr, Qi);
    // the next row's right-hand side is read before the current row's
    // results are stored (the recurrence then waits on arithmetic only)
    double2 nxt = tq[tl + l];
    for (int r = 1; r < size; r++) {
      coeffs(r, a, b, c, rhs);
      const double2 sv = nxt;
      if (r + 1 < size) nxt = tq[(r + 1) * tl + l];
      double dr = 0.0, di = 0.0;
      if (rhs) { dr = sv.x / nxz; di = sv.y / nxz; }
      const double den = a + c * Pp;
      double P, nQr, nQi;
      if (fabs(den) < 1.0e-16) { P = 0.0; nQr = 0.0; nQi = 0.0; }
      else { P = -b / den; nQr = (dr - c * Qr) / den; nQi = (di - c * Qi) / den; }
      Pp = P; Qr = nQr; Qi = nQi;
      tp[r * tl + l] = P;
      tq[r * tl + l] = make_double2(Qr, Qi);
    }
    double xr = Qr, xi = Qi;
    double Pn = tp[(size - 2) * tl + l];
    double2 qn = tq[(size - 2) * tl + l];
    for (int r = size - 2; r >= 0; r--) {
      const double P = Pn;
      const double2 q = qn;
      if (r > 0) { Pn = tp[(r - 1) * tl + l]; qn = tq[(r - 1) * tl + l]; }
      xr = P * xr + q.x;
      xi = P * xi + q.y;
      tq[r * tl + l] = make_double2(xr, xi);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < size * tl; e += blockDim.x) {
    const int r = e / tl, l = e - r * tl;
    if (l < nl) base[(long)nxh * r + l] = tq[e];
  }
}

// lines per block of the LDS line solve: 64 KiB of LDS at most
int thomas_lines(int ny, int bytes_per_row) {
  const int per_line = (ny + 2) * bytes_per_row;
  return std::max(1, std::min(32, 65536 / per_line));
}

static int lines_for(int n) { return std::max(1, std::min(16, 2048 / n)); }

void f64_solve(Ctx& c, int kind, const double* q, double* x) {
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  int lx = lines_for(g.nx);
  long xlines = (long)g.ny * g.nz;
  size_t shx = sizeof(double2) * 2 * lx * g.nx;
  hipLaunchKernelGGL(k_fft_x_fwd, dim3(nblk(xlines, lx)), dim3(256), shx, s, g, q, c.spec, c.twx, lx);
  int tiy = std::max(1, std::min(16, 2048 / g.ny)), tiz = std::max(1, std::min(16, 2048 / g.nz));
  int nty = (g.nxh + tiy - 1) / tiy, ntz = (g.nxh + tiz - 1) / tiz;
  size_t shy = sizeof(double2) * 2 * tiy * g.ny, shz = sizeof(double2) * 2 * tiz * g.nz;
  if (!g.walls) {
    hipLaunchKernelGGL(k_fft_yz, dim3(nty * g.nz), dim3(256), shy, s, g, c.spec, c.twy, 1, 0, kind, 0, 0, tiy, c.kx2, c.ky2, c.kz2);
    hipLaunchKernelGGL(k_fft_yz, dim3(ntz * g.ny), dim3(256), shz, s, g, c.spec, c.twz, 2, 2, kind, 0, g.ny, tiz, c.kx2, c.ky2, c.kz2);
    hipLaunchKernelGGL(k_fft_yz, dim3(nty * g.nz), dim3(256), shy, s, g, c.spec, c.twy, 1, 1, kind, 0, 0, tiy, c.kx2, c.ky2, c.kz2);
  } else {
    hipLaunchKernelGGL(k_fft_yz, dim3(ntz * g.ny), dim3(256), shz, s, g, c.spec, c.twz, 2, 0, kind, 1, g.ny, tiz, c.kx2, c.ky2, c.kz2);
    const int tl = thomas_lines(g.ny, (int)(sizeof(double2) + sizeof(double)));
    const int ngrp = (g.nxh + tl - 1) / tl;
    hipLaunchKernelGGL(k_thomas, dim3(ngrp * g.nz), dim3(256), (size_t)(g.ny + 2) * tl * (sizeof(double2) + sizeof(double)),
                       s, g, c.spec, kind, tl, c.kx2, c.kz2);
    hipLaunchKernelGGL(k_fft_yz, dim3(ntz * g.ny), dim3(256), shz, s, g, c.spec, c.twz, 2, 1, kind, 1, g.ny, tiz, c.kx2, c.ky2, c.kz2);
  }
  hipLaunchKernelGGL(k_fft_x_inv, dim3(nblk(xlines, lx)), dim3(256), shx, s, g, c.spec, x, c.twx, lx);
}

// ======================================================= projection =======
// fft_poisson right-hand side (fenep.f90:1550-1552)
template <bool WL>
__global__ __launch_bounds__(256) void k_divergence(Grid g, const double* __restrict__ up, double* __restrict__ q) {
  const long L = xcd_cell();
  if (L >= g.n) return;
  long m;
  int i, j, k;
  band_cell(g, L, m, i, j, k);
  const double *Uu = up, *Vv = up + g.n, *Ww = up + 2 * g.n;
  q[m] = (((-ldbw<WL>(Uu, g, i - 1, j, k, BC_U) + ldbw<WL>(Uu, g, i, j, k, BC_U)) / g.dx
           + (-ldbw<WL>(Vv, g, i, j - 1, k, BC_V) + ldbw<WL>(Vv, g, i, j, k, BC_V)) / g.dy)
          + (-ldbw<WL>(Ww, g, i, j, k - 1, BC_W) + ldbw<WL>(Ww, g, i, j, k, BC_W)) / g.dz) / g.dt;
}

void f64_divergence(Ctx& c, const double* up, double* q) {
  hipLaunchKernelGGL(c.g.walls ? k_divergence<true> : k_divergence<false>, dim3(nblk(c.g.n, 256)), dim3(256), 0, c.stream, c.g, up, q);
}

// march / fft_march (fenep.f90:907-946, 1562-1604)
template <bool WL>
__global__ __launch_bounds__(256) void k_march(Grid g, int explicit_p, const double* __restrict__ up, const double* __restrict__ phi,
                                               double* __restrict__ U, double* __restrict__ V, double* __restrict__ W,
                                               double* __restrict__ P) {
  const long L = xcd_cell();
  if (L >= g.n) return;
  long m;
  int i, j, k;
  band_cell(g, L, m, i, j, k);
  const double dt = g.dt, dx = g.dx, dy = g.dy, dz = g.dz;
  const double *Uu = up, *Vv = up + g.n, *Ww = up + 2 * g.n;
#define PH(a, b, c) ldbw<WL>(phi, g, a, b, c, BC_P)
  double ph = phi[m];
  U[m] = ldbw<WL>(Uu, g, i, j, k, BC_U) - (dt * (-ph + PH(i + 1, j, k))) / dx;
  V[m] = ldbw<WL>(Vv, g, i, j, k, BC_V) - (dt * (-ph + PH(i, j + 1, k))) / dy;
  W[m] = ldbw<WL>(Ww, g, i, j, k, BC_W) - (dt * (-ph + PH(i, j, k + 1))) / dz;
  if (explicit_p) {
    P[m] = P[m] + ph;
  } else {
    const double coef = ((dt * g.beta) / g.re) / 2.0;
    double lap = ((((PH(i - 1, j, k) - 2.0 * ph) + PH(i + 1, j, k)) / (dx * dx)
                   + ((PH(i, j - 1, k) - 2.0 * ph) + PH(i, j + 1, k)) / (dy * dy))
                  + ((PH(i, j, k - 1) - 2.0 * ph) + PH(i, j, k + 1)) / (dz * dz));
    P[m] = (P[m] + ph) - coef * lap;
  }
#undef PH
}

// ============================================================ IBM ==========
// interpolation and marker forces of ibm_Helmholtz (fenep.f90:2076-2119,
// forces kept as in fenep_mpi.f90:1845-1853)
// w: the 27 weights ibm_weight of each marker and component (k_ibm_interp_w).
__global__ __launch_bounds__(256) void k_ibm_interp(Grid g, long nmk, int nc, int nl, const double* __restrict__ mk,
                                                    const double* __restrict__ w, const double* __restrict__ ua,
                                                    double* __restrict__ mb, double* __restrict__ mf) {
  long id = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (id >= nmk) return;
  int n = (int)(id / ((long)nc * nl)) + 1;
  double xc = mk[id], yc = mk[nmk + id];
  const double dx3 = (g.dx * g.dx) * g.dx;
  for (int c = 0; c < 3; c++) {
    int i0, j0; ibm_origin_d(g, c, xc, yc, i0, j0);
    const double* X = ua + c * g.n;
    const double* wm = w + (c * nmk + id) * 27;
    int kind = c == 0 ? BC_U : (c == 1 ? BC_V : BC_W);
    double acc = 0.0;
    int slot = 0;
    for (int kk = n - 1; kk <= n + 1; kk++)
      for (int jj = j0; jj <= j0 + 2; jj++)
        for (int ii = i0; ii <= i0 + 2; ii++, slot++) {
          double v = ldb(X, g, ii - 1, jj - 1, kk - 1, kind);
          acc = acc + (v * wm[slot]) * dx3;
        }
    mb[c * nmk + id] = acc;
    double uc = mk[(3 + c) * nmk + id];
    mf[c * nmk + id] = 0.0 + (uc - acc) / g.dt;
  }
}

// spreading (fenep.f90:2125-2162) as a gather over the CSR of contributions
// of each ghosted target, in the reference's marker loop order.
// (w: each entry's weight, k_ibm_spread_w).  The three components in one
// launch: component blockIdx.y, its CSR and its ghosted array ft0 + comp gn.
struct SpreadCsr {
  const int* rowp[3];
  const long* tgt[3];
  const int* ent[3];
  const double* w[3];
  int rows[3];
};
__global__ __launch_bounds__(256) void k_ibm_spread(Grid g, SpreadCsr cs, long nmk, const double* __restrict__ mf,
                                                    double* __restrict__ ft0, long gn) {
  const int comp = blockIdx.y;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= cs.rows[comp]) return;
  const int* __restrict__ rowp = cs.rowp[comp];
  const int* __restrict__ ent = cs.ent[comp];
  const double* __restrict__ w = cs.w[comp];
  long t = cs.tgt[comp][r];
  double acc = 0.0;
  for (int e = rowp[r]; e < rowp[r + 1]; e++) acc = acc + (mf[comp * nmk + ent[e]] * w[e]) * g.sw;
  ft0[comp * gn + t] = acc;
}

// The weights depend only on the markers and the grid: ibm_weight of every
// spreading entry (its marker at its row's target) and of every marker's 27
// interpolation points per component, evaluated once per marker set.
__global__ __launch_bounds__(256) void k_ibm_spread_w(Grid g, int comp, int rows, const int* __restrict__ rowp,
                                                      const long* __restrict__ tgt, const int* __restrict__ ent,
                                                      long nmk, const double* __restrict__ mk, double* __restrict__ w) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const long gx = g.nx + 2, gy = g.ny + 2;
  long t = tgt[r];
  int ii = (int)(t % gx), jj = (int)((t / gx) % gy), kk = (int)(t / (gx * gy));
  for (int e = rowp[r]; e < rowp[r + 1]; e++) {
    long id = ent[e];
    w[e] = ibm_weight(g, comp, ii, jj, kk, mk[id], mk[nmk + id], mk[2 * nmk + id]);
  }
}
__global__ __launch_bounds__(256) void k_ibm_interp_w(Grid g, long nmk, int nc, int nl, const double* __restrict__ mk,
                                                      double* __restrict__ w) {
  long id = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (id >= nmk) return;
  int n = (int)(id / ((long)nc * nl)) + 1;
  double xc = mk[id], yc = mk[nmk + id], zc = mk[2 * nmk + id];
  for (int c = 0; c < 3; c++) {
    int i0, j0; ibm_origin_d(g, c, xc, yc, i0, j0);
    double* wm = w + (c * nmk + id) * 27;
    int slot = 0;
    for (int kk = n - 1; kk <= n + 1; kk++)
      for (int jj = j0; jj <= j0 + 2; jj++)
        for (int ii = i0; ii <= i0 + 2; ii++, slot++) wm[slot] = ibm_weight(g, c, ii, jj, kk, xc, yc, zc);
  }
}

// ibm_PBM fold (fenep.f90:2195-2205), one axis per launch, the component
// (ghosted array ft + blockIdx.y gn) per grid row
__global__ __launch_bounds__(256) void k_ibm_fold(Grid g, int axis, double* __restrict__ ft0, long gn) {
  const long gx = g.nx + 2, gy = g.ny + 2, gz = g.nz + 2;
  double* __restrict__ ft = ft0 + blockIdx.y * gn;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  auto G = [&](long i, long j, long k) { return i + gx * (j + gy * k); };
  if (axis == 0) {
    if (t >= gy * gz) return;
    long j = t % gy, k = t / gy;
    ft[G(1, j, k)] = ft[G(1, j, k)] + ft[G(g.nx + 1, j, k)];
    ft[G(g.nx, j, k)] = ft[G(g.nx, j, k)] + ft[G(0, j, k)];
  } else if (axis == 1) {
    if (t >= gx * gz) return;
    long i = t % gx, k = t / gx;
    ft[G(i, 1, k)] = ft[G(i, 1, k)] + ft[G(i, g.ny + 1, k)];
    ft[G(i, g.ny, k)] = ft[G(i, g.ny, k)] + ft[G(i, 0, k)];
  } else {
    if (t >= gx * gy) return;
    long i = t % gx, j = t / gx;
    ft[G(i, j, 1)] = ft[G(i, j, 1)] + ft[G(i, j, g.nz + 1)];
    ft[G(i, j, g.nz)] = ft[G(i, j, g.nz)] + ft[G(i, j, 0)];
  }
}

// ibm_predict right-hand side (fenep.f90:2229): Ua + dt f - dt B/2
__global__ __launch_bounds__(256) void k_ibm_rhs(Grid g, const double* __restrict__ ua, const double* __restrict__ ft,
                                                 const double* __restrict__ B, double* __restrict__ q) {
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= 3 * g.n) return;
  int c = (int)(m / g.n);
  long mm = m % g.n;
  int i, j, k;
  cell_ijk(g, mm, i, j, k);
  const long gx = g.nx + 2, gy = g.ny + 2, gn = gx * gy * (g.nz + 2);
  double f = ft[c * gn + (i + 1) + gx * ((j + 1) + gy * (long)(k + 1))];
  int kind = c == 0 ? BC_U : (c == 1 ? BC_V : BC_W);
  double u = ldb(ua + c * g.n, g, i, j, k, kind);
  q[m] = (u + g.dt * f) - (g.dt * B[m]) / 2.0;
}

// spread of the marker values mfield (3 x nmk) onto the ghosted arrays ft
// (3 x gn) followed by the ghost folds; rows of the CSR that no marker
// reaches keep their previous (zero) value.
void f64_spread_markers(Ctx& c, const double* mfield, double* ftmp) {
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  const long gx = g.nx + 2, gy = g.ny + 2, gz = g.nz + 2;
  SpreadCsr cs{};
  int rmax = 0;
  for (int comp = 0; comp < 3; comp++) {
    cs.rowp[comp] = c.sp_row[comp];
    cs.tgt[comp] = c.sp_tgt[comp];
    cs.ent[comp] = c.sp_ent[comp];
    cs.w[comp] = c.sp_w[comp];
    cs.rows[comp] = c.sp_rows[comp];
    rmax = std::max(rmax, c.sp_rows[comp]);
  }
  if (rmax > 0)
    hipLaunchKernelGGL(k_ibm_spread, dim3(nblk(rmax, 256), 3), dim3(256), 0, s, g, cs, c.nmk, mfield, ftmp, c.gn);
  // the folds of the three components, x, then y, then z
  hipLaunchKernelGGL(k_ibm_fold, dim3(nblk(gy * gz, 256), 3), dim3(256), 0, s, g, 0, ftmp, c.gn);
  hipLaunchKernelGGL(k_ibm_fold, dim3(nblk(gx * gz, 256), 3), dim3(256), 0, s, g, 1, ftmp, c.gn);
  hipLaunchKernelGGL(k_ibm_fold, dim3(nblk(gx * gy, 256), 3), dim3(256), 0, s, g, 2, ftmp, c.gn);
}

void f64_ibm(Ctx& c) {
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  hipLaunchKernelGGL(k_ibm_interp, dim3(nblk(c.nmk, 256)), dim3(256), 0, s, g, c.nmk, c.nc, c.nl, c.mk, c.ib_w, c.Q, c.mb,
                     c.mf);
  f64_spread_markers(c, c.mf, c.ftmp);
  hipLaunchKernelGGL(k_ibm_rhs, dim3(nblk(3 * g.n, 256)), dim3(256), 0, s, g, c.Q, c.ftmp, c.B, c.Up);
}

// Host-side construction of the spreading CSR: for each ghosted target the
// contributing (marker, stencil slot) pairs in the order of the reference's
// loops (markers n, m, l; then k, j, i), extended by every point the ghost
// folds write to, so that one gather pass per step rewrites all of them.
void f64_ibm_setup(Ctx& c) {
  const Grid& g = c.g;
  std::vector<double> mk(6 * c.nmk);
  FP_HIP_CHECK(hipMemcpy(mk.data(), c.mk, sizeof(double) * mk.size(), hipMemcpyDeviceToHost));
  const long gx = g.nx + 2, gy = g.ny + 2, gz = g.nz + 2, gn = gx * gy * gz;
  for (int comp = 0; comp < 3; comp++) {
    std::map<long, std::vector<int>> rows;
    for (long id = 0; id < c.nmk; id++) {
      int n = (int)(id / ((long)c.nc * c.nl)) + 1;
      double xc = mk[id], yc = mk[c.nmk + id];
      int i0, j0;
      if (comp == 0) { i0 = (int)(xc / g.dx - 0.5); j0 = (int)(yc / g.dy); }
      else if (comp == 1) { i0 = (int)(xc / g.dx); j0 = (int)(yc / g.dy - 0.5); }
      else { i0 = (int)(xc / g.dx); j0 = (int)(yc / g.dy); }
      int slot = 0;
      for (int kk = n - 1; kk <= n + 1; kk++)
        for (int jj = j0; jj <= j0 + 2; jj++)
          for (int ii = i0; ii <= i0 + 2; ii++, slot++) {
            if (ii < 0 || ii > g.nx + 1 || jj < 0 || jj > g.ny + 1 || kk < 0 || kk > g.nz + 1)
              throw std::string("IBM stencil outside the ghosted grid");
            rows[ii + gx * (jj + gy * (long)kk)].push_back((int)id);
          }
    }
    // fold closure
    std::vector<char> mark(gn, 0);
    for (auto& kv : rows) mark[kv.first] = 1;
    auto G = [&](long i, long j, long k) { return i + gx * (j + gy * k); };
    for (long k = 0; k < gz; k++) for (long j = 0; j < gy; j++) {
      mark[G(1, j, k)] |= mark[G(g.nx + 1, j, k)]; mark[G(g.nx, j, k)] |= mark[G(0, j, k)];
    }
    for (long k = 0; k < gz; k++) for (long i = 0; i < gx; i++) {
      mark[G(i, 1, k)] |= mark[G(i, g.ny + 1, k)]; mark[G(i, g.ny, k)] |= mark[G(i, 0, k)];
    }
    for (long j = 0; j < gy; j++) for (long i = 0; i < gx; i++) {
      mark[G(i, j, 1)] |= mark[G(i, j, g.nz + 1)]; mark[G(i, j, g.nz)] |= mark[G(i, j, 0)];
    }
    std::vector<int> rowp(1, 0), ent;
    std::vector<long> tgt;
    for (long t = 0; t < gn; t++) {
      if (!mark[t]) continue;
      auto it = rows.find(t);
      if (it != rows.end()) for (int id : it->second) ent.push_back(id);
      tgt.push_back(t);
      rowp.push_back((int)ent.size());
    }
    if (c.sp_row[comp]) {
      (void)hipFree(c.sp_row[comp]); (void)hipFree(c.sp_tgt[comp]); (void)hipFree(c.sp_ent[comp]); (void)hipFree(c.sp_w[comp]);
    }
    c.sp_rows[comp] = (int)tgt.size();
    FP_HIP_CHECK(hipMalloc(&c.sp_row[comp], sizeof(int) * rowp.size()));
    FP_HIP_CHECK(hipMalloc(&c.sp_tgt[comp], sizeof(long) * std::max<size_t>(1, tgt.size())));
    FP_HIP_CHECK(hipMalloc(&c.sp_ent[comp], sizeof(int) * std::max<size_t>(1, ent.size())));
    FP_HIP_CHECK(hipMalloc(&c.sp_w[comp], sizeof(double) * std::max<size_t>(1, ent.size())));
    FP_HIP_CHECK(hipMemcpy(c.sp_row[comp], rowp.data(), sizeof(int) * rowp.size(), hipMemcpyHostToDevice));
    if (!tgt.empty()) FP_HIP_CHECK(hipMemcpy(c.sp_tgt[comp], tgt.data(), sizeof(long) * tgt.size(), hipMemcpyHostToDevice));
    if (!ent.empty()) FP_HIP_CHECK(hipMemcpy(c.sp_ent[comp], ent.data(), sizeof(int) * ent.size(), hipMemcpyHostToDevice));
    if (!tgt.empty())
      hipLaunchKernelGGL(k_ibm_spread_w, dim3(nblk((long)tgt.size(), 256)), dim3(256), 0, c.stream, g, comp,
                         (int)tgt.size(), c.sp_row[comp], c.sp_tgt[comp], c.sp_ent[comp], c.nmk, c.mk, c.sp_w[comp]);
  }
  if (c.ib_w) (void)hipFree(c.ib_w);
  FP_HIP_CHECK(hipMalloc(&c.ib_w, sizeof(double) * std::max<long>(1, 3 * 27 * c.nmk)));
  if (c.nmk > 0)
    hipLaunchKernelGGL(k_ibm_interp_w, dim3(nblk(c.nmk, 256)), dim3(256), 0, c.stream, g, c.nmk, c.nc, c.nl, c.mk, c.ib_w);
  FP_HIP_CHECK(hipGetLastError());
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
}

// ============================================================ SOR ==========
// Red-black SOR for the pressure-correction Poisson equation (the reference
// sweeps lexicographically with odd/even i per row, fenep.f90:885-901; the
// converged solution is the same to the solver tolerance).  Residuals are
// taken before each update, as E in the reference.
__global__ __launch_bounds__(256) void k_sor_color(Grid g, int color, double om, const double* __restrict__ q,
                                                   double* __restrict__ phi, double* __restrict__ e2) {
  long h = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long half = g.n / 2;
  double e = 0.0;
  if (h < half) {
    // enumerate cells of one colour: (i + j + k) % 2 == color
    long jk = h / (g.nx / 2);
    int j = (int)(jk % g.ny), k = (int)(jk / g.ny);
    int i = (int)(2 * (h % (g.nx / 2))) + ((j + k + color) & 1);
    long m = cidx(g, i, j, k);
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    double BX = 1.0 / (dx * dx), BY = 1.0 / (dy * dy), BZ = 1.0 / (dz * dz);
    double B0 = ((((BX + BX) + BY) + BY) + BZ) + BZ;
    double p0 = phi[m];
    e = (((((BZ * ldb(phi, g, i, j, k - 1, BC_P) + BY * ldb(phi, g, i, j - 1, k, BC_P)) + BX * ldb(phi, g, i - 1, j, k, BC_P))
           - B0 * p0) + BX * ldb(phi, g, i + 1, j, k, BC_P)) + BY * ldb(phi, g, i, j + 1, k, BC_P)
         + BZ * ldb(phi, g, i, j, k + 1, BC_P)) - q[m];
    phi[m] = p0 + (om * e) / B0;
  }
  // block partial of e^2 (the second colour adds to the first's partial)
  __shared__ double sh[256];
  double s = block_sum256(e * e, sh);
  if (threadIdx.x == 0) e2[blockIdx.x] = color == 0 ? s : e2[blockIdx.x] + s;
}

__global__ __launch_bounds__(256) void k_sumsq(long n, const double* __restrict__ q, double* __restrict__ part) {
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double v = m < n ? q[m] * q[m] : 0.0;
  __shared__ double sh[256];
  double s = block_sum256(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_zero_n(double* p, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

int f64_sor(Ctx& c, const double* q, double* phi) {
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  double om = c.p.sor_omega > 0.0 ? c.p.sor_omega : 2.0 / (1.0 + sin(3.14159265358979323846 * g.dx));
  const int check = 10;  // residual read back every `check` sweeps
  double* red = c.red;   // [0] = er0, [1 + s] = residual of sweep s
  double* part = c.part; // per-block partial sums
  hipLaunchKernelGGL(k_sumsq, dim3(nblk(g.n, 256)), dim3(256), 0, s, g.n, q, part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, part, nblk(g.n, 256), red, 0);
  double er0 = 0.0;
  FP_HIP_CHECK(hipMemcpyAsync(&er0, red, sizeof(double), hipMemcpyDeviceToHost, s));
  FP_HIP_CHECK(hipStreamSynchronize(s));
  if (er0 == 0.0) return 0;
  long half = g.n / 2;
  const int nbh = nblk(half, 256);
  int itr = 0;
  std::vector<double> res(check);
  while (itr < c.p.sor_itrmax) {
    int batch = std::min(check, c.p.sor_itrmax - itr);
    for (int b = 0; b < batch; b++) {
      hipLaunchKernelGGL(k_sor_color, dim3(nbh), dim3(256), 0, s, g, 0, om, q, phi, part);
      hipLaunchKernelGGL(k_sor_color, dim3(nbh), dim3(256), 0, s, g, 1, om, q, phi, part);
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, part, nbh, red + 1 + b, 0);
    }
    FP_HIP_CHECK(hipMemcpyAsync(res.data(), red + 1, sizeof(double) * batch, hipMemcpyDeviceToHost, s));
    FP_HIP_CHECK(hipStreamSynchronize(s));
    for (int b = 0; b < batch; b++) {
      itr++;
      if (std::sqrt(res[b] / er0) < c.p.sor_eps) return itr;
    }
  }
  return itr;
}

// ======================================================= per step =========
void f64_select(Ctx& c) {
  hipLaunchKernelGGL(c.g.walls ? k_select<true> : k_select<false>, dim3(nblk(c.g.n, 256)), dim3(256), 0, c.stream, c.g, c.C[c.cur], c.sel);
}

void f64_cstar_newton(Ctx& c) {
  hipLaunchKernelGGL(c.g.walls ? k_cstar_newton<true> : k_cstar_newton<false>, dim3(nblk(c.g.n, 128)), dim3(128), 0, c.stream, c.g, c.C[c.cur], c.sel, c.U, c.V, c.W,
                     c.C[1 - c.cur], c.fC, c.nitr, c.p.newton_dc, c.p.newton_eps, c.p.newton_itrmax);
  c.cur = 1 - c.cur;
}

void f64_predictor(Ctx& c) {
  const Grid& g = c.g;
  int polymer = c.p.beta != 1.0;
  double* Bout = c.p.method == FP_METHOD_IBM ? c.B : nullptr;
  hipLaunchKernelGGL(c.g.walls ? k_predictor<true> : k_predictor<false>, dim3(nblk(g.n, 256)), dim3(256), 0, c.stream, g, c.p.method, c.has_history ? 0 : 1,
                     polymer, c.U, c.V, c.W, c.P, c.C[c.cur], c.fC, c.A0, c.T0, c.B0, Bout, c.F,
                     c.p.body_force[0], c.p.body_force[1], c.p.body_force[2], c.Q);
}

void f64_march(Ctx& c) {
  hipLaunchKernelGGL(c.g.walls ? k_march<true> : k_march<false>, dim3(nblk(c.g.n, 256)), dim3(256), 0, c.stream, c.g,
                     c.p.method == FP_METHOD_EXPLICIT ? 1 : 0, c.Up, c.Phi, c.U, c.V, c.W, c.P);
}

void f64_zero_c(Ctx& c) {
  zero_words(c.C[c.cur], 12 * c.g.n, c.stream);
}

// =============================================== time-series statistics ===
// Per-cell terms of all_time_dif / all_time_dif_map (fenep_mpi.f90:2200-2564)
// for the velocity fluctuation u' = u - u_mean (u_mean = nullptr: none, the
// 'same' records; otherwise the 'diff' records): 0-2 regular-grid u'^2,
// 3-5 squared staggered gradients of u', 6 S:S with S = D + D^T of the
// 4-point gradient of u', 7 |u|^2 on the staggered grid (raw u, the
// K_staggered term), 8 log f(C), 9 tr C.  The ghosts of u' are those of u
// minus those of the mean, each by the boundary rules.
constexpr int TS_N = 10;
__device__ __forceinline__ void stat_cell(const Grid& g, const double* __restrict__ U, const double* __restrict__ V,
                                          const double* __restrict__ W, const double* __restrict__ Um,
                                          const double* __restrict__ C, int i, int j, int k, long m, double* s) {
  const double dx = g.dx, dy = g.dy, dz = g.dz;
  const double* Vm = Um ? Um + g.n : nullptr;
  const double* Wm = Um ? Um + 2 * g.n : nullptr;
  auto UG = [&](int a, int b, int c) {
    return ldb(U, g, i + a, j + b, k + c, BC_U) - (Um ? ldb(Um, g, i + a, j + b, k + c, BC_U) : 0.0);
  };
  auto VG = [&](int a, int b, int c) {
    return ldb(V, g, i + a, j + b, k + c, BC_V) - (Vm ? ldb(Vm, g, i + a, j + b, k + c, BC_V) : 0.0);
  };
  auto WG = [&](int a, int b, int c) {
    return ldb(W, g, i + a, j + b, k + c, BC_W) - (Wm ? ldb(Wm, g, i + a, j + b, k + c, BC_W) : 0.0);
  };
  const double u0 = UG(0, 0, 0), v0 = VG(0, 0, 0), w0 = WG(0, 0, 0);
  const double ur = (u0 + UG(-1, 0, 0)) / 2.0, vr = (v0 + VG(0, -1, 0)) / 2.0, wr = (w0 + WG(0, 0, -1)) / 2.0;
  s[0] = ur * ur; s[1] = vr * vr; s[2] = wr * wr;
  double D[3][3];
  D[0][0] = (u0 - UG(-1, 0, 0)) / dx;
  D[0][1] = (UG(0, 1, 0) - UG(0, -1, 0) + UG(-1, 1, 0) - UG(-1, -1, 0)) / (4 * dy);
  D[0][2] = (UG(0, 0, 1) - UG(0, 0, -1) + UG(-1, 0, 1) - UG(-1, 0, -1)) / (4 * dz);
  D[1][0] = (VG(1, 0, 0) - VG(-1, 0, 0) + VG(1, -1, 0) - VG(-1, -1, 0)) / (4 * dx);
  D[1][1] = (v0 - VG(0, -1, 0)) / dy;
  D[1][2] = (VG(0, 0, 1) - VG(0, 0, -1) + VG(0, -1, 1) - VG(0, -1, -1)) / (4 * dz);
  D[2][0] = (WG(1, 0, 0) - WG(-1, 0, 0) + WG(1, 0, -1) - WG(-1, 0, -1)) / (4 * dx);
  D[2][1] = (WG(0, 1, 0) - WG(0, -1, 0) + WG(0, 1, -1) - WG(0, -1, -1)) / (4 * dy);
  D[2][2] = (w0 - WG(0, 0, -1)) / dz;
  s[3] = D[0][0] * D[0][0]; s[4] = D[1][1] * D[1][1]; s[5] = D[2][2] * D[2][2];
  s[6] = 0.0;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) { const double sab = D[a][b] + D[b][a]; s[6] += sab * sab; }
  s[7] = U[m] * U[m] + V[m] * V[m] + W[m] * W[m];
  const double t = (C[m] + C[3 * g.n + m]) + C[5 * g.n + m];
  const double L2 = g.lp * g.lp;
  s[8] = log((L2 - 3.0) / (L2 - t));
  s[9] = t;
}

__global__ __launch_bounds__(256) void k_time_stats(Grid g, const double* __restrict__ U, const double* __restrict__ V,
                                                    const double* __restrict__ W, const double* __restrict__ Um,
                                                    const double* __restrict__ C, double* __restrict__ part, int nb) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double s[TS_N];
  for (int q = 0; q < TS_N; q++) s[q] = 0.0;
  if (m < g.n) {
    int i, j, k;
    cell_ijk(g, m, i, j, k);
    stat_cell(g, U, V, W, Um, C, i, j, k, m, s);
  }
  __shared__ double sh[256];
  for (int q = 0; q < TS_N; q++) {
    const double v = block_sum256(s[q], sh);
    if (threadIdx.x == 0) part[(long)q * nb + blockIdx.x] = v;
    __syncthreads();
  }
}

// all_time_dif_map (fenep_mpi.f90:2350-2500): the per-cell terms summed over
// z for each (i, j) column, in k order, then the z variance of tr C about the
// column mean.  col[q * nx*ny + i + nx*j]: q = 0..9 as stat_cell (7 unused),
// q = 10 the sum of (tr C - column mean)^2.
constexpr int COL_N = 11;
__global__ __launch_bounds__(256) void k_column_stats(Grid g, const double* __restrict__ U, const double* __restrict__ V,
                                                      const double* __restrict__ W, const double* __restrict__ Um,
                                                      const double* __restrict__ C, double* __restrict__ col) {
  const long ij = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nxy = (long)g.nx * g.ny;
  if (ij >= nxy) return;
  const int i = (int)(ij & (g.nx - 1)), j = (int)(ij >> g.lnx);
  double acc[TS_N], s[TS_N];
  for (int q = 0; q < TS_N; q++) acc[q] = 0.0;
  for (int k = 0; k < g.nz; k++) {
    const long m = ij + nxy * k;
    stat_cell(g, U, V, W, Um, C, i, j, k, m, s);
    for (int q = 0; q < TS_N; q++) acc[q] += s[q];
  }
  const double mean = acc[9] / (double)g.nz;
  double var = 0.0;
  for (int k = 0; k < g.nz; k++) {
    const long m = ij + nxy * k;
    const double t = (C[m] + C[3 * g.n + m]) + C[5 * g.n + m] - mean;
    var += t * t;
  }
  for (int q = 0; q < TS_N; q++) col[q * nxy + ij] = acc[q];
  col[TS_N * nxy + ij] = var;
}

// fft_energy_dif_map (fenep_mpi.f90:2670-2797): the z transform of the
// regular-grid fluctuation velocity of every (i, j) column, normalised by
// 1/NZ (fft_z, fenep_mpi.f90:1442-1463), as a direct sum over z; energy
// (|u^|^2 + |v^|^2 + |w^|^2)/2 per |kz| = 0..NZ/2, doubled for 0 < kz < NZ/2
// (the +-kz pair).  E[kz * nx*ny + i + nx*j].
__global__ __launch_bounds__(256) void k_zspec_columns(Grid g, const double* __restrict__ U, const double* __restrict__ V,
                                                       const double* __restrict__ W, const double* __restrict__ Um,
                                                       double* __restrict__ E) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nxy = (long)g.nx * g.ny;
  const int nk = g.nz / 2 + 1;
  if (t >= nxy * nk) return;
  const long ij = t % nxy;
  const int kz = (int)(t / nxy);
  const int i = (int)(ij & (g.nx - 1)), j = (int)(ij >> g.lnx);
  const double* Vm = Um ? Um + g.n : nullptr;
  const double* Wm = Um ? Um + 2 * g.n : nullptr;
  double re[3] = {0, 0, 0}, im[3] = {0, 0, 0};
  for (int k = 0; k < g.nz; k++) {
    const double ur = ((ldb(U, g, i, j, k, BC_U) - (Um ? ldb(Um, g, i, j, k, BC_U) : 0.0))
                       + (ldb(U, g, i - 1, j, k, BC_U) - (Um ? ldb(Um, g, i - 1, j, k, BC_U) : 0.0))) / 2.0;
    const double vr = ((ldb(V, g, i, j, k, BC_V) - (Vm ? ldb(Vm, g, i, j, k, BC_V) : 0.0))
                       + (ldb(V, g, i, j - 1, k, BC_V) - (Vm ? ldb(Vm, g, i, j - 1, k, BC_V) : 0.0))) / 2.0;
    const double wr = ((ldb(W, g, i, j, k, BC_W) - (Wm ? ldb(Wm, g, i, j, k, BC_W) : 0.0))
                       + (ldb(W, g, i, j, k - 1, BC_W) - (Wm ? ldb(Wm, g, i, j, k - 1, BC_W) : 0.0))) / 2.0;
    // e^{-2 pi i kz k / NZ}; the product kz*k is reduced mod NZ first
    const double a = 2.0 * (double)((long)kz * k & (g.nz - 1)) / (double)g.nz;
    double sn, cs;
    sincospi(a, &sn, &cs);
    re[0] += ur * cs; im[0] -= ur * sn;
    re[1] += vr * cs; im[1] -= vr * sn;
    re[2] += wr * cs; im[2] -= wr * sn;
  }
  const double nz = (double)g.nz;
  double e = 0.0;
  for (int q = 0; q < 3; q++) e += (re[q] / nz) * (re[q] / nz) + (im[q] / nz) * (im[q] / nz);
  E[t] = (kz == 0 || kz == g.nz / 2) ? e / 2.0 : e;
}

// masked average over the (i, j) plane of each kz row: sum(E * mask)/sum(mask)
__global__ __launch_bounds__(256) void k_masked_rows(long nxy, const double* __restrict__ E,
                                                     const double* __restrict__ mask, double msum,
                                                     double* __restrict__ out) {
  const double* row = E + nxy * blockIdx.x;
  double v = 0.0;
  for (long ij = threadIdx.x; ij < nxy; ij += blockDim.x) v += row[ij] * mask[ij];
  __shared__ double sh[256];
  v = block_sum256(v, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = v / msum;
}

// time-mean accumulation of the velocity (fenep_mpi.f90:2969-2971)
template <typename T>
__global__ __launch_bounds__(256) void k_mean_acc(long n, const T* __restrict__ U, const T* __restrict__ V,
                                                  const T* __restrict__ W, double* __restrict__ acc) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= n) return;
  acc[m] += (double)U[m];
  acc[n + m] += (double)V[m];
  acc[2 * n + m] += (double)W[m];
}

// mean = acc / count, then (zavg) its z average for every (i, j) (ave_mean,
// fenep_mpi.f90:2010-2033), in k order
__global__ __launch_bounds__(256) void k_mean_finish(Grid g, const double* __restrict__ acc, double count, int zavg,
                                                     double* __restrict__ mean) {
  const long nxy = (long)g.nx * g.ny;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 3 * nxy) return;
  const int comp = (int)(t / nxy);
  const long ij = t % nxy;
  const double* a = acc + comp * g.n;
  double* o = mean + comp * g.n;
  if (!zavg) {
    for (int k = 0; k < g.nz; k++) o[ij + nxy * k] = a[ij + nxy * k] / count;
    return;
  }
  double s = 0.0;
  for (int k = 0; k < g.nz; k++) s += a[ij + nxy * k] / count;
  s = s / (double)g.nz;
  for (int k = 0; k < g.nz; k++) o[ij + nxy * k] = s;
}

// sum of (tr C - mean)^2, per block
__global__ __launch_bounds__(256) void k_trc_var(Grid g, const double* __restrict__ C, const double* __restrict__ mean,
                                                 double* __restrict__ part) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double v = 0.0;
  if (m < g.n) {
    const double t = (C[m] + C[3 * g.n + m]) + C[5 * g.n + m] - mean[0];
    v = t * t;
  }
  __shared__ double sh[256];
  const double s = block_sum256(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

void f64_time_stats(Ctx& c, double* out) {
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  const int nb = nblk(g.n, 256);
  double* part = nullptr;
  double* red = nullptr;
  FP_HIP_CHECK(hipMalloc((void**)&part, sizeof(double) * (size_t)(TS_N + 1) * nb));
  FP_HIP_CHECK(hipMalloc((void**)&red, sizeof(double) * (TS_N + 2)));
  double r[TS_N + 2];
  try {
    hipLaunchKernelGGL(k_time_stats, dim3(nb), dim3(256), 0, s, g, c.U, c.V, c.W, c.mean_u, c.C[c.cur], part, nb);
    for (int q = 0; q < TS_N; q++)
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, part + (long)q * nb, nb, red + q, 0);
    FP_HIP_CHECK(hipMemcpyAsync(r, red, sizeof(double) * TS_N, hipMemcpyDeviceToHost, s));
    FP_HIP_CHECK(hipStreamSynchronize(s));
    const double mean = r[9] / (double)g.n;
    FP_HIP_CHECK(hipMemcpyAsync(red + TS_N, &mean, sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_trc_var, dim3(nb), dim3(256), 0, s, g, c.C[c.cur], red + TS_N, part + (long)TS_N * nb);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, part + (long)TS_N * nb, nb, red + TS_N + 1, 0);
    FP_HIP_CHECK(hipMemcpyAsync(r + TS_N + 1, red + TS_N + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    FP_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(part); (void)hipFree(red);
    throw;
  }
  (void)hipFree(part);
  (void)hipFree(red);
  const double N = (double)g.n, nu = 1.0 / c.p.re, L2 = c.p.lp * c.p.lp;
  const double urms = r[0] / N, vrms = r[1] / N, wrms = r[2] / N;
  const double all_rms = (urms + vrms + wrms) / 3.0, all_grad = (r[3] / N + r[4] / N + r[5] / N) / 3.0;
  const double eps = nu / 2.0 * r[6] / N;
  const double lambda = std::sqrt(all_rms / all_grad);
  out[0] = (double)c.step;
  out[1] = c.step * c.p.dt;
  out[2] = (urms + vrms + wrms) / 2.0;
  out[3] = r[7] / (2.0 * N);
  out[4] = lambda;
  out[5] = std::sqrt(all_rms) * lambda / nu;
  out[6] = eps;
  out[7] = std::pow(nu * nu * nu / eps, 0.25);
  out[8] = std::sqrt(nu / eps);
  out[9] = std::pow(nu * eps, 0.25);
  out[10] = 0.5 * (1.0 - c.p.beta) / c.p.re / c.p.wi * (L2 - 3.0) * r[8] / N;
  out[11] = r[9] / N;
  out[12] = std::sqrt(r[TS_N + 1] / N);
}

// all_time_dif_map (fenep_mpi.f90:2350-2564): column sums on the device, the
// per-(i, j) statistics and their masked plane averages here.  mask(i, j) = 0
// within 1.5 radii of a cylinder centre (fenep_mpi.f90:2530-2537, with the
// centres given instead of the reference's four fixed ones, and dY for y).
static std::vector<double> plane_mask(const Grid& g, int ncyl, const double* xy, double diameter) {
  std::vector<double> mask((size_t)g.nx * g.ny, 1.0);
  const double r = diameter * 1.5 / 2.0;
  for (int j = 0; j < g.ny; j++)
    for (int i = 0; i < g.nx; i++)
      for (int c = 0; c < ncyl; c++) {
        const double dxc = (i + 0.5) * g.dx - xy[2 * c], dyc = (j + 0.5) * g.dy - xy[2 * c + 1];
        if (dxc * dxc + dyc * dyc < r * r) mask[i + (size_t)g.nx * j] = 0.0;
      }
  return mask;
}

int f64_time_stats_map(Ctx& c, int ncyl, const double* xy, double diameter, double* map18, double* out13) {
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  const long nxy = (long)g.nx * g.ny;
  std::vector<double> col((size_t)COL_N * nxy);
  double* dcol = nullptr;
  FP_HIP_CHECK(hipMalloc((void**)&dcol, sizeof(double) * col.size()));
  try {
    hipLaunchKernelGGL(k_column_stats, dim3(nblk(nxy, 256)), dim3(256), 0, s, g, c.U, c.V, c.W, c.mean_u, c.C[c.cur], dcol);
    FP_HIP_CHECK(hipMemcpyAsync(col.data(), dcol, sizeof(double) * col.size(), hipMemcpyDeviceToHost, s));
    FP_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(dcol);
    throw;
  }
  (void)hipFree(dcol);
  const std::vector<double> mask = plane_mask(g, ncyl, xy, diameter);
  double msum = 0.0;
  for (double m : mask) msum += m;
  if (msum == 0.0) return 1;
  const double NZ = (double)g.nz, nu = 1.0 / c.p.re, L2 = c.p.lp * c.p.lp;
  const double epc = 0.5 * (1.0 - c.p.beta) / c.p.re / c.p.wi * (L2 - 3.0);
  double acc[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (long ij = 0; ij < nxy; ij++) {
    auto C_ = [&](int q) { return col[(size_t)q * nxy + ij]; };
    const double urms = C_(0) / NZ, vrms = C_(1) / NZ, wrms = C_(2) / NZ;
    const double all_rms = (urms + vrms + wrms) / 3.0;
    const double all_grad = (C_(3) / NZ + C_(4) / NZ + C_(5) / NZ) / 3.0;
    const double eps = nu / 2.0 * C_(6) / NZ;
    const double Lk = std::pow(nu * nu * nu / eps, 0.25), Tk = std::sqrt(nu / eps), Uk = std::pow(nu * eps, 0.25);
    const double lam = std::sqrt(all_rms / all_grad);
    const double relam = std::sqrt(all_rms) * lam / nu;
    const double K = (urms + vrms + wrms) / 2.0;
    const double Ep = epc * (C_(8) / NZ);
    const double trm = C_(9) / NZ, trs = std::sqrt(C_(10) / NZ);
    const double vals[11] = {K, lam, relam, eps, Lk, Tk, Uk, Ep, trm, trs, mask[ij]};
    const int i = (int)(ij % g.nx), j = (int)(ij / g.nx);
    double* row = map18 ? map18 + 18 * ij : nullptr;
    if (row) {
      row[0] = (i + 0.5) * g.dx;
      row[1] = (j + 0.5) * g.dy;
      for (int q = 0; q < 11; q++) row[2 + q] = vals[q];
      for (int q = 13; q < 18; q++) row[q] = 0.0;
    }
    for (int q = 0; q < 10; q++) acc[q] += vals[q] * mask[ij];
  }
  if (out13) {
    out13[0] = (double)c.step;
    out13[1] = c.step * c.p.dt;
    out13[2] = acc[0] / msum;   // K
    out13[3] = 0.0;             // no K_staggered in the masked record
    out13[4] = acc[1] / msum;   // lambda
    out13[5] = acc[2] / msum;   // Re_lambda
    out13[6] = acc[3] / msum;   // epsilon
    out13[7] = acc[4] / msum;   // Kolmogorov length, time, velocity
    out13[8] = acc[5] / msum;
    out13[9] = acc[6] / msum;
    out13[10] = acc[7] / msum;  // Ep
    out13[11] = acc[8] / msum;  // tr C mean
    out13[12] = acc[9] / msum;  // tr C std
  }
  return 0;
}

// fft_energy_dif_map (fenep_mpi.f90:2670-2797) of the current state:
// Emask[kz] = masked plane average of the column z spectra, Efirst[kz] = the
// spectrum of column (1, 1); kz = 0..NZ-1 (zero above NZ/2).
int f64_energy_spectrum_z(Ctx& c, int ncyl, const double* xy, double diameter, double* Emask, double* Efirst) {
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  const long nxy = (long)g.nx * g.ny;
  const int nk = g.nz / 2 + 1;
  const std::vector<double> mask = plane_mask(g, ncyl, xy, diameter);
  double msum = 0.0;
  for (double m : mask) msum += m;
  if (msum == 0.0) return 1;
  double *dE = nullptr, *dmask = nullptr, *dout = nullptr;
  std::vector<double> em(nk), ef(nk);
  try {
    FP_HIP_CHECK(hipMalloc((void**)&dE, sizeof(double) * nxy * nk));
    FP_HIP_CHECK(hipMalloc((void**)&dmask, sizeof(double) * nxy));
    FP_HIP_CHECK(hipMalloc((void**)&dout, sizeof(double) * nk));
    FP_HIP_CHECK(hipMemcpyAsync(dmask, mask.data(), sizeof(double) * nxy, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_zspec_columns, dim3(nblk(nxy * nk, 256)), dim3(256), 0, s, g, c.U, c.V, c.W, c.mean_u, dE);
    hipLaunchKernelGGL(k_masked_rows, dim3(nk), dim3(256), 0, s, nxy, dE, dmask, msum, dout);
    FP_HIP_CHECK(hipMemcpyAsync(em.data(), dout, sizeof(double) * nk, hipMemcpyDeviceToHost, s));
    FP_HIP_CHECK(hipMemcpy2DAsync(ef.data(), sizeof(double), dE, sizeof(double) * nxy, sizeof(double), nk,
                                  hipMemcpyDeviceToHost, s));
    FP_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(dE); (void)hipFree(dmask); (void)hipFree(dout);
    throw;
  }
  (void)hipFree(dE); (void)hipFree(dmask); (void)hipFree(dout);
  for (int k = 0; k < g.nz; k++) {
    if (Emask) Emask[k] = k < nk ? em[k] : 0.0;
    if (Efirst) Efirst[k] = k < nk ? ef[k] : 0.0;
  }
  return 0;
}

// time mean of the velocity for the mean-subtracted statistics
void f64_mean_accumulate(Ctx& c, const float* u32, const float* v32, const float* w32) {
  const Grid& g = c.g;
  if (!c.mean_acc) {
    FP_HIP_CHECK(hipMalloc((void**)&c.mean_acc, sizeof(double) * 3 * g.n));
    zero_words(c.mean_acc, 6 * g.n, c.stream);
    c.mean_count = 0;
  }
  if (u32)
    hipLaunchKernelGGL(k_mean_acc<float>, dim3(nblk(g.n, 256)), dim3(256), 0, c.stream, g.n, u32, v32, w32, c.mean_acc);
  else
    hipLaunchKernelGGL(k_mean_acc<double>, dim3(nblk(g.n, 256)), dim3(256), 0, c.stream, g.n, (const double*)c.U,
                       (const double*)c.V, (const double*)c.W, c.mean_acc);
  c.mean_count++;
}

int f64_mean_apply(Ctx& c, int mode) {
  const Grid& g = c.g;
  if (mode == 0) {
    if (c.mean_u) (void)hipFree(c.mean_u);
    c.mean_u = nullptr;
  } else {
    if (!c.mean_acc || c.mean_count == 0) return 1;
    if (!c.mean_u) FP_HIP_CHECK(hipMalloc((void**)&c.mean_u, sizeof(double) * 3 * g.n));
    const long nxy = (long)g.nx * g.ny;
    hipLaunchKernelGGL(k_mean_finish, dim3(nblk(3 * nxy, 256)), dim3(256), 0, c.stream, g, c.mean_acc,
                       (double)c.mean_count, mode == 1 ? 1 : 0, c.mean_u);
  }
  if (c.mean_acc) {
    FP_HIP_CHECK(hipStreamSynchronize(c.stream));
    (void)hipFree(c.mean_acc);
    c.mean_acc = nullptr;
  }
  c.mean_count = 0;
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

// ===================================================== energy spectrum =====
// fft_energy_dif (fenep_mpi.f90:2566-2665) without a subtracted mean: the
// regular-grid velocity, its forward transform / N, |u^|^2 summed over the
// three components into integer shells nint(|k|), half weight on the kx = 0
// and Nyquist planes of the half spectrum.
__global__ __launch_bounds__(256) void k_reg_velocity(Grid g, const double* __restrict__ U, const double* __restrict__ V,
                                                      const double* __restrict__ W, const double* __restrict__ Um,
                                                      double* __restrict__ out) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= g.n) return;
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  out[m] = (U[m] + ldb(U, g, i - 1, j, k, BC_U)) / 2.0;
  out[g.n + m] = (V[m] + ldb(V, g, i, j - 1, k, BC_V)) / 2.0;
  out[2 * g.n + m] = (W[m] + ldb(W, g, i, j, k - 1, BC_W)) / 2.0;
  if (Um) {  // the fluctuation about the subtracted mean
    const double *Vm = Um + g.n, *Wm = Um + 2 * g.n;
    out[m] = ((U[m] - Um[m]) + (ldb(U, g, i - 1, j, k, BC_U) - ldb(Um, g, i - 1, j, k, BC_U))) / 2.0;
    out[g.n + m] = ((V[m] - Vm[m]) + (ldb(V, g, i, j - 1, k, BC_V) - ldb(Vm, g, i, j - 1, k, BC_V))) / 2.0;
    out[2 * g.n + m] = ((W[m] - Wm[m]) + (ldb(W, g, i, j, k - 1, BC_W) - ldb(Wm, g, i, j, k - 1, BC_W))) / 2.0;
  }
}

__global__ __launch_bounds__(256) void k_spec_shells(Grid g, const double2* __restrict__ spec, int nbins,
                                                     double* __restrict__ E) {
  extern __shared__ double bins[];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) bins[b] = 0.0;
  __syncthreads();
  const long total = (long)g.nxh * g.ny * g.nz;
  const double N = (double)g.n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.nxh);
    const int j = (int)((e / g.nxh) % g.ny);
    const int k = (int)(e / ((long)g.nxh * g.ny));
    const double kx = (double)i;
    const double ky = (double)(j <= g.ny / 2 ? j : j - g.ny);
    const double kz = (double)(k <= g.nz / 2 ? k : k - g.nz);
    const int idx = (int)floor(sqrt(kx * kx + ky * ky + kz * kz) + 0.5);
    const double2 s = spec[e];
    const double a = (s.x / N) * (s.x / N) + (s.y / N) * (s.y / N);
    const double w = (i == 0 || i == g.nxh - 1) ? 0.5 : 1.0;
    if (idx < nbins) atomicAdd(&bins[idx], a * w);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x)
    if (bins[b] != 0.0) atomicAdd(&E[b], bins[b]);
}

void f64_energy_spectrum(Ctx& c, double* E, int nbins) {
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  double* dE = nullptr;
  FP_HIP_CHECK(hipMalloc((void**)&dE, sizeof(double) * nbins));
  try {
    zero_words(dE, 2L * nbins, s);
    hipLaunchKernelGGL(k_reg_velocity, dim3(nblk(g.n, 256)), dim3(256), 0, s, g, c.U, c.V, c.W, c.mean_u, c.Q);
    const int lx = lines_for(g.nx);
    const long xlines = (long)g.ny * g.nz;
    const size_t shx = sizeof(double2) * 2 * lx * g.nx;
    const int tiy = std::max(1, std::min(16, 2048 / g.ny)), tiz = std::max(1, std::min(16, 2048 / g.nz));
    const int nty = (g.nxh + tiy - 1) / tiy, ntz = (g.nxh + tiz - 1) / tiz;
    const size_t shy = sizeof(double2) * 2 * tiy * g.ny, shz = sizeof(double2) * 2 * tiz * g.nz;
    const long total = (long)g.nxh * g.ny * g.nz;
    const int sb = (int)std::min<long>(1024, (total + 255) / 256);
    for (int comp = 0; comp < 3; comp++) {
      hipLaunchKernelGGL(k_fft_x_fwd, dim3(nblk(xlines, lx)), dim3(256), shx, s, g, c.Q + comp * g.n, c.spec, c.twx, lx);
      hipLaunchKernelGGL(k_fft_yz, dim3(nty * g.nz), dim3(256), shy, s, g, c.spec, c.twy, 1, 0, 3, 0, 0, tiy, c.kx2, c.ky2,
                         c.kz2);
      hipLaunchKernelGGL(k_fft_yz, dim3(ntz * g.ny), dim3(256), shz, s, g, c.spec, c.twz, 2, 0, 3, 0, g.ny, tiz, c.kx2,
                         c.ky2, c.kz2);
      hipLaunchKernelGGL(k_spec_shells, dim3(sb), dim3(256), sizeof(double) * nbins, s, g, c.spec, nbins, dE);
    }
    FP_HIP_CHECK(hipMemcpyAsync(E, dE, sizeof(double) * nbins, hipMemcpyDeviceToHost, s));
    FP_HIP_CHECK(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipFree(dE);
    throw;
  }
  (void)hipFree(dE);
}

// ===================================================== diagnostics ========
__global__ __launch_bounds__(256) void k_diag(Grid g, const double* __restrict__ U, const double* __restrict__ V,
                                              const double* __restrict__ W, const double* __restrict__ C,
                                              const int* __restrict__ nitr, int polymer, double* __restrict__ red) {
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double s[6] = {0, 0, 0, 0, 0, 0};  // ke, trsum, trsq, nitr, nan, Newton stopped at the cap
  double mx[6] = {-1e300, -1e300, -1e300, -1e300, 1e300, 0.0};  // umax vmax wmax trmax -trmin divmax
  if (m < g.n) {
    int i, j, k;
  cell_ijk(g, m, i, j, k);
    double u = ldb(U, g, i, j, k, BC_U), v = ldb(V, g, i, j, k, BC_V), w = ldb(W, g, i, j, k, BC_W);
    double ur = (u + ldb(U, g, i - 1, j, k, BC_U)) / 2.0;
    double vr = (v + ldb(V, g, i, j - 1, k, BC_V)) / 2.0;
    double wr = (w + ldb(W, g, i, j, k - 1, BC_W)) / 2.0;
    s[0] = ur * ur + vr * vr + wr * wr;
    double tr = (C[m] + C[3 * g.n + m]) + C[5 * g.n + m];
    s[1] = tr; s[2] = tr * tr;
    s[3] = polymer ? (double)abs(nitr[m]) : 0.0;
    s[5] = (polymer && nitr[m] < 0) ? 1.0 : 0.0;
    s[4] = isnan(U[m]) ? 1.0 : 0.0;
    mx[0] = u; mx[1] = v; mx[2] = w; mx[3] = tr; mx[4] = tr;
    double dv = ((u - ldb(U, g, i - 1, j, k, BC_U)) / g.dx + (v - ldb(V, g, i, j - 1, k, BC_V)) / g.dy)
                + (w - ldb(W, g, i, j, k - 1, BC_W)) / g.dz;
    mx[5] = fabs(dv);
  }
  __shared__ double sh[256];
  for (int a = 0; a < 12; a++) {
    double v = a < 6 ? s[a] : mx[a - 6];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if (threadIdx.x < o) {
        double x = sh[threadIdx.x], y = sh[threadIdx.x + o];
        if (a < 6) sh[threadIdx.x] = x + y;
        else if (a == 10) sh[threadIdx.x] = fmin(x, y);
        else sh[threadIdx.x] = fmax(x, y);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) red[a * gridDim.x + blockIdx.x] = sh[0];
    __syncthreads();
  }
}

// the 12 per-block partials of k_diag reduced to one value each: sums for
// a < 6, the minimum of tr C for a = 10, maxima otherwise (one block per a)
__global__ __launch_bounds__(1024) void k_diag_reduce(const double* __restrict__ part, int nb, double* __restrict__ out) {
  const int a = blockIdx.x;
  const double* row = part + (long)a * nb;
  __shared__ double sh[1024];
  const double init = a < 6 ? 0.0 : (a == 10 ? 1e300 : -1e300);
  double v = init;
  for (int b = threadIdx.x; b < nb; b += 1024) {
    const double x = row[b];
    v = a < 6 ? v + x : (a == 10 ? fmin(v, x) : fmax(v, x));
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double x = sh[threadIdx.x], y = sh[threadIdx.x + o];
      sh[threadIdx.x] = a < 6 ? x + y : (a == 10 ? fmin(x, y) : fmax(x, y));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[a] = sh[0];
}

// sums of the x and y marker forces (drag and lift before the -dX^3 factor)
__global__ __launch_bounds__(1024) void k_marker_sums(const double* __restrict__ mf, long nmk, double* __restrict__ out) {
  __shared__ double sh[1024];
  for (int comp = 0; comp < 2; comp++) {
    double v = 0.0;
    for (long id = threadIdx.x; id < nmk; id += 1024) v += mf[comp * nmk + id];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
      if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[comp] = sh[0];
    __syncthreads();
  }
}

size_t diag_scratch_words(const Grid& g) { return (size_t)12 * nblk(g.n, 256) + 16; }

// log_progress quantities (fenep.f90:2279-2311) and the drag and lift of the
// marker forces mf (nullptr: none), reduced on the device in `scratch`
// (diag_scratch_words doubles, owned by the caller); 14 doubles come back.
void f64_diag(Ctx& c, fp_diag* d, double* scratch, const double* mf) {
  const Grid& g = c.g;
  const int nb = nblk(g.n, 256);
  double* out = scratch + (size_t)12 * nb;
  hipLaunchKernelGGL(k_diag, dim3(nb), dim3(256), 0, c.stream, g, c.U, c.V, c.W, c.C[c.cur], c.nitr,
                     c.p.beta != 1.0 ? 1 : 0, scratch);
  hipLaunchKernelGGL(k_diag_reduce, dim3(12), dim3(1024), 0, c.stream, scratch, nb, out);
  if (mf && c.nmk) hipLaunchKernelGGL(k_marker_sums, dim3(1), dim3(1024), 0, c.stream, mf, c.nmk, out + 12);
  double h[14] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  FP_HIP_CHECK(hipMemcpyAsync(h, out, sizeof(double) * ((mf && c.nmk) ? 14 : 12), hipMemcpyDeviceToHost, c.stream));
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
  const double N = (double)g.n;
  d->ke = (h[0] / 2.0) / N;
  d->cfl = std::max(h[6] * g.dt / g.dx, std::max(h[7] * g.dt / g.dy, h[8] * g.dt / g.dz)) * 6.0;
  d->trc_mean = h[1] / N;
  d->trc_std = std::sqrt(std::max(0.0, h[2] / N - d->trc_mean * d->trc_mean));
  d->trc_max = h[9];
  d->trc_min = h[10];
  d->div_max = h[11];
  d->newton_itr_mean = h[3] / N;
  d->newton_unconverged = h[5];
  d->nan_flag = h[4] > 0 ? 1 : 0;
  if (mf && c.nmk) {
    d->drag = -(h[12] * g.sw);
    d->lift = -(h[13] * g.sw);
  }
}

}  // namespace fp
