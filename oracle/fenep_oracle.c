/*
 * fenep_oracle.c -- FP64 CPU restatement of the FENE-P / IBM SMAC time step of
 * machapin/fortran8.  TEST INFRASTRUCTURE ONLY (the checker, never shipped).
 *
 * Every expression is evaluated left to right exactly as written in the
 * Fortran source it cites, with no fused multiply-add (build with
 * -ffp-contract=off), so that a device implementation with the same
 * operation order reproduces it to the last bit apart from the libm
 * transcendentals (acos/cos in the eigenvalue test).  Choices where the
 * reference has no single answer are marked DECISION.
 *
 * Sources followed:
 *   fenep.f90:147-166   PBM / C_PBM ghost copies
 *   fenep.f90:169-382   CpxCnx / CpyCny / CpzCnz, fenep.f90:425-444 Vieta
 *   fenep.f90:464-507   Cstar
 *   fenep.f90:510-643   f, Lyapunov, Lyapunov_func, gauss
 *   fenep.f90:645-691   polymer_stress
 *   fenep.f90:694-770   convection, viscous
 *   fenep.f90:772-946   navier, poisson (SOR), march   (method 0)
 *   fenep.f90:1400-1604 fft_solve, fft_navier, fft_poisson, fft_march (method 1)
 *   fenep.f90:1970-2257 ibm_preliminary, delta/weight, ibm_Helmholtz, ibm_PBM,
 *                       ibm_predict (method 2); forces as fenep_mpi.f90:1845-1853
 *   fenep_2d.f90:1098-1383 wall ghost rules, tdma, wall solves (walls_y=1)
 *   fenep.f90:2477-2527 step order
 */
#include "fenep_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

/* The state scalar: double, or in the C++ complex-step build (ORC_CS) the
 * complex-step type of cstep.h, whose real parts are computed by the same
 * double operations.  Geometry (spacings, twiddles, modified wavenumbers,
 * marker positions) stays double. */
#ifdef ORC_CS
#include "cstep.h"
typedef cs_t real_t;
#else
typedef double real_t;
static inline double rv(double x) { return x; }
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { BC_U = 0, BC_V = 1, BC_W = 2, BC_P = 3, BC_C = 4 };

typedef struct {
  fp_params p;
  int nx, ny, nz, nxh;
  long gx, gy, gz, gn;  /* ghosted dims and count */
  long n;               /* interior count */
  double dx, dy, dz;
  int walls;
  real_t wi, beta, lp;  /* Wi, beta, Lp of the step (p holds their real values) */
  real_t *U, *V, *W, *P, *Phi;
  real_t *C[6], *C0[6];
  real_t *A[3], *B[3], *T[3], *A0[3], *B0[3], *T0[3];
  real_t *Up[3], *Ua[3];
  real_t *F[3];
  real_t *fint[3];
  real_t *ftmp;
  real_t *Cp[3][6], *Cn[3][6];
  real_t *Cx[6];
  real_t *q;            /* interior scratch */
  real_t *e;            /* SOR residual */
  /* spectral */
  double *twx, *twy, *twz, *kx2, *ky2, *kz2;
  real_t *spec;         /* complex (nxh)(sy)(nz) */
  long sy;              /* spectral rows in y: ny (periodic) or ny+2 (walls) */
  int jo;               /* row offset of interior j=1 in spec: 0 or 1 */
  /* markers */
  int nc, nl, nmk;
  double *xc, *yc, *zc, *uc, *vc, *wc;
  real_t *fxc, *fyc, *fzc, *ubm, *vbm, *wbm;
  int has_history, step;
  double newton_itr_total;
  double sor_itr;
  double sel_count[4];
} orc_t;

#define GI(o, i, j, k) ((long)(i) + (o)->gx * ((long)(j) + (o)->gy * (long)(k)))
#define II(o, i, j, k) ((long)((i) - 1) + (long)(o)->nx * ((long)((j) - 1) + (long)(o)->ny * (long)((k) - 1)))

static real_t* zalloc(long n) { return (real_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(real_t)); }
static double* dzalloc(long n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* ------------------------------------------------------------------ ghosts */
/* PBM (fenep.f90:147-155) in x and z; in y either periodic or the wall rules
 * of fenep_2d.f90:1098-1143 with NA=0, NB=NY (stationary walls, Uwall=0).
 * The order x, then y, then z makes edge/corner ghosts consistent. */
static void fill_ghosts(orc_t* o, real_t* A, int kind) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  for (long k = 0; k < o->gz; k++)
    for (long j = 0; j < o->gy; j++) {
      A[GI(o, 0, j, k)] = A[GI(o, nx, j, k)];
      A[GI(o, nx + 1, j, k)] = A[GI(o, 1, j, k)];
    }
  for (long k = 0; k < o->gz; k++)
    for (long i = 0; i < o->gx; i++) {
      if (!o->walls) {
        A[GI(o, i, 0, k)] = A[GI(o, i, ny, k)];
        A[GI(o, i, ny + 1, k)] = A[GI(o, i, 1, k)];
      } else {
        switch (kind) {
          case BC_U: case BC_W:  /* U_kabe: U(NA)=-U(NA+1), U(NB+1)=2Uwall-U(NB) */
            A[GI(o, i, 0, k)] = -A[GI(o, i, 1, k)];
            A[GI(o, i, ny + 1, k)] = 2.0 * 0.0 - A[GI(o, i, ny, k)];
            break;
          case BC_V:             /* V_kabe: V(NA)=0, V(NB)=0, V(NB+1:)=0 */
            A[GI(o, i, 0, k)] = 0.0;
            A[GI(o, i, ny, k)] = 0.0;
            A[GI(o, i, ny + 1, k)] = 0.0;
            break;
          case BC_P:             /* P_kabe: Neumann */
            A[GI(o, i, 0, k)] = A[GI(o, i, 1, k)];
            A[GI(o, i, ny + 1, k)] = A[GI(o, i, ny, k)];
            break;
          default:               /* C_kabe: linear extrapolation */
            A[GI(o, i, 0, k)] = 2.0 * A[GI(o, i, 1, k)] - A[GI(o, i, 2, k)];
            A[GI(o, i, ny + 1, k)] = 2.0 * A[GI(o, i, ny, k)] - A[GI(o, i, ny - 1, k)];
            break;
        }
      }
    }
  for (long j = 0; j < o->gy; j++)
    for (long i = 0; i < o->gx; i++) {
      A[GI(o, i, j, 0)] = A[GI(o, i, j, nz)];
      A[GI(o, i, j, nz + 1)] = A[GI(o, i, j, 1)];
    }
}

/* ------------------------------------------------------------- FENE-P math */
/* Vieta, fenep.f90:425-444: smallest eigenvalue of the symmetric matrix a
 * (xx,xy,xz,yy,yz,zz) by the trigonometric root with +2pi/3; -100 when the
 * discriminant t >= 0. */
static real_t vieta_r(const real_t* a) {
  real_t a2 = -((a[0] + a[3]) + a[5]);
  real_t a1 = (((((a[0] * a[3]) + (a[3] * a[5])) + (a[5] * a[0])) - (a[4] * a[4])) - (a[2] * a[2])) - (a[1] * a[1]);
  real_t a0 = -((((((a[0] * a[3]) * a[5] + (a[1] * a[4]) * a[2]) + (a[1] * a[4]) * a[2]) - (a[0] * a[4]) * a[4]) - (a[1] * a[1]) * a[5]) - (a[2] * a[3]) * a[2]);
  real_t p = a1 - (a2 * a2) / 3.0;
  real_t q = (a0 - (a1 * a2) / 3.0) + (2.0 * ((a2 * a2) * a2)) / 27.0;
  real_t q2 = q / 2.0, p3 = p / 3.0;
  real_t t = q2 * q2 + (p3 * p3) * p3;
  if (t >= 0.0) return -100.0;
  real_t arg = ((3.0 * q) / (2.0 * p)) * sqrt(-(3.0 / p));
  real_t ang = acos(arg) / 3.0 + (2.0 * M_PI) / 3.0;
  return (2.0 * sqrt(-(p / 3.0))) * cos(ang) - a2 / 3.0;
}

/* minval of two values with NaN ignored unless both are NaN (gfortran MINVAL) */
static inline real_t minval2(real_t a, real_t b) {
  if (isnan(a)) return b;
  if (isnan(b)) return a;
  return (b < a) ? b : a;
}

/* Candidate face values and selection of one cell along one axis,
 * fenep.f90:179-230.  Returns the 1-based index (4 = zero slope). */
static int select_r(const real_t* cm, const real_t* c, const real_t* cp, real_t* cpo, real_t* cno) {
  real_t Cpt[3][6], Cnt[3][6], ml[4];
  for (int m = 0; m < 6; m++) {
    Cpt[0][m] = (c[m] * 3.0) / 2.0 - cp[m] / 2.0;
    Cnt[0][m] = c[m] / 2.0 + cp[m] / 2.0;
    Cpt[1][m] = cm[m] / 2.0 + c[m] / 2.0;
    Cnt[1][m] = -(cm[m] / 2.0) + (c[m] * 3.0) / 2.0;
    Cpt[2][m] = (cm[m] / 4.0 + c[m]) - cp[m] / 4.0;
    Cnt[2][m] = (-(cm[m] / 4.0) + c[m]) + cp[m] / 4.0;
  }
  for (int l = 0; l < 3; l++) ml[l] = minval2(vieta_r(Cpt[l]), vieta_r(Cnt[l]));
  ml[3] = 0.0;
  int idx = -1; real_t best = 0.0;
  for (int l = 0; l < 4; l++) {   /* maxloc: first maximum, NaN skipped */
    if (isnan(ml[l])) continue;
    if (idx < 0 || ml[l] > best) { idx = l; best = ml[l]; }
  }
  if (idx == 3) {
    for (int m = 0; m < 6; m++) { cpo[m] = c[m]; cno[m] = c[m]; }
  } else {
    for (int m = 0; m < 6; m++) { cpo[m] = Cpt[idx][m]; cno[m] = Cnt[idx][m]; }
  }
  return idx + 1;
}

/* gauss, fenep.f90:608-643: 6x6 Gaussian elimination with partial pivoting,
 * Gauss-Jordan form; x = 0 when a zero pivot is met.  a is column-major. */
static void gauss6_r(const real_t* a, const real_t* b, real_t* x) {
  real_t Ah[6][7];
  for (int r = 0; r < 6; r++) {
    for (int c = 0; c < 6; c++) Ah[r][c] = a[r + 6 * c];
    Ah[r][6] = b[r];
  }
  for (int k = 0; k < 6; k++) {
    int mi = k; real_t ma = fabs(Ah[k][k]);
    for (int r = k + 1; r < 6; r++) if (fabs(Ah[r][k]) > ma) { ma = fabs(Ah[r][k]); mi = r; }
    if (ma == 0.0) { for (int r = 0; r < 6; r++) Ah[r][6] = 0.0; break; }
    if (mi != k) for (int c = k; c < 7; c++) { real_t t = Ah[k][c]; Ah[k][c] = Ah[mi][c]; Ah[mi][c] = t; }
    real_t piv = Ah[k][k];
    for (int c = k + 1; c < 7; c++) Ah[k][c] = Ah[k][c] / piv;
    Ah[k][k] = 1.0;
    for (int r = 0; r < 6; r++) {
      if (r == k) continue;
      real_t f = Ah[r][k];
      for (int c = k + 1; c < 7; c++) Ah[r][c] = Ah[r][c] - f * Ah[k][c];
      Ah[r][k] = 0.0;
    }
  }
  for (int r = 0; r < 6; r++) x[r] = Ah[r][6];
}

/* Lyapunov_func, fenep.f90:588-605: l = M(x) x - b with
 * A = A0 + diag((1 + f(x) dt/Wi)/2) and f the Peterlin factor (fenep.f90:513). */
static inline void lyap_func(const real_t* A0, const real_t* b, const real_t* x,
                             double dt, real_t wi, real_t lp, real_t* l) {
  real_t A[9];
  for (int m = 0; m < 9; m++) A[m] = A0[m];
  real_t f = (lp * lp - 3.0) / (((lp * lp - x[0]) - x[3]) - x[5]);
  real_t dg = (1.0 + (f * dt) / wi) / 2.0;
#define AM(r, c) A[((r) - 1) + 3 * ((c) - 1)]
  AM(1, 1) = AM(1, 1) + dg;
  AM(2, 2) = AM(2, 2) + dg;
  AM(3, 3) = AM(3, 3) + dg;
  l[0] = ((((AM(1, 1) + AM(1, 1)) * x[0] + AM(1, 2) * x[1]) + AM(1, 3) * x[2]) + AM(1, 2) * x[1]) + AM(1, 3) * x[2] - b[0];
  l[1] = ((((AM(2, 1) * x[0] + (AM(2, 2) + AM(1, 1)) * x[1]) + AM(2, 3) * x[2]) + AM(1, 2) * x[3]) + AM(1, 3) * x[4]) - b[1];
  l[2] = ((((AM(3, 1) * x[0] + AM(3, 2) * x[1]) + (AM(3, 3) + AM(1, 1)) * x[2]) + AM(1, 2) * x[4]) + AM(1, 3) * x[5]) - b[2];
  l[3] = ((((AM(2, 1) * x[1] + AM(2, 1) * x[1]) + (AM(2, 2) + AM(2, 2)) * x[3]) + AM(2, 3) * x[4]) + AM(2, 3) * x[4]) - b[3];
  l[4] = ((((AM(2, 1) * x[2] + AM(3, 1) * x[1]) + AM(3, 2) * x[3]) + (AM(3, 3) + AM(2, 2)) * x[4]) + AM(2, 3) * x[5]) - b[4];
  l[5] = ((((AM(3, 1) * x[2] + AM(3, 2) * x[4]) + AM(3, 1) * x[2]) + AM(3, 2) * x[4]) + (AM(3, 3) + AM(3, 3)) * x[5]) - b[5];
#undef AM
}

/* One cell of Lyapunov (fenep.f90:554-578): Newton from x0 with a central
 * finite-difference Jacobian; returns the iteration count. */
static int newton_cell_r(const real_t* A0, const real_t* b, const real_t* x0, double dt,
                         real_t wi, real_t lp, double dc, double eps, int itrmax, real_t* xo) {
  real_t x[6], x1[6], x2[6], l[6], l1[6], l2[6], J[36], r[6];
  for (int m = 0; m < 6; m++) x[m] = x0[m];
  int itr;
  for (itr = 1; itr <= itrmax; itr++) {
    for (int col = 0; col < 6; col++) {
      for (int m = 0; m < 6; m++) { x1[m] = x[m]; x2[m] = x[m]; }
      x1[col] = x1[col] + dc;
      x2[col] = x2[col] - dc;
      lyap_func(A0, b, x1, dt, wi, lp, l1);
      lyap_func(A0, b, x2, dt, wi, lp, l2);
      for (int row = 0; row < 6; row++) J[row + 6 * col] = (l1[row] - l2[row]) / (2.0 * dc);
    }
    lyap_func(A0, b, x, dt, wi, lp, l);
    gauss6_r(J, l, r);
    for (int m = 0; m < 6; m++) x[m] = x[m] - r[m];
    real_t sr = 0.0, sx = 0.0;
    for (int m = 0; m < 6; m++) sr = sr + r[m] * r[m];
    for (int m = 0; m < 6; m++) sx = sx + x[m] * x[m];
    if (sqrt(sr / sx) < eps) break;
  }
  if (itr > itrmax) itr = itrmax;
  for (int m = 0; m < 6; m++) xo[m] = x[m];
  return itr;
}

/* ------------------------------------------------------------------- FFT */
/* Stockham radix-2 autosort DFT, forward e^{-2 pi i jk/n}, unnormalised;
 * inverse with the conjugate twiddles.  tw holds (cos, -sin)(2 pi j/n),
 * j < n/2, computed once on the host. Returns the buffer with the result. */
static real_t* stockham(int n, real_t* x, real_t* y, const double* tw, int inv) {
  int len = n, s = 1;
  while (len > 1) {
    int m = len / 2;
    for (int p = 0; p < m; p++) {
      double wr = tw[2 * (p * s)], wi = tw[2 * (p * s) + 1];
      if (inv) wi = -wi;
      for (int q = 0; q < s; q++) {
        const real_t* a = x + 2 * (q + s * p);
        const real_t* b = x + 2 * (q + s * (p + m));
        real_t* y0 = y + 2 * (q + s * (2 * p));
        real_t* y1 = y + 2 * (q + s * (2 * p + 1));
        real_t ar = a[0], ai = a[1], br = b[0], bi = b[1];
        y0[0] = ar + br;
        y0[1] = ai + bi;
        real_t dr = ar - br, di = ai - bi;
        y1[0] = dr * wr - di * wi;
        y1[1] = dr * wi + di * wr;
      }
    }
    real_t* t = x; x = y; y = t;
    len = m; s *= 2;
  }
  return x;
}

static double* twiddles(int n) {
  double* tw = dzalloc(n > 1 ? n : 2);
  for (int j = 0; j < n / 2; j++) {
    double ang = (2.0 * M_PI * (double)j) / (double)n;
    tw[2 * j] = cos(ang);
    tw[2 * j + 1] = -sin(ang);
  }
  return tw;
}

/* modified wavenumber of the 3-point second difference: 2(1-cos th)/h^2 with
 * th = 2 pi m/N (DECISION: fenep.f90:1458 uses (i-1)*dX, equal only for
 * L = 2 pi) */
static double* modwav(int n, double h, int count) {
  double* k2 = dzalloc(count);
  for (int m = 0; m < count; m++) {
    double th = (2.0 * M_PI * (double)m) / (double)n;
    k2[m] = (2.0 * (1.0 - cos(th))) / (h * h);
  }
  return k2;
}

#ifndef ORC_CS
static void fft_r(int n, real_t* data, int inverse) {
  double* tw = twiddles(n);
  real_t* w = zalloc(2 * n);
  real_t* r = stockham(n, data, w, tw, inverse);
  if (r != data) memcpy(data, r, sizeof(real_t) * 2 * n);
  free(w); free(tw);
}
#endif

/* real line (n) -> half spectrum (n/2+1) */
static void rfft_line(orc_t* o, const real_t* in, long in_stride, real_t* out, long out_stride,
                      real_t* b0, real_t* b1) {
  int n = o->nx;
  for (int i = 0; i < n; i++) { b0[2 * i] = in[i * in_stride]; b0[2 * i + 1] = 0.0; }
  real_t* r = stockham(n, b0, b1, o->twx, 0);
  for (int i = 0; i <= n / 2; i++) { out[2 * i * out_stride] = r[2 * i]; out[2 * i * out_stride + 1] = r[2 * i + 1]; }
}
/* half spectrum -> real line (c2r: imaginary parts of DC and Nyquist ignored) */
static void irfft_line(orc_t* o, const real_t* in, long in_stride, real_t* out, long out_stride,
                       real_t* b0, real_t* b1) {
  int n = o->nx, h = n / 2;
  for (int i = 0; i <= h; i++) { b0[2 * i] = in[2 * i * in_stride]; b0[2 * i + 1] = in[2 * i * in_stride + 1]; }
  b0[1] = 0.0;
  b0[2 * h + 1] = 0.0;
  for (int i = 1; i < h; i++) { b0[2 * (n - i)] = b0[2 * i]; b0[2 * (n - i) + 1] = -b0[2 * i + 1]; }
  real_t* r = stockham(n, b0, b1, o->twx, 1);
  for (int i = 0; i < n; i++) out[i * out_stride] = r[2 * i];
}
/* complex lines along an axis of the spectral array */
static void cfft_axis(orc_t* o, int axis, int inv, long j0, long j1) {
  long nxh = o->nxh, sy = o->sy, nz = o->nz;
  int n = axis == 1 ? o->ny : o->nz;
  const double* tw = axis == 1 ? o->twy : o->twz;
  real_t* b0 = zalloc(2 * n); real_t* b1 = zalloc(2 * n);
  if (axis == 1) {
    for (long k = 0; k < nz; k++)
      for (long i = 0; i < nxh; i++) {
        real_t* base = o->spec + 2 * (i + nxh * (o->jo + sy * k));
        for (int j = 0; j < n; j++) { b0[2 * j] = base[2 * nxh * j]; b0[2 * j + 1] = base[2 * nxh * j + 1]; }
        real_t* r = stockham(n, b0, b1, tw, inv);
        for (int j = 0; j < n; j++) { base[2 * nxh * j] = r[2 * j]; base[2 * nxh * j + 1] = r[2 * j + 1]; }
      }
  } else {
    for (long j = j0; j < j1; j++)
      for (long i = 0; i < nxh; i++) {
        real_t* base = o->spec + 2 * (i + nxh * j);
        long st = 2 * nxh * sy;
        for (int k = 0; k < n; k++) { b0[2 * k] = base[st * k]; b0[2 * k + 1] = base[st * k + 1]; }
        real_t* r = stockham(n, b0, b1, tw, inv);
        for (int k = 0; k < n; k++) { base[st * k] = r[2 * k]; base[st * k + 1] = r[2 * k + 1]; }
      }
  }
  free(b0); free(b1);
}

/* tdma, fenep_2d.f90:1145-1177 with real coefficients (a diagonal, b upper,
 * c lower) and complex right-hand side; pivots below 1e-16 give P=Q=0. */
static void tdma_r(int n, const real_t* a, const real_t* b, const real_t* c, const real_t* dre,
                   const real_t* dim, real_t* xre, real_t* xim) {
  real_t* P = zalloc(n); real_t* Qr = zalloc(n); real_t* Qi = zalloc(n);
  P[0] = -b[0] / a[0];
  Qr[0] = dre[0] / a[0];
  Qi[0] = dim[0] / a[0];
  for (int i = 1; i < n; i++) {
    real_t den = a[i] + c[i] * P[i - 1];
    if (fabs(den) < 1.0e-16) { P[i] = 0.0; Qr[i] = 0.0; Qi[i] = 0.0; }
    else {
      P[i] = -b[i] / den;
      Qr[i] = (dre[i] - c[i] * Qr[i - 1]) / den;
      Qi[i] = (dim[i] - c[i] * Qi[i - 1]) / den;
    }
  }
  xre[n - 1] = Qr[n - 1]; xim[n - 1] = Qi[n - 1];
  for (int i = n - 2; i >= 0; i--) {
    xre[i] = P[i] * xre[i + 1] + Qr[i];
    xim[i] = P[i] * xim[i + 1] + Qi[i];
  }
  free(P); free(Qr); free(Qi);
}

/* Solve on the grid. kind 0..2: Helmholtz [1-(dt beta/2Re) lap_h] x = q for
 * the u,v,w component; kind 3: Poisson lap_h x = q.  Periodic: fft_solve
 * (fenep.f90:1400-1428) with the LHS of fenep.f90:1455-1463 / 1536-1544.
 * Walls: FFT in x,z, then the tridiagonal line solve with the boundary rows
 * of fenep_2d.f90:1222-1237 (U,W), 1262-1276 (V), 1318-1331 (Phi). */
static void solve_grid(orc_t* o, int kind, const real_t* q, real_t* x) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const long nxh = o->nxh, sy = o->sy;
  real_t* b0 = zalloc(2 * nx); real_t* b1 = zalloc(2 * nx);
  /* forward x */
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      rfft_line(o, q + II(o, 1, j, k), 1, o->spec + 2 * (nxh * ((j - 1 + o->jo) + sy * (k - 1))), 1, b0, b1);
  const real_t coef = ((o->p.dt * o->beta) / o->p.re) / 2.0;
  if (!o->walls) {
    cfft_axis(o, 1, 0, 0, 0);
    cfft_axis(o, 2, 0, 0, sy);
    const double ntot = (double)((long)nx * ny * nz);
    for (long k = 0; k < nz; k++)
      for (long j = 0; j < ny; j++)
        for (long i = 0; i < nxh; i++) {
          real_t lhs;
          if (kind == 3) lhs = -((o->kx2[i] + o->ky2[j]) + o->kz2[k]);
          else lhs = 1.0 + coef * ((o->kx2[i] + o->ky2[j]) + o->kz2[k]);
          real_t* s = o->spec + 2 * (i + nxh * (j + sy * k));
          real_t re = s[0] / ntot, im = s[1] / ntot;
          if (fabs(lhs) < 1.0e-16) { s[0] = 0.0; s[1] = 0.0; }
          else { s[0] = re / lhs; s[1] = im / lhs; }
        }
    cfft_axis(o, 2, 1, 0, sy);
    cfft_axis(o, 1, 1, 0, 0);
  } else {
    cfft_axis(o, 2, 0, 1, ny + 1);
    const double nxz = (double)((long)nx * nz);
    const int size = ny + 2;
    real_t *a = zalloc(size), *b = zalloc(size), *c = zalloc(size);
    real_t *dr = zalloc(size), *di = zalloc(size), *xr = zalloc(size), *xi = zalloc(size);
    const double idy2 = o->dy * o->dy;
    for (long k = 0; k < nz; k++)
      for (long i = 0; i < nxh; i++) {
        real_t kk = o->kx2[i] + o->kz2[k];
        for (int r = 1; r <= ny; r++) {
          real_t* s = o->spec + 2 * (i + nxh * (r + sy * k));
          dr[r] = s[0] / nxz; di[r] = s[1] / nxz;
        }
        if (kind == 3) {
          for (int r = 1; r <= ny; r++) { a[r] = -(kk + 2.0 / idy2); b[r] = 1.0 / idy2; c[r] = 1.0 / idy2; }
          a[0] = 1.0; b[0] = -1.0; c[0] = 0.0; dr[0] = 0.0; di[0] = 0.0;
          a[size - 1] = 1.0; c[size - 1] = -1.0; b[size - 1] = 0.0; dr[size - 1] = 0.0; di[size - 1] = 0.0;
        } else if (kind == 1) {
          for (int r = 1; r <= ny - 1; r++) { a[r] = 1.0 + coef * (kk + 2.0 / idy2); b[r] = -(coef / idy2); c[r] = -(coef / idy2); }
          a[0] = 1.0; b[0] = 0.0; c[0] = 0.0; dr[0] = 0.0; di[0] = 0.0;
          for (int r = ny; r <= ny + 1; r++) { a[r] = 1.0; b[r] = 0.0; c[r] = 0.0; dr[r] = 0.0; di[r] = 0.0; }
        } else {
          for (int r = 1; r <= ny; r++) { a[r] = 1.0 + coef * (kk + 2.0 / idy2); b[r] = -(coef / idy2); c[r] = -(coef / idy2); }
          a[0] = 1.0; b[0] = 1.0; c[0] = 0.0; dr[0] = 0.0; di[0] = 0.0;
          a[size - 1] = 1.0; c[size - 1] = 1.0; b[size - 1] = 0.0; dr[size - 1] = 0.0; di[size - 1] = 0.0;
        }
        tdma_r(size, a, b, c, dr, di, xr, xi);
        for (int r = 0; r < size; r++) {
          real_t* s = o->spec + 2 * (i + nxh * (r + sy * k));
          s[0] = xr[r]; s[1] = xi[r];
        }
      }
    free(a); free(b); free(c); free(dr); free(di); free(xr); free(xi);
    cfft_axis(o, 2, 1, 1, ny + 1);
  }
  /* inverse x */
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      irfft_line(o, o->spec + 2 * (nxh * ((j - 1 + o->jo) + sy * (k - 1))), 1, x + II(o, 1, j, k), 1, b0, b1);
  free(b0); free(b1);
}

/* ------------------------------------------------------------ operators */
static void cpm_axis(orc_t* o, int d) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const long s = d == 0 ? 1 : (d == 1 ? o->gx : o->gx * o->gy);
  const long fn[3] = {nx + 1, ny + 1, nz + 1};
  double cnt[4] = {0, 0, 0, 0};
#pragma omp parallel for reduction(+ : cnt[:4]) schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long m = GI(o, i, j, k);
        real_t cm[6], c[6], cp[6], po[6], no[6];
        for (int q = 0; q < 6; q++) { cm[q] = o->C[q][m - s]; c[q] = o->C[q][m]; cp[q] = o->C[q][m + s]; }
        int idx = select_r(cm, c, cp, po, no);
        cnt[idx - 1] += 1.0;
        long fl, fr;  /* faces a-1 (Cp) and a (Cn) along the axis */
        if (d == 0) { fl = (i - 1) + fn[0] * ((j - 1) + (long)ny * (k - 1)); fr = fl + 1; }
        else if (d == 1) { fl = (i - 1) + (long)nx * ((j - 1) + fn[1] * (k - 1)); fr = fl + nx; }
        else { fl = (i - 1) + (long)nx * ((j - 1) + (long)ny * (k - 1)); fr = fl + (long)nx * ny; }
        for (int q = 0; q < 6; q++) { o->Cp[d][q][fl] = po[q]; o->Cn[d][q][fr] = no[q]; }
      }
  for (int l = 0; l < 4; l++) o->sel_count[l] += cnt[l];
  /* periodic wrap of the end faces (fenep.f90:240-241) */
  for (int q = 0; q < 6; q++) {
    real_t* Cp = o->Cp[d][q]; real_t* Cn = o->Cn[d][q];
    if (d == 0) {
      for (long r = 0; r < (long)ny * nz; r++) { Cp[nx + fn[0] * r] = Cp[0 + fn[0] * r]; Cn[0 + fn[0] * r] = Cn[nx + fn[0] * r]; }
    } else if (d == 1) {
      for (long k = 0; k < nz; k++) for (long i = 0; i < nx; i++) {
        Cp[i + nx * (ny + fn[1] * k)] = Cp[i + nx * (0 + fn[1] * k)];
        Cn[i + nx * (0 + fn[1] * k)] = Cn[i + nx * (ny + fn[1] * k)];
      }
    } else {
      long pl = (long)nx * ny;
      for (long r = 0; r < pl; r++) { Cp[r + pl * nz] = Cp[r]; Cn[r] = Cn[r + pl * nz]; }
    }
  }
}

/* Cstar, fenep.f90:464-507 */
static void cstar(orc_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const real_t dt = o->p.dt, dx = o->dx, dy = o->dy, dz = o->dz;
  const long fx = nx + 1, fy = ny + 1;
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        real_t u1 = o->U[GI(o, i - 1, j, k)], u2 = o->U[GI(o, i, j, k)];
        real_t v1 = o->V[GI(o, i, j - 1, k)], v2 = o->V[GI(o, i, j, k)];
        real_t w1 = o->W[GI(o, i, j, k - 1)], w2 = o->W[GI(o, i, j, k)];
        real_t s1 = ((-u2 + fabs(u2)) * dt) / (2.0 * dx);
        real_t s2 = 1.0 / 6.0 + ((-u2 - fabs(u2)) * dt) / (2.0 * dx);
        real_t s3 = 1.0 / 6.0 + ((u1 - fabs(u1)) * dt) / (2.0 * dx);
        real_t s4 = ((u1 + fabs(u1)) * dt) / (2.0 * dx);
        real_t s5 = ((-v2 + fabs(v2)) * dt) / (2.0 * dy);
        real_t s6 = 1.0 / 6.0 + ((-v2 - fabs(v2)) * dt) / (2.0 * dy);
        real_t s7 = 1.0 / 6.0 + ((v1 - fabs(v1)) * dt) / (2.0 * dy);
        real_t s8 = ((v1 + fabs(v1)) * dt) / (2.0 * dy);
        real_t s9 = ((-w2 + fabs(w2)) * dt) / (2.0 * dz);
        real_t s10 = 1.0 / 6.0 + ((-w2 - fabs(w2)) * dt) / (2.0 * dz);
        real_t s11 = 1.0 / 6.0 + ((w1 - fabs(w1)) * dt) / (2.0 * dz);
        real_t s12 = ((w1 + fabs(w1)) * dt) / (2.0 * dz);
        long xl = (i - 1) + fx * ((j - 1) + (long)ny * (k - 1)), xr = xl + 1;
        long yl = (i - 1) + (long)nx * ((j - 1) + fy * (k - 1)), yr = yl + nx;
        long zl = (i - 1) + (long)nx * ((j - 1) + (long)ny * (k - 1)), zr = zl + (long)nx * ny;
        long m = II(o, i, j, k);
        for (int q = 0; q < 6; q++) {
          o->Cx[q][m] = (((((((((((s3 * o->Cp[0][q][xl] + s1 * o->Cp[0][q][xr]) + s4 * o->Cn[0][q][xl]) + s2 * o->Cn[0][q][xr])
                       + s7 * o->Cp[1][q][yl]) + s5 * o->Cp[1][q][yr]) + s8 * o->Cn[1][q][yl]) + s6 * o->Cn[1][q][yr])
                       + s11 * o->Cp[2][q][zl]) + s9 * o->Cp[2][q][zr]) + s12 * o->Cn[2][q][zl]) + s10 * o->Cn[2][q][zr]);
        }
      }
}

/* Lyapunov, fenep.f90:518-585 */
static void lyapunov(orc_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const double dt = o->p.dt, dx = o->dx, dy = o->dy, dz = o->dz;
  const real_t wi = o->wi;
  double itr_total = 0.0;
#pragma omp parallel for reduction(+ : itr_total) schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        const real_t *U = o->U, *V = o->V, *W = o->W;
#define Ug(a, b, c) U[GI(o, a, b, c)]
#define Vg(a, b, c) V[GI(o, a, b, c)]
#define Wg(a, b, c) W[GI(o, a, b, c)]
        long m = II(o, i, j, k);
        real_t bb[6];
        for (int q = 0; q < 6; q++) bb[q] = o->Cx[q][m];
        bb[0] = bb[0] + dt / wi;
        bb[3] = bb[3] + dt / wi;
        bb[5] = bb[5] + dt / wi;
        real_t A[9];  /* A0 built then transposed: A[r + 3c] = A0(c, r) */
        real_t a11 = -((dt * (-Ug(i - 1, j, k) + Ug(i, j, k))) / dx);
        real_t a12 = -((dt * (((-Vg(i - 1, j - 1, k) + Vg(i + 1, j - 1, k)) - Vg(i - 1, j, k)) + Vg(i + 1, j, k))) / (4.0 * dx));
        real_t a13 = -((dt * (((-Wg(i - 1, j, k - 1) + Wg(i + 1, j, k - 1)) - Wg(i - 1, j, k)) + Wg(i + 1, j, k))) / (4.0 * dx));
        real_t a21 = -((dt * (((-Ug(i - 1, j - 1, k) - Ug(i, j - 1, k)) + Ug(i - 1, j + 1, k)) + Ug(i, j + 1, k))) / (4.0 * dy));
        real_t a22 = -((dt * (-Vg(i, j - 1, k) + Vg(i, j, k))) / dy);
        real_t a23 = -((dt * (((-Wg(i, j - 1, k - 1) + Wg(i, j + 1, k - 1)) - Wg(i, j - 1, k)) + Wg(i, j + 1, k))) / (4.0 * dy));
        real_t a31 = -((dt * (((-Ug(i - 1, j, k - 1) - Ug(i, j, k - 1)) + Ug(i - 1, j, k + 1)) + Ug(i, j, k + 1))) / (4.0 * dz));
        real_t a32 = -((dt * (((-Vg(i, j - 1, k - 1) - Vg(i, j, k - 1)) + Vg(i, j - 1, k + 1)) + Vg(i, j, k + 1))) / (4.0 * dz));
        real_t a33 = -((dt * (-Wg(i, j, k - 1) + Wg(i, j, k))) / dz);
#undef Ug
#undef Vg
#undef Wg
        /* transpose(A0): A(r,c) = A0(c,r) */
        A[0] = a11; A[3] = a21; A[6] = a31;
        A[1] = a12; A[4] = a22; A[7] = a32;
        A[2] = a13; A[5] = a23; A[8] = a33;
        real_t x0[6], x[6];
        long g = GI(o, i, j, k);
        for (int q = 0; q < 6; q++) x0[q] = o->C[q][g];
        int itr = newton_cell_r(A, bb, x0, dt, wi, o->lp, o->p.newton_dc, o->p.newton_eps,
                                o->p.newton_itrmax, x);
        itr_total += (double)itr;
        for (int q = 0; q < 6; q++) o->Cx[q][m] = x[q];  /* stage; C is read by neighbours */
      }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        for (int q = 0; q < 6; q++) o->C[q][GI(o, i, j, k)] = o->Cx[q][II(o, i, j, k)];
  for (int q = 0; q < 6; q++) fill_ghosts(o, o->C[q], BC_C);
  o->newton_itr_total = itr_total;
}

/* polymer_stress, fenep.f90:645-691 (C0 pointwise on the ghosted array) */
static void polymer_stress(orc_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const double dx = o->dx, dy = o->dy, dz = o->dz;
  const real_t lp = o->lp;
  static const double E[6] = {1.0, 0.0, 0.0, 1.0, 0.0, 1.0};
  for (long g = 0; g < o->gn; g++) {
    real_t f = (lp * lp - 3.0) / (((lp * lp - o->C[0][g]) - o->C[3][g]) - o->C[5][g]);
    for (int q = 0; q < 6; q++) o->C0[q][g] = f * o->C[q][g] - E[q];
  }
  const real_t coef = (1.0 - o->beta) / (o->p.re * o->wi);
#define C0g(q, a, b, c) o->C0[q][GI(o, a, b, c)]
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long m = II(o, i, j, k);
        real_t tx = ((-C0g(0, i, j, k) + C0g(0, i + 1, j, k)) / dx
                     + (((-C0g(1, i, j - 1, k) - C0g(1, i + 1, j - 1, k)) + C0g(1, i, j + 1, k)) + C0g(1, i + 1, j + 1, k)) / (4.0 * dy))
                    + (((-C0g(2, i, j, k - 1) - C0g(2, i + 1, j, k - 1)) + C0g(2, i, j, k + 1)) + C0g(2, i + 1, j, k + 1)) / (4.0 * dz);
        real_t ty = ((((-C0g(1, i - 1, j, k) + C0g(1, i + 1, j, k)) - C0g(1, i - 1, j + 1, k)) + C0g(1, i + 1, j + 1, k)) / (4.0 * dx)
                     + (-C0g(3, i, j, k) + C0g(3, i, j + 1, k)) / dy)
                    + (((-C0g(4, i, j, k - 1) - C0g(4, i, j + 1, k - 1)) + C0g(4, i, j, k + 1)) + C0g(4, i, j + 1, k + 1)) / (4.0 * dz);
        real_t tz = ((((-C0g(2, i - 1, j, k) + C0g(2, i + 1, j, k)) - C0g(2, i - 1, j, k + 1)) + C0g(2, i + 1, j, k + 1)) / (4.0 * dx)
                     + (((-C0g(4, i, j - 1, k) + C0g(4, i, j + 1, k)) - C0g(4, i, j - 1, k + 1)) + C0g(4, i, j + 1, k + 1)) / (4.0 * dy))
                    + (-C0g(5, i, j, k) + C0g(5, i, j, k + 1)) / dz;
        o->T[0][m] = coef * tx;
        o->T[1][m] = coef * ty;
        o->T[2][m] = coef * tz;
      }
#undef C0g
}

/* convection, fenep.f90:694-734 */
static void convection(orc_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const real_t dx = o->dx, dy = o->dy, dz = o->dz;
#define Ug(a, b, c) o->U[GI(o, a, b, c)]
#define Vg(a, b, c) o->V[GI(o, a, b, c)]
#define Wg(a, b, c) o->W[GI(o, a, b, c)]
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long m = II(o, i, j, k);
        real_t h1 = (Ug(i - 1, j, k) + Ug(i, j, k)) / 2.0, h2 = (Ug(i, j, k) + Ug(i + 1, j, k)) / 2.0;
        real_t ax = ((-(h1 * h1) + h2 * h2) / dx
                     + (-((((Vg(i, j - 1, k) + Vg(i + 1, j - 1, k)) / 2.0) * (Ug(i, j - 1, k) + Ug(i, j, k))) / 2.0)
                        + (((Vg(i, j, k) + Vg(i + 1, j, k)) / 2.0) * (Ug(i, j, k) + Ug(i, j + 1, k))) / 2.0) / dy)
                    + (-((((Wg(i, j, k - 1) + Wg(i + 1, j, k - 1)) / 2.0) * (Ug(i, j, k - 1) + Ug(i, j, k))) / 2.0)
                       + (((Wg(i, j, k) + Wg(i + 1, j, k)) / 2.0) * (Ug(i, j, k) + Ug(i, j, k + 1))) / 2.0) / dz;
        real_t g1 = (Vg(i, j - 1, k) + Vg(i, j, k)) / 2.0, g2 = (Vg(i, j, k) + Vg(i, j + 1, k)) / 2.0;
        real_t ay = ((-((((Ug(i - 1, j, k) + Ug(i - 1, j + 1, k)) / 2.0) * (Vg(i - 1, j, k) + Vg(i, j, k))) / 2.0)
                      + (((Ug(i, j, k) + Ug(i, j + 1, k)) / 2.0) * (Vg(i, j, k) + Vg(i + 1, j, k))) / 2.0) / dx
                     + (-(g1 * g1) + g2 * g2) / dy)
                    + (-((((Wg(i, j, k - 1) + Wg(i, j + 1, k - 1)) / 2.0) * (Vg(i, j, k - 1) + Vg(i, j, k))) / 2.0)
                       + (((Wg(i, j, k) + Wg(i, j + 1, k)) / 2.0) * (Vg(i, j, k) + Vg(i, j, k + 1))) / 2.0) / dz;
        real_t e1 = (Wg(i, j, k - 1) + Wg(i, j, k)) / 2.0, e2 = (Wg(i, j, k) + Wg(i, j, k + 1)) / 2.0;
        real_t az = ((-((((Ug(i - 1, j, k) + Ug(i - 1, j, k + 1)) / 2.0) * (Wg(i - 1, j, k) + Wg(i, j, k))) / 2.0)
                      + (((Ug(i, j, k) + Ug(i, j, k + 1)) / 2.0) * (Wg(i, j, k) + Wg(i + 1, j, k))) / 2.0) / dx
                     + (-((((Vg(i, j - 1, k) + Vg(i, j - 1, k + 1)) / 2.0) * (Wg(i, j - 1, k) + Wg(i, j, k))) / 2.0)
                        + (((Vg(i, j, k) + Vg(i, j, k + 1)) / 2.0) * (Wg(i, j, k) + Wg(i, j + 1, k))) / 2.0) / dy)
                    + (-(e1 * e1) + e2 * e2) / dz;
        o->A[0][m] = -ax;
        o->A[1][m] = -ay;
        o->A[2][m] = -az;
      }
}

/* viscous, fenep.f90:736-770 */
static void viscous(orc_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const real_t dx = o->dx, dy = o->dy, dz = o->dz;
  const real_t cb = o->beta / o->p.re;
  real_t* F[3] = {o->U, o->V, o->W};
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long m = II(o, i, j, k);
        for (int c = 0; c < 3; c++) {
          const real_t* X = F[c];
          real_t x0 = X[GI(o, i, j, k)];
          real_t b = (((X[GI(o, i - 1, j, k)] - 2.0 * x0) + X[GI(o, i + 1, j, k)]) / (dx * dx)
                      + ((X[GI(o, i, j - 1, k)] - 2.0 * x0) + X[GI(o, i, j + 1, k)]) / (dy * dy))
                     + ((X[GI(o, i, j, k - 1)] - 2.0 * x0) + X[GI(o, i, j, k + 1)]) / (dz * dz);
          o->B[c][m] = cb * b;
        }
      }
#undef Ug
#undef Vg
#undef Wg
}

static inline real_t fval(orc_t* o, int c, long m) {
  return (o->F[c] ? o->F[c][m] : 0.0) + o->p.body_force[c];
}

/* pressure gradient term dt*(-P(i)+P(i+1))/h for component c at cell (i,j,k) */
static inline real_t dpdx(orc_t* o, const real_t* P, int c, int i, int j, int k) {
  real_t h = c == 0 ? o->dx : (c == 1 ? o->dy : o->dz);
  real_t p0 = P[GI(o, i, j, k)];
  real_t p1 = c == 0 ? P[GI(o, i + 1, j, k)] : (c == 1 ? P[GI(o, i, j + 1, k)] : P[GI(o, i, j, k + 1)]);
  return (o->p.dt * (-p0 + p1)) / h;
}

static const int BCK[3] = {BC_U, BC_V, BC_W};

/* right-hand side of the predictor, per method */
static void predictor_rhs(orc_t* o, int c, real_t* Q) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const real_t dt = o->p.dt;
  real_t* Uc = c == 0 ? o->U : (c == 1 ? o->V : o->W);
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long m = II(o, i, j, k);
        real_t u = Uc[GI(o, i, j, k)];
        real_t r;
        if (o->p.method == FP_METHOD_EXPLICIT) {
          /* navier, fenep.f90:803-806 */
          r = (((u - dpdx(o, o->P, c, i, j, k))
                + (dt * (3.0 * (o->A[c][m] + o->B[c][m]) - (o->A0[c][m] + o->B0[c][m]))) / 2.0)
               + (dt * (o->T[c][m] + o->T0[c][m])) / 2.0)
              + dt * fval(o, c, m);
        } else if (o->p.method == FP_METHOD_FFT) {
          /* fft_navier, fenep.f90:1469-1474 */
          r = ((((u - dpdx(o, o->P, c, i, j, k)) + (dt * (3.0 * o->A[c][m] - o->A0[c][m])) / 2.0)
                + (dt * o->B[c][m]) / 2.0)
               + (dt * (o->T[c][m] + o->T0[c][m])) / 2.0)
              + dt * fval(o, c, m);
        } else {
          /* ibm_preliminary, fenep.f90:1997-2000; DECISION: + dt*F (the 3D
           * reference omits the body force here, fenep.f90:1994-2013) */
          r = ((((u - dpdx(o, o->P, c, i, j, k)) + (dt * (3.0 * o->A[c][m] - o->A0[c][m])) / 2.0)
                + dt * o->B[c][m])
               + (dt * (o->T[c][m] + o->T0[c][m])) / 2.0)
              + dt * fval(o, c, m);
        }
        Q[m] = r;
      }
}

static void interior_to_ghosted(orc_t* o, const real_t* in, real_t* out) {
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++) out[GI(o, i, j, k)] = in[II(o, i, j, k)];
}

/* ------------------------------------------------------------------- IBM */
/* weight, fenep.f90:2045-2055: 3-point kernel of Roma, Peskin & Berger */
static double roma_weight(double r) {
  double a = fabs(r);
  if (a < 1.5 && a >= 0.5) return ((5.0 - 3.0 * a) - sqrt(-(3.0 * ((1.0 - a) * (1.0 - a))) + 1.0)) / 6.0;
  if (a < 0.5) return (1.0 + sqrt(-(3.0 * (r * r)) + 1.0)) / 3.0;
  return 0.0;
}
/* delta, fenep.f90:2039-2043 */
static inline double delta3(orc_t* o, double x, double y, double z, double xc, double yc, double zc) {
  return ((roma_weight((x - xc) / o->dx) * roma_weight((y - yc) / o->dy)) * roma_weight((z - zc) / o->dz))
         / ((o->dx * o->dy) * o->dz);
}
/* stencil origin and sample position of component c for marker (xc,yc) at
 * grid index (i,j,k) (fenep.f90:2080-2104): X(i)=(i-0.5)dX, +0.5dX for the
 * staggered component. */
static inline void ibm_origin(orc_t* o, int c, double xc, double yc, int* i0, int* j0) {
  if (c == 0) { *i0 = (int)(xc / o->dx - 0.5); *j0 = (int)(yc / o->dy); }
  else if (c == 1) { *i0 = (int)(xc / o->dx); *j0 = (int)(yc / o->dy - 0.5); }
  else { *i0 = (int)(xc / o->dx); *j0 = (int)(yc / o->dy); }
}
static inline void ibm_pos(orc_t* o, int c, int i, int j, int k, double* x, double* y, double* z) {
  *x = ((double)i - 0.5) * o->dx; *y = ((double)j - 0.5) * o->dy; *z = ((double)k - 0.5) * o->dz;
  if (c == 0) *x = *x + 0.5 * o->dx;
  else if (c == 1) *y = *y + 0.5 * o->dy;
  else *z = *z + 0.5 * o->dz;
}

/* The spreading weight per marker, also the weight of the drag sum:
 * dX^3 (fenep.f90:2125-2162) unless params.ibm_weight > 0 (smac.f90:1428
 * spreads with its dV). */
static double spread_weight(const orc_t* o) {
  return o->p.ibm_weight > 0.0 ? o->p.ibm_weight : (o->dx * o->dx) * o->dx;
}

/* ibm_Helmholtz, fenep.f90:2057-2193, with the marker forces kept as
 * computed (fenep_mpi.f90:1845-1853; fenep.f90:2121-2123 overwrites them
 * with 1 for debugging).  Interpolation weight dX^3, spreading weight
 * spread_weight(). */
static void ibm_helmholtz(orc_t* o) {
  const int nc = o->nc, nl = o->nl, nz = o->nz;
  const double dx3 = (o->dx * o->dx) * o->dx, dt = o->p.dt;
  real_t* Fm[3] = {o->fxc, o->fyc, o->fzc};
  real_t* Bm[3] = {o->ubm, o->vbm, o->wbm};
  const double* Cm[3] = {o->uc, o->vc, o->wc};
  for (int n = 1; n <= nz; n++)
    for (int mm = 1; mm <= nl; mm++)
      for (int l = 1; l <= nc; l++) {
        long id = (l - 1) + (long)nc * ((mm - 1) + (long)nl * (n - 1));
        double xc = o->xc[id], yc = o->yc[id], zc = o->zc[id];
        for (int c = 0; c < 3; c++) {
          int i0, j0; ibm_origin(o, c, xc, yc, &i0, &j0);
          real_t acc = 0.0;
          for (int k = n - 1; k <= n + 1; k++)
            for (int j = j0; j <= j0 + 2; j++)
              for (int i = i0; i <= i0 + 2; i++) {
                double x, y, z; ibm_pos(o, c, i, j, k, &x, &y, &z);
                acc = acc + (o->Up[c][GI(o, i, j, k)] * delta3(o, x, y, z, xc, yc, zc)) * dx3;
              }
          Bm[c][id] = acc;
        }
      }
  for (long id = 0; id < o->nmk; id++)
    for (int c = 0; c < 3; c++) Fm[c][id] = Fm[c][id] + (Cm[c][id] - Bm[c][id]) / dt;
  const int nx = o->nx, ny = o->ny;
  const double sw = spread_weight(o);
  for (int c = 0; c < 3; c++) {
    real_t* ft = o->ftmp;
    memset((void*)ft, 0, sizeof(real_t) * o->gn);
    for (int n = 1; n <= nz; n++)
      for (int mm = 1; mm <= nl; mm++)
        for (int l = 1; l <= nc; l++) {
          long id = (l - 1) + (long)nc * ((mm - 1) + (long)nl * (n - 1));
          double xc = o->xc[id], yc = o->yc[id], zc = o->zc[id];
          int i0, j0; ibm_origin(o, c, xc, yc, &i0, &j0);
          for (int k = n - 1; k <= n + 1; k++)
            for (int j = j0; j <= j0 + 2; j++)
              for (int i = i0; i <= i0 + 2; i++) {
                double x, y, z; ibm_pos(o, c, i, j, k, &x, &y, &z);
                long g = GI(o, i, j, k);
                ft[g] = ft[g] + (Fm[c][id] * delta3(o, x, y, z, xc, yc, zc)) * sw;
              }
        }
    /* ibm_PBM, fenep.f90:2195-2205: fold ghost layers back additively */
    for (long k = 0; k < o->gz; k++)
      for (long j = 0; j < o->gy; j++) {
        ft[GI(o, 1, j, k)] = ft[GI(o, 1, j, k)] + ft[GI(o, nx + 1, j, k)];
        ft[GI(o, nx, j, k)] = ft[GI(o, nx, j, k)] + ft[GI(o, 0, j, k)];
      }
    for (long k = 0; k < o->gz; k++)
      for (long i = 0; i < o->gx; i++) {
        ft[GI(o, i, 1, k)] = ft[GI(o, i, 1, k)] + ft[GI(o, i, ny + 1, k)];
        ft[GI(o, i, ny, k)] = ft[GI(o, i, ny, k)] + ft[GI(o, i, 0, k)];
      }
    for (long j = 0; j < o->gy; j++)
      for (long i = 0; i < o->gx; i++) {
        ft[GI(o, i, j, 1)] = ft[GI(o, i, j, 1)] + ft[GI(o, i, j, nz + 1)];
        ft[GI(o, i, j, nz)] = ft[GI(o, i, j, nz)] + ft[GI(o, i, j, 0)];
      }
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) o->fint[c][II(o, i, j, k)] = ft[GI(o, i, j, k)];
  }
}

/* ------------------------------------------------------------------- SOR */
/* poisson, fenep.f90:848-905: lexicographic SOR, odd then even i per row */
static int sor_solve(orc_t* o, const real_t* Q, real_t* Phi) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const double dx = o->dx, dy = o->dy, dz = o->dz;
  double om = o->p.sor_omega > 0.0 ? o->p.sor_omega : 2.0 / (1.0 + sin(M_PI * dx));
  double BXM = 1.0 / (dx * dx), BXP = 1.0 / (dx * dx), BYM = 1.0 / (dy * dy), BYP = 1.0 / (dy * dy);
  double BZM = 1.0 / (dz * dz), BZP = 1.0 / (dz * dz);
  double B0 = ((((BXM + BXP) + BYM) + BYP) + BZM) + BZP;
  real_t er0 = 0.0;
  for (long m = 0; m < o->n; m++) er0 = er0 + Q[m] * Q[m];
  real_t* E = o->e;
  int itr;
  for (itr = 1; itr <= o->p.sor_itrmax; itr++) {
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int l = 1; l <= 2; l++)
          for (int i = l; i <= nx; i += 2) {
            long g = GI(o, i, j, k), m = II(o, i, j, k);
            real_t e = (((((BZM * Phi[GI(o, i, j, k - 1)] + BYM * Phi[GI(o, i, j - 1, k)]) + BXM * Phi[GI(o, i - 1, j, k)])
                          - B0 * Phi[g]) + BXP * Phi[GI(o, i + 1, j, k)]) + BYP * Phi[GI(o, i, j + 1, k)]
                        + BZP * Phi[GI(o, i, j, k + 1)]) - Q[m];
            E[m] = e;
            Phi[g] = Phi[g] + (om * e) / B0;
          }
    fill_ghosts(o, Phi, BC_P);
    real_t er = 0.0;
    for (long m = 0; m < o->n; m++) er = er + E[m] * E[m];
    if (sqrt(er / er0) < o->p.sor_eps) break;
  }
  if (itr > o->p.sor_itrmax) itr = o->p.sor_itrmax;
  return itr;
}

/* divergence RHS of fft_poisson / poisson, fenep.f90:1550-1552 */
static void div_rhs(orc_t* o, real_t* Q) {
  const real_t dt = o->p.dt;
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++)
        Q[II(o, i, j, k)] = (((-o->Up[0][GI(o, i - 1, j, k)] + o->Up[0][GI(o, i, j, k)]) / o->dx
                              + (-o->Up[1][GI(o, i, j - 1, k)] + o->Up[1][GI(o, i, j, k)]) / o->dy)
                             + (-o->Up[2][GI(o, i, j, k - 1)] + o->Up[2][GI(o, i, j, k)]) / o->dz) / dt;
}

/* ------------------------------------------------------------------ step */
static void one_step(orc_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const real_t dt = o->p.dt;
  const int first = !o->has_history;
  /* fenep.f90:2479-2497 */
  if (o->p.beta == 1.0) {
    for (int q = 0; q < 6; q++) memset((void*)o->C[q], 0, sizeof(real_t) * o->gn);
    for (int c = 0; c < 3; c++) memset((void*)o->T[c], 0, sizeof(real_t) * o->n);
  } else {
    cpm_axis(o, 0); cpm_axis(o, 1); cpm_axis(o, 2);
    cstar(o);
    lyapunov(o);
    polymer_stress(o);
  }
  convection(o);
  viscous(o);
  if (first) {
    for (int c = 0; c < 3; c++) {
      memcpy(o->A0[c], o->A[c], sizeof(real_t) * o->n);
      memcpy(o->T0[c], o->T[c], sizeof(real_t) * o->n);
      memcpy(o->B0[c], o->B[c], sizeof(real_t) * o->n);
    }
  }
  real_t* Q = o->q;
  if (o->p.method == FP_METHOD_EXPLICIT || o->p.method == FP_METHOD_FFT) {
    for (int c = 0; c < 3; c++) {
      predictor_rhs(o, c, Q);
      if (o->p.method == FP_METHOD_FFT) {
        real_t* tmp = zalloc(o->n);
        solve_grid(o, c, Q, tmp);
        interior_to_ghosted(o, tmp, o->Up[c]);
        free(tmp);
      } else {
        interior_to_ghosted(o, Q, o->Up[c]);
      }
      fill_ghosts(o, o->Up[c], BCK[c]);
    }
  } else {
    for (int c = 0; c < 3; c++) {
      predictor_rhs(o, c, Q);
      interior_to_ghosted(o, Q, o->Ua[c]);
      fill_ghosts(o, o->Ua[c], BCK[c]);
      memcpy(o->Up[c], o->Ua[c], sizeof(real_t) * o->gn);
    }
    for (long id = 0; id < o->nmk; id++) { o->fxc[id] = 0.0; o->fyc[id] = 0.0; o->fzc[id] = 0.0; }
    ibm_helmholtz(o);
    /* ibm_predict, fenep.f90:2226-2256 */
    real_t* tmp = zalloc(o->n);
    for (int c = 0; c < 3; c++) {
      for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
          for (int i = 1; i <= nx; i++) {
            long m = II(o, i, j, k);
            Q[m] = (o->Ua[c][GI(o, i, j, k)] + dt * o->fint[c][m]) - (dt * o->B[c][m]) / 2.0;
          }
      solve_grid(o, c, Q, tmp);
      interior_to_ghosted(o, tmp, o->Up[c]);
      fill_ghosts(o, o->Up[c], BCK[c]);
    }
    free(tmp);
  }
  /* histories (fenep.f90:837-845, 1521-1526, 2020-2025) */
  for (int c = 0; c < 3; c++) {
    memcpy(o->A0[c], o->A[c], sizeof(real_t) * o->n);
    memcpy(o->T0[c], o->T[c], sizeof(real_t) * o->n);
    memcpy(o->B0[c], o->B[c], sizeof(real_t) * o->n);
  }
  /* Poisson */
  div_rhs(o, Q);
  if (o->p.poisson_variant == FP_POISSON_SOR) {
    o->sor_itr = (double)sor_solve(o, Q, o->Phi);
  } else {
    real_t* tmp = zalloc(o->n);
    solve_grid(o, 3, Q, tmp);
    interior_to_ghosted(o, tmp, o->Phi);
    free(tmp);
    fill_ghosts(o, o->Phi, BC_P);
    o->sor_itr = -1.0;
  }
  /* march / fft_march (fenep.f90:907-946, 1562-1604) */
  const real_t coef = ((dt * o->beta) / o->p.re) / 2.0;
  const real_t dx = o->dx, dy = o->dy, dz = o->dz;
  real_t* Ph = o->Phi;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        long g = GI(o, i, j, k);
        o->U[g] = o->Up[0][g] - (dt * (-Ph[g] + Ph[GI(o, i + 1, j, k)])) / dx;
        o->V[g] = o->Up[1][g] - (dt * (-Ph[g] + Ph[GI(o, i, j + 1, k)])) / dy;
        o->W[g] = o->Up[2][g] - (dt * (-Ph[g] + Ph[GI(o, i, j, k + 1)])) / dz;
        if (o->p.method == FP_METHOD_EXPLICIT) {
          o->P[g] = o->P[g] + Ph[g];
        } else {
          real_t lap = ((((Ph[GI(o, i - 1, j, k)] - 2.0 * Ph[g]) + Ph[GI(o, i + 1, j, k)]) / (dx * dx)
                         + ((Ph[GI(o, i, j - 1, k)] - 2.0 * Ph[g]) + Ph[GI(o, i, j + 1, k)]) / (dy * dy))
                        + ((Ph[GI(o, i, j, k - 1)] - 2.0 * Ph[g]) + Ph[GI(o, i, j, k + 1)]) / (dz * dz));
          o->P[g] = (o->P[g] + Ph[g]) - coef * lap;
        }
      }
  fill_ghosts(o, o->U, BC_U);
  fill_ghosts(o, o->V, BC_V);
  fill_ghosts(o, o->W, BC_W);
  fill_ghosts(o, o->P, BC_P);
  o->has_history = 1;
  o->step++;
}

/* ------------------------------------------------------------------- API */
static int pow2(int n) { return n >= 4 && (n & (n - 1)) == 0; }

void* orc_create(const fp_params* p) {
  if (!p || !pow2(p->nx) || !pow2(p->ny) || !pow2(p->nz)) return NULL;
  orc_t* o = (orc_t*)calloc(1, sizeof(orc_t));
  o->p = *p;
  o->nx = p->nx; o->ny = p->ny; o->nz = p->nz; o->nxh = p->nx / 2 + 1;
  o->gx = p->nx + 2; o->gy = p->ny + 2; o->gz = p->nz + 2;
  o->gn = o->gx * o->gy * o->gz;
  o->n = (long)p->nx * p->ny * p->nz;
  o->dx = p->lx / p->nx; o->dy = p->ly / p->ny; o->dz = p->lz / p->nz;
  o->walls = p->walls_y != 0;
  o->wi = p->wi; o->beta = p->beta; o->lp = p->lp;
  o->U = zalloc(o->gn); o->V = zalloc(o->gn); o->W = zalloc(o->gn); o->P = zalloc(o->gn); o->Phi = zalloc(o->gn);
  for (int q = 0; q < 6; q++) { o->C[q] = zalloc(o->gn); o->C0[q] = zalloc(o->gn); o->Cx[q] = zalloc(o->n); }
  for (int c = 0; c < 3; c++) {
    o->A[c] = zalloc(o->n); o->B[c] = zalloc(o->n); o->T[c] = zalloc(o->n);
    o->A0[c] = zalloc(o->n); o->B0[c] = zalloc(o->n); o->T0[c] = zalloc(o->n);
    o->Up[c] = zalloc(o->gn); o->Ua[c] = zalloc(o->gn); o->fint[c] = zalloc(o->n);
  }
  long fsz[3] = {(long)(p->nx + 1) * p->ny * p->nz, (long)p->nx * (p->ny + 1) * p->nz, (long)p->nx * p->ny * (p->nz + 1)};
  for (int d = 0; d < 3; d++)
    for (int q = 0; q < 6; q++) { o->Cp[d][q] = zalloc(fsz[d]); o->Cn[d][q] = zalloc(fsz[d]); }
  o->ftmp = zalloc(o->gn);
  o->q = zalloc(o->n);
  o->e = zalloc(o->n);
  o->twx = twiddles(p->nx); o->twy = twiddles(p->ny); o->twz = twiddles(p->nz);
  o->kx2 = modwav(p->nx, o->dx, o->nxh); o->ky2 = modwav(p->ny, o->dy, p->ny); o->kz2 = modwav(p->nz, o->dz, p->nz);
  o->sy = o->walls ? p->ny + 2 : p->ny;
  o->jo = o->walls ? 1 : 0;
  o->spec = zalloc(2 * (long)o->nxh * o->sy * p->nz);
  return o;
}

void orc_destroy(void* vo) {
  orc_t* o = (orc_t*)vo;
  if (!o) return;
  free(o->U); free(o->V); free(o->W); free(o->P); free(o->Phi);
  for (int q = 0; q < 6; q++) { free(o->C[q]); free(o->C0[q]); free(o->Cx[q]); }
  for (int c = 0; c < 3; c++) {
    free(o->A[c]); free(o->B[c]); free(o->T[c]); free(o->A0[c]); free(o->B0[c]); free(o->T0[c]);
    free(o->Up[c]); free(o->Ua[c]); free(o->fint[c]); free(o->F[c]);
  }
  for (int d = 0; d < 3; d++) for (int q = 0; q < 6; q++) { free(o->Cp[d][q]); free(o->Cn[d][q]); }
  free(o->ftmp); free(o->q); free(o->e); free(o->twx); free(o->twy); free(o->twz);
  free(o->kx2); free(o->ky2); free(o->kz2); free(o->spec);
  free(o->xc); free(o->yc); free(o->zc); free(o->uc); free(o->vc); free(o->wc);
  free(o->fxc); free(o->fyc); free(o->fzc); free(o->ubm); free(o->vbm); free(o->wbm);
  free(o);
}

/* state in and out of the restatement's scalar (identity in the FP64 build) */
static void load_r(real_t* d, const double* src, long n) { for (long i = 0; i < n; i++) d[i] = src[i]; }
static void store_r(double* d, const real_t* src, long n) { for (long i = 0; i < n; i++) d[i] = rv(src[i]); }

int orc_set_state(void* vo, const double* U, const double* V, const double* W, const double* P,
                  const double* C6, const double* hist, int has_history) {
  orc_t* o = (orc_t*)vo;
  load_r(o->U, U, o->gn);
  load_r(o->V, V, o->gn);
  load_r(o->W, W, o->gn);
  load_r(o->P, P, o->gn);
  for (long g = 0; g < o->gn; g++)
    for (int q = 0; q < 6; q++) o->C[q][g] = C6[6 * g + q];
  fill_ghosts(o, o->U, BC_U); fill_ghosts(o, o->V, BC_V); fill_ghosts(o, o->W, BC_W);
  fill_ghosts(o, o->P, BC_P);
  for (int q = 0; q < 6; q++) fill_ghosts(o, o->C[q], BC_C);
  if (hist) {
    for (int c = 0; c < 3; c++) {
      load_r(o->A0[c], hist + (long)c * o->n, o->n);
      load_r(o->T0[c], hist + (long)(3 + c) * o->n, o->n);
      load_r(o->B0[c], hist + (long)(6 + c) * o->n, o->n);
    }
  }
  o->has_history = has_history && hist;
  memset((void*)o->Phi, 0, sizeof(real_t) * o->gn);
  return 0;
}

int orc_get_state(void* vo, double* U, double* V, double* W, double* P, double* C6, double* hist) {
  orc_t* o = (orc_t*)vo;
  if (U) store_r(U, o->U, o->gn);
  if (V) store_r(V, o->V, o->gn);
  if (W) store_r(W, o->W, o->gn);
  if (P) store_r(P, o->P, o->gn);
  if (C6)
    for (long g = 0; g < o->gn; g++)
      for (int q = 0; q < 6; q++) C6[6 * g + q] = rv(o->C[q][g]);
  if (hist)
    for (int c = 0; c < 3; c++) {
      store_r(hist + (long)c * o->n, o->A0[c], o->n);
      store_r(hist + (long)(3 + c) * o->n, o->T0[c], o->n);
      store_r(hist + (long)(6 + c) * o->n, o->B0[c], o->n);
    }
  return 0;
}

int orc_get_phi(void* vo, double* Phi) {
  orc_t* o = (orc_t*)vo;
  store_r(Phi, o->Phi, o->gn);
  return 0;
}

int orc_set_body_force(void* vo, const double* fx, const double* fy, const double* fz) {
  orc_t* o = (orc_t*)vo;
  const double* f[3] = {fx, fy, fz};
  for (int c = 0; c < 3; c++) {
    free(o->F[c]); o->F[c] = NULL;
    if (f[c]) { o->F[c] = zalloc(o->n); load_r(o->F[c], f[c], o->n); }
  }
  return 0;
}

int orc_set_markers(void* vo, int nc, int nl, const double* xc, const double* yc, const double* zc,
                    const double* uc, const double* vc, const double* wc) {
  orc_t* o = (orc_t*)vo;
  o->nc = nc; o->nl = nl; o->nmk = nc * nl * o->nz;
  long n = o->nmk;
  double** pos[6] = {&o->xc, &o->yc, &o->zc, &o->uc, &o->vc, &o->wc};
  const double* src[6] = {xc, yc, zc, uc, vc, wc};
  for (int a = 0; a < 6; a++) { free(*pos[a]); *pos[a] = dzalloc(n); memcpy(*pos[a], src[a], sizeof(double) * n); }
  real_t** frc[6] = {&o->fxc, &o->fyc, &o->fzc, &o->ubm, &o->vbm, &o->wbm};
  for (int a = 0; a < 6; a++) { free(*frc[a]); *frc[a] = zalloc(n); }
  return 0;
}

int orc_get_marker_forces(void* vo, double* fx, double* fy, double* fz) {
  orc_t* o = (orc_t*)vo;
  store_r(fx, o->fxc, o->nmk);
  store_r(fy, o->fyc, o->nmk);
  store_r(fz, o->fzc, o->nmk);
  return 0;
}

int orc_step(void* vo, int nsteps) {
  orc_t* o = (orc_t*)vo;
  if (o->p.method == FP_METHOD_IBM && o->nmk == 0) return FP_E_STATE;
  for (int s = 0; s < nsteps; s++) one_step(o);
  return 0;
}

#ifndef ORC_CS
int orc_diagnostics(void* vo, fp_diag* d) {
  orc_t* o = (orc_t*)vo;
  memset(d, 0, sizeof(*d));
  real_t su = 0, sv = 0, sw = 0, umax = -INFINITY, vmax = -INFINITY, wmax = -INFINITY;
  real_t tmin = INFINITY, tmax = -INFINITY, tsum = 0, dmax = 0;
  int nan = 0;
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++) {
        long g = GI(o, i, j, k);
        real_t ur = ((o->U[g] + o->U[GI(o, i - 1, j, k)]) * 1.0) / 2.0;
        real_t vr = ((o->V[g] + o->V[GI(o, i, j - 1, k)]) * 1.0) / 2.0;
        real_t wr = ((o->W[g] + o->W[GI(o, i, j, k - 1)]) * 1.0) / 2.0;
        su += ur * ur; sv += vr * vr; sw += wr * wr;
        if (isnan(o->U[g])) nan = 1;
        if (o->U[g] > umax) umax = o->U[g];
        if (o->V[g] > vmax) vmax = o->V[g];
        if (o->W[g] > wmax) wmax = o->W[g];
        real_t tr = (o->C[0][g] + o->C[3][g]) + o->C[5][g];
        if (tr < tmin) tmin = tr;
        if (tr > tmax) tmax = tr;
        tsum += tr;
        real_t dv = ((o->U[g] - o->U[GI(o, i - 1, j, k)]) / o->dx + (o->V[g] - o->V[GI(o, i, j - 1, k)]) / o->dy)
                    + (o->W[g] - o->W[GI(o, i, j, k - 1)]) / o->dz;
        if (fabs(dv) > dmax) dmax = fabs(dv);
      }
  real_t N = (real_t)o->n;
  d->ke = ((su + sv + sw) / 2.0) / N;
  real_t c1 = umax * o->p.dt / o->dx, c2 = vmax * o->p.dt / o->dy, c3 = wmax * o->p.dt / o->dz;
  d->cfl = fmax(c1, fmax(c2, c3)) * 6.0;
  d->trc_min = tmin; d->trc_max = tmax; d->trc_mean = tsum / N;
  real_t ss = 0;
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++) {
        long g = GI(o, i, j, k);
        real_t tr = (o->C[0][g] + o->C[3][g]) + o->C[5][g];
        ss += (tr - d->trc_mean) * (tr - d->trc_mean);
      }
  d->trc_std = sqrt(ss / N);
  d->div_max = dmax;
  const double wgt = spread_weight(o);
  real_t dr = 0, lf = 0;
  for (long id = 0; id < o->nmk; id++) { dr += o->fxc[id]; lf += o->fyc[id]; }
  d->drag = rv(-(dr * wgt)); d->lift = rv(-(lf * wgt));
  d->newton_itr_mean = o->newton_itr_total / N;
  d->sor_itr = o->sor_itr;
  d->n_dc = 0;
  d->nan_flag = nan;
  d->step = o->step;
  return 0;
}

/* Time-series statistics of all_time_dif (fenep_mpi.f90:2200-2348) without
 * a subtracted mean (the 'same' file) and with unit reference scales
 * (U_C = L_C = 1, nu = 1/Re): out = step, step*dt, K, K_staggered, lambda,
 * Re_lambda, epsilon, eta, tau_eta, v_eta, Ep, trC mean, trC std. */
int orc_time_stats(void* vo, real_t* out) {
  orc_t* o = (orc_t*)vo;
  const real_t dx = o->dx, dy = o->dy, dz = o->dz;
  real_t ur2 = 0, vr2 = 0, wr2 = 0, ug2 = 0, vg2 = 0, wg2 = 0, s2 = 0, kst = 0, ep = 0, tr = 0;
  const real_t L2 = o->p.lp * o->p.lp;
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++) {
        const real_t* U = o->U;
        const real_t* V = o->V;
        const real_t* W = o->W;
        long g = GI(o, i, j, k);
        real_t ur = (U[g] + U[GI(o, i - 1, j, k)]) / 2.0;
        real_t vr = (V[g] + V[GI(o, i, j - 1, k)]) / 2.0;
        real_t wr = (W[g] + W[GI(o, i, j, k - 1)]) / 2.0;
        ur2 += ur * ur; vr2 += vr * vr; wr2 += wr * wr;
        real_t ug = (U[g] - U[GI(o, i - 1, j, k)]) / dx;
        real_t vg = (V[g] - V[GI(o, i, j - 1, k)]) / dy;
        real_t wg = (W[g] - W[GI(o, i, j, k - 1)]) / dz;
        ug2 += ug * ug; vg2 += vg * vg; wg2 += wg * wg;
        real_t D[3][3];
        D[0][0] = (U[g] - U[GI(o, i - 1, j, k)]) / dx;
        D[0][1] = (U[GI(o, i, j + 1, k)] - U[GI(o, i, j - 1, k)] + U[GI(o, i - 1, j + 1, k)] - U[GI(o, i - 1, j - 1, k)]) / (4 * dy);
        D[0][2] = (U[GI(o, i, j, k + 1)] - U[GI(o, i, j, k - 1)] + U[GI(o, i - 1, j, k + 1)] - U[GI(o, i - 1, j, k - 1)]) / (4 * dz);
        D[1][0] = (V[GI(o, i + 1, j, k)] - V[GI(o, i - 1, j, k)] + V[GI(o, i + 1, j - 1, k)] - V[GI(o, i - 1, j - 1, k)]) / (4 * dx);
        D[1][1] = (V[g] - V[GI(o, i, j - 1, k)]) / dy;
        D[1][2] = (V[GI(o, i, j, k + 1)] - V[GI(o, i, j, k - 1)] + V[GI(o, i, j - 1, k + 1)] - V[GI(o, i, j - 1, k - 1)]) / (4 * dz);
        D[2][0] = (W[GI(o, i + 1, j, k)] - W[GI(o, i - 1, j, k)] + W[GI(o, i + 1, j, k - 1)] - W[GI(o, i - 1, j, k - 1)]) / (4 * dx);
        D[2][1] = (W[GI(o, i, j + 1, k)] - W[GI(o, i, j - 1, k)] + W[GI(o, i, j + 1, k - 1)] - W[GI(o, i, j - 1, k - 1)]) / (4 * dy);
        D[2][2] = (W[g] - W[GI(o, i, j, k - 1)]) / dz;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            real_t sab = D[a][b] + D[b][a];
            s2 += sab * sab;
          }
        kst += U[g] * U[g] + V[g] * V[g] + W[g] * W[g];
        real_t t = (o->C[0][g] + o->C[3][g]) + o->C[5][g];
        tr += t;
        ep += log((L2 - 3.0) / (L2 - t));
      }
  const real_t N = (real_t)o->n, nu = 1.0 / o->p.re;
  const real_t urms = ur2 / N, vrms = vr2 / N, wrms = wr2 / N;
  const real_t ugr = ug2 / N, vgr = vg2 / N, wgr = wg2 / N;
  const real_t all_rms = (urms + vrms + wrms) / 3.0, all_grad = (ugr + vgr + wgr) / 3.0;
  const real_t eps = nu / 2.0 * s2 / N;
  const real_t lambda = sqrt(all_rms / all_grad);
  real_t mean = tr / N, var = 0;
  for (int k = 1; k <= o->nz; k++)
    for (int j = 1; j <= o->ny; j++)
      for (int i = 1; i <= o->nx; i++) {
        long g = GI(o, i, j, k);
        real_t t = (o->C[0][g] + o->C[3][g]) + o->C[5][g];
        var += (t - mean) * (t - mean);
      }
  out[0] = (real_t)o->step;
  out[1] = o->step * o->p.dt;
  out[2] = (urms + vrms + wrms) / 2.0;
  out[3] = kst / (2.0 * N);
  out[4] = lambda;
  out[5] = sqrt(all_rms) * lambda / nu;
  out[6] = eps;
  out[7] = pow(nu * nu * nu / eps, 0.25);
  out[8] = sqrt(nu / eps);
  out[9] = pow(nu * eps, 0.25);
  out[10] = 0.5 * (1.0 - o->p.beta) / o->p.re / o->p.wi * (L2 - 3.0) * ep / N;
  out[11] = mean;
  out[12] = sqrt(var / N);
  return 0;
}

int orc_selection_counts(void* vo, double* c4) {
  orc_t* o = (orc_t*)vo;
  for (int l = 0; l < 4; l++) c4[l] = o->sel_count[l];
  return 0;
}

int orc_eval_operator(void* vo, int which, real_t* ox, real_t* oy, real_t* oz) {
  orc_t* o = (orc_t*)vo;
  real_t** src;
  if (which == 0) { convection(o); src = o->A; }
  else if (which == 1) { viscous(o); src = o->B; }
  else { polymer_stress(o); src = o->T; }
  memcpy(ox, src[0], sizeof(real_t) * o->n);
  memcpy(oy, src[1], sizeof(real_t) * o->n);
  memcpy(oz, src[2], sizeof(real_t) * o->n);
  return 0;
}

int orc_solve(void* vo, int kind, const real_t* q, real_t* x) {
  orc_t* o = (orc_t*)vo;
  if (kind < 0 || kind > 3) return FP_E_ARG;
  solve_grid(o, kind, q, x);
  return 0;
}

int orc_sor(void* vo, const real_t* q, real_t* phi) {
  orc_t* o = (orc_t*)vo;
  fill_ghosts(o, phi, BC_P);
  return sor_solve(o, q, phi);
}

/* ibm_Helmholtz alone on a given ghosted velocity (u*,v*,w*): returns the
 * spread force (interior), marker forces and interpolated marker velocities
 * (for the adjointness / conservation tests). */
int orc_ibm_apply(void* vo, const real_t* up3, real_t* fint3, real_t* fm3, real_t* ub3) {
  orc_t* o = (orc_t*)vo;
  if (!o->nmk) return FP_E_STATE;
  for (int c = 0; c < 3; c++) memcpy(o->Up[c], up3 + (long)c * o->gn, sizeof(real_t) * o->gn);
  for (long id = 0; id < o->nmk; id++) { o->fxc[id] = 0.0; o->fyc[id] = 0.0; o->fzc[id] = 0.0; }
  ibm_helmholtz(o);
  real_t* Fm[3] = {o->fxc, o->fyc, o->fzc};
  real_t* Bm[3] = {o->ubm, o->vbm, o->wbm};
  for (int c = 0; c < 3; c++) {
    memcpy(fint3 + (long)c * o->n, o->fint[c], sizeof(real_t) * o->n);
    memcpy(fm3 + (long)c * o->nmk, Fm[c], sizeof(real_t) * o->nmk);
    memcpy(ub3 + (long)c * o->nmk, Bm[c], sizeof(real_t) * o->nmk);
  }
  return 0;
}

/* ---- unit-level entry points (pinning tests of the FP64 build) */
double orc_vieta(const double* a) { return vieta_r(a); }
int orc_select(const double* cm, const double* c, const double* cp, double* cpo, double* cno) {
  return select_r(cm, c, cp, cpo, cno);
}
void orc_gauss6(const double* a, const double* b, double* x) { gauss6_r(a, b, x); }
int orc_newton_cell(const double* A0, const double* b, const double* x0, double dt, double wi, double lp,
                    double dc, double eps, int itrmax, double* x) {
  return newton_cell_r(A0, b, x0, dt, wi, lp, dc, eps, itrmax, x);
}
double orc_roma_weight(double r) { return roma_weight(r); }
void orc_tdma(int n, const double* a, const double* b, const double* c, const double* dre, const double* dim,
              double* xre, double* xim) {
  tdma_r(n, a, b, c, dre, dim, xre, xim);
}
void orc_fft(int n, double* data, int inverse) { fft_r(n, data, inverse); }

#else  /* ORC_CS: the complex-step build */

/* Seed one parameter: Wi (which 0), beta (1) or L^2 (2) takes imaginary part
 * h; the others none.  Lp = sqrt(L^2) carries h / (2 Lp), so Lp*Lp has
 * imaginary part h. */
int orc_cs_seed_param(void* vo, int which, double h) {
  orc_t* o = (orc_t*)vo;
  o->wi = cs_t(o->p.wi, which == 0 ? h : 0.0);
  o->beta = cs_t(o->p.beta, which == 1 ? h : 0.0);
  o->lp = cs_t(o->p.lp, which == 2 ? h / (2.0 * o->p.lp) : 0.0);
  return 0;
}

/* Seed the body-force field along a direction dF (3 interior planes): F
 * takes imaginary part h dF.  The field must have been set. */
int orc_cs_seed_force(void* vo, const double* dF, double h) {
  orc_t* o = (orc_t*)vo;
  for (int c = 0; c < 3; c++) {
    if (!o->F[c]) return FP_E_STATE;
    for (long m = 0; m < o->n; m++) o->F[c][m].im = h * dF[(long)c * o->n + m];
  }
  return 0;
}

/* Drag of the last step, D = -sum_markers Fxc W (W = spread_weight()):
 * out[0] = D, out[1] = its imaginary part. */
int orc_cs_drag(void* vo, double* out2) {
  orc_t* o = (orc_t*)vo;
  const double sw = spread_weight(o);
  real_t dr = 0.0;
  for (long id = 0; id < o->nmk; id++) dr += o->fxc[id];
  const real_t d = -(dr * sw);
  out2[0] = d.re;
  out2[1] = d.im;
  return 0;
}
#endif
