// Device helpers shared by the FP64 kernels: boundary-rule accessors and the
// per-cell FENE-P algebra.  Every expression keeps the left-to-right order of
// the Fortran statement it restates (the file that includes this is compiled
// without fused multiply-add for the FP64 parity path).
#pragma once
#include "fp_internal.h"

namespace fp {

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Value of a scalar field at (i,j,k), |offset| <= 1 outside the interior:
// periodic copies (PBM, fenep.f90:147-155) in x and z; in y either periodic or
// the wall rules U_kabe/V_kabe/P_kabe/C_kabe (fenep_2d.f90:1098-1143, NA=0,
// NB=NY, Uwall=0).
template <bool WALLS, typename T>
__device__ __forceinline__ T ldbw(const T* __restrict__ A, const Grid& g, int i, int j, int k, int kind) {
  // cell indices by shifts (power-of-two axes, n <= 2^30 cells)
  i = wrapi(i, g.nx);
  k = wrapi(k, g.nz);
  const int lp = g.lnx + g.lny;
  if constexpr (!WALLS) {
    j = wrapi(j, g.ny);
    return A[(unsigned)(i + (j << g.lnx) + (k << lp))];
  } else {
    // y-walls (U_kabe .. C_kabe, fenep_2d.f90:1098-1143), without branches
    // so that a caller's loads can be issued together: every rule is one
    // load of the nearest interior row, negated or replaced by zero, or (C)
    // the linear extrapolation from the two nearest rows
    const T* col = A + (unsigned)(i + (k << lp));
    const int ny = g.ny;
    const bool lo = j < 0, hi = j >= ny, ghost = lo || hi;
    const int jc = lo ? 0 : (hi ? ny - 1 : j);
    if (kind == BC_U || kind == BC_V || kind == BC_W || kind == BC_P) {
      const T v = col[(unsigned)(jc << g.lnx)];
      // U, W ghosts: -U_1 below, 2*0 - U_NY above (the reference's forms,
      // which differ for a zero in the sign of the result)
      if (kind == BC_U || kind == BC_W) return lo ? -v : (hi ? T(2) * T(0) - v : v);
      const bool zero = kind == BC_V && (ghost || j == ny - 1);
      return zero ? T(0) : v;
    }
    const int j2 = lo ? 1 : (hi ? ny - 2 : jc);
    const T a = col[(unsigned)(jc << g.lnx)], b = col[(unsigned)(j2 << g.lnx)];
    return ghost ? T(2) * a - b : a;
  }
}
// the same with the wall flag read at run time
template <typename T>
__device__ __forceinline__ T ldb(const T* __restrict__ A, const Grid& g, int i, int j, int k, int kind) {
  return g.walls ? ldbw<true>(A, g, i, j, k, kind) : ldbw<false>(A, g, i, j, k, kind);
}

// Workgroups of a launch are handed to the 8 XCDs of the chip in turn, and
// each XCD has its own L2.  Stencil kernels renumber their workgroups so that
// each XCD walks one contiguous range of cells: the neighbouring y rows and z
// planes a stencil reads are then fetched into one L2 instead of into all
// eight.  The renumbering is a bijection for any number of workgroups.
__device__ __forceinline__ long xcd_block() {
  const long nb = gridDim.x, b = blockIdx.x;
  const long q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}
__device__ __forceinline__ long xcd_cell() { return xcd_block() * (long)blockDim.x + threadIdx.x; }

// Zero n 32-bit words.  A kernel rather than a memset, so that a caller may
// capture the step in a HIP graph.
static __global__ __launch_bounds__(256) void k_zero_words(long n, unsigned* __restrict__ p) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0u;
}
static inline void zero_words(void* p, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, (unsigned*)p);
}

// (i, j, k) of interior cell m (all dimensions are powers of two)
__device__ __forceinline__ void cell_ijk(const Grid& g, long m, int& i, int& j, int& k) {
  const unsigned u = (unsigned)m;
  i = (int)(u & (unsigned)(g.nx - 1));
  j = (int)((u >> g.lnx) & (unsigned)(g.ny - 1));
  k = (int)(u >> (g.lnx + g.lny));
}

__device__ __forceinline__ long cidx(const Grid& g, int i, int j, int k) {
  return (long)(unsigned)(i + (j << g.lnx) + (k << (g.lnx + g.lny)));
}

// Band order of the cells of a stencil kernel: x fastest, then the 16
// rows of a y-band, then z, then the bands.  With the XCD-contiguous block
// numbering each XCD sweeps whole bands plane by plane, so the planes k-2..k+2
// a stencil reads stay in that XCD's L2 (a plane-major walk keeps five full
// planes of every field live, more than one L2 holds).
__device__ __forceinline__ void band_cell(const Grid& g, long L, long& m, int& i, int& j, int& k) {
  const int lb = g.lny < 4 ? g.lny : 4;
  i = (int)(L & (g.nx - 1));
  long t = L >> g.lnx;
  const int jj = (int)(t & ((1 << lb) - 1));
  t >>= lb;
  const int lnz = __ffs(g.nz) - 1;
  k = (int)(t & (g.nz - 1));
  j = ((int)(t >> lnz) << lb) + jj;
  m = cidx(g, i, j, k);
}


// ---- FENE-P algebra (FP64) ------------------------------------------------
// Vieta, fenep.f90:425-444
// The reference formulas below are evaluated exactly as written, also in
// translation units compiled with contraction on (the mixed path re-decides
// selections with them and must reproduce the FP64 path's bits).
__device__ __forceinline__ double vieta_d(const double* a) {
#pragma clang fp contract(off)
  double a2 = -((a[0] + a[3]) + a[5]);
  double a1 = (((((a[0] * a[3]) + (a[3] * a[5])) + (a[5] * a[0])) - (a[4] * a[4])) - (a[2] * a[2])) - (a[1] * a[1]);
  double a0 = -((((((a[0] * a[3]) * a[5] + (a[1] * a[4]) * a[2]) + (a[1] * a[4]) * a[2]) - (a[0] * a[4]) * a[4]) -
                  (a[1] * a[1]) * a[5]) - (a[2] * a[3]) * a[2]);
  double p = a1 - (a2 * a2) / 3.0;
  double q = (a0 - (a1 * a2) / 3.0) + (2.0 * ((a2 * a2) * a2)) / 27.0;
  double q2 = q / 2.0, p3 = p / 3.0;
  double t = q2 * q2 + (p3 * p3) * p3;
  // evaluated without a branch (keeps the six tests of an axis in one
  // scheduling region); the t >= 0 sentinel is selected afterwards
  double arg = ((3.0 * q) / (2.0 * p)) * sqrt(-(3.0 / p));
  double ang = acos(arg) / 3.0 + (2.0 * 3.14159265358979323846) / 3.0;
  double x = (2.0 * sqrt(-(p / 3.0))) * cos(ang) - a2 / 3.0;
  return t >= 0.0 ? -100.0 : x;
}

__device__ __forceinline__ double minval2_d(double a, double b) {
  if (isnan(a)) return b;
  if (isnan(b)) return a;
  return (b < a) ? b : a;
}

// candidate face values of one cell (fenep.f90:179-184): left face (Cp) and
// right face (Cn) for candidate l in 0..2; l == 3 is the zero slope.
__device__ __forceinline__ double cand_p(int l, double cm, double c, double cp) {
#pragma clang fp contract(off)
  if (l == 0) return (c * 3.0) / 2.0 - cp / 2.0;
  if (l == 1) return cm / 2.0 + c / 2.0;
  if (l == 2) return (cm / 4.0 + c) - cp / 4.0;
  return c;
}
__device__ __forceinline__ double cand_n(int l, double cm, double c, double cp) {
#pragma clang fp contract(off)
  if (l == 0) return c / 2.0 + cp / 2.0;
  if (l == 1) return -(cm / 2.0) + (c * 3.0) / 2.0;
  if (l == 2) return (-(cm / 4.0) + c) + cp / 4.0;
  return c;
}

// selection of fenep.f90:186-221 (maxloc over {min eig of candidate l} and 0),
// returns 0..3 (3 = reference index 4, zero slope)
__device__ __forceinline__ int select_d(const double* cm, const double* c, const double* cp) {
  double best = 0.0;
  int idx = -1;
#pragma unroll
  for (int l = 0; l < 3; l++) {
    double ap[6], an[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
      ap[q] = cand_p(l, cm[q], c[q], cp[q]);
      an[q] = cand_n(l, cm[q], c[q], cp[q]);
    }
    double v = minval2_d(vieta_d(ap), vieta_d(an));
    if (!isnan(v) && (idx < 0 || v > best)) { idx = l; best = v; }
  }
  if (idx < 0 || 0.0 > best) idx = 3;
  return idx;
}

// Lyapunov_func, fenep.f90:588-605 (A column-major, already transposed)
__device__ __forceinline__ void lyap_func_d(const double* A0, const double* b, const double* x, double dt,
                                            double wi, double lp, double* l) {
  double f = (lp * lp - 3.0) / (((lp * lp - x[0]) - x[3]) - x[5]);
  double dg = (1.0 + (f * dt) / wi) / 2.0;
  double a11 = A0[0] + dg, a22 = A0[4] + dg, a33 = A0[8] + dg;
  double a21 = A0[1], a31 = A0[2], a12 = A0[3], a32 = A0[5], a13 = A0[6], a23 = A0[7];
  l[0] = ((((a11 + a11) * x[0] + a12 * x[1]) + a13 * x[2]) + a12 * x[1]) + a13 * x[2] - b[0];
  l[1] = ((((a21 * x[0] + (a22 + a11) * x[1]) + a23 * x[2]) + a12 * x[3]) + a13 * x[4]) - b[1];
  l[2] = ((((a31 * x[0] + a32 * x[1]) + (a33 + a11) * x[2]) + a12 * x[4]) + a13 * x[5]) - b[2];
  l[3] = ((((a21 * x[1] + a21 * x[1]) + (a22 + a22) * x[3]) + a23 * x[4]) + a23 * x[4]) - b[3];
  l[4] = ((((a21 * x[2] + a31 * x[1]) + a32 * x[3]) + (a33 + a22) * x[4]) + a23 * x[5]) - b[4];
  l[5] = ((((a31 * x[2] + a32 * x[4]) + a31 * x[2]) + a32 * x[4]) + (a33 + a33) * x[5]) - b[5];
}

// gauss, fenep.f90:608-643, fully unrolled so that row swaps are selects.
__device__ __forceinline__ void gauss6_d(double (&Ah)[6][7], double* x) {
  bool zero = false;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (!zero) {
      int mi = k;
      double ma = fabs(Ah[k][k]);
#pragma unroll
      for (int r = k + 1; r < 6; r++) {
        double v = fabs(Ah[r][k]);
        if (v > ma) { ma = v; mi = r; }
      }
      if (ma == 0.0) {
        zero = true;
      } else {
#pragma unroll
        for (int r = k + 1; r < 6; r++) {
          if (r == mi) {
#pragma unroll
            for (int c = k; c < 7; c++) { double t = Ah[k][c]; Ah[k][c] = Ah[r][c]; Ah[r][c] = t; }
          }
        }
        double piv = Ah[k][k];
#pragma unroll
        for (int c = k + 1; c < 7; c++) Ah[k][c] = Ah[k][c] / piv;
        Ah[k][k] = 1.0;
#pragma unroll
        for (int r = 0; r < 6; r++) {
          if (r != k) {
            double f = Ah[r][k];
#pragma unroll
            for (int c = k + 1; c < 7; c++) Ah[r][c] = Ah[r][c] - f * Ah[k][c];
            Ah[r][k] = 0.0;
          }
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 6; r++) x[r] = zero ? 0.0 : Ah[r][6];
}

// Newton of Lyapunov (fenep.f90:554-578) with the central-difference Jacobian.
// Returns the iterations taken, itrmax + 1 when the tolerance was not met.
__device__ __forceinline__ int newton_d(const double* A0, const double* b, double* x, double dt, double wi,
                                        double lp, double dc, double eps, int itrmax) {
  int itr;
  for (itr = 1; itr <= itrmax; itr++) {
    double Ah[6][7];
#pragma unroll
    for (int col = 0; col < 6; col++) {
      double x1[6], x2[6], l1[6], l2[6];
#pragma unroll
      for (int m = 0; m < 6; m++) { x1[m] = x[m]; x2[m] = x[m]; }
      x1[col] = x1[col] + dc;
      x2[col] = x2[col] - dc;
      lyap_func_d(A0, b, x1, dt, wi, lp, l1);
      lyap_func_d(A0, b, x2, dt, wi, lp, l2);
#pragma unroll
      for (int row = 0; row < 6; row++) Ah[row][col] = (l1[row] - l2[row]) / (2.0 * dc);
    }
    double l[6], r[6];
    lyap_func_d(A0, b, x, dt, wi, lp, l);
#pragma unroll
    for (int m = 0; m < 6; m++) Ah[m][6] = l[m];
    gauss6_d(Ah, r);
#pragma unroll
    for (int m = 0; m < 6; m++) x[m] = x[m] - r[m];
    double sr = 0.0, sx = 0.0;
#pragma unroll
    for (int m = 0; m < 6; m++) sr = sr + r[m] * r[m];
#pragma unroll
    for (int m = 0; m < 6; m++) sx = sx + x[m] * x[m];
    if (sqrt(sr / sx) < eps) break;
  }
  return itr;
}

__device__ __forceinline__ double peterlin_d(double lp, double c0, double c3, double c5) {
  return (lp * lp - 3.0) / (((lp * lp - c0) - c3) - c5);
}

// ---- IBM (FP64) -------------------------------------------------------------
// 3-point kernel of Roma, Peskin & Berger (delta/weight, fenep.f90:2039-2055)
__device__ __forceinline__ double roma_d(double r) {
  double a = fabs(r);
  if (a < 1.5 && a >= 0.5) return ((5.0 - 3.0 * a) - sqrt(-(3.0 * ((1.0 - a) * (1.0 - a))) + 1.0)) / 6.0;
  if (a < 0.5) return (1.0 + sqrt(-(3.0 * (r * r)) + 1.0)) / 3.0;
  return 0.0;
}

// delta (fenep.f90:2039-2043) * dX^3 between marker (xc,yc,zc) and the
// sample of component c at grid index (ii,jj,kk) (1-based, ghosts included).
__device__ __forceinline__ double ibm_weight(const Grid& g, int c, int ii, int jj, int kk, double xc, double yc, double zc) {
  double x = ((double)ii - 0.5) * g.dx, y = ((double)jj - 0.5) * g.dy, z = ((double)kk - 0.5) * g.dz;
  if (c == 0) x = x + 0.5 * g.dx;
  else if (c == 1) y = y + 0.5 * g.dy;
  else z = z + 0.5 * g.dz;
  double d = ((roma_d((x - xc) / g.dx) * roma_d((y - yc) / g.dy)) * roma_d((z - zc) / g.dz)) / ((g.dx * g.dy) * g.dz);
  return d;
}
__device__ __forceinline__ void ibm_origin_d(const Grid& g, int c, double xc, double yc, int& i0, int& j0) {
  if (c == 0) { i0 = (int)(xc / g.dx - 0.5); j0 = (int)(yc / g.dy); }
  else if (c == 1) { i0 = (int)(xc / g.dx); j0 = (int)(yc / g.dy - 0.5); }
  else { i0 = (int)(xc / g.dx); j0 = (int)(yc / g.dy); }
}

// Deterministic two-level reductions: each block stores its partial sum,
// one block adds the partials (no contended atomics on a single address).
__device__ __forceinline__ double block_sum256(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  return sh[0];
}

// out[0] = sum(part[0..n)) (accumulate = 0) or out[0] += sum (accumulate = 1)
static __global__ void __launch_bounds__(1024) k_sum_partials(const double* __restrict__ part, int n, double* __restrict__ out,
                                                       int accumulate) {
  __shared__ double sh[1024];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) v += part[i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = accumulate ? out[0] + sh[0] : sh[0];
}

}  // namespace fp
