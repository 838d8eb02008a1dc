// Recording and discrete adjoint of the FP64 step on periodic grids (FFT
// Poisson, Crank-Nicolson predictor with or without immersed boundary).
//
// Forward (fp_adj_record): the step runs unchanged; every `every` steps the
// full state is checkpointed, the C+- selections of every step and the drag
// D = -sum_markers Fxc dX^3 of every step are kept, and (FP64 step) the last
// segment is taped as it runs.
//
// Reverse (fp_adj_backward): segments are recomputed from their checkpoint
// with the recorded selections (a tape of the state, u* and Phi of each step)
// and traversed backwards; the segment the tape already holds is not
// recomputed.  After a mixed-precision recording the segments are recomputed
// with the mixed step, and the tape holds its states widened to FP64.  Each forward kernel has an adjoint kernel here:
//   fft_march          -> k_adj_march            (transposed gradient/Laplacian)
//   fft_poisson solve  -> the same spectral solve (symmetric operator)
//   divergence         -> k_adj_div
//   Helmholtz solves   -> the same spectral solves (symmetric)
//   IBM interp/spread  -> k_adj_ibm_markers (interp) + the spread kernels
//   predictor          -> k_adj_pred_local, k_adj_pred_u (convection in flux
//                         form, viscous, pressure), k_adj_pred_tau (polymer
//                         stress and Peterlin factor)
//   Lyapunov Newton    -> k_adj_newton: implicit-function adjoint at the
//                         converged root, J^T lambda = xbar with the exact
//                         Jacobian of the Lyapunov residual
//   Cstar              -> k_adj_cstar (face reconstructions with the frozen
//                         selection, upwind weights) and k_adj_vel_grad
// The C+- selection is a discrete choice (maxloc) and is frozen at the
// recorded run; the derivative is the one of the step with that choice.
// Gradients: d/dWi, d/dbeta, d/d(L^2) of sum_s w_s D_s (+ <bar_final, state>),
// the body-force field and the initial state.
#include "fp_device.h"
#include "fp_mixed.h"
#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <utility>
#include <vector>

namespace fp {

namespace {

inline int nblk(long n, int t) { return (int)((n + t - 1) / t); }

// periodic index of (i+di, j+dj, k+dk) for |d| <= n
__device__ __forceinline__ long pidx(const Grid& g, int i, int j, int k) {
  return wrapi(i, g.nx) + (long)g.nx * (wrapi(j, g.ny) + (long)g.ny * wrapi(k, g.nz));
}

// ------------------------------------------------------------- reductions
// quantity q of the parameter accumulators
enum { ACC_BETA = 0, ACC_COEF = 1, ACC_WI = 2, ACC_L2 = 3, NACC = 4 };

// ---------------------------------------------------------- fft_march^T
// U' = u* - dt grad Phi; P' = P + Phi - eps lap Phi (eps = dt beta / 2Re)
__global__ __launch_bounds__(256) void k_adj_march(Grid g, const double* __restrict__ aU, const double* __restrict__ aP,
                                                   const double* __restrict__ phi, double* __restrict__ ubs,
                                                   double* __restrict__ phib, double* __restrict__ part) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double sb = 0.0;
  if (m < g.n) {
    int i, j, k;
    cell_ijk(g, m, i, j, k);
    const long n = g.n;
    const double dt = g.dt, eps = dt * g.beta / g.re / 2.0;
    const double ix2 = 1.0 / (g.dx * g.dx), iy2 = 1.0 / (g.dy * g.dy), iz2 = 1.0 / (g.dz * g.dz);
    for (int c = 0; c < 3; c++) ubs[c * n + m] = aU[c * n + m];
    const long mx = pidx(g, i - 1, j, k), my = pidx(g, i, j - 1, k), mz = pidx(g, i, j, k - 1);
    const long px = pidx(g, i + 1, j, k), py = pidx(g, i, j + 1, k), pz = pidx(g, i, j, k + 1);
    const double lapP = (aP[mx] - 2.0 * aP[m] + aP[px]) * ix2 + (aP[my] - 2.0 * aP[m] + aP[py]) * iy2 +
                        (aP[mz] - 2.0 * aP[m] + aP[pz]) * iz2;
    phib[m] = dt / g.dx * (aU[m] - aU[mx]) + dt / g.dy * (aU[n + m] - aU[n + my]) +
              dt / g.dz * (aU[2 * n + m] - aU[2 * n + mz]) + aP[m] - eps * lapP;
    const double lapPhi = (phi[mx] - 2.0 * phi[m] + phi[px]) * ix2 + (phi[my] - 2.0 * phi[m] + phi[py]) * iy2 +
                          (phi[mz] - 2.0 * phi[m] + phi[pz]) * iz2;
    sb = -(dt / (2.0 * g.re)) * aP[m] * lapPhi;
  }
  __shared__ double sh[256];
  const double s = block_sum256(sb, sh);
  if (threadIdx.x == 0) part[blockIdx.x] += s;
}

// ------------------------------------------------------ divergence^T
// b = div_h u* / dt: u*(m) enters b(m) with +1/(h dt) and b(m+e) with -1/(h dt)
__global__ __launch_bounds__(256) void k_adj_div(Grid g, const double* __restrict__ bb, double* __restrict__ ubs) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= g.n) return;
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  const long n = g.n;
  const double b0 = bb[m];
  ubs[m] += (b0 - bb[pidx(g, i + 1, j, k)]) / (g.dx * g.dt);
  ubs[n + m] += (b0 - bb[pidx(g, i, j + 1, k)]) / (g.dy * g.dt);
  ubs[2 * n + m] += (b0 - bb[pidx(g, i, j, k + 1)]) / (g.dz * g.dt);
}

// d/dbeta through the Helmholtz operator H = 1 - (dt beta/2Re) lap:
// (dt/2Re) <Rbar_c, lap u*_c>
__global__ __launch_bounds__(256) void k_adj_helm_beta(Grid g, const double* __restrict__ Rb, const double* __restrict__ up,
                                                       double* __restrict__ part) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double sb = 0.0;
  if (m < g.n) {
    int i, j, k;
    cell_ijk(g, m, i, j, k);
    const long n = g.n;
    const double ix2 = 1.0 / (g.dx * g.dx), iy2 = 1.0 / (g.dy * g.dy), iz2 = 1.0 / (g.dz * g.dz);
    const long mx = pidx(g, i - 1, j, k), my = pidx(g, i, j - 1, k), mz = pidx(g, i, j, k - 1);
    const long px = pidx(g, i + 1, j, k), py = pidx(g, i, j + 1, k), pz = pidx(g, i, j, k + 1);
    for (int c = 0; c < 3; c++) {
      const double* X = up + c * n;
      const double lap = (X[mx] - 2.0 * X[m] + X[px]) * ix2 + (X[my] - 2.0 * X[m] + X[py]) * iy2 +
                         (X[mz] - 2.0 * X[m] + X[pz]) * iz2;
      sb += Rb[c * n + m] * lap;
    }
    sb *= g.dt / (2.0 * g.re);
  }
  __shared__ double sh[256];
  const double s = block_sum256(sb, sh);
  if (threadIdx.x == 0) part[blockIdx.x] += s;
}

// ------------------------------------------------------------- IBM^T
// Forward: Ub = interp(Ua); Fc = (Uc - Ub)/dt; f = spread(Fc);
// R = Ua + dt f - dt B/2; D = -sum Fcx W (W the spreading weight g.sw, dX^3
// by default).  With Rbar given: Fcbar = dt spread^T(Rbar) - w W e_x and
// Uabar = Rbar + interp^T(-Fcbar/dt).
// This kernel writes gm = -Fcbar/dt (times dX^3 / W) per marker and component.
__global__ __launch_bounds__(256) void k_adj_ibm_markers(Grid g, long nmk, int nc, int nl, const double* __restrict__ mk,
                                                         const double* __restrict__ w, const double* __restrict__ Rb,
                                                         double wdrag, double* __restrict__ gm) {
  const long id = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (id >= nmk) return;
  const int n = (int)(id / ((long)nc * nl)) + 1;
  const double xc = mk[id], yc = mk[nmk + id];
  for (int c = 0; c < 3; c++) {
    int i0, j0;
    ibm_origin_d(g, c, xc, yc, i0, j0);
    const double* X = Rb + c * g.n;
    const double* wm = w + (c * nmk + id) * 27;   // the interpolation weights
    double acc = 0.0;
    int slot = 0;
    for (int kk = n - 1; kk <= n + 1; kk++)
      for (int jj = j0; jj <= j0 + 2; jj++)
        for (int ii = i0; ii <= i0 + 2; ii++, slot++)
          acc += X[pidx(g, ii - 1, jj - 1, kk - 1)] * wm[slot] * g.sw;
    const double fbar = g.dt * acc - (c == 0 ? g.sw * wdrag : 0.0);
    // interp^T(x) = spread(x) dX^3 / W: the spread kernels apply W
    gm[c * nmk + id] = (-fbar / g.dt) * (((g.dx * g.dx) * g.dx) / g.sw);
  }
}

__global__ __launch_bounds__(256) void k_adj_ibm_q(Grid g, const double* __restrict__ Rb, const double* __restrict__ ft,
                                                   double* __restrict__ Qb) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= 3 * g.n) return;
  const int c = (int)(e / g.n);
  const long m = e % g.n;
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  const long gx = g.nx + 2, gy = g.ny + 2, gn = gx * gy * (g.nz + 2);
  Qb[e] = Rb[e] + ft[c * gn + (i + 1) + gx * ((j + 1) + gy * (long)(k + 1))];
}

// ------------------------------------------------------- predictor^T
// Q = u - dt grad p + dt(3A - A0)/2 + cB dt B + dt(T + T0)/2 + dt F with
// cB = 1 (IBM form) or 1/2 (fft_navier); histories A0' = A, T0' = T.
// In: Qbar, Rbar (IBM: B also enters R with -dt/2), the adjoints of the
// output histories in aA0/aT0.  Out: Abar, Bbar, Tbar (ABT, 9n), the adjoints
// of the input histories (in place in aA0/aT0), Fbar += dt Qbar.
__global__ __launch_bounds__(256) void k_adj_pred_local(long n3, double dt, int method, int first,
                                                        const double* __restrict__ Qb, const double* __restrict__ Rb,
                                                        double* __restrict__ aA0, double* __restrict__ aT0,
                                                        double* __restrict__ ABT, double* __restrict__ Fb) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= n3) return;
  const double q = Qb[e];
  double ab = 1.5 * dt * q + aA0[e];
  double tb = 0.5 * dt * q + aT0[e];
  if (first) {  // A0 = A, T0 = T on the first step
    ab += -0.5 * dt * q;
    tb += 0.5 * dt * q;
  }
  const double bb = method == FP_METHOD_IBM ? dt * q - 0.5 * dt * Rb[e] : 0.5 * dt * q;
  aA0[e] = first ? 0.0 : -0.5 * dt * q;
  aT0[e] = first ? 0.0 : 0.5 * dt * q;
  Fb[e] += dt * q;
  ABT[e] = ab;
  ABT[n3 + e] = bb;
  ABT[2 * n3 + e] = tb;
}

// Velocity and pressure adjoints of the predictor: identity term, convection
// (divergence form with face-averaged fluxes), viscous, pressure gradient.
// Convection in flux form: a_x = -[(Cxx(i+1)-Cxx(i))/dx + (Fxy(j)-Fxy(j-1))/dy
// + (Fxz(k)-Fxz(k-1))/dz] with Cxx = h^2 at centres, Fxy = Vav*Uav at x-y
// edges (shared with a_y), Fxz at x-z edges (shared with a_z), Fyz at y-z
// edges (shared by a_y and a_z).
__global__ __launch_bounds__(256) void k_adj_pred_u(Grid g, const double* __restrict__ Qb, const double* __restrict__ ABT,
                                                    const double* __restrict__ U, const double* __restrict__ V,
                                                    const double* __restrict__ W, double* __restrict__ aU,
                                                    double* __restrict__ aP, double* __restrict__ part) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double sb = 0.0;
  if (m < g.n) {
    int i, j, k;
    cell_ijk(g, m, i, j, k);
    const long n = g.n;
    const double dx = g.dx, dy = g.dy, dz = g.dz, dt = g.dt;
    const double* Ax = ABT;
    const double* Ay = ABT + n;
    const double* Az = ABT + 2 * n;
    const double* Bb = ABT + 3 * n;
#define AT(A, a, b, c) A[pidx(g, i + (a), j + (b), k + (c))]
    // flux adjoints
    auto Fxy = [&](int a, int b) {  // edge (i+a+1/2, j+b+1/2)
      return (-AT(Ax, a, b, 0) + AT(Ax, a, b + 1, 0)) / dy + (-AT(Ay, a, b, 0) + AT(Ay, a + 1, b, 0)) / dx;
    };
    auto Fxz = [&](int a, int c) {  // edge (i+a+1/2, k+c+1/2)
      return (-AT(Ax, a, 0, c) + AT(Ax, a, 0, c + 1)) / dz + (-AT(Az, a, 0, c) + AT(Az, a + 1, 0, c)) / dx;
    };
    auto Fyz = [&](int b, int c) {  // edge (j+b+1/2, k+c+1/2)
      return (-AT(Ay, 0, b, c) + AT(Ay, 0, b, c + 1)) / dz + (-AT(Az, 0, b, c) + AT(Az, 0, b + 1, c)) / dy;
    };
    // U at (i,j,k)
    double ub;
    {
      const double cxx0 = (AT(Ax, 0, 0, 0) - AT(Ax, -1, 0, 0)) / dx, cxx1 = (AT(Ax, 1, 0, 0) - AT(Ax, 0, 0, 0)) / dx;
      const double h0 = (AT(U, -1, 0, 0) + AT(U, 0, 0, 0)) / 2.0, h1 = (AT(U, 0, 0, 0) + AT(U, 1, 0, 0)) / 2.0;
      ub = cxx0 * h0 + cxx1 * h1;
      ub += Fxy(0, 0) * (AT(V, 0, 0, 0) + AT(V, 1, 0, 0)) / 4.0 + Fxy(0, -1) * (AT(V, 0, -1, 0) + AT(V, 1, -1, 0)) / 4.0;
      ub += Fxz(0, 0) * (AT(W, 0, 0, 0) + AT(W, 1, 0, 0)) / 4.0 + Fxz(0, -1) * (AT(W, 0, 0, -1) + AT(W, 1, 0, -1)) / 4.0;
    }
    double vb;
    {
      const double cyy0 = (AT(Ay, 0, 0, 0) - AT(Ay, 0, -1, 0)) / dy, cyy1 = (AT(Ay, 0, 1, 0) - AT(Ay, 0, 0, 0)) / dy;
      const double g0 = (AT(V, 0, -1, 0) + AT(V, 0, 0, 0)) / 2.0, g1 = (AT(V, 0, 0, 0) + AT(V, 0, 1, 0)) / 2.0;
      vb = cyy0 * g0 + cyy1 * g1;
      vb += Fxy(0, 0) * (AT(U, 0, 0, 0) + AT(U, 0, 1, 0)) / 4.0 + Fxy(-1, 0) * (AT(U, -1, 0, 0) + AT(U, -1, 1, 0)) / 4.0;
      vb += Fyz(0, 0) * (AT(W, 0, 0, 0) + AT(W, 0, 1, 0)) / 4.0 + Fyz(0, -1) * (AT(W, 0, 0, -1) + AT(W, 0, 1, -1)) / 4.0;
    }
    double wb;
    {
      const double czz0 = (AT(Az, 0, 0, 0) - AT(Az, 0, 0, -1)) / dz, czz1 = (AT(Az, 0, 0, 1) - AT(Az, 0, 0, 0)) / dz;
      const double e0 = (AT(W, 0, 0, -1) + AT(W, 0, 0, 0)) / 2.0, e1 = (AT(W, 0, 0, 0) + AT(W, 0, 0, 1)) / 2.0;
      wb = czz0 * e0 + czz1 * e1;
      wb += Fxz(0, 0) * (AT(U, 0, 0, 0) + AT(U, 0, 0, 1)) / 4.0 + Fxz(-1, 0) * (AT(U, -1, 0, 0) + AT(U, -1, 0, 1)) / 4.0;
      wb += Fyz(0, 0) * (AT(V, 0, 0, 0) + AT(V, 0, 0, 1)) / 4.0 + Fyz(-1, 0) * (AT(V, 0, -1, 0) + AT(V, 0, -1, 1)) / 4.0;
    }
    // viscous (beta/Re) lap, identity term, pressure gradient
    const double cb = g.beta / g.re;
    const double ix2 = 1.0 / (dx * dx), iy2 = 1.0 / (dy * dy), iz2 = 1.0 / (dz * dz);
    double out[3] = {ub, vb, wb};
    for (int c = 0; c < 3; c++) {
      const double* B = Bb + c * n;
      const double* X = c == 0 ? U : (c == 1 ? V : W);
      const double lapB = (AT(B, -1, 0, 0) - 2.0 * B[m] + AT(B, 1, 0, 0)) * ix2 +
                          (AT(B, 0, -1, 0) - 2.0 * B[m] + AT(B, 0, 1, 0)) * iy2 +
                          (AT(B, 0, 0, -1) - 2.0 * B[m] + AT(B, 0, 0, 1)) * iz2;
      const double lapX = (AT(X, -1, 0, 0) - 2.0 * X[m] + AT(X, 1, 0, 0)) * ix2 +
                          (AT(X, 0, -1, 0) - 2.0 * X[m] + AT(X, 0, 1, 0)) * iy2 +
                          (AT(X, 0, 0, -1) - 2.0 * X[m] + AT(X, 0, 0, 1)) * iz2;
      out[c] += cb * lapB + Qb[c * n + m];
      sb += B[m] * lapX / g.re;
    }
    for (int c = 0; c < 3; c++) aU[c * n + m] = out[c];
    const double* Qx = Qb;
    const double* Qy = Qb + n;
    const double* Qz = Qb + 2 * n;
    aP[m] += dt / dx * (Qx[m] - AT(Qx, -1, 0, 0)) + dt / dy * (Qy[m] - AT(Qy, 0, -1, 0)) +
             dt / dz * (Qz[m] - AT(Qz, 0, 0, -1));
#undef AT
  }
  __shared__ double sh[256];
  const double s = block_sum256(sb, sh);
  if (threadIdx.x == 0) part[blockIdx.x] += s;
}

// Polymer stress T = coef div(tau), tau = f(C) C - I (C the new
// conformation), coef = (1-beta)/(Re Wi).  Out: Xbar = Cbar_state + dT/dC^T
// Tbar (including the Peterlin factor), partial sums of <tau_raw_bar, tau>
// (coefficient adjoint) and of fbar df/dL^2.
__global__ __launch_bounds__(256) void k_adj_pred_tau(Grid g, const double* __restrict__ Tb, const double* __restrict__ C1,
                                                      const double* __restrict__ aC, double* __restrict__ Xb,
                                                      double* __restrict__ part, int nb) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double s_coef = 0.0, s_l2 = 0.0;
  if (m < g.n) {
    int i, j, k;
    cell_ijk(g, m, i, j, k);
    const long n = g.n;
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    const double* Tx = Tb;
    const double* Ty = Tb + n;
    const double* Tz = Tb + 2 * n;
#define AT(A, a, b, c) A[pidx(g, i + (a), j + (b), k + (c))]
    double tr[6];
    tr[0] = (AT(Tx, -1, 0, 0) - Tx[m]) / dx;
    tr[3] = (AT(Ty, 0, -1, 0) - Ty[m]) / dy;
    tr[5] = (AT(Tz, 0, 0, -1) - Tz[m]) / dz;
    tr[1] = (-AT(Tx, 0, 1, 0) - AT(Tx, -1, 1, 0) + AT(Tx, 0, -1, 0) + AT(Tx, -1, -1, 0)) / (4.0 * dy) +
            (-AT(Ty, 1, 0, 0) - AT(Ty, 1, -1, 0) + AT(Ty, -1, 0, 0) + AT(Ty, -1, -1, 0)) / (4.0 * dx);
    tr[2] = (-AT(Tx, 0, 0, 1) - AT(Tx, -1, 0, 1) + AT(Tx, 0, 0, -1) + AT(Tx, -1, 0, -1)) / (4.0 * dz) +
            (-AT(Tz, 1, 0, 0) - AT(Tz, 1, 0, -1) + AT(Tz, -1, 0, 0) + AT(Tz, -1, 0, -1)) / (4.0 * dx);
    tr[4] = (-AT(Ty, 0, 0, 1) - AT(Ty, 0, -1, 1) + AT(Ty, 0, 0, -1) + AT(Ty, 0, -1, -1)) / (4.0 * dz) +
            (-AT(Tz, 0, 1, 0) - AT(Tz, 0, 1, -1) + AT(Tz, 0, -1, 0) + AT(Tz, 0, -1, -1)) / (4.0 * dy);
#undef AT
    double x[6];
    for (int q = 0; q < 6; q++) x[q] = C1[q * n + m];
    const double L2 = g.lp * g.lp;
    const double trc = x[0] + x[3] + x[5];
    const double f = (L2 - 3.0) / (L2 - trc);
    const double coef = (1.0 - g.beta) / (g.re * g.wi);
    double fbar = 0.0;
    for (int q = 0; q < 6; q++) {
      const double E = (q == 0 || q == 3 || q == 5) ? 1.0 : 0.0;
      s_coef += tr[q] * (f * x[q] - E);
      fbar += coef * tr[q] * x[q];
    }
    const double dfdc = f * f / (L2 - 3.0);
    const double dfdl2 = (3.0 - trc) / ((L2 - trc) * (L2 - trc));
    for (int q = 0; q < 6; q++) {
      const double E = (q == 0 || q == 3 || q == 5) ? 1.0 : 0.0;
      Xb[q * n + m] = aC[q * n + m] + f * coef * tr[q] + fbar * dfdc * E;
    }
    s_l2 = fbar * dfdl2;
  }
  __shared__ double sh[256];
  const double a = block_sum256(s_coef, sh);
  __syncthreads();
  const double b = block_sum256(s_l2, sh);
  if (threadIdx.x == 0) {
    part[ACC_COEF * nb + blockIdx.x] += a;
    part[ACC_L2 * nb + blockIdx.x] += b;
  }
}

// ---------------------------------------------------- Lyapunov Newton^T
// Velocity-gradient matrix of the Lyapunov equation (k_cstar_newton), in
// the order A[r + 3c] = M_rc of l(x) = M X + X M^T + (1 + f dt/Wi) X - b.
__device__ __forceinline__ void vel_grad(const Grid& g, const double* __restrict__ U, const double* __restrict__ V,
                                         const double* __restrict__ W, int i, int j, int k, double* A) {
  const double dt = g.dt, dx = g.dx, dy = g.dy, dz = g.dz;
#define UL(a, b, c) U[pidx(g, a, b, c)]
#define VL(a, b, c) V[pidx(g, a, b, c)]
#define WL(a, b, c) W[pidx(g, a, b, c)]
  A[0] = -(dt * (-UL(i - 1, j, k) + UL(i, j, k))) / dx;
  A[1] = -(dt * (-VL(i - 1, j - 1, k) + VL(i + 1, j - 1, k) - VL(i - 1, j, k) + VL(i + 1, j, k))) / (4.0 * dx);
  A[2] = -(dt * (-WL(i - 1, j, k - 1) + WL(i + 1, j, k - 1) - WL(i - 1, j, k) + WL(i + 1, j, k))) / (4.0 * dx);
  A[3] = -(dt * (-UL(i - 1, j - 1, k) - UL(i, j - 1, k) + UL(i - 1, j + 1, k) + UL(i, j + 1, k))) / (4.0 * dy);
  A[4] = -(dt * (-VL(i, j - 1, k) + VL(i, j, k))) / dy;
  A[5] = -(dt * (-WL(i, j - 1, k - 1) + WL(i, j + 1, k - 1) - WL(i, j - 1, k) + WL(i, j + 1, k))) / (4.0 * dy);
  A[6] = -(dt * (-UL(i - 1, j, k - 1) - UL(i, j, k - 1) + UL(i - 1, j, k + 1) + UL(i, j, k + 1))) / (4.0 * dz);
  A[7] = -(dt * (-VL(i, j - 1, k - 1) - VL(i, j, k - 1) + VL(i, j - 1, k + 1) + VL(i, j, k + 1))) / (4.0 * dz);
  A[8] = -(dt * (-WL(i, j, k - 1) + WL(i, j, k))) / dz;
#undef UL
#undef VL
#undef WL
}

// Solve the 6x6 system a y = b by Gaussian elimination with partial
// pivoting (a and b are overwritten).
__device__ __forceinline__ void solve6(double (&a)[6][6], double (&b)[6]) {
#pragma unroll
  for (int c = 0; c < 6; c++) {
    int p = c;
    double mv = fabs(a[c][c]);
#pragma unroll
    for (int r = c + 1; r < 6; r++)
      if (fabs(a[r][c]) > mv) { mv = fabs(a[r][c]); p = r; }
#pragma unroll
    for (int r = c + 1; r < 6; r++) {
      if (r == p) {
#pragma unroll
        for (int q = 0; q < 6; q++) { const double t = a[c][q]; a[c][q] = a[r][q]; a[r][q] = t; }
        const double t = b[c]; b[c] = b[r]; b[r] = t;
      }
    }
    const double inv = 1.0 / a[c][c];
#pragma unroll
    for (int r = c + 1; r < 6; r++) {
      const double f = a[r][c] * inv;
#pragma unroll
      for (int q = c; q < 6; q++) a[r][q] -= f * a[c][q];
      b[r] -= f * b[c];
    }
  }
#pragma unroll
  for (int r = 5; r >= 0; r--) {
    double s = b[r];
#pragma unroll
    for (int q = r + 1; q < 6; q++) s -= a[r][q] * b[q];
    b[r] = s / a[r][r];
  }
}

// G(X) = (1 + f(X) dt/Wi) X + M X + X M^T - C* - (dt/Wi) I = 0 at X = C^{n+1}.
// lambda = J^{-T} Xbar; C*bar = lambda; Mbar = -2 Lambda X (Lambda the
// symmetric matrix of lambda with halved off-diagonals); Wi and L^2 partials.
__global__ __launch_bounds__(256) void k_adj_newton(Grid g, const double* __restrict__ Xb, const double* __restrict__ C1,
                                                    const double* __restrict__ U, const double* __restrict__ V,
                                                    const double* __restrict__ W, double* __restrict__ lam,
                                                    double* __restrict__ Mb, double* __restrict__ part, int nb) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double s_wi = 0.0, s_l2 = 0.0;
  if (m < g.n) {
    int i, j, k;
    cell_ijk(g, m, i, j, k);
    const long n = g.n;
    const double dt = g.dt, wi = g.wi, L2 = g.lp * g.lp;
    double A[9];
    vel_grad(g, U, V, W, i, j, k, A);
    const double a11 = A[0], a21 = A[1], a31 = A[2], a12 = A[3], a22 = A[4], a32 = A[5], a13 = A[6], a23 = A[7],
                 a33 = A[8];
    double x[6], xb[6];
    for (int q = 0; q < 6; q++) { x[q] = C1[q * n + m]; xb[q] = Xb[q * n + m]; }
    const double trc = x[0] + x[3] + x[5];
    const double f = (L2 - 3.0) / (L2 - trc);
    const double dfdc = f * f / (L2 - 3.0);
    const double dg = 1.0 + f * dt / wi;
    // J[p][q] = dG_p/dx_q ; stored transposed for the solve: JT[q][p]
    const double lin[6][6] = {
        {2 * a11, 2 * a12, 2 * a13, 0, 0, 0},
        {a21, a22 + a11, a23, a12, a13, 0},
        {a31, a32, a33 + a11, 0, a12, a13},
        {0, 2 * a21, 0, 2 * a22, 2 * a23, 0},
        {0, a31, a21, a32, a33 + a22, a23},
        {0, 0, 2 * a31, 0, 2 * a32, 2 * a33}};
    double JT[6][6];
#pragma unroll
    for (int p = 0; p < 6; p++)
#pragma unroll
      for (int q = 0; q < 6; q++) {
        const double dq = (q == 0 || q == 3 || q == 5) ? dfdc : 0.0;
        JT[q][p] = lin[p][q] + (p == q ? dg : 0.0) + (dt / wi) * x[p] * dq;
      }
    solve6(JT, xb);  // xb <- lambda
    for (int q = 0; q < 6; q++) lam[q * n + m] = xb[q];
    const double Lm[3][3] = {{xb[0], 0.5 * xb[1], 0.5 * xb[2]},
                             {0.5 * xb[1], xb[3], 0.5 * xb[4]},
                             {0.5 * xb[2], 0.5 * xb[4], xb[5]}};
    const double Xm[3][3] = {{x[0], x[1], x[2]}, {x[1], x[3], x[4]}, {x[2], x[4], x[5]}};
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int q = 0; q < 3; q++) s += Lm[r][q] * Xm[q][c];
        Mb[(r + 3 * c) * n + m] = -2.0 * s;
      }
    double lx = 0.0;
    for (int q = 0; q < 6; q++) lx += xb[q] * x[q];
    s_wi = (f * dt / (wi * wi)) * lx - (dt / (wi * wi)) * (xb[0] + xb[3] + xb[5]);
    s_l2 = -(dt / wi) * ((3.0 - trc) / ((L2 - trc) * (L2 - trc))) * lx;
  }
  __shared__ double sh[256];
  const double a = block_sum256(s_wi, sh);
  __syncthreads();
  const double b = block_sum256(s_l2, sh);
  if (threadIdx.x == 0) {
    part[ACC_WI * nb + blockIdx.x] += a;
    part[ACC_L2 * nb + blockIdx.x] += b;
  }
}

// ------------------------------------------------------------- Cstar^T
// coefficients of the face reconstructions (cand_p / cand_n) of candidate l
// on C(o-1), C(o), C(o+1): pos = -1, 0, +1
__device__ __forceinline__ double coefP(int l, int pos) {
  if (l == 0) return pos == 0 ? 1.5 : (pos == 1 ? -0.5 : 0.0);
  if (l == 1) return pos == -1 ? 0.5 : (pos == 0 ? 0.5 : 0.0);
  if (l == 2) return pos == -1 ? 0.25 : (pos == 0 ? 1.0 : -0.25);
  return pos == 0 ? 1.0 : 0.0;
}
__device__ __forceinline__ double coefN(int l, int pos) {
  if (l == 0) return pos == 0 ? 0.5 : (pos == 1 ? 0.5 : 0.0);
  if (l == 1) return pos == -1 ? -0.5 : (pos == 0 ? 1.5 : 0.0);
  if (l == 2) return pos == -1 ? -0.25 : (pos == 0 ? 1.0 : 0.25);
  return pos == 0 ? 1.0 : 0.0;
}

// C*(a) = sum_d s3 Cp(a) + s1 Cp(a+e) + s4 Cn(a-e) + s2 Cn(a) with upwind
// weights of the face velocities v1 = vel(a-e), v2 = vel(a).  Writes the
// adjoint of C^n (overwrites aC) and adds the face-velocity adjoints of the
// weights to aU (the face between m and m+e belongs to cell m).
__global__ __launch_bounds__(256) void k_adj_cstar(Grid g, const double* __restrict__ lam, const double* __restrict__ C0,
                                                   const uint8_t* __restrict__ sel, const double* __restrict__ U,
                                                   const double* __restrict__ V, const double* __restrict__ W,
                                                   double* __restrict__ aC, double* __restrict__ aU) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= g.n) return;
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  const long n = g.n;
  const double dt = g.dt;
  double cb[6] = {0, 0, 0, 0, 0, 0};
  for (int d = 0; d < 3; d++) {
    const int di = d == 0, dj = d == 1, dk = d == 2;
    const double h = d == 0 ? g.dx : (d == 1 ? g.dy : g.dz);
    const double* Vd = d == 0 ? U : (d == 1 ? V : W);
    auto at = [&](int o) { return pidx(g, i + o * di, j + o * dj, k + o * dk); };
    // face velocities vel(m + o e), o = -2..1
    double vel[4];
    for (int o = -2; o <= 1; o++) vel[o + 2] = Vd[at(o)];
    auto s1 = [&](int o) { const double v = vel[o + 2]; return ((-v + fabs(v)) * dt) / (2.0 * h); };
    auto s2 = [&](int o) { const double v = vel[o + 2]; return 1.0 / 6.0 + ((-v - fabs(v)) * dt) / (2.0 * h); };
    auto s3 = [&](int o) { const double v = vel[o + 1]; return 1.0 / 6.0 + ((v - fabs(v)) * dt) / (2.0 * h); };
    auto s4 = [&](int o) { const double v = vel[o + 1]; return ((v + fabs(v)) * dt) / (2.0 * h); };
    // lambda at m + o e, o = -2..2
    double la[5][6];
    for (int o = -2; o <= 2; o++) {
      const long a = at(o);
      for (int q = 0; q < 6; q++) la[o + 2][q] = lam[q * n + a];
    }
    int ls[3];
    for (int o = -1; o <= 1; o++) ls[o + 1] = (sel[at(o)] >> (2 * d)) & 3;
    // owners o = -1, 0, 1: CPbar(o) = s3(o) lam(o) + s1(o-1) lam(o-1);
    // CNbar(o) = s2(o) lam(o) + s4(o+1) lam(o+1)
    for (int o = -1; o <= 1; o++) {
      const double w3 = s3(o), w1 = o - 1 >= -2 ? s1(o - 1) : 0.0;
      const double w2 = s2(o), w4 = s4(o + 1);
      const double cp = coefP(ls[o + 1], -o), cn = coefN(ls[o + 1], -o);
      if (cp == 0.0 && cn == 0.0) continue;
      for (int q = 0; q < 6; q++) {
        const double pb = w3 * la[o + 2][q] + w1 * la[o + 1][q];
        const double nbv = w2 * la[o + 2][q] + w4 * la[o + 3][q];
        cb[q] += cp * pb + cn * nbv;
      }
    }
    // face between m and m+e: v = vel(m); Cp(m+e) and Cn(m)
    {
      const double v = vel[2];
      const double sg = (v > 0.0) - (v < 0.0);
      const double kk = dt / (2.0 * h);
      const double d1 = (-1.0 + sg) * kk, d2 = (-1.0 - sg) * kk, d3 = (1.0 - sg) * kk, d4 = (1.0 + sg) * kk;
      const int lm = ls[1], lp = ls[2];
      double vb = 0.0;
      const long a_m1 = at(-1), a_0 = m, a_1 = at(1), a_2 = at(2);
      for (int q = 0; q < 6; q++) {
        const double* Cq = C0 + q * n;
        const double cpR = coefP(lp, -1) * Cq[a_0] + coefP(lp, 0) * Cq[a_1] + coefP(lp, 1) * Cq[a_2];
        const double cnL = coefN(lm, -1) * Cq[a_m1] + coefN(lm, 0) * Cq[a_0] + coefN(lm, 1) * Cq[a_1];
        vb += la[2][q] * (d1 * cpR + d2 * cnL) + la[3][q] * (d3 * cpR + d4 * cnL);
      }
      aU[d * n + m] += vb;
    }
  }
  for (int q = 0; q < 6; q++) aC[q * n + m] = cb[q];
}

// Velocity adjoints of the velocity-gradient matrix (vel_grad): gather of
// Mbar over the cells whose stencil reads each face velocity.
__global__ __launch_bounds__(256) void k_adj_vel_grad(Grid g, const double* __restrict__ Mb, double* __restrict__ aU) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= g.n) return;
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  const long n = g.n;
  const double t = g.dt, dx = g.dx, dy = g.dy, dz = g.dz;
#define MB(q, a, b, c) Mb[(q) * n + pidx(g, i + (a), j + (b), k + (c))]
  // A[0] = a11 (U, x), A[1] = a12 (V, x), A[2] = a13 (W, x), A[3] = a21 (U, y),
  // A[4] = a22 (V, y), A[5] = a23 (W, y), A[6] = a31 (U, z), A[7] = a32 (V, z),
  // A[8] = a33 (W, z)
  const double ub = -t / dx * (MB(0, 0, 0, 0) - MB(0, 1, 0, 0)) +
                    t / (4.0 * dy) * (MB(3, 1, 1, 0) + MB(3, 0, 1, 0) - MB(3, 1, -1, 0) - MB(3, 0, -1, 0)) +
                    t / (4.0 * dz) * (MB(6, 1, 0, 1) + MB(6, 0, 0, 1) - MB(6, 1, 0, -1) - MB(6, 0, 0, -1));
  const double vb = -t / dy * (MB(4, 0, 0, 0) - MB(4, 0, 1, 0)) +
                    t / (4.0 * dx) * (MB(1, 1, 1, 0) - MB(1, -1, 1, 0) + MB(1, 1, 0, 0) - MB(1, -1, 0, 0)) +
                    t / (4.0 * dz) * (MB(7, 0, 1, 1) + MB(7, 0, 0, 1) - MB(7, 0, 1, -1) - MB(7, 0, 0, -1));
  const double wb = -t / dz * (MB(8, 0, 0, 0) - MB(8, 0, 0, 1)) +
                    t / (4.0 * dx) * (MB(2, 1, 0, 1) - MB(2, -1, 0, 1) + MB(2, 1, 0, 0) - MB(2, -1, 0, 0)) +
                    t / (4.0 * dy) * (MB(5, 0, 1, 1) - MB(5, 0, -1, 1) + MB(5, 0, 1, 0) - MB(5, 0, -1, 0));
#undef MB
  aU[m] += ub;
  aU[n + m] += vb;
  aU[2 * n + m] += wb;
}

// drag of the last step: D = -sum Fxc dX^3 (one block)
__global__ __launch_bounds__(1024) void k_drag(const double* __restrict__ mf, long nmk, double wd, double* __restrict__ out) {
  __shared__ double sh[1024];
  double v = 0.0;
  for (long t = threadIdx.x; t < nmk; t += 1024) v += mf[t];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = -(sh[0] * wd);
}

// ------------------------------------------------------------ host state
constexpr int CK_WORDS = 19;    // U V W P C6 A0(3) T0(3) B0(3)
constexpr int TAPE_WORDS = 14;  // U V W P C6 u*(3) Phi
constexpr int BAR_WORDS = 16;   // U V W P C6 A0(3) T0(3)

struct Adj {
  long n = 0;
  int nsteps = 0, every = 0, nck = 0, sel_steps = 0;
  int recorded = 0;
  double* ck = nullptr;            // (nck + 1) checkpoints; the last is the final state
  std::vector<int> ck_hist;
  uint8_t* selh = nullptr;         // sel_steps * n
  int mixed_rec = 0;               // recorded by the mixed step
  double* drag = nullptr;          // nsteps
  double* tape = nullptr;          // every * TAPE_WORDS * n + 6n (C after the segment)
  std::vector<int> tape_first;
  int tape_seg = -1;               // the segment the tape holds (-1: none)
  double *bar = nullptr, *ubs = nullptr, *phib = nullptr, *bb = nullptr, *Rb = nullptr, *Qb = nullptr;
  double *ABT = nullptr, *Xb = nullptr, *lam = nullptr, *Mb = nullptr, *gm = nullptr, *Fb = nullptr;
  double *acc = nullptr, *part = nullptr;
  int tape_every = 0, ck_count = 0;
};

template <typename T>
void alloc0(T** p, size_t count) {
  FP_HIP_CHECK(hipMalloc((void**)p, sizeof(T) * (count ? count : 1)));
  FP_HIP_CHECK(hipMemset(*p, 0, sizeof(T) * (count ? count : 1)));
  FP_HIP_CHECK(hipDeviceSynchronize());
}
template <typename T>
void free0(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

Adj& adj_of(Ctx& c) {
  if (!c.adj) c.adj = new Adj();
  return *(Adj*)c.adj;
}

void d2d(Ctx& c, void* dst, const void* src, size_t bytes) {
  FP_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c.stream));
}

// up to eight arrays copied one after another into one range, in one launch
struct Gather {
  const double* src[8];
  long len[8];
  int count;
};
__global__ __launch_bounds__(256) void k_gather(double* __restrict__ dst, Gather gt, long total) {
  for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    long off = e;
    int q = 0;
    while (q < gt.count - 1 && off >= gt.len[q]) {
      off -= gt.len[q];
      q++;
    }
    dst[e] = gt.src[q][off];
  }
}
void gather(Ctx& c, double* dst, std::initializer_list<std::pair<const double*, long>> parts) {
  Gather gt{};
  long total = 0;
  for (const auto& pr : parts) {
    gt.src[gt.count] = pr.first;
    gt.len[gt.count] = pr.second;
    gt.count++;
    total += pr.second;
  }
  const long nb = std::min((long)nblk(total, 256), 2048L);
  hipLaunchKernelGGL(k_gather, dim3((unsigned)nb), dim3(256), 0, c.stream, dst, gt, total);
}
// tape entry of one step: the state before it, then (after the step) u* and Phi
void tape_before(Ctx& c, Adj& a, double* E, int t) {
  const long n = c.g.n;
  gather(c, E, {{c.U, n}, {c.V, n}, {c.W, n}, {c.P, n}, {c.C[c.cur], 6 * n}});
  a.tape_first[t] = c.has_history ? 0 : 1;
}
void tape_after(Ctx& c, double* E) {
  const long n = c.g.n;
  gather(c, E + 10 * n, {{c.Up, 3 * n}, {c.Phi, n}});
}
void ensure_tape(Adj& a, int every, long n) {
  if (!a.tape || a.tape_every != every) {
    free0(a.tape);
    alloc0(&a.tape, (size_t)every * TAPE_WORDS * n + 6 * n);
    a.tape_every = every;
    a.tape_seg = -1;
  }
  if ((int)a.tape_first.size() != every) a.tape_first.assign(every, 0);
}

void save_state(Ctx& c, double* dst) {
  const long n = c.g.n;
  d2d(c, dst, c.U, sizeof(double) * n);
  d2d(c, dst + n, c.V, sizeof(double) * n);
  d2d(c, dst + 2 * n, c.W, sizeof(double) * n);
  d2d(c, dst + 3 * n, c.P, sizeof(double) * n);
  d2d(c, dst + 4 * n, c.C[c.cur], sizeof(double) * 6 * n);
  d2d(c, dst + 10 * n, c.A0, sizeof(double) * 3 * n);
  d2d(c, dst + 13 * n, c.T0, sizeof(double) * 3 * n);
  d2d(c, dst + 16 * n, c.B0, sizeof(double) * 3 * n);
}

void load_state(Ctx& c, const double* src) {
  const long n = c.g.n;
  d2d(c, c.U, src, sizeof(double) * n);
  d2d(c, c.V, src + n, sizeof(double) * n);
  d2d(c, c.W, src + 2 * n, sizeof(double) * n);
  d2d(c, c.P, src + 3 * n, sizeof(double) * n);
  d2d(c, c.C[c.cur], src + 4 * n, sizeof(double) * 6 * n);
  d2d(c, c.A0, src + 10 * n, sizeof(double) * 3 * n);
  d2d(c, c.T0, src + 13 * n, sizeof(double) * 3 * n);
  d2d(c, c.B0, src + 16 * n, sizeof(double) * 3 * n);
}

void check_supported(Ctx& c) {
  if (c.g.walls) throw std::string("E_ARG:the adjoint supports periodic grids only");
  if (c.p.poisson_variant != FP_POISSON_FFT) throw std::string("E_ARG:the adjoint needs the FFT Poisson solve");
  if (c.p.method != FP_METHOD_FFT && c.p.method != FP_METHOD_IBM)
    throw std::string("E_ARG:the adjoint supports methods 1 (fft_navier) and 2 (IBM)");
}

void sum_part(Ctx& c, Adj& a, int q, int nb) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, c.stream, a.part + (long)q * nb, nb, a.acc + q, 1);
}

// one adjoint step of tape entry t (step s of the record)
void adjoint_step(Ctx& c, Adj& a, int t, int seg_len, int s, double wdrag) {
  const Grid& g = c.g;
  const long n = g.n;
  hipStream_t st = c.stream;
  const int nb = nblk(n, 256);
  const double* E = a.tape + (long)t * TAPE_WORDS * n;
  const double *U = E, *V = E + n, *W = E + 2 * n, *C0 = E + 4 * n, *Up = E + 10 * n, *Phi = E + 13 * n;
  const double* C1 = t + 1 < seg_len ? a.tape + (long)(t + 1) * TAPE_WORDS * n + 4 * n
                                     : a.tape + (long)a.tape_every * TAPE_WORDS * n;
  const uint8_t* sel = a.selh + (long)s * n;
  double* aU = a.bar;
  double* aP = a.bar + 3 * n;
  double* aC = a.bar + 4 * n;
  double* aA0 = a.bar + 10 * n;
  double* aT0 = a.bar + 13 * n;
  const int polymer = c.p.beta != 1.0;
  // projection
  hipLaunchKernelGGL(k_adj_march, dim3(nb), dim3(256), 0, st, g, aU, aP, Phi, a.ubs, a.phib, a.part);
  f64_solve(c, 3, a.phib, a.bb);
  hipLaunchKernelGGL(k_adj_div, dim3(nb), dim3(256), 0, st, g, a.bb, a.ubs);
  // Helmholtz
  for (int comp = 0; comp < 3; comp++) f64_solve(c, comp, a.ubs + comp * n, a.Rb + comp * n);
  hipLaunchKernelGGL(k_adj_helm_beta, dim3(nb), dim3(256), 0, st, g, a.Rb, Up, a.part);
  // immersed boundary
  const double* Qb = a.Rb;
  if (c.p.method == FP_METHOD_IBM) {
    hipLaunchKernelGGL(k_adj_ibm_markers, dim3(nblk(c.nmk, 256)), dim3(256), 0, st, g, c.nmk, c.nc, c.nl, c.mk, c.ib_w,
                       a.Rb, wdrag, a.gm);
    f64_spread_markers(c, a.gm, c.ftmp);
    hipLaunchKernelGGL(k_adj_ibm_q, dim3(nblk(3 * n, 256)), dim3(256), 0, st, g, a.Rb, c.ftmp, a.Qb);
    Qb = a.Qb;
  }
  // predictor
  hipLaunchKernelGGL(k_adj_pred_local, dim3(nblk(3 * n, 256)), dim3(256), 0, st, 3 * n, g.dt, c.p.method,
                     a.tape_first[t], Qb, a.Rb, aA0, aT0, a.ABT, a.Fb);
  hipLaunchKernelGGL(k_adj_pred_u, dim3(nb), dim3(256), 0, st, g, Qb, a.ABT, U, V, W, aU, aP, a.part);
  if (polymer) {
    hipLaunchKernelGGL(k_adj_pred_tau, dim3(nb), dim3(256), 0, st, g, a.ABT + 6 * n, C1, aC, a.Xb, a.part, nb);
    hipLaunchKernelGGL(k_adj_newton, dim3(nb), dim3(256), 0, st, g, a.Xb, C1, U, V, W, a.lam, a.Mb, a.part, nb);
    hipLaunchKernelGGL(k_adj_cstar, dim3(nb), dim3(256), 0, st, g, a.lam, C0, sel, U, V, W, aC, aU);
    hipLaunchKernelGGL(k_adj_vel_grad, dim3(nb), dim3(256), 0, st, g, a.Mb, aU);
  } else {
    zero_words(aC, 12 * n, st);
  }
}

}  // namespace

// A change of the step's physics, force or markers after a recording: the
// reverse sweep recomputes every segment (with the settings then current), as
// it would without the taped segment.
void adj_drop_tape(Ctx& c) {
  if (c.adj) ((Adj*)c.adj)->tape_seg = -1;
}

void adj_free(Ctx& c) {
  if (!c.adj) return;
  Adj& a = *(Adj*)c.adj;
  free0(a.ck); free0(a.selh); free0(a.drag); free0(a.tape);
  free0(a.bar); free0(a.ubs); free0(a.phib); free0(a.bb); free0(a.Rb); free0(a.Qb);
  free0(a.ABT); free0(a.Xb); free0(a.lam); free0(a.Mb); free0(a.gm); free0(a.Fb);
  free0(a.acc); free0(a.part);
  delete (Adj*)c.adj;
  c.adj = nullptr;
}

void adj_record(Ctx& c, int nsteps, int every, int replay, double* drag_out) {
  check_supported(c);
  if (nsteps < 1 || every < 1) throw std::string("E_ARG:nsteps and every must be >= 1");
  if (c.p.method == FP_METHOD_IBM && c.nmk == 0) throw std::string("E_STATE:IBM method without markers");
  Adj& a = adj_of(c);
  const long n = c.g.n;
  const int polymer = c.p.beta != 1.0;
  if (replay && c.mixed) throw std::string("E_ARG:replay of recorded selections runs on the FP64 path");
  if (replay && (!a.selh || a.sel_steps < nsteps || a.n != n || a.mixed_rec))
    throw std::string("E_STATE:replay needs a recorded run of at least as many steps");
  a.n = n;
  const int nck = (nsteps + every - 1) / every;
  if (a.ck_count != nck + 1) {
    free0(a.ck);
    alloc0(&a.ck, (size_t)(nck + 1) * CK_WORDS * n);
    a.ck_count = nck + 1;
  }
  if (!replay && a.sel_steps != nsteps) {
    free0(a.selh);
    alloc0(&a.selh, (size_t)nsteps * n);
    a.sel_steps = nsteps;
  }
  free0(a.drag);
  alloc0(&a.drag, (size_t)nsteps);
  a.ck_hist.assign(nck + 1, 0);
  // the FP64 recording tapes its last segment, which the reverse sweep then
  // need not recompute (the recomputation replays the recorded selections,
  // so it is the same computation)
  const int tape_from = c.mixed ? nsteps : (nck - 1) * every;
  a.tape_seg = -1;
  if (!c.mixed) ensure_tape(a, every, n);
  const double wd = c.g.sw;   // the drag sum's weight
  const int step0 = c.step;
  // Mixed precision: the checkpoints and the drag are those of the mixed
  // step (its FP32 state widened to FP64); the reverse sweep recomputes each
  // segment from its checkpoint with the FP64 step and its own selections.
  a.mixed_rec = c.mixed ? 1 : 0;
  for (int s = 0; s < nsteps; s++) {
    if (s % every == 0) {
      if (c.mixed) mixed_store_to_f64(c);
      save_state(c, a.ck + (size_t)(s / every) * CK_WORDS * n);
      a.ck_hist[s / every] = c.has_history;
    }
    if (c.mixed) {
      mixed_step(c);
      if (c.p.method == FP_METHOD_IBM)
        hipLaunchKernelGGL(k_drag, dim3(1), dim3(1024), 0, c.stream, mixed_marker_forces(c), c.nmk, wd, a.drag + s);
      continue;
    }
    double* E = s >= tape_from ? a.tape + (size_t)(s - tape_from) * TAPE_WORDS * n : nullptr;
    if (E) tape_before(c, a, E, s - tape_from);
    c.sel_replay = (replay && polymer) ? a.selh + (size_t)s * n : nullptr;
    step_f64(c);
    c.sel_replay = nullptr;
    if (E) tape_after(c, E);
    if (!replay && polymer) d2d(c, a.selh + (size_t)s * n, c.sel, n);
    if (c.p.method == FP_METHOD_IBM)
      hipLaunchKernelGGL(k_drag, dim3(1), dim3(1024), 0, c.stream, c.mf, c.nmk, wd, a.drag + s);
  }
  if (!c.mixed) {
    d2d(c, a.tape + (size_t)every * TAPE_WORDS * n, c.C[c.cur], sizeof(double) * 6 * n);
    a.tape_seg = nck - 1;
  }
  if (c.mixed) {
    mixed_join(c);
    mixed_store_to_f64(c);
  }
  save_state(c, a.ck + (size_t)nck * CK_WORDS * n);
  a.ck_hist[nck] = c.has_history;
  c.step = step0 + nsteps;
  a.nsteps = nsteps;
  a.every = every;
  a.nck = nck;
  a.recorded = 1;
  if (drag_out) FP_HIP_CHECK(hipMemcpyAsync(drag_out, a.drag, sizeof(double) * nsteps, hipMemcpyDeviceToHost, c.stream));
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
}

void adj_backward(Ctx& c, const double* weights, const double* bar_final, double* bar_initial, double* bar_params,
                  double* bar_force) {
  check_supported(c);
  if (!c.adj || !((Adj*)c.adj)->recorded) throw std::string("E_STATE:fp_adj_backward before fp_adj_record");
  Adj& a = *(Adj*)c.adj;
  const Grid& g = c.g;
  const long n = g.n;
  const int every = a.every;
  ensure_tape(a, every, n);
  if (!a.bar) {
    alloc0(&a.bar, (size_t)BAR_WORDS * n);
    alloc0(&a.ubs, 3 * n); alloc0(&a.phib, n); alloc0(&a.bb, n); alloc0(&a.Rb, 3 * n); alloc0(&a.Qb, 3 * n);
    alloc0(&a.ABT, 9 * n); alloc0(&a.Xb, 6 * n); alloc0(&a.lam, 6 * n); alloc0(&a.Mb, 9 * n);
    alloc0(&a.gm, 3 * (c.nmk > 0 ? c.nmk : 1)); alloc0(&a.Fb, 3 * n);
    alloc0(&a.acc, NACC); alloc0(&a.part, (size_t)NACC * nblk(n, 256));
  }
  hipStream_t st = c.stream;
  if (c.mixed) mixed_join(c);   // no selection made ahead still writes c.sel
  if (bar_final) d2d(c, a.bar, bar_final, sizeof(double) * BAR_WORDS * n);
  else zero_words(a.bar, 2 * BAR_WORDS * n, st);
  zero_words(a.Fb, 6 * n, st);
  zero_words(a.acc, 2 * NACC, st);
  // the kernels add their per-block partials of the parameter gradients over
  // all steps; they are summed once, after the sweep
  const int nbp = nblk(n, 256);
  zero_words(a.part, 2 * NACC * nbp, st);
  const int step0 = c.step;
  for (int seg = a.nck - 1; seg >= 0; seg--) {
    const int s0 = seg * every, s1 = std::min(s0 + every, a.nsteps);
    if (seg != a.tape_seg) {
      load_state(c, a.ck + (size_t)seg * CK_WORDS * n);
      c.has_history = a.ck_hist[seg];
      if (a.mixed_rec) {
        // a mixed recording: the segment is recomputed with the mixed step
        // from its checkpoint (FP32 values, so exactly), and the tape holds
        // its states widened; the FP64 adjoint steps are linearised there
        mixed_load_from_f64(c);
        for (int s = s0; s < s1; s++) {
          double* E = a.tape + (size_t)(s - s0) * TAPE_WORDS * n;
          a.tape_first[s - s0] = c.has_history ? 0 : 1;
          mixed_tape_in(c, E, c.p.beta != 1.0 ? a.selh + (size_t)s * n : nullptr);
          mixed_step(c);
          mixed_tape_out(c, E);
        }
        mixed_join(c);
        mixed_widen_c(c, a.tape + (size_t)every * TAPE_WORDS * n);
      } else {
        for (int s = s0; s < s1; s++) {
          double* E = a.tape + (size_t)(s - s0) * TAPE_WORDS * n;
          tape_before(c, a, E, s - s0);
          c.sel_replay = c.p.beta != 1.0 ? a.selh + (size_t)s * n : nullptr;
          step_f64(c);
          c.sel_replay = nullptr;
          tape_after(c, E);
        }
        d2d(c, a.tape + (size_t)every * TAPE_WORDS * n, c.C[c.cur], sizeof(double) * 6 * n);
      }
      a.tape_seg = seg;
    }
    // tape entry t holds the state before step s0 + t; the C after the
    // segment follows the last entry
    const int len = s1 - s0;
    for (int s = s1 - 1; s >= s0; s--) adjoint_step(c, a, s - s0, len, s, weights ? weights[s] : 0.0);
  }
  // leave the context at the recorded final state
  load_state(c, a.ck + (size_t)a.nck * CK_WORDS * n);
  c.has_history = a.ck_hist[a.nck];
  c.step = step0;
  if (c.mixed) mixed_load_from_f64(c);
  if (bar_initial) d2d(c, bar_initial, a.bar, sizeof(double) * BAR_WORDS * n);
  if (bar_force) d2d(c, bar_force, a.Fb, sizeof(double) * 3 * n);
  for (int q = 0; q < NACC; q++) sum_part(c, a, q, nbp);
  double acc[NACC];
  FP_HIP_CHECK(hipMemcpyAsync(acc, a.acc, sizeof(acc), hipMemcpyDeviceToHost, st));
  FP_HIP_CHECK(hipStreamSynchronize(st));
  if (bar_params) {
    const double re = c.p.re, wi = c.p.wi, beta = c.p.beta;
    bar_params[0] = acc[ACC_WI] - acc[ACC_COEF] * (1.0 - beta) / (re * wi * wi);
    bar_params[1] = acc[ACC_BETA] - acc[ACC_COEF] / (re * wi);
    bar_params[2] = acc[ACC_L2];
  }
}

}  // namespace fp
