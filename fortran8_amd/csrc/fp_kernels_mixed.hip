// Mixed-precision path of the FENE-P / IBM SMAC step.
//
// Storage and stencils in FP32 (U, V, W, P, C, histories, FFT spectra, IBM
// weights); FP64 where the accuracy demands it:
//   * the pressure-correction Poisson solve is iteratively refined: the
//     residual b - L_h phi and the potential phi are FP64, each correction is
//     an FP32 spectral (+ tridiagonal) solve;
//   * the Peterlin factor f = (L^2-3)/(L^2-trC) when tr C is within 1% of L^2;
//   * candidate selections of the C+- reconstruction whose FP32 margin is
//     inside a guard band are re-decided with the FP64 eigenvalue test;
//   * drag / energy reductions.
// The Lyapunov Newton uses the analytic Jacobian J = M(f) + (dt/Wi) f' x e^T
// (e = [1,0,0,1,0,1]) instead of the reference's central differences, which
// carry no information in FP32 (fenep.f90:558-568 uses dc = 1e-5).
#include "fp_mixed.h"
#include "fp_device.h"
#include <algorithm>
#include <cmath>
#include <map>
#include <unordered_map>
#include <vector>

namespace fp {

namespace {

struct Mixed {
  // The C+- selection of the C a step produces is made on a second stream
  // (default priority) beside the rest of that step; the next step's
  // CStar/Newton waits for it (sel_done).
  hipStream_t side = nullptr;
  hipEvent_t ev_cstar = nullptr, sel_done = nullptr;
  int sel_ready = 0;              // c.sel holds the selection of the current C
  float *U = nullptr, *V = nullptr, *W = nullptr, *P = nullptr;
  float* C[2] = {nullptr, nullptr};
  int cur = 0;
  float* f = nullptr;
  float *A0 = nullptr, *T0 = nullptr, *B = nullptr;
  float* Q = nullptr;             // predictor increment d_q (+ dt f of the IBM spread)
  float* Up = nullptr;            // Helmholtz increment d: u* = u^n + d
  float* F = nullptr;
  float2* spec = nullptr;
  float* thp = nullptr;
  float2 *twx = nullptr, *twy = nullptr, *twz = nullptr;
  float *kx2 = nullptr, *ky2 = nullptr, *kz2 = nullptr;
  double *phi = nullptr, *b = nullptr;
  float* r32 = nullptr;
  float* e32 = nullptr;           // SOR: the FP32 correction e
  double* sor_pb = nullptr;       // SOR: per-block |b|^2
  int* sor_flag = nullptr;        // SOR: converged, sweeps, FP64 check (m_sor_decide)
  int* sor_host = nullptr;        // pinned copy of sor_flag
  double* red = nullptr;          // [0]=|b|^2, [1]=|r|^2 of the last sweep
  double* part = nullptr;         // per-block partial sums
  int* skip = nullptr;            // defect correction converged: skip the sweep
  double* partb = nullptr;        // per-block |b|^2 of the fused divergence
  int slab = 0;                   // walls: Poisson through m_poisson_slab
  float* pbuf = nullptr;          // its line-solve coefficients
  int* dc_host = nullptr;         // pinned copy of the last periodic decision
  hipEvent_t dc_ev = nullptr;
  bool dc_pending = false, dc_correct = false;
  float* X2 = nullptr;            // Helmholtz iterate
  int helm_iters = 0;             // fixed-point sweeps (0: spectral solve)
  unsigned* amb = nullptr;        // uncertified selections, re-decided in FP64
  unsigned* amb_list = nullptr;   // per-block regions of m * 3 + axis entries
  unsigned* amb_idx = nullptr;    // the entries' places in amb_list, packed (3 n)
  float* amb_data = nullptr;      // staged cell values, 18 * SEL_STAGE per select block
  unsigned* chk = nullptr;        // selection check: differing decisions
  // IBM
  float* wi = nullptr;            // interpolation weights, nmk * 3 * 27
  int4* org = nullptr;            // per marker and component: i0, j0, n
  float* uc = nullptr;            // target velocities, 3 planes
  float* mf = nullptr;            // marker forces, 3 planes
  double* mf64 = nullptr;
  float* ftmp = nullptr;          // 3 ghosted raw spread arrays
  float* sw[3] = {nullptr, nullptr, nullptr};  // CSR entry weights
  // spreading folded onto interior targets (all three components)
  int* fr_row = nullptr;          // sliced ELL: offset of each 64-row slice
  int* fr_len = nullptr;          // entries per row of each slice (padded)
  long* fr_tgt = nullptr;         // c*n + interior index
  int* fr_ent = nullptr;          // c*nmk + marker
  float* fr_w = nullptr;
  int fr_rows = 0;
  // markers repeated along z (every layer is layer 0 moved up by whole
  // planes, checked on the weights themselves): the interpolation weights of
  // layer 0 and the spreading rows of plane 0 serve every plane
  int zinv = 0;
  long zP = 0;                    // markers per layer
  float* zi_w = nullptr;          // interpolation weights of layer 0, 3 * 27 * zP
  int2* zi_org = nullptr;         // i0, j0 per component and layer-0 marker
  int* zs_off = nullptr;          // plane rows as sliced ELL (as fr_*)
  int* zs_len = nullptr;
  long* zs_tgt = nullptr;         // c*n + in-plane interior index
  int* zs_ent = nullptr;          // 4 * (c*nmk + layer-0 marker) + layer offset + 1
  float* zs_w = nullptr;
  int zs_rows = 0, zs_rows64 = 0;
  double last_ndc = 0.0;
};

inline Mixed& MX(Ctx& c) { return *(Mixed*)c.mixed; }
inline int nblk(long n, int t) { return (int)((n + t - 1) / t); }
// cells in the 2R rows next to the y-walls (the GEN = true instances)
inline long wall_rows(const Grid& g, int R) { return (long)g.nx * 2 * R * g.nz; }

template <typename T>
void dalloc(T** p, size_t count) {
  FP_HIP_CHECK(hipMalloc((void**)p, sizeof(T) * (count ? count : 1)));
  FP_HIP_CHECK(hipMemset(*p, 0, sizeof(T) * (count ? count : 1)));
  // the zero fill runs on the null stream; kernels that initialise the
  // buffer next run on the context's non-blocking stream
  FP_HIP_CHECK(hipDeviceSynchronize());
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace

// ================================================================ FENE-P ===
// Wrapped offset of cell a+d along an axis of n cells and stride s (|d| <= 2).
// Every axis and stride is a power of two: the wrap is a mask and the stride
// a shift (no compare, select or 32-bit multiply per offset).
__device__ __forceinline__ long woff(int a, int d, int n, long s) {
  return (long)((unsigned long)(long)(((a + d) & (n - 1)) - a) << __builtin_ctzl((unsigned long)s));
}
// Element e of an array of n elements through a raw buffer resource: the
// base and the size sit in scalar registers and each load or store takes one
// 32-bit byte offset per lane instead of a 64-bit address (n <= 2^30 cells
// keeps 4 n below 2^32).  An offset beyond the size would read zero and
// drop the write; every index here is a cell index below n.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)bytes, 0x00020000);
}
template <typename V>
__device__ __forceinline__ V ldu(const V* base, long n, unsigned e) {
  static_assert(sizeof(V) == 4 || sizeof(V) == 1, "4-byte or 1-byte elements");
  const __amdgpu_buffer_rsrc_t r = brsrc(base, (unsigned)n * (unsigned)sizeof(V));
  if constexpr (sizeof(V) == 4)
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(e * 4u), 0, 0));
  else
    return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b8(r, (int)e, 0, 0));
}
template <typename V>
__device__ __forceinline__ void stu(V* base, long n, unsigned e, V v) {
  static_assert(sizeof(V) == 4, "4-byte elements");
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), brsrc(base, (unsigned)n * 4u), (int)(e * 4u), 0, 0);
}
// wrapped neighbour offset along an axis in 32 bits
__device__ __forceinline__ int woff32(int a, int d, int n, int s) {
  return (int)((unsigned)(((a + d) & (n - 1)) - a) << __builtin_ctz((unsigned)s));
}

// ---- C+- selection in FP32 with certified decisions ----------------------
// The selection of fenep.f90:186-221 takes, per cell and axis, the maxloc over
// {v_0, v_1, v_2, 0} of v_l = min(lambda_min(Cp_l), lambda_min(Cn_l)), with
// lambda_min from the trigonometric root of Vieta (fenep.f90:425-444) and the
// sentinel -100 where the discriminant t >= 0.  The FP32 kernel evaluates
// lambda_min on the deviatoric part B = A - (tr A/3) I (so that p and q do not
// cancel) together with a bound on |lambda_32 - lambda_ref|, where lambda_ref
// is the reference formula evaluated in FP64 on the same stored FP32 values.
// A decision is kept only if the interval of the winner is strictly separated
// from every other interval and no candidate is near the sentinel boundary;
// otherwise the (cell, axis) is re-decided with the reference formula in FP64
// (m_select_fix).  The final selection is therefore the reference's selection
// of the stored field.
//
// With m = tr A/3, r = sqrt(-p/3), arg = det B/(2 r^3) = cos(theta),
// lambda_min = m + 2 r psi(s),  psi(s) = cos(acos(1 - s^2)/3 + 2 pi/3),
// s = sqrt(1 - arg).  psi is analytic in s on [0, sqrt 2] (the square-root
// singularity of acos at arg = 1 is absorbed by s, and psi is even in
// pi - theta at arg = -1), so it is a polynomial of degree 8 (Chebyshev
// least-squares fit at 4000 Chebyshev nodes, converted to powers of s;
// max |error| 2.4e-8 on [0, sqrt 2] with the coefficients as printed, below
// half an FP32 ulp of |psi| >= 0.5).
#define PSI0 -5.000000000e-01f
#define PSI1 -4.082478881e-01f
#define PSI2 5.554784089e-02f
#define PSI3 -1.884038374e-02f
#define PSI4 7.990322076e-03f
#define PSI5 -3.478206461e-03f
#define PSI6 1.290685032e-03f
#define PSI7 -3.277156502e-04f
#define PSI8 3.995704537e-05f

// Error model, rho = S/r with S = |m| + 2 r (a bound on |lambda|):
//   |arg_32 - arg_exact|  <= eps32 (K1 + K2 rho)    (B's diagonal carries
//                                                    eps32 |m| absolute error)
//   |arg_ref - arg_exact| <= eps64 K3 rho^3         (p and q of the reference
//                                                    cancel on the full matrix)
//   d lambda / d arg      <= r / sqrt(1 - arg)      (0.82 r sqrt(delta) in the
//                                                    square-root regime)
//   plus Km eps32 S for m, r and psi, and Kc eps32 Smax for forming the
//   candidates in FP32 (Smax bounds the entries of the three cells).
// K1 = K2 = Km = Kc = 8, K3 = 64: in a float32 emulation over 10^6 random
// symmetric matrices (definite, indefinite, near double roots at either end,
// near isotropic, exactly uniaxial) |lambda_32 - lambda_ref| stays below 0.37
// of the bound, and no reference sentinel or NaN falls outside the flagged
// zone |arg| > 1 - 2 delta.
#define SEL_K1E 4.76837158e-07f    /* 8 eps32 */
#define SEL_K2E 4.76837158e-07f    /* 8 eps32 */
#define SEL_K3E 7.10542736e-15f    /* 64 eps64 */
#define SEL_KME 4.76837158e-07f    /* 8 eps32 */
#define SEL_KCE 4.76837158e-07f    /* 8 eps32 */

// lambda_min of one candidate and the half-width of its interval without
// the term of the cell's entries (ein, added by the caller): neighbouring
// cells that share a candidate share its evaluation.
struct Lmin {
  float lam, e0;
  bool amb;   // near the sentinel boundary (or degenerate): decide in FP64
};

__device__ __forceinline__ Lmin lmin32(const float* a) {
  const float third = 1.0f / 3.0f;
  const float m = ((a[0] + a[3]) + a[5]) * third;
  const float b11 = a[0] - m, b22 = a[3] - m, b33 = a[5] - m;
  const float o2 = fmaf(a[1], a[1], fmaf(a[2], a[2], a[4] * a[4]));
  const float hp = fmaf(b11, b11, fmaf(b22, b22, b33 * b33));
  // -p/3 >= 0; 0 for an isotropic A, whose arg is then NaN (ambiguous)
  const float r2 = fmaf(hp, 1.0f / 6.0f, o2 * third);
  const float t1 = fmaf(b22, b33, -(a[4] * a[4]));
  const float t2 = fmaf(a[1], b33, -(a[4] * a[2]));
  const float t3 = fmaf(a[1], a[4], -(b22 * a[2]));
  const float det = fmaf(b11, t1, fmaf(-a[1], t2, a[2] * t3));
  // two reciprocal square roots (quarter-rate) per eigenvalue
  const float ir = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * ir;
  const float ir3 = (ir * ir) * ir;
  const float arg = (0.5f * det) * ir3;
  const float s2 = 1.0f - arg;
  // rho = (|m| + 2 r)/r; d > 1 only where the candidate is flagged below
  const float rho = fmaf(fabsf(m), ir, 2.0f);
  const float d = fmaf(rho, fmaf(SEL_K3E, rho * rho, SEL_K2E), SEL_K1E);
  // 1/sqrt(max(1 - arg, delta)); where the decision is certified, 1 - arg >=
  // 2 delta and s = (1 - arg) v is sqrt(1 - arg) (s beyond sqrt 2, arg < -1,
  // only where the candidate is ambiguous)
  const float v = __builtin_amdgcn_rsqf(fmaxf(s2, d));
  const float s = s2 * v;
  // 2 psi(s): the coefficients doubled (exactly), lambda = m + r (2 psi)
  float ps = 2.0f * PSI8;
  ps = fmaf(ps, s, 2.0f * PSI7);
  ps = fmaf(ps, s, 2.0f * PSI6);
  ps = fmaf(ps, s, 2.0f * PSI5);
  ps = fmaf(ps, s, 2.0f * PSI4);
  ps = fmaf(ps, s, 2.0f * PSI3);
  ps = fmaf(ps, s, 2.0f * PSI2);
  ps = fmaf(ps, s, 2.0f * PSI1);
  ps = fmaf(ps, s, 2.0f * PSI0);
  Lmin o;
  o.lam = fmaf(r, ps, m);
  // written so that a NaN arg counts as ambiguous
  o.amb = !(fabsf(arg) <= fmaf(-2.0f, d, 1.0f));
  o.e0 = r * fmaf(d, v, SEL_KME * rho);
  return o;
}

// bound on the entries of one cell's C: |diag| + 2 |offdiag| summed
__device__ __forceinline__ float entry_bound(const float* c) {
  return (fabsf(c[0]) + fabsf(c[3]) + fabsf(c[5])) + 2.0f * ((fabsf(c[1]) + fabsf(c[2])) + fabsf(c[4]));
}

// candidate face values (fenep.f90:179-184) in FP32, each with at most two
// roundings: left face Cp_l and right face Cn_l, l = 0..2.  The halves and
// quarters are exact, so the midpoint c/2 + cp/2 of one cell's right face is
// bit for bit the midpoint cm/2 + c/2 of the next cell's left face.
__device__ __forceinline__ float cand_pf(int l, float cm, float c, float cp) {
  if (l == 0) return fmaf(c, 1.5f, -(0.5f * cp));
  if (l == 1) return fmaf(c, 0.5f, 0.5f * cm);
  return fmaf(cp, -0.25f, fmaf(cm, 0.25f, c));
}
__device__ __forceinline__ float cand_nf(int l, float cm, float c, float cp) {
  if (l == 0) return fmaf(c, 0.5f, 0.5f * cp);
  if (l == 1) return fmaf(c, 1.5f, -(0.5f * cm));
  return fmaf(cm, -0.25f, fmaf(cp, 0.25f, c));
}
// the six candidates of one axis line (cm, c, cp) evaluated
__device__ __forceinline__ Lmin cand_lmin(bool right, int l, const float (&v)[3][6]) {
  float a[6];
#pragma unroll
  for (int q = 0; q < 6; q++) a[q] = right ? cand_nf(l, v[0][q], v[1][q], v[2][q]) : cand_pf(l, v[0][q], v[1][q], v[2][q]);
  return lmin32(a);
}

// maxloc over {v_0, v_1, v_2, 0} of v_l = min(lambda(Cp_l), lambda(Cn_l))
// (the first maximum); unsure unless the winner's interval clears every
// other option's and no candidate is ambiguous
__device__ __forceinline__ int sel_decide(const Lmin (&P)[3], const Lmin (&N)[3], float ein, bool& unsure) {
  float vals[4], errs[4];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    vals[l] = fminf(P[l].lam, N[l].lam);
    errs[l] = fmaxf(P[l].e0, N[l].e0) + ein;
    unsure = unsure || P[l].amb || N[l].amb;
  }
  vals[3] = 0.0f;
  errs[3] = 0.0f;
  int idx = 0;
  float best = vals[0], eb = errs[0];
  if (vals[1] > best) { best = vals[1]; eb = errs[1]; idx = 1; }
  if (vals[2] > best) { best = vals[2]; eb = errs[2]; idx = 2; }
  if (0.0f > best) { best = 0.0f; eb = 0.0f; idx = 3; }
  // certified only if the winner's interval clears every other interval
#pragma unroll
  for (int l = 0; l < 4; l++)
    if (l != idx) unsure = unsure || !(best - vals[l] > eb + errs[l]);
  return idx;
}

// C at cells a-r..a+r along axis d (r = 1 or 2) for the 6 components,
// through wrapped offsets.  With y-walls a point of the line on the ghost row
// takes the linear extrapolation 2 C_0 - C_1 of the two nearest rows
// (C_kabe), and a point beyond it the centre (it only feeds faces whose
// upwind weight is exactly zero); the wrapped rows loaded there are replaced.
template <int R>
__device__ __forceinline__ void load_line(const float* __restrict__ C, const Grid& g, long m, int i, int j, int k, int d,
                                          float (&v)[2 * R + 1][6]) {
  // 32-bit cell indices from the plane bases (n <= 2^30 cells)
  unsigned e[2 * R + 1];
#pragma unroll
  for (int o = -R; o <= R; o++) {
    int off;
    if (d == 0) off = woff32(i, o, g.nx, 1);
    else if (d == 1) off = woff32(j, o, g.ny, g.nx);
    else off = woff32(k, o, g.nz, g.nx * g.ny);
    e[o + R] = (unsigned)m + (unsigned)off;
  }
#pragma unroll
  for (int o = 0; o < 2 * R + 1; o++)
#pragma unroll
    for (int q = 0; q < 6; q++) v[o][q] = ldu(C + q * g.n, g.n, e[o]);
  // (the lanes of a wavefront share their row in the callers' cell orders,
  // so the interior rows branch past the patches)
  if (R == 1 && d == 1 && g.walls && (j == 0 || j == g.ny - 1)) {
    const bool lo = j == 0, hi = j == g.ny - 1;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      v[0][q] = lo ? 2.0f * v[1][q] - v[2][q] : v[0][q];
      v[2][q] = hi ? 2.0f * v[1][q] - v[0][q] : v[2][q];
    }
  }
  if (R == 2 && d == 1 && g.walls && (j < 2 || j >= g.ny - 2)) {
    // rows 0, 1 at the low wall and ny-2, ny-1 at the high wall; each
    // replacement reads only entries that no earlier one changes
    const bool l0 = j == 0, l1 = j == 1, h1 = j == g.ny - 2, h0 = j == g.ny - 1;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      v[1][q] = l0 ? 2.0f * v[2][q] - v[3][q] : v[1][q];
      v[0][q] = l0 ? v[2][q] : (l1 ? 2.0f * v[1][q] - v[2][q] : v[0][q]);
      v[3][q] = h0 ? 2.0f * v[2][q] - v[1][q] : v[3][q];
      v[4][q] = h0 ? v[2][q] : (h1 ? 2.0f * v[3][q] - v[2][q] : v[4][q]);
    }
  }
}

// Stencil kernels come in two instances.  GEN = false covers the cells whose
// stencils of radius R stay clear of the y-wall ghosts (every cell of a
// periodic grid) and reads neighbours through wrapped offsets only;
// GEN = true covers the 2R rows next to the walls through the boundary-rule
// accessor.  Separate kernels keep the bulk kernel's register budget, and so
// its occupancy, free of the wall path.  The bulk instance walks the cells
// in band order (band_cell).
template <bool GEN, int R>
__device__ __forceinline__ bool row_cell(const Grid& g, long& m, int& i, int& j, int& k) {
  if (!GEN) {
    const long L = xcd_cell();
    if (L >= g.n) return false;
    band_cell(g, L, m, i, j, k);
    return !g.walls || (j >= R && j < g.ny - R);
  }
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long per = (long)g.nx * 2 * R;
  if (t >= per * g.nz) return false;
  i = (int)(t % g.nx);
  const int r = (int)((t / g.nx) % (2 * R));
  k = (int)(t / per);
  j = r < R ? r : g.ny - 2 * R + r;
  m = cidx(g, i, j, k);
  return true;
}

// Uncertified (cell, axis) pairs go to a list in per-block regions of 3
// entries per cell of the block (block b owns entries [region b, region b +
// cnt[b]))): the slots are taken with LDS atomics, so the kernel issues no
// device-scope atomic.  The first `cap` entries of a block also keep the 18
// values the selection read (the three cells' six components, plane-major in
// the block's staging area of 18 cap values), so that their re-decision
// reads them contiguously instead of gathering them again; later entries
// gather.
#define SEL_BLOCK 256
#define SEL_REGION (3 * SEL_BLOCK)
#define SEL_STAGE 64

// one LDS atomic per wavefront: the first unsure lane reserves the slots.
// Called by every lane of the wavefront.
__device__ __forceinline__ void sel_record(bool unsure, unsigned entry, const float (&v)[3][6], unsigned* s_cnt,
                                           unsigned* __restrict__ region, float* __restrict__ stage, unsigned cap) {
  const unsigned long long ball = __ballot(unsure);
  if (ball) {
    const int lane = __lane_id();
    const int leader = __ffsll((long long)ball) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(s_cnt, (unsigned)__popcll(ball));
    base = __shfl(base, leader);
    if (unsure) {
      const unsigned slot = base + (unsigned)__popcll(ball & ((1ull << lane) - 1ull));
      region[slot] = entry;
      if (slot < cap) {
#pragma unroll
        for (int o = 0; o < 3; o++)
#pragma unroll
          for (int q = 0; q < 6; q++) stage[(o * 6 + q) * cap + slot] = v[o][q];
      }
    }
  }
}

// The block's entries join the packed list of all uncertified entries: one
// device-scope atomic per block reserves [base, base + count), and the
// places of the block's entries in amb_list (block * region + slot) go there,
// so that the FP64 re-decision walks one dense list.  Called by every thread
// of the block once its entries are recorded.
__device__ __forceinline__ void sel_publish(unsigned* s_cnt, unsigned* s_base, long blk, unsigned region,
                                            unsigned* __restrict__ total, unsigned* __restrict__ idx) {
  __syncthreads();
  const unsigned c = *s_cnt;
  if (threadIdx.x == 0) *s_base = c ? atomicAdd(total, c) : 0u;
  __syncthreads();
  const unsigned base = *s_base;
  for (unsigned sl = threadIdx.x; sl < c; sl += blockDim.x) idx[base + sl] = (unsigned)blk * region + sl;
}

// One cell per thread, any grid shape.
__global__ __launch_bounds__(SEL_BLOCK) void m_select(Grid g, const float* __restrict__ C, uint8_t* __restrict__ sel,
                                                      unsigned* __restrict__ total, unsigned* __restrict__ amb_idx,
                                                      unsigned* __restrict__ amb_list, float* __restrict__ amb_data) {
  __shared__ unsigned s_cnt, s_base;
  if (threadIdx.x == 0) s_cnt = 0u;
  __syncthreads();
  long m;
  int i, j, k;
  const bool valid = row_cell<false, 0>(g, m, i, j, k);   // the wall rows too (see load_line)
  unsigned* region = amb_list + (size_t)blockIdx.x * SEL_REGION;
  float* stage = amb_data + (size_t)blockIdx.x * 18 * SEL_STAGE;
  uint8_t s = 0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    bool unsure = false;
    int idx = 0;
    float v[3][6];
    if (valid) {
      load_line<1>(C, g, m, i, j, k, d, v);
      const float ein = SEL_KCE * fmaxf(fmaxf(entry_bound(v[0]), entry_bound(v[1])), entry_bound(v[2]));
      Lmin P[3], N[3];
#pragma unroll
      for (int l = 0; l < 3; l++) { P[l] = cand_lmin(false, l, v); N[l] = cand_lmin(true, l, v); }
      idx = sel_decide(P, N, ein, unsure);
    }
    sel_record(unsure, (unsigned)(m * 3 + d), v, &s_cnt, region, stage, SEL_STAGE);
    s |= (uint8_t)(idx << (2 * d));
  }
  if (valid) sel[m] = s;
  sel_publish(&s_cnt, &s_base, blockIdx.x, SEL_REGION, total, amb_idx);
}

// The selection streamed along z, for nx = 64, 128 or 256 and nz a multiple
// of SZ_K: a block of 256 threads holds 256/nx whole x-rows and walks SZ_K
// planes.  The midpoint candidate of a face, C_a/2 + C_{a+1}/2, is the right
// face's candidate of level 0 for cell a and the left face's of level 1 for
// cell a+1 (fenep.f90:179-184), bit for bit.  Along x a thread evaluates it
// for its right face and hands it to its neighbour through LDS (the rows are
// whole, so the periodic neighbour is in the block); along z it carries it
// to the next plane in registers.  16 of the 18 eigenvalue evaluations of a
// cell remain, plus one per column for its first plane, and the z line is
// loaded one plane per step.
#define SZ_K 4
__global__ __launch_bounds__(SEL_BLOCK) void m_select_z(Grid g, const float* __restrict__ C, uint8_t* __restrict__ sel,
                                                        unsigned* __restrict__ total, unsigned* __restrict__ amb_idx,
                                                        unsigned* __restrict__ amb_list, float* __restrict__ amb_data) {
  __shared__ unsigned s_cnt, s_base;
  __shared__ float s_lam[2][SEL_BLOCK], s_e0[2][SEL_BLOCK];
  __shared__ int s_amb[2][SEL_BLOCK];
  const int t = threadIdx.x;
  if (t == 0) s_cnt = 0u;
  const long xb = xcd_block();
  const int nkg = g.nz / SZ_K;
  const int kg = (int)(xb % nkg), jg = (int)(xb / nkg);
  const int i = t & (g.nx - 1);
  const int j = (jg << (8 - g.lnx)) + (t >> g.lnx);
  const int tl = (t & ~(g.nx - 1)) | ((i - 1) & (g.nx - 1));   // the x neighbour on the left
  const int k0 = kg * SZ_K;
  const unsigned cap = SEL_STAGE * SZ_K;
  unsigned* region = amb_list + (size_t)blockIdx.x * (SEL_REGION * SZ_K);
  float* stage = amb_data + (size_t)blockIdx.x * 18 * cap;
  const long n = g.n;
  const int sz = g.nx * g.ny;
  unsigned m = (unsigned)cidx(g, i, j, k0);
  const int exm = woff32(i, -1, g.nx, 1), exp = woff32(i, 1, g.nx, 1);
  const int eym = woff32(j, -1, g.ny, g.nx), eyp = woff32(j, 1, g.ny, g.nx);
  const bool ylo = g.walls && j == 0, yhi = g.walls && j == g.ny - 1;
  // the z line (k-1, k, k+1) and the entry bounds of its cells
  float z[3][6];
  {
    const unsigned e = m + (unsigned)woff32(k0, -1, g.nz, sz);
#pragma unroll
    for (int q = 0; q < 6; q++) {
      z[0][q] = ldu(C + q * n, n, e);
      z[1][q] = ldu(C + q * n, n, m);
    }
  }
  float ebm = entry_bound(z[0]), ebc = entry_bound(z[1]);
  // Cp_1 along z of the first plane; later planes take the previous Cn_0
  Lmin zl = cand_lmin(false, 1, z);
  __syncthreads();
  for (int kk = 0; kk < SZ_K; kk++) {
    const int par = kk & 1;
    {
      const unsigned e = m + (unsigned)woff32(k0 + kk, 1, g.nz, sz);
#pragma unroll
      for (int q = 0; q < 6; q++) z[2][q] = ldu(C + q * n, n, e);
    }
    const float ebp = entry_bound(z[2]);
    // x: the right-face midpoint first, for the neighbour
    float x[3][6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
      x[0][q] = ldu(C + q * n, n, m + (unsigned)exm);
      x[1][q] = z[1][q];
      x[2][q] = ldu(C + q * n, n, m + (unsigned)exp);
    }
    Lmin XP[3], XN[3];
    XN[0] = cand_lmin(true, 0, x);
    s_lam[par][t] = XN[0].lam;
    s_e0[par][t] = XN[0].e0;
    s_amb[par][t] = XN[0].amb;
    XP[0] = cand_lmin(false, 0, x);
    XN[1] = cand_lmin(true, 1, x);
    XP[2] = cand_lmin(false, 2, x);
    XN[2] = cand_lmin(true, 2, x);
    const float einx = SEL_KCE * fmaxf(fmaxf(entry_bound(x[0]), ebc), entry_bound(x[2]));
    uint8_t sb = 0;
    // z: Cp_1 carried in
    {
      Lmin P[3], N[3];
      P[1] = zl;
      P[0] = cand_lmin(false, 0, z);
      N[0] = cand_lmin(true, 0, z);
      N[1] = cand_lmin(true, 1, z);
      P[2] = cand_lmin(false, 2, z);
      N[2] = cand_lmin(true, 2, z);
      bool unsure = false;
      const int idx = sel_decide(P, N, SEL_KCE * fmaxf(fmaxf(ebm, ebc), ebp), unsure);
      sel_record(unsure, m * 3u + 2u, z, &s_cnt, region, stage, cap);
      sb |= (uint8_t)(idx << 4);
      zl = N[0];
    }
    // y: C_kabe ghosts at the walls (as load_line)
    {
      float y[3][6];
#pragma unroll
      for (int q = 0; q < 6; q++) {
        y[0][q] = ldu(C + q * n, n, m + (unsigned)eym);
        y[1][q] = z[1][q];
        y[2][q] = ldu(C + q * n, n, m + (unsigned)eyp);
      }
      if (ylo || yhi) {   // a wavefront holds one row: interior rows skip this
#pragma unroll
        for (int q = 0; q < 6; q++) {
          y[0][q] = ylo ? 2.0f * y[1][q] - y[2][q] : y[0][q];
          y[2][q] = yhi ? 2.0f * y[1][q] - y[0][q] : y[2][q];
        }
      }
      Lmin P[3], N[3];
#pragma unroll
      for (int l = 0; l < 3; l++) { P[l] = cand_lmin(false, l, y); N[l] = cand_lmin(true, l, y); }
      bool unsure = false;
      const int idx = sel_decide(P, N, SEL_KCE * fmaxf(fmaxf(entry_bound(y[0]), ebc), entry_bound(y[2])), unsure);
      sel_record(unsure, m * 3u + 1u, y, &s_cnt, region, stage, cap);
      sb |= (uint8_t)(idx << 2);
    }
    __syncthreads();
    XP[1].lam = s_lam[par][tl];
    XP[1].e0 = s_e0[par][tl];
    XP[1].amb = s_amb[par][tl] != 0;
    {
      bool unsure = false;
      const int idx = sel_decide(XP, XN, einx, unsure);
      sel_record(unsure, m * 3u, x, &s_cnt, region, stage, cap);
      sb |= (uint8_t)idx;
    }
    sel[m] = sb;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      z[0][q] = z[1][q];
      z[1][q] = z[2][q];
    }
    ebm = ebc;
    ebc = ebp;
    m += (unsigned)sz;
  }
  sel_publish(&s_cnt, &s_base, blockIdx.x, SEL_REGION * SZ_K, total, amb_idx);
}

// the reference's selection (select_d, evaluated as written) of one (cell,
// axis) of the stored FP32 field, wall ghosts by the C_kabe rule
template <bool WL>
__device__ __forceinline__ int select_ref(const Grid& g, const float* __restrict__ C, long m, int d) {
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  const int di = d == 0, dj = d == 1, dk = d == 2;
  double cm[6], c[6], cp[6];
  for (int q = 0; q < 6; q++) {
    const float* Cq = C + q * g.n;
    cm[q] = (double)ldbw<WL>(Cq, g, i - di, j - dj, k - dk, BC_C);
    c[q] = (double)Cq[m];
    cp[q] = (double)ldbw<WL>(Cq, g, i + di, j + dj, k + dk, BC_C);
  }
  return select_d(cm, c, cp);
}

// The reference's selection of one (cell, axis), evaluating only the
// candidates that can decide it.  The FP32 intervals of the six candidates
// are recomputed as m_select makes them.  A candidate whose interval lies
// above its partner's (the other face of its level) is not the minimum of its
// level, and a level whose interval lies below another option's (another
// level, or the zero slope) is not the maximum; the reference values keep
// these strict orders, so the first maximum over the evaluated options is
// select_d's answer.  A candidate near the sentinel boundary bounds nothing:
// its level is evaluated in full.  Each lane walks its own needed candidates
// (2.4 of 6 on average in the throughput case), one reference Vieta per trip.
// The three cells of (cell m, axis d), wall ghosts by the C_kabe rule (the
// values m_select reads through load_line).
template <bool WL>
__device__ __forceinline__ void sel_cells(const Grid& g, const float* __restrict__ C, long m, int d, float (&fm)[6],
                                          float (&fc)[6], float (&fp)[6]) {
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  const int di = d == 0, dj = d == 1, dk = d == 2;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const float* Cq = C + q * g.n;
    fm[q] = ldbw<WL>(Cq, g, i - di, j - dj, k - dk, BC_C);
    fc[q] = Cq[m];
    fp[q] = ldbw<WL>(Cq, g, i + di, j + dj, k + dk, BC_C);
  }
}

__device__ __forceinline__ int select_pruned(const float (&fm)[6], const float (&fc)[6], const float (&fp)[6]) {
  const float ein = SEL_KCE * fmaxf(fmaxf(entry_bound(fm), entry_bound(fc)), entry_bound(fp));
  // one level per trip (a rolled loop keeps the register count, and so the
  // occupancy that hides the FP64 latencies, down)
  float lam[6], err[6];
  bool amb[3];
#pragma unroll 1
  for (int l = 0; l < 3; l++) {
    float ap[6], an[6];
#pragma unroll
    for (int q = 0; q < 6; q++) { ap[q] = cand_pf(l, fm[q], fc[q], fp[q]); an[q] = cand_nf(l, fm[q], fc[q], fp[q]); }
    const Lmin P = lmin32(ap), N = lmin32(an);
#pragma unroll
    for (int u = 0; u < 3; u++)
      if (u == l) {
        lam[2 * u] = P.lam; err[2 * u] = P.e0 + ein;
        lam[2 * u + 1] = N.lam; err[2 * u + 1] = N.e0 + ein;
        amb[u] = P.amb || N.amb;
      }
  }
  float lo[4], hi[3];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    const float v = fminf(lam[2 * l], lam[2 * l + 1]), e = fmaxf(err[2 * l], err[2 * l + 1]);
    lo[l] = amb[l] ? -INFINITY : v - e;
    hi[l] = amb[l] ? INFINITY : v + e;
  }
  lo[3] = 0.0f;
  unsigned need = 0u;
#pragma unroll
  for (int l = 0; l < 3; l++) {
    float other = -INFINITY;
#pragma unroll
    for (int o = 0; o < 4; o++)
      if (o != l) other = fmaxf(other, lo[o]);
    if (!(other > hi[l])) {
      const bool pP = amb[l] || !(lam[2 * l] - err[2 * l] > lam[2 * l + 1] + err[2 * l + 1]);
      const bool pN = amb[l] || !(lam[2 * l + 1] - err[2 * l + 1] > lam[2 * l] + err[2 * l]);
      need |= (pP ? 1u : 0u) << (2 * l);
      need |= (pN ? 1u : 0u) << (2 * l + 1);
    }
  }
  // Candidate t = 2 l + (right face) as (wa ua) sa + z + (wc uc) sc, the
  // terms of cand_p / cand_n in their order ((c 3)/2 as (3 c) 0.5: the same
  // rounding).  Levels are finished in increasing order, at their last
  // needed candidate, which keeps select_d's first maximum.
  int idx = -1;
  double best = 0.0, vleft = 0.0;
  unsigned rest = need;
  while (rest) {
    const int t = __builtin_ctz(rest);
    rest &= rest - 1u;
    const bool a_c = t < 2, c_c = t == 2 || t == 3, three = t >= 4;
    const double wa = t == 0 ? 3.0 : (t == 3 || t == 5 ? -1.0 : 1.0);
    const double wc = t == 3 ? 3.0 : (t == 0 || t == 4 ? -1.0 : 1.0);
    const double sh = t >= 4 ? 0.25 : 0.5;
    double a[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
#pragma clang fp contract(off)
      const double cq = (double)fc[q];
      const double A = (wa * (a_c ? cq : (double)fm[q])) * sh;
      const double Cc = (wc * (c_c ? cq : (double)fp[q])) * sh;
      a[q] = (A + (three ? cq : -0.0)) + Cc;
    }
    const double v = vieta_d(a);
    const bool right = t & 1;
    const bool pair = (need >> (t ^ 1)) & 1u;   // the other face of the level is needed too
    if (!right && pair) {
      vleft = v;
      continue;
    }
    const double lv = right && pair ? minval2_d(vleft, v) : v;
    if (!isnan(lv) && (idx < 0 || lv > best)) { idx = t >> 1; best = lv; }
  }
  if (idx < 0 || 0.0 > best) idx = 3;
  return idx;
}

// FP64 re-decision of the uncertified selections: grid-stride over the
// packed list; an entry's 18 input values come from its select block's
// staging area for the block's first `cap` entries and are gathered for the
// rest
template <bool WL>
__global__ __launch_bounds__(64) void m_select_fix(Grid g, const float* __restrict__ C, uint8_t* __restrict__ sel,
                                                    const unsigned* __restrict__ total,
                                                    const unsigned* __restrict__ amb_idx,
                                                    const unsigned* __restrict__ amb_list,
                                                    const float* __restrict__ amb_data, unsigned region, unsigned cap) {
  const long cnt = (long)*total;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < cnt; t += (long)gridDim.x * blockDim.x) {
    const unsigned at = amb_idx[t];
    const unsigned blk = at / region, slot = at - blk * region;
    const unsigned e = amb_list[at];
    const long m = (long)(e / 3u);
    const int d = (int)(e % 3u);
    float fm[6], fc[6], fp[6];
    if (slot < cap) {
      const float* src = amb_data + (size_t)blk * 18 * cap + slot;
#pragma unroll
      for (int q = 0; q < 6; q++) {
        fm[q] = src[q * cap];
        fc[q] = src[(6 + q) * cap];
        fp[q] = src[(12 + q) * cap];
      }
    } else {
      sel_cells<WL>(g, C, m, d, fm, fc, fp);
    }
    const int idx = select_pruned(fm, fc, fp);
    // the byte holds three axes: update only this axis
    unsigned* word = (unsigned*)((size_t)(sel + m) & ~(size_t)3);
    const int sh = 8 * (int)((size_t)(sel + m) & 3) + 2 * d;
    const unsigned mask = 3u << sh;
    unsigned old = *word, assumed;
    do {
      assumed = old;
      const unsigned nv = (assumed & ~mask) | ((unsigned)idx << sh);
      old = atomicCAS(word, assumed, nv);
    } while (old != assumed);
  }
}

// count of (cell, axis) whose selection differs from the reference's
// selection of the stored field
__global__ __launch_bounds__(256) void m_select_check(Grid g, const float* __restrict__ C,
                                                      const uint8_t* __restrict__ sel, unsigned* __restrict__ bad) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  unsigned nb = 0;
  if (m < g.n) {
    const unsigned s = sel[m];
    for (int d = 0; d < 3; d++)
      nb += (unsigned)((g.walls ? select_ref<true>(g, C, m, d) : select_ref<false>(g, C, m, d)) != (int)((s >> (2 * d)) & 3u));
  }
  for (int o = 32; o > 0; o >>= 1) nb += __shfl_xor(nb, o);
  if (__lane_id() == 0 && nb) atomicAdd(bad, nb);
}

// Peterlin factor: FP32 unless tr C is within 1% of L^2 (then FP64)
__device__ __forceinline__ float peterlin_f(float L2, float s) {
  if (s > 0.99f * L2) return (float)(((double)L2 - 3.0) / ((double)L2 - (double)s));
  return (L2 - 3.0f) / (L2 - s);
}
__device__ __forceinline__ float dpeterlin_f(float L2, float s) {
  if (s > 0.99f * L2) {
    double den = (double)L2 - (double)s;
    return (float)(((double)L2 - 3.0) / (den * den));
  }
  float den = L2 - s;
  return (L2 - 3.0f) / (den * den);
}

// wrapped neighbour offsets of one cell (periodic in every axis; callers use
// the boundary-rule accessor for wall rows)
struct Nbr {
  long x[3], y[3], z[3];
};
__device__ __forceinline__ Nbr nbr(const Grid& g, int i, int j, int k) {
  Nbr o;
  const long sy = g.nx, sz = (long)g.nx * g.ny;
#pragma unroll
  for (int d = -1; d <= 1; d++) {
    o.x[d + 1] = woff(i, d, g.nx, 1);
    o.y[d + 1] = woff(j, d, g.ny, sy);
    o.z[d + 1] = woff(k, d, g.nz, sz);
  }
  return o;
}

// coefficients (of c_{a-1}, c_a, c_{a+1}) of the candidate face values of
// fenep.f90:179-184: left face Cp_l and right face Cn_l, l = 3 the zero slope
template <typename T>
__device__ __forceinline__ void cand_coef_p(int l, T& a, T& b, T& c) {
  a = l == 1 ? T(0.5) : (l == 2 ? T(0.25) : T(0));
  b = l == 0 ? T(1.5) : (l == 1 ? T(0.5) : T(1));
  c = l == 0 ? T(-0.5) : (l == 2 ? T(-0.25) : T(0));
}
template <typename T>
__device__ __forceinline__ void cand_coef_n(int l, T& a, T& b, T& c) {
  a = l == 1 ? T(-0.5) : (l == 2 ? T(-0.25) : T(0));
  b = l == 0 ? T(0.5) : (l == 1 ? T(1.5) : T(1));
  c = l == 0 ? T(0.5) : (l == 2 ? T(0.25) : T(0));
}
__device__ __forceinline__ float rcp_t(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double rcp_t(double v) { return 1.0 / v; }
// Peterlin factor and its derivative: in FP64 within 1 % of L^2, otherwise
// through the hardware reciprocal (within an ulp; the IEEE division costs
// about ten instructions in this VALU-bound kernel)
template <typename T>
__device__ __forceinline__ T peterlin_t(T L2, T s) {
  if (s > T(0.99) * L2) return (T)(((double)L2 - 3.0) / ((double)L2 - (double)s));
  return (L2 - T(3)) * rcp_t(L2 - s);
}
template <typename T>
__device__ __forceinline__ T dpeterlin_t(T L2, T s) {
  if (s > T(0.99) * L2) {
    const double den = (double)L2 - (double)s;
    return (T)(((double)L2 - 3.0) / (den * den));
  }
  const T r = rcp_t(L2 - s);
  return (L2 - T(3)) * (r * r);
}

// Cstar + Lyapunov Newton (analytic Jacobian) + Peterlin factor of the new C
// for one cell, in the arithmetic type T (storage stays FP32; the step uses
// FP32: an FP64 evaluation of the strongly stretched cells was measured to
// leave the error against the FP64 step unchanged).
//
// Increment form: C* = C^n + D* and C^{n+1} = C^n + delta, with D* and the
// Newton residual summed from O(dt) terms only, so that C is rounded to FP32
// once per step.  Cstar's weights hold 1/6 on the own cell's two faces
// (fenep.f90:484-497); both faces of every candidate sum to 2 C^n, so those
// parts add up to C^n exactly and D* keeps only the velocity-dependent
// parts.  (Summed as written in FP32, 6 fl(1/6) = 1 + 3e-8 multiplies C by
// that factor every step: a drift of 3e-5 in 1000 steps where C ~ I.)
template <typename T>
__device__ __forceinline__ void cstar_newton_cell(const Grid& g, const float* __restrict__ C,
                                                  const uint8_t* __restrict__ sel, const float* __restrict__ U,
                                                  const float* __restrict__ V, const float* __restrict__ W,
                                                  float* __restrict__ Cn, float* __restrict__ fC, int* __restrict__ nitr,
                                                  T eps, int itrmax, long m, int i, int j, int k) {
  const T dt = (T)g.dt;
  // 32-bit neighbour offsets and cell index (n <= 2^30 cells)
  struct {
    int x[3], y[3], z[3];
  } o;
#pragma unroll
  for (int d = -1; d <= 1; d++) {
    o.x[d + 1] = woff32(i, d, g.nx, 1);
    o.y[d + 1] = woff32(j, d, g.ny, g.nx);
    o.z[d + 1] = woff32(k, d, g.nz, g.nx * g.ny);
  }
  const unsigned mu = (unsigned)m;
  // the two wall rows: the values the boundary rules give replace the
  // wrapped rows in the few reads that reach a ghost row or the wall face of V
  const bool wlo = g.walls && j == 0, whi = g.walls && j == g.ny - 1;
  T cx[6];
#pragma unroll
  for (int q = 0; q < 6; q++) cx[q] = T(0);
#pragma unroll 1
  for (int d = 0; d < 3; d++) {
    const T h = (T)(d == 0 ? g.dx : (d == 1 ? g.dy : g.dz));
    T v1, v2;
    if (d == 0) { v1 = ldu(U, g.n, mu + o.x[0]); v2 = ldu(U, g.n, mu); }
    else if (d == 1) { v1 = wlo ? T(0) : (T)ldu(V, g.n, mu + o.y[0]); v2 = whi ? T(0) : (T)ldu(V, g.n, mu); }   // V_kabe
    else { v1 = ldu(W, g.n, mu + o.z[0]); v2 = ldu(W, g.n, mu); }
    const T r2h = dt * (T(0.5) * rcp_t(h));
    // the weights less their 1/6 parts (s2 and s3 of the reference)
    const T s1 = (-v2 + fabs(v2)) * r2h;
    const T s2 = (-v2 - fabs(v2)) * r2h;
    const T s3 = (v1 - fabs(v1)) * r2h;
    const T s4 = (v1 + fabs(v1)) * r2h;
    const int om = d == 0 ? o.x[0] : (d == 1 ? o.y[0] : o.z[0]);
    const int op = d == 0 ? o.x[2] : (d == 1 ? o.y[2] : o.z[2]);
    const int lm = (ldu(sel, g.n, mu + om) >> (2 * d)) & 3, l0 = (ldu(sel, g.n, mu) >> (2 * d)) & 3,
              lp = (ldu(sel, g.n, mu + op) >> (2 * d)) & 3;
    // The four face values are linear in the line c0..c4:
    //   pl = Cp_l0(c1,c2,c3)  own left face      nr = Cn_l0(c1,c2,c3) own right face
    //   pr = Cp_lp(c2,c3,c4)  left face of a+1   nl = Cn_lm(c0,c1,c2) right face of a-1
    // so s3 pl + s1 pr + s4 nl + s2 nr = sum_p w_p c_p with five weights per
    // axis shared by the six components (the 1/6 parts, 1/6 (pl + nr) =
    // c2 / 3 per axis, make the C^n of C* = C^n + D*).
    T ap, bp, gp_, an, bn, gn, aq, bq, gq, am, bm, gm;
    cand_coef_p<T>(l0, ap, bp, gp_);
    cand_coef_n<T>(l0, an, bn, gn);
    cand_coef_p<T>(lp, aq, bq, gq);
    cand_coef_n<T>(lm, am, bm, gm);
    const T w0 = s4 * am;
    const T w1 = s3 * ap + s2 * an + s4 * bm;
    const T w2 = s3 * bp + s2 * bn + s1 * aq + s4 * gm;
    const T w3 = s3 * gp_ + s2 * gn + s1 * bq;
    const T w4 = s1 * gq;
    float v[5][6];
    load_line<2>(C, g, m, i, j, k, d, v);
#pragma unroll
    for (int q = 0; q < 6; q++)
      cx[q] += w0 * (T)v[0][q] + w1 * (T)v[1][q] + w2 * (T)v[2][q] + w3 * (T)v[3][q] + w4 * (T)v[4][q];
  }
  const T dtw = dt * rcp_t((T)g.wi);
  // right-hand side less C^n: D* + (dt/Wi) I
  const T rhs[6] = {cx[0] + dtw, cx[1], cx[2], cx[3] + dtw, cx[4], cx[5] + dtw};
  const T dx = (T)g.dx, dy = (T)g.dy, dz = (T)g.dz;
  // A = transpose(A0) = -dt L with L_ij = du_i/dx_j (fenep.f90:541-550)
  T a11, a12, a13, a21, a22, a23, a31, a32, a33;
  {
#define UR(a, bb, c) ((T)ldu(U, g.n, mu + (unsigned)(o.x[a] + o.y[bb] + o.z[c])))
#define VR(a, bb, c) ((T)ldu(V, g.n, mu + (unsigned)(o.x[a] + o.y[bb] + o.z[c])))
#define WR(a, bb, c) ((T)ldu(W, g.n, mu + (unsigned)(o.x[a] + o.y[bb] + o.z[c])))
    // the boundary rules (U_kabe, V_kabe, W_kabe): U, W ghosts -v below and
    // 2*0 - v above; V zero at a ghost row and at the wall face j = ny-1
#define UF(a, bb, c) ((bb) == 0 && wlo ? -UR(a, 1, c) : ((bb) == 2 && whi ? T(2) * T(0) - UR(a, 1, c) : UR(a, bb, c)))
#define WF(a, bb, c) ((bb) == 0 && wlo ? -WR(a, 1, c) : ((bb) == 2 && whi ? T(2) * T(0) - WR(a, 1, c) : WR(a, bb, c)))
#define VF(a, bb, c) (((bb) == 0 && wlo) || ((bb) == 1 && whi) ? T(0) : VR(a, bb, c))
    // the grid spacings through reciprocals (no division per term)
    const T idx = rcp_t(dx), idy = rcp_t(dy), idz = rcp_t(dz);
    const T qdx = T(0.25) * idx, qdy = T(0.25) * idy, qdz = T(0.25) * idz;
#define GRAD(UU, VV, WW)                                                                        \
  a11 = -dt * (UU(1, 1, 1) - UU(0, 1, 1)) * idx;                                                \
  a21 = -dt * ((VV(2, 0, 1) - VV(0, 0, 1)) + (VV(2, 1, 1) - VV(0, 1, 1))) * qdx;                \
  a31 = -dt * ((WW(2, 1, 0) - WW(0, 1, 0)) + (WW(2, 1, 1) - WW(0, 1, 1))) * qdx;                \
  a12 = -dt * ((UU(0, 2, 1) + UU(1, 2, 1)) - (UU(0, 0, 1) + UU(1, 0, 1))) * qdy;                \
  a22 = -dt * (VV(1, 1, 1) - VV(1, 0, 1)) * idy;                                                \
  a32 = -dt * ((WW(1, 2, 0) - WW(1, 0, 0)) + (WW(1, 2, 1) - WW(1, 0, 1))) * qdy;                \
  a13 = -dt * ((UU(0, 1, 2) + UU(1, 1, 2)) - (UU(0, 1, 0) + UU(1, 1, 0))) * qdz;                \
  a23 = -dt * ((VV(1, 0, 2) + VV(1, 1, 2)) - (VV(1, 0, 0) + VV(1, 1, 0))) * qdz;                \
  a33 = -dt * (WW(1, 1, 1) - WW(1, 1, 0)) * idz;
    // the interior rows read plainly; the two wall rows (a wavefront shares
    // its row) again with the boundary rules
    if (wlo || whi) {
      GRAD(UF, VF, WF)
    } else {
      GRAD(UR, VR, WR)
    }
#undef GRAD
#undef UF
#undef VF
#undef WF
#undef UR
#undef VR
#undef WR
  }
  const T L2 = (T)(g.lp * g.lp);
  T c0[6], x[6], dl[6];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    c0[q] = ldu(C + q * g.n, g.n, mu);
    x[q] = c0[q];
    dl[q] = T(0);
  }
  int it;
  // Chord iterations: the Jacobian of the first iterate (the linear operator
  // plus the rank-one Peterlin term) is factored once, its multipliers kept
  // below the diagonal; the second iteration takes a fresh residual through
  // the same factors.  The Jacobian moves with x only through f and f', by
  // about 1e-5 relative between iterates, so the chord converges like Newton
  // here (two iterations in the throughput case) at half the cost of the
  // second.  Near the pole (tr C -> L^2) f varies fast with tr C and the
  // chord can stall: from the third iteration on the Jacobian is factored
  // again at the current iterate (full Newton).
  //
  // The LU takes no pivots.  J = J0 + gp x e^T: J0, the linear operator, has
  // diagonal entries 1 + f dt/Wi + O(dt |grad u|) and off-diagonal entries
  // O(dt |grad u|); the Peterlin term adds gp x_r (gp = dt/Wi f' > 0) to
  // columns 0, 3, 5 of row r, which on the diagonal of the rows 0, 3, 5 is
  // gp times a diagonal entry of the SPD iterate, > 0.  Eliminating a column
  // subtracts from a later diagonal pivot at most the part of its own
  // rank-one entry that the earlier rows carry, so every pivot stays above
  // 1 + f dt/Wi - O(dt |grad u|).
  T J[6][6];
  for (it = 1; it <= itrmax; it++) {
    const T s = x[0] + x[3] + x[5];
    const T f = peterlin_t<T>(L2, s);
    const T dh = T(0.5) * (f * dtw);   // the diagonal less its 1/2
    const T B11 = a11 + dh, B22 = a22 + dh, B33 = a33 + dh;
    // residual l = M x - b (fenep.f90:598-603) from O(dt) terms: M = I + M'
    // and b = C^n + rhs, so l = M' x + (delta - rhs)
    T l[6];
    l[0] = T(2) * (B11 * x[0] + a12 * x[1] + a13 * x[2]) + (dl[0] - rhs[0]);
    l[1] = a21 * x[0] + (B22 + B11) * x[1] + a23 * x[2] + a12 * x[3] + a13 * x[4] + (dl[1] - rhs[1]);
    l[2] = a31 * x[0] + a32 * x[1] + (B33 + B11) * x[2] + a12 * x[4] + a13 * x[5] + (dl[2] - rhs[2]);
    l[3] = T(2) * (a21 * x[1] + B22 * x[3] + a23 * x[4]) + (dl[3] - rhs[3]);
    l[4] = a31 * x[1] + a21 * x[2] + a32 * x[3] + (B33 + B22) * x[4] + a23 * x[5] + (dl[4] - rhs[4]);
    l[5] = T(2) * (a31 * x[2] + a32 * x[4] + B33 * x[5]) + (dl[5] - rhs[5]);
    if (it == 1 || it >= 3) {
      const T A11 = B11 + T(0.5), A22 = B22 + T(0.5), A33 = B33 + T(0.5);
      // Jacobian: linear operator + rank one Peterlin term
      const T gp = dtw * dpeterlin_t<T>(L2, s);
      const T J0[6][6] = {
          {T(2) * A11, T(2) * a12, T(2) * a13, T(0), T(0), T(0)},
          {a21, A22 + A11, a23, a12, a13, T(0)},
          {a31, a32, A33 + A11, T(0), a12, a13},
          {T(0), T(2) * a21, T(0), T(2) * A22, T(2) * a23, T(0)},
          {T(0), a31, a21, a32, A33 + A22, a23},
          {T(0), T(0), T(2) * a31, T(0), T(2) * a32, T(2) * A33}};
#pragma unroll
      for (int r = 0; r < 6; r++) {
        const T gx = gp * x[r];
#pragma unroll
        for (int cc = 0; cc < 6; cc++) J[r][cc] = J0[r][cc] + ((cc == 0 || cc == 3 || cc == 5) ? gx : T(0));
      }
      // LU without pivoting (diagonally dominant): multipliers below the
      // diagonal, reciprocal pivots on it
#pragma unroll
      for (int kk = 0; kk < 6; kk++) {
        const T inv = rcp_t(J[kk][kk]);
#pragma unroll
        for (int r = kk + 1; r < 6; r++) {
          const T fct = J[r][kk] * inv;
#pragma unroll
          for (int cc = kk + 1; cc < 6; cc++) J[r][cc] -= fct * J[kk][cc];
          J[r][kk] = fct;
        }
        J[kk][kk] = inv;
      }
    }
#pragma unroll
    for (int kk = 0; kk < 6; kk++)
#pragma unroll
      for (int r = kk + 1; r < 6; r++) l[r] -= J[r][kk] * l[kk];
    T rr[6];
#pragma unroll
    for (int r = 5; r >= 0; r--) {
      T acc = l[r];
#pragma unroll
      for (int cc = r + 1; cc < 6; cc++) acc -= J[r][cc] * rr[cc];
      rr[r] = acc * J[r][r];
    }
    T sr = T(0), sx = T(0);
#pragma unroll
    for (int q = 0; q < 6; q++) {
      dl[q] -= rr[q];
      x[q] = c0[q] + dl[q];
      sr += rr[q] * rr[q];
      sx += x[q] * x[q];
    }
    if (sr < eps * eps * sx) break;
  }
  if (it > itrmax) it = -itrmax;   // stopped at the cap without meeting eps
#pragma unroll
  for (int q = 0; q < 6; q++) stu(Cn + q * g.n, g.n, mu, (float)x[q]);
  stu(fC, g.n, mu, (float)peterlin_t<T>(L2, x[0] + x[3] + x[5]));
  stu(nitr, g.n, mu, it);
}

// one instance over every cell (the wall rows through the boundary-rule
// patches in cstar_newton_cell and load_line), in band order.  The kernel
// waits on its loads more than it issues: it is held to five waves per SIMD
// (96 registers, no spills; 145 -> 140 us at 256x128x128 against four waves;
// six waves spill).  Blocks of 256 threads: 2 us less per step than 128, 64
// slower
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) void m_cstar_newton(Grid g, const float* __restrict__ C, const uint8_t* __restrict__ sel,
                                                      const float* __restrict__ U, const float* __restrict__ V,
                                                      const float* __restrict__ W, float* __restrict__ Cn,
                                                      float* __restrict__ fC, int* __restrict__ nitr, float eps,
                                                      int itrmax) {
  long m;
  int i, j, k;
  if (!row_cell<false, 0>(g, m, i, j, k)) return;
  cstar_newton_cell<float>(g, C, sel, U, V, W, Cn, fC, nitr, eps, itrmax, m, i, j, k);
}

// ============================================================= predictor ===
__device__ __forceinline__ float c0_atf(const Grid& g, const float* __restrict__ C, const float* __restrict__ fC,
                                        int q, int i, int j, int k) {
  const float E = (q == 0 || q == 3 || q == 5) ? 1.0f : 0.0f;
  if (g.walls && (j < 0 || j >= g.ny)) {
    float c0 = ldbw<true>(C, g, i, j, k, BC_C), c3 = ldbw<true>(C + 3 * g.n, g, i, j, k, BC_C), c5 = ldbw<true>(C + 5 * g.n, g, i, j, k, BC_C);
    float f = peterlin_f((float)(g.lp * g.lp), c0 + c3 + c5);
    return f * ldbw<true>(C + q * g.n, g, i, j, k, BC_C) - E;
  }
  long m = cidx(g, wrapi(i, g.nx), wrapi(j, g.ny), wrapi(k, g.nz));
  return fC[m] * C[q * g.n + m] - E;
}

// Stencil access of the predictor, neighbour (a, b, c) in {-1, 0, 1}^3 of the
// cell: u, v, w, p and c0(q) = f C_q - delta_q (the polymer stress deviation
// of component q).  PredFast reads through wrapped offsets (no wall row in
// the stencil), PredWalls through the boundary-rule accessor.
__device__ __forceinline__ float kron6(int q) { return (q == 0 || q == 3 || q == 5) ? 1.0f : 0.0f; }

struct PredFast {
  const Grid& g;
  long m;
  Nbr o;
  const float *U, *V, *W, *P, *C, *fC;
  __device__ __forceinline__ long at(int a, int b, int c) const { return m + o.x[a + 1] + o.y[b + 1] + o.z[c + 1]; }
  __device__ __forceinline__ float u(int a, int b, int c) const { return U[at(a, b, c)]; }
  __device__ __forceinline__ float v(int a, int b, int c) const { return V[at(a, b, c)]; }
  __device__ __forceinline__ float w(int a, int b, int c) const { return W[at(a, b, c)]; }
  __device__ __forceinline__ float p(int a, int b, int c) const { return P[at(a, b, c)]; }
  __device__ __forceinline__ float c0(int q, int a, int b, int c) const {
    const long mm = at(a, b, c);
    return fC[mm] * C[q * g.n + mm] - kron6(q);
  }
};

struct PredWalls {
  const Grid& g;
  int i, j, k;
  const float *U, *V, *W, *P, *C, *fC;
  __device__ __forceinline__ float u(int a, int b, int c) const { return ldbw<true>(U, g, i + a, j + b, k + c, BC_U); }
  __device__ __forceinline__ float v(int a, int b, int c) const { return ldbw<true>(V, g, i + a, j + b, k + c, BC_V); }
  __device__ __forceinline__ float w(int a, int b, int c) const { return ldbw<true>(W, g, i + a, j + b, k + c, BC_W); }
  __device__ __forceinline__ float p(int a, int b, int c) const { return ldbw<true>(P, g, i + a, j + b, k + c, BC_P); }
  __device__ __forceinline__ float c0(int q, int a, int b, int c) const { return c0_atf(g, C, fC, q, i + a, j + b, k + c); }
};

// the cell's own inputs besides the stencil: the AB2/CN histories and the
// forcing (body force added)
struct PredCell {
  float a0[3], t0[3], fv[3];
};
__device__ __forceinline__ void pred_cell_load(const Grid& g, long m, int first, const float* __restrict__ A0,
                                               const float* __restrict__ T0, const float* __restrict__ F, float bf0,
                                               float bf1, float bf2, PredCell& in) {
  const float bfs[3] = {bf0, bf1, bf2};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const long oo = c * g.n + m;
    in.a0[c] = first ? 0.0f : A0[oo];
    in.t0[c] = first ? 0.0f : T0[oo];
    in.fv[c] = (F ? F[oo] : 0.0f) + bfs[c];
  }
}

// predictor body over a stencil accessor S and the cell's own inputs.
// Increment form: Q receives d_q = u~ - u^n, the predictor's change of the
// velocity, summed in FP32 before it meets u^n.  The Helmholtz solve then
// works on the increment d = u* - u^n and the correction adds u^n + d -
// dt grad Phi in FP64, so the velocity is rounded to FP32 once per step.
// (Adding each term to u ~ 1 in FP32 rounds every near-uniform term, such as
// dt times a constant body force, the same way in every cell of a mean flow:
// a bias that grows linearly in time.)
//   fft_navier (fenep.f90:1466-1477): [1 - eps L] u* = u^n + r + dt B/2, with
//     u* = u^n + d:  [1 - eps L] d = r + dt B  (eps L u^n = dt B / 2)
//   ibm_preliminary + ibm_predict (fenep.f90:1994-2013, 2229):
//     u~ = u^n + r + dt B,  [1 - eps L] u* = u~ + dt f - dt B/2,
//     so [1 - eps L] d = r + dt B + dt f
// with r = -dt grad p + dt (3A - A0)/2 + dt (T + T0)/2 + dt F: the same
// d_q = r + dt B for both methods; the IBM spread adds dt f to it.
template <class S_t>
__device__ __forceinline__ void predictor_body(const Grid& g, const S_t& S, const PredCell& in, long m, int j,
                                               int method, int first, int polymer, float* __restrict__ A0,
                                               float* __restrict__ T0, float* __restrict__ Q) {
  const float dx = (float)g.dx, dy = (float)g.dy, dz = (float)g.dz, dt = (float)g.dt;
  const float idx_ = 1.0f / dx, idy = 1.0f / dy, idz = 1.0f / dz;
#define Ug(a, b, c) S.u(a, b, c)
#define Vg(a, b, c) S.v(a, b, c)
#define Wg(a, b, c) S.w(a, b, c)
#define Pg(a, b, c) S.p(a, b, c)
  float a[3], bv[3], t[3];
  const float uc = Ug(0, 0, 0), vc = Vg(0, 0, 0), wc = Wg(0, 0, 0);
  {
    // convection (fenep.f90:694-734): -div(u u) with face averages
    float h1 = 0.5f * (Ug(-1, 0, 0) + uc), h2 = 0.5f * (uc + Ug(1, 0, 0));
    float ax = (h2 * h2 - h1 * h1) * idx_
               + 0.25f * ((vc + Vg(1, 0, 0)) * (uc + Ug(0, 1, 0)) - (Vg(0, -1, 0) + Vg(1, -1, 0)) * (Ug(0, -1, 0) + uc)) * idy
               + 0.25f * ((wc + Wg(1, 0, 0)) * (uc + Ug(0, 0, 1)) - (Wg(0, 0, -1) + Wg(1, 0, -1)) * (Ug(0, 0, -1) + uc)) * idz;
    float g1 = 0.5f * (Vg(0, -1, 0) + vc), g2 = 0.5f * (vc + Vg(0, 1, 0));
    float ay = 0.25f * ((uc + Ug(0, 1, 0)) * (vc + Vg(1, 0, 0)) - (Ug(-1, 0, 0) + Ug(-1, 1, 0)) * (Vg(-1, 0, 0) + vc)) * idx_
               + (g2 * g2 - g1 * g1) * idy
               + 0.25f * ((wc + Wg(0, 1, 0)) * (vc + Vg(0, 0, 1)) - (Wg(0, 0, -1) + Wg(0, 1, -1)) * (Vg(0, 0, -1) + vc)) * idz;
    float e1 = 0.5f * (Wg(0, 0, -1) + wc), e2 = 0.5f * (wc + Wg(0, 0, 1));
    float az = 0.25f * ((uc + Ug(0, 0, 1)) * (wc + Wg(1, 0, 0)) - (Ug(-1, 0, 0) + Ug(-1, 0, 1)) * (Wg(-1, 0, 0) + wc)) * idx_
               + 0.25f * ((vc + Vg(0, 0, 1)) * (wc + Wg(0, 1, 0)) - (Vg(0, -1, 0) + Vg(0, -1, 1)) * (Wg(0, -1, 0) + wc)) * idy
               + (e2 * e2 - e1 * e1) * idz;
    a[0] = -ax; a[1] = -ay; a[2] = -az;
    // viscous (fenep.f90:736-770)
    const float cb = (float)(g.beta / g.re);
    const float ix2 = idx_ * idx_, iy2 = idy * idy, iz2 = idz * idz;
    bv[0] = cb * ((Ug(-1, 0, 0) - 2.0f * uc + Ug(1, 0, 0)) * ix2 + (Ug(0, -1, 0) - 2.0f * uc + Ug(0, 1, 0)) * iy2 +
                  (Ug(0, 0, -1) - 2.0f * uc + Ug(0, 0, 1)) * iz2);
    bv[1] = cb * ((Vg(-1, 0, 0) - 2.0f * vc + Vg(1, 0, 0)) * ix2 + (Vg(0, -1, 0) - 2.0f * vc + Vg(0, 1, 0)) * iy2 +
                  (Vg(0, 0, -1) - 2.0f * vc + Vg(0, 0, 1)) * iz2);
    bv[2] = cb * ((Wg(-1, 0, 0) - 2.0f * wc + Wg(1, 0, 0)) * ix2 + (Wg(0, -1, 0) - 2.0f * wc + Wg(0, 1, 0)) * iy2 +
                  (Wg(0, 0, -1) - 2.0f * wc + Wg(0, 0, 1)) * iz2);
  }
  if (polymer) {
#define C0g(q, aa, bb, cc) S.c0(q, aa, bb, cc)
    const float tx = (C0g(0, 1, 0, 0) - C0g(0, 0, 0, 0)) * idx_
                     + 0.25f * ((C0g(1, 0, 1, 0) + C0g(1, 1, 1, 0)) - (C0g(1, 0, -1, 0) + C0g(1, 1, -1, 0))) * idy
                     + 0.25f * ((C0g(2, 0, 0, 1) + C0g(2, 1, 0, 1)) - (C0g(2, 0, 0, -1) + C0g(2, 1, 0, -1))) * idz;
    const float ty = 0.25f * ((C0g(1, 1, 0, 0) - C0g(1, -1, 0, 0)) + (C0g(1, 1, 1, 0) - C0g(1, -1, 1, 0))) * idx_
                     + (C0g(3, 0, 1, 0) - C0g(3, 0, 0, 0)) * idy
                     + 0.25f * ((C0g(4, 0, 0, 1) + C0g(4, 0, 1, 1)) - (C0g(4, 0, 0, -1) + C0g(4, 0, 1, -1))) * idz;
    const float tz = 0.25f * ((C0g(2, 1, 0, 0) - C0g(2, -1, 0, 0)) + (C0g(2, 1, 0, 1) - C0g(2, -1, 0, 1))) * idx_
                     + 0.25f * ((C0g(4, 0, 1, 0) - C0g(4, 0, -1, 0)) + (C0g(4, 0, 1, 1) - C0g(4, 0, -1, 1))) * idy
                     + (C0g(5, 0, 0, 1) - C0g(5, 0, 0, 0)) * idz;
#undef C0g
    const float coef = (float)((1.0 - g.beta) / (g.re * g.wi));
    t[0] = coef * tx; t[1] = coef * ty; t[2] = coef * tz;
  } else {
    t[0] = t[1] = t[2] = 0.0f;
  }
  (void)method;
  const float p0 = Pg(0, 0, 0);
  const float pn[3] = {Pg(1, 0, 0), Pg(0, 1, 0), Pg(0, 0, 1)};
  const float ih[3] = {idx_, idy, idz};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    long oo = c * g.n + m;
    float a0 = first ? a[c] : in.a0[c];
    float t0 = first ? t[c] : in.t0[c];
    const float dq = -(dt * (pn[c] - p0) * ih[c]) + 0.5f * dt * (3.0f * a[c] - a0) + dt * bv[c] +
                     0.5f * dt * (t[c] + t0) + dt * in.fv[c];
    // the wall face of V stays 0 (V_kabe)
    Q[oo] = (c == 1 && g.walls && j == g.ny - 1) ? 0.0f : dq;
    A0[oo] = a[c];
    T0[oo] = t[c];
  }
#undef Ug
#undef Vg
#undef Wg
#undef Pg
}

// At most four waves per SIMD: with the registers that frees, the compiler
// issues the ~70 stencil loads of a cell in one batch instead of ~15 dependent
// groups of two to four (the kernel waits on L2 latency, not on bandwidth).
template <bool GEN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 4))) void m_predictor(Grid g, int method, int first, int polymer,
                                                   const float* __restrict__ U, const float* __restrict__ V,
                                                   const float* __restrict__ W, const float* __restrict__ P,
                                                   const float* __restrict__ C, const float* __restrict__ fC,
                                                   float* __restrict__ A0, float* __restrict__ T0,
                                                   const float* __restrict__ F,
                                                   float bf0, float bf1, float bf2, float* __restrict__ Q) {
  long m;
  int i, j, k;
  if (!row_cell<GEN, 1>(g, m, i, j, k)) return;
  PredCell in;
  pred_cell_load(g, m, first, A0, T0, F, bf0, bf1, bf2, in);
  if (GEN) {
    const PredWalls S{g, i, j, k, U, V, W, P, C, fC};
    predictor_body(g, S, in, m, j, method, first, polymer, A0, T0, Q);
  } else {
    const PredFast S{g, m, nbr(g, i, j, k), U, V, W, P, C, fC};
    predictor_body(g, S, in, m, j, method, first, polymer, A0, T0, Q);
  }
}

// ---- predictor streamed along z through LDS -----------------------------
// A block owns a PZ_TX x PZ_TY tile of cells in (x, y) and walks its column
// of planes k0 .. k0 + kc - 1.  The LDS holds three planes (k-1, k, k+1) of
// the tile with a one-cell halo for the ten stencil fields (u, v, w, p and
// c0 of the six components); every neighbour the predictor reads lies there.
// Plane k+2 is loaded into registers while plane k is computed, then takes
// the place of plane k-1.  Each value is fetched from L2 about 1.3 times at
// 64 x 16 tiles and 16 planes (halo rows and chunk ends), in whole rows,
// where the per-cell stencil of m_predictor fetches it about 6 times.  At
// 256x128x128 that is one 1024-thread block per CU (142 KB of LDS); 64 x 4
// tiles over 32 planes (1.6 fetches, two 256-thread blocks per CU) took
// 11 us more per step, 64 x 8 over 16 planes 4 us more.  The
// values in LDS are those of the boundary-rule accessor (wall ghosts and the
// wall face of V included), so one kernel covers the wall rows too.
#define PZ_TX 64
#define PZ_TY 16
#define PZ_HX (PZ_TX + 2)
#define PZ_PTS (PZ_HX * (PZ_TY + 2))
#define PZ_NF 10
#define PZ_THREADS (PZ_TX * PZ_TY)
#define PZ_PER ((PZ_PTS + PZ_THREADS - 1) / PZ_THREADS)
#define PZ_KC 16                  // planes per block

struct PredLds {
  const float *pm, *p0, *pp;   // planes k-1, k, k+1: PZ_NF fields of PZ_PTS
  int c;                       // the cell's halo-tile index
  __device__ __forceinline__ float at(int f, int a, int b, int cz) const {
    const float* pl = cz < 0 ? pm : (cz > 0 ? pp : p0);
    return pl[f * PZ_PTS + c + b * PZ_HX + a];
  }
  __device__ __forceinline__ float u(int a, int b, int c_) const { return at(0, a, b, c_); }
  __device__ __forceinline__ float v(int a, int b, int c_) const { return at(1, a, b, c_); }
  __device__ __forceinline__ float w(int a, int b, int c_) const { return at(2, a, b, c_); }
  __device__ __forceinline__ float p(int a, int b, int c_) const { return at(3, a, b, c_); }
  __device__ __forceinline__ float c0(int q, int a, int b, int c_) const { return at(4 + q, a, b, c_); }
};

// raw loads of the halo-tile points this thread fills (one plane)
struct PzRegs {
  float u[PZ_PER], v[PZ_PER], w[PZ_PER], p[PZ_PER], f[PZ_PER], ca[PZ_PER][6], cb[PZ_PER][6];
};

template <bool WL>
__device__ __forceinline__ void pz_fetch(const Grid& g, int i0, int j0, int kk, const float* __restrict__ U,
                                         const float* __restrict__ V, const float* __restrict__ W,
                                         const float* __restrict__ P, const float* __restrict__ C,
                                         const float* __restrict__ fC, PzRegs& r) {
  const long plane = (long)g.nx * g.ny;
  const int kz = wrapi(kk, g.nz);
#pragma unroll
  for (int e = 0; e < PZ_PER; e++) {
    const int pt = threadIdx.x + PZ_THREADS * e;
    if (pt < PZ_PTS) {
      const int hx = pt % PZ_HX, hy = pt / PZ_HX;
      const int ii = wrapi(i0 + hx - 1, g.nx);
      int jj = j0 + hy - 1;
      bool ghost = false;
      int j2 = jj;
      if (WL) {
        const bool lo = jj < 0, hi = jj >= g.ny;
        ghost = lo || hi;
        jj = lo ? 0 : (hi ? g.ny - 1 : jj);
        j2 = lo ? 1 : (hi ? g.ny - 2 : jj);
      } else {
        jj = wrapi(jj, g.ny);
      }
      const long b = ii + (long)g.nx * jj + plane * kz;
      r.u[e] = U[b];
      r.v[e] = V[b];
      r.w[e] = W[b];
      r.p[e] = P[b];
      r.f[e] = fC[b];
#pragma unroll
      for (int q = 0; q < 6; q++) r.ca[e][q] = C[q * g.n + b];
      if (WL && ghost) {
        const long b2 = ii + (long)g.nx * j2 + plane * kz;
#pragma unroll
        for (int q = 0; q < 6; q++) r.cb[e][q] = C[q * g.n + b2];
      }
    }
  }
}

// the fetched plane into an LDS slot, with the rules of the boundary-rule
// accessor (ldbw, c0_atf) applied
template <bool WL>
__device__ __forceinline__ void pz_put(const Grid& g, int j0, const PzRegs& r, float* __restrict__ slot) {
#pragma unroll
  for (int e = 0; e < PZ_PER; e++) {
    const int pt = threadIdx.x + PZ_THREADS * e;
    if (pt < PZ_PTS) {
      const int jj = j0 + pt / PZ_HX - 1;
      float u = r.u[e], v = r.v[e], w = r.w[e], pv = r.p[e];
      float c0[6];
      if (WL && (jj < 0 || jj >= g.ny)) {
        const bool lo = jj < 0;
        u = lo ? -u : 2.0f * 0.0f - u;
        w = lo ? -w : 2.0f * 0.0f - w;
        v = 0.0f;
        float cg[6];
#pragma unroll
        for (int q = 0; q < 6; q++) cg[q] = 2.0f * r.ca[e][q] - r.cb[e][q];
        const float f = peterlin_f((float)(g.lp * g.lp), cg[0] + cg[3] + cg[5]);
#pragma unroll
        for (int q = 0; q < 6; q++) c0[q] = f * cg[q] - kron6(q);
      } else {
        if (WL && jj == g.ny - 1) v = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++) c0[q] = r.f[e] * r.ca[e][q] - kron6(q);
      }
      slot[0 * PZ_PTS + pt] = u;
      slot[1 * PZ_PTS + pt] = v;
      slot[2 * PZ_PTS + pt] = w;
      slot[3 * PZ_PTS + pt] = pv;
#pragma unroll
      for (int q = 0; q < 6; q++) slot[(4 + q) * PZ_PTS + pt] = c0[q];
    }
  }
}

// grid: (nx / PZ_TX) (ny / PZ_TY) ceil(nz / kc) blocks of PZ_THREADS
template <bool WL>
__global__ __launch_bounds__(PZ_THREADS) void m_predictor_z(Grid g, int kc, int method, int first, int polymer,
                                                            const float* __restrict__ U, const float* __restrict__ V,
                                                            const float* __restrict__ W, const float* __restrict__ P,
                                                            const float* __restrict__ C, const float* __restrict__ fC,
                                                            float* __restrict__ A0, float* __restrict__ T0,
                                                            const float* __restrict__ F,
                                                            float bf0, float bf1, float bf2, float* __restrict__ Q) {
  __shared__ float lds[3][PZ_NF * PZ_PTS];
  const int ntx = g.nx / PZ_TX, nty = g.ny / PZ_TY;
  const long b = xcd_block();
  const int bx = (int)(b % ntx), by = (int)((b / ntx) % nty), bz = (int)(b / ((long)ntx * nty));
  const int i0 = bx * PZ_TX, j0 = by * PZ_TY, k0 = bz * kc;
  const int k1 = min(k0 + kc, g.nz);
  const int tx = threadIdx.x % PZ_TX, ty = threadIdx.x / PZ_TX;
  const int i = i0 + tx, j = j0 + ty;
  PzRegs r;
  // planes k0-1, k0, k0+1 into slots 0, 1, 2; plane kk lives in slot
  // (kk - k0 + 1) mod 3
#pragma unroll 1
  for (int s = 0; s < 3; s++) {
    pz_fetch<WL>(g, i0, j0, k0 - 1 + s, U, V, W, P, C, fC, r);
    pz_put<WL>(g, j0, r, lds[s]);
  }
  if (k0 + 2 <= k1) pz_fetch<WL>(g, i0, j0, k0 + 2, U, V, W, P, C, fC, r);
  int sm = 0, s0 = 1, sp = 2;
#pragma unroll 1
  for (int kk = k0; kk < k1; kk++) {
    __syncthreads();
    const long m = cidx(g, i, j, kk);
    PredCell in;
    pred_cell_load(g, m, first, A0, T0, F, bf0, bf1, bf2, in);
    const PredLds S{lds[sm], lds[s0], lds[sp], (ty + 1) * PZ_HX + tx + 1};
    predictor_body(g, S, in, m, j, method, first, polymer, A0, T0, Q);
    __syncthreads();
    // plane kk+2 replaces plane kk-1; fetch plane kk+3
    if (kk + 2 <= k1) pz_put<WL>(g, j0, r, lds[sm]);
    if (kk + 3 <= k1) pz_fetch<WL>(g, i0, j0, kk + 3, U, V, W, P, C, fC, r);
    const int t = sm; sm = s0; s0 = sp; sp = t;
  }
}

// ================================================================== FFT ===
// Stockham autosort (forward e^{-2 pi i jk/n}, unnormalised; the inverse
// uses the conjugate twiddles) over nl lines held in LDS with line pitch
// `pitch` (>= n, padded where lines are staged column-wise so that the
// staging is free of bank conflicts): radix-4 stages (decimation in
// frequency), after one radix-2 stage when log2 n is odd, so a 256-point line
// takes four barriers instead of eight.  tws: the n/2 twiddles (cos, -sin)(2
// pi j/n), already in LDS; W^(j + n/2) = -W^j gives the rest.  Returns the
// buffer holding the result.
__device__ __forceinline__ float2 tw_at(const float2* tws, int half, int j, bool inv) {
  const float2 w = j < half ? tws[j] : make_float2(-tws[j - half].x, -tws[j - half].y);
  return inv ? make_float2(w.x, -w.y) : w;
}
__device__ __forceinline__ float2 cmul_f(float2 a, float2 w) {
  return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

__device__ float2* stockham_f(float2* x, float2* y, int n, int logn, int nl, int pitch, const float2* tws, bool inv) {
  int len = n, s = 1, lgs = 0;
  const int half = n >> 1;
  if (logn & 1) {
    const int m = len >> 1;
    for (int b = threadIdx.x; b < nl * half; b += blockDim.x) {
      const int line = b >> (logn - 1);
      const int t = b & (half - 1);
      const int p = t >> lgs, q = t & (s - 1);
      const float2 w = tw_at(tws, half, p * s, inv);
      const float2* xl = x + line * pitch;
      float2* yl = y + line * pitch;
      const float2 a = xl[q + s * p], bb = xl[q + s * (p + m)];
      yl[q + s * (2 * p)] = make_float2(a.x + bb.x, a.y + bb.y);
      yl[q + s * (2 * p + 1)] = cmul_f(make_float2(a.x - bb.x, a.y - bb.y), w);
    }
    __syncthreads();
    float2* tt = x; x = y; y = tt;
    len = m; s = 2; lgs = 1;
  }
  const int quarter = n >> 2;
  while (len > 1) {
    const int m = len >> 2;
    for (int b = threadIdx.x; b < nl * quarter; b += blockDim.x) {
      const int line = b >> (logn - 2);
      const int t = b & (quarter - 1);
      const int p = t >> lgs, q = t & (s - 1);
      const float2* xl = x + line * pitch;
      float2* yl = y + line * pitch;
      const float2 a0 = xl[q + s * p], a1 = xl[q + s * (p + m)], a2 = xl[q + s * (p + 2 * m)],
                   a3 = xl[q + s * (p + 3 * m)];
      const float2 b0 = make_float2(a0.x + a2.x, a0.y + a2.y), b1 = make_float2(a0.x - a2.x, a0.y - a2.y);
      const float2 b2 = make_float2(a1.x + a3.x, a1.y + a3.y), d = make_float2(a1.x - a3.x, a1.y - a3.y);
      // (a1 - a3) times -i (forward) or +i (inverse)
      const float2 b3 = inv ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
      const int j1 = p * s;
      yl[q + s * (4 * p)] = make_float2(b0.x + b2.x, b0.y + b2.y);
      yl[q + s * (4 * p + 1)] = cmul_f(make_float2(b1.x + b3.x, b1.y + b3.y), tw_at(tws, half, j1, inv));
      yl[q + s * (4 * p + 2)] = cmul_f(make_float2(b0.x - b2.x, b0.y - b2.y), tw_at(tws, half, 2 * j1, inv));
      yl[q + s * (4 * p + 3)] = cmul_f(make_float2(b1.x - b3.x, b1.y - b3.y), tw_at(tws, half, 3 * j1, inv));
    }
    __syncthreads();
    float2* tt = x; x = y; y = tt;
    len = m; s <<= 2; lgs += 2;
  }
  return x;
}

__device__ __forceinline__ int ilog2f(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

__device__ __forceinline__ void load_twiddles(float2* tws, const float2* __restrict__ tw, int n) {
  for (int e = threadIdx.x; e < n / 2; e += blockDim.x) tws[e] = tw[e];
}

// Half-spectrum element (kx, row, kz) in the x-major layout (kx fastest) or
// the transposed layout of the slab solve (row fastest, then kz, then kx: one
// kx slab is contiguous).
// x-line L = j + ny k -> its row j and plane k (ny is a power of two)
__device__ __forceinline__ long line_row(const Grid& g, long L) { return L & (long)(g.ny - 1); }
__device__ __forceinline__ long line_plane(const Grid& g, long L) { return L >> g.lny; }
__device__ __forceinline__ long hs_index(const Grid& g, int tr, long k, long row, long kk) {
  return tr ? row + (long)g.sy * (kk + (long)g.nz * k) : k + (long)g.nxh * (row + (long)g.sy * kk);
}

// Store the half spectra of nl real x-lines (two per complex transform r:
// A[k] = (Z[k] + conj Z[n-k]) / 2, B[k] = (Z[k] - conj Z[n-k]) / 2i).  In the
// transposed layout consecutive threads take consecutive lines of one kx, so
// each kx receives a contiguous run of rows.
__device__ __forceinline__ void store_half_spectra(const Grid& g, const float2* r, float2* __restrict__ spec, int nl,
                                                   long l0, int tr) {
  const int n = g.nx, nc = nl / 2, nxh = g.nxh;
  const long nlines = (long)g.ny * g.nz;
  for (int e = threadIdx.x; e < nc * nxh; e += blockDim.x) {
    int c, k;
    if (tr) { k = e >> __builtin_ctz(nc); c = e & (nc - 1); }   // nl: 4, 8 or 16 (lines_for32)
    else { c = e / nxh; k = e - c * nxh; }
    const float2 Z = r[c * n + k], W = r[c * n + ((n - k) & (n - 1))];
    const long La = l0 + 2 * c, Lb = La + 1;
    if (La < nlines)
      spec[hs_index(g, tr, k, line_row(g, La) + g.jo, line_plane(g, La))] = make_float2(0.5f * (Z.x + W.x), 0.5f * (Z.y - W.y));
    if (Lb < nlines)
      spec[hs_index(g, tr, k, line_row(g, Lb) + g.jo, line_plane(g, Lb))] = make_float2(0.5f * (Z.y + W.y), 0.5f * (W.x - Z.x));
  }
}

// Real x-lines -> half-spectrum rows, two real lines per complex transform:
// z = a + i b, A[k] = (Z[k] + conj Z[n-k]) / 2, B[k] = (Z[k] - conj Z[n-k]) / 2i.
// nl (even) real lines per block.
__global__ __launch_bounds__(256) void m_fft_x_fwd(Grid g, const float* __restrict__ q, float2* __restrict__ spec,
                                                   const float2* __restrict__ tw, int nl, const int* __restrict__ skip,
                                                   int tr) {
  if (skip && *skip) return;
  extern __shared__ float2 ldsf[];
  const int n = g.nx, nc = nl / 2;
  float2* x = ldsf;
  float2* y = ldsf + nc * n;
  float2* tws = ldsf + 2 * nc * n;
  load_twiddles(tws, tw, n);
  const long nlines = (long)g.ny * g.nz;
  const long l0 = (long)blockIdx.x * nl;
  for (int e = threadIdx.x; e < nc * n; e += blockDim.x) {
    const int c = e >> g.lnx, i = e & (n - 1);   // n = nx, a power of two
    const long La = l0 + 2 * c, Lb = La + 1;
    x[e] = make_float2(La < nlines ? q[La * n + i] : 0.0f, Lb < nlines ? q[Lb * n + i] : 0.0f);
  }
  __syncthreads();
  const float2* r = stockham_f(x, y, n, ilog2f(n), nc, n, tws, false);
  store_half_spectra(g, r, spec, nl, l0, tr);
}

// Half-spectrum rows -> real x-lines (unnormalised), two lines per complex
// transform: Z = A + i B over the full Hermitian spectra; a = Re z, b = Im z.
__global__ __launch_bounds__(256) void m_fft_x_inv(Grid g, const float2* __restrict__ spec, float* __restrict__ out,
                                                   const float2* __restrict__ tw, int nl, const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ float2 ldsf[];
  const int n = g.nx, h = n / 2, nxh = g.nxh, nc = nl / 2;
  float2* x = ldsf;
  float2* y = ldsf + nc * n;
  float2* tws = ldsf + 2 * nc * n;
  load_twiddles(tws, tw, n);
  const long nlines = (long)g.ny * g.nz;
  const long l0 = (long)blockIdx.x * nl;
  for (int e = threadIdx.x; e < nc * n; e += blockDim.x) {
    const int c = e >> g.lnx, i = e & (n - 1);   // n = nx, a power of two
    const long La = l0 + 2 * c, Lb = La + 1;
    const int src = i <= h ? i : n - i;
    float2 A = make_float2(0.0f, 0.0f), B = make_float2(0.0f, 0.0f);
    if (La < nlines) A = spec[src + (long)nxh * ((line_row(g, La) + g.jo) + g.sy * (line_plane(g, La)))];
    if (Lb < nlines) B = spec[src + (long)nxh * ((line_row(g, Lb) + g.jo) + g.sy * (line_plane(g, Lb)))];
    if (src == 0 || src == h) { A.y = 0.0f; B.y = 0.0f; }
    if (i > h) { A.y = -A.y; B.y = -B.y; }
    x[e] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  const float2* r = stockham_f(x, y, n, ilog2f(n), nc, n, tws, true);
  for (int e = threadIdx.x; e < nc * n; e += blockDim.x) {
    const int c = e >> g.lnx, i = e & (n - 1);   // n = nx, a power of two
    const long La = l0 + 2 * c, Lb = La + 1;
    const float2 z = r[e];
    if (La < nlines) out[La * n + i] = z.x;
    if (Lb < nlines) out[Lb * n + i] = z.y;
  }
}

// u* = u^n + d at cell (i, j, k), exact in FP64, each part through the
// boundary rules (linear and homogeneous, so that the sum takes them too)
template <bool WL>
__device__ __forceinline__ double ustar(const Grid& g, const float* un, const float* d, int i, int j, int k, int kind) {
  return (double)ldbw<WL>(un, g, i, j, k, kind) + (double)ldbw<WL>(d, g, i, j, k, kind);
}

// b = div_h(u*)/dt at cell (i, j, k) in FP64 from u* = u^n + d (fft_poisson,
// fenep.f90:1550-1552; poisson, :869-877): U, V, W the state, D the three
// planes of the increment.  idx.. are the FP64 reciprocals of the spacings
// and of dt (an FP64 division is some twenty instructions; the products
// differ from the quotients by an FP64 ulp).
template <bool WL>
__device__ __forceinline__ double div_rhs(const Grid& g, const float* U, const float* V, const float* W, const float* D,
                                          int i, int j, int k, double idx, double idy, double idz, double idt) {
  const float *Du = D, *Dv = D + g.n, *Dw = D + 2 * g.n;
  return ((ustar<WL>(g, U, Du, i, j, k, BC_U) - ustar<WL>(g, U, Du, i - 1, j, k, BC_U)) * idx +
          (ustar<WL>(g, V, Dv, i, j, k, BC_V) - ustar<WL>(g, V, Dv, i, j - 1, k, BC_V)) * idy +
          (ustar<WL>(g, W, Dw, i, j, k, BC_W) - ustar<WL>(g, W, Dw, i, j, k - 1, BC_W)) * idz) * idt;
}

// The Poisson right-hand side b = div_h(u*)/dt (div_rhs) computed in the load stage of
// the forward x transform, two real lines per complex transform as in
// m_fft_x_fwd.  want_b: also store b (FP64) and the per-block |b|^2 for the
// defect check.
template <bool WL>
__global__ __launch_bounds__(256) void m_div_fft_x_fwd(Grid g, const float* __restrict__ U, const float* __restrict__ V,
                                                       const float* __restrict__ W, const float* __restrict__ up,
                                                       float2* __restrict__ spec,
                                                       const float2* __restrict__ tw, int nl, double* __restrict__ b,
                                                       double* __restrict__ nrm, int want_b, int tr) {
  extern __shared__ float2 ldsf[];
  const int n = g.nx, nc = nl / 2;
  float2* x = ldsf;
  float2* y = ldsf + nc * n;
  float2* tws = ldsf + 2 * nc * n;
  load_twiddles(tws, tw, n);
  const long nlines = (long)g.ny * g.nz;
  const long l0 = (long)blockIdx.x * nl;
  double ss = 0.0;
  const double idx = 1.0 / g.dx, idy = 1.0 / g.dy, idz = 1.0 / g.dz, idt = 1.0 / g.dt;
#pragma unroll 2
  for (int e = threadIdx.x; e < nc * n; e += blockDim.x) {
    const int c = e >> g.lnx, i = e & (n - 1);   // n = nx, a power of two
    float2 v2 = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const long L = l0 + 2 * c + h;
      if (L >= nlines) continue;
      const int j = (int)(line_row(g, L)), k = (int)(line_plane(g, L));
      const double v = div_rhs<WL>(g, U, V, W, up, i, j, k, idx, idy, idz, idt);
      if (want_b) {
        b[L * n + i] = v;
        ss += v * v;
      }
      if (h == 0) v2.x = (float)v;
      else v2.y = (float)v;
    }
    x[e] = v2;
  }
  if (want_b) {
    __shared__ double sh[256];
    const double sum = block_sum256(ss, sh);
    if (threadIdx.x == 0) nrm[blockIdx.x] = sum;
  }
  __syncthreads();
  const float2* r = stockham_f(x, y, n, ilog2f(n), nc, n, tws, false);
  store_half_spectra(g, r, spec, nl, l0, tr);
}

// Inverse x transform of the Poisson correction d with the FP64
// accumulation of the potential in the store stage: phi = d (first sweep) or
// phi + d.
__global__ __launch_bounds__(256) void m_fft_x_inv_accum(Grid g, const float2* __restrict__ spec, double* __restrict__ phi,
                                                         const float2* __restrict__ tw, int nl, int first,
                                                         const int* __restrict__ skip, int tr) {
  if (skip && *skip) return;
  extern __shared__ float2 ldsf[];
  const int n = g.nx, h = n / 2, nxh = g.nxh, nc = nl / 2;
  float2* x = ldsf;
  float2* y = ldsf + nc * n;
  float2* tws = ldsf + 2 * nc * n;
  float2* hs = tws + n / 2;        // transposed layout: [line][kx] staging
  load_twiddles(tws, tw, n);
  const long nlines = (long)g.ny * g.nz;
  const long l0 = (long)blockIdx.x * nl;
  if (tr) {
    // consecutive threads read consecutive rows of one kx (coalesced); the
    // unrolled loop keeps several loads of a thread in flight
#pragma unroll 4
    for (int e = threadIdx.x; e < nl * nxh; e += blockDim.x) {
      const int k = e >> __builtin_ctz(nl), jl = e & (nl - 1);   // nl: 4, 8 or 16 (lines_for32)
      const long L = l0 + jl;
      hs[jl * nxh + k] = L < nlines ? spec[hs_index(g, 1, k, line_row(g, L) + g.jo, line_plane(g, L))] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < nc * n; e += blockDim.x) {
    const int c = e >> g.lnx, i = e & (n - 1);   // n = nx, a power of two
    const long La = l0 + 2 * c, Lb = La + 1;
    const int src = i <= h ? i : n - i;
    float2 A = make_float2(0.0f, 0.0f), B = make_float2(0.0f, 0.0f);
    if (tr) {
      A = hs[(2 * c) * nxh + src];
      B = hs[(2 * c + 1) * nxh + src];
    } else {
      if (La < nlines) A = spec[src + (long)nxh * ((line_row(g, La) + g.jo) + g.sy * (line_plane(g, La)))];
      if (Lb < nlines) B = spec[src + (long)nxh * ((line_row(g, Lb) + g.jo) + g.sy * (line_plane(g, Lb)))];
    }
    if (src == 0 || src == h) { A.y = 0.0f; B.y = 0.0f; }
    if (i > h) { A.y = -A.y; B.y = -B.y; }
    x[e] = make_float2(A.x - B.y, A.y + B.x);
  }
  __syncthreads();
  const float2* r = stockham_f(x, y, n, ilog2f(n), nc, n, tws, true);
  for (int e = threadIdx.x; e < nc * n; e += blockDim.x) {
    const int c = e >> g.lnx, i = e & (n - 1);   // n = nx, a power of two
    const long La = l0 + 2 * c, Lb = La + 1;
    const float2 z = r[e];
    if (La < nlines) phi[La * n + i] = first ? (double)z.x : phi[La * n + i] + (double)z.x;
    if (Lb < nlines) phi[Lb * n + i] = first ? (double)z.y : phi[Lb * n + i] + (double)z.y;
  }
}

// y or z transforms of ti consecutive kx columns; lines are staged
// column-wise into LDS with pitch n + 1.  mode 0 forward, 1 inverse, 2
// forward + spectral divide (Helmholtz or Poisson symbol) + inverse.
__global__ __launch_bounds__(256) void m_fft_yz(Grid g, float2* __restrict__ spec, const float2* __restrict__ tw, int axis,
                                                int mode, int kind, int row0, int ti, const float* __restrict__ kx2,
                                                const float* __restrict__ ky2, const float* __restrict__ kz2,
                                                const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ float2 ldsf[];
  const int n = axis == 1 ? g.ny : g.nz;
  const int pitch = n + 1;
  const int nxh = g.nxh;
  const long stride = axis == 1 ? (long)nxh : (long)nxh * g.sy;
  const int ntile = (nxh + ti - 1) / ti;
  const int tile = blockIdx.x % ntile;
  const long other = blockIdx.x / ntile;
  const int i0 = tile * ti;
  float2* base = axis == 1 ? spec + (long)nxh * (g.jo + g.sy * other) : spec + (long)nxh * (row0 + other);
  float2* x = ldsf;
  float2* y = ldsf + ti * pitch;
  float2* tws = ldsf + 2 * ti * pitch;
  load_twiddles(tws, tw, n);
  for (int e = threadIdx.x; e < ti * n; e += blockDim.x) {
    const int l = e % ti, r = e / ti;
    const int i = i0 + l;
    x[l * pitch + r] = i < nxh ? base[i + stride * r] : make_float2(0.0f, 0.0f);
  }
  __syncthreads();
  const int logn = ilog2f(n);
  float2* res = stockham_f(x, y, n, logn, ti, pitch, tws, mode == 1);
  if (mode == 2) {
    float2* other_buf = res == x ? y : x;
    const float intot = (float)(1.0 / ((double)g.nx * g.ny * g.nz));
    const float coef = (float)(((g.dt * g.beta) / g.re) / 2.0);
    for (int e = threadIdx.x; e < ti * n; e += blockDim.x) {
      const int l = e / n, kk = e - l * n;
      const int i = i0 + l;
      if (i >= nxh) continue;
      const long jj = other;
      const float lhs = kind == 3 ? -((kx2[i] + ky2[jj]) + kz2[kk]) : 1.0f + coef * ((kx2[i] + ky2[jj]) + kz2[kk]);
      const float2 sv = res[l * pitch + kk];
      const float sc = fabsf(lhs) < 1.0e-16f ? 0.0f : intot / lhs;
      res[l * pitch + kk] = make_float2(sv.x * sc, sv.y * sc);
    }
    __syncthreads();
    res = stockham_f(res, other_buf, n, logn, ti, pitch, tws, true);
  }
  for (int e = threadIdx.x; e < ti * n; e += blockDim.x) {
    const int l = e % ti, r = e / ti;
    const int i = i0 + l;
    if (i < nxh) base[i + stride * r] = res[l * pitch + r];
  }
}

// In-place radix-2 transforms along the slow index of an LDS array held
// [point][row] (rows contiguous, point pitch P), for nr rows (a power of two,
// lnr = log2 nr) of n points.  Consecutive threads take consecutive rows of
// one butterfly: each LDS access of a wavefront is one contiguous run and its
// twiddles are broadcasts.  Two stages per pass: the four elements of two
// consecutive stages stay in registers.  dif = false: decimation in time
// (bit-reversed input, natural-order output); dif = true: decimation in
// frequency (natural input, bit-reversed output).  tws[t] = exp(-2 pi i t /
// n), t < n/2; inv: conjugate twiddles (unnormalised).  Called by every
// thread of the block; ends with a barrier.
__device__ __forceinline__ void bfly2(float2& a, float2& b, float2 w, bool dif) {
  if (dif) {
    const float2 d = make_float2(a.x - b.x, a.y - b.y);
    a = make_float2(a.x + b.x, a.y + b.y);
    b = cmul_f(d, w);
  } else {
    const float2 t = cmul_f(b, w);
    b = make_float2(a.x - t.x, a.y - t.y);
    a = make_float2(a.x + t.x, a.y + t.y);
  }
}
__device__ void fft_cols_inplace(float2* S, int P, int nr, int lnr, int n, int logn, const float2* tws, bool inv,
                                 bool dif) {
  const int half = n >> 1;
  // stage spans in order: DIT 1, 2, ..., n/2; DIF n/2, ..., 2, 1.  A lone
  // radix-2 stage (log2 n odd) is the last DIT stage or the first DIF stage.
  int done = 0;
  if (dif && (logn & 1)) {
    for (int b = threadIdx.x; b < nr * half; b += blockDim.x) {
      const int r = b & (nr - 1), i = b >> lnr;
      float2 x0 = S[(size_t)i * P + r], x1 = S[(size_t)(i + half) * P + r];
      bfly2(x0, x1, tw_at(tws, half, i, inv), true);
      S[(size_t)i * P + r] = x0;
      S[(size_t)(i + half) * P + r] = x1;
    }
    __syncthreads();
    done = 1;
  }
  for (; done + 2 <= logn; done += 2) {
    // spans h (first) and 2h (DIT) or 2h (first) and h (DIF)
    const int lh = dif ? logn - done - 2 : done, h = 1 << lh;
    for (int b = threadIdx.x; b < nr * (n >> 2); b += blockDim.x) {
      const int r = b & (nr - 1), t = b >> lnr;
      const int k = t & (h - 1);
      const int i0 = ((t >> lh) << (lh + 2)) + k;
      float2* x = S + (size_t)i0 * P + r;
      float2 a0 = x[0], a1 = x[(size_t)h * P], a2 = x[(size_t)2 * h * P], a3 = x[(size_t)3 * h * P];
      // twiddles of span h (index k n/(2h)) and span 2h (k n/(4h), (k+h) n/(4h))
      const float2 w1 = tw_at(tws, half, k << (logn - lh - 1), inv);
      const float2 wa = tw_at(tws, half, k << (logn - lh - 2), inv), wb = tw_at(tws, half, (k + h) << (logn - lh - 2), inv);
      if (dif) {
        bfly2(a0, a2, wa, true);
        bfly2(a1, a3, wb, true);
        bfly2(a0, a1, w1, true);
        bfly2(a2, a3, w1, true);
      } else {
        bfly2(a0, a1, w1, false);
        bfly2(a2, a3, w1, false);
        bfly2(a0, a2, wa, false);
        bfly2(a1, a3, wb, false);
      }
      x[0] = a0;
      x[(size_t)h * P] = a1;
      x[(size_t)2 * h * P] = a2;
      x[(size_t)3 * h * P] = a3;
    }
    __syncthreads();
  }
  if (done < logn) {
    // the last DIT stage, span n/2
    for (int b = threadIdx.x; b < nr * half; b += blockDim.x) {
      const int r = b & (nr - 1), i = b >> lnr;
      float2 x0 = S[(size_t)i * P + r], x1 = S[(size_t)(i + half) * P + r];
      bfly2(x0, x1, tw_at(tws, half, i, inv), false);
      S[(size_t)i * P + r] = x0;
      S[(size_t)(i + half) * P + r] = x1;
    }
    __syncthreads();
  }
}

// Pressure Poisson solve of one kx slab with y-walls (fft_poisson_kabe and
// tdma, fenep_2d.f90:1145-1177, 1179-1342): forward z transforms of the
// interior rows, the line solve along y of every kz (boundary rows x_0 = x_1,
// x_{NY+1} = x_NY), inverse z transforms.  The slab is held in LDS as it lies
// in the transposed half spectrum, [kz][row] (rows contiguous, kz pitch the
// odd number slab_pitch >= ny + 2), so its loads and stores are straight
// copies.  The z transforms run across rows (fft_cols_inplace); their bit
// reversals are folded into the copies: the load puts kz at its bit-reversed
// place for the decimation-in-time forward transform, and the store reads
// the decimation-in-frequency inverse's bit-reversed output.  The line solve
// keeps its P coefficients in pbuf (L2).
static int slab_pitch(const Grid& g) { return g.sy | 1; }
static size_t lds_slab(const Grid& g) { return sizeof(float2) * ((size_t)g.nz * slab_pitch(g) + g.nz / 2); }
__global__ __launch_bounds__(1024) void m_poisson_slab(Grid g, float2* __restrict__ spec, const float2* __restrict__ tw,
                                                      const float* __restrict__ kx2, const float* __restrict__ kz2,
                                                      float* __restrict__ pbuf, const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ float2 lds[];
  const int nz = g.nz, sy = g.sy, ny = g.ny, pitch = g.sy | 1;
  const long kx = blockIdx.x;
  const int logn = ilog2f(nz);
  float2* S = lds;                              // [kz][pitch]
  float2* tws = lds + (size_t)nz * pitch;
  load_twiddles(tws, tw, nz);
  float2* slab = spec + (size_t)sy * nz * kx;   // [kz][row]
  // one workgroup per CU: issue a thread's loads together, then store them
  constexpr int LU = 16;
  for (int e0 = 0; e0 < ny * nz; e0 += 1024 * LU) {
    float2 v[LU];
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int e = e0 + 1024 * u + (int)threadIdx.x;
      if (e < ny * nz) {
        const int kk = e >> g.lny, j = e & (ny - 1);
        v[u] = slab[(size_t)sy * kk + j + 1];
      }
    }
#pragma unroll
    for (int u = 0; u < LU; u++) {
      const int e = e0 + 1024 * u + (int)threadIdx.x;
      if (e < ny * nz) {
        const int kk = e >> g.lny, j = e & (ny - 1);
        const int rk = (int)(__brev((unsigned)kk) >> (32 - logn));
        S[(size_t)rk * pitch + j + 1] = v[u];
      }
    }
  }
  __syncthreads();
  fft_cols_inplace(S + 1, pitch, ny, g.lny, nz, logn, tws, false, false);
  // line solve along y, one thread per kz; rows go in batches of TB with the
  // next batch's operands loaded while the current one is computed (the back
  // substitution reads P from L2)
  if ((int)threadIdx.x < nz) {
    constexpr int TB = 8;
    const int kk = threadIdx.x;
    float2* L = S + (size_t)kk * pitch;
    float* P = pbuf + (size_t)sy * nz * kx;   // [row][kz]
    const float inxz = (float)(1.0 / ((double)g.nx * g.nz));
    const float idy2 = (float)(1.0 / (g.dy * g.dy));
    const float kq = kx2[kx] + kz2[kk];
    const float ai = -(kq + 2.0f * idy2);
    float Pp = 1.0f, Qr = 0.0f, Qi = 0.0f;   // row 0: a = 1, b = -1
    P[kk] = Pp;
    L[0] = make_float2(0.0f, 0.0f);
    float2 cur[TB];
#pragma unroll
    for (int t = 0; t < TB; t++) cur[t] = 1 + t < sy ? L[1 + t] : make_float2(0.0f, 0.0f);
    for (int r0 = 1; r0 < sy; r0 += TB) {
      float2 nxt[TB];
#pragma unroll
      for (int t = 0; t < TB; t++) nxt[t] = r0 + TB + t < sy ? L[r0 + TB + t] : make_float2(0.0f, 0.0f);
#pragma unroll
      for (int t = 0; t < TB; t++) {
        const int r = r0 + t;
        if (r < sy) {
          const bool last = r == sy - 1;
          const float a = last ? 1.0f : ai;
          const float bb = last ? 0.0f : idy2;
          const float cc = last ? -1.0f : idy2;
          const float dr = last ? 0.0f : cur[t].x * inxz, di = last ? 0.0f : cur[t].y * inxz;
          const float den = a + cc * Pp;
          if (fabsf(den) < 1.0e-16f * fabsf(a) + 1.0e-30f) { Pp = 0.0f; Qr = 0.0f; Qi = 0.0f; }
          else { const float id = __builtin_amdgcn_rcpf(den); Pp = -bb * id; Qr = (dr - cc * Qr) * id; Qi = (di - cc * Qi) * id; }
          P[r * nz + kk] = Pp;
          L[r] = make_float2(Qr, Qi);
        }
      }
#pragma unroll
      for (int t = 0; t < TB; t++) cur[t] = nxt[t];
    }
    // back substitution x_r = P_r x_{r+1} + Q_r, r = sy-2 .. 0
    float xr = Qr, xi = Qi;
    float pc[TB];
    float2 qc[TB];
#pragma unroll
    for (int t = 0; t < TB; t++) {
      const int r = sy - 2 - t;
      pc[t] = r >= 0 ? P[r * nz + kk] : 0.0f;
      qc[t] = r >= 0 ? L[r] : make_float2(0.0f, 0.0f);
    }
    for (int r0 = sy - 2; r0 >= 0; r0 -= TB) {
      float pn[TB];
      float2 qn[TB];
#pragma unroll
      for (int t = 0; t < TB; t++) {
        const int r = r0 - TB - t;
        pn[t] = r >= 0 ? P[r * nz + kk] : 0.0f;
        qn[t] = r >= 0 ? L[r] : make_float2(0.0f, 0.0f);
      }
#pragma unroll
      for (int t = 0; t < TB; t++) {
        const int r = r0 - t;
        if (r >= 0) {
          xr = pc[t] * xr + qc[t].x;
          xi = pc[t] * xi + qc[t].y;
          L[r] = make_float2(xr, xi);
        }
      }
#pragma unroll
      for (int t = 0; t < TB; t++) { pc[t] = pn[t]; qc[t] = qn[t]; }
    }
  }
  __syncthreads();
  fft_cols_inplace(S + 1, pitch, ny, g.lny, nz, logn, tws, true, true);
  for (int e = threadIdx.x; e < ny * nz; e += blockDim.x) {
    const int kk = e >> g.lny, j = e & (ny - 1);
    const int rk = (int)(__brev((unsigned)kk) >> (32 - logn));
    slab[(size_t)sy * kk + j + 1] = S[(size_t)rk * pitch + j + 1];
  }
}

// Line solve along y of each (kx, kz) mode (FP32), staged through LDS like
// the FP64 k_thomas: coalesced row loads, one thread per line for the
// recurrence, coalesced stores.
__global__ __launch_bounds__(256) void m_thomas(Grid g, float2* __restrict__ spec, int kind, int tl,
                                                const float* __restrict__ kx2, const float* __restrict__ kz2,
                                                const int* __restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ float2 sq[];                 // [size][tl]
  const int nxh = g.nxh, ny = g.ny, size = ny + 2;
  float* sp = (float*)(sq + (long)size * tl);    // [size][tl]
  const int ngrp = (nxh + tl - 1) / tl;
  const int i0 = (int)(blockIdx.x % ngrp) * tl;
  const long k = blockIdx.x / ngrp;
  const int nl = min(tl, nxh - i0);
  float2* base = spec + (long)nxh * g.sy * k + i0;
  for (int e = threadIdx.x; e < size * tl; e += blockDim.x) {
    const int r = e / tl, l = e - r * tl;
    if (l < nl) sq[e] = base[(long)nxh * r + l];
  }
  __syncthreads();
  if (threadIdx.x < nl) {
    const int l = threadIdx.x, i = i0 + l;
    const float inxz = (float)(1.0 / ((double)g.nx * g.nz));
    const float coef = (float)(((g.dt * g.beta) / g.re) / 2.0);
    const float idy2 = (float)(1.0 / (g.dy * g.dy));
    const float kk = kx2[i] + kz2[k];
    auto coeffs = [&](int r, float& a, float& b, float& c, bool& rhs) {
      rhs = false;
      if (kind == 3) {
        if (r == 0) { a = 1.0f; b = -1.0f; c = 0.0f; }
        else if (r == size - 1) { a = 1.0f; b = 0.0f; c = -1.0f; }
        else { a = -(kk + 2.0f * idy2); b = idy2; c = idy2; rhs = true; }
      } else if (kind == 1) {
        if (r == 0 || r >= ny) { a = 1.0f; b = 0.0f; c = 0.0f; }
        else { a = 1.0f + coef * (kk + 2.0f * idy2); b = -coef * idy2; c = b; rhs = true; }
      } else {
        if (r == 0) { a = 1.0f; b = 1.0f; c = 0.0f; }
        else if (r == size - 1) { a = 1.0f; b = 0.0f; c = 1.0f; }
        else { a = 1.0f + coef * (kk + 2.0f * idy2); b = -coef * idy2; c = b; rhs = true; }
      }
    };
    float a, b, c; bool rhs;
    coeffs(0, a, b, c, rhs);
    float Pp = -b / a, Qr = 0.0f, Qi = 0.0f;
    sp[l] = Pp;
    sq[l] = make_float2(0.0f, 0.0f);
    // prefetch the next row before storing the current one: the recurrence
    // then waits on arithmetic only
    float2 nxt = sq[tl + l];
    for (int r = 1; r < size; r++) {
      coeffs(r, a, b, c, rhs);
      const float2 sv = nxt;
      if (r + 1 < size) nxt = sq[(r + 1) * tl + l];
      float dr = 0.0f, di = 0.0f;
      if (rhs) { dr = sv.x * inxz; di = sv.y * inxz; }
      const float den = a + c * Pp;
      if (fabsf(den) < 1.0e-16f * fabsf(a) + 1.0e-30f) { Pp = 0.0f; Qr = 0.0f; Qi = 0.0f; }
      else { const float id = __builtin_amdgcn_rcpf(den); Pp = -b * id; Qr = (dr - c * Qr) * id; Qi = (di - c * Qi) * id; }
      sp[r * tl + l] = Pp;
      sq[r * tl + l] = make_float2(Qr, Qi);
    }
    float xr = Qr, xi = Qi;
    float Pn = sp[(size - 2) * tl + l];
    float2 qn = sq[(size - 2) * tl + l];
    for (int r = size - 2; r >= 0; r--) {
      const float P = Pn;
      const float2 q = qn;
      if (r > 0) { Pn = sp[(r - 1) * tl + l]; qn = sq[(r - 1) * tl + l]; }
      xr = P * xr + q.x;
      xi = P * xi + q.y;
      sq[r * tl + l] = make_float2(xr, xi);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < size * tl; e += blockDim.x) {
    const int r = e / tl, l = e - r * tl;
    if (l < nl) base[(long)nxh * r + l] = sq[e];
  }
}

// real x-lines per block (even: the x transforms pack two real lines per
// complex line); LDS 8 B * n * lines + n/2 twiddles.  A power of two for the
// power-of-two n (4, 8 or 16): the spectrum stages index lines by shifts.
static int lines_for32(int n) { return std::max(2, std::min(16, 4096 / n)) & ~1; }
// kx columns per block of the y/z transforms (pitch n + 1)
static int cols_for32(int n) { return std::max(1, std::min(16, 3968 / (n + 1))); }

// y and z transforms (periodic y) or z transforms and the y line solve
// (walls) of the half spectrum, with the spectral divide of the solve kind
static void solve32_yz(Ctx& c, int kind, const int* skip, int tr = 0) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  if (tr) {
    // transposed half spectrum, walls, pressure Poisson
    hipLaunchKernelGGL(m_poisson_slab, dim3(g.nxh), dim3(1024), lds_slab(g), s, g, X.spec, X.twz, X.kx2, X.kz2, X.pbuf,
                       skip);
    return;
  }
  int tiy = cols_for32(g.ny), tiz = cols_for32(g.nz);
  int nty = (g.nxh + tiy - 1) / tiy, ntz = (g.nxh + tiz - 1) / tiz;
  size_t shy = sizeof(float2) * (2 * (size_t)tiy * (g.ny + 1) + g.ny / 2);
  size_t shz = sizeof(float2) * (2 * (size_t)tiz * (g.nz + 1) + g.nz / 2);
  if (!g.walls) {
    hipLaunchKernelGGL(m_fft_yz, dim3(nty * g.nz), dim3(256), shy, s, g, X.spec, X.twy, 1, 0, kind, 0, tiy, X.kx2, X.ky2, X.kz2, skip);
    hipLaunchKernelGGL(m_fft_yz, dim3(ntz * g.ny), dim3(256), shz, s, g, X.spec, X.twz, 2, 2, kind, 0, tiz, X.kx2, X.ky2, X.kz2, skip);
    hipLaunchKernelGGL(m_fft_yz, dim3(nty * g.nz), dim3(256), shy, s, g, X.spec, X.twy, 1, 1, kind, 0, tiy, X.kx2, X.ky2, X.kz2, skip);
  } else {
    hipLaunchKernelGGL(m_fft_yz, dim3(ntz * g.ny), dim3(256), shz, s, g, X.spec, X.twz, 2, 0, kind, 1, tiz, X.kx2, X.ky2, X.kz2, skip);
    const int tl = thomas_lines(g.ny, (int)(sizeof(float2) + sizeof(float)));
    const int ngrp = (g.nxh + tl - 1) / tl;
    hipLaunchKernelGGL(m_thomas, dim3(ngrp * g.nz), dim3(256), (size_t)(g.ny + 2) * tl * (sizeof(float2) + sizeof(float)),
                       s, g, X.spec, kind, tl, X.kx2, X.kz2, skip);
    hipLaunchKernelGGL(m_fft_yz, dim3(ntz * g.ny), dim3(256), shz, s, g, X.spec, X.twz, 2, 1, kind, 1, tiz, X.kx2, X.ky2, X.kz2, skip);
  }
}

static size_t lds_x32(const Grid& g) { return sizeof(float2) * ((size_t)lines_for32(g.nx) * g.nx + g.nx / 2); }
static int blocks_x32(const Grid& g) { return nblk((long)g.ny * g.nz, lines_for32(g.nx)); }
// the inverse x transform reading the transposed layout stages its lines' half
// spectra in LDS as well
static size_t lds_x32_tr(const Grid& g) { return lds_x32(g) + sizeof(float2) * (size_t)lines_for32(g.nx) * g.nxh; }

static void solve32(Ctx& c, int kind, const float* q, float* x, const int* skip = nullptr) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  const int lx = lines_for32(g.nx);
  hipLaunchKernelGGL(m_fft_x_fwd, dim3(blocks_x32(g)), dim3(256), lds_x32(g), s, g, q, X.spec, X.twx, lx, skip, 0);
  solve32_yz(c, kind, skip);
  hipLaunchKernelGGL(m_fft_x_inv, dim3(blocks_x32(g)), dim3(256), lds_x32(g), s, g, X.spec, x, X.twx, lx, skip);
}

// ============================================================ projection ===
// r = b - L_h phi (FP64, boundary rules of Phi), rounded to FP32 for the
// correction solve; |r|^2 accumulated
template <bool WL>
__global__ __launch_bounds__(256) void m_residual(Grid g, const double* __restrict__ b, const double* __restrict__ phi,
                                                  float* __restrict__ r32, double* __restrict__ nrm) {
  long m = xcd_cell();
  double r = 0.0;
  if (m < g.n) {
    int i, j, k;
  cell_ijk(g, m, i, j, k);
    double p = phi[m];
#define PH(a, bb, c) ldbw<WL>(phi, g, a, bb, c, BC_P)
    double lap = (PH(i - 1, j, k) - 2.0 * p + PH(i + 1, j, k)) / (g.dx * g.dx) +
                 (PH(i, j - 1, k) - 2.0 * p + PH(i, j + 1, k)) / (g.dy * g.dy) +
                 (PH(i, j, k - 1) - 2.0 * p + PH(i, j, k + 1)) / (g.dz * g.dz);
#undef PH
    r = b[m] - lap;
    r32[m] = (float)r;
  }
  __shared__ double sh[256];
  double sum = block_sum256(r * r, sh);
  if (threadIdx.x == 0) nrm[blockIdx.x] = sum;
}

// skip the correction sweep once |r| <= tol |b|
// |b|^2 and |r|^2 from the per-block partials, and the decision to skip
// the next defect-correction sweep (|r| <= tol |b|)
__global__ __launch_bounds__(1024) void m_sum_decide(const double* __restrict__ pb, int nbb,
                                                     const double* __restrict__ pr, int nbr, double* __restrict__ red,
                                                     double tol, int* __restrict__ skip) {
  __shared__ double sh[1024];
  double vb = 0.0, vr = 0.0;
#pragma unroll 8
  for (int i = threadIdx.x; i < nbb; i += 1024) vb += pb[i];
#pragma unroll 8
  for (int i = threadIdx.x; i < nbr; i += 1024) vr += pr[i];
  sh[threadIdx.x] = vb;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double sb = sh[0];
  __syncthreads();
  sh[threadIdx.x] = vr;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    red[0] = sb;
    red[1] = sh[0];
    skip[0] = sh[0] <= tol * tol * sb ? 1 : 0;
  }
}

// ============================================================ SOR ==========
// The SOR variant (poisson, fenep.f90:848-905) as FP32 red-black sweeps of a
// correction e on the FP64 potential: with r = b - L_h phi in FP64, the
// sweeps solve L_h e = r from e = 0, so that phi + e is the iterate of the
// sweeps on L_h phi = b, and the residual a sweep takes before updating a
// cell, E = L_h e - r, is the reference's E = L_h(phi + e) - Q to FP32
// rounding of the correction.  The stopping test sqrt(sum E^2 / sum Q^2) < eps
// is the reference's, summed in FP64.  Every few sweeps phi += e in FP64 and
// the defect is taken again; once the sweeps stop, the FP64 residual of the
// new phi is held to the same tolerance, and sweeps go on if it is not met.
//
// flag[0]: converged (sweeps return at once), flag[1]: sweeps counted as the
// reference counts them, flag[2]: the FP64 check passed.

// b (FP64), r = b - L_h phi (FP64, stored in FP32), e = 0, per-block |b|^2
template <bool WL>
__global__ __launch_bounds__(256) void m_sor_start(Grid g, const float* __restrict__ U, const float* __restrict__ V,
                                                   const float* __restrict__ W, const float* __restrict__ up,
                                                   const double* __restrict__ phi,
                                                   double* __restrict__ b, float* __restrict__ r32,
                                                   float* __restrict__ e, double* __restrict__ pb) {
  const long m = xcd_cell();
  double bb = 0.0;
  if (m < g.n) {
    int i, j, k;
    cell_ijk(g, m, i, j, k);
    const double idx = 1.0 / g.dx, idy = 1.0 / g.dy, idz = 1.0 / g.dz, idt = 1.0 / g.dt;
    bb = div_rhs<WL>(g, U, V, W, up, i, j, k, idx, idy, idz, idt);
    const double p = phi[m];
#define PH(a, bq, c) ldbw<WL>(phi, g, a, bq, c, BC_P)
    const double lap = (PH(i - 1, j, k) - 2.0 * p + PH(i + 1, j, k)) / (g.dx * g.dx) +
                       (PH(i, j - 1, k) - 2.0 * p + PH(i, j + 1, k)) / (g.dy * g.dy) +
                       (PH(i, j, k - 1) - 2.0 * p + PH(i, j, k + 1)) / (g.dz * g.dz);
#undef PH
    b[m] = bb;
    r32[m] = (float)(bb - lap);
    e[m] = 0.0f;
  }
  __shared__ double sh[256];
  const double sum = block_sum256(bb * bb, sh);
  if (threadIdx.x == 0) pb[blockIdx.x] = sum;
}

// One colour of a sweep: cells with (i + j + k) % 2 == color, i fastest, as
// k_sor_color of the FP64 path.  The per-block sum of E^2 is stored by the
// first colour and added to by the second.
template <bool WL>
__global__ __launch_bounds__(256) void m_sor_sweep(Grid g, int color, float om_b0, float bx, float by, float bz, float b0,
                                                   const float* __restrict__ r32, float* __restrict__ e,
                                                   double* __restrict__ e2, const int* __restrict__ flag) {
  if (flag[0]) return;   // converged: the same for every thread of the launch
  const long h = blockIdx.x * (long)blockDim.x + threadIdx.x;
  float E = 0.0f;
  if (h < g.n / 2) {
    const long jk = h >> (g.lnx - 1);
    const int j = (int)(jk & (g.ny - 1)), k = (int)(jk >> g.lny);
    const int i = (int)(2 * (h & ((g.nx >> 1) - 1))) + ((j + k + color) & 1);
    const long m = cidx(g, i, j, k);
    const float e0 = e[m];
#define EE(a, bq, c) ldbw<WL>(e, g, a, bq, c, BC_P)
    E = (((((bz * EE(i, j, k - 1) + by * EE(i, j - 1, k)) + bx * EE(i - 1, j, k)) - b0 * e0) + bx * EE(i + 1, j, k)) +
         by * EE(i, j + 1, k)) + bz * EE(i, j, k + 1) - r32[m];
#undef EE
    e[m] = e0 + om_b0 * E;
  }
  __shared__ double sh[256];
  const double s = block_sum256((double)E * (double)E, sh);
  if (threadIdx.x == 0) e2[blockIdx.x] = color == 0 ? s : e2[blockIdx.x] + s;
}

// sum of the per-block partials of n blocks, in a fixed order
__device__ __forceinline__ double sum_partials1024(const double* __restrict__ p, int n, double* sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 1024) v += p[i];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  return sh[0];
}

// mode 0 (after m_sor_start): red[0] = sum Q^2, flags cleared (converged at
// once when Q = 0).  mode 1 (after a sweep): count the sweep and test it.
// mode 2 (after the FP64 residual of phi + e, only once the sweeps have met
// the test): red[1] = |b - L_h phi|^2, flag[2] = the FP64 test; the sweeps
// resume when it fails.
__global__ __launch_bounds__(1024) void m_sor_decide(const double* __restrict__ part, int nb, double* __restrict__ red,
                                                     double eps, int mode, int* __restrict__ flag) {
  if (mode == 2 && !flag[0]) return;   // the same for every thread
  __shared__ double sh[1024];
  const double sum = sum_partials1024(part, nb, sh);
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    red[0] = sum;
    flag[0] = sum == 0.0 ? 1 : 0;
    flag[1] = 0;
    flag[2] = 0;
  } else if (mode == 1) {
    if (!flag[0]) {
      flag[1] += 1;
      if (sqrt(sum / red[0]) < eps) flag[0] = 1;
    }
  } else {
    red[1] = sum;
    const int ok = red[0] == 0.0 || sqrt(sum / red[0]) < eps;
    flag[2] = ok;
    flag[0] = ok;
  }
}

// phi += e (FP64), e = 0
__global__ __launch_bounds__(256) void m_sor_accum(long n, double* __restrict__ phi, float* __restrict__ e) {
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m < n) {
    phi[m] += (double)e[m];
    e[m] = 0.0f;
  }
}

// One fixed-point sweep of the Crank-Nicolson Helmholtz problem
// [1 - eps lap_h] x = q for the three velocity components at once:
// x_out = q + eps lap_h x_in (x_in = q on the first sweep).  With
// eps * |lap_h| << 1 the error contracts by eps*lambda_max per sweep.
template <bool WL>
__global__ __launch_bounds__(256) void m_helm_sweep(Grid g, float eps, const float* __restrict__ q,
                                                    const float* __restrict__ xin, float* __restrict__ xout) {
  long m = xcd_cell();
  if (m >= g.n) return;
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  const float ix2 = (float)(1.0 / (g.dx * g.dx)), iy2 = (float)(1.0 / (g.dy * g.dy)), iz2 = (float)(1.0 / (g.dz * g.dz));
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const float* X = xin + c * g.n;
    int kind = c == 0 ? BC_U : (c == 1 ? BC_V : BC_W);
    float x0 = ldbw<WL>(X, g, i, j, k, kind);
    float lap = (ldbw<WL>(X, g, i - 1, j, k, kind) - 2.0f * x0 + ldbw<WL>(X, g, i + 1, j, k, kind)) * ix2 +
                (ldbw<WL>(X, g, i, j - 1, k, kind) - 2.0f * x0 + ldbw<WL>(X, g, i, j + 1, k, kind)) * iy2 +
                (ldbw<WL>(X, g, i, j, k - 1, kind) - 2.0f * x0 + ldbw<WL>(X, g, i, j, k + 1, kind)) * iz2;
    float v = q[c * g.n + m] + eps * lap;
    if (c == 1 && WL && j == g.ny - 1) v = 0.0f;   // wall face of V
    xout[c * g.n + m] = v;
  }
}

// Two sweeps in one pass, streamed along z: x1 = q + eps L q on the tile and
// a one-cell halo, one plane ahead, then x2 = q + eps L x1 on the tile.  The
// LDS holds four planes of q with a two-cell halo (the plane being filled
// among them) and three planes of x1 with a one-cell halo, the three
// components each.  Values follow the boundary-rule accessor: q's wall
// ghosts and the wall face of V on filling, and x1's ghost rows by the same
// rules from x1 at the nearest row, which the ghost's thread computes itself.
// HBM traffic: q read once and x2 written once (24 B a cell for both sweeps
// instead of 60).
#define HZ_TX 64
#define HZ_TY 4
#define HZ_THREADS (HZ_TX * HZ_TY)
#define HZ_QX (HZ_TX + 4)
#define HZ_QP (HZ_QX * (HZ_TY + 4))
#define HZ_XX (HZ_TX + 2)
#define HZ_XP (HZ_XX * (HZ_TY + 2))
#define HZ_QPER ((HZ_QP + HZ_THREADS - 1) / HZ_THREADS)
#define HZ_KC 32

struct HzRegs {
  float v[HZ_QPER][3];
};

template <bool WL>
__device__ __forceinline__ void hz_fetch(const Grid& g, int i0, int j0, int kk, const float* __restrict__ q, HzRegs& r) {
  const long plane = (long)g.nx * g.ny;
  const int kz = wrapi(kk, g.nz);
#pragma unroll
  for (int e = 0; e < HZ_QPER; e++) {
    const int pt = threadIdx.x + HZ_THREADS * e;
    if (pt < HZ_QP) {
      const int ii = wrapi(i0 + pt % HZ_QX - 2, g.nx);
      int jj = j0 + pt / HZ_QX - 2;
      jj = WL ? (jj < 0 ? 0 : (jj >= g.ny ? g.ny - 1 : jj)) : wrapi(jj, g.ny);
      const long b = ii + (long)g.nx * jj + plane * kz;
#pragma unroll
      for (int c = 0; c < 3; c++) r.v[e][c] = q[c * g.n + b];
    }
  }
}

// q values as the boundary-rule accessor returns them (U_kabe, V_kabe, W_kabe)
template <bool WL>
__device__ __forceinline__ float hz_rule(const Grid& g, int c, int jj, float v) {
  if (!WL) return v;
  const bool lo = jj < 0, hi = jj >= g.ny;
  if (c == 1) return (lo || hi || jj == g.ny - 1) ? 0.0f : v;
  return lo ? -v : (hi ? 2.0f * 0.0f - v : v);
}

template <bool WL>
__device__ __forceinline__ void hz_put(const Grid& g, int j0, const HzRegs& r, float* __restrict__ slot) {
#pragma unroll
  for (int e = 0; e < HZ_QPER; e++) {
    const int pt = threadIdx.x + HZ_THREADS * e;
    if (pt < HZ_QP) {
      const int jj = j0 + pt / HZ_QX - 2;
#pragma unroll
      for (int c = 0; c < 3; c++) slot[c * HZ_QP + pt] = hz_rule<WL>(g, c, jj, r.v[e][c]);
    }
  }
}

// one sweep at the x1-halo points (hx, hy) of plane p from the q planes
// p-1, p, p+1: q + eps L q, the wall face of V set to zero, ghost rows by the
// rules from the nearest row.  A thread's points are pt = threadIdx.x +
// e HZ_THREADS, e = 0, 1; qce holds their q-tile offsets, taken where no point
// is a ghost (WL = false).
static_assert(HZ_XP <= 2 * HZ_THREADS, "two x1 points per thread");
__device__ __forceinline__ void hz_x1_offsets(int (&qce)[2]) {
#pragma unroll
  for (int e = 0; e < 2; e++) {
    const int pt = threadIdx.x + HZ_THREADS * e;
    qce[e] = (pt / HZ_XX + 1) * HZ_QX + pt % HZ_XX + 1;
  }
}
template <bool WL>
__device__ __forceinline__ void hz_x1(const Grid& g, int j0, float eps, float ix2, float iy2, float iz2,
                                      const float* __restrict__ qm, const float* __restrict__ q0,
                                      const float* __restrict__ qp, float* __restrict__ xs, const int (&qce)[2]) {
#pragma unroll
  for (int e = 0; e < 2; e++) {
    const int pt = threadIdx.x + HZ_THREADS * e;
    if (pt >= HZ_XP) break;
    const int hx = pt % HZ_XX, hy = pt / HZ_XX;
    const int jj = j0 + hy - 1;
    const bool lo = WL && jj < 0, hi = WL && jj >= g.ny;
    // the row whose sweep value this point takes (itself, or the nearest
    // interior row of a ghost)
    const int hyc = lo ? hy + 1 : (hi ? hy - 1 : hy);
    const int jc = j0 + hyc - 1;
    const int qc = WL ? (hyc + 1) * HZ_QX + hx + 1 : qce[e];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float* Qm = qm + c * HZ_QP;
      const float* Q0 = q0 + c * HZ_QP;
      const float* Qp = qp + c * HZ_QP;
      const float x0 = Q0[qc];
      const float lap = (Q0[qc - 1] - 2.0f * x0 + Q0[qc + 1]) * ix2 + (Q0[qc - HZ_QX] - 2.0f * x0 + Q0[qc + HZ_QX]) * iy2 +
                        (Qm[qc] - 2.0f * x0 + Qp[qc]) * iz2;
      float v = x0 + eps * lap;
      if (c == 1 && WL && jc == g.ny - 1) v = 0.0f;   // wall face of V
      xs[c * HZ_XP + pt] = (lo || hi) ? hz_rule<WL>(g, c, jj, v) : v;
    }
  }
}

template <bool WL>
__device__ __forceinline__ void helm_2sweeps_tile(const Grid& g, float eps, const float* __restrict__ q,
                                                  float* __restrict__ x, int i0, int j0, int k0, float (*qs)[3 * HZ_QP],
                                                  float (*xs)[3 * HZ_XP]) {
  const int k1 = min(k0 + HZ_KC, g.nz);
  const int tx = threadIdx.x % HZ_TX, ty = threadIdx.x / HZ_TX;
  const float ix2 = (float)(1.0 / (g.dx * g.dx)), iy2 = (float)(1.0 / (g.dy * g.dy)), iz2 = (float)(1.0 / (g.dz * g.dz));
  // q plane p in slot (p - k0 + 2) mod 4, x1 plane p in slot (p - k0 + 1) mod 3
  HzRegs r;
#pragma unroll 1
  for (int s = 0; s < 4; s++) {
    hz_fetch<WL>(g, i0, j0, k0 - 2 + s, q, r);
    hz_put<WL>(g, j0, r, qs[s]);
  }
  if (k0 + 2 <= k1 + 1) hz_fetch<WL>(g, i0, j0, k0 + 2, q, r);
  __syncthreads();
  int qce[2];
  hz_x1_offsets(qce);
  hz_x1<WL>(g, j0, eps, ix2, iy2, iz2, qs[0], qs[1], qs[2], xs[0], qce);   // x1 at k0-1
  hz_x1<WL>(g, j0, eps, ix2, iy2, iz2, qs[1], qs[2], qs[3], xs[1], qce);   // x1 at k0
  __syncthreads();
  if (k0 + 2 <= k1 + 1) hz_put<WL>(g, j0, r, qs[0]);                  // q at k0+2
  if (k0 + 3 <= k1 + 1) hz_fetch<WL>(g, i0, j0, k0 + 3, q, r);
  __syncthreads();
#pragma unroll 1
  for (int kk = k0; kk < k1; kk++) {
    const int sq = kk - k0 + 2, sx = kk - k0 + 1;
    // x1 at kk+1 from q at kk, kk+1, kk+2 (into the slot of x1 at kk-2)
    hz_x1<WL>(g, j0, eps, ix2, iy2, iz2, qs[sq & 3], qs[(sq + 1) & 3], qs[(sq + 2) & 3], xs[(sx + 1) % 3], qce);
    __syncthreads();
    {
      // x2 at kk from x1 at kk-1, kk, kk+1 and q at kk
      const float* Xm = xs[(sx + 2) % 3];
      const float* X0 = xs[sx % 3];
      const float* Xp = xs[(sx + 1) % 3];
      const int xc = (ty + 1) * HZ_XX + tx + 1, qc = (ty + 2) * HZ_QX + tx + 2;
      const int j = j0 + ty;
      const long m = cidx(g, i0 + tx, j, kk);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const float* A = X0 + c * HZ_XP;
        const float x0 = A[xc];
        const float lap = (A[xc - 1] - 2.0f * x0 + A[xc + 1]) * ix2 + (A[xc - HZ_XX] - 2.0f * x0 + A[xc + HZ_XX]) * iy2 +
                          (Xm[c * HZ_XP + xc] - 2.0f * x0 + Xp[c * HZ_XP + xc]) * iz2;
        float v = qs[sq & 3][c * HZ_QP + qc] + eps * lap;
        if (c == 1 && WL && j == g.ny - 1) v = 0.0f;   // wall face of V
        x[c * g.n + m] = v;
      }
    }
    // q at kk+3 into the fourth slot (read from the next plane on), then
    // fetch q at kk+4
    if (kk + 3 <= k1 + 1) hz_put<WL>(g, j0, r, qs[(sq + 3) & 3]);
    if (kk + 4 <= k1 + 1) hz_fetch<WL>(g, i0, j0, kk + 4, q, r);
    __syncthreads();
  }
}

// A tile whose q rows (j0-2 .. j0+HZ_TY+1) all lie inside the walls and
// short of the wall face of V (row ny-1) takes no boundary rule: it runs the
// periodic path, whose values are then the same.
template <bool WL>
__global__ __launch_bounds__(HZ_THREADS) void m_helm_2sweeps_z(Grid g, float eps, const float* __restrict__ q,
                                                              float* __restrict__ x) {
  __shared__ float qs[4][3 * HZ_QP];
  __shared__ float xs[3][3 * HZ_XP];
  const int ntx = g.nx / HZ_TX, nty = g.ny / HZ_TY;
  const long b = xcd_block();
  const int bx = (int)(b % ntx), by = (int)((b / ntx) % nty), bz = (int)(b / ((long)ntx * nty));
  const int i0 = bx * HZ_TX, j0 = by * HZ_TY, k0 = bz * HZ_KC;
  if (WL && (j0 < 2 || j0 + HZ_TY + 1 >= g.ny - 1)) helm_2sweeps_tile<true>(g, eps, q, x, i0, j0, k0, qs, xs);
  else helm_2sweeps_tile<false>(g, eps, q, x, i0, j0, k0, qs, xs);
}

// Helmholtz solves of the three predictor components: fixed-point sweeps when
// eps*lambda_max is small, the spectral (+ line) solve otherwise.  q and the
// result x are 3-plane buffers; x may not alias q.
static void helmholtz32(Ctx& c, const float* q, float* x) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  const long n = g.n;
  if (X.helm_iters <= 0) {
    for (int k = 0; k < 3; k++) solve32(c, k, q + k * n, x + k * n);
    return;
  }
  const float eps = (float)(((g.dt * g.beta) / g.re) / 2.0);
  if (X.helm_iters == 2 && g.nx % HZ_TX == 0 && g.ny % HZ_TY == 0 && g.nz >= 3) {
    const long grid = (long)(g.nx / HZ_TX) * (g.ny / HZ_TY) * ((g.nz + HZ_KC - 1) / HZ_KC);
    hipLaunchKernelGGL(g.walls ? m_helm_2sweeps_z<true> : m_helm_2sweeps_z<false>, dim3(grid), dim3(HZ_THREADS), 0,
                       c.stream, g, eps, q, x);
    return;
  }
  const float* in = q;
  for (int it = 0; it < X.helm_iters; it++) {
    // the last sweep writes x; earlier ones alternate so that it does
    float* out = ((X.helm_iters - 1 - it) % 2 == 0) ? x : X.X2;
    if (g.walls) hipLaunchKernelGGL(m_helm_sweep<true>, dim3(nblk(n, 256)), dim3(256), 0, c.stream, g, eps, q, in, out);
    else hipLaunchKernelGGL(m_helm_sweep<false>, dim3(nblk(n, 256)), dim3(256), 0, c.stream, g, eps, q, in, out);
    in = out;
  }
}

// fft_march (fenep.f90:1562-1604) with FP64 potential, FP32 storage
template <bool WL>
__global__ __launch_bounds__(256) void m_march(Grid g, const float* __restrict__ up, const double* __restrict__ phi,
                                               float* U, float* V, float* W, float* __restrict__ P) {
  long m = xcd_cell();
  if (m >= g.n) return;
  int i, j, k;
  cell_ijk(g, m, i, j, k);
  const double dt = g.dt;
  const float *Du = up, *Dv = up + g.n, *Dw = up + 2 * g.n;
#define PH(a, b, c) ldbw<WL>(phi, g, a, b, c, BC_P)
  double ph = phi[m];
  // u^{n+1} = (u^n + d) - dt grad Phi in FP64, one rounding; each thread
  // reads and writes its own cell of U, V, W only
  const double us = ustar<WL>(g, U, Du, i, j, k, BC_U), vs = ustar<WL>(g, V, Dv, i, j, k, BC_V),
               ws = ustar<WL>(g, W, Dw, i, j, k, BC_W);
  U[m] = (float)(us - dt * (PH(i + 1, j, k) - ph) / g.dx);
  V[m] = (float)(vs - dt * (PH(i, j + 1, k) - ph) / g.dy);
  W[m] = (float)(ws - dt * (PH(i, j, k + 1) - ph) / g.dz);
  const double coef = ((dt * g.beta) / g.re) / 2.0;
  double lap = (PH(i - 1, j, k) - 2.0 * ph + PH(i + 1, j, k)) / (g.dx * g.dx) +
               (PH(i, j - 1, k) - 2.0 * ph + PH(i, j + 1, k)) / (g.dy * g.dy) +
               (PH(i, j, k - 1) - 2.0 * ph + PH(i, j, k + 1)) / (g.dz * g.dz);
#undef PH
  P[m] = (float)(((double)P[m] + ph) - coef * lap);
}

// The same correction streamed along z: a 64 x 16 tile walks 8 planes with
// three planes of Phi (FP64, one-cell halo, the boundary rule of P applied)
// in LDS and the next plane prefetched; u^n, the increment d, P and the
// outputs are the cell's own, read and written once.
#define MZ_TX 64
#define MZ_TY 16
#define MZ_THREADS (MZ_TX * MZ_TY)
#define MZ_HX (MZ_TX + 2)
#define MZ_PTS (MZ_HX * (MZ_TY + 2))
#define MZ_PER ((MZ_PTS + MZ_THREADS - 1) / MZ_THREADS)
#define MZ_KC 8                   // planes per block (64 x 4 tiles: +3 us per step; 16 planes: +3)

template <bool WL>
__device__ __forceinline__ void mz_fetch(const Grid& g, int i0, int j0, int kk, const double* __restrict__ phi,
                                         double (&r)[MZ_PER]) {
  const long plane = (long)g.nx * g.ny;
  const int kz = wrapi(kk, g.nz);
#pragma unroll
  for (int e = 0; e < MZ_PER; e++) {
    const int pt = threadIdx.x + MZ_THREADS * e;
    if (pt < MZ_PTS) {
      const int ii = wrapi(i0 + pt % MZ_HX - 1, g.nx);
      int jj = j0 + pt / MZ_HX - 1;
      // P_kabe: a ghost row takes the nearest row (ldbw, BC_P)
      jj = WL ? (jj < 0 ? 0 : (jj >= g.ny ? g.ny - 1 : jj)) : wrapi(jj, g.ny);
      r[e] = phi[ii + (long)g.nx * jj + plane * kz];
    }
  }
}

__device__ __forceinline__ void mz_put(const double (&r)[MZ_PER], double* __restrict__ slot) {
#pragma unroll
  for (int e = 0; e < MZ_PER; e++) {
    const int pt = threadIdx.x + MZ_THREADS * e;
    if (pt < MZ_PTS) slot[pt] = r[e];
  }
}

// grid: (nx / MZ_TX) (ny / MZ_TY) ceil(nz / MZ_KC) blocks of MZ_THREADS
template <bool WL>
__global__ __launch_bounds__(MZ_THREADS) void m_march_z(Grid g, const float* __restrict__ up,
                                                        const double* __restrict__ phi, float* __restrict__ U,
                                                        float* __restrict__ V, float* __restrict__ W,
                                                        float* __restrict__ P) {
  __shared__ double ps[3][MZ_PTS];
  const int ntx = g.nx / MZ_TX, nty = g.ny / MZ_TY;
  const long b = xcd_block();
  const int bx = (int)(b % ntx), by = (int)((b / ntx) % nty), bz = (int)(b / ((long)ntx * nty));
  const int i0 = bx * MZ_TX, j0 = by * MZ_TY, k0 = bz * MZ_KC;
  const int k1 = min(k0 + MZ_KC, g.nz);
  const int tx = threadIdx.x % MZ_TX, ty = threadIdx.x / MZ_TX;
  const int i = i0 + tx, j = j0 + ty;
  const double dt = g.dt;
  const double coef = ((dt * g.beta) / g.re) / 2.0;
  // reciprocals of the spacings (FP64 divisions cost some twenty instructions)
  const double idx = 1.0 / g.dx, idy = 1.0 / g.dy, idz = 1.0 / g.dz;
  const double idx2 = 1.0 / (g.dx * g.dx), idy2 = 1.0 / (g.dy * g.dy), idz2 = 1.0 / (g.dz * g.dz);
  const float *Uu = up, *Vv = up + g.n, *Ww = up + 2 * g.n;
  double r[MZ_PER];
  // plane p in slot (p - k0 + 1) mod 3
#pragma unroll 1
  for (int s = 0; s < 3; s++) {
    mz_fetch<WL>(g, i0, j0, k0 - 1 + s, phi, r);
    mz_put(r, ps[s]);
  }
  if (k0 + 2 <= k1) mz_fetch<WL>(g, i0, j0, k0 + 2, phi, r);
  const int c = (ty + 1) * MZ_HX + tx + 1;
  int sm = 0, s0 = 1, sp = 2;
#pragma unroll 1
  for (int kk = k0; kk < k1; kk++) {
    __syncthreads();
    const long m = cidx(g, i, j, kk);
    // u* = u^n + d, exact in FP64 (V_kabe: the wall face reads 0)
    const double u = (double)U[m] + (double)Uu[m];
    const double v = (WL && j == g.ny - 1) ? 0.0 : (double)V[m] + (double)Vv[m];
    const double w = (double)W[m] + (double)Ww[m];
    const float pold = P[m];
    const double* A = ps[s0];
    const double ph = A[c];
    U[m] = (float)(u - dt * (A[c + 1] - ph) * idx);
    V[m] = (float)(v - dt * (A[c + MZ_HX] - ph) * idy);
    W[m] = (float)(w - dt * (ps[sp][c] - ph) * idz);
    const double lap = (A[c - 1] - 2.0 * ph + A[c + 1]) * idx2 + (A[c - MZ_HX] - 2.0 * ph + A[c + MZ_HX]) * idy2 +
                       (ps[sm][c] - 2.0 * ph + ps[sp][c]) * idz2;
    P[m] = (float)(((double)pold + ph) - coef * lap);
    __syncthreads();
    if (kk + 2 <= k1) mz_put(r, ps[sm]);
    if (kk + 3 <= k1) mz_fetch<WL>(g, i0, j0, kk + 3, phi, r);
    const int t = sm; sm = s0; s0 = sp; sp = t;
  }
}

// ================================================================== IBM ===
__device__ __forceinline__ double roma64(double r) {
  double a = fabs(r);
  if (a < 1.5 && a >= 0.5) return ((5.0 - 3.0 * a) - sqrt(-(3.0 * ((1.0 - a) * (1.0 - a))) + 1.0)) / 6.0;
  if (a < 0.5) return (1.0 + sqrt(-(3.0 * (r * r)) + 1.0)) / 3.0;
  return 0.0;
}
// delta_h of the marker at a staggered point times the weight w (dX^3 for
// the interpolation; the spreading weight g.sw for the spread)
__device__ __forceinline__ double ibm_w64(const Grid& g, int c, int ii, int jj, int kk, double xc, double yc, double zc,
                                          double w) {
  double x = ((double)ii - 0.5) * g.dx, y = ((double)jj - 0.5) * g.dy, z = ((double)kk - 0.5) * g.dz;
  if (c == 0) x += 0.5 * g.dx;
  else if (c == 1) y += 0.5 * g.dy;
  else z += 0.5 * g.dz;
  return roma64((x - xc) / g.dx) * roma64((y - yc) / g.dy) * roma64((z - zc) / g.dz) / (g.dx * g.dy * g.dz) * w;
}

__global__ void m_ibm_setup_interp(Grid g, long nmk, int nc, int nl, const double* __restrict__ mk,
                                   float* __restrict__ wi, int4* __restrict__ org, float* __restrict__ uc) {
  long id = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (id >= nmk) return;
  int n = (int)(id / ((long)nc * nl)) + 1;
  double xc = mk[id], yc = mk[nmk + id], zc = mk[2 * nmk + id];
  for (int c = 0; c < 3; c++) {
    int i0, j0;
    if (c == 0) { i0 = (int)(xc / g.dx - 0.5); j0 = (int)(yc / g.dy); }
    else if (c == 1) { i0 = (int)(xc / g.dx); j0 = (int)(yc / g.dy - 0.5); }
    else { i0 = (int)(xc / g.dx); j0 = (int)(yc / g.dy); }
    org[c * nmk + id] = make_int4(i0, j0, n, 0);
    int s = 0;
    // slot-major: weight s of (component c, marker id) at (c*27 + s)*nmk + id
    for (int kk = n - 1; kk <= n + 1; kk++)
      for (int jj = j0; jj <= j0 + 2; jj++)
        for (int ii = i0; ii <= i0 + 2; ii++, s++)
          wi[(long)(c * 27 + s) * nmk + id] = (float)ibm_w64(g, c, ii, jj, kk, xc, yc, zc, g.dx * g.dx * g.dx);
    uc[c * nmk + id] = (float)mk[(3 + c) * nmk + id];
  }
}

__global__ void m_ibm_setup_spread(Grid g, int comp, int rows, const int* __restrict__ rowp, const long* __restrict__ tgt,
                                   const int* __restrict__ ent, long nmk, const double* __restrict__ mk,
                                   float* __restrict__ sw) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const long gx = g.nx + 2, gy = g.ny + 2;
  long t = tgt[r];
  int ii = (int)(t % gx), jj = (int)((t / gx) % gy), kk = (int)(t / (gx * gy));
  for (int e = rowp[r]; e < rowp[r + 1]; e++) {
    long id = ent[e];
    sw[e] = (float)ibm_w64(g, comp, ii, jj, kk, mk[id], mk[nmk + id], mk[2 * nmk + id], g.sw);
  }
}

// interpolation of Ua = u^n + d_q at the markers (the two summed apart) and
// the direct forcing Fc = (Uc - Ub)/dt (ibm_Helmholtz, fenep.f90:2076-2119),
// one thread per (component, marker)
template <bool WL>
__global__ __launch_bounds__(256) void m_ibm_interp(Grid g, long nmk, const float* __restrict__ wi, const int4* __restrict__ org,
                                                    const float* __restrict__ uc, const float* __restrict__ U,
                                                    const float* __restrict__ V, const float* __restrict__ W,
                                                    const float* __restrict__ dq, float* __restrict__ mf) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 3 * nmk) return;
  const int c = (int)(t / nmk);
  const long id = t - c * nmk;
  const float idt = (float)(1.0 / g.dt);
  const int4 o = org[t];
  const float* X = c == 0 ? U : (c == 1 ? V : W);   // three separate planes of the state
  const float* D = dq + c * g.n;
  const int kind = c == 0 ? BC_U : (c == 1 ? BC_V : BC_W);
  const float* w = wi + (long)c * 27 * nmk + id;
  float acc = 0.0f, accd = 0.0f;
  int s = 0;
#pragma unroll
  for (int kk = 0; kk < 3; kk++)
#pragma unroll
    for (int jj = 0; jj < 3; jj++)
#pragma unroll
      for (int ii = 0; ii < 3; ii++, s++) {
        const float wv = w[(long)s * nmk];
        acc += wv * ldbw<WL>(X, g, o.x + ii - 1, o.y + jj - 1, o.z - 1 + kk - 1, kind);
        accd += wv * ldbw<WL>(D, g, o.x + ii - 1, o.y + jj - 1, o.z - 1 + kk - 1, kind);
      }
  mf[t] = ((uc[t] - acc) - accd) * idt;
}

// entries of a sliced-ELL slice come in groups of this many
#define ELL_UNROLL 8

// spreading gathered onto interior targets with the ghost images folded in
// (ibm_PBM, fenep.f90:2195-2205), FP32 accumulation
__global__ __launch_bounds__(256) void m_ibm_spread_folded(int rows, const int* __restrict__ slice_off,
                                                           const int* __restrict__ slice_len, const long* __restrict__ tgt,
                                                           const int* __restrict__ ent, const float* __restrict__ w,
                                                           const float* __restrict__ mf, float dt,
                                                           float* __restrict__ rhs) {
  // sliced ELL: the rows of a 64-row slice store their entries slot-major
  // (entry k of lane l at off + 64 k + l), padded with zero weights to a
  // multiple of ELL_UNROLL entries, so each wave reads its entries with
  // contiguous loads, ELL_UNROLL entries' loads in flight together
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int sl = r >> 6, lane = r & 63;
  const int off = slice_off[sl], len = slice_len[sl];
  float acc = 0.0f;
  for (int k = 0; k < len; k += ELL_UNROLL) {
    int pk[ELL_UNROLL];
    float wv[ELL_UNROLL];
#pragma unroll
    for (int u = 0; u < ELL_UNROLL; u++) {
      const int e = off + 64 * (k + u) + lane;
      pk[u] = ent[e];
      wv[u] = w[e];
    }
#pragma unroll
    for (int u = 0; u < ELL_UNROLL; u++) acc += mf[pk[u]] * wv[u];
  }
  rhs[tgt[r]] += dt * acc;   // each target appears in one row only
}

// The same two passes for markers repeated along z.  Interpolation: the
// weights and the in-plane origin of the layer-0 marker, the layer from the
// marker index (markers are numbered layer by layer, zP to a layer).  One
// lane per (component, marker), consecutive lanes on consecutive markers:
// spreading the 27 points over three lanes or three waves per marker measured
// 14 us slower per step (more cache lines per load instruction).
template <bool WL>
__global__ __launch_bounds__(256) void m_ibm_interp_z(Grid g, long nmk, long zP, const float* __restrict__ wz,
                                                      const int2* __restrict__ orgz, const float* __restrict__ uc,
                                                      const float* __restrict__ U, const float* __restrict__ V,
                                                      const float* __restrict__ W, const float* __restrict__ dq,
                                                      float* __restrict__ mf) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 3 * nmk) return;
  const int c = (int)(t / nmk);
  const long id = t - c * nmk;
  const int L = (int)(id / zP);
  const long p = id - (long)L * zP;
  const float idt = (float)(1.0 / g.dt);
  const int2 o = orgz[c * zP + p];
  const float* X = c == 0 ? U : (c == 1 ? V : W);   // three separate planes of the state
  const float* D = dq + c * g.n;
  const int kind = c == 0 ? BC_U : (c == 1 ? BC_V : BC_W);
  const float* w = wz + (long)c * 27 * zP + p;
  float acc = 0.0f, accd = 0.0f;
  int s = 0;
#pragma unroll
  for (int kk = 0; kk < 3; kk++)
#pragma unroll
    for (int jj = 0; jj < 3; jj++)
#pragma unroll
      for (int ii = 0; ii < 3; ii++, s++) {
        const float wv = w[(long)s * zP];
        acc += wv * ldbw<WL>(X, g, o.x + ii - 1, o.y + jj - 1, L + kk - 1, kind);
        accd += wv * ldbw<WL>(D, g, o.x + ii - 1, o.y + jj - 1, L + kk - 1, kind);
      }
  mf[t] = ((uc[t] - acc) - accd) * idt;
}

// The interpolation of LC consecutive layers of one layer-0 marker and
// component in one lane: the 3 x 3 x 3 stencils of layers L and L+1 share two
// planes, so the lane loads each of the LC + 2 planes' 3 x 3 values of u^n and
// of the increment once (18 LC + 36 loads for LC markers instead of 54 LC)
// and adds each marker's 27 terms in the order of m_ibm_interp_z (plane by
// plane, then j, then i), so that the forces are the same bit for bit.
#define IZ_LC 4
template <bool WL>
__global__ __launch_bounds__(256) void m_ibm_interp_zc(Grid g, long nmk, long zP, const float* __restrict__ wz,
                                                       const int2* __restrict__ orgz, const float* __restrict__ uc,
                                                       const float* __restrict__ U, const float* __restrict__ V,
                                                       const float* __restrict__ W, const float* __restrict__ dq,
                                                       float* __restrict__ mf) {
  const long nq = g.nz / IZ_LC;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 3 * nq * zP) return;
  const int c = (int)(t / (nq * zP));
  const long rem = t - (long)c * nq * zP;
  const int q = (int)(rem / zP);
  const long p = rem - (long)q * zP;
  const float idt = (float)(1.0 / g.dt);
  const int2 o = orgz[c * zP + p];
  const float* X = c == 0 ? U : (c == 1 ? V : W);
  const float* D = dq + c * g.n;
  const int kind = c == 0 ? BC_U : (c == 1 ? BC_V : BC_W);
  float w[27];
#pragma unroll
  for (int s = 0; s < 27; s++) w[s] = wz[((long)c * 27 + s) * zP + p];
  const int L0 = q * IZ_LC;
  float acc[IZ_LC], accd[IZ_LC];
#pragma unroll
  for (int l = 0; l < IZ_LC; l++) { acc[l] = 0.0f; accd[l] = 0.0f; }
  // plane P = L0 - 1 + pp serves the markers L = P + 1 - kk, kk = 0, 1, 2
#pragma unroll
  for (int pp = 0; pp < IZ_LC + 2; pp++) {
    float xv[9], dv[9];
#pragma unroll
    for (int jj = 0; jj < 3; jj++)
#pragma unroll
      for (int ii = 0; ii < 3; ii++) {
        xv[jj * 3 + ii] = ldbw<WL>(X, g, o.x + ii - 1, o.y + jj - 1, L0 - 1 + pp, kind);
        dv[jj * 3 + ii] = ldbw<WL>(D, g, o.x + ii - 1, o.y + jj - 1, L0 - 1 + pp, kind);
      }
#pragma unroll
    for (int kk = 2; kk >= 0; kk--) {
      const int l = pp - kk;   // marker L0 + l takes this plane as its plane kk
      if (l < 0 || l >= IZ_LC) continue;
#pragma unroll
      for (int e = 0; e < 9; e++) {
        acc[l] += w[kk * 9 + e] * xv[e];
        accd[l] += w[kk * 9 + e] * dv[e];
      }
    }
  }
#pragma unroll
  for (int l = 0; l < IZ_LC; l++) {
    const long tm = (long)c * nmk + (long)(L0 + l) * zP + p;
    mf[tm] = ((uc[tm] - acc[l]) - accd[l]) * idt;
  }
}

// Spreading: thread (plane k, row r of plane 0); an entry names the layer-0
// marker and the layer offset d in {-1, 0, 1} (periodic in z), so the marker
// it takes is that of layer k + d.  Rows are padded to whole slices per plane.
__global__ __launch_bounds__(256) void m_ibm_spread_z(int rows, int rows64, int nz, long zP, long plane,
                                                      const int* __restrict__ slice_off,
                                                      const int* __restrict__ slice_len, const long* __restrict__ tgt,
                                                      const int* __restrict__ ent, const float* __restrict__ w,
                                                      const float* __restrict__ mf, float dt, float* __restrict__ rhs) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int k = (int)(t / rows64);
  const int r = (int)(t - (long)k * rows64);
  if (k >= nz || r >= rows) return;
  const int sl = r >> 6, lane = r & 63;
  const int off = slice_off[sl], len = slice_len[sl];
  float acc = 0.0f;
  for (int q = 0; q < len; q += ELL_UNROLL) {
    int pk[ELL_UNROLL];
    float wv[ELL_UNROLL];
#pragma unroll
    for (int u = 0; u < ELL_UNROLL; u++) {
      const int e = off + 64 * (q + u) + lane;
      pk[u] = ent[e];
      wv[u] = w[e];
    }
#pragma unroll
    for (int u = 0; u < ELL_UNROLL; u++) {
      int L = k + (pk[u] & 3) - 1;
      L = L < 0 ? L + nz : (L >= nz ? L - nz : L);
      acc += mf[(long)(pk[u] >> 2) + (long)L * zP] * wv[u];
    }
  }
  rhs[tgt[r] + (long)k * plane] += dt * acc;   // each target appears in one row only
}

// ========================================================== conversions ===
__global__ void m_to32(long n, const double* __restrict__ a, float* __restrict__ b) {
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m < n) b[m] = (float)a[m];
}
__global__ void m_to64(long n, const float* __restrict__ a, double* __restrict__ b) {
  long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m < n) b[m] = (double)a[m];
}

// ================================================================= host ===
// Fixed-point Helmholtz sweeps x <- q + eps Lap x from x = q: the error
// starts at eps Lap x* (<= rho |x*|, rho = eps lambda_max) and contracts by
// rho per sweep, so k sweeps leave rho^(k+1) |x*|; enough sweeps to bring that
// below the FP32 unit roundoff 2^-24 of the stored result, 0 (the spectral
// solve) if more than 6 would be needed.
static int helm_sweeps(const Grid& g, int requested) {
  if (requested > 0) return std::min(requested, 16);
  if (requested < 0) return 0;
  const double eps = ((g.dt * g.beta) / g.re) / 2.0;
  const double lam = 4.0 / (g.dx * g.dx) + 4.0 / (g.dy * g.dy) + 4.0 / (g.dz * g.dz);
  const double rho = eps * lam;
  int m = rho > 0.0 ? (int)std::ceil(std::log(std::ldexp(1.0, -24)) / std::log(rho)) - 1 : 1;
  return (rho < 0.25 && m <= 6) ? std::max(1, m) : 0;
}

void mixed_create(Ctx& c) {
  if (c.p.method == FP_METHOD_EXPLICIT)
    throw std::string("mixed precision supports the FFT and IBM methods");
  Mixed* X = new Mixed();
  c.mixed = X;
  const Grid& g = c.g;
  const long n = g.n;
  FP_HIP_CHECK(hipStreamCreateWithFlags(&X->side, hipStreamNonBlocking));
  // the two streams' handoffs stay on the device: a device-scope release
  // instead of the system-scope one
  FP_HIP_CHECK(hipEventCreateWithFlags(&X->ev_cstar, hipEventDisableTiming | hipEventReleaseToDevice));
  FP_HIP_CHECK(hipEventCreateWithFlags(&X->sel_done, hipEventDisableTiming | hipEventReleaseToDevice));
  dalloc(&X->U, n); dalloc(&X->V, n); dalloc(&X->W, n); dalloc(&X->P, n);
  dalloc(&X->C[0], 6 * n); dalloc(&X->C[1], 6 * n); dalloc(&X->f, n);
  dalloc(&X->A0, 3 * n); dalloc(&X->T0, 3 * n); dalloc(&X->B, 3 * n);
  dalloc(&X->Q, 3 * n); dalloc(&X->Up, 3 * n);
  dalloc(&X->spec, (size_t)g.nxh * g.sy * g.nz);
  if (g.walls) dalloc(&X->thp, (size_t)(g.ny + 2) * g.nxh * g.nz);
  dalloc(&X->phi, n); dalloc(&X->b, n); dalloc(&X->r32, n);
  if (c.p.poisson_variant == FP_POISSON_SOR) {
    dalloc(&X->e32, n); dalloc(&X->sor_pb, (size_t)nblk(n, 256)); dalloc(&X->sor_flag, 4);
    FP_HIP_CHECK(hipHostMalloc((void**)&X->sor_host, 4 * sizeof(int), hipHostMallocDefault));
  }
  dalloc(&X->red, 8);
  dalloc(&X->part, (size_t)(n / 256 + 1024));
  dalloc(&X->partb, (size_t)blocks_x32(g) + 1);
  if (g.walls) {
    // the slab solve holds one kx slab in LDS: use it where the device's
    // per-block LDS takes it
    int lds_max = 0, dev = 0;
    FP_HIP_CHECK(hipGetDevice(&dev));
    FP_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (lds_slab(g) <= (size_t)lds_max && g.nz <= 1024) {
      FP_HIP_CHECK(hipFuncSetAttribute((const void*)m_poisson_slab, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_slab(g)));
      X->slab = 1;
      dalloc(&X->pbuf, (size_t)g.nxh * g.sy * g.nz);
    }
  }
  FP_HIP_CHECK(hipHostMalloc((void**)&X->dc_host, sizeof(int), hipHostMallocDefault));
  *X->dc_host = 1;
  FP_HIP_CHECK(hipEventCreateWithFlags(&X->dc_ev, hipEventDisableTiming));
  dalloc(&X->skip, 1);
  dalloc(&X->X2, 3 * n);
  X->helm_iters = helm_sweeps(g, c.p.helm_sweeps);
  // every (cell, axis) may need the FP64 decision: a region of 3 * 256
  // entries per select block (3 n for n >= 256; m * 3 + axis < 2^32 as every
  // axis has at most 1024 cells)
  {
    const long nb = nblk(n, SEL_BLOCK);
    dalloc(&X->amb, 1); dalloc(&X->amb_list, nb * SEL_REGION); dalloc(&X->chk, 1);
    dalloc(&X->amb_idx, nb * SEL_REGION);
    dalloc(&X->amb_data, nb * 18 * SEL_STAGE);
  }
  // FP32 twiddles and wavenumbers from the FP64 host tables
  auto tw32 = [&](const double2* d, int nn, float2** out) {
    std::vector<double2> h(nn / 2);
    FP_HIP_CHECK(hipMemcpy(h.data(), d, sizeof(double2) * h.size(), hipMemcpyDeviceToHost));
    std::vector<float2> f(h.size());
    for (size_t a = 0; a < h.size(); a++) f[a] = make_float2((float)h[a].x, (float)h[a].y);
    dalloc(out, f.size());
    FP_HIP_CHECK(hipMemcpy(*out, f.data(), sizeof(float2) * f.size(), hipMemcpyHostToDevice));
  };
  auto k32 = [&](const double* d, int nn, float** out) {
    std::vector<double> h(nn);
    FP_HIP_CHECK(hipMemcpy(h.data(), d, sizeof(double) * nn, hipMemcpyDeviceToHost));
    std::vector<float> f(nn);
    for (int a = 0; a < nn; a++) f[a] = (float)h[a];
    dalloc(out, nn);
    FP_HIP_CHECK(hipMemcpy(*out, f.data(), sizeof(float) * nn, hipMemcpyHostToDevice));
  };
  tw32(c.twx, g.nx, &X->twx); tw32(c.twy, g.ny, &X->twy); tw32(c.twz, g.nz, &X->twz);
  k32(c.kx2, g.nxh, &X->kx2); k32(c.ky2, g.ny, &X->ky2); k32(c.kz2, g.nz, &X->kz2);
}

void mixed_destroy(Ctx& c) {
  Mixed* X = (Mixed*)c.mixed;
  if (!X) return;
  if (X->side) (void)hipStreamSynchronize(X->side);
  if (X->ev_cstar) (void)hipEventDestroy(X->ev_cstar);
  if (X->sel_done) (void)hipEventDestroy(X->sel_done);
  if (X->side) (void)hipStreamDestroy(X->side);
  dfree(X->U); dfree(X->V); dfree(X->W); dfree(X->P); dfree(X->C[0]); dfree(X->C[1]); dfree(X->f);
  dfree(X->A0); dfree(X->T0); dfree(X->B); dfree(X->Q); dfree(X->Up); dfree(X->F);
  dfree(X->spec); dfree(X->thp); dfree(X->twx); dfree(X->twy); dfree(X->twz);
  dfree(X->kx2); dfree(X->ky2); dfree(X->kz2);
  dfree(X->phi); dfree(X->b); dfree(X->r32); dfree(X->red); dfree(X->part); dfree(X->skip);
  dfree(X->e32); dfree(X->sor_pb); dfree(X->sor_flag);
  if (X->sor_host) (void)hipHostFree(X->sor_host);
  dfree(X->partb);
  dfree(X->pbuf);
  if (X->dc_ev) (void)hipEventDestroy(X->dc_ev);
  if (X->dc_host) (void)hipHostFree(X->dc_host);
  dfree(X->X2); dfree(X->amb); dfree(X->amb_list); dfree(X->chk);
  dfree(X->amb_idx); dfree(X->amb_data);
  dfree(X->wi); dfree(X->org); dfree(X->uc); dfree(X->mf); dfree(X->mf64); dfree(X->ftmp);
  for (int k = 0; k < 3; k++) dfree(X->sw[k]);
  dfree(X->fr_row); dfree(X->fr_len); dfree(X->fr_tgt); dfree(X->fr_ent); dfree(X->fr_w);
  dfree(X->zi_w); dfree(X->zi_org);
  dfree(X->zs_off); dfree(X->zs_len); dfree(X->zs_tgt); dfree(X->zs_ent); dfree(X->zs_w);
  delete X;
  c.mixed = nullptr;
}

void mixed_load_from_f64(Ctx& c) {
  Mixed& X = MX(c);
  const long n = c.g.n;
  hipStream_t s = c.stream;
  // a selection made ahead still reads the old C; it no longer applies
  if (X.sel_ready) FP_HIP_CHECK(hipStreamWaitEvent(s, X.sel_done, 0));
  X.sel_ready = 0;
  hipLaunchKernelGGL(m_to32, dim3(nblk(n, 256)), dim3(256), 0, s, n, c.U, X.U);
  hipLaunchKernelGGL(m_to32, dim3(nblk(n, 256)), dim3(256), 0, s, n, c.V, X.V);
  hipLaunchKernelGGL(m_to32, dim3(nblk(n, 256)), dim3(256), 0, s, n, c.W, X.W);
  hipLaunchKernelGGL(m_to32, dim3(nblk(n, 256)), dim3(256), 0, s, n, c.P, X.P);
  X.cur = 0;
  hipLaunchKernelGGL(m_to32, dim3(nblk(6 * n, 256)), dim3(256), 0, s, 6 * n, c.C[c.cur], X.C[0]);
  hipLaunchKernelGGL(m_to32, dim3(nblk(3 * n, 256)), dim3(256), 0, s, 3 * n, c.A0, X.A0);
  hipLaunchKernelGGL(m_to32, dim3(nblk(3 * n, 256)), dim3(256), 0, s, 3 * n, c.T0, X.T0);
  // the potential: the SOR variant starts its sweeps from it
  FP_HIP_CHECK(hipMemcpyAsync(X.phi, c.Phi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  FP_HIP_CHECK(hipStreamSynchronize(s));
}

void mixed_phi_to_f64(Ctx& c) {
  FP_HIP_CHECK(hipMemcpyAsync(c.Phi, MX(c).phi, sizeof(double) * c.g.n, hipMemcpyDeviceToDevice, c.stream));
}

void mixed_store_to_f64(Ctx& c) {
  Mixed& X = MX(c);
  const long n = c.g.n;
  hipStream_t s = c.stream;
  hipLaunchKernelGGL(m_to64, dim3(nblk(n, 256)), dim3(256), 0, s, n, X.U, c.U);
  hipLaunchKernelGGL(m_to64, dim3(nblk(n, 256)), dim3(256), 0, s, n, X.V, c.V);
  hipLaunchKernelGGL(m_to64, dim3(nblk(n, 256)), dim3(256), 0, s, n, X.W, c.W);
  hipLaunchKernelGGL(m_to64, dim3(nblk(n, 256)), dim3(256), 0, s, n, X.P, c.P);
  hipLaunchKernelGGL(m_to64, dim3(nblk(6 * n, 256)), dim3(256), 0, s, 6 * n, X.C[X.cur], c.C[c.cur]);
  hipLaunchKernelGGL(m_to64, dim3(nblk(3 * n, 256)), dim3(256), 0, s, 3 * n, X.A0, c.A0);
  hipLaunchKernelGGL(m_to64, dim3(nblk(3 * n, 256)), dim3(256), 0, s, 3 * n, X.T0, c.T0);
  FP_HIP_CHECK(hipMemcpyAsync(c.Phi, X.phi, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  if (c.nmk) hipLaunchKernelGGL(m_to64, dim3(nblk(3 * c.nmk, 256)), dim3(256), 0, s, 3 * c.nmk, X.mf, X.mf64);
  FP_HIP_CHECK(hipStreamSynchronize(s));
}

void mixed_set_force(Ctx& c) {
  Mixed& X = MX(c);
  dfree(X.F);
  if (!c.F) return;
  dalloc(&X.F, 3 * c.g.n);
  hipLaunchKernelGGL(m_to32, dim3(nblk(3 * c.g.n, 256)), dim3(256), 0, c.stream, 3 * c.g.n, c.F, X.F);
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
}

void mixed_set_markers(Ctx& c) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  dfree(X.wi); dfree(X.org); dfree(X.uc); dfree(X.mf); dfree(X.mf64); dfree(X.ftmp);
  dalloc(&X.wi, (size_t)c.nmk * 81); dalloc(&X.org, (size_t)c.nmk * 3); dalloc(&X.uc, 3 * c.nmk);
  dalloc(&X.mf, 3 * c.nmk); dalloc(&X.mf64, 3 * c.nmk);
  hipLaunchKernelGGL(m_ibm_setup_interp, dim3(nblk(c.nmk, 128)), dim3(128), 0, c.stream, g, c.nmk, c.nc, c.nl, c.mk, X.wi,
                     X.org, X.uc);
  for (int k = 0; k < 3; k++) {
    dfree(X.sw[k]);
    std::vector<int> rowp(c.sp_rows[k] + 1);
    FP_HIP_CHECK(hipMemcpy(rowp.data(), c.sp_row[k], sizeof(int) * rowp.size(), hipMemcpyDeviceToHost));
    dalloc(&X.sw[k], std::max(1, rowp.back()));
    if (c.sp_rows[k])
      hipLaunchKernelGGL(m_ibm_setup_spread, dim3(nblk(c.sp_rows[k], 128)), dim3(128), 0, c.stream, g, k, c.sp_rows[k],
                         c.sp_row[k], c.sp_tgt[k], c.sp_ent[k], c.nmk, c.mk, X.sw[k]);
  }
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
  // fold the ghosted targets onto their periodic interior images
  const long gx = g.nx + 2, gy = g.ny + 2;
  std::map<long, std::vector<std::pair<int, float>>> rows;
  for (int k = 0; k < 3; k++) {
    const int nr = c.sp_rows[k];
    if (!nr) continue;
    std::vector<int> rowp(nr + 1);
    std::vector<long> tgt(nr);
    FP_HIP_CHECK(hipMemcpy(rowp.data(), c.sp_row[k], sizeof(int) * rowp.size(), hipMemcpyDeviceToHost));
    FP_HIP_CHECK(hipMemcpy(tgt.data(), c.sp_tgt[k], sizeof(long) * nr, hipMemcpyDeviceToHost));
    std::vector<int> ent(std::max(1, rowp.back()));
    std::vector<float> w(std::max(1, rowp.back()));
    if (rowp.back()) {
      FP_HIP_CHECK(hipMemcpy(ent.data(), c.sp_ent[k], sizeof(int) * rowp.back(), hipMemcpyDeviceToHost));
      FP_HIP_CHECK(hipMemcpy(w.data(), X.sw[k], sizeof(float) * rowp.back(), hipMemcpyDeviceToHost));
    }
    for (int r = 0; r < nr; r++) {
      if (rowp[r] == rowp[r + 1]) continue;
      long t = tgt[r];
      int ii = (int)(t % gx) - 1, jj = (int)((t / gx) % gy) - 1, kk = (int)(t / (gx * gy)) - 1;
      ii = (ii + g.nx) % g.nx; jj = (jj + g.ny) % g.ny; kk = (kk + g.nz) % g.nz;
      long lin = (long)k * g.n + ii + (long)g.nx * (jj + (long)g.ny * kk);
      auto& dst = rows[lin];
      for (int e = rowp[r]; e < rowp[r + 1]; e++) dst.push_back({(int)(k * c.nmk + ent[e]), w[e]});
    }
  }
  dfree(X.fr_row); dfree(X.fr_len); dfree(X.fr_tgt); dfree(X.fr_ent); dfree(X.fr_w);
  dfree(X.zi_w); dfree(X.zi_org);
  dfree(X.zs_off); dfree(X.zs_len); dfree(X.zs_tgt); dfree(X.zs_ent); dfree(X.zs_w);
  X.fr_rows = 0;
  X.zinv = 0;
  X.zs_rows = X.zs_rows64 = 0;
  // CSR -> sliced ELL (64-row slices, slot-major, padded with zero weights and
  // the entry `pad`), uploaded to the given buffers
  auto upload_ell = [&](const std::vector<int>& frow, const std::vector<int>& fent, const std::vector<float>& fw,
                        const std::vector<long>& ftgt, int pad, int*& d_off, int*& d_len, long*& d_tgt, int*& d_ent,
                        float*& d_w) {
    const int nrows = (int)ftgt.size();
    const int nsl = (nrows + 63) / 64;
    std::vector<int> soff(std::max(1, nsl)), slen(std::max(1, nsl));
    long total = 0;
    for (int sl = 0; sl < nsl; sl++) {
      int mx = 0;
      for (int r = sl * 64; r < std::min(nrows, sl * 64 + 64); r++) mx = std::max(mx, frow[r + 1] - frow[r]);
      mx = (mx + ELL_UNROLL - 1) / ELL_UNROLL * ELL_UNROLL;
      soff[sl] = (int)total;
      slen[sl] = mx;
      total += 64L * mx;
    }
    std::vector<int> eent(std::max<long>(1, total), pad);
    std::vector<float> ew(std::max<long>(1, total), 0.0f);
    for (int r = 0; r < nrows; r++) {
      const int sl = r >> 6, lane = r & 63;
      for (int e = frow[r]; e < frow[r + 1]; e++) {
        const long pos = soff[sl] + 64L * (e - frow[r]) + lane;
        eent[pos] = fent[e];
        ew[pos] = fw[e];
      }
    }
    dalloc(&d_off, soff.size()); dalloc(&d_len, slen.size()); dalloc(&d_tgt, std::max<size_t>(1, ftgt.size()));
    dalloc(&d_ent, eent.size()); dalloc(&d_w, ew.size());
    FP_HIP_CHECK(hipMemcpy(d_off, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice));
    FP_HIP_CHECK(hipMemcpy(d_len, slen.data(), sizeof(int) * slen.size(), hipMemcpyHostToDevice));
    FP_HIP_CHECK(hipMemcpy(d_ent, eent.data(), sizeof(int) * eent.size(), hipMemcpyHostToDevice));
    FP_HIP_CHECK(hipMemcpy(d_w, ew.data(), sizeof(float) * ew.size(), hipMemcpyHostToDevice));
    if (!ftgt.empty())
      FP_HIP_CHECK(hipMemcpy(d_tgt, ftgt.data(), sizeof(long) * ftgt.size(), hipMemcpyHostToDevice));
  };

  // Markers repeated along z (ibm_init, fenep.f90:1826-1895, places every
  // layer at the same x, y with Zc = (k - 0.25) dZ): checked on the tables
  // themselves, every marker's interpolation origin and weights must be those
  // of the layer-0 marker moved up by its layer, and every spreading row of
  // plane k that of plane 0 with each marker moved up k layers (weights equal
  // to within 2^-20, the rounding of the FP64 weights to FP32).
  const long nmk = c.nmk, plane = (long)g.nx * g.ny;
  const long P = nmk / g.nz;
  bool zinv = g.nz >= 3 && P > 0 && P * g.nz == nmk;
  auto close = [](float a, float b) { return a == b || fabsf(a - b) <= 9.5367431640625e-07f * fmaxf(fabsf(a), fabsf(b)); };
  std::vector<float> wz;
  std::vector<int2> oz;
  if (zinv) {
    std::vector<float> wih((size_t)nmk * 81);
    std::vector<int4> orgh((size_t)nmk * 3);
    FP_HIP_CHECK(hipMemcpy(wih.data(), X.wi, sizeof(float) * wih.size(), hipMemcpyDeviceToHost));
    FP_HIP_CHECK(hipMemcpy(orgh.data(), X.org, sizeof(int4) * orgh.size(), hipMemcpyDeviceToHost));
    for (int cc = 0; cc < 3 && zinv; cc++)
      for (long id = 0; id < nmk && zinv; id++) {
        const long L = id / P, pm = id - L * P;
        const int4 a = orgh[cc * nmk + id], b = orgh[cc * nmk + pm];
        if (a.x != b.x || a.y != b.y || (long)a.z != (long)b.z + L) zinv = false;
        for (int s2 = 0; s2 < 27 && zinv; s2++)
          if (!close(wih[(size_t)(cc * 27 + s2) * nmk + id], wih[(size_t)(cc * 27 + s2) * nmk + pm])) zinv = false;
      }
    if (zinv) {
      wz.resize((size_t)81 * P);
      oz.resize((size_t)3 * P);
      for (int cs = 0; cs < 81; cs++)
        for (long pm = 0; pm < P; pm++) wz[(size_t)cs * P + pm] = wih[(size_t)cs * nmk + pm];
      for (int cc = 0; cc < 3; cc++)
        for (long pm = 0; pm < P; pm++) oz[cc * P + pm] = make_int2(orgh[cc * nmk + pm].x, orgh[cc * nmk + pm].y);
    }
  }
  // a row of plane k in layer-0 form: entries 4 (c nmk + p) + d + 1 for the
  // marker p of layer k + d, d in {-1, 0, 1}, sorted
  using Row = std::vector<std::pair<int, float>>;
  auto plane_form = [&](long k, const Row& src, Row& out) -> bool {
    out.clear();
    for (auto& pr : src) {
      const long cc = pr.first / nmk, id = pr.first - cc * nmk;
      const long L = id / P, pm = id - L * P;
      const long d = ((L - k) % g.nz + g.nz) % g.nz;
      const int dl = d == 0 ? 0 : (d == 1 ? 1 : (d == g.nz - 1 ? -1 : 2));
      if (dl == 2) return false;
      out.push_back({(int)(4 * (cc * nmk + pm) + dl + 1), pr.second});
    }
    std::sort(out.begin(), out.end());
    return true;
  };
  std::vector<Row> prow;
  std::vector<long> ptgt;
  if (zinv) {
    std::unordered_map<long, int> key0;   // c * plane + in-plane index -> plane-0 row
    Row tmp;
    for (auto& kv : rows) {
      const long cc = kv.first / g.n, rem = kv.first - cc * g.n;
      if (rem / plane != 0) continue;
      if (!plane_form(0, kv.second, tmp)) { zinv = false; break; }
      key0[cc * plane + rem] = (int)prow.size();
      prow.push_back(tmp);
      ptgt.push_back(kv.first);
    }
    std::vector<long> count(g.nz, 0);
    for (auto& kv : rows) {
      if (!zinv) break;
      const long cc = kv.first / g.n, rem = kv.first - cc * g.n, k = rem / plane;
      const auto it = key0.find(cc * plane + (rem - k * plane));
      if (it == key0.end() || !plane_form(k, kv.second, tmp)) { zinv = false; break; }
      const Row& ref = prow[it->second];
      if (ref.size() != tmp.size()) { zinv = false; break; }
      for (size_t e = 0; e < tmp.size() && zinv; e++)
        if (tmp[e].first != ref[e].first || !close(tmp[e].second, ref[e].second)) zinv = false;
      count[k]++;
    }
    for (long k = 0; k < g.nz && zinv; k++)
      if (count[k] != (long)prow.size()) zinv = false;
  }
  if (zinv) {
    std::vector<int> frow(1, 0), fent;
    std::vector<float> fw;
    for (auto& r : prow) {
      for (auto& pr : r) { fent.push_back(pr.first); fw.push_back(pr.second); }
      frow.push_back((int)fent.size());
    }
    // padding: marker 0 of component 0 at offset 0, weight 0
    upload_ell(frow, fent, fw, ptgt, 1, X.zs_off, X.zs_len, X.zs_tgt, X.zs_ent, X.zs_w);
    dalloc(&X.zi_w, wz.size()); dalloc(&X.zi_org, oz.size());
    FP_HIP_CHECK(hipMemcpy(X.zi_w, wz.data(), sizeof(float) * wz.size(), hipMemcpyHostToDevice));
    FP_HIP_CHECK(hipMemcpy(X.zi_org, oz.data(), sizeof(int2) * oz.size(), hipMemcpyHostToDevice));
    X.zinv = 1;
    X.zP = P;
    X.zs_rows = (int)prow.size();
    X.zs_rows64 = (X.zs_rows + 63) / 64 * 64;
    return;
  }
  std::vector<int> frow(1, 0), fent;
  std::vector<long> ftgt;
  std::vector<float> fw;
  for (auto& kv : rows) {
    ftgt.push_back(kv.first);
    for (auto& pr : kv.second) { fent.push_back(pr.first); fw.push_back(pr.second); }
    frow.push_back((int)fent.size());
  }
  upload_ell(frow, fent, fw, ftgt, 0, X.fr_row, X.fr_len, X.fr_tgt, X.fr_ent, X.fr_w);
  X.fr_rows = (int)ftgt.size();
}

const double* mixed_marker_forces(Ctx& c) {
  Mixed& X = MX(c);
  hipLaunchKernelGGL(m_to64, dim3(nblk(3 * c.nmk, 256)), dim3(256), 0, c.stream, 3 * c.nmk, X.mf, X.mf64);
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
  return X.mf64;
}

// defect-correction sweeps the last step ran: the later sweeps are skipped on
// the device once the FP64 residual is within dc_tol
double mixed_last_ndc(Ctx& c) {
  const int ndc = std::max(1, c.p.n_dc_max);
  if (ndc == 1) return 1.0;
  int skip = 0;
  FP_HIP_CHECK(hipMemcpy(&skip, MX(c).skip, sizeof(int), hipMemcpyDeviceToHost));
  return skip ? 1.0 : (double)ndc;
}

// C+- selection of the current stored field: certified FP32 decisions, the
// per-block list counts scanned, the rest re-decided in FP64.  tm != 0 times
// the two phases.
static void select_fenep(Ctx& c, int tm, hipStream_t s);

// Selection of the current stored field as the step makes it (certified FP32
// decisions, FP64 re-decision of the rest), compared with the reference's
// selection of every (cell, axis).  out[0] = re-decided in FP64, out[1] =
// differing decisions, out[2] = decisions (3 n).  Leaves the selection bytes
// as the next step recomputes them anyway.
int mixed_check_selection(Ctx& c, double* out3) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  out3[0] = out3[1] = 0.0;
  out3[2] = 3.0 * (double)g.n;
  if (c.p.beta == 1.0) return 0;   // no polymer stress: no selection
  if (X.sel_ready) FP_HIP_CHECK(hipStreamWaitEvent(s, X.sel_done, 0));
  zero_words(X.chk, 1, s);
  select_fenep(c, 0, s);
  X.sel_ready = 1;
  hipLaunchKernelGGL(m_select_check, dim3(nblk(g.n, 256)), dim3(256), 0, s, g, X.C[X.cur], c.sel, X.chk);
  FP_HIP_CHECK(hipGetLastError());
  unsigned h[2] = {0, 0};
  FP_HIP_CHECK(hipMemcpyAsync(&h[0], X.amb, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FP_HIP_CHECK(hipMemcpyAsync(&h[1], X.chk, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  FP_HIP_CHECK(hipStreamSynchronize(s));
  out3[0] = (double)h[0];
  out3[1] = (double)h[1];
  return 0;
}

// Tape entry of one mixed step for the adjoint's reverse sweep (FP64 words
// U V W P C6 u*(3) Phi, as the FP64 step's): before the step, the stored
// state widened and the selection the step will take (made now, or the one
// made ahead); after it, u* = u^n + d from the step's FP32 increment d and
// the FP64 potential.
__global__ __launch_bounds__(256) void m_ustar64(long n3, const double* __restrict__ un, const float* __restrict__ d,
                                                 double* __restrict__ us) {
  const long e = blockIdx.x * 256L + threadIdx.x;
  if (e < n3) us[e] = un[e] + (double)d[e];
}
void mixed_tape_in(Ctx& c, double* E, uint8_t* sel_out) {
  Mixed& X = MX(c);
  const long n = c.g.n;
  hipStream_t s = c.stream;
  if (sel_out) {
    if (X.sel_ready) FP_HIP_CHECK(hipStreamWaitEvent(s, X.sel_done, 0));
    else select_fenep(c, 0, s);
    X.sel_ready = 1;
    FP_HIP_CHECK(hipMemcpyAsync(sel_out, c.sel, (size_t)n, hipMemcpyDeviceToDevice, s));
  }
  hipLaunchKernelGGL(m_to64, dim3(nblk(n, 256)), dim3(256), 0, s, n, X.U, E);
  hipLaunchKernelGGL(m_to64, dim3(nblk(n, 256)), dim3(256), 0, s, n, X.V, E + n);
  hipLaunchKernelGGL(m_to64, dim3(nblk(n, 256)), dim3(256), 0, s, n, X.W, E + 2 * n);
  hipLaunchKernelGGL(m_to64, dim3(nblk(n, 256)), dim3(256), 0, s, n, X.P, E + 3 * n);
  hipLaunchKernelGGL(m_to64, dim3(nblk(6 * n, 256)), dim3(256), 0, s, 6 * n, X.C[X.cur], E + 4 * n);
}
void mixed_tape_out(Ctx& c, double* E) {
  Mixed& X = MX(c);
  const long n = c.g.n;
  hipLaunchKernelGGL(m_ustar64, dim3(nblk(3 * n, 256)), dim3(256), 0, c.stream, 3 * n, E, X.Up, E + 10 * n);
  FP_HIP_CHECK(hipMemcpyAsync(E + 13 * n, X.phi, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
}
void mixed_widen_c(Ctx& c, double* dst) {
  Mixed& X = MX(c);
  const long n = c.g.n;
  hipLaunchKernelGGL(m_to64, dim3(nblk(6 * n, 256)), dim3(256), 0, c.stream, 6 * n, X.C[X.cur], dst);
}

// the selection of the current stored field (as mixed_check_selection makes
// it), copied to host memory
int mixed_get_selection(Ctx& c, uint8_t* out) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  if (c.p.beta == 1.0) {
    std::memset(out, 0, (size_t)g.n);
    return 0;
  }
  if (X.sel_ready) FP_HIP_CHECK(hipStreamWaitEvent(s, X.sel_done, 0));
  select_fenep(c, 0, s);
  X.sel_ready = 1;
  FP_HIP_CHECK(hipMemcpyAsync(out, c.sel, (size_t)g.n, hipMemcpyDeviceToHost, s));
  FP_HIP_CHECK(hipStreamSynchronize(s));
  return 0;
}

// statistics of the last step: [0] |b|^2, [1] |r|^2 before the last
// correction, [2] guarded selections re-decided in FP64
// The context's stream waits for work made ahead on the side stream, so that
// synchronizing the context's stream covers it.
void mixed_physics(Ctx& c) {
  mixed_join(c);
  MX(c).helm_iters = helm_sweeps(c.g, c.p.helm_sweeps);
}

void mixed_velocity(Ctx& c, const float** u, const float** v, const float** w) {
  Mixed& X = MX(c);
  *u = X.U; *v = X.V; *w = X.W;
}

void mixed_join(Ctx& c) {
  Mixed& X = MX(c);
  if (X.sel_ready) FP_HIP_CHECK(hipStreamWaitEvent(c.stream, X.sel_done, 0));
}

int mixed_stats(Ctx& c, double* out3) {
  Mixed& X = MX(c);
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
  double h[8];
  FP_HIP_CHECK(hipMemcpy(h, X.red, sizeof(h), hipMemcpyDeviceToHost));
  unsigned amb = 0;
  FP_HIP_CHECK(hipMemcpy(&amb, X.amb, sizeof(unsigned), hipMemcpyDeviceToHost));
  out3[0] = h[0]; out3[1] = h[1]; out3[2] = (double)amb;
  return 0;
}

namespace {
struct MTimer {
  Ctx& c;
  const char* name;   // nullptr: not timed
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  MTimer(Ctx& cc, const char* nm, hipStream_t s = nullptr) : c(cc), name(nm), st(s ? s : cc.stream) {
    if (!phase_timed(c, name)) return;
    a = ev_get(c); b = ev_get(c);
    (void)hipEventRecord(a, st);
  }
  ~MTimer() {
    if (!phase_timed(c, name)) return;
    (void)hipEventRecord(b, st);
    c.ev.push_back({std::string(name) + "#a", a});
    c.ev.push_back({std::string(name) + "#b", b});
  }
};
}  // namespace

#define DC_EVERY 8

// The FP64 defect check (residual, |r| <= dc_tol |b|, and the correction
// sweep it gates on the device) runs every DC_EVERY steps.  Its decision is
// copied to pinned host memory; at the next periodic check the host reads it
// (the copy finished DC_EVERY steps earlier) and, if a correction was needed,
// checks every step until a periodic check passes again.
static bool dc_check_due(Ctx& c) {
  Mixed& X = MX(c);
  if (c.step % DC_EVERY == 0) {
    if (X.dc_pending) {
      FP_HIP_CHECK(hipEventSynchronize(X.dc_ev));
      X.dc_correct = *X.dc_host == 0;
      X.dc_pending = false;
    }
    return true;
  }
  return X.dc_correct;
}
static void dc_record(Ctx& c) {
  Mixed& X = MX(c);
  FP_HIP_CHECK(hipMemcpyAsync(X.dc_host, X.skip, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  FP_HIP_CHECK(hipEventRecord(X.dc_ev, c.stream));
  X.dc_pending = true;
}

// the z-streamed selection's shapes: whole x-rows in a block of 256 threads
static bool select_streamed(const Grid& g) {
  return (g.nx == 64 || g.nx == 128 || g.nx == 256) && g.nz % SZ_K == 0;
}

static void select_fenep(Ctx& c, int tm, hipStream_t s) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  const long nb = nblk(g.n, SEL_BLOCK);
  // select blocks, and each block's region and staging entries (the same
  // totals either way: 3 n entries, 18 n / 4 staged values)
  const bool zs = select_streamed(g);
  const long nbs = zs ? g.n / (SEL_BLOCK * SZ_K) : nb;
  const unsigned region = zs ? SEL_REGION * SZ_K : SEL_REGION, cap = zs ? SEL_STAGE * SZ_K : SEL_STAGE;
  zero_words(X.amb, 1, s);
  {
    MTimer t(c, tm ? "fenep_select" : nullptr, s);
    if (zs)
      hipLaunchKernelGGL(m_select_z, dim3(nbs), dim3(SEL_BLOCK), 0, s, g, X.C[X.cur], c.sel, X.amb, X.amb_idx,
                         X.amb_list, X.amb_data);
    else
      hipLaunchKernelGGL(m_select, dim3(nb), dim3(SEL_BLOCK), 0, s, g, X.C[X.cur], c.sel, X.amb, X.amb_idx,
                         X.amb_list, X.amb_data);
  }
  {
    MTimer t(c, tm ? "fenep_select_fp64_fix" : nullptr, s);
    // n / 4 threads, one wavefront per block: 16 per SIMD at 256x128x128,
    // enough to hide the FP64 latencies
    hipLaunchKernelGGL(g.walls ? m_select_fix<true> : m_select_fix<false>, dim3(nb), dim3(64), 0, s, g, X.C[X.cur], c.sel,
                       X.amb, X.amb_idx, X.amb_list, X.amb_data, region, cap);
  }
}

// Start the selection of the C this step produced on the side stream: it
// waits for everything enqueued on the context's stream so far and runs
// beside what follows (its VALU work beside the memory-bound kernels and the
// Poisson slab solve, which leaves half of the CUs free).
static void select_ahead(Ctx& c) {
  Mixed& X = MX(c);
  FP_HIP_CHECK(hipEventRecord(X.ev_cstar, c.stream));
  FP_HIP_CHECK(hipStreamWaitEvent(X.side, X.ev_cstar, 0));
  select_fenep(c, 1, X.side);
  FP_HIP_CHECK(hipEventRecord(X.sel_done, X.side));
  X.sel_ready = 1;
}

// The SOR variant's solve of L_h phi = div u*/dt (m_sor_* above).  Every
// SOR_BATCH sweeps the correction is moved into phi (FP64) and the defect
// taken again in FP64, so that the FP32 sweeps only ever carry the change of
// the last few sweeps: a smooth correction that grew over many sweeps would
// make L_h e cancel in FP32 down to the size of E and stall the stopping test.
// (SOR is affine in its iterate: the restarted sweeps are the same iteration.)
// The host reads the flags back after each batch, as the FP64 path reads its
// residuals; sweeps launched after convergence return at once.
static void mixed_sor(Ctx& c) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  const long n = g.n;
  const int nb = nblk(n, 256), nbh = nblk(n / 2, 256);
  const double om = c.p.sor_omega > 0.0 ? c.p.sor_omega : 2.0 / (1.0 + sin(3.14159265358979323846 * g.dx));
  const double BX = 1.0 / (g.dx * g.dx), BY = 1.0 / (g.dy * g.dy), BZ = 1.0 / (g.dz * g.dz);
  const double B0 = ((((BX + BX) + BY) + BY) + BZ) + BZ;
  const float om_b0 = (float)(om / B0);
  const int SOR_BATCH = 10;
  hipLaunchKernelGGL(g.walls ? m_sor_start<true> : m_sor_start<false>, dim3(nb), dim3(256), 0, s, g, X.U, X.V, X.W, X.Up, X.phi, X.b,
                     X.r32, X.e32, X.sor_pb);
  hipLaunchKernelGGL(m_sor_decide, dim3(1), dim3(1024), 0, s, X.sor_pb, nb, X.red, c.p.sor_eps, 0, X.sor_flag);
  int launched = 0;
  while (launched < c.p.sor_itrmax) {
    const int batch = std::min(SOR_BATCH, c.p.sor_itrmax - launched);
    for (int b = 0; b < batch; b++) {
      for (int color = 0; color < 2; color++)
        hipLaunchKernelGGL(g.walls ? m_sor_sweep<true> : m_sor_sweep<false>, dim3(nbh), dim3(256), 0, s, g, color, om_b0,
                           (float)BX, (float)BY, (float)BZ, (float)B0, X.r32, X.e32, X.part, X.sor_flag);
      hipLaunchKernelGGL(m_sor_decide, dim3(1), dim3(1024), 0, s, X.part, nbh, X.red, c.p.sor_eps, 1, X.sor_flag);
    }
    launched += batch;
    // phi += e, e = 0, the FP64 defect of the new phi; once the sweeps have
    // met the test, the same test on the FP64 residual (else they go on)
    hipLaunchKernelGGL(m_sor_accum, dim3(nb), dim3(256), 0, s, n, X.phi, X.e32);
    hipLaunchKernelGGL(g.walls ? m_residual<true> : m_residual<false>, dim3(nb), dim3(256), 0, s, g, X.b, X.phi, X.r32,
                       X.part);
    hipLaunchKernelGGL(m_sor_decide, dim3(1), dim3(1024), 0, s, X.part, nb, X.red, c.p.sor_eps, 2, X.sor_flag);
    FP_HIP_CHECK(hipMemcpyAsync(X.sor_host, X.sor_flag, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
    FP_HIP_CHECK(hipStreamSynchronize(s));
    if (X.sor_host[0]) break;
  }
  c.sor_itr = (double)X.sor_host[1];
}

void mixed_step(Ctx& c) {
  Mixed& X = MX(c);
  const Grid& g = c.g;
  hipStream_t s = c.stream;
  const long n = g.n;
  const int polymer = c.p.beta != 1.0;
  if (!polymer) {
    zero_words(X.C[X.cur], 6 * n, s);
  } else {
    // the selection of this C: made ahead beside the previous step, or now
    if (X.sel_ready) FP_HIP_CHECK(hipStreamWaitEvent(s, X.sel_done, 0));
    else select_fenep(c, 1, s);
    const int itmax = std::min(c.p.newton_itrmax, 16);
    {
      MTimer t(c, "fenep_cstar_newton");
      hipLaunchKernelGGL(m_cstar_newton, dim3(nblk(n, 256)), dim3(256), 0, s, g, X.C[X.cur], c.sel, X.U, X.V, X.W,
                         X.C[1 - X.cur], X.f, c.nitr, (float)c.p.newton_eps, itmax);
    }
    X.cur = 1 - X.cur;
  }
  if (g.nx % PZ_TX == 0 && g.ny % PZ_TY == 0) {
    // streamed along z, wall rows included
    MTimer t(c, "predictor");
    const long nbk = (g.nz + PZ_KC - 1) / PZ_KC;
    const long grid = (long)(g.nx / PZ_TX) * (g.ny / PZ_TY) * nbk;
    hipLaunchKernelGGL(g.walls ? m_predictor_z<true> : m_predictor_z<false>, dim3(grid), dim3(PZ_THREADS), 0, s, g, PZ_KC,
                       c.p.method, c.has_history ? 0 : 1, polymer, X.U, X.V, X.W, X.P, X.C[X.cur], X.f, X.A0, X.T0,
                       X.F, (float)c.p.body_force[0], (float)c.p.body_force[1], (float)c.p.body_force[2], X.Q);
  } else {
    MTimer t(c, "predictor");
    hipLaunchKernelGGL(m_predictor<false>, dim3(nblk(n, 256)), dim3(256), 0, s, g, c.p.method, c.has_history ? 0 : 1, polymer,
                       X.U, X.V, X.W, X.P, X.C[X.cur], X.f, X.A0, X.T0, X.F, (float)c.p.body_force[0],
                       (float)c.p.body_force[1], (float)c.p.body_force[2], X.Q);
  }
  if (g.walls && !(g.nx % PZ_TX == 0 && g.ny % PZ_TY == 0)) {
    MTimer t(c, "predictor_walls");
    hipLaunchKernelGGL(m_predictor<true>, dim3(nblk(wall_rows(g, 1), 256)), dim3(256), 0, s, g, c.p.method, c.has_history ? 0 : 1, polymer,
                       X.U, X.V, X.W, X.P, X.C[X.cur], X.f, X.A0, X.T0, X.F, (float)c.p.body_force[0],
                       (float)c.p.body_force[1], (float)c.p.body_force[2], X.Q);
  }
  // The next step's selection (it reads only the new C) forks onto the second
  // stream here, beside the Helmholtz sweeps and the Poisson solve.  Forking
  // before the IBM phase measured slower, forking after the sweeps leaves the
  // selection chain longer than what overlaps it.
  if (c.p.method == FP_METHOD_FFT) {
    if (polymer) select_ahead(c);
    MTimer t(c, "helmholtz");
    helmholtz32(c, X.Q, X.Up);
  } else {
    {
      MTimer t(c, "ibm");
      if (X.zinv) {
        if (g.nz % IZ_LC == 0)
          hipLaunchKernelGGL(g.walls ? m_ibm_interp_zc<true> : m_ibm_interp_zc<false>,
                             dim3(nblk(3 * (g.nz / IZ_LC) * X.zP, 256)), dim3(256), 0, s, g, c.nmk, X.zP, X.zi_w, X.zi_org,
                             X.uc, X.U, X.V, X.W, X.Q, X.mf);
        else
          hipLaunchKernelGGL(g.walls ? m_ibm_interp_z<true> : m_ibm_interp_z<false>, dim3(nblk(3 * c.nmk, 256)), dim3(256),
                             0, s, g, c.nmk, X.zP, X.zi_w, X.zi_org, X.uc, X.U, X.V, X.W, X.Q, X.mf);
        if (X.zs_rows)
          hipLaunchKernelGGL(m_ibm_spread_z, dim3(nblk((long)g.nz * X.zs_rows64, 256)), dim3(256), 0, s, X.zs_rows,
                             X.zs_rows64, g.nz, X.zP, (long)g.nx * g.ny, X.zs_off, X.zs_len, X.zs_tgt, X.zs_ent, X.zs_w,
                             X.mf, (float)g.dt, X.Q);
      } else {
        hipLaunchKernelGGL(g.walls ? m_ibm_interp<true> : m_ibm_interp<false>, dim3(nblk(3 * c.nmk, 256)), dim3(256), 0, s,
                           g, c.nmk, X.wi, X.org, X.uc, X.U, X.V, X.W, X.Q, X.mf);
        if (X.fr_rows)
          hipLaunchKernelGGL(m_ibm_spread_folded, dim3(nblk(X.fr_rows, 256)), dim3(256), 0, s, X.fr_rows, X.fr_row,
                             X.fr_len, X.fr_tgt, X.fr_ent, X.fr_w, X.mf, (float)g.dt, X.Q);
      }
    }
    if (polymer) select_ahead(c);
    MTimer t(c, "helmholtz");
    helmholtz32(c, X.Q, X.Up);
  }
  if (c.p.poisson_variant == FP_POISSON_SOR) {
    MTimer t(c, "poisson");
    mixed_sor(c);
  } else {
    MTimer t(c, "poisson");
    // first sweep: b = div u*/dt in the load stage of the x transform, the
    // FP32 spectral + line solve, Phi accumulated in FP64 in the store stage
    const int nb = nblk(n, 256), nbx = blocks_x32(g), lx = lines_for32(g.nx);
    const int ndc = std::max(1, c.p.n_dc_max);
    const bool check = ndc > 1 && dc_check_due(c);
    const int tr = X.slab;
    const size_t lds_inv = tr ? lds_x32_tr(g) : lds_x32(g);
    hipLaunchKernelGGL(g.walls ? m_div_fft_x_fwd<true> : m_div_fft_x_fwd<false>, dim3(nbx), dim3(256), lds_x32(g), s, g,
                       X.U, X.V, X.W, X.Up, X.spec, X.twx, lx, X.b, X.partb, check ? 1 : 0, tr);
    solve32_yz(c, 3, nullptr, tr);
    hipLaunchKernelGGL(m_fft_x_inv_accum, dim3(nbx), dim3(256), lds_inv, s, g, X.spec, X.phi, X.twx, lx, 1, nullptr, tr);
    if (check) {
      // FP64 defect r = b - L_h phi; the next sweep is skipped on the device
      // once |r| <= dc_tol |b|
      for (int it = 1; it < ndc; it++) {
        hipLaunchKernelGGL(g.walls ? m_residual<true> : m_residual<false>, dim3(nb), dim3(256), 0, s, g, X.b, X.phi, X.r32,
                           X.part);
        hipLaunchKernelGGL(m_sum_decide, dim3(1), dim3(1024), 0, s, X.partb, nbx, X.part, nb, X.red, c.p.dc_tol, X.skip);
        if (it == 1 && c.step % DC_EVERY == 0) dc_record(c);
        hipLaunchKernelGGL(m_fft_x_fwd, dim3(nbx), dim3(256), lds_x32(g), s, g, X.r32, X.spec, X.twx, lx, X.skip, tr);
        solve32_yz(c, 3, X.skip, tr);
        hipLaunchKernelGGL(m_fft_x_inv_accum, dim3(nbx), dim3(256), lds_inv, s, g, X.spec, X.phi, X.twx, lx, 0, X.skip, tr);
      }
    }
  }
  {
    MTimer t(c, "march");
    if (g.nx % MZ_TX == 0 && g.ny % MZ_TY == 0) {
      const long grid = (long)(g.nx / MZ_TX) * (g.ny / MZ_TY) * ((g.nz + MZ_KC - 1) / MZ_KC);
      hipLaunchKernelGGL(g.walls ? m_march_z<true> : m_march_z<false>, dim3(grid), dim3(MZ_THREADS), 0, s, g, X.Up,
                         X.phi, X.U, X.V, X.W, X.P);
    } else {
      hipLaunchKernelGGL(g.walls ? m_march<true> : m_march<false>, dim3(nblk(n, 256)), dim3(256), 0, s, g, X.Up, X.phi,
                         X.U, X.V, X.W, X.P);
    }
  }
  c.has_history = 1;
  c.step++;
}

}  // namespace fp
