// Internal definitions shared by the host C-ABI (fp_api.hip) and the kernel
// translation units.  Device layout: interior cells only, x fastest,
// idx = i + nx*(j + ny*k) with 0-based i,j,k (reference index i_ref = i+1);
// ghost values are produced on the fly by the boundary rules in `Grid::ld*`.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/fenep_gpu.h"

namespace fp {

enum BcKind { BC_U = 0, BC_V = 1, BC_W = 2, BC_P = 3, BC_C = 4 };

// Grid description passed by value to every kernel.
struct Grid {
  int nx, ny, nz, nxh;
  int lnx, lny;     // log2(nx), log2(ny): cell index -> (i, j, k) by shifts
  int walls;
  long n;           // nx*ny*nz
  double dx, dy, dz;
  double dt, re, beta, wi, lp;
  double sw;        // IBM spreading and drag weight per marker (dX^3 or ibm_weight)
  long sy;          // spectral rows in y (ny or ny+2)
  int jo;           // spectral row of interior j=0 (0 or 1)
};

// Device buffers of one context.
struct Ctx {
  fp_params p;
  Grid g;
  hipStream_t stream = nullptr;
  std::string err;
  int has_history = 0;
  int state_set = 0;
  int step = 0;
  int timing = 0;
  // plain FP64 state
  double *U = nullptr, *V = nullptr, *W = nullptr, *P = nullptr, *Phi = nullptr;
  double *C[2] = {nullptr, nullptr};  // 6 planes each
  int cur = 0;
  double *fC = nullptr;               // Peterlin factor of the current C
  uint8_t* sel = nullptr;
  double *A0 = nullptr, *T0 = nullptr, *B0 = nullptr, *B = nullptr;  // 3 planes each
  double *Q = nullptr;                // 3 planes: predictor RHS / Ua
  double *Up = nullptr;               // 3 planes: u*
  double *F = nullptr;                // 3 planes body-force field (nullable)
  double2* spec = nullptr;            // nxh*sy*nz complex
  double* thp = nullptr;              // Thomas P scratch (ny+2)*nxh*nz
  double2 *twx = nullptr, *twy = nullptr, *twz = nullptr;
  double *kx2 = nullptr, *ky2 = nullptr, *kz2 = nullptr;
  double* red = nullptr;              // reduction scratch
  double* part = nullptr;             // per-block partial sums
  int* nitr = nullptr;                // Newton iteration counts (n)
  // IBM
  int nc = 0, nl = 0;
  long nmk = 0;
  double* mk = nullptr;               // 6 planes: xc,yc,zc,uc,vc,wc
  double* mf = nullptr;               // 3 planes: Fxc,Fyc,Fzc
  double* mb = nullptr;               // 3 planes: Ub,Vb,Wb
  double* ftmp = nullptr;             // 3 ghosted raw spread arrays
  long gn = 0;
  int* sp_row[3] = {nullptr, nullptr, nullptr};   // CSR row pointers
  long* sp_tgt[3] = {nullptr, nullptr, nullptr};  // ghosted target index
  int* sp_ent[3] = {nullptr, nullptr, nullptr};   // marker of each entry
  int sp_rows[3] = {0, 0, 0};
  double* sp_w[3] = {nullptr, nullptr, nullptr};  // weight of each entry (ibm_weight)
  double* ib_w = nullptr;             // interpolation weights, (3 nmk) x 27 stencil points
  // SOR
  double sor_itr = -1.0;
  // timing: recorded (name, event) pairs and a pool of created events, so
  // that timed steps do not create or destroy events
  std::vector<std::pair<std::string, hipEvent_t>> ev;
  std::vector<hipEvent_t> evpool;
  std::vector<std::pair<std::string, double>> ktime;
  // mixed precision state (allocated only for precision == 1)
  void* mixed = nullptr;
  // selections to use instead of computing them (frozen-branch replays of
  // a recorded run); nullptr: computed by the step
  const uint8_t* sel_replay = nullptr;
  // recording and adjoint state (fp_adjoint.hip)
  void* adj = nullptr;
  // statistics: time-mean accumulator of U, V, W and the mean subtracted by
  // the statistics (3 planes each; nullptr: none)
  double* mean_acc = nullptr;
  double* mean_u = nullptr;
  long mean_count = 0;
};

// ---- launchers (fp_kernels_f64.hip)
void f64_select(Ctx& c);
void f64_cstar_newton(Ctx& c);
void f64_predictor(Ctx& c);
void f64_solve(Ctx& c, int kind, const double* q, double* x);   // kind 0..2 Helmholtz, 3 Poisson
void f64_divergence(Ctx& c, const double* up, double* q);
void f64_march(Ctx& c);
void f64_ibm(Ctx& c);        // interp + forces + spread + predictor RHS into Q
void f64_ibm_setup(Ctx& c);  // CSR build (host) after markers are set
int  f64_sor(Ctx& c, const double* q, double* phi);
void f64_zero_c(Ctx& c);
void f64_diag(Ctx& c, fp_diag* d, double* scratch, const double* mf);  // scratch: diag_scratch_words
size_t diag_scratch_words(const Grid& g);
void f64_time_stats(Ctx& c, double* out13);  // all_time_dif statistics
void f64_energy_spectrum(Ctx& c, double* E, int nbins);  // shell spectrum (periodic)
int f64_time_stats_map(Ctx& c, int ncyl, const double* xy, double diameter, double* map18, double* out13);
int f64_energy_spectrum_z(Ctx& c, int ncyl, const double* xy, double diameter, double* Emask, double* Efirst);
void f64_mean_accumulate(Ctx& c, const float* u32, const float* v32, const float* w32);
int f64_mean_apply(Ctx& c, int mode);
void f64_fill_f(Ctx& c);     // Peterlin factor of the current C
void f64_spread_markers(Ctx& c, const double* mfield, double* ftmp);  // spread + ghost folds
int thomas_lines(int ny, int bytes_per_row);  // lines per block of the LDS line solve
void step_f64(Ctx& c);       // one FP64 time step (fp_api.hip)

// ---- recording and discrete adjoint (fp_adjoint.hip)
void adj_record(Ctx& c, int nsteps, int every, int replay, double* drag_out);
void adj_backward(Ctx& c, const double* weights, const double* bar_final, double* bar_initial, double* bar_params,
                  double* bar_force);
void adj_free(Ctx& c);
void adj_drop_tape(Ctx& c);

// an event from the context's pool (created on first use)
// Phase timing: level 1 records every phase, level 2 only the predictor (the
// kernel whose roofline the benchmark reports), so that a timed run carries
// two events per step, level 3 the predictor of every 8th step (c.step a
// multiple of 8): each event idles the stream for a few microseconds.
inline bool phase_timed(const Ctx& c, const char* name) {
  if (!name) return false;
  if (c.timing == 1) return true;
  const bool pred = std::strcmp(name, "predictor") == 0;
  return pred && (c.timing == 2 || (c.timing == 3 && c.step % 8 == 0));
}
// Timing events skip the system-scope fence on record (its cache writeback
// idles the GPU for about 6 us per event); they are read only after the
// stream has been synchronized.
inline hipEvent_t ev_get(Ctx& c) {
  if (c.evpool.empty()) {
    hipEvent_t e = nullptr;
    (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    return e;
  }
  hipEvent_t e = c.evpool.back();
  c.evpool.pop_back();
  return e;
}

}  // namespace fp

#define FP_HIP_CHECK(expr)                                                   \
  do {                                                                       \
    hipError_t e_ = (expr);                                                  \
    if (e_ != hipSuccess) throw std::string(hipGetErrorString(e_)) + " @ " + #expr; \
  } while (0)
