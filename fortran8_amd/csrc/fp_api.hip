// C-ABI of the MI355X FENE-P / IBM SMAC step (include/fenep_gpu.h).
// Host orchestration: one context = one HIP stream + device-resident state;
// each fp_step enqueues the kernels of one reference time step
// (fenep.f90:2477-2527) per step.
#include "fp_internal.h"
#include "fp_mixed.h"
#include <cmath>
#include <cstring>
#include <cstdio>
#include <mutex>
#include <new>

using namespace fp;

namespace {

const double kPi = 3.14159265358979323846;

bool pow2(int n) { return n >= 4 && (n & (n - 1)) == 0; }

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

// twiddles (cos, -sin)(2 pi j/n), j < n/2, and modified wavenumbers
// 2(1-cos(2 pi m/N))/h^2, computed on the host exactly as the CPU restatement
std::vector<double2> twiddles(int n) {
  std::vector<double2> tw(n / 2 > 0 ? n / 2 : 1);
  for (int j = 0; j < n / 2; j++) {
    double ang = (2.0 * kPi * (double)j) / (double)n;
    tw[j] = make_double2(std::cos(ang), -std::sin(ang));
  }
  return tw;
}
std::vector<double> modwav(int n, double h, int count) {
  std::vector<double> k2(count);
  for (int m = 0; m < count; m++) {
    double th = (2.0 * kPi * (double)m) / (double)n;
    k2[m] = (2.0 * (1.0 - std::cos(th))) / (h * h);
  }
  return k2;
}

// Setters replace state that an fp_step(sync = 0) still in flight reads:
// join the selection stream and drain the context's stream first.
void drain(Ctx& c) {
  if (c.mixed) mixed_join(c);
  FP_HIP_CHECK(hipStreamSynchronize(c.stream));
}

int fail(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// Host copy of the boundary rules (used to fill ghosts of arrays handed out).
void host_fill(const Grid& g, double* A, int kind) {
  const long gx = g.nx + 2, gy = g.ny + 2, gz = g.nz + 2;
  auto G = [&](long i, long j, long k) { return i + gx * (j + gy * k); };
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  for (long k = 0; k < gz; k++)
    for (long j = 0; j < gy; j++) { A[G(0, j, k)] = A[G(nx, j, k)]; A[G(nx + 1, j, k)] = A[G(1, j, k)]; }
  for (long k = 0; k < gz; k++)
    for (long i = 0; i < gx; i++) {
      if (!g.walls) { A[G(i, 0, k)] = A[G(i, ny, k)]; A[G(i, ny + 1, k)] = A[G(i, 1, k)]; continue; }
      switch (kind) {
        case BC_U: case BC_W:
          A[G(i, 0, k)] = -A[G(i, 1, k)]; A[G(i, ny + 1, k)] = 2.0 * 0.0 - A[G(i, ny, k)]; break;
        case BC_V:
          A[G(i, 0, k)] = 0.0; A[G(i, ny, k)] = 0.0; A[G(i, ny + 1, k)] = 0.0; break;
        case BC_P:
          A[G(i, 0, k)] = A[G(i, 1, k)]; A[G(i, ny + 1, k)] = A[G(i, ny, k)]; break;
        default:
          A[G(i, 0, k)] = 2.0 * A[G(i, 1, k)] - A[G(i, 2, k)];
          A[G(i, ny + 1, k)] = 2.0 * A[G(i, ny, k)] - A[G(i, ny - 1, k)];
      }
    }
  for (long j = 0; j < gy; j++)
    for (long i = 0; i < gx; i++) { A[G(i, j, 0)] = A[G(i, j, nz)]; A[G(i, j, nz + 1)] = A[G(i, j, 1)]; }
}

void ghosted_to_interior(const Grid& g, const double* src, long sstride, long soff, double* dst) {
  const long gx = g.nx + 2, gy = g.ny + 2;
  for (long k = 0; k < g.nz; k++)
    for (long j = 0; j < g.ny; j++)
      for (long i = 0; i < g.nx; i++)
        dst[i + g.nx * (j + (long)g.ny * k)] = src[sstride * ((i + 1) + gx * ((j + 1) + gy * (k + 1))) + soff];
}
void interior_to_ghosted(const Grid& g, const double* src, double* dst, long dstride, long doff) {
  const long gx = g.nx + 2, gy = g.ny + 2;
  for (long k = 0; k < g.nz; k++)
    for (long j = 0; j < g.ny; j++)
      for (long i = 0; i < g.nx; i++)
        dst[dstride * ((i + 1) + gx * ((j + 1) + gy * (k + 1))) + doff] = src[i + g.nx * (j + (long)g.ny * k)];
}

struct Timer {
  Ctx& c;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  Timer(Ctx& cc, const char* nm) : c(cc), name(nm) {
    if (!phase_timed(c, name)) return;
    a = ev_get(c); b = ev_get(c);
    (void)hipEventRecord(a, c.stream);
  }
  ~Timer() {
    if (!phase_timed(c, name)) return;
    (void)hipEventRecord(b, c.stream);
    c.ev.push_back({std::string(name) + "#a", a});
    c.ev.push_back({std::string(name) + "#b", b});
  }
};

void flush_timers(Ctx& c) {
  if (c.ev.empty()) return;
  (void)hipStreamSynchronize(c.stream);
  for (size_t e = 0; e + 1 < c.ev.size(); e += 2) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c.ev[e].second, c.ev[e + 1].second);
    std::string nm = c.ev[e].first.substr(0, c.ev[e].first.size() - 2);
    bool found = false;
    for (auto& kv : c.ktime) if (kv.first == nm) { kv.second += ms; found = true; }
    if (!found) c.ktime.push_back({nm, (double)ms});
    c.evpool.push_back(c.ev[e].second);
    c.evpool.push_back(c.ev[e + 1].second);
  }
  c.ev.clear();
}

}  // namespace

void fp::step_f64(Ctx& c) {
  const Grid& g = c.g;
  if (c.p.beta == 1.0) {
    f64_zero_c(c);
  } else {
    if (c.sel_replay) {
      FP_HIP_CHECK(hipMemcpyAsync(c.sel, c.sel_replay, g.n, hipMemcpyDeviceToDevice, c.stream));
    } else {
      Timer t(c, "fenep_select");
      f64_select(c);
    }
    { Timer t(c, "fenep_cstar_newton"); f64_cstar_newton(c); }
  }
  { Timer t(c, "predictor"); f64_predictor(c); }
  if (c.p.method == FP_METHOD_FFT) {
    Timer t(c, "helmholtz");
    for (int comp = 0; comp < 3; comp++) f64_solve(c, comp, c.Q + comp * g.n, c.Up + comp * g.n);
  } else if (c.p.method == FP_METHOD_EXPLICIT) {
    FP_HIP_CHECK(hipMemcpyAsync(c.Up, c.Q, sizeof(double) * 3 * g.n, hipMemcpyDeviceToDevice, c.stream));
  } else {
    { Timer t(c, "ibm"); f64_ibm(c); }
    Timer t(c, "helmholtz");
    for (int comp = 0; comp < 3; comp++) f64_solve(c, comp, c.Up + comp * g.n, c.Up + comp * g.n);
  }
  {
    Timer t(c, "poisson");
    f64_divergence(c, c.Up, c.Q);
    if (c.p.poisson_variant == FP_POISSON_SOR) c.sor_itr = (double)f64_sor(c, c.Q, c.Phi);
    else { f64_solve(c, 3, c.Q, c.Phi); c.sor_itr = -1.0; }
  }
  { Timer t(c, "march"); f64_march(c); }
  c.has_history = 1;
  c.step++;
}

extern "C" {

int fp_default_params(fp_params* p) {
  if (!p) return FP_E_ARG;
  std::memset(p, 0, sizeof(*p));
  p->nx = p->ny = p->nz = 32;
  p->lx = p->ly = p->lz = 2.0 * kPi;
  p->dt = 0.001; p->re = 2.0; p->beta = 0.9; p->wi = 2.0; p->lp = 100.0;
  p->method = FP_METHOD_FFT; p->walls_y = 0; p->poisson_variant = FP_POISSON_FFT; p->precision = FP_PREC_F64;
  p->newton_itrmax = 10000; p->newton_eps = 1.0e-6; p->newton_dc = 1.0e-5;
  p->sor_itrmax = 10000; p->sor_eps = 1.0e-5; p->sor_omega = 0.0;
  p->n_dc_max = 2; p->dc_tol = 1.0e-5;
  return FP_OK;
}

int fp_create(const fp_params* p, void** out) {
  if (!p || !out) return FP_E_ARG;
  *out = nullptr;
  if (!pow2(p->nx) || !pow2(p->ny) || !pow2(p->nz) || p->nx > 1024 || p->ny > 1024 || p->nz > 1024)
    return FP_E_ARG;
  if (p->method < 0 || p->method > 2 || p->precision < 0 || p->precision > 1 || p->dt <= 0 || p->re <= 0 ||
      p->wi <= 0 || p->lx <= 0 || p->ly <= 0 || p->lz <= 0)
    return FP_E_ARG;
  if (p->walls_y && p->ny < 4) return FP_E_ARG;
  // the mixed kernels address each field plane with 32-bit byte offsets
  if (p->precision == FP_PREC_MIXED &&
      (p->method == FP_METHOD_EXPLICIT || (long)p->nx * p->ny * p->nz >= (1L << 30)))
    return FP_E_ARG;
  Ctx* c = new (std::nothrow) Ctx();
  if (!c) return FP_E_NOMEM;
  try {
    c->p = *p;
    Grid& g = c->g;
    g.nx = p->nx; g.ny = p->ny; g.nz = p->nz; g.nxh = p->nx / 2 + 1;
    g.lnx = 0; while ((1 << g.lnx) < g.nx) g.lnx++;
    g.lny = 0; while ((1 << g.lny) < g.ny) g.lny++;
    g.walls = p->walls_y ? 1 : 0;
    g.n = (long)p->nx * p->ny * p->nz;
    g.dx = p->lx / p->nx; g.dy = p->ly / p->ny; g.dz = p->lz / p->nz;
    g.dt = p->dt; g.re = p->re; g.beta = p->beta; g.wi = p->wi; g.lp = p->lp;
    g.sw = p->ibm_weight > 0.0 ? p->ibm_weight : (g.dx * g.dx) * g.dx;
    g.sy = g.walls ? p->ny + 2 : p->ny;
    g.jo = g.walls ? 1 : 0;
    FP_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const long n = g.n;
    dalloc(&c->U, n); dalloc(&c->V, n); dalloc(&c->W, n); dalloc(&c->P, n); dalloc(&c->Phi, n);
    dalloc(&c->C[0], 6 * n); dalloc(&c->C[1], 6 * n);
    dalloc(&c->fC, n); dalloc(&c->sel, n); dalloc(&c->nitr, n);
    dalloc(&c->A0, 3 * n); dalloc(&c->T0, 3 * n); dalloc(&c->B0, 3 * n); dalloc(&c->B, 3 * n);
    dalloc(&c->Q, 3 * n); dalloc(&c->Up, 3 * n);
    dalloc(&c->spec, (size_t)g.nxh * g.sy * g.nz);
    if (g.walls) dalloc(&c->thp, (size_t)(g.ny + 2) * g.nxh * g.nz);
    dalloc(&c->red, 64);
    dalloc(&c->part, (size_t)(g.n / 256 + 1024));
    auto up2 = [&](double2** d, const std::vector<double2>& h) {
      dalloc(d, h.size());
      FP_HIP_CHECK(hipMemcpy(*d, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice));
    };
    auto up1 = [&](double** d, const std::vector<double>& h) {
      dalloc(d, h.size());
      FP_HIP_CHECK(hipMemcpy(*d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    };
    up2(&c->twx, twiddles(g.nx)); up2(&c->twy, twiddles(g.ny)); up2(&c->twz, twiddles(g.nz));
    up1(&c->kx2, modwav(g.nx, g.dx, g.nxh)); up1(&c->ky2, modwav(g.ny, g.dy, g.ny)); up1(&c->kz2, modwav(g.nz, g.dz, g.nz));
    c->gn = (long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    if (p->precision == FP_PREC_MIXED) mixed_create(*c);
    FP_HIP_CHECK(hipDeviceSynchronize());
  } catch (const std::string& e) {
    fp_destroy(c);
    return FP_E_HIP;
  } catch (const std::bad_alloc&) {
    fp_destroy(c);
    return FP_E_NOMEM;
  }
  *out = c;
  return FP_OK;
}

int fp_destroy(void* vc) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->ev) (void)hipEventDestroy(kv.second);
  for (auto e : c->evpool) (void)hipEventDestroy(e);
  if (c->mixed) mixed_destroy(*c);
  adj_free(*c);
  dfree(c->U); dfree(c->V); dfree(c->W); dfree(c->P); dfree(c->Phi);
  dfree(c->C[0]); dfree(c->C[1]); dfree(c->fC); dfree(c->sel); dfree(c->nitr);
  dfree(c->A0); dfree(c->T0); dfree(c->B0); dfree(c->B); dfree(c->Q); dfree(c->Up); dfree(c->F);
  dfree(c->spec); dfree(c->thp); dfree(c->red); dfree(c->part);
  dfree(c->twx); dfree(c->twy); dfree(c->twz); dfree(c->kx2); dfree(c->ky2); dfree(c->kz2);
  dfree(c->mk); dfree(c->mf); dfree(c->mb); dfree(c->ftmp);
  dfree(c->mean_acc); dfree(c->mean_u);
  for (int k = 0; k < 3; k++) { dfree(c->sp_row[k]); dfree(c->sp_tgt[k]); dfree(c->sp_ent[k]); dfree(c->sp_w[k]); }
  dfree(c->ib_w);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FP_OK;
}

const char* fp_last_error(void* vc) {
  Ctx* c = (Ctx*)vc;
  return c ? c->err.c_str() : "null context";
}

int fp_set_state(void* vc, const double* U, const double* V, const double* W, const double* P, const double* C6,
                 const double* hist, int has_history) {
  Ctx* c = (Ctx*)vc;
  if (!c || !U || !V || !W || !P || !C6) return fail(c, FP_E_ARG, "null argument");
  try {
    drain(*c);
    const Grid& g = c->g;
    const long n = g.n;
    std::vector<double> buf(6 * n);
    ghosted_to_interior(g, U, 1, 0, buf.data());
    ghosted_to_interior(g, V, 1, 0, buf.data() + n);
    ghosted_to_interior(g, W, 1, 0, buf.data() + 2 * n);
    ghosted_to_interior(g, P, 1, 0, buf.data() + 3 * n);
    if (g.walls)  // wall face of V (V_kabe)
      for (long k = 0; k < g.nz; k++)
        for (long i = 0; i < g.nx; i++) buf[n + i + g.nx * ((g.ny - 1) + (long)g.ny * k)] = 0.0;
    FP_HIP_CHECK(hipMemcpy(c->U, buf.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    FP_HIP_CHECK(hipMemcpy(c->V, buf.data() + n, sizeof(double) * n, hipMemcpyHostToDevice));
    FP_HIP_CHECK(hipMemcpy(c->W, buf.data() + 2 * n, sizeof(double) * n, hipMemcpyHostToDevice));
    FP_HIP_CHECK(hipMemcpy(c->P, buf.data() + 3 * n, sizeof(double) * n, hipMemcpyHostToDevice));
    for (int q = 0; q < 6; q++) ghosted_to_interior(g, C6, 6, q, buf.data() + q * n);
    c->cur = 0;
    FP_HIP_CHECK(hipMemcpy(c->C[0], buf.data(), sizeof(double) * 6 * n, hipMemcpyHostToDevice));
    if (hist) {
      FP_HIP_CHECK(hipMemcpy(c->A0, hist, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
      FP_HIP_CHECK(hipMemcpy(c->T0, hist + 3 * n, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
      FP_HIP_CHECK(hipMemcpy(c->B0, hist + 6 * n, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
    }
    FP_HIP_CHECK(hipMemset(c->Phi, 0, sizeof(double) * n));
    c->has_history = (has_history && hist) ? 1 : 0;
    c->state_set = 1;
    if (c->mixed) mixed_load_from_f64(*c);
    FP_HIP_CHECK(hipDeviceSynchronize());
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_get_state(void* vc, double* U, double* V, double* W, double* P, double* C6, double* hist) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  try {
    if (c->mixed) mixed_store_to_f64(*c);
    FP_HIP_CHECK(hipStreamSynchronize(c->stream));
    const Grid& g = c->g;
    const long n = g.n;
    std::vector<double> buf(6 * n);
    double* outs[4] = {U, V, W, P};
    double* srcs[4] = {c->U, c->V, c->W, c->P};
    int kinds[4] = {BC_U, BC_V, BC_W, BC_P};
    for (int a = 0; a < 4; a++) {
      if (!outs[a]) continue;
      FP_HIP_CHECK(hipMemcpy(buf.data(), srcs[a], sizeof(double) * n, hipMemcpyDeviceToHost));
      interior_to_ghosted(g, buf.data(), outs[a], 1, 0);
      host_fill(g, outs[a], kinds[a]);
    }
    if (C6) {
      FP_HIP_CHECK(hipMemcpy(buf.data(), c->C[c->cur], sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
      std::vector<double> tmp(c->gn);
      for (int q = 0; q < 6; q++) {
        interior_to_ghosted(g, buf.data() + q * n, tmp.data(), 1, 0);
        host_fill(g, tmp.data(), BC_C);
        for (long t = 0; t < c->gn; t++) C6[6 * t + q] = tmp[t];
      }
    }
    if (hist) {
      FP_HIP_CHECK(hipMemcpy(hist, c->A0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
      FP_HIP_CHECK(hipMemcpy(hist + 3 * n, c->T0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
      FP_HIP_CHECK(hipMemcpy(hist + 6 * n, c->B0, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    }
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_get_phi(void* vc, double* Phi) {
  Ctx* c = (Ctx*)vc;
  if (!c || !Phi) return FP_E_ARG;
  try {
    if (c->mixed) mixed_store_to_f64(*c);
    FP_HIP_CHECK(hipStreamSynchronize(c->stream));
    std::vector<double> buf(c->g.n);
    FP_HIP_CHECK(hipMemcpy(buf.data(), c->Phi, sizeof(double) * c->g.n, hipMemcpyDeviceToHost));
    interior_to_ghosted(c->g, buf.data(), Phi, 1, 0);
    host_fill(c->g, Phi, BC_P);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_set_body_force(void* vc, const double* fx, const double* fy, const double* fz) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  try {
    drain(*c);
    adj_drop_tape(*c);
    const long n = c->g.n;
    if (!fx && !fy && !fz) { dfree(c->F); if (c->mixed) mixed_set_force(*c); return FP_OK; }
    if (!c->F) dalloc(&c->F, 3 * n);
    const double* f[3] = {fx, fy, fz};
    for (int k = 0; k < 3; k++) {
      if (f[k]) FP_HIP_CHECK(hipMemcpy(c->F + k * n, f[k], sizeof(double) * n, hipMemcpyHostToDevice));
      else FP_HIP_CHECK(hipMemset(c->F + k * n, 0, sizeof(double) * n));
    }
    if (c->mixed) mixed_set_force(*c);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_set_markers(void* vc, int nc, int nl, const double* xc, const double* yc, const double* zc, const double* uc,
                   const double* vc_, const double* wc) {
  Ctx* c = (Ctx*)vc;
  if (!c || nc <= 0 || nl <= 0 || !xc || !yc || !zc || !uc || !vc_ || !wc) return fail(c, FP_E_ARG, "bad markers");
  try {
    drain(*c);
    adj_drop_tape(*c);
    const Grid& g = c->g;
    long nmk = (long)nc * nl * g.nz;
    // every stencil must stay within one ghost layer (and off the walls)
    for (long id = 0; id < nmk; id++) {
      double x = xc[id], y = yc[id];
      if (!(x >= 0.5 * g.dx && x <= (g.nx - 0.5) * g.dx && y >= 1.5 * g.dy && y <= (g.ny - 1.5) * g.dy))
        return fail(c, FP_E_ARG, "marker too close to the domain boundary");
    }
    dfree(c->mk); dfree(c->mf); dfree(c->mb); dfree(c->ftmp);
    c->nc = nc; c->nl = nl; c->nmk = nmk;
    dalloc(&c->mk, 6 * nmk); dalloc(&c->mf, 3 * nmk); dalloc(&c->mb, 3 * nmk);
    dalloc(&c->ftmp, 3 * c->gn);
    const double* src[6] = {xc, yc, zc, uc, vc_, wc};
    for (int a = 0; a < 6; a++) FP_HIP_CHECK(hipMemcpy(c->mk + a * nmk, src[a], sizeof(double) * nmk, hipMemcpyHostToDevice));
    f64_ibm_setup(*c);
    if (c->mixed) mixed_set_markers(*c);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_get_marker_forces(void* vc, double* fx, double* fy, double* fz) {
  Ctx* c = (Ctx*)vc;
  if (!c || !c->nmk) return FP_E_STATE;
  try {
    FP_HIP_CHECK(hipStreamSynchronize(c->stream));
    const double* mf = c->mixed ? mixed_marker_forces(*c) : c->mf;
    double* out[3] = {fx, fy, fz};
    for (int a = 0; a < 3; a++)
      if (out[a]) FP_HIP_CHECK(hipMemcpy(out[a], mf + a * c->nmk, sizeof(double) * c->nmk, hipMemcpyDeviceToHost));
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_step(void* vc, int nsteps, int sync) {
  Ctx* c = (Ctx*)vc;
  if (!c || nsteps < 0) return FP_E_ARG;
  if (!c->state_set) return fail(c, FP_E_STATE, "fp_step before fp_set_state");
  if (c->p.method == FP_METHOD_IBM && c->nmk == 0) return fail(c, FP_E_STATE, "IBM method without markers");
  try {
    for (int s = 0; s < nsteps; s++) {
      if (c->mixed) mixed_step(*c);
      else step_f64(*c);
    }
    if (c->mixed) mixed_join(*c);
    if (sync || c->timing) FP_HIP_CHECK(hipStreamSynchronize(c->stream));
    flush_timers(*c);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

// Body of fp_diagnostics and fp_diagnostics_static: the reductions run in
// `scratch` (diag_scratch_words(c.g) doubles of device memory).
static void diagnostics_into(Ctx& c, fp_diag* d, double* scratch) {
  std::memset(d, 0, sizeof(*d));
  if (c.mixed) mixed_store_to_f64(c);
  const double* mf = c.nmk ? (c.mixed ? mixed_marker_forces(c) : c.mf) : nullptr;
  f64_diag(c, d, scratch, mf);
  d->sor_itr = c.sor_itr;
  d->n_dc = c.mixed ? mixed_last_ndc(c) : 0.0;
  if (c.mixed) {
    double st[3];
    mixed_stats(c, st);
    d->dc_residual = st[0] > 0 ? std::sqrt(st[1] / st[0]) : 0.0;
    d->sel_guarded = st[2];
  }
  d->step = c.step;
}

int fp_diagnostics(void* vc, fp_diag* d) {
  Ctx* c = (Ctx*)vc;
  if (!c || !d) return FP_E_ARG;
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  double* scratch = nullptr;
  try {
    FP_HIP_CHECK(hipMalloc((void**)&scratch, sizeof(double) * diag_scratch_words(c->g)));
    diagnostics_into(*c, d, scratch);
  } catch (const std::string& e) {
    if (scratch) (void)hipFree(scratch);
    return fail(c, FP_E_HIP, e);
  }
  (void)hipFree(scratch);
  return d->nan_flag ? FP_E_NAN : FP_OK;
}

}  // extern "C"

// ---- fp_diagnostics_static: the diagnostics without any allocation.
// Its reductions use this file-scope device array instead of a buffer from
// hipMalloc.  Each GPU holds its own copy (6.3 MB, for grids up to
// FP_DIAG_STATIC_MAX_CELLS cells) until the process ends.  In host code the
// name `scratch` is not a device address: the address is taken once per
// device with hipGetSymbolAddress and only that pointer is used.
namespace {
constexpr size_t kScratchWords = (size_t)12 * (FP_DIAG_STATIC_MAX_CELLS / 256) + 16;
static __device__ __align__(256) double scratch[kScratchWords];
constexpr int kMaxDevices = 64;
double* g_scratch_ptr[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

// the current device's copy of `scratch`, looked up on its first use
int scratch_ptr(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FP_E_HIP;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FP_E_HIP;
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return FP_OK;
}
}  // namespace

extern "C" {

int fp_diagnostics_static(void* vc, fp_diag* d) {
  Ctx* c = (Ctx*)vc;
  if (!c || !d) return FP_E_ARG;
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  if (c->g.n > (long)FP_DIAG_STATIC_MAX_CELLS || diag_scratch_words(c->g) > kScratchWords)
    return fail(c, FP_E_ARG, "fp_diagnostics_static: more than FP_DIAG_STATIC_MAX_CELLS cells");
  double* ptr = nullptr;
  if (scratch_ptr(&ptr) != FP_OK) return fail(c, FP_E_HIP, "hipGetSymbolAddress of the static scratch failed");
  try {
    diagnostics_into(*c, d, ptr);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return d->nan_flag ? FP_E_NAN : FP_OK;
}

int fp_set_state_device(void* vc, const double* U, const double* V, const double* W, const double* P,
                        const double* C6, int has_history, void* stream) {
  Ctx* c = (Ctx*)vc;
  if (!c || !U || !V || !W || !P || !C6) return FP_E_ARG;
  try {
    drain(*c);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const long n = c->g.n;
    FP_HIP_CHECK(hipMemcpyAsync(c->U, U, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FP_HIP_CHECK(hipMemcpyAsync(c->V, V, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FP_HIP_CHECK(hipMemcpyAsync(c->W, W, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    FP_HIP_CHECK(hipMemcpyAsync(c->P, P, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    c->cur = 0;
    FP_HIP_CHECK(hipMemcpyAsync(c->C[0], C6, sizeof(double) * 6 * n, hipMemcpyDeviceToDevice, s));
    if (has_history == 2) {
      // the potential stays: the SOR variant's sweeps start from it
      if (c->mixed) mixed_phi_to_f64(*c);
    } else {
      FP_HIP_CHECK(hipMemsetAsync(c->Phi, 0, sizeof(double) * n, s));
    }
    FP_HIP_CHECK(hipStreamSynchronize(s));
    FP_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->has_history = has_history ? 1 : 0;
    c->state_set = 1;
    if (c->mixed) mixed_load_from_f64(*c);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_get_state_device(void* vc, double* U, double* V, double* W, double* P, double* C6, void* stream) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  try {
    if (c->mixed) mixed_store_to_f64(*c);
    FP_HIP_CHECK(hipStreamSynchronize(c->stream));
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const long n = c->g.n;
    if (U) FP_HIP_CHECK(hipMemcpyAsync(U, c->U, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (V) FP_HIP_CHECK(hipMemcpyAsync(V, c->V, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (W) FP_HIP_CHECK(hipMemcpyAsync(W, c->W, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (P) FP_HIP_CHECK(hipMemcpyAsync(P, c->P, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    if (C6) FP_HIP_CHECK(hipMemcpyAsync(C6, c->C[c->cur], sizeof(double) * 6 * n, hipMemcpyDeviceToDevice, s));
    FP_HIP_CHECK(hipStreamSynchronize(s));
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_enable_timing(void* vc, int on) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  if (on < 0 || on > 3) return fail(c, FP_E_ARG, "timing level is 0, 1, 2 or 3");
  c->timing = on;
  c->ktime.clear();
  return FP_OK;
}

int fp_kernel_times(void* vc, double* ms, int max_n, char* names, int names_len) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  std::string all;
  int k = 0;
  for (auto& kv : c->ktime) {
    if (k >= max_n) break;
    ms[k++] = kv.second;
    if (!all.empty()) all += ",";
    all += kv.first;
  }
  if (names && names_len > 0) {
    std::strncpy(names, all.c_str(), names_len - 1);
    names[names_len - 1] = 0;
  }
  return k;
}

#ifndef FP_SOURCE_HASH
#define FP_SOURCE_HASH "unknown"
#endif
const char* fp_build_id(void) { return FP_SOURCE_HASH; }

int fp_seeded_uniform(uint64_t seed, uint64_t offset, size_t n, double* out) {
  if (!out) return FP_E_ARG;
  for (size_t i = 0; i < n; i++) {
    uint64_t z = seed + (offset + i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    out[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  }
  return FP_OK;
}

int fp_time_stats(void* vc, double* out) {
  Ctx* c = (Ctx*)vc;
  if (!c || !out) return FP_E_ARG;
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  try {
    if (c->mixed) mixed_store_to_f64(*c);
    f64_time_stats(*c, out);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_energy_spectrum(void* vc, double* E, int nbins) {
  Ctx* c = (Ctx*)vc;
  if (!c || !E || nbins < 1) return FP_E_ARG;
  if (c->g.walls) return fail(c, FP_E_ARG, "the shell spectrum needs a periodic grid");
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  try {
    if (c->mixed) mixed_store_to_f64(*c);
    f64_energy_spectrum(*c, E, nbins);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_time_stats_map(void* vc, int ncyl, const double* xy, double diameter, double* map18, double* out13) {
  Ctx* c = (Ctx*)vc;
  if (!c || ncyl < 0 || (ncyl > 0 && !xy) || !(diameter >= 0.0) || (!map18 && !out13)) return fail(c, FP_E_ARG, "bad arguments");
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  try {
    if (c->mixed) mixed_store_to_f64(*c);
    if (f64_time_stats_map(*c, ncyl, xy, diameter, map18, out13)) return fail(c, FP_E_ARG, "the mask covers the whole plane");
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_energy_spectrum_z(void* vc, int ncyl, const double* xy, double diameter, double* Emask, double* Efirst) {
  Ctx* c = (Ctx*)vc;
  if (!c || ncyl < 0 || (ncyl > 0 && !xy) || !(diameter >= 0.0) || (!Emask && !Efirst)) return fail(c, FP_E_ARG, "bad arguments");
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  try {
    if (c->mixed) mixed_store_to_f64(*c);
    if (f64_energy_spectrum_z(*c, ncyl, xy, diameter, Emask, Efirst)) return fail(c, FP_E_ARG, "the mask covers the whole plane");
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_mean_accumulate(void* vc) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  try {
    const float *u = nullptr, *v = nullptr, *w = nullptr;
    if (c->mixed) mixed_velocity(*c, &u, &v, &w);
    f64_mean_accumulate(*c, u, v, w);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_mean_apply(void* vc, int mode) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  if (mode < 0 || mode > 2) return fail(c, FP_E_ARG, "mode is 0, 1 or 2");
  try {
    if (f64_mean_apply(*c, mode)) return fail(c, FP_E_STATE, "no accumulated velocity to average");
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_check_selection(void* vc, double* out) {
  Ctx* c = (Ctx*)vc;
  if (!c || !out) return FP_E_ARG;
  if (!c->mixed) return fail(c, FP_E_ARG, "the selection check applies to the mixed-precision path");
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  try {
    mixed_check_selection(*c, out);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

int fp_get_selection(void* vc, unsigned char* out) {
  Ctx* c = (Ctx*)vc;
  if (!c || !out) return FP_E_ARG;
  if (!c->mixed) return fail(c, FP_E_ARG, "the selection bytes are those of the mixed-precision path");
  if (!c->state_set) return fail(c, FP_E_STATE, "state not set");
  try {
    mixed_get_selection(*c, out);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  return FP_OK;
}

// Errors of the recording/adjoint layer arrive as "E_ARG:..." or
// "E_STATE:..." (argument / call order) or as HIP errors.
static int adj_fail(Ctx* c, const std::string& e) {
  if (e.rfind("E_ARG:", 0) == 0) return fail(c, FP_E_ARG, e.substr(6));
  if (e.rfind("E_STATE:", 0) == 0) return fail(c, FP_E_STATE, e.substr(8));
  return fail(c, FP_E_HIP, e);
}

int fp_set_physics(void* vc, double wi, double beta, double lp) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  if (!(wi > 0.0) || !(beta > 0.0 && beta <= 1.0) || !(lp * lp > 3.0))
    return fail(c, FP_E_ARG, "need Wi > 0, 0 < beta <= 1 and Lp^2 > 3");
  try {
    drain(*c);
  } catch (const std::string& e) {
    return fail(c, FP_E_HIP, e);
  }
  adj_drop_tape(*c);
  c->p.wi = wi; c->p.beta = beta; c->p.lp = lp;
  c->g.wi = wi; c->g.beta = beta; c->g.lp = lp;
  if (c->mixed) {
    try {
      mixed_physics(*c);
    } catch (const std::string& e) {
      return fail(c, FP_E_HIP, e);
    }
  }
  return FP_OK;
}

int fp_adj_record(void* vc, int nsteps, int every, int replay, double* drag_out) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  if (!c->state_set) return fail(c, FP_E_STATE, "fp_adj_record before fp_set_state");
  try {
    adj_record(*c, nsteps, every, replay, drag_out);
  } catch (const std::string& e) {
    return adj_fail(c, e);
  }
  return FP_OK;
}

int fp_adj_backward(void* vc, const double* weights, const double* bar_final, double* bar_initial,
                    double* bar_params, double* bar_force) {
  Ctx* c = (Ctx*)vc;
  if (!c) return FP_E_ARG;
  try {
    adj_backward(*c, weights, bar_final, bar_initial, bar_params, bar_force);
  } catch (const std::string& e) {
    return adj_fail(c, e);
  }
  return FP_OK;
}

}  // extern "C"
