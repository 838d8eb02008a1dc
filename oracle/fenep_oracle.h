/*
 * fenep_oracle.h -- CPU restatement (FP64) of the reference time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product (fortran8_amd/libfenep_gpu.so) never links or calls it.
 *
 * Layouts follow the reference (Fortran column-major, x fastest):
 * ghosted scalars (0:NX+1,0:NY+1,0:NZ+1), C as (6,0:NX+1,0:NY+1,0:NZ+1),
 * histories (1:NX,1:NY,1:NZ).
 */
#ifndef FENEP_ORACLE_H
#define FENEP_ORACLE_H
#include "../include/fenep_gpu.h"
#ifdef __cplusplus
extern "C" {
#endif

void* orc_create(const fp_params* p);
void  orc_destroy(void* o);
int   orc_set_state(void* o, const double* U, const double* V, const double* W,
                    const double* P, const double* C6, const double* hist,
                    int has_history);
int   orc_get_state(void* o, double* U, double* V, double* W, double* P,
                    double* C6, double* hist);
int   orc_get_phi(void* o, double* Phi);
int   orc_set_body_force(void* o, const double* fx, const double* fy,
                         const double* fz);
int   orc_set_markers(void* o, int nc, int nl, const double* xc,
                      const double* yc, const double* zc, const double* uc,
                      const double* vc, const double* wc);
int   orc_get_marker_forces(void* o, double* fx, double* fy, double* fz);
int   orc_step(void* o, int nsteps);
int   orc_diagnostics(void* o, fp_diag* d);
/* counts of C+- selections by index 1..4 accumulated over all steps */
int   orc_selection_counts(void* o, double* counts4);

/* ---- operator-level entry points (for pinning tests) ---- */
/* evaluate convection (which=0), viscous (1), polymer stress (2) of the
 * current state into three interior arrays */
int   orc_eval_operator(void* o, int which, double* ox, double* oy, double* oz);
/* spectral/line solve on the context's grid: kind 0,1,2 = Helmholtz for the
 * u,v,w component [1-(dt beta/2Re) lap_h] x = q, kind 3 = Poisson lap_h x = q.
 * q and x are interior arrays. */
int   orc_solve(void* o, int kind, const double* q, double* x);
/* SOR Poisson (fenep.f90:848-905) on the context's grid; returns iterations */
int   orc_sor(void* o, const double* q, double* phi_ghosted);

int    orc_ibm_apply(void* o, const double* up3_ghosted, double* fint3,
                     double* fm3, double* ub3);

int orc_time_stats(void* o, double* out13);
double orc_vieta(const double* a6);
int    orc_select(const double* cm1, const double* c0, const double* cp1,
                  double* cp_out, double* cn_out);
void   orc_gauss6(const double* a_colmajor, const double* b, double* x);
int    orc_newton_cell(const double* a0t_colmajor, const double* b6,
                       const double* x0, double dt, double wi, double lp,
                       double dc, double eps, int itrmax, double* x);
double orc_roma_weight(double r);
void   orc_tdma(int n, const double* a, const double* b, const double* c,
                const double* dre, const double* dim, double* xre, double* xim);
void   orc_fft(int n, double* data_interleaved, int inverse);

/* ---- complex-step build only (liboracle_cs.so, oracle/cstep.h): the same
 * state entry points, and the seeds and drag of a complex-step run */
int orc_cs_seed_param(void* o, int which, double h);
int orc_cs_seed_force(void* o, const double* dF3, double h);
int orc_cs_drag(void* o, double* out2);

#ifdef __cplusplus
}
#endif
#endif
