/*
 * howmhd_oracle.h — CPU restatement of the HOW-MHD time step (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This library is the parity reference of the tests and the timed CPU
 * baseline.  Only the test suite, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg load it; the product (the HIP library and the host driver)
 * never does.
 *
 * It restates, in plain C with the reference's operation order, the serial
 * (one-rank) paths of:
 *   src/ssprk.cpp    54-140, 381-1329 (live code; dF summed in wave order 0..6)
 *   src/fluxct.cpp   14-285, 481-552, 753-899
 *   src/bound.cpp    19-107, 295-317
 *   src/tstep.cpp    11-85, 123-135
 *   src/prot.cpp     10-51
 * State is kept exactly as the reference keeps it: q(stage,var,i,j,k) with
 * 3 ghost cells per side (k fastest), 6 stages, face arrays bxb/byb/bzb(stage,
 * i,j,k) and the 12 CT work arrays.
 */
#ifndef HOWMHD_ORACLE_H
#define HOWMHD_ORACLE_H

#include "../include/howmhd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct howmhd_ref_ctx howmhd_ref_ctx;

int howmhd_ref_create(const howmhd_params* p, howmhd_ref_ctx** out);
int howmhd_ref_destroy(howmhd_ref_ctx* c);
int howmhd_ref_upload_q0(howmhd_ref_ctx* c, const double* q0);
int howmhd_ref_upload_faces(howmhd_ref_ctx* c, const double* b);
int howmhd_ref_init_faces(howmhd_ref_ctx* c);
int howmhd_ref_tstep(howmhd_ref_ctx* c, double* dt);
int howmhd_ref_ssprk(howmhd_ref_ctx* c, double dt);
int howmhd_ref_download_q0(howmhd_ref_ctx* c, double* q0);
int howmhd_ref_download_faces(howmhd_ref_ctx* c, double* b);
int howmhd_ref_set_threads(int n);
/* Test hook: cells whose density and pressure floors the stage-5 prot of the
 * last ssprk applied (value below the floor or not a number); 0 before the
 * first ssprk. */
int howmhd_ref_floor_counts(const howmhd_ref_ctx* c, long long* n_rho, long long* n_p);

/* Test hook: the corner-EMF construction of one CT update on given flux
 * arrays (src/fluxct.cpp:795-842, the code fluxct runs).  ct6 holds the
 * padded arrays fsy, fsz, gsx, gsz, hsx, hsy (6 x (nx+6)(ny+6)(nz+6)); o3
 * receives the smoothed corner EMFs Ox, Oy, Oz after their ghost fill.  A
 * context with fewer than two active axes returns HOWMHD_EINVAL. */
int howmhd_ref_ct_emf(howmhd_ref_ctx* c, const double* ct6, double* o3);

/* Test hook: the first nst (1..5) stages of one ssprk step with the given dt;
 * q_out (8 x cells) and b_out (3 x cells) receive stage nst's padded state.
 * q(0) and the faces of stage 0 are left as they are. */
int howmhd_ref_ssprk_stages(howmhd_ref_ctx* c, double dt, int nst, double* q_out, double* b_out);

/* --- single-line and single-interface pieces, for the oracle's own checks --- */

/* One directional sweep of one line (src/ssprk.cpp:458-997): qline holds
 * 8 x (nn+6) conserved values in sweep order (rho, Mn, Mt1, Mt2, Bn, Bt1, Bt2,
 * E).  On return qline[:, 3..nn+2] holds the increments -dtdl*(dF(i)-dF(i-1))
 * (normal-B slot 0), and bsy/bsz (nn+1 each) the CT flux terms. */
int howmhd_ref_line_sweep(double gam, double rhomin, double pgmin, int nn,
                          double dtdl, double* qline, double* bsy, double* bsz,
                          double* dF_out /* 7 x (nn+1), may be NULL */);

/* Interface eigensystem between two primitive states
 * (rho, vn, vt1, vt2, Bn, Bt1, Bt2, p, H) (src/ssprk.cpp:602-819).
 * L and R are 7x7 row-major: L[row=wave][col], R[row=variable][col=wave]. */
void howmhd_ref_eigensystem(double gam, const double* ul, const double* ur,
                            double* L, double* R);

/* Characteristic flux of one wave from its 6 projected values
 * (src/ssprk.cpp:883-937). */
double howmhd_ref_weno_wave(const double* Fsk, const double* qsk, double amx);

/* Same, also returning the nonlinear weights (w0,w1,w2) of the plus and the
 * minus side (src/ssprk.cpp:903-908, 926-931). */
double howmhd_ref_weno_wave_weights(const double* Fsk, const double* qsk,
                                    double amx, double* wplus, double* wminus);

/* --- learned-indicator form (no counterpart in the reference) --- */

/* Linear upwind-5 LLF flux of one component from 6 stencil values. */
double howmhd_ref_upwind5(const double* F, const double* U, double alpha);

/* 0: characteristic WENO5 at every interface (default); 1: linear upwind-5
 * at every interface (the indicator form with no interface flagged). */
int howmhd_ref_set_route(howmhd_ref_ctx* c, int route);
/* Test hook: WENO weight evaluations so far with some t_k = (eps + IS_k)^2
 * above 1e200 or not finite (reset != 0: start again from zero). */
long long howmhd_ref_large_t_count(int reset);
/* Test hook: sweep square-root operands so far outside [2^-767, inf) and not
 * +0 (reset != 0: start again from zero). */
long long howmhd_ref_odd_sqrt_count(int reset);

/* Indicator label source for one line (qline as in howmhd_ref_line_sweep,
 * not modified): maxdev[i], i = 0..nn, the largest deviation of the WENO
 * nonlinear weights from the ideal weights at interface i. */
int howmhd_ref_line_maxdev(double gam, double rhomin, double pgmin, int nn, const double* qline, double* maxdev);

#ifdef __cplusplus
}
#endif

#endif
