/*
 * mpulse_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (float64, single-threaded unless built with OpenMP) of the MPulse
 * time-stepping path.  It is the checker for the HIP library: only tests/, the smoke
 * entry and the CPU-baseline leg of the benchmark may load it.  No product code links it.
 */
#ifndef MPULSE_ORACLE_H
#define MPULSE_ORACLE_H

#include <stddef.h>
#include "../include/mpulse_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_sim or_sim;

or_sim *or_create(const mpx_desc *desc, char *err, size_t errlen);
void    or_destroy(or_sim *s);
int     or_step(or_sim *s, int nsteps);
int     or_step_init(or_sim *s);
int     or_get_time(or_sim *s);
/* Storage rounding applied after every update of E, B and the CPML auxiliaries: 0 float64
   (the reference), 1 float32, 2 float32 + 10-bit residual word (see store_round). */
int     or_set_storage(or_sim *s, int mode);
double  or_store_round(double x, int mode);
void    or_set_time(or_sim *s, int t);
int     or_get_grid(or_sim *s, const char *name, double *out, size_t count);
int     or_set_grid(or_sim *s, const char *name, const double *in, size_t count);
int     or_get_layer(or_sim *s, const char *name, int dir, double *out, size_t count,
                     int lo[3], int hi[3]);
int     or_get_line(or_sim *s, const char *name, double *out, size_t count);
double  or_energy(or_sim *s);

/* unit-level entry points, used by the pinning tests */
void    or_faddeeva(double re, double im, double out[2]);
void    or_pole_constants(const double l_eps[3], const double l_delta[3], const double l_om2[3],
                          double dt, double crec[6], double dchi0[6], double dxi0[6],
                          double sums[2]);
double  or_kerr_newton(double A, double C, double E0, int *iterations);
/* short-pulse source function on its own: params = {length,width,om0,tshift,zshift,phase,amp,eps} */
void    or_pulse_efunc(const double params[8], int normalise, double x, double y, double z,
                       double t, double out[2]);
void    or_pulse_bfunc(const double params[8], int normalise, double x, double y, double z,
                       double t, int bx, double out[2]);
double  or_pulse_amp(const double params[8]);
/* focused-pulse plane: comp 0 = E_x, 1 = B_x, 2 = B_y at distance z from the focus, Yee
   offsets (sx, sy); out = (nx+2)(ny+2) interleaved (re, im) */
/* The GaussInject sample block in force at step `time` for the E (isH 0) or H (isH 1) sheet:
   F1, F2 receive Nt x (nx+2) x (ny+2) values ([kh][i][j]); returns Nt, or 0 when the source is
   inactive at that step or on a bad argument (F1/F2 may be NULL to ask for Nt only). */
int     or_gauss_block(const mpx_factory *f, int isH, int nx, int ny, double dx, double dy, double dt, int lowz,
                       int highz, int time, double *F1, double *F2, int *block_toff);
void    or_focused_plane(const mpx_factory *f, int nx, int ny, double dx, double dy, double z, int comp,
                         double sx, double sy, double *out);
/* CPML coefficient table of one face current (isH, dir) as built by makeCoeff */
int     or_cpml_coeff(or_sim *s, int isH, int dir, double *b, double *c, int *lo, int *hi);

#ifdef __cplusplus
}
#endif
#endif
