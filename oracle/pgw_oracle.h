/*
 * pgw_oracle.h - CPU FP64 restatement of the polgraw-allsky network search core.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU library is compared
 * against, and the timed CPU baseline.  It is never linked into the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Each function names the reference file:line it restates.  Parity pinning:
 *   lin2ast, spline, splint, splintpad, var, FStat - against the compiled
 *     reference auxi.c (oracle/_ref) and against library routines;
 *   fft - against numpy.fft (pocketfft) and closed-form DFTs;
 *   modvir/rogcvir - against an independent detector-tensor computation
 *     (a^2+b^2 = F+^2 + Fx^2);
 *   job_core composition - resampling against band-limited closed forms,
 *     threshold against a brute-force run scan, and end-to-end injection
 *     recovery (candidate at the injected frequency/spindown/sky);
 *   fstatnet - noise-free injection (F = rho^2/2 + 2 at the true point, amplitude
 *     estimates equal the injected amplitudes), agreement with the grid search's F
 *     at the grid candidate, gradient against central differences, phase against
 *     an mpmath evaluation; amoeba - converges to the injected point; mads - its
 *     mesh order against the reference's array (hand-derived), its trajectory against the
 *     product's batched mesh search and an improvement from every start.
 */
#ifndef PGW_ORACLE_H
#define PGW_ORACLE_H

#include <complex.h>
#include "../include/pgw.h"

#ifdef __cplusplus
extern "C" {
#endif

/* FFT plan: Stockham autosort, radix 4 with one radix-2 stage when log2 n is odd */
typedef struct orc_fft_plan {
  int n;
  double complex *tw;    /* W_n^m = exp(-2 pi i m / n), m < n */
  double complex *work;  /* scratch, n */
} orc_fft_plan;

int  orc_fft_plan_init(orc_fft_plan *p, int n);
void orc_fft_plan_free(orc_fft_plan *p);
/* In-place DFT of x (length p->n). dir = -1: forward e^{-}, dir = +1: backward e^{+}; unnormalised.
 * work must be a distinct buffer of n elements (NULL: use the plan's own). */
void orc_fft_exec(const orc_fft_plan *p, double complex *x, int dir, double complex *work);

void orc_lin2ast(double be1, double be2, int pm, double sepsm, double cepsm,
                 double *sinal, double *cosal, double *sindel, double *cosdel);
void orc_rogcvir(double ephi, double egam, double c[9]);
void orc_modvir(double sinal, double cosal, double sindel, double cosdel,
                double sphir, double cphir, const double c[9], int Np,
                const double *cosmodf, const double *sinmodf, double *aa, double *bb);
void orc_spline(const double complex *y, int n, double complex *y2, double complex *u);
double complex orc_splint(const double complex *ya, const double complex *y2a, int n, double x);
void orc_splintpad(const double complex *ya, const double *shftf, int N, int interpftpad,
                   double complex *out);
double orc_var(const double *x, int n);
void orc_fstat_norm(double *F, int nfft, int nav);
int orc_settings(double fpo, double dt, int N, int fftpad, pgw_settings *sett);
void orc_resample(const orc_fft_plan *pl_int, const orc_fft_plan *pl_inv, const double complex *ina,
                  const double complex *inb, const double *shftf, int N, int nfft, int interpftpad,
                  double complex *xa, double complex *xb, double complex *work,
                  double complex *outa, double complex *outb);
int orc_runmax(const double *F, int nmin, int nmax, double trl, int *ii_out, double *F_out, int cap);

/* Search state for one segment (what init_arrays() builds, init.c:326-457). */
typedef struct orc_search orc_search;

orc_search *orc_create(const pgw_settings *sett, const pgw_detector *ifo, int nifo, int nthreads);
void orc_destroy(orc_search *s);

/* job_core restatement (jobcore.c:189-785).  nthreads > 1 parallelises like
 * src-openmp/jobcore.c: samples in the sky-point stage, spindowns in the template loop;
 * candidates are always returned in the serial (ss, bin) order.
 * F (optional, 2*nfft doubles) receives the last template's normalised F as in the reference.
 * Returns 0, or -2 if sgnlv_cap would be exceeded (*sgnlc then holds the count needed). */
int orc_job_core(orc_search *s, int pm, int mm, int nn, const pgw_range *s_range,
                 const pgw_opts *opts, double *F, int *sgnlc, double *sgnlv, int sgnlv_cap,
                 int *FNum);

/* F[nmin..nmax) (normalised, as thresholded) for one template (pm, mm, nn, ss). */
int orc_fstat_one(orc_search *s, int pm, int mm, int nn, int ss, const pgw_opts *opts, double *F);

/* resampled, noise-weighted xDatma/xDatmb of one detector for one sky point (jobcore.c:270-393) */
int orc_resampled(orc_search *s, int pm, int mm, int nn, int det, double complex *xa, double complex *xb);

/* spindown range of a sky point as job_core computes it (jobcore.c:204,424-444) */
void orc_spindown_range(const pgw_settings *sett, int mm, int nn, const pgw_range *s_range,
                        const pgw_opts *opts, int *smin, int *smax);

/* Fstatnet (followup.c:413-534) at continuous pt = (omega, omegadot, delta, alpha): out[11] as the
 * reference (out[5] = -F); grad (optional) = dF/dpt of F = -out[5].  Thread-safe. */
int orc_fstatnet(const orc_search *s, const double pt[4], double out[11], double grad[4]);
/* n points pts[4n] -> out[11n], grad[4n] (optional); OpenMP over points */
int orc_fstatnet_batch(const orc_search *s, const double *pts, int n, double *out, double *grad);
/* invertedMADS (followup.c:537-754): mesh search from point with final mesh sizes p2 */
int orc_mads(const orc_search *s, double point[4], const double p2[4], double nod, int limit, double out[11]);
/* its 80 mesh offsets (signs) in the reference's array order */
void orc_mads_stencil(signed char st[80][4]);
/* followup's Nelder-Mead (amoeba, followup.c:971-1173) on -F from point with initial steps pc2 */
int orc_amoeba(const orc_search *s, double point[4], const double pc2[4], double tol, int maxeval,
               double out[11]);

#ifdef __cplusplus
}
#endif
#endif
