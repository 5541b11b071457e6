/*
 * pgw_oracle.c - CPU FP64 restatement of the polgraw-allsky network search core
 * (search/network/src-cpu/{jobcore.c,auxi.c,settings.c,init.c}) with the OpenMP
 * structure of search/network/src-openmp/jobcore.c.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the CPU
 * baseline that bench.py times.  Nothing in the product links it.
 *
 * Written from the reference's documented behaviour, not copied: the FFT is
 * this file's own Stockham radix-4/2 (the reference uses FFTW 3.3.4, which is
 * absent here), sin/cos come from glibc (the reference's GNUSINCOS build), and
 * there is no hidden static state (the reference's spline() keeps one).
 */
#define _GNU_SOURCE
#include "pgw_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SQR(x) ((x) * (x))
#define C_OMEGA_R 7.2921151467064e-5 /* settings.h:31 */

/* ------------------------------------------------------------------ FFT -- */

int orc_fft_plan_init(orc_fft_plan *p, int n) {
  if (n < 2 || (n & (n - 1))) return -1;
  p->n = n;
  p->tw = malloc(sizeof(double complex) * (size_t)n);
  p->work = malloc(sizeof(double complex) * (size_t)n);
  if (!p->tw || !p->work) return -4;
  for (int m = 0; m < n; m++) {
    double s, c;
    sincos(2.0 * M_PI * (double)m / (double)n, &s, &c);
    p->tw[m] = c - I * s;
  }
  return 0;
}

void orc_fft_plan_free(orc_fft_plan *p) {
  free(p->tw);
  free(p->work);
  p->tw = p->work = NULL;
}

/* One Stockham pass: radix R (2 or 4), Ns = product of the radices already done.
 * out[(j/Ns)*Ns*R + k + r*Ns] = sum_q in[j + q*m] W_{Ns R}^{k q} W_R^{q r},  k = j mod Ns. */
static void stockham_pass(const orc_fft_plan *p, const double complex *in, double complex *out,
                          int Ns, int R, int dir) {
  const int n = p->n, m = n / R, step = n / (Ns * R);
  for (int jb = 0; jb < m / Ns; jb++) {
    for (int k = 0; k < Ns; k++) {
      const int j = jb * Ns + k;
      const int idx = jb * Ns * R + k;
      if (R == 4) {
        double complex v0 = in[j], v1 = in[j + m], v2 = in[j + 2 * m], v3 = in[j + 3 * m];
        if (k) {
          double complex w1 = p->tw[k * step], w2 = p->tw[2 * k * step], w3 = p->tw[3 * k * step];
          if (dir > 0) { w1 = conj(w1); w2 = conj(w2); w3 = conj(w3); }
          v1 *= w1; v2 *= w2; v3 *= w3;
        }
        double complex t0 = v0 + v2, t1 = v0 - v2, t2 = v1 + v3, d = v1 - v3;
        /* d * W_4 with W_4 = -i (forward) or +i (backward) */
        double complex t3 = (dir < 0) ? (cimag(d) - I * creal(d)) : (-cimag(d) + I * creal(d));
        out[idx] = t0 + t2;
        out[idx + Ns] = t1 + t3;
        out[idx + 2 * Ns] = t0 - t2;
        out[idx + 3 * Ns] = t1 - t3;
      } else {
        double complex v0 = in[j], v1 = in[j + m];
        if (k) {
          double complex w1 = p->tw[k * step];
          if (dir > 0) w1 = conj(w1);
          v1 *= w1;
        }
        out[idx] = v0 + v1;
        out[idx + Ns] = v0 - v1;
      }
    }
  }
}

void orc_fft_exec(const orc_fft_plan *p, double complex *x, int dir, double complex *work) {
  const int n = p->n;
  double complex *a = x, *b = work ? work : p->work;
  for (int Ns = 1; Ns < n;) {
    int R = (n / Ns >= 4) ? 4 : 2;
    stockham_pass(p, a, b, Ns, R, dir);
    double complex *t = a; a = b; b = t;
    Ns *= R;
  }
  if (a != x) memcpy(x, a, sizeof(double complex) * (size_t)n);
}

/* ------------------------------------------------------- small helpers -- */

/* lin2ast, auxi.c:15-24 */
void orc_lin2ast(double be1, double be2, int pm, double sepsm, double cepsm,
                 double *sinal, double *cosal, double *sindel, double *cosdel) {
  *sindel = be1 * sepsm - (2 * pm - 3) * sqrt(1. - SQR(be1) - SQR(be2)) * cepsm;
  *cosdel = sqrt(1. - SQR(*sindel));
  *sinal = (be1 - sepsm * (*sindel)) / (cepsm * (*cosdel));
  *cosal = be2 / (*cosdel);
}

/* rogcvir, settings.c:253-275 (JKS98 Eqs. 12-13 coefficients) */
void orc_rogcvir(double ephi, double egam, double c[9]) {
  c[0] = .25 * sin(2. * egam) * (1 + SQR(sin(ephi)));
  c[1] = -.5 * cos(2. * egam) * sin(ephi);
  c[2] = .5 * sin(2. * egam) * sin(2. * ephi);
  c[3] = -cos(2. * egam) * cos(ephi);
  c[4] = .75 * sin(2. * egam) * SQR(cos(ephi));
  c[5] = cos(2. * egam) * sin(ephi);
  c[6] = .5 * sin(2. * egam) * (1. + SQR(sin(ephi)));
  c[7] = cos(2. * egam) * cos(ephi);
  c[8] = .5 * sin(2. * egam) * sin(2. * ephi);
}

/* modvir, settings.c:281-325 */
void orc_modvir(double sinal, double cosal, double sindel, double cosdel,
                double sphir, double cphir, const double c[9], int Np,
                const double *cosmodf, const double *sinmodf, double *aa, double *bb) {
  const double cosalfr = cosal * cphir + sinal * sphir;
  const double sinalfr = sinal * cphir - cosal * sphir;
  const double c2d = SQR(cosdel), c2sd = sindel * cosdel;
  for (int t = 0; t < Np; t++) {
    double cc = cosalfr * cosmodf[t] + sinalfr * sinmodf[t];
    double ss = sinalfr * cosmodf[t] - cosalfr * sinmodf[t];
    double c2s = 2. * SQR(cc), cs = cc * ss;
    aa[t] = c[0] * (2. - c2d) * c2s + c[1] * (2. - c2d) * 2. * cs + c[2] * c2sd * cc +
            c[3] * c2sd * ss - c[0] * (2. - c2d) + c[4] * c2d;
    bb[t] = c[5] * sindel * c2s + c[6] * sindel * 2. * cs + c[7] * cosdel * cc +
            c[8] * cosdel * ss - c[5] * sindel;
  }
}

/* Natural cubic spline on unit-spaced knots, auxi.c:55-81 (Thomas recursion).
 * u is caller scratch of n-1 entries (the reference keeps a static one). */
void orc_spline(const double complex *y, int n, double complex *y2, double complex *u) {
  y2[0] = u[0] = 0.;
  for (int i = 1; i < n - 1; ++i) {
    double complex invp = 2. / (y2[i - 1] + 4.);
    y2[i] = -.5 * invp;
    u[i] = y[i - 1] - 2. * y[i] + y[i + 1];
    u[i] = (-.5 * u[i - 1] + 3. * u[i]) * invp;
  }
  y2[n - 1] = 0.;
  for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

/* auxi.c:83-96; outside [0, n-1] the value is exactly 0.  At x == n-1 the
 * reference reads ya[n] and y2a[n] (past the knots) times a zero weight; here
 * those terms are taken as 0. */
double complex orc_splint(const double complex *ya, const double complex *y2a, int n, double x) {
  if (x < 0 || x > n - 1) return 0.;
  int klo = (int)floor(x), khi = klo + 1;
  double a = khi - x, b = x - klo;
  double complex yhi = (khi < n) ? ya[khi] : 0., y2hi = (khi < n) ? y2a[khi] : 0.;
  return a * ya[klo] + b * yhi + ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2hi) / 6.0;
}

/* auxi.c:98-122: spline on the first interpftpad*N samples, evaluated at interpftpad*(i - shftf[i]) */
void orc_splintpad(const double complex *ya, const double *shftf, int N, int interpftpad,
                   double complex *out) {
  const int n = interpftpad * N;
  double complex *y2 = malloc(sizeof(double complex) * (size_t)n);
  double complex *u = malloc(sizeof(double complex) * (size_t)n);
  orc_spline(ya, n, y2, u);
  for (int i = 0; i < N; ++i) {
    double x = interpftpad * (i - shftf[i]);
    out[i] = orc_splint(ya, y2, n, x);
  }
  free(u);
  free(y2);
}

/* auxi.c:124-139 */
double orc_var(const double *x, int n) {
  double mean = 0., variance = 0.;
  for (int i = 0; i < n; i++) mean += x[i];
  mean /= n;
  for (int i = 0; i < n; i++) variance += SQR(x[i] - mean);
  variance /= (n - 1);
  return variance;
}

/* FStat, auxi.c:214-240: floor(nfft/nav) blocks normalised by mean/2; the tail stays raw */
void orc_fstat_norm(double *F, int nfft, int nav) {
  double *fr = F;
  for (int j = 0; j < nfft / nav; j++) {
    double mu = 0.;
    for (int i = 0; i < nav; i++) mu += *fr++;
    mu /= 2. * nav;
    fr -= nav;
    for (int i = 0; i < nav; i++) *fr++ /= mu;
  }
}

/* search_settings (settings.c:16-81) with N given, plus the sizes set in plan_fftw (init.c:835-844) */
int orc_settings(double fpo, double dt, int N, int fftpad, pgw_settings *sett) {
  if (!sett || N < 16 || dt <= 0 || fftpad < 1) return -1;
  memset(sett, 0, sizeof(*sett));
  sett->fpo = fpo;
  sett->dt = dt;
  sett->B = 0.5 / dt;
  sett->oms = 2. * M_PI * fpo * dt;
  sett->omr = C_OMEGA_R * dt;
  sett->N = N;
  sett->nfft = 1 << (int)ceil(log(N) / log(2.));
  sett->Smax = 2. * M_PI * 0.5e-8 * dt * dt;
  sett->Smin = 2. * M_PI * 0.5e-9 * dt * dt;
  sett->nd = 2;
  sett->interpftpad = 2;
  sett->fftpad = fftpad;
  sett->nmin = sett->fftpad * PGW_NAVB * sett->B;
  sett->nmax = (sett->nfft / 2 - PGW_NAVB * sett->B) * sett->fftpad;
  sett->Ninterp = sett->interpftpad * sett->nfft;
  sett->nfftf = sett->fftpad * sett->nfft;
  sett->sepsm = sin(0.40909280422232897);
  sett->cepsm = cos(0.40909280422232897);
  return 0;
}

/* ------------------------------------------------------------- search -- */

typedef struct orc_det {
  double *xDat, *DetSSB, sig2, sphir, cphir, amod[9];
  double *aa, *bb, *shft, *shftf, *tmp1;
  double complex *xDatma, *xDatmb;
} orc_det;

typedef struct orc_thr {
  double complex *xa, *xb, *work;
  double *F;
} orc_thr;

struct orc_search {
  pgw_settings sett;
  int nifo, nthreads, arr_len;
  orc_det det[PGW_MAX_DETECTORS];
  double *t2, *cosmodf, *sinmodf;
  orc_fft_plan plan, pl_int, pl_inv;
  orc_thr *thr;
  /* sky-point state shared with the template loop */
  double aa, bb, het0, sgnlt2, sgnlt3;
};

void orc_destroy(orc_search *s) {
  if (!s) return;
  for (int n = 0; n < s->nifo; n++) {
    orc_det *d = &s->det[n];
    free(d->xDat); free(d->DetSSB); free(d->aa); free(d->bb); free(d->shft);
    free(d->shftf); free(d->tmp1); free(d->xDatma); free(d->xDatmb);
  }
  free(s->t2); free(s->cosmodf); free(s->sinmodf);
  orc_fft_plan_free(&s->plan); orc_fft_plan_free(&s->pl_int); orc_fft_plan_free(&s->pl_inv);
  if (s->thr)
    for (int t = 0; t < s->nthreads; t++) {
      free(s->thr[t].xa); free(s->thr[t].xb); free(s->thr[t].work); free(s->thr[t].F);
    }
  free(s->thr);
  free(s);
}

orc_search *orc_create(const pgw_settings *sett, const pgw_detector *ifo, int nifo, int nthreads) {
  if (!sett || !ifo || nifo < 1 || nifo > PGW_MAX_DETECTORS) return NULL;
  orc_search *s = calloc(1, sizeof(*s));
  if (!s) return NULL;
  s->sett = *sett;
  s->sett.nifo = nifo;
  s->nifo = nifo;
  s->nthreads = nthreads < 1 ? 1 : nthreads;
  const int N = sett->N;
  const size_t nd = sizeof(double) * (size_t)N, nc = sizeof(double complex) * (size_t)N;
  for (int n = 0; n < nifo; n++) {
    orc_det *d = &s->det[n];
    d->xDat = malloc(nd); d->DetSSB = malloc(3 * nd);
    d->aa = malloc(nd); d->bb = malloc(nd); d->shft = malloc(nd); d->shftf = malloc(nd);
    d->tmp1 = malloc(nd); d->xDatma = malloc(nc); d->xDatmb = malloc(nc);
    if (!d->xDat || !d->DetSSB || !d->aa || !d->bb || !d->shft || !d->shftf || !d->tmp1 ||
        !d->xDatma || !d->xDatmb) { orc_destroy(s); return NULL; }
    memcpy(d->xDat, ifo[n].xDat, nd);
    memcpy(d->DetSSB, ifo[n].DetSSB, 3 * nd);
    d->sig2 = ifo[n].sig2; d->sphir = ifo[n].sphir; d->cphir = ifo[n].cphir;
    memcpy(d->amod, ifo[n].amod, sizeof(d->amod));
  }
  /* auxiliary arrays, init.c:445-455 */
  s->t2 = malloc(nd); s->cosmodf = malloc(nd); s->sinmodf = malloc(nd);
  if (!s->t2 || !s->cosmodf || !s->sinmodf) { orc_destroy(s); return NULL; }
  for (int i = 0; i < N; i++) {
    double omrt = sett->omr * i;
    s->t2[i] = SQR((double)i);
    s->cosmodf[i] = cos(omrt);
    s->sinmodf[i] = sin(omrt);
  }
  if (orc_fft_plan_init(&s->plan, sett->nfftf) || orc_fft_plan_init(&s->pl_int, sett->nfft) ||
      orc_fft_plan_init(&s->pl_inv, sett->Ninterp)) { orc_destroy(s); return NULL; }
  s->arr_len = sett->nfftf > sett->Ninterp ? sett->nfftf : sett->Ninterp;
  s->thr = calloc((size_t)s->nthreads, sizeof(orc_thr));
  if (!s->thr) { orc_destroy(s); return NULL; }
  for (int t = 0; t < s->nthreads; t++) {
    s->thr[t].xa = malloc(sizeof(double complex) * (size_t)s->arr_len);
    s->thr[t].xb = malloc(sizeof(double complex) * (size_t)s->arr_len);
    s->thr[t].work = malloc(sizeof(double complex) * (size_t)s->arr_len);
    s->thr[t].F = calloc((size_t)(2 * sett->nfft), sizeof(double));
    if (!s->thr[t].xa || !s->thr[t].xb || !s->thr[t].work || !s->thr[t].F) { orc_destroy(s); return NULL; }
  }
  return s;
}

/* FFT resampling + spline, jobcore.c:319-370: zero-pad N -> nfft, forward FFT, move bins
 * [nyqst, nfft) to the top of an Ninterp = interpftpad*nfft spectrum and zero the gap, backward
 * FFT, scale by interpftpad/Ninterp, then splintpad at interpftpad*(i - shftf[i]).
 * xa, xb: scratch of Ninterp; outa/outb may alias ina/inb. */
void orc_resample(const orc_fft_plan *pl_int, const orc_fft_plan *pl_inv, const double complex *ina,
                  const double complex *inb, const double *shftf, int N, int nfft, int interpftpad,
                  double complex *xa, double complex *xb, double complex *work,
                  double complex *outa, double complex *outb) {
  const int Ninterp = interpftpad * nfft, nyqst = nfft / 2 + 1;
  for (int i = 0; i < N; ++i) { xa[i] = ina[i]; xb[i] = inb[i]; }
  for (int i = N; i < nfft; ++i) xa[i] = xb[i] = 0.;
  orc_fft_exec(pl_int, xa, -1, work);
  orc_fft_exec(pl_int, xb, -1, work);
  for (int i = nyqst + Ninterp - nfft, j = nyqst; i < Ninterp; ++i, ++j) xa[i] = xa[j];
  for (int i = nyqst; i < nyqst + Ninterp - nfft; ++i) xa[i] = 0.;
  for (int i = nyqst + Ninterp - nfft, j = nyqst; i < Ninterp; ++i, ++j) xb[i] = xb[j];
  for (int i = nyqst; i < nyqst + Ninterp - nfft; ++i) xb[i] = 0.;
  orc_fft_exec(pl_inv, xa, +1, work);
  orc_fft_exec(pl_inv, xb, +1, work);
  const double ft = (double)interpftpad / Ninterp;
  for (int i = 0; i < Ninterp; ++i) { xa[i] *= ft; xb[i] *= ft; }
  orc_splintpad(xa, shftf, N, interpftpad, outa);
  orc_splintpad(xb, shftf, N, interpftpad, outb);
}

/* Run-max scan of jobcore.c:659-669 over F[nmin..nmax): one (ii, Fc) per maximal run of F > trl,
 * ii the last index holding the run's maximum.  Returns the number of runs (entries beyond cap
 * are counted but not stored). */
int orc_runmax(const double *F, int nmin, int nmax, double trl, int *ii_out, double *F_out, int cap) {
  int cnt = 0;
  for (int i = nmin; i < nmax; i++) {
    double Fc;
    if ((Fc = F[i]) > trl) {
      int ii = i;
      while (++i < nmax && F[i] > trl) {
        if (F[i] >= Fc) { ii = i; Fc = F[i]; }
      }
      if (cnt < cap) { ii_out[cnt] = ii; F_out[cnt] = Fc; }
      cnt++;
    }
  }
  return cnt;
}

/* Sky-point stage, jobcore.c:243-393.  Returns 0 if the point is outside the disk. */
static int sky_stage(orc_search *s, int pm, int mm, int nn) {
  const pgw_settings *sett = &s->sett;
  const double *M = sett->M;
  const int N = sett->N;
  double sinalt, cosalt, sindelt, cosdelt, nSource[3];

  double al1 = nn * M[10] + mm * M[14];
  double al2 = nn * M[11] + mm * M[15];
  if ((SQR(al1) + SQR(al2)) / SQR(sett->oms) > 1.) return 0;

  orc_lin2ast(al1 / sett->oms, al2 / sett->oms, pm, sett->sepsm, sett->cepsm,
              &sinalt, &cosalt, &sindelt, &cosdelt);
  s->sgnlt2 = asin(sindelt);
  s->sgnlt3 = fmod(atan2(sinalt, cosalt) + 2. * M_PI, 2. * M_PI);
  s->het0 = fmod(nn * M[8] + mm * M[12], M[0]);
  double complex *xa = s->thr[0].xa, *xb = s->thr[0].xb, *work = s->thr[0].work;

  for (int n = 0; n < s->nifo; ++n) {
    orc_det *d = &s->det[n];
    orc_modvir(sinalt, cosalt, sindelt, cosdelt, d->sphir, d->cphir, d->amod, N,
               s->cosmodf, s->sinmodf, d->aa, d->bb);
    nSource[0] = cosalt * cosdelt;
    nSource[1] = sinalt * cosdelt;
    nSource[2] = sindelt;
    const double shft1 = nSource[0] * d->DetSSB[0] + nSource[1] * d->DetSSB[1] + nSource[2] * d->DetSSB[2];
    const double het0 = s->het0, oms = sett->oms;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int i = 0; i < N; ++i) {
      d->shft[i] = nSource[0] * d->DetSSB[i * 3] + nSource[1] * d->DetSSB[i * 3 + 1] +
                   nSource[2] * d->DetSSB[i * 3 + 2];
      d->shftf[i] = d->shft[i] - shft1;
      d->tmp1[i] = s->t2[i] + (double)(2 * i) * d->shft[i];
      double sp, cp;
      sincos(het0 * i + oms * d->shft[i], &sp, &cp);
      double complex exph = cp - I * sp;
      d->xDatma[i] = d->xDat[i] * d->aa[i] * exph;
      d->xDatmb[i] = d->xDat[i] * d->bb[i] * exph;
    }
    orc_resample(&s->pl_int, &s->pl_inv, d->xDatma, d->xDatmb, d->shftf, N, sett->nfft,
                 sett->interpftpad, xa, xb, work, d->xDatma, d->xDatmb);
  }
  /* noise weighting, jobcore.c:374-393 */
  double aa = 0., bb = 0.;
  for (int n = 0; n < s->nifo; ++n) {
    orc_det *d = &s->det[n];
    double aatemp = 0., bbtemp = 0.;
    for (int i = 0; i < N; ++i) { aatemp += SQR(d->aa[i]); bbtemp += SQR(d->bb[i]); }
    for (int i = 0; i < N; ++i) { d->xDatma[i] /= d->sig2; d->xDatmb[i] /= d->sig2; }
    aa += aatemp / d->sig2;
    bb += bbtemp / d->sig2;
  }
  s->aa = aa;
  s->bb = bb;
  return 1;
}

void orc_spindown_range(const pgw_settings *sett, int mm, int nn, const pgw_range *s_range,
                        const pgw_opts *opts, int *smin_out, int *smax_out) {
  const double *M = sett->M;
  int smin = s_range->sst, smax = s_range->spndr[1];
  if (!opts->use_range) {
    smin = trunc((sett->Smin - nn * M[9] - mm * M[13]) / M[5]);
    smax = trunc(-(nn * M[9] + mm * M[13] + sett->Smax) / M[5]);
    if (smin > smax) { int t = smin; smin = smax; smax = t; }
  }
  if (opts->s0_flag) smin = smax;
  *smin_out = smin;
  *smax_out = smax;
}

/* Demodulation + F-statistic FFT + F + FStat for one spindown, jobcore.c:455-651.
 * Leaves the normalised F in t->F[nmin..nmax). Returns sgnl0 and sgnlt[1]. */
static void template_stage(orc_search *s, orc_thr *t, int mm, int nn, int ss, const pgw_opts *opts,
                           double *sgnl0_out, double *sgnlt1_out) {
  const pgw_settings *sett = &s->sett;
  const double *M = sett->M;
  const int N = sett->N;
  double complex *xa = t->xa, *xb = t->xb;
  double *F = t->F;
  double sgnlt1 = (opts->s0_flag ? 0. : ss * M[5] + nn * M[9] + mm * M[13]);
  double het1 = fmod(ss * M[4], M[0]);
  if (het1 < 0) het1 += M[0];
  *sgnl0_out = s->het0 + het1;
  *sgnlt1_out = sgnlt1;
  for (int n = 0; n < s->nifo; ++n) {
    const orc_det *d = &s->det[n];
    for (int i = N - 1; i != -1; --i) {
      double sp, cp;
      sincos(het1 * i + sgnlt1 * d->tmp1[i], &sp, &cp);
      double complex exph = cp - I * sp;
      if (n == 0) { xa[i] = d->xDatma[i] * exph; xb[i] = d->xDatmb[i] * exph; }
      else { xa[i] += d->xDatma[i] * exph; xb[i] += d->xDatmb[i] * exph; }
    }
  }
  for (int i = sett->fftpad * sett->nfft - 1; i != N - 1; --i) xa[i] = xb[i] = 0.;
  orc_fft_exec(&s->plan, xa, -1, t->work);
  orc_fft_exec(&s->plan, xb, -1, t->work);
  for (int i = sett->nmin; i < sett->nmax; i++)
    F[i] = (SQR(creal(xa[i])) + SQR(cimag(xa[i]))) / s->aa + (SQR(creal(xb[i])) + SQR(cimag(xb[i]))) / s->bb;
  if (!opts->white_flag) orc_fstat_norm(F + sett->nmin, sett->nmax - sett->nmin, PGW_NAV);
}

typedef struct cand { int ss, ii; double rec[PGW_NPAR]; } cand;

static int cand_cmp(const void *a, const void *b) {
  const cand *x = a, *y = b;
  if (x->ss != y->ss) return x->ss < y->ss ? -1 : 1;
  return (x->ii > y->ii) - (x->ii < y->ii);
}

/* threshold + candidate records, jobcore.c:659-695; appends to *list */
static void threshold(const orc_search *s, const double *F, int ss, double sgnl0, double sgnlt1,
                      const pgw_opts *opts, cand **list, int *cnt, int *cap) {
  const pgw_settings *sett = &s->sett;
  const int maxruns = (sett->nmax - sett->nmin + 1) / 2 + 1;
  int *iis = malloc(sizeof(int) * (size_t)maxruns);
  double *Fcs = malloc(sizeof(double) * (size_t)maxruns);
  const int nr = orc_runmax(F, sett->nmin, sett->nmax, opts->trl, iis, Fcs, maxruns);
  for (int r = 0; r < nr; r++) {
    const int ii = iis[r];
    const double Fc = Fcs[r];
    double sgnlt[PGW_NPAR];
    sgnlt[0] = 2. * M_PI * ii / ((double)sett->fftpad * sett->nfft) + sgnl0;
    sgnlt[1] = sgnlt1;
    sgnlt[2] = s->sgnlt2;
    sgnlt[3] = s->sgnlt3;
    sgnlt[4] = sqrt(2. * (Fc - sett->nd));
    int veto = 0;
    for (int k = 0; k < sett->numlines_band; k++)
      if (sgnlt[0] >= sett->lines[k][0] && sgnlt[0] <= sett->lines[k][1]) { veto = 1; break; }
    if (!veto) {
      if (*cnt == *cap) {
        *cap = *cap ? 2 * *cap : 64;
        *list = realloc(*list, sizeof(cand) * (size_t)*cap);
      }
      cand *c = &(*list)[(*cnt)++];
      c->ss = ss; c->ii = ii;
      memcpy(c->rec, sgnlt, sizeof(sgnlt));
    }
  }
  free(iis);
  free(Fcs);
}

int orc_job_core(orc_search *s, int pm, int mm, int nn, const pgw_range *s_range,
                 const pgw_opts *opts, double *F, int *sgnlc, double *sgnlv, int sgnlv_cap,
                 int *FNum) {
  if (!sky_stage(s, pm, mm, nn)) return 0;
  int smin, smax;
  orc_spindown_range(&s->sett, mm, nn, s_range, opts, &smin, &smax);
  const int nss = smax - smin + 1;
  cand *all = NULL;
  int nall = 0, capall = 0;
  if (nss > 0) {
#pragma omp parallel num_threads(s->nthreads)
    {
      int tid = 0;
#ifdef _OPENMP
      tid = omp_get_thread_num();
#endif
      orc_thr *t = &s->thr[tid];
      cand *mine = NULL;
      int nmine = 0, capmine = 0;
#pragma omp for schedule(static, 4)
      for (int ss = smin; ss <= smax; ++ss) {
        double sgnl0, sgnlt1;
        template_stage(s, t, mm, nn, ss, opts, &sgnl0, &sgnlt1);
        threshold(s, t->F, ss, sgnl0, sgnlt1, opts, &mine, &nmine, &capmine);
        if (F && ss == smax) memcpy(F, t->F, sizeof(double) * 2 * (size_t)s->sett.nfft);
      }
#pragma omp critical
      {
        if (nmine) {
          all = realloc(all, sizeof(cand) * (size_t)(nall + nmine));
          memcpy(all + nall, mine, sizeof(cand) * (size_t)nmine);
          nall += nmine;
        }
      }
      free(mine);
    }
    (void)capall;
    *FNum += nss;
  }
  qsort(all, (size_t)nall, sizeof(cand), cand_cmp);
  int status = 0;
  if (*sgnlc + nall > sgnlv_cap) {
    *sgnlc += nall;
    status = -2;
  } else {
    for (int c = 0; c < nall; c++)
      memcpy(sgnlv + (size_t)PGW_NPAR * (size_t)(*sgnlc + c), all[c].rec, sizeof(double) * PGW_NPAR);
    *sgnlc += nall;
  }
  free(all);
  return status;
}

int orc_fstat_one(orc_search *s, int pm, int mm, int nn, int ss, const pgw_opts *opts, double *F) {
  if (!sky_stage(s, pm, mm, nn)) return 1;
  double sgnl0, sgnlt1;
  template_stage(s, &s->thr[0], mm, nn, ss, opts, &sgnl0, &sgnlt1);
  memcpy(F, s->thr[0].F + s->sett.nmin, sizeof(double) * (size_t)(s->sett.nmax - s->sett.nmin));
  return 0;
}

int orc_resampled(orc_search *s, int pm, int mm, int nn, int det, double complex *xa, double complex *xb) {
  if (det < 0 || det >= s->nifo) return -1;
  if (!sky_stage(s, pm, mm, nn)) return 1;
  memcpy(xa, s->det[det].xDatma, sizeof(double complex) * (size_t)s->sett.N);
  memcpy(xb, s->det[det].xDatmb, sizeof(double complex) * (size_t)s->sett.N);
  return 0;
}

/* ------------------------------------------------------------ Fstatnet (followup) */

/* modvir at one sample plus the derivatives of a and b with respect to delta and alpha
 * (followup settings.c:275-318 differentiated: dc/dalpha = -s, ds/dalpha = c,
 * d cos^2(delta)/d delta = -2 sin cos, d(sin cos)/d delta = cos 2 delta). */
static void modvir_d(const double c[9], double sindel, double cosdel, double cosalfr, double sinalfr,
                     double cm, double sm, double *a, double *b, double *da_d, double *da_a,
                     double *db_d, double *db_a) {
  const double c2d = SQR(cosdel), c2sd = sindel * cosdel, cos2d = SQR(cosdel) - SQR(sindel);
  const double cc = cosalfr * cm + sinalfr * sm, ss = sinalfr * cm - cosalfr * sm;
  const double c2s = 2. * SQR(cc), cs = cc * ss;
  *a = c[0] * (2. - c2d) * c2s + c[1] * (2. - c2d) * 2. * cs + c[2] * c2sd * cc + c[3] * c2sd * ss -
       c[0] * (2. - c2d) + c[4] * c2d;
  *b = c[5] * sindel * c2s + c[6] * sindel * 2. * cs + c[7] * cosdel * cc + c[8] * cosdel * ss -
       c[5] * sindel;
  /* d/d delta: c2d' = -2 c2sd, c2sd' = cos 2 delta */
  *da_d = c[0] * 2. * c2sd * c2s + c[1] * 2. * c2sd * 2. * cs + c[2] * cos2d * cc + c[3] * cos2d * ss -
          c[0] * 2. * c2sd - 2. * c[4] * c2sd;
  *db_d = c[5] * cosdel * c2s + c[6] * cosdel * 2. * cs - c[7] * sindel * cc - c[8] * sindel * ss -
          c[5] * cosdel;
  /* d/d alpha: c' = -s, s' = c, (c^2)' = -2 c s, (c s)' = c^2 - s^2 */
  *da_a = c[0] * (2. - c2d) * (-4. * cs) + c[1] * (2. - c2d) * 2. * (SQR(cc) - SQR(ss)) - c[2] * c2sd * ss +
          c[3] * c2sd * cc;
  *db_a = c[5] * sindel * (-4. * cs) + c[6] * sindel * 2. * (SQR(cc) - SQR(ss)) - c[7] * cosdel * ss +
          c[8] * cosdel * cc;
}

/* Fstatnet, followup.c:413-534: the F-statistic of the whole network at continuous parameters
 * pt = (omega, omegadot, delta, alpha), with the detector-response arrays from modvir at
 * (delta, alpha) as the callers build them (followup.c:676-685).  out[11] as the reference:
 * [0..3] amplitude estimates, [4] sqrt(2(F-2)), [5] -F, [6..9] the point, [10] amplitude SNR.
 * grad (optional): dF/d(omega, omegadot, delta, alpha) of F = -out[5], accumulated in the same
 * sample order (analytic; checked against central differences of out[5]).
 * Sums run per detector in sample order, as the reference's. */
int orc_fstatnet(const orc_search *s, const double pt[4], double out[11], double grad[4]) {
  const int N = s->sett.N, nifo = s->nifo;
  const double oms = s->sett.oms;
  const double sinalt = sin(pt[3]), cosalt = cos(pt[3]), sindelt = sin(pt[2]), cosdelt = cos(pt[2]);
  const double nS[3] = {cosalt * cosdelt, sinalt * cosdelt, sindelt};
  const double nSd[3] = {-cosalt * sindelt, -sinalt * sindelt, cosdelt};  /* d nS / d delta */
  const double nSa[3] = {-sinalt * cosdelt, cosalt * cosdelt, 0.};         /* d nS / d alpha */
  double complex *xa = malloc(sizeof(double complex) * (size_t)N);
  double complex *xb = malloc(sizeof(double complex) * (size_t)N);
  double *sa = malloc(sizeof(double) * (size_t)N), *sb = malloc(sizeof(double) * (size_t)N);
  if (!xa || !xb || !sa || !sb) { free(xa); free(xb); free(sa); free(sb); return -1; }
  double complex xasum = 0, xbsum = 0, dxa[4] = {0, 0, 0, 0}, dxb[4] = {0, 0, 0, 0};
  double aa = 0., bb = 0., daa[2] = {0, 0}, dbb[2] = {0, 0};
  for (int n = 0; n < nifo; ++n) {
    const orc_det *d = &s->det[n];
    const double cosalfr = cosalt * d->cphir + sinalt * d->sphir;
    const double sinalfr = sinalt * d->cphir - cosalt * d->sphir;
    double aatemp = 0., bbtemp = 0., daat[2] = {0, 0}, dbbt[2] = {0, 0};
    double complex dxat[4] = {0, 0, 0, 0}, dxbt[4] = {0, 0, 0, 0};
    for (int i = 0; i < N; ++i) {
      const double *R = d->DetSSB + 3 * (size_t)i;
      const double shft = nS[0] * R[0] + nS[1] * R[1] + nS[2] * R[2];
      const double phase = pt[0] * (double)(i + shft) + (pt[1] * i * i) + ((oms + 2. * pt[1] * i) * shft);
      const double complex exph = cos(phase) - I * sin(phase);
      double a, b, dad, daa_, dbd, dba;
      modvir_d(d->amod, sindelt, cosdelt, cosalfr, sinalfr, s->cosmodf[i], s->sinmodf[i], &a, &b, &dad,
               &daa_, &dbd, &dba);
      sa[i] = a;
      sb[i] = b;
      xa[i] = d->xDat[i] * a * exph;
      xb[i] = d->xDat[i] * b * exph;
      if (grad) {
        /* phase derivatives: i + shft, i^2 + 2 i shft, (omega + oms + 2 omegadot i) dshft */
        const double K = pt[0] + oms + 2. * pt[1] * i;
        const double ph[4] = {i + shft, (double)i * i + 2. * i * shft,
                              K * (nSd[0] * R[0] + nSd[1] * R[1] + nSd[2] * R[2]),
                              K * (nSa[0] * R[0] + nSa[1] * R[1] + nSa[2] * R[2])};
        const double dav[4] = {0, 0, dad, daa_}, dbv[4] = {0, 0, dbd, dba};
        const double complex xe = d->xDat[i] * exph / d->sig2;
        for (int k = 0; k < 4; k++) {
          dxat[k] += xe * (dav[k] - I * a * ph[k]);
          dxbt[k] += xe * (dbv[k] - I * b * ph[k]);
        }
        daat[0] += 2. * a * dad; daat[1] += 2. * a * daa_;
        dbbt[0] += 2. * b * dbd; dbbt[1] += 2. * b * dba;
      }
    }
    for (int i = 0; i < N; i++) {
      aatemp += SQR(sa[i]);
      bbtemp += SQR(sb[i]);
    }
    for (int i = 0; i < N; ++i) {
      xa[i] /= d->sig2;
      xb[i] /= d->sig2;
      xasum += xa[i];
      xbsum += xb[i];
    }
    aa += aatemp / d->sig2;
    bb += bbtemp / d->sig2;
    for (int k = 0; k < 4; k++) { dxa[k] += dxat[k]; dxb[k] += dxbt[k]; }
    for (int k = 0; k < 2; k++) { daa[k] += daat[k] / d->sig2; dbb[k] += dbbt[k] / d->sig2; }
  }
  const double Fa = SQR(creal(xasum)) + SQR(cimag(xasum)), Fb = SQR(creal(xbsum)) + SQR(cimag(xbsum));
  out[5] = -((Fa / aa) + (Fb / bb));
  out[0] = 2 * creal(xasum) / aa;
  out[1] = 2 * creal(xbsum) / bb;
  out[2] = -2 * cimag(xasum) / aa;
  out[3] = -2 * cimag(xbsum) / bb;
  out[4] = sqrt(2 * (-out[5] - 2));
  for (int k = 0; k < 4; k++) out[6 + k] = pt[k];
  out[10] = sqrt(SQR(2 * creal(xasum)) + SQR(2 * creal(xbsum)) + SQR(2 * cimag(xasum)) + SQR(2 * cimag(xbsum)));
  if (grad) {
    for (int k = 0; k < 4; k++) {
      const double dA = k >= 2 ? daa[k - 2] : 0., dB = k >= 2 ? dbb[k - 2] : 0.;
      grad[k] = 2. * creal(conj(xasum) * dxa[k]) / aa - Fa * dA / SQR(aa) + 2. * creal(conj(xbsum) * dxb[k]) / bb -
                Fb * dB / SQR(bb);
    }
  }
  free(xa); free(xb); free(sa); free(sb);
  return 0;
}

int orc_fstatnet_batch(const orc_search *s, const double *pts, int n, double *out, double *grad) {
  int st = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(s->nthreads) reduction(| : st)
  for (int p = 0; p < n; p++) st |= orc_fstatnet(s, pts + 4 * (size_t)p, out + 11 * (size_t)p, grad ? grad + 4 * (size_t)p : NULL);
  return st;
}

/* Nelder-Mead on -F as followup's amoeba (followup.c:971-1173): initial simplex point + pc2[i] e_i,
 * reflection (-1), expansion (-2) from the same centroid and line, contraction (0.5), else
 * shrink toward the best vertex; stop when |f_hi - f_lo| < (|f_hi| + |f_lo|) tol + 1e-14.
 * point is replaced by the best vertex, out[11] receives its Fstatnet values.  Returns the number
 * of Fstatnet evaluations, or -1 if maxeval is reached first (the reference has no such limit). */
/* update_simplex, followup.c:1055-1087: vertex ihi moves to mid + scale line if that lowers -F */
static int try_vertex(const orc_search *s, double sx[5][4], double fx[5][11], int ihi, const double mid[4],
                      const double line[4], double scale, int *nev) {
  double next[4], f2[11];
  const double fmax = fx[ihi][5];
  for (int j = 0; j < 4; j++) next[j] = mid[j] + scale * line[j];
  orc_fstatnet(s, next, f2, NULL);
  (*nev)++;
  if (!(f2[5] < fmax)) return 0;
  memcpy(sx[ihi], next, sizeof next);
  memcpy(fx[ihi], f2, sizeof f2);
  return 1;
}

int orc_amoeba(const orc_search *s, double point[4], const double pc2[4], double tol, int maxeval,
               double out[11]) {
  enum { D = 4 };
  double sx[D + 1][D], fx[D + 1][11], mid[D], line[D];
  int nev = 0;
  for (int i = 0; i < D + 1; i++)
    for (int j = 0; j < D; j++) sx[i][j] = point[j];
  for (int i = 0; i < D; i++) sx[i][i] += pc2[i];
  for (int i = 0; i < D + 1; i++, nev++) orc_fstatnet(s, sx[i], fx[i], NULL);
  int ilo = 0;
  for (;;) {
    int ihi, inhi;
    if (fx[0][5] > fx[1][5]) { ihi = 0; ilo = inhi = 1; }
    else { ihi = 1; ilo = inhi = 0; }
    for (int i = 2; i < D + 1; i++) {
      if (fx[i][5] <= fx[ilo][5]) ilo = i;
      else if (fx[i][5] > fx[ihi][5]) { inhi = ihi; ihi = i; }
      else if (fx[i][5] > fx[inhi][5]) inhi = i;
    }
    for (int j = 0; j < D; j++) mid[j] = 0.;
    for (int i = 0; i < D + 1; i++)
      if (i != ihi)
        for (int j = 0; j < D; j++) mid[j] += sx[i][j];
    for (int j = 0; j < D; j++) { mid[j] /= D; line[j] = sx[ihi][j] - mid[j]; }
    const double fmax = fx[ihi][5], fmin = fx[ilo][5];
    if (fabs(fmax - fmin) < (fabs(fmax) + fabs(fmin)) * tol + 1e-14) break;
    if (nev >= maxeval) return -1;
    /* reflection; then expansion if the reflected vertex is the new best, else a contraction
     * when it is still the worst, and a shrink if that contraction fails (followup.c:1147-1156) */
    try_vertex(s, sx, fx, ihi, mid, line, -1.0, &nev);
    if (fx[ihi][5] < fx[ilo][5]) {
      try_vertex(s, sx, fx, ihi, mid, line, -2.0, &nev);
    } else if (fx[ihi][5] > fx[inhi][5]) {
      if (!try_vertex(s, sx, fx, ihi, mid, line, 0.5, &nev)) {
        for (int i = 0; i < D + 1; i++) {
          if (i == ilo) continue;
          for (int j = 0; j < D; j++) sx[i][j] = (sx[ilo][j] + sx[i][j]) * 0.5;
          orc_fstatnet(s, sx[i], fx[i], NULL);
          nev++;
        }
      }
    }
  }
  for (int j = 0; j < D; j++) point[j] = sx[ilo][j];
  memcpy(out, fx[ilo], sizeof(double) * 11);
  return nev;
}

/* The 80 points of the 3^4 mesh around the centre as sign vectors (+1, 0, -1) x the mesh sizes of
 * (omega, omegadot, delta, alpha), in the order invertedMADS fills array[1..80]
 * (followup.c:573-672): the 16 vertices (every parameter moved; omega slowest, + before -), the
 * 32 edges (one parameter kept, in the order omega, omegadot, delta, alpha; the other three in the
 * vertices' order), the 24 faces (two kept) and the 8 cells (one moved: omegadot, alpha, delta,
 * omega).  The order decides only exact ties (the first strict minimum wins here; the reference's
 * OpenMP loop leaves them to its schedule). */
static const signed char mads_faces[24][4] = {
  {0, 0, 1, 1}, {0, 0, 1, -1}, {0, 0, -1, 1}, {0, 0, -1, -1},
  {0, 1, 0, 1}, {0, 1, 0, -1}, {0, 1, 1, 0}, {0, 1, -1, 0},
  {0, -1, 0, 1}, {0, -1, 0, -1}, {0, -1, 1, 0}, {0, -1, -1, 0},
  {1, 0, 0, 1}, {1, 0, 0, -1}, {1, 0, 1, 0}, {1, 0, -1, 0}, {1, 1, 0, 0}, {1, -1, 0, 0},
  {-1, 0, 0, 1}, {-1, 0, 0, -1}, {-1, 0, 1, 0}, {-1, 0, -1, 0}, {-1, 1, 0, 0}, {-1, -1, 0, 0}};
static const signed char mads_cells[8][4] = {
  {0, 1, 0, 0}, {0, -1, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, -1}, {0, 0, 1, 0}, {0, 0, -1, 0}, {1, 0, 0, 0}, {-1, 0, 0, 0}};

void orc_mads_stencil(signed char st[80][4]) {
  int j = 0;
  for (int v = 0; v < 16; v++, j++)          /* bit 3 omega ... bit 0 alpha; 0 = +, 1 = - */
    for (int k = 0; k < 4; k++) st[j][k] = (v >> (3 - k)) & 1 ? -1 : 1;
  for (int z = 0; z < 4; z++)                /* edges: parameter z kept */
    for (int v = 0; v < 8; v++, j++) {
      int b = 2;
      for (int k = 0; k < 4; k++) st[j][k] = k == z ? 0 : ((v >> b--) & 1 ? -1 : 1);
    }
  for (int f = 0; f < 24; f++, j++) memcpy(st[j], mads_faces[f], 4);
  for (int c = 0; c < 8; c++, j++) memcpy(st[j], mads_cells[c], 4);
}

/* invertedMADS, followup.c:537-754, from point with final mesh sizes p2 over a segment of nod
 * days: mesh sizes start at (1 - scale) p2 with scale = 1 - 0.06 / nod; each round evaluates the
 * centre and its 80 mesh points; a point with strictly lower -F than the centre (the first such
 * minimum in mesh order) becomes the centre and the mesh restarts, otherwise every mesh size grows
 * by (1 - scale) p2; it ends when every size exceeds its p2, or after `limit` rounds.  point and
 * out[11] receive the final centre and its Fstatnet values; returns the rounds. */
int orc_mads(const orc_search *s, double point[4], const double p2[4], double nod, int limit, double out[11]) {
  signed char st[80][4];
  orc_mads_stencil(st);
  const double scale = 1.0 - (0.01 * 6.0 / nod);
  double param[4], start[4], centre[4], mesh[80][4], res[80][11], cur[11];
  for (int r = 0; r < 4; r++) start[r] = param[r] = (1.0 - scale) * p2[r];
  memcpy(centre, point, sizeof centre);
  int rounds = 0;
  while (param[0] <= p2[0] || param[1] <= p2[1] || param[2] <= p2[2] || param[3] <= p2[3]) {
    rounds++;
    orc_fstatnet(s, centre, cur, NULL);
    for (int j = 0; j < 80; j++)
      for (int k = 0; k < 4; k++) mesh[j][k] = centre[k] + st[j][k] * param[k];
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int j = 0; j < 80; j++) orc_fstatnet(s, mesh[j], res[j], NULL);
    int best = -1;
    double fbest = cur[5];
    for (int j = 0; j < 80; j++)
      if (res[j][5] < fbest) { fbest = res[j][5]; best = j; }
    if (best < 0) {
      for (int r = 0; r < 4; r++) param[r] += (1.0 - scale) * p2[r];
    } else {
      memcpy(centre, mesh[best], sizeof centre);
      for (int r = 0; r < 4; r++) param[r] = start[r];
    }
    if (rounds >= limit) break;
  }
  orc_fstatnet(s, centre, out, NULL);
  memcpy(point, centre, sizeof centre);
  return rounds;
}

