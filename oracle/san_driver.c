/*
 * san_driver.c - a small standalone run of the CPU oracle for the sanitizer builds (test
 * infrastructure, CPU only): one 2-detector segment of N = 6000 seeded samples, its OpenMP
 * template loop over 40 spindowns of one sky point on 4 threads (job_core), a batch of Fstatnet
 * points (OpenMP over points) and one mesh search.  Values are not checked here; the oracle tests
 * check them.  Built and run by `make -C oracle sanitize`.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pgw_oracle.h"

static unsigned long long rs = 88172645463325252ull;
static double urand(void) {  /* xorshift64, uniform in [0, 1) */
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return (double)(rs >> 11) / 9007199254740992.0;
}

int main(void) {
  enum { N = 6000, NIFO = 2 };
  pgw_settings S;
  if (orc_settings(171.3, 2.0, N, 1, &S)) return 1;
  /* grid matrix with the zero pattern of the reference's grid files, steps scaled to N */
  memset(S.M, 0, sizeof S.M);
  const double r = 86164.0 / N;
  S.M[0] = 2 * M_PI / (S.fftpad * S.nfft);
  S.M[4] = 2.2268 * S.M[0];
  S.M[8] = 181.77 * S.M[0];
  S.M[12] = -215.90 * S.M[0];
  S.M[5] = -1.07995468e-09 * r * r;
  S.M[9] = 6.07898667e-10 * r * r;
  S.M[13] = 1.30440156e-09 * r * r;
  S.M[10] = -8.96932636e+01 * r;
  S.M[14] = 8.72851497e+01 * r;
  S.M[15] = 1.11152781e+02 * r;
  S.sepsm = sin(0.40909280422232897);
  S.cepsm = cos(0.40909280422232897);
  S.nifo = NIFO;
  static double xd[NIFO][N], ssb[NIFO][3 * N];
  pgw_detector ifo[NIFO];
  for (int d = 0; d < NIFO; d++) {
    for (int i = 0; i < N; i++) {
      const double u1 = urand() + 1e-300, u2 = urand();
      xd[d][i] = sqrt(-2 * log(u1)) * cos(2 * M_PI * u2);
      if (i % 997 < 40) xd[d][i] = 0.0;  /* gaps */
      const double t = 2.0 * i;
      ssb[d][3 * i] = 249.5 * cos(1.0 + 2e-7 * t);
      ssb[d][3 * i + 1] = 229.0 * sin(1.0 + 2e-7 * t);
      ssb[d][3 * i + 2] = 99.2 * sin(1.0 + 2e-7 * t);
    }
    memset(&ifo[d], 0, sizeof ifo[d]);
    ifo[d].xDat = xd[d];
    ifo[d].DetSSB = ssb[d];
    ifo[d].sig2 = orc_var(xd[d], N);
    ifo[d].sphir = sin(0.3 + d);
    ifo[d].cphir = cos(0.3 + d);
    orc_rogcvir(0.8 - 0.2 * d, 2.9 - 0.5 * d, ifo[d].amod);
  }
  orc_search *s = orc_create(&S, ifo, NIFO, 4);
  if (!s) return 2;
  pgw_range rg;
  memset(&rg, 0, sizeof rg);
  rg.pmr[0] = rg.pmr[1] = 1;
  rg.spndr[0] = -20;
  rg.spndr[1] = 19;
  rg.sst = -20;
  pgw_opts o = {5.0, 0, 0, 1};
  static double sgnlv[5 * 100000];
  int sgnlc = 0, FNum = 0;
  const int st = orc_job_core(s, 1, 0, 0, &rg, &o, NULL, &sgnlc, sgnlv, 100000, &FNum);
  double pts[4 * 16], out[11 * 16], grad[4 * 16];
  for (int p = 0; p < 16; p++) {
    pts[4 * p] = 0.3 + 0.01 * p;
    pts[4 * p + 1] = -1e-9;
    pts[4 * p + 2] = 0.2;
    pts[4 * p + 3] = 1.1;
  }
  const int st2 = orc_fstatnet_batch(s, pts, 16, out, grad);
  double x[4] = {0.31, -1e-9, 0.2, 1.1}, p2[4] = {1e-4, 1e-10, 1e-2, 1e-2}, res[11];
  const int rounds = orc_mads(s, x, p2, 0.1, 20, res);
  printf("job_core status %d, %d templates, %d records; fstatnet status %d; mads %d rounds\n", st, FNum, sgnlc, st2,
         rounds);
  orc_destroy(s);
  return st || st2;
}
