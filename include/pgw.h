/*
 * pgw.h - C ABI of the MI355X coherent F-statistic search core.
 *
 * This library is a drop-in for the sky-position / spindown loop of the
 * polgraw-allsky network search (`gwsearch`): everything `job_core()` does for
 * one sky point (amplitude modulation, barycentric resampling, spindown
 * demodulation, zero-padded F-statistic FFT, FStat normalisation, threshold and
 * candidate emission) runs as HIP kernels on one GPU.  The host keeps the
 * reference's call order main -> search -> job_core; only the body of
 * job_core moves to the device.
 *
 * All signatures use plain C types.  Every function returns an int status
 * (PGW_OK = 0, negative on error) and never calls exit().  Units and
 * conventions are the reference's: frequencies in rad/sample, spindowns in
 * rad/sample^2, DetSSB in samples (light travel time / dt), forward FFT sign
 * e^{-2 pi i jk/n}, unnormalised.
 *
 * Threading: one host thread per handle; distinct handles are independent.
 */
#ifndef PGW_H
#define PGW_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGW_NPAR 5              /* record fields, search/network/src-cpu/settings.h:7 (NPAR) */
#define PGW_NAV 4096            /* FStat block length, search/network/src-cpu/auxi.h:13 (NAVFSTAT) */
#define PGW_NAVB 4096           /* band-edge constant, search/network/src-cpu/settings.h:12 (NAV) */
#define PGW_MAX_DETECTORS 8     /* search/network/src-cpu/struct.h:7 */
#define PGW_MAXL 2048           /* max known lines, search/network/src-cpu/struct.h:13 */

/* status codes */
#define PGW_OK 0
#define PGW_ERR_ARG (-1)        /* invalid argument or settings */
#define PGW_ERR_CAPACITY (-2)   /* caller's candidate buffer too small; *sgnlc holds the count needed */
#define PGW_ERR_DEVICE (-3)     /* a HIP runtime call failed */
#define PGW_ERR_NOMEM (-4)      /* host or device allocation failed */
#define PGW_ERR_UNSUPPORTED (-5)/* size outside what the kernels are built for */
#define PGW_ERR_NODEVICE (-6)   /* no HIP device visible */
#define PGW_ERR_IO (-7)         /* a trigger or state file could not be written or read */
#define PGW_ERR_SCRATCH (-8)    /* input larger than the allocation-free entry point's scratch */

/* Device scratch of pgw_fstatnet_noalloc, per GPU: 8 MiB, i.e. npts * ceil(N / 4096) <= 40329
 * (1833 points at N = 86400, 630 at N = 259200). */
#define PGW_FSN_SCRATCH_BYTES (8u << 20)

/* arithmetic of the modulated series and the FFTs; phases are FP64 in both */
#define PGW_PREC_FP64 0
#define PGW_PREC_FP32 1

/* Search settings: the fields of Search_settings (search/network/src-cpu/struct.h:103-140)
 * that the search core reads.  Fill with pgw_search_settings(). */
typedef struct pgw_settings {
  double fpo;          /* band start frequency [Hz] */
  double dt;           /* sampling time [s] */
  double B;            /* bandwidth 0.5/dt */
  double oms;          /* 2 pi fpo dt */
  double omr;          /* Earth rotation C_OMEGA_R dt */
  double Smin, Smax;   /* spindown bounds */
  double sepsm, cepsm; /* sin/cos of the ecliptic obliquity (from DetSSB.bin) */
  int nfft;            /* 2^ceil(log2 N) */
  int N;               /* samples per segment */
  int nfftf;           /* fftpad * nfft = F-statistic FFT length L */
  int nmax, nmin;      /* F-statistic bin range [nmin, nmax) */
  int nd;              /* degrees of freedom (2) */
  int interpftpad;     /* resampling factor (2) */
  int fftpad;          /* zero padding (from grid.bin) */
  int Ninterp;         /* interpftpad * nfft */
  int nifo;            /* detectors */
  double M[16];        /* grid-generating matrix, column-major as in grid.bin */
  int numlines_band;   /* known lines to veto */
  double lines[PGW_MAXL][2];
} pgw_settings;

/* One detector's data: the fields of Detector_settings/Signals
 * (search/network/src-cpu/struct.h:36-55,154-171) that job_core reads. */
typedef struct pgw_detector {
  const double *xDat;    /* N samples, exact zeros mark gaps */
  const double *DetSSB;  /* 3N doubles: detector position w.r.t. the SSB, in samples */
  double sig2;           /* crf0 * var(xDat), see pgw_sig2() */
  double sphir, cphir;   /* sin/cos of the sidereal phase phir at t=0 */
  double amod[9];        /* c1..c9 of the amplitude modulation, see pgw_rogcvir() */
} pgw_detector;

/* Search_range, search/network/src-cpu/struct.h:70-73 */
typedef struct pgw_range {
  int pmr[2], mr[2], nr[2], spndr[2];
  int pst, mst, nst, sst;
} pgw_range;

/* The Command_line_opts fields read by job_core (search/network/src-cpu/struct.h:15-32). */
typedef struct pgw_opts {
  double trl;        /* F threshold (default 20) */
  int white_flag;    /* --whitenoise: skip FStat normalisation */
  int s0_flag;       /* --nospindown */
  int use_range;     /* nonzero when -range or -addsig was given: spindowns [sst, spndr[1]],
                        otherwise the range follows from Smin/Smax (jobcore.c:424-439) */
} pgw_opts;

typedef struct pgw_handle pgw_handle;

/* Where pgw_search writes (the Command_line_opts fields of search(), jobcore.c:51-184). */
typedef struct pgw_output {
  const char *prefix;    /* directory of the trigger files (opts->prefix) */
  int ident, band;       /* frame and band numbers in the file names (opts->ident, opts->band) */
  const char *label;     /* file-name label (opts->label); NULL or "" for none */
  const char *qname;     /* checkpoint state file (opts->qname); NULL: no checkpointing */
  int checkpoint_every;  /* sky points between checkpoints and trigger flushes (<= 0: 64) */
  int max_sky;           /* stop after this many sky points, state saved (<= 0: no limit) */
  int float_records;     /* nonzero: records of 5 floats, as a COMP_FLOAT build writes them
                            (FLOAT_TYPE = float, search/network/src-cpu/auxi.h:16-23) */
} pgw_output;

/* search_settings() + the size part of plan_fftw()
 * (search/network/src-cpu/settings.c:16-81, init.c:835-844) for a segment of N samples.
 * N replaces round(nod*C_SIDDAY/dt) so that solar-day segment lengths can be used. */
int pgw_search_settings(double fpo, double dt, int N, int fftpad, pgw_settings *sett);

/* rogcvir(), search/network/src-cpu/settings.c:253-275: amplitude modulation
 * coefficients c1..c9 from latitude ephi and orientation egam [rad]. */
int pgw_rogcvir(double ephi, double egam, double amod[9]);

/* crf0*var(xDat) as in init_arrays(), search/network/src-cpu/init.c:359-370. */
double pgw_sig2(const double *xDat, int N);

/* Allocate device memory, upload the detector data and settings once.
 * Replaces the allocation part of init_arrays()/plan_fftw() (search/network/src-cpu/init.c:326-457,810-862).
 * precision: PGW_PREC_FP64 (parity path) or PGW_PREC_FP32 (FP32 series and FFTs, FP64 phases). */
int pgw_init(pgw_handle **h, const pgw_settings *sett, const pgw_detector *ifo, int nifo, int precision);

/* One sky point, all its spindowns: the drop-in for job_core()
 * (search/network/src-cpu/jobcore.h:22-35, body jobcore.c:189-785).
 * Candidates are appended to sgnlv (PGW_NPAR doubles each) at index *sgnlc, in the
 * reference's (ss, bin) order; *FNum is incremented once per template.
 * Returns PGW_ERR_CAPACITY (and leaves *sgnlc at the count needed) if sgnlv_cap records
 * would be exceeded.  A sky point outside the grid disk returns PGW_OK with no work. */
int pgw_job_core(pgw_handle *h, int pm, int mm, int nn, pgw_range *s_range, const pgw_opts *opts,
                 int *sgnlc, double *sgnlv, int sgnlv_cap, int *FNum);

/* Many sky points in one call: sky[3*i..3*i+2] = (pm, mm, nn).  Equivalent to calling
 * pgw_job_core for each in order (the spotlight driver's loop,
 * search/spotlight/src/jobcore.c:79-122).  Up to 16 consecutive sky points form a group whose
 * templates share kernel launches; the host builds the records of one group while the device
 * works on the next. */
int pgw_job_batch(pgw_handle *h, const int *sky, int nsky, pgw_range *s_range, const pgw_opts *opts,
                  int *sgnlc, double *sgnlv, int sgnlv_cap, int *FNum);

/* pgw_job_core and pgw_job_batch for a driver built with COMP_FLOAT (FLOAT_TYPE = float,
 * search/network/src-cpu/auxi.h:16-23): records of PGW_NPAR floats, each field the double the
 * double-record entry points write rounded once to float, as job_core's assignment to its
 * FLOAT_TYPE sgnlv does (jobcore.c:686-687).  Same semantics and status codes otherwise. */
int pgw_job_core_f(pgw_handle *h, int pm, int mm, int nn, pgw_range *s_range, const pgw_opts *opts,
                   int *sgnlc, float *sgnlv, int sgnlv_cap, int *FNum);
int pgw_job_batch_f(pgw_handle *h, const int *sky, int nsky, pgw_range *s_range, const pgw_opts *opts,
                    int *sgnlc, float *sgnlv, int sgnlv_cap, int *FNum);

/* The spotlight search (search/spotlight/src/jobcore.c:51-135 and its job_core, :137-640): sky
 * positions (mm[i], nn[i]) of hemisphere pm, each searched over its own list of noss[i] spindowns
 * (the lists concatenated in ss, in the order of the spotlight file read at
 * search/spotlight/src/init.c:392-442).  Records are appended per sky position in list order, as
 * the spotlight driver writes them; one FNum per template.  Sky points outside the grid disk are
 * skipped (the spotlight job_core does not test for them).  No line veto is applied unless the
 * handle's settings carry lines. */
int pgw_job_spotlight(pgw_handle *h, int pm, const int *mm, const int *nn, const int *noss, const int *ss,
                      int nsky, const pgw_opts *opts, int *sgnlc, double *sgnlv, int sgnlv_cap, int *FNum);

/* Several segments (time frames) of one network in one handle: sett[nseg], ifo[nseg*nifo]
 * (detector d of segment k at ifo[k*nifo + d]).  Segments may have different N; those with the
 * same nfft and fftpad share kernel launches.  pgw_init is the nseg = 1 case.  No counterpart
 * in the reference, which searches one frame per process (search/network/src-cpu/init.c:81-99). */
int pgw_init_multi(pgw_handle **h, const pgw_settings *sett, const pgw_detector *ifo, int nseg, int nifo,
                   int precision);

/* Work items over segments: work[4*i..4*i+3] = (segment, pm, mm, nn), processed in order with
 * pgw_job_core's semantics per item.  rec_work (optional, sgnlv_cap entries) receives the work
 * index of each record. */
int pgw_job_batch_multi(pgw_handle *h, const int *work, int nwork, pgw_range *s_range,
                        const pgw_opts *opts, int *sgnlc, double *sgnlv, int sgnlv_cap, int *rec_work,
                        int *FNum);

/* Exact network F-statistic at continuous parameters: the followup's Fstatnet
 * (followup/src/followup.c:413-534, detector responses from modvir at the point as
 * followup.c:676-685 builds them) for npts points pts[4*i..4*i+3] = (omega [rad/sample],
 * omegadot [rad/sample^2], delta, alpha [rad]) over segment `seg`.  out[11*i..] receives the
 * reference's 11 values: amplitude estimates [0..3], sqrt(2(F-2)) [4], -F [5], the point [6..9],
 * amplitude SNR [10].  grad (optional, 4*npts) receives dF/d(omega, omegadot, delta, alpha) of
 * F = -out[5] (the analytic gradient; no counterpart in the reference, whose followup searches with
 * Nelder-Mead and MADS, followup.c:537-1173).  FP64 on the device. */
int pgw_fstatnet(pgw_handle *h, int seg, const double *pts, int npts, double *out, double *grad);

/* pgw_fstatnet for points of several segments in one call: point i (pts[4*i..], out[11*i..],
 * grad[4*i..]) is evaluated over segment seg[i].  The followup's candidates of many time frames
 * (or many injections) go to the device together; no counterpart in the reference, which follows
 * up one frame per process. */
int pgw_fstatnet_multi(pgw_handle *h, const int *seg, const double *pts, int npts, double *out, double *grad);

/* Allocation-free pgw_fstatnet.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  pts (4 * npts), out (11 * npts) and grad (4 * npts, optional) are DEVICE addresses
 * owned by the caller, with pgw_fstatnet's layouts and values.  The partial sums live in a fixed
 * device array of PGW_FSN_SCRATCH_BYTES per GPU, set up when the handle is created: no hipMalloc,
 * hipFree or allocating library call happens here.  npts * ceil(N / 4096) * 26 doubles must fit
 * (else PGW_ERR_SCRATCH, before any device work).  The kernels run on the handle's stream and the
 * call returns when they are done.  All handles and calls on one GPU share the scratch: calls
 * that overlap in time on one GPU (other host threads) race and must be serialised by the caller. */
int pgw_fstatnet_noalloc(pgw_handle *h, int seg, const double *pts, int npts, double *out, double *grad);

/* Autotuner (no counterpart in the reference): times a small search around sky point
 * (pm, mm, nn) of segment `seg` (which must lie inside the sky disk; opts->trl is not used) for
 * every kernel variant (templates per launch 64/128/256;
 * row pass with 16 elements per thread and both spectra together or in turn, or, in the FP64
 * form, with 8 elements per thread; column pass LDS sharing; in the FP32 form at column lengths
 * of 128 and 256, loads-first demodulation; spectra written with or without non-temporal
 * stores) and keeps the fastest for this handle.  The chosen values are returned (pointers may
 * be NULL): *row_na is 1 or 2 (16 elements per thread, spectra in turn or together) or 8 (8
 * elements per thread); the loads-first and store choices are kept by the handle and not
 * returned. */
int pgw_autotune(pgw_handle *h, int seg, int pm, int mm, int nn, const pgw_opts *opts, int *tb,
                 int *row_na, int *col_seq);

/* pgw_autotune over calls shaped like the caller's (no counterpart in the reference): each timed
 * call holds nsky work items (1..64, copies of the sky point), each with the spindowns that
 * s_range (and opts->use_range) give there, as pgw_job_batch would, instead of four items of 512,
 * and the searches use opts->trl (candidates and records formed as in the caller's search).
 * Small calls favour other variants than large ones (their spectra can stay in the device's
 * last-level cache between the passes; their launches have fewer neighbours to overlap with). */
int pgw_autotune_range(pgw_handle *h, int seg, int pm, int mm, int nn, const pgw_opts *opts,
                       const pgw_range *s_range, int nsky, int *tb, int *row_na, int *col_seq);

/* The handle's kernel variants (no counterpart in the reference) as a PGW_TUNE string, e.g.
 * "tb=128,row_na=1,row8=0,col_seq=1,col_lf=0,col_nt=0,row_nt=1,f_nt=1,rev=1,half=1,spl_k=4,
 * group=16,thr_fuse=1,l1max=0,candcap=1048576": after pgw_autotune, the choice it made; setting
 * PGW_TUNE to it before
 * pgw_init reproduces it.  buf receives a NUL-terminated string; PGW_ERR_ARG when len is too
 * small (256 bytes always suffice). */
int pgw_tune_string(const pgw_handle *h, char *buf, int len);

/* Debug tap: the normalised F-statistic F[nmin..nmax) of one template (pm, mm, nn, ss),
 * as job_core holds it just before thresholding (jobcore.c:635-651).  F has nmax-nmin entries. */
int pgw_debug_fstat(pgw_handle *h, int pm, int mm, int nn, int ss, const pgw_opts *opts, double *F);

/* Debug tap: the resampled, noise-weighted series xDatma/xDatmb of detector `det`
 * for sky point (pm, mm, nn) (jobcore.c:367-393); each array holds N complex values
 * as interleaved (re, im) doubles. */
int pgw_debug_resampled(pgw_handle *h, int pm, int mm, int nn, int det, double *xDatma, double *xDatmb);

/* Profiling tap (no counterpart in the reference): runs the sky-point stage of (pm, mm, nn)
 * once, then `reps` times the template stage for spindowns [s0, s0+nt) (nt is capped at the
 * handle's templates per launch), with HIP events around each kernel on the handle's stream.
 * Average milliseconds per launch: ms[0] sky-point stage, ms[1] column pass (demodulation +
 * column FFT), ms[2] row pass (row FFT + F + FStat partials), ms[3] threshold; ms[4] = the
 * number of templates per launch used. */
int pgw_time_stages(pgw_handle *h, int pm, int mm, int nn, int s0, int nt, const pgw_opts *opts,
                    int reps, double ms[5]);

/* Kernel timing over real searches (no counterpart in the reference): with enable = k >= 1,
 * every k-th template launch of later pgw_job_core / pgw_job_batch / pgw_job_batch_multi /
 * pgw_job_spotlight / pgw_search calls records HIP events (without system-scope fences) around
 * its column pass, row pass and threshold kernels on the handle's stream; enable = 0 stops
 * recording.  Each call returns in out[5] the totals since the previous call, then resets them:
 * out[0..2] milliseconds of the column, row and threshold kernels summed over the recorded
 * launches, out[3] those launches, out[4] their templates. */
int pgw_kernel_times(pgw_handle *h, int enable, double out[5]);

/* Calibration helper (no counterpart in the reference): a plain device copy kernel of `bytes`
 * bytes (read + write), run `reps` times; *ms = average milliseconds per launch.  Used to scale
 * hardware byte counters. */
int pgw_copy_bench(size_t bytes, int reps, double *ms);

/* Measurement helper (no counterpart in the reference): the HBM bandwidth this device reaches with
 * simple streaming kernels over a buffer of `bytes` bytes.  kind 0: copy (bytes read and bytes
 * written), 1: read only, 2: write only (non-temporal stores); the fastest of three grid sizes,
 * `reps` launches each; *ms = milliseconds per launch. */
int pgw_hbm_probe(int kind, size_t bytes, int reps, double *ms);

/* Elapsed device time (ms) of the last pgw_job_core / pgw_job_batch call, from HIP events. */
double pgw_last_gpu_ms(const pgw_handle *h);

/* The settings of segment `seg` as the handle holds them (pgw_init's input). */
int pgw_get_settings(const pgw_handle *h, int seg, pgw_settings *sett);

/* search(), search/network/src-cpu/jobcore.c:51-184: hemispheres pm = pst..pmr[1], then
 * mm = mst..mr[1], nn = nst..nr[1] (the starts fall back to mr[0], nr[0] and sst to spndr[0]
 * after the first), every sky point through pgw_job_core's semantics.  Records are appended to
 * <prefix>/triggers_<ident %03d>_<band %04d><label>_<pm>.bin (5 doubles each, 5 floats with
 * out->float_records) under an fcntl
 * write lock whenever more than nfft are buffered and at the end of each hemisphere (the file is
 * created even when empty), as the reference does.  Sky points run in batches of up to
 * checkpoint_every through the batched engine.  With a state file, after each batch the buffered
 * records are appended first and then "pm mm nn sst FNum" of the next sky point is written (the
 * reference writes the state before its buffer is flushed, so a restart loses the buffered
 * records); at the end the state file is emptied (search/network/src-cpu/main.c:141-147).
 * s_range's start fields are updated as the reference's are. */
int pgw_search(pgw_handle *h, pgw_range *s_range, const pgw_opts *opts, const pgw_output *out, int *FNum);

/* read_checkpoints(), search/network/src-cpu/init.c:868-917, for the state file `qname`:
 * no file: starts at the range's beginning, *FNum = 0, returns PGW_OK; a state: pst, mst, nst,
 * sst and *FNum from it, returns PGW_OK; an empty file (a finished run): returns 1;
 * unreadable contents: PGW_ERR_IO. */
int pgw_read_checkpoint(const char *qname, pgw_range *s_range, int *FNum);

/* Release everything owned by the handle.  Safe on NULL. */
void pgw_destroy(pgw_handle *h);

/* Human-readable text for a status code. */
const char *pgw_strerror(int status);

#ifdef __cplusplus
}
#endif

#endif /* PGW_H */
