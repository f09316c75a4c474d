/*
 * CPU oracle for the TDFstat coherent-search hot path (test infrastructure).
 *
 * This is a plain-C restatement of the reference CPU search core
 * (search/network/openmp/jobcore.c:193-630 and its helpers) used ONLY as the
 * checker by tests/, by __graft_entry__.smoke() and as the CPU timing baseline
 * in bench.py.  Product code never links or calls it.
 *
 * Two arithmetic modes are provided:
 *   OR_PREC_DOUBLE : the reference built with PREC=DOUBLE SIMD=NONE
 *                    (FLOAT_TYPE=double, libm sincos) -- the FP64 oracle.
 *   OR_PREC_FLOAT  : the reference's production PREC=FLOAT semantics
 *                    (FLOAT_TYPE=float, FP32 spindown phase, sincosf,
 *                    single-precision F-statistic FFT).  Used only as the CPU
 *                    timing baseline; SLEEF's u35 sincos is replaced by libm
 *                    sincosf (same contract, different last bits).
 *
 * Pinning status of every function is listed in tdf_oracle.c's header.
 */
#ifndef TDF_ORACLE_H
#define TDF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_PREC_DOUBLE 0
#define OR_PREC_FLOAT 1

typedef struct {
    int N, nfft, Ninterp, nfftf, interpftpad, nifo;
    double oms, omr, dt, sepsm, cepsm, Smin, Smax;
    double M[16];
} or_sett;

typedef struct {
    const float *xDat;      /* [N] */
    const double *DetSSB;   /* [3N] */
    double sig2, cphir, sphir;
    double c[9];            /* amod c1..c9 */
} or_det;

/* per-sky results (jobcore.c:242-411) */
typedef struct {
    int visible;            /* 0 if the grid point fails the (al1^2+al2^2)/oms^2 cut */
    double ra, dec;         /* sgnlt[3], sgnlt[2] (FLOAT_TYPE in the reference) */
    double het0;
    double aa, bb;          /* network normalisation sums (FLOAT_TYPE=double) */
    float aa_f, bb_f;       /* the same sums accumulated as FLOAT_TYPE=float */
    float smin, smax;       /* spindown range used by the spindown loop */
    int iss_size;
} or_sky;

typedef struct {
    int gtype_allsky;       /* strcmp(opts->gtype,"allsky")==0 */
    float sst, spndr1, sstep;
    int fr0, fr1;
    double thr;
    int fstat_norm;         /* 1 = "blocks_avg" */
    int nav;                /* FStat block length (NAVFSTAT) */
    int numlines;
    const double *lines;    /* [numlines][2] */
    int bufsize;            /* ffstat floats per spindown record */
} or_opts;

/* ---- helpers (each restates one reference function) ---- */
void or_lin2ast(double be1, double be2, int pm, double sepsm, double cepsm,
                double *sinal, double *cosal, double *sindel, double *cosdel);
void or_rogcvir(double ephi, double egam, double c[9]);
void or_modvir(double sinal, double cosal, double sindel, double cosdel,
               int Np, const or_det *d, const double *cosmodf,
               const double *sinmodf, double *aa, double *bb);
/* complex arrays are interleaved (re,im) doubles */
void or_fft(double *x, int n, int sign);       /* unnormalised, in place */
void or_fftf(float *x, int n, int sign);
void or_spline(const double *y, int n, double *y2);
void or_splintpad(double *ya, const double *shftf, int N, int interpftpad,
                  double *out);
double or_fstat_norm(double *F, int nfft, int nav);
double or_fstat_normf(float *F, int nfft, int nav);

/* ---- per-sky stage: jobcore.c:242-411 ----
 * xDatma/xDatmb: [nifo][N] complex (interleaved) after resampling and /sig2
 * tmp1: [nifo][N] (double; the FLOAT build rounds it to float on use)
 * shft/shftf/aaf/bbf optional (may be NULL): [nifo][N] */
int or_sky_stage(const or_sett *s, const or_det *dets, int pm, float mm,
                 float nn, const or_opts *o, or_sky *sky, double *xDatma,
                 double *xDatmb, double *tmp1, double *shft, double *shftf,
                 double *aaf, double *bbf);

/* ---- one spindown template: jobcore.c:437-603 ----
 * F (optional): [nfftf] F-statistic after optional FStat normalisation
 * trig_ic/f/F: triggers in ascending bin order; returns trigger count,
 * or -1 when more than bufsize/2 triggers would be written
 * (jobcore.c:586-590). */
int or_spindown(const or_sett *s, const or_sky *sky, int prec, float mm,
                float nn, int iss, const or_opts *o, const double *xDatma,
                const double *xDatmb, const double *tmp1, double *F,
                int max_trig, int *trig_ic, float *trig_f, float *trig_F,
                float *rec6 /* m n s ra dec fdot */);

/* MAX_ALG 3 trigger selection alone (jobcore.c:553-603), FP64 */
int or_triggers(const double *F, const or_opts *o, int nfftf, double sgnl0,
                int max_trig, int *tic, float *tf, float *tF);

/* Whole job_core for one sky point with all of its spindowns, OpenMP over
 * spindowns (the CPU baseline).  Triggers are appended as
 * (iss, ic, f, F); returns total trigger count, -1 on buffer overflow,
 * -2 on allocation failure. */
long or_job_core(const or_sett *s, const or_det *dets, int pm, float mm,
                 float nn, const or_opts *o, int prec, int iss_begin,
                 int iss_count, or_sky *sky, long max_trig, int *trig_iss,
                 int *trig_ic, float *trig_f, float *trig_F, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
