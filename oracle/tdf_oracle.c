/*
 * CPU oracle for the TDFstat coherent-search hot path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker / the timed CPU baseline.
 *
 * Every function restates the reference (search/network/openmp/...) at the
 * cited lines.  The reference itself cannot be compiled in this image
 * (fftw3.h and GSL are absent), so each function is pinned instead by
 * independent checks in tests/test_oracle_*.py:
 *
 *   or_lin2ast      auxi.c:22-29      docs KAT (ra/dec/fdot of a published
 *                                     trigger record) + sphere identities
 *   or_rogcvir      settings.c:226-248  } JKS 1998 eqs 12-13 closed form
 *   or_modvir       settings.c:254-294  } evaluated in a different algebraic
 *                                         form; gap samples -> 0
 *   or_fft/or_fftf  (FFTW_FORWARD/BACKWARD semantics, init.c:741-752)
 *                                     numpy.fft and brute-force DFT
 *   or_spline       auxi.c:93-117     scipy CubicSpline(bc_type='natural')
 *   or_splintpad    auxi.c:120-169    scipy natural spline evaluated at the
 *                                     same abscissae; zero outside [0,n-1]
 *   or_fstat_norm   auxi.c:336-363    block means become exactly 2; tail
 *                                     untouched (closed form)
 *   or_sky_stage    jobcore.c:242-411 band-limited tone round trip (closed
 *                                     form: pure tone demodulates to DC and
 *                                     the FFT interpolation reproduces a trig
 *                                     polynomial), aa/bb sums vs numpy,
 *                                     spindown range vs docs KAT
 *   or_spindown     jobcore.c:437-603 brute-force DFT of the demodulated
 *                                     series at tiny sizes; hand-built F
 *                                     vectors with known MAX_ALG-3 triggers
 *                                     (ties, edges, veto, threshold)
 *   or_job_core     composition of the two above (OpenMP over spindowns)
 *
 * Known deviation: splint reads y2[n] when x == n-1 exactly (auxi.c:128-131
 * with y2 of length n); here y2[n] is taken as 0 (its coefficient b^3-b is 0).
 */
#define _GNU_SOURCE
#include "tdf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SQR(x) ((x) * (x))

/* ------------------------------------------------------------------ */
/* auxi.c:22-29 */
void or_lin2ast(double be1, double be2, int pm, double sepsm, double cepsm,
                double *sinal, double *cosal, double *sindel, double *cosdel)
{
    *sindel = be1 * sepsm - (2 * pm - 3) * sqrt(1. - SQR(be1) - SQR(be2)) * cepsm;
    *cosdel = sqrt(1. - SQR(*sindel));
    *sinal = (be1 - sepsm * (*sindel)) / (cepsm * (*cosdel));
    *cosal = be2 / (*cosdel);
}

/* settings.c:226-248, coefficients of JKS 1998 eqs 12-13 */
void or_rogcvir(double ephi, double egam, double c[9])
{
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

/* settings.c:254-294 */
void or_modvir(double sinal, double cosal, double sindel, double cosdel,
               int Np, const or_det *d, const double *cosmodf,
               const double *sinmodf, double *aa, double *bb)
{
    const double *c = d->c;
    double cosalfr = cosal * d->cphir + sinal * d->sphir;
    double sinalfr = sinal * d->cphir - cosal * d->sphir;
    double c2d = SQR(cosdel);
    double c2sd = sindel * cosdel;
    for (int t = 0; t < Np; t++) {
        if (fabs(d->xDat[t]) > DBL_MIN) {
            double cc = cosalfr * cosmodf[t] + sinalfr * sinmodf[t];
            double ss = sinalfr * cosmodf[t] - cosalfr * sinmodf[t];
            double c2s = 2. * SQR(cc);
            double cs = cc * ss;
            aa[t] = c[0] * (2. - c2d) * c2s + c[1] * (2. - c2d) * 2. * cs +
                    c[2] * c2sd * cc + c[3] * c2sd * ss - c[0] * (2. - c2d) +
                    c[4] * c2d;
            bb[t] = c[5] * sindel * c2s + c[6] * sindel * 2. * cs +
                    c[7] * cosdel * cc + c[8] * cosdel * ss - c[5] * sindel;
        } else {
            aa[t] = 0.;
            bb[t] = 0.;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Unnormalised complex DFT, FFTW sign convention (sign=-1 forward).
 * Iterative radix-2 with exact per-size twiddle tables. */

static int ilog2(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

static double *tw_d[32];
static float *tw_f[32];

static const double *twiddles_d(int n)
{
    int l = ilog2(n);
    double *t = __atomic_load_n(&tw_d[l], __ATOMIC_ACQUIRE);
    if (t) return t;
#pragma omp critical(or_tw)
    {
        if (!tw_d[l]) {
            double *w = (double *)malloc(sizeof(double) * (size_t)n);
            for (int k = 0; k < n / 2; k++) {
                double s, c;
                sincos(2. * M_PI * (double)k / (double)n, &s, &c);
                w[2 * k] = c;
                w[2 * k + 1] = s;
            }
            __atomic_store_n(&tw_d[l], w, __ATOMIC_RELEASE);
        }
    }
    return tw_d[l];
}

static const float *twiddles_f(int n)
{
    int l = ilog2(n);
    float *t = __atomic_load_n(&tw_f[l], __ATOMIC_ACQUIRE);
    if (t) return t;
    const double *wd = twiddles_d(n);
#pragma omp critical(or_twf)
    {
        if (!tw_f[l]) {
            float *w = (float *)malloc(sizeof(float) * (size_t)n);
            for (int k = 0; k < n; k++) w[k] = (float)wd[k];
            __atomic_store_n(&tw_f[l], w, __ATOMIC_RELEASE);
        }
    }
    return tw_f[l];
}

#define DEFINE_FFT(NAME, T, TWFN)                                              \
    void NAME(T *x, int n, int sign)                                           \
    {                                                                          \
        const T *w = TWFN(n);                                                  \
        /* bit reversal */                                                     \
        for (int i = 1, j = 0; i < n; i++) {                                   \
            int bit = n >> 1;                                                  \
            for (; j & bit; bit >>= 1) j ^= bit;                               \
            j ^= bit;                                                          \
            if (i < j) {                                                       \
                T tr = x[2 * i], ti = x[2 * i + 1];                            \
                x[2 * i] = x[2 * j];                                           \
                x[2 * i + 1] = x[2 * j + 1];                                   \
                x[2 * j] = tr;                                                 \
                x[2 * j + 1] = ti;                                             \
            }                                                                  \
        }                                                                      \
        for (int len = 2; len <= n; len <<= 1) {                               \
            int half = len >> 1, step = n / len;                               \
            for (int i = 0; i < n; i += len) {                                 \
                for (int k = 0; k < half; k++) {                               \
                    T wr = w[2 * k * step];                                    \
                    T wi = (T)sign * w[2 * k * step + 1];                      \
                    T *a = x + 2 * (i + k), *b = x + 2 * (i + k + half);       \
                    T br = b[0] * wr - b[1] * wi;                              \
                    T bi = b[0] * wi + b[1] * wr;                              \
                    b[0] = a[0] - br;                                          \
                    b[1] = a[1] - bi;                                          \
                    a[0] += br;                                                \
                    a[1] += bi;                                                \
                }                                                              \
            }                                                                  \
        }                                                                      \
    }

DEFINE_FFT(or_fft, double, twiddles_d)
DEFINE_FFT(or_fftf, float, twiddles_f)

/* ------------------------------------------------------------------ */
/* auxi.c:93-117: natural cubic spline, unit spacing, Thomas sweep.
 * The reference keeps y2 complex; its first-sweep values are real
 * (-0.5*invp), so they are carried here in a real array cf[]. */
void or_spline(const double *y, int n, double *y2)
{
    double *cf = (double *)malloc(sizeof(double) * (size_t)n);
    double *u = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    cf[0] = 0.;
    u[0] = u[1] = 0.;
    for (int i = 1; i < n - 1; ++i) {
        double invp = 2. / (cf[i - 1] + 4.);
        cf[i] = -.5 * invp;
        for (int c = 0; c < 2; c++) {
            double ui = y[2 * (i - 1) + c] - 2. * y[2 * i + c] + y[2 * (i + 1) + c];
            u[2 * i + c] = (-.5 * u[2 * (i - 1) + c] + 3. * ui) * invp;
        }
    }
    y2[2 * (n - 1)] = 0.;
    y2[2 * (n - 1) + 1] = 0.;
    for (int k = n - 2; k >= 0; --k)
        for (int c = 0; c < 2; c++)
            y2[2 * k + c] = cf[k] * y2[2 * (k + 1) + c] + u[2 * k + c];
    free(cf);
    free(u);
}

/* auxi.c:120-133 */
static inline void splint(const double *ya, const double *y2a, int n,
                          double x, double *out)
{
    if (x < 0 || x > n - 1) {
        out[0] = out[1] = 0.;
        return;
    }
    int klo = (int)floor(x);
    int khi = klo + 1;
    double a = khi - x;
    double b = x - klo;
    for (int c = 0; c < 2; c++) {
        double y2hi = (khi < n) ? y2a[2 * khi + c] : 0.;
        out[c] = a * ya[2 * klo + c] + b * ya[2 * khi + c] +
                 ((a * a * a - a) * y2a[2 * klo + c] + (b * b * b - b) * y2hi) / 6.0;
    }
}

/* auxi.c:136-169 */
void or_splintpad(double *ya, const double *shftf, int N, int interpftpad,
                  double *out)
{
    int n = interpftpad * N;
    double *y2 = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    or_spline(ya, n, y2);
    for (int i = 0; i < N; ++i) {
        double x = interpftpad * (i - shftf[i]);
        splint(ya, y2, n, x, out + 2 * i);
    }
    free(y2);
}

/* auxi.c:336-363 (FStat) -- block-mean normalisation */
double or_fstat_norm(double *F, int nfft, int nav)
{
    double pxout = 100.;
    double *fr = F;
    for (int j = 0; j < nfft / nav; j++) {
        double mu = 0.;
        for (int i = 0; i < nav; i++) mu += *fr++;
        mu /= 2. * nav;
        if (mu < 0.6) pxout = mu;
        fr -= nav;
        for (int i = 0; i < nav; i++) *fr++ /= mu;
    }
    return pxout;
}

double or_fstat_normf(float *F, int nfft, int nav)
{
    float mu, pxout = 100.f;
    float *fr = F;
    for (int j = 0; j < nfft / nav; j++) {
        mu = 0.f;
        for (int i = 0; i < nav; i++) mu += *fr++;
        mu /= 2. * nav;
        if (mu < 0.6) pxout = mu;
        fr -= nav;
        for (int i = 0; i < nav; i++) *fr++ /= mu;
    }
    return pxout;
}

/* ------------------------------------------------------------------ */
/* jobcore.c:242-411 -- per-sky stage */
int or_sky_stage(const or_sett *s, const or_det *dets, int pm, float mm,
                 float nn, const or_opts *o, or_sky *sky, double *xDatma,
                 double *xDatmb, double *tmp1, double *shft_out,
                 double *shftf_out, double *aa_out, double *bb_out)
{
    const int N = s->N, nfft = s->nfft, Ninterp = s->Ninterp;
    const double *M = s->M;
    double al1 = nn * M[10] + mm * M[14];
    double al2 = nn * M[11] + mm * M[15];
    double cof = s->oms;
    memset(sky, 0, sizeof(*sky));
    if ((SQR(al1) + SQR(al2)) / SQR(cof) > 1.) {
        sky->visible = 0;
        return 0;
    }
    sky->visible = 1;
    double sinalt, cosalt, sindelt, cosdelt;
    or_lin2ast(al1 / cof, al2 / cof, pm, s->sepsm, s->cepsm, &sinalt, &cosalt,
               &sindelt, &cosdelt);
    sky->dec = asin(sindelt);
    sky->ra = fmod(atan2(sinalt, cosalt) + 2. * M_PI, 2. * M_PI);
    double het0 = fmod(nn * M[8] + mm * M[12], M[0]);
    sky->het0 = het0;

    int nyqst = nfft / 2 + 1;
    size_t alen = (size_t)(Ninterp > s->nfftf ? Ninterp : s->nfftf);
    double *xa = (double *)calloc(2 * alen, sizeof(double));
    double *xb = (double *)calloc(2 * alen, sizeof(double));
    double *aa = (double *)malloc(sizeof(double) * N);
    double *bb = (double *)malloc(sizeof(double) * N);
    double *shft = (double *)malloc(sizeof(double) * N);
    double *shftf = (double *)malloc(sizeof(double) * N);
    double *cosm = (double *)malloc(sizeof(double) * N);
    double *sinm = (double *)malloc(sizeof(double) * N);
    if (!xa || !xb || !aa || !bb || !shft || !shftf || !cosm || !sinm) return -2;
    /* init.c:225-238 (init_aux_array) */
    for (int i = 0; i < N; i++) {
        double omrt = s->omr * i;
        cosm[i] = cos(omrt);
        sinm[i] = sin(omrt);
    }
    double aasum = 0., bbsum = 0.; /* FLOAT_TYPE aa,bb of the DOUBLE build */
    float aasumf = 0.f, bbsumf = 0.f; /* ... and of the FLOAT build */
    for (int n = 0; n < s->nifo; ++n) {
        const or_det *d = &dets[n];
        or_modvir(sinalt, cosalt, sindelt, cosdelt, N, d, cosm, sinm, aa, bb);
        double ns0 = cosalt * cosdelt, ns1 = sinalt * cosdelt, ns2 = sindelt;
        double shft1 = ns0 * d->DetSSB[0] + ns1 * d->DetSSB[1] + ns2 * d->DetSSB[2];
        double *t1 = tmp1 + (size_t)n * N;
        for (int i = 0; i < N; ++i) {
            shft[i] = ns0 * d->DetSSB[i * 3] + ns1 * d->DetSSB[i * 3 + 1] +
                      ns2 * d->DetSSB[i * 3 + 2];
            shftf[i] = shft[i] - shft1;
            t1[i] = SQR((double)i) + (double)(2 * i) * shft[i];
        }
        for (int i = 0; i < N; ++i) {
            double phase = het0 * i + s->oms * shft[i];
            double sp, cp;
            sincos(phase, &sp, &cp);
            double xaa = (double)d->xDat[i] * aa[i];
            double xbb = (double)d->xDat[i] * bb[i];
            xa[2 * i] = xaa * cp;
            xa[2 * i + 1] = -xaa * sp;
            xb[2 * i] = xbb * cp;
            xb[2 * i + 1] = -xbb * sp;
        }
        for (int i = N; i < nfft; ++i) xa[2 * i] = xa[2 * i + 1] = xb[2 * i] = xb[2 * i + 1] = 0.;
        or_fft(xa, nfft, -1);
        or_fft(xb, nfft, -1);
        /* jobcore.c:340-350 spectrum move */
        for (int i = nyqst + Ninterp - nfft, j = nyqst; i < Ninterp; ++i, ++j) {
            xa[2 * i] = xa[2 * j];
            xa[2 * i + 1] = xa[2 * j + 1];
            xb[2 * i] = xb[2 * j];
            xb[2 * i + 1] = xb[2 * j + 1];
        }
        for (int i = nyqst; i < nyqst + Ninterp - nfft; ++i)
            xa[2 * i] = xa[2 * i + 1] = xb[2 * i] = xb[2 * i + 1] = 0.;
        or_fft(xa, Ninterp, +1);
        or_fft(xb, Ninterp, +1);
        double ft = (double)s->interpftpad / Ninterp;
        for (int i = 0; i < 2 * Ninterp; ++i) {
            xa[i] *= ft;
            xb[i] *= ft;
        }
        double *oa = xDatma + 2 * (size_t)n * N;
        double *ob = xDatmb + 2 * (size_t)n * N;
        or_splintpad(xa, shftf, N, s->interpftpad, oa);
        or_splintpad(xb, shftf, N, s->interpftpad, ob);
        if (shft_out) memcpy(shft_out + (size_t)n * N, shft, sizeof(double) * N);
        if (shftf_out) memcpy(shftf_out + (size_t)n * N, shftf, sizeof(double) * N);
        if (aa_out) memcpy(aa_out + (size_t)n * N, aa, sizeof(double) * N);
        if (bb_out) memcpy(bb_out + (size_t)n * N, bb, sizeof(double) * N);
        /* jobcore.c:373-391 */
        double aat = 0., bbt = 0.;
        for (int i = 0; i < N; ++i) {
            aat += SQR(aa[i]);
            bbt += SQR(bb[i]);
        }
        aasum += aat / d->sig2;
        bbsum += bbt / d->sig2;
        aasumf += aat / d->sig2;
        bbsumf += bbt / d->sig2;
        for (int i = 0; i < 2 * N; ++i) {
            oa[i] /= d->sig2;
            ob[i] /= d->sig2;
        }
    }
    sky->aa = aasum;
    sky->bb = bbsum;
    sky->aa_f = aasumf;
    sky->bb_f = bbsumf;

    /* jobcore.c:399-411 spindown range */
    float smin = o->sst, smax = o->spndr1;
    if (o->gtype_allsky) {
        smin = trunc((s->Smin - nn * M[9] - mm * M[13]) / M[5]);
        smax = trunc(-(nn * M[9] + mm * M[13] + s->Smax) / M[5]);
        if (smin > smax) {
            smin = smin + smax;
            smax = smin - smax;
            smin = smin - smax;
        }
    }
    sky->smin = smin;
    sky->smax = smax;
    sky->iss_size = (smax - smin) / o->sstep + 1;

    free(xa); free(xb); free(aa); free(bb); free(shft); free(shftf);
    free(cosm); free(sinm);
    return 0;
}

/* ------------------------------------------------------------------ */
/* MAX_ALG 3 trigger selection, jobcore.c:553-603 (sequential semantics). */
#define DEFINE_TRIGGERS(NAME, T)                                               \
    static int NAME(const T *F, const or_opts *o, int nfftf, T sgnl0,          \
                    int max_trig, int *tic, float *tf, float *tF)              \
    {                                                                          \
        int itrig = 0;                                                         \
        for (int i = o->fr0 + 1; i < o->fr1 - 1; ++i) {                        \
            if (F[i] < o->thr || F[i - 1] > F[i] || F[i] < F[i + 1]) continue; \
            int ic = i;                                                        \
            T Fc = F[ic];                                                      \
            i++;                                                               \
            T f_ = (T)(2 * ic) / (T)nfftf * M_PI + sgnl0;                      \
            int veto = 0;                                                      \
            for (int k = 0; k < o->numlines; k++)                              \
                if (f_ >= o->lines[2 * k] && f_ <= o->lines[2 * k + 1]) {      \
                    veto = 1;                                                  \
                    break;                                                     \
                }                                                              \
            if (veto) continue;                                                \
            if (itrig >= o->bufsize / 2) return -1;                            \
            if (itrig < max_trig) {                                            \
                tic[itrig] = ic;                                               \
                tf[itrig] = (float)f_;                                         \
                tF[itrig] = (float)Fc;                                         \
            }                                                                  \
            itrig++;                                                           \
        }                                                                      \
        return itrig;                                                          \
    }

DEFINE_TRIGGERS(triggers_d, double)
DEFINE_TRIGGERS(triggers_f, float)

/* Exported for tests: MAX_ALG 3 selection on a given F vector (double). */
int or_triggers(const double *F, const or_opts *o, int nfftf, double sgnl0,
                int max_trig, int *tic, float *tf, float *tF)
{
    return triggers_d(F, o, nfftf, sgnl0, max_trig, tic, tf, tF);
}

/* jobcore.c:437-603, one spindown template.  work arrays are allocated per
 * call so that the function is reentrant. */
int or_spindown(const or_sett *s, const or_sky *sky, int prec, float mm,
                float nn, int iss, const or_opts *o, const double *xDatma,
                const double *xDatmb, const double *tmp1, double *Fout,
                int max_trig, int *trig_ic, float *trig_f, float *trig_F,
                float *rec6)
{
    const int N = s->N, nfftf = s->nfftf, nifo = s->nifo;
    const double *M = s->M;
    float ss = sky->smin + iss * o->sstep;
    int ntrig;
    if (prec == OR_PREC_DOUBLE) {
        double spnd = ss * M[5] + nn * M[9] + mm * M[13];
        double het1 = fmod(ss * M[4], M[0]);
        if (het1 < 0) het1 += M[0];
        double sgnl0 = sky->het0 + het1;
        if (rec6) {
            rec6[0] = mm; rec6[1] = nn; rec6[2] = ss;
            rec6[3] = (float)sky->ra; rec6[4] = (float)sky->dec;
            rec6[5] = (float)(spnd / (M_PI * s->dt * s->dt));
        }
        double *fxa = (double *)calloc(2 * (size_t)nfftf, sizeof(double));
        double *fxb = (double *)calloc(2 * (size_t)nfftf, sizeof(double));
        double *F = (double *)malloc(sizeof(double) * (size_t)nfftf);
        if (!fxa || !fxb || !F) return -2;
        /* spinmod.c:490-514 (GNUSINCOS, double) */
        for (int n = 0; n < nifo; ++n) {
            const double *ma = xDatma + 2 * (size_t)n * N;
            const double *mb = xDatmb + 2 * (size_t)n * N;
            const double *t1 = tmp1 + (size_t)n * N;
            for (int i = N - 1; i > -1; --i) {
                double phase = het1 * i + spnd * t1[i];
                double sp, cp;
                sincos(phase, &sp, &cp);
                /* (x + iy)(cp - i sp) */
                double ar = ma[2 * i] * cp + ma[2 * i + 1] * sp;
                double ai = ma[2 * i + 1] * cp - ma[2 * i] * sp;
                double br = mb[2 * i] * cp + mb[2 * i + 1] * sp;
                double bi = mb[2 * i + 1] * cp - mb[2 * i] * sp;
                if (n == 0) {
                    fxa[2 * i] = ar; fxa[2 * i + 1] = ai;
                    fxb[2 * i] = br; fxb[2 * i + 1] = bi;
                } else {
                    fxa[2 * i] += ar; fxa[2 * i + 1] += ai;
                    fxb[2 * i] += br; fxb[2 * i + 1] += bi;
                }
            }
        }
        or_fft(fxa, nfftf, -1);
        or_fft(fxb, nfftf, -1);
        for (int i = 0; i < nfftf; i++) F[i] = 0.;
        for (int i = o->fr0; i <= o->fr1; ++i)
            F[i] = (SQR(fxa[2 * i]) + SQR(fxa[2 * i + 1])) / sky->aa +
                   (SQR(fxb[2 * i]) + SQR(fxb[2 * i + 1])) / sky->bb;
        if (o->fstat_norm) or_fstat_norm(F + o->fr0, o->fr1 - o->fr0, o->nav);
        ntrig = triggers_d(F, o, nfftf, sgnl0, max_trig, trig_ic, trig_f, trig_F);
        if (Fout) memcpy(Fout, F, sizeof(double) * (size_t)nfftf);
        free(fxa); free(fxb); free(F);
    } else {
        /* PREC=FLOAT semantics: FLOAT_TYPE spnd/het1/sgnl0/aa/bb/tmp1/F */
        float spnd = ss * M[5] + nn * M[9] + mm * M[13];
        float het1 = fmod(ss * M[4], M[0]);
        if (het1 < 0) het1 += M[0];
        float sgnl0 = sky->het0 + het1;
        float aa = sky->aa_f, bb = sky->bb_f;
        if (rec6) {
            rec6[0] = mm; rec6[1] = nn; rec6[2] = ss;
            rec6[3] = (float)sky->ra; rec6[4] = (float)sky->dec;
            rec6[5] = (float)(spnd / (M_PI * s->dt * s->dt));
        }
        float *fxa = (float *)calloc(2 * (size_t)nfftf, sizeof(float));
        float *fxb = (float *)calloc(2 * (size_t)nfftf, sizeof(float));
        float *F = (float *)malloc(sizeof(float) * (size_t)nfftf);
        if (!fxa || !fxb || !F) return -2;
        /* spinmod.c:490-514 (GNUSINCOS, COMP_FLOAT -> sincosf) */
        for (int n = 0; n < nifo; ++n) {
            const double *ma = xDatma + 2 * (size_t)n * N;
            const double *mb = xDatmb + 2 * (size_t)n * N;
            const double *t1 = tmp1 + (size_t)n * N;
            for (int i = N - 1; i > -1; --i) {
                float tf = (float)t1[i];
                float phase = het1 * i + spnd * tf;
                float sp, cp;
                sincosf(phase, &sp, &cp);
                double ar = ma[2 * i] * cp + ma[2 * i + 1] * sp;
                double ai = ma[2 * i + 1] * cp - ma[2 * i] * sp;
                double br = mb[2 * i] * cp + mb[2 * i + 1] * sp;
                double bi = mb[2 * i + 1] * cp - mb[2 * i] * sp;
                if (n == 0) {
                    fxa[2 * i] = (float)ar; fxa[2 * i + 1] = (float)ai;
                    fxb[2 * i] = (float)br; fxb[2 * i + 1] = (float)bi;
                } else {
                    fxa[2 * i] = (float)(fxa[2 * i] + ar);
                    fxa[2 * i + 1] = (float)(fxa[2 * i + 1] + ai);
                    fxb[2 * i] = (float)(fxb[2 * i] + br);
                    fxb[2 * i + 1] = (float)(fxb[2 * i + 1] + bi);
                }
            }
        }
        or_fftf(fxa, nfftf, -1);
        or_fftf(fxb, nfftf, -1);
        for (int i = 0; i < nfftf; i++) F[i] = 0.f;
        for (int i = o->fr0; i <= o->fr1; ++i)
            F[i] = (fxa[2 * i] * fxa[2 * i] + fxa[2 * i + 1] * fxa[2 * i + 1]) / aa +
                   (fxb[2 * i] * fxb[2 * i] + fxb[2 * i + 1] * fxb[2 * i + 1]) / bb;
        if (o->fstat_norm) or_fstat_normf(F + o->fr0, o->fr1 - o->fr0, o->nav);
        ntrig = triggers_f(F, o, nfftf, sgnl0, max_trig, trig_ic, trig_f, trig_F);
        if (Fout)
            for (int i = 0; i < nfftf; i++) Fout[i] = F[i];
        free(fxa); free(fxb); free(F);
    }
    return ntrig;
}

/* ------------------------------------------------------------------ */
long or_job_core(const or_sett *s, const or_det *dets, int pm, float mm,
                 float nn, const or_opts *o, int prec, int iss_begin,
                 int iss_count, or_sky *sky, long max_trig, int *trig_iss,
                 int *trig_ic, float *trig_f, float *trig_F, int nthreads)
{
    const int N = s->N, nifo = s->nifo;
    double *xDatma = (double *)malloc(sizeof(double) * 2 * (size_t)nifo * N);
    double *xDatmb = (double *)malloc(sizeof(double) * 2 * (size_t)nifo * N);
    double *tmp1 = (double *)malloc(sizeof(double) * (size_t)nifo * N);
    if (!xDatma || !xDatmb || !tmp1) return -2;
    int rc = or_sky_stage(s, dets, pm, mm, nn, o, sky, xDatma, xDatmb, tmp1,
                          NULL, NULL, NULL, NULL);
    if (rc < 0 || !sky->visible) {
        free(xDatma); free(xDatmb); free(tmp1);
        return rc < 0 ? rc : 0;
    }
    int iss_end = iss_count < 0 ? sky->iss_size : iss_begin + iss_count;
    if (iss_end > sky->iss_size) iss_end = sky->iss_size;
    int nss = iss_end - iss_begin;
    if (nss <= 0) {
        free(xDatma); free(xDatmb); free(tmp1);
        return 0;
    }
    int cap = o->bufsize / 2 + 1;
    int *cnt = (int *)calloc((size_t)nss, sizeof(int));
    int *tic = (int *)malloc(sizeof(int) * (size_t)nss * cap);
    float *tf = (float *)malloc(sizeof(float) * (size_t)nss * cap);
    float *tF = (float *)malloc(sizeof(float) * (size_t)nss * cap);
    if (!cnt || !tic || !tf || !tF) return -2;
    int fail = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int k = 0; k < nss; k++) {
        int r = or_spindown(s, sky, prec, mm, nn, iss_begin + k, o, xDatma,
                            xDatmb, tmp1, NULL, cap, tic + (size_t)k * cap,
                            tf + (size_t)k * cap, tF + (size_t)k * cap, NULL);
        if (r < 0) {
#pragma omp atomic write
            fail = r;
            r = 0;
        }
        cnt[k] = r;
    }
    long total = 0;
    if (!fail) {
        for (int k = 0; k < nss; k++) {
            for (int t = 0; t < cnt[k]; t++) {
                if (total < max_trig) {
                    trig_iss[total] = iss_begin + k;
                    trig_ic[total] = tic[(size_t)k * cap + t];
                    trig_f[total] = tf[(size_t)k * cap + t];
                    trig_F[total] = tF[(size_t)k * cap + t];
                }
                total++;
            }
        }
    }
    free(cnt); free(tic); free(tf); free(tF);
    free(xDatma); free(xDatmb); free(tmp1);
    return fail ? fail : total;
}
