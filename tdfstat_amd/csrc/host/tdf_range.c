/* Search ranges (set_search_range, init.c:510-705; gridr and
 * dimless_to_code, auxi.c:63-90,264-332) and the checkpoint reader
 * (read_checkpoints, init.c:761-809). */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "tdfstat_search.h"

/* Solves A x = b for the 4x4 row-major A by Gaussian elimination with
 * partial pivoting (the reference uses an LU decomposition with partial
 * pivoting for both solves).  Returns 0, or -1 for a singular A. */
static int solve4(const double A_[16], const double b_[4], double x[4])
{
    double A[4][5];
    for (int i = 0; i < 4; i++) {
        for (int j = 0; j < 4; j++) A[i][j] = A_[4 * i + j];
        A[i][4] = b_[i];
    }
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int i = c + 1; i < 4; i++)
            if (fabs(A[i][c]) > fabs(A[p][c])) p = i;
        if (A[p][c] == 0.) return -1;
        if (p != c)
            for (int j = 0; j < 5; j++) {
                const double t = A[c][j];
                A[c][j] = A[p][j];
                A[p][j] = t;
            }
        for (int i = c + 1; i < 4; i++) {
            const double f = A[i][c] / A[c][c];
            for (int j = c; j < 5; j++) A[i][j] -= f * A[c][j];
        }
    }
    for (int i = 3; i >= 0; i--) {
        double s = A[i][4];
        for (int j = i + 1; j < 4; j++) s -= A[i][j] * x[j];
        x[i] = s / A[i][i];
    }
    return 0;
}

/* M^T (row-major) of the grid matrix M (row-major 4x4) */
static void transpose(const double M[16], double Mt[16])
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) Mt[4 * i + j] = M[4 * j + i];
}

void tdf_gridr(const double M[16], double oms, double Smax, float spndr[2], float nr[2], float mr[2])
{
    /* the 16 corners of [-pi, pi] x [-Smax, 0] x [-cof, cof]^2 (auxi.c:286-306) */
    const double Ob = M_PI, cof = oms + Ob;
    double Mt[16];
    transpose(M, Mt);
    spndr[0] = nr[0] = mr[0] = 16384;
    spndr[1] = nr[1] = mr[1] = -16384;
    for (int k = 0; k < 16; k++) {
        const double b[4] = {k < 8 ? Ob : -Ob, ((k >> 2) & 1) ? 0. : -Smax, (k & 1) ? -cof : cof,
                             ((k >> 1) & 1) ? -cof : cof};
        double x[4];
        if (solve4(Mt, b, x)) return;
        if (floor(x[1]) < spndr[0]) spndr[0] = floor(x[1]);
        if (ceil(x[1]) > spndr[1]) spndr[1] = ceil(x[1]);
        if (floor(x[2]) < nr[0]) nr[0] = floor(x[2]);
        if (ceil(x[2]) > nr[1]) nr[1] = ceil(x[2]);
        if (floor(x[3]) < mr[0]) mr[0] = floor(x[3]);
        if (ceil(x[3]) > mr[1]) mr[1] = ceil(x[3]);
    }
}

int tdf_dimless_to_code(const double M[16], const double pos_dimless[4], float pos_code[4])
{
    double Mt[16], x[4];
    transpose(M, Mt);
    if (solve4(Mt, pos_dimless, x)) return TDF_ERR_ARG;
    for (int i = 0; i < 4; i++) pos_code[i] = (float)x[i];
    return TDF_OK;
}

/* ast2lin (auxi.c:32-58) */
static int ast2lin(double alfa, double delta, double epsm, double be[2])
{
    be[0] = cos(epsm) * sin(alfa) * cos(delta) + sin(epsm) * sin(delta);
    be[1] = cos(alfa) * cos(delta);
    const double d1 = asin(be[0] * sin(epsm) + sqrt(1. - be[0] * be[0] - be[1] * be[1]) * cos(epsm)) - delta;
    return fabs(d1) < 10. * DBL_EPSILON ? 1 : 2;
}

int tdf_set_search_range(const tdf_run_info *run, int gtype, int hemi, const float gcenter[4],
                         const float gsizes[4], const float gsteps[4], tdf_range *r)
{
    if (!run || !r || gtype < 0 || gtype > 2) return TDF_ERR_ARG;
    memset(r, 0, sizeof *r);
    r->pmr[0] = hemi ? hemi : 1;
    r->pmr[1] = hemi ? hemi : 2;
    r->mstep = r->nstep = r->sstep = 1.f;
    if (gtype == 0) {
        tdf_gridr(run->M, run->oms, run->Smax, r->spndr, r->nr, r->mr);
        r->fr[0] = run->nmin;
        r->fr[1] = run->nmax;
        return TDF_OK;
    }
    if (!gcenter || !gsizes || !gsteps) return TDF_ERR_ARG;
    const float fr = gcenter[0], sr = gcenter[1], alphar = gcenter[2], deltar = gcenter[3];
    int fr_gsize = (int)gsizes[0];
    float sr_gsize = gsizes[1], mr_gsize = gsizes[2], nr_gsize = gsizes[3];
    const float sr_step = gsteps[1], mr_step = gsteps[2], nr_step = gsteps[3];
    if (!(sr_step > 0) || !(mr_step > 0) || !(nr_step > 0)) return TDF_ERR_ARG;
    /* physical -> dimensionless -> code units (init.c:609-627) */
    double sgnlol[4];
    sgnlol[0] = 0.;
    sgnlol[1] = M_PI * sr * run->dt * run->dt;
    const double cof = run->oms;
    double be[2];
    /* the obliquity constant, not sett's sepsm/cepsm (init.c:615) */
    const double epsm = 84381.448 / 3600. / (180 / M_PI);
    r->pmr[0] = r->pmr[1] = ast2lin((double)alphar, (double)deltar, epsm, be);
    sgnlol[2] = be[0] * cof;
    sgnlol[3] = be[1] * cof;
    float code[4];
    if (tdf_dimless_to_code(run->M, sgnlol, code)) return TDF_ERR_ARG;
    /* frequency bin of fr at this grid point (init.c:630-639) */
    const double *M = run->M;
    const double het0 = fmod(code[2] * M[8] + code[3] * M[12], M[0]);
    double het1 = fmod(code[1] * M[4], M[0]);
    if (het1 < 0) het1 += M[0];
    const float sgnl0 = (float)(het0 + het1);
    code[0] = (float)((fr - run->fpo) / run->B * run->nfftf / 2 - sgnl0 * run->nfftf / (2 * M_PI));
    r->mstep = mr_step;
    r->nstep = nr_step;
    r->sstep = sr_step;
    fr_gsize = (int)round(fr_gsize);
    sr_gsize = roundf(sr_gsize / sr_step) * sr_step;
    nr_gsize = roundf(nr_gsize / nr_step) * nr_step;
    mr_gsize = roundf(mr_gsize / mr_step) * mr_step;
    int fmid;
    float smid, nmid, mmid;
    if (gtype == 1) {
        fmid = (int)round(code[0]);
        smid = roundf(code[1]);
        nmid = roundf(code[2]);
        mmid = roundf(code[3]);
    } else {
        fmid = (int)code[0];
        smid = code[1];
        nmid = code[2];
        mmid = code[3];
    }
    r->fr[0] = fmid - fr_gsize;
    r->fr[1] = fmid + fr_gsize;
    r->spndr[0] = smid - sr_gsize;
    r->spndr[1] = smid + sr_gsize;
    r->nr[0] = nmid - nr_gsize;
    r->nr[1] = nmid + nr_gsize;
    r->mr[0] = mmid - mr_gsize;
    r->mr[1] = mmid + mr_gsize;
    return TDF_OK;
}

int tdf_read_checkpoint(const char *state_file, int checkp_flag, tdf_range *r, long long *fnum)
{
    if (!r || !fnum) return TDF_ERR_ARG;
    if (checkp_flag && state_file) {
        FILE *f = fopen(state_file, "r");
        if (f) {
            int pst, fn;
            float mst, nst, sst;
            const int got = fscanf(f, "%d %f %f %f %d", &pst, &mst, &nst, &sst, &fn);
            fclose(f);
            if (got != 5) return TDF_ERR_STATE; /* empty state file: the run had finished */
            r->pst = pst;
            r->mst = mst;
            r->nst = nst;
            r->sst = sst;
            *fnum = fn;
            return TDF_OK;
        }
    }
    r->pst = r->pmr[0];
    r->mst = r->mr[0];
    r->nst = r->nr[0];
    r->sst = r->spndr[0];
    *fnum = 0;
    return TDF_OK;
}
