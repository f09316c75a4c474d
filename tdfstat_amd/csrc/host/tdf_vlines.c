/* Veto lines of a band: narrow_down_band, extract_band_vlines,
 * read_band_vlines, lines_veto_fraction, line_in_band (common.c:155-553)
 * and the fdot range they use (get_fdot_range, common.c:22-33). */
#include <glob.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tdfstat_search.h"

#define VL_MAXLINE 1024
#define C_YEARSEC_ (365.25 * 86400.)

void tdf_fdot_range(double fpo, double dt, double *fdot_min, double *fdot_max)
{
    const double B = 1. / (2 * dt);
    if (fpo < 200.) {
        *fdot_min = 2. * (fpo + B) / (2. * 1000. * C_YEARSEC_);
        *fdot_max = 0.;
    } else {
        *fdot_min = 2e-10;
        *fdot_max = 2e-11;
    }
}

int tdf_narrow_down_band(double narrowdown, double lines[][2], int *numlines_band)
{
    if (!lines || !numlines_band) return TDF_ERR_ARG;
    lines[0][0] = 0;
    lines[0][1] = M_PI_2 - narrowdown;
    lines[1][0] = M_PI_2 + narrowdown;
    lines[1][1] = M_PI;
    *numlines_band = 2;
    return TDF_OK;
}

/* clip [fl, fr] (Hz) to the band [fpo, fpo + B]; 0 if outside */
static int line_in_band(double *fl, double *fr, double fpo, double B)
{
    const double bs = fpo, be = fpo + B;
    if (!(*fr < bs || *fl > be)) {
        if (*fl < bs) *fl = bs;
        if (*fr > be) *fr = be;
        return 1;
    }
    return 0;
}

static int cmp2(const void *a, const void *b)
{
    const double *da = (const double *)a, *db = (const double *)b;
    const int d1 = (da[0] > db[0]) - (da[0] < db[0]);
    if (d1 != 0) return d1;
    return (da[1] > db[1]) - (da[1] < db[1]);
}

float tdf_lines_veto_fraction(double lines[][2], int lf, int le)
{
    /* sorted by start then end, then the gaps between the merged intervals */
    if (le > lf) qsort(&lines[lf], (size_t)(le - lf), 2 * sizeof(double), cmp2);
    double ll = 0., gap = 0.;
    for (int i = lf; i < le; i++) {
        if (lines[i][0] >= ll) {
            gap += lines[i][0] - ll;
            ll = lines[i][1];
        } else if (ll < lines[i][1]) {
            ll = lines[i][1];
        }
    }
    if (ll < M_PI) gap += M_PI - ll;
    return (float)(M_PI - gap) / M_PI;
}

/* one veto-file record: comb spacing (or line frequency), type (0 singlet,
 * 1 comb of fixed width, 2 comb of scaling width), offset of harmonic 0,
 * first and last visible harmonic, left and right widths (Hz) */
typedef struct {
    double f, offset, lwidth, rwidth;
    int type, iharm1, iharm2, det, nline;
    char vfile[VL_MAXLINE];
} vrec;

static int add_line(double lines[][2], int *j, int maxl, double fl, double fr, double fpo, double B, char info[][VL_MAXLINE],
                    const char *desc)
{
    if (!line_in_band(&fl, &fr, fpo, B)) return TDF_OK;
    if (*j >= maxl) return TDF_ERR_CAPACITY;
    lines[*j][0] = fl;
    lines[*j][1] = fr;
    if (info) snprintf(info[*j], VL_MAXLINE, "%s", desc);
    (*j)++;
    return TDF_OK;
}

int tdf_extract_band_vlines(const tdf_run_info *run, const double *const *detssb, double lines[][2], int maxl,
                            int *numlines_band, int *nvlines_all_inband, const char *band_vl_file,
                            float *veto_fraction)
{
    if (!run || !detssb || !lines || !numlines_band || !nvlines_all_inband || run->nifo < 1 || run->nifo > 8 ||
        maxl < *numlines_band)
        return TDF_ERR_ARG;
    const int N = run->N;
    double fdotmin, fdotmax;
    tdf_fdot_range(run->fpo, run->dt, &fdotmin, &fdotmax);
    /* per-detector maximum line broadening by the demodulation: squared
     * norms of the step of DetSSB and of the step of t*DetSSB (common.c:193-210) */
    double normdE[8], normdtE[8];
    for (int d = 0; d < run->nifo; d++) {
        const double *e = detssb[d];
        normdE[d] = normdtE[d] = 0.;
        for (int i = 0; i < N - 1; i++) {
            double dE = 0, dtE = 0;
            for (int k = 0; k < 3; k++) {
                dE += pow(fabs(e[(i + 1) * 3 + k] - e[i * 3 + k]), 2);
                dtE += pow(fabs(e[(i + 1) * 3 + k] * (i + 1) - e[i * 3 + k] * i) * run->dt, 2);
            }
            if (dE > normdE[d]) normdE[d] = dE;
            if (dtE > normdtE[d]) normdtE[d] = dtE;
        }
    }
    /* veto files <indir>/lines/<det>lines*.csv, comma separated, '%' comments */
    size_t cap = 256, nrec = 0;
    vrec *rec = (vrec *)malloc(cap * sizeof(vrec));
    if (!rec) return TDF_ERR_NOMEM;
    int rc = TDF_OK;
    for (int d = 0; d < run->nifo && rc == TDF_OK; d++) {
        char pat[4096];
        snprintf(pat, sizeof pat, "%s/lines/%slines*.csv", run->indir ? run->indir : ".",
                 run->det_name[d] ? run->det_name[d] : "");
        glob_t gb;
        memset(&gb, 0, sizeof gb);
        if (glob(pat, 0, NULL, &gb) != 0) {
            globfree(&gb);
            continue;
        }
        for (size_t f = 0; f < gb.gl_pathc && rc == TDF_OK; f++) {
            FILE *fp = fopen(gb.gl_pathv[f], "r");
            if (!fp) {
                rc = TDF_ERR_ARG;
                break;
            }
            char line[VL_MAXLINE];
            int nline = 0;
            while (fgets(line, sizeof line, fp)) {
                nline++;
                if (*line == '%') continue;
                char *t[7];
                t[0] = strtok(line, ",");
                for (int k = 1; k < 7; k++) t[k] = t[k - 1] ? strtok(NULL, ",") : NULL;
                if (!t[6]) {
                    rc = TDF_ERR_ARG;  /* the reference would read garbage here */
                    break;
                }
                if (nrec == cap) {
                    vrec *nr = (vrec *)realloc(rec, 2 * cap * sizeof(vrec));
                    if (!nr) {
                        rc = TDF_ERR_NOMEM;
                        break;
                    }
                    rec = nr;
                    cap *= 2;
                }
                vrec *v = &rec[nrec++];
                v->f = atof(t[0]);
                v->type = atoi(t[1]);
                v->offset = atof(t[2]);
                v->iharm1 = atoi(t[3]);
                v->iharm2 = atoi(t[4]);
                v->lwidth = atof(t[5]);
                v->rwidth = atof(t[6]);
                v->det = d;
                v->nline = nline;
                snprintf(v->vfile, sizeof v->vfile, "%s", gb.gl_pathv[f]);
            }
            fclose(fp);
        }
        globfree(&gb);
    }
    /* line widths (common.c:276-354): the veto lines follow the
     * narrowdown lines when the band is narrowed down */
    char(*info)[VL_MAXLINE] = (char(*)[VL_MAXLINE])calloc((size_t)maxl, VL_MAXLINE);
    int j = (run->narrowdown < 0.5 * M_PI) ? *numlines_band : 0;
    const double fpo = run->fpo, B = run->B, Tobs = run->N * run->dt;
    if (!info) rc = TDF_ERR_NOMEM;
    for (size_t i = 0; i < nrec && rc == TDF_OK; i++) {
        const vrec *v = &rec[i];
        char desc[VL_MAXLINE];
        if (v->type == 0) {
            const double D = 2. * fdotmax * (Tobs + sqrt(normdtE[v->det])) + v->f * sqrt(normdE[v->det]);
            snprintf(desc, sizeof desc, "singlet          [l.%d]%s", v->nline, v->vfile);
            rc = add_line(lines, &j, maxl, v->f - v->lwidth - D, v->f + v->rwidth + D, fpo, B, info, desc);
        } else if (v->type == 1 || v->type == 2) {
            for (int k = v->iharm1; k <= v->iharm2 && rc == TDF_OK; k++) {
                const double lf = v->offset + k * v->f;
                const double D = 2. * fdotmax * (Tobs + sqrt(normdtE[v->det])) + lf * sqrt(normdE[v->det]);
                const double wl = v->type == 1 ? v->lwidth : k * v->lwidth;
                const double wr = v->type == 1 ? v->rwidth : k * v->rwidth;
                snprintf(desc, sizeof desc, "%s rank %4d  [l.%d]%s", v->type == 1 ? "fcomb" : "scomb", k, v->nline,
                         v->vfile);
                rc = add_line(lines, &j, maxl, lf - wl - D, lf + wr + D, fpo, B, info, desc);
            }
        }
    }
    free(rec);
    if (rc != TDF_OK) {
        free(info);
        return rc;
    }
    *nvlines_all_inband = j;
    /* to radians (the narrowdown lines already are), then the band file */
    FILE *out = band_vl_file ? fopen(band_vl_file, "w") : NULL;
    if (band_vl_file && !out) {
        free(info);
        return TDF_ERR_ARG;
    }
    if (out) fprintf(out, "# fl[rad]  fr[rad]   fl[Hz]  fr[Hz]   line info\n");
    for (int i = *numlines_band; i < j; i++) {
        const double fl = lines[i][0], fr = lines[i][1];
        lines[i][0] = (lines[i][0] - fpo) / B * M_PI;
        lines[i][1] = (lines[i][1] - fpo) / B * M_PI;
        if (out) fprintf(out, "   %.15f  %.15f  %.15f  %.15f  %s\n", lines[i][0], lines[i][1], fl, fr, info[i]);
    }
    free(info);
    const float vf = tdf_lines_veto_fraction(lines, *numlines_band, j);
    if (out) {
        fprintf(out, "#band_veto_fraction= %6.4f\n", vf);
        fclose(out);
    }
    if (veto_fraction) *veto_fraction = vf;
    if (run->veto_flag) *numlines_band = j;
    return TDF_OK;
}

int tdf_read_band_vlines(const char *band_vl_file, double narrowdown, int veto_flag, double lines[][2], int maxl,
                         int *numlines_band, int *nvlines_all_inband, float *veto_fraction)
{
    if (!band_vl_file || !lines || !numlines_band || !nvlines_all_inband) return TDF_ERR_ARG;
    static const char tag[] = "#band_veto_fraction=";
    FILE *fp = fopen(band_vl_file, "r");
    if (!fp) return TDF_ERR_ARG;
    int i = (narrowdown < 0.5 * M_PI) ? *numlines_band : 0;
    float vf = -1.f;
    int rc = TDF_OK;
    char line[VL_MAXLINE];
    while (fgets(line, sizeof line, fp)) {
        if (!strncmp(line, tag, sizeof tag - 1)) {
            if (sscanf(line + sizeof tag - 1, "%f", &vf) != 1) {
                rc = TDF_ERR_ARG;
                break;
            }
            continue;
        }
        if (*line == '#') continue;
        if (i >= maxl) {
            rc = TDF_ERR_CAPACITY;
            break;
        }
        if (sscanf(line, "%lf %lf", &lines[i][0], &lines[i][1]) != 2) {
            rc = TDF_ERR_ARG;
            break;
        }
        i++;
    }
    fclose(fp);
    if (rc != TDF_OK) return rc;
    if (vf < 0.f) return TDF_ERR_ARG; /* the veto fraction is missing */
    *nvlines_all_inband = i;
    if (veto_fraction) *veto_fraction = vf;
    if (veto_flag) *numlines_band = i;
    return TDF_OK;
}
