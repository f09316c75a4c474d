// Internal declarations shared by the kernels (tdf_kernels.hip) and the
// C-ABI implementation (tdf_capi.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "tdf_device.h"

namespace tdf {

constexpr int MAXDET = 8;
constexpr int TW_LO_LOG2 = 11;
constexpr int TW_LO = 1 << TW_LO_LOG2;

// LDS tile of a FFT pass holds ELEMS complex values; NT threads.
template <typename T>
struct FftCfg;
template <>
struct FftCfg<float> {
    static constexpr int ELEMS = 4096, NT = 256, IS_D = 0;
};
template <>
struct FftCfg<double> {
    static constexpr int ELEMS = 2048, NT = 128, IS_D = 1;
};

// natural bins per trigger / FStat tile (LDS <= 64 KB)
template <typename T>
struct TrigCfg {
    static constexpr int TB = sizeof(T) == 4 ? 8192 : 4096;
};

// spline taps: K two-sided taps, boundary decay over BK nodes
template <typename T>
struct SplineCfg;
template <>
struct SplineCfg<float> {
    static constexpr int K = 16, BK = 24;
};
template <>
struct SplineCfg<double> {
    static constexpr int K = 32, BK = 48;
};

struct DetParams {
    double c[9];
    double sig2, cphir, sphir;
};

struct SkyPrepArgs {
    const float *xDat;      // [nifo][N]
    const double *DetSSB;   // [nifo][3N]
    const DetParams *det;   // [nifo]
    double *shftf, *tmp1;   // [nifo][N] out
    double *partials;       // [nifo][2][nblk] out
    int N, nfft;
    double omr, oms, het0;
    double ns0, ns1, ns2;
    double c2d, c2sd, sindel, cosdel;
    double shft1[MAXDET], cosalfr[MAXDET], sinalfr[MAXDET];
};

// Four-step factorisation n = n1 * n2: pass 1 = n1 columns of length n2,
// pass 2 = n2 columns of length n1.
struct FftPlan {
    int n, n1, n2, l1, l2;
    const void *twL[2][16];  // [is_double][log2 L] -> W_L^m, m < L
    const void *tw_lo[2], *tw_hi[2];
};

struct TmplDesc {
    double het1, spnd, sgnl0;
};

struct SpinArgs {
    const void *xm;       // [nifo][2][N] complex T
    const double *tmp1;   // [nifo][N]
    const TmplDesc *desc; // [nslots]
    void *Z;              // [nslots][2][n]
    void *FT;             // [nslots][n]
    const double *sky;    // aa, bb
    int N, nifo, slot0, nslots, fr0, fr1;
};

struct FstatArgs {
    int n, n1, n2, fr0, fr1, nblk, nav, nblk_cap;
};

struct TrigArgs {
    int n, n1, n2, nfftf, fr0, fr1;
    double thr;
    int fstat_norm, nblk, nav, nblk_cap, numlines, cap;
    const double *lines;
    const TmplDesc *desc;
};

struct TrigEntry {
    int slot, ic;
    float f, F;
};

template <typename T>
hipError_t launch_sky_prep(const SkyPrepArgs &a, void *xab, int nblk, int nifo, hipStream_t st);
hipError_t launch_sky_reduce(const double *partials, int nblk, int nifo, const double *sig2, double *out,
                             hipStream_t st);
template <typename T>
hipError_t launch_fft_natural_to_T(const FftPlan &p, const void *in, void *tmp, void *outT, int npairs, int sign,
                                   int gap0, int gap1, double scale, hipStream_t st);
template <typename T>
hipError_t launch_transpose(const FftPlan &p, const void *inT, void *out, int nitems, size_t in_stride,
                            size_t out_stride, int split, int shift, int kmax, hipStream_t st);
template <typename T>
hipError_t launch_spline(const void *Y, void *Y2, int n, size_t ystride, size_t y2stride, int nitems,
                         const double *taps, hipStream_t st);
template <typename T>
hipError_t launch_splint(const void *Y, const void *Y2, const double *shftf, const double *sig2, void *xm, int N,
                         int n, int ipad, size_t ystride, size_t y2stride, int nitems, hipStream_t st);
template <typename T>
hipError_t launch_spin_pass1(const FftPlan &p, const SpinArgs &s, hipStream_t st);
template <typename T>
hipError_t launch_spin_pass2(const FftPlan &p, const SpinArgs &s, hipStream_t st);
template <typename T>
hipError_t launch_fstat_blocks(const void *FT, void *mu, const FstatArgs &a, int nslots, hipStream_t st);
template <typename T>
hipError_t launch_triggers(const void *FT, const void *mu, const TrigArgs &a, int nslots, TrigEntry *out,
                           int *counters, hipStream_t st);

}  // namespace tdf
