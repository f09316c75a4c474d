"""Pins the CPU oracle (oracle/tdf_oracle.c) with checks that do not retype
its formulas: textbook/library routines (numpy.fft, brute-force DFT,
scipy's natural CubicSpline), the published closed form of the amplitude
modulation (JKS 1998 eqs 12-13, PRD 58 063001), the docs-held known
answers (docs/search_for_candidates.md:377-438), invariants of FStat, and
hand-built F vectors with known MAX_ALG-3 triggers."""
import json
import math
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle as O
from tdfstat_amd import search as TS
from tdfstat_amd import settings as S
from tdfstat_amd import synth


# ---------------- FFT ----------------
@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 1024, 1 << 14])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_vs_numpy(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    assert np.abs(O.fft(x, sign) - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    ref32 = np.abs(O.fftf(x.astype(np.complex64), sign) - ref).max() / np.abs(ref).max()
    assert ref32 < 1e-5


@pytest.mark.parametrize("n", [8, 32, 64])
def test_fft_vs_bruteforce_dft(n):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    np.testing.assert_allclose(O.fft(x, -1), W @ x, atol=1e-12)
    np.testing.assert_allclose(O.fft(x, +1), W.conj() @ x, atol=1e-12)


# ---------------- spline ----------------
@pytest.mark.parametrize("n", [5, 40, 257, 3000])
def test_spline_second_derivatives_vs_scipy(n):
    rng = np.random.default_rng(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y2 = O.spline(y)
    xs = np.arange(n)
    ref = (CubicSpline(xs, y.real, bc_type="natural")(xs, 2)
           + 1j * CubicSpline(xs, y.imag, bc_type="natural")(xs, 2))
    np.testing.assert_allclose(y2, ref, atol=1e-11)
    assert y2[0] == 0 and y2[-1] == 0


def test_splintpad_vs_scipy_and_boundaries():
    rng = np.random.default_rng(5)
    N = 400
    n = 2 * N
    ya = rng.standard_normal(n + 8) + 1j * rng.standard_normal(n + 8)  # reference array is longer
    shftf = 3.0 * np.sin(np.arange(N) / 50.0) + 0.37
    out = O.splintpad(ya, shftf, N, 2)
    xs = np.arange(n)
    cs_r = CubicSpline(xs, ya[:n].real, bc_type="natural")
    cs_i = CubicSpline(xs, ya[:n].imag, bc_type="natural")
    x = 2 * (np.arange(N) - shftf)
    inside = (x >= 0) & (x <= n - 1)
    assert (~inside).any() and inside.sum() > N - 20  # both regimes exercised
    np.testing.assert_allclose(out[inside], cs_r(x[inside]) + 1j * cs_i(x[inside]), atol=1e-11)
    assert np.all(out[~inside] == 0)
    # linear data is reproduced exactly by a natural spline
    lin = (0.5 + 0.25j) * np.arange(n + 8)
    out2 = O.splintpad(lin, shftf, N, 2)
    np.testing.assert_allclose(out2[inside], (0.5 + 0.25j) * x[inside], atol=1e-9)


# ---------------- FStat ----------------
@pytest.mark.parametrize("nav,length", [(16, 100), (16, 16 * 7), (4096, 4096 * 3 + 17)])
def test_fstat_block_means_become_two(nav, length):
    rng = np.random.default_rng(nav + length)
    F = rng.exponential(2.0, length) * rng.uniform(0.5, 3.0)
    G = O.fstat_norm(F, nav)
    nb = length // nav
    for j in range(nb):
        assert np.mean(G[j * nav:(j + 1) * nav]) == pytest.approx(2.0, rel=1e-12)
        # shape inside the block is preserved (F / mu)
        r = G[j * nav:(j + 1) * nav] / F[j * nav:(j + 1) * nav]
        assert np.ptp(r) <= 1e-12 * r.max()
    np.testing.assert_array_equal(G[nb * nav:], F[nb * nav:])  # tail untouched (auxi.c:351)


# ---------------- amplitude modulation (JKS 1998 eqs 12-13) ----------------
def _jks_ab(alpha, delta, lam, gam, phir, omr, t):
    """a(t), b(t) of Jaranowski, Krolak & Schutz 1998, eqs 12-13, written
    in their published trigonometric form with x = alpha - phir - Omega_r t."""
    x = alpha - phir - omr * t
    s2g, c2g = math.sin(2 * gam), math.cos(2 * gam)
    a = (1 / 16 * s2g * (3 - math.cos(2 * lam)) * (3 - math.cos(2 * delta)) * np.cos(2 * x)
         - 1 / 4 * c2g * math.sin(lam) * (3 - math.cos(2 * delta)) * np.sin(2 * x)
         + 1 / 4 * s2g * math.sin(2 * lam) * math.sin(2 * delta) * np.cos(x)
         - 1 / 2 * c2g * math.cos(lam) * math.sin(2 * delta) * np.sin(x)
         + 3 / 4 * s2g * math.cos(lam) ** 2 * math.cos(delta) ** 2)
    b = (c2g * math.sin(lam) * math.sin(delta) * np.cos(2 * x)
         + 1 / 4 * s2g * (3 - math.cos(2 * lam)) * math.sin(delta) * np.sin(2 * x)
         + c2g * math.cos(lam) * math.cos(delta) * np.cos(x)
         + 1 / 2 * s2g * math.sin(2 * lam) * math.cos(delta) * np.sin(x))
    return a, b


@pytest.mark.parametrize("det", ["H1", "L1", "V1"])
def test_modvir_matches_jks_closed_form(det):
    import ctypes as C
    lat, lon, h, gam = S.DETECTORS[det]
    rng = np.random.default_rng(len(det))
    N = 5000
    omr = S.C_OMEGA_R * 2.0
    x = rng.standard_normal(N).astype(np.float32)
    x[100:200] = 0.0
    for _ in range(3):
        alpha, delta, phir = rng.uniform(0, 2 * np.pi), rng.uniform(-1.5, 1.5), rng.uniform(0, 6.28)
        d = O.OrDet()
        d.xDat = x.ctypes.data_as(C.POINTER(C.c_float))
        d.cphir, d.sphir = math.cos(phir), math.sin(phir)
        for j, v in enumerate(O.rogcvir(lat, gam)):
            d.c[j] = v
        t = np.arange(N, dtype=np.float64)
        cm, sm = np.cos(omr * t), np.sin(omr * t)
        aa, bb = np.zeros(N), np.zeros(N)
        O.load().or_modvir(math.sin(alpha), math.cos(alpha), math.sin(delta), math.cos(delta), N, C.byref(d),
                           cm.ctypes.data_as(O._DP), sm.ctypes.data_as(O._DP), aa.ctypes.data_as(O._DP),
                           bb.ctypes.data_as(O._DP))
        ja, jb = _jks_ab(alpha, delta, lat, gam, phir, omr, t)
        nz = x != 0
        np.testing.assert_allclose(aa[nz], ja[nz], atol=1e-13)
        np.testing.assert_allclose(bb[nz], jb[nz], atol=1e-13)
        assert np.all(aa[~nz] == 0) and np.all(bb[~nz] == 0)


def test_rogcvir_python_host_matches_oracle():
    for det, (lat, lon, h, gam) in S.DETECTORS.items():
        np.testing.assert_allclose(S.rogcvir(lat, gam), O.rogcvir(lat, gam), rtol=0, atol=1e-16)


# ---------------- docs known answers (tests/golden/docs_known_answers.json) ----------------
with open(os.path.join(os.path.dirname(__file__), "golden", "docs_known_answers.json")) as _f:
    KAT = json.load(_f)
DOCS_M = np.array(KAT["grid_matrix"]["M"])


def test_search_settings_docs_kat():
    """search_settings of the documented example (band 17, overlap 0.0625,
    dt 0.5, nod 6, fftpad 1), docs/search_for_candidates.md:377-403."""
    k = KAT["search_settings"]
    i, o = k["input"], k["output"]
    sett = S.SearchSettings(fpo=S.fpo(i["band"], i["overlap"], i["dt"]), dt=i["dt"], nod=i["nod"],
                            fftpad=i["fftpad"])
    S.search_settings(sett)
    assert sett.fpo == o["fpo"]
    for name in ("N", "nfft", "nfftf", "nmin", "nmax", "dd", "Ninterp", "interpftpad"):
        assert getattr(sett, name) == o[name], name
    assert sett.Smax == pytest.approx(o["Smax"], rel=1e-5)
    assert sett.Smin == o["Smin"]
    assert sett.oms == pytest.approx(o["oms"], rel=1e-6)
    assert sett.omr == pytest.approx(o["omr"], rel=1e-5)
    assert sett.nfftf // 16 == o["bufsize_allsky"]  # jobcore.c:59-61
    assert math.sin(S.C_EPSMA) == pytest.approx(o["sepsm"], abs=1e-6)
    assert math.cos(S.C_EPSMA) == pytest.approx(o["cepsm"], abs=1e-6)


def test_trigger_record_docs_kat():
    """Record (m=-8, n=0, s=-136) of docs/search_for_candidates.md:430-438
    computed by the oracle's job_core scalar path (lin2ast, jobcore.c:242-262,
    437-447) from the documented grid matrix (6 printed digits, hence the
    tolerance)."""
    k = KAT["trigger_record"]
    i, o = k["input"], k["output"]
    seg = synth.make_segment(nod=6, dt=0.5, band=17, overlap=0.0625, N=6000, det_names=("H1", "L1"),
                             seed=3, M=DOCS_M)
    sett = seg.sett
    opts = TS.JobOpts(fr0=1000, fr1=3000, thr=1e9, sst=i["s"], spndr1=i["s"], sstep=1, bufsize=8192)
    orc = O.Oracle(sett, seg.dets)
    orc.sky_stage(i["pm"], float(i["m"]), float(i["n"]), opts)
    rec = orc.spindown(0, opts)["rec"]
    assert rec[0] == i["m"] and rec[1] == i["n"] and rec[2] == i["s"]
    assert rec[3] == pytest.approx(o["ra"], abs=3e-6)
    assert rec[4] == pytest.approx(o["dec"], abs=3e-6)
    assert rec[5] == pytest.approx(o["fdot"], rel=2e-5)


def test_two_day_grid_matrix_is_the_documented_one():
    np.testing.assert_array_equal(synth.M_2DAY_DT2, np.array(KAT["grid_2day_dt2"]["M"]))


def test_lin2ast_is_on_the_sphere_and_inverts_ast2lin():
    rng = np.random.default_rng(9)
    eps = S.C_EPSMA
    for _ in range(50):
        ra, dec = rng.uniform(0, 2 * np.pi), math.asin(rng.uniform(-1, 1))
        be1 = math.cos(eps) * math.sin(ra) * math.cos(dec) + math.sin(eps) * math.sin(dec)
        be2 = math.cos(ra) * math.cos(dec)
        got = [O.lin2ast(be1, be2, pm, math.sin(eps), math.cos(eps)) for pm in (1, 2)]
        ok = [abs(math.atan2(g[0], g[1]) % (2 * np.pi) - ra) < 1e-9 and abs(math.asin(g[2]) - dec) < 1e-9
              for g in got]
        assert any(ok)
        for sa, ca, sd, cd in got:
            assert sa * sa + ca * ca == pytest.approx(1.0, abs=1e-12)
            assert sd * sd + cd * cd == pytest.approx(1.0, abs=1e-12)


# ---------------- per-sky stage composition ----------------
def test_sky_stage_vs_numpy_scipy_composition():
    """xDatma/xDatmb of the oracle equal an independent composition:
    numpy demodulation, numpy zero-padded FFT interpolation (the spectrum
    move of jobcore.c:340-350 done with np.fft), scipy natural spline at
    2(i - shftf[i]), divided by sig2."""
    seg = synth.make_segment(nod=2, N=3000, det_names=("H1", "V1"), seed=4, gap_fracs=(0.1, 0.0))
    sett = seg.sett
    opts = TS.JobOpts(fr0=100, fr1=1500, sst=0, spndr1=0)
    orc = O.Oracle(sett, seg.dets)
    mm, nn = 1.0, -1.0
    r = orc.sky_stage(2, mm, nn, opts)
    sky = r["sky"]
    M = sett.M
    het0 = math.fmod(nn * M[8] + mm * M[12], M[0])
    assert sky.het0 == het0
    sa, ca = math.sin(sky.ra), math.cos(sky.ra)
    sd, cd = math.sin(sky.dec), math.cos(sky.dec)
    ns = np.array([ca * cd, sa * cd, sd])
    N, nfft, Ni = sett.N, sett.nfft, sett.Ninterp
    i = np.arange(N)
    aasum = bbsum = 0.0
    for k, d in enumerate(seg.dets):
        lat, lon, h, gam = S.DETECTORS[d.name]
        ja, jb = _jks_ab(sky.ra, sky.dec, lat, gam, d.phir, sett.omr, i.astype(float))
        nz = d.xDat != 0
        ja, jb = np.where(nz, ja, 0), np.where(nz, jb, 0)
        sh = d.DetSSB.reshape(-1, 3) @ ns
        shf = sh - sh[0]
        np.testing.assert_allclose(r["shftf"][k], shf, atol=1e-12)
        np.testing.assert_allclose(r["tmp1"][k], i.astype(float) ** 2 + 2 * i * sh, rtol=0, atol=1e-14 * N * N)
        ph = het0 * i + sett.oms * sh
        for amp, key in ((ja, "xDatma"), (jb, "xDatmb")):
            x = d.xDat.astype(np.float64) * amp * np.exp(-1j * ph)
            X = np.fft.fft(np.concatenate([x, np.zeros(nfft - N)]))
            Y = np.zeros(Ni, complex)
            ny = nfft // 2 + 1
            Y[:ny] = X[:ny]
            Y[ny + Ni - nfft:] = X[ny:]
            y = np.fft.ifft(Y) * 2  # interpftpad/Ninterp scaling of an unnormalised inverse
            n = 2 * N
            xs = 2 * (i - shf)
            cr = CubicSpline(np.arange(n), y[:n].real, bc_type="natural")
            ci = CubicSpline(np.arange(n), y[:n].imag, bc_type="natural")
            inside = (xs >= 0) & (xs <= n - 1)
            want = np.where(inside, cr(xs) + 1j * ci(xs), 0) / d.sig2
            got = r[key][k]
            # phase oms*shft ~ 3e5 rad: its last-bit rounding (6e-11 rad) bounds the agreement
            assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
        aasum += np.sum(ja ** 2) / d.sig2
        bbsum += np.sum(jb ** 2) / d.sig2
    assert sky.aa == pytest.approx(aasum, rel=1e-12)
    assert sky.bb == pytest.approx(bbsum, rel=1e-12)


# ---------------- spindown template: brute-force DFT ----------------
def test_spindown_vs_bruteforce_dft():
    """F of one template equals a direct O(N*nfftf) DFT of
    sum_det xDatm*exp(-i(het1 i + spnd tmp1)) on a tiny segment."""
    import ctypes as C
    rng = np.random.default_rng(21)
    N, nfftf, nifo = 300, 512, 2

    class _S:
        pass
    sett = _S()
    sett.N, sett.nfft, sett.Ninterp, sett.nfftf, sett.interpftpad, sett.nifo = N, 512, 1024, nfftf, 2, nifo
    sett.oms, sett.omr, sett.dt, sett.sepsm, sett.cepsm, sett.Smin, sett.Smax = 10.0, 1e-4, 2.0, 0.4, 0.9, 0, 1
    M = np.zeros(16)
    M[0] = 2 * np.pi / nfftf
    M[4], M[5], M[9], M[13] = 3.3e-3, 2.1e-7, 1e-8, -2e-8
    sett.M = M
    s = O.make_sett(sett)
    sky = O.OrSky()
    sky.visible, sky.het0, sky.aa, sky.bb, sky.smin = 1, 0.0123, 37.0, 51.0, -3.0
    xa = rng.standard_normal((nifo, N)) + 1j * rng.standard_normal((nifo, N))
    xb = rng.standard_normal((nifo, N)) + 1j * rng.standard_normal((nifo, N))
    tmp1 = np.stack([np.arange(N) ** 2 + 2 * np.arange(N) * rng.uniform(-5, 5, N) for _ in range(nifo)])
    opts = TS.JobOpts(fr0=10, fr1=400, thr=1e30, sst=-3, spndr1=3, sstep=1, bufsize=1024)
    o = O.make_opts(opts)
    F = np.zeros(nfftf)
    dummy_i = np.zeros(8, dtype=np.int32)
    dummy_f = np.zeros(8, dtype=np.float32)
    mm, nn, iss = 0.5, -1.5, 2
    n = O.load().or_spindown(C.byref(s), C.byref(sky), 0, mm, nn, iss, C.byref(o),
                             O._dp(np.ascontiguousarray(xa).view(np.float64)),
                             O._dp(np.ascontiguousarray(xb).view(np.float64)), O._dp(np.ascontiguousarray(tmp1)),
                             O._dp(F), 8, dummy_i.ctypes.data_as(O._IP), dummy_f.ctypes.data_as(O._FP),
                             dummy_f.ctypes.data_as(O._FP), dummy_f.ctypes.data_as(O._FP))
    assert n == 0
    ss = np.float32(np.float32(-3.0) + np.float32(iss * 1.0))
    spnd = float(ss) * M[5] + nn * M[9] + mm * M[13]
    het1 = math.fmod(float(ss) * M[4], M[0])
    het1 = het1 + M[0] if het1 < 0 else het1
    i = np.arange(N)
    sa = sum(xa[d] * np.exp(-1j * (het1 * i + spnd * tmp1[d])) for d in range(nifo))
    sb = sum(xb[d] * np.exp(-1j * (het1 * i + spnd * tmp1[d])) for d in range(nifo))
    k = np.arange(opts.fr0, opts.fr1 + 1)
    W = np.exp(-2j * np.pi * np.outer(k, i) / nfftf)
    Fw = np.abs(W @ sa) ** 2 / 37.0 + np.abs(W @ sb) ** 2 / 51.0
    np.testing.assert_allclose(F[opts.fr0:opts.fr1 + 1], Fw, rtol=1e-10)
    assert np.all(F[:opts.fr0] == 0) and np.all(F[opts.fr1 + 1:] == 0)


# ---------------- MAX_ALG 3 trigger selection ----------------
def _opts(fr0, fr1, thr, lines=(), bufsize=1000):
    return TS.JobOpts(fr0=fr0, fr1=fr1, thr=thr, lines=lines, bufsize=bufsize)


def test_triggers_hand_built():
    nfftf = 64
    F = np.full(nfftf, 1.0)
    F[5] = 30.0                       # isolated maximum
    F[10], F[11] = 25.0, 25.0         # tie plateau of 2: only the first (skip rule)
    F[20], F[21], F[22] = 22.0, 22.0, 22.0  # plateau of 3: 20 and 22 (alternating)
    F[30] = 20.0                      # exactly thr: accepted (F < thr rejects)
    F[31] = 19.999                    # below thr
    F[40] = 50.0                      # vetoed by a line
    F[2] = 99.0                       # fr0 + 1 = 2 is the first candidate
    F[61] = 99.0                      # fr1 - 1: outside the loop (i < fr1 - 1)
    F[60] = 98.0                      # fr1 - 2: last candidate, F[61] > F[60] so not a max
    sgnl0 = 0.1
    f = lambda ic: (2 * ic) / nfftf * math.pi + sgnl0  # noqa: E731
    lines = [(f(40) - 1e-6, f(40) + 1e-6)]
    n, ic, tf, tF = O.triggers(F, _opts(1, 62, 20.0, lines), nfftf, sgnl0)
    assert list(ic) == [2, 5, 10, 20, 22, 30]
    np.testing.assert_allclose(tf, [np.float32(f(k)) for k in ic])
    np.testing.assert_array_equal(tF, F[ic].astype(np.float32))


def test_triggers_overflow_returns_error():
    nfftf = 64
    F = np.zeros(nfftf)
    F[2:60:2] = 30.0
    n, *_ = O.triggers(F, _opts(0, 63, 20.0, bufsize=20), nfftf, 0.0)
    assert n == -1  # more than bufsize/2 = 10 triggers (jobcore.c:586-590)
    n, *_ = O.triggers(F, _opts(0, 63, 20.0, bufsize=200), nfftf, 0.0)
    assert n == 29
