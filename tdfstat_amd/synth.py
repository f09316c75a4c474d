"""Seeded synthetic inputs for the search (the input side of the path).

* Gaussian noise per detector with its own sigma and zeroed gap masks
  (the role of genseg/gauss-xdat-mask.c, seeded instead of /dev/random).
* A synthetic detector ephemeris DetSSB: circular 1-AU orbit in the
  ecliptic plus sidereal rotation at the detector's latitude/longitude, in
  units of samples (light-travel time / dt), followed by phir and epsm
  (genseg/genseg_hdf.c:627-636 layout).
* Signal injection restating add_signal (init.c:365-504) in "amp" mode, with
  h0 chosen by the generator for a target network SNR (the reference's
  "snr" mode computes h0 with ifo[nifo], i.e. h0 = 0, init.c:462-464).
* grid.bin matrices: the 2-day dt=2 s normalised matrix published in
  docs/grid_generation.md:31-35, and a rescaled version for longer
  segments / zero padding.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import settings as S

# docs/grid_generation.md:31-35 (rows of the normalised grid matrix = M[0..15])
M_2DAY_DT2 = np.array([
    4.793689962143e-05, 0.0, 0.0, 0.0,
    -1.742438091399e-04, 2.161320668089e-09, 0.0, 0.0,
    -2.479053689156e-02, -1.684378936606e-09, 2.558634258126e+02, 0.0,
    9.606645586405e-03, 3.102095432504e-09, -1.321384079961e+02, 1.941709922254e+02,
])


def scaled_grid(N, nfftf, N0=86164, nfftf0=1 << 17, M0=M_2DAY_DT2):
    """Grid matrix for a segment of N samples and F-statistic FFT nfftf,
    scaled from the 2-day matrix: frequency entries by nfftf0/nfftf,
    spindown entries by (N0/N)^2, sky entries unchanged."""
    M = np.array(M0, dtype=np.float64)
    fs = nfftf0 / nfftf
    ss = (N0 / N) ** 2
    for k in (0, 4, 8, 12):
        M[k] *= fs
    for k in (5, 9, 13):
        M[k] *= ss
    return M


@dataclass
class Injection:
    snr: float
    f: float          # dimensionless frequency in the band, rad/sample (0..pi)
    fdot: float       # Hz/s
    ra: float
    dec: float
    iota: float
    psi: float
    phase: float
    h0: float = 0.0   # filled by inject()


@dataclass
class Detector:
    name: str
    xDat: np.ndarray
    DetSSB: np.ndarray
    phir: float
    epsm: float
    sigma: float
    amod: np.ndarray = field(default=None)
    sig2: float = 0.0
    start_time: float = 0.0

    @property
    def cphir(self):
        return math.cos(self.phir)

    @property
    def sphir(self):
        return math.sin(self.phir)


def ephemeris(name, N, dt, gmst0, orb0):
    """DetSSB[3N] in samples (equatorial frame) and phir for detector name."""
    lat, lon, _h, _g = S.DETECTORS[name]
    t = np.arange(N, dtype=np.float64) * dt
    om_orb = 2.0 * math.pi / S.C_YEARSEC
    lam = orb0 + om_orb * t
    eps = S.C_EPSMA
    au = S.C_AU / S.C_SPEED_OF_LIGHT  # light seconds
    xe, ye = au * np.cos(lam), au * np.sin(lam)
    theta0 = gmst0 + lon
    th = theta0 + S.C_OMEGA_R * t
    re = 6378.140 / S.C_SPEED_OF_LIGHT
    r = np.empty((N, 3))
    r[:, 0] = xe + re * math.cos(lat) * np.cos(th)
    r[:, 1] = ye * math.cos(eps) + re * math.cos(lat) * np.sin(th)
    r[:, 2] = ye * math.sin(eps) + re * math.sin(lat)
    return (r / dt).reshape(-1), theta0 % (2 * math.pi), eps


def modvir_np(sinal, cosal, sindel, cosdel, N, c, cphir, sphir, omr, xDat):
    """a(t), b(t) of settings.c:254-294 (used by the injection)."""
    t = np.arange(N)
    cm, sm = np.cos(omr * t), np.sin(omr * t)
    cosalfr = cosal * cphir + sinal * sphir
    sinalfr = sinal * cphir - cosal * sphir
    c2d = cosdel ** 2
    c2sd = sindel * cosdel
    cc = cosalfr * cm + sinalfr * sm
    ss = sinalfr * cm - cosalfr * sm
    c2s = 2.0 * cc ** 2
    cs = cc * ss
    aa = (c[0] * (2.0 - c2d) * c2s + c[1] * (2.0 - c2d) * 2.0 * cs + c[2] * c2sd * cc
          + c[3] * c2sd * ss - c[0] * (2.0 - c2d) + c[4] * c2d)
    bb = c[5] * sindel * c2s + c[6] * sindel * 2.0 * cs + c[7] * cosdel * cc + c[8] * cosdel * ss - c[5] * sindel
    nz = np.abs(xDat.astype(np.float64)) > np.finfo(np.float64).tiny
    return np.where(nz, aa, 0.0), np.where(nz, bb, 0.0)


def signal_template(sett, dets, inj):
    """Per-detector unit-amplitude signal (init.c:383-457 with reffr = seg)."""
    f_hz = sett.fpo + inj.f / (2 * math.pi * sett.dt)
    freqo0 = 2 * math.pi * f_hz * sett.dt
    freqo1 = math.pi * inj.fdot * sett.dt * sett.dt
    sa, ca, sd, cd = math.sin(inj.ra), math.cos(inj.ra), math.sin(inj.dec), math.cos(inj.dec)
    cosi = math.cos(inj.iota)
    cosip = (1.0 + cosi * cosi) / 2.0
    p2, ph = 2.0 * inj.psi, inj.phase
    amp = [math.cos(p2) * cosip * math.cos(ph) - math.sin(p2) * cosi * math.sin(ph),
           math.sin(p2) * cosip * math.cos(ph) + math.cos(p2) * cosi * math.sin(ph),
           -math.cos(p2) * cosip * math.sin(ph) - math.sin(p2) * cosi * math.cos(ph),
           -math.sin(p2) * cosip * math.sin(ph) + math.cos(p2) * cosi * math.cos(ph)]
    ns = np.array([ca * cd, sa * cd, sd])
    i = np.arange(sett.N, dtype=np.float64)
    out = []
    for d in dets:
        aa, bb = modvir_np(sa, ca, sd, cd, sett.N, d.amod, d.cphir, d.sphir, sett.omr, d.xDat)
        sh = d.DetSSB.reshape(-1, 3) @ ns
        ph_ = freqo0 * (i + sh) + freqo1 * (2.0 * sh * i + i * i + sh * sh)
        ph_ -= 2 * math.pi * sett.dt * sett.fpo * i
        out.append(amp[0] * aa * np.cos(ph_) + amp[1] * bb * np.cos(ph_)
                   + amp[2] * aa * np.sin(ph_) + amp[3] * bb * np.sin(ph_))
    return out


def inject(sett, dets, inj):
    """Add h0*signal where xDat != 0 (init.c:466-473); h0 from target SNR."""
    sig = signal_template(sett, dets, inj)
    rho2 = sum(float(np.sum(s[d.xDat != 0] ** 2)) / d.sigma ** 2 for s, d in zip(sig, dets))
    inj.h0 = inj.snr / math.sqrt(rho2) if rho2 > 0 else 0.0
    for s, d in zip(sig, dets):
        nz = d.xDat != 0
        x = d.xDat.astype(np.float64)
        x[nz] += inj.h0 * s[nz]
        d.xDat = x.astype(np.float32)
    return inj


def random_injection(rng, snr, fdot_range=(-1e-9, 0.0), f_range=(0.1 * math.pi, 0.9 * math.pi)):
    return Injection(snr=snr, f=rng.uniform(*f_range), fdot=rng.uniform(*fdot_range),
                     ra=rng.uniform(0, 2 * math.pi), dec=math.asin(rng.uniform(-1, 1)),
                     iota=math.acos(rng.uniform(-1, 1)), psi=rng.uniform(0, 2 * math.pi),
                     phase=rng.uniform(0, 2 * math.pi))


@dataclass
class Segment:
    sett: S.SearchSettings
    dets: list
    injections: list


def override_length(sett, N):
    """Settings of a segment of N samples instead of nod sidereal days
    (small tests): the lengths and band edges search_settings derives
    from N."""
    sett.N = int(N)
    sett.nfft = 1 << int(math.ceil(math.log2(sett.N)))
    sett.Ninterp = 2 * sett.nfft
    sett.nfftf = sett.fftpad * sett.nfft
    sett.nmin = int(sett.fftpad * S.NAV * sett.B)
    sett.nmax = int((sett.nfft // 2 - S.NAV * sett.B) * sett.fftpad)


def make_segment(nod=2, dt=2.0, band=372, overlap=0.03125, det_names=("H1", "L1"), sigmas=None,
                 gap_fracs=None, fftpad=1, seed=1, injections=(), M=None, N=None):
    """Seeded segment.  N overrides the sidereal-day length (small tests).
    sig2 is computed from the noise-and-gap data before injection, in the
    order main.c uses (init_arrays, then add_signal)."""
    rng = np.random.default_rng(seed)
    sett = S.SearchSettings(fpo=S.fpo(band, overlap, dt), dt=dt, nod=nod, fftpad=fftpad)
    S.search_settings(sett)
    if N is not None:
        override_length(sett, N)
    sett.nifo = len(det_names)
    sett.M = np.array(M if M is not None else scaled_grid(sett.N, sett.nfftf), dtype=np.float64)
    gmst0 = rng.uniform(0, 2 * math.pi)
    orb0 = rng.uniform(0, 2 * math.pi)
    dets = []
    for k, name in enumerate(det_names):
        sub = np.random.default_rng([seed, k])
        sigma = 1.0 if sigmas is None else float(sigmas[k])
        x = sub.normal(0.0, sigma, sett.N).astype(np.float32)
        if gap_fracs is not None and gap_fracs[k] > 0:
            # contiguous gaps of ~1/50 of the segment until the fraction is reached
            mask = np.zeros(sett.N, dtype=bool)
            glen = max(1, sett.N // 50)
            while mask.mean() < gap_fracs[k]:
                s0 = int(sub.integers(0, sett.N - glen))
                mask[s0:s0 + glen] = True
            x[mask] = 0.0
        ssb, phir, epsm = ephemeris(name, sett.N, dt, gmst0, orb0)
        lat, lon, h, gam = S.DETECTORS[name]
        d = Detector(name=name, xDat=x, DetSSB=ssb, phir=phir, epsm=epsm, sigma=sigma,
                     amod=S.rogcvir(lat, gam))
        d.sig2 = S.xdat_sig2(d.xDat)[2]
        dets.append(d)
    sett.sepsm = math.sin(dets[0].epsm)
    sett.cepsm = math.cos(dets[0].epsm)
    injs = [inject(sett, dets, inj) for inj in injections]
    return Segment(sett=sett, dets=dets, injections=injs)


def sky_to_grid(sett, ra, dec):
    """Fractional (mm, nn) and hemisphere of sky position (ra, dec):
    ast2lin (auxi.c:32-58) then the inverse of al = (nn*M10 + mm*M14,
    nn*M11 + mm*M15) (jobcore.c:243-244)."""
    eps = math.atan2(sett.sepsm, sett.cepsm)
    be0 = math.cos(eps) * math.sin(ra) * math.cos(dec) + math.sin(eps) * math.sin(dec)
    be1 = math.cos(ra) * math.cos(dec)
    root = math.sqrt(max(0.0, 1 - be0 ** 2 - be1 ** 2))
    # lin2ast (auxi.c:25) gives sindel = be0 sin(eps) -(2pm-3) root cos(eps)
    s1 = be0 * math.sin(eps) + root * math.cos(eps)
    s2 = be0 * math.sin(eps) - root * math.cos(eps)
    pm = 1 if abs(s1 - math.sin(dec)) <= abs(s2 - math.sin(dec)) else 2
    al = np.array([be0, be1]) * sett.oms
    M = sett.M
    A = np.array([[M[10], M[14]], [M[11], M[15]]])
    nn, mm = np.linalg.solve(A, al)
    return pm, float(mm), float(nn)


def spindown_to_grid(sett, fdot, mm, nn):
    """Fractional ss with spnd = ss*M5 + nn*M9 + mm*M13 = pi*fdot*dt^2."""
    M = sett.M
    spnd = math.pi * fdot * sett.dt ** 2
    return (spnd - nn * M[9] - mm * M[13]) / M[5]
