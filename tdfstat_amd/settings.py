"""Host-side setup of a search segment (Python mirror of the reference host).

Restates, for the hot path's inputs:
  common.h:8-29          constants
  common.c:14-33         fpo, get_fdot_range
  common.c:42-80         read_xdat (zero count, sig2 = N/(N-Nzeros) * var)
  common.c:88-131        read_detssb (3N doubles + phir + epsm)
  settings.c:20-105      search_settings
  settings.c:178-213     detector geometry (H1, L1, V1)
  settings.c:226-248     rogcvir
  common.c:433-445       narrow_down_band
  init.c:148-197         read_grid (int fftpad + 16 doubles)
  auxi.c:245-260         var
"""
import math
import os
from dataclasses import dataclass, field

import numpy as np

C_OMEGA_R = 7.2921151467064e-5
C_SIDDAY = 2.0 * math.pi / C_OMEGA_R
C_TAIDAY = 86400.0
C_YEARSEC = 365.25 * C_TAIDAY
RAD_TO_DEG = 180.0 / math.pi
C_EPSMA = 84381.448 / 3600.0 / RAD_TO_DEG
C_SPEED_OF_LIGHT = 299792.458  # km/s
C_AU = 1.49597870691e8  # km
NAV = 4096        # settings.h:12
NAVFSTAT = 4096   # auxi.h:12

# settings.c:178-213: (latitude, longitude, height, orientation gamma), radians / m
DETECTORS = {
    "V1": ((43.0 + 37.0 / 60.0 + 53.0880 / 3600.0) / RAD_TO_DEG,
           (10.0 + 30.0 / 60.0 + 16.1885 / 3600.0) / RAD_TO_DEG,
           51.884,
           (135.0 - (19.0 + 25.0 / 60.0 + 57.96 / 3600.0)) / RAD_TO_DEG),
    "H1": ((46 + (27 + 18.528 / 60.0) / 60.0) / RAD_TO_DEG,
           -(119 + (24 + 27.5657 / 60.0) / 60.0) / RAD_TO_DEG,
           142.554,
           170.9994 / RAD_TO_DEG),
    "L1": ((30 + (33 + 46.4196 / 60.0) / 60.0) / RAD_TO_DEG,
           -(90 + (46 + 27.2654 / 60.0) / 60.0) / RAD_TO_DEG,
           -6.574,
           242.7165 / RAD_TO_DEG),
}
DET_ORDER = ("H1", "L1", "V1")  # discovery order (settings.c:122)


def c_round(x):
    """round() as redefined in auxi.h:13: floor(x + 0.5)."""
    return math.floor(x + 0.5)


def fpo(band, overlap, dt):
    """common.c:14-18"""
    return 10.0 + (1.0 - overlap) * band / (2 * dt)


def get_fdot_range(fpo_, dt):
    """common.c:22-33 -> (fdot_min, fdot_max) in Hz/s."""
    B = 1.0 / (2 * dt)
    if fpo_ < 200.0:
        return 2.0 * (fpo_ + B) / (2.0 * 1000.0 * C_YEARSEC), 0.0
    return 2e-10, 2e-11


@dataclass
class SearchSettings:
    """Field names follow Search_settings (struct.h:105-145)."""
    fpo: float
    dt: float
    nod: int
    fftpad: int = 1
    B: float = 0.0
    oms: float = 0.0
    omr: float = 0.0
    Smin: float = 0.0
    Smax: float = 0.0
    sepsm: float = math.sin(C_EPSMA)
    cepsm: float = math.cos(C_EPSMA)
    nfft: int = 0
    N: int = 0
    nfftf: int = 0
    nmax: int = 0
    nmin: int = 0
    s: int = 1
    nd: int = 2
    interpftpad: int = 2
    Ninterp: int = 0
    nifo: int = 0
    numlines_band: int = 0
    nvlines_all_inband: int = 0
    bufsize: int = 0
    dd: int = 0
    M: np.ndarray = field(default_factory=lambda: np.zeros(16))
    lines: list = field(default_factory=list)


def search_settings(sett, nav=NAV):
    """settings.c:20-105 (N from the sidereal day, nfft = 2^ceil(log2 N))."""
    dt = sett.dt
    sett.B = 0.5 / dt
    sett.oms = 2.0 * math.pi * sett.fpo * dt
    sett.omr = C_OMEGA_R * dt
    sett.N = int(c_round(sett.nod * C_SIDDAY / dt))
    sett.nfft = 1 << int(math.ceil(math.log(sett.N) / math.log(2.0)))
    fdotmin, fdotmax = get_fdot_range(sett.fpo, dt)
    sett.Smax = math.pi * fdotmin * dt * dt
    sett.Smin = math.pi * fdotmax * dt * dt
    sett.nd = 2
    sett.interpftpad = 2
    sett.s = 1
    sett.Ninterp = sett.interpftpad * sett.nfft
    sett.nfftf = sett.fftpad * sett.nfft
    sett.nmin = int(sett.fftpad * nav * sett.B)
    sett.nmax = int((sett.nfft // 2 - nav * sett.B) * sett.fftpad)
    dayfbins = int(1.0 / C_SIDDAY * sett.nfftf / (2 * sett.B))
    sett.dd = dayfbins - 1
    sett.numlines_band = 0
    return sett


def rogcvir(ephi, egam):
    """settings.c:226-248 -> amod c1..c9."""
    s2g, c2g = math.sin(2.0 * egam), math.cos(2.0 * egam)
    sp, cp = math.sin(ephi), math.cos(ephi)
    return np.array([
        0.25 * s2g * (1 + sp * sp),
        -0.5 * c2g * sp,
        0.5 * s2g * math.sin(2.0 * ephi),
        -c2g * cp,
        0.75 * s2g * cp * cp,
        c2g * sp,
        0.5 * s2g * (1.0 + sp * sp),
        c2g * cp,
        0.5 * s2g * math.sin(2.0 * ephi),
    ])


def var(x):
    """auxi.c:245-260: double accumulation, n-1 normalisation."""
    x = np.asarray(x, dtype=np.float64)
    mean = float(np.sum(x)) / x.size
    return float(np.sum((x - mean) ** 2)) / (x.size - 1)


def xdat_sig2(xDat):
    """read_xdat (common.c:69-78): (Nzeros, crf0, sig2)."""
    N = xDat.size
    nz = int(np.count_nonzero(xDat == 0))
    crf0 = N / (N - nz)
    return nz, crf0, crf0 * var(xDat)


def narrow_down_band(narrowdown):
    """common.c:433-445: two excluded edge ranges (radians)."""
    return [(0.0, math.pi / 2 - narrowdown), (math.pi / 2 + narrowdown, math.pi)]


def default_narrowdown(overlap):
    """init.c:136-139: narrowdown = (1-overlap)/2 * pi when not given."""
    return (1.0 - overlap) / 2.0 * math.pi


# ---- the reference's input file formats ----

def write_grid(path, fftpad, M):
    with open(path, "wb") as f:
        f.write(np.int32(fftpad).tobytes())
        f.write(np.asarray(M, dtype=np.float64).tobytes())


def read_grid(path):
    """init.c:182-186: int fftpad then double M[16]."""
    with open(path, "rb") as f:
        fftpad = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
        M = np.frombuffer(f.read(16 * 8), dtype=np.float64).copy()
    return fftpad, M


def write_detssb(path, DetSSB, phir, epsm):
    with open(path, "wb") as f:
        f.write(np.asarray(DetSSB, dtype=np.float64).tobytes())
        f.write(np.float64(phir).tobytes())
        f.write(np.float64(epsm).tobytes())


def read_detssb(path, N):
    """common.c:106-113 -> (DetSSB[3N], phir, epsm)."""
    a = np.fromfile(path, dtype=np.float64, count=3 * N + 2)
    return a[:3 * N].copy(), float(a[3 * N]), float(a[3 * N + 1])


def read_xdat(path, N, read_o3=False):
    """common.c:51-61: N float32 (or float64 with read_O3)."""
    if read_o3:
        return np.fromfile(path, dtype=np.float64, count=N).astype(np.float32)
    return np.fromfile(path, dtype=np.float32, count=N)


def segment_paths(indir, seg, band, det):
    d = os.path.join(indir, f"{seg:03d}", det)
    return (os.path.join(d, f"xdat_{seg:03d}_{band:04d}.bin"), os.path.join(d, "DetSSB.bin"),
            os.path.join(d, "starting_date"))
