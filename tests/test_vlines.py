"""CPU: veto lines of a band (common.c:155-553) through libtdfstat_search.so:
line files -> widened lines in radians of the band, the band file and its
read-back, the veto fraction, and the narrowdown edges."""
import math

import numpy as np
import pytest

from tdfstat_amd import _lib as L
from tdfstat_amd import search as TS
from tdfstat_amd import settings as S

FPO, DT = 100.09375, 2.0
B = 1.0 / (2 * DT)


def _sett(N=1000):
    st = S.SearchSettings(fpo=FPO, dt=DT, nod=2)
    st.B, st.N, st.nifo = B, N, 2
    st.M = np.eye(4).reshape(-1)
    return st


def _static_ssb(N, v=(0.0, 0.0, 0.0)):
    """DetSSB moving by v (samples per sample) each step: |step|^2 = |v|^2."""
    i = np.arange(N)[:, None]
    return (np.array([1.0, 2.0, 3.0])[None, :] + i * np.array(v)[None, :]).reshape(-1)


def _run(tmp_path, files, N=1000, v=((0, 0, 0), (0, 0, 0)), veto=1, nd=0.5):
    (tmp_path / "lines").mkdir(exist_ok=True)
    for name, text in files.items():
        (tmp_path / "lines" / name).write_text(text)
    ri = TS.run_info(_sett(N), band=372, det_names=("H1", "L1"), indir=str(tmp_path), narrowdown=nd)
    ri.veto_flag = veto
    ssb = [_static_ssb(N, v[0]), _static_ssb(N, v[1])]
    out = str(tmp_path / "vlines_0372.dat")
    return TS.band_vlines(ri, ssb, band_vl_file=out), out


def rad(f):
    return (f - FPO) / B * math.pi


def test_fdot_range_matches_settings():
    lib = L.load_search()
    import ctypes as C
    lo, hi = C.c_double(), C.c_double()
    lib.tdf_fdot_range(FPO, DT, C.byref(lo), C.byref(hi))
    assert (lo.value, hi.value) == S.get_fdot_range(FPO, DT)
    lib.tdf_fdot_range(250.0, DT, C.byref(lo), C.byref(hi))
    assert (lo.value, hi.value) == (2e-10, 2e-11)


def test_singlet_and_clipping(tmp_path):
    """A still detector (no broadening; fdotmax = 0 below 200 Hz): a singlet
    is [f - lw, f + rw] in radians; a line across the band edge is clipped;
    one outside the band is dropped; '%' lines are comments."""
    f1, f2 = FPO + 0.10, FPO + 0.245
    txt = (f"% comment\n{f1},0,0,0,0,0.001,0.002,a\n{f2},0,0,0,0,0.01,0.02,edge\n"
           f"{FPO + 0.5},0,0,0,0,0.001,0.001,outside\n")
    (lines, nb, nv, vf), out = _run(tmp_path, {"H1lines_test.csv": txt})
    assert nb == nv == 4  # 2 narrowdown + 2 veto lines (veto applied)
    np.testing.assert_allclose(lines[:2], [[0, math.pi / 2 - 0.5], [math.pi / 2 + 0.5, math.pi]])
    got = sorted(map(tuple, lines[2:]))
    want = sorted([(rad(f1 - 0.001), rad(f1 + 0.002)), (rad(f2 - 0.01), math.pi)])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    # the veto fraction is the covered part of [0, pi] of the veto lines alone
    cov = sum(b - a for a, b in want) / math.pi
    assert abs(vf - cov) < 1e-4
    # the band file reads back to the same lines and fraction
    rl, rnb, rnv, rvf = TS.read_band_vlines(out, 0.5, veto_flag=1)
    assert (rnb, rnv) == (nb, nv) and rvf == pytest.approx(vf, abs=5e-5)
    np.testing.assert_allclose(sorted(map(tuple, rl[2:])), sorted(map(tuple, lines[2:])), atol=1e-14)


def test_combs(tmp_path):
    """Type 1: harmonics offset + k*spacing with fixed widths; type 2: widths
    scale with k.  Only harmonics inside the band are kept."""
    sp, off = 0.05, FPO - 0.01
    txt = f"{sp},1,{off},0,10,0.001,0.001,fixed\n{sp},2,{off},0,10,0.0001,0.0002,scaling\n"
    (lines, nb, nv, vf), _ = _run(tmp_path, {"L1lines_c.csv": txt})
    want = []
    for k in range(11):
        lf = off + k * sp
        for wl, wr in ((0.001, 0.001), (k * 0.0001, k * 0.0002)):
            a, b = lf - wl, lf + wr
            if not (b < FPO or a > FPO + B):
                want.append((rad(max(a, FPO)), rad(min(b, FPO + B))))
    got = sorted(map(tuple, lines[nb - (nv - 2):]))
    assert nv - 2 == len(want)
    np.testing.assert_allclose(got, sorted(want), atol=1e-12)


def test_broadening_from_detector_motion(tmp_path):
    """A detector moving |v| samples per sample broadens a line at f by
    f*|v| on both sides (the fdot term vanishes below 200 Hz)."""
    v = (3e-5, 4e-5, 0.0)  # |v| = 5e-5
    f = FPO + 0.12
    (lines, nb, nv, vf), _ = _run(tmp_path, {"H1lines_m.csv": f"{f},0,0,0,0,0,0,moving\n"},
                                   v=(v, (0, 0, 0)))
    assert nv == 3
    np.testing.assert_allclose(lines[2], [rad(f - f * 5e-5), rad(f + f * 5e-5)], atol=1e-9)


def test_without_veto_flag_lines_are_listed_not_applied(tmp_path):
    (lines, nb, nv, vf), _ = _run(tmp_path, {"H1lines_x.csv": f"{FPO + 0.1},0,0,0,0,0.001,0.001,x\n"}, veto=0)
    assert nb == 2 and nv == 3


def test_veto_fraction_merges_overlaps():
    lib = L.load_search()
    lines = np.array([[0.5, 1.0], [0.8, 1.5], [2.0, 2.5], [2.5, 2.6]], dtype=np.float64)
    vf = lib.tdf_lines_veto_fraction(L.dptr(lines), 0, 4)
    assert vf == pytest.approx((1.0 + 0.6) / math.pi, abs=1e-6)
    assert lib.tdf_lines_veto_fraction(L.dptr(lines), 0, 0) == pytest.approx(0.0)
