"""GPU parity on uneven segments (the batched-segments configuration): three
detectors (H1/L1/V1), segments of 2, 4 and 6 days (N = 86164, 172328,
258492; nfft 2^17, 2^18, 2^18), gap masks covering 10-30 % of each
detector's samples and sigma per detector in [0.8, 1.5], all seeded.  For
each segment an SNR-12 injection is searched at its nearest grid point over
three spindowns; F and triggers match the oracle (tolerances as in
test_gpu_parity)."""
import numpy as np
import pytest

from oracle import oracle as O
from tdfstat_amd import _lib as L
from tdfstat_amd import search as TS
from tdfstat_amd import synth
from test_gpu_parity import PRECS, _compare_templates, _flush

pytestmark = pytest.mark.gpu

SEGMENTS = [2, 4, 6]


@pytest.mark.parametrize("nod", SEGMENTS)
@pytest.mark.parametrize("prec,tol", PRECS)
def test_uneven_segment(nod, prec, tol):
    _flush(f"segment nod={nod} prec={prec}")
    rng = np.random.default_rng(100 + nod)
    gaps = tuple(float(g) for g in rng.uniform(0.1, 0.3, 3))
    sig = tuple(float(s) for s in rng.uniform(0.8, 1.5, 3))
    inj = synth.random_injection(rng, 12.0)
    seg = synth.make_segment(nod=nod, det_names=("H1", "L1", "V1"), seed=100 + nod, gap_fracs=gaps,
                             sigmas=sig, injections=[inj])
    sett = seg.sett
    pm, mm, nn = synth.sky_to_grid(sett, inj.ra, inj.dec)
    mm, nn = float(round(mm)), float(round(nn))
    s0 = round(synth.spindown_to_grid(sett, inj.fdot, mm, nn))
    opts = TS.default_opts(sett, thr=20.0, sst=s0 - 1, spndr1=s0 + 1, sstep=1)
    orc = O.Oracle(sett, seg.dets)
    orc.sky_stage(pm, mm, nn, opts)
    with TS.Context(sett, seg.dets, prec=prec, batch=8) as ctx:
        ctx.autotune()
        sky, recs = ctx.job_core(pm, mm, nn, opts)
        assert len(recs) == 3
        worst = _compare_templates(ctx, orc, recs, opts, 3, tol, 20.0)
    _flush(f"  N={sett.N} nfftf={sett.nfftf} max rel F err={worst:.3e}, "
           f"{sum(len(r.ffstat) for r in recs)} triggers")
