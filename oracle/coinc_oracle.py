"""Test infrastructure only (imported by tests/): a plain-Python
restatement of the coincidence search between segments,
coincidences/coincidences.c:39-515 (select_goodcands at 430-515), with the
reference's float32 roundings.  It is the checker of libtdfstat_search.so's
tdf_coincidences, never part of the product path.

Ties: the reference's qsort leaves the order of equal keys open (records
inside a cell, coincidences of equal w and avg_snr); this restatement takes
record order inside a cell and discovery order for the final ties, as the
library does.
"""
import math

import numpy as np

f32 = np.float32


def good_candidates(seg, refr, cthr):
    """coincidences.c:430-515: per record the (f at refr, F) pairs kept."""
    out = []
    for r in seg["records"]:
        fdot_lin = f32(float(r.fdot) * math.pi * seg["dt"] ** 2)
        kept = []
        for f, F in np.asarray(r.ffstat, dtype=np.float32).reshape(-1, 2):
            if F < cthr:
                continue
            if f <= math.pi / 2 - seg["narrowdown"] or f >= math.pi / 2 + seg["narrowdown"]:
                continue
            if any(a <= f <= b for a, b in seg.get("lines", ())):
                continue
            f2 = f32(float(f) + 2.0 * float(fdot_lin) * seg["N"] * (refr - seg["seg"]))
            if f2 < 0 or f2 > math.pi:
                continue
            kept.append((f2, F))
        out.append(kept)
    return out


def coincidences(segments, refr, cthr, scalem=1, scalen=1, scales=4, scalef=32, mincoin=3, shift=-1):
    nseg = len(segments)
    recs0 = segments[0]["records"]
    good = [good_candidates(sg, refr, cthr) for sg in segments]
    nfftf = segments[-1]["nfftf"]
    nfc = nfftf // scalef
    res, uniq = [], []
    for ish in ([shift] if shift >= 0 else range(16)):
        sh = [(ish >> i) & 1 for i in range(4)]  # f, s, n, m
        shm, shn, shs, shf = (f32(sh[3] * 0.5), f32(sh[2] * 0.5), f32(sh[1] * 0.5), f32(sh[0] * 0.5))
        cells = {}
        for j, r in enumerate(recs0):
            key = (int(math.floor(f32(f32(r.m) / f32(scalem)) - shm)),
                   int(math.floor(f32(f32(r.n) / f32(scalen)) - shn)),
                   int(math.floor(f32(f32(r.s) / f32(scales)) - shs)))
            cells.setdefault(key, []).append(j)
        found, un = [], [0] * nseg
        for key in sorted(cells):
            best = {}  # (s, fic) -> [snr2, cidx, f, ntrigs]
            for cidx in cells[key]:
                for s in range(nseg):
                    for f, F in good[s][cidx]:
                        snr2 = f32(2.0 * float(F) - 4.0)
                        fic = int(math.floor(f32(float(f) / (math.pi / nfftf) / scalef - float(shf))))
                        if fic < 0 or fic >= nfc:
                            continue
                        b = best.setdefault((s, fic), [f32(0), -1, f32(0), 0])
                        b[3] += 1
                        if snr2 > b[0]:
                            b[0], b[1], b[2] = snr2, cidx, f
            for fic in sorted({fc for (_, fc) in best}):
                mem, ssnr, sf, sfd, sra, sdec = [], f32(0), f32(0), f32(0), f32(0), f32(0)
                for s in range(nseg):
                    b = best.get((s, fic))
                    if b is None or b[1] < 0:
                        continue
                    r = recs0[b[1]]
                    un[s] += 1
                    ssnr = f32(ssnr + b[0])
                    sf = f32(sf + b[2])
                    sfd = f32(sfd + f32(r.fdot))
                    sra = f32(sra + f32(r.ra))
                    sdec = f32(sdec + f32(r.dec))
                    mem.append((s, b[1], b[3]))
                w = len(mem)
                if w < mincoin:
                    continue
                found.append(dict(shift="".join(str(sh[3 - i]) for i in range(4)), w=w,
                                  avg_snr=float(np.sqrt(f32(ssnr / f32(w)))), avg_f=float(f32(sf / f32(w))),
                                  avg_fdot=float(f32(sfd / f32(w))), avg_ra=float(f32(sra / f32(w))),
                                  avg_dec=float(f32(sdec / f32(w))), members=mem))
        order = sorted(range(len(found)), key=lambda k: (-found[k]["w"], -found[k]["avg_snr"], k))
        res += [found[k] for k in order]
        uniq.append(un)
    return res, [sum(len(x) for x in g) for g in good], uniq
