"""OIFITS oracle -- TEST INFRASTRUCTURE ONLY (tests load it as the checker; the
product never imports it).

A second, independent reading of OIFITS files and a plain restatement of the
reference's extraction, gpair1/trunk/getoifits.c:65-357 (get_oi_fits_data), with
its linear scan over the stored uv points and compare_uv
(gpair1/trunk/getoifits.c:17-62) in the same float / double arithmetic:

* FITS parsing here maps each binary table onto a numpy big-endian structured
  dtype built from TFORMn (the product decodes cells one by one in C);
* target / band / flag selection, the +u flip, the NaN-amplitude rule and the
  first-match-in-index-order deduplication follow the reference line by line.

The reference loader itself links cfitsio (read_fits.c), an external library,
so it is not built here; this restatement is pinned instead by the checks in
tests/test_oifits.py (table sizes of the reference's own data file, triangle
closure of every T3's legs, threshold cases on both sides of 5e-5, round trips
of files with known records).
"""
from __future__ import annotations

import math

import numpy as np

BLOCK = 2880
_TYPES = {"L": "S1", "B": "u1", "I": ">i2", "J": ">i4", "K": ">i8", "A": "S", "E": ">f4", "D": ">f8"}


def _value(text):
    t = text.strip()
    if t.startswith("'"):
        out, i = [], 1
        while i < len(t):
            if t[i] == "'":
                if i + 1 < len(t) and t[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(t[i])
            i += 1
        return "".join(out).rstrip()
    t = t.split("/")[0].strip()
    if t in ("T", "F"):
        return t == "T"
    try:
        return int(t)
    except ValueError:
        return float(t.replace("D", "E"))


def parse(path):
    """[(header dict, data bytes)] of every HDU"""
    raw = open(path, "rb").read()
    pos, hdus = 0, []
    while pos + BLOCK <= len(raw):
        hdr, end = {}, False
        while not end:
            card = raw[pos:pos + 80].decode("ascii")
            pos += 80
            if card.startswith("END") and card[3:].strip() == "":
                end = True
            elif card[8:10] == "= ":
                hdr[card[:8].strip()] = _value(card[10:])
        pos = -(-pos // BLOCK) * BLOCK
        n = 1 if hdr.get("NAXIS", 0) > 0 else 0
        for a in range(1, hdr.get("NAXIS", 0) + 1):
            n *= hdr["NAXIS%d" % a]
        size = abs(hdr["BITPIX"]) // 8 * hdr.get("GCOUNT", 1) * (hdr.get("PCOUNT", 0) + n)
        hdus.append((hdr, raw[pos:pos + size]))
        pos += -(-size // BLOCK) * BLOCK
    return hdus


def table(hdr, data):
    """binary table as a numpy structured array"""
    fields = []
    for i in range(1, hdr["TFIELDS"] + 1):
        form = hdr["TFORM%d" % i].strip()
        k = 0
        while k < len(form) and form[k].isdigit():
            k += 1
        rep = int(form[:k]) if k else 1
        code = form[k]
        name = hdr.get("TTYPE%d" % i, "col%d" % i).upper()
        if code == "A":
            fields.append((name, "S%d" % rep))
        else:
            fields.append((name, _TYPES[code], (rep,)) if rep != 1 else (name, _TYPES[code]))
    dt = np.dtype(fields)
    assert dt.itemsize == hdr["NAXIS1"]
    return np.frombuffer(data[:dt.itemsize * hdr["NAXIS2"]], dtype=dt)


def _col(rec, name):
    return np.atleast_1d(rec[name])


def _f32(x):
    return np.float32(x)


def compare_uv(u, v, su, sv, thresh=np.float32(5.0e-5)):
    """gpair1/trunk/getoifits.c:17-62 with its evaluation types: float differences
    and sums, ((2.0 * diff) * diff) / sum in double (IEEE division), stored as float"""
    u, v, su, sv = _f32(u), _f32(v), _f32(su), _f32(sv)
    sq = _f32(thresh * thresh)

    def ratio(d, a, b):
        s = _f32(_f32(a * a) + _f32(b * b))
        with np.errstate(all="ignore"):
            return _f32(np.float64(2.0) * np.float64(d) * np.float64(d) / np.float64(s))
    with np.errstate(all="ignore"):
        du, dv, au, av = _f32(u - su), _f32(v - sv), _f32(u + su), _f32(v + sv)
    pu, pv, mu, mv = ratio(du, u, su), ratio(dv, v, sv), ratio(au, u, su), ratio(av, v, sv)
    if du == 0.0:
        pu = _f32(0.0)
    if dv == 0.0:
        pv = _f32(0.0)
    if au == 0.0:
        mu = _f32(0.0)
    if av == 0.0:
        mv = _f32(0.0)
    if pu < sq and pv < sq:
        return 1
    if mu < sq and mv < sq:
        return -1
    return 0


def read(path, target_id=-1, minband=-1.0, maxband=-1.0):
    """the reference's extraction with a linear uv scan; returns a dict of arrays"""
    hdus = parse(path)
    ext = [(h, d) for h, d in hdus[1:] if h.get("XTENSION") == "BINTABLE"]
    tgt = [table(h, d) for h, d in ext if h.get("EXTNAME") == "OI_TARGET"][0]
    if target_id < 0:
        assert tgt.size == 1
        target_id = int(tgt["TARGET_ID"][0])
    if minband < 0 or maxband <= minband:
        minband, maxband = 1.0, 50000.0
    minband, maxband = float(_f32(minband)), float(_f32(maxband))

    def wave_of(ins):
        for h, d in ext:
            if h.get("EXTNAME") == "OI_WAVELENGTH" and h.get("INSNAME") == ins:
                t = table(h, d)
                return t["EFF_WAVE"].astype(np.float32), t["EFF_BAND"].astype(np.float32)
        raise KeyError(ins)

    def ok(w):
        return float(w) * 1.0e9 > minband and float(w) * 1.0e9 < maxband

    pw, pe, uu, vv, wl, bw = [], [], [], [], [], []
    for h, d in ext:
        if h.get("EXTNAME") != "OI_VIS2" or h["NAXIS2"] < 1:
            continue
        t = table(h, d)
        if int(t["TARGET_ID"][0]) != target_id:
            continue
        ew, eb = wave_of(h["INSNAME"])
        for rec in t:
            data, err, flag = _col(rec, "VIS2DATA"), _col(rec, "VIS2ERR"), _col(rec, "FLAG")
            for k in range(data.size):
                if not ok(ew[k]) or flag[k] == b"T":
                    continue
                pw.append(_f32(float(data[k])))
                pe.append(_f32(float(err[k])))
                u = _f32(float(rec["UCOORD"]) / float(ew[k]))
                v = _f32(float(rec["VCOORD"]) / float(ew[k]))
                if u < 0.0:
                    u, v = -u, -v
                uu.append(u), vv.append(v), wl.append(ew[k]), bw.append(eb[k])
    amp, ampe, phs, phse, uvp, sgn = [], [], [], [], [], []
    for h, d in ext:
        if h.get("EXTNAME") != "OI_T3" or h["NAXIS2"] < 1:
            continue
        t = table(h, d)
        if int(t["TARGET_ID"][0]) != target_id:
            continue
        ew, eb = wave_of(h["INSNAME"])
        for rec in t:
            a, ae, p, pe_ = (_col(rec, n) for n in ("T3AMP", "T3AMPERR", "T3PHI", "T3PHIERR"))
            flag = _col(rec, "FLAG")
            for k in range(a.size):
                if not ok(ew[k]) or flag[k] == b"T":
                    continue
                if math.isnan(float(a[k])):
                    amp.append(_f32(1.0)), ampe.append(_f32(np.inf))   # (float)1e99
                else:
                    amp.append(_f32(float(a[k]))), ampe.append(_f32(float(ae[k])))
                phs.append(_f32(float(p[k]))), phse.append(_f32(float(pe_[k])))
                w = float(ew[k])
                l1 = (_f32(float(rec["U1COORD"]) / w), _f32(float(rec["V1COORD"]) / w))
                l2 = (_f32(float(rec["U2COORD"]) / w), _f32(float(rec["V2COORD"]) / w))
                l3 = (-_f32(l1[0] + l2[0]), -_f32(l1[1] + l2[1]))
                refs, signs = [], []
                for lu, lv in (l1, l2, l3):
                    found = -1
                    for i in range(len(uu)):          # linear scan, first match
                        m = compare_uv(lu, lv, uu[i], vv[i])
                        if m != 0 and ew[k] == wl[i]:
                            found, s = i, m
                            break
                    if found < 0:
                        found, s = len(uu), 1
                        if lu < 0.0:
                            lu, lv, s = -lu, -lv, -1
                        uu.append(lu), vv.append(lv), wl.append(ew[k]), bw.append(eb[k])
                    refs.append(found), signs.append(s)
                uvp.append(refs), sgn.append(signs)
    f = lambda x: np.array(x, np.float32)
    return dict(pow=f(pw), powerr=f(pe), bisamp=f(amp), bisamperr=f(ampe), bisphs=f(phs), bisphserr=f(phse),
                u=f(uu), v=f(vv), wavelength=f(wl), bandwidth=f(bw),
                uvpnt=np.array(uvp, np.int64).reshape(-1, 3), sign=np.array(sgn, np.int16).reshape(-1, 3),
                target_id=target_id)
