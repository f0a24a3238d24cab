"""Writes small FITS / OIFITS files for the I/O tests with numpy (a second
writer, independent of gpair_oifits_write): any number of targets and tables,
per-record target ids, chosen column types, BSCALE / BZERO images."""
from __future__ import annotations

import numpy as np

BLOCK = 2880


def _card(key, value, comment=None):
    if isinstance(value, bool):
        v = "%20s" % ("T" if value else "F")
    elif isinstance(value, int):
        v = "%20d" % value
    elif isinstance(value, float):
        v = "%20.13E" % value
    else:
        v = "'%-8s'" % str(value).replace("'", "''")
    text = "%-8s= %s" % (key, v)
    if comment:
        text += " / " + comment
    return text.ljust(80)[:80].encode("ascii")


def _header(cards):
    raw = b"".join(_card(*c) for c in cards) + b"END".ljust(80)
    return raw + b" " * (-len(raw) % BLOCK)


def _pad(data):
    return data + b"\0" * (-len(data) % BLOCK)


def primary(extend=True):
    return _header([("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 0), ("EXTEND", extend)])


def bintable(extname, columns, keywords=()):
    """columns: [(name, tform, numpy array [nrows] or [nrows, repeat])]"""
    nrows = len(columns[0][2])
    dt, cards = [], [("XTENSION", "BINTABLE"), ("BITPIX", 8), ("NAXIS", 2)]
    for name, form, _ in columns:
        code = form.lstrip("0123456789")
        rep = int(form[:len(form) - len(code)] or 1)
        base = {"I": ">i2", "J": ">i4", "K": ">i8", "E": ">f4", "D": ">f8", "L": "S1", "B": "u1"}.get(code)
        if code == "A":
            dt.append((name, "S%d" % rep))
        else:
            dt.append((name, base, (rep,)) if rep > 1 else (name, base))
    rec = np.zeros(nrows, dtype=np.dtype(dt))
    for name, form, arr in columns:
        if form.endswith("L"):
            arr = np.where(np.asarray(arr).astype(bool), b"T", b"F")
        rec[name] = arr
    cards += [("NAXIS1", rec.dtype.itemsize), ("NAXIS2", nrows), ("PCOUNT", 0), ("GCOUNT", 1),
              ("TFIELDS", len(columns))]
    for i, (name, form, _) in enumerate(columns):
        cards += [("TTYPE%d" % (i + 1), name), ("TFORM%d" % (i + 1), form)]
    cards += [("EXTNAME", extname)] + list(keywords)
    return _header(cards) + _pad(rec.tobytes())


def oi_target(ids, names):
    n = len(ids)
    z = np.zeros(n)
    cols = [("TARGET_ID", "1I", np.asarray(ids)), ("TARGET", "16A", np.array(names, "S16")),
            ("RAEP0", "1D", z), ("DECEP0", "1D", z), ("EQUINOX", "1E", z + 2000), ("RA_ERR", "1D", z),
            ("DEC_ERR", "1D", z), ("SYSVEL", "1D", z), ("VELTYP", "8A", np.array(["UNKNOWN"] * n, "S8")),
            ("VELDEF", "8A", np.array(["OPTICAL"] * n, "S8")), ("PMRA", "1D", z), ("PMDEC", "1D", z),
            ("PMRA_ERR", "1D", z), ("PMDEC_ERR", "1D", z), ("PARALLAX", "1E", z), ("PARA_ERR", "1E", z),
            ("SPECTYP", "16A", np.array(["UNKNOWN"] * n, "S16"))]
    return bintable("OI_TARGET", cols, [("OI_REVN", 1)])


def oi_wavelength(insname, wave, band):
    return bintable("OI_WAVELENGTH", [("EFF_WAVE", "1E", np.asarray(wave, np.float32)),
                                      ("EFF_BAND", "1E", np.asarray(band, np.float32))],
                    [("OI_REVN", 1), ("INSNAME", insname)])


def oi_vis2(insname, target_ids, u, v, data, err, flag=None, data_form="D"):
    n = len(u)
    nw = np.asarray(data).reshape(n, -1).shape[1]
    f = np.zeros((n, nw), bool) if flag is None else np.asarray(flag, bool).reshape(n, nw)
    z = np.zeros(n)
    cols = [("TARGET_ID", "1I", np.asarray(target_ids)), ("TIME", "1D", z), ("MJD", "1D", z),
            ("INT_TIME", "1D", z), ("VIS2DATA", "%d%s" % (nw, data_form), np.asarray(data).reshape(n, nw).squeeze(-1)
                                    if nw == 1 else np.asarray(data).reshape(n, nw)),
            ("VIS2ERR", "%d%s" % (nw, data_form), np.asarray(err).reshape(n, nw).squeeze(-1) if nw == 1
             else np.asarray(err).reshape(n, nw)),
            ("UCOORD", "1D", np.asarray(u)), ("VCOORD", "1D", np.asarray(v)),
            ("STA_INDEX", "2I", np.tile([1, 2], (n, 1))),
            ("FLAG", "%dL" % nw, f.squeeze(-1) if nw == 1 else f)]
    return bintable("OI_VIS2", cols, [("OI_REVN", 1), ("DATE-OBS", "2000-01-01"), ("INSNAME", insname)])


def oi_t3(insname, target_ids, u1, v1, u2, v2, amp, amperr, phi, phierr, flag=None):
    n = len(u1)
    nw = np.asarray(amp).reshape(n, -1).shape[1]
    f = np.zeros((n, nw), bool) if flag is None else np.asarray(flag, bool).reshape(n, nw)
    z = np.zeros(n)

    def m(a):
        a = np.asarray(a, np.float64).reshape(n, nw)
        return a.squeeze(-1) if nw == 1 else a
    cols = [("TARGET_ID", "1I", np.asarray(target_ids)), ("TIME", "1D", z), ("MJD", "1D", z),
            ("INT_TIME", "1D", z), ("T3AMP", "%dD" % nw, m(amp)), ("T3AMPERR", "%dD" % nw, m(amperr)),
            ("T3PHI", "%dD" % nw, m(phi)), ("T3PHIERR", "%dD" % nw, m(phierr)),
            ("U1COORD", "1D", np.asarray(u1)), ("V1COORD", "1D", np.asarray(v1)),
            ("U2COORD", "1D", np.asarray(u2)), ("V2COORD", "1D", np.asarray(v2)),
            ("STA_INDEX", "3I", np.tile([1, 2, 3], (n, 1))), ("FLAG", "%dL" % nw, f.squeeze(-1) if nw == 1 else f)]
    return bintable("OI_T3", cols, [("OI_REVN", 1), ("DATE-OBS", "2000-01-01"), ("INSNAME", insname)])


def image(arr, bitpix=16, bscale=1.0, bzero=0.0):
    a = np.asarray(arr)
    h, w = a.shape
    cards = [("SIMPLE", True), ("BITPIX", bitpix), ("NAXIS", 2), ("NAXIS1", w), ("NAXIS2", h)]
    if bscale != 1.0 or bzero != 0.0:
        cards += [("BSCALE", float(bscale)), ("BZERO", float(bzero))]
    dt = {8: "u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
    return _header(cards) + _pad(a.astype(dt).tobytes())


def write(path, *hdus):
    with open(path, "wb") as f:
        for h in hdus:
            f.write(h)
