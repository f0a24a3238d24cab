"""OIFITS loading / writing and FITS images (include/gpair_io.h) against the
OIFITS oracle (oracle/oifits_oracle.py, the reference's extraction with its
linear uv scan) -- CPU only.

The oracle is pinned by: the table sizes of the reference's own data file
(gpair1/trunk/2004contest1.oifits: 195 V2 rows, 130 T3 rows, one channel),
closure of every triangle's three legs through the stored points and signs,
threshold cases on both sides of 5e-5 with known answers, and files written
from known records whose uv points and leg references are fixed by construction.
Product and oracle must agree bit for bit (integer and float arrays)."""
import os

import numpy as np
import pytest

import fits_files as FF
import oifits_oracle as OO
import oracle as O
from gpair_amd import oifits, synth

# the reference's own data file (gpair1/trunk/2004contest1.oifits), kept as a fixture
REF_FILE = os.path.join(os.path.dirname(__file__), "golden", "2004contest1.oifits")
FIELDS = ("pow", "powerr", "bisamp", "bisamperr", "bisphs", "bisphserr", "u", "v", "wavelength", "bandwidth",
          "uvpnt", "sign")


def _same(prod, orc):
    for k in FIELDS:
        a, b = getattr(prod, k), orc[k]
        assert a.shape == b.shape, k
        assert a.dtype == b.dtype, k
        if a.dtype.kind == "f":
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k   # bit for bit
        else:
            assert np.array_equal(a, b), k


def _closure(d):
    """max over T3 of |sum_j sign_j uv[leg_j]| / max_j |uv[leg_j]|"""
    if d.nbis == 0:
        return 0.0
    uv = np.stack([d.u, d.v], 1).astype(np.float64)
    legs = uv[d.uvpnt] * d.sign[..., None]
    return float((np.abs(legs.sum(1)).max(1) / np.abs(legs).max((1, 2))).max())


@pytest.mark.skipif(not os.path.exists(REF_FILE), reason="reference data file absent")
def test_reference_data_file():
    hd = OO.parse(REF_FILE)
    rows = {h.get("EXTNAME"): h.get("NAXIS2") for h, _ in hd[1:]}
    assert rows["OI_VIS2"] == 195 and rows["OI_T3"] == 130
    d = oifits.read(REF_FILE)
    assert (d.npow, d.nbis, d.target_id) == (195, 130, 0)
    assert oifits.targets(REF_FILE) == [(0, "FKV0497")]
    # every leg found an existing point here: the legs are the V2 baselines
    assert d.nuv == 195 and np.all(d.u >= 0)
    assert _closure(d) < 2e-4
    _same(d, OO.read(REF_FILE))
    # the hash made a few comparisons per leg; the linear scan of the reference
    # stops at the first match, index l + 1 comparisons
    linear = int((d.uvpnt + 1).sum())
    print(f"uv comparisons: hashed {d.uv_compares}, linear scan {linear}")
    assert d.uv_compares <= 2 * 3 * d.nbis


def _write_cases(path):
    lam = 1.0e-6
    base_u = np.array([10.0, 25.0, 0.0, 40.0]) * lam * 1e6
    base_v = np.array([3.0, -7.0, 5.0, 0.0]) * lam * 1e6
    n = base_u.size
    # T3 legs in metres: leg1 = point 0 stretched by 4e-5 (match) or 6e-5 (new);
    # leg2 = the conjugate of point 1; leg3 follows from closure
    u1 = np.array([base_u[0] * (1 + 4e-5), base_u[0] * (1 + 6e-5), 0.0, 0.0])
    v1 = np.array([base_v[0] * (1 + 4e-5), base_v[0] * (1 + 6e-5), 5.0, -5.0])
    u2 = np.array([-base_u[1], -base_u[1], base_u[3], -base_u[3]])
    v2 = np.array([-base_v[1], -base_v[1], 0.0, 0.0])
    nt = u1.size
    FF.write(path, FF.primary(), FF.oi_target([7], ["STAR"]), FF.oi_wavelength("INS", [lam], [1e-8]),
             FF.oi_vis2("INS", [7] * n, base_u, base_v, np.full(n, 0.5), np.full(n, 0.01)),
             FF.oi_t3("INS", [7] * nt, u1, v1, u2, v2, np.full(nt, 0.1), np.full(nt, 0.01), np.zeros(nt),
                      np.ones(nt)))
    return path


def _write_multi(path):
    """two targets, two instruments of three channels, flags, a NaN T3 amplitude"""
    rng = np.random.default_rng(3)
    wa = np.array([1.5e-6, 1.65e-6, 1.8e-6])
    wb = np.array([0.7e-6, 0.8e-6, 0.9e-6])
    hd = [FF.primary(), FF.oi_target([1, 2], ["ALPHA", "BETA"]), FF.oi_wavelength("A", wa, wa * 0.02),
          FF.oi_wavelength("B", wb, wb * 0.02)]
    for ins, tgt, nrec in (("A", 1, 6), ("B", 1, 4), ("A", 2, 5)):
        u, v = rng.uniform(-80, 80, nrec), rng.uniform(-80, 80, nrec)
        flag = rng.random((nrec, 3)) < 0.2
        ids = [tgt] * nrec
        hd.append(FF.oi_vis2(ins, ids, u, v, rng.random((nrec, 3)), rng.random((nrec, 3)) * 0.1, flag,
                             data_form="E" if ins == "B" else "D"))
        # T3 over pairs of those baselines (legs match V2 points of the same channel)
        i = np.arange(nrec - 1)
        amp = rng.random((nrec - 1, 3))
        amp[0, 1] = np.nan
        hd.append(FF.oi_t3(ins, [tgt] * (nrec - 1), u[i], v[i], u[i + 1], v[i + 1], amp, amp * 0.1,
                           rng.uniform(-180, 180, (nrec - 1, 3)), np.ones((nrec - 1, 3)),
                           rng.random((nrec - 1, 3)) < 0.2))
    FF.write(path, *hd)
    return path


def test_threshold_conjugate_and_zero_cases(tmp_path):
    """legs that match a stored point within the relative threshold 5e-5 (same or
    conjugate direction), legs just outside it, and zero coordinates"""
    path = _write_cases(str(tmp_path / "cases.oifits"))
    n = 4
    d = oifits.read(path)
    _same(d, OO.read(path))
    assert d.uvpnt[0, 0] == 0 and d.sign[0, 0] == 1              # within 4e-5: point 0
    assert d.uvpnt[1, 0] >= n                                    # 6e-5 out: a new point
    assert d.uvpnt[0, 1] == 1 and d.sign[0, 1] == -1             # conjugate of point 1
    assert d.uvpnt[2, 0] == 2 and d.sign[2, 0] == 1              # (0, 5): point 2
    assert d.uvpnt[3, 0] == 2 and d.sign[3, 0] == -1             # (0, -5): its conjugate
    assert d.uvpnt[2, 1] == 3 and d.sign[2, 1] == 1              # (40, 0)
    assert d.uvpnt[3, 1] == 3 and d.sign[3, 1] == -1             # (-40, 0)
    assert _closure(d) < 2e-4
    # the oracle's comparison, the reference's arithmetic: both sides of the threshold
    assert OO.compare_uv(1.0 + 4e-5, 2.0, 1.0, 2.0) == 1
    assert OO.compare_uv(1.0 + 6e-5, 2.0, 1.0, 2.0) == 0
    assert OO.compare_uv(-1.0, -2.0 * (1 + 3e-5), 1.0, 2.0) == -1
    assert OO.compare_uv(0.0, 0.0, 0.0, 0.0) == 1
    assert OO.compare_uv(1e-30, 0.0, 0.0, 0.0) == 0


def test_selection_flags_band_and_nan_amplitudes(tmp_path):
    """two targets (tables kept by their first record's id), two instruments with
    three channels, flagged data, a waveband cut and a NaN T3 amplitude"""
    path = _write_multi(str(tmp_path / "multi.oifits"))
    assert oifits.targets(path) == [(1, "ALPHA"), (2, "BETA")]
    with pytest.raises(oifits.OifitsError, match="2 targets"):
        oifits.read(path)
    with pytest.raises(oifits.OifitsError, match="not in OI_TARGET"):
        oifits.read(path, target_id=5)
    for tid, lo, hi in ((1, -1, -1), (2, -1, -1), (1, 1000, 2000), (1, 600, 850)):
        d = oifits.read(path, tid, lo, hi)
        o = OO.read(path, tid, lo, hi)
        _same(d, o)
        assert _closure(d) < 2e-4
        assert np.all(d.u >= 0)
        if tid == 1 and lo < 0:
            assert np.any(np.isinf(d.bisamperr)) and np.any(d.bisamp == 1.0)   # the NaN amplitude
    with pytest.raises(oifits.OifitsError, match="no data"):
        oifits.read(path, 1, 100, 200)


def test_write_problem_round_trip(tmp_path):
    """a synthetic problem written as OIFITS and read back: the leg references and
    signs are the layout's, the uv points agree to float rounding, and init_data
    packs the problem's data"""
    prob = synth.make_problem("cfg1", seed=4, ntel=5, nch=3, nsnap=2)
    path = str(tmp_path / "prob.oifits")
    oifits.write_problem(path, prob)
    d = oifits.read(path)
    _same(d, OO.read(path))
    assert (d.npow, d.nbis, d.nuv) == (prob.npow, prob.nbis, prob.nuv)
    pv, pt = _file_order(prob)
    assert np.array_equal(d.uvpnt[pt], pv[prob.uvpnt]) and np.array_equal(d.sign[pt], prob.sign)
    assert np.abs(d.u[pv] - prob.u).max() <= 4e-7 * np.abs(prob.u).max()
    assert np.abs(d.v[pv] - prob.v).max() <= 4e-7 * np.abs(prob.v).max()
    data, err, phasor, ndof = oifits.pack(d)
    assert ndof == prob.ndof
    order = np.concatenate([pv, prob.npow + (2 * pt[:, None] + np.arange(2)).ravel()])
    assert np.allclose(data[order], prob.data, rtol=1e-6, atol=0)
    assert np.allclose(err[order], prob.err, rtol=1e-5, atol=0)
    assert np.abs(phasor[pt] - prob.phasor).max() < 1e-5
    # the same through the one-call set-up, with a Gaussian prior from set_model
    p2 = oifits.problem(path, prob.W, prob.pixellation, model_type=3, model_width=5.0)
    assert p2.nuv == prob.nuv and np.array_equal(p2.data, data) and abs(p2.model.sum() - prob.W ** 2) < 1.0


def _file_order(prob):
    """index in the file of each layout V2 point and T3: the layout orders
    [snapshot][channel][baseline | triangle], the file's records run over
    [snapshot][baseline | triangle] with the channels inside each record"""
    m = prob.meta
    s, c, b, t = m["nsnap"], m["nch"], m["nbase"], m["ntri"]
    S, Cc, B = np.meshgrid(np.arange(s), np.arange(c), np.arange(b), indexing="ij")
    pv = ((S * b + B) * c + Cc).ravel()
    S, Cc, T = np.meshgrid(np.arange(s), np.arange(c), np.arange(t), indexing="ij")
    pt = ((S * t + T) * c + Cc).ravel()
    return pv, pt


def test_larger_file_matches_the_layout(tmp_path):
    """9 telescopes, 6 channels, 12 snapshots (2,592 V2, 6,048 T3): the hash gives
    the layout's references; comparisons stay a few per leg"""
    prob = synth.make_problem("cfg1", seed=5, ntel=9, nch=6, nsnap=12)
    path = str(tmp_path / "big.oifits")
    oifits.write_problem(path, prob)
    d = oifits.read(path)
    pv, pt = _file_order(prob)
    assert np.array_equal(d.uvpnt[pt], pv[prob.uvpnt]) and np.array_equal(d.sign[pt], prob.sign)
    assert d.uv_compares <= 4 * 3 * d.nbis
    assert _closure(d) < 2e-4


def test_fits_images(tmp_path):
    rng = np.random.default_rng(2)
    img = rng.random((64, 64)).astype(np.float32)
    for bp in (-64, -32):
        path = str(tmp_path / f"img{bp}.fits")
        oifits.write_image(path, img, pixsize=0.25, bitpix=bp, datafile="data.oifits", target="STAR")
        back = oifits.read_image(path)
        assert back.shape == (64, 64) and np.array_equal(back, img)
        hdr = OO.parse(path)[0][0]
        assert hdr["BITPIX"] == bp and hdr["WIDTH"] == 64 and hdr["DATAFILE"] == "data.oifits"
        assert hdr["TARGET"] == "STAR" and np.float32(hdr["PIXSIZE"]) == np.float32(0.25)
    # integer images with BSCALE / BZERO
    q = rng.integers(-3000, 3000, (8, 16))
    path = str(tmp_path / "scaled.fits")
    FF.write(path, FF.image(q, 16, bscale=0.5, bzero=100.0))
    back = oifits.read_image(path)
    assert back.shape == (8, 16) and np.array_equal(back, (0.5 * q + 100.0).astype(np.float32))
    with pytest.raises(oifits.OifitsError):
        oifits.read_image(str(tmp_path / "missing.fits"))
    bad = tmp_path / "bad.fits"
    bad.write_bytes(b"not a fits file" * 300)
    with pytest.raises(oifits.OifitsError):
        oifits.read_image(str(bad))


@pytest.mark.skipif(not os.path.exists(O.REF_OIFITS), reason="reference loader not built (oracle/_ref)")
def test_reference_loader_agrees(tmp_path):
    """the reference's own loader (get_oi_fits_selection + get_oi_fits_data,
    gpair1/trunk/getoifits.c:65-714 with read_fits.c and cfitsio, compiled into
    oracle/_ref by oracle/Makefile) on every kind of test file -- the reference's
    data file, the threshold / conjugate / zero cases, two targets with two
    instruments, flags, band cuts and a NaN amplitude, and synthetic layouts
    written by gpair_oifits_write -- gives the same arrays as the product and the
    numpy restatement, bit for bit (uv order, first-match references, signs)"""
    cases = []
    if os.path.exists(REF_FILE):
        cases.append((REF_FILE, -1, -1.0, -1.0))
    cases.append((_write_cases(str(tmp_path / "cases.oifits")), -1, -1.0, -1.0))
    multi = _write_multi(str(tmp_path / "multi.oifits"))
    cases += [(multi, 1, -1.0, -1.0), (multi, 2, -1.0, -1.0), (multi, 1, 1000.0, 2000.0), (multi, 1, 600.0, 850.0)]
    for k, kw in enumerate((dict(ntel=5, nch=3, nsnap=2), dict(ntel=9, nch=6, nsnap=12))):
        prob = synth.make_problem("cfg1", seed=4 + k, **kw)
        path = str(tmp_path / f"prob{k}.oifits")
        oifits.write_problem(path, prob)
        cases.append((path, -1, -1.0, -1.0))
    for path, tid, lo, hi in cases:
        ref = O.ref_oifits_read(path, tid, lo, hi, workdir=str(tmp_path))
        d = oifits.read(path, tid, lo, hi)
        print(f"{os.path.basename(path)} target {tid} band {lo}-{hi}: {ref['pow'].size} V2, {ref['bisamp'].size} T3, "
              f"{ref['u'].size} uv", flush=True)
        _same(d, ref)
        if ref["bisamp"].size < 1000:      # the numpy linear scan is slow on the 6,048-T3 layout
            _same(d, OO.read(path, tid, lo, hi))
