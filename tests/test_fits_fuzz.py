"""FITS / OIFITS parsing under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only).  gpair_io.c reads untrusted bytes: the sizes in a header decide how
far the readers index into the file buffer.  tests/fits_fuzz_main.c hands every
file to every reader; it is built here with gcc -fsanitize=address,undefined
(host code only) and run over
  * valid files (the reference's own 2004contest1.oifits, files written by the
    numpy writer of tests/fits_files.py),
  * crafted headers whose sizes pass a naive bounds check (GCOUNT = 0, negative
    PCOUNT, NAXIS3 = 0, an overflowing NAXIS product, out-of-range values, a
    binary table longer than its data, truncated data, zero-repeat columns),
  * seeded mutations of the valid files (header values replaced by extreme
    numbers, random bytes in headers and data).
Refusing a file is fine; a sanitizer report (non-zero exit) fails the test."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import fits_files as FF

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_FILE = os.path.join(HERE, "golden", "2004contest1.oifits")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("ff") / "fits_fuzz")
    cmd = ["gcc", "-std=gnu11", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-o", out, os.path.join(HERE, "fits_fuzz_main.c"), os.path.join(ROOT, "gpair_amd", "csrc", "gpair_io.c"),
           "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("sanitizer build unavailable: " + r.stderr[-400:])
    return out


def _run(harness, paths):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    # one process per chunk keeps the failing file identifiable
    for i in range(0, len(paths), 64):
        chunk = paths[i:i + 64]
        r = subprocess.run([harness, *chunk], capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0, f"sanitizer error on one of {chunk}:\n{r.stderr[-3000:]}"


def _valid_files(d):
    paths = []
    if os.path.exists(REF_FILE):
        paths.append(REF_FILE)
    rng = np.random.default_rng(0)
    n, nw = 6, 3
    u, v = rng.uniform(-1e2, 1e2, n), rng.uniform(-1e2, 1e2, n)
    f = os.path.join(d, "small.oifits")
    FF.write(f, FF.primary(), FF.oi_target([1], ["STAR"]), FF.oi_wavelength("I", [1.6e-6, 1.7e-6, 1.8e-6], [1e-8] * 3),
             FF.oi_vis2("I", [1] * n, u, v, rng.random((n, nw)), 0.01 + rng.random((n, nw))),
             FF.oi_t3("I", [1] * 2, u[:2], v[:2], u[2:4], v[2:4], rng.random((2, nw)), rng.random((2, nw)),
                      rng.random((2, nw)), rng.random((2, nw))))
    paths.append(f)
    for bp in (8, 16, 32, 64, -32, -64):
        g = os.path.join(d, f"img{bp}.fits")
        FF.write(g, FF.image(rng.integers(0, 100, (64, 64)), bp))
        paths.append(g)
    return paths


def _crafted(d):
    cases = {
        "gcount0": [("SIMPLE", True), ("BITPIX", -32), ("NAXIS", 2), ("NAXIS1", 64), ("NAXIS2", 64),
                    ("GCOUNT", 0)],
        "pcount_neg": [("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 128), ("NAXIS2", 128),
                       ("PCOUNT", -16384)],
        "naxis3_zero": [("SIMPLE", True), ("BITPIX", -64), ("NAXIS", 3), ("NAXIS1", 64), ("NAXIS2", 64),
                        ("NAXIS3", 0)],
        "naxis_overflow": [("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 4294967296),
                           ("NAXIS2", 4294967296)],
        "naxis_huge_float": [("SIMPLE", True), ("BITPIX", 16), ("NAXIS", 2), ("NAXIS1", 1e300), ("NAXIS2", 2)],
        "naxis_nan": [("SIMPLE", True), ("BITPIX", 16), ("NAXIS", 2), ("NAXIS1", "NaN"), ("NAXIS2", 2)],
        "bitpix_bad": [("SIMPLE", True), ("BITPIX", 7), ("NAXIS", 2), ("NAXIS1", 4), ("NAXIS2", 4)],
        "naxis_neg": [("SIMPLE", True), ("BITPIX", 8), ("NAXIS", -2)],
        "truncated": [("SIMPLE", True), ("BITPIX", -64), ("NAXIS", 2), ("NAXIS1", 512), ("NAXIS2", 512)],
    }
    paths = []
    for name, cards in cases.items():
        # header only (a data block would be parsed as the next header and refused):
        # the sizes alone must stop the readers
        p = os.path.join(d, name + ".fits")
        with open(p, "wb") as fh:
            fh.write(FF._header(cards) + (b"\0" * 100 if name == "truncated" else b""))
        paths.append(p)
    # binary tables: more rows than data, a huge repeat count, zero-repeat columns
    tab = [("XTENSION", "BINTABLE"), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 16), ("NAXIS2", 1 << 40),
           ("PCOUNT", 0), ("GCOUNT", 1), ("TFIELDS", 2), ("TTYPE1", "EFF_WAVE"), ("TFORM1", "1D"),
           ("TTYPE2", "EFF_BAND"), ("TFORM2", "1D"), ("EXTNAME", "OI_WAVELENGTH"), ("INSNAME", "I")]
    rep = [("XTENSION", "BINTABLE"), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 8), ("NAXIS2", 2),
           ("PCOUNT", 0), ("GCOUNT", 1), ("TFIELDS", 1), ("TTYPE1", "TARGET_ID"), ("TFORM1", "99999999999I"),
           ("EXTNAME", "OI_TARGET")]
    zero = [("XTENSION", "BINTABLE"), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", 8), ("NAXIS2", 3),
            ("PCOUNT", 0), ("GCOUNT", 1), ("TFIELDS", 3), ("TTYPE1", "TARGET_ID"), ("TFORM1", "0I"),
            ("TTYPE2", "TARGET"), ("TFORM2", "0A"), ("TTYPE3", "X"), ("TFORM3", "1D"), ("EXTNAME", "OI_TARGET")]
    for name, cards, nbytes in (("table_rows", tab, 32), ("table_repeat", rep, 16), ("table_zero_repeat", zero, 24)):
        p = os.path.join(d, name + ".fits")
        with open(p, "wb") as fh:
            fh.write(FF.primary() + FF._header(cards) + FF._pad(b"\x01" * nbytes))
        paths.append(p)
    return paths


def _mutants(d, sources, count, seed):
    rng = np.random.default_rng(seed)
    values = [b"0", b"-1", b"1", b"-2147483649", b"2147483648", b"9223372036854775807", b"1E300", b"-1E300",
              b"NaN", b"3", b"4294967296", b"65536"]
    paths = []
    blobs = [bytearray(open(s, "rb").read()) for s in sources]
    for k in range(count):
        b = bytearray(blobs[k % len(blobs)])
        cards = [i for i in range(0, min(len(b), 2880 * 40), 80) if b[i + 8:i + 10] == b"= "]
        for _ in range(int(rng.integers(1, 4))):
            kind = rng.integers(0, 3)
            if kind == 0 and cards:        # a header value replaced
                i = cards[int(rng.integers(0, len(cards)))]
                b[i + 10:i + 30] = values[int(rng.integers(0, len(values)))].rjust(20)
            elif kind == 1:                # a random header byte
                i = int(rng.integers(0, min(len(b), 2880 * 4)))
                b[i] = int(rng.integers(32, 127))
            else:                          # a random data byte
                i = int(rng.integers(0, len(b)))
                b[i] = int(rng.integers(0, 256))
        if rng.random() < 0.2:
            del b[int(rng.integers(2880, max(len(b), 2881))):]   # truncated
        p = os.path.join(d, f"mut{k:04d}.fits")
        with open(p, "wb") as fh:
            fh.write(bytes(b))
        paths.append(p)
    return paths


def test_valid_and_crafted_files(harness, tmp_path):
    _run(harness, _valid_files(str(tmp_path)) + _crafted(str(tmp_path)))


def test_crafted_headers_are_refused(tmp_path):
    """the library refuses the crafted headers with an error instead of reading past its buffer"""
    from gpair_amd import oifits
    for p in _crafted(str(tmp_path)):
        name = os.path.basename(p)
        if name.startswith("table"):
            continue
        with pytest.raises(oifits.OifitsError):
            oifits.read_image(p)


def test_mutated_files(harness, tmp_path):
    src = _valid_files(str(tmp_path))
    _run(harness, _mutants(str(tmp_path), src, 600, seed=1))
