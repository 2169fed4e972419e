"""LAS reader/writer restatement (oracle) pinned without a GPU, and the
product's host-side header parse and ldecimal against it.

Pins:
  * header and record layouts: images packed by tests/las_spec.py from the
    published LAS 1.2 / 1.4 specifications, with the field values known;
  * the point-count repair of LasHeader::openRead (las.cpp:354-402): cases
    worked out by hand, including the example of its comment (986 points:
    377 233 144 89 55 34 21 13 8 5 3 2 1 1 0);
  * degtobin / bintodeg, manygcd, ldecimal: oracle/_ref (the reference's own
    angle.cpp, manygcd.cpp, ldecimal.cpp) bit for bit;
  * joinPointFormat: the specification's feature table;
  * the file plan of CloudOutput::writeFiles and the duplicate merge of
    OctBuffer::put: brute-force restatements on small inputs.
"""
import ctypes as C
import struct

import numpy as np
import pytest

import las_spec
import oracle_las as OL
import oracle_lib as O
import wolkenbase_amd as W
from wolkenbase_amd import las as WL


def product_header(tmp_path, img):
    p = tmp_path / "h.las"
    p.write_bytes(img)
    return WL.read_header(p)


def oracle_header(img):
    return WL.header_dict(OL.parse(img))


# ---------------------------------------------------------------- headers
@pytest.mark.parametrize("version,fmt", [((1, 2), 1), ((1, 2), 3), ((1, 3), 1), ((1, 4), 1),
                                         ((1, 4), 6), ((1, 4), 10), ((1, 0), 0)])
def test_header_known_answers(tmp_path, version, fmt):
    rng = np.random.default_rng(fmt + 10 * version[1])
    f = las_spec.random_fields(rng, 7, fmt)
    for r in f:
        r["ret"] = 1 + r["ret"] % 5
    img = las_spec.image(fmt, f, version=version, scale=(0.001, 0.002, 0.01),
                         offset=(500000.0, 4000000.0, 100.0), extra=3, vlr_bytes=54, n_vlr=1)
    want_size = {0: 227, 2: 227, 3: 235, 4: 375}[version[1]]
    assert oracle_header(img) == product_header(tmp_path, img)
    for h in (oracle_header(img), product_header(tmp_path, img)):
        assert (h["version_major"], h["version_minor"]) == version
        assert h["header_size"] == want_size
        assert h["point_offset"] == want_size + 54
        assert h["n_vlr"] == 1
        assert h["point_format"] == fmt
        assert h["point_length"] == las_spec.LENGTH[fmt] + 3
        assert h["scale"] == [0.001, 0.002, 0.01]
        assert h["offset"] == [500000.0, 4000000.0, 100.0]
        assert h["system_id"] == "TEST" and h["software"] == "las_spec"
        assert (h["creation_day"], h["creation_year"], h["source_id"]) == (1, 2020, 7)
        if version[1] == 3:
            # a 1.3 header (235 bytes) is longer than 0xe3, so openRead also
            # reads the 1.4 counts, from bytes 235-375 that hold VLR and record
            # data here: they do not add up and no count is trusted
            assert h["n_points"][0] == 0
        else:
            assert h["n_points"][0] == 7
            by = [sum(1 for r in f if r["ret"] == k) for k in range(1, 6)]
            assert h["n_points"][1:6] == by
        xs = [500000.0 + 0.001 * r["xi"] for r in f]
        assert h["min"][0] == min(xs) and h["max"][0] == max(xs)
        # LasHeader::isValid: minor version 0 is not valid; formats >= 6 need
        # the 1.4 header
        assert h["valid"] == int(version[1] > 0 and (want_size > 0xe3 or fmt < 6))
        assert h["zipped"] == 0


def counts_image(version, fmt, legacy, legacy_by, n14=None, n14_by=None):
    return las_spec.header(version, fmt, las_spec.LENGTH[fmt], legacy, legacy_by,
                           (0.01,) * 3, (0.0,) * 3, (0,) * 3, (1,) * 3, n14=n14,
                           n14_by_ret=n14_by)


FIB = [377, 233, 144, 89, 55, 34, 21, 13, 8, 5, 3, 2, 1, 1, 0]


@pytest.mark.parametrize("case", [
    # consistent LAS 1.2 counts: kept
    dict(img=((1, 2), 1, 100, [60, 30, 10, 0, 0]), want=[100, 60, 30, 10, 0, 0]),
    # the comment's example in a LAS 1.4 file: the 64-bit counts are used
    dict(img=((1, 4), 6, 0, [0] * 5, 986, FIB), want=[986] + FIB),
    # a total but no per-return counts (photogrammetry): all first returns
    dict(img=((1, 2), 1, 100, [0] * 5), want=[100, 100, 0, 0, 0, 0]),
    dict(img=((1, 4), 6, 0, [0] * 5, 100, [0] * 15), want=[100, 100] + [0] * 14),
    # per-return counts that do not add up: no points are trusted
    dict(img=((1, 2), 1, 100, [60, 30, 0, 0, 0]), want=[0] * 6),
    # both sets present and equal: kept
    dict(img=((1, 4), 1, 50, [20, 30, 0, 0, 0], 50, [20, 30] + [0] * 13),
         want=[50, 20, 30] + [0] * 13),
    # both sets present and disagreeing: each check disowns the other set,
    # so neither is trusted
    dict(img=((1, 4), 1, 50, [20, 30, 0, 0, 0], 60, [30, 30] + [0] * 13), want=[0] * 6),
    # legacy counts zero in a 1.4 file with returns above 5: the 64-bit set
    dict(img=((1, 4), 1, 0, [0] * 5, 20, [10, 5, 2, 1, 1, 1] + [0] * 9),
         want=[20, 10, 5, 2, 1, 1, 1]),
])
def test_point_count_repair(tmp_path, case):
    img = counts_image(*case["img"])
    want = case["want"]
    for h in (oracle_header(img), product_header(tmp_path, img)):
        assert h["n_points"][:len(want)] == want, h["n_points"]


def test_header_fuzz_product_equals_oracle(tmp_path):
    # random bytes after a LASF signature, random count fields: the product's
    # host parse and the oracle's agree field by field
    rng = np.random.default_rng(5)
    for t in range(300):
        body = bytearray(rng.integers(0, 256, 400, dtype=np.uint8).tobytes())
        body[0:4] = b"LASF"
        if t % 2:
            body[94:96] = struct.pack("<H", [227, 235, 375][t % 3])
        if t % 3 == 0:
            body[107:131] = struct.pack("<6I", *rng.integers(0, 50, 6))
        if t % 5 == 0:
            body[247:375] = struct.pack("<16Q", *rng.integers(0, 50, 16))
        img = bytes(body[:int(rng.integers(100, 401))])
        assert product_header(tmp_path, img) == oracle_header(img)


def test_laszip_signature_detected(tmp_path):
    h = las_spec.header((1, 4), 1, 28, 0, [0] * 5, (0.01,) * 3, (0,) * 3, (0,) * 3, (1,) * 3,
                        n14=0, n14_by_ret=[0] * 15, point_offset=375 + 64)
    img = bytearray(h)
    img[375:383] = b"\0\0laszip"
    for hd in (oracle_header(bytes(img)), product_header(tmp_path, bytes(img))):
        assert hd["zipped"] == 1


# ---------------------------------------------------------------- records
def ref_degtobin(a):
    if O.ref_available():
        R = O.ref()
        R.ref_degtobin.argtypes = [C.c_double]
        R.ref_degtobin.restype = C.c_int
        return R.ref_degtobin(float(a))
    return OL.degtobin(a)


@pytest.mark.parametrize("fmt", range(11))
@pytest.mark.parametrize("unit", [1.0, 0.3048])
def test_record_decode_known_answers(fmt, unit):
    rng = np.random.default_rng(100 + fmt)
    f = las_spec.random_fields(rng, 40, fmt)
    img = las_spec.image(fmt, f, version=(1, 4), extra=5 if fmt % 3 == 0 else 0)
    h = OL.parse(img)
    pts = OL.read_points(img, h, unit)
    sc, of = (0.01, 0.01, 0.01), (1000.0, 2000.0, 0.0)
    for r, p in zip(f, pts):
        assert p.x == (of[0] + sc[0] * r["xi"]) * unit
        assert p.y == (of[1] + sc[1] * r["yi"]) * unit
        assert p.z == (of[2] + sc[2] * r["zi"]) * unit
        assert (p.intensity, p.ret, p.nret, p.scan_dir, p.edge) == \
            (r["intensity"], r["ret"], r["nret"], r["scan_dir"], r["edge"])
        assert (p.cls, p.cls_flags, p.user, p.source) == \
            (r["cls"], r["cls_flags"], r["user"], r["source"])
        assert p.channel == (r["channel"] if fmt >= 6 else 0)
        want_angle = ref_degtobin(r["scan_angle"] if fmt < 6 else r["scan_angle"] * 0.006)
        assert p.scan_angle == want_angle
        assert p.gps == (r["gps"] if fmt in las_spec.GPS else 0.0)
        assert (p.red, p.green, p.blue) == ((r["red"], r["green"], r["blue"])
                                            if fmt in las_spec.RGB else (0, 0, 0))
        assert p.nir == (r["nir"] if fmt in las_spec.NIR else 0)


@pytest.mark.parametrize("fmt", range(11))
def test_record_write_read_round_trip(fmt):
    # writePoint of a decoded record under the file's own transform gives
    # back the record (waveform fields are written as zeros; 1-byte scan
    # angle ranks and the 0.006-degree shorts survive the binary angle)
    rng = np.random.default_rng(200 + fmt)
    f = las_spec.random_fields(rng, 60, fmt)
    for r in f:
        r["wave"] = (0, 0, 0, 0.0, 0.0, 0.0, 0.0)
    img = las_spec.image(fmt, f, version=(1, 4))
    h = OL.parse(img)
    pts = OL.read_points(img, h)
    L = OL.lib()
    out = OL.Out()
    sc = np.array(list(h.scale))
    of = np.array(list(h.offset))
    L.wbo_las_out_init(C.byref(out), fmt, sc.ctypes.data_as(OL._dp), of.ctypes.data_as(OL._dp),
                       1.0)
    for i, p in enumerate(pts):
        rec = C.create_string_buffer(67)
        L.wbo_las_write_point(rec, C.byref(out), C.byref(p))
        off = h.point_offset + i * h.point_length
        assert rec.raw[:las_spec.LENGTH[fmt]] == img[off:off + las_spec.LENGTH[fmt]], (fmt, i)


def test_angle_functions_match_reference():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    R = O.ref()
    R.ref_degtobin.argtypes = [C.c_double]
    R.ref_degtobin.restype = C.c_int
    R.ref_bintodeg.argtypes = [C.c_int]
    R.ref_bintodeg.restype = C.c_double
    L = OL.lib()
    rng = np.random.default_rng(9)
    vals = list(range(-128, 128)) + list(np.arange(-32768, 32768, 7) * 0.006) + \
        list(rng.uniform(-1000, 1000, 3000)) + [179.99999999, -180.0, 180.0, 359.5, 540.0]
    for v in vals:
        assert L.wbo_degtobin(float(v)) == R.ref_degtobin(float(v)), v
    for a in list(rng.integers(-2**31, 2**31, 3000)) + [0, 1, -1, 2**31 - 1, -2**31]:
        assert L.wbo_bintodeg(int(a)) == R.ref_bintodeg(int(a))


# ---------------------------------------------------------------- formats, scales
def test_join_point_format_feature_table():
    # LAS 1.4 R15 tables: 1 = 0 + GPS, 2 = 0 + RGB, 3 = GPS + RGB, 4/5 add
    # waves to 1/3, 6-10 are the extended records (7 RGB, 8 RGB + NIR,
    # 9 waves, 10 RGB + NIR + waves)
    cases = {(0,): 0, (1,): 1, (1, 2): 3, (0, 3): 3, (4, 2): 5, (0, 6): 6, (2, 6): 7,
             (3, 6): 7, (7, 8): 8, (4, 6): 9, (9, 7): 10, (5, 8): 10, (): 0, (1, 1, 1): 1}
    for fs, want in cases.items():
        assert OL.join_point_format(list(fs)) == want, fs


def test_manygcd_and_combine_scales():
    L = OL.lib()
    rng = np.random.default_rng(3)
    ref = O.ref() if O.ref_available() else None
    if ref is not None:
        ref.ref_manygcd.argtypes = [OL._dp, C.c_int, C.c_double]
        ref.ref_manygcd.restype = C.c_double
    for t in range(2000):
        n = int(rng.integers(1, 6))
        s = 10.0 ** -int(rng.integers(1, 4))
        v = (rng.integers(1, 10**6, n) * s * rng.choice([1, -1], n)).astype(np.float64)
        if t % 3 == 0:
            v = v + rng.standard_normal(n) * s * 1e-9
        got = L.wbo_manygcd(v.ctypes.data_as(OL._dp), n, s / 1e6)
        if ref is not None:
            assert got == ref.ref_manygcd(v.ctypes.data_as(OL._dp), n, s / 1e6)
    v = np.array([0.003, 0.005])
    assert abs(L.wbo_manygcd(v.ctypes.data_as(OL._dp), 2, 1e-9) - 0.001) < 1e-12
    assert OL.combine1_scale([0.01], [500000.0]) == 0.01
    assert OL.combine1_scale([0.01, 0.001], [0.0, 0.0]) == pytest.approx(0.001, rel=1e-12)
    # offsets off the scale's lattice: no common spacing
    assert OL.combine1_scale([0.01], [0.005]) == 0.005 or OL.combine1_scale([0.01], [0.005]) == 0


def test_ldecimal_matches_reference():
    rng = np.random.default_rng(11)
    vals = [0.0, -0.0, 1.0, -1.0, 2.0, 0.1, 1e-5, 123456.789, 1e300, -1e-300, 5e-324,
            float("inf"), float("nan"), 100.0, 1000.0, 1e5, 0.001, 12.5, 1e21, 1e22]
    vals += list(rng.standard_normal(600) * 10.0 ** rng.integers(-8, 9, 600))
    vals += list(np.round(rng.uniform(-1e6, 1e6, 300), 3)) + list(rng.integers(-99999, 99999, 200) / 64.0)
    ref = O.ref() if O.ref_available() else None
    if ref is not None:
        ref.ref_ldecimal.argtypes = [C.c_double, C.c_double, C.c_char_p, C.c_int]
    for tol in (0.0, 1e-3, 0.5):
        for v in vals:
            o = OL.ldecimal(v, tol)
            assert W.ldecimal(v, tol) == o, (v, tol)
            if ref is not None:
                b = C.create_string_buffer(128)
                ref.ref_ldecimal(float(v), tol, b, 128)
                assert b.value.decode() == o, (v, tol)
    # the forms ldecimal gives: no leading zero, no needless exponent
    assert [OL.ldecimal(v) for v in (0.5, 100.0, 1e6, -0.25, 1e-7, 2.0)] == \
        [".5", "100", "1e6", "-.25", "1e-7", "2"]


# ---------------------------------------------------------------- writer plan, dedupe
def brute_plan(leaf_of, cls, ppf, separate):
    totals = np.bincount(cls, minlength=256)
    grand = int(totals.sum())
    keys = [k for k in range(256) if totals[k]] if separate else [0]
    nfiles = {}
    for k in keys:
        t = totals[k] if separate else grand
        nfiles[k] = (t + ppf - 1) // ppf if ppf else 1
    fill = {k: [0] * nfiles[k] for k in keys}
    nxt = {k: 0 for k in keys}
    file_of, pos_of = [], []
    i = 0
    n = len(leaf_of)
    while i < n:
        j = i
        while j < n and leaf_of[j] == leaf_of[i]:
            j += 1
        for k in keys:
            mn = grand
            for q in range(nfiles[k]):
                if fill[k][q] < mn:
                    nxt[k], mn = q, fill[k][q]
        for m in range(i, j):
            k = int(cls[m]) if separate else 0
            file_of.append(nxt[k])
            pos_of.append(fill[k][nxt[k]])
            fill[k][nxt[k]] += 1
        i = j
    return nfiles, file_of, pos_of


@pytest.mark.parametrize("separate", [True, False])
def test_output_plan_matches_brute_force(separate):
    rng = np.random.default_rng(21)
    for t in range(40):
        nl = int(rng.integers(1, 60))
        counts = rng.integers(1, 40, nl)
        leaf_of = np.repeat(np.arange(nl), counts)
        cls = rng.choice([1, 2, 2, 7], len(leaf_of)).astype(np.uint8)
        ppf = int(rng.choice([0, 7, 50, 100000]))
        totals = np.bincount(cls, minlength=256)
        ndig, nfiles, file_of, pos_of = OL.output_plan(leaf_of, cls, totals, ppf, separate)
        bn, bf, bp = brute_plan(leaf_of, cls, ppf, separate)
        assert {k: int(nfiles[k]) for k in range(256) if nfiles[k]} == bn
        assert list(file_of) == bf and list(pos_of) == bp
        if ppf:
            q = (len(cls) + ppf - 1) // ppf
            q = max(q - 1, 1)
            assert ndig == len(str(q))


def test_dedupe_matches_brute_force():
    rng = np.random.default_rng(4)
    for t in range(30):
        n = int(rng.integers(0, 300))
        base = rng.integers(0, 20, (n, 3)).astype(np.float64) * 0.5
        if n:
            base[rng.integers(0, n, n // 5)] = -0.0  # signed zeros compare equal
        nd, first, last = OL.dedupe(base[:, 0], base[:, 1], base[:, 2])
        seen = {}
        for i in range(n):
            key = tuple(float(v) + 0.0 for v in base[i])
            seen.setdefault(key, []).append(i)
        for i in range(n):
            grp = seen[tuple(float(v) + 0.0 for v in base[i])]
            assert first[i] == grp[0] and last[i] == grp[-1]
        assert nd == n - len(seen)
