"""Host-side logic of the library, without a GPU: ABI exports, the variation
database and DSL front end, the packed layout, JSON / flam3 import and the
offline hiprtc compile of generated kernels."""
import base64
import ctypes as C
import json
import re
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle
from refraktor_amd import _lib
from refraktor_amd import flame as fl
from refraktor_amd import render

ROOT = Path(__file__).resolve().parents[1]
REFERENCE_CONFIG = Path("/root/reference/config")
# the reference's flam3 test flame (assets/flames_test/), kept as data
REFERENCE_FLAM3 = Path(__file__).resolve().parent / "golden" / "electricsheep.247.50049.flam3"


def test_header_symbols_exported():
    header = (ROOT / "include" / "refrakt.h").read_text()
    declared = set(re.findall(r"\b(rfk_[a-z0-9_]+)\s*\(", header))
    lib = C.CDLL(str(_lib.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), f"{name} declared but not exported"
    assert declared == set(_lib.EXPORTED_SYMBOLS), declared ^ set(_lib.EXPORTED_SYMBOLS)


def test_product_does_not_touch_oracle():
    for p in list((ROOT / "refraktor_amd").rglob("*.py")) + list((ROOT / "refraktor_amd").rglob("*.cpp")) + \
            list((ROOT / "refraktor_amd").rglob("*.hip")) + list((ROOT / "refraktor_amd").rglob("*.h")):
        text = p.read_text()
        assert "import oracle" not in text and "from oracle" not in text and "liboracle" not in text, p


def test_own_database_loads(db):
    assert db.errors == []
    names = db.variations
    for v in fl.CONFIG_VARIATIONS + ["linear", "sinusoidal", "spherical", "fisheye"]:
        assert v in names
    assert names == sorted(names)


@pytest.mark.skipif(not REFERENCE_CONFIG.exists(), reason="reference tree not mounted")
def test_reference_database_loads_unchanged():
    """drop-in: librefrakt's own variation files parse and type-check"""
    rdb = render.VariationDB(REFERENCE_CONFIG)
    assert rdb.errors == []
    assert len(rdb.variations) == 99


def _db_from(tmp_path, variations: str, common: str = "sumsq: p.x * p.x + p.y * p.y\n"):
    (tmp_path / "variations_new.yml").write_text(variations)
    (tmp_path / "common.yml").write_text(common)
    return render.VariationDB(tmp_path)


def test_dsl_errors_are_reported_and_skipped(tmp_path):
    src = """
good: result += weight * p;
syntax: result += weight * ;
unknown_fn: result += weight * frob(p);
bad_type: result += weight;
write_p: |-
  p = result;
unknown_param: result += param.nope * p;
bad_common: result += common.missing * p;
scoped: |-
  a := 1.0;
  if (a > 0.5) {
    b := a * 2.0;
    result += b * p;
  } else result -= a * p;
"""
    d = _db_from(tmp_path, src)
    assert d.variations == ["good", "scoped"]
    assert len(d.errors) == 6


def test_dsl_emits_fma_like_the_reference(db):
    f = fl.make_flame([fl.xform(1.0, 0, 0.5, [fl.vlink([1, 0, 0, 1, 0, 0], {"swirl": 1.0})])],
                      [[0, 0, 1]] * 256)
    src = db.flame(f).source()
    # result.x += weight * (sc.x * p.x - sc.y * p.y)  ->  nested fma with the
    # subtraction folded into the inner one (flame_compiler.cpp:114-140)
    assert "res.x = fl::fma(P[" in src and "fl::fma(u_sc.x, p.x, -(u_sc.y * p.y))" in src


def test_layout_small_by_hand(db):
    f = db.flame(fl.flame_small(1))
    # 13 header reals + 3 xforms x (4 + 6 affine + 1 weight) = 46
    assert f.size_reals() == 46
    assert f.affine_indices() == [0, 6, 17, 28, 39]


@pytest.mark.parametrize("maker,seed", [(fl.flame_small, 3), (fl.flame_1080, 1), (fl.flame_1080, 7),
                                        (fl.flame_4k, 2), (fl.flame_keyframe, 4)])
def test_layout_and_pack_match_oracle(db, maker, seed):
    fd = maker(seed)
    f = db.flame(fd)
    L = oracle.Layout(fd, oracle.load_variation_params(render.CONFIG_DIR))
    assert f.size_reals() == L.size_reals
    assert f.affine_indices() == L.affine_indices
    w, h = 1280, 720
    got = f.pack_samples(w, h, 0.0, 0.1, 4).reshape(4, -1)
    want = oracle.pack_sample(fd, L, w, h)
    for i in range(4):
        np.testing.assert_allclose(got[i], want, rtol=1e-13, atol=1e-13)


def test_fix_rotation_matches_oracle(db):
    fd = fl.flame_1080(3)
    f = db.flame(fd)
    n = f.size_reals()
    ss = n + 768
    rng = np.random.default_rng(5)
    s = np.concatenate([f.pack_samples(64, 64) for _ in range(1)])
    s = s.reshape(4, ss)
    # rotate affines and hues differently per sample so the unwrap paths run
    for i in range(4):
        for a in f.affine_indices():
            s[i, a:a + 4] = rng.normal(size=4)
        s[i, n::3] = rng.uniform(-400, 400, size=256)
    s = s.ravel().copy()
    idx = (C.c_size_t * len(f.affine_indices()))(*f.affine_indices())
    got = s.copy()
    _lib.check(_lib.lib().rfk_fix_rotation(got.ctypes.data_as(C.POINTER(C.c_double)), ss, n, idx,
                                          len(f.affine_indices()), 4))
    want = oracle.fix_rotation(s, ss, n, f.affine_indices(), 4)
    np.testing.assert_allclose(got, want, rtol=1e-15, atol=1e-15)
    # unwrapped angles never jump by more than pi between samples
    g = got.reshape(4, ss)
    for a in f.affine_indices():
        assert np.all(np.abs(np.diff(g[:, a:a + 2], axis=0)) <= np.pi + 1e-9)


def test_json_round_trip(db):
    fd = fl.flame_1080(2)
    f = db.flame(fd)
    back = f.to_json()
    f2 = db.flame(back)
    np.testing.assert_array_equal(f.pack_samples(320, 180), f2.pack_samples(320, 180))


def test_palette_blob_is_accepted(db):
    """librefrakt stores the palette as base64url(size_t length + zlib(doubles))"""
    fd = fl.flame_small(4)
    raw = np.array(fd["palette"], dtype=np.float64).tobytes()
    blob = struct.pack("<Q", len(raw)) + zlib.compress(raw)
    fd2 = dict(fd)
    fd2["palette"] = base64.urlsafe_b64encode(blob).decode().rstrip("=")
    a, b = db.flame(fd), db.flame(fd2)
    np.testing.assert_array_equal(a.pack_samples(64, 64), b.pack_samples(64, 64))


def test_bad_flames_are_rejected(db):
    fd = fl.flame_small(1)
    bad = [dict(fd, palette=fd["palette"][:10]), dict(fd, xforms=[]), "{not json"]
    x = json.loads(json.dumps(fd))
    x["xforms"][0]["vchain"][0]["variations"]["no_such_variation"] = 1.0
    bad.append(x)
    y = json.loads(json.dumps(fd))
    y["xforms"][0]["vchain"][0]["variations"]["curl"] = {"weight": 1.0, "parameters": {"c7": 1.0}}
    bad.append(y)
    for b in bad:
        with pytest.raises(_lib.RfkError):
            db.flame(b)


FLAM3 = """<flames><flame name="t" size="640 480" center="0.1 -0.2" scale="120" rotate="30" brightness="4"
 gamma="2.5" vibrancy="1" highlight_power="1" gamma_threshold="0.01">
  <xform weight="0.5" color="0" color_speed="0.5" spherical="0.7" julian="0.3" julian_power="3" coefs="1 0 0 1 0 0"/>
  <xform weight="0.25" color="1" pre_blur="0.2" linear="1" coefs="0.5 0.1 -0.1 0.5 0.2 0.3" post="1 0 0 1 0.5 0"/>
  <finalxform color="0" color_speed="0" linear="1" coefs="0.9 0 0 0.9 0 0"/>
  <palette count="256" format="RGB">""" + "FF0000" * 256 + """</palette>
</flame></flames>"""


def test_flam3_import(db):
    f = db.flame_from_flam3(FLAM3)
    js = f.to_json()
    assert len(js["xforms"]) == 2 and "final_xform" in js
    x0, x1 = js["xforms"]
    assert set(x0["vchain"][0]["variations"]) == {"spherical", "julian"}
    assert x0["vchain"][0]["variations"]["julian"]["parameters"]["power"] == 3.0
    # pre_blur -> gaussian_blur on an identity vlink; post -> a linear vlink
    assert len(x1["vchain"]) == 3 and "gaussian_blur" in x1["vchain"][0]["variations"]
    assert x1["vchain"][2]["transform"] == [1, 0, 0, 1, 0.5, 0]
    assert abs(js["scale"] - 120 / 480) < 1e-12
    h, s, v = js["palette"][0]
    assert (h, s, v) == (0.0, 1.0, 1.0)


def test_reference_flam3_fixture_imports(db):
    f = db.flame_from_flam3(REFERENCE_FLAM3.read_text())
    js = f.to_json()
    assert len(js["xforms"]) == 4 and len(js["palette"]) == 256
    assert [len(x["vchain"]) for x in js["xforms"]] == [1, 2, 2, 1]
    assert f.compile_check() > 0


@pytest.mark.parametrize("maker", [fl.flame_small, fl.flame_1080, fl.flame_4k])
def test_generated_kernels_compile_offline(db, maker):
    f = db.flame(maker(1))
    assert f.compile_check() > 0
    assert f.compile_check(threads_per_block=1024, atomic_width=64) > 0
    assert f.compile_check(precision=1) > 0
    assert f.compile_check(threads_per_block=512, atomic_width=64, bin_queue=True) > 0


def test_every_variation_compiles(db):
    """one flame that uses every variation of the shipped database"""
    names = db.variations
    xfs = []
    for i in range(0, len(names), 4):
        group = names[i:i + 4]
        vars_ = {n: 0.25 for n in group}
        xfs.append(fl.xform(1.0, 0.5, 0.5, [fl.vlink([1, 0, 0, 1, 0, 0], vars_)]))
    fd = fl.make_flame(xfs, [[0, 0, 1]] * 256)
    # parameterised variations need the object form
    vp = oracle.load_variation_params(render.CONFIG_DIR)
    for x in fd["xforms"]:
        vs = x["vchain"][0]["variations"]
        for n in list(vs):
            if vp[n][0]:
                vs[n] = {"weight": 0.25, "parameters": {p: 0.5 for p in vp[n][0]}}
    f = db.flame(fd)
    assert f.compile_check() > 0


def test_chaos_flame_compiles(db):
    fd = fl.flame_small(2)
    fd["chaos"] = [[1.0, 0.0, 1.0], [0.5, 1.0, 0.0], [1.0, 1.0, 1.0]]
    f = db.flame(fd)
    assert f.size_reals() == 46 + 3 * 4
    assert "rfk_select_xform_chaos" in f.source()
    assert f.compile_check() > 0


@pytest.mark.skipif(not REFERENCE_CONFIG.exists(), reason="reference tree not mounted")
def test_reference_variations_compile():
    rdb = render.VariationDB(REFERENCE_CONFIG)
    names = rdb.variations
    y = __import__("yaml").safe_load((REFERENCE_CONFIG / "variations_new.yml").read_text())
    xfs = []
    for i in range(0, len(names), 5):
        vars_ = {}
        for n in names[i:i + 5]:
            params = [p for p, d in ((y[n].get("param") or {}).items() if isinstance(y[n], dict) else [])
                      if "precalc" not in ((d or {}).get("tags") or [])]
            vars_[n] = {"weight": 0.2, "parameters": {p: 0.5 for p in params}} if params else 0.2
        xfs.append(fl.xform(1.0, 0.5, 0.5, [fl.vlink([1, 0, 0, 1, 0, 0], vars_)]))
    f = rdb.flame(fl.make_flame(xfs, [[0, 0, 1]] * 256))
    assert f.compile_check() > 0


def test_align_keyframes(db):
    a, b = fl.flame_keyframe(1), fl.flame_keyframe(2)
    fa, fb = db.flame(a), db.flame(b)
    arr = (C.c_void_p * 2)(fa.handle, fb.handle)
    out = (C.c_void_p * 2)()
    _lib.check(_lib.lib().rfk_flames_align(db.handle, arr, 2, out))
    oa, ob = render.Flame(db, None, handle=out[0]), render.Flame(db, None, handle=out[1])  # take ownership
    ja, jb = oa.to_json(), ob.to_json()
    n = len(a["xforms"]) + len(b["xforms"])
    assert len(ja["xforms"]) == n == len(jb["xforms"])
    wa = [x["weight"] for x in ja["xforms"]]
    assert all(w == 0 for w in wa[len(a["xforms"]):]) and any(w > 0 for w in wa[:len(a["xforms"])])
    # identical structure -> one compiled kernel serves every keyframe
    assert oa.size_reals() == ob.size_reals()
    assert oa.source() == ob.source()
