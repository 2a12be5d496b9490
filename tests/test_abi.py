"""The C ABI library: it loads, exports every entry point of include/tbg.h,
its host-compiled primitives pass their self-test, and its error paths are
reachable from the host without a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

import tb8gen_amd as tbg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "tbg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tbg_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = tbg.lib()
    names = declared_functions()
    assert len(names) >= 19
    for n in names:
        assert hasattr(L, n), n
    assert sorted(tbg.ABI_SYMBOLS) == names


def test_host_selftest_of_device_primitives():
    r, msg = tbg.host_selftest()
    assert r == 0, msg


def test_king_slice_numbering_equals_the_reference_table():
    """tbg_king_slice / tbg_slice_kings against every entry of the
    reference's KKIdx and InvTriangle tables (src/probe.c:327-329, 441-522,
    committed as data), and KKMap's symmetry classes over all king pairs."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "kkidx_probe_c.json")) as f:
        ref = json.load(f)
    seen = set()
    for tri, wk in enumerate(ref["inv_triangle"]):
        for bk, s in enumerate(ref["kkidx"][tri]):
            if s < 0:
                continue
            assert tbg.king_slice(wk, bk) == s, (wk, bk)
            assert tbg.slice_kings(s) == (wk, bk), s
            seen.add(s)
    assert seen == set(range(462))
    f = lambda x: x ^ 7          # file mirror
    r = lambda x: x ^ 56         # rank mirror
    d = lambda x: ((x & 7) << 3) | (x >> 3)  # a1-h8 flip
    for wk in range(64):
        for bk in range(64):
            s = tbg.king_slice(wk, bk)
            touching = max(abs((wk & 7) - (bk & 7)), abs((wk >> 3) - (bk >> 3))) <= 1
            assert (s == -1) == touching
            if s >= 0:
                for g in (f, r, d):
                    assert tbg.king_slice(g(wk), g(bk)) == s
    assert tbg.king_slice(-1, 5) == -1 and tbg.king_slice(3, 64) == -1
    with pytest.raises(tbg.TbgError):
        tbg.slice_kings(462)


def test_error_codes_have_messages():
    L = tbg.lib()
    for code in (0, -1, -2, -3, -4, -5, -6, -7, -8):
        assert L.tbg_strerror(code)


def test_geometry_is_validated_on_the_host():
    """Launch geometries are checked before anything reaches a device; the
    process default round-trips."""
    L = tbg.lib()
    g0 = tbg.default_geometry()
    assert g0.verify_unit_words % 256 == 0 and g0.sweep_unit_words % 512 == 0
    for field, bad in [("verify_unit_words", 100), ("verify_unit_words", 0),
                       ("sweep_unit_words", 256), ("sweep_unit_words", 1 << 20),
                       ("pred_wg_per_cu", 0), ("verify_wg_per_cu", 65), ("scatter", 7)]:
        g = tbg.Geometry(**g0.as_dict())
        setattr(g, field, bad)
        assert L.tbg_set_geometry(None, ctypes.byref(g)) == -1, field
    g = tbg.Geometry(**g0.as_dict())
    g.verify_unit_words, g.scatter = 1024, tbg.SCATTER_GLOBAL
    tbg.set_default_geometry(g)
    assert tbg.default_geometry().as_dict() == g.as_dict()
    tbg.set_default_geometry(g0)
    assert tbg.default_geometry().as_dict() == g0.as_dict()
    assert L.tbg_autotune(None, 0, None, None, None) == -1


def test_registry_arguments_are_checked():
    L = tbg.lib()
    r = ctypes.c_void_p()
    assert L.tbg_registry_create(None, 2) == -1
    assert L.tbg_registry_create(ctypes.byref(r), 9) == -1
    assert L.tbg_registry_count(None) == 0
    L.tbg_registry_destroy(None)
    assert L.tbg_registry_get(None, b"KRvK", None) == -1


def test_calls_before_init_are_rejected():
    L = tbg.lib()
    if tbg._inited:
        pytest.skip("library already initialised in this process")
    p = ctypes.c_void_p()
    assert L.tbg_create(ctypes.byref(p), b"KRvK") == -5  # TBG_E_STATE
    assert L.tbg_get_info(None, None) == -1


def test_init_without_device_fails_loudly():
    r = subprocess.run(["rocminfo"], capture_output=True, text=True) if os.path.exists(
        "/opt/rocm/bin/rocminfo") else None
    has_gpu = r is not None and "gfx9" in (r.stdout or "")
    if has_gpu:
        pytest.skip("a GPU is present")
    assert tbg.lib().tbg_init(0) == -7  # TBG_E_NODEV
    with pytest.raises(tbg.TbgError):
        tbg.Table("KRvK")


def test_driver_binary_usage():
    exe = tbg.DRIVER_PATH
    assert os.path.exists(exe)
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode != 0 and "No tablebase specified" in r.stdout


def test_material_helpers():
    assert tbg.canonical("KvKR") == "KRvK"
    assert tbg.canonical("KRvKQ") == "KQvKR"
    assert tbg.canonical("KNvKBN") == "KBNvKN"
    assert tbg.sub_materials("KQRvKR") == ["KRvKR", "KQvKR", "KQRvK"]
    assert tbg.sub_materials("KRvK") == []
