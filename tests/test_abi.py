"""The C ABI without a GPU: every declared entry point is exported and bound,
and the host-only entry points agree bit for bit with the oracle.

The host-only functions restate BaseModel::model2UnitCube
(BaseModel.cpp:84-92), BaseModel::setUniformBoundingBox (BaseModel.cpp:24-51)
and the shell classification of ThinShells::singlePointQuery
(ThinShells.cpp:1412-1417).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from its_amd import _capi, meshes

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "its_c.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(its_\w+)\s*\(", text, flags=re.M)))


def test_header_parses():
    names = declared_functions()
    assert len(names) >= 30
    assert "its_bspline_eval_device" in names and "its_lscg_solve" in names


def test_library_exports_every_declared_symbol():
    L = _capi.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    bound = set(_capi.SIGNATURES) | set(_capi._RET)
    assert set(declared_functions()) == bound


def test_version_and_error_text():
    L = _capi.lib()
    assert b"gfx950" in L.its_version()
    with pytest.raises(_capi.ItsError) as e:
        _capi.tune_set(_capi.TUNE_EVAL_ILP, 3)
    assert e.value.code == -5 and "ILP" in str(e.value)
    with pytest.raises(_capi.ItsError):
        _capi.tune_set(99, 1)


def test_tunables_roundtrip():
    for key in (_capi.TUNE_DENSE_CELLS, _capi.TUNE_EVAL_BLOCKS_PER_CU, _capi.TUNE_COARSE_VERTS, _capi.TUNE_EVAL_ILP):
        old = _capi.tune_get(key)
        _capi.tune_set(key, old)
        assert _capi.tune_get(key) == old
    with pytest.raises(_capi.ItsError):
        _capi.tune_set(_capi.TUNE_COARSE_VERTS, 12289)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_unit_cube_and_bbox_bit_equal(seed):
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(500, 3)) * rng.uniform(0.1, 10, size=3) + rng.uniform(-5, 5, size=3)
    a = _capi.model_to_unit_cube(V)
    b = O.normalize(V)
    assert np.array_equal(a, b)
    for got, want in zip(_capi.uniform_bbox(a), O.uniform_bbox(b)):
        assert np.array_equal(np.asarray(got), np.asarray(want))


@pytest.mark.parametrize("k", [0, 1, 2])
def test_bbox_known_answers(k):
    """The library's setUniformBoundingBox on the hand-derived boxes
    (BaseModel.cpp:24-51: binary32 longest side and padding, binary64 1/11)."""
    from helpers import bbox_known_answers
    V, o, e, w = bbox_known_answers()[k]
    go, ge, gw = _capi.uniform_bbox(V)
    for got, want in ((go, o), (ge, e), (gw, w)):
        np.testing.assert_array_equal(np.asarray(got).view(np.uint64), want.view(np.uint64))


def test_unit_cube_known_answer():
    # a box [2,4]x[1,2]x[0,1]: scale 2, centre (3,1.5,0.5) maps to (0.5,0.5,0.5)
    # and the longest side to [0,1]
    V = np.array([[2, 1, 0], [4, 2, 1], [3, 1.5, 0.5]], np.float64)
    W = _capi.model_to_unit_cube(V)
    assert np.array_equal(W, np.array([[0, 0.25, 0.25], [1, 0.75, 0.75], [0.5, 0.5, 0.5]]))


def test_classify_thresholds():
    # thresholds exact in binary32, so the widened f compares exactly
    f = np.array([-1.0, -0.25, 0.0, 0.125, 0.25, 0.5, np.nan], np.float32)
    c = _capi.classify(f, -0.25, 0.25)
    # >= outer -> 1, <= inner -> -1, otherwise 0 (NaN compares false)
    assert c.tolist() == [-1, -1, 0, 0, 1, 1, 0]


def test_product_fails_loudly_without_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_capi, "_lib", None)
    monkeypatch.setattr(_capi, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(_capi.LibraryMissing):
        _capi.lib()


def test_triangles_layout():
    V, F = meshes.noisy_icosphere(seed=1)
    T = meshes.triangles(V, F)
    assert T.shape == (len(F), 9)
    assert np.array_equal(T[:, 3:6], V[F[:, 1]])
