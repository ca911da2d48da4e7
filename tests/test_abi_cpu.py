"""The C-ABI library on a machine without a GPU: it loads, exports every
function include/kgmt.h declares, fills and validates configurations, and
refuses to create a context without a device. No compute calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from kgmt_amd import _lib
from kgmt_amd import workloads as W
from kgmt_amd.abi import KGMT_E_ARG, KGMT_E_UNSUPPORTED, KgmtConfig, KgmtResult
from kgmt_amd.abi import KGMT_OK

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "kgmt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kgmt_[a-z_0-9]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    # and the Python bindings cover the same set
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_struct_layouts():
    assert C.sizeof(KgmtConfig) == 168 and C.sizeof(KgmtResult) == 40
    assert KgmtConfig.query_id.offset == 148 and KgmtConfig.uni_length.offset == 152
    assert KgmtConfig.seed.offset == 160


def test_presets_match_python_workloads():
    L = _lib.lib()
    for name, p in W.PRESET_OF.items():
        c = KgmtConfig()
        _lib.check(L.kgmt_config_preset(p, C.byref(c)))
        assert c.as_dict() == W.config(name).as_dict(), name
        assert L.kgmt_config_check(C.byref(c)) == KGMT_OK


def test_config_validation():
    L = _lib.lib()
    good = W.config("di6d")
    bad = [
        dict(model=9),
        dict(model=-1),
        dict(reserved=1),
        dict(capacity=1),
        dict(dt=0.0),
        dict(steps_min=5, steps_max=4),
        dict(ws_max=0.0),
        dict(r1_len=0),
        dict(r1_len=21),
        dict(repeat_sparse=64),
        dict(repeat_sparse=0),
        dict(repeat_dense=0),
        dict(children_per_entry=0),
        dict(max_iter=0),
    ]
    for kw in bad:
        rc = L.kgmt_config_check(C.byref(good.replace(**kw)))
        assert rc in (KGMT_E_ARG, KGMT_E_UNSUPPORTED), kw
        assert L.kgmt_last_error()
    assert L.kgmt_config_check(None) == KGMT_E_ARG
    uni = W.config("unicycle2d")
    assert L.kgmt_config_check(C.byref(uni)) == KGMT_OK
    # planar grids: r1_len^2 regions, so 21 is fine there; the wheel base must be positive
    assert L.kgmt_config_check(C.byref(uni.replace(r1_len=21))) == KGMT_OK
    for kw in (dict(uni_length=0.0), dict(uni_length=float("nan")), dict(r1_len=91)):
        assert L.kgmt_config_check(C.byref(uni.replace(**kw))) == KGMT_E_ARG, kw


def test_model_dims():
    L = _lib.lib()
    s, c, d = C.c_int32(), C.c_int32(), C.c_int32()
    for m, want in [(0, (4, 2, 7)), (1, (6, 3, 10)), (2, (6, 3, 10)), (3, (12, 4, 17)), (4, (5, 3, 9))]:
        assert L.kgmt_model_dims(m, C.byref(s), C.byref(c), C.byref(d)) == KGMT_OK
        assert (s.value, c.value, d.value) == want
    assert L.kgmt_model_dims(5, C.byref(s), C.byref(c), C.byref(d)) != KGMT_OK


def test_null_arguments_are_rejected():
    L = _lib.lib()
    assert L.kgmt_create(None, None) == KGMT_E_ARG
    assert L.kgmt_plan(None, None, None, None, 0, 0, None) == KGMT_E_ARG
    assert L.kgmt_get_tree(None, None, None, None, 0, None) == KGMT_E_ARG
    assert L.kgmt_batch_create(None, 0, None) == KGMT_E_ARG
    assert L.kgmt_destroy(None) == KGMT_OK


def test_no_device_is_reported():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    L = _lib.lib()
    h = C.c_void_p()
    rc = L.kgmt_create(C.byref(W.config("di6d")), C.byref(h))
    assert rc == -6 and not h.value
    assert b"device" in L.kgmt_last_error()


def test_obstacle_csv_fixtures():
    import json
    gd = os.path.join(ROOT, "tests", "golden", "obstacles")
    meta = json.load(open(os.path.join(gd, "obstacles.json")))["boxes"]
    for name, n in meta.items():
        boxes = W.read_obstacles_csv(os.path.join(gd, f"{name}.csv"))
        assert boxes.shape == (n, 6), name
    # the reference's house set holds eight boxes with xmax < xmin (rows 38-41,
    # 46, 47, 50, 51); rows are passed through as read, like readObstaclesFromCSV
    house = W.read_obstacles_csv(os.path.join(gd, "house.csv"))
    inverted = np.flatnonzero(~np.all(house[:, 3:] >= house[:, :3], axis=1))
    assert list(inverted) == [38, 39, 40, 41, 46, 47, 50, 51]
    # narrowPassage: two boxes, the last line without a newline
    np_ = W.read_obstacles_csv(os.path.join(gd, "narrowPassage.csv"))
    assert np.allclose(np_, [[0.3, 0, 0, 0.5, 0.49, 1.0], [0.3, 0.51, 0, 0.5, 1.0, 1.0]])


def test_csv_export_roundtrip(tmp_path):
    """export.write_plan_csv on a stand-in planner object (host arrays only)."""
    from kgmt_amd import export
    from kgmt_amd.abi import KgmtResult

    class Stub:
        h_itr_ = 3
        h_treeSize_ = 4
        h_pathToGoal_ = 3
        result = KgmtResult()

        def tree(self):
            s = np.arange(40, dtype=np.float32).reshape(4, 10) / 7
            return s, np.array([-1, 0, 0, 2], np.int32), np.zeros(4, np.float32)

        def regions(self):
            return np.arange(8, dtype=np.int32), np.arange(8, dtype=np.int32) // 2, np.ones(8, np.float32)

    st = Stub()
    st.h_controlPathToGoal_ = st.tree()[0][[3, 2, 0]]
    export.write_plan_csv(st, str(tmp_path), 0)
    s = export.read_rows(tmp_path / "Samples" / "Samples0" / "samples3.csv")
    assert np.allclose(s, st.tree()[0], atol=5e-7)
    p = export.read_rows(tmp_path / "Parents" / "Parents0" / "parents3.csv")
    assert list(p[:, 0]) == [-1, 0, 0, 2]
    path = export.read_rows(tmp_path / "ControlPathToGoal" / "ControlPathToGoal0" / "controlPathToGoal.csv")
    assert path.shape == (3, 10)
