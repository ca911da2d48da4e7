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
    """export.write_iteration on a stand-in planner object (host arrays only):
    capacity rows of samples and parents, one row per iteration of counters,
    flags and scores, the path when found."""
    from kgmt_amd import export
    from kgmt_amd.abi import KgmtResult

    class Stub:
        h_itr_ = 3
        h_treeSize_ = 4
        h_pathToGoal_ = 3
        result = KgmtResult()

        class cfg:
            capacity = 6

        def tree(self):
            s = np.arange(40, dtype=np.float32).reshape(4, 10) / 7
            return s, np.array([-1, 0, 0, 2], np.int32), np.zeros(4, np.float32)

        def regions(self):
            return np.arange(8, dtype=np.int32), np.arange(8, dtype=np.int32) // 2, np.ones(8, np.float32)

        def frontier(self):
            return np.array([0, 1, 1, 0], np.uint8), np.array([0, 15, 1, 0], np.uint32)

    st = Stub()
    st.h_controlPathToGoal_ = st.tree()[0][[3, 2, 0]]
    export.write_iteration(st, str(tmp_path), 0, frontier_size=2, expanded=64)
    s = export.read_rows(tmp_path / "Samples" / "Samples0" / "samples3.csv")
    assert s.shape == (6, 10) and np.allclose(s[:4], st.tree()[0], atol=5e-7) and np.all(s[4:] == 0)
    p = (tmp_path / "Parents" / "Parents0" / "parents3.csv").read_text().split()
    assert p == ["-1", "0", "0", "2", "-1", "-1"]
    fr = (tmp_path / "Frontier" / "Frontier0" / "frontier.csv").read_text()
    assert fr == "0,1,1,0,0,0\n"
    rep = (tmp_path / "FrontierRepeatCount" / "FrontierRepeatCount0" / "frontierRepeatCount.csv").read_text()
    assert rep == "0,15,1,0,0,0\n"
    assert (tmp_path / "ExpandedNodes" / "ExpandedNodes0" / "expandedNodes.csv").read_text() == "64\n"
    path = export.read_rows(tmp_path / "ControlPathToGoal" / "ControlPathToGoal0" / "controlPathToGoal.csv")
    assert path.shape == (3, 10)


WRITER_SRC = r"""
#include "kgmt_planner.hpp"
int main(int argc, char** argv)
{
    std::vector<float> f = {0.0f, -0.0f, 1.0f / 3.0f, -2.5e-7f, 3.0e-7f, 123456.789f, -1.0e-3f, 0.5e-6f, 1.5e-6f,
                            2.5e-6f, 1e30f, -7.0000005f};
    std::vector<int> i = {-1, 0, 7, 2147483647};
    std::vector<unsigned int> u = {0u, 15u, 1u, 4000000000u};
    kgmt::writeVectorToCSV(f, std::string(argv[1]) + "/f.csv", 3, 4, false);
    kgmt::writeVectorToCSV(i, std::string(argv[1]) + "/i.csv", 4, 1, false);
    kgmt::writeVectorToCSV(u, std::string(argv[1]) + "/u.csv", 1, 4, true);
    kgmt::writeValueToCSV(12345u, std::string(argv[1]) + "/v.csv");
    kgmt::writeValueToCSV(0.25, std::string(argv[1]) + "/v.csv");
    return 0;
}
"""


def test_csv_writers_match_the_iostream_format(tmp_path):
    """The Python writers of kgmt_amd.export produce the bytes of the
    reference's iostream writers (helper.cuh:56-109: std::fixed with
    setprecision 6 for arrays, 10 for single values; integers plain), here
    the same writers in include/kgmt_planner.hpp compiled with g++."""
    import subprocess
    from kgmt_amd import export
    src = tmp_path / "w.cpp"
    src.write_text(WRITER_SRC)
    exe = tmp_path / "w"
    subprocess.run(["g++", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)], check=True)
    cdir = tmp_path / "c"
    cdir.mkdir()
    subprocess.run([str(exe), str(cdir)], check=True)
    f = np.array([0.0, -0.0, 1.0 / 3.0, -2.5e-7, 3.0e-7, 123456.789, -1.0e-3, 0.5e-6, 1.5e-6, 2.5e-6, 1e30,
                  -7.0000005], np.float32)
    assert (cdir / "f.csv").read_text() == export._rows_text(f.reshape(3, 4), "%.6f")
    assert (cdir / "i.csv").read_text() == export._rows_text(np.array([-1, 0, 7, 2147483647])[:, None], "%d")
    assert (cdir / "u.csv").read_text() == export._rows_text(np.array([[0, 15, 1, 4000000000]], np.uint32), "%d")
    pdir = tmp_path / "p"
    pdir.mkdir()
    export._value(str(pdir / "v.csv"), np.uint32(12345))
    export._value(str(pdir / "v.csv"), 0.25)
    assert (cdir / "v.csv").read_text() == (pdir / "v.csv").read_text()
