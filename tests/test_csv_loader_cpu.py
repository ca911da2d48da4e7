"""Obstacle CSV reading, C++ and Python, against the reference's parsing
rules (src/helper/helper.cu:13-39: `ss >> value` per line, one comma after
a number skipped, a line ends at the first token that is not a number,
count = values / (2 * W_DIM)): include/kgmt_planner.hpp's
readObstaclesFromCSV is compiled with g++ (it needs no GPU and no library)
and run on the reference's obstacle files and on hand-made edge cases; the
Python reader (workloads.read_obstacles_csv) must return the same values."""
import glob
import os
import subprocess

import numpy as np
import pytest

from kgmt_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "csv_main.cpp")
FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "obstacles", "*.csv")))

EDGE_CASES = {
    "no_trailing_newline": "0.1,0.2,0.3,0.4,0.5,0.6\n0.7,0.8,0.9,1.0,1.1,1.2",
    "spaces_and_commas": "  1, 2 ,3,4,5,6\n7 8 9 10 11 12\n",
    "blank_lines": "\n1,2,3,4,5,6\n\n\n",
    "header_line_skipped": "xmin,ymin,zmin,xmax,ymax,zmax\n1,2,3,4,5,6\n",
    "stops_at_text": "1,2,3,abc,4,5\n6,7,8,9,10,11,12,13,14\n",
    "double_comma": "1,,2,3\n4,5,6,7,8,9,10,11\n",
    "exponents_signs": "-1e-3,+2.5E2,.5,5.,-0,1e1\n",
    "partial_box": "1,2,3,4,5,6,7,8\n",
    "trailing_comma": "1,2,3,4,5,6,\n",
    "crlf": "1,2,3,4,5,6\r\n7,8,9,10,11,12\r\n",
    "empty": "",
}


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = tmp_path_factory.mktemp("csv") / "csv_main"
    subprocess.run(["g++", "-O1", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", SRC, "-o", str(out)],
                   check=True)
    return str(out)


def cpp_read(exe, path, wdim):
    r = subprocess.run([exe, path, str(wdim)], capture_output=True, text=True, timeout=30)
    lines = r.stdout.split()
    if lines[0] == "error":
        return None, None
    n = int(lines[0])
    vals = np.array([float(x) for x in lines[2:]], np.float32)
    assert len(vals) == int(lines[1])
    return n, vals


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_reference_obstacle_files(path, exe):
    n, vals = cpp_read(exe, path, 3)
    boxes = W.read_obstacles_csv(path)
    assert n == len(boxes)
    assert np.array_equal(vals[: 6 * n].reshape(n, 6), boxes)


@pytest.mark.parametrize("name", sorted(EDGE_CASES))
def test_edge_cases(name, exe, tmp_path):
    p = tmp_path / f"{name}.csv"
    p.write_bytes(EDGE_CASES[name].encode())
    for wdim in (2, 3):
        n, vals = cpp_read(exe, str(p), wdim)
        boxes = W.read_obstacles_csv(str(p), wdim)
        k = 2 * wdim
        assert n == len(vals) // k == len(boxes), (name, wdim)
        assert np.array_equal(vals[: n * k].reshape(n, k), boxes), (name, wdim)


def test_known_counts(exe, tmp_path):
    p = tmp_path / "x.csv"
    p.write_text(EDGE_CASES["stops_at_text"])
    n, vals = cpp_read(exe, str(p), 3)
    assert vals.tolist() == [1, 2, 3] + list(range(6, 15)) and n == 2
    p.write_text(EDGE_CASES["double_comma"])
    assert cpp_read(exe, str(p), 3)[1].tolist() == [1] + list(range(4, 12))


def test_missing_file_is_an_error(exe, tmp_path):
    assert cpp_read(exe, str(tmp_path / "none.csv"), 3) == (None, None)
    with pytest.raises(FileNotFoundError):
        W.read_obstacles_csv(str(tmp_path / "none.csv"))
