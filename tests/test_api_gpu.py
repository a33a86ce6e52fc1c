"""The oneapi::dal::kmeans C++ drop-in surface on the device: the reference's
gold dataset, empty-cluster fixture and smoke-data inference through
dal::train / dal::infer (tests/cpp/test_api_gpu.cpp)."""

import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.gpu


def build_exe():
    src = os.path.join(ROOT, "tests", "cpp", "test_api_gpu.cpp")
    exe = os.path.join(ROOT, "tests", "cpp", "build", "test_api_gpu")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    lib_dir = os.path.join(ROOT, "onedal_amd", "lib")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-L", lib_dir,
                    "-lkmeans_hip", "-Wl,-rpath," + lib_dir, "-o", exe], check=True)
    return exe


def test_cpp_api_on_device(tmp_path):
    with open(os.path.join(ROOT, "tests", "golden", "gold_dataset.json")) as f:
        g = json.load(f)
    txt = tmp_path / "gold.txt"
    vals = [g["row_count"], g["column_count"], g["cluster_count"]] + g["data"] + \
        g["initial_centroids"] + g["expected_centroids"] + g["expected_responses"] + \
        [g["expected_objective"]]
    txt.write_text(" ".join(repr(v) for v in vals))
    exe = build_exe()
    res = subprocess.run([exe, str(txt)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
