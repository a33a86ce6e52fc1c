"""Extract the reference's k-means golden vectors into small data fixtures.

Reads, as text, the reference's own test data and example data:

* ``cpp/oneapi/dal/algo/kmeans/test/data.hpp:26-184`` -- the "gold dataset"
  (100 x 3 points, k = 3, given initial centroids, expected centroids, expected
  responses and expected objective 241.3593613), used by
  ``test/fixture.hpp:156-203`` with max_iteration_count = 100, threshold 0.
* ``data/kmeans_dense_train_data.csv`` / ``kmeans_dense_train_centroids.csv`` --
  the example inputs of ``examples/oneapi/cpp/source/kmeans/kmeans_lloyd_dense_batch.cpp``.

Writes ``gold_dataset.json`` and ``kmeans_dense_example.npz`` next to this
script.  Only numbers are copied; no reference source text is kept.

Usage: python tests/golden/make_gold_fixtures.py [/root/reference]
"""

import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _arrays_in_function(text, func_name):
    """Return the float literals of the `data = {...}` initialiser inside func."""
    start = text.index(func_name)
    brace = text.index("data = {", start)
    end = text.index("};", brace)
    body = text[brace + len("data = {"):end]
    body = re.sub(r"//[^\n]*", "", body)
    return [float(tok) for tok in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", body)]


def extract_gold(ref_root):
    path = os.path.join(ref_root, "cpp/oneapi/dal/algo/kmeans/test/data.hpp")
    with open(path) as f:
        text = f.read()
    rows = int(re.search(r"row_count = (\d+);", text).group(1))
    cols = int(re.search(r"column_count = (\d+);", text).group(1))
    k = int(re.search(r"cluster_count = (\d+);", text).group(1))
    obj = float(re.search(r"get_expected_objective\(\) \{\s*return ([-+0-9.eE]+);", text).group(1))
    data = _arrays_in_function(text, "get_data()")
    init = _arrays_in_function(text, "get_initial_centroids()")
    cent = _arrays_in_function(text, "get_expected_centroids()")
    resp = _arrays_in_function(text, "get_expected_responses()")
    assert len(data) == rows * cols, (len(data), rows, cols)
    assert len(init) == k * cols and len(cent) == k * cols and len(resp) == rows
    return {
        "source": "cpp/oneapi/dal/algo/kmeans/test/data.hpp:26-184",
        "row_count": rows,
        "column_count": cols,
        "cluster_count": k,
        "max_iteration_count": 100,
        "accuracy_threshold": 0.0,
        "data": data,
        "initial_centroids": init,
        "expected_centroids": cent,
        "expected_responses": [int(r) for r in resp],
        "expected_objective": obj,
    }


def extract_example(ref_root):
    base = os.path.join(ref_root, "data")
    x = np.loadtxt(os.path.join(base, "kmeans_dense_train_data.csv"), delimiter=",", dtype=np.float32)
    c = np.loadtxt(os.path.join(base, "kmeans_dense_train_centroids.csv"), delimiter=",", dtype=np.float32)
    return x, c


def main():
    ref_root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    gold = extract_gold(ref_root)
    with open(os.path.join(HERE, "gold_dataset.json"), "w") as f:
        json.dump(gold, f, indent=1)
    x, c = extract_example(ref_root)
    np.savez_compressed(os.path.join(HERE, "kmeans_dense_example.npz"), data=x, initial_centroids=c)
    print("wrote gold_dataset.json (%dx%d, k=%d) and kmeans_dense_example.npz %s %s"
          % (gold["row_count"], gold["column_count"], gold["cluster_count"], x.shape, c.shape))


if __name__ == "__main__":
    main()
