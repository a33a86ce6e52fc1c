"""Argument checks of the CSR binding (no device needed): array lengths that the
C ABI would trust are verified before any call."""
import numpy as np
import pytest

from onedal_amd.engine import _check_csr


def test_host_arrays_pass():
    assert _check_csr(np.ones(3, np.float32), np.zeros(3, np.int64), np.array([0, 1, 3])) is False


def test_length_mismatch_is_refused():
    with pytest.raises(ValueError, match="same length"):
        _check_csr(np.ones(3, np.float32), np.zeros(2, np.int64), np.array([0, 3]))
    with pytest.raises(ValueError, match="two entries"):
        _check_csr(np.ones(0, np.float32), np.zeros(0, np.int64), np.array([0]))
    with pytest.raises(ValueError, match="one-dimensional"):
        _check_csr(np.ones((3, 1), np.float32), np.zeros(3, np.int64), np.array([0, 3]))


class _FakeCudaTensor:
    """Duck-typed stand-in for a CUDA tensor: the checks read only metadata."""

    is_cuda = True

    def __init__(self, shape, dtype="torch.float32", contiguous=True, device="cuda:0"):
        self.shape = shape
        self.dtype = dtype
        self._contig = contiguous
        self.device = device

    def data_ptr(self):
        return 0

    def is_contiguous(self):
        return self._contig


def test_centroid_shape_is_checked_on_host():
    from onedal_amd.engine import _check_centroids

    _check_centroids(np.zeros((4, 3), np.float32), 3, np.float32, False)
    _check_centroids([[1.0, 2.0, 3.0]], 3, np.float64, False)
    with pytest.raises(ValueError, match=r"shape \(k, 3\)"):
        _check_centroids(np.zeros((4, 2), np.float32), 3, np.float32, False)
    with pytest.raises(ValueError, match=r"shape \(k, 3\)"):
        _check_centroids(np.zeros(12, np.float32), 3, np.float32, False)
    with pytest.raises(TypeError, match="CUDA tensor exactly"):
        _check_centroids(_FakeCudaTensor((4, 3)), 3, np.float32, False)


def test_device_centroids_must_match_type_layout_and_device():
    from onedal_amd.engine import _check_centroids

    _check_centroids(_FakeCudaTensor((4, 3)), 3, np.float32, True, "cuda:0")
    with pytest.raises(TypeError, match="CUDA tensor exactly"):
        _check_centroids(np.zeros((4, 3), np.float32), 3, np.float32, True)
    with pytest.raises(TypeError, match="data's type float64"):
        _check_centroids(_FakeCudaTensor((4, 3)), 3, np.float64, True)
    with pytest.raises(ValueError, match="contiguous"):
        _check_centroids(_FakeCudaTensor((4, 3), contiguous=False), 3, np.float32, True)
    with pytest.raises(ValueError, match="device"):
        _check_centroids(_FakeCudaTensor((4, 3), device="cuda:1"), 3, np.float32, True, "cuda:0")
    with pytest.raises(ValueError, match=r"shape \(k, 5\)"):
        _check_centroids(_FakeCudaTensor((4, 3)), 5, np.float32, True)


def test_device_rows_must_be_2d_contiguous_float():
    from onedal_amd.engine import _check_device_rows

    _check_device_rows(_FakeCudaTensor((10, 3)))
    with pytest.raises(ValueError, match="two-dimensional"):
        _check_device_rows(_FakeCudaTensor((10,)))
    with pytest.raises(ValueError, match="contiguous"):
        _check_device_rows(_FakeCudaTensor((10, 3), contiguous=False))
    with pytest.raises(TypeError, match="float32 and float64"):
        _check_device_rows(_FakeCudaTensor((10, 3), dtype="torch.float16"))
