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
