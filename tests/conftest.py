"""Shared pytest configuration: the ``gpu`` marker and common fixtures."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def grid_coordinates(region, shape=None, spacing=None, extra=None):
    """Regular grid with inclusive endpoints, meshgrid shape (n_north, n_east),
    as bordado.grid_coordinates returns (used throughout the reference tests)."""
    w, e, s, n = region
    if shape is None:
        n_east = int(round((e - w) / spacing)) + 1
        n_north = int(round((n - s) / spacing)) + 1
    else:
        n_north, n_east = shape
    east = np.linspace(w, e, n_east)
    north = np.linspace(s, n, n_north)
    E, N = np.meshgrid(east, north)
    if extra is None:
        return E, N
    return E, N, np.full_like(E, extra, dtype=float)


@pytest.fixture
def gpu_lib():
    """The HIP library, skipping when no GPU is visible."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from harmonica_amd import _lib

    return _lib.lib()
