"""The gfx950 library loads on a CPU-only host and exports every entry point include/mpmsnow_gpu.h declares.

No compute calls are made here (there may be no GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mpmsnow_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ["mpm_create", "mpm_sim_advance", "mpm_grid_update_grid_velocities", "mpm_grid_force_differentials",
              "mpm_model_update_particle_data", "mpm_particles_set"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    import mpmsnow_amd

    L = ctypes.CDLL(mpmsnow_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared_symbols()) == set(mpmsnow_amd.SIGNATURES)
    assert L.mpm_version() == 1


def test_create_without_device_fails_loudly():
    """On a host without a GPU, mpm_create reports MPM_ERR_NO_DEVICE instead of falling back to the CPU."""
    import torch  # noqa: F401  (only to check whether this host has a GPU)
    if torch.cuda.is_available():
        return
    import mpmsnow_amd

    import pytest

    with pytest.raises(mpmsnow_amd.MpmError):
        mpmsnow_amd.Context(0)
