"""Host-side checks of the torch bridge and the Lame entry points: argument errors are raised before any device
work, so these run without a GPU."""
import ctypes

import numpy as np
import pytest

from litepgo_amd import _native


def test_lame_entry_points_reject_bad_arguments():
    lib = _native.lib()
    buf = (ctypes.c_double * 4)()
    assert lib.lpg_lame_device(buf, buf, -1, buf, None) == _native.LPG_INVALID_ARGUMENT
    assert lib.lpg_lame_device(None, buf, 2, buf, None) == _native.LPG_INVALID_ARGUMENT
    assert lib.lpg_lame_vjp_device(buf, buf, None, 2, buf, buf, None) == _native.LPG_INVALID_ARGUMENT
    # n == 0 enqueues nothing and succeeds.
    assert lib.lpg_lame_device(None, None, 0, None, None) == _native.LPG_OK
    assert lib.lpg_lame_vjp_device(None, None, None, 0, None, None, None) == _native.LPG_OK


def test_vjp_entry_point_rejects_a_null_handle():
    lib = _native.lib()
    assert lib.lpg_tet_energy_vjp_device(None, None, None, None, None, None, None, None) == \
        _native.LPG_INVALID_ARGUMENT


def test_bridge_argument_errors():
    torch = pytest.importorskip("torch")
    from litepgo_amd import torch as lt

    with pytest.raises(TypeError):
        lt.lame(np.ones(3), torch.ones(3, dtype=torch.float64))
    with pytest.raises(ValueError):
        lt.lame(torch.ones(3, dtype=torch.float64), torch.ones(4, dtype=torch.float64))
    with pytest.raises(TypeError):
        lt.energy(None, np.zeros(3))
    with pytest.raises(TypeError):
        lt.force(None, [0.0])
    with pytest.raises(TypeError):
        lt.energy_and_force(None, 1.0)
    with pytest.raises(TypeError):
        lt.force_fp(None, np.zeros(9), torch.zeros(3, dtype=torch.float64))


def test_force_fp_entry_points_reject_a_null_handle():
    lib = _native.lib()
    assert lib.lpg_tet_energy_force_fp_device(None, None, None, None, None, None) == _native.LPG_INVALID_ARGUMENT
    assert lib.lpg_tet_energy_force_vjp_fp_device(None, None, None, None, None, None, None) == \
        _native.LPG_INVALID_ARGUMENT
    assert lib.lpg_tet_energy_force_jvp_fp_device(None, None, None, None, None, None, None) == \
        _native.LPG_INVALID_ARGUMENT
