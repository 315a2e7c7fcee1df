"""Host-side argument checks of the Python layer (no GPU needed): device
tensors handed to the kernels as raw pointers must be contiguous float64 (or
uint8 flags) arrays of the right shape, all on one device."""
import pytest
import torch

import taichi_collisional_amd as tca


def _state(n, **over):
    f8 = torch.float64
    st = dict(pos=torch.zeros(n, 3, dtype=f8), pos_e=torch.zeros(n, 3, dtype=f8),
              vel=torch.zeros(n, 3, dtype=f8), mass=torch.ones(n, dtype=f8),
              timestep=torch.zeros(n, dtype=f8), acc=torch.zeros(n, 3, dtype=f8),
              pot=torch.zeros(n, dtype=f8), acc_old=torch.zeros(n, dtype=f8))
    st.update(over)
    return st


def _specs(st):
    vec = ("pos", "pos_e", "vel", "acc")
    return [(k, t, torch.float64, (3,) if k in vec else ()) for k, t in st.items()]


def test_check_tensors_accepts_good_arrays():
    assert tca.check_tensors(_specs(_state(10)), 10, device_type="cpu") == torch.device("cpu")


@pytest.mark.parametrize("bad,exc", [
    (dict(mass=torch.ones(10, dtype=torch.float32)), TypeError),
    (dict(pos=torch.zeros(10, 2, dtype=torch.float64)), ValueError),
    (dict(acc_old=torch.zeros(9, dtype=torch.float64)), ValueError),
    (dict(vel=torch.zeros(3, 10, dtype=torch.float64).t()), TypeError),
    (dict(pot=[0.0] * 10), TypeError),
])
def test_check_tensors_rejects(bad, exc):
    with pytest.raises(exc):
        tca.check_tensors(_specs(_state(10, **bad)), 10, device_type="cpu")


def test_host_tensors_never_reach_the_kernels():
    """A CPU state is refused before any C call (no context is needed)."""
    f = tca.Fmm.__new__(tca.Fmm)
    f.h = None
    with pytest.raises(TypeError, match="cuda"):
        f.master_evolve(_state(10), 0.01)
    with pytest.raises(TypeError, match="lacks"):
        st = _state(10)
        del st["pos_e"]
        f.master_evolve(st, 0.01)
    x = torch.zeros(10, 3, dtype=torch.float64)
    with pytest.raises(TypeError, match="cuda"):
        f.eval(x, x, torch.ones(10, dtype=torch.float64))
    with pytest.raises(TypeError, match="cuda"):
        f.master_eval(x, x, torch.ones(10, dtype=torch.float64))
