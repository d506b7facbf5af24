"""GPU FFI smoke tests of the boundary, restated from the reference's
python/tests/test_sim_core_cuda_poc.py:45-123 (crates/sim-core/src/cuda.rs:51-68,
crates/cuda-kernels/src/scale.cu:22-34): a device buffer allocated on the
Python side is handed to the library as a raw address taken from the public
``__cuda_array_interface__`` (and from ``data_ptr()``), the library's kernel
mutates it in place, and Python reads the mutation back."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import host

pytestmark = pytest.mark.gpu


def _cai_ptr(t) -> int:
    return int(t.__cuda_array_interface__["data"][0])


def test_array_interface_pointer_matches_data_ptr(torch_gpu):
    """The public protocol and the tensor's own address agree (the
    reference's pointer-provenance safeguard), and boilsim.interop reads the
    same address."""
    torch = torch_gpu
    from boilsim import interop

    t = torch.ones(64, dtype=torch.float32, device="cuda")
    assert _cai_ptr(t) == t.data_ptr()
    assert interop.device_ptr(t, shape=(64,), dtype=np.float32, writable=True) == t.data_ptr()


def test_scale_array_doubles_in_place(simcore, torch_gpu):
    torch = torch_gpu
    n = 4096
    t = torch.ones(n, dtype=torch.float32, device="cuda")
    simcore.scale_array(_cai_ptr(t), n, 2.0)
    torch.cuda.synchronize()
    back = host(t)
    assert back.dtype == np.float32 and back.shape == (n,)
    assert (back == 2.0).all()


def test_scale_array_arbitrary_factor_bit_exact(simcore, torch_gpu):
    torch = torch_gpu
    n = 1024
    src = np.linspace(0.0, 1.0, n, dtype=np.float32)
    t = torch.from_numpy(src.copy()).cuda()
    simcore.scale_array(t.data_ptr(), n, 0.5)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(t), src * np.float32(0.5))
    # A factor that is not a power of two: one correctly rounded multiply.
    u = torch.from_numpy(src.copy()).cuda()
    simcore.scale_array(u.data_ptr(), n, 0.3)
    torch.cuda.synchronize()
    assert np.array_equal(host(u).view(np.uint32), (src * np.float32(0.3)).view(np.uint32))


def test_scale_array_zero_n_is_noop(simcore, torch_gpu):
    torch = torch_gpu
    t = torch.tensor([3.0, 4.0, 5.0], dtype=torch.float32, device="cuda")
    simcore.scale_array(_cai_ptr(t), 0, 99.0)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(host(t), [3.0, 4.0, 5.0])


def test_scale_array_composes_across_calls(simcore, torch_gpu):
    torch = torch_gpu
    n = 256
    t = torch.ones(n, dtype=torch.float32, device="cuda")
    for s in (2.0, 3.0, 5.0):
        simcore.scale_array(_cai_ptr(t), n, s)
    torch.cuda.synchronize()
    assert (host(t) == 30.0).all()


def test_scale_array_rejects_bad_arguments(simcore):
    """usize and pointer checks of the shim (crates/cuda-kernels/src/lib.rs:250-252)."""
    with pytest.raises(OverflowError):
        simcore.scale_array(0, -1, 2.0)
    with pytest.raises(RuntimeError, match="overflows i32"):
        simcore.scale_array(0, 2**31, 2.0)
    with pytest.raises(OverflowError):
        simcore.scale_array(-5, 4, 2.0)
