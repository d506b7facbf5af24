"""Host handle of the learned preconditioner network (the U-Net of
boiling-sim_amd/csrc/unet.hip): FP32 master parameters, their packed BF16
MFMA fragments, workspaces, and save/load.  The reference has no learned
component; this is new."""

from __future__ import annotations

import os

import numpy as np

CONVS = [(8, 8), (8, 16), (16, 32), (32, 32), (32, 16), (16, 8), (8, 8), (8, 1)]
NUM_PARAMS = sum(27 * ci * co + co for ci, co in CONVS)
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_pot_110.npz")


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (fan-in 27*cin), zero biases, last layer scaled by 0.1."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (ci, co) in enumerate(CONVS):
        std = np.sqrt(2.0 / (27 * ci)) * (0.1 if li == len(CONVS) - 1 else 1.0)
        out.append(rng.standard_normal(27 * ci * co) * std)
        out.append(np.zeros(co))
    return np.concatenate(out).astype(np.float32)


class UNet:
    """Parameters on the device plus their packed fragments."""

    def __init__(self, params: np.ndarray | None = None, device: str = "cuda", seed: int = 0):
        import torch

        from sim_core import _lib

        self._lib = _lib
        lib = _lib.load()
        assert int(lib.bs_unet_num_params()) == NUM_PARAMS, "network layout mismatch with the library"
        if params is None:
            params = init_params(seed)
        params = np.ascontiguousarray(params, dtype=np.float32)
        assert params.size == NUM_PARAMS
        self.device = device
        self.params = torch.from_numpy(params.copy()).to(device)
        self.packed = torch.zeros(int(lib.bs_unet_packed_bytes()), dtype=torch.uint8, device=device)
        self.pack()

    def pack(self, stream: int = 0) -> None:
        self._lib.call("bs_unet_pack", self.params.data_ptr(), self.packed.data_ptr(), stream)

    @staticmethod
    def workspace_bytes(shape, training: bool = False) -> int:
        from sim_core import _lib

        return int(_lib.load().bs_unet_workspace_bytes(*shape, 1 if training else 0))

    def forward(self, r, code, n_fluid: int, z, ws, training: bool = False, stream: int = 0) -> None:
        shape = tuple(r.shape)
        self._lib.call("bs_unet_forward", self.params.data_ptr(), self.packed.data_ptr(), r.data_ptr(),
                       code.data_ptr(), *shape, float(n_fluid), z.data_ptr(), ws.data_ptr(), ws.numel(),
                       1 if training else 0, stream)

    def backward(self, r, e, code, grads, loss2, ws, stream: int = 0) -> None:
        shape = tuple(r.shape)
        self._lib.call("bs_unet_backward", self.params.data_ptr(), self.packed.data_ptr(), r.data_ptr(),
                       e.data_ptr(), code.data_ptr(), *shape, grads.data_ptr(), loss2.data_ptr(),
                       ws.data_ptr(), ws.numel(), stream)

    def params_numpy(self) -> np.ndarray:
        return self.params.detach().cpu().numpy().copy()

    def save(self, path: str, **meta) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.params_numpy(), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS, device: str = "cuda") -> "UNet":
        with np.load(path, allow_pickle=False) as f:
            return cls(f["params"], device=device)


def build_codes(mat, stream: int = 0):
    """uint8 class codes of an int32 mask tensor (0 air, 1 solid, 2+#solid fluid)."""
    import torch

    from sim_core import _lib

    code = torch.empty(tuple(mat.shape), dtype=torch.uint8, device=mat.device)
    _lib.call("bs_build_codes", mat.data_ptr(), *mat.shape, 0, 2, code.data_ptr(), stream)
    return code
