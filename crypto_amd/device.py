"""Device-resident helpers over the kinet_hip_* entry points.

PyTorch is used only as plumbing: device allocations (uint8 tensors in HBM)
and the stream handle.  All arithmetic runs in the HIP library.
"""
from __future__ import annotations

import torch

from .lib import GPUKIT_OK, GPUKIT_STATUS, KinetError, point_bytes, require_device


def affine_words(curve: int) -> int:
    """Words per prepared (device-form) affine point record."""
    return 28 if curve == 2 else 20


def stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def to_device(data: bytes, device="cuda") -> torch.Tensor:
    t = torch.frombuffer(bytearray(data), dtype=torch.uint8)
    return t.to(device)


class PreparedPoints:
    """Points uploaded once and converted to the device's Montgomery affine form."""

    def __init__(self, curve: int, wire: bytes | torch.Tensor, device="cuda"):
        lib = require_device()
        self.curve = curve
        pb = point_bytes(curve)
        d_wire = wire if isinstance(wire, torch.Tensor) else to_device(wire, device)
        self.n = d_wire.numel() // pb
        self.pts = torch.empty(self.n * affine_words(curve) * 4 + 16, dtype=torch.uint8, device=device)
        self.inf = torch.empty(self.n + 16, dtype=torch.uint8, device=device)
        rc = lib.kinet_hip_points_prepare(curve, d_wire.data_ptr(), self.n, self.pts.data_ptr(),
                                          self.inf.data_ptr(), stream_handle())
        if rc != GPUKIT_OK:
            raise KinetError("kinet_hip_points_prepare", rc, GPUKIT_STATUS)
        torch.cuda.current_stream().synchronize()


def msm_prepared(pp: PreparedPoints, d_scalars: torch.Tensor, n: int | None = None, window_bits: int = 0,
                 out: torch.Tensor | None = None, stream=None, sync: bool = True, chunk: int = 0) -> torch.Tensor:
    """MSM over prepared points; returns the device output record (wire || compressed)."""
    lib = require_device()
    n = pp.n if n is None else n
    pb = point_bytes(pp.curve)
    if out is None:
        out = torch.empty(pb + pb // 2, dtype=torch.uint8, device=d_scalars.device)
    rc = lib.kinet_hip_msm_prepared(pp.curve, d_scalars.data_ptr(), pp.pts.data_ptr(), pp.inf.data_ptr(), n,
                                    window_bits, chunk, out.data_ptr(), stream_handle(stream))
    if rc != GPUKIT_OK:
        raise KinetError("kinet_hip_msm_prepared", rc, GPUKIT_STATUS)
    if sync:
        (stream or torch.cuda.current_stream()).synchronize()
    return out
