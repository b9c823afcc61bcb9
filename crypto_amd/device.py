"""Device-resident helpers over the kinet_hip_* entry points.

PyTorch is used only as plumbing: device allocations (uint8 tensors in HBM)
and the stream handle.  All arithmetic runs in the HIP library.
"""
from __future__ import annotations

import torch

from .lib import GPUKIT_OK, GPUKIT_STATUS, KinetError, point_bytes, require_device


def affine_words(curve: int) -> int:
    """Words per prepared (device-form) affine point record."""
    return 32 if curve in (2, 3) else 20


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
        self.pts = torch.empty(lib.kinet_hip_points_bytes(curve, self.n) + 16, dtype=torch.uint8, device=device)
        self.inf = torch.empty(self.n + 16, dtype=torch.uint8, device=device)
        rc = lib.kinet_hip_points_prepare(curve, d_wire.data_ptr(), self.n, self.pts.data_ptr(),
                                          self.inf.data_ptr(), stream_handle())
        if rc != GPUKIT_OK:
            raise KinetError("kinet_hip_points_prepare", rc, GPUKIT_STATUS)
        torch.cuda.current_stream().synchronize()


def msm_prepared(pp: PreparedPoints, d_scalars: torch.Tensor, n: int | None = None, window_bits: int = 0,
                 out: torch.Tensor | None = None, stream=None, sync: bool = True, chunk: int = 0,
                 g1: bool = False) -> torch.Tensor:
    """MSM over prepared points; returns the device output record (wire || compressed).
    g1=True: the caller guarantees the points lie in G1 (kinet_hip_msm_prepared_g1)."""
    lib = require_device()
    n = pp.n if n is None else n
    pb = point_bytes(pp.curve)
    if out is None:
        out = torch.empty(pb + pb // 2, dtype=torch.uint8, device=d_scalars.device)
    fn = lib.kinet_hip_msm_prepared_g1 if g1 else lib.kinet_hip_msm_prepared
    rc = fn(pp.curve, d_scalars.data_ptr(), pp.pts.data_ptr(), pp.inf.data_ptr(), n, window_bits, chunk,
            out.data_ptr(), stream_handle(stream))
    if rc != GPUKIT_OK:
        raise KinetError(fn.__name__, rc, GPUKIT_STATUS)
    if sync:
        (stream or torch.cuda.current_stream()).synchronize()
    return out


class PointsTable:
    """Window table of prepared points (kinet_hip_points_table_build): row w
    holds 2^(c w) P_i, so an MSM over it reduces one window and combines none."""

    def __init__(self, pp: PreparedPoints, window_bits: int = 0, stream=None):
        lib = require_device()
        self.pp = pp
        self.curve, self.n = pp.curve, pp.n
        self.c = window_bits or lib.kinet_hip_table_window_bits(pp.n)
        nbytes = lib.kinet_hip_table_bytes(pp.curve, pp.n, self.c)
        if nbytes == 0 and pp.n:
            raise KinetError("kinet_hip_table_bytes", -2, GPUKIT_STATUS)
        self.table = torch.empty(max(nbytes, 16), dtype=torch.uint8, device=pp.pts.device)
        rc = lib.kinet_hip_points_table_build(pp.curve, pp.pts.data_ptr(), pp.inf.data_ptr(), pp.n, self.c,
                                              self.table.data_ptr(), stream_handle(stream))
        if rc != GPUKIT_OK:
            raise KinetError("kinet_hip_points_table_build", rc, GPUKIT_STATUS)
        (stream or torch.cuda.current_stream()).synchronize()

    def msm(self, d_scalars: torch.Tensor, n: int | None = None, out: torch.Tensor | None = None, stream=None,
            sync: bool = True) -> torch.Tensor:
        """MSM of the first n table points (default all) with the scalars; device output record."""
        lib = require_device()
        n = self.n if n is None else n
        pb = point_bytes(self.curve)
        if out is None:
            out = torch.empty(pb + pb // 2, dtype=torch.uint8, device=d_scalars.device)
        rc = lib.kinet_hip_msm_table(self.curve, d_scalars.data_ptr(), self.table.data_ptr(),
                                     self.pp.inf.data_ptr(), self.n, n, self.c, out.data_ptr(), stream_handle(stream))
        if rc != GPUKIT_OK:
            raise KinetError("kinet_hip_msm_table", rc, GPUKIT_STATUS)
        if sync:
            (stream or torch.cuda.current_stream()).synchronize()
        return out


def check_g1(pp: PreparedPoints) -> bool:
    """True if every non-identity prepared point is on the curve and in G1."""
    lib = require_device()
    bad = torch.zeros(1, dtype=torch.int32, device=pp.pts.device)
    rc = lib.kinet_hip_points_check_g1(pp.curve, pp.pts.data_ptr(), pp.inf.data_ptr(), pp.n, bad.data_ptr(),
                                       stream_handle())
    if rc != GPUKIT_OK:
        raise KinetError("kinet_hip_points_check_g1", rc, GPUKIT_STATUS)
    return int(bad.item()) == 0


def points_decompress(curve: int, d_comp: torch.Tensor, stream=None):
    """Compressed G1 encodings on the device (n x 48 / 32 bytes) -> (wire
    records n x 96 / 64 bytes, all valid?) through kinet_hip_points_decompress."""
    lib = require_device()
    cb = point_bytes(curve) // 2
    n = d_comp.numel() // cb
    wire = torch.empty(n * 2 * cb, dtype=torch.uint8, device=d_comp.device)
    bad = torch.zeros(1, dtype=torch.int32, device=d_comp.device)
    rc = lib.kinet_hip_points_decompress(curve, d_comp.data_ptr(), n, wire.data_ptr(), bad.data_ptr(),
                                         stream_handle(stream))
    if rc != GPUKIT_OK:
        raise KinetError("kinet_hip_points_decompress", rc, GPUKIT_STATUS)
    (stream or torch.cuda.current_stream()).synchronize()
    return wire, int(bad.item()) == 0


class MsmBatch:
    """A plan of many MSMs run as one pipeline (kinet_hip_msm_batch_*).

    sizes[j] points per job; sc_off / pt_off as in the C header (None: scalars
    back to back, every job from point 0)."""

    def __init__(self, curve: int, sizes, sc_off=None, pt_off=None, window_bits: int = 0):
        import ctypes

        lib = require_device()
        self.curve = curve
        self.count = len(sizes)
        n_arr = (ctypes.c_size_t * self.count)(*sizes)
        sc_arr = (ctypes.c_uint64 * self.count)(*sc_off) if sc_off is not None else None
        pt_arr = (ctypes.c_uint32 * self.count)(*pt_off) if pt_off is not None else None
        h = ctypes.c_void_p()
        rc = lib.kinet_hip_msm_batch_create(curve, n_arr, sc_arr, pt_arr, self.count, window_bits, ctypes.byref(h))
        if rc != GPUKIT_OK:
            raise KinetError("kinet_hip_msm_batch_create", rc, GPUKIT_STATUS)
        self.handle = h.value
        self.record = point_bytes(curve) + point_bytes(curve) // 2

    def window_bits(self, job: int) -> int:
        return require_device().kinet_hip_msm_batch_window_bits(self.handle, job)

    def run(self, pp: PreparedPoints, d_scalars: torch.Tensor, out: torch.Tensor | None = None,
            use_graph: bool = False, stream=None, sync: bool = True) -> torch.Tensor:
        """Runs every job; returns the device records (count x record bytes)."""
        lib = require_device()
        if out is None:
            out = torch.empty(self.count * self.record, dtype=torch.uint8, device=d_scalars.device)
        rc = lib.kinet_hip_msm_batch_run(self.handle, d_scalars.data_ptr(), pp.pts.data_ptr(), pp.inf.data_ptr(),
                                         out.data_ptr(), 1 if use_graph else 0, stream_handle(stream))
        if rc != GPUKIT_OK:
            raise KinetError("kinet_hip_msm_batch_run", rc, GPUKIT_STATUS)
        if sync:
            (stream or torch.cuda.current_stream()).synchronize()
        return out

    def close(self):
        if getattr(self, "handle", None):
            require_device().kinet_hip_msm_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class JobSet:
    """A fixed set of MSMs and Fr NTT groups run as one HIP graph launch
    (kinet_hip_job_set_*).  groups: [(log2 n, count, first element offset)]."""

    def __init__(self, curve: int, msm_sizes, field: int, groups, inverse: bool = False, sc_off=None, pt_off=None):
        import ctypes

        from .lib import KinetNttGroup

        lib = require_device()
        self.curve = curve
        self.count = len(msm_sizes)
        n_arr = (ctypes.c_size_t * max(1, self.count))(*msm_sizes)
        sc_arr = (ctypes.c_uint64 * self.count)(*sc_off) if sc_off is not None else None
        pt_arr = (ctypes.c_uint32 * self.count)(*pt_off) if pt_off is not None else None
        g_arr = (KinetNttGroup * max(1, len(groups)))(*[KinetNttGroup(lg, c, o) for lg, c, o in groups])
        h = ctypes.c_void_p()
        rc = lib.kinet_hip_job_set_create(curve, n_arr, sc_arr, pt_arr, self.count, field, 1 if inverse else 0,
                                          g_arr, len(groups), ctypes.byref(h))
        if rc != GPUKIT_OK:
            raise KinetError("kinet_hip_job_set_create", rc, GPUKIT_STATUS)
        self.handle = h.value
        self.record = point_bytes(curve) + point_bytes(curve) // 2

    def run(self, pp: PreparedPoints | None, d_scalars, out, d_ntt_in, d_ntt_out, d_ntt_tmp, d_err, stream=None,
            sync: bool = True):
        lib = require_device()

        def ptr(t):
            return t.data_ptr() if t is not None else None

        s = stream if stream is not None else torch.cuda.current_stream()
        rc = lib.kinet_hip_job_set_run(self.handle, ptr(d_scalars), pp.pts.data_ptr() if pp else None,
                                       pp.inf.data_ptr() if pp else None, ptr(out), ptr(d_ntt_in), ptr(d_ntt_out),
                                       ptr(d_ntt_tmp), ptr(d_err), s.cuda_stream)
        if rc != GPUKIT_OK:
            raise KinetError("kinet_hip_job_set_run", rc, GPUKIT_STATUS)
        if sync:
            s.synchronize()

    def close(self):
        if getattr(self, "handle", None):
            require_device().kinet_hip_job_set_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ntt_batch(field: int, d_in: torch.Tensor, n: int, count: int, inverse: bool = False,
              d_out: torch.Tensor | None = None, d_tmp: torch.Tensor | None = None, d_err: torch.Tensor | None = None,
              stream=None, sync: bool = True) -> torch.Tensor:
    """count same-size Fr transforms back to back (kinet_hip_ntt_fr_batch)."""
    from .lib import CRYPTO_STATUS

    lib = require_device()
    if d_out is None:
        d_out = torch.empty(count * n * 32, dtype=torch.uint8, device=d_in.device)
    if d_tmp is None:
        d_tmp = torch.empty(count * n * 32, dtype=torch.uint8, device=d_in.device)
    own_err = d_err is None
    if own_err:
        d_err = torch.zeros(1, dtype=torch.int32, device=d_in.device)
    rc = lib.kinet_hip_ntt_fr_batch(field, 1 if inverse else 0, d_in.data_ptr(), d_out.data_ptr(), d_tmp.data_ptr(),
                                    n, count, d_err.data_ptr(), stream_handle(stream))
    if rc != 0:
        raise KinetError("kinet_hip_ntt_fr_batch", rc, CRYPTO_STATUS)
    if sync:
        (stream or torch.cuda.current_stream()).synchronize()
        if own_err and int(d_err.item()) != 0:
            raise KinetError("kinet_hip_ntt_fr_batch", -1, CRYPTO_STATUS)
    return d_out
