"""Helpers for GPU tests: device buffers through torch (plumbing only) and
device-to-host copies through the HIP runtime that torch already loaded."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        h = C.CDLL("libamdhip64.so.7")
        h.hipMemcpy.restype = C.c_int
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.restype = C.c_int
        _hip = h
    return _hip


def d2h(ptr: int, count: int, dtype) -> np.ndarray:
    out = np.empty(count, dtype)
    if count:
        assert ptr, "null device pointer"
        rc = hip().hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, 2)  # hipMemcpyDeviceToHost
        assert rc == 0, f"hipMemcpy failed: {rc}"
    return out


def to_device(torch, data: np.ndarray, offset: int = 0):
    """Copy bytes to a fresh device buffer; returns (tensor, device pointer of byte 0)."""
    t = torch.empty(data.size + offset + 64, dtype=torch.uint8, device="cuda")
    if data.size:
        t[offset : offset + data.size].copy_(torch.from_numpy(np.ascontiguousarray(data)))
    torch.cuda.synchronize()
    return t, t.data_ptr() + offset


def say(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def gpu_full(torch, scanner, data: np.ndarray, offset: int = 0, want_pairs: bool = True, run_pairs: bool = True):
    """Run scan_device and copy every output back."""
    t, ptr = to_device(torch, data, offset)
    stream = torch.cuda.current_stream().cuda_stream
    r = scanner.scan_device(ptr, data.size, stream, want_pairs=want_pairs, run_pairs=run_pairs)
    out = {
        "n_struct": r.n_struct,
        "n_brackets": r.n_brackets,
        "positions": d2h(r.structural, r.n_struct, np.uint32),
        "bracket_pos": d2h(r.bracket_pos, r.n_brackets, np.uint32),
        "bracket_char": d2h(r.bracket_char, r.n_brackets, np.uint8),
        "balanced": r.balanced,
        "well_formed": r.well_formed,
        "max_depth": r.max_depth,
        "unclosed_mode": r.unclosed_mode,
    }
    if r.chars:
        out["chars"] = d2h(r.chars, r.n_struct, np.uint8)
    if r.struct_bitmap:
        out["struct_bitmap"] = d2h(r.struct_bitmap, (data.size + 63) // 64, np.uint64)
    if run_pairs and r.bracket_pair:
        out["pair"] = d2h(r.bracket_pair, r.n_brackets, np.int32)
    if want_pairs:
        out["pair_pos"] = d2h(r.pair_pos, r.n_struct, np.int32)
    del t
    return out
