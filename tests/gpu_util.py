"""Helpers for the GPU tests: device buffers (torch as plumbing) and copies
of raw device pointers through the HIP runtime."""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import torch  # noqa: F401  (torch's HIP runtime first, then bind it by soname)

        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


def d2h(ptr: int, count: int, dtype) -> np.ndarray:
    """Copy `count` elements of `dtype` from a device address."""
    out = np.empty(count, dtype=dtype)
    if count:
        rc = hip().hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2)  # hipMemcpyDeviceToHost
        assert rc == 0, f"hipMemcpy failed ({rc})"
    return out


def dev(a: np.ndarray, dtype=None):
    """numpy -> torch device tensor (complex128 by default for complex input)."""
    import torch

    t = torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype)))
    return t.to("cuda")


def host(t) -> np.ndarray:
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def empty(n, dtype="complex128"):
    import torch

    return torch.zeros(n, dtype=getattr(torch, dtype), device="cuda")


def rel_err(got, ref) -> float:
    ref = np.asarray(ref)
    scale = np.max(np.abs(ref)) if ref.size else 1.0
    return float(np.max(np.abs(np.asarray(got) - ref)) / (scale if scale > 0 else 1.0))
