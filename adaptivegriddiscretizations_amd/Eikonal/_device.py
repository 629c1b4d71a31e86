"""Device-array plumbing: torch tensors on the current HIP device hold the
solver's buffers, and their raw pointers are handed to libagd_hip.so."""
import numpy as np

try:
    import torch
except ImportError as e:  # pragma: no cover - torch is part of the image
    raise ImportError("torch (ROCm build) is required for device buffers") from e

_TORCH_DTYPES = {
    np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
    np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
    np.dtype(np.uint8): torch.uint8, np.dtype(np.bool_): torch.bool,
}


def tdtype(dtype):
    return _TORCH_DTYPES[np.dtype(dtype)]


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("No HIP device available: the eikonal solver runs on the GPU only")
    return torch.device('cuda', torch.cuda.current_device())


def stream_handle():
    return torch.cuda.current_stream().cuda_stream


def device_tensor(x, dtype):
    """Contiguous device tensor of the given numpy dtype (copies host data)."""
    dev = device()
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=dev, dtype=tdtype(dtype)).contiguous()
    arr = np.asarray(host_array(x), dtype=dtype)
    if not arr.flags.c_contiguous or not arr.flags.writeable:  # (np.ascontiguousarray turns 0-d into 1-d)
        arr = arr.copy(order='C')
    return torch.from_numpy(arr).to(dev)


def host_array(x):
    """numpy view/copy of x (torch tensors, and arrays of other device
    libraries that expose .get(), are copied to the host)."""
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    if not isinstance(x, np.ndarray) and hasattr(x, 'get') and hasattr(x, 'shape'):
        return np.asarray(x.get())
    return np.asarray(x)


def empty(shape, dtype):
    return torch.empty(shape, dtype=tdtype(dtype), device=device())


def zeros(shape, dtype):
    return torch.zeros(shape, dtype=tdtype(dtype), device=device())


def full(shape, value, dtype):
    return torch.full(shape, value, dtype=tdtype(dtype), device=device())


def ptr(t):
    """Raw device pointer of a contiguous tensor (None for None)."""
    if t is None:
        return None
    assert t.is_contiguous()
    return t.data_ptr()
