"""Tiled (bi-level) layout helpers on host arrays.

Same conventions as the reference (agd/FiniteDifferences.py:156-222): the
inner (tile) axes are moved last, the array is padded up to a multiple of
the tile shape.  The device-side equivalents are agd_block_expand /
agd_block_squeeze in libagd_hip.so; these numpy versions serve host-side
bookkeeping (seed indices, tile counts).
"""
import numpy as np


def round_up_ratio(num, den):
    """Smallest integer >= num/den, elementwise (num, den integers, den>0)."""
    num, den = np.asarray(num), np.asarray(den)
    return (num + den - 1) // den


def block_expand(arr, shape_i, shape_o=None, **kwargs):
    """Pad and reshape so that the tile axes come last; kwargs go to np.pad."""
    ndim = len(shape_i)
    if ndim == 0:
        return arr
    shape_pre = arr.shape[:-ndim]
    ndim_pre = len(shape_pre)
    shape_tot = np.array(arr.shape[-ndim:])
    shape_i = np.array(shape_i)
    if shape_o is None:
        shape_o = round_up_ratio(shape_tot, shape_i)
    shape_pad = (0,) * ndim_pre + tuple(shape_o * shape_i - shape_tot)
    arr = np.pad(arr, tuple((0, s) for s in shape_pad), **kwargs)
    shape_interleaved = tuple(np.stack((shape_o, shape_i), axis=1).flatten())
    arr = arr.reshape(shape_pre + shape_interleaved)
    rg = np.arange(ndim)
    return np.moveaxis(arr, ndim_pre + 1 + 2 * rg, ndim_pre + ndim + rg)


def block_squeeze(arr, shape):
    """Inverse of block_expand."""
    ndim = len(shape)
    if ndim == 0:
        return arr
    shape_pre = arr.shape[:-2 * ndim]
    ndim_pre = len(shape_pre)
    shape_o = arr.shape[-2 * ndim:-ndim]
    shape_i = arr.shape[-ndim:]
    rg = np.arange(ndim)
    arr = np.moveaxis(arr, ndim_pre + ndim + rg, ndim_pre + 1 + 2 * rg)
    arr = arr.reshape(shape_pre + tuple(so * si for so, si in zip(shape_o, shape_i)))
    return arr[tuple(slice(0, s) for s in shape_pre + tuple(shape))]


def tiled_index(index, shape_i, shape_o):
    """Linear tiled index n_o*size_i + n_i of integer positions index[..., d]."""
    index = np.asarray(index, dtype=np.int64)
    shape_i = np.asarray(shape_i, dtype=np.int64)
    shape_o = np.asarray(shape_o, dtype=np.int64)
    xo, xi = index // shape_i, index % shape_i
    n_o = np.zeros(index.shape[:-1], dtype=np.int64)
    n_i = np.zeros(index.shape[:-1], dtype=np.int64)
    for k in range(index.shape[-1]):
        n_o = n_o * shape_o[k] + xo[..., k]
        n_i = n_i * shape_i[k] + xi[..., k]
    return n_o * int(np.prod(shape_i)) + n_i
