"""Index arithmetic of the tiled (bi-level) layout on the host.

The solver stores every per-cell field tile after tile: a grid of shape
`shape` is padded up to whole tiles of shape `shape_i`, and cell x lives at
n_o * prod(shape_i) + n_i, with n_o the row-major index of its tile
x // shape_i in the tile grid shape_o and n_i the row-major index of x %
shape_i inside the tile (the layout of agd/FiniteDifferences.py:156-222 and
Update.h:60-66).  The data movement itself runs on the device
(agd_block_expand / agd_block_squeeze in libagd_hip.so); the host only needs
tile counts and the tiled index of a few cells (seeds, receivers).
"""
import numpy as np


def round_up_ratio(num, den):
    """Smallest integer >= num/den, elementwise (num, den integers, den>0)."""
    num, den = np.asarray(num), np.asarray(den)
    return (num + den - 1) // den


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
