"""Tilted transversally isotropic norms (agd/Metrics/Seismic/tti.py).

The dual unit ball is l(X^2 [+ Y^2], Z^2) + 1/2 q(X^2 [+ Y^2], Z^2) = 1 with
l linear, q quadratic and (X, [Y,] Z) the coefficients of A^-T p, where A is
the `inverse_transformation` of the primal norm N(x) = N0(A x).  The eikonal
solver consumes `flatten(transposed_transformation=True)`:
[l0, l1, q00, q10, q11, A^-T row-major]; the grid scale is applied on the
device (A -> A diag(h), Metrics/base.py:175-185).
"""
import numpy as np

from .. import Base, _np, flatten_symmetric_matrix, expand_symmetric_matrix


class TTI(Base):
    """linear (2, *shape), quadratic (2, 2, *shape), optional
    inverse_transformation (d, d, *shape); vdim 2 or 3."""

    def __init__(self, linear, quadratic, vdim=None, inverse_transformation=None):
        self.linear = linear
        self.quadratic = quadratic
        self.inverse_transformation = inverse_transformation
        if np.shape(linear)[0] != 2 or tuple(np.shape(quadratic)[:2]) != (2, 2):
            raise ValueError("TTI expects linear of shape (2, ...) and quadratic of shape (2, 2, ...)")
        if vdim is None:
            if inverse_transformation is not None:
                vdim = np.shape(inverse_transformation)[0]
            elif np.ndim(linear) > 1:
                vdim = np.ndim(linear) - 1
            else:
                raise ValueError("Unspecified dimension")
        self._vdim = int(vdim)

    @property
    def vdim(self):
        return self._vdim

    @property
    def shape(self):
        return tuple(np.shape(self.linear)[1:])

    def model_HFM(self):
        return f"TTI{self.vdim}"

    @classmethod
    def from_hexagonal(cls, c11, _, c13, c33, c44, vdim=3):
        """Hexagonal Hooke tensor coefficients (tti.py:141-146)."""
        linear = [c11 + c44, c33 + c44]
        mixed = c13 ** 2 - c11 * c33 + 2. * c13 * c44
        quadratic = [[-2. * c11 * c44, mixed], [mixed, -2. * c33 * c44]]
        return cls(np.asarray(linear, dtype=float), np.asarray(quadratic, dtype=float), vdim=vdim)

    @classmethod
    def from_cast(cls, metric):
        if isinstance(metric, cls):
            return metric
        raise ValueError("No casting operations supported towards the TTI model")

    def inv_transform(self, a):
        """Compose the inverse transformation with a (implicit_base.py:57-63)."""
        a = _np(a)
        t = a if self.inverse_transformation is None else np.einsum(
            'ij...,jk...->ik...', _np(self.inverse_transformation), a)
        return TTI(self.linear, self.quadratic, self.vdim, t)

    def with_costs(self, costs):
        costs = np.asarray(_np(costs), dtype=float)
        d = self.vdim
        a = np.zeros((d, d) + costs.shape[1:])
        for i in range(d):
            a[i, i] = costs[i]
        return self.inv_transform(a)

    def flatten(self, transposed_transformation=False):
        lin = np.asarray(_np(self.linear), dtype=float)
        quad = np.asarray(_np(self.quadratic), dtype=float)
        shape = np.broadcast_shapes(lin.shape[1:], quad.shape[2:],
                                    () if self.inverse_transformation is None
                                    else np.shape(self.inverse_transformation)[2:])
        d = self.vdim
        trans = (np.broadcast_to(np.eye(d).reshape((d, d) + (1,) * len(shape)), (d, d) + shape)
                 if self.inverse_transformation is None
                 else np.broadcast_to(_np(self.inverse_transformation), (d, d) + shape))
        if transposed_transformation:
            inv = np.moveaxis(np.linalg.inv(np.moveaxis(trans, (0, 1), (-2, -1))), (-2, -1), (0, 1))
            trans = np.swapaxes(inv, 0, 1)
        lin = np.broadcast_to(lin.reshape(lin.shape[:1] + (1,) * (len(shape) - lin.ndim + 1) + lin.shape[1:]),
                              (2,) + shape)
        quad = np.broadcast_to(quad.reshape(quad.shape[:2] + (1,) * (len(shape) - quad.ndim + 2) + quad.shape[2:]),
                               (2, 2) + shape)
        return np.concatenate([lin, flatten_symmetric_matrix(quad), trans.reshape((d * d,) + shape)], axis=0)

    @classmethod
    def expand(cls, arr):
        arr = np.asarray(_np(arr))
        vdim = int(round(np.sqrt(len(arr) - 5)))
        if vdim * vdim + 5 != len(arr):
            raise ValueError("TTI.expand expects 5 + d*d components")
        return cls(arr[0:2], expand_symmetric_matrix(arr[2:5]), vdim=vdim,
                   inverse_transformation=arr[5:].reshape((vdim, vdim) + arr.shape[1:]))

    @classmethod
    def from_HFM(cls, arr):
        return cls.expand(arr)
