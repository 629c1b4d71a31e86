"""denseAD: value and first-order coefficients (Dense.py:14-44 of agd).

`value` has shape S, `coef` has shape S + (size_ad,): coef[..., i] is the
derivative of value along the i-th symbolic perturbation.  Arrays may be
numpy arrays or torch tensors (on the HIP device for the solver's outputs).
"""


class denseAD:
    """Dense first-order forward AD variable (carrier only)."""

    def __init__(self, value, coef=None):
        if isinstance(value, denseAD):
            if coef is not None:
                raise ValueError("denseAD(denseAD, coef) is ambiguous")
            value, coef = value.value, value.coef
        self.value = value
        if coef is None:
            coef = _zeros_like(value, tuple(value.shape) + (0,))
        if tuple(coef.shape[:-1]) != tuple(value.shape):
            raise ValueError(f"coef shape {tuple(coef.shape)} does not extend value shape "
                             f"{tuple(value.shape)}")
        self.coef = coef

    @property
    def shape(self):
        return tuple(self.value.shape)

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def size_ad(self):
        return int(self.coef.shape[-1])

    def gradient(self, i=None):
        """coef moved to the front: gradient()[i] == coef[..., i]."""
        c = self.coef
        if hasattr(c, 'movedim'):
            g = c.movedim(-1, 0)
        else:
            import numpy as np
            g = np.moveaxis(c, -1, 0)
        return g if i is None else g[i]

    def __getitem__(self, key):
        return denseAD(self.value[key], self.coef[key])

    def __repr__(self):
        return f"denseAD({self.value!r},{self.coef!r})"


def _zeros_like(value, shape):
    if hasattr(value, 'new_zeros'):
        return value.new_zeros(shape)
    import numpy as np
    return np.zeros(shape, dtype=np.asarray(value).dtype)
