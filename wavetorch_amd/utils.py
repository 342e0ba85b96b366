"""dtype plumbing and the loss head (reference: wavetorch/utils.py:6-36)."""
from __future__ import annotations

import numpy as np
import torch


def to_tensor(x, dtype=None):
    """utils.py:6-11: numpy arrays via from_numpy, everything else via torch.tensor,
    in the default dtype unless given."""
    dtype = dtype if dtype is not None else torch.get_default_dtype()
    if type(x) is np.ndarray:
        return torch.from_numpy(x).type(dtype=dtype)
    if isinstance(x, torch.Tensor):  # what torch.tensor(x) yields, without its warning
        return x.detach().clone().to(dtype=dtype)
    return torch.tensor(x, dtype=dtype)


def set_dtype(dtype=None):
    """utils.py:14-20.  Modules built after set_dtype('float64') hold float64
    buffers and run the double-precision time loop (one launch per step,
    every field kept); float32 runs the temporally blocked kernels."""
    if dtype == 'float32' or dtype is None:
        torch.set_default_dtype(torch.float32)
    elif dtype == 'float64':
        torch.set_default_dtype(torch.float64)
    else:
        raise ValueError('Unsupported data type: %s; should be either float32 or float64' % dtype)


def window_data(X, window_length):
    """utils.py:23-26: crop to `window_length` centred at the middle."""
    return X[int(len(X) / 2 - window_length / 2):int(len(X) / 2 + window_length / 2)]


def accuracy_onehot(y_pred, y_label):
    """utils.py:29-32."""
    return (y_pred.argmax(dim=1) == y_label).float().mean().item()


def normalize_power(X):
    """utils.py:35-36: probe powers normalised to sum to one over the probes."""
    return X / torch.sum(X, dim=1, keepdim=True)
