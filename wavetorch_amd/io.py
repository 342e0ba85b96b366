"""Saved-model files (the reference's format: wavetorch/io.py:13-87).

File format
-----------
One ``torch.save`` of a dict with these entries (the reference reads and
writes the first five; a reference loader ignores the sixth):

``model_geom_class_str``  name of the geometry class in ``wavetorch_amd.geom``
``model_state``           ``WaveRNN.state_dict()``: ``cell.dt``,
                          ``cell.satdamp_b0``, ``cell.satdamp_uth``,
                          ``cell.c_nl``, ``cell.geom.*``, then
                          ``sources.<i>.x|y`` and ``probes.<i>.x|y``
``history``               training history (a pandas DataFrame from
                          ``train``) or None
``history_geom_state``    list of geometry constructor-argument dicts, one
                          per saved iteration (``state_reconstruction_args``)
``cfg``                   the run configuration (its ``dtype`` entry sets the
                          default float type on load) or None
``model_ext``             this package's cell extensions (``c_clamp``)

Rebuilding a model from a file: the geometry from one entry of
``history_geom_state``; the cell from ``cell.dt`` and the nonlinear
coefficients ``cell.satdamp_b0``, ``cell.satdamp_uth``, ``cell.c_nl`` (the
reference rebuilds a linear cell here, io.py:82, and so silently drops a
nonlinear model's coefficients); one ``WaveSource`` per saved source index
and one ``WaveIntensityProbe`` per saved probe index, in index order (the
state holds no probe kind, so like the reference every probe comes back as
an intensity probe).  A source saved with several pixels (a line source)
comes back as a ``WaveSource`` over those pixels.

Loading unpickles the file.  By default only torch's restricted loader runs
(``weights_only=True``): tensors, containers and plain values, which covers
files saved without a pandas history.  A file holding a pandas history (what
``train`` saves) needs ``weights_only=False``, which runs code from the file:
only for files you trust.
"""
from __future__ import annotations

import os
import re
from copy import deepcopy

import torch

from . import geom as _geom
from .cell import WaveCell
from .probe import WaveIntensityProbe
from .rnn import WaveRNN
from .source import WaveSource
from .utils import set_dtype

_COORD_KEY = re.compile(r"^(sources|probes)\.(\d+)\.(x|y)$")
_NONLINEAR_KEYS = ("satdamp_b0", "satdamp_uth", "c_nl")


def _file_record(model, history, history_geom_state, cfg):
    g = model.cell.geom
    if history_geom_state is None:
        history_geom_state = [g.state_reconstruction_args()]
    clamp = getattr(model.cell, "c_clamp", None)
    return {
        "model_geom_class_str": type(g).__name__,
        "model_state": {k: t.detach().cpu() for k, t in model.state_dict().items()},
        "history": history,
        "history_geom_state": history_geom_state,
        "cfg": cfg,
        "model_ext": {"c_clamp": None if clamp is None else [float(v) for v in clamp]},
    }


def save_model(model, name, savedir='./study/', history=None, history_geom_state=None,
               cfg=None, verbose=True):
    """Writes ``savedir + name + '.pt'`` (io.py:13-40: same path rule, same
    default geometry history of one entry, the current geometry)."""
    os.makedirs(savedir, exist_ok=True)
    path = savedir + name + '.pt'
    if verbose:
        print("Saving model to %s" % path)
    torch.save(_file_record(model, history, history_geom_state, cfg), path)


def new_geometry(class_str, state):
    """Geometry of class ``class_str`` from a copy of its constructor
    arguments ``state`` (io.py:43-46)."""
    return getattr(_geom, class_str)(**deepcopy(state))


def _coords(state, kind):
    """{index: [x, y]} of the ``kind`` ('sources' or 'probes') buffers in a
    WaveRNN state dict."""
    out = {}
    for key, val in state.items():
        m = _COORD_KEY.match(key)
        if m and m.group(1) == kind:
            out.setdefault(int(m.group(2)), [None, None])["xy".index(m.group(3))] = val
    return [out[i] for i in sorted(out)]


def _as_coord(t):
    t = torch.as_tensor(t)
    return t.item() if t.numel() == 1 and t.dim() == 0 else t.tolist()


def model_from_state(class_str, geom_state, state, ext=None):
    """WaveRNN rebuilt from a geometry class name, its constructor arguments
    and a ``WaveRNN.state_dict`` (the reconstruction of io.py:49-87 with the
    nonlinear coefficients kept), in eval mode."""
    geometry = new_geometry(class_str, geom_state)
    nl = {k: float(state[f"cell.{k}"]) for k in _NONLINEAR_KEYS if f"cell.{k}" in state}
    clamp = (ext or {}).get("c_clamp")
    cell = WaveCell(float(state["cell.dt"]), geometry, c_clamp=clamp, **nl)
    sources = [WaveSource(_as_coord(x), _as_coord(y)) for x, y in _coords(state, "sources")]
    probes = [WaveIntensityProbe(_as_coord(x), _as_coord(y)) for x, y in _coords(state, "probes")]
    model = WaveRNN(cell, sources, probes)
    model.eval()
    return model


def load_model(str_filename, which_iteration=-1, weights_only=True):
    """(model, history, history_geom_state, cfg) from a file written by
    ``save_model`` or by the reference (io.py:49-87).  The geometry comes from
    ``history_geom_state[which_iteration]``; ``cfg['dtype']``, when present,
    sets the default float type first, as in the reference."""
    print("Loading model from %s" % str_filename)
    rec = torch.load(str_filename, map_location="cpu", weights_only=weights_only)
    cfg = rec["cfg"]
    if isinstance(cfg, dict) and "dtype" in cfg:
        set_dtype(cfg["dtype"])
    model = model_from_state(rec["model_geom_class_str"],
                             rec["history_geom_state"][which_iteration], rec["model_state"],
                             rec.get("model_ext"))
    return model, rec["history"], rec["history_geom_state"], cfg
