"""Model save / load in the reference's file format (reference: wavetorch/io.py:13-87).

A saved file is ``torch.save`` of the dict the reference writes:
``model_geom_class_str``, ``model_state`` (the WaveRNN ``state_dict``, whose
keys this package keeps), ``history``, ``history_geom_state`` and ``cfg``.
``load_model`` rebuilds the geometry from one entry of the geometry history,
then the cell, sources and probes as the reference does (io.py:49-87).

Loading unpickles the file.  By default only torch's restricted loader runs
(``weights_only=True``): it takes tensors, containers and plain values, which
covers files saved without a pandas history; a file holding a pandas history
(what ``train`` saves) needs ``weights_only=False``, which runs arbitrary code
from the file -- pass it only for files you trust.
"""
from __future__ import annotations

import copy
import os

import torch

from . import geom
from .cell import WaveCell
from .probe import WaveIntensityProbe
from .rnn import WaveRNN
from .source import WaveSource
from .utils import set_dtype


def save_model(model, name, savedir='./study/', history=None, history_geom_state=None,
               cfg=None, verbose=True):
    """Writes `savedir` + `name` + '.pt' (io.py:13-40)."""
    str_filename = name + '.pt'
    if not os.path.exists(savedir):
        os.makedirs(savedir)
    str_savepath = savedir + str_filename
    if history_geom_state is None:
        history_geom_state = [model.cell.geom.state_reconstruction_args()]
    data = {'model_geom_class_str': model.cell.geom.__class__.__name__,
            'model_state': {k: v.detach().cpu() for k, v in model.state_dict().items()},
            'history': history,
            'history_geom_state': history_geom_state,
            'cfg': cfg}
    if verbose:
        print("Saving model to %s" % str_savepath)
    torch.save(data, str_savepath)


def new_geometry(class_str, state):
    """The geometry class named `class_str` built from its reconstruction
    arguments (io.py:43-46)."""
    WaveGeometryClass = getattr(geom, class_str)
    geom_state = copy.deepcopy(state)
    return WaveGeometryClass(**geom_state)


def load_model(str_filename, which_iteration=-1, weights_only=True):
    """(model, history, history_geom_state, cfg) from a saved file
    (io.py:49-87): sources become WaveSource and probes WaveIntensityProbe
    at the saved coordinates, as in the reference.  The default dtype is set
    from cfg['dtype'] when the file has one (the reference requires it)."""
    print("Loading model from %s" % str_filename)
    data = torch.load(str_filename, map_location="cpu", weights_only=weights_only)
    cfg = data['cfg']
    if isinstance(cfg, dict) and 'dtype' in cfg:
        set_dtype(cfg['dtype'])
    new_geom = new_geometry(data['model_geom_class_str'], data['history_geom_state'][which_iteration])
    model_state = copy.deepcopy(data['model_state'])
    px = [model_state[k].item() for k in model_state if 'probes' in k and 'x' in k]
    py = [model_state[k].item() for k in model_state if 'probes' in k and 'y' in k]
    sx = [model_state[k].item() for k in model_state if 'sources' in k and 'x' in k]
    sy = [model_state[k].item() for k in model_state if 'sources' in k and 'y' in k]
    new_probes = [WaveIntensityProbe(x, y) for x, y in zip(px, py)]
    new_sources = [WaveSource(x, y) for x, y in zip(sx, sy)]
    new_cell = WaveCell(model_state['cell.dt'].item(), new_geom)
    new_model = WaveRNN(new_cell, new_sources, new_probes)
    new_model.eval()
    return new_model, data['history'], data['history_geom_state'], cfg
