"""Training loop (reference: wavetorch/train.py:13-133) and a training step
captured as one HIP graph.

``train`` keeps the reference's signature and per-epoch behaviour: epoch 0
only characterises the starting structure, later epochs take one optimizer
step per batch through a closure and re-apply the design-region constraint;
per epoch it records losses, accuracies and confusion matrices in a pandas
history with the reference's columns and, given a ``name``, saves the model
(``io.save_model``).  The criterion is evaluated on
``normalize_power(model(xb).sum(dim=1))`` as in train.py:61-62; for the
default ``torch.nn.CrossEntropyLoss`` the fused loss head
(``loss.power_cross_entropy``) computes the same value in three short kernel
launches.

``GraphedTrainStep`` captures forward, loss head, checkpointed adjoint,
optimizer step and design-region constraint of one batch shape in a single
HIP graph (torch.cuda.CUDAGraph) and replays it per batch: every operation
on that path launches on the current stream, reads nothing back to the host
and records no memset.
"""
from __future__ import annotations

import copy
import time

import numpy as np
import torch

from .loss import is_default_cross_entropy, power_cross_entropy
from .utils import accuracy_onehot, normalize_power


def criterion_on_probes(criterion, y, labels):
    """criterion(normalize_power(y.sum(dim=1)), labels) (train.py:61-62): the
    fused loss head when `criterion` is the default CrossEntropyLoss and y is a
    (B, T, P) series on the device, else the torch composition."""
    if is_default_cross_entropy(criterion) and y.dim() == 3 and y.is_cuda:
        return power_cross_entropy(y, labels)
    return criterion(normalize_power(y.sum(dim=1)), labels)


def train(model, optimizer, criterion, train_dl, test_dl, N_epochs: int, batch_size: int,
          history=None, history_model_state=None, fold=None, name=None, savedir=None, cfg=None,
          accuracy=None):
    """Trains `model` for `N_epochs` epochs (train.py:13-133); returns
    (history, history_model_state).

    As in the reference: epoch 0 takes no optimizer step; batches come from
    `train_dl` / `test_dl` as (xb, yb) with one-hot yb; the train accuracy uses
    `accuracy` when given and the test accuracy uses `accuracy_onehot` when
    `accuracy` is given (train.py:103-104); `batch_size` is accepted and not
    used.  Differences: `history_model_state` defaults to a new list (the
    reference's default list is shared between calls), rows are added with
    pandas.concat (DataFrame.append is gone from pandas 2), and losses and
    predictions are moved to the host before numpy reductions, so device
    tensors work."""
    import pandas as pd
    from sklearn.metrics import confusion_matrix

    if history_model_state is None:
        history_model_state = []
    columns = ['time', 'epoch', 'fold', 'loss_train', 'loss_test', 'acc_train', 'acc_test',
               'cm_train', 'cm_test']
    if history is None:
        history = pd.DataFrame(columns=columns)

    t_start = time.time()
    for epoch in range(0, N_epochs + 1):
        t_epoch = time.time()
        loss_iter = []
        for xb, yb in train_dl:
            def closure():
                optimizer.zero_grad()
                loss = criterion_on_probes(criterion, model(xb), yb.argmax(dim=1))
                loss.backward()
                return loss

            if epoch == 0:  # characterise the starting structure only
                with torch.no_grad():
                    loss = criterion_on_probes(criterion, model(xb), yb.argmax(dim=1))
            else:
                loss = optimizer.step(closure)
                model.cell.geom.constrain_to_design_region()
            loss_iter.append(float(loss.item()))

        with torch.no_grad():
            acc_train_tmp = []
            preds, truths = [], []
            for xb, yb in train_dl:
                yb_pred = normalize_power(model(xb).sum(dim=1))
                preds.append(yb_pred)
                truths.append(yb)
                if accuracy is not None:
                    acc_train_tmp.append(accuracy(yb_pred, yb.argmax(dim=1)))
            y_pred = torch.cat(preds, dim=0).cpu()
            y_truth = torch.cat(truths, dim=0).cpu()
            cm_train = confusion_matrix(y_truth.argmax(dim=1).numpy(), y_pred.argmax(dim=1).numpy())

            acc_test_tmp, loss_test_tmp = [], []
            cm_test = None
            if test_dl is not None:
                preds, truths = [], []
                for xb, yb in test_dl:
                    yb_pred = normalize_power(model(xb).sum(dim=1))
                    preds.append(yb_pred)
                    truths.append(yb)
                    loss_test_tmp.append(float(criterion(yb_pred, yb.argmax(dim=1)).item()))
                    if accuracy is not None:
                        acc_test_tmp.append(accuracy_onehot(yb_pred, yb.argmax(dim=1)))
                y_pred = torch.cat(preds, dim=0).cpu()
                y_truth = torch.cat(truths, dim=0).cpu()
                cm_test = confusion_matrix(y_truth.argmax(dim=1).numpy(),
                                           y_pred.argmax(dim=1).numpy())

        def _mean(v):
            return float(np.mean(v)) if len(v) else float("nan")

        print('Epoch %2d/%2d --- Elapsed Time:  %4.2f min | Training Loss:  %.4e | Testing '
              'Loss:  %.4e | Training Accuracy:  %.4f | Testing Accuracy:  %.4f' %
              (epoch, N_epochs, (time.time() - t_epoch) / 60, _mean(loss_iter),
               _mean(loss_test_tmp), _mean(acc_train_tmp), _mean(acc_test_tmp)))

        row = pd.DataFrame([{'time': pd.to_datetime('now'), 'epoch': epoch, 'fold': fold,
                             'loss_train': _mean(loss_iter), 'loss_test': _mean(loss_test_tmp),
                             'acc_train': _mean(acc_train_tmp), 'acc_test': _mean(acc_test_tmp),
                             'cm_train': cm_train, 'cm_test': cm_test}], columns=columns)
        history = row if len(history) == 0 else pd.concat([history, row], ignore_index=True)
        history_model_state.append(copy.deepcopy(model.cell.geom.state_reconstruction_args()))

        if name is not None:
            from .io import save_model
            save_model(model, name, savedir, history, history_model_state, cfg, verbose=False)

    print('Total Time: %.2f min\n' % ((time.time() - t_start) / 60))
    return history, history_model_state


class GraphedTrainStep:
    """One training step of `model` on batches shaped like `x` -- forward,
    fused loss head (the default CrossEntropyLoss of train.py:61-62), the
    checkpointed adjoint, `optimizer.step()` and
    ``geom.constrain_to_design_region()`` (train.py:71-72) -- captured as
    one HIP graph.  ``step(x, labels)`` copies the batch into the graph's
    input buffers, replays the graph on the current stream and returns the
    loss as a device tensor (overwritten by the next replay).

    `optimizer` must be torch.optim.Adam or AdamW built with
    ``capturable=True`` (its step count then lives on the device).  Capture
    needs `warmup` eager steps first (plans, lazily created optimizer state);
    the parameters and optimizer state they changed are restored afterwards,
    so building the step leaves the model as it was.  One object per batch
    shape; its graph holds the step's device memory until it is deleted."""

    def __init__(self, model, optimizer, x, labels, warmup: int = 2, constrain: bool = True):
        if not isinstance(optimizer, (torch.optim.Adam, torch.optim.AdamW)) or \
                not all(g.get("capturable", False) for g in optimizer.param_groups):
            raise ValueError("GraphedTrainStep: the optimizer must be torch.optim.Adam or AdamW "
                             "with capturable=True")
        if not x.is_cuda:
            raise RuntimeError("GraphedTrainStep: x must be on the HIP device")
        self.model, self.opt, self.constrain = model, optimizer, constrain
        self.x = x.detach().clone()
        self.labels = labels.detach().to(device=x.device, dtype=torch.int64).clone()
        params = [p for g in optimizer.param_groups for p in g["params"]]
        p0 = [p.detach().clone() for p in params]
        st0 = {p: {k: (v.detach().clone() if torch.is_tensor(v) else copy.deepcopy(v))
                   for k, v in optimizer.state[p].items()}
               for p in params if p in optimizer.state}

        side = torch.cuda.Stream(device=x.device)
        side.wait_stream(torch.cuda.current_stream(x.device))
        with torch.cuda.stream(side):
            for _ in range(max(1, warmup)):
                optimizer.zero_grad(set_to_none=True)
                self._body()
        torch.cuda.current_stream(x.device).wait_stream(side)

        self.graph = torch.cuda.CUDAGraph()
        optimizer.zero_grad(set_to_none=True)
        with torch.cuda.graph(self.graph):
            self.loss = self._body()

        with torch.no_grad():  # undo the warm-up steps
            for p, v in zip(params, p0):
                p.copy_(v)
                if p.grad is not None:  # the graph's gradient buffer: defined
                    p.grad.zero_()      # before the first replay
            for p in params:
                old = st0.get(p, {})
                for k, v in optimizer.state.get(p, {}).items():
                    if torch.is_tensor(v):
                        if k in old:
                            v.copy_(old[k])
                        else:  # created by the warm-up: Adam's initial state is zero
                            v.zero_()

    def _body(self):
        loss = power_cross_entropy(self.model(self.x), self.labels)
        loss.backward()
        self.opt.step()
        if self.constrain:
            self.model.cell.geom.constrain_to_design_region()
        return loss

    def step(self, x, labels):
        if tuple(x.shape) != tuple(self.x.shape) or tuple(labels.shape) != tuple(self.labels.shape):
            raise ValueError(f"GraphedTrainStep: captured for x {tuple(self.x.shape)}, labels "
                             f"{tuple(self.labels.shape)}; got {tuple(x.shape)}, "
                             f"{tuple(labels.shape)}")
        self.x.copy_(x)
        self.labels.copy_(labels)
        self.graph.replay()
        return self.loss

    __call__ = step
