"""Vowel recordings for the training workloads, without librosa (the
reference's data path is wavetorch/data/vowels.py:12-113).

Reading: a PCM WAV file through ``scipy.io.wavfile``.  Integer samples are
scaled to [-1, 1) as float32 (16-bit: ``s / 32768``, the value librosa's
soundfile reader returns), channels are averaged to mono (librosa's
``mono=True``), and another rate ``sr`` is reached with
``scipy.signal.resample_poly`` (polyphase FIR, Kaiser window, ratio reduced
by the gcd: 16 kHz -> 10 kHz is up 5 / down 8).  librosa uses soxr's
high-quality filter instead, so resampled waveforms differ from the
reference's at the level of the two filters' passband ripple and transition
band (both keep the band below about 0.45 of the new rate); at the file's own
rate (``sr=None``) the samples are the reference's exactly.

The set: per gender a list of recordings (waveform, one-hot label, path),
collected class by class in sorted file order (files ``w*<class><ext>`` for
women, ``m*<class><ext>`` for men); with ``max_samples`` each gender is cut
to a stratified subset drawn by scikit-learn's ``train_test_split`` with the
reference's arguments (so a given ``random_state`` picks the reference's
files), and the genders are concatenated men first.
"""
from __future__ import annotations

import glob
import math
import os
import random
from math import gcd

import numpy as np
import torch

_PREFIX = {"women": "w", "men": "m"}


def normalize_vowel(wav_data):
    """The waveform scaled to unit total power (vowels.py:12-17), with the
    power summed in the waveform's own dtype."""
    w = np.asarray(wav_data)
    return w / np.sqrt(np.sum(w * w))


_INT_SCALE = {np.dtype(np.int16): 32768.0, np.dtype(np.int32): 2147483648.0}


def _pcm_to_float(data):
    if data.dtype in _INT_SCALE:
        out = (data.astype(np.float64) / _INT_SCALE[data.dtype]).astype(np.float32)
    elif data.dtype == np.uint8:
        out = (data.astype(np.float32) - 128.0) / 128.0
    else:
        out = data.astype(np.float32)
    return out.mean(axis=1, dtype=np.float32) if out.ndim == 2 else out


def load_vowel(file, sr=None, normalize=True):
    """One recording as float32 at rate ``sr`` (the file's own when None),
    optionally at unit power (vowels.py:20-29)."""
    from scipy.io import wavfile
    rate, raw = wavfile.read(file)
    wave = _pcm_to_float(raw)
    if sr is not None and int(sr) != int(rate):
        from scipy.signal import resample_poly
        g = gcd(int(sr), int(rate))
        wave = resample_poly(wave, int(sr) // g, int(rate) // g).astype(np.float32)
    return normalize_vowel(wave) if normalize else wave


def _recordings(root, prefix, classes, ext, sr, normalize):
    """[(waveform, one-hot label, path)] of one gender, class by class."""
    eye = np.eye(len(classes))
    out = []
    for k, name in enumerate(classes):
        for path in sorted(glob.glob(os.path.join(root, prefix + "*" + name + ext))):
            out.append((load_vowel(path, sr=sr, normalize=normalize), eye[k], path))
    return out


def _stratified_subset(recs, n, n_classes, random_state):
    """The training part of a stratified split of ``recs`` with ``n``
    recordings kept and ``n_classes`` held out (the split of vowels.py:75-81,
    made on indices: the permutation depends only on the count, the labels and
    the random state)."""
    from sklearn.model_selection import train_test_split
    labels = [r[1] for r in recs]
    keep, _ = train_test_split(np.arange(len(recs)), train_size=n, test_size=n_classes,
                               stratify=labels, shuffle=True, random_state=random_state)
    return [recs[i] for i in keep]


def load_all_vowels(str_classes, gender='both', sr=None, normalize=True, dir='data/vowels/',
                    ext='.wav', max_samples=None, random_state=None):
    """Every recording of the classes ``str_classes`` for ``gender``
    ('women', 'men' or 'both') as lists X (float tensors), Y (one-hot float
    tensors) and F (paths) (vowels.py:32-99).  ``max_samples`` limits the
    total (split floor/ceil between men and women for 'both')."""
    assert gender in ['women', 'men', 'both'], "gender must be either 'women', 'men', or 'both'"
    want = ("men", "women") if gender == "both" else (gender,)
    quota = {"men": None, "women": None}
    if max_samples is not None:
        if gender == "both":
            quota = {"men": math.floor(max_samples / 2), "women": math.ceil(max_samples / 2)}
        else:
            quota = {"men": max_samples, "women": max_samples}
    recs = []
    for g in want:
        part = _recordings(dir, _PREFIX[g], str_classes, ext, sr, normalize)
        if quota[g] is not None:
            part = _stratified_subset(part, quota[g], len(str_classes), random_state)
        recs += part
    dt = torch.get_default_dtype()
    X = [torch.tensor(w, dtype=dt) for w, _, _ in recs]
    Y = [torch.tensor(y, dtype=dt) for _, y, _ in recs]
    F = [p for _, _, p in recs]
    print("Dataset: %d vowel samples" % len(X))
    return X, Y, F


def select_vowel_sample(X, Y, F, y_class, ind=None):
    """Recording number ``ind`` (drawn with ``random.random()`` when None)
    among those of class ``y_class``, as a (1, T) input and a (1, C) label
    (vowels.py:102-113)."""
    members = [i for i, y in enumerate(Y) if int(torch.argmax(y)) == y_class]
    if ind is None:
        ind = int(random.random() * len(members))
    pick = members[ind]
    print('Selected sample %d, corresponding to file %s' % (ind, F[pick]))
    return X[pick][None], Y[pick][None]
