"""Vowel recordings for the training workloads (reference:
wavetorch/data/vowels.py:12-113), without librosa.

``load_vowel`` reads a PCM WAV file with ``scipy.io.wavfile``: 16-bit samples
become float32 ``s / 32768`` (what librosa's soundfile reader returns for
16-bit PCM), several channels are averaged to mono (librosa's ``mono=True``),
and a different target rate ``sr`` is reached with ``scipy.signal.resample_poly``
(polyphase FIR, Kaiser window beta 5, rate ratio reduced by the gcd; 16 kHz ->
10 kHz is up 5 / down 8).  librosa resamples with soxr's high-quality filter
instead, so resampled waveforms differ from the reference's at the level of
the two filters' passband ripple and transition band (both keep the band
below about 0.45 of the new rate); at the native rate (``sr=None``) the
samples are the reference's exactly.  ``load_all_vowels`` and
``select_vowel_sample`` keep the reference's file order, one-hot labels,
gender selection and stratified subsetting.
"""
from __future__ import annotations

import glob
import math
import os
import random
from math import gcd

import numpy as np
import torch


def normalize_vowel(wav_data):
    """vowels.py:12-17: scale to unit total power."""
    total_power = np.square(wav_data).sum()
    return wav_data / np.sqrt(total_power)


def _read_wav(file):
    from scipy.io import wavfile
    rate, data = wavfile.read(file)
    if data.dtype == np.int16:
        x = data.astype(np.float32) / 32768.0
    elif data.dtype == np.int32:
        x = (data.astype(np.float64) / 2147483648.0).astype(np.float32)
    elif data.dtype == np.uint8:
        x = (data.astype(np.float32) - 128.0) / 128.0
    else:
        x = data.astype(np.float32)
    if x.ndim == 2:
        x = x.mean(axis=1, dtype=np.float32)
    return x, int(rate)


def load_vowel(file, sr=None, normalize=True):
    """vowels.py:20-29: one recording as float32 at rate `sr` (the file's own
    rate when None), optionally normalised to unit power."""
    data, rate = _read_wav(file)
    if sr is not None and int(sr) != rate:
        from scipy.signal import resample_poly
        g = gcd(int(sr), rate)
        data = resample_poly(data, int(sr) // g, rate // g).astype(np.float32)
    if normalize:
        data = normalize_vowel(data)
    return data


def load_all_vowels(str_classes, gender='both', sr=None, normalize=True, dir='data/vowels/',
                    ext='.wav', max_samples=None, random_state=None):
    """vowels.py:32-99: every recording of the classes `str_classes` (file
    names w*<class><ext> for women, m*<class><ext> for men), as lists X of
    float tensors, Y of one-hot labels and F of file names; men first, then
    women, each in sorted file order per class.  `max_samples` limits the
    total with a stratified subset per gender (half each for 'both')."""
    assert gender in ['women', 'men', 'both'], "gender must be either 'women', 'men', or 'both'"
    x_w, y_w, F_w, x_m, y_m, F_m = [], [], [], [], [], []
    for i, str_class in enumerate(str_classes):
        y = np.eye(len(str_classes))[i]
        for file in sorted(glob.glob(os.path.join(dir, 'w*' + str_class + ext))):
            x_w.append(load_vowel(file, sr=sr, normalize=normalize))
            F_w.append(file)
            y_w.append(y)
        for file in sorted(glob.glob(os.path.join(dir, 'm*' + str_class + ext))):
            x_m.append(load_vowel(file, sr=sr, normalize=normalize))
            F_m.append(file)
            y_m.append(y)

    if max_samples is not None:
        from sklearn.model_selection import train_test_split
        if gender == 'both':
            num_samples_men = math.floor(max_samples / 2)
            num_samples_women = math.ceil(max_samples / 2)
        else:
            num_samples_men = num_samples_women = max_samples
        # a stratified subset: the training split of train_test_split, the
        # test split dropped (vowels.py:75-81)
        x_m, _, y_m, _, F_m, _ = train_test_split(
            x_m, y_m, F_m, train_size=num_samples_men, test_size=len(str_classes),
            stratify=y_m, shuffle=True, random_state=random_state)
        x_w, _, y_w, _, F_w, _ = train_test_split(
            x_w, y_w, F_w, train_size=num_samples_women, test_size=len(str_classes),
            stratify=y_w, shuffle=True, random_state=random_state)

    dt = torch.get_default_dtype()
    if gender == 'both':
        xs, ys, F = x_m + x_w, y_m + y_w, F_m + F_w
    elif gender == 'women':
        xs, ys, F = x_w, y_w, F_w
    else:
        xs, ys, F = x_m, y_m, F_m
    X = [torch.tensor(x, dtype=dt) for x in xs]
    Y = [torch.tensor(y, dtype=dt) for y in ys]
    print("Dataset: %d vowel samples" % len(X))
    return X, Y, F


def select_vowel_sample(X, Y, F, y_class, ind=None):
    """vowels.py:102-113: one sample of class `y_class` (random when `ind` is
    None), as (1, T) input and (1, C) label."""
    labels_ints = [y.argmax().item() for y in Y]
    inds_this_clss = [i for i in range(len(labels_ints)) if labels_ints[i] == y_class]
    if ind is None:
        ind = int(random.random() * len(inds_this_clss))
    print('Selected sample %d, corresponding to file %s' % (ind, F[inds_this_clss[ind]]))
    return X[inds_this_clss[ind]].unsqueeze(0), Y[inds_this_clss[ind]].unsqueeze(0)
