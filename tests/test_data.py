"""CPU tests of the vowel data path (wavetorch_amd.data, reference:
wavetorch/data/vowels.py:12-113) on WAV files written here.

* native rate: 16-bit samples come back as s / 32768 exactly (librosa's value
  for 16-bit PCM), stereo is averaged, normalisation gives unit power;
* resampling 16 kHz -> 10 kHz (scipy polyphase, in place of librosa's soxr):
  a 440 Hz tone matches the analytic tone at the new sample times to 1e-3 away
  from the ends; the exact match to librosa's filter is not pinned (librosa is
  absent), see the module docstring;
* load_all_vowels: the reference's order (men then women, sorted per class),
  one-hot labels, gender selection, stratified max_samples subsets;
  select_vowel_sample shapes.
"""
import os

import numpy as np
import pytest
import torch
from scipy.io import wavfile

from wavetorch_amd.data import load_all_vowels, load_vowel, normalize_vowel, select_vowel_sample


def test_native_rate_samples_and_normalisation(tmp_path):
    rng = np.random.default_rng(0)
    s = rng.integers(-32768, 32767, 5000, dtype=np.int16)
    f = str(tmp_path / "a.wav")
    wavfile.write(f, 16000, s)
    x = load_vowel(f, normalize=False)
    assert x.dtype == np.float32 and np.array_equal(x, s.astype(np.float32) / 32768.0)
    xn = load_vowel(f)
    assert abs(float(np.square(xn.astype(np.float64)).sum()) - 1.0) < 1e-5
    np.testing.assert_allclose(xn, normalize_vowel(x), rtol=0, atol=0)
    st = np.stack([s, (s // 2).astype(np.int16)], axis=1)
    f2 = str(tmp_path / "b.wav")
    wavfile.write(f2, 16000, st)
    x2 = load_vowel(f2, normalize=False)
    want = (s.astype(np.float32) / 32768.0 + (s // 2).astype(np.float32) / 32768.0) / 2
    np.testing.assert_allclose(x2, want, rtol=0, atol=1e-7)


def test_resample_16k_to_10k_tone(tmp_path):
    n = np.arange(32000)
    tone = 0.5 * np.sin(2 * np.pi * 440.0 * n / 16000.0)
    f = str(tmp_path / "t.wav")
    wavfile.write(f, 16000, np.round(tone * 32767).astype(np.int16))
    y = load_vowel(f, sr=10000, normalize=False)
    assert len(y) == 20000
    m = np.arange(len(y))
    want = 0.5 * np.sin(2 * np.pi * 440.0 * m / 10000.0)
    assert np.max(np.abs(y[200:-200] - want[200:-200])) < 1e-3


def _write_set(d, names, rate=16000, n=800):
    rng = np.random.default_rng(1)
    for nm in names:
        wavfile.write(os.path.join(d, nm + ".wav"), rate,
                      rng.integers(-3000, 3000, n, dtype=np.int16))


def test_load_all_vowels_order_labels_and_subsets(tmp_path):
    d = str(tmp_path)
    names = ["w01ae", "w02ae", "w03ae", "m01ae", "m02ae", "m03ae",
             "w01ei", "w02ei", "w03ei", "m01ei", "m02ei", "m04ei",
             "w01iy", "w05iy", "w06iy", "m07iy", "m08iy", "m09iy"]
    _write_set(d, names)
    X, Y, F = load_all_vowels(["ae", "ei", "iy"], dir=d + "/")
    base = [os.path.basename(f)[:-4] for f in F]
    assert base == ["m01ae", "m02ae", "m03ae", "m01ei", "m02ei", "m04ei", "m07iy", "m08iy",
                    "m09iy", "w01ae", "w02ae", "w03ae", "w01ei", "w02ei", "w03ei", "w01iy",
                    "w05iy", "w06iy"]
    assert all(torch.equal(y, torch.eye(3)[["ae", "ei", "iy"].index(b[-2:])])
               for y, b in zip(Y, base))
    assert all(abs(float((x.double() ** 2).sum()) - 1.0) < 1e-5 for x in X)
    Xw, Yw, Fw = load_all_vowels(["ae", "ei", "iy"], gender="women", dir=d + "/")
    assert [os.path.basename(f)[0] for f in Fw] == ["w"] * 9
    Xs, Ys, Fs = load_all_vowels(["ae", "ei", "iy"], dir=d + "/", max_samples=6,
                                 random_state=0)
    assert len(Xs) == 6 and sum(os.path.basename(f)[0] == "m" for f in Fs) == 3
    labs = sorted(int(y.argmax()) for y in Ys[:3])
    assert labs == [0, 1, 2]  # stratified: one man per class
    x1, y1 = select_vowel_sample(X, Y, F, 1, ind=2)
    assert x1.shape == (1, 800) and y1.shape == (1, 3) and int(y1.argmax()) == 1


def test_stratified_subset_picks_the_reference_files(tmp_path):
    """max_samples subsets: the files are those the reference's split of the
    per-gender lists themselves (vowels.py:75-81: waveforms, labels and paths
    passed together, stratified on the labels) keeps, in the same order, for
    several random states."""
    from sklearn.model_selection import train_test_split
    d = str(tmp_path)
    names = [f"{g}{k:02d}{c}" for g in "wm" for c in ("ae", "ei", "iy") for k in range(1, 6)]
    _write_set(d, names)
    classes = ["ae", "ei", "iy"]
    X, Y, F = load_all_vowels(classes, dir=d + "/")
    xm, ym, fm = X[:15], [y.numpy() for y in Y[:15]], F[:15]
    xw, yw, fw = X[15:], [y.numpy() for y in Y[15:]], F[15:]
    for rs in (0, 1, 7, 123):
        _, _, Fs = load_all_vowels(classes, dir=d + "/", max_samples=7, random_state=rs)
        _, _, _, _, fm_k, _ = train_test_split(xm, ym, fm, train_size=3, test_size=3,
                                               stratify=ym, shuffle=True, random_state=rs)
        _, _, _, _, fw_k, _ = train_test_split(xw, yw, fw, train_size=4, test_size=3,
                                               stratify=yw, shuffle=True, random_state=rs)
        assert Fs == fm_k + fw_k, rs


REF_VOWELS = "/root/reference/data/vowels"


@pytest.mark.skipif(not os.path.isdir(REF_VOWELS), reason="reference vowel files not present")
def test_reference_recordings_native_rate():
    """The reference's own recordings (16 kHz, 16-bit) load at the native rate
    as s / 32768, normalised to unit power; at 10 kHz the length is the
    16 -> 10 kHz ratio of the original."""
    f = os.path.join(REF_VOWELS, "m09er.wav")
    rate, s = wavfile.read(f)
    x = load_vowel(f, normalize=False)
    assert rate == 16000 and np.array_equal(x, s.astype(np.float32) / 32768.0)
    y = load_vowel(f, sr=10000)
    assert len(y) == -(-len(s) * 5 // 8)
    assert abs(float(np.square(y.astype(np.float64)).sum()) - 1.0) < 1e-5
