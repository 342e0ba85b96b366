"""Vowel data path (reference: wavetorch/data/__init__.py:1)."""
from .vowels import load_all_vowels, load_vowel, normalize_vowel, select_vowel_sample

__all__ = ["load_all_vowels", "load_vowel", "normalize_vowel", "select_vowel_sample"]
