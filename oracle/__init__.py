"""The oracle: CPU restatements of the reference's WaveRNN hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything here, and only as
the checker (or, for the baseline, as the CPU code being timed).  The
product package ``wavetorch_amd`` never imports this package.

Contents
--------
``wave_oracle.c`` (loaded through :mod:`oracle.c_oracle`)
    Plain-C forward loop (fp32 in the reference's exact operation order, and
    fp64) and the fp64 discrete adjoint (of its own fp64 forward, or taken at
    a given fp32 trajectory).
``ref_torch.py``
    Torch-CPU restatement of the reference modules, op for op, including the
    hand-written ``TimeStep.backward`` replayed by autograd.

Both are pinned in ``tests/test_oracle.py`` against the golden vectors that
``tests/golden/make_golden.py`` writes from the reference's own code, and
against closed forms (Dirichlet eigenmodes, energy, PML ramp, blur and
projection), central differences and the dot-product identity; its
docstring lists which check pins which function.
"""
