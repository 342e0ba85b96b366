"""The oracle: CPU restatements of the reference's WaveRNN hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything here, and only as
the checker (or, for the baseline, as the CPU code being timed).  The
product package ``wavetorch_amd`` never imports this package.

Contents
--------
``wave_oracle.c`` (loaded through :mod:`oracle.c_oracle`)
    Plain-C forward loop (fp32 in the reference's exact operation order, and
    fp64) and the fp64 discrete adjoint.
``ref_torch.py``
    Torch-CPU restatement of the reference modules, op for op, including the
    hand-written ``TimeStep.backward`` replayed by autograd.
``np_oracle.py``
    Small numpy helpers with closed forms (Dirichlet eigenmodes, energy)
    used to pin the two restatements above.

Which check pins which function is listed in each test module's docstring.
"""
