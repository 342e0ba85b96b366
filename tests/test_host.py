"""CPU tests of the host side: the C ABI library loads and exports every
declared symbol, host-only plan resolution (validation, CFL, tiling
heuristics, workspace sizes), and the reference's module API surface
(constructors, buffers, state_dict keys, errors).  No kernel is launched."""
from __future__ import annotations

import math
import os
import re

import numpy as np
import pytest
import torch

from wavetorch_amd import _lib
from wavetorch_amd._loop import LoopConfig, resolve

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names += re.findall(r"^int\s+(wt_\w+)\s*\(", txt, flags=re.M)
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/ but not exported"
    # every binding in the ctypes table is declared in the header
    assert set(_lib._SIGS) <= set(names)
    assert L.wt_abi_version() == _lib.ABI_VERSION


def test_last_error_is_nul_terminated():
    cfg = LoopConfig(Nx=2, Ny=2, dt=0.5, h=1.0, src_px=(), taps=())
    with pytest.raises(ValueError, match="invalid shape"):
        resolve(cfg, 1, 10, 1.0)
    import ctypes
    buf = ctypes.create_string_buffer(8)
    n = _lib.lib().wt_last_error(buf, 8)
    assert n > 8 and len(buf.value) == 7


def test_entry_points_reject_null_arguments():
    """Every compute and query entry point checks its arguments before any
    device call: null plans, buffers or descriptors give WT_EINVAL (-1) and a
    message naming the entry point (include/wavetorch_hip.h status codes)."""
    import ctypes
    L = _lib.lib()
    names = ["wt_laplacian", "wt_check_finite", "wt_check_finite_f64", "wt_coef_linear",
             "wt_step_fwd", "wt_step_bwd", "wt_plan_resolve",
             "wt_plan_create", "wt_plan_tuning", "wt_workspace_bytes", "wt_loop_fwd",
             "wt_loop_adj", "wt_geom_freeform_c", "wt_geom_freeform_c_bwd",
             "wt_geom_holey_c", "wt_geom_holey_c_bwd", "wt_geom_holey_workspace_bytes",
             "wt_loss_power_ce", "wt_loss_power_ce_workspace_bytes", "wt_batch_plan_create",
             "wt_batch_plan_tuning", "wt_batch_workspace_bytes",
             "wt_batch_fwd", "wt_batch_adj", "wt_batch_fwd_adj", "wt_graph_launch",
             "wt_autotune", "wt_loss_power_ce_static", "wt_workspace_bytes_f64",
             "wt_loop_fwd_f64", "wt_loop_adj_f64", "wt_laplacian_f64", "wt_step_fwd_f64",
             "wt_step_bwd_f64", "wt_geom_freeform_c_f64", "wt_geom_freeform_c_bwd_f64",
             "wt_geom_holey_c_f64", "wt_geom_holey_c_bwd_f64",
             "wt_geom_holey_workspace_bytes_f64"]
    for n in names:
        argtypes = _lib._SIGS[n][1]
        args = []
        for t in argtypes:
            if t in (ctypes.c_float, ctypes.c_double):
                args.append(1.0)
            elif t in (ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t):
                args.append(1)
            else:
                args.append(None)
        rc = getattr(L, n)(*args)
        assert rc == -1, (n, rc)
        buf = ctypes.create_string_buffer(256)
        L.wt_last_error(buf, 256)
        assert buf.value.decode().startswith(n), (n, buf.value)


def c2_cfg(N=150):
    return LoopConfig(Nx=N, Ny=N, dt=0.707, h=1.0, src_px=((40, 75, 0),),
                      taps=((110, 55, 1), (110, 75, 1), (110, 95, 1)))


def test_resolve_training_shape_heuristic():
    """Batch 64 on 150x150: 2x2 tiles per sample (256 workgroups), the
    deepest blocks that fit 1024 threads -- forward 32 steps at 12 rows per
    thread, adjoint 24 steps at 10 (what the autotuner also finds best)."""
    t, saved, temp = resolve(c2_cfg(), 64, 3000, 1.0)
    assert (t["plain"], t["kind"], t["strips"], t["tcols"]) == (0, 0, 2, 2)
    assert (t["tblock"], t["nt"], t["cpt"]) == (32, 1024, 12)
    assert (t["a_strips"], t["a_tcols"], t["a_tblock"], t["a_nt"], t["a_cpt"]) == \
        (2, 2, 24, 1024, 10)
    assert t["ck_every"] == 64
    F = 150 * 150
    # checkpoints (u_n, u_{n-1}) for ceil(T/K) segments and the last segment's
    # record (lap u of K steps, rows padded to Ny + 1), which the forward
    # writes, dominate the saved bytes
    ck = math.ceil(3000 / 64) * 2 * 64 * F * 4
    rec = 64 * 64 * 150 * 151 * 4
    assert saved >= ck + rec
    assert saved < ck + rec + 8 * 2 ** 20


def nl_cfg():
    return LoopConfig(Nx=150, Ny=150, dt=0.64, h=1.0, src_px=((40, 75, 0),),
                      taps=((110, 55, 1), (110, 75, 1), (110, 95, 1)),
                      sat_b0=0.5, sat_uth=math.sqrt(0.1), c_nl=0.05, c_clamp=(0.4, 1.1))


def test_resolve_rectangular_tilings():
    """The heuristic takes the fewest tiles per sample that give 256
    workgroups, near-square tiles first: the nonlinear batch-32 step gets 4x2
    tiles with 8-step forward blocks at 6 rows per thread and 5-step adjoint
    blocks (the deepest that fit 4 rows per thread); a 300x100 grid gets 8x2
    tiles of 37x50; batch 4 on 100x100 stays at 3x3, the most tiles whose own
    rows are twice the 16-step halo."""
    t, _, _ = resolve(nl_cfg(), 32, 3000, 1.0)
    assert (t["strips"], t["tcols"], t["tblock"], t["nt"], t["cpt"]) == (4, 2, 8, 1024, 6)
    assert (t["a_strips"], t["a_tcols"], t["a_tblock"], t["a_nt"], t["a_cpt"]) == \
        (4, 2, 5, 1024, 4)
    rect = LoopConfig(Nx=300, Ny=100, dt=0.707, h=1.0, src_px=((100, 50, 0),),
                      taps=((200, 50, 1),))
    t, _, _ = resolve(rect, 16, 1000, 1.0)
    assert (t["strips"], t["tcols"]) == (8, 2)
    sq = LoopConfig(Nx=100, Ny=100, dt=0.707, h=1.0, src_px=((20, 50, 0),),
                    taps=((80, 30, 1), (80, 50, 1), (80, 70, 1)))
    t, _, _ = resolve(sq, 4, 1000, 1.0)
    assert (t["strips"], t["tcols"], t["tblock"]) == (3, 3, 16)


def test_autotune_tiling_options():
    from wavetorch_amd import autotune
    opts = autotune.tilings(150, 150, 4, 2)
    assert (4, 2) in opts and (2, 4) in opts and (4, 4) in opts and (2, 2) in opts
    for s, c in opts:  # tiles at most 2:1, tile count within 4x of 8
        assert max((150 / s) / (150 / c), (150 / c) / (150 / s)) <= 2.0 + 1e-9
        assert 2 <= s * c <= 32
    st = {"tiling": (4, 2), "tblock": 8, "cpt": 6, "a_tiling": (2, 4), "a_tblock": 5,
          "a_cpt": 4}
    tun = autotune._to_tuning(st)
    assert (tun["strips"], tun["tcols"], tun["a_strips"], tun["a_tcols"]) == (4, 2, 2, 4)
    assert autotune._from_tuning(tun) == st


def test_tuner_checkpoint_interval_options():
    """With tune_k the single-plan search tries K0, its doublings up to T, and
    T (one segment, no recompute); without it the tuning carries no K."""
    from wavetorch_amd import autotune
    assert autotune.k_options(64, 1000) == (64, 128, 256, 512, 1000)
    assert autotune.k_options(64, 64) == (64,)
    st = {"tiling": (6, 5), "tblock": 16, "cpt": 4, "a_tiling": (5, 6), "a_tblock": 16,
          "a_cpt": 4}
    assert "ck_every" not in autotune._to_tuning(st)
    t = autotune._to_tuning(dict(st, ck_every=1000))
    assert t["ck_every"] == 1000 and autotune._from_tuning(t, with_k=True)["ck_every"] == 1000


def test_checkpoint_interval_override():
    """A checkpoint interval set on the model wins over a tuning's ck_every;
    unset, the tuning's applies, else 64.  Tunings whose search chose K are
    cached under K = "open"."""
    from wavetorch_amd import autotune
    m = _model()
    assert m.checkpoint_every == 64
    m.tuning = {"ck_every": 1000}
    assert m.checkpoint_every == 1000
    m.checkpoint_every = 32
    assert m.checkpoint_every == 32
    m.checkpoint_every = None
    assert m.checkpoint_every == 1000
    assert autotune.key(100, 100, 4, 1000, "open", "gfx950") == "gfx950:100x100:B4:T1000:Kopen"
    t = autotune.cached("propagate_100_b4")
    assert t is not None and t["ck_every"] == 1000


def test_batch_candidates_checkpoint_interval():
    """Batch tuning candidates carry K = 64 and the smallest K >= 64 that both
    block depths divide (segments then end on full launches)."""
    from wavetorch_amd import autotune
    assert autotune._aligned_k(24, 24) == 72 and autotune._aligned_k(16, 16) == 64
    assert autotune._aligned_k(32, 24) == 96 and autotune._aligned_k(24, 16) == 96
    for c in autotune.BATCH_CANDIDATES:
        k = c["ck_every"]
        assert k == 64 or (k % c["tblock"] == 0 and k % c["a_tblock"] == 0 and k <= 256)
    assert {(c["tblock"], c["a_tblock"], c["ck_every"]) for c in autotune.BATCH_CANDIDATES} \
        >= {(24, 24, 64), (24, 24, 72)}


def test_resolve_single_tile_runs_whole_segments():
    """One tile per sample has no halo: one launch per checkpoint segment."""
    cfg = LoopConfig(Nx=40, Ny=36, dt=0.707, h=1.0, src_px=((9, 18, 0),), taps=((30, 12, 1),))
    t, _, _ = resolve(cfg, 2, 120, 1.0, {"ck_every": 32})
    assert (t["strips"], t["tcols"], t["tblock"], t["a_tblock"]) == (1, 1, 32, 32)


def test_resolve_keeps_explicit_tuning():
    tun = {"strips": 6, "tcols": 1, "tblock": 8, "nt": 1024, "cpt": 8, "a_tblock": 8}
    t, _, _ = resolve(c2_cfg(), 8, 200, 1.0, tun)
    assert (t["strips"], t["tcols"], t["tblock"], t["nt"], t["cpt"]) == (6, 1, 8, 1024, 8)
    assert t["a_tblock"] == 8 and (t["a_nt"], t["a_cpt"]) == (1024, 8)
    t, _, _ = resolve(c2_cfg(), 8, 200, 1.0, {"plain": 1})
    assert t["plain"] == 1 and t["tblock"] == 1


def test_resolve_rejects_infeasible_tuning():
    # 1 strip of 150 rows at 4 rows per thread needs 38 x 150 threads
    with pytest.raises(RuntimeError, match="does not fit"):
        resolve(c2_cfg(), 8, 200, 1.0, {"strips": 1, "tcols": 1, "nt": 1024, "cpt": 4})


def test_resolve_cfl_and_indices():
    with pytest.raises(ValueError, match="CFL"):
        resolve(c2_cfg(), 4, 100, 1.5)  # dt 0.707 > 1/(1.5*sqrt 2)
    bad = LoopConfig(Nx=50, Ny=50, dt=0.5, h=1.0, src_px=((50, 3, 0),), taps=())
    with pytest.raises(ValueError):
        resolve(bad, 1, 10, 1.0)
    bad = LoopConfig(Nx=50, Ny=50, dt=0.5, h=1.0, src_px=((3, 3, 0),), taps=((3, -1, 1),))
    with pytest.raises(ValueError):
        resolve(bad, 1, 10, 1.0)


def test_checkpoint_memory_vs_K():
    """Saved + temporary bytes of the nonlinear 150x150 batch-32 loop for K in
    {16, 64, 256}: checkpoints scale as T/K, the segment buffer as K."""
    cfg = nl_cfg()
    F = 32 * 150 * 150 * 4
    tot = {}
    for K in (16, 64, 256):
        _, saved, temp = resolve(cfg, 32, 3000, 1.0, {"ck_every": K})
        assert saved >= math.ceil(3000 / K) * 2 * F
        tot[K] = saved + temp
    assert tot[64] < tot[16] and tot[64] < tot[256]
    assert tot[64] < 1.2e9  # vs ~8.6 GB for every field of the 3000 steps


# ---------------------------------------------------------------- module API

def _model():
    from wavetorch_amd import (WaveCell, WaveGeometryFreeForm, WaveIntensityProbe, WaveRNN,
                               WaveSource)
    geom = WaveGeometryFreeForm((60, 50), 1.0, c0=1.0, c1=0.5, abs_N=10, abs_sig=1.0,
                                abs_p=2.0, rho="rand")
    cell = WaveCell(0.5, geom)
    return WaveRNN(cell, [WaveSource(20, 25)],
                   [WaveIntensityProbe(40, 10), WaveIntensityProbe(40, 30)])


def test_state_dict_keys_match_reference_order():
    keys = list(_model().state_dict().keys())
    want = ["cell.dt", "cell.satdamp_b0", "cell.satdamp_uth", "cell.c_nl",
            "cell.geom.rho", "cell.geom.h", "cell.geom.c0", "cell.geom.c1", "cell.geom.abs_N",
            "cell.geom.abs_sig", "cell.geom.abs_p", "cell.geom._b", "cell.geom.eta",
            "cell.geom.beta", "cell.geom.design_region", "cell.geom.blur_kernel",
            "cell.geom.blur_N", "cell.geom.blur_radius",
            "sources.0.x", "sources.0.y", "probes.0.x", "probes.0.y", "probes.1.x",
            "probes.1.y"]
    assert sorted(keys) == sorted(want)
    sd = _model().state_dict()
    assert sd["cell.geom.abs_N"].dtype == torch.uint8
    assert sd["cell.geom.abs_p"].dtype == torch.uint8
    assert sd["cell.geom.design_region"].dtype == torch.uint8
    assert sd["cell.geom.blur_N"].dtype == torch.int32
    assert sd["sources.0.x"].dtype == torch.int64
    assert tuple(sd["cell.geom.blur_kernel"].shape) == (1, 1, 3, 3)
    # blur radius 1 selects the full 3x3 (a 1/9 box)
    assert torch.allclose(sd["cell.geom.blur_kernel"], torch.full((1, 1, 3, 3), 1 / 9))


def test_cell_cfl_raises_value_error():
    from wavetorch_amd import WaveCell, WaveGeometryFreeForm
    geom = WaveGeometryFreeForm((40, 40), 1.0, c0=1.0, c1=0.5, abs_N=5)
    WaveCell(0.70, geom)
    with pytest.raises(ValueError, match="CFL"):
        WaveCell(0.71, geom)


def test_parameters_are_geometry_parameters():
    m = _model()
    ps = list(m.parameters())
    assert len(ps) == 1 and ps[0] is m.cell.geom.rho


def test_rho_constrained_to_design_region_and_pml():
    m = _model()
    g = m.cell.geom
    assert torch.all(g.rho.detach()[g.b > 0] == 0)


def test_constrain_equals_mask_assignments():
    """geom.py:201-205 zeroes rho by two boolean-mask assignments; the select
    form gives the same values, NaN and infinities included."""
    import numpy as np
    from wavetorch_amd import WaveGeometryFreeForm
    dr = np.zeros((30, 40), dtype=np.uint8)
    dr[5:20, 8:30] = 1
    g = WaveGeometryFreeForm((30, 40), 1.0, c0=1.0, c1=0.5, abs_N=4, abs_sig=1.0, abs_p=2,
                             rho="rand", design_region=dr)
    r = torch.rand(30, 40) * 3 - 1
    r[0, 0] = float("nan")
    r[10, 10] = float("inf")
    r[12, 12] = float("nan")
    with torch.no_grad():
        g.rho.copy_(r)
    g.constrain_to_design_region()
    ref = r.clone()
    ref[torch.from_numpy(dr) == 0] = 0.0
    ref[g.b > 0] = 0.0
    assert torch.equal(torch.nan_to_num(g.rho.detach(), nan=7.0), torch.nan_to_num(ref, nan=7.0))
    assert torch.isnan(g.rho.detach()[12, 12]) and g.rho.detach()[0, 0] == 0


def test_save_load_model_round_trip(tmp_path):
    """io.py:13-87 file layout: the dict keys of the reference, the WaveRNN
    state_dict, and a model rebuilt from the last geometry state; with no
    pandas history the restricted loader reads it."""
    from wavetorch_amd.io import load_model, save_model
    m = _model()
    with torch.no_grad():
        m.cell.geom.rho.mul_(0.5)
    save_model(m, "m1", str(tmp_path) + "/", cfg={"dtype": "float32"}, verbose=False)
    raw = torch.load(str(tmp_path / "m1.pt"), weights_only=True)
    assert set(raw) == {"model_geom_class_str", "model_state", "history", "history_geom_state",
                        "cfg", "model_ext"}
    assert raw["model_geom_class_str"] == "WaveGeometryFreeForm"
    m2, hist, hgs, cfg = load_model(str(tmp_path / "m1.pt"))
    assert hist is None and cfg == {"dtype": "float32"} and len(hgs) == 1
    assert torch.equal(m2.cell.geom.rho.detach(), m.cell.geom.rho.detach())
    assert float(m2.cell.dt) == float(m.cell.dt)
    assert [(p.x.item(), p.y.item()) for p in m2.probes] == \
        [(p.x.item(), p.y.item()) for p in m.probes]
    assert [(q.x.item(), q.y.item()) for q in m2.sources] == \
        [(q.x.item(), q.y.item()) for q in m.sources]


def test_save_load_nonlinear_round_trip(tmp_path):
    """A nonlinear model (saturable damping, Kerr speed, clamp extension) and
    a line source come back as saved: the cell's nonlinear buffers, its
    loop configuration and the source's pixel list."""
    import math
    from wavetorch_amd import (WaveCell, WaveGeometryFreeForm, WaveIntensityProbe,
                               WaveLineSource, WaveRNN)
    from wavetorch_amd.io import load_model, save_model
    geom = WaveGeometryFreeForm((40, 40), 1.0, c0=1.0, c1=0.5, abs_N=8, abs_sig=1.0,
                                abs_p=2.0, rho="rand")
    cell = WaveCell(0.6, geom, satdamp_b0=0.5, satdamp_uth=math.sqrt(0.1), c_nl=0.05,
                    c_clamp=(0.4, 1.1))
    m = WaveRNN(cell, [WaveLineSource(12, 10, 12, 29)],
                [WaveIntensityProbe(30, 15), WaveIntensityProbe(30, 25)])
    save_model(m, "nl", str(tmp_path) + "/", cfg={"dtype": "float32"}, verbose=False)
    m2, _, _, _ = load_model(str(tmp_path / "nl.pt"))
    for k in ("dt", "satdamp_b0", "satdamp_uth", "c_nl"):
        assert torch.equal(getattr(m2.cell, k), getattr(m.cell, k)), k
    assert m2.cell.c_clamp == (0.4, 1.1)
    c1, c2 = m.loop_config(), m2.loop_config()
    assert c2.nonlinear and c2 == c1
    assert torch.equal(m2.sources[0].x, m.sources[0].x)
    assert torch.equal(m2.sources[0].y, m.sources[0].y)
    assert torch.equal(m2.cell.geom.rho.detach(), m.cell.geom.rho.detach())


def test_load_reference_style_file(tmp_path):
    """A file in the reference's layout (five entries, no model_ext, a
    linear cell) loads: coordinates by index order, linear cell."""
    from wavetorch_amd.io import load_model
    m = _model()
    rec = {"model_geom_class_str": "WaveGeometryFreeForm", "model_state": m.state_dict(),
           "history": None,
           "history_geom_state": [m.cell.geom.state_reconstruction_args()],
           "cfg": {"dtype": "float32"}}
    torch.save(rec, str(tmp_path / "ref.pt"))
    m2, _, _, _ = load_model(str(tmp_path / "ref.pt"))
    assert not m2.loop_config().nonlinear and m2.cell.c_clamp is None
    assert [(p.x.item(), p.y.item()) for p in m2.probes] == [(40, 10), (40, 30)]
    assert [(q.x.item(), q.y.item()) for q in m2.sources] == [(20, 25)]


def test_product_path_has_no_cpu_fallback():
    m = _model()
    with pytest.raises(RuntimeError, match="HIP device only"):
        m(torch.zeros(1, 10))
    from wavetorch_amd import TimeStep
    z = torch.zeros(1, 8, 8)
    with pytest.raises(RuntimeError, match="HIP device only"):
        TimeStep.apply(torch.zeros(8, 8), torch.ones(8, 8), z, z, torch.tensor(0.5),
                       torch.tensor(1.0))


def test_line_source_axis_aligned_and_diagonal():
    from wavetorch_amd import WaveLineSource
    s = WaveLineSource(5, 3, 5, 9)
    assert s.pixels() == [(5, c) for c in range(3, 10)]
    s = WaveLineSource(10, 4, 2, 4)
    assert s.pixels() == [(r, 4) for r in range(10, 1, -1)]
    s = WaveLineSource(0, 0, 4, 4)
    assert s.pixels() == [(i, i) for i in range(5)]
    # a shallow line: one pixel per column, endpoints included, rows monotone
    s = WaveLineSource(0, 0, 3, 10)
    px = s.pixels()
    assert len(px) == 11 and px[0] == (0, 0) and px[-1] == (3, 10)
    rows = [p[0] for p in px]
    assert rows == sorted(rows) and all(b - a <= 1 for a, b in zip(rows, rows[1:]))


def test_probe_taps_order_matches_stack():
    """torch.stack(dim=-1) of multi-pixel probes: column = l * P + p."""
    from wavetorch_amd import (WaveCell, WaveGeometryFreeForm, WaveProbe, WaveIntensityProbe,
                               WaveRNN, WaveSource)
    geom = WaveGeometryFreeForm((40, 40), 1.0, c0=1.0, c1=0.5, abs_N=5)
    m = WaveRNN(WaveCell(0.5, geom), WaveSource(10, 10),
                [WaveProbe([30, 31], [5, 6]), WaveIntensityProbe([30, 31], [20, 21])])
    cfg = m.loop_config()
    assert cfg.taps == ((30, 5, 0), (30, 20, 1), (31, 6, 0), (31, 21, 1))
    assert cfg.src_px == ((10, 10, 0),)


def test_pml_profile_closed_form():
    from wavetorch_amd import WaveGeometryFreeForm
    N, aN, sig, p = 50, 8, 1.3, 2
    g = WaveGeometryFreeForm((N, N), 1.0, c0=1.0, c1=0.5, abs_N=aN, abs_sig=sig, abs_p=p)
    b = g.b.numpy().astype(np.float64)
    ramp = sig * (np.linspace(0, 1, aN + 1) ** p)[::-1]
    bx = np.zeros(N)
    bx[:aN + 1] = ramp
    bx[N - aN - 1:] = ramp[::-1]
    want = np.sqrt(bx[:, None] ** 2 + bx[None, :] ** 2)
    np.testing.assert_allclose(b, want, rtol=1e-6, atol=1e-7)
    assert b[aN, aN] == 0 and abs(b[0, 0] - math.sqrt(2) * sig) < 1e-6


def test_holey_oracle_closed_forms():
    """The oracle's restatement of geom.py:109-135, rho = P(sum_i exp(-|p - p_i| / r_i))
    with the tanh projection P (P(0) = 0, P(1) = 1): c = c1 at a lone hole's
    centre, c -> c0 far from every hole; two coincident holes add before P;
    autograd through it gives NaN x/y gradients for a centre on a grid point,
    as the reference does, and for off-grid centres agrees in fp64 with
    central differences of the same function."""
    import numpy as np
    from oracle import ref_torch as rt

    def t(v, dt=torch.float32):
        return torch.tensor(v, dtype=dt, requires_grad=True)
    c = rt.holey_c(t([12.0]), t([9.0]), t([1.5]), (40, 30), 0.5, 100.0, 1.0, 0.5).detach()
    assert abs(c[12, 9].item() - 0.5) < 1e-6
    assert abs(c[38, 28].item() - 1.0) < 1e-6
    d = np.exp(-2.0 / 1.5)  # exp(-d/r) at distance 2 from the centre, projected
    P = (np.tanh(50.0) + np.tanh(100.0 * (d - 0.5))) / (np.tanh(50.0) + np.tanh(50.0))
    rho = rt.holey_rho(t([12.0]), t([9.0]), t([1.5]), (40, 30), 0.5, 100.0)
    assert abs(rho[14, 9].item() - P) < 1e-5
    hx, hy, hr = t([12.0, 12.0]), t([9.0, 9.0]), t([3.0, 3.0])
    s2 = 2 * np.exp(-1.0 / 3.0)
    P2 = (np.tanh(2.0) + np.tanh(4.0 * (s2 - 0.5))) / (2 * np.tanh(2.0))
    assert abs(rt.holey_rho(hx, hy, hr, (40, 30), 0.5, 4.0)[13, 9].item() - P2) < 1e-5
    rt.holey_c(hx, hy, hr, (40, 30), 0.5, 4.0, 1.0, 0.5).sum().backward()
    assert torch.isnan(hx.grad).all() and torch.isnan(hy.grad).all()
    assert torch.isfinite(hr.grad).all()
    # off-grid centres, fp64: autograd = central differences
    p64 = [t([12.3, 20.6], torch.float64), t([9.4, 17.1], torch.float64),
           t([3.0, 2.0], torch.float64)]
    w = torch.from_numpy(np.random.default_rng(2).standard_normal((40, 30)))

    def J(a, b, r):
        return (rt.holey_c(a, b, r, (40, 30), 0.5, 4.0, 1.0, 0.5) * w).sum()
    J(*p64).backward()
    eps = 1e-6
    for k in range(3):
        for i in range(2):
            args_p = [q.detach().clone() for q in p64]
            args_m = [q.detach().clone() for q in p64]
            args_p[k][i] += eps
            args_m[k][i] -= eps
            fd = (J(*args_p) - J(*args_m)).item() / (2 * eps)
            assert abs(p64[k].grad[i].item() - fd) <= 1e-6 * max(1.0, abs(fd)), (k, i)


def test_batch_candidates_one_segment_within_budget():
    """One-segment candidates (K >= the longest length, a multiple of both
    depths) are offered only when their record fits the budget."""
    from wavetorch_amd import autotune
    from wavetorch_amd.batch import Instance
    cfg = LoopConfig(Nx=40, Ny=36, dt=0.707, h=1.0, src_px=((9, 18, 0),), taps=((30, 12, 1),))
    insts = [Instance(cfg, torch.zeros(40, 36), T, 1.0) for T in (150, 260, 97)]
    assert autotune.batch_record_bytes(insts, 10) == 3 * 4 * 10 * 40 * 37
    base = autotune.batch_candidates(insts, 0)
    assert base == autotune.BATCH_CANDIDATES
    big = autotune.batch_candidates(insts, float("inf"))
    extra = big[len(base):]
    assert extra and all(c["ck_every"] >= 260 and c["ck_every"] % c["tblock"] == 0
                         and c["ck_every"] % c["a_tblock"] == 0 for c in extra)
    k = autotune._aligned_k(24, 24, kmin=260)
    assert k == 264
    mid = autotune.batch_candidates(insts, autotune.batch_record_bytes(insts, 264))
    assert any(c["ck_every"] == 264 for c in mid[len(base):])
    assert all(autotune.batch_record_bytes(insts, c["ck_every"])
               <= autotune.batch_record_bytes(insts, 264) for c in mid[len(base):])
