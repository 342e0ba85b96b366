"""Material parameterisation (reference: wavetorch/geom.py).

Same classes, constructor signatures, buffers and ``state_dict`` keys as the
reference.  The PML damping field ``b`` is built once at construction
(geom.py:63-85).  ``WaveGeometryFreeForm.c`` -- blur, tanh projection and
the linear map to wave speed (geom.py:207-233) -- runs as gfx950 kernels
(``wt_geom_freeform_c``) with a hand-written backward, so autograd carries
dJ/dc from the time loop back to ``rho``; ``WaveGeometryHoley.c`` -- the sum
of exponential holes and the same projection (geom.py:109-135) -- likewise
(``wt_geom_holey_c``), back to the hole centres and radii.
"""
from __future__ import annotations

from copy import deepcopy
from typing import Tuple

import numpy as np
import torch

from . import _lib
from .utils import to_tensor


class WaveGeometry(torch.nn.Module):
    """Base geometry: domain shape, spacing h, speeds c0/c1 and the PML (geom.py:11-85)."""

    def __init__(self, domain_shape: Tuple, h: float, c0: float, c1: float, abs_N: int = 20,
                 abs_sig: float = 11, abs_p: float = 4.0):
        super().__init__()
        assert len(domain_shape) == 2, \
            "len(domain_shape) must be equal to 2: only two-dimensional (2D) domains are supported"
        self.domain_shape = tuple(domain_shape)
        self.register_buffer("h", to_tensor(h))
        self.register_buffer("c0", to_tensor(c0))
        self.register_buffer("c1", to_tensor(c1))
        self.register_buffer("abs_N", to_tensor(abs_N, dtype=torch.uint8))
        self.register_buffer("abs_sig", to_tensor(abs_sig))
        self.register_buffer("abs_p", to_tensor(abs_p, dtype=torch.uint8))
        self._init_b(abs_N, abs_sig, abs_p)

    def state_reconstruction_args(self):
        return {"domain_shape": self.domain_shape,
                "h": self.h.item(),
                "c0": self.c0.item(),
                "c1": self.c1.item(),
                "abs_N": self.abs_N.item(),
                "abs_sig": self.abs_sig.item(),
                "abs_p": self.abs_p.item()}

    def __repr__(self):
        return "WaveGeometry shape={}, h={}".format(self.domain_shape, self.h)

    def forward(self):
        raise NotImplementedError(
            "WaveGeometry forward() is not implemented. Although WaveGeometry is a subclass of a "
            "torch.nn.Module, its forward() method should never be called. It only exists as a "
            "torch.nn.Module to hook into pytorch as a component of a WaveCell.")

    @property
    def c(self):
        raise NotImplementedError

    @property
    def b(self):
        return self._b

    @property
    def cmax(self):
        """Maximum wave speed for the CFL check (geom.py:55-58)."""
        return np.max([self.c0.item(), self.c1.item()])

    def constrain_to_design_region(self):
        pass

    def _init_b(self, abs_N: int, abs_sig: float, abs_p: float):
        """Polynomial damping ramp of width abs_N on all four sides (geom.py:63-85):
        b = sqrt(b_x^2 + b_y^2) with b_vals = abs_sig * linspace(0,1,abs_N+1)**abs_p."""
        Nx, Ny = self.domain_shape
        assert Nx > 2 * abs_N + 1, \
            "The domain isn't large enough in the x-direction to fit absorbing layer. " \
            "Nx = {} and N = {}".format(Nx, abs_N)
        assert Ny > 2 * abs_N + 1, \
            "The domain isn't large enough in the y-direction to fit absorbing layer. " \
            "Ny = {} and N = {}".format(Ny, abs_N)
        b_vals = abs_sig * torch.linspace(0.0, 1.0, abs_N + 1) ** abs_p
        b_x = torch.zeros(Nx, Ny)
        b_y = torch.zeros(Nx, Ny)
        if abs_N > 0:
            b_x[0:abs_N + 1, :] = torch.flip(b_vals, [0]).repeat(Ny, 1).transpose(0, 1)
            b_x[(Nx - abs_N - 1):Nx, :] = b_vals.repeat(Ny, 1).transpose(0, 1)
            b_y[:, 0:abs_N + 1] = torch.flip(b_vals, [0]).repeat(Nx, 1)
            b_y[:, (Ny - abs_N - 1):Ny] = b_vals.repeat(Nx, 1)
        self.register_buffer("_b", torch.sqrt(b_x ** 2 + b_y ** 2))


class WaveGeometryHoley(WaveGeometry):
    """Density from a sum of exponential 'holes' (geom.py:88-133)."""

    def __init__(self, domain_shape: Tuple, h: float, c0: float, c1: float, abs_N: int = 20,
                 abs_sig: float = 11, abs_p: float = 4.0, eta: float = 0.5, beta: float = 100.0,
                 x=None, y=None, r=None):
        super().__init__(domain_shape, h, c0, c1, abs_N, abs_sig, abs_p)
        self.x = torch.nn.Parameter(to_tensor(x))
        self.y = torch.nn.Parameter(to_tensor(y))
        self.r = torch.nn.Parameter(to_tensor(r))
        self.register_buffer("eta", to_tensor(eta))
        self.register_buffer("beta", to_tensor(beta))

    def state_reconstruction_args(self):
        my_args = {"eta": self.eta.item(),
                   "beta": self.beta.item(),
                   "x": deepcopy(self.x.detach()),
                   "y": deepcopy(self.y.detach()),
                   "r": deepcopy(self.r.detach())}
        return {**super().state_reconstruction_args(), **my_args}

    def _host_params(self):
        """(eta, beta, c0, c1) as host numbers, re-read from the device buffers
        only when one of them was replaced or modified (data_ptr, _version):
        reading them on every call would synchronise the host with the device
        once per training step."""
        ts = (self.eta, self.beta, self.c0, self.c1)
        key = tuple((t.data_ptr(), t._version) for t in ts)
        cached = getattr(self, "_host_cache", None)
        if cached is None or cached[0] != key:
            cached = (key, tuple(float(t.item()) for t in ts))
            self._host_cache = cached
        return cached[1]

    def _holes(self, c0, c1):
        _require_hip(self.x, "WaveGeometryHoley")
        eta, beta, _, _ = self._host_params()
        Nx, Ny = self.domain_shape
        return _HoleyC.apply(self.x, self.y, self.r, Nx, Ny, eta, beta, c0, c1)

    def _rho(self):
        """Projected density P(sum_h exp(-d_h / r_h)) (geom.py:109-128), gfx950
        kernels (the same kernels as ``c`` with c0 = 0, c1 = 1)."""
        return self._holes(0.0, 1.0)

    @property
    def rho(self):
        return self._rho()

    @property
    def c(self):
        """c0 + (c1 - c0) * rho (geom.py:133-135), gfx950 kernels."""
        _, _, c0, c1 = self._host_params()
        return self._holes(c0, c1)


class _HoleyC(torch.autograd.Function):
    """c = c0 + (c1-c0) * P(sum_h exp(-d_h / r_h)) on the GPU (geom.py:109-135),
    with the hand-written backward for the hole centres and radii."""

    @staticmethod
    def forward(ctx, hx, hy, hr, Nx, Ny, eta, beta, c0, c1):
        shapes = (hx.shape, hy.shape, hr.shape)
        hx, hy, hr = (t.reshape(-1).contiguous() for t in (hx, hy, hr))
        H = hx.numel()
        if hy.numel() != H or hr.numel() != H:
            raise ValueError("WaveGeometryHoley: x, y and r must have one entry per hole")
        for t in (hy, hr):
            _require_hip(t, "WaveGeometryHoley")
            if t.dtype != hx.dtype:
                raise TypeError("WaveGeometryHoley: x, y and r must share one dtype")
        c = torch.empty((Nx, Ny), dtype=hx.dtype, device=hx.device)
        raw = torch.empty_like(c)
        name = "wt_geom_holey_c_f64" if hx.dtype == torch.float64 else "wt_geom_holey_c"
        _lib.check(getattr(_lib.lib(), name)(
            hx.data_ptr(), hy.data_ptr(), hr.data_ptr(), H, Nx, Ny, eta, beta, c0, c1,
            c.data_ptr(), raw.data_ptr(), _lib.stream_ptr(hx.device)), name)
        ctx.save_for_backward(raw, hx, hy, hr)
        ctx.params = (Nx, Ny, eta, beta, c0, c1)
        ctx.shapes = shapes
        return c

    @staticmethod
    def backward(ctx, grad_c):
        import ctypes
        raw, hx, hy, hr = ctx.saved_tensors
        Nx, Ny, eta, beta, c0, c1 = ctx.params
        H = hx.numel()
        need = ctx.needs_input_grad[:3]
        out = [torch.empty_like(t) if n else None for t, n in zip((hx, hy, hr), need)]
        L = _lib.lib()
        suffix = "_f64" if raw.dtype == torch.float64 else ""
        nbytes = ctypes.c_int64()
        _lib.check(getattr(L, "wt_geom_holey_workspace_bytes" + suffix)(
            H, Nx, Ny, ctypes.byref(nbytes)), "wt_geom_holey_workspace_bytes" + suffix)
        work = torch.empty((nbytes.value + 7) // 8, dtype=torch.float64, device=raw.device)
        _lib.check(getattr(L, "wt_geom_holey_c_bwd" + suffix)(
            grad_c.to(raw.dtype).contiguous().data_ptr(), raw.data_ptr(), hx.data_ptr(),
            hy.data_ptr(), hr.data_ptr(), H, Nx, Ny, eta, beta, c0, c1,
            *(_lib.ptr(o) for o in out), work.data_ptr(), _lib.stream_ptr(raw.device)),
            "wt_geom_holey_c_bwd" + suffix)
        grads = [o.reshape(s) if o is not None else None for o, s in zip(out, ctx.shapes)]
        return (*grads, None, None, None, None, None, None)


def _disk_pixels(radius: int) -> np.ndarray:
    """Pixels of a disk of radius ``radius+1`` centred at (radius, radius) inside
    the (2r+1)^2 box, by the strict rule (dr/R)^2 + (dc/R)^2 < 1 with R = radius+1
    -- the rule of the scikit-image ``circle`` call in geom.py:149."""
    R = radius + 1
    m = np.zeros((2 * radius + 1, 2 * radius + 1), dtype=bool)
    for i in range(2 * radius + 1):
        for j in range(2 * radius + 1):
            if ((i - radius) / R) ** 2 + ((j - radius) / R) ** 2 < 1.0:
                m[i, j] = True
    return m


class _FreeFormC(torch.autograd.Function):
    """c = c0 + (c1-c0) * P(Blur^N(rho)) on the GPU (geom.py:207-233)."""

    @staticmethod
    def forward(ctx, rho, kernel, radius, blur_N, eta, beta, c0, c1):
        rho = rho.contiguous()
        kernel = kernel.contiguous()
        Nx, Ny = rho.shape
        if kernel.dtype != rho.dtype:
            raise TypeError("WaveGeometryFreeForm: blur kernel and rho must share one dtype")
        c = torch.empty_like(rho)
        blur = torch.empty_like(rho)
        work = torch.empty(2 * Nx * Ny, dtype=rho.dtype, device=rho.device)
        s = _lib.stream_ptr(rho.device)
        name = "wt_geom_freeform_c_f64" if rho.dtype == torch.float64 else "wt_geom_freeform_c"
        _lib.check(getattr(_lib.lib(), name)(
            rho.data_ptr(), kernel.data_ptr(), radius, blur_N, Nx, Ny, eta, beta, c0, c1,
            c.data_ptr(), None, blur.data_ptr(), work.data_ptr(), s), name)
        ctx.save_for_backward(blur, kernel)
        ctx.params = (radius, blur_N, eta, beta, c0, c1)
        return c

    @staticmethod
    def backward(ctx, grad_c):
        blur, kernel = ctx.saved_tensors
        radius, blur_N, eta, beta, c0, c1 = ctx.params
        grad_c = grad_c.to(blur.dtype).contiguous()
        Nx, Ny = blur.shape
        g = torch.empty_like(blur)
        work = torch.empty(2 * Nx * Ny, dtype=blur.dtype, device=blur.device)
        s = _lib.stream_ptr(blur.device)
        name = ("wt_geom_freeform_c_bwd_f64" if blur.dtype == torch.float64
                else "wt_geom_freeform_c_bwd")
        _lib.check(getattr(_lib.lib(), name)(
            grad_c.data_ptr(), blur.data_ptr(), kernel.data_ptr(), radius, blur_N, Nx, Ny, eta,
            beta, c0, c1, g.data_ptr(), work.data_ptr(), s), name)
        return g, None, None, None, None, None, None, None


def _require_hip(t, what):
    if not t.is_cuda:
        raise RuntimeError(
            f"wavetorch_amd: {what} runs on the HIP device only; move the model with .to('cuda')")
    if t.dtype not in (torch.float32, torch.float64):
        raise NotImplementedError(f"wavetorch_amd: {what} computes in float32 or float64 "
                                  f"(got {t.dtype})")


class WaveGeometryFreeForm(WaveGeometry):
    """Free-form density with blur and tanh projection (geom.py:136-233)."""

    def __init__(self, domain_shape: Tuple, h: float, c0: float, c1: float, abs_N: int = 20,
                 abs_sig: float = 11, abs_p: float = 4.0, eta: float = 0.5, beta: float = 100.0,
                 design_region=None, rho='half', blur_radius: int = 1, blur_N: int = 1):
        super().__init__(domain_shape, h, c0, c1, abs_N, abs_sig, abs_p)
        self.register_buffer("eta", to_tensor(eta))
        self.register_buffer("beta", to_tensor(beta))
        self._init_design_region(design_region, domain_shape)
        self._init_rho(rho, domain_shape)
        m = _disk_pixels(blur_radius)
        blur_kernel = torch.zeros((2 * blur_radius + 1, 2 * blur_radius + 1),
                                  dtype=torch.get_default_dtype())
        blur_kernel[torch.from_numpy(m)] = 1
        blur_kernel = blur_kernel / blur_kernel.sum()
        self.register_buffer("blur_kernel", blur_kernel.unsqueeze(0).unsqueeze(0))
        self.register_buffer("blur_N", to_tensor(blur_N, dtype=torch.int))
        # the reference stores blur_N under this name (geom.py:156); kept for
        # state_dict compatibility
        self.register_buffer("blur_radius", to_tensor(blur_N, dtype=torch.int))
        self._blur_radius = int(blur_radius)
        self.constrain_to_design_region()

    def state_reconstruction_args(self):
        my_args = {"eta": self.eta.item(),
                   "beta": self.beta.item(),
                   "design_region": deepcopy(self.design_region),
                   "rho": deepcopy(self.rho.detach()),
                   "blur_radius": self.blur_radius.item(),
                   "blur_N": self.blur_N.item()}
        return {**super().state_reconstruction_args(), **my_args}

    def __repr__(self):
        return super().__repr__() + ", " + str(self.design_region.sum().item()) + " DOFs"

    def _init_design_region(self, design_region, domain_shape):
        if design_region is not None:
            assert tuple(design_region.shape) == tuple(domain_shape), \
                "The design region shape must match domain shape; design_region.shape = {} " \
                "domain_shape = {}".format(design_region.shape, domain_shape)
            if type(design_region) is np.ndarray:
                design_region = torch.from_numpy(design_region).to(torch.uint8)
        else:
            design_region = torch.ones(domain_shape, dtype=torch.uint8)
        self.register_buffer("design_region", design_region)

    def _init_rho(self, rho, domain_shape):
        if isinstance(rho, torch.Tensor) | isinstance(rho, np.ndarray):
            assert tuple(rho.shape) == tuple(domain_shape)
            self.rho = torch.nn.Parameter(to_tensor(rho))
        elif isinstance(rho, str):
            if rho == 'rand':
                self.rho = torch.nn.Parameter(torch.round(torch.rand(domain_shape)))
            elif rho == 'half':
                self.rho = torch.nn.Parameter(torch.ones(domain_shape) * 0.5)
            elif rho == 'blank':
                self.rho = torch.nn.Parameter(torch.zeros(domain_shape))
            else:
                raise ValueError('The domain initialization defined by `rho = %s` is invalid' % rho)
        else:
            raise ValueError('The domain initialization is invalid')

    def constrain_to_design_region(self):
        """rho = 0 outside the design region and wherever b > 0 (geom.py:201-205).
        One elementwise select in place of the reference's two boolean-mask
        assignments (same values; a mask assignment on the device reads the
        mask's count back to the host, which also bars graph capture)."""
        with torch.no_grad():
            keep = (self.design_region != 0) & ~(self.b > 0)
            self.rho.copy_(torch.where(keep, self.rho, 0.0))

    def _rho_model(self):
        """Blurred and projected density P(Blur^N(rho)) (geom.py:224-229); the
        kernels fuse it into ``c``, so it is recovered from c here."""
        c0, c1 = self.c0.item(), self.c1.item()
        return (self.c - c0) / (c1 - c0)

    def _host_params(self):
        """(blur_N, eta, beta, c0, c1) as host numbers, re-read from the
        device buffers only when one of them was replaced or modified
        (data_ptr, _version): reading them every step would synchronise the
        host with the device once per training step."""
        ts = (self.blur_N, self.eta, self.beta, self.c0, self.c1)
        key = tuple((t.data_ptr(), t._version) for t in ts)
        cached = getattr(self, "_host_cache", None)
        if cached is None or cached[0] != key:
            vals = (int(self.blur_N.item()), float(self.eta.item()), float(self.beta.item()),
                    float(self.c0.item()), float(self.c1.item()))
            cached = (key, vals)
            self._host_cache = cached
        return cached[1]

    @property
    def c(self):
        """c0 + (c1-c0) * P(Blur^N(rho)) (geom.py:231-233), gfx950 kernels in
        the geometry's dtype (float32, or float64 after set_dtype)."""
        _require_hip(self.rho, "WaveGeometryFreeForm.c")
        k = self.blur_kernel[0, 0].to(device=self.rho.device, dtype=self.rho.dtype)
        blur_N, eta, beta, c0, c1 = self._host_params()
        return _FreeFormC.apply(self.rho, k, self._blur_radius, blur_N, eta, beta, c0, c1)
