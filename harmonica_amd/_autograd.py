"""
Differentiable g_z of a prism layer (torch.autograd) and a Gauss-Newton
inversion for densities and top heights built on it.

The reference's forward models are not differentiable; this module adds
derivatives of DatasetAccessorPrismLayer.gravity(coordinates, "g_z")
(src/harmonica/_forward/prisms/layer.py:314-433) with respect to the layer's
densities and top heights, computed by hand-written kernels
(hmx_prism_layer_gz_vjp / _jvp): the Jacobian-transpose product is
recomputed on the fly and never stored; the top-height derivative is the
closed-form top-face term (sum over the top vertices of the signed
arctangent atan(x y / (z r))).
"""

import ctypes

import numpy as np
import torch

from ._forward import layer_fields_device, to_device
from ._lib import check, lib, ptr, stream_handle
from .constants import GRAVITATIONAL_CONST

# d g_z[mGal] / d (G-free vertex sums): g_z = -1e5 * g_u (gravity.py:228-232)
_GZ_FACTOR = -1e5 * GRAVITATIONAL_CONST


class LayerGeometry:
    """Fixed part of a differentiable prism layer: prism centres (1d),
    bottoms (n_north, n_east), observation points and thickness threshold,
    all resident on the GPU."""

    def __init__(self, easting, northing, bottom, coordinates, thickness_threshold=None):
        self.easting = to_device(np.asarray(easting, dtype=float))
        self.northing = to_device(np.asarray(northing, dtype=float))
        self.bottom = to_device(np.asarray(bottom, dtype=float))
        cast = np.broadcast(*coordinates[:3])
        self.shape = cast.shape
        self.coords = tuple(
            to_device(np.broadcast_to(np.asarray(c, dtype=float), cast.shape).ravel())
            for c in coordinates[:3]
        )
        self.threshold = 0.0 if thickness_threshold is None else float(thickness_threshold)
        self.n_obs = self.coords[0].numel()
        self.n_prisms = self.easting.numel() * self.northing.numel()
        nbytes = ctypes.c_size_t(0)
        check(lib().hmx_gz_grad_workspace_bytes(self.n_obs, self.n_prisms, ctypes.byref(nbytes)))
        self._ws = torch.empty(max(nbytes.value, 8), dtype=torch.uint8, device=self.easting.device)
        self._ws_bytes = nbytes.value

    def forward(self, density, top):
        """g_z (mGal) of the layer, a (n_obs,) tensor."""
        return layer_fields_device(*self.coords, self.easting, self.northing, self.bottom,
                                   top.contiguous(), density.contiguous(), ("g_z",),
                                   thickness_threshold=self.threshold)[0]

    def jvp(self, density, top, d_density, d_top):
        """Directional derivative of g_z (mGal) along (d_density, d_top)."""
        wa = (d_density * _GZ_FACTOR).contiguous()
        wb = (density * d_top * _GZ_FACTOR).contiguous()
        out = torch.empty(self.n_obs, dtype=torch.float64, device=self.easting.device)
        check(lib().hmx_prism_layer_gz_jvp(
            *(ptr(c) for c in self.coords), self.n_obs, ptr(self.easting), self.easting.numel(),
            ptr(self.northing), self.northing.numel(), ptr(self.bottom), ptr(top.contiguous()),
            ptr(density.contiguous()), self.threshold, ptr(wa), ptr(wb), ptr(out), ptr(self._ws),
            self._ws_bytes, stream_handle(None)))
        return out

    def vjp(self, density, top, v):
        """(d/d density, d/d top) of sum_i v_i g_z(i) (mGal units)."""
        shape = (self.northing.numel(), self.easting.numel())
        gs = torch.empty(shape, dtype=torch.float64, device=self.easting.device)
        gf = torch.empty(shape, dtype=torch.float64, device=self.easting.device)
        check(lib().hmx_prism_layer_gz_vjp(
            *(ptr(c) for c in self.coords), ptr(v.contiguous()), self.n_obs, ptr(self.easting),
            self.easting.numel(), ptr(self.northing), self.northing.numel(), ptr(self.bottom),
            ptr(top.contiguous()), ptr(density.contiguous()), self.threshold, ptr(gs), ptr(gf),
            ptr(self._ws), self._ws_bytes, stream_handle(None)))
        # The kernel returns zero sums for the cells the layer loop skips
        # (NaN density, NaN top or bottom, thin cells); the density weight of
        # the top derivative is applied only where it is finite, so a NaN
        # density gives a zero derivative instead of 0 * NaN.
        weight = torch.where(torch.isfinite(density), density, torch.zeros_like(density))
        return gs * _GZ_FACTOR, gf * weight * _GZ_FACTOR


class _LayerGz(torch.autograd.Function):
    @staticmethod
    def forward(ctx, density, top, geometry):
        ctx.geometry = geometry
        ctx.save_for_backward(density, top)
        return geometry.forward(density, top)

    @staticmethod
    def backward(ctx, grad_out):
        density, top = ctx.saved_tensors
        g_density, g_top = ctx.geometry.vjp(density, top, grad_out.contiguous())
        return g_density, g_top, None


def layer_gz(density, top, geometry):
    """Differentiable g_z (mGal) of a prism layer with respect to its
    densities and tops ((n_north, n_east) float64 CUDA tensors)."""
    return _LayerGz.apply(density, top, geometry)


def gauss_newton(geometry, data, density0, top0, n_iterations=20, damping=1e-2, cg_iterations=10,
                 density_scale=100.0, top_scale=1000.0, density_ref=None, top_ref=None):
    """Damped Gauss-Newton inversion of g_z data (mGal) for the densities and
    top heights of a prism layer.

    Minimises |g(rho, T) - d|^2 + damping (|(rho - rho_ref)/s_rho|^2 +
    |(T - T_ref)/s_T|^2) over rho and T, each step solving the normal
    equations of the linearised problem by conjugate gradients with
    matrix-free products (hmx_prism_layer_gz_jvp and _vjp).  Returns
    (density, top, misfit history) with misfit = RMS of the residual.
    """
    rho = density0.clone()
    top = top0.clone()
    rho_ref = rho.clone() if density_ref is None else density_ref
    top_ref = top.clone() if top_ref is None else top_ref
    # Cells with a NaN density or top are holes the forward model skips:
    # they stay fixed (their update is zero), and their regularisation
    # terms are left out so that NaN does not reach the step.
    free = torch.isfinite(rho) & torch.isfinite(top)
    zero = torch.zeros_like(rho)
    history = []
    for _ in range(n_iterations):
        r = geometry.forward(rho, top) - data
        history.append(float(torch.sqrt(torch.mean(r * r))))
        # gradient of the objective in scaled variables (x = rho/s_rho, y = T/s_T)
        g_rho, g_top = geometry.vjp(rho, top, r)
        b_rho = torch.where(free, -(g_rho * density_scale + damping * (rho - rho_ref) / density_scale), zero)
        b_top = torch.where(free, -(g_top * top_scale + damping * (top - top_ref) / top_scale), zero)

        def normal(p_rho, p_top):
            jv = geometry.jvp(rho, top, p_rho * density_scale, p_top * top_scale)
            a_rho, a_top = geometry.vjp(rho, top, jv)
            return (torch.where(free, a_rho * density_scale + damping * p_rho, zero),
                    torch.where(free, a_top * top_scale + damping * p_top, zero))

        x_rho = torch.zeros_like(rho)
        x_top = torch.zeros_like(top)
        res_rho, res_top = b_rho.clone(), b_top.clone()
        p_rho, p_top = res_rho.clone(), res_top.clone()
        rr = float((res_rho * res_rho).sum() + (res_top * res_top).sum())
        for _ in range(cg_iterations):
            q_rho, q_top = normal(p_rho, p_top)
            pq = float((p_rho * q_rho).sum() + (p_top * q_top).sum())
            if pq <= 0:
                break
            alpha = rr / pq
            x_rho += alpha * p_rho
            x_top += alpha * p_top
            res_rho -= alpha * q_rho
            res_top -= alpha * q_top
            rr_new = float((res_rho * res_rho).sum() + (res_top * res_top).sum())
            p_rho = res_rho + (rr_new / rr) * p_rho
            p_top = res_top + (rr_new / rr) * p_top
            rr = rr_new
        rho = torch.where(free, rho + x_rho * density_scale, rho)
        top = torch.where(free, top + x_top * top_scale, top)
    r = geometry.forward(rho, top) - data
    history.append(float(torch.sqrt(torch.mean(r * r))))
    return rho, top, history


__all__ = ["LayerGeometry", "gauss_newton", "layer_gz"]
