"""Dense, differentiable torch restatement of the forward path (TEST INFRASTRUCTURE).

Only the test suite may import this module: it is the checker of the backward
pass, never the product.  It renders small problems (about a thousand
Gaussians, a few 64x64 cameras) by evaluating every (pixel, Gaussian) pair at
once in any floating dtype, so torch autograd in float64 gives reference
gradients and central finite differences in float64 can pin them.

The forward it restates is the reference screen-space path
(project_visible_kernel ``renderer_cuda.cu:614-1055``, bin_tiles_kernel
``1691-1829``, composite_tiles_kernel ``1973-2305``), the same semantics the C
oracle (``oracle/vgr_oracle.c``) implements, plus SH bands 1..3 on the view
direction (3DGS basis) and the covariance of ``precompute_covariances``
(``renderer_cuda.cu:492-540``).

Discrete decisions (visibility, tile rectangles, depth order, which splats a
pixel blends before its transmittance runs out, alpha and colour clamps) are
computed once at the evaluation point and can be frozen (``structure``) so
finite differences see the smooth function those decisions select.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch

ALPHA_THRESHOLD = 1.0 / 255.0
MAX_ALPHA = 0.99
TRANSMITTANCE_THRESHOLD = 1.0e-4
MIN_ANTIALIAS_COMPENSATION = 0.005

SH_C1 = 0.4886025119029199
SH_C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396)
SH_C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
         -0.4570457994644658, 1.445305721320277, -0.5900435899266435)


def covariances(scales: torch.Tensor, quats: torch.Tensor) -> torch.Tensor:
    """Upper triangle (xx, xy, xz, yy, yz, zz) of R(q/|q|) diag(s^2) R^T (renderer_cuda.cu:492-540)."""
    q = quats / torch.sqrt(torch.clamp((quats * quats).sum(-1, keepdim=True), min=1e-20))
    w, x, y, z = q.unbind(-1)
    r = torch.stack((
        1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
        2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
        2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)), dim=-1).reshape(-1, 3, 3)
    sigma = r @ torch.diag_embed(scales * scales) @ r.transpose(1, 2)
    return torch.stack((sigma[:, 0, 0], sigma[:, 0, 1], sigma[:, 0, 2], sigma[:, 1, 1], sigma[:, 1, 2],
                        sigma[:, 2, 2]), dim=-1)


def sh_basis(direction: torch.Tensor, degree: int) -> torch.Tensor:
    """[N, (degree+1)^2] real SH basis at unit directions (3DGS constants; band 0 weight 1)."""
    x, y, z = direction.unbind(-1)
    out = [torch.ones_like(x)]
    if degree >= 1:
        out += [-SH_C1 * y, SH_C1 * z, -SH_C1 * x]
    if degree >= 2:
        xx, yy, zz = x * x, y * y, z * z
        out += [SH_C2[0] * x * y, SH_C2[1] * y * z, SH_C2[2] * (2 * zz - xx - yy), SH_C2[3] * x * z,
                SH_C2[4] * (xx - yy)]
    if degree >= 3:
        out += [SH_C3[0] * y * (3 * xx - yy), SH_C3[1] * x * y * z, SH_C3[2] * y * (4 * zz - xx - yy),
                SH_C3[3] * z * (2 * zz - 3 * xx - 3 * yy), SH_C3[4] * x * (4 * zz - xx - yy),
                SH_C3[5] * z * (xx - yy), SH_C3[6] * x * (xx - 3 * yy)]
    return torch.stack(out, dim=-1)


@dataclass
class CameraStructure:
    """Frozen discrete decisions of one camera."""
    order: torch.Tensor        # visible Gaussian indices in (depth, id) order
    blended: torch.Tensor      # [P, V] bool: pixel blends splat (in order)
    alpha_clamped: torch.Tensor  # [P, V] bool: alpha hit the 0.99 cap
    colour_live: torch.Tensor  # [V, 3] bool: colour not clamped at zero
    inside_x: torch.Tensor     # [V] bool: tan-FOV clamp inactive in x
    inside_y: torch.Tensor
    compensation_live: torch.Tensor  # [V] bool (antialiased only)


@dataclass
class Options:
    height: int
    width: int
    near_plane: float = 0.01
    far_plane: float = 100.0
    support_sigma: float = 3.33
    covariance_epsilon: float = 0.3
    antialiased: bool = False
    tile_size: int = 16


def _camera_geometry(means, cov6, opacities, view, intrinsics, env, opt: Options):
    """Per-Gaussian projected quantities of one camera, in the dtype of the inputs."""
    e_r, e_t = env[:3, :3], env[:3, 3]
    v_r, v_t = view[:3, :3], view[:3, 3]
    p_env = means @ e_r.T + e_t
    p_cam = p_env @ v_r.T + v_t
    x, y, z = p_cam.unbind(-1)
    fx, fy, cx, cy = intrinsics[0, 0], intrinsics[1, 1], intrinsics[0, 2], intrinsics[1, 2]
    inverse_z = 1.0 / z
    mean_x = fx * x * inverse_z + cx
    mean_y = fy * y * inverse_z + cy
    w = v_r @ e_r
    sigma = torch.stack((cov6[:, 0], cov6[:, 1], cov6[:, 2], cov6[:, 1], cov6[:, 3], cov6[:, 4],
                         cov6[:, 2], cov6[:, 4], cov6[:, 5]), dim=-1).reshape(-1, 3, 3)
    sigma_cam = w @ sigma @ w.T
    tan_x = 0.5 * opt.width / fx
    tan_y = 0.5 * opt.height / fy
    lim_xp = (opt.width - cx) / fx + 0.3 * tan_x
    lim_xn = cx / fx + 0.3 * tan_x
    lim_yp = (opt.height - cy) / fy + 0.3 * tan_y
    lim_yn = cy / fy + 0.3 * tan_y
    tx = x * inverse_z
    ty = y * inverse_z
    return dict(p_cam=p_cam, x=x, y=y, z=z, inverse_z=inverse_z, mean_x=mean_x, mean_y=mean_y, sigma_cam=sigma_cam,
                tx=tx, ty=ty, lim=(lim_xp, lim_xn, lim_yp, lim_yn), fx=fx, fy=fy, w=w, opacities=opacities)


def _splat(geo, opt: Options, inside_x, inside_y, compensation_live):
    """EWA footprint of visible Gaussians given the frozen clamp flags."""
    x, y, z, inverse_z = geo["x"], geo["y"], geo["z"], geo["inverse_z"]
    lim_xp, lim_xn, lim_yp, lim_yn = geo["lim"]
    fx, fy = geo["fx"], geo["fy"]
    tx_clamped = torch.where(inside_x, geo["tx"], torch.clamp(geo["tx"], -lim_xn, lim_xp).detach())
    ty_clamped = torch.where(inside_y, geo["ty"], torch.clamp(geo["ty"], -lim_yn, lim_yp).detach())
    clamped_x = z * tx_clamped
    clamped_y = z * ty_clamped
    j00 = fx * inverse_z
    j02 = -fx * clamped_x * inverse_z * inverse_z
    j11 = fy * inverse_z
    j12 = -fy * clamped_y * inverse_z * inverse_z
    s = geo["sigma_cam"]
    c00, c01, c02, c11, c12, c22 = s[:, 0, 0], s[:, 0, 1], s[:, 0, 2], s[:, 1, 1], s[:, 1, 2], s[:, 2, 2]
    raw00 = j00 * j00 * c00 + 2 * j00 * j02 * c02 + j02 * j02 * c22
    cov01 = j00 * j11 * c01 + j00 * j12 * c02 + j02 * j11 * c12 + j02 * j12 * c22
    raw11 = j11 * j11 * c11 + 2 * j11 * j12 * c12 + j12 * j12 * c22
    cov00 = raw00 + opt.covariance_epsilon
    cov11 = raw11 + opt.covariance_epsilon
    det = cov00 * cov11 - cov01 * cov01
    opacity = geo["opacities"]
    ratio = (raw00 * raw11 - cov01 * cov01) / det
    if opt.antialiased:
        floor2 = MIN_ANTIALIAS_COMPENSATION ** 2
        compensation = torch.sqrt(torch.where(compensation_live, ratio, torch.full_like(ratio, floor2)))
        opacity = opacity * compensation
    conic = torch.stack((cov11 / det, -cov01 / det, cov00 / det), dim=-1)
    return dict(cov00=cov00, cov11=cov11, det=det, opacity=opacity, conic=conic, ratio=ratio)


def render(means, scales, quats, opacities, features, viewmats, intrinsics, opt: Options, *, sh_degree: int,
           env_xforms=None, structure: list[CameraStructure] | None = None, covariance6=None):
    """Renders every camera; returns ({rgb [C,H,W,3], depth [C,H,W], alpha [C,H,W]}, structure).

    ``covariance6`` replaces (scales, quats) when given.  With ``structure``
    the discrete decisions of an earlier call are reused.
    """
    dtype = means.dtype
    cov6 = covariances(scales, quats) if covariance6 is None else covariance6
    cameras = viewmats.shape[0]
    if env_xforms is None:
        env_xforms = torch.eye(4, dtype=dtype).expand(cameras, 4, 4)
    height, width, ts = opt.height, opt.width, opt.tile_size
    tiles_x = (width + ts - 1) // ts
    py, px = torch.meshgrid(torch.arange(height, dtype=dtype) + 0.5, torch.arange(width, dtype=dtype) + 0.5,
                            indexing="ij")
    px, py = px.reshape(-1), py.reshape(-1)
    pixel_tile_x = (torch.arange(width).repeat(height)) // ts
    pixel_tile_y = (torch.arange(height).repeat_interleave(width)) // ts
    bands = (sh_degree + 1) ** 2
    out_rgb, out_depth, out_alpha, structures = [], [], [], []
    for c in range(cameras):
        geo = _camera_geometry(means, cov6, opacities, viewmats[c], intrinsics[c], env_xforms[c], opt)
        if structure is None:
            st = _decide(geo, opt, features, sh_degree, env_xforms[c], viewmats[c], means, px, py, pixel_tile_x,
                         pixel_tile_y)
        else:
            st = structure[c]
        structures.append(st)
        order = st.order
        per_gaussian = ("p_cam", "x", "y", "z", "inverse_z", "mean_x", "mean_y", "sigma_cam", "tx", "ty", "opacities")
        sub = {k: (v[order] if k in per_gaussian else v) for k, v in geo.items()}
        sp = _splat(sub, opt, st.inside_x, st.inside_y, st.compensation_live)
        record_opacity = sp["opacity"]  # clamp(., 0, 1) is the identity on visible splats here
        # Colour at the view direction from the camera centre (scene coordinates).
        w = sub["w"]
        view = viewmats[c]
        env = env_xforms[c]
        t = view[:3, :3] @ env[:3, 3] + view[:3, 3]
        origin = -torch.linalg.solve(w, t)
        f = features[order][:, : 3 * bands].reshape(-1, bands, 3)
        if sh_degree > 0:
            d = means[order] - origin
            norm = torch.sqrt((d * d).sum(-1, keepdim=True))
            basis = sh_basis(d / norm, sh_degree)
            colour = (basis.unsqueeze(-1) * f).sum(1)
            colour = torch.where(st.colour_live, colour, torch.zeros_like(colour))
        else:
            colour = f[:, 0, :]
        dx = px[:, None] - sub["mean_x"][None, :]
        dy = py[:, None] - sub["mean_y"][None, :]
        k = sp["conic"]
        power = -0.5 * (k[None, :, 0] * dx * dx + 2.0 * k[None, :, 1] * dx * dy + k[None, :, 2] * dy * dy)
        raw_alpha = record_opacity[None, :] * torch.exp(power)
        alpha = torch.where(st.alpha_clamped, torch.full_like(raw_alpha, MAX_ALPHA), raw_alpha)
        alpha = torch.where(st.blended, alpha, torch.zeros_like(alpha))
        keep = 1.0 - alpha
        before = torch.cumprod(torch.cat((torch.ones_like(keep[:, :1]), keep[:, :-1]), dim=1), dim=1)
        weight = alpha * before
        rgb = weight @ colour
        depth_sum = weight @ sub["z"]
        accumulated = 1.0 - torch.prod(keep, dim=1)
        depth = torch.where(accumulated > 1e-8, depth_sum / torch.clamp(accumulated, min=1e-30),
                            torch.full_like(depth_sum, math.inf))
        out_rgb.append(rgb.reshape(height, width, 3))
        out_depth.append(depth.reshape(height, width))
        out_alpha.append(accumulated.reshape(height, width))
    return {"rgb": torch.stack(out_rgb), "depth": torch.stack(out_depth), "alpha": torch.stack(out_alpha)}, structures


def _decide(geo, opt: Options, features, sh_degree, env, view, means, px, py, pixel_tile_x, pixel_tile_y):
    """Discrete decisions of one camera at the current parameters (no gradient)."""
    with torch.no_grad():
        z = geo["z"]
        opacity = geo["opacities"]
        lim_xp, lim_xn, lim_yp, lim_yn = geo["lim"]
        inside_x = (geo["tx"] >= -lim_xn) & (geo["tx"] <= lim_xp)
        inside_y = (geo["ty"] >= -lim_yn) & (geo["ty"] <= lim_yp)
        candidate = (opacity >= ALPHA_THRESHOLD) & (z >= opt.near_plane) & (z <= opt.far_plane)
        live = torch.ones_like(inside_x)
        if opt.antialiased:
            ratio = _splat(geo, opt, inside_x, inside_y, live)["ratio"]
            live = ratio > MIN_ANTIALIAS_COMPENSATION ** 2
        sp = _splat(geo, opt, inside_x, inside_y, live)
        det = sp["det"]
        eff = sp["opacity"]
        candidate &= det > 0
        candidate &= eff >= ALPHA_THRESHOLD
        extend = torch.clamp(torch.sqrt(2.0 * torch.log(torch.clamp(eff, min=1e-30) / ALPHA_THRESHOLD)),
                             max=opt.support_sigma)
        radius_x = torch.ceil(extend * torch.sqrt(torch.clamp(sp["cov00"], min=0)))
        radius_y = torch.ceil(extend * torch.sqrt(torch.clamp(sp["cov11"], min=0)))
        mx, my = geo["mean_x"], geo["mean_y"]
        candidate &= (radius_x > 0) | (radius_y > 0)
        candidate &= ~((mx + radius_x <= 0) | (my + radius_y <= 0) | (mx - radius_x >= opt.width)
                       | (my - radius_y >= opt.height))
        index = torch.nonzero(candidate).squeeze(-1)
        depth32 = z[index].float()
        order = index[torch.argsort(depth32, stable=True)]
        ts = float(opt.tile_size)
        tiles_x = (opt.width + opt.tile_size - 1) // opt.tile_size
        tiles_y = (opt.height + opt.tile_size - 1) // opt.tile_size
        rx, ry = radius_x[order], radius_y[order]
        ox, oy = mx[order], my[order]
        min_x = torch.clamp(torch.floor(ox / ts - rx / ts), min=0)
        min_y = torch.clamp(torch.floor(oy / ts - ry / ts), min=0)
        max_x = torch.clamp(torch.ceil(ox / ts + rx / ts), max=tiles_x)
        max_y = torch.clamp(torch.ceil(oy / ts + ry / ts), max=tiles_y)
        in_rect = ((pixel_tile_x[:, None] >= min_x[None, :]) & (pixel_tile_x[:, None] < max_x[None, :])
                   & (pixel_tile_y[:, None] >= min_y[None, :]) & (pixel_tile_y[:, None] < max_y[None, :]))
        k = sp["conic"][order]
        dx = px[:, None] - ox[None, :]
        dy = py[:, None] - oy[None, :]
        power = -0.5 * (k[None, :, 0] * dx * dx + 2.0 * k[None, :, 1] * dx * dy + k[None, :, 2] * dy * dy)
        raw_alpha = eff[order][None, :] * torch.exp(power)
        alpha = torch.clamp(raw_alpha, max=MAX_ALPHA)
        valid = in_rect & (power <= 0) & (alpha >= ALPHA_THRESHOLD)
        a = torch.where(valid, alpha, torch.zeros_like(alpha))
        after = torch.cumprod(1.0 - a, dim=1)
        blended = valid & (after > TRANSMITTANCE_THRESHOLD)
        bands = (sh_degree + 1) ** 2
        colour_live = torch.ones((order.numel(), 3), dtype=torch.bool)
        if sh_degree > 0:
            w = geo["w"]
            t = view[:3, :3] @ env[:3, 3] + view[:3, 3]
            origin = -torch.linalg.solve(w, t)
            d = means[order] - origin
            basis = sh_basis(d / torch.sqrt((d * d).sum(-1, keepdim=True)), sh_degree)
            f = features[order][:, : 3 * bands].reshape(-1, bands, 3)
            colour_live = (basis.unsqueeze(-1) * f).sum(1) > 0
        return CameraStructure(order=order, blended=blended, alpha_clamped=raw_alpha > MAX_ALPHA,
                               colour_live=colour_live, inside_x=inside_x[order], inside_y=inside_y[order],
                               compensation_live=live[order])
