"""Differentiable rendering of material colours: a torch.autograd.Function over the renderer's forward frame
and its path-replay adjoint (rayna_adjoint), and the inverse-rendering loop that recovers hidden colours.

    f = ColourRenderer(scene, camera, opts)
    img = f(colours)                      # (height, width, 3) float32, the frame's per-pixel sample mean
    loss = ((img - target) ** 2).mean()
    loss.backward()                       # colours.grad via path replay backpropagation on the GPU

The engine itself is not differentiable; the forward frame is Renderer::render (render/renderer.rs:224-264)
and the colours are the albedos (lambertian.rs, metal.rs, dielectric.rs, isotropic.rs) and light emissions
(light.rs) of the scene's materials.  Paths are retraced from the counter RNG in the adjoint, never stored.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np
import torch

from .renderer import Renderer


class _RenderColours(torch.autograd.Function):
    @staticmethod
    def forward(ctx, colours: torch.Tensor, owner: "ColourRenderer") -> torch.Tensor:
        img = owner._forward(colours.detach().to("cpu", torch.float32).numpy())
        ctx.owner = owner
        ctx.frame_token = owner._token
        return torch.from_numpy(img).to(colours.device)

    @staticmethod
    def backward(ctx, grad_img: torch.Tensor):
        owner = ctx.owner
        if owner._token != ctx.frame_token:
            raise RuntimeError("the renderer rendered another frame since this one; backward must come first")
        g = owner._adjoint(grad_img.detach().to("cpu", torch.float32).contiguous().numpy())
        return torch.from_numpy(g).to(device=grad_img.device, dtype=torch.float32), None


class ColourRenderer:
    """A Renderer whose material colours are a differentiable input.  Each call renders one frame of
    `opts.samples` samples with the given colours ((n_materials, 3), material-table order) and seed."""

    def __init__(self, scene, camera, opts):
        self.scene = scene
        self.opts = opts
        self.renderer = Renderer(scene, camera, opts)
        self._token = 0
        self.last_forward_stats = None
        self.last_adjoint_stats = None

    def close(self):
        self.renderer.close()

    def __call__(self, colours: torch.Tensor, seed: Optional[int] = None) -> torch.Tensor:
        if seed is not None and seed != self.opts.seed:
            self.opts = replace(self.opts, seed=int(seed))
            self.renderer.set_options(self.opts)
        return _RenderColours.apply(colours, self)

    def _forward(self, colours: np.ndarray) -> np.ndarray:
        self.renderer.set_material_colours(colours)  # clears accumulation: the frame below is frame 0
        out = self.renderer.render(layout="rowmajor")
        self._token += 1
        self.last_forward_stats = out.stats
        return out.img

    def _adjoint(self, d_image: np.ndarray) -> np.ndarray:
        grad, stats = self.renderer.adjoint(d_image, frame=0, layout="rowmajor")
        self.last_adjoint_stats = stats
        return grad.astype(np.float32)


@dataclass
class InverseResult:
    losses: List[float] = field(default_factory=list)
    param_rmse: List[float] = field(default_factory=list)   # over the identifiable hidden colours, per step
    forward_ms: List[float] = field(default_factory=list)   # device time of each forward frame
    adjoint_ms: List[float] = field(default_factory=list)   # device time of each adjoint
    identifiable: int = 0                                   # hidden colours with a non-zero gradient at step 0
    seconds: float = 0.0

    @property
    def adjoint_forward_ratio(self) -> float:
        return float(np.median(self.adjoint_ms) / np.median(self.forward_ms))


def recover_colours(steps: int = 200, lr: float = 0.02, samples: int = 16, target_samples: int = 1024,
                    init_seed: int = 7, log_every: int = 0) -> InverseResult:
    """The inverse-rendering workload: render the inverse_spheres scene at `target_samples` spp with its
    true colours, reset the hidden colours to U[0.1, 0.9], and fit them with Adam (learning rate `lr`) on
    the L2 image loss, one `samples`-spp forward frame plus its adjoint per step, a fresh seed per step."""
    from . import presets

    scene, cam, opts, hidden = presets.inverse_spheres(samples=samples)
    truth = scene.colours()
    tgt = Renderer(scene, cam, replace(opts, samples=target_samples, seed=10_000))
    target = torch.from_numpy(tgt.render(layout="rowmajor").img.copy())
    tgt.close()

    rng = np.random.Generator(np.random.PCG64(init_seed))
    init = truth.copy()
    init[hidden] = rng.uniform(0.1, 0.9, (len(hidden), 3)).astype(np.float32)
    params = torch.tensor(init[hidden], requires_grad=True)
    known = torch.from_numpy(truth)
    opt = torch.optim.Adam([params], lr=lr)
    f = ColourRenderer(scene, cam, opts)
    res = InverseResult()
    mask = None
    t0 = time.time()
    for step in range(steps):
        colours = torch.cat([params, known[len(hidden):]], 0)
        img = f(colours, seed=1000 + step)
        loss = ((img - target) ** 2).mean()
        opt.zero_grad()
        loss.backward()
        if mask is None:
            mask = (params.grad.abs() > 0).numpy()
            res.identifiable = int(mask.sum())
        opt.step()
        with torch.no_grad():
            params.clamp_(min=1e-3)
        err = (params.detach().numpy() - truth[hidden])[mask]
        res.losses.append(loss.item())
        res.param_rmse.append(float(np.sqrt((err ** 2).mean())))
        res.forward_ms.append(f.last_forward_stats.kernel_ms)
        res.adjoint_ms.append(f.last_adjoint_stats.kernel_ms)
        if log_every and (step % log_every == 0 or step == steps - 1):
            print(f"step {step}: loss {res.losses[-1]:.6f} param rmse {res.param_rmse[-1]:.4f}", flush=True)
    res.seconds = time.time() - t0
    f.close()
    return res
