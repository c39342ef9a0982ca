"""Differentiable rendering: a torch.autograd.Function over the forward render and its path-replay adjoint.

Parameters are per-material albedo and emission (M, 3), per-light colour (L, 3) and, optionally,
per-material roughness (M,) and the scene's texels (the textures' rgb values, any shape with N * 3
elements). The forward pass uploads them (IPathTracer::update_scene_delta, update_texels), renders the
frame (one full-frame tile per sample, render_tiles) and returns the HDR image (resolve_hdr: film sum *
1/count). The backward pass hands dLoss/d(film sum) = dLoss/d(image) / count to pt_hip_render_adjoint,
which replays every path with the same random numbers and accumulates the parameter gradients with
fp32 atomics — nothing about the forward paths is stored.

MultiViewRenderer does the same for several cameras on one scene through one batch (pt_batch_*):
every view's paths share one launch sequence per frame, forward and adjoint.
"""
from __future__ import annotations

import numpy as np
import torch

from .scene import RenderSettings, Scene
from .tracer import HipPathTracer


class _RenderFn(torch.autograd.Function):
    """The backward pass replays the paths at the parameters of its own forward pass: each upload bumps
    the renderer's version, the forward keeps its parameters and version, and a backward that finds
    the renderer holding other parameters (another forward ran in between) uploads its own again."""

    @staticmethod
    def forward(ctx, albedo, emission, light_color, roughness, texture, renderer):
        renderer.upload(albedo, emission, light_color, roughness, texture)
        image, count = renderer.render()
        ctx.renderer = renderer
        ctx.version = renderer.version
        ctx.has_roughness = roughness is not None
        ctx.texture_shape = None if texture is None else texture.shape
        ctx.save_for_backward(count, albedo, emission, light_color, roughness, texture)
        return image

    @staticmethod
    def backward(ctx, grad_image):
        count, albedo, emission, light_color, roughness, texture = ctx.saved_tensors
        r = ctx.renderer
        if r.version != ctx.version:
            r.upload(albedo, emission, light_color, roughness, texture)
        inv = 1.0 / torch.clamp(count, min=1).to(torch.float32)
        film_grad = (grad_image.to(torch.float32) * inv[..., None]).contiguous()
        dev = film_grad.device
        ga = torch.zeros((r.n_materials, 3), dtype=torch.float32, device=dev)
        ge = torch.zeros((r.n_materials, 3), dtype=torch.float32, device=dev)
        gl = torch.zeros((max(r.n_lights, 1), 3), dtype=torch.float32, device=dev)
        gr = torch.zeros(r.n_materials, dtype=torch.float32, device=dev) if ctx.has_roughness else None
        gt = None
        if ctx.texture_shape is not None:
            gt = torch.zeros(ctx.texture_shape, dtype=torch.float32, device=dev)
        r.adjoint(film_grad, ga, ge, gl[: r.n_lights], gr, gt)
        return ga, ge, gl[: r.n_lights], gr, gt, None


class DifferentiableRenderer:
    """Owns one HipPathTracer over `scene`; calling it with (albedo, emission, light_color) tensors
    renders the image and makes it differentiable with respect to them."""

    def __init__(self, scene: Scene, settings: RenderSettings, device: int = 0, frame_index: int = 0):
        self.scene = scene
        self.settings = settings
        self.frame_index = frame_index
        self.tracer = HipPathTracer(device)
        for st in (self.tracer.configure(settings), self.tracer.load_scene_snapshot(scene),
                   self.tracer.build_or_update_acceleration()):
            if not st:
                raise RuntimeError(st.message)
        self.n_materials = len(scene.materials)
        self.n_lights = len(scene.lights)
        self.tiles = [(0, 0, settings.width, settings.height, s, 0, 0) for s in range(settings.spp)]
        self.device = torch.device("cuda", device)
        self.version = 0  # bumped by every upload of parameters

    def initial_parameters(self):
        m, l = self.scene.materials, self.scene.lights
        t = lambda a: torch.tensor(np.asarray(a, np.float32), device=self.device)
        return t(m["albedo"]), t(m["emissive"]), t(l["color"]) if len(l) else t(np.zeros((0, 3)))

    def initial_texture(self):
        """The scene's texels as one (N, 3) float32 device tensor (textures in order)."""
        texels, _ = self.scene.texture_arrays()
        return torch.from_numpy(texels).to(self.device)

    def _records(self, albedo, emission, light_color, roughness):
        mats = self.scene.materials.copy()
        mats["albedo"] = albedo.detach().to("cpu", torch.float32).numpy()
        mats["emissive"] = emission.detach().to("cpu", torch.float32).numpy()
        if roughness is not None:
            mats["roughness"] = roughness.detach().to("cpu", torch.float32).numpy()
        lights = self.scene.lights.copy()
        if len(lights):
            lights["color"] = light_color.detach().to("cpu", torch.float32).numpy()
        return mats, lights

    def upload(self, albedo, emission, light_color, roughness=None, texture=None):
        mats, lights = self._records(albedo, emission, light_color, roughness)
        if not self.tracer.update_scene_delta(mats, lights if len(lights) else None):
            raise RuntimeError(self.tracer.status().last_error)
        if texture is not None:
            self.tracer.update_texels(texture.detach().to(self.device, torch.float32).contiguous())
        self.version += 1

    def adjoint(self, film_grad, ga, ge, gl, gr, gt):
        self.tracer.render_adjoint(self.tiles, film_grad, ga, ge, gl, self.frame_index, grad_roughness=gr,
                                   grad_texels=gt)

    def render(self):
        if not self.tracer.reset_accumulation() or not self.tracer.render_tiles(self.tiles, self.frame_index):
            raise RuntimeError(self.tracer.status().last_error)
        rgb, cnt = self.tracer.film()
        count = torch.from_numpy(cnt.astype(np.int32)).to(self.device)
        inv = (1.0 / np.maximum(cnt, 1).astype(np.float32))[..., None]
        image = torch.from_numpy(rgb * inv).to(self.device)
        return image, count

    def __call__(self, albedo, emission, light_color, roughness=None, texture=None):
        return _RenderFn.apply(albedo, emission, light_color, roughness, texture, self)

    def close(self):
        self.tracer.close()


class MultiViewRenderer(DifferentiableRenderer):
    """Several cameras on one scene, rendered (and differentiated) through one HipBatch: the scene plus
    one view per further camera. Calling it returns the images (V, H, W, 3), V = len(cameras)."""

    def __init__(self, scene: Scene, cameras, width: int, height: int, spp: int, settings: RenderSettings,
                 device: int = 0, frame_index: int = 0, use_graph: bool = False, leaf_size: int = 4):
        import dataclasses

        from .tracer import HipBatch

        self.scene = scene
        self.settings = settings
        self.frame_index = frame_index
        self.use_graph = use_graph
        self.shape = (len(cameras), height, width)
        self.batch = HipBatch(settings, device)
        self.batch.add_scene(dataclasses.replace(scene, camera=cameras[0]), width, height, spp, leaf_size=leaf_size)
        for cam in cameras[1:]:
            self.batch.add_view(0, cam, width, height, spp)
        self.batch.build()
        self.n_materials = len(scene.materials)
        self.n_lights = len(scene.lights)
        self.device = torch.device("cuda", device)
        self.version = 0

    def upload(self, albedo, emission, light_color, roughness=None, texture=None):
        mats, lights = self._records(albedo, emission, light_color, roughness)
        self.batch.update_scene_delta(mats, lights if len(lights) else None)
        if texture is not None:
            self.batch.update_texels(texture.detach().to(self.device, torch.float32).contiguous())
        self.version += 1

    def render(self):
        self.batch.reset_accumulation()
        self.batch.render(self.frame_index, use_graph=self.use_graph)
        films = self.batch.films()
        rgb = np.stack([f for f, _ in films])
        cnt = np.stack([c for _, c in films])
        count = torch.from_numpy(cnt.astype(np.int32)).to(self.device)
        image = torch.from_numpy(rgb / np.maximum(cnt, 1).astype(np.float32)[..., None]).to(self.device)
        return image, count

    def adjoint(self, film_grad, ga, ge, gl, gr, gt):
        self.batch.render_adjoint(film_grad, ga, ge, gl, self.frame_index, grad_roughness=gr, grad_texels=gt)

    def close(self):
        self.batch.close()
