"""Differentiable rendering: forward/backward pair over the C ABI.

forward() is lucent_forward: Renderer.traceSamples into a fresh film (the
image is the running mean of `spp` samples from seeds seed0..seed0+spp-1),
keeping each path's radiance on the device; backward() replays those same
samples (lucent_backward) and returns the gradient of
sum_pixels dL_dI · image with respect to every material's (albedo rgb,
roughness, metallic) and every environment texel channel.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _native
from .renderer import Renderer, RenderSettings, camera_struct
from .runtime import Context
from .scene import Scene, pack_materials

PARAM_NAMES = ("albedo_r", "albedo_g", "albedo_b", "roughness", "metallic")


class _DeviceArray:
    """A device buffer described by __cuda_array_interface__, so that torch
    can view it without a copy (torch.as_tensor(..., device="cuda"))."""

    def __init__(self, ptr: int, shape: tuple, typestr: str = "<f4"):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (ptr, False), "version": 2,
                                         "strides": None}


class DiffRenderer:
    def __init__(self, ctx: Context, scene: Scene, width: int, height: int, max_bounces: int = 6):
        self.ctx = ctx
        self.scene = scene
        self.width, self.height = width, height
        self.renderer = Renderer(ctx, width, height, RenderSettings(max_bounces=max_bounces, firefly_clamp=0.0))
        self.renderer.load_scene(scene)
        self.materials = list(self.renderer.packed.material_list)
        ew, eh, env = scene.env.map
        self.env_size = (ew, eh)
        self.env = np.asarray(env, np.float32).copy()
        self._version = 0      # bumped by every parameter change
        self.adjoint_occupancy = 4  # register variant of the replay kernel (2, 3, 4); 4 measured fastest
        self.adjoint_tile = (0, 0)  # job tile of the replay (0 = 8 x 8)
        self._record = None    # (spp, seed0, version, device buffer) of the last forward

    @property
    def settings(self):
        return self.renderer.settings

    def set_materials(self, materials) -> None:
        packed = pack_materials(materials)
        _native.check(_native.lib().lucent_scene_set_materials(self.ctx.handle, self.renderer.device_scene.handle,
                                                               packed.ctypes.data, len(materials)))
        self.materials = list(materials)
        self._version += 1

    def set_env_texels(self, rgb, rebuild_sampler: bool = False, on_device: int = 0) -> None:
        """New environment texels (same size).  The importance sampler stays
        the one built with the lighting unless `rebuild_sampler`: rebuilt
        from the new texels by the host build (on_device 0) or the FP64
        device build (1; the same table either way), or by the parallel FP64
        device build (2; the same distribution, other alias pairs)."""
        arr = np.ascontiguousarray(np.asarray(rgb, np.float32).reshape(-1))
        _native.check(_native.lib().lucent_scene_update_env_texels(self.ctx.handle, self.renderer.device_scene.handle,
                                                                   arr.ctypes.data, *self.env_size))
        if rebuild_sampler:
            self.renderer.device_scene.rebuild_env_sampler(on_device)
        self.env = arr.copy()
        self._version += 1

    def forward(self, spp: int, seed0: int = 0, record: bool = True, device: bool = False,
                stream: Optional[int] = None):
        """Image (H, W, 3) as float64, rows bottom-up (lucent_forward).  With
        `record`, each path's radiance is kept on the device so that a
        backward() with the same parameters and seeds need not re-trace.
        `stream`: a HIP stream to enqueue on (the call then returns without
        waiting; device=True is required)."""
        import torch
        r = self.renderer
        buf = None
        if record:
            buf = torch.empty(spp * self.height * self.width * 4, dtype=torch.float32, device=torch.device("cuda:0"))
        cam = camera_struct(self.scene.camera)
        st = self.settings.struct()
        st.firefly_clamp, st.flags, st.waves_per_cu = 0.0, 0, 0
        if stream is not None and not device:
            raise ValueError("forward(stream=...) returns the device image: pass device=True")
        if stream is None:
            torch.cuda.synchronize()
        _native.check(_native.lib().lucent_forward(self.ctx.handle, r.device_scene.handle, r.film.handle, C.byref(cam),
                                                   C.byref(st), spp, seed0, buf.data_ptr() if buf is not None else None,
                                                   stream))
        if stream is None:
            self.ctx.synchronize()
        self._record = (spp, seed0, self._version, buf) if record else None
        if device:
            return self.image_view()
        return r.film.read()["color"][..., :3].astype(np.float64)

    def image_view(self):
        """The film's colour target as a torch float32 (H, W, 4) view on the
        device (rgb, luminance²), valid until the next trace: forward(...,
        device=True) returns it so that a loss and its gradient can stay in
        device memory."""
        import torch
        ptr = self.renderer.film.device_ptrs()["color"]
        return torch.as_tensor(_DeviceArray(ptr, (self.height, self.width, 4)), device=torch.device("cuda:0"))

    def detached_forward(self, value_materials, spp: int, seed0: int = 0) -> np.ndarray:
        """Image (H, W, 3) of the detached estimator (lucent_detached_forward):
        sampling from the uploaded materials, BSDF values from
        `value_materials`."""
        import torch
        packed = pack_materials(value_materials)
        if len(value_materials) != len(self.materials):
            raise ValueError("material count differs from the uploaded scene")
        h, w = self.height, self.width
        out = torch.empty(h * w * 4, dtype=torch.float32, device=torch.device("cuda:0"))
        cam = camera_struct(self.scene.camera)
        st = _native.RenderSettingsStruct(self.settings.max_bounces, 0.0, 0, 0, 0)
        torch.cuda.synchronize()
        _native.check(_native.lib().lucent_detached_forward(
            self.ctx.handle, self.renderer.device_scene.handle, C.byref(cam), C.byref(st), w, h, spp, seed0,
            packed.ctypes.data, len(value_materials), out.data_ptr(), None))
        self.ctx.synchronize()
        return out.cpu().numpy().reshape(h, w, 4)[..., :3].astype(np.float64)

    def backward(self, dl_di, spp: int, seed0: int = 0, env: bool = True, use_record: bool = True,
                 device: bool = False, stream: Optional[int] = None):
        """Gradients (material_count x 5, env (h, w, 3)) of sum dl_di · image.
        dl_di: (h, w, 3) host array, or a torch float32 (h, w, 4) tensor on the
        device (its fourth channel is ignored).  Uses the radiance recorded by
        the last forward() when it was made with the same samples, seeds and
        parameters; otherwise re-traces.  device=True returns the float64
        gradients as device tensors instead of host arrays; `stream` enqueues
        on that HIP stream without waiting (device=True required)."""
        import torch
        dev = torch.device("cuda:0")
        h, w = self.height, self.width
        if isinstance(dl_di, torch.Tensor):
            if dl_di.device.type != "cuda" or dl_di.dtype != torch.float32 or tuple(dl_di.shape) != (h, w, 4):
                raise ValueError(f"device dl_di must be a float32 ({h}, {w}, 4) CUDA tensor")
            d_dl = dl_di.contiguous()
        else:
            d = np.zeros((h, w, 4), np.float32)
            d[..., :3] = dl_di
            d_dl = torch.from_numpy(d).to(dev)
        nm = len(self.materials)
        g_mat = torch.zeros(nm * 5, dtype=torch.float64, device=dev)
        ew, eh = self.env_size
        g_env = torch.zeros(ew * eh * 3, dtype=torch.float64, device=dev) if env else None
        cam = camera_struct(self.scene.camera)
        st = _native.RenderSettingsStruct(self.settings.max_bounces, 0.0, 0, 0, self.adjoint_occupancy,
                                          *self.adjoint_tile)
        if stream is not None and not device:
            raise ValueError("backward(stream=...) returns device gradients: pass device=True")
        if stream is None:
            torch.cuda.synchronize()
        rec = self._record
        use = use_record and rec is not None and rec[:3] == (spp, seed0, self._version)
        _native.check(_native.lib().lucent_backward(
            self.ctx.handle, self.renderer.device_scene.handle, C.byref(cam), C.byref(st), w, h, spp, seed0,
            d_dl.data_ptr(), rec[3].data_ptr() if use else None, g_mat.data_ptr(), g_env.data_ptr() if env else None,
            stream))
        if stream is None:
            self.ctx.synchronize()
        if device:
            return g_mat.reshape(nm, 5), (g_env.reshape(eh, ew, 3) if env else None)
        gm = g_mat.cpu().numpy().reshape(nm, 5)
        ge = g_env.cpu().numpy().reshape(eh, ew, 3) if env else None
        return gm, ge

    def close(self) -> None:
        self.renderer.close()


def l2_loss(image: np.ndarray, target: np.ndarray) -> float:
    return float(((image - target) ** 2).sum())
