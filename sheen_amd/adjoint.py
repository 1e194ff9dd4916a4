"""Differentiable rendering through the HIP adjoint kernel.

`DiffRenderer(scene)` keeps a scene resident on the GPU; `render(params)`
returns a (H, W, 4) float32 CUDA tensor and participates in autograd:
gradients reach the four parameter tensors through
`sheen_adjoint_backward` (path replay of the same frame, visibility held
fixed).  Parameters (float32 CUDA tensors):

    albedo     (materials, 3)   Material.diffuse_color
    emission   (materials, 3)   emitted radiance (emission_color * emission)
    intensity  (lights,)        Light.intensity
    texels     (texels,)        level-0 texels of every texture, RGB / 255

torch supplies device memory and the stream (its current stream, often the
null stream); the library orders its own stream after that stream on entry
and that stream after its own on return, so torch kernels that write the
parameters before a call and read the outputs after it need no
synchronisation.  torch ships its own HIP runtime (soname libamdhip64.so.7,
the same as the system ROCm one this library links): importing torch first
makes the dynamic loader resolve the library's HIP calls to torch's runtime,
so there is one runtime and torch's stream handles are valid here.
"""
import ctypes as C

import torch

from . import native
from .sheen import _scene_ptr


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class _Render(torch.autograd.Function):
    @staticmethod
    def forward(ctx, r, width, height, aa, albedo, emission, intensity, texels):
        albedo, emission, intensity, texels = (t.detach().contiguous() for t in (albedo, emission, intensity, texels))
        img = torch.empty((height, width, 4), device=albedo.device, dtype=torch.float32)
        native.check(r._lib.sheen_adjoint_forward(r._c, width, height, aa, _ptr(albedo), _ptr(emission), _ptr(intensity),
                                                  _ptr(texels), _ptr(img), _stream()))
        ctx.r, ctx.size = r, (width, height, aa)
        ctx.save_for_backward(albedo, emission, intensity, texels)
        return img

    @staticmethod
    def backward(ctx, grad):
        albedo, emission, intensity, texels = ctx.saved_tensors
        width, height, aa = ctx.size
        r = ctx.r
        grad = grad.detach().contiguous().float()
        d = [torch.empty_like(t) for t in (albedo, emission, intensity, texels)]
        native.check(r._lib.sheen_adjoint_backward(r._c, width, height, aa, _ptr(albedo), _ptr(emission),
                                                   _ptr(intensity), _ptr(texels), _ptr(grad), _ptr(d[0]), _ptr(d[1]),
                                                   _ptr(d[2]), _ptr(d[3]), _stream()))
        return (None, None, None, None) + tuple(d)


class DiffRenderer:
    """A scene on the GPU with differentiable albedo, emission, light
    intensities and texels."""

    def __init__(self, scene, width, height, aa):
        lib = native.capi()
        self._lib = lib
        self._c = lib.sheen_context_create(_scene_ptr(scene))
        if not self._c:
            raise RuntimeError(native.last_error())
        self.width, self.height, self.aa = width, height, aa
        m, l, t = C.c_int(), C.c_int(), C.c_longlong()
        native.check(lib.sheen_adjoint_sizes(self._c, C.byref(m), C.byref(l), C.byref(t)))
        self.sizes = {"materials": m.value, "lights": l.value, "texels": t.value}

    def params(self, device="cuda"):
        """The scene's own parameter values as CUDA tensors."""
        s = self.sizes
        p = {"albedo": torch.empty((s["materials"], 3), device=device),
             "emission": torch.empty((s["materials"], 3), device=device),
             "intensity": torch.empty((s["lights"],), device=device),
             "texels": torch.empty((s["texels"],), device=device)}
        native.check(self._lib.sheen_adjoint_get_params(self._c, _ptr(p["albedo"]), _ptr(p["emission"]),
                                                        _ptr(p["intensity"]), _ptr(p["texels"]), _stream()))
        return p

    def render(self, albedo, emission, intensity, texels):
        return _Render.apply(self, self.width, self.height, self.aa, albedo, emission, intensity, texels)

    def stream_fallbacks(self):
        """Caller-stream orderings that fell back to host synchronisation."""
        return self._lib.sheen_context_stream_fallbacks(self._c)

    def set_param(self, name, value):
        native.check(self._lib.sheen_context_set_param(self._c, name.encode(), int(value)))

    def close(self):
        if self._c:
            torch.cuda.synchronize()
            self._lib.sheen_context_destroy(self._c)
            self._c = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
