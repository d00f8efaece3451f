"""Differentiable frame (torch): images of a set of views as a function of the
instance transforms, with gradients through the analytic edge antialiasing.

    images = frame(transforms)        # transforms [T, 16] float32, column-major
    loss = sum((img - target) ** 2 for img, target in zip(images, targets))
    loss.backward()                   # transforms.grad

The forward uploads the transforms, renders the views (one batch frame,
two-pass Hi-Z culling) and returns each view's antialiased linear image
[H, W, 3] (vkgv_view_aa_image); the backward runs vkgv_view_aa_backward per
view into one float64 accumulator.  The rasterization itself has no
gradient: only the antialiasing term depends continuously on the
transforms (the kernels are in csrc/grad.hip).
"""
from __future__ import annotations

import numpy as np

from . import _native as N


class DifferentiableFrame:
    def __init__(self, renderer, views, cameras, flags: int = N.CULL_FRUSTUM | N.CULL_TWO_PASS):
        import torch
        self.r = renderer
        self.views = list(views)
        self.cams = list(cameras)
        if len(self.views) != len(self.cams):
            raise ValueError("one camera per view")
        self.flags = flags
        self.batch = renderer.create_batch(self.views)
        outer = self

        class _Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, transforms):
                return tuple(outer._forward(transforms))

            @staticmethod
            def backward(ctx, *g_images):
                return outer._backward(g_images)

        self._fn = _Fn

    def close(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None

    def _forward(self, transforms):
        import torch
        t = transforms.detach()
        if t.dim() != 2 or t.shape[1] != 16:
            raise ValueError("transforms must be [T, 16]")
        if t.is_cuda:
            # device to device on torch's stream: no host copy, no synchronization
            self.r.set_transforms_device(t.to(torch.float32).contiguous())
        else:
            self.r.set_transforms(np.ascontiguousarray(t.to(torch.float32).numpy()))
        self._count = t.shape[0]
        self._device = transforms.device
        self._dtype = transforms.dtype
        s = torch.cuda.current_stream()
        # the default torch stream (handle 0) means the context's own stream to the
        # library; aa_image below orders it with torch's stream
        self.batch.render(self.cams, self.flags, stream=s.cuda_stream or None)
        return [v.aa_image(c) for v, c in zip(self.views, self.cams)]

    def _backward(self, g_images):
        import torch
        g = torch.zeros((self._count, 16), dtype=torch.float64, device=self._device)
        for v, c, gi in zip(self.views, self.cams, g_images):
            if gi is None:
                continue
            v.aa_backward(c, gi.to(torch.float32).contiguous(), g)
        return g.to(self._dtype)

    def __call__(self, transforms):
        return self._fn.apply(transforms)
