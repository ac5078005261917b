"""ProbeShader (tina/probe.py:4-32): per-pixel face index and texcoord readback.

The reference records elmid[P] = f and texcoord[P] inside render_color.  Here
the shade launch leaves the winning global id in the packed key buffer;
post_shade copies the object's pixels (ids in [id_base, id_base + nfaces)) into
elmid as local face indices and writes their interpolated texcoords (zero
without texturing, as the reference passes V(0, 0); tina_gbuffer's texcoord
mode), which is what the reference leaves after the last object is shaded.
"""
from .field import Field
from .shader import IShader, gbuffer

__all__ = ['ProbeShader']


class ProbeShader(IShader):
    def __init__(self, res):
        import torch
        self.res = (int(res[0]), int(res[1])) if not isinstance(res, int) else (res, res)
        W, H = self.res
        self.elmid = Field(torch.full((H, W), -1, dtype=torch.int32, device='cuda'))
        self.texcoord = Field(torch.zeros((H, W, 2), dtype=torch.float32, device='cuda'))

    def clear_buffer(self):
        self.elmid.storage.fill_(-1)
        self.texcoord.storage.zero_()

    def post_shade(self, raster, id_base, nfaces):
        import torch
        ids, _ = raster.engine.resolve(want_depth=False)
        local = ids.long() - id_base
        inside = (ids >= 0) & (local >= 0) & (local < nfaces)
        self.elmid.storage.copy_(torch.where(inside, local.to(torch.int32), self.elmid.storage))
        gbuffer(raster, 2, self.texcoord.storage)
