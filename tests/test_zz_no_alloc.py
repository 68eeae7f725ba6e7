"""GPU test of rihvr_prepare_holograms_noalloc, the allocation-free variant of
rihvr_prepare_holograms whose scratch (the mean pass's partial sums and the frames' inverse
means) lives in a fixed per-device array: two calls in a row, the second with a smaller batch
(so stale scratch from the first call would show), each equal bit for bit to the allocating
entry point's output on the same frames, for 8-bit, 16-bit and float frames, with and without
a background; a batch beyond the scratch is refused before any device work."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fp(pkg, pad=0, center=(-1, -1), size=-1, remove_bg=False):
    fp = pkg.FileParams()
    fp.solver.zero_padding = pad
    fp.center_x, fp.center_y = center
    fp.roi_size = size
    fp.remove_background = 1 if remove_bg else 0
    return fp


def _frames(torch, dtype, n, h, w, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if dtype == torch.float32:
        return torch.rand((n, h, w), generator=g, device="cuda") * 300
    hi = 256 if dtype == torch.uint8 else 65536
    return torch.randint(0, hi, (n, h, w), generator=g, device="cuda").to(dtype)


@pytest.mark.parametrize("dtype_name", ["uint8", "uint16", "float32"])
@pytest.mark.parametrize("remove_bg", [False, True])
def test_noalloc_equals_allocating_entry_point(pkg, gpu_available, dtype_name, remove_bg):
    import torch
    dtype = getattr(torch, dtype_name)
    # first call: 24 frames of 512^2 (pad 8); second call: 5 frames of 200x200 cropped to 96^2
    calls = [dict(n=24, h=512, w=512, fp=dict(pad=8), seed=1),
             dict(n=5, h=200, w=200, fp=dict(center=(90, 110), size=96, pad=4), seed=2)]
    for c in calls:
        print(f"noalloc case: {dtype_name} bg={remove_bg} {c['n']}x{c['h']}x{c['w']} {c['fp']}", flush=True)
        fp = _fp(pkg, remove_bg=remove_bg, **c["fp"])
        fr = _frames(torch, dtype, c["n"], c["h"], c["w"], c["seed"])
        bg = (torch.rand((c["h"], c["w"]), device="cuda") * 200) if remove_bg else None
        exp = torch.from_numpy(pkg.prepare_holograms(fp, fr, bg)).cuda()  # device frames, host output
        got = torch.full_like(exp, float("nan"))
        pkg.prepare_holograms_noalloc(fp, fr, got, bg)
        torch.cuda.synchronize()
        assert torch.equal(got.view(torch.float32).view(torch.int32), exp.view(torch.float32).view(torch.int32))


def test_noalloc_refuses_a_batch_beyond_its_scratch(pkg, gpu_available):
    import torch
    # 600 float frames of 1024^2 need 600 * 256 * 8 bytes of partial sums: more than 1 MiB
    fp = _fp(pkg)
    fr = torch.zeros((600, 1024, 1024), dtype=torch.float32, device="cuda")
    out = torch.empty((600, 1024, 1024), dtype=torch.complex64, device="cuda")
    with pytest.raises(pkg.RihvrError) as e:
        pkg.prepare_holograms_noalloc(fp, fr, out)
    assert e.value.status == 7
    del fr, out
    torch.cuda.empty_cache()
