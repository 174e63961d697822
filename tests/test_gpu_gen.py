"""Device image generator (split_hip_gen_images) against the host generator.

Without noise the two are bit-identical (same hashes, IEEE double arithmetic
without contraction, exact nearest-site search).  With noise the device's
log/cos may differ from glibc's in the last bit, so a pixel may differ by
one float ulp; nearly all pixels still match exactly.
"""
import numpy as np
import pytest

import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu

SPECS = [
    W.ImageSpec(256, 256, seed=11, sites=100),
    W.ImageSpec(257, 131, seed=12, sites=700, shading=True),           # > one LDS chunk
    W.ImageSpec(300, 200, seed=13, sites=60, shading=True, noise_sigma=0.01),
    W.ImageSpec(64, 999, seed=14, sites=5, noise_uniform=0.05),
    W.ImageSpec(1, 77, seed=15, sites=3, shading=True, octaves=3, noise_sigma=0.02),
    W.ImageSpec(1024, 768, seed=16, sites=1500, shading=True, noise_sigma=0.01),
]


def test_gen_matches_host(torch_cuda):
    torch = torch_cuda
    say("device generator vs host generator")
    outs = [torch.empty((3, s.width * s.height), dtype=torch.float32, device="cuda")
            for s in SPECS]
    sa.gen_images(SPECS, outs)
    torch.cuda.synchronize()
    for s, o in zip(SPECS, outs):
        got = o.cpu().numpy()
        want = W.make_image(s)
        noisy = s.noise_sigma > 0 or s.noise_uniform > 0
        if not noisy:
            assert np.array_equal(got, want), s
        else:
            diff = np.abs(got.astype(np.float64) - want)
            assert diff.max() <= 2.0 ** -23, (s, diff.max())
            assert (got == want).mean() > 0.999, s


def test_gen_errors(torch_cuda):
    torch = torch_cuda
    o = torch.empty((3, 16), dtype=torch.float32, device="cuda")
    with pytest.raises(sa.SplitError):
        sa.gen_images([W.ImageSpec(4, 4, seed=1, sites=0)], [o])
