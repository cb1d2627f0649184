"""GPU checks of ib_collide_static, the entry point that allocates nothing: its scratch is a
static device region, so it must give the results of ib_collide (the same kernels) on
consecutive calls of decreasing size, reject batches whose scratch exceeds the region with a
ValueError before any launch, and leave device memory untouched."""
import ctypes

import numpy as np
import pytest

from test_gpu_collide import TOL64, _cfg, _data, _err, _smooth

pytestmark = pytest.mark.gpu

S3 = 3.624654714576

CASES = [
    # d, N, nr, M, s, precision, first batch, second batch, bar against ib_collide
    (2, 32, 32, 8, S3, "fp64", 64, 20, TOL64),   # fused FP64 2D 32^2 (the 1D x 2D cell shape)
    (3, 16, 8, 14, 3.0, "fp64", 48, 5, TOL64),   # staged FP64 3D 16^3, split pair sums at 5 cells
    (3, 32, 4, 6, S3, "fp32", 8, 3, 1e-6),       # fused FP32 3D 32^3 with the Nyquist-plane term
    (2, 64, 16, 8, 3.0, "fp32", 40, 7, 1e-6),    # fused FP32 2D 64^2
    (2, 32, 32, 8, S3, "fp32", 64, 20, 1e-6),    # real-data FP32 2D 32^2 with the Nyquist-line term
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "d%dN%d%s" % (c[0], c[1], c[5]))
def test_static_matches_allocating_entry_point(pkg, need_gpu, case):
    import torch
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    d, N, nr, M, s, prec, nb1, nb2, tol = case
    cfg = _cfg(d, N, nr, M, s, 0.8, 0.8, 1.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm, precision=prec)
    rng = np.random.default_rng(N + nb1)
    f1 = torch.from_numpy(_data(d, N, sm.L, nb1, seed=1)).cuda()
    f2 = torch.from_numpy(_smooth(d, N, sm.L, nb2, seed=2)).cuda()
    e1, e2 = rng.uniform(0.3, 1.0, nb1), rng.uniform(0.3, 1.0, nb2)
    # two static calls in a row, the second smaller, before any allocating call on the plan
    qs1 = coll.collide(f1, with_loss=True, e=e1, static_scratch=True).cpu().numpy()
    qs2 = coll.collide(f2, with_loss=True, heat_bath=0.05, e=e2, static_scratch=True).cpu().numpy()
    qa1 = coll.collide(f1, with_loss=True, e=e1).cpu().numpy()
    qa2 = coll.collide(f2, with_loss=True, heat_bath=0.05, e=e2).cpu().numpy()
    w1 = max(_err(qs1[c], qa1[c]) for c in range(nb1))
    w2 = max(_err(qs2[c], qa2[c]) for c in range(nb2))
    print(case, "static vs ib_collide", w1, w2)
    assert w1 <= tol and w2 <= tol, (w1, w2)
    # and again after the allocating calls (the plan's own buffers untouched by the static ones)
    qs3 = coll.collide(f2, with_loss=True, heat_bath=0.05, e=e2, static_scratch=True).cpu().numpy()
    assert max(_err(qs3[c], qa2[c]) for c in range(nb2)) <= tol


def test_static_rejects_oversized_batch_and_recovers(pkg, need_gpu):
    import torch
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    cfg = _cfg(3, 32, 32, 38, S3, 0.8, 0.8, 1.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm, precision="fp32")
    big = torch.from_numpy(_smooth(3, 32, sm.L, 40, seed=3)).cuda()  # ~7 MiB of scratch per cell
    with pytest.raises(ValueError, match="static region"):
        coll.collide(big, with_loss=True, static_scratch=True)
    q = coll.collide(big[:2], with_loss=True, static_scratch=True).cpu().numpy()
    qa = coll.collide(big[:2], with_loss=True).cpu().numpy()
    assert max(_err(q[c], qa[c]) for c in range(2)) <= 1e-6


def test_static_call_allocates_no_device_memory(pkg, need_gpu):
    """A fresh plan's first static call (about 27 MB of scratch at 4 cells of 3D 32^3 FP32, Nr = 32)
    leaves the device's free memory where it was; ib_collide on the same plan allocates."""
    import torch
    from proj_int_inelastic_amd import _lib
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    lib = _lib.load()
    cfg = _cfg(3, 32, 32, 14, S3, 0.8, 0.8, 1.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm, precision="fp32")
    f = torch.from_numpy(_smooth(3, 32, sm.L, 4, seed=4)).cuda()
    qs = torch.empty_like(f)
    qa = torch.empty_like(f)
    st = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    _lib.check(lib.ib_collide_static(coll._plan, f.data_ptr(), qs.data_ptr(), 4, None, 1, 0.0, st),
               "ib_collide_static")
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    _lib.check(lib.ib_collide(coll._plan, f.data_ptr(), qa.data_ptr(), 4, None, 1, 0.0, st),
               "ib_collide")
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    print("free memory drop: static call %d B, allocating call %d B" % (free0 - free1, free1 - free2))
    assert free0 - free1 < (4 << 20), free0 - free1
    assert free1 - free2 > (16 << 20), free1 - free2
    assert torch.equal(qs, qa) or float((qs - qa).abs().max() / qa.abs().max()) <= 1e-6
