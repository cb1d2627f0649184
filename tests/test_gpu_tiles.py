"""Tile shapes of the fused engines (ib_plan_set_tiles): every split of the angle loop and
every round count gives the reference operator within the precision's bar, a fixed tile
shape keeps a cell's result bit-identical in any batch, and the autotuner's tile candidates
reach the plan.

Bars (relative L-infinity against the oracle's max |Q|): FP64 forms <= 1e-12, FP32 forms
<= 1e-5.
"""
import numpy as np
import pytest

from test_gpu_collide import TOL32, TOL64, _cfg, _data, _err, _oracle

pytestmark = pytest.mark.gpu

# d, N, nr, M, s, precision: one case per engine with a tile knob
ENGINES = [
    (3, 32, 4, 14, 3.624654714576, "fp32"),   # k_fused3 + k_f3_nyq (real-data form)
    (3, 32, 3, 14, 3.624654714576, "fp32k"),  # the same, Kahan angular sums
    (3, 16, 4, 14, 3.0, "fp32"),              # k_fusedc3
    (2, 64, 6, 8, 3.0, "fp32"),               # k_fusedw<float, 64>
    (2, 32, 8, 8, 3.624654714576, "fp64"),    # k_fusedw<double, 32>
]


def _coll(d, N, nr, M, s, prec):
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    cfg = _cfg(d, N, nr, M, s, 1.0, 1.0, 1.0)
    sm = SpectralMesh(cfg)
    return FSInelasticVHSCollision(cfg, sm, precision=prec), sm


@pytest.mark.parametrize("case", ENGINES, ids=lambda c: "d%dN%d_%s" % (c[0], c[1], c[5]))
@pytest.mark.parametrize("split", [2, 3, 64])
def test_angle_split_matches_oracle(pkg, need_gpu, case, split):
    """Splits that divide M, that leave ragged ranges (3) and more splits than angles (64,
    acting as M): per-cell e, grid-scale noise (Nyquist content), gain + loss."""
    import torch
    d, N, nr, M, s, prec = case
    coll, sm = _coll(*case)
    coll.set_tiles(angle_split=split)
    assert coll.tiles()[0] == min(split, M)
    es = np.array([0.55, 0.95, 0.7])
    f = _data(d, N, sm.L, 3, seed=11)
    q = coll.collide(torch.from_numpy(f).cuda(), with_loss=True, e=es).cpu().numpy()
    tol = TOL64 if prec == "fp64" else TOL32
    for c in range(3):
        ref = _oracle(d, N, nr, M, s, es[c], es[c], 1.0).collide(f[c], with_loss=True)
        assert _err(q[c], ref) <= tol, (split, c, _err(q[c], ref))
    # the default tile shape agrees to the precision's rounding
    coll.set_tiles(angle_split=1)
    q1 = coll.collide(torch.from_numpy(f).cuda(), with_loss=True, e=es).cpu().numpy()
    assert _err(q, q1) <= (1e-14 if prec == "fp64" else 2e-6), _err(q, q1)


@pytest.mark.parametrize("case", ENGINES[:1] + ENGINES[2:], ids=lambda c: "d%dN%d_%s" % (c[0], c[1], c[5]))
def test_split_batch_invariance(pkg, need_gpu, case):
    """With a fixed split, 20 cells made of 4 distinct cells give bit-identical copies, equal
    to each distinct cell alone."""
    import torch
    d, N, nr, M, s, prec = case
    coll, sm = _coll(*case)
    coll.set_tiles(angle_split=4)
    base = _data(d, N, sm.L, 4, seed=5)
    eb = np.array([0.6, 0.8, 0.9, 1.0])
    idx = np.arange(20) % 4
    q = coll.collide(torch.from_numpy(base[idx]).cuda(), with_loss=True, e=eb[idx]).cpu().numpy()
    for c in range(4):
        copies = q[idx == c]
        assert np.array_equal(copies, np.broadcast_to(copies[0], copies.shape)), c
        alone = coll.collide(torch.from_numpy(base[c:c + 1]).cuda(), with_loss=True,
                             e=eb[c:c + 1]).cpu().numpy()
        assert np.array_equal(alone[0], copies[0]), c


@pytest.mark.parametrize("N,nr", [(16, 8), (32, 16)])
@pytest.mark.parametrize("rounds", [1, 4, 32])
def test_pair_rounds_match_oracle(pkg, need_gpu, N, nr, rounds):
    """Rounds per workgroup of the FP32 2D 16^2 / 32^2 form (more or fewer workgroups per cell)."""
    import torch
    coll, sm = _coll(2, N, nr, 8, 3.0, "fp32")
    coll.set_tiles(angle_split=1, pair_rounds=rounds)
    assert coll.tiles() == (1, rounds)
    es = np.array([0.35, 0.85])
    f = _data(2, N, sm.L, 2, seed=21)
    q = coll.collide(torch.from_numpy(f).cuda(), with_loss=True, e=es).cpu().numpy()
    for c in range(2):
        ref = _oracle(2, N, nr, 8, 3.0, es[c], es[c], 1.0).collide(f[c], with_loss=True)
        assert _err(q[c], ref) <= TOL32, (rounds, c, _err(q[c], ref))


def test_tile_shape_reaches_static_entry(pkg, need_gpu):
    """ib_collide_static runs the split form too (its scratch grows with the partials)."""
    import torch
    d, N, nr, M, s, prec = 3, 16, 4, 14, 3.0, "fp32"
    coll, sm = _coll(d, N, nr, M, s, prec)
    coll.set_tiles(angle_split=7)
    f = torch.from_numpy(_data(d, N, sm.L, 2, seed=2)).cuda()
    a = coll.collide(f, with_loss=True).cpu().numpy()
    b = coll.collide(f, with_loss=True, static_scratch=True).cpu().numpy()
    assert np.array_equal(a, b)


def test_autotuner_tries_tiles_on_small_batches(pkg, need_gpu, tmp_path, monkeypatch):
    """A one-cell 3D 32^3 batch gets split candidates; the chosen setting is applied."""
    import torch
    from proj_int_inelastic_amd import autotune
    monkeypatch.setenv("INELASTIC_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_MEM", {})
    coll, sm = _coll(3, 32, 4, 14, 3.624654714576, "fp32")
    cands = autotune.candidates(coll, 1)
    assert {c.get("angle_split", 1) for c in cands} == {1, 2, 4, 8}
    f = torch.from_numpy(_data(3, 32, sm.L, 1, seed=4)).cuda()
    best = autotune.tune(coll, f)
    assert best in cands
    assert coll.tiles()[0] == best.get("angle_split", 1)
    big = autotune.candidates(coll, 4096)
    assert {c.get("angle_split", 1) for c in big} == {1}
