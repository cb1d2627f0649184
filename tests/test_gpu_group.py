"""The grouped launch of the fused FP32 pair loops (ib_collide_group, ib_set_group_mode):
one kernel launch runs the pair loops of every fused grid shape (3D 32^3 real-data form,
3D 16^3, 2D 64^2, 2D 32^2, 2D 16^2), each shape's workgroups as segments of the launch.
Its results are bit-identical to one fused launch per plan (the same per-(cell, partial)
work and the same FP64 reduce order), for every tile shape and for batches on both sides
of the 8-cell workgroup-numbering switch, and match the FP64 oracle within the FP32 bar.
"""
import numpy as np
import pytest

from test_gpu_collide import TOL32, _cfg, _data, _err, _oracle

pytestmark = pytest.mark.gpu

# d, N, nr, M, s, cells
SHAPES = [(3, 32, 4, 14, 3.624654714576, 3), (3, 16, 4, 14, 3.0, 5), (2, 64, 6, 8, 3.0, 4),
          (2, 32, 8, 8, 3.0, 9), (2, 16, 8, 8, 2.5, 7)]


def _setup(prec, shapes, seed=3):
    import torch
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    rng = np.random.default_rng(seed)
    colls, fs, es, gs = [], [], [], []
    for d, N, nr, M, s, nb in shapes:
        cfg = _cfg(d, N, nr, M, s, 1.0, 1.0, 0.0)
        sm = SpectralMesh(cfg)
        colls.append(FSInelasticVHSCollision(cfg, sm, precision=prec))
        fs.append(torch.from_numpy(_data(d, N, sm.L, nb, seed=N + d + nb)).cuda())
        es.append(rng.uniform(0.3, 1.0, nb))
        gs.append(rng.integers(0, 2, nb).astype(float))
    return colls, fs, es, gs


def _both_modes(colls, fs, es, gs):
    from proj_int_inelastic_amd.collision.inelastic_collisions import collide_group, set_group_mode
    try:
        set_group_mode(False)
        per_plan = [q.cpu().numpy() for q in collide_group(colls, fs, with_loss=True, es=es, gammas=gs)]
        set_group_mode(True)
        grouped = [q.cpu().numpy() for q in collide_group(colls, fs, with_loss=True, es=es, gammas=gs)]
    finally:
        set_group_mode(True)
    return per_plan, grouped


@pytest.mark.parametrize("prec", ["fp32", "fp32k"])
@pytest.mark.parametrize("tiles", [(1, 0), (3, 4), (8, 1)])
def test_grouped_equals_per_plan_bitwise(pkg, need_gpu, prec, tiles):
    colls, fs, es, gs = _setup(prec, SHAPES)
    for c in colls:
        c.set_tiles(angle_split=tiles[0], pair_rounds=tiles[1])
    per_plan, grouped = _both_modes(colls, fs, es, gs)
    for k, (a, b) in enumerate(zip(per_plan, grouped)):
        assert np.array_equal(a, b), (SHAPES[k], np.abs(a - b).max())


def test_grouped_matches_oracle(pkg, need_gpu):
    """Every segment of one grouped launch against the FP64 oracle (last cell of each shape)."""
    from proj_int_inelastic_amd.collision.inelastic_collisions import collide_group
    colls, fs, es, gs = _setup("fp32", SHAPES, seed=9)
    qs = [q.cpu().numpy() for q in collide_group(colls, fs, with_loss=True, es=es, gammas=gs)]
    for k, (d, N, nr, M, s, nb) in enumerate(SHAPES):
        ref = _oracle(d, N, nr, M, s, es[k][-1], es[k][-1], gs[k][-1]).collide(
            fs[k][-1].cpu().numpy(), with_loss=True)
        assert _err(qs[k][-1], ref) <= TOL32, (SHAPES[k], _err(qs[k][-1], ref))


def test_grouped_large_batches_and_budget_fallback(pkg, need_gpu):
    """Batches of 8 cells and more (XCD-grouped numbering of the 3D 32^3 segment, several
    virtual blocks per workgroup with a ragged last workgroup), and a plan whose whole batch
    exceeds its scratch budget: it leaves the grouped launch and runs in its own
    sub-batches, with the same results."""
    shapes = [(3, 32, 4, 6, 3.624654714576, 19), (3, 16, 4, 6, 3.0, 33), (2, 64, 4, 8, 3.0, 21),
              (2, 32, 8, 8, 3.0, 67), (2, 16, 8, 8, 2.5, 45)]
    colls, fs, es, gs = _setup("fp32", shapes, seed=5)
    colls[1].set_tuning(scratch_budget=4 << 20)  # 33 cells of 3D 16^3 need more than 4 MiB
    per_plan, grouped = _both_modes(colls, fs, es, gs)
    for k, (a, b) in enumerate(zip(per_plan, grouped)):
        assert np.array_equal(a, b), (shapes[k], np.abs(a - b).max())


@pytest.mark.parametrize("with_loss,heat_bath", [(True, 0.3), (False, 0.0), (False, 0.2)])
def test_grouped_equals_per_plan_heat_bath_and_gain_only(pkg, need_gpu, with_loss, heat_bath):
    """The 2D 16^2 / 32^2 forms run their FP64 prologue and epilogue inside their own
    launches per plan (a front kernel, and the last workgroup of each cell), the grouped
    launch in the staged passes: both give the same bits with the heat-bath term and
    without the loss term."""
    from proj_int_inelastic_amd.collision.inelastic_collisions import collide_group, set_group_mode
    shapes = [(2, 32, 8, 8, 3.0, 11), (2, 16, 8, 8, 2.5, 6), (3, 16, 4, 6, 3.0, 2)]
    colls, fs, es, gs = _setup("fp32", shapes, seed=11)
    out = {}
    try:
        for mode in (False, True):
            set_group_mode(mode)
            out[mode] = [q.cpu().numpy() for q in collide_group(
                colls, fs, with_loss=with_loss, heat_bath=heat_bath, es=es, gammas=gs)]
    finally:
        set_group_mode(True)
    for k, (a, b) in enumerate(zip(out[False], out[True])):
        assert np.array_equal(a, b), (shapes[k], np.abs(a - b).max())
