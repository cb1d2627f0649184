"""GPU transport kernel (ib_tvd1d) through the C ABI: bit-exact against the
golden vectors of the reference and against the oracle on the 1D x 2D grid."""
import pathlib

import numpy as np
import pytest

from oracle import transport as T

pytestmark = pytest.mark.gpu
G = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "transport_golden.npz")


def _tvd(lib, f, v1, dtdx, order, limiter, lo, hi):
    import torch
    fd = torch.from_numpy(np.ascontiguousarray(f)).cuda()
    vd = torch.from_numpy(np.ascontiguousarray(v1).reshape(-1)).cuda()
    out = torch.empty_like(fd)
    rc = lib.ib_tvd1d(fd.data_ptr(), out.data_ptr(), f.shape[0], f[0].size, vd.data_ptr(), dtdx,
                      order, limiter, lo, hi, f.shape[1], torch.cuda.current_stream().cuda_stream)
    return rc, out.cpu().numpy()


def test_tvd_golden_bit_exact(pkg, need_gpu):
    lib = pkg._lib.load()
    f, v1, dx = G["f"], G["v1"], float(G["dx"][0])
    for c in range(8):
        bc, order, dt = G["case%d_params" % c]
        rc, df = _tvd(lib, f, v1, dt / dx, int(order), 0, int(bc), int(bc))
        assert rc == 0
        assert np.array_equal(df, G["case%d_df" % c]), c


@pytest.mark.parametrize("limiter", [0, 1])
@pytest.mark.parametrize("bc", [1, 2])
def test_tvd_oracle_config_grid(pkg, need_gpu, limiter, bc):
    lib = pkg._lib.load()
    nx, N, L = 200, 32, 8.0
    rng = np.random.default_rng(limiter * 10 + bc)
    dv = 2 * L / N
    v = -L + (np.arange(N) + 0.5) * dv
    v1 = np.repeat(v[:, None], N, axis=1)
    f = rng.random((nx, N, N))
    f[50:80] = 0.25
    dx = 1.0 / nx
    dt = 0.4 * dx / 8
    rc, df = _tvd(lib, f, v1, dt / dx, 2, limiter, bc, bc)
    assert rc == 0
    ref = T.df_advection(f, v1, dt, dx, bc, bc, 2, [T.minmod, T.mc][limiter])
    assert np.array_equal(df, ref)


def test_tvd_errors(pkg, need_gpu):
    lib = pkg._lib.load()
    f = np.zeros((4, 8, 8))
    v1 = np.ones((8, 8))
    rc, _ = _tvd(lib, f, v1, 0.1, 2, 0, 1, 1)
    assert rc == pkg._lib.IB_ERR_ARG  # nx + 2 < N, as the reference's broadcast fails
    rc, _ = _tvd(lib, np.zeros((10, 8, 8)), v1, 0.1, 2, 0, 0, 1)
    assert rc == pkg._lib.IB_ERR_BC
