"""CPU checks of the C ABI's argument validation and of the Python error mapping.

Every call here is rejected before any HIP call, so no GPU is needed: the
status codes, ib_last_error() messages and the exception types the package
raises at the reference's call points (ValueError, NotImplementedError,
Exception for a missing boundary condition).
"""
import ctypes

import numpy as np
import pytest

FAKE = ctypes.c_void_p(16)  # never dereferenced: every call below fails validation first


def _desc(pkg, **kw):
    L = pkg._lib
    d = L.ib_plan_desc()
    d.dim, d.n, d.nr, d.m, d.L, d.gamma, d.precision, d.n_param = 2, 16, 4, 4, 5.0, 0.0, 0, 1
    for k, v in kw.items():
        setattr(d, k, v)
    arr = np.linspace(0.1, 1.0, 16)
    p = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d.r = d.wr = d.sigma = d.wsigma = d.e_r = p
    return d, arr


@pytest.mark.parametrize("field,value,code,msg", [
    ("dim", 4, -1, "Dimension must be 2 or 3."),
    ("n", 12, -3, "n must be"),
    ("n", 64, -3, "n must be"),  # with dim = 3 below
    ("nr", 0, -1, "nr, m, n_param"),
    ("precision", 7, -1, "precision must be"),
])
def test_plan_create_rejects(pkg, field, value, code, msg):
    lib = pkg._lib.load()
    kw = {field: value}
    if field == "n" and value == 64:
        kw["dim"] = 3
    d, keep = _desc(pkg, **kw)
    h = ctypes.c_void_p()
    rc = lib.ib_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == code
    assert msg in pkg._lib.last_error()
    assert h.value is None


def test_plan_create_null_arrays(pkg):
    lib = pkg._lib.load()
    d, keep = _desc(pkg)
    d.e_r = None
    h = ctypes.c_void_p()
    assert lib.ib_plan_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert "null quadrature array" in pkg._lib.last_error()
    assert lib.ib_plan_create(None, ctypes.byref(h)) == -1


def test_transport_entry_points_reject(pkg):
    lib = pkg._lib.load()
    st = None
    # missing / unknown boundary conditions (solver.py:80-87, 100-107)
    assert lib.ib_tvd1d(FAKE, FAKE, 10, 16, FAKE, 0.1, 2, 0, 0, 1, 4, st) == -4
    assert pkg._lib.last_error() == "Lower boundary condition not specified."
    assert lib.ib_tvd1d(FAKE, FAKE, 10, 16, FAKE, 0.1, 2, 0, 1, 0, 4, st) == -4
    assert pkg._lib.last_error() == "Upper boundary condition not specified."
    assert lib.ib_tvd1d(FAKE, FAKE, 10, 16, FAKE, 0.1, 2, 0, 3, 1, 4, st) == -3
    assert lib.ib_tvd1d(FAKE, FAKE, 10, 16, FAKE, 0.1, 3, 0, 1, 1, 4, st) == -1
    assert lib.ib_tvd1d(FAKE, FAKE, 10, 16, FAKE, 0.1, 2, 5, 1, 1, 4, st) == -1
    # the reference's s[LL-1:UL-1] slice fails to broadcast when nx + 2 < N
    assert lib.ib_tvd1d(FAKE, FAKE, 10, 16, FAKE, 0.1, 2, 0, 1, 1, 32, st) == -1
    assert "broadcast" in pkg._lib.last_error()
    assert lib.ib_tvd_axis(FAKE, FAKE, 1, 10, 16, FAKE, 16, 0.1, 2, 0, 0, 1, 0, st) == -4
    assert lib.ib_tvd_axis(FAKE, FAKE, 0, 10, 16, FAKE, 16, 0.1, 2, 0, 1, 1, 0, st) == -1
    assert lib.ib_upwind_flux(FAKE, FAKE, FAKE, 10, 4, 0.1, 0.01, 4, st) == -1
    assert lib.ib_upwind_flux(FAKE, FAKE, FAKE, 3, 4, 0.1, 0.01, 1, st) == -1
    assert "at least 4 rows" in pkg._lib.last_error()
    assert lib.ib_upwind_flux(FAKE, FAKE, FAKE, 10, 0, 0.1, 0.01, 1, st) == 0  # nothing to do
    assert lib.ib_fast_sweep(FAKE, FAKE, FAKE, FAKE, 0, 4, 4, 0.1, 1.0, FAKE, st) == -1
    assert lib.ib_moments(FAKE, FAKE, 3, 4, 16, 5.0, st) == -1
    assert lib.ib_moments(FAKE, FAKE, 3, 2, 300, 5.0, st) == -1
    assert lib.ib_moments(None, None, 0, 2, 16, 5.0, st) == 0
    assert lib.ib_project_conservative(FAKE, -1, 2, 16, 5.0, st) == -1
    assert lib.ib_axpy(FAKE, FAKE, 1.0, -1, st) == -1
    assert lib.ib_axpy(None, None, 1.0, 0, st) == 0


def test_collision_entry_points_reject(pkg):
    lib = pkg._lib.load()
    assert lib.ib_collide(None, FAKE, FAKE, 1, None, 1, 0.0, None) == -1
    assert lib.ib_plan_set_tuning(None, 0, 0, 0) == -1
    assert lib.ib_plan_set_timing(None, 1) == -1
    assert lib.ib_collide_group(0, None, None, None, None, None, 1, 0.0, None) == -1
    e_r = np.ones(4)
    assert lib.ib_plan_set_params(None, 1, e_r.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  None, None) == -1
    assert lib.ib_plan_destroy(None) == 0


def test_group_mode_rejects(pkg):
    lib = pkg._lib.load()
    assert lib.ib_set_group_mode(2) == -1
    assert b"group mode" in lib.ib_last_error()
    assert lib.ib_set_group_mode(-1) == -1


@pytest.mark.parametrize("split,rounds", [(0, 0), (65, 0), (1, -1), (1, 1025)])
def test_tile_shape_rejects(pkg, split, rounds):
    """Tile shapes outside [1, 64] splits and [0, 1024] rounds are refused before the plan
    is touched (the fake plan pointer is never dereferenced)."""
    lib = pkg._lib.load()
    assert lib.ib_plan_set_tiles(FAKE, split, rounds) == -1
    assert b"tile shape" in lib.ib_last_error()
    assert lib.ib_plan_set_tiles(None, 1, 0) == -1
    a, r = ctypes.c_int(), ctypes.c_int()
    assert lib.ib_plan_get_tiles(None, ctypes.byref(a), ctypes.byref(r)) == -1
    assert lib.ib_plan_get_tiles(FAKE, None, ctypes.byref(r)) == -1


def test_status_to_exception_mapping(pkg):
    L = pkg._lib
    L.check(L.IB_OK)
    with pytest.raises(ValueError):
        L.check(L.IB_ERR_ARG, "x")
    with pytest.raises(NotImplementedError):
        L.check(L.IB_ERR_UNSUPPORTED, "x")
    with pytest.raises(Exception) as e:
        L.check(L.IB_ERR_BC, "x")
    assert type(e.value) is Exception
    with pytest.raises(RuntimeError):
        L.check(L.IB_ERR_HIP, "x")


def test_python_layer_rejects_before_the_device(pkg):
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision
    from proj_int_inelastic_amd.riemann import upwind
    from proj_int_inelastic_amd.solver import Solver
    with pytest.raises(ValueError):
        FSInelasticVHSCollision({"collision-model": {"dim": 2}}, None, precision="fp16")
    with pytest.raises(NotImplementedError):
        upwind.F_p_2(np.zeros((8, 2)), np.zeros((6, 2)), 0.1, 0.01, limiter=lambda r: r)

    class S(Solver):
        num_dim_x = 1
    s = S()
    with pytest.raises(Exception, match="Lower boundary condition not specified."):
        s._bc_codes(0)
    s.bc_lower[0] = 1
    with pytest.raises(Exception, match="Upper boundary condition not specified."):
        s._bc_codes(0)
    s.bc_upper[0] = 5
    with pytest.raises(NotImplementedError):
        s._bc_codes(0)


def test_server_rejects_requests_it_could_never_admit(pkg):
    """A request above the whole HBM budget or the per-class cell cap, or whose dt does
    not divide t_final, raises at submit instead of waiting in the queue forever."""
    from proj_int_inelastic_amd import serving
    f0 = np.zeros((16, 16))
    ok = serving.Request(rid=0, f0=f0, t_final=0.5, e=0.9, n=16, nr=16, dt=0.25)
    job = serving.Request(rid=1, f0=np.zeros((4, 16, 16)), t_final=0.5, e=0.9, n=16, nr=16,
                          dt=0.25, xlen=1.0)
    srv = serving.Server(hbm_budget=job.nbytes(), max_cells_per_class=1)
    srv.submit(ok)
    assert len(srv.queue) == 1
    with pytest.raises(ValueError, match="HBM budget"):
        serving.Server(hbm_budget=ok.nbytes() - 1).submit(ok)
    with pytest.raises(ValueError, match="max_cells_per_class"):
        srv.submit(job)
    with pytest.raises(Exception, match="dt does not divide"):
        srv.submit(serving.Request(rid=2, f0=f0, t_final=0.6, e=0.9, n=16, nr=16, dt=0.25))
    assert [r.rid for r in srv.queue] == [0]


def test_probe_library_rejects_bad_arguments():
    """The FMA-throughput probe the benchmark loads validates before any HIP call."""
    import pathlib
    so = pathlib.Path(__file__).resolve().parents[1] / "proj-int-inelastic_amd" / "libinelastic_probe.so"
    if not so.exists():
        pytest.skip("probe library not built")
    lib = ctypes.CDLL(str(so))
    fn = lib.ibp_fma_peak
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    fl, ms = ctypes.c_double(), ctypes.c_double()
    assert fn(0, 0, 1, 1, ctypes.byref(fl), ctypes.byref(ms)) == -1
    assert fn(0, 16, 0, 1, ctypes.byref(fl), ctypes.byref(ms)) == -1
    assert fn(2, 16, 1, 1, ctypes.byref(fl), ctypes.byref(ms)) == -1
    assert fn(0, 16, 1, 1, None, ctypes.byref(ms)) == -1
