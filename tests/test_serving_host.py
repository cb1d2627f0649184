"""Server.submit's validation (host only, no GPU): requests that no server state could run
fail at submit instead of aborting the in-flight batch later."""
import numpy as np
import pytest


def _req(Request, **kw):
    base = dict(rid=1, f0=np.zeros((16, 16)), t_final=0.5, e=0.9, n=16, nr=16, m=8, s=3.0)
    base.update(kw)
    return Request(**base)


def test_unknown_integrator_rejected(pkg):
    from proj_int_inelastic_amd import serving
    srv = serving.Server()
    with pytest.raises(ValueError, match="unknown time integrator"):
        srv.submit(_req(serving.Request, integrator="RK4"))
    assert srv.queue == []


def test_tpi3_needs_three_levels(pkg):
    from proj_int_inelastic_amd import serving
    srv = serving.Server()
    with pytest.raises(ValueError, match="3 inner levels"):
        srv.submit(_req(serving.Request, integrator="TPI3", inner_steps=(2, 2), inner_dt=(0.01, 0.05)))
    with pytest.raises(ValueError, match="inner levels"):
        srv.submit(_req(serving.Request, integrator="PFE", inner_steps=(0,), inner_dt=(0.01,)))
    srv.submit(_req(serving.Request, integrator="TPI3", inner_steps=(3, 3, 3), inner_dt=(0.01, 0.05, 0.1)))
    assert len(srv.queue) == 1


def test_step_must_divide_final_time(pkg):
    from proj_int_inelastic_amd import serving
    with pytest.raises(Exception, match="dt does not divide"):
        serving.Server().submit(_req(serving.Request, t_final=0.3, dt=0.25))
