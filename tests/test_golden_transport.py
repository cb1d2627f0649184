"""Oracle transport step against golden vectors written from the reference's
own modules (tests/golden/make_transport_golden.py) — CPU only, bit-exact."""
import pathlib

import numpy as np

from oracle import transport as T

G = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "transport_golden.npz")


def test_oracle_df_advection_bit_exact():
    f, v1, dx = G["f"], G["v1"], float(G["dx"][0])
    for c in range(8):
        bc, order, dt = G["case%d_params" % c]
        df = T.df_advection(f, v1, dt, dx, int(bc), int(bc), int(order), T.minmod)
        assert np.array_equal(df, G["case%d_df" % c]), c


def test_oracle_limit_bit_exact():
    out = T.limit(G["limit_wave"], G["limit_s"], T.minmod, 0.1)
    assert np.array_equal(out, G["limit_out"])


def test_oracle_flux_form_upwind_bit_exact():
    """oracle.transport restatement of src/riemann/upwind.py against its golden vectors."""
    f, v, v1 = G["up_f"], G["up_v"], G["up_v1"]
    dx, dt = G["up_dxdt"]
    assert np.array_equal(T.grad_vf2(f, v, dx, dt), G["up_grad_vf2"])
    assert np.array_equal(T.grad_vf2(f, v1, dx, dt), G["up_grad_vf2_b"])
    assert np.array_equal(T.grad_vf1(f, v), G["up_grad_vf1"])
    vp, vm = np.maximum(v, 0.0), np.minimum(v, 0.0)
    assert np.array_equal(T.F_p_2(f, vp[1:-1], dx, dt), G["up_F_p_2"])
    assert np.array_equal(T.F_m_2(f, vm[1:-1], dx, dt), G["up_F_m_2"])
