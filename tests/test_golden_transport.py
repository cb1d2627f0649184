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
