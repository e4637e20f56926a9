"""The Kulla-Conty tables the library uploads (host code, Microfacet.h:231-291)
equal the oracle's restatement bit for bit, and have the properties the
construction fixes: every directional albedo lies in [0, 1] (it is clamped,
Microfacet.h:252), the average albedo is not clamped and its 128-sample
quadrature overshoots 1 by a few 1e-3 at the smooth end, it decreases with
roughness beyond that end, and the smoothest row is ~1.
CPU only: the table is computed on the host."""
import ctypes as C

import numpy as np

import oracle_lib
from optixrender_amd import abi


def product_tables():
    b = np.zeros(64 * 64, np.float32)
    a = np.zeros(64, np.float32)
    assert abi.lib().oxr_kulla_conty(abi.ptr(b), abi.ptr(a)) == 0
    return b.reshape(64, 64), a


def test_tables_bit_exact_with_oracle():
    pb, pa = product_tables()
    ob, oa = oracle_lib.kulla_conty()
    assert np.array_equal(pb.view(np.uint32), ob.view(np.uint32))
    assert np.array_equal(pa.view(np.uint32), oa.view(np.uint32))


def test_table_properties():
    b, a = oracle_lib.kulla_conty()
    assert np.all((b >= 0) & (b <= 1))
    assert np.all((a >= 0) & (a <= 1.005))
    assert b[0, 32:].min() > 0.97          # alpha ~ 0.008: almost no energy loss
    assert a[-1] < a[0]
    assert np.all(np.diff(a[8:]) <= 1e-6)  # monotone loss with roughness
