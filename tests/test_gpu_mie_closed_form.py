"""The library's Chem_Mie query and tracer sum (aa_mie_query,
aa_get_aop_vector on the GPU) against the closed forms of
tests/_mie_closed_form.py: node values, RH-linear tables, one- and
two-tracer sums."""
import numpy as np
import pytest

from aeroapps_amd import chem_mie as AM
import _mie_closed_form as MC

pytestmark = pytest.mark.gpu


def lib_query(t, q, rh):
    mie = AM.MieTables.from_synthetic([t])
    try:
        o = mie.query(0, 1.0, q, rh, nmom=MC.NMOM, npol=1)
    finally:
        mie.close()
    return dict(tau=o["tau"], ssa=o["ssa"], g=o["gasym"], pmom=o["pmom"][:, :, 0])


def lib_get_aop(tables, vnames, qm, rh):
    mie = AM.MieTables.from_synthetic(tables)
    try:
        tau, ssa, g, pm, rc = AM.get_aop_vector(mie, (550.0,), vnames, 0, qm.T, rh.T, MC.NMOM, 1)
    finally:
        mie.close()
    assert rc == 0
    return tau.T, ssa.T, g.T, pm.T


def test_query_returns_node_values(gpu):
    print("case mie closed form: node values", flush=True)
    MC.check_node_values(lib_query)


def test_query_linear_table(gpu):
    print("case mie closed form: linear table", flush=True)
    MC.check_linear_table(lib_query)


def test_get_aop_single_tracer(gpu):
    print("case mie closed form: one tracer", flush=True)
    MC.check_single_tracer(lib_get_aop)


def test_get_aop_two_tracers_weighted_means(gpu):
    print("case mie closed form: two tracers", flush=True)
    MC.check_two_tracers(lib_get_aop)
