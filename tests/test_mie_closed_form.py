"""The Chem_Mie oracle (oracle/chem_mie.py) against closed forms of the
reference's algorithm (tests/_mie_closed_form.py): node values at RH nodes,
RH-linear tables, one-tracer and two-tracer optics sums."""
import numpy as np

from oracle import chem_mie as CM
import _mie_closed_form as MC


def oracle_query(t, q, rh):
    rhi, rha = CM.rh_map(t.rh)
    tau, ssa, g, pm = CM.query(t.bext, t.bsca, t.g, t.pmom, rhi, rha, 0, 0, q, rh)
    return dict(tau=tau, ssa=ssa, g=g, pmom=pm[..., 0])


def oracle_get_aop(tables, vnames, qm, rh):
    tau, ssa, g, pm, rc = CM.get_aop_vector(tables, (550.0,), tables[0].lambda_m, vnames, qm, rh, MC.NMOM, 1)
    assert rc == 0
    return tau, ssa, g, pm


def test_query_returns_node_values():
    MC.check_node_values(oracle_query)


def test_query_linear_table():
    MC.check_linear_table(oracle_query)


def test_get_aop_single_tracer():
    MC.check_single_tracer(oracle_get_aop)


def test_get_aop_two_tracers_weighted_means():
    MC.check_two_tracers(oracle_get_aop)
