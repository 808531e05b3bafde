"""The reference's own regression inputs (its proptest-regressions files):
the product's tree equals the oracle's, the oracle's solutions are well
formed (crates/solver/src/proptests.rs:224-259), and on the GPU the exact
engine reproduces the oracle bit for bit (the reference's GPU-vs-CPU engine
property, crates/solver/src/gpu/engine/tests.rs)."""
import numpy as np
import pytest

import pyoracle as O
from pts.native import EXACT, FAST
from pts.solver import Solver
from regression_cases import all_configs
from test_host_cpu import host_solver, product_nodes

CASES = all_configs()
IDS = [n for n, _ in CASES]


@pytest.mark.parametrize("name,cfg", CASES, ids=IDS)
def test_regression_tree_equals_the_oracle(name, cfg):
    from pts.native import lib
    o = O.OracleSolver(cfg)
    h = host_solver(cfg)
    try:
        a = product_nodes(h)
        b = o.nodes()
        for k in ("kind", "street", "num_children", "first_child", "offset"):
            assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(a["invest"], b["invest"])
        assert np.array_equal(a["rake"], b["rake"]) and np.array_equal(a["chance_factor"], b["chance_factor"])
    finally:
        lib().pts_solver_free(h)


@pytest.mark.parametrize("name,cfg", CASES, ids=IDS)
def test_regression_oracle_solutions_are_well_formed(name, cfg):
    o = O.OracleSolver(cfg)
    o.iterate(30)
    pot = cfg.root.pot
    assert o.exploitability_f64() >= -1e-3 * pot
    v = o.root_values_f64(2) if hasattr(o, "root_values_f64") else None
    if v is not None:
        assert abs(v[0] + v[1]) <= 1e-3 * pot


@pytest.mark.gpu
@pytest.mark.parametrize("name,cfg", CASES, ids=IDS)
def test_regression_gpu_bit_exact(name, cfg):
    g, c = Solver(cfg, precision=EXACT), O.OracleSolver(cfg)
    g.iterate(25)
    c.iterate(25)
    rg, sg = g.storage()
    rc, sc = c.storage()
    assert np.array_equal(rg, rc) and np.array_equal(sg, sc)
    assert g.exploitability() == c.exploitability_f64()


@pytest.mark.gpu
@pytest.mark.parametrize("name,cfg", CASES, ids=IDS)
def test_regression_gpu_fast_close(name, cfg):
    g, c = Solver(cfg, precision=FAST), O.OracleSolver(cfg)
    g.iterate(25)
    c.iterate(25)
    assert abs(g.exploitability() - c.exploitability_f64()) < 1e-4 * cfg.root.pot
