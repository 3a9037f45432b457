"""The C ABI from a plain C program: tests/c_abi/solve_fixture.c, compiled
against include/qiotk.h and include/qmc.h and linked with the in-tree
libraries, solves the reference's fixtures (data/ising1.json -> -10,
isingemptymixed.json -> -100 with {"1": -1, "10": 1}, puboemptymixed.json ->
-104, params1.json) through qiotk_solve_file; every returned configuration
has the returned cost under the oracle's cost function."""
import json
import os
import subprocess

import pytest

from tests import _c_abi
from tests.test_gpu_solvers import oracle_cost

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return _c_abi.build(str(tmp_path_factory.mktemp("c_abi")))


@pytest.mark.parametrize("fname,cost,config", [("ising1.json", -10, None),
                                               ("isingemptymixed.json", -100, {"1": -1, "10": 1}),
                                               ("puboemptymixed.json", -104, {"1": 1, "12": 1})])
@pytest.mark.parametrize("target", ["simulatedannealing.qiotoolkit", "paralleltempering.qiotoolkit",
                                    "populationannealing.cpu"])
def test_c_program_solves_reference_fixture(exe, fname, cost, config, target):
    path = os.path.join(GOLD, fname)
    r = subprocess.run([exe, target, os.path.join(GOLD, "params1.json"), path], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    doc = json.loads(r.stdout)
    sol = doc["solutions"]
    assert sol["cost"] == cost
    for k, v in (config or {}).items():
        assert sol["configuration"][k] == v
    fixture = json.load(open(path))
    assert oracle_cost(fixture, sol["configuration"]) == cost
    assert "benchmark" in doc
