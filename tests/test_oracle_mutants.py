"""Mutation test of the oracle's pins (no GPU).

The oracle is rebuilt with each planted mistake of hydro_oracle.cpp
(ORACLE_MUTANT = k: plausible slips in restating the reference, two of them
the reference's own CPU transliteration's, calculate.cpp:156 and :321-322)
and the pins of test_oracle_pins_kinds.py and test_oracle_pins.py are run
against it.  Every mutant must fail at least one pin; the unmutated oracle
passes them all (the pin modules themselves).
"""
import itertools
import types

import pytest

import test_oracle_pins as P0
import test_oracle_pins_kinds as P1

MUTANTS = {
    1: "branch 4 outflow with the neighbour's depth (calculate.cpp:156)",
    2: "branch 5 pressure over the neighbour's bed",
    3: "ladder tests the neighbour's shallowness before the own",
    4: "KLAS 10 momentum flux with the first guess of the depth",
    5: "KLAS 10 fixed point without relaxation",
    6: "KLAS 1 with the tangential velocity (calculate.cpp:321-322)",
    7: "KLAS 5 pressure scaled by (1 - JL) once",
    8: "KLAS 6 free overflow with the submerged coefficient",
    9: "QD free flow on the downstream head",
    10: "LAQP without interpolation",
    11: "KLAS 7 slot not rewritten when the breach opens",
    12: "branch 1 inflow with the own depth",
}

# pins that take the oracle as their only fixture (plus parametrize)
PINS = [(P1, n) for n in dir(P1) if n.startswith("test_")] + \
       [(P0, n) for n in ("test_lake_at_rest", "test_inflow_mass_budget", "test_wall_flux_is_hydrostatic",
                          "test_level_boundary_equilibrium_and_variants")]


def _cases(fn):
    """Every argument combination of a test's parametrize marks."""
    grids = []
    for mark in getattr(fn, "pytestmark", []):
        if mark.name != "parametrize":
            continue
        names, values = mark.args[0], mark.args[1]
        names = [x.strip() for x in names.split(",")]
        grids.append([dict(zip(names, v if len(names) > 1 else (v,))) for v in values])
    for combo in itertools.product(*grids) if grids else [()]:
        kw = {}
        for d in combo:
            kw.update(d)
        yield kw


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_every_mutant_fails_a_pin(k, oracle, mutant_dir):
    lib = oracle.load_mutant(k, mutant_dir)
    proxy = types.SimpleNamespace(**{n: getattr(oracle, n) for n in dir(oracle) if not n.startswith("__")})
    proxy.OracleRun = lambda mesh, variant=0: oracle.OracleRun(mesh, variant=variant, lib=lib)
    caught = []
    for mod, name in PINS:
        fn = getattr(mod, name)
        for kw in _cases(fn):
            try:
                fn(proxy, **kw)
            except AssertionError:
                caught.append(f"{name}{kw or ''}")
                break
    assert caught, f"mutant {k} ({MUTANTS[k]}) passes every pin"
    print(f"mutant {k} ({MUTANTS[k]}): caught by {', '.join(caught)}")
