"""Each planted mistake of the oracle (a build with -DORC_MUTATION=k, listed in dem_oracle.c) fails
at least one physics pin: the pins can see a swapped stiffness, a flipped sign, a dropped factor,
a transposed rotation, a wrong lever arm and the other mistakes listed there."""
import pytest

from oracle import oracle as O
from oracle_pins import PINS

MUTATIONS = {
    1: "k5/k6 swapped in the moment on i",
    2: "sign of the k3 term of the bond shear force",
    3: "sqrt(5/6) dropped from the Hertz damping",
    4: "Hertz contact moment F x r",
    5: "torsional shear stress r / I",
    6: "sign of global damping",
    7: "bond forces scattered with R instead of R^-1",
    8: "Hertz contact point at r + delta",
    9: "sign of the Hertz contact velocity",
    10: "bond normal stress from F_i",
    11: "quaternion increment scaled by dt",
    12: "bond torsion stiffness without 1 / (1 + nu)",
    13: "position advanced with the new velocity",
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


@pytest.mark.parametrize("k", sorted(MUTATIONS), ids=lambda k: f"mut{k}")
def test_planted_mistake_fails_a_pin(k, mutant_dir):
    lib = O.load_path(O.build_mutant(k, mutant_dir))
    caught = []
    for pin in PINS:
        try:
            pin(lib)
        except (AssertionError, RuntimeError, FloatingPointError):
            caught.append(pin.__name__)
    print(f"[mutation {k}: {MUTATIONS[k]}] caught by {caught}", flush=True)
    assert caught, f"no pin rejects mutation {k} ({MUTATIONS[k]})"
