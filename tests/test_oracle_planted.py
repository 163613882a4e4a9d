"""The closed-form pins have teeth: copies of the oracle built with one
planted mistake each (oracle/Makefile `planted`, ORA_PLANT = 1..5) must fail
the pin that covers it, while the oracle proper passes them all
(tests/test_oracle_closed_forms.py).

  1 sheen lobe doubled                      -> Disney sheen pin
  2 subsurface clamp of cos_i + cos_o dropped -> Disney diffuse/subsurface pin
  3 eta^2 dropped from the glass Jacobian   -> rough-glass pdf pin
  4 emitter-pick uniform not rescaled       -> emitter-picking pin
  5 spotlight falloff linear, not smoothstep -> spotlight NEE pin
"""
import pytest

import oracle
from test_oracle_closed_forms import CHECKS

PLANTED = {1: "disney_sheen", 2: "disney_diffuse_subsurface", 3: "rough_glass_pdf", 4: "emitter_picking",
           5: "spot_nee"}


@pytest.fixture(scope="module")
def planted_libs():
    return oracle.build_planted()


@pytest.mark.parametrize("plant", sorted(PLANTED))
def test_planted_mistake_fails_its_pin(planted_libs, plant):
    with oracle.using_library(planted_libs[plant - 1]):
        with pytest.raises(AssertionError):
            CHECKS[PLANTED[plant]]()
