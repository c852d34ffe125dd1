"""The oracle's pins catch planted slips: each slip below is planted into a copy of the oracle's transport
(`oracle/transport.py`, the restatement of `src/sosfuse2/backends/numpy_reference.py:467-643`), on the CPU, and at
least one of the independent FP64 pins of `tests/test_warm_pins.py` must then fail. The pins compute their answers
from the physics (closed forms and FP64 quadratures), not from the oracle's formulas, so a slip shared by the oracle
and the HIP kernels cannot hide behind the GPU parity tests.

The slips: the SH term of the warm tensor divided by 3 instead of 2 (numpy_reference.py:619); a permuted index of
the tensor rotation (:615); the ground boundary term not halved (:642); the first-light ground-pool sum doubled
(:555)."""

import types
from pathlib import Path

import numpy as np
import pytest

import oracle.transport
import test_warm_pins as pins

SLIPS = {
    'sh_term_third': ("tensor[:, :, k] += (aero * pu[:, k, None]) / F(2)",
                      "tensor[:, :, k] += (aero * pu[:, k, None]) / F(3)"),
    'rotation_index': ("rotated += q['T'][:, k, uu, None] * rm[:, :, uu]",
                       "rotated += q['T'][:, k, (uu + 1) % 4, None] * rm[:, :, uu]"),
    'boundary_unhalved': ("total[:, :, k] += (E * pu[:, k, None]) / F(2)",
                          "total[:, :, k] += (E * pu[:, k, None]) / F(1)"),
    'ground_pools_doubled': ("ground += wph[:, gi] * gv[:, gi]",
                             "ground += F(2) * (wph[:, gi] * gv[:, gi])"),
}


def mutant(name):
    """A copy of oracle.transport with one slip planted (its source text replaced once)."""
    path = Path(oracle.transport.__file__)
    src = path.read_text()
    old, new = SLIPS[name]
    assert src.count(old) == 1, name
    mod = types.ModuleType('oracle.transport_' + name)
    mod.__file__ = str(path)
    mod.__package__ = 'oracle'
    exec(compile(src.replace(old, new), str(path), 'exec'), mod.__dict__)
    return mod


@pytest.fixture(scope='module')
def case():
    return pins.make_warm_case()


def run_pins(case, which):
    """Run the named pins; True when all pass."""
    checks = {
        'isotropic': lambda: pins.test_isotropic_warm_source(case, np.float64, 2e-6),
        'rotation': lambda: pins.test_rotation_and_order_one_harmonic(case, np.float64, 2e-6),
        'boundary': lambda: pins.test_ground_boundary_term(case, np.float64, 1e-5),
        'first_light': lambda: pins.test_first_light_source_at_cell_centres(1, 1),
    }
    for w in which:
        try:
            checks[w]()
        except AssertionError:
            return False
    return True


@pytest.mark.parametrize('slip, which', [('sh_term_third', ['isotropic', 'rotation']),
                                         ('rotation_index', ['isotropic', 'rotation']),
                                         ('boundary_unhalved', ['boundary']),
                                         ('ground_pools_doubled', ['first_light'])])
def test_planted_slip_fails_a_pin(case, monkeypatch, slip, which):
    """With the oracle as it is the pins pass; with the slip planted at least one of them fails."""
    assert run_pins(case, which)
    monkeypatch.setattr(pins, 'ot', mutant(slip))
    assert not run_pins(case, which), f'no pin caught the planted slip {slip}'
