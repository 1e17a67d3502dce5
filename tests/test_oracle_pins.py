"""The oracle's phase field, CSF force and ABC forcing
against what the mathematics fixes (tests/oracle_pins.py): the ACDI
equilibrium of the tanh profile, the centred-difference eigenvalue of a
Fourier mode, second-order convergence of the staggered convective term, the
vanishing chemical potential of a plane interface, Laplace's law for a drop,
and the ABC forcing pointwise -- for both forms of the sharpening factor
(tanh of psi as written, and the identity the HIP path evaluates).  tests/test_oracle_mutants.py shows that each
of these fails when a plausible mistake is planted in the oracle."""
import pytest

import oracle_pins as P


@pytest.mark.parametrize("form", ["identity", "literal"])
@pytest.mark.parametrize("pin", sorted(P.ALL))
def test_oracle_pin(pin, form):
    P.ALL[pin](form=form)
