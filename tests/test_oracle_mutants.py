"""The pins of tests/oracle_pins.py catch plausible mistakes.

Each case plants one mistake in a copy of oracle/mhit_oracle.c (a sign, a
factor, a neighbour index, a missing term), builds that copy with the
oracle's own flags into a temporary directory, and requires the named pin to
fail on it.  The oracle itself is never modified; the planted copies exist
only for the duration of the test.
"""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle_pins as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "mhit_oracle.c")

# (name, exact text in the oracle, planted replacement, pin that must fail
# [, sharpening form the pin runs with; default the identity form])
MUTANTS = [
    ("sharpening_sign", "double r = rphi[id] - gamma * ((X + Y) + Z);",
     "double r = rphi[id] + gamma * ((X + Y) + Z);", "sphere_equilibrium"),
    ("tanh_factor", "a = 0.25 * (1.0 - th * th);",
     "a = 0.5 * (1.0 - th * th);", "sphere_equilibrium", "literal"),
    ("identity_factor", "a = (ps != ps) ? ps : (A * B) / (S * S);",
     "a = (ps != ps) ? ps : (2.0 * A * B) / (S * S);", "sphere_equilibrium"),
    ("tanh_argument", "orc_tanh((0.5 * psi[IDX(i, j, k, n)]) * epsi)",
     "orc_tanh(psi[IDX(i, j, k, n)] * epsi)", "sphere_equilibrium", "literal"),
    ("identity_sum", "double A = val + en, B = (1.0 - val) + en, S = A + B;",
     "double A = val + en, B = (1.0 - val) + en, S = A + 2.0 * B;", "sphere_equilibrium"),
    ("diffusion_y_no_eps", "double dyy = eps * ((fyp - 2.0 * f0) + fym) * ddxi;",
     "double dyy = ((fyp - 2.0 * f0) + fym) * ddxi;", "sphere_equilibrium"),
    ("normal_y_one_sided", "double gy = psi[IDX(i, jp, k, n)] - psi[IDX(i, jm, k, n)];",
     "double gy = 2.0 * (psi[IDX(i, jp, k, n)] - psi[IDX(i, j, k, n)]);", "sphere_equilibrium"),
    ("normal_no_sqrt", "double normod = 1.0 / (sqrt((gx * gx + gy * gy) + gz * gz) + tiny);",
     "double normod = 1.0 / (((gx * gx + gy * gy) + gz * gz) + tiny);", "sphere_equilibrium"),
    # psi enters the identity form only through its normalised gradient, so a
    # missing eps changes only the literal form's tanh argument
    ("psi_no_eps", "psi[q] = eps * orc_log(", "psi[q] = orc_log(", "sphere_equilibrium", "literal"),
    ("convection_x_staggering",
     "double A = u[IDX(ip, j, k, n)] * 0.5 * (fxp + f0) - u[IDX(i, j, k, n)] * 0.5 * (f0 + fxm);",
     "double A = u[IDX(i, j, k, n)] * 0.5 * (fxp + f0) - u[IDX(im, j, k, n)] * 0.5 * (f0 + fxm);",
     "convection_staggering"),
    ("convection_z_face",
     "double C = w[IDX(i, j, kp, n)] * 0.5 * (fzp + f0) - w[IDX(i, j, k, n)] * 0.5 * (f0 + fzm);",
     "double C = w[IDX(i, j, k, n)] * 0.5 * (fzp + f0) - w[IDX(i, j, km, n)] * 0.5 * (f0 + fzm);",
     "convection_staggering"),
    ("convection_average",
     "double A = u[IDX(ip, j, k, n)] * 0.5 * (fxp + f0) - u[IDX(i, j, k, n)] * 0.5 * (f0 + fxm);",
     "double A = u[IDX(ip, j, k, n)] * (fxp + f0) - u[IDX(i, j, k, n)] * (f0 + fxm);",
     "convection_mode"),
    ("phi_update_dt", "phin[id] = phi[id] + s->p.dt * r;",
     "phin[id] = phi[id] + 2.0 * s->p.dt * r;", "convection_mode"),
    ("csf_factor", "sig6 = 6.0 * s->p.sigma;", "sig6 = 3.0 * s->p.sigma;", "laplace"),
    ("chem_laplacian_sign", "double chem = t1 - eps * s6 * ddxi;",
     "double chem = t1 + eps * s6 * ddxi;", "csf_plane"),
    ("chem_well_sign", "double t1 = f0 * (1.0 - f0) * (1.0 - 2.0 * f0) * epsi;",
     "double t1 = f0 * (1.0 - f0) * (2.0 * f0 - 1.0) * epsi;", "csf_plane"),
    ("grad_z_one_sided", "double gz = 0.5 * (fzp - fzm) * dxi;",
     "double gz = (fzp - f0) * dxi;", "laplace"),
    ("csf_face_x", "ru = ru + 0.5 * (s->fx[IDX(im, j, k, n)] + s->fx[IDX(i, j, k, n)]) * rhoi;",
     "ru = ru + 0.5 * (s->fx[IDX(ip, j, k, n)] + s->fx[IDX(i, j, k, n)]) * rhoi;", "laplace"),
    ("forcing_u_swap", "ru = (ru + f3 * s->sinf[k]) + f2 * s->cosf[j];",
     "ru = (ru + f3 * s->sinf[j]) + f2 * s->cosf[k];", "forcing"),
    ("forcing_v_index", "rv = (rv + f1 * s->sinf[i]) + f3 * s->cosf[k];",
     "rv = (rv + f1 * s->sinf[j]) + f3 * s->cosf[k];", "forcing"),
    ("forcing_w_coeff", "rw = (rw + f2 * s->sinf[j]) + f1 * s->cosf[i];",
     "rw = (rw + f1 * s->sinf[j]) + f2 * s->cosf[i];", "forcing"),
]


def _cflags():
    mk = open(os.path.join(ROOT, "oracle", "Makefile")).read()
    m = re.search(r"CFLAGS\s*\?=\s*((?:.*\\\n)*.*)", mk)
    return m.group(1).replace("\\\n", " ").split()


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = open(SRC).read()
    base = tmp_path_factory.mktemp("mutants")
    shutil.copy(os.path.join(ROOT, "oracle", "mhit_oracle.h"), base)
    flags = _cflags()

    def build(m):
        name, old, new = m[:3]
        assert src.count(old) == 1, f"{name}: site not found exactly once in the oracle"
        c = base / f"{name}.c"
        c.write_text(src.replace(old, new))
        so = base / f"lib{name}.so"
        subprocess.run(["gcc", *flags, "-shared", "-o", str(so), str(c), "-lm"], check=True,
                       capture_output=True)
        return name, str(so)

    with ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build, MUTANTS))


@pytest.mark.slow
@pytest.mark.parametrize("m", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pin_catches_planted_mistake(mutant_libs, m):
    name, pin = m[0], m[3]
    form = m[4] if len(m) > 4 else "identity"
    with pytest.raises(AssertionError):
        P.ALL[pin](mutant_libs[name], form=form)


@pytest.mark.parametrize("pin", sorted(P.ALL))
def test_every_pin_has_a_mutant(pin):
    assert any(m[3] == pin for m in MUTANTS)
