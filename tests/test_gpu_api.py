"""GPU tests of the C-ABI behaviour around the step: field round trips, the
CFL / NaN stop (main.f90:945, and a NaN), Euler after a restart (main.f90:323),
the readwrite.f90 file layout and the Fortran driver."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from mhit36_amd import Solver, MhitError, MHIT_ERR_CFL, MHIT_ERR_NAN
from mhit36_amd._lib import MHIT_ERR_ARG, MHIT_ERR_IO
from oracle import orc

from gpu_util import random_velocity

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_set_get_roundtrip_and_mean_bookkeeping():
    n = 16
    u, v, w = random_velocity(n, seed=2)
    with Solver(nx=n, dt=1e-3, mu=0.01, phiflag=0, inflow=1) as s:
        for f, a in ((1, u), (2, v), (3, w)):
            s.set_field(f, a)
        for f, a in ((1, u), (2, v), (3, w)):
            assert np.array_equal(s.get_field(f), a)
        s.step(2)
        got = s.get_field(1)
        s.set_field(1, got)            # re-setting a field keeps the state
        assert np.array_equal(s.get_field(1), got)


def test_cfl_stop_matches_oracle():
    n = 16
    kw = dict(nx=n, dt=3.0, mu=0.0, phiflag=0)
    with Solver(**kw) as s:
        s.init_fields()
        d = s.step(5, raise_on_stop=False)
        assert d.status == MHIT_ERR_CFL and d.steps_done == 1 and d.t == 1
        o = orc.Oracle(**kw)
        o.init_ic()
        do = o.step(5)
        assert do["status"] == 5 and d.cou == do["cou"]
        assert np.array_equal(s.get_field(1), o.get_field(1))
        with pytest.raises(MhitError) as e:
            s.step(1)
        assert e.value.code == MHIT_ERR_CFL


def test_nan_stop_keeps_phi_parity():
    """A NaN halts mid-call; the fields are those of the last step that
    ran, identical to the oracle's."""
    n = 16
    kw = dict(nx=n, dt=1e-3, mu=0.01, phiflag=1, sigma=0.5)
    with Solver(**kw) as s:
        s.init_fields()
        o = orc.Oracle(**kw)
        o.init_ic()
        phi = o.get_field(5)
        phi[3, 3, 3] = -0.1
        s.set_field(5, phi)
        o.set_field(5, phi)
        d = s.step(6, raise_on_stop=False)
        do = o.step(6)
        assert d.status == MHIT_ERR_NAN and do["status"] == 6
        assert d.steps_done == do["steps_done"]
        a, b = s.get_field(5), o.get_field(5)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def test_restart_resets_to_euler():
    n = 16
    kw = dict(nx=n, dt=1e-3, mu=0.01, phiflag=0)
    with Solver(**kw) as s:
        s.init_fields()
        s.step(3)
        state = {f: s.get_field(f) for f in (1, 2, 3)}
        s.reset_time(3)
        d = s.step(2)
        assert d.t == 5
        o = orc.Oracle(**{**kw, "inflow": 1})
        for f, a in state.items():
            o.set_field(f, a)
        o.step(2)
        for f in (1, 2, 3):
            assert np.array_equal(s.get_field(f), o.get_field(f))


def test_restart_with_phase_field_restores_gumax():
    """A restart (main.f90:318-323) starts with gumax = 1.0, which sets the
    lagged gamma of the first phase-field step: after reset_time the device
    steps exactly as a fresh oracle seeded with the same fields."""
    n = 16
    kw = dict(nx=n, dt=1e-3, mu=0.01, phiflag=1, sigma=0.2, f1=0.5, f2=0.3, f3=0.2, gamma_factor=1.2)
    with Solver(**kw) as s:
        s.init_fields()
        s.step(3)
        state = {f: s.get_field(f) for f in (1, 2, 3, 5)}
        s.reset_time(3)
        d = s.step(2)
        assert d.t == 5
        o = orc.Oracle(**{**kw, "inflow": 1, "inphi": 1})
        for f, a in state.items():
            o.set_field(f, a)
        do = o.step(2)
        assert d.gumax == do["gumax"]
        for f in (1, 2, 3, 4, 5):
            assert np.array_equal(s.get_field(f), o.get_field(f)), f


def test_load_state_rejects_other_abi_and_length(tmp_path):
    n = 16
    kw = dict(nx=n, dt=1e-3, mu=0.01, phiflag=1)
    path = str(tmp_path / "s.state")
    with Solver(**kw) as s:
        s.init_fields()
        s.step(2)
        s.save_state(path)
        raw = open(path, "rb").read()
        with Solver(**kw) as t:
            t.load_state(path)                       # the file as written loads
            assert np.array_equal(t.get_field(5), s.get_field(5))
            bad = bytearray(raw)
            bad[8:12] = (12345).to_bytes(4, "little")    # abi field after the 8-byte magic
            open(path, "wb").write(bytes(bad))
            with pytest.raises(MhitError) as e:
                t.load_state(path)
            assert e.value.code == MHIT_ERR_ARG
            open(path, "wb").write(raw[:-8])
            with pytest.raises(MhitError) as e:
                t.load_state(path)
            assert e.value.code == MHIT_ERR_IO
            open(path, "wb").write(raw + b"\0" * 8)
            with pytest.raises(MhitError) as e:
                t.load_state(path)
            assert e.value.code == MHIT_ERR_IO


def test_field_files_follow_readwrite_layout(tmp_path):
    n = 16
    with Solver(nx=n, dt=1e-3, mu=0.01, phiflag=1) as s:
        s.init_fields()
        s.step(1)
        for f in (1, 2, 3, 4, 5):
            s.write_field(str(tmp_path), 7, f)
        for f, name in ((1, "u"), (2, "v"), (3, "w"), (4, "p"), (5, "phi")):
            path = tmp_path / f"{name}_00000007.dat"
            raw = np.fromfile(path, dtype="<f8")
            assert raw.size == n ** 3
            assert np.array_equal(raw.reshape((n, n, n), order="F"), s.get_field(f))
        u = s.get_field(1)
        s.set_field(1, np.zeros((n, n, n)))
        s.read_field(str(tmp_path / "u_00000007.dat"), 1)
        assert np.array_equal(s.get_field(1), u)


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "mhit36_amd", "fortran", "mhit36")),
                    reason="Fortran driver not built (amdflang missing)")
def test_fortran_driver_end_to_end(tmp_path):
    """The Fortran driver reads input.inp, runs 6 steps with dumps every 3 and
    writes readwrite.f90-layout files identical to the oracle's fields."""
    n = 16
    lines = open(os.path.join(ROOT, "tests", "golden", "input.inp")).read().splitlines()
    lines[2] = "6           ! Final time step"
    lines[3] = "3           ! Saving frequency"
    (tmp_path / "input.inp").write_text("\n".join(lines) + "\n")
    (tmp_path / "output").mkdir()
    exe = os.path.join(ROOT, "mhit36_amd", "fortran", "mhit36")
    r = subprocess.run([exe, str(n), "1"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Time step" in r.stdout and "CFL" in r.stdout
    o = orc.Oracle(nx=n, dt=0.0002, mu=0.006, rho=1.0, k0=2.0, radius=0.5, sigma=1.0, epsr=1.0, phiflag=1)
    o.init_ic()
    for t in (0, 3, 6):
        if t:
            o.step(3)
        for f, name in ((1, "u"), (2, "v"), (3, "w"), (5, "phi")):
            raw = np.fromfile(tmp_path / "output" / f"{name}_{t:08d}.dat", dtype="<f8").reshape((n, n, n), order="F")
            assert np.array_equal(raw, o.get_field(f)), (t, name)
    # drop count at every dump, as post/DropCountSize writes it
    counts = [tuple(int(x) for x in l.split()) for l in open(tmp_path / "output" / "drop_count.dat")]
    assert [t for t, _ in counts] == [0, 3, 6]
    for t, nd in counts:
        phi = np.fromfile(tmp_path / "output" / f"phi_{t:08d}.dat", dtype="<f8").reshape((n, n, n), order="F")
        assert nd == orc.drop_count(phi)[0] == 1


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "mhit36_amd", "fortran", "mhit36")),
                    reason="Fortran driver not built (amdflang missing)")
def test_fortran_driver_mixed_precision(tmp_path):
    """`mhit36 16 1 1` runs the mixed-precision step: its dumps equal the
    fields of the same box stepped through the Python binding."""
    n = 16
    lines = open(os.path.join(ROOT, "tests", "golden", "input.inp")).read().splitlines()
    lines[2] = "4           ! Final time step"
    lines[3] = "4           ! Saving frequency"
    (tmp_path / "input.inp").write_text("\n".join(lines) + "\n")
    (tmp_path / "output").mkdir()
    exe = os.path.join(ROOT, "mhit36_amd", "fortran", "mhit36")
    r = subprocess.run([exe, str(n), "1", "1"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    with Solver(nx=n, phiflag=1, precision=1) as s:
        s.init_fields()
        s.step(4)
        for f, name in ((1, "u"), (2, "v"), (3, "w"), (5, "phi")):
            raw = np.fromfile(tmp_path / "output" / f"{name}_00000004.dat", dtype="<f8").reshape((n, n, n), order="F")
            assert np.array_equal(raw, s.get_field(f)), name


@pytest.mark.parametrize("precision", [0, 1])
def test_checkpoint_resume_is_bitwise(tmp_path, precision):
    """mhit_save_state / mhit_load_state: 4 steps, checkpoint, a new context
    resumes and runs 5 more; the fields equal those of 9 uninterrupted steps
    (AB2 history, pending means and the lagged umax all restored)."""
    from mhit36_amd import ic
    n = 32
    u, v, w = ic.random_velocity(n, seed=9)
    _, dx = ic.grid(n)
    phi = ic.drops(n, ic.drop_centres(2, 0.6, dx, seed=2), 0.6, dx)
    kw = dict(nx=n, dt=2e-3, mu=0.01, f1=1.0, phiflag=1, inflow=1, inphi=1, radius=0.6, sigma=0.1,
              gamma_factor=1.2, precision=precision)
    fields = {1: u, 2: v, 3: w, 5: phi}
    with Solver(**kw) as a:
        for f, x in fields.items():
            a.set_field(f, x)
        a.step(9)
        ref = {f: a.get_field(f) for f in (1, 2, 3, 4, 5)}
    with Solver(**kw) as b:
        for f, x in fields.items():
            b.set_field(f, x)
        b.step(4)
        b.save_state(tmp_path / "box.state")
    with Solver(**kw) as c:
        c.load_state(tmp_path / "box.state")
        c.step(5)
        for f, x in ref.items():
            assert np.array_equal(c.get_field(f), x), f
    other = dict(kw, precision=1 - precision)
    with Solver(**other) as d:
        with pytest.raises(Exception):
            d.load_state(tmp_path / "box.state")
