"""The Fortran host program (shared_cfd_amd/fortran/main_shcfd.f90) drives
the HIP library through ISO_C_BINDING with the reference's routine names;
its RMS.dat and result.bin agree with the CPU oracle run the same way."""
import os
import subprocess

import numpy as np
import pytest

from shared_cfd_amd import cases as C

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "shared_cfd_amd", "main_shcfd")


def _run_driver(case, tmp_path, nsteps, out_period, fast):
    C.write_case(case, str(tmp_path / "case.inp"), max_step=nsteps, out_period=out_period, fast=fast)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "normal end" in r.stdout
    rms = np.loadtxt(tmp_path / "RMS.dat", ndmin=2)
    return rms, C.read_result_bin(str(tmp_path / "result.bin"), case)


def _run_oracle(case, nsteps, out_period):
    from oracle.oracle import Oracle

    o = Oracle(case)
    o.prepare()
    rms = []
    i = 1
    while True:
        if i % out_period == 0:
            rms.append((i, o.energy_rms()))
            if nsteps <= i:
                break
        assert o.step()[0] == 0
        i += 1
    return o, np.array(rms)


@pytest.mark.parametrize("name,fast", [("sod", 0), ("sod", 1), ("step", 1), ("step_nr", 0)])
def test_fortran_driver_matches_oracle(tmp_path, name, fast):
    if name == "sod":
        case, n, per = C.sod(64, 4), 20, 5
    elif name == "step_nr":
        case, n, per = C.forward_step(n_per_unit=40, scheme=C.SCHEME_NR, num_internal_loop=6), 3, 1
    else:
        case, n, per = C.forward_step(n_per_unit=40, num_internal_loop=6), 3, 1
    rms_d, (step, tt, blocks) = _run_driver(case, tmp_path, n, per, fast)
    o, rms_o = _run_oracle(case, n, per)
    assert rms_d.shape == rms_o.shape
    np.testing.assert_array_equal(rms_d[:, 0], rms_o[:, 0])
    np.testing.assert_allclose(rms_d[1:, 1], rms_o[1:, 1], rtol=1e-9)
    assert step == n - 1  # steps 1..n-1 run: the loop test at step n stops (check_convergence.f90:48)
    np.testing.assert_allclose(tt, o.tt(), rtol=1e-13)
    for b in range(case.nblocks):
        ni, nj = case.ni[b], case.nj[b]
        q_o = o.field("Q", b)[2:nj + 2, 2:ni + 2]
        q_d = blocks[b][1:nj + 1, 1:ni + 1, :case.dimq]
        np.testing.assert_allclose(q_d, q_o, rtol=1e-11, atol=1e-11 * np.max(np.abs(q_o)))


def test_fortran_driver_reacting(tmp_path):
    """Point-implicit LU-SGS with Chemkin thermo (init_therm, proceed_reaction)
    on a small mixing layer of the no27 mechanism plus H2O2."""
    from shared_cfd_amd import chem

    from chem_cases import NO27_9, no27_plus_h2o2

    p9 = no27_plus_h2o2(str(tmp_path / "mech"))
    m = chem.parse(p9)
    case = C.mixing_layer(m, ni=24, nj=12, seed=2, species=NO27_9, reaction=2, T_ox=1200.0)
    n = 4
    C.write_case(case, str(tmp_path / "case.inp"), max_step=n, out_period=1, fast=True,
                 mech_files=(p9, chem.MW_INP, None))
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "normal end" in r.stdout, r.stdout + r.stderr
    step, tt, blocks = C.read_result_bin(str(tmp_path / "result.bin"), case)
    o, _ = _run_oracle(case, n, 1)
    q_o = o.field("Q", 0)[2:14, 2:26]
    q_d = blocks[0][1:13, 1:25, :case.dimq]
    np.testing.assert_allclose(q_d, q_o, rtol=1e-10, atol=1e-10 * np.max(np.abs(q_o)))


@pytest.mark.parametrize("scheme", [C.SCHEME_LUSGS, C.SCHEME_PRECON_LUSGS])
def test_fortran_driver_cold_flow(tmp_path, scheme):
    """Cold flow with NASA-9 thermodynamics (init_pack_cold_flow) on a small
    fuel/oxidizer layer with the tables of the reference's sample no28."""
    from shared_cfd_amd import nasa9 as N

    g = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    files = (os.path.join(g, "no28_chem.inp"), os.path.join(g, "nasa9_thermo.inp"),
             os.path.join(g, "nasa9_trans.inp"), os.path.join(g, "no28_control_chem.inp"))
    t, pt = N.parse(*files)
    case = N.mixing_case(t, pt, ni=24, nj=12, seed=2, scheme=scheme)
    n = 4
    C.write_case(case, str(tmp_path / "case.inp"), max_step=n, out_period=1, fast=True, mech_files=files)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "normal end" in r.stdout, r.stdout + r.stderr
    step, tt, blocks = C.read_result_bin(str(tmp_path / "result.bin"), case)
    o, _ = _run_oracle(case, n, 1)
    q_o = o.field("Q", 0)[2:14, 2:26]
    q_d = blocks[0][1:13, 1:25, :case.dimq]
    np.testing.assert_allclose(q_d, q_o, rtol=1e-11, atol=1e-11 * np.max(np.abs(q_o)))
