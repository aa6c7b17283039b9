"""Restart files (REVIVE, REVCON; utility_pack.f:1772-1862) against the
reference's own example outputs (data fixtures), and the Fortran output
editing they use against the flang runtime in this image."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN

FLANG = "/opt/rocm/llvm/bin/flang"


def test_revive_round_trip_is_byte_identical(restart_example, tmp_path):
    from fastmc_amd import inputs
    src = os.path.join(restart_example["dir"], "REVIVE")
    rev = inputs.read_revive(src, restart_example["field"])
    assert rev.nummols == [90, 36] and rev.natms == 1854 and rev.is_production
    assert len(rev.chainstats) == 15 and rev.chainstats[0] == 99999800.0
    out = tmp_path / "REVIVE"
    inputs.write_revive(str(out), rev)
    assert out.read_bytes() == open(src, "rb").read()


def test_revive_reader_rejects_a_foreign_field(restart_example, tmp_path):
    """A REVIVE that does not fit the FIELD's molecule types is refused."""
    import copy
    from fastmc_amd import inputs
    f = copy.deepcopy(restart_example["field"])
    f.molecules[1].names = f.molecules[1].names[:-1]
    with pytest.raises(ValueError):
        inputs.read_revive(os.path.join(restart_example["dir"], "REVIVE"), f)
    bad = tmp_path / "REVIVE"
    bad.write_bytes(open(os.path.join(restart_example["dir"], "REVIVE"), "rb").read()[:-3])
    with pytest.raises(ValueError):
        inputs.read_revive(str(bad), restart_example["field"])


def test_revcon_regenerated_byte_for_byte(restart_example, tmp_path):
    """The reference wrote REVCON and REVIVE in one call: the REVCON written
    here from the REVIVE coordinates equals the reference's, byte for byte
    (atom order, names, G20.10/E20.10/F20.12 editing)."""
    from fastmc_amd import inputs
    ex = restart_example
    rev = ex["revive"]
    atoms = inputs.revcon_atoms(ex["field"], rev.coords)
    out = tmp_path / "REVCON"
    inputs.write_revcon(str(out), ex["config"].title, ex["config"].imcon,
                        ex["config"].cell, atoms, eng=0.0)
    ref = open(os.path.join(ex["dir"], "REVCON")).read()
    assert out.read_text() == ref


def test_revcon_starts_the_binary_example(restart_example):
    """The reference's binary example starts a CO2 + TIP4P-Ew run from the
    unary REVCON used as CONFIG, with the CO2 NUMMOLS set to 90: the
    coordinates read back are the REVIVE's."""
    from fastmc_amd import inputs
    d = os.path.join(GOLDEN, "binary_example")
    field = inputs.read_field(os.path.join(d, "FIELD"))
    assert [m.guest for m in field.molecules] == [True, True, False]
    field.molecules[0].nummols = 90
    cfg = inputs.read_config(os.path.join(restart_example["dir"], "REVCON"), field)
    rev = restart_example["revive"]
    got = np.array(cfg.coords[0])
    assert got.shape == (270, 3) and cfg.coords[1] == []
    for c in range(3):
        # G20.10 keeps 10 significant digits
        assert np.allclose(got[:, c], rev.coords[0][c], rtol=1e-9, atol=1e-9)
    fw = np.array(cfg.coords[2])
    assert fw.shape == (1584, 3)
    assert np.allclose(fw[:, 0], rev.coords[1][0], rtol=1e-9, atol=1e-9)
    ctl = inputs.read_control(os.path.join(d, "CONTROL"))
    assert sorted(ctl.guests) == [1, 2]


VALUES = [0.0, 1.0, -1.0, 0.1, -0.1, 0.09999999999, 0.099999999994, 0.0999999999951,
          1e-5, -3.25e-7, 12345.678901234, 9999999999.4, 9999999999.6, 1e10, -2.5e12,
          35.6552, -12.425126850515, 0.6562186900, 0.0559972839614, 123456789.987654321,
          7.5e-300, 1.0e99]


def test_fortran_editing_matches_flang(tmp_path):
    """G20.10 and E20.10 as written by a Fortran runtime (flang's), on
    rounding and range edges."""
    if not os.path.exists(FLANG):
        pytest.skip("flang not in this image")
    from fastmc_amd import inputs
    vals = VALUES
    src = tmp_path / "g.f90"
    # the exact doubles travel through a stream file
    lines = ["program g", "  implicit none", f"  real(8) :: v({len(vals)})",
             "  integer :: i",
             "  open(1, file='v.bin', access='stream', form='unformatted')",
             "  read(1) v", "  close(1)",
             f"  do i = 1, {len(vals)}",
             "    write(*, '(g20.10)') v(i)",
             "    write(*, '(e20.10)') v(i)",
             "  end do", "end program g"]
    src.write_text("\n".join(lines) + "\n")
    np.array(vals, "<f8").tofile(str(tmp_path / "v.bin"))
    subprocess.check_call([FLANG, "-o", str(tmp_path / "g"), str(src)], cwd=str(tmp_path))
    out = subprocess.run([str(tmp_path / "g")], cwd=str(tmp_path), capture_output=True,
                         text=True, check=True).stdout.split("\n")
    for i, v in enumerate(vals):
        assert inputs.fortran_g(v) == out[2 * i], (v, out[2 * i])
        assert inputs.fortran_e(v) == out[2 * i + 1], (v, out[2 * i + 1])


def binary_example(restart_example):
    """The reference's binary example as its README sets it up: the binary
    FIELD with CO2 NUMMOLS 90 and the unary REVCON as CONFIG."""
    from fastmc_amd import inputs
    d = os.path.join(GOLDEN, "binary_example")
    field = inputs.read_field(os.path.join(d, "FIELD"))
    field.molecules[0].nummols = 90
    config = inputs.read_config(os.path.join(restart_example["dir"], "REVCON"), field)
    control = inputs.read_control(os.path.join(d, "CONTROL"))
    return field, config, control


def test_binary_example_starts_from_known_answer_2(oracle, restart_example):
    """Adding TIP4P-Ew as a second, empty guest type leaves the energy of
    the 90-CO2 configuration at known answer 2 (-731.5072668 kcal/mol);
    the REVCON's 10-digit coordinates move it by less than 1e-6."""
    field, config, control = binary_example(restart_example)
    assert [g.press_pa for _, g in sorted(control.guests.items())] == [99036.4, 441.1]
    s, _, _ = oracle.system_from_inputs(field, config, control)
    e = s.energy_full()
    spe = sum(e[1:6]) / field.engunit
    assert abs(spe - restart_example["kat"]["revive_energy_kcal"]) < 1e-6
