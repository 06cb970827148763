"""The reference's file formats (host only, no GPU): the parameter and
geometry files it reads (SIMPLE.cpp:31-60, 517-551), the restart matrices
(SIMPLE.cpp:203-232) and the files saveAll writes (output.cpp:47-266). The
reference ships no sample files, so the inputs are written here in the
layout its loaders parse, and the outputs are checked against the formulas
of output.cpp."""
from __future__ import annotations

import numpy as np
import pytest

import cpbind as cb


def test_read_params_core_and_optional_columns(tmp_path):
    f = tmp_path / "fluid_params.txt"
    base = cb.default_params()
    base.Ht = 7e-5
    rho0, mu0 = base.rho, base.mu
    f.write_text("50 200 5e-05 4e-05 0.1 10 20\n")
    p, ni, no = cb.read_params(f, base)
    assert (p.M, p.N, p.hy, p.hx, p.U_in, ni, no) == (50, 200, 5e-5, 4e-5, 0.1, 10, 20)
    assert p.Ht == 7e-5  # absent: unchanged
    f.write_text("50 200 5e-05 4e-05 0.1 10 20 0.0002\n")
    p, _, _ = cb.read_params(f, base)
    assert p.Ht == 2e-4
    f.write_text("64 128 2.5e-05 2.5e-05 0.2 0 0 0.0002 997 1e-06\n")
    p, _, _ = cb.read_params(f, base)
    assert (p.M, p.N, p.Ht) == (64, 128, 2e-4)
    assert (p.rho, p.mu) == (rho0, mu0)  # file fluid properties are ignored
    f.write_text("50 200 5e-05\n")
    with pytest.raises(cb.CpError):
        cb.read_params(f, base)


def test_read_geometry_formats(tmp_path):
    f = tmp_path / "geometry_fluid.txt"
    ref = np.array([[1, 1, 1, 1], [0, 0.5, 0.25, 0], [1, 1, 1, 1]], dtype=float)
    f.write_text("1 1 1 1\n0 0.5 0.25 0\n1 1 1 1\n")
    ct, found = cb.read_geometry(f, 3, 4)
    assert found
    np.testing.assert_array_equal(ct, ref)
    f.write_text("1 1 1 1\n0 0.5\n")  # short: the rest stays fluid
    ct, _ = cb.read_geometry(f, 3, 4)
    np.testing.assert_array_equal(ct.ravel()[:6], [1, 1, 1, 1, 0, 0.5])
    assert np.all(ct.ravel()[6:] == 0)
    ct, found = cb.read_geometry(tmp_path / "missing.txt", 3, 4)
    assert not found and np.all(ct == 0)


def test_matrix_round_trip_and_reference_digits(tmp_path):
    a = np.random.default_rng(1).standard_normal((7, 9)) * 1e3
    f = tmp_path / "m.txt"
    cb.write_matrix(f, a, precision=17)
    np.testing.assert_array_equal(cb.read_matrix(f, 7, 9), a)
    cb.write_matrix(f, np.array([[1.0 / 3.0, 123456.789, 0.1]]))
    assert f.read_text() == "0.333333\t123457\t0.1\n"  # float, 6 significant digits
    with pytest.raises(cb.CpError):
        cb.read_matrix(f, 2, 3)  # incomplete


def test_save_all_files(tmp_path):
    M, N, ni, no = 5, 8, 2, 1
    rng = np.random.default_rng(3)
    u, v, p = rng.standard_normal((M + 1, N)), rng.standard_normal((M, N + 1)), rng.standard_normal((M + 1, N + 1))
    ct = (rng.random((M, N)) > 0.5).astype(float)
    gam, alp = 1 - ct, ct * 1e9
    prm = cb.default_params()
    prm.M, prm.N, prm.hx, prm.hy = M, N, 2e-5, 3e-5
    cb.save_all(tmp_path, prm, u, v, p, ct, gam, alp, ni, no, precision=17)
    np.testing.assert_array_equal(cb.read_matrix(tmp_path / "u.txt", M + 1, N), u)
    np.testing.assert_array_equal(cb.read_matrix(tmp_path / "p.txt", M + 1, N + 1), p)
    uc = 0.5 * (u[:-1, :] + u[1:, :])
    vc = 0.5 * (v[:, :-1] + v[:, 1:])
    rows = np.clip(np.arange(M) + 1, 1, M - 1)
    cols = np.clip(np.arange(N) + 1, 1, N - 1)
    pc = p[np.ix_(rows, cols)]
    np.testing.assert_allclose(cb.read_matrix(tmp_path / "u_full.txt", M, N), uc, rtol=1e-15)
    np.testing.assert_allclose(cb.read_matrix(tmp_path / "v_full.txt", M, N), vc, rtol=1e-15)
    np.testing.assert_array_equal(cb.read_matrix(tmp_path / "pressure_full.txt", M, N), pc)
    gx = np.gradient(pc, prm.hx, axis=1, edge_order=1)
    gy = np.gradient(pc, prm.hy, axis=0, edge_order=1)
    nt = N - ni - no
    np.testing.assert_allclose(cb.read_matrix(tmp_path / "u_thermal.txt", M, nt), uc[:, ni:ni + nt], rtol=1e-15)
    np.testing.assert_allclose(cb.read_matrix(tmp_path / "p_gradient.txt", M, nt),
                               np.hypot(gx, gy)[:, ni:ni + nt], rtol=1e-12)
    vtk = (tmp_path / "fluid_results.vtk").read_text().splitlines()
    assert vtk[:5] == ["# vtk DataFile Version 3.0", "SIMPLE CFD Results", "ASCII", "DATASET STRUCTURED_POINTS",
                       f"DIMENSIONS {N} {M} 1"]
    assert vtk[7] == f"POINT_DATA {M * N}"
    names = [l.split()[1] for l in vtk if l.startswith(("SCALARS", "VECTORS"))]
    assert names == ["pressure", "velocity", "cellType", "Density", "Alpha", "PressureGradient", "pressure_gradient"]
    assert len(vtk) == 8 + 7 * M * N + 5 * 2 + 2
