"""Scattering-source contraction (computeFSRSources, src/CPUSolver.cpp:
1939-2023) on its three paths: the FP64 loops, fp16 MFMA on hi/lo split
operands with FP32 accumulation, and the FP32 vector fallback.

The expected sources are the FP64 restatement below (numpy over the same
dest-major matrices).  The split-fp16 error model: every operand is scaled by
a power of two to a largest entry in [0.5, 1) per FSR row and per matrix
column, and represented as hi + lo to about 2^-21 of that largest entry, so
with K <= 80 products an entry is off by at most about
80 * 2^-21 * rowmax * colmax < 4e-5 * rowmax * colmax (before the 1/4 pi).
A transposed fragment, a dropped lo term or a wrong row/column scale breaks
that bound by orders of magnitude."""
import numpy as np
import pytest

import openmoc_amd as om

pytestmark = pytest.mark.gpu

FOUR_PI = 4.0 * np.pi


def fp64_sources(tg, phi, k=1.0, iteration=0):
    G = tg.getNumEnergyGroups()
    mat = tg.fsr_material
    S = tg.sigma_s.reshape(-1, G, G)[mat]                # [N, dest, origin]
    nsf, chi = tg.nu_sigma_f[mat], tg.chi[mat]
    fis = (nsf * phi).sum(1)
    fissionable = (tg.nu_sigma_f != 0).any(1)[mat]
    q = np.einsum("rdg,rg->rd", S, phi)
    q += np.where(fissionable[:, None], chi * fis[:, None] / k, 0.0)
    q /= FOUR_PI
    if iteration < 30:
        q = np.where(q < 0.0, 1e-25, q)
    return q


def error_bound(tg, phi, k=1.0, rel=4e-5):
    G = tg.getNumEnergyGroups()
    mat = tg.fsr_material
    S = np.abs(tg.sigma_s.reshape(-1, G, G))
    colmax = S.max(2)[mat]                               # [N, dest]
    fmax = np.abs(tg.nu_sigma_f).max(1)[mat]             # [N]
    rowmax = np.abs(phi).max(1)[:, None]
    return rel * rowmax * (colmax + tg.chi[mat] * fmax[:, None] / k) / FOUR_PI


def wide_range_flux(tg, seed):
    rng = np.random.default_rng(seed)
    N, G = tg.getNumFSRs(), tg.getNumEnergyGroups()
    # six decades across groups and FSRs
    return np.exp(rng.uniform(-9.0, 5.0, (N, G)))


@pytest.mark.parametrize("name", ["pin_cell_70g", "pin_cell_7g"])
def test_fp64_path_matches_restatement(tracks, name):
    tg = tracks(name)
    phi = wide_range_flux(tg, 1)
    s = om.HipSolver(tg)
    assert s.setScatterPrecision(om.SCATTER_FP64) == om.SCATTER_FP64
    s.setFluxes(phi)
    q = s.computeFSRSources(0)
    np.testing.assert_allclose(q, fp64_sources(tg, phi), rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("name", ["pin_cell_70g", "pin_cell_7g"])
def test_mfma_path_within_split_fp16_bound(tracks, name):
    tg = tracks(name)
    phi = wide_range_flux(tg, 2)
    s = om.HipSolver(tg)
    assert s.setScatterPrecision(om.SCATTER_MFMA_F16) == om.SCATTER_MFMA_F16
    s.setFluxes(phi)
    q = s.computeFSRSources(30)
    ref = fp64_sources(tg, phi, iteration=30)
    err = np.abs(q - ref)
    bound = error_bound(tg, phi)
    assert np.all(err <= bound), float((err / bound).max())
    # typical error is far below the worst case
    assert np.linalg.norm(q - ref) / np.linalg.norm(ref) < 1e-6


def test_fp32_path_within_fp32_bound(tracks):
    tg = tracks("pin_cell_70g")
    phi = wide_range_flux(tg, 3)
    s = om.HipSolver(tg)
    assert s.setScatterPrecision(om.SCATTER_FP32) == om.SCATTER_FP32
    s.setFluxes(phi)
    q = s.computeFSRSources(30)
    ref = fp64_sources(tg, phi, iteration=30)
    assert np.all(np.abs(q - ref) <= error_bound(tg, phi, rel=1e-5))


def test_auto_picks_matrix_cores_from_32_groups(tracks):
    assert om.HipSolver(tracks("pin_cell_70g")).setScatterPrecision() == om.SCATTER_MFMA_F16
    assert om.HipSolver(tracks("pin_cell_7g")).setScatterPrecision() == om.SCATTER_FP64


def test_negative_sources_clamped_before_iteration_30(tracks):
    tg = tracks("pin_cell_70g")
    rng = np.random.default_rng(4)
    phi = rng.uniform(-1.0, 1.0, (tg.getNumFSRs(), tg.getNumEnergyGroups()))
    s = om.HipSolver(tg)
    s.setFluxes(phi)
    for it in (0, 30):
        s.setScatterPrecision(om.SCATTER_FP64)
        s.setFluxes(phi)
        np.testing.assert_allclose(s.computeFSRSources(it), fp64_sources(tg, phi, iteration=it),
                                   rtol=1e-12, atol=1e-14)
    s.setScatterPrecision(om.SCATTER_MFMA_F16)
    s.setFluxes(phi)
    q = s.computeFSRSources(0)
    ref = fp64_sources(tg, phi, iteration=30)           # unclamped
    bound = error_bound(tg, phi)
    clear = np.abs(ref) > bound                          # sign decided beyond the error
    np.testing.assert_array_equal(q[clear & (ref < 0)], 1e-25)
    assert np.all(np.abs(q - ref)[clear & (ref > 0)] <= bound[clear & (ref > 0)])


@pytest.mark.parametrize("mode", [om.SCATTER_MFMA_F16, om.SCATTER_FP32])
def test_eigenvalue_with_reduced_precision_contraction(tracks, mode):
    tg = tracks("pin_cell_70g")
    ref = om.HipSolver(tg)
    ref.setScatterPrecision(om.SCATTER_FP64)
    ref.computeEigenvalue(1000)
    s = om.HipSolver(tg)
    s.setScatterPrecision(mode)
    s.computeEigenvalue(1000)
    assert abs(s.getKeff() - ref.getKeff()) < 1e-6      # 0.1 pcm
    f, fr = s.getFluxes(), ref.getFluxes()
    assert np.linalg.norm(f - fr) / np.linalg.norm(fr) < 1e-5


def test_linear_source_with_matrix_cores_matches_fp64(tracks):
    tg = tracks("pin_cell_70g")
    ref = om.HipLSSolver(tg)
    ref.setScatterPrecision(om.SCATTER_FP64)
    ref.computeEigenvalue(1000)
    s = om.HipLSSolver(tg)
    assert s.setScatterPrecision() == om.SCATTER_MFMA_F16
    s.computeEigenvalue(1000)
    assert abs(s.getKeff() - ref.getKeff()) < 1e-6
    np.testing.assert_allclose(s.getFluxMoments(), ref.getFluxMoments(),
                               atol=1e-4 * np.abs(ref.getFluxMoments()).max())


def test_krylov_scatter_sweep_with_matrix_cores(tracks):
    tg = tracks("pin_cell_70g")
    out = []
    for mode in (om.SCATTER_FP64, om.SCATTER_MFMA_F16):
        s = om.HipSolver(tg)
        s.setScatterPrecision(mode)
        s.computeEigenvalue(1000)
        s.scatterTransportSweep()
        out.append(s.getFluxes())
    assert np.linalg.norm(out[1] - out[0]) / np.linalg.norm(out[0]) < 1e-5
