"""The seeded workload generators (jfem_amd/workloads.py), CPU only: the Gaussian random
field behind the log-normal coefficients has the squared-exponential covariance it states,
in the reference's cell order, and the contrast it is rescaled to."""
import numpy as np

from jfem_amd.workloads import (SplitMix64, gaussian_field_raw, gaussian_random_field, se_covariance_1d,
                                sinusoid_terms, splitmix_normals)


def test_splitmix_normals_match_the_scalar_stream():
    """the vectorised normals are the scalar generator's Box-Muller pairs"""
    r = SplitMix64(77)
    ref = []
    for _ in range(5):
        u1, u2 = max(r.uniform(), 1e-300), r.uniform()
        ref.append(np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2))
    assert np.allclose(splitmix_normals(77, 5), ref, rtol=0, atol=1e-15)


def test_sinusoid_terms_ranges():
    kw, amp = sinusoid_terms(2, 16, 3)
    assert kw.shape == (16, 3) and kw.min() >= 1 and kw.max() <= 8 and amp.shape == (16,)
    assert np.array_equal(sinusoid_terms(2, 16, 3)[0], kw)


def test_gaussian_field_covariance():
    """the empirical covariance over 4000 seeds of the raw field on 5 x 9 cells equals
    K_0 (x) K_1 (last axis fastest) to Monte-Carlo accuracy (entry sd <= 0.023); the
    transposed Kronecker order, or twice the correlation length, is far off"""
    cells, length, N = (5, 9), 0.2, 4000
    G = np.stack([gaussian_field_raw(1000 + s, cells, length) for s in range(N)])
    C = G.T @ G / N
    K = np.kron(se_covariance_1d(5, length), se_covariance_1d(9, length))
    assert np.max(np.abs(C - K)) < 0.12
    assert np.max(np.abs(np.mean(G, axis=0))) < 0.12
    Kt = np.kron(se_covariance_1d(9, length), se_covariance_1d(5, length))
    assert np.max(np.abs(C - Kt)) > 0.3
    K2 = np.kron(se_covariance_1d(5, 2 * length), se_covariance_1d(9, 2 * length))
    assert np.max(np.abs(C - K2)) > 0.3


def test_gaussian_random_field_contrast_and_determinism():
    """max g - min g = ln(contrast) and mean 0, for the C3 (1e4, 3D) and C4 (1e3, 2D) settings;
    the same seed gives the same field"""
    for cells, contrast in (((12, 12, 12), 1e4), ((32, 32), 1e3)):
        g = gaussian_random_field(3, cells, 0.1, contrast)
        assert g.shape == (int(np.prod(cells)),)
        assert abs((g.max() - g.min()) - np.log(contrast)) < 1e-12
        assert abs(g.mean()) < 1e-12
        a = np.exp(g)
        assert abs(a.max() / a.min() / contrast - 1) < 1e-10
        assert np.array_equal(g, gaussian_random_field(3, cells, 0.1, contrast))
        assert not np.array_equal(g, gaussian_random_field(4, cells, 0.1, contrast))
