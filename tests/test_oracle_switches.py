"""The 3D operator's switches in the oracle (magnetostatics_onelab.hpp:796-828,
1893-1941, 2359-2392, 3204-3240): the divergence gauge's codifferential
against the DEC oracle's, the gauge-plus-curl-curl operator against the
7-point vector Laplacian it reduces to on a unit grid, symmetry and
definiteness, the assembled matrices against the operators, and the
y-response sign."""
import numpy as np
import pytest

import oracle_lib as orc


def grid(shape, rng, uniform=False):
    if uniform:
        return [np.arange(float(n)) for n in shape]
    return [np.cumsum(np.concatenate([[rng.uniform(-1, 1)], rng.uniform(0.5, 1.5, n - 1)])) for n in shape]


def interior_mask(shape):
    nx, ny, nz = shape
    m = np.zeros((nz, ny, nx, 3), dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.ravel()


def dense(apply, n):
    return np.stack([apply(e) for e in np.eye(n)], axis=1)


@pytest.mark.parametrize("shape", [(5, 6, 7), (4, 4, 4), (9, 5, 6)])
def test_codifferential_matches_the_dec_oracle(shape):
    """δA of the divergence gauge (interior DOFs) is the DEC codifferential of
    the 1-form with its boundary DOFs zeroed, at every node"""
    rng = np.random.default_rng(sum(shape))
    coords = grid(shape, rng)
    nx, ny, nz = shape
    a = rng.standard_normal((nz, ny, nx, 3))
    a[~interior_mask(shape).reshape(a.shape)] = 0.0
    got = orc.ms_codifferential(coords, a.ravel()).reshape(nz, ny, nx)
    geom = orc.Geometry([nx, ny, nz], axes=coords)
    ref = orc.codifferential(geom, 1, np.transpose(a, (2, 1, 0, 3)).copy(), [nx, ny, nz])[..., 0]
    ref = np.transpose(ref, (2, 1, 0))
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_gauge_and_curl_curl_make_the_vector_laplacian():
    """unit spacing, ν = 1, gauge factor 1: curlᵀcurl + (grad div) gauge is
    the 7-point −Δ on each component wherever the stencils stay interior"""
    shape = (9, 9, 9)
    rng = np.random.default_rng(5)
    coords = grid(shape, rng, uniform=True)
    nx, ny, nz = shape
    a = np.zeros((nz, ny, nx, 3))
    a[2:-2, 2:-2, 2:-2] = rng.standard_normal((nz - 4, ny - 4, nx - 4, 3))
    with orc.ms_switches(divergence_gauge=True):
        out = orc.ms_apply(coords, np.ones(nx * ny * nz), 1.0, a.ravel()).reshape(a.shape)
    lap = 6 * a.copy()
    for ax in range(3):
        lap -= np.roll(a, 1, axis=ax) + np.roll(a, -1, axis=ax)
    core = (slice(2, -2),) * 3
    assert np.max(np.abs(out[core] - lap[core])) <= 1e-12 * np.max(np.abs(lap[core]))
    # without the switch the penalty term stays
    plain = orc.ms_apply(coords, np.ones(nx * ny * nz), 1.0, a.ravel()).reshape(a.shape)
    assert np.max(np.abs(plain[core] - lap[core])) > 1e-3


@pytest.mark.parametrize("gauge_sign", [1.0, -1.0])
def test_divergence_gauge_is_symmetric_and_definite(gauge_sign):
    shape = (5, 5, 6)
    rng = np.random.default_rng(7)
    coords = grid(shape, rng)
    n = 3 * np.prod(shape)
    mu = rng.uniform(1, 3, np.prod(shape))
    inner = interior_mask(shape)
    with orc.ms_switches(divergence_gauge=True, gauge_sign=gauge_sign):
        A = dense(lambda e: orc.ms_apply(coords, mu, 20.0, e), n)
        G = A - dense(lambda e: orc.ms_apply(coords, mu, 0.0, e), n)
    Ai, Gi = A[np.ix_(inner, inner)], G[np.ix_(inner, inner)]
    assert np.max(np.abs(Gi - Gi.T)) <= 1e-12 * np.max(np.abs(Gi))
    ev = np.linalg.eigvalsh(0.5 * (Gi + Gi.T))
    if gauge_sign > 0:
        assert ev.min() >= -1e-10 * ev.max()  # δᵀWδ: positive semidefinite
        eva = np.linalg.eigvalsh(0.5 * (Ai + Ai.T))
        assert eva.min() > 0  # curl-curl + grad-div on interior DOFs: definite
    else:
        assert ev.max() <= 1e-10 * abs(ev.min())


@pytest.mark.parametrize("switches", [dict(divergence_gauge=True), dict(divergence_gauge=True, gauge_sign=-1.0),
                                      dict(y_sign=-1.0), dict(y_sign=-1.0, divergence_gauge=True)])
def test_assembled_matrix_is_the_operator(switches):
    shape = (5, 6, 5)
    rng = np.random.default_rng(11)
    coords = grid(shape, rng)
    n = 3 * np.prod(shape)
    mu = rng.uniform(1, 3, np.prod(shape))
    with orc.ms_switches(**switches):
        rp, cc, vv = orc.ms_assemble_csr(coords, mu, 20.0)
        A = dense(lambda e: orc.ms_apply(coords, mu, 20.0, e), n)
        D = orc.ms_diagonal(coords, mu, 20.0)
    M = np.zeros((n, n))
    for r in range(n):
        M[r, cc[rp[r]:rp[r + 1]]] = vv[rp[r]:rp[r + 1]]
        assert np.all(np.diff(cc[rp[r]:rp[r + 1]]) > 0)
    assert np.max(np.abs(M - A)) <= 1e-12 * np.max(np.abs(A))
    assert np.max(np.abs(D - np.diag(A))) <= 1e-12 * np.max(np.abs(A))
    if switches.get("divergence_gauge"):
        assert max(rp[r + 1] - rp[r] for r in range(n)) <= 15


def test_y_response_sign_flips_the_y_moment():
    """y_sign = −1: curlᵀ diag(w ν (1, −1, 1)) curl, the same as the explicit
    argument of the linear apply and as the dense product of the moments"""
    shape = (5, 5, 5)
    rng = np.random.default_rng(2)
    coords = grid(shape, rng)
    mu = rng.uniform(1, 3, np.prod(shape))
    x = rng.standard_normal(3 * np.prod(shape))
    with orc.ms_switches(y_sign=-1.0):
        got = orc.ms_apply(coords, mu, 0.0, x)
    assert np.array_equal(got, orc.ms_apply(coords, mu, 0.0, x, y_sign=-1.0))
    plus = orc.ms_apply(coords, mu, 0.0, x)
    assert np.max(np.abs(got - plus)) > 1e-6 * np.max(np.abs(plus))


@pytest.mark.parametrize("switches", [dict(divergence_gauge=True), dict(y_sign=-1.0)])
def test_nonlinear_jacobian_assembly_and_linearisation(switches):
    """the nonlinear operator's Jacobian (assembled and applied) follows the
    switches: it is the finite-difference derivative of F"""
    shape = (5, 5, 6)
    rng = np.random.default_rng(4)
    coords = grid(shape, rng)
    nn = np.prod(shape)
    mu = rng.uniform(1, 3, nn) * 4e-7 * np.pi
    ferro = (rng.random(nn) < 0.5).astype(float)
    bb, hh = [0.0, 0.5, 1.0, 1.5, 2.0], [0.0, 80.0, 200.0, 900.0, 5000.0]
    op = orc.Nonlinear(coords, mu, ferro, bb, hh, 20.0) if hasattr(orc, "Nonlinear") else None
    if op is None:
        pytest.skip("oracle_lib has no Nonlinear wrapper")
    x = rng.standard_normal(3 * nn) * 0.3
    v = rng.standard_normal(3 * nn)
    with orc.ms_switches(**switches):
        jv = op.jacobian_apply(x, v)
        eps = 1e-6
        fd = (op.apply(x + eps * v) - op.apply(x - eps * v)) / (2 * eps)
        rp, cc, vv = op.jacobian_csr(x)
    assert np.max(np.abs(jv - fd)) <= 1e-5 * np.max(np.abs(jv))
    M = np.zeros((3 * nn, 3 * nn))
    for r in range(3 * nn):
        M[r, cc[rp[r]:rp[r + 1]]] = vv[rp[r]:rp[r + 1]]
    assert np.max(np.abs(M @ v - jv)) <= 1e-11 * np.max(np.abs(jv))
