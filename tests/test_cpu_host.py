"""CPU tests: the C ABI library loads and exports every symbol of
include/cosahip.h, host bookkeeping matches the reference's conventions, and
the oracle satisfies what the mathematics fixes (no GPU needed)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from cosa_amd import _lib
from cosa_amd import case as cases
from cosa_amd.config import CosahipConfig, hb_operator
from cosa_amd.layout import Cut, Patch, bcdata_of, bctopo_level, cutdata_of, cutopo_level, level_nodes
from oracle_bind import Oracle, OrCase, dp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "cosahip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cosahip_\w+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert header_functions() == sorted(_lib.exported_symbols())


def test_library_exports_every_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libcosahip.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name
    _lib.load()


def test_config_struct_layout():
    # 38 int32 members then 37 doubles, the viscous block (10 int32, 14
    # doubles), the CFL-ramp block (6 int32, 6 doubles) (include/cosahip.h)
    assert ctypes.sizeof(CosahipConfig) == 38 * 4 + 37 * 8 + 10 * 4 + 14 * 8 + 6 * 4 + 6 * 8
    assert CosahipConfig.gamma.offset == 152


def test_level_nodes_and_bc_records():
    assert level_nodes(193, 1) == 193 and level_nodes(193, 2) == 97 and level_nodes(193, 3) == 49
    p = Patch((1, 1, 1), (193, 1, 5), 14)
    bd = bcdata_of(p)
    assert bd == [14, 2, 1, 1, 192, 1, 1, 1, 4]
    assert bctopo_level(bd, 2) == [14, 2, 1, 1, 96, 1, 1, 1, 2, 0]
    q = Patch((1, 49, 1), (193, 49, 5), 30)
    assert bctopo_level(bcdata_of(q), 3)[2] == 49 // 4 + 1


def test_cut_records_self_cut():
    c = Cut(1, (1, 1, 1), (1, 49, 5), 1, (193, 1, 1), (193, 49, 5), (1, 2, 3))
    a, b = cutdata_of(c)
    assert a[:9] == [1, 1, 1, 1, 1, 1, 48, 1, 4]
    assert a[9:18] == [1, 1, 193, 193, 193, 1, 48, 1, 4]
    assert a[18:] == [1, 2, 3] and b[18:] == [1, 2, 3]
    assert cutopo_level(a, 2)[11] == 193 // 2 + 1


def test_uneven_dims_seed0():
    dims = cases.uneven_dims(96, 0)
    cells = sum((a - 1) * (b - 1) * (c - 1) for a, b, c in dims)
    assert cells == 24285184
    assert min((a - 1) * (b - 1) * (c - 1) for a, b, c in dims) == 4096


@pytest.mark.parametrize("N", [1, 2, 4])
def test_hb_operator_closed_form(N):
    """mathb's D equals D_ij = 1/2 (-1)^(i-j) / sin(pi (i-j) / (2N+1)),
    D_ii = 0, for unit frequency; D is skew, D 1 = 0, and D differentiates
    trigonometric polynomials of degree <= N exactly."""
    om = 1.0
    D = hb_operator(N, om)
    n = 2 * N + 1
    ref = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                ref[i, j] = 0.5 * (-1) ** (i - j) / math.sin(math.pi * (i - j) / n)
    assert np.max(np.abs(D - ref)) < 2e-15
    assert np.max(np.abs(D + D.T)) < 2e-15
    assert np.max(np.abs(D @ np.ones(n))) < 1e-14
    t = np.arange(n) * 2 * math.pi / n
    for k in range(1, N + 1):
        assert np.max(np.abs(D @ np.sin(k * t) - k * np.cos(k * t))) < 1e-13


def test_oracle_tridi_against_dense_solve():
    o = Oracle()
    rng = np.random.default_rng(0)
    for n in (2, 5, 33):
        eps = rng.uniform(0.1, 1.0, n)
        a, b, c = -eps.copy(), 1 + 2 * eps, -eps.copy()
        a[0] = 0
        c[-1] = 0
        x = rng.uniform(-1, 1, n)
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        want = np.linalg.solve(A, x)
        got = x.copy()
        o.or_tridi(dp(a), dp(b), dp(c), dp(got), n)
        assert np.max(np.abs(got - want)) < 1e-14


def test_oracle_freestream_preserved():
    """Uniform flow on a closed distorted grid without walls: zero residual."""
    c = cases.two_block_case(ni=13, nj=9, nk=9, distort=0.3, wall=30, alpha_deg=0.0)
    o = Oracle()
    o.set_config(c.config)
    oc = OrCase(c)
    cp = oc.ptr()
    q = oc[1]["q"]
    o.or_cons2prim(cp, 1, dp(q))
    o.or_bc(cp, 1, dp(q))
    o.or_cutman_q(cp, 1, dp(q))
    o.or_residual(cp, 1)
    assert np.max(np.abs(oc[1]["res"])) < 1e-15


def test_oracle_roe_consistency():
    """F(q, q) is the physical flux times the face area (equal states give
    zero wave strengths)."""
    o = Oracle()
    c = cases.two_block_case(ni=9, nj=9, nk=5)
    o.set_config(c.config)
    dims = (5, 4, 4)
    n3 = 8 * 7 * 7
    rng = np.random.default_rng(1)
    q = np.empty((5, n3))
    q[0] = rng.uniform(0.8, 1.2, n3)
    q[1:4] = rng.uniform(-0.4, 0.4, (3, n3))
    q[4] = rng.uniform(0.6, 0.8, n3)
    qf = q.reshape(-1).copy()
    n2 = 7 * 6 * 6
    s = np.zeros((n2, 4))
    nv = rng.normal(size=(n2, 3))
    s[:, :3] = nv / np.linalg.norm(nv, axis=1)[:, None]
    s[:, 3] = rng.uniform(0.5, 2.0, n2)
    sf = s.reshape(-1).copy()
    flux = np.zeros(5 * 5 * 4 * 4)
    o.or_roflux_b(1, dp(qf), dp(qf), dp(flux), dp(sf), *dims)
    f = flux.reshape(5, 4, 4, 5)
    for k in range(1, 4):
        for j in range(1, 4):
            for i in range(1, 6):
                x = (i + 1) + 8 * ((j + 1) + 7 * (k + 1))
                m = i + 7 * (j + 6 * k)
                rho, u, v, w, p = q[:, x]
                kx, ky, kz, ds = s[m]
                un = u * kx + v * ky + w * kz
                h0 = 1.4 / 0.4 * p / rho + (u * u + v * v + w * w) / 2
                phys = np.array([rho * un, rho * u * un + p * kx, rho * v * un + p * ky, rho * w * un + p * kz,
                                 rho * h0 * un]) * ds
                got = f[:, k - 1, j - 1, i - 1]
                assert np.allclose(got, phys, rtol=1e-13, atol=1e-15)
