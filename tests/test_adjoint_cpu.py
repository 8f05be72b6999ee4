"""Reverse-mode bounce derivatives (cuda-rt_amd/csrc/rt_adjoint.h) against
central finite differences of double-precision forwards, on the CPU.

The header is compiled with g++ through tests/native/adjoint_probe.cpp.
Each check perturbs every input of one primitive: a dropped term, a sign
error or a swapped operand in an adjoint shows up as a mismatch in the
component it touches. The Moller-Trumbore forward is also checked against
numpy's linear solve of o + t d = A + alpha (B-A) + beta (C-A).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
c_d = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    out = tmp_path_factory.mktemp("adj") / "libadjprobe.so"
    src = os.path.join(HERE, "native", "adjoint_probe.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", src, "-o", str(out)],
                   check=True)
    L = ctypes.CDLL(str(out))
    for f in ("probe_hit_fwd", "probe_hit_adj", "probe_frame_fwd", "probe_frame_adj", "probe_shade_fwd",
              "probe_shade_adj"):
        getattr(L, f).restype = None
    L.probe_shade_adj.argtypes = [c_d, ctypes.c_double, c_d, c_d]
    return L


def P(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(c_d)


def hit_case(rng):
    while True:
        A, B, C = rng.normal(size=(3, 3))
        w = rng.dirichlet([2, 2, 2])
        target = w[0] * A + w[1] * B + w[2] * C
        o = target + rng.normal(size=3) * 2
        d = target - o
        d /= np.linalg.norm(d)
        n = np.cross(B - A, C - A)
        if abs(np.dot(n, d)) > 0.3 * np.linalg.norm(n):
            return np.concatenate([A, B, C, o, d])


def fd_grad(f, x, h):
    g = np.zeros_like(x)
    for i in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (f(xp) - f(xm)) / (2 * h)
    return g


def test_hit_forward_is_linear_solve(probe):
    rng = np.random.default_rng(1)
    for _ in range(20):
        x = hit_case(rng)
        out = np.empty(3)
        probe.probe_hit_fwd(P(x), out.ctypes.data_as(c_d))
        A, B, C, o, d = x.reshape(5, 3)
        M = np.stack([-d, B - A, C - A], axis=1)
        t, al, be = np.linalg.solve(M, o - A)
        np.testing.assert_allclose(out, [t, al, be], rtol=1e-10, atol=1e-12)


def test_hit_adjoint_matches_fd(probe):
    rng = np.random.default_rng(2)
    for _ in range(20):
        x = hit_case(rng)
        bar = rng.normal(size=3)

        def f(xx):
            out = np.empty(3)
            probe.probe_hit_fwd(P(xx), out.ctypes.data_as(c_d))
            return float(np.dot(bar, out))

        g = np.empty(15)
        probe.probe_hit_adj(P(x), P(bar), g.ctypes.data_as(c_d))
        np.testing.assert_allclose(g, fd_grad(f, x, 1e-6), rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_frame_adjoint_matches_fd(probe, axis):
    # the smallest |component| picks the replaced axis; cover all three
    rng = np.random.default_rng(3 + axis)
    for _ in range(20):
        n = rng.normal(size=3)
        n[axis] *= 0.1
        n /= np.linalg.norm(n)
        u1, u2 = rng.uniform(0.05, 0.95, size=2)
        s = np.array([np.sqrt(u1) * np.cos(2 * np.pi * u2), np.sqrt(1 - u1), np.sqrt(u1) * np.sin(2 * np.pi * u2)])
        dbar = rng.normal(size=3)

        def f(nn):
            out = np.empty(3)
            probe.probe_frame_fwd(P(nn), P(s), out.ctypes.data_as(c_d))
            return float(np.dot(dbar, out))

        g = np.empty(3)
        probe.probe_frame_adj(P(n), P(s), P(dbar), g.ctypes.data_as(c_d))
        np.testing.assert_allclose(g, fd_grad(f, n, 1e-6), rtol=1e-5, atol=1e-7)


def shade_case(rng, clamped):
    o = rng.normal(size=3)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    t = rng.uniform(0.5, 2.0)
    al, be = rng.dirichlet([2, 2, 2])[:2]
    q = o + t * d
    nrm = rng.normal(size=(3, 3)) * 0.2 + rng.normal(size=3)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    nsm = nrm[0] * (1 - al - be) + nrm[1] * al + nrm[2] * be
    nsm /= np.linalg.norm(nsm)
    # light in the lit hemisphere
    lc = q + nsm * rng.uniform(0.3, 3.0) + rng.normal(size=3) * 0.2
    lp = lc + rng.uniform(-0.2, 0.2, size=3)
    # falloff 1/(c0 + c1 d + c2 d^2): above 1 (clamped to 1) or below 1
    c = (0.5, 0.0, 0.0) if clamped else (1.0, 0.5, 0.25)
    return np.concatenate([o, d, [t, al, be], nrm.ravel(), lp, lc, c])


@pytest.mark.parametrize("clamped", [False, True])
def test_shade_adjoint_matches_fd(probe, clamped):
    rng = np.random.default_rng(7 + clamped)
    checked = 0
    for _ in range(40):
        x = shade_case(rng, clamped)
        out = np.empty(4)
        probe.probe_shade_fwd(P(x), out.ctypes.data_as(c_d))
        if out[0] <= 1e-3:
            continue
        gbar = rng.normal()
        nbar = rng.normal(size=3)

        def f(xx):
            o = np.empty(4)
            probe.probe_shade_fwd(P(xx), o.ctypes.data_as(c_d))
            return gbar * o[0] + float(np.dot(nbar, o[1:]))

        a = np.empty(5)
        probe.probe_shade_adj(P(x), gbar, P(nbar), a.ctypes.data_as(c_d))
        g = fd_grad(f, x, 1e-6)
        # q enters only through o (q = o + t d, d and t fixed), so dL/dq = dL/do
        np.testing.assert_allclose(a[:3], g[0:3], rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(a[3:], g[7:9], rtol=1e-5, atol=1e-7)
        checked += 1
    assert checked >= 20
