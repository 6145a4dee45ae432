"""The device Jacobi eigen-decomposition (csrc/ep_math.h, sym_eig) compiled
for the host with g++: eigenvalues and eigenvectors against numpy for
random symmetric matrices in FP64 and FP32, and finite results for matrices
with tiny entries and equal diagonals, where a rotation formula built on
β² + a_pq² underflows (the FP32 path decomposes b - I, whose entries can be
that small).  The FP64 step's fast form (one division per rotation, no
confirming sweep) is held to the same bars on random matrices and on
left Cauchy-Green tensors b = F Fᵀ near the identity, including repeated
eigenvalues."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "elastoplasm.jl_amd" / "csrc"

PROG = r"""
#define EP_CPU_CHECK 1
#include "ep_cpu_check.h"
#include "ep_math.h"
#include <cstdio>
template <int D, typename T, bool FAST = false> void run() {
    int n;
    if (scanf("%d", &n) != 1) return;
    for (int c = 0; c < n; c++) {
        T a[D][D], v[D][D];
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) { double x; if (scanf("%lf", &x) != 1) return; a[i][j] = T(x); }
        ep::sym_eig<D, T, FAST>(a, v);
        for (int i = 0; i < D; i++) printf("%.17g ", double(a[i][i]));
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) printf("%.17g ", double(v[i][j]));
        printf("\n");
    }
}
int main(int argc, char **argv) {
    const int d = argv[1][0] - '0';
    const char p = argv[2][0];   // f: FP32, d: FP64, x: FP64 fast form
    if (d == 2) { if (p == 'f') run<2, float>(); else if (p == 'x') run<2, double, true>(); else run<2, double>(); }
    else { if (p == 'f') run<3, float>(); else if (p == 'x') run<3, double, true>(); else run<3, double>(); }
}
"""


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("eig")
    src = d / "eig.cpp"
    src.write_text(PROG)
    out = d / "eig"
    subprocess.run(["g++", "-std=c++20", "-O2", f"-I{CSRC}", f"-I{ROOT / 'include'}", str(src), "-o", str(out),
                    "-lpthread"], check=True, capture_output=True)
    return out


def _eig(exe, mats, dim, prec):
    inp = f"{len(mats)}\n" + "\n".join(" ".join(f"{x:.17g}" for x in m.ravel()) for m in mats) + "\n"
    r = subprocess.run([str(exe), str(dim), prec], input=inp, capture_output=True, text=True, check=True)
    rows = [np.array(line.split(), dtype=np.float64) for line in r.stdout.strip().splitlines()]
    return [(row[:dim], row[dim:].reshape(dim, dim)) for row in rows]


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", ["d", "f", "x"])
def test_random_symmetric(exe, dim, prec):
    rng = np.random.default_rng(31 + dim)
    mats = []
    for _ in range(40):
        A = rng.normal(size=(dim, dim))
        mats.append(0.5 * (A + A.T) * 10.0 ** rng.uniform(-6, 2))
    tol = 2e-6 if prec == "f" else 1e-13
    for A, (lam, V) in zip(mats, _eig(exe, mats, dim, prec)):
        scale = np.abs(A).max()
        assert np.allclose(np.sort(lam), np.linalg.eigvalsh(A), atol=tol * scale)
        assert np.abs(V @ np.diag(lam) @ V.T - A).max() <= 10 * tol * scale
        assert np.abs(V.T @ V - np.eye(dim)).max() <= 10 * tol


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("prec", ["d", "f"])
def test_tiny_entries_stay_finite(exe, dim, prec):
    mats = []
    for e in (1e-23, 1e-30, 1e-38):
        A = np.full((dim, dim), e)
        np.fill_diagonal(A, 0.0)
        mats.append(A)
        B = np.eye(dim) * 2e-23 + A
        mats.append(B)
    for A, (lam, V) in zip(mats, _eig(exe, mats, dim, prec)):
        assert np.isfinite(lam).all() and np.isfinite(V).all(), A
        assert np.abs(V.T @ V - np.eye(dim)).max() <= 1e-5


@pytest.mark.parametrize("dim", [2, 3])
def test_fast_form_on_trial_b(exe, dim):
    """b = F Fᵀ with F = I + small (the step's trial b): the fast form's
    log(b) = V diag(log λ) Vᵀ matches scipy's matrix logarithm, also for
    repeated eigenvalues (pure stretches along the axes, F = diag)."""
    from scipy.linalg import logm
    rng = np.random.default_rng(7 + dim)
    mats = []
    for k in range(60):
        if k % 3 == 0:
            F = np.diag(1.0 + rng.normal(scale=1e-3, size=dim).round(4))
            F[0, 0] = F[-1, -1]   # a repeated eigenvalue
        else:
            F = np.eye(dim) + rng.normal(scale=10.0 ** rng.uniform(-6, -1), size=(dim, dim))
        mats.append(F @ F.T)
    for b, (lam, V) in zip(mats, _eig(exe, mats, dim, "x")):
        L = V @ np.diag(np.log(lam)) @ V.T
        assert np.abs(L - logm(b).real).max() <= 1e-13, b
        assert np.abs(V.T @ V - np.eye(dim)).max() <= 1e-13
