"""Host build of the assembly's shared-side logic (csrc/td_field.hpp,
compiled with g++ like the flop counter):

* shared_side on constructed triangle pairs: a side shared and run the other
  way, run the same way, one vertex shared, none, all three (a duplicate
  triangle), and degenerate inputs;
* td_pair_grad_shared over runs of consecutive source triangles against
  td_pair_grad (every side evaluated for every pair), on a planar mesh in
  its own order and a rough mesh with every third triangle's vertex order
  reversed (shared sides run both ways), half space: displacement gradients
  equal to 1e-12 of each pair's largest, for receivers at element centroids
  and at random points below the surface.  This pins the identity the
  kernel relies on (the free-surface correction of side BA is minus that of
  side AB, with the side frame's first two axes reversed) independently of
  the GPU.  Measured: 1.0e-13 (planar) and 5.1e-13 (rough, both
  orientations) at most, medians near 1e-15.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyquake3d_amd", "csrc")

PROG = r"""
#include <cstdio>
#include <cmath>
#include <vector>
#include "td_field.hpp"
using namespace pq3d;
int main(int argc, char** argv) {
    if (argc > 1 && argv[1][0] == 's') {
        // shared_side cases: A, B as 9 numbers each on stdin, one case per line
        double A[3][3], B[3][3];
        while (true) {
            for (int i = 0; i < 3; ++i) for (int k = 0; k < 3; ++k) if (scanf("%lf", &A[i][k]) != 1) return 0;
            for (int i = 0; i < 3; ++i) for (int k = 0; k < 3; ++k) if (scanf("%lf", &B[i][k]) != 1) return 0;
            int rev = -7;
            const int s = shared_side(A, B, &rev);
            printf("%d %d\n", s, rev);
        }
    }
    // pairs: stdin "ntri nrcv" then ntri x 9 vertex coordinates, nrcv x 3 receivers;
    // prints per (receiver, source) max |G_shared - G_plain| and max |G_plain|
    long long nt, nr;
    if (scanf("%lld %lld", &nt, &nr) != 2) return 1;
    std::vector<double> T(9 * nt), R(3 * nr);
    for (auto& v : T) if (scanf("%lf", &v) != 1) return 1;
    for (auto& v : R) if (scanf("%lf", &v) != 1) return 1;
    const double bT[2][3] = {{0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
    std::vector<SrcGeom<2>> g(nt);
    for (long long j = 0; j < nt; ++j) {
        src_setup<2>(&T[9 * j], &T[9 * j + 3], &T[9 * j + 6], bT, g[j]);
        double V[3][3], W[3][3];
        for (int v = 0; v < 3; ++v) for (int k = 0; k < 3; ++k) V[v][k] = T[9 * j + 3 * v + k];
        if (j > 0) {
            for (int v = 0; v < 3; ++v) for (int k = 0; k < 3; ++k) W[v][k] = T[9 * (j - 1) + 3 * v + k];
            g[j].prev_side = shared_side(V, W, &g[j].prev_rev);
        }
        if (j + 1 < nt) {
            int rev;
            for (int v = 0; v < 3; ++v) for (int k = 0; k < 3; ++k) W[v][k] = T[9 * (j + 1) + 3 * v + k];
            g[j].next_side = shared_side(V, W, &rev);
        }
        if (g[j].prev_side >= 0 && g[j].prev_side == g[j].next_side) g[j].prev_side = -1;
    }
    const double nu = 0.25;
    int shared = 0;
    for (long long j = 0; j < nt; ++j) shared += g[j].prev_side >= 0;
    printf("shared %d of %lld\n", shared, nt);
    for (long long r = 0; r < nr; ++r) {
        double carry[2][3][9];
        for (long long j = 0; j < nt; ++j) {
            const int cur = (int)(j & 1);
            double Gs[2][9], Gp[2][9];
            td_pair_grad_shared<double>(
                g[j], R[3 * r], R[3 * r + 1], R[3 * r + 2], nu, true, j > 0,
                [&](int k, int c, double v) { carry[cur][k][c] = v; },
                [&](int k, int c) { return carry[cur ^ 1][k][c]; }, Gs);
            td_pair_grad<double, 2, false>(g[j], R[3 * r], R[3 * r + 1], R[3 * r + 2], nu, true, Gp);
            double d = 0, m = 0;
            for (int b = 0; b < 2; ++b)
                for (int k = 0; k < 9; ++k) {
                    d = fmax(d, fabs(Gs[b][k] - Gp[b][k]));
                    m = fmax(m, fabs(Gp[b][k]));
                }
            printf("%.17g %.17g\n", d, m);
        }
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    d = tmp_path_factory.mktemp("shared")
    (d / "p.cpp").write_text(PROG)
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, str(d / "p.cpp"), "-o", str(d / "p")], check=True)
    return str(d / "p")


def test_shared_side_cases(exe):
    P = np.array([[0, 0, -1], [1, 0, -1], [0, 0, -2], [1, 0, -2], [2, 0, -1], [5, 5, -5]], float)
    tri = lambda a, b, c: np.concatenate([P[a], P[b], P[c]])
    cases = [
        (tri(0, 1, 2), tri(1, 0, 3), (0, 1)),   # side 0 (P0 -> P1), B runs P1 -> P0
        (tri(0, 1, 2), tri(0, 1, 3), (0, 0)),   # side 0, B runs it the same way
        (tri(1, 2, 0), tri(3, 0, 2), (1, 1)),   # side 1 of A (P2 -> P0), B runs P0 -> P2
        (tri(2, 0, 1), tri(1, 2, 3), (2, 0)),   # side 2 of A (P1 -> P2), B runs it the same way
        (tri(0, 1, 2), tri(1, 4, 3), (-1, 0)),  # one vertex shared
        (tri(0, 1, 2), tri(3, 4, 5), (-1, 0)),  # none
        (tri(0, 1, 2), tri(2, 0, 1), (-1, 0)),  # all three (a duplicate triangle)
        (tri(0, 0, 2), tri(0, 2, 3), (-1, 0)),  # degenerate A
    ]
    inp = "\n".join(" ".join(f"{v:.17g}" for v in np.r_[a, b]) for a, b, _ in cases)
    out = subprocess.run([exe, "s"], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    for (a, b, want), line in zip(cases, out):
        assert tuple(int(x) for x in line.split()) == want, (want, line)


def _run_pairs(exe, nodes, tri, rcv):
    T = np.concatenate([nodes[tri[:, 0] - 1], nodes[tri[:, 1] - 1], nodes[tri[:, 2] - 1]], axis=1)
    inp = f"{len(T)} {len(rcv)}\n" + "\n".join(" ".join(f"{v:.17g}" for v in row) for row in T) + "\n" + \
        "\n".join(" ".join(f"{v:.17g}" for v in row) for row in rcv)
    out = subprocess.run([exe, "p"], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    shared = int(out[0].split()[1])
    vals = np.array([[float(x) for x in line.split()] for line in out[1:] if line.strip()])
    return shared, vals


@pytest.mark.parametrize("kind", ["planar", "rough_mixed_orientation"])
def test_shared_pair_gradients_equal_every_side(exe, kind):
    from pyquake3d_amd import readmsh, synth
    if kind == "planar":
        nodes_km, tri = synth.planar_fault_mesh(6, 4, 12.0, 8.0)
    else:
        nodes_km, tri = synth.rough_fault_mesh(6, 4, length_km=12.0, depth_km=8.0, amp_km=0.5,
                                               wavelength_km=6.0, seed=4)
    tri = np.array(tri)
    if kind != "planar":
        tri[::3] = tri[::3, ::-1]
    nodes = nodes_km * 1e3
    _, xg = readmsh.get_eleVec(nodes, tri, False)
    rng = np.random.default_rng(2)
    rnd = np.c_[rng.uniform(-8e3, 8e3, 12), rng.uniform(-3e3, 3e3, 12), rng.uniform(-9e3, -100.0, 12)]
    rcv = np.r_[xg[::5], rnd]
    shared, vals = _run_pairs(exe, nodes, tri, rcv)
    assert shared >= 0.8 * len(tri)   # a run breaks where mesh order jumps to the next row
    d, m = vals[:, 0], vals[:, 1]
    ok = np.isfinite(m) & (m > 0)
    rel = d[ok] / m[ok]
    print(f"{kind}: {shared} of {len(tri)} sources share a side with the previous one; "
          f"max {rel.max():.2e}, median {np.median(rel):.2e} over {ok.sum()} pairs", flush=True)
    assert ok.sum() > 0.99 * len(m)
    assert rel.max() < 1e-12
