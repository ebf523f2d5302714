"""Accuracy of the ensemble RHS's exponential (csrc/rsf_math.hpp,
exp_clipped), built for the host from the same source with g++: fma, rint and
ldexp are exactly rounded on host and device alike, so the host results are
the device's bit for bit.  Checked against exp in 80-bit extended precision
over the clipped domain [-700, 700] (random arguments, the reduction's
boundaries k ln2 +- ln2 / 2, integers, zero): at most one ulp, and below half
an ulp on average."""
import os
import shutil
import subprocess

import numpy as np
import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyquake3d_amd", "csrc")

PROG = r"""
#include <cstdio>
#include <vector>
#include "rsf_math.hpp"
int main(int argc, char** argv) {
    FILE* f = std::fopen(argv[1], "rb");
    std::vector<double> x;
    double v;
    while (std::fread(&v, sizeof v, 1, f) == 1) x.push_back(v);
    std::fclose(f);
    FILE* g = std::fopen(argv[2], "wb");
    for (double a : x) {
        const double y = pq3d::exp_clipped(a);
        std::fwrite(&y, sizeof y, 1, g);
    }
    std::fclose(g);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_exp_clipped_within_one_ulp(tmp_path):
    src = tmp_path / "expcheck.cpp"
    src.write_text(PROG)
    exe = tmp_path / "expcheck"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, str(src), "-o", str(exe)],
                   check=True)
    rng = np.random.default_rng(11)
    k = np.arange(-1009, 1010)
    ln2 = np.log(2.0)
    x = np.concatenate([rng.uniform(-700.0, 700.0, 200000), rng.uniform(-1.0, 1.0, 20000),
                        np.nextafter(k * ln2 + ln2 / 2, 0.0), np.nextafter(k * ln2 - ln2 / 2, 0.0),
                        np.arange(-700.0, 701.0), [0.0, -0.0, 1e-300, -1e-300, 700.0, -700.0]])
    x = x[np.abs(x) <= 700.0]
    (tmp_path / "x.bin").write_bytes(x.astype(np.float64).tobytes())
    subprocess.run([str(exe), str(tmp_path / "x.bin"), str(tmp_path / "y.bin")], check=True)
    y = np.frombuffer((tmp_path / "y.bin").read_bytes(), dtype=np.float64)
    assert len(y) == len(x)
    ref = np.exp(x.astype(np.longdouble))
    ulp = np.spacing(ref.astype(np.float64)).astype(np.longdouble)
    err = np.abs(y.astype(np.longdouble) - ref) / ulp
    print(f"exp_clipped: max {float(err.max()):.3f} ulp, mean {float(err.mean()):.3f} ulp over {len(x)} arguments")
    assert float(err.max()) <= 1.0
    assert float(err.mean()) < 0.5
    assert y[x == 0.0].tolist() == [1.0] * int(np.sum(x == 0.0))
