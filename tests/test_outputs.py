"""The reference's command-line convention and text outputs, host only:
set_para_val (src/functions.cpp:46-80) through dmrgx_parse_args, and
print_energy / print_singular (src/useful.cpp:54-102) built from
dmrg_amd/csrc/output.cpp with the plain C++ compiler.  Expected strings are
written out by hand from those functions' stream settings (fixed, 10
decimals; the suffix in the default 6-significant-digit format)."""
import math
import shutil
import subprocess
from pathlib import Path

import pytest

from dmrg_amd import driver

SRC = Path(__file__).resolve().parents[1] / "dmrg_amd" / "csrc"


def test_parse_args_known_answers():
    L, m, n, s, p, suf = driver.parse_args(["xyz", "l=100", "m=512", "n=8", "jx=1", "jy=0.8", "jz=0.5"],
                                           ["jx", "jy", "jz"])
    assert (L, m, n, s, p) == (100, 512, 8, 0, [1.0, 0.8, 0.5])
    assert suf == "_L_100_m_512_n_8_s_0_jx_1_jy_0.8_jz_0.5_.txt"
    # defaults of the mains: l=4, m=10, n=-1, s=0, first parameter 1, others 0
    L, m, n, s, p, suf = driver.parse_args(["kitaev"], ["t", "p", "u"])
    assert (L, m, n, s, p) == (4, 10, -1, 0, [1.0, 0.0, 0.0])
    assert suf == "_L_4_m_10_n_-1_s_0_t_1_p_0_u_0_.txt"
    # integer keys truncate (atof then int); 6 significant digits in the suffix;
    # words without '=' change nothing; the last assignment wins
    L, m, n, s, p, suf = driver.parse_args(
        ["potts", "l=12.7", "m=81", "f=1", "s=1", "jx", "t=3.14159265359", "p=1e-7", "j=2", "j=0.25", "n=3"],
        ["f", "j", "p", "t"])
    assert (L, m, n, s) == (12, 81, 3, 1)
    assert p == [1.0, 0.25, 1e-7, 3.14159265359]
    assert suf == "_L_12_m_81_n_3_s_1_f_1_j_0.25_p_1e-07_t_3.14159_.txt"


DRIVER = r"""
#include <iostream>
#include <vector>
#include "output.hpp"
int main() {
  double e[6] = {-10.5, -9.25, -8.125, 1.0 / 3.0, 0.0, 2.0};
  dmrgx::write_energy_line(std::cout, 3, 4, e, 6);
  dmrgx::write_energy_line(std::cout, 0, 9, e, 2);
  std::vector<double> s = {0.8, 0.6};
  dmrgx::write_singular_line(std::cout, 1, 2, s, 1);
  std::vector<double> t(25);
  for (int i = 0; i < 25; ++i) t[i] = 1.0 / (i + 2);
  dmrgx::write_singular_line(std::cout, 5, 6, t, 30);
  std::cout << 0.1 << "\n";  // the stream's own settings are restored
  return 0;
}
"""


def test_output_lines(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    (tmp_path / "d.cpp").write_text(DRIVER)
    exe = tmp_path / "d"
    subprocess.run([gxx, "-O2", "-std=c++17", f"-I{SRC}", str(tmp_path / "d.cpp"), str(SRC / "output.cpp"), "-o",
                    str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    assert lines[0] == "3\t4\t-10.5000000000\t-9.2500000000\t-8.1250000000\t0.3333333333\t0.0000000000\t2.0000000000"
    assert lines[1] == "0\t9\t-10.5000000000\t-9.2500000000\t0\t0\t0\t0"
    ent = -(0.64 * math.log(0.64 + 1e-10) + 0.36 * math.log(0.36 + 1e-10))
    assert lines[2] == "1\t2\t%.10f\t0.8000000000\t0.6000000000" % ent + "\t0" * 18 + "\t%.10f" % (1 - 0.64)
    t = [1.0 / (i + 2) for i in range(25)]
    ent = -sum(x * x * math.log(x * x + 1e-10) for x in t)
    err = 1 - sum(x * x for x in t)  # cutoff 30 > 25 values: all of them
    assert lines[3] == "5\t6\t%.10f\t" % ent + "\t".join("%.10f" % x for x in t[:20]) + "\t%.10f" % err
    assert lines[4] == "0.1"
