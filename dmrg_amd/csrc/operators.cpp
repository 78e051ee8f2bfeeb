// Model MPOs, built exactly as src/operators.cpp builds them (block by block
// through qtensor::update), then uploaded to the device.
#include <cmath>

#include "qtensor.hpp"

namespace dmrgx {

// XYZ chain in Pauli matrices, Z2 = Sz parity (src/operators.cpp:361-391)
qtensor<cplx> H_XYZ(double Jx, double Jy, double Jz) {
  const cplx i(0, 1);
  tensor<cplx> t00(5, 1, 5, 1), t01(5, 1, 5, 1), t10(5, 1, 5, 1), t11(5, 1, 5, 1);
  t00.update(1.0, 0, 0, 0, 0);
  t00.update(1.0, 3, 0, 0, 0);
  t00.update(Jz, 4, 0, 3, 0);
  t00.update(1.0, 4, 0, 4, 0);
  t01.update(1.0, 1, 0, 0, 0);
  t01.update(-i, 2, 0, 0, 0);
  t01.update(Jx, 4, 0, 1, 0);
  t01.update(-i * Jy, 4, 0, 2, 0);
  t10.update(1.0, 1, 0, 0, 0);
  t10.update(i, 2, 0, 0, 0);
  t10.update(Jx, 4, 0, 1, 0);
  t10.update(i * Jy, 4, 0, 2, 0);
  t11.update(1.0, 0, 0, 0, 0);
  t11.update(-1.0, 3, 0, 0, 0);
  t11.update(-Jz, 4, 0, 3, 0);
  t11.update(1.0, 4, 0, 4, 0);
  qtensor<cplx> r(1, 2, 1, 2);
  r.update(t00, 0, 0, 0, 0);
  r.update(t01, 0, 0, 0, 1);
  r.update(t10, 0, 1, 0, 0);
  r.update(t11, 0, 1, 0, 1);
  return r;
}

// real Kitaev chain (src/operators.cpp:406-431), fermionic signs via add_sign
qtensor<double> H_Kitaev(double t, double p, double u) {
  tensor<double> t00(4, 1, 4, 1), t11(4, 1, 4, 1), t01(4, 1, 4, 1), t10(4, 1, 4, 1);
  t00.update(1, 0, 0, 0, 0);
  t00.update(1, 3, 0, 3, 0);
  t11.update(1, 0, 0, 0, 0);
  t11.update(-2 * u, 3, 0, 0, 0);
  t11.update(1, 3, 0, 3, 0);
  t01.update(1, 1, 0, 0, 0);
  t01.update(-p, 3, 0, 1, 0);
  t01.update(-t, 3, 0, 2, 0);
  t10.update(1, 2, 0, 0, 0);
  t10.update(t, 3, 0, 1, 0);
  t10.update(p, 3, 0, 2, 0);
  qtensor<double> r(1, 2, 1, 2);
  r.update(t00, 0, 0, 0, 0);
  r.update(t01, 0, 0, 0, 1);
  r.update(t10, 0, 1, 0, 0);
  r.update(t11, 0, 1, 0, 1);
  r.add_sign(0, -1, 0, 1);
  return r;
}

// chiral 3-state Potts chain, Z3 (src/operators.cpp:301-350), truncated pi
qtensor<cplx> H_Potts3(double f, double j, double p, double t) {
  auto diag_blk = [&](double phase) {
    tensor<cplx> x(4, 1, 4, 1);
    x.update(1.0, 0, 0, 0, 0);
    x.update(-2 * f * std::cos(phase), 3, 0, 0, 0);
    x.update(1.0, 3, 0, 3, 0);
    return x;
  };
  const cplx v1 = -j * cplx(std::cos(t), std::sin(t));
  const cplx v2 = -j * cplx(std::cos(t), -std::sin(t));
  auto s_blk = [&]() {
    tensor<cplx> x(4, 1, 4, 1);
    x.update(1.0, 1, 0, 0, 0);
    x.update(v1, 3, 0, 2, 0);
    return x;
  };
  auto sd_blk = [&]() {
    tensor<cplx> x(4, 1, 4, 1);
    x.update(1.0, 2, 0, 0, 0);
    x.update(v2, 3, 0, 1, 0);
    return x;
  };
  qtensor<cplx> r(1, 3, 1, 3);
  r.update(diag_blk(p), 0, 0, 0, 0);
  r.update(s_blk(), 0, 0, 0, 1);
  r.update(sd_blk(), 0, 0, 0, 2);
  r.update(sd_blk(), 0, 1, 0, 0);
  r.update(diag_blk(p + 2 * PI_REF / 3), 0, 1, 0, 1);
  r.update(s_blk(), 0, 1, 0, 2);
  r.update(s_blk(), 0, 2, 0, 0);
  r.update(sd_blk(), 0, 2, 0, 1);
  r.update(diag_blk(p + 4 * PI_REF / 3), 0, 2, 0, 2);
  return r;
}

// Z2 Potts / transverse Ising (src/operators.cpp:265-288)
qtensor<double> H_Potts2(double h, double j) {
  tensor<double> t00(3, 1, 3, 1), t11(3, 1, 3, 1), t01(3, 1, 3, 1), t10(3, 1, 3, 1);
  t00.update(1, 0, 0, 0, 0);
  t00.update(h, 2, 0, 0, 0);
  t00.update(1, 2, 0, 2, 0);
  t11.update(1, 0, 0, 0, 0);
  t11.update(-h, 2, 0, 0, 0);
  t11.update(1, 2, 0, 2, 0);
  t01.update(1, 1, 0, 0, 0);
  t01.update(j, 2, 0, 1, 0);
  t10.update(1, 1, 0, 0, 0);
  t10.update(j, 2, 0, 1, 0);
  qtensor<double> r(1, 2, 1, 2);
  r.update(t00, 0, 0, 0, 0);
  r.update(t01, 0, 0, 0, 1);
  r.update(t10, 0, 1, 0, 0);
  r.update(t11, 0, 1, 0, 1);
  return r;
}

// Ising without symmetry (src/operators.cpp:239-256)
qtensor<double> H_Ising(double h, double j) {
  tensor<double> x(3, 2, 3, 2);
  x.update(1, 0, 0, 0, 0);
  x.update(1, 0, 1, 0, 1);
  x.update(1, 1, 0, 0, 0);
  x.update(-1, 1, 1, 0, 1);
  x.update(h, 2, 0, 0, 1);
  x.update(h, 2, 1, 0, 0);
  x.update(j, 2, 0, 1, 0);
  x.update(-j, 2, 1, 1, 1);
  x.update(1, 2, 0, 2, 0);
  x.update(1, 2, 1, 2, 1);
  qtensor<double> r(1, 1, 1, 1);
  r.update(x, 0, 0, 0, 0);
  return r;
}

}  // namespace dmrgx
