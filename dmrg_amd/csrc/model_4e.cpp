// The 4e superconductor chain of 4e.cpp: two sites per MPO tensor (16 local
// states in the eight Z2 x Z2 x Z2 sectors, dimension 2 each), graded
// fermionic MPOs of bond 9, and the correlation functions its main prints.
// The tables below are the reference's own data (4e.cpp:203-384): each row
// {sector in, sector out, state in, state out, sign} is one matrix element of
// an operator on the two-site space.
#include <cmath>

#include "model_4e.hpp"
#include "observables.hpp"

namespace dmrgx {

namespace {

using Row = int[5];

// 4e.cpp:205-237 left c_up
const Row kLcu[32] = {{0, 4, 0, 0, 1},  {0, 4, 1, 1, 1},  {0, 5, 0, 0, 1},  {0, 5, 1, 1, -1}, {1, 4, 0, 0, 1},
                      {1, 4, 1, 1, -1}, {1, 5, 0, 0, 1},  {1, 5, 1, 1, 1},  {2, 6, 0, 0, 1},  {2, 6, 1, 1, -1},
                      {2, 7, 0, 0, 1},  {2, 7, 1, 1, 1},  {3, 6, 0, 0, 1},  {3, 6, 1, 1, -1}, {3, 7, 0, 0, 1},
                      {3, 7, 1, 1, 1},  {4, 0, 0, 0, 1},  {4, 0, 1, 1, 1},  {4, 1, 0, 0, -1}, {4, 1, 1, 1, 1},
                      {5, 0, 0, 0, -1}, {5, 0, 1, 1, 1},  {5, 1, 0, 0, 1},  {5, 1, 1, 1, 1},  {6, 2, 0, 0, 1},
                      {6, 2, 1, 1, -1}, {6, 3, 0, 0, -1}, {6, 3, 1, 1, 1},  {7, 2, 0, 0, -1}, {7, 2, 1, 1, -1},
                      {7, 3, 0, 0, 1},  {7, 3, 1, 1, 1}};
// 4e.cpp:239-271 left c_down
const Row kLcd[32] = {{0, 2, 0, 1, 1},  {0, 2, 1, 0, 1},  {0, 3, 0, 1, 1},  {0, 3, 1, 0, -1}, {1, 2, 0, 1, 1},
                      {1, 2, 1, 0, 1},  {1, 3, 0, 1, 1},  {1, 3, 1, 0, -1}, {2, 0, 0, 1, 1},  {2, 0, 1, 0, 1},
                      {2, 1, 0, 1, -1}, {2, 1, 1, 0, -1}, {3, 0, 0, 1, 1},  {3, 0, 1, 0, -1}, {3, 1, 0, 1, -1},
                      {3, 1, 1, 0, 1},  {4, 6, 0, 1, 1},  {4, 6, 1, 0, -1}, {4, 7, 0, 1, -1}, {4, 7, 1, 0, 1},
                      {5, 6, 0, 1, 1},  {5, 6, 1, 0, -1}, {5, 7, 0, 1, -1}, {5, 7, 1, 0, 1},  {6, 4, 0, 1, -1},
                      {6, 4, 1, 0, 1},  {6, 5, 0, 1, 1},  {6, 5, 1, 0, -1}, {7, 4, 0, 1, -1}, {7, 4, 1, 0, 1},
                      {7, 5, 0, 1, 1},  {7, 5, 1, 0, -1}};
// 4e.cpp:273-305 right c_up
const Row kRcu[32] = {{0, 4, 1, 0, -1}, {0, 4, 0, 1, 1},  {0, 5, 1, 0, 1},  {0, 5, 0, 1, 1},  {1, 4, 1, 0, 1},
                      {1, 4, 0, 1, 1},  {1, 5, 1, 0, -1}, {1, 5, 0, 1, 1},  {2, 6, 1, 0, 1},  {2, 6, 0, 1, 1},
                      {2, 7, 1, 0, -1}, {2, 7, 0, 1, 1},  {3, 6, 1, 0, 1},  {3, 6, 0, 1, 1},  {3, 7, 1, 0, -1},
                      {3, 7, 0, 1, 1},  {4, 0, 1, 0, 1},  {4, 0, 0, 1, -1}, {4, 1, 1, 0, -1}, {4, 1, 0, 1, -1},
                      {5, 0, 1, 0, -1}, {5, 0, 0, 1, -1}, {5, 1, 1, 0, 1},  {5, 1, 0, 1, -1}, {6, 2, 1, 0, 1},
                      {6, 2, 0, 1, 1},  {6, 3, 1, 0, -1}, {6, 3, 0, 1, -1}, {7, 2, 1, 0, -1}, {7, 2, 0, 1, 1},
                      {7, 3, 1, 0, 1},  {7, 3, 0, 1, -1}};
// 4e.cpp:307-339 right c_down
const Row kRcd[32] = {{0, 2, 0, 0, 1},   {0, 2, 1, 1, -1},  {0, 3, 0, 0, 1},   {0, 3, 1, 1, 1},   {1, 2, 0, 0, 1},
                      {1, 2, 1, 1, -1},  {1, 3, 0, 0, 1},   {1, 3, 1, 1, 1},   {2, 0, 0, 0, 1},   {2, 0, 1, 1, -1},
                      {2, 1, 0, 0, -1},  {2, 1, 1, 1, 1},   {3, 0, 0, 0, -1},  {3, 0, 1, 1, -1},  {3, 1, 0, 0, 1},
                      {3, 1, 1, 1, 1},   {4, 6, 0, 0, -1},  {4, 6, 1, 1, -1},  {4, 7, 0, 0, -1},  {4, 7, 1, 1, -1},
                      {5, 6, 0, 0, -1},  {5, 6, 1, 1, -1},  {5, 7, 0, 0, -1},  {5, 7, 1, 1, -1},  {6, 4, 0, 0, -1},
                      {6, 4, 1, 1, -1},  {6, 5, 0, 0, 1},   {6, 5, 1, 1, 1},   {7, 4, 0, 0, 1},   {7, 4, 1, 1, 1},
                      {7, 5, 0, 0, -1},  {7, 5, 1, 1, -1}};
// 4e.cpp:341-357 left c_up c_down
const Row kLp[16] = {{0, 6, 0, 1, 1},  {0, 7, 0, 1, -1}, {1, 6, 0, 1, 1},  {1, 7, 0, 1, -1},
                     {2, 4, 0, 1, -1}, {2, 5, 0, 1, 1},  {3, 4, 0, 1, -1}, {3, 5, 0, 1, 1},
                     {4, 2, 1, 0, 1},  {4, 3, 1, 0, -1}, {5, 2, 1, 0, 1},  {5, 3, 1, 0, -1},
                     {6, 0, 1, 0, -1}, {6, 1, 1, 0, 1},  {7, 0, 1, 0, -1}, {7, 1, 1, 0, 1}};
// 4e.cpp:359-375 right c_up c_down
const Row kRp[16] = {{0, 6, 0, 1, -1}, {0, 7, 0, 1, -1}, {1, 6, 0, 1, -1}, {1, 7, 0, 1, -1},
                     {2, 4, 1, 0, -1}, {2, 5, 1, 0, 1},  {3, 4, 1, 0, -1}, {3, 5, 1, 0, 1},
                     {4, 2, 0, 1, 1},  {4, 3, 0, 1, -1}, {5, 2, 0, 1, 1},  {5, 3, 0, 1, -1},
                     {6, 0, 1, 0, 1},  {6, 1, 1, 0, -1}, {7, 0, 1, 0, -1}, {7, 1, 1, 0, 1}};
// 4e.cpp:377-381 c c c c
const Row kPp[4] = {{0, 0, 0, 0, 1}, {0, 1, 0, 0, -1}, {1, 0, 0, 0, 1}, {1, 1, 0, 0, -1}};
// 4e.cpp:383-391 left / right n_up + n_down - 1
const Row kLn[8] = {{0, 1, 0, 0, -1}, {1, 0, 0, 0, -1}, {2, 3, 0, 0, -1}, {3, 2, 0, 0, -1},
                    {4, 5, 1, 1, -1}, {5, 4, 1, 1, -1}, {6, 7, 1, 1, -1}, {7, 6, 1, 1, -1}};
const Row kRn[8] = {{0, 1, 0, 0, -1}, {1, 0, 0, 0, -1}, {2, 3, 1, 1, -1}, {3, 2, 1, 1, -1},
                    {4, 5, 0, 0, -1}, {5, 4, 0, 0, -1}, {6, 7, 1, 1, 1},  {7, 6, 1, 1, 1}};

using Grid = vector<vector<tensor<cplx>>>;

Grid grid(int D) { return Grid(8, vector<tensor<cplx>>(8, tensor<cplx>(D, 2, D, 2))); }

// the 8 x 8 sector blocks into a graded operator (4e.cpp:478-485): block
// (0, j, 0, k) in the order k outer, j inner; dirs (0, -1, 0, 1); zero blocks dropped
qtensor<cplx> assemble(Grid& t, int D) {
  qtensor<cplx> r(1, 8, 1, 8);
  r.u_dim(D, 2, D, 2);
  for (int k = 0; k < 8; ++k)
    for (int j = 0; j < 8; ++j) r.update(t[j][k], 0, j, 0, k);
  r.add_sign(0, -1, 0, 1);
  return r.simplify();
}

// a bond-1 operator from one table (4e.cpp:587-779); `flip` swaps the roles
// of in and out (the daggered pair operators)
qtensor<cplx> single_op(const Row* tab, int n, double coef, bool flip) {
  Grid t = grid(1);
  for (int k = 0; k < n; ++k) {
    const int* q = tab[k];
    if (flip) t[q[1]][q[0]].update(coef * q[4], 0, q[3], 0, q[2]);
    else t[q[0]][q[1]].update(coef * q[4], 0, q[2], 0, q[3]);
  }
  return assemble(t, 1);
}

}  // namespace

// H_4e1 (4e.cpp:397-487) and H_4e2 (4e.cpp:490-580): identical but for which
// of the left / right c tables carries the operator that opens a term
qtensor<cplx> H_4e(bool second, double t1, double t3, double f0, double f1, double p1, double frac) {
  const cplx i(0, 1);
  Grid t = grid(9);
  for (int k = 0; k < 8; ++k) {
    t[k][k].update(1, 0, 0, 0, 0);
    t[k][k].update(1, 0, 1, 0, 1);
    t[k][k].update(1, 8, 0, 8, 0);
    t[k][k].update(1, 8, 1, 8, 1);
  }
  // 2i t1 c^d c - f0 (2n-1)(2n-1): (8, 0)
  for (int k : {0, 1}) {
    t[k][k].update(-2 * f0, 8, 0, 0, 0);
    t[k][k].update(2 * f0, 8, 1, 0, 1);
  }
  for (int k : {6, 7}) {
    t[k][k].update(2 * f0, 8, 0, 0, 0);
    t[k][k].update(-2 * f0, 8, 1, 0, 1);
  }
  for (int k : {2, 3}) {
    t[k][k].update(-2. * i * t1, 8, 0, 0, 1);
    t[k][k].update(2. * i * t1, 8, 1, 0, 0);
  }
  for (int k : {4, 5}) {
    t[k][k].update(2. * i * t1, 8, 0, 0, 1);
    t[k][k].update(-2. * i * t1, 8, 1, 0, 0);
  }
  t[6][6].update(4. * i * t1, 8, 0, 0, 1);
  t[6][6].update(-4. * i * t1, 8, 1, 0, 0);
  const Row* ou = second ? kRcu : kLcu;  // opens c_u: (1,0) and c_u^d: (2,0)
  const Row* cu = second ? kLcu : kRcu;  // closes them: (8,2) and (8,1)
  const Row* od = second ? kRcd : kLcd;
  const Row* cd = second ? kLcd : kRcd;
  for (int k = 0; k < 32; ++k) {
    t[ou[k][0]][ou[k][1]].update(0.5 * ou[k][4], 1, ou[k][2], 0, ou[k][3]);
    t[cu[k][0]][cu[k][1]].update(i * (t3 * cu[k][4]), 8, cu[k][2], 2, cu[k][3]);
    t[ou[k][1]][ou[k][0]].update(0.5 * ou[k][4], 2, ou[k][3], 0, ou[k][2]);
    t[cu[k][1]][cu[k][0]].update(i * (t3 * cu[k][4]), 8, cu[k][3], 1, cu[k][2]);
    t[od[k][0]][od[k][1]].update(0.5 * od[k][4], 3, od[k][2], 0, od[k][3]);
    t[cd[k][0]][cd[k][1]].update(i * (t3 * cd[k][4]), 8, cd[k][2], 4, cd[k][3]);
    t[od[k][1]][od[k][0]].update(0.5 * od[k][4], 4, od[k][3], 0, od[k][2]);
    t[cd[k][1]][cd[k][0]].update(i * (t3 * cd[k][4]), 8, cd[k][3], 3, cd[k][2]);
  }
  const cplx pc = p1 * cplx(std::cos(2 * PI_REF * frac), std::sin(2 * PI_REF * frac));
  const cplx pd = p1 * cplx(std::cos(2 * PI_REF * frac), -std::sin(2 * PI_REF * frac));
  for (int k = 0; k < 16; ++k) {
    t[kLp[k][0]][kLp[k][1]].update(0.5 * kLp[k][4], 5, kLp[k][2], 0, kLp[k][3]);
    t[kRp[k][0]][kRp[k][1]].update(pc * (4.0 * kRp[k][4]), 8, kRp[k][2], 5, kRp[k][3]);
    t[kLp[k][1]][kLp[k][0]].update(0.5 * kLp[k][4], 6, kLp[k][3], 0, kLp[k][2]);
    t[kRp[k][1]][kRp[k][0]].update(pd * (4.0 * kRp[k][4]), 8, kRp[k][3], 6, kRp[k][2]);
  }
  for (int k = 0; k < 8; ++k) {
    t[kLn[k][0]][kLn[k][1]].update(kLn[k][4], 7, kLn[k][2], 0, kLn[k][3]);
    t[kRn[k][0]][kRn[k][1]].update(-4 * f1 * kRn[k][4], 8, kRn[k][2], 7, kRn[k][3]);
  }
  return assemble(t, 9);
}

qtensor<cplx> op_4e(Op4e which) {
  switch (which) {
    case Op4e::up1: return single_op(kLcu, 32, 0.5, false);
    case Op4e::down1: return single_op(kLcd, 32, 0.5, false);
    case Op4e::up2: return single_op(kRcu, 32, 0.5, false);
    case Op4e::down2: return single_op(kRcd, 32, 0.5, false);
    case Op4e::pair1: return single_op(kLp, 16, 0.5, false);
    case Op4e::pair2: return single_op(kRp, 16, 0.5, false);
    case Op4e::pairpair: return single_op(kPp, 4, 0.5, false);
    case Op4e::pair_dagger1: return single_op(kLp, 16, -0.5, true);
    case Op4e::pair_dagger2: return single_op(kRp, 16, -0.5, true);
  }
  DMRGX_CHECK(false, "op_4e: unknown operator");
  return {};
}

// 4e.cpp:49-59: `sites` MPO tensors (two physical sites each), H_4e1 on
// even and H_4e2 on odd tensors with the pairing phase frac * i / sites
mpo<cplx> mpo_4e(int sites, const double* p) {
  const double t1 = p[0], t3 = p[1], f0 = p[2], f1 = p[3], frac = p[4];
  mpo<cplx> m(sites);
  for (int i = 1; i < sites - 1; ++i) m.operators_[i] = H_4e(i % 2 == 1, t1, t3, f0, f1, f1, frac * i / sites);
  m.edge(H_4e(false, t1, t3, f0, f1, f1, 0).left(),
         H_4e(true, t1, t3, f0, f1, f1, frac * (sites - 1) / sites).right());
  return m;
}

// 4e.cpp:64-68: the perturbation's boundary term, the right half of H_4e1 on
// the first tensor and the left half of H_4e2 on the last
void edges_4e(int sites, const double* p, qtensor<cplx>& first, qtensor<cplx>& last) {
  const double t3 = p[1], f1 = p[3], frac = p[4] * (sites - 1) / sites;
  first = H_4e(false, 0, -t3, 0, f1, f1, frac).right();
  last = H_4e(true, 0, -t3, 0, f1, f1, frac).left();
}

// 4e.cpp:71-196
vector<Corr4e> correlations_4e(const mps<cplx>& m, bool fermion) {
  const int sites = m.size();
  qtensor<cplx> u1 = op_4e(Op4e::up1), d1 = op_4e(Op4e::down1), u2 = op_4e(Op4e::up2), d2 = op_4e(Op4e::down2);
  auto dagger = [](const qtensor<cplx>& a) {
    qtensor<cplx> r = a.exchange(1, 3);
    r.conjugate();
    return r;
  };
  qtensor<cplx> u1d = dagger(u1), d1d = dagger(d1), u2d = dagger(u2), d2d = dagger(d2);
  qtensor<cplx> p1 = op_4e(Op4e::pair1), p1d = op_4e(Op4e::pair_dagger1), p2 = op_4e(Op4e::pair2),
                p2d = op_4e(Op4e::pair_dagger2), pp = op_4e(Op4e::pairpair);
  qtensor<cplx> id = p1.id();
  vector<Corr4e> out;
  mpo<cplx> o(sites);
  auto all_id = [&]() {
    for (int i = 1; i < sites - 1; ++i) o.operators_[i] = id;
  };
  auto ends = [&](const char* label, const qtensor<cplx>& a, const qtensor<cplx>& b) {
    all_id();
    o.edge(a.left(), b.right());
    out.push_back({label, (double)sites, calc(m, o, fermion)});
  };
  ends("L1 R1", p1d, p2);
  ends("L1 R2", p1d, p1);
  ends("L2 R1", p2d, p2);
  ends("L2 R2", p2d, p1);
  ends("up: L1 R1", u1d, u2);
  ends("up: L1 R2", u1d, u1);
  ends("up: L2 R1", u2d, u2);
  ends("up: L2 R2", u2d, u1);
  ends("down: L1 R1", d1d, d2);
  ends("down: L1 R2", d1d, d1);
  ends("down: L2 R1", d2d, d2);
  ends("down: L2 R2", d2d, d1);
  out.push_back({"Normalization", (double)sites, calc(m, fermion)});
  for (int q = 1; q < sites - 1; ++q) {
    all_id();
    o.edge(id.left(), id.right());
    o.operators_[q] = p1;
    out.push_back({"Cooper-pair", 2.0 * q + 1, calc(m, o, fermion)});
    o.operators_[q] = p2;
    out.push_back({"Cooper-pair", 2.0 * q + 2, calc(m, o, fermion)});
  }
  for (int q = 1; q < sites - 1; ++q) {
    all_id();
    o.edge(id.left(), id.right());
    o.operators_[q] = pp;
    out.push_back({"pair-pair", 2.0 * q + 1.5, calc(m, o, fermion)});
    if (q < sites - 2) {
      o.operators_[q] = p2;
      o.operators_[q + 1] = p1;
      out.push_back({"pair-pair", 2.0 * q + 2.5, calc(m, o, fermion)});
    }
  }
  return out;
}

}  // namespace dmrgx
