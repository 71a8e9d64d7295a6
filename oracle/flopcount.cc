// flopcount.cc -- counts the algorithmic floating-point operations of one
// track-layer step of the fit by running the oracle restatement (oracle_core.h)
// on a counting scalar type.  TEST / MEASUREMENT INFRASTRUCTURE ONLY.
//
// Convention: +, -, * and / each count 1 FLOP (an FMA is 2: one * and one +);
// sqrt, sin, cos, log count 1 each but are also reported separately as
// "special" operations; negation, abs, comparisons and copies count 0.  The
// FP64 determinant of the 2x2 inverse (2 mul, 1 sub, 1 div) is added by hand
// because the restatement computes it in double, outside the counting type.
//
// Output: one line per configuration, "name flops special div sqrt trans".
#define ORC_NN 1
#include <cstdio>
#include <cmath>

struct CF;
struct Cnt { long add = 0, mul = 0, div = 0, sqrt = 0, trans = 0; };
static Cnt g;

struct CF {
  double v;
  CF() : v(0) {}
  CF(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator float() const { return (float)v; }
  CF& operator+=(CF o) { ++g.add; v += o.v; return *this; }
  CF& operator-=(CF o) { ++g.add; v -= o.v; return *this; }
  CF& operator*=(CF o) { ++g.mul; v *= o.v; return *this; }
};
inline CF operator+(CF a, CF b) { ++g.add; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { ++g.add; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { ++g.mul; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { ++g.div; return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline bool operator>(CF a, CF b) { return a.v > b.v; }
inline bool operator<=(CF a, CF b) { return a.v <= b.v; }
inline bool operator>=(CF a, CF b) { return a.v >= b.v; }
namespace std {
inline CF sqrt(CF a) { ++g.sqrt; return CF(std::sqrt(a.v)); }
inline CF sin(CF a) { ++g.trans; return CF(std::sin(a.v)); }
inline CF cos(CF a) { ++g.trans; return CF(std::cos(a.v)); }
inline CF log(CF a) { ++g.trans; return CF(std::log(a.v)); }
inline CF abs(CF a) { return CF(std::fabs(a.v)); }
inline CF hypot(CF a, CF b) { g.mul += 2; g.add += 1; ++g.sqrt; return CF(std::hypot(a.v, b.v)); }
}  // namespace std

#include "oracle_core.h"

using namespace orc;

static void report(const char* name, const Cnt& c) {
  // FP64 determinant of the 2x2 inverse is counted by the caller where it applies.
  const long flops = c.add + c.mul + c.div + c.sqrt + c.trans;
  std::printf("%s flops=%ld add=%ld mul=%ld div=%ld sqrt=%ld trans=%ld\n", name, flops, c.add, c.mul, c.div, c.sqrt,
              c.trans);
}

int main() {
  float rl[300 * 120], xi[300 * 120];
  for (int i = 0; i < 300 * 120; ++i) { rl[i] = 0.02f; xi[i] = 4.4e-5f; }
  Ctx ctx{3.8f, rl, xi};
  // A representative barrel step: track at r ~ 20 cm going to r ~ 26 cm.
  CF inPar[6] = {CF(14.1), CF(14.2), CF(3.0), CF(0.5), CF(0.79), CF(1.2)};
  CF inErr[21];
  for (int i = 0; i < 6; ++i) inErr[sym(i, i)] = CF(1e-3);
  int chg[1] = {1};
  CF msRad[1] = {CF(26.0)};
  CF outErr[21], outPar[6];
  int fail[1];
  CF msPar[3] = {CF(18.4), CF(18.3), CF(3.6)};
  CF msErr[6] = {CF(5e-5), CF(-5e-5), CF(5e-5), CF(0), CF(0), CF(1e-2)};
  CF chi2[1];
  CF uErr[21], uPar[6];

  for (int mat = 0; mat < 2; ++mat) {
    g = Cnt();
    CF r[1] = {std::hypot(msPar[0], msPar[1])};  // MkBase.h:55 target radius
    propagate_helix_to_r(inErr, inPar, chg, r, outErr, outPar, 1, mat ? PF_apply_material : 0, nullptr, fail, ctx);
    Cnt prop = g;
    g = Cnt();
    kalman_operation(KFO_Calculate_Chi2 | KFO_Update_Params, outErr, outPar, msErr, msPar, uErr, uPar, chi2);
    Cnt kal = g;
    kal.add += 1; kal.mul += 2; kal.div += 1;  // FP64 determinant and its reciprocal
    report(mat ? "barrel_propagate_material" : "barrel_propagate", prop);
    report("barrel_kalman_chi2_update", kal);
    Cnt tot;
    tot.add = prop.add + kal.add + 1;  // + chi2 accumulation
    tot.mul = prop.mul + kal.mul; tot.div = prop.div + kal.div; tot.sqrt = prop.sqrt + kal.sqrt;
    tot.trans = prop.trans + kal.trans;
    report(mat ? "barrel_step_material" : "barrel_step", tot);
  }
  {
    // Finding: the chi2 of a hit alone (kalmanPropagateAndComputeChi2's update-less operation) and the update
    // alone (UpdateWithLastHit / FindCandidates' m_update_param_foo).
    g = Cnt();
    kalman_operation(KFO_Calculate_Chi2, outErr, outPar, msErr, msPar, uErr, uPar, chi2);
    Cnt c2 = g;
    c2.add += 1; c2.mul += 2; c2.div += 1;
    report("barrel_kalman_chi2", c2);
    g = Cnt();
    kalman_operation(KFO_Update_Params, outErr, outPar, msErr, msPar, uErr, uPar, chi2);
    Cnt up = g;
    up.add += 1; up.mul += 2; up.div += 1;
    report("barrel_kalman_update", up);
  }
  {
    g = Cnt();
    CF z[1] = {CF(60.0)};
    CF ip[6] = {CF(5.0), CF(5.0), CF(30.0), CF(0.5), CF(0.79), CF(0.3)};
    propagate_helix_to_z(inErr, ip, chg, z, outErr, outPar, 1, PF_apply_material, nullptr, ctx);
    Cnt prop = g;
    g = Cnt();
    kalman_operation_endcap(KFO_Calculate_Chi2 | KFO_Update_Params, outErr, outPar, msErr, msPar, uErr, uPar, chi2);
    Cnt kal = g;
    kal.add += 1; kal.mul += 2; kal.div += 1;
    report("endcap_propagate_material", prop);
    report("endcap_kalman_chi2_update", kal);
  }
  return 0;
}
