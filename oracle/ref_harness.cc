// ref_harness.cc -- C entry points over NimbleSM's own element, material and
// tensor code, compiled from the reference's sources (see oracle/Makefile,
// target "ref").  TEST INFRASTRUCTURE ONLY: the resulting oracle/_ref library
// is used by tests to pin the C restatement in nsm_oracle.c bit for bit.  It
// never ships with, and is never loaded by, the product.
//
// The wrappers reach the reference's protected *_impl templates through
// derived classes so that the Kokkos form of the element kernels (current
// coordinates x = X + u, src/nimble_element.h:430-502 and :539-625) is what
// gets exercised.
#include "nimble_element.h"
#include "nimble_material.h"
#include "nimble_utils.h"
#include "nimble_view.h"

#include <map>
#include <string>

namespace {

struct RefHex : public nimble::HexElement {
  void F(const double* X, const double* u, double* Fo) const
  {
    nimble::Viewify<2, const double> Xv(X, {8, 3}, {3, 1});
    nimble::Viewify<2, const double> uv(u, {8, 3}, {3, 1});
    nimble::Viewify<2>               Fv(Fo, {8, 9}, {9, 1});
    ComputeDeformationGradients_impl(Xv, uv, Fv);
  }
  void f(const double* X, const double* u, const double* s, double* fo) const
  {
    nimble::Viewify<2, const double> Xv(X, {8, 3}, {3, 1});
    nimble::Viewify<2, const double> uv(u, {8, 3}, {3, 1});
    nimble::Viewify<2, const double> sv(s, {8, 6}, {6, 1});
    nimble::Viewify<2>               fv(fo, {8, 3}, {3, 1});
    ComputeNodalForces_impl(Xv, uv, sv, fv);
  }
};

struct RefNeohookean : public nimble::NeohookeanMaterial {
  explicit RefNeohookean(const nimble::MaterialParameters& p) : nimble::NeohookeanMaterial(p) {}
  void stress(const double* F, double* s) const
  {
    nimble::Viewify<1, const double> null_view;
    nimble::Viewify<1, const double> fv(F, 9);
    GetStress(0.0, 0.0, null_view, fv, null_view, {s, 6});
  }
};

nimble::MaterialParameters make_params(double rho, double K, double G)
{
  std::map<std::string, std::string> sp;
  std::map<std::string, double>      dp;
  dp["density"]       = rho;
  dp["bulk_modulus"]  = K;
  dp["shear_modulus"] = G;
  return nimble::MaterialParameters("neohookean", sp, dp, 8);
}

}  // namespace

extern "C" {

void ref_deformation_gradients(long E, const double* Xg, const double* ug, double* F)
{
  RefHex h;
  for (long e = 0; e < E; e++) h.F(Xg + 24 * e, ug + 24 * e, F + 72 * e);
}

void ref_nodal_forces(long E, const double* Xg, const double* ug, const double* s, double* f)
{
  RefHex h;
  for (long e = 0; e < E; e++) h.f(Xg + 24 * e, ug + 24 * e, s + 48 * e, f + 24 * e);
}

void ref_lumped_mass(long E, double rho, const double* Xg, double* m)
{
  nimble::HexElement h;
  for (long e = 0; e < E; e++) h.ComputeLumpedMass(rho, Xg + 24 * e, m + 8 * e);
}

void ref_char_length(long E, const double* xg, double* out)
{
  nimble::HexElement h;
  for (long e = 0; e < E; e++) out[e] = h.ComputeCharacteristicLength(xg + 24 * e);
}

void ref_stress(long P, double K, double G, const double* F, double* s)
{
  RefNeohookean m(make_params(1.0, K, G));
  for (long p = 0; p < P; p++) m.stress(F + 9 * p, s + 6 * p);
}

void ref_polar_decomp(long P, const double* F, double* V, double* R)
{
  for (long p = 0; p < P; p++) ::Polar_Decomp(F + 9 * p, V + 6 * p, R + 9 * p);
}

double ref_invert_full33(const double* m, double* inv) { return ::Invert_Full33(m, inv); }

double ref_invert3x3(const double* m, double* inv)
{
  double a[3][3], ai[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i][j] = m[3 * i + j];
  double det = ::Invert3x3(a, ai);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) inv[3 * i + j] = ai[i][j];
  return det;
}

double ref_cos_of_acos_div3(double x) { return ::Cos_Of_Acos_Divided_By_3(x); }

}  // extern "C"
