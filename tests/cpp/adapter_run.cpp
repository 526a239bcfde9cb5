// One explicit run driven through the C++ adapter classes of
// include/nimblesm_amd.hpp (what a NimbleSM host would call), for the GPU
// test that compares it with the same run through the Python host layer.
//
// usage: adapter_run <n> <fp64|mixed> <steps> <dt_factor> <out> <rho1> <K1> <G1> <rho2> <K2> <G2>
//
// Mesh: the n^3 jittered unit cube (seed 1234, 10% jitter), blocks 1 and 2
// split at the middle z-layer; bottom face fixed, top face driven at v_z = 1;
// reaction on the top face.  Writes u, v, f ([N][3], mesh order) and the
// reaction history ([steps]) to <out> as raw doubles.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/nimblesm_amd.hpp"

int main(int argc, char** argv)
{
  if (argc != 12) {
    std::fprintf(stderr, "usage: adapter_run <n> <fp64|mixed> <steps> <dt_factor> <out> rho1 K1 G1 rho2 K2 G2\n");
    return 2;
  }
  const int     n         = std::atoi(argv[1]);
  const bool    fp64      = std::strcmp(argv[2], "fp64") == 0;
  const int64_t steps     = std::atoll(argv[3]);
  const double  dt_factor = std::atof(argv[4]);
  const char*   out_path  = argv[5];
  double        p[6];
  for (int i = 0; i < 6; i++) p[i] = std::strtod(argv[6 + i], nullptr);
  if (n < 2 || steps < 1) return 2;
  try {
    const int64_t    n1 = n + 1, N = n1 * n1 * n1, E = (int64_t)n * n * n;
    std::vector<double> X((size_t)(3 * N));
    std::vector<int>    conn((size_t)(8 * E));
    nimble_amd::check(nsm_jittered_cube(n, 1234, 0.1, X.data(), conn.data()));
    const int64_t half = (int64_t)(n / 2) * n * n;  // elements are k-slowest
    std::vector<nimble_amd::BlockSpec> blocks = {
        {1, std::vector<int>(conn.begin(), conn.begin() + 8 * half), nimble_amd::NeohookeanMaterial(p[0], p[1], p[2])},
        {2, std::vector<int>(conn.begin() + 8 * half, conn.end()), nimble_amd::NeohookeanMaterial(p[3], p[4], p[5])}};
    nimble_amd::ModelData model(X, blocks, fp64 ? NSM_FP64 : NSM_MIXED);
    std::vector<int> bottom, top;
    for (int64_t i = 0; i < n1 * n1; i++) {
      bottom.push_back((int)i);
      top.push_back((int)(i + n * n1 * n1));
    }
    for (int c = 0; c < 3; c++) model.AddBoundaryCondition(NSM_BC_PRESCRIBED_VELOCITY, bottom, c, 0.0);
    model.AddBoundaryCondition(NSM_BC_PRESCRIBED_VELOCITY, top, 2, 1.0);
    model.SetReactionNodes(top, 2);
    model.ComputeLumpedMass();
    const double dt = dt_factor * model.GetCriticalTimeStep();
    nimble_amd::ExplicitTimeIntegrator integ(model, 0.0, (double)steps * dt, steps, 0);
    integ.Initialize();
    integ.Integrate();
    std::vector<double> u = model.GetVectorNodeData(NSM_FIELD_DISPLACEMENT);
    std::vector<double> v = model.GetVectorNodeData(NSM_FIELD_VELOCITY);
    std::vector<double> f = model.GetVectorNodeData(NSM_FIELD_INTERNAL_FORCE);
    std::vector<double> R = model.GetReactionHistory(0, steps);
    std::FILE* fo = std::fopen(out_path, "wb");
    if (!fo) return 1;
    std::fwrite(u.data(), sizeof(double), u.size(), fo);
    std::fwrite(v.data(), sizeof(double), v.size(), fo);
    std::fwrite(f.data(), sizeof(double), f.size(), fo);
    std::fwrite(R.data(), sizeof(double), R.size(), fo);
    std::fclose(fo);
    std::printf("adapter_run: %s n=%d steps=%lld t=%.17g R_last=%.17g\n", argv[2], n, (long long)steps,
                integ.TimeCurrent(), R.back());
  } catch (const nimble_amd::Error& e) {
    std::fprintf(stderr, "adapter_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
