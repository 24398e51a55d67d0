// A C++ caller of mkpp::host::MechanismRegistry and of a generated
// SolverKernels type, compiled with g++ (no HIP compiler) against the
// header-only mkpp_host layer and linked to libmkpp_amd.so.
//
// usage: registry_main NAME IN OUT
//   IN:  the c_abi_main input format (int64 cells, LayoutLeft conc, temp,
//        pressure, air density, LayoutLeft rate parameters)
//   OUT: four runs, each int32 return flag then the integrated conc
//        (LayoutLeft) and int32 status per cell:
//     1. MechanismRegistry::execute on device memory: LayoutLeft state,
//        per-cell temperature and density, per-cell forcing given as a
//        LayoutRight [cells x parameters] view (strides P, 1);
//     2. the same through SolverKernels<>::integrate of the generated header;
//     3. MechanismRegistry::execute on host memory: a LayoutRight state
//        (strides species, 1), the scalar temperature and density of cell 0
//        and cell 0's forcing as the shared jvals vector;
//     4. an unknown mechanism name and a mismatched per-cell extent, which
//        must throw (flag 1 when both did).
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <vector>

#include "mkpp_host/mechanism_manager.hpp"
#include "saprc99.hpp"

using mkpp::host::BasicHostExecutionParams;
using mkpp::host::DeviceSpace;
using mkpp::host::HostSpace;
using mkpp::host::MechanismRegistry;
using mkpp::host::View;

#define CHECK(x)                                                                  \
  do {                                                                            \
    if ((x) != hipSuccess) {                                                      \
      std::fprintf(stderr, "HIP call failed: %s (line %d)\n", #x, __LINE__);      \
      std::exit(3);                                                               \
    }                                                                             \
  } while (0)

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  const std::string name = argv[1];
  FILE* in = std::fopen(argv[2], "rb");
  if (!in) return 2;
  int64_t n64 = 0;
  if (std::fread(&n64, sizeof n64, 1, in) != 1) return 2;
  const std::size_t n = static_cast<std::size_t>(n64);
  const auto info = MechanismRegistry::get_info(name);
  const std::size_t N = info.num_species;
  int ns = 0, nr = 0, nph = 0, npar = 0, env = 0;
  mkpp_amd_mechanism_info(name.c_str(), &ns, &nr, &nph, &npar, &env);
  const std::size_t P = npar;
  std::vector<double> conc(N * n), T(n), pres(n), rho(n), jv(P * n);
  auto rd = [&](std::vector<double>& v) { return std::fread(v.data(), sizeof(double), v.size(), in) == v.size(); };
  if (!rd(conc) || !rd(T) || !rd(pres) || !rd(rho) || (P && !rd(jv))) return 2;
  std::fclose(in);
  std::printf("registry: %zu mechanisms, %s registered %d, %zu species, %d reactions, requires_env %d\n",
              MechanismRegistry::get_registered_mechanisms().size(), name.c_str(),
              (int)MechanismRegistry::has_mechanism(name), N, info.num_reactions, (int)info.requires_env);

  FILE* out = std::fopen(argv[3], "wb");
  auto emit = [&](int32_t ok, const std::vector<double>& c, const std::vector<int>& st) {
    std::fwrite(&ok, sizeof ok, 1, out);
    std::fwrite(c.data(), sizeof(double), c.size(), out);
    std::vector<int32_t> s(st.begin(), st.end());
    s.resize(n, -99);
    std::fwrite(s.data(), sizeof(int32_t), n, out);
  };

  // LayoutRight copy of the forcing: jr[cell * P + slot]
  std::vector<double> jr(P * n);
  for (std::size_t c = 0; c < n; ++c)
    for (std::size_t q = 0; q < P; ++q) jr[c * P + q] = jv[c + n * q];

  double *d_conc, *d_T, *d_rho, *d_j;
  CHECK(hipMalloc(&d_conc, sizeof(double) * N * n));
  CHECK(hipMalloc(&d_T, sizeof(double) * n));
  CHECK(hipMalloc(&d_rho, sizeof(double) * n));
  CHECK(hipMalloc(&d_j, sizeof(double) * (P ? P : 1) * n));
  CHECK(hipMemcpy(d_T, T.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  CHECK(hipMemcpy(d_rho, rho.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  CHECK(hipMemcpy(d_j, jr.data(), sizeof(double) * P * n, hipMemcpyHostToDevice));

  BasicHostExecutionParams<DeviceSpace> dp;
  dp.dt = 3600.0;
  dp.cell_temperature = View<const double, 1, DeviceSpace>(d_T, n);
  dp.cell_density = View<const double, 1, DeviceSpace>(d_rho, n);
  dp.cell_jvals = View<const double, 2, DeviceSpace>(d_j, {n, P}, {P, 1});
  View<double, 2, DeviceSpace> dstate(d_conc, n, N);  // LayoutLeft

  // 1. MechanismRegistry::execute on device memory
  {
    CHECK(hipMemcpy(d_conc, conc.data(), sizeof(double) * N * n, hipMemcpyHostToDevice));
    MechanismRegistry::execute(name, dstate, dp);
    std::vector<double> got(N * n);
    CHECK(hipMemcpy(got.data(), d_conc, sizeof(double) * N * n, hipMemcpyDeviceToHost));
    emit(1, got, MechanismRegistry::last_results(name, n).status);
  }
  // 2. the generated SolverKernels type
  {
    CHECK(hipMemcpy(d_conc, conc.data(), sizeof(double) * N * n, hipMemcpyHostToDevice));
    mkpp::generated::saprc99::SolverKernels<> kernels;
    kernels.integrate(dstate, dp);
    std::vector<double> got(N * n);
    CHECK(hipMemcpy(got.data(), d_conc, sizeof(double) * N * n, hipMemcpyDeviceToHost));
    emit(1, got, MechanismRegistry::last_results(name, n).status);
  }
  // 3. host memory, LayoutRight state, scalar environment, shared forcing
  {
    std::vector<double> hr(N * n);
    for (std::size_t c = 0; c < n; ++c)
      for (std::size_t s = 0; s < N; ++s) hr[c * N + s] = conc[c + n * s];
    std::vector<double> shared(P);
    for (std::size_t q = 0; q < P; ++q) shared[q] = jv[n * q];
    BasicHostExecutionParams<HostSpace> hp;
    hp.dt = 3600.0;
    hp.temp = T[0];
    hp.air_density = rho[0];
    hp.jvals = View<const double, 1, HostSpace>(shared.data(), P);
    View<double, 2, HostSpace> hstate(hr.data(), {n, N}, {N, 1});
    MechanismRegistry::execute(name, hstate, hp);
    std::vector<double> got(N * n);
    for (std::size_t c = 0; c < n; ++c)
      for (std::size_t s = 0; s < N; ++s) got[c + n * s] = hr[c * N + s];
    emit(1, got, MechanismRegistry::last_results(name, n).status);
  }
  // 4. errors the reference raises
  {
    int thrown = 0;
    try {
      MechanismRegistry::execute("no_such_mechanism", dstate, dp);
    } catch (const std::runtime_error&) {
      thrown += 1;
    }
    BasicHostExecutionParams<DeviceSpace> bad = dp;
    bad.cell_temperature = View<const double, 1, DeviceSpace>(d_T, n - 1);
    try {
      MechanismRegistry::execute(name, dstate, bad);
    } catch (const std::invalid_argument&) {
      thrown += 1;
    }
    emit(thrown == 2 ? 1 : 0, conc, {});
  }
  std::fclose(out);
  CHECK(hipFree(d_conc));
  CHECK(hipFree(d_T));
  CHECK(hipFree(d_rho));
  CHECK(hipFree(d_j));
  std::printf("registry_main: %zu cells done\n", n);
  return 0;
}
