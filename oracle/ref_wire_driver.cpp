// ref_wire_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Links the reference's header-only wire format (topo_opt/fem_interface.h +
// topo_opt/serialization.h, included from /root/reference by oracle/Makefile)
// to produce and check byte streams:
//   wire_ref fixture <out> <scale>   FEMInputs of the interface test
//                                    (topo_opt/interface_test.cpp:45-166),
//                                    resolution divided by <scale>
//   wire_ref reencode <in> <out>     parse a FEMInputs stream with the
//                                    reference reader, write it back
//   wire_ref outputs <in>            parse a FEMOutputs stream, print fields
//   wire_ref fake <in> <out>         run the interface test's fake solver
//                                    (interface_test.cpp:7-34) on a stream
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "fem_interface.h"

using namespace fem_interface;

static std::vector<char> slurp(const char* path) {
  std::ifstream f(path, std::ios::binary);
  return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

static void dump(const char* path, const std::vector<uint8_t>& data) {
  FILE* f = fopen(path, "wb");
  fwrite(data.data(), 1, data.size(), f);
  fclose(f);
}

static std::vector<uint8_t> serialize_inputs(FEMInputs& param) {
  BinaryOutputSerializer ser;
  ser.initialize();
  param.io(ser);
  ser.finalize();
  return ser.data;
}

static int fixture(const char* out, int scale) {
  FEMInputs param;
  param.version_major = FEM_INTERFACE_VERSION_MAJOR;
  param.version_minor = FEM_INTERFACE_VERSION_MINOR;
  param.resolution[0] = 128 / scale;
  param.resolution[1] = 256 / scale;
  param.resolution[2] = 512 / scale;
  param.use_density_only = true;
  param.dx = 0.01f;
  for (int i = 0; i < 17; i++) {
    ForceOnNode f;
    f.coord[0] = i * 2 + 1;
    f.coord[1] = i * i + 14;
    f.coord[2] = i / 3 + 1;
    f.f[0] = std::sqrt(1.0 * i);
    f.f[1] = i * i;
    f.f[2] = i * i * i / 10000.0;
    param.forces.push_back(f);
  }
  DirichletOnNode d;
  d.coord[0] = 95; d.coord[1] = 6; d.coord[2] = 3; d.axis = 2; d.value = 0.1234;
  param.dirichlet_nodes.push_back(d);
  d.coord[0] = 94; d.coord[1] = 11; d.coord[2] = 16; d.axis = 1; d.value = 0.5432;
  param.dirichlet_nodes.push_back(d);
  param.krylov.tolerance = 1e-5;
  param.krylov.max_iterations = 10000;
  param.krylov.restart_iterations = 1000;
  param.krylov.print_residuals = true;
  param.caller_method = "topology_optimization";
  param.extra_ints["multigrid_num_pre_smoothing"] = 2;
  param.extra_ints["multigrid_num_post_smoothing"] = 3;
  param.extra_strings["log_file_path"] = "/tmp/fem_solver.log";
  param.extra_doubles["smoother_jacobi_damping"] = 0.66667f;
  param.global_mu = 2.0f;
  param.global_lambda = 1.0f;
  param.mg_level = 0;
  param.explicit_mg_level = 0;
  param.pre_and_post_smoothing_iter = 0;
  param.bottom_smoothing_iter = 0;
  param.jacobi_damping = 0;
  for (int i = 0; i < 3; i++) param.density.resolution[i] = param.resolution[i];
  int centers[3];
  for (int i = 0; i < 3; i++) centers[i] = param.resolution[i] / 2;
  int depth = 6;
  auto node_inside = [&](int* c) {
    for (int i = 0; i < 3; i++)
      if (centers[i] - depth < c[i] && c[i] < centers[i] + depth) return true;
    return false;
  };
  auto cell_inside = [&](int* c) {
    for (int i = 0; i < 3; i++)
      if (centers[i] - depth < c[i] && c[i] < centers[i] + depth - 1) return true;
    return false;
  };
  for (int i = 0; i < param.resolution[0]; i += 4)
    for (int j = 0; j < param.resolution[1]; j += 4)
      for (int k = 0; k < param.resolution[2]; k += 4) {
        bool active = false;
        for (int ii = 0; ii < 4; ii++)
          for (int jj = 0; jj < 4; jj++)
            for (int kk = 0; kk < 4; kk++) {
              int c[3] = {i + ii, j + jj, k + kk};
              if (node_inside(c)) active = true;
            }
        if (!active) continue;
        FEMInputs::ScalarGrid::Block b;
        b.base_coordinates[0] = i;
        b.base_coordinates[1] = j;
        b.base_coordinates[2] = k;
        b.reset_data();
        for (int ii = 0; ii < 4; ii++)
          for (int jj = 0; jj < 4; jj++)
            for (int kk = 0; kk < 4; kk++) {
              int c[3] = {i + ii, j + jj, k + kk};
              if (cell_inside(c)) b.get(ii, jj, kk) = i * jj + k - j + kk * ii;
            }
        param.density.blocks.push_back(b);
      }
  // FEMInterface::preserve_output (fem_interface.h:445-454)
  param.preserved_buffer_size = (1U << 20) + param.density.blocks.size() * sizeof(FEMOutputs::VectorGrid::Block);
  dump(out, serialize_inputs(param));
  printf("blocks=%zu\n", param.density.blocks.size());
  return 0;
}

static int reencode(const char* in, const char* out) {
  std::vector<char> raw = slurp(in);
  FEMInputs param;
  BinaryInputSerializer rd;
  rd.initialize(raw.data());
  param.io(rd);
  rd.finalize();
  dump(out, serialize_inputs(param));
  return 0;
}

static int outputs(const char* in) {
  std::vector<char> raw = slurp(in);
  FEMOutputs o;
  BinaryInputSerializer rd;
  rd.initialize(raw.data());
  rd("outputs", o);
  double sum = 0;
  for (auto& b : o.displacements.blocks)
    for (int e = 0; e < 64; ++e)
      for (int p = 0; p < 4; ++p) sum += b.data[e][p];
  printf("version=%d.%d blocks=%zu success=%d error_code=%d residuals=%zu info=%s checksum=%.17g head=%zu\n",
         o.version_major, o.version_minor, o.displacements.blocks.size(), (int)o.success, o.error_code,
         o.residuals.size(), o.info.c_str(), sum, rd.head);
  for (auto& kv : o.extra_ints) printf("extra_int %s=%d\n", kv.first.c_str(), kv.second);
  return 0;
}

static int fake(const char* in, const char* out) {
  std::vector<char> raw = slurp(in);
  FEMInputs param;
  BinaryInputSerializer rd;
  rd.initialize(raw.data());
  param.io(rd);
  FEMOutputs o;
  for (auto& db : param.density.blocks) {
    FEMOutputs::VectorGrid::Block vb;
    for (int i = 0; i < 3; i++) vb.base_coordinates[i] = db.base_coordinates[i];
    for (int ii = 0; ii < 4; ii++)
      for (int jj = 0; jj < 4; jj++)
        for (int kk = 0; kk < 4; kk++) {
          for (int p = 0; p < 3; p++) vb.get(ii, jj, kk)[p] = db.get(ii, jj, kk) + p;
          vb.get(ii, jj, kk)[3] = 0;
        }
    o.displacements.blocks.push_back(vb);
  }
  for (int i = 0; i < 3; i++) o.displacements.resolution[i] = param.resolution[i];
  for (int i = 0; i < 3; i++) o.residual_field.resolution[i] = 0;
  o.error_code = 0;
  o.success = true;
  BinaryOutputSerializer ser;
  ser.initialize();
  ser("outputs", o);
  ser.finalize();
  dump(out, ser.data);
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 3 && std::string(argv[1]) == "fixture") return fixture(argv[2], argc > 3 ? atoi(argv[3]) : 1);
  if (argc >= 4 && std::string(argv[1]) == "reencode") return reencode(argv[2], argv[3]);
  if (argc >= 3 && std::string(argv[1]) == "outputs") return outputs(argv[2]);
  if (argc >= 4 && std::string(argv[1]) == "fake") return fake(argv[2], argv[3]);
  fprintf(stderr, "usage: wire_ref fixture|reencode|outputs|fake ...\n");
  return 2;
}
