// fem_wire.hpp -- FEMInputs / FEMOutputs data model and the binary stream the
// reference exchanges through `fem_solve(void*, void*)`.
//
// Stream rules (topo_opt/serialization.h): a u64 total length (header
// included) followed by the fields in declaration order; PODs are raw
// little-endian bytes (bool = 1 byte); C arrays element by element;
// std::vector / std::string as u64 count then elements; std::map as u64
// count then (key, value) pairs in key order (serialization.h:317-327,330,
// 360,481,511).  Field order: FEMDataBase (fem_interface.h:182-190),
// FEMInputs (:252-276), FEMOutputs (:373-381), grids (:109-160).
#pragma once
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace spg {

constexpr int kFemVersionMajor = 3;
constexpr int kFemVersionMinor = 1;

struct WireError : std::runtime_error {
  int code;
  WireError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

struct ForceOnNode { int32_t coord[3]; double f[3]; };
struct DirichletOnNode { int32_t coord[3]; int32_t axis; double value; };
struct Krylov { double tolerance = 1e-4; int32_t max_iterations = 500; int32_t restart_iterations = 0; bool print_residuals = false; };

struct ScalarBlock { int32_t base[3]; double data[64]; };
struct VectorBlock { int32_t base[3]; double data[64][4]; };
struct ScalarGrid { int32_t resolution[3] = {0, 0, 0}; std::vector<ScalarBlock> blocks; };
struct VectorGrid { int32_t resolution[3] = {0, 0, 0}; std::vector<VectorBlock> blocks; };

struct FemBase {
  int32_t version_major = kFemVersionMajor;
  int32_t version_minor = kFemVersionMinor;
  std::map<std::string, std::string> extra_strings;
  std::map<std::string, int32_t> extra_ints;
  std::map<std::string, double> extra_doubles;

  int32_t get_int(const std::string& k, int32_t d) const {
    auto it = extra_ints.find(k);
    return it == extra_ints.end() ? d : it->second;
  }
  double get_double(const std::string& k, double d) const {
    auto it = extra_doubles.find(k);
    return it == extra_doubles.end() ? d : it->second;
  }
  bool has_int(const std::string& k) const { return extra_ints.count(k) != 0; }
  // opaque solver state split into two 32-bit ints (fem_interface.h:193-209)
  void* get_state_ptr() const {
    auto lo = extra_ints.find("state_ptr_lower");
    auto hi = extra_ints.find("state_ptr_higher");
    if (lo == extra_ints.end() || hi == extra_ints.end()) return nullptr;
    uint64_t v = ((uint64_t)(uint32_t)hi->second << 32) + (uint64_t)(uint32_t)lo->second;
    return reinterpret_cast<void*>(v);
  }
  void set_state_ptr(void* p) {
    uint64_t v = reinterpret_cast<uint64_t>(p);
    extra_ints["state_ptr_lower"] = (int32_t)(uint32_t)(v & 0xffffffffull);
    extra_ints["state_ptr_higher"] = (int32_t)(uint32_t)(v >> 32);
  }
};

struct FemInputs : FemBase {
  double dx = 0;
  uint64_t preserved_buffer_size = 0;
  bool use_density_only = true;
  double global_mu = 0, global_lambda = 0;
  ScalarGrid density, mu, lambda;
  int32_t mg_level = 0, explicit_mg_level = 0, pre_and_post_smoothing_iter = 1, bottom_smoothing_iter = 0;
  double jacobi_damping = 0;
  int32_t resolution[3] = {0, 0, 0};
  std::vector<ForceOnNode> forces;
  std::vector<DirichletOnNode> dirichlet_nodes, dirichlet_cells;
  Krylov krylov;
  std::string caller_method;
};

struct FemOutputs : FemBase {
  VectorGrid displacements;
  ScalarGrid residual_field;
  std::string info;
  std::vector<double> residuals;
  int32_t error_code = 0;
  bool success = false;
};

// ------------------------------------------------------------------ reader
class WireReader {
 public:
  explicit WireReader(const void* data) : p_(static_cast<const uint8_t*>(data)) {
    if (!p_) throw WireError("null input buffer", -1);
    std::memcpy(&total_, p_, sizeof(uint64_t));
    if (total_ < sizeof(uint64_t)) throw WireError("corrupt stream header", -1);
    head_ = sizeof(uint64_t);
  }
  template <class T> T pod() {
    need(sizeof(T));
    T v;
    std::memcpy(&v, p_ + head_, sizeof(T));
    head_ += sizeof(T);
    return v;
  }
  void raw(void* dst, size_t n) {
    need(n);
    std::memcpy(dst, p_ + head_, n);
    head_ += n;
  }
  uint64_t count(size_t elem_bytes) {
    uint64_t n = pod<uint64_t>();
    if (elem_bytes && n > (total_ - head_) / elem_bytes) throw WireError("vector length exceeds stream", -1);
    return n;
  }
  std::string str() {
    uint64_t n = count(1);
    std::string s(n, '\0');
    if (n) raw(&s[0], n);
    return s;
  }
  size_t head() const { return head_; }
  uint64_t total() const { return total_; }

 private:
  void need(size_t n) const {
    if (head_ + n > total_) throw WireError("stream truncated", -1);
  }
  const uint8_t* p_;
  uint64_t total_ = 0;
  size_t head_ = 0;
};

// ------------------------------------------------------------------ writer
class WireWriter {
 public:
  // with `length_header` the stream opens with its total length (finalize)
  explicit WireWriter(bool length_header = true) {
    if (length_header) buf_.resize(sizeof(uint64_t), 0);
  }
  template <class T> void pod(const T& v) {
    size_t h = buf_.size();
    buf_.resize(h + sizeof(T));
    std::memcpy(&buf_[h], &v, sizeof(T));
  }
  void raw(const void* src, size_t n) {
    size_t h = buf_.size();
    buf_.resize(h + n);
    if (n) std::memcpy(&buf_[h], src, n);
  }
  void str(const std::string& s) {
    pod<uint64_t>(s.size());
    raw(s.data(), s.size());
  }
  const std::vector<uint8_t>& bytes() const { return buf_; }
  std::vector<uint8_t>& finalize() {
    uint64_t n = buf_.size();
    std::memcpy(&buf_[0], &n, sizeof(uint64_t));
    return buf_;
  }

 private:
  std::vector<uint8_t> buf_;
};

// ------------------------------------------------------------- field order
inline void read_base(WireReader& r, FemBase& b) {
  b.version_major = r.pod<int32_t>();
  b.version_minor = r.pod<int32_t>();
  if (b.version_major != kFemVersionMajor || b.version_minor > kFemVersionMinor)
    throw WireError("FEM interface version " + std::to_string(b.version_major) + "." +
                        std::to_string(b.version_minor) + " not supported (need 3.x, x<=1)",
                    -7);
  uint64_t n = r.count(16);
  for (uint64_t i = 0; i < n; ++i) {
    std::string k = r.str();
    b.extra_strings[k] = r.str();
  }
  n = r.count(12);
  for (uint64_t i = 0; i < n; ++i) {
    std::string k = r.str();
    b.extra_ints[k] = r.pod<int32_t>();
  }
  n = r.count(16);
  for (uint64_t i = 0; i < n; ++i) {
    std::string k = r.str();
    b.extra_doubles[k] = r.pod<double>();
  }
}

inline void write_base(WireWriter& w, const FemBase& b) {
  w.pod(b.version_major);
  w.pod(b.version_minor);
  w.pod<uint64_t>(b.extra_strings.size());
  for (auto& kv : b.extra_strings) { w.str(kv.first); w.str(kv.second); }
  w.pod<uint64_t>(b.extra_ints.size());
  for (auto& kv : b.extra_ints) { w.str(kv.first); w.pod(kv.second); }
  w.pod<uint64_t>(b.extra_doubles.size());
  for (auto& kv : b.extra_doubles) { w.str(kv.first); w.pod(kv.second); }
}

inline void read_scalar_grid(WireReader& r, ScalarGrid& g) {
  r.raw(g.resolution, sizeof(g.resolution));
  uint64_t n = r.count(12 + 64 * 8);
  g.blocks.resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    r.raw(g.blocks[i].base, 12);
    r.raw(g.blocks[i].data, 64 * 8);
  }
}
inline void write_scalar_grid(WireWriter& w, const ScalarGrid& g) {
  w.raw(g.resolution, 12);
  w.pod<uint64_t>(g.blocks.size());
  for (auto& b : g.blocks) { w.raw(b.base, 12); w.raw(b.data, 64 * 8); }
}
inline void read_vector_grid(WireReader& r, VectorGrid& g) {
  r.raw(g.resolution, 12);
  uint64_t n = r.count(12 + 64 * 32);
  g.blocks.resize(n);
  for (uint64_t i = 0; i < n; ++i) {
    r.raw(g.blocks[i].base, 12);
    r.raw(g.blocks[i].data, 64 * 32);
  }
}
inline void write_vector_grid(WireWriter& w, const VectorGrid& g) {
  w.raw(g.resolution, 12);
  w.pod<uint64_t>(g.blocks.size());
  for (auto& b : g.blocks) { w.raw(b.base, 12); w.raw(b.data, 64 * 32); }
}

inline void read_inputs(WireReader& r, FemInputs& in) {
  read_base(r, in);
  in.dx = r.pod<double>();
  in.preserved_buffer_size = r.pod<uint64_t>();
  in.use_density_only = r.pod<uint8_t>() != 0;
  if (in.use_density_only) {
    in.global_mu = r.pod<double>();
    in.global_lambda = r.pod<double>();
    read_scalar_grid(r, in.density);
  } else {
    read_scalar_grid(r, in.mu);
    read_scalar_grid(r, in.lambda);
  }
  in.mg_level = r.pod<int32_t>();
  in.explicit_mg_level = r.pod<int32_t>();
  in.pre_and_post_smoothing_iter = r.pod<int32_t>();
  in.bottom_smoothing_iter = r.pod<int32_t>();
  in.jacobi_damping = r.pod<double>();
  r.raw(in.resolution, 12);
  uint64_t n = r.count(36);
  in.forces.resize(n);
  for (auto& f : in.forces) { r.raw(f.coord, 12); r.raw(f.f, 24); }
  for (auto* v : {&in.dirichlet_nodes, &in.dirichlet_cells}) {
    n = r.count(24);
    v->resize(n);
    for (auto& d : *v) {
      r.raw(d.coord, 12);
      d.axis = r.pod<int32_t>();
      d.value = r.pod<double>();
    }
  }
  in.krylov.tolerance = r.pod<double>();
  in.krylov.max_iterations = r.pod<int32_t>();
  in.krylov.restart_iterations = r.pod<int32_t>();
  in.krylov.print_residuals = r.pod<uint8_t>() != 0;
  in.caller_method = r.str();
  if (r.head() != r.total()) throw WireError("trailing bytes after FEMInputs", -1);
}

inline void write_inputs(WireWriter& w, const FemInputs& in) {
  write_base(w, in);
  w.pod(in.dx);
  w.pod(in.preserved_buffer_size);
  w.pod<uint8_t>(in.use_density_only ? 1 : 0);
  if (in.use_density_only) {
    w.pod(in.global_mu);
    w.pod(in.global_lambda);
    write_scalar_grid(w, in.density);
  } else {
    write_scalar_grid(w, in.mu);
    write_scalar_grid(w, in.lambda);
  }
  w.pod(in.mg_level);
  w.pod(in.explicit_mg_level);
  w.pod(in.pre_and_post_smoothing_iter);
  w.pod(in.bottom_smoothing_iter);
  w.pod(in.jacobi_damping);
  w.raw(in.resolution, 12);
  w.pod<uint64_t>(in.forces.size());
  for (auto& f : in.forces) { w.raw(f.coord, 12); w.raw(f.f, 24); }
  for (auto* v : {&in.dirichlet_nodes, &in.dirichlet_cells}) {
    w.pod<uint64_t>(v->size());
    for (auto& d : *v) { w.raw(d.coord, 12); w.pod(d.axis); w.pod(d.value); }
  }
  w.pod(in.krylov.tolerance);
  w.pod(in.krylov.max_iterations);
  w.pod(in.krylov.restart_iterations);
  w.pod<uint8_t>(in.krylov.print_residuals ? 1 : 0);
  w.str(in.caller_method);
}

// FEMOutputs = head, `nblocks` displacement block records (int32 base[3],
// then 64 nodes x 4 doubles), tail
inline void write_outputs_head(WireWriter& w, const FemOutputs& o, uint64_t nblocks) {
  write_base(w, o);
  w.raw(o.displacements.resolution, 12);
  w.pod<uint64_t>(nblocks);
}
inline void write_outputs_tail(WireWriter& w, const FemOutputs& o) {
  write_scalar_grid(w, o.residual_field);
  w.str(o.info);
  w.pod<uint64_t>(o.residuals.size());
  for (double d : o.residuals) w.pod(d);
  w.pod(o.error_code);
  w.pod<uint8_t>(o.success ? 1 : 0);
}
inline void write_outputs(WireWriter& w, const FemOutputs& o) {
  write_outputs_head(w, o, o.displacements.blocks.size());
  for (auto& b : o.displacements.blocks) { w.raw(b.base, 12); w.raw(b.data, 64 * 32); }
  write_outputs_tail(w, o);
}

inline void read_outputs(WireReader& r, FemOutputs& o) {
  read_base(r, o);
  read_vector_grid(r, o.displacements);
  read_scalar_grid(r, o.residual_field);
  o.info = r.str();
  uint64_t n = r.count(8);
  o.residuals.resize(n);
  for (auto& d : o.residuals) d = r.pod<double>();
  o.error_code = r.pod<int32_t>();
  o.success = r.pod<uint8_t>() != 0;
}

}  // namespace spg
