// Minimal HDF5 access for the instance files (the reference's src_util/hdf5_io.hpp
// primitives: scalar double / int / string attributes, 1-D double datasets, matrices as a
// group with "r", "c" attributes and "data" (real) or "re" / "im" (complex) datasets,
// column-major). libhdf5 is loaded on first use (QED_HDF5_LIB, else the soname, else
// /opt/conda/lib), so the rest of the library does not depend on it. Errors throw
// std::runtime_error.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace qed {
namespace h5 {

// an open HDF5 object (file or group); closed by the destructor
class Node {
 public:
  Node() = default;
  Node(int64_t id, bool file) : id_(id), file_(file) {}
  Node(Node&& o) noexcept : id_(o.id_), file_(o.file_) { o.id_ = -1; }
  Node& operator=(Node&& o) noexcept;
  Node(const Node&) = delete;
  Node& operator=(const Node&) = delete;
  ~Node();
  int64_t id() const { return id_; }

 private:
  int64_t id_ = -1;
  bool file_ = false;
};

// opens the file read-write if it exists, else creates it
Node open_or_create(const std::string& path);
Node open_read(const std::string& path);

Node create_group(const Node& at, const std::string& name);
Node open_group(const Node& at, const std::string& name);
bool exists(const Node& at, const std::string& name);
void unlink(const Node& at, const std::string& name);

void write_attr(const Node& at, const std::string& name, double v);
void write_attr(const Node& at, const std::string& name, int v);
void write_attr(const Node& at, const std::string& name, const std::string& v);
bool attr_exists(const Node& at, const std::string& name);
double read_attr_double(const Node& at, const std::string& name);
int read_attr_int(const Node& at, const std::string& name);
std::string read_attr_string(const Node& at, const std::string& name);

void write_vec(const Node& at, const std::string& name, const std::vector<double>& v);
std::vector<double> read_vec(const Node& at, const std::string& name);

// r x c column-major; the complex form writes "re" and "im"
void write_mat(const Node& at, const std::string& name, int r, int c, const double* v);
void write_mat_complex(const Node& at, const std::string& name, int r, int c, const double* re, const double* im);
// real part of a matrix stored either way; r and c returned
std::vector<double> read_mat_real(const Node& at, const std::string& name, int& r, int& c);

}  // namespace h5
}  // namespace qed
