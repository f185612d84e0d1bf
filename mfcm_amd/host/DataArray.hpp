// DataArray.hpp -- host-side containers of the C++ mirror over libmfsv_hip.so.
//
// Mirrors the roles of the reference's DataArray<T> (DataArray/DataArray.hpp:32-257),
// Field<T, Mesh, Cell> (Field/Field.hpp:39-251) and Cartesian2DMesh
// (Mesh/Cartesian2DMesh.hpp:28-204) for the Saint-Venant hot path.  The device
// copy of a field lives in an mfsv context (one structure-of-arrays buffer per
// field); a DataArray here holds the host copy and moves it across with
// mfsv_upload / mfsv_download, which replace DataArray::move_to_device /
// move_to_host (DataArray/DataArray.cpp:84-116).
//
// Errors: every mfsv status code becomes the exception type the reference
// throws for the same kind of failure, after a message on std::cerr --
// std::logic_error for misuse (bad argument, call in the wrong state) and
// std::runtime_error for device/runtime failures.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/mfsv.h"

namespace mfsv {

// Status code of an mfsv call -> exception (no-op for MFSV_OK).
inline void check(int rc, const mfsv_ctx* ctx, const char* what) {
  if (rc == MFSV_OK) return;
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
  if (ctx) {
    const char* e = mfsv_last_error(ctx);
    if (e && *e) msg += ": " + std::string(e);
  }
  std::cerr << msg << std::endl;
  if (rc == MFSV_ERR_ARG || rc == MFSV_ERR_STATE) throw std::logic_error(msg);
  throw std::runtime_error(msg);
}

// Owns one mfsv context (one catchment on one HIP stream).  Shared between
// the solver and the fields bound to it, as the reference shares its
// sycl::queue.
class Context {
 public:
  Context(const mfsv_grid& grid, int precision, int device) {
    mfsv_ctx* c = nullptr;
    check(mfsv_create(&grid, precision, device, &c), nullptr, "mfsv_create");
    ctx_ = c;
  }
  ~Context() {
    if (ctx_) mfsv_destroy(ctx_);
  }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  mfsv_ctx* get() const { return ctx_; }

 private:
  mfsv_ctx* ctx_ = nullptr;
};

// Structured 2-D mesh (zero rotation): nx x ny cells of dx x dy, cell
// i = y*nx + x, x-faces before y-faces (Mesh/Cartesian2DMesh.hpp:379-460).
class Cartesian2DMesh {
 public:
  Cartesian2DMesh(int64_t nx, int64_t ny, double dx, double dy, double ox = 0.0, double oy = 0.0)
      : nx_(nx), ny_(ny), dx_(dx), dy_(dy), ox_(ox), oy_(oy) {
    if (nx < 1 || ny < 1 || !(dx > 0) || !(dy > 0))
      throw std::runtime_error("Cartesian2DMesh: cell counts and sizes must be positive");
  }
  int64_t nx() const { return nx_; }
  int64_t ny() const { return ny_; }
  double dx() const { return dx_; }
  double dy() const { return dy_; }
  size_t cell_count() const { return size_t(nx_ * ny_); }
  size_t face_count() const { return size_t((nx_ + 1) * ny_ + nx_ * (ny_ + 1)); }
  mfsv_grid grid() const { return mfsv_grid{nx_, ny_, dx_, dy_, ox_, oy_}; }

 private:
  int64_t nx_, ny_;
  double dx_, dy_, ox_, oy_;
};

// A 1-D array with a host copy and (when bound) a device copy held by a
// context as field `field_id`.  Host values are FP64 whatever the context's
// precision; an FP32 context rounds on upload.
template <typename T>
class DataArray {
 public:
  DataArray() = default;
  explicit DataArray(size_t n, T fill_value = T(0)) : host_(n, fill_value) {}

  size_t size() const { return host_.size(); }
  T& operator[](size_t i) { return host_[i]; }
  const T& operator[](size_t i) const { return host_[i]; }
  T* data() { return host_.data(); }
  const T* data() const { return host_.data(); }
  void fill(T v) { std::fill(host_.begin(), host_.end(), v); }

  void bind(std::shared_ptr<Context> ctx, int field_id) {
    ctx_ = std::move(ctx);
    field_ = field_id;
  }
  bool on_device() const { return ctx_ != nullptr; }

  // DataArray::move_to_device (DataArray.cpp:84-99): host -> device field.
  void move_to_device() {
    require_bound();
    static_assert(sizeof(T) == 8, "host copies are FP64");
    check(mfsv_upload(ctx_->get(), field_, reinterpret_cast<const double*>(host_.data()),
                      int64_t(host_.size())),
          ctx_->get(), "mfsv_upload");
  }
  // DataArray::move_to_host (DataArray.cpp:101-116): device field -> host.
  void move_to_host() {
    require_bound();
    check(mfsv_download(ctx_->get(), field_, reinterpret_cast<double*>(host_.data()),
                        int64_t(host_.size())),
          ctx_->get(), "mfsv_download");
  }

 private:
  void require_bound() const {
    if (!ctx_) throw std::logic_error("DataArray is not bound to a device field");
  }
  std::vector<T> host_;
  std::shared_ptr<Context> ctx_;
  int field_ = -1;
};

// A named cell field on the mesh (Field<T, Mesh, MeshComponent::Cell>).
template <typename T>
class Field {
 public:
  Field(std::string name, const Cartesian2DMesh& mesh, T fill_value = T(0))
      : name_(std::move(name)), data_(mesh.cell_count(), fill_value), nx_(mesh.nx()) {}
  const std::string& name() const { return name_; }
  DataArray<T>& data() { return data_; }
  const DataArray<T>& data() const { return data_; }
  size_t size() const { return data_.size(); }
  T& operator()(int64_t x, int64_t y) { return data_[size_t(y * nx_ + x)]; }
  const T& operator()(int64_t x, int64_t y) const { return data_[size_t(y * nx_ + x)]; }

 private:
  std::string name_;
  DataArray<T> data_;
  int64_t nx_;
};

// SpatialDerivativeOperator<T, Mesh, Cell, Minmod3> (SpatialDerivative/
// SpatialDerivative.hpp:48-53, 104-122): the Minmod3 slope of a cell field
// along X (west to east) or Y (the reference's orientation: north as the
// left neighbour, so the result is the slope toward decreasing y), evaluated
// on the context's device.
enum class SpatialDerivativeAxis { X = 0, Y = 1 };

class SpatialDerivativeOperator {
 public:
  template <SpatialDerivativeAxis A>
  static void apply(const Context& ctx, const Field<double>& in, Field<double>& out) {
    if (in.size() != out.size()) throw std::logic_error("SpatialDerivativeOperator: field sizes differ");
    check(mfsv_spatial_derivative(ctx.get(), int(A), in.data().data(), out.data().data(),
                                  int64_t(in.size())),
          ctx.get(), "mfsv_spatial_derivative");
  }
};

}  // namespace mfsv
