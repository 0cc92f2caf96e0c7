// sayram_host.h -- host-side types with the interfaces of Sayram-2D's
// geometry and physics classes, for building the GPU solver without Eigen or
// xtensor.
//
// Where the reference's own headers are available, GpuSolver.h works with
// them directly (it is a template over the Mesh, Equation and 2-D array
// types); these stand-ins keep the same names and member signatures:
//   Ind, BoundaryID, BCType        source/Mesh.h:21-24, source/BCTypes.h:13,16
//   Grid2D                         source/Grid2D.h:16-68
//   Mesh (the accessors the step uses)   source/Mesh.h:32-96, source/Mesh.cc:36-65
//   Equation                       source/Equation.h:27-75
//   Array2D ~ xt::xtensor<double,2>: shape (nx, ny), row-major
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

struct Ind {
  std::size_t i;
  std::size_t j;
};

enum class BoundaryID { XMIN, XMAX, YMIN, YMAX };
enum class BCType { Dirichlet, ZeroFlux };

class Array2D {
 public:
  Array2D() = default;
  explicit Array2D(std::array<std::size_t, 2> shape) { resize(shape); }
  void resize(std::array<std::size_t, 2> shape) {
    shape_ = shape;
    v_.assign(shape[0] * shape[1], 0.0);
  }
  void fill(double x) { v_.assign(v_.size(), x); }
  double &operator()(std::size_t i, std::size_t j) { return v_[i * shape_[1] + j]; }
  double operator()(std::size_t i, std::size_t j) const { return v_[i * shape_[1] + j]; }
  const std::array<std::size_t, 2> &shape() const { return shape_; }
  double *data() { return v_.data(); }
  const double *data() const { return v_.data(); }
  std::size_t size() const { return v_.size(); }

 private:
  std::array<std::size_t, 2> shape_{{0, 0}};
  std::vector<double> v_;
};

class Grid2D {
 public:
  std::vector<double> x_edges, y_edges;
  Grid2D() = default;
  Grid2D(std::vector<double> xe, std::vector<double> ye)
      : x_edges(std::move(xe)), y_edges(std::move(ye)) {
    validate();
  }
  std::size_t nx() const { return x_edges.size() - 1; }
  std::size_t ny() const { return y_edges.size() - 1; }
  void validate() const {
    if (x_edges.size() < 2) throw std::runtime_error("Grid2D: x_edges must have size >= 2.");
    if (y_edges.size() < 2) throw std::runtime_error("Grid2D: y_edges must have size >= 2.");
    for (std::size_t i = 0; i + 1 < x_edges.size(); ++i)
      if (!(x_edges[i + 1] > x_edges[i]))
        throw std::runtime_error("Grid2D: x_edges must be strictly increasing at i=" +
                                 std::to_string(i));
    for (std::size_t j = 0; j + 1 < y_edges.size(); ++j)
      if (!(y_edges[j + 1] > y_edges[j]))
        throw std::runtime_error("Grid2D: y_edges must be strictly increasing at j=" +
                                 std::to_string(j));
  }
};

class Mesh {
 public:
  Mesh(const Grid2D &g, double dt) : xe_(g.x_edges), ye_(g.y_edges), dt_(dt) {
    for (std::size_t i = 0; i + 1 < xe_.size(); ++i) x_.push_back(0.5 * (xe_[i] + xe_[i + 1]));
    for (std::size_t j = 0; j + 1 < ye_.size(); ++j) y_.push_back(0.5 * (ye_[j] + ye_[j + 1]));
  }
  std::size_t nx() const { return x_.size(); }
  std::size_t ny() const { return y_.size(); }
  const std::vector<double> &x() const { return x_; }
  const std::vector<double> &y() const { return y_; }
  double x(std::size_t i) const { return x_[i]; }
  double y(std::size_t j) const { return y_[j]; }
  double x_edge(std::size_t i) const { return xe_[i]; }
  double y_edge(std::size_t j) const { return ye_[j]; }
  double dx(std::size_t i) const { return xe_[i + 1] - xe_[i]; }
  double dy(std::size_t j) const { return ye_[j + 1] - ye_[j]; }
  double dt() const { return dt_; }
  double cell_area_dt(const Ind &ind) const { return dx(ind.i) * dy(ind.j) / dt(); }
  std::size_t flatten_cell_index(const Ind &ind) const { return ind.j * nx() + ind.i; }

 private:
  std::vector<double> xe_, ye_, x_, y_;
  double dt_;
};

class Equation {
 public:
  explicit Equation(const Mesh &m) {
    const std::array<std::size_t, 2> shape = {m.nx(), m.ny()};
    G_.resize(shape);
    Dxx_.resize(shape);
    Dyy_.resize(shape);
    Dxy_.resize(shape);
    inv_tau_.resize(shape);
    inv_tau_.fill(0.0);
  }
  virtual ~Equation() = default;
  double G(const Ind &ind) const { return G_(ind.i, ind.j); }
  double Dxx(const Ind &ind) const { return Dxx_(ind.i, ind.j); }
  double Dyy(const Ind &ind) const { return Dyy_(ind.i, ind.j); }
  double Dxy(const Ind &ind) const { return Dxy_(ind.i, ind.j); }
  double inv_tau(const Ind &ind) const { return inv_tau_(ind.i, ind.j); }
  virtual BCType bc_type(BoundaryID side) const = 0;
  virtual double init_f(const Ind &ind) const = 0;
  virtual void update(double t) = 0;
  virtual bool dirichlet_vertex_value(BoundaryID side, std::size_t i, std::size_t j, double t,
                                      double *out_value) const {
    (void)side; (void)i; (void)j; (void)t; (void)out_value;
    return false;
  }

 protected:
  Array2D G_, Dxx_, Dyy_, Dxy_, inv_tau_;
};
