// GpuSolver.h -- the reference's Solver class (source/Solver.h:18-82) on top of
// the gfx950 C ABI (include/sayram_gpu.h).
//
// Drop-in for source/main.cc: `Solver solver(m, &eq);` becomes
// `GpuSolver solver(m, &eq);` (class template argument deduction picks the
// Mesh and Equation types); update(), t() and f() keep their meaning:
//   GpuSolver(m, eqp)   Solver::Solver + init     source/Solver.cc:14-55
//   update()            Solver::update            source/Solver.cc:270-290
//   t()                 istep * dt                source/Solver.h:23
//   f(), f(ind)         the state, (nx, ny)       source/Solver.h:24-25
//
// It borrows the Mesh and the Equation like Solver does. After every step it
// calls Equation::update(t) (source/Solver.cc:287), gathers the per-cell G, D
// and 1/tau again through the Equation accessors, as update_Lambda does
// (source/Solver.cc:57-65, 288), and hands the library only the arrays that
// changed (Lambda = G*D is formed on the device); the Dirichlet vertex values
// are gathered at t = 0 and after every step at the new t, as
// fill_vertex_from_bcs does (source/Solver.cc:385-422). An Equation known to
// keep its coefficients fixed may pass coeffs_constant = true to skip the
// per-step gather.
//
// Errors become std::runtime_error, as in the reference (source/Solver.cc:393).
// A step the library rejects (SG_ERR_NOCONV: the solve did not converge, broke
// down or left non-finite values; f and the step count are unchanged) throws
// too: the reference's direct solve never yields an unconverged step.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "sayram_gpu.h"

#ifndef SAYRAM_GPU_ARRAY_DEFAULT
#include "sayram_host.h"
#define SAYRAM_GPU_ARRAY_DEFAULT Array2D
#endif

template <class MeshT, class EquationT, class ArrayT = SAYRAM_GPU_ARRAY_DEFAULT>
class GpuSolver {
 public:
  GpuSolver(const MeshT &m_in, EquationT *eqp, const sg_opts *opts = nullptr,
            bool coeffs_constant = false)
      : m(m_in), eq(*eqp), constant_(coeffs_constant) {
    nx_ = m.nx();
    ny_ = m.ny();
    std::vector<double> xe(nx_ + 1), ye(ny_ + 1);
    for (std::size_t i = 0; i <= nx_; ++i) xe[i] = m.x_edge(i);
    for (std::size_t j = 0; j <= ny_; ++j) ye[j] = m.y_edge(j);
    std::vector<double> f0(nx_ * ny_);
    for (std::size_t i = 0; i < nx_; ++i)
      for (std::size_t j = 0; j < ny_; ++j) f0[i * ny_ + j] = eq.init_f({i, j});
    gather_coeffs();
    sg_bc bc{};
    gather_bc(0.0, bc);
    sg_grid grid{(int)nx_, (int)ny_, xe.data(), ye.data()};
    sg_coeffs c{G_.data(), Dxx_.data(), Dxy_.data(), Dyy_.data(), itau_.data()};
    check(sg_create(&grid, &c, &bc, f0.data(), m.dt(), opts, &h_));
    f_.resize({nx_, ny_});
    for (std::size_t i = 0; i < nx_; ++i)
      for (std::size_t j = 0; j < ny_; ++j) f_(i, j) = f0[i * ny_ + j];
    stale_ = false;
  }
  GpuSolver(const GpuSolver &) = delete;
  GpuSolver &operator=(const GpuSolver &) = delete;
  ~GpuSolver() { sg_destroy(h_); }

  void update() {
    const int rc = sg_step(h_, 1);
    last_status_ = rc;
    check(rc);
    istep_ += 1;
    stale_ = true;
    eq.update(t());
    if (!constant_) {
      // only the arrays whose values changed cross to the device
      std::vector<double> *now[5] = {&G_, &Dxx_, &Dxy_, &Dyy_, &itau_};
      for (int k = 0; k < 5; ++k) prev_[k].swap(*now[k]);
      gather_coeffs();
      const double *p[5];
      bool any = false;
      for (int k = 0; k < 5; ++k) {
        const bool same = prev_[k].size() == now[k]->size() &&
                          std::memcmp(prev_[k].data(), now[k]->data(), prev_[k].size() * 8) == 0;
        p[k] = same ? nullptr : now[k]->data();
        any |= !same;
      }
      if (any) {
        sg_coeffs c{p[0], p[1], p[2], p[3], p[4]};
        check(sg_set_coeffs(h_, &c));
      }
    }
    sg_bc bc{};
    gather_bc(t(), bc);
    for (int s = 0; s < 4; ++s)
      if (bc.type[s] == SG_DIRICHLET) check(sg_set_bc_values(h_, s, bc.values[s]));
  }

  double t() const { return (double)istep_ * m.dt(); }

  const ArrayT &f() const {
    if (stale_) {
      host_.resize(nx_ * ny_);
      check(sg_get_f(h_, host_.data()));
      for (std::size_t i = 0; i < nx_; ++i)
        for (std::size_t j = 0; j < ny_; ++j) f_(i, j) = host_[i * ny_ + j];
      stale_ = false;
    }
    return f_;
  }
  template <class IndT>
  double f(const IndT &ind) const { return f()(ind.i, ind.j); }

  // diagnostics beyond the reference: Krylov iterations, residual, min f
  sg_step_stats stats() const {
    sg_step_stats s{};
    check(sg_stats(h_, &s));
    return s;
  }
  int last_status() const { return last_status_; }

 private:
  const MeshT &m;
  EquationT &eq;
  bool constant_;
  std::size_t nx_ = 0, ny_ = 0;
  std::int64_t istep_ = 0;
  sg_solver *h_ = nullptr;
  int last_status_ = SG_OK;
  std::vector<double> G_, Dxx_, Dxy_, Dyy_, itau_;
  std::vector<double> prev_[5];   // the arrays of the previous gather
  std::array<std::vector<double>, 4> bcv_;
  mutable ArrayT f_;
  mutable std::vector<double> host_;
  mutable bool stale_ = true;

  static void check(int rc) {
    if (rc != SG_OK) throw std::runtime_error(std::string("sayram_gpu: ") + sg_last_error());
  }

  void gather_coeffs() {
    const std::size_t N = nx_ * ny_;
    G_.resize(N); Dxx_.resize(N); Dxy_.resize(N); Dyy_.resize(N); itau_.resize(N);
    for (std::size_t i = 0; i < nx_; ++i)
      for (std::size_t j = 0; j < ny_; ++j) {
        const std::size_t q = i * ny_ + j;
        G_[q] = eq.G({i, j});
        Dxx_[q] = eq.Dxx({i, j});
        Dxy_[q] = eq.Dxy({i, j});
        Dyy_[q] = eq.Dyy({i, j});
        itau_[q] = eq.inv_tau({i, j});
      }
  }

  // Dirichlet vertex values of every Dirichlet side at time tt
  void gather_bc(double tt, sg_bc &bc) {
    using BID = decltype(BoundaryID::XMIN);
    const BID sides[4] = {BoundaryID::XMIN, BoundaryID::XMAX, BoundaryID::YMIN, BoundaryID::YMAX};
    for (int s = 0; s < 4; ++s) {
      const bool dir = eq.bc_type(sides[s]) == BCType::Dirichlet;
      bc.type[s] = dir ? SG_DIRICHLET : SG_ZEROFLUX;
      bc.values[s] = nullptr;
      if (!dir) continue;
      const std::size_t n = s < 2 ? ny_ + 1 : nx_ + 1;
      bcv_[s].resize(n);
      for (std::size_t k = 0; k < n; ++k) {
        const std::size_t vi = s == 0 ? 0 : s == 1 ? nx_ : k;
        const std::size_t vj = s == 2 ? 0 : s == 3 ? ny_ : k;
        double u = 0.0;
        if (!eq.dirichlet_vertex_value(sides[s], vi, vj, tt, &u))
          throw std::runtime_error("Dirichlet BC: missing value.");
        bcv_[s][k] = u;
      }
      bc.values[s] = bcv_[s].data();
    }
  }
};
