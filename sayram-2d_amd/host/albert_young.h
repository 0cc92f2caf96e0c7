// albert_young.h -- the reference's Albert & Young (2005) chorus cases as
// Equation inputs for the gfx950 step: G, D, f0, 1/tau and the boundary data of
// source/Cases/Albert_Young.{h,cc} and source/Cases/Albert_Young_LC.{h,cc},
// with the diffusion table of source/Cases/Albert_Young_IO.cc.
//
// Every quantity is evaluated with the operations of the reference in the same
// order (helpers below name the reference lines they restate), so a build of
// this file and the Python restatement in sayram_cases.albert_young_case give
// identical inputs on one machine.
#pragma once
#include <cfloat>
#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <vector>

#include "sayram_host.h"

namespace ay {

constexpr double kPi = 3.141592653589793238462;   // source/common.h:39
constexpr double kE0 = 0.511875;                  // MeV, source/common.h:42
constexpr double kC = 1;                          // source/common.h:41
constexpr double kME = kE0 / (kC * kC);           // source/common.h:43
constexpr double kRE = 6371000;                   // m, source/common.h:44
constexpr double kEps = DBL_EPSILON;              // gEPS, source/common.h:38

// source/utils.h:12-22
inline double e2p(double E) { return std::sqrt(E * (E + 2 * kE0)) / kC; }
inline double p2e(double p) { return std::sqrt(p * p * kC * kC + kE0 * kE0) - kE0; }
inline double dlogE_dp(double logE) {
  const double E = std::exp(logE);
  return e2p(E) * kC * kC / (E * (E + kE0));
}

// Diffusion table D/<dID>.h5 (source/Cases/Albert_Young_IO.cc:17-37): alpha0
// in degrees (converted to radians here), E in MeV, and Daa, Dap, Dpp as
// (n_alpha0, n_E) arrays, row-major.
struct DTable {
  std::vector<double> x, y, daa, dap, dpp;
  std::size_t nx() const { return x.size(); }
  std::size_t ny() const { return y.size(); }
};

// Bilinear position in the table (Albert_Young::locate, Albert_Young.cc:94-111,
// and calWeight, source/utils.h:41-51). The reference keeps the index in a
// size_t, so its "i < 0" branch never runs: a position below the table is
// converted to an unsigned value, which on x86-64 is out of range and is
// clamped to the last interval with weight 0 (the table's last row). The grids
// of the shipped parameter files lie inside the table and never reach it.
struct Loc { std::size_t i0, j0; double wi, wj; };

inline void cal_weight(double pos, std::size_t n, std::size_t &i, double &w) {
  const double fl = std::floor(pos);
  i = fl < 0 ? n : (std::size_t)fl;
  if (i < n) {
    w = 1.0 - (pos - (double)i);
  } else {
    i = n - 1;
    w = 0.0;
  }
}

inline Loc locate(const DTable &t, double alpha0, double logE) {
  const double xmin = t.x.front(), xmax = t.x.back();
  const double ymin = t.y.front(), ymax = t.y.back();
  const double pos_x = (alpha0 - xmin) / (xmax - xmin) * (double)(t.nx() - 1);
  const double pos_y = (logE - std::log(ymin)) / (std::log(ymax) - std::log(ymin)) * (double)(t.ny() - 1);
  Loc l;
  cal_weight(pos_x, t.nx() - 1, l.i0, l.wi);
  cal_weight(pos_y, t.ny() - 1, l.j0, l.wj);
  return l;
}

// interp2D, source/utils.h:25-38
inline double interp2d(const std::vector<double> &raw, std::size_t ny, const Loc &l) {
  auto r = [&](std::size_t i, std::size_t j) { return raw[i * ny + j]; };
  return r(l.i0, l.j0) * l.wi * l.wj + r(l.i0 + 1, l.j0) * (1 - l.wi) * l.wj +
         r(l.i0 + 1, l.j0 + 1) * (1 - l.wi) * (1 - l.wj) + r(l.i0, l.j0 + 1) * l.wi * (1 - l.wj);
}

// Albert_Young (lc = false) or Albert_Young_LC (lc = true) on mesh m, with
// logEmin = log(Emin) of the parameter file (source/Parameters.cc:51).
class AlbertYoung : public Equation {
 public:
  AlbertYoung(const Mesh &m, const DTable &t, double logEmin, bool lc)
      : Equation(m), m_(m), logEmin_(logEmin), lc_(lc) {
    if (t.nx() < 2 || t.ny() < 2) throw std::runtime_error("D table: fewer than 2 points per axis");
    for (std::size_t i = 0; i < m.nx(); ++i)       // init(), Albert_Young.cc:29-40
      for (std::size_t j = 0; j < m.ny(); ++j) G_(i, j) = calc_G(m.x(i), m.y(j));
    if (lc) {                                      // Albert_Young_LC.cc:42-53
      const double L = 4.5;
      const double alpha0lc = std::asin(std::pow(std::pow(L, 5) * (4 * L - 3), -0.25));
      for (std::size_t i = 0; i < m.nx(); ++i) {
        const double alpha0 = m.x(i);
        if (alpha0 < alpha0lc)
          for (std::size_t j = 0; j < m.ny(); ++j) {
            const double p = e2p(std::exp(m.y(j)));
            inv_tau_(i, j) = 4.0 / bounce_period(alpha0, p, L);
          }
      }
    }
    // constructD, Albert_Young.cc:113-134
    const double denormalize_factor = kME * kME * kC * kC;
    const double second_to_day = 3600 * 24;
    for (std::size_t i = 0; i < m.nx(); ++i) {
      const double alpha0 = m.x(i);
      for (std::size_t j = 0; j < m.ny(); ++j) {
        const double logE = m.y(j);
        const double p = e2p(std::exp(logE));
        const Loc l = locate(t, alpha0, logE);
        Dxx_(i, j) = interp2d(t.daa, t.ny(), l) * denormalize_factor * second_to_day / (p * p);
        Dxy_(i, j) = interp2d(t.dap, t.ny(), l) * denormalize_factor * second_to_day * dlogE_dp(logE) / p;
        Dyy_(i, j) = interp2d(t.dpp, t.ny(), l) * denormalize_factor * second_to_day *
                     std::pow(dlogE_dp(logE), 2);
      }
    }
  }

  // Albert_Young.cc:42-59; Albert_Young_LC.cc:56-73
  BCType bc_type(BoundaryID side) const override {
    if (side == BoundaryID::XMAX || (lc_ && side == BoundaryID::XMIN)) return BCType::ZeroFlux;
    return BCType::Dirichlet;
  }
  double init_f(const Ind &ind) const override { return calc_init_f(m_.x(ind.i), m_.y(ind.j)); }
  void update(double) override {}   // time-independent (Albert_Young.cc:62)
  bool dirichlet_vertex_value(BoundaryID side, std::size_t i, std::size_t, double,
                              double *out) const override {
    switch (side) {
      case BoundaryID::XMIN:
        if (lc_) return false;
        *out = 0.0;
        return true;
      case BoundaryID::YMIN:
        *out = calc_init_f(m_.x_edge(i), logEmin_);
        return true;
      case BoundaryID::YMAX:
        *out = 0.0;
        return true;
      default:
        return false;
    }
  }

 private:
  // Albert_Young.h:37-40 and Albert_Young_LC.h:36-39
  double calc_init_f(double a, double logE) const {
    const double p = e2p(std::exp(logE));
    if (lc_) return std::exp(-(std::exp(logE) - 0.2) / 0.1) * std::sin(a) / (p * p) + kEps;
    return std::exp(-(std::exp(logE) - 0.2) / 0.1) * (std::sin(a) - std::sin(5 * kPi / 180)) / (p * p) +
           kEps;
  }
  // Albert_Young.h:42-45
  static double calc_G(double alpha, double logE) {
    const double t = 1.30 - 0.56 * std::sin(alpha);
    return std::pow(e2p(std::exp(logE)), 2) * t * std::sin(alpha) * std::cos(alpha) / dlogE_dp(logE);
  }
  // Albert_Young_LC.cc:150-160
  static double bounce_period(double a0, double p, double L) {
    const double T0 = 1.3802, T1 = 0.7405;
    const double y = std::sin(a0);
    const double Ty = T0 - 0.5 * (T0 - T1) * (y + std::sqrt(y));
    return 4 * L * kRE * ((kE0 + p2e(p)) / (kC * kC)) / p * Ty / (3e8 * 3600 * 24);
  }

  const Mesh &m_;
  double logEmin_;
  bool lc_;
};

}  // namespace ay
