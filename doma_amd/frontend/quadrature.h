// quadrature.h -- angular quadratures of the DOMA front end: Gauss-Legendre
// polar nodes (Newton), Chebyshev azimuths and the Gauss-Legendre-Chebyshev
// product set binned by octant.  Same public interface and construction order as
// the reference (include/AngularPolynomials.h:10-64, include/GLAngularQuadrature.h:
// 10-26, include/GCAngularQuadrature.h:12-49).
#pragma once

#include <array>
#include <vector>

#include "transport_base.h"

namespace doma
{
class LegendrePolynomial
{
public:
  explicit LegendrePolynomial(unsigned int degree);

  unsigned int degree() const { return _degree; }
  const double & root(unsigned int index) const { return _roots[index]; }
  const double & weight(unsigned int index) const { return _weights[index]; }
  const std::vector<double> & getRoots() const { return _roots; }
  const std::vector<double> & getWeights() const { return _weights; }

private:
  double evaluate(double x) const;
  double evaluateDerivative(double x) const;

  unsigned int _degree;
  std::vector<double> _roots;
  std::vector<double> _weights;
};

class ChebyshevPolynomial
{
public:
  explicit ChebyshevPolynomial(unsigned int degree);

  unsigned int degree() const { return _degree; }
  const double & root(unsigned int index) const { return _roots[index]; }
  const double & angularRoot(unsigned int index) const { return _angular_roots[index]; }
  const double & weight(unsigned int index) const { return _weights[index]; }

private:
  unsigned int _degree;
  std::vector<double> _roots;
  std::vector<double> _angular_roots;
  std::vector<double> _weights;
};

class GLAngularQuadrature
{
public:
  explicit GLAngularQuadrature(unsigned int n_l);

  unsigned int order() const { return _n_l; }
  double direction(unsigned int n) const { return _polar.root(n); }
  const double & weight(unsigned int n) const { return _polar.weight(n); }
  const LegendrePolynomial & getPolarLegendre() const { return _polar; }

private:
  unsigned int _n_l;
  LegendrePolynomial _polar;
};

class GCAngularQuadrature
{
public:
  // Throws std::invalid_argument when a direction has a zero component (odd
  // n_l or n_c), where the reference prints an error and exits.
  GCAngularQuadrature(unsigned int n_c, unsigned int n_l, unsigned int num_dims);

  unsigned int order(Octant oct) const;
  void direction(Octant oct, unsigned int n, double & mu, double & eta, double & xi) const;
  const double & weight(Octant oct, unsigned int n) const;
  unsigned int totalOrder() const { return 2u * _n_c * _n_l; }
  unsigned int legendreOrder() const { return _n_l; }
  unsigned int chebyshevOrder() const { return _n_c; }

private:
  unsigned int _n_c, _n_l, _n_d;
  LegendrePolynomial _polar;
  ChebyshevPolynomial _azimuthal;
  std::array<std::vector<double>, 8u> _mu, _eta, _xi, _w;
};
} // namespace doma
