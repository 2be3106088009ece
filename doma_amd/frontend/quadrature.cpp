// quadrature.cpp -- Gauss-Legendre / Chebyshev / product quadratures
// (AngularPolynomials.cpp:8-111, GLAngularQuadrature.cpp:3-20,
// GCAngularQuadrature.cpp:9-100): Newton iteration for the Legendre roots from
// cos(pi (i - 1/4) / (n + 1/2)) until |dr| <= 2e-16, roots descending; Chebyshev
// azimuths (2i-1) pi / (2n) with weights pi/n; product directions generated polar
// index major, antipodes interleaved in 3D, then binned by octant in insertion
// order.
#include "quadrature.h"

#include <cmath>
#include <stdexcept>

#include "doma_c.h"

namespace doma
{
LegendrePolynomial::LegendrePolynomial(unsigned int degree) : _degree(degree)
{
  _roots.assign(degree, 0.0);
  _weights.assign(degree, 0.0);
  for (unsigned int i = 1; i <= _degree; ++i)
  {
    double x = std::cos(M_PI * (static_cast<double>(i) - 0.25) / (static_cast<double>(_degree) + 0.5));
    double v = evaluate(x);
    double d = evaluateDerivative(x);
    double dr;
    do
    {
      dr = v / d;
      x -= dr;
      v = evaluate(x);
      d = evaluateDerivative(x);
    } while (std::abs(dr) > 2e-16);
    _roots[i - 1u] = x;
    _weights[i - 1u] = 2.0 / ((1.0 - x * x) * d * d);
  }
}

// Three-term recurrence (i) P_i = (2i-1) x P_{i-1} - (i-1) P_{i-2}.
double
LegendrePolynomial::evaluate(double x) const
{
  if (_degree == 0u)
    return 1.0;
  if (_degree == 1u)
    return x;
  double v = 1.0, v1 = x, v2 = 1.0;
  for (unsigned int i = 2; i <= _degree; ++i)
  {
    const double di = static_cast<double>(i);
    v = ((2.0 * di - 1.0) * x * v1 - (di - 1.0) * v2) / di;
    v2 = v1;
    v1 = v;
  }
  return v;
}

// P'_n(x) = n (x P_n - P_{n-1}) / (x^2 - 1).
double
LegendrePolynomial::evaluateDerivative(double x) const
{
  if (_degree == 0u)
    return 0.0;
  if (_degree == 1u)
    return 1.0;
  double v = 1.0, d = 0.0, v1 = x, v2 = 1.0;
  const double f = 1.0 / (x * x - 1.0);
  for (unsigned int i = 2; i <= _degree; ++i)
  {
    const double di = static_cast<double>(i);
    v = ((2.0 * di - 1.0) * x * v1 - (di - 1.0) * v2) / di;
    d = di * f * (x * v - v1);
    v2 = v1;
    v1 = v;
  }
  return d;
}

ChebyshevPolynomial::ChebyshevPolynomial(unsigned int degree) : _degree(degree)
{
  _roots.assign(_degree, 0.0);
  _angular_roots.assign(_degree, 0.0);
  _weights.assign(_degree, M_PI / static_cast<double>(_degree));
  for (unsigned int i = 1; i <= _degree; ++i)
  {
    _angular_roots[i - 1u] =
        (2 * static_cast<double>(i) - 1.0) / (2.0 * static_cast<double>(_degree)) * M_PI;
    _roots[i - 1u] = std::cos(_angular_roots[i - 1u]);
  }
}

GLAngularQuadrature::GLAngularQuadrature(unsigned int n_l) : _n_l(n_l), _polar(n_l) {}

namespace
{
int
octant_of(double mu, double eta, double xi)
{
  if (mu == 0.0 || eta == 0.0 || xi == 0.0)
    return -1;
  return (mu < 0.0 ? 4 : 0) | (eta < 0.0 ? 2 : 0) | (xi < 0.0 ? 1 : 0);
}
} // namespace

GCAngularQuadrature::GCAngularQuadrature(unsigned int n_c, unsigned int n_l, unsigned int num_dims)
  : _n_c(n_c), _n_l(n_l), _n_d(num_dims), _polar(n_l), _azimuthal(n_c)
{
  for (unsigned int i = 0u; i < _n_l; ++i)
  {
    const double mu = _polar.root(i);
    const double sin_theta = std::sqrt(1.0 - (mu * mu));
    for (unsigned int j = 0u; j < _n_c; ++j)
    {
      const double wgt = _polar.weight(i) * _azimuthal.weight(j);
      const double om = _azimuthal.angularRoot(j);
      const double eta = sin_theta * std::cos(om);
      const double xi = sin_theta * std::sin(om);
      // The reference pushes (mu, eta, xi) and, in 3D, its antipode right after;
      // binning keeps that order within every octant.
      for (int flip = 0; flip < (_n_d > 2 ? 2 : 1); ++flip)
      {
        const double s = flip ? -1.0 : 1.0;
        const int oct = octant_of(s * mu, s * eta, s * xi);
        if (oct < 0)
          throw std::invalid_argument(
              "Error: TransportSolver::classifyDirection(const double & mu, const double & eta, "
              "const double & xi).");
        _mu[oct].push_back(s * mu);
        _eta[oct].push_back(s * eta);
        _xi[oct].push_back(s * xi);
        _w[oct].push_back(wgt);
      }
    }
  }
}

unsigned int
GCAngularQuadrature::order(Octant oct) const
{
  return static_cast<unsigned int>(_w[static_cast<unsigned int>(oct)].size());
}

void
GCAngularQuadrature::direction(Octant oct, unsigned int n, double & mu, double & eta, double & xi) const
{
  const unsigned int o = static_cast<unsigned int>(oct);
  mu = _mu[o][n];
  eta = _eta[o][n];
  xi = _xi[o][n];
}

const double &
GCAngularQuadrature::weight(Octant oct, unsigned int n) const
{
  return _w[static_cast<unsigned int>(oct)][n];
}
} // namespace doma

// C-ABI helpers: the quadrature tables in the layout doma_problem_desc expects.
extern "C" int
doma_gc_quadrature(int32_t n_c, int32_t n_l, int32_t dims, double * mu, double * eta, double * xi,
                   double * w)
{
  if (n_c < 1 || n_l < 1 || (dims != 2 && dims != 3) || !mu || !eta || !xi || !w)
    return -1;
  try
  {
    doma::GCAngularQuadrature q(n_c, n_l, dims);
    const int per = (n_c * n_l) / 4;
    const int slots = dims == 3 ? 8 : 4;
    for (int s = 0; s < slots; ++s)
    {
      const auto oct = static_cast<doma::Octant>(dims == 3 ? s : 2 * s);
      if ((int)q.order(oct) != per)
        return -1;
      for (int n = 0; n < per; ++n)
      {
        q.direction(oct, n, mu[s * per + n], eta[s * per + n], xi[s * per + n]);
        w[s * per + n] = q.weight(oct, n);
      }
    }
    return per;
  }
  catch (...)
  {
    return -1;
  }
}

extern "C" int
doma_gl_quadrature(int32_t n, double * mu, double * w)
{
  if (n < 2 || (n & 1) || !mu || !w)
    return -1;
  doma::GLAngularQuadrature q(n);
  for (int i = 0; i < n; ++i)
  {
    mu[i] = q.direction(i);
    w[i] = q.weight(i);
  }
  return n / 2;
}
