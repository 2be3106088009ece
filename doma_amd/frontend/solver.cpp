// solver.cpp -- TransportSolver1D/2D/3D<T> over the C ABI.  Materials are the
// deck's blocks (MaterialProps per block id), mapped to dense material indices;
// quadratures are the reference's (GCAngularQuadrature(2na, 2np, 3|2),
// GLAngularQuadrature(2np): TransportSolver3D.cpp:10, TransportSolver2D.cpp:10,
// TransportSolver1D.cpp:12).  Console lines mirror the reference's summaries.
#include "solver.h"

#include <algorithm>
#include <iostream>
#include <map>

namespace doma
{
SolverCore::SolverCore(BrickMesh & mesh, const InputParameters & params, bool verbose, int scheme)
  : _mesh(mesh), _params(params), _verbose(verbose), _scheme(scheme)
{
  _mesh.setNumGroups(params._num_e_groups);
}

bool
SolverCore::refuse(const std::string & why)
{
  _error = why;
  std::cout << why << std::endl;
  return false;
}

bool
SolverCore::fixedSourceOnly()
{
  _error.clear();
  if (_params._mode == RunMode::FixedSrc)
    return true;
  return refuse("This build solves fixed-source problems only (solve_type=\"steady\").");
}

bool
SolverCore::run(const std::string & output_file_base, int dims, int dirs,
                const std::vector<double> & mu, const std::vector<double> & eta,
                const std::vector<double> & xi, const std::vector<double> & w, Mode mode)
{
  const bool eigen = mode == Mode::Eigen;
  _error.clear();
  const unsigned int G = _params._num_e_groups;
  if (G < 1 || G > DOMA_MAX_GROUPS)
  {
    _error = "num_groups must be in [1, 64]";
    std::cerr << _error << std::endl;
    return false;
  }
  // Dense material indices for the block ids in use, ascending by block id.
  std::map<unsigned int, int> index;
  for (const auto & kv : _params._block_mat_info)
    index.emplace(kv.first, 0);
  int m = 0;
  for (auto & kv : index)
    kv.second = m++;
  const int M = m;
  std::vector<double> sigt((size_t)M * G), sigs((size_t)M * G * G), src((size_t)M * G, 0.0);
  std::vector<uint8_t> has(M, 0);
  // Fission data: chi_p, nu_sigma_f and the prompt fraction 1 - sum_d beta of a
  // material with delayed groups (TransportSolver3D.cpp:377-387).
  std::vector<double> chi((size_t)M * G, 0.0), prod((size_t)M * G, 0.0), prompt((size_t)M * G, 1.0);
  std::vector<uint8_t> has_chi(M, 0), has_prod(M, 0);
  bool fission = false;
  for (const auto & kv : index)
  {
    const MaterialProps & p = _params._block_mat_info.at(kv.first);
    const int mi = kv.second;
    if (!p._g_chi_p.empty() && p._g_prod.empty())
    {
      _error = "Block " + std::to_string(kv.first) +
               " has a prompt fission spectrum (chi_p) but no nu_sigma_f!";
      std::cerr << _error << std::endl;
      return false;
    }
    if (!p._g_chi_p.empty() && p._num_d_groups > 0 &&
        p._g_n_beta.size() != (size_t)G * p._num_d_groups)
    {
      _error = "Not enough delayed fractions have been provided for block " +
               std::to_string(kv.first) + "!";
      std::cerr << _error << std::endl;
      return false;
    }
    if (!p._g_chi_p.empty())
    {
      has_chi[mi] = 1;
      for (unsigned int g = 0; g < G; ++g)
      {
        chi[(size_t)mi * G + g] = p._g_chi_p[g];
        double gb = 0.0;
        for (unsigned int d = 0; d < p._num_d_groups; ++d)
          gb += p._g_n_beta[g * p._num_d_groups + d];
        prompt[(size_t)mi * G + g] = 1.0 - gb;
      }
    }
    if (!p._g_prod.empty())
    {
      has_prod[mi] = 1;
      for (unsigned int g = 0; g < G; ++g)
        prod[(size_t)mi * G + g] = p._g_prod[g];
    }
    fission = fission || has_chi[mi] || has_prod[mi];
    for (unsigned int g = 0; g < G; ++g)
      sigt[(size_t)mi * G + g] = p._g_total[g];
    for (unsigned int e = 0; e < G * G; ++e)
      sigs[(size_t)mi * G * G + e] = p._g_g_scatter_mat[e];
    if (!p._g_src.empty())
    {
      has[mi] = 1;
      for (unsigned int g = 0; g < G; ++g)
        src[(size_t)mi * G + g] = p._g_src[g];
    }
  }
  const unsigned int N = _mesh.numCells();
  std::vector<int32_t> mat(N);
  for (unsigned int c = 0; c < N; ++c)
  {
    auto it = index.find(_mesh.blockIds()[c]);
    if (it == index.end())
    {
      _error = "Block " + std::to_string(_mesh.blockIds()[c]) +
               " (defined on the mesh) has no material properties!";
      std::cerr << _error << std::endl;
      return false;
    }
    mat[c] = it->second;
  }

  doma_problem_desc d{};
  d.dims = dims;
  d.nx = (int)_mesh.numX();
  d.ny = (int)_mesh.numY();
  d.nz = (int)_mesh.numZ();
  d.lx = _mesh.widthsX().data();
  d.ly = _mesh.widthsY().data();
  d.lz = _mesh.widthsZ().data();
  d.groups = (int)G;
  d.scheme = _scheme;
  d.precision = _opt.precision;
  d.dirs = dirs;
  d.mu = mu.data();
  d.eta = eta.data();
  d.xi = xi.data();
  d.w = w.data();

  std::cout << "Solving..." << std::endl;
  std::cout << "Initializing the solver..." << std::endl;
  doma_ctx * ctx = nullptr;
  int rc = doma_create(&d, &ctx);
  if (rc == DOMA_OK)
    rc = doma_set_materials(ctx, mat.data(), M, sigt.data(), sigs.data(), src.data(), has.data());
  if (rc == DOMA_OK && fission)
    rc = doma_set_fission(ctx, M, chi.data(), prod.data(), prompt.data(), has_chi.data(),
                          has_prod.data());
  if (rc == DOMA_OK && _use_dsa)
    rc = doma_set_dsa(ctx, 1);
  if (rc == DOMA_OK && mode == Mode::Transient)
  {
    // Kinetics data and step sources (MaterialProps, SourceStep) by material index.
    const size_t K = DOMA_MAX_DELAYED;
    std::vector<double> inv_v((size_t)M * G, 0.0), chi_d((size_t)M * G * K, 0.0),
        beta((size_t)M * G * K, 0.0), lambda((size_t)M * K, 0.0), step_in(M, 0.0),
        step_out(M, 0.0), step_src((size_t)M * G, 0.0);
    std::vector<int32_t> nd(M, 0), step_type(M, DOMA_STEP_NONE);
    for (const auto & kv : index)
    {
      const MaterialProps & p = _params._block_mat_info.at(kv.first);
      const int mi = kv.second;
      if (p._g_inv_v.size() != G || p._num_d_groups > DOMA_MAX_DELAYED)
      {
        _error = "Block " + std::to_string(kv.first) +
                 " needs inverse velocities and at most " + std::to_string(DOMA_MAX_DELAYED) +
                 " delayed groups for a transient";
        std::cerr << _error << std::endl;
        doma_destroy(ctx);
        return false;
      }
      nd[mi] = (int32_t)p._num_d_groups;
      for (unsigned int g = 0; g < G; ++g)
      {
        inv_v[(size_t)mi * G + g] = p._g_inv_v[g];
        for (unsigned int d = 0; d < p._num_d_groups; ++d)
        {
          chi_d[((size_t)mi * G + g) * K + d] = p._n_g_chi_d[g * p._num_d_groups + d];
          beta[((size_t)mi * G + g) * K + d] = p._g_n_beta[g * p._num_d_groups + d];
        }
      }
      for (unsigned int d = 0; d < p._num_d_groups; ++d)
        lambda[(size_t)mi * K + d] = p._n_lambda[d];
      auto st = _params._block_step_src.find(kv.first);
      if (st != _params._block_step_src.end())
      {
        step_type[mi] = st->second._type == StepType::Both
                            ? DOMA_STEP_BOTH
                            : (st->second._type == StepType::Insert ? DOMA_STEP_INSERT
                                                                    : DOMA_STEP_REMOVE);
        step_in[mi] = st->second._insert_time;
        step_out[mi] = st->second._remove_time;
        for (unsigned int g = 0; g < G && g < st->second._g_src.size(); ++g)
          step_src[(size_t)mi * G + g] = st->second._g_src[g];
      }
    }
    rc = doma_set_kinetics(ctx, M, inv_v.data(), nd.data(), chi_d.data(), beta.data(),
                           lambda.data());
    if (rc == DOMA_OK && !_params._block_step_src.empty())
      rc = doma_set_step_sources(ctx, M, step_type.data(), step_in.data(), step_out.data(),
                                 step_src.data());
  }
  if (rc == DOMA_OK && (_opt.tuning.brick || _opt.tuning.threads || _opt.tuning.exact_arith ||
                        _opt.tuning.schedule || _opt.tuning.octants || _opt.tuning.dirs_per_lane ||
                        _opt.tuning.shape))
    rc = doma_set_tuning(ctx, &_opt.tuning);
  if (rc == DOMA_OK && _opt.autotune)
    rc = doma_autotune(ctx, 0, nullptr, nullptr);
  if (rc == DOMA_OK && mode == Mode::Transient)
  {
    const bool ok = transientSteps(ctx, output_file_base);
    doma_destroy(ctx);
    return ok;
  }
  if (rc == DOMA_OK && eigen)
    rc = doma_solve_eigenvalue(ctx, _params._src_it_tol, (int)_params._num_src_it, _params._gs_tol,
                               (int)_params._num_mg_it, _params._k_tol, _mesh.flux().data(),
                               nullptr, &_stats);
  else if (rc == DOMA_OK)
    rc = doma_solve_fixed_source(ctx, _params._src_it_tol, (int)_params._num_src_it, _params._gs_tol,
                                 (int)_params._num_mg_it, _mesh.flux().data(), &_stats);
  if (rc != DOMA_OK)
  {
    _error = doma_last_error();
    std::cerr << "doma: " << _error << std::endl;
    doma_destroy(ctx);
    return false;
  }
  doma_destroy(ctx);

  for (unsigned int g = 0; g < G && g < DOMA_MAX_GROUPS; ++g)
  {
    if (_stats.si_iterations[g] == 0)
      break;
    const bool conv = _stats.si_iterations[g] < (int)_params._num_src_it;
    std::cout << (conv ? "SI converged after " : "SI failed to converge after ")
              << _stats.si_iterations[g] << " iterations with a residual of "
              << _stats.si_residual[g] << " for G" << g << std::endl;
  }
  if (eigen)
  {
    // solveEigenvalue's summary (TransportSolver3D.cpp:65-85).
    if (_stats.success)
    {
      std::cout << "PI converged after " << _stats.mg_iterations
                << " iterations with a flux shape residual of " << _stats.mg_residual
                << " and a k residual of " << _stats.k_residual << "." << std::endl
                << "Converged value of k = " << _stats.k_eff << std::endl;
      if (!output_file_base.empty())
        _mesh.dumpToTextFile(output_file_base);
    }
    else if (_stats.mg_iterations >= (int)_params._num_mg_it)
      std::cout << "PI failed to converge after " << _stats.mg_iterations
                << " iterations with a flux shape residual of " << _stats.mg_residual
                << " and a k residual of " << _stats.k_residual << "." << std::endl;
  }
  else if (_stats.success)
  {
    if (G > 1)
      std::cout << "MGI converged after " << _stats.mg_iterations << " iterations with a residual of "
                << _stats.mg_residual << std::endl;
    if (!output_file_base.empty())
      _mesh.dumpToTextFile(output_file_base);
  }
  else if (_stats.mg_iterations >= (int)_params._num_mg_it)
    std::cout << "MGI failed to converge after " << _stats.mg_iterations
              << " iterations with a residual of " << _stats.mg_residual << std::endl;
  if (_verbose)
    std::cout << "GPU: " << _stats.sweeps << " group sweeps, " << _stats.sweep_ms
              << " ms in sweep kernels, " << _stats.elapsed_ms << " ms total" << std::endl;
  return _stats.success != 0;
}

// solveTransient (TransportSolver3D.cpp:88-135) on a prepared context: zero
// initial condition, its dump, then per step the solve, the flux and precursor
// dumps at <base>_t<step>.
bool
SolverCore::transientSteps(doma_ctx * ctx, const std::string & output_file_base)
{
  const unsigned int G = _params._num_e_groups;
  std::cout << "Setting up initial conditions..." << std::endl;
  if (_params._ic != TransientIC::Zero)
  {
    // initSteadyIC runs solveFixedSource in transient mode before the previous-
    // step flux and precursor vectors are sized (TransportSolver2D.cpp:158-200
    // with :377-396), so its result is undefined; only ics="zero" is solved.
    return refuse("Transient mode supports ics=\"zero\" only: the reference's steady-state "
                  "initial condition reads unsized vectors.");
  }
  int rc = doma_transient_begin(ctx, _params._t0, _params._t1, (int)_params._num_steps);
  if (rc != DOMA_OK)
  {
    _error = doma_last_error();
    std::cerr << "doma: " << _error << std::endl;
    return false;
  }
  std::fill(_mesh.flux().begin(), _mesh.flux().end(), 0.0);
  if (!output_file_base.empty())
    _mesh.dumpToTextFile(output_file_base);
  const unsigned int nd = _mesh.maxDelayedGroups();
  int32_t nd_all = 0;
  doma_get_dnps(ctx, nullptr, &nd_all);
  std::vector<double> dnps((size_t)std::max<int32_t>(nd_all, 1) * _mesh.numCells());
  for (unsigned int step = 0; step < _params._num_steps; ++step)
  {
    std::cout << "Solving timestep " << step << "..." << std::endl;
    double t = 0.0;
    rc = doma_transient_step(ctx, _params._src_it_tol, (int)_params._num_src_it, _params._gs_tol,
                             (int)_params._num_mg_it, &t, &_stats);
    if (rc == DOMA_OK)
      rc = doma_get_flux(ctx, _mesh.flux().data());
    if (rc == DOMA_OK && nd_all > 0)
      rc = doma_get_dnps(ctx, dnps.data(), &nd_all);
    if (rc != DOMA_OK)
    {
      _error = doma_last_error();
      std::cerr << "doma: " << _error << std::endl;
      return false;
    }
    for (unsigned int g = 0; g < G && g < DOMA_MAX_GROUPS; ++g)
    {
      if (_stats.si_iterations[g] == 0)
        break;
      const bool conv = _stats.si_iterations[g] < (int)_params._num_src_it;
      std::cout << (conv ? "SI converged after " : "SI failed to converge after ")
                << _stats.si_iterations[g] << " iterations with a residual of "
                << _stats.si_residual[g] << " for G" << g << std::endl;
    }
    if (!_stats.success)
    {
      if (_stats.mg_iterations >= (int)_params._num_mg_it)
        std::cout << "MGI failed to converge after " << _stats.mg_iterations
                  << " iterations with a residual of " << _stats.mg_residual << std::endl;
      return false;
    }
    if (G > 1)
      std::cout << "MGI converged after " << _stats.mg_iterations << " iterations with a residual of "
                << _stats.mg_residual << std::endl;
    if (!output_file_base.empty())
    {
      const std::string b = output_file_base + "_t" + std::to_string(step);
      _mesh.dumpToTextFile(b, true);
      _mesh.dumpDNPsToTextFile(b, dnps, nd);
    }
  }
  return true;
}

namespace
{
// Flatten a product quadrature into [octant slot][n] arrays; 2D keeps the four
// slots with xi > 0 (PPP, PMP, MPP, MMP).
int
flatten(const GCAngularQuadrature & q, int dims, std::vector<double> & mu, std::vector<double> & eta,
        std::vector<double> & xi, std::vector<double> & w)
{
  const int slots = dims == 3 ? 8 : 4;
  const int per = (int)q.order(Octant::PPP);
  for (int s = 0; s < slots; ++s)
  {
    const auto oct = static_cast<Octant>(dims == 3 ? s : 2 * s);
    for (int n = 0; n < per; ++n)
    {
      double a, b, c;
      q.direction(oct, n, a, b, c);
      mu.push_back(a);
      eta.push_back(b);
      xi.push_back(c);
      w.push_back(q.weight(oct, n));
    }
  }
  return per;
}
} // namespace

// ---------------------------------------------------------------- 3D ------- //

template <typename T>
TransportSolver3D<T>::TransportSolver3D(BrickMesh3D & mesh, const InputParameters & params,
                                        bool verbose, unsigned int)
  : SolverCore(mesh, params, verbose, T::scheme),
    _angular_quad(2u * params._num_azimuthal, 2u * params._num_polar, 3u)
{
}

template <typename T>
bool
TransportSolver3D<T>::solveFixedSource(const std::string & output_file_base, const double &)
{
  if (!fixedSourceOnly())
    return false;
  std::vector<double> mu, eta, xi, w;
  const int per = flatten(_angular_quad, 3, mu, eta, xi, w);
  return run(output_file_base, 3, per, mu, eta, xi, w);
}

template <typename T>
bool
TransportSolver3D<T>::solveEigenvalue(const std::string & output_file_base)
{
  _error.clear();
  std::vector<double> mu, eta, xi, w;
  const int per = flatten(_angular_quad, 3, mu, eta, xi, w);
  return run(output_file_base, 3, per, mu, eta, xi, w, Mode::Eigen);
}

template <typename T>
bool
TransportSolver3D<T>::solveTransient(const std::string & output_file_base)
{
  _error.clear();
  std::vector<double> mu, eta, xi, w;
  const int per = flatten(_angular_quad, 3, mu, eta, xi, w);
  return run(output_file_base, 3, per, mu, eta, xi, w, Mode::Transient);
}

// ---------------------------------------------------------------- 2D ------- //

template <typename T>
TransportSolver2D<T>::TransportSolver2D(BrickMesh2D & mesh, const InputParameters & params,
                                        bool verbose, unsigned int)
  : SolverCore(mesh, params, verbose, T::scheme),
    _angular_quad(2u * params._num_azimuthal, 2u * params._num_polar, 2u)
{
}

template <typename T>
bool
TransportSolver2D<T>::solveFixedSource(const std::string & output_file_base, const double &)
{
  if (!fixedSourceOnly())
    return false;
  std::vector<double> mu, eta, xi, w;
  const int per = flatten(_angular_quad, 2, mu, eta, xi, w);
  return run(output_file_base, 2, per, mu, eta, xi, w);
}

template <typename T>
bool
TransportSolver2D<T>::solveEigenvalue(const std::string & output_file_base)
{
  _error.clear();
  std::vector<double> mu, eta, xi, w;
  const int per = flatten(_angular_quad, 2, mu, eta, xi, w);
  return run(output_file_base, 2, per, mu, eta, xi, w, Mode::Eigen);
}

template <typename T>
bool
TransportSolver2D<T>::solveTransient(const std::string & output_file_base)
{
  _error.clear();
  std::vector<double> mu, eta, xi, w;
  const int per = flatten(_angular_quad, 2, mu, eta, xi, w);
  return run(output_file_base, 2, per, mu, eta, xi, w, Mode::Transient);
}

// ---------------------------------------------------------------- 1D ------- //

template <typename T>
TransportSolver1D<T>::TransportSolver1D(BrickMesh1D & mesh, const InputParameters & params,
                                        bool verbose, bool dsa, unsigned int)
  : SolverCore(mesh, params, verbose, T::scheme), _angular_quad(2u * params._num_polar), _dsa(dsa)
{
  _use_dsa = dsa;
}

template <typename T>
bool
TransportSolver1D<T>::solveFixedSource(const std::string & output_file_base, const double &)
{
  if (!fixedSourceOnly())
    return false;
  const int n = (int)_angular_quad.order();
  std::vector<double> mu(n), eta(n, 0.0), xi(n, 0.0), w(n);
  for (int i = 0; i < n; ++i)
  {
    mu[i] = _angular_quad.direction(i);
    w[i] = _angular_quad.weight(i);
  }
  return run(output_file_base, 1, n / 2, mu, eta, xi, w);
}

template <typename T>
bool
TransportSolver1D<T>::solveEigenvalue(const std::string & output_file_base)
{
  _error.clear();
  const int n = (int)_angular_quad.order();
  std::vector<double> mu(n), eta(n, 0.0), xi(n, 0.0), w(n);
  for (int i = 0; i < n; ++i)
  {
    mu[i] = _angular_quad.direction(i);
    w[i] = _angular_quad.weight(i);
  }
  return run(output_file_base, 1, n / 2, mu, eta, xi, w, Mode::Eigen);
}

template <typename T>
bool
TransportSolver1D<T>::solveTransient(const std::string & output_file_base)
{
  _error.clear();
  const int n = (int)_angular_quad.order();
  std::vector<double> mu(n), eta(n, 0.0), xi(n, 0.0), w(n);
  for (int i = 0; i < n; ++i)
  {
    mu[i] = _angular_quad.direction(i);
    w[i] = _angular_quad.weight(i);
  }
  return run(output_file_base, 1, n / 2, mu, eta, xi, w, Mode::Transient);
}

template class TransportSolver3D<TWDiamondDifference3D>;
template class TransportSolver3D<DiamondDifference3D>;
template class TransportSolver2D<TWDiamondDifference2D>;
template class TransportSolver2D<DiamondDifference2D>;
template class TransportSolver1D<DiamondDifference1D>;
} // namespace doma
