// deck_capi.cpp -- C-ABI entry points over the C++ front end: read a
// <doma_inp> deck, and run a deck end to end the way the reference's main() does
// for fixed-source problems (src/main.cpp:51-298): mesh from the deck, solver by
// dimension and scheme, output files at <base>_*.txt.
#include <cstring>
#include <iostream>
#include <string>

#include "doma_c.h"
#include "internal.h"
#include "input.h"
#include "mesh.h"
#include "solver.h"

using namespace doma;

extern "C" int
doma_deck_read(const char * path, doma_deck_info * out)
{
  if (!path || !out)
    return fail(DOMA_ERR_INVALID, "doma_deck_read: null argument");
  try
  {
    const InputParameters p = parseInputParameters(path);
    std::memset(out, 0, sizeof(*out));
    out->dims = (int32_t)p._num_dims;
    out->groups = (int32_t)p._num_e_groups;
    out->scheme = p._eq_type == EquationType::TW_DD ? DOMA_SCHEME_TW_DD : DOMA_SCHEME_DD;
    out->mode = (int32_t)p._mode;
    out->np = (int32_t)p._num_polar;
    out->na = (int32_t)p._num_azimuthal;
    out->num_materials = (int32_t)p._block_mat_info.size();
    out->dsa = p._dsa ? 1 : 0;
    out->fs_tol = p._src_it_tol;
    out->mg_tol = p._gs_tol;
    out->fs_max_it = (int32_t)p._num_src_it;
    out->mg_max_it = (int32_t)p._num_mg_it;
    out->refinement = (int32_t)p._refinement;
    if (p._num_dims == 1)
    {
      BrickMesh1D m(p, {BoundaryCondition::Vacuum, BoundaryCondition::Vacuum}, 1);
      m.validateProps();
      out->nx = (int32_t)m.numX();
      out->ny = out->nz = 1;
      out->total_volume = m.totalVolume();
    }
    else if (p._num_dims == 2)
    {
      std::array<BoundaryCondition, 4u> b;
      b.fill(BoundaryCondition::Vacuum);
      BrickMesh2D m(p, b, 1);
      m.validateProps();
      out->nx = (int32_t)m.numX();
      out->ny = (int32_t)m.numY();
      out->nz = 1;
      out->total_volume = m.totalVolume();
    }
    else if (p._num_dims == 3)
    {
      std::array<BoundaryCondition, 6u> b;
      b.fill(BoundaryCondition::Vacuum);
      BrickMesh3D m(p, b, 1);
      m.validateProps();
      out->nx = (int32_t)m.numX();
      out->ny = (int32_t)m.numY();
      out->nz = (int32_t)m.numZ();
      out->total_volume = m.totalVolume();
    }
    else
      return fail(DOMA_ERR_INVALID, "mesh dims must be 1, 2 or 3");
    return DOMA_OK;
  }
  catch (const std::exception & e)
  {
    return fail(DOMA_ERR_INVALID, e.what());
  }
}

namespace
{
template <typename Solver, typename Mesh>
int
run_solver(Solver & s, Mesh & m, RunMode mode, const char * output_base, int32_t precision,
           const doma_tuning * tuning, int autotune, double * phi, int64_t phi_len, doma_stats * st,
           int32_t * converged)
{
  s.options().precision = precision;
  if (tuning)
    s.options().tuning = *tuning;
  s.options().autotune = autotune != 0;
  const std::string base = output_base ? output_base : "";
  const bool ok = mode == RunMode::Eigen       ? s.solveEigenvalue(base)
                  : mode == RunMode::Transient ? s.solveTransient(base)
                                               : s.solveFixedSource(base);
  if (!s.lastError().empty())
    return fail(DOMA_ERR_INVALID, s.lastError());
  if (converged)
    *converged = ok ? 1 : 0;
  if (st)
    *st = s.lastStats();
  if (phi)
  {
    if (phi_len < (int64_t)m.flux().size())
      return fail(DOMA_ERR_INVALID, "doma_run_deck: flux buffer too small");
    std::memcpy(phi, m.flux().data(), sizeof(double) * m.flux().size());
  }
  return DOMA_OK;
}
} // namespace

extern "C" int
doma_run_deck(const char * path, const char * output_base, int32_t precision,
              const doma_tuning * tuning, int32_t autotune, double * phi, int64_t phi_len,
              doma_stats * stats, int32_t * converged)
{
  if (!path)
    return fail(DOMA_ERR_INVALID, "doma_run_deck: null path");
  try
  {
    const InputParameters p = parseInputParameters(path);
    if (p._num_dims == 1)
    {
      if (p._eq_type == EquationType::TW_DD)
        return fail(DOMA_ERR_INVALID, "Theta-Weighted Diamond Differences are currently not "
                                      "supported in slab geometries!");
      BrickMesh1D m(p, {BoundaryCondition::Vacuum, BoundaryCondition::Vacuum}, 1);
      m.validateProps();
      DDTransportSolver1D s(m, p, false, p._dsa, 1);
      return run_solver(s, m, p._mode, output_base, precision, tuning, autotune, phi, phi_len, stats,
                        converged);
    }
    if (p._num_dims == 2)
    {
      std::array<BoundaryCondition, 4u> b;
      b.fill(BoundaryCondition::Vacuum);
      BrickMesh2D m(p, b, 1);
      m.validateProps();
      if (p._eq_type == EquationType::TW_DD)
      {
        TWDDTransportSolver2D s(m, p, false, 1);
        return run_solver(s, m, p._mode, output_base, precision, tuning, autotune, phi, phi_len, stats,
                          converged);
      }
      DDTransportSolver2D s(m, p, false, 1);
      return run_solver(s, m, p._mode, output_base, precision, tuning, autotune, phi, phi_len, stats,
                        converged);
    }
    if (p._num_dims == 3)
    {
      std::array<BoundaryCondition, 6u> b;
      b.fill(BoundaryCondition::Vacuum);
      BrickMesh3D m(p, b, 1);
      m.validateProps();
      if (p._eq_type == EquationType::TW_DD)
      {
        TWDDTransportSolver3D s(m, p, false, 1);
        return run_solver(s, m, p._mode, output_base, precision, tuning, autotune, phi, phi_len, stats,
                          converged);
      }
      DDTransportSolver3D s(m, p, false, 1);
      return run_solver(s, m, p._mode, output_base, precision, tuning, autotune, phi, phi_len, stats,
                        converged);
    }
    return fail(DOMA_ERR_INVALID, "mesh dims must be 1, 2 or 3");
  }
  catch (const std::exception & e)
  {
    return fail(DOMA_ERR_INVALID, e.what());
  }
}
