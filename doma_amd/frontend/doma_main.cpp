// doma_main.cpp -- command-line driver with the reference's interface
// (src/main.cpp:16-298): doma -i deck.xml [--verbose] [-n threads], solving
// fixed-source, eigenvalue and transient decks on the GPU and writing
// <deck dir>/<deck stem>_*.txt.
// Extra options of this build: --fp32 (FP32-sweep mode), --autotune.
#include <chrono>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>

#include "input.h"
#include "mesh.h"
#include "solver.h"

using namespace doma;

namespace
{
void
usage(const char * argv0)
{
  std::cerr << "Usage: " << argv0
            << " -i DECK.xml [--verbose] [-n N_THREADS] [--fp32] [--autotune]\n"
               "Discrete Ordinates Mini-App (MI355X build): fixed-source, k-eigenvalue and "
               "transient S_N solves on the GPU.\n";
}

template <typename S>
bool
solve(S & s, RunMode mode, bool fp32, bool autotune, const std::string & base)
{
  s.options().precision = fp32 ? DOMA_FP32_SWEEP : DOMA_FP64;
  s.options().autotune = autotune;
  return mode == RunMode::Eigen       ? s.solveEigenvalue(base)
         : mode == RunMode::Transient ? s.solveTransient(base)
                                      : s.solveFixedSource(base);
}
} // namespace

int
main(int argc, char ** argv)
{
  std::string input;
  bool verbose = false, fp32 = false, autotune = false;
  int threads = 1;
  for (int i = 1; i < argc; ++i)
  {
    const std::string a = argv[i];
    if ((a == "-i" || a == "--input_file") && i + 1 < argc)
      input = argv[++i];
    else if (a == "--verbose")
      verbose = true;
    else if ((a == "-n" || a == "--n-threads") && i + 1 < argc)
      threads = std::atoi(argv[++i]);
    else if (a == "--fp32")
      fp32 = true;
    else if (a == "--autotune")
      autotune = true;
    else if (a == "-h" || a == "--help")
    {
      usage(argv[0]);
      return 0;
    }
    else
    {
      std::cerr << "Unknown argument: " << a << std::endl;
      usage(argv[0]);
      return 1;
    }
  }
  if (input.empty())
  {
    std::cerr << "-i/--input_file: required." << std::endl;
    usage(argv[0]);
    return 1;
  }
  (void)threads;  // the sweep runs on the GPU; the thread count only mirrors the reference's CLI

  const std::filesystem::path inp(input);
  const std::string base = (inp.parent_path() / inp.stem()).string();
  const auto start = std::chrono::steady_clock::now();
  try
  {
    const InputParameters params = parseInputParameters(inp.string());
    bool ok = false;
    if (params._num_dims == 1)
    {
      if (params._eq_type == EquationType::TW_DD)
      {
        std::cerr << "Theta-Weighted Diamond Differences are currently not supported in slab "
                     "geometries!"
                  << std::endl;
        return 1;
      }
      BrickMesh1D mesh(params, {BoundaryCondition::Vacuum, BoundaryCondition::Vacuum}, threads);
      mesh.validateProps();
      DDTransportSolver1D s(mesh, params, verbose, params._dsa, threads);
      ok = solve(s, params._mode, fp32, autotune, base);
    }
    else if (params._num_dims == 2)
    {
      std::array<BoundaryCondition, 4u> bcs;
      bcs.fill(BoundaryCondition::Vacuum);
      BrickMesh2D mesh(params, bcs, threads);
      mesh.validateProps();
      if (params._eq_type == EquationType::DD)
      {
        DDTransportSolver2D s(mesh, params, verbose, threads);
        ok = solve(s, params._mode, fp32, autotune, base);
      }
      else
      {
        TWDDTransportSolver2D s(mesh, params, verbose, threads);
        ok = solve(s, params._mode, fp32, autotune, base);
      }
    }
    else if (params._num_dims == 3)
    {
      std::array<BoundaryCondition, 6u> bcs;
      bcs.fill(BoundaryCondition::Vacuum);
      BrickMesh3D mesh(params, bcs, threads);
      mesh.validateProps();
      if (params._eq_type == EquationType::DD)
      {
        DDTransportSolver3D s(mesh, params, verbose, threads);
        ok = solve(s, params._mode, fp32, autotune, base);
      }
      else
      {
        TWDDTransportSolver3D s(mesh, params, verbose, threads);
        ok = solve(s, params._mode, fp32, autotune, base);
      }
    }
    else
    {
      std::cerr << "mesh dims must be 1, 2 or 3" << std::endl;
      return 1;
    }
    if (!ok)
      return 2;
    std::cout << (params._mode == RunMode::Eigen       ? "Eigenvalue solver finished executing."
                  : params._mode == RunMode::Transient ? "Transient solver finished executing."
                                                       : "Fixed source solver finished executing.")
              << std::endl;
    const double s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    std::cout << "Time elapsed: " << s << " (s)." << std::endl;
    return 0;
  }
  catch (const std::exception & e)
  {
    std::cerr << e.what() << std::endl;
    return 1;
  }
}
