// solver.h -- TransportSolver1D/2D/3D<T> of the DOMA front end (reference
// interface: include/TransportSolver3D.h:15-28, TransportSolver2D.h,
// TransportSolver1D.h:25).  The cell-solve policy T selects the scheme as in the
// reference (DiamondDifference{1,2,3}D, TWDiamondDifference{2,3}D); the solve
// itself runs on the GPU through the C ABI (include/doma_c.h).  Results land in
// the mesh (BrickMesh::flux), as the reference leaves them in its cells.
#pragma once

#include <string>

#include "doma_c.h"
#include "mesh.h"
#include "quadrature.h"

namespace doma
{
// Cell-solve policies (include/DiamondDifference3D.h, TWDiamondDifference3D.h ...).
struct DiamondDifference1D { static constexpr int scheme = DOMA_SCHEME_DD; };
struct DiamondDifference2D { static constexpr int scheme = DOMA_SCHEME_DD; };
struct DiamondDifference3D { static constexpr int scheme = DOMA_SCHEME_DD; };
struct TWDiamondDifference2D { static constexpr int scheme = DOMA_SCHEME_TW_DD; };
struct TWDiamondDifference3D { static constexpr int scheme = DOMA_SCHEME_TW_DD; };

// Options of this build that the reference does not have.
struct SolverOptions
{
  int precision = DOMA_FP64;  // DOMA_FP32_SWEEP for the FP32-sweep mode
  doma_tuning tuning{};       // kernel shape (zero = default/autotuned)
  bool autotune = false;      // run the project autotuner before solving
};

// Shared driver: builds the C-ABI problem from a mesh + quadrature, solves,
// writes the flux back, prints the reference's summary lines.
class SolverCore
{
public:
  const doma_stats & lastStats() const { return _stats; }
  const std::string & lastError() const { return _error; }
  SolverOptions & options() { return _opt; }

protected:
  SolverCore(BrickMesh & mesh, const InputParameters & params, bool verbose, int scheme);
  enum class Mode
  {
    FixedSource,  // solveFixedSource
    Eigen,        // solveEigenvalue (power iteration)
    Transient     // solveTransient (zero initial condition)
  };
  bool run(const std::string & output_file_base, int dims, int dirs, const std::vector<double> & mu,
           const std::vector<double> & eta, const std::vector<double> & xi,
           const std::vector<double> & w, Mode mode = Mode::FixedSource);
  bool transientSteps(doma_ctx * ctx, const std::string & output_file_base);
  bool fixedSourceOnly();
  bool refuse(const std::string & why);

  BrickMesh & _mesh;
  const InputParameters & _params;
  bool _verbose;
  int _scheme;
  bool _use_dsa = false;  // 1D diffusion synthetic acceleration
  SolverOptions _opt;
  doma_stats _stats{};
  std::string _error;
};

template <typename T>
class TransportSolver3D : public SolverCore
{
public:
  TransportSolver3D(BrickMesh3D & mesh, const InputParameters & params, bool verbose,
                    unsigned int num_threads);
  // Solve the fixed-source problem (TransportSolver3D.cpp:262-316).
  bool solveFixedSource(const std::string & output_file_base = "", const double & t = 0.0);
  // k-eigenvalue power iteration (TransportSolver3D.cpp:29-86).
  bool solveEigenvalue(const std::string & output_file_base = "");
  // Transient from a zero initial condition (TransportSolver3D.cpp:88-135); the
  // steady-state initial condition is refused (see solver.cpp).
  bool solveTransient(const std::string & output_file_base = "");
  const GCAngularQuadrature & quadrature() const { return _angular_quad; }

private:
  GCAngularQuadrature _angular_quad;
};

template <typename T>
class TransportSolver2D : public SolverCore
{
public:
  TransportSolver2D(BrickMesh2D & mesh, const InputParameters & params, bool verbose,
                    unsigned int num_threads);
  bool solveFixedSource(const std::string & output_file_base = "", const double & t = 0.0);
  bool solveEigenvalue(const std::string & output_file_base = "");
  bool solveTransient(const std::string & output_file_base = "");

private:
  GCAngularQuadrature _angular_quad;
};

template <typename T>
class TransportSolver1D : public SolverCore
{
public:
  TransportSolver1D(BrickMesh1D & mesh, const InputParameters & params, bool verbose, bool dsa,
                    unsigned int num_threads);
  bool solveFixedSource(const std::string & output_file_base = "", const double & t = 0.0);
  bool solveEigenvalue(const std::string & output_file_base = "");
  bool solveTransient(const std::string & output_file_base = "");

private:
  GLAngularQuadrature _angular_quad;
  bool _dsa;
};

typedef TransportSolver3D<TWDiamondDifference3D> TWDDTransportSolver3D;
typedef TransportSolver3D<DiamondDifference3D> DDTransportSolver3D;
typedef TransportSolver2D<TWDiamondDifference2D> TWDDTransportSolver2D;
typedef TransportSolver2D<DiamondDifference2D> DDTransportSolver2D;
typedef TransportSolver1D<DiamondDifference1D> DDTransportSolver1D;
} // namespace doma
