// ref_harness.cpp -- drives the reference's own compiled TransportSolver2D/3D
// units (built from /root/reference/src by oracle/Makefile, target "ref") to
// write golden vectors for the oracle and the HIP path, and to time the
// reference's own OpenMP solve as the benchmark's CPU baseline.  TEST
// INFRASTRUCTURE: never part of the product.
//
// The harness fills InputParameters programmatically (the reference's XML
// reader needs pugixml, which is absent): every cell becomes its own mesh
// interval (ix = 1, dx = the cell width) and its own block id = material id, so
// BrickMesh2D/3D (BrickMesh3D.cpp:8-102) expands to exactly the requested cell
// widths and material map.  The scalar flux is read back through an explicit
// specialisation of TransportSolver3D/2D, which BrickMesh3D/2D declare as
// friends for every template argument (BrickMesh3D.h:46-47).  The log ends with
// the wall time of solveFixedSource ("solve_seconds <t>"), which the benchmark's
// CPU baseline reads when it times the reference's own OpenMP solve.
//
// Problem file (little endian):
//   int32 dims, nx, ny, nz, G, scheme, M, np, na, fs_max_it, mg_max_it, nthreads
//   f64   fs_tol, mg_tol
//   f64   lx[nx], ly[ny], lz[nz]
//   int32 mat[nx*ny*nz]            (k-major, i fastest)
//   f64   sigt[M*G], sigs[M*G*G] (dest-major), src[M*G]
//   uint8 has_src[M]
// Output file: int32 success; f64 phi[G][N]; the solver's stdout goes to
// <out>.log.
#include "TransportSolver2D.h"
#include "TransportSolver3D.h"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

struct RefProbe {};

template <>
class TransportSolver3D<RefProbe>
{
public:
  static const std::vector<CartesianCell3D> & cells(const BrickMesh3D & m) { return m._cells; }
};

template <>
class TransportSolver2D<RefProbe>
{
public:
  static const std::vector<CartesianCell2D> & cells(const BrickMesh2D & m) { return m._cells; }
};

namespace
{
template <typename T>
bool
readv(std::ifstream & f, std::vector<T> & v, size_t n)
{
  v.resize(n);
  f.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
  return static_cast<bool>(f);
}
} // namespace

int
main(int argc, char ** argv)
{
  if (argc != 3)
  {
    std::fprintf(stderr, "usage: %s problem.bin out.bin\n", argv[0]);
    return 2;
  }
  std::ifstream f(argv[1], std::ios::binary);
  std::vector<int32_t> h;
  std::vector<double> tol;
  if (!readv(f, h, 12) || !readv(f, tol, 2))
    return 3;
  const int dims = h[0], nx = h[1], ny = h[2], nz = h[3], G = h[4], scheme = h[5], M = h[6];
  const int np = h[7], na = h[8], fs_max_it = h[9], mg_max_it = h[10], T = h[11];
  const size_t N = static_cast<size_t>(nx) * ny * nz;
  std::vector<double> lx, ly, lz, sigt, sigs, src;
  std::vector<int32_t> mat;
  std::vector<uint8_t> has_src;
  if (!readv(f, lx, nx) || !readv(f, ly, ny) || !readv(f, lz, nz) || !readv(f, mat, N) ||
      !readv(f, sigt, static_cast<size_t>(M) * G) ||
      !readv(f, sigs, static_cast<size_t>(M) * G * G) ||
      !readv(f, src, static_cast<size_t>(M) * G) || !readv(f, has_src, M))
    return 4;

  InputParameters params;
  params._mode = RunMode::FixedSrc;
  params._num_e_groups = G;
  params._t0 = 0.0;
  params._t1 = 1.0;
  params._num_steps = 1;
  params._ic = TransientIC::Zero;
  params._src_it_tol = tol[0];
  params._gs_tol = tol[1];
  params._k_tol = 0.0;
  params._dsa = false;
  params._num_src_it = fs_max_it;
  params._num_mg_it = mg_max_it;
  params._num_polar = np;
  params._num_azimuthal = na;
  params._eq_type = scheme == 1 ? EquationType::TW_DD : EquationType::DD;
  params._num_dims = dims;
  params._refinement = 0;
  params._x_intervals.assign(nx, 1u);
  params._dx = lx;
  if (dims > 1)
  {
    params._y_intervals.assign(ny, 1u);
    params._dy = ly;
  }
  if (dims > 2)
  {
    params._z_intervals.assign(nz, 1u);
    params._dz = lz;
  }
  params._blocks.assign(mat.begin(), mat.end());
  for (int m = 0; m < M; ++m)
  {
    MaterialProps p;
    p._g_total.assign(sigt.begin() + m * G, sigt.begin() + (m + 1) * G);
    p._g_g_scatter_mat.assign(sigs.begin() + static_cast<size_t>(m) * G * G,
                              sigs.begin() + static_cast<size_t>(m + 1) * G * G);
    if (has_src[m])
      p._g_src.assign(src.begin() + m * G, src.begin() + (m + 1) * G);
    p._num_d_groups = 0;
    params._block_mat_info.emplace(static_cast<unsigned int>(m), p);
  }

  omp_set_num_threads(T);
  std::ostringstream log;
  auto * old = std::cout.rdbuf(log.rdbuf());
  bool ok = false;
  std::vector<double> phi(static_cast<size_t>(G) * N, 0.0);
  // Wall time of solveFixedSource alone (mesh construction excluded), written to
  // the log as "solve_seconds <t>" for the CPU baseline.
  double solve_s = 0.0;
  auto t_solve = [&solve_s](auto && solve) {
    const auto t0 = std::chrono::steady_clock::now();
    const bool r = solve();
    solve_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
  };
  if (dims == 3)
  {
    std::array<BoundaryCondition, 6u> bcs;
    bcs.fill(BoundaryCondition::Vacuum);
    BrickMesh3D mesh(params, bcs, T);
    mesh.validateProps();
    if (scheme == 1)
    {
      TransportSolver3D<TWDiamondDifference3D> s(mesh, params, true, T);
      ok = t_solve([&] { return s.solveFixedSource(""); });
    }
    else
    {
      TransportSolver3D<DiamondDifference3D> s(mesh, params, true, T);
      ok = t_solve([&] { return s.solveFixedSource(""); });
    }
    const auto & cells = TransportSolver3D<RefProbe>::cells(mesh);
    for (size_t c = 0; c < N; ++c)
      for (int g = 0; g < G; ++g)
        phi[g * N + c] = cells[c]._total_scalar_flux[g];
  }
  else if (dims == 2)
  {
    std::array<BoundaryCondition, 4u> bcs;
    bcs.fill(BoundaryCondition::Vacuum);
    BrickMesh2D mesh(params, bcs, T);
    mesh.validateProps();
    if (scheme == 1)
    {
      TransportSolver2D<TWDiamondDifference2D> s(mesh, params, true, T);
      ok = t_solve([&] { return s.solveFixedSource(""); });
    }
    else
    {
      TransportSolver2D<DiamondDifference2D> s(mesh, params, true, T);
      ok = t_solve([&] { return s.solveFixedSource(""); });
    }
    const auto & cells = TransportSolver2D<RefProbe>::cells(mesh);
    for (size_t c = 0; c < N; ++c)
      for (int g = 0; g < G; ++g)
        phi[g * N + c] = cells[c]._total_scalar_flux[g];
  }
  else
  {
    std::cout.rdbuf(old);
    std::fprintf(stderr, "dims %d is not buildable from the reference here\n", dims);
    return 5;
  }
  std::cout.rdbuf(old);

  std::ofstream o(argv[2], std::ios::binary);
  int32_t s = ok ? 1 : 0;
  o.write(reinterpret_cast<const char *>(&s), sizeof(s));
  o.write(reinterpret_cast<const char *>(phi.data()),
          static_cast<std::streamsize>(phi.size() * sizeof(double)));
  std::ofstream lg(std::string(argv[2]) + ".log");
  lg << log.str() << "solve_seconds " << solve_s << "\n";
  return 0;
}
