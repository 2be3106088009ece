// ref_deck.cpp -- drives the reference's own compiled mesh and solver units on
// a parsed input deck and lets the reference write its text outputs
// (BrickMesh3D.cpp:216-271, BrickMesh2D.cpp:171-200, BrickMesh1D.cpp:129-160).
// TEST INFRASTRUCTURE: built by oracle/Makefile (target "ref") from the sources
// under /root/reference, run only in the build container to write fixtures.
//
// The reference's XML reader needs pugixml, which is absent, so the deck is
// handed over as a whitespace-separated spec holding the InputParameters
// fields after parsing (ix already multiplied by refinement + 1, scatter
// matrices dest-major after any transpose):
//   dims G scheme np na fs_max_it mg_max_it nthreads run
//   fs_tol mg_tol k_tol
//   t0 t1 steps                                  (run = 3 only)
//   nxi dx[nxi] ix[nxi]     (and y, z for dims > 1, > 2)
//   nblocks blocks[nblocks]
//   M  then per material: id has_src has_chi has_prod nd has_inv sigt[G] sigs[G*G]
//      (src[G] if has_src) (chi_p[G] if has_chi) (nu_sigma_f[G] if has_prod)
//      beta[G*nd] chi_d[G*nd] lambda_d[nd] (inv_vel[G] if has_inv)
//   S  then per step source: block type insert_at remove_at src[G]
//      (type 0 both, 1 insert, 2 remove)
// run = 1 solves the fixed-source problem, run = 2 the eigenvalue problem, run = 3
// the transient from a zero initial condition (2D/3D; the reference's 1D solver
// needs Eigen) and dumps the outputs at <base> (per step <base>_t<step>_*.txt for
// a transient); run = 0 only builds the mesh and dumps it with zero flux.
// Output: <base>.bin = int32 success, f64 phi[G][N] (the last step's for a
// transient), then for run = 3 int32 nd_max and f64 dnps[nd_max][N] (zero where a
// cell has fewer delayed groups); the solver's stdout goes to <base>.log, printed
// at 17 significant digits (for k).
#include "BrickMesh1D.h"
#include "TransportSolver2D.h"
#include "TransportSolver3D.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

struct RefDeckProbe {};

template <>
class TransportSolver3D<RefDeckProbe>
{
public:
  static const std::vector<CartesianCell3D> & cells(const BrickMesh3D & m) { return m._cells; }
  static const std::vector<double> & dnps(const CartesianCell3D & c) { return c._current_t_dnps; }
};

template <>
class TransportSolver2D<RefDeckProbe>
{
public:
  static const std::vector<CartesianCell2D> & cells(const BrickMesh2D & m) { return m._cells; }
  static const std::vector<double> & dnps(const CartesianCell2D & c) { return c._current_t_dnps; }
};

// The 1D mesh sizes its per-cell flux vectors only inside the (Eigen-dependent)
// 1D solver; the mesh befriends every TransportSolver1D<T> (BrickMesh1D.h:46-47,
// CartesianCell1D.h:91-92), so a probe sets the group count and zero fluxes the
// way the solver's set-up would before the mesh is dumped.
template <typename T>
class TransportSolver1D;

template <>
class TransportSolver1D<RefDeckProbe>
{
public:
  static void zeroFlux(BrickMesh1D & m, unsigned int G)
  {
    m._num_groups = G;
    for (auto & c : m._cells)
      c._total_scalar_flux.assign(G, 0.0);
  }
};

namespace
{
template <typename T>
void
readn(std::istream & in, std::vector<T> & v, size_t n)
{
  v.resize(n);
  for (auto & x : v)
    in >> x;
}

// Precursor concentrations [nd_max][N] as the reference's dumpDNPsToTextFile
// lays them out (zero where a cell has fewer delayed groups).
template <typename Probe, typename Cells>
void
collect_dnps(const Cells & cells, std::vector<double> & out, int & nd_max)
{
  nd_max = 0;
  for (const auto & c : cells)
    nd_max = std::max<int>(nd_max, (int)Probe::dnps(c).size());
  const size_t N = cells.size();
  out.assign((size_t)nd_max * N, 0.0);
  for (size_t i = 0; i < N; ++i)
  {
    const auto & d = Probe::dnps(cells[i]);
    if ((int)d.size() == nd_max)
      for (int k = 0; k < nd_max; ++k)
        out[(size_t)k * N + i] = d[k];
  }
}

template <typename Cells>
void
collect(const Cells & cells, int G, std::vector<double> & phi)
{
  const size_t N = cells.size();
  phi.assign(static_cast<size_t>(G) * N, 0.0);
  for (size_t c = 0; c < N; ++c)
    for (int g = 0; g < G; ++g)
      phi[g * N + c] = cells[c]._total_scalar_flux[g];
}
} // namespace

int
main(int argc, char ** argv)
{
  if (argc != 3)
  {
    std::fprintf(stderr, "usage: %s deck.spec out_base\n", argv[0]);
    return 2;
  }
  std::ifstream in(argv[1]);
  const std::string base = argv[2];
  int dims, G, scheme, np, na, fs_max_it, mg_max_it, T, run;
  double fs_tol, mg_tol, k_tol;
  in >> dims >> G >> scheme >> np >> na >> fs_max_it >> mg_max_it >> T >> run >> fs_tol >> mg_tol >>
      k_tol;
  if (!in)
    return 3;
  double t0 = 0.0, t1 = 1.0;
  unsigned int steps = 1;
  if (run == 3)
    in >> t0 >> t1 >> steps;

  InputParameters params;
  params._mode = run == 2 ? RunMode::Eigen : (run == 3 ? RunMode::Transient : RunMode::FixedSrc);
  params._num_e_groups = G;
  params._t0 = t0;
  params._t1 = t1;
  params._num_steps = steps;
  params._ic = TransientIC::Zero;
  params._src_it_tol = fs_tol;
  params._gs_tol = mg_tol;
  params._k_tol = k_tol;
  params._dsa = false;
  params._num_src_it = fs_max_it;
  params._num_mg_it = mg_max_it;
  params._num_polar = np;
  params._num_azimuthal = na;
  params._eq_type = scheme == 1 ? EquationType::TW_DD : EquationType::DD;
  params._num_dims = dims;
  params._refinement = 0;
  auto axis = [&](std::vector<double> & d, std::vector<unsigned int> & n) {
    size_t k;
    in >> k;
    readn(in, d, k);
    readn(in, n, k);
  };
  axis(params._dx, params._x_intervals);
  if (dims > 1)
    axis(params._dy, params._y_intervals);
  if (dims > 2)
    axis(params._dz, params._z_intervals);
  size_t nb;
  in >> nb;
  readn(in, params._blocks, nb);
  int M;
  in >> M;
  for (int m = 0; m < M; ++m)
  {
    unsigned int id, nd;
    int has, has_chi, has_prod, has_inv;
    MaterialProps p;
    in >> id >> has >> has_chi >> has_prod >> nd >> has_inv;
    readn(in, p._g_total, G);
    readn(in, p._g_g_scatter_mat, static_cast<size_t>(G) * G);
    if (has)
      readn(in, p._g_src, G);
    if (has_chi)
      readn(in, p._g_chi_p, G);
    if (has_prod)
      readn(in, p._g_prod, G);
    p._num_d_groups = nd;
    readn(in, p._g_n_beta, static_cast<size_t>(G) * nd);
    readn(in, p._n_g_chi_d, static_cast<size_t>(G) * nd);
    readn(in, p._n_lambda, nd);
    if (has_inv)
      readn(in, p._g_inv_v, G);
    params._block_mat_info.emplace(id, p);
  }
  int S;
  in >> S;
  for (int k = 0; k < S; ++k)
  {
    unsigned int blk;
    int type;
    SourceStep st;
    in >> blk >> type >> st._insert_time >> st._remove_time;
    st._type = type == 0 ? StepType::Both : (type == 1 ? StepType::Insert : StepType::Remove);
    readn(in, st._g_src, G);
    params._block_step_src.emplace(blk, st);
  }
  if (!in)
    return 4;

  omp_set_num_threads(T);
  std::ostringstream log;
  auto * old = std::cout.rdbuf(log.rdbuf());
  std::cout.precision(17);
  bool ok = false;
  std::vector<double> phi, dnps;
  int nd_max = 0;
  if (dims == 3)
  {
    std::array<BoundaryCondition, 6u> bcs;
    bcs.fill(BoundaryCondition::Vacuum);
    BrickMesh3D mesh(params, bcs, T);
    mesh.validateProps();
    if (!run)
      mesh.dumpToTextFile(base);
    else if (scheme == 1)
    {
      TransportSolver3D<TWDiamondDifference3D> s(mesh, params, true, T);
      ok = run == 3 ? s.solveTransient(base)
                    : (run == 2 ? s.solveEigenvalue(base) : s.solveFixedSource(base));
    }
    else
    {
      TransportSolver3D<DiamondDifference3D> s(mesh, params, true, T);
      ok = run == 3 ? s.solveTransient(base)
                    : (run == 2 ? s.solveEigenvalue(base) : s.solveFixedSource(base));
    }
    collect(TransportSolver3D<RefDeckProbe>::cells(mesh), G, phi);
    if (run == 3)
      collect_dnps<TransportSolver3D<RefDeckProbe>>(TransportSolver3D<RefDeckProbe>::cells(mesh),
                                                    dnps, nd_max);
  }
  else if (dims == 2)
  {
    std::array<BoundaryCondition, 4u> bcs;
    bcs.fill(BoundaryCondition::Vacuum);
    BrickMesh2D mesh(params, bcs, T);
    mesh.validateProps();
    if (!run)
      mesh.dumpToTextFile(base);
    else if (scheme == 1)
    {
      TransportSolver2D<TWDiamondDifference2D> s(mesh, params, true, T);
      ok = run == 3 ? s.solveTransient(base)
                    : (run == 2 ? s.solveEigenvalue(base) : s.solveFixedSource(base));
    }
    else
    {
      TransportSolver2D<DiamondDifference2D> s(mesh, params, true, T);
      ok = run == 3 ? s.solveTransient(base)
                    : (run == 2 ? s.solveEigenvalue(base) : s.solveFixedSource(base));
    }
    collect(TransportSolver2D<RefDeckProbe>::cells(mesh), G, phi);
    if (run == 3)
      collect_dnps<TransportSolver2D<RefDeckProbe>>(TransportSolver2D<RefDeckProbe>::cells(mesh),
                                                    dnps, nd_max);
  }
  else if (dims == 1 && !run)
  {
    BrickMesh1D mesh(params, {BoundaryCondition::Vacuum, BoundaryCondition::Vacuum}, T);
    mesh.validateProps();
    TransportSolver1D<RefDeckProbe>::zeroFlux(mesh, G);
    mesh.dumpToTextFile(base);
  }
  else
  {
    std::cout.rdbuf(old);
    std::fprintf(stderr, "a 1D solve is not buildable from the reference here\n");
    return 5;
  }
  std::cout.rdbuf(old);

  std::ofstream o(base + ".bin", std::ios::binary);
  int32_t s = ok ? 1 : 0;
  o.write(reinterpret_cast<const char *>(&s), sizeof(s));
  o.write(reinterpret_cast<const char *>(phi.data()),
          static_cast<std::streamsize>(phi.size() * sizeof(double)));
  if (run == 3)
  {
    const int32_t nd = nd_max;
    o.write(reinterpret_cast<const char *>(&nd), sizeof(nd));
    o.write(reinterpret_cast<const char *>(dnps.data()),
            static_cast<std::streamsize>(dnps.size() * sizeof(double)));
  }
  std::ofstream lg(base + ".log");
  lg << log.str();
  return 0;
}
