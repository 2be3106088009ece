// mesh.cpp -- brick mesh construction and output (BrickMesh3D.cpp:8-271,
// BrickMesh2D.cpp, BrickMesh1D.cpp).  Only vacuum boundaries exist in the
// reference (CartesianCell3D.cpp:31-38; main.cpp:53, 108, 201), so the boundary
// arguments are accepted and must be vacuum.
#include "mesh.h"

#include <algorithm>
#include <fstream>
#include <iomanip>

namespace doma
{
namespace
{
// Interval expansion: each interval of width d split into n equal cells
// (BrickMesh3D.cpp:44-61).
void
expand(const std::vector<double> & d, const std::vector<unsigned int> & n, std::vector<double> & l)
{
  l.clear();
  for (size_t i = 0; i < n.size(); ++i)
  {
    const double w = d[i] / static_cast<double>(n[i]);
    for (unsigned int j = 0; j < n[i]; ++j)
      l.push_back(w);
  }
}

// Centroids as the reference accumulates them: start at half the first width,
// then add each cell's full width (BrickMesh3D.cpp:83-101).
void
centroids(const std::vector<double> & l, std::vector<double> & c)
{
  c.assign(l.size(), 0.0);
  double x = 0.5 * l[0];
  for (size_t i = 0; i < l.size(); ++i)
  {
    c[i] = x;
    x += l[i];
  }
}
} // namespace

BrickMesh::BrickMesh(const InputParameters & params, unsigned int dims) : _params(params), _dims(dims)
{
  const auto & nx = params._x_intervals;
  const auto & ny = params._y_intervals;
  const auto & nz = params._z_intervals;
  const size_t sx = nx.size(), sy = dims > 1 ? ny.size() : 1, sz = dims > 2 ? nz.size() : 1;
  if (sx * sy * sz != params._blocks.size())
    throw InputError("Error: The blocks vector is not equal to the number of subdivisions!");
  if (params._dx.size() != nx.size() || (dims > 1 && params._dy.size() != ny.size()) ||
      (dims > 2 && params._dz.size() != nz.size()))
    throw InputError("Error: every mesh interval needs a length and a division count!");
  for (auto v : nx)
    if (v == 0)
      throw InputError("Error: mesh intervals must hold at least one cell!");
  expand(params._dx, nx, _lx);
  if (dims > 1)
    expand(params._dy, ny, _ly);
  else
    _ly.assign(1, 1.0);
  if (dims > 2)
    expand(params._dz, nz, _lz);
  else
    _lz.assign(1, 1.0);
  if (_lx.empty() || _ly.empty() || _lz.empty())
    throw InputError("Error: the mesh has no cells!");
  _tot_num_x = static_cast<unsigned int>(_lx.size());
  _tot_num_y = static_cast<unsigned int>(_ly.size());
  _tot_num_z = static_cast<unsigned int>(_lz.size());
  _num_cells = _tot_num_x * _tot_num_y * _tot_num_z;
  centroids(_lx, _cx);
  centroids(_ly, _cy);
  centroids(_lz, _cz);

  // Block of every cell from its interval (BrickMesh3D.cpp:64-80).
  std::vector<unsigned int> ix, iy, iz;
  for (size_t i = 0; i < sx; ++i)
    ix.insert(ix.end(), nx[i], (unsigned int)i);
  if (dims > 1)
    for (size_t i = 0; i < sy; ++i)
      iy.insert(iy.end(), ny[i], (unsigned int)i);
  else
    iy.assign(1, 0);
  if (dims > 2)
    for (size_t i = 0; i < sz; ++i)
      iz.insert(iz.end(), nz[i], (unsigned int)i);
  else
    iz.assign(1, 0);
  _block.resize(_num_cells);
  for (unsigned int k = 0; k < _tot_num_z; ++k)
    for (unsigned int j = 0; j < _tot_num_y; ++j)
      for (unsigned int i = 0; i < _tot_num_x; ++i)
      {
        const size_t c = ((size_t)k * _tot_num_y + j) * _tot_num_x + i;
        _block[c] = params._blocks[(iz[k] * sy + iy[j]) * sx + ix[i]];
        _total_volume += dims == 3 ? _lx[i] * _ly[j] * _lz[k] : (dims == 2 ? _lx[i] * _ly[j] : _lx[i]);
      }
  _phi.assign(_num_cells, 0.0);
}

void
BrickMesh::setNumGroups(unsigned int g)
{
  _num_groups = g;
  _phi.assign((size_t)g * _num_cells, 0.0);
}

void
BrickMesh::validateProps() const
{
  for (unsigned int c = 0; c < _num_cells; ++c)
    if (_params._block_mat_info.count(_block[c]) == 0)
      throw InputError("Block " + std::to_string(_block[c]) +
                       " (defined on the mesh) has no material properties!");
}

// First cell (in storage order) whose closed box contains the point
// (CartesianCell3D.cpp:57-65, BrickMesh3D.cpp:201-213).
bool
BrickMesh::lookup(const double * pt, unsigned int g, double & returned_flux) const
{
  if (g >= _num_groups)
    return false;
  auto in = [](double c, double l, double x) { return (c - (0.5 * l)) <= x && (c + (0.5 * l)) >= x; };
  for (unsigned int k = 0; k < _tot_num_z; ++k)
  {
    if (_dims > 2 && !in(_cz[k], _lz[k], pt[2]))
      continue;
    for (unsigned int j = 0; j < _tot_num_y; ++j)
    {
      if (_dims > 1 && !in(_cy[j], _ly[j], pt[1]))
        continue;
      for (unsigned int i = 0; i < _tot_num_x; ++i)
        if (in(_cx[i], _lx[i], pt[0]))
        {
          returned_flux = _phi[(size_t)g * _num_cells + ((size_t)k * _tot_num_y + j) * _tot_num_x + i];
          return true;
        }
    }
  }
  return false;
}

void
BrickMesh::dumpToTextFile(const std::string & file_name, bool only_flux) const
{
  if (!only_flux)
  {
    std::ofstream dims(file_name + "_dims.txt");
    dims << "num_x: " << _tot_num_x << std::endl;
    if (_dims > 1)
      dims << "num_y: " << _tot_num_y << std::endl;
    if (_dims > 2)
      dims << "num_z: " << _tot_num_z << std::endl;
    dims << "num_g: " << _num_groups << std::endl;
    std::ofstream blocks(file_name + "_blocks.txt");
    std::ofstream x(file_name + "_meshx.txt");
    x << std::setprecision(6);
    std::ofstream y, z;
    if (_dims > 1)
    {
      y.open(file_name + "_meshy.txt");
      y << std::setprecision(6);
    }
    if (_dims > 2)
    {
      z.open(file_name + "_meshz.txt");
      z << std::setprecision(6);
    }
    for (unsigned int k = 0; k < _tot_num_z; ++k)
      for (unsigned int j = 0; j < _tot_num_y; ++j)
        for (unsigned int i = 0; i < _tot_num_x; ++i)
        {
          x << _cx[i] << std::endl;
          if (_dims > 1)
            y << _cy[j] << std::endl;
          if (_dims > 2)
            z << _cz[k] << std::endl;
          blocks << _block[((size_t)k * _tot_num_y + j) * _tot_num_x + i] << std::endl;
        }
  }
  for (unsigned int g = 0; g < _num_groups; ++g)
  {
    std::ofstream flux(file_name + "_g" + std::to_string(g) + "_flux.txt");
    flux << std::setprecision(6);
    for (unsigned int c = 0; c < _num_cells; ++c)
      flux << _phi[(size_t)g * _num_cells + c] << std::endl;
  }
}

unsigned int
BrickMesh::maxDelayedGroups() const
{
  unsigned int nd = 0;
  for (unsigned int c = 0; c < _num_cells; ++c)
    nd = std::max(nd, _params._block_mat_info.at(_block[c])._num_d_groups);
  return nd;
}

void
BrickMesh::dumpDNPsToTextFile(const std::string & file_name, const std::vector<double> & dnps,
                              unsigned int nd) const
{
  for (unsigned int d = 0; d < nd; ++d)
  {
    std::ofstream out(file_name + "_d" + std::to_string(d) + "_dnps.txt");
    out << std::setprecision(6);
    for (unsigned int c = 0; c < _num_cells; ++c)
    {
      if (_params._block_mat_info.at(_block[c])._num_d_groups == nd)
        out << dnps[(size_t)d * _num_cells + c] << std::endl;
      else
        out << 0.0 << std::endl;
    }
  }
}

namespace
{
template <size_t NB>
void
require_vacuum(const std::array<BoundaryCondition, NB> & bcs)
{
  for (auto b : bcs)
    if (b != BoundaryCondition::Vacuum)
      throw InputError("Only vacuum boundary conditions are supported.");
}
} // namespace

BrickMesh1D::BrickMesh1D(const InputParameters & params, const std::array<BoundaryCondition, 2u> & bcs,
                         unsigned int)
  : BrickMesh(params, 1)
{
  require_vacuum(bcs);
}

bool
BrickMesh1D::fluxAtPoint(const double & x, unsigned int g, double & returned_flux) const
{
  const double pt[3] = {x, 0.0, 0.0};
  return lookup(pt, g, returned_flux);
}

BrickMesh2D::BrickMesh2D(const InputParameters & params, const std::array<BoundaryCondition, 4u> & bcs,
                         unsigned int)
  : BrickMesh(params, 2)
{
  require_vacuum(bcs);
}

bool
BrickMesh2D::fluxAtPoint(const double & x, const double & y, unsigned int g,
                         double & returned_flux) const
{
  const double pt[3] = {x, y, 0.0};
  return lookup(pt, g, returned_flux);
}

BrickMesh3D::BrickMesh3D(const InputParameters & params, const std::array<BoundaryCondition, 6u> & bcs,
                         unsigned int)
  : BrickMesh(params, 3)
{
  require_vacuum(bcs);
}

bool
BrickMesh3D::fluxAtPoint(const double & x, const double & y, const double & z, unsigned int g,
                         double & returned_flux) const
{
  const double pt[3] = {x, y, z};
  return lookup(pt, g, returned_flux);
}
} // namespace doma
