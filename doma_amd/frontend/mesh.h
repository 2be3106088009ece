// mesh.h -- Cartesian brick meshes of the DOMA front end (BrickMesh1D/2D/3D,
// include/BrickMesh3D.h:14-99 and the 1D/2D counterparts).  Cells are stored as
// structure-of-arrays in the reference's order (k-major, i fastest); interval
// expansion, block assignment, centroids, point lookup and the text dumps
// follow src/BrickMesh3D.cpp:8-271, src/BrickMesh2D.cpp, src/BrickMesh1D.cpp.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "input.h"

namespace doma
{
class BrickMesh
{
public:
  unsigned int dims() const { return _dims; }
  unsigned int numX() const { return _tot_num_x; }
  unsigned int numY() const { return _tot_num_y; }
  unsigned int numZ() const { return _tot_num_z; }
  unsigned int numCells() const { return _num_cells; }
  unsigned int numGroups() const { return _num_groups; }
  void setNumGroups(unsigned int g);

  // Per-axis cell widths and centroids (the reference's running-sum centroids).
  const std::vector<double> & widthsX() const { return _lx; }
  const std::vector<double> & widthsY() const { return _ly; }
  const std::vector<double> & widthsZ() const { return _lz; }
  const std::vector<unsigned int> & blockIds() const { return _block; }
  const InputParameters & params() const { return _params; }

  // Total scalar flux, [group][cell].
  std::vector<double> & flux() { return _phi; }
  const std::vector<double> & flux() const { return _phi; }
  double totalVolume() const { return _total_volume; }

  // Every block on the mesh must have material properties (throws InputError).
  void validateProps() const;

  // Text dumps readable by the reference's plotting scripts.
  void dumpToTextFile(const std::string & file_name, bool only_flux = false) const;
  // Precursor dumps <file_name>_d<d>_dnps.txt (BrickMesh3D::dumpDNPsToTextFile):
  // dnps [nd][N]; a cell whose material has other than the mesh's largest number
  // of delayed groups prints 0.
  void dumpDNPsToTextFile(const std::string & file_name, const std::vector<double> & dnps,
                          unsigned int nd) const;
  // Largest delayed-group count over the materials of the mesh's cells.
  unsigned int maxDelayedGroups() const;

protected:
  BrickMesh(const InputParameters & params, unsigned int dims);
  bool lookup(const double * pt, unsigned int g, double & returned_flux) const;

  const InputParameters & _params;
  unsigned int _dims;
  unsigned int _tot_num_x = 1, _tot_num_y = 1, _tot_num_z = 1;
  unsigned int _num_cells = 0;
  unsigned int _num_groups = 1;
  std::vector<double> _lx, _ly, _lz;  // widths per column
  std::vector<double> _cx, _cy, _cz;  // centroids per column
  std::vector<unsigned int> _block;   // per cell
  std::vector<double> _phi;           // [G][N]
  double _total_volume = 0.0;
};

class BrickMesh1D : public BrickMesh
{
public:
  BrickMesh1D(const InputParameters & params, const std::array<BoundaryCondition, 2u> & bcs,
              unsigned int num_threads);
  bool fluxAtPoint(const double & x, unsigned int g, double & returned_flux) const;
};

class BrickMesh2D : public BrickMesh
{
public:
  BrickMesh2D(const InputParameters & params, const std::array<BoundaryCondition, 4u> & bcs,
              unsigned int num_threads);
  bool fluxAtPoint(const double & x, const double & y, unsigned int g, double & returned_flux) const;
};

class BrickMesh3D : public BrickMesh
{
public:
  BrickMesh3D(const InputParameters & params, const std::array<BoundaryCondition, 6u> & bcs,
              unsigned int num_threads);
  bool fluxAtPoint(const double & x, const double & y, const double & z, unsigned int g,
                   double & returned_flux) const;
};
} // namespace doma
