// input.h -- input deck of the DOMA front end: MaterialProps / InputParameters
// with the reference's field names and meaning (include/InputParameters.h:9-119),
// and parseInputParameters() over an own XML reader for the <doma_inp> schema
// (src/InputParameters.cpp:91-524).  Where the reference prints a message and
// calls std::exit(1), this reader throws doma::InputError with the same message.
#pragma once

#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "transport_base.h"

namespace doma
{
struct InputError : std::runtime_error
{
  using std::runtime_error::runtime_error;
};

struct SourceStep
{
  StepType _type = StepType::Both;
  double _insert_time = 0.0;
  double _remove_time = 0.0;
  std::vector<double> _g_src;
};

struct MaterialProps
{
  std::vector<double> _g_total;
  std::vector<double> _g_prod;
  std::vector<double> _g_chi_p;
  std::vector<double> _g_inv_v;
  std::vector<double> _g_g_scatter_mat;  // [g_dest * G + g_src]
  std::vector<double> _g_src;            // empty: no external source
  unsigned int _num_d_groups = 0;
  std::vector<double> _n_g_chi_d;
  std::vector<double> _g_n_beta;
  std::vector<double> _n_lambda;
};

struct InputParameters
{
  RunMode _mode = RunMode::FixedSrc;
  unsigned int _num_e_groups = 1;
  double _t0 = 0.0, _t1 = 0.0;
  unsigned int _num_steps = 0;
  TransientIC _ic = TransientIC::Zero;
  double _src_it_tol = 0.0;
  double _gs_tol = 0.0;
  double _k_tol = 0.0;
  bool _dsa = false;
  unsigned int _num_src_it = 0;
  unsigned int _num_mg_it = 0;
  unsigned int _num_polar = 0;
  unsigned int _num_azimuthal = 0;
  EquationType _eq_type = EquationType::DD;
  unsigned int _num_dims = 1;
  unsigned int _refinement = 0;
  std::vector<unsigned int> _x_intervals, _y_intervals, _z_intervals;
  std::vector<double> _dx, _dy, _dz;
  std::vector<unsigned int> _blocks;
  std::unordered_map<unsigned int, MaterialProps> _block_mat_info;
  std::unordered_map<unsigned int, SourceStep> _block_step_src;
};

// Parse a <doma_inp> deck file.  Throws InputError on malformed XML or on any
// condition the reference rejects.
InputParameters parseInputParameters(const std::string & file_path);

// Same, from an in-memory document.
InputParameters parseInputParametersString(const std::string & xml);

// Minimal XML element tree (elements, attributes; text and comments ignored).
struct XmlNode
{
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<XmlNode> children;

  const XmlNode * child(const std::string & n) const;
  bool has(const std::string & a) const;
  std::string attr(const std::string & a, const std::string & dflt = "") const;
};

XmlNode parseXml(const std::string & text);
} // namespace doma
