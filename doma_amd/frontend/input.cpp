// input.cpp -- <doma_inp> deck reader.  Deck semantics follow
// src/InputParameters.cpp:91-524 (solver, angular_quad, mesh with
// ix *= refinement + 1, materials with the scatter-size check and optional
// transpose, transients/source_step); messages match the reference's.
#include "input.h"

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace doma
{
// ------------------------------------------------------------------------- //
// XML                                                                        //
// ------------------------------------------------------------------------- //

const XmlNode *
XmlNode::child(const std::string & n) const
{
  for (const auto & c : children)
    if (c.name == n)
      return &c;
  return nullptr;
}

bool
XmlNode::has(const std::string & a) const
{
  for (const auto & kv : attrs)
    if (kv.first == a)
      return true;
  return false;
}

std::string
XmlNode::attr(const std::string & a, const std::string & dflt) const
{
  for (const auto & kv : attrs)
    if (kv.first == a)
      return kv.second;
  return dflt;
}

namespace
{
struct Cursor
{
  const std::string & s;
  size_t i = 0;
  explicit Cursor(const std::string & t) : s(t) {}
  bool eof() const { return i >= s.size(); }
  bool starts(const char * p) const { return s.compare(i, std::char_traits<char>::length(p), p) == 0; }
  void skip_ws()
  {
    while (!eof() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r'))
      ++i;
  }
  [[noreturn]] void error(const std::string & what) const
  {
    size_t line = 1 + std::count(s.begin(), s.begin() + std::min(i, s.size()), '\n');
    throw InputError("XML parse error (line " + std::to_string(line) + "): " + what);
  }
  void skip_until(const char * end)
  {
    size_t p = s.find(end, i);
    if (p == std::string::npos)
      error(std::string("unterminated construct, expected ") + end);
    i = p + std::char_traits<char>::length(end);
  }
  // Comments, processing instructions and doctype between elements.
  void skip_misc()
  {
    for (;;)
    {
      skip_ws();
      if (starts("<!--"))
        skip_until("-->");
      else if (starts("<?"))
        skip_until("?>");
      else if (starts("<!"))
        skip_until(">");
      else
        return;
    }
  }
  std::string name()
  {
    size_t b = i;
    while (!eof() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' ||
                      s[i] == ':' || s[i] == '.'))
      ++i;
    if (b == i)
      error("expected a name");
    return s.substr(b, i - b);
  }
};

std::string
unescape(const std::string & v)
{
  std::string o;
  for (size_t i = 0; i < v.size(); ++i)
  {
    if (v[i] != '&')
    {
      o += v[i];
      continue;
    }
    const size_t e = v.find(';', i);
    if (e == std::string::npos)
    {
      o += v[i];
      continue;
    }
    const std::string ent = v.substr(i + 1, e - i - 1);
    if (ent == "lt") o += '<';
    else if (ent == "gt") o += '>';
    else if (ent == "amp") o += '&';
    else if (ent == "quot") o += '"';
    else if (ent == "apos") o += '\'';
    else { o += v.substr(i, e - i + 1); }
    i = e;
  }
  return o;
}

XmlNode
parse_element(Cursor & c)
{
  if (c.eof() || c.s[c.i] != '<')
    c.error("expected '<'");
  ++c.i;
  XmlNode n;
  n.name = c.name();
  for (;;)
  {
    c.skip_ws();
    if (c.eof())
      c.error("unterminated start tag <" + n.name + ">");
    if (c.starts("/>"))
    {
      c.i += 2;
      return n;
    }
    if (c.s[c.i] == '>')
    {
      ++c.i;
      break;
    }
    std::string an = c.name();
    c.skip_ws();
    if (c.eof() || c.s[c.i] != '=')
      c.error("expected '=' after attribute " + an);
    ++c.i;
    c.skip_ws();
    if (c.eof() || (c.s[c.i] != '"' && c.s[c.i] != '\''))
      c.error("expected a quoted value for attribute " + an);
    const char q = c.s[c.i++];
    size_t e = c.s.find(q, c.i);
    if (e == std::string::npos)
      c.error("unterminated attribute value");
    n.attrs.emplace_back(an, unescape(c.s.substr(c.i, e - c.i)));
    c.i = e + 1;
  }
  // Content: child elements, text, comments, CDATA; then </name>.
  for (;;)
  {
    if (c.eof())
      c.error("missing </" + n.name + ">");
    if (c.starts("<!--"))
      c.skip_until("-->");
    else if (c.starts("<![CDATA["))
      c.skip_until("]]>");
    else if (c.starts("<?"))
      c.skip_until("?>");
    else if (c.starts("</"))
    {
      c.i += 2;
      const std::string en = c.name();
      if (en != n.name)
        c.error("mismatched </" + en + ">, expected </" + n.name + ">");
      c.skip_ws();
      if (c.eof() || c.s[c.i] != '>')
        c.error("malformed end tag");
      ++c.i;
      return n;
    }
    else if (c.s[c.i] == '<')
      n.children.push_back(parse_element(c));
    else
      ++c.i;  // character data is not part of the deck schema
  }
}
} // namespace

XmlNode
parseXml(const std::string & text)
{
  Cursor c(text);
  c.skip_misc();
  if (c.eof())
    throw InputError("XML parse error: no root element");
  XmlNode root = parse_element(c);
  c.skip_misc();
  if (!c.eof())
    c.error("content after the root element");
  return root;
}

// ------------------------------------------------------------------------- //
// Deck                                                                       //
// ------------------------------------------------------------------------- //

namespace
{
// Leading whitespace skipped, trailing characters ignored (strtod semantics of
// the reference's attribute and list parsing).
bool
to_double(const std::string & s, double & v)
{
  const char * b = s.c_str();
  char * e = nullptr;
  errno = 0;
  v = std::strtod(b, &e);
  return e != b && errno != ERANGE;
}

bool
to_uint(const std::string & s, unsigned int & v)
{
  size_t p = 0;
  while (p < s.size() && std::isspace((unsigned char)s[p]))
    ++p;
  if (p >= s.size() || !(std::isdigit((unsigned char)s[p]) || s[p] == '+'))
    return false;
  char * e = nullptr;
  errno = 0;
  const unsigned long long x = std::strtoull(s.c_str() + p, &e, 10);
  if (errno == ERANGE || x >= UINT_MAX)
    return false;
  v = static_cast<unsigned int>(x);
  return true;
}

// parseVecFromString (InputParameters.cpp:10-49): spaces and newlines removed,
// comma separated.
template <typename T>
std::vector<T>
parse_vec(const std::string & str, const char * what)
{
  std::string s;
  for (char ch : str)
    if (ch != ' ' && ch != '\n')
      s += ch;
  std::vector<T> out;
  size_t b = 0;
  for (;;)
  {
    const size_t e = s.find(',', b);
    const std::string item = s.substr(b, e == std::string::npos ? std::string::npos : e - b);
    if constexpr (std::is_same<T, double>::value)
    {
      double v;
      if (!to_double(item, v))
        throw InputError(std::string("invalid number '") + item + "' in " + what);
      out.push_back(v);
    }
    else
    {
      unsigned int v;
      if (!to_uint(item, v))
        throw InputError(std::string("invalid integer '") + item + "' in " + what);
      out.push_back(v);
    }
    if (e == std::string::npos)
      break;
    b = e + 1;
  }
  return out;
}

double
double_param(const XmlNode & n, const std::string & name)
{
  double v = -1.0;
  if (!n.has(name) || !to_double(n.attr(name), v) || !(v >= 0.0))
    throw InputError("You must provide a valid param in " + name + "!");
  return v;
}

unsigned int
uint_param(const XmlNode & n, const std::string & name)
{
  unsigned int v = 0;
  if (!n.has(name) || !to_uint(n.attr(name), v))
    throw InputError("You must provide a valid param in " + name + "!");
  return v;
}

std::string
string_param(const XmlNode & n, const std::string & name)
{
  const std::string v = n.attr(name);
  if (v.empty())
    throw InputError("You must provide a valid param in " + name + "!");
  return v;
}

void
check_size(const std::vector<double> & v, size_t want, bool may_be_empty, const std::string & msg)
{
  if (v.size() != want && !(may_be_empty && v.empty()))
    throw InputError(msg);
}
} // namespace

InputParameters
parseInputParametersString(const std::string & xml)
{
  const XmlNode root = parseXml(xml);
  if (root.name != "doma_inp")
    throw InputError("The root element must be <doma_inp>!");
  InputParameters params;

  const XmlNode * s = root.child("solver");
  if (!s)
    throw InputError("A solver block has not been provided!");
  params._dsa = s->attr("dsa") == "true";
  const std::string st = s->attr("solve_type");
  if (st == "steady")
    params._mode = RunMode::FixedSrc;
  else if (st == "eigen")
    params._mode = RunMode::Eigen;
  else if (st == "transient")
    params._mode = RunMode::Transient;
  else
    throw InputError("Unknown run_mode " + st);
  const std::string sc = s->attr("scheme");
  if (sc == "dd")
    params._eq_type = EquationType::DD;
  else if (sc == "tw_dd")
    params._eq_type = EquationType::TW_DD;
  else
    throw InputError("Unknown scheme " + sc);
  params._num_e_groups = uint_param(*s, "num_groups");
  params._src_it_tol = double_param(*s, "fs_tol");
  params._num_src_it = uint_param(*s, "fs_max_it");
  params._gs_tol = double_param(*s, "mg_tol");
  params._num_mg_it = uint_param(*s, "mg_max_it");
  if (params._mode == RunMode::Eigen)
    params._k_tol = double_param(*s, "k_tol");
  if (params._mode == RunMode::Transient)
  {
    params._t0 = double_param(*s, "t0");
    params._t1 = double_param(*s, "t1");
    params._num_steps = uint_param(*s, "t_steps");
    const std::string ic = s->attr("ics");
    if (ic == "zero")
      params._ic = TransientIC::Zero;
    else if (ic == "steady")
      params._ic = TransientIC::SteadyState;
    else
      throw InputError("Unknown initial condition " + ic);
  }

  const XmlNode * aq = root.child("angular_quad");
  if (!aq)
    throw InputError("An angular quadrature block has not been provided!");
  params._num_polar = uint_param(*aq, "np");
  params._num_azimuthal = uint_param(*aq, "na");

  const XmlNode * m = root.child("mesh");
  if (!m)
    throw InputError("A mesh block has not been provided!");
  params._num_dims = uint_param(*m, "dims");
  params._refinement = uint_param(*m, "refinement");
  auto axis = [&](const char * d, const char * iv, std::vector<double> & widths,
                  std::vector<unsigned int> & ivals, const char * axis_name) {
    const XmlNode * dn = m->child(d);
    if (!dn)
      throw InputError(std::string("Mesh ") + axis_name + " dimensions have not been provided!");
    widths = parse_vec<double>(string_param(*dn, "vals"), d);
    const XmlNode * in = m->child(iv);
    if (!in)
      throw InputError(std::string("Mesh ") + axis_name + " divisions have not been provided!");
    ivals = parse_vec<unsigned int>(string_param(*in, "vals"), iv);
    for (auto & i : ivals)
      i *= (params._refinement + 1);
  };
  axis("dx", "ix", params._dx, params._x_intervals, "x");
  if (params._num_dims > 1)
    axis("dy", "iy", params._dy, params._y_intervals, "y");
  if (params._num_dims > 2)
    axis("dz", "iz", params._dz, params._z_intervals, "z");
  const XmlNode * b = m->child("blocks");
  if (!b)
    throw InputError("Mesh blocks have not been provided!");
  params._blocks = parse_vec<unsigned int>(string_param(*b, "vals"), "blocks");

  const XmlNode * mats = root.child("materials");
  if (!mats)
    throw InputError("A materials block has not been provided!");
  const unsigned int G = params._num_e_groups;
  for (const auto & mat : mats->children)
  {
    const unsigned int blk = uint_param(mat, "block");
    if (params._block_mat_info.count(blk))
      throw InputError("A block can only contain one set of material properties, you have provided "
                       "multiple material properties for block " + std::to_string(blk) + "!");
    MaterialProps & p = params._block_mat_info[blk];
    p._num_d_groups = params._mode == RunMode::Transient ? uint_param(mat, "delayed_groups") : 0u;
    for (const auto & rxn : mat.children)
    {
      const std::string t = rxn.attr("type");
      auto vals = [&]() { return parse_vec<double>(string_param(rxn, "mgxs"), "mgxs"); };
      if (t == "total")
        p._g_total = vals();
      else if (t == "scatter")
      {
        p._g_g_scatter_mat = vals();
        if (p._g_g_scatter_mat.size() != (size_t)G * G)
          throw InputError("Not enough cross sections have been provided! The simulation requires " +
                           std::to_string(G * G) + " cross sections; 'scatter' only provides " +
                           std::to_string(p._g_g_scatter_mat.size()) + ".");
        const std::string tr = rxn.attr("transpose");
        if (tr == "true" || tr == "1" || tr == "yes" || tr == "True" || tr == "TRUE")
        {
          auto copy = p._g_g_scatter_mat;
          for (unsigned int r = 0; r < G; ++r)
            for (unsigned int cidx = 0; cidx < G; ++cidx)
              p._g_g_scatter_mat[r * G + cidx] = copy[cidx * G + r];
        }
      }
      else if (t == "source")
        p._g_src = vals();
      else if (t == "chi_p")
        p._g_chi_p = vals();
      else if (t == "nu_sigma_f")
        p._g_prod = vals();
      else if (t == "inv_vel")
        p._g_inv_v = vals();
      else if (t == "chi_d")
        p._n_g_chi_d = vals();
      else if (t == "beta")
        p._g_n_beta = vals();
      else if (t == "lambda_d")
        p._n_lambda = vals();
      else
        throw InputError("Unsupported reaction type " + t);
    }
    check_size(p._g_total, G, false,
               "Not enough cross sections have been provided! The simulation requires " +
                   std::to_string(G) + " cross sections; 'total' only provides " +
                   std::to_string(p._g_total.size()) + ".");
    check_size(p._g_g_scatter_mat, (size_t)G * G, false,
               "Not enough cross sections have been provided! The simulation requires " +
                   std::to_string(G * G) + " cross sections; 'scatter' only provides " +
                   std::to_string(p._g_g_scatter_mat.size()) + ".");
    check_size(p._g_src, G, true,
               "Not enough external sources have been provided! The simulation requires " +
                   std::to_string(G) + " sources; 'source' only provides " +
                   std::to_string(p._g_src.size()) + ".");
    check_size(p._g_chi_p, G, true, "Not enough prompt fission spectra values have been provided!");
    check_size(p._g_prod, G, true,
               "Not enough fission production cross sections have been provided!");
    if (params._mode == RunMode::Transient)
    {
      check_size(p._g_inv_v, G, false, "Not enough inverse velocities have been provided!");
      check_size(p._n_g_chi_d, (size_t)G * p._num_d_groups, false,
                 "Not enough delayed fission spectra values have been provided!");
      check_size(p._g_n_beta, (size_t)G * p._num_d_groups, false,
                 "Not enough delayed fractions have been provided!");
      check_size(p._n_lambda, p._num_d_groups, false,
                 "Not enough delayed decay constants have been provided!");
    }
    if (params._dsa)
      for (double t : p._g_total)
        if (t < 1e-6)
          throw InputError("Void region detected! You cannot use synthetic acceleration when cross "
                           "sections are close to zero.");
  }

  const XmlNode * tr = root.child("transients");
  if (tr && params._mode == RunMode::Transient)
  {
    for (const auto & node : tr->children)
    {
      if (node.name != "source_step")
        continue;
      const unsigned int blk = uint_param(node, "block");
      if (params._block_step_src.count(blk))
        throw InputError("A block can only contain one step source, you have provided multiple "
                         "step sources for block " + std::to_string(blk) + "!");
      SourceStep & step = params._block_step_src[blk];
      const bool ins = !node.attr("insert_at").empty(), rem = !node.attr("remove_at").empty();
      if (ins && rem)
      {
        step._type = StepType::Both;
        step._insert_time = double_param(node, "insert_at");
        step._remove_time = double_param(node, "remove_at");
        if (step._remove_time < step._insert_time)
          throw InputError("The step source for block " + std::to_string(blk) +
                           " is being removed before it is inserted!");
      }
      else if (ins)
      {
        step._type = StepType::Insert;
        step._insert_time = double_param(node, "insert_at");
      }
      else if (rem)
      {
        step._type = StepType::Remove;
        step._remove_time = double_param(node, "remove_at");
      }
      else
        throw InputError("Invalid step source for block " + std::to_string(blk) + "!");
      step._g_src = parse_vec<double>(string_param(node, "mgxs"), "mgxs");
      if (step._g_src.size() != G)
        throw InputError("Not enough external sources have been provided for the step source in "
                         "block " + std::to_string(blk) + "!");
    }
  }
  return params;
}

InputParameters
parseInputParameters(const std::string & file_path)
{
  std::ifstream f(file_path, std::ios::binary);
  if (!f)
    throw InputError("Failed to parse " + file_path + "\nfile not found");
  std::stringstream ss;
  ss << f.rdbuf();
  try
  {
    return parseInputParametersString(ss.str());
  }
  catch (const InputError & e)
  {
    const std::string w = e.what();
    if (w.rfind("XML parse error", 0) == 0)
      throw InputError("Failed to parse " + file_path + "\n" + w);
    throw;
  }
}
} // namespace doma
