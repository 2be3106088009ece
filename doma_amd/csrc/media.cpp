// media.cpp -- seeded synthetic media for the block-structured fixed-source
// workloads (random sigma_t / scattering ratio per block and group).  Every value
// is a pure function of (seed, block, group, field) through SplitMix64, so the
// CPU oracle, the test fixtures and the GPU path see identical media on any host.
#include <cmath>
#include <cstring>

#include "doma_c.h"
#include "internal.h"

namespace
{
inline uint64_t
splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Uniform double in [0, 1) keyed by (seed, a, b, field).
inline double
u01(uint64_t seed, uint64_t a, uint64_t b, uint64_t field)
{
  uint64_t k = splitmix64(seed);
  k = splitmix64(k ^ (a * 0xD1B54A32D192ED03ull + 1));
  k = splitmix64(k ^ (b * 0xABC98388FB8FAC03ull + 2));
  k = splitmix64(k ^ (field * 0x8CB92BA72F3D8DD7ull + 3));
  return static_cast<double>(k >> 11) * 0x1.0p-53;
}

enum Field : uint64_t
{
  F_SIGT = 1,
  F_C = 2,
  F_HAS_SRC = 3,
  F_Q = 4
};

// Fraction of a source group's scattering that goes to each destination group.
void
scatter_split(int rule, int G, int g, double * frac /* [G] */)
{
  for (int d = 0; d < G; ++d)
    frac[d] = 0.0;
  if (G == 1)
  {
    frac[0] = 1.0;
    return;
  }
  if (rule == DOMA_SCATTER_DOWN1)
  {
    frac[g] = 0.7;
    if (g + 1 < G)
      frac[g + 1] = 0.3;
    return;
  }
  if (rule == DOMA_SCATTER_INGROUP)
  {
    // Scattering stays in the group except a 2 % share to the next group; the
    // last group keeps all of it.
    frac[g] = g + 1 < G ? 0.98 : 1.0;
    if (g + 1 < G)
      frac[g + 1] = 0.02;
    return;
  }
  // DOMA_SCATTER_FULL
  if (g == G - 1)
  {
    frac[g] = 0.8;      // in-group
    frac[g - 1] = 0.2;  // up-scatter into the group above
    return;
  }
  if (g == G - 2)
  {
    frac[g] = 0.7;
    frac[g + 1] = 0.3;
    return;
  }
  // Fast groups: 0.6 in-group, 0.4 spread over all lower groups with weights
  // 2^-(d-g), normalised.
  frac[g] = 0.6;
  double norm = 0.0;
  for (int d = g + 1; d < G; ++d)
    norm += std::ldexp(1.0, -(d - g));
  for (int d = g + 1; d < G; ++d)
    frac[d] = 0.4 * std::ldexp(1.0, -(d - g)) / norm;
}
} // namespace

extern "C" int32_t
doma_media_num_materials(const doma_media_spec * s)
{
  if (!s || s->n <= 0 || s->block <= 0 || s->n % s->block != 0)
    return -1;
  const int nb = s->n / s->block;
  const int blocks = nb * nb * nb;
  return s->source_rule == DOMA_SOURCE_CENTRAL ? 2 * blocks : blocks;
}

extern "C" int
doma_media_generate(const doma_media_spec * s, int32_t * mat, double * sigt, double * sigs,
                    double * src, uint8_t * has_src)
{
  const int32_t M = doma_media_num_materials(s);
  if (M < 0 || s->groups < 1 || s->groups > DOMA_MAX_GROUPS || !mat || !sigt || !sigs || !src ||
      !has_src)
    return doma::fail(DOMA_ERR_INVALID, "doma_media_generate: invalid spec or null output");
  if (s->sigt_lo <= 0.0 || s->sigt_hi < s->sigt_lo || s->c_lo < 0.0 || s->c_hi < s->c_lo)
    return doma::fail(DOMA_ERR_INVALID, "doma_media_generate: invalid sigma_t / c ranges");
  const int n = s->n, b = s->block, nb = n / b, G = s->groups;
  const int blocks = nb * nb * nb;
  double frac[DOMA_MAX_GROUPS];

  std::memset(has_src, 0, static_cast<size_t>(M));
  std::memset(src, 0, sizeof(double) * static_cast<size_t>(M) * G);
  std::memset(sigs, 0, sizeof(double) * static_cast<size_t>(M) * G * G);

  // Which blocks carry a source (DOMA_SOURCE_BLOCKS); at least one always does.
  int min_block = 0;
  double min_u = 2.0;
  int any = 0;
  for (int blk = 0; blk < blocks; ++blk)
  {
    double u = u01(s->seed, blk, 0, F_HAS_SRC);
    if (u < min_u)
    {
      min_u = u;
      min_block = blk;
    }
    if (s->source_rule == DOMA_SOURCE_BLOCKS && u < s->source_fraction)
    {
      has_src[blk] = 1;
      any = 1;
    }
  }
  if (s->source_rule == DOMA_SOURCE_BLOCKS && !any)
    has_src[min_block] = 1;

  for (int blk = 0; blk < blocks; ++blk)
  {
    for (int g = 0; g < G; ++g)
    {
      const double u = u01(s->seed, blk, g, F_SIGT);
      double st;
      if (s->sigt_log)
        st = std::exp(std::log(s->sigt_lo) + u * (std::log(s->sigt_hi) - std::log(s->sigt_lo)));
      else
        st = s->sigt_lo + u * (s->sigt_hi - s->sigt_lo);
      const double c = s->c_lo + u01(s->seed, blk, g, F_C) * (s->c_hi - s->c_lo);
      sigt[static_cast<size_t>(blk) * G + g] = st;
      scatter_split(s->scatter_rule, G, g, frac);
      for (int d = 0; d < G; ++d)
        sigs[(static_cast<size_t>(blk) * G + d) * G + g] = frac[d] * c * st;
      if (s->source_rule == DOMA_SOURCE_BLOCKS && has_src[blk])
        src[static_cast<size_t>(blk) * G + g] = u01(s->seed, blk, g, F_Q);
    }
  }
  if (s->source_rule == DOMA_SOURCE_CENTRAL)
  {
    // Materials blocks..2*blocks-1 copy their block's cross sections and add a
    // unit source in group 0.
    for (int blk = 0; blk < blocks; ++blk)
    {
      const int m = blocks + blk;
      for (int g = 0; g < G; ++g)
        sigt[static_cast<size_t>(m) * G + g] = sigt[static_cast<size_t>(blk) * G + g];
      for (int e = 0; e < G * G; ++e)
        sigs[static_cast<size_t>(m) * G * G + e] = sigs[static_cast<size_t>(blk) * G * G + e];
      has_src[m] = 1;
      src[static_cast<size_t>(m) * G + 0] = 1.0;
    }
  }
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i)
      {
        const int blk = ((k / b) * nb + (j / b)) * nb + (i / b);
        int m = blk;
        if (s->source_rule == DOMA_SOURCE_CENTRAL && i >= s->src_lo && i < s->src_hi &&
            j >= s->src_lo && j < s->src_hi && k >= s->src_lo && k < s->src_hi)
          m = blocks + blk;
        mat[(static_cast<size_t>(k) * n + j) * n + i] = m;
      }
  return DOMA_OK;
}
