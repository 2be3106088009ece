// transport_base.h -- enumerations of the DOMA front end, with the reference's
// names and numbering (include/TransportBase.h:3-62).
#pragma once

namespace doma
{
enum class CertesianAxis
{
  X = 0u,
  Y = 1u,
  Z = 2u
};

enum class CertesianFaceSide
{
  Right = 0u,  // face normal along +x
  Left = 1u,   // -x
  Front = 2u,  // +y
  Back = 3u,   // -y
  Top = 4u,    // +z
  Bottom = 5u  // -z
};

enum class Octant
{
  PPP = 0u,
  PPM = 1u,
  PMP = 2u,
  PMM = 3u,
  MPP = 4u,
  MPM = 5u,
  MMP = 6u,
  MMM = 7u
};

enum class BoundaryCondition
{
  Vacuum = 0u,
  Reflective = 1u
};

enum class RunMode
{
  FixedSrc = 0u,
  Eigen = 1u,
  Transient = 2u
};

enum class EquationType
{
  DD = 0u,
  TW_DD = 1u
};

enum class TransientIC
{
  Zero = 0u,
  SteadyState = 1u
};

enum class StepType
{
  Insert = 0u,
  Remove = 1u,
  Both = 2u
};
} // namespace doma
