// OSPRay C++ wrapper: GeometricModel (ospray_cpp/GeometricModel.h of OSPRay 3.x).
#pragma once

#include "Geometry.h"

namespace ospray {
namespace cpp {

class GeometricModel : public ManagedObject<OSPGeometricModel, OSP_GEOMETRIC_MODEL>
{
 public:
  GeometricModel(const Geometry &geometry) : GeometricModel(geometry.handle()) {}
  GeometricModel(OSPGeometry geometry)
      : ManagedObject<OSPGeometricModel, OSP_GEOMETRIC_MODEL>(ospNewGeometricModel(geometry))
  {
  }
  GeometricModel(OSPGeometricModel existing = nullptr)
      : ManagedObject<OSPGeometricModel, OSP_GEOMETRIC_MODEL>(existing)
  {
  }
};

static_assert(sizeof(GeometricModel) == sizeof(OSPGeometricModel), "cpp::GeometricModel holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::GeometricModel, OSP_GEOMETRIC_MODEL);

} // namespace ospray
