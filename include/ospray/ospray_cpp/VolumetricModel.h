// OSPRay C++ wrapper: VolumetricModel (ospray_cpp/VolumetricModel.h of OSPRay 3.x).
#pragma once

#include "Volume.h"

namespace ospray {
namespace cpp {

class VolumetricModel : public ManagedObject<OSPVolumetricModel, OSP_VOLUMETRIC_MODEL>
{
 public:
  VolumetricModel(const Volume &volume) : VolumetricModel(volume.handle()) {}
  VolumetricModel(OSPVolume volume)
      : ManagedObject<OSPVolumetricModel, OSP_VOLUMETRIC_MODEL>(ospNewVolumetricModel(volume))
  {
  }
  VolumetricModel(OSPVolumetricModel existing = nullptr)
      : ManagedObject<OSPVolumetricModel, OSP_VOLUMETRIC_MODEL>(existing)
  {
  }
};

static_assert(sizeof(VolumetricModel) == sizeof(OSPVolumetricModel), "cpp::VolumetricModel holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::VolumetricModel, OSP_VOLUMETRIC_MODEL);

} // namespace ospray
