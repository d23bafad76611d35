// OSPRay C++ wrapper: Camera (ospray_cpp/Camera.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class Camera : public ManagedObject<OSPCamera, OSP_CAMERA>
{
 public:
  Camera(const std::string &type) : ManagedObject<OSPCamera, OSP_CAMERA>(ospNewCamera(type.c_str())) {}
  Camera(OSPCamera existing = nullptr) : ManagedObject<OSPCamera, OSP_CAMERA>(existing) {}
};

static_assert(sizeof(Camera) == sizeof(OSPCamera), "cpp::Camera holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::Camera, OSP_CAMERA);

} // namespace ospray
