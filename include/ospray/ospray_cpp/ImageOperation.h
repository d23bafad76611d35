// OSPRay C++ wrapper: ImageOperation (ospray_cpp/ImageOperation.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class ImageOperation : public ManagedObject<OSPImageOperation, OSP_IMAGE_OPERATION>
{
 public:
  ImageOperation(const std::string &type) : ManagedObject<OSPImageOperation, OSP_IMAGE_OPERATION>(ospNewImageOperation(type.c_str())) {}
  ImageOperation(OSPImageOperation existing = nullptr) : ManagedObject<OSPImageOperation, OSP_IMAGE_OPERATION>(existing) {}
};

static_assert(sizeof(ImageOperation) == sizeof(OSPImageOperation), "cpp::ImageOperation holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::ImageOperation, OSP_IMAGE_OPERATION);

} // namespace ospray
