// OSPRay C++ wrapper: one include for every class (ospray_cpp.h of OSPRay
// 3.x).  The classes are header-only owning handles over the C API of
// ospray/ospray.h; they work with any device, the "hip" device included.
#pragma once

#include "ospray_cpp/Camera.h"
#include "ospray_cpp/Data.h"
#include "ospray_cpp/Device.h"
#include "ospray_cpp/FrameBuffer.h"
#include "ospray_cpp/Future.h"
#include "ospray_cpp/GeometricModel.h"
#include "ospray_cpp/Geometry.h"
#include "ospray_cpp/Group.h"
#include "ospray_cpp/ImageOperation.h"
#include "ospray_cpp/Instance.h"
#include "ospray_cpp/Light.h"
#include "ospray_cpp/Material.h"
#include "ospray_cpp/Renderer.h"
#include "ospray_cpp/Texture.h"
#include "ospray_cpp/TransferFunction.h"
#include "ospray_cpp/Volume.h"
#include "ospray_cpp/VolumetricModel.h"
#include "ospray_cpp/World.h"
