// OSPRay C++ wrapper: the OSPDataType of a C++ type (ospray_cpp/Traits.h of
// OSPRay 3.x).  OSPTypeFor<T>::value is OSP_UNKNOWN unless specialised;
// OSPTYPEFOR_SPECIALIZATION(T, OSP_...) registers an application type (a
// math library's vector, say) whose memory layout matches the OSPDataType.
#pragma once

#include "ospray/ospray.h"

static_assert(sizeof(bool) == 1, "OSP_BOOL parameters are one byte");

namespace ospray {

template <typename T>
struct OSPTypeFor
{
  static constexpr OSPDataType value = OSP_UNKNOWN;
};

#define OSPTYPEFOR_SPECIALIZATION(cpp_type, osp_type)                                                                 \
  template <>                                                                                                          \
  struct OSPTypeFor<cpp_type>                                                                                          \
  {                                                                                                                    \
    static constexpr OSPDataType value = osp_type;                                                                     \
  };

// Out-of-line definition of a specialisation's value, for code that still
// needs one (C++11/14 odr-use); since C++17 the in-class member is inline
// and this only checks that the type was registered.
#if __cplusplus >= 201703L
#define OSPTYPEFOR_DEFINITION(cpp_type)                                                                                \
  static_assert(OSPTypeFor<cpp_type>::value != OSP_UNKNOWN, "OSPTYPEFOR_DEFINITION of an unregistered type")
#else
#define OSPTYPEFOR_DEFINITION(cpp_type) constexpr OSPDataType OSPTypeFor<cpp_type>::value
#endif

// scalars and strings
OSPTYPEFOR_SPECIALIZATION(void *, OSP_VOID_PTR);
OSPTYPEFOR_SPECIALIZATION(char *, OSP_STRING);
OSPTYPEFOR_SPECIALIZATION(const char *, OSP_STRING);
OSPTYPEFOR_SPECIALIZATION(const char[], OSP_STRING);
OSPTYPEFOR_SPECIALIZATION(bool, OSP_BOOL);
OSPTYPEFOR_SPECIALIZATION(char, OSP_CHAR);
OSPTYPEFOR_SPECIALIZATION(unsigned char, OSP_UCHAR);
OSPTYPEFOR_SPECIALIZATION(short, OSP_SHORT);
OSPTYPEFOR_SPECIALIZATION(unsigned short, OSP_USHORT);
OSPTYPEFOR_SPECIALIZATION(int, OSP_INT);
OSPTYPEFOR_SPECIALIZATION(unsigned int, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(long, OSP_LONG);
OSPTYPEFOR_SPECIALIZATION(unsigned long, OSP_ULONG);
OSPTYPEFOR_SPECIALIZATION(long long, OSP_LONG);
OSPTYPEFOR_SPECIALIZATION(unsigned long long, OSP_ULONG);
OSPTYPEFOR_SPECIALIZATION(float, OSP_FLOAT);
OSPTYPEFOR_SPECIALIZATION(double, OSP_DOUBLE);

// object handles
OSPTYPEFOR_SPECIALIZATION(OSPObject, OSP_OBJECT);
OSPTYPEFOR_SPECIALIZATION(OSPCamera, OSP_CAMERA);
OSPTYPEFOR_SPECIALIZATION(OSPData, OSP_DATA);
OSPTYPEFOR_SPECIALIZATION(OSPFrameBuffer, OSP_FRAMEBUFFER);
OSPTYPEFOR_SPECIALIZATION(OSPFuture, OSP_FUTURE);
OSPTYPEFOR_SPECIALIZATION(OSPGeometricModel, OSP_GEOMETRIC_MODEL);
OSPTYPEFOR_SPECIALIZATION(OSPGeometry, OSP_GEOMETRY);
OSPTYPEFOR_SPECIALIZATION(OSPGroup, OSP_GROUP);
OSPTYPEFOR_SPECIALIZATION(OSPImageOperation, OSP_IMAGE_OPERATION);
OSPTYPEFOR_SPECIALIZATION(OSPInstance, OSP_INSTANCE);
OSPTYPEFOR_SPECIALIZATION(OSPLight, OSP_LIGHT);
OSPTYPEFOR_SPECIALIZATION(OSPMaterial, OSP_MATERIAL);
OSPTYPEFOR_SPECIALIZATION(OSPRenderer, OSP_RENDERER);
OSPTYPEFOR_SPECIALIZATION(OSPTexture, OSP_TEXTURE);
OSPTYPEFOR_SPECIALIZATION(OSPTransferFunction, OSP_TRANSFER_FUNCTION);
OSPTYPEFOR_SPECIALIZATION(OSPVolume, OSP_VOLUME);
OSPTYPEFOR_SPECIALIZATION(OSPVolumetricModel, OSP_VOLUMETRIC_MODEL);
OSPTYPEFOR_SPECIALIZATION(OSPWorld, OSP_WORLD);

// enums travel as 32-bit unsigned values (cell types as bytes)
OSPTYPEFOR_SPECIALIZATION(OSPLogLevel, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPDeviceProperty, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPDataType, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPTextureFormat, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPTextureFilter, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPTextureWrapMode, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPError, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPFrameBufferFormat, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPFrameBufferChannel, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPSyncEvent, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPUnstructuredCellType, OSP_UCHAR);
OSPTYPEFOR_SPECIALIZATION(OSPStereoMode, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPShutterType, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPCurveType, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPCurveBasis, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPSubdivisionMode, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPAMRMethod, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPVolumeFilter, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPVolumeFormat, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPIntensityQuantity, OSP_UINT);
OSPTYPEFOR_SPECIALIZATION(OSPPixelFilterType, OSP_UINT);

// the wrapper's own math-free array types
OSPTYPEFOR_SPECIALIZATION(OSPBounds, OSP_BOX3F);

} // namespace ospray
