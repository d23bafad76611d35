/* OSPRay-compatible C API, served by libospray.so of this project.
 *
 * The declarations keep the reference signatures one for one
 * (ospray/include/ospray/ospray.h:151-352) so that an application written
 * against OSPRay 3.3 (e.g. apps/ospTutorial/ospTutorial.c) compiles and links
 * unchanged.  Behind the API sits a device chosen at ospInit()/ospNewDevice()
 * time; the device of this project is "hip" (libospray_module_hip.so,
 * entry point ospray_module_init_hip), which renders the "pathtracer" on an
 * MI355X through hand-written CDNA4 kernels.
 */
#pragma once

#ifndef NULL
#if defined(__cplusplus) && __cplusplus >= 201103L
#define NULL nullptr
#else
#define NULL 0
#endif
#endif

#include <stdint.h>
#include <sys/types.h>

#include "OSPEnums.h"

#define OSPRAY_VERSION_MAJOR 3
#define OSPRAY_VERSION_MINOR 3
#define OSPRAY_VERSION_PATCH 0

#define OSPRAY_INTERFACE __attribute__((visibility("default")))

#ifdef __cplusplus
#define OSP_DEFAULT_VAL(a) = a
#else
#define OSP_DEFAULT_VAL(a)
#endif

#ifdef __cplusplus
namespace osp {
struct Device;
struct ManagedObject {};
struct Camera : public ManagedObject {};
struct Data : public ManagedObject {};
struct FrameBuffer : public ManagedObject {};
struct Future : public ManagedObject {};
struct GeometricModel : public ManagedObject {};
struct Geometry : public ManagedObject {};
struct Group : public ManagedObject {};
struct ImageOperation : public ManagedObject {};
struct Instance : public ManagedObject {};
struct Light : public ManagedObject {};
struct Material : public ManagedObject {};
struct Renderer : public ManagedObject {};
struct Texture : public ManagedObject {};
struct TransferFunction : public ManagedObject {};
struct Volume : public ManagedObject {};
struct VolumetricModel : public ManagedObject {};
struct World : public ManagedObject {};
} // namespace osp

typedef osp::Device *OSPDevice;
typedef osp::Camera *OSPCamera;
typedef osp::Data *OSPData;
typedef osp::FrameBuffer *OSPFrameBuffer;
typedef osp::Future *OSPFuture;
typedef osp::GeometricModel *OSPGeometricModel;
typedef osp::Geometry *OSPGeometry;
typedef osp::Group *OSPGroup;
typedef osp::ImageOperation *OSPImageOperation;
typedef osp::Instance *OSPInstance;
typedef osp::Light *OSPLight;
typedef osp::ManagedObject *OSPObject;
typedef osp::Material *OSPMaterial;
typedef osp::Renderer *OSPRenderer;
typedef osp::Texture *OSPTexture;
typedef osp::TransferFunction *OSPTransferFunction;
typedef osp::Volume *OSPVolume;
typedef osp::VolumetricModel *OSPVolumetricModel;
typedef osp::World *OSPWorld;
#else
typedef void *OSPDevice;
typedef void _OSPManagedObject;
typedef _OSPManagedObject *OSPManagedObject, *OSPCamera, *OSPData,
    *OSPFrameBuffer, *OSPFuture, *OSPGeometricModel, *OSPGeometry, *OSPGroup,
    *OSPImageOperation, *OSPInstance, *OSPLight, *OSPMaterial, *OSPObject,
    *OSPRenderer, *OSPTexture, *OSPTransferFunction, *OSPVolume,
    *OSPVolumetricModel, *OSPWorld;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- initialisation and devices (ospray.h:151-214) ---------------------- */

OSPRAY_INTERFACE OSPError ospInit(
    int *argc OSP_DEFAULT_VAL(NULL), const char **argv OSP_DEFAULT_VAL(NULL));
OSPRAY_INTERFACE int64_t ospDeviceGetProperty(OSPDevice, OSPDeviceProperty);
OSPRAY_INTERFACE void ospShutdown();
OSPRAY_INTERFACE OSPDevice ospNewDevice(
    const char *deviceType OSP_DEFAULT_VAL("cpu"));
OSPRAY_INTERFACE void ospSetCurrentDevice(OSPDevice);
OSPRAY_INTERFACE OSPDevice ospGetCurrentDevice();
OSPRAY_INTERFACE void ospDeviceSetParam(
    OSPDevice, const char *id, OSPDataType, const void *mem);
OSPRAY_INTERFACE void ospDeviceRemoveParam(OSPDevice, const char *id);

typedef void (*OSPStatusCallback)(void *userData, const char *messageText);
OSPRAY_INTERFACE void ospDeviceSetStatusCallback(
    OSPDevice, OSPStatusCallback, void *userData);

typedef void (*OSPErrorCallback)(
    void *userData, OSPError, const char *errorDetails);
OSPRAY_INTERFACE void ospDeviceSetErrorCallback(
    OSPDevice, OSPErrorCallback, void *userData);

OSPRAY_INTERFACE OSPError ospDeviceGetLastErrorCode(OSPDevice);
OSPRAY_INTERFACE const char *ospDeviceGetLastErrorMsg(OSPDevice);
OSPRAY_INTERFACE void ospDeviceCommit(OSPDevice);
OSPRAY_INTERFACE void ospDeviceRelease(OSPDevice);
OSPRAY_INTERFACE void ospDeviceRetain(OSPDevice);

/* Loads libospray_module_<name>.so and calls ospray_module_init_<name>. */
OSPRAY_INTERFACE OSPError ospLoadModule(const char *name);

/* ---- data arrays (ospray.h:219-262) ------------------------------------- */

typedef void (*OSPDeleterCallback)(const void *userData, const void *sharedData);

OSPRAY_INTERFACE OSPData ospNewSharedData(const void *sharedData,
    OSPDataType,
    uint64_t numItems1,
    int64_t byteStride1 OSP_DEFAULT_VAL(0),
    uint64_t numItems2 OSP_DEFAULT_VAL(1),
    int64_t byteStride2 OSP_DEFAULT_VAL(0),
    uint64_t numItems3 OSP_DEFAULT_VAL(1),
    int64_t byteStride3 OSP_DEFAULT_VAL(0),
    OSPDeleterCallback OSP_DEFAULT_VAL(NULL),
    const void *userPtr OSP_DEFAULT_VAL(NULL));

OSPRAY_INTERFACE OSPData ospNewData(OSPDataType,
    uint64_t numItems1,
    uint64_t numItems2 OSP_DEFAULT_VAL(1),
    uint64_t numItems3 OSP_DEFAULT_VAL(1));

OSPRAY_INTERFACE void ospCopyData(const OSPData source,
    OSPData destination,
    uint64_t destinationIndex1 OSP_DEFAULT_VAL(0),
    uint64_t destinationIndex2 OSP_DEFAULT_VAL(0),
    uint64_t destinationIndex3 OSP_DEFAULT_VAL(0));

/* ---- scene objects (ospray.h:265-300) ----------------------------------- */

OSPRAY_INTERFACE OSPLight ospNewLight(const char *type);
OSPRAY_INTERFACE OSPCamera ospNewCamera(const char *type);
OSPRAY_INTERFACE OSPGeometry ospNewGeometry(const char *type);
OSPRAY_INTERFACE OSPVolume ospNewVolume(const char *type);
OSPRAY_INTERFACE OSPGeometricModel ospNewGeometricModel(
    OSPGeometry OSP_DEFAULT_VAL(NULL));
OSPRAY_INTERFACE OSPVolumetricModel ospNewVolumetricModel(
    OSPVolume OSP_DEFAULT_VAL(NULL));
OSPRAY_INTERFACE OSPMaterial ospNewMaterial(const char *materialType);
OSPRAY_INTERFACE OSPTransferFunction ospNewTransferFunction(const char *type);
OSPRAY_INTERFACE OSPTexture ospNewTexture(const char *type);
OSPRAY_INTERFACE OSPGroup ospNewGroup();
OSPRAY_INTERFACE OSPInstance ospNewInstance(OSPGroup OSP_DEFAULT_VAL(NULL));
OSPRAY_INTERFACE OSPWorld ospNewWorld();

typedef struct
{
  float lower[3];
  float upper[3];
} OSPBounds;

OSPRAY_INTERFACE OSPBounds ospGetBounds(OSPObject);

/* ---- parameters and lifetime (ospray.h:303-316) ------------------------- */

OSPRAY_INTERFACE void ospSetParam(
    OSPObject, const char *id, OSPDataType, const void *mem);
OSPRAY_INTERFACE void ospRemoveParam(OSPObject, const char *id);
OSPRAY_INTERFACE void ospCommit(OSPObject);
OSPRAY_INTERFACE void ospRelease(OSPObject);
OSPRAY_INTERFACE void ospRetain(OSPObject);

/* ---- framebuffer (ospray.h:319-336) ------------------------------------- */

OSPRAY_INTERFACE OSPFrameBuffer ospNewFrameBuffer(int size_x,
    int size_y,
    OSPFrameBufferFormat format OSP_DEFAULT_VAL(OSP_FB_SRGBA),
    uint32_t frameBufferChannels OSP_DEFAULT_VAL(OSP_FB_COLOR));
OSPRAY_INTERFACE OSPImageOperation ospNewImageOperation(const char *type);
OSPRAY_INTERFACE const void *ospMapFrameBuffer(
    OSPFrameBuffer, OSPFrameBufferChannel OSP_DEFAULT_VAL(OSP_FB_COLOR));
OSPRAY_INTERFACE void ospUnmapFrameBuffer(const void *mapped, OSPFrameBuffer);
OSPRAY_INTERFACE float ospGetVariance(OSPFrameBuffer);
OSPRAY_INTERFACE void ospResetAccumulation(OSPFrameBuffer);

/* ---- rendering (ospray.h:339-352) --------------------------------------- */

OSPRAY_INTERFACE OSPRenderer ospNewRenderer(const char *type);
OSPRAY_INTERFACE OSPFuture ospRenderFrame(
    OSPFrameBuffer, OSPRenderer, OSPCamera, OSPWorld);
OSPRAY_INTERFACE int ospIsReady(
    OSPFuture, OSPSyncEvent OSP_DEFAULT_VAL(OSP_TASK_FINISHED));
OSPRAY_INTERFACE void ospWait(
    OSPFuture, OSPSyncEvent OSP_DEFAULT_VAL(OSP_TASK_FINISHED));
OSPRAY_INTERFACE void ospCancel(OSPFuture);
OSPRAY_INTERFACE float ospGetProgress(OSPFuture);
OSPRAY_INTERFACE float ospGetTaskDuration(OSPFuture);

typedef struct
{
  int hasHit;
  float worldPosition[3];
  OSPInstance instance;
  OSPGeometricModel model;
  uint32_t primID;
} OSPPickResult;

OSPRAY_INTERFACE void ospPick(OSPPickResult *result,
    OSPFrameBuffer,
    OSPRenderer,
    OSPCamera,
    OSPWorld,
    float screenPos_x,
    float screenPos_y);

#ifdef __cplusplus
} /* extern "C" */
#endif
