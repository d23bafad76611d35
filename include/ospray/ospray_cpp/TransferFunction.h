// OSPRay C++ wrapper: TransferFunction (ospray_cpp/TransferFunction.h of OSPRay 3.x).
#pragma once

#include "ManagedObject.h"

namespace ospray {
namespace cpp {

class TransferFunction : public ManagedObject<OSPTransferFunction, OSP_TRANSFER_FUNCTION>
{
 public:
  TransferFunction(const std::string &type) : ManagedObject<OSPTransferFunction, OSP_TRANSFER_FUNCTION>(ospNewTransferFunction(type.c_str())) {}
  TransferFunction(OSPTransferFunction existing = nullptr) : ManagedObject<OSPTransferFunction, OSP_TRANSFER_FUNCTION>(existing) {}
};

static_assert(sizeof(TransferFunction) == sizeof(OSPTransferFunction), "cpp::TransferFunction holds only its handle");

} // namespace cpp

OSPTYPEFOR_SPECIALIZATION(cpp::TransferFunction, OSP_TRANSFER_FUNCTION);

} // namespace ospray
