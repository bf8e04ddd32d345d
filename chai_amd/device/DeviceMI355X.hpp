// DeviceMI355X.hpp -- the MI355X compute device of Chai's scheduler.
//
// A DeviceBase (runtime/DeviceBase.hpp:20-60) that evaluates a VectorTrace on
// one MI355X through the C ABI of libchai_mi355x.so (include/chai_mi355x.h),
// in place of the OpenCL Translator (runtime/Translator.cpp:366-411).  It is
// registered in DeviceMap (runtime/DeviceMap.cpp:26-36; here
// chai_amd/device/DeviceMap_mi355x.cpp) and called only from its scheduler
// thread (runtime/Scheduler.cpp:186) through Executor::dispatch.
//
// Memory stays the interpreter's: the device is initialised with a MemInterp
// (DeviceBase::initDevice(void), runtime/DeviceBase.cpp:72-78), so every
// array a statement assigns is host memory in the reference's own FrontMem /
// BackMem objects, and traces can move between this device and the
// interpreter freely.  sub_evaluate lowers the trace's statements to a
// chai_trace, uploads the arrays it reads, compiles and launches it, and
// writes every assigned array and every readout back to host memory.
//
// A trace the lowering or the back end cannot take makes sub_evaluate return
// false: the scheduler then re-queues it (runtime/Scheduler.cpp:221-254) and
// the interpreter evaluates it.
#ifndef CHAI_DEVICE_MI355X_HPP
#define CHAI_DEVICE_MI355X_HPP

#include <string>

#include "DeviceBase.hpp"

namespace chai_internal {

struct MI355XBackend;   // the C ABI entry points, resolved at initialisation

class DeviceMI355X : public DeviceBase
{
    const int      _ordinal;     // HIP device ordinal
    MI355XBackend* _be;          // NULL when the library or the GPU is missing
    std::string    _why;         // why the device cannot run (for logs)
    size_t         _traces, _fallbacks;

public:
    DeviceMI355X(const int deviceNum, const int ordinal);
    ~DeviceMI355X(void);

    void sub_initDevice(MemInterp&);
    bool sub_evaluate(VectorTrace&);

    // traces evaluated here / handed back to the scheduler
    size_t tracesEvaluated(void) const { return _traces; }
    size_t tracesRefused(void) const { return _fallbacks; }
};

}; // namespace chai_internal

#endif
