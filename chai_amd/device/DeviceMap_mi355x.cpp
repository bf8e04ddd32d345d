// DeviceMap_mi355x.cpp -- Chai's device table with MI355X devices.
//
// The registration point of compute devices (runtime/DeviceMap.hpp:18-35;
// the reference's runtime/DeviceMap.cpp:13-36 creates the interpreter and one
// OpenCL Translator per configured OpenCL device).  This definition of the
// same class keeps both and adds one DeviceMI355X per HIP ordinal listed in
// the environment variable CHAI_MI355X (comma separated, e.g. "0"), numbered
// after the OpenCL devices.  Linked ahead of the reference's library, it
// replaces runtime/DeviceMap.o; nothing else of the runtime changes.
#include <stdlib.h>

#include <sstream>
#include <string>

#include "DeviceMap.hpp"
#include "DeviceMI355X.hpp"
#include "Interpreter.hpp"
#include "OCLhacks.hpp"
#include "Translator.hpp"

using namespace std;

namespace chai_internal {

DeviceMap::DeviceMap(OCLinit& oclInit, istream& configSpec)
{
    // configuration first: it decides which devices exist
    OCLhacks::singleton().initHacks(oclInit, configSpec);

    int next = 0;
    if (OCLhacks::singleton().enableInterpreter()) {
        Interpreter* interp = new Interpreter(-1);
        _allDevnums.insert(-1);
        _deviceMap[-1] = interp;
        interp->initDevice();
    }

    const set< size_t >& ocl = OCLhacks::singleton().deviceIndexes();
    for (set< size_t >::const_iterator it = ocl.begin(); it != ocl.end(); it++) {
        const int devnum = (int)*it;
        Translator* jit = new Translator(devnum);
        _allDevnums.insert(devnum);
        _deviceMap[devnum] = jit;
        jit->initDevice(oclInit);
        if (devnum + 1 > next) next = devnum + 1;
    }

    // MI355X devices share the interpreter's host memory manager
    const char* spec = getenv("CHAI_MI355X");
    if (spec && *spec) {
        stringstream ss(spec);
        string tok;
        while (getline(ss, tok, ',')) {
            if (tok.empty()) continue;
            const int ordinal = atoi(tok.c_str());
            DeviceMI355X* gpu = new DeviceMI355X(next, ordinal);
            _allDevnums.insert(next);
            _deviceMap[next] = gpu;
            gpu->initDevice();
            next++;
        }
    }
}

DeviceMap::~DeviceMap(void)
{
    for (map< int, DeviceBase* >::const_iterator it = _deviceMap.begin(); it != _deviceMap.end(); it++)
        delete it->second;
}

// language extensions reach the devices that take them: the interpreter's
// handler and the OpenCL translators' (an MI355X device refuses traces with
// opcodes it does not know, and the interpreter evaluates them)
void DeviceMap::extendLanguage(const uint32_t opCode, const BaseInterp& interpHandler,
                               const BaseTrans& jitHandler)
{
    for (map< int, DeviceBase* >::const_iterator it = _deviceMap.begin(); it != _deviceMap.end(); it++) {
        if (it->first < 0)
            it->second->extendLanguage(opCode, interpHandler);
        else if (dynamic_cast< Translator* >(it->second))
            it->second->extendLanguage(opCode, jitHandler);
    }
}

const set< int >& DeviceMap::allDevnums(void) const
{
    return _allDevnums;
}

DeviceBase* DeviceMap::getDevice(const int deviceNum)
{
    return _deviceMap[deviceNum];
}

}; // namespace chai_internal
