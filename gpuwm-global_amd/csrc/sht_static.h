// The library-owned device array behind arwen_sht_sample_static (sht_static.cpp).
#pragma once

#include <hip/hip_runtime.h>

// Resolves this device's copy of the array (hipGetSymbolAddress), once per
// device; call with `device` current.  Called when a plan is created.
hipError_t sht_static_setup(int device);
// The resolved address of the device's copy, or NULL before sht_static_setup.
void* sht_static_scratch(int device);
