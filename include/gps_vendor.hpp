// gps_vendor.hpp - the reference's vendor-library comparator class
// (GPUPrefixSumsCUDA/CubDispatcher.cuh:21-76) on rocPRIM, implemented in
// libgps_vendor.so (not in the product library libgps.so).
//
// The class keeps the reference's name and method so callers switch
// unchanged. Like the reference it times the in-place inclusive scan of
// all-ones input, re-filled before every iteration, first iteration
// discarded; it returns keys/s. Unlike the reference it checks the result.
#pragma once

#include <stdint.h>

#include "gps_dispatchers.hpp"
#include "gps_vendor.h"

namespace gps {

class CubDispatcher {
   public:
    explicit CubDispatcher(uint32_t size);
    ~CubDispatcher();
    CubDispatcher(const CubDispatcher&) = delete;
    CubDispatcher& operator=(const CubDispatcher&) = delete;
    double BatchTimingCubChainedScan(uint32_t size, uint32_t batchCount);

   private:
    const uint32_t k_maxSize;
    uint32_t* m_scan = nullptr;
};

}  // namespace gps
