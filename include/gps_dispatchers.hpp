// gps_dispatchers.hpp - the reference's dispatcher classes, rebuilt over the
// C ABI in gps.h (implemented in libgps.so).
//
// Same class names, constructor argument and public methods as
//   GPUPrefixSumsCUDA/ChainedScanDecoupledLookbackDispatcher.cuh:16-163
//   GPUPrefixSumsCUDA/ReduceThenScanDispatcher.cuh:16-152
//   GPUPrefixSumsCUDA/EmulatedDeadlockingDispatcher.cuh:19-163
// plus public Inclusive/Exclusive entry points on caller buffers (the
// reference keeps its DispatchKernels* private and scans only its own
// buffer). Like the reference, the Test*/Batch* methods print their results;
// unlike it, nothing fails silently: errors throw gps::Error.
#pragma once

#include <stdint.h>

#include <stdexcept>
#include <string>

#include "gps.h"

namespace gps {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

// Shared implementation: owns m_scan (the in-place test/timing buffer, as in
// the reference), the C-ABI handle, and the last test tally.
class DispatcherBase {
   public:
    DispatcherBase(int algo, uint32_t maxSize, gps_tuning tuning);
    virtual ~DispatcherBase();
    DispatcherBase(const DispatcherBase&) = delete;
    DispatcherBase& operator=(const DispatcherBase&) = delete;

    // Scans d_in into d_out (d_in == d_out allowed) on `stream`.
    void Inclusive(const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream = nullptr);
    void Exclusive(const uint32_t* d_in, uint32_t* d_out, uint64_t n, void* stream = nullptr);

    uint32_t TestsPassed() const { return m_passed; }
    uint32_t TestsRun() const { return m_total; }
    uint32_t PartitionSize() const { return k_partitionSize; }
    gps_handle* Handle() const { return m_handle; }

   protected:
    // Reference TestAll* sweep (...Dispatcher.cuh:38-114), with its large-size
    // tests fixed to use 1 << i elements, plus a full-range random test.
    void TestAll(bool inclusive, const char* name);
    double BatchTiming(bool inclusive, uint32_t size, uint32_t batchCount, const char* name);

    const uint32_t k_maxSize;
    uint32_t k_partitionSize = 0;
    uint32_t* m_scan = nullptr;
    uint32_t* m_scratch = nullptr;  // second buffer for the random test
    gps_handle* m_handle = nullptr;
    uint32_t m_passed = 0, m_total = 0;
};

class ChainedScanDecoupledLookbackDispatcher : public DispatcherBase {
   public:
    explicit ChainedScanDecoupledLookbackDispatcher(uint32_t maxSize, gps_tuning tuning = gps_tuning{})
        : DispatcherBase(GPS_ALGO_CSDL, maxSize, tuning) {}
    void TestAllExclusive() { TestAll(false, "ChainedScanDecoupledLookback exclusive"); }
    void TestAllInclusive() { TestAll(true, "ChainedScanDecoupledLookback inclusive"); }
    double BatchTimingInclusive(uint32_t size, uint32_t batchCount) {
        return BatchTiming(true, size, batchCount, "ChainedScanDecoupledLookback inclusive");
    }
};

class ReduceThenScanDispatcher : public DispatcherBase {
   public:
    explicit ReduceThenScanDispatcher(uint32_t maxSize, gps_tuning tuning = gps_tuning{})
        : DispatcherBase(GPS_ALGO_RTS, maxSize, tuning) {}
    void TestAllExclusive() { TestAll(false, "ReduceThenScan exclusive"); }
    void TestAllInclusive() { TestAll(true, "ReduceThenScan inclusive"); }
    double BatchTimingInclusive(uint32_t size, uint32_t batchCount) {
        return BatchTiming(true, size, batchCount, "ReduceThenScan inclusive");
    }
};

// Decoupled fallback under emulated deadlock. The reference stalls tiles with
// !(idx & DEADLOCK_MASK), DEADLOCK_MASK = 7 (EmulatedDeadlocking.cuh:16,65),
// i.e. stall period 8; that is the default here. Its timing overload takes
// partMask/maxSpinCount but ignores them (EmulatedDeadlockingDispatcher.cuh:79-80);
// here they take effect (period = partMask + 1).
class EmulatedDeadlockingDispatcher : public DispatcherBase {
   public:
    explicit EmulatedDeadlockingDispatcher(uint32_t maxSize, uint32_t stallPeriod = 8, uint32_t maxSpinCount = 4);
    void TestAllInclusive() { TestAll(true, "EmulatedDeadlocking inclusive"); }
    void TestAllExclusive() { TestAll(false, "EmulatedDeadlocking exclusive"); }
    double BatchTimingInclusive(uint32_t size, uint32_t batchCount, uint32_t partMask, uint32_t maxSpinCount);
    double BatchTimingInclusive(uint32_t size, uint32_t batchCount) {
        return BatchTimingInclusive(size, batchCount, m_stallPeriod - 1, m_spinLimit);
    }
    // EmulatedDeadlockingDispatcher.cuh:128-138: keys/s at 2^28 for stall
    // periods 512..64, written to data.txt.
    void RecordTimingData(uint32_t spinCount);

   private:
    void Rebuild(uint32_t stallPeriod, uint32_t spinLimit);
    uint32_t m_stallPeriod, m_spinLimit;
};

}  // namespace gps
