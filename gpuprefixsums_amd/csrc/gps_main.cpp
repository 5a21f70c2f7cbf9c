// gps_main: the reference's test-and-time program (GPUPrefixSumsCUDA/GPUPrefixSumsCUDA.cu:14-38)
// on the MI355X library, through the C++ dispatcher classes.
//
//   gps_main [log2_max_size] [batch]     defaults 28 and 100
//
// For each dispatcher: TestAllExclusive / TestAllInclusive, then
// BatchTimingInclusive(1 << log2_max_size, batch); then the vendor comparator
// (GPUPrefixSumsCUDA/CubDispatcher.cuh:32-76), which on ROCm times rocPRIM.
#include <cstdio>
#include <cstdlib>

#include "../../include/gps_dispatchers.hpp"
#include "../../include/gps_vendor.hpp"

int main(int argc, char** argv) {
    const uint32_t log2n = argc > 1 ? (uint32_t)atoi(argv[1]) : 28u;
    const uint32_t batch = argc > 2 ? (uint32_t)atoi(argv[2]) : 100u;
    if (log2n < 12 || log2n > 30) {
        fprintf(stderr, "log2_max_size must be in [12, 30]\n");
        return 2;
    }
    const uint32_t size = 1u << log2n;
    int failures = 0;
    try {
        {
            gps::ChainedScanDecoupledLookbackDispatcher csdl(size);
            csdl.TestAllExclusive();
            failures += csdl.TestsPassed() != csdl.TestsRun();
            csdl.TestAllInclusive();
            failures += csdl.TestsPassed() != csdl.TestsRun();
            csdl.BatchTimingInclusive(size, batch);
        }
        {
            gps::ReduceThenScanDispatcher rts(size);
            rts.TestAllExclusive();
            failures += rts.TestsPassed() != rts.TestsRun();
            rts.TestAllInclusive();
            failures += rts.TestsPassed() != rts.TestsRun();
            rts.BatchTimingInclusive(size, batch);
        }
        {
            gps::EmulatedDeadlockingDispatcher ed(size);
            ed.TestAllInclusive();
            failures += ed.TestsPassed() != ed.TestsRun();
            ed.TestAllExclusive();
            failures += ed.TestsPassed() != ed.TestsRun();
            ed.BatchTimingInclusive(size, batch);
        }
        {
            gps::CubDispatcher cub(size);
            cub.BatchTimingCubChainedScan(size, batch);
        }
    } catch (const gps::Error& e) {
        fprintf(stderr, "gps error %d: %s\n", e.code, e.what());
        return 1;
    }
    return failures ? 1 : 0;
}
