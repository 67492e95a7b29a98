// CPU oracle (TEST INFRASTRUCTURE ONLY): the comparison sort of the
// reference CPU builder, for the timed CPU baseline.
//
// BVHCPUBuilder::build sorts its IndexedMorton array with std::sort ordered
// by the Morton key alone (minimum_lbvh.h:554): an introsort, unstable, so
// records with equal keys end in an unspecified order.  Records here are the
// oracle's u64 layout (key in the low word, triangle index in the high word),
// compared on the low word only, as the reference compares .morton.
#include <algorithm>
#include <cstdint>

extern "C" void orc_sort_records_comparison(uint64_t* records, int n)
{
    std::sort(records, records + n, [](uint64_t a, uint64_t b) { return (uint32_t)a < (uint32_t)b; });
}
