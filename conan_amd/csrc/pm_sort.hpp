// pm_sort.hpp -- host interface of the deterministic radix sort and scan.
#pragma once
#include "pm_common.hpp"

namespace cpm {

// Exclusive prefix sum of n uint32 values (in may equal out).  scratch must
// hold scan_scratch_elems(n) values.
size_t scan_scratch_elems(long n);
void exclusive_scan(const uint32_t* in, uint32_t* out, long n, uint32_t* scratch, hipStream_t st);

// Stable LSD radix sort of n (key, value) pairs on the low keybits bits,
// 9 or fewer bits per pass.
struct RadixSorter {
    long n;
    int keybits, passes, dbits, nblocks;
    long ncounts;
    uint32_t* counts = nullptr;
    uint32_t* scratch = nullptr;
    RadixSorter(long n, int keybits);
    ~RadixSorter();
    RadixSorter(const RadixSorter&) = delete;
    RadixSorter& operator=(const RadixSorter&) = delete;
    // Sorts (k0, v0) using (k1, v1) as the other buffer; returns 0 when the
    // result is in (k0, v0) and 1 when it is in (k1, v1).
    int sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, hipStream_t st);
};

}  // namespace cpm
