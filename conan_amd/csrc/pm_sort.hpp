// pm_sort.hpp -- host interface of the deterministic radix sort and scan.
#pragma once
#include "pm_common.hpp"

namespace cpm {

// Exclusive prefix sum of n uint32 values (in may equal out).  scratch must
// hold scan_scratch_elems(n) values.
size_t scan_scratch_elems(long n);
void exclusive_scan(const uint32_t* in, uint32_t* out, long n, uint32_t* scratch, hipStream_t st);

// Stable LSD radix sort of n (key, value) pairs on the low keybits bits,
// 8 or fewer bits per pass.
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

namespace cpm {

// Incremental re-sort of keys sorted one step ago (the gather of the FX step
// moves the particles into the sorted order, so the next step's keys are in
// nearly sorted order).  bprev: the group starts of the last sort (ngroups +
// 1), so the previous key of position s is the group that holds s;
// in_count / out_count: movers into / out of each group (keys_fx).
// resort_counts: stay[g] = stayers of g, tot[g] = stay[g] + in_count[g]
// (tot[ngroups] = 0; its exclusive scan gives the new group starts bnew);
// it then clears in_count, out_count and fill (the placement's counters), so
// after a re-sort the counts are zero for the next gather.
// resort_place, per brick of G = 2^R groups (np particles in all): perm[pos] = s;
// stayers keep their order at the front of their group, movers follow in
// atomic order (fill: per-group counters, zero on entry).
// group_keys: skeys[s] = the group of bstart that holds s (the sorted keys,
// for the kernels that compare new keys with them).
void resort_counts(const uint32_t* bprev, uint32_t* in_count, uint32_t* out_count, uint32_t* fill, long ngroups,
                   uint32_t* stay, uint32_t* tot, hipStream_t st);
void resort_place(const uint32_t* keys, const uint32_t* bprev, const uint32_t* bnew, const uint32_t* stay,
                  uint32_t* fill, long nbricks, int G, long np, uint32_t* perm, hipStream_t st);
void group_keys(const uint32_t* bstart, long nbricks, int G, uint32_t* skeys, hipStream_t st);
void zero_u32(uint32_t* p, long n, hipStream_t st);

}  // namespace cpm
