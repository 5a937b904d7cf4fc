// primitives.hpp — device-wide building blocks (stable radix sort, scans,
// fills) shared by the tree build, the walk set-up and the PM.
#pragma once

#include "common.hpp"

namespace chui {

// Scratch owned by the caller, sized with radix_sort_scratch_bytes().
std::size_t radix_sort_scratch_bytes(std::size_t n);

// Stable LSD radix sort of (key, value) pairs on bits [begin_bit, end_bit).
// keys/vals are sorted in place (ping-pong with the alt buffers; the result is
// always copied back to keys/vals). Equal keys keep their input order, the
// property the reference's Morton sort relies on (tree_ordering.cpp:13-71).
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                      std::size_t n, int begin_bit, int end_bit, void* scratch, hipStream_t s);
// Same for 32-bit keys.
void radix_sort_pairs32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                        std::size_t n, int begin_bit, int end_bit, void* scratch, hipStream_t s);

// Exclusive prefix sum of n uint32 values (in -> out, may alias); total written to
// *total_dev (device pointer, may be null). scratch >= scan_scratch_bytes(n).
std::size_t scan_scratch_bytes(std::size_t n);
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, std::size_t n, uint32_t* total_dev,
                        void* scratch, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, std::size_t n, hipStream_t s);
void fill_u64(uint64_t* p, uint64_t v, std::size_t n, hipStream_t s);
void fill_f64(double* p, double v, std::size_t n, hipStream_t s);
void fill_f32(float* p, float v, std::size_t n, hipStream_t s);
void iota_u32(uint32_t* p, std::size_t n, hipStream_t s);

}  // namespace chui
