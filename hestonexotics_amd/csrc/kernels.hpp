// kernels.hpp -- host-side launch wrappers of the HIP kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "layout.hpp"

namespace hexo {

enum class PathMode { Normal, Debug, Count };

// mart / geo select the FP32 kernel with the martingale-corrected step / the geometric control
// variate's accumulation (the FP64 kernel reads the request's flags itself)
const void* path_kernel_symbol(uint32_t precision, uint32_t ppt, PathMode mode, bool mart = false, bool geo = false);
size_t path_lds_bytes(uint32_t precision, uint32_t ppt, uint32_t block);
hipError_t launch_paths(uint32_t precision, uint32_t ppt, PathMode mode, uint32_t grid, uint32_t block,
                        hipStream_t st, const PathArgs& a, bool mart = false, bool geo = false);
// Exact prices of the geometric control variates (geo.hip); GeoArgs.state holds
// geo_state_doubles(n_items) doubles.
hipError_t launch_geo(const GeoArgs& a, hipStream_t st);
size_t geo_state_doubles(uint32_t n_items);
hipError_t launch_reduce(const ReduceArgs& a, uint64_t n_group_threads, uint64_t n_out_threads, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st);
hipError_t launch_zero_u64(unsigned long long* p, uint32_t n, hipStream_t st);
hipError_t launch_rng_fill(uint32_t* out, uint64_t seed, uint32_t stream, uint64_t path0, uint64_t n_paths,
                           uint64_t word0, uint64_t n_words, hipStream_t st);
hipError_t launch_philox(const uint32_t* ctr, const uint32_t* key, uint64_t n, uint32_t* out, hipStream_t st);
hipError_t launch_normals(const uint32_t* words, uint64_t n, uint32_t precision, double* out, hipStream_t st);
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, uint32_t grid, hipStream_t st);

}  // namespace hexo
