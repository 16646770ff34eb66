// Host launchers of the kernels compiled ahead of time into the library.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace rfk::sk {

// Catmull-Rom segment coefficients (assets/kernels/catmull.cu:11-42).
hipError_t catmull(const double* samples, uint32_t sample_size, uint32_t nseg, double* segments, hipStream_t s);

// Evaluate the cubic segments for every sample (block) at its own t, convert
// polar affines back to cartesian and build the RGBA8 palette
// (refactor.cu:105-151, 256-294). kind[i]: 0 plain, 1..4 component of an
// affine whose first real sits at i-(kind-1).
hipError_t eval_samples(const double* segments, uint32_t sample_size, uint32_t size_reals, uint32_t nseg,
                        uint32_t nsamples, const int8_t* kind, int precision, void* params, uint32_t param_stride,
                        uint32_t* palettes, hipStream_t s, const double* times = nullptr);

// graph-safe clear / fill (a kernel instead of hipMemsetAsync); the second
// form writes w instead of v into the m words from index at
hipError_t fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
hipError_t fill_u32_except(uint32_t* p, uint32_t v, size_t n, size_t at, size_t m, uint32_t w, hipStream_t s);

// Bin n plane-space points through a screen transform / palette.
// (hist64: the histogram holds u64 RGBA bins, else u32 RGBA)
hipError_t bin_points(const float* screen6, const uint32_t* palette, const float* pts, size_t n, uint32_t* hist,
                      bool hist64, uint32_t width, uint32_t height, hipStream_t s);

// calculate_histogram<4> (density_histo.cu:1-41) over the fixed-point histogram.
hipError_t density_histogram(const uint32_t* hist, bool hist64, size_t npix, unsigned long long* bins97,
                             unsigned long long* max_density, hipStream_t s);

// Density estimation: levels + gather. inv_z has `nlevels` entries.
hipError_t density_estimate(const uint32_t* hist, bool hist64, uint32_t width, uint32_t height, float max_radius,
                            float min_radius, float curve, const float* inv_z, int nlevels, int radius,
                            int* level_scratch, float* out_rgba, hipStream_t s);

// Tonemap (tonemap.cu:50-105). in_kind: 0 float4 draw units, 1 u32 RGBA
// fixed point, 2 u64 RGBA fixed point.
// the denoiser's 8 input channels (BF16) and FP32 base image from the bins
// (tonemap + features in one pass); with `hits` (device) the scale constant
// comes from the bin call's hit counter instead of scale_constant
hipError_t tonemap_features(const void* hist, bool hist64, size_t npix, float scale_constant,
                            const unsigned long long* hits, float brightness, float gamma, float vibrancy, void* feat,
                            float* base, hipStream_t s);
hipError_t tonemap(const void* in, int in_kind, size_t npix, void* out, int format, float scale_constant,
                   float brightness, float gamma, float vibrancy, hipStream_t s);

// Queue binning (queue_bin.hip): fold nparts * per_part draw records
// (partition p = the records of iterate block p, 64x64-pixel tile layout,
// see flame_kernels.h) into the histogram. counts: [ntiles][nparts] records
// per tile and partition, as the iterate kernel wrote them (overwritten).
// Scratch: sorted (8 bytes per record; holds the sorted colours and pixels),
// tile_scratch (2 * ntiles u32), segments
// (max_segments * 16 B), nseg (one u32). packed64: 32-bit bins updated as
// two 64-bit pairs.
uint32_t queue_tiles_x(uint32_t width);
uint32_t queue_tiles(uint32_t width, uint32_t height);
uint32_t queue_max_segments(size_t records, uint32_t ntiles);
hipError_t queue_bin(const unsigned long long* records, size_t per_part, uint32_t nparts, unsigned long long* sorted,
                     uint32_t* counts, uint32_t* tile_scratch, void* segments, uint32_t max_segments, uint32_t* nseg,
                     uint32_t width, uint32_t height, uint32_t* hist, bool hist64, bool packed64, hipStream_t s);
// The same for a chunk of several frames (animation batches): the parts come
// in groups of ncols, group g holding records of frame group_frame[g] only
// (tile fields inside that frame), counts [g * ntiles + tile][ncols];
// frame f's bins start at hist + f * frame_elems elements. tile_scratch
// holds 2 * ntiles * nparts / ncols u32.
hipError_t queue_bin_groups(const unsigned long long* records, size_t per_part, uint32_t nparts, uint32_t ncols,
                            unsigned long long* sorted, uint32_t* counts, uint32_t* tile_scratch, void* segments,
                            uint32_t max_segments, uint32_t* nseg, uint32_t width, uint32_t height,
                            const uint32_t* group_frame, size_t frame_elems, uint32_t* hist, bool hist64,
                            bool packed64, hipStream_t s);

// density_hdr (tonemap.cu:159-233) on a tonemapped float4 / half4 image,
// with the 97-bin density histogram (device u64) of the same render.
hipError_t density_hdr(const void* in, bool in_half, size_t npix, const unsigned long long* bins97, float* out,
                       hipStream_t s);

// convert_to (convert.cu:72-98): a half3 / half4 image (in_channels 3 / 4)
// to out_format 0 float4, 1 half4, 2 float3, 3 half3, 4 RGBA8, 5 RGB8.
hipError_t convert_image(const void* in, int in_channels, void* out, int out_format, uint32_t width, uint32_t height,
                         bool flip, bool swap, hipStream_t s);

// Random RGBA atomic increments for the binning roofline.
hipError_t atomic_peak(uint32_t* buf, uint64_t nbins, uint32_t grid, uint32_t block, uint32_t iters, uint32_t seed,
                       hipStream_t s);

}  // namespace rfk::sk
