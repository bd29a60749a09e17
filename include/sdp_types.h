/*
 * sdp_types.h -- data types shared by the actor C-ABI (sdp_actors.h).
 *
 * These types must match the reference actor library byte for byte, because
 * PREESM-generated schedules pass pointers into one static pool and the
 * actor reads them in place.  Layout sources:
 *   PRECISION / PRECISION2 / PRECISION3 / int2 :
 *       preesm_pipelines/simsdp_fft_imaging_pipeline/Code/include/common.h:62-168
 *   Config : .../Code/include/common.h:187-265
 *   IN / OUT tags : .../Code/include/preesm.h:14-15
 *
 * Only the reference's default SINGLE_PRECISION=1 build is supported
 * (PRECISION is float).  On x86-64 SysV the structure sizes are checked at
 * compile time below (sizeof(Config) == 360 and the offsets the actors use).
 */
#ifndef SDP_TYPES_H
#define SDP_TYPES_H

#include <stdbool.h>
#include <stddef.h>

#ifndef IN
#define IN
#endif
#ifndef OUT
#define OUT
#endif

#define SDP_SINGLE_PRECISION 1
#define PRECISION float
#define PRECISION2 float2
#define PRECISION3 float3

/* The HIP compiler provides its own vector types under these names; a
 * plain C (or non-GPU C++) includer gets the reference's struct layouts. */
#if !(defined(__cplusplus) && (defined(__HIPCC__) || defined(__HIP__)))
#ifndef SDP_HAVE_VECTOR_TYPES
#define SDP_HAVE_VECTOR_TYPES
typedef struct { int x; int y; } int2;
typedef struct { float x; float y; } float2;
typedef struct { float x; float y; float z; } float3;
#endif
#endif

/* Runtime configuration record, same field order and types as the
 * reference (common.h:187-265); the pipeline copies it between actors. */
typedef struct Config {
    bool retain_device_mem;
    int num_major_cycles;
    int num_recievers;
    int num_baselines;
    int grid_size;
    double cell_size;
    const char *output_path;
    const char *dirty_image_output;
    int gpu_max_threads_per_block;
    int gpu_max_threads_per_block_dimension;
    bool right_ascension;

    bool save_dirty_image;
    bool save_residual_image;
    bool save_extracted_sources;
    bool save_predicted_visibilities;
    bool save_estimated_gains;

    bool perform_system_test;
    const char *system_test_image;
    const char *system_test_sources;
    const char *system_test_visibilities;
    const char *system_test_gains;

    const char *default_gains_file;
    const char *output_gains_file;
    bool perform_gain_calibration;
    bool use_default_gains;
    int max_calibration_cycles;
    int number_cal_major_cycles;

    bool force_weight_to_one;
    double frequency_hz;
    int oversampling;
    double uv_scale;
    int num_visibilities;
    int num_kernels;
    int total_kernel_samples;
    double max_w;
    double w_scale;
    const char *kernel_real_file;
    const char *kernel_imag_file;
    const char *kernel_support_file;
    const char *degridding_kernel_real_file;
    const char *degridding_kernel_imag_file;
    const char *degridding_kernel_support_file;
    const char *visibility_source_file;

    unsigned int number_minor_cycles_cal;
    unsigned int number_minor_cycles_img;
    double loop_gain;
    double weak_source_percent_gc;
    double weak_source_percent_img;
    double noise_factor;
    double psf_max_value;
    const char *model_sources_output;
    const char *residual_image_output;
    const char *psf_input_file;

    const char *psf_image_output;
    const char *clean_psf_image_output;
    const char *model_image_output;
    const char *final_image_output;

    int num_sources;

    const char *predicted_vis_output;
} Config;

#if defined(__x86_64__) && !defined(__cplusplus)
_Static_assert(sizeof(Config) == 360, "Config must match the reference layout");
_Static_assert(offsetof(Config, cell_size) == 24, "Config.cell_size offset");
_Static_assert(offsetof(Config, frequency_hz) == 128, "Config.frequency_hz offset");
_Static_assert(offsetof(Config, oversampling) == 136, "Config.oversampling offset");
_Static_assert(offsetof(Config, uv_scale) == 144, "Config.uv_scale offset");
_Static_assert(offsetof(Config, max_w) == 168, "Config.max_w offset");
_Static_assert(offsetof(Config, w_scale) == 176, "Config.w_scale offset");
_Static_assert(offsetof(Config, loop_gain) == 248, "Config.loop_gain offset");
_Static_assert(offsetof(Config, weak_source_percent_img) == 264, "Config.weak_source_percent_img offset");
_Static_assert(offsetof(Config, noise_factor) == 272, "Config.noise_factor offset");
_Static_assert(offsetof(Config, psf_max_value) == 280, "Config.psf_max_value offset");
#endif

#endif /* SDP_TYPES_H */
