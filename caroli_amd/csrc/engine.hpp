// The GPU engine: batched multifrontal elimination of many transmission points on one device.
#pragma once
#include <cstdint>

#include "../../include/caroli_amd.h"
#include "system.hpp"

namespace caroli::gpu {

/// Transmission of `count` points (systems[p], sigmas[p]) in `precision`, level by level over
/// the elimination trees, all points of one structure per launch.
int32_t batch_transmission(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count,
                           int32_t precision, double* out_t, int32_t* out_status, cg_batch_stats* stats);

int32_t create_batch(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count, int32_t precision,
                     cg_batch** out);
int32_t run_batch(cg_batch* b, double* out_t, int32_t* out_status, cg_batch_stats* stats);
/// A batch of multi-terminal points: T between labels, nlab x nlab per point (row: into).
int32_t run_batch_labels(cg_batch* b, double* out_lab, int32_t* out_status, cg_batch_stats* stats);
void destroy_batch(cg_batch* b);
int32_t get_tuning(cg_tuning* out);
int32_t set_tuning(const cg_tuning* t);
int32_t autotune_batch(cg_batch* b, cg_tuning* out, double* out_seconds);
int32_t kernel_timing(int32_t enable);
int32_t front_schur(const double* f, int64_t n, int64_t ne, int64_t ne1, int32_t precision, double* out_s,
                    int32_t* out_status);
int32_t kernel_stats(cg_kernel_stat* out, int32_t max_records);
/// front_schur with its work area in a static device array (no device allocation).
constexpr int64_t kNoAllocMaxN = CG_FRONT_NOALLOC_MAX_N;
int32_t front_schur_noalloc(const double* f, int64_t n, int64_t ne, int64_t ne1, int32_t precision, double* out_s,
                            int32_t* out_status);

/// T and dT/d(eps_a) for every orbital, from the stored fronts of one point.
int32_t backward(const cg_system* s, const cg_sigma* g, double* out_t, double* out_grad);
/// The same for `count` points of one structure in one sweep (out_grad: count x n).
int32_t backward_batch(const cg_system* const* systems, const cg_sigma* const* sigmas, int64_t count, double* out_t,
                       double* out_grad, int32_t* out_status);

/// psi_L = G E_L W_L and psi_R = G E_R W_R on every orbital (cols[0], cols[1] channels, the
/// blocks' columns one after another); with both outputs NULL only the channel counts.
int32_t scattering_states(const cg_system* s, const cg_sigma* g, int64_t* cols, double* out_l, double* out_r);

/// A multi-terminal point, block b, from one complex f64 sweep with the factors kept: G on the
/// block's kept orbitals (green), and/or the states G W of every label's channels on all orbitals
/// (n x the block's channel columns, complex, column-major). CG_ERR_INVALID for delayed separators.
int32_t multi_block_solve(const cg_system* s, const cg_sigma* g, int b, caroli::CMat* green,
                          std::vector<caroli::cplx>* states);

bool device_available();

/// Frees the device copies cached for a system.
void release(const cg_system* s);

}  // namespace caroli::gpu
